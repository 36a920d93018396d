// Command-line runner: the reference's main() (main.cpp:13-31) with the
// scenario, storage variant, step semantics and step count as options.
//   plbpd_run [Couette2D|Couette3D|Poiseuille2D|Poiseuille3D|<scenario file>]
//             [--fused] [--variant f64|f32|f64a32|f32s|f16s] [--steps N]
//             [--out FILE] [--device-times]
// print_times reports the GPU time of each reference stage over the run;
// --device-times adds the per-stage times of the one device state.
// A scenario file (plbpd_scenario.hpp) may set the variant, the mode, the
// initial state and interior solids; command-line options take precedence.
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "Flow.hpp"
#include "PLBPD.hpp"
#include "plbpd_state.hpp"

int main(int argc, char** argv) {
    std::string flow_name = "Couette3D", out = "final_state.h5", variant = "f64";
    long steps = -1;
    bool fused = false, variant_given = false, device_times = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--fused") fused = true;
        else if (a == "--variant" && i + 1 < argc) variant = argv[++i], variant_given = true;
        else if (a == "--steps" && i + 1 < argc) steps = std::atol(argv[++i]);
        else if (a == "--out" && i + 1 < argc) out = argv[++i];
        else if (a == "--device-times") device_times = true;
        else if (a[0] != '-') flow_name = a;
        else {
            std::cerr << "unknown option " << a << "\n";
            return 2;
        }
    }
    plbpd::DeviceConfig cfg;
    cfg.mode = fused ? PLB_MODE_FUSED : PLB_MODE_LITERAL;
    if (variant == "f64") cfg.variant = PLB_F64;
    else if (variant == "f32") cfg.variant = PLB_F32;
    else if (variant == "f64a32") cfg.variant = PLB_F64_A32;
    else if (variant == "f32s") cfg.variant = PLB_F32S;
    else if (variant == "f16s") cfg.variant = PLB_F16S;
    else {
        std::cerr << "unknown variant " << variant << "\n";
        return 2;
    }
    try {
        const auto flow = std::make_unique<Flow>(flow_name);
        const plbpd::Scenario* sc = flow->get_scenario();
        if (sc && sc->has_variant && !variant_given) cfg.variant = sc->variant;
        if (sc && sc->has_mode && !fused) cfg.mode = sc->mode;
        plbpd::set_default_device_config(cfg);
        const auto lbmodel = flow->get_lbmodel();
        const auto domain = flow->get_domain();
        const auto boundary = flow->get_boundary();
        const auto problem = std::make_unique<PLBPD>(lbmodel, domain, boundary);
        SimData simdata(domain->get_dimensions(), lbmodel->get_num_directions());
        if (sc) {
            const std::vector<uint8_t> solids = plbpd::scenario_solids(*sc);
            if (!solids.empty()) boundary->set_solid_mask(simdata, solids.data());
        }
        problem->initialize(simdata);
        problem->march_in_time(steps >= 0 ? (size_t)steps : flow->get_num_timesteps(), simdata);
        simdata.write_state(out);
        problem->print_times();
        if (device_times) problem->print_device_times(simdata);
        std::cout << "wrote " << out << (plbpd::state_files_are_hdf5() ? " (HDF5)" : " (raw)") << "\n";
    } catch (const std::exception& e) {
        std::cerr << "plbpd_run: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
