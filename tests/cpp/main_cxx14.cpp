// A caller of the drop-in written to the reference's public interface and
// built as C++14, the reference's language standard (src/CMakeLists.txt:6).
// It drives the calls the reference's own main() makes (main.cpp:13-31):
// a Flow preset supplies the lattice, domain and boundary; PLBPD is built
// from them; a SimData is copy-initialised from a temporary -- which C++14
// accepts only when SimData has an accessible copy or move constructor --;
// the state is initialised and written, marched, written again, and the
// class timers are printed.
//
//   main_cxx14 [preset] [steps]      (defaults: Couette3D, the preset's count)
//
// It writes <preset>_begin.h5 and <preset>_end.h5 into the working directory
// and prints the FP64 mass sum of the final state.
#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>

#include "Boundary.hpp"
#include "CalcMoments.hpp"
#include "Domain.hpp"
#include "Flow.hpp"
#include "LBModel.hpp"
#include "PLBPD.hpp"
#include "SimData.hpp"

#if __cplusplus != 201402L
#error "this program checks the drop-in headers under C++14"
#endif

namespace {

SimData make_state(const Flow& flow) {
    const LBModel* model = flow.get_lbmodel();
    auto state = SimData(flow.get_domain()->get_dimensions(), model->get_num_directions());
    return state;
}

int run(const std::string& preset, long steps) {
    std::unique_ptr<Flow> flow(new Flow(preset));
    std::unique_ptr<PLBPD> solver(new PLBPD(flow->get_lbmodel(), flow->get_domain(), flow->get_boundary()));
    SimData state = make_state(*flow);
    solver->initialize(state);
    state.write_state(preset + "_begin.h5");
    const size_t n = steps > 0 ? (size_t)steps : flow->get_num_timesteps();
    solver->march_in_time(n, state);
    state.write_state(preset + "_end.h5");
    solver->print_times();
    const CalcMoments sums(flow->get_lbmodel());
    std::cout << preset << ": " << n << " steps, mass " << sums.global_moments(state)[0] << "\n";
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string preset = argc > 1 ? argv[1] : "Couette3D";
    const long steps = argc > 2 ? std::atol(argv[2]) : 0;
    try {
        return run(preset, steps);
    } catch (const std::exception& e) {
        std::cerr << "main_cxx14: " << e.what() << "\n";
        return 1;
    }
}
