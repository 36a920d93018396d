#include "PLBPD.hpp"

#include <chrono>
#include <iostream>

PLBPD::PLBPD(const LBModel* lbmodel, const Domain* domain, const Boundary* boundary)
    : _domain(domain), _boundary(boundary), _lbdynamics(new LBDynamics(lbmodel, domain)),
      _calc_moments(new CalcMoments(lbmodel)) {}

PLBPD::~PLBPD() {}

void PLBPD::initialize(SimData& simdata) const {
    _domain->initialize(simdata);
    _boundary->reset(simdata);
}

void PLBPD::march_in_time(size_t num_timesteps, SimData& simdata) const {
    const auto start = std::chrono::steady_clock::now();
    for (size_t istep = 0; istep < num_timesteps; ++istep) {
        _lbdynamics->collide(simdata);
        _lbdynamics->stream(simdata, "push");
        _boundary->apply_noslip(simdata);
        _boundary->apply_periodicity(simdata);
        (*_calc_moments)(simdata);
        _boundary->reset(simdata);
    }
    plbpd::check(plb_synchronize(simdata.device()));  // the loop time includes the GPU work
    const std::chrono::duration<double> elapsed = std::chrono::steady_clock::now() - start;
    std::cout << "Time taken by timeloop: " << elapsed.count() << "s" << std::endl;
}

void PLBPD::print_times() const {
    std::cout << "LBDynamics: " << _lbdynamics->get_total_time() << "s\n";
    std::cout << "-collide: " << _lbdynamics->get_time_collide() << "s\n";
    std::cout << "-stream: " << _lbdynamics->get_time_stream() << "s\n";
    std::cout << "CalcMoments: " << _calc_moments->get_total_time() << "s\n";
    std::cout << "Boundary: " << _boundary->get_total_time() << "s\n";
    std::cout << "-noslip: " << _boundary->get_time_noslip() << "s\n";
    std::cout << "-periodicity: " << _boundary->get_time_periodicity() << "s\n";
    std::cout << "-reset: " << _boundary->get_time_reset() << "s\n";
    std::cout << "--velocity-reset: " << _boundary->get_time_velocity_reset() << "s\n";
    std::cout << "--density-reset: " << _boundary->get_time_density_reset() << "s\n";
}

void PLBPD::print_device_times(SimData& simdata) const {
    double ms[PLB_NSTAGES];
    plbpd::check(plb_stage_times(simdata.device(), ms));
    static const char* names[PLB_NSTAGES] = {"collide", "stream", "noslip", "periodicity",
                                             "moments", "velocity-reset", "density-reset", "fused"};
    for (int i = 0; i < PLB_NSTAGES; ++i)
        if (ms[i] > 0) std::cout << "device " << names[i] << ": " << ms[i] * 1e-3 << "s\n";
}
