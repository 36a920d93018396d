#include "PLBPD.hpp"

#include <chrono>
#include <iostream>
#include <string>

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

// The class timers, one line each, nested by how the reference groups them
// (a leading '-' per level).  Their values are GPU seconds (HIP events).
void PLBPD::print_times() const {
    struct Row {
        int depth;
        const char* name;
        float seconds;
    };
    const Boundary& b = *_boundary;
    const Row rows[] = {{0, "LBDynamics", _lbdynamics->get_total_time()},
                        {1, "collide", _lbdynamics->get_time_collide()},
                        {1, "stream", _lbdynamics->get_time_stream()},
                        {0, "CalcMoments", _calc_moments->get_total_time()},
                        {0, "Boundary", b.get_total_time()},
                        {1, "noslip", b.get_time_noslip()},
                        {1, "periodicity", b.get_time_periodicity()},
                        {1, "reset", b.get_time_reset()},
                        {2, "velocity-reset", b.get_time_velocity_reset()},
                        {2, "density-reset", b.get_time_density_reset()}};
    for (const Row& r : rows) std::cout << std::string(r.depth, '-') << r.name << ": " << r.seconds << "s\n";
}

void PLBPD::print_device_times(SimData& simdata) const {
    double ms[PLB_NSTAGES];
    plbpd::check(plb_stage_times(simdata.device(), ms));
    static const char* names[PLB_NSTAGES] = {"collide", "stream", "noslip", "periodicity",
                                             "moments", "velocity-reset", "density-reset", "fused"};
    for (int i = 0; i < PLB_NSTAGES; ++i)
        if (ms[i] > 0) std::cout << "device " << names[i] << ": " << ms[i] * 1e-3 << "s\n";
}
