// Drop-in PLBPD driver (interface of the reference's PLBPD.hpp:14-28).  The
// step loop is the reference's: collide, push, no-slip, periodicity,
// moments, reset (PLBPD.cpp:23-30), each call reaching the GPU.
#ifndef PLBPD_AMD_HOST_PLBPD_HPP_
#define PLBPD_AMD_HOST_PLBPD_HPP_

#include <memory>

#include "Boundary.hpp"
#include "CalcMoments.hpp"
#include "Domain.hpp"
#include "LBDynamics.hpp"
#include "LBModel.hpp"
#include "SimData.hpp"

class PLBPD final {
 public:
    PLBPD(const LBModel* lbmodel, const Domain* domain, const Boundary* boundary);
    PLBPD(PLBPD&) = delete;
    PLBPD& operator=(PLBPD&) = delete;
    ~PLBPD();
    void initialize(SimData& simdata) const;
    void march_in_time(size_t num_timesteps, SimData& simdata) const;
    // The reference's per-class timers (PLBPD.cpp:37-48), in GPU seconds
    // summed over every device state of the process.
    void print_times() const;
    // GPU time per stage of this state alone (HIP events), in seconds,
    // accumulated while its timers are on (plbpd::DeviceConfig::device_timers,
    // the default).  No reference counterpart.
    void print_device_times(SimData& simdata) const;

 private:
    const Domain* _domain;
    const Boundary* _boundary;
    std::unique_ptr<LBDynamics> _lbdynamics;
    std::unique_ptr<CalcMoments> _calc_moments;
};

#endif  // PLBPD_AMD_HOST_PLBPD_HPP_
