// Drop-in CalcMoments (interface of the reference's CalcMoments.hpp:15-74):
// rho and u on the interior, on the device.  In fused mode the moments are
// computed from the populations when they are read.
#ifndef PLBPD_AMD_HOST_CALCMOMENTS_HPP_
#define PLBPD_AMD_HOST_CALCMOMENTS_HPP_

#include <array>

#include "LBModel.hpp"
#include "SimData.hpp"

class CalcMoments final {
 public:
    explicit CalcMoments(const LBModel* lbmodel);
    CalcMoments(const CalcMoments&) = delete;
    CalcMoments& operator=(const CalcMoments&) = delete;
    ~CalcMoments();
    float get_total_time() const;
    void operator()(SimData& simdata) const;
    // FP64 sums over fluid cells: {sum rho, sum rho ux, sum rho uy, sum rho uz}.
    std::array<double, 4> global_moments(SimData& simdata) const;
};

#endif  // PLBPD_AMD_HOST_CALCMOMENTS_HPP_
