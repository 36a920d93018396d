// Drop-in CollisionSRT (interface of the reference's CollisionSRT.hpp:11-45):
// BGK with the velocity-shift body force, on the device
// (plb_collide_srt).  `reference` selects the serial or the TBB loop in the
// reference (CollisionSRT.cpp:25-33); both loops call the same scalar kernel
// in a build without AVX2 (CollisionSRT.cpp:60-64, 153-176), so both values
// run the same GPU kernel, in that kernel's arithmetic (in FP32, its
// float/double promotions).  The AVX2 kernel of builds with -DAVX2 -mfma
// rounds differently and is not reproduced.
#ifndef PLBPD_AMD_HOST_COLLISIONSRT_HPP_
#define PLBPD_AMD_HOST_COLLISIONSRT_HPP_

#include <array>
#include <vector>

#include "Collision.hpp"
#include "Domain.hpp"
#include "LBModel.hpp"

class CollisionSRT final : public Collision {
 public:
    CollisionSRT() = delete;
    CollisionSRT(const LBModel* lbmodel, const Domain* domain);
    CollisionSRT(CollisionSRT&) = delete;
    CollisionSRT& operator=(CollisionSRT&) = delete;
    ~CollisionSRT();
    void operator()(SimData& simdata, bool reference = false) const;
    double get_tau() const { return tau_; }

 private:
    double tau_;
    std::array<double, 3> force_;
};

#endif  // PLBPD_AMD_HOST_COLLISIONSRT_HPP_
