// Drop-in CollisionSRT (interface of the reference's CollisionSRT.hpp:11-45):
// BGK with the velocity-shift body force, on the device
// (plb_collide_srt).  `reference` selects the serial or the TBB loop in the
// reference (CollisionSRT.cpp:25-33).  In FP64 the two agree and run the
// same GPU kernel.  In an FP32 literal state `reference` = true gives the
// serial kernel's arithmetic (CollisionSRT.cpp:153-176) and false, the
// default, that of the AVX2 kernel of the reference's default build (FP32
// fused multiply-adds for the first 8 * floor(q / 8) directions,
// CollisionSRT.cpp:93-149), through PLB_TUNE_LITERAL_AVX2.
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
