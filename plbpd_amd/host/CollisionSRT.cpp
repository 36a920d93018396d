#include "CollisionSRT.hpp"

namespace chrono = std::chrono;

CollisionSRT::CollisionSRT(const LBModel* lbmodel, const Domain* domain)
    : Collision(), tau_(domain->get_relaxation_time()), force_(domain->get_external_force_f64()) {
    (void)lbmodel;
}

CollisionSRT::~CollisionSRT() {}

void CollisionSRT::operator()(SimData& simdata, bool reference) const {
    const auto start = chrono::system_clock::now();
    // The reference's default loop (reference = false) is its AVX2 kernel in
    // the default build; FP32 literal states reproduce its rounding.
    const plbpd::DeviceConfig& cfg = simdata.device_config();
    if (cfg.mode == PLB_MODE_LITERAL && cfg.variant == PLB_F32)
        plbpd::check(plb_set_tuning(simdata.device(), PLB_TUNE_LITERAL_AVX2, reference ? 0 : 1));
    plbpd::check(plb_collide_srt(simdata.device(), tau_, force_.data()));
    const chrono::duration<float> elapsed = chrono::system_clock::now() - start;
    Collision::_time += elapsed.count();
}
