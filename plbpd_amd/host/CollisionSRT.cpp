#include "CollisionSRT.hpp"

CollisionSRT::CollisionSRT(const LBModel* lbmodel, const Domain* domain)
    : Collision(), tau_(domain->get_relaxation_time()), force_(domain->get_external_force_f64()) {
    (void)lbmodel;
}

CollisionSRT::~CollisionSRT() {}

// Its GPU time is counted by the state's timers (Collision::get_total_time).
void CollisionSRT::operator()(SimData& simdata, bool reference) const {
    (void)reference;  // the serial and parallel loops compute the same values
    plbpd::check(plb_collide_srt(simdata.device(), tau_, force_.data()));
}
