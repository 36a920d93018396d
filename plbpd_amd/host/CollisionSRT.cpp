#include "CollisionSRT.hpp"

namespace chrono = std::chrono;

CollisionSRT::CollisionSRT(const LBModel* lbmodel, const Domain* domain)
    : Collision(), tau_(domain->get_relaxation_time()), force_(domain->get_external_force_f64()) {
    (void)lbmodel;
}

CollisionSRT::~CollisionSRT() {}

void CollisionSRT::operator()(SimData& simdata, bool reference) const {
    (void)reference;
    const auto start = chrono::system_clock::now();
    plbpd::check(plb_collide_srt(simdata.device(), tau_, force_.data()));
    const chrono::duration<float> elapsed = chrono::system_clock::now() - start;
    Collision::_time += elapsed.count();
}
