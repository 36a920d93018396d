#include "CalcMoments.hpp"

CalcMoments::CalcMoments(const LBModel* lbmodel) { (void)lbmodel; }
CalcMoments::~CalcMoments() {}

// GPU seconds of the moment passes run alone (a literal pass folded into the
// next collide + push kernel counts under the collision).
float CalcMoments::get_total_time() const { return (float)plbpd::stage_seconds()[PLB_STAGE_MOMENTS]; }

void CalcMoments::operator()(SimData& simdata) const { plbpd::check(plb_calc_moments(simdata.device())); }

std::array<double, 4> CalcMoments::global_moments(SimData& simdata) const {
    std::array<double, 4> out{};
    plbpd::check(plb_global_moments(simdata.device(), out.data()));
    return out;
}
