#include "CalcMoments.hpp"

namespace chrono = std::chrono;

float CalcMoments::_time_calc_moment = 0.0f;

CalcMoments::CalcMoments(const LBModel* lbmodel) { (void)lbmodel; }
CalcMoments::~CalcMoments() {}
float CalcMoments::get_total_time() const { return _time_calc_moment; }

void CalcMoments::operator()(SimData& simdata) const {
    const auto start = chrono::system_clock::now();
    plbpd::check(plb_calc_moments(simdata.device()));
    const chrono::duration<float> elapsed = chrono::system_clock::now() - start;
    _time_calc_moment += elapsed.count();
}

std::array<double, 4> CalcMoments::global_moments(SimData& simdata) const {
    std::array<double, 4> out{};
    plbpd::check(plb_global_moments(simdata.device(), out.data()));
    return out;
}
