// Drop-in EqlbDist (interface of the reference's EqlbDist.hpp:10-39): the
// second-order equilibrium f_k = w_k rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u),
// with the bracket evaluated in double and the result stored as float, as
// the reference's double literals make it do.
#ifndef PLBPD_AMD_HOST_EQLBDIST_HPP_
#define PLBPD_AMD_HOST_EQLBDIST_HPP_

#include <array>
#include <cstddef>
#include <vector>

#include "LBModel.hpp"

class EqlbDist final {
 public:
    explicit EqlbDist(const LBModel* lbmodel)
        : c_(lbmodel->get_directional_velocities()), w_(lbmodel->get_directional_weights()) {}
    EqlbDist(const EqlbDist&) = delete;
    EqlbDist& operator=(const EqlbDist&) = delete;
    ~EqlbDist() {}

    void operator()(const float rholocal, const std::array<float, 3>& ulocal, std::vector<float>& nlocal) {
        const float usq = ulocal[0] * ulocal[0] + ulocal[1] * ulocal[1] + ulocal[2] * ulocal[2];
        for (size_t k = 0; k < c_.size(); ++k) {
            const float cu = c_[k][0] * ulocal[0] + c_[k][1] * ulocal[1] + c_[k][2] * ulocal[2];
            const float wr = w_[k] * rholocal;
            nlocal[k] = (float)((double)wr * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq));
        }
    }

 private:
    const std::vector<std::array<int32_t, 3> >& c_;
    const std::vector<float>& w_;
};

#endif  // PLBPD_AMD_HOST_EQLBDIST_HPP_
