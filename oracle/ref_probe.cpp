// Probe of the reference's own table, equilibrium, field-index and domain
// code -- TEST INFRASTRUCTURE ONLY (built into oracle/_ref/ by
// oracle/Makefile from the sources under the reference tree, never copied).
//
// The reference's hot-path files (CollisionSRT.cpp, Streaming.cpp,
// Boundary.cpp, CalcMoments.hpp, PLBPD.cpp) include tbb/tbb.h, which this
// image does not have, so they cannot be built; LBModel.cpp, EqlbDist.hpp,
// Field.hpp and Domain.cpp need only the standard library and are exercised
// here through a C interface that tests/test_oracle_ref.py loads.
#include <array>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "Domain.hpp"
#include "EqlbDist.hpp"
#include "Field.hpp"
#include "LBModel.hpp"

extern "C" {

// LBModel tables: returns q, fills c (q*3), w (q floats), rev (q), cs2.
int ref_tables(const char* name, int32_t* c, float* w, uint32_t* rev, float* cs2) {
    try {
        LBModel m(name);
        const int q = (int)m.get_num_directions();
        const auto& cv = m.get_directional_velocities();
        const auto& wv = m.get_directional_weights();
        const auto& rv = m.get_reverse();
        for (int k = 0; k < q; ++k) {
            for (int i = 0; i < 3; ++i) c[3 * k + i] = cv[k][i];
            w[k] = wv[k];
            rev[k] = rv[k];
        }
        *cs2 = m.get_speed_of_sound_squared();
        return q;
    } catch (const std::exception&) {
        return -1;
    }
}

// EqlbDist::operator() for one (rho, u).
int ref_feq(const char* name, float rho, const float* u, float* out) {
    try {
        LBModel m(name);
        EqlbDist eq(&m);
        std::vector<float> n(m.get_num_directions(), 0.0f);
        eq(rho, std::array<float, 3>{u[0], u[1], u[2]}, n);
        for (size_t k = 0; k < n.size(); ++k) out[k] = n[k];
        return (int)n.size();
    } catch (const std::exception&) {
        return -1;
    }
}

// Field::VectorField<float, NB>::sub2ind and extents for a field of interior
// size (z, y, x) and vector length v.  ext: zbegin zend ybegin yend xbegin xend.
int ref_field(int nbuf, uint32_t z, uint32_t y, uint32_t x, uint32_t v, uint32_t qz, uint32_t qy,
              uint32_t qx, uint32_t qv, uint32_t* index, uint32_t* ext) {
    try {
        if (nbuf == 0) {
            Field::VectorField<float, 0> f(z, y, x, v, 0.0f);
            *index = f.sub2ind(qz, qy, qx, qv);
            const auto& e = f.get_extents();
            const uint32_t vals[6] = {e.zbegin, e.zend, e.ybegin, e.yend, e.xbegin, e.xend};
            std::memcpy(ext, vals, sizeof(vals));
        } else {
            Field::VectorField<float, 1> f(z, y, x, v, 0.0f);
            *index = f.sub2ind(qz, qy, qx, qv);
            const auto& e = f.get_extents();
            const uint32_t vals[6] = {e.zbegin, e.zend, e.ybegin, e.yend, e.xbegin, e.xend};
            std::memcpy(ext, vals, sizeof(vals));
        }
        return 0;
    } catch (const std::invalid_argument&) {
        return 1;  // zero extent
    }
}

// Domain getters (Domain.cpp:24-44): dims as (x, y, z).
int ref_domain(const char* name, uint32_t x, uint32_t y, uint32_t z, float nu, float rho,
               const float* u0, const float* F, uint64_t* dims_out, float* nu_out, float* rho_out) {
    try {
        LBModel m(name);
        Domain d(&m, x, y, z, nu, rho, {u0[0], u0[1], u0[2]}, {F[0], F[1], F[2]});
        size_t a, b, c;
        std::tie(a, b, c) = d.get_dimensions();
        dims_out[0] = a;
        dims_out[1] = b;
        dims_out[2] = c;
        *nu_out = d.get_fluid_viscosity();
        *rho_out = d.get_fluid_density();
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

}  // extern "C"
