// plbpd CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference lattice-Boltzmann BGK step of
// pchakraborty/plbpd, used by tests/ (and by bench.py's cpu_baseline leg) as
// the reference the HIP product is compared with.  Nothing in plbpd_amd/ links, loads or
// calls this file.  Built with -ffp-contract=off so that every expression is
// evaluated exactly in the order written below.
//
// Two step semantics are restated:
//
//  * literal: the reference's staged step on halo'd AoS arrays, in the order
//    of src/PLBPD.cpp:23-30 (collide, push+swap, no-slip, periodic, moments,
//    reset), for R = double (the FP64 contract) and R = float with the
//    reference's own mixed float/double promotions.
//  * corrected: one fused collide + push step with periodic wraps and
//    halfway bounce-back driven by a per-cell flag, on compact (halo-free)
//    AoS arrays, for every storage variant the product offers
//    (f64, f64-storage/f32-math, f32, f32 shifted, f16 shifted+scaled).
//
// Pinning (what checks each function; see tests/):
//  * tables, feq: golden vectors from src/utest/TestLBModel.cpp and
//    src/utest/TestEqlbDist.cpp, and the reference's own LBModel.cpp /
//    EqlbDist.hpp compiled into oracle/_ref (tests/test_oracle_ref.py).
//  * field index: src/utest/TestField.cpp:106-107 known answers.
//  * collide: feq fixed point, mass conservation, momentum gain exactly F.
//  * push: tagged-value bijection on tiny lattices.
//  * no-slip / periodic: tagged tiny lattices with hand-derived expected
//    destinations, and the stale-value quirks at periodic seams next to a wall
//    and at doubly periodic corner lines (hand-derived values at step 1).
//  * moments: moments of feq return (rho, u).
//  * whole step: transient Poiseuille series, Couette profile, Taylor-Green
//    decay rate, literal == corrected in a closed box, axis-permutation
//    symmetry, mass conservation in corrected mode.
//  * f16 conversion: against numpy.float16 (IEEE round-to-nearest-even).
//  * storage variants (corrected step): each within its velocity L2 bound of
//    the FP64 run; FP64 storage with FP32 arithmetic equal to FP32 storage;
//    exact stored values of the shifted forms (float(f - w), half(2^15 (f - w)));
//    shifted rest state stored as zero (tests/test_oracle_variants.py).
//  * solid voxels: a solid plane in a periodic box equals the wall-face
//    channel; mass conserved around solids; mirror symmetry of a symmetric
//    mask.
//  Parity unpinned: the Q5 race of _apply_noslip (Boundary.cpp:199-220) has no
//  single reference answer; the serial loop order restated here is one valid
//  schedule of it.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <type_traits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------- lattices --
struct Lattice {
    int q;
    int c[27][3];
    double w[27];
    int rev[27];
};

// D2Q9 in the x-z plane (c_y == 0): LBModel.cpp:23-49
const Lattice kD2Q9 = {
    9,
    {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 0, 1}, {0, 0, -1},
     {1, 0, 1}, {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}},
    {4. / 9., 1. / 9., 1. / 9., 1. / 9., 1. / 9., 1. / 36., 1. / 36., 1. / 36., 1. / 36.},
    {0, 2, 1, 4, 3, 6, 5, 8, 7}};

// D3Q19: LBModel.cpp:119-165
const Lattice kD3Q19 = {
    19,
    {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
     {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0},
     {1, 0, 1}, {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
     {0, 1, 1}, {0, -1, -1}, {0, 1, -1}, {0, -1, 1}},
    {1. / 3., 1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18.,
     1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.,
     1. / 36., 1. / 36., 1. / 36., 1. / 36.},
    {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17}};

// D3Q27: LBModel.cpp:51-117
const Lattice kD3Q27 = {
    27,
    {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
     {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0},
     {1, 0, 1}, {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
     {0, 1, 1}, {0, -1, -1}, {0, 1, -1}, {0, -1, 1},
     {1, 1, 1}, {-1, -1, -1}, {1, 1, -1}, {-1, -1, 1},
     {1, -1, 1}, {-1, 1, -1}, {-1, 1, 1}, {1, -1, -1}},
    {8. / 27., 2. / 27., 2. / 27., 2. / 27., 2. / 27., 2. / 27., 2. / 27.,
     1. / 54., 1. / 54., 1. / 54., 1. / 54., 1. / 54., 1. / 54.,
     1. / 54., 1. / 54., 1. / 54., 1. / 54., 1. / 54., 1. / 54.,
     1. / 216., 1. / 216., 1. / 216., 1. / 216., 1. / 216., 1. / 216., 1. / 216., 1. / 216.},
    {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17,
     20, 19, 22, 21, 24, 23, 26, 25}};

const Lattice* lattice(int q) {
    if (q == 9) return &kD2Q9;
    if (q == 19) return &kD3Q19;
    if (q == 27) return &kD3Q27;
    return nullptr;
}

// ------------------------------------------------------------ half floats --
// IEEE binary16 conversions written from the format definition
// (round to nearest, ties to even; subnormals kept; overflow to infinity).
uint16_t f32_to_f16(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t a = x & 0x7fffffffu;
    if (a >= 0x7f800000u) {  // inf / nan
        if (a == 0x7f800000u) return (uint16_t)(sign | 0x7c00u);
        return (uint16_t)(sign | 0x7e00u | ((a >> 13) & 0x3ffu));
    }
    if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // >= 65520 -> inf
    if (a >= 0x38800000u) {                                   // normal half
        uint32_t h = (a - 0x38000000u) >> 13;
        const uint32_t rem = a & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
        return (uint16_t)(sign | h);
    }
    // subnormal half (or zero): value / 2^-24, rounded to nearest even
    const int e = (int)(a >> 23);
    if (e < 102) return (uint16_t)sign;  // < 2^-25: rounds to zero
    const uint32_t mant = (a & 0x7fffffu) | 0x800000u;
    const int shift = 126 - e;  // >= 14
    uint32_t h = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u);
    const uint32_t halfway = 1u << (shift - 1);
    if (rem > halfway || (rem == halfway && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
}

float f16_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu;
    const uint32_t m = h & 0x3ffu;
    uint32_t x;
    if (e == 0) {
        if (m == 0) {
            x = sign;
        } else {  // subnormal: normalise
            int ee = -1;
            uint32_t mm = m;
            do {
                ++ee;
                mm <<= 1;
            } while ((mm & 0x400u) == 0);
            x = sign | ((uint32_t)(127 - 15 - ee) << 23) | ((mm & 0x3ffu) << 13);
        }
    } else if (e == 31) {
        x = sign | 0x7f800000u | (m << 13);
    } else {
        x = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

// ---------------------------------------------------------- initial state --
// Counter-based generator: splitmix64 finaliser (Steele, Lea, Flood 2014).
uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

double unit_uniform(uint64_t seed, uint64_t cell, int comp) {
    const uint64_t h = mix64(mix64(seed) + 3ull * cell + (uint64_t)comp);
    return (double)(h >> 11) * 0x1.0p-53;
}

}  // namespace

extern "C" {

// Boundary description in the reference's face order (Boundary.cpp:17-18):
// 0 east (+x), 1 west (-x), 2 north (+y), 3 south (-y), 4 up (+z), 5 down (-z).
typedef struct {
    int type[6];     // 0 unset, 1 periodic, 2 noslip
    int has_vel[6];  // velocity prescribed for the face
    double vel[6][3];
    double rho_s;    // solid density used by reset (Boundary.cpp:58-89)
} orc_bc;

// Initial velocity field: 0 uniform u0 (Domain.cpp:46-64); 1 u0 plus a seeded
// uniform perturbation amp*U(-1,1) per component per cell; 2 Taylor-Green in
// the x-z plane (walls along z): ux = u0x sin(kx) cos(kz) cos(ky),
// uz = -u0x cos(kx) sin(kz) cos(ky), uy = 0, with k = 2 pi / L per axis and
// cell-centred coordinates (i + 0.5).
typedef struct {
    int kind;
    unsigned long long seed;
    double amp;
} orc_init;

}  // extern "C"

namespace {

enum { kUnset = 0, kPeriodic = 1, kNoslip = 2 };

// initial velocity of interior cell (xi, yi, zi), 0-based
void cell_velocity(const orc_init* in, int X, int Y, int Z, const double u0[3],
                   int xi, int yi, int zi, const std::vector<double>* tg, double u[3]) {
    const int kind = in ? in->kind : 0;
    if (kind == 1) {
        const uint64_t cell = ((uint64_t)zi * Y + yi) * X + xi;
        for (int i = 0; i < 3; ++i) {
            const double r = unit_uniform(in->seed, cell, i);
            u[i] = u0[i] + in->amp * (2.0 * r - 1.0);
        }
    } else if (kind == 2) {
        // tg: sx[X] cx[X] sy[Y] cy[Y] sz[Z] cz[Z]
        const double* sx = tg->data();
        const double* cx = sx + X;
        const double* cy = cx + X + Y;
        const double* sz = cy + Y;
        const double* cz = sz + Z;
        u[0] = ((u0[0] * sx[xi]) * cz[zi]) * cy[yi];
        u[1] = 0.0;
        u[2] = ((-u0[0] * cx[xi]) * sz[zi]) * cy[yi];
    } else {
        u[0] = u0[0];
        u[1] = u0[1];
        u[2] = u0[2];
    }
}

std::vector<double> tg_tables(int X, int Y, int Z) {
    std::vector<double> t(2 * (X + Y + Z));
    double* p = t.data();
    const int L[3] = {X, Y, Z};
    for (int a = 0; a < 3; ++a) {
        const double k = 2.0 * M_PI / (double)L[a];
        for (int i = 0; i < L[a]; ++i) p[i] = std::sin(k * ((double)i + 0.5));
        for (int i = 0; i < L[a]; ++i) p[L[a] + i] = std::cos(k * ((double)i + 0.5));
        p += 2 * L[a];
    }
    return t;
}

// feq restated from EqlbDist.hpp:26-38.  For R = float the reference's
// promotions are kept: usq and cu in float, the bracket in double (double
// literals), w*rho in float, result rounded to float.
template <class R>
void feq_ref(const Lattice& L, R rho, const R u[3], R* out) {
    const R usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    for (int k = 0; k < L.q; ++k) {
        const R cu = (R)L.c[k][0] * u[0] + (R)L.c[k][1] * u[1] + (R)L.c[k][2] * u[2];
        const R wr = (R)L.w[k] * rho;
        out[k] = (R)((double)wr * (1.0 + 3.0 * (double)cu + 4.5 * (double)cu * (double)cu -
                                   1.5 * (double)usq));
    }
}

// ------------------------------------------------------------ literal mode --
template <class R>
struct Lit {
    const Lattice* L;
    int X, Y, Z;
    size_t nx, ny, nz;  // halo'd extents
    size_t cell(int z, int y, int x) const { return ((size_t)z * ny + y) * nx + x; }
    size_t ncell() const { return nx * ny * nz; }
};

// Domain::initialize (Domain.cpp:46-64) + SimData zero-filled ntmp (SimData.cpp:53-56)
template <class R>
void lit_init(const Lit<R>& g, double rho0, const double u0[3], const orc_init* in,
              R* rho, R* u, R* n, R* ntmp) {
    const int q = g.L->q;
    std::vector<double> tg;
    if (in && in->kind == 2) tg = tg_tables(g.X, g.Y, g.Z);
    for (int z = 0; z < g.Z + 2; ++z)
        for (int y = 0; y < g.Y + 2; ++y)
            for (int x = 0; x < g.X + 2; ++x) {
                const int xi = (x - 1 + g.X) % g.X, yi = (y - 1 + g.Y) % g.Y,
                          zi = (z - 1 + g.Z) % g.Z;
                double uc[3];
                cell_velocity(in, g.X, g.Y, g.Z, u0, xi, yi, zi, &tg, uc);
                const size_t c = g.cell(z, y, x);
                rho[c] = (R)rho0;
                R ur[3] = {(R)uc[0], (R)uc[1], (R)uc[2]};
                for (int i = 0; i < 3; ++i) u[3 * c + i] = ur[i];
                feq_ref<R>(*g.L, (R)rho0, ur, n + q * c);
                for (int k = 0; k < q; ++k) ntmp[q * c + k] = (R)0;
            }
}

// CollisionSRT::_collision_kernel (CollisionSRT.cpp:153-176) with
// _get_updated_u (72-81) and _get_cu (84-91), every interior cell.  In FP32
// the reference's promotions: w_k * rho in float, the bracket in double, the
// blend's first product in float and its sum in double.
template <class R>
void lit_collide(const Lit<R>& g, double tau, const double F[3], const R* rho, const R* u, R* n) {
    const int q = g.L->q;
    const Lattice& L = *g.L;
    R tau_r, om, omm;
    if (std::is_same<R, float>::value) {  // CollisionSRT.cpp:19-20
        tau_r = (R)tau;
        om = (R)(1. / (double)tau_r);
        omm = (R)1 - om;
    } else {
        tau_r = (R)tau;
        om = (R)(1.0 / tau);
        omm = (R)1 - om;
    }
    const R Fr[3] = {(R)F[0], (R)F[1], (R)F[2]};
#pragma omp parallel for schedule(static)
    for (int z = 1; z <= g.Z; ++z)
        for (int y = 1; y <= g.Y; ++y)
            for (int x = 1; x <= g.X; ++x) {
                const size_t c = g.cell(z, y, x);
                const R rl = rho[c];
                const R tri = tau_r / rl;
                R uu[3];
                for (int i = 0; i < 3; ++i) uu[i] = u[3 * c + i] + Fr[i] * tri;
                const R usq = uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2];
                R* nl = n + q * c;
                for (int k = 0; k < q; ++k) {
                    const R cu = (k == 0) ? (R)0
                                          : (R)L.c[k][0] * uu[0] + (R)L.c[k][1] * uu[1] +
                                                (R)L.c[k][2] * uu[2];
                    if (std::is_same<R, float>::value) {
                        const float wr = (float)L.w[k] * (float)rl;
                        const double neq = (double)wr * (1.0 + 3.0 * (double)cu +
                                                         4.5 * (double)cu * (double)cu -
                                                         1.5 * (double)usq);
                        const float a = (float)omm * (float)nl[k];
                        nl[k] = (R)((double)a + (double)om * neq);
                    } else {
                        const double neq =
                            ((double)L.w[k] * (double)rl) *
                            (1.0 + 3.0 * (double)cu + 4.5 * (double)cu * (double)cu -
                             1.5 * (double)usq);
                        nl[k] = (R)((double)omm * (double)nl[k] + (double)om * neq);
                    }
                }
            }
}

// Streaming::_push_kernel (Streaming.cpp:63-74): ntmp(x + c_k, k) = n(x, k)
template <class R>
void lit_push(const Lit<R>& g, const R* n, R* ntmp) {
    const int q = g.L->q;
    const Lattice& L = *g.L;
#pragma omp parallel for schedule(static)
    for (int z = 1; z <= g.Z; ++z)
        for (int y = 1; y <= g.Y; ++y)
            for (int x = 1; x <= g.X; ++x) {
                const size_t c = g.cell(z, y, x);
                for (int k = 0; k < q; ++k) {
                    const size_t d = g.cell(z + L.c[k][2], y + L.c[k][1], x + L.c[k][0]);
                    ntmp[q * d + k] = n[q * c + k];
                }
            }
}

// Boundary::_get_boundary_extent (Boundary.cpp:113-148): [lo, hi) per axis
void face_extent(int face, int X, int Y, int Z, int lo[3], int hi[3]) {
    lo[0] = 1; hi[0] = X + 1;
    lo[1] = 1; hi[1] = Y + 1;
    lo[2] = 1; hi[2] = Z + 1;
    switch (face) {
        case 0: lo[0] = X; hi[0] = X + 1; break;  // east
        case 1: lo[0] = 1; hi[0] = 2; break;      // west
        case 2: lo[1] = Y; hi[1] = Y + 1; break;  // north
        case 3: lo[1] = 1; hi[1] = 2; break;      // south
        case 4: lo[2] = Z; hi[2] = Z + 1; break;  // up
        default: lo[2] = 1; hi[2] = 2; break;     // down
    }
}

// Boundary::apply_noslip -> _apply_noslip (Boundary.cpp:101-111, 189-221),
// faces in the order east, west, north, south, up, down; the face loop is
// the serial z, y, x, kp order (one valid schedule of the TBB loop).
template <class R>
int lit_noslip(const Lit<R>& g, const orc_bc* bc, const R* rho, R* n) {
    const int q = g.L->q;
    const Lattice& L = *g.L;
    const R cs2inv = (R)1 / (R)(1.0 / 3.0);
    for (int f = 0; f < 6; ++f) {
        if (bc->type[f] != kNoslip) continue;
        if (!bc->has_vel[f]) return -3;  // std::out_of_range from _velocity.at
        const R ub[3] = {(R)bc->vel[f][0], (R)bc->vel[f][1], (R)bc->vel[f][2]};
        int lo[3], hi[3];
        face_extent(f, g.X, g.Y, g.Z, lo, hi);
        for (int z = lo[2]; z < hi[2]; ++z)
            for (int y = lo[1]; y < hi[1]; ++y)
                for (int x = lo[0]; x < hi[0]; ++x)
                    for (int kp = 1; kp < q; ++kp) {
                        const size_t nb = g.cell(z + L.c[kp][2], y + L.c[kp][1], x + L.c[kp][0]);
                        const R rhonbr = rho[nb];
                        const int k = L.rev[kp];
                        const R cu = (R)L.c[k][0] * ub[0] + (R)L.c[k][1] * ub[1] +
                                     (R)L.c[k][2] * ub[2];
                        n[q * nb + kp] =
                            n[q * g.cell(z, y, x) + k] - (R)2 * (R)L.w[k] * rhonbr * cu * cs2inv;
                    }
    }
    return 0;
}

// Boundary::apply_periodicity (Boundary.cpp:91-99, 223-281): x then y, no z.
template <class R>
void lit_periodic(const Lit<R>& g, const orc_bc* bc, R* n) {
    const int q = g.L->q;
    const Lattice& L = *g.L;
    if (bc->type[0] == kPeriodic && bc->type[1] == kPeriodic) {
        for (int z = 1; z <= g.Z; ++z)
            for (int y = 1; y <= g.Y; ++y)
                for (int k = 1; k < q; ++k) {
                    if (L.c[k][0] > 0) n[q * g.cell(z, y, 1) + k] = n[q * g.cell(z, y, g.X + 1) + k];
                    if (L.c[k][0] < 0) n[q * g.cell(z, y, g.X) + k] = n[q * g.cell(z, y, 0) + k];
                }
    }
    if (bc->type[2] == kPeriodic && bc->type[3] == kPeriodic) {
        for (int z = 1; z <= g.Z; ++z)
            for (int x = 1; x <= g.X; ++x)
                for (int k = 1; k < q; ++k) {
                    if (L.c[k][1] > 0) n[q * g.cell(z, 1, x) + k] = n[q * g.cell(z, g.Y + 1, x) + k];
                    if (L.c[k][1] < 0) n[q * g.cell(z, g.Y, x) + k] = n[q * g.cell(z, 0, x) + k];
                }
    }
}

// CalcMoments::operator() / _get_local_moments (CalcMoments.hpp:22-37, 46-73)
template <class R>
void lit_moments(const Lit<R>& g, const R* n, R* rho, R* u) {
    const int q = g.L->q;
    const Lattice& L = *g.L;
#pragma omp parallel for schedule(static)
    for (int z = 1; z <= g.Z; ++z)
        for (int y = 1; y <= g.Y; ++y)
            for (int x = 1; x <= g.X; ++x) {
                const size_t c = g.cell(z, y, x);
                R rl = (R)0;
                R ul[3] = {(R)0, (R)0, (R)0};
                for (int k = 0; k < q; ++k) {
                    const R nk = n[q * c + k];
                    rl += nk;
                    for (int i = 0; i < 3; ++i) ul[i] += nk * (R)L.c[k][i];
                }
                const R ri = (R)1 / rl;
                rho[c] = rl;
                for (int i = 0; i < 3; ++i) u[3 * c + i] = ul[i] * ri;
            }
}

// Boundary::reset = reset_velocity then reset_density (Boundary.cpp:58-89, 150-187)
template <class R>
void lit_reset(const Lit<R>& g, const orc_bc* bc, R* rho, R* u) {
    for (int f = 0; f < 6; ++f) {
        if (!bc->has_vel[f]) continue;
        int lo[3], hi[3];
        face_extent(f, g.X, g.Y, g.Z, lo, hi);
        for (int z = lo[2]; z < hi[2]; ++z)
            for (int y = lo[1]; y < hi[1]; ++y)
                for (int x = lo[0]; x < hi[0]; ++x)
                    for (int i = 0; i < 3; ++i) u[3 * g.cell(z, y, x) + i] = (R)bc->vel[f][i];
    }
    for (int f = 0; f < 6; ++f) {
        if (bc->type[f] != kNoslip) continue;
        int lo[3], hi[3];
        face_extent(f, g.X, g.Y, g.Z, lo, hi);
        for (int z = lo[2]; z < hi[2]; ++z)
            for (int y = lo[1]; y < hi[1]; ++y)
                for (int x = lo[0]; x < hi[0]; ++x) rho[g.cell(z, y, x)] = (R)bc->rho_s;
    }
}

// PLBPD::march_in_time (PLBPD.cpp:21-33).  bufs[0] holds n on entry; the
// return value says which buffer holds n on exit (the std::swap of
// Streaming.cpp:93 alternates them).
template <class R>
int lit_run(const Lit<R>& g, double tau, const double F[3], const orc_bc* bc, int nsteps,
            R* rho, R* u, R* bufA, R* bufB) {
    R* n = bufA;
    R* t = bufB;
    for (int s = 0; s < nsteps; ++s) {
        lit_collide<R>(g, tau, F, rho, u, n);
        lit_push<R>(g, n, t);
        R* tmp = n;
        n = t;
        t = tmp;
        const int rc = lit_noslip<R>(g, bc, rho, n);
        if (rc) return rc;
        lit_periodic<R>(g, bc, n);
        lit_moments<R>(g, n, rho, u);
        lit_reset<R>(g, bc, rho, u);
    }
    return n == bufA ? 0 : 1;
}

template <class R>
bool make_lit(int q, int X, int Y, int Z, Lit<R>* g) {
    g->L = lattice(q);
    if (!g->L || X < 1 || Y < 1 || Z < 1) return false;
    g->X = X;
    g->Y = Y;
    g->Z = Z;
    g->nx = X + 2;
    g->ny = Y + 2;
    g->nz = Z + 2;
    return true;
}

// ---------------------------------------------------------- corrected mode --
// Storage variants.  T is the collision arithmetic type, S the stored type.
// Shifted variants store g = f - w_k (rho0 = 1); the f16 variant stores
// half(g * 2^15) (Lehmann, Krause, Amati, Sega, Harting, Gekle,
// Phys. Rev. E 106, 015308 (2022): DDF shifting plus scaled FP16).
struct VF64 {
    using S = double;
    using T = double;
    static constexpr bool shifted = false;
    static T load(S s) { return s; }
    static S store(T t) { return t; }
    static S from_double(double v) { return v; }
};
struct VF64A32 {
    using S = double;
    using T = float;
    static constexpr bool shifted = false;
    static T load(S s) { return (float)s; }
    static S store(T t) { return (double)t; }
    static S from_double(double v) { return v; }
};
struct VF32 {
    using S = float;
    using T = float;
    static constexpr bool shifted = false;
    static T load(S s) { return s; }
    static S store(T t) { return t; }
    static S from_double(double v) { return (float)v; }
};
struct VF32S {
    using S = float;
    using T = float;
    static constexpr bool shifted = true;
    static T load(S s) { return s; }
    static S store(T t) { return t; }
    static S from_double(double v) { return (float)v; }
};
struct VF16S {
    using S = uint16_t;
    using T = float;
    static constexpr bool shifted = true;
    static T load(S s) { return f16_to_f32(s) * 0x1.0p-15f; }
    static S store(T t) { return f32_to_f16(t * 0x1.0p15f); }
    static S from_double(double v) { return f32_to_f16((float)v * 0x1.0p15f); }
};

struct Cor {
    const Lattice* L;
    int X, Y, Z;
    size_t cell(int z, int y, int x) const { return ((size_t)z * Y + y) * X + x; }
    size_t ncell() const { return (size_t)X * Y * Z; }
};

int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Per-cell flag: 0 fluid, 1 + face for a no-slip face layer (the last face in
// reference order wins, as in _apply_noslip and reset), 7 interior solid.
int cor_flags(const Cor& g, const orc_bc* bc, const uint8_t* solid, uint8_t* flags) {
    const Lattice& L = *g.L;
    const int dims[3] = {g.X, g.Y, g.Z};
    for (int a = 0; a < 3; ++a) {
        bool moves = false;
        for (int k = 0; k < L.q; ++k) moves |= L.c[k][a] != 0;
        if (!moves) continue;
        const int hi = bc->type[2 * a], lo = bc->type[2 * a + 1];
        if (!((hi == kPeriodic && lo == kPeriodic) || (hi == kNoslip && lo == kNoslip)))
            return -2;
        if (hi == kNoslip && dims[a] < 3) return -2;
    }
    for (int f = 0; f < 6; ++f)
        if (bc->type[f] == kNoslip && !bc->has_vel[f]) return -3;
    for (int z = 0; z < g.Z; ++z)
        for (int y = 0; y < g.Y; ++y)
            for (int x = 0; x < g.X; ++x) {
                const size_t c = g.cell(z, y, x);
                int f8 = (solid && solid[c]) ? 7 : 0;
                const bool on[6] = {x == g.X - 1, x == 0, y == g.Y - 1, y == 0, z == g.Z - 1, z == 0};
                for (int f = 0; f < 6; ++f)
                    if (bc->type[f] == kNoslip && on[f]) f8 = 1 + f;
                flags[c] = (uint8_t)f8;
            }
    return 0;
}

template <class V>
void cor_init(const Cor& g, double rho0, const double u0[3], const orc_init* in, typename V::S* f) {
    const Lattice& L = *g.L;
    const int q = L.q;
    std::vector<double> tg;
    if (in && in->kind == 2) tg = tg_tables(g.X, g.Y, g.Z);
    for (int z = 0; z < g.Z; ++z)
        for (int y = 0; y < g.Y; ++y)
            for (int x = 0; x < g.X; ++x) {
                double uc[3], fe[27];
                cell_velocity(in, g.X, g.Y, g.Z, u0, x, y, z, &tg, uc);
                feq_ref<double>(L, rho0, uc, fe);
                const size_t c = g.cell(z, y, x);
                for (int k = 0; k < q; ++k)
                    f[q * c + k] = V::from_double(V::shifted ? fe[k] - L.w[k] : fe[k]);
            }
}

// Local moments in arithmetic T (the order of CalcMoments.hpp:27-35, k
// ascending); shifted variants sum deviations and add 1 afterwards.  The
// corrected step sums only the nonzero components of c_k: a product with a
// zero component is an exact +-0 that changes at most the sign of a zero sum.
template <class V>
void cor_local_moments(const Lattice& L, const typename V::T* f, typename V::T* rho,
                       typename V::T* drho, typename V::T u[3]) {
    using T = typename V::T;
    T r = (T)0;
    T j[3] = {(T)0, (T)0, (T)0};
    for (int k = 0; k < L.q; ++k) {
        r += f[k];
        for (int i = 0; i < 3; ++i)
            if (L.c[k][i] != 0) j[i] += f[k] * (T)L.c[k][i];
    }
    *drho = r;
    if (V::shifted) r = (T)1 + r;
    *rho = r;
    const T ri = (T)1 / r;
    for (int i = 0; i < 3; ++i) u[i] = j[i] * ri;
}

struct CorParams {
    double tau;
    double F[3];
    double vel[8][3];  // by flag value
    bool moving;       // some wall face has a nonzero velocity
};

template <class V>
void cor_step(const Cor& g, const CorParams& p, const uint8_t* flags, const typename V::S* A,
              typename V::S* B) {
    using T = typename V::T;
    const Lattice& L = *g.L;
    const int q = L.q;
    const T tau = (T)p.tau;
    T om, omm;
    if (std::is_same<T, float>::value) {
        om = (T)(1. / (double)tau);
    } else {
        om = (T)(1.0 / p.tau);
    }
    omm = (T)1 - om;
    const T F[3] = {(T)p.F[0], (T)p.F[1], (T)p.F[2]};
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g.Z; ++z)
        for (int y = 0; y < g.Y; ++y)
            for (int x = 0; x < g.X; ++x) {
                const size_t c = g.cell(z, y, x);
                if (flags[c]) continue;
                T f[27];
                for (int k = 0; k < q; ++k) f[k] = V::load(A[q * c + k]);
                T rho, drho, u[3];
                cor_local_moments<V>(L, f, &rho, &drho, u);
                const T tri = tau / rho;
                T uu[3];
                for (int i = 0; i < 3; ++i) uu[i] = u[i] + F[i] * tri;
                const T usq = uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2];
                for (int k = 0; k < q; ++k) {
                    // c_k . u over the nonzero components, x, y, z order
                    T cu = (T)0;
                    bool any = false;
                    for (int i = 0; i < 3; ++i)
                        if (L.c[k][i] != 0) {
                            const T t = (T)L.c[k][i] * uu[i];
                            cu = any ? cu + t : t;
                            any = true;
                        }
                    const T w = (T)L.w[k];
                    T fs;
                    if (V::shifted) {
                        const T geq = w * (drho + rho * ((((T)3 * cu) + ((T)4.5 * cu) * cu) -
                                                         (T)1.5 * usq));
                        fs = omm * f[k] + om * geq;
                    } else {
                        const T neq = (w * rho) * ((((T)1 + (T)3 * cu) + ((T)4.5 * cu) * cu) -
                                                   (T)1.5 * usq);
                        fs = omm * f[k] + om * neq;
                    }
                    const int xd = wrap(x + L.c[k][0], g.X), yd = wrap(y + L.c[k][1], g.Y),
                              zd = wrap(z + L.c[k][2], g.Z);
                    const size_t d = g.cell(zd, yd, xd);
                    const int fl = flags[d];
                    if (fl) {
                        // halfway bounce-back; the moving-wall term of
                        // Boundary.cpp:213-217 where a wall face (flags 1-6,
                        // not an interior solid, 7) moves
                        T v = fs;
                        if (p.moving && fl <= 6) {
                            const double* ubd = p.vel[fl];
                            const T ub[3] = {(T)ubd[0], (T)ubd[1], (T)ubd[2]};
                            const T cub = (T)L.c[k][0] * ub[0] + (T)L.c[k][1] * ub[1] +
                                          (T)L.c[k][2] * ub[2];
                            const T term = (T)2 * w * rho * cub * (T)3;
                            v = fs - term;
                        }
                        B[q * c + L.rev[k]] = V::store(v);
                    } else {
                        B[q * d + k] = V::store(fs);
                    }
                }
            }
}

template <class V>
void cor_moments(const Cor& g, const CorParams& p, const double rho_s, const uint8_t* flags,
                 const typename V::S* f, double* rho, double* u) {
    using T = typename V::T;
    const Lattice& L = *g.L;
    const int q = L.q;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g.Z; ++z)
        for (int y = 0; y < g.Y; ++y)
            for (int x = 0; x < g.X; ++x) {
                const size_t c = g.cell(z, y, x);
                if (flags[c]) {
                    rho[c] = rho_s;
                    for (int i = 0; i < 3; ++i) u[3 * c + i] = p.vel[flags[c]][i];
                    continue;
                }
                T fl[27];
                for (int k = 0; k < q; ++k) fl[k] = V::load(f[q * c + k]);
                T r, dr, uu[3];
                cor_local_moments<V>(L, fl, &r, &dr, uu);
                rho[c] = (double)r;
                for (int i = 0; i < 3; ++i) u[3 * c + i] = (double)uu[i];
            }
}

template <class V>
void cor_populations(const Cor& g, const typename V::S* f, double* out) {
    const Lattice& L = *g.L;
    const size_t n = g.ncell() * L.q;
    for (size_t i = 0; i < n; ++i) {
        const double v = (double)V::load(f[i]);
        out[i] = V::shifted ? v + L.w[i % L.q] : v;
    }
}

bool make_cor(int q, int X, int Y, int Z, Cor* g) {
    g->L = lattice(q);
    if (!g->L || X < 1 || Y < 1 || Z < 1) return false;
    g->X = X;
    g->Y = Y;
    g->Z = Z;
    return true;
}

void cor_params(const orc_bc* bc, double tau, const double F[3], CorParams* p) {
    std::memset(p, 0, sizeof(*p));
    p->tau = tau;
    for (int i = 0; i < 3; ++i) p->F[i] = F[i];
    p->moving = false;
    for (int f = 0; f < 6; ++f)
        for (int i = 0; i < 3; ++i) {
            p->vel[1 + f][i] = bc->has_vel[f] ? bc->vel[f][i] : 0.0;
            if (bc->type[f] == kNoslip && p->vel[1 + f][i] != 0.0) p->moving = true;
        }
}

template <class Fn>
int dispatch_variant(int variant, Fn&& fn) {
    switch (variant) {
        case 0: fn(VF64()); return 0;
        case 1: fn(VF32()); return 0;
        case 2: fn(VF64A32()); return 0;
        case 3: fn(VF32S()); return 0;
        case 4: fn(VF16S()); return 0;
        default: return -1;
    }
}

size_t storage_bytes(int variant) {
    switch (variant) {
        case 0: case 2: return 8;
        case 1: case 3: return 4;
        case 4: return 2;
        default: return 0;
    }
}

}  // namespace

// =================================================================== C ABI ==
extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// Lattice tables: c (q*3 ints), w (q doubles), w as float (q), rev (q ints).
int orc_tables(int q, int* c, double* w, float* wf, int* rev) {
    const Lattice* L = lattice(q);
    if (!L) return -1;
    for (int k = 0; k < q; ++k) {
        for (int i = 0; i < 3; ++i) c[3 * k + i] = L->c[k][i];
        w[k] = L->w[k];
        wf[k] = (float)L->w[k];
        rev[k] = L->rev[k];
    }
    return 0;
}

int orc_feq_f64(int q, double rho, const double* u, double* out) {
    const Lattice* L = lattice(q);
    if (!L) return -1;
    feq_ref<double>(*L, rho, u, out);
    return 0;
}

int orc_feq_f32(int q, float rho, const float* u, float* out) {
    const Lattice* L = lattice(q);
    if (!L) return -1;
    feq_ref<float>(*L, rho, u, out);
    return 0;
}

// Field::VectorField::sub2ind (Field.hpp:91-93) for an nbuf-layer field of
// interior size (zlen, ylen, xlen) and vector length vlen.
unsigned orc_sub2ind(unsigned zlen, unsigned ylen, unsigned xlen, unsigned vlen, unsigned nbuf,
                     unsigned z, unsigned y, unsigned x, unsigned v) {
    const unsigned Y = ylen + 2 * nbuf, X = xlen + 2 * nbuf;
    (void)zlen;
    return z * (vlen * X * Y) + y * (vlen * X) + x * vlen + v;
}

uint16_t orc_f32_to_f16(float f) { return f32_to_f16(f); }
float orc_f16_to_f32(uint16_t h) { return f16_to_f32(h); }

double orc_unit_uniform(unsigned long long seed, unsigned long long cell, int comp) {
    return unit_uniform(seed, cell, comp);
}

// ---- literal mode, halo'd AoS arrays ((Z+2)(Y+2)(X+2) cells) ----
#define ORC_LIT_API(SUF, R)                                                                     \
    int orc_lit_init_##SUF(int q, int X, int Y, int Z, double rho0, const double* u0,         \
                           const orc_init* in, const orc_bc* bc, R* rho, R* u, R* n, R* ntmp) { \
        Lit<R> g;                                                                             \
        if (!make_lit<R>(q, X, Y, Z, &g)) return -1;                                          \
        lit_init<R>(g, rho0, u0, in, rho, u, n, ntmp);                                        \
        if (bc) lit_reset<R>(g, bc, rho, u);                                                  \
        return 0;                                                                             \
    }                                                                                         \
    int orc_lit_collide_##SUF(int q, int X, int Y, int Z, double tau, const double* F,        \
                              const R* rho, const R* u, R* n) {                               \
        Lit<R> g;                                                                             \
        if (!make_lit<R>(q, X, Y, Z, &g)) return -1;                                          \
        lit_collide<R>(g, tau, F, rho, u, n);                                                 \
        return 0;                                                                             \
    }                                                                                         \
    int orc_lit_push_##SUF(int q, int X, int Y, int Z, const R* n, R* ntmp) {                 \
        Lit<R> g;                                                                             \
        if (!make_lit<R>(q, X, Y, Z, &g)) return -1;                                          \
        lit_push<R>(g, n, ntmp);                                                              \
        return 0;                                                                             \
    }                                                                                         \
    int orc_lit_noslip_##SUF(int q, int X, int Y, int Z, const orc_bc* bc, const R* rho, R* n) { \
        Lit<R> g;                                                                             \
        if (!make_lit<R>(q, X, Y, Z, &g)) return -1;                                          \
        return lit_noslip<R>(g, bc, rho, n);                                                  \
    }                                                                                         \
    int orc_lit_periodic_##SUF(int q, int X, int Y, int Z, const orc_bc* bc, R* n) {          \
        Lit<R> g;                                                                             \
        if (!make_lit<R>(q, X, Y, Z, &g)) return -1;                                          \
        lit_periodic<R>(g, bc, n);                                                            \
        return 0;                                                                             \
    }                                                                                         \
    int orc_lit_moments_##SUF(int q, int X, int Y, int Z, const R* n, R* rho, R* u) {         \
        Lit<R> g;                                                                             \
        if (!make_lit<R>(q, X, Y, Z, &g)) return -1;                                          \
        lit_moments<R>(g, n, rho, u);                                                         \
        return 0;                                                                             \
    }                                                                                         \
    int orc_lit_reset_##SUF(int q, int X, int Y, int Z, const orc_bc* bc, R* rho, R* u) {     \
        Lit<R> g;                                                                             \
        if (!make_lit<R>(q, X, Y, Z, &g)) return -1;                                          \
        lit_reset<R>(g, bc, rho, u);                                                          \
        return 0;                                                                             \
    }                                                                                         \
    int orc_lit_run_##SUF(int q, int X, int Y, int Z, double tau, const double* F,            \
                          const orc_bc* bc, int nsteps, R* rho, R* u, R* bufA, R* bufB) {     \
        Lit<R> g;                                                                             \
        if (!make_lit<R>(q, X, Y, Z, &g)) return -1;                                          \
        return lit_run<R>(g, tau, F, bc, nsteps, rho, u, bufA, bufB);                         \
    }

ORC_LIT_API(f64, double)
ORC_LIT_API(f32, float)

// ---- corrected mode, compact AoS arrays (Z*Y*X cells), storage by variant ----
// variant: 0 f64, 1 f32, 2 f64 storage with f32 arithmetic, 3 f32 shifted,
// 4 f16 shifted (stored as uint16 bit patterns).
size_t orc_storage_bytes(int variant) { return storage_bytes(variant); }

int orc_cor_flags(int q, int X, int Y, int Z, const orc_bc* bc, const uint8_t* solid,
                  uint8_t* flags) {
    Cor g;
    if (!make_cor(q, X, Y, Z, &g)) return -1;
    return cor_flags(g, bc, solid, flags);
}

int orc_cor_init(int variant, int q, int X, int Y, int Z, double rho0, const double* u0,
                 const orc_init* in, void* f) {
    Cor g;
    if (!make_cor(q, X, Y, Z, &g)) return -1;
    return dispatch_variant(variant, [&](auto v) {
        using V = decltype(v);
        cor_init<V>(g, rho0, u0, in, (typename V::S*)f);
    });
}

// Runs nsteps fused steps alternating bufA -> bufB -> bufA ...; returns the
// index of the buffer holding the current populations, or < 0 on error.
int orc_cor_run(int variant, int q, int X, int Y, int Z, double tau, const double* F,
                const orc_bc* bc, const uint8_t* solid, int nsteps, void* bufA, void* bufB) {
    Cor g;
    if (!make_cor(q, X, Y, Z, &g)) return -1;
    std::vector<uint8_t> flags(g.ncell());
    const int rc = cor_flags(g, bc, solid, flags.data());
    if (rc) return rc;
    CorParams p;
    cor_params(bc, tau, F, &p);
    int cur = 0;
    const int r2 = dispatch_variant(variant, [&](auto v) {
        using V = decltype(v);
        using S = typename V::S;
        S* bufs[2] = {(S*)bufA, (S*)bufB};
        for (int s = 0; s < nsteps; ++s) {
            cor_step<V>(g, p, flags.data(), bufs[cur], bufs[cur ^ 1]);
            cur ^= 1;
        }
    });
    return r2 ? r2 : cur;
}

int orc_cor_moments(int variant, int q, int X, int Y, int Z, const orc_bc* bc,
                    const uint8_t* solid, const void* f, double* rho, double* u) {
    Cor g;
    if (!make_cor(q, X, Y, Z, &g)) return -1;
    std::vector<uint8_t> flags(g.ncell());
    const int rc = cor_flags(g, bc, solid, flags.data());
    if (rc) return rc;
    CorParams p;
    const double F0[3] = {0, 0, 0};
    cor_params(bc, 1.0, F0, &p);
    return dispatch_variant(variant, [&](auto v) {
        using V = decltype(v);
        cor_moments<V>(g, p, bc->rho_s, flags.data(), (const typename V::S*)f, rho, u);
    });
}

// Populations as plain f in double (shift added back).
int orc_cor_populations(int variant, int q, int X, int Y, int Z, const void* f, double* out) {
    Cor g;
    if (!make_cor(q, X, Y, Z, &g)) return -1;
    return dispatch_variant(variant, [&](auto v) {
        using V = decltype(v);
        cor_populations<V>(g, (const typename V::S*)f, out);
    });
}

}  // extern "C"
