// Literal mode: the reference's staged step (PLBPD.cpp:23-30), one kernel per
// reference operator, on halo'd structure-of-arrays buffers:
//   n[k * ncell + c],  c = (z*(Y+2) + y)*(X+2) + x,  x in [0, X+1] etc.
// A collision and the push after it can run as one kernel (k_collide_push),
// which can also take the previous step's moments and resets
// (k_moments_collide_push); every buffer ends as after the separate kernels.
// R = double (FP64 contract) or float (the reference's own rounding, with its
// float/double promotions).  Compiled with -ffp-contract=off: every
// expression below is evaluated in the order written, as in the reference.
#include <mutex>
#include <string>
#include <type_traits>

#include "state.hpp"

namespace plb {
namespace {

constexpr unsigned kThreads = 256;

struct LitDims {
    int X, Y, Z;
    int nx, ny;      // X+2, Y+2
    uint64_t ncell;  // (X+2)(Y+2)(Z+2)
    __host__ __device__ uint64_t cell(int z, int y, int x) const {
        return ((uint64_t)z * ny + y) * nx + x;
    }
};

LitDims dims_of(const plb_state* s) {
    LitDims d;
    d.X = s->X;
    d.Y = s->Y;
    d.Z = s->Z;
    d.nx = s->X + 2;
    d.ny = s->Y + 2;
    d.ncell = s->ncell;
    return d;
}

template <class Fn>
int with_q(int q, Fn&& fn) {
    switch (q) {
        case 9: return fn(std::integral_constant<int, 9>());
        case 19: return fn(std::integral_constant<int, 19>());
        case 27: return fn(std::integral_constant<int, 27>());
        default: return fail(PLB_ERR_LOGIC, "Model with q=" + std::to_string(q) + " not yet implemented");
    }
}

template <class Fn>
int with_qr(const plb_state* s, Fn&& fn) {
    return with_q(s->q, [&](auto qc) {
        if (s->variant == kF64) return fn(qc, double());
        return fn(qc, float());
    });
}

// ---------------------------------------------------------------- kernels --
// Domain::initialize (Domain.cpp:46-64): every cell, halo included, gets
// rho0, u and feq(rho0, u); ntmp is zero (SimData.cpp:53-56).
template <int Q, class R>
__global__ void k_init(LitDims g, R* n, R* ntmp, R* rho, R* u, double rho0, double u0x,
                       double u0y, double u0z, int kind, uint64_t seed, double amp,
                       const double* tg) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.ncell) return;
    const int x = (int)(c % g.nx);
    const int y = (int)((c / g.nx) % g.ny);
    const int z = (int)(c / ((uint64_t)g.nx * g.ny));
    const int xi = (x - 1 + g.X) % g.X, yi = (y - 1 + g.Y) % g.Y, zi = (z - 1 + g.Z) % g.Z;
    const double u0[3] = {u0x, u0y, u0z};
    double uc[3];
    cell_velocity(kind, seed, amp, g.X, g.Y, g.Z, u0, tg, xi, yi, zi, uc);
    const R ur[3] = {(R)uc[0], (R)uc[1], (R)uc[2]};
    const R r0 = (R)rho0;
    rho[c] = r0;
    for (int i = 0; i < 3; ++i) u[i * g.ncell + c] = ur[i];
    const R usq = ur[0] * ur[0] + ur[1] * ur[1] + ur[2] * ur[2];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        n[k * g.ncell + c] = feq_ref<Q, R>(k, r0, ur, usq);
        ntmp[k * g.ncell + c] = (R)0;
    }
}

// CollisionSRT::_collision_kernel (CollisionSRT.cpp:153-176) for population
// k of one cell, with _get_cu (84-91): the post-collision value of old.  In
// FP32 the reference's promotions are kept: w_k * rho in float, the bracket in
// double (double literals), the blend's first product in float, its sum in
// double, the result rounded to float.
template <int Q, class R>
__device__ inline R collide_value(int k, R old, R rl, const R (&uu)[3], R usq, R om, R omm) {
    const R cu = (k == 0) ? (R)0
                          : (R)Lat<Q>::cx(k) * uu[0] + (R)Lat<Q>::cy(k) * uu[1] + (R)Lat<Q>::cz(k) * uu[2];
    if (std::is_same<R, float>::value) {
        const float wr = (float)Lat<Q>::w(k) * (float)rl;
        const double neq =
            (double)wr * (1.0 + 3.0 * (double)cu + 4.5 * (double)cu * (double)cu - 1.5 * (double)usq);
        const float a = (float)omm * (float)old;
        return (R)((double)a + (double)om * neq);
    } else {
        const double neq =
            (Lat<Q>::w(k) * (double)rl) * (1.0 + 3.0 * (double)cu + 4.5 * (double)cu * (double)cu - 1.5 * (double)usq);
        return (R)((double)omm * (double)old + (double)om * neq);
    }
}

// _get_updated_u (CollisionSRT.cpp:72-81): u* = u + tau F / rho, and |u*|^2.
template <class R>
__device__ inline void updated_u(R u0, R u1, R u2, R rl, R tau, R F0, R F1, R F2, R (&uu)[3], R& usq) {
    const R tri = tau / rl;
    uu[0] = u0 + F0 * tri;
    uu[1] = u1 + F1 * tri;
    uu[2] = u2 + F2 * tri;
    usq = uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2];
}

// CollisionSRT::operator() (CollisionSRT.cpp:25-176): in place, every interior cell.
template <int Q, class R>
__global__ void k_collide(LitDims g, R* __restrict__ n, const R* __restrict__ rho,
                          const R* __restrict__ u, R tau, R om, R omm, R F0, R F1, R F2) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int y = blockIdx.y + 1, z = blockIdx.z + 1;
    if (x > g.X) return;
    const uint64_t c = g.cell(z, y, x);
    const R rl = rho[c];
    R uu[3], usq;
    updated_u(u[c], u[g.ncell + c], u[2 * g.ncell + c], rl, tau, F0, F1, F2, uu, usq);
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        R* p = n + k * g.ncell + c;
        *p = collide_value<Q, R>(k, *p, rl, uu, usq, om, omm);
    }
}

// Whether interior cell (x, y, z) lies on the outer interior layer (a face
// layer of Boundary.cpp:113-148 on any axis).  The collide+push kernels store
// their in-place post-collision values only there.  Off that layer, the
// value the reference leaves in place (it becomes ntmp after the swap) is
// the one the same kernel pushed to c + c_k, and nothing reads it before the
// next push overwrites it: the next push writes ntmp(c', k) for every c' with
// c' - c_k interior, which is every cell off the outer layer.  Between the
// two pushes only a read of ntmp or a change of n at those pushed slots
// (a collision in place, an upload) needs the values, and k_ntmp_restore
// copies them back from n first.  No-slip and periodic copies never change
// a slot n(c + c_k, k) of such a cell: no-slip writes n(f + c_kp, kp) for face
// cells f, periodic copies write n(c', k) with c' - c_k in the halo.
__device__ inline bool outer_layer(const LitDims& g, int x, int y, int z) {
    return x == 1 || x == g.X || y == 1 || y == g.Y || z == 1 || z == g.Z;
}

// ntmp(c, k) = n(c + c_k, k) for the interior cells off the outer layer: the
// in-place values the collide+push kernels did not store (see outer_layer).
template <int Q, class R>
__global__ void k_ntmp_restore(LitDims g, const R* __restrict__ n, R* __restrict__ t) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int y = blockIdx.y + 1, z = blockIdx.z + 1;
    if (x > g.X || outer_layer(g, x, y, z)) return;
    const uint64_t c = g.cell(z, y, x);
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const int64_t off = (int64_t)Lat<Q>::cx(k) + (int64_t)Lat<Q>::cy(k) * g.nx +
                            (int64_t)Lat<Q>::cz(k) * g.nx * g.ny;
        t[k * g.ncell + c] = n[k * g.ncell + (uint64_t)((int64_t)c + off)];
    }
}

// The collision followed by Streaming::_push_kernel (Streaming.cpp:63-74) in
// one pass: each post-collision value is stored at x + c_k in t, as k_push
// copies it, and in place on the outer layer, as k_collide leaves it (it
// becomes ntmp after the swap; the next push leaves those entries, and the
// reference's stage order reads some of them back, Q1).  Every value a caller
// or a later stage can read equals the one k_collide then k_push give (see
// outer_layer); the push's re-read of n and the in-place stores off the outer
// layer are saved.
template <int Q, class R>
__global__ void k_collide_push(LitDims g, R* __restrict__ n, const R* __restrict__ rho,
                               const R* __restrict__ u, R* __restrict__ t, R tau, R om, R omm, R F0, R F1,
                               R F2) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int y = blockIdx.y + 1, z = blockIdx.z + 1;
    if (x > g.X) return;
    const uint64_t c = g.cell(z, y, x);
    const R rl = rho[c];
    const bool keep = outer_layer(g, x, y, z);
    R uu[3], usq;
    updated_u(u[c], u[g.ncell + c], u[2 * g.ncell + c], rl, tau, F0, F1, F2, uu, usq);
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const R v = collide_value<Q, R>(k, n[k * g.ncell + c], rl, uu, usq, om, omm);
        if (keep) n[k * g.ncell + c] = v;
        const int64_t off = (int64_t)Lat<Q>::cx(k) + (int64_t)Lat<Q>::cy(k) * g.nx +
                            (int64_t)Lat<Q>::cz(k) * g.nx * g.ny;
        t[k * g.ncell + (uint64_t)((int64_t)c + off)] = v;
    }
}

// Streaming::_push_kernel (Streaming.cpp:63-74): ntmp(x + c_k, k) = n(x, k).
template <int Q, class R>
__global__ void k_push(LitDims g, const R* __restrict__ n, R* __restrict__ t) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int y = blockIdx.y + 1, z = blockIdx.z + 1;
    if (x > g.X) return;
    const uint64_t c = g.cell(z, y, x);
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const int64_t off = (int64_t)Lat<Q>::cx(k) + (int64_t)Lat<Q>::cy(k) * g.nx +
                            (int64_t)Lat<Q>::cz(k) * g.nx * g.ny;
        t[k * g.ncell + (uint64_t)((int64_t)c + off)] = n[k * g.ncell + c];
    }
}

struct FaceBox {
    int lo[3], ex[3];  // first cell and extent per axis (halo'd coordinates)
    __host__ __device__ uint64_t count() const { return (uint64_t)ex[0] * ex[1] * ex[2]; }
};

// Boundary::_get_boundary_extent (Boundary.cpp:113-148)
FaceBox face_box(int face, int X, int Y, int Z) {
    FaceBox b;
    b.lo[0] = 1; b.ex[0] = X;
    b.lo[1] = 1; b.ex[1] = Y;
    b.lo[2] = 1; b.ex[2] = Z;
    const int axis = face / 2;
    const int hi = (face % 2) == 0;  // east, north, up are the high faces
    const int dim = axis == 0 ? X : (axis == 1 ? Y : Z);
    b.lo[axis] = hi ? dim : 1;
    b.ex[axis] = 1;
    return b;
}

// Boundary::_apply_noslip (Boundary.cpp:189-221) for one face, as the serial
// z, y, x, kp loop computes it.  A destination (y, kp) of face cell c is
// written once; its value reads n(c, k = rev kp).  When y is itself in the
// face and comes before c in the loop order, op (y, rev kp) has already
// replaced n(c, k) by n(y, kp) - t(c): the value is then
// (n(y, kp) - t(c)) - t(y).  Otherwise n(c, k) is still the value from before
// the face pass.  Every read below is of the values from before the pass, so
// the gather kernel writes to a staging buffer and a second kernel scatters.
template <int Q, class R>
__global__ void k_noslip_gather(LitDims g, FaceBox b, const R* __restrict__ n,
                                const R* __restrict__ rho, R ub0, R ub1, R ub2, R cs2inv,
                                R* __restrict__ stage, uint64_t total) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint64_t fi = t / (Q - 1);
    const int kp = 1 + (int)(t % (Q - 1));
    const int x = b.lo[0] + (int)(fi % b.ex[0]);
    const int y = b.lo[1] + (int)((fi / b.ex[0]) % b.ex[1]);
    const int z = b.lo[2] + (int)(fi / ((uint64_t)b.ex[0] * b.ex[1]));
    int cxp = 0, cyp = 0, czp = 0, k = 0;
#pragma unroll
    for (int j = 1; j < Q; ++j)
        if (j == kp) {
            cxp = Lat<Q>::cx(j);
            cyp = Lat<Q>::cy(j);
            czp = Lat<Q>::cz(j);
            k = rev<Q>(j);
        }
    const int yx = x + cxp, yy = y + cyp, yz = z + czp;
    const uint64_t cc = g.cell(z, y, x), cy = g.cell(yz, yy, yx);
    const R wk = (R)Lat<Q>::w(k);  // w_k == w_kp
    // c_k = -c_kp
    const R cu_k = (R)(-cxp) * ub0 + (R)(-cyp) * ub1 + (R)(-czp) * ub2;
    const R t_y = (R)2 * wk * rho[cy] * cu_k * cs2inv;
    const bool in_face = yx >= b.lo[0] && yx < b.lo[0] + b.ex[0] && yy >= b.lo[1] &&
                         yy < b.lo[1] + b.ex[1] && yz >= b.lo[2] && yz < b.lo[2] + b.ex[2];
    R v;
    if (in_face && cy < cc) {
        const R cu_kp = (R)cxp * ub0 + (R)cyp * ub1 + (R)czp * ub2;
        const R t_c = (R)2 * wk * rho[cc] * cu_kp * cs2inv;
        v = (n[kp * g.ncell + cy] - t_c) - t_y;
    } else {
        v = n[k * g.ncell + cc] - t_y;
    }
    stage[t] = v;
}

template <int Q, class R>
__global__ void k_noslip_scatter(LitDims g, FaceBox b, R* __restrict__ n,
                                 const R* __restrict__ stage, uint64_t total) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint64_t fi = t / (Q - 1);
    const int kp = 1 + (int)(t % (Q - 1));
    const int x = b.lo[0] + (int)(fi % b.ex[0]);
    const int y = b.lo[1] + (int)((fi / b.ex[0]) % b.ex[1]);
    const int z = b.lo[2] + (int)(fi / ((uint64_t)b.ex[0] * b.ex[1]));
    int cxp = 0, cyp = 0, czp = 0;
#pragma unroll
    for (int j = 1; j < Q; ++j)
        if (j == kp) {
            cxp = Lat<Q>::cx(j);
            cyp = Lat<Q>::cy(j);
            czp = Lat<Q>::cz(j);
        }
    n[kp * g.ncell + g.cell(z + czp, y + cyp, x + cxp)] = stage[t];
}

// Boundary::_apply_periodicity_east_west (Boundary.cpp:234-251): for
// z in 1..Z, y in 1..Y: n(1, k) = n(X+1, k) for c_kx > 0, n(X, k) = n(0, k)
// for c_kx < 0.  axis 1 is _north_south (264-281) with x in 1..X.
template <int Q, class R>
__global__ void k_periodic(LitDims g, int axis, R* __restrict__ n) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int other = axis == 0 ? g.Y : g.X;
    const uint64_t total = (uint64_t)g.Z * other * (Q - 1);
    if (t >= total) return;
    const int k = 1 + (int)(t % (Q - 1));
    const int a = 1 + (int)((t / (Q - 1)) % other);
    const int z = 1 + (int)(t / ((uint64_t)(Q - 1) * other));
    int ck = 0;
#pragma unroll
    for (int j = 1; j < Q; ++j)
        if (j == k) ck = axis == 0 ? Lat<Q>::cx(j) : Lat<Q>::cy(j);
    const int hi = axis == 0 ? g.X : g.Y;
    uint64_t dst, src;
    if (ck > 0) {
        dst = axis == 0 ? g.cell(z, a, 1) : g.cell(z, 1, a);
        src = axis == 0 ? g.cell(z, a, hi + 1) : g.cell(z, hi + 1, a);
    } else if (ck < 0) {
        dst = axis == 0 ? g.cell(z, a, hi) : g.cell(z, hi, a);
        src = axis == 0 ? g.cell(z, a, 0) : g.cell(z, 0, a);
    } else {
        return;
    }
    n[k * g.ncell + dst] = n[k * g.ncell + src];
}

// CalcMoments::_get_local_moments (CalcMoments.hpp:22-37): rho = sum_k n_k,
// u = (sum_k n_k c_k) * (1 / rho), k ascending.
template <int Q, class R>
struct MomentSum {
    R rl = (R)0, u0 = (R)0, u1 = (R)0, u2 = (R)0;
    __device__ void add(int k, R nk) {
        rl += nk;
        u0 += nk * (R)Lat<Q>::cx(k);
        u1 += nk * (R)Lat<Q>::cy(k);
        u2 += nk * (R)Lat<Q>::cz(k);
    }
    __device__ void finish(R& rho, R (&u)[3]) const {
        const R ri = (R)1 / rl;
        rho = rl;
        u[0] = u0 * ri;
        u[1] = u1 * ri;
        u[2] = u2 * ri;
    }
};

// CalcMoments::_get_local_moments (CalcMoments.hpp:22-37) on interior cells.
// The populations are read with non-temporal loads (measured 3-9% faster
// for this stage on MI355X).
template <int Q, class R>
__global__ void k_moments(LitDims g, const R* __restrict__ n, R* __restrict__ rho,
                          R* __restrict__ u) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int y = blockIdx.y + 1, z = blockIdx.z + 1;
    if (x > g.X) return;
    const uint64_t c = g.cell(z, y, x);
    MomentSum<Q, R> m;
#pragma unroll
    for (int k = 0; k < Q; ++k) m.add(k, __builtin_nontemporal_load(n + k * g.ncell + c));
    R rl, um[3];
    m.finish(rl, um);
    rho[c] = rl;
    u[c] = um[0];
    u[g.ncell + c] = um[1];
    u[2 * g.ncell + c] = um[2];
}

// Boundary::_set_velocity / _set_density (Boundary.cpp:150-187) on one face.
template <class R>
__global__ void k_set_face(LitDims g, FaceBox b, R* __restrict__ dst, int vlen, R v0, R v1, R v2) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.count()) return;
    const int x = b.lo[0] + (int)(t % b.ex[0]);
    const int y = b.lo[1] + (int)((t / b.ex[0]) % b.ex[1]);
    const int z = b.lo[2] + (int)(t / ((uint64_t)b.ex[0] * b.ex[1]));
    const uint64_t c = g.cell(z, y, x);
    dst[c] = v0;
    if (vlen == 3) {
        dst[g.ncell + c] = v1;
        dst[2 * g.ncell + c] = v2;
    }
}

// Boundary::reset (Boundary.cpp:58-89, 150-187) for one interior cell: the
// faces in order, a later face overwriting an earlier one, as the per-face
// k_set_face launches of lit_reset_velocity and lit_reset_density do.  Face f
// is the layer at coordinate X|Y|Z (f even) or 1 (f odd) on axis f / 2, all
// interior cells of the other two axes (Boundary.cpp:113-148).
template <class R>
struct FaceResets {
    int has_vel[6], noslip[6];
    R vel[6][3];
    R rho_s;
};

template <class R>
__device__ inline void apply_resets(const FaceResets<R>& fr, const LitDims& g, int x, int y, int z, R& rho,
                                    R (&u)[3]) {
    const int co[3] = {x, y, z}, dim[3] = {g.X, g.Y, g.Z};
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        if (!fr.has_vel[f] || co[f >> 1] != ((f & 1) == 0 ? dim[f >> 1] : 1)) continue;
        u[0] = fr.vel[f][0];
        u[1] = fr.vel[f][1];
        u[2] = fr.vel[f][2];
    }
#pragma unroll
    for (int f = 0; f < 6; ++f)
        if (fr.noslip[f] && co[f >> 1] == ((f & 1) == 0 ? dim[f >> 1] : 1)) rho = fr.rho_s;
}

// The previous step's moments and resets (CalcMoments, Boundary::reset)
// folded into this step's collision and push, inside plb_step or for staged
// calls that the C ABI deferred until the push.
// rho and u are computed from n exactly as k_moments and the resets leave
// them, stored (no-slip reads rho before the next moments), and used by the
// collision at once; n and t end as k_collide_push leaves them.
template <int Q, class R>
__global__ void k_moments_collide_push(LitDims g, R* __restrict__ n, R* __restrict__ rho, R* __restrict__ u,
                                       R* __restrict__ t, FaceResets<R> fr, R tau, R om, R omm, R F0, R F1,
                                       R F2) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int y = blockIdx.y + 1, z = blockIdx.z + 1;
    if (x > g.X) return;
    const uint64_t c = g.cell(z, y, x);
    R nv[Q];
    MomentSum<Q, R> m;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        nv[k] = n[k * g.ncell + c];
        m.add(k, nv[k]);
    }
    R rl, um[3];
    m.finish(rl, um);
    apply_resets(fr, g, x, y, z, rl, um);
    rho[c] = rl;
    u[c] = um[0];
    u[g.ncell + c] = um[1];
    u[2 * g.ncell + c] = um[2];
    const bool keep = outer_layer(g, x, y, z);
    R uu[3], usq;
    updated_u(um[0], um[1], um[2], rl, tau, F0, F1, F2, uu, usq);
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const R v = collide_value<Q, R>(k, nv[k], rl, uu, usq, om, omm);
        if (keep) n[k * g.ncell + c] = v;
        const int64_t off = (int64_t)Lat<Q>::cx(k) + (int64_t)Lat<Q>::cy(k) * g.nx +
                            (int64_t)Lat<Q>::cz(k) * g.nx * g.ny;
        t[k * g.ncell + (uint64_t)((int64_t)c + off)] = v;
    }
}

// Host AoS <-> device SoA with dtype conversion.  compact: host cells are
// interior only (X*Y*Z); otherwise the full halo'd array.
template <class H, class R>
__global__ void k_transfer(LitDims g, H* __restrict__ host, R* __restrict__ dev, int vlen,
                           int compact, int upload, uint64_t total) {
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const int v = (int)(t % vlen);
        const uint64_t hc = t / vlen;
        uint64_t c = hc;
        if (compact) {
            const int x = (int)(hc % g.X), y = (int)((hc / g.X) % g.Y), z = (int)(hc / ((uint64_t)g.X * g.Y));
            c = g.cell(z + 1, y + 1, x + 1);
        }
        if (upload)
            dev[v * g.ncell + c] = (R)host[t];
        else
            host[t] = (H)dev[v * g.ncell + c];
    }
}

// FP64 sums over interior cells of rho and rho*u (two deterministic passes).
template <class R>
__global__ void k_global_partial(LitDims g, const R* __restrict__ rho, const R* __restrict__ u,
                                 double* __restrict__ partial) {
    __shared__ double sh[4][kThreads];
    double acc[4] = {0, 0, 0, 0};
    const uint64_t total = (uint64_t)g.X * g.Y * g.Z;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % g.X) + 1, y = (int)((i / g.X) % g.Y) + 1,
                  z = (int)(i / ((uint64_t)g.X * g.Y)) + 1;
        const uint64_t c = g.cell(z, y, x);
        const double r = (double)rho[c];
        acc[0] += r;
        for (int d = 0; d < 3; ++d) acc[1 + d] += r * (double)u[d * g.ncell + c];
    }
    for (int d = 0; d < 4; ++d) sh[d][threadIdx.x] = acc[d];
    __syncthreads();
    for (unsigned s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int d = 0; d < 4; ++d) sh[d][threadIdx.x] += sh[d][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 4; ++d) partial[4 * blockIdx.x + d] = sh[d][0];
}

__global__ void k_global_final(const double* __restrict__ partial, int nblocks, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a[4] = {0, 0, 0, 0};
    for (int b = 0; b < nblocks; ++b)
        for (int d = 0; d < 4; ++d) a[d] += partial[4 * b + d];
    for (int d = 0; d < 4; ++d) out[d] = a[d];
}

int ensure_scratch(plb_state* s, size_t bytes) {
    if (s->scratch_bytes >= bytes) return PLB_OK;
    if (s->scratch) PLB_HIP_TRY(hipFree(s->scratch));
    s->scratch = nullptr;
    s->scratch_bytes = 0;
    PLB_HIP_TRY(hipMalloc(&s->scratch, bytes));
    s->scratch_bytes = bytes;
    return PLB_OK;
}

// The per-cell stages (collide, push, moments) launch one block per x-row
// piece of up to 256 interior cells (whole wavefronts), the row's y and z
// taken from the block index.
unsigned interior_block(const plb_state* s) {
    const unsigned w = (unsigned)((s->X + 63) / 64) * 64u;
    return w < 256u ? w : 256u;
}
dim3 interior_grid(const plb_state* s) {
    const unsigned b = interior_block(s);
    return dim3((unsigned)((s->X + b - 1) / b), (unsigned)s->Y, (unsigned)s->Z);
}

// Stores the in-place values the last collide+push kernel left out (see
// outer_layer), before anything reads ntmp or changes n.
int lit_ntmp_restore(plb_state* s) {
    if (!s->lit_ntmp_folded) return PLB_OK;
    const LitDims g = dims_of(s);
    const int rc = with_qr(s, [&](auto qc, auto rr) {
        constexpr int Q = decltype(qc)::value;
        using R = decltype(rr);
        hipLaunchKernelGGL((k_ntmp_restore<Q, R>), interior_grid(s), dim3(interior_block(s)), 0, s->stream, g,
                           (const R*)s->buf[s->cur], (R*)s->buf[s->cur ^ 1]);
        PLB_HIP_TRY(hipGetLastError());
        return PLB_OK;
    });
    if (rc == PLB_OK) s->lit_ntmp_folded = false;
    return rc;
}

}  // namespace

// ------------------------------------------------------------ launchers ----
int lit_initialize(plb_state* s, double rho0, const double u0[3], const plb_init_spec* init,
                   const double* tg_dev) {
    const LitDims g = dims_of(s);
    const int kind = init ? init->kind : PLB_INIT_UNIFORM;
    const uint64_t seed = init ? init->seed : 0;
    const double amp = init ? init->amp : 0.0;
    return with_qr(s, [&](auto qc, auto rr) {
        constexpr int Q = decltype(qc)::value;
        using R = decltype(rr);
        hipLaunchKernelGGL((k_init<Q, R>), dim3(blocks_for(g.ncell, kThreads)), dim3(kThreads), 0,
                           s->stream, g, (R*)s->buf[s->cur], (R*)s->buf[s->cur ^ 1], (R*)s->rho,
                           (R*)s->u, rho0, u0[0], u0[1], u0[2], kind, seed, amp, tg_dev);
        PLB_HIP_TRY(hipGetLastError());
        s->lit_ntmp_folded = false;
        return PLB_OK;
    });
}

namespace {
// tau and omega = 1/tau as CollisionSRT.cpp:19-20 hold them, and 1 - omega.
template <class R>
void relaxation(double tau, R* tr, R* om, R* omm) {
    *tr = (R)tau;
    if (std::is_same<R, float>::value)
        *om = (R)(1. / (double)*tr);
    else
        *om = (R)(1.0 / tau);
    *omm = (R)1 - *om;
}
}  // namespace

int lit_collide(plb_state* s, double tau, const double F[3]) {
    const LitDims g = dims_of(s);
    const int rc = lit_ntmp_restore(s);  // the collision changes the n that ntmp is kept in
    if (rc) return rc;
    return with_qr(s, [&](auto qc, auto rr) {
        constexpr int Q = decltype(qc)::value;
        using R = decltype(rr);
        R tr, om, omm;
        relaxation(tau, &tr, &om, &omm);
        hipLaunchKernelGGL((k_collide<Q, R>), interior_grid(s), dim3(interior_block(s)), 0, s->stream, g,
                           (R*)s->buf[s->cur], (const R*)s->rho, (const R*)s->u, tr, om, omm, (R)F[0], (R)F[1],
                           (R)F[2]);
        PLB_HIP_TRY(hipGetLastError());
        return PLB_OK;
    });
}

int lit_collide_push(plb_state* s, double tau, const double F[3]) {
    const LitDims g = dims_of(s);
    const int rc = with_qr(s, [&](auto qc, auto rr) {
        constexpr int Q = decltype(qc)::value;
        using R = decltype(rr);
        R tr, om, omm;
        relaxation(tau, &tr, &om, &omm);
        hipLaunchKernelGGL((k_collide_push<Q, R>), interior_grid(s), dim3(interior_block(s)), 0, s->stream, g,
                           (R*)s->buf[s->cur], (const R*)s->rho, (const R*)s->u, (R*)s->buf[s->cur ^ 1], tr, om,
                           omm, (R)F[0], (R)F[1], (R)F[2]);
        PLB_HIP_TRY(hipGetLastError());
        return PLB_OK;
    });
    if (rc == PLB_OK) {
        s->cur ^= 1;  // std::swap(simdata.ntmp, simdata.n)
        s->lit_ntmp_folded = true;
    }
    return rc;
}

int lit_moments_collide_push(plb_state* s, double tau, const double F[3], const plb_boundary* bc_vel,
                             const plb_boundary* bc_rho) {
    const LitDims g = dims_of(s);
    const int rc = with_qr(s, [&](auto qc, auto rr) {
        constexpr int Q = decltype(qc)::value;
        using R = decltype(rr);
        R tr, om, omm;
        relaxation(tau, &tr, &om, &omm);
        FaceResets<R> fr;
        for (int f = 0; f < 6; ++f) {
            fr.has_vel[f] = bc_vel->has_vel[f] != 0;
            fr.noslip[f] = bc_rho->type[f] == kNoslip;
            for (int i = 0; i < 3; ++i) fr.vel[f][i] = (R)bc_vel->vel[f][i];
        }
        fr.rho_s = (R)bc_rho->rho_s;
        hipLaunchKernelGGL((k_moments_collide_push<Q, R>), interior_grid(s), dim3(interior_block(s)), 0, s->stream,
                           g, (R*)s->buf[s->cur], (R*)s->rho, (R*)s->u, (R*)s->buf[s->cur ^ 1], fr, tr, om, omm,
                           (R)F[0], (R)F[1], (R)F[2]);
        PLB_HIP_TRY(hipGetLastError());
        return PLB_OK;
    });
    if (rc == PLB_OK) {
        s->cur ^= 1;  // std::swap(simdata.ntmp, simdata.n)
        s->lit_ntmp_folded = true;
    }
    return rc;
}

int lit_push(plb_state* s) {
    const LitDims g = dims_of(s);
    const int rc = with_qr(s, [&](auto qc, auto rr) {
        constexpr int Q = decltype(qc)::value;
        using R = decltype(rr);
        hipLaunchKernelGGL((k_push<Q, R>), interior_grid(s), dim3(interior_block(s)), 0, s->stream, g,
                           (const R*)s->buf[s->cur], (R*)s->buf[s->cur ^ 1]);
        PLB_HIP_TRY(hipGetLastError());
        return PLB_OK;
    });
    if (rc == PLB_OK) {
        s->cur ^= 1;  // std::swap(simdata.ntmp, simdata.n)
        s->lit_ntmp_folded = false;  // the push wrote every cell off the outer layer
    }
    return rc;
}

// Staging of plb_apply_noslip_noalloc: one face's gathered values, held by
// every GPU the library runs on until the process ends.  Host code uses the
// device address hipGetSymbolAddress returns for it (noalloc_stage), never
// the array's name.
constexpr size_t kNoAllocStageBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char g_noalloc_stage[kNoAllocStageBytes];

namespace {
// The no-slip pass with a given staging buffer (stage_bytes of device
// memory), or with the state's scratch buffer, grown as needed, when stage is
// null.
int noslip_with(plb_state* s, const plb_boundary* bc, void* stage, size_t stage_bytes) {
    const LitDims g = dims_of(s);
    for (int f = 0; f < 6; ++f) {
        if (bc->type[f] != kNoslip) continue;
        if (!bc->has_vel[f])
            return fail(PLB_ERR_OUT_OF_RANGE, "no-slip face without a prescribed velocity");
        const FaceBox b = face_box(f, s->X, s->Y, s->Z);
        const int rc = with_qr(s, [&](auto qc, auto rr) {
            constexpr int Q = decltype(qc)::value;
            using R = decltype(rr);
            const uint64_t total = b.count() * (Q - 1);
            void* buf = stage;
            if (!buf) {
                const int e = ensure_scratch(s, total * sizeof(R));
                if (e) return e;
                buf = s->scratch;
            }
            const R cs2inv = (R)1 / (R)(1.0 / 3.0);
            hipLaunchKernelGGL((k_noslip_gather<Q, R>), dim3(blocks_for(total, kThreads)),
                               dim3(kThreads), 0, s->stream, g, b, (const R*)s->buf[s->cur],
                               (const R*)s->rho, (R)bc->vel[f][0], (R)bc->vel[f][1],
                               (R)bc->vel[f][2], cs2inv, (R*)buf, total);
            PLB_HIP_TRY(hipGetLastError());
            hipLaunchKernelGGL((k_noslip_scatter<Q, R>), dim3(blocks_for(total, kThreads)),
                               dim3(kThreads), 0, s->stream, g, b, (R*)s->buf[s->cur],
                               (const R*)buf, total);
            PLB_HIP_TRY(hipGetLastError());
            return PLB_OK;
        });
        if (rc) return rc;
    }
    (void)stage_bytes;
    return PLB_OK;
}

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices];

// The device address of g_noalloc_stage on the current device (s->device),
// looked up on the first call on that device.
int noalloc_stage(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return fail(PLB_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(g_noalloc_stage));
        if (e != hipSuccess || !p)
            return fail(PLB_ERR_HIP, std::string("no-allocation staging: ") + hipGetErrorString(e));
        g_noalloc_ptr[device] = p;
    }
    *out = g_noalloc_ptr[device];
    return PLB_OK;
}
}  // namespace

int lit_noslip(plb_state* s, const plb_boundary* bc) { return noslip_with(s, bc, nullptr, 0); }

int lit_noslip_noalloc(plb_state* s, const plb_boundary* bc) {
    // every face must fit the staging before anything runs
    const size_t rbytes = s->variant == kF64 ? sizeof(double) : sizeof(float);
    for (int f = 0; f < 6; ++f) {
        if (bc->type[f] != kNoslip) continue;
        const uint64_t need = face_box(f, s->X, s->Y, s->Z).count() * (uint64_t)(s->q - 1) * rbytes;
        if (need > kNoAllocStageBytes)
            return fail(PLB_ERR_INVALID_ARGUMENT,
                        "no-slip face of " + std::to_string(need) + " bytes exceeds the " +
                            std::to_string(kNoAllocStageBytes >> 20) +
                            " MiB no-allocation staging; use plb_apply_noslip");
    }
    void* stage = nullptr;
    const int rc = noalloc_stage(s->device, &stage);
    if (rc) return rc;
    return noslip_with(s, bc, stage, kNoAllocStageBytes);
}

int lit_periodic(plb_state* s, const plb_boundary* bc) {
    const LitDims g = dims_of(s);
    for (int axis = 0; axis < 2; ++axis) {
        if (!(bc->type[2 * axis] == kPeriodic && bc->type[2 * axis + 1] == kPeriodic)) continue;
        const int rc = with_qr(s, [&](auto qc, auto rr) {
            constexpr int Q = decltype(qc)::value;
            using R = decltype(rr);
            const uint64_t total = (uint64_t)s->Z * (axis == 0 ? s->Y : s->X) * (Q - 1);
            hipLaunchKernelGGL((k_periodic<Q, R>), dim3(blocks_for(total, kThreads)), dim3(kThreads),
                               0, s->stream, g, axis, (R*)s->buf[s->cur]);
            PLB_HIP_TRY(hipGetLastError());
            return PLB_OK;
        });
        if (rc) return rc;
    }
    return PLB_OK;
}

int lit_moments(plb_state* s) {
    const LitDims g = dims_of(s);
    return with_qr(s, [&](auto qc, auto rr) {
        constexpr int Q = decltype(qc)::value;
        using R = decltype(rr);
        hipLaunchKernelGGL((k_moments<Q, R>), interior_grid(s), dim3(interior_block(s)), 0, s->stream, g,
                           (const R*)s->buf[s->cur], (R*)s->rho, (R*)s->u);
        PLB_HIP_TRY(hipGetLastError());
        return PLB_OK;
    });
}

int lit_reset_velocity(plb_state* s, const plb_boundary* bc) {
    const LitDims g = dims_of(s);
    for (int f = 0; f < 6; ++f) {
        if (!bc->has_vel[f]) continue;
        const FaceBox b = face_box(f, s->X, s->Y, s->Z);
        const int rc = with_qr(s, [&](auto, auto rr) {
            using R = decltype(rr);
            hipLaunchKernelGGL((k_set_face<R>), dim3(blocks_for(b.count(), kThreads)), dim3(kThreads),
                               0, s->stream, g, b, (R*)s->u, 3, (R)bc->vel[f][0], (R)bc->vel[f][1],
                               (R)bc->vel[f][2]);
            PLB_HIP_TRY(hipGetLastError());
            return PLB_OK;
        });
        if (rc) return rc;
    }
    return PLB_OK;
}

int lit_reset_density(plb_state* s, const plb_boundary* bc) {
    const LitDims g = dims_of(s);
    for (int f = 0; f < 6; ++f) {
        if (bc->type[f] != kNoslip) continue;
        const FaceBox b = face_box(f, s->X, s->Y, s->Z);
        const int rc = with_qr(s, [&](auto, auto rr) {
            using R = decltype(rr);
            hipLaunchKernelGGL((k_set_face<R>), dim3(blocks_for(b.count(), kThreads)), dim3(kThreads),
                               0, s->stream, g, b, (R*)s->rho, 1, (R)bc->rho_s, (R)0, (R)0);
            PLB_HIP_TRY(hipGetLastError());
            return PLB_OK;
        });
        if (rc) return rc;
    }
    return PLB_OK;
}

int lit_transfer(plb_state* s, int field, int layout, int dtype, void* host, bool upload) {
    const LitDims g = dims_of(s);
    if ((field == PLB_FIELD_NTMP && !upload) || (field == PLB_FIELD_N && upload)) {
        const int r = lit_ntmp_restore(s);
        if (r) return r;
    }
    int vlen;
    void* dev;
    switch (field) {
        case PLB_FIELD_RHO: vlen = 1; dev = s->rho; break;
        case PLB_FIELD_U: vlen = 3; dev = s->u; break;
        case PLB_FIELD_N: vlen = s->q; dev = s->buf[s->cur]; break;
        case PLB_FIELD_NTMP: vlen = s->q; dev = s->buf[s->cur ^ 1]; break;
        default: return fail(PLB_ERR_INVALID_ARGUMENT, "unknown field");
    }
    const int compact = layout == PLB_LAYOUT_COMPACT;
    const uint64_t cells = compact ? (uint64_t)s->X * s->Y * s->Z : s->ncell;
    const uint64_t total = cells * vlen;
    const size_t hbytes = total * (dtype == PLB_DTYPE_F64 ? 8 : 4);
    void* stage = nullptr;
    PLB_HIP_TRY(hipMalloc(&stage, hbytes));
    int rc = PLB_OK;
    if (upload) {
        hipError_t e = hipMemcpyAsync(stage, host, hbytes, hipMemcpyHostToDevice, s->stream);
        if (e != hipSuccess) rc = fail(PLB_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
    }
    if (rc == PLB_OK) {
        rc = with_qr(s, [&](auto, auto rr) {
            using R = decltype(rr);
            if (dtype == PLB_DTYPE_F64)
                hipLaunchKernelGGL((k_transfer<double, R>), dim3(stride_blocks(total, kThreads)),
                                   dim3(kThreads), 0, s->stream, g, (double*)stage, (R*)dev, vlen,
                                   compact, upload ? 1 : 0, total);
            else
                hipLaunchKernelGGL((k_transfer<float, R>), dim3(stride_blocks(total, kThreads)),
                                   dim3(kThreads), 0, s->stream, g, (float*)stage, (R*)dev, vlen,
                                   compact, upload ? 1 : 0, total);
            PLB_HIP_TRY(hipGetLastError());
            return PLB_OK;
        });
    }
    if (rc == PLB_OK && !upload) {
        hipError_t e = hipMemcpyAsync(host, stage, hbytes, hipMemcpyDeviceToHost, s->stream);
        if (e != hipSuccess) rc = fail(PLB_ERR_HIP, std::string("download: ") + hipGetErrorString(e));
    }
    hipError_t e = hipStreamSynchronize(s->stream);
    (void)hipFree(stage);
    if (rc == PLB_OK && e != hipSuccess) rc = fail(PLB_ERR_HIP, hipGetErrorString(e));
    if (rc == PLB_OK && upload && field == PLB_FIELD_NTMP) s->lit_ntmp_folded = false;
    return rc;
}

int lit_global(plb_state* s, double out[4]) {
    const LitDims g = dims_of(s);
    const int nb = 512;
    int rc = ensure_scratch(s, (size_t)(4 * nb + 4) * sizeof(double));
    if (rc) return rc;
    double* partial = (double*)s->scratch;
    rc = with_qr(s, [&](auto, auto rr) {
        using R = decltype(rr);
        hipLaunchKernelGGL((k_global_partial<R>), dim3(nb), dim3(kThreads), 0, s->stream, g,
                           (const R*)s->rho, (const R*)s->u, partial);
        PLB_HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_global_final, dim3(1), dim3(64), 0, s->stream, partial, nb, partial + 4 * nb);
        PLB_HIP_TRY(hipGetLastError());
        return PLB_OK;
    });
    if (rc) return rc;
    PLB_HIP_TRY(hipMemcpyAsync(out, partial + 4 * nb, 4 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    PLB_HIP_TRY(hipStreamSynchronize(s->stream));
    return PLB_OK;
}

}  // namespace plb
