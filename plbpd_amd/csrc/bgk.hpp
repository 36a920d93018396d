// The BGK update of the fused and batched kernels: moments, the
// velocity-shift force (CollisionSRT.cpp:72-91) and the blend with the
// equilibrium (CollisionSRT.cpp:170-175, EqlbDist.hpp:26-38), in the storage
// variant's arithmetic T.
//
// Sums over directions skip the zero components of c_k (a product with a
// zero component would be an exact +-0 whose only effect is on the sign of a
// zero sum), so a D3Q19 moment or c.u costs its nonzero terms only; the
// nonzero terms are taken in ascending k (moments) and x, y, z order (c.u),
// the order of the reference's expressions.  The CPU oracle's corrected step
// sums in exactly this way.  The literal kernels keep the reference's dense
// expressions.
#pragma once

#include "lattice.hpp"

namespace plb {

// rho (or its deviation, shifted storage) and momentum j = sum_k f_k c_k.
template <int Q, class T>
__device__ inline void moment_sums(const T (&f)[Q], T& r, T& j0, T& j1, T& j2) {
    r = (T)0;
    j0 = (T)0;
    j1 = (T)0;
    j2 = (T)0;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        r += f[k];
        if (Lat<Q>::cx(k) != 0) j0 += f[k] * (T)Lat<Q>::cx(k);
        if (Lat<Q>::cy(k) != 0) j1 += f[k] * (T)Lat<Q>::cy(k);
        if (Lat<Q>::cz(k) != 0) j2 += f[k] * (T)Lat<Q>::cz(k);
    }
}

// c_k . u over the nonzero components of c_k, in x, y, z order (0 for k = 0).
template <int Q, class T>
__device__ inline T c_dot(int k, T u0, T u1, T u2) {
    const int cx = Lat<Q>::cx(k), cy = Lat<Q>::cy(k), cz = Lat<Q>::cz(k);
    T cu = (T)0;
    bool any = false;
    if (cx != 0) {
        cu = (T)cx * u0;
        any = true;
    }
    if (cy != 0) {
        cu = any ? cu + (T)cy * u1 : (T)cy * u1;
        any = true;
    }
    if (cz != 0) cu = any ? cu + (T)cz * u2 : (T)cz * u2;
    return cu;
}

// Weight w_k times the storage scale (exact: a power of two).
template <int Q, class P>
__device__ inline typename P::T w_scaled(int k) {
    return (typename P::T)Lat<Q>::w(k) * P::kScale;
}

// Directions 2j - 1 and 2j are opposite with equal weights (LBModel.cpp's
// ordering), which the pair loop of bgk_update relies on.
template <int Q>
constexpr bool pairs_opposite() {
    if (Q % 2 != 1) return false;
    for (int k = 1; k < Q; k += 2)
        if (Lat<Q>::cx(k + 1) != -Lat<Q>::cx(k) || Lat<Q>::cy(k + 1) != -Lat<Q>::cy(k) ||
            Lat<Q>::cz(k + 1) != -Lat<Q>::cz(k) || Lat<Q>::w(k + 1) != Lat<Q>::w(k))
            return false;
    return true;
}

// Moments and, when `collide`, the BGK update in place, on f in the storage
// scale (P::load_raw: kScale * f, kScale a power of two).  Returns rho.
// Shifted storage sums deviations and adds 1 afterwards; its equilibrium is
// w (drho + rho (3 cu + 4.5 cu^2 - 1.5 u^2)).  Every scaled quantity is
// kScale times its unscaled value exactly, so the stored results equal those
// of the same arithmetic on unscaled values.
template <int Q, class P>
__device__ inline typename P::T bgk_update(typename P::T (&f)[Q], typename P::T tau, typename P::T om,
                                           typename P::T omm, typename P::T F0, typename P::T F1,
                                           typename P::T F2, bool collide) {
    using T = typename P::T;
    static_assert(pairs_opposite<Q>(), "bgk_update pairs direction k with k + 1 = rev(k)");
    T r, j0, j1, j2;
    moment_sums<Q, T>(f, r, j0, j1, j2);
    if (P::kScale != (T)1) {
        const T inv = (T)1 / P::kScale;
        r *= inv;
        j0 *= inv;
        j1 *= inv;
        j2 *= inv;
    }
    const T drho = r;
    if (P::shifted) r = (T)1 + r;
    if (!collide) return r;
    const T ri = (T)1 / r;
    const T tri = tau / r;
    const T u0 = j0 * ri + F0 * tri;
    const T u1 = j1 * ri + F1 * tri;
    const T u2 = j2 * ri + F2 * tri;
    const T usq = u0 * u0 + u1 * u1 + u2 * u2;
    const T u15 = (T)1.5 * usq;
    // rest direction: c_0 . u = 0
    {
        const T w = w_scaled<Q, P>(0);
        const T cu = c_dot<Q, T>(0, u0, u1, u2);
        const T t3 = (T)3 * cu, q = ((T)4.5 * cu) * cu;
        if (P::shifted)
            f[0] = omm * f[0] + om * (w * (drho + r * ((t3 + q) - u15)));
        else
            f[0] = omm * f[0] + om * ((w * r) * ((((T)1 + t3) + q) - u15));
    }
    // Directions k and k + 1 = rev(k) are opposite, with equal weights.  Each
    // rounding is symmetric under negation, so c_{k+1}.u == -(c_k.u),
    // 3 c_{k+1}.u == -(3 c_k.u) and (4.5 c_{k+1}.u) c_{k+1}.u == (4.5 c_k.u) c_k.u
    // exactly (a zero c.u gives +0 sums either way): the pair shares its
    // products, and every result is the one the per-direction expression
    // (3 cu + 4.5 cu cu) - 1.5 u^2 gives.
#pragma unroll
    for (int k = 1; k < Q; k += 2) {
        const T w = w_scaled<Q, P>(k);
        const T cu = c_dot<Q, T>(k, u0, u1, u2);
        const T t3 = (T)3 * cu, q = ((T)4.5 * cu) * cu;
        if (P::shifted) {
            f[k] = omm * f[k] + om * (w * (drho + r * ((t3 + q) - u15)));
            f[k + 1] = omm * f[k + 1] + om * (w * (drho + r * ((q - t3) - u15)));
        } else {
            const T wr = w * r;
            f[k] = omm * f[k] + om * (wr * ((((T)1 + t3) + q) - u15));
            f[k + 1] = omm * f[k + 1] + om * (wr * ((((T)1 - t3) + q) - u15));
        }
    }
    return r;
}

}  // namespace plb
