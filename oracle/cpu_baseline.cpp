// CPU baseline -- TEST/BENCH INFRASTRUCTURE ONLY (bench.py's cpu_baseline leg).
//
// A restatement of the reference's own FP32 CPU step as it runs with TBB and
// AVX2 (src/PLBPD.cpp:21-33), timed on the host: float AoS fields with one halo
// layer (Field.hpp:60-71), a parallel-for over z-slabs (CollisionSRT.cpp:49-69,
// Streaming.cpp:85-94, CalcMoments.hpp:46-73) -- OpenMP here, because the
// image has no TBB headers -- the 8-wide AVX2/FMA collision with its scalar
// double-promoted tail (CollisionSRT.cpp:93-148), a scalar push with an A/B
// swap, no-slip faces (Boundary.cpp:189-221), serial periodic copies
// (Boundary.cpp:223-281), moments and reset, with per-phase wall timers like
// PLBPD::print_times (PLBPD.cpp:37-48).
#include <immintrin.h>
#include <omp.h>

#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

struct Tables {
    int q;
    int c[27][3];
    float w[27];
    int rev[27];
};

void tables(int q, Tables* t) {
    static const int c19[19][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0},
                                   {0, 0, 1}, {0, 0, -1}, {1, 1, 0}, {-1, -1, 0}, {1, -1, 0},
                                   {-1, 1, 0}, {1, 0, 1}, {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
                                   {0, 1, 1}, {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
    static const int c27x[8][3] = {{1, 1, 1}, {-1, -1, -1}, {1, 1, -1}, {-1, -1, 1},
                                   {1, -1, 1}, {-1, 1, -1}, {-1, 1, 1}, {1, -1, -1}};
    t->q = q;
    for (int k = 0; k < 19 && k < q; ++k)
        for (int i = 0; i < 3; ++i) t->c[k][i] = c19[k][i];
    for (int k = 19; k < q; ++k)
        for (int i = 0; i < 3; ++i) t->c[k][i] = c27x[k - 19][i];
    for (int k = 0; k < q; ++k) {
        t->rev[k] = k == 0 ? 0 : ((k & 1) ? k + 1 : k - 1);
        double w;
        if (q == 19)
            w = k == 0 ? 1. / 3. : (k < 7 ? 1. / 18. : 1. / 36.);
        else
            w = k == 0 ? 8. / 27. : (k < 7 ? 2. / 27. : (k < 19 ? 1. / 54. : 1. / 216.));
        t->w[k] = (float)w;
    }
}

struct Sim {
    Tables L;
    int X, Y, Z, nx, ny, nz;
    float tau, omega;
    float F[3];
    int type[6];  // 0 unset 1 periodic 2 noslip; east west north south up down
    float vel[6][3];
    float rho_s;
    std::vector<float> rho, u, n, ntmp;
    float* N;
    float* T;
    size_t cell(int z, int y, int x) const { return ((size_t)z * ny + y) * nx + x; }
};

inline void collide_cell(const Sim& s, size_t c, float* cu) {
    const int q = s.L.q;
    const float rl = s.rho[c];
    const float* ul = &s.u[3 * c];
    float* nl = s.N + (size_t)q * c;
    const float tri = s.tau / rl;
    float uu[3];
    for (int i = 0; i < 3; ++i) uu[i] = ul[i] + s.F[i] * tri;
    const float usq = uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2];
    cu[0] = 0.0f;
    for (int k = 1; k < q; ++k)
        cu[k] = s.L.c[k][0] * uu[0] + s.L.c[k][1] * uu[1] + s.L.c[k][2] * uu[2];
    const __m256 one = _mm256_set1_ps(1.0f), three = _mm256_set1_ps(3.0f),
                 four5 = _mm256_set1_ps(4.5f), om = _mm256_set1_ps(s.omega),
                 omm = _mm256_set1_ps(1.0f - s.omega), vr = _mm256_set1_ps(rl),
                 vk = _mm256_set1_ps(-1.5f * usq);
    const int full = (q / 8) * 8;
    for (int k = 0; k < full; k += 8) {
        const __m256 c8 = _mm256_loadu_ps(cu + k);
        __m256 e = _mm256_fmadd_ps(three, c8, one);
        e = _mm256_fmadd_ps(four5, _mm256_mul_ps(c8, c8), e);
        e = _mm256_mul_ps(_mm256_mul_ps(_mm256_add_ps(e, vk), vr), _mm256_loadu_ps(s.L.w + k));
        const __m256 nv = _mm256_fmadd_ps(omm, _mm256_loadu_ps(nl + k), _mm256_mul_ps(om, e));
        _mm256_storeu_ps(nl + k, nv);
    }
    for (int k = full; k < q; ++k) {
        const double neq = s.L.w[k] * rl * (1.0 + 3.0 * cu[k] + 4.5 * cu[k] * cu[k] - 1.5 * usq);
        nl[k] = (1.0f - s.omega) * nl[k] + s.omega * neq;
    }
}

void collide(Sim& s) {
#pragma omp parallel
    {
        alignas(32) float cu[32];
#pragma omp for schedule(static)
        for (int z = 1; z <= s.Z; ++z)
            for (int y = 1; y <= s.Y; ++y)
                for (int x = 1; x <= s.X; ++x) collide_cell(s, s.cell(z, y, x), cu);
    }
}

void push(Sim& s) {
    const int q = s.L.q;
#pragma omp parallel for schedule(static)
    for (int z = 1; z <= s.Z; ++z)
        for (int y = 1; y <= s.Y; ++y)
            for (int x = 1; x <= s.X; ++x) {
                const size_t c = s.cell(z, y, x);
                for (int k = 0; k < q; ++k)
                    s.T[(size_t)q * s.cell(z + s.L.c[k][2], y + s.L.c[k][1], x + s.L.c[k][0]) + k] =
                        s.N[(size_t)q * c + k];
            }
    std::swap(s.N, s.T);
}

void face_extent(const Sim& s, int f, int lo[3], int hi[3]) {
    lo[0] = 1; hi[0] = s.X + 1;
    lo[1] = 1; hi[1] = s.Y + 1;
    lo[2] = 1; hi[2] = s.Z + 1;
    const int a = f / 2, dim = a == 0 ? s.X : (a == 1 ? s.Y : s.Z);
    lo[a] = (f % 2 == 0) ? dim : 1;
    hi[a] = lo[a] + 1;
}

void noslip(Sim& s) {
    const int q = s.L.q;
    for (int f = 0; f < 6; ++f) {
        if (s.type[f] != 2) continue;
        int lo[3], hi[3];
        face_extent(s, f, lo, hi);
        const float* ub = s.vel[f];
#pragma omp parallel for schedule(static)
        for (int z = lo[2]; z < hi[2]; ++z)
            for (int y = lo[1]; y < hi[1]; ++y)
                for (int x = lo[0]; x < hi[0]; ++x)
                    for (int kp = 1; kp < q; ++kp) {
                        const size_t nb = s.cell(z + s.L.c[kp][2], y + s.L.c[kp][1], x + s.L.c[kp][0]);
                        const int k = s.L.rev[kp];
                        const float cu = s.L.c[k][0] * ub[0] + s.L.c[k][1] * ub[1] + s.L.c[k][2] * ub[2];
                        s.N[(size_t)q * nb + kp] =
                            s.N[(size_t)q * s.cell(z, y, x) + k] - 2.0f * s.L.w[k] * s.rho[nb] * cu * 3.0f;
                    }
    }
}

void periodic(Sim& s) {
    const int q = s.L.q;
    if (s.type[0] == 1 && s.type[1] == 1)
        for (int z = 1; z <= s.Z; ++z)
            for (int y = 1; y <= s.Y; ++y)
                for (int k = 1; k < q; ++k) {
                    if (s.L.c[k][0] > 0) s.N[(size_t)q * s.cell(z, y, 1) + k] = s.N[(size_t)q * s.cell(z, y, s.X + 1) + k];
                    if (s.L.c[k][0] < 0) s.N[(size_t)q * s.cell(z, y, s.X) + k] = s.N[(size_t)q * s.cell(z, y, 0) + k];
                }
    if (s.type[2] == 1 && s.type[3] == 1)
        for (int z = 1; z <= s.Z; ++z)
            for (int x = 1; x <= s.X; ++x)
                for (int k = 1; k < q; ++k) {
                    if (s.L.c[k][1] > 0) s.N[(size_t)q * s.cell(z, 1, x) + k] = s.N[(size_t)q * s.cell(z, s.Y + 1, x) + k];
                    if (s.L.c[k][1] < 0) s.N[(size_t)q * s.cell(z, s.Y, x) + k] = s.N[(size_t)q * s.cell(z, 0, x) + k];
                }
}

void moments(Sim& s) {
    const int q = s.L.q;
#pragma omp parallel for schedule(static)
    for (int z = 1; z <= s.Z; ++z)
        for (int y = 1; y <= s.Y; ++y)
            for (int x = 1; x <= s.X; ++x) {
                const size_t c = s.cell(z, y, x);
                float r = 0.f, u0 = 0.f, u1 = 0.f, u2 = 0.f;
                for (int k = 0; k < q; ++k) {
                    const float nk = s.N[(size_t)q * c + k];
                    r += nk;
                    u0 += nk * s.L.c[k][0];
                    u1 += nk * s.L.c[k][1];
                    u2 += nk * s.L.c[k][2];
                }
                const float ri = 1.0f / r;
                s.rho[c] = r;
                s.u[3 * c] = u0 * ri;
                s.u[3 * c + 1] = u1 * ri;
                s.u[3 * c + 2] = u2 * ri;
            }
}

void reset(Sim& s) {
    for (int pass = 0; pass < 2; ++pass)
        for (int f = 0; f < 6; ++f) {
            if (s.type[f] != 2) continue;
            int lo[3], hi[3];
            face_extent(s, f, lo, hi);
#pragma omp parallel for schedule(static)
            for (int z = lo[2]; z < hi[2]; ++z)
                for (int y = lo[1]; y < hi[1]; ++y)
                    for (int x = lo[0]; x < hi[0]; ++x) {
                        const size_t c = s.cell(z, y, x);
                        if (pass == 0)
                            for (int i = 0; i < 3; ++i) s.u[3 * c + i] = s.vel[f][i];
                        else
                            s.rho[c] = s.rho_s;
                    }
        }
}

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

// Runs `warmup` untimed and `steps` timed steps of the staged FP32 step on a
// (x, y, z) lattice with walls on up/down and periodic x, y, from rest.
// phase[0..5]: seconds in collide, stream, noslip, periodic, moments, reset.
// Returns the wall time of the timed loop in seconds (< 0 on bad input).
double cpubase_run(int q, int X, int Y, int Z, float nu, const float* F, int warmup, int steps,
                   int threads, double* phase) {
    if ((q != 19 && q != 27) || X < 1 || Y < 1 || Z < 3) return -1.0;
    if (threads > 0) omp_set_num_threads(threads);
    Sim s;
    tables(q, &s.L);
    s.X = X;
    s.Y = Y;
    s.Z = Z;
    s.nx = X + 2;
    s.ny = Y + 2;
    s.nz = Z + 2;
    s.tau = 3.0f * nu + 0.5f;  // CollisionSRT.cpp:19-20
    s.omega = (float)(1. / s.tau);
    for (int i = 0; i < 3; ++i) s.F[i] = F[i];
    const int types[6] = {1, 1, 1, 1, 2, 2};
    std::memcpy(s.type, types, sizeof(types));
    std::memset(s.vel, 0, sizeof(s.vel));
    s.rho_s = 1.0f;
    const size_t nc = (size_t)s.nx * s.ny * s.nz;
    s.rho.assign(nc, 1.0f);
    s.u.assign(3 * nc, 0.0f);
    s.n.assign(nc * q, 0.0f);
    s.ntmp.assign(nc * q, 0.0f);
    // first-touch the populations from the threads that will use them
#pragma omp parallel for schedule(static)
    for (int z = 0; z < s.nz; ++z)
        for (size_t i = (size_t)z * s.nx * s.ny * q; i < (size_t)(z + 1) * s.nx * s.ny * q; ++i) {
            s.n[i] = s.L.w[i % q];
            s.ntmp[i] = 0.0f;
        }
    s.N = s.n.data();
    s.T = s.ntmp.data();
    reset(s);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    double t_loop = 0.0;
    for (int it = 0; it < warmup + steps; ++it) {
        const bool timed = it >= warmup;
        const double t0 = now();
        collide(s);
        const double t1 = now();
        push(s);
        const double t2 = now();
        noslip(s);
        const double t3 = now();
        periodic(s);
        const double t4 = now();
        moments(s);
        const double t5 = now();
        reset(s);
        const double t6 = now();
        if (timed) {
            acc[0] += t1 - t0;
            acc[1] += t2 - t1;
            acc[2] += t3 - t2;
            acc[3] += t4 - t3;
            acc[4] += t5 - t4;
            acc[5] += t6 - t5;
            t_loop += t6 - t0;
        }
    }
    if (phase)
        for (int i = 0; i < 6; ++i) phase[i] = acc[i];
    return t_loop;
}

int cpubase_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
