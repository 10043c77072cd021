// Per-particle device operations shared by the single-grid step kernels
// (step.hip) and the batched ensemble (ensemble.hip): CIC field gather,
// leapfrog push and the counter-based particle initialiser.
#pragma once

#include "pic_common.hpp"
#include "pic_kernels.hpp"

namespace pichip {
namespace {

// CIC interpolation of (Ex, Ey) at a position, as gatherFieldsSoA
// (reference src/deposition/deposition.cpp:335-356) in FP32.
__device__ __forceinline__ void gather_cic(float x, float y, const float* __restrict__ Ex,
                                           const float* __restrict__ Ey, const DevDomain& d, float& ex, float& ey) {
    int cx, cy;
    float wx, wy;
    cell_coord(x, d.inv_dx, d.dx, d.exact, cx, wx);
    cell_coord(y, d.inv_dy, d.dy, d.exact, cy, wy);
    const int ix = wrap_index(cx, d.Nx), iy = wrap_index(cy, d.Ny);
    const int ixp = ix + 1 == d.Nx ? 0 : ix + 1;
    const int iyp = iy + 1 == d.Ny ? 0 : iy + 1;
    const size_t r0 = (size_t)iy * d.Nx, r1 = (size_t)iyp * d.Nx;
    const float ux = 1.0f - wx, uy = 1.0f - wy;
    ex = Ex[r0 + ix] * ux * uy + Ex[r0 + ixp] * wx * uy + Ex[r1 + ix] * ux * wy + Ex[r1 + ixp] * wx * wy;
    ey = Ey[r0 + ix] * ux * uy + Ey[r0 + ixp] * wx * uy + Ey[r1 + ix] * ux * wy + Ey[r1 + ixp] * wx * wy;
}

// x <- fmod(x + v dt, L) wrapped into [0, L) (pushParticlesSoA,
// deposition.cpp:367-374).  The displacement is reduced modulo L before it is
// added, so |v dt| >> L does not cost FP32 position bits; in exact arithmetic
// this equals fmod(x + v dt, L).
__device__ __forceinline__ float push_coord(float x, float v, float dt, float L) {
    const float s = v * dt;
    if (fabsf(s) < L) {  // fmodf(s, L) == s
        float t = x + s;
        // fmodf(t, L) is t on (-L, L) and t - L (exact) on [L, 2L)
        if (t >= 0.0f && t < L) return t;
        if (t >= L && t < L + L) return t - L;
        if (t < 0.0f && t > -L) return t + L;
        t = fmodf(t, L);
        return t < 0.0f ? t + L : t;
    }
    x += fmodf(s, L);
    x = fmodf(x, L);
    if (x < 0.0f) x += L;
    return x;
}

// ---- Philox4x32-10 (Salmon et al., SC'11) ----------------------------------
struct U4 {
    uint32_t v[4];
};

__device__ __forceinline__ U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c0;
        const uint64_t p1 = (uint64_t)M1 * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += W0;
        k1 += W1;
    }
    U4 o;
    o.v[0] = c0;
    o.v[1] = c1;
    o.v[2] = c2;
    o.v[3] = c3;
    return o;
}

// 53-bit uniform in [0, 1) from two words.
__device__ __forceinline__ double u01(uint32_t a, uint32_t b) {
    const uint64_t r = (((uint64_t)a << 32) | b) >> 11;
    return (double)r * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ float wrap_pos(double x, double L) {
    float f = (float)x;
    const float Lf = (float)L;
    if (f >= Lf) f -= Lf;
    if (f < 0.0f) f += Lf;
    if (f >= Lf) f = 0.0f;
    return f;
}

struct ParticleInit {
    float x, y, vx, vy, q, m;
};

// Particle i of n drawn from Philox(i, seed): kind 0 uniform positions,
// kind 1 x from the inverse CDF of 1 + amp cos(k x), kind 2 two beams +-amp
// (first half +, second half -); velocities U(-vmax, vmax) and/or Maxwellian
// vth, plus pert * sin(k x) along x.
__device__ __forceinline__ ParticleInit init_particle(int64_t i, int64_t n, const DevDomain& d,
                                                      const InitParams& p) {
    const uint32_t k0 = (uint32_t)p.seed, k1 = (uint32_t)(p.seed >> 32);
    const U4 a = philox((uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0u, 0u, k0, k1);
    const U4 b = philox((uint32_t)i, (uint32_t)((uint64_t)i >> 32), 1u, 0u, k0, k1);
    const double ux = u01(a.v[0], a.v[1]);
    const double uy = u01(a.v[2], a.v[3]);
    const double kx = 2.0 * M_PI * (double)p.mode / d.Lx;
    double x = ux * d.Lx;
    if (p.kind == 1 && p.amp != 0.0) {
        // Inverse CDF of n(x) ~ 1 + amp cos(kx) on [0, Lx) by Newton iteration.
        const double target = ux * d.Lx;
        for (int it = 0; it < 8; ++it) {
            const double F = x + p.amp / kx * sin(kx * x) - target;
            const double dF = 1.0 + p.amp * cos(kx * x);
            x -= F / dF;
        }
    }
    const double y = uy * d.Ly;
    double vx = 0.0, vy = 0.0;
    if (p.vmax > 0.0) {
        vx = (2.0 * u01(b.v[0], b.v[1]) - 1.0) * p.vmax;
        vy = (2.0 * u01(b.v[2], b.v[3]) - 1.0) * p.vmax;
    }
    if (p.vth > 0.0) {
        // Box-Muller on (1 - u) in (0, 1].
        const double r = sqrt(-2.0 * log(1.0 - u01(b.v[0], b.v[1])));
        const double th = 2.0 * M_PI * u01(b.v[2], b.v[3]);
        vx += p.vth * r * cos(th);
        vy += p.vth * r * sin(th);
    }
    if (p.kind == 2) vx += (i < n / 2 ? 1.0 : -1.0) * p.amp;
    if (p.pert != 0.0) vx += p.pert * sin(kx * x);
    ParticleInit o;
    o.x = wrap_pos(x, d.Lx);
    o.y = wrap_pos(y, d.Ly);
    o.vx = (float)vx;
    o.vy = (float)vy;
    o.q = (float)p.qp;
    o.m = (float)p.mp;
    return o;
}

}  // namespace
}  // namespace pichip
