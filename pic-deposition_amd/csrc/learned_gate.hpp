// Gate and renormalisation of a learned density, evaluated by one wavefront
// (the first pass of the LDS Poisson solve runs it, see LearnedApply): the
// per-block FP64 sums of the output convolution (sum of u = 1 + out /
// out_scale, and of (mean u - rho_plain)^2) and of the particles' charge
// moments (sum q, sum q^2) are added in a fixed order, so every block takes the
// same decision.  The residual ratio ||rho_learned - rho_plain||^2 / (kappa
// sum(q^2) / cv^2) must not exceed tol; the scale makes sum(rho) * cv equal
// sum(q).
#pragma once

#include "pic_kernels.hpp"

namespace pichip {

struct LearnedGate {
    bool ok;
    double mean, scale, ratio, q1;
};

__device__ __forceinline__ LearnedGate learned_gate_wave(const LearnedApply& la, int64_t n) {
    const int lane = threadIdx.x & 63;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    // 16 pairs per lane in flight at once (the sums are small and L2-resident;
    // one dependent load per pair would make this loop latency-bound)
    constexpr int U = 16;
    const double2* st = reinterpret_cast<const double2*>(la.stats);
    for (int b0 = 0; b0 < la.nbs; b0 += 64 * U) {
        double2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int b = b0 + u * 64 + lane;
            v[u] = b < la.nbs ? st[b] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            a[0] += v[u].x;
            a[1] += v[u].y;
        }
    }
    const double2* mo = reinterpret_cast<const double2*>(la.moments);
    for (int b0 = 0; b0 < la.nbm; b0 += 64 * U) {
        double2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int b = b0 + u * 64 + lane;
            v[u] = b < la.nbm ? mo[b] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            a[2] += v[u].x;
            a[3] += v[u].y;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) a[k] += __shfl_xor(a[k], o, 64);
    LearnedGate g;
    g.mean = *la.sum * la.sum_scale / (double)n;
    const double noise = la.kappa * a[3] / (la.cv * la.cv);  // expected squared shot-noise norm of the plain deposit
    g.ratio = noise > 0.0 ? a[1] / noise : INFINITY;
    g.scale = (a[2] / la.cv) / (g.mean * a[0]);
    g.q1 = a[2];
    g.ok = g.mean != 0.0 && isfinite(g.ratio) && isfinite(g.scale) && g.ratio <= la.tol;
    return g;
}

}  // namespace pichip
