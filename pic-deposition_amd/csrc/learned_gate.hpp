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
    // All pairs of both sums requested before the first add (the sums are
    // small and L2-resident; U pairs of each per lane in one round trip cover
    // up to 64 U blocks, longer lists take more rounds).
    constexpr int U = 16;
    const double2* st = reinterpret_cast<const double2*>(la.stats);
    const double2* mo = reinterpret_cast<const double2*>(la.moments);
    const int nmax = la.nbs > la.nbm ? la.nbs : la.nbm;
    const double sum = *la.sum;
    for (int b0 = 0; b0 < nmax; b0 += 64 * U) {
        double2 v[U], w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int b = b0 + u * 64 + lane;
            v[u] = b < la.nbs ? st[b] : make_double2(0.0, 0.0);
            w[u] = b < la.nbm ? mo[b] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            a[0] += v[u].x;
            a[1] += v[u].y;
            a[2] += w[u].x;
            a[3] += w[u].y;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) a[k] += __shfl_xor(a[k], o, 64);
    LearnedGate g;
    g.mean = sum * la.sum_scale / (double)n;
    const double noise = la.kappa * a[3] / (la.cv * la.cv);  // expected squared shot-noise norm of the plain deposit
    g.ratio = noise > 0.0 ? a[1] / noise : INFINITY;
    g.scale = (a[2] / la.cv) / (g.mean * a[0]);
    g.q1 = a[2];
    g.ok = g.mean != 0.0 && isfinite(g.ratio) && isfinite(g.scale) && g.ratio <= la.tol;
    return g;
}

}  // namespace pichip
