// Spectral Poisson solve in three launches for power-of-two grids up to 1024
// cells per axis (Nx >= 4, Ny >= 1), with every 1D transform done in LDS by
// one workgroup or one wavefront.
//
// The solve is the one of poisson_fft.cpp:39-66 (rho * cv with the mean
// removed, phi_hat = rho_hat / (k^2 Nx Ny), k = 0 zeroed) followed by the
// spectral forward differences for E (the same operator as the reference's
// real-space difference of phi in exact arithmetic):
//
//  * poisson_rows_kernel: the mean-removed rows 2p and 2p+1 as the real and
//    imaginary part of one complex transform of length Nx; the two half
//    spectra are separated with the conjugate symmetry of real input (a grid
//    of one row, the quasi-1D geometry, transforms it alone).
//  * poisson_cols_kernel: one half-spectrum column: forward transform along y
//    and the Green's function and difference symbols by three wavefronts
//    together, then the inverse transforms along y of phi, Ex and Ey, one
//    wavefront each, written back as row half spectra.
//  * poisson_rows_inv_kernel: one row: the half spectra of two real outputs X
//    and Y combined into the full spectrum of X + iY, one inverse transform,
//    real part to X and imaginary part to Y.  The imaginary parts of the DC and
//    Nyquist bins are dropped, as a complex-to-real transform does.
//
// Transforms are unnormalised (forward e^{-i}, inverse e^{+i}); the 1/(Nx Ny)
// of the solve sits in the Green's function.  The 1D transform is the
// Stockham autosort form: radix-4 passes and a final radix-2 pass when the
// length is not a power of four, twiddles from one table e^{2 pi i m / L}
// tabulated in FP64 and rounded to FP32.

#include "cnn_device.hpp"
#include "learned_gate.hpp"
#include "pic_common.hpp"
#include "pic_kernels.hpp"

namespace pichip {

namespace {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// LDS ordering among the lanes of one wavefront.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Unnormalised DFT of a[0, L), SIGN = -1 forward, +1 inverse, by the whole
// workgroup (BLOCK) or by the calling wavefront alone.  b is scratch of the
// same length and tw[m] = e^{2 pi i m / L}.  Returns the buffer that holds the
// result.
template <int SIGN, bool BLOCK = true>
__device__ float2* wave_fft(float2* a, float2* b, int L, const float2* tw) {
    const int lane = BLOCK ? (int)threadIdx.x : (int)(threadIdx.x & 63);
    const int nthr = BLOCK ? (int)blockDim.x : 64;
    for (int Ls = 1; Ls < L;) {
        const int R = (L / Ls >= 4) ? 4 : 2;
        const int NR = L / R;
        const int tstep = L / (Ls * R);
        for (int j = lane; j < NR; j += nthr) {
            const int k = j & (Ls - 1);
            const int o = (j - k) * R + k;
            if (R == 4) {
                float2 x0 = a[j], x1 = a[j + NR], x2 = a[j + 2 * NR], x3 = a[j + 3 * NR];
                if (k != 0) {
                    float2 w1 = tw[k * tstep], w2 = tw[2 * k * tstep], w3 = tw[3 * k * tstep];
                    if (SIGN < 0) {
                        w1.y = -w1.y;
                        w2.y = -w2.y;
                        w3.y = -w3.y;
                    }
                    x1 = cmul(x1, w1);
                    x2 = cmul(x2, w2);
                    x3 = cmul(x3, w3);
                }
                const float2 s02 = make_float2(x0.x + x2.x, x0.y + x2.y);
                const float2 d02 = make_float2(x0.x - x2.x, x0.y - x2.y);
                const float2 s13 = make_float2(x1.x + x3.x, x1.y + x3.y);
                const float2 d13 = make_float2(x1.x - x3.x, x1.y - x3.y);
                // (SIGN i) * d13
                const float2 r13 = SIGN > 0 ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
                b[o] = make_float2(s02.x + s13.x, s02.y + s13.y);
                b[o + Ls] = make_float2(d02.x + r13.x, d02.y + r13.y);
                b[o + 2 * Ls] = make_float2(s02.x - s13.x, s02.y - s13.y);
                b[o + 3 * Ls] = make_float2(d02.x - r13.x, d02.y - r13.y);
            } else {
                float2 x0 = a[j], x1 = a[j + NR];
                if (k != 0) {
                    float2 w1 = tw[k * tstep];
                    if (SIGN < 0) w1.y = -w1.y;
                    x1 = cmul(x1, w1);
                }
                b[o] = make_float2(x0.x + x1.x, x0.y + x1.y);
                b[o + Ls] = make_float2(x0.x - x1.x, x0.y - x1.y);
            }
        }
        if (BLOCK)
            __syncthreads();
        else
            wave_lds_sync();
        float2* t = a;
        a = b;
        b = t;
        Ls *= R;
    }
    return a;
}

// The buffer in which wave_fft(a, b, L, ...) leaves its result (one swap per
// stage).
__device__ __forceinline__ float2* wave_fft_result(float2* a, float2* b, int L) {
    int stages = 0;
    for (int Ls = 1; Ls < L; Ls *= (L / Ls >= 4) ? 4 : 2) ++stages;
    return (stages & 1) ? b : a;
}

// With a deferred learned density (la.net_out set), each block first takes the
// gate decision and then writes its rows of rho (and la.out64) as
// denoise_apply_kernel would, plus la.bg, before transforming them.
__global__ __launch_bounds__(64) void poisson_rows_kernel(float* __restrict__ rho, float2* __restrict__ spec,
                                                          int W, int H, int64_t G, double bg, double cv,
                                                          const double* __restrict__ rho_sum, double sum_scale,
                                                          double sum_add, const float2* __restrict__ tw,
                                                          LearnedApply la) {
    extern __shared__ float2 sm[];
    float2* a = sm;
    float2* b = sm + W;
    float2* t = sm + 2 * W;
    const int lane = threadIdx.x;
    const int r0 = 2 * blockIdx.x, r1 = r0 + 1;
    const bool pair = r1 < H;  // a grid of one row transforms it alone
    const double mean = (*rho_sum * sum_scale + sum_add + bg * (double)G) * cv / (double)G;
    // With a deferred learned density, the rows' plain and learned values of
    // up to kPre nodes per lane are requested before the gate's sums, so their
    // loads overlap its reduction.
    constexpr int kPre = 4;
    float pre_r[2][kPre], pre_n[2][kPre];
    const bool prefetch = la.net_out && W <= 64 * kPre;
    if (prefetch) {
#pragma unroll
        for (int j = 0; j < kPre; ++j) {
            const int n = lane + 64 * j;
            const bool in = n < W;
            pre_r[0][j] = in ? rho[(size_t)r0 * W + n] : 0.0f;
            pre_n[0][j] = in ? la.net_out[(size_t)r0 * W + n] : 0.0f;
            pre_r[1][j] = in && pair ? rho[(size_t)r1 * W + n] : 0.0f;
            pre_n[1][j] = in && pair ? la.net_out[(size_t)r1 * W + n] : 0.0f;
        }
    }
    LearnedGate gt{};
    if (la.net_out) {
        gt = learned_gate_wave(la, G);
        if (blockIdx.x == 0 && lane == 0) {
            la.result[0] = gt.ratio;
            la.result[1] = gt.ok ? 0.0 : 1.0;
            la.result[2] = gt.scale;
            la.result[3] = gt.q1;
            if (!gt.ok && la.fallback_count) *la.fallback_count += 1;
        }
    }
    // rho of node i (plain value pr, learned output nv), replaced first by the
    // learned density when one is deferred
    auto node_v = [&](size_t i, float pr, float nv) -> float {
        float v = pr;
        if (la.net_out) {
            if (gt.ok) {
                const double lv = gt.mean * (1.0 + (double)nv * la.inv_scale) * gt.scale;
                la.out64[i] = lv;
                v = (float)lv;
            } else {
                la.out64[i] = (double)v;  // fallback: the plain deposit stays
            }
            if (la.bg != 0.0f) v += la.bg;
            rho[i] = v;
        }
        return v;
    };
    auto node = [&](size_t i) -> float { return node_v(i, rho[i], la.net_out ? la.net_out[i] : 0.0f); };
    if (prefetch) {
#pragma unroll
        for (int j = 0; j < kPre; ++j) {
            const int n = lane + 64 * j;
            if (n < W) {
                const float u =
                    (float)(((double)node_v((size_t)r0 * W + n, pre_r[0][j], pre_n[0][j]) + bg) * cv - mean);
                const float v =
                    pair ? (float)(((double)node_v((size_t)r1 * W + n, pre_r[1][j], pre_n[1][j]) + bg) * cv - mean)
                         : 0.0f;
                a[n] = make_float2(u, v);
                t[n] = tw[n];
            }
        }
    } else {
        for (int n = lane; n < W; n += 64) {
            const float u = (float)(((double)node((size_t)r0 * W + n) + bg) * cv - mean);
            const float v = pair ? (float)(((double)node((size_t)r1 * W + n) + bg) * cv - mean) : 0.0f;
            a[n] = make_float2(u, v);
            t[n] = tw[n];
        }
    }
    __syncthreads();
    const float2* Z = wave_fft<-1, false>(a, b, W, t);
    const int hx = W / 2 + 1;
    for (int k = lane; k < hx; k += 64) {
        const float2 p = Z[k], m = Z[(W - k) & (W - 1)];
        // A = (Z[k] + conj Z[-k]) / 2, B = (Z[k] - conj Z[-k]) / (2i)
        spec[(size_t)r0 * hx + k] = make_float2(0.5f * (p.x + m.x), 0.5f * (p.y - m.y));
        if (pair) spec[(size_t)r1 * hx + k] = make_float2(0.5f * (p.y + m.y), -0.5f * (p.x - m.x));
    }
}

// Row pass of a learned step that runs the network's output layer (32 -> 1
// channels, 3x3, dilation la.dil4) itself: one workgroup of 16 waves per pair
// of rows.  The waves compute the output for the two rows on MFMA (a tile of
// 16 pixels per instruction group, nine taps of K = 32 channels) and store it
// (la.net_out); the per-block FP64 gate sums are added in a fixed order; wave
// 0 transforms the plain rows exactly as poisson_rows_kernel does (the same
// pairing, so a fallback step is bit for bit the plain solve) and wave 1 the
// learned rows' fluctuation mean * out / out_scale (the renormalising scale,
// a real factor, is applied by the column pass; the mean only moves the zero
// mode, which the solve drops).
constexpr int kRowsNetThreads = 1024;

__global__ __launch_bounds__(kRowsNetThreads) void poisson_rows_net_kernel(
    const float* __restrict__ rho, float2* __restrict__ spec, int W, int H, int64_t G, double bg, double cv,
    const double* __restrict__ rho_sum, double sum_scale, double sum_add, const float2* __restrict__ tw,
    LearnedApply la) {
    extern __shared__ float2 sm[];
    float2* a = sm;           // plain rows r0 + i r1
    float2* b = sm + W;       // scratch
    float2* l = sm + 2 * W;   // learned rows
    float2* lb = sm + 3 * W;  // scratch
    float2* t = sm + 4 * W;
    float* orow = reinterpret_cast<float*>(sm + 5 * W);  // [2][W] network output
    __shared__ double red[2][kRowsNetThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int grp = lane >> 4, col = lane & 15;
    const int r0 = 2 * blockIdx.x, r1 = r0 + 1;
    const bool pair = r1 < H;
    // plain values of this thread's column (W <= kRowsNetThreads)
    const bool has = tid < W;
    const float p0 = has ? rho[(size_t)r0 * W + tid] : 0.0f;
    const float p1 = has && pair ? rho[(size_t)r1 * W + tid] : 0.0f;
    // output layer, two 16-pixel tiles per pass of a wave
    const bf16x8* wpk = static_cast<const bf16x8*>(la.w4);
    const __bf16* act = static_cast<const __bf16*>(la.a3);
    bf16x8 bw[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) bw[k] = wpk[k * 64 + lane];
    const float bias = *la.bias4;
    const int tpr = W / 16;
    const int ntile = (pair ? 2 : 1) * tpr;
    constexpr int nw = kRowsNetThreads / 64;
    for (int tb = wv; tb < ntile; tb += 2 * nw) {
        bf16x8 av[2][9];
        int yt[2], xt[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int tile = min(tb + u * nw, ntile - 1);  // a missing second tile repeats the last (not stored)
            const int rl = tile / tpr;
            yt[u] = r0 + rl;
            xt[u] = (tile - rl * tpr) * 16;
            const int x = xt[u] + col;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const int yy = wrapn(yt[u] + (k / 3 - 1) * la.dil4, H);
                const int xx = wrapn(x + (k % 3 - 1) * la.dil4, W);
                av[u][k] = *reinterpret_cast<const bf16x8*>(act + ((size_t)yy * W + xx) * 32 + grp * 8);
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            if (tb + u * nw >= ntile) break;  // wave-uniform
            f32x4 acc{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int k = 0; k < 9; ++k) acc = mfma(av[u][k], bw[k], acc);
            if (col == 0) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int xo = xt[u] + grp * 4 + i;
                    const float o = acc[i] + bias;
                    orow[(yt[u] - r0) * W + xo] = o;
                    la.net_out[(size_t)yt[u] * W + xo] = o;
                }
            }
        }
    }
    __syncthreads();
    // gate sums (u = 1 + out / out_scale, (mean u - rho_plain)^2) and the
    // transform inputs
    const double mean = (*rho_sum * sum_scale + sum_add + bg * (double)G) * cv / (double)G;
    const double mean_l = *la.sum * la.sum_scale / (double)G;
    double su = 0.0, sr = 0.0;
    if (has) {
        const float o0 = orow[tid];
        const double u0 = 1.0 + (double)o0 * la.inv_scale;
        const double e0 = mean_l * u0 - (double)p0;
        su += u0;
        sr += e0 * e0;
        float v0 = p0, v1 = p1;
        if (la.bg != 0.0f) {
            v0 += la.bg;
            v1 += la.bg;
        }
        float lv1 = 0.0f;
        if (pair) {
            const float o1 = orow[W + tid];
            const double u1 = 1.0 + (double)o1 * la.inv_scale;
            const double e1 = mean_l * u1 - (double)p1;
            su += u1;
            sr += e1 * e1;
            lv1 = (float)(mean_l * la.inv_scale * cv * (double)o1);
        }
        a[tid] = make_float2((float)(((double)v0 + bg) * cv - mean),
                             pair ? (float)(((double)v1 + bg) * cv - mean) : 0.0f);
        l[tid] = make_float2((float)(mean_l * la.inv_scale * cv * (double)o0), lv1);
        t[tid] = tw[tid];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        su += __shfl_xor(su, o, 64);
        sr += __shfl_xor(sr, o, 64);
    }
    if (lane == 0) {
        red[0][wv] = su;
        red[1][wv] = sr;
    }
    __syncthreads();
    if (tid == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int w = 0; w < nw; ++w) {
            s0 += red[0][w];
            s1 += red[1][w];
        }
        la.stats[2 * blockIdx.x] = s0;
        la.stats[2 * blockIdx.x + 1] = s1;
    }
    const int hx = W / 2 + 1;
    if (wv < 2) {  // wave-uniform: the plain pair on wave 0, the learned pair on wave 1
        float2* src = wv == 0 ? a : l;
        float2* dst = wv == 0 ? spec : la.spec_l;
        const float2* Z = wave_fft<-1, false>(src, wv == 0 ? b : lb, W, t);
        for (int k = lane; k < hx; k += 64) {
            const float2 p = Z[k], m = Z[(W - k) & (W - 1)];
            dst[(size_t)r0 * hx + k] = make_float2(0.5f * (p.x + m.x), 0.5f * (p.y - m.y));
            if (pair) dst[(size_t)r1 * hx + k] = make_float2(0.5f * (p.y + m.y), -0.5f * (p.x - m.x));
        }
    }
}

// Three wavefronts per column: the forward transform and the Green's function
// by all of them, then one inverse transform each (phi, Ex, Ey) side by side.
constexpr int kColsThreads = 192;

// With the output layer in the row pass (la.a3), every wave takes the gate
// decision (fixed-order sums: all blocks agree), block 0 records it, and the
// column is read from the learned spectrum times the renormalising scale, or
// from the plain one when the gate falls back.
__global__ __launch_bounds__(kColsThreads) void poisson_cols_kernel(float2* __restrict__ spec,
                                                                    float2* __restrict__ sex,
                                                                    float2* __restrict__ sey, PoissonTables tb,
                                                                    DevDomain d, const float2* __restrict__ tw,
                                                                    LearnedApply la) {
    extern __shared__ float2 sm[];
    const int H = d.Ny;
    float2* a = sm;
    float2* b = sm + H;
    float2* ex = sm + 2 * H;
    float2* ey = sm + 3 * H;
    float2* s1 = sm + 4 * H;
    float2* s2 = sm + 5 * H;
    float2* t = sm + 6 * H;
    const int tid = threadIdx.x;
    const int hx = d.Nx / 2 + 1;
    const int kx = blockIdx.x;
    float2* F;
    if (la.a3) {
        // Both columns are transformed (one wave each) while the third wave
        // takes the gate decision, then the chosen transform is kept, the
        // learned one times the renormalising scale (the transform is linear;
        // a fallback keeps the plain transform, formed exactly as without a
        // learned density).
        __shared__ int gate_ok;
        __shared__ float gate_scale;
        for (int y = tid; y < H; y += kColsThreads) {
            a[y] = spec[(size_t)y * hx + kx];
            s1[y] = la.spec_l[(size_t)y * hx + kx];
            t[y] = tw[y];
        }
        __syncthreads();
        const int wv = tid >> 6;  // wave-uniform
        float2* fp = nullptr;
        float2* fl = nullptr;
        if (wv == 0) {
            fp = wave_fft<-1, false>(a, b, H, t);
        } else if (wv == 1) {
            fl = wave_fft<-1, false>(s1, s2, H, t);
        } else {
            const LearnedGate gt = learned_gate_wave(la, (int64_t)d.Nx * d.Ny);
            if ((tid & 63) == 0) {
                gate_ok = gt.ok ? 1 : 0;
                gate_scale = (float)gt.scale;
                if (blockIdx.x == 0) {
                    la.result[0] = gt.ratio;
                    la.result[1] = gt.ok ? 0.0 : 1.0;
                    la.result[2] = gt.scale;
                    la.result[3] = gt.q1;
                    if (!gt.ok && la.fallback_count) *la.fallback_count += 1;
                }
            }
        }
        // every wave learns where the transforms landed (both took the same stages)
        if (wv != 0) fp = wave_fft_result(a, b, H);
        if (wv != 1) fl = wave_fft_result(s1, s2, H);
        __syncthreads();
        F = fp;
        if (gate_ok) {
            const float sc = gate_scale;
            for (int y = tid; y < H; y += kColsThreads) F[y] = make_float2(fl[y].x * sc, fl[y].y * sc);
        }
        __syncthreads();
    } else {
        for (int y = tid; y < H; y += kColsThreads) {
            a[y] = spec[(size_t)y * hx + kx];
            t[y] = tw[y];
        }
        __syncthreads();
        F = wave_fft<-1>(a, b, H, t);
    }
    float2* S = F == a ? b : a;  // free buffer
    const float2 sx = tb.dx_sym[kx];
    for (int ky = tid; ky < H; ky += kColsThreads) {
        float2 p = make_float2(0.0f, 0.0f);
        if (kx != 0 || ky != 0) {
            const float f = (float)(1.0 / ((tb.k2x[kx] + tb.k2y[ky]) * (double)d.Nx * (double)d.Ny));
            const float2 v = F[ky];
            p = make_float2(v.x * f, v.y * f);
        }
        F[ky] = p;
        const float2 sy = tb.dy_sym[ky];
        ex[ky] = make_float2(-(sx.x * p.x - sx.y * p.y), -(sx.x * p.y + sx.y * p.x));
        ey[ky] = make_float2(-(sy.x * p.x - sy.y * p.y), -(sy.x * p.y + sy.y * p.x));
    }
    __syncthreads();
    const int wv = tid >> 6, lane = tid & 63;  // wave-uniform branches below
    if (wv == 0) {
        const float2* P = wave_fft<1, false>(F, S, H, t);
        for (int y = lane; y < H; y += 64) spec[(size_t)y * hx + kx] = P[y];
    } else if (wv == 1) {
        const float2* X = wave_fft<1, false>(ex, s1, H, t);
        for (int y = lane; y < H; y += 64) sex[(size_t)y * hx + kx] = X[y];
    } else {
        const float2* Y = wave_fft<1, false>(ey, s2, H, t);
        for (int y = lane; y < H; y += 64) sey[(size_t)y * hx + kx] = Y[y];
    }
}

// With la.a3 set it also writes the row's density as the gate decided (the
// column pass's record): the learned value mean (1 + out / out_scale) scale in
// FP64 to la.out64 and, rounded, to rho, or the plain value; plus la.bg.
__global__ __launch_bounds__(64) void poisson_rows_inv_kernel(const float2* __restrict__ X,
                                                              const float2* __restrict__ Y, float* __restrict__ ox,
                                                              float* __restrict__ oy, int W,
                                                              const float2* __restrict__ tw, float* __restrict__ rho,
                                                              LearnedApply la) {
    extern __shared__ float2 sm[];
    float2* a = sm;
    float2* b = sm + W;
    float2* t = sm + 2 * W;
    const int lane = threadIdx.x;
    const int hx = W / 2 + 1;
    const size_t row = blockIdx.x;
    const float2* xr = X + row * hx;
    const float2* yr = Y ? Y + row * hx : nullptr;
    for (int k = lane; k < hx; k += 64) {
        const float2 u = xr[k];
        const float2 v = yr ? yr[k] : make_float2(0.0f, 0.0f);
        if (k == 0 || k == W / 2) {
            a[k] = make_float2(u.x, v.x);  // DC and Nyquist: real parts only
        } else {
            a[k] = make_float2(u.x - v.y, u.y + v.x);           // X + iY
            a[W - k] = make_float2(u.x + v.y, -u.y + v.x);      // conj X + i conj Y
        }
    }
    for (int n = lane; n < W; n += 64) t[n] = tw[n];
    __syncthreads();
    // The learned step's density: the row's plain values and network output
    // are requested before the transform, which their loads then overlap.
    constexpr int kPre = 16;  // W <= 1024
    float pr[kPre], po[kPre];
    double ok_d = 0.0, scale = 0.0, lsum = 0.0;
    if (la.a3) {
        ok_d = la.result[1];
        scale = la.result[2];
        lsum = *la.sum;
#pragma unroll
        for (int j = 0; j < kPre; ++j) {
            const int n = lane + 64 * j;
            pr[j] = n < W ? rho[row * W + n] : 0.0f;
            po[j] = n < W ? la.net_out[row * W + n] : 0.0f;
        }
    }
    const float2* z = wave_fft<1, false>(a, b, W, t);
    for (int n = lane; n < W; n += 64) {
        const float2 v = z[n];
        ox[row * W + n] = v.x;
        if (oy) oy[row * W + n] = v.y;
    }
    if (la.a3) {
        const bool ok = ok_d == 0.0;
        const double mean = lsum * la.sum_scale / ((double)W * (double)gridDim.x);
#pragma unroll
        for (int j = 0; j < kPre; ++j) {
            const int n = lane + 64 * j;
            if (n < W) {
                const size_t i = row * W + n;
                float v = pr[j];
                if (ok) {
                    const double lv = mean * (1.0 + (double)po[j] * la.inv_scale) * scale;
                    la.out64[i] = lv;
                    v = (float)lv;
                } else {
                    la.out64[i] = (double)v;  // fallback: the plain deposit stays
                }
                if (la.bg != 0.0f) v += la.bg;
                rho[i] = v;
            }
        }
    }
}

}  // namespace

bool poisson_lds_supported(int Nx, int Ny) {
    auto pow2 = [](int v, int lo) { return v >= lo && v <= 1024 && (v & (v - 1)) == 0; };
    return pow2(Nx, 4) && pow2(Ny, 1);
}

void launch_poisson_lds(float* rho, int64_t G, double bg, double cv, const double* rho_sum, double sum_scale,
                        double sum_add, float2* spec, float2* spec_ex, float2* spec_ey, const PoissonTables& t,
                        const DevDomain& d, const float2* tw_x, const float2* tw_y, hipStream_t s,
                        const LearnedApply& la) {
    const int W = d.Nx, H = d.Ny;
    if (la.a3)
        hipLaunchKernelGGL(poisson_rows_net_kernel, dim3((unsigned)((H + 1) / 2)), dim3(kRowsNetThreads),
                           (size_t)5 * W * sizeof(float2) + (size_t)2 * W * sizeof(float), s, rho, spec, W, H, G, bg,
                           cv, rho_sum, sum_scale, sum_add, tw_x, la);
    else
        hipLaunchKernelGGL(poisson_rows_kernel, dim3((unsigned)((H + 1) / 2)), dim3(64),
                           (size_t)3 * W * sizeof(float2), s, rho, spec, W, H, G, bg, cv, rho_sum, sum_scale, sum_add,
                           tw_x, la);
    hipLaunchKernelGGL(poisson_cols_kernel, dim3((unsigned)(W / 2 + 1)), dim3(kColsThreads),
                       (size_t)7 * H * sizeof(float2), s, spec, spec_ex, spec_ey, t, d, tw_y, la);
}

bool poisson_rows_can_run_output_layer(int Nx, int Ny, int dil) {
    return poisson_lds_supported(Nx, Ny) && Nx % 16 == 0 && Nx <= kRowsNetThreads && dil < Nx && dil < Ny;
}

void launch_poisson_lds_rows(const float2* X, const float2* Y, float* ox, float* oy, const DevDomain& d,
                             const float2* tw_x, hipStream_t s, float* rho, const LearnedApply& la) {
    hipLaunchKernelGGL(poisson_rows_inv_kernel, dim3((unsigned)d.Ny), dim3(64), (size_t)3 * d.Nx * sizeof(float2), s,
                       X, Y, ox, oy, d.Nx, tw_x, rho, la);
}

}  // namespace pichip
