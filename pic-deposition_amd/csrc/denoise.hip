// Learned density denoiser kernels for gfx950: a 4-layer periodic 3x3
// convolution network (1 -> 32 -> 32 -> 32 -> 1 channels, ReLU between layers,
// per-layer dilation) with BF16 weights and activations and FP32
// accumulation on the matrix cores (v_mfma_f32_16x16x32_bf16).
//
// Layouts: hidden activations and their gradients are BF16 NHWC [B][H][W][32];
// the network input, output and output gradient are FP32 [B][H][W].  A 3x3
// tap t = 3 * ty + tx reads pixel (y + (ty - 1) * dil, x + (tx - 1) * dil),
// wrapped periodically.
//
// (MFMA operand layout: cnn_device.hpp.)
//
//  * conv3x3_kernel: implicit GEMM with M = 16 pixels of one row per tile,
//    N = output channels, K = (tap, input channel) = 32 per MFMA.  For a
//    single input channel K is the 9 taps (padded to 32), so one MFMA covers
//    the whole 3x3 stencil.  The same kernel computes the input gradient
//    (a convolution with flipped, transposed weights) followed by the ReLU
//    mask of the layer below.
//  * wgrad_kernel: weight gradient as a GEMM over pixels, M = input channels
//    (or taps), N = output channels, K = 32 pixels of one row per MFMA; every
//    wave writes its own partial and a fixed-order FP64 reduction finishes it.
//  * inference (launch_denoise): the first layer normalises rho on load, the
//    last writes per-block FP64 sums for the gate, and denoise_apply_kernel
//    re-adds those sums in a fixed order in every block, decides the gate and
//    renormalises: five launches from deposited rho to learned rho.

#include "cnn_device.hpp"
#include "deposit_device.hpp"
#include "pic_common.hpp"
#include "pic_kernels.hpp"

namespace pichip {

namespace {


__device__ __forceinline__ int64_t pix(const CnnGeom& g, int b, int y, int x) {
    return ((int64_t)b * g.H + y) * g.W + x;
}

template <int N>
__device__ __forceinline__ void block_sum_d(double (&v)[N]) {
    __shared__ double red[N][256];
#pragma unroll
    for (int k = 0; k < N; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
#pragma unroll
            for (int k = 0; k < N; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + o];
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = red[k][0];
}

enum { kHidden = 0, kOutput = 1, kBackward = 2 };

// Optional on-the-fly normalisation of a single-channel FP32 input:
// x -> scale * (x / mean - 1), mean = *sum * sum_scale * inv_count.
struct InNorm {
    const double* sum;  // null: the input is used as stored
    double sum_scale, inv_count;
    double scale;
};

// Optional per-block FP64 sums over the output pixels of the last layer
// (kOutput): u = 1 + out * inv_scale and (mean * u - rho)^2, mean as in InNorm,
// into partial[2 * block + {0, 1}].
struct OutStats {
    double* partial;  // null: no sums
    const float* rho;
    const double* sum;
    double sum_scale, inv_count;
    double inv_scale;
};

// Optional clear of the deposit's FP64 accumulation grid by a hidden layer
// (after a fused merge read it): node p is zeroed with the output pixel p when
// the deposit marked the grid (as the merge would have after folding it).
struct AccClear {
    double* acc;  // null: nothing to clear
    const int* mark;
    int epoch;
};

template <int CIN, int NT, int MODE, int MT>
__global__ __launch_bounds__(256) void conv3x3_kernel(const void* __restrict__ in_, const bf16x8* __restrict__ wpk,
                                                      const float* __restrict__ bias, void* __restrict__ out_,
                                                      const __bf16* __restrict__ mask, CnnGeom g, int dil, InNorm nrm,
                                                      OutStats ost, AccClear clr) {
    const int lane = threadIdx.x & 63;
    const bool clear = MODE == kHidden && clr.acc != nullptr && (clr.mark == nullptr || *clr.mark == clr.epoch);
    // Single-channel input transform x * mul + add (identity without nrm.sum).
    float mul = 1.0f, add = 0.0f;
    const bool norm_in = CIN == 1 && nrm.sum != nullptr;
    if (norm_in) {
        const double mean = *nrm.sum * nrm.sum_scale * nrm.inv_count;
        mul = (float)(nrm.scale / mean);
        add = (float)(-nrm.scale);
    }
    const bool stats = MODE == kOutput && ost.partial != nullptr;
    double st_mean = 0.0, st_u = 0.0, st_r2 = 0.0;
    if (stats) st_mean = *ost.sum * ost.sum_scale * ost.inv_count;
    const int grp = lane >> 4, col = lane & 15;
    const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t nwaves = (int64_t)gridDim.x * 4;
    constexpr int NTAP = CIN == 32 ? 9 : 1;
    bf16x8 bw[NTAP][NT];
#pragma unroll
    for (int t = 0; t < NTAP; ++t)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) bw[t][nt] = wpk[(t * NT + nt) * 64 + lane];
    const int nseg = g.W / (16 * MT);
    const int64_t ntile = (int64_t)g.B * g.H * nseg;
    for (int64_t tile = wave; tile < ntile; tile += nwaves) {
        const int seg = (int)(tile % nseg);
        const int64_t by = tile / nseg;
        const int y = (int)(by % g.H), b = (int)(by / g.H);
        const int x0 = seg * 16 * MT;
        f32x4 acc[MT][NT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const int x = x0 + mt * 16 + col;  // A row: this lane's pixel
            if (CIN == 32) {
                const __bf16* in = static_cast<const __bf16*>(in_);
                bf16x8 a[9];
#pragma unroll
                for (int t = 0; t < 9; ++t) {
                    const int yy = wrapn(y + (t / 3 - 1) * dil, g.H);
                    const int xx = wrapn(x + (t % 3 - 1) * dil, g.W);
                    a[t] = *reinterpret_cast<const bf16x8*>(in + pix(g, b, yy, xx) * 32 + grp * 8);
                }
#pragma unroll
                for (int t = 0; t < 9; ++t)
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = mfma(a[t], bw[t][nt], acc[mt][nt]);
            } else {
                const float* in = static_cast<const float*>(in_);
                bf16x8 a;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int k = grp * 8 + j;  // tap
                    float v = 0.0f;
                    if (k < 9) {
                        const int yy = wrapn(y + (k / 3 - 1) * dil, g.H);
                        const int xx = wrapn(x + (k % 3 - 1) * dil, g.W);
                        v = in[pix(g, b, yy, xx)];
                        if (norm_in) v = fmaf(v, mul, add);
                    }
                    a[j] = (__bf16)v;
                }
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = mfma(a, bw[0][nt], acc[mt][nt]);
            }
        }
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int c = nt * 16 + col;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int x = x0 + mt * 16 + grp * 4 + i;
                    const int64_t p = pix(g, b, y, x);
                    const float v = acc[mt][nt][i];
                    if (MODE == kHidden) {
                        static_cast<__bf16*>(out_)[p * 32 + c] = (__bf16)fmaxf(v + bias[c], 0.0f);
                        if (clear && c == 0) clr.acc[p] = 0.0;
                    } else if (MODE == kOutput) {
                        if (c == 0) {
                            const float o = v + bias[0];
                            static_cast<float*>(out_)[p] = o;
                            if (stats) {
                                const double u = 1.0 + (double)o * ost.inv_scale;
                                const double r = st_mean * u - (double)ost.rho[p];
                                st_u += u;
                                st_r2 += r * r;
                            }
                        }
                    } else {
                        const float m = (float)mask[p * 32 + c];
                        static_cast<__bf16*>(out_)[p * 32 + c] = (__bf16)(m > 0.0f ? v : 0.0f);
                    }
                }
            }
        }
    }
    if (stats) {
        __shared__ double sh[4];
        const double su = block_sum256(st_u, sh);
        const double sr = block_sum256(st_r2, sh);
        if (threadIdx.x == 0) {
            ost.partial[2 * blockIdx.x] = su;
            ost.partial[2 * blockIdx.x + 1] = sr;
        }
    }
}

// First layer (one input channel, dilation 1) fused with the deposit's window
// merge, for the learned step loop.  Workgroup y merges rows y-1, y, y+1 of
// rho into LDS node by node exactly as deposit_merge_kernel adds them (the
// FP64 accumulation grid folded in but not cleared: the neighbouring
// workgroups read the same rows; the second layer clears it), one node per
// thread, writes row y of rho and computes the layer's output for row y from
// LDS.
constexpr int kMergeConvThreads = 768;  // one node of rows y-1, y, y+1 per thread at W = 256

template <int B>
__global__ __launch_bounds__(kMergeConvThreads) void merge_conv1_kernel(const float* __restrict__ partial,
                                                                        const double* __restrict__ gacc, DevDomain d,
                                                                        TileGeom tg, float* __restrict__ rho,
                                                                        const bf16x8* __restrict__ wpk,
                                                                        const float* __restrict__ bias,
                                                                        __bf16* __restrict__ out, InNorm nrm) {
    extern __shared__ float rows[];  // [3][W]: rows y-1, y, y+1
    const int W = d.Nx, H = d.Ny;
    const int y = blockIdx.x;
    const bool fold = tg.acc_mark == nullptr || *tg.acc_mark == tg.epoch;
    for (int i = threadIdx.x; i < 3 * W; i += kMergeConvThreads) {
        const int r = i / W, x = i - r * W;
        const int gy = wrapn(y - 1 + r, H);
        // the overflow grid's value is requested with the windows' (used when marked)
        const double o = gacc[(size_t)gy * W + x];
        double s = merge_windows<B>(x, gy, partial, d, tg);
        if (fold && o != 0.0) s += o;
        rows[i] = (float)s;
        if (r == 1) rho[(size_t)y * W + x] = (float)s;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, grp = lane >> 4, col = lane & 15;
    const double mean = *nrm.sum * nrm.sum_scale * nrm.inv_count;
    const float mul = (float)(nrm.scale / mean), add = (float)(-nrm.scale);
    const bf16x8 bw0 = wpk[lane], bw1 = wpk[64 + lane];
    for (int tile = threadIdx.x >> 6; tile < W / 16; tile += kMergeConvThreads / 64) {
        const int x = tile * 16 + col;  // A row: this lane's pixel
        bf16x8 a;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = grp * 8 + j;  // tap
            float v = 0.0f;
            if (k < 9) v = fmaf(rows[(k / 3) * W + wrapn(x + k % 3 - 1, W)], mul, add);
            a[j] = (__bf16)v;
        }
        const f32x4 z{0.0f, 0.0f, 0.0f, 0.0f};
        const f32x4 acc0 = mfma(a, bw0, z), acc1 = mfma(a, bw1, z);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const size_t p = (size_t)y * W + tile * 16 + grp * 4 + i;
            out[p * 32 + col] = (__bf16)fmaxf(acc0[i] + bias[col], 0.0f);
            out[p * 32 + 16 + col] = (__bf16)fmaxf(acc1[i] + bias[16 + col], 0.0f);
        }
    }
}

// Per-wave partial of dW[tap] = sum over pixels of act(p + off(tap)) (x) dz(p):
// rows = input channel (CI = 32) or tap (CI = 1), columns = output channel.
// Partial layout [slot][gridDim.y][32][32] (unused rows / columns unwritten).
template <int CI, int CO>
__global__ __launch_bounds__(256) void wgrad_kernel(const void* __restrict__ act_, const void* __restrict__ dz_,
                                                    CnnGeom g, int dil, float* __restrict__ partial) {
    const int lane = threadIdx.x & 63;
    const int grp = lane >> 4, col = lane & 15;
    const int tap = blockIdx.y;
    const int64_t slot = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t nslots = (int64_t)gridDim.x * 4;
    constexpr int MTI = CI == 32 ? 2 : 1;
    constexpr int NTO = CO == 32 ? 2 : 1;
    f32x4 acc[MTI][NTO];
#pragma unroll
    for (int mt = 0; mt < MTI; ++mt)
#pragma unroll
        for (int nt = 0; nt < NTO; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int nkr = g.W / 32;
    const int64_t nk = (int64_t)g.B * g.H * nkr;
    for (int64_t k = slot; k < nk; k += nslots) {
        const int xs = (int)(k % nkr);
        const int64_t by = k / nkr;
        const int y = (int)(by % g.H), b = (int)(by / g.H);
        const int xb = xs * 32 + grp * 8;
        bf16x8 bo[NTO];
        if (CO == 32) {
            const __bf16* dz = static_cast<const __bf16*>(dz_);
#pragma unroll
            for (int nt = 0; nt < NTO; ++nt)
#pragma unroll
                for (int j = 0; j < 8; ++j) bo[nt][j] = dz[pix(g, b, y, xb + j) * 32 + nt * 16 + col];
        } else {
            const float* dz = static_cast<const float*>(dz_);
#pragma unroll
            for (int j = 0; j < 8; ++j) bo[0][j] = (__bf16)(col == 0 ? dz[pix(g, b, y, xb + j)] : 0.0f);
        }
        bf16x8 ao[MTI];
        if (CI == 32) {
            const __bf16* act = static_cast<const __bf16*>(act_);
            const int yy = wrapn(y + (tap / 3 - 1) * dil, g.H);
            const int dx = (tap % 3 - 1) * dil;
#pragma unroll
            for (int mt = 0; mt < MTI; ++mt)
#pragma unroll
                for (int j = 0; j < 8; ++j) ao[mt][j] = act[pix(g, b, yy, wrapn(xb + j + dx, g.W)) * 32 + mt * 16 + col];
        } else {
            const float* act = static_cast<const float*>(act_);
            const int t = col;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                float v = 0.0f;
                if (t < 9) v = act[pix(g, b, wrapn(y + (t / 3 - 1) * dil, g.H), wrapn(xb + j + (t % 3 - 1) * dil, g.W))];
                ao[0][j] = (__bf16)v;
            }
        }
#pragma unroll
        for (int mt = 0; mt < MTI; ++mt)
#pragma unroll
            for (int nt = 0; nt < NTO; ++nt) acc[mt][nt] = mfma(ao[mt], bo[nt], acc[mt][nt]);
    }
    float* P = partial + (slot * gridDim.y + tap) * 1024;
#pragma unroll
    for (int mt = 0; mt < MTI; ++mt)
#pragma unroll
        for (int nt = 0; nt < NTO; ++nt)
#pragma unroll
            for (int i = 0; i < 4; ++i) P[(mt * 16 + grp * 4 + i) * 32 + nt * 16 + col] = acc[mt][nt][i];
}

// Fixed-order FP64 sum of the wave partials into the FP32 gradient vector.
// kind 0: 32 -> 32 layer at offset w (W[t][co][ci]); 1: first layer (W[co][t]);
// 2: last layer (W[t][ci]).
__global__ __launch_bounds__(256) void wgrad_reduce_kernel(const float* __restrict__ partial, int nslots, int ny,
                                                           int kind, float* __restrict__ grad) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    int t, row, colc, count;
    if (kind == 0) {
        count = 9 * 32 * 32;
        t = e / 1024;
        colc = (e / 32) % 32;  // co
        row = e % 32;          // ci
    } else if (kind == 1) {
        count = 32 * 9;
        colc = e / 9;  // co
        row = e % 9;   // tap
        t = 0;
    } else {
        count = 9 * 32;
        t = e / 32;
        row = e % 32;  // ci
        colc = 0;
    }
    if (e >= count) return;
    double s = 0.0;
    for (int k = 0; k < nslots; ++k) s += (double)partial[((int64_t)k * ny + t) * 1024 + row * 32 + colc];
    grad[e] = (float)s;
}

// Bias gradient: per-channel sums of dz (32-channel BF16 or 1-channel FP32).
template <int CO>
__global__ __launch_bounds__(256) void bias_grad_partial_kernel(const void* __restrict__ dz_, int64_t npix,
                                                                double* __restrict__ partial) {
    constexpr int R = 256 / CO;  // pixel lanes per block
    const int c = threadIdx.x % CO;
    const int r = threadIdx.x / CO;
    double s = 0.0;
    for (int64_t p = (int64_t)blockIdx.x * R + r; p < npix; p += (int64_t)gridDim.x * R) {
        if (CO == 32)
            s += (double)(float)static_cast<const __bf16*>(dz_)[p * 32 + c];
        else
            s += (double)static_cast<const float*>(dz_)[p];
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < CO) {
        double t = 0.0;
        for (int k = 0; k < R; ++k) t += red[k * CO + threadIdx.x];
        partial[(int64_t)blockIdx.x * CO + threadIdx.x] = t;
    }
}

__global__ void bias_grad_final_kernel(const double* __restrict__ partial, int nb, int co, float* __restrict__ grad) {
    const int c = threadIdx.x;
    if (c >= co) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += partial[(int64_t)b * co + c];
    grad[c] = (float)s;
}

// dz = 2 (out - target) / n, per-block FP64 sums of (out - target)^2.
__global__ __launch_bounds__(256) void loss_kernel(const float* __restrict__ out, const float* __restrict__ target,
                                                   int64_t n, float* __restrict__ dz, double* __restrict__ partial) {
    double v[1] = {0.0};
    const float sc = (float)(2.0 / (double)n);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float d = out[i] - target[i];
        dz[i] = sc * d;
        v[0] += (double)d * (double)d;
    }
    block_sum_d<1>(v);
    if (threadIdx.x == 0) partial[blockIdx.x] = v[0];
}

__global__ void sum_partials_kernel(const double* __restrict__ partial, int nb, double scale,
                                    double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += partial[b];
    *out = s * scale;
}

// Adam (Kingma & Ba 2015) with bias correction; FP32 master weights.
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ w, const float* __restrict__ grad,
                                                   float* __restrict__ m, float* __restrict__ v, int n, float lr,
                                                   float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 MFMA B operands of every conv launch from the FP32 parameters.
__global__ __launch_bounds__(256) void pack_kernel(const float* __restrict__ w, __bf16* __restrict__ pk) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;  // element = vec * 8 + j
    if (e >= cnn::kPackVecs * 8) return;
    const int vec = e / 8, j = e % 8;
    const int lane = vec % 64, grp = lane >> 4, col = lane & 15;
    const int k = grp * 8 + j;
    float v = 0.0f;
    if (vec < cnn::kPackF2) {  // first layer, forward: K = taps, N = co
        const int nt = vec / 64;
        if (k < 9) v = w[cnn::kW1 + (nt * 16 + col) * 9 + k];
    } else if (vec < cnn::kPackF4) {  // hidden layers, forward: W[t][co][ci]
        const int base = vec < cnn::kPackF3 ? cnn::kW2 : cnn::kW3;
        const int r = vec - (vec < cnn::kPackF3 ? cnn::kPackF2 : cnn::kPackF3);
        const int t = r / 128, nt = (r / 64) % 2;
        v = w[base + (t * 32 + nt * 16 + col) * 32 + k];
    } else if (vec < cnn::kPackB4) {  // last layer, forward: column 0 only
        const int t = (vec - cnn::kPackF4) / 64;
        if (col == 0) v = w[cnn::kW4 + t * 32 + k];
    } else if (vec < cnn::kPackB3) {  // last layer, input gradient: K = flipped taps, N = ci
        const int nt = (vec - cnn::kPackB4) / 64;
        if (k < 9) v = w[cnn::kW4 + (8 - k) * 32 + nt * 16 + col];
    } else {  // hidden layers, input gradient: new K = co, new N = ci, flipped taps
        const int base = vec < cnn::kPackB2 ? cnn::kW3 : cnn::kW2;
        const int r = vec - (vec < cnn::kPackB2 ? cnn::kPackB3 : cnn::kPackB2);
        const int t = r / 128, nt = (r / 64) % 2;
        v = w[base + ((8 - t) * 32 + k) * 32 + nt * 16 + col];
    }
    pk[e] = (__bf16)v;
}

// delta = scale * (rho / mean - 1) with mean = *rho_sum / n (FP64).
__global__ __launch_bounds__(256) void normalize_kernel(const float* __restrict__ rho, int64_t n,
                                                        const double* __restrict__ rho_sum, double scale,
                                                        float* __restrict__ delta) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double mean = *rho_sum / (double)n;
    delta[i] = (float)(scale * ((double)rho[i] / mean - 1.0));
}

// Training batch: image b = pair items[b].x rolled by (items[b].y, items[b].z),
// input and target multiplied by their scales.
__global__ __launch_bounds__(256) void batch_kernel(const float* __restrict__ inputs,
                                                    const float* __restrict__ targets, const int4* __restrict__ items,
                                                    CnnGeom g, float sa, float sb, float* __restrict__ in_b,
                                                    float* __restrict__ tg_b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t plane = (int64_t)g.H * g.W;
    if (i >= plane * g.B) return;
    const int b = (int)(i / plane);
    const int y = (int)((i % plane) / g.W), x = (int)(i % g.W);
    const int4 it = items[b];
    const int64_t src = (int64_t)it.x * plane + (int64_t)((y + it.y) % g.H) * g.W + (x + it.z) % g.W;
    in_b[i] = sa * inputs[src];
    tg_b[i] = sb * targets[src];
}

// Per-block FP64 sums of q and q^2 over particles.
__global__ __launch_bounds__(256) void charge_moments_kernel(const float* __restrict__ q, int64_t n,
                                                             double* __restrict__ partial) {
    double v[2] = {0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double a = q[i];
        v[0] += a;
        v[1] += a * a;
    }
    block_sum_d<2>(v);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = v[0];
        partial[2 * blockIdx.x + 1] = v[1];
    }
}

// Gate and renormalisation, one thread per node.  Every block first adds the
// per-block sums of the output layer (u, squared residual) and of the charge
// moments (q, q^2) in the same fixed order, so all blocks take the same
// decision: the residual ratio ||rho_learned - rho_plain||^2 / (kappa sum q^2 /
// cv^2) must not exceed tol, and the renormalisation scale makes sum(rho) * cv
// equal sum(q).  Block 0 writes result = {ratio, fallback, scale, sum q} and
// counts a fallback.
__global__ __launch_bounds__(256) void denoise_apply_kernel(const float* __restrict__ net_out,
                                                            float* __restrict__ rho, double* __restrict__ out64,
                                                            int64_t n, const double* __restrict__ sum,
                                                            double sum_scale, double inv_scale,
                                                            const double* __restrict__ stats, int nbs,
                                                            const double* __restrict__ moments, int nbm, double cv,
                                                            double kappa, double tol, double* __restrict__ result,
                                                            int* __restrict__ fallback_count) {
    __shared__ double sh[4];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nbs; b += blockDim.x) {
        a[0] += stats[2 * b];
        a[1] += stats[2 * b + 1];
    }
    for (int b = threadIdx.x; b < nbm; b += blockDim.x) {
        a[2] += moments[2 * b];
        a[3] += moments[2 * b + 1];
    }
    const double s1 = block_sum256(a[0], sh), s2 = block_sum256(a[1], sh);
    const double q1 = block_sum256(a[2], sh), q2 = block_sum256(a[3], sh);
    const double mean = *sum * sum_scale / (double)n;
    const double noise = kappa * q2 / (cv * cv);  // expected squared shot-noise norm of the plain deposit
    const double ratio = noise > 0.0 ? s2 / noise : INFINITY;
    const double scale = (q1 / cv) / (mean * s1);  // makes sum(rho_out) * cv == sum(q)
    const bool ok = mean != 0.0 && isfinite(ratio) && isfinite(scale) && ratio <= tol;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        result[0] = ratio;
        result[1] = ok ? 0.0 : 1.0;
        result[2] = scale;
        result[3] = q1;
        if (!ok && fallback_count) *fallback_count += 1;
    }
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!ok) {  // fallback: the plain deposit stays
        out64[i] = (double)rho[i];
        return;
    }
    const double v = mean * (1.0 + (double)net_out[i] * inv_scale) * scale;
    out64[i] = v;
    rho[i] = (float)v;
}

// Launches one layer; returns the number of workgroups.
template <int CIN, int NT, int MODE>
int conv_launch(const void* in, const void* wpk, const float* bias, void* out, const void* mask, const CnnGeom& g,
                int dil, int ncu, hipStream_t s, const InNorm& nrm = InNorm{}, const OutStats& ost = OutStats{},
                const AccClear& clr = AccClear{}) {
    const int64_t rowtiles16 = (int64_t)g.B * g.H * (g.W / 16);
    // Four 16-pixel tiles per wave when there is enough work.  Otherwise the
    // 32 -> 32 layers take two (every wave loads the layer's 18 KB of packed
    // weights, so two tiles halve that traffic: 8.8 -> 7.2 us at 256^2) and
    // the single-channel layers one (more, shorter waves are faster there).
    const bool wide = (g.W % 64) == 0 && rowtiles16 >= (int64_t)ncu * 64;
    const int mt = wide ? 4 : (CIN == 32 && NT == 2 && (g.W % 32) == 0) ? 2 : 1;
    const int64_t tiles = rowtiles16 / mt;
    const int64_t blocks = std::min<int64_t>((tiles + 3) / 4, (int64_t)ncu * 8);
    const dim3 grid((unsigned)std::max<int64_t>(blocks, 1));
    const auto* w = static_cast<const bf16x8*>(wpk);
    const auto* m = static_cast<const __bf16*>(mask);
    if (mt == 4)
        hipLaunchKernelGGL((conv3x3_kernel<CIN, NT, MODE, 4>), grid, dim3(256), 0, s, in, w, bias, out, m, g, dil, nrm,
                           ost, clr);
    else if (mt == 2)
        hipLaunchKernelGGL((conv3x3_kernel<CIN, NT, MODE, 2>), grid, dim3(256), 0, s, in, w, bias, out, m, g, dil, nrm,
                           ost, clr);
    else
        hipLaunchKernelGGL((conv3x3_kernel<CIN, NT, MODE, 1>), grid, dim3(256), 0, s, in, w, bias, out, m, g, dil, nrm,
                           ost, clr);
    return (int)grid.x;
}

}  // namespace

void launch_cnn_forward(const CnnBuffers& c, const CnnGeom& g, const int (&dil)[4], int ncu, hipStream_t s) {
    const __bf16* pk = reinterpret_cast<const __bf16*>(c.packed);
    conv_launch<1, 2, kHidden>(c.x0, pk + 8 * 64 * 0, c.w + cnn::kB1, c.a1, nullptr, g, dil[0], ncu, s);
    conv_launch<32, 2, kHidden>(c.a1, pk + 8 * cnn::kPackF2, c.w + cnn::kB2, c.a2, nullptr, g, dil[1], ncu, s);
    conv_launch<32, 2, kHidden>(c.a2, pk + 8 * cnn::kPackF3, c.w + cnn::kB3, c.a3, nullptr, g, dil[2], ncu, s);
    conv_launch<32, 1, kOutput>(c.a3, pk + 8 * cnn::kPackF4, c.w + cnn::kB4, c.out, nullptr, g, dil[3], ncu, s);
}

void launch_cnn_backward(const CnnBuffers& c, const CnnGeom& g, const int (&dil)[4], int ncu, hipStream_t s) {
    const __bf16* pk = reinterpret_cast<const __bf16*>(c.packed);
    const int64_t npix = (int64_t)g.B * g.H * g.W;
    const int wblocks = std::max(1, std::min(ncu, (int)((npix / 32 + 3) / 4)));
    const int nslots = wblocks * 4;
    const int bb = 256;
    auto bias_grad = [&](const void* dz, int co, float* out) {
        if (co == 32)
            hipLaunchKernelGGL(bias_grad_partial_kernel<32>, dim3(bb), dim3(256), 0, s, dz, npix, c.bias_part);
        else
            hipLaunchKernelGGL(bias_grad_partial_kernel<1>, dim3(bb), dim3(256), 0, s, dz, npix, c.bias_part);
        hipLaunchKernelGGL(bias_grad_final_kernel, dim3(1), dim3(64), 0, s, c.bias_part, bb, co, out);
    };
    // Layer 4 (32 -> 1): dz4 = c.dout (FP32).
    hipLaunchKernelGGL((wgrad_kernel<32, 1>), dim3(wblocks, 9), dim3(256), 0, s, c.a3, c.dout, g, dil[3], c.wpart);
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((288 + 255) / 256), dim3(256), 0, s, c.wpart, nslots, 9, 2,
                       c.grad + cnn::kW4);
    bias_grad(c.dout, 1, c.grad + cnn::kB4);
    conv_launch<1, 2, kBackward>(c.dout, pk + 8 * cnn::kPackB4, nullptr, c.d3, c.a3, g, dil[3], ncu, s);
    // Layer 3 (32 -> 32): dz3 = c.d3.
    hipLaunchKernelGGL((wgrad_kernel<32, 32>), dim3(wblocks, 9), dim3(256), 0, s, c.a2, c.d3, g, dil[2], c.wpart);
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(9216 / 256), dim3(256), 0, s, c.wpart, nslots, 9, 0,
                       c.grad + cnn::kW3);
    bias_grad(c.d3, 32, c.grad + cnn::kB3);
    conv_launch<32, 2, kBackward>(c.d3, pk + 8 * cnn::kPackB3, nullptr, c.d2, c.a2, g, dil[2], ncu, s);
    // Layer 2: dz2 = c.d2.
    hipLaunchKernelGGL((wgrad_kernel<32, 32>), dim3(wblocks, 9), dim3(256), 0, s, c.a1, c.d2, g, dil[1], c.wpart);
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(9216 / 256), dim3(256), 0, s, c.wpart, nslots, 9, 0,
                       c.grad + cnn::kW2);
    bias_grad(c.d2, 32, c.grad + cnn::kB2);
    conv_launch<32, 2, kBackward>(c.d2, pk + 8 * cnn::kPackB2, nullptr, c.d3, c.a1, g, dil[1], ncu, s);  // dz1 in d3
    // Layer 1 (1 -> 32): rows = taps.
    hipLaunchKernelGGL((wgrad_kernel<1, 32>), dim3(wblocks, 1), dim3(256), 0, s, c.x0, c.d3, g, dil[0], c.wpart);
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((288 + 255) / 256), dim3(256), 0, s, c.wpart, nslots, 1, 1,
                       c.grad + cnn::kW1);
    bias_grad(c.d3, 32, c.grad + cnn::kB1);
}

size_t cnn_wpart_floats(int ncu) { return (size_t)ncu * 4 * 9 * 1024; }

void launch_cnn_loss(const CnnBuffers& c, const CnnGeom& g, double* loss, hipStream_t s) {
    const int64_t n = (int64_t)g.B * g.H * g.W;
    const int nb = 256;
    hipLaunchKernelGGL(loss_kernel, dim3(nb), dim3(256), 0, s, c.out, c.target, n, c.dout, c.loss_part);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(64), 0, s, c.loss_part, nb, 1.0 / (double)n, loss);
}

void launch_cnn_adam(const CnnBuffers& c, float lr, float b1, float b2, float eps, int step, hipStream_t s) {
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((cnn::kParams + 255) / 256), dim3(256), 0, s, c.w, c.grad, c.m, c.v,
                       cnn::kParams, lr, b1, b2, eps, c1, c2);
}

void launch_cnn_pack(const CnnBuffers& c, hipStream_t s) {
    hipLaunchKernelGGL(pack_kernel, dim3((cnn::kPackVecs * 8 + 255) / 256), dim3(256), 0, s, c.w,
                       reinterpret_cast<__bf16*>(c.packed));
}

void launch_cnn_batch(const float* inputs, const float* targets, const int4* items, const CnnGeom& g, float sa,
                      float sb, float* in_b, float* tg_b, hipStream_t s) {
    const int64_t n = (int64_t)g.B * g.H * g.W;
    hipLaunchKernelGGL(batch_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, inputs, targets, items, g,
                       sa, sb, in_b, tg_b);
}

void launch_normalize(const float* rho, int64_t n, const double* rho_sum, double scale, float* delta,
                      hipStream_t s) {
    hipLaunchKernelGGL(normalize_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rho, n, rho_sum, scale,
                       delta);
}

void launch_moments(const float* v, int64_t n, double* partial, int nb, hipStream_t s) {
    hipLaunchKernelGGL(charge_moments_kernel, dim3(nb), dim3(256), 0, s, v, n, partial);
}

bool denoise_can_merge(const CnnGeom& g, const int (&dil)[4]) {
    return g.B == 1 && dil[0] == 1 && g.W % 16 == 0 && g.W <= 2048;
}

void launch_denoise(const CnnBuffers& c, const CnnGeom& g, const int (&dil)[4], int ncu, const DenoiseArgs& a,
                    hipStream_t s, LearnedApply* defer, const TileMerge* merge, float2* spec_l) {
    const __bf16* pk = reinterpret_cast<const __bf16*>(c.packed);
    const double inv_g = 1.0 / (double)a.G;
    const InNorm nrm{a.sum, a.sum_scale, inv_g, a.in_scale};
    const OutStats ost{a.stats, a.rho, a.sum, a.sum_scale, inv_g, a.inv_out_scale};
    AccClear clr{};
    if (merge) {
        const auto* w1 = reinterpret_cast<const bf16x8*>(pk);
        const dim3 grid((unsigned)g.H);
        const size_t lds = (size_t)3 * g.W * sizeof(float);
        const TileMerge& m = *merge;
        __bf16* a1 = static_cast<__bf16*>(c.a1);
        if (m.geom.mbatch <= 4)
            hipLaunchKernelGGL(merge_conv1_kernel<4>, grid, dim3(kMergeConvThreads), lds, s, m.partial, m.acc, m.dom, m.geom, a.rho,
                               w1, c.w + cnn::kB1, a1, nrm);
        else if (m.geom.mbatch <= 8)
            hipLaunchKernelGGL(merge_conv1_kernel<8>, grid, dim3(kMergeConvThreads), lds, s, m.partial, m.acc, m.dom,
                               m.geom, a.rho, w1, c.w + cnn::kB1, a1, nrm);
        else if (m.geom.mbatch <= 16)
            hipLaunchKernelGGL(merge_conv1_kernel<16>, grid, dim3(kMergeConvThreads), lds, s, m.partial, m.acc, m.dom,
                               m.geom, a.rho, w1, c.w + cnn::kB1, a1, nrm);
        else
            hipLaunchKernelGGL(merge_conv1_kernel<32>, grid, dim3(kMergeConvThreads), lds, s, m.partial, m.acc, m.dom,
                               m.geom, a.rho, w1, c.w + cnn::kB1, a1, nrm);
        clr = AccClear{m.acc, m.geom.acc_mark, m.geom.epoch};
    } else {
        conv_launch<1, 2, kHidden>(a.rho, pk + 8 * 64 * 0, c.w + cnn::kB1, c.a1, nullptr, g, dil[0], ncu, s, nrm);
    }
    conv_launch<32, 2, kHidden>(c.a1, pk + 8 * cnn::kPackF2, c.w + cnn::kB2, c.a2, nullptr, g, dil[1], ncu, s,
                                InNorm{}, OutStats{}, clr);
    conv_launch<32, 2, kHidden>(c.a2, pk + 8 * cnn::kPackF3, c.w + cnn::kB3, c.a3, nullptr, g, dil[2], ncu, s);
    if (defer && spec_l) {  // the output layer runs in the solve's row pass
        LearnedApply la{c.out, a.out64, a.sum, a.sum_scale, a.inv_out_scale, a.stats, (g.H + 1) / 2, a.moments,
                        a.nbm, a.cv, a.kappa, a.tol, a.result, a.fallback_count, 0.0f};
        la.a3 = c.a3;
        la.w4 = pk + 8 * cnn::kPackF4;
        la.bias4 = c.w + cnn::kB4;
        la.dil4 = dil[3];
        la.spec_l = spec_l;
        *defer = la;
        return;
    }
    const int nbs =
        conv_launch<32, 1, kOutput>(c.a3, pk + 8 * cnn::kPackF4, c.w + cnn::kB4, c.out, nullptr, g, dil[3], ncu, s,
                                    InNorm{}, ost);
    if (defer) {
        *defer = LearnedApply{c.out, a.out64, a.sum, a.sum_scale, a.inv_out_scale, a.stats, nbs, a.moments, a.nbm,
                              a.cv, a.kappa, a.tol, a.result, a.fallback_count, 0.0f};
        return;
    }
    hipLaunchKernelGGL(denoise_apply_kernel, dim3((unsigned)((a.G + 255) / 256)), dim3(256), 0, s, c.out, a.rho,
                       a.out64, a.G, a.sum, a.sum_scale, a.inv_out_scale, a.stats, nbs, a.moments, a.nbm, a.cv,
                       a.kappa, a.tol, a.result, a.fallback_count);
}

size_t denoise_stats_doubles(int ncu) { return (size_t)2 * 8 * std::max(ncu, 1); }

}  // namespace pichip
