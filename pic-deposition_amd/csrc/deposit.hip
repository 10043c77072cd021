// Charge-deposition kernels for gfx950.
//
//  * deposit_atomic_kernel  - "GPU_Atomics": one particle per lane, FP64
//    hardware atomics into an accumulation grid that fold_acc_kernel rounds
//    into rho (and clears).  Works on any particle order.  Counterpart of the
//    reference depositKernelSoA (src/deposition/cuda/deposition_cuda.cu:205-210).
//  * deposit_tile_kernel    - "GPU_Priv": needs cell-sorted particles and the
//    cell offset table produced by the sort.  One wavefront owns one tile of
//    T x R cells and a private FP64 LDS window of (T+3+2D) x (R+3+2D) nodes
//    (stencil halo -1..+2 plus a drift halo of D cells).  Each lane streams K
//    consecutive particles with float4 loads, accumulates same-cell runs in
//    registers, the wave then combines runs that straddle lanes with a DPP
//    segmented scan, and only one lane per cell adds into LDS (one
//    non-returning FP64 LDS atomic per stencil value).  Particles whose
//    current cell is outside the window's reach (drift since the last sort)
//    take an overflow path with FP64 global atomics.
//  * deposit_merge_kernel   - writes rho node by node by adding, in FP64 and
//    in a fixed order, the windows of the tiles that cover it and the
//    accumulation grid.  Together with the per-wave windows this makes the
//    sorted path's result independent of scheduling.
// The reference's own privatised kernel (deposition_cuda.cu:212-241) scans all
// particles per tile; this design reads each particle once.

#include "deposit_device.hpp"

#include <algorithm>

namespace pichip {

namespace {

// A wave's particles that share a stencil origin and sit in consecutive lanes
// (cell-sorted input: a wave spans one to a few cells) are summed with a
// segmented DPP scan and added by the run's last lane: one FP64 atomic per
// node and run instead of one per particle, which on sorted input would all
// hit the same few nodes and serialise.  Waves with many short runs
// (unordered input) add particle by particle.
constexpr int kAtomicMaxRuns = 16;

template <int S>
__global__ __launch_bounds__(256) void deposit_atomic_kernel(const float* __restrict__ px, const float* __restrict__ py,
                                                             const float* __restrict__ pvx,
                                                             const float* __restrict__ pvy,
                                                             const float* __restrict__ pq, int64_t n, DevDomain d,
                                                             float dt, double* __restrict__ gacc,
                                                             int* __restrict__ status) {
    using St = Stencil<S>;
    constexpr int V = St::V;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;  // every lane stays for the wave-wide scan
    float x = 0.0f, y = 0.0f, vx = 0.0f, vy = 0.0f, qv = 0.0f;
    if (live) {
        x = px[i];
        y = py[i];
        if (St::kNeedsVelocity) {
            vx = pvx[i];
            vy = pvy[i];
        }
        qv = pq[i] * d.inv_vol;
    }
    float w[V];
    int cx = 0, cy = 0;
    const bool ok = live && St::weights(x, y, vx, vy, qv, dt, d, cx, cy, w);
    // Only the Esirkepov density can fail the fixed stencil (long moves).
    if (live && !ok) scatter_esk_long(x, y, vx, vy, qv, dt, d, gacc, status);
    const int key = ok ? wrap_index(cy, d.Ny) * d.Nx + wrap_index(cx, d.Nx) : -1;
    const bool cont = key >= 0 && key == dpp_i<kWaveShr1, 0xf>(-1, key);
    const int runs = __popcll(__ballot(key >= 0 && !cont));
    if (runs > kAtomicMaxRuns) {  // wave-uniform
        if (ok) {
#pragma unroll
            for (int k = 0; k < V; ++k)
                if (w[k] != 0.0f)
                    acc_add(gacc, (size_t)wrap_index(cy + St::OY + k / St::SX, d.Ny) * d.Nx +
                                      wrap_index(cx + St::OX + k % St::SX, d.Nx), w[k]);
        }
        return;
    }
    float z[V];
#pragma unroll
    for (int k = 0; k < V; ++k) z[k] = ok ? w[k] : 0.0f;
    wave_segmented_scan<V>(z, cont ? 1.0f : 0.0f);
    const int cont_next = dpp_i<kWaveShl1, 0xf>(0, cont ? 1 : 0);
    if (key >= 0 && !cont_next) {
#pragma unroll
        for (int k = 0; k < V; ++k)
            if (z[k] != 0.0f)
                acc_add(gacc, (size_t)wrap_index(cy + St::OY + k / St::SX, d.Ny) * d.Nx +
                                  wrap_index(cx + St::OX + k % St::SX, d.Nx), z[k]);
    }
}

// rho = FP32(acc); acc = 0 (closes the FP64 atomic sums).
__global__ __launch_bounds__(256) void fold_acc_kernel(float* __restrict__ rho, double* __restrict__ gacc, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        rho[i] = (float)gacc[i];
        gacc[i] = 0.0;
    }
}

template <int S, int K, int EX, int PIPE, int LA>
__global__ __launch_bounds__(256) void deposit_tile_kernel(const float* __restrict__ px, const float* __restrict__ py,
                                                           const float* __restrict__ pvx,
                                                           const float* __restrict__ pvy,
                                                           const float* __restrict__ pq,
                                                           const int* __restrict__ cell_off, DevDomain d, TileGeom g,
                                                           float dt, float* __restrict__ partial,
                                                           double* __restrict__ gacc, int* __restrict__ status) {
    extern __shared__ double smem_d[];
    float* smem = reinterpret_cast<float*>(smem_d);
    // wave-uniform (scalar registers: the tile's index arithmetic runs on the SALU)
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile = blockIdx.x * (blockDim.x >> 6) + wv;
    if (tile >= g.ntiles) return;  // wave-uniform
    deposit_one_tile<S, K, EX, PIPE, LA>(tile, smem + (size_t)wv * g.lds_per_wave, px, py, pvx, pvy, pq, cell_off, d,
                                         g, dt, partial, gacc, status);
}

template <int B>
__global__ __launch_bounds__(256) void deposit_merge_kernel(const float* __restrict__ partial,
                                                            double* __restrict__ gacc, DevDomain d, TileGeom g,
                                                            float* __restrict__ rho) {
    const int gx = blockIdx.x * blockDim.x + threadIdx.x;
    const int gy = blockIdx.y;
    if (gx < d.Nx) rho[(size_t)gy * d.Nx + gx] = merge_node<B>(gx, gy, partial, gacc, d, g);
}

__global__ __launch_bounds__(256) void fill_kernel(float* __restrict__ p, size_t n, float v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <int S>
void launch_atomic(const DepositArgs& a) {
    const int64_t blocks = (a.n + 255) / 256;
    if (a.ev_begin) (void)hipEventRecord(a.ev_begin, a.stream);
    if (a.n > 0)
        hipLaunchKernelGGL(deposit_atomic_kernel<S>, dim3((unsigned)blocks), dim3(256), 0, a.stream, a.x, a.y, a.vx,
                           a.vy, a.q, a.n, a.dom, a.dt, a.acc, a.status);
    if (a.ev_end) (void)hipEventRecord(a.ev_end, a.stream);
    const size_t G = (size_t)a.dom.Nx * a.dom.Ny;
    hipLaunchKernelGGL(fold_acc_kernel, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, a.stream, a.rho, a.acc, G);
}

}  // namespace

int moment_lds_per_wave(const TileGeom& g, int ppl) {
    const int during = 2 * Stencil<TSC_MOM>::V * g.T * g.R + 4 * g.R + 2 + 3 * mom_ring_cap(ppl);
    // The FP64 node window overlays the moments (and, once they are done
    // with, the row tables and the ring) for at most 64 home nodes, and
    // follows the ring otherwise.
    return g.T * g.R <= 64 ? std::max(during, 2 * g.wsize) : during + 2 * g.wsize;
}

bool tsc_uses_moments(const TileGeom& g, int ppl) {
    return g.D == 0 && (size_t)4 * moment_lds_per_wave(g, ppl) * sizeof(float) <= 64 * 1024;
}

namespace {

template <int S, int K, int EX, int PIPE, int LA>
void launch_tile(const DepositArgs& a) {
    TileGeom g = a.geom;
    if (S == TSC_MOM) g.lds_per_wave = moment_lds_per_wave(g, K);
    const int waves_per_block = 4;
    const unsigned blocks = (unsigned)((g.ntiles + waves_per_block - 1) / waves_per_block);
    const size_t lds = (size_t)waves_per_block * g.lds_per_wave * sizeof(float);
    if (a.ev_begin) (void)hipEventRecord(a.ev_begin, a.stream);
    hipLaunchKernelGGL((deposit_tile_kernel<S, K, EX, PIPE, LA>), dim3(blocks), dim3(64 * waves_per_block), lds, a.stream,
                       a.x, a.y, a.vx, a.vy, a.q, a.cell_off, a.dom, g, a.dt, a.partial, a.acc, a.status);
    if (a.ev_end) (void)hipEventRecord(a.ev_end, a.stream);
    if (a.skip_merge) return;
    const dim3 mgrid((unsigned)((a.dom.Nx + 255) / 256), (unsigned)a.dom.Ny);
    if (g.mbatch <= 4)
        hipLaunchKernelGGL(deposit_merge_kernel<4>, mgrid, dim3(256), 0, a.stream, a.partial, a.acc, a.dom, g, a.rho);
    else if (g.mbatch <= 8)
        hipLaunchKernelGGL(deposit_merge_kernel<8>, mgrid, dim3(256), 0, a.stream, a.partial, a.acc, a.dom, g, a.rho);
    else if (g.mbatch <= 16)
        hipLaunchKernelGGL(deposit_merge_kernel<16>, mgrid, dim3(256), 0, a.stream, a.partial, a.acc, a.dom, g, a.rho);
    else
        hipLaunchKernelGGL(deposit_merge_kernel<32>, mgrid, dim3(256), 0, a.stream, a.partial, a.acc, a.dom, g, a.rho);
}

// With a drift halo the particles have moved since the sort.  The lookahead
// variant (isolated strays go to the window alone and runs stay open; no
// register double-buffering) pays off when cells hold long runs; with a few
// particles per cell the plain chunk loop is faster (256^2, 8 per cell, TSC:
// 21.8 vs 23.2 us per deposit).
constexpr int64_t kLookaheadMinPerCell = 16;

template <int S, int K>
void launch_tile_kp(const DepositArgs& a) {
    const bool la = a.geom.D > 0 && a.n >= kLookaheadMinPerCell * (int64_t)a.dom.Nx * a.dom.Ny;
    const bool pipe = a.pipe && !la;
    if (a.dom.exact) {
        if (la) launch_tile<S, K, 1, 0, 1>(a);
        else if (pipe) launch_tile<S, K, 1, 1, 0>(a);
        else launch_tile<S, K, 1, 0, 0>(a);
    } else {
        if (la) launch_tile<S, K, 0, 0, 1>(a);
        else if (pipe) launch_tile<S, K, 0, 1, 0>(a);
        else launch_tile<S, K, 0, 0, 0>(a);
    }
}

template <int S>
void launch_tile_k(const DepositArgs& a) {
    if (a.ppl == 4) launch_tile_kp<S, 4>(a);
    else launch_tile_kp<S, 8>(a);
}

// TSC on the node-keyed order without a drift halo: the moment form (no
// lookahead or register double-buffering variants: the latter, 134 VGPRs,
// was 8-10% slower).
template <int K>
void launch_tile_moments_k(const DepositArgs& a) {
    if (a.pipe) {
        if (a.dom.exact) launch_tile<TSC_MOM, K, 1, 1, 0>(a);
        else launch_tile<TSC_MOM, K, 0, 1, 0>(a);
    } else {
        if (a.dom.exact) launch_tile<TSC_MOM, K, 1, 0, 0>(a);
        else launch_tile<TSC_MOM, K, 0, 0, 0>(a);
    }
}

void launch_tile_moments(const DepositArgs& a) {
    if (a.ppl == 4) launch_tile_moments_k<4>(a);
    else launch_tile_moments_k<8>(a);
}

// Bernoulli thinning by hash: particle i (in the current order) is kept when
// the top 32 bits of splitmix64(i, salt) fall below 2^32 / stride, i.e. with
// probability 1/stride, independently of where the particle sits.
__device__ __forceinline__ bool keep_particle(int64_t i, uint64_t salt, uint32_t threshold) {
    uint64_t z = (uint64_t)i + salt * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint32_t)(z >> 32) < threshold;
}

template <int S>
__global__ __launch_bounds__(256) void deposit_subsample_kernel(const float* __restrict__ px,
                                                                const float* __restrict__ py,
                                                                const float* __restrict__ pvx,
                                                                const float* __restrict__ pvy,
                                                                const float* __restrict__ pq, int64_t n, DevDomain d,
                                                                float dt, uint32_t threshold, uint64_t salt,
                                                                float scale, double* __restrict__ gacc,
                                                                int* __restrict__ status) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep_particle(i, salt, threshold)) return;
    const float vx = Stencil<S>::kNeedsVelocity ? pvx[i] : 0.0f;
    const float vy = Stencil<S>::kNeedsVelocity ? pvy[i] : 0.0f;
    scatter_global<S>(px[i], py[i], vx, vy, pq[i] * scale * d.inv_vol, dt, d, gacc, status);
}

template <int S>
void launch_subsample(const DepositArgs& a, uint32_t threshold, uint64_t salt, float scale) {
    if (a.n > 0)
        hipLaunchKernelGGL(deposit_subsample_kernel<S>, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, a.stream,
                           a.x, a.y, a.vx, a.vy, a.q, a.n, a.dom, a.dt, threshold, salt, scale, a.acc, a.status);
    const size_t G = (size_t)a.dom.Nx * a.dom.Ny;
    hipLaunchKernelGGL(fold_acc_kernel, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, a.stream, a.rho, a.acc, G);
}

}  // namespace

hipError_t launch_deposit_subsample(const DepositArgs& a, uint32_t stride, uint64_t salt, float scale) {
    const uint32_t threshold = (uint32_t)((1ull << 32) / stride);
    switch (a.scheme) {
        case NGP: launch_subsample<NGP>(a, threshold, salt, scale); break;
        case CIC: launch_subsample<CIC>(a, threshold, salt, scale); break;
        case TSC: launch_subsample<TSC>(a, threshold, salt, scale); break;
        default: launch_subsample<ESK>(a, threshold, salt, scale); break;
    }
    return hipGetLastError();
}

TileGeom make_tile_geom(int Nx, int Ny, int T, int R, int D) {
    TileGeom g;
    g.D = D < 0 ? 0 : (D > kMaxDriftHalo ? kMaxDriftHalo : D);
    g.T = T < 1 ? 1 : (T > Nx ? Nx : T);
    g.R = R < 1 ? 1 : (R > Ny ? Ny : R);
    if (g.T > 4096 - 2 * kMaxDriftHalo) g.T = 4096 - 2 * kMaxDriftHalo;  // window column fits 12 key bits
    if (g.R > 64) g.R = 64;
    g.ntx = (Nx + g.T - 1) / g.T;
    g.nty = (Ny + g.R - 1) / g.R;
    g.ntiles = g.ntx * g.nty;
    g.WX = g.T + 3 + 2 * g.D;
    g.WY = g.R + 3 + 2 * g.D;
    g.wsize = g.WX * g.WY;
    g.lds_per_wave = 2 * g.wsize + 4 * g.R + 2;  // FP64 window + row tables, even for 8-byte alignment
    const int rx = Nx % g.T, ry = Ny % g.R;
    g.fast = g.D == 0 && g.T >= 3 && g.R >= 3 && g.ntx >= 3 && g.nty >= 3 && (rx == 0 || rx >= 3) &&
             (ry == 0 || ry >= 3);
    // At most (T + 2 + 2D) / T + 1 consecutive tiles cover a node along an
    // axis; the range path needs them distinct and each window to hold one
    // image of the node.
    const bool ux = rx == 0 && g.WX <= Nx && (g.T + 2 + 2 * g.D) / g.T + 1 <= g.ntx;
    const bool uy = ry == 0 && g.WY <= Ny && (g.R + 2 + 2 * g.D) / g.R + 1 <= g.nty;
    g.uniform = ux && uy;
    g.fast_rows = !g.fast && g.D == 0 && g.T >= 3 && g.R < 3 && g.ntx >= 3 && (rx == 0 || rx >= 3) && uy;
    const int most = ((g.T + 2 + 2 * g.D) / g.T + 1) * ((g.R + 2 + 2 * g.D) / g.R + 1);
    g.mbatch = most <= 4 ? 4 : most <= 8 ? 8 : most <= 16 ? 16 : 32;
    return g;
}

void fill_async(float* p, size_t n, float v, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

hipError_t launch_deposit(const DepositArgs& a) {
    if (a.backend == 1) {
        switch (a.scheme) {
            case NGP: launch_atomic<NGP>(a); break;
            case CIC: launch_atomic<CIC>(a); break;
            case TSC: launch_atomic<TSC>(a); break;
            default: launch_atomic<ESK>(a); break;
        }
    } else {
        switch (a.scheme) {
            case NGP: launch_tile_k<NGP>(a); break;
            case CIC: launch_tile_k<CIC>(a); break;
            case TSC: launch_tile_k<TSC>(a); break;
            case NGP_NODE: launch_tile_k<NGP_NODE>(a); break;
            case TSC_NODE:
                if (tsc_uses_moments(a.geom, a.ppl))
                    launch_tile_moments(a);
                else
                    launch_tile_k<TSC_NODE>(a);
                break;
            default: launch_tile_k<ESK>(a); break;
        }
    }
    return hipGetLastError();
}

}  // namespace pichip
