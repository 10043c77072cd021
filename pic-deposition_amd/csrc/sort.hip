// Cell sort for gfx950: exact cell keys, stable LSD radix sort of (key, index)
// pairs with 8-bit digits, permutation of the particle streams, and the
// cell-offset table that the tiled deposit reads.
//
// Keys follow particleCellIndex (reference src/sorting/cell_sort.cpp:8-13) in
// FP64, so they are bit-identical to the reference's.  The reference orders
// with std::sort (cell_sort.cpp:28-33), whose order among equal keys is an
// introsort artefact; this sort is stable, i.e. equal to std::stable_sort.

#include "pic_common.hpp"
#include "pic_kernels.hpp"

namespace pichip {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 16;                       // items per thread per block
constexpr int kBlockItems = kThreads * kItems;  // 4096
constexpr int kDigits = 256;
constexpr int kWaves = kThreads / 64;

__global__ __launch_bounds__(256) void keys_kernel(const float* __restrict__ x, const float* __restrict__ y, int64_t n,
                                                   DevDomain d, int* __restrict__ keys, int node) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = node ? node_key_exact(x[i], y[i], d) : cell_key_exact(x[i], y[i], d);
}

// Per-block digit histogram.  Keys re-sorted between steps arrive nearly in
// order, so a wave mostly holds one or two digits: the first two distinct
// digits of each wave are counted with one LDS add each (ballot + popcount),
// the rest with one LDS atomic per lane.
template <int ITEMS>
__global__ __launch_bounds__(256) void radix_hist_kernel(const int* __restrict__ keys, int64_t n, int shift,
                                                         int nblocks, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kDigits];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t base = (int64_t)blockIdx.x * (kThreads * ITEMS);
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const int64_t i = base + (int64_t)j * kThreads + threadIdx.x;
        const bool valid = i < n;
        const unsigned dg = valid ? ((unsigned)keys[i] >> shift) & (kDigits - 1) : 0u;
        bool pending = valid;
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            const uint64_t act = __ballot(pending);
            if (act == 0) break;  // wave-uniform
            const int leader = __ffsll((unsigned long long)act) - 1;
            const unsigned d0 = (unsigned)__shfl((int)dg, leader, 64);
            const bool same = pending && dg == d0;
            const uint64_t m = __ballot(same);
            if (lane == leader) atomicAdd(&h[d0], (uint32_t)__popcll(m));
            if (same) pending = false;
        }
        if (pending) atomicAdd(&h[dg], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: items of a block are processed in ITEMS slots of 256
// consecutive items; within a slot, equal digits are ranked in thread order
// with wave ballots, across waves and slots with LDS counters.
template <int ITEMS>
__global__ __launch_bounds__(256) void radix_scatter_kernel(const int* __restrict__ keys_in,
                                                            const int* __restrict__ vals_in, int64_t n, int shift,
                                                            int nblocks, const uint32_t* __restrict__ hist_scan,
                                                            int* __restrict__ keys_out, int* __restrict__ vals_out) {
    __shared__ uint32_t gbase[kDigits];
    __shared__ uint32_t run[kDigits];
    __shared__ uint32_t cnt[kWaves][kDigits];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    gbase[t] = hist_scan[(size_t)t * nblocks + blockIdx.x];
    run[t] = 0;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t base = (int64_t)blockIdx.x * (kThreads * ITEMS);
    for (int j = 0; j < ITEMS; ++j) {
        const int64_t i = base + (int64_t)j * kThreads + t;
        const bool valid = i < n;
        int key = 0, val = 0;
        if (valid) {
            key = keys_in[i];
            val = vals_in ? vals_in[i] : (int)i;
        }
        const unsigned dg = ((unsigned)key >> shift) & (kDigits - 1);
#pragma unroll
        for (int ww = 0; ww < kWaves; ++ww) cnt[ww][t] = 0;
        __syncthreads();
        // Lanes of this wave holding the same digit (invalid lanes match nobody).
        uint64_t match = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t bal = __ballot(valid && ((dg >> b) & 1u));
            match &= ((dg >> b) & 1u) ? bal : ~bal;
        }
        const int wrank = __popcll(match & lt_mask);
        if (valid && wrank == 0) cnt[w][dg] = (uint32_t)__popcll(match);
        __syncthreads();
        {
            uint32_t acc = run[t];
#pragma unroll
            for (int ww = 0; ww < kWaves; ++ww) {
                const uint32_t c = cnt[ww][t];
                cnt[ww][t] = acc;
                acc += c;
            }
            run[t] = acc;
        }
        __syncthreads();
        if (valid) {
            const uint32_t pos = gbase[dg] + cnt[w][dg] + (uint32_t)wrank;
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
    }
}

// ---- exclusive scan of uint32 (3 kernels) ---------------------------------

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t& total) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(incl, o, 64);
        if (lane >= o) incl += u;
    }
    if (lane == 63) s_wave[w] = incl;
    __syncthreads();
    uint32_t wpre = 0, tot = 0;
#pragma unroll
    for (int ww = 0; ww < kWaves; ++ww) {
        const uint32_t c = s_wave[ww];
        if (ww < w) wpre += c;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return wpre + incl - v;
}

__global__ __launch_bounds__(256) void scan_reduce_kernel(const uint32_t* __restrict__ in, size_t m,
                                                          uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_wave[kWaves];
    const size_t base = (size_t)blockIdx.x * kBlockItems + (size_t)threadIdx.x * kItems;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j)
        if (base + j < m) s += in[base + j];
    uint32_t total;
    block_exclusive_scan(s, s_wave, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void scan_top_kernel(uint32_t* __restrict__ sums, size_t nb) {
    __shared__ uint32_t s_wave[kWaves];
    uint32_t carry = 0;
    for (size_t off = 0; off < nb; off += kThreads) {
        const size_t i = off + threadIdx.x;
        const uint32_t v = i < nb ? sums[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, s_wave, total);
        if (i < nb) sums[i] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(256) void scan_down_kernel(const uint32_t* __restrict__ in, size_t m,
                                                        const uint32_t* __restrict__ sums,
                                                        uint32_t* __restrict__ out) {
    __shared__ uint32_t s_wave[kWaves];
    const size_t base = (size_t)blockIdx.x * kBlockItems + (size_t)threadIdx.x * kItems;
    uint32_t v[kItems];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        v[j] = (base + j < m) ? in[base + j] : 0u;
        s += v[j];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, s_wave, total) + sums[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        if (base + j < m) out[base + j] = run;
        run += v[j];
    }
}

void exclusive_scan(const uint32_t* in, uint32_t* out, size_t m, uint32_t* tmp, hipStream_t s) {
    const size_t nb = (m + kBlockItems - 1) / kBlockItems;
    if (nb == 0) return;
    hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nb), dim3(kThreads), 0, s, in, m, tmp);
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(kThreads), 0, s, tmp, nb);
    hipLaunchKernelGGL(scan_down_kernel, dim3((unsigned)nb), dim3(kThreads), 0, s, in, m, tmp, out);
}

// off[c] = first sorted index with key >= c, off[G] = n.
__global__ __launch_bounds__(256) void cell_offsets_kernel(const int* __restrict__ keys, int64_t n, int G,
                                                           int* __restrict__ off) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    const int lo = (i == 0) ? -1 : keys[i - 1];
    const int hi = (i == n) ? G : keys[i];
    for (int c = lo + 1; c <= hi; ++c) off[c] = (int)i;
}

__global__ __launch_bounds__(256) void gather_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                     const int* __restrict__ perm, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

// Keys of the current order, and per block the number of particles whose key
// moved by more than `far` from the key at the same position after the last
// sort (`keys` holds those on entry).  Many far moves mean the permutation
// will be far from the identity.
__global__ __launch_bounds__(256) void keys_far_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                       int64_t n, DevDomain d, int* __restrict__ keys, int far,
                                                       int node, unsigned* __restrict__ far_part) {
    __shared__ unsigned wave_far[4];
    unsigned cnt = 0;  // wave total, kept by every lane
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
        const int64_t i = i0 + threadIdx.x;
        bool f = false;
        if (i < n) {
            const int old = keys[i];
            const int k = node ? node_key_exact(x[i], y[i], d) : cell_key_exact(x[i], y[i], d);
            keys[i] = k;
            const long long dk = (long long)k - (long long)old;
            f = dk > far || dk < -far;
        }
        cnt += (unsigned)__popcll(__ballot(f));
    }
    if ((threadIdx.x & 63) == 0) wave_far[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) far_part[blockIdx.x] = wave_far[0] + wave_far[1] + wave_far[2] + wave_far[3];
}

// Permutation mode of this sort: 1 (staged gather through packed records)
// when more than a quarter of the particles moved far, else 0.  One block
// sums the `nparts` per-block counts.
__global__ __launch_bounds__(256) void permute_mode_kernel(const unsigned* __restrict__ far_part, int nparts,
                                                           int64_t n, int* __restrict__ mode) {
    __shared__ unsigned long long sh[256];
    unsigned long long c = 0;
    for (int i = threadIdx.x; i < nparts; i += 256) c += far_part[i];
    sh[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *mode = (int64_t)sh[0] > n / 4 ? 1 : 0;
}

// Mode 1: the permuted streams of particle i packed into one or two float4
// records, so that the gather reads one or two lines per particle instead of
// one per stream.
__global__ __launch_bounds__(256) void pack_streams_kernel(StreamSet ss, int64_t n, const int* __restrict__ mode,
                                                           float4* __restrict__ rec) {
    if (*mode == 0) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < kMaxStreams; ++k)
        if (k < ss.count) v[k] = ss.src[k][i];
    const int w = ss.count > 4 ? 2 : 1;
    rec[w * i] = make_float4(v[0], v[1], v[2], v[3]);
    if (w == 2) rec[w * i + 1] = make_float4(v[4], v[5], v[6], v[7]);
}

__global__ __launch_bounds__(256) void gather_unpack_kernel(StreamSet ss, const int* __restrict__ perm, int64_t n,
                                                            const int* __restrict__ mode,
                                                            const float4* __restrict__ rec) {
    if (*mode == 0) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t p = perm[i];
    const int w = ss.count > 4 ? 2 : 1;
    const float4 a = rec[w * p];
    const float4 b = w == 2 ? rec[w * p + 1] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int k = 0; k < kMaxStreams; ++k)
        if (k < ss.count) ss.dst[k][i] = v[k];
}

// dst_k[i] = src_k[perm[i]] for up to kMaxStreams streams: the permutation is
// read once for all of them.  With perm == nullptr it is a plain copy.  With mode
// set to 1 the staged gather does this sort and the kernel returns.
__global__ __launch_bounds__(256) void permute_streams_kernel(StreamSet ss, const int* __restrict__ perm, int64_t n,
                                                              const int* __restrict__ mode) {
    if (mode && *mode != 0) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t p = perm ? (int64_t)perm[i] : i;
    float v[kMaxStreams];
#pragma unroll
    for (int k = 0; k < kMaxStreams; ++k)
        if (k < ss.count) v[k] = ss.src[k][p];
#pragma unroll
    for (int k = 0; k < kMaxStreams; ++k)
        if (k < ss.count) ss.dst[k][i] = v[k];
}

int key_bits(int G) {
    int bits = 0;
    while (bits < 31 && (1ll << bits) < (long long)G) ++bits;
    return bits;
}

}  // namespace

// Radix blocks: 16 items per thread, or 4 for sets of up to kSmallSort
// particles, whose passes would otherwise run on a few hundred workgroups
// with 16 barrier-separated slots each.
constexpr int64_t kSmallSort = int64_t(1) << 21;
static int radix_items(int64_t n) { return n <= kSmallSort ? 4 : kItems; }

size_t sort_hist_elems(int64_t n) {
    const size_t large = (size_t)((n + kBlockItems - 1) / kBlockItems);
    const size_t small = (size_t)((std::min(n, kSmallSort) + kThreads * 4 - 1) / (kThreads * 4));
    return std::max<size_t>(std::max(large, small), 1) * kDigits;
}

size_t sort_scan_tmp_elems(size_t m) { return (m + kBlockItems - 1) / kBlockItems + 1; }

void launch_keys(const float* x, const float* y, int64_t n, const DevDomain& d, int* keys, hipStream_t s, bool node) {
    if (n > 0)
        hipLaunchKernelGGL(keys_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, n, d, keys,
                           node ? 1 : 0);
}

hipError_t launch_sort_precomputed(int64_t n, int key_range, SortBuffers& sb, hipStream_t s) {
    if (n > 0) {
        const int items = radix_items(n);
        const int nblocks = (int)((n + kThreads * items - 1) / (kThreads * items));
        const size_t m = (size_t)nblocks * kDigits;
        const int bits = key_bits(key_range);
        const int passes = bits == 0 ? 1 : (bits + 7) / 8;
        int* kin = sb.keys;
        int* kout = sb.keys_alt;
        int* vin = nullptr;  // identity permutation on the first pass
        int* vout = sb.perm;
        for (int p = 0; p < passes; ++p) {
            const int shift = 8 * p;
            if (items == 4)
                hipLaunchKernelGGL(radix_hist_kernel<4>, dim3(nblocks), dim3(kThreads), 0, s, kin, n, shift, nblocks,
                                   sb.hist);
            else
                hipLaunchKernelGGL(radix_hist_kernel<kItems>, dim3(nblocks), dim3(kThreads), 0, s, kin, n, shift,
                                   nblocks, sb.hist);
            exclusive_scan(sb.hist, sb.hist, m, sb.scan_tmp, s);
            if (items == 4)
                hipLaunchKernelGGL(radix_scatter_kernel<4>, dim3(nblocks), dim3(kThreads), 0, s, kin, vin, n, shift,
                                   nblocks, sb.hist, kout, vout);
            else
                hipLaunchKernelGGL(radix_scatter_kernel<kItems>, dim3(nblocks), dim3(kThreads), 0, s, kin, vin, n,
                                   shift, nblocks, sb.hist, kout, vout);
            // next pass reads what this one wrote
            int* kt = kin;
            kin = kout;
            kout = kt;
            vin = vout;
            vout = (vout == sb.perm) ? sb.perm_alt : sb.perm;
        }
        // Leave sorted keys in sb.keys and the permutation in sb.perm.
        if (kin != sb.keys) {
            int* t = sb.keys;
            sb.keys = sb.keys_alt;
            sb.keys_alt = t;
        }
        if (vin != sb.perm) {
            int* t = sb.perm;
            sb.perm = sb.perm_alt;
            sb.perm_alt = t;
        }
    }
    hipLaunchKernelGGL(cell_offsets_kernel, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0, s, sb.keys, n,
                       key_range, sb.cell_off);
    return hipGetLastError();
}

hipError_t launch_sort_keys_perm(const float* x, const float* y, int64_t n, const DevDomain& d, SortBuffers& sb,
                                 hipStream_t s) {
    launch_keys(x, y, n, d, sb.keys, s);
    return launch_sort_precomputed(n, d.Nx * d.Ny, sb, s);
}

void launch_permute_streams(const StreamSet& ss, const int* perm, int64_t n, hipStream_t s, const int* mode) {
    if (n > 0)
        hipLaunchKernelGGL(permute_streams_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ss, perm, n,
                           mode);
}

void launch_keys_far(const float* x, const float* y, int64_t n, const DevDomain& d, int* keys, int far, bool node,
                     unsigned* far_part, int* mode, hipStream_t s) {
    const int blocks = (int)std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), kFarParts);
    if (n > 0)
        hipLaunchKernelGGL(keys_far_kernel, dim3((unsigned)blocks), dim3(256), 0, s, x, y, n, d, keys, far,
                           node ? 1 : 0, far_part);
    hipLaunchKernelGGL(permute_mode_kernel, dim3(1), dim3(256), 0, s, far_part, n > 0 ? blocks : 0, n, mode);
}

void launch_pack_streams(const StreamSet& ss, int64_t n, const int* mode, float4* rec, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(pack_streams_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ss, n, mode, rec);
}

void launch_gather_unpack(const StreamSet& ss, const int* perm, int64_t n, const int* mode, const float4* rec,
                          hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(gather_unpack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ss, perm, n, mode,
                           rec);
}

void launch_gather_f32(const float* src, float* dst, const int* perm, int64_t n, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(gather_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, perm, n);
}

}  // namespace pichip
