// Cell sort for gfx950: exact cell keys, a stable LSD radix sort with digits
// of up to 9 bits (three passes for 2^20 cells, two for 2^18), and the
// cell-offset table that the tiled deposit reads.
//
// Keys follow particleCellIndex (reference src/sorting/cell_sort.cpp:8-13) in
// FP64, so they are bit-identical to the reference's.  The reference orders
// with std::sort (cell_sort.cpp:28-33), whose order among equal keys is an
// introsort artefact; this sort is stable, i.e. equal to std::stable_sort.
//
// One radix pass = a histogram kernel (per block and digit), an exclusive
// scan of that table (digit-major), and a scatter kernel.  A scatter block of
// 4096 items ranks its items stably in LDS (each wave ranks a contiguous
// part of the block in item order against a wave-private digit histogram; a
// block-wide scan then combines the waves), stages them in digit order in
// LDS and writes them out so that consecutive threads write consecutive
// addresses of a digit's run.  The context's sort (rrec_kernel) moves the
// particle streams themselves through the passes, one stream at a time
// through the same stage; the (key, index) form (rscatter_kernel) serves the
// permutation output of pic_hip_sort_perm and sets below 2^22 particles,
// whose streams are then gathered once.  Blocks map to item ranges
// XCD-contiguously (blocks are dispatched round-robin over the 8 XCDs; block
// b takes range (b mod 8) * q + b / 8), so the runs of neighbouring ranges,
// which are adjacent in the output, are written through the same L2.

#include "pic_common.hpp"
#include "pic_kernels.hpp"

namespace pichip {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kBlockItems = kThreads * kItems;  // 4096 (scan blocks)
constexpr int kWaves = kThreads / 64;

constexpr int kRT = 256;        // radix block threads
constexpr int kRW = kRT / 64;   // waves per radix block
constexpr int kMaxBins = 1024;     // LDS histogram size (digits of up to 10 bits)
constexpr int kMaxDigitBits = 9;   // digit width the passes use at most
constexpr int kXcds = 8;        // MI355X: workgroups are dispatched round-robin over 8 XCDs

__global__ __launch_bounds__(256) void keys_kernel(const float* __restrict__ x, const float* __restrict__ y, int64_t n,
                                                   DevDomain d, int* __restrict__ keys, int node) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = node ? node_key_exact(x[i], y[i], d) : cell_key_exact(x[i], y[i], d);
}

// Logical block (item range) of this workgroup; ranges >= nblocks are idle.
__device__ __forceinline__ int logical_block(int q) { return (int)(blockIdx.x % kXcds) * q + (int)(blockIdx.x / kXcds); }

// Adds the digits of one slot of 64 items to an LDS histogram.  Keys
// re-sorted between steps arrive nearly in order, so a wave often holds one
// or two digits: the first two distinct digits are counted with one LDS add
// each (ballot + popcount), the rest with one LDS atomic per lane (a k-way
// address conflict costs about k cycles, cheap for the few digits left).
__device__ __forceinline__ void hist_add(uint32_t* __restrict__ h, bool valid, unsigned dg) {
    const int lane = threadIdx.x & 63;
    uint64_t pending = __ballot(valid);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        if (pending == 0) break;  // wave-uniform
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const unsigned d0 = (unsigned)__builtin_amdgcn_readlane((int)dg, leader);
        const uint64_t m = __ballot(dg == d0) & pending;
        if (lane == leader) atomicAdd(&h[d0], (uint32_t)__popcll(m));
        pending &= ~m;
    }
    if ((pending >> lane) & 1ull) atomicAdd(&h[dg], 1u);
}

// Lanes of this wave holding the same digit as this lane (a 64-bit lane
// mask; empty for an invalid lane).  Up to twelve distinct digits are taken
// one at a time (a read of the first pending lane's digit and one ballot
// each: a nearly sorted slot holds a few), until a digit holds a single
// lane; lanes still pending after that are matched with one ballot per digit
// bit.
template <int MAXBITS>
__device__ __forceinline__ uint64_t digit_peers(bool valid, unsigned dg, int nbits) {
    const int lane = threadIdx.x & 63;
    uint64_t pending = __ballot(valid);
    uint64_t m = 0;
    for (int it = 0; it < 12 && pending != 0; ++it) {  // wave-uniform
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const unsigned d0 = (unsigned)__builtin_amdgcn_readlane((int)dg, leader);
        const uint64_t same = __ballot(dg == d0) & pending;
        if (dg == d0 && ((pending >> lane) & 1ull)) m = same;
        pending &= ~same;
        if (__popcll(same) < 2) break;  // a lone digit: an unordered slot
    }
    if (pending != 0) {  // wave-uniform
        const bool mine = (pending >> lane) & 1ull;
        uint32_t lo = (uint32_t)pending, hi = (uint32_t)(pending >> 32);
#pragma unroll
        for (int b = 0; b < MAXBITS; ++b) {
            if (b < nbits) {
                // sgn: all ones where this lane's bit b is set; the lanes
                // agreeing with it on bit b are ~(ballot ^ sgn)
                const uint32_t sgn = (uint32_t)((int)(dg << (31 - b)) >> 31);
                const uint64_t bal = __ballot(sgn != 0u) & pending;
                lo &= ~((uint32_t)bal ^ sgn);
                hi &= ~((uint32_t)(bal >> 32) ^ sgn);
            }
        }
        if (mine) m = ((uint64_t)hi << 32) | lo;
    }
    return m;
}

// Per-(digit, block) counts of a pass: table[d * nblocks + block].
template <int I>
__global__ __launch_bounds__(kRT) void rhist_kernel(const int* __restrict__ keys, int64_t n, int shift, int bins,
                                                    int nblocks, int q, uint32_t* __restrict__ table,
                                                    const int* __restrict__ mode) {
    if (mode && *mode == 2) return;  // already in order: no pass runs
    const int lb = logical_block(q);
    if (lb >= nblocks) return;
    __shared__ uint32_t h[kMaxBins];
    for (int i = threadIdx.x; i < bins; i += kRT) h[i] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t wb = (int64_t)lb * (kRT * I) + (int64_t)w * (64 * I);
    int k[I];
#pragma unroll
    for (int j = 0; j < I; ++j) {
        const int64_t i = wb + j * 64 + lane;
        k[j] = i < n ? keys[i] : -1;
    }
#pragma unroll
    for (int j = 0; j < I; ++j) hist_add(h, k[j] >= 0, ((unsigned)k[j] >> shift) & (unsigned)(bins - 1));
    __syncthreads();
    for (int d = threadIdx.x; d < bins; d += kRT) table[(size_t)d * nblocks + lb] = h[d];
}

// Keys of the current order (into `keys`, which holds the keys of the last
// sort's order on entry) and the first pass's histogram table, with per
// block (unless part is null): part[2 lb] the particles whose key moved by
// more than `far` from the key at the same position after the last sort,
// part[2 lb + 1] the positions whose key is below the previous position's (0
// everywhere: already sorted).
template <int I, int NT = kRT>
__global__ __launch_bounds__(NT) void keys_rhist_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                         int64_t n, DevDomain d, int node, int* __restrict__ keys,
                                                         int far, int shift, int bins, int nblocks, int q,
                                                         uint32_t* __restrict__ table, unsigned* __restrict__ part,
                                                         int write_keys) {
    const int lb = logical_block(q);
    if (lb >= nblocks) return;
    __shared__ uint32_t h[kMaxBins];
    constexpr int NW = NT / 64;
    __shared__ unsigned cnt[2][NW];
    for (int i = threadIdx.x; i < bins; i += NT) h[i] = 0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t wb = (int64_t)lb * (NT * I) + (int64_t)w * (64 * I);
    int k[I];
    unsigned nfar = 0, ndesc = 0;  // wave totals, kept by every lane
    // all loads first (clamped to the last particle, so none is conditional)
    float xs[I], ys[I];
    int olds[I];
    const bool cmp_old = far >= 0;  // far < 0: no comparison with the last sort's keys (not read)
#pragma unroll
    for (int j = 0; j < I; ++j) {
        const int64_t i = min(wb + j * 64 + lane, n - 1);
        xs[j] = x[i];
        ys[j] = y[i];
        olds[j] = cmp_old ? keys[i] : 0;
    }
#pragma unroll
    for (int j = 0; j < I; ++j) {
        const int64_t i = wb + j * 64 + lane;
        k[j] = -1;
        bool f = false;
        if (i < n) {
            k[j] = node ? node_key_exact(xs[j], ys[j], d) : cell_key_exact(xs[j], ys[j], d);
            if (write_keys) keys[i] = k[j];
            const long long dk = (long long)k[j] - (long long)olds[j];
            f = cmp_old && (dk > far || dk < -far);
        }
        nfar += (unsigned)__popcll(__ballot(f));
    }
    // predecessor of the wave's first item (the previous wave's or block's last)
    int first_prev = -1;
    if (lane == 0 && wb > 0 && wb <= n)
        first_prev = node ? node_key_exact(x[wb - 1], y[wb - 1], d) : cell_key_exact(x[wb - 1], y[wb - 1], d);
#pragma unroll
    for (int j = 0; j < I; ++j) {
        const int up = __shfl_up(k[j], 1, 64);
        const int last_prev = j > 0 ? __shfl(k[j - 1], 63, 64) : __shfl(first_prev, 0, 64);
        const int prev = lane > 0 ? up : last_prev;
        ndesc += (unsigned)__popcll(__ballot(k[j] >= 0 && prev >= 0 && k[j] < prev));
    }
    __syncthreads();  // zeroed h
#pragma unroll
    for (int j = 0; j < I; ++j) hist_add(h, k[j] >= 0, ((unsigned)k[j] >> shift) & (unsigned)(bins - 1));
    if (lane == 0) {
        cnt[0][w] = nfar;
        cnt[1][w] = ndesc;
    }
    __syncthreads();
    for (int dd = threadIdx.x; dd < bins; dd += NT) table[(size_t)dd * nblocks + lb] = h[dd];
    if (part && threadIdx.x < 2) {
        unsigned s = 0;
        for (int ww = 0; ww < NW; ++ww) s += cnt[threadIdx.x][ww];
        part[2 * (size_t)lb + threadIdx.x] = s;
    }
}

// Stable ranking of one block's items by digit, shared by the scatter
// kernels.  Each wave ranks its contiguous quarter of the block slot by slot:
// lanes holding the same digit are matched (digit_peers), the group's first
// lane adds the group size to the wave's LDS histogram (returning the count
// before it) and the others read it from that lane, so ranks follow item
// order.  Then, per digit, the waves' counts become exclusive offsets and the
// digit totals a block-wide exclusive scan.  On return rank[j] is item j's
// slot in the digit-ordered block (items with key < 0 excluded), and
// gbase[d] + i the global position of the digit-ordered block's item i when
// its digit is d.  wh (kRW * 2^nbits words) is dead after the final barrier.
template <int I, int NT = kRT>
__device__ __forceinline__ void block_digit_rank(const int (&key)[I], int shift, int nbits, int nblocks, int lb,
                                                 const uint32_t* __restrict__ scan, uint32_t* wh, uint32_t* gbase,
                                                 uint32_t* bpre, uint32_t* wtot, uint32_t (&rank)[I]) {
    const int bins = 1 << nbits;
    const unsigned dmask = (unsigned)(bins - 1);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    constexpr int NW = NT / 64;
    for (int i = threadIdx.x; i < NW * bins; i += NT) wh[i] = 0;
    for (int dd = threadIdx.x; dd < bins; dd += NT) gbase[dd] = scan[(size_t)dd * nblocks + lb];
    __syncthreads();  // zeroed wave histograms, gbase loaded
    uint32_t* myh = wh + w * bins;
    // Ranking four slots at a time, each group in three phases so that the
    // slots' latencies overlap: the lanes sharing a digit (no memory access),
    // then the group leaders' histogram adds in slot order (LDS atomics of
    // one wave execute in program order, which keeps the ranks in item
    // order), then the broadcasts of the returned counts.  Groups of four
    // keep the peer masks of only four slots live (register pressure sets
    // the kernels' occupancy).
    constexpr int GR = I < 4 ? I : 4;
#pragma unroll
    for (int j0 = 0; j0 < I; j0 += GR) {
        uint64_t mm[GR];
        uint32_t old[GR], below[GR];
#pragma unroll
        for (int g = 0; g < GR; ++g) {
            const int j = j0 + g;
            mm[g] = digit_peers<10>(key[j] >= 0, ((unsigned)key[j] >> shift) & dmask, nbits);
        }
#pragma unroll
        for (int g = 0; g < GR; ++g) {
            const int j = j0 + g;
            // group members below this lane: its rank in the group
            below[g] =
                __builtin_amdgcn_mbcnt_hi((uint32_t)(mm[g] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mm[g], 0u));
            old[g] = 0;
            const unsigned dg = ((unsigned)key[j] >> shift) & dmask;
            if (key[j] >= 0 && below[g] == 0) old[g] = atomicAdd(&myh[dg], (uint32_t)__popcll(mm[g]));
        }
#pragma unroll
        for (int g = 0; g < GR; ++g) {
            const int j = j0 + g;
            const int leader = key[j] >= 0 ? __ffsll((unsigned long long)mm[g]) - 1 : lane;
            rank[j] = (uint32_t)__shfl((int)old[g], leader, 64) + below[g];
        }
    }
    __syncthreads();
    // wave counts -> exclusive offsets per digit; digit totals into bpre
    for (int dd = threadIdx.x; dd < bins; dd += NT) {
        uint32_t s = 0;
#pragma unroll
        for (int ww = 0; ww < NW; ++ww) {
            const uint32_t c = wh[ww * bins + dd];
            wh[ww * bins + dd] = s;
            s += c;
        }
        bpre[dd] = s;
    }
    __syncthreads();
    // block-wide exclusive scan of the digit totals, kDpt consecutive digits
    // per thread
    {
        constexpr int kDpt = kMaxBins / NT;
        const int d0 = kDpt * threadIdx.x;
        uint32_t t[kDpt];
        uint32_t v = 0;
#pragma unroll
        for (int e = 0; e < kDpt; ++e) {
            t[e] = d0 + e < bins ? bpre[d0 + e] : 0u;
            v += t[e];
        }
        uint32_t incl = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(incl, o, 64);
            if (lane >= o) incl += u;
        }
        if (lane == 63) wtot[w] = incl;
        __syncthreads();
        uint32_t ex = incl - v;
#pragma unroll
        for (int ww = 0; ww < NW; ++ww)
            if (ww < w) ex += wtot[ww];
#pragma unroll
        for (int e = 0; e < kDpt; ++e) {
            if (d0 + e < bins) {
                bpre[d0 + e] = ex;
                gbase[d0 + e] -= ex;
            }
            ex += t[e];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < I; ++j) {
        const unsigned dg = ((unsigned)key[j] >> shift) & dmask;
        rank[j] += key[j] >= 0 ? bpre[dg] + myh[dg] : 0u;  // local position in the digit-ordered block
    }
    __syncthreads();  // the wave histograms are read: their memory may be reused
}

// Stable scatter of (key, value) pairs for one pass (block_digit_rank, then
// every item to its digit-ordered slot of an LDS stage, from which the block
// writes out run by run).  On an order already sorted (*mode == 2) the pass
// only copies the keys through when it is the last of an odd number of
// passes (so that the sorted keys end in the same buffer).
template <int I>
__global__ __launch_bounds__(kRT, 2) void rscatter_kernel(const int* __restrict__ keys_in,
                                                       const int* __restrict__ vals_in, int64_t n, int shift,
                                                       int nbits, int nblocks, int q,
                                                       const uint32_t* __restrict__ scan, int* __restrict__ keys_out,
                                                       int* __restrict__ vals_out, const int* __restrict__ mode,
                                                       int identity_copy) {
    constexpr int B = kRT * I;
    const int lb = logical_block(q);
    if (lb >= nblocks) return;
    const int64_t b0 = (int64_t)lb * B;
    const int nb = (int)min((int64_t)B, n - b0);
    if (mode && *mode == 2) {
        if (identity_copy)
            for (int i = threadIdx.x; i < nb; i += kRT) keys_out[b0 + i] = keys_in[b0 + i];
        return;
    }
    // The wave histograms are dead once every item holds its local position,
    // so the stage reuses their memory (40 KiB per block: four blocks per CU).
    union Lds {
        int2 stage[B];
        uint32_t wh[kRW * kMaxBins];
    };
    __shared__ Lds lds;
    __shared__ uint32_t gbase[kMaxBins];
    __shared__ uint32_t bpre[kMaxBins];
    __shared__ uint32_t wtot[kRW];
    const unsigned dmask = (unsigned)((1 << nbits) - 1);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t wb = b0 + (int64_t)w * (64 * I);
    int key[I], val[I];
#pragma unroll
    for (int j = 0; j < I; ++j) {
        const int64_t i = wb + j * 64 + lane;
        key[j] = -1;
        val[j] = 0;
        if (i < n) {
            key[j] = keys_in[i];
            val[j] = vals_in ? vals_in[i] : (int)i;
        }
    }
    uint32_t rank[I];
    block_digit_rank<I>(key, shift, nbits, nblocks, lb, scan, lds.wh, gbase, bpre, wtot, rank);
#pragma unroll
    for (int j = 0; j < I; ++j)
        if (key[j] >= 0) lds.stage[rank[j]] = make_int2(key[j], val[j]);
    __syncthreads();
    // write-out: item i of the digit-ordered stage goes to gbase[digit] + i
    int2 kv[I];
#pragma unroll
    for (int j = 0; j < I; ++j) {
        const int i = threadIdx.x + j * kRT;
        kv[j] = i < nb ? lds.stage[i] : make_int2(0, 0);
    }
    uint32_t off[I];
#pragma unroll
    for (int j = 0; j < I; ++j) off[j] = gbase[((unsigned)kv[j].x >> shift) & dmask];
#pragma unroll
    for (int j = 0; j < I; ++j) {
        const int i = threadIdx.x + j * kRT;
        if (i < nb) {
            const uint32_t pos = off[j] + (uint32_t)i;
            keys_out[pos] = kv[j].x;
            vals_out[pos] = kv[j].y;
        }
    }
}

// Stable scatter of one pass that moves the particle streams with the keys
// (no permutation array, no gather afterwards).  After block_digit_rank the
// keys go through a 16 KiB LDS stage and out run by run, and then each stream
// in turn through the same stage, to the positions the keys took.  io.src /
// io.dst: this pass's input and output streams.  With kx (the positions of
// the unsorted order) the first pass forms the keys itself, as
// launch_keys_hist did, instead of reading them.  On an order already sorted
// (*mode == 2) only the last pass runs: it copies orig (the unsorted
// streams, which are already in order) to io.dst, and writes the keys (from
// kx, ky; without them copies keys_in when copy_keys, an odd number of
// passes).
template <int I, int NT>
__global__ __launch_bounds__(NT) void rrec_kernel(const int* __restrict__ keys_in, int64_t n, int shift, int nbits,
                                                   int nblocks, int q, const uint32_t* __restrict__ scan,
                                                   int* __restrict__ keys_out, StreamSet io, StreamSet orig,
                                                   const int* __restrict__ mode, int last, int copy_keys,
                                                   const float* __restrict__ kx, const float* __restrict__ ky,
                                                   DevDomain d, int node, int first) {
    constexpr int B = NT * I;
    constexpr int NW = NT / 64;
    const int lb = logical_block(q);
    if (lb >= nblocks) return;
    const int64_t b0 = (int64_t)lb * B;
    const int nb = (int)min((int64_t)B, n - b0);
    if (mode && *mode == 2) {
        if (last) {
            // 16-byte copies (block ranges start at multiples of B; the
            // stream buffers are 256-byte aligned)
            const int nv = nb >> 2;
            for (int s = 0; s < io.count; ++s) {
                const float4* src = reinterpret_cast<const float4*>(orig.src[s] + b0);
                float4* dst = reinterpret_cast<float4*>(io.dst[s] + b0);
                for (int i = threadIdx.x; i < nv; i += NT) dst[i] = src[i];
                for (int i = 4 * nv + threadIdx.x; i < nb; i += NT) io.dst[s][b0 + i] = orig.src[s][b0 + i];
            }
            if (kx) {
                for (int i = threadIdx.x; i < nb; i += NT) {
                    const float xv = kx[b0 + i], yv = ky[b0 + i];
                    keys_out[b0 + i] = node ? node_key_exact(xv, yv, d) : cell_key_exact(xv, yv, d);
                }
            } else if (copy_keys) {
                for (int i = threadIdx.x; i < nb; i += NT) keys_out[b0 + i] = keys_in[b0 + i];
            }
        }
        return;
    }
    union Lds {
        uint32_t stage[B];
        uint32_t wh[NW * kMaxBins];
    };
    __shared__ Lds lds;
    __shared__ uint32_t gbase[kMaxBins];
    __shared__ uint32_t bpre[kMaxBins];
    __shared__ uint32_t wtot[NW];
    const unsigned dmask = (unsigned)((1 << nbits) - 1);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // 32-bit item indices (n < 2^31): loads and stores address as a scalar
    // base plus a 32-bit lane offset, which keeps 64-bit addresses out of
    // the registers
    const uint32_t n32 = (uint32_t)n;
    const uint32_t wb = (uint32_t)b0 + (uint32_t)w * (64 * I);
    int key[I];
    if (first && kx) {
        float xs[I], ys[I];
#pragma unroll
        for (int j = 0; j < I; ++j) {
            const uint32_t i = min(wb + j * 64 + lane, n32 - 1);
            xs[j] = kx[i];
            ys[j] = ky[i];
        }
#pragma unroll
        for (int j = 0; j < I; ++j) {
            const uint32_t i = wb + j * 64 + lane;
            key[j] = i < n32 ? (node ? node_key_exact(xs[j], ys[j], d) : cell_key_exact(xs[j], ys[j], d)) : -1;
        }
    } else {
#pragma unroll
        for (int j = 0; j < I; ++j) {
            const uint32_t i = wb + j * 64 + lane;
            key[j] = i < n32 ? keys_in[i] : -1;
        }
    }
    uint32_t rank[I];
    block_digit_rank<I, NT>(key, shift, nbits, nblocks, lb, scan, lds.wh, gbase, bpre, wtot, rank);
#pragma unroll
    for (int j = 0; j < I; ++j) {
        if (key[j] >= 0)
            lds.stage[rank[j]] = (uint32_t)key[j];
        else
            rank[j] = ~0u;  // no item: from here on rank alone marks it
    }
    __syncthreads();
    uint32_t pos[I];
#pragma unroll
    for (int j = 0; j < I; ++j) {
        const int i = threadIdx.x + j * NT;
        const uint32_t kv = i < nb ? lds.stage[i] : 0u;
        pos[j] = gbase[(kv >> shift) & dmask] + (uint32_t)i;
        if (i < nb) keys_out[pos[j]] = (int)kv;
    }
    for (int s = 0; s < io.count; ++s) {
        float v[I];
#pragma unroll
        for (int j = 0; j < I; ++j) v[j] = io.src[s][min(wb + j * 64 + lane, n32 - 1)];
        __syncthreads();  // the stage's previous contents are read
#pragma unroll
        for (int j = 0; j < I; ++j)
            if (rank[j] != ~0u) lds.stage[rank[j]] = __float_as_uint(v[j]);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < I; ++j) {
            const int i = threadIdx.x + j * NT;
            if (i < nb) io.dst[s][pos[j]] = __uint_as_float(lds.stage[i]);
        }
    }
}

// ---- exclusive scan of uint32 (3 kernels) ---------------------------------

template <int NW = kWaves>
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
    for (int ww = 0; ww < NW; ++ww) {
        const uint32_t c = s_wave[ww];
        if (ww < w) wpre += c;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return wpre + incl - v;
}

// Each thread takes 16 consecutive values (four 16-byte loads; the table
// length is padded to a multiple of 4 by the caller's allocation and the
// values past m are read as 0).
__device__ __forceinline__ void load16(const uint32_t* __restrict__ in, size_t base, size_t m, uint32_t (&v)[kItems]) {
#pragma unroll
    for (int q4 = 0; q4 < kItems / 4; ++q4) {
        const size_t b = base + 4 * q4;
        if (b + 4 <= m) {
            const uint4 u = *reinterpret_cast<const uint4*>(in + b);
            v[4 * q4] = u.x;
            v[4 * q4 + 1] = u.y;
            v[4 * q4 + 2] = u.z;
            v[4 * q4 + 3] = u.w;
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) v[4 * q4 + e] = b + e < m ? in[b + e] : 0u;
        }
    }
}

__global__ __launch_bounds__(256) void scan_reduce_kernel(const uint32_t* __restrict__ in, size_t m,
                                                          uint32_t* __restrict__ sums, const int* __restrict__ mode) {
    if (mode && *mode == 2) return;
    __shared__ uint32_t s_wave[kWaves];
    const size_t base = (size_t)blockIdx.x * kBlockItems + (size_t)threadIdx.x * kItems;
    uint32_t v[kItems];
    load16(in, base, m, v);
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) s += v[j];
    uint32_t total;
    block_exclusive_scan(s, s_wave, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void scan_top_kernel(uint32_t* __restrict__ sums, size_t nb,
                                                       const int* __restrict__ mode) {
    if (mode && *mode == 2) return;
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

// LOCAL: `sums` holds the blocks' totals (at most kThreads of them), and each
// block adds those before it itself; otherwise `sums` holds their exclusive
// prefix (scan_top_kernel).
template <bool LOCAL>
__global__ __launch_bounds__(256) void scan_down_kernel(const uint32_t* __restrict__ in, size_t m,
                                                        const uint32_t* __restrict__ sums,
                                                        uint32_t* __restrict__ out, const int* __restrict__ mode) {
    if (mode && *mode == 2) return;
    __shared__ uint32_t s_wave[kWaves];
    const size_t base = (size_t)blockIdx.x * kBlockItems + (size_t)threadIdx.x * kItems;
    uint32_t v[kItems];
    load16(in, base, m, v);
    uint32_t carry;
    if (LOCAL) {
        const uint32_t c = threadIdx.x < blockIdx.x ? sums[threadIdx.x] : 0u;
        uint32_t before;
        (void)block_exclusive_scan(c, s_wave, before);
        carry = before;
    } else {
        carry = sums[blockIdx.x];
    }
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) s += v[j];
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, s_wave, total) + carry;
#pragma unroll
    for (int q4 = 0; q4 < kItems / 4; ++q4) {
        const size_t b = base + 4 * q4;
        uint32_t o[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            o[e] = run;
            run += v[4 * q4 + e];
        }
        if (b + 4 <= m)
            *reinterpret_cast<uint4*>(out + b) = make_uint4(o[0], o[1], o[2], o[3]);
        else
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (b + e < m) out[b + e] = o[e];
    }
}

void exclusive_scan(const uint32_t* in, uint32_t* out, size_t m, uint32_t* tmp, hipStream_t s, const int* mode) {
    const size_t nb = (m + kBlockItems - 1) / kBlockItems;
    if (nb == 0) return;
    hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nb), dim3(kThreads), 0, s, in, m, tmp, mode);
    if (nb <= kThreads) {  // every down-sweep block adds the totals before it itself: no top-level launch
        hipLaunchKernelGGL(scan_down_kernel<true>, dim3((unsigned)nb), dim3(kThreads), 0, s, in, m, tmp, out, mode);
        return;
    }
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(kThreads), 0, s, tmp, nb, mode);
    hipLaunchKernelGGL(scan_down_kernel<false>, dim3((unsigned)nb), dim3(kThreads), 0, s, in, m, tmp, out, mode);
}

// off[c] = first sorted index with key >= c, off[G] = n.  Four keys per thread.
__global__ __launch_bounds__(256) void cell_offsets_kernel(const int* __restrict__ keys, int64_t n, int G,
                                                           int* __restrict__ off) {
    const int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i0 > n) return;
    int k[4];
    if (i0 + 4 <= n) {
        const int4 v = *reinterpret_cast<const int4*>(keys + i0);
        k[0] = v.x;
        k[1] = v.y;
        k[2] = v.z;
        k[3] = v.w;
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) k[e] = (i0 + e < n) ? keys[i0 + e] : G;
    }
    int lo = i0 == 0 ? -1 : keys[i0 - 1];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int64_t i = i0 + e;
        if (i > n) break;
        const int hi = (i == n) ? G : k[e];
        for (int c = lo + 1; c <= hi; ++c) off[c] = (int)i;
        lo = hi;
    }
}

__global__ __launch_bounds__(256) void gather_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                     const int* __restrict__ perm, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

// Permutation mode of this sort from the per-block counts of the keys
// kernel: 2 when no key is below its predecessor's (the order is already
// sorted: the passes are skipped and the streams copied), else 1 (staged
// gather through packed records) when more than an eighth of the particles
// moved by more than `far` keys, else 0.  One block sums the counts.
__global__ __launch_bounds__(1024) void permute_mode_kernel(const unsigned* __restrict__ part, int nparts,
                                                            int64_t n, int* __restrict__ mode) {
    __shared__ unsigned long long sh[2][16];
    unsigned long long f = 0, dsc = 0;
    const uint2* pairs = reinterpret_cast<const uint2*>(part);
    for (int i = threadIdx.x; i < nparts; i += 1024) {
        const uint2 v = pairs[i];
        f += v.x;
        dsc += v.y;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        f += __shfl_xor(f, o, 64);
        dsc += __shfl_xor(dsc, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = f;
        sh[1][threadIdx.x >> 6] = dsc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tf = 0, td = 0;
        for (int w = 0; w < 16; ++w) {
            tf += sh[0][w];
            td += sh[1][w];
        }
        *mode = td == 0 ? 2 : ((int64_t)tf > n / 8 ? 1 : 0);
    }
}

// The stream kernels below take one item per thread: their reads are random
// (the permutation) or line-sized, and the most waves in flight hide them best.
unsigned stream_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

// Mode 1: the permuted streams of particle i packed into W float4 records
// (W = 2 for five or six streams), so that the gather reads one or two lines
// per particle instead of one per stream.
template <int W>
__global__ __launch_bounds__(256) void pack_streams_kernel(StreamSet ss, int64_t n, const int* __restrict__ mode,
                                                           float4* __restrict__ rec) {
    if (*mode != 1) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v[4 * W];
#pragma unroll
    for (int k = 0; k < 4 * W; ++k) v[k] = k < ss.count ? ss.src[k][i] : 0.0f;
    rec[W * i] = make_float4(v[0], v[1], v[2], v[3]);
    if (W == 2) rec[W * i + 1] = make_float4(v[4 % (4 * W)], v[5 % (4 * W)], v[6 % (4 * W)], v[7 % (4 * W)]);
}

// dst_k[i] = src_k[perm[i]] for up to kMaxStreams streams, by mode: 0 a
// gather of each stream (the permutation read once for all of them), 1 the
// gather of the packed records, 2 (already sorted) or perm == nullptr a copy.
template <int W>
__global__ __launch_bounds__(256) void permute_streams_kernel(StreamSet ss, const int* __restrict__ perm, int64_t n,
                                                              const int* __restrict__ mode,
                                                              const float4* __restrict__ rec) {
    const int md = mode ? *mode : 0;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v[kMaxStreams > 4 * W ? kMaxStreams : 4 * W];
    if (md == 1) {
        const int64_t p = perm[i];
#pragma unroll
        for (int h = 0; h < W; ++h) {
            const float4 r = rec[W * p + h];
            v[4 * h] = r.x;
            v[4 * h + 1] = r.y;
            v[4 * h + 2] = r.z;
            v[4 * h + 3] = r.w;
        }
    } else {
        const int64_t p = (perm == nullptr || md == 2) ? i : (int64_t)perm[i];
#pragma unroll
        for (int k = 0; k < kMaxStreams; ++k)
            if (k < ss.count) v[k] = ss.src[k][p];
    }
#pragma unroll
    for (int k = 0; k < kMaxStreams; ++k)
        if (k < ss.count) ss.dst[k][i] = v[k];
}

// The (key, index) sort's last step: dst_k[i] = src_k[perm[i]] and the
// cell-offset table of the sorted keys (thread i writes off[c] = i for the
// keys c in (keys[i-1], keys[i]]; thread n the tail up to G).
__global__ __launch_bounds__(256) void permute_offsets_kernel(StreamSet ss, const int* __restrict__ perm,
                                                              const int* __restrict__ keys, int64_t n, int G,
                                                              int* __restrict__ off) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    const int lo = i == 0 ? -1 : keys[i - 1];
    const int hi = i == n ? G : keys[i];
    for (int c = lo + 1; c <= hi; ++c) off[c] = (int)i;
    if (i == n) return;
    const int64_t p = perm[i];
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

// Blocking of the radix passes: 16 items per thread (8192-item blocks), or 4
// for sets of up to kSmallSort particles, whose passes would otherwise run on
// a few hundred workgroups.
constexpr int64_t kSmallSort = int64_t(1) << 21;

struct RadixPlan {
    int items;    // per thread
    int nblocks;  // item ranges
    int q;        // ranges per XCD
    int passes;
    int bits;     // per pass
};

RadixPlan radix_plan(int64_t n, int key_range) {
    RadixPlan p;
    p.items = n <= kSmallSort ? 4 : 16;
    const int64_t bi = (int64_t)kRT * p.items;
    p.nblocks = (int)std::max<int64_t>(1, (n + bi - 1) / bi);
    p.q = (p.nblocks + kXcds - 1) / kXcds;
    const int kb = key_bits(key_range);
    // Digits of at most 9 bits: a 4096-item block then holds 8 or more items
    // per digit on average, so its runs reach whole 32-byte sectors of each
    // stream (10-bit digits leave runs of 4 items: a 2^20-key sort of a
    // random order ran 25% slower in two 10-bit passes than in three of 7).
    p.passes = kb == 0 ? 1 : (kb + kMaxDigitBits - 1) / kMaxDigitBits;
    p.bits = kb == 0 ? 1 : (kb + p.passes - 1) / p.passes;
    return p;
}

}  // namespace

size_t sort_hist_elems(int64_t n) {
    const size_t small = (size_t)((std::min(n, kSmallSort) + kRT * 4 - 1) / (kRT * 4));
    const size_t large = (size_t)((n + kRT * 16 - 1) / (kRT * 16));
    return std::max<size_t>(std::max(small, large), 1) * kMaxBins;
}

size_t sort_scan_tmp_elems(size_t m) { return (m + kBlockItems - 1) / kBlockItems + 1; }

size_t sort_part_elems(int64_t n) { return 2 * (sort_hist_elems(n) / kMaxBins); }

void launch_keys(const float* x, const float* y, int64_t n, const DevDomain& d, int* keys, hipStream_t s, bool node) {
    if (n > 0)
        hipLaunchKernelGGL(keys_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, n, d, keys,
                           node ? 1 : 0);
}

hipError_t launch_sort_precomputed(int64_t n, int key_range, SortBuffers& sb, hipStream_t s, const int* mode,
                                   bool first_hist_done, bool offsets) {
    if (n > 0) {
        const RadixPlan p = radix_plan(n, key_range);
        const int bins = 1 << p.bits;
        const size_t m = (size_t)p.nblocks * bins;
        const dim3 grid((unsigned)(p.q * kXcds));
        int* kin = sb.keys;
        int* kout = sb.keys_alt;
        int* vin = nullptr;  // identity permutation on the first pass
        int* vout = sb.perm;
        for (int pass = 0; pass < p.passes; ++pass) {
            const int shift = p.bits * pass;
            if (pass > 0 || !first_hist_done) {
                if (p.items == 4)
                    hipLaunchKernelGGL(rhist_kernel<4>, grid, dim3(kRT), 0, s, kin, n, shift, bins, p.nblocks, p.q,
                                       sb.hist, mode);
                else
                    hipLaunchKernelGGL(rhist_kernel<16>, grid, dim3(kRT), 0, s, kin, n, shift, bins, p.nblocks, p.q,
                                       sb.hist, mode);
            }
            exclusive_scan(sb.hist, sb.hist, m, sb.scan_tmp, s, mode);
            const int copy = (pass == p.passes - 1 && (p.passes & 1)) ? 1 : 0;
            if (p.items == 4)
                hipLaunchKernelGGL(rscatter_kernel<4>, grid, dim3(kRT), 0, s, kin, vin, n, shift, p.bits, p.nblocks,
                                   p.q, sb.hist, kout, vout, mode, copy);
            else
                hipLaunchKernelGGL(rscatter_kernel<16>, grid, dim3(kRT), 0, s, kin, vin, n, shift, p.bits,
                                   p.nblocks, p.q, sb.hist, kout, vout, mode, copy);
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
    if (offsets)
        hipLaunchKernelGGL(cell_offsets_kernel, dim3((unsigned)((n / 4 + 1 + 255) / 256)), dim3(256), 0, s, sb.keys,
                           n, key_range, sb.cell_off);
    return hipGetLastError();
}

hipError_t launch_sort_keys_perm(const float* x, const float* y, int64_t n, const DevDomain& d, SortBuffers& sb,
                                 hipStream_t s) {
    launch_keys(x, y, n, d, sb.keys, s);
    return launch_sort_precomputed(n, d.Nx * d.Ny, sb, s);
}

hipError_t launch_sort_records(int64_t n, int key_range, SortBuffers& sb, const StreamSet& ss, float* tmp,
                               size_t stride, hipStream_t s, const int* mode, bool first_hist_done,
                               const float* kx, const float* ky, const DevDomain* d, bool node) {
    const DevDomain dd = d ? *d : DevDomain{};
    if (n > 0) {
        const RadixPlan p = radix_plan(n, key_range);
        const int bins = 1 << p.bits;
        const size_t m = (size_t)p.nblocks * bins;
        const dim3 grid((unsigned)(p.q * kXcds));
        int* kin = sb.keys;
        int* kout = sb.keys_alt;
        const float* in[kMaxStreams] = {};
        float* tmps[kMaxStreams] = {};
        for (int k = 0; k < ss.count; ++k) {
            in[k] = ss.src[k];
            tmps[k] = tmp + (size_t)k * stride;
        }
        for (int pass = 0; pass < p.passes; ++pass) {
            const int shift = p.bits * pass;
            if (pass > 0 || !first_hist_done) {
                if (p.items == 4)
                    hipLaunchKernelGGL(rhist_kernel<4>, grid, dim3(kRT), 0, s, kin, n, shift, bins, p.nblocks, p.q,
                                       sb.hist, mode);
                else
                    hipLaunchKernelGGL(rhist_kernel<16>, grid, dim3(kRT), 0, s, kin, n, shift, bins, p.nblocks, p.q,
                                       sb.hist, mode);
            }
            exclusive_scan(sb.hist, sb.hist, m, sb.scan_tmp, s, mode);
            const int last = pass == p.passes - 1 ? 1 : 0;
            const int copy_keys = (last && (p.passes & 1)) ? 1 : 0;
            // the last pass writes ss.dst, the ones before alternate with tmp
            const bool to_dst = ((p.passes - 1 - pass) & 1) == 0;
            StreamSet io{};
            io.count = ss.count;
            for (int k = 0; k < ss.count; ++k) {
                io.src[k] = in[k];
                io.dst[k] = to_dst ? ss.dst[k] : tmps[k];
            }
            // blocks of 256 * items: 4096-item blocks run 512 threads of 8
            // items (fewer registers per thread, twice the waves per CU)
            const int first = pass == 0 ? 1 : 0;
            if (p.items == 4)
                hipLaunchKernelGGL((rrec_kernel<4, kRT>), grid, dim3(kRT), 0, s, kin, n, shift, p.bits, p.nblocks,
                                   p.q, sb.hist, kout, io, ss, mode, last, copy_keys, kx, ky, dd, node ? 1 : 0,
                                   first);
            else
                hipLaunchKernelGGL((rrec_kernel<8, 2 * kRT>), grid, dim3(2 * kRT), 0, s, kin, n, shift, p.bits,
                                   p.nblocks, p.q, sb.hist, kout, io, ss, mode, last, copy_keys, kx, ky, dd,
                                   node ? 1 : 0, first);
            int* kt = kin;
            kin = kout;
            kout = kt;
            for (int k = 0; k < ss.count; ++k) in[k] = io.dst[k];
        }
        if (kin != sb.keys) {
            int* t = sb.keys;
            sb.keys = sb.keys_alt;
            sb.keys_alt = t;
        }
    }
    hipLaunchKernelGGL(cell_offsets_kernel, dim3((unsigned)((n / 4 + 1 + 255) / 256)), dim3(256), 0, s, sb.keys, n,
                       key_range, sb.cell_off);
    return hipGetLastError();
}

void launch_permute_streams(const StreamSet& ss, const int* perm, int64_t n, hipStream_t s, const int* mode,
                            const float4* rec) {
    if (n > 0)
        hipLaunchKernelGGL(ss.count > 4 ? permute_streams_kernel<2> : permute_streams_kernel<1>,
                           dim3(stream_grid(n)), dim3(256), 0, s, ss, perm, n, mode, rec);
}

void launch_keys_hist(const float* x, const float* y, int64_t n, const DevDomain& d, int key_range, SortBuffers& sb,
                      int far, bool node, unsigned* part, int* mode, hipStream_t s, bool write_keys) {
    const RadixPlan p = radix_plan(n, key_range);
    const int bins = 1 << p.bits;
    const dim3 grid((unsigned)(p.q * kXcds));
    if (n > 0) {
        if (p.items == 4)
            hipLaunchKernelGGL(keys_rhist_kernel<4>, grid, dim3(kRT), 0, s, x, y, n, d, node ? 1 : 0, sb.keys, far, 0,
                               bins, p.nblocks, p.q, sb.hist, part, write_keys ? 1 : 0);
        else
            hipLaunchKernelGGL((keys_rhist_kernel<8, 2 * kRT>), grid, dim3(2 * kRT), 0, s, x, y, n, d, node ? 1 : 0,
                               sb.keys, far,
                               0, bins, p.nblocks, p.q, sb.hist, part, write_keys ? 1 : 0);
    }
    hipLaunchKernelGGL(permute_mode_kernel, dim3(1), dim3(1024), 0, s, part, n > 0 ? p.nblocks : 0, n, mode);
}

hipError_t launch_sort_gather(const float* x, const float* y, int64_t n, const DevDomain& d, int key_range,
                              bool node, SortBuffers& sb, const StreamSet& ss, hipStream_t s) {
    if (n > 0) {
        const RadixPlan p = radix_plan(n, key_range);
        const int bins = 1 << p.bits;
        const dim3 grid((unsigned)(p.q * kXcds));
        if (p.items == 4)
            hipLaunchKernelGGL(keys_rhist_kernel<4>, grid, dim3(kRT), 0, s, x, y, n, d, node ? 1 : 0, sb.keys, -1, 0,
                               bins, p.nblocks, p.q, sb.hist, nullptr, 1);
        else
            hipLaunchKernelGGL((keys_rhist_kernel<8, 2 * kRT>), grid, dim3(2 * kRT), 0, s, x, y, n, d, node ? 1 : 0,
                               sb.keys, -1, 0, bins, p.nblocks, p.q, sb.hist, nullptr, 1);
        const hipError_t e = launch_sort_precomputed(n, key_range, sb, s, nullptr, true, false);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(permute_offsets_kernel, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0, s, ss, sb.perm,
                       sb.keys, n, key_range, sb.cell_off);
    return hipGetLastError();
}

void launch_permute_mode(const unsigned* part, int nparts, int64_t n, int* mode, hipStream_t s) {
    hipLaunchKernelGGL(permute_mode_kernel, dim3(1), dim3(1024), 0, s, part, nparts, n, mode);
}

void launch_pack_streams(const StreamSet& ss, int64_t n, const int* mode, float4* rec, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(ss.count > 4 ? pack_streams_kernel<2> : pack_streams_kernel<1>, dim3(stream_grid(n)),
                           dim3(256), 0, s, ss, n, mode, rec);
}

void launch_gather_f32(const float* src, float* dst, const int* perm, int64_t n, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(gather_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, perm, n);
}



}  // namespace pichip
