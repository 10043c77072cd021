// Batched ensemble kernels: many independent periodic PIC instances
// ("members") of uneven size share one set of particle streams and grids.
//
//  * Per-particle kernels run one workgroup per chunk of kEnsChunk slots;
//    member ranges are chunk aligned, so the member of a workgroup is found
//    once (binary search over chunk starts, uniform per workgroup).
//  * ens_deposit_kernel is persistent: its waves pull (member, tile) work
//    items from a device counter, largest tiles first, and deposit each with
//    the same per-wave LDS tile code as the single-grid kernel.
//  * Grid kernels (merge, FFT preparation, field sums) run one workgroup per
//    256 nodes of one member; FP64 per-member sums go through per-block
//    partials summed by one block per member in a fixed order.

#include "deposit_device.hpp"
#include "particle_device.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

namespace pichip {

namespace {

// Largest m with start[m] <= b (start non-decreasing, start[0] = 0).
__device__ __forceinline__ int find_member(const int64_t* __restrict__ start, int M, int64_t b) {
    int lo = 0, hi = M - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Fixed-shape block tree of N FP64 values (256 threads); thread 0 returns the totals.
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N]) {
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

__global__ __launch_bounds__(256) void ens_init_kernel(EnsParticles p, EnsTables t) {
    const int64_t c = blockIdx.x;
    const int mi = find_member(t.pchunk, t.M, c);
    const EnsMember& m = t.mem[mi];
    for (int j = 0; j < kEnsChunk / 256; ++j) {
        const int64_t i = c * kEnsChunk + j * 256 + threadIdx.x;
        const int64_t local = i - m.pbase;
        ParticleInit o{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};  // padding slot
        if (local < m.n) o = init_particle(local, m.n, m.d, m.init);
        p.x[i] = o.x;
        p.y[i] = o.y;
        p.vx[i] = o.vx;
        p.vy[i] = o.vy;
        p.q[i] = o.q;
        p.m[i] = o.m;
    }
}

// Member-offset keys: cbase + particleCellIndex, or for NGP and TSC members
// cbase + nearest node (the node-keyed order their tile kernels take);
// padding slots get the member's last key cbase + G, so a stable sort keeps
// them at its end.
__device__ __forceinline__ int ens_key_at(const float* __restrict__ x, const float* __restrict__ y,
                                          const EnsMember& m, int64_t i) {
    const int64_t local = i - m.pbase;
    if (local >= m.n) return (int)(m.cbase + m.G);
    const bool node = m.scheme == NGP || m.scheme == TSC;
    const int k = node ? node_key_exact(x[i], y[i], m.d) : cell_key_exact(x[i], y[i], m.d);
    return (int)(m.cbase + k);
}

// One block per chunk (every chunk belongs to one member); each wave takes a
// contiguous quarter of it.  With `part`, also per chunk: part[2c] the slots
// whose key differs from the key at the same position after the last sort
// (`keys` holds those on entry), part[2c + 1] the slots whose key is below
// the previous slot's (0 everywhere: the order is already sorted).
__global__ __launch_bounds__(256) void ens_keys_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                       EnsTables t, int* __restrict__ keys,
                                                       unsigned* __restrict__ part) {
    constexpr int J = kEnsChunk / 256;
    const int64_t c = blockIdx.x;
    const EnsMember& m = t.mem[find_member(t.pchunk, t.M, c)];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t wb = c * kEnsChunk + (int64_t)w * (64 * J);
    int kk[J];
    unsigned nchg = 0;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int64_t i = wb + j * 64 + lane;
        kk[j] = ens_key_at(x, y, m, i);
        if (part) nchg += (unsigned)__popcll(__ballot(kk[j] != keys[i]));
        keys[i] = kk[j];
    }
    if (!part) return;
    int first_prev = -1;  // the slot before the wave's first (another chunk's for w == 0)
    if (lane == 0 && wb > 0) {
        const EnsMember& mp = w > 0 ? m : t.mem[find_member(t.pchunk, t.M, c - 1)];
        first_prev = ens_key_at(x, y, mp, wb - 1);
    }
    unsigned ndesc = 0;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int up = __shfl_up(kk[j], 1, 64);
        const int last_prev = j > 0 ? __shfl(kk[j - 1], 63, 64) : __shfl(first_prev, 0, 64);
        const int prev = lane > 0 ? up : last_prev;
        ndesc += (unsigned)__popcll(__ballot(kk[j] < prev));
    }
    __shared__ unsigned cnt[2][4];
    if (lane == 0) {
        cnt[0][w] = nchg;
        cnt[1][w] = ndesc;
    }
    __syncthreads();
    if (threadIdx.x < 2)
        part[2 * c + threadIdx.x] = cnt[threadIdx.x][0] + cnt[threadIdx.x][1] + cnt[threadIdx.x][2] + cnt[threadIdx.x][3];
}

__global__ void set_ints_kernel(int* __restrict__ p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <int S, int K, int EX, int LA>
__global__ __launch_bounds__(256) void ens_deposit_kernel(EnsParticles p, const int* __restrict__ cell_off,
                                                          EnsTables t, const int2* __restrict__ items, int nitems,
                                                          int* __restrict__ counter, float dt,
                                                          float* __restrict__ partial, double* __restrict__ gacc,
                                                          int* __restrict__ status,
                                                          int lds_per_wave) {
    extern __shared__ double smem_d[];
    const int lane = threadIdx.x & 63;
    float* lds = reinterpret_cast<float*>(smem_d) + (size_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * lds_per_wave;
    for (;;) {
        int it = 0;
        if (lane == 0) it = atomicAdd(counter, 1);
        it = __builtin_amdgcn_readfirstlane(it);
        if (it >= nitems) break;  // wave-uniform exit once the list is drained
        const int2 w = items[it];
        const EnsMember& m = t.mem[w.x];
        const DevDomain d = m.d;
        const TileGeom g = m.g;
        deposit_one_tile<S, K, EX, 0, LA>(w.y, lds, p.x, p.y, p.vx, p.vy, p.q, cell_off + m.cbase, d, g, dt,
                                          partial + m.partial_base, gacc + m.gbase, status);
    }
}

__global__ __launch_bounds__(256) void ens_merge_kernel(const float* __restrict__ partial,
                                                        double* __restrict__ gacc,
                                                        EnsTables t, int64_t blk0, float* __restrict__ rho,
                                                        double* __restrict__ rho_part) {
    const int64_t b = blk0 + blockIdx.x;
    const int mi = find_member(t.gblk, t.M, b);
    const EnsMember& m = t.mem[mi];
    const int64_t node = (b - t.gblk[mi]) * 256 + threadIdx.x;
    double v[1] = {0.0};
    if (node < m.G) {
        const DevDomain d = m.d;
        const TileGeom g = m.g;
        const int gx = (int)(node % d.Nx), gy = (int)(node / d.Nx);
        const float r = merge_node(gx, gy, partial + m.partial_base, gacc + m.gbase, d, g);
        rho[m.gbase + node] = r;
        v[0] = r;
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) rho_part[b] = v[0];
}

__global__ __launch_bounds__(256) void ens_member_grid_sum_kernel(const double* __restrict__ part, int stride, int k,
                                                                  EnsTables t, double* __restrict__ out) {
    const int mi = blockIdx.x;
    const int64_t b0 = t.gblk[mi], b1 = t.gblk[mi + 1];
    double v[1] = {0.0};
    for (int64_t b = b0 + threadIdx.x; b < b1; b += 256) v[0] += part[b * stride + k];
    block_sum<1>(v);
    if (threadIdx.x == 0) out[mi] = v[0];
}

// in = rho * cv - mean(rho * cv) per member (FP64 mean; poisson_fft.cpp:45-66
// zeroes the k = 0 coefficient, here the mean never enters the FP32 FFT).
__global__ __launch_bounds__(256) void ens_prep_kernel(const float* __restrict__ rho, float* __restrict__ in,
                                                       EnsTables t, const double* __restrict__ rho_sum) {
    const int64_t b = blockIdx.x;
    const int mi = find_member(t.gblk, t.M, b);
    const EnsMember& m = t.mem[mi];
    const int64_t node = (b - t.gblk[mi]) * 256 + threadIdx.x;
    if (node >= m.G) return;
    const double cv = m.d.dx * m.d.dy;
    const double mean = rho_sum[mi] * cv / (double)m.G;
    in[m.gbase + node] = (float)((double)rho[m.gbase + node] * cv - mean);
}

// Ex_hat = -Dx rho_hat / (k^2 Nx Ny), Ey_hat likewise (k = 0 zeroed), with the
// member's k^2 and forward-difference symbol tables.
__global__ __launch_bounds__(256) void ens_spectral_kernel(const float2* __restrict__ spec, float2* __restrict__ sex,
                                                           float2* __restrict__ sey, const double* __restrict__ k2,
                                                           const float2* __restrict__ sym, EnsTables t) {
    const int64_t b = blockIdx.x;
    const int mi = find_member(t.sblk, t.M, b);
    const EnsMember& m = t.mem[mi];
    const int hx = m.d.Nx / 2 + 1;
    const int64_t e = (b - t.sblk[mi]) * 256 + threadIdx.x;
    if (e >= (int64_t)hx * m.d.Ny) return;
    const int i = (int)(e % hx), j = (int)(e / hx);
    float2 ph = make_float2(0.0f, 0.0f);
    if (i != 0 || j != 0) {
        const double* kt = k2 + m.tbase;
        const float f = (float)(1.0 / ((kt[i] + kt[hx + j]) * (double)m.d.Nx * (double)m.d.Ny));
        const float2 v = spec[m.sbase + e];
        ph = make_float2(v.x * f, v.y * f);
    }
    const float2 a = sym[m.tbase + i], c = sym[m.tbase + hx + j];
    sex[m.sbase + e] = make_float2(-(a.x * ph.x - a.y * ph.y), -(a.x * ph.y + a.y * ph.x));
    sey[m.sbase + e] = make_float2(-(c.x * ph.x - c.y * ph.y), -(c.x * ph.y + c.y * ph.x));
}

__global__ __launch_bounds__(256) void ens_field_sums_kernel(const float* __restrict__ Ex,
                                                             const float* __restrict__ Ey,
                                                             const double* __restrict__ cosk,
                                                             const double* __restrict__ sink, EnsTables t,
                                                             double* __restrict__ field_part) {
    const int64_t b = blockIdx.x;
    const int mi = find_member(t.gblk, t.M, b);
    const EnsMember& m = t.mem[mi];
    const int64_t node = (b - t.gblk[mi]) * 256 + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (node < m.G) {
        const double ex = Ex[m.gbase + node], ey = Ey[m.gbase + node];
        const int ix = (int)(node % m.d.Nx);
        v[0] = ex * ex + ey * ey;
        v[1] = ex * cosk[m.xbase + ix];
        v[2] = ex * sink[m.xbase + ix];
    }
    block_sum<3>(v);
    if (threadIdx.x == 0) {
        field_part[3 * b] = v[0];
        field_part[3 * b + 1] = v[1];
        field_part[3 * b + 2] = v[2];
    }
}

__global__ __launch_bounds__(256) void ens_gather_push_kernel(EnsParticles p, const float* __restrict__ Ex,
                                                              const float* __restrict__ Ey, EnsTables t, float dt,
                                                              int do_gather, const int* __restrict__ push_flag,
                                                              double* __restrict__ part_p) {
    const int64_t c = blockIdx.x;
    const int mi = find_member(t.pchunk, t.M, c);
    const EnsMember& m = t.mem[mi];
    const DevDomain d = m.d;
    const float* ex_g = Ex + m.gbase;
    const float* ey_g = Ey + m.gbase;
    const bool do_push = *push_flag != 0;
    const float Lx = (float)d.Lx, Ly = (float)d.Ly;
    const int64_t end = m.pbase + m.n;
    double v[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < kEnsChunk / 256; ++j) {
        const int64_t i = c * kEnsChunk + j * 256 + threadIdx.x;
        if (i >= end) break;
        const float x = p.x[i], y = p.y[i];
        float vx = p.vx[i], vy = p.vy[i];
        const float q = p.q[i], mass = p.m[i];
        if (do_gather) {
            float ex, ey;
            gather_cic(x, y, ex_g, ey_g, d, ex, ey);
            const float qm = -q / mass;
            vx += qm * ex * dt;
            vy += qm * ey * dt;
            p.vx[i] = vx;
            p.vy[i] = vy;
        }
        if (do_push) {
            p.x[i] = push_coord(x, vx, dt, Lx);
            p.y[i] = push_coord(y, vy, dt, Ly);
        }
        const double a = vx, bb = vy;
        v[0] += 0.5 * (double)mass * (a * a + bb * bb);
        v[1] += (double)q;
        v[2] += fabs((double)q);
    }
    block_sum<3>(v);
    if (threadIdx.x == 0) {
        part_p[3 * c] = v[0];
        part_p[3 * c + 1] = v[1];
        part_p[3 * c + 2] = v[2];
    }
}

__global__ __launch_bounds__(256) void ens_diag_kernel(const double* __restrict__ rho_part,
                                                       const double* __restrict__ field_part,
                                                       const double* __restrict__ part_p, EnsTables t,
                                                       int* __restrict__ counters, int record_from, int cap,
                                                       StepRecords rec) {
    const int mi = blockIdx.x;
    const EnsMember& m = t.mem[mi];
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int64_t b = t.gblk[mi] + threadIdx.x; b < t.gblk[mi + 1]; b += 256) {
        v[0] += rho_part[b];
        v[1] += field_part[3 * b];
        v[2] += field_part[3 * b + 1];
        v[3] += field_part[3 * b + 2];
    }
    for (int64_t c = t.pchunk[mi] + threadIdx.x; c < t.pchunk[mi + 1]; c += 256) {
        v[4] += part_p[3 * c];
        v[5] += part_p[3 * c + 1];
        v[6] += part_p[3 * c + 2];
    }
    block_sum<7>(v);
    if (threadIdx.x == 0) {
        const int step = counters[mi];
        const int slot = step - record_from;
        if (slot >= 0 && slot < cap) {
            const double cv = m.d.dx * m.d.dy;
            const double G = (double)m.G;
            const size_t o = (size_t)m.user * cap + slot;
            const double re = v[2] / G, im = v[3] / G;
            if (rec.field_energy) rec.field_energy[o] = 0.5 * v[1] * cv;
            if (rec.mode1) rec.mode1[o] = sqrt(re * re + im * im);
            if (rec.kinetic) rec.kinetic[o] = v[4];
            if (rec.charge_error) rec.charge_error[o] = fabs(v[0] * cv - v[5]) / fmax(v[6], 1e-12);
        }
        counters[mi] = step + 1;
    }
}

// ---- persistent deposit dispatch ---------------------------------------------

using EnsKernel = void (*)(EnsParticles, const int*, EnsTables, const int2*, int, int*, float, float*, double*, int*,
                           int);

// drift: the members' windows have a drift halo, so the lookahead variant of
// the chunk processing runs (see process_chunk).
template <int S, int K>
EnsKernel pick_ex(int exact, int drift) {
    if (drift)
        return exact ? (EnsKernel)ens_deposit_kernel<S, K, 1, 1> : (EnsKernel)ens_deposit_kernel<S, K, 0, 1>;
    return exact ? (EnsKernel)ens_deposit_kernel<S, K, 1, 0> : (EnsKernel)ens_deposit_kernel<S, K, 0, 0>;
}

template <int S>
EnsKernel pick_k(int ppl, int exact, int drift) {
    return ppl == 4 ? pick_ex<S, 4>(exact, drift) : pick_ex<S, 8>(exact, drift);
}

EnsKernel pick_kernel(int scheme, int ppl, int exact, int drift) {
    switch (scheme) {
        case NGP:  // members sorted by nearest node
            return pick_k<NGP_NODE>(ppl, exact, drift);
        case CIC:
            return pick_k<CIC>(ppl, exact, drift);
        case TSC:
            return pick_k<TSC_NODE>(ppl, exact, drift);
        case TSC_MOM:  // only without a drift halo
            if (ppl == 4)
                return exact ? (EnsKernel)ens_deposit_kernel<TSC_MOM, 4, 1, 0> : (EnsKernel)ens_deposit_kernel<TSC_MOM, 4, 0, 0>;
            return exact ? (EnsKernel)ens_deposit_kernel<TSC_MOM, 8, 1, 0> : (EnsKernel)ens_deposit_kernel<TSC_MOM, 8, 0, 0>;
        default:
            return pick_ex<ESK, 4>(exact, drift);  // Esirkepov runs 4 particles per lane
    }
}

// Resident workgroups per CU of one instantiation at a given LDS size.
int blocks_per_cu(EnsKernel k, size_t shmem) {
    static std::mutex mu;
    static std::map<std::pair<const void*, size_t>, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_pair((const void*)k, shmem);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)k, 256, shmem) != hipSuccess || nb < 1) nb = 1;
    cache[key] = nb;
    return nb;
}

}  // namespace

void launch_ens_init(const EnsParticles& p, const EnsTables& t, int64_t nchunks, hipStream_t s) {
    if (nchunks > 0) hipLaunchKernelGGL(ens_init_kernel, dim3((unsigned)nchunks), dim3(256), 0, s, p, t);
}

void launch_ens_keys(const EnsParticles& p, const EnsTables& t, int64_t nchunks, int* keys, hipStream_t s,
                     unsigned* part) {
    if (nchunks > 0)
        hipLaunchKernelGGL(ens_keys_kernel, dim3((unsigned)nchunks), dim3(256), 0, s, p.x, p.y, t, keys, part);
}

void launch_set_ints(int* p, int n, int v, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(set_ints_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

int ens_deposit_grid(const EnsWork& w, int ncu) {
    const EnsKernel k = pick_kernel(w.scheme, w.ppl, w.exact, w.drift);
    const size_t shmem = (size_t)4 * w.lds_per_wave * sizeof(float);
    const int64_t want = (int64_t)blocks_per_cu(k, shmem) * ncu;
    const int64_t need = ((int64_t)w.nitems + 3) / 4;  // four waves per workgroup
    return (int)std::max<int64_t>(1, std::min(want, need));
}

hipError_t launch_ens_deposit(const EnsWork* groups, int ngroups, const EnsParticles& p, const int* cell_off,
                              const EnsTables& t, float dt, float* partial, double* gacc,
                              int* status, int* counters, int ncu, hipStream_t s) {
    (void)ncu;
    launch_set_ints(counters, ngroups, 0, s);
    for (int gi = 0; gi < ngroups; ++gi) {
        const EnsWork& w = groups[gi];
        if (w.nitems <= 0) continue;
        const EnsKernel k = pick_kernel(w.scheme, w.ppl, w.exact, w.drift);
        const size_t shmem = (size_t)4 * w.lds_per_wave * sizeof(float);
        hipLaunchKernelGGL(k, dim3((unsigned)w.grid), dim3(256), shmem, s, p, cell_off, t, w.items, w.nitems,
                           counters + gi, dt, partial, gacc, status, w.lds_per_wave);
    }
    return hipGetLastError();
}

void launch_ens_merge(const float* partial, double* gacc, const EnsTables& t, int64_t blk0,
                      int64_t nblk, float* rho, double* rho_part, hipStream_t s) {
    if (nblk > 0)
        hipLaunchKernelGGL(ens_merge_kernel, dim3((unsigned)nblk), dim3(256), 0, s, partial, gacc, t, blk0,
                           rho, rho_part);
}

void launch_ens_member_grid_sum(const double* part, int stride, int k, const EnsTables& t, double* out,
                                hipStream_t s) {
    hipLaunchKernelGGL(ens_member_grid_sum_kernel, dim3((unsigned)t.M), dim3(256), 0, s, part, stride, k, t, out);
}

void launch_ens_prep(const float* rho, float* in, const EnsTables& t, int64_t nblk, const double* rho_sum,
                     hipStream_t s) {
    if (nblk > 0) hipLaunchKernelGGL(ens_prep_kernel, dim3((unsigned)nblk), dim3(256), 0, s, rho, in, t, rho_sum);
}

void launch_ens_spectral(const float2* spec, float2* spec_ex, float2* spec_ey, const double* k2, const float2* sym,
                         const EnsTables& t, int64_t nblk, hipStream_t s) {
    if (nblk > 0)
        hipLaunchKernelGGL(ens_spectral_kernel, dim3((unsigned)nblk), dim3(256), 0, s, spec, spec_ex, spec_ey, k2, sym,
                           t);
}

void launch_ens_field_sums(const float* Ex, const float* Ey, const double* cosk, const double* sink,
                           const EnsTables& t, int64_t nblk, double* field_part, hipStream_t s) {
    if (nblk > 0)
        hipLaunchKernelGGL(ens_field_sums_kernel, dim3((unsigned)nblk), dim3(256), 0, s, Ex, Ey, cosk, sink, t,
                           field_part);
}

void launch_ens_gather_push(const EnsParticles& p, const float* Ex, const float* Ey, const EnsTables& t,
                            int64_t nchunks, float dt, int do_gather, const int* push_flag, double* part_p,
                            hipStream_t s) {
    if (nchunks > 0)
        hipLaunchKernelGGL(ens_gather_push_kernel, dim3((unsigned)nchunks), dim3(256), 0, s, p, Ex, Ey, t, dt,
                           do_gather, push_flag, part_p);
}

void launch_ens_diag(const double* rho_part, const double* field_part, const double* part_p, const EnsTables& t,
                     int* counters, int record_from, int cap, const StepRecords& rec, hipStream_t s) {
    hipLaunchKernelGGL(ens_diag_kernel, dim3((unsigned)t.M), dim3(256), 0, s, rho_part, field_part, part_p, t,
                       counters, record_from, cap, rec);
}

}  // namespace pichip
