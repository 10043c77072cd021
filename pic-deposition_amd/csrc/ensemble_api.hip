// C ABI of the batched ensemble (pic_hip_ensemble_*, declared in
// include/pic_hip.h).
//
// Members are stored in an internal order sorted by (grid, scheme, particles
// per cell), so that members of one grid size are adjacent (one batched hipFFT
// plan per size) and the autotuner's classes are contiguous.  Per-member
// offsets into the shared buffers live in a device table (EnsMember); the
// caller's member index is kept in it for the records.

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pic_hip.h"
#include "api_util.hpp"
#include "pic_common.hpp"
#include "pic_kernels.hpp"

using namespace pichip;

namespace {

struct FftClass {
    int first = 0, count = 0;  // internal member range
    hipfftHandle r2c = 0, c2r = 0;
};

}  // namespace

struct pic_hip_ensemble {
    int M = 0;
    std::vector<int> user_of;  // internal -> caller index
    std::vector<int> pos_of;   // caller -> internal index
    std::vector<EnsMember> mem;
    std::vector<int64_t> pchunk, gblk, sblk;  // M + 1 block starts
    int64_t nslots = 0, nchunks = 0, Gtot = 0, Stot = 0, keys_total = 0, gblocks = 0, sblocks = 0;
    int64_t tab_total = 0, x_total = 0;
    int ncu = 1;
    size_t cap = 0;  // slots of each stream buffer (nslots plus a float4 over-read pad)
    hipStream_t stream = nullptr;
    float* s[kStreams] = {};
    float* alt[kStreams] = {};
    SortBuffers sb{};
    bool sorted = false;
    float *rho = nullptr, *Ex = nullptr, *Ey = nullptr, *fft_in = nullptr;
    double* acc64 = nullptr;  // FP64 global-atomic accumulation grid of all members, zero between deposits
    float2 *spec = nullptr, *spec_ex = nullptr, *spec_ey = nullptr;
    double *k2 = nullptr, *cosk = nullptr, *sink = nullptr;
    float2* sym = nullptr;
    float* partial = nullptr;
    size_t partial_cap = 0;
    int* status = nullptr;
    // sort: per-chunk key-change / descent counts, the device-chosen mode
    // (2: still in order), and the intermediate streams of the radix passes
    // (eight floats per slot of cap)
    unsigned* sort_part = nullptr;
    int* permute_mode = nullptr;
    float4* sort_rec = nullptr;
    int* counters = nullptr;       // one per work group
    int* diag_counters = nullptr;  // one per member
    int* push_flag = nullptr;
    double *rho_part = nullptr, *field_part = nullptr, *part_p = nullptr, *rho_sum = nullptr;
    int* h_status = nullptr;
    EnsMember* d_mem = nullptr;
    int64_t *d_pchunk = nullptr, *d_gblk = nullptr, *d_sblk = nullptr;
    std::vector<EnsWork> groups;
    int2* d_items = nullptr;
    size_t items_cap = 0;
    bool dirty = true;  // member table / work lists must be rebuilt before the next deposit
    std::vector<FftClass> classes;
    bool fft_ready = false;
    int halo = 0;  // drift halo of every member's tile windows

    ~pic_hip_ensemble() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& c : classes) {
            if (c.r2c) (void)hipfftDestroy(c.r2c);
            if (c.c2r) (void)hipfftDestroy(c.c2r);
        }
        for (int i = 0; i < kStreams; ++i) {
            (void)hipFree(s[i]);
            (void)hipFree(alt[i]);
        }
        void* bufs[] = {sb.keys,   sb.keys_alt, sb.perm,     sb.perm_alt,    sb.hist,    sb.scan_tmp, sb.cell_off,
                        rho,       acc64,       Ex,          Ey,             fft_in,     spec,
                        spec_ex,   spec_ey,     k2,          cosk,           sink,       sym,         partial,
                        status,    counters,    diag_counters, push_flag,    rho_part,   field_part,  part_p,
                        rho_sum,   d_mem,       d_pchunk,    d_gblk,         d_sblk,     d_items,
                        sort_part, permute_mode, sort_rec};
        for (void* b : bufs) (void)hipFree(b);
        (void)hipHostFree(h_status);
        if (stream) (void)hipStreamDestroy(stream);
    }

    EnsTables tables() const { return EnsTables{d_mem, M, d_pchunk, d_gblk, d_sblk}; }
    EnsParticles particles() const { return EnsParticles{s[0], s[1], s[2], s[3], s[4], s[5]}; }
};

namespace {

pic_status check_ens(pic_hip_ensemble* e) {
    if (!e) return fail(PIC_ERR_INVALID_ARGUMENT, "null ensemble");
    return PIC_OK;
}

pic_status check_member(pic_hip_ensemble* e, int32_t member) {
    if (check_ens(e) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (member < 0 || member >= e->M) return fail(PIC_ERR_INVALID_ARGUMENT, "member index out of range");
    return PIC_OK;
}

pic_status set_member_tiling(pic_hip_ensemble* e, int i, int tx, int ty, int ppl) {
    EnsMember& m = e->mem[(size_t)i];
    if (tx < 1 || ty < 1) return fail(PIC_ERR_INVALID_ARGUMENT, "tile sizes must be >= 1");
    if (ppl != 4 && ppl != 8) return fail(PIC_ERR_INVALID_ARGUMENT, "particles_per_lane must be 4 or 8");
    if (m.scheme == PIC_SCHEME_ESIRKEPOV && ppl != 4)
        return fail(PIC_ERR_INVALID_ARGUMENT, "ensemble Esirkepov members run 4 particles per lane");
    const TileGeom g = make_tile_geom(m.d.Nx, m.d.Ny, tx, ty, e->halo);
    if ((size_t)4 * g.lds_per_wave * sizeof(float) > kMaxDepositLds)
        return fail(PIC_ERR_INVALID_ARGUMENT, "tile too large for LDS (4 windows must fit 64 KiB)");
    m.g = g;
    m.ppl = ppl;
    e->dirty = true;
    return PIC_OK;
}

// Lays out the tile windows, uploads the member table and builds the work
// lists of the members in [m0, m1) (all members by default).
pic_status rebuild_work(pic_hip_ensemble* e, int m0 = 0, int m1 = -1) {
    if (m1 < 0) m1 = e->M;
    size_t pb = 0;
    for (auto& m : e->mem) {
        m.partial_base = (int64_t)pb;
        pb += (size_t)m.g.ntiles * m.g.wsize;
    }
    if (pb > e->partial_cap) {
        PIC_HIP_TRY(hipStreamSynchronize(e->stream));
        (void)hipFree(e->partial);
        e->partial = nullptr;
        e->partial_cap = 0;
        PIC_HIP_TRY(dmalloc(&e->partial, pb));
        e->partial_cap = pb;
    }
    PIC_HIP_TRY(hipMemcpyAsync(e->d_mem, e->mem.data(), e->mem.size() * sizeof(EnsMember), hipMemcpyHostToDevice,
                               e->stream));
    // Groups by (stencil, particles per lane, coordinate path); members inside a
    // group ordered by particles per tile, largest first.  TSC members whose
    // windows take the moment form (as a single context's deposit would) form
    // groups of their own.
    struct Key {
        int scheme, ppl, exact;
        bool operator<(const Key& o) const {
            return std::tie(scheme, ppl, exact) < std::tie(o.scheme, o.ppl, o.exact);
        }
    };
    std::map<Key, std::vector<int>> by;
    for (int i = m0; i < m1; ++i) {
        const EnsMember& m = e->mem[(size_t)i];
        const int st = (m.scheme == PIC_SCHEME_TSC && tsc_uses_moments(m.g, m.ppl)) ? (int)TSC_MOM : m.scheme;
        by[Key{st, m.ppl, m.d.exact}].push_back(i);
    }
    std::vector<int2> items;
    std::vector<size_t> offs;
    e->groups.clear();
    for (auto& kv : by) {
        auto& list = kv.second;
        std::stable_sort(list.begin(), list.end(), [&](int a, int b) {
            const EnsMember &A = e->mem[(size_t)a], &B = e->mem[(size_t)b];
            return (double)A.n / A.g.ntiles > (double)B.n / B.g.ntiles;
        });
        EnsWork w{};
        w.scheme = kv.first.scheme;
        w.ppl = kv.first.ppl;
        w.exact = kv.first.exact;
        w.drift = e->halo > 0 ? 1 : 0;
        w.lds_per_wave = 0;
        offs.push_back(items.size());
        for (int i : list) {
            const EnsMember& m = e->mem[(size_t)i];
            w.lds_per_wave = std::max(w.lds_per_wave, w.scheme == TSC_MOM ? moment_lds_per_wave(m.g, m.ppl) : m.g.lds_per_wave);
            for (int t = 0; t < m.g.ntiles; ++t) items.push_back(make_int2(i, t));
        }
        w.nitems = (int)(items.size() - offs.back());
        w.grid = ens_deposit_grid(w, e->ncu);
        e->groups.push_back(w);
    }
    if (e->groups.size() > 16) return fail(PIC_ERR_HIP, "too many work groups");
    if (items.size() > e->items_cap) {
        PIC_HIP_TRY(hipStreamSynchronize(e->stream));
        (void)hipFree(e->d_items);
        e->d_items = nullptr;
        e->items_cap = 0;
        PIC_HIP_TRY(dmalloc(&e->d_items, items.size()));
        e->items_cap = items.size();
    }
    if (!items.empty())
        PIC_HIP_TRY(hipMemcpyAsync(e->d_items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice,
                                   e->stream));
    for (size_t g = 0; g < e->groups.size(); ++g) e->groups[g].items = e->d_items + offs[g];
    PIC_HIP_TRY(hipStreamSynchronize(e->stream));  // host staging vectors go out of scope
    e->dirty = (m0 != 0 || m1 != e->M);
    return PIC_OK;
}

pic_status ensure_work(pic_hip_ensemble* e) { return e->dirty ? rebuild_work(e) : PIC_OK; }

pic_status ensure_fft(pic_hip_ensemble* e) {
    if (e->fft_ready) return PIC_OK;
    for (auto& c : e->classes) {
        const EnsMember& m = e->mem[(size_t)c.first];
        // Explicit layouts: real [Ny][Nx] and half-complex [Ny][Nx/2+1] per member, back to back.
        const int hx = m.d.Nx / 2 + 1;
        int n[2] = {m.d.Ny, m.d.Nx};
        int rem[2] = {m.d.Ny, m.d.Nx};
        int cem[2] = {m.d.Ny, hx};
        const int rdist = m.d.Nx * m.d.Ny, cdist = hx * m.d.Ny;
        PIC_FFT_TRY(hipfftPlanMany(&c.r2c, 2, n, rem, 1, rdist, cem, 1, cdist, HIPFFT_R2C, c.count));
        PIC_FFT_TRY(hipfftPlanMany(&c.c2r, 2, n, cem, 1, cdist, rem, 1, rdist, HIPFFT_C2R, c.count));
        PIC_FFT_TRY(hipfftSetStream(c.r2c, e->stream));
        PIC_FFT_TRY(hipfftSetStream(c.c2r, e->stream));
    }
    e->fft_ready = true;
    return PIC_OK;
}

// As for a context: the radix passes carry the six streams with the keys
// (the staging buffer holds the intermediate set), and an order that is
// still sorted (decided on the device) is copied.
pic_status enqueue_sort(pic_hip_ensemble* e, bool copy_back) {
    StreamSet fwd{}, back{};
    fwd.count = back.count = kStreams;
    for (int i = 0; i < kStreams; ++i) {
        fwd.src[i] = e->s[i];
        fwd.dst[i] = e->alt[i];
        back.src[i] = e->alt[i];
        back.dst[i] = e->s[i];
    }
    launch_ens_keys(e->particles(), e->tables(), e->nchunks, e->sb.keys, e->stream, e->sort_part);
    launch_permute_mode(e->sort_part, (int)e->nchunks, e->nslots, e->permute_mode, e->stream);
    PIC_HIP_TRY(launch_sort_records(e->nslots, (int)e->keys_total, e->sb, fwd, reinterpret_cast<float*>(e->sort_rec),
                                    (size_t)e->cap, e->stream, e->permute_mode, false));
    if (copy_back)
        launch_permute_streams(back, nullptr, e->nslots, e->stream);
    else
        for (int i = 0; i < kStreams; ++i) std::swap(e->s[i], e->alt[i]);
    PIC_HIP_TRY(hipGetLastError());
    e->sorted = true;
    return PIC_OK;
}

// Persistent deposit kernels over the current work lists, then the merge of
// grid blocks [blk0, blk0 + nblk).
pic_status enqueue_deposit(pic_hip_ensemble* e, float dt, int64_t blk0, int64_t nblk) {
    PIC_HIP_TRY(launch_ens_deposit(e->groups.data(), (int)e->groups.size(), e->particles(), e->sb.cell_off,
                                   e->tables(), dt, e->partial, e->acc64, e->status, e->counters, e->ncu,
                                   e->stream));
    launch_ens_merge(e->partial, e->acc64, e->tables(), blk0, nblk, e->rho, e->rho_part, e->stream);
    PIC_HIP_TRY(hipGetLastError());
    return PIC_OK;
}

pic_status enqueue_poisson(pic_hip_ensemble* e) {
    const EnsTables t = e->tables();
    launch_ens_member_grid_sum(e->rho_part, 1, 0, t, e->rho_sum, e->stream);
    launch_ens_prep(e->rho, e->fft_in, t, e->gblocks, e->rho_sum, e->stream);
    for (auto& c : e->classes) {
        const EnsMember& m = e->mem[(size_t)c.first];
        PIC_FFT_TRY(hipfftExecR2C(c.r2c, e->fft_in + m.gbase, reinterpret_cast<hipfftComplex*>(e->spec + m.sbase)));
    }
    launch_ens_spectral(e->spec, e->spec_ex, e->spec_ey, e->k2, e->sym, t, e->sblocks, e->stream);
    for (auto& c : e->classes) {
        const EnsMember& m = e->mem[(size_t)c.first];
        PIC_FFT_TRY(hipfftExecC2R(c.c2r, reinterpret_cast<hipfftComplex*>(e->spec_ex + m.sbase), e->Ex + m.gbase));
        PIC_FFT_TRY(hipfftExecC2R(c.c2r, reinterpret_cast<hipfftComplex*>(e->spec_ey + m.sbase), e->Ey + m.gbase));
    }
    PIC_HIP_TRY(hipGetLastError());
    return PIC_OK;
}

// One step after the (prologue or previous step's) push: [sort], deposit,
// Poisson, gather + push (push skipped when *push_flag == 0), records.
pic_status enqueue_step(pic_hip_ensemble* e, float dt, bool with_sort, bool copy_back, int record_from, int cap,
                        const StepRecords& rec) {
    pic_status st;
    if (with_sort && (st = enqueue_sort(e, copy_back)) != PIC_OK) return st;
    if ((st = enqueue_deposit(e, dt, 0, e->gblocks)) != PIC_OK) return st;
    if ((st = enqueue_poisson(e)) != PIC_OK) return st;
    const EnsTables t = e->tables();
    launch_ens_gather_push(e->particles(), e->Ex, e->Ey, t, e->nchunks, dt, 1, e->push_flag, e->part_p, e->stream);
    launch_ens_field_sums(e->Ex, e->Ey, e->cosk, e->sink, t, e->gblocks, e->field_part, e->stream);
    launch_ens_diag(e->rho_part, e->field_part, e->part_p, t, e->diag_counters, record_from, cap, rec, e->stream);
    PIC_HIP_TRY(hipGetLastError());
    return PIC_OK;
}

struct Graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~Graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

// Captures one step.  A sort inside permutes into the other buffer set and
// swaps the host pointers (graphs for both sets are captured in turn).
pic_status capture(pic_hip_ensemble* e, float dt, bool with_sort, int record_from, int cap, const StepRecords& rec,
                   Graph& g) {
    PIC_HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    const pic_status st = enqueue_step(e, dt, with_sort, false, record_from, cap, rec);
    hipGraph_t graph = nullptr;
    const hipError_t err = hipStreamEndCapture(e->stream, &graph);
    if (st != PIC_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    PIC_HIP_TRY(err);
    g.graph = graph;
    PIC_HIP_TRY(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
    return PIC_OK;
}

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

}  // namespace

extern "C" {

pic_status pic_hip_ensemble_create(const pic_ensemble_member* members, int32_t count, pic_hip_ensemble** out) {
    if (!out) return fail(PIC_ERR_INVALID_ARGUMENT, "null output pointer");
    *out = nullptr;
    if (!members || count < 1) return fail(PIC_ERR_INVALID_ARGUMENT, "need at least one member");
    for (int i = 0; i < count; ++i) {
        const pic_ensemble_member& u = members[i];
        if (!valid_domain(&u.domain))
            return fail(PIC_ERR_INVALID_ARGUMENT, "member " + std::to_string(i) + ": invalid domain");
        if (u.scheme < 0 || u.scheme > 3)
            return fail(PIC_ERR_INVALID_ARGUMENT, "member " + std::to_string(i) + ": unknown scheme");
        if (u.num_particles < 0)
            return fail(PIC_ERR_INVALID_ARGUMENT, "member " + std::to_string(i) + ": negative particle count");
        if (u.init.kind < 0 || u.init.kind > 2 || (u.init.kind == 1 && std::fabs(u.init.amp) >= 1.0))
            return fail(PIC_ERR_INVALID_ARGUMENT, "member " + std::to_string(i) + ": invalid init params");
    }
    std::unique_ptr<pic_hip_ensemble> e(new pic_hip_ensemble());
    const int M = count;
    e->M = M;
    e->user_of.resize((size_t)M);
    std::iota(e->user_of.begin(), e->user_of.end(), 0);
    auto cls = [&](int i) {
        const pic_ensemble_member& u = members[i];
        return std::make_tuple(u.domain.Nx, u.domain.Ny, u.domain.Lx, u.domain.Ly, u.scheme,
                               ppc_bucket(u.num_particles, (int64_t)u.domain.Nx * u.domain.Ny));
    };
    std::stable_sort(e->user_of.begin(), e->user_of.end(), [&](int a, int b) { return cls(a) < cls(b); });
    e->pos_of.resize((size_t)M);
    for (int i = 0; i < M; ++i) e->pos_of[(size_t)e->user_of[(size_t)i]] = i;

    // Layout.
    e->mem.resize((size_t)M);
    e->pchunk.assign((size_t)M + 1, 0);
    e->gblk.assign((size_t)M + 1, 0);
    e->sblk.assign((size_t)M + 1, 0);
    int64_t slots = 0, keys = 0, G = 0, S = 0, tab = 0, xt = 0;
    for (int i = 0; i < M; ++i) {
        const pic_ensemble_member& u = members[e->user_of[(size_t)i]];
        EnsMember& m = e->mem[(size_t)i];
        m = EnsMember{};
        m.d = make_dev_domain(u.domain);
        m.scheme = u.scheme;
        m.user = e->user_of[(size_t)i];
        m.n = u.num_particles;
        m.init.kind = u.init.kind;
        m.init.seed = u.init.seed;
        m.init.amp = u.init.amp;
        m.init.pert = u.init.pert;
        m.init.vth = u.init.vth;
        m.init.vmax = u.init.vmax;
        m.init.mode = u.init.mode;
        m.init.qp = u.init.q;
        m.init.mp = u.init.m;
        const int hx = u.domain.Nx / 2 + 1;
        m.G = (int64_t)u.domain.Nx * u.domain.Ny;
        m.pbase = slots;
        m.cbase = keys;
        m.gbase = G;
        m.sbase = S;
        m.tbase = tab;
        m.xbase = xt;
        const int64_t chunks = (m.n + kEnsChunk - 1) / kEnsChunk;
        slots += chunks * kEnsChunk;
        keys += m.G + 1;
        G += m.G;
        S += (int64_t)hx * u.domain.Ny;
        tab += hx + u.domain.Ny;
        xt += u.domain.Nx;
        e->pchunk[(size_t)i + 1] = e->pchunk[(size_t)i] + chunks;
        e->gblk[(size_t)i + 1] = e->gblk[(size_t)i] + (m.G + 255) / 256;
        e->sblk[(size_t)i + 1] = e->sblk[(size_t)i] + ((int64_t)hx * u.domain.Ny + 255) / 256;
        // Members share the persistent launch, so only the work per tile
        // matters: at most 32768 particles per tile.
        const int* dt = kDefaultTiling[m.scheme];
        int tx = dt[0], ty = dt[1];
        default_tile_shape(m.scheme, m.d.Nx, m.d.Ny, m.n, 1, 32768, 512, tx, ty);
        m.g = make_tile_geom(m.d.Nx, m.d.Ny, tx, ty);
        m.ppl = dt[2];
    }
    if (slots + 64 > 0x7fffffc0ll) return fail(PIC_ERR_INVALID_ARGUMENT, "ensemble above 2^31 - 64 particle slots");
    if (keys >= (1ll << 30)) return fail(PIC_ERR_INVALID_ARGUMENT, "ensemble above 2^30 cells");
    e->nslots = slots;
    e->nchunks = e->pchunk[(size_t)M];
    e->keys_total = keys;
    e->Gtot = G;
    e->Stot = S;
    e->gblocks = e->gblk[(size_t)M];
    e->sblocks = e->sblk[(size_t)M];
    e->tab_total = tab;
    e->x_total = xt;
    // FFT classes: runs of equal grid shape.
    for (int i = 0; i < M; ++i) {
        const EnsMember& m = e->mem[(size_t)i];
        if (e->classes.empty() || e->mem[(size_t)e->classes.back().first].d.Nx != m.d.Nx ||
            e->mem[(size_t)e->classes.back().first].d.Ny != m.d.Ny) {
            FftClass c;
            c.first = i;
            e->classes.push_back(c);
        }
        ++e->classes.back().count;
    }

    int dev = 0;
    hipDeviceProp_t prop;
    PIC_HIP_TRY(hipGetDevice(&dev));
    PIC_HIP_TRY(hipGetDeviceProperties(&prop, dev));
    e->ncu = std::max(1, prop.multiProcessorCount);

    PIC_HIP_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    const size_t cap = (size_t)slots + 64;  // float4 over-read pad
    e->cap = cap;
    for (int i = 0; i < kStreams; ++i) {
        PIC_HIP_TRY(dmalloc(&e->s[i], cap));
        PIC_HIP_TRY(dmalloc(&e->alt[i], cap));
        PIC_HIP_TRY(hipMemsetAsync(e->s[i], 0, cap * sizeof(float), e->stream));
        PIC_HIP_TRY(hipMemsetAsync(e->alt[i], 0, cap * sizeof(float), e->stream));
    }
    PIC_HIP_TRY(dmalloc(&e->sb.keys, cap));
    PIC_HIP_TRY(dmalloc(&e->sb.keys_alt, cap));
    // no permutation array: the radix passes move the streams
    e->sb.hist_cap = sort_hist_elems((int64_t)cap);
    PIC_HIP_TRY(dmalloc(&e->sb.hist, e->sb.hist_cap));
    e->sb.scan_tmp_cap = sort_scan_tmp_elems(e->sb.hist_cap);
    PIC_HIP_TRY(dmalloc(&e->sb.scan_tmp, e->sb.scan_tmp_cap));
    PIC_HIP_TRY(dmalloc(&e->sb.cell_off, (size_t)keys + 1));
    PIC_HIP_TRY(dmalloc(&e->sort_part, 2 * (size_t)std::max<int64_t>(e->nchunks, 1)));
    PIC_HIP_TRY(dmalloc(&e->permute_mode, 1));
    PIC_HIP_TRY(dmalloc(&e->sort_rec, 2 * (size_t)cap));
    PIC_HIP_TRY(dmalloc(&e->acc64, (size_t)G));
    PIC_HIP_TRY(hipMemsetAsync(e->acc64, 0, (size_t)G * sizeof(double), e->stream));
    float** grids[] = {&e->rho, &e->Ex, &e->Ey, &e->fft_in};
    for (float** g : grids) {
        PIC_HIP_TRY(dmalloc(g, (size_t)G));
        PIC_HIP_TRY(hipMemsetAsync(*g, 0, (size_t)G * sizeof(float), e->stream));
    }
    PIC_HIP_TRY(dmalloc(&e->spec, (size_t)S));
    PIC_HIP_TRY(dmalloc(&e->spec_ex, (size_t)S));
    PIC_HIP_TRY(dmalloc(&e->spec_ey, (size_t)S));
    PIC_HIP_TRY(dmalloc(&e->status, 1));
    PIC_HIP_TRY(hipMemsetAsync(e->status, 0, sizeof(int), e->stream));
    PIC_HIP_TRY(dmalloc(&e->counters, 16));
    PIC_HIP_TRY(dmalloc(&e->diag_counters, (size_t)M));
    PIC_HIP_TRY(dmalloc(&e->push_flag, 1));
    PIC_HIP_TRY(dmalloc(&e->rho_part, (size_t)e->gblocks));
    PIC_HIP_TRY(dmalloc(&e->field_part, (size_t)3 * e->gblocks));
    PIC_HIP_TRY(dmalloc(&e->part_p, (size_t)3 * std::max<int64_t>(e->nchunks, 1)));
    PIC_HIP_TRY(hipMemsetAsync(e->part_p, 0, (size_t)3 * std::max<int64_t>(e->nchunks, 1) * sizeof(double),
                               e->stream));
    PIC_HIP_TRY(dmalloc(&e->rho_sum, (size_t)M));
    PIC_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&e->h_status), sizeof(int) * 4, hipHostMallocDefault));
    PIC_HIP_TRY(dmalloc(&e->d_mem, (size_t)M));
    PIC_HIP_TRY(dmalloc(&e->d_pchunk, (size_t)M + 1));
    PIC_HIP_TRY(dmalloc(&e->d_gblk, (size_t)M + 1));
    PIC_HIP_TRY(dmalloc(&e->d_sblk, (size_t)M + 1));
    PIC_HIP_TRY(hipMemcpy(e->d_pchunk, e->pchunk.data(), (M + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    PIC_HIP_TRY(hipMemcpy(e->d_gblk, e->gblk.data(), (M + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    PIC_HIP_TRY(hipMemcpy(e->d_sblk, e->sblk.data(), (M + 1) * sizeof(int64_t), hipMemcpyHostToDevice));

    // Per-member spectral tables (poisson_fft.cpp:23-35 k^2; forward-difference
    // symbols (e^{i k d} - 1)/d) and mode-1 cos / sin at cell centres
    // (extractFourierModeAmplitude, validation.cpp:99-119), in FP64.
    std::vector<double> k2((size_t)tab), cs((size_t)xt), sn((size_t)xt);
    std::vector<float2> sym((size_t)tab);
    for (const EnsMember& m : e->mem) {
        const int nx = m.d.Nx, ny = m.d.Ny, hx = nx / 2 + 1;
        for (int i = 0; i < hx; ++i) {
            const double kx = 2.0 * M_PI * ((i <= nx / 2) ? i : i - nx) / m.d.Lx;
            k2[(size_t)(m.tbase + i)] = kx * kx;
            const double th = 2.0 * M_PI * (double)i / nx;
            sym[(size_t)(m.tbase + i)] =
                make_float2((float)((std::cos(th) - 1.0) / m.d.dx), (float)(std::sin(th) / m.d.dx));
        }
        for (int j = 0; j < ny; ++j) {
            const double ky = 2.0 * M_PI * ((j <= ny / 2) ? j : j - ny) / m.d.Ly;
            k2[(size_t)(m.tbase + hx + j)] = ky * ky;
            const double th = 2.0 * M_PI * (double)j / ny;
            sym[(size_t)(m.tbase + hx + j)] =
                make_float2((float)((std::cos(th) - 1.0) / m.d.dy), (float)(std::sin(th) / m.d.dy));
        }
        const double k = 2.0 * M_PI / m.d.Lx;
        for (int i = 0; i < nx; ++i) {
            const double x = (static_cast<double>(i) + 0.5) * m.d.dx;
            cs[(size_t)(m.xbase + i)] = std::cos(k * x);
            sn[(size_t)(m.xbase + i)] = std::sin(k * x);
        }
    }
    PIC_HIP_TRY(dmalloc(&e->k2, (size_t)tab));
    PIC_HIP_TRY(dmalloc(&e->sym, (size_t)tab));
    PIC_HIP_TRY(dmalloc(&e->cosk, (size_t)xt));
    PIC_HIP_TRY(dmalloc(&e->sink, (size_t)xt));
    PIC_HIP_TRY(hipMemcpy(e->k2, k2.data(), tab * sizeof(double), hipMemcpyHostToDevice));
    PIC_HIP_TRY(hipMemcpy(e->sym, sym.data(), tab * sizeof(float2), hipMemcpyHostToDevice));
    PIC_HIP_TRY(hipMemcpy(e->cosk, cs.data(), xt * sizeof(double), hipMemcpyHostToDevice));
    PIC_HIP_TRY(hipMemcpy(e->sink, sn.data(), xt * sizeof(double), hipMemcpyHostToDevice));
    pic_status st = rebuild_work(e.get());
    if (st != PIC_OK) return st;
    PIC_HIP_TRY(hipStreamSynchronize(e->stream));
    *out = e.release();
    return PIC_OK;
}

pic_status pic_hip_ensemble_destroy(pic_hip_ensemble* ens) {
    delete ens;
    return PIC_OK;
}

int32_t pic_hip_ensemble_count(pic_hip_ensemble* ens) { return ens ? ens->M : 0; }

pic_status pic_hip_ensemble_sync(pic_hip_ensemble* ens) {
    if (check_ens(ens) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    PIC_HIP_TRY(hipMemcpyAsync(ens->h_status, ens->status, sizeof(int), hipMemcpyDeviceToHost, ens->stream));
    PIC_HIP_TRY(hipStreamSynchronize(ens->stream));
    if (ens->h_status[0] != 0) {
        const int st = ens->h_status[0];
        PIC_HIP_TRY(hipMemsetAsync(ens->status, 0, sizeof(int), ens->stream));
        PIC_HIP_TRY(hipStreamSynchronize(ens->stream));
        if (st & 1)
            return fail(PIC_ERR_UNSUPPORTED,
                        "Esirkepov deposit: a particle moved more than 15 cells in one step (not supported)");
        return fail(PIC_ERR_HIP, "device status flag set");
    }
    return PIC_OK;
}

pic_status pic_hip_ensemble_init(pic_hip_ensemble* ens) {
    if (check_ens(ens) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    pic_status st = ensure_work(ens);  // member table on the device
    if (st != PIC_OK) return st;
    launch_ens_init(ens->particles(), ens->tables(), ens->nchunks, ens->stream);
    PIC_HIP_TRY(hipGetLastError());
    ens->sorted = false;
    return PIC_OK;
}

pic_status pic_hip_ensemble_upload(pic_hip_ensemble* ens, int32_t member, const float* x, const float* y,
                                   const float* vx, const float* vy, const float* q, const float* m) {
    if (check_member(ens, member) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    const EnsMember& mm = ens->mem[(size_t)ens->pos_of[(size_t)member]];
    if (mm.n > 0 && (!x || !y || !q)) return fail(PIC_ERR_INVALID_ARGUMENT, "x, y and q are required");
    const float* src[kStreams] = {x, y, vx, vy, q, m};
    for (int i = 0; i < kStreams && mm.n > 0; ++i) {
        float* dst = ens->s[i] + mm.pbase;
        if (src[i])
            PIC_HIP_TRY(hipMemcpyAsync(dst, src[i], (size_t)mm.n * sizeof(float), hipMemcpyHostToDevice, ens->stream));
        else
            fill_async(dst, (size_t)mm.n, i == 5 ? 1.0f : 0.0f, ens->stream);
    }
    ens->sorted = false;
    PIC_HIP_TRY(hipStreamSynchronize(ens->stream));
    return PIC_OK;
}

pic_status pic_hip_ensemble_download(pic_hip_ensemble* ens, int32_t member, float* x, float* y, float* vx, float* vy,
                                     float* q, float* m) {
    if (check_member(ens, member) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    const EnsMember& mm = ens->mem[(size_t)ens->pos_of[(size_t)member]];
    float* dst[kStreams] = {x, y, vx, vy, q, m};
    for (int i = 0; i < kStreams; ++i)
        if (dst[i] && mm.n > 0)
            PIC_HIP_TRY(hipMemcpyAsync(dst[i], ens->s[i] + mm.pbase, (size_t)mm.n * sizeof(float),
                                       hipMemcpyDeviceToHost, ens->stream));
    PIC_HIP_TRY(hipStreamSynchronize(ens->stream));
    return PIC_OK;
}

pic_status pic_hip_ensemble_sort(pic_hip_ensemble* ens) {
    if (check_ens(ens) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    pic_status st = ensure_work(ens);
    if (st != PIC_OK) return st;
    return enqueue_sort(ens, false);
}

pic_status pic_hip_ensemble_deposit(pic_hip_ensemble* ens, double dt) {
    if (check_ens(ens) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!std::isfinite(dt)) return fail(PIC_ERR_INVALID_ARGUMENT, "dt must be finite");
    if (!ens->sorted) return fail(PIC_ERR_NOT_SORTED, "the ensemble deposit needs a cell order: call pic_hip_ensemble_sort");
    pic_status st = ensure_work(ens);
    if (st != PIC_OK) return st;
    return enqueue_deposit(ens, (float)dt, 0, ens->gblocks);
}

pic_status pic_hip_ensemble_get_rho(pic_hip_ensemble* ens, int32_t member, float* rho) {
    if (check_member(ens, member) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!rho) return fail(PIC_ERR_INVALID_ARGUMENT, "null rho");
    const EnsMember& mm = ens->mem[(size_t)ens->pos_of[(size_t)member]];
    PIC_HIP_TRY(hipMemcpyAsync(rho, ens->rho + mm.gbase, (size_t)mm.G * sizeof(float), hipMemcpyDeviceToHost,
                               ens->stream));
    return pic_hip_ensemble_sync(ens);
}

pic_status pic_hip_ensemble_poisson(pic_hip_ensemble* ens) {
    if (check_ens(ens) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    pic_status st = ensure_work(ens);
    if (st == PIC_OK) st = ensure_fft(ens);
    if (st != PIC_OK) return st;
    return enqueue_poisson(ens);
}

pic_status pic_hip_ensemble_get_fields(pic_hip_ensemble* ens, int32_t member, float* Ex, float* Ey) {
    if (check_member(ens, member) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    const EnsMember& mm = ens->mem[(size_t)ens->pos_of[(size_t)member]];
    const size_t bytes = (size_t)mm.G * sizeof(float);
    if (Ex) PIC_HIP_TRY(hipMemcpyAsync(Ex, ens->Ex + mm.gbase, bytes, hipMemcpyDeviceToHost, ens->stream));
    if (Ey) PIC_HIP_TRY(hipMemcpyAsync(Ey, ens->Ey + mm.gbase, bytes, hipMemcpyDeviceToHost, ens->stream));
    return pic_hip_ensemble_sync(ens);
}

pic_status pic_hip_ensemble_set_tiling(pic_hip_ensemble* ens, int32_t member, int tile_x, int tile_y,
                                       int particles_per_lane) {
    if (check_ens(ens) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (member < -1 || member >= ens->M) return fail(PIC_ERR_INVALID_ARGUMENT, "member index out of range");
    PIC_HIP_TRY(hipStreamSynchronize(ens->stream));  // the member table is rewritten
    for (int u = 0; u < ens->M; ++u) {
        if (member >= 0 && u != member) continue;
        const pic_status st = set_member_tiling(ens, ens->pos_of[(size_t)u], tile_x, tile_y, particles_per_lane);
        if (st != PIC_OK) return st;
    }
    return PIC_OK;
}

pic_status pic_hip_ensemble_set_tile_halo(pic_hip_ensemble* ens, int cells) {
    if (check_ens(ens) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (cells < 0 || cells > kMaxDriftHalo)
        return fail(PIC_ERR_INVALID_ARGUMENT, "tile halo must be 0.." + std::to_string(kMaxDriftHalo) + " cells");
    PIC_HIP_TRY(hipStreamSynchronize(ens->stream));
    const int old = ens->halo;
    ens->halo = cells;
    for (int i = 0; i < ens->M; ++i) {
        const EnsMember& m = ens->mem[(size_t)i];
        const pic_status st = set_member_tiling(ens, i, m.g.T, m.g.R, m.ppl);
        if (st != PIC_OK) {
            ens->halo = old;
            for (int j = 0; j < ens->M; ++j) {
                const EnsMember& mj = ens->mem[(size_t)j];
                (void)set_member_tiling(ens, j, mj.g.T, mj.g.R, mj.ppl);
            }
            return st;
        }
    }
    return PIC_OK;
}

pic_status pic_hip_ensemble_get_tiling(pic_hip_ensemble* ens, int32_t member, int* tile_x, int* tile_y,
                                       int* particles_per_lane) {
    if (check_member(ens, member) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    const EnsMember& mm = ens->mem[(size_t)ens->pos_of[(size_t)member]];
    if (tile_x) *tile_x = mm.g.T;
    if (tile_y) *tile_y = mm.g.R;
    if (particles_per_lane) *particles_per_lane = mm.ppl;
    return PIC_OK;
}

pic_status pic_hip_ensemble_autotune(pic_hip_ensemble* ens, const char* cache_path, int reps) {
    if (check_ens(ens) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!ens->sorted) return fail(PIC_ERR_NOT_SORTED, "autotune times the tiled deposit: call pic_hip_ensemble_sort");
    if (reps < 1) reps = 5;
    PIC_HIP_TRY(hipStreamSynchronize(ens->stream));
    const auto cached = tuning_cache_read(cache_path);
    EventPair ev;
    PIC_HIP_TRY(hipEventCreate(&ev.a));
    PIC_HIP_TRY(hipEventCreate(&ev.b));
    // Median device time of the deposit of members [m0, m1) (the whole ensemble
    // by default) with the current tilings.
    auto time_deposit = [&](int m0, int m1, double& med) -> pic_status {
        pic_status st = rebuild_work(ens, m0, m1);
        if (st != PIC_OK) return st;
        const int64_t b0 = ens->gblk[(size_t)m0], nb = ens->gblk[(size_t)m1] - b0;
        if ((st = enqueue_deposit(ens, 0.0f, b0, nb)) != PIC_OK) return st;  // warm-up
        std::vector<double> t;
        for (int r = 0; r < reps; ++r) {
            PIC_HIP_TRY(hipEventRecord(ev.a, ens->stream));
            if ((st = enqueue_deposit(ens, 0.0f, b0, nb)) != PIC_OK) return st;
            PIC_HIP_TRY(hipEventRecord(ev.b, ens->stream));
            PIC_HIP_TRY(hipEventSynchronize(ev.b));
            float ms = 0.0f;
            PIC_HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
            t.push_back(ms);
        }
        med = median_of(t);
        return PIC_OK;
    };
    // The tilings before tuning, kept if the tuned set is not faster as a whole
    // (classes are timed alone, and run together in the persistent kernels).
    std::vector<TilingChoice> before;
    for (const EnsMember& m : ens->mem) before.push_back({m.g.T, m.g.R, m.ppl});
    double before_ms = 0.0;
    if (pic_status st0 = time_deposit(0, ens->M, before_ms); st0 != PIC_OK) return st0;
    // Classes: runs of members with equal grid, scheme and particles per cell bucket.
    int i0 = 0;
    while (i0 < ens->M) {
        const EnsMember& a = ens->mem[(size_t)i0];
        const int64_t bucket = ppc_bucket(a.n, a.G);
        int i1 = i0 + 1;
        while (i1 < ens->M) {
            const EnsMember& b = ens->mem[(size_t)i1];
            if (b.d.Nx != a.d.Nx || b.d.Ny != a.d.Ny || b.d.Lx != a.d.Lx || b.d.Ly != a.d.Ly ||
                b.scheme != a.scheme || ppc_bucket(b.n, b.G) != bucket)
                break;
            ++i1;
        }
        const std::string key = tuning_key(a.scheme, a.d.Nx, a.d.Ny, bucket, ens->halo);
        auto apply = [&](const TilingChoice& c) -> pic_status {
            for (int i = i0; i < i1; ++i) {
                const pic_status st = set_member_tiling(ens, i, c.tx, c.ty, c.ppl);
                if (st != PIC_OK) return st;
            }
            return PIC_OK;
        };
        auto hit = cached.find(key);
        if (hit != cached.end() && apply(hit->second) == PIC_OK) {
            i0 = i1;
            continue;
        }
        TilingChoice best{};
        double best_ms = -1.0;
        for (const TilingChoice& c : tiling_candidates(a.d.Nx, a.d.Ny, a.scheme == PIC_SCHEME_ESIRKEPOV, ens->halo)) {
            if (apply(c) != PIC_OK) continue;
            double med = 0.0;
            if (pic_status st = time_deposit(i0, i1, med); st != PIC_OK) return st;
            if (best_ms < 0.0 || med < best_ms) {
                best_ms = med;
                best = c;
            }
        }
        if (best_ms < 0.0) return fail(PIC_ERR_INVALID_ARGUMENT, "no tiling candidate fits member grid " + key);
        pic_status st = apply(best);
        if (st != PIC_OK) return st;
        if (!tuning_cache_store(cache_path, key, best))
            return fail(PIC_ERR_INVALID_ARGUMENT, std::string("cannot write tuning cache ") + cache_path);
        i0 = i1;
    }
    double after_ms = 0.0;
    pic_status st = time_deposit(0, ens->M, after_ms);
    if (st != PIC_OK) return st;
    if (after_ms > before_ms) {
        for (int i = 0; i < ens->M; ++i)
            if ((st = set_member_tiling(ens, i, before[(size_t)i].tx, before[(size_t)i].ty, before[(size_t)i].ppl)) !=
                PIC_OK)
                return st;
    }
    if ((st = rebuild_work(ens)) != PIC_OK) return st;
    return pic_hip_ensemble_sync(ens);
}

pic_status pic_hip_ensemble_run(pic_hip_ensemble* ens, const pic_ensemble_run_config* cfg, int32_t nsteps,
                                int32_t record_from, double* field_energy, double* mode1_amplitude,
                                double* kinetic_energy, double* charge_error) {
    if (check_ens(ens) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!cfg) return fail(PIC_ERR_INVALID_ARGUMENT, "null run config");
    if (nsteps < 0 || record_from < 0) return fail(PIC_ERR_INVALID_ARGUMENT, "negative step counts");
    if (cfg->sort_interval < 1) return fail(PIC_ERR_INVALID_ARGUMENT, "sort_interval must be >= 1");
    if (!std::isfinite(cfg->dt)) return fail(PIC_ERR_INVALID_ARGUMENT, "dt must be finite");
    pic_status st = ensure_work(ens);
    if (st == PIC_OK) st = ensure_fft(ens);
    if (st != PIC_OK) return st;
    const int M = ens->M;
    const int cap = nsteps > record_from ? nsteps - record_from : 0;
    double* dev_rec = nullptr;
    PIC_HIP_TRY(dmalloc(&dev_rec, (size_t)4 * M * std::max(cap, 1)));
    std::unique_ptr<double, hipError_t (*)(void*)> rec_guard(dev_rec, hipFree);
    const size_t plane = (size_t)M * cap;
    StepRecords rec{field_energy ? dev_rec : nullptr, mode1_amplitude ? dev_rec + plane : nullptr,
                    kinetic_energy ? dev_rec + 2 * plane : nullptr, charge_error ? dev_rec + 3 * plane : nullptr};
    const float dt = (float)cfg->dt;
    launch_set_ints(ens->diag_counters, M, 0, ens->stream);
    launch_set_ints(ens->push_flag, 1, 1, ens->stream);
    // Prologue push (the push that opens step 0); each step then ends with the
    // gather of its own fields fused with the next step's push.
    if (nsteps > 0)
        launch_ens_gather_push(ens->particles(), ens->Ex, ens->Ey, ens->tables(), ens->nchunks, dt, 0,
                               ens->push_flag, ens->part_p, ens->stream);
    if (cfg->use_graph && nsteps > 0) {
        // gs[p] sorts out of buffer set p, gp[p] steps on set p (see pic_hip_run).
        Graph gs[2], gp[2];
        const bool plain = cfg->sort_interval > 1 && nsteps > 1;
        if ((st = capture(ens, dt, true, record_from, cap, rec, gs[0])) != PIC_OK) return st;
        if (plain && (st = capture(ens, dt, false, record_from, cap, rec, gp[1])) != PIC_OK) return st;
        if ((st = capture(ens, dt, true, record_from, cap, rec, gs[1])) != PIC_OK) return st;
        if (plain && (st = capture(ens, dt, false, record_from, cap, rec, gp[0])) != PIC_OK) return st;
        int parity = 0;
        for (int step = 0; step < nsteps; ++step) {
            if (step == nsteps - 1) launch_set_ints(ens->push_flag, 1, 0, ens->stream);
            if (step % cfg->sort_interval == 0) {
                PIC_HIP_TRY(hipGraphLaunch(gs[parity].exec, ens->stream));
                parity ^= 1;
            } else {
                PIC_HIP_TRY(hipGraphLaunch(gp[parity].exec, ens->stream));
            }
        }
        if (parity)
            for (int i = 0; i < kStreams; ++i) std::swap(ens->s[i], ens->alt[i]);
        ens->sorted = true;
    } else {
        for (int step = 0; step < nsteps; ++step) {
            if (step == nsteps - 1) launch_set_ints(ens->push_flag, 1, 0, ens->stream);
            if ((st = enqueue_step(ens, dt, step % cfg->sort_interval == 0, false, record_from, cap, rec)) != PIC_OK)
                return st;
        }
    }
    double* outs[4] = {field_energy, mode1_amplitude, kinetic_energy, charge_error};
    for (int k = 0; k < 4; ++k)
        if (outs[k] && cap)
            PIC_HIP_TRY(hipMemcpyAsync(outs[k], dev_rec + (size_t)k * plane, plane * sizeof(double),
                                       hipMemcpyDeviceToHost, ens->stream));
    return pic_hip_ensemble_sync(ens);
}

pic_status pic_hip_ensemble_time(pic_hip_ensemble* ens, int what, int reps, double dt, double* mean_ms) {
    if (check_ens(ens) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (what < 0 || what > 3) return fail(PIC_ERR_INVALID_ARGUMENT, "what must be 0..3");
    if (reps < 1) return fail(PIC_ERR_INVALID_ARGUMENT, "reps must be >= 1");
    if (!std::isfinite(dt)) return fail(PIC_ERR_INVALID_ARGUMENT, "dt must be finite");
    pic_status st = ensure_work(ens);
    if (st == PIC_OK && what >= 2) st = ensure_fft(ens);
    if (st != PIC_OK) return st;
    if ((what == 0 || what == 2) && !ens->sorted)
        return fail(PIC_ERR_NOT_SORTED, "call pic_hip_ensemble_sort first");
    const float fdt = (float)dt;
    Graph g[2];
    const StepRecords none{nullptr, nullptr, nullptr, nullptr};
    int parity = 0;
    if (what >= 2) {
        launch_set_ints(ens->push_flag, 1, 1, ens->stream);
        if ((st = capture(ens, fdt, what == 3, 0, 0, none, g[0])) != PIC_OK) return st;
        if (what == 3 && (st = capture(ens, fdt, true, 0, 0, none, g[1])) != PIC_OK) return st;
    }
    auto once = [&]() -> pic_status {
        if (what == 0) return enqueue_deposit(ens, fdt, 0, ens->gblocks);
        if (what == 1) return enqueue_sort(ens, false);
        PIC_HIP_TRY(hipGraphLaunch(g[parity].exec, ens->stream));
        if (what == 3) parity ^= 1;  // each sort graph permutes into the other buffer set
        return PIC_OK;
    };
    if ((st = once()) != PIC_OK) return st;  // warm-up
    EventPair ev;
    PIC_HIP_TRY(hipEventCreate(&ev.a));
    PIC_HIP_TRY(hipEventCreate(&ev.b));
    PIC_HIP_TRY(hipEventRecord(ev.a, ens->stream));
    for (int r = 0; r < reps; ++r)
        if ((st = once()) != PIC_OK) return st;
    PIC_HIP_TRY(hipEventRecord(ev.b, ens->stream));
    PIC_HIP_TRY(hipEventSynchronize(ev.b));
    float ms = 0.0f;
    PIC_HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
    if (mean_ms) *mean_ms = ms / reps;
    if (parity)
        for (int i = 0; i < kStreams; ++i) std::swap(ens->s[i], ens->alt[i]);
    if (what >= 2) ens->sorted = true;
    return pic_hip_ensemble_sync(ens);
}

}  // extern "C"
