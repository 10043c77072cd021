// C ABI of the gfx950 charge-deposition library (declared in include/pic_hip.h).
//
// The context owns every device buffer of one PIC instance: the FP32 SoA
// particle streams (and their sort double-buffers), the cell-sort scratch, the
// FP32 rho / phi / Ex / Ey grids, the tile partials of the LDS deposit and the
// hipFFT plans.  All work is enqueued on the context's stream.

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pic_hip.h"
#include "pic_common.hpp"
#include "pic_kernels.hpp"

using namespace pichip;

namespace {

thread_local std::string g_last_error;

pic_status fail(pic_status s, const std::string& msg) {
    g_last_error = msg;
    return s;
}

#define PIC_HIP_TRY(call)                                                                                   \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess) {                                                                             \
            return fail(e_ == hipErrorOutOfMemory ? PIC_ERR_OUT_OF_MEMORY : PIC_ERR_HIP,                     \
                        std::string("HIP error: ") + hipGetErrorString(e_) + " (" #call ")");                \
        }                                                                                                   \
    } while (0)

#define PIC_FFT_TRY(call)                                                                        \
    do {                                                                                         \
        hipfftResult r_ = (call);                                                                \
        if (r_ != HIPFFT_SUCCESS)                                                                \
            return fail(PIC_ERR_FFT, std::string("hipFFT error ") + std::to_string((int)r_) + " (" #call ")"); \
    } while (0)

constexpr int kStreams = 6;  // x, y, vx, vy, q, m

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

bool valid_domain(const pic_domain* d) {
    return d && d->Nx > 0 && d->Ny > 0 && std::isfinite(d->Lx) && std::isfinite(d->Ly) && d->Lx > 0.0 &&
           d->Ly > 0.0 && (int64_t)d->Nx * d->Ny < (1ll << 30);
}

DevDomain make_dev_domain(const pic_domain& d) {
    DevDomain dd;
    dd.Nx = d.Nx;
    dd.Ny = d.Ny;
    dd.Lx = d.Lx;
    dd.Ly = d.Ly;
    dd.dx = d.Lx / d.Nx;  // Domain::updateDerived, domain.hpp:17-20
    dd.dy = d.Ly / d.Ny;
    dd.inv_dx = (float)(1.0 / dd.dx);
    dd.inv_dy = (float)(1.0 / dd.dy);
    dd.inv_vol = (float)(1.0 / (dd.dx * dd.dy));
    return dd;
}

}  // namespace

struct pic_hip_ctx {
    pic_domain dom{};
    DevDomain dev{};
    int64_t G = 0;
    size_t cap = 0;  // particle capacity (rounded, plus float4 over-read pad)
    size_t n = 0;
    hipStream_t stream = nullptr;
    float* s[kStreams] = {};
    float* alt[kStreams] = {};
    SortBuffers sb{};
    bool order_valid = false;
    float* rho = nullptr;
    float* overflow = nullptr;
    int* ovf_count = nullptr;
    int* status = nullptr;
    float* partial = nullptr;
    size_t partial_cap = 0;
    TileGeom geom{};
    int ppl = 8;
    float* phi = nullptr;
    float* Ex = nullptr;
    float* Ey = nullptr;
    float* fft_in = nullptr;
    float2* spec = nullptr;
    hipfftHandle plan_r2c = 0, plan_c2r = 0;
    bool fft_ready = false;
    double* diag_partials = nullptr;
    double* diag_out = nullptr;
    int diag_blocks = 0;
    int* h_status = nullptr;     // pinned
    double* h_diag = nullptr;    // pinned
    std::vector<hipEvent_t> prof_events;  // pairs (begin, end)
    int prof_max = 0, prof_used = 0;
    bool profiling = false;

    ~pic_hip_ctx() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipEvent_t e : prof_events) (void)hipEventDestroy(e);
        if (plan_r2c) (void)hipfftDestroy(plan_r2c);
        if (plan_c2r) (void)hipfftDestroy(plan_c2r);
        for (int i = 0; i < kStreams; ++i) {
            (void)hipFree(s[i]);
            (void)hipFree(alt[i]);
        }
        (void)hipFree(sb.keys);
        (void)hipFree(sb.keys_alt);
        (void)hipFree(sb.perm);
        (void)hipFree(sb.perm_alt);
        (void)hipFree(sb.hist);
        (void)hipFree(sb.scan_tmp);
        (void)hipFree(sb.cell_off);
        (void)hipFree(rho);
        (void)hipFree(overflow);
        (void)hipFree(ovf_count);
        (void)hipFree(status);
        (void)hipFree(partial);
        (void)hipFree(phi);
        (void)hipFree(Ex);
        (void)hipFree(Ey);
        (void)hipFree(fft_in);
        (void)hipFree(spec);
        (void)hipFree(diag_partials);
        (void)hipFree(diag_out);
        (void)hipHostFree(h_status);
        (void)hipHostFree(h_diag);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <typename T>
hipError_t dmalloc(T** p, size_t elems) {
    *p = nullptr;
    return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(elems, 1) * sizeof(T));
}

pic_status ensure_partial(pic_hip_ctx* c) {
    const size_t need = (size_t)c->geom.ntiles * c->geom.wsize;
    if (need <= c->partial_cap) return PIC_OK;
    (void)hipFree(c->partial);
    c->partial = nullptr;
    c->partial_cap = 0;
    PIC_HIP_TRY(dmalloc(&c->partial, need));
    c->partial_cap = need;
    return PIC_OK;
}

pic_status apply_tiling(pic_hip_ctx* c, int tx, int ty, int ppl) {
    if (tx < 1 || ty < 1) return fail(PIC_ERR_INVALID_ARGUMENT, "tile sizes must be >= 1");
    if (ppl != 4 && ppl != 8 && ppl != 16) return fail(PIC_ERR_INVALID_ARGUMENT, "particles_per_lane must be 4, 8 or 16");
    TileGeom g = make_tile_geom(c->dom.Nx, c->dom.Ny, tx, ty);
    const size_t lds = (size_t)4 * g.lds_per_wave * sizeof(float);
    if (lds > 64 * 1024) return fail(PIC_ERR_INVALID_ARGUMENT, "tile too large for LDS (4 windows must fit 64 KiB)");
    c->geom = g;
    c->ppl = ppl;
    return ensure_partial(c);
}

pic_status ensure_fft(pic_hip_ctx* c) {
    if (c->fft_ready) return PIC_OK;
    const int hx = c->dom.Nx / 2 + 1;
    PIC_HIP_TRY(dmalloc(&c->phi, (size_t)c->G));
    PIC_HIP_TRY(dmalloc(&c->fft_in, (size_t)c->G));
    PIC_HIP_TRY(dmalloc(&c->spec, (size_t)c->dom.Ny * hx));
    PIC_FFT_TRY(hipfftPlan2d(&c->plan_r2c, c->dom.Ny, c->dom.Nx, HIPFFT_R2C));
    PIC_FFT_TRY(hipfftPlan2d(&c->plan_c2r, c->dom.Ny, c->dom.Nx, HIPFFT_C2R));
    PIC_FFT_TRY(hipfftSetStream(c->plan_r2c, c->stream));
    PIC_FFT_TRY(hipfftSetStream(c->plan_c2r, c->stream));
    c->fft_ready = true;
    return PIC_OK;
}

pic_status check_ctx(pic_hip_ctx* c) {
    if (!c) return fail(PIC_ERR_INVALID_ARGUMENT, "null context");
    return PIC_OK;
}

}  // namespace

extern "C" {

const char* pic_hip_last_error(void) { return g_last_error.c_str(); }
const char* pic_hip_version(void) { return "pic_hip 0.1 gfx950"; }

pic_status pic_hip_create(const pic_domain* domain, size_t capacity, pic_hip_ctx** out) {
    if (!out) return fail(PIC_ERR_INVALID_ARGUMENT, "null output pointer");
    *out = nullptr;
    if (!valid_domain(domain)) return fail(PIC_ERR_INVALID_ARGUMENT, "invalid domain (need Nx, Ny > 0, L > 0, Nx*Ny < 2^30)");
    if (capacity > (size_t)0x7fffffc0) return fail(PIC_ERR_INVALID_ARGUMENT, "capacity above 2^31 - 64 particles");
    std::unique_ptr<pic_hip_ctx> c(new pic_hip_ctx());
    c->dom = *domain;
    c->dev = make_dev_domain(*domain);
    c->G = (int64_t)domain->Nx * domain->Ny;
    c->cap = round_up(std::max<size_t>(capacity, 1), 64) + 64;
    PIC_HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (int i = 0; i < kStreams; ++i) {
        PIC_HIP_TRY(dmalloc(&c->s[i], c->cap));
        PIC_HIP_TRY(dmalloc(&c->alt[i], c->cap));
        PIC_HIP_TRY(hipMemsetAsync(c->s[i], 0, c->cap * sizeof(float), c->stream));
        PIC_HIP_TRY(hipMemsetAsync(c->alt[i], 0, c->cap * sizeof(float), c->stream));
    }
    PIC_HIP_TRY(dmalloc(&c->sb.keys, c->cap));
    PIC_HIP_TRY(dmalloc(&c->sb.keys_alt, c->cap));
    PIC_HIP_TRY(dmalloc(&c->sb.perm, c->cap));
    PIC_HIP_TRY(dmalloc(&c->sb.perm_alt, c->cap));
    c->sb.hist_cap = sort_hist_elems((int64_t)c->cap);
    PIC_HIP_TRY(dmalloc(&c->sb.hist, c->sb.hist_cap));
    c->sb.scan_tmp_cap = sort_scan_tmp_elems(c->sb.hist_cap);
    PIC_HIP_TRY(dmalloc(&c->sb.scan_tmp, c->sb.scan_tmp_cap));
    PIC_HIP_TRY(dmalloc(&c->sb.cell_off, (size_t)c->G + 1));
    PIC_HIP_TRY(dmalloc(&c->rho, (size_t)c->G));
    PIC_HIP_TRY(dmalloc(&c->overflow, (size_t)c->G));
    PIC_HIP_TRY(hipMemsetAsync(c->rho, 0, (size_t)c->G * sizeof(float), c->stream));
    PIC_HIP_TRY(hipMemsetAsync(c->overflow, 0, (size_t)c->G * sizeof(float), c->stream));
    PIC_HIP_TRY(dmalloc(&c->ovf_count, 1));
    PIC_HIP_TRY(dmalloc(&c->status, 1));
    PIC_HIP_TRY(hipMemsetAsync(c->status, 0, sizeof(int), c->stream));
    PIC_HIP_TRY(dmalloc(&c->Ex, (size_t)c->G));
    PIC_HIP_TRY(dmalloc(&c->Ey, (size_t)c->G));
    PIC_HIP_TRY(hipMemsetAsync(c->Ex, 0, (size_t)c->G * sizeof(float), c->stream));
    PIC_HIP_TRY(hipMemsetAsync(c->Ey, 0, (size_t)c->G * sizeof(float), c->stream));
    c->diag_blocks = 1024;
    PIC_HIP_TRY(dmalloc(&c->diag_partials, (size_t)5 * c->diag_blocks));
    PIC_HIP_TRY(dmalloc(&c->diag_out, 8));
    PIC_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&c->h_status), sizeof(int) * 4, hipHostMallocDefault));
    PIC_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&c->h_diag), sizeof(double) * 8, hipHostMallocDefault));
    pic_status st = apply_tiling(c.get(), 32, 8, 8);
    if (st != PIC_OK) return st;
    PIC_HIP_TRY(hipStreamSynchronize(c->stream));
    *out = c.release();
    return PIC_OK;
}

pic_status pic_hip_destroy(pic_hip_ctx* ctx) {
    delete ctx;
    return PIC_OK;
}

void* pic_hip_stream(pic_hip_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

size_t pic_hip_count(pic_hip_ctx* ctx) { return ctx ? ctx->n : 0; }

pic_status pic_hip_sync(pic_hip_ctx* ctx) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    PIC_HIP_TRY(hipMemcpyAsync(ctx->h_status, ctx->status, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->h_status[0] != 0) {
        const int st = ctx->h_status[0];
        PIC_HIP_TRY(hipMemsetAsync(ctx->status, 0, sizeof(int), ctx->stream));
        PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (st & 1)
            return fail(PIC_ERR_UNSUPPORTED,
                        "Esirkepov deposit: a particle moved more than 15 cells in one step (not supported)");
        return fail(PIC_ERR_HIP, "device status flag set");
    }
    return PIC_OK;
}

pic_status pic_hip_set_tiling(pic_hip_ctx* ctx, int tile_x, int tile_y, int particles_per_lane) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    return apply_tiling(ctx, tile_x, tile_y, particles_per_lane);
}

pic_status pic_hip_get_tiling(pic_hip_ctx* ctx, int* tile_x, int* tile_y, int* ppl) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (tile_x) *tile_x = ctx->geom.T;
    if (tile_y) *tile_y = ctx->geom.R;
    if (ppl) *ppl = ctx->ppl;
    return PIC_OK;
}

pic_status pic_hip_upload(pic_hip_ctx* ctx, const float* x, const float* y, const float* vx, const float* vy,
                          const float* q, const float* m, size_t n) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (n > 0 && (!x || !y || !q)) return fail(PIC_ERR_INVALID_ARGUMENT, "x, y and q are required");
    if (n + 64 > ctx->cap) return fail(PIC_ERR_INVALID_ARGUMENT, "particle count exceeds context capacity");
    const float* src[kStreams] = {x, y, vx, vy, q, m};
    for (int i = 0; i < kStreams; ++i) {
        if (n == 0) break;
        if (src[i]) {
            PIC_HIP_TRY(hipMemcpyAsync(ctx->s[i], src[i], n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        } else {
            fill_async(ctx->s[i], n, i == 5 ? 1.0f : 0.0f, ctx->stream);
        }
    }
    ctx->n = n;
    ctx->order_valid = false;
    PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

pic_status pic_hip_download(pic_hip_ctx* ctx, float* x, float* y, float* vx, float* vy, float* q, float* m) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    float* dst[kStreams] = {x, y, vx, vy, q, m};
    for (int i = 0; i < kStreams; ++i)
        if (dst[i] && ctx->n)
            PIC_HIP_TRY(hipMemcpyAsync(dst[i], ctx->s[i], ctx->n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

pic_status pic_hip_init_particles(pic_hip_ctx* ctx, const pic_init_params* p, size_t n) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!p) return fail(PIC_ERR_INVALID_ARGUMENT, "null init params");
    if (n + 64 > ctx->cap) return fail(PIC_ERR_INVALID_ARGUMENT, "particle count exceeds context capacity");
    if (p->kind < 0 || p->kind > 2) return fail(PIC_ERR_INVALID_ARGUMENT, "init kind must be 0, 1 or 2");
    if (p->kind == 1 && std::fabs(p->amp) >= 1.0)
        return fail(PIC_ERR_INVALID_ARGUMENT, "density perturbation amplitude must be below 1");
    InitParams ip;
    ip.kind = p->kind;
    ip.seed = p->seed;
    ip.amp = p->amp;
    ip.pert = p->pert;
    ip.vth = p->vth;
    ip.vmax = p->vmax;
    ip.mode = p->mode;
    ip.qp = p->q;
    ip.mp = p->m;
    launch_init(ctx->s[0], ctx->s[1], ctx->s[2], ctx->s[3], ctx->s[4], ctx->s[5], (int64_t)n, ctx->dev, ip,
                ctx->stream);
    PIC_HIP_TRY(hipGetLastError());
    ctx->n = n;
    ctx->order_valid = false;
    return PIC_OK;
}

pic_status pic_hip_sort(pic_hip_ctx* ctx) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    PIC_HIP_TRY(launch_sort_keys_perm(ctx->s[0], ctx->s[1], (int64_t)ctx->n, ctx->dev, ctx->sb, ctx->stream));
    for (int i = 0; i < kStreams; ++i) {
        launch_gather_f32(ctx->s[i], ctx->alt[i], ctx->sb.perm, (int64_t)ctx->n, ctx->stream);
        std::swap(ctx->s[i], ctx->alt[i]);
    }
    PIC_HIP_TRY(hipGetLastError());
    ctx->order_valid = true;
    return PIC_OK;
}

pic_status pic_hip_sort_perm(pic_hip_ctx* ctx, int32_t* keys_sorted, int32_t* perm, int32_t* cell_offsets) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    PIC_HIP_TRY(launch_sort_keys_perm(ctx->s[0], ctx->s[1], (int64_t)ctx->n, ctx->dev, ctx->sb, ctx->stream));
    ctx->order_valid = false;  // offsets now describe an order the particles are not in
    if (keys_sorted && ctx->n)
        PIC_HIP_TRY(hipMemcpyAsync(keys_sorted, ctx->sb.keys, ctx->n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    if (perm && ctx->n)
        PIC_HIP_TRY(hipMemcpyAsync(perm, ctx->sb.perm, ctx->n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    if (cell_offsets)
        PIC_HIP_TRY(hipMemcpyAsync(cell_offsets, ctx->sb.cell_off, ((size_t)ctx->G + 1) * sizeof(int),
                                   hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

pic_status pic_hip_get_keys(pic_hip_ctx* ctx, int32_t* keys) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!keys) return fail(PIC_ERR_INVALID_ARGUMENT, "null keys");
    if (ctx->n == 0) return PIC_OK;
    launch_keys(ctx->s[0], ctx->s[1], (int64_t)ctx->n, ctx->dev, ctx->sb.keys_alt, ctx->stream);
    PIC_HIP_TRY(hipMemcpyAsync(keys, ctx->sb.keys_alt, ctx->n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

pic_status pic_hip_deposit(pic_hip_ctx* ctx, int scheme, int backend, double dt) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (scheme < 0 || scheme > 3) return fail(PIC_ERR_INVALID_ARGUMENT, "unknown scheme (0 NGP, 1 CIC, 2 TSC, 3 Esirkepov)");
    if (backend == PIC_BACKEND_CPU)
        return fail(PIC_ERR_UNSUPPORTED, "the CPU backend is served by the host library, not by pic_hip");
    if (backend != PIC_BACKEND_GPU_ATOMICS && backend != PIC_BACKEND_GPU_PRIV)
        return fail(PIC_ERR_INVALID_ARGUMENT, "unknown backend (1 GPU_Atomics, 2 GPU_Priv)");
    if (backend == PIC_BACKEND_GPU_PRIV && !ctx->order_valid)
        return fail(PIC_ERR_NOT_SORTED, "GPU_Priv deposit needs a cell order: call pic_hip_sort first");
    if (!std::isfinite(dt)) return fail(PIC_ERR_INVALID_ARGUMENT, "dt must be finite");
    DepositArgs a;
    a.x = ctx->s[0];
    a.y = ctx->s[1];
    a.vx = ctx->s[2];
    a.vy = ctx->s[3];
    a.q = ctx->s[4];
    a.n = (int64_t)ctx->n;
    a.cell_off = ctx->sb.cell_off;
    a.dom = ctx->dev;
    a.scheme = scheme;
    a.backend = backend;
    a.ppl = ctx->ppl;
    a.geom = ctx->geom;
    a.dt = (float)dt;
    a.rho = ctx->rho;
    a.partial = ctx->partial;
    a.overflow = ctx->overflow;
    a.ovf_count = ctx->ovf_count;
    a.status = ctx->status;
    a.stream = ctx->stream;
    if (ctx->profiling && ctx->prof_used < ctx->prof_max) {
        a.ev_begin = ctx->prof_events[2 * ctx->prof_used];
        a.ev_end = ctx->prof_events[2 * ctx->prof_used + 1];
        ++ctx->prof_used;
    }
    PIC_HIP_TRY(launch_deposit(a));
    return PIC_OK;
}

pic_status pic_hip_profile_begin(pic_hip_ctx* ctx, int max_launches) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (max_launches < 1 || max_launches > 100000) return fail(PIC_ERR_INVALID_ARGUMENT, "max_launches out of range");
    while ((int)ctx->prof_events.size() < 2 * max_launches) {
        hipEvent_t e;
        PIC_HIP_TRY(hipEventCreate(&e));
        ctx->prof_events.push_back(e);
    }
    ctx->prof_max = max_launches;
    ctx->prof_used = 0;
    ctx->profiling = true;
    return PIC_OK;
}

pic_status pic_hip_profile_end(pic_hip_ctx* ctx, double* total_ms, int* launches) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!ctx->profiling) return fail(PIC_ERR_INVALID_ARGUMENT, "profiling was not started");
    ctx->profiling = false;
    PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    double sum = 0.0;
    for (int i = 0; i < ctx->prof_used; ++i) {
        float ms = 0.0f;
        PIC_HIP_TRY(hipEventElapsedTime(&ms, ctx->prof_events[2 * i], ctx->prof_events[2 * i + 1]));
        sum += ms;
    }
    if (total_ms) *total_ms = sum;
    if (launches) *launches = ctx->prof_used;
    return PIC_OK;
}

pic_status pic_hip_get_rho(pic_hip_ctx* ctx, float* rho) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!rho) return fail(PIC_ERR_INVALID_ARGUMENT, "null rho");
    PIC_HIP_TRY(hipMemcpyAsync(rho, ctx->rho, (size_t)ctx->G * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    return pic_hip_sync(ctx);
}

pic_status pic_hip_get_rho_f64(pic_hip_ctx* ctx, double* rho) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!rho) return fail(PIC_ERR_INVALID_ARGUMENT, "null rho");
    std::vector<float> tmp((size_t)ctx->G);
    pic_status st = pic_hip_get_rho(ctx, tmp.data());
    if (st != PIC_OK) return st;
    for (size_t i = 0; i < tmp.size(); ++i) rho[i] = tmp[i];
    return PIC_OK;
}

pic_status pic_hip_set_rho(pic_hip_ctx* ctx, const float* rho) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!rho) return fail(PIC_ERR_INVALID_ARGUMENT, "null rho");
    PIC_HIP_TRY(hipMemcpyAsync(ctx->rho, rho, (size_t)ctx->G * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

pic_status pic_hip_add_background(pic_hip_ctx* ctx, double value) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    launch_add_scalar(ctx->rho, ctx->G, (float)value, ctx->stream);
    PIC_HIP_TRY(hipGetLastError());
    return PIC_OK;
}

pic_status pic_hip_poisson(pic_hip_ctx* ctx) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    pic_status st = ensure_fft(ctx);
    if (st != PIC_OK) return st;
    const double cv = ctx->dev.dx * ctx->dev.dy;
    launch_diagnostics(ctx->rho, nullptr, nullptr, ctx->G, nullptr, nullptr, nullptr, nullptr, 0, ctx->diag_partials,
                       ctx->diag_blocks, ctx->diag_out, ctx->stream);
    launch_poisson_prep(ctx->rho, ctx->fft_in, ctx->G, 0.0, cv, ctx->diag_out, ctx->stream);
    PIC_FFT_TRY(hipfftExecR2C(ctx->plan_r2c, ctx->fft_in, reinterpret_cast<hipfftComplex*>(ctx->spec)));
    launch_poisson_scale(ctx->spec, ctx->dev, ctx->stream);
    PIC_FFT_TRY(hipfftExecC2R(ctx->plan_c2r, reinterpret_cast<hipfftComplex*>(ctx->spec), ctx->phi));
    launch_efield(ctx->phi, ctx->Ex, ctx->Ey, ctx->dev, ctx->stream);
    PIC_HIP_TRY(hipGetLastError());
    return PIC_OK;
}

pic_status pic_hip_get_fields(pic_hip_ctx* ctx, float* phi, float* Ex, float* Ey) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    const size_t bytes = (size_t)ctx->G * sizeof(float);
    if (phi) {
        if (!ctx->phi) return fail(PIC_ERR_INVALID_ARGUMENT, "phi not computed yet (call pic_hip_poisson)");
        PIC_HIP_TRY(hipMemcpyAsync(phi, ctx->phi, bytes, hipMemcpyDeviceToHost, ctx->stream));
    }
    if (Ex) PIC_HIP_TRY(hipMemcpyAsync(Ex, ctx->Ex, bytes, hipMemcpyDeviceToHost, ctx->stream));
    if (Ey) PIC_HIP_TRY(hipMemcpyAsync(Ey, ctx->Ey, bytes, hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

pic_status pic_hip_set_fields(pic_hip_ctx* ctx, const float* Ex, const float* Ey) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!Ex || !Ey) return fail(PIC_ERR_INVALID_ARGUMENT, "null field");
    const size_t bytes = (size_t)ctx->G * sizeof(float);
    PIC_HIP_TRY(hipMemcpyAsync(ctx->Ex, Ex, bytes, hipMemcpyHostToDevice, ctx->stream));
    PIC_HIP_TRY(hipMemcpyAsync(ctx->Ey, Ey, bytes, hipMemcpyHostToDevice, ctx->stream));
    PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

pic_status pic_hip_gather_push(pic_hip_ctx* ctx, double dt, int do_gather, int do_push) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!std::isfinite(dt)) return fail(PIC_ERR_INVALID_ARGUMENT, "dt must be finite");
    launch_gather_push(ctx->s[0], ctx->s[1], ctx->s[2], ctx->s[3], ctx->s[4], ctx->s[5], (int64_t)ctx->n, ctx->Ex,
                       ctx->Ey, ctx->dev, (float)dt, do_gather, do_push, ctx->stream);
    PIC_HIP_TRY(hipGetLastError());
    return PIC_OK;
}

pic_status pic_hip_diagnostics(pic_hip_ctx* ctx, pic_diagnostics* out) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!out) return fail(PIC_ERR_INVALID_ARGUMENT, "null output");
    launch_diagnostics(ctx->rho, ctx->Ex, ctx->Ey, ctx->G, ctx->s[2], ctx->s[3], ctx->s[4], ctx->s[5],
                       (int64_t)ctx->n, ctx->diag_partials, ctx->diag_blocks, ctx->diag_out, ctx->stream);
    PIC_HIP_TRY(hipMemcpyAsync(ctx->h_diag, ctx->diag_out, 5 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    const double cv = ctx->dev.dx * ctx->dev.dy;
    out->rho_sum = ctx->h_diag[0];
    out->integrated_rho = ctx->h_diag[0] * cv;
    out->field_energy = 0.5 * ctx->h_diag[1] * cv;
    out->kinetic_energy = ctx->h_diag[2];
    out->total_charge = ctx->h_diag[3];
    out->total_abs_charge = ctx->h_diag[4];
    out->charge_error =
        std::fabs(out->integrated_rho - out->total_charge) / std::max(out->total_abs_charge, 1e-12);
    return PIC_OK;
}

// ---- drop-in seam -----------------------------------------------------------

pic_status pic_hip_deposit_host(const double* x, const double* y, const double* vx, const double* vy,
                                const double* q, size_t n, const pic_domain* domain, int scheme, int backend,
                                int sorted, double dt, double* rho_out) {
    if (!valid_domain(domain)) return fail(PIC_ERR_INVALID_ARGUMENT, "invalid domain");
    if (!rho_out) return fail(PIC_ERR_INVALID_ARGUMENT, "null rho_out");
    if (n > 0 && (!x || !y || !q)) return fail(PIC_ERR_INVALID_ARGUMENT, "x, y and q are required");
    if (scheme == PIC_SCHEME_ESIRKEPOV && n > 0 && (!vx || !vy))
        return fail(PIC_ERR_INVALID_ARGUMENT, "the Esirkepov scheme needs vx and vy");
    // A per-thread context, grown on demand (the reference keeps a static
    // grow-only device pool, deposition_cuda.cu:243-273).
    thread_local std::unique_ptr<pic_hip_ctx, pic_status (*)(pic_hip_ctx*)> cache(nullptr, pic_hip_destroy);
    if (!cache || cache->dom.Nx != domain->Nx || cache->dom.Ny != domain->Ny || cache->dom.Lx != domain->Lx ||
        cache->dom.Ly != domain->Ly || cache->cap < n + 64) {
        cache.reset();
        pic_hip_ctx* c = nullptr;
        pic_status st = pic_hip_create(domain, n, &c);
        if (st != PIC_OK) return st;
        cache.reset(c);
    }
    pic_hip_ctx* c = cache.get();
    std::vector<float> buf[5];
    const double* src[5] = {x, y, vx, vy, q};
    for (int i = 0; i < 5; ++i) {
        if (!src[i]) continue;
        buf[i].resize(n);
        for (size_t k = 0; k < n; ++k) buf[i][k] = (float)src[i][k];
    }
    pic_status st = pic_hip_upload(c, n ? buf[0].data() : nullptr, n ? buf[1].data() : nullptr,
                                   buf[2].empty() ? nullptr : buf[2].data(), buf[3].empty() ? nullptr : buf[3].data(),
                                   n ? buf[4].data() : nullptr, nullptr, n);
    if (st != PIC_OK) return st;
    if (sorted || backend == PIC_BACKEND_GPU_PRIV) {
        st = pic_hip_sort(c);
        if (st != PIC_OK) return st;
    }
    st = pic_hip_deposit(c, scheme, backend, dt);
    if (st != PIC_OK) return st;
    return pic_hip_get_rho_f64(c, rho_out);
}

}  // extern "C"
