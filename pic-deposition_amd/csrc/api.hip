// C ABI of the gfx950 charge-deposition library (declared in include/pic_hip.h):
// the single-grid context (layout in context.hpp), the drop-in seam, the
// timestep driver and the tiling autotuner.

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <random>
#include <thread>
#include <vector>

#include "../../include/pic_hip.h"
#include "pic_common.hpp"
#include "pic_kernels.hpp"
#include "api_util.hpp"
#include "context.hpp"

using namespace pichip;


namespace {

pic_status ensure_partial(pic_hip_ctx* c) {
    size_t need = (size_t)c->tsc_cell_geom.ntiles * c->tsc_cell_geom.wsize;
    for (int k = 0; k < 4; ++k) need = std::max(need, (size_t)c->geom[k].ntiles * c->geom[k].wsize);
    if (need <= c->partial_cap) return PIC_OK;
    (void)hipFree(c->partial);
    c->partial = nullptr;
    c->partial_cap = 0;
    PIC_HIP_TRY(dmalloc(&c->partial, need));
    c->partial_cap = need;
    return PIC_OK;
}

pic_status apply_tiling(pic_hip_ctx* c, int scheme, int tx, int ty, int ppl, int pipe) {
    if (scheme < -1 || scheme > 3) return fail(PIC_ERR_INVALID_ARGUMENT, "scheme must be -1 (all) or 0..3");
    if (tx < 1 || ty < 1) return fail(PIC_ERR_INVALID_ARGUMENT, "tile sizes must be >= 1");
    if (ppl != 4 && ppl != 8) return fail(PIC_ERR_INVALID_ARGUMENT, "particles_per_lane must be 4 or 8");
    TileGeom g = make_tile_geom(c->dom.Nx, c->dom.Ny, tx, ty, c->halo);
    const size_t lds = (size_t)4 * g.lds_per_wave * sizeof(float);
    if (lds > kMaxDepositLds) return fail(PIC_ERR_INVALID_ARGUMENT, "tile too large for LDS (4 windows must fit 64 KiB)");
    for (int k = 0; k < 4; ++k) {
        if (scheme >= 0 && k != scheme) continue;
        c->geom[k] = g;
        c->ppl[k] = ppl;
        c->pipe[k] = pipe ? 1 : 0;
    }
    return ensure_partial(c);
}

// Tiling of TSC on the cell-keyed order (pic_hip_ctx::tsc_cell_geom): the TSC
// tiling, or while that is the create-time one without a drift halo, the
// cell-keyed long rows where the grid holds 4096 of them with at least 1024
// particles each.  Called after every change of the TSC tiling or the halo.
pic_status sync_tsc_cell(pic_hip_ctx* c) {
    c->tsc_cell_geom = c->geom[PIC_SCHEME_TSC];
    c->tsc_cell_ppl = c->ppl[PIC_SCHEME_TSC];
    c->tsc_cell_pipe = c->pipe[PIC_SCHEME_TSC];
    if (c->tiling_default[PIC_SCHEME_TSC] && c->halo == 0) {
        const int rx = kRowTilingTscCell[0], ry = kRowTilingTscCell[1];
        const int Nx = c->dom.Nx, Ny = c->dom.Ny;
        const int64_t tiles = (int64_t)((Nx + rx - 1) / rx) * ((Ny + ry - 1) / ry);
        const int64_t per_tile = c->tile_capacity * rx * ry / std::max<int64_t>(c->G, 1);
        const TileGeom g = make_tile_geom(Nx, Ny, rx, ry, 0);
        if (rx <= Nx && tiles >= 4096 && per_tile >= 1024 &&
            (size_t)4 * g.lds_per_wave * sizeof(float) <= kMaxDepositLds) {
            c->tsc_cell_geom = g;
            c->tsc_cell_ppl = kRowTilingTscCell[2];
            c->tsc_cell_pipe = 0;
        }
    }
    return ensure_partial(c);
}

// Create-time tiling of scheme k for a drift halo: long-row tiles
// (kRowTiling) on grids that hold at least 4096 of them, without a halo (a
// halo around a one- or two-row tile is mostly halo); otherwise the shape of
// default_tile_shape, with 4 particles per lane below 1024 particles per tile.
void default_tiling(const pic_hip_ctx* c, int k, int64_t capacity, int halo, int& tx, int& ty, int& ppl,
                    int& pipe) {
    const int Nx = c->dom.Nx, Ny = c->dom.Ny;
    const int rx = kRowTiling[k][0], ry = kRowTiling[k][1];
    const int64_t row_tiles = (int64_t)((Nx + rx - 1) / rx) * ((Ny + ry - 1) / ry);
    const int64_t row_per_tile = capacity * rx * ry / std::max<int64_t>(c->G, 1);
    pipe = 0;
    if (halo == 0 && rx <= Nx && row_tiles >= 4096 && row_per_tile >= 1024) {
        tx = rx;
        ty = ry;
        ppl = kRowTiling[k][2];
        pipe = kRowTiling[k][3];
        return;
    }
    tx = kDefaultTiling[k][0];
    ty = kDefaultTiling[k][1];
    // With a drift halo a wave's window is mostly halo and the tile kernel is
    // latency-bound at low particle counts: tiles down to 128 particles, for
    // more waves (256^2 at 8 per cell, halo 4: TSC step loop 89.8 -> 86.7 us).
    default_tile_shape(k, Nx, Ny, capacity, 4096, INT64_MAX, halo > 0 ? 128 : 256, tx, ty);
    const int64_t per_tile = capacity * tx * ty / std::max<int64_t>(c->G, 1);
    ppl = per_tile < 1024 ? 4 : kDefaultTiling[k][2];
}

pic_status ensure_fft(pic_hip_ctx* c) {
    if (c->fft_ready) return PIC_OK;
    const int hx = c->dom.Nx / 2 + 1;
    PIC_HIP_TRY(dmalloc(&c->phi, (size_t)c->G));
    PIC_HIP_TRY(dmalloc(&c->fft_in, (size_t)c->G));
    PIC_HIP_TRY(dmalloc(&c->spec, (size_t)c->dom.Ny * hx));
    PIC_HIP_TRY(dmalloc(&c->spec_ex, (size_t)2 * c->dom.Ny * hx));  // Ex then Ey spectrum
    c->spec_ey = c->spec_ex + (size_t)c->dom.Ny * hx;
    {
        // k^2 per axis (poisson_fft.cpp:23-35) and the forward-difference
        // symbols (e^{i k d} - 1)/d, tabulated in FP64.
        const int nx = c->dom.Nx, ny = c->dom.Ny;
        std::vector<double> k2x((size_t)hx), k2y((size_t)ny);
        std::vector<float2> sx((size_t)hx), sy((size_t)ny);
        for (int i = 0; i < hx; ++i) {
            const double kx = 2.0 * M_PI * ((i <= nx / 2) ? i : i - nx) / c->dom.Lx;
            k2x[(size_t)i] = kx * kx;
            const double th = 2.0 * M_PI * (double)i / nx;
            sx[(size_t)i] = make_float2((float)((std::cos(th) - 1.0) / c->dev.dx), (float)(std::sin(th) / c->dev.dx));
        }
        for (int j = 0; j < ny; ++j) {
            const double ky = 2.0 * M_PI * ((j <= ny / 2) ? j : j - ny) / c->dom.Ly;
            k2y[(size_t)j] = ky * ky;
            const double th = 2.0 * M_PI * (double)j / ny;
            sy[(size_t)j] = make_float2((float)((std::cos(th) - 1.0) / c->dev.dy), (float)(std::sin(th) / c->dev.dy));
        }
        PIC_HIP_TRY(dmalloc(&c->ptab.k2x, (size_t)hx));
        PIC_HIP_TRY(dmalloc(&c->ptab.k2y, (size_t)ny));
        PIC_HIP_TRY(dmalloc(&c->ptab.dx_sym, (size_t)hx));
        PIC_HIP_TRY(dmalloc(&c->ptab.dy_sym, (size_t)ny));
        PIC_HIP_TRY(hipMemcpy(c->ptab.k2x, k2x.data(), hx * sizeof(double), hipMemcpyHostToDevice));
        PIC_HIP_TRY(hipMemcpy(c->ptab.k2y, k2y.data(), ny * sizeof(double), hipMemcpyHostToDevice));
        PIC_HIP_TRY(hipMemcpy(c->ptab.dx_sym, sx.data(), hx * sizeof(float2), hipMemcpyHostToDevice));
        PIC_HIP_TRY(hipMemcpy(c->ptab.dy_sym, sy.data(), ny * sizeof(float2), hipMemcpyHostToDevice));
    }
    c->lds_fft = poisson_lds_supported(c->dom.Nx, c->dom.Ny);
    if (c->lds_fft) {
        auto table = [](int L) {
            std::vector<float2> t((size_t)L);
            for (int m = 0; m < L; ++m) {
                const double a = 2.0 * M_PI * (double)m / (double)L;
                t[(size_t)m] = make_float2((float)std::cos(a), (float)std::sin(a));
            }
            return t;
        };
        const std::vector<float2> tx = table(c->dom.Nx), ty = table(c->dom.Ny);
        PIC_HIP_TRY(dmalloc(&c->tw_x, tx.size()));
        PIC_HIP_TRY(dmalloc(&c->tw_y, ty.size()));
        PIC_HIP_TRY(hipMemcpy(c->tw_x, tx.data(), tx.size() * sizeof(float2), hipMemcpyHostToDevice));
        PIC_HIP_TRY(hipMemcpy(c->tw_y, ty.data(), ty.size() * sizeof(float2), hipMemcpyHostToDevice));
        c->fft_ready = true;
        return PIC_OK;
    }
    PIC_FFT_TRY(hipfftPlan2d(&c->plan_r2c, c->dom.Ny, c->dom.Nx, HIPFFT_R2C));
    PIC_FFT_TRY(hipfftPlan2d(&c->plan_c2r, c->dom.Ny, c->dom.Nx, HIPFFT_C2R));
    {
        // Ex and Ey in one batched inverse: spectra spec_ex, spec_ey -> grids Ex, Ey.
        int dims[2] = {c->dom.Ny, c->dom.Nx};
        int inembed[2] = {c->dom.Ny, hx};
        int onembed[2] = {c->dom.Ny, c->dom.Nx};
        PIC_FFT_TRY(hipfftPlanMany(&c->plan_c2r_e, 2, dims, inembed, 1, c->dom.Ny * hx, onembed, 1, (int)c->G,
                                   HIPFFT_C2R, 2));
    }
    PIC_FFT_TRY(hipfftSetStream(c->plan_r2c, c->stream));
    PIC_FFT_TRY(hipfftSetStream(c->plan_c2r, c->stream));
    PIC_FFT_TRY(hipfftSetStream(c->plan_c2r_e, c->stream));
    c->fft_ready = true;
    return PIC_OK;
}

pic_status check_ctx(pic_hip_ctx* c) {
    if (!c) return fail(PIC_ERR_INVALID_ARGUMENT, "null context");
    return PIC_OK;
}

}  // namespace

extern "C" {

const char* pic_hip_last_error(void) { return last_error_slot().c_str(); }
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
    PIC_HIP_TRY(dmalloc(&c->acc64, (size_t)c->G));
    PIC_HIP_TRY(hipMemsetAsync(c->rho, 0, (size_t)c->G * sizeof(float), c->stream));
    PIC_HIP_TRY(hipMemsetAsync(c->acc64, 0, (size_t)c->G * sizeof(double), c->stream));
    PIC_HIP_TRY(dmalloc(&c->status, 1));
    PIC_HIP_TRY(dmalloc(&c->acc_mark, 1));
    PIC_HIP_TRY(hipMemsetAsync(c->acc_mark, 0, sizeof(int), c->stream));
    PIC_HIP_TRY(hipMemsetAsync(c->status, 0, sizeof(int), c->stream));
    PIC_HIP_TRY(dmalloc(&c->Ex, (size_t)2 * c->G));  // Ex then Ey
    c->Ey = c->Ex + c->G;
    PIC_HIP_TRY(hipMemsetAsync(c->Ex, 0, (size_t)c->G * sizeof(float), c->stream));
    PIC_HIP_TRY(hipMemsetAsync(c->Ey, 0, (size_t)c->G * sizeof(float), c->stream));
    c->diag_blocks = 1024;
    PIC_HIP_TRY(dmalloc(&c->diag_partials, (size_t)5 * c->diag_blocks));
    PIC_HIP_TRY(dmalloc(&c->diag_out, 8));
    PIC_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&c->h_status), sizeof(int) * 4, hipHostMallocDefault));
    PIC_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&c->h_diag), sizeof(double) * 8, hipHostMallocDefault));
    // Defaults from the MI355X tiling sweeps (1024^2 / 64 ppc, 256^2 / 8-512 ppc):
    // at least 4096 tiles where the particles allow it; tiles of fewer than
    // two 512-particle chunks take 4 particles per lane (256-particle chunks).
    // Grids that hold at least 4096 of them take long-row tiles (kRowTiling):
    // a wave then streams one or two long cell ranges (1024^2 / 64 ppc: tile
    // kernel CIC 0.165 -> 0.159 ms, NGP 0.168 -> 0.153, TSC on the node order
    // 0.215 -> 0.203, Esirkepov 0.389 -> 0.375).
    c->tile_capacity = (int64_t)capacity;
    for (int k = 0; k < 4; ++k) {
        int tx, ty, ppl, pipe;
        default_tiling(c.get(), k, c->tile_capacity, 0, tx, ty, ppl, pipe);
        const pic_status st = apply_tiling(c.get(), k, tx, ty, ppl, pipe);
        if (st != PIC_OK) return st;
        c->tiling_default[k] = true;
    }
    {
        const pic_status st = sync_tsc_cell(c.get());
        if (st != PIC_OK) return st;
    }
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
        if (st & 2)
            return fail(PIC_ERR_UNSUPPORTED, "Esirkepov current: a particle moved one cell or more in one step");
        return fail(PIC_ERR_HIP, "device status flag set");
    }
    return PIC_OK;
}

pic_status pic_hip_set_tiling(pic_hip_ctx* ctx, int scheme, int tile_x, int tile_y, int particles_per_lane,
                              int pipelined) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    const pic_status st = apply_tiling(ctx, scheme, tile_x, tile_y, particles_per_lane, pipelined);
    if (st != PIC_OK) return st;
    for (int k = 0; k < 4; ++k)
        if (scheme < 0 || k == scheme) ctx->tiling_default[k] = false;
    return sync_tsc_cell(ctx);
}

pic_status pic_hip_set_tile_halo(pic_hip_ctx* ctx, int cells) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (cells < 0 || cells > kMaxDriftHalo)
        return fail(PIC_ERR_INVALID_ARGUMENT, "tile halo must be 0.." + std::to_string(kMaxDriftHalo) + " cells");
    // Every scheme's tiling as it stands, restored as a whole if any scheme's
    // tiling does not fit the new halo.
    const int old = ctx->halo;
    int saved[4][4];
    for (int k = 0; k < 4; ++k) {
        saved[k][0] = ctx->geom[k].T;
        saved[k][1] = ctx->geom[k].R;
        saved[k][2] = ctx->ppl[k];
        saved[k][3] = ctx->pipe[k];
    }
    ctx->halo = cells;
    for (int k = 0; k < 4; ++k) {
        // a scheme still on its create-time tiling takes the default for this halo
        int tx = saved[k][0], ty = saved[k][1], ppl = saved[k][2], pipe = saved[k][3];
        if (ctx->tiling_default[k]) default_tiling(ctx, k, ctx->tile_capacity, cells, tx, ty, ppl, pipe);
        const pic_status st = apply_tiling(ctx, k, tx, ty, ppl, pipe);
        if (st != PIC_OK) {
            const std::string msg = pic_hip_last_error();
            ctx->halo = old;
            for (int j = 0; j < 4; ++j) (void)apply_tiling(ctx, j, saved[j][0], saved[j][1], saved[j][2], saved[j][3]);
            (void)sync_tsc_cell(ctx);
            return fail(st, msg);
        }
    }
    return sync_tsc_cell(ctx);
}

int pic_hip_get_tile_halo(pic_hip_ctx* ctx) { return ctx ? ctx->halo : 0; }

pic_status pic_hip_get_tiling(pic_hip_ctx* ctx, int scheme, int* tile_x, int* tile_y, int* ppl, int* pipelined) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (scheme < 0 || scheme > 3) return fail(PIC_ERR_INVALID_ARGUMENT, "scheme must be 0..3");
    if (tile_x) *tile_x = ctx->geom[scheme].T;
    if (tile_y) *tile_y = ctx->geom[scheme].R;
    if (ppl) *ppl = ctx->ppl[scheme];
    if (pipelined) *pipelined = ctx->pipe[scheme];
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
    for (int i = 0; i < kStreams; ++i) {
        bool u = i >= 4;  // only charges and masses are candidates
        if (u && src[i])
            for (size_t k = 1; k < n && u; ++k) u = src[i][k] == src[i][0];
        ctx->uniform[i] = u;
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

pic_status pic_hip_upload_aos(pic_hip_ctx* ctx, const double* records, size_t n) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (n > 0 && !records) return fail(PIC_ERR_INVALID_ARGUMENT, "null records");
    if (n > 0 && (reinterpret_cast<uintptr_t>(records) & 15) != 0)
        return fail(PIC_ERR_INVALID_ARGUMENT, "records must be 16-byte aligned");
    if (n + 64 > ctx->cap) return fail(PIC_ERR_INVALID_ARGUMENT, "particle count exceeds context capacity");
    const size_t chunk = std::min<size_t>(n, kAosChunk);
    double* dev = nullptr;
    if (n > 0) PIC_HIP_TRY(dmalloc(&dev, chunk * kAosFields));
    std::unique_ptr<double, hipError_t (*)(void*)> guard(dev, hipFree);
    for (size_t b = 0; b < n; b += chunk) {
        const size_t m = std::min(chunk, n - b);
        PIC_HIP_TRY(hipMemcpyAsync(dev, records + b * kAosFields, m * kAosFields * sizeof(double),
                                   hipMemcpyHostToDevice, ctx->stream));
        float* const dst[6] = {ctx->s[0] + b, ctx->s[1] + b, ctx->s[2] + b, ctx->s[3] + b, ctx->s[4] + b,
                               ctx->s[5] + b};
        launch_aos_unpack(dev, (int64_t)m, dst, ctx->stream);
        PIC_HIP_TRY(hipGetLastError());
    }
    for (int i = 0; i < kStreams; ++i) {
        bool u = i >= 4;  // only charges and masses are candidates
        if (u) {
            const float v0 = n > 0 ? (float)records[i] : 0.0f;
            for (size_t k = 1; k < n && u; ++k) u = (float)records[k * kAosFields + i] == v0;
        }
        ctx->uniform[i] = u;
    }
    ctx->n = n;
    ctx->order_valid = false;
    PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

pic_status pic_hip_download_aos(pic_hip_ctx* ctx, double* records) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    const size_t n = ctx->n;
    if (n == 0) return PIC_OK;
    if (!records) return fail(PIC_ERR_INVALID_ARGUMENT, "null records");
    const size_t chunk = std::min<size_t>(n, kAosChunk);
    double* dev = nullptr;
    PIC_HIP_TRY(dmalloc(&dev, chunk * kAosFields));
    std::unique_ptr<double, hipError_t (*)(void*)> guard(dev, hipFree);
    for (size_t b = 0; b < n; b += chunk) {
        const size_t m = std::min(chunk, n - b);
        const float* const src[6] = {ctx->s[0] + b, ctx->s[1] + b, ctx->s[2] + b, ctx->s[3] + b, ctx->s[4] + b,
                                     ctx->s[5] + b};
        launch_aos_pack(src, (int64_t)m, dev, ctx->stream);
        PIC_HIP_TRY(hipGetLastError());
        PIC_HIP_TRY(hipMemcpyAsync(records + b * kAosFields, dev, m * kAosFields * sizeof(double),
                                   hipMemcpyDeviceToHost, ctx->stream));
    }
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
    for (int i = 0; i < kStreams; ++i) ctx->uniform[i] = i >= 4;  // one charge and one mass for all
    ctx->n = n;
    ctx->order_valid = false;
    return PIC_OK;
}

}  // extern "C"

namespace {

// Sort by cell and permute the six streams.  copy_back keeps the stream
// pointers fixed (needed inside captured graphs, which bake pointers in).
// Swaps the particle buffers of every stream a sort permutes.
void swap_permuted(pic_hip_ctx* ctx) {
    for (int i = 0; i < kStreams; ++i)
        if (!ctx->uniform[i]) std::swap(ctx->s[i], ctx->alt[i]);
}

// Buffers of the staged sort: the intermediate streams of the radix passes,
// the per-block counts and the mode word (outside stream capture).
pic_status ensure_sort_staging(pic_hip_ctx* ctx) {
    if (ctx->sort_rec) return PIC_OK;
    PIC_HIP_TRY(dmalloc(&ctx->far_part, sort_part_elems((int64_t)ctx->cap)));
    PIC_HIP_TRY(dmalloc(&ctx->permute_mode, 1));
    PIC_HIP_TRY(hipMemsetAsync(ctx->permute_mode, 0, sizeof(int), ctx->stream));
    PIC_HIP_TRY(dmalloc(&ctx->sort_rec, 2 * ctx->cap));
    return PIC_OK;
}

// Stencils whose tile kernel prefers the node-keyed order (a 9-slot TSC and
// a 1-slot NGP instead of 16 and 4 slots keyed by cell).
bool prefers_node_order(int scheme) { return scheme == NGP || scheme == TSC; }

// Sorts by cell key (particleCellIndex) or, with `node`, by nearest node.
pic_status sort_particles(pic_hip_ctx* ctx, bool copy_back, bool node) {
    StreamSet fwd{}, back{};
    for (int i = 0; i < kStreams; ++i) {
        if (ctx->uniform[i]) continue;  // every particle holds the same value
        fwd.src[fwd.count] = ctx->s[i];
        fwd.dst[fwd.count++] = ctx->alt[i];
        back.src[back.count] = ctx->alt[i];
        back.dst[back.count++] = ctx->s[i];
    }
    const int64_t n = (int64_t)ctx->n;
    ctx->last_sort_staged = ctx->sort_rec && n >= kStagedSortMin;
    if (ctx->sort_rec && n >= kStagedSortMin) {
        // Keys, the first digit's histogram and the already-sorted test (mode
        // 2: the passes only copy); then the radix passes move the streams
        // themselves (the staging buffer holds the intermediate set), so no
        // permutation is gathered.  Below kStagedSortMin particles the sort
        // keeps (key, index) passes and one gather, whose few launches cost
        // less than the staged path's.
        launch_keys_hist(ctx->s[0], ctx->s[1], n, ctx->dev, (int)ctx->G, ctx->sb, -1, node,
                         ctx->far_part, ctx->permute_mode, ctx->stream, false);
        PIC_HIP_TRY(launch_sort_records(n, (int)ctx->G, ctx->sb, fwd, reinterpret_cast<float*>(ctx->sort_rec),
                                        ctx->cap, ctx->stream, ctx->permute_mode, true, ctx->s[0], ctx->s[1],
                                        &ctx->dev, node));
    } else {
        PIC_HIP_TRY(launch_sort_gather(ctx->s[0], ctx->s[1], n, ctx->dev, (int)ctx->G, node, ctx->sb, fwd,
                                       ctx->stream));
    }
    if (copy_back)
        launch_permute_streams(back, nullptr, (int64_t)ctx->n, ctx->stream);
    else
        swap_permuted(ctx);
    PIC_HIP_TRY(hipGetLastError());
    ctx->order_valid = true;
    ctx->order_node = node;
    return PIC_OK;
}

}  // namespace

extern "C" {

pic_status pic_hip_sort(pic_hip_ctx* ctx) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    const pic_status st = ensure_sort_staging(ctx);
    if (st != PIC_OK) return st;
    return sort_particles(ctx, false, false);
}

pic_status pic_hip_sort_nodes(pic_hip_ctx* ctx) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    const pic_status st = ensure_sort_staging(ctx);
    if (st != PIC_OK) return st;
    return sort_particles(ctx, false, true);
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

pic_status pic_hip_sort_path(pic_hip_ctx* ctx, int32_t* path) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!path) return fail(PIC_ERR_INVALID_ARGUMENT, "null path");
    *path = -1;
    if (!ctx->last_sort_staged || !ctx->permute_mode) return PIC_OK;
    int m = -1;
    PIC_HIP_TRY(hipMemcpyAsync(&m, ctx->permute_mode, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    *path = m;
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
    // On the node-keyed order NGP and TSC run their node-centred stencils
    // (same tiling and windows as the cell-keyed kernels).
    a.scheme = backend == PIC_BACKEND_GPU_PRIV && ctx->order_node && prefers_node_order(scheme)
                   ? (scheme == NGP ? NGP_NODE : TSC_NODE)
                   : scheme;
    a.backend = backend;
    a.ppl = ctx->ppl[scheme];
    a.pipe = ctx->pipe[scheme];
    a.geom = ctx->geom[scheme];
    if (a.scheme == TSC && backend == PIC_BACKEND_GPU_PRIV) {  // TSC on the cell-keyed order
        a.ppl = ctx->tsc_cell_ppl;
        a.pipe = ctx->tsc_cell_pipe;
        a.geom = ctx->tsc_cell_geom;
    }
    a.geom.acc_mark = ctx->acc_mark;
    a.geom.epoch = ctx->deposit_epoch = ctx->deposit_epoch % 0x7ffffff0 + 1;  // never 0 (the initial marker)
    a.dt = (float)dt;
    a.rho = ctx->rho;
    a.partial = ctx->partial;
    a.acc = ctx->acc64;
    a.status = ctx->status;
    a.stream = ctx->stream;
    if (ctx->profiling && ctx->prof_used < ctx->prof_max) {
        a.ev_begin = ctx->prof_events[2 * ctx->prof_used];
        a.ev_end = ctx->prof_events[2 * ctx->prof_used + 1];
        ++ctx->prof_used;
    }
    if (ctx->defer_merge && backend == PIC_BACKEND_GPU_PRIV) {
        a.skip_merge = true;
        ctx->pending_merge = TileMerge{a.partial, a.acc, a.dom, a.geom};
        ctx->merge_pending = true;
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

}  // extern "C"

namespace {

// Spectral solve of ctx->rho on the context's stream (kernel launches and FFT
// executions only).  The mean removed before the transform comes from the
// device FP64 sum `sum` (see launch_poisson_prep).  Without `with_phi` the
// potential stays as its spectrum in ctx->spec until phi is read.
pic_status enqueue_solve(pic_hip_ctx* ctx, const double* sum, double sum_scale, double sum_add, bool with_phi,
                         const LearnedApply& la = LearnedApply{}) {
    const double cv = ctx->dev.dx * ctx->dev.dy;
    if (ctx->lds_fft) {
        launch_poisson_lds(ctx->rho, ctx->G, 0.0, cv, sum, sum_scale, sum_add, ctx->spec, ctx->spec_ex, ctx->spec_ey,
                           ctx->ptab, ctx->dev, ctx->tw_x, ctx->tw_y, ctx->stream, la);
        if (ctx->dom.Ny > 1) {
            launch_poisson_lds_rows(ctx->spec_ex, ctx->spec_ey, ctx->Ex, ctx->Ey, ctx->dev, ctx->tw_x, ctx->stream,
                                    ctx->rho, la);
        } else {  // one row: Ey is exactly zero (phi does not vary along y)
            launch_poisson_lds_rows(ctx->spec_ex, nullptr, ctx->Ex, nullptr, ctx->dev, ctx->tw_x, ctx->stream, ctx->rho,
                                    la);
            fill_async(ctx->Ey, (size_t)ctx->G, 0.0f, ctx->stream);
        }
        if (with_phi) launch_poisson_lds_rows(ctx->spec, nullptr, ctx->phi, nullptr, ctx->dev, ctx->tw_x, ctx->stream);
        ctx->phi_pending = !with_phi;
        PIC_HIP_TRY(hipGetLastError());
        return PIC_OK;
    }
    launch_poisson_prep(ctx->rho, ctx->fft_in, ctx->G, 0.0, cv, sum, sum_scale, sum_add, ctx->stream);
    PIC_FFT_TRY(hipfftExecR2C(ctx->plan_r2c, ctx->fft_in, reinterpret_cast<hipfftComplex*>(ctx->spec)));
    launch_poisson_spectral(ctx->spec, ctx->spec_ex, ctx->spec_ey, ctx->ptab, ctx->dev, ctx->stream);
    if (with_phi) PIC_FFT_TRY(hipfftExecC2R(ctx->plan_c2r, reinterpret_cast<hipfftComplex*>(ctx->spec), ctx->phi));
    PIC_FFT_TRY(hipfftExecC2R(ctx->plan_c2r_e, reinterpret_cast<hipfftComplex*>(ctx->spec_ex), ctx->Ex));
    ctx->phi_pending = !with_phi;
    PIC_HIP_TRY(hipGetLastError());
    return PIC_OK;
}

}  // namespace

extern "C" {

pic_status pic_hip_poisson(pic_hip_ctx* ctx) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    pic_status st = ensure_fft(ctx);
    if (st != PIC_OK) return st;
    launch_diagnostics(ctx->rho, nullptr, nullptr, ctx->G, nullptr, nullptr, nullptr, nullptr, 0, ctx->diag_partials,
                       ctx->diag_blocks, ctx->diag_out, ctx->stream);
    return enqueue_solve(ctx, ctx->diag_out, 1.0, 0.0, true);
}

pic_status pic_hip_get_fields(pic_hip_ctx* ctx, float* phi, float* Ex, float* Ey) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    const size_t bytes = (size_t)ctx->G * sizeof(float);
    if (phi) {
        if (!ctx->phi) return fail(PIC_ERR_INVALID_ARGUMENT, "phi not computed yet (call pic_hip_poisson)");
        if (ctx->phi_pending) {  // the step loop's last solve left the spectrum
            if (ctx->lds_fft)
                launch_poisson_lds_rows(ctx->spec, nullptr, ctx->phi, nullptr, ctx->dev, ctx->tw_x, ctx->stream);
            else
                PIC_FFT_TRY(hipfftExecC2R(ctx->plan_c2r, reinterpret_cast<hipfftComplex*>(ctx->spec), ctx->phi));
            ctx->phi_pending = false;
        }
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
                       ctx->Ey, ctx->dev, (float)dt, do_gather, do_push, ctx->stream, nullptr, ctx->uniform[4],
                       ctx->uniform[5]);
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

// ---- timestep driver ----------------------------------------------------------

}  // extern "C"

namespace {

pic_status ensure_step_buffers(pic_hip_ctx* ctx) {
    if (ctx->cosk) return PIC_OK;
    const int nx = ctx->dom.Nx;
    std::vector<double> c((size_t)nx), s((size_t)nx);
    const double k = 2.0 * M_PI / ctx->dom.Lx;
    for (int i = 0; i < nx; ++i) {
        const double x = (static_cast<double>(i) + 0.5) * ctx->dev.dx;
        c[(size_t)i] = std::cos(k * x);
        s[(size_t)i] = std::sin(k * x);
    }
    PIC_HIP_TRY(dmalloc(&ctx->cosk, (size_t)nx));
    PIC_HIP_TRY(dmalloc(&ctx->sink, (size_t)nx));
    PIC_HIP_TRY(dmalloc(&ctx->step_counter, 1));
    PIC_HIP_TRY(dmalloc(&ctx->push_flag, 1));
    PIC_HIP_TRY(dmalloc(&ctx->step_partials, (size_t)7 * ctx->diag_blocks));
    PIC_HIP_TRY(dmalloc(&ctx->run_sums, 5));
    PIC_HIP_TRY(hipMemcpyAsync(ctx->cosk, c.data(), nx * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    PIC_HIP_TRY(hipMemcpyAsync(ctx->sink, s.data(), nx * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

// Enqueues one step (see pic_hip_run); `diag` adds the step's diagnostics
// record.  The Poisson solve takes the mean of rho from the particles' charge
// sum in ctx->run_sums (set at the start of the run).
pic_status enqueue_step(pic_hip_ctx* ctx, const pic_step_config& cfg, bool with_sort, bool copy_back, bool diag,
                        int record_from, int cap, const StepRecords& rec, const StepHook& post_deposit) {
    pic_status st = PIC_OK;
    const float dt = (float)cfg.dt;
    auto push = [&]() {
        launch_gather_push(ctx->s[0], ctx->s[1], ctx->s[2], ctx->s[3], ctx->s[4], ctx->s[5], (int64_t)ctx->n, ctx->Ex,
                           ctx->Ey, ctx->dev, dt, 0, 1, ctx->stream);
    };
    auto deposit = [&]() -> pic_status {
        if (with_sort) {
            pic_status s2 = sort_particles(ctx, copy_back, prefers_node_order(cfg.scheme));
            if (s2 != PIC_OK) return s2;
        }
        pic_status s3 = pic_hip_deposit(ctx, cfg.scheme, cfg.backend, cfg.dt);
        if (s3 == PIC_OK && post_deposit) s3 = post_deposit();
        return s3;
    };
    auto solve = [&]() -> pic_status {
        // a learned density deferred by the post-deposit hook is written, and
        // the background added to it, by the solve's first pass
        LearnedApply la = ctx->pending_apply;
        ctx->pending_apply = LearnedApply{};
        if (la.net_out)
            la.bg = (float)cfg.background;
        else if (cfg.background != 0.0)
            launch_add_scalar(ctx->rho, ctx->G, (float)cfg.background, ctx->stream);
        const double cv = ctx->dev.dx * ctx->dev.dy;
        return enqueue_solve(ctx, ctx->run_sums + 3, 1.0 / cv, (double)(float)cfg.background * (double)ctx->G, false,
                             la);
    };
    if (cfg.order == PIC_ORDER_PUSH_FIRST) {
        // The push that opens this step ran fused with the previous step's
        // gather (or as the run's prologue); this step's gather is fused with
        // the next push, which ctx->push_flag disables after the last step.
        if ((st = deposit()) != PIC_OK) return st;
        if ((st = solve()) != PIC_OK) return st;
        launch_gather_push(ctx->s[0], ctx->s[1], ctx->s[2], ctx->s[3], ctx->s[4], ctx->s[5], (int64_t)ctx->n, ctx->Ex,
                           ctx->Ey, ctx->dev, dt, 1, 1, ctx->stream, ctx->push_flag, ctx->uniform[4], ctx->uniform[5]);
        const bool parts = cfg.record_particles != 0;
        if (diag)
            launch_step_diag(ctx->rho, ctx->Ex, ctx->Ey, ctx->cosk, ctx->sink, ctx->dev, ctx->s[2], ctx->s[3],
                             ctx->s[4], ctx->s[5], parts ? (int64_t)ctx->n : 0, ctx->step_partials, ctx->diag_blocks,
                             ctx->step_counter, record_from, cap, rec, ctx->stream);
        PIC_HIP_TRY(hipGetLastError());
        return PIC_OK;
    } else if (cfg.order == PIC_ORDER_DEPOSIT_SOLVE_PUSH) {
        if ((st = deposit()) != PIC_OK) return st;
        if ((st = solve()) != PIC_OK) return st;
        push();
    } else {
        if ((st = deposit()) != PIC_OK) return st;
        push();
        if ((st = solve()) != PIC_OK) return st;
    }
    launch_gather_push(ctx->s[0], ctx->s[1], ctx->s[2], ctx->s[3], ctx->s[4], ctx->s[5], (int64_t)ctx->n, ctx->Ex,
                       ctx->Ey, ctx->dev, dt, 1, 0, ctx->stream, nullptr, ctx->uniform[4], ctx->uniform[5]);
    const bool parts = cfg.record_particles != 0;
    if (diag)
        launch_step_diag(ctx->rho, ctx->Ex, ctx->Ey, ctx->cosk, ctx->sink, ctx->dev, ctx->s[2], ctx->s[3], ctx->s[4],
                         ctx->s[5], parts ? (int64_t)ctx->n : 0, ctx->step_partials, ctx->diag_blocks,
                         ctx->step_counter, record_from, cap, rec, ctx->stream);
    PIC_HIP_TRY(hipGetLastError());
    return PIC_OK;
}

struct GraphPair {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    ~GraphPair() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

pic_status capture_step(pic_hip_ctx* ctx, const pic_step_config& cfg, bool with_sort, bool diag, int record_from,
                        int cap, const StepRecords& rec, GraphPair& g, const StepHook& post_deposit) {
    PIC_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    pic_status st = enqueue_step(ctx, cfg, with_sort, false, diag, record_from, cap, rec, post_deposit);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (st != PIC_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    PIC_HIP_TRY(e);
    g.graph = graph;
    PIC_HIP_TRY(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
    return PIC_OK;
}

constexpr int kMaxIntervalGraph = 32;

// One sort interval in one graph: the sorting step and the `len` - 1 steps
// after it, so that a step loop of small sets launches one graph per
// interval instead of one per step.
pic_status capture_chunk(pic_hip_ctx* ctx, const pic_step_config& cfg, int len, bool diag, int record_from, int cap,
                         const StepRecords& rec, GraphPair& g, const StepHook& post_deposit) {
    PIC_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    pic_status st = PIC_OK;
    for (int k = 0; k < len && st == PIC_OK; ++k)
        st = enqueue_step(ctx, cfg, k == 0, false, diag, record_from, cap, rec, post_deposit);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (st != PIC_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    PIC_HIP_TRY(e);
    g.graph = graph;
    PIC_HIP_TRY(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
    return PIC_OK;
}

}  // namespace

extern "C" {

pic_status pic_hip_run(pic_hip_ctx* ctx, const pic_step_config* cfg, int32_t nsteps, int32_t record_from,
                       double* field_energy, double* mode1_amplitude, double* kinetic_energy, double* charge_error) {
    return run_steps(ctx, cfg, nsteps, record_from, field_energy, mode1_amplitude, kinetic_energy, charge_error,
                     StepHook());
}

}  // extern "C"

pic_status run_steps(pic_hip_ctx* ctx, const pic_step_config* cfg, int32_t nsteps, int32_t record_from,
                     double* field_energy, double* mode1_amplitude, double* kinetic_energy, double* charge_error,
                     const StepHook& post_deposit) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!cfg) return fail(PIC_ERR_INVALID_ARGUMENT, "null step config");
    if (nsteps < 0 || record_from < 0) return fail(PIC_ERR_INVALID_ARGUMENT, "negative step counts");
    if (cfg->order < 0 || cfg->order > 2) return fail(PIC_ERR_INVALID_ARGUMENT, "unknown step order");
    if (cfg->backend == PIC_BACKEND_GPU_PRIV && cfg->sort_interval < 1)
        return fail(PIC_ERR_INVALID_ARGUMENT, "GPU_Priv steps need sort_interval >= 1");
    if (cfg->scheme < 0 || cfg->scheme > 3) return fail(PIC_ERR_INVALID_ARGUMENT, "unknown scheme");
    if (cfg->backend != PIC_BACKEND_GPU_ATOMICS && cfg->backend != PIC_BACKEND_GPU_PRIV)
        return fail(PIC_ERR_INVALID_ARGUMENT, "unknown backend");
    if (!std::isfinite(cfg->dt)) return fail(PIC_ERR_INVALID_ARGUMENT, "dt must be finite");
    pic_status st = ensure_step_buffers(ctx);
    if (st != PIC_OK) return st;
    if ((st = ensure_fft(ctx)) != PIC_OK) return st;
    if ((st = ensure_sort_staging(ctx)) != PIC_OK) return st;
    ctx->pending_apply = LearnedApply{};  // nothing left over from a failed enqueue
    const int cap = nsteps > record_from ? nsteps - record_from : 0;
    double* dev_rec = nullptr;
    PIC_HIP_TRY(dmalloc(&dev_rec, (size_t)4 * std::max(cap, 1)));
    std::unique_ptr<double, hipError_t (*)(void*)> rec_guard(dev_rec, hipFree);
    StepRecords rec{field_energy ? dev_rec : nullptr, mode1_amplitude ? dev_rec + cap : nullptr,
                    kinetic_energy ? dev_rec + 2 * cap : nullptr, charge_error ? dev_rec + 3 * cap : nullptr};
    // Diagnostics run only on the steps that are recorded; the device step
    // counter starts at the first of them.
    const bool want_rec = cap > 0 && (field_energy || mode1_amplitude || kinetic_energy || charge_error);
    const int diag_from = want_rec ? record_from : nsteps;
    launch_set_int(ctx->step_counter, diag_from, ctx->stream);
    // The particles' total charge, from which every solve of the run takes the
    // mean of rho (the deposit conserves charge).
    launch_diagnostics(nullptr, nullptr, nullptr, 0, ctx->s[2], ctx->s[3], ctx->s[4], ctx->s[5], (int64_t)ctx->n,
                       ctx->diag_partials, ctx->diag_blocks, ctx->run_sums, ctx->stream);
    const bool priv = cfg->backend == PIC_BACKEND_GPU_PRIV;
    const bool fused = cfg->order == PIC_ORDER_PUSH_FIRST;
    if (fused && nsteps > 0) {
        launch_set_int(ctx->push_flag, 1, ctx->stream);
        launch_gather_push(ctx->s[0], ctx->s[1], ctx->s[2], ctx->s[3], ctx->s[4], ctx->s[5], (int64_t)ctx->n, ctx->Ex,
                           ctx->Ey, ctx->dev, (float)cfg->dt, 0, 1, ctx->stream);  // the first step's push
    }
    if (cfg->use_graph && nsteps > 0) {
        // A sort permutes the particles into the other buffer set, so graphs
        // exist for both: gs[v][p] sorts out of set p, gp[v][p] steps on set p,
        // v = 1 with the diagnostics record.  Capturing gs[v][0] then gs[v][1]
        // leaves the host pointers where they began.
        GraphPair gs[2][2], gp[2][2];
        const bool plain = !priv || cfg->sort_interval > 1;
        for (int v = 0; v < 2; ++v) {
            if (v == 0 ? diag_from == 0 : diag_from >= nsteps) continue;  // variant unused
            const bool d = v == 1;
            if (priv) {
                if ((st = capture_step(ctx, *cfg, true, d, record_from, cap, rec, gs[v][0], post_deposit)) != PIC_OK)
                    return st;
                if (plain &&
                    (st = capture_step(ctx, *cfg, false, d, record_from, cap, rec, gp[v][1], post_deposit)) != PIC_OK)
                    return st;
                if ((st = capture_step(ctx, *cfg, true, d, record_from, cap, rec, gs[v][1], post_deposit)) != PIC_OK)
                    return st;
            }
            if (plain &&
                (st = capture_step(ctx, *cfg, false, d, record_from, cap, rec, gp[v][0], post_deposit)) != PIC_OK)
                return st;
        }
        // Whole sort intervals as one graph each (gc[v][p], out of set p), for
        // the intervals that end before the last step (whose push the host
        // switches off) and lie on one side of the first recorded step.
        // Intervals longer than kMaxIntervalGraph steps replay step by step
        // (bounded graph size and capture time).
        GraphPair gc[2][2];
        const int si = cfg->sort_interval;
        const auto variant = [&](int step) { return step >= diag_from ? 1 : 0; };
        if (priv && si >= 2 && si <= kMaxIntervalGraph) {
            bool used[2] = {false, false};
            for (int step = 0; step + si < nsteps; step += si)
                if (variant(step) == variant(step + si - 1)) used[variant(step)] = true;
            for (int v = 0; v < 2; ++v) {
                if (!used[v]) continue;
                for (int p = 0; p < 2; ++p)
                    if ((st = capture_chunk(ctx, *cfg, si, v == 1, record_from, cap, rec, gc[v][p], post_deposit)) !=
                        PIC_OK)
                        return st;
            }
        }
        int parity = 0;
        for (int step = 0; step < nsteps; ++step) {
            if (fused && step == nsteps - 1) launch_set_int(ctx->push_flag, 0, ctx->stream);
            const int v = variant(step);
            if (priv && step % si == 0 && step + si < nsteps && gc[v][parity].exec && variant(step + si - 1) == v) {
                PIC_HIP_TRY(hipGraphLaunch(gc[v][parity].exec, ctx->stream));
                parity ^= 1;
                step += si - 1;
            } else if (priv && step % si == 0) {
                PIC_HIP_TRY(hipGraphLaunch(gs[v][parity].exec, ctx->stream));
                parity ^= 1;
            } else {
                PIC_HIP_TRY(hipGraphLaunch(gp[v][parity].exec, ctx->stream));
            }
        }
        PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (parity) swap_permuted(ctx);
    } else {
        for (int step = 0; step < nsteps; ++step) {
            if (fused && step == nsteps - 1) launch_set_int(ctx->push_flag, 0, ctx->stream);
            const bool sort_now = priv && step % cfg->sort_interval == 0;
            if ((st = enqueue_step(ctx, *cfg, sort_now, false, step >= diag_from, record_from, cap, rec,
                                   post_deposit)) != PIC_OK)
                return st;
        }
    }
    double* outs[4] = {field_energy, mode1_amplitude, kinetic_energy, charge_error};
    for (int k = 0; k < 4; ++k)
        if (outs[k] && cap)
            PIC_HIP_TRY(hipMemcpyAsync(outs[k], dev_rec + (size_t)k * cap, cap * sizeof(double),
                                       hipMemcpyDeviceToHost, ctx->stream));
    return pic_hip_sync(ctx);
}

extern "C" {

// Simulation::runProfiled (src/pic/simulation.cpp:147-292), SoA branch, on the
// device: per step [sort] -> push -> deposit -> background -> Poisson ->
// gather (Esirkepov: [sort] -> deposit -> background -> Poisson -> push ->
// gather), each phase bracketed by HIP events on the context's stream and
// executed eagerly (no graph, no fused gather + push), so the phase times are
// kernel times plus the launch gaps between them.  The reference deposits a
// copy of the particles taken before the push (simulation.cpp:191,203-208);
// here the deposit reads the pushed particles, as a timestep does.
pic_status pic_hip_run_profiled(pic_hip_ctx* ctx, const pic_step_config* cfg, int32_t warmup_steps,
                                int32_t timed_steps, pic_timestep_profile* out) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!cfg || !out) return fail(PIC_ERR_INVALID_ARGUMENT, "null step config or output");
    if (warmup_steps < 0 || timed_steps < 1) return fail(PIC_ERR_INVALID_ARGUMENT, "need timed_steps >= 1");
    if (cfg->scheme < 0 || cfg->scheme > 3) return fail(PIC_ERR_INVALID_ARGUMENT, "unknown scheme");
    if (cfg->backend != PIC_BACKEND_GPU_ATOMICS && cfg->backend != PIC_BACKEND_GPU_PRIV)
        return fail(PIC_ERR_INVALID_ARGUMENT, "unknown backend");
    if (cfg->backend == PIC_BACKEND_GPU_PRIV && cfg->sort_interval < 1)
        return fail(PIC_ERR_INVALID_ARGUMENT, "GPU_Priv steps need sort_interval >= 1");
    if (!std::isfinite(cfg->dt)) return fail(PIC_ERR_INVALID_ARGUMENT, "dt must be finite");
    pic_status st = ensure_fft(ctx);
    if (st != PIC_OK) return st;
    if ((st = ensure_sort_staging(ctx)) != PIC_OK) return st;
    // events e[0..7] per timed step mark the phase boundaries: sort | push
    // (not Esirkepov) | deposit | background (untimed, as in the reference) |
    // Poisson | push (Esirkepov) | gather
    constexpr int kEv = 8;
    std::vector<hipEvent_t> ev((size_t)kEv * timed_steps, nullptr);
    struct EventGuard {
        std::vector<hipEvent_t>& v;
        ~EventGuard() {
            for (hipEvent_t e : v)
                if (e) (void)hipEventDestroy(e);
        }
    } guard{ev};
    for (auto& e : ev) PIC_HIP_TRY(hipEventCreate(&e));
    const bool priv = cfg->backend == PIC_BACKEND_GPU_PRIV;
    const bool esk = cfg->scheme == ESK;
    // GPU_Atomics sorts only when asked to (sort_interval >= 1), like
    // SimulationConfig::sorted; GPU_Priv always does
    const int interval = std::max(1, cfg->sort_interval);
    const bool sorted = priv || cfg->sort_interval >= 1;
    const float dt = (float)cfg->dt;
    auto push = [&]() {
        launch_gather_push(ctx->s[0], ctx->s[1], ctx->s[2], ctx->s[3], ctx->s[4], ctx->s[5], (int64_t)ctx->n, ctx->Ex,
                           ctx->Ey, ctx->dev, dt, 0, 1, ctx->stream);
    };
    for (int step = 0; step < warmup_steps + timed_steps; ++step) {
        hipEvent_t* e = step >= warmup_steps ? &ev[(size_t)kEv * (step - warmup_steps)] : nullptr;
        auto mark = [&](int k) -> hipError_t { return e ? hipEventRecord(e[k], ctx->stream) : hipSuccess; };
        PIC_HIP_TRY(mark(0));
        if (sorted && step % interval == 0)
            if ((st = sort_particles(ctx, false, priv && prefers_node_order(cfg->scheme))) != PIC_OK) return st;
        PIC_HIP_TRY(mark(1));
        if (!esk) push();
        PIC_HIP_TRY(mark(2));
        if ((st = pic_hip_deposit(ctx, cfg->scheme, cfg->backend, cfg->dt)) != PIC_OK) return st;
        PIC_HIP_TRY(mark(3));
        if (cfg->background != 0.0) launch_add_scalar(ctx->rho, ctx->G, (float)cfg->background, ctx->stream);
        PIC_HIP_TRY(mark(4));
        if ((st = pic_hip_poisson(ctx)) != PIC_OK) return st;
        PIC_HIP_TRY(mark(5));
        if (esk) push();
        PIC_HIP_TRY(mark(6));
        launch_gather_push(ctx->s[0], ctx->s[1], ctx->s[2], ctx->s[3], ctx->s[4], ctx->s[5], (int64_t)ctx->n, ctx->Ex,
                           ctx->Ey, ctx->dev, dt, 1, 0, ctx->stream, nullptr, ctx->uniform[4], ctx->uniform[5]);
        PIC_HIP_TRY(mark(7));
        PIC_HIP_TRY(hipGetLastError());
    }
    if ((st = pic_hip_sync(ctx)) != PIC_OK) return st;
    // (phase, begin event, end event)
    const int spans[5][3] = {{0, 0, 1}, {1, esk ? 5 : 1, esk ? 6 : 2}, {2, 2, 3}, {3, 4, 5}, {4, 6, 7}};
    double acc[5] = {0, 0, 0, 0, 0};
    for (int t = 0; t < timed_steps; ++t) {
        const hipEvent_t* e = &ev[(size_t)kEv * t];
        for (const auto& sp : spans) {
            float ms = 0.0f;
            PIC_HIP_TRY(hipEventElapsedTime(&ms, e[sp[1]], e[sp[2]]));
            acc[sp[0]] += ms;
        }
    }
    const double n = (double)timed_steps;
    out->sort_ms = acc[0] / n;
    out->push_ms = acc[1] / n;
    out->deposit_ms = acc[2] / n;
    out->poisson_ms = acc[3] / n;
    out->gather_ms = acc[4] / n;
    out->total_ms = out->push_ms + out->sort_ms + out->deposit_ms + out->poisson_ms + out->gather_ms;
    out->sort_interval = interval;
    return PIC_OK;
}

double pic_hip_dominant_frequency(const double* signal, int64_t n, double dt) {
    if (!signal || n < 8) return 0.0;
    double mean = 0.0;
    for (int64_t i = 0; i < n; ++i) mean += signal[i];
    mean /= static_cast<double>(n);
    const int nn = static_cast<int>(n);
    double best_power = 0.0;
    int best_k = 1;
    for (int k = 1; k < nn / 2; ++k) {
        double re = 0.0, im = 0.0;
        for (int t = 0; t < nn; ++t) {
            const double phase = 2.0 * M_PI * static_cast<double>(k * t) / static_cast<double>(nn);
            const double c = signal[t] - mean;
            re += c * std::cos(phase);
            im -= c * std::sin(phase);
        }
        const double power = re * re + im * im;
        if (power > best_power) {
            best_power = power;
            best_k = k;
        }
    }
    return static_cast<double>(best_k) / (static_cast<double>(nn) * dt);
}

double pic_hip_growth_rate(const double* signal, int64_t n, double dt, int32_t record_start_step) {
    if (!signal || n < 8) return 0.0;
    const int64_t start = n / 4, end = (3 * n) / 4;
    if (end <= start + 2) return 0.0;
    double st = 0.0, sl = 0.0, stt = 0.0, stl = 0.0;
    int count = 0;
    for (int64_t i = start; i < end; ++i) {
        const double t = static_cast<double>(record_start_step + static_cast<int>(i)) * dt;
        const double l = std::log(std::max(signal[i], 1e-30));
        st += t;
        sl += l;
        stt += t * t;
        stl += t * l;
        ++count;
    }
    const double denom = static_cast<double>(count) * stt - st * st;
    if (std::abs(denom) < 1e-30) return 0.0;
    return (static_cast<double>(count) * stl - st * sl) / denom;
}

double pic_hip_plasma_frequency(const pic_domain* domain, int64_t num_particles) {
    if (!domain || num_particles <= 0) return 0.0;
    const double n0 = static_cast<double>(num_particles) / (domain->Lx * domain->Ly);
    const double q = 1.0 / static_cast<double>(num_particles);
    return std::sqrt(n0 * q * q);
}

double pic_hip_cold_two_stream_growth_rate(const pic_domain* domain, int64_t num_particles) {
    return pic_hip_plasma_frequency(domain, num_particles) / (2.0 * std::sqrt(2.0));
}

double pic_hip_resonant_beam_velocity(const pic_domain* domain, int64_t num_particles, int32_t mode) {
    if (!domain) return 0.0;
    const double k = 2.0 * M_PI * static_cast<double>(mode) / domain->Lx;
    return k < 1e-30 ? 0.0 : pic_hip_plasma_frequency(domain, num_particles) / k;
}

pic_status pic_hip_reference_particles(const pic_domain* domain, const pic_reference_init* init, size_t n, double* x,
                                       double* y, double* vx, double* vy, double* q, double* m) {
    if (!valid_domain(domain)) return fail(PIC_ERR_INVALID_ARGUMENT, "invalid domain");
    if (!init) return fail(PIC_ERR_INVALID_ARGUMENT, "null init");
    if (init->kind < PIC_REF_MAXWELLIAN || init->kind > PIC_REF_TWO_STREAM_1D)
        return fail(PIC_ERR_INVALID_ARGUMENT, "unknown reference initialiser");
    if (n > 0 && (!x || !y || !vx || !vy || !q || !m)) return fail(PIC_ERR_INVALID_ARGUMENT, "null output array");
    const double Lx = domain->Lx, Ly = domain->Ly;
    std::mt19937 gen(init->seed);
    std::uniform_real_distribution<double> along_x(0.0, Lx), along_y(0.0, Ly);
    // A new distribution object per draw: no second polar-method value is
    // carried from one draw to the next.
    auto gauss = [&gen](double sd) { return std::normal_distribution<double>(0.0, sd)(gen); };
    const double sd = std::sqrt(init->temperature);
    const double per_particle = n ? -1.0 / static_cast<double>(n) : 0.0;
    switch (init->kind) {
        case PIC_REF_MAXWELLIAN:
            for (size_t i = 0; i < n; ++i) {
                x[i] = along_x(gen);
                y[i] = along_y(gen);
                vx[i] = gauss(sd);
                vy[i] = gauss(sd);
                m[i] = 1.0;
                q[i] = i % 2 == 0 ? 1.0 : -1.0;
            }
            break;
        case PIC_REF_LANGMUIR:
        case PIC_REF_WARM_LANGMUIR: {
            const double k = 2.0 * M_PI * static_cast<double>(init->mode) / Lx;
            const bool warm = init->kind == PIC_REF_WARM_LANGMUIR;
            for (size_t i = 0; i < n; ++i) {
                x[i] = Lx * static_cast<double>(i) / static_cast<double>(n);
                y[i] = 0.5 * Ly;
                vx[i] = init->amplitude * std::sin(k * x[i]);
                vy[i] = 0.0;
                if (warm) {
                    vx[i] += gauss(sd);
                    vy[i] = gauss(sd);
                }
                m[i] = 1.0;
                q[i] = per_particle;
            }
            break;
        }
        default: {  // the two beams
            const double k = 2.0 * M_PI / Lx;
            const bool one_row = init->kind == PIC_REF_TWO_STREAM_1D;
            for (size_t i = 0; i < n; ++i) {
                x[i] = along_x(gen);
                y[i] = one_row ? 0.5 * Ly : along_y(gen);
                const double beam = i < n / 2 ? init->beam_velocity : -init->beam_velocity;
                vx[i] = beam + init->perturbation * std::sin(k * x[i]);
                vy[i] = 0.0;
                m[i] = 1.0;
                q[i] = per_particle;
            }
            break;
        }
    }
    return PIC_OK;
}

pic_status pic_hip_esirkepov_current(pic_hip_ctx* ctx, double dt) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!std::isfinite(dt) || dt <= 0.0) return fail(PIC_ERR_INVALID_ARGUMENT, "dt must be finite and positive");
    if (!ctx->cur.Jx) {
        float* base = nullptr;
        PIC_HIP_TRY(dmalloc(&base, (size_t)8 * ctx->G));
        PIC_HIP_TRY(hipMemsetAsync(base, 0, (size_t)8 * ctx->G * sizeof(float), ctx->stream));
        float** f[8] = {&ctx->cur.Jx, &ctx->cur.Jy, &ctx->cur.rho_old, &ctx->cur.rho_new,
                        &ctx->cur.Jx_c, &ctx->cur.Jy_c, &ctx->cur.rho_old_c, &ctx->cur.rho_new_c};
        for (int k = 0; k < 8; ++k) *f[k] = base + (size_t)k * ctx->G;
    }
    if (ctx->order_valid && !ctx->order_node && ctx->n > 0) {
        // Cell-sorted particles: tiles of 32 x 2 cells (smaller grids: the
        // grid), the size at which the kernel's LDS and registers leave room
        // for 8 waves per CU (64 x 2: 6 waves, 0.76 ms at 67M particles; 32 x
        // 2: 0.60 ms; 16 x 2: 0.64 ms).
        const TileGeom tg = make_tile_geom(ctx->dom.Nx, ctx->dom.Ny, 32, 2, 0);
        const size_t need = current_partial_doubles(tg);
        if (need > ctx->cur_partial_cap) {
            (void)hipFree(ctx->cur_partial);
            ctx->cur_partial = nullptr;
            ctx->cur_partial_cap = 0;
            PIC_HIP_TRY(dmalloc(&ctx->cur_partial, need));
            ctx->cur_partial_cap = need;
        }
        launch_esirkepov_current_tiled(ctx->s[0], ctx->s[1], ctx->s[2], ctx->s[3], ctx->s[4], ctx->sb.cell_off,
                                       ctx->dev, tg, (float)dt, ctx->cur_partial, ctx->cur, ctx->status, ctx->stream);
    } else {
        launch_esirkepov_current(ctx->s[0], ctx->s[1], ctx->s[2], ctx->s[3], ctx->s[4], (int64_t)ctx->n, ctx->dev,
                                 (float)dt, ctx->cur, ctx->status, ctx->stream);
    }
    PIC_HIP_TRY(hipGetLastError());
    return PIC_OK;
}

pic_status pic_hip_get_current(pic_hip_ctx* ctx, float* Jx, float* Jy, float* rho_old, float* rho_new) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!ctx->cur.Jx) return fail(PIC_ERR_INVALID_ARGUMENT, "no current computed yet (call pic_hip_esirkepov_current)");
    const size_t bytes = (size_t)ctx->G * sizeof(float);
    float* dst[4] = {Jx, Jy, rho_old, rho_new};
    float* src[4] = {ctx->cur.Jx, ctx->cur.Jy, ctx->cur.rho_old, ctx->cur.rho_new};
    for (int k = 0; k < 4; ++k)
        if (dst[k]) PIC_HIP_TRY(hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyDeviceToHost, ctx->stream));
    return pic_hip_sync(ctx);
}

pic_status pic_hip_continuity_residual(pic_hip_ctx* ctx, double dt, double* max_residual, double* max_drho_dt) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!ctx->cur.Jx) return fail(PIC_ERR_INVALID_ARGUMENT, "no current computed yet (call pic_hip_esirkepov_current)");
    const int nb = 512;
    double* part = nullptr;
    PIC_HIP_TRY(dmalloc(&part, (size_t)2 * nb + 2));
    std::unique_ptr<double, hipError_t (*)(void*)> guard(part, hipFree);
    launch_continuity_residual(ctx->cur, ctx->dev, dt, part, nb, part + 2 * nb, ctx->stream);
    PIC_HIP_TRY(hipMemcpyAsync(ctx->h_diag, part + 2 * nb, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    const pic_status st = pic_hip_sync(ctx);
    if (st != PIC_OK) return st;
    if (max_residual) *max_residual = ctx->h_diag[0];
    if (max_drho_dt) *max_drho_dt = ctx->h_diag[1];
    return PIC_OK;
}

pic_status pic_hip_autotune(pic_hip_ctx* ctx, int scheme, const char* cache_path, int reps, int* tile_x,
                            int* tile_y, int* particles_per_lane, double* ms_per_deposit) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (scheme < 0 || scheme > 3) return fail(PIC_ERR_INVALID_ARGUMENT, "scheme must be 0..3");
    if (!ctx->order_valid) return fail(PIC_ERR_NOT_SORTED, "autotune times the tiled deposit: call pic_hip_sort first");
    if (reps < 1) reps = 5;
    const std::string key = tuning_key(scheme, ctx->dom.Nx, ctx->dom.Ny, ppc_bucket((int64_t)ctx->n, ctx->G), ctx->halo);
    TilingChoice best{ctx->geom[scheme].T, ctx->geom[scheme].R, ctx->ppl[scheme]};
    double best_ms = 0.0;
    const auto cached = tuning_cache_read(cache_path);
    auto hit = cached.find(key);
    // the tilings tried apply to both orders of TSC
    ctx->tiling_default[scheme] = false;
    auto apply = [&](int tx, int ty, int ppl) {
        const pic_status st = apply_tiling(ctx, scheme, tx, ty, ppl, 0);
        return st != PIC_OK ? st : sync_tsc_cell(ctx);
    };
    if (hit != cached.end() && apply(hit->second.tx, hit->second.ty, hit->second.ppl) == PIC_OK) {
        best = hit->second;
    } else {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        PIC_HIP_TRY(hipEventCreate(&e0));
        std::unique_ptr<ihipEvent_t, hipError_t (*)(hipEvent_t)> g0(e0, hipEventDestroy);
        PIC_HIP_TRY(hipEventCreate(&e1));
        std::unique_ptr<ihipEvent_t, hipError_t (*)(hipEvent_t)> g1(e1, hipEventDestroy);
        best_ms = -1.0;
        for (const TilingChoice& c : tiling_candidates(ctx->dom.Nx, ctx->dom.Ny, false, ctx->halo)) {
            if (apply(c.tx, c.ty, c.ppl) != PIC_OK) continue;
            pic_status st = pic_hip_deposit(ctx, scheme, PIC_BACKEND_GPU_PRIV, 0.0);  // warm-up
            if (st != PIC_OK) return st;
            std::vector<double> t;
            for (int r = 0; r < reps; ++r) {
                PIC_HIP_TRY(hipEventRecord(e0, ctx->stream));
                if ((st = pic_hip_deposit(ctx, scheme, PIC_BACKEND_GPU_PRIV, 0.0)) != PIC_OK) return st;
                PIC_HIP_TRY(hipEventRecord(e1, ctx->stream));
                PIC_HIP_TRY(hipEventSynchronize(e1));
                float ms = 0.0f;
                PIC_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
                t.push_back(ms);
            }
            const double m = median_of(t);
            if (best_ms < 0.0 || m < best_ms) {
                best_ms = m;
                best = c;
            }
        }
        if (best_ms < 0.0) return fail(PIC_ERR_INVALID_ARGUMENT, "no tiling candidate fits this grid");
        if (!tuning_cache_store(cache_path, key, best))
            return fail(PIC_ERR_INVALID_ARGUMENT, std::string("cannot write tuning cache ") + cache_path);
    }
    pic_status st = apply(best.tx, best.ty, best.ppl);
    if (st != PIC_OK) return st;
    if (tile_x) *tile_x = best.tx;
    if (tile_y) *tile_y = best.ty;
    if (particles_per_lane) *particles_per_lane = best.ppl;
    if (ms_per_deposit) *ms_per_deposit = best_ms;
    return pic_hip_sync(ctx);
}

pic_status pic_hip_copy_probe(pic_hip_ctx* ctx, size_t bytes) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    const size_t n = bytes / sizeof(float);
    float *a = nullptr, *b = nullptr;
    PIC_HIP_TRY(dmalloc(&a, n));
    std::unique_ptr<float, hipError_t (*)(void*)> ga(a, hipFree);
    PIC_HIP_TRY(dmalloc(&b, n));
    std::unique_ptr<float, hipError_t (*)(void*)> gb(b, hipFree);
    fill_async(a, n, 1.0f, ctx->stream);
    launch_copy_f32(a, b, (int64_t)n, ctx->stream);
    PIC_HIP_TRY(hipGetLastError());
    PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

pic_status pic_hip_read_probe(pic_hip_ctx* ctx, size_t bytes, int reps, double* gbs) {
    if (check_ctx(ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!gbs || reps < 1 || bytes < 16) return fail(PIC_ERR_INVALID_ARGUMENT, "need gbs, reps >= 1, bytes >= 16");
    const int64_t n4 = (int64_t)(bytes / 16);
    int dev = 0, ncu = 1;
    PIC_HIP_TRY(hipGetDevice(&dev));
    PIC_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int blocks = std::max(1, ncu) * 8;
    float *a = nullptr, *part = nullptr;
    PIC_HIP_TRY(dmalloc(&a, (size_t)n4 * 4));
    std::unique_ptr<float, hipError_t (*)(void*)> ga(a, hipFree);
    PIC_HIP_TRY(dmalloc(&part, (size_t)blocks));
    std::unique_ptr<float, hipError_t (*)(void*)> gp(part, hipFree);
    fill_async(a, (size_t)n4 * 4, 1.0f, ctx->stream);
    launch_read_probe(a, n4, part, blocks, ctx->stream);  // warm-up
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PIC_HIP_TRY(hipEventCreate(&e0));
    PIC_HIP_TRY(hipEventCreate(&e1));
    std::vector<double> t;
    hipError_t err = hipSuccess;
    for (int r = 0; r < reps && err == hipSuccess; ++r) {
        (void)hipEventRecord(e0, ctx->stream);
        launch_read_probe(a, n4, part, blocks, ctx->stream);
        (void)hipEventRecord(e1, ctx->stream);
        err = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        t.push_back(ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    PIC_HIP_TRY(err);
    *gbs = (double)n4 * 16.0 / (median_of(t) * 1e-3) / 1e9;
    return PIC_OK;
}

// ---- drop-in seam -----------------------------------------------------------

// The seam of both layouts: field a of particle k is src[a][k * stride] (SoA:
// stride 1; AoS records: field pointers into the first record, stride 6).
static pic_status seam_deposit(const double* const src_in[5], size_t stride, size_t n, const pic_domain* domain,
                        int scheme, int backend, int sorted, double dt, double* rho_out) {
    if (!valid_domain(domain)) return fail(PIC_ERR_INVALID_ARGUMENT, "invalid domain");
    if (!rho_out) return fail(PIC_ERR_INVALID_ARGUMENT, "null rho_out");
    // A per-thread context and pinned staging buffer, grown on demand (the
    // reference keeps a static grow-only device pool, deposition_cuda.cu:243-273).
    struct Staging {
        float* p = nullptr;
        size_t cap = 0;
        ~Staging() {
            if (p) (void)hipHostFree(p);
        }
    };
    thread_local std::unique_ptr<pic_hip_ctx, pic_status (*)(pic_hip_ctx*)> cache(nullptr, pic_hip_destroy);
    thread_local Staging stage;
    if (!cache || cache->dom.Nx != domain->Nx || cache->dom.Ny != domain->Ny || cache->dom.Lx != domain->Lx ||
        cache->dom.Ly != domain->Ly || cache->cap < n + 64) {
        cache.reset();
        pic_hip_ctx* c = nullptr;
        pic_status st = pic_hip_create(domain, n, &c);
        if (st != PIC_OK) return st;
        cache.reset(c);
    }
    pic_hip_ctx* c = cache.get();
    const double* src[5] = {src_in[0], src_in[1], src_in[2], src_in[3], src_in[4]};
    if (stage.cap < 5 * n) {
        if (stage.p) (void)hipHostFree(stage.p);
        stage.p = nullptr;
        stage.cap = 0;
        PIC_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&stage.p), std::max<size_t>(5 * n, 1) * sizeof(float),
                                  hipHostMallocDefault));
        stage.cap = 5 * n;
    }
    // FP64 -> FP32 into pinned memory, split over host threads for large sets,
    // then DMA copies on the context's stream.  (The workers see the staging
    // buffer through `sp`: `stage` names a different object in each thread.)
    float* const sp = stage.p;
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const unsigned nt = n >= ((size_t)1 << 18) ? hw : 1;
    std::vector<char> q_same(nt, 1);  // per worker: its charges all equal the first particle's
    auto convert = [sp, &src, n, stride, &q_same](unsigned t, size_t b, size_t e) {
        for (int a = 0; a < 5; ++a) {
            if (!src[a]) continue;
            float* d = sp + (size_t)a * n;
            const double* f = src[a];
            for (size_t k = b; k < e; ++k) d[k] = (float)f[k * stride];
        }
        const float q0 = (float)src[4][0];
        bool same = true;
        for (size_t k = b; k < e && same; ++k) same = sp[4 * n + k] == q0;
        q_same[t] = same ? 1 : 0;
    };
    if (n > 0) {
        if (nt > 1) {
            std::vector<std::thread> pool;
            const size_t chunk = (n + nt - 1) / nt;
            for (unsigned t = 1; t < nt; ++t) {
                const size_t b = std::min(n, t * chunk), e = std::min(n, b + chunk);
                pool.emplace_back(convert, t, b, e);
            }
            convert(0, 0, std::min(n, chunk));
            for (auto& th : pool) th.join();
        } else {
            convert(0, 0, n);
        }
    }
    for (int a = 0; a < 5 && n > 0; ++a) {
        float* d = c->s[a];
        if (src[a])
            PIC_HIP_TRY(hipMemcpyAsync(d, sp + (size_t)a * n, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        else
            fill_async(d, n, 0.0f, c->stream);
    }
    if (n > 0) fill_async(c->s[5], n, 1.0f, c->stream);  // mass
    for (int i = 0; i < kStreams; ++i) c->uniform[i] = false;
    c->uniform[5] = true;
    c->uniform[4] = n > 0 && std::all_of(q_same.begin(), q_same.end(), [](char v) { return v != 0; });
    c->n = n;
    c->order_valid = false;
    pic_status st = PIC_OK;
    if (sorted || backend == PIC_BACKEND_GPU_PRIV) {
        // The caller's order: usually far from sorted.  The tiled NGP and TSC
        // kernels take the node-keyed order.
        if ((st = ensure_sort_staging(c)) != PIC_OK) return st;
        st = sort_particles(c, false, backend == PIC_BACKEND_GPU_PRIV && prefers_node_order(scheme));
        if (st != PIC_OK) return st;
    }
    st = pic_hip_deposit(c, scheme, backend, dt);
    if (st != PIC_OK) return st;
    return pic_hip_get_rho_f64(c, rho_out);
}

pic_status pic_hip_deposit_host(const double* x, const double* y, const double* vx, const double* vy,
                                const double* q, size_t n, const pic_domain* domain, int scheme, int backend,
                                int sorted, double dt, double* rho_out) {
    if (n > 0 && (!x || !y || !q)) return fail(PIC_ERR_INVALID_ARGUMENT, "x, y and q are required");
    if (scheme == PIC_SCHEME_ESIRKEPOV && n > 0 && (!vx || !vy))
        return fail(PIC_ERR_INVALID_ARGUMENT, "the Esirkepov scheme needs vx and vy");
    const double* src[5] = {x, y, vx, vy, q};
    return seam_deposit(src, 1, n, domain, scheme, backend, sorted, dt, rho_out);
}

pic_status pic_hip_deposit_host_aos(const double* records, size_t n, const pic_domain* domain, int scheme,
                                    int backend, int sorted, double dt, double* rho_out) {
    if (n > 0 && !records) return fail(PIC_ERR_INVALID_ARGUMENT, "null records");
    const bool esk = scheme == PIC_SCHEME_ESIRKEPOV;
    const double* src[5] = {records, records + 1, esk ? records + 2 : nullptr, esk ? records + 3 : nullptr,
                            records + 4};
    return seam_deposit(src, kAosFields, n, domain, scheme, backend, sorted, dt, rho_out);
}

}  // extern "C"
