// Internal layout of pic_hip_ctx (the opaque handle of include/pic_hip.h),
// shared by the C ABI translation units.
//
// The context owns every device buffer of one PIC instance: the FP32 SoA
// particle streams (and their sort double-buffers), the cell-sort scratch, the
// FP32 rho / phi / Ex / Ey grids, the tile partials of the LDS deposit and the
// hipFFT plans.  All work is enqueued on the context's stream.
#pragma once

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <functional>
#include <vector>

#include "../../include/pic_hip.h"
#include "api_util.hpp"
#include "pic_common.hpp"
#include "pic_kernels.hpp"

using pichip::CurrentGrids;
using pichip::DevDomain;
using pichip::kStreams;
using pichip::PoissonTables;
using pichip::SortBuffers;
using pichip::TileGeom;

struct pic_hip_ctx {
    pic_domain dom{};
    DevDomain dev{};
    int64_t G = 0;
    size_t cap = 0;  // particle capacity (rounded, plus float4 over-read pad)
    size_t n = 0;
    hipStream_t stream = nullptr;
    float* s[kStreams] = {};
    float* alt[kStreams] = {};
    // Stream k holds one value for every particle (charges and masses of a
    // single species): a permutation leaves it unchanged, so sorts skip it.
    bool uniform[kStreams] = {};
    // Staged sort (at least kStagedSortMin particles): the radix passes carry
    // the streams, the intermediate set in sort_rec (8 floats per particle of
    // capacity); far_part holds per-block counts from which the device sets
    // permute_mode (2: the order is still sorted, the passes only copy);
    // allocated before the first sort.
    float4* sort_rec = nullptr;
    unsigned* far_part = nullptr;  // sort_part_elems(cap): per-block far / descent counts
    int* permute_mode = nullptr;
    SortBuffers sb{};
    bool last_sort_staged = false;  // the last sort set permute_mode (pic_hip_sort_path)
    bool order_valid = false;
    bool order_node = false;  // the valid order is the node-keyed one (pic_hip_sort_nodes)
    float* rho = nullptr;
    double* acc64 = nullptr;  // FP64 global-atomic accumulation grid, zero between deposits
    int* status = nullptr;
    int* acc_mark = nullptr;  // overflow marker of the tiled deposit (TileGeom::acc_mark)
    int deposit_epoch = 0;
    float* partial = nullptr;
    size_t partial_cap = 0;
    // LDS tiling per scheme (NGP, CIC, TSC, Esirkepov); the partial buffer is
    // sized for the largest.
    TileGeom geom[4]{};
    int ppl[4] = {8, 8, 8, 4};
    int halo = 0;  // drift halo of every scheme's tile windows (cells)
    // Scheme still on its create-time tiling (not set by pic_hip_set_tiling or
    // pic_hip_autotune): pic_hip_set_tile_halo re-picks it for the new halo.
    bool tiling_default[4] = {};
    int64_t tile_capacity = 0;  // capacity the create-time tilings were chosen for
    int pipe[4] = {1, 1, 0, 0};
    // Tiling of TSC on the cell-keyed order: the TSC tiling, except while TSC
    // is on its create-time tiling without a drift halo, when the cell-keyed
    // kernel takes its own long rows (the create-time long rows are chosen for
    // the node-keyed order's moment form).
    TileGeom tsc_cell_geom{};
    int tsc_cell_ppl = 8, tsc_cell_pipe = 0;
    float* phi = nullptr;
    float* Ex = nullptr;
    float* Ey = nullptr;
    float* fft_in = nullptr;
    float2* spec = nullptr;     // phi spectrum [Ny][Nx/2+1]
    float2* spec_ex = nullptr;  // Ex, Ey spectra (one allocation, Ey after Ex)
    float2* spec_ey = nullptr;
    PoissonTables ptab{};
    hipfftHandle plan_r2c = 0, plan_c2r = 0, plan_c2r_e = 0;  // plan_c2r_e: Ex and Ey batched
    // Power-of-two grids up to 1024 per axis solve with the LDS transforms
    // (twiddle tables tw_x, tw_y); the hipFFT plans serve the other grids.
    bool lds_fft = false;
    float2* tw_x = nullptr;
    float2* tw_y = nullptr;
    // The step loop leaves the potential's spectrum in `spec` and transforms it
    // only when phi is read.
    bool phi_pending = false;
    double* run_sums = nullptr;  // per run: FP64 sums of the particle charges (diagnostics layout)
    bool fft_ready = false;
    double* diag_partials = nullptr;
    double* diag_out = nullptr;
    int diag_blocks = 0;
    int* h_status = nullptr;     // pinned
    double* h_diag = nullptr;    // pinned
    CurrentGrids cur{};           // Esirkepov current buffers (allocated on first use)
    double* cur_partial = nullptr;  // tile windows of the tiled current (4 planes)
    size_t cur_partial_cap = 0;
    double* cosk = nullptr;      // cos/sin(2 pi (i + 1/2) dx / Lx), mode-1 diagnostic tables
    double* sink = nullptr;
    int* step_counter = nullptr;
    int* push_flag = nullptr;    // device flag: the fused gather + push of a step pushes
    double* step_partials = nullptr;
    // A learned density's gate / renormalisation enqueued by a step's
    // post-deposit hook and not yet launched: the step's Poisson solve takes
    // it into its first pass (net_out null: none pending).
    pichip::LearnedApply pending_apply{};
    // Learned step loop: the tiled deposit leaves its window merge to the
    // denoiser's first layer (defer_merge), which takes pending_merge.
    bool defer_merge = false;
    bool merge_pending = false;
    pichip::TileMerge pending_merge{};
    std::vector<hipEvent_t> prof_events;  // pairs (begin, end)
    int prof_max = 0, prof_used = 0;
    bool profiling = false;

    ~pic_hip_ctx() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipEvent_t e : prof_events) (void)hipEventDestroy(e);
        if (plan_r2c) (void)hipfftDestroy(plan_r2c);
        if (plan_c2r) (void)hipfftDestroy(plan_c2r);
        if (plan_c2r_e) (void)hipfftDestroy(plan_c2r_e);
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
        (void)hipFree(acc64);
        (void)hipFree(status);
        (void)hipFree(partial);
        (void)hipFree(phi);
        (void)hipFree(Ex);  // Ey shares the allocation
        (void)hipFree(fft_in);
        (void)hipFree(spec);
        (void)hipFree(spec_ex);  // spec_ey shares the allocation
        (void)hipFree(ptab.k2x);
        (void)hipFree(ptab.k2y);
        (void)hipFree(ptab.dx_sym);
        (void)hipFree(ptab.dy_sym);
        (void)hipFree(diag_partials);
        (void)hipFree(diag_out);
        (void)hipFree(cur.Jx);  // one allocation holds all eight grids
        (void)hipFree(cur_partial);
        (void)hipFree(cosk);
        (void)hipFree(sink);
        (void)hipFree(step_counter);
        (void)hipFree(push_flag);
        (void)hipFree(step_partials);
        (void)hipFree(run_sums);
        (void)hipFree(tw_x);
        (void)hipFree(tw_y);
        (void)hipFree(sort_rec);
        (void)hipFree(far_part);
        (void)hipFree(permute_mode);
        (void)hipFree(acc_mark);
        (void)hipHostFree(h_status);
        (void)hipHostFree(h_diag);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Work enqueued on the context's stream right after each step's deposit
// (kernel launches only, so that it can be captured in a graph).
using StepHook = std::function<pic_status()>;

// pic_hip_run with an optional hook after the deposit.
pic_status run_steps(pic_hip_ctx* ctx, const pic_step_config* cfg, int32_t nsteps, int32_t record_from,
                     double* field_energy, double* mode1_amplitude, double* kinetic_energy, double* charge_error,
                     const StepHook& post_deposit);
