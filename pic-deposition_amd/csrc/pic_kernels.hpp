// Host-side launch interfaces of the HIP kernels (one translation unit each:
// deposit.hip, sort.hip, step.hip, current.hip, ensemble.hip).  Used by the
// C ABI files (api.hip, ensemble_api.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "pic_common.hpp"

namespace pichip {

// Largest drift halo (cells) of the tile windows.
constexpr int kMaxDriftHalo = 4;

// Tiling of the grid for the LDS-privatised deposit: tiles of T cells along x
// by R cells along y.  A tile's window takes the particles sorted into it whose
// current cell lies within D cells of the tile (the drift halo: particles move
// between sorts), so it spans nodes (x0-1-D .. x0+T+1+D) x (y0-1-D .. y0+R+1+D),
// i.e. WX = T+3+2D by WY = R+3+2D; particles further out take the overflow path.
struct TileGeom {
    int T, R;
    int D;  // drift halo, 0 .. kMaxDriftHalo
    int ntx, nty, ntiles;
    int WX, WY, wsize;
    int lds_per_wave;  // floats: 2 windows + row tables
    int fast;          // merge fast path: no halo, tiles >= 3 cells on each axis, >= 3 tiles per axis
    int fast_rows;     // merge row path: as fast along x; tiles of 1 or 2 rows dividing the y axis
    int uniform;       // merge range path: equal tiles dividing each axis, one window image per node
    int mbatch = 16;   // range path: windows per node loaded together (4, 8, 16 or 32: the most per node, up to 32)
    // Overflow marker of the single-context deposit: a tile that sends a
    // particle to the FP64 accumulation grid stores `epoch` at *acc_mark, and
    // the merge reads that grid only when *acc_mark == epoch.  A marker left
    // by an earlier deposit with the same epoch (graph replays) costs a read of
    // a zero grid, never a missed contribution.  nullptr: the grid is always read.
    int* acc_mark = nullptr;
    int epoch = 0;
};

TileGeom make_tile_geom(int Nx, int Ny, int T, int R, int D = 0);

// LDS floats per wave of the TSC moment stencil: the nine FP64 moments of
// every home node of the tile, the row tables, the ring of particles left to
// the per-particle path and the FP64 node window the moments map to.
int moment_lds_per_wave(const TileGeom& g, int ppl);
// TSC on the node-keyed order runs as charge moments: no drift halo and the
// moment windows of four waves fit 64 KiB of LDS.
bool tsc_uses_moments(const TileGeom& g, int ppl);

struct DepositArgs {
    const float *x, *y, *vx, *vy, *q;
    int64_t n;
    const int* cell_off;  // G+1 offsets of the cell-sorted order (tile backend only)
    DevDomain dom;
    int scheme;   // Scheme
    int backend;  // 1 = global atomics, 2 = LDS tiles
    int ppl;      // particles per lane per chunk (4 or 8)
    int pipe;     // 1: two register sets, next chunk's loads overlap processing
    TileGeom geom;
    float dt;
    float* rho;       // [Ny][Nx] FP32 out
    float* partial;   // ntiles * wsize
    double* acc;      // [Ny][Nx] FP64 global-atomic accumulation grid, kept zero between calls
    int* status;      // bit 0: Esirkepov move longer than the general walk supports
    hipStream_t stream;
    hipEvent_t ev_begin = nullptr;  // if set, recorded around the main scatter kernel
    hipEvent_t ev_end = nullptr;
    bool skip_merge = false;  // tile backend: leave the window merge to the caller (TileMerge)
};

// What a deferred window merge needs: the tile windows, the FP64 accumulation
// grid, the deposit's domain and tiling (with its overflow marker and epoch).
struct TileMerge {
    const float* partial;
    double* acc;
    DevDomain dom;
    TileGeom geom;
};

hipError_t launch_deposit(const DepositArgs& a);
// Global-atomic deposit of a Bernoulli(1/stride) subset of the particles
// (selected by a hash of index and salt) with charges scaled by `scale`.
hipError_t launch_deposit_subsample(const DepositArgs& a, uint32_t stride, uint64_t salt, float scale);
void fill_async(float* p, size_t n, float v, hipStream_t s);

// ---- sort.hip -------------------------------------------------------------
struct SortBuffers {
    int* keys;       // [n] cell key of each particle, current order
    int* keys_alt;   // [n]
    int* perm;       // [n]
    int* perm_alt;   // [n]
    uint32_t* hist;  // radix histogram scratch (digit-major, blocks minor)
    size_t hist_cap;
    uint32_t* scan_tmp;  // scan scratch
    size_t scan_tmp_cap;
    int* cell_off;   // [G+1]
};

// Bytes of histogram / scan scratch the sort needs for n particles.
size_t sort_hist_elems(int64_t n);
size_t sort_scan_tmp_elems(size_t m);

// Computes keys for the current order and a stable permutation sorting them;
// on return perm[i] is the source index of the i-th particle in sorted order,
// keys_sorted (== sb.keys) the sorted keys, and cell_off the G+1 offsets.
hipError_t launch_sort_keys_perm(const float* x, const float* y, int64_t n, const DevDomain& d, SortBuffers& sb,
                                 hipStream_t s);
// Stable sort of precomputed keys in sb.keys (each in [0, key_range)):
// same outputs as launch_sort_keys_perm, cell_off with key_range + 1 entries.
// With `mode` (set by launch_keys_hist), mode 2 (keys already in order) skips
// the radix passes and leaves sb.perm unset; first_hist_done: the first
// pass's histogram table is already in sb.hist (launch_keys_hist); offsets
// false: cell_off is left to the caller.
hipError_t launch_sort_precomputed(int64_t n, int key_range, SortBuffers& sb, hipStream_t s,
                                   const int* mode = nullptr, bool first_hist_done = false, bool offsets = true);
// dst[i] = src[perm[i]] for one float stream.
void launch_gather_f32(const float* src, float* dst, const int* perm, int64_t n, hipStream_t s);
// dst_k[i] = src_k[perm[i]] for ss.count streams in one pass (perm == nullptr: copy).
constexpr int kMaxStreams = 6;
struct StreamSet {
    const float* src[kMaxStreams];
    float* dst[kMaxStreams];
    int count;
};
// dst[k][i] = src[k][perm[i]] (perm null or *mode == 2: identity); with
// *mode == 1 from the records of launch_pack_streams (rec[W perm[i] + h],
// W = 1 for up to four streams, else 2).
void launch_permute_streams(const StreamSet& ss, const int* perm, int64_t n, hipStream_t s,
                            const int* mode = nullptr, const float4* rec = nullptr);
// Stable sort of the keys in sb.keys (each in [0, key_range)), the particle
// streams moving with the keys through the radix passes (ss.src -> ss.dst;
// tmp: ss.count intermediate streams of `stride` floats each), then cell_off.
// Mode 2 (already in order, set by launch_keys_hist / launch_permute_mode)
// copies the streams.  first_hist_done: launch_keys_hist left the first
// pass's histogram in sb.hist.  With kx, ky (the positions, in the unsorted
// order) and the domain, the keys are formed from them (cell keys, or node
// keys with `node`) instead of read from sb.keys.
hipError_t launch_sort_records(int64_t n, int key_range, SortBuffers& sb, const StreamSet& ss, float* tmp,
                               size_t stride, hipStream_t s, const int* mode, bool first_hist_done = true,
                               const float* kx = nullptr, const float* ky = nullptr, const DevDomain* d = nullptr,
                               bool node = false);
// The (key, index) sort with the streams gathered once: keys and the first
// pass's histogram in one kernel, the radix passes on (key, index) pairs, then
// one kernel that gathers ss.src into ss.dst by the permutation and writes
// cell_off.  Fewest launches, for sets too small to fill the GPU.
hipError_t launch_sort_gather(const float* x, const float* y, int64_t n, const DevDomain& d, int key_range,
                              bool node, SortBuffers& sb, const StreamSet& ss, hipStream_t s);
// Cell keys (node keys with `node`) of the current order into sb.keys (which
// holds the keys of the last sort's order), the first radix pass's histogram
// table, and from per-block counts (`part`, sort_part_elems(n) values) the
// permutation mode *mode: 2 when the keys are already in order (the passes
// are skipped and the streams copied), else 1 when more than an eighth of the
// particles moved by more than `far` keys (the permutation goes through
// launch_pack_streams before the radix passes and the records after them),
// else 0 (launch_permute_streams gathers stream by stream).
size_t sort_part_elems(int64_t n);
// *mode from nparts pairs (changed, descents) of per-block counts.
void launch_permute_mode(const unsigned* part, int nparts, int64_t n, int* mode, hipStream_t s);
// far < 0: no comparison with the last sort's keys; write_keys false: the
// keys are not stored (launch_sort_records forms them again from x, y).
void launch_keys_hist(const float* x, const float* y, int64_t n, const DevDomain& d, int key_range, SortBuffers& sb,
                      int far, bool node, unsigned* part, int* mode, hipStream_t s, bool write_keys = true);

// rec: n (up to 4 streams) or 2n (5 or 6 streams) float4 records.
void launch_pack_streams(const StreamSet& ss, int64_t n, const int* mode, float4* rec, hipStream_t s);
// Cell keys (particleCellIndex), or with `node` the keys of the node-keyed
// order (nearest node, node_key_exact).
void launch_keys(const float* x, const float* y, int64_t n, const DevDomain& d, int* keys, hipStream_t s,
                 bool node = false);

// ---- step.hip -------------------------------------------------------------
// push_flag (device int, may be null) overrides do_push when set.
void launch_gather_push(float* x, float* y, float* vx, float* vy, const float* q, const float* m, int64_t n,
                        const float* Ex, const float* Ey, const DevDomain& d, float dt, int do_gather, int do_push,
                        hipStream_t s, const int* push_flag = nullptr, bool q_one = false, bool m_one = false);

// FP64 sums: out[0] = sum(rho), out[1] = sum(Ex^2+Ey^2), out[2] = sum(0.5 m v^2),
// out[3] = sum(q), out[4] = sum(|q|).  `partials` holds 5 * nblocks doubles.
void launch_diagnostics(const float* rho, const float* Ex, const float* Ey, int64_t G, const float* vx,
                        const float* vy, const float* q, const float* m, int64_t n, double* partials, int nblocks,
                        double* out, hipStream_t s);

// Particle initialisers (Philox4x32-10 counter-based draws; see step.hip).
struct InitParams {
    int kind;       // 0 uniform positions, 1 x drawn from n(x) ~ 1 + amp cos(k x), 2 two beams +-amp
    uint64_t seed;
    double amp;     // density perturbation amplitude (kind 1) or beam speed (kind 2)
    double pert;    // velocity perturbation: vx += pert * sin(k x)
    double vth;     // Maxwellian spread per velocity component
    double vmax;    // uniform velocity U(-vmax, vmax) per component
    int mode;       // k = 2 pi mode / Lx
    double qp, mp;  // per-particle charge and mass
};
void launch_init(float* x, float* y, float* vx, float* vy, float* q, float* m, int64_t n, const DevDomain& d,
                 const InitParams& p, hipStream_t s);

// Spectral Poisson solve helpers (hipFFT plans live in api.hip).
// in[i] = (rho[i] + bg) * cv - mean, where mean * G / cv is the FP64 sum of the
// stored rho plus bg * G, taken as *rho_sum * sum_scale + sum_add (a sum of rho
// has scale 1; the particles' charge sum, which the deposit conserves, has
// scale 1 / cv).
void launch_poisson_prep(const float* rho, float* in, int64_t G, double bg, double cv, const double* rho_sum,
                         double sum_scale, double sum_add, hipStream_t s);
struct PoissonTables {
    double* k2x;     // [Nx/2+1] kx^2
    double* k2y;     // [Ny]     ky^2
    float2* dx_sym;  // [Nx/2+1] (e^{2 pi i i/Nx} - 1) / dx: forward difference along x
    float2* dy_sym;  // [Ny]     (e^{2 pi i j/Ny} - 1) / dy
};
// On the half spectrum [Ny][Nx/2+1]: phi_hat = rho_hat / (|k|^2 Nx Ny) with
// k = 0 zeroed; Ex_hat = -Dx phi_hat, Ey_hat = -Dy phi_hat (the spectral form
// of E = -forward difference of phi, identical in exact arithmetic and free of
// the cancellation of differencing an FP32 phi).
void launch_poisson_spectral(float2* spec, float2* spec_ex, float2* spec_ey, const PoissonTables& t,
                             const DevDomain& d, hipStream_t s);

// The same solve in three launches with the 1D transforms in LDS (csrc/poisson.hip),
// for power-of-two grids of 4 to 1024 cells along x and 1 to 1024 along y.  launch_poisson_lds takes
// rho (mean removed as in launch_poisson_prep) to the row half spectra
// [Ny][Nx/2+1] of phi (in spec), Ex and Ey; launch_poisson_lds_rows turns row
// half spectra X (and Y, may be null) into the real grids ox (and oy).
// tw_x, tw_y: e^{2 pi i m / Nx}, m < Nx, and e^{2 pi i m / Ny}, m < Ny.
bool poisson_lds_supported(int Nx, int Ny);
// The gate, renormalisation and write of a learned density (the last step of
// launch_denoise), deferred into the first pass of launch_poisson_lds when
// net_out is set: rho (and out64) become the learned density, or stay the
// plain deposit when the gate rejects it, then rho += bg, before the solve
// reads them.  One launch fewer per step of a learned run.
struct LearnedApply {
    float* net_out;  // null: nothing deferred
    double* out64;
    const double* sum;
    double sum_scale, inv_scale;
    double* stats;  // per-block gate sums (written by the layer that produces them)
    int nbs;
    const double* moments;
    int nbm;
    double cv, kappa, tol;
    double* result;
    int* fallback_count;
    float bg;
    // Output layer run by the solve's row pass (a3 set): it computes the
    // network output (net_out) for its rows from the third layer's
    // activations (BF16 NHWC, 32 channels) with the packed weights w4 and
    // bias *bias4 at dilation dil4, writes per-block gate sums (stats, nbs =
    // row blocks) and transforms the plain rows and the learned rows (into
    // spec_l, without the renormalising scale); the column pass takes the gate
    // decision and picks the spectrum, and the inverse-row pass writes rho and
    // out64.
    const void* a3 = nullptr;
    const void* w4 = nullptr;
    const float* bias4 = nullptr;
    int dil4 = 1;
    float2* spec_l = nullptr;
};
// The row pass can run the output layer (LearnedApply::a3) on this grid.
bool poisson_rows_can_run_output_layer(int Nx, int Ny, int dil);
void launch_poisson_lds(float* rho, int64_t G, double bg, double cv, const double* rho_sum, double sum_scale,
                        double sum_add, float2* spec, float2* spec_ex, float2* spec_ey, const PoissonTables& t,
                        const DevDomain& d, const float2* tw_x, const float2* tw_y, hipStream_t s,
                        const LearnedApply& la = LearnedApply{});
// With la.a3 set (the Ex / Ey call of a learned step), also writes the
// learned or plain density into rho and la.out64 as the gate decided.
void launch_poisson_lds_rows(const float2* X, const float2* Y, float* ox, float* oy, const DevDomain& d,
                             const float2* tw_x, hipStream_t s, float* rho = nullptr,
                             const LearnedApply& la = LearnedApply{});
// Ex = -(phi(i+1,j) - phi(i,j))/dx, Ey = -(phi(i,j+1) - phi(i,j))/dy.
void launch_efield(const float* phi, float* Ex, float* Ey, const DevDomain& d, hipStream_t s);
void launch_add_scalar(float* p, int64_t n, float v, hipStream_t s);

// Per-step FP64 records (any pointer may be null).
struct StepRecords {
    double* field_energy;  // 0.5 * sum(Ex^2+Ey^2) * cellVolume
    double* mode1;         // |mode-1 Fourier amplitude of Ex| (extractFourierModeAmplitude)
    double* kinetic;       // sum 0.5 m v^2
    double* charge_error;  // |integrated rho - sum q| / sum |q|
};
// Sums over the grid (and particles when n > 0), appends one record at
// *counter - record_from (if inside [0, cap)) and increments *counter.
void launch_step_diag(const float* rho, const float* Ex, const float* Ey, const double* cosk, const double* sink,
                      const DevDomain& d, const float* vx, const float* vy, const float* q, const float* m, int64_t n,
                      double* partials, int nblocks, int* counter, int record_from, int cap, const StepRecords& rec,
                      hipStream_t s);
void launch_set_int(int* p, int v, hipStream_t s);

// ---- current.hip -------------------------------------------------------------
// Charge-conserving current and the CIC densities at both ends of the move;
// the *_c grids are TwoSum compensation terms (zero between calls).
struct CurrentGrids {
    float *Jx, *Jy, *rho_old, *rho_new;
    float *Jx_c, *Jy_c, *rho_old_c, *rho_new_c;
};
void launch_esirkepov_current(const float* x, const float* y, const float* vx, const float* vy, const float* q,
                              int64_t n, const DevDomain& d, float dt, const CurrentGrids& g, int* status,
                              hipStream_t s);
// The same on cell-sorted particles with their cell offsets: one wavefront
// per tile of tg (at most 128 cells, no drift halo), FP64 LDS windows written to
// `partial` (current_partial_doubles(tg) doubles), then a merge.
size_t current_partial_doubles(const TileGeom& tg);
size_t current_tile_lds(const TileGeom& tg);
void launch_esirkepov_current_tiled(const float* x, const float* y, const float* vx, const float* vy, const float* q,
                                    const int* cell_off, const DevDomain& d, const TileGeom& tg, float dt,
                                    double* partial, const CurrentGrids& g, int* status, hipStream_t s);
// out[0] = max |(rho_new - rho_old)/dt + Dx Jx + Dy Jy|, out[1] = max |(rho_new - rho_old)/dt| (FP64).
void launch_continuity_residual(const CurrentGrids& g, const DevDomain& d, double dt, double* partial, int nblocks,
                                double* out, hipStream_t s);
void launch_copy_f32(const float* src, float* dst, int64_t n, hipStream_t s);
// 48-byte FP64 particle records (x, y, vx, vy, q, m; rec 16-byte aligned)
// <-> the six FP32 streams.
void launch_aos_unpack(const double* rec, int64_t n, float* const dst[6], hipStream_t s);
void launch_aos_pack(const float* const src[6], int64_t n, double* rec, hipStream_t s);
// Streaming read of n4 float4 values, one partial sum per block.
void launch_read_probe(const float* src, int64_t n4, float* partial, int blocks, hipStream_t s);

// ---- ensemble.hip -----------------------------------------------------------
// Particle slots per chunk; each member's slot range is chunk aligned, so a
// chunk (one workgroup of the per-particle kernels) belongs to one member.
constexpr int kEnsChunk = 4096;

struct EnsMember {
    DevDomain d;
    TileGeom g;
    InitParams init;
    int scheme, ppl, user;  // deposit scheme, particles per lane, index in the caller's member list
    int64_t n;              // particles
    int64_t pbase;          // first particle slot (multiple of kEnsChunk)
    int64_t cbase;          // first cell key; key cbase + G collects the member's padding slots
    int64_t gbase;          // first node of the member's rho / Ex / Ey / overflow grids
    int64_t G;              // Nx * Ny
    int64_t partial_base;   // first float of the member's tile windows
    int64_t sbase;          // first element of its half spectrum [Ny][Nx/2+1]
    int64_t tbase;          // first entry of its k^2 and difference-symbol tables [Nx/2+1 + Ny]
    int64_t xbase;          // first entry of its mode-1 cos / sin tables [Nx]
};

// Member m owns blocks [start[m], start[m+1]) of each launch family.
struct EnsTables {
    const EnsMember* mem;
    int M;
    const int64_t* pchunk;  // particle chunks
    const int64_t* gblk;    // grid blocks of 256 nodes
    const int64_t* sblk;    // spectrum blocks of 256 elements
};

struct EnsParticles {
    float *x, *y, *vx, *vy, *q, *m;
};

// One persistent deposit launch: the (member, tile) items of the members
// with this scheme, particles per lane and coordinate path.
struct EnsWork {
    int scheme, ppl, exact;  // scheme: a C-ABI scheme, or TSC_MOM for TSC members in the moment form
    int drift;  // the windows have a drift halo: lookahead chunk processing
    const int2* items;
    int nitems;
    int lds_per_wave;  // floats, the largest of the group's members
    int grid;          // workgroups of the persistent launch (set by ens_deposit_grid)
};

// Workgroups for a persistent launch of group w: the resident workgroups per
// CU of its kernel at its LDS size times ncu, at most one per four items.
// Queries the runtime, so it is called outside stream capture.
int ens_deposit_grid(const EnsWork& w, int ncu);

void launch_ens_init(const EnsParticles& p, const EnsTables& t, int64_t nchunks, hipStream_t s);
// With `part` (2 * nchunks values), also the per-chunk counts that
// launch_permute_mode turns into a permutation mode (see launch_keys_hist).
void launch_ens_keys(const EnsParticles& p, const EnsTables& t, int64_t nchunks, int* keys, hipStream_t s,
                     unsigned* part = nullptr);
// counters: one int per group, reset by the launch itself.
hipError_t launch_ens_deposit(const EnsWork* groups, int ngroups, const EnsParticles& p, const int* cell_off,
                              const EnsTables& t, float dt, float* partial, double* gacc,
                              int* status, int* counters, int ncu, hipStream_t s);
// rho of grid blocks [blk0, blk0 + nblk) and their FP64 block sums rho_part[b].
void launch_ens_merge(const float* partial, double* gacc, const EnsTables& t, int64_t blk0,
                      int64_t nblk, float* rho, double* rho_part, hipStream_t s);
// out[m] = sum of part[b * stride + k] over member m's grid blocks (k fixed).
void launch_ens_member_grid_sum(const double* part, int stride, int k, const EnsTables& t, double* out,
                                hipStream_t s);
void launch_ens_prep(const float* rho, float* in, const EnsTables& t, int64_t nblk, const double* rho_sum,
                     hipStream_t s);
void launch_ens_spectral(const float2* spec, float2* spec_ex, float2* spec_ey, const double* k2, const float2* sym,
                         const EnsTables& t, int64_t nblk, hipStream_t s);
// Per grid block: sum(Ex^2 + Ey^2), sum Ex cos(k x_i), sum Ex sin(k x_i).
void launch_ens_field_sums(const float* Ex, const float* Ey, const double* cosk, const double* sink,
                           const EnsTables& t, int64_t nblk, double* field_part, hipStream_t s);
// Gather (v += -(q/m) E dt) then, if *push_flag, push; per chunk FP64 sums of
// 0.5 m v^2, q and |q| into part_p[3 c + k].
void launch_ens_gather_push(const EnsParticles& p, const float* Ex, const float* Ey, const EnsTables& t,
                            int64_t nchunks, float dt, int do_gather, const int* push_flag, double* part_p,
                            hipStream_t s);
// One block per member: records field energy, mode-1 amplitude, kinetic
// energy and charge error at [user][counter[m] - record_from] if inside
// [0, cap), then counter[m] += 1.
void launch_ens_diag(const double* rho_part, const double* field_part, const double* part_p, const EnsTables& t,
                     int* counters, int record_from, int cap, const StepRecords& rec, hipStream_t s);
void launch_set_ints(int* p, int n, int v, hipStream_t s);

// ---- denoise.hip --------------------------------------------------------------
// Parameter vector of the 1 -> 32 -> 32 -> 32 -> 1 network (FP32 master copy):
// W1[co][tap], b1, W2[tap][co][ci], b2, W3[tap][co][ci], b3, W4[tap][ci], b4.
namespace cnn {
constexpr int kW1 = 0, kB1 = 288, kW2 = 320, kB2 = 9536, kW3 = 9568, kB3 = 18784, kW4 = 18816, kB4 = 19104;
constexpr int kParams = 19105;
// Packed BF16 B operands (units of 8 BF16 = one lane's operand, 64 per block):
// forward layers 1-4, then the input-gradient operands of layers 4, 3, 2.
constexpr int kPackF2 = 128, kPackF3 = 1280, kPackF4 = 2432, kPackB4 = 3008, kPackB3 = 3136, kPackB2 = 4288;
constexpr int kPackVecs = 5440;
}  // namespace cnn

struct CnnGeom {
    int B, H, W;  // images, rows, columns (W a multiple of 32)
};

struct CnnBuffers {
    float *w, *grad, *m, *v;  // [kParams]
    void* packed;             // [kPackVecs * 8] BF16
    float* x0;                // input  [B][H][W]
    float* out;               // output [B][H][W]
    float* target;            // [B][H][W]
    float* dout;              // d loss / d out [B][H][W]
    void *a1, *a2, *a3;       // BF16 [B][H][W][32] activations
    void *d2, *d3;            // BF16 [B][H][W][32] gradients
    float* wpart;             // cnn_wpart_floats(ncu)
    double* bias_part;        // 256 * 32
    double* loss_part;        // 256
};

void launch_cnn_forward(const CnnBuffers& c, const CnnGeom& g, const int (&dil)[4], int ncu, hipStream_t s);
// Gradients of the loss into c.grad from c.dout (needs the forward activations).
void launch_cnn_backward(const CnnBuffers& c, const CnnGeom& g, const int (&dil)[4], int ncu, hipStream_t s);
size_t cnn_wpart_floats(int ncu);
// Mean-squared error of c.out against c.target into *loss, and c.dout.
void launch_cnn_loss(const CnnBuffers& c, const CnnGeom& g, double* loss, hipStream_t s);
void launch_cnn_adam(const CnnBuffers& c, float lr, float b1, float b2, float eps, int step, hipStream_t s);
void launch_cnn_pack(const CnnBuffers& c, hipStream_t s);
// Training batch: image b = pair items[b].x of the sets, rolled by (items[b].y,
// items[b].z); inputs times sa, targets times sb.
void launch_cnn_batch(const float* inputs, const float* targets, const int4* items, const CnnGeom& g, float sa,
                      float sb, float* in_b, float* tg_b, hipStream_t s);
// delta = scale * (rho / (rho_sum / n) - 1).
void launch_normalize(const float* rho, int64_t n, const double* rho_sum, double scale, float* delta,
                      hipStream_t s);
// Per-block FP64 sums of v and v^2 into partial[2 * block + {0, 1}].
void launch_moments(const float* v, int64_t n, double* partial, int nb, hipStream_t s);
// Learned density of one [H][W] grid: rho normalised on the fly (mean of rho =
// *sum * sum_scale / G), the network, then the gate and renormalisation:
// result = {ratio, fallback, scale, sum q}.  When the gate passes, rho (FP32)
// and out64 hold the renormalised learned density; otherwise rho is left as
// deposited, out64 is its FP64 copy and *fallback_count (if not null) is
// incremented.  `moments` holds nbm per-block charge moments (launch_moments).
struct DenoiseArgs {
    float* rho;
    double* out64;
    int64_t G;
    const double* sum;
    double sum_scale;
    double in_scale, inv_out_scale;  // network input scale, 1 / output scale
    const double* moments;
    int nbm;
    double cv, kappa, tol;
    double* stats;  // denoise_stats_doubles(ncu)
    double* result;
    int* fallback_count;
};
// With `defer`, the final gate / renormalisation launch is not made: its
// arguments are written to *defer for launch_poisson_lds (bg left 0).
// With `merge`, rho is not yet written: the first layer merges the deposit's
// tile windows itself (writing rho) and the second clears the FP64
// accumulation grid the merge folded (needs dil[0] == 1, see
// denoise_can_merge).
// With `spec_l` (and `defer`), the output layer is left to the solve's row
// pass as well (poisson_rows_can_run_output_layer; spec_l holds the learned
// half spectrum).
void launch_denoise(const CnnBuffers& c, const CnnGeom& g, const int (&dil)[4], int ncu, const DenoiseArgs& a,
                    hipStream_t s, LearnedApply* defer = nullptr, const TileMerge* merge = nullptr,
                    float2* spec_l = nullptr);
bool denoise_can_merge(const CnnGeom& g, const int (&dil)[4]);
size_t denoise_stats_doubles(int ncu);

}  // namespace pichip
