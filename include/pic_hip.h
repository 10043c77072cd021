/*
 * pic_hip.h - C ABI of the MI355X (gfx950) charge-deposition library.
 *
 * Plain pointers and sizes only.  Every call returns a pic_status; on failure
 * pic_hip_last_error() holds a message for the calling thread.  No call throws.
 *
 * Reference interfaces replaced (paths relative to the ehocine/pic-deposition
 * repository root):
 *   - depositChargeCudaSoA / depositChargeCudaAoS
 *       src/deposition/cuda/deposition_cuda.hpp:11-12,
 *       src/deposition/cuda/deposition_cuda.cu:288-337
 *     -> pic_hip_deposit_host, pic_hip_deposit_host_aos (synchronous drop-ins:
 *        FP64 host in, FP64 rho out).
 *   - the device side of depositChargeSoA / depositChargeTimedSoA
 *       src/deposition/deposition.cpp:120-132,181-200,257-310
 *     -> pic_hip_create / pic_hip_upload / pic_hip_deposit / pic_hip_get_rho
 *        (device-resident particles, asynchronous on the context's stream).
 *   - sortParticlesByCell / particleCellIndex
 *       src/sorting/cell_sort.cpp:8-46, cell_sort.hpp:8-11
 *     -> pic_hip_sort, pic_hip_sort_perm, pic_hip_get_keys.
 *   - gatherFieldsSoA / pushParticlesSoA   src/deposition/deposition.cpp:335-374
 *     -> pic_hip_gather_push.
 *   - PoissonSolverFFT::solve               src/pic/poisson_fft.cpp:45-94
 *     -> pic_hip_poisson.
 *   - FieldGrid::integratedRho / fieldEnergy, ParticlesSoA::totalCharge /
 *     kineticEnergy, totalAbsChargeSoA      src/pic/field_grid.cpp:24-38,
 *     src/pic/particles_aos.cpp:182-196, src/pic/diagnostics.cpp:18-51
 *     -> pic_hip_diagnostics (FP64 sums on the device).
 *   - applyNeutralizingBackground           src/pic/validation.cpp:56-61,
 *     src/pic/simulation.cpp:52-59
 *     -> pic_hip_add_background.
 */
#ifndef PIC_HIP_H
#define PIC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PIC_OK = 0,
    PIC_ERR_INVALID_ARGUMENT = 1,
    PIC_ERR_HIP = 2,             /* HIP runtime error (message has the HIP error string) */
    PIC_ERR_OUT_OF_MEMORY = 3,
    PIC_ERR_UNSUPPORTED = 4,     /* e.g. Esirkepov move longer than 15 cells */
    PIC_ERR_NOT_SORTED = 5,      /* GPU_Priv deposit without a cell order (call pic_hip_sort) */
    PIC_ERR_FFT = 6
} pic_status;

/* DepositionScheme, reference src/deposition/deposition.hpp:11, same order as
 * schemeToInt (deposition_cuda.cu:275-286).  3 is the reference's
 * trajectory-averaged Esirkepov density (ngp.cpp:13-84), computed for real
 * (the reference CUDA path substitutes CIC). */
typedef enum { PIC_SCHEME_NGP = 0, PIC_SCHEME_CIC = 1, PIC_SCHEME_TSC = 2, PIC_SCHEME_ESIRKEPOV = 3 } pic_scheme;

/* DepositionBackend, reference deposition.hpp:13 (names CPU / GPU_Atomics /
 * GPU_Priv, deposition.cpp:390-415).  CPU is not served by this library. */
typedef enum { PIC_BACKEND_CPU = 0, PIC_BACKEND_GPU_ATOMICS = 1, PIC_BACKEND_GPU_PRIV = 2 } pic_backend;

/* pic::Domain (src/pic/domain.hpp:7-44): dx = Lx / Nx, dy = Ly / Ny. */
typedef struct {
    double Lx, Ly;
    int32_t Nx, Ny;
} pic_domain;

/* FP64 device sums, see pic_hip_diagnostics. */
typedef struct {
    double rho_sum;          /* sum of rho over cells */
    double integrated_rho;   /* rho_sum * cellVolume          (field_grid.cpp:24-30) */
    double field_energy;     /* 0.5 * sum(Ex^2+Ey^2) * cellVolume (field_grid.cpp:32-38) */
    double kinetic_energy;   /* sum 0.5 m v^2                 (particles_aos.cpp:190-196) */
    double total_charge;     /* sum q                          (particles_aos.cpp:182-188) */
    double total_abs_charge; /* sum |q|                        (diagnostics.cpp:18-24) */
    double charge_error;     /* |integrated_rho - total_charge| / max(sum|q|, 1e-12) (diagnostics.cpp:26-29) */
} pic_diagnostics;

/* On-device particle initialiser (build-defined; Philox4x32-10 draws keyed by
 * seed, counter = particle index, so the result does not depend on launch
 * shape).  kind 0: x, y uniform; kind 1: x drawn from n(x) ~ 1 + amp*cos(k x)
 * (k = 2 pi mode / Lx), y uniform; kind 2: x, y uniform, first half of the
 * particles at vx = +amp, second half at -amp (the split of
 * initializeTwoStream, particles_aos.cpp:142-160).  Every kind then adds
 * pert*sin(k x) to vx, N(0, vth^2) to vx and vy, and U(-vmax, vmax) to vx and
 * vy.  Positions are rounded to FP32 and wrapped into [0, L). */
typedef struct {
    int32_t kind;
    uint64_t seed;
    double amp, pert, vth, vmax;
    int32_t mode;
    double q, m;
} pic_init_params;

typedef struct pic_hip_ctx pic_hip_ctx;

/* Message of the last failed call on this thread ("" if none). */
const char* pic_hip_last_error(void);
const char* pic_hip_version(void);

/* ---- drop-in seam --------------------------------------------------------
 * Replaces depositChargeCudaSoA (deposition_cuda.cu:335-337 -> launchDeposition
 * :288-326).  Synchronous; caller-owned FP64 SoA in (never reordered), FP64 rho
 * out (all Nx*Ny cells overwritten, as the reference's device memset + D2H).
 * Positions and charges are rounded to FP32 on the way in.  backend
 * GPU_ATOMICS deposits in the given order; GPU_PRIV sorts a device copy by
 * cell first.  `sorted` mirrors DepositionConfig::sorted: the device copy is
 * cell-sorted before depositing (for GPU_PRIV it always is).  vx, vy and dt
 * are read only for the Esirkepov scheme (DepositionConfig::esirkepov_dt). */
pic_status pic_hip_deposit_host(const double* x, const double* y, const double* vx, const double* vy,
                                const double* q, size_t n, const pic_domain* domain, int scheme, int backend,
                                int sorted, double dt, double* rho_out);
/* Replaces depositChargeCudaAoS (deposition_cuda.cu:330-333, which copies the
 * records into a new FP64 SoA with toSoA before the SoA path).  `records` holds
 * n particles in the layout of ParticleAoS (src/pic/particles.hpp:13-20): six
 * FP64 values x, y, vx, vy, q, m per 48-byte record.  Same semantics and
 * results as pic_hip_deposit_host on the same particles; the host threads
 * round straight from the records, without the intermediate SoA copy. */
pic_status pic_hip_deposit_host_aos(const double* records, size_t n, const pic_domain* domain, int scheme,
                                    int backend, int sorted, double dt, double* rho_out);

/* ---- device-resident context --------------------------------------------
 * One HIP stream per context; calls enqueue work asynchronously on it and are
 * ordered; pic_hip_sync waits and reports kernel-side failures.  A context is
 * not thread-safe; use one per host thread. */
pic_status pic_hip_create(const pic_domain* domain, size_t capacity, pic_hip_ctx** out);
pic_status pic_hip_destroy(pic_hip_ctx* ctx);
/* The context's hipStream_t, for event timing by the caller. */
void* pic_hip_stream(pic_hip_ctx* ctx);
pic_status pic_hip_sync(pic_hip_ctx* ctx);
size_t pic_hip_count(pic_hip_ctx* ctx);

/* LDS tiling of the GPU_PRIV deposit for one scheme (or all with scheme -1):
 * tiles of tile_x by tile_y cells, particles_per_lane in {4, 8}, pipelined =
 * overlap the next chunk's loads with the current chunk (two register sets).
 * Build-defined tuning knobs (no reference counterpart); defaults are chosen
 * at create time and pic_hip_autotune searches them. */
pic_status pic_hip_set_tiling(pic_hip_ctx* ctx, int scheme, int tile_x, int tile_y, int particles_per_lane,
                              int pipelined);
/* Drift halo (0..4 cells, default 0) of every scheme's tile windows: a tile
 * deposits into LDS the particles sorted into it whose current cell is within
 * `cells` of the tile, so that particles that moved since the last sort stay
 * off the global-atomic overflow path.  Costs window LDS and merge reads.
 * Schemes still on their create-time tiling take the default shape for the
 * new halo (long-row tiles on large grids only at halo 0); tilings set with
 * pic_hip_set_tiling or pic_hip_autotune are kept. */
pic_status pic_hip_set_tile_halo(pic_hip_ctx* ctx, int cells);
int pic_hip_get_tile_halo(pic_hip_ctx* ctx);
pic_status pic_hip_get_tiling(pic_hip_ctx* ctx, int scheme, int* tile_x, int* tile_y, int* particles_per_lane,
                              int* pipelined);

/* Particles: FP32 SoA x, y, vx, vy, q, m.  NULL vx / vy upload zeros, NULL m
 * uploads ones; x, y, q are required.  Upload clears the cell order. */
pic_status pic_hip_upload(pic_hip_ctx* ctx, const float* x, const float* y, const float* vx, const float* vy,
                          const float* q, const float* m, size_t n);
/* Copies the current device order back; NULL pointers are skipped.  Synchronous. */
pic_status pic_hip_download(pic_hip_ctx* ctx, float* x, float* y, float* vx, float* vy, float* q, float* m);
/* AoS layout on the device side (ParticleAoS records, src/pic/particles.hpp:
 * 13-20; the reference's device path accepts AoS only through toSoA,
 * deposition_cuda.cu:330-333).  pic_hip_upload_aos copies n FP64 records
 * (records 16-byte aligned) to the device as they are and converts them there
 * into the FP32 streams (rounded to nearest: the same streams as
 * pic_hip_upload of the values cast to float).  pic_hip_download_aos writes the
 * current device order back as FP64 records (exact promotion).  Both are
 * synchronous and stage through device memory in chunks of up to 2^22 records. */
pic_status pic_hip_upload_aos(pic_hip_ctx* ctx, const double* records, size_t n);
pic_status pic_hip_download_aos(pic_hip_ctx* ctx, double* records);
pic_status pic_hip_init_particles(pic_hip_ctx* ctx, const pic_init_params* params, size_t n);

/* Cell sort.  pic_hip_sort computes every particle's cell key
 * (particleCellIndex), a stable permutation by key, applies it to all six
 * streams, and builds the cell offset table.  pic_hip_sort_perm computes the
 * same without moving the particles and copies out (synchronously) the sorted
 * keys, perm[i] = source index of the i-th sorted particle, and the Nx*Ny+1
 * cell offsets; NULL outputs are skipped.  The first pic_hip_sort or
 * pic_hip_run of a context allocates 32 bytes per particle of capacity for
 * the intermediate streams of the radix passes. */
pic_status pic_hip_sort(pic_hip_ctx* ctx);
/* Sort by nearest node instead of by cell (key floor(x/dx) + [fraction >= 1/2],
 * wrapped; build-defined, no reference counterpart).  On this order the
 * GPU_PRIV NGP and TSC deposits run node-centred stencils of 1 and 9 slots
 * instead of 4 and 16 (same weights, same results to FP32 rounding); CIC and
 * Esirkepov deposits stay exact on it through their windows and overflow path
 * but prefer pic_hip_sort's order.  pic_hip_run sorts NGP and TSC runs this way;
 * pic_hip_deposit_host does for GPU_PRIV NGP and TSC. */
pic_status pic_hip_sort_nodes(pic_hip_ctx* ctx);
pic_status pic_hip_sort_perm(pic_hip_ctx* ctx, int32_t* keys_sorted, int32_t* perm, int32_t* cell_offsets);
/* How the context's last sort (pic_hip_sort, pic_hip_sort_nodes or a sort
 * inside pic_hip_run) moved the particles, decided on the device: 0 the radix
 * passes, which carry the streams with the keys, 2 a copy (the order was
 * still sorted); -1 before any sort, or for sets below 2^22 particles, which
 * always take (key, index) radix passes and one gather.  Synchronous.
 * Build-defined, no reference counterpart (the reference's std::sort has no
 * such choice). */
pic_status pic_hip_sort_path(pic_hip_ctx* ctx, int32_t* path);
/* Keys of the current order (synchronous copy out). */
pic_status pic_hip_get_keys(pic_hip_ctx* ctx, int32_t* keys);

/* Deposit into the context's FP32 rho (all cells overwritten).  GPU_PRIV
 * needs a cell order from pic_hip_sort; particles that moved out of their
 * cell since then are still deposited exactly (global-atomic overflow path). */
pic_status pic_hip_deposit(pic_hip_ctx* ctx, int scheme, int backend, double dt);
pic_status pic_hip_get_rho(pic_hip_ctx* ctx, float* rho);
pic_status pic_hip_get_rho_f64(pic_hip_ctx* ctx, double* rho);
pic_status pic_hip_set_rho(pic_hip_ctx* ctx, const float* rho);

/* rho += value on every cell (neutralising background 1/V). */
pic_status pic_hip_add_background(pic_hip_ctx* ctx, double value);
/* phi, Ex, Ey from rho: spectral solve with k = 0 removed (FP64 mean), FP32
 * transforms (in LDS for power-of-two grids up to 1024 cells per axis, hipFFT
 * otherwise), E = -forward difference of phi. */
pic_status pic_hip_poisson(pic_hip_ctx* ctx);
pic_status pic_hip_get_fields(pic_hip_ctx* ctx, float* phi, float* Ex, float* Ey);
pic_status pic_hip_set_fields(pic_hip_ctx* ctx, const float* Ex, const float* Ey);
/* do_gather: v += -(q/m) E(x) dt (CIC interpolation); then do_push:
 * x += v dt wrapped periodically. */
pic_status pic_hip_gather_push(pic_hip_ctx* ctx, double dt, int do_gather, int do_push);
/* FP64 sums over rho, Ex/Ey and the particles (synchronous). */
pic_status pic_hip_diagnostics(pic_hip_ctx* ctx, pic_diagnostics* out);

/* ---- charge-conserving current (build-defined; no reference counterpart) --
 * Esirkepov (Comput. Phys. Commun. 135 (2001) 144) current of the move
 * x -> x + v dt with CIC shape factors: Jx on x faces (i + 1/2, j), Jy on y
 * faces (i, j + 1/2), plus the CIC densities at both ends, so that
 * (rho_new - rho_old)/dt + Dx Jx + Dy Jy = 0 with backward differences.
 * Moves must be shorter than one cell (PIC_ERR_UNSUPPORTED otherwise, at the
 * next pic_hip_sync).  Any particle order. */
pic_status pic_hip_esirkepov_current(pic_hip_ctx* ctx, double dt);
pic_status pic_hip_get_current(pic_hip_ctx* ctx, float* Jx, float* Jy, float* rho_old, float* rho_new);
/* FP64 max over cells of the continuity residual and of |d rho / dt|. */
pic_status pic_hip_continuity_residual(pic_hip_ctx* ctx, double dt, double* max_residual, double* max_drho_dt);

/* ---- timestep driver ------------------------------------------------------
 * Runs nsteps of the electrostatic loop on the device.  Step orders follow
 * the reference loops:
 *   PIC_ORDER_PUSH_FIRST          push, deposit, +background, Poisson, gather
 *       (Simulation::run for NGP/CIC/TSC, simulation.cpp:85-96,112-122;
 *        validation.cpp:225-235 and :375-385)
 *   PIC_ORDER_DEPOSIT_SOLVE_PUSH  deposit, +background, Poisson, push, gather
 *       (Esirkepov in Simulation::run and the Langmuir loop, validation.cpp:226-233)
 *   PIC_ORDER_DEPOSIT_PUSH_SOLVE  deposit, push, +background, Poisson, gather
 *       (Esirkepov in the two-stream loop, validation.cpp:375-379)
 * After each step the FP64 diagnostics are recorded for steps >= record_from:
 * field energy (FieldGrid::fieldEnergy), |mode-1 Fourier amplitude of Ex|
 * (extractFourierModeAmplitude, validation.cpp:99-119) and, when
 * record_particles is set, kinetic energy and charge error
 * (computeDiagnostics, diagnostics.cpp:42-51).  Output arrays hold
 * nsteps - record_from values; NULL skips.  With GPU_PRIV the particles are
 * re-sorted by cell before the deposit of every sort_interval-th step (first
 * step included); between sorts the deposit stays exact through the drift
 * halo of its tile windows and its overflow path.  use_graph captures one step
 * (with and without sort, with and without the diagnostics record) as
 * hipGraphs and replays them.  Diagnostics run only on recorded steps.  The
 * run's Poisson solves take the mean of rho from the particles' total charge
 * (which the deposit conserves) and leave phi as a spectrum: pic_hip_get_fields
 * transforms it when phi is read. */
enum { PIC_ORDER_PUSH_FIRST = 0, PIC_ORDER_DEPOSIT_SOLVE_PUSH = 1, PIC_ORDER_DEPOSIT_PUSH_SOLVE = 2 };

typedef struct {
    int32_t scheme;
    int32_t backend;
    int32_t sort_interval;
    int32_t order;
    double dt;
    double background;
    int32_t record_particles;
    int32_t use_graph;
} pic_step_config;

pic_status pic_hip_run(pic_hip_ctx* ctx, const pic_step_config* cfg, int32_t nsteps, int32_t record_from,
                       double* field_energy, double* mode1_amplitude, double* kinetic_energy, double* charge_error);

/* ---- allocation-free deposit ----------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (work under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the global-atomic deposit of pic_hip_deposit with backend
 * GPU_ATOMICS, on caller-owned device arrays, enqueued on `stream` (a
 * hipStream_t; NULL is the null stream) of the current device, asynchronous.
 * x, y, q (and vx, vy for the Esirkepov scheme; otherwise they may be NULL)
 * hold n FP32 particles in any order; rho (Nx * Ny FP32, row-major) is
 * overwritten.  The call allocates nothing: its FP64 accumulation grid is a
 * file-scope device array of PIC_NOALLOC_MAX_CELLS cells (16 MiB plus one
 * 256-byte status line per device, held until the process ends); larger grids
 * are rejected with PIC_ERR_INVALID_ARGUMENT.  The array is cleared at the
 * start of every call; calls that overlap on one device (several streams or
 * host threads) race on it and must be serialised by the caller.  An
 * Esirkepov move longer than 15 cells, which pic_hip_sync reports for a
 * context, sets every cell of rho to NaN instead. */
#define PIC_NOALLOC_MAX_CELLS 2097152
pic_status pic_hip_deposit_noalloc(const pic_domain* domain, const float* x, const float* y, const float* vx,
                                   const float* vy, const float* q, size_t n, int scheme, double dt, float* rho,
                                   void* stream);

/* Per-phase timing of the timestep, replacing Simulation::runProfiled
 * (src/pic/simulation.cpp:147-292, TimestepProfile simulation.hpp:40-48; rows
 * written by benchmark_runner.cpp:224-306).  Runs warmup_steps + timed_steps
 * eager steps of the reference's SoA order ([sort every sort_interval-th
 * step] -> push -> deposit -> +background -> Poisson -> gather; Esirkepov
 * pushes after the Poisson solve) and returns the mean milliseconds per timed
 * step of each phase, measured with HIP events on the context's stream.
 * GPU_Atomics sorts only with sort_interval >= 1 (SimulationConfig::sorted);
 * NGP and TSC GPU_Priv steps sort by nearest node.  The deposit reads the
 * pushed particles (the reference deposits a copy taken before the push,
 * simulation.cpp:191,203-208).  use_graph and record_particles are ignored.
 * Synchronous. */
typedef struct {
    double push_ms, sort_ms, deposit_ms, poisson_ms, gather_ms, total_ms;
    int32_t sort_interval;
} pic_timestep_profile;
pic_status pic_hip_run_profiled(pic_hip_ctx* ctx, const pic_step_config* cfg, int32_t warmup_steps,
                                int32_t timed_steps, pic_timestep_profile* out);

/* Physics estimators on recorded series (host FP64; reference
 * src/pic/validation.cpp).  Dominant frequency: DFT peak bin of the
 * mean-removed series, k / (n dt) (:19-54).  Growth rate: least-squares slope
 * of ln(max(s, 1e-30)) over the middle half, times t = (record_start + i) dt
 * (:121-152).  Plasma frequency sqrt(n0 q^2 / m) with |q| = 1/N, n0 = N/V
 * (:66-71). */
double pic_hip_dominant_frequency(const double* signal, int64_t n, double dt);
double pic_hip_growth_rate(const double* signal, int64_t n, double dt, int32_t record_start_step);
double pic_hip_plasma_frequency(const pic_domain* domain, int64_t num_particles);
/* coldTwoStreamGrowthRate (validation.cpp:73-76): omega_p / (2 sqrt 2).
 * resonantBeamVelocity (validation.cpp:78-85): omega_p / k, k = 2 pi mode / Lx. */
double pic_hip_cold_two_stream_growth_rate(const pic_domain* domain, int64_t num_particles);
double pic_hip_resonant_beam_velocity(const pic_domain* domain, int64_t num_particles, int32_t mode);

/* The reference's particle initialisers (ParticlesSoA::initialize*,
 * src/pic/particles_aos.cpp:67-180), drawn on the host with the C++ standard
 * library's std::mt19937(seed), uniform_real_distribution over [0, L) and a
 * fresh normal_distribution per velocity draw, in the reference's order; FP64
 * out (the caller rounds to FP32 for pic_hip_upload).  The draws are the
 * reference's wherever both are built with the same standard library (the
 * uniform draws with any; libstdc++ here).
 *   PIC_REF_MAXWELLIAN      x, y uniform; vx, vy ~ N(0, temperature); q = +1, -1
 *                           alternating; m = 1 (:67-80)
 *   PIC_REF_LANGMUIR        lattice x = Lx i / n, y = Ly / 2, vx = amplitude
 *                           sin(2 pi mode x / Lx), vy = 0, q = -1/n (:82-95)
 *   PIC_REF_WARM_LANGMUIR   the lattice plus N(0, temperature) draws on vx then
 *                           vy per particle (:97-112)
 *   PIC_REF_TWO_STREAM      x, y uniform; vx = +-beam_velocity (first half +)
 *                           + perturbation sin(2 pi x / Lx); q = -1/n (:142-160)
 *   PIC_REF_TWO_STREAM_1D   as TWO_STREAM with y = Ly / 2, x the only draw
 *                           (:162-180) */
enum {
    PIC_REF_MAXWELLIAN = 0,
    PIC_REF_LANGMUIR = 1,
    PIC_REF_WARM_LANGMUIR = 2,
    PIC_REF_TWO_STREAM = 3,
    PIC_REF_TWO_STREAM_1D = 4
};
typedef struct {
    int32_t kind;
    uint32_t seed;
    double temperature;
    double amplitude;
    int32_t mode;
    double beam_velocity;
    double perturbation;
} pic_reference_init;
pic_status pic_hip_reference_particles(const pic_domain* domain, const pic_reference_init* init, size_t n, double* x,
                                       double* y, double* vx, double* vy, double* q, double* m);

/* ---- kernel timing (build-defined; no reference counterpart) -------------
 * Between begin and end, every deposit records a pair of HIP events on the
 * context's stream around its main scatter kernel (the tiled kernel for
 * GPU_PRIV, the atomic kernel for GPU_ATOMICS).  end waits for the stream and
 * returns the summed kernel time and the number of launches recorded
 * (at most max_launches). */
pic_status pic_hip_profile_begin(pic_hip_ctx* ctx, int max_launches);
pic_status pic_hip_profile_end(pic_hip_ctx* ctx, double* total_ms, int* launches);
/* Copies `bytes` (rounded down to 4) between two scratch device buffers with
 * a plain streaming kernel (copy_f32_kernel): a known-size reference for
 * calibrating hardware byte counters.  Synchronous. */
pic_status pic_hip_copy_probe(pic_hip_ctx* ctx, size_t bytes);
/* Streaming-read bandwidth of this device: `reps` timed reads (HIP events,
 * median) of a `bytes` scratch buffer with 16-byte loads; *gbs in GB/s.  The
 * measured ceiling that HBM-bound kernels are compared with beside the
 * published peak.  Synchronous. */
pic_status pic_hip_read_probe(pic_hip_ctx* ctx, size_t bytes, int reps, double* gbs);

/* ---- tiling autotuner (build-defined; no reference counterpart) ----------
 * Times every LDS tiling candidate (tile shapes from 16x8 to 128x4 cells,
 * 4 or 8 particles per lane) for `scheme` on the context's current, sorted
 * particles (tiled kernel + merge, HIP events, median of `reps` deposits) and
 * keeps the fastest for that scheme.  With a cache_path, a JSON file keyed by
 * device architecture, scheme, grid and particles per cell (rounded to a power
 * of two) is consulted first and updated after a measurement; NULL disables
 * the cache.  Outputs (any may be NULL): the chosen tiling and its time per
 * deposit (0 when it came from the cache). */
pic_status pic_hip_autotune(pic_hip_ctx* ctx, int scheme, const char* cache_path, int reps, int* tile_x,
                            int* tile_y, int* particles_per_lane, double* ms_per_deposit);

/* ---- batched ensemble (build-defined; no reference counterpart) ----------
 * Many independent periodic PIC instances ("members") of uneven grid size,
 * particle count, scheme and seed, held in one set of device buffers:
 *   - particles of all members in shared FP32 SoA streams, member by member,
 *     each member's range padded to a multiple of 4096 slots;
 *   - one cell sort over all members (keys offset per member, so members
 *     stay contiguous);
 *   - the tiled deposit of every member in persistent kernels (one per
 *     scheme / particles-per-lane / coordinate-path group) that pull
 *     (member, tile) work items from a device counter, largest tiles first,
 *     then one merge over all members' nodes;
 *   - per-member FP64 mean removal and batched hipFFT plans (one per grid
 *     size) for the spectral Poisson solve, fused gather + push.
 * pic_hip_ensemble_run captures a step in hipGraphs and replays them.  Each
 * member's results equal those of a pic_hip_ctx with the same particles and
 * tiling (deposit: bit for bit). */
typedef struct {
    pic_domain domain;
    int32_t scheme;        /* PIC_SCHEME_*: the member's deposit scheme */
    int64_t num_particles;
    pic_init_params init;  /* used by pic_hip_ensemble_init */
} pic_ensemble_member;

typedef struct pic_hip_ensemble pic_hip_ensemble;

pic_status pic_hip_ensemble_create(const pic_ensemble_member* members, int32_t count, pic_hip_ensemble** out);
pic_status pic_hip_ensemble_destroy(pic_hip_ensemble* ens);
int32_t pic_hip_ensemble_count(pic_hip_ensemble* ens);
/* Waits for the ensemble's stream; reports device-side errors as pic_hip_sync. */
pic_status pic_hip_ensemble_sync(pic_hip_ensemble* ens);
/* Initialises every member's particles on the device from its init params. */
pic_status pic_hip_ensemble_init(pic_hip_ensemble* ens);
/* Copies one member's particles in / out (num_particles FP32 values per
 * stream, synchronous; NULL vx/vy upload zeros, NULL m uploads ones, NULL
 * outputs are skipped). */
pic_status pic_hip_ensemble_upload(pic_hip_ensemble* ens, int32_t member, const float* x, const float* y,
                                   const float* vx, const float* vy, const float* q, const float* m);
pic_status pic_hip_ensemble_download(pic_hip_ensemble* ens, int32_t member, float* x, float* y, float* vx, float* vy,
                                     float* q, float* m);
/* Sort of all members in one radix sort: CIC and Esirkepov members by cell,
 * NGP and TSC members by nearest node (the order of pic_hip_sort_nodes). */
pic_status pic_hip_ensemble_sort(pic_hip_ensemble* ens);
/* Tiled deposit of every member with its own scheme (needs a sort). */
pic_status pic_hip_ensemble_deposit(pic_hip_ensemble* ens, double dt);
pic_status pic_hip_ensemble_get_rho(pic_hip_ensemble* ens, int32_t member, float* rho);
/* Ex, Ey of every member from its rho (k = 0 removed per member in FP64). */
pic_status pic_hip_ensemble_poisson(pic_hip_ensemble* ens);
pic_status pic_hip_ensemble_get_fields(pic_hip_ensemble* ens, int32_t member, float* Ex, float* Ey);
/* Tiling of one member (member -1: all members). */
pic_status pic_hip_ensemble_set_tiling(pic_hip_ensemble* ens, int32_t member, int tile_x, int tile_y,
                                       int particles_per_lane);
pic_status pic_hip_ensemble_get_tiling(pic_hip_ensemble* ens, int32_t member, int* tile_x, int* tile_y,
                                       int* particles_per_lane);
/* Drift halo of every member's tile windows (as pic_hip_set_tile_halo). */
pic_status pic_hip_ensemble_set_tile_halo(pic_hip_ensemble* ens, int cells);
/* Autotunes the members by class (grid, scheme, particles per cell rounded
 * to a power of two): each class's candidates are timed with only that
 * class's work items in the persistent kernels.  The whole ensemble's deposit
 * is then timed with the tuned and with the previous tilings, and the
 * previous ones stay if they are faster.  Needs a sort.  cache_path as
 * pic_hip_autotune. */
pic_status pic_hip_ensemble_autotune(pic_hip_ensemble* ens, const char* cache_path, int reps);

/* Step loop of every member (PIC_ORDER_PUSH_FIRST: push, [sort], deposit,
 * Poisson, gather), no background (the solve removes each member's mean).
 * Records per member for steps >= record_from, laid out
 * [member][nsteps - record_from]: field energy, |mode-1 amplitude of Ex|,
 * kinetic energy, charge error (as pic_hip_run with record_particles).
 * Sort before the deposit of every sort_interval-th step (first included). */
typedef struct {
    double dt;
    int32_t sort_interval;
    int32_t use_graph;
} pic_ensemble_run_config;

pic_status pic_hip_ensemble_run(pic_hip_ensemble* ens, const pic_ensemble_run_config* cfg, int32_t nsteps,
                                int32_t record_from, double* field_energy, double* mode1_amplitude,
                                double* kinetic_energy, double* charge_error);
/* Mean device time (HIP events on the ensemble's stream, `reps` runs after
 * one warm-up) of: what = 0 the deposit (persistent kernels + merge),
 * 1 the sort, 2 one step replayed from its hipGraph without sort, 3 the same
 * with sort.  Particles are left as the runs leave them. */
pic_status pic_hip_ensemble_time(pic_hip_ensemble* ens, int what, int reps, double dt, double* mean_ms);

/* ---- learned density denoiser (build-defined; no reference counterpart) --
 * A periodic convolution network maps a low-particles-per-cell density to
 * the density a high-particles-per-cell deposit would give.  It works on the
 * relative fluctuation delta = rho / mean(rho) - 1: four 3x3 convolution
 * layers 1 -> 32 -> 32 -> 32 -> 1 with ReLU between them and per-layer
 * dilation (default 1, 3, 9, 27: an 81 x 81 receptive field in which every
 * offset is reached by exactly one tap path), 19,105
 * parameters, BF16 weights and activations with FP32 accumulation on the
 * matrix cores, FP32 master weights trained with Adam on mean-squared error.
 * Training pairs are deposits made by this library (pic_hip_denoiser_add_sample
 * normalises a context's current rho into the input or the target set).
 *
 * pic_hip_denoise replaces a context's rho by the network's estimate,
 * renormalised in FP64 so that sum(rho) * cellVolume equals the particles'
 * total charge to FP64 round-off, unless the gate rejects it.  The gate is
 * the Gauss-law residual of the learned field against the deposited charge,
 * ||div E_learned - rho_plain||^2 = ||rho_learned - rho_plain||^2, divided by
 * its shot-noise expectation kappa * sum(q^2) / cellVolume^2 (kappa = sum of
 * squared node weights per particle: 1 NGP, 4/9 CIC and Esirkepov, 0.3025
 * TSC); above `tolerance`, or for a non-finite result or zero mean charge,
 * the plain deposit is kept.  The grid must have Nx a multiple of 32. */
enum { PIC_DENOISER_PARAMS = 19105 };
typedef struct pic_hip_denoiser pic_hip_denoiser;

/* dilation: 4 values (NULL: 1, 3, 9, 27), each below min(Nx, Ny).  Weights
 * get He-normal initialisation from seed (the last layer scaled by 0.1). */
pic_status pic_hip_denoiser_create(int32_t Nx, int32_t Ny, int32_t max_batch, int32_t max_pairs,
                                   const int32_t* dilation, uint64_t seed, pic_hip_denoiser** out);
pic_status pic_hip_denoiser_destroy(pic_hip_denoiser* net);
/* FP32 master parameters: W1[32][9], b1[32], W2[9][32][32] (tap, out, in),
 * b2[32], W3[9][32][32], b3[32], W4[9][32], b4[1]; tap = 3 * dy + dx. */
pic_status pic_hip_denoiser_get_params(pic_hip_denoiser* net, float* params);
pic_status pic_hip_denoiser_set_params(pic_hip_denoiser* net, const float* params);
/* Standardisation: the network sees in_scale * delta and predicts
 * out_scale * delta.  Both are 1 until the first pic_hip_denoiser_train,
 * which sets them to 1 / RMS of the input and target sets unless
 * pic_hip_denoiser_set_scales fixed them before. */
pic_status pic_hip_denoiser_get_scales(pic_hip_denoiser* net, double* in_scale, double* out_scale);
pic_status pic_hip_denoiser_set_scales(pic_hip_denoiser* net, double in_scale, double out_scale);
/* The raw network on host images [batch][Ny][Nx] (no scales), synchronous. */
pic_status pic_hip_denoiser_forward_host(pic_hip_denoiser* net, const float* input, float* output, int32_t batch);
/* Mean-squared-error loss of the raw network and its gradient with respect
 * to the parameters (PIC_DENOISER_PARAMS values) for one host batch,
 * synchronous. */
pic_status pic_hip_denoiser_gradient_host(pic_hip_denoiser* net, const float* input, const float* target,
                                          int32_t batch, double* loss, float* grad);
/* Appends the context's current rho (as delta) to the input (is_target = 0)
 * or target (1) set; the k-th input pairs with the k-th target. */
pic_status pic_hip_denoiser_add_sample(pic_hip_denoiser* net, pic_hip_ctx* ctx, int32_t is_target);
int32_t pic_hip_denoiser_pairs(pic_hip_denoiser* net);
pic_status pic_hip_denoiser_clear_samples(pic_hip_denoiser* net);
/* Adam (beta 0.9 / 0.999, eps 1e-8) on batches of random pairs, each rolled
 * by a random periodic shift; learning rate cosine-decayed from lr to
 * lr / 20.  loss_history (NULL to skip) receives one loss per iteration. */
pic_status pic_hip_denoiser_train(pic_hip_denoiser* net, int32_t iterations, int32_t batch, double lr,
                                  uint64_t seed, double* loss_history);
/* Learned density for the context (see above).  fallback / residual_ratio
 * (NULL to skip; non-NULL makes the call synchronous) report the gate. */
pic_status pic_hip_denoise(pic_hip_ctx* ctx, pic_hip_denoiser* net, int scheme, double tolerance, int32_t* fallback,
                           double* residual_ratio);
/* pic_hip_run with the learned density: after every deposit the context's
 * rho goes through pic_hip_denoise(ctx, net, cfg->scheme, tolerance) (gate
 * decided on the device, no host round trip), before the background and the
 * Poisson solve; graph capture covers it.  fallbacks (NULL to skip) receives
 * the number of steps whose gate kept the plain deposit. */
pic_status pic_hip_run_learned(pic_hip_ctx* ctx, pic_hip_denoiser* net, const pic_step_config* cfg, double tolerance,
                               int32_t nsteps, int32_t record_from, double* field_energy, double* mode1_amplitude,
                               double* kinetic_energy, double* charge_error, int32_t* fallbacks);
/* FP64 density of the last pic_hip_denoise (renormalised, or the plain
 * deposit when the gate fell back), Nx * Ny values. */
pic_status pic_hip_denoiser_get_output_f64(pic_hip_denoiser* net, double* rho);
/* Deposit (global atomics) of a hash-selected Bernoulli(1/stride) subset of
 * the context's particles with charges multiplied by stride: a density with
 * stride times fewer particles per cell of the same state, for training
 * pairs.  Any particle order. */
pic_status pic_hip_deposit_subsample(pic_hip_ctx* ctx, int scheme, uint32_t stride, uint64_t salt);

#ifdef __cplusplus
}
#endif

#endif /* PIC_HIP_H */
