// Charge-conserving current deposit (Esirkepov, Comput. Phys. Commun. 135
// (2001) 144, eqs. 21-23, in 2D with first-order (CIC) shape factors) and the
// discrete continuity residual.  The reference has no current deposit: its
// "Esirkepov" scheme deposits a trajectory-averaged density (see
// pic_common.hpp); this kernel is the build's charge-conserving counterpart.
//
// Layout: Jx(i, j) lives on the x face (i + 1/2, j), Jy(i, j) on the y face
// (i, j + 1/2); both stored row-major like rho.  With backward differences
// (Dx Jx)(i, j) = (Jx(i, j) - Jx(i-1, j)) / dx the deposit satisfies
//   (rho_new - rho_old) / dt + Dx Jx + Dy Jy = 0
// exactly in real arithmetic, where rho_old / rho_new are the CIC densities
// at x and x + v dt.  Moves must stay under one cell per step.

#include "deposit_device.hpp"
#include "pic_common.hpp"
#include "pic_kernels.hpp"

namespace pichip {

namespace {

__device__ __forceinline__ void comp_add(float* __restrict__ g, float* __restrict__ c, size_t i, float v) {
    if (v != 0.0f) {
        const float old = unsafeAtomicAdd(g + i, v);
        unsafeAtomicAdd(c + i, two_sum_err(old, v));
    }
}

// One axis: old and new CIC shapes on a 4-node frame starting at c0 - 1.
// Returns false if the move reaches a cell boundary beyond the adjacent one.
template <int EX>
__device__ __forceinline__ bool frame_1d(float x, float v, float dt, float inv_d, double dd, int exact, int& base,
                                         float (&S0)[4], float (&S1)[4]) {
    int c0;
    float f0;
    cell_coord<EX>(x, inv_d, dd, exact, c0, f0);
    const float dc = cells<EX>(v * dt, inv_d, dd, exact);
    if (!(fabsf(dc) < 1.0f)) return false;
    const float f1 = f0 + dc;  // new position relative to the old cell's lower face, in (-1, 2)
    const float fl1 = floorf(f1);
    const int i1 = (int)fl1;   // -1, 0 or 1
    const float g1 = f1 - fl1;
    base = c0 - 1;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        S0[a] = 0.0f;
        S1[a] = 0.0f;
    }
    S0[1] = 1.0f - f0;
    S0[2] = f0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (a == i1 + 1) {
            S1[a] = 1.0f - g1;
            S1[a + 1] = g1;
        }
    }
    return true;
}

// One particle's current and densities added into the global grids with
// compensated FP32 atomics (any particle order).
template <int EX>
__device__ __forceinline__ void current_particle_global(float x, float y, float vx, float vy, float q,
                                                        const DevDomain& d, float dt, const CurrentGrids& g,
                                                        int* __restrict__ status) {
    int bx, by;
    float S0x[4], S1x[4], S0y[4], S1y[4];
    const bool okx = frame_1d<EX>(x, vx, dt, d.inv_dx, d.dx, d.exact, bx, S0x, S1x);
    const bool oky = frame_1d<EX>(y, vy, dt, d.inv_dy, d.dy, d.exact, by, S0y, S1y);
    if (!okx || !oky) {
        atomicOr(status, 2);
        return;
    }
    const float qv = q * d.inv_vol;
    const float cx = -qv * (float)(d.dx / (double)dt);  // -q dx / (vol dt)
    const float cy = -qv * (float)(d.dy / (double)dt);
    int gx[4], gy[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) gx[a] = wrap_index(bx + a, d.Nx);
#pragma unroll
    for (int b = 0; b < 4; ++b) gy[b] = wrap_index(by + b, d.Ny);
    float dSx[4], dSy[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) dSx[a] = S1x[a] - S0x[a];
#pragma unroll
    for (int b = 0; b < 4; ++b) dSy[b] = S1y[b] - S0y[b];
    // Jx: running sum along x of -q dx/dt * W_x, W_x = dSx (S0y + dSy / 2).
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const float wy = S0y[b] + 0.5f * dSy[b];
        float acc = 0.0f;
#pragma unroll
        for (int a = 0; a < 3; ++a) {  // the 4th partial sum is zero (sum of dSx vanishes)
            acc = fmaf(cx * dSx[a], wy, acc);
            comp_add(g.Jx, g.Jx_c, (size_t)gy[b] * d.Nx + gx[a], acc);
        }
    }
    // Jy: running sum along y of -q dy/dt * W_y, W_y = dSy (S0x + dSx / 2).
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const float wx = S0x[a] + 0.5f * dSx[a];
        float acc = 0.0f;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            acc = fmaf(cy * dSy[b], wx, acc);
            comp_add(g.Jy, g.Jy_c, (size_t)gy[b] * d.Nx + gx[a], acc);
        }
    }
    // CIC densities at the old and new positions.
#pragma unroll
    for (int b = 0; b < 4; ++b) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const size_t node = (size_t)gy[b] * d.Nx + gx[a];
            comp_add(g.rho_old, g.rho_old_c, node, qv * S0x[a] * S0y[b]);
            comp_add(g.rho_new, g.rho_new_c, node, qv * S1x[a] * S1y[b]);
        }
    }
}

// The tile kernel's rare path (a particle outside its sorted cell or moving
// a cell or more), kept out of line so that the tile loop stays small.
template <int EX>
__device__ __noinline__ void current_particle_global_call(float x, float y, float vx, float vy, float q,
                                                          const DevDomain& d, float dt, const CurrentGrids& g,
                                                          int* __restrict__ status) {
    current_particle_global<EX>(x, y, vx, vy, q, d, dt, g, status);
}

template <int EX>
__global__ __launch_bounds__(256) void esirkepov_current_kernel(const float* __restrict__ px,
                                                                const float* __restrict__ py,
                                                                const float* __restrict__ pvx,
                                                                const float* __restrict__ pvy,
                                                                const float* __restrict__ pq, int64_t n, DevDomain d,
                                                                float dt, CurrentGrids g, int* __restrict__ status) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    current_particle_global<EX>(px[p], py[p], pvx[p], pvy[p], pq[p], d, dt, g, status);
}

// Cell-sorted particles, one wavefront per tile of T x R cells (R <= 2, no
// drift halo).  The tile's work is cut into items of at most 8 consecutive
// particles of one cell (a cell of n particles gives ceil(n / 8) items), one
// item per lane and step.  A step's 64 items cover at most two contiguous
// particle ranges (the tile's rows), which the wave copies into an LDS stage
// with 16-byte loads, consecutive lanes on consecutive quads; the loads of
// the next step are issued before this step's arithmetic, so that their
// latency hides behind it.  All particles of a cell share the 4 x 4 node
// frame that starts one node below the cell: a lane sums its item in 28 FP32
// registers (CurSums), one FMA per register and particle, a segmented wave
// scan adds the items of each cell, and the cell's last lane adds the frame
// into the wave's four FP64 LDS windows ((T + 3) x (R + 3) nodes each, the
// deposit's layout), written out in FP64 for the merge.  A particle outside
// its sorted cell (moved since the sort) or with a move of a cell or more
// takes the global path.
constexpr int kCurItem = 8;                                // particles per item
constexpr int kCurStep = 64 * kCurItem;                    // particles staged per step
constexpr int kCurStage = kCurStep + kCurStep / kCurItem;  // words per staged stream (a pad word per item)
constexpr int kCurQuads = (kCurStep / 4 + 4 + 63) / 64;    // 16-byte quads per lane and stream per step
constexpr int kCurCells = 128;                             // cells per tile at most (T x R = 64 x 2)
constexpr int kCurWaves = 2;                               // waves (tiles) per workgroup

// Stage word of staged particle s: a pad word after every kCurItem, so that
// lanes reading items kCurItem apart hit different LDS banks.
__device__ __forceinline__ int cur_stage_word(int s) { return s + (int)((unsigned)s / kCurItem); }

// Per-wave LDS of the tile kernel, in doubles: the four FP64 windows, the
// cell tables (3 kCurCells + 1 ints) and the five staged streams.
__host__ __device__ constexpr int cur_wave_doubles(int wsize) {
    return 4 * wsize + (3 * kCurCells + 2) / 2 + (5 * kCurStage + 1) / 2;
}

// Sums over a cell's particles, per particle (q' = q / cell volume, S0 / S1
// the CIC shapes before / after the move, frame slots a, b = 0..3):
//   gx[a][b] += q' (S1x[a] - S0x[a]) (S0y[b] + S1y[b])   (a < 3)
//   gy[a][b] += q' (S0x[a] + S1x[a]) (S1y[b] - S0y[b])   (b < 3)
//   ro[a][b] += q' S0x[a] S0y[b]                         (a, b in {1, 2})
// Esirkepov's Jx is -(dx / dt) / 2 times the running sum of gx over a (Jy
// likewise over b); the fourth slot of each is minus the sum of the first
// three (sum_a (S1x - S0x) = 0).  Since q' S1x S1y - q' S0x S0y = (gx + gy) / 2
// slot by slot (with the fourth slots), rho_new = rho_old + (gx + gy) / 2:
// the frame satisfies the discrete continuity equation by construction.
// Packed for the wave scan: gx at 0 (a * 4 + b), gy at 12 (a * 3 + b), ro at
// 24 ((b - 1) * 2 + a - 1).
constexpr int kCurV = 28;

// Adds a cell's frame (cell (lx, ly) of the tile) into the four windows:
// frame node (a, b) is window node (lx + a, ly + b).  hx, hy: -(dx / dt) / 2,
// -(dy / dt) / 2.
__device__ __forceinline__ void cur_run_emit(double* __restrict__ win, int wsize, int WX, int lx, int ly,
                                             const float (&r)[kCurV], float hx, float hy) {
    const int base = ly * WX + lx;
    float gx[16], gy[16];  // all four slots: gx[a * 4 + b], gy[a * 4 + b]
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        gx[12 + b] = -(r[b] + r[4 + b] + r[8 + b]);
#pragma unroll
        for (int a = 0; a < 3; ++a) gx[a * 4 + b] = r[a * 4 + b];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gy[a * 4 + 3] = -(r[12 + a * 3] + r[12 + a * 3 + 1] + r[12 + a * 3 + 2]);
#pragma unroll
        for (int b = 0; b < 3; ++b) gy[a * 4 + b] = r[12 + a * 3 + b];
    }
    float jx[12], jy[12];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        float acc = 0.0f;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            acc += gx[a * 4 + b];
            jx[a * 4 + b] = hx * acc;
        }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        float acc = 0.0f;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            acc += gy[a * 4 + b];
            jy[a * 3 + b] = hy * acc;
        }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int w = base + b * WX + a;
            const bool inner = a >= 1 && a <= 2 && b >= 1 && b <= 2;
            const float ro = inner ? r[24 + (b - 1) * 2 + a - 1] : 0.0f;
            if (a < 3) atomicAdd(win + w, (double)jx[a * 4 + b]);
            if (b < 3) atomicAdd(win + wsize + w, (double)jy[a * 3 + b]);
            if (inner) atomicAdd(win + 2 * wsize + w, (double)ro);
            atomicAdd(win + 3 * wsize + w, (double)ro + 0.5 * ((double)gx[a * 4 + b] + (double)gy[a * 4 + b]));
        }
}

// One axis of a particle in the tile kernel: cell c, in-cell coordinate f0
// in [0, 1) and f1 = f0 + the move in cells; false for a move of a cell or
// more (the global path reports it).
// dt_inv = dt / cell size: with a power-of-two cell size (EX = 1) v * dt_inv
// equals (v * dt) / cell size exactly (scaling by a power of two commutes with
// the rounding of v * dt).
template <int EX>
__device__ __forceinline__ bool cur_axis(float x, float v, float dt, float dt_inv, float inv_d, double dd, int exact,
                                         int& c, float& f0, float& f1) {
    cell_coord<EX>(x, inv_d, dd, exact, c, f0);
    const float dc = EX == 1 ? v * dt_inv : cells<EX>(v * dt, inv_d, dd, exact);
    f1 = f0 + dc;
    return fabsf(dc) < 1.0f;
}

// A step of the tile loop: this lane's item (local cell lc, particles [p0,
// p1), empty past the tile's items) and the two particle ranges the step's
// items cover, [s0b, s0e) in the tile's first row and [s1b, s1e) in its second.
struct CurPlan {
    int lc, p0, p1;
    int s0b, s0e, s1b, s1e;
    int nq0, nq1;  // 16-byte quads covering each range
};

__device__ __forceinline__ CurPlan cur_plan(int base, int total, int ncell, int tw, const int* __restrict__ icum,
                                            const int* __restrict__ cbeg, const int* __restrict__ cend) {
    CurPlan pl;
    const int lane = threadIdx.x & 63;
    const int item = base + lane;
    const bool live = item < total;
    // cell of this item: the last lc with icum[lc] <= item
    int lo = 0, hi = ncell - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (icum[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    pl.lc = lo;
    pl.p0 = cbeg[lo] + (item - icum[lo]) * kCurItem;
    pl.p1 = live ? min(pl.p0 + kCurItem, cend[lo]) : pl.p0;
    // The step's items in row 0 (lanes [0, L0)) and row 1 (lanes [L0, NL))
    // each cover one contiguous range: a cell's items cover its particles in
    // order, and the next cell of a row starts where the last one ends.
    const int L0 = __popcll(__ballot(live && lo < tw));
    const int NL = __popcll(__ballot(live));
    pl.s0b = L0 > 0 ? __shfl(pl.p0, 0, 64) : 0;
    pl.s0e = L0 > 0 ? __shfl(pl.p1, L0 - 1, 64) : 0;
    pl.s1b = NL > L0 ? __shfl(pl.p0, L0, 64) : 0;
    pl.s1e = NL > L0 ? __shfl(pl.p1, NL - 1, 64) : 0;
    pl.nq0 = pl.s0e > pl.s0b ? ((pl.s0e - 1) >> 2) - (pl.s0b >> 2) + 1 : 0;
    pl.nq1 = pl.s1e > pl.s1b ? ((pl.s1e - 1) >> 2) - (pl.s1b >> 2) + 1 : 0;
    return pl;
}

// The quads of a step's ranges, lane-strided (quad qi = lane + 64 j); reads
// up to 3 particles past a range end stay inside the padded streams.
struct CurQuads {
    float4 v[5][kCurQuads];
};

__device__ __forceinline__ void cur_load(const CurPlan& pl, const float* __restrict__ px, const float* __restrict__ py,
                                         const float* __restrict__ pvx, const float* __restrict__ pvy,
                                         const float* __restrict__ pq, CurQuads& qd) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < kCurQuads; ++j) {
        const int qi = lane + 64 * j;
        if (qi < pl.nq0 + pl.nq1) {
            const bool r1 = qi >= pl.nq0;
            const int qq = ((r1 ? pl.s1b : pl.s0b) >> 2) + (r1 ? qi - pl.nq0 : qi);
            qd.v[0][j] = reinterpret_cast<const float4*>(px)[qq];
            qd.v[1][j] = reinterpret_cast<const float4*>(py)[qq];
            qd.v[2][j] = reinterpret_cast<const float4*>(pvx)[qq];
            qd.v[3][j] = reinterpret_cast<const float4*>(pvy)[qq];
            qd.v[4][j] = reinterpret_cast<const float4*>(pq)[qq];
        }
    }
}

// Writes the in-range particles of the loaded quads to the stage: range 0 at
// staged index 0, range 1 after it.
__device__ __forceinline__ void cur_stage(const CurPlan& pl, const CurQuads& qd, float* __restrict__ stage) {
    const int lane = threadIdx.x & 63;
    const int len0 = pl.s0e - pl.s0b;
#pragma unroll
    for (int j = 0; j < kCurQuads; ++j) {
        const int qi = lane + 64 * j;
        if (qi < pl.nq0 + pl.nq1) {
            const bool r1 = qi >= pl.nq0;
            const int sb = r1 ? pl.s1b : pl.s0b, se = r1 ? pl.s1e : pl.s0e, off = r1 ? len0 : 0;
            const int qq = (sb >> 2) + (r1 ? qi - pl.nq0 : qi);
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const float e4[4] = {qd.v[k][j].x, qd.v[k][j].y, qd.v[k][j].z, qd.v[k][j].w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int pp = 4 * qq + e;
                    if (pp >= sb && pp < se) stage[k * kCurStage + cur_stage_word(pp - sb + off)] = e4[e];
                }
            }
        }
    }
}

template <int EX>
__global__ __launch_bounds__(64 * kCurWaves) void current_tile_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pvx,
    const float* __restrict__ pvy, const float* __restrict__ pq, const int* __restrict__ cell_off, DevDomain d,
    TileGeom tg, float dt, double* __restrict__ partial, CurrentGrids g, int* __restrict__ status) {
    extern __shared__ double cur_lds[];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: scalar tile arithmetic
    const int tile = blockIdx.x * kCurWaves + wv;
    if (tile >= tg.ntiles) return;  // wave-uniform
    double* win = cur_lds + (size_t)wv * cur_wave_doubles(tg.wsize);  // Jx, Jy, rho_old, rho_new
    int* icum = reinterpret_cast<int*>(win + 4 * tg.wsize);  // [cells + 1] items before local cell c
    int* cbeg = icum + kCurCells + 1;                          // [cells] first particle of local cell c
    int* cend = cbeg + kCurCells;                              // [cells] one past its last
    float* stage = reinterpret_cast<float*>(win + 4 * tg.wsize + (3 * kCurCells + 2) / 2);  // [5][kCurStage]
    const int tx = tile % tg.ntx, ty = tile / tg.ntx;
    const int x0 = tx * tg.T, y0 = ty * tg.R;
    const int tw = min(tg.T, d.Nx - x0), th = min(tg.R, d.Ny - y0);
    const int ncell = tw * th;  // local cell lc = ly * tw + lx
    for (int i = lane; i < 4 * tg.wsize; i += 64) win[i] = 0.0;
    // Items per cell and their prefix (two cells per lane).
    int cnt[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int lc = 2 * lane + k;
        cnt[k] = 0;
        if (lc < ncell) {
            const int c = (y0 + lc / tw) * d.Nx + x0 + lc % tw;
            const int b = cell_off[c], e = cell_off[c + 1];
            cbeg[lc] = b;
            cend[lc] = e;
            cnt[k] = (e - b + kCurItem - 1) / kCurItem;
        }
    }
    int incl = cnt[0] + cnt[1];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(incl, o, 64);
        if (lane >= o) incl += u;
    }
    if (2 * lane < ncell) icum[2 * lane] = incl - cnt[0] - cnt[1];
    if (2 * lane + 1 < ncell) icum[2 * lane + 1] = incl - cnt[1];
    if (lane == 63) icum[ncell] = incl;
    wave_sync();
    const int total = icum[ncell];
    const float hx = -0.5f * (float)(d.dx / (double)dt), hy = -0.5f * (float)(d.dy / (double)dt);
    const float dtx = dt * d.inv_dx, dty = dt * d.inv_dy;
    CurQuads qd;
    CurPlan cur;
    if (total > 0) {
        cur = cur_plan(0, total, ncell, tw, icum, cbeg, cend);
        cur_load(cur, px, py, pvx, pvy, pq, qd);
        cur_stage(cur, qd, stage);
        wave_sync();
    }
    for (int base = 0; base < total; base += 64) {
        const bool more = base + 64 < total;  // wave-uniform
        CurPlan nxt;
        if (more) {  // the next step's loads, in flight during this step's arithmetic
            nxt = cur_plan(base + 64, total, ncell, tw, icum, cbeg, cend);
            cur_load(nxt, px, py, pvx, pvy, pq, qd);
        }
        const bool live = base + lane < total;
        const int lc = cur.lc;
        const int ly = lc >= tw ? 1 : 0;
        const int lx = lc - ly * tw;
        const int sl = cur.p0 - (ly ? cur.s1b : cur.s0b) + (ly ? cur.s0e - cur.s0b : 0);  // staged index of p0
        float r[kCurV];
#pragma unroll
        for (int k = 0; k < kCurV; ++k) r[k] = 0.0f;
        // the item's particles from the stage, all read before the arithmetic
        // (slots past the item read its first particle)
        float xs[kCurItem], ys[kCurItem], vxs[kCurItem], vys[kCurItem], qs[kCurItem];
#pragma unroll
        for (int v = 0; v < kCurItem; ++v) {
            const int w = cur_stage_word(cur.p0 + v < cur.p1 ? sl + v : (live ? sl : 0));
            xs[v] = stage[w];
            ys[v] = stage[kCurStage + w];
            vxs[v] = stage[2 * kCurStage + w];
            vys[v] = stage[3 * kCurStage + w];
            qs[v] = stage[4 * kCurStage + w];
        }
#pragma unroll
        for (int v = 0; v < kCurItem; ++v) {
            if (cur.p0 + v >= cur.p1) break;
            const float xv = xs[v], yv = ys[v], vxv = vxs[v], vyv = vys[v], qq = qs[v];
            int cx, cy;
            float f0x, f1x, f0y, f1y;
            const bool okx = cur_axis<EX>(xv, vxv, dt, dtx, d.inv_dx, d.dx, d.exact, cx, f0x, f1x);
            const bool oky = cur_axis<EX>(yv, vyv, dt, dty, d.inv_dy, d.dy, d.exact, cy, f0y, f1y);
            if (!(okx && oky && cx == x0 + lx && cy == y0 + ly)) {
                current_particle_global_call<EX>(xv, yv, vxv, vyv, qq, d, dt, g, status);
                continue;
            }
            const float qv = qq * d.inv_vol;
            // S0 = (0, 1 - f0, f0, 0) and S1 the CIC hat at f1 on the frame
            // nodes -1 .. 2 of the cell; A0 = q' S0x, A1 = q' S1x
            float S1x[4], S1y[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                S1x[a] = fmaxf(0.0f, 1.0f - fabsf(f1x + (float)(1 - a)));
                S1y[a] = fmaxf(0.0f, 1.0f - fabsf(f1y + (float)(1 - a)));
            }
            const float S0y1 = 1.0f - f0y, S0y2 = f0y;
            float A1[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) A1[a] = qv * S1x[a];
            const float A01 = qv * (1.0f - f0x), A02 = qv * f0x;
            const float dA[3] = {A1[0], A1[1] - A01, A1[2] - A02};
            const float sA[4] = {A1[0], A1[1] + A01, A1[2] + A02, A1[3]};
            const float dB[3] = {S1y[0], S1y[1] - S0y1, S1y[2] - S0y2};
            const float sB[4] = {S1y[0], S1y[1] + S0y1, S1y[2] + S0y2, S1y[3]};
            const float S0y[4] = {0.0f, S0y1, S0y2, 0.0f};
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) r[a * 4 + b] = fmaf(dA[a], sB[b], r[a * 4 + b]);
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) r[12 + a * 3 + b] = fmaf(sA[a], dB[b], r[12 + a * 3 + b]);
            r[24] = fmaf(A01, S0y[1], r[24]);
            r[25] = fmaf(A02, S0y[1], r[25]);
            r[26] = fmaf(A01, S0y[2], r[26]);
            r[27] = fmaf(A02, S0y[2], r[27]);
        }
        // Items of one cell sit in consecutive lanes: a segmented scan within
        // each row of 16 lanes adds them, and the last lane of each cell in a
        // row adds the sum to the windows (a cell's items span about half a
        // row at 64 particles per cell; four scan steps instead of six).
        const int key = live ? lc : -1;
        const int key_prev = __shfl_up(key, 1, 64);
        const int key_next = __shfl_down(key, 1, 64);
        const float A = (live && lane > 0 && key_prev == key) ? 1.0f : 0.0f;
        wave_segmented_scan<kCurV, 4>(r, A);
        if (live && ((lane & 15) == 15 || key_next != key)) cur_run_emit(win, tg.wsize, tg.WX, lx, ly, r, hx, hy);
        if (more) {
            wave_sync();  // this step's stage reads are done
            cur_stage(nxt, qd, stage);
            wave_sync();
            cur = nxt;
        }
    }
    wave_sync();
    const size_t plane = (size_t)tg.ntiles * tg.wsize;
    for (int i = lane; i < tg.wsize; i += 64) {
#pragma unroll
        for (int k = 0; k < 4; ++k) partial[k * plane + (size_t)tile * tg.wsize + i] = win[k * tg.wsize + i];
    }
}

// Final grids: the four planes of tile windows added in FP64 per node (fixed
// order), plus the compensated global sums of the particles that took the
// global path; the compensation grids are cleared.
template <int B>
__global__ __launch_bounds__(256) void current_merge_kernel(const double* __restrict__ partial, DevDomain d,
                                                            TileGeom tg, CurrentGrids g) {
    const int gx = blockIdx.x * blockDim.x + threadIdx.x;
    const int gy = blockIdx.y;
    if (gx >= d.Nx) return;
    const size_t node = (size_t)gy * d.Nx + gx;
    const size_t plane = (size_t)tg.ntiles * tg.wsize;
    float* grid[4] = {g.Jx, g.Jy, g.rho_old, g.rho_new};
    float* comp[4] = {g.Jx_c, g.Jy_c, g.rho_old_c, g.rho_new_c};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double s = merge_windows<B, double>(gx, gy, partial + k * plane, d, tg);
        grid[k][node] = (float)(s + ((double)grid[k][node] + (double)comp[k][node]));
        comp[k][node] = 0.0f;
    }
}

// g += c; c = 0 for the four compensated grids.
__global__ __launch_bounds__(256) void fold_current_kernel(CurrentGrids g, size_t G) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G) return;
    g.Jx[i] += g.Jx_c[i];
    g.Jy[i] += g.Jy_c[i];
    g.rho_old[i] += g.rho_old_c[i];
    g.rho_new[i] += g.rho_new_c[i];
    g.Jx_c[i] = 0.0f;
    g.Jy_c[i] = 0.0f;
    g.rho_old_c[i] = 0.0f;
    g.rho_new_c[i] = 0.0f;
}

__global__ __launch_bounds__(256) void zero4_kernel(CurrentGrids g, size_t G) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G) return;
    g.Jx[i] = 0.0f;
    g.Jy[i] = 0.0f;
    g.rho_old[i] = 0.0f;
    g.rho_new[i] = 0.0f;
}

// Block maxima of |(rho_new - rho_old)/dt + Dx Jx + Dy Jy| and |(rho_new - rho_old)/dt| in FP64.
__global__ __launch_bounds__(256) void continuity_partial_kernel(CurrentGrids g, DevDomain d, double dt,
                                                                 double* __restrict__ partial) {
    double r = 0.0, m = 0.0;
    const int64_t G = (int64_t)d.Nx * d.Ny;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < G; i += (int64_t)gridDim.x * blockDim.x) {
        const int ix = (int)(i % d.Nx), iy = (int)(i / d.Nx);
        const int ixm = ix == 0 ? d.Nx - 1 : ix - 1;
        const int iym = iy == 0 ? d.Ny - 1 : iy - 1;
        const double drho = ((double)g.rho_new[i] - (double)g.rho_old[i]) / dt;
        const double div = ((double)g.Jx[i] - (double)g.Jx[(size_t)iy * d.Nx + ixm]) / d.dx +
                           ((double)g.Jy[i] - (double)g.Jy[(size_t)iym * d.Nx + ix]) / d.dy;
        r = fmax(r, fabs(drho + div));
        m = fmax(m, fabs(drho));
    }
    __shared__ double sr[256], sm[256];
    sr[threadIdx.x] = r;
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + o]);
            sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sr[0];
        partial[2 * blockIdx.x + 1] = sm[0];
    }
}

__global__ void continuity_final_kernel(const double* __restrict__ partial, int nb, double* __restrict__ out) {
    double r = 0.0, m = 0.0;
    for (int b = 0; b < nb; ++b) {
        r = fmax(r, partial[2 * b]);
        m = fmax(m, partial[2 * b + 1]);
    }
    out[0] = r;
    out[1] = m;
}

}  // namespace

void launch_esirkepov_current(const float* x, const float* y, const float* vx, const float* vy, const float* q,
                              int64_t n, const DevDomain& d, float dt, const CurrentGrids& g, int* status,
                              hipStream_t s) {
    const size_t G = (size_t)d.Nx * d.Ny;
    const unsigned gb = (unsigned)((G + 255) / 256);
    hipLaunchKernelGGL(zero4_kernel, dim3(gb), dim3(256), 0, s, g, G);
    if (n > 0) {
        const unsigned pb = (unsigned)((n + 255) / 256);
        if (d.exact)
            hipLaunchKernelGGL(esirkepov_current_kernel<1>, dim3(pb), dim3(256), 0, s, x, y, vx, vy, q, n, d, dt, g,
                               status);
        else
            hipLaunchKernelGGL(esirkepov_current_kernel<0>, dim3(pb), dim3(256), 0, s, x, y, vx, vy, q, n, d, dt, g,
                               status);
    }
    hipLaunchKernelGGL(fold_current_kernel, dim3(gb), dim3(256), 0, s, g, G);
}

size_t current_partial_doubles(const TileGeom& tg) { return (size_t)4 * tg.ntiles * tg.wsize; }

size_t current_tile_lds(const TileGeom& tg) {
    return (size_t)kCurWaves * cur_wave_doubles(tg.wsize) * sizeof(double);
}

void launch_esirkepov_current_tiled(const float* x, const float* y, const float* vx, const float* vy, const float* q,
                                    const int* cell_off, const DevDomain& d, const TileGeom& tg, float dt,
                                    double* partial, const CurrentGrids& g, int* status, hipStream_t s) {
    const size_t G = (size_t)d.Nx * d.Ny;
    const unsigned gb = (unsigned)((G + 255) / 256);
    hipLaunchKernelGGL(zero4_kernel, dim3(gb), dim3(256), 0, s, g, G);
    const unsigned blocks = (unsigned)((tg.ntiles + kCurWaves - 1) / kCurWaves);
    const size_t lds = current_tile_lds(tg);
    if (d.exact)
        hipLaunchKernelGGL(current_tile_kernel<1>, dim3(blocks), dim3(64 * kCurWaves), lds, s, x, y, vx, vy, q, cell_off, d, tg,
                           dt, partial, g, status);
    else
        hipLaunchKernelGGL(current_tile_kernel<0>, dim3(blocks), dim3(64 * kCurWaves), lds, s, x, y, vx, vy, q, cell_off, d, tg,
                           dt, partial, g, status);
    const dim3 mgrid((unsigned)((d.Nx + 255) / 256), (unsigned)d.Ny);
    if (tg.mbatch <= 4)
        hipLaunchKernelGGL(current_merge_kernel<4>, mgrid, dim3(256), 0, s, partial, d, tg, g);
    else if (tg.mbatch <= 8)
        hipLaunchKernelGGL(current_merge_kernel<8>, mgrid, dim3(256), 0, s, partial, d, tg, g);
    else if (tg.mbatch <= 16)
        hipLaunchKernelGGL(current_merge_kernel<16>, mgrid, dim3(256), 0, s, partial, d, tg, g);
    else
        hipLaunchKernelGGL(current_merge_kernel<32>, mgrid, dim3(256), 0, s, partial, d, tg, g);
}

void launch_continuity_residual(const CurrentGrids& g, const DevDomain& d, double dt, double* partial, int nblocks,
                                double* out, hipStream_t s) {
    hipLaunchKernelGGL(continuity_partial_kernel, dim3(nblocks), dim3(256), 0, s, g, d, dt, partial);
    hipLaunchKernelGGL(continuity_final_kernel, dim3(1), dim3(1), 0, s, partial, nblocks, out);
}

}  // namespace pichip
