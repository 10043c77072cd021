// Device-side building blocks of the tiled charge deposit: DPP wave scans,
// compensated LDS and global adds, the per-wave tile deposit and the
// per-node window merge.  Shared by the single-grid kernels (deposit.hip) and
// the batched ensemble kernels (ensemble.hip).
#pragma once

#include "pic_common.hpp"
#include "pic_kernels.hpp"

namespace pichip {
namespace {

// LDS visibility between lanes of one wavefront (in-order LDS queue + no
// compiler reordering across this point).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DPP controls (GFX9 encoding).
constexpr int kRowShr = 0x110;
constexpr int kRowBcast15 = 0x142;
constexpr int kRowBcast31 = 0x143;
constexpr int kWaveShr1 = 0x138;
constexpr int kWaveShl1 = 0x130;

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f(float old, float v) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROW_MASK, 0xf, false);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_i(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROW_MASK, 0xf, false);
}
// Shifted value with zero where the source lane is outside the row or wave.
// With every row enabled this is the bound_ctrl form, which needs no zeroed
// destination first and folds into the arithmetic that consumes it (one
// v_fmac_f32_dpp instead of a zeroing move, a DPP move and the FMA).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_z(float v) {
    if constexpr (ROW_MASK == 0xf) return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xf, 0xf, true));
    else return __builtin_amdgcn_update_dpp(0.0f, v, CTRL, ROW_MASK, 0xf, false);
}

// One Hillis-Steele step of the linear recurrence B_i = b_i + a_i * B_{i-1}.
template <int CTRL, int ROW_MASK, int V>
__device__ __forceinline__ void lin_scan_step(float (&B)[V], float& A) {
    const float As = dpp_f<CTRL, ROW_MASK>(1.0f, A);
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const float Bs = dpp_z<CTRL, ROW_MASK>(B[k]);
        B[k] = fmaf(Bs, A, B[k]);
    }
    A *= As;
}

// Inclusive wave-wide scan of B_i = b_i + a_i * B_{i-1} with a_i in {0, 1}:
// a segmented sum where a_i = 0 starts a new segment at lane i.  With
// STEPS < 6 the scan stays within groups of 2^STEPS lanes (16 at 4 steps, 32
// at 5): a group's first lane starts a segment whatever its a_i.
template <int V, int STEPS = 6>
__device__ __forceinline__ void wave_segmented_scan(float (&B)[V], float A) {
    lin_scan_step<kRowShr + 1, 0xf>(B, A);
    if constexpr (STEPS >= 2) lin_scan_step<kRowShr + 2, 0xf>(B, A);
    if constexpr (STEPS >= 3) lin_scan_step<kRowShr + 4, 0xf>(B, A);
    if constexpr (STEPS >= 4) lin_scan_step<kRowShr + 8, 0xf>(B, A);
    if constexpr (STEPS >= 5) lin_scan_step<kRowBcast15, 0xa>(B, A);
    if constexpr (STEPS >= 6) lin_scan_step<kRowBcast31, 0xc>(B, A);
}

// Adds a cell's V stencil values into the wave's FP64 LDS window (one
// non-returning LDS atomic per value; FP64 keeps the sum of thousands of
// FP32 contributions exact to FP32 after the final rounding).
template <int SX, int V>
__device__ __forceinline__ void emit_cell(double* __restrict__ win, int wbase, int WX, const float (&z)[V]) {
#pragma unroll
    for (int k = 0; k < V; ++k) atomicAdd(win + wbase + (k / SX) * WX + (k % SX), (double)z[k]);
}

// General reference-Esirkepov walk (esirkepovWeights1DSparse,
// src/deposition/ngp.cpp:36-84) for moves of any length, measured in cells
// from the home cell's lower face.  Returns the number of support cells
// (<= kEskMaxCells) or -1 if the move is longer.
constexpr int kEskMaxCells = 16;
__device__ __noinline__ int esk_walk_1d(float x0, float v, float dt, float inv_d, double dd, int exact, int* cell,
                                        float* w) {
    int c0;
    float fr0;
    cell_coord(x0, inv_d, dd, exact, c0, fr0);
    const float disp = v * dt;
    if (fabsf(disp) < 1e-15f) {
        cell[0] = c0;
        w[0] = 1.0f - fr0;
        cell[1] = c0 + 1;
        w[1] = fr0;
        return 2;
    }
    const float end = fr0 + cells(disp, inv_d, dd, exact);  // relative to the home cell's lower face
    int n = 0;
    float total = 0.0f;
    float cur = fr0;
    int c = 0;
    if (disp > 0.0f) {
        while (cur < end) {
            if (n == kEskMaxCells) return -1;
            const float nxt = fminf(end, (float)(c + 1));
            cell[n] = c0 + c;
            w[n] = nxt - cur;
            total += w[n];
            ++n;
            cur = nxt;
            ++c;
        }
    } else {
        while (cur > end) {
            if (n == kEskMaxCells) return -1;
            const float nxt = fmaxf(end, (float)c);
            cell[n] = c0 + c;
            w[n] = cur - nxt;
            total += w[n];
            ++n;
            cur = nxt;
            --c;
        }
    }
    if (n == 0) {  // displacement below FP32 resolution: all weight on the home cell
        cell[0] = c0;
        w[0] = 1.0f;
        return 1;
    }
    if (total > 1e-30f) {
        const float inv = 1.0f / total;
        for (int i = 0; i < n; ++i) w[i] *= inv;
    }
    return n;
}

// Global accumulation in FP64 (one non-returning hardware atomic per value;
// sums stay exact to FP64 round-off whatever their order).
__device__ __forceinline__ void acc_add(double* __restrict__ gacc, size_t i, float v) {
    unsafeAtomicAdd(gacc + i, (double)v);
}

// Rare path of scatter_global: Esirkepov moves longer than one cell.
__device__ __noinline__ void scatter_esk_long(float x, float y, float vx, float vy, float qv, float dt,
                                              const DevDomain& d, double* __restrict__ gacc,
                                              int* __restrict__ status) {
    int cxs[kEskMaxCells], cys[kEskMaxCells];
    float wxs[kEskMaxCells], wys[kEskMaxCells];
    const int nx = esk_walk_1d(x, vx, dt, d.inv_dx, d.dx, d.exact, cxs, wxs);
    const int ny = esk_walk_1d(y, vy, dt, d.inv_dy, d.dy, d.exact, cys, wys);
    if (nx < 0 || ny < 0) {
        atomicOr(status, 1);
        return;
    }
    for (int i = 0; i < nx; ++i) {
        const int gx = wrap_index(cxs[i], d.Nx);
        for (int j = 0; j < ny; ++j) {
            const int gy = wrap_index(cys[j], d.Ny);
            acc_add(gacc, (size_t)gy * d.Nx + gx, qv * wxs[i] * wys[j]);
        }
    }
}

// Scatter one particle with FP64 global atomics (used by GPU_Atomics and by
// the tile kernel's overflow path).
template <int S>
__device__ __forceinline__ void scatter_global(float x, float y, float vx, float vy, float qv, float dt,
                                               const DevDomain& d, double* __restrict__ gacc,
                                               int* __restrict__ status) {
    using St = Stencil<S>;
    float w[St::V];
    int cx, cy;
    const bool ok = St::weights(x, y, vx, vy, qv, dt, d, cx, cy, w);
    if (ok) {
#pragma unroll
        for (int k = 0; k < St::V; ++k) {
            if (w[k] != 0.0f) {
                const int gx = wrap_index(cx + St::OX + k % St::SX, d.Nx);
                const int gy = wrap_index(cy + St::OY + k / St::SX, d.Ny);
                acc_add(gacc, (size_t)gy * d.Nx + gx, w[k]);
            }
        }
        return;
    }
    // Only the Esirkepov density can fail the fixed stencil (long moves).
    scatter_esk_long(x, y, vx, vy, qv, dt, d, gacc, status);
}

// Q consecutive particle quads of one lane: physical start index of each quad
// and the valid particle range [lo, hi) of the row it belongs to.
template <int Q>
struct Chunk {
    float4 X[Q], Y[Q], QQ[Q], VX[Q], VY[Q];
    int p0[Q], lo[Q], hi[Q];
};

// Maps the lane's virtual quads vq0 .. vq0+Q-1 (rows of the tile laid end to
// end) to physical quads with the wave's row table and issues the loads.
template <int Q, bool VEL>
__device__ __forceinline__ void load_chunk(int vq0, int total_q, int th, const int* __restrict__ t_qcum,
                                           const int* __restrict__ t_qstart, const int* __restrict__ t_rs,
                                           const int* __restrict__ t_re, const float* __restrict__ px,
                                           const float* __restrict__ py, const float* __restrict__ pq,
                                           const float* __restrict__ pvx, const float* __restrict__ pvy,
                                           Chunk<Q>& c) {
    int r = 0;
    while (r + 1 < th && vq0 >= t_qcum[r + 1]) ++r;
    int next = t_qcum[r + 1];
    int delta = t_qstart[r] - t_qcum[r];
    int lo = t_rs[r], hi = t_re[r];
#pragma unroll
    for (int j = 0; j < Q; ++j) {
        const int vq = vq0 + j;
        if (vq < total_q) {
            if (vq >= next) {  // lane straddles a row boundary (rare)
                while (vq >= t_qcum[r + 1]) ++r;
                next = t_qcum[r + 1];
                delta = t_qstart[r] - t_qcum[r];
                lo = t_rs[r];
                hi = t_re[r];
            }
            const int p0 = (vq + delta) * 4;
            c.p0[j] = p0;
            c.lo[j] = lo;
            c.hi[j] = hi;
            c.X[j] = *reinterpret_cast<const float4*>(px + p0);
            c.Y[j] = *reinterpret_cast<const float4*>(py + p0);
            c.QQ[j] = *reinterpret_cast<const float4*>(pq + p0);
            if (VEL) {
                c.VX[j] = *reinterpret_cast<const float4*>(pvx + p0);
                c.VY[j] = *reinterpret_cast<const float4*>(pvy + p0);
            }
        } else {
            c.p0[j] = 0;
            c.lo[j] = 1;  // empty range: no particle of this quad is valid
            c.hi[j] = 0;
        }
    }
}

// A particle whose current cell is outside its tile (drift since the last
// sort): its already computed stencil goes to the FP64 accumulation grid
// with global atomics.  `ok` false (an Esirkepov move longer than a
// cell) takes the general walk.
template <int S>
__device__ __forceinline__ void overflow_scatter(bool ok, int cx, int cy, const float (&w)[Stencil<S>::V], float x,
                                                 float y, float vx, float vy, float qv, float dt, const DevDomain& d,
                                                 double* __restrict__ gacc, int* __restrict__ status) {
    using St = Stencil<S>;
    if (ok) {
#pragma unroll
        for (int k = 0; k < St::V; ++k) {
            const int gx = wrap_index(cx + St::OX + k % St::SX, d.Nx);
            const int gy = wrap_index(cy + St::OY + k / St::SX, d.Ny);
            acc_add(gacc, (size_t)gy * d.Nx + gx, w[k]);
        }
    } else {
        scatter_esk_long(x, y, vx, vy, qv, dt, d, gacc, status);
    }
}

// Per-wave state of the tiled deposit (one tile per wavefront).
struct TileCtx {
    double* win;  // FP64 LDS window
    int* acc_mark;  // overflow marker (TileGeom::acc_mark), nullptr for none
    int epoch;
    int WX;
    int x0, y0, tw, th;  // the tile's own cells
    int ax0, ay0, aw, ah;  // cells deposited into the window: the tile widened by the drift halo
    int mrow;              // moment stencils: home nodes per row of the moment window (the tile width)
};

// Adds a run of window key k into LDS: the V stencil values at the stencil's
// nodes of the window, or, for a moment stencil, the V moments at the home
// node's slot of the moment window.
template <int S, int V>
__device__ __forceinline__ void emit_run(const TileCtx& t, int k, const float (&z)[V]) {
    using St = Stencil<S>;
    if constexpr (S == TSC_MOM) {
        const int base = ((k >> 12) * t.mrow + (k & 4095)) * V;
#pragma unroll
        for (int i = 0; i < V; ++i) atomicAdd(t.win + base + i, (double)z[i]);
    } else {
        const int wb = ((k >> 12) + 1 + St::OY) * t.WX + (k & 4095) + 1 + St::OX;
        emit_cell<St::SX, V>(t.win, wb, t.WX, z);
    }
}

// emit_run with the values negated (the negation folds into the FP64 conversion).
template <int S, int V>
__device__ __forceinline__ void emit_run_neg(const TileCtx& t, int k, const float (&z)[V]) {
    float m[V];
#pragma unroll
    for (int i = 0; i < V; ++i) m[i] = -z[i];
    emit_run<S, V>(t, k, m);
}

// Window key ((ly << 12) | lx) of a particle's home cell, -1 when the particle
// is not this lane's (outside its row range) or not in the window.
template <int S, int EX>
__device__ __forceinline__ int window_key(float xx, float yy, bool valid, const TileCtx& t, const DevDomain& d) {
    int cx, cy;
    float fx, fy;
    Stencil<S>::template home<EX>(xx, yy, d, cx, cy, fx, fy);
    const int lx = cx - t.ax0, ly = cy - t.ay0;
    const bool in_win = (unsigned)lx < (unsigned)t.aw && (unsigned)ly < (unsigned)t.ah;
    return (valid && in_win) ? ((ly << 12) | lx) : -1;
}

// Processes one chunk: K consecutive particles per lane accumulated as runs of
// equal cell in registers, then one segmented wave scan, then one FP64 LDS
// add per (cell, stencil slot).
//
// LA (lookahead, for particles that drifted since the sort): a particle whose
// cell differs from the open run while the next particle returns to it is
// added to the window on its own and the run stays open, so the scattered
// strays of a drifted chunk do not cut its runs into pieces that every lane
// then adds to the same few nodes.
template <int S, int Q, int EX, int LA>
__device__ __forceinline__ void process_chunk(const Chunk<Q>& cur, const TileCtx& t, const DevDomain& d, float dt,
                                              double* __restrict__ gacc, int* __restrict__ status) {
    static_assert(S != TSC_MOM, "the moment stencil has its own chunk processing (mom_chunk)");
    using St = Stencil<S>;
    constexpr int V = St::V;
    constexpr int K = 4 * Q;
    float acc[V], H[V];
#pragma unroll
    for (int k = 0; k < V; ++k) {
        acc[k] = 0.0f;
        H[k] = 0.0f;
    }
    int hk = -1, ck = -1;
    bool multi = false;
    int nk[K];  // lookahead keys (LA only)
    if (LA) {
#pragma unroll
        for (int j = 0; j < Q; ++j) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int p = cur.p0[j] + e;
                const bool valid = p >= cur.lo[j] && p < cur.hi[j];
                const float xx = e == 0 ? cur.X[j].x : e == 1 ? cur.X[j].y : e == 2 ? cur.X[j].z : cur.X[j].w;
                const float yy = e == 0 ? cur.Y[j].x : e == 1 ? cur.Y[j].y : e == 2 ? cur.Y[j].z : cur.Y[j].w;
                nk[4 * j + e] = window_key<S, EX>(xx, yy, valid, t, d);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < Q; ++j) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int p = cur.p0[j] + e;
            // Particles of the quad outside this row's range belong to other
            // tiles: they contribute zero and leave the current run open.
            const bool valid = p >= cur.lo[j] && p < cur.hi[j];
            const float xx = e == 0 ? cur.X[j].x : e == 1 ? cur.X[j].y : e == 2 ? cur.X[j].z : cur.X[j].w;
            const float yy = e == 0 ? cur.Y[j].x : e == 1 ? cur.Y[j].y : e == 2 ? cur.Y[j].z : cur.Y[j].w;
            const float qq = e == 0 ? cur.QQ[j].x : e == 1 ? cur.QQ[j].y : e == 2 ? cur.QQ[j].z : cur.QQ[j].w;
            float vxx = 0.0f, vyy = 0.0f;
            if (St::kNeedsVelocity) {
                vxx = e == 0 ? cur.VX[j].x : e == 1 ? cur.VX[j].y : e == 2 ? cur.VX[j].z : cur.VX[j].w;
                vyy = e == 0 ? cur.VY[j].x : e == 1 ? cur.VY[j].y : e == 2 ? cur.VY[j].z : cur.VY[j].w;
            }
            // Home cell first: the window test decides whether the weights go
            // to the window, so the charge of any other particle is zeroed
            // before the (linear in q) weights are formed.  Unwrapped cell:
            // positions outside [0, L) fall outside every tile and take the
            // (wrapping) overflow path.
            int cx, cy;
            float fx, fy;
            St::template home<EX>(xx, yy, d, cx, cy, fx, fy);
            const int lx = cx - t.ax0, ly = cy - t.ay0;
            const bool in_win = (unsigned)lx < (unsigned)t.aw && (unsigned)ly < (unsigned)t.ah;
            const float qv = valid ? qq * d.inv_vol : 0.0f;
            // Slots past the tile's particles hold unloaded registers and
            // quads may reach into the array's pad: a particle that is not
            // deposited here gets zero charge and finite fractions, so its
            // weights are exact zeros.
            const bool take0 = valid && in_win;
            float w[V];
            const bool ok = St::template fill<EX>(take0 ? fx : 0.0f, take0 ? fy : 0.0f, vxx, vyy, take0 ? qv : 0.0f,
                                                  dt, d, w);
            const bool inside = ok && in_win;
            if (valid && !inside) {
                // overflow (rare): weights with the particle's charge
                float wo[V];
                const bool ok_o = St::template fill<EX>(fx, fy, vxx, vyy, qv, dt, d, wo);
                overflow_scatter<S>(ok_o, cx, cy, wo, xx, yy, vxx, vyy, qv, dt, d, gacc, status);
                if (t.acc_mark) *t.acc_mark = t.epoch;  // the merge must fold the grid
            }
            const int lk = (valid && inside) ? ((ly << 12) | lx) : ck;
            bool nseg = lk != ck;
            // w is zero unless the particle is valid and in the window; only an
            // Esirkepov move longer than a cell (ok false) must be masked.
            const bool take = St::kAlwaysOk ? true : (valid && inside);
            if (LA) {
                const int idx = 4 * j + e;
                const bool stray = nseg && ck >= 0 && idx + 1 < K && nk[idx + 1] == ck;
                if (stray) {  // alone into the window; the run stays open
                    emit_run<S, V>(t, lk, w);
                    continue;
                }
            }
            if (nseg && ck >= 0) {
                if (!multi) {
#pragma unroll
                    for (int k = 0; k < V; ++k) H[k] = acc[k];
                    multi = true;
                } else {
                    emit_run<S, V>(t, ck, acc);
                }
            }
            const float keep = nseg ? 0.0f : 1.0f;
#pragma unroll
            for (int k = 0; k < V; ++k) acc[k] = fmaf(acc[k], keep, take ? w[k] : 0.0f);
            if (nseg) {
                ck = lk;
                if (hk < 0) hk = lk;
            }
        }
    }
    // Combine the runs that straddle lanes within each row of 16 lanes.  Lane
    // i's outgoing run is acc with key ck; it continues lane i-1's outgoing
    // run when hk == ck(i-1) and both lanes are in one row (a run crossing a
    // row boundary is added once per row: four scan steps instead of six).
    const int ck_prev = dpp_i<kWaveShr1, 0xf>(-2, ck);
    const bool cont = hk >= 0 && hk == ck_prev && (threadIdx.x & 15) != 0;
    const float a = (cont && !multi) ? 1.0f : 0.0f;
    float Z[V];
#pragma unroll
    for (int k = 0; k < V; ++k) Z[k] = acc[k];
    wave_segmented_scan<V, 4>(Z, a);
    float Zp[V];
#pragma unroll
    for (int k = 0; k < V; ++k) Zp[k] = dpp_z<kWaveShr1, 0xf>(Z[k]);
    if (multi) {
#pragma unroll
        for (int k = 0; k < V; ++k) H[k] = cont ? H[k] + Zp[k] : H[k];
        emit_run<S, V>(t, hk, H);
    }
    const int cont_next = dpp_i<kWaveShl1, 0xf>(0, cont ? 1 : 0);
    if (ck >= 0 && !cont_next) {
        emit_run<S, V>(t, ck, Z);
    }
}

// ---- TSC moments on the node-keyed order (no drift halo) -------------------
//
// Each lane sums the moments of its K particles about one reference node, the
// nearest node of its first particle: per particle the two offsets from that
// node, the |offset| <= 1/2 test and the nine moment sums, with no per-particle
// key or run bookkeeping.  On node-sorted input a lane's particles share that
// node except where a node's run ends inside the lane; such particles (and any
// that drifted since the sort, or whose reference node is not one of the
// tile's) are queued in an LDS ring and deposited 64 at a time, one per
// lane, by the general per-particle path.  The lanes then
// combine equal reference nodes with one segmented wave scan and one FP64 LDS
// add per moment and run.  Window moments carry q without 1 / cell volume (the
// tile's node values are scaled once when they are formed).
// Ring capacity (particles) per wave for K particles per lane: the at most
// 63 + 64 K entries pending after a chunk.  Entries hold the particle's x, y
// and q (three float arrays of this length), so the per-particle path
// re-reads nothing from global memory.
__host__ __device__ constexpr int mom_ring_cap(int K) { return 64 + 64 * K; }

// Offset in cells of coordinate x from node n.  With a power-of-two cell size
// x * inv_d is exact in FP32 and so is the difference (the same value as the
// floor-cell fraction minus 0 or 1); otherwise the product with the FP64
// reciprocal.
template <int EX>
__device__ __forceinline__ float node_offset(float x, float inv_d, double rinv, float n) {
    if constexpr (EX == 1) return fmaf(x, inv_d, -n);
    else return (float)fma((double)x, rinv, -(double)n);
}

// Fast path over one chunk.  Appends the indices of the particles it leaves to
// the ring (`tail` counts appended entries, wave-uniform).
template <int Q, int EX>
__device__ __forceinline__ void mom_chunk(Chunk<Q>& c, const TileCtx& t, const DevDomain& d, double rdx, double rdy,
                                          float* __restrict__ ring, int& tail) {
    constexpr int CAP = mom_ring_cap(4 * Q);
    const int lane = threadIdx.x & 63;
    using St = Stencil<TSC_MOM>;
    constexpr int V = St::V;
    // Quad slots outside the row ranges (a row's first and last quads, lanes
    // past the tile's particles) get zero charge and the coordinates of the
    // lane's first particle in range, so they pass the node test with it and
    // add exact zeros; a lane with no particle in range is not live.
    bool partial = false;
#pragma unroll
    for (int j = 0; j < Q; ++j) partial |= !(c.p0[j] >= c.lo[j] && c.p0[j] + 4 <= c.hi[j]);
    bool live = true;
    if (partial) {
        float x0 = 0.0f, y0 = 0.0f;
        bool found = false;
#pragma unroll
        for (int k = 4 * Q - 1; k >= 0; --k) {
            const int j = k >> 2, e = k & 3;
            const int p = c.p0[j] + e;
            const bool valid = p >= c.lo[j] && p < c.hi[j];
            const float4 X = c.X[j], Y = c.Y[j];
            const float xx = e == 0 ? X.x : e == 1 ? X.y : e == 2 ? X.z : X.w;
            const float yy = e == 0 ? Y.x : e == 1 ? Y.y : e == 2 ? Y.z : Y.w;
            x0 = valid ? xx : x0;
            y0 = valid ? yy : y0;
            found |= valid;
        }
        live = found;
#pragma unroll
        for (int j = 0; j < Q; ++j) {
            const bool v0 = c.p0[j] >= c.lo[j] && c.p0[j] < c.hi[j];
            const bool v1 = c.p0[j] + 1 >= c.lo[j] && c.p0[j] + 1 < c.hi[j];
            const bool v2 = c.p0[j] + 2 >= c.lo[j] && c.p0[j] + 2 < c.hi[j];
            const bool v3 = c.p0[j] + 3 >= c.lo[j] && c.p0[j] + 3 < c.hi[j];
            c.X[j] = make_float4(v0 ? c.X[j].x : x0, v1 ? c.X[j].y : x0, v2 ? c.X[j].z : x0, v3 ? c.X[j].w : x0);
            c.Y[j] = make_float4(v0 ? c.Y[j].x : y0, v1 ? c.Y[j].y : y0, v2 ? c.Y[j].z : y0, v3 ? c.Y[j].w : y0);
            c.QQ[j] = make_float4(v0 ? c.QQ[j].x : 0.0f, v1 ? c.QQ[j].y : 0.0f, v2 ? c.QQ[j].z : 0.0f,
                                  v3 ? c.QQ[j].w : 0.0f);
        }
    }
    // Reference node: the nearest node of the first slot (node_cell), kept
    // unwrapped for the offsets; its window key, -1 outside the tile.
    int cx, cy;
    float fx, fy;
    cell_coord<EX>(c.X[0].x, d.inv_dx, d.dx, d.exact, cx, fx);
    cell_coord<EX>(c.Y[0].x, d.inv_dy, d.dy, d.exact, cy, fy);
    const int nux = cx + (fx >= 0.5f ? 1 : 0), nuy = cy + (fy >= 0.5f ? 1 : 0);
    const int nwx = (nux == d.Nx && cx == d.Nx - 1) ? 0 : nux;
    const int nwy = (nuy == d.Ny && cy == d.Ny - 1) ? 0 : nuy;
    const int lx = nwx - t.ax0, ly = nwy - t.ay0;
    const bool inw = (unsigned)lx < (unsigned)t.aw && (unsigned)ly < (unsigned)t.ah;
    const int key = !live ? -2 : inw ? ((ly << 12) | lx) : -1;
    const bool fast = key >= 0;
    const float nfx = (float)nux, nfy = (float)nuy;
    // A particle is taken when both offsets are within 1/2 of the reference
    // node (lim = -1 takes none); the ring gets the untaken particles of the
    // lanes with a reference node: those past one of its runs, or drifted.
    const float lim = fast ? 0.5f : -1.0f;
    const uint64_t fast_lanes = __builtin_amdgcn_ballot_w64(fast);

    float s[V];
#pragma unroll
    for (int k = 0; k < V; ++k) s[k] = 0.0f;
#pragma unroll
    for (int j = 0; j < Q; ++j) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float xx = e == 0 ? c.X[j].x : e == 1 ? c.X[j].y : e == 2 ? c.X[j].z : c.X[j].w;
            const float yy = e == 0 ? c.Y[j].x : e == 1 ? c.Y[j].y : e == 2 ? c.Y[j].z : c.Y[j].w;
            const float qq = e == 0 ? c.QQ[j].x : e == 1 ? c.QQ[j].y : e == 2 ? c.QQ[j].z : c.QQ[j].w;
            const float ox = node_offset<EX>(xx, d.inv_dx, rdx, nfx);
            const float oy = node_offset<EX>(yy, d.inv_dy, rdy, nfy);
            // |offset| = 1/2 (a tie between two nodes) gives the same TSC
            // weights on either node, so it may stay with the reference node.
            const bool untaken = fmaxf(fabsf(ox), fabsf(oy)) > lim;
            const float qf = untaken ? 0.0f : qq;
            const float m10 = qf * ox, m01 = qf * oy;
            const float m20 = m10 * ox, m11 = m10 * oy, m21 = m20 * oy;
            s[0] += qf;
            s[1] += m10;
            s[2] += m20;
            s[3] += m01;
            s[4] += m11;
            s[5] += m21;
            s[6] = fmaf(m01, oy, s[6]);
            s[7] = fmaf(m11, oy, s[7]);
            s[8] = fmaf(m21, oy, s[8]);
            // Slots outside the row ranges are taken in a lane with a reference
            // node (they repeat its first particle), so only particles are queued.
            const uint64_t qm = __builtin_amdgcn_ballot_w64(untaken) & fast_lanes;
            if (qm) {
                if (untaken && fast) {
                    const int r = tail + __builtin_amdgcn_mbcnt_hi((unsigned)(qm >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((unsigned)qm, 0u));
                    ring[r] = xx;
                    ring[CAP + r] = yy;
                    ring[2 * CAP + r] = qq;
                }
                tail += (int)__popcll(qm);
            }
        }
    }
    // A lane whose reference node is not the tile's (a particle that drifted
    // or lies outside [0, L)) leaves every particle in range to the ring.
    if (__builtin_amdgcn_ballot_w64(live && !fast)) {
#pragma unroll
        for (int k = 0; k < 4 * Q; ++k) {
            const int j = k >> 2, e = k & 3;
            const int p = c.p0[j] + e;
            const bool queue = live && !fast && p >= c.lo[j] && p < c.hi[j];
            const uint64_t qm = __builtin_amdgcn_ballot_w64(queue);
            if (qm) {
                if (queue) {
                    const int r = tail + __builtin_amdgcn_mbcnt_hi((unsigned)(qm >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((unsigned)qm, 0u));
                    ring[r] = e == 0 ? c.X[j].x : e == 1 ? c.X[j].y : e == 2 ? c.X[j].z : c.X[j].w;
                    ring[CAP + r] = e == 0 ? c.Y[j].x : e == 1 ? c.Y[j].y : e == 2 ? c.Y[j].z : c.Y[j].w;
                    ring[2 * CAP + r] = e == 0 ? c.QQ[j].x : e == 1 ? c.QQ[j].y : e == 2 ? c.QQ[j].z : c.QQ[j].w;
                }
                tail += (int)__popcll(qm);
            }
        }
    }
    // Lanes of equal reference node are consecutive on sorted input.  Within
    // each 16-lane row, P = the inclusive prefix of the lanes' sums (one
    // fused DPP add per value and step); a run of lanes [s, e] of one node
    // sums to P_e - P_{s-1}.  So the last lane of each run of equal keys adds
    // +P_e to its node and, when another run follows in the row, -P_e to the
    // next run's node (runs of lanes without a reference node add exact zeros
    // to P and only pass -P on).
#pragma unroll
    for (int k = 0; k < V; ++k) {
        s[k] += dpp_z<kRowShr + 1, 0xf>(s[k]);
        s[k] += dpp_z<kRowShr + 2, 0xf>(s[k]);
        s[k] += dpp_z<kRowShr + 4, 0xf>(s[k]);
        s[k] += dpp_z<kRowShr + 8, 0xf>(s[k]);
    }
    const int knext = dpp_i<kWaveShl1, 0xf>(-3, key);
    const bool row_end = (lane & 15) == 15;
    const bool run_end = row_end || knext != key;
    if (run_end && fast) emit_run<TSC_MOM, V>(t, key, s);
    if (run_end && !row_end && knext >= 0) emit_run_neg<TSC_MOM, V>(t, knext, s);
}

// Deposits ring entries [first, first + cnt), cnt <= 64, one per lane, with
// the general path: nearest node, window test, moments into the window or
// node weights to the FP64 accumulation grid.
template <int EX, int CAP>
__device__ __forceinline__ void mom_ring_batch(const TileCtx& t, const float* __restrict__ ring, int first, int cnt,
                                               const DevDomain& d, double* __restrict__ gacc,
                                               int* __restrict__ status) {
    const int lane = threadIdx.x & 63;
    if (lane >= cnt) return;
    const float xx = ring[first + lane], yy = ring[CAP + first + lane], qq = ring[2 * CAP + first + lane];
    int cx, cy;
    float fx, fy;
    node_cell<EX>(xx, yy, d, cx, cy, fx, fy);
    const int lx = cx - t.ax0, ly = cy - t.ay0;
    if ((unsigned)lx < (unsigned)t.aw && (unsigned)ly < (unsigned)t.ah) {
        float w[9];
        Stencil<TSC_MOM>::template fill<EX>(fx, fy, 0.0f, 0.0f, qq, 0.0f, d, w);
        emit_run<TSC_MOM, 9>(t, (ly << 12) | lx, w);
    } else {
        const float qv = qq * d.inv_vol;
        float w[9];
        Stencil<TSC_NODE>::template fill<EX>(fx, fy, 0.0f, 0.0f, qv, 0.0f, d, w);
        overflow_scatter<TSC_NODE>(true, cx, cy, w, xx, yy, 0.0f, 0.0f, qv, 0.0f, d, gacc, status);
        if (t.acc_mark) *t.acc_mark = t.epoch;  // the merge must fold the grid
    }
}

// Deposits the ring's full batches of 64 and moves the (< 64) entries left
// to its front; returns their count.
template <int EX, int CAP>
__device__ __forceinline__ int mom_ring_drain(const TileCtx& t, float* __restrict__ ring, int tail,
                                              const DevDomain& d, double* __restrict__ gacc,
                                              int* __restrict__ status) {
    const int lane = threadIdx.x & 63;
    wave_sync();
    int first = 0;
    for (; tail - first >= 64; first += 64) mom_ring_batch<EX, CAP>(t, ring, first, 64, d, gacc, status);
    const int left = tail - first;
    if (first > 0 && left > 0) {
        float a = 0.0f, b = 0.0f, q = 0.0f;
        if (lane < left) {
            a = ring[first + lane];
            b = ring[CAP + first + lane];
            q = ring[2 * CAP + first + lane];
        }
        wave_sync();
        if (lane < left) {
            ring[lane] = a;
            ring[CAP + lane] = b;
            ring[2 * CAP + lane] = q;
        }
    }
    wave_sync();
    return left;
}

// Deposits one tile with one wavefront: LDS window `lds` (g.lds_per_wave
// floats, 8-byte aligned), cell offsets and partial / accumulation buffers
// already offset to the tile's instance.
template <int S, int K, int EX, int PIPE, int LA = 0>
__device__ __forceinline__ void deposit_one_tile(int tile, float* __restrict__ lds, const float* __restrict__ px,
                                                 const float* __restrict__ py, const float* __restrict__ pvx,
                                                 const float* __restrict__ pvy, const float* __restrict__ pq,
                                                 const int* __restrict__ cell_off, const DevDomain& d,
                                                 const TileGeom& g, float dt, float* __restrict__ partial,
                                                 double* __restrict__ gacc, int* __restrict__ status) {
    using St = Stencil<S>;
    constexpr int Q = K / 4;
    static_assert(K % 4 == 0, "K must be a multiple of 4 (float4 loads)");
    const int lane = threadIdx.x & 63;

    TileCtx t;
    t.win = reinterpret_cast<double*>(lds);
    t.acc_mark = g.acc_mark;
    t.epoch = g.epoch;
    t.mrow = g.T;
    // LDS window: node values, or for a moment stencil (no drift halo) the
    // V moments of every home node of the tile.
    const int wlen = S == TSC_MOM ? St::V * g.T * g.R : g.wsize;
    int* tab = reinterpret_cast<int*>(t.win + wlen);
    int* t_qstart = tab;          // [R]   first aligned quad of row r
    int* t_rs = tab + g.R;        // [R]   first particle of row r
    int* t_re = tab + 2 * g.R;    // [R]   one past the last particle of row r
    int* t_qcum = tab + 3 * g.R;  // [R+1] cumulative quad counts

    const int tx = tile % g.ntx, ty = tile / g.ntx;
    t.x0 = tx * g.T;
    t.y0 = ty * g.R;
    t.tw = min(g.T, d.Nx - t.x0);
    t.th = min(g.R, d.Ny - t.y0);
    t.ax0 = t.x0 - g.D;
    t.ay0 = t.y0 - g.D;
    t.aw = t.tw + 2 * g.D;
    t.ah = t.th + 2 * g.D;
    t.WX = g.WX;
    const int th = t.th;

    for (int i = lane; i < wlen; i += 64) {
        t.win[i] = 0.0;
    }
    // Row table: lane r < th describes row y0 + r of the tile.
    int nq = 0;
    if (lane < th) {
        const int c = (t.y0 + lane) * d.Nx + t.x0;
        const int s = cell_off[c], e = cell_off[c + t.tw];
        const int qs = s >> 2, qe = (e + 3) >> 2;
        nq = e > s ? qe - qs : 0;
        t_qstart[lane] = qs;
        t_rs[lane] = s;
        t_re[lane] = e;
    }
    // Inclusive prefix of nq over the th (at most 64) row lanes.
    int incl = nq;
    for (int o = 1; o < th; o <<= 1) {
        const int u = __shfl_up(incl, o, 64);
        if (lane >= o) incl += u;
    }
    if (lane < th) t_qcum[lane + 1] = incl;
    if (lane == 0) t_qcum[0] = 0;
    wave_sync();
    const int total_q = t_qcum[th];
    const int step = 64 * Q;

    if constexpr (S == TSC_MOM) {
        constexpr int CAP = mom_ring_cap(K);
        float* ring = reinterpret_cast<float*>(t_qcum + g.R + 2);  // x, y, q arrays of CAP; 8-byte aligned
        int tail = 0;
        const double rdx = EX == 1 ? 0.0 : 1.0 / d.dx, rdy = EX == 1 ? 0.0 : 1.0 / d.dy;
        // Full batches of ring entries after each chunk, the rest at the end.
        auto chunk = [&](Chunk<Q>& A) {
            mom_chunk<Q, EX>(A, t, d, rdx, rdy, ring, tail);
            tail = __builtin_amdgcn_readfirstlane(tail);
            if (tail >= 64) tail = mom_ring_drain<EX, CAP>(t, ring, tail, d, gacc, status);
        };
        if (PIPE) {
            Chunk<Q> A, B;
            int base = 0;
            if (base < total_q)
                load_chunk<Q, false>(lane * Q, total_q, th, t_qcum, t_qstart, t_rs, t_re, px, py, pq, pvx, pvy, A);
            while (base < total_q) {
                const int b1 = base + step;
                if (b1 < total_q)
                    load_chunk<Q, false>(b1 + lane * Q, total_q, th, t_qcum, t_qstart, t_rs, t_re, px, py, pq, pvx,
                                         pvy, B);
                chunk(A);
                if (b1 >= total_q) break;
                const int b2 = b1 + step;
                if (b2 < total_q)
                    load_chunk<Q, false>(b2 + lane * Q, total_q, th, t_qcum, t_qstart, t_rs, t_re, px, py, pq, pvx,
                                         pvy, A);
                chunk(B);
                base = b2;
            }
        } else {
            for (int base = 0; base < total_q; base += step) {
                Chunk<Q> A;
                load_chunk<Q, false>(base + lane * Q, total_q, th, t_qcum, t_qstart, t_rs, t_re, px, py, pq, pvx, pvy,
                                     A);
                chunk(A);
            }
        }
        if (tail > 0) {
            wave_sync();
            mom_ring_batch<EX, CAP>(t, ring, 0, tail, d, gacc, status);
        }
    } else if (PIPE) {
        // Two register sets: the loads of one chunk are in flight while the
        // other is processed.
        Chunk<Q> A, B;
        int base = 0;
        if (base < total_q)
            load_chunk<Q, St::kNeedsVelocity>(lane * Q, total_q, th, t_qcum, t_qstart, t_rs, t_re, px, py, pq, pvx,
                                              pvy, A);
        while (base < total_q) {
            const int b1 = base + step;
            if (b1 < total_q)
                load_chunk<Q, St::kNeedsVelocity>(b1 + lane * Q, total_q, th, t_qcum, t_qstart, t_rs, t_re, px, py,
                                                  pq, pvx, pvy, B);
            process_chunk<S, Q, EX, LA>(A, t, d, dt, gacc, status);
            if (b1 >= total_q) break;
            const int b2 = b1 + step;
            if (b2 < total_q)
                load_chunk<Q, St::kNeedsVelocity>(b2 + lane * Q, total_q, th, t_qcum, t_qstart, t_rs, t_re, px, py,
                                                  pq, pvx, pvy, A);
            process_chunk<S, Q, EX, LA>(B, t, d, dt, gacc, status);
            base = b2;
        }
    } else {
        for (int base = 0; base < total_q; base += step) {
            Chunk<Q> A;
            load_chunk<Q, St::kNeedsVelocity>(base + lane * Q, total_q, th, t_qcum, t_qstart, t_rs, t_re, px, py, pq,
                                              pvx, pvy, A);
            process_chunk<S, Q, EX, LA>(A, t, d, dt, gacc, status);
        }
    }

    wave_sync();
    float* out = partial + (size_t)tile * g.wsize;
    if constexpr (S == TSC_MOM) {
        // Node values of the window from the home nodes' moments: each lane
        // maps one home node's nine moments to its 3 x 3 node values (the
        // TSC weight polynomials, separable in x and y) and adds them to the
        // FP64 node window; in each of the nine adds the lanes' target nodes
        // differ.  Window node (wx, wy) is tile node (wx - 1, wy - 1).
        // With at most 64 home nodes every lane holds its node's moments
        // before the window is cleared, so the node window overlays them.
        const bool overlay = g.T * g.R <= 64;
        double* nw = overlay ? t.win : reinterpret_cast<double*>(t_qcum + g.R + 2 + 3 * mom_ring_cap(K));
        double mm[9];
        if (overlay && lane < g.T * g.R) {
#pragma unroll
            for (int k = 0; k < 9; ++k) mm[k] = t.win[(size_t)lane * St::V + k];
        }
        wave_sync();
        for (int i = lane; i < g.wsize; i += 64) nw[i] = 0.0;
        wave_sync();
        for (int h = lane; h < g.T * g.R; h += 64) {
            const int hy = h / g.T, hx = h - hy * g.T;
            if (hx < t.tw && hy < t.th) {
                if (!overlay) {
#pragma unroll
                    for (int k = 0; k < 9; ++k) mm[k] = t.win[(size_t)h * St::V + k];
                }
                // x: node h-1 1/8 m0 - 1/2 m1 + 1/2 m2, node h 3/4 m0 - m2,
                // node h+1 1/8 m0 + 1/2 m1 + 1/2 m2 (m_i: moment order i in x)
                double u[3][3];  // [x node a][y moment order j]
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const double A = fma(0.125, mm[3 * j], 0.5 * mm[3 * j + 2]), B = 0.5 * mm[3 * j + 1];
                    u[0][j] = A - B;
                    u[1][j] = fma(0.75, mm[3 * j], -mm[3 * j + 2]);
                    u[2][j] = A + B;
                }
                const int base = hy * g.WX + hx;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const double A = fma(0.125, u[a][0], 0.5 * u[a][2]), B = 0.5 * u[a][1];
                    atomicAdd(nw + base + a, A - B);
                    atomicAdd(nw + base + g.WX + a, fma(0.75, u[a][0], -u[a][2]));
                    atomicAdd(nw + base + 2 * g.WX + a, A + B);
                }
            }
        }
        wave_sync();
        const double iv = (double)d.inv_vol;
        for (int i = lane; i < g.wsize; i += 64) out[i] = (float)(nw[i] * iv);
    } else {
        for (int i = lane; i < g.wsize; i += 64) out[i] = (float)t.win[i];
    }
    wave_sync();  // the window may be reused for the wave's next tile
}

// Tiles (along one axis) whose window can hold node g: those owning cells
// g-2 .. g+1 (window halo is 1 before and 2 after the tile).  Writes up to 4
// distinct tile indices in a fixed order.
__device__ __forceinline__ int window_candidates(int gnode, int T, int nt, int n, int D,
                                                 int (&t)[4 + 2 * kMaxDriftHalo]) {
    int cnt = 0;
    for (int o = -2 - D; o <= 1 + D; ++o) {
        const int c = wrap_index(gnode + o, n);
        const int ti = c / T;
        bool dup = false;
        for (int i = 0; i < cnt; ++i) dup |= (t[i] == ti);
        if (!dup) t[cnt++] = ti;
    }
    (void)nt;
    return cnt;
}

// Fast path (every tile at least 3 cells wide and tall, at least 3 tiles per
// axis): along each axis a node lies in its own tile's window and, near the
// tile's edge, in one neighbour's.  Writes (tile, window index) pairs.
__device__ __forceinline__ int axis_pairs(int gn, int T, int nt, int n, int (&t)[2], int (&w)[2]) {
    const int t0 = gn / T;
    const int x0 = t0 * T;
    const int tw0 = min(T, n - x0);
    const int l = gn - x0;
    t[0] = t0;
    w[0] = l + 1;
    if (l < 2) {  // the left neighbour's window reaches two nodes past its end
        const int tl = t0 == 0 ? nt - 1 : t0 - 1;
        const int twl = min(T, n - tl * T);
        t[1] = tl;
        w[1] = l + twl + 1;
        return 2;
    }
    if (l >= tw0 - 1) {  // the right neighbour's window starts one node before it
        t[1] = t0 == nt - 1 ? 0 : t0 + 1;
        w[1] = l - tw0 + 1;
        return 2;
    }
    return 1;
}

// Equal tiles of T cells dividing the axis, windows [tT - 1 - D, tT + T + 2 + D):
// the tiles whose windows hold node gn are the consecutive (unwrapped) tiles
// floor((gn - 2 - D) / T) .. floor((gn + 1 + D) / T), the same tiles in the same
// order as window_candidates finds them.  Returns the first one's (wrapped)
// tile index t0 and window index w0 and their count.
__device__ __forceinline__ void axis_first(int gn, int T, int nt, int D, int& t0, int& w0, int& cnt) {
    const int a = gn - 2 - D;
    const int tlo = a >= 0 ? a / T : -((-a + T - 1) / T);
    const int thi = (gn + 1 + D) / T;
    cnt = min(thi - tlo + 1, 4 + 2 * kMaxDriftHalo);
    t0 = ((tlo % nt) + nt) % nt;
    w0 = gn - (tlo * T - 1 - D);
}

// Sum at node (gx, gy) of the windows covering it (FP32 or FP64 window
// values P), added in FP64 in a fixed order.  B: candidates loaded together
// on the range path (TileGeom::mbatch).
template <int B = 16, typename P = float>
__device__ __forceinline__ double merge_windows(int gx, int gy, const P* __restrict__ partial,
                                                const DevDomain& d, const TileGeom& g) {
    double s = 0.0;
    // All window loads of a node are issued before the first add (a load per
    // add would wait out one L2 round trip per window); the adds keep the
    // row-major order of the tiles.
    if (g.fast) {
        int txs[2], wxs[2], tys[2], wys[2];
        const int ncx = axis_pairs(gx, g.T, g.ntx, d.Nx, txs, wxs);
        const int ncy = axis_pairs(gy, g.R, g.nty, d.Ny, tys, wys);
        // A missing second candidate repeats the first: every slot loads a
        // valid window value and the adds mask by value (s starts at +0 and
        // never becomes -0, so adding +0 leaves it unchanged).
        if (ncx < 2) {
            txs[1] = txs[0];
            wxs[1] = wxs[0];
        }
        if (ncy < 2) {
            tys[1] = tys[0];
            wys[1] = wys[0];
        }
        P v[2][2];
#pragma unroll
        for (int iy = 0; iy < 2; ++iy)
#pragma unroll
            for (int ix = 0; ix < 2; ++ix)
                v[iy][ix] = partial[(size_t)(tys[iy] * g.ntx + txs[ix]) * g.wsize + (size_t)wys[iy] * g.WX + wxs[ix]];
#pragma unroll
        for (int iy = 0; iy < 2; ++iy)
#pragma unroll
            for (int ix = 0; ix < 2; ++ix) s += (iy < ncy && ix < ncx) ? (double)v[iy][ix] : 0.0;
    } else if (g.fast_rows) {
        // Tiles of one or two rows (no halo): along x as in the fast path,
        // along y up to four consecutive tile rows, the same for every node
        // of a grid row.
        int txs[2], wxs[2];
        const int ncx = axis_pairs(gx, g.T, g.ntx, d.Nx, txs, wxs);
        if (ncx < 2) {
            txs[1] = txs[0];
            wxs[1] = wxs[0];
        }
        int ty0, wy0, ncy;
        axis_first(gy, g.R, g.nty, 0, ty0, wy0, ncy);
        P v[4][2];
        int ty = ty0, wy = wy0;
#pragma unroll
        for (int iy = 0; iy < 4; ++iy) {
            const int ry = iy < ncy ? ty : ty0, sy = iy < ncy ? wy : wy0;
#pragma unroll
            for (int ix = 0; ix < 2; ++ix)
                v[iy][ix] = partial[((size_t)ry * g.ntx + txs[ix]) * g.wsize + (size_t)sy * g.WX + wxs[ix]];
            ty = ty + 1 == g.nty ? 0 : ty + 1;
            wy -= g.R;
        }
#pragma unroll
        for (int iy = 0; iy < 4; ++iy)
#pragma unroll
            for (int ix = 0; ix < 2; ++ix) s += (iy < ncy && ix < ncx) ? (double)v[iy][ix] : 0.0;
    } else if (g.uniform) {
        int tx0, wx0, ncx, ty0, wy0, ncy;
        axis_first(gx, g.T, g.ntx, g.D, tx0, wx0, ncx);
        axis_first(gy, g.R, g.nty, g.D, ty0, wy0, ncy);
        const int total = ncx * ncy;
        const size_t a0 = ((size_t)ty0 * g.ntx + tx0) * g.wsize + (size_t)wy0 * g.WX + wx0;
        // Candidates in row-major order: along an axis the next candidate is
        // the next tile (wrapped) and the window index drops by the tile size.
        // The first B are loaded together (slots past the end load the first
        // candidate and add +0, as in the fast path); larger sets add the
        // rest one by one.
        constexpr int kBatch = B;
        int tx = tx0, wx = wx0, ty = ty0, wy = wy0, ix = 0;
        auto advance = [&]() {
            tx = tx + 1 == g.ntx ? 0 : tx + 1;
            wx -= g.T;
            if (++ix == ncx) {
                ix = 0;
                tx = tx0;
                wx = wx0;
                ty = ty + 1 == g.nty ? 0 : ty + 1;
                wy -= g.R;
            }
        };
        P v[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            const bool ok = k < total;
            v[k] = partial[ok ? ((size_t)ty * g.ntx + tx) * g.wsize + (size_t)wy * g.WX + wx : a0];
            if (ok) advance();
        }
#pragma unroll
        for (int k = 0; k < kBatch; ++k) s += k < total ? (double)v[k] : 0.0;
        for (int k = kBatch; k < total; ++k) {
            s += (double)partial[((size_t)ty * g.ntx + tx) * g.wsize + (size_t)wy * g.WX + wx];
            advance();
        }
    } else {
        int txs[4 + 2 * kMaxDriftHalo], tys[4 + 2 * kMaxDriftHalo];
        const int D = g.D;
        const int ncx = window_candidates(gx, g.T, g.ntx, d.Nx, D, txs);
        const int ncy = window_candidates(gy, g.R, g.nty, d.Ny, D, tys);
        for (int iy = 0; iy < ncy; ++iy) {
            const int ty = tys[iy];
            const int y0 = ty * g.R;
            const int th = min(g.R, d.Ny - y0);
            // images gy' = gy + m*Ny inside [y0 - 1 - D, y0 + th + 2 + D)
            int gyi = gy;
            while (gyi - d.Ny >= y0 - 1 - D) gyi -= d.Ny;
            while (gyi < y0 - 1 - D) gyi += d.Ny;
            for (; gyi < y0 + th + 2 + D; gyi += d.Ny) {
                const int wy = gyi - (y0 - 1 - D);
                for (int ix = 0; ix < ncx; ++ix) {
                    const int tx = txs[ix];
                    const int x0 = tx * g.T;
                    const int tw = min(g.T, d.Nx - x0);
                    int gxi = gx;
                    while (gxi - d.Nx >= x0 - 1 - D) gxi -= d.Nx;
                    while (gxi < x0 - 1 - D) gxi += d.Nx;
                    for (; gxi < x0 + tw + 2 + D; gxi += d.Nx) {
                        const int wx = gxi - (x0 - 1 - D);
                        const P v = partial[(size_t)(ty * g.ntx + tx) * g.wsize + (size_t)wy * g.WX + wx];
                        s += (double)v;
                    }
                }
            }
        }
    }
    return s;
}

// rho at node (gx, gy): the windows covering it, added in FP64 in a fixed
// order, plus the FP64 accumulation grid where it is non-zero (then cleared).
template <int B = 16>
__device__ __forceinline__ float merge_node(int gx, int gy, const float* __restrict__ partial,
                                            double* __restrict__ gacc, const DevDomain& d, const TileGeom& g) {
    // The marker is read before the windows so that its latency overlaps theirs.
    const bool fold = g.acc_mark == nullptr || *g.acc_mark == g.epoch;
    double s = merge_windows<B>(gx, gy, partial, d, g);
    // The accumulation grid is zero except where drifted particles were
    // scattered (and untouched when no tile marked this deposit); fold those
    // nodes in and clear them.
    if (fold) {
        const size_t node = (size_t)gy * d.Nx + gx;
        const double o = gacc[node];
        if (o != 0.0) {
            s += o;
            gacc[node] = 0.0;
        }
    }
    return (float)s;
}

}  // namespace
}  // namespace pichip
