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
// a segmented sum where a_i = 0 starts a new segment at lane i.
template <int V>
__device__ __forceinline__ void wave_segmented_scan(float (&B)[V], float A) {
    lin_scan_step<kRowShr + 1, 0xf>(B, A);
    lin_scan_step<kRowShr + 2, 0xf>(B, A);
    lin_scan_step<kRowShr + 4, 0xf>(B, A);
    lin_scan_step<kRowShr + 8, 0xf>(B, A);
    lin_scan_step<kRowBcast15, 0xa>(B, A);
    lin_scan_step<kRowBcast31, 0xc>(B, A);
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
                constexpr int SO = S == TSC_MOM ? (int)TSC_NODE : S;  // node weights, not moments
                float wo[V];
                const bool ok_o = Stencil<SO>::template fill<EX>(fx, fy, vxx, vyy, qv, dt, d, wo);
                overflow_scatter<SO>(ok_o, cx, cy, wo, xx, yy, vxx, vyy, qv, dt, d, gacc, status);
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
    // Combine the runs that straddle lanes.  Lane i's outgoing run is acc
    // with key ck; it continues lane i-1's outgoing run when hk == ck(i-1).
    const int ck_prev = dpp_i<kWaveShr1, 0xf>(-2, ck);
    const bool cont = hk >= 0 && hk == ck_prev;
    const float a = (cont && !multi) ? 1.0f : 0.0f;
    float Z[V];
#pragma unroll
    for (int k = 0; k < V; ++k) Z[k] = acc[k];
    wave_segmented_scan<V>(Z, a);
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
    // Inclusive prefix of nq over the (at most 64) row lanes.
    int incl = nq;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(incl, o, 64);
        if (lane >= o) incl += u;
    }
    if (lane < th) t_qcum[lane + 1] = incl;
    if (lane == 0) t_qcum[0] = 0;
    wave_sync();
    const int total_q = t_qcum[th];
    const int step = 64 * Q;

    if (PIPE) {
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
        // Window node (wx, wy) is tile node (wx - 1, wy - 1): the sum over the
        // home nodes within one node of it, in a fixed order, of the weight
        // of its offset mapped from the home node's moments.
        for (int i = lane; i < g.wsize; i += 64) {
            const int wy = i / g.WX, wx = i - wy * g.WX;
            double sum = 0.0;
            for (int b = 0; b < 3; ++b) {
                const int hy = wy - b;
                if (hy < 0 || hy >= t.th) continue;
                for (int a = 0; a < 3; ++a) {
                    const int hx = wx - a;
                    if (hx < 0 || hx >= t.tw) continue;
                    const double* m = t.win + (hy * t.mrow + hx) * St::V;
                    double col[3];
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        col[j] = tsc_moment_coeff(a, 0) * m[3 * j] + tsc_moment_coeff(a, 1) * m[3 * j + 1] +
                                 tsc_moment_coeff(a, 2) * m[3 * j + 2];
                    sum += tsc_moment_coeff(b, 0) * col[0] + tsc_moment_coeff(b, 1) * col[1] +
                           tsc_moment_coeff(b, 2) * col[2];
                }
            }
            out[i] = (float)sum;
        }
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
