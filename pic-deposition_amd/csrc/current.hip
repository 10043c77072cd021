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

// Cell-sorted particles, one wavefront per tile of T x R cells (no drift
// halo).  The tile's work is cut into items of at most 16 consecutive
// particles of one cell (a cell of n particles gives ceil(n / 16) items), one
// item per lane and step, loaded 8 particles at a time.  All particles of a cell share the 4 x 4 node frame
// that starts one node below the cell, so a lane adds its item in 44 FP32
// registers (Jx 3 x 4, Jy 4 x 3, rho_old 2 x 2 and rho_new 4 x 4 frame slots,
// each particle's values formed as on the global path) and then adds them
// once into the wave's four FP64 LDS windows ((T + 3) x (R + 3) nodes each,
// the deposit's layout), written out in FP64 for the merge.  Items never
// cross a cell, so no lane changes frame inside a step.  A particle outside
// its sorted cell (moved since the sort) or with a move of a cell or more
// takes the global path.
constexpr int kCurItem = 16;  // particles per item
constexpr int kCurLoad = 8;   // particles loaded together
constexpr int kCurCells = 128;  // cells per tile at most (T x R = 64 x 2)

struct CurRun {
    float jx[12], jy[12], ro[4], rn[16];  // jx[a * 4 + b], jy[a * 3 + b], ro[(b-1) * 2 + a-1], rn[b * 4 + a]
};

// Adds an item of cell (lx, ly) of the tile into the four windows: frame node
// (a, b) is window node (lx + a, ly + b).
__device__ __forceinline__ void cur_run_emit(double* __restrict__ win, int wsize, int WX, int lx, int ly,
                                             const CurRun& r) {
    const int base = ly * WX + lx;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int w = base + b * WX + a;
            if (a < 3) atomicAdd(win + w, (double)r.jx[a * 4 + b]);
            if (b < 3) atomicAdd(win + wsize + w, (double)r.jy[a * 3 + b]);
            if (a >= 1 && a <= 2 && b >= 1 && b <= 2) atomicAdd(win + 2 * wsize + w, (double)r.ro[(b - 1) * 2 + a - 1]);
            atomicAdd(win + 3 * wsize + w, (double)r.rn[b * 4 + a]);
        }
}

template <int EX>
__global__ __launch_bounds__(256) void current_tile_kernel(const float* __restrict__ px, const float* __restrict__ py,
                                                           const float* __restrict__ pvx,
                                                           const float* __restrict__ pvy,
                                                           const float* __restrict__ pq,
                                                           const int* __restrict__ cell_off, DevDomain d, TileGeom tg,
                                                           float dt, double* __restrict__ partial, CurrentGrids g,
                                                           int* __restrict__ status) {
    extern __shared__ double cur_lds[];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: scalar tile arithmetic
    const int tile = blockIdx.x * (blockDim.x >> 6) + wv;
    if (tile >= tg.ntiles) return;  // wave-uniform
    double* win = cur_lds + (size_t)wv * (4 * tg.wsize + (3 * kCurCells + 2) / 2);  // Jx, Jy, rho_old, rho_new
    int* icum = reinterpret_cast<int*>(win + 4 * tg.wsize);  // [cells + 1] items before local cell c
    int* cbeg = icum + kCurCells + 1;                          // [cells] first particle of local cell c
    int* cend = cbeg + kCurCells;                              // [cells] one past its last
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
    const float fdx = (float)(d.dx / (double)dt), fdy = (float)(d.dy / (double)dt);
    for (int base = 0; base < total; base += 64) {
        const int item = base + lane;
        const bool live = item < total;
        // cell of this item: the last lc with icum[lc] <= item
        int lo = 0, hi = ncell - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (icum[mid] <= item) lo = mid;
            else hi = mid - 1;
        }
        const int lc = lo;
        const int ly = lc >= tw ? 1 : 0;  // tiles of at most two rows
        const int lx = lc - ly * tw;
        const int p0 = cbeg[lc] + (item - icum[lc]) * kCurItem;
        const int p1 = live ? min(p0 + kCurItem, cend[lc]) : p0;
        CurRun r;
#pragma unroll
        for (int k = 0; k < 12; ++k) r.jx[k] = r.jy[k] = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) r.ro[k] = 0.0f;
#pragma unroll
        for (int k = 0; k < 16; ++k) r.rn[k] = 0.0f;
        for (int h = 0; h < kCurItem; h += kCurLoad) {
            if (p0 + h >= p1) break;
            float xs[kCurLoad], ys[kCurLoad], vxs[kCurLoad], vys[kCurLoad], qs[kCurLoad];
#pragma unroll
            for (int v = 0; v < kCurLoad; ++v) {
                const int i = p0 + h + v < p1 ? p0 + h + v : p0;
                xs[v] = px[i];
                ys[v] = py[i];
                vxs[v] = pvx[i];
                vys[v] = pvy[i];
                qs[v] = pq[i];
            }
    #pragma unroll
            for (int v = 0; v < kCurLoad; ++v) {
                const int u = h + v;
                if (p0 + u >= p1) break;
                int bx, by;
                float S0x[4], S1x[4], S0y[4], S1y[4];
                const bool okx = frame_1d<EX>(xs[v], vxs[v], dt, d.inv_dx, d.dx, d.exact, bx, S0x, S1x);
                const bool oky = frame_1d<EX>(ys[v], vys[v], dt, d.inv_dy, d.dy, d.exact, by, S0y, S1y);
                if (!(okx && oky && bx + 1 == x0 + lx && by + 1 == y0 + ly)) {
                    current_particle_global<EX>(xs[v], ys[v], vxs[v], vys[v], qs[v], d, dt, g, status);
                    continue;
                }
                const float qv = qs[v] * d.inv_vol;
                const float cxq = -qv * fdx, cyq = -qv * fdy;
                float dSx[4], dSy[4];
    #pragma unroll
                for (int a = 0; a < 4; ++a) {
                    dSx[a] = S1x[a] - S0x[a];
                    dSy[a] = S1y[a] - S0y[a];
                }
    #pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const float wy = S0y[b] + 0.5f * dSy[b];
                    float acc = 0.0f;
    #pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        acc = fmaf(cxq * dSx[a], wy, acc);
                        r.jx[a * 4 + b] += acc;
                    }
                }
    #pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const float wx = S0x[a] + 0.5f * dSx[a];
                    float acc = 0.0f;
    #pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        acc = fmaf(cyq * dSy[b], wx, acc);
                        r.jy[a * 3 + b] += acc;
                    }
                }
                const float o1 = qv * S0x[1], o2 = qv * S0x[2];
                r.ro[0] += o1 * S0y[1];
                r.ro[1] += o2 * S0y[1];
                r.ro[2] += o1 * S0y[2];
                r.ro[3] += o2 * S0y[2];
    #pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const float un = qv * S1x[a];
    #pragma unroll
                    for (int b = 0; b < 4; ++b) r.rn[b * 4 + a] += un * S1y[b];
                }
            }
        }
        if (p1 > p0) cur_run_emit(win, tg.wsize, tg.WX, lx, ly, r);
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
    return (size_t)4 * (4 * tg.wsize + (3 * kCurCells + 2) / 2) * sizeof(double);
}

void launch_esirkepov_current_tiled(const float* x, const float* y, const float* vx, const float* vy, const float* q,
                                    const int* cell_off, const DevDomain& d, const TileGeom& tg, float dt,
                                    double* partial, const CurrentGrids& g, int* status, hipStream_t s) {
    const size_t G = (size_t)d.Nx * d.Ny;
    const unsigned gb = (unsigned)((G + 255) / 256);
    hipLaunchKernelGGL(zero4_kernel, dim3(gb), dim3(256), 0, s, g, G);
    const unsigned blocks = (unsigned)((tg.ntiles + 3) / 4);
    const size_t lds = current_tile_lds(tg);
    if (d.exact)
        hipLaunchKernelGGL(current_tile_kernel<1>, dim3(blocks), dim3(256), lds, s, x, y, vx, vy, q, cell_off, d, tg,
                           dt, partial, g, status);
    else
        hipLaunchKernelGGL(current_tile_kernel<0>, dim3(blocks), dim3(256), lds, s, x, y, vx, vy, q, cell_off, d, tg,
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
