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

template <int EX>
__global__ __launch_bounds__(256) void esirkepov_current_kernel(const float* __restrict__ px,
                                                                const float* __restrict__ py,
                                                                const float* __restrict__ pvx,
                                                                const float* __restrict__ pvy,
                                                                const float* __restrict__ pq, int64_t n, DevDomain d,
                                                                float dt, CurrentGrids g, int* __restrict__ status) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int bx, by;
    float S0x[4], S1x[4], S0y[4], S1y[4];
    const bool okx = frame_1d<EX>(px[p], pvx[p], dt, d.inv_dx, d.dx, d.exact, bx, S0x, S1x);
    const bool oky = frame_1d<EX>(py[p], pvy[p], dt, d.inv_dy, d.dy, d.exact, by, S0y, S1y);
    if (!okx || !oky) {
        atomicOr(status, 2);
        return;
    }
    const float qv = pq[p] * d.inv_vol;
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

void launch_continuity_residual(const CurrentGrids& g, const DevDomain& d, double dt, double* partial, int nblocks,
                                double* out, hipStream_t s) {
    hipLaunchKernelGGL(continuity_partial_kernel, dim3(nblocks), dim3(256), 0, s, g, d, dt, partial);
    hipLaunchKernelGGL(continuity_final_kernel, dim3(1), dim3(1), 0, s, partial, nblocks, out);
}

}  // namespace pichip
