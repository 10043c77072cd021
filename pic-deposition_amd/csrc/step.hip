// Kernels that close the timestep: CIC gather + leapfrog push, the spectral
// Poisson solve around hipFFT, FP64 diagnostics, and on-device particle
// initialisers.

#include <algorithm>

#include "pic_common.hpp"
#include "pic_kernels.hpp"
#include "particle_device.hpp"

namespace pichip {

namespace {

// Gather (v += -(q/m) E dt) then push (x += v dt) of one particle per lane;
// a non-null push_flag replaces do_push by its device value.
__global__ __launch_bounds__(256) void gather_push_kernel(float* __restrict__ px, float* __restrict__ py,
                                                          float* __restrict__ pvx, float* __restrict__ pvy,
                                                          const float* __restrict__ pq, const float* __restrict__ pm,
                                                          int64_t n, const float* __restrict__ Ex,
                                                          const float* __restrict__ Ey, DevDomain d, float dt,
                                                          int do_gather, int do_push,
                                                          const int* __restrict__ push_flag, int q_one, int m_one) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (push_flag) do_push = *push_flag;
    float x = px[i], y = py[i];
    float vx = pvx[i], vy = pvy[i];
    if (do_gather) {
        float ex, ey;
        gather_cic(x, y, Ex, Ey, d, ex, ey);
        // A stream that holds one value for every particle (a single
        // species' charge or mass) is read once, not per particle.
        const float qm = -pq[q_one ? 0 : i] / pm[m_one ? 0 : i];
        vx += qm * ex * dt;
        vy += qm * ey * dt;
        pvx[i] = vx;
        pvy[i] = vy;
    }
    if (do_push) {
        px[i] = push_coord(x, vx, dt, (float)d.Lx);
        py[i] = push_coord(y, vy, dt, (float)d.Ly);
    }
}

// Five FP64 sums in one pass; fixed-shape tree per block, then a single block
// adds the block partials in index order (deterministic).
__global__ __launch_bounds__(256) void diag_partial_kernel(const float* __restrict__ rho,
                                                           const float* __restrict__ Ex,
                                                           const float* __restrict__ Ey, int64_t G,
                                                           const float* __restrict__ vx,
                                                           const float* __restrict__ vy,
                                                           const float* __restrict__ q,
                                                           const float* __restrict__ m, int64_t n,
                                                           double* __restrict__ partials) {
    double s[5] = {0, 0, 0, 0, 0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < G; i += stride) {
        if (rho) s[0] += (double)rho[i];
        if (Ex) {
            const double ex = Ex[i], ey = Ey[i];
            s[1] += ex * ex + ey * ey;
        }
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double a = vx[i], b = vy[i], mm = m[i], qq = q[i];
        s[2] += 0.5 * mm * (a * a + b * b);
        s[3] += qq;
        s[4] += fabs(qq);
    }
    __shared__ double red[5][256];
#pragma unroll
    for (int k = 0; k < 5; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
#pragma unroll
            for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 5) partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void diag_final_kernel(const double* __restrict__ partials, int nblocks,
                                                         double* __restrict__ out) {
    __shared__ double sh[4];
    for (int k = 0; k < 5; ++k) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += partials[(size_t)k * nblocks + b];
        s = block_sum256(s, sh);
        if (threadIdx.x == 0) out[k] = s;
    }
}

__global__ __launch_bounds__(256) void init_kernel(float* __restrict__ px, float* __restrict__ py,
                                                   float* __restrict__ pvx, float* __restrict__ pvy,
                                                   float* __restrict__ pq, float* __restrict__ pm, int64_t n,
                                                   DevDomain d, InitParams p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ParticleInit o = init_particle(i, n, d, p);
    px[i] = o.x;
    py[i] = o.y;
    pvx[i] = o.vx;
    pvy[i] = o.vy;
    pq[i] = o.q;
    pm[i] = o.m;
}


// ---- spectral Poisson (reference src/pic/poisson_fft.cpp:45-94) ------------
// The reference zeroes the k = 0 coefficient; here the mean of rho*cv is
// removed in FP64 before the FP32 transform as well, so the DC term never
// enters the single-precision FFT.
__global__ __launch_bounds__(256) void poisson_prep_kernel(const float* __restrict__ rho, float* __restrict__ in,
                                                           int64_t G, double bg, double cv,
                                                           const double* __restrict__ rho_sum, double sum_scale,
                                                           double sum_add) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G) return;
    const double mean = (*rho_sum * sum_scale + sum_add + bg * (double)G) * cv / (double)G;
    in[i] = (float)(((double)rho[i] + bg) * cv - mean);
}

// phi_hat = rho_hat / (k^2 Nx Ny) (poisson_fft.cpp:59-66, k^2 table :23-35),
// then Ex_hat, Ey_hat from the forward-difference symbols.
__global__ __launch_bounds__(256) void poisson_spectral_kernel(float2* __restrict__ spec, float2* __restrict__ sex,
                                                               float2* __restrict__ sey, PoissonTables t, DevDomain d) {
    const int hx = d.Nx / 2 + 1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= hx) return;
    const size_t id = (size_t)j * hx + i;
    float2 p = make_float2(0.0f, 0.0f);
    if (i != 0 || j != 0) {
        const float f = (float)(1.0 / ((t.k2x[i] + t.k2y[j]) * (double)d.Nx * (double)d.Ny));
        const float2 v = spec[id];
        p = make_float2(v.x * f, v.y * f);
    }
    spec[id] = p;
    const float2 a = t.dx_sym[i], b = t.dy_sym[j];
    sex[id] = make_float2(-(a.x * p.x - a.y * p.y), -(a.x * p.y + a.y * p.x));
    sey[id] = make_float2(-(b.x * p.x - b.y * p.y), -(b.x * p.y + b.y * p.x));
}

__global__ __launch_bounds__(256) void efield_kernel(const float* __restrict__ phi, float* __restrict__ Ex,
                                                     float* __restrict__ Ey, DevDomain d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= d.Nx) return;
    const int ip = i + 1 == d.Nx ? 0 : i + 1;
    const int jp = j + 1 == d.Ny ? 0 : j + 1;
    const size_t id = (size_t)j * d.Nx + i;
    const float p = phi[id];
    Ex[id] = -((phi[(size_t)j * d.Nx + ip] - p) * (float)(1.0 / d.dx));
    Ey[id] = -((phi[(size_t)jp * d.Nx + i] - p) * (float)(1.0 / d.dy));
}

__global__ __launch_bounds__(256) void add_scalar_kernel(float* __restrict__ p, int64_t n, float v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] += v;
}


// ---- per-step diagnostics (recorded into device arrays) ---------------------
// Partial sums per block: 0 sum rho, 1 sum(Ex^2+Ey^2), 2 sum Ex cos(k x_i),
// 3 sum Ex sin(k x_i) (x_i = (i + 1/2) dx, extractFourierModeAmplitude,
// validation.cpp:99-119), 4 sum 0.5 m v^2, 5 sum q, 6 sum |q|.
constexpr int kStepSums = 7;

__global__ __launch_bounds__(256) void step_diag_partial_kernel(const float* __restrict__ rho,
                                                                const float* __restrict__ Ex,
                                                                const float* __restrict__ Ey,
                                                                const double* __restrict__ cosk,
                                                                const double* __restrict__ sink, int Nx, int64_t G,
                                                                const float* __restrict__ vx,
                                                                const float* __restrict__ vy,
                                                                const float* __restrict__ q,
                                                                const float* __restrict__ m, int64_t n,
                                                                double* __restrict__ partials) {
    double s[kStepSums] = {0, 0, 0, 0, 0, 0, 0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < G; i += stride) {
        const double ex = Ex[i], ey = Ey[i];
        s[0] += (double)rho[i];
        s[1] += ex * ex + ey * ey;
        const int ix = (int)(i % Nx);
        s[2] += ex * cosk[ix];
        s[3] += ex * sink[ix];
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double a = vx[i], b = vy[i], qq = q[i];
        s[4] += 0.5 * (double)m[i] * (a * a + b * b);
        s[5] += qq;
        s[6] += fabs(qq);
    }
    __shared__ double red[kStepSums][256];
#pragma unroll
    for (int k = 0; k < kStepSums; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
#pragma unroll
            for (int k = 0; k < kStepSums; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < kStepSums) partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = red[threadIdx.x][0];
}

// Adds the block partials in index order and appends one record at the
// device-side step counter; then advances the counter.
__global__ __launch_bounds__(256) void step_diag_final_kernel(const double* __restrict__ partials, int nblocks,
                                                              double cv, int64_t G, int* __restrict__ counter,
                                                              int record_from, int cap, StepRecords rec) {
    __shared__ double sh[4];
    double tot[kStepSums];
    for (int k = 0; k < kStepSums; ++k) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += partials[(size_t)k * nblocks + b];
        tot[k] = block_sum256(s, sh);
    }
    if (threadIdx.x == 0) {
        const int step = *counter;
        const int slot = step - record_from;
        if (slot >= 0 && slot < cap) {
            const double re = tot[2] / (double)G, im = tot[3] / (double)G;
            const double irho = tot[0] * cv;
            if (rec.field_energy) rec.field_energy[slot] = 0.5 * tot[1] * cv;
            if (rec.mode1) rec.mode1[slot] = sqrt(re * re + im * im);
            if (rec.kinetic) rec.kinetic[slot] = tot[4];
            if (rec.charge_error) rec.charge_error[slot] = fabs(irho - tot[5]) / fmax(tot[6], 1e-12);
        }
        *counter = step + 1;
    }
}

__global__ void set_int_kernel(int* __restrict__ p, int v) { *p = v; }

__global__ __launch_bounds__(256) void copy_f32_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                       int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Streaming read of n float4 values: every thread adds its grid-stride share,
// one partial per block (so the loads cannot be dropped).
__global__ __launch_bounds__(256) void read_probe_kernel(const float4* __restrict__ src, int64_t n,
                                                         float* __restrict__ partial) {
    float s = 0.0f;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float4 v = src[i];
        s += (v.x + v.y) + (v.z + v.w);
    }
    __shared__ float sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// Particle records in the layout of ParticleAoS (src/pic/particles.hpp:13-20):
// x, y, vx, vy, q, m as FP64, 48 bytes each.  Consecutive lanes read
// consecutive records (three 16-byte loads) and write consecutive elements of
// the six FP32 streams, rounded to nearest as a host (float) conversion is.
__global__ __launch_bounds__(256) void aos_unpack_kernel(const double2* __restrict__ rec, int64_t n,
                                                         float* __restrict__ x, float* __restrict__ y,
                                                         float* __restrict__ vx, float* __restrict__ vy,
                                                         float* __restrict__ q, float* __restrict__ m) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 a = rec[3 * i], b = rec[3 * i + 1], c = rec[3 * i + 2];
    x[i] = (float)a.x;
    y[i] = (float)a.y;
    vx[i] = (float)b.x;
    vy[i] = (float)b.y;
    q[i] = (float)c.x;
    m[i] = (float)c.y;
}

// The six FP32 streams back into FP64 records (exact promotion).
__global__ __launch_bounds__(256) void aos_pack_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                       const float* __restrict__ vx, const float* __restrict__ vy,
                                                       const float* __restrict__ q, const float* __restrict__ m,
                                                       int64_t n, double2* __restrict__ rec) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rec[3 * i] = make_double2((double)x[i], (double)y[i]);
    rec[3 * i + 1] = make_double2((double)vx[i], (double)vy[i]);
    rec[3 * i + 2] = make_double2((double)q[i], (double)m[i]);
}

}  // namespace

void launch_read_probe(const float* src, int64_t n4, float* partial, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(read_probe_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                       reinterpret_cast<const float4*>(src), n4, partial);
}

void launch_poisson_prep(const float* rho, float* in, int64_t G, double bg, double cv, const double* rho_sum,
                         double sum_scale, double sum_add, hipStream_t s) {
    hipLaunchKernelGGL(poisson_prep_kernel, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, s, rho, in, G, bg, cv,
                       rho_sum, sum_scale, sum_add);
}

void launch_poisson_spectral(float2* spec, float2* spec_ex, float2* spec_ey, const PoissonTables& t,
                             const DevDomain& d, hipStream_t s) {
    const int hx = d.Nx / 2 + 1;
    hipLaunchKernelGGL(poisson_spectral_kernel, dim3((unsigned)((hx + 255) / 256), (unsigned)d.Ny), dim3(256), 0, s,
                       spec, spec_ex, spec_ey, t, d);
}

void launch_efield(const float* phi, float* Ex, float* Ey, const DevDomain& d, hipStream_t s) {
    hipLaunchKernelGGL(efield_kernel, dim3((unsigned)((d.Nx + 255) / 256), (unsigned)d.Ny), dim3(256), 0, s, phi, Ex,
                       Ey, d);
}

void launch_add_scalar(float* p, int64_t n, float v, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(add_scalar_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

void launch_gather_push(float* x, float* y, float* vx, float* vy, const float* q, const float* m, int64_t n,
                        const float* Ex, const float* Ey, const DevDomain& d, float dt, int do_gather, int do_push,
                        hipStream_t s, const int* push_flag, bool q_one, bool m_one) {
    if (n > 0)
        hipLaunchKernelGGL(gather_push_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, vx, vy, q, m,
                           n, Ex, Ey, d, dt, do_gather, do_push, push_flag, q_one ? 1 : 0, m_one ? 1 : 0);
}

// Partial-sum blocks for a reduction over `items` values: one per 2048 (the
// grid-stride loops then give each thread about eight), at most `cap`.
static int reduction_blocks(int64_t items, int cap) {
    const int64_t nb = (items + 2047) / 2048;
    return (int)std::max<int64_t>(1, std::min<int64_t>(nb, cap));
}

void launch_diagnostics(const float* rho, const float* Ex, const float* Ey, int64_t G, const float* vx,
                        const float* vy, const float* q, const float* m, int64_t n, double* partials, int nblocks,
                        double* out, hipStream_t s) {
    nblocks = reduction_blocks(std::max(G, n), nblocks);
    hipLaunchKernelGGL(diag_partial_kernel, dim3(nblocks), dim3(256), 0, s, rho, Ex, Ey, G, vx, vy, q, m, n,
                       partials);
    hipLaunchKernelGGL(diag_final_kernel, dim3(1), dim3(256), 0, s, partials, nblocks, out);
}

void launch_init(float* x, float* y, float* vx, float* vy, float* q, float* m, int64_t n, const DevDomain& d,
                 const InitParams& p, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(init_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, vx, vy, q, m, n, d,
                           p);
}

void launch_step_diag(const float* rho, const float* Ex, const float* Ey, const double* cosk, const double* sink,
                      const DevDomain& d, const float* vx, const float* vy, const float* q, const float* m, int64_t n,
                      double* partials, int nblocks, int* counter, int record_from, int cap, const StepRecords& rec,
                      hipStream_t s) {
    const int64_t G = (int64_t)d.Nx * d.Ny;
    nblocks = reduction_blocks(std::max(G, n), nblocks);
    hipLaunchKernelGGL(step_diag_partial_kernel, dim3(nblocks), dim3(256), 0, s, rho, Ex, Ey, cosk, sink, d.Nx, G, vx,
                       vy, q, m, n, partials);
    hipLaunchKernelGGL(step_diag_final_kernel, dim3(1), dim3(256), 0, s, partials, nblocks, d.dx * d.dy, G, counter,
                       record_from, cap, rec);
}

void launch_set_int(int* p, int v, hipStream_t s) { hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, p, v); }

void launch_copy_f32(const float* src, float* dst, int64_t n, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(copy_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
}

void launch_aos_unpack(const double* rec, int64_t n, float* const dst[6], hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(aos_unpack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           reinterpret_cast<const double2*>(rec), n, dst[0], dst[1], dst[2], dst[3], dst[4], dst[5]);
}

void launch_aos_pack(const float* const src[6], int64_t n, double* rec, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(aos_pack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src[0], src[1],
                           src[2], src[3], src[4], src[5], n, reinterpret_cast<double2*>(rec));
}

}  // namespace pichip
