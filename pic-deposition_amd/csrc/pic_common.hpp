// Device-side building blocks shared by the deposition, sort and step kernels.
//
// Geometry follows pic::Domain (reference src/pic/domain.hpp:7-44): a periodic
// Lx x Ly box of Nx x Ny cells, cell (ix, iy) has row-major key iy*Nx + ix, a
// particle's cell is floor(x/dx) wrapped into [0, Nx).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pichip {

enum Scheme { NGP = 0, CIC = 1, TSC = 2, ESK = 3 };
// Tile-kernel stencils for the node-keyed order (particles sorted by nearest
// node instead of by cell): the same NGP / TSC weights on a stencil centred
// on that node.  Internal kernel variants, not schemes of the C ABI.
enum NodeScheme { NGP_NODE = 4, TSC_NODE = 5, TSC_MOM = 6 };

struct DevDomain {
    int Nx, Ny;
    double dx, dy;        // FP64 cell sizes: cell keys are computed exactly as the reference does
    double Lx, Ly;
    float inv_dx, inv_dy;  // FP32 reciprocals for shape-function weights
    float inv_vol;         // 1 / (dx*dy) rounded to FP32
    int exact;             // 1: dx and dy are powers of two, so x * inv_dx is exact in FP32
};

// Cell index (unwrapped floor) and in-cell fraction of a coordinate.  With a
// power-of-two cell size the FP32 product is exact; otherwise x / dx is formed
// in FP64 (an FP32 product would carry an error of ~Nx ulps in the fraction).
// EX = 1 / 0 fixes the path at compile time, EX = -1 reads d.exact.
template <int EX = -1>
__device__ __forceinline__ void cell_coord(float x, float inv_d, double d, int exact, int& c, float& f) {
    if (EX == 1 || (EX < 0 && exact)) {
        const float fx = x * inv_d;
        const float fl = floorf(fx);
        c = (int)fl;
        f = fx - fl;
    } else {
        const double fx = (double)x / d;
        const double fl = floor(fx);
        c = (int)fl;
        f = (float)(fx - fl);
    }
}

// A displacement in cell units.
template <int EX = -1>
__device__ __forceinline__ float cells(float disp, float inv_d, double d, int exact) {
    return (EX == 1 || (EX < 0 && exact)) ? disp * inv_d : (float)((double)disp / d);
}

// Reference wrapIndex (src/deposition/ngp.cpp:9): ((i % n) + n) % n, with the
// common in-range case taken without integer division.
__device__ __forceinline__ int wrap_index(int i, int n) {
    if ((unsigned)i < (unsigned)n) return i;
    if (i >= n && i < 2 * n) return i - n;
    if (i < 0 && i >= -n) return i + n;
    int r = i % n;
    return r < 0 ? r + n : r;
}

// x / dx in FP64 as Domain::cellCoords divides (domain.hpp:34-43).  With a
// power-of-two cell size the product with its (exact) reciprocal is the same
// FP64 value (both are the exact quotient, which needs no rounding).
__device__ __forceinline__ double cell_units_exact(float x, double dd, float inv, int exact) {
    return exact ? (double)x * (double)inv : (double)x / dd;
}

// Cell key of a position exactly as particleCellIndex (src/sorting/cell_sort.cpp:8-13)
// via Domain::cellCoords (domain.hpp:34-43): FP64 division, floor, wrap.
__device__ __forceinline__ int cell_key_exact(float x, float y, const DevDomain& d) {
    const double fx = cell_units_exact(x, d.dx, d.inv_dx, d.exact);
    const double fy = cell_units_exact(y, d.dy, d.inv_dy, d.exact);
    const int ix = wrap_index((int)floor(fx), d.Nx);
    const int iy = wrap_index((int)floor(fy), d.Ny);
    return iy * d.Nx + ix;
}

// Sum of one double per thread over a 256-thread block in a fixed order
// (butterfly within each wave, then the four wave totals in order); every
// thread receives the total.  `sh` holds 4 doubles of shared memory.
__device__ __forceinline__ double block_sum256(double v, double* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    const double t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return t;
}

// Knuth TwoSum error term: (a + b) == fl(a + b) + two_sum_err(a, b) exactly.
__device__ __forceinline__ float two_sum_err(float a, float b) {
    const float s = a + b;
    const float bb = s - a;
    return (a - (s - bb)) + (b - bb);
}

// ---------------------------------------------------------------------------
// Shape functions.  Each scheme maps one particle to a home cell (cx, cy) =
// unwrapped floor of the position in cell units, and V weights on a fixed
// SX x SY node stencil whose lower-left corner is (cx + OX, cy + OY).  Slot k
// is node (cx + OX + k % SX, cy + OY + k / SX).  Weights include q / cellVolume.
// ---------------------------------------------------------------------------
// Home cell and in-cell fractions of a particle: the cell every stencil is
// anchored to (weights = fill(fractions), so the window test can run first).
template <int EX = -1>
__device__ __forceinline__ void home_cell(float x, float y, const DevDomain& d, int& cx, int& cy, float& fx,
                                          float& fy) {
    cell_coord<EX>(x, d.inv_dx, d.dx, d.exact, cx, fx);
    cell_coord<EX>(y, d.inv_dy, d.dy, d.exact, cy, fy);
}

// Nearest node of a position (the node-keyed order's home): floor cell plus
// one when the in-cell fraction is at least 1/2 (the NGP decision of
// ngp.cpp:108-120), with the node past the last cell taken as node 0 of the
// periodic grid.  fx, fy stay the floor-cell fractions.
template <int EX = -1>
__device__ __forceinline__ void node_cell(float x, float y, const DevDomain& d, int& nx, int& ny, float& fx,
                                          float& fy) {
    int cx, cy;
    cell_coord<EX>(x, d.inv_dx, d.dx, d.exact, cx, fx);
    cell_coord<EX>(y, d.inv_dy, d.dy, d.exact, cy, fy);
    nx = cx + (fx >= 0.5f ? 1 : 0);
    ny = cy + (fy >= 0.5f ? 1 : 0);
    if (nx == d.Nx && cx == d.Nx - 1) nx = 0;
    if (ny == d.Ny && cy == d.Ny - 1) ny = 0;
}

// Sort key of the node-keyed order: the wrapped nearest node, from the same
// FP64 division as cell_key_exact.
__device__ __forceinline__ int node_key_exact(float x, float y, const DevDomain& d) {
    const double fx = cell_units_exact(x, d.dx, d.inv_dx, d.exact);
    const double fy = cell_units_exact(y, d.dy, d.inv_dy, d.exact);
    const double flx = floor(fx), fly = floor(fy);
    const int ix = wrap_index((int)flx + ((float)(fx - flx) >= 0.5f ? 1 : 0), d.Nx);
    const int iy = wrap_index((int)fly + ((float)(fy - fly) >= 0.5f ? 1 : 0), d.Ny);
    return iy * d.Nx + ix;
}

template <int S>
struct Stencil;

// NGP: src/deposition/ngp.cpp:108-120.  Nearest node is ix or ix+1 per axis
// (wx >= 0.5 picks ix+1), so the stencil is 2x2 with one non-zero slot.
template <>
struct Stencil<NGP> {
    static constexpr int SX = 2, SY = 2, V = 4, OX = 0, OY = 0;
    static constexpr bool kNeedsVelocity = false;
    static constexpr bool kAlwaysOk = true;  // fill never rejects a particle
    template <int EX = -1>
    __device__ __forceinline__ static bool fill(float fx, float fy, float, float, float qv, float, const DevDomain&,
                                                float (&w)[V]) {
        const bool rx = fx >= 0.5f, ry = fy >= 0.5f;
        w[0] = (!rx && !ry) ? qv : 0.0f;
        w[1] = (rx && !ry) ? qv : 0.0f;
        w[2] = (!rx && ry) ? qv : 0.0f;
        w[3] = (rx && ry) ? qv : 0.0f;
        return true;
    }
    // Home cell the window keys and the stencil offsets refer to.
    template <int EX = -1>
    __device__ __forceinline__ static void home(float x, float y, const DevDomain& d, int& cx, int& cy, float& fx,
                                                float& fy) {
        home_cell<EX>(x, y, d, cx, cy, fx, fy);
    }
    template <int EX = -1>
    __device__ __forceinline__ static bool weights(float x, float y, float vx, float vy, float qv, float dt,
                                                   const DevDomain& d, int& cx, int& cy, float (&w)[V]) {
        float fx, fy;
        home_cell<EX>(x, y, d, cx, cy, fx, fy);
        return fill<EX>(fx, fy, vx, vy, qv, dt, d, w);
    }
};

// CIC: src/deposition/ngp.cpp:140-157, same product order q*(1-wx)*(1-wy).
template <>
struct Stencil<CIC> {
    static constexpr int SX = 2, SY = 2, V = 4, OX = 0, OY = 0;
    static constexpr bool kNeedsVelocity = false;
    static constexpr bool kAlwaysOk = true;  // fill never rejects a particle
    template <int EX = -1>
    __device__ __forceinline__ static bool fill(float wx, float wy, float, float, float qv, float, const DevDomain&,
                                                float (&w)[V]) {
        // q(1-wx) and q wx first, then the y factors: 6 operations, each
        // weight within 2 ulp of q*(1-wx)*(1-wy) etc.
        const float a1 = qv * wx, a0 = qv - a1;
        w[2] = a0 * wy;
        w[3] = a1 * wy;
        w[0] = a0 - w[2];
        w[1] = a1 - w[3];
        return true;
    }
    // Home cell the window keys and the stencil offsets refer to.
    template <int EX = -1>
    __device__ __forceinline__ static void home(float x, float y, const DevDomain& d, int& cx, int& cy, float& fx,
                                                float& fy) {
        home_cell<EX>(x, y, d, cx, cy, fx, fy);
    }
    template <int EX = -1>
    __device__ __forceinline__ static bool weights(float x, float y, float vx, float vy, float qv, float dt,
                                                   const DevDomain& d, int& cx, int& cy, float (&w)[V]) {
        float fx, fy;
        home_cell<EX>(x, y, d, cx, cy, fx, fy);
        return fill<EX>(fx, fy, vx, vy, qv, dt, d, w);
    }
};

// TSC: tscWeights (src/deposition/ngp.cpp:28-34) with i0 = floor(fx - 0.5).
// Written in terms of the in-cell fraction f = fx - floor(fx): for f >= 0.5,
// i0 = cx and x' = f; for f < 0.5, i0 = cx - 1 and x' = f + 1.  The quadratic
// B-spline values are the same polynomials re-expanded around f, which keeps
// every subtraction exact in FP32.  Slots cover node offsets -1..2 per axis.
__device__ __forceinline__ void tsc_1d(float f, float (&w)[4]) {
    // f >= 0.5: a = 1.5 - f, b = f - 1, c = f - 0.5 into slots 1..3;
    // f <  0.5: a = 0.5 - f, b = f,     c = f + 0.5 into slots 0..2.
    // One set of operations with per-case constants (the same FP32 operations
    // as either case alone), then the slot placement.
    const bool hi = f >= 0.5f;
    const float a = (hi ? 1.5f : 0.5f) - f;
    const float b = f - (hi ? 1.0f : 0.0f);
    const float c = f - (hi ? 0.5f : -0.5f);
    const float w0 = 0.5f * a * a, w1 = 0.75f - b * b, w2 = 0.5f * c * c;
    w[0] = hi ? 0.0f : w0;
    w[1] = hi ? w0 : w1;
    w[2] = hi ? w1 : w2;
    w[3] = hi ? w2 : 0.0f;
}

template <>
struct Stencil<TSC> {
    static constexpr int SX = 4, SY = 4, V = 16, OX = -1, OY = -1;
    static constexpr bool kNeedsVelocity = false;
    static constexpr bool kAlwaysOk = true;  // fill never rejects a particle
    // Home cell the window keys and the stencil offsets refer to.
    template <int EX = -1>
    __device__ __forceinline__ static void home(float x, float y, const DevDomain& d, int& cx, int& cy, float& fx,
                                                float& fy) {
        home_cell<EX>(x, y, d, cx, cy, fx, fy);
    }
    template <int EX = -1>
    __device__ __forceinline__ static bool weights(float x, float y, float vx, float vy, float qv, float dt,
                                                   const DevDomain& d, int& cx, int& cy, float (&w)[V]) {
        float fx, fy;
        home_cell<EX>(x, y, d, cx, cy, fx, fy);
        return fill<EX>(fx, fy, vx, vy, qv, dt, d, w);
    }
    template <int EX = -1>
    __device__ __forceinline__ static bool fill(float fx, float fy, float, float, float qv, float, const DevDomain&,
                                                float (&w)[V]) {
        float wx[4], wy[4];
        tsc_1d(fx, wx);
        tsc_1d(fy, wy);
#pragma unroll
        for (int a = 0; a < 4; ++a) wx[a] *= qv;
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int a = 0; a < 4; ++a) w[b * 4 + a] = wx[a] * wy[b];
        return true;
    }
};

// Reference "Esirkepov" density (esirkepovWeights1DSparse, src/deposition/ngp.cpp:36-84):
// per axis, path-length fractions of the straight move x0 -> x0 + v*dt over
// the floor cells it crosses, renormalised to sum 1; a static particle
// (|v*dt| < 1e-15) gets the CIC pair.  Everything is measured from the home
// cell's faces (fr0 = in-cell fraction, dc = displacement in cells), so no
// FP32 rounding of x0 + v*dt enters the fractions.  For moves shorter than a
// cell the support is within offsets -1..1 of the home cell.  Returns false
// when a move crosses more than one face (handled by the general walk).
template <int EX = -1>
__device__ __forceinline__ bool esk_1d_from(float fr0, float v, float dt, float inv_d, double dd, int exact,
                                            float (&w)[3]) {
    const float disp = v * dt;
    // The reference tests |x1 - x0| < 1e-15 in FP64; the displacement itself
    // is the same quantity without the cancellation.  A static particle gets
    // the CIC pair.
    const bool still = fabsf(disp) < 1e-15f;
    // A moving particle: wa = path inside the home cell up to the face it
    // heads for; a move past that face adds wb in the neighbour cell, and the
    // pair is renormalised.  Both directions and both outcomes are evaluated
    // with the same FP32 operations as the case alone and then selected.
    const float dc = cells<EX>(disp, inv_d, dd, exact);
    const bool up = dc > 0.0f;
    const float adc = up ? dc : -dc;
    const float wa = up ? 1.0f - fr0 : fr0;
    const bool cross = !(adc <= wa);
    const float wb = adc - wa;
    // Renormalisation by the hardware reciprocal (v_rcp_f32, about 1 ulp): the
    // pair sums to 1 within a few 1e-8 instead of the division's 0.5 ulp.
    const float inv = __builtin_amdgcn_rcpf(wa + wb);
    const float wh = cross ? wa * inv : 1.0f;
    const float wn = cross ? wb * inv : 0.0f;
    w[0] = still || up ? 0.0f : wn;
    w[1] = still ? 1.0f - fr0 : wh;
    w[2] = still ? fr0 : (up ? wn : 0.0f);
    return still || !(cross && wb > 1.0f);
}

template <>
struct Stencil<ESK> {
    static constexpr int SX = 3, SY = 3, V = 9, OX = -1, OY = -1;
    static constexpr bool kNeedsVelocity = true;
    static constexpr bool kAlwaysOk = false;  // moves across more than one face take the walk
    // Home cell the window keys and the stencil offsets refer to.
    template <int EX = -1>
    __device__ __forceinline__ static void home(float x, float y, const DevDomain& d, int& cx, int& cy, float& fx,
                                                float& fy) {
        home_cell<EX>(x, y, d, cx, cy, fx, fy);
    }
    template <int EX = -1>
    __device__ __forceinline__ static bool weights(float x, float y, float vx, float vy, float qv, float dt,
                                                   const DevDomain& d, int& cx, int& cy, float (&w)[V]) {
        float fx, fy;
        home_cell<EX>(x, y, d, cx, cy, fx, fy);
        return fill<EX>(fx, fy, vx, vy, qv, dt, d, w);
    }
    template <int EX = -1>
    __device__ __forceinline__ static bool fill(float fx, float fy, float vx, float vy, float qv, float dt,
                                                const DevDomain& d, float (&w)[V]) {
        float wx[3], wy[3];
        const bool okx = esk_1d_from<EX>(fx, vx, dt, d.inv_dx, d.dx, d.exact, wx);
        const bool oky = esk_1d_from<EX>(fy, vy, dt, d.inv_dy, d.dy, d.exact, wy);
#pragma unroll
        for (int a = 0; a < 3; ++a) wx[a] *= qv;
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int a = 0; a < 3; ++a) w[b * 3 + a] = wx[a] * wy[b];
        return okx && oky;
    }
};

// NGP on the node-keyed order: the home is the nearest node itself.
template <>
struct Stencil<NGP_NODE> {
    static constexpr int SX = 1, SY = 1, V = 1, OX = 0, OY = 0;
    static constexpr bool kNeedsVelocity = false;
    static constexpr bool kAlwaysOk = true;
    template <int EX = -1>
    __device__ __forceinline__ static void home(float x, float y, const DevDomain& d, int& cx, int& cy, float& fx,
                                                float& fy) {
        node_cell<EX>(x, y, d, cx, cy, fx, fy);
    }
    template <int EX = -1>
    __device__ __forceinline__ static bool fill(float, float, float, float, float qv, float, const DevDomain&,
                                                float (&w)[V]) {
        w[0] = qv;
        return true;
    }
};

// TSC on the node-keyed order: nodes n-1, n, n+1 around the nearest node n.
// With f the floor-cell fraction these are the three non-zero weights of
// tsc_1d (the same FP32 operations), without the fourth, empty slot.
__device__ __forceinline__ void tsc_node_1d(float f, float (&w)[3]) {
    const bool hi = f >= 0.5f;
    const float a = (hi ? 1.5f : 0.5f) - f;
    const float b = f - (hi ? 1.0f : 0.0f);
    const float c = f - (hi ? 0.5f : -0.5f);
    w[0] = 0.5f * a * a;
    w[1] = 0.75f - b * b;
    w[2] = 0.5f * c * c;
}

template <>
struct Stencil<TSC_NODE> {
    static constexpr int SX = 3, SY = 3, V = 9, OX = -1, OY = -1;
    static constexpr bool kNeedsVelocity = false;
    static constexpr bool kAlwaysOk = true;
    template <int EX = -1>
    __device__ __forceinline__ static void home(float x, float y, const DevDomain& d, int& cx, int& cy, float& fx,
                                                float& fy) {
        node_cell<EX>(x, y, d, cx, cy, fx, fy);
    }
    template <int EX = -1>
    __device__ __forceinline__ static bool fill(float fx, float fy, float, float, float qv, float, const DevDomain&,
                                                float (&w)[V]) {
        float wx[3], wy[3];
        tsc_node_1d(fx, wx);
        tsc_node_1d(fy, wy);
#pragma unroll
        for (int a = 0; a < 3; ++a) wx[a] *= qv;
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int a = 0; a < 3; ++a) w[b * 3 + a] = wx[a] * wy[b];
        return true;
    }
};

// TSC on the node-keyed order carried as charge moments.  Slot i + 3 j holds
// q d_x^i d_y^j, d the particle's offset from its home node in cells (the
// `b` of tsc_node_1d, in [-1/2, 1/2)).  The node weights are polynomials in d
// (tsc_node_1d: node h-1 0.5 (1/2 - d)^2, node h 3/4 - d^2, node h+1
// 0.5 (1/2 + d)^2), so the 3 x 3 node values of any sum of particles of one
// home node follow linearly from its nine moment sums.  The tile kernel adds
// moments per home node and maps them to node values once per tile: fewer
// operations per particle than forming the weights.
template <>
struct Stencil<TSC_MOM> {
    static constexpr int SX = 3, SY = 3, V = 9, OX = -1, OY = -1;
    static constexpr bool kNeedsVelocity = false;
    static constexpr bool kAlwaysOk = true;
    template <int EX = -1>
    __device__ __forceinline__ static void home(float x, float y, const DevDomain& d, int& cx, int& cy, float& fx,
                                                float& fy) {
        node_cell<EX>(x, y, d, cx, cy, fx, fy);
    }
    template <int EX = -1>
    __device__ __forceinline__ static bool fill(float fx, float fy, float, float, float qv, float, const DevDomain&,
                                                float (&w)[V]) {
        const float dx = fx - (fx >= 0.5f ? 1.0f : 0.0f);
        const float dy = fy - (fy >= 0.5f ? 1.0f : 0.0f);
        const float m10 = qv * dx, m01 = qv * dy;
        const float m20 = m10 * dx, m11 = m10 * dy, m21 = m20 * dy;
        w[0] = qv;
        w[1] = m10;
        w[2] = m20;
        w[3] = m01;
        w[4] = m11;
        w[5] = m21;
        w[6] = m01 * dy;
        w[7] = m11 * dy;
        w[8] = m21 * dy;
        return true;
    }
};


}  // namespace pichip
