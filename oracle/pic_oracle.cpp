// FP64 CPU oracle for the charge-deposition hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: the parity tests, the
// smoke check and the CPU-baseline leg of the benchmark call it; the HIP product
// library never links or calls it.
//
// It restates, in double precision and in the reference's arithmetic order, the
// algorithms of ehocine/pic-deposition listed below (paths are relative to the
// reference repository root).  Every exported function names the lines it
// follows.  Deliberate deviations are marked DEVIATION.
//
// Pinning (which independent check fixes each function) is listed in the header
// comment of oracle.py next to the Python binding.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include <omp.h>

extern "C" {

// Mirrors pic::Domain (src/pic/domain.hpp:7-44).  dx, dy are derived exactly as
// Domain::updateDerived (domain.hpp:17-20): dx = Lx / Nx in double.
typedef struct {
    double Lx, Ly;
    int Nx, Ny;
} or_domain;

}  // extern "C"

namespace {

struct Dom {
    double Lx, Ly, dx, dy;
    int Nx, Ny;
    explicit Dom(const or_domain* d)
        : Lx(d->Lx), Ly(d->Ly), dx(d->Lx / d->Nx), dy(d->Ly / d->Ny), Nx(d->Nx), Ny(d->Ny) {}
    double cellVolume() const { return dx * dy; }          // domain.hpp:22
    double domainVolume() const { return Lx * Ly; }        // domain.hpp:23
    int cellIndex(int ix, int iy) const { return iy * Nx + ix; }  // domain.hpp:25
    // domain.hpp:27-32
    void wrapPosition(double& x, double& y) const {
        x = std::fmod(x, Lx);
        if (x < 0.0) x += Lx;
        y = std::fmod(y, Ly);
        if (y < 0.0) y += Ly;
    }
    // domain.hpp:34-43
    void cellCoords(double x, double y, int& ix, int& iy, double& wx, double& wy) const {
        wx = x / dx;
        wy = y / dy;
        ix = static_cast<int>(std::floor(wx));
        iy = static_cast<int>(std::floor(wy));
        wx -= ix;
        wy -= iy;
        ix = (ix % Nx + Nx) % Nx;
        iy = (iy % Ny + Ny) % Ny;
    }
};

// src/deposition/ngp.cpp:9
inline int wrapIndex(int i, int n) { return (i % n + n) % n; }

// src/deposition/ngp.cpp:28-34
void tscWeights(double xi, double w[3], int& i0) {
    i0 = static_cast<int>(std::floor(xi - 0.5));
    const double x = xi - static_cast<double>(i0);
    w[0] = 0.5 * (1.5 - x) * (1.5 - x);
    w[1] = 0.75 - (x - 1.0) * (x - 1.0);
    w[2] = 0.5 * (x - 0.5) * (x - 0.5);
}

struct EW {
    int index;
    double weight;
};

// src/deposition/ngp.cpp:36-84 (trajectory path fractions on floor cells,
// renormalised to a partition of unity; exact CIC pair when |x1-x0| < 1e-15).
void esirkepovWeights1DSparse(double x0, double x1, double dx, int n, std::vector<EW>& weights) {
    weights.clear();
    if (std::abs(x1 - x0) < 1e-15) {
        int ix = static_cast<int>(std::floor(x0 / dx));
        const double xi = x0 / dx - static_cast<double>(ix);
        weights.push_back({wrapIndex(ix, n), 1.0 - xi});
        weights.push_back({wrapIndex(ix + 1, n), xi});
        return;
    }
    const double inv_dx = 1.0 / dx;
    double x_curr = x0;
    const double x_end = x1;
    int cell = static_cast<int>(std::floor(x_curr * inv_dx));
    if (x_end > x0) {
        while (x_curr < x_end - 1e-15) {
            const double x_next = std::min(x_end, (cell + 1) * dx);
            const double w = (x_next - x_curr) * inv_dx;
            weights.push_back({wrapIndex(cell, n), w});
            x_curr = x_next;
            cell++;
        }
    } else {
        while (x_curr > x_end + 1e-15) {
            const double x_next = std::max(x_end, cell * dx);
            const double w = (x_curr - x_next) * inv_dx;
            weights.push_back({wrapIndex(cell, n), w});
            x_curr = x_next;
            cell--;
        }
    }
    double total = 0.0;
    for (const auto& e : weights) total += e.weight;
    if (total > 1e-300) {
        const double inv_total = 1.0 / total;
        for (auto& e : weights) e.weight *= inv_total;
    }
}

// Per-particle scatter of each scheme into `rho`, in the serial-path order of
// src/deposition/ngp.cpp.  `inv_vol_mul` selects between the serial NGP form
// (q / cellVolume, ngp.cpp:118) and the OpenMP form (q * inv_vol,
// deposition_parallel.cpp:63).
inline void depositNGPOne(const Dom& d, double x, double y, double q, double* rho, bool inv_vol_mul) {
    int ix, iy;
    double wx, wy;
    d.cellCoords(x, y, ix, iy, wx, wy);
    const int jx = wx >= 0.5 ? wrapIndex(ix + 1, d.Nx) : ix;
    const int jy = wy >= 0.5 ? wrapIndex(iy + 1, d.Ny) : iy;
    const double v = inv_vol_mul ? q * (1.0 / d.cellVolume()) : q / d.cellVolume();
    rho[static_cast<size_t>(d.cellIndex(jx, jy))] += v;
}

// src/deposition/ngp.cpp:140-157
inline void depositCICOne(const Dom& d, double inv_vol, double x, double y, double qp, double* rho) {
    int ix, iy;
    double wx, wy;
    d.cellCoords(x, y, ix, iy, wx, wy);
    const double q = qp * inv_vol;
    const int ixp = wrapIndex(ix + 1, d.Nx);
    const int iyp = wrapIndex(iy + 1, d.Ny);
    rho[static_cast<size_t>(d.cellIndex(ix, iy))] += q * (1.0 - wx) * (1.0 - wy);
    rho[static_cast<size_t>(d.cellIndex(ixp, iy))] += q * wx * (1.0 - wy);
    rho[static_cast<size_t>(d.cellIndex(ix, iyp))] += q * (1.0 - wx) * wy;
    rho[static_cast<size_t>(d.cellIndex(ixp, iyp))] += q * wx * wy;
}

// src/deposition/ngp.cpp:181-202 (serial-path periodic wrap)
inline void depositTSCOne(const Dom& d, double inv_vol, double x, double y, double qp, double* rho) {
    const double fx = x / d.dx;
    const double fy = y / d.dy;
    int ix0, iy0;
    double wx[3], wy[3];
    tscWeights(fx, wx, ix0);
    tscWeights(fy, wy, iy0);
    const double q = qp * inv_vol;
    for (int dy = 0; dy < 3; ++dy) {
        const int jy = wrapIndex(iy0 + dy, d.Ny);
        for (int dx = 0; dx < 3; ++dx) {
            const int jx = wrapIndex(ix0 + dx, d.Nx);
            rho[static_cast<size_t>(d.cellIndex(jx, jy))] += q * wx[dx] * wy[dy];
        }
    }
}

// src/deposition/ngp.cpp:13-24
inline void depositEsirkepovOne(const Dom& d, double inv_vol, double x, double y, double vx, double vy, double q,
                                double dt, double* rho, std::vector<EW>& xs, std::vector<EW>& ys) {
    esirkepovWeights1DSparse(x, x + vx * dt, d.dx, d.Nx, xs);
    esirkepovWeights1DSparse(y, y + vy * dt, d.dy, d.Ny, ys);
    for (const auto& sx : xs)
        for (const auto& sy : ys)
            rho[static_cast<size_t>(d.cellIndex(sx.index, sy.index))] += q * sx.weight * sy.weight * inv_vol;
}

void depositRange(const Dom& d, int scheme, const double* x, const double* y, const double* vx, const double* vy,
                  const double* q, int64_t begin, int64_t end, double dt, double* rho, bool parallel_forms) {
    const double inv_vol = 1.0 / d.cellVolume();
    std::vector<EW> xs, ys;
    xs.reserve(8);
    ys.reserve(8);
    for (int64_t i = begin; i < end; ++i) {
        switch (scheme) {
            case 0: depositNGPOne(d, x[i], y[i], q[i], rho, parallel_forms); break;
            case 1: depositCICOne(d, inv_vol, x[i], y[i], q[i], rho); break;
            case 2: depositTSCOne(d, inv_vol, x[i], y[i], q[i], rho); break;
            default:
                // deposition_parallel.cpp:552-557 multiplies q * wx * wy * inv_vol with q
                // unscaled, the same expression as ngp.cpp:20-21.
                depositEsirkepovOne(d, inv_vol, x[i], y[i], vx ? vx[i] : 0.0, vy ? vy[i] : 0.0, q[i], dt, rho, xs,
                                    ys);
                break;
        }
    }
}


}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Deposition
// ---------------------------------------------------------------------------

// Serial deposition, src/deposition/ngp.cpp:108-228 (SoA variants), with
// FieldGrid::clearRho first (field_grid.cpp:22, deposition.cpp:121).
// scheme: 0 NGP, 1 CIC, 2 TSC, 3 Esirkepov (density variant, needs vx, vy, dt).
int or_deposit(int scheme, const double* x, const double* y, const double* vx, const double* vy, const double* q,
               int64_t n, const or_domain* dom, double dt, double* rho) {
    if (scheme < 0 || scheme > 3 || !dom || dom->Nx <= 0 || dom->Ny <= 0) return 1;
    Dom d(dom);
    std::fill(rho, rho + static_cast<size_t>(d.Nx) * d.Ny, 0.0);
    depositRange(d, scheme, x, y, vx, vy, q, 0, n, dt, rho, false);
    return 0;
}

// OpenMP privatised deposition, src/deposition/deposition_parallel.cpp:18-47:
// static contiguous chunk per thread, private full-grid rho_local, merge under
// a critical section.  DEVIATION: TSC and NGP use the serial-path periodic
// wrap (ngp.cpp:9) because deposition_parallel.cpp:108-111 indexes rho_local[-1]
// for x < 0.5*dx.  The merge is done in thread order (deterministic) instead of
// critical-section arrival order.
int or_deposit_omp(int scheme, const double* x, const double* y, const double* vx, const double* vy,
                   const double* q, int64_t n, const or_domain* dom, double dt, int num_threads, double* rho) {
    if (scheme < 0 || scheme > 3 || !dom || dom->Nx <= 0 || dom->Ny <= 0) return 1;
    Dom d(dom);
    const size_t ncells = static_cast<size_t>(d.Nx) * d.Ny;
    std::fill(rho, rho + ncells, 0.0);
    if (num_threads <= 1) {
        depositRange(d, scheme, x, y, vx, vy, q, 0, n, dt, rho, true);
        return 0;
    }
    std::vector<std::vector<double>> locals(static_cast<size_t>(num_threads));
#pragma omp parallel num_threads(num_threads)
    {
        const int tid = omp_get_thread_num();
        const int nth = omp_get_num_threads();
        const int64_t chunk = (n + nth - 1) / nth;
        const int64_t begin = static_cast<int64_t>(tid) * chunk;
        const int64_t end = std::min<int64_t>(begin + chunk, n);
        std::vector<double>& local = locals[static_cast<size_t>(tid)];
        local.assign(ncells, 0.0);
        if (begin < end) depositRange(d, scheme, x, y, vx, vy, q, begin, end, dt, local.data(), true);
    }
    for (int t = 0; t < num_threads; ++t) {
        const std::vector<double>& local = locals[static_cast<size_t>(t)];
        if (local.empty()) continue;
        for (size_t i = 0; i < ncells; ++i) rho[i] += local[i];
    }
    return 0;
}

// Timed OpenMP deposition in the manner of depositChargeTimedSoA
// (deposition.cpp:257-310): 3 warm-up deposits, then `repeats` timed ones.
// Returns mean milliseconds; *std_ms receives the sample standard deviation.
double or_deposit_omp_timed(int scheme, const double* x, const double* y, const double* vx, const double* vy,
                            const double* q, int64_t n, const or_domain* dom, double dt, int num_threads,
                            int repeats, double* rho, double* std_ms) {
    for (int i = 0; i < 3; ++i) or_deposit_omp(scheme, x, y, vx, vy, q, n, dom, dt, num_threads, rho);
    std::vector<double> samples;
    for (int i = 0; i < repeats; ++i) {
        const double t0 = omp_get_wtime();
        or_deposit_omp(scheme, x, y, vx, vy, q, n, dom, dt, num_threads, rho);
        samples.push_back((omp_get_wtime() - t0) * 1000.0);
    }
    double sum = 0.0;
    for (double v : samples) sum += v;
    const double mean = samples.empty() ? 0.0 : sum / samples.size();
    double var = 0.0;
    for (double v : samples) var += (v - mean) * (v - mean);
    if (std_ms) *std_ms = samples.size() > 1 ? std::sqrt(var / (samples.size() - 1)) : 0.0;
    return mean;
}

// ---------------------------------------------------------------------------
// Cell sort (src/sorting/cell_sort.cpp)
// ---------------------------------------------------------------------------

// particleCellIndex, cell_sort.cpp:8-13
void or_cell_keys(const double* x, const double* y, int64_t n, const or_domain* dom, int32_t* keys) {
    Dom d(dom);
    for (int64_t i = 0; i < n; ++i) {
        int ix, iy;
        double wx, wy;
        d.cellCoords(x[i], y[i], ix, iy, wx, wy);
        keys[i] = d.cellIndex(ix, iy);
    }
}

// Permutation that sorts particles by cell key, cell_sort.cpp:22-46.
// DEVIATION: stable (std::stable_sort) because the reference's std::sort tie
// order is an introsort artefact.  perm[i] = source index of the i-th sorted particle.
// Non-negative keys (cell and node indices) take a counting sort: per-key
// counts, their exclusive prefix, then particles placed in index order, which
// is the order std::stable_sort gives.  Other keys fall back to std::stable_sort.
void or_sort_perm(const int32_t* keys, int64_t n, int32_t* perm) {
    int32_t lo = 0, hi = -1;
    for (int64_t i = 0; i < n; ++i) {
        lo = std::min(lo, keys[i]);
        hi = std::max(hi, keys[i]);
    }
    if (lo < 0) {
        std::vector<int32_t> order(static_cast<size_t>(n));
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return keys[a] < keys[b]; });
        std::copy(order.begin(), order.end(), perm);
        return;
    }
    std::vector<int64_t> next(static_cast<size_t>(hi) + 2, 0);
    for (int64_t i = 0; i < n; ++i) next[static_cast<size_t>(keys[i]) + 1]++;
    for (size_t k = 1; k < next.size(); ++k) next[k] += next[k - 1];
    for (int64_t i = 0; i < n; ++i) perm[next[static_cast<size_t>(keys[i])]++] = static_cast<int32_t>(i);
}

// Exclusive-prefix cell offsets of a key array: off[c] = number of keys < c,
// off[G] = n.  (Build-defined table; the reference has none.)
void or_cell_offsets(const int32_t* keys, int64_t n, int ncells, int32_t* off) {
    std::vector<int64_t> cnt(static_cast<size_t>(ncells) + 1, 0);
    for (int64_t i = 0; i < n; ++i) cnt[static_cast<size_t>(keys[i]) + 1]++;
    for (int c = 0; c < ncells; ++c) cnt[static_cast<size_t>(c) + 1] += cnt[static_cast<size_t>(c)];
    for (int c = 0; c <= ncells; ++c) off[c] = static_cast<int32_t>(cnt[static_cast<size_t>(c)]);
}

// ---------------------------------------------------------------------------
// Gather and push (src/deposition/deposition.cpp:335-374)
// ---------------------------------------------------------------------------

// gatherFieldsSoA, deposition.cpp:335-356: CIC interpolation of Ex, Ey at
// cellCoords, v += -(q/m) E dt.
void or_gather(const double* x, const double* y, double* vx, double* vy, const double* q, const double* m, int64_t n,
               const double* Ex, const double* Ey, const or_domain* dom, double dt) {
    Dom d(dom);
    for (int64_t i = 0; i < n; ++i) {
        int ix, iy;
        double wx, wy;
        d.cellCoords(x[i], y[i], ix, iy, wx, wy);
        const int ixp = (ix + 1) % d.Nx;
        const int iyp = (iy + 1) % d.Ny;
        const size_t i00 = static_cast<size_t>(d.cellIndex(ix, iy));
        const size_t i10 = static_cast<size_t>(d.cellIndex(ixp, iy));
        const size_t i01 = static_cast<size_t>(d.cellIndex(ix, iyp));
        const size_t i11 = static_cast<size_t>(d.cellIndex(ixp, iyp));
        const double ex = Ex[i00] * (1.0 - wx) * (1.0 - wy) + Ex[i10] * wx * (1.0 - wy) +
                          Ex[i01] * (1.0 - wx) * wy + Ex[i11] * wx * wy;
        const double ey = Ey[i00] * (1.0 - wx) * (1.0 - wy) + Ey[i10] * wx * (1.0 - wy) +
                          Ey[i01] * (1.0 - wx) * wy + Ey[i11] * wx * wy;
        vx[i] += -q[i] / m[i] * ex * dt;
        vy[i] += -q[i] / m[i] * ey * dt;
    }
}

// pushParticlesSoA, deposition.cpp:367-374: drift then fmod wrap.
void or_push(double* x, double* y, const double* vx, const double* vy, int64_t n, const or_domain* dom, double dt) {
    Dom d(dom);
    for (int64_t i = 0; i < n; ++i) {
        x[i] += vx[i] * dt;
        y[i] += vy[i] * dt;
        d.wrapPosition(x[i], y[i]);
    }
}

// ---------------------------------------------------------------------------
// Reductions and diagnostics (field_grid.cpp:24-38, diagnostics.cpp:10-51,
// particles_aos.cpp:182-196)
// ---------------------------------------------------------------------------

double or_integrated_rho(const double* rho, const or_domain* dom) {
    Dom d(dom);
    double sum = 0.0;
    const size_t nc = static_cast<size_t>(d.Nx) * d.Ny;
    for (size_t i = 0; i < nc; ++i) sum += rho[i];
    return sum * d.cellVolume();
}

double or_field_energy(const double* Ex, const double* Ey, const or_domain* dom) {
    Dom d(dom);
    double sum = 0.0;
    const size_t nc = static_cast<size_t>(d.Nx) * d.Ny;
    for (size_t i = 0; i < nc; ++i) sum += Ex[i] * Ex[i] + Ey[i] * Ey[i];
    return 0.5 * sum * d.cellVolume();
}

double or_kinetic_energy(const double* vx, const double* vy, const double* m, int64_t n) {
    double sum = 0.0;
    for (int64_t i = 0; i < n; ++i) sum += 0.5 * m[i] * (vx[i] * vx[i] + vy[i] * vy[i]);
    return sum;
}

double or_total_charge(const double* q, int64_t n) {
    double sum = 0.0;
    for (int64_t i = 0; i < n; ++i) sum += q[i];
    return sum;
}

double or_total_abs_charge(const double* q, int64_t n) {
    double sum = 0.0;
    for (int64_t i = 0; i < n; ++i) sum += std::abs(q[i]);
    return sum;
}

// diagnostics.cpp:26-29
double or_charge_error(double integrated_rho, double total_charge, double abs_charge_sum) {
    const double denom = std::max(abs_charge_sum, 1e-12);
    return std::abs(integrated_rho - total_charge) / denom;
}

// ---------------------------------------------------------------------------
// Initialisers (src/pic/particles_aos.cpp), libstdc++ <random> draws
// ---------------------------------------------------------------------------

// particles_aos.cpp:67-81
void or_init_uniform_maxwellian(int64_t n, const or_domain* dom, double T, unsigned seed, double* x, double* y,
                                double* vx, double* vy, double* q, double* m) {
    Dom d(dom);
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> ux(0.0, d.Lx);
    std::uniform_real_distribution<double> uy(0.0, d.Ly);
    const double stdv = std::sqrt(T);
    for (int64_t i = 0; i < n; ++i) {
        x[i] = ux(rng);
        y[i] = uy(rng);
        {
            std::normal_distribution<double> dist(0.0, stdv);  // sampleNormal, particles_aos.cpp:11-14
            vx[i] = dist(rng);
        }
        {
            std::normal_distribution<double> dist(0.0, stdv);
            vy[i] = dist(rng);
        }
        m[i] = 1.0;
        q[i] = (i % 2 == 0) ? 1.0 : -1.0;
    }
}

// particles_aos.cpp:83-96 (deterministic lattice; the seed is unused)
void or_init_langmuir(int64_t n, const or_domain* dom, double amplitude, int mode, double* x, double* y, double* vx,
                      double* vy, double* q, double* m) {
    Dom d(dom);
    const double k = 2.0 * M_PI * static_cast<double>(mode) / d.Lx;
    const double qp = -1.0 / static_cast<double>(n);
    for (int64_t i = 0; i < n; ++i) {
        x[i] = d.Lx * static_cast<double>(i) / static_cast<double>(n);
        y[i] = 0.5 * d.Ly;
        vx[i] = amplitude * std::sin(k * x[i]);
        vy[i] = 0.0;
        m[i] = 1.0;
        q[i] = qp;
    }
}

// particles_aos.cpp:142-160 (quasi1d != 0: particles_aos.cpp:162-180)
void or_init_two_stream(int64_t n, const or_domain* dom, double beam_velocity, double perturbation, unsigned seed,
                        int quasi1d, double* x, double* y, double* vx, double* vy, double* q, double* m) {
    Dom d(dom);
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> ux(0.0, d.Lx);
    std::uniform_real_distribution<double> uy(0.0, d.Ly);
    const double qp = -1.0 / static_cast<double>(n);
    const double k = 2.0 * M_PI / d.Lx;
    for (int64_t i = 0; i < n; ++i) {
        x[i] = ux(rng);
        y[i] = quasi1d ? 0.5 * d.Ly : uy(rng);
        const double sign = (i < n / 2) ? 1.0 : -1.0;
        const double pert = perturbation * std::sin(k * x[i]);
        vx[i] = sign * beam_velocity + pert;
        vy[i] = 0.0;
        m[i] = 1.0;
        q[i] = qp;
    }
}

// Raw mt19937 words, for pinning the Philox/mt19937 helpers in tests.
void or_mt19937_words(unsigned seed, int64_t n, uint32_t* out) {
    std::mt19937 rng(seed);
    for (int64_t i = 0; i < n; ++i) out[i] = static_cast<uint32_t>(rng());
}

// ---------------------------------------------------------------------------
// FFT Poisson (src/pic/poisson_fft.cpp:10-94)
// ---------------------------------------------------------------------------
// FFTW is absent here.  The reference's separable R2HC/HC2R solve divides
// every half-complex coefficient by k^2 with |k| taken from its frequency,
// which in exact arithmetic is the complex spectral solve
//   phi_hat(k) = cv * rho_hat(k) / |k|^2,  phi_hat(0) = 0,
// followed by E = -forward difference of phi.  This oracle computes that with
// its own FP64 radix-2 complex FFT (power-of-two grids) or a direct DFT.

// Twiddles w_k = exp(-+2 pi i k / n), k < n, evaluated directly (no recurrence).
static std::vector<std::complex<double>> twiddles(size_t n, bool inverse) {
    std::vector<std::complex<double>> w(n);
    for (size_t k = 0; k < n; ++k) {
        const double ang = 2.0 * M_PI * static_cast<double>(k) / static_cast<double>(n) * (inverse ? 1.0 : -1.0);
        w[k] = std::complex<double>(std::cos(ang), std::sin(ang));
    }
    return w;
}

static void fft1d(std::vector<std::complex<double>>& a, const std::vector<std::complex<double>>& tw) {
    const size_t n = a.size();
    if (n <= 1) return;
    if ((n & (n - 1)) == 0) {
        for (size_t i = 1, j = 0; i < n; ++i) {
            size_t bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) std::swap(a[i], a[j]);
        }
        for (size_t len = 2; len <= n; len <<= 1) {
            const size_t step = n / len;
            for (size_t i = 0; i < n; i += len) {
                for (size_t k = 0; k < len / 2; ++k) {
                    const std::complex<double> u = a[i + k];
                    const std::complex<double> v = a[i + k + len / 2] * tw[k * step];
                    a[i + k] = u + v;
                    a[i + k + len / 2] = u - v;
                }
            }
        }
        return;
    }
    std::vector<std::complex<double>> out(n);
    for (size_t k = 0; k < n; ++k) {
        std::complex<double> s(0.0, 0.0);
        for (size_t t = 0; t < n; ++t) s += a[t] * tw[(k * t) % n];
        out[k] = s;
    }
    a.swap(out);
}

static void fft2d(std::vector<std::complex<double>>& a, int nx, int ny, bool inverse) {
    const auto twx = twiddles(static_cast<size_t>(nx), inverse);
    const auto twy = twiddles(static_cast<size_t>(ny), inverse);
    std::vector<std::complex<double>> row(static_cast<size_t>(nx)), col(static_cast<size_t>(ny));
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) row[static_cast<size_t>(i)] = a[static_cast<size_t>(j) * nx + i];
        fft1d(row, twx);
        for (int i = 0; i < nx; ++i) a[static_cast<size_t>(j) * nx + i] = row[static_cast<size_t>(i)];
    }
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) col[static_cast<size_t>(j)] = a[static_cast<size_t>(j) * nx + i];
        fft1d(col, twy);
        for (int j = 0; j < ny; ++j) a[static_cast<size_t>(j) * nx + i] = col[static_cast<size_t>(j)];
    }
}

// PoissonSolverFFT::solve, poisson_fft.cpp:45-94 (k^2 table :23-35).
void or_poisson(const double* rho, const or_domain* dom, double* phi, double* Ex, double* Ey) {
    Dom d(dom);
    const int nx = d.Nx, ny = d.Ny;
    const size_t n = static_cast<size_t>(nx) * ny;
    const double cell_vol = d.cellVolume();
    std::vector<std::complex<double>> a(n);
    for (size_t i = 0; i < n; ++i) a[i] = std::complex<double>(rho[i] * cell_vol, 0.0);
    fft2d(a, nx, ny, false);
    a[0] = 0.0;
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            if (i == 0 && j == 0) continue;
            const double kx = 2.0 * M_PI * ((i <= nx / 2) ? i : i - nx) / d.Lx;
            const double ky = 2.0 * M_PI * ((j <= ny / 2) ? j : j - ny) / d.Ly;
            a[static_cast<size_t>(j) * nx + i] /= (kx * kx + ky * ky);
        }
    }
    fft2d(a, nx, ny, true);
    const double norm = 1.0 / static_cast<double>(n);
    for (size_t i = 0; i < n; ++i) phi[i] = a[i].real() * norm;
    const double idx_dx = 1.0 / d.dx;
    const double idx_dy = 1.0 / d.dy;
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            const int ip = (i + 1) % nx;
            const int jp = (j + 1) % ny;
            const size_t id = static_cast<size_t>(j) * nx + i;
            Ex[id] = -((phi[static_cast<size_t>(j) * nx + ip] - phi[id]) * idx_dx);
            Ey[id] = -((phi[static_cast<size_t>(jp) * nx + i] - phi[id]) * idx_dy);
        }
    }
}

// ---------------------------------------------------------------------------
// Physics estimators (src/pic/validation.cpp)
// ---------------------------------------------------------------------------

// validation.cpp:19-54
double or_dominant_frequency(const double* signal, int64_t n, double dt) {
    if (n < 8) return 0.0;
    double mean = 0.0;
    for (int64_t i = 0; i < n; ++i) mean += signal[i];
    mean /= static_cast<double>(n);
    std::vector<double> c(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) c[static_cast<size_t>(i)] = signal[i] - mean;
    const int nn = static_cast<int>(n);
    double best_power = 0.0;
    int best_k = 1;
    for (int k = 1; k < nn / 2; ++k) {
        double re = 0.0, im = 0.0;
        for (int t = 0; t < nn; ++t) {
            const double phase = 2.0 * M_PI * static_cast<double>(k * t) / static_cast<double>(nn);
            re += c[static_cast<size_t>(t)] * std::cos(phase);
            im -= c[static_cast<size_t>(t)] * std::sin(phase);
        }
        const double power = re * re + im * im;
        if (power > best_power) {
            best_power = power;
            best_k = k;
        }
    }
    return static_cast<double>(best_k) / (static_cast<double>(nn) * dt);
}

// validation.cpp:66-71
double or_plasma_frequency(const or_domain* dom, int64_t num_particles) {
    Dom d(dom);
    const double n0 = static_cast<double>(num_particles) / d.domainVolume();
    const double q_macro = 1.0 / static_cast<double>(num_particles);
    return std::sqrt(n0 * q_macro * q_macro);
}

// validation.cpp:78-85
double or_resonant_beam_velocity(const or_domain* dom, int64_t num_particles, int mode) {
    Dom d(dom);
    const double omega_p = or_plasma_frequency(dom, num_particles);
    const double k = 2.0 * M_PI * static_cast<double>(mode) / d.Lx;
    if (k < 1e-30) return 0.0;
    return omega_p / k;
}

// validation.cpp:99-119
double or_fourier_mode_amplitude(const double* Ex, const or_domain* dom, int mode_x) {
    Dom d(dom);
    const double k = 2.0 * M_PI * static_cast<double>(mode_x) / d.Lx;
    double re = 0.0, im = 0.0;
    for (int j = 0; j < d.Ny; ++j) {
        for (int i = 0; i < d.Nx; ++i) {
            const double xx = (static_cast<double>(i) + 0.5) * d.dx;
            const double ex = Ex[static_cast<size_t>(j) * d.Nx + i];
            re += ex * std::cos(k * xx);
            im += ex * std::sin(k * xx);
        }
    }
    const double norm = static_cast<double>(d.Nx * d.Ny);
    re /= norm;
    im /= norm;
    return std::sqrt(re * re + im * im);
}

// validation.cpp:121-152
double or_growth_rate(const double* signal, int64_t n, double dt, int record_start_step) {
    if (n < 8) return 0.0;
    const int64_t start = n / 4;
    const int64_t end = (3 * n) / 4;
    if (end <= start + 2) return 0.0;
    double st = 0.0, sl = 0.0, stt = 0.0, stl = 0.0;
    int count = 0;
    for (int64_t i = start; i < end; ++i) {
        const double amp = std::max(signal[i], 1e-30);
        const double t = static_cast<double>(record_start_step + static_cast<int>(i)) * dt;
        const double l = std::log(amp);
        st += t;
        sl += l;
        stt += t * t;
        stl += t * l;
        ++count;
    }
    const double denom = static_cast<double>(count) * stt - st * st;
    if (std::abs(denom) < 1e-30) return 0.0;
    return (static_cast<double>(count) * stl - st * sl) / denom;
}

// ---------------------------------------------------------------------------
// Timestep loop (src/pic/simulation.cpp:69-145, src/pic/validation.cpp:223-241
// and :373-387).  order 0: push, deposit, +bg, solve, gather; order 1:
// deposit, +bg, solve, push, gather; order 2: deposit, push, +bg, solve,
// gather.  The deposit is unsorted (the validation loops set sorted=false).
// After each step with index >= record_from it records field energy,
// |mode-1 amplitude of Ex|, kinetic energy and charge error.  Particles are
// advanced in place.
// ---------------------------------------------------------------------------
int or_run_loop(double* x, double* y, double* vx, double* vy, const double* q, const double* m, int64_t n,
                const or_domain* dom, int scheme, int order, double dt, double background, int steps,
                int record_from, double* field_energy, double* mode1, double* kinetic, double* charge_err) {
    Dom d(dom);
    const size_t G = static_cast<size_t>(d.Nx) * d.Ny;
    std::vector<double> rho(G), phi(G), Ex(G, 0.0), Ey(G, 0.0);
    const double abs_q = or_total_abs_charge(q, n);
    const double tot_q = or_total_charge(q, n);
    for (int step = 0; step < steps; ++step) {
        auto deposit = [&]() { return or_deposit(scheme, x, y, vx, vy, q, n, dom, dt, rho.data()); };
        auto solve = [&]() {
            for (double& v : rho) v += background;
            or_poisson(rho.data(), dom, phi.data(), Ex.data(), Ey.data());
        };
        if (order == 0) {
            or_push(x, y, vx, vy, n, dom, dt);
            if (deposit()) return 1;
            solve();
        } else if (order == 1) {
            if (deposit()) return 1;
            solve();
            or_push(x, y, vx, vy, n, dom, dt);
        } else {
            if (deposit()) return 1;
            or_push(x, y, vx, vy, n, dom, dt);
            solve();
        }
        or_gather(x, y, vx, vy, q, m, n, Ex.data(), Ey.data(), dom, dt);
        if (step >= record_from) {
            const int s = step - record_from;
            if (field_energy) field_energy[s] = or_field_energy(Ex.data(), Ey.data(), dom);
            if (mode1) mode1[s] = or_fourier_mode_amplitude(Ex.data(), dom, 1);
            if (kinetic) kinetic[s] = or_kinetic_energy(vx, vy, m, n);
            if (charge_err) charge_err[s] = or_charge_error(or_integrated_rho(rho.data(), dom), tot_q, abs_q);
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Esirkepov charge-conserving current (build-defined; no reference
// counterpart).  Esirkepov, Comput. Phys. Commun. 135 (2001) 144, eqs. (21)-(23)
// specialised to 2D with CIC (first-order) shape factors S at the old and new
// positions.  Jx is stored on the x-face at (i+1/2, j) in array slot (i, j),
// Jy on the y-face at (i, j+1/2) in slot (i, j).  rho_old / rho_new receive
// the CIC densities at x0 and x1 = x0 + v dt (no wrap of x1 needed: indices
// are wrapped).  Continuity: (rho_new - rho_old)/dt + Dx Jx + Dy Jy = 0 with
// backward differences Dx Jx(i,j) = (Jx(i,j) - Jx(i-1,j))/dx.
// ---------------------------------------------------------------------------

static void cicShape1D(double xi, int& i0, double w[2]) {
    i0 = static_cast<int>(std::floor(xi));
    const double f = xi - i0;
    w[0] = 1.0 - f;
    w[1] = f;
}

int or_esirkepov_current(const double* x, const double* y, const double* vx, const double* vy, const double* q,
                         int64_t n, const or_domain* dom, double dt, double* rho_old, double* rho_new, double* Jx,
                         double* Jy) {
    Dom d(dom);
    const size_t nc = static_cast<size_t>(d.Nx) * d.Ny;
    std::fill(rho_old, rho_old + nc, 0.0);
    std::fill(rho_new, rho_new + nc, 0.0);
    std::fill(Jx, Jx + nc, 0.0);
    std::fill(Jy, Jy + nc, 0.0);
    const double inv_vol = 1.0 / d.cellVolume();
    for (int64_t p = 0; p < n; ++p) {
        const double fx0 = x[p] / d.dx, fy0 = y[p] / d.dy;
        const double fx1 = (x[p] + vx[p] * dt) / d.dx, fy1 = (y[p] + vy[p] * dt) / d.dy;
        if (std::abs(fx1 - fx0) > 1.0 || std::abs(fy1 - fy0) > 1.0) return 2;
        // Common 3x3 frame anchored at the old floor cell minus one.
        const int bx = static_cast<int>(std::floor(fx0)) - 1;
        const int by = static_cast<int>(std::floor(fy0)) - 1;
        double S0x[4] = {0, 0, 0, 0}, S1x[4] = {0, 0, 0, 0}, S0y[4] = {0, 0, 0, 0}, S1y[4] = {0, 0, 0, 0};
        int i0;
        double w[2];
        cicShape1D(fx0, i0, w);
        S0x[i0 - bx] += w[0];
        S0x[i0 - bx + 1] += w[1];
        cicShape1D(fx1, i0, w);
        S1x[i0 - bx] += w[0];
        S1x[i0 - bx + 1] += w[1];
        cicShape1D(fy0, i0, w);
        S0y[i0 - by] += w[0];
        S0y[i0 - by + 1] += w[1];
        cicShape1D(fy1, i0, w);
        S1y[i0 - by] += w[0];
        S1y[i0 - by + 1] += w[1];
        const double qv = q[p] * inv_vol;
        // Esirkepov decomposition W (eq. 21 in 2D).
        double Wx[4][4], Wy[4][4];
        for (int a = 0; a < 4; ++a) {
            for (int b = 0; b < 4; ++b) {
                const double dSx = S1x[a] - S0x[a];
                const double dSy = S1y[b] - S0y[b];
                Wx[a][b] = dSx * (S0y[b] + 0.5 * dSy);
                Wy[a][b] = dSy * (S0x[a] + 0.5 * dSx);
            }
        }
        for (int b = 0; b < 4; ++b) {
            double acc = 0.0;
            for (int a = 0; a < 4; ++a) {
                acc -= qv * d.dx / dt * Wx[a][b];
                const int gx = wrapIndex(bx + a, d.Nx), gy = wrapIndex(by + b, d.Ny);
                Jx[static_cast<size_t>(gy) * d.Nx + gx] += acc;
            }
        }
        for (int a = 0; a < 4; ++a) {
            double acc = 0.0;
            for (int b = 0; b < 4; ++b) {
                acc -= qv * d.dy / dt * Wy[a][b];
                const int gx = wrapIndex(bx + a, d.Nx), gy = wrapIndex(by + b, d.Ny);
                Jy[static_cast<size_t>(gy) * d.Nx + gx] += acc;
            }
        }
        for (int a = 0; a < 4; ++a) {
            for (int b = 0; b < 4; ++b) {
                const int gx = wrapIndex(bx + a, d.Nx), gy = wrapIndex(by + b, d.Ny);
                rho_old[static_cast<size_t>(gy) * d.Nx + gx] += qv * S0x[a] * S0y[b];
                rho_new[static_cast<size_t>(gy) * d.Nx + gx] += qv * S1x[a] * S1y[b];
            }
        }
    }
    return 0;
}

}  // extern "C"

// ---- denoiser network (build-defined; no reference counterpart) -------------
// FP64 restatement of the 1 -> 32 -> 32 -> 32 -> 1 periodic 3x3 convolution
// network: z_l(p, co) = b_l(co) + sum_t sum_ci W_l(t, co, ci) a_{l-1}(p + d_l off(t), ci),
// a_l = relu(z_l) for l < 4, out = z_4; tap t = 3 ty + tx, off(t) = (ty - 1, tx - 1);
// loss = mean((out - target)^2).  Parameter layout as the device library:
// W1[co][t], b1, W2[t][co][ci], b2, W3[t][co][ci], b3, W4[t][ci], b4.
// With emulate_bf16 the values the device keeps in BF16 are rounded the same
// way (round to nearest even through FP32): weights, the input image,
// hidden activations, the output gradient where it feeds a product, and the
// masked hidden gradients; sums stay in FP64.

namespace {

constexpr int kCW1 = 0, kCB1 = 288, kCW2 = 320, kCB2 = 9536, kCW3 = 9568, kCB3 = 18784, kCW4 = 18816, kCB4 = 19104;
constexpr int kCParams = 19105;

double bf16_round(double v) {
    float f = static_cast<float>(v);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return f;
}

struct Net {
    int B, H, W;
    int dil[4];
    bool emu;
    std::vector<double> w;  // possibly BF16-rounded weights
    size_t npix() const { return static_cast<size_t>(B) * H * W; }
    size_t shifted(size_t p, int t, int d) const {
        const int x = static_cast<int>(p % W), y = static_cast<int>((p / W) % H), b = static_cast<int>(p / (static_cast<size_t>(W) * H));
        int yy = (y + (t / 3 - 1) * d) % H, xx = (x + (t % 3 - 1) * d) % W;
        if (yy < 0) yy += H;
        if (xx < 0) xx += W;
        return (static_cast<size_t>(b) * H + yy) * W + xx;
    }
    double r(double v) const { return emu ? bf16_round(v) : v; }
    // layer l (1..4) forward: in has cin channels, out cout channels (NHWC)
    void conv(int l, const std::vector<double>& in, int cin, std::vector<double>& out, int cout) const {
        const size_t P = npix();
        out.assign(P * cout, 0.0);
        for (size_t p = 0; p < P; ++p)
            for (int t = 0; t < 9; ++t) {
                const size_t q = shifted(p, t, dil[l - 1]);
                for (int co = 0; co < cout; ++co) {
                    double s = 0.0;
                    for (int ci = 0; ci < cin; ++ci) s += weight(l, t, co, ci) * in[q * cin + ci];
                    out[p * cout + co] += s;
                }
            }
        for (size_t p = 0; p < P; ++p)
            for (int co = 0; co < cout; ++co) out[p * cout + co] += bias(l, co);
    }
    double weight(int l, int t, int co, int ci) const {
        switch (l) {
            case 1: return w[kCW1 + co * 9 + t];
            case 2: return w[kCW2 + (t * 32 + co) * 32 + ci];
            case 3: return w[kCW3 + (t * 32 + co) * 32 + ci];
            default: return w[kCW4 + t * 32 + ci];
        }
    }
    static int weight_index(int l, int t, int co, int ci) {
        switch (l) {
            case 1: return kCW1 + co * 9 + t;
            case 2: return kCW2 + (t * 32 + co) * 32 + ci;
            case 3: return kCW3 + (t * 32 + co) * 32 + ci;
            default: return kCW4 + t * 32 + ci;
        }
    }
    double bias(int l, int co) const {
        switch (l) {
            case 1: return w[kCB1 + co];
            case 2: return w[kCB2 + co];
            case 3: return w[kCB3 + co];
            default: return w[kCB4];
        }
    }
    static int bias_index(int l, int co) {
        switch (l) {
            case 1: return kCB1 + co;
            case 2: return kCB2 + co;
            case 3: return kCB3 + co;
            default: return kCB4;
        }
    }
};

Net make_net(const double* params, int B, int H, int W, const int* dil, int emu) {
    Net n{B, H, W, {1, 1, 1, 1}, emu != 0, std::vector<double>(params, params + kCParams)};
    for (int k = 0; k < 4; ++k) n.dil[k] = dil ? dil[k] : 1;
    if (n.emu)
        for (int i = 0; i < kCParams; ++i)
            if (!(i >= kCB1 && i < kCW2) && !(i >= kCB2 && i < kCW3) && !(i >= kCB3 && i < kCW4) && i != kCB4)
                n.w[static_cast<size_t>(i)] = bf16_round(n.w[static_cast<size_t>(i)]);
    return n;
}

// a[0] = input (1 channel), a[1..3] = hidden activations, out = output.
void net_forward(const Net& n, const double* in, std::vector<double> (&a)[4], std::vector<double>& out) {
    const size_t P = n.npix();
    a[0].resize(P);
    for (size_t p = 0; p < P; ++p) a[0][p] = n.r(in[p]);
    std::vector<double> z;
    for (int l = 1; l <= 3; ++l) {
        n.conv(l, a[l - 1], l == 1 ? 1 : 32, z, 32);
        a[l].resize(z.size());
        for (size_t i = 0; i < z.size(); ++i) a[l][i] = n.r(std::max(z[i], 0.0));
    }
    n.conv(4, a[3], 32, out, 1);
}

}  // namespace

extern "C" {

int or_cnn_forward(const double* params, const double* in, int B, int H, int W, const int* dil, int emulate_bf16,
                   double* out) {
    if (B < 1 || H < 1 || W < 1) return 1;
    const Net n = make_net(params, B, H, W, dil, emulate_bf16);
    std::vector<double> a[4], o;
    net_forward(n, in, a, o);
    std::copy(o.begin(), o.end(), out);
    return 0;
}

// Backpropagation through net_forward: loss and d loss / d params.
int or_cnn_backward(const double* params, const double* in, const double* target, int B, int H, int W,
                    const int* dil, int emulate_bf16, double* loss, double* grad) {
    if (B < 1 || H < 1 || W < 1) return 1;
    const Net n = make_net(params, B, H, W, dil, emulate_bf16);
    const size_t P = n.npix();
    std::vector<double> a[4], out;
    net_forward(n, in, a, out);
    double L = 0.0;
    std::vector<double> dout(P), dz(P);  // dz: the output gradient as it enters products
    for (size_t p = 0; p < P; ++p) {
        const double d = out[p] - target[p];
        L += d * d;
        dout[p] = 2.0 * d / static_cast<double>(P);
        dz[p] = n.r(dout[p]);
    }
    *loss = L / static_cast<double>(P);
    std::fill(grad, grad + kCParams, 0.0);
    std::vector<double> g = dz;  // gradient at the current layer's output (cout channels)
    std::vector<double> gb = dout;  // the same, as the bias gradient reads it
    for (int l = 4; l >= 1; --l) {
        const int cout = l == 4 ? 1 : 32, cin = l == 1 ? 1 : 32;
        const int d = n.dil[l - 1];
        for (size_t p = 0; p < P; ++p)
            for (int co = 0; co < cout; ++co) grad[Net::bias_index(l, co)] += gb[p * cout + co];
        for (size_t p = 0; p < P; ++p)
            for (int t = 0; t < 9; ++t) {
                const size_t q = n.shifted(p, t, d);
                for (int co = 0; co < cout; ++co) {
                    const double gv = g[p * cout + co];
                    if (gv == 0.0) continue;
                    for (int ci = 0; ci < cin; ++ci) grad[Net::weight_index(l, t, co, ci)] += gv * a[l - 1][q * cin + ci];
                }
            }
        if (l == 1) break;
        // gradient at the input activations, masked by ReLU'
        std::vector<double> ga(P * 32, 0.0);
        for (size_t p = 0; p < P; ++p)
            for (int t = 0; t < 9; ++t) {
                const size_t q = n.shifted(p, t, d);
                for (int co = 0; co < cout; ++co) {
                    const double gv = g[p * cout + co];
                    if (gv == 0.0) continue;
                    for (int ci = 0; ci < 32; ++ci) ga[q * 32 + ci] += n.weight(l, t, co, ci) * gv;
                }
            }
        for (size_t i = 0; i < ga.size(); ++i) ga[i] = a[l - 1][i] > 0.0 ? n.r(ga[i]) : 0.0;
        g = ga;
        gb = ga;
    }
    return 0;
}

}  // extern "C"
