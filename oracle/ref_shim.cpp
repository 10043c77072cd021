// C binding over the reference's own CPU sources, compiled from where they lie
// in the reference checkout by oracle/Makefile (target `ref`).  TEST
// INFRASTRUCTURE ONLY: used to validate the FP64 restatement in pic_oracle.cpp
// and as the "reference" CPU baseline.  The translation units that need FFTW
// (src/pic/diagnostics.cpp, src/pic/poisson_fft.cpp) cannot be built here (no
// fftw3.h); depositChargeTimed* reference symbols from diagnostics.cpp and stay
// unresolved, so the library must be opened with RTLD_LAZY and those two
// functions are never called.

#include "deposition/deposition.hpp"
#include "deposition/deposition_impl.hpp"
#include "deposition/deposition_parallel.hpp"
#include "pic/domain.hpp"
#include "pic/field_grid.hpp"
#include "pic/particles.hpp"
#include "sorting/cell_sort.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

pic::Domain makeDomain(double Lx, double Ly, int Nx, int Ny) {
    pic::Domain d;
    d.Lx = Lx;
    d.Ly = Ly;
    d.Nx = Nx;
    d.Ny = Ny;
    d.updateDerived();
    return d;
}

pic::ParticlesSoA makeSoA(const double* x, const double* y, const double* vx, const double* vy, const double* q,
                          const double* m, int64_t n) {
    pic::ParticlesSoA p(static_cast<std::size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        p.x()[i] = x[i];
        p.y()[i] = y[i];
        p.vx()[i] = vx ? vx[i] : 0.0;
        p.vy()[i] = vy ? vy[i] : 0.0;
        p.q()[i] = q[i];
        p.m()[i] = m ? m[i] : 1.0;
    }
    return p;
}

void unpackSoA(const pic::ParticlesSoA& p, double* x, double* y, double* vx, double* vy, double* q, double* m) {
    const std::size_t n = p.size();
    std::copy(p.x().begin(), p.x().end(), x);
    std::copy(p.y().begin(), p.y().end(), y);
    std::copy(p.vx().begin(), p.vx().end(), vx);
    std::copy(p.vy().begin(), p.vy().end(), vy);
    std::copy(p.q().begin(), p.q().end(), q);
    std::copy(p.m().begin(), p.m().end(), m);
    (void)n;
}

}  // namespace

extern "C" {

// depositChargeSoA (deposition.cpp:181-200) with backend CPU.
int ref_deposit_soa(int scheme, int num_threads, int sorted, const double* x, const double* y, const double* vx,
                    const double* vy, const double* q, int64_t n, double Lx, double Ly, int Nx, int Ny, double dt,
                    double* rho) {
    pic::Domain d = makeDomain(Lx, Ly, Nx, Ny);
    pic::FieldGrid grid(d);
    pic::ParticlesSoA p = makeSoA(x, y, vx, vy, q, nullptr, n);
    pic::DepositionConfig cfg;
    cfg.scheme = static_cast<pic::DepositionScheme>(scheme);
    cfg.layout = pic::ParticleLayout::SoA;
    cfg.backend = pic::DepositionBackend::CPU;
    cfg.sorted = sorted != 0;
    cfg.num_threads = num_threads;
    cfg.esirkepov_dt = dt;
    pic::depositChargeSoA(p, grid, cfg);
    std::copy(grid.rho().begin(), grid.rho().end(), rho);
    return 0;
}

// depositChargeAoS (deposition.cpp:160-179) with backend CPU.
int ref_deposit_aos(int scheme, int num_threads, const double* x, const double* y, const double* vx,
                    const double* vy, const double* q, int64_t n, double Lx, double Ly, int Nx, int Ny, double dt,
                    double* rho) {
    pic::Domain d = makeDomain(Lx, Ly, Nx, Ny);
    pic::FieldGrid grid(d);
    pic::ParticlesAoS p = pic::toAoS(makeSoA(x, y, vx, vy, q, nullptr, n));
    pic::DepositionConfig cfg;
    cfg.scheme = static_cast<pic::DepositionScheme>(scheme);
    cfg.layout = pic::ParticleLayout::AoS;
    cfg.num_threads = num_threads;
    cfg.esirkepov_dt = dt;
    pic::depositChargeAoS(p, grid, cfg);
    std::copy(grid.rho().begin(), grid.rho().end(), rho);
    return 0;
}

// particleCellIndex (cell_sort.cpp:8-13)
void ref_cell_index(const double* x, const double* y, int64_t n, double Lx, double Ly, int Nx, int Ny,
                    int32_t* keys) {
    pic::Domain d = makeDomain(Lx, Ly, Nx, Ny);
    for (int64_t i = 0; i < n; ++i) keys[i] = pic::particleCellIndex(x[i], y[i], d);
}

// sortParticlesByCell SoA (cell_sort.cpp:22-46), in place on the six arrays.
void ref_sort_soa(double* x, double* y, double* vx, double* vy, double* q, double* m, int64_t n, double Lx,
                  double Ly, int Nx, int Ny) {
    pic::Domain d = makeDomain(Lx, Ly, Nx, Ny);
    pic::ParticlesSoA p = makeSoA(x, y, vx, vy, q, m, n);
    pic::sortParticlesByCell(p, d);
    unpackSoA(p, x, y, vx, vy, q, m);
}

// gatherFieldsSoA / pushParticlesSoA (deposition.cpp:335-374), in place.
void ref_gather_push_soa(int do_gather, int do_push, double* x, double* y, double* vx, double* vy, double* q,
                         double* m, int64_t n, double Lx, double Ly, int Nx, int Ny, const double* Ex,
                         const double* Ey, double dt) {
    pic::Domain d = makeDomain(Lx, Ly, Nx, Ny);
    pic::FieldGrid grid(d);
    std::copy(Ex, Ex + static_cast<std::size_t>(Nx) * Ny, grid.Ex().begin());
    std::copy(Ey, Ey + static_cast<std::size_t>(Nx) * Ny, grid.Ey().begin());
    pic::ParticlesSoA p = makeSoA(x, y, vx, vy, q, m, n);
    if (do_gather) pic::gatherFieldsSoA(p, grid, dt);
    if (do_push) pic::pushParticlesSoA(p, grid, dt);
    unpackSoA(p, x, y, vx, vy, q, m);
}

// Initialisers (particles_aos.cpp).  kind: 0 uniform Maxwellian (p1 = T),
// 1 Langmuir lattice (p1 = amplitude, mode), 2 two-stream (p1 = beam v,
// p2 = perturbation), 3 quasi-1D two-stream, 4 warm Langmuir (p1 = amplitude,
// p2 = temperature, mode).
void ref_init_soa(int kind, int64_t n, double Lx, double Ly, int Nx, int Ny, double p1, double p2, int mode,
                  unsigned seed, double* x, double* y, double* vx, double* vy, double* q, double* m) {
    pic::Domain d = makeDomain(Lx, Ly, Nx, Ny);
    pic::ParticlesSoA p(static_cast<std::size_t>(n));
    switch (kind) {
        case 0: p.initializeUniformMaxwellian(d, p1, seed); break;
        case 1: p.initializeLangmuirWave(d, p1, mode, seed); break;
        case 2: p.initializeTwoStream(d, p1, p2, seed); break;
        case 3: p.initializeTwoStreamQuasi1D(d, p1, p2, seed); break;
        default: p.initializeWarmLangmuirWave(d, p1, mode, p2, seed); break;
    }
    unpackSoA(p, x, y, vx, vy, q, m);
}

// The timed loop of depositChargeTimedSoA (deposition.cpp:257-301) around
// depositKernelSoA's CPU branch (deposition.cpp:120-126): clearRho, then the
// serial or OpenMP scheme function; 3 warm-ups, `repeats` timed deposits with
// steady_clock.  The particle set is built once, outside the timed region.
double ref_deposit_kernel_timed(int scheme, int num_threads, const double* x, const double* y, const double* vx,
                                const double* vy, const double* q, int64_t n, double Lx, double Ly, int Nx, int Ny,
                                double dt, int repeats, double* std_ms, double* rho) {
    pic::Domain d = makeDomain(Lx, Ly, Nx, Ny);
    pic::FieldGrid grid(d);
    pic::ParticlesSoA p = makeSoA(x, y, vx, vy, q, nullptr, n);
    pic::DepositionConfig cfg;
    cfg.scheme = static_cast<pic::DepositionScheme>(scheme);
    cfg.num_threads = num_threads;
    cfg.esirkepov_dt = dt;
    auto once = [&]() {
        grid.clearRho();
        if (num_threads > 1) {
            switch (cfg.scheme) {
                case pic::DepositionScheme::NGP: pic::depositNGP_SoA_Parallel(p, grid, num_threads); break;
                case pic::DepositionScheme::CIC: pic::depositCIC_SoA_Parallel(p, grid, num_threads); break;
                case pic::DepositionScheme::TSC: pic::depositTSC_SoA_Parallel(p, grid, num_threads); break;
                default: pic::depositEsirkepov_SoA_Parallel(p, grid, dt, num_threads); break;
            }
        } else {
            switch (cfg.scheme) {
                case pic::DepositionScheme::NGP: pic::depositNGP_SoA(p, grid); break;
                case pic::DepositionScheme::CIC: pic::depositCIC_SoA(p, grid); break;
                case pic::DepositionScheme::TSC: pic::depositTSC_SoA(p, grid); break;
                default: pic::depositEsirkepov_SoA(p, grid, dt); break;
            }
        }
    };
    for (int i = 0; i < 3; ++i) once();
    std::vector<double> samples;
    for (int i = 0; i < repeats; ++i) {
        const auto t0 = std::chrono::steady_clock::now();
        once();
        const auto t1 = std::chrono::steady_clock::now();
        samples.push_back(std::chrono::duration<double>(t1 - t0).count() * 1000.0);
    }
    double sum = 0.0;
    for (double v : samples) sum += v;
    const double mean = samples.empty() ? 0.0 : sum / samples.size();
    double var = 0.0;
    for (double v : samples) var += (v - mean) * (v - mean);
    if (std_ms) *std_ms = samples.size() > 1 ? std::sqrt(var / (samples.size() - 1)) : 0.0;
    if (rho) std::copy(grid.rho().begin(), grid.rho().end(), rho);
    return mean;
}

// FieldGrid::integratedRho and FieldGrid::fieldEnergy (field_grid.cpp:24-38)
// on a grid holding the given rho, Ex, Ey.
double ref_integrated_rho(const double* rho, double Lx, double Ly, int Nx, int Ny) {
    pic::FieldGrid grid(makeDomain(Lx, Ly, Nx, Ny));
    std::copy(rho, rho + static_cast<std::size_t>(Nx) * Ny, grid.rho().begin());
    return grid.integratedRho();
}

double ref_field_energy(const double* Ex, const double* Ey, double Lx, double Ly, int Nx, int Ny) {
    pic::FieldGrid grid(makeDomain(Lx, Ly, Nx, Ny));
    std::copy(Ex, Ex + static_cast<std::size_t>(Nx) * Ny, grid.Ex().begin());
    std::copy(Ey, Ey + static_cast<std::size_t>(Nx) * Ny, grid.Ey().begin());
    return grid.fieldEnergy();
}

// ParticlesSoA::totalCharge and ParticlesSoA::kineticEnergy
// (particles_aos.cpp:182-196).
double ref_total_charge(const double* q, int64_t n) {
    return makeSoA(q, q, nullptr, nullptr, q, nullptr, n).totalCharge();
}

double ref_kinetic_energy(const double* vx, const double* vy, const double* m, int64_t n) {
    return makeSoA(vx, vy, vx, vy, vx, m, n).kineticEnergy();
}

// tscWeights (ngp.cpp:28-34) and esirkepovWeights1D (ngp.cpp:86-93).
void ref_tsc_weights(double xi, double* w, int* i0) { pic::tscWeights(xi, w, *i0); }

void ref_esirkepov_weights_1d(double x0, double x1, double dx, int n, double* dense) {
    std::vector<double> w;
    pic::esirkepovWeights1D(x0, x1, dx, n, w);
    std::copy(w.begin(), w.end(), dense);
}

}  // extern "C"
