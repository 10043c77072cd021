// HIP seam of the reference's deposition dispatch (see deposition_hip.hpp).

#include "deposition_hip.hpp"

#include <cstddef>
#include <stdexcept>
#include <string>

#include "pic_hip.h"

namespace pic {

namespace {

// schemeToInt, src/deposition/cuda/deposition_cuda.cu:275-286, except that
// Esirkepov (3) is deposited with its own shape instead of CIC.
int schemeCode(DepositionScheme s) {
    switch (s) {
        case DepositionScheme::NGP: return PIC_SCHEME_NGP;
        case DepositionScheme::CIC: return PIC_SCHEME_CIC;
        case DepositionScheme::TSC: return PIC_SCHEME_TSC;
        case DepositionScheme::Esirkepov: return PIC_SCHEME_ESIRKEPOV;
    }
    return PIC_SCHEME_CIC;
}

int backendCode(DepositionBackend b) {
    switch (b) {
        case DepositionBackend::GPUAtomics: return PIC_BACKEND_GPU_ATOMICS;
        case DepositionBackend::GPUPrivatized: return PIC_BACKEND_GPU_PRIV;
        case DepositionBackend::CPU: break;
    }
    throw std::runtime_error("HIP error: the CPU backend is not a device backend");
}

void check(pic_status s) {
    if (s != PIC_OK) throw std::runtime_error(std::string("HIP error: ") + pic_hip_last_error());
}

}  // namespace

void depositChargeHipSoA(const ParticlesSoA& particles, FieldGrid& grid, const DepositionConfig& config) {
    const Domain& d = grid.domain();
    const pic_domain pd{d.Lx, d.Ly, d.Nx, d.Ny};
    const bool esk = config.scheme == DepositionScheme::Esirkepov;
    check(pic_hip_deposit_host(particles.x().data(), particles.y().data(), esk ? particles.vx().data() : nullptr,
                               esk ? particles.vy().data() : nullptr, particles.q().data(), particles.size(), &pd,
                               schemeCode(config.scheme), backendCode(config.backend), config.sorted ? 1 : 0,
                               config.esirkepov_dt, grid.rho().data()));
}

// The records go to the library as they lie in memory: six FP64 fields in the
// order x, y, vx, vy, q, m (src/pic/particles.hpp:13-20).
static_assert(sizeof(ParticleAoS) == 6 * sizeof(double), "ParticleAoS must be six packed doubles");
static_assert(offsetof(ParticleAoS, x) == 0 && offsetof(ParticleAoS, y) == 8 && offsetof(ParticleAoS, vx) == 16 &&
                  offsetof(ParticleAoS, vy) == 24 && offsetof(ParticleAoS, q) == 32 && offsetof(ParticleAoS, m) == 40,
              "ParticleAoS field order");

void depositChargeHipAoS(const ParticlesAoS& particles, FieldGrid& grid, const DepositionConfig& config) {
    // The reference's AoS device path builds an FP64 SoA copy first
    // (deposition_cuda.cu:330-333); here the records are rounded to FP32 straight
    // from the caller's array.
    const Domain& d = grid.domain();
    const pic_domain pd{d.Lx, d.Ly, d.Nx, d.Ny};
    const double* rec = particles.size() ? &particles.data()[0].x : nullptr;
    check(pic_hip_deposit_host_aos(rec, particles.size(), &pd, schemeCode(config.scheme), backendCode(config.backend),
                                   config.sorted ? 1 : 0, config.esirkepov_dt, grid.rho().data()));
}

}  // namespace pic
