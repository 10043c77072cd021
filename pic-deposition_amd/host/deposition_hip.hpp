// HIP seam of the reference's deposition dispatch.
//
// Drop-in for src/deposition/cuda/deposition_cuda.hpp:11-12
// (depositChargeCudaSoA / depositChargeCudaAoS): same arguments, same
// semantics (caller-owned particles are not reordered; every cell of
// grid.rho() is overwritten), errors thrown as std::runtime_error("HIP error:
// ...") like CUDA_CHECK (deposition_cuda.cu:12-18).  Compiled against the
// reference's own headers (deposition/deposition.hpp, pic/field_grid.hpp,
// pic/particles.hpp); forwards to the C ABI in include/pic_hip.h.
#pragma once

#include "deposition/deposition.hpp"
#include "pic/field_grid.hpp"
#include "pic/particles.hpp"

namespace pic {

void depositChargeHipSoA(const ParticlesSoA& particles, FieldGrid& grid, const DepositionConfig& config);
void depositChargeHipAoS(const ParticlesAoS& particles, FieldGrid& grid, const DepositionConfig& config);

}  // namespace pic
