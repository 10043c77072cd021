// Drop-in test of the HIP seam inside the reference's own dispatch code.
//
// Built (by oracle/Makefile, target `seam`) together with the reference's CPU
// sources compiled with -DPIC_BUILD_CUDA, so that the reference's unmodified
// depositChargeSoA / depositChargeAoS (src/deposition/deposition.cpp:160-200)
// route their GPU backends through the two seam functions defined below into
// the HIP library.  The checks mirror src/main_test.cpp:158-196 (GPU charge
// conservation) and add rho parity against the reference's own CPU backend.
// Test infrastructure only.

#include "deposition/cuda/deposition_cuda.hpp"
#include "deposition/deposition.hpp"
#include "deposition_hip.hpp"
#include "pic/field_grid.hpp"
#include "pic/particles.hpp"

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace pic {

// The reference's CUDA seam names bound to the HIP adapter: the whole change a
// PIC_BUILD_CUDA build of deposition.cpp needs.
void depositChargeCudaSoA(const ParticlesSoA& p, FieldGrid& g, const DepositionConfig& c) {
    depositChargeHipSoA(p, g, c);
}
void depositChargeCudaAoS(const ParticlesAoS& p, FieldGrid& g, const DepositionConfig& c) {
    depositChargeHipAoS(p, g, c);
}

}  // namespace pic

namespace {

int g_fail = 0;

void expect(bool ok, const std::string& name, double value) {
    std::printf("%s %s (%.3e)\n", ok ? "PASS" : "FAIL", name.c_str(), value);
    std::fflush(stdout);
    if (!ok) ++g_fail;
}

pic::Domain makeDomain(int n) {
    pic::Domain d;
    d.Nx = n;
    d.Ny = n;
    d.updateDerived();
    return d;
}

// Round every particle attribute to FP32 so that the CPU backend sees exactly
// the values the device computes with.
void roundToFloat(pic::ParticlesSoA& p) {
    for (auto* v : {&p.x(), &p.y(), &p.vx(), &p.vy(), &p.q()})
        for (double& e : *v) e = static_cast<double>(static_cast<float>(e));
}

double chargeError(const pic::FieldGrid& g, const pic::ParticlesSoA& p) {
    double abs_sum = 0.0;
    for (double q : p.q()) abs_sum += std::abs(q);
    return std::abs(g.integratedRho() - p.totalCharge()) / std::max(abs_sum, 1e-12);
}

double relL2(const std::vector<double>& a, const std::vector<double>& b) {
    double num = 0.0, den = 0.0;
    for (std::size_t i = 0; i < a.size(); ++i) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return std::sqrt(num / std::max(den, 1e-300));
}

}  // namespace

extern "C" int seam_run_tests() {
    g_fail = 0;
    const pic::DepositionScheme schemes[] = {pic::DepositionScheme::NGP, pic::DepositionScheme::CIC,
                                             pic::DepositionScheme::TSC, pic::DepositionScheme::Esirkepov};
    const pic::DepositionBackend backends[] = {pic::DepositionBackend::GPUAtomics,
                                               pic::DepositionBackend::GPUPrivatized};
    try {
        // main_test.cpp:158-180: 128x128, 5000 Maxwellian particles, seed 123.
        for (auto backend : backends) {
            pic::Domain d = makeDomain(128);
            pic::FieldGrid grid(d);
            pic::ParticlesSoA p(5000);
            p.initializeUniformMaxwellian(d, 1.0, 123);
            pic::DepositionConfig cfg;
            cfg.scheme = pic::DepositionScheme::CIC;
            cfg.backend = backend;
            pic::depositChargeSoA(p, grid, cfg);
            const double err = chargeError(grid, p);
            expect(err <= 1e-6, "GPU CIC " + pic::backendName(backend) + " charge conservation", err);
        }
        // rho parity against the reference CPU backend, both layouts, sorted and not.
        for (auto scheme : schemes) {
            pic::Domain d = makeDomain(64);
            pic::ParticlesSoA p(64 * 64 * 16);
            p.initializeUniformMaxwellian(d, 1.0, 42);
            for (double& v : p.vx()) v *= 0.3;  // keep Esirkepov moves (v*dt) under one cell
            for (double& v : p.vy()) v *= 0.3;
            roundToFloat(p);
            pic::DepositionConfig cpu;
            cpu.scheme = scheme;
            cpu.esirkepov_dt = 0.01;
            pic::FieldGrid ref(d);
            pic::depositChargeSoA(p, ref, cpu);
            for (auto backend : backends) {
                std::vector<double> soa_unsorted;
                for (int sorted = 0; sorted < 2; ++sorted) {
                    pic::DepositionConfig cfg = cpu;
                    cfg.backend = backend;
                    cfg.sorted = sorted != 0;
                    pic::FieldGrid g(d);
                    pic::depositChargeSoA(p, g, cfg);
                    if (!sorted) soa_unsorted = g.rho();
                    const double e = relL2(g.rho(), ref.rho());
                    expect(e <= 1e-6, pic::schemeName(scheme) + " SoA " + pic::backendName(backend) +
                                          (sorted ? " sorted" : " unsorted") + " vs CPU rel L2",
                           e);
                }
                pic::DepositionConfig cfg = cpu;
                cfg.backend = backend;
                cfg.layout = pic::ParticleLayout::AoS;
                pic::FieldGrid g(d);
                pic::depositChargeAoS(pic::toAoS(p), g, cfg);
                const double e = relL2(g.rho(), ref.rho());
                expect(e <= 1e-6, pic::schemeName(scheme) + " AoS " + pic::backendName(backend) + " vs CPU rel L2",
                       e);
                if (backend == pic::DepositionBackend::GPUPrivatized)  // deterministic: the same bits as SoA
                    expect(g.rho() == soa_unsorted, pic::schemeName(scheme) + " AoS GPU_Priv == SoA bitwise", 0.0);
            }
        }
        // errors surface as std::runtime_error("HIP error: ...")
        {
            pic::Domain d = makeDomain(16);
            d.Nx = 0;
            pic::FieldGrid grid(makeDomain(16));
            pic::ParticlesSoA p(10);
            bool thrown = false;
            try {
                pic::depositChargeHipSoA(p, grid, pic::DepositionConfig{});
            } catch (const std::runtime_error& e) {
                thrown = std::string(e.what()).rfind("HIP error:", 0) == 0;
            }
            expect(thrown, "CPU backend through the device seam throws HIP error", 0.0);
        }
    } catch (const std::exception& e) {
        std::printf("FAIL exception: %s\n", e.what());
        ++g_fail;
    }
    std::printf("%d failures\n", g_fail);
    std::fflush(stdout);
    return g_fail;
}
