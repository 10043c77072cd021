// Host-side helpers shared by the C ABI files (api.hip, ensemble_api.hip):
// error reporting, allocation, domain conversion and the tiling autotuner's
// candidate list and JSON cache.
#pragma once

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pic_hip.h"
#include "pic_common.hpp"
#include "pic_kernels.hpp"

namespace pichip {

// Message of the last failed call on this thread (one slot per thread for the
// whole library).
inline std::string& last_error_slot() {
    thread_local std::string msg;
    return msg;
}

inline pic_status fail(pic_status s, const std::string& msg) {
    last_error_slot() = msg;
    return s;
}

#define PIC_HIP_TRY(call)                                                                                   \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess) {                                                                             \
            return ::pichip::fail(e_ == hipErrorOutOfMemory ? PIC_ERR_OUT_OF_MEMORY : PIC_ERR_HIP,           \
                                  std::string("HIP error: ") + hipGetErrorString(e_) + " (" #call ")");      \
        }                                                                                                   \
    } while (0)

#define PIC_FFT_TRY(call)                                                                                    \
    do {                                                                                                     \
        hipfftResult r_ = (call);                                                                            \
        if (r_ != HIPFFT_SUCCESS)                                                                            \
            return ::pichip::fail(PIC_ERR_FFT, std::string("hipFFT error ") + std::to_string((int)r_) +      \
                                                   " (" #call ")");                                          \
    } while (0)

constexpr int kStreams = 6;  // x, y, vx, vy, q, m
constexpr size_t kAosFields = 6;            // FP64 values per ParticleAoS record
constexpr size_t kAosChunk = size_t(1) << 22;  // records staged per device copy
constexpr int64_t kStagedSortMin = int64_t(1) << 22;  // particles from which a sort may stage records

inline size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

template <typename T>
hipError_t dmalloc(T** p, size_t elems) {
    *p = nullptr;
    return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(elems, 1) * sizeof(T));
}

inline bool valid_domain(const pic_domain* d) {
    return d && d->Nx > 0 && d->Ny > 0 && std::isfinite(d->Lx) && std::isfinite(d->Ly) && d->Lx > 0.0 &&
           d->Ly > 0.0 && (int64_t)d->Nx * d->Ny < (1ll << 30);
}

inline DevDomain make_dev_domain(const pic_domain& d) {
    DevDomain dd;
    dd.Nx = d.Nx;
    dd.Ny = d.Ny;
    dd.Lx = d.Lx;
    dd.Ly = d.Ly;
    dd.dx = d.Lx / d.Nx;  // Domain::updateDerived, domain.hpp:17-20
    dd.dy = d.Ly / d.Ny;
    dd.inv_dx = (float)(1.0 / dd.dx);
    dd.inv_dy = (float)(1.0 / dd.dy);
    dd.inv_vol = (float)(1.0 / (dd.dx * dd.dy));
    int ex = 0, ey = 0;
    dd.exact = (std::frexp(dd.dx, &ex) == 0.5 && std::frexp(dd.dy, &ey) == 0.5) ? 1 : 0;
    return dd;
}

// Largest LDS a deposit workgroup (four waves) may take.
constexpr size_t kMaxDepositLds = 64 * 1024;

// Default tiling per scheme (NGP, CIC, TSC, Esirkepov) from the MI355X sweep at
// 1024^2 / 64 particles per cell: {tile_x, tile_y, particles per lane}.
constexpr int kDefaultTiling[4][3] = {{64, 4, 8}, {64, 4, 8}, {32, 4, 8}, {32, 4, 4}};
// Long-row tiling of a single grid with at least 4096 such tiles (the MI355X
// sweep at 1024^2 / 64 ppc, interleaved repeats): {tile_x, tile_y, particles
// per lane, pipelined}.  TSC takes short rows: its moment form (node-keyed
// order) is bound by VALU issue and per-tile LDS, and 16 x 1 tiles keep the
// most waves resident; with 48 VGPRs it double-buffers its loads.
constexpr int kRowTiling[4][4] = {{128, 2, 4, 0}, {128, 2, 4, 0}, {16, 1, 4, 1}, {128, 1, 4, 0}};
// Long rows of TSC on the cell-keyed order (16-slot stencil).
constexpr int kRowTilingTscCell[3] = {128, 1, 8};

// Default tile shape for a grid: the first of the scheme's shapes (largest
// first, down to 8 x 2) that gives at least `min_tiles` tiles and at most
// `max_per_tile` particles per tile, without going below `min_per_tile`
// particles per tile (a wave chunk is 256 or 512 particles) when the particle
// count n is known.  Few large tiles leave the GPU idle behind a handful of
// long waves (a 256^2 grid at 512 particles per cell deposits 3.5x faster in
// 8 x 2 tiles than in 32 x 4); tiles with fewer particles than a chunk leave
// lanes idle and only add window halo to the merge.
inline void default_tile_shape(int scheme, int Nx, int Ny, int64_t n, int min_tiles, int64_t max_per_tile,
                               int64_t min_per_tile, int& tx, int& ty) {
    static const int wide[][2] = {{64, 4}, {32, 4}, {16, 4}, {8, 4}, {8, 2}};
    static const int narrow[][2] = {{32, 4}, {16, 4}, {8, 4}, {8, 2}};
    const bool w = scheme == 0 || scheme == 1;
    const int count = w ? 5 : 4;
    const int64_t G = std::max<int64_t>((int64_t)Nx * Ny, 1);
    bool have = false;
    for (int i = 0; i < count; ++i) {
        const int cx = w ? wide[i][0] : narrow[i][0];
        const int cy = w ? wide[i][1] : narrow[i][1];
        const int64_t tiles = (int64_t)((Nx + cx - 1) / cx) * ((Ny + cy - 1) / cy);
        const int64_t per_tile = n * cx * cy / G;
        if (n > 0 && have && per_tile < min_per_tile) return;  // keep the previous, larger shape
        tx = cx;
        ty = cy;
        have = true;
        if (tiles >= min_tiles && (n <= 0 || per_tile <= max_per_tile)) return;
    }
}

// ---- autotuner ----------------------------------------------------------------

struct TilingChoice {
    int tx, ty, ppl;
};

// Candidate tilings for a grid: tile shapes that fit the grid (a tile larger
// than the grid only adds halo), 4 or 8 particles per lane (4 only when
// `ppl4_only`), windows within the workgroup LDS budget.
inline std::vector<TilingChoice> tiling_candidates(int Nx, int Ny, bool ppl4_only, int halo) {
    static const int shapes[][2] = {{8, 2},  {8, 4},   {16, 4},  {16, 8},  {32, 4},  {32, 8},  {64, 4},
                                    {64, 8}, {128, 4}, {16, 16}, {32, 16}, {128, 2}, {128, 1}};
    std::vector<TilingChoice> out;
    for (const auto& sh : shapes) {
        if (sh[0] > std::max(Nx, 16) || sh[1] > std::max(Ny, 2)) continue;
        const TileGeom g = make_tile_geom(Nx, Ny, sh[0], sh[1], halo);
        if ((size_t)4 * g.lds_per_wave * sizeof(float) > kMaxDepositLds) continue;
        for (int ppl : {4, 8}) {
            if (ppl4_only && ppl != 4) continue;
            out.push_back({sh[0], sh[1], ppl});
        }
    }
    return out;
}

inline const char* scheme_tag(int scheme) {
    static const char* names[4] = {"NGP", "CIC", "TSC", "Esirkepov"};
    return (scheme >= 0 && scheme < 4) ? names[scheme] : "unknown";
}

// Particles per cell rounded to the nearest power of two (at least 1).
inline int64_t ppc_bucket(int64_t n, int64_t G) {
    const double p = G > 0 ? (double)n / (double)G : 0.0;
    if (p <= 1.0) return 1;
    return (int64_t)1 << (int)std::lround(std::log2(p));
}

inline std::string device_arch() {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    std::string a = prop.gcnArchName;
    const size_t c = a.find(':');
    return c == std::string::npos ? a : a.substr(0, c);
}

inline std::string tuning_key(int scheme, int Nx, int Ny, int64_t ppc, int halo) {
    std::ostringstream k;
    k << device_arch() << "/" << scheme_tag(scheme) << "/" << Nx << "x" << Ny << "/ppc" << ppc;
    if (halo) k << "/halo" << halo;
    return k.str();
}

// The cache is a flat JSON object {"<key>": [tile_x, tile_y, particles_per_lane], ...}
// written by tuning_cache_store; the reader accepts exactly that shape.
inline std::map<std::string, TilingChoice> tuning_cache_read(const char* path) {
    std::map<std::string, TilingChoice> out;
    if (!path || !*path) return out;
    std::ifstream f(path);
    if (!f) return out;
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string s = ss.str();
    size_t i = 0;
    while ((i = s.find('"', i)) != std::string::npos) {
        const size_t j = s.find('"', i + 1);
        if (j == std::string::npos) break;
        const std::string key = s.substr(i + 1, j - i - 1);
        const size_t lb = s.find('[', j), rb = s.find(']', j);
        const size_t nq = s.find('"', j + 1);
        if (lb == std::string::npos || rb == std::string::npos || (nq != std::string::npos && nq < lb)) {
            i = j + 1;
            continue;
        }
        int v[3] = {0, 0, 0};
        if (std::sscanf(s.c_str() + lb, "[%d , %d , %d]", &v[0], &v[1], &v[2]) == 3 && v[0] > 0 && v[1] > 0 &&
            (v[2] == 4 || v[2] == 8))
            out[key] = {v[0], v[1], v[2]};
        i = rb + 1;
    }
    return out;
}

inline bool tuning_cache_store(const char* path, const std::string& key, const TilingChoice& t) {
    if (!path || !*path) return true;
    auto all = tuning_cache_read(path);
    all[key] = t;
    const std::string tmp = std::string(path) + ".tmp";
    {
        std::ofstream f(tmp, std::ios::trunc);
        if (!f) return false;
        f << "{\n";
        size_t k = 0;
        for (const auto& kv : all) {
            f << "  \"" << kv.first << "\": [" << kv.second.tx << ", " << kv.second.ty << ", " << kv.second.ppl << "]"
              << (++k < all.size() ? ",\n" : "\n");
        }
        f << "}\n";
        if (!f) return false;
    }
    return std::rename(tmp.c_str(), path) == 0;
}

inline double median_of(std::vector<double> v) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const size_t n = v.size();
    return n % 2 ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

}  // namespace pichip
