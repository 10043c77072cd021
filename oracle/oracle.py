"""Python binding of the FP64 CPU oracle (liboracle.so) and of the reference
build (_ref/libpicref.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this module, and only as the checker.

Pinning of each oracle function (what independent evidence fixes it):

===========================  =================================================
oracle function              pinned by
===========================  =================================================
or_deposit (NGP/CIC/TSC/     bit-for-bit against the reference build
Esirkepov density, serial)   (ref_deposit_soa, ngp.cpp compiled from the
                             reference checkout); closed-form single-particle
                             KATs (tests/test_oracle.py); partition of unity
                             and first moments; charge conservation
                             (main_test.cpp:25-72 tolerances).
or_deposit_omp               against the reference build's OpenMP path for
                             NGP/CIC/Esirkepov (deposition_parallel.cpp) and
                             against or_deposit to 1e-12 for all schemes.
or_cell_keys                 bit-exact vs ref_cell_index (particleCellIndex).
or_sort_perm                 sorted keys and permuted particles identical to
                             ref_sort_soa up to the order inside a cell (the
                             reference std::sort is unstable); stability
                             checked directly against numpy's stable argsort.
or_cell_offsets              histogram closed form (numpy bincount / cumsum).
or_gather / or_push          bit-exact vs ref_gather_push_soa
                             (deposition.cpp:335-374); constant and linear
                             field closed forms.
or_init_*                    bit-exact vs ref_init_soa (particles_aos.cpp).
or_poisson                   single Fourier mode closed form
                             phi = cv*rho0*cos(kx)/k^2; numpy FFT spectral
                             solve; E forward difference.  FFTW itself is
                             absent: parity at the FFTW boundary is unpinned.
or_dominant_frequency,       synthetic sinusoid (peak bin) and exponential
or_growth_rate,              (log slope) closed forms; mode amplitude of a
or_fourier_mode_amplitude    pure cos/sin field.
or_esirkepov_current         discrete continuity identity (rho_new-rho_old)/dt
                             + div J = 0 to round-off; zero motion gives
                             J = 0 and rho_old = rho_new = CIC.
or_integrated_rho,           bit-for-bit against the reference build
or_field_energy,             (FieldGrid::integratedRho / fieldEnergy,
or_total_charge,             ParticlesSoA::totalCharge / kineticEnergy:
or_kinetic_energy            ref_* in ref_shim.cpp) on random grids and
                             particles; closed forms: single Fourier mode
                             field energy A^2 Lx Ly / 4, uniform-velocity
                             kinetic energy N m v^2 / 2, uniform rho.
or_total_abs_charge,         closed forms (diagnostics.cpp:18-29 is not
or_charge_error              buildable here: it includes fftw3.h): alternating
                             +-q sets, the 1e-12 denominator floor.
or_cnn_forward               an independent numpy restatement with np.roll
                             (periodic correlation) to 1e-12; a centred delta
                             kernel is the identity and a shifted one a roll.
or_cnn_backward              central finite differences of or_cnn_forward's
                             loss (FP64, no BF16 rounding) on sampled
                             parameters of every layer, weights and biases.
===========================  =================================================
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(_HERE, "liboracle.so")
REF_LIB = os.path.join(_HERE, "_ref", "libpicref.so")


class ODomain(ctypes.Structure):
    _fields_ = [("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("Nx", ctypes.c_int32), ("Ny", ctypes.c_int32)]


_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)
_U = ctypes.POINTER(ctypes.c_uint32)
_L = ctypes.c_int64
_DOM = ctypes.POINTER(ODomain)

_OR_SIGS = {
    "or_deposit": (ctypes.c_int, [ctypes.c_int, _D, _D, _D, _D, _D, _L, _DOM, ctypes.c_double, _D]),
    "or_deposit_omp": (ctypes.c_int, [ctypes.c_int, _D, _D, _D, _D, _D, _L, _DOM, ctypes.c_double, ctypes.c_int, _D]),
    "or_deposit_omp_timed": (ctypes.c_double, [ctypes.c_int, _D, _D, _D, _D, _D, _L, _DOM, ctypes.c_double,
                                               ctypes.c_int, ctypes.c_int, _D, _D]),
    "or_cell_keys": (None, [_D, _D, _L, _DOM, _I]),
    "or_sort_perm": (None, [_I, _L, _I]),
    "or_cell_offsets": (None, [_I, _L, ctypes.c_int, _I]),
    "or_gather": (None, [_D, _D, _D, _D, _D, _D, _L, _D, _D, _DOM, ctypes.c_double]),
    "or_push": (None, [_D, _D, _D, _D, _L, _DOM, ctypes.c_double]),
    "or_integrated_rho": (ctypes.c_double, [_D, _DOM]),
    "or_field_energy": (ctypes.c_double, [_D, _D, _DOM]),
    "or_kinetic_energy": (ctypes.c_double, [_D, _D, _D, _L]),
    "or_total_charge": (ctypes.c_double, [_D, _L]),
    "or_total_abs_charge": (ctypes.c_double, [_D, _L]),
    "or_charge_error": (ctypes.c_double, [ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "or_init_uniform_maxwellian": (None, [_L, _DOM, ctypes.c_double, ctypes.c_uint, _D, _D, _D, _D, _D, _D]),
    "or_init_langmuir": (None, [_L, _DOM, ctypes.c_double, ctypes.c_int, _D, _D, _D, _D, _D, _D]),
    "or_init_two_stream": (None, [_L, _DOM, ctypes.c_double, ctypes.c_double, ctypes.c_uint, ctypes.c_int,
                                  _D, _D, _D, _D, _D, _D]),
    "or_mt19937_words": (None, [ctypes.c_uint, _L, _U]),
    "or_poisson": (None, [_D, _DOM, _D, _D, _D]),
    "or_dominant_frequency": (ctypes.c_double, [_D, _L, ctypes.c_double]),
    "or_plasma_frequency": (ctypes.c_double, [_DOM, _L]),
    "or_resonant_beam_velocity": (ctypes.c_double, [_DOM, _L, ctypes.c_int]),
    "or_fourier_mode_amplitude": (ctypes.c_double, [_D, _DOM, ctypes.c_int]),
    "or_growth_rate": (ctypes.c_double, [_D, _L, ctypes.c_double, ctypes.c_int]),
    "or_esirkepov_current": (ctypes.c_int, [_D, _D, _D, _D, _D, _L, _DOM, ctypes.c_double, _D, _D, _D, _D]),
    "or_cnn_forward": (ctypes.c_int, [_D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I, ctypes.c_int, _D]),
    "or_cnn_backward": (ctypes.c_int, [_D, _D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I, ctypes.c_int, _D,
                                       _D]),
    "or_run_loop": (ctypes.c_int, [_D, _D, _D, _D, _D, _D, _L, _DOM, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D]),
}


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise ImportError(f"{ORACLE_LIB} not built (make -C oracle)")
        _lib = _bind(ctypes.CDLL(ORACLE_LIB), _OR_SIGS)
    return _lib


_REF_SIGS = {
    "ref_deposit_soa": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, _D, _L,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       _D]),
    "ref_deposit_aos": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, _D, _L, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double, _D]),
    "ref_cell_index": (None, [_D, _D, _L, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, _I]),
    "ref_sort_soa": (None, [_D, _D, _D, _D, _D, _D, _L, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                            ctypes.c_int]),
    "ref_gather_push_soa": (None, [ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, _D, _D, _L, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_double]),
    "ref_init_soa": (None, [ctypes.c_int, _L, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                            ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_uint, _D, _D, _D, _D, _D, _D]),
    "ref_deposit_kernel_timed": (ctypes.c_double, [ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, _D, _L,
                                                   ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                   ctypes.c_double, ctypes.c_int, _D, _D]),
    "ref_integrated_rho": (ctypes.c_double, [_D, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]),
    "ref_field_energy": (ctypes.c_double, [_D, _D, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]),
    "ref_total_charge": (ctypes.c_double, [_D, _L]),
    "ref_kinetic_energy": (ctypes.c_double, [_D, _D, _D, _L]),
    "ref_tsc_weights": (None, [ctypes.c_double, _D, _I]),
    "ref_esirkepov_weights_1d": (None, [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, _D]),
}


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref():
    """The reference's own CPU code (opened lazily: its FFTW-dependent symbols
    stay unresolved and are never called)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise ImportError(f"{REF_LIB} not built (make -C oracle ref, needs the reference checkout)")
        _ref = _bind(ctypes.CDLL(REF_LIB, mode=os.RTLD_LAZY), _REF_SIGS)
    return _ref


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _pd(a):
    return None if a is None else a.ctypes.data_as(_D)


def _pi(a):
    return a.ctypes.data_as(_I)


def domain(Lx, Ly, Nx, Ny) -> ODomain:
    return ODomain(float(Lx), float(Ly), int(Nx), int(Ny))


def deposit(scheme, x, y, q, dom: ODomain, vx=None, vy=None, dt=0.0, threads=0) -> np.ndarray:
    """FP64 rho[Ny, Nx].  threads=0: serial path (ngp.cpp); threads>=1: OpenMP
    privatised path (deposition_parallel.cpp)."""
    x, y, q, vx, vy = _d(x), _d(y), _d(q), _d(vx), _d(vy)
    rho = np.zeros(dom.Nx * dom.Ny)
    if threads:
        rc = lib().or_deposit_omp(int(scheme), _pd(x), _pd(y), _pd(vx), _pd(vy), _pd(q), len(x), ctypes.byref(dom),
                                  float(dt), int(threads), _pd(rho))
    else:
        rc = lib().or_deposit(int(scheme), _pd(x), _pd(y), _pd(vx), _pd(vy), _pd(q), len(x), ctypes.byref(dom),
                              float(dt), _pd(rho))
    if rc:
        raise ValueError(f"or_deposit failed ({rc})")
    return rho.reshape(dom.Ny, dom.Nx)


def deposit_omp_timed(scheme, x, y, q, dom, threads, repeats, vx=None, vy=None, dt=0.0):
    x, y, q, vx, vy = _d(x), _d(y), _d(q), _d(vx), _d(vy)
    rho = np.zeros(dom.Nx * dom.Ny)
    sd = ctypes.c_double()
    mean = lib().or_deposit_omp_timed(int(scheme), _pd(x), _pd(y), _pd(vx), _pd(vy), _pd(q), len(x),
                                      ctypes.byref(dom), float(dt), int(threads), int(repeats), _pd(rho),
                                      ctypes.byref(sd))
    return mean, sd.value, rho.reshape(dom.Ny, dom.Nx)


def cell_keys(x, y, dom) -> np.ndarray:
    x, y = _d(x), _d(y)
    k = np.empty(len(x), dtype=np.int32)
    lib().or_cell_keys(_pd(x), _pd(y), len(x), ctypes.byref(dom), _pi(k))
    return k


def sort_perm(keys) -> np.ndarray:
    keys = np.ascontiguousarray(keys, dtype=np.int32)
    p = np.empty(len(keys), dtype=np.int32)
    lib().or_sort_perm(_pi(keys), len(keys), _pi(p))
    return p


def cell_offsets(keys, ncells) -> np.ndarray:
    keys = np.ascontiguousarray(keys, dtype=np.int32)
    off = np.empty(ncells + 1, dtype=np.int32)
    lib().or_cell_offsets(_pi(keys), len(keys), int(ncells), _pi(off))
    return off


def gather(x, y, vx, vy, q, m, Ex, Ey, dom, dt):
    """Returns updated (vx, vy)."""
    x, y, q, m = _d(x), _d(y), _d(q), _d(m)
    vx, vy = _d(vx).copy(), _d(vy).copy()
    Ex, Ey = _d(np.ravel(Ex)), _d(np.ravel(Ey))
    lib().or_gather(_pd(x), _pd(y), _pd(vx), _pd(vy), _pd(q), _pd(m), len(x), _pd(Ex), _pd(Ey), ctypes.byref(dom),
                    float(dt))
    return vx, vy


def push(x, y, vx, vy, dom, dt):
    x, y = _d(x).copy(), _d(y).copy()
    vx, vy = _d(vx), _d(vy)
    lib().or_push(_pd(x), _pd(y), _pd(vx), _pd(vy), len(x), ctypes.byref(dom), float(dt))
    return x, y


def init(kind, n, dom, seed=42, T=1.0, amplitude=0.01, mode=1, beam_velocity=0.3, perturbation=0.01,
         quasi1d=False):
    """Reference initialisers: kind 'maxwellian' | 'langmuir' | 'two_stream'."""
    arrs = [np.empty(n) for _ in range(6)]
    ps = [_pd(a) for a in arrs]
    if kind == "maxwellian":
        lib().or_init_uniform_maxwellian(n, ctypes.byref(dom), float(T), int(seed), *ps)
    elif kind == "langmuir":
        lib().or_init_langmuir(n, ctypes.byref(dom), float(amplitude), int(mode), *ps)
    elif kind == "two_stream":
        lib().or_init_two_stream(n, ctypes.byref(dom), float(beam_velocity), float(perturbation), int(seed),
                                 int(quasi1d), *ps)
    else:
        raise ValueError(kind)
    return dict(zip(("x", "y", "vx", "vy", "q", "m"), arrs))


def poisson(rho, dom):
    rho = _d(np.ravel(rho))
    G = dom.Nx * dom.Ny
    phi, ex, ey = np.empty(G), np.empty(G), np.empty(G)
    lib().or_poisson(_pd(rho), ctypes.byref(dom), _pd(phi), _pd(ex), _pd(ey))
    shp = (dom.Ny, dom.Nx)
    return phi.reshape(shp), ex.reshape(shp), ey.reshape(shp)


def integrated_rho(rho, dom):
    return lib().or_integrated_rho(_pd(_d(np.ravel(rho))), ctypes.byref(dom))


def field_energy(Ex, Ey, dom):
    return lib().or_field_energy(_pd(_d(np.ravel(Ex))), _pd(_d(np.ravel(Ey))), ctypes.byref(dom))


def kinetic_energy(vx, vy, m):
    vx, vy, m = _d(vx), _d(vy), _d(m)
    return lib().or_kinetic_energy(_pd(vx), _pd(vy), _pd(m), len(vx))


def total_charge(q):
    q = _d(q)
    return lib().or_total_charge(_pd(q), len(q))


def total_abs_charge(q):
    q = _d(q)
    return lib().or_total_abs_charge(_pd(q), len(q))


def charge_error(integrated, total, abs_sum):
    return lib().or_charge_error(float(integrated), float(total), float(abs_sum))


def dominant_frequency(signal, dt):
    s = _d(signal)
    return lib().or_dominant_frequency(_pd(s), len(s), float(dt))


def plasma_frequency(dom, n):
    return lib().or_plasma_frequency(ctypes.byref(dom), int(n))


def resonant_beam_velocity(dom, n, mode=1):
    return lib().or_resonant_beam_velocity(ctypes.byref(dom), int(n), int(mode))


def fourier_mode_amplitude(Ex, dom, mode=1):
    return lib().or_fourier_mode_amplitude(_pd(_d(np.ravel(Ex))), ctypes.byref(dom), int(mode))


def growth_rate(signal, dt, record_start):
    s = _d(signal)
    return lib().or_growth_rate(_pd(s), len(s), float(dt), int(record_start))


def run_loop(p: dict, dom, scheme, order, dt, background, steps, record_from):
    """FP64 timestep loop (simulation.cpp / validation.cpp order); returns the
    advanced particles and the recorded series."""
    arrs = {k: _d(p[k]).copy() for k in ("x", "y", "vx", "vy", "q", "m")}
    cap = max(steps - record_from, 0)
    rec = {k: np.zeros(cap) for k in ("field_energy", "mode1", "kinetic", "charge_error")}
    rc = lib().or_run_loop(*[_pd(arrs[k]) for k in ("x", "y", "vx", "vy", "q", "m")], len(arrs["x"]),
                           ctypes.byref(dom), int(scheme), int(order), float(dt), float(background), int(steps),
                           int(record_from), *[_pd(rec[k]) for k in ("field_energy", "mode1", "kinetic",
                                                                      "charge_error")])
    if rc:
        raise ValueError("or_run_loop failed")
    return arrs, rec


def esirkepov_current(x, y, vx, vy, q, dom, dt):
    x, y, vx, vy, q = _d(x), _d(y), _d(vx), _d(vy), _d(q)
    G = dom.Nx * dom.Ny
    outs = [np.empty(G) for _ in range(4)]
    rc = lib().or_esirkepov_current(_pd(x), _pd(y), _pd(vx), _pd(vy), _pd(q), len(x), ctypes.byref(dom), float(dt),
                                    *[_pd(o) for o in outs])
    if rc:
        raise ValueError("move longer than one cell")
    shp = (dom.Ny, dom.Nx)
    return tuple(o.reshape(shp) for o in outs)  # rho_old, rho_new, Jx, Jy


# ---- reference build wrappers ------------------------------------------------

def ref_deposit(scheme, x, y, q, dom, vx=None, vy=None, dt=0.0, threads=1, sorted_=False, layout="SoA"):
    x, y, q, vx, vy = _d(x), _d(y), _d(q), _d(vx), _d(vy)
    rho = np.zeros(dom.Nx * dom.Ny)
    if layout == "SoA":
        ref().ref_deposit_soa(int(scheme), int(threads), int(sorted_), _pd(x), _pd(y), _pd(vx), _pd(vy), _pd(q),
                              len(x), dom.Lx, dom.Ly, dom.Nx, dom.Ny, float(dt), _pd(rho))
    else:
        ref().ref_deposit_aos(int(scheme), int(threads), _pd(x), _pd(y), _pd(vx), _pd(vy), _pd(q), len(x), dom.Lx,
                              dom.Ly, dom.Nx, dom.Ny, float(dt), _pd(rho))
    return rho.reshape(dom.Ny, dom.Nx)


def ref_deposit_timed(scheme, x, y, q, dom, threads, repeats, vx=None, vy=None, dt=0.0):
    """(mean ms, std ms, rho) of the reference's own CPU deposit, timed like
    depositChargeTimedSoA."""
    x, y, q, vx, vy = _d(x), _d(y), _d(q), _d(vx), _d(vy)
    rho = np.zeros(dom.Nx * dom.Ny)
    sd = ctypes.c_double()
    mean = ref().ref_deposit_kernel_timed(int(scheme), int(threads), _pd(x), _pd(y), _pd(vx), _pd(vy), _pd(q), len(x),
                                          dom.Lx, dom.Ly, dom.Nx, dom.Ny, float(dt), int(repeats), ctypes.byref(sd),
                                          _pd(rho))
    return mean, sd.value, rho.reshape(dom.Ny, dom.Nx)


def ref_cell_index(x, y, dom):
    x, y = _d(x), _d(y)
    k = np.empty(len(x), dtype=np.int32)
    ref().ref_cell_index(_pd(x), _pd(y), len(x), dom.Lx, dom.Ly, dom.Nx, dom.Ny, _pi(k))
    return k


def ref_sort(p: dict, dom):
    arrs = [_d(p[k]).copy() for k in ("x", "y", "vx", "vy", "q", "m")]
    ref().ref_sort_soa(*[_pd(a) for a in arrs], len(arrs[0]), dom.Lx, dom.Ly, dom.Nx, dom.Ny)
    return dict(zip(("x", "y", "vx", "vy", "q", "m"), arrs))


def ref_gather_push(p: dict, Ex, Ey, dom, dt, gather=True, push=True):
    arrs = [_d(p[k]).copy() for k in ("x", "y", "vx", "vy", "q", "m")]
    Ex, Ey = _d(np.ravel(Ex)), _d(np.ravel(Ey))
    ref().ref_gather_push_soa(int(gather), int(push), *[_pd(a) for a in arrs], len(arrs[0]), dom.Lx, dom.Ly, dom.Nx,
                              dom.Ny, _pd(Ex), _pd(Ey), float(dt))
    return dict(zip(("x", "y", "vx", "vy", "q", "m"), arrs))


def ref_init(kind, n, dom, p1=0.0, p2=0.0, mode=1, seed=42):
    """kind: 0 Maxwellian(T=p1), 1 Langmuir(amp=p1), 2 two-stream(v=p1, pert=p2),
    3 quasi-1D two-stream, 4 warm Langmuir(amp=p1, T=p2)."""
    arrs = [np.empty(n) for _ in range(6)]
    ref().ref_init_soa(int(kind), int(n), dom.Lx, dom.Ly, dom.Nx, dom.Ny, float(p1), float(p2), int(mode), int(seed),
                       *[_pd(a) for a in arrs])
    return dict(zip(("x", "y", "vx", "vy", "q", "m"), arrs))


def ref_integrated_rho(rho, dom):
    """FieldGrid::integratedRho of the reference build."""
    return ref().ref_integrated_rho(_pd(_d(np.ravel(rho))), dom.Lx, dom.Ly, dom.Nx, dom.Ny)


def ref_field_energy(Ex, Ey, dom):
    """FieldGrid::fieldEnergy of the reference build."""
    return ref().ref_field_energy(_pd(_d(np.ravel(Ex))), _pd(_d(np.ravel(Ey))), dom.Lx, dom.Ly, dom.Nx, dom.Ny)


def ref_total_charge(q):
    """ParticlesSoA::totalCharge of the reference build."""
    q = _d(q)
    return ref().ref_total_charge(_pd(q), len(q))


def ref_kinetic_energy(vx, vy, m):
    """ParticlesSoA::kineticEnergy of the reference build."""
    vx, vy, m = _d(vx), _d(vy), _d(m)
    return ref().ref_kinetic_energy(_pd(vx), _pd(vy), _pd(m), len(vx))


def ref_tsc_weights(xi):
    w = np.empty(3)
    i0 = ctypes.c_int32()
    ref().ref_tsc_weights(float(xi), _pd(w), ctypes.byref(i0))
    return i0.value, w


def ref_esirkepov_weights_1d(x0, x1, dx, n):
    w = np.zeros(n)
    ref().ref_esirkepov_weights_1d(float(x0), float(x1), float(dx), int(n), _pd(w))
    return w


def _r2hc(a, axis):
    """FFTW_R2HC along one axis (forward sign): out[k] = Re X_k for k <= n/2,
    out[n-k] = Im X_k for 0 < k < n/2 (fftw3 documentation, "The Halfcomplex-format
    DFT")."""
    n = a.shape[axis]
    F = np.fft.rfft(a, axis=axis)
    out = np.empty(a.shape, dtype=np.float64)
    idx = [slice(None)] * a.ndim
    for k in range(n // 2 + 1):
        idx[axis] = k
        src = [slice(None)] * a.ndim
        src[axis] = k
        out[tuple(idx)] = F[tuple(src)].real
    for k in range(1, (n + 1) // 2):
        idx[axis] = n - k
        src = [slice(None)] * a.ndim
        src[axis] = k
        out[tuple(idx)] = F[tuple(src)].imag
    return out


def spectral_noise_level(rho):
    """spectralNoiseLevel (src/pic/diagnostics.cpp:53-102): fraction of the
    separable R2HC x R2HC power with |kx| >= Nx/3 or |ky| >= Ny/3."""
    rho = np.asarray(rho, dtype=np.float64)
    ny, nx = rho.shape
    out = _r2hc(_r2hc(rho, 1), 0)
    power = out * out
    kx = np.array([i if i <= nx // 2 else nx - i for i in range(nx)])
    ky = np.array([j if j <= ny // 2 else ny - j for j in range(ny)])
    high = (kx[None, :] >= max(1, nx // 3)) | (ky[:, None] >= max(1, ny // 3))
    total = power.sum()
    if total <= 1e-30:
        return 0.0
    return float(power[high].sum() / total)


CNN_PARAMS = 19105


def cnn_forward(params, images, dilation=(1, 2, 4, 8), emulate_bf16=False) -> np.ndarray:
    """FP64 network output for images [B, H, W] (delta in, delta out)."""
    x = np.ascontiguousarray(images, dtype=np.float64)
    B, H, W = x.shape
    p = _d(params)
    dil = np.ascontiguousarray(dilation, dtype=np.int32)
    out = np.zeros_like(x)
    if lib().or_cnn_forward(_pd(p), _pd(x), B, H, W, _pi(dil), int(bool(emulate_bf16)), _pd(out)):
        raise ValueError("or_cnn_forward failed")
    return out


def cnn_backward(params, images, targets, dilation=(1, 2, 4, 8), emulate_bf16=False):
    """(loss, gradient[CNN_PARAMS]) of the mean-squared error."""
    x = np.ascontiguousarray(images, dtype=np.float64)
    t = np.ascontiguousarray(targets, dtype=np.float64)
    B, H, W = x.shape
    p = _d(params)
    dil = np.ascontiguousarray(dilation, dtype=np.int32)
    loss = np.zeros(1)
    grad = np.zeros(CNN_PARAMS)
    if lib().or_cnn_backward(_pd(p), _pd(x), _pd(t), B, H, W, _pi(dil), int(bool(emulate_bf16)), _pd(loss),
                             _pd(grad)):
        raise ValueError("or_cnn_backward failed")
    return float(loss[0]), grad
