"""ctypes binding of libpic_hip.so (the C ABI in include/pic_hip.h).

This is the Python face of the HIP path: every call goes to the gfx950
kernels.  There is no CPU fallback: if the shared library is missing or cannot
be loaded, importing this module raises.

Names mirror the reference's API (ehocine/pic-deposition):
``DepositionScheme`` / ``DepositionBackend`` (src/deposition/deposition.hpp:11-13),
``scheme_from_string`` / ``backend_from_string`` (deposition.cpp:402-415),
``deposit_charge`` (depositChargeSoA, deposition.cpp:181-200, GPU branch).
"""
from __future__ import annotations

import ctypes
import enum
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpic_hip.so")


class PicError(RuntimeError):
    """A pic_hip call returned a non-zero pic_status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"pic_hip error {status}: {message}")
        self.status = status


class DepositionScheme(enum.IntEnum):
    NGP = 0
    CIC = 1
    TSC = 2
    Esirkepov = 3


class DepositionBackend(enum.IntEnum):
    CPU = 0
    GPUAtomics = 1
    GPUPrivatized = 2


_SCHEME_NAMES = {"NGP": 0, "CIC": 1, "TSC": 2, "Esirkepov": 3}
_BACKEND_NAMES = {"CPU": 0, "GPU_Atomics": 1, "GPU_Priv": 2}


def scheme_from_string(name: str) -> DepositionScheme:
    """schemeFromString (deposition.cpp:402-408): unknown names raise."""
    if name not in _SCHEME_NAMES:
        raise RuntimeError("Unknown deposition scheme: " + name)
    return DepositionScheme(_SCHEME_NAMES[name])


def backend_from_string(name: str) -> DepositionBackend:
    """backendFromString (deposition.cpp:410-415): unknown names raise."""
    if name not in _BACKEND_NAMES:
        raise RuntimeError("Unknown backend: " + name)
    return DepositionBackend(_BACKEND_NAMES[name])


def scheme_name(s: int) -> str:
    return {0: "NGP", 1: "CIC", 2: "TSC", 3: "Esirkepov"}.get(int(s), "Unknown")


def backend_name(b: int) -> str:
    return {0: "CPU", 1: "GPU_Atomics", 2: "GPU_Priv"}.get(int(b), "Unknown")


class Domain(ctypes.Structure):
    """pic_domain / pic::Domain: dx = Lx / Nx."""

    _fields_ = [("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("Nx", ctypes.c_int32), ("Ny", ctypes.c_int32)]

    @property
    def dx(self) -> float:
        return self.Lx / self.Nx

    @property
    def dy(self) -> float:
        return self.Ly / self.Ny

    def cell_volume(self) -> float:
        return self.dx * self.dy


class StepConfig(ctypes.Structure):
    _fields_ = [("scheme", ctypes.c_int32), ("backend", ctypes.c_int32), ("sort_interval", ctypes.c_int32),
                ("order", ctypes.c_int32), ("dt", ctypes.c_double), ("background", ctypes.c_double),
                ("record_particles", ctypes.c_int32), ("use_graph", ctypes.c_int32)]


ORDER_PUSH_FIRST = 0
ORDER_DEPOSIT_SOLVE_PUSH = 1
ORDER_DEPOSIT_PUSH_SOLVE = 2


class Diagnostics(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("rho_sum", "integrated_rho", "field_energy", "kinetic_energy",
                                                "total_charge", "total_abs_charge", "charge_error")]


class TimestepProfile(ctypes.Structure):
    """TimestepProfile (reference src/pic/simulation.hpp:40-48)."""
    _fields_ = [("push_ms", ctypes.c_double), ("sort_ms", ctypes.c_double), ("deposit_ms", ctypes.c_double),
                ("poisson_ms", ctypes.c_double), ("gather_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("sort_interval", ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class InitParams(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("seed", ctypes.c_uint64), ("amp", ctypes.c_double),
                ("pert", ctypes.c_double), ("vth", ctypes.c_double), ("vmax", ctypes.c_double),
                ("mode", ctypes.c_int32), ("q", ctypes.c_double), ("m", ctypes.c_double)]


class ReferenceInit(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("seed", ctypes.c_uint32), ("temperature", ctypes.c_double),
                ("amplitude", ctypes.c_double), ("mode", ctypes.c_int32), ("beam_velocity", ctypes.c_double),
                ("perturbation", ctypes.c_double)]


REFERENCE_INIT_KINDS = {"maxwellian": 0, "langmuir": 1, "warm_langmuir": 2, "two_stream": 3, "two_stream_1d": 4}


class EnsembleMember(ctypes.Structure):
    _fields_ = [("domain", Domain), ("scheme", ctypes.c_int32), ("num_particles", ctypes.c_int64),
                ("init", InitParams)]


class EnsembleRunConfig(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("sort_interval", ctypes.c_int32), ("use_graph", ctypes.c_int32)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `make -C pic-deposition_amd` (or __graft_entry__.build())")
    return ctypes.CDLL(LIB_PATH)


_lib = _load()

_P = ctypes.c_void_p
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)

_SIGS = {
    "pic_hip_last_error": (ctypes.c_char_p, []),
    "pic_hip_version": (ctypes.c_char_p, []),
    "pic_hip_deposit_host": (ctypes.c_int, [_f64p, _f64p, _f64p, _f64p, _f64p, ctypes.c_size_t,
                                            ctypes.POINTER(Domain), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_double, _f64p]),
    "pic_hip_deposit_host_aos": (ctypes.c_int, [_f64p, ctypes.c_size_t, ctypes.POINTER(Domain), ctypes.c_int,
                                                ctypes.c_int, ctypes.c_int, ctypes.c_double, _f64p]),
    "pic_hip_reference_particles": (ctypes.c_int, [ctypes.POINTER(Domain), ctypes.POINTER(ReferenceInit),
                                                   ctypes.c_size_t, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]),
    "pic_hip_cold_two_stream_growth_rate": (ctypes.c_double, [ctypes.POINTER(Domain), ctypes.c_int64]),
    "pic_hip_resonant_beam_velocity": (ctypes.c_double, [ctypes.POINTER(Domain), ctypes.c_int64, ctypes.c_int32]),
    "pic_hip_create": (ctypes.c_int, [ctypes.POINTER(Domain), ctypes.c_size_t, ctypes.POINTER(_P)]),
    "pic_hip_destroy": (ctypes.c_int, [_P]),
    "pic_hip_stream": (_P, [_P]),
    "pic_hip_sync": (ctypes.c_int, [_P]),
    "pic_hip_count": (ctypes.c_size_t, [_P]),
    "pic_hip_set_tiling": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "pic_hip_set_tile_halo": (ctypes.c_int, [_P, ctypes.c_int]),
    "pic_hip_get_tile_halo": (ctypes.c_int, [_P]),
    "pic_hip_ensemble_set_tile_halo": (ctypes.c_int, [_P, ctypes.c_int]),
    "pic_hip_get_tiling": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "pic_hip_upload": (ctypes.c_int, [_P, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_size_t]),
    "pic_hip_download": (ctypes.c_int, [_P, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]),
    "pic_hip_upload_aos": (ctypes.c_int, [_P, _f64p, ctypes.c_size_t]),
    "pic_hip_download_aos": (ctypes.c_int, [_P, _f64p]),
    "pic_hip_init_particles": (ctypes.c_int, [_P, ctypes.POINTER(InitParams), ctypes.c_size_t]),
    "pic_hip_sort": (ctypes.c_int, [_P]),
    "pic_hip_sort_nodes": (ctypes.c_int, [_P]),
    "pic_hip_sort_perm": (ctypes.c_int, [_P, _i32p, _i32p, _i32p]),
    "pic_hip_get_keys": (ctypes.c_int, [_P, _i32p]),
    "pic_hip_sort_path": (ctypes.c_int, [_P, _i32p]),
    "pic_hip_deposit": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
    "pic_hip_get_rho": (ctypes.c_int, [_P, _f32p]),
    "pic_hip_get_rho_f64": (ctypes.c_int, [_P, _f64p]),
    "pic_hip_set_rho": (ctypes.c_int, [_P, _f32p]),
    "pic_hip_add_background": (ctypes.c_int, [_P, ctypes.c_double]),
    "pic_hip_poisson": (ctypes.c_int, [_P]),
    "pic_hip_get_fields": (ctypes.c_int, [_P, _f32p, _f32p, _f32p]),
    "pic_hip_set_fields": (ctypes.c_int, [_P, _f32p, _f32p]),
    "pic_hip_gather_push": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_int, ctypes.c_int]),
    "pic_hip_diagnostics": (ctypes.c_int, [_P, ctypes.POINTER(Diagnostics)]),
    "pic_hip_run": (ctypes.c_int, [_P, ctypes.POINTER(StepConfig), ctypes.c_int32, ctypes.c_int32, _f64p, _f64p,
                                   _f64p, _f64p]),
    "pic_hip_deposit_noalloc": (ctypes.c_int, [ctypes.POINTER(Domain), _P, _P, _P, _P, _P, ctypes.c_size_t,
                                               ctypes.c_int, ctypes.c_double, _P, _P]),
    "pic_hip_run_profiled": (ctypes.c_int, [_P, ctypes.POINTER(StepConfig), ctypes.c_int32, ctypes.c_int32,
                                            ctypes.POINTER(TimestepProfile)]),
    "pic_hip_dominant_frequency": (ctypes.c_double, [_f64p, ctypes.c_int64, ctypes.c_double]),
    "pic_hip_growth_rate": (ctypes.c_double, [_f64p, ctypes.c_int64, ctypes.c_double, ctypes.c_int32]),
    "pic_hip_plasma_frequency": (ctypes.c_double, [ctypes.POINTER(Domain), ctypes.c_int64]),
    "pic_hip_copy_probe": (ctypes.c_int, [_P, ctypes.c_size_t]),
    "pic_hip_read_probe": (ctypes.c_int, [_P, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "pic_hip_esirkepov_current": (ctypes.c_int, [_P, ctypes.c_double]),
    "pic_hip_get_current": (ctypes.c_int, [_P, _f32p, _f32p, _f32p, _f32p]),
    "pic_hip_continuity_residual": (ctypes.c_int, [_P, ctypes.c_double, _f64p, _f64p]),
    "pic_hip_profile_begin": (ctypes.c_int, [_P, ctypes.c_int]),
    "pic_hip_profile_end": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]),
    "pic_hip_autotune": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _f64p]),
    "pic_hip_ensemble_create": (ctypes.c_int, [ctypes.POINTER(EnsembleMember), ctypes.c_int32, ctypes.POINTER(_P)]),
    "pic_hip_ensemble_destroy": (ctypes.c_int, [_P]),
    "pic_hip_ensemble_count": (ctypes.c_int32, [_P]),
    "pic_hip_ensemble_sync": (ctypes.c_int, [_P]),
    "pic_hip_ensemble_init": (ctypes.c_int, [_P]),
    "pic_hip_ensemble_upload": (ctypes.c_int, [_P, ctypes.c_int32, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]),
    "pic_hip_ensemble_download": (ctypes.c_int, [_P, ctypes.c_int32, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]),
    "pic_hip_ensemble_sort": (ctypes.c_int, [_P]),
    "pic_hip_ensemble_deposit": (ctypes.c_int, [_P, ctypes.c_double]),
    "pic_hip_ensemble_get_rho": (ctypes.c_int, [_P, ctypes.c_int32, _f32p]),
    "pic_hip_ensemble_poisson": (ctypes.c_int, [_P]),
    "pic_hip_ensemble_get_fields": (ctypes.c_int, [_P, ctypes.c_int32, _f32p, _f32p]),
    "pic_hip_ensemble_set_tiling": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "pic_hip_ensemble_get_tiling": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "pic_hip_ensemble_autotune": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int]),
    "pic_hip_ensemble_run": (ctypes.c_int, [_P, ctypes.POINTER(EnsembleRunConfig), ctypes.c_int32, ctypes.c_int32,
                                            _f64p, _f64p, _f64p, _f64p]),
    "pic_hip_ensemble_time": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_double, _f64p]),
    "pic_hip_denoiser_create": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _i32p,
                                               ctypes.c_uint64, ctypes.POINTER(_P)]),
    "pic_hip_denoiser_destroy": (ctypes.c_int, [_P]),
    "pic_hip_denoiser_get_params": (ctypes.c_int, [_P, _f32p]),
    "pic_hip_denoiser_set_params": (ctypes.c_int, [_P, _f32p]),
    "pic_hip_denoiser_forward_host": (ctypes.c_int, [_P, _f32p, _f32p, ctypes.c_int32]),
    "pic_hip_denoiser_gradient_host": (ctypes.c_int, [_P, _f32p, _f32p, ctypes.c_int32, _f64p, _f32p]),
    "pic_hip_denoiser_add_sample": (ctypes.c_int, [_P, _P, ctypes.c_int32]),
    "pic_hip_denoiser_pairs": (ctypes.c_int32, [_P]),
    "pic_hip_denoiser_get_scales": (ctypes.c_int, [_P, _f64p, _f64p]),
    "pic_hip_denoiser_set_scales": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double]),
    "pic_hip_denoiser_clear_samples": (ctypes.c_int, [_P]),
    "pic_hip_denoiser_train": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_uint64,
                                              _f64p]),
    "pic_hip_denoise": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_double, _i32p, _f64p]),
    "pic_hip_denoiser_get_output_f64": (ctypes.c_int, [_P, _f64p]),
    "pic_hip_run_learned": (ctypes.c_int, [_P, _P, ctypes.POINTER(StepConfig), ctypes.c_double, ctypes.c_int32,
                                           ctypes.c_int32, _f64p, _f64p, _f64p, _f64p, _i32p]),
    "pic_hip_deposit_subsample": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64]),
}

EXPORTED = tuple(_SIGS)

for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(_lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def _check(status: int) -> None:
    if status != 0:
        raise PicError(status, _lib.pic_hip_last_error().decode())


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def version() -> str:
    return _lib.pic_hip_version().decode()


NOALLOC_MAX_CELLS = 2097152


def deposit_noalloc(domain: Domain, x, y, q, n: int, rho, scheme=DepositionScheme.CIC, vx=None, vy=None,
                    dt: float = 0.0, stream=None) -> None:
    """pic_hip_deposit_noalloc on device addresses (ints) of FP32 arrays."""
    _check(_lib.pic_hip_deposit_noalloc(ctypes.byref(domain), x, y, vx, vy, q, int(n), int(scheme), float(dt), rho,
                                        stream))


def make_domain(Lx: float, Ly: float, Nx: int, Ny: int) -> Domain:
    return Domain(float(Lx), float(Ly), int(Nx), int(Ny))


def deposit_host(x, y, q, domain: Domain, scheme=DepositionScheme.CIC, backend=DepositionBackend.GPUAtomics,
                 sorted_=False, vx=None, vy=None, dt=0.0) -> np.ndarray:
    """pic_hip_deposit_host: FP64 host arrays in, FP64 rho[Ny, Nx] out (synchronous)."""
    n = len(x)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    vx = None if vx is None else np.ascontiguousarray(vx, dtype=np.float64)
    vy = None if vy is None else np.ascontiguousarray(vy, dtype=np.float64)
    rho = np.zeros(domain.Nx * domain.Ny, dtype=np.float64)
    _check(_lib.pic_hip_deposit_host(_ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double), _ptr(vx, ctypes.c_double),
                                     _ptr(vy, ctypes.c_double), _ptr(q, ctypes.c_double), n, ctypes.byref(domain),
                                     int(scheme), int(backend), int(bool(sorted_)), float(dt),
                                     _ptr(rho, ctypes.c_double)))
    return rho.reshape(domain.Ny, domain.Nx)


AOS_FIELDS = ("x", "y", "vx", "vy", "q", "m")  # ParticleAoS, src/pic/particles.hpp:13-20


def to_aos(particles: dict) -> np.ndarray:
    """FP64 records [n, 6] (x, y, vx, vy, q, m) from a dict of columns; missing
    velocities are 0, a missing mass 1 (ParticleAoS defaults)."""
    n = len(particles["x"])
    rec = np.empty((n, 6), dtype=np.float64)
    for j, k in enumerate(AOS_FIELDS):
        rec[:, j] = particles[k] if particles.get(k) is not None else (1.0 if k == "m" else 0.0)
    return rec


def _records(records) -> np.ndarray:
    rec = np.ascontiguousarray(records, dtype=np.float64)
    if rec.ndim != 2 or rec.shape[1] != 6:
        raise ValueError("particle records must have shape [n, 6] (x, y, vx, vy, q, m)")
    if rec.ctypes.data % 16:
        rec = np.require(np.array(rec, copy=True), requirements=["C", "A"])
    return rec


def deposit_host_aos(records, domain: Domain, scheme=DepositionScheme.CIC, backend=DepositionBackend.GPUAtomics,
                     sorted_=False, dt=0.0) -> np.ndarray:
    """pic_hip_deposit_host_aos: FP64 particle records [n, 6] in, FP64 rho[Ny, Nx] out."""
    rec = _records(records)
    rho = np.zeros(domain.Nx * domain.Ny, dtype=np.float64)
    _check(_lib.pic_hip_deposit_host_aos(_ptr(rec, ctypes.c_double), rec.shape[0], ctypes.byref(domain), int(scheme),
                                         int(backend), int(bool(sorted_)), float(dt), _ptr(rho, ctypes.c_double)))
    return rho.reshape(domain.Ny, domain.Nx)


@dataclass
class Tiling:
    tile_x: int
    tile_y: int
    particles_per_lane: int
    pipelined: int


class Context:
    """A device-resident PIC instance (pic_hip_ctx)."""

    def __init__(self, domain: Domain, capacity: int):
        self.domain = domain
        h = _P()
        _check(_lib.pic_hip_create(ctypes.byref(domain), int(capacity), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.pic_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    @property
    def stream(self) -> int:
        return _lib.pic_hip_stream(self._h) or 0

    @property
    def count(self) -> int:
        return _lib.pic_hip_count(self._h)

    def sync(self) -> None:
        _check(_lib.pic_hip_sync(self._h))

    def set_tiling(self, tile_x: int, tile_y: int, particles_per_lane: int = 8, pipelined: int = 0,
                   scheme: int = -1) -> None:
        _check(_lib.pic_hip_set_tiling(self._h, int(scheme), tile_x, tile_y, particles_per_lane, int(pipelined)))

    def set_tile_halo(self, cells: int) -> None:
        _check(_lib.pic_hip_set_tile_halo(self._h, int(cells)))

    @property
    def tile_halo(self) -> int:
        return _lib.pic_hip_get_tile_halo(self._h)

    def tiling(self, scheme: int = 1) -> Tiling:
        a, b, c, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(_lib.pic_hip_get_tiling(self._h, int(scheme), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                       ctypes.byref(d)))
        return Tiling(a.value, b.value, c.value, d.value)

    def upload(self, x, y, q, vx=None, vy=None, m=None) -> None:
        arrs = [_f32(a) for a in (x, y, vx, vy, q, m)]
        n = len(arrs[0])
        for a in arrs:
            if a is not None and len(a) != n:
                raise ValueError("all particle arrays must have the same length")
        _check(_lib.pic_hip_upload(self._h, *[_ptr(a, ctypes.c_float) for a in arrs], n))

    def download(self):
        n = self.count
        out = [np.empty(n, dtype=np.float32) for _ in range(6)]
        _check(_lib.pic_hip_download(self._h, *[_ptr(a, ctypes.c_float) for a in out]))
        return dict(zip(("x", "y", "vx", "vy", "q", "m"), out))

    def upload_aos(self, records) -> None:
        """FP64 particle records [n, 6] (x, y, vx, vy, q, m), converted on the device."""
        rec = _records(records)
        _check(_lib.pic_hip_upload_aos(self._h, _ptr(rec, ctypes.c_double), rec.shape[0]))

    def download_aos(self) -> np.ndarray:
        rec = np.empty((self.count, 6), dtype=np.float64)
        _check(_lib.pic_hip_download_aos(self._h, _ptr(rec, ctypes.c_double)))
        return rec

    def init_particles(self, n: int, kind: int = 0, seed: int = 42, amp: float = 0.0, pert: float = 0.0,
                       vth: float = 0.0, vmax: float = 0.0, mode: int = 1, q: float = 1.0, m: float = 1.0) -> None:
        p = InitParams(kind, seed, amp, pert, vth, vmax, mode, q, m)
        _check(_lib.pic_hip_init_particles(self._h, ctypes.byref(p), int(n)))

    def sort(self) -> None:
        _check(_lib.pic_hip_sort(self._h))

    def sort_nodes(self) -> None:
        """pic_hip_sort_nodes: the node-keyed order (NGP / TSC tile kernels)."""
        _check(_lib.pic_hip_sort_nodes(self._h))

    def sort_perm(self):
        n = self.count
        keys = np.empty(n, dtype=np.int32)
        perm = np.empty(n, dtype=np.int32)
        off = np.empty(self.domain.Nx * self.domain.Ny + 1, dtype=np.int32)
        _check(_lib.pic_hip_sort_perm(self._h, _ptr(keys, ctypes.c_int32), _ptr(perm, ctypes.c_int32),
                                      _ptr(off, ctypes.c_int32)))
        return keys, perm, off

    def sort_path(self) -> int:
        """pic_hip_sort_path: how the last sort moved the particles (0 radix
        passes carrying the streams, 2 copy of a sorted order, -1 unstaged /
        none)."""
        v = ctypes.c_int32(-1)
        _check(_lib.pic_hip_sort_path(self._h, ctypes.byref(v)))
        return int(v.value)

    def keys(self) -> np.ndarray:
        k = np.empty(self.count, dtype=np.int32)
        _check(_lib.pic_hip_get_keys(self._h, _ptr(k, ctypes.c_int32)))
        return k

    def deposit(self, scheme=DepositionScheme.CIC, backend=DepositionBackend.GPUPrivatized, dt: float = 0.0) -> None:
        _check(_lib.pic_hip_deposit(self._h, int(scheme), int(backend), float(dt)))

    def rho(self) -> np.ndarray:
        r = np.empty(self.domain.Nx * self.domain.Ny, dtype=np.float32)
        _check(_lib.pic_hip_get_rho(self._h, _ptr(r, ctypes.c_float)))
        return r.reshape(self.domain.Ny, self.domain.Nx)

    def set_rho(self, rho) -> None:
        r = np.ascontiguousarray(rho, dtype=np.float32).ravel()
        _check(_lib.pic_hip_set_rho(self._h, _ptr(r, ctypes.c_float)))

    def add_background(self, value: float) -> None:
        _check(_lib.pic_hip_add_background(self._h, float(value)))

    def poisson(self) -> None:
        _check(_lib.pic_hip_poisson(self._h))

    def fields(self):
        G = self.domain.Nx * self.domain.Ny
        phi, ex, ey = (np.empty(G, dtype=np.float32) for _ in range(3))
        _check(_lib.pic_hip_get_fields(self._h, _ptr(phi, ctypes.c_float), _ptr(ex, ctypes.c_float),
                                       _ptr(ey, ctypes.c_float)))
        shp = (self.domain.Ny, self.domain.Nx)
        return phi.reshape(shp), ex.reshape(shp), ey.reshape(shp)

    def set_fields(self, ex, ey) -> None:
        a, b = _f32(np.ravel(ex)), _f32(np.ravel(ey))
        _check(_lib.pic_hip_set_fields(self._h, _ptr(a, ctypes.c_float), _ptr(b, ctypes.c_float)))

    def gather_push(self, dt: float, gather: bool = True, push: bool = True) -> None:
        _check(_lib.pic_hip_gather_push(self._h, float(dt), int(gather), int(push)))

    def esirkepov_current(self, dt: float) -> None:
        _check(_lib.pic_hip_esirkepov_current(self._h, float(dt)))

    def current(self):
        """(Jx, Jy, rho_old, rho_new) as [Ny, Nx] FP32 arrays."""
        G = self.domain.Nx * self.domain.Ny
        out = [np.empty(G, dtype=np.float32) for _ in range(4)]
        _check(_lib.pic_hip_get_current(self._h, *[_ptr(a, ctypes.c_float) for a in out]))
        return tuple(a.reshape(self.domain.Ny, self.domain.Nx) for a in out)

    def continuity_residual(self, dt: float):
        r, m = np.zeros(1), np.zeros(1)
        _check(_lib.pic_hip_continuity_residual(self._h, float(dt), _ptr(r, ctypes.c_double), _ptr(m, ctypes.c_double)))
        return float(r[0]), float(m[0])

    def copy_probe(self, nbytes: int) -> None:
        _check(_lib.pic_hip_copy_probe(self._h, int(nbytes)))

    def read_probe(self, nbytes: int, reps: int = 10) -> float:
        """pic_hip_read_probe: measured streaming-read bandwidth in GB/s."""
        g = ctypes.c_double()
        _check(_lib.pic_hip_read_probe(self._h, int(nbytes), int(reps), ctypes.byref(g)))
        return float(g.value)

    def profile_begin(self, max_launches: int) -> None:
        _check(_lib.pic_hip_profile_begin(self._h, int(max_launches)))

    def profile_end(self):
        """(summed main-kernel milliseconds, launches) since profile_begin."""
        ms, k = ctypes.c_double(), ctypes.c_int()
        _check(_lib.pic_hip_profile_end(self._h, ctypes.byref(ms), ctypes.byref(k)))
        return ms.value, k.value

    def run(self, steps: int, scheme=DepositionScheme.CIC, backend=DepositionBackend.GPUPrivatized, dt=0.01,
            background=0.0, order=ORDER_PUSH_FIRST, sort_interval=1, record_from=0, record_particles=False,
            use_graph=False):
        """pic_hip_run: returns dict of recorded FP64 series (length steps - record_from)."""
        cfg = StepConfig(int(scheme), int(backend), int(sort_interval), int(order), float(dt), float(background),
                         int(bool(record_particles)), int(bool(use_graph)))
        cap = max(steps - record_from, 0)
        out = {k: np.zeros(cap) for k in ("field_energy", "mode1", "kinetic", "charge_error")}
        ptrs = [_ptr(out[k], ctypes.c_double) for k in ("field_energy", "mode1")]
        ptrs += [_ptr(out[k], ctypes.c_double) if record_particles else None for k in ("kinetic", "charge_error")]
        _check(_lib.pic_hip_run(self._h, ctypes.byref(cfg), int(steps), int(record_from), *ptrs))
        if not record_particles:
            del out["kinetic"], out["charge_error"]
        return out

    def run_profiled(self, warmup: int, timed: int, scheme=DepositionScheme.CIC,
                     backend=DepositionBackend.GPUPrivatized, dt=0.01, background=0.0, sort_interval=1) -> dict:
        """pic_hip_run_profiled (Simulation::runProfiled): mean ms per timed
        step of push, sort, deposit, Poisson and gather, plus their total."""
        cfg = StepConfig(int(scheme), int(backend), int(sort_interval), ORDER_PUSH_FIRST, float(dt), float(background),
                         0, 0)
        out = TimestepProfile()
        _check(_lib.pic_hip_run_profiled(self._h, ctypes.byref(cfg), int(warmup), int(timed), ctypes.byref(out)))
        return out.as_dict()

    def diagnostics(self) -> Diagnostics:
        d = Diagnostics()
        _check(_lib.pic_hip_diagnostics(self._h, ctypes.byref(d)))
        return d

    def deposit_subsample(self, scheme, stride: int, salt: int) -> None:
        """rho from a hash-selected 1/stride subset of the particles, charges x stride."""
        _check(_lib.pic_hip_deposit_subsample(self._h, int(scheme), int(stride), int(salt)))

    def autotune(self, scheme=DepositionScheme.CIC, cache_path=None, reps: int = 5):
        """pic_hip_autotune: (Tiling chosen, ms per deposit or 0.0 when read from the cache)."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = np.zeros(1)
        path = None if cache_path is None else os.fsencode(cache_path)
        _check(_lib.pic_hip_autotune(self._h, int(scheme), path, int(reps), ctypes.byref(a), ctypes.byref(b),
                                     ctypes.byref(c), _ptr(ms, ctypes.c_double)))
        return Tiling(a.value, b.value, c.value, 0), float(ms[0])


@dataclass
class Member:
    """One ensemble member (pic_ensemble_member)."""
    Nx: int
    Ny: int
    num_particles: int
    scheme: int = DepositionScheme.CIC
    Lx: float = 1.0
    Ly: float = 1.0
    kind: int = 1
    seed: int = 42
    amp: float = 0.0
    pert: float = 0.0
    vth: float = 0.0
    vmax: float = 0.0
    mode: int = 1
    q: float = 1.0
    m: float = 1.0

    def domain(self) -> Domain:
        return make_domain(self.Lx, self.Ly, self.Nx, self.Ny)

    def _c(self) -> EnsembleMember:
        return EnsembleMember(self.domain(), int(self.scheme), int(self.num_particles),
                              InitParams(self.kind, self.seed, self.amp, self.pert, self.vth, self.vmax, self.mode,
                                         self.q, self.m))


def unit_plasma_charge(n: int, Nx: int, Ny: int) -> float:
    """Electron charge giving plasma frequency 1 with m = 1: the Poisson solve
    follows the reference's normalisation (phi = cellVolume * rho / k^2 per
    mode), so omega_p^2 = n0 q^2 cellVolume / m = n q^2 / (Nx Ny m)."""
    return -float(np.sqrt(Nx * Ny / n))


def langmuir_ensemble(count: int = 64, grids=(64, 128, 256, 512), ppc=(16, 256), amp=(0.001, 0.05),
                      schemes=(DepositionScheme.CIC, DepositionScheme.TSC), seed: int = 2024, mode: int = 1,
                      unit_plasma_frequency: bool = False):
    """Members of a Langmuir ensemble: square grids drawn from `grids`, particles
    per cell uniform in [ppc[0], ppc[1]], density perturbation amplitude
    uniform in [amp[0], amp[1]] on mode `mode`, scheme drawn from `schemes`,
    seeds from one generator; unit box, cold electrons with m = 1 and
    q = -1/N (the reference's Langmuir normalisation, validation.cpp:200-261)
    or, with unit_plasma_frequency, the charge of unit_plasma_charge."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        g = int(rng.choice(grids))
        n = g * g * int(rng.integers(ppc[0], ppc[1] + 1))
        out.append(Member(Nx=g, Ny=g, num_particles=n, scheme=int(rng.choice(schemes)), kind=1,
                          seed=int(rng.integers(1, 2**62)), amp=float(rng.uniform(*amp)), mode=mode,
                          q=unit_plasma_charge(n, g, g) if unit_plasma_frequency else -1.0 / n, m=1.0))
    return out


class Ensemble:
    """A batch of independent PIC instances on one device (pic_hip_ensemble)."""

    def __init__(self, members):
        self.members = list(members)
        arr = (EnsembleMember * len(self.members))(*[m._c() for m in self.members])
        h = _P()
        _check(_lib.pic_hip_ensemble_create(arr, len(self.members), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.pic_hip_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self) -> int:
        return _lib.pic_hip_ensemble_count(self._h)

    def sync(self) -> None:
        _check(_lib.pic_hip_ensemble_sync(self._h))

    def init(self) -> None:
        _check(_lib.pic_hip_ensemble_init(self._h))

    def upload(self, i: int, x, y, q, vx=None, vy=None, m=None) -> None:
        n = self.members[i].num_particles
        arrs = [_f32(a) for a in (x, y, vx, vy, q, m)]
        for a in arrs:
            if a is not None and len(a) != n:
                raise ValueError(f"member {i} holds {n} particles")
        _check(_lib.pic_hip_ensemble_upload(self._h, int(i), *[_ptr(a, ctypes.c_float) for a in arrs]))

    def download(self, i: int):
        n = self.members[i].num_particles
        out = [np.empty(n, dtype=np.float32) for _ in range(6)]
        _check(_lib.pic_hip_ensemble_download(self._h, int(i), *[_ptr(a, ctypes.c_float) for a in out]))
        return dict(zip(("x", "y", "vx", "vy", "q", "m"), out))

    def sort(self) -> None:
        _check(_lib.pic_hip_ensemble_sort(self._h))

    def deposit(self, dt: float = 0.0) -> None:
        _check(_lib.pic_hip_ensemble_deposit(self._h, float(dt)))

    def rho(self, i: int) -> np.ndarray:
        mb = self.members[i]
        r = np.empty(mb.Nx * mb.Ny, dtype=np.float32)
        _check(_lib.pic_hip_ensemble_get_rho(self._h, int(i), _ptr(r, ctypes.c_float)))
        return r.reshape(mb.Ny, mb.Nx)

    def poisson(self) -> None:
        _check(_lib.pic_hip_ensemble_poisson(self._h))

    def fields(self, i: int):
        mb = self.members[i]
        ex, ey = (np.empty(mb.Nx * mb.Ny, dtype=np.float32) for _ in range(2))
        _check(_lib.pic_hip_ensemble_get_fields(self._h, int(i), _ptr(ex, ctypes.c_float), _ptr(ey, ctypes.c_float)))
        return ex.reshape(mb.Ny, mb.Nx), ey.reshape(mb.Ny, mb.Nx)

    def set_tiling(self, tile_x: int, tile_y: int, particles_per_lane: int = 8, member: int = -1) -> None:
        _check(_lib.pic_hip_ensemble_set_tiling(self._h, int(member), tile_x, tile_y, particles_per_lane))

    def set_tile_halo(self, cells: int) -> None:
        _check(_lib.pic_hip_ensemble_set_tile_halo(self._h, int(cells)))

    def tiling(self, i: int) -> Tiling:
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(_lib.pic_hip_ensemble_get_tiling(self._h, int(i), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return Tiling(a.value, b.value, c.value, 0)

    def autotune(self, cache_path=None, reps: int = 5) -> None:
        path = None if cache_path is None else os.fsencode(cache_path)
        _check(_lib.pic_hip_ensemble_autotune(self._h, path, int(reps)))

    def run(self, steps: int, dt: float, sort_interval: int = 1, record_from: int = 0, use_graph: bool = True):
        """pic_hip_ensemble_run: dict of [members, steps - record_from] FP64 records."""
        cfg = EnsembleRunConfig(float(dt), int(sort_interval), int(bool(use_graph)))
        cap = max(steps - record_from, 0)
        M = len(self.members)
        out = {k: np.zeros((M, cap)) for k in ("field_energy", "mode1", "kinetic", "charge_error")}
        _check(_lib.pic_hip_ensemble_run(self._h, ctypes.byref(cfg), int(steps), int(record_from),
                                         *[_ptr(out[k], ctypes.c_double) for k in out]))
        return out

    def time(self, what: str, reps: int = 10, dt: float = 0.0) -> float:
        """Mean milliseconds of 'deposit', 'sort', 'step' or 'step_sort' (HIP events)."""
        code = {"deposit": 0, "sort": 1, "step": 2, "step_sort": 3}[what]
        ms = np.zeros(1)
        _check(_lib.pic_hip_ensemble_time(self._h, code, int(reps), float(dt), _ptr(ms, ctypes.c_double)))
        return float(ms[0])


def dominant_frequency(signal, dt: float) -> float:
    s = np.ascontiguousarray(signal, dtype=np.float64)
    return _lib.pic_hip_dominant_frequency(_ptr(s, ctypes.c_double), len(s), float(dt))


def growth_rate(signal, dt: float, record_start: int) -> float:
    s = np.ascontiguousarray(signal, dtype=np.float64)
    return _lib.pic_hip_growth_rate(_ptr(s, ctypes.c_double), len(s), float(dt), int(record_start))


def plasma_frequency(domain: Domain, n: int) -> float:
    return _lib.pic_hip_plasma_frequency(ctypes.byref(domain), int(n))


def cold_two_stream_growth_rate(domain: Domain, n: int) -> float:
    return _lib.pic_hip_cold_two_stream_growth_rate(ctypes.byref(domain), int(n))


def resonant_beam_velocity(domain: Domain, n: int, mode: int = 1) -> float:
    return _lib.pic_hip_resonant_beam_velocity(ctypes.byref(domain), int(n), int(mode))


def reference_particles(kind: str, n: int, domain: Domain, seed: int = 42, temperature: float = 1.0,
                        amplitude: float = 0.0, mode: int = 1, beam_velocity: float = 0.0,
                        perturbation: float = 0.0) -> dict:
    """The reference's initialisers (particles_aos.cpp:67-180), FP64 columns
    x, y, vx, vy, q, m; kind in REFERENCE_INIT_KINDS."""
    p = ReferenceInit(REFERENCE_INIT_KINDS[kind], int(seed), float(temperature), float(amplitude), int(mode),
                      float(beam_velocity), float(perturbation))
    cols = [np.empty(n, dtype=np.float64) for _ in range(6)]
    _check(_lib.pic_hip_reference_particles(ctypes.byref(domain), ctypes.byref(p), int(n),
                                            *[_ptr(a, ctypes.c_double) for a in cols]))
    return dict(zip(("x", "y", "vx", "vy", "q", "m"), cols))


def deposit_charge(particles: dict, domain: Domain, scheme="CIC", backend="GPU_Priv", sorted_=False,
                   esirkepov_dt=0.0) -> np.ndarray:
    """GPU branch of depositChargeSoA (deposition.cpp:181-200): host FP64 SoA in,
    rho (FP64) out; the caller's particle order is not changed."""
    s = scheme_from_string(scheme) if isinstance(scheme, str) else DepositionScheme(scheme)
    b = backend_from_string(backend) if isinstance(backend, str) else DepositionBackend(backend)
    if b == DepositionBackend.CPU:
        raise RuntimeError("CPU backend requested from the HIP library; the CPU schemes live in the host code")
    return deposit_host(particles["x"], particles["y"], particles["q"], domain, s, b, sorted_,
                        particles.get("vx"), particles.get("vy"), esirkepov_dt)


def deposit_charge_aos(records, domain: Domain, scheme="CIC", backend="GPU_Priv", sorted_=False,
                       esirkepov_dt=0.0) -> np.ndarray:
    """GPU branch of depositChargeAoS (deposition.cpp:160-179): FP64 particle
    records [n, 6] in, rho (FP64) out; the records are not changed."""
    s = scheme_from_string(scheme) if isinstance(scheme, str) else DepositionScheme(scheme)
    b = backend_from_string(backend) if isinstance(backend, str) else DepositionBackend(backend)
    if b == DepositionBackend.CPU:
        raise RuntimeError("CPU backend requested from the HIP library; the CPU schemes live in the host code")
    return deposit_host_aos(records, domain, s, b, sorted_, esirkepov_dt)


DENOISER_PARAMS = 19105


class Denoiser:
    """Learned density denoiser (pic_hip_denoiser): delta = rho / mean - 1 in and out."""

    def __init__(self, Nx: int, Ny: int, max_batch: int = 8, max_pairs: int = 64, dilation=(1, 3, 9, 27),
                 seed: int = 1):
        self.Nx, self.Ny = int(Nx), int(Ny)
        d = np.ascontiguousarray(dilation, dtype=np.int32)
        h = _P()
        _check(_lib.pic_hip_denoiser_create(self.Nx, self.Ny, int(max_batch), int(max_pairs),
                                            _ptr(d, ctypes.c_int32), int(seed), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.pic_hip_denoiser_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def params(self) -> np.ndarray:
        p = np.empty(DENOISER_PARAMS, dtype=np.float32)
        _check(_lib.pic_hip_denoiser_get_params(self._h, _ptr(p, ctypes.c_float)))
        return p

    def set_params(self, p) -> None:
        a = np.ascontiguousarray(p, dtype=np.float32)
        if a.size != DENOISER_PARAMS:
            raise ValueError(f"expected {DENOISER_PARAMS} parameters")
        _check(_lib.pic_hip_denoiser_set_params(self._h, _ptr(a, ctypes.c_float)))

    def forward(self, images) -> np.ndarray:
        x = np.ascontiguousarray(images, dtype=np.float32)
        out = np.empty_like(x)
        _check(_lib.pic_hip_denoiser_forward_host(self._h, _ptr(x, ctypes.c_float), _ptr(out, ctypes.c_float),
                                                  x.shape[0]))
        return out

    def gradient(self, images, targets):
        x = np.ascontiguousarray(images, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        loss = np.zeros(1)
        g = np.empty(DENOISER_PARAMS, dtype=np.float32)
        _check(_lib.pic_hip_denoiser_gradient_host(self._h, _ptr(x, ctypes.c_float), _ptr(t, ctypes.c_float),
                                                   x.shape[0], _ptr(loss, ctypes.c_double), _ptr(g, ctypes.c_float)))
        return float(loss[0]), g

    def add_sample(self, ctx: Context, is_target: bool) -> None:
        _check(_lib.pic_hip_denoiser_add_sample(self._h, ctx.handle, int(bool(is_target))))

    def scales(self):
        a, b = np.zeros(1), np.zeros(1)
        _check(_lib.pic_hip_denoiser_get_scales(self._h, _ptr(a, ctypes.c_double), _ptr(b, ctypes.c_double)))
        return float(a[0]), float(b[0])

    def set_scales(self, in_scale: float, out_scale: float) -> None:
        _check(_lib.pic_hip_denoiser_set_scales(self._h, float(in_scale), float(out_scale)))

    @property
    def pairs(self) -> int:
        return _lib.pic_hip_denoiser_pairs(self._h)

    def clear(self) -> None:
        _check(_lib.pic_hip_denoiser_clear_samples(self._h))

    def train(self, iterations: int, batch: int = 8, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        hist = np.zeros(int(iterations))
        _check(_lib.pic_hip_denoiser_train(self._h, int(iterations), int(batch), float(lr), int(seed),
                                           _ptr(hist, ctypes.c_double)))
        return hist

    def denoise(self, ctx: Context, scheme=DepositionScheme.TSC, tolerance: float = 4.0, report: bool = True):
        """Replaces ctx's rho by the learned density unless the gate falls back.
        With report, returns (fallback, residual_ratio) (synchronous)."""
        if not report:
            _check(_lib.pic_hip_denoise(ctx.handle, self._h, int(scheme), float(tolerance), None, None))
            return None
        fb = np.zeros(1, dtype=np.int32)
        r = np.zeros(1)
        _check(_lib.pic_hip_denoise(ctx.handle, self._h, int(scheme), float(tolerance), _ptr(fb, ctypes.c_int32),
                                    _ptr(r, ctypes.c_double)))
        return bool(fb[0]), float(r[0])

    def run(self, ctx: Context, steps: int, scheme=DepositionScheme.TSC, backend=DepositionBackend.GPUPrivatized,
            dt=0.1, background=0.0, order=ORDER_PUSH_FIRST, sort_interval=10, record_from=0, use_graph=True,
            tolerance: float = 4.0):
        """pic_hip_run_learned: the step loop with the learned density after each
        deposit.  Returns (records dict as Context.run with particles, fallbacks)."""
        cfg = StepConfig(int(scheme), int(backend), int(sort_interval), int(order), float(dt), float(background), 1,
                         int(bool(use_graph)))
        cap = max(steps - record_from, 0)
        out = {k: np.zeros(cap) for k in ("field_energy", "mode1", "kinetic", "charge_error")}
        fb = np.zeros(1, dtype=np.int32)
        _check(_lib.pic_hip_run_learned(ctx.handle, self._h, ctypes.byref(cfg), float(tolerance), int(steps),
                                        int(record_from), *[_ptr(out[k], ctypes.c_double) for k in out],
                                        _ptr(fb, ctypes.c_int32)))
        return out, int(fb[0])

    def output_f64(self) -> np.ndarray:
        out = np.empty(self.Nx * self.Ny)
        _check(_lib.pic_hip_denoiser_get_output_f64(self._h, _ptr(out, ctypes.c_double)))
        return out.reshape(self.Ny, self.Nx)


def langmuir_delta_truth(Nx: int, Ny: int, amp: float, mode: int, scheme=DepositionScheme.TSC) -> np.ndarray:
    """Expected delta = rho / mean - 1 of a CIC or TSC deposit of particles
    drawn from n(x) ~ 1 + amp cos(2 pi mode x / Lx) (init kind 1): node i sits
    at x = i dx and the shape function's Fourier factor is sinc(k dx / 2)^2
    (CIC) or ^3 (TSC)."""
    order = {DepositionScheme.CIC: 2, DepositionScheme.TSC: 3}[int(scheme)]
    kdx = 2 * np.pi * mode / Nx
    form = (np.sin(kdx / 2) / (kdx / 2)) ** order
    row = amp * form * np.cos(kdx * np.arange(Nx))
    return np.tile(row, (Ny, 1))
