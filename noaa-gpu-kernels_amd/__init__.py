"""Python host bindings for libtrmc.so -- Muskingum-Cunge routing on MI355X.

This is a thin ctypes layer over the C-ABI declared in include/trmc.h.  Every
routing call runs the hand-written gfx950 HIP kernels; there is no CPU
fallback, and a missing or unloadable library raises ``TrmcError``.

Names mirror the reference's host interface (consigcody94/noaa-gpu-kernels,
water-prediction/t-route/wavefront/wavefront_mc.cu):

=========================  ==================================================
``wavefront_fp64``         cpu_wavefront_fp64 (wavefront_mc.cu:495-564)
``wavefront_f64``          same call, FP64 inputs end to end
``reach_parallel``         cpu_mc / kernel_mc_gpu (troute_mc_final.cu:142-252)
``mc_solve``               mc_solve_d (wavefront_mc.cu:215-337), batched
``load_topo``              load_topo (wavefront_mc.cu:599-632)
``load_params``            params.bin read (wavefront_mc.cu:654-656)
``load_forcings``          forcings.bin read (wavefront_mc.cu:658-672)
``compute_stats``          accuracy block (wavefront_mc.cu:822-858)
``Network``                [new] device-resident handle (trmc_network_*)
=========================  ==================================================
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtrmc.so")

OK = 0
ERR_ARG, ERR_TOPOLOGY, ERR_HIP, ERR_NO_DEVICE, ERR_TIMEOUT = -1, -2, -3, -4, -5
ERR_NOMEM, ERR_IO, ERR_STATE, ERR_NOT_CONVERGED = -6, -7, -8, -9
MODE_REFERENCE, MODE_CONVERGED, MODE_LEARNED, MODE_LINEAR, MODE_REFERENCE_F32 = 0, 1, 2, 3, 4
SCHED_AUTO, SCHED_SKEWED, SCHED_LEVEL, SCHED_DATAFLOW = 0, 1, 2, 3

PARAM_FIELDS = ("dx", "bw", "tw", "twcc", "n_ch", "ncc", "cs", "s0")


class TrmcError(RuntimeError):
    """A C-ABI call returned a negative status (code in ``.code``)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"trmc error {code}: {msg}")
        self.code = code


class ChParamsF32(C.Structure):
    _fields_ = [(n, C.c_float) for n in PARAM_FIELDS]


class ChParamsF64(C.Structure):
    _fields_ = [(n, C.c_double) for n in PARAM_FIELDS]


class Csr(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_int), ("rowptr", C.POINTER(C.c_int)), ("col", C.POINTER(C.c_int)),
                ("val", C.POINTER(C.c_float))]


class Options(C.Structure):
    _fields_ = [
        ("mode", C.c_int), ("schedule", C.c_int), ("unit_segments", C.c_int),
        ("block_size", C.c_int), ("n_members", C.c_int), ("record_history", C.c_int),
        ("tol", C.c_double), ("max_iter", C.c_int), ("learned_budget", C.c_int),
        ("grid_blocks", C.c_int), ("dataflow_variant", C.c_int), ("device", C.c_int),
        ("trace", C.c_int), ("batch_iters", C.c_int), ("guess_model", C.c_int),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double), ("wall_ms", C.c_double), ("seg_steps", C.c_longlong),
        ("secant_iters", C.c_longlong), ("fallbacks", C.c_longlong),
        ("not_converged", C.c_longlong), ("launches", C.c_int), ("barriers", C.c_int),
        ("super_levels", C.c_int), ("grid_blocks", C.c_int), ("block_size", C.c_int),
        ("barrier_wait_frac", C.c_double), ("mlp_calls", C.c_longlong), ("mlp_frac", C.c_double),
        ("certified_solves", C.c_longlong), ("noncanonical", C.c_longlong),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Topo(C.Structure):
    _fields_ = [("n_reaches", C.c_int), ("n_segments", C.c_int), ("n_levels", C.c_int),
                ("max_upstream", C.c_int)] + [
        (n, C.POINTER(C.c_int)) for n in (
            "reach_seg_start", "reach_seg_len", "reach_level", "reach_n_up",
            "reach_up_ptr", "reach_up_idx", "level_ptr", "level_reach")]


_lib = None


def lib() -> C.CDLL:
    """Load libtrmc.so (built by ``make -C noaa-gpu-kernels_amd/csrc``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TrmcError(ERR_STATE, f"{LIB_PATH} is not built")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _p(t):
    return C.POINTER(t)


def _declare(L: C.CDLL) -> None:
    I, D, V = C.c_int, C.c_double, C.c_void_p
    pI, pD, pF = _p(C.c_int), _p(C.c_double), _p(C.c_float)
    topo = [pI] * 6 + [I, I, I]
    sig = {
        "trmc_last_error": (C.c_char_p, []),
        "trmc_version": (I, []),
        "trmc_device_count": (I, [pI]),
        "trmc_default_options": (None, [_p(Options)]),
        "trmc_device_synchronize": (I, [I]),
        "trmc_wavefront_f32in_f64": (I, [_p(ChParamsF32)] + topo + [pF, pF, pF, pF, pD, pD, I, D]),
        "trmc_wavefront_f32": (I, [_p(ChParamsF32)] + topo + [pF, pF, pF, pF, pF, pF, I, C.c_float]),
        "trmc_wavefront_f64": (I, [_p(ChParamsF64)] + topo + [pD, pD, pD, pD, pD, pD, I, D, _p(Options)]),
        "trmc_reach_parallel_f64": (I, [_p(ChParamsF64), pD, pI, pI, pD, pD, pD, pD, pD, pD, D, D, I]),
        "trmc_mc_solve_f64": (I, [_p(ChParamsF64), pD, pD, pI, I, D, I]),
        "trmc_plan_build": (I, topo + [I, _p(V)]),
        "trmc_plan_destroy": (None, [V]),
        "trmc_plan_info": (I, [V, I, I, _p(C.c_int64), I]),
        "trmc_plan_export": (I, [V, I, pI, I]),
        "trmc_network_create": (I, topo + [_p(Options), _p(V)]),
        "trmc_network_destroy": (None, [V]),
        "trmc_network_set_params_f64": (I, [V, _p(ChParamsF64)]),
        "trmc_network_set_params_f32": (I, [V, _p(ChParamsF32)]),
        "trmc_network_set_state_f64": (I, [V, pD, pD, pD]),
        "trmc_network_set_forcing_f64": (I, [V, pD, I]),
        "trmc_network_set_forcing_f32": (I, [V, pF, I]),
        "trmc_network_run": (I, [V, I, D, _p(RunStats)]),
        "trmc_network_set_forcing_mapped": (I, [V, pD, I]),
        "trmc_network_order_forcing_f64": (I, [V, pD, I, pD]),
        "trmc_host_alloc": (I, [C.c_size_t, _p(V)]),
        "trmc_host_free": (None, [V]),
        "trmc_network_reset": (I, [V]),
        "trmc_network_get_state_f64": (I, [V, pD, pD, pD, pD]),
        "trmc_network_get_history_f64": (I, [V, pD]),
        "trmc_network_info": (I, [V, I, _p(C.c_int64), I]),
        "trmc_network_get_trace": (I, [V, pD, pI, I]),
        "trmc_load_topo": (I, [C.c_char_p, _p(Topo)]),
        "trmc_save_topo": (I, [C.c_char_p, _p(Topo)]),
        "trmc_free_topo": (None, [_p(Topo)]),
        "trmc_load_params": (I, [C.c_char_p, I, _p(ChParamsF32)]),
        "trmc_save_params": (I, [C.c_char_p, I, _p(ChParamsF32)]),
        "trmc_load_forcings": (I, [C.c_char_p, I, I, pI, pF, pF, pF, pF]),
        "trmc_save_forcings": (I, [C.c_char_p, I, I, I, pF, pF, pF, pF]),
        "trmc_free": (None, [V]),
        "trmc_matrix_coeffs_f32": (I, [_p(ChParamsF32), I, C.c_float] + [pF] * 4),
        "trmc_matrix_build": (I, [I, pI, pI] + [pF] * 4 + [C.c_float, _p(Csr), _p(Csr)]),
        "trmc_matrix_secant_coeffs_f32": (I, [_p(ChParamsF32), I, pI, pI, pI, pI, I, pF, I, C.c_float] + [pF] * 4),
        "trmc_free_csr": (None, [_p(Csr)]),
        "trmc_matrix_route_f32": (I, [_p(Csr), _p(Csr), pF, pF, I, I, pF]),
        "trmc_matrix_route_timed_f32": (I, [_p(Csr), _p(Csr), pF, pF, I, I, pF, I, pD]),
        "trmc_build_flowpath_network": (I, [_p(C.c_int64), _p(C.c_int64), I, _p(C.c_int64), pD, I, _p(Topo),
                                            _p(_p(ChParamsF32)), _p(_p(C.c_int64))]),
        "trmc_generate_network": (I, [I, I, I, C.c_uint64, _p(Topo)]),
        "trmc_generate_params": (I, [I, C.c_uint64, _p(ChParamsF64)]),
        "trmc_generate_forcing": (I, [I, I, I, C.c_uint64, pD]),
        "trmc_generate_forcing_member": (I, [I, I, I, C.c_uint64, pD]),
        "trmc_generate_state": (I, [I, _p(ChParamsF64), D, pD, pD]),
        "trmc_compute_stats": (I, [pD, pD, I, pD]),
        "trmc_network_set_mlp": (I, [V, pF]),
        "trmc_network_set_linear_coeffs": (I, [V, pF, pF, pF, pF]),
        "trmc_linear_coeffs_f32": (I, [_p(ChParamsF32), I, C.c_float, pF, pF, pF, pF]),
        "trmc_linear_mc_f64": (I, [pF] * 4 + topo + [pF, pF, pD, I]),
        "trmc_diffusive_tridiag_f32": (I, [V] + [pF] * 5 + [I]),
        "trmc_diffusive_tridiag_f32_noalloc": (I, [V] * 6 + [I, V]),
        "trmc_ingest_nex_csv": (I, [C.c_char_p, C.POINTER(C.c_int64), I, I, I, pF, C.POINTER(C.c_int64)]),
        "trmc_ingest_nex_csv_to_forcings": (I, [C.c_char_p] * 3 + [I, I, C.POINTER(C.c_int64)]),
        "trmc_ingest_nex_csv_rules": (I, [C.c_char_p, C.POINTER(C.c_int64), I, I, I, I, pF, C.POINTER(C.c_int64)]),
        "trmc_ingest_nex_csv_to_forcings_rules": (I, [C.c_char_p] * 3 + [I, I, I, C.POINTER(C.c_int64)]),
        "trmc_diffusive_tridiag_timed_f32": (I, [V] + [pF] * 5 + [I, I, pD]),
        "trmc_network_generate_forcing": (I, [V, I, C.c_uint64]),
        "trmc_stream_copy": (I, [I, C.c_size_t, I, pD]),
        "trmc_topo_concat": (I, [_p(Topo), I, _p(Topo), pI, pI]),
        "trmc_autotune": (I, topo + [_p(ChParamsF64), pD, pD, pD, pD, C.c_uint64, I, D, pF, _p(Options),
                                     _p(Options), pD]),
        "trmc_mlp_make_dataset": (I, [I, C.c_uint64, pF, pF, pD]),
        "trmc_mlp_init_weights": (None, [C.c_uint64, pF]),
        "trmc_mlp_train": (I, [pF, I, I, C.c_float, C.c_uint64, pF]),
        "trmc_mlp_forward_backward": (I, [pF, pF, pF, I, pF, pF]),
        "trmc_mlp_predict": (I, [pF, pF, I, pF, I]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _check(rc: int) -> int:
    if rc < 0:
        msg = lib().trmc_last_error()
        raise TrmcError(rc, msg.decode() if msg else "")
    return rc


# ---------------------------------------------------------------- arrays
def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _need(a: np.ndarray, n: int, what: str, at_least: bool = False) -> None:
    """The C side reads n values of a: refuse a shorter (or, unless
    at_least, a differently sized) array before any device work."""
    if (a.size < n) if at_least else (a.size != n):
        raise TrmcError(ERR_ARG, f"{what} holds {a.size} values, needs {'at least ' if at_least else ''}{n}")


def params_f64(p: np.ndarray) -> np.ndarray:
    """[N, 8] parameter matrix in ChParams field order -> contiguous float64."""
    p = _f64(p)
    if p.ndim != 2 or p.shape[1] != 8:
        raise TrmcError(ERR_ARG, "params must be [n_segments, 8]")
    return p


def params_f32(p: np.ndarray) -> np.ndarray:
    p = _f32(p)
    if p.ndim != 2 or p.shape[1] != 8:
        raise TrmcError(ERR_ARG, "params must be [n_segments, 8]")
    return p


@dataclass
class Topology:
    """Reach CSR + level layout of topo.bin (gen_network.py:10-37)."""

    n_reaches: int
    n_segments: int
    n_levels: int
    max_upstream: int
    reach_seg_start: np.ndarray
    reach_seg_len: np.ndarray
    reach_level: np.ndarray
    reach_n_up: np.ndarray
    reach_up_ptr: np.ndarray
    reach_up_idx: np.ndarray
    level_ptr: np.ndarray
    level_reach: np.ndarray

    def arrays(self):
        """The six arrays and three sizes in the reference call order."""
        a = [_i32(self.reach_seg_start), _i32(self.reach_seg_len), _i32(self.reach_up_ptr),
             _i32(self.reach_up_idx), _i32(self.level_ptr), _i32(self.level_reach)]
        return a, (int(self.n_reaches), int(self.n_levels), int(self.n_segments))

    def c_args(self):
        a, sizes = self.arrays()
        self._keep = a
        return [_ptr(x, C.c_int) for x in a] + list(sizes)

    @classmethod
    def from_c(cls, t: Topo) -> "Topology":
        R, L = t.n_reaches, t.n_levels
        nnz = t.reach_up_ptr[R] if R >= 0 else 0

        def take(ptr, n):
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(np.int32).copy() if n > 0 else np.zeros(0, np.int32)

        return cls(R, t.n_segments, L, t.max_upstream,
                   take(t.reach_seg_start, R), take(t.reach_seg_len, R), take(t.reach_level, R),
                   take(t.reach_n_up, R), take(t.reach_up_ptr, R + 1), take(t.reach_up_idx, nnz),
                   take(t.level_ptr, L + 1), take(t.level_reach, R))

    def to_c(self) -> Topo:
        t = Topo()
        t.n_reaches, t.n_segments, t.n_levels, t.max_upstream = (
            int(self.n_reaches), int(self.n_segments), int(self.n_levels), int(self.max_upstream))
        keep = []
        for name in ("reach_seg_start", "reach_seg_len", "reach_level", "reach_n_up",
                     "reach_up_ptr", "reach_up_idx", "level_ptr", "level_reach"):
            a = _i32(getattr(self, name))
            keep.append(a)
            setattr(t, name, _ptr(a, C.c_int))
        t._keep = keep
        return t


# ------------------------------------------------------------ formats
def load_topo(path: str) -> Topology:
    t = Topo()
    _check(lib().trmc_load_topo(path.encode(), C.byref(t)))
    try:
        return Topology.from_c(t)
    finally:
        lib().trmc_free_topo(C.byref(t))


def save_topo(path: str, topo: Topology) -> None:
    t = topo.to_c()
    _check(lib().trmc_save_topo(path.encode(), C.byref(t)))


def load_params(path: str, n_segments: int) -> np.ndarray:
    out = np.zeros((n_segments, 8), np.float32)
    _check(lib().trmc_load_params(path.encode(), n_segments, _ptr(out, ChParamsF32)))
    return out


def save_params(path: str, params: np.ndarray) -> None:
    p = params_f32(params)
    _check(lib().trmc_save_params(path.encode(), p.shape[0], _ptr(p, ChParamsF32)))


def load_forcings(path: str, n_reaches: int, n_segments: int):
    """Returns (qlat [T, N], qup0 [R], qdp0 [N], dp0 [N]) as float32."""
    T = C.c_int()
    none = C.POINTER(C.c_float)()
    _check(lib().trmc_load_forcings(path.encode(), n_reaches, n_segments, C.byref(T), none, none, none, none))
    qlat = np.zeros((T.value, n_segments), np.float32)
    qup0 = np.zeros(n_reaches, np.float32)
    qdp0 = np.zeros(n_segments, np.float32)
    dp0 = np.zeros(n_segments, np.float32)
    _check(lib().trmc_load_forcings(path.encode(), n_reaches, n_segments, C.byref(T),
                                    _ptr(qlat, C.c_float), _ptr(qup0, C.c_float),
                                    _ptr(qdp0, C.c_float), _ptr(dp0, C.c_float)))
    return qlat, qup0, qdp0, dp0


def save_forcings(path: str, qlat, qup0, qdp0, dp0) -> None:
    qlat = _f32(qlat)
    T, N = qlat.shape
    qup0, qdp0, dp0 = _f32(qup0), _f32(qdp0), _f32(dp0)
    _check(lib().trmc_save_forcings(path.encode(), qup0.shape[0], N, T, _ptr(qlat, C.c_float),
                                    _ptr(qup0, C.c_float), _ptr(qdp0, C.c_float), _ptr(dp0, C.c_float)))


# ---------------------------------------------------------- generator
def generate_network(kind: int, n_segments: int, n_reaches: int = 0, seed: int = 42) -> Topology:
    t = Topo()
    _check(lib().trmc_generate_network(kind, n_segments, n_reaches, seed, C.byref(t)))
    try:
        return Topology.from_c(t)
    finally:
        lib().trmc_free_topo(C.byref(t))


def generate_params(n_segments: int, seed: int = 42) -> np.ndarray:
    out = np.zeros((n_segments, 8), np.float64)
    _check(lib().trmc_generate_params(n_segments, seed, _ptr(out, ChParamsF64)))
    return out


def generate_forcing(n_segments: int, n_timesteps: int, n_members: int = 1, seed: int = 42) -> np.ndarray:
    out = np.zeros((n_members, n_timesteps, n_segments), np.float64)
    _check(lib().trmc_generate_forcing(n_segments, n_timesteps, n_members, seed, _ptr(out, C.c_double)))
    return out


def generate_forcing_member(n_segments: int, n_timesteps: int, member: int, seed: int = 42) -> np.ndarray:
    """Member `member` of generate_forcing(..., n_members > member, seed), [T, N]."""
    out = np.zeros((n_timesteps, n_segments), np.float64)
    _check(lib().trmc_generate_forcing_member(n_segments, n_timesteps, member, seed, _ptr(out, C.c_double)))
    return out


def generate_state(params: np.ndarray, q0: float = 1.0):
    p = params_f64(params)
    n = p.shape[0]
    qdp0 = np.zeros(n, np.float64)
    dp0 = np.zeros(n, np.float64)
    _check(lib().trmc_generate_state(n, _ptr(p, ChParamsF64), q0, _ptr(qdp0, C.c_double), _ptr(dp0, C.c_double)))
    return qdp0, dp0


def concat_topologies(topos):
    """Disjoint union of networks (trmc_topo_concat) for one launch.
    Returns (Topology, segment offsets [n+1], reach offsets [n+1])."""
    n = len(topos)
    arr = (Topo * max(n, 1))()
    keep = []
    for i, t in enumerate(topos):
        c = t.to_c()
        keep.append(c)
        arr[i] = c
    out = Topo()
    so = np.zeros(n + 1, np.int32)
    ro = np.zeros(n + 1, np.int32)
    _check(lib().trmc_topo_concat(arr, n, C.byref(out), _ptr(so, C.c_int), _ptr(ro, C.c_int)))
    try:
        topo = Topology.from_c(out)
    finally:
        lib().trmc_free_topo(C.byref(out))
    return topo, so, ro


def batch_sizes(n_networks: int, lo: int, hi: int, seed: int) -> np.ndarray:
    """Network sizes log-uniform in [lo, hi] (seeded)."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n_networks)).astype(np.int64)


def make_batch_workload(sizes, n_timesteps: int, n_members: int = 1, seed: int = 42, forcing: bool = True):
    """Independent Shreve-style networks (kind 2, network i seeded seed + i,
    about 0.3 reaches per segment) routed as one union network.  Parameters
    and state for the union; lateral inflow [M, T, N] on the host unless
    forcing=False (then use Network.generate_forcing with seed ^ 0xA1A1)."""
    nets = [generate_network(2, int(n), int(0.3 * n), seed + i) for i, n in enumerate(sizes)]
    topo, so, ro = concat_topologies(nets)
    params = generate_params(topo.n_segments, seed ^ 0x5A5A)
    qdp0, dp0 = generate_state(params)
    qlat = generate_forcing(topo.n_segments, n_timesteps, n_members, seed ^ 0xA1A1) if forcing else None
    return Workload(topo, params, qlat, qdp0, dp0, reach_last0_from(topo, qdp0)), so, ro


def reach_last0_from(topo: Topology, qdp0: np.ndarray) -> np.ndarray:
    """reach_last0[r] = qdp0 of reach r's last segment (wavefront_mc.cu:676-680)."""
    last = topo.reach_seg_start + topo.reach_seg_len - 1
    return np.asarray(qdp0)[last].copy()


@dataclass
class Workload:
    topo: Topology
    params: np.ndarray      # [N, 8] float64 (float-representable)
    qlat: np.ndarray        # [M, T, N] float64
    qdp0: np.ndarray        # [N]
    dp0: np.ndarray         # [N]
    reach_last0: np.ndarray  # [R]


def make_workload(kind: int, n_segments: int, n_timesteps: int, n_reaches: int = 0,
                  n_members: int = 1, seed: int = 42, forcing: bool = True) -> Workload:
    """Seeded synthetic network, parameters, lateral inflow and initial state
    (forcing=False: no host lateral inflow; Network.generate_forcing with
    seed ^ 0xA1A1 makes the same values on the device)."""
    topo = generate_network(kind, n_segments, n_reaches, seed)
    params = generate_params(topo.n_segments, seed ^ 0x5A5A)
    qlat = generate_forcing(topo.n_segments, n_timesteps, n_members, seed ^ 0xA1A1) if forcing else None
    qdp0, dp0 = generate_state(params, 1.0)
    return Workload(topo, params, qlat, qdp0, dp0, reach_last0_from(topo, qdp0))


def compute_stats(pred: np.ndarray, truth: np.ndarray) -> dict:
    pred, truth = _f64(pred).ravel(), _f64(truth).ravel()
    out = np.zeros(9, np.float64)
    _check(lib().trmc_compute_stats(_ptr(pred, C.c_double), _ptr(truth, C.c_double), pred.size,
                                    _ptr(out, C.c_double)))
    keys = ("max_abs", "p50", "p90", "p99", "p99_9", "max_rel", "within_1pct", "within_10pct", "nans")
    return dict(zip(keys, out.tolist()))


# ---------------------------------------------------------------- plan
PLAN_ARRAYS = ("level_ptr", "unit_level", "unit_seg", "up_ptr", "up_idx", "unit_flags",
               "unit_reach", "unit_init", "seg_of_dev", "reach_last_unit", "unit_rec")


def plan(topo: Topology, unit_segments: int = 1, n_timesteps: int = 1, n_members: int = 1) -> dict:
    """Host-only schedule plan (no device needed): info + exported arrays."""
    h = C.c_void_p()
    _check(lib().trmc_plan_build(*topo.c_args(), unit_segments, C.byref(h)))
    try:
        info = np.zeros(8, np.int64)
        _check(lib().trmc_plan_info(h, n_timesteps, n_members, _ptr(info, C.c_int64), 8))
        keys = ("n_units", "n_levels", "ring", "super_levels", "max_items", "n_dev_segments",
                "max_span", "n_links")
        out = dict(zip(keys, info.tolist()))
        cap = int(max(topo.n_segments, topo.n_reaches, 8 * out["n_units"], out["n_links"]) + out["n_levels"] + 2)
        buf = np.zeros(cap, np.int32)
        for i, name in enumerate(PLAN_ARRAYS):
            n = _check(lib().trmc_plan_export(h, i, _ptr(buf, C.c_int), cap))
            out[name] = buf[:n].copy()
        return out
    finally:
        lib().trmc_plan_destroy(h)


# -------------------------------------------------------------- routing
def default_options(**kw) -> Options:
    o = Options()
    lib().trmc_default_options(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TrmcError(ERR_ARG, f"unknown option {k}")
        setattr(o, k, v)
    return o


def stream_copy(n_bytes: int, reps: int = 5, device: int = 0) -> float:
    """GB/s of a plain device copy of n_bytes (read + write counted)."""
    g = C.c_double()
    _check(lib().trmc_stream_copy(int(device), int(n_bytes), int(reps), C.byref(g)))
    return g.value


def synchronize(device: int = 0) -> None:
    _check(lib().trmc_device_synchronize(device))


def device_count() -> int:
    n = C.c_int()
    _check(lib().trmc_device_count(C.byref(n)))
    return n.value


def _check_route_inputs(topo, p, q, qd, dp, rl, n_timesteps):
    N, R = topo.n_segments, topo.n_reaches
    _need(p, 8 * N, "params")
    _need(q, int(n_timesteps) * N, "qlat_ts [T, N]", at_least=True)
    _need(qd, N, "qdp0_seg")
    _need(dp, N, "dp0_seg")
    _need(rl, R, "reach_last0")


def wavefront_fp64(params32, topo: Topology, qlat_ts, qdp0_seg, dp0_seg, reach_last0,
                   n_timesteps: int, dt: float = 300.0):
    """Drop-in for cpu_wavefront_fp64: float32 inputs, FP64 routing on the GPU.

    Returns (qdc_seg_out_d [N], reach_last_out_d [R]) as float64."""
    p = params_f32(params32)
    q, qd, dp, rl = _f32(qlat_ts), _f32(qdp0_seg), _f32(dp0_seg), _f32(reach_last0)
    _check_route_inputs(topo, p, q, qd, dp, rl, n_timesteps)
    qout = np.zeros(topo.n_segments, np.float64)
    rout = np.zeros(topo.n_reaches, np.float64)
    _check(lib().trmc_wavefront_f32in_f64(
        _ptr(p, ChParamsF32), *topo.c_args(), _ptr(q, C.c_float), _ptr(qd, C.c_float),
        _ptr(dp, C.c_float), _ptr(rl, C.c_float), _ptr(qout, C.c_double), _ptr(rout, C.c_double),
        int(n_timesteps), float(dt)))
    return qout, rout


def wavefront_f32(params32, topo: Topology, qlat_ts, qdp0_seg, dp0_seg, reach_last0,
                  n_timesteps: int, dt: float = 300.0):
    """Drop-in for cpu_wavefront (wavefront_mc.cu:413-427): the FP32 rule end to
    end on the GPU (every powf correctly rounded).

    Returns (qdc_seg_out [N], reach_last_out [R]) as float32."""
    p = params_f32(params32)
    q, qd, dp, rl = _f32(qlat_ts), _f32(qdp0_seg), _f32(dp0_seg), _f32(reach_last0)
    _check_route_inputs(topo, p, q, qd, dp, rl, n_timesteps)
    qout = np.zeros(topo.n_segments, np.float32)
    rout = np.zeros(topo.n_reaches, np.float32)
    _check(lib().trmc_wavefront_f32(
        _ptr(p, ChParamsF32), *topo.c_args(), _ptr(q, C.c_float), _ptr(qd, C.c_float),
        _ptr(dp, C.c_float), _ptr(rl, C.c_float), _ptr(qout, C.c_float), _ptr(rout, C.c_float),
        int(n_timesteps), float(dt)))
    return qout, rout


def linear_coeffs(params32, dt: float = 300.0):
    """compute_mc_coeffs on the GPU (linear_mc.cu:57-113): (C1, C2, C3, C4/ql),
    float32 [N] each."""
    p = params_f32(params32)
    n = len(p)
    out = [np.zeros(n, np.float32) for _ in range(4)]
    _check(lib().trmc_linear_coeffs_f32(_ptr(p, ChParamsF32), n, float(dt), *[_ptr(a, C.c_float) for a in out]))
    return tuple(out)


def linear_mc(coeffs, topo: Topology, qlat_ts, Q0, n_timesteps: int):
    """Drop-in for cpu_linear_mc_fp64 (linear_mc.cu:238-289): float32
    coefficients, forcing [T, N] and initial flow; FP64 routing on the GPU.
    Returns Q_final [N] float64."""
    cs = [_f32(c) for c in coeffs]
    q, q0 = _f32(qlat_ts), _f32(Q0)
    N = topo.n_segments
    if len(cs) != 4:
        raise TrmcError(ERR_ARG, "coeffs must be (C1, C2, C3, C4/ql)")
    for i, c in enumerate(cs):
        _need(c, N, f"coefficient {i + 1}")
    _need(q, int(n_timesteps) * N, "qlat_ts [T, N]", at_least=True)
    _need(q0, N, "Q0")
    out = np.zeros(topo.n_segments, np.float64)
    _check(lib().trmc_linear_mc_f64(*[_ptr(c, C.c_float) for c in cs], *topo.c_args(), _ptr(q, C.c_float),
                                    _ptr(q0, C.c_float), _ptr(out, C.c_double), int(n_timesteps)))
    return out


INGEST_CPP, INGEST_RUST = 0, 1


def ingest_nex_csv(csv_dir: str, topo_to_wbid, n_timesteps: int, n_threads: int = 0, rules: int = INGEST_CPP):
    """Lateral inflow [T, R] float32 from per-nexus CSVs, under the parsing rules
    of parse_qlat_fast_cpp.cpp (INGEST_CPP) or rust_preprocessor (INGEST_RUST);
    returns (qlat, files read, files skipped)."""
    w = np.ascontiguousarray(topo_to_wbid, np.int64)
    q = np.zeros((int(n_timesteps), len(w)), np.float32)
    cnt = np.zeros(2, np.int64)
    _check(lib().trmc_ingest_nex_csv_rules(os.fsencode(csv_dir), w.ctypes.data_as(C.POINTER(C.c_int64)), len(w),
                                           int(n_timesteps), int(n_threads), int(rules), _ptr(q, C.c_float),
                                           cnt.ctypes.data_as(C.POINTER(C.c_int64))))
    return q, int(cnt[0]), int(cnt[1])


def ingest_nex_csv_to_forcings(csv_dir: str, wbid_npy: str, out_path: str, n_timesteps: int = 24,
                               n_threads: int = 0, rules: int = INGEST_CPP):
    """The parse_qlat_fast program's path (or the Rust twin's, INGEST_RUST): writes
    forcings.bin; returns (files read, skipped)."""
    cnt = np.zeros(2, np.int64)
    _check(lib().trmc_ingest_nex_csv_to_forcings_rules(os.fsencode(csv_dir), os.fsencode(wbid_npy),
                                                       os.fsencode(out_path), int(n_timesteps), int(n_threads),
                                                       int(rules), cnt.ctypes.data_as(C.POINTER(C.c_int64))))
    return int(cnt[0]), int(cnt[1])


def matrix_coeffs(params, dt: float = 300.0):
    """compute_mc_coeffs_cpu (linear_mc_matrix.cu:58-84) on the GPU: the
    linearized MC coefficients with X = 0.25, float [n] each."""
    p = params_f32(params)
    n = p.shape[0]
    out = [np.zeros(n, np.float32) for _ in range(4)]
    _check(lib().trmc_matrix_coeffs_f32(_ptr(p, ChParamsF32), n, float(dt), *[_ptr(o, C.c_float) for o in out]))
    return tuple(out)


def matrix_secant_coeffs(params, topo: "Topology", qlat0, dt: float = 300.0):
    """The secant-fitted coefficients of linear_mc_matrix_secant.cu (one segment
    per reach): (C1, C2, C3, C4/ql), float [n_reaches] each, from the first
    step's lateral inflow qlat0 [n_segments]."""
    p = params_f32(params)
    q = _f32(qlat0)
    R = topo.n_reaches
    _need(p, 8 * R, "params (one segment per reach)")
    _need(q, topo.n_segments, "qlat0")
    out = [np.zeros(R, np.float32) for _ in range(4)]
    ptr, lptr, lrch = _i32(topo.reach_up_ptr), _i32(topo.level_ptr), _i32(topo.level_reach)
    idx = _i32(topo.reach_up_idx) if len(topo.reach_up_idx) else np.zeros(1, np.int32)
    _check(lib().trmc_matrix_secant_coeffs_f32(
        _ptr(p, ChParamsF32), R, _ptr(ptr, C.c_int), _ptr(idx, C.c_int), _ptr(lptr, C.c_int),
        _ptr(lrch, C.c_int), topo.n_levels, _ptr(q, C.c_float), len(q), float(dt),
        *[_ptr(o, C.c_float) for o in out]))
    return tuple(out)


def matrix_build(up_ptr, up_idx, coeffs, tol: float = 1e-6):
    """build_propagation_matrices (linear_mc_matrix.cu:118-229): (P_Q, P_q), each
    (rowptr, col, val) numpy arrays."""
    ptr, idx = _i32(up_ptr), _i32(up_idx)
    if ptr.size < 1 or ptr[0] != 0 or int(ptr[-1]) > idx.size or np.any(np.diff(ptr) < 0):
        raise TrmcError(ERR_ARG, "up_ptr must be a CSR row pointer over up_idx")
    if len(idx) == 0:
        idx = np.zeros(1, np.int32)
    cs = [_f32(c) for c in coeffs]
    if len(cs) != 4:
        raise TrmcError(ERR_ARG, "coeffs must be (C1, C2, C3, C4/ql)")
    for i, c in enumerate(cs):
        _need(c, len(ptr) - 1, f"coefficient {i + 1}")
    mQ, mq = Csr(), Csr()
    _check(lib().trmc_matrix_build(len(ptr) - 1, _ptr(ptr, C.c_int), _ptr(idx, C.c_int),
                                   *[_ptr(c, C.c_float) for c in cs], float(tol), C.byref(mQ), C.byref(mq)))
    try:
        out = []
        for m in (mQ, mq):
            rp = np.ctypeslib.as_array(m.rowptr, (m.n + 1,)).copy()
            col = np.ctypeslib.as_array(m.col, (m.nnz,)).copy() if m.nnz else np.zeros(0, np.int32)
            val = np.ctypeslib.as_array(m.val, (m.nnz,)).copy() if m.nnz else np.zeros(0, np.float32)
            out.append((rp, col, val))
    finally:
        lib().trmc_free_csr(C.byref(mQ))
        lib().trmc_free_csr(C.byref(mq))
    return tuple(out)


def _csr_view(m):
    rp, col, val = _i32(m[0]), _i32(m[1]), _f32(m[2])
    keep = (rp, col if len(col) else np.zeros(1, np.int32), val if len(val) else np.zeros(1, np.float32))
    c = Csr(len(rp) - 1, len(col), _ptr(keep[0], C.c_int), _ptr(keep[1], C.c_int), _ptr(keep[2], C.c_float))
    return c, keep


def matrix_route(P_Q, P_q, qlat, Q0, n_timesteps: int, lanes_per_row: int = 0, reps: int = 0):
    """The matrix-form time loop on the GPU (linear_mc_matrix_graph.cu:355-392):
    Q after n_timesteps steps, float [n]; with reps > 0 also the mean time of
    one whole replay in ms (inputs resident)."""
    cQ, kQ = _csr_view(P_Q)
    cq, kq = _csr_view(P_q)
    q, q0 = _f32(qlat), _f32(Q0)
    if cq.n != cQ.n:
        raise TrmcError(ERR_ARG, "P_Q and P_q must have the same number of rows")
    _need(q, int(n_timesteps) * cQ.n, "qlat [T, n]", at_least=True)
    _need(q0, cQ.n, "Q0")
    out = np.zeros(cQ.n, np.float32)
    if reps > 0:
        ms = C.c_double()
        _check(lib().trmc_matrix_route_timed_f32(C.byref(cQ), C.byref(cq), _ptr(q, C.c_float), _ptr(q0, C.c_float),
                                                 int(n_timesteps), int(lanes_per_row), _ptr(out, C.c_float),
                                                 int(reps), C.byref(ms)))
        return out, ms.value
    _check(lib().trmc_matrix_route_f32(C.byref(cQ), C.byref(cq), _ptr(q, C.c_float), _ptr(q0, C.c_float),
                                       int(n_timesteps), int(lanes_per_row), _ptr(out, C.c_float)))
    return out


HYDROFABRIC_ATTRS = ("Length_m", "n", "nCC", "BtmWdth", "TopWdth", "TopWdthCC", "ChSlp", "So")


def hydrofabric_id(s):
    """The number after the first '-' of "wb-12", "nex-12", "tnx-12" as Python's
    int() reads it, or None (parse_gpkg.py:30-40)."""
    if s is None:
        return None
    parts = s.split("-", 1)
    if len(parts) != 2:
        return None
    try:
        return int(parts[1])
    except ValueError:
        return None


def build_flowpath_network(ids, toids, links, vals):
    """Network from the flowpaths rows (text ids, None for NULL) and the
    flowpath-attributes rows (link text, vals [rows, 8] float with NaN for NULL
    in HYDROFABRIC_ATTRS order), by the rules of parse_gpkg.py:43-270.
    Returns (Topology, params [n, 8] float32, topo_to_wbid [n] int64)."""
    wid, tid = [], []
    for i, t in zip(ids, toids):
        w = hydrofabric_id(i)
        if w is None:
            continue
        k = hydrofabric_id(t)
        wid.append(w)
        tid.append(-1 if k is None or k < 0 else k)
    keep = [j for j, l in enumerate(links) if hydrofabric_id(l) is not None]
    lid = np.array([hydrofabric_id(links[j]) for j in keep], np.int64)
    av = np.ascontiguousarray(np.asarray(vals, np.float64).reshape(-1, 8)[keep], np.float64)
    wid, tid = np.array(wid, np.int64), np.array(tid, np.int64)
    t = Topo()
    pp = C.POINTER(ChParamsF32)()
    pw = C.POINTER(C.c_int64)()
    i64 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
    _check(lib().trmc_build_flowpath_network(i64(wid), i64(tid), len(wid), i64(lid), _ptr(av, C.c_double), len(lid),
                                             C.byref(t), C.byref(pp), C.byref(pw)))
    try:
        n = t.n_reaches
        topo = Topology.from_c(t)
        params = np.ctypeslib.as_array(C.cast(pp, C.POINTER(C.c_float)), (n * 8,)).reshape(n, 8).copy() if n else \
            np.zeros((0, 8), np.float32)
        wbid = np.ctypeslib.as_array(pw, (n,)).copy() if n else np.zeros(0, np.int64)
    finally:
        lib().trmc_free_topo(C.byref(t))
        lib().trmc_free(pp)
        lib().trmc_free(pw)
    return topo, params, wbid


def read_hydrofabric(gpkg_path: str):
    """The two GeoPackage tables parse_gpkg.py reads (:52-53, :183-186) through
    sqlite3: (ids, toids, links, vals [rows, 8] with NaN for NULL)."""
    import sqlite3
    if not os.path.exists(gpkg_path):
        raise TrmcError(ERR_IO, f"no such GeoPackage: {gpkg_path}")
    con = sqlite3.connect(f"file:{gpkg_path}?mode=ro", uri=True)
    try:
        rows = con.execute("SELECT id, toid FROM flowpaths").fetchall()
        cols = ", ".join(f'"{c}"' for c in HYDROFABRIC_ATTRS)
        attrs = con.execute(f'SELECT link, {cols} FROM "flowpath-attributes"').fetchall()
    finally:
        con.close()
    # numpy reads a NULL (None) as NaN in a float array
    vals = np.array([a[1:] for a in attrs], np.float64).reshape(-1, 8)
    return [r[0] for r in rows], [r[1] for r in rows], [a[0] for a in attrs], vals


def build_hydrofabric(gpkg_path: str, out_dir: str = None):
    """parse_gpkg.py's whole path: the network of a GeoPackage, and, when
    out_dir is given, its topo.bin, params.bin and topo_to_wbid.npy there.
    Returns (Topology, params, topo_to_wbid)."""
    topo, params, wbid = build_flowpath_network(*read_hydrofabric(gpkg_path))
    if out_dir is not None:
        os.makedirs(out_dir, exist_ok=True)
        save_topo(os.path.join(out_dir, "topo.bin"), topo)
        save_params(os.path.join(out_dir, "params.bin"), params)
        np.save(os.path.join(out_dir, "topo_to_wbid.npy"), wbid)
    return topo, params, wbid


DIFF_MAX_NODES = 500
DIFF_REACH = np.dtype([("ncomp", np.int32), ("theta", np.float32), ("dtini", np.float32)])


def diffusive_tridiag(rp, dx, diffusivity, celerity, qp_prev, qp_out=None, reps: int = 0):
    """Drop-in for cpu_diffusive_tridiag (owp_extended_kernels.cu:35-94) on the
    GPU.  rp: DIFF_REACH [nreach]; node arrays float32 [nreach * 500].  Returns
    qp_out (entries beyond ncomp as given, zeros by default); with reps > 0
    returns (qp_out, average kernel ms over reps resident launches)."""
    rp = np.ascontiguousarray(rp, DIFF_REACH)
    arrs = [_f32(a) for a in (dx, diffusivity, celerity, qp_prev)]
    out = np.zeros_like(arrs[3]) if qp_out is None else np.array(qp_out, np.float32, copy=True)
    if any(a.size != len(rp) * DIFF_MAX_NODES for a in arrs + [out]):
        raise TrmcError(ERR_ARG, "node arrays must hold nreach * 500 floats")
    ptrs = [_ptr(a, C.c_float) for a in arrs] + [_ptr(out, C.c_float)]
    if reps > 0:
        ms = C.c_double()
        _check(lib().trmc_diffusive_tridiag_timed_f32(rp.ctypes.data, *ptrs, len(rp), int(reps), C.byref(ms)))
        return out, ms.value
    _check(lib().trmc_diffusive_tridiag_f32(rp.ctypes.data, *ptrs, len(rp)))
    return out


DIFF_NOALLOC_MAX_REACHES = 16384


def diffusive_tridiag_noalloc(d_rp: int, d_dx: int, d_diffusivity: int, d_celerity: int, d_qp_prev: int,
                              d_qp_out: int, nreach: int, stream: int = 0) -> None:
    """trmc_diffusive_tridiag_f32_noalloc: the diffusive step on device-resident
    arrays (device addresses as ints, e.g. a torch tensor's data_ptr()),
    allocating nothing; at most DIFF_NOALLOC_MAX_REACHES reaches."""
    _check(lib().trmc_diffusive_tridiag_f32_noalloc(d_rp, d_dx, d_diffusivity, d_celerity, d_qp_prev, d_qp_out,
                                                    int(nreach), stream or None))


def wavefront_f64(params, topo: Topology, qlat_ts, qdp0_seg, dp0_seg, reach_last0,
                  n_timesteps: int, dt: float = 300.0, **options):
    """FP64-end-to-end variant of the drop-in; options as trmc_options fields."""
    p = params_f64(params)
    q, qd, dp, rl = _f64(qlat_ts), _f64(qdp0_seg), _f64(dp0_seg), _f64(reach_last0)
    _check_route_inputs(topo, p, q, qd, dp, rl, n_timesteps)
    qout = np.zeros(topo.n_segments, np.float64)
    rout = np.zeros(topo.n_reaches, np.float64)
    o = default_options(**options)
    _check(lib().trmc_wavefront_f64(
        _ptr(p, ChParamsF64), *topo.c_args(), _ptr(q, C.c_double), _ptr(qd, C.c_double),
        _ptr(dp, C.c_double), _ptr(rl, C.c_double), _ptr(qout, C.c_double), _ptr(rout, C.c_double),
        int(n_timesteps), float(dt), C.byref(o)))
    return qout, rout


def reach_parallel(params, ql, rs, rl, qu, qd, dp, uq: float = 10.0, dt: float = 300.0):
    """Drop-in for cpu_mc / kernel_mc_gpu: returns (qo, vo, dpo)."""
    p = params_f64(params)
    ql, qu, qd, dp = _f64(ql), _f64(qu), _f64(qd), _f64(dp)
    rs, rl = _i32(rs), _i32(rl)
    n = p.shape[0]
    for a, what in ((ql, "ql"), (qu, "qu"), (qd, "qd"), (dp, "dp")):
        _need(a, n, what)
    _need(rl, rs.size, "rl")
    if rs.size and (rs.min() < 0 or rl.min() < 0 or int((rs.astype(np.int64) + rl).max()) > n):
        raise TrmcError(ERR_ARG, "reach segments lie outside the segment arrays")
    qo, vo, dpo = np.zeros(n), np.zeros(n), np.zeros(n)
    _check(lib().trmc_reach_parallel_f64(
        _ptr(p, ChParamsF64), _ptr(ql, C.c_double), _ptr(rs, C.c_int), _ptr(rl, C.c_int),
        _ptr(qu, C.c_double), _ptr(qd, C.c_double), _ptr(dp, C.c_double), _ptr(qo, C.c_double),
        _ptr(vo, C.c_double), _ptr(dpo, C.c_double), float(uq), float(dt), len(rs)))
    return qo, vo, dpo


def mc_solve(params, inputs, dt: float = 300.0, mode: int = MODE_REFERENCE):
    """Batched mc_solve_d on the GPU.  inputs [n, 5] = qup, quc, qdp, ql, depthp.

    Returns (out [n, 3] = qdc, depth, velocity; iters [n])."""
    p = params_f64(params)
    x = _f64(inputs).reshape(-1, 5)
    n = x.shape[0]
    _need(p, 8 * n, "params (one row per solve)")
    out = np.zeros((n, 3), np.float64)
    it = np.zeros(n, np.int32)
    _check(lib().trmc_mc_solve_f64(_ptr(p, ChParamsF64), _ptr(x, C.c_double), _ptr(out, C.c_double),
                                   _ptr(it, C.c_int), n, float(dt), int(mode)))
    return out, it


# ------------------------------------------------------------ learned guess
MLP_NPARAM = 1281


def mlp_make_dataset(n: int, seed: int = 1, raw: bool = False):
    """Training pairs generated on the GPU (kept there for mlp_train).
    Returns (features [n, 5] f32, target [n] f32, raw [n, 9] f64 or None)."""
    X = np.zeros((n, 5), np.float32)
    Y = np.zeros(n, np.float32)
    R = np.zeros((n, 9), np.float64) if raw else None
    _check(lib().trmc_mlp_make_dataset(int(n), int(seed), _ptr(X, C.c_float), _ptr(Y, C.c_float),
                                       _ptr(R, C.c_double) if raw else C.POINTER(C.c_double)()))
    return X, Y, R


def mlp_init_weights(seed: int = 1) -> np.ndarray:
    W = np.zeros(MLP_NPARAM, np.float32)
    lib().trmc_mlp_init_weights(int(seed), _ptr(W, C.c_float))
    return W


def mlp_train(weights, epochs: int = 30, batch: int = 512, lr: float = 3e-3, seed: int = 1):
    """Adam on the last generated training set.  Returns (weights, loss per epoch)."""
    W = _f32(np.array(weights, np.float32, copy=True))
    loss = np.zeros(int(epochs), np.float32)
    _check(lib().trmc_mlp_train(_ptr(W, C.c_float), int(epochs), int(batch), float(lr), int(seed),
                                _ptr(loss, C.c_float)))
    return W, loss


def mlp_forward_backward(weights, features, target):
    W, X, Y = _f32(weights), _f32(features), _f32(target)
    n = len(Y)
    pred = np.zeros(n, np.float32)
    grad = np.zeros(MLP_NPARAM, np.float32)
    _check(lib().trmc_mlp_forward_backward(_ptr(W, C.c_float), _ptr(X, C.c_float), _ptr(Y, C.c_float), n,
                                           _ptr(pred, C.c_float), _ptr(grad, C.c_float)))
    return pred, grad


def mlp_predict(weights, features, use_mfma=True) -> np.ndarray:
    """use_mfma: True/1 MFMA wave forward, False/0 scalar, 2 the routing
    kernel's packed pairs (NaN for the pairs of lanes i with i % 3 == 1)."""
    W, X = _f32(weights), _f32(np.asarray(features, np.float32).reshape(-1, 5))
    out = np.zeros(len(X), np.float32)
    _check(lib().trmc_mlp_predict(_ptr(W, C.c_float), _ptr(X, C.c_float), len(X), _ptr(out, C.c_float),
                                  int(use_mfma)))
    return out


def autotune(topo: Topology, params, qdp0, dp0, reach_last0=None, n_members: int = 1, n_probe: int = 16,
             qlat=None, forcing_seed: int = 0, dt: float = 300.0, mlp_weights=None, **options):
    """trmc_autotune: the fastest schedule (tile / whole item / wavefront),
    unit_segments, batch_iters and block_size on the first n_probe steps.
    Returns (options dict, ms per step)."""
    base = default_options(n_members=n_members, **options)
    p = params_f64(params)
    M, N = int(n_members), topo.n_segments
    q = _f64(np.broadcast_to(np.asarray(qdp0, np.float64), (M, N)))
    d = _f64(np.broadcast_to(np.asarray(dp0, np.float64), (M, N)))
    rl = None if reach_last0 is None else _f64(np.broadcast_to(np.asarray(reach_last0, np.float64),
                                                                (M, topo.n_reaches)))
    ql = None if qlat is None else _f64(np.asarray(qlat)[:, :n_probe])
    W = None if mlp_weights is None else _f32(mlp_weights)
    best = Options()
    ms = C.c_double()
    _check(lib().trmc_autotune(*topo.c_args(), _ptr(p, ChParamsF64), _ptr(q, C.c_double), _ptr(d, C.c_double),
                               _ptr(rl, C.c_double) if rl is not None else C.POINTER(C.c_double)(),
                               _ptr(ql, C.c_double) if ql is not None else C.POINTER(C.c_double)(),
                               int(forcing_seed), int(n_probe), float(dt),
                               _ptr(W, C.c_float) if W is not None else C.POINTER(C.c_float)(),
                               C.byref(base), C.byref(best), C.byref(ms)))
    keys = [f[0] for f in Options._fields_]
    return {k: getattr(best, k) for k in keys}, ms.value


class _HostArray:
    """float64 array in page-locked host memory (trmc_host_alloc), freed with
    the object."""

    def __init__(self, shape):
        self.shape = tuple(int(x) for x in shape)
        n = int(np.prod(self.shape))
        self._p = C.c_void_p()
        _check(lib().trmc_host_alloc(max(n, 1) * 8, C.byref(self._p)))
        self._a = np.ctypeslib.as_array(C.cast(self._p, C.POINTER(C.c_double)), (max(n, 1),))[:n].reshape(self.shape)

    def array(self) -> np.ndarray:
        return self._a

    def __del__(self):
        if getattr(self, "_p", None) and self._p.value:
            self._a = None
            lib().trmc_host_free(self._p)
            self._p = C.c_void_p()


class Network:
    """Device-resident network (trmc_network_*): topology, parameters, forcing
    and state stay in HBM across runs."""

    def __init__(self, topo: Topology, **options):
        self.topo = topo
        self.options = default_options(**options)
        self.n_members = int(self.options.n_members)
        self._h = C.c_void_p()
        _check(lib().trmc_network_create(*topo.c_args(), C.byref(self.options), C.byref(self._h)))
        self.last_stats: dict = {}

    def close(self) -> None:
        if self._h:
            lib().trmc_network_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params) -> None:
        p = np.asarray(params)
        if p.dtype == np.float32:
            p = params_f32(p)
            _check(lib().trmc_network_set_params_f32(self._h, _ptr(p, ChParamsF32)))
        else:
            p = params_f64(p)
            _check(lib().trmc_network_set_params_f64(self._h, _ptr(p, ChParamsF64)))

    def set_state(self, qdp0, dp0, reach_last0=None) -> None:
        M = self.n_members
        q = _f64(np.broadcast_to(np.asarray(qdp0, np.float64), (M, self.topo.n_segments)))
        d = _f64(np.broadcast_to(np.asarray(dp0, np.float64), (M, self.topo.n_segments)))
        rl = None
        if reach_last0 is not None:
            rl = _f64(np.broadcast_to(np.asarray(reach_last0, np.float64), (M, self.topo.n_reaches)))
        _check(lib().trmc_network_set_state_f64(
            self._h, _ptr(q, C.c_double), _ptr(d, C.c_double),
            _ptr(rl, C.c_double) if rl is not None else C.POINTER(C.c_double)()))

    def set_forcing(self, qlat) -> None:
        q = np.asarray(qlat)
        if q.ndim == 2:
            q = q[None]
        if q.shape[0] != self.n_members or q.shape[2] != self.topo.n_segments:
            raise TrmcError(ERR_ARG, "forcing must be [n_members, T, n_segments]")
        if q.dtype == np.float32:
            q = _f32(q)
            _check(lib().trmc_network_set_forcing_f32(self._h, _ptr(q, C.c_float), q.shape[1]))
        else:
            q = _f64(q)
            _check(lib().trmc_network_set_forcing_f64(self._h, _ptr(q, C.c_double), q.shape[1]))
        self._mapped = None

    def generate_forcing(self, n_timesteps: int, seed: int) -> None:
        """Synthetic lateral inflow generated on the device: the values
        generate_forcing(n_segments, n_timesteps, n_members, seed) gives."""
        _check(lib().trmc_network_generate_forcing(self._h, int(n_timesteps), int(seed)))

    def set_linear_coeffs(self, coeffs) -> None:
        """LINEAR mode coefficients (C1, C2, C3, C4/ql), float32 [N] each."""
        cs = [_f32(c) for c in coeffs]
        if any(len(c) != self.topo.n_segments for c in cs):
            raise TrmcError(ERR_ARG, "coefficients must have one value per segment")
        _check(lib().trmc_network_set_linear_coeffs(self._h, *[_ptr(c, C.c_float) for c in cs]))

    def set_mlp(self, weights) -> None:
        """Weights of the LEARNED depth guess (MLP_NPARAM float32)."""
        W = _f32(weights)
        if W.size != MLP_NPARAM:
            raise TrmcError(ERR_ARG, f"need {MLP_NPARAM} weights")
        _check(lib().trmc_network_set_mlp(self._h, _ptr(W, C.c_float)))

    def run(self, n_timesteps: int, dt: float = 300.0) -> dict:
        st = RunStats()
        _check(lib().trmc_network_run(self._h, int(n_timesteps), float(dt), C.byref(st)))
        self.last_stats = st.as_dict()
        return self.last_stats

    def order_forcing(self, qlat, pinned: bool = True):
        """The forcing [M, T, N] (float64) in the network's device order,
        [T, n_dev_segments, M], for run_streamed; in page-locked host memory
        (pinned) so the step copies run at the link's rate."""
        q = _f64(qlat)
        M, N = self.n_members, self.topo.n_segments
        if q.ndim == 2:
            q = q[None]
        if q.shape[0] != M or q.shape[2] != N:
            raise TrmcError(ERR_ARG, f"qlat must be [{M}, T, {N}]")
        T = q.shape[1]
        nd = self.info(T)["n_dev_segments"]
        out = _HostArray((T, nd, M)) if pinned else np.zeros((T, nd, M), np.float64)
        dst = out.array() if pinned else out
        _check(lib().trmc_network_order_forcing_f64(self._h, _ptr(q, C.c_double), T, _ptr(dst, C.c_double)))
        return out

    def set_forcing_mapped(self, qlat_dev_order) -> None:
        """trmc_network_set_forcing_mapped: the routing reads the forcing in
        place from page-locked host memory (qlat_dev_order from order_forcing,
        kept alive here while it is the network's forcing)."""
        if not isinstance(qlat_dev_order, _HostArray):
            raise TrmcError(ERR_ARG, "mapped forcing must be page-locked (Network.order_forcing)")
        arr = qlat_dev_order.array()
        if arr.shape[2] != self.n_members:
            raise TrmcError(ERR_ARG, "forcing in device order has another member count")
        _check(lib().trmc_network_set_forcing_mapped(self._h, _ptr(arr, C.c_double), arr.shape[0]))
        self._mapped = qlat_dev_order

    def reset(self) -> None:
        """Back to the initial state and the first forcing step (device copies)."""
        _check(lib().trmc_network_reset(self._h))

    def state(self):
        """(qdc, depth, velocity) [M, N] and reach_last [M, R], reference order."""
        M, N, R = self.n_members, self.topo.n_segments, self.topo.n_reaches
        q, d, v = np.zeros((M, N)), np.zeros((M, N)), np.zeros((M, N))
        rl = np.zeros((M, R))
        _check(lib().trmc_network_get_state_f64(self._h, _ptr(q, C.c_double), _ptr(d, C.c_double),
                                                _ptr(v, C.c_double), _ptr(rl, C.c_double)))
        return q, d, v, rl

    def history(self, n_timesteps: int) -> np.ndarray:
        out = np.zeros((self.n_members, n_timesteps, self.topo.n_reaches))
        _check(lib().trmc_network_get_history_f64(self._h, _ptr(out, C.c_double)))
        return out

    def trace(self, n: int):
        """Per-super-level (microseconds, max secant iterations) of the last run."""
        us = np.zeros(n)
        it = np.zeros(n, np.int32)
        m = _check(lib().trmc_network_get_trace(self._h, _ptr(us, C.c_double), _ptr(it, C.c_int), n))
        return us[:m], it[:m]

    def info(self, n_timesteps: int) -> dict:
        info = np.zeros(8, np.int64)
        _check(lib().trmc_network_info(self._h, n_timesteps, _ptr(info, C.c_int64), 8))
        keys = ("n_units", "n_levels", "ring", "super_levels", "max_items", "n_dev_segments",
                "max_span", "n_links")
        return dict(zip(keys, info.tolist()))
