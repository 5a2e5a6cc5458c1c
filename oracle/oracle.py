"""ctypes bindings of the CPU oracle (oracle/liborc.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
module, and only as the checker or the timed CPU baseline; the routing product
(noaa-gpu-kernels_amd/libtrmc.so) never calls it.

``load_ref()`` additionally binds oracle/_ref/libtroute_ref.so, the reference's
own FP64 host routines compiled from the reference sources by
oracle/build_ref.sh (present only where that build ran).
"""
from __future__ import annotations

import ctypes as C
import ctypes.util as ctypes_util
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
D, I, LL = C.c_double, C.c_int, C.c_longlong
pD, pI, pF = C.POINTER(D), C.POINTER(I), C.POINTER(C.c_float)
FIELDS = ("dx", "bw", "tw", "twcc", "n_ch", "ncc", "cs", "s0")


class P64(C.Structure):
    _fields_ = [(n, D) for n in FIELDS]


class P32(C.Structure):
    _fields_ = [(n, C.c_float) for n in FIELDS]


MODE_REFERENCE, MODE_CONVERGED = 0, 1
# "libm": pow of the host libm, as the reference (liborc.so);
# "cr":   correctly rounded pow (liborc_cr.so), the GPU library's arithmetic.
VARIANTS = {"libm": "liborc.so", "cr": "liborc_cr.so"}
_orc = {}


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def lib(variant: str = "libm") -> C.CDLL:
    if variant not in _orc:
        path = os.path.join(HERE, VARIANTS[variant])
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built (make -C oracle)")
        L = C.CDLL(path)
        pP64, pP32 = C.POINTER(P64), C.POINTER(P32)
        topo = [pI] * 6 + [I, I, I]
        L.orc_mc_solve.restype = D
        L.orc_mc_solve.argtypes = [D] * 5 + [pP64, D, pD, pD, pI]
        L.orc_wavefront_f32in.restype = None
        L.orc_wavefront_f32in.argtypes = [pP32] + topo + [pF] * 4 + [pD, pD, I, D, C.POINTER(LL)]
        L.orc_wavefront_f32in_mt.restype = None
        L.orc_wavefront_f32in_mt.argtypes = [pP32] + topo + [pF] * 4 + [pD, pD, I, D, C.POINTER(LL), I]
        L.orc_wavefront_f64.restype = None
        L.orc_wavefront_f64.argtypes = [pP64] + topo + [pD] * 4 + [pD] * 5 + [I, D, I, C.POINTER(LL)]
        L.orc_wavefront_f64_mt.restype = None
        L.orc_wavefront_f64_mt.argtypes = [pP64] + topo + [pD] * 4 + [pD] * 5 + [I, D, I, C.POINTER(LL), I]
        L.orc_reach_parallel.restype = None
        L.orc_reach_parallel.argtypes = [pP64, pD, pI, pI, pD, pD, pD, pD, pD, pD, D, D, I, C.POINTER(LL)]
        L.orc_mc_residual.restype = D
        L.orc_mc_residual.argtypes = [D] * 5 + [pP64, D, pD]
        L.orc_mc_solve_converged.restype = D
        L.orc_mc_solve_converged.argtypes = [D] * 5 + [pP64, D, pD, pD]
        L.orc_residual_enclosure.restype = None
        L.orc_residual_enclosure.argtypes = [D] * 5 + [pP64, D, D, pD, pD]
        L.orc_residual_slope_enclosure.restype = I
        L.orc_residual_slope_enclosure.argtypes = [D] * 5 + [pP64, D, D, pD, pD]
        L.orc_normal_depth.restype = D
        L.orc_normal_depth.argtypes = [D, pP64]
        L.orc_manning_flow.restype = D
        L.orc_manning_flow.argtypes = [D, pP64]
        L.orc_compute_stats.restype = None
        L.orc_compute_stats.argtypes = [pD, pD, I, pD]
        L.orc_pow_fn.restype = D
        L.orc_pow_fn.argtypes = [D, D]
        L.orc_powf_fn.restype = C.c_float
        L.orc_powf_fn.argtypes = [C.c_float, C.c_float]
        L.orc_mc_solve_f32.restype = C.c_float
        L.orc_mc_solve_f32.argtypes = [C.c_float] * 5 + [pP32, C.c_float, pF, pF, pI]
        L.orc_wavefront_f32.restype = None
        L.orc_wavefront_f32.argtypes = [pP32] + topo + [pF] * 4 + [pF] * 4 + [I, C.c_float, C.POINTER(LL), I]
        L.orc_linear_coeffs.restype = None
        L.orc_linear_coeffs.argtypes = [pP32, I, C.c_float] + [pF] * 4
        L.orc_diffusive_tridiag.restype = None
        L.orc_diffusive_tridiag.argtypes = [C.c_void_p] + [pF] * 5 + [I]
        L.orc_linear_mc_fp64.restype = None
        L.orc_linear_mc_fp64.argtypes = [pF] * 4 + topo + [pF, pF, pD, I]
        L.orc_matrix_coeffs.restype = None
        L.orc_matrix_coeffs.argtypes = [pP32, I, C.c_float] + [pF] * 4
        L.orc_matrix_build.restype = I
        L.orc_matrix_build.argtypes = _MATRIX_BUILD_ARGS
        L.orc_free.restype = None
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_matrix_route.restype = None
        L.orc_matrix_route.argtypes = [I, pI, pI, pF, pI, pI, pF, pF, pF, I, I]
        _orc[variant] = L
    return _orc[variant]


def _topo_args(topo):
    keep = [np.ascontiguousarray(topo[k], np.int32) for k in
            ("reach_seg_start", "reach_seg_len", "reach_up_ptr", "reach_up_idx", "level_ptr", "level_reach")]
    return keep, [_ptr(a, I) for a in keep] + [int(topo["n_reaches"]), int(topo["n_levels"]),
                                               int(topo["n_segments"])]


def as_topo(t) -> dict:
    """dict view of a topology (dict, npz-like mapping or object with attributes)."""
    keys = ("n_reaches", "n_segments", "n_levels", "reach_seg_start", "reach_seg_len",
            "reach_up_ptr", "reach_up_idx", "level_ptr", "level_reach")
    if isinstance(t, dict):
        return {k: t[k] for k in keys}
    return {k: getattr(t, k) for k in keys}


def mc_solve(params8, qup, quc, qdp, ql, depthp, dt=300.0, variant="libm"):
    """One segment: returns (qdc, depth, velocity, iterations)."""
    p = P64(*map(float, params8))
    d, v, it = D(), D(), I()
    q = lib(variant).orc_mc_solve(dt, qup, quc, qdp, ql, C.byref(p), depthp, C.byref(d), C.byref(v), C.byref(it))
    return q, d.value, v.value, it.value


def mc_solve_many(params, inputs, dt=300.0, mode=MODE_REFERENCE, variant="libm"):
    """params [n, 8], inputs [n, 5] (qup, quc, qdp, ql, depthp) -> out [n, 3], iters [n]."""
    params = np.asarray(params, np.float64)
    inputs = np.asarray(inputs, np.float64)
    out = np.zeros((len(params), 3))
    its = np.zeros(len(params), np.int32)
    for i in range(len(params)):
        p = P64(*params[i])
        d, v, it = D(), D(), I()
        x = inputs[i]
        if mode == MODE_REFERENCE:
            q = lib(variant).orc_mc_solve(dt, x[0], x[1], x[2], x[3], C.byref(p), x[4], C.byref(d), C.byref(v), C.byref(it))
        else:
            q = lib(variant).orc_mc_solve_converged(dt, x[0], x[1], x[2], x[3], C.byref(p), x[4], C.byref(d), C.byref(v))
        out[i] = (q, d.value, v.value)
        its[i] = it.value
    return out, its


def wavefront_f32in(topo, params32, qlat32, qdp0, dp0, reach_last0, n_timesteps, dt=300.0, variant="libm",
                    threads=1):
    """cpu_wavefront_fp64 restated: returns (qdc_out [N], reach_last_out [R], iterations).
    threads > 1 shares each level's reaches over host threads (same bits)."""
    t = as_topo(topo)
    keep, targs = _topo_args(t)
    p = np.ascontiguousarray(params32, np.float32)
    q = np.ascontiguousarray(qlat32, np.float32)
    q0, d0, r0 = (np.ascontiguousarray(x, np.float32) for x in (qdp0, dp0, reach_last0))
    qo = np.zeros(int(t["n_segments"]))
    ro = np.zeros(int(t["n_reaches"]))
    its = LL(0)
    args = [_ptr(p, P32), *targs, _ptr(q, C.c_float), _ptr(q0, C.c_float), _ptr(d0, C.c_float),
            _ptr(r0, C.c_float), _ptr(qo, D), _ptr(ro, D), int(n_timesteps), float(dt), C.byref(its)]
    if threads > 1:
        lib(variant).orc_wavefront_f32in_mt(*args, int(threads))
    else:
        lib(variant).orc_wavefront_f32in(*args)
    return qo, ro, its.value


def wavefront_f64(topo, params, qlat, qdp0, dp0, reach_last0, n_timesteps, dt=300.0,
                  mode=MODE_REFERENCE, history=False, variant="libm", threads=1):
    """All-double DAG routing.  Returns dict with qdc, depth, vel, reach_last,
    iterations and (history=True) reach_hist [T, R]."""
    t = as_topo(topo)
    keep, targs = _topo_args(t)
    N, R = int(t["n_segments"]), int(t["n_reaches"])
    p = np.ascontiguousarray(params, np.float64)
    q = np.ascontiguousarray(qlat, np.float64)
    q0, d0, r0 = (np.ascontiguousarray(x, np.float64) for x in (qdp0, dp0, reach_last0))
    qo, do, vo, ro = np.zeros(N), np.zeros(N), np.zeros(N), np.zeros(R)
    hist = np.zeros((int(n_timesteps), R)) if history else None
    its = LL(0)
    args = [_ptr(p, P64), *targs, _ptr(q, D), _ptr(q0, D), _ptr(d0, D), _ptr(r0, D),
            _ptr(qo, D), _ptr(ro, D), _ptr(do, D), _ptr(vo, D),
            _ptr(hist, D) if history else pD(), int(n_timesteps), float(dt), int(mode), C.byref(its)]
    if threads > 1:
        lib(variant).orc_wavefront_f64_mt(*args, int(threads))
    else:
        lib(variant).orc_wavefront_f64(*args)
    out = dict(qdc=qo, depth=do, vel=vo, reach_last=ro, iterations=its.value)
    if history:
        out["reach_hist"] = hist
    return out


def reach_parallel(params, ql, rs, rl, qu, qd, dp, uq=10.0, dt=300.0, variant="libm"):
    p = np.ascontiguousarray(params, np.float64)
    ql, qu, qd, dp = (np.ascontiguousarray(x, np.float64) for x in (ql, qu, qd, dp))
    rs, rl = np.ascontiguousarray(rs, np.int32), np.ascontiguousarray(rl, np.int32)
    n = len(p)
    qo, vo, dpo = np.zeros(n), np.zeros(n), np.zeros(n)
    lib(variant).orc_reach_parallel(_ptr(p, P64), _ptr(ql, D), _ptr(rs, I), _ptr(rl, I), _ptr(qu, D), _ptr(qd, D),
                             _ptr(dp, D), _ptr(qo, D), _ptr(vo, D), _ptr(dpo, D), uq, dt, len(rs), None)
    return qo, vo, dpo


def residual(params8, qup, quc, qdp, ql, h, dt=300.0, variant="libm"):
    p = P64(*map(float, params8))
    qmc = D()
    f = lib(variant).orc_mc_residual(dt, qup, quc, qdp, ql, C.byref(p), h, C.byref(qmc))
    return f, qmc.value


def residual_enclosure(params8, qup, quc, qdp, ql, a, b, dt=300.0, variant="libm"):
    """Enclosure (lo, hi) of the converged residual over depths [a, b]."""
    p = P64(*map(float, params8))
    lo, hi = D(), D()
    lib(variant).orc_residual_enclosure(dt, qup, quc, qdp, ql, C.byref(p), a, b, C.byref(lo), C.byref(hi))
    return lo.value, hi.value


def residual_slope_enclosure(params8, qup, quc, qdp, ql, a, b, dt=300.0, variant="libm"):
    """Enclosure (lo, hi) of dF/dh over [a, b], or None where none is formed."""
    p = P64(*map(float, params8))
    lo, hi = D(), D()
    ok = lib(variant).orc_residual_slope_enclosure(dt, qup, quc, qdp, ql, C.byref(p), a, b, C.byref(lo), C.byref(hi))
    return (lo.value, hi.value) if ok else None


def _canon_lib(variant):
    L = lib(variant)
    if not getattr(L, "_canon_bound", False):
        pP64 = C.POINTER(P64)
        L.orc_canonical_root.restype = D
        L.orc_canonical_root.argtypes = [D] * 5 + [pP64, D, pD]
        L.orc_canon_residual.restype = D
        L.orc_canon_residual.argtypes = [D] * 5 + [pP64, D, pD]
        L.orc_pow2353_fast.restype = None
        L.orc_pow2353_fast.argtypes = [D, pD, pD]
        L.orc_noncanon_count.restype = LL
        L.orc_noncanon_reset.restype = None
        L._canon_bound = True
    return L


def canonical_root(params8, qup, quc, qdp, ql, r, dt=300.0, variant="libm"):
    """(h*, Q*) of the canonical cell next to the estimate r, or (nan, nan)."""
    p = P64(*map(float, params8))
    q = D(np.nan)
    h = _canon_lib(variant).orc_canonical_root(dt, qup, quc, qdp, ql, C.byref(p), r, C.byref(q))
    return h, q.value


def canon_residual(params8, qup, quc, qdp, ql, h, dt=300.0, variant="libm"):
    """(Fc(h), outflow of Fc's coefficients at h)."""
    p = P64(*map(float, params8))
    q = D()
    f = _canon_lib(variant).orc_canon_residual(dt, qup, quc, qdp, ql, C.byref(p), h, C.byref(q))
    return f, q.value


def pow2353_fast(x, variant="libm"):
    a, b = D(), D()
    _canon_lib(variant).orc_pow2353_fast(float(x), C.byref(a), C.byref(b))
    return a.value, b.value


def noncanon_count(variant="libm", reset=False):
    L = _canon_lib(variant)
    n = int(L.orc_noncanon_count())
    if reset:
        L.orc_noncanon_reset()
    return n


def solve_converged(params8, qup, quc, qdp, ql, depthp, dt=300.0, variant="libm"):
    p = P64(*map(float, params8))
    d, v = D(), D()
    q = lib(variant).orc_mc_solve_converged(dt, qup, quc, qdp, ql, C.byref(p), depthp, C.byref(d), C.byref(v))
    return q, d.value, v.value


def normal_depth(params8, Q, variant="libm"):
    p = P64(*map(float, params8))
    return lib(variant).orc_normal_depth(Q, C.byref(p))


def manning_flow(params8, h, variant="libm"):
    p = P64(*map(float, params8))
    return lib(variant).orc_manning_flow(h, C.byref(p))


def pow_fn(x, y, variant="libm"):
    return lib(variant).orc_pow_fn(float(x), float(y))


def mc_solve_f32(params8, qup, quc, qdp, ql, depthp, dt=300.0, variant="libm"):
    """The FP32 REFERENCE solve (wavefront_mc.cu:66-193), float in and out:
    returns (qdc, depth, velocity, iterations) as Python floats of float32 values."""
    p = P32(*map(float, params8))
    d, v, it = C.c_float(), C.c_float(), I()
    q = lib(variant).orc_mc_solve_f32(float(dt), float(qup), float(quc), float(qdp), float(ql), C.byref(p),
                                      float(depthp), C.byref(d), C.byref(v), C.byref(it))
    return q, d.value, v.value, it.value


def wavefront_f32(topo, params32, qlat32, qdp0, dp0, reach_last0, n_timesteps, dt=300.0, variant="libm",
                  threads=1, history=False):
    """cpu_wavefront restated (wavefront_mc.cu:413-491), FP32 end to end: returns a
    dict of float32 arrays qdc [N], reach_last [R], depth [N] (final step),
    reach_hist [T, R] (history=True) and the iteration total."""
    t = as_topo(topo)
    keep, targs = _topo_args(t)
    n, r = int(t["n_segments"]), int(t["n_reaches"])
    p = np.ascontiguousarray(params32, np.float32)
    q = np.ascontiguousarray(qlat32, np.float32)
    q0, d0, r0 = (np.ascontiguousarray(x, np.float32) for x in (qdp0, dp0, reach_last0))
    qo, ro, do = np.zeros(n, np.float32), np.zeros(r, np.float32), np.zeros(n, np.float32)
    hist = np.zeros((int(n_timesteps), r), np.float32) if history else None
    its = LL(0)
    lib(variant).orc_wavefront_f32(_ptr(p, P32), *targs, _ptr(q, C.c_float), _ptr(q0, C.c_float),
                                   _ptr(d0, C.c_float), _ptr(r0, C.c_float), _ptr(qo, C.c_float),
                                   _ptr(ro, C.c_float), _ptr(do, C.c_float),
                                   _ptr(hist, C.c_float) if history else None, int(n_timesteps), float(dt),
                                   C.byref(its), int(max(1, threads)))
    out = {"qdc": qo, "reach_last": ro, "depth": do, "iters": its.value}
    if history:
        out["reach_hist"] = hist
    return out


def powf_fn(x, y, variant="libm"):
    return lib(variant).orc_powf_fn(float(x), float(y))


def load_ref_f32():
    """The reference's own FP32 host routines (oracle/_ref/libtroute_ref_f32.so:
    mc_solve and cpu_wavefront of wavefront_mc.cu), or None if not built."""
    path = os.path.join(HERE, "_ref", "libtroute_ref_f32.so")
    if not os.path.exists(path):
        return None
    ref = C.CDLL(path)
    F = C.c_float
    ref.ref_mc_solve_f32.restype = F
    ref.ref_mc_solve_f32.argtypes = [F] * 14 + [pF, pF]
    ref.ref_cpu_wavefront_f32.restype = None
    ref.ref_cpu_wavefront_f32.argtypes = [C.c_void_p] + [pI] * 6 + [I, I, I] + [pF] * 4 + [pF, pF, I, F]
    return ref


def compute_stats(pred, truth):
    pred = np.ascontiguousarray(pred, np.float64).ravel()
    truth = np.ascontiguousarray(truth, np.float64).ravel()
    out = np.zeros(9)
    lib().orc_compute_stats(_ptr(pred, D), _ptr(truth, D), pred.size, _ptr(out, D))
    keys = ("max_abs", "p50", "p90", "p99", "p99_9", "max_rel", "within_1pct", "within_10pct", "nans")
    return dict(zip(keys, out.tolist()))


_ppI, _ppF = C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.POINTER(C.c_float))
_MATRIX_BUILD_ARGS = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)] + [C.POINTER(C.c_float)] * 4 + \
    [C.c_float, _ppI, _ppI, _ppF, _ppI, _ppI, _ppF, C.POINTER(C.c_int)]


def load_ref():
    """The reference's own host routines (oracle/_ref), or None if not built."""
    path = os.path.join(HERE, "_ref", "libtroute_ref.so")
    if not os.path.exists(path):
        return None
    ref = C.CDLL(path)
    ref.ref_mc_solve_d.restype = D
    ref.ref_mc_solve_d.argtypes = [D] * 14 + [pD, pD]
    ref.ref_cpu_wavefront_fp64.restype = None
    ref.ref_cpu_wavefront_fp64.argtypes = [C.c_void_p] + [pI] * 6 + [I, I, I] + [pF] * 4 + [pD, pD, I, D]
    ref.ref_cpu_mc.restype = None
    ref.ref_cpu_mc.argtypes = [C.c_void_p, pD, pI, pI, pD, pD, pD, pD, pD, pD, D, D, I]
    if hasattr(ref, "ref_cpu_diffusive_tridiag"):
        ref.ref_cpu_diffusive_tridiag.restype = None
        ref.ref_cpu_diffusive_tridiag.argtypes = [C.c_void_p] + [pF] * 5 + [I]
        ref.ref_gen_diff_data.restype = None
        ref.ref_gen_diff_data.argtypes = [C.c_void_p] + [pF] * 4 + [I, C.c_uint]
    if hasattr(ref, "ref_matrix_build"):
        ref.ref_matrix_coeffs.restype = None
        ref.ref_matrix_coeffs.argtypes = [C.c_void_p, I, C.c_float] + [pF] * 4
        ref.ref_matrix_build.restype = None
        ref.ref_matrix_build.argtypes = _MATRIX_BUILD_ARGS
        ref.ref_free.restype = None
        ref.ref_free.argtypes = [C.c_void_p]
    if hasattr(ref, "ref_matrix_secant_solve"):
        ref.ref_matrix_secant_solve.restype = None
        ref.ref_matrix_secant_solve.argtypes = [I, C.c_float] + [pF] * 10
    if hasattr(ref, "ref_cpu_linear_mc_fp64"):
        ref.ref_cpu_linear_mc_fp64.restype = None
        ref.ref_cpu_linear_mc_fp64.argtypes = [pF] * 4 + [pI] * 6 + [I, I, I] + [pF, pF, pD, I]
    return ref


def linear_coeffs(params32, dt=300.0):
    """compute_mc_coeffs restated (linear_mc.cu:57-100): (C1, C2, C3, C4/ql), float [n] each."""
    p = np.ascontiguousarray(params32, np.float32).reshape(-1, 8)
    n = len(p)
    out = [np.zeros(n, np.float32) for _ in range(4)]
    lib().orc_linear_coeffs(p.ctypes.data_as(C.POINTER(P32)), n, float(dt), *[_ptr(a, C.c_float) for a in out])
    return tuple(out)


def linear_mc_fp64(coeffs, topo, qlat32, Q0, n_timesteps, reference=False):
    """cpu_linear_mc_fp64 restated (linear_mc.cu:238-289); reference=True runs the
    reference's own routine from oracle/_ref instead.  Returns Q_final [N] (double)."""
    t = as_topo(topo)
    keep, targs = _topo_args(t)
    cs = [np.ascontiguousarray(c, np.float32) for c in coeffs]
    q = np.ascontiguousarray(qlat32, np.float32)
    q0 = np.ascontiguousarray(Q0, np.float32)
    out = np.zeros(int(t["n_segments"]))
    fn = load_ref().ref_cpu_linear_mc_fp64 if reference else lib().orc_linear_mc_fp64
    fn(*[_ptr(c, C.c_float) for c in cs], *targs, _ptr(q, C.c_float), _ptr(q0, C.c_float), _ptr(out, D),
       int(n_timesteps))
    return out


DIFF_MAX_NODES = 500
DIFF_REACH = np.dtype([("ncomp", np.int32), ("theta", np.float32), ("dtini", np.float32)])


def diffusive_tridiag(rp, dx, diffusivity, celerity, qp_prev, reference=False):
    """cpu_diffusive_tridiag restated (owp_extended_kernels.cu:35-94); reference=True
    runs the reference's own routine (oracle/_ref).  rp: DIFF_REACH [nreach];
    node arrays float32 [nreach * 500].  Returns qp_out float32."""
    rp = np.ascontiguousarray(rp, DIFF_REACH)
    arrs = [np.ascontiguousarray(a, np.float32) for a in (dx, diffusivity, celerity, qp_prev)]
    out = np.zeros_like(arrs[3])
    fn = load_ref().ref_cpu_diffusive_tridiag if reference else lib().orc_diffusive_tridiag
    fn(rp.ctypes.data, *[_ptr(a, C.c_float) for a in arrs], _ptr(out, C.c_float), len(rp))
    return out


def gen_diff_data_reference(nreach, seed=42):
    """The reference's own gen_diff_data (owp_extended_kernels.cu:347-362, glibc
    rand) from oracle/_ref: (rp, dx, diffusivity, celerity, qp)."""
    rp = np.zeros(nreach, DIFF_REACH)
    arrs = [np.zeros(nreach * DIFF_MAX_NODES, np.float32) for _ in range(4)]
    load_ref().ref_gen_diff_data(rp.ctypes.data, *[_ptr(a, C.c_float) for a in arrs], nreach, seed)
    return (rp, *arrs)


# ------------------------------------------------------------ matrix form
def matrix_coeffs(params32, dt=300.0, reference=False):
    """compute_mc_coeffs_cpu restated (linear_mc_matrix.cu:58-84), X = 0.25;
    reference=True runs the reference's own routine from oracle/_ref."""
    p = np.ascontiguousarray(params32, np.float32).reshape(-1, 8)
    n = len(p)
    out = [np.zeros(n, np.float32) for _ in range(4)]
    if reference:
        load_ref().ref_matrix_coeffs(p.ctypes.data, n, float(dt), *[_ptr(a, C.c_float) for a in out])
    else:
        lib().orc_matrix_coeffs(p.ctypes.data_as(C.POINTER(P32)), n, float(dt), *[_ptr(a, C.c_float) for a in out])
    return tuple(out)


def matrix_build(up_ptr, up_idx, coeffs, tol, reference=False):
    """build_propagation_matrices restated (linear_mc_matrix.cu:118-229); reference=True
    runs the reference's own routine.  Returns (P_Q, P_q), each (rowptr, col, val)."""
    ptr = np.ascontiguousarray(up_ptr, np.int32)
    idx = np.ascontiguousarray(up_idx, np.int32)
    n = len(ptr) - 1
    cs = [np.ascontiguousarray(c, np.float32) for c in coeffs]
    outs = [C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_float)(),
            C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_float)()]
    nnz = np.zeros(2, np.int32)
    args = [n, _ptr(ptr, C.c_int), _ptr(idx if len(idx) else np.zeros(1, np.int32), C.c_int),
            *[_ptr(c, C.c_float) for c in cs], float(tol), *[C.byref(o) for o in outs], _ptr(nnz, C.c_int)]
    if reference:
        R = load_ref()
        R.ref_matrix_build(*args)
        free = R.ref_free
    else:
        if lib().orc_matrix_build(*args) != 0:
            raise ValueError("an upstream reach index is not below its reach")
        free = lib().orc_free

    def take(i, count, dt):
        a = np.ctypeslib.as_array(outs[i], (count,)).copy() if count else np.zeros(0, dt)
        free(outs[i])
        return a
    PQ = (take(0, n + 1, np.int32), take(1, int(nnz[0]), np.int32), take(2, int(nnz[0]), np.float32))
    Pq = (take(3, n + 1, np.int32), take(4, int(nnz[1]), np.int32), take(5, int(nnz[1]), np.float32))
    return PQ, Pq


def matrix_route(PQ, Pq, qlat32, Q0, n_timesteps, lanes=1):
    """The matrix-form steps in float over each row's union of its P_Q and P_q
    entries, with the summation order of `lanes` lanes per row (1 = sequential;
    0 = by union length U, the smallest power of two G with 2 G >= U, at most
    64).  Returns Q after n_timesteps steps, float [n]."""
    n = len(PQ[0]) - 1
    arr = [np.ascontiguousarray(a) for a in (*PQ, *Pq)]
    q = np.ascontiguousarray(qlat32, np.float32)
    Q = np.array(Q0, np.float32, copy=True)
    z = lambda a, t: _ptr(a if len(a) else np.zeros(1, a.dtype), t)
    lib().orc_matrix_route(n, z(arr[0], C.c_int), z(arr[1], C.c_int), z(arr[2], C.c_float),
                           z(arr[3], C.c_int), z(arr[4], C.c_int), z(arr[5], C.c_float),
                           z(q.reshape(-1), C.c_float), _ptr(Q, C.c_float), int(n_timesteps), int(lanes))
    return Q


def matrix_secant_inputs(topo, params32, qlat0):
    """The per-reach secant inputs of linear_mc_matrix_secant.cu main (:358-412)
    restated in float32: the mean first-step inflow (sequential double sum over
    all segments), the steady reference flows in level order (upstream sum in
    CSR order + own inflow + mean), then qup = quc = max(upstream sum, 0.1),
    qdp = max(reference flow, 0.1), ql = max(own inflow, 0.01) and
    depthp = max(0.1, (qdp / (tw + 1))^0.6).  Returns (qup, quc, qdp, ql,
    depthp, reference flows), float32 [n_reaches] each.  Python loops: small
    networks only."""
    t = as_topo(topo)
    f = np.float32
    q0 = np.ascontiguousarray(qlat0, np.float32)
    p = np.ascontiguousarray(params32, np.float32).reshape(-1, 8)
    R, L = int(t["n_reaches"]), int(t["n_levels"])
    mean = f(np.cumsum(q0.astype(np.float64))[-1] / len(q0))
    ptr, idx = np.asarray(t["reach_up_ptr"]), np.asarray(t["reach_up_idx"])
    lp, lr = np.asarray(t["level_ptr"]), np.asarray(t["level_reach"])
    refq = np.zeros(R, np.float32)
    upsum = np.zeros(R, np.float32)
    for lv in range(L):
        for r in lr[lp[lv]:lp[lv + 1]]:
            s_ = f(0.0)
            for k in range(ptr[r], ptr[r + 1]):
                s_ = f(s_ + refq[idx[k]])
            upsum[r] = s_
            refq[r] = f(f(s_ + q0[r]) + mean)
    qdp = np.fmax(refq, f(0.1))
    qup = np.fmax(upsum, f(0.1))
    ql = np.fmax(q0[:R], f(0.01))
    # the reference's std::pow(float, float) is the C library's powf; numpy's
    # float32 power rounds differently in about a fifth of cases
    libm = C.CDLL(ctypes_util.find_library("m"))
    libm.powf.restype, libm.powf.argtypes = C.c_float, [C.c_float, C.c_float]
    base = (qdp / (p[:R, 2] + f(1.0))).astype(np.float32)
    pw = np.array([libm.powf(float(b), 0.6) for b in base], np.float32)
    depthp = np.fmax(f(0.1), pw)
    return qup, qup.copy(), qdp, ql, depthp, refq


def matrix_secant_solve_reference(inputs, params32, dt=300.0):
    """The reference's own mc_secant_solve_cpu (oracle/_ref) per reach on the
    given inputs: (C1, C2, C3, C4/ql)."""
    qup, quc, qdp, ql, depthp = [np.ascontiguousarray(a, np.float32) for a in inputs[:5]]
    p = np.ascontiguousarray(params32, np.float32).reshape(-1, 8)
    n = len(qup)
    out = [np.zeros(n, np.float32) for _ in range(4)]
    load_ref().ref_matrix_secant_solve(n, float(dt), *[_ptr(a, C.c_float) for a in (qup, quc, qdp, ql)],
                                       _ptr(p, C.c_float), _ptr(depthp, C.c_float),
                                       *[_ptr(o, C.c_float) for o in out])
    return tuple(out)
