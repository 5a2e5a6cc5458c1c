"""Writes the FP32-rule golden vectors under tests/golden/ from the reference's
own FP32 host routines (oracle/_ref/libtroute_ref_f32.so: mc_solve
wavefront_mc.cu:66-193 and cpu_wavefront :413-491, built by
oracle/build_ref.sh from the reference sources).  Run from the repository root:

    make -C oracle && python tests/golden/make_golden_f32.py

Fixtures (inputs + expected outputs, numpy .npz):
  mc_solve_f32_cases.npz  mc_solve on the inputs of mc_solve_cases.npz, as float
  wavefront_f32_c1.npz    cpu_wavefront on the seeded 1,000-segment binary-
                          confluence tree of wavefront_c1.npz, 288 steps, dt = 300 s
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.dirname(os.path.abspath(__file__))
F, I = C.c_float, C.c_int
pF, pI = C.POINTER(F), C.POINTER(I)


def load_ref():
    ref = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libtroute_ref_f32.so"))
    ref.ref_mc_solve_f32.restype = F
    ref.ref_mc_solve_f32.argtypes = [F] * 14 + [pF, pF]
    ref.ref_cpu_wavefront_f32.restype = None
    ref.ref_cpu_wavefront_f32.argtypes = [C.c_void_p] + [pI] * 6 + [I, I, I] + [pF] * 4 + [pF, pF, I, F]
    return ref


def ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def solve_cases(ref):
    z = np.load(os.path.join(OUT, "mc_solve_cases.npz"))
    params = z["params"].astype(np.float32)
    inputs = z["inputs"].astype(np.float32)
    out = np.zeros((len(params), 3), np.float32)
    for i in range(len(params)):
        d, v = F(), F()
        x = inputs[i]
        q = ref.ref_mc_solve_f32(300.0, x[0], x[1], x[2], x[3], *params[i], x[4], C.byref(d), C.byref(v))
        out[i] = (q, d.value, v.value)
    np.savez_compressed(os.path.join(OUT, "mc_solve_f32_cases.npz"), params=params, inputs=inputs,
                        expected=out, dt=300.0)
    print("mc_solve_f32_cases:", len(params))


def wavefront_c1(ref):
    z = np.load(os.path.join(OUT, "wavefront_c1.npz"))
    R, L, N = int(z["n_reaches"]), int(z["n_levels"]), int(z["n_segments"])
    T = int(z["n_timesteps"])
    arrs = [np.ascontiguousarray(z[k], np.int32) for k in
            ("reach_seg_start", "reach_seg_len", "reach_up_ptr", "reach_up_idx", "level_ptr", "level_reach")]
    p = np.ascontiguousarray(z["params"], np.float32)
    ql = np.ascontiguousarray(z["qlat"], np.float32)
    q0, d0, r0 = (np.ascontiguousarray(z[k], np.float32) for k in ("qdp0", "dp0", "reach_last0"))
    res = {}
    for steps in (T, 8):
        qout, rout = np.zeros(N, np.float32), np.zeros(R, np.float32)
        ref.ref_cpu_wavefront_f32(p.ctypes.data, *[ptr(a, I) for a in arrs], R, L, N, ptr(ql, F), ptr(q0, F),
                                  ptr(d0, F), ptr(r0, F), ptr(qout, F), ptr(rout, F), steps, 300.0)
        res[steps] = (qout, rout)
    np.savez_compressed(os.path.join(OUT, "wavefront_f32_c1.npz"), n_timesteps=T,
                        qdc_out=res[T][0], reach_last_out=res[T][1], qdc_out_8=res[8][0],
                        reach_last_out_8=res[8][1])
    print("wavefront_f32_c1: N", N, "T", T)


if __name__ == "__main__":
    ref = load_ref()
    solve_cases(ref)
    wavefront_c1(ref)
