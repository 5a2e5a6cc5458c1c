"""Benchmark: FP64 Muskingum-Cunge routing of the 1,000,000-segment Shreve-style
river tree over 288 steps (dt = 300 s) on one MI355X.

A "step" of this benchmark is one full routing pass: reset to the initial state
(device-to-device copies) and route all 288 timesteps of lateral inflow that
already sit in HBM.  The metric is segment-steps per second
(1e6 segments x 288 timesteps per pass / seconds per pass).

Prints one JSON line with the driver's fields plus
  roofline      HBM roofline of the routing kernel (112 algorithmic bytes per
                segment-step: params 64, qlat 8, previous flow + depth 16,
                writes of flow, depth, velocity 24)
  cpu_baseline  the reference's own FP64 host routine cpu_wavefront_fp64
                (oracle/_ref, built from the reference sources) timed on one
                host core over the first 24 steps of the same workload
  parity        GPU vs that CPU run over those 24 steps: max relative error and
                the fraction within 1e-12 (the GPU uses a correctly rounded pow,
                the reference the host libm's; see tests/test_gpu_parity.py)
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
BYTES_PER_SEG_STEP = 112.0
HBM_PEAK_GBS = 8000.0


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


WORKLOADS = {
    # name: (generator kind, segments, reaches, timesteps, members, seed)
    "shreve-1m": (2, 1_000_000, 300_000, 288, 1, 42),
    "tree-1k": (1, 1000, 0, 288, 1, 42),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(w, T_sample, dt):
    """Times the reference routine (or the oracle restatement when the
    reference build is absent) on one core; returns (record, outputs)."""
    orc = _load("trmc_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
    ref = orc.load_ref()
    topo = orc.as_topo(w.topo)
    p32 = np.ascontiguousarray(w.params, np.float32)
    q32 = np.ascontiguousarray(w.qlat[0][:T_sample], np.float32)
    q0, d0, r0 = (np.ascontiguousarray(x, np.float32) for x in (w.qdp0, w.dp0, w.reach_last0))
    N, R = w.topo.n_segments, w.topo.n_reaches
    t0 = time.perf_counter()
    if ref is not None:
        import ctypes as C

        keep = [np.ascontiguousarray(topo[k], np.int32) for k in
                ("reach_seg_start", "reach_seg_len", "reach_up_ptr", "reach_up_idx", "level_ptr", "level_reach")]
        qo, ro = np.zeros(N), np.zeros(R)
        ptr = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        ref.ref_cpu_wavefront_fp64(p32.ctypes.data, *[ptr(a, C.c_int) for a in keep], R,
                                   int(topo["n_levels"]), N, ptr(q32, C.c_float), ptr(q0, C.c_float),
                                   ptr(d0, C.c_float), ptr(r0, C.c_float), ptr(qo, C.c_double),
                                   ptr(ro, C.c_double), T_sample, dt)
        kind = "reference"
    else:
        qo, ro, _ = orc.wavefront_f32in(topo, p32, q32, q0, d0, r0, T_sample, dt)
        kind = "port"
    secs = time.perf_counter() - t0
    rec = {"value": N * T_sample / secs, "unit": "segment-steps/s", "cores": 1, "kind": kind,
           "sample": f"first {T_sample} of the {w.qlat.shape[1]} timesteps, all {N} segments, "
                     f"one thread ({secs:.1f} s); host has {os.cpu_count()} logical CPUs"}
    return rec, (qo, ro)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="shreve-1m", choices=sorted(WORKLOADS))
    ap.add_argument("--unit-segments", type=int, default=4)
    ap.add_argument("--block", type=int, default=64)
    ap.add_argument("--batch-iters", type=int, default=8)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--schedule", type=int, default=0, help="0 auto, 1 skewed, 2 level, 3 dataflow")
    ap.add_argument("--cpu-steps", type=int, default=24, help="timesteps in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build routes on one GPU only; run with --gpus 1"}))
        sys.exit(2)

    trmc = _load("noaa_gpu_kernels_amd", os.path.join(ROOT, "noaa-gpu-kernels_amd", "__init__.py"))
    kind, N, R, T, M, seed = WORKLOADS[args.workload]
    dt = 300.0
    t0 = time.perf_counter()
    w = trmc.make_workload(kind, N, T, n_reaches=R, n_members=M, seed=seed)
    log(f"[bench] workload {args.workload}: {w.topo.n_segments} segments, {w.topo.n_reaches} reaches, "
        f"{w.topo.n_levels} levels, {T} steps (generated in {time.perf_counter() - t0:.1f} s)")

    net = trmc.Network(w.topo, schedule=args.schedule, unit_segments=args.unit_segments,
                       block_size=args.block, grid_blocks=args.grid, n_members=M,
                       batch_iters=args.batch_iters)
    net.set_params(w.params)
    net.set_state(w.qdp0, w.dp0, w.reach_last0)
    net.set_forcing(w.qlat)
    info = net.info(T)
    for _ in range(args.warmup):
        net.reset()
        net.run(T, dt)
    trmc.synchronize()
    kernel_ms, iters, stats = [], 0, {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        net.reset()
        stats = net.run(T, dt)
        kernel_ms.append(stats["kernel_ms"])
        iters += stats["secant_iters"]
    trmc.synchronize()
    secs = time.perf_counter() - t0
    ms_per_step = 1e3 * secs / args.steps
    seg_steps = w.topo.n_segments * T * M
    value = seg_steps * args.steps / secs
    k_avg_s = float(np.mean(kernel_ms)) / 1e3
    achieved = BYTES_PER_SEG_STEP * seg_steps / k_avg_s / 1e9
    log(f"[bench] {ms_per_step:.2f} ms/pass, kernel {1e3 * k_avg_s:.2f} ms, {value / 1e9:.3f} G seg-steps/s, "
        f"iters/seg-step {iters / (seg_steps * args.steps):.3f}, barrier wait {stats['barrier_wait_frac']:.3f}")

    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")

    out = {
        "metric": "segment-steps/sec, 1M-seg FP64 MC; learned-guess vs plain secant speedup @1e-9",
        "value": value,
        "unit": "segment-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator; lateral inflow resident in HBM)",
        "config": {
            "workload": f"{args.workload}: FP64 MC routing, {w.topo.n_segments} segments, "
                        f"{w.topo.n_reaches} reaches, {w.topo.n_levels} reach levels, {T} steps, dt=300 s",
            "segments": w.topo.n_segments, "reaches": w.topo.n_reaches, "reach_levels": w.topo.n_levels,
            "timesteps": T, "members": M, "seed": seed, "stopping_rule": "reference (1% / 1 cm)",
            "schedule": {0: "auto (dataflow)", 1: "time-skewed persistent, device barrier",
                         2: "level launches, hipGraph per step", 3: "dataflow persistent"}[args.schedule],
            "unit_segments": args.unit_segments, "unit_levels": info["n_levels"],
            "super_levels": info["super_levels"], "grid_blocks": stats.get("grid_blocks"),
            "block": args.block, "batch_iters": args.batch_iters, "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
        },
        "kernel_ms": 1e3 * k_avg_s,
        "secant_iters_per_seg_step": iters / (seg_steps * args.steps),
        "barrier_wait_frac": stats.get("barrier_wait_frac"),
    }
    if not args.no_cpu:
        Tc = min(args.cpu_steps, T)
        log(f"[bench] cpu baseline: {Tc} steps on one core")
        rec, (qo, ro) = cpu_baseline(w, Tc, dt)
        out["cpu_baseline"] = rec
        net.reset()
        net.run(Tc, dt)
        q, _, _, rl = net.state()
        rel_q = np.abs(q[0] - qo) / np.maximum(np.abs(qo), 1e-300)
        rel_r = np.abs(rl[0] - ro) / np.maximum(np.abs(ro), 1e-300)
        out["parity"] = {"steps": Tc, "against": rec["kind"],
                         "max_rel_qdc": float(rel_q.max()), "max_rel_reach_last": float(rel_r.max()),
                         "frac_within_1e-12": float(np.mean(np.concatenate([rel_q, rel_r]) <= 1e-12)),
                         "bitwise_equal_frac": float(np.mean(q[0].view(np.uint64) == qo.view(np.uint64)))}
        out["speedup_vs_cpu"] = value / rec["value"]
        log(f"[bench] cpu {rec['value'] / 1e6:.2f} M seg-steps/s; parity {out['parity']}")
    net.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
