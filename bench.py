"""Benchmark: FP64 Muskingum-Cunge routing of the 1,000,000-segment Shreve-style
river tree over 288 steps (dt = 300 s) on one MI355X, with the learned depth
guess against the plain secant, both converged to 1e-13 in depth (1e-9 bar).

A "step" of this benchmark is one full routing pass: reset to the initial state
(device-to-device copies) and route all 288 timesteps of lateral inflow that
already sit in HBM.  The metric is segment-steps per second
(1e6 segments x 288 timesteps per pass / seconds per pass) of the LEARNED mode
(MLP guess, FP32 on MFMA inside the routing kernel, FP64 secant refinement).
The same passes are timed for
  plain       CONVERGED mode: the same FP64 secant from the reference's bracket
              (0.67 d, 1.33 d + 0.01); speedup_vs_plain = learned / plain
  reference   the drop-in REFERENCE rule (the reference's 1 % / 1 cm stop)
The MLP is trained first (not timed) by the library's HIP kernels on pairs
from its FP64 normal-depth solver.

Prints one JSON line with the driver's fields plus
  roofline      HBM roofline of the routing kernel (112 algorithmic bytes per
                segment-step: params 64, qlat 8, previous flow + depth 16,
                writes of flow, depth, velocity 24)
  cpu_baseline  the reference's own FP64 host routine cpu_wavefront_fp64
                (oracle/_ref, built from the reference sources) timed on one
                host core over the first 24 steps of the same workload
  parity        the reference-rule GPU pass vs that CPU run over those 24
                steps, and the learned pass's first step re-solved per segment
                by the FP64 oracle (converged rule) on a sample of segments
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


MODES = {"learned": 2, "plain": 1, "reference": 0}


def time_mode(trmc, w, T, dt, mode, args, weights=None):
    """Warm-up then timed passes of one mode; returns (record, network)."""
    M = w.qlat.shape[0]
    net = trmc.Network(w.topo, mode=mode, schedule=args.schedule, unit_segments=args.unit_segments,
                       block_size=args.block, grid_blocks=args.grid, n_members=M,
                       batch_iters=args.batch_iters)
    net.set_params(w.params)
    net.set_state(w.qdp0, w.dp0, w.reach_last0)
    net.set_forcing(w.qlat)
    if weights is not None:
        net.set_mlp(weights)
    for _ in range(args.warmup):
        net.reset()
        net.run(T, dt)
    trmc.synchronize()
    kernel_ms, iters, fallbacks, capped, stats = [], 0, 0, 0, {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        net.reset()
        stats = net.run(T, dt)
        kernel_ms.append(stats["kernel_ms"])
        iters += stats["secant_iters"]
        fallbacks += stats["fallbacks"]
        capped += stats["not_converged"]
    trmc.synchronize()
    secs = time.perf_counter() - t0
    seg_steps = w.topo.n_segments * T * M
    rec = {"value": seg_steps * args.steps / secs, "ms_per_step": 1e3 * secs / args.steps,
           "kernel_ms": float(np.mean(kernel_ms)),
           "evaluations_per_seg_step": iters / (seg_steps * args.steps),
           "fallbacks": fallbacks // args.steps, "not_converged": capped // args.steps,
           "grid_blocks": stats.get("grid_blocks")}
    log(f"[bench] mode {mode}: {rec['ms_per_step']:.2f} ms/pass, kernel {rec['kernel_ms']:.2f} ms, "
        f"{rec['value'] / 1e9:.3f} G seg-steps/s, evaluations/seg-step {rec['evaluations_per_seg_step']:.3f}, "
        f"fallbacks {rec['fallbacks']}, capped {rec['not_converged']}")
    return rec, net


def train_mlp(trmc, args):
    t0 = time.perf_counter()
    trmc.mlp_make_dataset(args.train_pairs, seed=7)
    t1 = time.perf_counter()
    W, loss = trmc.mlp_train(trmc.mlp_init_weights(3), epochs=args.train_epochs, batch=args.train_batch,
                             lr=args.train_lr)
    t2 = time.perf_counter()
    Xh, Yh, _ = trmc.mlp_make_dataset(1 << 16, seed=99)
    rmse = float(np.sqrt(np.mean((trmc.mlp_predict(W, Xh) - Yh) ** 2)))
    rec = {"pairs": args.train_pairs, "epochs": args.train_epochs, "batch": args.train_batch,
           "dataset_s": t1 - t0, "train_s": t2 - t1, "loss_first": float(loss[0]), "loss_last": float(loss[-1]),
           "heldout_rmse_log_depth": rmse}
    log(f"[bench] mlp: {args.train_pairs} pairs in {t1 - t0:.2f} s, {args.train_epochs} epochs in "
        f"{t2 - t1:.2f} s, loss {loss[0]:.4g} -> {loss[-1]:.4g}, held-out rmse {rmse:.4f}")
    return W, rec


def oracle_sample_check(orc, w, net, dt, n_sample):
    """Re-solve a sample of segments of the first step with the FP64 oracle
    (converged rule) from the GPU's own inputs; returns the max relative flow
    and depth errors."""
    topo = w.topo
    net.reset()
    net.run(1, dt)
    q1, d1, _, rl1 = net.state()
    q1, d1, rl1 = q1[0], d1[0], rl1[0]
    q0, d0, rl0 = w.qdp0, w.dp0, w.reach_last0
    starts, lens = np.asarray(topo.reach_seg_start), np.asarray(topo.reach_seg_len)
    offs = np.arange(int(lens.sum())) - np.repeat(np.cumsum(lens) - lens, lens)
    seg_reach = np.full(topo.n_segments, -1)
    seg_reach[np.repeat(starts, lens) + offs] = np.repeat(np.arange(topo.n_reaches), lens)
    first = np.zeros(topo.n_segments, bool)
    first[starts[lens > 0]] = True
    rng = np.random.default_rng(5)
    covered = np.nonzero(seg_reach >= 0)[0]
    idx = np.sort(rng.choice(covered, min(n_sample, len(covered)), replace=False))
    worst_q = worst_d = 0.0
    for s in idx:
        if first[s]:
            r = seg_reach[s]
            ups = topo.reach_up_idx[topo.reach_up_ptr[r]:topo.reach_up_ptr[r + 1]]
            qup = float(np.sum(rl0[ups])) if len(ups) else 0.0
            quc = float(np.sum(rl1[ups])) if len(ups) else 0.0
        else:
            qup, quc = float(q0[s - 1]), float(q1[s - 1])
        q, h, _ = orc.solve_converged(w.params[s], qup, quc, float(q0[s]), float(w.qlat[0][0][s]),
                                      float(d0[s]), dt, variant="cr")
        worst_q = max(worst_q, abs(q1[s] - q) / max(abs(q), 1e-300) if q != 0 else abs(q1[s]))
        worst_d = max(worst_d, abs(d1[s] - h) / max(abs(h), 1e-300) if h != 0 else abs(d1[s]))
    return worst_q, worst_d, len(idx)


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
    ap.add_argument("--mode", default="learned", choices=sorted(MODES),
                    help="mode of the headline value; the other two are timed beside it")
    ap.add_argument("--only", action="store_true", help="time the headline mode only")
    ap.add_argument("--train-pairs", type=int, default=2_000_000)
    ap.add_argument("--train-epochs", type=int, default=8)
    ap.add_argument("--train-batch", type=int, default=1024)
    ap.add_argument("--train-lr", type=float, default=3e-3)
    ap.add_argument("--oracle-sample", type=int, default=20000,
                    help="segments of the first learned step re-solved by the FP64 oracle")
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

    seg_steps = w.topo.n_segments * T * M
    weights, mlp_rec = None, None
    order = [args.mode] + ([] if args.only else [m for m in ("learned", "plain", "reference") if m != args.mode])
    if "learned" in order:
        weights, mlp_rec = train_mlp(trmc, args)
    recs, nets, finals = {}, {}, {}
    for name in order:
        recs[name], nets[name] = time_mode(trmc, w, T, dt, MODES[name], args,
                                           weights if name == "learned" else None)
        finals[name] = nets[name].state()
        if name not in ("learned", "reference"):
            nets.pop(name).close()
    head = recs[args.mode]
    info = nets[args.mode].info(T)
    value, ms_per_step = head["value"], head["ms_per_step"]
    k_avg_s = head["kernel_ms"] / 1e3
    achieved = BYTES_PER_SEG_STEP * seg_steps / k_avg_s / 1e9
    if "learned" in finals and "plain" in finals:
        # both converge to 1e-13 in depth: their final states after 288 steps agree to the 1e-9 bar
        a_, b_ = finals["learned"], finals["plain"]
        rel = lambda x, y: float(np.max(np.abs(x - y) / np.maximum(np.abs(y), 1e-3)))
        recs["learned"]["max_rel_vs_plain"] = {"qdc": rel(a_[0], b_[0]), "depth": rel(a_[1], b_[1]),
                                               "reach_last": rel(a_[3], b_[3])}
        log(f"[bench] learned vs plain after {T} steps: {recs['learned']['max_rel_vs_plain']}")

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
            "timesteps": T, "members": M, "seed": seed,
            "mode": {"learned": "LEARNED: MLP depth guess (FP32, MFMA, in-kernel) + FP64 secant to 1e-13",
                     "plain": "CONVERGED: FP64 secant from the reference bracket to 1e-13",
                     "reference": "REFERENCE rule (1% / 1 cm stop), drop-in"}[args.mode],
            "schedule": {0: "auto (dataflow)", 1: "time-skewed persistent, device barrier",
                         2: "level launches, hipGraph per step", 3: "dataflow persistent"}[args.schedule],
            "unit_segments": args.unit_segments, "unit_levels": info["n_levels"],
            "super_levels": info["super_levels"], "grid_blocks": head.get("grid_blocks"),
            "block": args.block, "batch_iters": args.batch_iters, "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
        },
        "kernel_ms": 1e3 * k_avg_s,
        "evaluations_per_seg_step": head["evaluations_per_seg_step"],
        "modes": recs,
    }
    if "learned" in recs and "plain" in recs:
        out["speedup_vs_plain"] = recs["learned"]["value"] / recs["plain"]["value"]
        log(f"[bench] learned vs plain secant: {out['speedup_vs_plain']:.3f}x")
    if mlp_rec is not None:
        out["mlp"] = mlp_rec
    if not args.no_cpu:
        Tc = min(args.cpu_steps, T)
        log(f"[bench] cpu baseline: {Tc} steps on one core")
        rec, (qo, ro) = cpu_baseline(w, Tc, dt)
        out["cpu_baseline"] = rec
        rnet = nets.get("reference")
        if rnet is None:
            rnet = trmc.Network(w.topo, unit_segments=args.unit_segments, block_size=args.block)
            rnet.set_params(w.params)
            rnet.set_state(w.qdp0, w.dp0, w.reach_last0)
            rnet.set_forcing(w.qlat)
            nets["reference"] = rnet
        rnet.reset()
        rnet.run(Tc, dt)
        q, _, _, rl = rnet.state()
        rel_q = np.abs(q[0] - qo) / np.maximum(np.abs(qo), 1e-300)
        rel_r = np.abs(rl[0] - ro) / np.maximum(np.abs(ro), 1e-300)
        out["parity"] = {"steps": Tc, "against": rec["kind"],
                         "max_rel_qdc": float(rel_q.max()), "max_rel_reach_last": float(rel_r.max()),
                         "frac_within_1e-12": float(np.mean(np.concatenate([rel_q, rel_r]) <= 1e-12)),
                         "bitwise_equal_frac": float(np.mean(q[0].view(np.uint64) == qo.view(np.uint64)))}
        out["speedup_vs_cpu"] = value / rec["value"]
        log(f"[bench] cpu {rec['value'] / 1e6:.2f} M seg-steps/s; parity {out['parity']}")
        if "learned" in nets and args.oracle_sample > 0:
            orc = _load("trmc_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
            t0 = time.perf_counter()
            eq, ed, n = oracle_sample_check(orc, w, nets["learned"], dt, args.oracle_sample)
            out["parity"]["learned_step1_oracle_sample"] = {
                "segments": n, "max_rel_qdc": eq, "max_rel_depth": ed, "bar": 1e-9}
            log(f"[bench] learned step 1 vs FP64 oracle on {n} segments: max rel q {eq:.3e}, "
                f"depth {ed:.3e} ({time.perf_counter() - t0:.1f} s)")
    for n_ in nets.values():
        n_.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
