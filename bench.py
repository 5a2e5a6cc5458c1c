"""Benchmark: FP64 Muskingum-Cunge routing on one MI355X.

Default workload (shreve-1m): the 1,000,000-segment Shreve-style river tree,
288 steps of dt = 300 s, with the LEARNED depth guess against the plain
secant, both ending every root in its canonical cell (the same bits as each
other and as the FP64 oracle).  A "step" of this
benchmark is one full routing pass: reset to the initial state (device-to-
device copies) and route all 288 timesteps of lateral inflow that already sit
in HBM.  The metric is segment-steps per second of the headline mode
(segments x members x 288 timesteps per pass / seconds per pass).  Beside it,
the same passes are timed for
  plain       CONVERGED mode: the same FP64 secant from the reference's bracket
              (0.67 d, 1.33 d + 0.01); speedup_vs_plain = learned / plain
  powerlaw    LEARNED without the network: the same anchored start with the
              wide-channel exponent 3/5 alone (ablation of the network)
  reference   the drop-in REFERENCE rule (the reference's 1 % / 1 cm stop)
The MLP is trained first (not timed) by the library's HIP kernels on pairs
from its FP64 normal-depth solver.

Other workloads (--workload):
  batch-64x8     64 independent Shreve networks, sizes log-uniform in
                 [1e4, 2e5] segments, x 8 lateral-inflow members, 288 steps,
                 one launch, schedule picked by the library's autotuner;
                 headline mode REFERENCE (plain FP64 MC)
  shreve-1m-x16  the 1M network x 16 members, LEARNED, one launch
  tree-1k        the reference's own 1,000-segment binary tree
  linear-1m      the linearized MC (wavefront/linear_mc.cu) on the 1M network
  diffusive-10k  the diffusive-wave tridiagonal step on 10,000 reaches
  matrix-309k    the matrix form of the linearized MC (linear_mc_matrix*.cu):
                 309,000 one-segment reaches, 24 steps, tol 1e-6
  matrix-secant-309k  the same with the secant-fitted coefficients of
                 linear_mc_matrix_secant.cu, tol 1e-7
  wavefront-1m-4 the FP32 REFERENCE rule (the cpu_wavefront drop-in) on a
                 network shaped like the reference's own results/run_1m_4.txt
                 run (its generator's construction at 1M target reaches:
                 ~872k reaches, ~10.5M segments, ~70 levels, 4 steps, its
                 parameter and forcing distributions), beside the FP64
                 REFERENCE rule on the same inputs; CPU baseline: the
                 reference's own float cpu_wavefront over all 4 steps

Prints one JSON line with the driver's fields plus
  roofline      HBM roofline of the routing kernel (112 algorithmic bytes per
                segment-step for one member: params 64, qlat 8, previous flow +
                depth 16, writes of flow, depth, velocity 24; 48 + 64/M per
                member-segment-step with M members sharing params)
  cpu_baseline  the reference's own FP64 host routine cpu_wavefront_fp64
                (oracle/_ref, built from the reference sources) timed on one
                host core over the first steps of member 0
  parity        the reference-rule GPU pass vs that CPU run over those steps,
                and the learned pass's first step re-solved per segment by
                the FP64 oracle (converged rule) on a sample of segments
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


# kind: generator (1 binary tree, 2 Shreve, "batch" = union of Shreve
# networks); n: segments (batch: networks); r: reaches (batch: size range);
# T timesteps, M members, seed, headline mode, CPU sample steps
WORKLOADS = {
    "shreve-1m": dict(kind=2, n=1_000_000, r=300_000, T=288, M=1, seed=42, mode="learned", cpu_steps=24),
    "shreve-1m-x16": dict(kind=2, n=1_000_000, r=300_000, T=288, M=16, seed=42, mode="learned", cpu_steps=24,
                          conv_units=1),
    "batch-64x8": dict(kind="batch", n=64, r=(10_000, 200_000), T=288, M=8, seed=3, mode="reference", cpu_steps=5),
    "tree-1k": dict(kind=1, n=1000, r=0, T=288, M=1, seed=42, mode="learned", cpu_steps=288),
    "linear-1m": dict(kind=2, n=1_000_000, r=300_000, T=288, M=1, seed=42, mode="linear", cpu_steps=288, units=8),
    "diffusive-10k": dict(kind="diffusive", n=10_000, r=0, T=1, M=1, seed=42, mode="diffusive", cpu_steps=1),
    "matrix-309k": dict(kind=2, n=309_000, r=309_000, T=24, M=1, seed=42, mode="matrix", cpu_steps=24),
    "matrix-secant-309k": dict(kind=2, n=309_000, r=309_000, T=24, M=1, seed=42, mode="matrix", cpu_steps=24,
                               coeffs="secant", tol=1e-7),
    "wavefront-1m-4": dict(kind=3, n=0, r=1_000_000, T=4, M=1, seed=42, mode="reference_f32", cpu_steps=4),
}
MODES = {"learned": 2, "plain": 1, "reference": 0, "powerlaw": 2, "reference_f32": 4}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def bytes_per_seg_step(M):
    return 48.0 + 64.0 / M


def host_forcing(N, T, M, seed):
    """Lateral inflow [M, T, N] float32 for ensembles: lognormal, mean 0.1
    m3/s, sigma_ln 1, member m drawn from numpy's generator seeded (seed, m)
    (member threads in parallel)."""
    from concurrent.futures import ThreadPoolExecutor

    out = np.empty((M, T, N), np.float32)
    mu = np.float32(np.log(0.1) - 0.5)

    def fill(m):
        rng = np.random.default_rng([seed, m])
        for t in range(T):
            x = rng.standard_normal(N, dtype=np.float32)
            np.exp(x + mu, out=out[m, t])

    with ThreadPoolExecutor(max(1, min(M, 16))) as ex:
        list(ex.map(fill, range(M)))
    return out


def make(trmc, spec):
    """(workload, description).  One member: the library's seeded generator
    (float64); ensembles and batches: host_forcing (float32)."""
    if spec["kind"] == "batch":
        sizes = trmc.batch_sizes(spec["n"], spec["r"][0], spec["r"][1], spec["seed"])
        w, so, ro = trmc.make_batch_workload(sizes, spec["T"], spec["M"], seed=spec["seed"], forcing=False)
        desc = (f"{spec['n']} independent Shreve networks ({int(sizes.min())}-{int(sizes.max())} segments, "
                f"log-uniform) x {spec['M']} members, one launch")
    else:
        w = trmc.make_workload(spec["kind"], spec["n"], spec["T"], n_reaches=spec["r"], n_members=1,
                               seed=spec["seed"], forcing=spec["M"] == 1)
        desc = "Shreve-style tree" if spec["kind"] == 2 else "binary-confluence tree"
        if spec["M"] > 1:
            desc += f" x {spec['M']} members"
    if w.qlat is None:
        w.qlat = host_forcing(w.topo.n_segments, spec["T"], spec["M"], spec["seed"] ^ 0xA1A1)
    return w, desc


def new_net(trmc, w, M, mode, opts, T, weights=None, guess_model=0):
    net = trmc.Network(w.topo, mode=mode, n_members=M, guess_model=guess_model, **opts)
    net.set_params(w.params)
    net.set_state(w.qdp0, w.dp0, w.reach_last0)
    net.set_forcing(w.qlat if T == w.qlat.shape[1] else np.ascontiguousarray(w.qlat[:, :T]))
    if weights is not None:
        net.set_mlp(weights)
    return net


def time_mode(trmc, w, M, T, dt, name, opts, args, weights=None):
    """Warm-up then timed passes of one mode; returns (record, network)."""
    net = new_net(trmc, w, M, MODES[name], opts, T, weights if name == "learned" else None,
                  guess_model=1 if name == "powerlaw" else 0)
    for _ in range(args.warmup):
        net.reset()
        net.run(T, dt)
    trmc.synchronize()
    kernel_ms, iters, fallbacks, capped, mlp_calls, mlp_frac, stats = [], 0, 0, 0, 0, 0.0, {}
    certified = noncanon = 0
    wait_frac = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        net.reset()
        stats = net.run(T, dt)
        kernel_ms.append(stats["kernel_ms"])
        iters += stats["secant_iters"]
        fallbacks += stats["fallbacks"]
        capped += stats["not_converged"]
        mlp_calls += stats["mlp_calls"]
        certified += stats.get("certified_solves", 0)
        noncanon += stats.get("noncanonical", 0)
        mlp_frac += stats["mlp_frac"]
        wait_frac += stats["barrier_wait_frac"]
    trmc.synchronize()
    secs = time.perf_counter() - t0
    seg_steps = w.topo.n_segments * T * M
    rec = {"value": seg_steps * args.steps / secs, "ms_per_step": 1e3 * secs / args.steps,
           "kernel_ms": float(np.mean(kernel_ms)),
           "evaluations_per_seg_step": iters / (seg_steps * args.steps),
           "fallbacks": fallbacks // args.steps, "not_converged": capped // args.steps,
           "certified_solves": certified // args.steps,
           "noncanonical_solves": noncanon // args.steps,
           "grid_blocks": stats.get("grid_blocks"),
           # the dataflow schedule has no grid barrier: the share of lane time
           # spent waiting for upstream outflows takes the barrier stall's
           # place (sampled on the first lane of every block)
           "upstream_wait_frac": wait_frac / args.steps}
    if name == "learned":
        rec["network_forwards"] = mlp_calls // args.steps
        rec["network_time_frac"] = mlp_frac / args.steps
    log(f"[bench] {name}: {rec['ms_per_step']:.2f} ms/pass, kernel {rec['kernel_ms']:.2f} ms, "
        f"{rec['value'] / 1e9:.3f} G seg-steps/s, evaluations/seg-step {rec['evaluations_per_seg_step']:.3f}, "
        f"fallbacks {rec['fallbacks']}, capped {rec['not_converged']}, certified {rec['certified_solves']}, "
        f"not canonical {rec['noncanonical_solves']}")
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


def cpu_baseline(trmc, w, T_sample, dt, t_first=0, state0=None):
    """Times the reference routine (or the oracle restatement when the
    reference build is absent) on one core over member 0's steps t_first ..
    t_first + T_sample - 1, from the initial state (t_first = 0) or from
    state0 = (qdc, depth, reach_last) after t_first steps (the routine takes
    them as float, as its initial state); returns (record, outputs, oracle
    module)."""
    orc = _load("trmc_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
    ref = orc.load_ref()
    topo = orc.as_topo(w.topo)
    N, R = w.topo.n_segments, w.topo.n_reaches
    p32 = np.ascontiguousarray(w.params, np.float32)
    q32 = np.ascontiguousarray(w.qlat[0][t_first:t_first + T_sample], np.float32)
    s0 = (w.qdp0, w.dp0, w.reach_last0) if state0 is None else state0
    q0, d0, r0 = (np.ascontiguousarray(x, np.float32) for x in s0)
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
    window = f"first {T_sample} timesteps" if t_first == 0 else f"timesteps {t_first + 1}-{t_first + T_sample}"
    rec = {"value": N * T_sample / secs, "unit": "segment-steps/s", "cores": 1, "kind": kind,
           "sample": f"member 0, {window}, all {N} segments, one thread ({secs:.1f} s); "
                     f"host has {os.cpu_count()} logical CPUs"}
    return rec, (qo, ro), orc


def oracle_sample_check(orc, w, net, dt, n_sample):
    """Re-solve a sample of member 0's segments in the first step with the
    FP64 oracle (converged rule) from the GPU's own inputs; returns the max
    relative flow and depth errors."""
    topo = w.topo
    net.reset()
    net.run(1, dt)
    q1, d1, _, rl1 = net.state()
    q1, d1, rl1 = q1[0], d1[0], rl1[0]
    q0, d0, rl0 = w.qdp0, w.dp0, w.reach_last0
    ql = np.asarray(w.qlat[0][0], np.float64)
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
        q, h, _ = orc.solve_converged(w.params[s], qup, quc, float(q0[s]), float(ql[s]), float(d0[s]), dt,
                                      variant="cr")
        worst_q = max(worst_q, abs(q1[s] - q) / max(abs(q), 1e-3))
        worst_d = max(worst_d, abs(d1[s] - h) / max(abs(h), 1e-2))
    return worst_q, worst_d, len(idx)


def bench_linear(trmc, spec, args):
    """The linearized MC (wavefront/linear_mc.cu) on the same 1M network:
    coefficients from the library's kernel, then LINEAR passes (reset + 288
    steps, inputs resident).  56 algorithmic bytes per segment-step: four
    coefficients 32, lateral inflow 8, previous flow 8, new flow 8.  The CPU
    baseline is the reference's cpu_linear_mc_fp64 over all steps, fed the
    same coefficients, which the GPU must match bit for bit."""
    T, dt = spec["T"], 300.0
    w = trmc.make_workload(spec["kind"], spec["n"], T, n_reaches=spec["r"], seed=spec["seed"])
    topo = w.topo
    p32 = np.ascontiguousarray(w.params, np.float32)
    q0 = np.asarray(w.qdp0, np.float64)
    last = np.array([q0[s + n - 1] for s, n in zip(topo.reach_seg_start, topo.reach_seg_len)])
    t0 = time.perf_counter()
    coeffs = trmc.linear_coeffs(p32, dt)
    coeff_s = time.perf_counter() - t0
    opts = dict(unit_segments=args.unit_segments, block_size=args.block, grid_blocks=args.grid,
                batch_iters=args.batch_iters)
    net = trmc.Network(topo, mode=trmc.MODE_LINEAR, **opts)
    net.set_linear_coeffs(coeffs)
    net.set_state(q0, np.zeros_like(q0), last)
    net.set_forcing(w.qlat)
    for _ in range(args.warmup):
        net.reset()
        net.run(T, dt)
    trmc.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        net.reset()
        kms.append(net.run(T, dt)["kernel_ms"])
    trmc.synchronize()
    secs = (time.perf_counter() - t0) / args.steps
    seg_steps = topo.n_segments * T
    q_gpu = net.state()[0][0].copy()
    k_s = float(np.mean(kms)) / 1e3
    achieved = 56.0 * seg_steps / k_s / 1e9
    log(f"[bench] linear: {1e3 * secs:.2f} ms/pass, kernel {1e3 * k_s:.2f} ms, {seg_steps / secs / 1e9:.3f} G seg-steps/s")
    out = {
        "metric": "segment-steps/sec, 1M-seg FP64 linearized MC (linear_mc.cu)",
        "value": seg_steps / secs, "unit": "segment-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * secs, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded generators; lateral inflow resident in HBM)",
        "config": {"workload": f"linear-1m: FP64 linearized MC, Shreve-style tree, {topo.n_segments} segments, "
                               f"{topo.n_reaches} reaches, {T} steps, coefficients at half bankfull",
                   "segments": topo.n_segments, "reaches": topo.n_reaches, "timesteps": T, "members": 1,
                   "seed": spec["seed"], "mode": "LINEAR", "parallelism": "1 GPU",
                   "unit_segments": args.unit_segments, "block": args.block, "coefficients_s": coeff_s},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "kernel_ms": 1e3 * k_s,
    }
    net.close()
    if not args.no_cpu:
        orc = _load("trmc_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
        use_ref = orc.load_ref() is not None and hasattr(orc.load_ref(), "ref_cpu_linear_mc_fp64")
        q32 = np.ascontiguousarray(w.qlat[0], np.float32)
        t0 = time.perf_counter()
        q_cpu = orc.linear_mc_fp64(coeffs, topo, q32, np.asarray(w.qdp0, np.float32), T, reference=use_ref)
        cs = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": seg_steps / cs, "unit": "segment-steps/s", "cores": 1,
                               "kind": "reference" if use_ref else "port",
                               "sample": f"all {T} steps of all {topo.n_segments} segments, one thread ({cs:.1f} s)"}
        out["parity"] = {"against": "reference" if use_ref else "port", "steps": T,
                         "bitwise_equal_frac": float(np.mean(q_gpu.view(np.uint64) == q_cpu.view(np.uint64))),
                         "max_rel": float(np.max(np.abs(q_gpu - q_cpu) / np.maximum(np.abs(q_cpu), 1e-300)))}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        log(f"[bench] linear cpu {out['cpu_baseline']['value'] / 1e6:.1f} M seg-steps/s; parity {out['parity']}")
    print(json.dumps(out), flush=True)


def f32_workload(trmc, spec):
    """The kind-3 merge tree with the distributions of the reference's network
    generator (wavefront/gen_network.py:253-273): dx U[300,5000], bw U[5,50],
    tw = bw + U[2,20], twcc = tw + U[5,100], n U[0.025,0.08], ncc = n U[1.2,2],
    cs U[0.5,3], s0 U[1e-4,5e-3]; lateral inflow U[0,2]; previous flow
    U[0,15] and depth U[0.3,3]; reach_last0 from the last segments
    (wavefront_mc.cu:676-680).  All float32."""
    T = spec["T"]
    topo = trmc.generate_network(3, 1, spec["r"], spec["seed"])
    N = topo.n_segments
    rng = np.random.default_rng(spec["seed"] ^ 0x5A5A)
    dx, bw = rng.uniform(300.0, 5000.0, N), rng.uniform(5.0, 50.0, N)
    tw = bw + rng.uniform(2.0, 20.0, N)
    twcc = tw + rng.uniform(5.0, 100.0, N)
    n_ch = rng.uniform(0.025, 0.08, N)
    ncc = n_ch * rng.uniform(1.2, 2.0, N)
    cs, s0 = rng.uniform(0.5, 3.0, N), rng.uniform(1e-4, 5e-3, N)
    params = np.stack([dx, bw, tw, twcc, n_ch, ncc, cs, s0], 1).astype(np.float32)
    rng = np.random.default_rng(spec["seed"] ^ 0xA1A1)
    qlat = rng.uniform(0.0, 2.0, (T, N)).astype(np.float32)
    qdp0 = rng.uniform(0.0, 15.0, N).astype(np.float32)
    dp0 = rng.uniform(0.3, 3.0, N).astype(np.float32)
    return trmc.Workload(topo, params, qlat[None], qdp0, dp0, trmc.reach_last0_from(topo, qdp0))


def bench_f32(trmc, spec, args):
    """FP32 REFERENCE rule on the run_1m_4-shaped network: resident-input
    passes of all T steps, the FP64 REFERENCE rule beside it, the reference's
    own float cpu_wavefront as the CPU baseline, and bitwise parity against
    the oracle's float restatement with the correctly rounded powf."""
    T, dt = spec["T"], 300.0
    t0 = time.perf_counter()
    w = f32_workload(trmc, spec)
    N, R = w.topo.n_segments, w.topo.n_reaches
    log(f"[bench] workload wavefront-1m-4: {N} segments, {R} reaches, {w.topo.n_levels} levels, "
        f"max upstream {w.topo.max_upstream}, {T} steps (generated in {time.perf_counter() - t0:.1f} s)")
    opts = dict(schedule=3, unit_segments=args.unit_segments, block_size=args.block, grid_blocks=args.grid,
                batch_iters=args.batch_iters)
    recs, finals = {}, {}
    for name in ("reference_f32", "reference"):
        recs[name], net = time_mode(trmc, w, 1, T, dt, name, opts, args)
        finals[name] = net.state()
        if name == "reference_f32":
            trmc.synchronize()
            t1 = time.perf_counter()
            net.set_forcing(w.qlat)             # the forcing from host memory, as the reference's timed loop
            t2 = time.perf_counter()
            net.reset()
            net.run(T, dt)
            trmc.synchronize()
            t3 = time.perf_counter()
            pcie = {"value": N * T / (t3 - t1), "unit": "segment-steps/s", "upload_s": t2 - t1,
                    "upload_bytes": int(w.qlat.nbytes), "host_dtype": "float32", "pass_s": t3 - t2}
        net.close()
    head = recs["reference_f32"]
    seg_steps = N * T
    # algorithmic bytes of the float drop-in per segment-step: params 32,
    # lateral inflow 4, previous flow and depth 8, writes of flow, depth and
    # velocity 12
    bps = 56.0
    achieved = bps * seg_steps / (head["kernel_ms"] / 1e3) / 1e9
    copy_gbps = trmc.stream_copy(1 << 30, reps=5)
    out = {
        "metric": "segment-steps/sec, FP32 REFERENCE rule (cpu_wavefront drop-in), run_1m_4-shaped network",
        "value": head["value"], "unit": "segment-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (the reference generator's construction and distributions; inputs resident in HBM)",
        "config": {"workload": f"wavefront-1m-4: FP32 REFERENCE rule on a {N}-segment, {R}-reach, "
                               f"{w.topo.n_levels}-level merge tree (confluences of up to 3), {T} steps, dt=300 s",
                   "segments": N, "reaches": R, "reach_levels": w.topo.n_levels, "timesteps": T,
                   "seed": spec["seed"], "unit_segments": args.unit_segments, "block": args.block},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": _traffic("wavefront-1m-4"),
                     "copy_reference_gbps": copy_gbps, "bytes_per_seg_step": bps},
        "kernel_ms": head["kernel_ms"],
        "modes": recs,
        "pcie_inclusive": pcie,
        "reference_gpu_fp32": {"value": 10466683 * 4 / 0.282841, "unit": "segment-steps/s",
                               "hardware": "RTX 3060, per-level launches, per-step host upload",
                               "source": "water-prediction/t-route/wavefront/results/run_1m_4.txt:2-6"},
    }
    if not args.no_cpu:
        orc = _load("trmc_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
        ref = orc.load_ref_f32()
        topo_o = orc.as_topo(w.topo)
        keep, targs = orc._topo_args(topo_o)
        F = orc.C.c_float
        ql = np.ascontiguousarray(w.qlat[0], np.float32)
        qo, ro = np.zeros(N, np.float32), np.zeros(R, np.float32)
        t0 = time.perf_counter()
        if ref is not None:
            ref.ref_cpu_wavefront_f32(w.params.ctypes.data, *targs,
                                      *(orc._ptr(a, F) for a in (ql, w.qdp0, w.dp0, w.reach_last0)),
                                      orc._ptr(qo, F), orc._ptr(ro, F), T, dt)
            kind = "reference"
        else:
            r_ = orc.wavefront_f32(topo_o, w.params, ql, w.qdp0, w.dp0, w.reach_last0, T, dt)
            qo, ro, kind = r_["qdc"], r_["reach_last"], "port"
        secs = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": seg_steps / secs, "unit": "segment-steps/s", "cores": 1, "kind": kind,
                               "sample": f"all {T} steps of the workload, one core ({secs:.1f} s)"}
        out["speedup_vs_cpu"] = head["value"] / out["cpu_baseline"]["value"]
        t0 = time.perf_counter()
        cr = orc.wavefront_f32(topo_o, w.params, ql, w.qdp0, w.dp0, w.reach_last0, T, dt, variant="cr",
                               threads=min(16, os.cpu_count() or 1))
        q, _, _, rl = finals["reference_f32"]
        q32, rl32 = q[0].astype(np.float32), rl[0].astype(np.float32)
        rel = np.abs(q32.astype(np.float64) - qo) / np.maximum(np.abs(qo.astype(np.float64)), 1e-3)
        out["parity"] = {
            "steps": T, "vs_oracle_cr_powf_bitwise": bool(np.array_equal(q32.view(np.uint32), cr["qdc"].view(np.uint32))
                                                         and np.array_equal(rl32.view(np.uint32),
                                                                            cr["reach_last"].view(np.uint32))),
            "oracle_seconds": time.perf_counter() - t0,
            "vs_reference_libm_powf": {"bitwise_equal_frac": float(np.mean(q32 == qo)),
                                       "median_rel": float(np.median(rel)), "max_rel": float(rel.max()),
                                       "within_1pct": float(np.mean(rel <= 0.01))}}
        log(f"[bench] cpu {out['cpu_baseline']}; parity {out['parity']}")
    print(json.dumps(out), flush=True)


def bench_diffusive(trmc, spec, args):
    """The diffusive-wave step (owp_extended_kernels.cu:23-135) on the
    reference's largest benchmark size, 10,000 reaches of 20-99 nodes: inputs
    resident, 200 timed launches.  16 algorithmic bytes per node: spacing,
    diffusivity and previous discharge read, new discharge written.  The CPU
    baseline is the reference's cpu_diffusive_tridiag on one core; the GPU
    must match it bit for bit."""
    n, reps = spec["n"], 200
    rng = np.random.default_rng(spec["seed"])
    rp = np.zeros(n, trmc.DIFF_REACH)
    rp["ncomp"] = rng.integers(20, 100, n)
    rp["theta"] = 0.6
    rp["dtini"] = 60.0
    size = n * trmc.DIFF_MAX_NODES
    dx, df, ce, qp = ((lo + w * rng.random(size)).astype(np.float32)
                      for lo, w in ((100, 500), (10, 100), (0.5, 2), (1, 50)))
    t0 = time.perf_counter()
    out, kms = trmc.diffusive_tridiag(rp, dx, df, ce, qp, reps=reps)
    wall = time.perf_counter() - t0
    nodes = int(rp["ncomp"].sum())
    achieved = 16.0 * nodes / (kms * 1e-3) / 1e9
    log(f"[bench] diffusive: {n} reaches, {nodes} nodes, kernel {1e3 * kms:.1f} us per step")
    res = {
        "metric": "node-steps/sec, diffusive-wave tridiagonal step (owp_extended_kernels.cu)",
        "value": nodes / (kms * 1e-3), "unit": "node-steps/s", "n_gpus": 1, "steps": reps, "warmup": 1,
        "ms_per_step": kms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded, the reference generator's distributions; inputs resident)",
        "config": {"workload": f"diffusive-10k: {n} reaches, {nodes} nodes, theta 0.6, dt 60 s", "reaches": n,
                   "nodes": nodes, "parallelism": "1 GPU", "call_wall_s": wall},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "kernel_ms": kms,
    }
    if not args.no_cpu:
        orc = _load("trmc_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
        use_ref = orc.load_ref() is not None and hasattr(orc.load_ref(), "ref_cpu_diffusive_tridiag")
        t0 = time.perf_counter()
        k = 0
        while True:
            want = orc.diffusive_tridiag(rp, dx, df, ce, qp, reference=use_ref)
            k += 1
            if time.perf_counter() - t0 > 2.0:
                break
        cs = (time.perf_counter() - t0) / k
        res["cpu_baseline"] = {"value": nodes / cs, "unit": "node-steps/s", "cores": 1,
                               "kind": "reference" if use_ref else "port",
                               "sample": f"the whole step, {k} repetitions ({cs * 1e3:.1f} ms each), one thread"}
        res["parity"] = {"against": "reference" if use_ref else "port",
                         "bitwise_equal_frac": float(np.mean(out.view(np.uint32) == want.view(np.uint32)))}
        res["speedup_vs_cpu"] = res["value"] / res["cpu_baseline"]["value"]
        log(f"[bench] diffusive cpu {res['cpu_baseline']['value'] / 1e6:.1f} M node-steps/s; parity {res['parity']}")
    print(json.dumps(res), flush=True)


def _traffic(workload):
    """HBM bytes per launch of the workload's kernel from the committed counter
    summary (two rocprofv3 --pmc runs scaled by a plain copy), or None."""
    tpath = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def bench_matrix(trmc, spec, args):
    """The matrix form of the linearized MC (linear_mc_matrix.cu,
    linear_mc_matrix_graph.cu) at the reference's real-network size: 309,000
    one-segment reaches, 24 steps, entries below 1e-6 dropped.  Coefficients
    (GPU) and propagation matrices (host) are set-up, timed apart; the metric
    is reach-steps per second of the 24 steps replayed as one HIP graph with
    every input resident.  Algorithmic bytes per step: 8 per matrix entry
    of the union of the two matrices (column, P_Q and P_q values), 16 per reach
    for its row record, and one pass over the state pairs {flow, lateral
    inflow} read (8) and written (8) and the next step's lateral inflow (4).
    The CPU baseline is the oracle's sequential float step loop on one core;
    the GPU must equal the oracle's loop in its own summation order bit for
    bit, and stay within 1 % of the FP64 level-scheduled recurrence."""
    n, T, tol = spec["n"], spec["T"], spec.get("tol", 1e-6)
    secant = spec.get("coeffs") == "secant"
    orc = _load("trmc_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
    w = trmc.make_workload(spec["kind"], n, 1, n_reaches=spec["r"], seed=spec["seed"])
    t = w.topo
    p32 = np.asarray(w.params, np.float32)
    rng = np.random.default_rng(spec["seed"])
    qlat = rng.lognormal(np.log(0.1) - 0.5, 1.0, (T, n)).astype(np.float32)
    Q0 = np.asarray(w.qdp0, np.float32)
    t0 = time.perf_counter()
    c = trmc.matrix_secant_coeffs(p32, t, qlat[0]) if secant else trmc.matrix_coeffs(p32)
    t1 = time.perf_counter()
    PQ, Pq = trmc.matrix_build(t.reach_up_ptr, t.reach_up_idx, c, tol)
    t2 = time.perf_counter()
    nnz = (len(PQ[1]), len(Pq[1]))
    copy_gbps = trmc.stream_copy(1 << 30, reps=3)
    sweep = {}
    for lanes in (() if args.only else (8, 16, 32)):
        _, ms = trmc.matrix_route(PQ, Pq, qlat, Q0, T, lanes_per_row=lanes, reps=args.steps)
        sweep[lanes] = ms
        log(f"[bench] matrix: lanes {lanes}: {ms:.3f} ms per {T}-step pass")
    lanes = 0
    out, ms = trmc.matrix_route(PQ, Pq, qlat, Q0, T, lanes_per_row=lanes, reps=args.steps)
    step_ms = ms / T
    rows = lambda m: np.repeat(np.arange(n, dtype=np.int64), np.diff(m[0]))
    union = len(np.union1d(rows(PQ) * n + PQ[1], rows(Pq) * n + Pq[1]))
    step_bytes = 12.0 * union + 16.0 * n + 20.0 * n
    achieved = step_bytes / (step_ms * 1e-3) / 1e9
    log(f"[bench] matrix: {n} reaches, nnz {nnz}, {step_ms * 1e3:.1f} us per step, {achieved:.0f} GB/s")
    res = {
        "metric": "reach-steps/sec, matrix-form linearized MC (linear_mc_matrix_graph.cu)",
        "value": n * T / (ms * 1e-3), "unit": "reach-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded Shreve-style tree, one segment per reach; inputs resident)",
        "config": {"workload": f"{args.workload}: {n} one-segment reaches, {t.n_levels} levels, {T} steps of "
                               f"dt=300 s, tol {tol:g}, "
                               + ("secant-fitted coefficients" if secant else "half-bankfull coefficients (X=0.25)"),
                   "reaches": n, "levels": t.n_levels, "timesteps": T,
                   "tol": tol, "nnz_P_Q": nnz[0], "nnz_P_q": nnz[1], "union_entries": union,
                   "lanes_per_row": "by union row length U (smallest power of two G with 2 G >= U, <= 64)",
                   "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": _traffic(args.workload),
                     "copy_reference_gbps": copy_gbps},
        "ms_per_timestep": step_ms,
        "lanes_sweep_ms_per_pass": sweep,
        "setup_s": {("coefficients_host" if secant else "coefficients_gpu"): t1 - t0, "matrices_host": t2 - t1},
    }
    if not args.no_cpu:
        t0 = time.perf_counter()
        want = orc.matrix_route(PQ, Pq, qlat, Q0, T, lanes=lanes)
        cs = time.perf_counter() - t0
        seq = orc.matrix_route(PQ, Pq, qlat, Q0, T, lanes=1)
        truth = orc.linear_mc_fp64(c, t, qlat, Q0, T)
        rel = np.abs(out.astype(np.float64) - truth) / np.abs(truth)
        res["cpu_baseline"] = {"value": n * T / cs, "unit": "reach-steps/s", "cores": 1, "kind": "port",
                               "sample": f"the whole {T}-step pass ({cs:.2f} s), oracle step loop, one thread"}
        res["parity"] = {"against": "port", "bitwise_equal_frac": float(np.mean(out.view(np.uint32) ==
                                                                             want.view(np.uint32))),
                         "max_rel_vs_sequential_order": float(np.max(np.abs(out - seq) / np.abs(seq))),
                         "vs_fp64_recurrence": {"max_rel": float(rel.max()),
                                                "frac_within_1pct": float(np.mean(rel < 0.01))}}
        res["speedup_vs_cpu"] = res["value"] / res["cpu_baseline"]["value"]
        if orc.load_ref() is not None and hasattr(orc.load_ref(), "ref_matrix_build"):
            t0 = time.perf_counter()
            rQ, rq = orc.matrix_build(t.reach_up_ptr, t.reach_up_idx, c, tol, reference=True)
            res["setup_s"]["reference_matrices_host"] = time.perf_counter() - t0
            res["parity"]["matrices_equal_reference_builder"] = all(
                np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))
                for m, r in ((PQ, rQ), (Pq, rq)) for a, b in zip(m, r))
        log(f"[bench] matrix cpu {res['cpu_baseline']['value'] / 1e6:.1f} M reach-steps/s; parity {res['parity']}")
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="shreve-1m", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default=None, choices=sorted(MODES), help="headline mode (workload default)")
    ap.add_argument("--only", action="store_true", help="time the headline mode only")
    ap.add_argument("--unit-segments", type=int, default=None,
                    help="segments per unit (default: the workload's, 4; 8 for linear-1m)")
    ap.add_argument("--block", type=int, default=64)
    ap.add_argument("--batch-iters", type=int, default=12)
    ap.add_argument("--tol", type=float, default=1e-10,
                    help="CONVERGED/LEARNED tolerance of the root estimate handed to the canonical cell")
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--variant", type=int, default=0,
                    help="DATAFLOW kernel: 0 interleaved, 1 whole item per pass")
    ap.add_argument("--schedule", type=int, default=0, help="0 auto, 1 skewed, 2 level, 3 dataflow")
    ap.add_argument("--autotune", type=int, default=-1,
                    help="probe steps for the library autotuner (0 off; default: on for batch workloads)")
    ap.add_argument("--cpu-steps", type=int, default=0, help="timesteps in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--train-pairs", type=int, default=10_000_000)
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
    spec = WORKLOADS[args.workload]
    # segments per unit: the workload's, else (measured best) 2 for the rules
    # that solve to convergence (1 with 16 members) and 4 for the REFERENCE rule
    units_given = args.unit_segments is not None
    if args.unit_segments is None:
        args.unit_segments = spec.get("units", 4)
    if spec["mode"] == "linear":
        bench_linear(trmc, spec, args)
        return
    if spec["mode"] == "diffusive":
        bench_diffusive(trmc, spec, args)
        return
    if spec["mode"] == "matrix":
        bench_matrix(trmc, spec, args)
        return
    if spec["mode"] == "reference_f32":
        bench_f32(trmc, spec, args)
        return
    T, M = spec["T"], spec["M"]
    head = args.mode or spec["mode"]
    dt = 300.0
    t0 = time.perf_counter()
    w, desc = make(trmc, spec)
    log(f"[bench] workload {args.workload}: {w.topo.n_segments} segments, {w.topo.n_reaches} reaches, "
        f"{w.topo.n_levels} levels, {M} members, {T} steps (generated in {time.perf_counter() - t0:.1f} s)")
    seg_steps = w.topo.n_segments * T * M

    weights, mlp_rec = None, None
    if args.only:
        order = [head]
    elif head == "learned":
        order = ["learned", "plain", "powerlaw", "reference"]
    else:
        order = [head] + [m for m in ("reference", "plain") if m != head]
    if "learned" in order:
        weights, mlp_rec = train_mlp(trmc, args)

    opts = dict(schedule=args.schedule, unit_segments=args.unit_segments, block_size=args.block,
                grid_blocks=args.grid, batch_iters=args.batch_iters, tol=args.tol,
                dataflow_variant=args.variant)
    tuned = None
    n_probe = args.autotune if args.autotune >= 0 else (16 if spec["kind"] == "batch" else 0)
    if n_probe > 0:
        t1 = time.perf_counter()
        best, ms = trmc.autotune(w.topo, w.params, w.qdp0, w.dp0, w.reach_last0, n_members=M, n_probe=n_probe,
                                 qlat=w.qlat[:, :n_probe], mlp_weights=weights, mode=MODES[head],
                                 schedule=args.schedule)
        for k in ("unit_segments", "block_size", "batch_iters", "schedule", "dataflow_variant"):
            opts[k] = int(best[k])
        tuned = {"probe_steps": n_probe, "ms_per_probe_step": ms, "seconds": time.perf_counter() - t1,
                 "unit_segments": opts["unit_segments"], "block": opts["block_size"],
                 "batch_iters": opts["batch_iters"], "schedule": opts["schedule"],
                 "dataflow_variant": opts["dataflow_variant"]}
        log(f"[bench] autotune: {tuned}")

    def mode_opts(name):
        o = dict(opts)
        if not units_given and tuned is None and "units" not in spec:
            o["unit_segments"] = 4 if name == "reference" else spec.get("conv_units", 2)
        return o

    recs, nets, finals = {}, {}, {}
    for name in order:
        recs[name], nets[name] = time_mode(trmc, w, M, T, dt, name, mode_opts(name), args, weights)
        recs[name]["unit_segments"] = mode_opts(name)["unit_segments"]
        finals[name] = nets[name].state()
        if name not in ("learned", head):
            nets.pop(name).close()     # keep one full forcing in HBM at a time
    hrec = recs[head]
    info = nets[head].info(T)
    value, ms_per_step = hrec["value"], hrec["ms_per_step"]
    k_avg_s = hrec["kernel_ms"] / 1e3
    achieved = bytes_per_seg_step(M) * seg_steps / k_avg_s / 1e9
    if "learned" in finals and "plain" in finals:
        # both end every root in its canonical cell: compare their states after all steps
        a_, b_ = finals["learned"], finals["plain"]
        rel = np.abs(a_[0] - b_[0]) / np.maximum(np.abs(b_[0]), 1e-3)
        same = np.ascontiguousarray(a_[0]).view(np.uint64) == np.ascontiguousarray(b_[0]).view(np.uint64)
        recs["learned"]["vs_plain_after_all_steps"] = {
            "max_rel_qdc": float(rel.max()), "segments_beyond_1e-9": int((rel > 1e-9).sum()),
            "segments_beyond_1e-12": int((rel > 1e-12).sum()), "segments_not_bitwise": int((~same).sum()),
            "segments": int(rel.size)}
        log(f"[bench] learned vs plain after {T} steps: {recs['learned']['vs_plain_after_all_steps']}")

    # plain streaming copy of 1 GiB: the achievable HBM rate on this box, and
    # the known-size reference the traffic counters are scaled by
    copy_gbps = trmc.stream_copy(1 << 30, reps=5)
    log(f"[bench] 1 GiB device copy: {copy_gbps:.0f} GB/s")
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
        "data": "synthetic (seeded generators; lateral inflow resident in HBM)",
        "config": {
            "workload": f"{args.workload}: FP64 MC routing, {desc}, {w.topo.n_segments} segments, "
                        f"{w.topo.n_reaches} reaches, {w.topo.n_levels} reach levels, {T} steps, dt=300 s",
            "segments": w.topo.n_segments, "reaches": w.topo.n_reaches, "reach_levels": w.topo.n_levels,
            "timesteps": T, "members": M, "seed": spec["seed"],
            "mode": {"learned": "LEARNED: MLP depth guess (FP32, MFMA, in-kernel) + FP64 secant, canonical cell",
                     "plain": "CONVERGED: FP64 secant from the reference bracket, canonical cell",
                     "powerlaw": "LEARNED start without the network (h ~ Q^(3/5)) + FP64 secant, canonical cell",
                     "reference": "REFERENCE rule (1% / 1 cm stop), drop-in"}[head],
            "schedule": {0: "auto (dataflow)", 1: "time-skewed persistent, device barrier",
                         2: "level launches, hipGraph per step", 3: "dataflow persistent"}[args.schedule],
            "unit_segments": mode_opts(head)["unit_segments"], "unit_levels": info["n_levels"],
            "super_levels": info["super_levels"], "grid_blocks": hrec.get("grid_blocks"),
            "block": opts["block_size"], "batch_iters": opts["batch_iters"], "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "copy_reference_gbps": copy_gbps,
        },
        "kernel_ms": 1e3 * k_avg_s,
        "evaluations_per_seg_step": hrec["evaluations_per_seg_step"],
        "modes": recs,
    }
    if tuned is not None:
        out["autotune"] = tuned
    if "learned" in recs and "plain" in recs:
        out["speedup_vs_plain"] = recs["learned"]["value"] / recs["plain"]["value"]
        log(f"[bench] learned vs plain secant: {out['speedup_vs_plain']:.3f}x")
    if "powerlaw" in recs and "plain" in recs:
        out["speedup_powerlaw_vs_plain"] = recs["powerlaw"]["value"] / recs["plain"]["value"]
    if mlp_rec is not None:
        out["mlp"] = mlp_rec
    if not args.no_cpu:
        Tc = min(args.cpu_steps or spec["cpu_steps"], T)
        log(f"[bench] cpu baseline: {Tc} steps on one core")
        rec, (qo, ro), orc = cpu_baseline(trmc, w, Tc, dt)
        out["cpu_baseline"] = rec
        rnet = nets.get("reference")
        if rnet is None:
            rnet = new_net(trmc, w, M, 0, mode_opts("reference"), Tc)
            nets["reference"] = rnet
        rnet.reset()
        rnet.run(Tc, dt)
        q, _, _, rl = rnet.state()
        rel_q = np.abs(q[0] - qo) / np.maximum(np.abs(qo), 1e-300)
        rel_r = np.abs(rl[0] - ro) / np.maximum(np.abs(ro), 1e-300)
        out["parity"] = {"steps": Tc, "against": rec["kind"], "member": 0,
                         "max_rel_qdc": float(rel_q.max()), "max_rel_reach_last": float(rel_r.max()),
                         "frac_within_1e-12": float(np.mean(np.concatenate([rel_q, rel_r]) <= 1e-12)),
                         "bitwise_equal_frac": float(np.mean(q[0].view(np.uint64) == qo.view(np.uint64)))}
        # the same steps on the oracle with a correctly rounded pow: the GPU
        # REFERENCE rule computes exactly that, so it must agree bit for bit;
        # the reference's own glibc pow differs in about 1 call in 1,000
        # (its pow is slower: the first 8 steps)
        t0 = time.perf_counter()
        Tcr = min(Tc, 8)
        topo_o = orc.as_topo(w.topo)
        qcr, rcr, _ = orc.wavefront_f32in(topo_o, np.ascontiguousarray(w.params, np.float32),
                                          np.ascontiguousarray(w.qlat[0][:Tcr], np.float32), w.qdp0, w.dp0,
                                          w.reach_last0, Tcr, dt, variant="cr")
        qr8, rr8, _ = orc.wavefront_f32in(topo_o, np.ascontiguousarray(w.params, np.float32),
                                          np.ascontiguousarray(w.qlat[0][:Tcr], np.float32), w.qdp0, w.dp0,
                                          w.reach_last0, Tcr, dt, variant="libm")
        rnet.reset()
        rnet.run(Tcr, dt)
        q8, _, _, rl8 = rnet.state()
        rel_cq = np.abs(q8[0] - qcr) / np.maximum(np.abs(qcr), 1e-300)
        rel_cr = np.abs(rl8[0] - rcr) / np.maximum(np.abs(rcr), 1e-300)
        out["parity"]["vs_oracle_cr_pow"] = {
            "steps": Tcr, "max_rel_qdc": float(rel_cq.max()), "max_rel_reach_last": float(rel_cr.max()),
            "bitwise_equal_frac": float(np.mean(np.concatenate([q8[0].view(np.uint64) == qcr.view(np.uint64),
                                                                rl8[0].view(np.uint64) == rcr.view(np.uint64)]))),
            "libm_pow_vs_cr_pow_max_rel": float(np.max(np.abs(qr8 - qcr) / np.maximum(np.abs(qcr), 1e-300))),
            "seconds": time.perf_counter() - t0}
        out["speedup_vs_cpu"] = value / rec["value"]
        # The sample is the run's first steps.  The same number of steps at
        # the end of the run, from the state the GPU REFERENCE rule reaches
        # there, is timed too, and the REFERENCE rule's evaluations per
        # segment-step in both windows are set beside its whole-run mean, so
        # the sample's work per segment-step can be compared with the pass's.
        if T > Tc:
            t_late = T - Tc
            rnet.reset()
            st_a = rnet.run(Tc, dt)
            lnet = new_net(trmc, w, M, 0, mode_opts("reference"), T)    # the whole forcing
            lnet.reset()
            lnet.run(t_late, dt)
            ql_, dl_, _, rll_ = lnet.state()
            st_b = lnet.run(Tc, dt)
            lnet.close()
            rec_late, _, _ = cpu_baseline(trmc, w, Tc, dt, t_first=t_late, state0=(ql_[0], dl_[0], rll_[0]))
            ref_all = recs.get("reference")
            denom = w.topo.n_segments * M * Tc
            out["cpu_baseline_late_window"] = dict(
                rec_late,
                gpu_reference_evaluations_per_seg_step={
                    "first_window": st_a["secant_iters"] / denom, "late_window": st_b["secant_iters"] / denom,
                    "whole_run": ref_all["evaluations_per_seg_step"] if ref_all else None})
            log(f"[bench] cpu late window {rec_late['value'] / 1e6:.2f} M seg-steps/s; "
                f"{out['cpu_baseline_late_window']['gpu_reference_evaluations_per_seg_step']}")
        # a second, labelled baseline: the oracle restatement (same bits as the
        # one-core run) with each level's reaches on all host threads
        nthr = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        qmt, rmt, _ = orc.wavefront_f32in(topo_o, np.ascontiguousarray(w.params, np.float32),
                                          np.ascontiguousarray(w.qlat[0][:Tc], np.float32), w.qdp0, w.dp0,
                                          w.reach_last0, Tc, dt, variant="libm", threads=nthr)
        smt = time.perf_counter() - t0
        out["cpu_baseline_all_cores"] = {
            "value": w.topo.n_segments * Tc / smt, "unit": "segment-steps/s", "cores": nthr, "kind": "port",
            "sample": f"member 0, first {Tc} timesteps, {nthr} threads, level-parallel ({smt:.1f} s)",
            "equal_to_one_core_frac": float(np.mean(qmt.view(np.uint64) == qo.view(np.uint64)))}
        log(f"[bench] cpu all cores {out['cpu_baseline_all_cores']['value'] / 1e6:.2f} M seg-steps/s")
        log(f"[bench] cpu {rec['value'] / 1e6:.2f} M seg-steps/s; parity {out['parity']}")
        if "learned" in nets and args.oracle_sample > 0:
            t0 = time.perf_counter()
            eq, ed, n = oracle_sample_check(orc, w, nets["learned"], dt, args.oracle_sample)
            out["parity"]["learned_step1_oracle_sample"] = {
                "segments": n, "max_rel_qdc": eq, "max_rel_depth": ed, "bar": 1e-9}
            log(f"[bench] learned step 1 vs FP64 oracle on {n} segments: max rel q {eq:.3e}, "
                f"depth {ed:.3e} ({time.perf_counter() - t0:.1f} s)")
    # the same pass with the lateral inflow uploaded from host memory first
    # (the reference times its per-step host-to-device copies); reported beside
    # the resident-input value, never as it
    hnet = nets[head]
    trmc.synchronize()
    t0 = time.perf_counter()
    hnet.set_forcing(w.qlat)
    t1 = time.perf_counter()
    hnet.reset()
    hnet.run(T, dt)
    trmc.synchronize()
    t2 = time.perf_counter()
    out["pcie_inclusive"] = {"value": seg_steps / (t2 - t0), "unit": "segment-steps/s",
                             "upload_s": t1 - t0, "upload_bytes": int(w.qlat.nbytes),
                             "host_dtype": str(w.qlat.dtype), "pass_s": t2 - t1}
    log(f"[bench] with forcing upload: {out['pcie_inclusive']}")
    # the reference copies each step's inflow inside its time loop
    # (wavefront_mc.cu:739); here the forcing stays in page-locked host memory,
    # mapped into the device's address space, and the routing kernel reads each
    # segment-step's inflow over the host link as it routes it
    # (the REFERENCE rule: the converged rules refuse forcing in host memory)
    if w.qlat.nbytes <= (8 << 30) and "reference" in recs and not args.no_cpu:
        mnet = new_net(trmc, w, M, MODES["reference"], mode_opts("reference"), T)
        mnet.run(T, dt)
        ref_state = mnet.state()
        t0 = time.perf_counter()
        qd = mnet.order_forcing(w.qlat)          # device order, page-locked (set-up, not timed)
        mnet.set_forcing_mapped(qd)
        t1 = time.perf_counter()
        mnet.reset()
        st_s = mnet.run(T, dt)
        trmc.synchronize()
        t2 = time.perf_counter()
        same = all(np.array_equal(a.view(np.uint64), b.view(np.uint64)) for a, b in zip(ref_state, mnet.state()))
        out["pcie_mapped_reference"] = {
            "value": seg_steps / (t2 - t1), "unit": "segment-steps/s", "rule": "REFERENCE",
            "resident_ms_per_pass": recs["reference"]["ms_per_step"], "pass_s": t2 - t1,
            "kernel_ms": st_s["kernel_ms"], "order_s": t1 - t0, "host_bytes_read": int(w.qlat.nbytes),
            "host_memory": "page-locked, mapped", "same_bits_as_resident": bool(same)}
        log(f"[bench] REFERENCE, forcing read from host memory during the pass: {out['pcie_mapped_reference']}")
        mnet.close()
        del qd
    for n_ in nets.values():
        n_.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
