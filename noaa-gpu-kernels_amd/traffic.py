"""Memory traffic per kernel launch from two rocprofv3 counter runs.

    rocprofv3 --pmc FETCH_SIZE -d <dirF> -o run --output-format csv -- python3 bench.py ...
    rocprofv3 --pmc WRITE_SIZE -d <dirW> -o run --output-format csv -- python3 bench.py ...
    python noaa-gpu-kernels_amd/traffic.py <dirF>/run_counter_collection.csv \\
        <dirW>/run_counter_collection.csv --kernel k_route_dataflow_il --copy-bytes 1073741824 \\
        [--json out.json]

FETCH_SIZE and WRITE_SIZE are in KiB per dispatch.  Each is scaled by the
reading of the same counter for k_stream_copy (the library's plain copy of
--copy-bytes bytes, read once and written once): factor = bytes / counter.
Prints per-kernel averages and the scaled traffic of the named kernel.
"""
import argparse
import csv
import json
from collections import defaultdict


def per_kernel(path):
    vals = defaultdict(list)
    for r in csv.DictReader(open(path)):
        vals[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return vals


def pick(vals, name):
    keys = [k for k in vals if name in k]
    if not keys:
        raise SystemExit(f"no dispatch of {name}")
    return keys, [v for k in keys for v in vals[k]]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("fetch_csv")
    ap.add_argument("write_csv")
    ap.add_argument("--kernel", required=True)
    ap.add_argument("--copy-bytes", type=float, required=True)
    ap.add_argument("--json")
    args = ap.parse_args()
    F, W = per_kernel(args.fetch_csv), per_kernel(args.write_csv)
    # the copy's warm-up dispatch is excluded (first reading)
    _, cf = pick(F, "k_stream_copy")
    _, cw = pick(W, "k_stream_copy")
    cf, cw = cf[1:] or cf, cw[1:] or cw
    kib = args.copy_bytes / 1024.0
    f_fac = kib / (sum(cf) / len(cf))
    w_fac = kib / (sum(cw) / len(cw))
    print(f"copy of {args.copy_bytes:.0f} B: FETCH_SIZE {sum(cf) / len(cf):.0f} KiB (factor {f_fac:.3f}), "
          f"WRITE_SIZE {sum(cw) / len(cw):.0f} KiB (factor {w_fac:.3f})")
    print(f"{'dispatches':>10} {'FETCH_KiB':>14} {'WRITE_KiB':>14}  kernel")
    for k in sorted(F, key=lambda k: -sum(F[k]) / len(F[k]))[:12]:
        fw = W.get(k, [0.0])
        print(f"{len(F[k]):>10} {sum(F[k]) / len(F[k]):>14.1f} {sum(fw) / len(fw):>14.1f}  {k[:90]}")
    keys, kf = pick(F, args.kernel)
    _, kw = pick(W, args.kernel)
    fetch = sum(kf) / len(kf) * 1024.0 * f_fac
    write = sum(kw) / len(kw) * 1024.0 * w_fac
    out = {"kernel": keys[0], "dispatches": len(kf), "fetch_bytes": fetch, "write_bytes": write,
           "bytes_per_launch": fetch + write, "copy_bytes": args.copy_bytes,
           "fetch_factor": f_fac, "write_factor": w_fac}
    print(json.dumps(out, indent=1))
    if args.json:
        with open(args.json, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
