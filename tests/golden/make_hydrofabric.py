"""Writes tests/golden/hydrofabric.npz: a synthetic hydrofabric GeoPackage's
two tables (our data) and the topology/params/id files that the reference's
own builder, water-prediction/t-route/wavefront/data_parsers/parse_gpkg.py
(main, :43-270), writes from it.  The builder's output order follows its
stack-ordered Kahn sort, which invariants alone do not fix, so the reference
itself is run here on the synthetic file (its input and output paths are
fixed relative to the working directory, so it runs in a scratch directory
with the file placed where it reads it).  Run from the repository root:

    python tests/golden/make_hydrofabric.py

Fixture keys:
  fp_id, fp_toid, fp_hydroseq, fp_lengthkm, fp_areasqkm   flowpaths rows (text ids;
                                                         "" stands for NULL)
  at_link, at_vals [rows, 8], at_null [rows, 8]          flowpath-attributes rows
                                                         (Length_m, n, nCC, BtmWdth,
                                                          TopWdth, TopWdthCC, ChSlp, So)
  topo_bin, params_bin                                   the builder's output bytes
  topo_to_wbid                                           its id table
"""
import os
import runpy
import shutil
import sqlite3
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.dirname(os.path.abspath(__file__))
BUILDER = "/root/reference/water-prediction/t-route/wavefront/data_parsers/parse_gpkg.py"
ATTR_COLS = ("Length_m", "n", "nCC", "BtmWdth", "TopWdth", "TopWdthCC", "ChSlp", "So")


def synthetic_tables(n=4000, seed=7):
    """A random river forest of n waterbodies "wb-<id>" draining through
    nexuses "nex-<downstream id>" (terminals "tnx-<k>"), with the irregular
    rows a real extract holds: duplicate ids, NULL and unparsable ids, NULL
    hydroseq/length/area, attributes missing, NULL, zero or negative."""
    rng = np.random.default_rng(seed)
    ids = rng.choice(np.arange(10, 5_000_000), n, replace=False)
    rows = []
    for k, wid in enumerate(ids):
        if k < 25 or rng.random() < 0.02:
            toid = f"tnx-{int(rng.integers(1, 10**6))}"            # outlet
        else:
            down = ids[int(rng.integers(0, k))]                      # an earlier wb: a forest
            toid = f"nex-{int(down)}"
        if rng.random() < 0.01:
            toid = None
        hyd = None if rng.random() < 0.01 else int(rng.integers(1, 10**7))
        ln = None if rng.random() < 0.02 else float(rng.uniform(0.01, 8.0))
        ar = None if rng.random() < 0.02 else float(rng.uniform(0.1, 50.0))
        rows.append((f"wb-{int(wid)}", toid, hyd, ln, ar))
    for k in range(30):                                              # duplicates: the first one is kept
        r = rows[int(rng.integers(0, n))]
        rows.append((r[0], f"tnx-{k}", r[2], 1.0, 1.0))
    rows += [("wb-x12", "nex-5", 1, 1.0, 1.0), (None, "nex-5", 1, 1.0, 1.0), ("bad", "nex-1", 1, 1.0, 1.0)]
    order = rng.permutation(len(rows))
    rows = [rows[i] for i in order]
    attrs = []
    for wid in ids:
        if rng.random() < 0.05:
            continue                                                 # no attribute row: defaults
        vals = [float(rng.uniform(50, 6000)), float(rng.uniform(0.005, 0.2)), float(rng.uniform(0.01, 0.3)),
                float(rng.uniform(0.5, 60)), float(rng.uniform(0.5, 90)), float(rng.uniform(1, 300)),
                float(rng.uniform(0.05, 4)), float(rng.lognormal(np.log(1e-3), 1.5))]
        for c in range(8):
            u = rng.random()
            if u < 0.03:
                vals[c] = None
            elif u < 0.05:
                vals[c] = 0.0
            elif u < 0.06:
                vals[c] = -abs(vals[c])
        attrs.append([f"wb-{int(wid)}"] + vals)
    attrs.append(["wb-999999999"] + [1.0] * 8)                       # unknown link
    return rows, attrs


def write_gpkg(path, rows, attrs):
    con = sqlite3.connect(path)
    cur = con.cursor()
    cur.execute("CREATE TABLE flowpaths (id TEXT, toid TEXT, hydroseq INTEGER, lengthkm REAL, areasqkm REAL)")
    cur.executemany("INSERT INTO flowpaths VALUES (?, ?, ?, ?, ?)", rows)
    cur.execute('CREATE TABLE "flowpath-attributes" (link TEXT, ' + ", ".join(f'"{c}" REAL' for c in ATTR_COLS) + ")")
    cur.executemany('INSERT INTO "flowpath-attributes" VALUES (?, ?, ?, ?, ?, ?, ?, ?, ?)', attrs)
    con.commit()
    con.close()


def main():
    rows, attrs = synthetic_tables()
    work = tempfile.mkdtemp(prefix="trmc_gpkg_")
    try:
        gdir = os.path.join(work, "extract", "big_one_fixed_1", "config")
        os.makedirs(gdir)
        write_gpkg(os.path.join(gdir, "big_one_fixed_subset.gpkg"), rows, attrs)
        cwd = os.getcwd()
        os.chdir(work)
        try:
            runpy.run_path(BUILDER, run_name="__main__")
        finally:
            os.chdir(cwd)
        out = os.path.join(work, "net_real_309k")
        topo_bin = np.fromfile(os.path.join(out, "topo.bin"), np.uint8)
        params_bin = np.fromfile(os.path.join(out, "params.bin"), np.uint8)
        wbid = np.load(os.path.join(out, "topo_to_wbid.npy"), allow_pickle=False)
    finally:
        shutil.rmtree(work)
    s = lambda v: "" if v is None else v
    f = lambda v: np.nan if v is None else v
    np.savez_compressed(
        os.path.join(OUT, "hydrofabric.npz"),
        fp_id=np.array([s(r[0]) for r in rows]), fp_toid=np.array([s(r[1]) for r in rows]),
        fp_hydroseq=np.array([-1 if r[2] is None else r[2] for r in rows], np.int64),
        fp_lengthkm=np.array([f(r[3]) for r in rows]), fp_areasqkm=np.array([f(r[4]) for r in rows]),
        at_link=np.array([a[0] for a in attrs]),
        at_vals=np.array([[0.0 if v is None else v for v in a[1:]] for a in attrs]),
        at_null=np.array([[v is None for v in a[1:]] for a in attrs]),
        topo_bin=topo_bin, params_bin=params_bin, topo_to_wbid=wbid)
    print(f"wrote {OUT}/hydrofabric.npz: {len(rows)} flowpaths, {len(attrs)} attribute rows, "
          f"{len(wbid)} reaches")


if __name__ == "__main__":
    sys.exit(main())
