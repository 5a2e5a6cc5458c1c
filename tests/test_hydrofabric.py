"""Hydrofabric network builder (data_parsers/parse_gpkg.py:43-270): the
library's builder against the reference's own output files for a synthetic
GeoPackage with duplicate, NULL and unparsable ids and NULL, zero and negative
attributes (tests/golden/hydrofabric.npz, written by make_hydrofabric.py), byte
for byte, and against what a river network fixes independently of the order
the builder picks: every edge of the table kept, upstream reaches first, level
= longest upstream path."""
import sqlite3

import numpy as np
import pytest

from conftest import golden, load_trmc

trmc = load_trmc()


def _fixture():
    z = golden("hydrofabric.npz")
    none = lambda a: [None if v == "" else str(v) for v in a]
    vals = np.where(z["at_null"], np.nan, z["at_vals"])
    return z, none(z["fp_id"]), none(z["fp_toid"]), [str(v) for v in z["at_link"]], vals


def _bytes(path):
    with open(path, "rb") as f:
        return f.read()


def _write_gpkg(path, z, ids, toids, links, vals):
    con = sqlite3.connect(path)
    con.execute("CREATE TABLE flowpaths (id TEXT, toid TEXT, hydroseq INTEGER, lengthkm REAL, areasqkm REAL)")
    f = lambda v: None if np.isnan(v) else float(v)
    con.executemany("INSERT INTO flowpaths VALUES (?, ?, ?, ?, ?)",
                    [(i, t, None if h < 0 else int(h), f(ln), f(ar)) for i, t, h, ln, ar in
                     zip(ids, toids, z["fp_hydroseq"], z["fp_lengthkm"], z["fp_areasqkm"])])
    con.execute('CREATE TABLE "flowpath-attributes" (link TEXT, '
                + ", ".join(f'"{c}" REAL' for c in trmc.HYDROFABRIC_ATTRS) + ")")
    con.executemany('INSERT INTO "flowpath-attributes" VALUES (?, ?, ?, ?, ?, ?, ?, ?, ?)',
                    [(l, *[f(v) for v in row]) for l, row in zip(links, vals)])
    con.commit()
    con.close()


def test_builder_writes_the_reference_files_byte_for_byte(tmp_path):
    z, ids, toids, links, vals = _fixture()
    topo, params, wbid = trmc.build_flowpath_network(ids, toids, links, vals)
    trmc.save_topo(str(tmp_path / "topo.bin"), topo)
    trmc.save_params(str(tmp_path / "params.bin"), params)
    assert _bytes(tmp_path / "topo.bin") == z["topo_bin"].tobytes()
    assert _bytes(tmp_path / "params.bin") == z["params_bin"].tobytes()
    assert np.array_equal(wbid, z["topo_to_wbid"])


def test_geopackage_path_writes_the_reference_files(tmp_path):
    """The whole path: sqlite3 tables in, the three files out."""
    z, ids, toids, links, vals = _fixture()
    gpkg = str(tmp_path / "subset.gpkg")
    _write_gpkg(gpkg, z, ids, toids, links, vals)
    out = tmp_path / "net"
    trmc.build_hydrofabric(gpkg, str(out))
    assert _bytes(out / "topo.bin") == z["topo_bin"].tobytes()
    assert _bytes(out / "params.bin") == z["params_bin"].tobytes()
    assert np.array_equal(np.load(out / "topo_to_wbid.npy"), z["topo_to_wbid"])


def test_network_keeps_every_edge_in_a_valid_order():
    """Independent of the sort's tie order: the reach graph is the table's
    graph (first row of each id, nexus m feeding waterbody m), upstreams come
    first, and each level is one more than the deepest upstream's."""
    _, ids, toids, links, vals = _fixture()
    topo, params, wbid = trmc.build_flowpath_network(ids, toids, links, vals)
    first = {}
    for i, t in zip(ids, toids):
        w = trmc.hydrofabric_id(i)
        if w is not None and w not in first:
            first[w] = trmc.hydrofabric_id(t)
    assert sorted(first) == sorted(wbid.tolist())
    want = sorted((u, d) for u, d in first.items() if d is not None and d >= 0 and d in first)
    got = []
    n = topo.n_reaches
    for r in range(n):
        ups = topo.reach_up_idx[topo.reach_up_ptr[r]:topo.reach_up_ptr[r + 1]]
        assert np.all(ups < r) and np.all(np.diff(ups) > 0)
        got += [(int(wbid[u]), int(wbid[r])) for u in ups]
        deepest = max((topo.reach_level[u] for u in ups), default=-1)
        assert topo.reach_level[r] == deepest + 1
    assert sorted(got) == want
    assert np.array_equal(topo.level_reach, np.argsort(topo.reach_level, kind="stable"))
    assert np.array_equal(topo.reach_seg_start, np.arange(n)) and np.all(topo.reach_seg_len == 1)
    # channel parameters stay inside the reference's clamps
    dx, bw, tw, twcc, nch, ncc, cs, s0 = params.T
    assert np.all(dx >= 100) and np.all(tw > bw) and np.all(twcc > tw)
    assert np.all(nch >= 0.01) and np.all(ncc >= 0.02) and np.all(cs >= 0.1) and np.all(s0 >= 1e-5)


def test_attribute_rules_on_single_rows():
    """Hand-worked rows: NULL, zero and negative values keep the defaults; a top
    width not above the bottom width becomes 1.2 x bottom width, a
    compound-channel width not above the top width 3 x top width; length, n,
    nCC, side slope and bed slope are clamped from below."""
    nan = np.nan
    ids = ["wb-1", "wb-2", "wb-3", "wb-4"]
    toids = ["nex-2", "tnx-9", None, "nex-3"]
    links = ["wb-1", "wb-2", "wb-3", "wb-77", "x"]
    vals = np.array([[50.0, 0.001, 0.001, 10.0, 8.0, 5.0, 0.01, 1e-9],
                     [nan, 0.0, -1.0, nan, 30.0, 200.0, 1.5, 0.002],
                     [2500.0, 0.05, 0.1, 12.0, 12.0, nan, nan, nan],
                     [1.0] * 8, [1.0] * 8])
    topo, params, wbid = trmc.build_flowpath_network(ids, toids, links, vals)
    p = {int(w): params[i] for i, w in enumerate(wbid)}
    f = np.float32
    assert list(p[1]) == [100.0, 10.0, f(10.0) * f(1.2), f(f(10.0) * f(1.2)) * f(3.0), 0.01, 0.02, 0.1, f(1e-5)]
    assert list(p[2]) == [1000.0, 20.0, 30.0, 200.0, f(0.06), f(0.12), 1.5, f(0.002)]
    assert list(p[3]) == [2500.0, 12.0, f(12.0) * f(1.2), f(f(12.0) * f(1.2)) * f(3.0), f(0.05), f(0.1), 2.0, f(0.001)]
    assert list(p[4]) == [1000.0, 20.0, 25.0, 60.0, f(0.06), f(0.12), 2.0, f(0.001)]
    # edges wb-1 -> wb-2 and wb-4 -> wb-3: two levels
    assert topo.n_reaches == 4 and topo.n_levels == 2 and sorted(topo.reach_n_up) == [0, 0, 1, 1]


def test_cycle_is_an_error():
    with pytest.raises(trmc.TrmcError) as e:
        trmc.build_flowpath_network(["wb-1", "wb-2", "wb-3"], ["nex-2", "nex-3", "nex-1"], [], np.zeros((0, 8)))
    assert e.value.code == trmc.ERR_TOPOLOGY


def test_ids_follow_the_reference_parse():
    h = trmc.hydrofabric_id
    assert h("wb-12") == 12 and h("tnx-7") == 7 and h("nex--3") == -3 and h("wb-1_000") == 1000
    assert h(None) is None and h("wb12") is None and h("wb-x") is None and h("cat-1-2") is None


def test_missing_geopackage_is_an_error(tmp_path):
    with pytest.raises(trmc.TrmcError):
        trmc.build_hydrofabric(str(tmp_path / "absent.gpkg"))
