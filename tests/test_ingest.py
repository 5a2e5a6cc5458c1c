"""Forcing ingest (data_parsers/parse_qlat_fast_cpp.cpp): per-nexus CSVs
pivoted into forcings.bin.  The library's native ingest is compared byte for
byte with the reference program compiled from its own source into
oracle/_ref, and with a direct restatement of its parsing rules."""
import os
import subprocess

import numpy as np
import pytest

from conftest import load_trmc

trmc = load_trmc()
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "parse_qlat_fast")


def _write_case(d, R=300, T=24, seed=0):
    """R reaches named by nexus ids; files for most of them, rows in the
    reference's "ts, time, value" form, plus the irregular cases its parser
    handles: unknown ids, names without a number, files of other names,
    blank and short lines, timesteps outside [0, T), rows without a value."""
    rng = np.random.default_rng(seed)
    wbid = rng.choice(np.arange(1000, 10_000_000), R, replace=False).astype(np.int64)
    np.save(os.path.join(d, "topo_to_wbid.npy"), wbid)
    want = np.zeros((T, R), np.float32)
    for r in range(R):
        if r % 17 == 5:
            continue                                   # reach without a file
        rows = []
        for t in range(T + 2):                         # two rows beyond T are ignored
            v = float(rng.lognormal(np.log(0.1), 1.0))
            txt = f"{v:.8g}" if r % 3 else f"{v:.17e}"
            rows.append(f"{t}, 2017-01-01 {t:02d}:00:00, {txt}")
            if t < T:
                want[t, r] = np.float32(float(txt))
        if r % 11 == 0:
            rows.insert(3, "")                         # blank line
            rows.insert(5, "x,y")                      # short line: skipped
        if r % 13 == 0:
            rows.append("7, 2017-01-01 07:00:00")      # no value column: atof of the time text after
            want[7, r] = np.float32(2017.0)            # the last comma reads its year
        with open(os.path.join(d, f"nex-{int(wbid[r])}_output.csv"), "w") as f:
            f.write("\n".join(rows) + "\n")
    for name in ("nex-99999999999_output.csv", "nex-abc_output.csv", "nex-12.txt", "other-5_output.csv"):
        with open(os.path.join(d, name), "w") as f:
            f.write("0, t, 1.5\n")
    return wbid, want


def _read_forcings(path, R):
    with open(path, "rb") as f:
        T = int(np.frombuffer(f.read(4), np.int32)[0])
        rest = np.frombuffer(f.read(), np.float32)
    return T, rest[:T * R].reshape(T, R), rest[T * R:]


def test_ingest_pivots_like_the_parsing_rules(tmp_path):
    wbid, want = _write_case(str(tmp_path))
    q, read, skipped = trmc.ingest_nex_csv(str(tmp_path), wbid, 24, n_threads=4)
    assert np.array_equal(q.view(np.uint32), want.view(np.uint32))
    # files of every reach but the 18 without one, plus the unknown id; skipped:
    # the unknown id, and the name without a number
    assert read == 300 - 18 and skipped == 2


def test_ingest_matches_reference_program_byte_for_byte(tmp_path):
    if not os.path.exists(REF):
        pytest.skip("reference build (oracle/_ref) not present")
    _write_case(str(tmp_path), R=500, T=24, seed=1)
    npy = os.path.join(str(tmp_path), "topo_to_wbid.npy")
    ref_out = os.path.join(str(tmp_path), "ref_forcings.bin")
    our_out = os.path.join(str(tmp_path), "our_forcings.bin")
    subprocess.run([REF, str(tmp_path), npy, ref_out, "24", "3"], check=True, capture_output=True, timeout=60)
    trmc.ingest_nex_csv_to_forcings(str(tmp_path), npy, our_out, 24, n_threads=5)
    with open(ref_out, "rb") as a, open(our_out, "rb") as b:
        assert a.read() == b.read()
    T, q, tail = _read_forcings(our_out, 500)
    assert T == 24 and np.all(tail[:1000] == 0) and np.all(tail[1000:] == 0.5)


def test_ingest_result_does_not_depend_on_threads(tmp_path):
    wbid, _ = _write_case(str(tmp_path), R=200, T=12, seed=2)
    a = trmc.ingest_nex_csv(str(tmp_path), wbid, 12, n_threads=1)[0]
    b = trmc.ingest_nex_csv(str(tmp_path), wbid, 12, n_threads=7)[0]
    assert np.array_equal(a, b)


def test_ingest_missing_directory_is_an_error(tmp_path):
    with pytest.raises(trmc.TrmcError):
        trmc.ingest_nex_csv(str(tmp_path / "absent"), np.arange(3), 4)


# ------------------------------------------------------ the Rust twin's rules
def test_rust_rules_match_the_cpp_program_on_well_formed_files(tmp_path):
    """On files every row of which both programs read (values printed from
    floats, names nex-<id>_output.csv), the two rule sets give the same
    forcings.bin bytes, and the C++ rules are the reference program's."""
    d = str(tmp_path)
    rng = np.random.default_rng(4)
    wbid = rng.choice(np.arange(1000, 10_000_000), 120, replace=False).astype(np.int64)
    np.save(os.path.join(d, "topo_to_wbid.npy"), wbid)
    for r in range(120):
        vals = rng.lognormal(np.log(0.1), 1.0, 24).astype(np.float32)
        with open(os.path.join(d, f"nex-{int(wbid[r])}_output.csv"), "w") as f:
            f.write("".join(f"{t}, 2017-01-01 {t:02d}:00:00, {float(v):.9g}\n" for t, v in enumerate(vals)))
    a, b = os.path.join(d, "cpp.bin"), os.path.join(d, "rust.bin")
    trmc.ingest_nex_csv_to_forcings(d, os.path.join(d, "topo_to_wbid.npy"), a, 24)
    trmc.ingest_nex_csv_to_forcings(d, os.path.join(d, "topo_to_wbid.npy"), b, 24, rules=trmc.INGEST_RUST)
    with open(a, "rb") as fa, open(b, "rb") as fb:
        assert fa.read() == fb.read()


def test_rust_rules_row_and_name_cases(tmp_path):
    """main.rs:28-91: the timestep is the leading digits (spaces first, no
    sign); the value after the last comma must read in Rust's float grammar
    (sign, digits with an optional point and exponent, inf / infinity / nan)
    and is correctly rounded to float; a row that does not read is skipped
    (the C++ program's atof would store 0 or a prefix); names must end in
    _output.csv and hold a whole integer."""
    d = str(tmp_path)
    wbid = np.array([11, 22, 33, -4, 55], np.int64)
    rows = {
        11: ["0, t, 1.5", "  1, t, +2.25e1", "2, t, .5", "3, t, 7.", "4, t, abc", "5, t, 0x1p3",
             "6, t, 1e", "7, t,   -inf  ", "8, t, 0.1\r", "-9, t, 3.0", "10 t, 4.0", "11, t"],
        22: ["0, t, 1.0", "1, t, NaN", "2, t, 2.0 extra", "3,t,3.0", "00000000000000000004, t, 4.0"],
    }
    for wid, rs in rows.items():
        with open(os.path.join(d, f"nex-{wid}_output.csv"), "w") as f:
            f.write("\n".join(rs) + "\n")
    for name in ("nex-33_output.csv.bak", "nex-3x_output.csv", "nex--4_output.csv", "nex-+55.csv"):
        with open(os.path.join(d, name), "w") as f:
            f.write("0, t, 9.0\n")
    q, read, skipped = trmc.ingest_nex_csv(d, wbid, 12, rules=trmc.INGEST_RUST)
    f = np.float32
    want11 = {0: 1.5, 1: 22.5, 2: 0.5, 3: 7.0, 7: -np.inf, 8: f(0.1), 10: 4.0}
    for t in range(12):
        if t in want11:
            assert q[t, 0] == f(want11[t]), t
        else:
            assert q[t, 0] == 0.0, t                      # rows 4, 5, 6, 9, 11 skipped
    assert q[0, 1] == 1.0 and np.isnan(q[1, 1]) and q[2, 1] == 0.0 and q[3, 1] == 3.0 and q[4, 1] == 4.0
    assert q[0, 3] == 9.0                                 # nex--4: a whole signed integer
    assert np.all(q[:, 2] == 0) and np.all(q[:, 4] == 0)  # .bak and "nex-+55.csv" are not read
    # files read: 11, 22, -4; skipped: "nex-3x" (no number)
    assert read == 3 and skipped == 1
    # the C++ rules read "4, t, abc" as 0 and "2, t, 2.0 extra" as 2
    qc = trmc.ingest_nex_csv(d, wbid, 12)[0]
    assert qc[4, 0] == 0.0 and qc[2, 1] == 2.0


def test_unknown_rules_are_an_error(tmp_path):
    with pytest.raises(trmc.TrmcError) as e:
        trmc.ingest_nex_csv(str(tmp_path), np.arange(3), 4, rules=7)
    assert e.value.code == trmc.ERR_ARG


def _write_shared(d, T=24):
    """Three reaches; reach 1 is named by "nex-5_output.csv" and by
    "nex-05_output.csv" (the same id, so the same column), and reach 2 by
    "nex-7_output.csv" and "nex-7_output.csv.bak" (the C++ filter accepts any
    name holding "_output.csv")."""
    wbid = np.array([3, 5, 7], np.int64)
    np.save(os.path.join(d, "topo_to_wbid.npy"), wbid)

    def put(name, rows):
        with open(os.path.join(d, name), "w") as f:
            f.write("".join(f"{t}, 2017-01-01, {v}\n" for t, v in rows))

    put("nex-3_output.csv", [(t, 1.0 + t) for t in range(T)])
    put("nex-5_output.csv", [(t, 10.0 + t) for t in range(0, T, 2)])     # even timesteps
    put("nex-05_output.csv", [(t, 20.0 + t) for t in range(1, T, 2)])    # odd timesteps
    put("nex-7_output.csv", [(t, 30.0 + t) for t in range(T)])
    put("nex-7_output.csv.bak", [(t, 40.0 + t) for t in range(0, T, 3)])  # every third, later by name
    return wbid


def test_ingest_applies_every_file_of_a_reach_cell_by_cell(tmp_path):
    """C++ rules: every file writes its own rows (parse_qlat_fast_cpp.cpp:
    164-176), so the rows of two files naming one reach all land; where both
    have a timestep the later file by name wins."""
    T = 24
    wbid = _write_shared(str(tmp_path), T)
    t = np.arange(T, dtype=np.float32)
    want = np.zeros((T, 3), np.float32)
    want[:, 0] = 1.0 + t
    want[:, 1] = np.where(t % 2 == 0, 10.0 + t, 20.0 + t)
    want[:, 2] = np.where(t % 3 == 0, 40.0 + t, 30.0 + t)
    for threads in (1, 2, 5):
        q, read, skipped = trmc.ingest_nex_csv(str(tmp_path), wbid, T, n_threads=threads)
        assert np.array_equal(q, want), threads
        assert read == 5 and skipped == 0


def test_ingest_shared_reach_without_overlap_matches_reference_program(tmp_path):
    """Where the files of one reach hold disjoint timesteps the reference's
    result does not depend on its thread order: byte for byte."""
    if not os.path.exists(REF):
        pytest.skip("reference build (oracle/_ref) not present")
    d = str(tmp_path)
    _write_shared(d, 24)
    os.remove(os.path.join(d, "nex-7_output.csv.bak"))     # the overlapping pair
    npy = os.path.join(d, "topo_to_wbid.npy")
    ref_out, our_out = os.path.join(d, "ref.bin"), os.path.join(d, "our.bin")
    subprocess.run([REF, d, npy, ref_out, "24", "2"], check=True, capture_output=True, timeout=60)
    trmc.ingest_nex_csv_to_forcings(d, npy, our_out, 24, n_threads=3)
    with open(ref_out, "rb") as a, open(our_out, "rb") as b:
        assert a.read() == b.read()


def test_ingest_rust_rules_keep_one_file_per_reach(tmp_path):
    """Rust rules: one file's whole column per reach (the last by name here)."""
    T = 24
    wbid = _write_shared(str(tmp_path), T)
    q, _, _ = trmc.ingest_nex_csv(str(tmp_path), wbid, T, n_threads=2, rules=trmc.INGEST_RUST)
    t = np.arange(T, dtype=np.float32)
    # "nex-05" sorts before "nex-5": the column of nex-5 (even steps only);
    # nex-7_output.csv.bak does not end in _output.csv and is not read
    assert np.array_equal(q[:, 1], np.where(t % 2 == 0, 10.0 + t, 0.0).astype(np.float32))
    assert np.array_equal(q[:, 2], 30.0 + t)
