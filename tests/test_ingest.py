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
