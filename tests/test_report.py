"""Result rows in the reference's CSV schema (dynlp_amd.report against bench.py:36-38, 58-70,
123-133, 152-165 of the reference)."""
import csv
import json
import math
import os

import numpy as np
import pytest

from dynlp_amd import report

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fields_match_the_published_csvs():
    """tests/golden/reference_results_header.csv is the header line of the reference's published
    results/local_*.csv; those runs predate the run_s column, which its bench.py now writes
    between ref_ms and delta."""
    with open(os.path.join(HERE, "golden", "reference_results_header.csv")) as fh:
        header = next(csv.reader(fh))
    assert [f for f in report.FIELDS if f != "run_s"] == header
    assert report.FIELDS.index("run_s") == report.FIELDS.index("ref_ms") + 1


def test_predict_two_class_and_argmax():
    F = np.array([[0.7], [0.5], [0.2]])
    assert report.predict(F, 2).tolist() == [1, 0, 0]
    G = np.array([[0.1, 0.6, 0.3], [0.5, 0.2, 0.3]])
    assert report.predict(G, 3).tolist() == [1, 0]


def test_metrics():
    F = np.array([[0.1, 0.9], [0.8, 0.2], [0.4, 0.6], [0.3, 0.7]])
    y = np.array([1, 0, 0, -1])
    acc, err, agree, cert = report.metrics(F, y, 3)
    assert acc == pytest.approx(2 / 3) and math.isnan(err) and math.isnan(agree) and math.isnan(cert)
    Fs = F + np.array([[0, 0], [0, 0], [0.3, -0.3], [0, 1e-3]])
    acc, err, agree, cert = report.metrics(F, y, 3, Fstar=Fs, rho_true=1e-6, hmax_ref=50.0)
    assert err == pytest.approx(0.3) and agree == pytest.approx(0.75) and cert == pytest.approx(5e-5)


def test_row_fills_nan_and_rejects_unknown_columns():
    r = report.row(solver="dynlp+inc", ms=1.5)
    assert list(r) == report.FIELDS and r["solver"] == "dynlp+inc" and math.isnan(r["err_max"])
    with pytest.raises(KeyError):
        report.row(latency=3)


def test_write_appends_with_one_header_and_json(tmp_path):
    out = tmp_path / "sub" / "run.csv"
    rows = [report.row(tag="t", solver="dynlp+inc", batch=i, ms=2.0, iters=6, acc_true=0.9) for i in range(3)]
    report.write(rows[:2], str(out), args={"steps": 2}, run_s=1.0)
    report.write(rows[2:], str(out), args={"steps": 1}, run_s=2.0)
    with open(out) as fh:
        got = list(csv.DictReader(fh))
    assert [int(r["batch"]) for r in got] == [0, 1, 2]
    assert list(got[0]) == report.FIELDS and got[0]["err_max"] == "nan"
    rec = json.load(open(tmp_path / "sub" / "run.json"))
    assert rec["args"] == {"steps": 1} and rec["run_s"] == 2.0 and "argv" in rec


def test_summarize():
    rows = [report.row(solver="dynlp", ms=200.0, iters=10, acc_true=0.8),
            report.row(solver="dynlp+inc", ms=20.0, iters=3, acc_true=0.9, cert_post=1e-4),
            report.row(solver="dynlp+inc", ms=30.0, iters=4, acc_true=0.7)]
    text = report.summarize(rows)
    lines = text.splitlines()
    assert lines[0].split()[:2] == ["solver", "time"]
    inc = next(l for l in lines if l.startswith("dynlp+inc")).split()
    assert inc[1] == "0.05" and inc[2] == "4.00" and inc[3] == "7" and inc[5] == "1.00e-04"
    assert float(inc[6]) == pytest.approx(0.8)


def _rows():
    nan = float("nan")
    r = report.row
    return [
        # tag b: dynlp twice as slow as dynlp+inc; a solver outside the method order
        r(tag="b", batch=0, solver="dynlp+inc", ms=100.0, iters=3, sweeps=1.5, err_max=1e-7, cert_post=2e-7,
          acc_true=0.9, path="push", build_ms=40.0),
        r(tag="b", batch=1, solver="dynlp+inc", ms=200.0, iters=4, sweeps=2.5, err_max=3e-7, cert_post=nan,
          acc_true=nan, path="pcg", build_ms=60.0),
        r(tag="b", batch=0, solver="dynlp", ms=400.0, iters=10, sweeps=10, err_max=nan, cert_post=nan,
          acc_true=0.8, path="frontier-jacobi", build_ms=40.0),
        r(tag="b", batch=1, solver="dynlp", ms=200.0, iters=12, sweeps=12, err_max=nan, cert_post=nan,
          acc_true=1.0, path="frontier-jacobi", build_ms=60.0),
        r(tag="b", batch=0, solver="custom", ms=50.0, iters=1, sweeps=1, path="pcg", build_ms=40.0),
        r(tag="b", batch=1, solver="custom", ms=50.0, iters=1, sweeps=1, path="amg", build_ms=60.0),
        # tag a: no dynlp row, so no speedup column
        r(tag="a", batch=0, solver="itlp", ms=1000.0, iters=7, sweeps=7, path="jacobi", build_ms=5.0, run_s=12.0),
    ]


def test_matrix_aggregates_hand_example():
    recs, build = report.matrix(_rows())
    # sorted tags; in a tag the method order first, then other solvers by name
    assert [(m["tag"], m["solver"]) for m in recs] == [("a", "itlp"), ("b", "dynlp"), ("b", "dynlp+inc"),
                                                      ("b", "custom")]
    m = {(x["tag"], x["solver"]): x for x in recs}
    inc, dyn, cus, it = m[("b", "dynlp+inc")], m[("b", "dynlp")], m[("b", "custom")], m[("a", "itlp")]
    assert inc["time_s"] == pytest.approx(0.3) and dyn["time_s"] == pytest.approx(0.6)
    assert inc["x_dynlp"] == pytest.approx(2.0) and dyn["x_dynlp"] == pytest.approx(1.0)
    assert cus["x_dynlp"] == pytest.approx(6.0) and math.isnan(it["x_dynlp"])
    assert inc["iters"] == 7 and inc["sweeps"] == pytest.approx(4.0)
    # NaN cells skipped: worst error over both batches, bound and accuracy from batch 0 only
    assert inc["err"] == pytest.approx(3e-7) and inc["bound"] == pytest.approx(2e-7) and inc["acc"] == pytest.approx(0.9)
    assert math.isnan(dyn["err"]) and dyn["acc"] == pytest.approx(0.9)
    # most frequent path, ties to the first in sorted order
    assert dyn["path"] == "frontier-jacobi" and inc["path"] == "pcg" and cus["path"] == "amg"
    # batch assembly once per (tag, batch)
    assert build == {"b": pytest.approx(0.1), "a": pytest.approx(0.005)}
    assert it["run_s"] == 12.0 and math.isnan(inc["run_s"])


def test_write_matrix_reads_result_csvs_only(tmp_path):
    rows = _rows()
    report.write([r for r in rows if r["tag"] == "a"], str(tmp_path / "one.csv"))
    report.write([r for r in rows if r["tag"] == "b"], str(tmp_path / "two.csv"))
    (tmp_path / "other.csv").write_text("x,y\n1,2\n")          # not a result file: skipped
    md = report.write_matrix(str(tmp_path))
    with open(tmp_path / "matrix.csv", newline="") as fh:
        got = list(csv.DictReader(fh))
    assert [g["solver"] for g in got] == ["itlp", "dynlp", "dynlp+inc", "custom"]
    assert list(got[0]) == report.MATRIX_FIELDS
    assert float(got[2]["x_dynlp"]) == pytest.approx(2.0)
    assert "| dynlp+inc | 0.30 s / 0.900 |" not in md     # tag a has no dynlp+inc cell ...
    assert "| dynlp+inc |  | 0.30 s / 0.900 |" in md       # ... tag b has
    assert "batch assembly 0.10 s" in md
    # a second run reads its own matrix.csv back as no result file
    assert report.write_matrix(str(tmp_path)) == md
