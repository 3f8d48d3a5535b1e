"""bench.py end to end on a small SBM (n = 20,000, k = 64): the one JSON line with its roofline
and CPU-baseline objects, and the per-request CSV rows in the reference's schema."""
import csv
import json
import math

import pytest

pytestmark = pytest.mark.gpu


def test_bench_line_and_csv_rows(tmp_path, capsys):
    import bench
    from dynlp_amd import report
    out = tmp_path / "rows.csv"
    rc = bench.main(["--scale", "0.02", "--steps", "3", "--warmup", "1", "--batch-edges", "500",
                     "--fp64-steps", "1", "--cpu-batches", "1", "--out", str(out), "--tag", "small",
                     "--open-loop-rates", "100,400", "--open-loop-seconds", "0.1"])
    assert rc == 0
    line = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert line["metric"].startswith("edge-updates/s") and line["value"] > 0 and line["n_gpus"] == 1
    assert line["steps"] == 3 and line["higher_is_better"] is True
    roof = line["roofline"]
    assert roof["bound"] == "hbm" and roof["unit"] == "GB/s" and roof["peak"] == 8000.0
    assert roof["achieved"] > 0 and roof["frac"] == pytest.approx(roof["achieved"] / roof["peak"], abs=1e-4)
    ver = line["verify_pass"]
    assert ver["ms"] > 0 and 0 < ver["frac"] < 1 and ver["achieved"] == pytest.approx(
        ver["algorithmic_bytes"] / (ver["ms"] * 1e-3) / 1e9, rel=5e-3)   # ms is rounded to 4 decimals
    cpu = line["cpu_baseline"]
    assert cpu["kind"] == "port" and cpu["cores"] == 1 and cpu["value"] > 0
    assert line["fp64_path"]["value"] > 0 and line["rho_max"] <= 1e-6 * 1.0001
    # open loop: every request served and certified; fused repairs never outnumber requests
    pts = line["open_loop"]["points"]
    assert [p["rate_req_s"] for p in pts] == [100.0, 400.0]
    for p in pts:
        assert p["requests"] >= 2 and p["errors"] == 0 and p["repairs"] <= p["requests"]
        assert 0 < p["p50_ms"] <= p["p99_ms"] and p["max_rho"] <= 1e-6 * 1.0001 and p["edge_updates_per_s"] > 0
    assert {k["kernel"].split()[0] for k in line["kernels"]} >= {"spmm_axpy", "frontier_jacobi", "push_scatter",
                                                                   "amg"}
    assert all(k["achieved"] > 0 and 0 < k["frac"] < 1 for k in line["kernels"])
    with open(out) as fh:
        rows = list(csv.DictReader(fh))
    assert len(rows) == 3 and list(rows[0]) == report.FIELDS
    for i, r in enumerate(rows):
        assert r["tag"] == "small" and r["solver"] == "dynlp+inc" and int(r["batch"]) == 2 + i
        assert int(r["inserted"]) + int(r["deleted"]) == 500 and int(r["iters"]) >= 1
        cert, hmax = float(r["cert_self"]), float(r["hmax"])
        assert math.isfinite(cert) and hmax >= 1 and cert <= 1e-6 * hmax * 1.0001
        assert 0.0 <= float(r["acc_true"]) <= 1.0 and math.isnan(float(r["err_max"]))
    assert json.load(open(tmp_path / "rows.json"))["args"]["tag"] == "small"
