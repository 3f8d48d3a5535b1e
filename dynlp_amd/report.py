"""Per-batch result rows in the reference's CSV schema, so results of this package line up
column for column with the reference's published CSVs.

The column names and their order are the reference's `FIELDS` (bench.py:36-38); a file is
appended to with a header only when new, and a JSON record of the run (arguments, environment,
argv, wall time) is written next to it (bench.py:123-133).  `metrics` follows bench.py:58-70:
accuracy of the predictions (solvers.py:166-167) on the rows with known ground truth, and,
when an exact solution F* is given, the largest error, the argmax agreement and the
certificate of the true residual.  Quantities a run did not compute are NaN, as in the
reference's CSVs from runs without the reference solve.
"""
from __future__ import annotations

import csv
import json
import math
import os
import sys

import numpy as np

FIELDS = ["tag", "dataset", "K", "batch", "solver", "n_active", "n_u", "nnz", "inserted", "deleted", "iters",
          "edges", "sweeps", "ms", "setup_ms", "path", "err_max", "cert_self", "cert_post", "hmax", "hmax_ref",
          "agree", "acc_true", "build_ms", "ref_ms", "run_s", "delta", "tol", "dtype", "group"]

NAN = float("nan")


def predict(F, K):
    """Class of every row: F[:, 0] > 0.5 for K = 2 (one solved column), else the argmax."""
    F = np.asarray(F)
    return (F[:, 0] > 0.5).astype(np.int32) if K == 2 else F.argmax(axis=1).astype(np.int32)


def metrics(F, y, K, Fstar=None, rho_true=None, hmax_ref=None):
    """(acc_true, err_max, agree, cert_post) of the scores F (solved rows, label columns) against
    the ground truth y of the same rows (y < 0: unknown) and, when given, the exact solution
    Fstar of the same rows; cert_post = rho_true (max_c rho_c of the true FP64 residual) times
    hmax_ref (the exact solve's h maximum)."""
    pred, known = predict(F, K), np.asarray(y) >= 0
    acc = float((pred[known] == np.asarray(y)[known]).mean()) if known.any() else NAN
    if Fstar is None or F.shape[0] == 0:
        return acc, NAN, NAN, NAN
    err = float(np.abs(np.asarray(F, dtype=np.float64) - Fstar).max())
    agree = float((pred == predict(Fstar, K)).mean())
    cert = float(rho_true) * float(hmax_ref) if rho_true is not None and hmax_ref is not None else NAN
    return acc, err, agree, cert


def row(**kw):
    """One CSV row: every field of FIELDS, NaN where not given; unknown keys are an error."""
    bad = set(kw) - set(FIELDS)
    if bad:
        raise KeyError(f"not columns of the schema: {sorted(bad)}")
    return {f: kw.get(f, NAN) for f in FIELDS}


def write(rows, out, args=None, env=None, run_s=None):
    """Appends rows to the CSV `out` (header only when the file is new) and writes the run's
    JSON record next to it (same stem, .json)."""
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    new = not os.path.exists(out)
    with open(out, "a", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=FIELDS)
        if new:
            w.writeheader()
        w.writerows(rows)
    with open(os.path.splitext(out)[0] + ".json", "w") as fh:
        json.dump({"args": args or {}, "env": env or {}, "argv": sys.argv, "run_s": run_s}, fh, indent=2)


def summarize(rows, names=None):
    """Per solver over all batches (bench.py:152-165): total time, speedup against "dynlp" when
    present, iterations, worst error and post-hoc bound, mean accuracy; returns the table."""
    names = names or list(dict.fromkeys(r["solver"] for r in rows))
    fin = lambda v: [x for x in v if not (isinstance(x, float) and math.isnan(x))]  # noqa: E731
    tot = {nm: sum(r["ms"] for r in rows if r["solver"] == nm) for nm in names}
    base = tot.get("dynlp")
    hdr = f"{'solver':<11} {'time s':>8} {'x dynlp':>8} {'iters':>8} {'max err':>9} {'max bound':>10} {'acc':>7}"
    out = [hdr, "-" * len(hdr)]
    for nm in names:
        rs = [r for r in rows if r["solver"] == nm]
        err, bnd, acc = (fin([r[k] for r in rs]) for k in ("err_max", "cert_post", "acc_true"))
        x = base / tot[nm] if base and tot[nm] else NAN
        out.append(f"{nm:<11} {tot[nm] / 1e3:8.2f} {x:8.2f} {sum(int(r['iters']) for r in rs):8d} "
                   f"{max(err, default=NAN):9.2e} {max(bnd, default=NAN):10.2e} "
                   f"{sum(acc) / len(acc) if acc else NAN:7.4f}")
    return "\n".join(out)


# Method order of the matrix (the reference's scripts/matrix.py ORDER); other solvers follow.
ORDER = ["itlp", "dynlp", "dynlp-fh", "dynlp+pcg", "dynlp+amg", "dynlp+auto", "dynlp+inc"]
MATRIX_FIELDS = ["tag", "solver", "time_s", "iters", "sweeps", "err", "bound", "acc", "run_s", "path", "x_dynlp"]


def _num(v):
    """CSV cell as float (empty or unparsable -> NaN)."""
    try:
        return float(v)
    except (TypeError, ValueError):
        return NAN


def _finite(vals):
    return [v for v in vals if not math.isnan(v)]


def read_rows(paths):
    """Rows of result CSVs (the FIELDS schema; files without a solver column are skipped)."""
    rows = []
    for p in paths:
        with open(p, newline="") as fh:
            rd = csv.DictReader(fh)
            if "solver" in (rd.fieldnames or []):
                rows.extend(rd)
    return rows


def matrix(rows):
    """Method matrix of per-batch result rows (the reference's scripts/matrix.py:20-32): per
    (tag, solver) the total solve time in s, summed iterations and sweeps, the worst error and
    post-hoc bound, the mean accuracy, the longest run, the most frequent path (ties: first in
    sorted order) and the speedup over "dynlp" of the same tag; NaN cells are skipped by every
    aggregate.  Returns (records in MATRIX_FIELDS, tag -> batch assembly s)."""
    groups = {}
    for r in rows:
        groups.setdefault((r["tag"], r["solver"]), []).append(r)
    out = []
    for (tag, solver), rs in groups.items():
        col = lambda k: [_num(r.get(k)) for r in rs]  # noqa: E731
        paths = sorted(str(r.get("path", "")) for r in rs)
        mode = max(paths, key=lambda p: (paths.count(p), -paths.index(p)))
        err, bnd, acc = _finite(col("err_max")), _finite(col("cert_post")), _finite(col("acc_true"))
        out.append({"tag": tag, "solver": solver,
                    "time_s": sum(_finite(col("ms"))) / 1e3,
                    "iters": int(sum(_finite(col("iters")))), "sweeps": sum(_finite(col("sweeps"))),
                    "err": max(err) if err else NAN, "bound": max(bnd) if bnd else NAN,
                    "acc": sum(acc) / len(acc) if acc else NAN,
                    "run_s": max(_finite(col("run_s")), default=NAN), "path": mode})
    base = {m["tag"]: m["time_s"] for m in out if m["solver"] == "dynlp"}
    for m in out:
        b, t = base.get(m["tag"], NAN), m["time_s"]
        m["x_dynlp"] = b / t if t else (math.copysign(math.inf, b) if b and not math.isnan(b) else NAN)
    # tags in sorted order; within a tag the ORDER methods first, then the others by name
    rank = {s: i for i, s in enumerate(ORDER)}
    out.sort(key=lambda m: (m["tag"], rank.get(m["solver"], len(ORDER)), m["solver"]))
    build, seen = {}, set()
    for r in rows:
        if (r["tag"], r["batch"]) not in seen:
            seen.add((r["tag"], r["batch"]))
            v = _num(r.get("build_ms"))
            build[r["tag"]] = build.get(r["tag"], 0.0) + (0.0 if math.isnan(v) else v / 1e3)
    return out, build


def write_matrix(directory):
    """matrix.csv and matrix.md for every result CSV in `directory` (the reference's
    scripts/matrix.py command); returns the markdown."""
    import glob
    paths = [p for p in sorted(glob.glob(os.path.join(directory, "*.csv"))) if not p.endswith("matrix.csv")]
    rows = read_rows(paths)
    recs, build = matrix(rows)
    with open(os.path.join(directory, "matrix.csv"), "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=MATRIX_FIELDS)
        w.writeheader()
        w.writerows(recs)
    fmt = lambda v, f: "–" if isinstance(v, float) and math.isnan(v) else format(v, f)  # noqa: E731
    tags = list(dict.fromkeys(m["tag"] for m in recs))
    rank = {s: i for i, s in enumerate(ORDER)}
    solvers = list(dict.fromkeys(m["solver"] for m in sorted(recs, key=lambda m: rank.get(m["solver"], len(ORDER)))))
    cell = {(m["tag"], m["solver"]): m for m in recs}

    def table(head, body):
        return ["| " + " | ".join(head) + " |", "|" + "---|" * len(head)] + ["| " + " | ".join(r) + " |" for r in body]

    delta, tol = (_num(rows[0].get(k)) if rows else NAN for k in ("delta", "tol"))
    md = [f"# Method matrix: `{directory}`", "",
          f"δ = {fmt(delta, 'g')} (DynLP/ItLP), tol = {fmt(tol, 'g')} (DynLP+). Cell: solve time / accuracy.", ""]
    md += table(["method"] + tags, [[s] + [f"{fmt(c['time_s'], '.2f')} s / {fmt(c['acc'], '.3f')}"
                                           if (c := cell.get((t, s))) else "" for t in tags] for s in solvers])
    for t in tags:
        rs = [cell[(t, s)] for s in solvers if (t, s) in cell]
        md += ["", f"## {t}", f"wall {fmt(rs[0]['run_s'], '.0f')} s per run, batch assembly {build[t]:.2f} s", ""]
        md += table(["method", "time s", "× dynlp", "iters", "sweeps", "max err", "bound", "acc", "path"],
                    [[r["solver"], fmt(r["time_s"], ".2f"), fmt(r["x_dynlp"], ".2f"), f"{r['iters']:.0f}",
                      fmt(r["sweeps"], ".0f"), fmt(r["err"], ".1e"), fmt(r["bound"], ".1e"), fmt(r["acc"], ".4f"),
                      r["path"]] for r in rs])
    text = "\n".join(md) + "\n"
    with open(os.path.join(directory, "matrix.md"), "w", encoding="utf-8") as fh:
        fh.write(text)
    return text


if __name__ == "__main__":
    print(write_matrix(sys.argv[1] if len(sys.argv) > 1 else "results"))
