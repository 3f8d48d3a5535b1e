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
