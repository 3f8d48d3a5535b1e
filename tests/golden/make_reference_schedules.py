"""Writes reference_schedules.json: the per-batch vertex counts of the reference's published
stream runs (results/*.csv with the run arguments in results/*.json, as its bench.py writes
them), for the schedule test.  Only the columns the random draws do not decide are kept:
batch, n_active, inserted, deleted.

    python tests/golden/make_reference_schedules.py [/root/reference]
"""
import csv
import json
import os
import sys

KEEP_ARGS = ("dataset", "n", "classes", "label_frac", "init_frac", "batches", "del_frac", "seed", "backend")


def main(ref):
    out = []
    res = os.path.join(ref, "results")
    for name in sorted(os.listdir(res)):
        if not name.endswith(".csv") or not os.path.exists(os.path.join(res, name[:-4] + ".json")):
            continue
        with open(os.path.join(res, name[:-4] + ".json")) as fh:
            args = json.load(fh)["args"]
        rows = {}
        with open(os.path.join(res, name), newline="") as fh:
            for r in csv.DictReader(fh):
                rows.setdefault((r["dataset"], int(r["batch"])), set()).add(
                    (int(r["n_active"]), int(r["inserted"]), int(r["deleted"])))
        for (ds, _), v in rows.items():
            assert len(v) == 1, (name, ds, v)      # every solver of a batch saw the same system
        datasets = sorted({ds for ds, _ in rows})
        for ds in datasets:
            out.append({"source": f"results/{name}", "dataset_name": ds,
                        "args": {k: args.get(k) for k in KEEP_ARGS},
                        "rows": [[b, *next(iter(rows[(d, b)]))] for d, b in sorted(rows) if d == ds]})
    with open(os.path.join(os.path.dirname(__file__), "reference_schedules.json"), "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
