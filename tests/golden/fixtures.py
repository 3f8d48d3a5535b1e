"""Loading the golden fixtures (written by make_fixtures.py) into graphs and batches."""
from __future__ import annotations

import os

import numpy as np

from dynlp_amd.graphs import EdgeBatch, Graph

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name):
    with np.load(os.path.join(HERE, name + ".npz")) as f:
        return {k: f[k] for k in f.files}


def graph(fx, name="golden"):
    return Graph(name, int(fx["n"]), fx["u"].astype(np.int64), fx["v"].astype(np.int64), fx["w"].astype(np.float32),
                 fx["labeled"].astype(bool), fx["y"].astype(np.int32), int(fx["K"]))


def batches(fx):
    return [EdgeBatch(fx[f"b{t}_ins_u"], fx[f"b{t}_ins_v"], fx[f"b{t}_ins_w"], fx[f"b{t}_del_u"], fx[f"b{t}_del_v"])
            for t in range(int(fx["n_batches"]))]
