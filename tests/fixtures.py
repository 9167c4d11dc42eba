"""Load the golden fixtures written by tests/golden/make_golden.py."""
import glob
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def names():
    """The stepping fixtures (the vector-frame fixtures, real_vec_*, hold FRVD payloads)."""
    return sorted(n for n in (os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))
                  if not n.startswith("real_vec_"))


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    prm = json.loads(str(z["params"]))
    mode = 2 if int(prm.get("complex_gauge", 0)) else (1 if int(prm.get("complex_phi", 0)) else 0)
    kw = {}
    for k in ("damp_width", "damp_rate", "kappa_h", "alpha_cs", "beta_h", "kappa_gamma"):
        if k in prm:
            kw[k] = float(prm[k])
    kw["bc_type"] = int(prm.get("bc_type", 2))
    if "gradient_margin" in prm:
        kw["gradient_margin"] = int(prm["gradient_margin"])
    if "mode" in prm:
        kw["rmode"] = int(prm["mode"])
    if "g_gauge" in prm:
        kw["g"] = float(prm["g_gauge"])
    return dict(name=name, mode=mode, N=int(prm["N"]), L=float(prm["L"]),
                dt_factor=float(prm["dt_factor"]), T=float(prm["T"]), diag_dt=float(prm["diag_dt"]),
                kw=kw, frame0=z["frame0"], final=z["final"], diag=str(z["diag"]),
                stdout=str(z["stdout"]), kvmd=json.loads(str(z["kvmd"])), params=prm)


def parse_diag(text):
    lines = [l for l in text.strip().split("\n") if l]
    head = lines[0].split("\t")
    rows = [[float(x) for x in l.split("\t")] for l in lines[1:]]
    return head, rows


def nsteps(fx):
    from oracle import core
    p = core.make_params(fx["N"], fx["mode"], fx["L"], fx["dt_factor"], **fx["kw"])
    return int(round(fx["T"] / p.dt)), p
