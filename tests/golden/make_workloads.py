"""Writes workloads.json: the parameter draws of the benchmark configurations,
generated once with numpy's default_rng so that every machine uses the same
values.

  oscillatory_16d  config 3: u ~ U(0,1), a_i ~ U(0,5), i < 16, seed 1
  mixture_6d       config 4: theta = [w (4) ~ U(0.5,1.5), s (4) ~ U(0.03,0.15),
                   mu (4x6) ~ U(0.2,0.8)], seed 2
  batch_512        config 5: 512 integrals, d ~ U{2..12}, family uniform in
                   {gaussian, oscillatory, corner}, Gaussian sigma ~ U(0.02,0.2)
                   and centre ~ U(0.3,0.7)^d, oscillatory u ~ U(0,1) and
                   a_i ~ U(0,5), corner a_i ~ U(0.5,3), neval log-uniform in
                   [1e4, 1e7], seed 3
"""
import json
import math
import os

import numpy as np


def main():
    out = {}
    rng = np.random.default_rng(1)
    u = float(rng.uniform(0, 1))
    a = rng.uniform(0, 5, 16).tolist()
    out["oscillatory_16d"] = {"u": u, "a": a}

    rng = np.random.default_rng(2)
    w = rng.uniform(0.5, 1.5, 4)
    s = rng.uniform(0.03, 0.15, 4)
    mu = rng.uniform(0.2, 0.8, (4, 6))
    out["mixture_6d"] = {"theta": [*w.tolist(), *s.tolist(), *mu.ravel().tolist()],
                         "layout": "w[4], s[4], mu[4][6]"}

    rng = np.random.default_rng(3)
    items = []
    for i in range(512):
        d = int(rng.integers(2, 13))
        fam = ["gaussian", "oscillatory", "corner"][int(rng.integers(0, 3))]
        neval = int(round(math.exp(rng.uniform(math.log(1e4), math.log(1e7)))))
        if fam == "gaussian":
            p = {"sigma": float(rng.uniform(0.02, 0.2)), "centre": rng.uniform(0.3, 0.7, d).tolist()}
        elif fam == "oscillatory":
            p = {"u": float(rng.uniform(0, 1)), "a": rng.uniform(0, 5, d).tolist()}
        else:
            p = {"a": rng.uniform(0.5, 3.0, d).tolist()}
        items.append({"d": d, "family": fam, "neval": neval, **p})
    out["batch_512"] = items
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "workloads.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
