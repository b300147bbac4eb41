"""Shared helpers of the tests: build an oracle and an engine on the same
set-up and compare them."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O
from sparc_amd import abi
from sparc_amd import driver as D


def rel_err(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    den = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b))) / (den if den > 0 else 1.0)


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    den = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (den if den > 0 else 1.0)


def make_oracle(W, S, mode=abi.FORWARD, nsteps=None):
    cfg = D.make_config(W.params, S, mode, nsteps if nsteps is not None else W.nsteps)
    o = O.Oracle(cfg)
    D.load_setup(o, S)
    return o, cfg


def small_workload(name="tiny", precision=abi.FP64, **kw):
    W = D.workload(name, precision=precision)
    for k, v in kw.items():
        setattr(W.params, k, v)
    return W
