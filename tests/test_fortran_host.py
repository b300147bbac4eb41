"""The Fortran host (include/sparc_gpu_mod.f90 + sparc_amd/fortran/
timestepping.f90): TIMESTEPPING's forward loop with `call step()` replaced by
sparc_step through ISO_C_BINDING.  On the GPU its v_z records equal the
Python host's through ctypes bit for bit (same engine, same calls); without a
GPU it fails loudly at sparc_create."""
import os
import subprocess

import numpy as np
import pytest

from sparc_amd import abi
from sparc_amd import build
from sparc_amd import driver as D

EXE = build.FORTRAN_EXE


def _bundle(tmp_path, nsteps=31):
    W = D.workload("tiny")
    S = D.setup_for(W)
    cfg = D.make_config(W.params, S, abi.FORWARD, nsteps)
    mask, series = D.forward_source(W, nsteps + 8)
    path = str(tmp_path / "setup.bin")
    D.write_bundle(path, cfg, S, abi.SRC_SERIES, mask, series, 10 ** 9, nsteps - 1, W.params.steps)
    return W, S, cfg, mask, series, path


def _run(path, out):
    return subprocess.run([EXE, path, out], capture_output=True, text=True, timeout=120)


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built (no amdflang)")
def test_fortran_host_reads_bundle_and_fails_loudly_without_gpu(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    *_, path = _bundle(tmp_path)
    r = _run(path, str(tmp_path / "vz.bin"))
    assert r.returncode != 0
    assert "sparc_create failed" in r.stdout + r.stderr


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built (no amdflang)")
def test_fortran_host_matches_python_host(tmp_path):
    W, S, cfg, mask, series, path = _bundle(tmp_path)
    out = str(tmp_path / "vz.bin")
    r = _run(path, out)
    assert r.returncode == 0, r.stdout + r.stderr
    P = W.params
    steps = P.steps
    nrec = (31 - 1) // steps + 1
    got = np.fromfile(out, dtype="<f8").reshape(nrec, P.ny, P.nx)
    eng = abi.Engine(cfg)
    try:
        D.load_setup(eng, S)
        eng.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
        want = []
        for i in range(31):
            eng.step(i, 1)
            if i % steps == 0:
                want.append(eng.plane(abi.V_Z, cfg.o_rad).T)
    finally:
        eng.close()
    want = np.array(want)
    assert np.max(np.abs(want)) > 0
    assert np.array_equal(got, want)
    assert r.stdout.count("norm(v_z)") == nrec
