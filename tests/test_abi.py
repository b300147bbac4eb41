"""The C ABI library loads (no GPU needed) and exports every entry point the
header declares; host-only calls behave."""
import ctypes as C
import os
import re

import pytest

from sparc_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "sparc_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sparc_[a-z_0-9]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert header_functions() == sorted(abi.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = abi.lib()
    missing = [s for s in header_functions() if not hasattr(L, s)]
    assert not missing, missing


def test_config_struct_layout():
    # 17 int32 + 7 reserved int32, then 9 doubles + 7 reserved doubles
    assert C.sizeof(abi.SparcConfig) == 24 * 4 + 16 * 8


def _cfg(**kw):
    c = abi.SparcConfig()
    c.abi_version = abi.ABI_VERSION
    c.nx, c.ny, c.nz = 64, 64, 100
    c.option = abi.OPT_3D
    c.precision = abi.FP32
    c.npmlbot = c.npmltop = 10
    c.damp_waves = 1
    c.deltat = 1e-5
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def test_footprint_host_only():
    b32 = abi.footprint(_cfg())
    b64 = abi.footprint(_cfg(precision=abi.FP64))
    assert b64 > b32 > 64 * 64 * 100 * 4 * 12


@pytest.mark.parametrize("bad", [dict(abi_version=99), dict(option=3), dict(nx=100),
                                 dict(nz=8), dict(npmlbot=1), dict(deltat=0.0),
                                 dict(option=abi.OPT_2D)])
def test_footprint_rejects_bad_configs(bad):
    with pytest.raises(abi.SparcError):
        abi.footprint(_cfg(**bad))
