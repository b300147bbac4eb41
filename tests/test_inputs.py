"""Readers of the reference's input files (sparc_amd/inputs.py): params.i
as Fortran text, the Instruction files, master.pixels, the ridge windows and
the frequency bands, on the reference's own input files (tests/golden/
reference_data) and on params.i-format text written here."""
import os

import numpy as np
import pytest

from sparc_amd import abi
from sparc_amd import inputs as I

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_data")
REF_PARAMS = "/root/reference/params.i"

PARAMS_TEXT = """
! a params.i in the reference's own format
integer nx, ny, nz
parameter (nx = 128, ny = 96, nz = 150)   ! grid
real*8 xlength, ylength, timestep
parameter (xlength = 100.0 * 10**(8), ylength = xlength*3/4, timestep = 1.5)
real*8 cadforcing
parameter( cadforcing = 30.0)
logical, parameter :: DAMP_WAVES = .TRUE.
logical, parameter :: magnetic = .false.
logical, parameter :: FLOWS = .FALSE.
logical :: TEST_IN_2D != .false.
real*8 nupeak, nuwidth
integer st_z, fi_z, nz_kern
parameter(st_z = 1, fi_z = nz, nz_kern = (fi_z - st_z+1),nupeak=0.0017, nuwidth = 0.0055d0)
logical, parameter :: DISPL = .truE.
logical USE_PML
parameter(USE_PML = .true.)
integer npmltop, npmlbot,nzpml
parameter(npmlbot = 10, npmltop = 10, nzpml = (npmlbot + npmltop))
real *8 solartime
parameter(solartime = 2.0)
real*8 obsheight
parameter (obsheight = 20000000.0)
real *8 outputcad
parameter (outputcad = 30.0)
real*8 excitdep
parameter( excitdep = - 15000000.0)
character (LEN = *), PARAMETER :: directory = '/data/run/'  ! with a ! in no string
integer, parameter :: halves = 7/2, neg = -7/2, p3 = 2**3**2
"""


def test_params_text_values_and_single_precision_literals():
    v = I.parse_params_i(PARAMS_TEXT)
    assert (v["nx"], v["ny"], v["nz"]) == (128, 96, 150)
    assert v["xlength"] == 1e10 and v["ylength"] == 7.5e9
    assert v["timestep"] == 1.5 and v["nz_kern"] == 150 and v["nzpml"] == 20
    # default reals are single precision (no -r8): 0.0017 -> float32, 0.0055d0 stays
    assert v["nupeak"] == float(np.float32(0.0017)) and v["nupeak"] != 0.0017
    assert v["nuwidth"] == 0.0055
    assert v["damp_waves"] is True and v["magnetic"] is False and "test_in_2d" not in v
    assert v["directory"] == "/data/run/"
    assert v["excitdep"] == -1.5e7
    # Fortran integer arithmetic: truncation toward zero, right-associative **
    assert (v["halves"], v["neg"], v["p3"]) == (3, -3, 512)
    P = I.params_from_reference(PARAMS_TEXT)
    assert P.option == abi.OPT_3D and (P.nx, P.ny, P.nz) == (128, 96, 150)
    assert P.npmlbot == P.npmltop == 10 and P.excitdep == -1.5e7
    assert I.params_from_reference(PARAMS_TEXT, test_in_2d=True).option == abi.OPT_2D


@pytest.mark.parametrize("magnetic,flows,pml,displ,two_d,want", [
    (False, False, True, True, False, 4), (False, False, True, False, False, 2),
    (True, False, True, True, False, 10), (True, False, True, False, False, 8),
    (False, False, True, True, True, 17), (True, False, True, True, True, 18),
    (False, False, False, True, False, 3), (False, True, True, True, False, 16)])
def test_parse_options(magnetic, flows, pml, displ, two_d, want):
    assert I.option_of(magnetic, flows, pml, displ, two_d) == want


def test_step_counts_of_the_reference_run():
    """2 h of solar time at 2 s steps with nupeak = 0.0017 (single-rounded):
    4185 steps and 280 forcing samples at the 30 s cadence."""
    sc = I.step_counts(2.0, float(np.float32(0.0017)), 2.0, 30.0)
    assert (sc.cadforcing_step, sc.maxtime, sc.forcing_length) == (15, 4185, 280)


def test_reference_params_i_when_present():
    if not os.path.exists(REF_PARAMS):
        pytest.skip("reference tree absent")
    v = I.parse_params_i(open(REF_PARAMS).read())
    assert (v["nx"], v["ny"], v["nz"]) == (256, 1, 300)
    assert v["xlength"] == 4e10 and v["timestep"] == 2.0
    assert (v["npmlbot"], v["npmltop"]) == (10, 0)
    assert v["magnetic"] is True and v["displ"] is True and v["use_pml"] is True
    assert v["obsheight"] == 2e7 and v["excitdep"] == -5e7
    P = I.params_from_reference(open(REF_PARAMS).read(), test_in_2d=True)
    assert P.option == abi.OPT_MHD2D


def test_instruction_files(tmp_path):
    sp = I.read_instruction(os.path.join(GOLD, "Spectral"))
    assert sp.kind == "forward" and sp.test_in_2d and sp.initcond is True
    ad = I.read_instruction(os.path.join(GOLD, "Adjoint"))
    assert ad.kind == "adjoint" and ad.test_in_2d and ad.initcond is None
    assert I.read_instruction(str(tmp_path / "absent")).kind == "kernels"
    assert I.instruction_name(1, 0) == "Instruction_src01_ls00"
    bad = tmp_path / "bad"
    bad.write_text("forward\n")
    with pytest.raises(ValueError):
        I.read_instruction(str(bad))


def test_data_directory_files():
    d = I.read_data_dir(GOLD)
    assert d.ridge[0] == I.RidgeParams(10.0, 80.0, 10.0)
    assert d.ridge[1] == I.RidgeParams(20.0, 100.0, 10.0)
    assert 2 not in d.ridge
    assert d.bands == {0: (1.0, 4.0), 1: (2.0, 5.0)}
    assert I.read_filter_params(os.path.join(GOLD, "filter.params")) == (2.0, 3.5)
    assert np.array_equal(d.pixels, [-5.0])
    assert I.master_pixel(os.path.join(GOLD, "master.pixels"), 1) == -5.0
    with pytest.raises(IndexError):
        I.master_pixel(os.path.join(GOLD, "master.pixels"), 2)
