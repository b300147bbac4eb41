"""The reference's on-disk formats for what this path produces, so that its
unchanged Python drivers can read the engine's outputs:

* FITS primary images (writefits_3D, all_modules.f90:296-378): BITPIX -64,
  NAXIS1..3 = (nx, ny, nz) with x fastest, big-endian float64, an integer
  TIME keyword; read back like readfits (all_modules.f90:100-140).
* vz_cc.bin (write_binary_seq_2D, all_modules.f90:947-1006): direct-access
  records of nx*ny float64 values, one record per sample, no record markers.
* kernel/misfit_SS_LL text lines (driver.f90:906-911): ``indexnum nmeas
  misfit`` in list-directed form, the third column being what misfit.py and
  lsmisfit.py sum.
* The kernel files of PRODUCE_KERNELS (kernels.f90:465-539) with their
  scalings, named ``<kind>_SS.fits`` as grad.py expects.

The FITS writer follows the FITS standard (80-character cards in 2880-byte
blocks) rather than cfitsio's exact card set; any FITS reader, cfitsio and
pyfits included, reads what it writes.
"""
from __future__ import annotations

import os

import numpy as np

from .background import DIMB, DIMC, DIML, sp

BLOCK = 2880
CARD = 80
# 10.0**(-8.) and 10.0**(-5.) of kernels.f90:470,521 are default-real literals
MM_PER_CM = sp(1e-8)
KMS_PER_CMS = sp(1e-5)


# --------------------------------------------------------------------------
# FITS
# --------------------------------------------------------------------------
def _card(key: str, value=None, comment: str | None = None) -> bytes:
    if value is None:
        s = key.ljust(CARD)
    else:
        if isinstance(value, bool):
            v = ("T" if value else "F").rjust(20)
        elif isinstance(value, (int, np.integer)):
            v = str(int(value)).rjust(20)
        elif isinstance(value, (float, np.floating)):
            v = ("%.16G" % float(value)).rjust(20)
        else:
            v = "'%s'" % str(value).replace("'", "''").ljust(8)
        s = key.ljust(8)[:8] + "= " + v
        if comment:
            s += " / " + comment
    return s[:CARD].ljust(CARD).encode("ascii")


def fits_write(path: str, arr: np.ndarray, time: int | None = None) -> None:
    """Primary-HDU image of arr, axis 1 = arr's first index (Fortran order)."""
    a = np.asarray(arr, dtype=np.float64)
    cards = [_card("SIMPLE", True, "file does conform to FITS standard"),
             _card("BITPIX", -64, "number of bits per data pixel"),
             _card("NAXIS", a.ndim, "number of data axes")]
    for i, n in enumerate(a.shape):
        cards.append(_card("NAXIS%d" % (i + 1), n, "length of data axis %d" % (i + 1)))
    cards.append(_card("EXTEND", False))
    if time is not None:
        cards.append(_card("TIME", int(time), "Also, init = time+1"))
    cards.append(_card("END"))
    head = b"".join(cards)
    head += b" " * (-len(head) % BLOCK)
    data = a.ravel(order="F").astype(">f8").tobytes()
    data += b"\0" * (-len(data) % BLOCK)
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(head)
        f.write(data)
    os.replace(tmp, path)


def fits_header(path: str) -> dict:
    """Keywords of the primary header (values parsed as bool, int, float or str)."""
    out = {}
    with open(path, "rb") as f:
        while True:
            blk = f.read(BLOCK)
            if len(blk) < BLOCK:
                raise ValueError(f"{path}: truncated FITS header")
            for i in range(0, BLOCK, CARD):
                c = blk[i:i + CARD].decode("ascii")
                key = c[:8].strip()
                if key == "END":
                    out["_data_offset"] = f.tell()
                    return out
                if c[8:10] != "= ":
                    continue
                v = c[10:].strip()
                if v.startswith("'"):
                    val = v[1:v.index("'", 1)].rstrip()
                else:
                    v = v.split("/")[0].strip()
                    if v in ("T", "F"):
                        val = v == "T"
                    else:
                        try:
                            val = int(v)
                        except ValueError:
                            val = float(v.replace("D", "E"))
                out[key] = val


def fits_read(path: str) -> np.ndarray:
    """The primary image as float64, shape (NAXIS1, NAXIS2, ...), Fortran order."""
    h = fits_header(path)
    bitpix = h["BITPIX"]
    shape = tuple(h["NAXIS%d" % (i + 1)] for i in range(h["NAXIS"]))
    dt = {-64: ">f8", -32: ">f4", 8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8"}[bitpix]
    n = int(np.prod(shape))
    with open(path, "rb") as f:
        f.seek(h["_data_offset"])
        a = np.frombuffer(f.read(n * np.dtype(dt).itemsize), dtype=dt).astype(np.float64)
    a = a * h.get("BSCALE", 1.0) + h.get("BZERO", 0.0)
    return a.reshape(shape, order="F")


# --------------------------------------------------------------------------
# vz_cc.bin and misfit lines
# --------------------------------------------------------------------------
def vzcc_write(path: str, samples: np.ndarray, first_record: int = 1) -> None:
    """Records first_record.. of (nx, ny) float64 planes; samples is (nx, ny, nt)."""
    s = np.asarray(samples, dtype=np.float64)
    rec = s.shape[0] * s.shape[1] * 8
    mode = "r+b" if os.path.exists(path) else "wb"
    with open(path, mode) as f:
        for k in range(s.shape[2]):
            f.seek((first_record - 1 + k) * rec)
            f.write(np.ascontiguousarray(s[:, :, k].T).tobytes())


def vzcc_read(path: str, nx: int, ny: int) -> np.ndarray:
    a = np.fromfile(path, dtype=np.float64)
    nt = a.size // (nx * ny)
    return a[: nt * nx * ny].reshape((nt, ny, nx)).transpose(2, 1, 0)


def misfit_line(indexnum: int, nmeas: int, misfit: float) -> str:
    return "%12d%12d %25.16E\n" % (indexnum, nmeas, misfit)


# --------------------------------------------------------------------------
# kernel files
# --------------------------------------------------------------------------
def kernel_files(raw: dict, rho0: np.ndarray, c2: np.ndarray, stepskern: int, timestep: float,
                 magnetic: bool = False) -> dict:
    """The 3-D kernel files of PRODUCE_KERNELS from the accumulated sums
    (keys abi.KERN_*), with the reference's scalings (kernels.f90:465-539):

      hessian          sum |acc|^2                                  (:465)
      kernel_c         -2 rho0 c2 sum * stepskern*timestep * dimen   (:537)
      kernel_rho       -rho0 sum * stepskern*timestep * dimen        (:530)
      kernel_v{x,y,z}  -2 rho0 sum * stepskern*timestep * dimen_v    (:485)
      kernel_b{x,y,z}  sum * stepskern*timestep * dimen_b            (:504)
      kernel_vectorpsi y-curl of the scaled kernel_b                 (:509)

    dimen = (diml 1e-8)^-3, dimen_v = dimen (dimc 1e-5)^-1, dimen_b = dimen
    dimb^-1 * 1000 (with the single-rounded literals of the reference);
    timestep in seconds; rho0, c2 nondimensional.  The engine's
    SPARC_KERN_PSI is the y curl of the unscaled magnetic sums, so it scales
    like kernel_b."""
    from . import abi
    dimen = (DIML * MM_PER_CM) ** -3.0
    cad = stepskern * timestep
    out = {}
    if abi.KERN_HESSIAN in raw:
        out["hessian"] = raw[abi.KERN_HESSIAN]
    if abi.KERN_C2 in raw:
        out["kernel_c"] = -raw[abi.KERN_C2] * rho0 * c2 * cad * dimen * 2.0
    if abi.KERN_RHO in raw:
        out["kernel_rho"] = -rho0 * raw[abi.KERN_RHO] * cad * dimen
    dv = dimen * (DIMC * KMS_PER_CMS) ** -1.0
    for k, name in ((abi.KERN_VX, "kernel_vx"), (abi.KERN_VY, "kernel_vy"),
                    (abi.KERN_VZ, "kernel_vz")):
        if k in raw:
            out[name] = -2.0 * raw[k] * rho0 * cad * dv
    if magnetic:
        db = dimen / DIMB * 1000.0
        for k, name in ((abi.KERN_BX, "kernel_bx"), (abi.KERN_BY, "kernel_by"),
                        (abi.KERN_BZ, "kernel_bz")):
            if k in raw:
                out[name] = raw[k] * cad * db
        if abi.KERN_PSI in raw:
            out["kernel_vectorpsi"] = raw[abi.KERN_PSI] * cad * db
    return out


def write_kernel_files(directory: str, contrib: int, files: dict, misfit: float | None = None,
                       nmeas: int = 0) -> list:
    """<directory>/<name>_SS.fits for each entry of kernel_files(...), and
    <directory>/misfit_SS_00 when a misfit is given.  Returns the paths."""
    os.makedirs(directory, exist_ok=True)
    tag = "%02d" % contrib
    paths = []
    for name, arr in files.items():
        p = os.path.join(directory, f"{name}_{tag}.fits")
        fits_write(p, arr)
        paths.append(p)
    if misfit is not None:
        p = os.path.join(directory, f"misfit_{tag}_00")
        with open(p, "w") as f:
            f.write(misfit_line(contrib, nmeas, misfit))
        paths.append(p)
    return paths
