"""The reference's own input files, read as a drop-in replacement for its
start-up: the compile-time configuration ``params.i`` (as text), the run
mode from ``Instruction_srcSS_lsLL`` (READ_PARAMETERS, initialize.f90:
964-1087), the source pixels of ``master.pixels`` (driver.f90:608-614), the
ridge-measurement windows ``params.{0,1,2}`` (driver.f90:1424-1437,
1562-1570) and the frequency bands ``filter.params.{0,1}``
(driver.f90:1382-1406).

params.i is Fortran: ``parameter (name = expr, ...)`` and
``type, parameter :: name = expr`` statements whose expressions may use
earlier names.  Default-real literals are single precision in the
reference (no -r8, makefile:27) and are rounded to float32 here the same
way, with single-precision arithmetic between them (so nupeak = 0.0017 is
0.0017000000225380063, as the reference computes with it).
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass

import numpy as np

from . import abi


# --------------------------------------------------------------------------
# params.i
# --------------------------------------------------------------------------
class _Real:
    """A Fortran REAL value with its kind (4 or 8)."""
    __slots__ = ("v", "kind")

    def __init__(self, v: float, kind: int):
        self.v = float(np.float32(v)) if kind == 4 else float(v)
        self.kind = kind


def _num(x):
    return x.v if isinstance(x, _Real) else x


def _arith(op: str, a, b):
    if isinstance(a, bool) or isinstance(b, bool):
        raise ValueError("arithmetic on a logical")
    if isinstance(a, int) and isinstance(b, int):
        if op == "+":
            return a + b
        if op == "-":
            return a - b
        if op == "*":
            return a * b
        if op == "/":
            q = abs(a) // abs(b)  # Fortran integer division truncates toward zero
            return q if (a >= 0) == (b >= 0) else -q
        if op == "**":
            return a ** b
    kind = max(x.kind for x in (a, b) if isinstance(x, _Real))
    x, y = _num(a), _num(b)
    if op == "**" and isinstance(b, int):
        r = x ** b
    else:
        r = {"+": lambda: x + y, "-": lambda: x - y, "*": lambda: x * y, "/": lambda: x / y,
             "**": lambda: x ** y}[op]()
    return _Real(r, kind)


_TOKEN = re.compile(r"""\s*(?:
    (?P<real>(?:\d+\.\d*|\.\d+|\d+)(?:[eEdD][+-]?\d+)?(?:_\d+)?) |
    (?P<logic>\.(?:true|false)\.) |
    (?P<str>'[^']*'|"[^"]*") |
    (?P<name>[a-z_][a-z_0-9]*) |
    (?P<op>\*\*|[-+*/(),])
)""", re.X | re.I)


def _tokens(expr: str) -> list:
    out, pos = [], 0
    while pos < len(expr):
        if expr[pos:].strip() == "":
            break
        m = _TOKEN.match(expr, pos)
        if not m:
            raise ValueError(f"cannot parse {expr!r} at {expr[pos:]!r}")
        pos = m.end()
        kind = m.lastgroup
        tok = m.group(kind)
        if kind == "real":
            t = tok.lower()
            if re.fullmatch(r"\d+", t):
                out.append(("val", int(t)))
            else:
                k = 8 if "d" in t or t.endswith("_8") else 4
                t = re.sub(r"_\d+$", "", t).replace("d", "e")
                out.append(("val", _Real(float(t), k)))
        elif kind == "logic":
            out.append(("val", tok.lower() == ".true."))
        elif kind == "str":
            out.append(("val", tok[1:-1]))
        elif kind == "name":
            out.append(("name", tok.lower()))
        else:
            out.append(("op", tok))
    return out


class _Parser:
    def __init__(self, toks, env):
        self.t, self.i, self.env = toks, 0, env

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else (None, None)

    def take(self):
        tok = self.peek()
        self.i += 1
        return tok

    def expr(self):
        kind, v = self.peek()
        sign = 1
        if kind == "op" and v in "+-":
            self.take()
            sign = -1 if v == "-" else 1
        x = self.term()
        if sign < 0:
            x = _arith("-", 0, x) if not isinstance(x, _Real) else _Real(-x.v, x.kind)
        while self.peek()[0] == "op" and self.peek()[1] in ("+", "-"):
            op = self.take()[1]
            x = _arith(op, x, self.term())
        return x

    def term(self):
        x = self.power()
        while self.peek()[0] == "op" and self.peek()[1] in ("*", "/"):
            op = self.take()[1]
            x = _arith(op, x, self.power())
        return x

    def power(self):
        x = self.atom()
        if self.peek() == ("op", "**"):
            self.take()
            kind, v = self.peek()
            if kind == "op" and v in "+-":  # signed exponent
                self.take()
                y = self.power()
                y = -y if isinstance(y, int) else _Real(-y.v, y.kind)
            else:
                y = self.power()  # right-associative
            x = _arith("**", x, y)
        return x

    def atom(self):
        kind, v = self.take()
        if kind == "val":
            return v
        if kind == "name":
            if v not in self.env:
                raise KeyError(f"params.i: {v} used before it is set")
            return self.env[v]
        if (kind, v) == ("op", "("):
            x = self.expr()
            if self.take() != ("op", ")"):
                raise ValueError("unbalanced parentheses")
            return x
        raise ValueError(f"unexpected token {v!r}")


def _split_top(s: str) -> list[str]:
    """Splits 'a = 1, b = f(2, 3)' at commas outside parentheses and quotes."""
    parts, depth, cur, q = [], 0, "", None
    for ch in s:
        if q:
            cur += ch
            if ch == q:
                q = None
            continue
        if ch in "'\"":
            q = ch
        elif ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        elif ch == "," and depth == 0:
            parts.append(cur)
            cur = ""
            continue
        cur += ch
    parts.append(cur)
    return [p.strip() for p in parts if p.strip()]


def _strip_comment(line: str) -> str:
    q = None
    for i, ch in enumerate(line):
        if q:
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
        elif ch == "!":
            return line[:i]
    return line


def parse_params_i(text: str) -> dict:
    """name (lower case) -> value of every parameter and initialised
    declaration in a params.i text: int, bool, str, or float (REAL values
    carry their single- or double-precision rounding)."""
    env: dict = {}
    real8: set = set()
    logical_stmt = []
    for raw in text.splitlines():
        line = _strip_comment(raw).strip()
        if not line:
            continue
        low = line.lower()
        # declarations that fix a kind: real*8 a, b
        m = re.match(r"real\s*\*\s*8\s+(?!.*::)(.*)$", low)
        if m:
            real8.update(n.strip() for n in m.group(1).split(","))
            continue
        m = re.match(r"parameter\s*\((.*)\)\s*$", line, re.I)
        if m:
            logical_stmt.extend(_split_top(m.group(1)))
        else:
            m = re.match(r"([a-z][^:]*)::(.*)$", line, re.I)
            if not m:
                continue
            decl = m.group(1).lower()
            if "real*8" in decl.replace(" ", "") or "real(8)" in decl.replace(" ", ""):
                for item in _split_top(m.group(2)):
                    real8.add(item.split("=")[0].strip().lower())
            logical_stmt.extend(i for i in _split_top(m.group(2)) if "=" in i)
        while logical_stmt:
            item = logical_stmt.pop(0)
            name, expr = item.split("=", 1)
            name = name.strip().lower()
            val = _Parser(_tokens(expr), env).expr()
            if name in real8 and isinstance(val, (int, _Real)):
                val = _Real(_num(val), 8)
            env[name] = val
    return {k: (v.v if isinstance(v, _Real) else v) for k, v in env.items()}


def option_of(magnetic: bool, flows: bool, use_pml: bool, displ: bool, test_in_2d: bool,
              kernel_mode: bool = True) -> int:
    """PARSE_OPTIONS (initialize.f90:778-824), then READ_PARAMETERS' 2-D
    override (initialize.f90:1079-1082)."""
    option = 0
    if not magnetic and not flows:
        if not test_in_2d:
            option = {(False, False): 1, (True, False): 2, (False, True): 3,
                      (True, True): 4}[(use_pml, displ)]
        else:
            option = 17 if (use_pml and kernel_mode) else (6 if use_pml else 5)
    if magnetic:
        if not test_in_2d:
            option = {(False, False): 7, (True, False): 8, (False, True): 9,
                      (True, True): 10}[(use_pml, displ)]
        else:
            option = 18 if (use_pml and kernel_mode) else (12 if use_pml else 11)
    if flows:
        option = {(False, False): 13, (True, False): 14, (False, True): 15,
                  (True, True): 16}[(use_pml, displ)]
    if test_in_2d:
        option = 18 if magnetic else 17
    return option


@dataclass
class StepCounts:
    """Step counts of READ_PARAMETERS (initialize.f90:977-981)."""
    cadforcing_step: int
    maxtime: int
    forcing_length: int


def step_counts(solartime_h: float, nupeak: float, timestep: float, cadforcing: float) -> StepCounts:
    cad = int(math.floor(cadforcing / timestep))
    maxtime = int(math.floor((solartime_h * 3600.0 + 2.0 / nupeak) / timestep)) + 2
    maxtime = cad * int(math.floor(maxtime / cad))
    return StepCounts(cad, maxtime, int(math.floor(maxtime / cad)) + 1)


def params_from_reference(text: str, test_in_2d: bool = False, precision: int = abi.FP64):
    """The build's Params (driver.Params) from a params.i text.  test_in_2d
    comes from the Instruction file (its second line, initialize.f90:1071)."""
    from .driver import Params
    v = parse_params_i(text)
    option = option_of(v.get("magnetic", False), v.get("flows", False),
                       v.get("use_pml", True), v.get("displ", True), test_in_2d,
                       v.get("kernel_mode", True))
    supported = (abi.OPT_RHOVP, abi.OPT_3D, abi.OPT_MHD8, abi.OPT_MHD, abi.OPT_2D, abi.OPT_MHD2D)
    if option not in supported:
        raise ValueError(f"params.i selects option {option}, not one of {supported}")
    return Params(nx=int(v["nx"]), ny=int(v["ny"]), nz=int(v["nz"]),
                  xlength=float(v["xlength"]), ylength=float(v["ylength"]),
                  timestep=float(v["timestep"]), option=option,
                  npmlbot=int(v.get("npmlbot", 10)), npmltop=int(v.get("npmltop", 0)),
                  damp_waves=bool(v.get("damp_waves", True)),
                  outputcad=float(v.get("outputcad", 30.0)),
                  cadforcing=float(v.get("cadforcing", 30.0)),
                  excitdep=float(v.get("excitdep", -1.5e7)),
                  obsheight=float(v.get("obsheight", 0.0)), precision=precision)


# --------------------------------------------------------------------------
# run mode, source pixels, measurement parameters
# --------------------------------------------------------------------------
@dataclass
class RunMode:
    """What READ_PARAMETERS decides from the Instruction file."""
    kind: str               # "forward" (spectral), "adjoint" or "kernels" (no file)
    test_in_2d: bool
    initcond: bool | None = None  # third line of a spectral instruction

    @property
    def compute_forward(self) -> bool:
        return self.kind == "forward"

    @property
    def compute_adjoint(self) -> bool:
        return self.kind == "adjoint"


def instruction_name(src: int | str, ls: int | str) -> str:
    """Instruction_srcSS_lsLL with the two-digit argv strings of the driver
    scripts (data_forward.py:48-52, linesearch.py:46-52)."""
    s = src if isinstance(src, str) else f"{src:02d}"
    l = ls if isinstance(ls, str) else f"{ls:02d}"
    return f"Instruction_src{s}_ls{l}"


def read_instruction(path: str) -> RunMode:
    """READ_PARAMETERS (initialize.f90:1006-1063): an absent file means
    kernel construction; line 1 is 'spectral' or 'adjoint', line 2 '2D' for
    the 2-D options, and a spectral run reads an initial-condition logical
    on line 3."""
    if not os.path.exists(path):
        return RunMode("kernels", False)
    with open(path) as f:
        lines = [ln.strip() for ln in f.read().splitlines()]
    calc = lines[0].strip().lower() if lines else ""
    if calc not in ("spectral", "adjoint"):
        raise ValueError(f"{path}: calculation type {lines[0] if lines else ''!r} "
                         "is neither 'spectral' nor 'adjoint'")
    two_d = len(lines) > 1 and lines[1].split()[:1] == ["2D"]
    init = None
    if calc == "spectral" and len(lines) > 2 and lines[2]:
        tok = lines[2].split()[0].upper().strip(".")
        init = tok in ("T", "TRUE")
    return RunMode("forward" if calc == "spectral" else "adjoint", two_d, init)


def _numbers(path: str) -> list[float]:
    with open(path) as f:
        vals = []
        for ln in f:
            ln = ln.strip()
            if ln:
                vals.append(float(ln.split()[0].replace("d", "e").replace("D", "e")))
    return vals


def read_master_pixels(path: str) -> np.ndarray:
    """Source x positions in Mm, one per line (driver.f90:608-614); their
    count is nmasters (initialize.f90:990-995, at most 100 lines read)."""
    return np.array(_numbers(path)[:100])


def master_pixel(path: str, index: int) -> float:
    """x (Mm) of source `index` (1-based, the integer value of argv 1)."""
    px = read_master_pixels(path)
    if not 1 <= index <= px.size:
        raise IndexError(f"source {index} not in {path} ({px.size} sources)")
    return float(px[index - 1])


@dataclass
class RidgeParams:
    """data/params.N (driver.f90:1424-1437): receivers at source distances in
    (mindist, maxdist) Mm, measured over a window of `window` minutes."""
    mindist: float
    maxdist: float
    window: float


def read_ridge_params(path: str) -> RidgeParams:
    v = _numbers(path)
    return RidgeParams(v[0], v[1], v[2])


def read_filter_params(path: str) -> tuple[float, float]:
    """data/filter.params.N: the frequency band (mHz) of FREQ_FILTER for the
    f mode (N = 0) or the p1 mode (N = 1) (driver.f90:1382-1406)."""
    v = _numbers(path)
    return v[0], v[1]


@dataclass
class DataDir:
    """The reference's data directory (params.i `directory`) as the
    measurement reads it: ridge windows params.{0,1,2} and frequency bands
    filter.params.{0,1} where present, and the source pixels."""
    ridge: dict
    bands: dict
    pixels: np.ndarray


def read_data_dir(directory: str) -> DataDir:
    ridge = {}
    for p in range(3):
        f = os.path.join(directory, f"params.{p}")
        if os.path.exists(f):
            ridge[p] = read_ridge_params(f)
    bands = {}
    for p in range(2):
        f = os.path.join(directory, f"filter.params.{p}")
        if os.path.exists(f):
            bands[p] = read_filter_params(f)
    pix = os.path.join(directory, "master.pixels")
    return DataDir(ridge, bands, read_master_pixels(pix) if os.path.exists(pix) else np.zeros(0))
