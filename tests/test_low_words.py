"""The compensated state's 16-bit low word (sparc_amd/csrc/lowword.h), checked
on the host: the header the engine's update kernel and its state transfers
use is compiled into a small host program (tests/native/lowword_check.hip)
that encodes 2 million FP64 values over the FP32 exponent range and checks
the one-step (2^-39 relative) bound, the exactness of hi + lo in FP64, the
idempotence of a second encoding, and subnormal high words."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not found")
def test_low_word_encoding_bounds(tmp_path):
    exe = str(tmp_path / "lowword_check")
    src = os.path.join(HERE, "native", "lowword_check.hip")
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-o", exe, src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("OK "), out.stdout
