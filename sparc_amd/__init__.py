"""MI355X-native forward/adjoint time-stepping engine for the SPARC
linearized Euler helioseismic wave solver (jishnub/sparc).

The hot path (one LDDRK step of the displacement-form RHS, its derivatives,
PML, damping, filters and the in-loop sensitivity-kernel accumulation) runs as
hand-written gfx950 HIP kernels in ``libsparc_gpu.so`` behind the C ABI of
``include/sparc_gpu.h``; this package is the host side: set-up (FP64), the
driver sequence of driver.f90, travel-time measurement and adjoint sources,
an autograd wrapper and a batched serving loop.
"""
from . import abi, background  # noqa: F401

__all__ = ["abi", "background"]
