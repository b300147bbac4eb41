"""Host-side inversion steps (the behaviour of grad.py, linesearch.py and
misfit.py): checked against closed forms and hand-computed cases."""
import numpy as np

from sparc_amd import inversion as INV


def test_filterx_is_a_periodic_gaussian_convolution():
    nx, ny, nz = 64, 2, 3
    k = np.zeros((nx, ny, nz))
    k[10] = 1.0
    INV.filterx(k)
    x = np.fft.fftfreq(nx) * nx
    g = np.exp(-x ** 2 / (2 * 4.0 ** 2))
    want = np.roll(g, 10)
    assert np.allclose(k[:, 0, 0], want, atol=1e-12)


def test_filterz_smooth_keeps_constants_and_kills_nyquist():
    a = np.full((2, 2, 30), 3.0)
    INV.filterz(a, algo="smooth")
    # the coefficients sum to one; planes 4 and nz-1 (0-based) are never
    # written into the temporary and come back as zero, as in grad.py:86-91
    keep = np.ones(30, bool)
    keep[[4, 29]] = False
    assert np.allclose(a[:, :, keep], 3.0)
    assert np.all(a[:, :, ~keep] == 0.0)
    b = np.broadcast_to((-1.0) ** np.arange(30), (2, 2, 30)).copy()
    INV.filterz(b, algo="smooth")
    assert np.allclose(b[:, :, 5:25], 0.0, atol=1e-12)
    assert np.all(b[:, :, 0] == (-1.0) ** 0)  # plane 1 is left alone


def test_antisymmetrize_and_rms():
    a = np.arange(8.0)[:, None, None] * np.ones((8, 1, 2))
    INV.antisymmetrize(a)
    assert np.allclose(a[:, 0, 0], 0.5 * (np.arange(8.0) - np.arange(8.0)[::-1]))
    assert np.allclose(a, -a[::-1])
    assert INV.rms(np.full((2, 2, 2), 3.0)) == 3.0


def test_precondition_normalises_and_floors():
    h1 = np.zeros((2, 2, 4))
    h1[0, 0, 0] = 2.0
    h2 = -np.ones((2, 2, 4)) * 1e-4
    rho = np.array([1.0, 0.5, 0.25, 0.125])
    h = INV.precondition([h1, h2], rho)
    assert h.max() == 1.0
    assert h.min() == 5e-3
    assert np.isclose(h[0, 0, 0], 1.0)


def test_conjugate_gradient_coefficient():
    g1 = np.array([1.0, 2.0, 3.0])
    g0 = np.array([1.0, 1.0, 1.0])
    u0 = np.array([0.5, 0.5, 0.5])
    # (1*0 + 2*1 + 3*2) / 3
    assert INV.cg_coefficient(g1, g0) == 8.0 / 3.0
    assert np.allclose(INV.update_direction(g1, u0, 8.0 / 3.0), g1 + 8.0 / 3.0 * u0)
    # a negative coefficient restarts with steepest descent
    assert INV.cg_coefficient(g0, g1) == 0.0
    assert np.allclose(INV.update_direction(g1), g1)


def test_linesearch_models_and_jobs():
    last = np.full((2, 2, 2), 2.0)
    upd = np.zeros((2, 2, 2))
    upd[0, 0, 0] = -4.0
    c = np.ones((2, 2, 2))
    ms = INV.linesearch_models(last, upd, last_c=c, n=5, eps=1e-7)
    assert len(ms) == 5
    assert ms[2][0] is c
    # rms(last) = 2, update normalised to -1 at (0,0,0)
    assert np.isclose(ms[2][1][0, 0, 0], 2.0 - 3e-7 * 2.0)
    assert np.all(ms[2][1][1:] == 2.0)
    jobs = [INV.linesearch_job(p, 3, 5) for p in range(16)]
    assert jobs[:4] == [(1, 1), (1, 2), (1, 3), (2, 1)]
    assert jobs[14] == (5, 3) and jobs[15] is None
    assert INV.total_misfit([1.0, 2.5]) == 3.5
    assert INV.linesearch_totals(np.arange(15.0), 3) == [3.0, 12.0, 21.0, 30.0, 39.0]


def test_kernel_files_round_trip_through_grad_inputs(tmp_path):
    """Kernel files written as PRODUCE_KERNELS names them are what the
    gradient step reads back (grad.py:158-164)."""
    from types import SimpleNamespace
    from sparc_amd import abi
    rng = np.random.default_rng(2)
    shape = (6, 5, 20)
    setup = SimpleNamespace(rho0=np.ones(shape) * 2.0, c2=np.ones(shape) * 3.0)
    spec = SimpleNamespace(steps=15, receivers=np.zeros((4, 2)),
                           W=SimpleNamespace(params=SimpleNamespace(timestep=2.0)))
    res = []
    for s in range(2):
        raw = {k: rng.standard_normal(shape) for k in (abi.KERN_C2, abi.KERN_RHO, abi.KERN_HESSIAN)}
        r = SimpleNamespace(kernel_raw=raw, misfit=0.5 * (s + 1))
        INV.write_solve_files(str(tmp_path), s + 1, r, setup, spec)
        res.append(r)
    kc, kp, hs = INV.read_gradient_inputs(str(tmp_path), 2, psi=False)
    assert kp is None
    for s in range(2):
        assert np.array_equal(hs[s], res[s].kernel_raw[abi.KERN_HESSIAN])
        assert np.allclose(kc[s] / res[s].kernel_raw[abi.KERN_C2], kc[0] / res[0].kernel_raw[abi.KERN_C2])
    lines = [np.loadtxt(str(tmp_path / ("misfit_%02d_00" % (s + 1))), usecols=[2]) for s in range(2)]
    assert INV.total_misfit(lines) == 1.5
