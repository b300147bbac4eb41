"""CPU checks of the computation schedules two HIP kernels are built on,
against independent methods (the GPU parity tests run the kernels
themselves):

* the exact segmented tridiagonal solve of the z derivatives (k_zsolve_b
  and the LDS variant of k_zsolve): local sweeps per segment coupled through
  products of the LU factors must equal the sequential LU solve;
* the length-256 x-line transform of k_xw256: radix-4 stages over four
  registers per lane with register/lane digit swaps by pairwise lane
  exchanges must equal numpy's FFT, leave the spectrum in the documented
  digit-reversed lane layout and invert back to natural order.
"""
import numpy as np
import pytest

# ---------------------------------------------------------------------------
# segmented solve
# ---------------------------------------------------------------------------


def _sequential(r, lw, ib, gm):
    n = len(r)
    u = np.zeros(n)
    p = 0.0
    for k in range(n):
        p = (r[k] - lw[k] * p) * ib[k]
        u[k] = p
    x = np.zeros(n)
    x[n - 1] = u[n - 1]
    for k in range(n - 2, -1, -1):
        x[k] = u[k] - gm[k + 1] * x[k + 1]
    return x


def _coupling(lw, ib, gm, nseg):
    """fcum, gcum as the host builds them for nseg segments of L rows."""
    n = len(lw)
    L = (n + nseg - 1) // nseg
    fc, gc = np.zeros(n), np.zeros(n)
    for a in range(0, n, L):
        b = min(n, a + L)
        p = 1.0
        for k in range(a, b):
            p *= -lw[k] * ib[k]
            fc[k] = p
        p = 1.0
        for k in range(b - 1, a - 1, -1):
            if b == n:
                gc[k] = 0.0
            else:
                p *= -gm[k + 1]
                gc[k] = p
    return fc, gc, L


def _segmented(r, lw, ib, gm, nseg, chain=True):
    """The kernels' phases: local forward sweeps, the chain of segment ends,
    corrected local backward sweeps, the chain of segment starts (chain=False
    leaves the segments uncoupled)."""
    n = len(r)
    fc, gc, L = _coupling(lw, ib, gm, nseg)
    segs = [(a, min(n, a + L)) for a in range(0, n, L)]
    X = np.zeros(n)
    for a, b in segs:  # forward from u = 0 before the segment
        u = 0.0
        for k in range(a, b):
            u = (r[k] - lw[k] * u) * ib[k]
            X[k] = u
    U = []
    for s in range(len(segs)):  # true u on the previous segment's last row
        v = 0.0
        for j in range(s):
            bj = (j + 1) * L - 1
            v = X[bj] + fc[bj] * v
        U.append(v if chain else 0.0)
    for s, (a, b) in enumerate(segs):  # local backward sweep, in place
        x = X[b - 1] + fc[b - 1] * U[s]
        X[b - 1] = x
        for k in range(b - 2, a - 1, -1):
            x = X[k] + fc[k] * U[s] - gm[k + 1] * x
            X[k] = x
    out = np.zeros(n)
    for s, (a, b) in enumerate(segs):  # true x on the next segment's first row
        xn = 0.0
        for j in range(len(segs) - 1, s, -1):
            xn = X[j * L] + gc[j * L] * xn
        out[a:b] = X[a:b] + gc[a:b] * (xn if chain else 0.0)
    return out


@pytest.mark.parametrize("n", [12, 13, 40, 200, 300])
@pytest.mark.parametrize("nseg", [1, 4, 8, 16, 32])
def test_segment_coupling_equals_sequential_lu(n, nseg):
    rng = np.random.default_rng(n * 100 + nseg)
    # LU factors with the contraction of the compact scheme's (|lw ib|, |gm| < 1)
    lw = rng.uniform(0.2, 0.4, n)
    ib = rng.uniform(0.6, 1.2, n)
    gm = rng.uniform(0.2, 0.4, n)
    r = rng.standard_normal(n)
    want = _sequential(r, lw, ib, gm)
    got = _segmented(r, lw, ib, gm, nseg)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


def test_segment_coupling_catches_a_dropped_chain():
    """Uncoupled segments are not the solve: the check above can fail."""
    n, nseg = 300, 16
    rng = np.random.default_rng(7)
    lw, ib, gm = rng.uniform(0.2, 0.4, n), rng.uniform(0.6, 1.2, n), rng.uniform(0.2, 0.4, n)
    r = rng.standard_normal(n)
    want = _sequential(r, lw, ib, gm)
    bad = _segmented(r, lw, ib, gm, nseg, chain=False)
    assert np.abs(bad - want).max() > 1e-3 * np.abs(want).max()


# ---------------------------------------------------------------------------
# length-256 x lines, one wavefront per line
# ---------------------------------------------------------------------------
N = 256
LANE = np.arange(64)
W256 = np.exp(-2j * np.pi * np.arange(N) / N)


def _dft4(v, sign):
    w = np.exp(sign * 2j * np.pi * np.outer(np.arange(4), np.arange(4)) / 4)
    return w @ v


def _swap(v, b):
    """Register digit <-> lane digit at bits log2(b), log2(b)+1: two rounds
    of exchanges with lane ^ 2b and lane ^ b, each lane sending the values
    whose register bit differs from its lane bit."""
    v = v.copy()
    for h in (1, 0):
        m = b << h
        hi = (LANE & m) != 0
        for o in (0, 1):
            i0 = o if h else 2 * o
            i1 = i0 + (1 << h)
            snd = np.where(hi, v[i0], v[i1])
            rcv = snd[LANE ^ m]
            v[i0], v[i1] = np.where(hi, rcv, v[i0]), np.where(hi, v[i1], rcv)
    return v


def _twiddle(v, w, conj):
    v = v.copy()
    for j in range(1, 4):
        v[j] = v[j] * (np.conj(w[j - 1]) if conj else w[j - 1])
    return v


_W1 = [W256[(LANE * k) & 255] for k in (1, 2, 3)]
_W2 = [W256[(4 * (LANE & 15) * k) & 255] for k in (1, 2, 3)]
_W3 = [W256[(16 * (LANE & 3) * k) & 255] for k in (1, 2, 3)]
# lane l holds X[kl + 64 j] after the forward stages
KL = (LANE >> 4) + 4 * ((LANE >> 2) & 3) + 16 * (LANE & 3)


def _fwd_lanes(v):
    """lane l, register j holds x[l + 64 j] -> X[KL[l] + 64 j]"""
    v = _swap(_twiddle(_dft4(v, -1), _W1, False), 16)
    v = _swap(_twiddle(_dft4(v, -1), _W2, False), 4)
    v = _swap(_twiddle(_dft4(v, -1), _W3, False), 1)
    return _dft4(v, -1)


def _inv_lanes(v):
    """X[KL[l] + 64 j] -> unnormalised inverse at l + 64 j"""
    v = _twiddle(_swap(_dft4(v, 1), 1), _W3, True)
    v = _twiddle(_swap(_dft4(v, 1), 4), _W2, True)
    v = _twiddle(_swap(_dft4(v, 1), 16), _W1, True)
    return _dft4(v, 1)


def _forward(x):
    return _fwd_lanes(np.stack([x[LANE + 64 * j] for j in range(4)]))


def _inverse(v):
    v = _inv_lanes(v)
    y = np.zeros(N, complex)
    for j in range(4):
        y[LANE + 64 * j] = v[j]
    return y


def test_lane_shuffle_forward_is_the_fft_in_digit_reversed_lanes():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    v = _forward(x)
    X = np.fft.fft(x)
    for j in range(4):
        np.testing.assert_allclose(v[j], X[KL + 64 * j], rtol=0, atol=1e-12 * np.abs(X).max())
    # the digit reversal is a permutation of the 64 lanes
    assert sorted(KL.tolist()) == list(range(64))


def test_lane_shuffle_derivative_round_trip():
    """Spectral multiplier applied in the digit-reversed layout, inverse back
    to natural order: the unnormalised inverse of (i m) X."""
    rng = np.random.default_rng(4)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    m = rng.standard_normal(N)
    v = _forward(x)
    v = np.stack([1j * m[KL + 64 * j] * v[j] for j in range(4)])
    got = _inverse(v)
    want = np.fft.ifft(1j * m * np.fft.fft(x)) * N
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-11 * np.abs(want).max())


def test_lane_shuffle_damping_layout_round_trip():
    """The damping forward pass stores the spectrum in natural order and the
    inverse pass loads it back into the digit-reversed lanes."""
    rng = np.random.default_rng(5)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    v = _forward(x)
    spec = np.zeros(N, complex)
    for j in range(4):
        spec[KL + 64 * j] = v[j]
    np.testing.assert_allclose(spec, np.fft.fft(x), rtol=0, atol=1e-12 * np.abs(spec).max())
    back = np.stack([spec[KL + 64 * j] for j in range(4)])
    np.testing.assert_allclose(_inverse(back) / N, x, rtol=0, atol=1e-13 * np.abs(x).max())


def _odd_multiplier(rng):
    """An odd multiplier with the Nyquist bin zero, like the derivative's."""
    k = np.fft.fftfreq(N) * N
    m = rng.standard_normal(N // 2 + 1)
    mult = np.sign(k) * m[np.abs(k).astype(int)]
    mult[N // 2] = 0.0
    return mult


def test_single_inverse_for_scaled_rows_and_damping():
    """The RHS_vx pass on interior plane pairs: rows scaled by -rhoinv (one
    constant per row) before the forward transform, (i m) X minus the
    damping spectrum, one inverse transform; each row must be its own
    scaled derivative minus its own damping, as with separate inverses."""
    rng = np.random.default_rng(8)
    p0, p1 = rng.standard_normal(N), rng.standard_normal(N)
    s0, s1 = -1.7, -0.4
    mult = _odd_multiplier(rng)
    d0, d1 = rng.standard_normal(N), rng.standard_normal(N)  # damping rows
    rates = np.abs(rng.standard_normal(N))
    rates = 0.5 * (rates + rates[(-np.arange(N)) % N])  # even in k
    B = rates * np.fft.fft(d0 + 1j * d1)  # the damping spectrum of the pair
    # one inverse, through the lane schedule
    v = _forward(s0 * p0 + 1j * s1 * p1)
    v = np.stack([1j * mult[KL + 64 * j] * v[j] - B[KL + 64 * j] for j in range(4)])
    got = _inverse(v) / N
    # separate real rows
    deriv = lambda f: np.fft.ifft(1j * mult * np.fft.fft(f)).real
    damp = np.fft.ifft(B)
    want0 = s0 * deriv(p0) - damp.real
    want1 = s1 * deriv(p1) - damp.imag
    np.testing.assert_allclose(got.real, want0, rtol=0, atol=1e-12 * np.abs(want0).max())
    np.testing.assert_allclose(got.imag, want1, rtol=0, atol=1e-12 * np.abs(want1).max())
    assert np.abs(np.fft.ifft(1j * mult * np.fft.fft(p0)).imag).max() < 1e-12



# ---------------------------------------------------------------------------
# one 256 x 256 plane of the damping per block (k_damp2d)
# ---------------------------------------------------------------------------


def _chunk_member(c, kx):
    return ((kx <= 64) | (kx >= 192)) if c == 0 else ((kx >= 65) & (kx <= 191))


def _chunk_col(c, kx):
    """column of the LDS tile holding x-bin kx in chunk c"""
    if c == 0:
        return np.where(kx <= 64, kx, kx - 127)
    return kx - 65


def _damp2d(v, rates):
    """The kernel's schedule on a real plane v[y, x]: row pairs q as complex
    rows, forward x transforms in the lane layout, two chunks of x-bins
    closed under kx -> 256 - kx staged in a tile [column, q], the y lines of
    kx = 0..128 unpacked from z_q[kx] and z_q[-kx], transformed, scaled by
    rates[ky, kx] / 2, transformed back and repacked into both bins, then
    inverse x transforms.  Returns the unnormalised IDFT2(rates DFT2(v))."""
    regs = np.stack([_forward(v[2 * q] + 1j * v[2 * q + 1]) for q in range(128)])
    kxs = KL[None, :] + 64 * np.arange(4)[:, None]
    y = LANE[None, :] + 64 * np.arange(4)[:, None]
    qy, even = y >> 1, (y & 1) == 0
    ky = kxs
    for c in (0, 1):
        mem = _chunk_member(c, kxs)
        col_of = _chunk_col(c, kxs)
        tile = np.full((129, 128), np.nan + 0j)
        for q in range(128):
            tile[col_of[mem], q] = regs[q][mem]
        for kx in (range(0, 65) if c == 0 else range(65, 129)):
            mk = (256 - kx) % 256
            ca, cb = int(_chunk_col(c, np.array(kx))), int(_chunk_col(c, np.array(mk)))
            A, B = tile[ca, qy], tile[cb, qy]
            col = np.where(even, (A.real + B.real) + 1j * (A.imag - B.imag),
                           (A.imag + B.imag) - 1j * (A.real - B.real))
            col = _fwd_lanes(col)
            col = col * 0.5 * rates[ky, kx]
            col = _inv_lanes(col)
            P = col[:, LANE ^ 1]
            E, O = np.where(even, col, P), np.where(even, P, col)
            if mk == kx:
                tile[ca, qy[even]] = (E.real + 1j * O.real)[even]
            else:
                tile[ca, qy[even]] = (E + 1j * O)[even]
                tile[cb, qy[~even]] = (np.conj(E) + 1j * np.conj(O))[~even]
        assert not np.isnan(tile[: (129 if c == 0 else 127)]).any()
        for q in range(128):
            regs[q][mem] = tile[col_of[mem], q]
    out = np.zeros((256, 256))
    for q in range(128):
        z = _inv_lanes(regs[q])
        for j in range(4):
            out[2 * q, LANE + 64 * j] = z[j].real
            out[2 * q + 1, LANE + 64 * j] = z[j].imag
    return out


def test_plane_damping_schedule_is_the_2d_transform():
    """The per-plane damping pass equals IDFT2(r DFT2(v)) with rates even in
    ky and indexed by min(kx, 256 - kx), as the three-pass chain computes."""
    rng = np.random.default_rng(11)
    v = rng.standard_normal((256, 256))
    r = rng.uniform(0.1, 1.0, (256, 129))
    r = 0.5 * (r + r[(-np.arange(256)) % 256])  # even in ky
    got = _damp2d(v, r)
    kx = np.arange(256)
    R = r[:, np.minimum(kx, 256 - kx)]
    want = np.fft.ifft2(R * np.fft.fft2(v)) * 65536
    assert np.abs(want.imag).max() < 1e-9 * np.abs(want.real).max()
    np.testing.assert_allclose(got, want.real, rtol=0, atol=1e-11 * np.abs(want.real).max())


def test_plane_damping_chunks_by_register():
    """The kernel tests chunk membership by register index (registers 0 and 3
    and bin 64 in chunk 0, registers 1 and 2 without it in chunk 1): the same
    sets as the bin ranges, each closed under kx -> 256 - kx."""
    kxs = KL[None, :] + 64 * np.arange(4)[:, None]
    j = np.arange(4)[:, None] + 0 * KL[None, :]
    by_reg = [(j == 0) | (j == 3) | (kxs == 64), ((j == 1) | (j == 2)) & (kxs != 64)]
    for c in (0, 1):
        np.testing.assert_array_equal(by_reg[c], _chunk_member(c, kxs))
        ks = set(kxs[by_reg[c]].tolist())
        assert ks == {(256 - k) % 256 for k in ks}
        cols = _chunk_col(c, kxs[by_reg[c]])
        assert sorted(cols.tolist()) == list(range(len(ks)))
