"""CPU pinning of the oracle's channeliser: FIR coefficients (closed forms),
FIR ring order (impulse response), DFT (numpy FFT), tone placement
(fftshift), Stokes-I and integration."""
import numpy as np
import pytest


def dirichlet(m, n_ones, big_m):
    """1 + 2 sum_{k=1}^{T-1} cos(2 pi m k / M) in closed form."""
    m = np.asarray(m, np.float64)
    out = np.empty_like(m)
    z = (np.mod(m, big_m) == 0)
    out[z] = 2 * n_ones - 1
    a = np.pi * m[~z] / big_m
    out[~z] = np.sin((2 * n_ones - 1) * a) / np.sin(a)
    return out


@pytest.mark.parametrize("n_taps,n_ch", [(8, 16), (8, 64), (4, 32), (8, 256)])
def test_fir2_dirichlet_closed_form(orc, n_taps, n_ch):
    r = orc.fir2_unwindowed(n_taps, n_ch)
    nt = n_taps * n_ch
    big_m = 4 * nt                      # 4 * grid_n, grid_n = nextPow2(nT)
    m = 2 * np.arange(nt) - (nt - 1)   # odd oversampled positions
    exp = dirichlet(m, n_taps, big_m) / dirichlet([1], n_taps, big_m)[0]
    assert np.allclose(r, exp, rtol=1e-12, atol=1e-13)
    assert np.allclose(r, r[::-1], rtol=1e-12, atol=1e-14)      # symmetric
    assert r[nt // 2 - 1] == 1.0                                # normalised by result[n/2]


@pytest.mark.parametrize("window", ["kaiser", "hamming", "hann", "gaussian"])
def test_coefficients_window_and_sign(orc, window):
    n_taps, n_ch = 8, 16
    nt = n_taps * n_ch
    h = orc.ppf_coefficients(n_taps, n_ch, window)
    r = orc.fir2_unwindowed(n_taps, n_ch)
    i = np.arange(nt)
    if window == "kaiser":
        w = np.kaiser(nt, 9.0695)       # exact I0; the reference's I0 is a 1e-7 polynomial
        rtol = 2e-6
    elif window == "hamming":
        w = np.hamming(nt)
        rtol = 1e-12
    elif window == "hann":
        w = np.hanning(nt)
        rtol = 1e-12
    else:
        w = np.exp(-0.5 * (3.5 / nt * (2 * i - (nt - 1))) ** 2)
        rtol = 1e-12
    unsigned = r * w / (r[nt // 2 - 1] * w[nt // 2 - 1]) / n_ch
    sign = np.where(i % n_ch % 2 == 0, 1.0, -1.0)
    assert np.allclose(h, sign * unsigned, rtol=rtol, atol=1e-15)


def test_blackman_integer_division_quirk(orc):
    # PolyphaseCoefficients.cpp:160: k = i/m in unsigned arithmetic is 0 or 1,
    # so every window point is 0.42 - 0.5 + 0.08 (a constant ~ -1.4e-17); the
    # centre-tap normalisation cancels it and Blackman acts as a rectangle.
    n_taps, n_ch = 8, 16
    h = orc.ppf_coefficients(n_taps, n_ch, "blackman")
    r = orc.fir2_unwindowed(n_taps, n_ch)
    i = np.arange(n_taps * n_ch)
    sign = np.where(i % n_ch % 2 == 0, 1.0, -1.0)
    assert np.allclose(h, sign * r / n_ch, rtol=1e-12, atol=1e-15)


def test_thread_assignment_known_answers(orc):
    # PPF_ChanneliserTest.cpp:281-355
    cases = [((62, 2, 0), (0, 31)), ((62, 2, 1), (31, 62)), ((1, 2, 0), (0, 1)), ((1, 2, 1), (0, 0)),
             ((3, 2, 0), (0, 2)), ((3, 2, 1), (2, 3)), ((4, 2, 0), (0, 2)), ((4, 2, 1), (2, 4)),
             ((4, 3, 0), (0, 2)), ((4, 3, 1), (2, 3)), ((4, 3, 2), (3, 4))]
    for (nv, nthr, tid), exp in cases:
        assert orc.assign_threads(nv, nthr, tid) == exp


def _numpy_ppf64(x, h, n_ch, n_taps):
    """FP64 PPF with numpy FFT: x complex [nBlk*nCh] of one stream."""
    blk = x.reshape(-1, n_ch)
    hist = np.vstack([np.zeros((n_taps - 1, n_ch), complex), blk])
    hh = h.reshape(n_taps, n_ch)
    out = np.empty_like(blk)
    for b in range(blk.shape[0]):
        f = np.sum(hist[b:b + n_taps] * hh, axis=0)
        out[b] = np.fft.fft(f)
    return out


def test_ppf64_matches_numpy_fft(orc):
    n_sub, n_ch, n_taps, n_blk = 2, 16, 8, 40
    h = orc.ppf_coefficients(n_taps, n_ch, "kaiser")
    rng = np.random.default_rng(5)
    x = rng.integers(-300, 300, size=(n_sub, 2, n_blk * n_ch, 2)).astype(np.int16)
    p = orc.PPF(n_sub, 2, n_ch, n_taps, h)
    out = p.run_f64(x, n_blk)
    for s in range(n_sub):
        for pol in range(2):
            z = x[s, pol, :, 0] + 1j * x[s, pol, :, 1].astype(np.float64)
            exp = _numpy_ppf64(z, h, n_ch, n_taps)
            assert np.allclose(out[:, s, pol, :], exp, rtol=1e-10, atol=1e-9)


def test_fir32_impulse_response_and_history(orc):
    """An impulse at block 3, channel 5 walks through the taps oldest-first."""
    n_ch, n_taps, n_blk = 16, 8, 24
    h = orc.ppf_coefficients(n_taps, n_ch, "hann")
    x = np.zeros((1, 2, n_blk * n_ch, 2), np.int16)
    x[0, 0, 3 * n_ch + 5, 0] = 1000
    p = orc.PPF(1, 2, n_ch, n_taps, h)
    # feed in two calls: the FIR history must carry across the boundary
    a = p.fir_only_f32(x[:, :, :6 * n_ch].copy(), 6)
    b = p.fir_only_f32(x[:, :, 6 * n_ch:].copy(), n_blk - 6)
    f = np.concatenate([a, b])[:, 0, 0, 5, 0]
    h32 = h.astype(np.float32).reshape(n_taps, n_ch)[:, 5]
    exp = np.zeros(n_blk, np.float32)
    for k in range(n_taps):
        exp[3 + k] = np.float32(1000.0) * h32[n_taps - 1 - k]   # newest block takes the last tap
    assert np.array_equal(f, exp)


def test_fir32_matches_float32_recurrence(orc):
    n_sub, n_ch, n_taps, n_blk = 2, 16, 8, 30
    h = orc.ppf_coefficients(n_taps, n_ch, "kaiser")
    h32 = h.astype(np.float32).reshape(n_taps, n_ch)
    rng = np.random.default_rng(6)
    x = rng.integers(-2000, 2000, size=(n_sub, 2, n_blk * n_ch, 2)).astype(np.int16)
    p = orc.PPF(n_sub, 2, n_ch, n_taps, h)
    f = p.fir_only_f32(x, n_blk)
    for s in range(n_sub):
        for pol in range(2):
            blk = x[s, pol].reshape(n_blk, n_ch, 2).astype(np.float32)
            hist = np.vstack([np.zeros((n_taps - 1, n_ch, 2), np.float32), blk])
            for b in range(n_blk):
                acc = np.zeros((n_ch, 2), np.float32)
                for t in range(n_taps):
                    acc = acc + hist[b + t] * h32[t][:, None]
                assert np.array_equal(f[b, s, pol], acc)


def test_tone_lands_in_fftshifted_bin(orc):
    # PPF_ChanneliserTest.cpp:195-255 setup: a tone at a quarter of the sample
    # rate with 128 channels peaks in channel 96 after the (-1)^c shift.
    n_ch, n_taps, n_blk = 128, 8, 32
    h = orc.ppf_coefficients(n_taps, n_ch, "kaiser")
    m = np.arange(n_blk * n_ch)
    z = 1000 * np.exp(2j * np.pi * 0.25 * m)
    x = np.zeros((1, 2, n_blk * n_ch, 2), np.int16)
    x[0, :, :, 0] = np.rint(z.real)
    x[0, :, :, 1] = np.rint(z.imag)
    p = orc.PPF(1, 2, n_ch, n_taps, h)
    out = p.run_f64(x, n_blk)
    power = np.abs(out[-1, 0, 0]) ** 2
    assert int(np.argmax(power)) == 96
    assert power[96] > 1e6 * np.median(power)


def test_stokes_and_integration_float32(orc):
    rng = np.random.default_rng(7)
    n_blk, n_sub, n_ch = 12, 3, 16
    spec = rng.normal(size=(n_blk, n_sub, 2, n_ch, 2)).astype(np.float32) * 100
    I = orc.stokes_i_f32(spec)
    X, Y = spec[:, :, 0], spec[:, :, 1]
    exp = (X[..., 0] * X[..., 0] + X[..., 1] * X[..., 1]) + (Y[..., 0] * Y[..., 0] + Y[..., 1] * Y[..., 1])
    assert np.array_equal(I[:, :, 0], exp)
    integ = orc.integrate(I, 4)
    e = np.zeros((n_blk // 4, n_sub, 1, n_ch), np.float32)
    for u in range(n_blk // 4):
        for w in range(4):
            e[u] = e[u] + I[4 * u + w]
    assert np.array_equal(integ, e)
    # channel binning: window-major, then channels of the bin
    integ2 = orc.integrate(I, 2, 4)
    e2 = np.zeros((n_blk // 2, n_sub, 1, n_ch // 4), np.float32)
    for u in range(n_blk // 2):
        for w in range(2):
            for c in range(4):
                e2[u] = e2[u] + I[2 * u + w].reshape(n_sub, 1, n_ch // 4, 4)[..., c]
    assert np.array_equal(integ2, e2)
