"""Pins the clipper oracle (oracle/clipper.cpp) with checks that do not restate it:

* the bandpass known answers of the reference's BandPassTest.cpp:35-101;
* the reference's own bandpass files (tests/golden);
* properties the reference's code fixes (src/lib/src/RFI_Clipper.cpp):
  - spectra are rejected while the mean buffer holds fewer than 1000 entries (:415-420);
  - every polarisation of a rejected spectrum is -0.2 (clipSample, :127-150);
  - in the steady state dataModel is 0 (:305), so polarisation 0 of a rejected spectrum is
    exactly float32(-0.2) - bandpass (:488);
  - kept spectra have unit RMS, and zero mean under zero-DMing (:498-532);
  - BandPass::setMedian's double count of the mean on the primary map (BandPass.cpp:48-61);
  - weights are 1 wherever the clipper writes (:143,332);
  - a bright narrowband channel is replaced (:311-342).
"""
import os

import numpy as np
import pytest


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_bandpass_rebin_known_answers(orc):
    # BandPassTest::test_reBin (BandPassTest.cpp:35-75), without the kill zone
    # (BandPass::killBand is not on the clipper's path).
    start, width, n = 137.304688, -0.0007628, 7936
    params = [4460.84130843, -24.8135957376]
    rms, median = 44.3413175435, 1128.9281113
    m0, r0, a, _ = orc.bandpass_probe(n, start, width, params, rms, median)
    assert m0 == pytest.approx(median, abs=1e-4) and r0 == pytest.approx(rms, abs=1e-5)
    m, r, a2, _ = orc.bandpass_probe(n, start, width, params, rms, median, rebin=(2 * n, start, width / 2.0))
    # the reference asserts 1e-6 absolute; the values are stored as float,
    # whose spacing at 62.7 is 3.8e-6, so the bound is one float step
    assert m == pytest.approx(median / 2.0, abs=1e-6 * median)
    assert r == pytest.approx(rms * np.sqrt(2.0), abs=4e-6)
    assert a2 == pytest.approx(a / 2.0, abs=0.001)


def test_bandpass_set_median_known_answers(orc):
    # BandPassTest::test_setMedian (BandPassTest.cpp:77-101)
    start, width, n = 137.304688, -0.0007628, 7936
    params = [4460.84130843, -24.8135957376]
    rms = 44.3413175435
    median, _, _, _ = orc.bandpass_probe(n, start, width, params, rms)
    half = median / 2.0
    m, _, _, p0 = orc.bandpass_probe(n, start, width, params, rms, half)
    assert m == pytest.approx(half, abs=1e-6 * half)
    m2, _, _, _ = orc.bandpass_probe(n, start, width, [p0, params[1]], rms)
    assert m2 == pytest.approx(m, abs=0.001)
    # a linear bandpass over an even number of bins: the median is the value
    # of bin n/2 of the sorted values, i.e. at frequency start + (n/2)*width
    # for a decreasing function of frequency along increasing bin index
    v = np.float32(start + width * np.arange(n))
    vals = np.float32(params[0]) + np.float32(params[1]) * v.astype(np.float64)
    assert median == pytest.approx(np.sort(vals.astype(np.float32))[n // 2], rel=1e-6)


@pytest.mark.parametrize("name,n_chan", [("t191_BAND.bp", 7936), ("band31.bp", 1984)])
def test_reference_bp_files(orc, name, n_chan):
    """BandPassAdapter::deserialise (BandPassAdapter.cpp:76-100) drops the
    first five numbers, so the polynomial starts with the median value. It
    then moves params[0] so that the median of the bandpass is the file's
    median. The clipper built from the file starts from that bandpass."""
    v = orc.read_bp(os.path.join(GOLDEN, name))
    assert int(v[0]) == n_chan and v.size >= 7
    c = orc.Clipper(v, history=10)
    bp = c.bandpass(n_chan)  # the primary map, before any spectrum
    median = float(np.float32(v[5]))
    assert np.float32(np.sort(bp)[n_chan // 2]) == pytest.approx(median, abs=0.02)
    m, r, b0, _ = orc.bandpass_probe(n_chan, float(np.float32(v[1])), float(np.float32(v[3])),
                                     list(v[5:]), float(v[4]), median)
    assert bp[0] == b0 and m == np.float32(median) and r == np.float32(v[4])
    c.run(np.ones((3, n_chan // 64, 1, 64), np.float32))  # runs on the file's bin count


def _band31_stream(orc, n_t, seed, spikes=(), impulses=()):
    """Spectra following the shape of band31.bp's bandpass plus an offset of
    500 (the file's polynomial, as the reference reads it, spans -328..416
    over the band), Gaussian noise of the file's RMS (40.4), +20000 on the
    spike channels and +50 on every channel of an impulse spectrum."""
    v = orc.read_bp(os.path.join(GOLDEN, "band31.bp"))
    level = orc.Clipper(v).bandpass(1984).astype(np.float64) + 500.0
    rng = np.random.default_rng(seed)
    x = level[None, :] + rng.normal(0.0, v[4], size=(n_t, 1984))
    x[:, list(spikes)] += 20000.0
    x[list(impulses)] += 50.0
    return v, x.astype(np.float32).reshape(n_t, 31, 1, 64)


@pytest.mark.parametrize("zero_dm", [True, False])
def test_clipper_invariants_band31(orc, zero_dm):
    n_t, hist = 2400, 1000
    impulses = list(range(1500, 2400, 100))
    spikes = (17, 900, 1700)
    v, x = _band31_stream(orc, n_t, seed=8, spikes=spikes, impulses=impulses)
    x2 = np.repeat(x, 2, axis=2)  # two polarisations, identical
    c = orc.Clipper(v, 10.0, 4.0, hist, zero_dm)
    W = np.zeros_like(x2)
    out = x2.copy()
    # calls split at the impulses: the bandpass the clipper holds between
    # calls is the one the next call's first spectrum is flattened with
    cuts = [0] + impulses + [n_t]
    good = np.zeros(n_t, np.uint8)
    nclip = np.zeros(n_t, np.int32)
    for a, b in zip(cuts[:-1], cuts[1:]):
        bp = c.bandpass(1984) if a > 0 else None
        seg = np.ascontiguousarray(out[a:b])
        wseg = np.zeros_like(seg)
        c.run(seg, wseg)
        out[a:b], W[a:b] = seg, wseg
        good[a:b], nclip[a:b] = c.last_flags(b - a)
        if bp is not None:
            # a steady-state rejected spectrum is exactly -0.2 - bandpass (dataModel = 0)
            assert good[a] == 0, a
            assert np.array_equal(out[a, :, 0, :], (np.float32(-0.2) - bp).reshape(31, 64)), a
    # the mean buffer already holds the current spectrum when the rule is
    # checked (:376-392 before :415), so spectra 0..998 are rejected by it
    assert not good[:999].any()
    rej = np.nonzero(good == 0)[0]
    assert np.all(out[rej, :, 1, :] == np.float32(-0.2))
    # weights are set to 1 where the clipper writes (a rejected spectrum, a
    # clipped channel) and left as they were elsewhere (the blob's reset sets
    # them all to 1 beforehand, WeightedSpectrumDataSet.cpp:13-78)
    assert np.all(W[rej] == 1.0)
    steady = np.arange(hist, n_t)
    kept = steady[good[steady] == 1]
    assert kept.size > 0.8 * (steady.size - len(impulses))
    # kept spectra: unit RMS before the > crFactor cut, which this noise never
    # reaches.  Here the flattened spectra sit near -8300 (the bandpass mean
    # is double-counted, see test_set_median_double_count) and the RMS is
    # taken as sqrt(sumSq/n - mean^2) in float32, so unit RMS holds to ~10 %;
    # test_clipper_flat_bandpass_normalisation checks it tightly
    y = out[kept, :, 0, :].reshape(kept.size, -1).astype(np.float64)
    assert np.all(y <= 10.0)
    assert np.allclose(y.std(axis=1), 1.0, atol=0.1)
    if zero_dm:
        assert np.allclose(y.mean(axis=1), 0.0, atol=1e-3)
    # the bright channels are clipped in every kept steady-state spectrum
    assert np.all(nclip[kept] >= len(spikes))
    wk = W[kept].reshape(kept.size, 31, 2, 64)
    for b in spikes:
        assert np.all(wk[:, b // 64, :, b % 64] == 1.0)
    assert np.count_nonzero(wk) == 2 * nclip[kept].sum()
    # polarisation 1 of a kept spectrum is only touched by channel clipping
    ok = [b for b in range(1984) if b not in spikes][:50]
    assert np.array_equal(out[kept, :, 1, :].reshape(kept.size, -1)[:, ok],
                          x2[kept, :, 1, :].reshape(kept.size, -1)[:, ok])


def test_set_median_double_count(orc):
    """BandPass::setMedian (BandPass.cpp:48-61) adds delta to _mean[current]
    and delta/scale to _mean[primary]; on the primary map these are the same
    entry, so the bookkept mean moves by 2*delta while the values move by
    delta.  The clipper's later setMean calls (RFI_Clipper.cpp:562) then keep
    the actual bandpass mean offset from the bookkept one by -delta."""
    v = orc.read_bp(os.path.join(GOLDEN, "band31.bp"))
    start, width = float(np.float32(v[1])), float(np.float32(v[3]))
    m_free, _, _, _ = orc.bandpass_probe(1984, start, width, list(v[5:]), float(v[4]))
    delta = float(np.float32(v[5])) - m_free
    c = orc.Clipper(v, history=10)
    c.run(np.full((1, 31, 1, 64), 500.0, np.float32))
    st, _ = c.state()
    actual = c.bandpass(1984).astype(np.float64).mean()
    assert abs(delta) > 1000.0
    assert actual - st[3] == pytest.approx(-delta, rel=1e-3)


@pytest.mark.parametrize("zero_dm", [True, False])
def test_clipper_flat_bandpass_normalisation(orc, zero_dm):
    """With a flat bandpass (.bp polynomial [median, 0]) setMedian does not
    move anything, the flattened spectra have near-zero mean, and a kept
    spectrum comes out with unit RMS (zero mean with zero-DMing) to float32
    accuracy."""
    from synth import bandpass_values
    n_t = 1600
    v = bandpass_values(128, 141.5625, -0.01220703125, 1000.0, 40.0, poly=(0.0,))
    rng = np.random.default_rng(4)
    x = (1000.0 + rng.normal(0.0, 40.0, size=(n_t, 8, 1, 16))).astype(np.float32)
    c = orc.Clipper(v, 10.0, 4.0, 1000, zero_dm)
    c.run(x)
    good, _ = c.last_flags(n_t)
    kept = np.nonzero(good)[0]
    assert kept.size > 500 and kept.min() >= 999
    y = x[kept].reshape(kept.size, -1).astype(np.float64)
    assert np.allclose(y.std(axis=1), 1.0, atol=2e-5)
    if zero_dm:
        assert np.allclose(y.mean(axis=1), 0.0, atol=2e-6)
    else:
        # the running mean is subtracted instead: spectra keep their own offsets
        assert np.std(y.mean(axis=1)) > 0.01


def test_learning_phase_data_model_known_answer(orc):
    """Worked by hand from RFI_Clipper.cpp:214-278, with a flat bandpass at
    1000 and spectra whose channels alternate 1400, 1600 (every band: mean
    1500, sigma exactly 100, minimum 1400).

    Spectrum 0: the band minima minus the model minima are 400, spectrumRMS
    is the smallest band sigma (100) and rmsRunAve = 100, so
    dataModel = 400 + (4/sqrt(8))*100.  Spectrum 1: the model has moved to
    the running mean (1500, :562), so dataModel = -100 + (4/sqrt(8))*100.
    Both are rejected (fewer than 1000 in the mean buffer), so polarisation 0
    ends at -0.2 - (bandpass + dataModel) and polarisation 1 at -0.2."""
    from synth import bandpass_values
    v = bandpass_values(128, 141.5625, -0.01220703125, 1000.0, 40.0, poly=(0.0,))
    c = orc.Clipper(v, history=1000)
    x = np.empty((2, 8, 2, 16), np.float32)
    x[..., 0::2] = 1400.0
    x[..., 1::2] = 1600.0
    c.run(x)
    f = np.float32
    k = f(4.0 / np.sqrt(8.0))
    for t, (bp, dmin) in enumerate([(f(1000.0), f(400.0)), (f(1500.0), f(-100.0))]):
        data_model = dmin + k * f(100.0)
        assert np.all(x[t, :, 0, :] == f(-0.2) - (bp + data_model)), t
    assert np.all(x[:, :, 1, :] == f(-0.2))
    st, sizes = c.state()
    assert st[0] == 1500.0 and st[1] == 100.0 and list(sizes) == [2, 2]


def test_clipper_nbins_not_multiple_of_8(orc):
    c = orc.Clipper(orc.read_bp(os.path.join(GOLDEN, "band31.bp")))
    with pytest.raises(ValueError):
        c.run(np.ones((2, 3, 1, 5), np.float32))


# ---- steady-state known answers (RFI_Clipper.cpp:288-440), worked by hand.
#
# A flat bandpass and spectra whose channels alternate 1400, 1600 (128 bins,
# 8 bands of 16): every band has mean 1500 and sigma exactly 100.  After the
# 1000-spectrum learning phase the state is exactly meanRunAve = 1500,
# rmsRunAve = 100, meanOverRMS = 15 and the bandpass sits at 1500, so in the
# steady state (dataModel = 0, :305):
#  * margin = crFactor * rmsRunAve = 1000: a channel 2000 above its level is
#    clipped, one at +/-100 kept (:311-342);
#  * tol = srFactor * bpRms / sqrt(nBins) = 400 / sqrt(128) = 35.355 (:421-425):
#    a spectrum whose kept mean is 35 above the running mean is kept, 36 above
#    rejected (:415-420);
#  * the running mean adds the newest ring entry before this spectrum's sum is
#    pushed (:394-399), so a kept spectrum moves meanRunAve one spectrum late;
#    a rejected one is popped and replaced by a copy of the previous entry
#    (:432-435) and never moves it;
#  * a band with fewer than 80 % kept channels scores 1, an empty band 4 more,
#    fewer than 80 % kept over the spectrum 4 more, and a score of 4 rejects
#    (:355-372);
#  * a kept spectrum's clipped channels take the last good spectrum's values
#    (:319-335), and kept spectra come out zero-DMed with unit RMS: -1, +1;
#    a rejected one is -0.2 minus the bandpass, -1500.2 (:127-150, 488).
KA_HIST = 1000


def _ka_normal(n):
    x = np.empty((n, 8, 1, 16), np.float32)
    x[..., 0::2] = 1400.0
    x[..., 1::2] = 1600.0
    return x


def _ka_bright(per_band):
    """One spectrum with the first per_band[b] channels of band b 2000 higher."""
    s = _ka_normal(1).reshape(128)
    for b, n in enumerate(per_band):
        s[b * 16:b * 16 + n] += 2000.0
    return s.reshape(1, 8, 1, 16)


def _ka_offset(d):
    return _ka_normal(1) + np.float32(d)


# name -> (case spectrum, kept?)
KA_CASES = {
    "mean +35": (_ka_offset(35.0), True),
    "mean +36": (_ka_offset(36.0), False),
    "25% bright in 3 bands": (_ka_bright([4, 4, 4, 0, 0, 0, 0, 0]), True),
    "25% bright in 4 bands": (_ka_bright([4, 4, 4, 4, 0, 0, 0, 0]), False),
    "one band all bright": (_ka_bright([16, 0, 0, 0, 0, 0, 0, 0]), False),
    "26 bright channels": (_ka_bright([4, 4, 3, 3, 3, 3, 3, 3]), False),
    "25 bright channels": (_ka_bright([4, 3, 3, 3, 3, 3, 3, 3]), True),
}


def ka_stream(case):
    """Learning phase, 20 steady spectra, the case spectrum, 3 more normal ones."""
    return np.concatenate([_ka_normal(KA_HIST + 20), KA_CASES[case][0], _ka_normal(3)])


def _ka_bp():
    from synth import bandpass_values
    return bandpass_values(128, 141.5625, -0.01220703125, 1000.0, 40.0, poly=(0.0,))


def test_steady_state_exact_state(orc):
    c = orc.Clipper(_ka_bp(), 10.0, 4.0, KA_HIST, True)
    x = _ka_normal(KA_HIST + 20)
    c.run(x)
    st, sizes = c.state()
    assert (st[0], st[1], st[4]) == (1500.0, 100.0, 15.0) and list(sizes) == [KA_HIST, KA_HIST]
    assert np.all(c.bandpass(128) == 1500.0)
    good, _ = c.last_flags(x.shape[0])
    # rejected while the mean buffer holds fewer than 1000 entries: the
    # 1000th spectrum is the first kept one
    assert not good[:KA_HIST - 1].any() and good[KA_HIST - 1:].all()
    assert np.all(x[KA_HIST:, :, 0, 0::2] == -1.0) and np.all(x[KA_HIST:, :, 0, 1::2] == 1.0)


@pytest.mark.parametrize("case", sorted(KA_CASES))
def test_steady_state_known_answers(orc, case):
    spec, kept = KA_CASES[case]
    t = KA_HIST + 20
    x = ka_stream(case)
    c = orc.Clipper(_ka_bp(), 10.0, 4.0, KA_HIST, True)
    c.run(x[:t + 1])
    good, nclip = c.last_flags(t + 1)
    assert bool(good[t]) == kept
    n_bright = int(np.count_nonzero(spec.reshape(128) > 2000.0))
    assert nclip[t] == n_bright
    # the running mean is unchanged right after the case spectrum either way:
    # its update reads the previous ring entry
    assert c.state()[0][0] == 1500.0
    rest = x[t + 1:].copy()
    c.run(rest)
    m = c.state()[0][0]
    control = orc.Clipper(_ka_bp(), 10.0, 4.0, KA_HIST, True)
    cx = np.concatenate([_ka_normal(KA_HIST + 20 + 1), _ka_normal(3)])
    control.run(cx)
    if kept:
        kept_mean = spec.reshape(128)[spec.reshape(128) < 2000.0].astype(np.float64).mean()
        assert m == pytest.approx(1500.0 + (kept_mean - 1500.0) / KA_HIST, abs=2e-4)
        # the clipped channels took the last good values: the spectrum is a
        # normal one again
        assert np.all(x[t, :, 0, 0::2] == -1.0) and np.all(x[t, :, 0, 1::2] == 1.0)
    else:
        # rolled back: the state and every later output equal the control's
        assert m == 1500.0 and np.array_equal(c.state()[0], control.state()[0])
        assert np.all(x[t, :, 0, :] == np.float32(-0.2) - np.float32(1500.0))
        assert np.array_equal(rest, cx[t + 1:])


# ---- kill lines (BandPassAdapter.cpp:58-71,101-106; BandPass.cpp:111-157)
def write_bp(path, n_chan, start, width, rms, median, poly, kill_lines=()):
    with open(path, "w") as f:
        f.write("# test bandpass\n%d\n%r\n%r\n%r\n%r\n%r\n" % (n_chan, start, start + width * (n_chan - 1), width,
                                                            rms, median))
        for c in poly:
            f.write("%r\n" % c)
        for k in kill_lines:
            f.write(k + "\n")
    return path


def test_kill_lines_known_answer(orc, tmp_path):
    """A map with frequencies equal to bin numbers (start 0, width 1), so the
    adapter's check binIndex(index) == index holds.  killChannel(k) kills
    [binStart(k), binEnd(k)] = [k - 1/2, k + 1/2], and _zeroChannelsMap zeroes
    binIndex(k - 1/2) = k through binIndex(k + 1/2) = k + 1: two bins per
    kill line.  The second field of a kill line is not read (the reference
    takes the first number for the frequency too) and a repeated index
    counts once."""
    path = write_bp(str(tmp_path / "k.bp"), 64, 0.0, 1.0, 40.0, 1000.0, [0.0, 2.0],
                    ["K 5 5.0", "K 20 123.4", "K 5 7"])
    v, kills = orc.read_bp(path, with_kills=True)
    assert kills == [5, 20]
    with pytest.raises(ValueError):
        orc.read_bp(path)  # refused unless the caller takes the kills
    x = (1000.0 + np.random.default_rng(3).normal(0, 10, size=(4, 8, 1, 8))).astype(np.float32)
    plain = orc.Clipper(v, history=1000)
    plain.run(x.copy())
    killed = orc.Clipper(v, history=1000, kill=kills)
    killed.run(x.copy())
    a, b = plain.bandpass(64), killed.bandpass(64)
    zero = [5, 6, 20, 21]
    assert np.all(b[zero] == 0.0) and np.all(a[zero] != 0.0)
    rest = np.setdiff1d(np.arange(64), zero)
    assert np.array_equal(a[rest], b[rest])


def test_kill_line_errors(orc, tmp_path):
    # the reference's frequency check fails on a real map (frequency 5 is far
    # outside 141.5-142.3 MHz), and the last bin's range ends at nChan
    p = write_bp(str(tmp_path / "a.bp"), 64, 141.5, -0.0122, 40.0, 1000.0, [0.0], ["K 5 141.44"])
    v, kills = orc.read_bp(p, with_kills=True)
    with pytest.raises(ValueError):
        orc.Clipper(v, kill=kills)
    p = write_bp(str(tmp_path / "b.bp"), 64, 0.0, 1.0, 40.0, 1000.0, [0.0], ["K 63 63"])
    v, kills = orc.read_bp(p, with_kills=True)
    with pytest.raises(ValueError):
        orc.Clipper(v, kill=kills)
    for bad in ("K 5", "K x 5", "K -3 3"):
        p = write_bp(str(tmp_path / "c.bp"), 64, 0.0, 1.0, 40.0, 1000.0, [0.0], [bad])
        with pytest.raises(ValueError):
            orc.read_bp(p, with_kills=True)
