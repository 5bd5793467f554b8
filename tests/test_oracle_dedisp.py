"""CPU pinning of the oracle's dedispersion arithmetic (plan integers, delay
table, brute-force sums, FP16/u8 conversions, buffer overlap, detection)."""
from fractions import Fraction

import numpy as np
import pytest

from synth import shift_table

CONFIGS = {
    # name: nChan, nsampBuf, tdms, dmStep, tsamp, fch1, foff -> maxshift, rem, nOut
    "small": (128, 4096, 128, 50 / 128, 5.12e-6 * 16 * 4, 141.5625, -0.01220703125, 697, 39, 3360),
    "lofar": (3968, 65536, 2048, 500 / 2048, 5.12e-6 * 64 * 4, 150.0, -0.1953125 / 64, 12886, 10, 52640),
    "lofar256": (15872, 65536, 2048, 500 / 2048, 5.12e-6 * 256 * 4, 150.0, -0.1953125 / 256, 3222, 42,
                 62272),
}


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_plan_integers(orc, name):
    n, ns, tdms, step, tsamp, fch1, foff, maxshift, rem, nout = CONFIGS[name]
    p = orc.dedisp_plan(n, ns, tdms, 0.0, step, tsamp, fch1, foff)
    assert p["rc"] == 0
    assert (p["maxshift"], p["rem"], p["nout"]) == (maxshift, rem, nout)
    # the defining identities: nOut + maxshift + rem == nsampBuf, nOut % 224 == 0
    assert p["nout"] + p["maxshift"] + p["rem"] == ns
    assert p["nout"] % 224 == 0
    # shift table is the FP64 formula rounded to float32
    ref = shift_table(n, fch1, foff).astype(np.float32)
    assert np.array_equal(p["shift"], ref)


def test_maxshift_error(orc):
    # DedispersionModule.cpp:175-178: maxshift must be below the buffer size
    p = orc.dedisp_plan(3968, 4096, 2048, 0.0, 500 / 2048, 5.12e-6 * 64 * 4, 150.0, -0.1953125 / 64)
    assert p["rc"] == -1


def _exact_delay(shift_f32, dm, mdmstep_f32):
    """trunc(fl32(shift*fl32(dm*mdmstep))) with exact rational rounding to float32."""
    prod = np.float32(np.float32(dm) * mdmstep_f32)  # numpy float32 multiply is IEEE
    v = np.float32(shift_f32 * prod)
    return int(Fraction(float(v)).__trunc__())


def _round_f32(x):
    """Round a Fraction to the nearest float32, ties to even (exactly)."""
    if x == 0:
        return np.float32(0.0)
    sign = -1 if x < 0 else 1
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    q = a / Fraction(2) ** (e - 23)  # in [2^23, 2^24)
    n, r = divmod(q.numerator, q.denominator)
    if 2 * r > q.denominator or (2 * r == q.denominator and n % 2 == 1):
        n += 1
    return np.float32(sign * float(Fraction(n) * Fraction(2) ** (e - 23)))


def test_fused_trial_value_pinned_and_counted(orc):
    """DedispersionKernel.cu:29 forms mstartdm + idx*mdmstep, which nvcc may
    contract into one fused multiply-add.  The oracle's fused table is pinned
    by exact rational rounding of dm*mdmstep + mstartdm; with dmLow == 0 it
    equals the unfused table entry for entry (the sum is exact), and with
    dmLow > 0 the tables differ in a few entries: those delays are parity
    unpinned against the reference (the product reports how many)."""
    n, ns, tdms, step, tsamp, fch1, foff = CONFIGS["lofar"][:7]
    p0 = orc.dedisp_plan(n, ns, tdms, 0.0, step, tsamp, fch1, foff)
    a = orc.dedisp_delays(p0["shift"], tdms, p0["mstartdm"], p0["mdmstep"])
    b = orc.dedisp_delays_fused(p0["shift"], tdms, p0["mstartdm"], p0["mdmstep"])
    assert np.array_equal(a, b)
    p1 = orc.dedisp_plan(n, ns, 1024, 37.3, step, tsamp, fch1, foff)
    ms, st = np.float32(p1["mstartdm"]), np.float32(p1["mdmstep"])
    assert ms > 0
    a = orc.dedisp_delays(p1["shift"], 1024, ms, st)
    b = orc.dedisp_delays_fused(p1["shift"], 1024, ms, st)
    diff = np.argwhere(a != b)
    assert 0 < len(diff) < a.size // 100, len(diff)
    # fused entries against exact rounding: every differing entry and a sample of the rest
    rng = np.random.default_rng(5)
    picks = [tuple(x) for x in diff[:40]] + [(int(rng.integers(1024)), int(rng.integers(n))) for _ in range(40)]
    for dm, c in picks:
        v = _round_f32(Fraction(int(dm)) * Fraction(float(st)) + Fraction(float(ms)))
        assert b[dm, c] == int(Fraction(float(np.float32(p1["shift"][c] * v))).__trunc__())
        u = np.float32(ms + np.float32(np.float32(dm) * st))
        assert a[dm, c] == int(Fraction(float(np.float32(p1["shift"][c] * u))).__trunc__())
    # a differing entry differs by one sample
    assert np.max(np.abs(a - b)) == 1


def test_delay_table_exact(orc):
    n, ns, tdms, step, tsamp, fch1, foff = CONFIGS["small"][:7]
    p = orc.dedisp_plan(n, ns, tdms, 0.0, step, tsamp, fch1, foff)
    d = orc.dedisp_delays(p["shift"], tdms, p["mstartdm"], p["mdmstep"])
    assert p["mstartdm"] == 0.0
    assert p["mdmstep"] == np.float32(np.float32(step) / np.float32(tsamp))
    for dm in (0, 1, 37, 100, tdms - 1):
        for c in (0, 5, 64, n - 1):
            assert d[dm, c] == _exact_delay(p["shift"][c], dm, np.float32(p["mdmstep"]))
    # delays are non-decreasing in DM and in channel (descending frequencies)
    assert np.all(np.diff(d, axis=0) >= 0)
    assert np.all(np.diff(d, axis=1) >= 0)
    assert d[-1, -1] <= p["maxshift"] + p["rem"]


def test_known_answer_dm10(orc):
    """DedispersionModuleTest.cpp:167-168 with DedispersionDataGenerator.cpp:47-81:
    a unit dispersed signal at DM 10 sums to nSub*nCh = 2048 at (t=0, DM index 100)."""
    nsub, nch = 32, 64
    n = nsub * nch
    fch1, foff, tsamp = 150.0, -6.0 / n, 0.00032768
    p = orc.dedisp_plan(n, 32768, 200, 0.0, 0.1, tsamp, fch1, foff)
    assert p["rc"] == 0
    d = orc.dedisp_delays(p["shift"], 200, p["mstartdm"], p["mdmstep"])
    # generator's integer start index, in FP64 (DedispersionDataGenerator.cpp:64-65)
    gen = np.array([int(10.0 * (4148.741601 * ((1.0 / (fch1 + foff * c) / (fch1 + foff * c)) -
                                              (1.0 / fch1 / fch1)) / tsamp)) for c in range(n)])
    inp = np.zeros((n, 32768), np.float32)
    for c in range(n):
        inp[c, gen[c]:gen[c] + 10] = 1.0
    out = orc.dedisperse(inp, d, p["nout"])
    # dmIndex(10.0) = int(0.5 + 10/0.1) (DedispersionSpectra.cpp:46-48, BinMap.cpp:49-52)
    dm_index = int(0.5 + (np.float32(10.0) - np.float32(0.0)) / np.float32(0.1))
    assert dm_index == 100
    assert out[dm_index, 0] == 2048.0
    assert gen.max() == 478


def test_brute_force_matches_direct_sum(orc):
    rng = np.random.default_rng(1)
    n, ns, tdms = 37, 700, 19
    inp = rng.normal(size=(n, ns)).astype(np.float32)
    d = np.sort(rng.integers(0, 200, size=(tdms, n)), axis=0).astype(np.int32)
    nout = 480
    out = orc.dedisperse(inp, d, nout)
    # sequential float32 sum in channel order
    exp = np.zeros((tdms, nout), np.float32)
    for c in range(n):
        for dm in range(tdms):
            exp[dm] = exp[dm] + inp[c, d[dm, c]:d[dm, c] + nout]
    assert np.array_equal(out, exp)
    # a permuted (non-channel-order) sum differs somewhere: the order matters
    exp2 = np.zeros((tdms, nout), np.float32)
    for c in reversed(range(n)):
        for dm in range(tdms):
            exp2[dm] = exp2[dm] + inp[c, d[dm, c]:d[dm, c] + nout]
    assert not np.array_equal(out, exp2)


def test_f16_conversion_matches_numpy(orc):
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.normal(scale=s, size=4000) for s in (1e-6, 1e-3, 1.0, 100.0, 3e4)])
    x = np.concatenate([x, [0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e9, -1e9, 2.0 ** -24, 2.0 ** -25,
                            1.5 * 2.0 ** -25, 6.1e-5, np.inf, -np.inf]]).astype(np.float32)
    bits = orc.f32_to_f16_bits(x)
    assert np.array_equal(bits, x.astype(np.float16).view(np.uint16))


def test_f16_and_u8_dedispersion(orc):
    rng = np.random.default_rng(3)
    n, ns, tdms, nout = 24, 400, 9, 256
    x = rng.normal(size=(n, ns)).astype(np.float32)
    d = np.sort(rng.integers(0, 100, size=(tdms, n)), axis=0).astype(np.int32)
    h = orc.f32_to_f16_bits(x)
    out16 = orc.dedisperse(h, d, nout)
    xf = h.view(np.float16).astype(np.float32)
    exp = np.zeros((tdms, nout), np.float32)
    for c in range(n):
        for dm in range(tdms):
            exp[dm] = exp[dm] + xf[c, d[dm, c]:d[dm, c] + nout]
    assert np.array_equal(out16, exp)
    inv = np.float32(127.0 / 10.0)
    q = orc.quantise_u8(x, inv)
    qe = (np.clip(np.rint(x * inv), -128, 127) + 128).astype(np.uint8)
    assert np.array_equal(q, qe)
    out8 = orc.dedisperse(q, d, nout)
    e8 = np.zeros((tdms, nout), np.int64)
    for c in range(n):
        for dm in range(tdms):
            e8[dm] += q[c, d[dm, c]:d[dm, c] + nout]
    assert np.array_equal(out8, e8)


def test_buffer_overlap_stream(orc):
    n, ns, ov = 5, 64, 20
    buf = orc.DDBuffer(n, ns, ov)
    stream = np.arange(300 * n, dtype=np.float32).reshape(300, n)
    outs = []
    for a in range(0, 300, 37):  # blob sizes that do not divide the buffer
        outs.extend(list(buf.add(stream[a:a + 37])))
    assert len(outs) == 1 + (300 - ns) // (ns - ov)
    for k, o in enumerate(outs):
        first = k * (ns - ov)
        assert np.array_equal(o, stream[first:first + ns].T)


def test_detection_invariants(orc):
    rng = np.random.default_rng(4)
    tdms, nsamp, nch, nsub = 6, 64 * 5 + 17, 16, 8
    rms = np.float32(np.sqrt(np.float32(nch) * np.float32(nsub)))
    data = (rng.normal(size=(tdms, nsamp)) * rms).astype(np.float32)
    data[2, 70:78] += 40 * rms
    data[4, 200] += 9 * rms
    ev, vals = orc.detect(data, nch, nsub, 6.0, 6)
    assert tuple(ev[0]) == (0, 0, 1) and vals[0] == 0.0
    ev, vals = ev[1:], vals[1:]
    thr = np.float32(6.0) * rms
    # level 0: exactly the samples >= threshold (trailing nsamp % 64 ignored)
    lv0 = ev[:, 2] == 1
    exp0 = np.argwhere(data[:, :nsamp - nsamp % 64] >= thr)
    assert np.array_equal(ev[lv0][:, :2], exp0)
    # every event value is its window's dyadic sum, above its level threshold
    for (dm, t, bf), v in zip(ev, vals):
        assert abs(float(v) - float(np.sum(data[dm, t:t + bf], dtype=np.float64))) <= 1e-4 * (
            1 + abs(float(v)))
        assert v >= np.float32(thr * np.float32(np.sqrt(np.float32(bf))))
    # order: DM, 64-sample segment, level, position
    key = np.stack([ev[:, 0], ev[:, 1] // 64, np.log2(ev[:, 2]).astype(int), ev[:, 1]], 1)
    assert all(tuple(key[i]) <= tuple(key[i + 1]) for i in range(len(key) - 1))
    # the strong 8-sample pulse is found at every level up to 8
    found = {(int(dm), int(bf)) for dm, t, bf in ev}
    for bf in (1, 2, 4, 8):
        assert (2, bf) in found


def test_reference_single_event(orc):
    """DedispersionAnalyserTest::test_singleEvent (DedispersionAnalyserTest.cpp:50-77)
    with DedispersionDataGenerator::dedispersionData(10) (:86-120): a DM-10
    signal, 2 x 10 / 0.1 = 200 trials of step 0.1 over a 6400-sample buffer,
    default analyser (threshold 6, 2^6 segments): more events than DM bins."""
    from synth import generator_blob
    n_sub, n_ch, nsamp = 32, 64, 6400
    tdms = int(2 * 10 / 0.1)
    buf = generator_blob(10.0)
    p = orc.dedisp_plan(n_sub * n_ch, nsamp, tdms, 0.0, 0.1, 0.00032768, 150.0, -6.0 / (n_sub * n_ch))
    assert p["rc"] == 0
    d = orc.dedisp_delays(p["shift"], tdms, p["mstartdm"], p["mdmstep"])
    out = orc.dedisperse(buf, d, p["nout"])
    ev, vals = orc.detect(out, n_ch, n_sub, 6.0, 6)
    assert len(ev) - 1 > tdms
    # single-sample events (binFactor 1) peak at 2048, the whole band summed
    # in one sample, and DM index 100 (DM 10) reaches that peak
    ev = np.asarray(ev)
    one = np.nonzero(ev[1:, 2] == 1)[0] + 1
    assert vals[one].max() == 2048.0
    assert np.any((ev[one, 0] == 100) & (vals[one] == 2048.0))


def test_u8_rows_quantiser_and_series_pinned(orc):
    """The oracle's u8 buffer with exceptional rows and its series, against a
    direct numpy statement: rows (columns of [nChan][nsamp]) with more than
    half their samples outside [-128, 127] after rint(x * inv) are stored
    relative to their middle channel; the series adds each such row's offset
    once per channel that lands on it.  A row with a third of its samples
    saturated stays as it is.  Rows with no exception quantise as before."""
    rng = np.random.default_rng(9)
    n, ns, inv = 24, 300, np.float32(12.7)
    x = rng.normal(size=(n, ns)).astype(np.float32)
    q0, off0 = orc.quantise_u8_rows(x, inv)
    assert np.array_equal(q0, orc.quantise_u8(x, inv)) and not off0.any()
    x[:, [5, 6, 200]] = -900.0 + rng.random(size=(n, 3)).astype(np.float32)
    x[: n // 3, 100] = 400.0
    q, off = orc.quantise_u8_rows(x, inv)
    assert list(np.nonzero(off)[0]) == [5, 6, 200]
    assert np.array_equal(off[[5, 6, 200]], x[n // 2, [5, 6, 200]])
    y = np.rint((x - off[None, :]) * inv)
    want = (np.clip(np.where(np.isnan(y), -128, y), -128, 127) + 128).astype(np.uint8)
    assert np.array_equal(q, want)
    # series: random delays non-decreasing in channel, against enumeration
    tdms, n_out = 7, 200
    d = np.sort(rng.integers(0, 90, size=(tdms, n)), axis=1).astype(np.int32)
    d[:, 0] = 0
    sums = orc.dedisperse(q, d, n_out)
    got = orc.u8_series(sums, off, d, n_out, np.float32(1.0 / inv))
    rows = [e for e in range(ns) if off[e] != 0]
    for dm in range(tdms):
        for t in range(n_out):
            corr = np.float32(0.0)
            for e in rows:
                cnt = int(np.count_nonzero(t + d[dm] == e))
                if cnt:
                    corr = np.float32(corr + np.float32(off[e] * np.float32(cnt)))
            v = np.float32(np.float32(sums[dm, t] - 128 * n) * np.float32(1.0 / inv)) + corr
            assert got[dm, t] == np.float32(v), (dm, t)


def _dyadic_sum(row, t, bf):
    """A window's sum as the analyser's binning tree defines it
    (DedispersionAnalyser.cpp:105-128): the sum of a window of 2^k samples is
    the float sum of its two halves' sums."""
    if bf == 1:
        return np.float32(row[t])
    h = bf // 2
    return np.float32(_dyadic_sum(row, t, h) + _dyadic_sum(row, t + h, h))


@pytest.mark.parametrize("integer", [True, False])
def test_detection_brute_force_every_level(orc, integer):
    """Every dyadic window of every level, enumerated: window [t, t + 2^k)
    with t a multiple of 2^k inside each whole 64-sample segment, k = 0..6,
    is an event exactly when its sum reaches threshold * rms * sqrt(2^k) in
    float (rms = sqrt(nChannels * nSubbands), DedispersionAnalyser.cpp:
    87-128).  The event set and values must be these windows and sums, at
    every level.  On integer data every summation order gives the same sums,
    so the enumeration alone decides; on float data the sums follow the
    halving tree."""
    rng = np.random.default_rng(31 if integer else 32)
    tdms, nsamp, nch, nsub, bin_pow2, thr = 3, 64 * 4 + 23, 4, 4, 6, 6.0
    if integer:
        data = rng.integers(-6, 10, size=(tdms, nsamp)).astype(np.float32)
        data[0, 5] = 30.0
        data[1, 100:104] = 25.0
    else:
        data = rng.normal(2.0, 4.0, size=(tdms, nsamp)).astype(np.float32)
        data[0, 10] = 40.5
        data[1, 66:68] = 25.25
        data[2, 130:146] += 9.0
    rms = np.float32(np.sqrt(np.float32(nch) * np.float32(nsub)))
    ev, vals = orc.detect(data, nch, nsub, thr, bin_pow2)
    got = {(int(d), int(t), int(b)): np.float32(v) for (d, t, b), v in zip(ev[1:], vals[1:])}
    want = {}
    for dm in range(tdms):
        for seg in range(nsamp // 64):
            for k in range(bin_pow2 + 1):
                bf = 1 << k
                level = np.float32(np.float32(thr) * rms) * np.float32(np.sqrt(np.float32(bf)))
                for t in range(seg * 64, seg * 64 + 64, bf):
                    s = (np.float32(np.sum(data[dm, t:t + bf], dtype=np.float64)) if integer
                         else _dyadic_sum(data[dm], t, bf))
                    if s >= level:
                        want[(dm, t, bf)] = s
    assert len(ev) - 1 == len(got)  # no window reported twice
    assert set(got) == set(want)
    assert all(got[k] == want[k] for k in want)
    # every level has events, so a detector dropping one fails here
    assert {b for (_, _, b) in want} == {1 << k for k in range(bin_pow2 + 1)}
