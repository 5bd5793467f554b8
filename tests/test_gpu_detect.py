"""GPU parity of the dyadic boxcar detection with the oracle: identical
event lists (indices, binning, order) and values."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _data(tdms, nsamp, nch, nsub, seed, pulses=()):
    rng = np.random.default_rng(seed)
    rms = np.float32(np.sqrt(np.float32(nch) * np.float32(nsub)))
    x = (rng.normal(size=(tdms, nsamp)) * rms).astype(np.float32)
    for dm, t, w, a in pulses:
        x[dm, t:t + w] += a * rms
    return x


@pytest.mark.parametrize("bin_pow2", [6, 5, 4, 2, 0])
@pytest.mark.parametrize("tail", [19, 12])
def test_events_identical(orc, ppl, bin_pow2, tail):
    """Rows of a length that is a multiple of four take the 16-byte count
    pass (segments of 16..64 samples), others the one-sample pass."""
    x = _data(64, 64 * 40 + tail, 16, 8, 1, [(10, 300, 9, 20), (11, 301, 9, 15), (40, 2000, 64, 3)])
    ev_o, val_o = orc.detect(x, 16, 8, 6.0, bin_pow2)
    ev_g, val_g, n = ppl.detect(torch.from_numpy(x).cuda(), 16, 8, 6.0, bin_pow2)
    assert n == len(ev_o)
    assert np.array_equal(ev_g, ev_o)
    assert np.array_equal(val_g, val_o)


def test_no_events_only_dummy(orc, ppl):
    x = np.zeros((8, 640), np.float32)
    ev, val, n = ppl.detect(torch.from_numpy(x).cuda(), 16, 8)
    assert n == 1 and tuple(ev[0]) == (0, 0, 1)


def test_truncated_event_buffer(orc, ppl):
    x = _data(16, 64 * 8, 16, 8, 2, [(3, 10, 30, 30)])
    ev_o, _ = orc.detect(x, 16, 8)
    ev_g, _, n = ppl.detect(torch.from_numpy(x).cuda(), 16, 8, max_events=5)
    assert n == len(ev_o) and len(ev_g) == 5
    assert np.array_equal(ev_g, ev_o[:5])


def test_u8_detection_dequantises(orc, ppl):
    rng = np.random.default_rng(3)
    n_chan, tdms, nsamp = 128, 16, 64 * 10
    s = rng.integers(128 * n_chan - 3000, 128 * n_chan + 3000, size=(tdms, nsamp)).astype(np.int32)
    s[5, 100:108] += 20000
    scale = np.float32(10.0 / 127.0)
    deq = ((s - 128 * n_chan).astype(np.float32) * scale).astype(np.float32)
    ev_o, val_o = orc.detect(deq, 16, 8)
    ev_g, val_g, n = ppl.detect_u8(torch.from_numpy(s).cuda(), n_chan, float(scale), 16, 8)
    assert np.array_equal(ev_g, ev_o) and np.array_equal(val_g, val_o)


def test_reference_single_event(orc, ppl):
    """DedispersionAnalyserTest::test_singleEvent (DedispersionAnalyserTest.cpp:50-77)
    through the device plan, dedispersion and detection: more events than DM
    bins, the same events and values as the oracle."""
    from synth import generator_blob
    n_sub, n_ch, nsamp, tdms = 32, 64, 6400, 200
    buf = generator_blob(10.0)
    plan = ppl.DedispPlan(n_sub * n_ch, nsamp, tdms, 0.0, 0.1, 0.00032768, 150.0, -6.0 / (n_sub * n_ch))
    out = plan.run(torch.from_numpy(buf).cuda())
    ev, vals, n = ppl.detect(out, n_ch, n_sub)
    assert n - 1 > tdms
    p = orc.dedisp_plan(n_sub * n_ch, nsamp, tdms, 0.0, 0.1, 0.00032768, 150.0, -6.0 / (n_sub * n_ch))
    d = orc.dedisp_delays(p["shift"], tdms, p["mstartdm"], p["mdmstep"])
    ev_o, vals_o = orc.detect(orc.dedisperse(buf, d, p["nout"]), n_ch, n_sub, 6.0, 6)
    assert np.array_equal(ev, ev_o) and np.array_equal(vals, vals_o)


@pytest.mark.parametrize("threshold", [2.5, 4.0])
@pytest.mark.parametrize("tail", [37, 36])
def test_events_identical_many_workgroups(orc, ppl, threshold, tail):
    """Low thresholds on noise: hits in most segments of 512 DM rows x 600
    segments (4800 workgroups of 64 segments), and a truncated list of the
    same events."""
    x = _data(512, 64 * 600 + tail, 16, 8, 5)
    ev_o, val_o = orc.detect(x, 16, 8, threshold, 6, max_events=1 << 22)
    d = torch.from_numpy(x).cuda()
    ev_g, val_g, n = ppl.detect(d, 16, 8, threshold, 6, max_events=1 << 22)
    assert n == len(ev_o) > (100000 if threshold < 3 else 1000)
    assert np.array_equal(ev_g, ev_o)
    assert np.array_equal(val_g, val_o)
    ev_t, val_t, n_t = ppl.detect(d, 16, 8, threshold, 6, max_events=777)
    assert n_t == n and np.array_equal(ev_t, ev_o[:777]) and np.array_equal(val_t, val_o[:777])


def test_events_identical_two_scan_rounds(orc, ppl):
    """1024 DM rows x 1100 segments: 17600 workgroups, more than one
    16384-entry round of the workgroup scan."""
    x = _data(1024, 64 * 1100 + 36, 16, 8, 6, [(700, 40000, 20, 9)])
    ev_o, val_o = orc.detect(x, 16, 8, 4.0, 6, max_events=1 << 22)
    ev_g, val_g, n = ppl.detect(torch.from_numpy(x).cuda(), 16, 8, 4.0, 6, max_events=1 << 22)
    assert n == len(ev_o) > 1000
    assert np.array_equal(ev_g, ev_o) and np.array_equal(val_g, val_o)


def test_count_only_and_empty_shapes(orc, ppl):
    """max_events = 0 returns the count alone; no DM rows, or rows shorter
    than one segment, give the dummy event alone."""
    x = _data(16, 64 * 8, 16, 8, 2, [(3, 10, 30, 30)])
    ev_o, _ = orc.detect(x, 16, 8)
    _, _, n = ppl.detect(torch.from_numpy(x).cuda(), 16, 8, max_events=0)
    assert n == len(ev_o)
    for shape in [(0, 640), (4, 63)]:
        ev, val, n = ppl.detect(torch.zeros(shape, device="cuda"), 16, 8)
        assert n == 1 and tuple(ev[0]) == (0, 0, 1) and val[0] == 0.0
