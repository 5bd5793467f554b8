"""End-to-end chain at the CPU-runnable configuration (8 subbands x 2 pol,
complex int16, sigma 64, one DM-30 pulse of 2 ms at 5 sigma per channel, 16
channels x 8 taps, Stokes-I x4, clipper on default thresholds, 128 DMs 0-50
over 4096-sample buffers, 6-sigma detection).

Two checks:
- stage-wise: the oracle fed the GPU's own Stokes-I output reproduces the
  GPU clipper output, dedispersion buffer contents, dedispersed series and
  event list bit for bit;
- end to end: the oracle's own chain (FP32 FIR + FP64 DFT) and the GPU
  chain find the same events, apart from events within 1e-3 of a level
  threshold, and both find the injected pulse at its DM and time.
"""
import numpy as np
import pytest
import torch

from synth import inject_pulse_samples, noise_voltages, tone_amplitude

pytestmark = pytest.mark.gpu

N_SUB, N_CH, N_TAPS, INTEG = 8, 16, 8, 4
TSAMP = 5.12e-6 * N_CH * INTEG
FCH1, FOFF = 141.5625, -0.01220703125
TDMS, DM_STEP, NSAMP_BUF = 128, 50.0 / 128, 4096
CHUNK = 16384            # input samples per subband and polarisation per call
N_CHUNKS = 56            # 14336 integrated spectra: learning (10000) then steady state
PULSE = dict(dm=30.0, t0=12000.0, width=2e-3 / TSAMP, excess=5.0)


def _inputs(orc):
    coeffs = orc.ppf_coefficients(N_TAPS, N_CH, "hann")
    n_samp = CHUNK * N_CHUNKS
    x = noise_voltages(N_SUB, n_samp, 64.0, seed=2024)
    amp = tone_amplitude(coeffs, N_CH, 64.0, PULSE["excess"], INTEG)
    inject_pulse_samples(x, N_CH, FCH1, FOFF, TSAMP, N_CH * INTEG, PULSE["dm"], PULSE["t0"],
                         PULSE["width"], amp, seed=5)
    e_h = float(np.sum(coeffs * coeffs))
    level = 2 * INTEG * 2.0 * 64.0 * 64.0 * e_h
    n_chan = N_SUB * N_CH
    bp = np.array([n_chan, FCH1, FCH1 + FOFF * (n_chan - 1), FOFF, np.sqrt(8) * level / 8, level])
    return coeffs, x, bp


def _gpu_chain(ppl, coeffs, x, bp):
    n_chan = N_SUB * N_CH
    plan = ppl.DedispPlan(n_chan, NSAMP_BUF, TDMS, 0.0, DM_STEP, TSAMP, FCH1, FOFF)
    ppf = ppl.PPF(N_SUB, N_CH, N_TAPS, coeffs.astype(np.float32), integrate=INTEG)
    clip = ppl.Clipper(bp)
    buf = ppl.DDBuffer(n_chan, NSAMP_BUF, plan.maxshift + plan.rem)
    xd = torch.from_numpy(x).cuda()
    stokes_raw, stokes_clipped, buffers, series, events = [], [], [], [], []
    for k in range(N_CHUNKS):
        chunk = xd[:, :, k * CHUNK:(k + 1) * CHUNK].contiguous()
        st = ppf.stokes_i16(chunk, CHUNK // N_CH)
        stokes_raw.append(st.cpu().numpy())
        clip.run(st)
        stokes_clipped.append(st.cpu().numpy())
        flat = st.view(st.shape[0], n_chan)
        first = 0
        while first < flat.shape[0]:
            used, full = buf.add(flat, first)
            first += used
            if full:
                buffers.append(buf.tensor().cpu().numpy().copy())
                out = plan.run(buf.tensor())
                series.append(out.cpu().numpy())
                events.append(ppl.detect(out, N_CH, N_SUB, 6.0, 6)[:2])
                buf.rotate()
    return plan, stokes_raw, stokes_clipped, buffers, series, events


@pytest.fixture(scope="module")
def chain(orc, ppl):
    coeffs, x, bp = _inputs(orc)
    return coeffs, x, bp, _gpu_chain(ppl, coeffs, x, bp)


def test_stagewise_bit_exact(orc, chain):
    coeffs, x, bp, (plan, raw, clipped, buffers, series, events) = chain
    n_chan = N_SUB * N_CH
    oc = orc.Clipper(bp)
    p = orc.dedisp_plan(n_chan, NSAMP_BUF, TDMS, 0.0, DM_STEP, TSAMP, FCH1, FOFF)
    delays = orc.dedisp_delays(p["shift"], TDMS, p["mstartdm"], p["mdmstep"])
    ob = orc.DDBuffer(n_chan, NSAMP_BUF, p["maxshift"] + p["rem"])
    o_buffers = []
    for k in range(N_CHUNKS):
        I = raw[k].copy()
        oc.run(I)
        assert np.array_equal(I, clipped[k]), k
        o_buffers.extend(list(ob.add(I.reshape(I.shape[0], n_chan))))
    assert len(o_buffers) == len(buffers) >= 4
    for b, (ob_k, gb_k) in enumerate(zip(o_buffers, buffers)):
        assert np.array_equal(ob_k, gb_k), b
        o_series = orc.dedisperse(ob_k, delays, p["nout"])
        assert np.array_equal(o_series, series[b]), b
        ev, vals = orc.detect(o_series, N_CH, N_SUB, 6.0, 6)
        assert np.array_equal(ev, events[b][0]) and np.array_equal(vals, events[b][1]), b


def _event_set(ev, vals, margin=1e-3):
    rms = np.float32(np.sqrt(np.float32(N_CH) * np.float32(N_SUB)))
    thr = 6.0 * rms * np.sqrt(ev[:, 2].astype(np.float64))
    clear = vals >= thr * (1 + margin)
    near = (vals >= thr * (1 - margin)) & ~clear
    return {tuple(e) for e in ev[clear]}, {tuple(e) for e in ev[near]}


def test_end_to_end_event_lists(orc, chain):
    coeffs, x, bp, (plan, raw, clipped, buffers, series, events) = chain
    n_chan = N_SUB * N_CH
    # the oracle's own chain
    op = orc.PPF(N_SUB, 2, N_CH, N_TAPS, coeffs)
    oc = orc.Clipper(bp)
    p = orc.dedisp_plan(n_chan, NSAMP_BUF, TDMS, 0.0, DM_STEP, TSAMP, FCH1, FOFF)
    delays = orc.dedisp_delays(p["shift"], TDMS, p["mstartdm"], p["mdmstep"])
    ob = orc.DDBuffer(n_chan, NSAMP_BUF, p["maxshift"] + p["rem"])
    o_events = []
    for k in range(N_CHUNKS):
        spec = op.run_fir32(np.ascontiguousarray(x[:, :, k * CHUNK:(k + 1) * CHUNK]), CHUNK // N_CH)
        sp32 = np.stack([spec.real, spec.imag], -1).astype(np.float32)
        I = orc.integrate(orc.stokes_i_f32(sp32), INTEG)
        oc.run(I)
        for b in ob.add(I.reshape(I.shape[0], n_chan)):
            o_events.append(orc.detect(orc.dedisperse(b, delays, p["nout"]), N_CH, N_SUB, 6.0, 6))
    assert len(o_events) == len(events)
    found_pulse = False
    dm_idx = int(round(PULSE["dm"] / DM_STEP))
    for b, ((ge, gv), (oe, ov)) in enumerate(zip(events, o_events)):
        g_clear, g_near = _event_set(ge, gv)
        o_clear, o_near = _event_set(oe, ov)
        assert g_clear - o_clear <= o_near, (b, sorted(g_clear - o_clear)[:5])
        assert o_clear - g_clear <= g_near, (b, sorted(o_clear - g_clear)[:5])
        start = b * p["nout"]
        for dm, t, bf in g_clear:
            if abs(dm - dm_idx) <= 2 and abs(start + t - PULSE["t0"]) <= 16:
                found_pulse = True
    assert found_pulse


def test_u8_chain_series_and_events_follow_f32(ppl):
    """The u8 chain against the FP32 chain on one stream with broadband
    impulses (which the clipper rejects and leaves thousands of units below
    the data): the u8 series stays within the quantisation bound of the FP32
    series everywhere -- the rejected spectra enter at their own level -- and
    the event lists agree (recall and precision 0.99 or better)."""
    from synth import beam_bandpass, device_voltages
    n_sub, n_ch, nsamp, tdms, dm_hi = 8, 16, 4096, 128, 50.0
    foff = -0.1953125 / n_ch
    h = ppl.ppf_coefficients(8, n_ch, "hann")
    bp = beam_bandpass(h, n_sub, n_ch, 4, 150.0, foff)
    tsamp = 5.12e-6 * n_ch * 4
    chains = {dt: ppl.BeamChain(n_sub, n_ch, 8, h, 4, bp, nsamp, tdms, 0.0, dm_hi / tdms, tsamp, 150.0, foff,
                                dtype=dt, history=1000) for dt in ("f32", "u8")}
    c32 = chains["f32"]
    amp = tone_amplitude(h, n_ch, 64.0, 20.0, 4)
    spc = 4 * n_ch
    n_steps = 4
    v = device_voltages(torch, n_sub, n_ch, 4, (c32.overlap + n_steps * c32.nout) * spc, 77, 150.0, foff,
                        pulses=[(20.0 + 2 * j, 1500.0 + 500.0 * j, 4.0, amp) for j in range(24)],
                        impulse_frac=0.003, wrap=False)
    prime = v[:, :, :c32.overlap * spc].contiguous()
    for ch in chains.values():
        ch.prime(prime)
    inv = chains["u8"].inv_scale
    hits = miss = extra = 0
    rejected_reach = 0.0
    for k in range(n_steps):
        a = c32.overlap * spc + k * c32.nout * spc
        step_in = v[:, :, a:a + c32.nout * spc].contiguous()
        for ch in chains.values():
            ch.step(step_in)
        ref, s8 = c32.out, chains["u8"].series
        bound = n_sub * n_ch * 0.5 / inv * 1.001 + 1e-5 * ref.abs()
        assert bool(torch.all((s8 - ref).abs() <= bound)), k
        rejected_reach = min(rejected_reach, float(ref.min()))
        e32 = {tuple(e) for e in c32.detect()[0][1:]}
        e8 = {tuple(e) for e in chains["u8"].detect()[0][1:]}
        hits += len(e32 & e8)
        miss += len(e32 - e8)
        extra += len(e8 - e32)
    assert rejected_reach < -1000.0  # rejected spectra did land in the series
    assert hits > 50
    assert hits / (hits + miss) >= 0.99 and hits / (hits + extra) >= 0.99, (hits, miss, extra)
