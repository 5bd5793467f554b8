"""The batched and streaming forms at the workload shapes they are built
for, checked against the oracle and against size-independent properties.

Batched (24 heterogeneous beams, one launch per stage for all beams):
- at full size (31-62 subbands x 16/64/256 channels, 512-4096 DM trials over
  0-100 .. 0-1000 pc/cm^3, 2^14-2^17-sample buffers), every beam's output of
  one batched dedispersion launch against properties that fix it: all-ones
  input sums to nChan exactly, two hot samples land at s - delay[dm][c] with
  their values (the delay table checked against the oracle's), and sums of
  small integers are linear, exactly;
- with the same channel structures at buffers the CPU covers (4096 samples,
  64 trials), the whole batched chain -- clipper, buffer, dedispersion,
  detection, two steps after the overlap prime -- against the oracle's
  clipper, buffer, brute-force sums and detector, bit for bit.

Streaming (beams of the LOFAR shape: 62 x 64 channels, 2048 trials over
0-500, 65536-sample buffers, 8192-spectrum chunks): each chunk's clipped
spectra against the oracle clipper, the overlap carried in the buffer against
the oracle's buffer, the dedispersed series against the same launch on the
oracle's buffer, and the service's plan against the full-size properties."""
import numpy as np
import pytest
import torch

from synth import beam_bandpass, device_voltages, draw_multibeam, shift_table, tone_amplitude

pytestmark = pytest.mark.gpu

TS_IN, INTEG, FCH1 = 5.12e-6, 4, 150.0


def _plan(ppl, n_sub, n_ch, nsamp, tdms, dm_hi, dtype="f32"):
    tsamp = TS_IN * n_ch * INTEG
    return ppl.DedispPlan(n_sub * n_ch, nsamp, tdms, 0.0, dm_hi / tdms, tsamp, FCH1, -0.1953125 / n_ch, dtype=dtype)


def _oracle_delays(orc, plan):
    c = plan.cfg
    p = orc.dedisp_plan(c.nChan, c.nsampBuf, c.tdms, c.dmLow, c.dmStep, c.tsamp, c.fch1, c.foff)
    assert p["rc"] == 0 and (p["maxshift"], p["rem"], p["nout"]) == (plan.maxshift, plan.rem, plan.nout)
    return orc.dedisp_delays(p["shift"], c.tdms, p["mstartdm"], p["mdmstep"])


def _hot_expected(delays, nout, hot):
    """[tdms][nOut] of a buffer that is zero but for hot = [(c, s, v)]."""
    d = torch.from_numpy(np.ascontiguousarray(delays)).cuda().long()
    exp = torch.zeros((d.shape[0], nout), device="cuda")
    rows = torch.arange(d.shape[0], device="cuda")
    for c, s, v in hot:
        t = s - d[:, c]
        ok = (t >= 0) & (t < nout)
        exp.index_put_((rows[ok], t[ok]), torch.full((int(ok.sum()),), float(v), device="cuda"), accumulate=True)
    return exp


@pytest.mark.parametrize("dtype", ["f32", "f16"])
def test_multibeam_batch_full_size_properties(orc, ppl, dtype):
    beams = draw_multibeam(2024, 24)
    plans = [_plan(ppl, b["n_sub"], b["n_ch"], b["nsamp"], b["tdms"], b["dm_hi"], dtype) for b in beams]
    tdt = torch.float32 if dtype == "f32" else torch.float16
    shapes = [(p.cfg.nChan, p.cfg.nsampBuf) for p in plans]
    delays = [_oracle_delays(orc, p) for p in plans]
    for p, d in zip(plans, delays):
        assert np.array_equal(p.delays(), d)
    for tile in ("own", (16, 512, 0, 4)):
        # all ones: every trial and sample sums to nChan exactly
        ins = [torch.ones(s, dtype=tdt, device="cuda") for s in shapes]
        outs = ppl.dedisperse_batch(plans, ins, tile=tile)
        for k, (p, o) in enumerate(zip(plans, outs)):
            assert o.shape == (p.cfg.tdms, p.nout)
            assert bool(torch.all(o == float(p.cfg.nChan))), (tile, k)
        del ins, outs
        # two hot samples: channel 0 (delay 0) at the last output sample, and a
        # third of the band in, part way through the buffer
        hots = []
        ins = []
        for p in plans:
            n_chan, ns = p.cfg.nChan, p.cfg.nsampBuf
            hot = [(0, p.nout - 1, 2.0), (n_chan // 3, p.nout // 2 + p.maxshift // 2, 1.0)]
            x = torch.zeros((n_chan, ns), dtype=tdt, device="cuda")
            for c, s, v in hot:
                x[c, s] = v
            hots.append(hot)
            ins.append(x)
        outs = ppl.dedisperse_batch(plans, ins, tile=tile)
        del ins
        for k, (p, o) in enumerate(zip(plans, outs)):
            exp = _hot_expected(delays[k], p.nout, hots[k])
            assert torch.equal(o, exp), (tile, k)
            assert float(o[:, p.nout - 1].min()) >= 2.0  # channel 0 reaches the last sample in every trial
            del exp
        del outs
    if dtype == "f32":
        # linearity on small integers (every partial sum exact in float32)
        g = torch.Generator(device="cuda").manual_seed(5)
        a = [torch.randint(-3, 4, s, generator=g, device="cuda").float() for s in shapes]
        b = [torch.randint(-3, 4, s, generator=g, device="cuda").float() for s in shapes]
        oa = ppl.dedisperse_batch(plans, a)
        ob = ppl.dedisperse_batch(plans, b)
        for x, y in zip(a, b):
            x.add_(y)
        del b
        oab = ppl.dedisperse_batch(plans, a)
        for k in range(len(plans)):
            assert torch.equal(oa[k] + ob[k], oab[k]), k


def _reduced_specs(ppl, beams, nsamp=4096, tdms=64, history=1000):
    """The drawn beams' channel structures with a DM range whose sweep fills
    about 30 % of a 4096-sample buffer."""
    specs = []
    for b in beams:
        n_sub, n_ch = b["n_sub"], b["n_ch"]
        foff = -0.1953125 / n_ch
        tsamp = TS_IN * n_ch * INTEG
        dm_hi = 0.3 * nsamp * tsamp / shift_table(n_sub * n_ch, FCH1, foff)[-1]
        h = ppl.ppf_coefficients(8, n_ch, "hann")
        bp = beam_bandpass(h, n_sub, n_ch, INTEG, FCH1, foff)
        specs.append(ppl.BeamSpec(n_sub, n_ch, 8, h, INTEG, bp, nsamp, tdms, 0.0, dm_hi / tdms, FCH1, foff,
                                  history=history))
    return specs


def _channelise_like(ppl_obj, volt, n_spec, m, n_ch, out):
    """MultiBeam's chunk-by-chunk channelisation of a periodic chunk."""
    done = 0
    while done < n_spec:
        k = min(m, n_spec - done)
        src = volt if k == m else volt[:, :, :k * INTEG * n_ch].contiguous()
        ppl_obj.stokes_i16(src, k * INTEG, d_out=out[done:done + k])
        done += k


def test_multibeam_reduced_buffers_vs_oracle(orc, ppl):
    beams = draw_multibeam(2024, 24)
    specs = _reduced_specs(ppl, beams)
    mb = ppl.MultiBeam(specs, max_chunk_samples=1 << 15)
    volts, twins, clips = [], [], []
    for b, (bm, sp) in enumerate(zip(beams, specs)):
        m = mb.chunk_spectra[b]
        amp = tone_amplitude(sp.coeffs, sp.n_ch, 64.0, 10.0, INTEG)
        pulses = [(0.5 * sp.dm_step * sp.tdms, 0.3 * m, 4.0, amp)]
        volts.append(device_voltages(torch, sp.n_sub, sp.n_ch, INTEG, mb.chunk_samples(b), bm["seed"], FCH1, sp.foff,
                                     pulses=pulses, impulse_frac=0.004))
        # the same channeliser on the same stream gives the spectra the batch clips
        twins.append(ppl.PPF(sp.n_sub, sp.n_ch, 8, sp.coeffs, integrate=INTEG))
        clips.append(orc.Clipper(sp.bp_values, history=sp.history))
    # the oracle's chain, spectrum by spectrum in order: prime, then two steps
    clipped = [[] for _ in specs]

    def oracle_clip(b, n_spec):
        sp, m = specs[b], mb.chunk_spectra[b]
        st = torch.empty((n_spec, sp.n_sub, 1, sp.n_ch), device="cuda")
        _channelise_like(twins[b], volts[b], n_spec, m, sp.n_ch, st)
        x = st.cpu().numpy()
        clips[b].run(x)
        clipped[b].append(x.reshape(n_spec, sp.n_sub * sp.n_ch))
        return x

    mb.prime(volts)
    for b in range(len(specs)):
        oracle_clip(b, mb.overlap[b])
    n_events = 0
    for k in range(2):
        mb.step(volts)
        for b, sp in enumerate(specs):
            x = oracle_clip(b, mb.nout[b])
            assert np.array_equal(mb.stokes[b].cpu().numpy(), x, equal_nan=True), (k, b)
            allsp = np.concatenate(clipped[b])
            nout, ov = mb.nout[b], mb.overlap[b]
            buf = np.ascontiguousarray(allsp[k * nout:k * nout + sp.nsamp_buf].T)
            plan = mb.plans[b]
            d = _oracle_delays(orc, plan)
            series = orc.dedisperse(buf, d, nout)
            assert np.array_equal(mb.outs[b].cpu().numpy(), series), (k, b)
            # the carried overlap now starts the beam's next buffer
            carried = mb.bufs[b].tensor()[:, :ov].cpu().numpy()
            assert np.array_equal(carried, allsp[(k + 1) * nout:(k + 1) * nout + ov].T), (k, b)
            ev, vals, _ = mb.detect(b)
            ev_o, vals_o = orc.detect(series, sp.n_ch, sp.n_sub)
            assert np.array_equal(ev, ev_o) and np.array_equal(vals, vals_o), (k, b)
            n_events += len(ev) - 1
    assert n_events > 0


def test_service_lofar_beams_vs_oracle(orc, ppl):
    n_sub, n_ch, nsamp, tdms, dm_hi, m = 62, 64, 65536, 2048, 500.0, 8192
    foff = -0.1953125 / n_ch
    h = ppl.ppf_coefficients(8, n_ch, "hann")
    bp = beam_bandpass(h, n_sub, n_ch, INTEG, FCH1, foff)
    sp = ppl.BeamSpec(n_sub, n_ch, 8, h, INTEG, bp, nsamp, tdms, 0.0, dm_hi / tdms, FCH1, foff, history=10000)
    svc = ppl.StreamService(sp, m, budget_bytes=1 << 36)
    n_chan = n_sub * n_ch
    plan = svc.plan
    delays = _oracle_delays(orc, plan)
    assert np.array_equal(plan.delays(), delays)
    keys = ["a", "b"]
    for key in keys:
        assert svc.admit(key)
    twins = {k: ppl.PPF(n_sub, n_ch, 8, h, integrate=INTEG) for k in keys}
    clips = {k: orc.Clipper(bp, history=10000) for k in keys}
    spectra = {k: [] for k in keys}
    amp = tone_amplitude(h, n_ch, 64.0, 12.0, INTEG)
    done = {k: 0 for k in keys}
    n_rounds = nsamp // m + 2  # the first buffer fills after nsamp spectra
    for r in range(n_rounds):
        ready = []
        for j, key in enumerate(keys):
            pulses = [(100.0 + 150.0 * j, 0.4 * m, 4.0, amp)]
            v = device_voltages(torch, n_sub, n_ch, INTEG, m * INTEG * n_ch, 1000 + 37 * r + j, FCH1, foff,
                                pulses=pulses, impulse_frac=0.002, wrap=False)
            st = torch.empty((m, n_sub, 1, n_ch), device="cuda")
            twins[key].stokes_i16(v, m * INTEG, d_out=st)
            x = st.cpu().numpy()
            clips[key].run(x)
            spectra[key].append(x.reshape(m, n_chan))
            ready.append((key, v))
        results = svc.process(ready)
        for key in keys:
            assert np.array_equal(svc.beams[key].stokes.cpu().numpy().reshape(m, n_chan), spectra[key][-1],
                                  equal_nan=True), (r, key)
        for key, idx, ints, vals in results:
            allsp = np.concatenate(spectra[key])
            start = idx * plan.nout
            buf = np.ascontiguousarray(allsp[start:start + nsamp].T)
            # the service's series equal the same launch on the oracle's buffer
            ref = plan.run(torch.from_numpy(buf).cuda())
            assert torch.equal(svc.beams[key].out, ref), (r, key)
            ev, vv, _ = ppl.detect(ref, n_ch, n_sub)
            assert np.array_equal(ev, ints) and np.array_equal(vv, vals), (r, key)
            # the carried overlap starts the next buffer
            ov = plan.maxshift + plan.rem
            nxt = allsp[start + plan.nout:start + plan.nout + ov].T
            assert np.array_equal(svc.beams[key].buf.tensor()[:, :ov].cpu().numpy(), nxt), (r, key)
            done[key] += 1
    assert all(done[k] >= 1 for k in keys)
    # the plan at this size: all ones, two hot samples, exact
    out = plan.run(torch.ones((n_chan, nsamp), device="cuda"))
    assert bool(torch.all(out == float(n_chan)))
    hot = [(0, plan.nout - 1, 2.0), (n_chan - 1, plan.maxshift + 7, 1.0)]
    x = torch.zeros((n_chan, nsamp), device="cuda")
    for c, s, v in hot:
        x[c, s] = v
    assert torch.equal(plan.run(x), _hot_expected(delays, plan.nout, hot))
