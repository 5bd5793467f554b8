"""The streaming service: admission by device footprint, and chunked
processing in rounds (chunks that straddle buffer boundaries, beams that join
late) that gives, bit for bit, the dedispersed buffers and events of a
single-beam chain on the same voltage stream."""
import numpy as np
import pytest
import torch

from synth import beam_bandpass, device_voltages, tone_amplitude

pytestmark = pytest.mark.gpu

N_SUB, N_CH, NSAMP, TDMS, DM_HI = 8, 16, 4096, 128, 50.0


def _spec(ppl):
    h = ppl.ppf_coefficients(8, N_CH, "hann")
    foff = -0.1953125 / N_CH
    bp = beam_bandpass(h, N_SUB, N_CH, 4, 150.0, foff)
    return ppl.BeamSpec(N_SUB, N_CH, 8, h, 4, bp, NSAMP, TDMS, 0.0, DM_HI / TDMS, 150.0, foff, history=1000)


def test_admission_by_footprint(ppl):
    sp = _spec(ppl)
    svc = ppl.StreamService(sp, 1000, budget_bytes=0)
    fp = svc.footprint()
    svc = ppl.StreamService(sp, 1000, budget_bytes=int(2.5 * fp))
    assert svc.admit("a") and svc.admit("b")
    assert not svc.admit("c")  # a third beam does not fit
    svc.release("a")
    assert svc.admit("c")
    with pytest.raises(ppl.PPLError):
        svc.admit("b")


@pytest.mark.parametrize("dtype", ["f32", "f16"])
def test_service_rounds_equal_single_beam_chain(ppl, dtype):
    sp = _spec(ppl)
    m = 1000  # spectra per chunk: does not divide nOut, chunks straddle buffers
    svc = ppl.StreamService(sp, m, budget_bytes=1 << 34, dtype=dtype)
    n_chunks = 12
    spc = m * 4 * N_CH
    streams = {}
    for k, key in enumerate(["x", "y"]):
        amp = tone_amplitude(sp.coeffs, N_CH, 64.0, 12.0, 4)
        pulses = [(25.0, 300.0 + 700 * j, 3.0, amp) for j in range(16)]
        v = device_voltages(torch, N_SUB, N_CH, 4, n_chunks * spc, 90 + k, 150.0, sp.foff, pulses=pulses)
        streams[key] = v
    assert svc.admit("x")
    got = {"x": [], "y": []}
    outs = {"x": [], "y": []}
    for r in range(n_chunks):
        if r == 3:
            assert svc.admit("y")  # joins three chunks later
        ready = []
        for key in ("x", "y"):
            if key in svc.beams:
                j = r if key == "x" else r - 3
                ready.append((key, streams[key][:, :, j * spc:(j + 1) * spc].contiguous()))
        for key, idx, ints, vals in svc.process(ready):
            got[key].append((idx, ints, vals))
            outs[key].append(svc.beams[key].out.clone())
    n_events = 0
    for key, n_used in (("x", n_chunks), ("y", n_chunks - 3)):
        ch = ppl.BeamChain(N_SUB, N_CH, 8, sp.coeffs, 4, sp.bp_values, NSAMP, TDMS, 0.0, sp.dm_step, sp.tsamp,
                           150.0, sp.foff, dtype=dtype, history=1000)
        v = streams[key]
        ov, nout = ch.overlap, ch.nout
        ch.prime(v[:, :, :ov * 4 * N_CH].contiguous())
        pos = ov * 4 * N_CH
        assert len(got[key]) == (n_used * m - ov) // nout and len(got[key]) >= 2
        for idx, ints, vals in got[key]:
            ch.step(v[:, :, pos:pos + nout * 4 * N_CH].contiguous())
            pos += nout * 4 * N_CH
            assert torch.equal(ch.out, outs[key][idx]), (key, idx)
            ei, ev, _ = ch.detect()
            assert np.array_equal(ei, ints) and np.array_equal(ev, vals), (key, idx)
            n_events += len(ints) - 1
    assert n_events > 0  # the pulses are found
