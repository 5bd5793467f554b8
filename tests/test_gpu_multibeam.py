"""Many beams of different shapes stepped together (batched clipper and
dedispersion launches) give, bit for bit, what independent single-beam chains
give on the same voltage streams, eagerly and replayed from HIP graphs."""
import numpy as np
import pytest
import torch

from synth import beam_bandpass, device_voltages

pytestmark = pytest.mark.gpu

SHAPES = [(8, 16, 4096, 128, 50.0), (4, 64, 8192, 64, 100.0), (2, 256, 2048, 32, 200.0)]


def _specs(ppl, dtype):
    specs = []
    for n_sub, n_ch, nsamp, tdms, dm_hi in SHAPES:
        h = ppl.ppf_coefficients(8, n_ch, "hann")
        foff = -0.1953125 / n_ch
        bp = beam_bandpass(h, n_sub, n_ch, 4, 150.0, foff)
        specs.append(ppl.BeamSpec(n_sub, n_ch, 8, h, 4, bp, nsamp, tdms, 0.0, dm_hi / tdms, 150.0, foff,
                                  history=1000))
    return specs


@pytest.mark.parametrize("dtype,dft", [("f32", "fp32"), ("f16", "fp32"), ("f16", "bf16x3"), ("u8", "bf16x3")])
def test_multibeam_equals_single_beam_chains(ppl, dtype, dft):
    specs = _specs(ppl, dtype)
    mb = ppl.MultiBeam(specs, dtype=dtype, max_chunk_samples=1 << 15, dft=dft)
    volts = []
    for b, sp in enumerate(specs):
        pulses = [(0.6 * sp.dm_step * sp.tdms, 100.0, 3.0, 40.0)]
        volts.append(device_voltages(torch, sp.n_sub, sp.n_ch, 4, mb.chunk_samples(b), 70 + b, sp.fch1, sp.foff,
                                     pulses=pulses, impulse_frac=0.002))
    singles = []
    for b, sp in enumerate(specs):
        prec = dft if sp.n_ch in (64, 256) else "fp32"
        ch = ppl.BeamChain(sp.n_sub, sp.n_ch, 8, sp.coeffs, 4, sp.bp_values, sp.nsamp_buf, sp.tdms, 0.0,
                           sp.dm_step, sp.tsamp, sp.fch1, sp.foff, dtype=dtype, history=1000, dft=prec)
        singles.append(ch)
    # the single chains read the same stream: MultiBeam's chunks, each a
    # prefix of the beam's voltage chunk, one after the other
    def stream(b, n_spec):
        per = 4 * specs[b].n_ch
        return torch.cat([volts[b][:, :, :k * per] for k in mb.chunk_lengths(b, n_spec)], dim=2).contiguous()

    step_in = [stream(b, mb.nout[b]) for b in range(len(specs))]
    prime_in = [stream(b, mb.overlap[b]) for b in range(len(specs))]
    mb.prime(volts)
    for b, ch in enumerate(singles):
        ch.prime(prime_in[b])
    mb_in = volts
    for k in range(2):
        mb.step(mb_in)
        for b, ch in enumerate(singles):
            ch.step(step_in[b])
            assert torch.equal(mb.outs[b], ch.out), (k, b)
            ev_m, val_m, _ = mb.detect(b)
            ev_s, val_s, _ = ch.detect()
            assert np.array_equal(ev_m, ev_s) and np.array_equal(val_m, val_s), (k, b)
    # graph replays continue the same streams
    mb.capture_steps(mb_in)
    for b, ch in enumerate(singles):
        ch.capture_steps(step_in[b])
    for k in range(3):
        mb.replay()
        for b, ch in enumerate(singles):
            ch.replay()
        torch.cuda.synchronize()
        for b, ch in enumerate(singles):
            assert torch.equal(mb.outs[b], ch.out), ("graph", k, b)


@pytest.mark.parametrize("per_beam", [False, True])
def test_multibeam_autotune_keeps_results(ppl, per_beam):
    specs = _specs(ppl, "f32")
    mb = ppl.MultiBeam(specs, max_chunk_samples=1 << 15)
    ms = mb.autotune(per_beam=per_beam, candidates=[(16, 512, 4), (32, 256, 8), (8, 512, 8)] if per_beam else None)
    assert ms > 0 and (mb.tile == "own" or mb.tile[0] % 4 == 0)
    outs = [torch.randn((sp.n_sub * sp.n_ch, sp.nsamp_buf), device="cuda") for sp in specs]
    got = ppl.dedisperse_batch(mb.plans, outs, tile=mb.tile)
    for p, x, g in zip(mb.plans, outs, got):
        assert torch.equal(p.run(x), g)
