"""Many channelisers in one batched call (ppl_ppf_stokes_i16_batch): every
job's integrated Stokes-I equals the same plan's own ppl_ppf_stokes_i16 call
bit for bit, across consecutive calls (FIR history carried), for mixed
channel counts, tap counts, integrations and DFT methods, chunks that are
prefixes of a longer voltage buffer, and skipped jobs; the FP32 jobs also
against the oracle's FP64 chain within the channeliser's tolerance."""
import numpy as np
import pytest
import torch

from synth import noise_voltages

pytestmark = pytest.mark.gpu

# (n_sub, n_ch, n_taps, integrate, precision)
BEAMS = [(3, 16, 8, 4, "fp32"), (2, 64, 8, 4, "fp32"), (2, 256, 8, 4, "fp32"), (4, 8, 4, 1, "fp32"),
         (2, 32, 16, 2, "fp32"), (2, 64, 8, 4, "bf16x3"), (1, 256, 8, 4, "fp16x2"), (3, 16, 8, 4, "fp32")]


def _plans(orc, ppl):
    out = []
    for n_sub, n_ch, n_taps, integ, prec in BEAMS:
        h = orc.ppf_coefficients(n_taps, n_ch, "kaiser").astype(np.float32)
        mk = lambda: ppl.PPF(n_sub, n_ch, n_taps, h, precision=prec, integrate=integ)  # noqa: E731
        out.append((mk(), mk(), h))
    return out


def test_batch_equals_single_calls(orc, ppl):
    plans = _plans(orc, ppl)
    rng = np.random.default_rng(11)
    # per call and beam: blocks to channelise (a multiple of the integration;
    # 0 skips the beam) out of a buffer with spare blocks after them
    calls = [[40, 96, 24, 13, 48, 64, 16, 200], [8, 0, 160, 7, 10, 32, 40, 4], [120, 32, 8, 64, 2, 96, 8, 60]]
    whole = [[] for _ in BEAMS]
    for c, blks in enumerate(calls):
        jobs, singles = [], []
        for b, ((n_sub, n_ch, n_taps, integ, prec), (pb, ps, h)) in enumerate(zip(BEAMS, plans)):
            n_blk = blks[b] - blks[b] % integ
            spare = int(rng.integers(0, 9))
            x = torch.from_numpy(noise_voltages(n_sub, (n_blk + spare) * n_ch, 64.0, seed=100 * c + b)).cuda()
            out = torch.full((n_blk // integ, n_sub, 1, n_ch), float("nan"), device="cuda")
            jobs.append((pb, x, n_blk, out))
            if n_blk:
                xs = x[:, :, :n_blk * n_ch].contiguous()
                singles.append((b, ps.stokes_i16(xs, n_blk), xs))
        ppl.ppf_stokes_i16_batch(jobs)
        for b, ref, xs in singles:
            assert torch.equal(jobs[b][3], ref), (c, b)
            whole[b].append(xs)
    # the FP32 jobs' streams against the FP64 chain (all calls in one piece)
    for b, (n_sub, n_ch, n_taps, integ, prec) in enumerate(BEAMS):
        if prec != "fp32":
            continue
        x = torch.cat(whole[b], dim=2).cpu().numpy()
        n_blk = x.shape[2] // n_ch
        h = plans[b][2].astype(np.float64)
        ref = orc.PPF(n_sub, 2, n_ch, n_taps, h).run_f64(x, n_blk)
        I64 = orc.integrate(orc.stokes_i_f64(ref), integ)
        p = ppl.PPF(n_sub, n_ch, n_taps, plans[b][2], integrate=integ)
        got = p.stokes_i16(torch.from_numpy(x).cuda(), n_blk).cpu().numpy()
        assert np.max(np.abs(got - I64)) / np.mean(I64) < 1e-4, b


def test_batch_at_the_batched_workload_shape(orc, ppl):
    """Beams of 31-62 subbands and 16/64/256 channels with chunks of 2^16
    samples per stream: the batched launch fills the GPU with short runs; the
    result equals each plan's own call."""
    shapes = [(62, 64), (31, 16), (45, 256), (52, 16), (38, 64), (60, 256)]
    plans, jobs, refs = [], [], []
    for b, (n_sub, n_ch) in enumerate(shapes):
        h = orc.ppf_coefficients(8, n_ch, "hann").astype(np.float32)
        pb, ps = ppl.PPF(n_sub, n_ch, 8, h, integrate=4), ppl.PPF(n_sub, n_ch, 8, h, integrate=4)
        n_blk = (1 << 16) // n_ch
        gen = torch.Generator(device="cuda").manual_seed(b)
        x = torch.randint(-300, 300, (n_sub, 2, n_blk * n_ch, 2), dtype=torch.int16, device="cuda", generator=gen)
        out = torch.empty((n_blk // 4, n_sub, 1, n_ch), device="cuda")
        plans.append((pb, ps))
        jobs.append((pb, x, n_blk, out))
        refs.append((ps, x, n_blk))
    for rep in range(2):
        ppl.ppf_stokes_i16_batch(jobs)
        for b, (ps, x, n_blk) in enumerate(refs):
            assert torch.equal(jobs[b][3], ps.stokes_i16(x, n_blk)), (rep, b)


def test_batch_refusals(orc, ppl):
    h = orc.ppf_coefficients(8, 64, "kaiser").astype(np.float32)
    p = ppl.PPF(2, 64, 8, h, integrate=4)
    q = ppl.PPF(2, 64, 8, h, integrate=4)
    x = torch.zeros((2, 2, 64 * 64, 2), dtype=torch.int16, device="cuda")
    out = torch.empty((16, 2, 1, 64), device="cuda")
    with pytest.raises(ppl.PPLError, match="again"):
        ppl.ppf_stokes_i16_batch([(p, x, 64, out), (p, x, 64, out)])
    with pytest.raises(ppl.PPLError, match="multiple of the integration"):
        ppl.ppf_stokes_i16_batch([(q, x, 62, out)])
    with pytest.raises(ppl.PPLError, match="blocks from a buffer"):
        ppl.ppf_stokes_i16_batch([(q, x, 128, out)])
    with pytest.raises(ppl.PPLError):
        ppl.ppf_stokes_i16_batch([])
    # nothing ran: both plans still give the same first chunk
    a = p.stokes_i16(x, 64)
    b = q.stokes_i16(x, 64)
    assert torch.equal(a, b)
