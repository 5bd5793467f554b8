"""Chain steps replayed from HIP graphs against the same chain run eagerly,
at the CPU-runnable configuration: every replayed buffer's dedispersed
series and event list equal the eager ones bit for bit, with eager work and
synchronisations between the replays."""
import numpy as np
import pytest
import torch

from synth import noise_voltages
from test_gpu_chain import DM_STEP, FCH1, FOFF, INTEG, N_CH, N_SUB, N_TAPS, NSAMP_BUF, TDMS, TSAMP, _inputs

pytestmark = pytest.mark.gpu


def _chain(ppl, coeffs, bp):
    return ppl.BeamChain(N_SUB, N_CH, N_TAPS, coeffs, INTEG, bp, NSAMP_BUF, TDMS, 0.0, DM_STEP, TSAMP, FCH1,
                         FOFF, history=1000)


def test_graph_replays_match_eager(orc, ppl):
    coeffs, _, bp = _inputs(orc)
    eager, graph = _chain(ppl, coeffs, bp), _chain(ppl, coeffs, bp)
    per = eager.samples_per_step()
    n_prime = eager.overlap * INTEG * N_CH
    n_steps = 8
    x = noise_voltages(N_SUB, n_prime + n_steps * per, 64.0, seed=31)
    xd = torch.from_numpy(x).cuda()
    prime = xd[:, :, :n_prime].contiguous()
    eager.prime(prime)
    graph.prime(prime)
    staging = torch.empty((N_SUB, 2, per, 2), dtype=torch.int16, device="cuda")
    for k in range(n_steps):
        chunk = xd[:, :, n_prime + k * per:n_prime + (k + 1) * per].contiguous()
        eager.step(chunk)
        want = eager.out.cpu().numpy().copy()
        want_ev = eager.detect()[:2]
        staging.copy_(chunk)
        if k < 2:
            graph.step(staging)           # eager warm-up (allocations) before capture
        else:
            graph.capture_steps(staging)  # no-op after the first capture
            graph.replay()
        torch.cuda.synchronize()
        got = graph.out.cpu().numpy()
        assert np.array_equal(got, want), k
        got_ev = graph.detect()[:2]
        assert np.array_equal(got_ev[0], want_ev[0]) and np.array_equal(got_ev[1], want_ev[1]), k


@pytest.mark.parametrize("dtype", ["f32", "u8"])
def test_pipelined_steps_match_eager(orc, ppl, dtype):
    """The two-stream schedule (front: PPF, clipper, buffer, carry; back:
    dedispersion, detection; step k's front overlapping step k-1's
    dedispersion) gives every step's events bit for bit as step() does, and
    leaves the last series in out."""
    coeffs, _, bp = _inputs(orc)
    mk = lambda: ppl.BeamChain(N_SUB, N_CH, N_TAPS, coeffs, INTEG, bp, NSAMP_BUF, TDMS, 0.0, DM_STEP, TSAMP,
                               FCH1, FOFF, history=1000, dtype=dtype)
    eager, piped = mk(), mk()
    per = eager.samples_per_step()
    n_prime = eager.overlap * INTEG * N_CH
    n_steps = 7
    x = noise_voltages(N_SUB, n_prime + n_steps * per, 64.0, seed=37)
    xd = torch.from_numpy(x).cuda()
    prime = xd[:, :, :n_prime].contiguous()
    eager.prime(prime)
    piped.prime(prime)
    chunks = [xd[:, :, n_prime + k * per:n_prime + (k + 1) * per].contiguous() for k in range(n_steps)]
    want = []
    for c in chunks:
        eager.step(c)
        want.append(eager.detect()[:2])
    last = eager.out.cpu().numpy().copy()
    got = [r[:2] for r in piped.run_pipelined(chunks)]
    torch.cuda.synchronize()
    assert len(got) == n_steps
    for k, ((gi, gv), (wi, wv)) in enumerate(zip(got, want)):
        assert np.array_equal(gi, wi) and np.array_equal(gv, wv), k
    assert np.array_equal(piped.out.cpu().numpy(), last)
    # and the chain carries on with step() after a pipelined run
    tail = noise_voltages(N_SUB, per, 64.0, seed=38)
    tc = torch.from_numpy(tail).cuda()
    eager.step(tc)
    piped.step(tc)
    assert np.array_equal(piped.out.cpu().numpy(), eager.out.cpu().numpy())


@pytest.mark.parametrize("dtype", ["f32", "u8"])
def test_cu_split_steps_match_eager(orc, ppl, dtype):
    """The CU-split schedule (clipper, buffer fill and carry on 4 CU slots of
    every XCD, the rest of the chain on the others; step k's clipper beside
    step k-1's dedispersion) gives every step's events bit for bit as step()
    does, leaves the last series in out, and the chain carries on with
    step() afterwards."""
    coeffs, _, bp = _inputs(orc)
    mk = lambda: ppl.BeamChain(N_SUB, N_CH, N_TAPS, coeffs, INTEG, bp, NSAMP_BUF, TDMS, 0.0, DM_STEP, TSAMP,
                               FCH1, FOFF, history=1000, dtype=dtype)
    eager, split = mk(), mk()
    per = eager.samples_per_step()
    n_prime = eager.overlap * INTEG * N_CH
    n_steps = 6
    x = noise_voltages(N_SUB, n_prime + n_steps * per, 64.0, seed=41)
    xd = torch.from_numpy(x).cuda()
    prime = xd[:, :, :n_prime].contiguous()
    eager.prime(prime)
    split.prime(prime)
    chunks = [xd[:, :, n_prime + k * per:n_prime + (k + 1) * per].contiguous() for k in range(n_steps)]
    want = []
    for c in chunks:
        eager.step(c)
        want.append(eager.detect()[:2])
    last = eager.out.cpu().numpy().copy()
    got = [r[:2] for r in split.run_split(chunks, small_cus=4)]
    torch.cuda.synchronize()
    assert len(got) == n_steps
    for k, ((gi, gv), (wi, wv)) in enumerate(zip(got, want)):
        assert np.array_equal(gi, wi) and np.array_equal(gv, wv), k
    assert np.array_equal(split.out.cpu().numpy(), last)
    tail = noise_voltages(N_SUB, per, 64.0, seed=42)
    tc = torch.from_numpy(tail).cuda()
    eager.step(tc)
    split.step(tc)
    assert np.array_equal(split.out.cpu().numpy(), eager.out.cpu().numpy())


@pytest.mark.parametrize("dtype", ["f32", "u8"])
def test_overlapped_steps_match_eager(orc, ppl, dtype):
    """The overlapped schedule (step k's clipper walk on CU slot 0 of every XCD
    beside step k-1's dedispersion on CU slots 1.. of every XCD) gives every
    step's events bit for bit as step() does, leaves the last series in out,
    and the chain carries on with step() afterwards."""
    coeffs, _, bp = _inputs(orc)
    mk = lambda: ppl.BeamChain(N_SUB, N_CH, N_TAPS, coeffs, INTEG, bp, NSAMP_BUF, TDMS, 0.0, DM_STEP, TSAMP,
                               FCH1, FOFF, history=1000, dtype=dtype)
    eager, ov = mk(), mk()
    per = eager.samples_per_step()
    n_prime = eager.overlap * INTEG * N_CH
    n_steps = 7
    x = noise_voltages(N_SUB, n_prime + n_steps * per, 64.0, seed=43)
    xd = torch.from_numpy(x).cuda()
    prime = xd[:, :, :n_prime].contiguous()
    eager.prime(prime)
    ov.prime(prime)
    chunks = [xd[:, :, n_prime + k * per:n_prime + (k + 1) * per].contiguous() for k in range(n_steps)]
    want = []
    for c in chunks:
        eager.step(c)
        want.append(eager.detect()[:2])
    last = eager.out.cpu().numpy().copy()
    got = [r[:2] for r in ov.run_overlapped(chunks)]
    torch.cuda.synchronize()
    assert len(got) == n_steps
    for k, ((gi, gv), (wi, wv)) in enumerate(zip(got, want)):
        assert np.array_equal(gi, wi) and np.array_equal(gv, wv), k
    assert np.array_equal(ov.out.cpu().numpy(), last)
    assert ov.clip.stats() == eager.clip.stats()
    tail = noise_voltages(N_SUB, per, 64.0, seed=44)
    tc = torch.from_numpy(tail).cuda()
    eager.step(tc)
    ov.step(tc)
    assert np.array_equal(ov.out.cpu().numpy(), eager.out.cpu().numpy())


def test_cu_range_stream_rejects_bad_ranges(ppl):
    """CU ranges outside the slots of an XCD, or empty, are refused."""
    per_xcd = torch.cuda.get_device_properties(0).multi_processor_count // 8
    for first, count in [(-1, 2), (0, 0), (per_xcd - 1, 2), (0, per_xcd + 1)]:
        with pytest.raises(ppl.PPLError):
            ppl.CuRangeStream(first, count)
    s = ppl.CuRangeStream(0, per_xcd)  # the whole GPU is a valid range
    x = torch.ones(1024, device="cuda")
    with torch.cuda.stream(s.stream):
        y = x * 2
    s.stream.synchronize()
    assert float(y.sum()) == 2048.0
    s.close()
