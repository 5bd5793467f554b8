"""ppl_detect_noalloc, the entry point that allocates nothing, against
ppl_detect (the allocating entry point it varies): the same events, values
and count, bit for bit, on two calls in a row -- the second on a smaller
input, so stale scratch from the first call would show -- and a refused
size beyond its static scratch.  Called through the C ABI."""
import ctypes as C

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


def _noalloc(ppl, d, nch, nsub, threshold, bin_pow2, max_events):
    L = ppl.lib()
    ints = torch.full((max_events, 3), -7, dtype=torch.int32, device="cuda")
    vals = torch.full((max_events,), -7.0, dtype=torch.float32, device="cuda")
    count = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    tdms, nsamp = d.shape
    rc = L.ppl_detect_noalloc(C.c_void_p(d.data_ptr()), tdms, nsamp, nch, nsub, threshold, bin_pow2,
                              C.c_void_p(ints.data_ptr()), C.c_void_p(vals.data_ptr()), max_events,
                              C.c_void_p(count.data_ptr()), None)
    assert rc == 0, L.ppl_last_error()
    torch.cuda.synchronize()
    n = int(count.item())
    m = min(n, max_events)
    return ints[:m].cpu().numpy(), vals[:m].cpu().numpy(), n


@pytest.mark.parametrize("threshold,bin_pow2", [(3.0, 6), (6.0, 4)])
def test_noalloc_equals_detect_on_two_calls(ppl, threshold, bin_pow2):
    big = _data(512, 64 * 700 + 11, 16, 8, 41, [(100, 5000, 40, 12), (300, 20000, 3, 30)])
    small = _data(96, 64 * 90 + 5, 16, 8, 42, [(7, 300, 9, 20)])
    for x in (big, small):
        d = torch.from_numpy(x).cuda()
        ev, val, n = ppl.detect(d, 16, 8, threshold, bin_pow2, max_events=1 << 21)
        ev_n, val_n, n_n = _noalloc(ppl, d, 16, 8, threshold, bin_pow2, 1 << 21)
        assert n_n == n > 1
        assert np.array_equal(ev_n, ev) and np.array_equal(val_n, val)
        # a truncated event buffer keeps the first events and the full count
        ev_t, val_t, n_t = _noalloc(ppl, d, 16, 8, threshold, bin_pow2, 5)
        assert n_t == n and np.array_equal(ev_t, ev[:5]) and np.array_equal(val_t, val[:5])


def test_noalloc_refuses_more_segments_than_its_scratch(ppl):
    L = ppl.lib()
    limit = ppl.DETECT_NOALLOC_MAX_SEGMENTS
    d = torch.zeros((1, 64), device="cuda")
    out_i = torch.zeros((4, 3), dtype=torch.int32, device="cuda")
    out_v = torch.zeros((4,), device="cuda")
    cnt = torch.zeros((1,), dtype=torch.int32, device="cuda")
    # tdms * (nsamp >> 6) one past the limit: refused before any launch
    tdms, nsamp = limit // 1024 + 1, 1024 * 64
    rc = L.ppl_detect_noalloc(C.c_void_p(d.data_ptr()), tdms, nsamp, 16, 8, 6.0, 6,
                              C.c_void_p(out_i.data_ptr()), C.c_void_p(out_v.data_ptr()), 4,
                              C.c_void_p(cnt.data_ptr()), None)
    assert rc != 0 and b"no-allocation scratch" in L.ppl_last_error()
