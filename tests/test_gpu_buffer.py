"""GPU dedispersion buffer (DedispersionBuffer::addSamples + copy) against
the oracle's buffer: ragged blob sizes, buffers that fill mid-blob, the
overlap carry, Stokes blobs with 4 polarisations (only I is taken), and the
FP16 / u8 conversions on the way in."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _feed(ppl, buf, blobs, use_blob):
    out = []
    for b in blobs:
        d = torch.from_numpy(b).cuda()
        first = 0
        while first < b.shape[0]:
            if use_blob:
                used, full = buf.add_blob(d, first)
            else:
                used, full = buf.add(d.view(b.shape[0], -1), first)
            first += used
            if full:
                out.append(buf.tensor().cpu().numpy().copy())
                buf.rotate()
    return out


@pytest.mark.parametrize("n_pol,use_blob", [(1, False), (1, True), (4, True)])
def test_buffer_matches_oracle(orc, ppl, n_pol, use_blob):
    n_sub, n_ch, nsamp, overlap = 6, 16, 1000, 237
    rng = np.random.default_rng(3)
    sizes = [1, 63, 64, 65, 700, 1500, 2, 999, 333]
    blobs = [rng.normal(size=(n, n_sub, n_pol, n_ch)).astype(np.float32) for n in sizes]
    buf = ppl.DDBuffer(n_sub * n_ch, nsamp, overlap)
    got = _feed(ppl, buf, blobs, use_blob)
    ob = orc.DDBuffer(n_sub * n_ch, nsamp, overlap)
    want = []
    for b in blobs:
        want.extend(list(ob.add(np.ascontiguousarray(b[:, :, 0, :]).reshape(b.shape[0], -1))))
    assert len(got) == len(want) >= 4
    for k, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(g, w), k


@pytest.mark.parametrize("dtype", ["f16", "u8"])
def test_buffer_converts_exactly(orc, ppl, dtype):
    n_chan, nsamp, overlap, inv = 96, 512, 100, 127.0 / 10.0
    rng = np.random.default_rng(4)
    blobs = [(rng.normal(size=(n, n_chan)) * 4).astype(np.float32) for n in (300, 513, 77, 600)]
    buf = ppl.DDBuffer(n_chan, nsamp, overlap, dtype=dtype, inv_scale=inv)
    got = []
    for b in blobs:
        d = torch.from_numpy(b).cuda()
        first = 0
        while first < b.shape[0]:
            used, full = buf.add(d, first)
            first += used
            if full:
                got.append(buf.tensor().cpu().numpy().copy())
                buf.rotate()
    ob = orc.DDBuffer(n_chan, nsamp, overlap)
    want = []
    for b in blobs:
        want.extend(list(ob.add(b)))
    assert len(got) == len(want) >= 2
    for g, w in zip(got, want):
        if dtype == "f16":
            assert np.array_equal(g.view(np.uint16), orc.f32_to_f16_bits(w))
        else:
            assert np.array_equal(g, orc.quantise_u8(w, inv))


def test_buffer_rejects_mismatched_blob(ppl):
    buf = ppl.DDBuffer(96, 512, 100)
    with pytest.raises(ppl.PPLError):
        buf.add_blob(torch.zeros((10, 5, 1, 16), device="cuda"))  # 80 channels, buffer has 96


def test_u8_exceptional_rows_match_oracle(orc, ppl):
    """u8 buffers: spectra most of whose samples fall outside the 8-bit range
    (the clipper's rejected spectra, thousands of units below the data) are
    stored relative to their middle channel's value, which is kept per row
    and carried with the overlap; a spectrum with fewer saturated samples
    than half is stored as it stands.  Codes and offsets equal the oracle's,
    buffer after buffer."""
    n_sub, n_ch, nsamp, overlap, inv = 6, 16, 512, 100, 127.0 / 10.0
    n_chan = n_sub * n_ch
    rng = np.random.default_rng(44)
    blobs = []
    for n in (300, 513, 77, 600):
        b = rng.normal(size=(n, n_chan)).astype(np.float32)
        rej = rng.choice(n, size=max(2, n // 40), replace=False)
        b[rej] = (-3000.0 - 40.0 * rng.random(size=(len(rej), n_chan))).astype(np.float32)
        b[rng.integers(n), : n_chan // 3] = 500.0  # a third saturated: not exceptional
        blobs.append(b)
    buf = ppl.DDBuffer(n_chan, nsamp, overlap, dtype="u8", inv_scale=inv)
    got, offs = [], []
    for b in blobs:
        d = torch.from_numpy(b).cuda()
        first = 0
        while first < b.shape[0]:
            used, full = buf.add(d, first)
            first += used
            if full:
                got.append(buf.tensor().cpu().numpy().copy())
                offs.append(buf.row_offsets().cpu().numpy().copy())
                tail = offs[-1][nsamp - overlap:].copy()
                buf.rotate()
                assert np.array_equal(buf.row_offsets()[:overlap].cpu().numpy(), tail)
    ob = orc.DDBuffer(n_chan, nsamp, overlap)
    want = []
    for b in blobs:
        want.extend(list(ob.add(b)))
    assert len(got) == len(want) >= 2
    n_exc = 0
    for g, o, w in zip(got, offs, want):
        q, off = orc.quantise_u8_rows(w, inv)
        assert np.array_equal(o, off)
        assert np.array_equal(g, q)
        n_exc += int(np.count_nonzero(off))
    assert n_exc >= 8
