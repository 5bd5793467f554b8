"""Packet ingest (AdapterTimeSeriesDataSet::_readData 8/16-bit branches):
the device transpose of [sub][t][pol] packet payloads into the blob layout
against a numpy restatement, with headers between payloads, and fed straight
into the channeliser."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _packets(rng, n_packets, n_sub, n_spp, n_pol, bits, header):
    dt = np.int16 if bits == 16 else np.int8
    lo, hi = (-32768, 32768) if bits == 16 else (-128, 128)
    payload = rng.integers(lo, hi, size=(n_packets, n_sub, n_spp, n_pol, 2)).astype(dt)
    raw = np.zeros((n_packets, header + payload[0].nbytes), np.uint8)
    raw[:, :header] = rng.integers(0, 256, size=(n_packets, header))
    raw[:, header:] = payload.reshape(n_packets, -1).view(np.uint8)
    return raw, payload


def _expected(payload):
    n_packets, n_sub, n_spp, n_pol, _ = payload.shape
    # time = packet * n_spp + t ; blob [sub][pol][time][re,im]
    return payload.transpose(1, 3, 0, 2, 4).reshape(n_sub, n_pol, n_packets * n_spp, 2)


@pytest.mark.parametrize("bits,header", [(16, 16), (8, 16), (16, 0)])
def test_ingest_layouts_exact(ppl, bits, header):
    rng = np.random.default_rng(bits + header)
    raw, payload = _packets(rng, 37, 61, 16, 2, bits, header)
    d = torch.from_numpy(raw).cuda()
    want = _expected(payload)
    c32 = ppl.ingest(d, 37, 61, 16, 2, bits, header, out="c32").cpu().numpy()
    assert np.array_equal(c32, want.astype(np.float32))
    i16 = ppl.ingest(d, 37, 61, 16, 2, bits, header, out="i16").cpu().numpy()
    assert np.array_equal(i16, want.astype(np.int16))


def test_ingest_feeds_channeliser(orc, ppl):
    n_sub, n_ch, n_taps, n_packets, n_spp = 4, 16, 8, 64, 16
    n_blk = n_packets * n_spp // n_ch
    rng = np.random.default_rng(9)
    raw, payload = _packets(rng, n_packets, n_sub, n_spp, 2, 16, 16)
    coeffs = orc.ppf_coefficients(n_taps, n_ch, "hann").astype(np.float32)
    d = torch.from_numpy(raw).cuda()
    direct = torch.from_numpy(np.ascontiguousarray(_expected(payload).astype(np.int16))).cuda()
    ref = ppl.PPF(n_sub, n_ch, n_taps, coeffs).spectra_i16(direct, n_blk).cpu().numpy()
    volt = ppl.ingest(d, n_packets, n_sub, n_spp, 2, 16, 16, out="i16")
    got_i16 = ppl.PPF(n_sub, n_ch, n_taps, coeffs).spectra_i16(volt, n_blk).cpu().numpy()
    blob = ppl.ingest(d, n_packets, n_sub, n_spp, 2, 16, 16, out="c32")
    got_c32 = ppl.PPF(n_sub, n_ch, n_taps, coeffs).spectra_c32(blob, n_blk).cpu().numpy()
    assert np.array_equal(got_i16, ref) and np.array_equal(got_c32, ref)


def test_ingest_rejects_bad_bits(ppl):
    d = torch.zeros(1024, dtype=torch.uint8, device="cuda")
    with pytest.raises(ppl.PPLError):
        ppl.ingest(d, 1, 2, 4, 2, sample_bits=12)


@pytest.mark.parametrize("bits", [16, 8])
def test_ingest_slots_with_lost_packets(ppl, bits):
    """A received stream with gaps, duplicates and bad timestamps, placed in
    chunk slots by the host sequencer (LofarDataSplittingChunker), then
    ingested slot by slot: lost slots are zeros (the chunker's empty
    packets), the rest the packets' payloads, as the oracle lays them out."""
    import packets as orc_packets
    from test_packets import N_SPP, _random_stream
    rng = np.random.default_rng(bits)
    stream = _random_stream(rng, 120)
    n_sub, n_pol, header = 31, 2, 16
    raw, payload = _packets(rng, len(stream), n_sub, N_SPP, n_pol, bits, header)
    for k, (s, b) in enumerate(stream):
        raw[k, :16] = np.frombuffer(orc_packets.header(s, b), np.uint8)
    st = ppl.PacketSeq()
    slots, used = ppl.sequence_packets(st, raw, N_SPP, 100)
    st_o = dict(startTime=0, startBlockid=0, accepted=0, rejected=0, lost=0)
    want_slots, _ = orc_packets.sequence(st_o, stream, N_SPP, 100)
    assert list(slots) == want_slots and (slots < 0).any()
    d = torch.from_numpy(raw).cuda()
    d_slots = torch.from_numpy(slots.astype(np.int32)).cuda()
    want = orc_packets.chunk(payload, want_slots)
    c32 = ppl.ingest(d, len(stream), n_sub, N_SPP, n_pol, bits, header, out="c32", slots=d_slots).cpu().numpy()
    assert np.array_equal(c32, want.astype(np.float32))
    i16 = ppl.ingest(d, len(stream), n_sub, N_SPP, n_pol, bits, header, out="i16", slots=d_slots).cpu().numpy()
    assert np.array_equal(i16, want.astype(np.int16))
    # an index past the packets given reads as a lost packet
    bad = torch.tensor([0, len(stream), -1], dtype=torch.int32, device="cuda")
    out = ppl.ingest(d, len(stream), n_sub, N_SPP, n_pol, bits, header, out="i16", slots=bad).cpu().numpy()
    assert np.array_equal(out[:, :, :N_SPP], payload[0].transpose(0, 2, 1, 3).astype(np.int16))
    assert not out[:, :, N_SPP:].any()
