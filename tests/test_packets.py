"""Packet sequencing (LofarDataSplittingChunker::next,
src/lib/src/LofarDataSplittingChunker.cpp:186-325): the oracle's
restatement on hand-built arrival sequences whose chunk layout is worked out
by hand (a stream's first packet is always dropped as a duplicate of
itself; gaps become empty slots; duplicates and untrustworthy timestamps are
rejected; block numbers wrap at 195313/195312 blocks per second at 200 MHz
and 156250 at 160 MHz; a packet after gap slots that filled the chunk is
dropped), then the product's host sequencer (ppl_packet_sequence, in the
C-ABI library, no GPU needed) against the oracle on random streams."""
import numpy as np
import pytest

import packets as orc_packets
from conftest import load_ppl

N_SPP = 16


def _new():
    return dict(startTime=0, startBlockid=0, accepted=0, rejected=0, lost=0)


def _raw(seq_block, n_bytes=64):
    raw = np.zeros((len(seq_block), n_bytes), np.uint8)
    for k, (s, b) in enumerate(seq_block):
        raw[k, :16] = np.frombuffer(orc_packets.header(s, b), np.uint8)
    return raw


def test_oracle_contiguous_stream_drops_its_first_packet():
    st = _new()
    slots, used = orc_packets.sequence(st, [(100, 16 * k) for k in range(6)], N_SPP, 5)
    assert slots == [1, 2, 3, 4, 5] and used == 6
    assert (st["accepted"], st["rejected"], st["lost"]) == (5, 1, 0)
    assert (st["startTime"], st["startBlockid"]) == (100, 80)


def test_oracle_gap_duplicate_and_bad_timestamps():
    st = _new()
    seq = [(100, 0), (100, 16), (100, 32), (100, 32), (0xFFFFFFFF, 48), (200, 48), (100, 80), (100, 96)]
    slots, used = orc_packets.sequence(st, seq, N_SPP, 6)
    # 0: first (dropped); 1, 2 kept; 3 duplicate; 4 bad timestamp; 5 more
    # than 10 s ahead; 6 after a gap of two packets (blocks 48, 64)
    assert slots == [1, 2, -1, -1, 6, 7] and used == 8
    assert (st["accepted"], st["rejected"], st["lost"]) == (4, 4, 2)


def test_oracle_block_wrap_200_and_160_mhz():
    # 200 MHz: an even second has 195313 blocks, an odd one 195312
    st = dict(_new(), startTime=100, startBlockid=195297)
    slots, _ = orc_packets.sequence(st, [(101, 0), (101, 16)], N_SPP, 2)
    assert slots == [0, 1] and st["lost"] == 0
    st = dict(_new(), startTime=101, startBlockid=195300)
    slots, _ = orc_packets.sequence(st, [(102, 4)], N_SPP, 1)
    assert slots == [0]                     # 195312 - 195300 + 4 = 16
    st = dict(_new(), startTime=7, startBlockid=156240)
    slots, _ = orc_packets.sequence(st, [(8, 6)], N_SPP, 1, clock_mhz=160)
    assert slots == [0]                     # 156250 - 156240 + 6 = 16
    # a gap across the wrap: the empty packets' seconds advance as they wrap
    st = dict(_new(), startTime=100, startBlockid=195281)
    slots, _ = orc_packets.sequence(st, [(101, 16)], N_SPP, 3)
    assert slots == [-1, -1, 0] and (st["startTime"], st["startBlockid"]) == (101, 16)


def test_oracle_gap_filling_the_chunk_drops_the_packet():
    st = dict(_new(), startTime=100, startBlockid=32)
    slots, used = orc_packets.sequence(st, [(100, 48), (100, 128)], N_SPP, 3)
    # 48 fills slot 0; 128 is 5 packets on: 4 empty slots, of which 2 fit
    assert slots == [0, -1, -1] and used == 2
    assert st["accepted"] == 1 and st["lost"] == 2
    assert (st["startTime"], st["startBlockid"]) == (100, 80)
    # the next chunk continues from block 80: 128 is now 3 packets on
    slots, _ = orc_packets.sequence(st, [(100, 128)], N_SPP, 3)
    assert slots == [-1, -1, 0]


def test_oracle_chunk_layout():
    payload = np.arange(3 * 2 * 4 * 2 * 2, dtype=np.int16).reshape(3, 2, 4, 2, 2)  # [pkt][sub][t][pol][ri]
    out = orc_packets.chunk(payload, [2, -1, 0])
    assert out.shape == (2, 2, 12, 2)
    assert np.array_equal(out[1, 0, 0:4], payload[2, 1, :, 0])
    assert not out[:, :, 4:8].any()
    assert np.array_equal(out[0, 1, 8:12], payload[0, 0, :, 1])


def _random_stream(rng, n, start_seq=1000):
    seq, blk, out = start_seq, int(rng.integers(0, 195000)), []
    for _ in range(n):
        r = rng.random()
        if r < 0.05:
            out.append((0xFFFFFFFF, blk))           # untrustworthy
            continue
        if r < 0.10 and out:
            out.append(out[-1])                     # duplicate
            continue
        step = 1 if r < 0.85 else int(rng.integers(2, 6))   # gaps of 1-4 packets
        for _ in range(step):
            tot = 195313 if seq % 2 == 0 else 195312
            if blk + N_SPP >= tot:
                seq += 1
            blk = (blk + N_SPP) % tot
        out.append((seq, blk))
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_product_sequencer_matches_oracle(seed):
    ppl = load_ppl()
    rng = np.random.default_rng(seed)
    stream = _random_stream(rng, 400)
    raw = _raw(stream)
    st_o, st_p = _new(), ppl.PacketSeq()
    pos = 0
    for n_slots in (37, 64, 5, 100, 64, 64, 64):
        want, used_o = orc_packets.sequence(st_o, stream[pos:], N_SPP, n_slots)
        got, used_p = ppl.sequence_packets(st_p, raw[pos:], N_SPP, n_slots)
        assert list(got) == want and used_p == used_o
        assert (st_p.startTime, st_p.startBlockid) == (st_o["startTime"], st_o["startBlockid"])
        assert (st_p.accepted, st_p.rejected, st_p.lost) == (st_o["accepted"], st_o["rejected"], st_o["lost"])
        pos += used_o
        if pos >= len(stream):
            break


def test_product_sequencer_argument_errors():
    ppl = load_ppl()
    st = ppl.PacketSeq()
    with pytest.raises(ppl.PPLError):
        ppl.sequence_packets(st, _raw([(1, 0)]), N_SPP, 4, clock_mhz=100)
    with pytest.raises(ppl.PPLError):
        ppl.sequence_packets(st, np.zeros((2, 8), np.uint8), N_SPP, 4)   # no room for a header
