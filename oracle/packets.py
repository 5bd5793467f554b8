"""CPU restatement of LofarDataSplittingChunker's packet sequencing
(src/lib/src/LofarDataSplittingChunker.cpp:186-325) and of the chunk it
writes, in plain Python loops (small cases).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
never imports this module.
"""
import struct

import numpy as np

U32 = 0xFFFFFFFF


def header(seqid, blockid, n_beamlets=0, n_blocks=16):
    """A 16-byte LOFAR UDP header (src/lib/LofarUdpHeader.h:25-36): version,
    sourceInfo, configuration, station, nrBeamlets, nrBlocks, timestamp,
    blockSequenceNumber, little endian."""
    return struct.pack("<BBHHBBII", 3, 0, 0, 0, n_beamlets, n_blocks, seqid & U32, blockid & U32)


def sequence(state, seq_block, n_spp, n_slots, clock_mhz=200):
    """state: dict(startTime, startBlockid, accepted, rejected, lost), updated.
    seq_block: list of (timestamp, blockSequenceNumber) in arrival order.
    Returns (slot sources, packets used): slot s holds packet index or -1."""
    prev_seq, prev_blk = state["startTime"], state["startBlockid"]
    slots = []
    k = 0
    while len(slots) < n_slots and k < len(seq_block):
        seqid, blockid = seq_block[k]
        cur = k
        k += 1
        i = len(slots)
        if i == 0 and state["startTime"] == 0:                      # :212-216
            prev_seq = state["startTime"] = seqid
            if state["startBlockid"] == 0:
                state["startBlockid"] = blockid
            prev_blk = state["startBlockid"]
        if seqid == U32 or ((prev_seq + 10) & U32) < seqid:         # :218-225
            state["rejected"] += 1
            continue
        tot = 156250 if clock_mhz == 160 else (195313 if prev_seq % 2 == 0 else 195312)
        diff = blockid - prev_blk if blockid >= prev_blk else (blockid + tot - prev_blk) & U32
        lost = 0
        if diff < n_spp:                                            # duplicate
            state["rejected"] += 1
            continue
        if diff > n_spp:
            lost = diff // n_spp - 1
        pc = 0
        while pc < lost and i + pc < n_slots:                        # :264-284
            prev_seq = prev_seq if prev_blk + n_spp < tot else (prev_seq + 1) & U32
            prev_blk = (prev_blk + n_spp) % tot
            slots.append(-1)
            state["lost"] += 1
            pc += 1
        if len(slots) != n_slots:                                   # :287-305
            slots.append(cur)
            state["accepted"] += 1
            prev_seq, prev_blk = seqid, blockid
    state["startTime"], state["startBlockid"] = prev_seq, prev_blk
    return slots, k


def chunk(payload, slots):
    """The chunk's samples [sub][pol][slot*n_spp + t][re, im] from packet
    payloads [packet][sub][t][pol][re, im]; lost slots are zero (the
    chunker's empty packets)."""
    n_sub, n_spp, n_pol = payload.shape[1:4]
    out = np.zeros((n_sub, n_pol, len(slots) * n_spp, 2), payload.dtype)
    for s, src in enumerate(slots):
        if src >= 0:
            out[:, :, s * n_spp:(s + 1) * n_spp, :] = payload[src].transpose(0, 2, 1, 3)
    return out
