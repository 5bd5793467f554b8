// ppl_ingest.hip -- packet payloads to the channeliser's input layouts
// (AdapterTimeSeriesDataSet::_readData, src/lib/src/AdapterTimeSeriesDataSet.cpp:
// 217-293).
//
// A packet carries nSamplesPerPacket time samples of every subband, samples
// [sub][t][pol] as complex 8- or 16-bit integers (re, im).  Packet k covers
// stream samples k*nSamplesPerPacket ... of the chunk.  The complex-float
// output is the TimeSeriesDataSetC32 layout [sub][pol][block][t] with blocks
// of nSamplesPerTimeBlock samples; the int16 output is the fused
// channeliser's [sub][pol][t] layout (the same order with blocks
// concatenated).  One thread per (sub, t, pol) complex sample, reads
// coalesced along the packet.
#include "ppl_internal.h"

namespace {

using namespace ppl;

template <typename Tin, typename Tout>
__global__ void ingest_kernel(const uint8_t* __restrict__ packets, long packetBytes, int headerBytes, int nSlots,
                              const int* __restrict__ slotSource, int nPackets, int nSub, int nSpp, int nPol,
                              Tout* __restrict__ out, long nSampOut)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long perPacket = (long)nSub * nSpp * nPol;
    if (i >= perPacket * nSlots) return;
    const int k = (int)(i / perPacket);
    const long r = i % perPacket;
    const int s = (int)(r / ((long)nSpp * nPol));
    const int t = (int)((r / nPol) % nSpp);
    const int p = (int)(r % nPol);
    // slot k holds packet slotSource[k], or zeros for a lost packet (< 0);
    // an index past the packets given reads as lost too
    const int src = slotSource ? slotSource[k] : k;
    Tout re = (Tout)0, im = (Tout)0;
    if (src >= 0 && src < nPackets) {
        const Tin* q = reinterpret_cast<const Tin*>(packets + (long)src * packetBytes + headerBytes) + 2 * r;
        re = (Tout)q[0];
        im = (Tout)q[1];
    }
    const long time = (long)k * nSpp + t;
    // [sub][pol][time]: the block layout [sub][pol][block][t] is the same
    // order with time = block * nSamplesPerTimeBlock + t
    const long o = ((long)s * nPol + p) * nSampOut + time;
    out[2 * o] = re;
    out[2 * o + 1] = im;
}

template <typename Tout>
int launch(const void* d_packets, size_t packetBytes, unsigned headerBytes, unsigned nSlots, const int* d_slotSource,
           unsigned nPackets, unsigned nSub, unsigned nSpp, unsigned nPol, int sampleBits, Tout* d_out, hipStream_t st)
{
    const long n = (long)nSub * nSpp * nPol * nSlots;
    if (n == 0) return PPL_OK;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    const long nSampOut = (long)nSpp * nSlots;
    if (sampleBits == 16)
        hipLaunchKernelGGL((ingest_kernel<int16_t, Tout>), dim3(blocks), dim3(256), 0, st,
                           static_cast<const uint8_t*>(d_packets), (long)packetBytes, (int)headerBytes, (int)nSlots,
                           d_slotSource, (int)nPackets, (int)nSub, (int)nSpp, (int)nPol, d_out, nSampOut);
    else
        hipLaunchKernelGGL((ingest_kernel<int8_t, Tout>), dim3(blocks), dim3(256), 0, st,
                           static_cast<const uint8_t*>(d_packets), (long)packetBytes, (int)headerBytes, (int)nSlots,
                           d_slotSource, (int)nPackets, (int)nSub, (int)nSpp, (int)nPol, d_out, nSampOut);
    PPL_LAUNCHED();
    return PPL_OK;
}

int check(const void* d_packets, size_t packetBytes, unsigned headerBytes, unsigned nSub, unsigned nSpp,
          unsigned nPol, int sampleBits, const void* d_out)
{
    PPL_CHECK_ARG(d_packets && d_out, "null argument");
    PPL_CHECK_ARG(sampleBits == 8 || sampleBits == 16, "Bits per sample (%d) unsupported", sampleBits);
    PPL_CHECK_ARG(nSub > 0 && nSpp > 0 && nPol > 0, "empty packet shape");
    const size_t payload = (size_t)nSub * nSpp * nPol * 2 * (sampleBits / 8);
    PPL_CHECK_ARG(packetBytes >= headerBytes + payload, "packet of %zu bytes cannot hold %u header + %zu data bytes",
                  packetBytes, headerBytes, payload);
    PPL_CHECK_ARG(sampleBits == 8 || headerBytes % 2 == 0, "16-bit samples need an even header size");
    PPL_CHECK_ARG(sampleBits == 8 || packetBytes % 2 == 0, "16-bit samples need an even packet size");
    return PPL_OK;
}

}  // namespace

extern "C" int ppl_ingest_c32(const void* d_packets, size_t packetBytes, unsigned headerBytes, unsigned nPackets,
                              unsigned nSub, unsigned nSamplesPerPacket, unsigned nPol, int sampleBits,
                              float* d_blob, void* stream)
{
    int rc = check(d_packets, packetBytes, headerBytes, nSub, nSamplesPerPacket, nPol, sampleBits, d_blob);
    if (rc != PPL_OK) return rc;
    return launch<float>(d_packets, packetBytes, headerBytes, nPackets, nullptr, nPackets, nSub, nSamplesPerPacket,
                         nPol, sampleBits, d_blob, stream_of(stream));
}

extern "C" int ppl_ingest_i16(const void* d_packets, size_t packetBytes, unsigned headerBytes, unsigned nPackets,
                              unsigned nSub, unsigned nSamplesPerPacket, unsigned nPol, int sampleBits,
                              int16_t* d_volt, void* stream)
{
    int rc = check(d_packets, packetBytes, headerBytes, nSub, nSamplesPerPacket, nPol, sampleBits, d_volt);
    if (rc != PPL_OK) return rc;
    return launch<int16_t>(d_packets, packetBytes, headerBytes, nPackets, nullptr, nPackets, nSub, nSamplesPerPacket,
                           nPol, sampleBits, d_volt, stream_of(stream));
}

extern "C" int ppl_ingest_c32_slots(const void* d_packets, size_t packetBytes, unsigned headerBytes,
                                    unsigned nPackets, const int* d_slotSource, unsigned nSlots, unsigned nSub,
                                    unsigned nSamplesPerPacket, unsigned nPol, int sampleBits, float* d_blob, void* stream)
{
    int rc = check(d_packets, packetBytes, headerBytes, nSub, nSamplesPerPacket, nPol, sampleBits, d_blob);
    if (rc != PPL_OK) return rc;
    PPL_CHECK_ARG(d_slotSource || nSlots == 0, "null slot map");
    return launch<float>(d_packets, packetBytes, headerBytes, nSlots, d_slotSource, nPackets, nSub,
                         nSamplesPerPacket, nPol, sampleBits, d_blob, stream_of(stream));
}

extern "C" int ppl_ingest_i16_slots(const void* d_packets, size_t packetBytes, unsigned headerBytes,
                                    unsigned nPackets, const int* d_slotSource, unsigned nSlots, unsigned nSub,
                                    unsigned nSamplesPerPacket, unsigned nPol, int sampleBits, int16_t* d_volt, void* stream)
{
    int rc = check(d_packets, packetBytes, headerBytes, nSub, nSamplesPerPacket, nPol, sampleBits, d_volt);
    if (rc != PPL_OK) return rc;
    PPL_CHECK_ARG(d_slotSource || nSlots == 0, "null slot map");
    return launch<int16_t>(d_packets, packetBytes, headerBytes, nSlots, d_slotSource, nPackets, nSub,
                           nSamplesPerPacket, nPol, sampleBits, d_volt, stream_of(stream));
}

// ---- packet sequencing (host code): LofarDataSplittingChunker::next
// (src/lib/src/LofarDataSplittingChunker.cpp:186-325) over packets already
// received, in arrival order, each starting with the 16-byte LOFAR UDP
// header (src/lib/LofarUdpHeader.h:25-36: timestamp at byte 8,
// blockSequenceNumber at byte 12, little endian).
extern "C" int ppl_packet_sequence(ppl_packet_seq* st, const void* h_packets, size_t packetBytes, unsigned nPackets,
                                   unsigned nSamplesPerPacket, int clockMHz, int* h_slotSource, unsigned nSlots,
                                   unsigned* packetsUsed, unsigned* slotsFilled)
{
    PPL_CHECK_ARG(st && packetsUsed && slotsFilled && (h_packets || nPackets == 0) && (h_slotSource || nSlots == 0),
                  "null argument");
    PPL_CHECK_ARG(packetBytes >= 16, "a packet of %zu bytes has no LOFAR header", packetBytes);
    PPL_CHECK_ARG(nSamplesPerPacket > 0, "no samples per packet");
    PPL_CHECK_ARG(clockMHz == 160 || clockMHz == 200, "clock %d MHz (160 or 200)", clockMHz);
    const unsigned char* base = static_cast<const unsigned char*>(h_packets);
    auto word = [&](unsigned k, int off) {
        const unsigned char* b = base + (size_t)k * packetBytes + off;
        return (unsigned)b[0] | ((unsigned)b[1] << 8) | ((unsigned)b[2] << 16) | ((unsigned)b[3] << 24);
    };
    const unsigned nS = nSamplesPerPacket;
    unsigned prevSeqid = st->startTime, prevBlockid = st->startBlockid;
    unsigned k = 0;  // next received packet
    unsigned i = 0;  // next slot
    while (i < nSlots && k < nPackets) {
        const unsigned seqid = word(k, 8), blockid = word(k, 12);
        const unsigned cur = k++;
        // the first packet ever sets the stream's start (:212-216)
        if (i == 0 && st->startTime == 0) {
            prevSeqid = st->startTime = seqid;
            prevBlockid = st->startBlockid = st->startBlockid == 0 ? blockid : st->startBlockid;
        }
        // an untrustworthy seconds counter (:218-225)
        if (seqid == ~0u || prevSeqid + 10u < seqid) {
            ++st->rejected;
            continue;
        }
        // contiguity (:229-252): blocks per second alternate 195313/195312 at 200 MHz
        const unsigned totBlocks = clockMHz == 160 ? 156250u : (prevSeqid % 2 == 0 ? 195313u : 195312u);
        unsigned lostPackets = 0;
        const unsigned diff = blockid >= prevBlockid ? blockid - prevBlockid : blockid + totBlocks - prevBlockid;
        if (diff < nS) {  // duplicate
            ++st->rejected;
            continue;
        } else if (diff > nS) {
            lostPackets = diff / nS - 1;
        }
        // empty packets for the lost ones (:264-284), as far as the chunk allows
        unsigned pc = 0;
        for (; pc < lostPackets && i + pc < nSlots; ++pc) {
            prevSeqid = (prevBlockid + nS < totBlocks) ? prevSeqid : prevSeqid + 1;
            prevBlockid = (prevBlockid + nS) % totBlocks;
            h_slotSource[i + pc] = -1;
            ++st->lost;
        }
        i += pc;
        // the packet itself, unless the empty ones filled the chunk (:287-305)
        if (i != nSlots) {
            h_slotSource[i++] = (int)cur;
            ++st->accepted;
            prevSeqid = seqid;
            prevBlockid = blockid;
        }
    }
    st->startTime = prevSeqid;
    st->startBlockid = prevBlockid;
    *packetsUsed = k;
    *slotsFilled = i;
    return PPL_OK;
}
