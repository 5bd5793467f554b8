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
__global__ void ingest_kernel(const uint8_t* __restrict__ packets, long packetBytes, int headerBytes, int nPackets,
                              int nSub, int nSpp, int nPol, Tout* __restrict__ out, long nSampOut)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long perPacket = (long)nSub * nSpp * nPol;
    if (i >= perPacket * nPackets) return;
    const int k = (int)(i / perPacket);
    const long r = i % perPacket;
    const int s = (int)(r / ((long)nSpp * nPol));
    const int t = (int)((r / nPol) % nSpp);
    const int p = (int)(r % nPol);
    const Tin* src = reinterpret_cast<const Tin*>(packets + (long)k * packetBytes + headerBytes) + 2 * r;
    const long time = (long)k * nSpp + t;
    // [sub][pol][time]: the block layout [sub][pol][block][t] is the same
    // order with time = block * nSamplesPerTimeBlock + t
    const long o = ((long)s * nPol + p) * nSampOut + time;
    out[2 * o] = (Tout)src[0];
    out[2 * o + 1] = (Tout)src[1];
}

template <typename Tout>
int launch(const void* d_packets, size_t packetBytes, unsigned headerBytes, unsigned nPackets, unsigned nSub,
           unsigned nSpp, unsigned nPol, int sampleBits, Tout* d_out, hipStream_t st)
{
    const long n = (long)nSub * nSpp * nPol * nPackets;
    if (n == 0) return PPL_OK;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    const long nSampOut = (long)nSpp * nPackets;
    if (sampleBits == 16)
        hipLaunchKernelGGL((ingest_kernel<int16_t, Tout>), dim3(blocks), dim3(256), 0, st,
                           static_cast<const uint8_t*>(d_packets), (long)packetBytes, (int)headerBytes,
                           (int)nPackets, (int)nSub, (int)nSpp, (int)nPol, d_out, nSampOut);
    else
        hipLaunchKernelGGL((ingest_kernel<int8_t, Tout>), dim3(blocks), dim3(256), 0, st,
                           static_cast<const uint8_t*>(d_packets), (long)packetBytes, (int)headerBytes,
                           (int)nPackets, (int)nSub, (int)nSpp, (int)nPol, d_out, nSampOut);
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
    return launch<float>(d_packets, packetBytes, headerBytes, nPackets, nSub, nSamplesPerPacket, nPol, sampleBits,
                         d_blob, stream_of(stream));
}

extern "C" int ppl_ingest_i16(const void* d_packets, size_t packetBytes, unsigned headerBytes, unsigned nPackets,
                              unsigned nSub, unsigned nSamplesPerPacket, unsigned nPol, int sampleBits,
                              int16_t* d_volt, void* stream)
{
    int rc = check(d_packets, packetBytes, headerBytes, nSub, nSamplesPerPacket, nPol, sampleBits, d_volt);
    if (rc != PPL_OK) return rc;
    return launch<int16_t>(d_packets, packetBytes, headerBytes, nPackets, nSub, nSamplesPerPacket, nPol, sampleBits,
                           d_volt, stream_of(stream));
}
