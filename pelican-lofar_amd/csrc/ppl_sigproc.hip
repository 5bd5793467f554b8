// ppl_sigproc.hip -- the data block of a Sigproc filterbank file
// (SigprocStokesWriter::sendStream, src/lib/src/SigprocStokesWriter.cpp:184-240,
// _float2int :285-290), formed on the device so that one contiguous copy
// brings back exactly the bytes the file holds.
//
// Order: for t, for p < nPolsToWrite, subbands from the last to the first,
// channels from the last to the first (the filterbank runs from the highest
// frequency down).  32 bits: the float itself.  8 bits: x = min(v, cropMax);
// 0 if x < cropMin, else int(rint((x - cropMin) * nRange / (cropMax - cropMin)))
// with nRange = 255, all in float, written as the int's low byte.
//
// One thread per output value: writes are coalesced, and each wave reads one
// channel row backwards (still one contiguous span).  HBM-bound: 4 B read and
// 4 or 1 B written per value.
#include "ppl_internal.h"

#include <math.h>

namespace {

using namespace ppl;

template <typename Tout>
__global__ __launch_bounds__(256) void sigproc_pack_kernel(const float* __restrict__ I, Tout* __restrict__ out,
                                                           long n, int nSub, int nPol, int nCh, int nPolW,
                                                           float cropMin, float cropMax, float nRange,
                                                           float scaleDelta)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int cr = (int)(i % nCh);
    long r = i / nCh;
    const int sr = (int)(r % nSub);
    r /= nSub;
    const int p = (int)(r % nPolW);
    const long t = r / nPolW;
    const int s = nSub - 1 - sr, c = nCh - 1 - cr;
    const float v = I[((t * nSub + s) * nPol + p) * nCh + c];
    if constexpr (sizeof(Tout) == 4) {
        out[i] = v;
    } else {
        const float x = (v > cropMax) ? cropMax : v;
        int q = 0;
        if (!(x < cropMin)) {
            const float y = rintf(((x - cropMin) * nRange) / scaleDelta);
            // the float-to-int conversion of the reference, NaN and out of
            // range giving INT_MIN (x86 cvttss2si), whose low byte is 0
            q = (y >= -2147483648.0f && y < 2147483648.0f) ? (int)y : (int)0x80000000;
        }
        out[i] = (uint8_t)(q & 0xff);
    }
}

}  // namespace

extern "C" int ppl_sigproc_pack(const float* d_stokes, unsigned nSamples, unsigned nSubbands, unsigned nPolarisations,
                                unsigned nChannels, unsigned nPolsToWrite, int nBits, float cropMin, float cropMax,
                                void* d_out, void* stream)
{
    PPL_CHECK_ARG(nBits == 32 || nBits == 8, "SigprocStokesWriter: %d bit datafiles not yet supported", nBits);
    PPL_CHECK_ARG(nSubbands > 0 && nChannels > 0 && nPolarisations > 0 && nPolsToWrite > 0,
                  "empty blob shape (%u subbands, %u channels, %u polarisations, %u to write)", nSubbands, nChannels,
                  nPolarisations, nPolsToWrite);
    PPL_CHECK_ARG(nPolsToWrite <= nPolarisations, "%u polarisations to write, the blob has %u", nPolsToWrite,
                  nPolarisations);
    const long n = (long)nSamples * nPolsToWrite * nSubbands * nChannels;
    if (n == 0) return PPL_OK;
    PPL_CHECK_ARG(d_stokes && d_out, "null argument");
    PPL_CHECK_ARG((long)nSamples * nPolarisations * nSubbands * nChannels < (1L << 40), "blob too large");
    const unsigned blocks = (unsigned)((n + 255) / 256);
    // nRange = (int)pow(2, 8) - 1.0 and scaleDelta = cropMax - cropMin, as floats
    const float nRange = 255.0f, scaleDelta = cropMax - cropMin;
    if (nBits == 32)
        hipLaunchKernelGGL((sigproc_pack_kernel<float>), dim3(blocks), dim3(256), 0, stream_of(stream), d_stokes,
                           static_cast<float*>(d_out), n, (int)nSubbands, (int)nPolarisations, (int)nChannels,
                           (int)nPolsToWrite, cropMin, cropMax, nRange, scaleDelta);
    else
        hipLaunchKernelGGL((sigproc_pack_kernel<uint8_t>), dim3(blocks), dim3(256), 0, stream_of(stream), d_stokes,
                           static_cast<uint8_t*>(d_out), n, (int)nSubbands, (int)nPolarisations, (int)nChannels,
                           (int)nPolsToWrite, cropMin, cropMax, nRange, scaleDelta);
    PPL_LAUNCHED();
    return PPL_OK;
}
