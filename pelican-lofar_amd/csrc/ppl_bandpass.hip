// ppl_bandpass.hip -- the integrated spectrum of BandPassRecorder::run
// (src/lib/src/BandPassRecorder.cpp:85-97): for every bin b = s*nChannels + c,
// sum[b] += I[t][s][0][c] for t = 0 .. nSamples-1 in order, in float, so the
// sums equal the reference's sequential loop bit for bit whatever the
// number and size of the blobs fed in.
//
// One thread per bin keeps each bin's additions in time order, so the
// parallelism is the number of bins (15872 at 62 x 256: 248 waves).  Each
// workgroup is one wave, so those waves land on as many CUs; lanes hold
// consecutive channels, so every time step is one coalesced 256-byte row
// read; 32 time steps are loaded before they are added, which keeps 32 loads
// per lane in flight (the sum is HBM-bound: 4 B read per bin and sample).
#include "ppl_internal.h"

namespace {

using namespace ppl;

constexpr int kThreads = 64;
constexpr int kUnroll = 32;

__global__ __launch_bounds__(kThreads) void bp_accumulate_kernel(const float* __restrict__ I, float* __restrict__ sum,
                                                                   int nT, int nSub, int nPol, int nCh)
{
    const int b = blockIdx.x * kThreads + threadIdx.x;
    if (b >= nSub * nCh) return;
    const int s = b / nCh;
    const int c = b - s * nCh;
    const long row = (long)nSub * nPol * nCh;  // floats per time step
    const float* p = I + (long)s * nPol * nCh + c;
    float acc = sum[b];
    int t = 0;
    for (; t + kUnroll <= nT; t += kUnroll) {
        float v[kUnroll];
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) v[k] = __builtin_nontemporal_load(p + (long)(t + k) * row);
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) acc += v[k];
    }
    for (; t < nT; ++t) acc += p[(long)t * row];
    sum[b] = acc;
}

}  // namespace

extern "C" int ppl_bandpass_accumulate(const float* d_stokes, int nSamples, int nSubbands, int nPolarisations,
                                       int nChannels, float* d_sum, void* stream)
{
    PPL_CHECK_ARG(d_sum && (d_stokes || nSamples == 0), "null argument");
    PPL_CHECK_ARG(nSamples >= 0 && nSubbands > 0 && nPolarisations > 0 && nChannels > 0,
                  "bad blob shape (%d samples, %d subbands, %d polarisations, %d channels)", nSamples, nSubbands,
                  nPolarisations, nChannels);
    PPL_CHECK_ARG((long)nSubbands * nChannels <= 0x7fffffffL - kThreads &&
                      (long)nSamples * nSubbands * nPolarisations * nChannels < (1L << 62),
                  "blob too large");
    if (nSamples == 0) return PPL_OK;
    const int nBins = nSubbands * nChannels;
    hipLaunchKernelGGL(bp_accumulate_kernel, dim3((nBins + kThreads - 1) / kThreads), dim3(kThreads), 0,
                       stream_of(stream), d_stokes, d_sum, nSamples, nSubbands, nPolarisations, nChannels);
    PPL_LAUNCHED();
    return PPL_OK;
}
