// ppl_detect.hip -- dyadic boxcar threshold detection
// (DedispersionAnalyser::analyse, src/lib/src/DedispersionAnalyser.cpp:30-203).
//
// One wave per 64-sample segment of one DM row: lane j holds sample j of
// the segment; the level-n bin j stays on lane j * 2^n, the sum of the
// level n-1 bins on that lane and 2^(n-1) lanes above (binned[n][j] =
// binned[n-1][2j] + binned[n-1][2j+1], the reference's pair order), and
// ballots give the per-level hit masks.  Pass 1 counts the hits
// of every (DM, segment); an exclusive scan of the counts (row-major, the
// reference's event order) gives each segment its output offset; pass 2
// writes the events of the segments that have any.
#include "ppl_internal.h"

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <vector>

namespace {

using namespace ppl;

constexpr int kMaxLevels = 7;  // binPow2 <= 6: 64-sample segments

struct Thresholds {
    float v[kMaxLevels];
};

template <typename T>
__device__ __forceinline__ float sample_value(const T* data, long idx, int offset, float scale);
template <>
__device__ __forceinline__ float sample_value<float>(const float* data, long idx, int, float)
{
    return data[idx];
}
template <>
__device__ __forceinline__ float sample_value<int32_t>(const int32_t* data, long idx, int offset, float scale)
{
    return (float)(data[idx] - offset) * scale;
}

// Value of lane + D (within the row of 16 lanes for D <= 8: DPP row_shl;
// across rows by ds_bpermute).
template <int D>
__device__ __forceinline__ float from_lane_above(float x)
{
    if constexpr (D <= 8)
        return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x100 + D, 0xf, 0xf,
                                                                     true));
    else
        return __shfl_down(x, D, 64);
}

// Level hit masks of one segment.  The level-n bin j stays on lane j * 2^n
// (valid for lanes below W that are multiples of 2^n): it is the level n-1
// bin on that lane plus the one 2^(n-1) lanes above, the reference's pair
// order (binned[n][j] = binned[n-1][2j] + binned[n-1][2j+1]).  A hit on
// lane L of level n is bin L >> n, sample L of the segment.
template <int N>
__device__ __forceinline__ void level_up(float& cur, int lane, int W, const Thresholds& th, uint64_t* masks,
                                         float* vals)
{
    constexpr int D = 1 << (N - 1);
    cur = cur + from_lane_above<D>(cur);
    vals[N] = cur;
    masks[N] = __ballot(lane < W && (lane & (2 * D - 1)) == 0 && cur >= th.v[N]);
}

template <typename T>
__device__ __forceinline__ void segment_masks(const T* data, long rowBase, int seg, int P, const Thresholds& th,
                                              int offset, float scale, uint64_t* masks, float* vals)
{
    const int lane = threadIdx.x & 63;
    const int W = 1 << P;
    float v = 0.0f;
    if (lane < W) v = sample_value<T>(data, rowBase + (long)seg * W + lane, offset, scale);
    vals[0] = v;
    masks[0] = __ballot(lane < W && v >= th.v[0]);
    float cur = v;
    if (P >= 1) level_up<1>(cur, lane, W, th, masks, vals);
    if (P >= 2) level_up<2>(cur, lane, W, th, masks, vals);
    if (P >= 3) level_up<3>(cur, lane, W, th, masks, vals);
    if (P >= 4) level_up<4>(cur, lane, W, th, masks, vals);
    if (P >= 5) level_up<5>(cur, lane, W, th, masks, vals);
    if (P >= 6) level_up<6>(cur, lane, W, th, masks, vals);
}

template <typename T>
__global__ __launch_bounds__(256) void detect_count_kernel(const T* __restrict__ data, int tdms, int nsamp,
                                                           int P, Thresholds th, int offset, float scale,
                                                           int* __restrict__ counts)
{
    const long waveId = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nSeg = nsamp >> P;
    const long total = (long)tdms * nSeg;
    if (waveId >= total) return;
    const int dm = (int)(waveId / nSeg), seg = (int)(waveId % nSeg);
    uint64_t masks[kMaxLevels];
    float vals[kMaxLevels];
    segment_masks<T>(data, (long)dm * nsamp, seg, P, th, offset, scale, masks, vals);
    int c = 0;
    for (int n = 0; n <= P; ++n) c += __popcll(masks[n]);
    if ((threadIdx.x & 63) == 0) counts[waveId] = c;
}

template <typename T>
__global__ __launch_bounds__(256) void detect_write_kernel(const T* __restrict__ data, int tdms, int nsamp,
                                                           int P, Thresholds th, int offset, float scale,
                                                           const int* __restrict__ counts,
                                                           const int* __restrict__ offsets, int maxEvents,
                                                           int* __restrict__ evInts, float* __restrict__ evVals)
{
    const long waveId = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nSeg = nsamp >> P;
    const long total = (long)tdms * nSeg;
    if (waveId >= total) return;
    if (counts[waveId] == 0) return;
    const int lane = threadIdx.x & 63;
    const int dm = (int)(waveId / nSeg), seg = (int)(waveId % nSeg);
    uint64_t masks[kMaxLevels];
    float vals[kMaxLevels];
    segment_masks<T>(data, (long)dm * nsamp, seg, P, th, offset, scale, masks, vals);
    int base = offsets[waveId] + 1;  // slot 0 holds the dummy event
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int n = 0; n <= P; ++n) {
        const uint64_t m = masks[n];
        if ((m >> lane) & 1ull) {
            // bins in order of j = lane >> n: the rank among the mask's lanes
            const int k = base + __popcll(m & below);
            if (k < maxEvents) {
                evInts[3 * k + 0] = dm;
                evInts[3 * k + 1] = seg * (1 << P) + lane;  // binFactor * j
                evInts[3 * k + 2] = 1 << n;
                evVals[k] = vals[n];
            }
        }
        base += __popcll(m);
    }
}

template <typename T>
int run_detect(const T* d_data, int tdms, int nsamp, int offset, float scale, int nChannels, int nSubbands,
               float threshold, unsigned binPow2, int32_t* h_ints, float* h_vals, int max_events, int* count,
               hipStream_t st)
{
    PPL_CHECK_ARG(d_data && count, "null argument");
    PPL_CHECK_ARG(binPow2 <= 6, "power2ForBinning %u > 6 is not supported", binPow2);
    PPL_CHECK_ARG(tdms >= 0 && nsamp >= 0, "negative shape");
    PPL_CHECK_ARG(max_events <= 0 || (h_ints && h_vals), "null event buffers");
    // thresholds exactly as the reference forms them (float products)
    Thresholds th;
    const float rms = std::sqrt((float)nChannels * (float)nSubbands);
    th.v[0] = threshold * rms;
    for (unsigned n = 1; n <= binPow2; ++n) {
        const int binFactor = 1 << n;
        th.v[n] = threshold * rms * std::sqrt((float)binFactor);
    }
    const int P = (int)binPow2;
    const int nSeg = nsamp >> P;
    const long total = (long)tdms * nSeg;
    int nEvents = 1;
    if (max_events > 0) {
        h_ints[0] = 0;
        h_ints[1] = 0;
        h_ints[2] = 1;
        h_vals[0] = 0.0f;
    }
    if (total > 0) {
        int *d_counts = nullptr, *d_offsets = nullptr, *d_ints = nullptr;
        float* d_vals = nullptr;
        void* d_tmp = nullptr;
        size_t tmpBytes = 0;
        PPL_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmpBytes, d_counts, d_offsets, (int)total, st));
        PPL_HIP(hipMallocAsync((void**)&d_counts, sizeof(int) * (total + 1), st));
        PPL_HIP(hipMallocAsync((void**)&d_offsets, sizeof(int) * (total + 1), st));
        PPL_HIP(hipMallocAsync(&d_tmp, tmpBytes + 16, st));
        const unsigned blocks = (unsigned)((total + 3) / 4);
        hipLaunchKernelGGL(detect_count_kernel<T>, dim3(blocks), dim3(256), 0, st, d_data, tdms, nsamp, P, th,
                           offset, scale, d_counts);
        PPL_LAUNCHED();
        PPL_HIP(hipcub::DeviceScan::ExclusiveSum(d_tmp, tmpBytes, d_counts, d_offsets, (int)total, st));
        int last[2] = {0, 0};
        PPL_HIP(hipMemcpyAsync(&last[0], d_offsets + total - 1, sizeof(int), hipMemcpyDeviceToHost, st));
        PPL_HIP(hipMemcpyAsync(&last[1], d_counts + total - 1, sizeof(int), hipMemcpyDeviceToHost, st));
        PPL_HIP(hipStreamSynchronize(st));
        const int found = last[0] + last[1];
        nEvents = 1 + found;
        const int cap = std::min(nEvents, std::max(max_events, 0));
        if (found > 0 && cap > 1) {
            PPL_HIP(hipMallocAsync((void**)&d_ints, sizeof(int32_t) * 3 * cap, st));
            PPL_HIP(hipMallocAsync((void**)&d_vals, sizeof(float) * cap, st));
            hipLaunchKernelGGL(detect_write_kernel<T>, dim3(blocks), dim3(256), 0, st, d_data, tdms, nsamp, P,
                               th, offset, scale, d_counts, d_offsets, cap, d_ints, d_vals);
            PPL_LAUNCHED();
            PPL_HIP(hipMemcpyAsync(h_ints + 3, d_ints + 3, sizeof(int32_t) * 3 * (cap - 1), hipMemcpyDeviceToHost,
                                   st));
            PPL_HIP(hipMemcpyAsync(h_vals + 1, d_vals + 1, sizeof(float) * (cap - 1), hipMemcpyDeviceToHost, st));
            PPL_HIP(hipFreeAsync(d_ints, st));
            PPL_HIP(hipFreeAsync(d_vals, st));
        }
        PPL_HIP(hipFreeAsync(d_counts, st));
        PPL_HIP(hipFreeAsync(d_offsets, st));
        PPL_HIP(hipFreeAsync(d_tmp, st));
        PPL_HIP(hipStreamSynchronize(st));
    }
    *count = nEvents;
    return PPL_OK;
}

}  // namespace

extern "C" int ppl_detect(const float* d_data, int tdms, int nsamp, int nChannels, int nSubbands, float threshold,
                          unsigned binPow2, int32_t* h_ev_ints, float* h_ev_vals, int max_events, int* count,
                          void* stream)
{
    return run_detect<float>(d_data, tdms, nsamp, 0, 1.0f, nChannels, nSubbands, threshold, binPow2, h_ev_ints,
                             h_ev_vals, max_events, count, stream_of(stream));
}

extern "C" int ppl_detect_u8(const int32_t* d_data, int tdms, int nsamp, int nChan, float scale, int nChannels,
                             int nSubbands, float threshold, unsigned binPow2, int32_t* h_ev_ints,
                             float* h_ev_vals, int max_events, int* count, void* stream)
{
    return run_detect<int32_t>(d_data, tdms, nsamp, 128 * nChan, scale, nChannels, nSubbands, threshold, binPow2,
                               h_ev_ints, h_ev_vals, max_events, count, stream_of(stream));
}

namespace {
__global__ void dequantise_kernel(const int32_t* __restrict__ in, float* __restrict__ out, long n, int offset,
                                  float scale)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)(in[i] - offset) * scale;
}
}  // namespace

extern "C" int ppl_dequantise_sums(const int32_t* d_in, float* d_out, size_t n, unsigned nChan, float scale,
                                   void* stream)
{
    PPL_CHECK_ARG(d_in && d_out, "null argument");
    if (n == 0) return PPL_OK;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(dequantise_kernel, dim3(blocks), dim3(256), 0, stream_of(stream), d_in, d_out, (long)n,
                       (int)(128 * nChan), scale);
    PPL_LAUNCHED();
    return PPL_OK;
}
