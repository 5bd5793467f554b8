// ppl_detect.hip -- dyadic boxcar threshold detection
// (DedispersionAnalyser::analyse, src/lib/src/DedispersionAnalyser.cpp:30-203).
//
// One wave per 64-sample segment of one DM row: lane j holds sample j of
// the segment; the level-n bin j stays on lane j * 2^n, the sum of the
// level n-1 bins on that lane and 2^(n-1) lanes above (binned[n][j] =
// binned[n-1][2j] + binned[n-1][2j+1], the reference's pair order), and
// ballots give the per-level hit masks.  Pass 1 counts the hits of every
// (DM, segment) and of every workgroup's 64 segments; pass 2 (one
// workgroup) scans the workgroup counts (row-major, the reference's event
// order); pass 3 re-reads only the segments that have hits and writes their
// events at their offsets.  No host synchronisation separates the passes.
#include "ppl_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <system_error>
#include <thread>
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

// Segments are numbered row-major (id = dm * nSeg + seg), the reference's
// event order.  A workgroup of four waves takes 64 consecutive ids, 16 per
// wave; the ids a wave owns are loaded before any is reduced, so 16 row
// segments are in flight per wave.
constexpr int kSegPerWave = 16;
constexpr int kSegPerBlock = 4 * kSegPerWave;

// Lanes that hold a level-n bin of a 2^P-sample segment: the multiples of
// 2^n below 2^P.
template <int P>
constexpr uint64_t level_lanes(int n)
{
    uint64_t m = 0;
    for (int l = 0; l < (1 << P); l += 1 << n) m |= 1ull << l;
    return m;
}

// Hits of one segment at every level (the row's value on this lane is v):
// per level one add of the lane 2^(n-1) above (a DPP move folded into the
// add for n <= 4) and one compare whose lane mask is the ballot; the lanes
// that hold no bin are masked off by a constant.
template <int P>
__device__ __forceinline__ int segment_hits(float v, const Thresholds& th)
{
    int c = __popcll(__builtin_amdgcn_ballot_w64(v >= th.v[0]) & level_lanes<P>(0));
    float cur = v;
    if constexpr (P >= 1) {
        cur = cur + from_lane_above<1>(cur);
        c += __popcll(__builtin_amdgcn_ballot_w64(cur >= th.v[1]) & level_lanes<P>(1));
    }
    if constexpr (P >= 2) {
        cur = cur + from_lane_above<2>(cur);
        c += __popcll(__builtin_amdgcn_ballot_w64(cur >= th.v[2]) & level_lanes<P>(2));
    }
    if constexpr (P >= 3) {
        cur = cur + from_lane_above<4>(cur);
        c += __popcll(__builtin_amdgcn_ballot_w64(cur >= th.v[3]) & level_lanes<P>(3));
    }
    if constexpr (P >= 4) {
        cur = cur + from_lane_above<8>(cur);
        c += __popcll(__builtin_amdgcn_ballot_w64(cur >= th.v[4]) & level_lanes<P>(4));
    }
    if constexpr (P >= 5) {
        cur = cur + from_lane_above<16>(cur);
        c += __popcll(__builtin_amdgcn_ballot_w64(cur >= th.v[5]) & level_lanes<P>(5));
    }
    if constexpr (P >= 6) {
        cur = cur + from_lane_above<32>(cur);
        c += __popcll(__builtin_amdgcn_ballot_w64(cur >= th.v[6]) & level_lanes<P>(6));
    }
    return c;
}

// Pass 1: hits per segment (at most 127, one byte each) and per workgroup.
// P (the levels above 0) is a template parameter, so the level loop and
// the lane masks are fixed at compile time.
template <typename T, int P>
__global__ __launch_bounds__(256) void detect_count_kernel(const T* __restrict__ data, int nsamp, int nSeg,
                                                           int total, Thresholds th, int offset,
                                                           float scale, uint8_t* __restrict__ counts,
                                                           int* __restrict__ blockSums)
{
    __shared__ int part[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int W = 1 << P;
    const int first = blockIdx.x * kSegPerBlock + wave * kSegPerWave;
    int dm = first / nSeg, seg = first - dm * nSeg;
    float v[kSegPerWave];
#pragma unroll
    for (int k = 0; k < kSegPerWave; ++k) {
        v[k] = 0.0f;
        if (first + k < total && lane < W)
            v[k] = sample_value<T>(data, (long)dm * nsamp + (long)seg * W + lane, offset, scale);
        if (++seg == nSeg) {
            seg = 0;
            ++dm;
        }
    }
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kSegPerWave; ++k) {
        if (first + k < total) {
            const int c = segment_hits<P>(v[k], th);
            if (lane == 0) counts[first + k] = (uint8_t)c;
            sum += c;
        }
    }
    if (lane == 0) part[wave] = sum;
    __syncthreads();
    if (threadIdx.x == 0) blockSums[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Four consecutive samples of a row as floats (16-byte load).
template <typename T>
__device__ __forceinline__ float4 sample_quad(const T* data, long idx, int offset, float scale);
template <>
__device__ __forceinline__ float4 sample_quad<float>(const float* data, long idx, int, float)
{
    return *reinterpret_cast<const float4*>(data + idx);
}
template <>
__device__ __forceinline__ float4 sample_quad<int32_t>(const int32_t* data, long idx, int offset, float scale)
{
    const int4 v = *reinterpret_cast<const int4*>(data + idx);
    return make_float4((float)(v.x - offset) * scale, (float)(v.y - offset) * scale, (float)(v.z - offset) * scale,
                       (float)(v.w - offset) * scale);
}

// Value of lane + D (D <= 8, inside a row of 16 lanes; 0 past the row).
__device__ __forceinline__ int int_from_lane_above(int x, int D)
{
    switch (D) {
        case 1: return __builtin_amdgcn_update_dpp(0, x, 0x101, 0xf, 0xf, true);
        case 2: return __builtin_amdgcn_update_dpp(0, x, 0x102, 0xf, 0xf, true);
        case 4: return __builtin_amdgcn_update_dpp(0, x, 0x104, 0xf, 0xf, true);
        default: return __builtin_amdgcn_update_dpp(0, x, 0x108, 0xf, 0xf, true);
    }
}

// Pass 1 with 16-byte loads, for segments of 16..64 samples (P = 4..6):
// lane m of a segment's LS = 2^(P-2) lanes holds its samples 4m..4m+3.
// Levels 0..2 are summed in the lane (x+y, z+w, then their sum, the
// reference's pair order); level n > 2 adds the level n-1 bin of the lane
// 2^(n-3) above (DPP row shift), kept on lanes m that are multiples of
// 2^(n-2).  Each lane counts its own bins' hits and the segment's lanes
// reduce them onto its first lane.  Same counts as detect_count_kernel.
template <typename T, int P>
__global__ __launch_bounds__(256) void detect_count4_kernel(const T* __restrict__ data, int nsamp, int nSeg,
                                                            int total, Thresholds th, int offset, float scale,
                                                            uint8_t* __restrict__ counts, int* __restrict__ blockSums)
{
    static_assert(P >= 4 && P <= 6, "16..64-sample segments");
    constexpr int LS = 1 << (P - 2);        // lanes per segment
    constexpr int SPL = 64 / LS;            // segments per load
    constexpr int NL = kSegPerWave / SPL;   // loads per lane
    __shared__ int part[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & (LS - 1);
    const int first = blockIdx.x * kSegPerBlock + wave * kSegPerWave;
    float4 v[NL];
#pragma unroll
    for (int L = 0; L < NL; ++L) {
        const int id = first + L * SPL + lane / LS;
        v[L] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (id < total) {
            const int dm = id / nSeg, seg = id - dm * nSeg;
            v[L] = sample_quad<T>(data, (long)dm * nsamp + ((long)seg << P) + 4 * m, offset, scale);
        }
    }
    int sum = 0;
#pragma unroll
    for (int L = 0; L < NL; ++L) {
        const float4 x = v[L];
        int c = (x.x >= th.v[0]) + (x.y >= th.v[0]) + (x.z >= th.v[0]) + (x.w >= th.v[0]);
        const float a = x.x + x.y, b = x.z + x.w;
        c += (a >= th.v[1]) + (b >= th.v[1]);
        float r = a + b;
        c += r >= th.v[2];
        r = r + from_lane_above<1>(r);
        c += (m & 1) == 0 && r >= th.v[3];
        r = r + from_lane_above<2>(r);
        c += (m & 3) == 0 && r >= th.v[4];
        if constexpr (P >= 5) {
            r = r + from_lane_above<4>(r);
            c += (m & 7) == 0 && r >= th.v[5];
        }
        if constexpr (P >= 6) {
            r = r + from_lane_above<8>(r);
            c += (m & 15) == 0 && r >= th.v[6];
        }
#pragma unroll
        for (int d = 1; d < LS; d <<= 1) c += int_from_lane_above(c, d);
        const int id = first + L * SPL + lane / LS;
        if (m == 0 && id < total) {
            counts[id] = (uint8_t)c;
            sum += c;
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
    if (lane == 0) part[wave] = sum;
    __syncthreads();
    if (threadIdx.x == 0) blockSums[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Pass 2 (one workgroup): exclusive prefix of the workgroup sums, in place,
// 16384 at a time (sixteen consecutive per thread), with off[nBlocks] = their total;
// *evTotal = 1 + total, and the dummy event (0, 0, 1, 0) in slot 0
// (DedispersionAnalyser.cpp:69).
__global__ __launch_bounds__(1024) void detect_scan_kernel(int* __restrict__ off, int nBlocks,
                                                           int* __restrict__ evTotal, int maxEvents,
                                                           int* __restrict__ evInts, float* __restrict__ evVals)
{
    __shared__ int waveTot[16];
    __shared__ int chunkTot;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    constexpr int K = 16;
    int carry = 0;
    for (int base = 0; base < nBlocks; base += 1024 * K) {
        const int i0 = base + K * t;
        int x[K];
        if (i0 + K <= nBlocks) {  // off is 256-byte aligned and i0 a multiple of 16
#pragma unroll
            for (int k = 0; k < K; k += 4) {
                const int4 q = *reinterpret_cast<const int4*>(off + i0 + k);
                x[k] = q.x;
                x[k + 1] = q.y;
                x[k + 2] = q.z;
                x[k + 3] = q.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k) x[k] = i0 + k < nBlocks ? off[i0 + k] : 0;
        }
        int s = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) s += x[k];
        int incl = s;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(incl, d, 64);
            if (lane >= d) incl += y;
        }
        if (lane == 63) waveTot[wave] = incl;
        __syncthreads();
        if (wave == 0) {
            const int w = lane < 16 ? waveTot[lane] : 0;
            int wi = w;
#pragma unroll
            for (int d = 1; d < 16; d <<= 1) {
                const int y = __shfl_up(wi, d, 64);
                if (lane >= d) wi += y;
            }
            if (lane < 16) waveTot[lane] = wi - w;
            if (lane == 15) chunkTot = wi;
        }
        __syncthreads();
        int e = carry + waveTot[wave] + incl - s;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (i0 + k < nBlocks) off[i0 + k] = e;
            e += x[k];
        }
        carry += chunkTot;
        __syncthreads();
    }
    if (t == 0) {
        off[nBlocks] = carry;
        *evTotal = 1 + carry;
        if (maxEvents > 0) {
            evInts[0] = 0;
            evInts[1] = 0;
            evInts[2] = 1;
            evVals[0] = 0.0f;
        }
    }
}

// Pass 3: the events of every segment with hits, at 1 + its offset (slot 0
// holds the dummy event).
template <typename T>
__global__ __launch_bounds__(256) void detect_write_kernel(const T* __restrict__ data, int nsamp, int nSeg,
                                                           int total, int P, Thresholds th, int offset,
                                                           float scale, const uint8_t* __restrict__ counts,
                                                           const int* __restrict__ off, int maxEvents,
                                                           int* __restrict__ evInts, float* __restrict__ evVals)
{
    const int b = blockIdx.x;
    const int off0 = off[b];
    if (off[b + 1] == off0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int blockFirst = b * kSegPerBlock;
    // lane l: hits of segment blockFirst + l and the hits before it in the workgroup
    const int c = blockFirst + lane < total ? (int)counts[blockFirst + lane] : 0;
    int incl = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
    }
    const int excl = incl - c;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int k = 0; k < kSegPerWave; ++k) {
        const int l = wave * kSegPerWave + k;
        if (__shfl(c, l, 64) == 0) continue;
        const int id = blockFirst + l;
        const int dm = id / nSeg, seg = id - dm * nSeg;
        uint64_t masks[kMaxLevels];
        float vals[kMaxLevels];
        segment_masks<T>(data, (long)dm * nsamp, seg, P, th, offset, scale, masks, vals);
        int base = 1 + off0 + __shfl(excl, l, 64);
        for (int n = 0; n <= P; ++n) {
            const uint64_t m = masks[n];
            if ((m >> lane) & 1ull) {
                // bins in order of j = lane >> n: the rank among the mask's lanes
                const int e = base + __popcll(m & below);
                if (e < maxEvents) {
                    evInts[3 * e + 0] = dm;
                    evInts[3 * e + 1] = seg * (1 << P) + lane;  // binFactor * j
                    evInts[3 * e + 2] = 1 << n;
                    evVals[e] = vals[n];
                }
            }
            base += __popcll(m);
        }
    }
}

// The reference's thresholds, formed exactly as it forms them (float products).
Thresholds make_thresholds(int nChannels, int nSubbands, float threshold, unsigned binPow2)
{
    Thresholds th{};
    const float rms = std::sqrt((float)nChannels * (float)nSubbands);
    th.v[0] = threshold * rms;
    for (unsigned n = 1; n <= binPow2; ++n) {
        const int binFactor = 1 << n;
        th.v[n] = threshold * rms * std::sqrt((float)binFactor);
    }
    return th;
}

// The three passes on a stream, into device-accessible event buffers of
// maxEvents slots and a device-accessible total; scratch holds
// detect_scratch_bytes(total) bytes.
size_t detect_scratch_bytes(long total)
{
    const long nBlocks = (total + kSegPerBlock - 1) / kSegPerBlock;
    return (size_t)((total + 255) / 256 * 256) + sizeof(int) * (size_t)(nBlocks + 1);
}

template <typename T>
int launch_detect(const T* d_data, int tdms, int nsamp, int offset, float scale, const Thresholds& th, int P,
                  void* scratch, int32_t* evInts, float* evVals, int maxEvents, int* evTotal, hipStream_t st)
{
    const int nSeg = nsamp >> P;
    const int total = tdms * nSeg;
    const int nBlocks = (total + kSegPerBlock - 1) / kSegPerBlock;
    uint8_t* counts = static_cast<uint8_t*>(scratch);
    int* off = reinterpret_cast<int*>(counts + (size_t)(total + 255) / 256 * 256);
    // 16-byte loads where the rows are 16-byte aligned and segments hold 16..64 samples
    const bool quads = P >= 4 && nsamp % 4 == 0 && (reinterpret_cast<uintptr_t>(d_data) & 15) == 0;
    if (total > 0 && quads) {
#define PPL_COUNT4(PP)                                                                                       \
    case PP:                                                                                                 \
        hipLaunchKernelGGL((detect_count4_kernel<T, PP>), dim3(nBlocks), dim3(256), 0, st, d_data, nsamp, nSeg, \
                           total, th, offset, scale, counts, off);                                           \
        break;
        switch (P) {
            PPL_COUNT4(4)
            PPL_COUNT4(5)
            PPL_COUNT4(6)
        }
#undef PPL_COUNT4
        PPL_LAUNCHED();
    } else if (total > 0) {
#define PPL_COUNT(PP)                                                                                         \
    case PP:                                                                                                  \
        hipLaunchKernelGGL((detect_count_kernel<T, PP>), dim3(nBlocks), dim3(256), 0, st, d_data, nsamp, nSeg, \
                           total, th, offset, scale, counts, off);                                            \
        break;
        switch (P) {
            PPL_COUNT(0)
            PPL_COUNT(1)
            PPL_COUNT(2)
            PPL_COUNT(3)
            PPL_COUNT(4)
            PPL_COUNT(5)
            PPL_COUNT(6)
            default:
                set_error("detect: binPow2 %d outside 0..6", P);
                return PPL_E_ARG;
        }
#undef PPL_COUNT
        PPL_LAUNCHED();
    }
    hipLaunchKernelGGL(detect_scan_kernel, dim3(1), dim3(1024), 0, st, off, nBlocks, evTotal, maxEvents, evInts,
                       evVals);
    PPL_LAUNCHED();
    if (total == 0) return PPL_OK;  // the dummy event alone
    hipLaunchKernelGGL(detect_write_kernel<T>, dim3(nBlocks), dim3(256), 0, st, d_data, nsamp, nSeg, total, P, th,
                       offset, scale, counts, off, maxEvents, evInts, evVals);
    PPL_LAUNCHED();
    return PPL_OK;
}

// Scratch of ppl_detect_noalloc: the hit count of every segment (a byte
// each) and the workgroup offsets, for up to PPL_DETECT_NOALLOC_MAX_SEGMENTS
// segments.  Every device the library runs on holds its own copy until the
// process ends.  Host code never passes the array's name to a kernel: it
// uses the address hipGetSymbolAddress returns for the current device,
// looked up once per device (at ppl_init, or at the first call on a device
// that ppl_init did not set up).
constexpr long kNoAllocSegments = PPL_DETECT_NOALLOC_MAX_SEGMENTS;
constexpr size_t kNoAllocBytes =
    (size_t)((kNoAllocSegments + 255) / 256 * 256) +
    sizeof(int) * (size_t)((kNoAllocSegments + kSegPerBlock - 1) / kSegPerBlock + 1);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
void* g_scratchPtr[kMaxDevices] = {};

// The current device's copy of the scratch array.
int scratch_of_current_device(void** out)
{
    int dev = 0;
    PPL_HIP(hipGetDevice(&dev));
    PPL_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d the library tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        PPL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratchPtr[dev] = p;
    }
    *out = g_scratchPtr[dev];
    return PPL_OK;
}

// Pinned host staging for the events of ppl_detect / ppl_detect_u8: the
// write pass stores straight into it, so a call synchronises once.  It grows
// to the largest max_events asked for and is shared by the calls of a
// process under a lock (each call is synchronous).
struct Staging {
    std::mutex mu;
    void* host = nullptr;
    size_t cap = 0;  // event slots
};
Staging g_staging;

int staging_reserve(size_t events)
{
    if (events <= g_staging.cap && g_staging.host) return PPL_OK;
    if (g_staging.host) PPL_HIP(hipHostFree(g_staging.host));
    g_staging.host = nullptr;
    g_staging.cap = 0;
    const size_t bytes = 256 + events * (3 * sizeof(int32_t) + sizeof(float));
    PPL_HIP(hipHostMalloc(&g_staging.host, bytes, hipHostMallocDefault));
    g_staging.cap = events;
    return PPL_OK;
}

// Copy n events out of the pinned staging.  Long lists are split over a few
// threads: the destination is usually fresh pageable memory, and faulting
// its pages in is most of the time.
void copy_events(int32_t* h_ints, float* h_vals, const int32_t* sInts, const float* sVals, size_t n)
{
    constexpr size_t kPerThread = size_t(1) << 17;  // events
    const size_t T = std::min<size_t>(8, (n + kPerThread - 1) / kPerThread);
    auto part = [&](size_t k) {
        const size_t a = n * k / T, b = n * (k + 1) / T;
        std::memcpy(h_ints + 3 * a, sInts + 3 * a, sizeof(int32_t) * 3 * (b - a));
        std::memcpy(h_vals + a, sVals + a, sizeof(float) * (b - a));
    };
    if (T <= 1) {
        std::memcpy(h_ints, sInts, sizeof(int32_t) * 3 * n);
        std::memcpy(h_vals, sVals, sizeof(float) * n);
        return;
    }
    std::vector<std::thread> pool;
    size_t k = 1;
    try {
        for (; k < T; ++k) pool.emplace_back(part, k);
    } catch (const std::system_error&) {
        // no thread to be had: this thread copies the parts not started
    }
    for (size_t r = k; r < T; ++r) part(r);
    part(0);
    for (auto& t : pool) t.join();
}

template <typename T>
int run_detect(const T* d_data, int tdms, int nsamp, int offset, float scale, int nChannels, int nSubbands,
               float threshold, unsigned binPow2, int32_t* h_ints, float* h_vals, int max_events, int* count,
               hipStream_t st)
{
    PPL_CHECK_ARG(count, "null argument");
    PPL_CHECK_ARG(binPow2 <= 6, "power2ForBinning %u > 6 is not supported", binPow2);
    PPL_CHECK_ARG(tdms >= 0 && nsamp >= 0, "negative shape");
    PPL_CHECK_ARG(max_events <= 0 || (h_ints && h_vals), "null event buffers");
    const int P = (int)binPow2;
    const long total = (long)tdms * (nsamp >> P);
    PPL_CHECK_ARG(d_data || total == 0, "null argument");
    PPL_CHECK_ARG(total < (1L << 30), "%ld segments (at most 2^30)", total);
    const Thresholds th = make_thresholds(nChannels, nSubbands, threshold, binPow2);
    std::lock_guard<std::mutex> lock(g_staging.mu);
    const size_t slots = (size_t)std::max(max_events, 1);
    int rc = staging_reserve(slots);
    if (rc != PPL_OK) return rc;
    int* hTotal = static_cast<int*>(g_staging.host);
    int32_t* sInts = reinterpret_cast<int32_t*>(static_cast<char*>(g_staging.host) + 256);
    float* sVals = reinterpret_cast<float*>(sInts + 3 * slots);
    void *dTotal = nullptr, *dInts = nullptr, *dVals = nullptr;
    PPL_HIP(hipHostGetDevicePointer(&dTotal, hTotal, 0));
    PPL_HIP(hipHostGetDevicePointer(&dInts, sInts, 0));
    PPL_HIP(hipHostGetDevicePointer(&dVals, sVals, 0));
    void* scratch = nullptr;
    PPL_HIP(hipMallocAsync(&scratch, detect_scratch_bytes(total), st));
    rc = launch_detect<T>(d_data, tdms, nsamp, offset, scale, th, P, scratch, static_cast<int32_t*>(dInts),
                          static_cast<float*>(dVals), std::max(max_events, 0), static_cast<int*>(dTotal), st);
    PPL_HIP(hipFreeAsync(scratch, st));
    if (rc != PPL_OK) return rc;
    PPL_HIP(hipStreamSynchronize(st));
    const int nEvents = *hTotal;
    const int n = std::min(nEvents, std::max(max_events, 0));
    if (n > 0) copy_events(h_ints, h_vals, sInts, sVals, (size_t)n);
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

// Set-up for the current device (ppl_init): looks up its scratch address.
int ppl::detect_noalloc_setup()
{
    void* p = nullptr;
    return scratch_of_current_device(&p);
}

extern "C" int ppl_detect_noalloc(const float* d_data, int tdms, int nsamp, int nChannels, int nSubbands,
                                  float threshold, unsigned binPow2, int32_t* d_ev_ints, float* d_ev_vals,
                                  int max_events, int* d_count, void* stream)
{
    PPL_CHECK_ARG(d_count, "null argument");
    PPL_CHECK_ARG(binPow2 <= 6, "power2ForBinning %u > 6 is not supported", binPow2);
    PPL_CHECK_ARG(tdms >= 0 && nsamp >= 0, "negative shape");
    PPL_CHECK_ARG(max_events <= 0 || (d_ev_ints && d_ev_vals), "null event buffers");
    const int P = (int)binPow2;
    const long total = (long)tdms * (nsamp >> P);
    PPL_CHECK_ARG(d_data || total == 0, "null argument");
    PPL_CHECK_ARG(total <= kNoAllocSegments,
                  "%ld segments of 2^%u samples exceed the %ld the no-allocation scratch holds", total, binPow2,
                  kNoAllocSegments);
    void* sc = nullptr;
    const int rc = scratch_of_current_device(&sc);
    if (rc != PPL_OK) return rc;
    const Thresholds th = make_thresholds(nChannels, nSubbands, threshold, binPow2);
    return launch_detect<float>(d_data, tdms, nsamp, 0, 1.0f, th, P, sc, d_ev_ints, d_ev_vals,
                                std::max(max_events, 0), d_count, stream_of(stream));
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
