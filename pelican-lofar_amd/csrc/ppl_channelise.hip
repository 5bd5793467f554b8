// ppl_channelise.hip -- polyphase channeliser, Stokes-I and time integration.
//
// Plain FP32 method.  A wave streams one run of L consecutive blocks of one
// subband, both polarisations.  Each lane owns R = N/G channels of a block
// (G = min(N, 64) lanes per block; for N < 64 a wave carries 64/N runs side
// by side).  The FIR keeps the last nTaps blocks of its channels in
// registers, so every input sample is read from HBM once (plus nTaps-1
// history blocks per run), and the polyphase sum is evaluated in FP32 in the
// reference's order (oldest block first, src/lib/src/PPFChanneliser.cpp:
// 276-289), each tap one fused multiply-add.  The FIR and the FFT are the
// one part of the chain held to a tolerance rather than to bits (the
// reference's FFTW butterflies cannot be reproduced), so they contract
// multiply-adds: one rounding instead of two, and a third fewer VALU
// instructions in this VALU-bound kernel.  Stokes-I and the integration keep
// the reference's separate roundings.  The N-point forward DFT is a
// Cooley-Tukey split N = R x G: an R-point DFT in registers, twiddles, then
// a G-point radix-2 DIF FFT across lanes with xor shuffles.  Stokes-I
// (src/lib/src/StokesGenerator.cpp:93-99) and the integration over
// `integrate` consecutive blocks (src/lib/src/StokesIntegrator.cpp:97-114)
// are fused behind the FFT, so only int16 voltages are read and only the
// integrated Stokes-I spectra are written.
#include "ppl_internal.h"
#include "ppl_ppf_batch.h"

#include <cmath>
#include <numeric>
#include <vector>

struct ppl_ppf_plan {
    ppl_ppf_config cfg;
    float* d_coeffs = nullptr;    // [nTaps][nCh]
    float* d_twiddle = nullptr;   // W_N^k, k < N, (re, im) pairs
    uint32_t* d_hist[2] = {nullptr, nullptr};  // [sub][pol][nTaps-1][nCh] complex int16
    float* d_histf[2] = {nullptr, nullptr};    // same for complex float input
    int cur = 0;                  // which history buffer holds the latest blocks
    void* mfma = nullptr;         // split-precision DFT state (ppl_channelise_mfma.hip)
};

namespace {

using namespace ppl;

struct C2 {
    float re, im;
};

__device__ __forceinline__ C2 load_sample(const uint32_t* __restrict__ p, long idx)
{
    const uint32_t v = p[idx];
    const int16_t re = (int16_t)(v & 0xffffu);
    const int16_t im = (int16_t)(v >> 16);
    return {(float)re, (float)im};
}
__device__ __forceinline__ C2 load_sample(const float2* __restrict__ p, long idx)
{
    const float2 v = p[idx];
    return {v.x, v.y};
}
// The same in two steps: the raw load (held in flight a step ahead) and the
// conversion at use.
__device__ __forceinline__ uint32_t load_raw(const uint32_t* __restrict__ p, long idx) { return p[idx]; }
__device__ __forceinline__ float2 load_raw(const float2* __restrict__ p, long idx) { return p[idx]; }
__device__ __forceinline__ C2 to_c2(uint32_t v) { return {(float)(int16_t)(v & 0xffffu), (float)(int16_t)(v >> 16)}; }
__device__ __forceinline__ C2 to_c2(float2 v) { return {v.x, v.y}; }

template <typename Tin> struct HistT;
template <> struct HistT<uint32_t> { using type = uint32_t; };
template <> struct HistT<float2> { using type = float2; };

// Value of lane l ^ M.  M = 1, 2 and 8 are DPP moves (quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_ror:8); the rest go through ds_bpermute.
template <int M>
__device__ __forceinline__ float xor_lane(float v)
{
    if constexpr (M == 1)
        return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, true));
    else if constexpr (M == 2)
        return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, true));
    else if constexpr (M == 8)
        return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x128, 0xF, 0xF, true));
    else
        return __shfl_xor(v, M, 64);
}

// Per-lane constants of the lane FFT, stage s (m = G/2 >> s): sgn = +1 on
// the lane that keeps the sum (lane & m == 0), -1 on its partner; the
// partner's twiddle W_{2m}^j, j = lane & (m-1), and (1, 0) on the sum lane.
struct LaneTw {
    float sgn[6], twr[6], twi[6];
};

// G-point radix-2 decimation-in-frequency FFT across the lanes of a lane
// group; lane l ends with X[bitrev(l)].  Branch-free butterflies: both lanes
// form pr + sgn*x in one rounding (the sum, or the partner's difference),
// then multiply by their twiddle ((1, 0) on the sum lane, an exact no-op up
// to the sign of a zero).  The last stage's twiddle is 1 on every lane.
// Every fused multiply-add is written out (the file is built without
// contraction), so each kernel built from this code rounds alike.
template <int G, int S = 0>
__device__ __forceinline__ void lane_fft(float& re, float& im, const LaneTw& L)
{
    constexpr int m = G >> (S + 1);
    if constexpr (m >= 1) {
        const float pr = xor_lane<m>(re);
        const float pi = xor_lane<m>(im);
        const float dr = __builtin_fmaf(L.sgn[S], re, pr);
        const float di = __builtin_fmaf(L.sgn[S], im, pi);
        if constexpr (m == 1) {
            re = dr;
            im = di;
        } else {
            re = __builtin_fmaf(dr, L.twr[S], -(di * L.twi[S]));
            im = __builtin_fmaf(dr, L.twi[S], di * L.twr[S]);
        }
        lane_fft<G, S + 1>(re, im, L);
    }
}

template <int G>
__device__ __forceinline__ int bitrev(int l)
{
    int r = 0;
#pragma unroll
    for (int b = 1; b < G; b <<= 1) {
        r <<= 1;
        r |= (l & b) ? 1 : 0;
    }
    return r;
}

// R-point DFT in registers (R = 1, 2, 4), forward.
template <int R>
__device__ __forceinline__ void reg_dft(C2* x)
{
    if constexpr (R == 2) {
        const C2 a = x[0], b = x[1];
        x[0] = {a.re + b.re, a.im + b.im};
        x[1] = {a.re - b.re, a.im - b.im};
    } else if constexpr (R == 4) {
        const C2 a = x[0], b = x[1], c = x[2], d = x[3];
        const C2 s02 = {a.re + c.re, a.im + c.im}, d02 = {a.re - c.re, a.im - c.im};
        const C2 s13 = {b.re + d.re, b.im + d.im}, d13 = {b.re - d.re, b.im - d.im};
        x[0] = {s02.re + s13.re, s02.im + s13.im};
        x[2] = {s02.re - s13.re, s02.im - s13.im};
        // X1 = d02 - i d13 ; X3 = d02 + i d13
        x[1] = {d02.re + d13.im, d02.im - d13.re};
        x[3] = {d02.re - d13.im, d02.im + d13.re};
    }
}

// Wide channelisers with long filters hold TAPS x N/64 complex samples per
// lane and polarisation in their FIR windows; where both polarisations'
// windows would take more than about 128 registers (two waves per SIMD), each
// polarisation of a run gets a wave of its own instead (pol split).
constexpr bool pol_split(int N, int TAPS) { return (N < 64 ? 1 : N / 64) * TAPS >= 32; }

// OUT: 0 = integrated Stokes-I [nBlk/integ][sub][1][N]; 1 = complex spectra
// [nBlk][sub][2][N].  TAPS is the FIR length (compile-time, so the sliding
// windows stay in registers).
// The runs of one wave: waveId counts the waves of the launch (or of one
// job of a batched launch); `stride` is the samples per (sub, pol) stream of
// `in`, of which the first nBlk*N are read.
// PS (pol split, Stokes-I only, N >= 64): waveId counts pairs of waves, two
// pairs per workgroup; wave (threadIdx.x >> 6) & 1 of a pair channelises that
// polarisation of the pair's run, and the second hands its power to the
// first through LDS at every block, where I = powerX + powerY is summed as
// in the one-wave form.  The workgroup meets at a barrier every block.
template <typename Tin, int N, int TAPS, int OUT, bool PS = false>
__device__ __forceinline__ void ppf_wave(const Tin* __restrict__ in, const typename HistT<Tin>::type* __restrict__ hist,
                                         float* __restrict__ out, const float* __restrict__ coeffs,
                                         const float* __restrict__ tw, int nSub, long nBlk, long stride, int runLen,
                                         int nRuns, int integ, long waveId)
{
    constexpr int G = N < 64 ? N : 64;
    constexpr int R = N / G;
    constexpr int RUNS_PER_WAVE = 64 / G;
    constexpr int H = TAPS - 1;
    static_assert(!PS || (OUT == 0 && G == 64), "pol split: Stokes-I, one run per pair of waves");
    const int lane = threadIdx.x & 63;
    const int grp = lane / G;
    const int l = lane % G;
    const int pol = PS ? (int)((threadIdx.x >> 6) & 1) : 0;  // the wave's polarisation (pol split)
    const long wavesPerSub = (nRuns + RUNS_PER_WAVE - 1) / RUNS_PER_WAVE;
    const int sub = (int)(waveId / wavesPerSub);
    if constexpr (!PS) {
        if (sub >= nSub) return;
    }
    // a pair past the last subband still meets its workgroup's barriers
    const bool subOk = sub < nSub;
    const int sb = subOk ? sub : 0;
    const int run0 = (int)(waveId % wavesPerSub) * RUNS_PER_WAVE;  // the wave's first run
    const int run = run0 + grp;
    const bool active = subOk && run < nRuns;
    const long T = stride;
    // pol split: x0 and h0 are the wave's polarisation, x1 and h1 unused
    const Tin* x0 = in + (long)(2 * sb + pol) * T;
    const Tin* x1 = in + (long)(2 * sb + 1) * T;
    const auto* h0 = hist + (long)(2 * sb + pol) * H * N;
    const auto* h1 = hist + (long)(2 * sb + 1) * H * N;

    // coefficients of this lane's channels, twiddles of this lane
    float hc[R][TAPS];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int t = 0; t < TAPS; ++t) hc[r][t] = coeffs[t * N + l + G * r];
    LaneTw ltw;
    {
        int s = 0;
#pragma unroll
        for (int m = G / 2; m >= 1; m >>= 1, ++s) {
            const int j = l & (m - 1);
            const int k = j * (N / (2 * m));  // W_{2m}^j = W_N^{j N/(2m)}
            const bool sumLane = (l & m) == 0;
            ltw.sgn[s] = sumLane ? 1.0f : -1.0f;
            ltw.twr[s] = sumLane ? 1.0f : tw[2 * k];
            ltw.twi[s] = sumLane ? 0.0f : tw[2 * k + 1];
        }
    }
    C2 rtw[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = (l * r) % N;
        rtw[r] = {tw[2 * k], tw[2 * k + 1]};
    }

    // The FIR window is a ring of TAPS blocks per polarisation: step s writes
    // its new block into slot s % TAPS, and tap t (oldest first) is slot
    // (s + 1 + t) % TAPS.  The step loop is unrolled by TAPS, so every slot
    // index is a compile-time constant and the window never moves between
    // registers.  Before the first step slots 1..TAPS-1 hold the TAPS-1
    // blocks preceding the run (from this chunk or from the stream's saved
    // history).
    C2 w0[TAPS][R], w1[TAPS][R];
    const long b0 = (long)run * runLen;
#pragma unroll
    for (int t = 0; t < TAPS; ++t)
#pragma unroll
        for (int r = 0; r < R; ++r) w0[t][r] = w1[t][r] = {0.f, 0.f};
#pragma unroll
    for (int t = 0; t < H; ++t) {
        const long b = b0 - H + t;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int n = l + G * r;
            C2 a = {0.f, 0.f}, c = {0.f, 0.f};
            if (active) {
                if (b >= 0) {
                    a = load_sample(x0, b * N + n);
                    if constexpr (!PS) c = load_sample(x1, b * N + n);
                } else {
                    a = load_sample(h0, (long)(H + b) * N + n);
                    if constexpr (!PS) c = load_sample(h1, (long)(H + b) * N + n);
                }
            }
            w0[t + 1][r] = a;
            w1[t + 1][r] = c;
        }
    }
    float acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0f;
    const int lastRunLen = (int)(nBlk - (long)(nRuns - 1) * runLen);
    const int len = active ? (run == nRuns - 1 ? lastRunLen : runLen) : 0;
    // steps of the wave (uniform): a full run unless its only run is the last
    int waveLen = run0 < nRuns - 1 ? runLen : lastRunLen;
    if constexpr (PS) {
        // both pairs of the workgroup take the same steps
        const long first = waveId - ((threadIdx.x >> 7) & 1);
        auto steps = [&](long id) {
            return id / nRuns >= nSub ? 0 : (int)(id % nRuns) == nRuns - 1 ? lastRunLen : runLen;
        };
        waveLen = max(steps(first), steps(first + 1));
    }
    const int kLane = bitrev<G>(l);
    int inWindow = 0;  // blocks summed into the current integration window
    // output row of the current window: b0 is a multiple of integ (runLen is),
    // so row (b0 + s) / integ is counted instead of divided
    long uo = OUT == 0 ? b0 / integ : 0;
    // every lane runs every step: the lane FFT shuffles need the whole wave.
    // Steps past a run's length (a shorter last run, and up to the next
    // multiple of TAPS) are not live: they store nothing.
    // Every lane loads, from a block clamped into the chunk: a load under a
    // branch would be waited for on its own.  The steps past a run's length
    // follow all of its live steps and store nothing (integration windows
    // never straddle the end of a run), so what they load reaches no output.
    // A step's samples are loaded a step ahead, so the load is in flight
    // while the previous block is transformed.
    using Raw = decltype(load_raw(x0, 0));
    Raw n0[R], n1[R];
    auto load_step = [&](int s) {
        const long b = b0 + s;
        const long bl = b < nBlk ? b : nBlk - 1;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int n = l + G * r;
            n0[r] = load_raw(x0, bl * N + n);
            if constexpr (!PS) n1[r] = load_raw(x1, bl * N + n);
        }
    };
    load_step(0);
    for (int s0 = 0; s0 < waveLen; s0 += TAPS) {
#pragma unroll
        for (int u = 0; u < TAPS; ++u) {
            const int s = s0 + u;
            const bool live = s < len;
            const long b = b0 + s;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                w0[u][r] = to_c2(n0[r]);
                if constexpr (!PS) w1[u][r] = to_c2(n1[r]);
            }
            load_step(s + 1);
            // FIR, oldest block first, coefficient index t*N + n, one
            // multiply-add per tap (written out, as in lane_fft)
            C2 f0[R], f1[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                f0[r] = f1[r] = {0.f, 0.f};
#pragma unroll
                for (int t = 0; t < TAPS; ++t) {
                    const int slot = (u + 1 + t) % TAPS;
                    const float coeff = hc[r][t];
                    f0[r].re = __builtin_fmaf(w0[slot][r].re, coeff, f0[r].re);
                    f0[r].im = __builtin_fmaf(w0[slot][r].im, coeff, f0[r].im);
                    f1[r].re = __builtin_fmaf(w1[slot][r].re, coeff, f1[r].re);
                    f1[r].im = __builtin_fmaf(w1[slot][r].im, coeff, f1[r].im);
                }
            }
            // DFT: registers, twiddle, lanes
            reg_dft<R>(f0);
            reg_dft<R>(f1);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const C2 a = f0[r], c = f1[r], t = rtw[r];
                f0[r] = {__builtin_fmaf(a.re, t.re, -(a.im * t.im)), __builtin_fmaf(a.re, t.im, a.im * t.re)};
                f1[r] = {__builtin_fmaf(c.re, t.re, -(c.im * t.im)), __builtin_fmaf(c.re, t.im, c.im * t.re)};
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                lane_fft<G>(f0[r].re, f0[r].im, ltw);
                lane_fft<G>(f1[r].re, f1[r].im, ltw);
            }
            if constexpr (OUT == 1) {
                if (live) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int k = r + R * kLane;
                        float2* o = reinterpret_cast<float2*>(out) + ((long)(b * nSub + sub) * 2) * N + k;
                        o[0] = make_float2(f0[r].re, f0[r].im);
                        o[N] = make_float2(f1[r].re, f1[r].im);
                    }
                }
            } else if constexpr (PS) {
                // the second wave's power to the first; two buffers by step
                // parity, so one barrier per step orders every write and read
                __shared__ float xch[2][2][R][64];
                const int pr = (threadIdx.x >> 7) & 1, par = s & 1;
                float pw[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    pw[r] = f0[r].re * f0[r].re + f0[r].im * f0[r].im;
                    if (pol == 1) xch[pr][par][r][lane] = pw[r];
                }
                __syncthreads();
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const float I = pw[r] + xch[pr][par][r][lane];  // powerX + powerY on the first wave
                    acc[r] = acc[r] + I;
                }
                if (++inWindow == integ) {  // (s + 1) % integ == 0
                    inWindow = 0;
                    if (live && pol == 0) {
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            const int k = r + R * kLane;
                            out[(uo * nSub + sub) * N + k] = acc[r];
                        }
                    }
                    ++uo;
#pragma unroll
                    for (int r = 0; r < R; ++r) acc[r] = 0.0f;
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const float powerX = f0[r].re * f0[r].re + f0[r].im * f0[r].im;
                    const float powerY = f1[r].re * f1[r].re + f1[r].im * f1[r].im;
                    const float I = powerX + powerY;
                    acc[r] = acc[r] + I;
                }
                if (++inWindow == integ) {  // (s + 1) % integ == 0
                    inWindow = 0;
                    if (live) {
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            const int k = r + R * kLane;
                            out[(uo * nSub + sub) * N + k] = acc[r];
                        }
                    }
                    ++uo;
#pragma unroll
                    for (int r = 0; r < R; ++r) acc[r] = 0.0f;
                }
            }
        }
    }
}

template <typename Tin, int N, int TAPS, int OUT>
__global__ __launch_bounds__(256) void ppf_kernel(
    const Tin* __restrict__ in, const typename HistT<Tin>::type* __restrict__ hist, float* __restrict__ out,
    const float* __restrict__ coeffs, const float* __restrict__ tw, int nSub, long nBlk, int runLen,
    int nRuns, int integ)
{
    const long waveId = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    ppf_wave<Tin, N, TAPS, OUT>(in, hist, out, coeffs, tw, nSub, nBlk, nBlk * N, runLen, nRuns, integ, waveId);
}

// The pol-split form of ppf_kernel (Stokes-I): two runs per workgroup.
template <typename Tin, int N, int TAPS>
__global__ __launch_bounds__(256) void ppf_pair_kernel(
    const Tin* __restrict__ in, const typename HistT<Tin>::type* __restrict__ hist, float* __restrict__ out,
    const float* __restrict__ coeffs, const float* __restrict__ tw, int nSub, long nBlk, int runLen,
    int nRuns, int integ)
{
    const long pairId = (long)blockIdx.x * 2 + (threadIdx.x >> 7);
    ppf_wave<Tin, N, TAPS, 0, true>(in, hist, out, coeffs, tw, nSub, nBlk, nBlk * N, runLen, nRuns, integ, pairId);
}

// Many plans' fused PPF -> Stokes-I -> integration of one channel count and
// tap count in one launch: each workgroup finds its job in the table.
template <int N, int TAPS>
__global__ __launch_bounds__(256) void ppf_batch_kernel(const PPFBatch B)
{
    int local;
    const PPFJob& J = find_batch_job(B, local);
    if constexpr (pol_split(N, TAPS)) {
        const long pairId = (long)local * 2 + (threadIdx.x >> 7);
        ppf_wave<uint32_t, N, TAPS, 0, true>(J.in, J.hist, J.out, J.coeffs, B.tw, J.nSub, J.nBlk, J.stride,
                                             J.runLen, J.nRuns, J.integ, pairId);
    } else {
        const long waveId = (long)local * 4 + (threadIdx.x >> 6);
        ppf_wave<uint32_t, N, TAPS, 0>(J.in, J.hist, J.out, J.coeffs, B.tw, J.nSub, J.nBlk, J.stride, J.runLen,
                                       J.nRuns, J.integ, waveId);
    }
}

// Save the last nTaps-1 blocks of every (sub, pol) stream (new history),
// taking them from the chunk or, for short chunks, from the old history.
template <typename Tin>
__global__ void save_history_kernel(const Tin* __restrict__ in, const Tin* __restrict__ oldh,
                                    Tin* __restrict__ newh, int nStreams, int H, int N, long nBlk)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = (long)H * N;
    if (i >= per * nStreams) return;
    const long stream = i / per;
    const long j = i % per;
    const long slot = j / N, n = j % N;
    const long b = nBlk - H + slot;
    newh[i] = b >= 0 ? in[stream * nBlk * N + b * N + n] : oldh[stream * per + (H + b) * N + n];
}

// History update of every job of a batched launch.
__global__ void save_history_batch_kernel(const HistBatch B)
{
    int local;
    const HistJob& J = find_batch_job(B, local);
    const long i = (long)local * blockDim.x + threadIdx.x;
    const long per = (long)J.H * J.N;
    if (i >= per * J.nStreams) return;
    const long stream = i / per;
    const long j = i % per;
    const long slot = j / J.N, n = j % J.N;
    const long b = J.nBlk - J.H + slot;
    J.newh[i] = b >= 0 ? J.in[stream * J.stride + b * J.N + n] : J.oldh[stream * per + (J.H + b) * J.N + n];
}

// The Stokes-I launch of one channel count and tap count: the pol-split
// form where it applies (blocks: one wave per run; pairBlocks: a pair).
template <typename Tin, int N, int TAPS>
void launch_stokes_i(const Tin* in, const typename HistT<Tin>::type* hist, float* out, const float* coeffs,
                     const float* tw, int S, long nBlk, int L, int nRuns, int integ, unsigned blocks,
                     unsigned pairBlocks, hipStream_t st)
{
    if constexpr (pol_split(N, TAPS))
        hipLaunchKernelGGL((ppf_pair_kernel<Tin, N, TAPS>), dim3(pairBlocks), dim3(256), 0, st, in, hist, out,
                           coeffs, tw, S, nBlk, L, nRuns, integ);
    else
        hipLaunchKernelGGL((ppf_kernel<Tin, N, TAPS, 0>), dim3(blocks), dim3(256), 0, st, in, hist, out, coeffs,
                           tw, S, nBlk, L, nRuns, integ);
}

template <typename Tin>
int run_ppf(ppl_ppf_plan* p, const Tin* in, typename HistT<Tin>::type* histCur,
            typename HistT<Tin>::type* histNext, unsigned nBlk, float* out, int mode, hipStream_t st)
{
    const int N = (int)p->cfg.nCh, S = (int)p->cfg.nSub, T = (int)p->cfg.nTaps;
    const int integ = mode == 0 ? (int)p->cfg.integrate : 1;
    PPL_CHECK_ARG(nBlk > 0, "Empty time data blob");
    PPL_CHECK_ARG(nBlk % integ == 0, "nBlk (%u) must be a multiple of the integration (%d)", nBlk, integ);
    // runs of L blocks, L a multiple of the integration window and of the tap count
    const int G = N < 64 ? N : 64;
    const long nb = nBlk;
    const int mult = std::lcm(integ, T);
    // pol split: two waves per run, so half the runs fill the GPU
    const bool ps = mode == 0 && pol_split(N, T);
    const int L = ppf_run_length(&nb, &S, 1, 64 / G, ps ? kFillWaves / 2 : kFillWaves, mult, 4 * T);
    const int nRuns = (int)cdiv(nBlk, L);
    const long wavesPerSub = (nRuns + (64 / G) - 1) / (64 / G);
    const long waves = wavesPerSub * S;
    const unsigned blocks = (unsigned)cdiv((unsigned)waves, 4);
    const unsigned pairBlocks = (unsigned)cdiv((unsigned)waves, 2);  // pol split: a run per pair of waves
#define PPL_PPF_LAUNCH(N_, T_)                                                                          \
    if (N == N_ && T == T_) {                                                                           \
        if (mode == 0)                                                                                  \
            launch_stokes_i<Tin, N_, T_>(in, histCur, out, p->d_coeffs, p->d_twiddle, S, (long)nBlk, L,   \
                                         nRuns, integ, blocks, pairBlocks, st);                         \
        else                                                                                            \
            hipLaunchKernelGGL((ppf_kernel<Tin, N_, T_, 1>), dim3(blocks), dim3(256), 0, st, in, histCur,\
                               out, p->d_coeffs, p->d_twiddle, S, (long)nBlk, L, nRuns, 1);            \
        PPL_LAUNCHED();                                                                                 \
        launched = true;                                                                                \
    }
#define PPL_PPF_CASE(N_) PPL_PPF_LAUNCH(N_, 4) PPL_PPF_LAUNCH(N_, 8) PPL_PPF_LAUNCH(N_, 16)
    bool launched = false;
    PPL_PPF_CASE(8)
    PPL_PPF_CASE(16)
    PPL_PPF_CASE(32)
    PPL_PPF_CASE(64)
    PPL_PPF_CASE(128)
    PPL_PPF_CASE(256)
#undef PPL_PPF_CASE
#undef PPL_PPF_LAUNCH
    if (!launched) {
        set_error("channeliser: no kernel for %d channels x %d taps", N, T);
        return PPL_E_ARG;
    }
    const int H = T - 1;
    if (H > 0) {
        const long n = (long)S * 2 * H * N;
        hipLaunchKernelGGL((save_history_kernel<typename HistT<Tin>::type>), dim3((unsigned)cdiv((unsigned)n, 256)),
                           dim3(256), 0, st, reinterpret_cast<const typename HistT<Tin>::type*>(in),
                           histCur, histNext, S * 2, H, N, (long)nBlk);
        PPL_LAUNCHED();
    }
    return PPL_OK;
}

// FP32 batched launch of a group of jobs sharing N and TAPS: sets each
// job's runs and first workgroup.
template <int N, int TAPS>
int launch_fp32_batch(PPFBatch& B, hipStream_t st)
{
    constexpr int G = N < 64 ? N : 64;
    constexpr int RPW = 64 / G;
    long nb[kMaxPPFJobs];
    int ns[kMaxPPFJobs];
    int mult = TAPS;
    for (int k = 0; k < B.n; ++k) {
        nb[k] = B.job[k].nBlk;
        ns[k] = B.job[k].nSub;
        mult = std::lcm(mult, B.job[k].integ);
    }
    // pol split: two waves per run, so half the runs fill the GPU
    constexpr bool PS = pol_split(N, TAPS);
    const int L = ppf_run_length(nb, ns, B.n, RPW, PS ? kFillWaves / 2 : kFillWaves, mult, 4 * TAPS);
    long blocks = 0;
    for (int k = 0; k < B.n; ++k) {
        PPFJob& J = B.job[k];
        J.runLen = L;
        J.nRuns = (J.nBlk + L - 1) / L;
        const long waves = (long)J.nSub * ((J.nRuns + RPW - 1) / RPW);  // pairs of waves under pol split
        J.blockStart = (int)blocks;
        blocks += PS ? (waves + 1) / 2 : (waves + 3) / 4;
        if (blocks >= (1L << 31)) {
            set_error("channeliser batch: %ld workgroups exceed the grid", blocks);
            return PPL_E_ARG;
        }
    }
    hipLaunchKernelGGL((ppf_batch_kernel<N, TAPS>), dim3((unsigned)blocks), dim3(256), 0, st, B);
    PPL_LAUNCHED();
    return PPL_OK;
}

int launch_fp32_batch_any(PPFBatch& B, int N, int taps, hipStream_t st)
{
#define PPL_PPFB_CASE(N_)                                           \
    if (N == N_) {                                                  \
        if (taps == 4) return launch_fp32_batch<N_, 4>(B, st);      \
        if (taps == 8) return launch_fp32_batch<N_, 8>(B, st);      \
        if (taps == 16) return launch_fp32_batch<N_, 16>(B, st);    \
    }
    PPL_PPFB_CASE(8)
    PPL_PPFB_CASE(16)
    PPL_PPFB_CASE(32)
    PPL_PPFB_CASE(64)
    PPL_PPFB_CASE(128)
    PPL_PPFB_CASE(256)
#undef PPL_PPFB_CASE
    set_error("channeliser: no kernel for %d channels x %d taps", N, taps);
    return PPL_E_ARG;
}

}  // namespace

int ppl::ppf_run_length(const long* nBlk, const int* nSub, int n, int unitsPerRun, long fill, int mult, int minLen)
{
    auto units = [&](long L) {
        long u = 0;
        for (int j = 0; j < n; ++j) u += (long)nSub[j] * (((nBlk[j] + L - 1) / L + unitsPerRun - 1) / unitsPerRun);
        return u;
    };
    auto round_up = [&](long v) { return (v + mult - 1) / mult * mult; };
    long L = round_up(256);
    const long floorLen = round_up(minLen);
    while (units(L) < fill) {
        const long h = round_up(L / 2);
        if (h >= L || h < floorLen) break;
        L = h;
    }
    return (int)L;
}

namespace {

bool pow2_supported(unsigned n) { return n >= 8 && n <= 256 && (n & (n - 1)) == 0; }

// Stokes parameters from [nBlk][nSub][2][nCh] complex spectra,
// src/lib/src/StokesGenerator.cpp:88-104.
__global__ void stokes_kernel(const float2* __restrict__ in, float* __restrict__ out, long nSpec,
                              int nCh, int nStokes)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;  // over nBlk*nSub*nCh
    if (i >= nSpec * nCh) return;
    const long spec = i / nCh;
    const int c = (int)(i % nCh);
    const float2 X = in[(spec * 2 + 0) * nCh + c];
    const float2 Y = in[(spec * 2 + 1) * nCh + c];
    const float Xr = X.x, Xi = X.y, Yr = Y.x, Yi = Y.y;
    const float XxYstarReal = Xr * Yr + Xi * Yi;
    const float XxYstarImag = Xi * Yr - Xr * Yi;
    const float powerX = Xr * Xr + Xi * Xi;
    const float powerY = Yr * Yr + Yi * Yi;
    float* o = out + spec * nStokes * nCh + c;
    o[0] = powerX + powerY;
    if (nStokes == 4) {
        o[nCh] = powerX - powerY;
        o[2 * nCh] = 2.0f * XxYstarReal;
        o[3 * nCh] = 2.0f * XxYstarImag;
    }
}

// src/lib/src/StokesIntegrator.cpp:97-114: out starts at 0 and adds, for
// each time bin of the window in order, the binCh channels of the bin.
__global__ void integrate_kernel(const float* __restrict__ in, float* __restrict__ out, long newT,
                                 int nSub, int nPol, int nCh, int window, int binCh)
{
    const int newC = nCh / binCh;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;  // over newT*nSub*nPol*newC
    const long total = newT * nSub * nPol * newC;
    if (i >= total) return;
    const int nc = (int)(i % newC);
    const long rest = i / newC;  // (u*nSub + s)*nPol + p
    const int p = (int)(rest % nPol);
    const long us = rest / nPol;
    const int s = (int)(us % nSub);
    const long u = us / nSub;
    float acc = 0.0f;
    for (int w = 0; w < window; ++w) {
        const long t = u * window + w;
        const float* v = in + ((t * nSub + s) * nPol + p) * nCh + (long)nc * binCh;
        for (int c = 0; c < binCh; ++c) acc = acc + v[c];
    }
    out[i] = acc;
}

}  // namespace

// split-precision DFT path (ppl_channelise_mfma.hip)
int ppl_mfma_create(ppl_ppf_plan* p);
void ppl_mfma_destroy(ppl_ppf_plan* p);
int ppl_mfma_run(ppl_ppf_plan* p, const int16_t* d_in, const uint32_t* histCur, unsigned nBlk,
                 float* d_out, int mode, hipStream_t st);
int ppl_mfma_run_batch(ppl::PPFBatch& B, int nCh, int taps, int prec, hipStream_t st);

extern "C" int ppl_ppf_plan_create(ppl_ppf_plan** plan, const ppl_ppf_config* cfg)
{
    PPL_CHECK_ARG(plan && cfg && cfg->coeffs, "null argument");
    PPL_CHECK_ARG(cfg->nPol == 2, "the channeliser kernels take 2 polarisations (got %u)", cfg->nPol);
    PPL_CHECK_ARG(pow2_supported(cfg->nCh),
                  "outputChannelsPerSubband %u: a power of two in [8, 256] is required", cfg->nCh);
    PPL_CHECK_ARG(cfg->nTaps == 4 || cfg->nTaps == 8 || cfg->nTaps == 16,
                  "nTaps %u: the channeliser is built for 4, 8 or 16 taps", cfg->nTaps);
    PPL_CHECK_ARG(cfg->nSub >= 1, "no subbands");
    PPL_CHECK_ARG(cfg->integrate >= 1, "integrateTimeBins must be >= 1");
    PPL_CHECK_ARG(cfg->precision == PPL_DFT_FP32 || cfg->precision == PPL_DFT_BF16X3 ||
                      cfg->precision == PPL_DFT_FP16X2,
                  "unknown precision %d", cfg->precision);
    ppl_ppf_plan* p = new ppl_ppf_plan;
    p->cfg = *cfg;
    p->cfg.coeffs = nullptr;
    const unsigned N = cfg->nCh, T = cfg->nTaps;
    std::vector<float> tw(2 * N);
    for (unsigned k = 0; k < N; ++k) {
        const double a = -2.0 * M_PI * (double)k / (double)N;
        tw[2 * k] = (float)std::cos(a);
        tw[2 * k + 1] = (float)std::sin(a);
    }
    auto fail = [&](int rc) {
        ppl_ppf_plan_destroy(p);
        return rc;
    };
    if (hipMalloc(&p->d_coeffs, sizeof(float) * N * T) != hipSuccess ||
        hipMalloc(&p->d_twiddle, sizeof(float) * 2 * N) != hipSuccess) {
        set_error("channeliser: device allocation failed");
        return fail(PPL_E_NOMEM);
    }
    const size_t hist = (size_t)cfg->nSub * 2 * (T > 1 ? T - 1 : 1) * N;
    for (int i = 0; i < 2; ++i) {
        if (hipMalloc(&p->d_hist[i], hist * sizeof(uint32_t)) != hipSuccess ||
            hipMalloc(&p->d_histf[i], hist * sizeof(float2)) != hipSuccess) {
            set_error("channeliser: history allocation failed");
            return fail(PPL_E_NOMEM);
        }
        if (hipMemset(p->d_hist[i], 0, hist * sizeof(uint32_t)) != hipSuccess ||
            hipMemset(p->d_histf[i], 0, hist * sizeof(float2)) != hipSuccess)
            return fail(PPL_E_HIP);
    }
    if (hipMemcpy(p->d_coeffs, cfg->coeffs, sizeof(float) * N * T, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->d_twiddle, tw.data(), sizeof(float) * 2 * N, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("channeliser: upload failed");
        return fail(PPL_E_HIP);
    }
    if (cfg->precision != PPL_DFT_FP32) {
        int rc = ppl_mfma_create(p);
        if (rc != PPL_OK) return fail(rc);
    }
    *plan = p;
    return PPL_OK;
}

extern "C" int ppl_ppf_plan_destroy(ppl_ppf_plan* p)
{
    if (!p) return PPL_OK;
    if (p->mfma) ppl_mfma_destroy(p);
    (void)hipFree(p->d_coeffs);
    (void)hipFree(p->d_twiddle);
    for (int i = 0; i < 2; ++i) {
        (void)hipFree(p->d_hist[i]);
        (void)hipFree(p->d_histf[i]);
    }
    delete p;
    return PPL_OK;
}

extern "C" int ppl_ppf_reset(ppl_ppf_plan* p, void* stream)
{
    PPL_CHECK_ARG(p, "null plan");
    const size_t hist = (size_t)p->cfg.nSub * 2 * (p->cfg.nTaps > 1 ? p->cfg.nTaps - 1 : 1) * p->cfg.nCh;
    for (int i = 0; i < 2; ++i) {
        PPL_HIP(hipMemsetAsync(p->d_hist[i], 0, hist * sizeof(uint32_t), stream_of(stream)));
        PPL_HIP(hipMemsetAsync(p->d_histf[i], 0, hist * sizeof(float2), stream_of(stream)));
    }
    return PPL_OK;
}

extern "C" int ppl_ppf_run_i16(ppl_ppf_plan* p, const int16_t* d_in, unsigned nBlk, float* d_spectra,
                               void* stream)
{
    PPL_CHECK_ARG(p && d_in && d_spectra, "null argument");
    int rc;
    if (p->cfg.precision != PPL_DFT_FP32)
        rc = ppl_mfma_run(p, d_in, p->d_hist[p->cur], nBlk, d_spectra, 1, stream_of(stream));
    else
        rc = run_ppf<uint32_t>(p, reinterpret_cast<const uint32_t*>(d_in), p->d_hist[p->cur],
                               p->d_hist[p->cur ^ 1], nBlk, d_spectra, 1, stream_of(stream));
    if (rc == PPL_OK) p->cur ^= 1;
    return rc;
}

extern "C" int ppl_ppf_run_c32(ppl_ppf_plan* p, const float* d_in, unsigned nBlk, float* d_spectra,
                               void* stream)
{
    PPL_CHECK_ARG(p && d_in && d_spectra, "null argument");
    PPL_CHECK_ARG(p->cfg.precision == PPL_DFT_FP32, "complex-float input runs the FP32 method only");
    float2* hc = reinterpret_cast<float2*>(p->d_histf[p->cur]);
    float2* hn = reinterpret_cast<float2*>(p->d_histf[p->cur ^ 1]);
    int rc = run_ppf<float2>(p, reinterpret_cast<const float2*>(d_in), hc, hn, nBlk, d_spectra, 1,
                             stream_of(stream));
    if (rc == PPL_OK) p->cur ^= 1;
    return rc;
}

extern "C" int ppl_ppf_stokes_i16(ppl_ppf_plan* p, const int16_t* d_in, unsigned nBlk, float* d_stokes,
                                  void* stream)
{
    PPL_CHECK_ARG(p && d_in && d_stokes, "null argument");
    int rc;
    if (p->cfg.precision != PPL_DFT_FP32)
        rc = ppl_mfma_run(p, d_in, p->d_hist[p->cur], nBlk, d_stokes, 0, stream_of(stream));
    else
        rc = run_ppf<uint32_t>(p, reinterpret_cast<const uint32_t*>(d_in), p->d_hist[p->cur],
                               p->d_hist[p->cur ^ 1], nBlk, d_stokes, 0, stream_of(stream));
    if (rc == PPL_OK) p->cur ^= 1;
    return rc;
}

extern "C" int ppl_ppf_stokes_i16_batch(const ppl_ppf_job* jobs, int n, void* stream)
{
    PPL_CHECK_ARG(jobs && n >= 1 && n <= kMaxPPFJobs, "%d channeliser jobs (1..%d)", n, kMaxPPFJobs);
    const hipStream_t st = stream_of(stream);
    for (int k = 0; k < n; ++k) {
        const ppl_ppf_job& j = jobs[k];
        PPL_CHECK_ARG(j.plan, "job %d: null plan", k);
        if (!j.nBlk) continue;
        PPL_CHECK_ARG(j.d_in && j.d_stokes, "job %d: null argument", k);
        PPL_CHECK_ARG(j.nBlk < (1u << 31), "job %d: %u blocks", k, j.nBlk);
        PPL_CHECK_ARG(j.nBlk % j.plan->cfg.integrate == 0,
                      "job %d: nBlk (%u) must be a multiple of the integration (%u)", k, j.nBlk,
                      j.plan->cfg.integrate);
        PPL_CHECK_ARG(j.strideBlk == 0 || j.strideBlk >= j.nBlk, "job %d: stride of %u blocks below nBlk %u", k,
                      j.strideBlk, j.nBlk);
        for (int m = 0; m < k; ++m)
            PPL_CHECK_ARG(jobs[m].plan != j.plan || !jobs[m].nBlk,
                          "job %d: the plan of job %d again (a plan's chunks run in order)", k, m);
    }
    auto fill = [&](const ppl_ppf_job& j) {
        const ppl_ppf_plan* p = j.plan;
        PPFJob J{};
        J.in = reinterpret_cast<const uint32_t*>(j.d_in);
        J.hist = p->d_hist[p->cur];
        J.out = j.d_stokes;
        J.coeffs = p->d_coeffs;
        J.stride = (long)(j.strideBlk ? j.strideBlk : j.nBlk) * p->cfg.nCh;
        J.nBlk = (int)j.nBlk;
        J.nSub = (int)p->cfg.nSub;
        J.integ = (int)p->cfg.integrate;
        return J;
    };
    // one launch per (channel count, taps, DFT method)
    bool done[kMaxPPFJobs] = {};
    for (int k = 0; k < n; ++k) {
        if (done[k] || !jobs[k].nBlk) continue;
        const ppl_ppf_config& c = jobs[k].plan->cfg;
        PPFBatch B{};
        B.tw = jobs[k].plan->d_twiddle;
        for (int m = k; m < n; ++m) {
            const ppl_ppf_config& d = jobs[m].plan->cfg;
            if (done[m] || !jobs[m].nBlk || d.nCh != c.nCh || d.nTaps != c.nTaps || d.precision != c.precision)
                continue;
            B.job[B.n++] = fill(jobs[m]);
            done[m] = true;
        }
        const int rc = c.precision == PPL_DFT_FP32 ? launch_fp32_batch_any(B, (int)c.nCh, (int)c.nTaps, st)
                                                  : ppl_mfma_run_batch(B, (int)c.nCh, (int)c.nTaps, c.precision, st);
        if (rc != PPL_OK) return rc;
    }
    // every job's new history in one launch
    HistBatch H{};
    long blocks = 0;
    for (int k = 0; k < n; ++k) {
        const ppl_ppf_job& j = jobs[k];
        if (!j.nBlk) continue;
        const ppl_ppf_plan* p = j.plan;
        HistJob& h = H.job[H.n++];
        h.in = reinterpret_cast<const uint32_t*>(j.d_in);
        h.oldh = p->d_hist[p->cur];
        h.newh = p->d_hist[p->cur ^ 1];
        h.N = (int)p->cfg.nCh;
        h.H = (int)p->cfg.nTaps - 1;
        h.nStreams = 2 * (int)p->cfg.nSub;
        h.stride = (long)(j.strideBlk ? j.strideBlk : j.nBlk) * h.N;
        h.nBlk = j.nBlk;
        h.blockStart = (int)blocks;
        blocks += ((long)h.nStreams * h.H * h.N + 255) / 256;
    }
    if (!H.n) return PPL_OK;
    hipLaunchKernelGGL(save_history_batch_kernel, dim3((unsigned)blocks), dim3(256), 0, st, H);
    PPL_LAUNCHED();
    for (int k = 0; k < n; ++k)
        if (jobs[k].nBlk) jobs[k].plan->cur ^= 1;
    return PPL_OK;
}

// history buffers for the split-precision path
uint32_t* ppl_ppf_hist_next(ppl_ppf_plan* p) { return p->d_hist[p->cur ^ 1]; }
const ppl_ppf_config& ppl_ppf_cfg(const ppl_ppf_plan* p) { return p->cfg; }
const float* ppl_ppf_coeffs(const ppl_ppf_plan* p) { return p->d_coeffs; }
const float* ppl_ppf_twiddle(const ppl_ppf_plan* p) { return p->d_twiddle; }
void*& ppl_ppf_mfma_slot(ppl_ppf_plan* p) { return p->mfma; }
int ppl_ppf_save_history(ppl_ppf_plan* p, const int16_t* d_in, unsigned nBlk, hipStream_t st)
{
    const int N = (int)p->cfg.nCh, S = (int)p->cfg.nSub, H = (int)p->cfg.nTaps - 1;
    if (H <= 0) return PPL_OK;
    const long n = (long)S * 2 * H * N;
    hipLaunchKernelGGL(save_history_kernel<uint32_t>, dim3((unsigned)cdiv((unsigned)n, 256)), dim3(256), 0, st,
                       reinterpret_cast<const uint32_t*>(d_in), p->d_hist[p->cur], p->d_hist[p->cur ^ 1],
                       S * 2, H, N, (long)nBlk);
    PPL_LAUNCHED();
    return PPL_OK;
}

extern "C" int ppl_stokes(const float* d_spectra, unsigned nBlk, unsigned nSub, unsigned nCh,
                          unsigned nStokes, float* d_out, void* stream)
{
    PPL_CHECK_ARG(d_spectra && d_out, "null argument");
    PPL_CHECK_ARG(nStokes == 1 || nStokes == 4,
                  "You can either generate 1 or 4 Stokes parameters (got %u)", nStokes);
    const long nSpec = (long)nBlk * nSub;
    const long n = nSpec * nCh;
    if (!n) return PPL_OK;
    hipLaunchKernelGGL(stokes_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream_of(stream),
                       reinterpret_cast<const float2*>(d_spectra), d_out, nSpec, (int)nCh, (int)nStokes);
    PPL_LAUNCHED();
    return PPL_OK;
}

extern "C" int ppl_integrate(const float* d_in, unsigned nT, unsigned nSub, unsigned nPol, unsigned nCh,
                             unsigned window, unsigned binCh, float* d_out, void* stream)
{
    PPL_CHECK_ARG(d_in && d_out, "null argument");
    PPL_CHECK_ARG(window >= 1 && binCh >= 1, "window and channel bin must be >= 1");
    if (window > nT) window = nT;  // StokesIntegrator.cpp:53-56
    if (!window) return PPL_OK;
    const long newT = nT / window;
    const long n = newT * nSub * nPol * (nCh / binCh);
    if (!n) return PPL_OK;
    hipLaunchKernelGGL(integrate_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream_of(stream),
                       d_in, d_out, newT, (int)nSub, (int)nPol, (int)nCh, (int)window, (int)binCh);
    PPL_LAUNCHED();
    return PPL_OK;
}
