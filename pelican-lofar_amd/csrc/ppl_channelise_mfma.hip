// ppl_channelise_mfma.hip -- polyphase channeliser whose DFT runs on the
// matrix cores as split-precision GEMMs with FP32 accumulation.
//
// The FIR is the plain method's: FP32, oldest block first (src/lib/src/
// PPFChanneliser.cpp:276-289), each tap one fused multiply-add (the spectra
// are held to a tolerance against FP64, not to the reference's bits).  The N-point DFT of each
// filtered block, N = 16 R, is split Cooley-Tukey style with n = n1 + R n2 and
// k = k2 + 16 k1:
//
//   G[n1][k2] = sum_n2 F[n1 + R n2] W16^(n2 k2)          (stage 1, MFMA)
//   G'        = G[n1][k2] W_N^(n1 k2)                     (twiddle, FP32)
//   Y[k2 + 16 k1] = sum_n1 G'[n1][k2] W_R^(n1 k1)         (stage 2)
//
// Stage 1 is a real GEMM [rows = (block, n1)] x [K = 32: Re, Im of 16 n2] x
// [32: Re, Im of 16 k2] in v_mfma_f32_16x16x16 tiles.  Stage 2 is a 4-point
// DFT in registers for R = 4 (a lane holds the four n1 of one block), and a
// second MFMA GEMM for R = 16: the accumulator layout of a 16x16x16 MFMA
// (lane l: column l % 16, rows 4 (l / 16) .. +3) is the B-operand layout of
// the next one, so G' feeds stage 2 without leaving the registers.
//
// Precision (each product a*b of FP32 values):
//   BF16x3: a = ah + al, b = bh + bl in bfloat16 (ah = RN(a), al = RN(a - ah));
//           a*b ~ ah*bh + ah*bl + al*bh, three MFMAs, FP32 accumulation.
//   FP16x2: the data split in binary16 (in range unscaled, see kS1), the DFT
//           matrices rounded once: a*b ~ (ah + al)*RN(b), two MFMAs.
// Measured error against an all-FP64 chain, in units of the spectrum's RMS
// (noise input): BF16x3 at most 2e-5 (RMS 4e-6), FP16x2 at most 7e-4 (RMS
// 1.4e-4); the FP32 method: 6e-7.
// Stokes-I (StokesGenerator.cpp:93-99) and the time integration
// (StokesIntegrator.cpp:97-114) follow in FP32 in the reference's order.
#include "ppl_internal.h"
#include "ppl_ppf_batch.h"

#include <cmath>

struct ppl_ppf_plan;
const ppl_ppf_config& ppl_ppf_cfg(const ppl_ppf_plan* p);
const float* ppl_ppf_coeffs(const ppl_ppf_plan* p);
const float* ppl_ppf_twiddle(const ppl_ppf_plan* p);
uint32_t* ppl_ppf_hist_next(ppl_ppf_plan* p);
int ppl_ppf_save_history(ppl_ppf_plan* p, const int16_t* d_in, unsigned nBlk, hipStream_t st);

namespace {

using namespace ppl;

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 h16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
// blocks per tile: 16, and 8 at 256 channels so that four workgroups share a
// CU's LDS (the filtered tile and the per-block Stokes-I take 40 KB)
template <int N> constexpr int tile_blocks() { return N == 256 ? 8 : 16; }

struct C2 {
    float re, im;
};

// One operand fragment (four values of a 16x16x16 MFMA operand) in the
// split precision PREC: hi and lo parts (lo unused by the rounded-once side).
template <int PREC> struct Frag;
template <> struct Frag<PPL_DFT_BF16X3> {
    s16x4 hi, lo;
};
template <> struct Frag<PPL_DFT_FP16X2> {
    h16x4 hi, lo;
};

__device__ __forceinline__ short bf16_bits(float x)
{
    const __bf16 b = (__bf16)x;  // round to nearest even
    return __builtin_bit_cast(short, b);
}
__device__ __forceinline__ float bf16_val(short s) { return (float)__builtin_bit_cast(__bf16, s); }

// Split four FP32 values (scale: exact power of two applied first, FP16 only).
template <int PREC>
__device__ __forceinline__ Frag<PREC> split4(const float (&v)[4], float scale)
{
    Frag<PREC> f;
    if constexpr (PREC == PPL_DFT_BF16X3) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const short h = bf16_bits(v[i]);
            f.hi[i] = h;
            f.lo[i] = bf16_bits(v[i] - bf16_val(h));
        }
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float x = v[i] * scale;
            const _Float16 h = (_Float16)x;
            f.hi[i] = h;
            f.lo[i] = (_Float16)(x - (float)h);
        }
    }
    return f;
}

// acc += a * b in split precision; b is the constant (DFT matrix) side.
template <int PREC>
__device__ __forceinline__ f32x4 mma(const Frag<PREC>& a, const Frag<PREC>& b, f32x4 acc)
{
    if constexpr (PREC == PPL_DFT_BF16X3) {
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.lo, b.hi, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, b.lo, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, b.hi, acc, 0, 0, 0);
    } else {
        acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a.lo, b.hi, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.hi, acc, 0, 0, 0);
    }
    return acc;
}
// The same with the data on the B side (stage 2: A = DFT matrix).
template <int PREC>
__device__ __forceinline__ f32x4 mma_b(const Frag<PREC>& a, const Frag<PREC>& b, f32x4 acc)
{
    if constexpr (PREC == PPL_DFT_BF16X3) {
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.lo, b.hi, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, b.lo, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, b.hi, acc, 0, 0, 0);
    } else {
        acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.lo, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.hi, acc, 0, 0, 0);
    }
    return acc;
}

__device__ __forceinline__ C2 load_c(const uint32_t* __restrict__ p, long idx)
{
    const uint32_t v = p[idx];
    return {(float)(int16_t)(v & 0xffffu), (float)(int16_t)(v >> 16)};
}

// OUT: 0 = integrated Stokes-I [nBlk/integ][sub][1][N]; 1 = complex spectra
// [nBlk][sub][2][N].  Workgroup `group` (of the launch, or of one job of a
// batched launch) runs the blocks [run*runLen, +runLen) of one subband, both
// polarisations, in tiles of tile_blocks<N>() blocks; `stride` is the samples
// per (sub, pol) stream of `in`, of which the first nBlk*N are read.
template <int N, int TAPS, int PREC, int OUT>
__device__ __forceinline__ void ppf_mfma_group(const uint32_t* __restrict__ in, const uint32_t* __restrict__ hist,
                                               float* __restrict__ out, const float* __restrict__ coeffs,
                                               const float* __restrict__ tw, int nSub, long nBlk, long stride,
                                               int runLen, int nRuns, int integ, int group)
{
    constexpr int R = N / 16;
    static_assert(R == 4 || R == 16, "split-precision DFT: 64 or 256 channels");
    constexpr int kTB = tile_blocks<N>();
    constexpr int H = TAPS - 1;
    constexpr int ITEMS = (2 * N + kThreads - 1) / kThreads;  // (pol, channel) streams per thread
    constexpr int ROWT = kTB * R / 16;                        // 16-row tiles per polarisation
    // Scales applied before the FP16 split and removed after each stage (exact
    // powers of two).  |filtered sample| <= 32768 * sum_t |h| <= 8192 for the
    // coefficients of PolyphaseCoefficients (peak 1/nCh, at most 16 taps) and
    // |stage-1 output| <= 16 times that, so both stay inside binary16 unscaled,
    // and unscaled noise keeps the low parts out of the subnormal range.
    constexpr float kS1 = 1.0f;
    constexpr float kS2 = 1.0f;
    __shared__ float2 F[2][kTB][N];   // filtered blocks
    __shared__ float P[kTB][N];       // Stokes-I per block (OUT 0)

    const int sub = group / nRuns;
    const int run = group % nRuns;
    if (sub >= nSub) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long T = stride;
    const long b0 = (long)run * runLen;
    const int len = (int)min((long)runLen, nBlk - b0);

    // ---- FIR state: thread item q handles stream (pol, n) = divmod(tid + 256 q, N)
    C2 w[ITEMS][TAPS];
    float hc[ITEMS][TAPS];
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
        const int it = tid + kThreads * q;
        const bool on = it < 2 * N;
        const int pol = on ? it / N : 0, n = on ? it % N : 0;
        const uint32_t* x = in + (long)(2 * sub + pol) * T;
        const uint32_t* h = hist + (long)(2 * sub + pol) * H * N;
#pragma unroll
        for (int t = 0; t < TAPS; ++t) {
            hc[q][t] = on ? coeffs[t * N + n] : 0.0f;
            w[q][t] = {0.f, 0.f};
        }
#pragma unroll
        for (int t = 0; t < H; ++t) {
            const long b = b0 - H + t;
            C2 a = {0.f, 0.f};
            if (on) a = b >= 0 ? load_c(x, b * N + n) : load_c(h, (long)(H + b) * N + n);
            w[q][t + 1] = a;
        }
    }

    // ---- DFT constants of this lane
    const int col = lane & 15, grp = lane >> 4;
    auto W = [&](int m) -> C2 {  // W_N^m
        m &= N - 1;
        return {tw[2 * m], tw[2 * m + 1]};
    };
    // stage 1 B operands: rows n2 = 4 grp + v, column k2 = col of W16
    Frag<PREC> bR, bI, bNI;
    {
        float r[4], i[4], ni[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const C2 c = W(((4 * grp + v) * col) * R);
            r[v] = c.re;
            i[v] = c.im;
            ni[v] = -c.im;
        }
        bR = split4<PREC>(r, 1.0f);
        bI = split4<PREC>(i, 1.0f);
        bNI = split4<PREC>(ni, 1.0f);
    }
    // stage 2 A operands (R = 16): row k1 = col, columns n1 = 4 grp + v of W16
    Frag<PREC> aR, aI, aNI;
    if constexpr (R == 16) {
        float r[4], i[4], ni[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const C2 c = W((col * (4 * grp + v)) * R);
            r[v] = c.re;
            i[v] = c.im;
            ni[v] = -c.im;
        }
        aR = split4<PREC>(r, 1.0f);
        aI = split4<PREC>(i, 1.0f);
        aNI = split4<PREC>(ni, 1.0f);
    }

    for (int s0 = 0; s0 < len; s0 += kTB) {
        const int nb = min(kTB, len - s0);
        // ---- the tile's input, all loads in flight together
        uint32_t xin[ITEMS][kTB];
#pragma unroll
        for (int q = 0; q < ITEMS; ++q) {
            const int it = tid + kThreads * q;
            const bool on = it < 2 * N;
            const int pol = on ? it / N : 0, n = on ? it % N : 0;
            const uint32_t* x = in + (long)(2 * sub + pol) * T + (b0 + s0) * N + n;
#pragma unroll
            for (int s = 0; s < kTB; ++s) xin[q][s] = (on && s < nb) ? x[(long)s * N] : 0u;
        }
        // ---- FIR of the tile's blocks into LDS
#pragma unroll
        for (int s = 0; s < kTB; ++s) {
#pragma unroll
            for (int q = 0; q < ITEMS; ++q) {
                const int it = tid + kThreads * q;
                const bool on = it < 2 * N;
                const int pol = on ? it / N : 0, n = on ? it % N : 0;
#pragma unroll
                for (int t = 0; t < TAPS - 1; ++t) w[q][t] = w[q][t + 1];
                const uint32_t xv = xin[q][s];
                w[q][TAPS - 1] = {(float)(int16_t)(xv & 0xffffu), (float)(int16_t)(xv >> 16)};
                C2 f = {0.f, 0.f};
#pragma unroll
                for (int t = 0; t < TAPS; ++t) {
#pragma clang fp contract(fast)
                    const float cre = w[q][t].re * hc[q][t], cim = w[q][t].im * hc[q][t];
                    f.re = f.re + cre;
                    f.im = f.im + cim;
                }
                if (on) F[pol][s][n] = make_float2(f.re, f.im);
            }
        }
        __syncthreads();
        // ---- DFT: row tiles of polarisation 0, then of polarisation 1
        for (int pol = 0; pol < 2; ++pol) {
        for (int rt = wave; rt < ROWT; rt += kThreads / 64) {
            // stage 1 A operands: row (block, n1) = rt*16 + col, columns n2 = 4 grp + v
            const int gr = rt * 16 + col;
            const int sA = gr / R, n1A = gr % R;
            float fr[4], fi[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float2 x = F[pol][sA][n1A + R * (4 * grp + v)];
                fr[v] = x.x;
                fi[v] = x.y;
            }
            const Frag<PREC> aFr = split4<PREC>(fr, kS1), aFi = split4<PREC>(fi, kS1);
            f32x4 gR = {0.f, 0.f, 0.f, 0.f}, gI = {0.f, 0.f, 0.f, 0.f};
            gR = mma<PREC>(aFr, bR, gR);
            gR = mma<PREC>(aFi, bNI, gR);
            gI = mma<PREC>(aFr, bI, gI);
            gI = mma<PREC>(aFi, bR, gI);
            // G' = G W_N^(n1 k2): lane holds column k2 = col, rows rt*16 + 4 grp + v
            float yr[4], yi[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = rt * 16 + 4 * grp + v;
                const int n1 = row % R;
                const C2 t = W(n1 * col);
                const float a = gR[v] / kS1, c = gI[v] / kS1;
                yr[v] = a * t.re - c * t.im;
                yi[v] = a * t.im + c * t.re;
            }
            int sOut;  // block of this lane's outputs
            int kOut[4];
            if constexpr (R == 4) {
                // rows 4 grp + v = (block grp, n1 = v): 4-point DFT over v in registers
                sOut = rt * 4 + grp;
                const C2 a0 = {yr[0], yi[0]}, a1 = {yr[1], yi[1]}, a2 = {yr[2], yi[2]}, a3 = {yr[3], yi[3]};
                const C2 s02 = {a0.re + a2.re, a0.im + a2.im}, d02 = {a0.re - a2.re, a0.im - a2.im};
                const C2 s13 = {a1.re + a3.re, a1.im + a3.im}, d13 = {a1.re - a3.re, a1.im - a3.im};
                yr[0] = s02.re + s13.re;
                yi[0] = s02.im + s13.im;
                yr[2] = s02.re - s13.re;
                yi[2] = s02.im - s13.im;
                yr[1] = d02.re + d13.im;  // X1 = d02 - i d13
                yi[1] = d02.im - d13.re;
                yr[3] = d02.re - d13.im;  // X3 = d02 + i d13
                yi[3] = d02.im + d13.re;
#pragma unroll
                for (int v = 0; v < 4; ++v) kOut[v] = col + 16 * v;
            } else {
                // stage 2 on the matrix cores: Y = W16 G' (rows k1, columns k2)
                sOut = rt;
                const Frag<PREC> bGr = split4<PREC>(yr, kS2), bGi = split4<PREC>(yi, kS2);
                f32x4 oR = {0.f, 0.f, 0.f, 0.f}, oI = {0.f, 0.f, 0.f, 0.f};
                oR = mma_b<PREC>(aR, bGr, oR);
                oR = mma_b<PREC>(aNI, bGi, oR);
                oI = mma_b<PREC>(aI, bGr, oI);
                oI = mma_b<PREC>(aR, bGi, oI);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    yr[v] = oR[v] / kS2;
                    yi[v] = oI[v] / kS2;
                    kOut[v] = col + 16 * (4 * grp + v);
                }
            }
            const bool live = sOut < nb;
            if constexpr (OUT == 1) {
                if (live) {
                    const long b = b0 + s0 + sOut;
                    float2* o = reinterpret_cast<float2*>(out) + ((b * nSub + sub) * 2 + pol) * (long)N;
#pragma unroll
                    for (int v = 0; v < 4; ++v) o[kOut[v]] = make_float2(yr[v], yi[v]);
                }
            } else {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const float pw = yr[v] * yr[v] + yi[v] * yi[v];
                    if (pol == 0) P[sOut][kOut[v]] = pw;
                    else P[sOut][kOut[v]] = P[sOut][kOut[v]] + pw;  // I = powerX + powerY
                }
            }
        }
            __syncthreads();  // polarisation 0's powers are in P before polarisation 1 adds
        }
        if constexpr (OUT == 0) {
            __syncthreads();
            // integration over `integ` consecutive blocks, in time order
            const int nu = nb / integ;
            for (int e = tid; e < nu * N; e += kThreads) {
                const int u = e / N, k = e % N;
                float acc = 0.0f;
                for (int j = 0; j < integ; ++j) acc = acc + P[u * integ + j][k];
                const long ug = (b0 + s0) / integ + u;
                out[(ug * nSub + sub) * (long)N + k] = acc;
            }
        }
        __syncthreads();
    }
}

template <int N, int TAPS, int PREC, int OUT>
__global__ __launch_bounds__(kThreads) void ppf_mfma_kernel(const uint32_t* __restrict__ in,
                                                            const uint32_t* __restrict__ hist,
                                                            float* __restrict__ out,
                                                            const float* __restrict__ coeffs,
                                                            const float* __restrict__ tw, int nSub, long nBlk,
                                                            int runLen, int nRuns, int integ)
{
    ppf_mfma_group<N, TAPS, PREC, OUT>(in, hist, out, coeffs, tw, nSub, nBlk, nBlk * N, runLen, nRuns, integ,
                                       (int)blockIdx.x);
}

// Many plans' fused channeliser of one channel count, tap count and
// precision in one launch: each workgroup finds its job in the table.
template <int N, int TAPS, int PREC>
__global__ __launch_bounds__(kThreads) void ppf_mfma_batch_kernel(const PPFBatch B)
{
    int local;
    const PPFJob& J = find_batch_job(B, local);
    ppf_mfma_group<N, TAPS, PREC, 0>(J.in, J.hist, J.out, J.coeffs, B.tw, J.nSub, J.nBlk, J.stride, J.runLen,
                                     J.nRuns, J.integ, local);
}

template <int N, int TAPS, int PREC>
int launch(const ppl_ppf_config& c, const uint32_t* in, const uint32_t* hist, float* out, const float* coeffs,
           const float* tw, long nBlk, int mode, hipStream_t st)
{
    const int nSub = (int)c.nSub;
    const int runLen = ppf_run_length(&nBlk, &nSub, 1, 1, kFillGroups, tile_blocks<N>(), 4 * tile_blocks<N>());
    const int nRuns = (int)((nBlk + runLen - 1) / runLen);
    const long grid = (long)c.nSub * nRuns;
    if (grid >= (1L << 31)) {
        set_error("channeliser: %ld workgroups exceed the grid", grid);
        return PPL_E_ARG;
    }
    if (mode == 0)
        hipLaunchKernelGGL((ppf_mfma_kernel<N, TAPS, PREC, 0>), dim3((unsigned)grid), dim3(kThreads), 0, st, in,
                           hist, out, coeffs, tw, (int)c.nSub, nBlk, runLen, nRuns, (int)c.integrate);
    else
        hipLaunchKernelGGL((ppf_mfma_kernel<N, TAPS, PREC, 1>), dim3((unsigned)grid), dim3(kThreads), 0, st, in,
                           hist, out, coeffs, tw, (int)c.nSub, nBlk, runLen, nRuns, 1);
    PPL_LAUNCHED();
    return PPL_OK;
}

template <int N, int PREC>
int launch_taps(const ppl_ppf_config& c, const uint32_t* in, const uint32_t* hist, float* out, const float* coeffs,
                const float* tw, long nBlk, int mode, hipStream_t st)
{
    switch (c.nTaps) {
        case 4: return launch<N, 4, PREC>(c, in, hist, out, coeffs, tw, nBlk, mode, st);
        case 8: return launch<N, 8, PREC>(c, in, hist, out, coeffs, tw, nBlk, mode, st);
        case 16: return launch<N, 16, PREC>(c, in, hist, out, coeffs, tw, nBlk, mode, st);
    }
    set_error("channeliser: %u taps", c.nTaps);
    return PPL_E_ARG;
}

// One batched launch of a group of jobs sharing N, TAPS and PREC: sets each
// job's runs and first workgroup.
template <int N, int TAPS, int PREC>
int launch_batch(PPFBatch& B, hipStream_t st)
{
    long nb[kMaxPPFJobs];
    int ns[kMaxPPFJobs];
    for (int k = 0; k < B.n; ++k) {
        nb[k] = B.job[k].nBlk;
        ns[k] = B.job[k].nSub;
    }
    const int L = ppf_run_length(nb, ns, B.n, 1, kFillGroups, tile_blocks<N>(), 4 * tile_blocks<N>());
    long blocks = 0;
    for (int k = 0; k < B.n; ++k) {
        PPFJob& J = B.job[k];
        J.runLen = L;
        J.nRuns = (J.nBlk + L - 1) / L;
        J.blockStart = (int)blocks;
        blocks += (long)J.nSub * J.nRuns;
        if (blocks >= (1L << 31)) {
            set_error("channeliser batch: %ld workgroups exceed the grid", blocks);
            return PPL_E_ARG;
        }
    }
    hipLaunchKernelGGL((ppf_mfma_batch_kernel<N, TAPS, PREC>), dim3((unsigned)blocks), dim3(kThreads), 0, st, B);
    PPL_LAUNCHED();
    return PPL_OK;
}

template <int N, int PREC>
int launch_batch_taps(PPFBatch& B, int taps, hipStream_t st)
{
    switch (taps) {
        case 4: return launch_batch<N, 4, PREC>(B, st);
        case 8: return launch_batch<N, 8, PREC>(B, st);
        case 16: return launch_batch<N, 16, PREC>(B, st);
    }
    set_error("channeliser: %d taps", taps);
    return PPL_E_ARG;
}

}  // namespace

// The split-precision part of ppl_ppf_stokes_i16_batch: jobs of one channel
// count (64 or 256), tap count and precision.
int ppl_mfma_run_batch(ppl::PPFBatch& B, int nCh, int taps, int prec, hipStream_t st)
{
    if (nCh == 64)
        return prec == PPL_DFT_BF16X3 ? launch_batch_taps<64, PPL_DFT_BF16X3>(B, taps, st)
                                      : launch_batch_taps<64, PPL_DFT_FP16X2>(B, taps, st);
    if (nCh == 256)
        return prec == PPL_DFT_BF16X3 ? launch_batch_taps<256, PPL_DFT_BF16X3>(B, taps, st)
                                      : launch_batch_taps<256, PPL_DFT_FP16X2>(B, taps, st);
    set_error("split-precision DFT channeliser: %d channels", nCh);
    return PPL_E_CONFIG;
}

int ppl_mfma_create(ppl_ppf_plan* p)
{
    const ppl_ppf_config& c = ppl_ppf_cfg(p);
    if (c.nCh != 64 && c.nCh != 256) {
        set_error("split-precision DFT channeliser: built for 64 and 256 channels per subband (got %u)", c.nCh);
        return PPL_E_CONFIG;
    }
    const unsigned tb = c.nCh == 256 ? tile_blocks<256>() : tile_blocks<64>();
    if (c.integrate == 0 || tb % c.integrate != 0) {
        set_error("split-precision DFT channeliser: integrateTimeBins must divide %u (got %u)", tb, c.integrate);
        return PPL_E_CONFIG;
    }
    return PPL_OK;
}

void ppl_mfma_destroy(ppl_ppf_plan*) {}

int ppl_mfma_run(ppl_ppf_plan* p, const int16_t* d_in, const uint32_t* histCur, unsigned nBlk, float* d_out,
                 int mode, hipStream_t st)
{
    const ppl_ppf_config& c = ppl_ppf_cfg(p);
    if (mode == 0 && nBlk % c.integrate != 0) {
        set_error("channeliser: %u blocks is not a multiple of integrateTimeBins %u", nBlk, c.integrate);
        return PPL_E_ARG;
    }
    if (nBlk == 0) return PPL_OK;
    const uint32_t* in = reinterpret_cast<const uint32_t*>(d_in);
    const float* coeffs = ppl_ppf_coeffs(p);
    const float* tw = ppl_ppf_twiddle(p);
    int rc = PPL_E_CONFIG;
    if (c.nCh == 64)
        rc = c.precision == PPL_DFT_BF16X3
                 ? launch_taps<64, PPL_DFT_BF16X3>(c, in, histCur, d_out, coeffs, tw, nBlk, mode, st)
                 : launch_taps<64, PPL_DFT_FP16X2>(c, in, histCur, d_out, coeffs, tw, nBlk, mode, st);
    else if (c.nCh == 256)
        rc = c.precision == PPL_DFT_BF16X3
                 ? launch_taps<256, PPL_DFT_BF16X3>(c, in, histCur, d_out, coeffs, tw, nBlk, mode, st)
                 : launch_taps<256, PPL_DFT_FP16X2>(c, in, histCur, d_out, coeffs, tw, nBlk, mode, st);
    if (rc != PPL_OK) return rc;
    return ppl_ppf_save_history(p, d_in, nBlk, st);
}
