// ppl_dedisperse.hip -- brute-force incoherent dedispersion for gfx950.
//
// out[dm][t] = sum_{c=0}^{nChan-1} in[c][t + delay[dm][c]],  t < nOut,
// summed in channel order (the FP32 result is bit-identical to the
// reference's per-thread channel loop, src/lib/DedispersionKernel.cu:32-43).
//
// Kernel shape.  One 256-thread workgroup owns a tile of TD = 4*ND DM trials
// by TT = 64*NT output samples.  Wave w of the workgroup accumulates ND DM
// trials; lane l accumulates samples t0 + l + 64*i (i < NT), so every LDS
// read of a wave is 64 consecutive words (bank-conflict free) and every
// store is a coalesced 256-byte row segment.  Channels are streamed in
// blocks of CB: for each channel the workgroup stages the window
// [t0 + delay[dm0][c], t0 + delay[dm0+TD-1][c] + TT) of the input row into
// LDS with coalesced loads, then each wave reads its ND delayed 64-wide
// slices from LDS.  Delays are wave-uniform and come from a transposed
// table delayT[c][dm] through the scalar cache.
#include "ppl_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

struct ppl_dedisp_plan {
    ppl_dedisp_config cfg;
    int maxshift = 0, rem = 0, nOut = 0;
    float mstartdm = 0, mdmstep = 0;
    std::vector<float> shift;
    std::vector<int32_t> delays;  // host [tdms][nChan]
    int32_t* d_delayT = nullptr;  // device [nChan][tdmsPad]
    int tdmsPad = 0;
    int nd = 8, nt = 8;           // per-thread DM x sample register tile
    int span = 0;                 // max delay spread inside a DM tile
};

namespace {

using namespace ppl;

template <typename T> struct Acc;
template <> struct Acc<float> { using type = float; };
template <> struct Acc<_Float16> { using type = float; };
template <> struct Acc<uint8_t> { using type = int32_t; };

template <typename Tin, int ND, int NT>
__global__ __launch_bounds__(256) void dedisperse_kernel(
    const Tin* __restrict__ in, typename Acc<Tin>::type* __restrict__ out,
    const int32_t* __restrict__ delayT, int nChan, int nsamp, int tdms, int tdmsPad, int nOut,
    int W, int CB, int nDMTiles)
{
    using TA = typename Acc<Tin>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Tin* lds = reinterpret_cast<Tin*>(smem_raw);
    constexpr int TT = 64 * NT;
    constexpr int TD = 4 * ND;

    const int bid = blockIdx.x;
    const int dmTile = bid % nDMTiles;
    const int tTile = bid / nDMTiles;
    const int t0 = tTile * TT;
    const int dm0 = dmTile * TD;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int dmw = dm0 + wave * ND;

    TA acc[ND][NT];
#pragma unroll
    for (int j = 0; j < ND; ++j)
#pragma unroll
        for (int i = 0; i < NT; ++i) acc[j][i] = TA(0);

    for (int cb0 = 0; cb0 < nChan; cb0 += CB) {
        const int nc = min(CB, nChan - cb0);
        __syncthreads();
        for (int cc = 0; cc < nc; ++cc) {
            const int c = cb0 + cc;
            const int dmin = delayT[(size_t)c * tdmsPad + dm0];
            const long base = (long)t0 + dmin;
            const Tin* src = in + (size_t)c * nsamp + base;
            const long lim = (long)nsamp - base;
            Tin* dst = lds + cc * W;
            for (int k = tid; k < W; k += 256) dst[k] = (k < lim) ? src[k] : Tin(0);
        }
        __syncthreads();
        for (int cc = 0; cc < nc; ++cc) {
            const int c = cb0 + cc;
            const int32_t* dl = delayT + (size_t)c * tdmsPad;
            const int dmin = dl[dm0];
            const Tin* row = lds + cc * W + lane;
#pragma unroll
            for (int j = 0; j < ND; ++j) {
                const Tin* p = row + (dl[dmw + j] - dmin);
#pragma unroll
                for (int i = 0; i < NT; ++i) acc[j][i] = acc[j][i] + (TA)p[64 * i];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < ND; ++j) {
        const int dm = dmw + j;
        if (dm < tdms) {
#pragma unroll
            for (int i = 0; i < NT; ++i) {
                const int t = t0 + lane + 64 * i;
                if (t < nOut) out[(size_t)dm * nOut + t] = acc[j][i];
            }
        }
    }
}

// delayT[c][dm] = trunc(shift[c] * (mstartdm + dm*mdmstep)) for dm < tdms,
// rows padded with the last trial (compiled without contraction, so the
// product and sum round separately, as in the host table).
__global__ void build_delays_kernel(const float* __restrict__ shift, int nChan, int tdms,
                                    int tdmsPad, float mstartdm, float mdmstep,
                                    int32_t* __restrict__ delayT)
{
    const int dm = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (dm >= tdmsPad || c >= nChan) return;
    const int d = dm < tdms ? dm : tdms - 1;
    const float prod = (float)d * mdmstep;
    const float dmval = mstartdm + prod;
    const float p = shift[c] * dmval;
    delayT[(size_t)c * tdmsPad + dm] = (int32_t)p;
}

struct Shape {
    int nd, nt;
};

constexpr int kMaxLDS = 64 * 1024;

template <typename Tin>
int launch_shape(const Shape& s, const Tin* in, void* out, const int32_t* delayT, int nChan,
                 int nsamp, int tdms, int tdmsPad, int nOut, int span, hipStream_t stream)
{
    using TA = typename Acc<Tin>::type;
    const int TT = 64 * s.nt, TD = 4 * s.nd;
    int W = TT + span;
    W = (W + 15) & ~15;
    const int rowBytes = W * (int)sizeof(Tin);
    if (rowBytes > kMaxLDS) {
        set_error("dedisperse: delay spread %d too large for tile %dx%d", span, TD, TT);
        return PPL_E_CONFIG;
    }
    int CB = std::min(16, kMaxLDS / rowBytes);
    CB = std::max(CB, 1);
    const int nDMTiles = (tdms + TD - 1) / TD;
    const int nTTiles = (nOut + TT - 1) / TT;
    dim3 grid((unsigned)nDMTiles * nTTiles), block(256);
    const size_t lds = (size_t)CB * rowBytes;
    TA* o = static_cast<TA*>(out);
#define PPL_DD_CASE(ND_, NT_)                                                                   \
    if (s.nd == ND_ && s.nt == NT_) {                                                           \
        hipLaunchKernelGGL((dedisperse_kernel<Tin, ND_, NT_>), grid, block, lds, stream, in, o, \
                           delayT, nChan, nsamp, tdms, tdmsPad, nOut, W, CB, nDMTiles);         \
        PPL_LAUNCHED();                                                                         \
        return PPL_OK;                                                                          \
    }
    PPL_DD_CASE(8, 8)
    PPL_DD_CASE(8, 4)
    PPL_DD_CASE(4, 8)
    PPL_DD_CASE(4, 4)
    PPL_DD_CASE(2, 8)
    PPL_DD_CASE(16, 4)
    PPL_DD_CASE(16, 2)
#undef PPL_DD_CASE
    set_error("dedisperse: unsupported tile %dx%d", s.nd, s.nt);
    return PPL_E_ARG;
}

const Shape kShapes[] = {{8, 8}, {8, 4}, {4, 8}, {4, 4}, {2, 8}, {16, 4}, {16, 2}};

bool shape_known(int nd, int nt)
{
    for (const Shape& s : kShapes)
        if (s.nd == nd && s.nt == nt) return true;
    return false;
}

// Largest delay spread inside any DM tile of TD trials.
int tile_span(const std::vector<int32_t>& d, int tdms, int nChan, int TD)
{
    int span = 0;
    for (int dm0 = 0; dm0 < tdms; dm0 += TD) {
        const int dm1 = std::min(dm0 + TD - 1, tdms - 1);
        const int32_t* a = &d[(size_t)dm0 * nChan];
        const int32_t* b = &d[(size_t)dm1 * nChan];
        for (int c = 0; c < nChan; ++c) span = std::max(span, b[c] - a[c]);
    }
    return span;
}

int upload_delays(ppl_dedisp_plan* p)
{
    const int TD = 4 * p->nd;
    const int tdms = (int)p->cfg.tdms, nChan = (int)p->cfg.nChan;
    const int pad = ((tdms + TD - 1) / TD) * TD;
    std::vector<int32_t> t((size_t)nChan * pad);
    for (int c = 0; c < nChan; ++c)
        for (int dm = 0; dm < pad; ++dm)
            t[(size_t)c * pad + dm] = p->delays[(size_t)std::min(dm, tdms - 1) * nChan + c];
    if (p->d_delayT) PPL_HIP(hipFree(p->d_delayT));
    p->d_delayT = nullptr;
    PPL_HIP(hipMalloc(&p->d_delayT, t.size() * sizeof(int32_t)));
    PPL_HIP(hipMemcpy(p->d_delayT, t.data(), t.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    p->tdmsPad = pad;
    p->span = tile_span(p->delays, tdms, nChan, TD);
    return PPL_OK;
}

}  // namespace

extern "C" int ppl_dedisp_plan_create(ppl_dedisp_plan** plan, const ppl_dedisp_config* cfg)
{
    PPL_CHECK_ARG(plan && cfg, "null argument");
    PPL_CHECK_ARG(cfg->nChan > 0 && cfg->nsampBuf > 0 && cfg->tdms > 0, "empty dedispersion shape");
    PPL_CHECK_ARG(cfg->dtype == PPL_F32 || cfg->dtype == PPL_F16 || cfg->dtype == PPL_U8,
                  "unknown dtype %d", cfg->dtype);
    PPL_CHECK_ARG(cfg->fch1 != 0.0, "frequencyChannel1 must be a positive number");
    PPL_CHECK_ARG(cfg->foff < 0.0,
                  "channelBandwidth must be negative (descending channel order)");
    ppl_dedisp_plan* p = new ppl_dedisp_plan;
    p->cfg = *cfg;
    if (p->cfg.dmLow < 0.0f) p->cfg.dmLow = 0.0f;  // DedispersionModule.cpp:57
    const unsigned N = cfg->nChan;
    // DedispersionModule::resize :152-157 (invert = 0)
    p->shift.resize(N);
    for (int c = 0; c < (int)N; ++c) {
        const double f = cfg->fch1 + (cfg->foff * c);
        float val = 4148.741601 * ((1.0 / f / f) - (1.0 / cfg->fch1 / cfg->fch1));
        p->shift[c] = val;
    }
    // :159 (float products, then the double division, truncated to int)
    const float top = p->cfg.dmLow + p->cfg.dmStep * (float)(cfg->tdms - 1);
    p->maxshift = (int)((double)(top * p->shift[N - 1]) / cfg->tsamp);
    if ((int)cfg->nsampBuf <= p->maxshift) {
        set_error("DedispersionModule: maxshift requirements (%d) are bigger than the number "
                  "of samples (%u)", p->maxshift, cfg->nsampBuf);
        delete p;
        return PPL_E_CONFIG;
    }
    p->rem = (int)((cfg->nsampBuf - (unsigned)p->maxshift) % (14u * 16u));  // :162
    p->nOut = (int)cfg->nsampBuf - p->maxshift - p->rem;                      // :274
    // DedispersionKernel members are float (DedispersionModule.h:50-52), :363-364
    const float tsamp_f = (float)cfg->tsamp;
    p->mstartdm = p->cfg.dmLow / tsamp_f;
    p->mdmstep = p->cfg.dmStep / tsamp_f;
    p->delays.resize((size_t)cfg->tdms * N);
    int maxdelay = 0;
    for (unsigned dm = 0; dm < cfg->tdms; ++dm) {
        const float prod = (float)dm * p->mdmstep;
        const float dmval = p->mstartdm + prod;
        for (unsigned c = 0; c < N; ++c) {
            const float v = p->shift[c] * dmval;
            const int32_t d = (int32_t)v;
            p->delays[(size_t)dm * N + c] = d;
            maxdelay = std::max(maxdelay, d);
        }
    }
    if (p->nOut <= 0 || p->nOut - 1 + maxdelay >= (int)cfg->nsampBuf) {
        set_error("dedispersion: delayed reads (max delay %d, nOut %d) leave the %u-sample buffer",
                  maxdelay, p->nOut, cfg->nsampBuf);
        delete p;
        return PPL_E_CONFIG;
    }
    // default tile: 32 DMs x 512 samples (f32/f16), narrower DM tiles when the
    // delay spread makes the LDS window too wide
    const int elem = cfg->dtype == PPL_F32 ? 4 : cfg->dtype == PPL_F16 ? 2 : 1;
    const Shape prefs[] = {{8, 8}, {4, 8}, {2, 8}};
    for (const Shape& s : prefs) {
        p->nd = s.nd;
        p->nt = s.nt;
        const int TD = 4 * s.nd;
        const int span = tile_span(p->delays, (int)cfg->tdms, (int)N, TD);
        if ((64 * s.nt + span + 16) * elem * 4 <= kMaxLDS) break;
    }
    int rc = upload_delays(p);
    if (rc != PPL_OK) {
        delete p;
        return rc;
    }
    *plan = p;
    return PPL_OK;
}

extern "C" int ppl_dedisp_plan_destroy(ppl_dedisp_plan* plan)
{
    if (!plan) return PPL_OK;
    if (plan->d_delayT) (void)hipFree(plan->d_delayT);
    delete plan;
    return PPL_OK;
}

extern "C" int ppl_dedisp_plan_info(const ppl_dedisp_plan* p, ppl_dedisp_info* info)
{
    PPL_CHECK_ARG(p && info, "null argument");
    info->maxshift = p->maxshift;
    info->remainingSamples = p->rem;
    info->nOut = p->nOut;
    info->mstartdm = p->mstartdm;
    info->mdmstep = p->mdmstep;
    info->tileDM = 4 * p->nd;
    info->tileT = 64 * p->nt;
    return PPL_OK;
}

extern "C" int ppl_dedisp_set_tile(ppl_dedisp_plan* p, int tileDM, int tileT)
{
    PPL_CHECK_ARG(p, "null plan");
    if (tileDM == 0 && tileT == 0) return PPL_OK;
    PPL_CHECK_ARG(tileDM % 4 == 0 && tileT % 64 == 0, "tile %dx%d not a multiple of 4x64", tileDM,
                  tileT);
    const int nd = tileDM / 4, nt = tileT / 64;
    PPL_CHECK_ARG(shape_known(nd, nt), "tile %dx%d not instantiated", tileDM, tileT);
    p->nd = nd;
    p->nt = nt;
    return upload_delays(p);
}

extern "C" int ppl_dedisp_delays(const ppl_dedisp_plan* p, int32_t* host_out)
{
    PPL_CHECK_ARG(p && host_out, "null argument");
    std::copy(p->delays.begin(), p->delays.end(), host_out);
    return PPL_OK;
}

extern "C" int ppl_dedisperse(const ppl_dedisp_plan* p, const void* d_in, void* d_out, void* stream)
{
    PPL_CHECK_ARG(p && d_in && d_out, "null argument");
    const Shape s{p->nd, p->nt};
    const int nChan = (int)p->cfg.nChan, nsamp = (int)p->cfg.nsampBuf, tdms = (int)p->cfg.tdms;
    hipStream_t st = stream_of(stream);
    switch (p->cfg.dtype) {
        case PPL_F32:
            return launch_shape<float>(s, static_cast<const float*>(d_in), d_out, p->d_delayT, nChan,
                                       nsamp, tdms, p->tdmsPad, p->nOut, p->span, st);
        case PPL_F16:
            return launch_shape<_Float16>(s, static_cast<const _Float16*>(d_in), d_out, p->d_delayT,
                                          nChan, nsamp, tdms, p->tdmsPad, p->nOut, p->span, st);
        case PPL_U8:
            return launch_shape<uint8_t>(s, static_cast<const uint8_t*>(d_in), d_out, p->d_delayT,
                                         nChan, nsamp, tdms, p->tdmsPad, p->nOut, p->span, st);
    }
    set_error("unknown dtype");
    return PPL_E_ARG;
}

extern "C" int ppl_cache_dedisperse_loop(float* outbuff, long outbufSize, float* buff, float mstartdm,
                                         float mdmstep, int tdms, int numSamples,
                                         const float* dmShift, int maxshift, int i_nchans,
                                         void* stream)
{
    PPL_CHECK_ARG(outbuff && buff && dmShift, "null argument");
    PPL_CHECK_ARG(tdms > 0 && i_nchans > 0 && numSamples > maxshift && maxshift >= 0,
                  "bad shape (tdms %d, nchans %d, numSamples %d, maxshift %d)", tdms, i_nchans,
                  numSamples, maxshift);
    const int nOut = numSamples - maxshift;
    PPL_CHECK_ARG((long)tdms * nOut * (long)sizeof(float) <= outbufSize,
                  "output buffer (%ld bytes) smaller than tdms*nOut floats", outbufSize);
    hipStream_t st = stream_of(stream);
    const Shape s{8, 8};
    const int TD = 4 * s.nd;
    const int pad = ((tdms + TD - 1) / TD) * TD;
    // delay table on the device, then the spread of the first/last trial
    // per tile is bounded by the spread of the whole DM range
    int32_t* d_delayT = nullptr;
    PPL_HIP(hipMallocAsync((void**)&d_delayT, (size_t)i_nchans * pad * sizeof(int32_t), st));
    dim3 g((unsigned)cdiv(pad, 256), (unsigned)i_nchans);
    hipLaunchKernelGGL(build_delays_kernel, g, dim3(256), 0, st, dmShift, i_nchans, tdms, pad,
                       mstartdm, mdmstep, d_delayT);
    PPL_LAUNCHED();
    std::vector<float> hshift(i_nchans);
    PPL_HIP(hipMemcpyAsync(hshift.data(), dmShift, i_nchans * sizeof(float), hipMemcpyDeviceToHost, st));
    PPL_HIP(hipStreamSynchronize(st));
    int span = 0;
    for (int c = 0; c < i_nchans; ++c) {
        for (int dm0 = 0; dm0 < tdms; dm0 += TD) {
            const int dm1 = std::min(dm0 + TD - 1, tdms - 1);
            const int a = (int)(hshift[c] * (mstartdm + (float)dm0 * mdmstep));
            const int b = (int)(hshift[c] * (mstartdm + (float)dm1 * mdmstep));
            span = std::max(span, b - a);
            if (a < 0 || (long)(nOut - 1) + b >= numSamples) {
                (void)hipFreeAsync(d_delayT, st);
                set_error("cacheDedisperseLoop: delayed read outside the buffer (channel %d)", c);
                return PPL_E_CONFIG;
            }
        }
    }
    const long used = (long)tdms * nOut * (long)sizeof(float);
    if (outbufSize > used)
        PPL_HIP(hipMemsetAsync(reinterpret_cast<char*>(outbuff) + used, 0, outbufSize - used, st));
    int rc = launch_shape<float>(s, buff, outbuff, d_delayT, i_nchans, numSamples, tdms, pad, nOut,
                                 span, st);
    (void)hipFreeAsync(d_delayT, st);
    return rc;
}

// ---------------------------------------------------------------- converters
namespace {

__global__ void to_f16_kernel(const float* __restrict__ in, _Float16* __restrict__ out, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (_Float16)in[i];  // v_cvt_f16_f32: round to nearest even
}

__global__ void quantise_kernel(const float* __restrict__ in, uint8_t* __restrict__ out, size_t n,
                                float inv)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float y = rintf(in[i] * inv);
        if (!(y >= -128.0f)) y = -128.0f;
        if (y > 127.0f) y = 127.0f;
        out[i] = (uint8_t)((int)y + 128);
    }
}

}  // namespace

extern "C" int ppl_convert_f16(const float* d_in, uint16_t* d_out, size_t n, void* stream)
{
    PPL_CHECK_ARG(d_in && d_out, "null argument");
    if (!n) return PPL_OK;
    hipLaunchKernelGGL(to_f16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       stream_of(stream), d_in, reinterpret_cast<_Float16*>(d_out), n);
    PPL_LAUNCHED();
    return PPL_OK;
}

extern "C" int ppl_quantise_u8(const float* d_in, uint8_t* d_out, size_t n, float inv_scale,
                               void* stream)
{
    PPL_CHECK_ARG(d_in && d_out, "null argument");
    if (!n) return PPL_OK;
    hipLaunchKernelGGL(quantise_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       stream_of(stream), d_in, d_out, n, inv_scale);
    PPL_LAUNCHED();
    return PPL_OK;
}
