// oracle/dedisp.cpp -- CPU restatement of the dedispersion plan arithmetic,
// the buffer assembly, brute-force dedispersion and event detection.
// TEST INFRASTRUCTURE ONLY (see oracle.h).
#include "oracle.h"

#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

// DedispersionModule::resize, src/lib/src/DedispersionModule.cpp:152-162,
// DedispersionKernel members (float) src/lib/DedispersionModule.h:50-52 and
// the kernel arguments :362-368.  Buffer order, invert = 0 (foff < 0).
extern "C" int orc_dedisp_plan(unsigned nChan, unsigned nsampBuf, unsigned tdms, float dmLow,
                               float dmStep, double tsamp, double fch1, double foff, float* shift,
                               float* ms, int* ints)
{
    for (int c = 0; c < (int)nChan; ++c) {
        float val = 4148.741601 * ((1.0 / (fch1 + (foff * c)) / (fch1 + (foff * c))) - (1.0 / fch1 / fch1));
        shift[c] = val;
    }
    // _maxshift = ((_dmLow + _dmStep * (_tdms - 1)) * _dmshifts[N-1]) / _tsamp  (int)
    float top = dmLow + dmStep * (float)(tdms - 1);
    int maxshift = (int)((double)(top * shift[nChan - 1]) / tsamp);
    if ((int)nsampBuf <= maxshift) return -1;
    int rem = (int)((nsampBuf - (unsigned)maxshift) % (14u * 16u));
    int nOut = (int)nsampBuf - maxshift - rem;
    float tsamp_f = (float)tsamp;
    ms[0] = dmLow / tsamp_f;
    ms[1] = dmStep / tsamp_f;
    ints[0] = maxshift;
    ints[1] = rem;
    ints[2] = nOut;
    // every read t + delay, t < nOut, must stay inside the channel row
    float dmval = ms[0] + (float)(tdms - 1) * ms[1];
    int maxdelay = (int)(shift[nChan - 1] * dmval);
    if (nOut <= 0 || nOut - 1 + maxdelay >= (int)nsampBuf) return -2;
    return 0;
}

// src/lib/DedispersionKernel.cu:29,36 -- trunc(shift[c] * (mstartdm + dm*mdmstep)),
// the sum left unfused (identical to the fused form when mstartdm == 0).
extern "C" void orc_dedisp_delays(const float* shift, unsigned nChan, unsigned tdms, float mstartdm,
                                  float mdmstep, int32_t* delay)
{
    for (unsigned dm = 0; dm < tdms; ++dm) {
        float prod = (float)dm * mdmstep;
        float dmval = mstartdm + prod;
        for (unsigned c = 0; c < nChan; ++c) {
            float p = shift[c] * dmval;
            delay[(size_t)dm * nChan + c] = (int32_t)p;  // truncation toward zero
        }
    }
}

// The same table with the trial value formed as one fused multiply-add,
// fmaf(dm, mdmstep, mstartdm): what DedispersionKernel.cu:29 computes when
// nvcc contracts `mstartdm + (idx * mdmstep)` (its default, -fmad=true).
extern "C" void orc_dedisp_delays_fused(const float* shift, unsigned nChan, unsigned tdms, float mstartdm,
                                        float mdmstep, int32_t* delay)
{
    for (unsigned dm = 0; dm < tdms; ++dm) {
        float dmval = std::fma((float)dm, mdmstep, mstartdm);
        for (unsigned c = 0; c < nChan; ++c) {
            float p = shift[c] * dmval;
            delay[(size_t)dm * nChan + c] = (int32_t)p;
        }
    }
}

template <typename Tin, typename Tacc, typename Conv>
static void brute(const Tin* in, unsigned nChan, unsigned nsampBuf, unsigned tdms, const int32_t* delay,
                  unsigned nOut, Tacc* out, int nthreads, Conv conv)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
#endif
    for (int dm = 0; dm < (int)tdms; ++dm) {
        std::vector<Tacc> acc(nOut, Tacc(0));
        const int32_t* d = delay + (size_t)dm * nChan;
        for (unsigned c = 0; c < nChan; ++c) {
            const Tin* row = in + (size_t)c * nsampBuf + d[c];
            for (unsigned t = 0; t < nOut; ++t) acc[t] = acc[t] + conv(row[t]);
        }
        std::memcpy(out + (size_t)dm * nOut, acc.data(), nOut * sizeof(Tacc));
    }
}

// out[dm][t] = sum_c in[c][t + delay[dm][c]] with FP32 adds in channel order
// (DedispersionKernel.cu:32-43).
extern "C" void orc_dedisperse_f32(const float* in, unsigned nChan, unsigned nsampBuf, unsigned tdms,
                                   const int32_t* delay, unsigned nOut, float* out, int nthreads)
{
    brute<float, float>(in, nChan, nsampBuf, tdms, delay, nOut, out, nthreads, [](float v) { return v; });
}

static float f16_to_f32(uint16_t h)
{
    uint32_t sign = (uint32_t)(h & 0x8000) << 16;
    uint32_t e = (h >> 10) & 0x1f, m = h & 0x3ff;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) bits = sign;
        else {
            float v = std::ldexp((float)m, -24);
            std::memcpy(&bits, &v, 4);
            bits |= sign;
        }
    } else if (e == 31) {
        bits = sign | 0x7f800000u | (m << 13);
    } else {
        bits = sign | ((e + 112) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

extern "C" void orc_dedisperse_f16(const uint16_t* in, unsigned nChan, unsigned nsampBuf, unsigned tdms,
                                   const int32_t* delay, unsigned nOut, float* out, int nthreads)
{
    brute<uint16_t, float>(in, nChan, nsampBuf, tdms, delay, nOut, out, nthreads, f16_to_f32);
}

extern "C" void orc_dedisperse_u8(const uint8_t* in, unsigned nChan, unsigned nsampBuf, unsigned tdms,
                                  const int32_t* delay, unsigned nOut, int32_t* out, int nthreads)
{
    brute<uint8_t, int32_t>(in, nChan, nsampBuf, tdms, delay, nOut, out, nthreads,
                            [](uint8_t v) { return (int32_t)v; });
}

// IEEE binary32 -> binary16, round to nearest, ties to even.
static uint16_t f32_to_f16_one(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u | ((ax >> 13) & 0x3ffu) : 0u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);
    if (ax < 0x38800000u) {
        float a;
        std::memcpy(&a, &ax, 4);
        float k = std::nearbyint(a * 16777216.0f);  // exact scaling by 2^24
        return (uint16_t)(sign | (uint32_t)k);
    }
    uint32_t e = (ax >> 23) - 127 + 15;
    uint32_t mant = ax & 0x7fffffu;
    uint32_t h = (e << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

extern "C" void orc_f32_to_f16(const float* in, uint16_t* out, size_t n)
{
    for (size_t i = 0; i < n; ++i) out[i] = f32_to_f16_one(in[i]);
}

extern "C" void orc_quantise_u8(const float* in, uint8_t* out, size_t n, float inv_scale)
{
    for (size_t i = 0; i < n; ++i) {
        float y = std::nearbyint(in[i] * inv_scale);
        if (!(y >= -128.0f)) y = -128.0f;
        if (y > 127.0f) y = 127.0f;
        out[i] = (uint8_t)((int)y + 128);
    }
}

// The u8 buffer's quantiser with exceptional rows (the product's u8 mode;
// the reference has no quantised mode): buffer x [nChan][nsamp]; a column
// (spectrum) t more than half of whose samples quantise outside [-128, 127]
// (NaN included) is quantised relative to off[t] = x[nChan/2][t]; every
// other column has off[t] = 0.
extern "C" void orc_quantise_u8_rows(const float* x, unsigned nChan, unsigned nsamp, float inv_scale, uint8_t* q,
                                     float* off)
{
    for (unsigned t = 0; t < nsamp; ++t) {
        unsigned n = 0;
        for (unsigned c = 0; c < nChan; ++c) {
            const float y = std::nearbyint(x[(size_t)c * nsamp + t] * inv_scale);
            n += (!(y >= -128.0f) || y > 127.0f) ? 1u : 0u;
        }
        off[t] = 2 * n > nChan ? x[(size_t)(nChan / 2) * nsamp + t] : 0.0f;
    }
    for (unsigned c = 0; c < nChan; ++c)
        for (unsigned t = 0; t < nsamp; ++t) {
            const float v = x[(size_t)c * nsamp + t];
            float y = std::nearbyint((off[t] != 0.0f ? v - off[t] : v) * inv_scale);
            if (!(y >= -128.0f)) y = -128.0f;
            if (y > 127.0f) y = 127.0f;
            q[(size_t)c * nsamp + t] = (uint8_t)((int)y + 128);
        }
}

// The u8 series: out[dm][t] = float(S - 128*nChan) * scale, plus, over the
// columns e with off[e] != 0 in increasing order and 0 <= e - t <=
// max delay of trial dm, off[e] * (number of channels with delay e - t),
// each product rounded before it is added.
extern "C" void orc_u8_series(const int32_t* sums, const float* off, const int32_t* delay, unsigned nChan,
                              unsigned nsamp, unsigned tdms, unsigned nOut, float scale, float* out)
{
    std::vector<int> rows;
    for (unsigned e = 0; e < nsamp; ++e)
        if (off[e] != 0.0f) rows.push_back((int)e);
    std::vector<int> len;
    for (unsigned dm = 0; dm < tdms; ++dm) {
        const int32_t* d = delay + (size_t)dm * nChan;
        int dmax = 0;
        for (unsigned c = 0; c < nChan; ++c) dmax = d[c] > dmax ? d[c] : dmax;
        len.assign((size_t)dmax + 1, 0);
        for (unsigned c = 0; c < nChan; ++c) ++len[(size_t)d[c]];
        for (unsigned t = 0; t < nOut; ++t) {
            float corr = 0.0f;
            for (int e : rows) {
                const int k = e - (int)t;
                if (k < 0) continue;
                if (k > dmax) break;
                if (len[(size_t)k] != 0) corr = corr + off[e] * (float)len[(size_t)k];
            }
            const size_t i = (size_t)dm * nOut + t;
            out[i] = (float)(sums[i] - (int32_t)(128 * nChan)) * scale + corr;
        }
    }
}

// DedispersionBuffer (src/lib/src/DedispersionBuffer.cpp:65-96,98-179),
// invert = 0, Stokes-I: spectra [t][c] are transposed into [c][t]; when the
// buffer holds nsampBuf samples it is emitted and its last `overlap`
// (= maxshift + remainingSamples, DedispersionModule.cpp:238) samples start
// the next buffer.  The weight term data - (W-1)*noise is the identity while
// the clipper's weights are 1 (RFI_Clipper.cpp:143,332) and is omitted.
struct orc_ddbuf {
    unsigned nChan, nsamp, overlap, fill;
    std::vector<float> buf;
};

extern "C" orc_ddbuf* orc_ddbuf_create(unsigned nChan, unsigned nsampBuf, unsigned overlap)
{
    orc_ddbuf* b = new orc_ddbuf;
    b->nChan = nChan; b->nsamp = nsampBuf; b->overlap = overlap; b->fill = 0;
    b->buf.assign((size_t)nChan * nsampBuf, 0.0f);
    return b;
}

extern "C" void orc_ddbuf_destroy(orc_ddbuf* b) { delete b; }

extern "C" int orc_ddbuf_add(orc_ddbuf* b, const float* spectra, unsigned nT, float* out, int max_out)
{
    int emitted = 0;
    unsigned t = 0;
    while (t < nT) {
        unsigned take = b->nsamp - b->fill;
        if (take > nT - t) take = nT - t;
        for (unsigned k = 0; k < take; ++k)
            for (unsigned c = 0; c < b->nChan; ++c)
                b->buf[(size_t)c * b->nsamp + b->fill + k] = spectra[(size_t)(t + k) * b->nChan + c];
        b->fill += take;
        t += take;
        if (b->fill == b->nsamp) {
            if (out && emitted < max_out)
                std::memcpy(out + (size_t)emitted * b->nChan * b->nsamp, b->buf.data(),
                            sizeof(float) * b->nChan * b->nsamp);
            ++emitted;
            for (unsigned c = 0; c < b->nChan; ++c)
                std::memmove(&b->buf[(size_t)c * b->nsamp], &b->buf[(size_t)c * b->nsamp + b->nsamp - b->overlap],
                             sizeof(float) * b->overlap);
            b->fill = b->overlap;
        }
    }
    return emitted;
}

// DedispersionAnalyser::analyse, src/lib/src/DedispersionAnalyser.cpp:30-203.
// Level thresholds use the float square root sqrtf((float)binFactor) (the
// reference's unqualified sqrt of a float: float or double overload is a
// header-inclusion question it leaves open; the two differ only for values
// within one float ulp of a threshold).
extern "C" int orc_detect(const float* data, int tdms, int nsamp, int nChannels, int nSubbands,
                          float threshold, unsigned binPow2, int32_t* ev_ints, float* ev_vals,
                          int max_events)
{
    int n = 0;
    auto add = [&](int dm, int t, int bf, float v) {
        if (n < max_events) {
            ev_ints[3 * n] = dm; ev_ints[3 * n + 1] = t; ev_ints[3 * n + 2] = bf; ev_vals[n] = v;
        }
        ++n;
    };
    float rms = std::sqrt((float)nChannels * (float)nSubbands);
    add(0, 0, 1, 0.0f);
    unsigned maxPow2 = 1u << binPow2;
    std::vector<std::vector<float>> binned(binPow2 + 1);
    unsigned cur = maxPow2;
    for (int lv = (int)binPow2; lv > -1; --lv) { binned[lv].resize(maxPow2 / cur); cur /= 2; }
    int widest = nsamp / (int)maxPow2;
    for (int dm = 0; dm < tdms; ++dm) {
        for (int i = 0; i < widest; ++i) {
            cur = maxPow2;
            for (int j = 0; j < (int)maxPow2; ++j) {
                int index = i * maxPow2 + j;
                binned[0][j] = data[(size_t)dm * nsamp + index];
                float detection = threshold * rms;
                if (binned[0][j] >= detection) add(dm, index, 1, binned[0][j]);
            }
            for (int lv = 1; lv < (int)binPow2 + 1; ++lv) {
                cur /= 2;
                for (int j = 0; j < (int)cur; ++j) {
                    int binFactor = maxPow2 / cur;
                    float detection = threshold * rms * std::sqrt((float)binFactor);
                    int index = i * maxPow2 + binFactor * j;
                    binned[lv][j] = binned[lv - 1][2 * j] + binned[lv - 1][2 * j + 1];
                    if (binned[lv][j] >= detection) add(dm, index, binFactor, binned[lv][j]);
                }
            }
        }
    }
    return n <= max_events ? n : -n;
}
