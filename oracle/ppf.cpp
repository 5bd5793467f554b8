// oracle/ppf.cpp -- CPU restatement of the polyphase channeliser, Stokes-I
// and time integration.  TEST INFRASTRUCTURE ONLY (see oracle.h).
#include "oracle.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <complex>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// PolyphaseCoefficients::_besselI0, src/lib/src/PolyphaseCoefficients.cpp:77-103
double besselI0(double x)
{
    const double p1 = 1.0, p2 = 3.5156229, p3 = 3.0899424, p4 = 1.2067492,
                 p5 = 0.2659732, p6 = 3.60768e-2, p7 = 4.5813e-3;
    const double q1 = 0.39894228, q2 = 1.328592e-2, q3 = 2.25319e-3,
                 q4 = -1.57565e-3, q5 = 9.16281e-3, q6 = -2.057706e-2,
                 q7 = 2.635537e-2, q8 = -1.647633e-2, q9 = 3.92377e-3;
    const double k1 = 3.75;
    double ax = std::fabs(x);
    if (ax < k1) {
        double xx = x / k1;
        double y = xx * xx;
        return p1 + y * (p2 + y * (p3 + y * (p4 + y * (p5 + y * (p6 + y * p7)))));
    }
    double y = k1 / ax;
    return (std::exp(ax) / std::sqrt(ax)) *
           (q1 + y * (q2 + y * (q3 + y * (q4 + y * (q5 + y * (q6 + y * (q7 + y * (q8 + y * q9))))))));
}

// _kaiser :114-127
void kaiser(int n, double beta, double* d)
{
    if (n == 1) { d[0] = 1.0; return; }
    int m = n - 1;
    for (int i = 0; i < n; i++) {
        double k = 2.0 * beta / m * std::sqrt(double(i * (m - i)));
        d[i] = besselI0(k) / besselI0(beta);
    }
}

// _gaussian :130-136
void gaussian(int n, double a, double* d)
{
    int index = 0;
    for (int i = -(n - 1); i <= n - 1; i += 2)
        d[index++] = std::exp(-0.5 * std::pow((a / n * i), 2));
}

// _hamming :139-149
void hamming(unsigned n, double* d)
{
    if (n == 1) { d[0] = 1.0; return; }
    unsigned m = n - 1;
    for (unsigned i = 0; i < n; i++) d[i] = 0.54 - 0.46 * std::cos((2.0 * M_PI * i) / m);
}

// _blackman :152-163, including its integer division k = i/m, which makes
// the window 0.0 everywhere except the last point (a reference quirk).
void blackman(unsigned n, double* d)
{
    if (n == 1) { d[0] = 1.0; return; }
    unsigned m = n - 1;
    for (unsigned i = 0; i < n; i++) {
        double k = i / m;
        d[i] = 0.42 - 0.5 * std::cos(2.0 * M_PI * k) + 0.08 * std::cos(4.0 * M_PI * k);
    }
}

// Hann window.  The reference has no Hann window; this library defines it as
// 0.5 - 0.5 cos(2 pi i/(n-1)), on the same fir2 path (parity unpinned by
// the reference; pinned by symmetry/endpoint tests).
void hann(unsigned n, double* d)
{
    if (n == 1) { d[0] = 1.0; return; }
    unsigned m = n - 1;
    for (unsigned i = 0; i < n; i++) d[i] = 0.5 - 0.5 * std::cos((2.0 * M_PI * i) / m);
}

// _interpolate :170-196
void interpolate(const double* x, const double* y, unsigned nX, unsigned n, double* result)
{
    unsigned nextX = 0, index = 0;
    for (double ix = 0.0; ix <= 1.0; ix += 1.0 / (n - 1), index++) {
        while (x[nextX] <= ix && nextX < nX - 1) nextX++;
        double rc = (y[nextX] - y[nextX - 1]) / (x[nextX] - x[nextX - 1]);
        result[index] = y[nextX - 1] + (ix - x[nextX - 1]) * rc;
    }
}

unsigned nextPowerOf2(unsigned n)
{
    unsigned res = 1;
    while (res < n) res *= 2;
    return res;
}

// _generateFirFilter :216-304 with the inverse transform evaluated as an
// FP64 DFT of the (real, symmetric) 4*grid_n input.  Only the real part of
// the needed outputs is formed, with the phase index reduced modulo M.
void fir2(unsigned n, double w, const double* window, double* result)
{
    unsigned grid_n = nextPowerOf2(n + 1);
    unsigned ramp_n = 2;
    // ramp_n/grid_n is integer division in the reference (0 for grid_n > 2)
    double f[] = {0.0, w - ramp_n / grid_n / 2.0, w, w + ramp_n / grid_n / 2.0, 1.0};
    double m[] = {1.0, 1.0, 0.0, 0.0, 0.0};
    std::vector<double> grid(grid_n + 1);
    interpolate(f, m, 5, grid_n + 1, grid.data());

    const unsigned M = 4 * grid_n;
    std::vector<double> cin(M, 0.0);
    for (unsigned i = 0; i < grid_n + 1; i++) cin[i] = double(float(grid[i]));
    for (unsigned i = grid_n - 1, idx = 0; i >= 1; i--, idx++)
        cin[grid_n * 3 + 1 + idx] = double(float(grid[i]));
    std::vector<unsigned> nz;
    for (unsigned k = 0; k < M; ++k) if (cin[k] != 0.0) nz.push_back(k);

    auto out_re = [&](unsigned i) {
        // backward transform: sum_k cin[k] * exp(+2 pi i k / M), real part
        long double acc = 0.0L;
        for (unsigned k : nz) {
            unsigned long long ph = ((unsigned long long)i * k) % M;
            acc += (long double)cin[k] * cosl(2.0L * (long double)M_PI * (long double)ph / (long double)M);
        }
        return (double)acc;
    };
    unsigned index = 0;
    for (unsigned i = M - n; i < M; i += 2) result[index++] = out_re(i);
    for (unsigned i = 1; i <= n; i += 2) result[index++] = out_re(i);

    for (unsigned i = 0; i <= n; i++) result[i] *= window[i];
    double factor = result[n / 2];
    for (unsigned i = 0; i <= n; i++) result[i] /= factor;
}

}  // namespace

extern "C" int orc_fir2_unwindowed(unsigned nTaps, unsigned nCh, double* out)
{
    unsigned n = nCh * nTaps;
    std::vector<double> win(n, 1.0);
    fir2(n - 1, 1.0 / nCh, win.data(), out);
    return 0;
}

// PolyphaseCoefficients::genereateFilter :308-364
extern "C" int orc_ppf_coefficients(unsigned nTaps, unsigned nCh, int window, double* out)
{
    unsigned n = nCh * nTaps;
    std::vector<double> win(n);
    switch (window) {
        case ORC_WIN_HAMMING: hamming(n, win.data()); break;
        case ORC_WIN_BLACKMAN: blackman(n, win.data()); break;
        case ORC_WIN_GAUSSIAN: gaussian(n, 3.5, win.data()); break;
        case ORC_WIN_KAISER: kaiser(n, 9.0695, win.data()); break;
        case ORC_WIN_HANN: hann(n, win.data()); break;
        default: return -1;
    }
    std::vector<double> result(n);
    fir2(n - 1, 1.0 / nCh, win.data(), result.data());
    for (unsigned t = 0; t < nTaps; ++t)
        for (unsigned c = 0; c < nCh; ++c) {
            unsigned i = t * nCh + c;
            out[i] = (c % 2 == 0) ? result[i] / nCh : -result[i] / nCh;
        }
    return 0;
}

// PPFChanneliser::_assign_threads, src/lib/src/PPFChanneliser.cpp:298-319
extern "C" int orc_assign_threads(unsigned* start, unsigned* end, unsigned nValues,
                                  unsigned nThreads, unsigned threadId)
{
    if (threadId >= nThreads) return -1;
    if (threadId >= nValues) { *start = *end = 0; return 0; }
    unsigned number = nValues / nThreads;
    unsigned remainder = nValues % nThreads;
    if (threadId < remainder) number++;
    *start = threadId * number;
    if (threadId >= remainder) *start += remainder;
    *end = *start + number;
    return 0;
}

struct orc_ppf {
    unsigned nSub, nPol, nCh, nTaps;
    std::vector<float> h32;
    std::vector<double> h64;
    // FIR history per (sub,pol): nTaps blocks of nCh complex samples, and the
    // ring position of the oldest block.  The reference keeps one ring index
    // per OpenMP thread (PPFChanneliser.cpp:40,250-253); with nBlk % nTaps == 0
    // every stream sees the same sequence, so one index per stream is
    // equivalent (asserted by the callers).
    std::vector<std::complex<float>> hist32;
    std::vector<std::complex<double>> hist64;
    std::vector<unsigned> oldest32, oldest64;
};

extern "C" orc_ppf* orc_ppf_create(unsigned nSub, unsigned nPol, unsigned nCh, unsigned nTaps,
                                   const float* c32, const double* c64)
{
    orc_ppf* p = new orc_ppf;
    p->nSub = nSub; p->nPol = nPol; p->nCh = nCh; p->nTaps = nTaps;
    p->h32.assign(c32, c32 + nTaps * nCh);
    p->h64.assign(c64, c64 + nTaps * nCh);
    p->hist32.assign((size_t)nSub * nPol * nTaps * nCh, {0.f, 0.f});
    p->hist64.assign((size_t)nSub * nPol * nTaps * nCh, {0.0, 0.0});
    p->oldest32.assign(nSub * nPol, 0);
    p->oldest64.assign(nSub * nPol, 0);
    return p;
}

extern "C" void orc_ppf_destroy(orc_ppf* p) { delete p; }

namespace {

// FP64 forward DFT, unnormalised (the closed form of FFTW_FORWARD), with the
// twiddles W_N^k tabulated in long double once per N.
void dft64(const std::complex<double>* x, unsigned N, std::complex<double>* y)
{
    static thread_local unsigned tabN = 0;
    static thread_local std::vector<long double> tc, ts;
    if (tabN != N) {
        tc.resize(N);
        ts.resize(N);
        for (unsigned k = 0; k < N; ++k) {
            long double a = -2.0L * (long double)M_PI * (long double)k / (long double)N;
            tc[k] = cosl(a);
            ts[k] = sinl(a);
        }
        tabN = N;
    }
    for (unsigned k = 0; k < N; ++k) {
        long double re = 0, im = 0;
        for (unsigned n = 0; n < N; ++n) {
            const unsigned ph = (unsigned)(((unsigned long long)n * k) % N);
            const long double c = tc[ph], s = ts[ph];
            re += (long double)x[n].real() * c - (long double)x[n].imag() * s;
            im += (long double)x[n].real() * s + (long double)x[n].imag() * c;
        }
        y[k] = {(double)re, (double)im};
    }
}

// One FP32 FIR step for one stream, PPFChanneliser::_updateBuffer :247-254
// then _filter :266-290 (separate multiply and add, oldest block first).
void fir_step32(orc_ppf* p, unsigned stream, const int16_t* blk, std::complex<float>* filt)
{
    const unsigned N = p->nCh, T = p->nTaps;
    std::complex<float>* buf = &p->hist32[(size_t)stream * T * N];
    unsigned& oldest = p->oldest32[stream];
    for (unsigned c = 0; c < N; ++c)
        buf[oldest * N + c] = {(float)blk[2 * c], (float)blk[2 * c + 1]};
    oldest = (oldest + 1) % T;
    for (unsigned c = 0; c < N; ++c) filt[c] = {0.f, 0.f};
    unsigned i = 0;
    for (unsigned t = 0; t < T; ++t) {
        unsigned ib = ((oldest + t) % T) * N;
        for (unsigned c = 0; c < N; ++c) {
            float coeff = p->h32[i];
            float re = buf[ib + c].real() * coeff;
            float im = buf[ib + c].imag() * coeff;
            filt[c] = {filt[c].real() + re, filt[c].imag() + im};
            i++;
        }
    }
}

void fir_step64(orc_ppf* p, unsigned stream, const int16_t* blk, std::complex<double>* filt)
{
    const unsigned N = p->nCh, T = p->nTaps;
    std::complex<double>* buf = &p->hist64[(size_t)stream * T * N];
    unsigned& oldest = p->oldest64[stream];
    for (unsigned c = 0; c < N; ++c)
        buf[oldest * N + c] = {(double)blk[2 * c], (double)blk[2 * c + 1]};
    oldest = (oldest + 1) % T;
    for (unsigned c = 0; c < N; ++c) filt[c] = {0.0, 0.0};
    unsigned i = 0;
    for (unsigned t = 0; t < T; ++t) {
        unsigned ib = ((oldest + t) % T) * N;
        for (unsigned c = 0; c < N; ++c) { filt[c] += buf[ib + c] * p->h64[i]; i++; }
    }
}

}  // namespace

extern "C" void orc_ppf_fir_only_f32(orc_ppf* p, const int16_t* in, unsigned nBlk, float* out)
{
    const unsigned N = p->nCh, S = p->nSub, P = p->nPol;
    std::vector<std::complex<float>> filt(N);
    for (unsigned s = 0; s < S; ++s)
        for (unsigned pol = 0; pol < P; ++pol) {
            unsigned stream = s * P + pol;
            const int16_t* x = in + 2 * (size_t)stream * nBlk * N;
            for (unsigned b = 0; b < nBlk; ++b) {
                fir_step32(p, stream, x + 2 * (size_t)b * N, filt.data());
                float* o = out + 2 * (size_t)N * (P * ((size_t)S * b + s) + pol);
                for (unsigned c = 0; c < N; ++c) { o[2 * c] = filt[c].real(); o[2 * c + 1] = filt[c].imag(); }
            }
        }
}

extern "C" void orc_ppf_run_fir32(orc_ppf* p, const int16_t* in, unsigned nBlk, double* out)
{
    const unsigned N = p->nCh, S = p->nSub, P = p->nPol;
    std::vector<std::complex<float>> filt(N);
    std::vector<std::complex<double>> x(N), y(N);
    for (unsigned s = 0; s < S; ++s)
        for (unsigned pol = 0; pol < P; ++pol) {
            unsigned stream = s * P + pol;
            const int16_t* xin = in + 2 * (size_t)stream * nBlk * N;
            for (unsigned b = 0; b < nBlk; ++b) {
                fir_step32(p, stream, xin + 2 * (size_t)b * N, filt.data());
                for (unsigned c = 0; c < N; ++c) x[c] = {filt[c].real(), filt[c].imag()};
                dft64(x.data(), N, y.data());
                double* o = out + 2 * (size_t)N * (P * ((size_t)S * b + s) + pol);
                for (unsigned c = 0; c < N; ++c) { o[2 * c] = y[c].real(); o[2 * c + 1] = y[c].imag(); }
            }
        }
}

extern "C" void orc_ppf_run_f64(orc_ppf* p, const int16_t* in, unsigned nBlk, double* out)
{
    const unsigned N = p->nCh, S = p->nSub, P = p->nPol;
    std::vector<std::complex<double>> filt(N), y(N);
    for (unsigned s = 0; s < S; ++s)
        for (unsigned pol = 0; pol < P; ++pol) {
            unsigned stream = s * P + pol;
            const int16_t* xin = in + 2 * (size_t)stream * nBlk * N;
            for (unsigned b = 0; b < nBlk; ++b) {
                fir_step64(p, stream, xin + 2 * (size_t)b * N, filt.data());
                dft64(filt.data(), N, y.data());
                double* o = out + 2 * (size_t)N * (P * ((size_t)S * b + s) + pol);
                for (unsigned c = 0; c < N; ++c) { o[2 * c] = y[c].real(); o[2 * c + 1] = y[c].imag(); }
            }
        }
}

// StokesGenerator::run (Stokes-I), src/lib/src/StokesGenerator.cpp:88-99
extern "C" void orc_stokes_i_f32(const float* in, unsigned nBlk, unsigned nSub, unsigned nCh, float* out)
{
    for (unsigned t = 0; t < nBlk; ++t)
        for (unsigned s = 0; s < nSub; ++s) {
            const float* X = in + 2 * (size_t)nCh * (2 * ((size_t)nSub * t + s) + 0);
            const float* Y = in + 2 * (size_t)nCh * (2 * ((size_t)nSub * t + s) + 1);
            float* I = out + (size_t)nCh * ((size_t)nSub * t + s);
            for (unsigned c = 0; c < nCh; ++c) {
                float Xr = X[2 * c], Xi = X[2 * c + 1], Yr = Y[2 * c], Yi = Y[2 * c + 1];
                float powerX = Xr * Xr + Xi * Xi;
                float powerY = Yr * Yr + Yi * Yi;
                I[c] = powerX + powerY;
            }
        }
}

extern "C" void orc_stokes_i_f64(const double* in, unsigned nBlk, unsigned nSub, unsigned nCh, double* out)
{
    for (unsigned t = 0; t < nBlk; ++t)
        for (unsigned s = 0; s < nSub; ++s) {
            const double* X = in + 2 * (size_t)nCh * (2 * ((size_t)nSub * t + s) + 0);
            const double* Y = in + 2 * (size_t)nCh * (2 * ((size_t)nSub * t + s) + 1);
            double* I = out + (size_t)nCh * ((size_t)nSub * t + s);
            for (unsigned c = 0; c < nCh; ++c)
                I[c] = (X[2 * c] * X[2 * c] + X[2 * c + 1] * X[2 * c + 1]) +
                       (Y[2 * c] * Y[2 * c] + Y[2 * c + 1] * Y[2 * c + 1]);
        }
}

// StokesIntegrator::run, src/lib/src/StokesIntegrator.cpp:53-114
template <typename T>
static void integrate(const T* in, unsigned nT, unsigned nSub, unsigned nPol, unsigned nCh,
                      unsigned window, unsigned binCh, T* out)
{
    if (window > nT) window = nT;
    unsigned newT = nT / window, newC = nCh / binCh;
    for (size_t i = 0; i < (size_t)newT * nSub * nPol * newC; ++i) out[i] = 0;
    unsigned tStart = 0;
    for (unsigned u = 0; u < newT; ++u) {
        for (unsigned t = tStart; t < window + tStart; ++t)
            for (unsigned s = 0; s < nSub; ++s)
                for (unsigned p = 0; p < nPol; ++p) {
                    const T* v = in + (size_t)nCh * (nPol * ((size_t)nSub * t + s) + p);
                    T* o = out + (size_t)newC * (nPol * ((size_t)nSub * u + s) + p);
                    unsigned cs = 0;
                    for (unsigned nc = 0; nc < newC; ++nc) {
                        for (unsigned c = cs; c < cs + binCh; ++c) o[nc] += v[c];
                        cs += binCh;
                    }
                }
        tStart += window;
    }
}

extern "C" void orc_integrate_f32(const float* in, unsigned nT, unsigned nSub, unsigned nPol, unsigned nCh,
                                  unsigned window, unsigned binCh, float* out)
{
    integrate<float>(in, nT, nSub, nPol, nCh, window, binCh, out);
}

extern "C" void orc_integrate_f64(const double* in, unsigned nT, unsigned nSub, unsigned nPol, unsigned nCh,
                                  unsigned window, unsigned binCh, double* out)
{
    integrate<double>(in, nT, nSub, nPol, nCh, window, binCh, out);
}

// FP32 radix-2 FFT (iterative, decimation in time) standing in for the
// reference's FFTW float plan when the CPU path is timed.
static void fft32(std::complex<float>* x, unsigned N, const std::vector<std::complex<float>>& tw)
{
    for (unsigned i = 1, j = 0; i < N; ++i) {
        unsigned bit = N >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(x[i], x[j]);
    }
    for (unsigned len = 2; len <= N; len <<= 1) {
        const unsigned step = N / len;
        for (unsigned i = 0; i < N; i += len)
            for (unsigned k = 0; k < len / 2; ++k) {
                const std::complex<float> w = tw[k * step];
                const std::complex<float> u = x[i + k], v = x[i + k + len / 2] * w;
                x[i + k] = u + v;
                x[i + k + len / 2] = u - v;
            }
    }
}

// PPFChanneliser::run with an FP32 FFT, OpenMP over streams (the reference
// parallelises over subbands, PPFChanneliser.cpp:139-183).  Writes complex
// float spectra [blk][sub][pol][nCh].  Used to time the CPU baseline.
extern "C" void orc_ppf_run_f32(orc_ppf* p, const int16_t* in, unsigned nBlk, float* out, int nthreads)
{
    const unsigned N = p->nCh, S = p->nSub, P = p->nPol;
    std::vector<std::complex<float>> tw(N);
    for (unsigned k = 0; k < N; ++k) {
        const double a = -2.0 * M_PI * k / N;
        tw[k] = {(float)std::cos(a), (float)std::sin(a)};
    }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
#endif
    for (int stream = 0; stream < (int)(S * P); ++stream) {
        const unsigned s = stream / P, pol = stream % P;
        std::vector<std::complex<float>> filt(N);
        const int16_t* x = in + 2 * (size_t)stream * nBlk * N;
        for (unsigned b = 0; b < nBlk; ++b) {
            fir_step32(p, stream, x + 2 * (size_t)b * N, filt.data());
            fft32(filt.data(), N, tw);
            float* o = out + 2 * (size_t)N * (P * ((size_t)S * b + s) + pol);
            for (unsigned c = 0; c < N; ++c) { o[2 * c] = filt[c].real(); o[2 * c + 1] = filt[c].imag(); }
        }
    }
}
