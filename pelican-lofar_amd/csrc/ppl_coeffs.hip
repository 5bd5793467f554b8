// ppl_coeffs.hip -- FP64 polyphase FIR coefficients, host code
// (PolyphaseCoefficients::genereateFilter, src/lib/src/PolyphaseCoefficients.cpp:308-364).
//
// Frequency-sampled low-pass (fir2 with cut-off 1/nCh on a grid of
// nextPow2(nTaps*nCh)+1 points, src/lib/src/PolyphaseCoefficients.cpp:216-304)
// times a window, normalised by the tap at n/2, scaled by 1/nCh and given
// the alternating sign that fftshifts the channeliser output.  The inverse
// transform of the 4*grid input is real and even, so each needed output is
// a cosine sum over the non-zero grid points, evaluated here in FP64 with
// the phase reduced modulo the transform length (the reference evaluates it
// with an FFTW single-precision plan).
#include "ppl_internal.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

double bessel_i0(double x)
{
    // polynomial approximation of I0 used by the reference (:77-103)
    const double ax = std::fabs(x);
    if (ax < 3.75) {
        const double y = (x / 3.75) * (x / 3.75);
        return 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 + y * (3.60768e-2 + y * 4.5813e-3)))));
    }
    const double y = 3.75 / ax;
    return (std::exp(ax) / std::sqrt(ax)) *
           (0.39894228 + y * (1.328592e-2 + y * (2.25319e-3 + y * (-1.57565e-3 + y * (9.16281e-3 + y * (-2.057706e-2 +
            y * (2.635537e-2 + y * (-1.647633e-2 + y * 3.92377e-3))))))));
}

bool make_window(int type, unsigned n, std::vector<double>& w)
{
    w.assign(n, 1.0);
    if (n == 1) return type >= 0 && type <= 4;
    const unsigned m = n - 1;
    switch (type) {
        case 0: {  // kaiser, beta 9.0695 (:114-127, :336)
            const double beta = 9.0695;
            for (unsigned i = 0; i < n; ++i) {
                const double k = 2.0 * beta / (int)m * std::sqrt(double((int)i * ((int)m - (int)i)));
                w[i] = bessel_i0(k) / bessel_i0(beta);
            }
            return true;
        }
        case 1: {  // gaussian, alpha 3.5 (:130-136)
            unsigned idx = 0;
            for (int i = -(int)(n - 1); i <= (int)n - 1; i += 2) {
                const double a = 3.5 / (int)n * i;
                w[idx++] = std::exp(-0.5 * a * a);
            }
            return true;
        }
        case 2:  // blackman with the reference's integer k = i/m (:152-163)
            for (unsigned i = 0; i < n; ++i) {
                const double k = (double)(i / m);
                w[i] = 0.42 - 0.5 * std::cos(2.0 * M_PI * k) + 0.08 * std::cos(4.0 * M_PI * k);
            }
            return true;
        case 3:  // hamming (:139-149)
            for (unsigned i = 0; i < n; ++i) w[i] = 0.54 - 0.46 * std::cos((2.0 * M_PI * i) / m);
            return true;
        case 4:  // hann: not in the reference; 0.5 - 0.5 cos(2 pi i/(n-1)) on the same path
            for (unsigned i = 0; i < n; ++i) w[i] = 0.5 - 0.5 * std::cos((2.0 * M_PI * i) / m);
            return true;
    }
    return false;
}

}  // namespace

extern "C" int ppl_ppf_coefficients(unsigned nTaps, unsigned nCh, int window, double* out)
{
    PPL_CHECK_ARG(out && nTaps > 0 && nCh > 0, "bad coefficient request");
    const unsigned n = nTaps * nCh;
    std::vector<double> win;
    PPL_CHECK_ARG(make_window(window, n, win), "Unknown coefficient window type %d", window);
    const unsigned order = n - 1;  // _generateFirFilter(n - 1, 1/nCh, ...)
    const double w = 1.0 / nCh;
    unsigned grid_n = 1;
    while (grid_n < order + 1) grid_n *= 2;
    // grid: linear interpolation of (0,1) (w,1) (w,0) (w,0) (1,0) at
    // 0, 1/grid_n, ... with the reference's accumulated abscissa (:170-196);
    // the ramp ramp_n/grid_n is integer division (0 for grid_n > 2)
    const double ramp = (double)(2u / grid_n) / 2.0;
    const double fx[5] = {0.0, w - ramp, w, w + ramp, 1.0};
    const double my[5] = {1.0, 1.0, 0.0, 0.0, 0.0};
    std::vector<double> grid(grid_n + 1, 0.0);
    {
        unsigned nextX = 0, idx = 0;
        for (double ix = 0.0; ix <= 1.0 && idx <= grid_n; ix += 1.0 / grid_n, ++idx) {
            while (fx[nextX] <= ix && nextX < 4) ++nextX;
            const double rc = (my[nextX] - my[nextX - 1]) / (fx[nextX] - fx[nextX - 1]);
            grid[idx] = my[nextX - 1] + (ix - fx[nextX - 1]) * rc;
        }
    }
    // the transform input is float(grid) mirrored: [g0..g_G, 0.., g_{G-1}..g_1]
    const unsigned M = 4 * grid_n;
    std::vector<unsigned> k_nz;
    std::vector<double> v_nz;
    for (unsigned k = 0; k <= grid_n; ++k) {
        const double g = (double)(float)grid[k];
        if (g != 0.0) {
            k_nz.push_back(k);
            v_nz.push_back(g);
            if (k != 0 && k != grid_n) {
                k_nz.push_back(M - k);
                v_nz.push_back(g);
            }
        }
    }
    auto out_re = [&](unsigned i) {
        double acc = 0.0;
        for (size_t j = 0; j < k_nz.size(); ++j) {
            const unsigned long long ph = ((unsigned long long)i * k_nz[j]) % M;
            acc += v_nz[j] * std::cos(2.0 * M_PI * (double)ph / (double)M);
        }
        return acc;
    };
    std::vector<double> result(n);
    unsigned index = 0;
    for (unsigned i = M - order; i < M; i += 2) result[index++] = out_re(i);
    for (unsigned i = 1; i <= order; i += 2) result[index++] = out_re(i);
    for (unsigned i = 0; i <= order; ++i) result[i] *= win[i];
    const double factor = result[order / 2];
    for (unsigned i = 0; i <= order; ++i) result[i] /= factor;
    for (unsigned t = 0; t < nTaps; ++t)
        for (unsigned c = 0; c < nCh; ++c) {
            const unsigned i = t * nCh + c;
            out[i] = (c % 2 == 0) ? result[i] / nCh : -result[i] / nCh;
        }
    return PPL_OK;
}
