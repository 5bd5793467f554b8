// BandPassRecorder.cpp -- bandpass measurement (src/lib/src/BandPassRecorder.cpp:26-187).
#include "BandPassRecorder.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "../../../include/ppl_amd.h"
#include "DeviceMirror.h"

namespace pelican {
namespace ampp {

namespace {
std::runtime_error err(const std::string& m) { return std::runtime_error("BandPassRecorder: " + m); }
}  // namespace

BandPassRecorder::BandPassRecorder(const ConfigNode& config) : AbstractModule(config)
{
    // (:26-42)
    _requiredSamples = config.getOption("requiredSamples", "value", "20000").toULong();
    _polyDegree = (int)config.getOption("fitParmaters", "coefficients", "3").toUInt();
    if (config.getOption("Band", "startFrequency").isEmpty()) throw err("<Band startFrequency=?> not defined");
    _startFrequency = config.getOption("Band", "startFrequency").toFloat();
    if (config.getOption("Band", "endFrequency").isEmpty()) throw err("<Band endFrequency=?> not defined");
    _endFrequency = config.getOption("Band", "endFrequency").toFloat();
    if (_polyDegree < 1) throw err("<fitParmaters coefficients=?> must be at least 1");
}

BandPassRecorder::~BandPassRecorder() { ppl_device_free(_dSum); }

BinMap BandPassRecorder::binMap(unsigned nBins) const
{
    BinMap map(nBins);
    map.setStart(_startFrequency);
    map.setEnd(_endFrequency);
    return map;
}

void BandPassRecorder::_reset(unsigned nBins)
{
    // (:47-65): zero the integrated spectrum
    if (nBins != _nBins) {
        ppl_device_free(_dSum);
        _dSum = nullptr;
        _nBins = 0;
        void* p = nullptr;
        ppl_check(ppl_device_alloc(&p, (size_t)nBins * sizeof(float)), "BandPassRecorder");
        _dSum = static_cast<float*>(p);
        _nBins = nBins;
    }
    ppl_check(ppl_memset_async(_dSum, 0, (size_t)nBins * sizeof(float), nullptr), "BandPassRecorder");
    _totalSamples = 0;
}

bool BandPassRecorder::run(SpectrumDataSetStokes* stokesI, BandPass* bp)
{
    // (:67-97)
    const unsigned nSamples = stokesI->nTimeBlocks(), nSubbands = stokesI->nSubbands();
    const unsigned nChannels = stokesI->nChannels(), nPolarisations = stokesI->nPolarisations();
    const unsigned nBins = nChannels * nSubbands;
    if (nBins == 0 || nPolarisations == 0) throw err("empty Stokes blob");
    if (nBins != _nBins || _totalSamples == 0) _reset(nBins);
    if (nSamples) {
        const float* d_I = static_cast<const float*>(stokesI->deviceData(nullptr));
        ppl_check(ppl_bandpass_accumulate(d_I, (int)nSamples, (int)nSubbands, (int)nPolarisations, (int)nChannels,
                                          _dSum, nullptr),
                  "BandPassRecorder");
    }
    _totalSamples += nSamples;
    if (_totalSamples < _requiredSamples) return false;  // collect more data

    std::vector<float> sum(nBins);
    ppl_check(ppl_memcpy_d2h(sum.data(), _dSum, (size_t)nBins * sizeof(float), nullptr), "BandPassRecorder");
    ppl_check(ppl_stream_sync(nullptr), "BandPassRecorder");
    // averaged spectrum (:101-103): float / float(total)
    for (unsigned b = 0; b < nBins; ++b) sum[b] /= (float)_totalSamples;
    _lastFits = fitSpectrum(sum, binMap(nBins), bp);
    _totalSamples = 0;  // the next run integrates afresh
    return true;
}

int BandPassRecorder::fitSpectrum(std::vector<float>& sum, const BinMap& map, BandPass* bp) const
{
    const unsigned nBins = map.numberBins();
    if (sum.size() != nBins) throw err("spectrum and map sizes differ");
    std::vector<float> freq(nBins);
    for (unsigned i = 0; i < nBins; ++i) freq[i] = (float)map.binAssignmentNumber(i);

    // fit, trim points more than 2 residual RMS above the fit to the fit's
    // value, refit until a pass trims nothing (:106-129; the reference never
    // resets its trim count and sets a trimmed point to fit * totalSamples,
    // :126, so it cannot end: each pass counts its own trims here, and a
    // trimmed point takes the fit's value)
    std::vector<float> fit, residual(nBins);
    int fits = 0;
    unsigned trimmed;
    do {
        fit = _polyFit(freq, sum);
        ++fits;
        float rsum = 0.0f, rsumSquared = 0.0f;
        for (unsigned s = 0; s < nBins; ++s) {
            residual[s] = sum[s] - _theFit(fit, freq[s]);
            rsum += residual[s];
            rsumSquared = (float)((double)rsumSquared + std::pow((double)residual[s], 2.0));
        }
        const float residualRms =
            (float)std::sqrt((double)(rsumSquared / (float)nBins) - std::pow((double)(rsum / (float)nBins), 2.0));
        const float margin = 2.0f * residualRms;
        trimmed = 0;
        for (unsigned s = 0; s < nBins; ++s) {
            if (residual[s] > margin) {
                ++trimmed;
                sum[s] = _theFit(fit, freq[s]);
            }
        }
    } while (trimmed && fits <= (int)nBins);

    // RMS of the cleaned spectrum (:131-137)
    float sumInt = 0.0f, sumSquared = 0.0f;
    for (unsigned s = 0; s < nBins; ++s) {
        sumInt += sum[s];
        sumSquared = (float)((double)sumSquared + std::pow((double)sum[s], 2.0));
    }
    bp->setData(map, fit);
    bp->setRMS(
        (float)std::sqrt((double)(sumSquared / (float)nBins) - std::pow((double)(sumInt / (float)nBins), 2.0)));
    return fits;
}

float BandPassRecorder::_theFit(const std::vector<float>& fit, float v)
{
    // (:141-148): float total, each term in double
    float tot = 0.0f;
    for (size_t i = 0; i < fit.size(); ++i) tot += fit[i] * std::pow((double)v, (double)i);
    return tot;
}

std::vector<float> BandPassRecorder::_polyFit(const std::vector<float>& freq, const std::vector<float>& y) const
{
    // The least-squares solution of A c = y, A[i][p] = float(freq[i]^p)
    // (the reference's design matrix, :56-62, for LAPACK sgels at
    // :150-187), here by Householder QR in double on unit-norm columns.
    const int n = (int)y.size(), m = _polyDegree;
    if (n < m) throw err("fewer bins than polynomial coefficients");
    std::vector<double> A((size_t)n * m), b(y.begin(), y.end()), scale(m);
    for (int p = 0; p < m; ++p) {
        double norm = 0.0;
        for (int i = 0; i < n; ++i) {
            const double a = (double)(float)std::pow((double)freq[i], (double)p);
            A[(size_t)p * n + i] = a;
            norm += a * a;
        }
        scale[p] = norm > 0.0 ? 1.0 / std::sqrt(norm) : 1.0;
        for (int i = 0; i < n; ++i) A[(size_t)p * n + i] *= scale[p];
    }
    std::vector<double> diag(m);
    for (int k = 0; k < m; ++k) {
        double* col = &A[(size_t)k * n];
        double alpha = 0.0;
        for (int i = k; i < n; ++i) alpha += col[i] * col[i];
        alpha = std::sqrt(alpha);
        if (col[k] > 0.0) alpha = -alpha;
        // v = col[k..] - alpha e_k, stored in place
        col[k] -= alpha;
        double vnorm2 = 0.0;
        for (int i = k; i < n; ++i) vnorm2 += col[i] * col[i];
        diag[k] = alpha;
        if (vnorm2 == 0.0) continue;
        auto reflect = [&](double* x) {
            double d = 0.0;
            for (int i = k; i < n; ++i) d += col[i] * x[i];
            d = 2.0 * d / vnorm2;
            for (int i = k; i < n; ++i) x[i] -= d * col[i];
        };
        for (int j = k + 1; j < m; ++j) reflect(&A[(size_t)j * n]);
        reflect(b.data());
    }
    double rmax = 0.0;
    for (int k = 0; k < m; ++k) rmax = std::max(rmax, std::fabs(diag[k]));
    for (int k = 0; k < m; ++k)
        if (!(std::fabs(diag[k]) > 1e-13 * rmax))
            throw err("the diagonal element " + std::to_string(k + 1) +
                      " of the triangular factor of A is zero, so that A does not have full rank; "
                      "the least squares solution could not be computed.");
    // back substitution R c = Q^T y (R's diagonal in diag, above it in A)
    std::vector<double> c(m);
    for (int k = m - 1; k >= 0; --k) {
        double s = b[k];
        for (int j = k + 1; j < m; ++j) s -= A[(size_t)j * n + k] * c[j];
        c[k] = s / diag[k];
    }
    std::vector<float> out(m);
    for (int p = 0; p < m; ++p) out[p] = (float)(c[p] * scale[p]);
    return out;
}

}  // namespace ampp
}  // namespace pelican
