#include "PPFChanneliser.h"

#include <stdexcept>

#include "../../../include/ppl_amd.h"

namespace pelican {
namespace ampp {

namespace {
std::runtime_error err(const std::string& m) { return std::runtime_error("PPFChanneliser: " + m); }
}  // namespace

PPFChanneliser::PPFChanneliser(const ConfigNode& config) : AbstractModule(config)
{
    _nChannels = config.getOption("outputChannelsPerSubband", "value", "512").toUInt();
    _nTaps = config.getOption("filter", "nTaps", "8").toUInt();
    const std::string window = config.getOption("filter", "filterWindow", "kaiser").toLower();
    const std::string prec = config.getOption("dftPrecision", "value", "fp32").toLower();
    if (_nChannels != 1 && _nChannels % 2 == 1) throw err("Number of channels needs to be even.");
    int w;
    if (window == "kaiser") w = PPL_WIN_KAISER;
    else if (window == "gaussian") w = PPL_WIN_GAUSSIAN;
    else if (window == "blackman") w = PPL_WIN_BLACKMAN;
    else if (window == "hamming") w = PPL_WIN_HAMMING;
    else if (window == "hann") w = PPL_WIN_HANN;
    else throw err("Unknown coefficient window type '" + window + "'.");
    if (prec == "fp32") _precision = PPL_DFT_FP32;
    else if (prec == "bf16x3") _precision = PPL_DFT_BF16X3;
    else if (prec == "fp16x2") _precision = PPL_DFT_FP16X2;
    else throw err("Unknown DFT precision '" + prec + "'.");
    // PolyphaseCoefficients::genereateFilter in FP64, then cast to float
    // (PPFChanneliser.cpp:195-217)
    std::vector<double> c((size_t)_nTaps * _nChannels);
    ppl_check(ppl_ppf_coefficients(_nTaps, _nChannels, w, c.data()), "PPFChanneliser: coefficients");
    _coeffs.resize(c.size());
    for (size_t i = 0; i < c.size(); ++i) _coeffs[i] = (float)c[i];
}

PPFChanneliser::~PPFChanneliser() { ppl_ppf_plan_destroy(_plan); }

void PPFChanneliser::_checkData(const TimeSeriesDataSetC32* timeData) const
{
    if (!timeData) throw err("Time stream data blob missing.");
    if (!timeData->size()) throw err("Empty time data blob");
    if (!timeData->nSubbands()) throw err("Empty time data blob");
    if (!timeData->nPolarisations()) throw err("Empty time data blob");
    if (!timeData->nTimeBlocks()) throw err("Empty time data blob");
    if (_nChannels > 1 && timeData->nTimesPerBlock() != _nChannels)
        throw err("Dimension mismatch: " + std::to_string(timeData->nTimesPerBlock()) +
                  " samples per block != number of output channels " + std::to_string(_nChannels) + ".");
}

void PPFChanneliser::run(const TimeSeriesDataSetC32* timeSeries, SpectrumDataSetC32* spectra)
{
    _checkData(timeSeries);
    const unsigned nSub = timeSeries->nSubbands(), nPol = timeSeries->nPolarisations();
    const unsigned nBlk = timeSeries->nTimeBlocks(), nTimes = timeSeries->nTimesPerBlock();
    spectra->setLofarTimestamp(timeSeries->getLofarTimestamp());
    spectra->setBlockRate(timeSeries->getBlockRate() * _nChannels);
    if (_nChannels == 1) {
        // passthrough (:108-130): every sample becomes a one-channel
        // spectrum; the output holds nTimeBlocks * nTimesPerBlock spectra
        spectra->resize(nBlk * nTimes, nSub, nPol, 1);
        const std::complex<float>* in = timeSeries->constData();
        std::complex<float>* out = spectra->data();
        for (unsigned s = 0; s < nSub; ++s)
            for (unsigned p = 0; p < nPol; ++p)
                for (unsigned b = 0; b < nBlk; ++b)
                    for (unsigned t = 0; t < nTimes; ++t)
                        out[SpectrumDataSetBase::index(s, nSub, p, nPol, nTimes * b + t, 1)] =
                            in[TimeSeriesDataSetC32::index(s, nTimes, p, nPol, b, nBlk) + t];
        return;
    }
    spectra->resize(nBlk, nSub, nPol, _nChannels);
    if (!_plan || _planSub != nSub || _planPol != nPol) {
        ppl_ppf_plan_destroy(_plan);
        _plan = nullptr;
        ppl_ppf_config cfg{};
        cfg.nSub = nSub;
        cfg.nPol = nPol;
        cfg.nCh = _nChannels;
        cfg.nTaps = _nTaps;
        cfg.coeffs = _coeffs.data();
        cfg.precision = _precision;
        cfg.integrate = 1;
        ppl_check(ppl_ppf_plan_create(&_plan, &cfg), "PPFChanneliser");
        _planSub = nSub;
        _planPol = nPol;
    }
    const float* d_in = static_cast<const float*>(timeSeries->deviceData(nullptr));
    float* d_out = static_cast<float*>(spectra->deviceDataForWrite());
    ppl_check(ppl_ppf_run_c32(_plan, d_in, nBlk, d_out, nullptr), "PPFChanneliser");
}

}  // namespace ampp
}  // namespace pelican
