// BandPassRecorder.h -- measures a bandpass from a Stokes-I stream
// (src/lib/BandPassRecorder.h:28-61, src/lib/src/BandPassRecorder.cpp:26-187).
//
// Configuration (the reference's tags and defaults):
//   <requiredSamples value="20000"/>      spectra to integrate
//   <fitParmaters coefficients="3"/>      polynomial coefficients (degree + 1)
//   <Band startFrequency=".." endFrequency=".."/>   required
//
// run(stokes, bp) adds the blob's Stokes-I spectra to a per-bin float sum
// kept on the device (ppl_bandpass_accumulate, the reference's sequential
// sums bit for bit) and returns false until requiredSamples spectra are in.
// Then it fits the averaged spectrum on the host (fitSpectrum) and fills
// bp: the polynomial over the map start..end of nSubbands*nChannels bins, and
// the RMS of the cleaned spectrum.  The next run starts a new integration.
#pragma once

#include <vector>

#include "../pelican/DataBlob.h"
#include "BandPass.h"
#include "DataSets.h"

namespace pelican {
namespace ampp {

class BandPassRecorder : public AbstractModule {
  public:
    explicit BandPassRecorder(const ConfigNode& config);
    ~BandPassRecorder() override;
    BandPassRecorder(const BandPassRecorder&) = delete;
    BandPassRecorder& operator=(const BandPassRecorder&) = delete;

    bool run(SpectrumDataSetStokes* stokesI, BandPass* bp);

    // The host stage of run, on an averaged spectrum (modified in place:
    // outliers are replaced by the fit): fits, trims values more than twice
    // the residual RMS above the fit, refits until none is left, and sets
    // bp's map, polynomial and RMS.  Returns the number of fits made.
    int fitSpectrum(std::vector<float>& spectrum, const BinMap& map, BandPass* bp) const;

    // The map run uses for nBins bins: start at startFrequency, width
    // (endFrequency - startFrequency) / nBins.
    BinMap binMap(unsigned nBins) const;
    unsigned long samplesIntegrated() const { return _totalSamples; }
    // fits the last completed run made
    int fitsMade() const { return _lastFits; }

  private:
    void _reset(unsigned nBins);
    // least-squares polynomial through (freq[i], y[i])
    std::vector<float> _polyFit(const std::vector<float>& freq, const std::vector<float>& y) const;
    static float _theFit(const std::vector<float>& fit, float v);

    unsigned long _requiredSamples = 20000, _totalSamples = 0;
    int _polyDegree = 3;
    float _startFrequency = 0.0f, _endFrequency = 0.0f;
    unsigned _nBins = 0;
    int _lastFits = 0;
    float* _dSum = nullptr;
};

}  // namespace ampp
}  // namespace pelican
