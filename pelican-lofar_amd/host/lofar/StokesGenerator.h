// StokesGenerator.h -- Stokes parameters from dual-polarisation spectra
// (src/lib/StokesGenerator.h:35-43, src/lib/src/StokesGenerator.cpp:15-108).
#pragma once

#include "../pelican/DataBlob.h"
#include "DataSets.h"

namespace pelican {
namespace ampp {

class StokesGenerator : public AbstractModule {
  public:
    // numberOfStokes/value: 1 (I) or 4 (I, Q, U, V), default 4.
    explicit StokesGenerator(const ConfigNode& config);
    void run(const SpectrumDataSetC32* channeliserOutput, SpectrumDataSetStokes* stokes);

  private:
    unsigned _numberOfStokes;
};

// StokesIntegrator (src/lib/src/StokesIntegrator.cpp:15-123): sums
// integrateTimeBins consecutive spectra and integrateFrequencyChannels
// adjacent channels.
class StokesIntegrator : public AbstractModule {
  public:
    explicit StokesIntegrator(const ConfigNode& config);
    void run(const SpectrumDataSetStokes* stokesGeneratorOutput, SpectrumDataSetStokes* intStokes);

  private:
    unsigned _windowSize, _binChannels;
};

PELICAN_DECLARE_MODULE(StokesGenerator)
PELICAN_DECLARE_MODULE(StokesIntegrator)

}  // namespace ampp
}  // namespace pelican
