#include "StokesGenerator.h"

#include <iostream>
#include <stdexcept>

#include "../../../include/ppl_amd.h"

namespace pelican {
namespace ampp {

StokesGenerator::StokesGenerator(const ConfigNode& config) : AbstractModule(config)
{
    _numberOfStokes = config.getOption("numberOfStokes", "value", "4").toUInt();
    // the reference prints a message and exits the process (:21-25); a
    // library throws instead
    if (_numberOfStokes != 1 && _numberOfStokes != 4)
        throw std::runtime_error(
            "You can either generate 1 or 4 Stokes parameters. Change numberOfStokes in xml file.");
}

void StokesGenerator::run(const SpectrumDataSetC32* in, SpectrumDataSetStokes* stokes)
{
    if (in->nPolarisations() < 2) throw std::runtime_error("StokesGenerator: two polarisations required");
    const unsigned nT = in->nTimeBlocks(), nSub = in->nSubbands(), nCh = in->nChannels();
    if (in->nPolarisations() != 2)
        throw std::runtime_error("StokesGenerator: spectra with exactly two polarisations are supported");
    stokes->setLofarTimestamp(in->getLofarTimestamp());
    stokes->setBlockRate(in->getBlockRate());
    stokes->resize(nT, nSub, _numberOfStokes, nCh);
    stokes->setRawData(in);
    const float* d_in = static_cast<const float*>(in->deviceData(nullptr));
    float* d_out = static_cast<float*>(stokes->deviceDataForWrite());
    ppl_check(ppl_stokes(d_in, nT, nSub, nCh, _numberOfStokes, d_out, nullptr), "StokesGenerator");
}

StokesIntegrator::StokesIntegrator(const ConfigNode& config) : AbstractModule(config)
{
    _windowSize = config.getOption("integrateTimeBins", "value", "1").toUInt();
    _binChannels = config.getOption("integrateFrequencyChannels", "value", "1").toUInt();
}

void StokesIntegrator::run(const SpectrumDataSetStokes* in, SpectrumDataSetStokes* out)
{
    const unsigned nSamples = in->nTimeBlocks(), nSub = in->nSubbands(), nCh = in->nChannels();
    const unsigned nPols = in->nPolarisations();
    if (_windowSize > nSamples) {
        // the reference shrinks the window for good (:50-53)
        std::cout << "Warning the window size has been reduced from" << _windowSize
                  << "to the total sample count" << nSamples << std::endl;
        _windowSize = nSamples;
    }
    const unsigned newSamples = nSamples / _windowSize, newChannels = nCh / _binChannels;
    out->resize(newSamples, nSub, nPols, newChannels);
    const float* d_in = static_cast<const float*>(in->deviceData(nullptr));
    float* d_out = static_cast<float*>(out->deviceDataForWrite());
    if (newSamples > 0)
        ppl_check(ppl_integrate(d_in, newSamples * _windowSize, nSub, nPols, nCh, _windowSize, _binChannels, d_out,
                                nullptr),
                  "StokesIntegrator");
    out->setLofarTimestamp(in->getLofarTimestamp());
    out->setBlockRate(in->getBlockRate() * _windowSize);
}

}  // namespace ampp
}  // namespace pelican
