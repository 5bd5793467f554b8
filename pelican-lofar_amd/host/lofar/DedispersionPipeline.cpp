#include "DedispersionPipeline.h"

namespace pelican {
namespace ampp {

DedispersionPipeline::DedispersionPipeline(const ConfigNode& config)
{
    const ConfigNode c = config.getNode("DedispersionPipeline");
    const unsigned history = c.getOption("history", "value", "10").toUInt();
    _minEventsFound = c.getOption("events", "min", "5").toUInt();
    _maxEventsFound = c.getOption("events", "max", "0").toUInt();
    _ppfChanneliser.reset(new PPFChanneliser(config.getNode("PPFChanneliser")));
    _stokesGenerator.reset(new StokesGenerator(config.getNode("StokesGenerator")));
    const ConfigNode integ = config.getNode("StokesIntegrator");
    if (!integ.tag().empty()) _stokesIntegrator.reset(new StokesIntegrator(integ));
    _rfiClipper.reset(new RFI_Clipper(config.getNode("RFI_Clipper")));
    _dedispersionModule.reset(new DedispersionModule(config.getNode("DedispersionModule")));
    _dedispersionAnalyser.reset(new DedispersionAnalyser(config.getNode("DedispersionAnalyser")));
    _dedispersionModule->connect([this](DataBlob* b) { dedispersionAnalysis(b); });
    _dedispersionModule->unlockCallback([this](const std::vector<DataBlob*>& d) { updateBufferLock(d); });
    std::vector<SpectrumDataSetStokes*> blobs;
    for (unsigned i = 0; i < history; ++i) {
        _stokesData.emplace_back(new SpectrumDataSetStokes());
        blobs.push_back(_stokesData.back().get());
    }
    _stokesBuffer.reset(new LockingPtrContainer<SpectrumDataSetStokes>(blobs));
}

DedispersionPipeline::~DedispersionPipeline()
{
    if (_dedispersionModule) _dedispersionModule->waitForJobCompletion();
}

void DedispersionPipeline::run(TimeSeriesDataSetC32* timeSeries)
{
    _ppfChanneliser->run(timeSeries, &_spectra);
    SpectrumDataSetStokes* stokes = _stokesBuffer->next();
    if (_stokesIntegrator) {
        _stokesGenerator->run(&_spectra, &_rawStokes);
        _stokesIntegrator->run(&_rawStokes, stokes);
    } else {
        _stokesGenerator->run(&_spectra, stokes);
    }
    _weightedIntStokes.reset(stokes);
    _rfiClipper->run(&_weightedIntStokes);
    _dedispersionModule->dedisperse(&_weightedIntStokes);
    // a blob that contributed no samples is never locked: back to the pool
    if (_dedispersionModule->lockNumber(stokes) == 0) _stokesBuffer->unlock(stokes);
}

void DedispersionPipeline::dedispersionAnalysis(DataBlob* blob)
{
    DedispersionDataAnalysis result;
    auto* data = static_cast<DedispersionSpectra*>(blob);
    const unsigned found = (unsigned)_dedispersionAnalyser->analyse(data, &result);
    ++_analysed;
    if (!found) return;
    if (_outputs) _outputs->send(&result, "TriggerInput");
    bool write;
    if (_minEventsFound >= _maxEventsFound) write = found >= _minEventsFound;
    else write = found >= _minEventsFound && found <= _maxEventsFound;
    if (!write) return;
    if (_onEvents) _onEvents(result);
    if (_outputs) {
        _outputs->send(&result, "DedispersionDataAnalysis");
        for (const SpectrumDataSetStokes* d : data->inputDataBlobs()) _outputs->send(d, "SignalFoundSpectrum");
    }
}

void DedispersionPipeline::updateBufferLock(const std::vector<DataBlob*>& freeData)
{
    for (DataBlob* b : freeData) _stokesBuffer->unlock(static_cast<SpectrumDataSetStokes*>(b));
}

}  // namespace ampp
}  // namespace pelican
