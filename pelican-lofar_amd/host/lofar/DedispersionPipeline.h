// DedispersionPipeline.h -- the pulsar-search chain of the udpstream
// pipeline (src/pipelines/udpstream/src/DedispersionPipeline.cpp:62-216):
// PPF -> Stokes -> [integration] -> RFI clipper -> dedispersion (async) ->
// analysis, with a pool of `history` Stokes blobs that the dedispersion
// module hands back through its unlock callback.
#pragma once

#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <vector>

#include "../pelican/OutputStreamManager.h"
#include "DataSets.h"
#include "DedispersionModule.h"
#include "PPFChanneliser.h"
#include "RFI_Clipper.h"
#include "StokesGenerator.h"

namespace pelican {
namespace ampp {

// LockingPtrContainer: a pool whose next() blocks until an element is free.
template <class T>
class LockingPtrContainer {
  public:
    explicit LockingPtrContainer(std::vector<T*> items)
    {
        for (T* t : items) _free.push_back(t);
    }
    T* next()
    {
        std::unique_lock<std::mutex> l(_m);
        _cv.wait(l, [this] { return !_free.empty(); });
        T* t = _free.front();
        _free.pop_front();
        return t;
    }
    void unlock(T* t)
    {
        {
            std::lock_guard<std::mutex> l(_m);
            _free.push_back(t);
        }
        _cv.notify_one();
    }
    size_t available()
    {
        std::lock_guard<std::mutex> l(_m);
        return _free.size();
    }

  private:
    std::mutex _m;
    std::condition_variable _cv;
    std::deque<T*> _free;
};

class DedispersionPipeline {
  public:
    using EventCallback = std::function<void(const DedispersionDataAnalysis&)>;
    // `config` holds a DedispersionPipeline element (history/value, events
    // min/max) and one element per module, named after the module; a
    // StokesIntegrator element adds the integration stage.
    explicit DedispersionPipeline(const ConfigNode& config);
    ~DedispersionPipeline();
    // One chunk of voltages through the chain (DedispersionPipeline::run).
    void run(TimeSeriesDataSetC32* timeSeries);
    // Called for every analysed dedispersion buffer whose event count passes
    // the events min/max rule (the "DedispersionDataAnalysis" output).
    void setEventCallback(EventCallback cb) { _onEvents = std::move(cb); }
    // The output streams the pipeline publishes to (dataOutput,
    // DedispersionPipeline.cpp:175-206): "TriggerInput" for every analysed
    // buffer with events, then, under the events min/max rule,
    // "DedispersionDataAnalysis" and "SignalFoundSpectrum" (each Stokes blob
    // that fed the buffer).  Not owned; null publishes nothing.
    void setOutputStreams(OutputStreamManager* outputs) { _outputs = outputs; }
    // Waits until every dedispersion buffer submitted so far is analysed.
    void waitForCompletion() { _dedispersionModule->waitForJobCompletion(); }
    DedispersionModule& dedispersionModule() { return *_dedispersionModule; }
    RFI_Clipper& rfiClipper() { return *_rfiClipper; }
    int buffersAnalysed() const { return _analysed; }
    // Stokes blobs back in the pool (all of them once every buffer is done).
    size_t freeStokesBlobs() { return _stokesBuffer->available(); }

  private:
    void dedispersionAnalysis(DataBlob* blob);
    void updateBufferLock(const std::vector<DataBlob*>& freeData);

    unsigned _minEventsFound, _maxEventsFound;
    std::unique_ptr<PPFChanneliser> _ppfChanneliser;
    std::unique_ptr<StokesGenerator> _stokesGenerator;
    std::unique_ptr<StokesIntegrator> _stokesIntegrator;
    std::unique_ptr<RFI_Clipper> _rfiClipper;
    std::unique_ptr<DedispersionModule> _dedispersionModule;
    std::unique_ptr<DedispersionAnalyser> _dedispersionAnalyser;
    SpectrumDataSetC32 _spectra;
    SpectrumDataSetStokes _rawStokes;
    std::vector<std::unique_ptr<SpectrumDataSetStokes>> _stokesData;
    std::unique_ptr<LockingPtrContainer<SpectrumDataSetStokes>> _stokesBuffer;
    WeightedSpectrumDataSet _weightedIntStokes;
    EventCallback _onEvents;
    OutputStreamManager* _outputs = nullptr;
    int _analysed = 0;
};

}  // namespace ampp
}  // namespace pelican
