// DedispersionModule.h -- brute-force dedispersion module
// (src/lib/DedispersionModule.h:40-150, src/lib/src/DedispersionModule.cpp)
// over the device plan and buffer.  Stokes-I spectra are transposed into a
// device-resident [channel][sample] buffer as they arrive; each full buffer
// is dedispersed on the device into one of numberOfBuffers output blobs,
// its overlap (maxshift + remaining samples) carried into the next buffer,
// and the output is exported to the connected functors from the worker
// thread once the device has finished.  Input blobs stay locked until every
// output that holds their samples has been exported.
#pragma once

#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <vector>

#include "AsyncronousModule.h"
#include "DataSets.h"

struct ppl_dedisp_plan;
struct ppl_ddbuf;

namespace pelican {
namespace ampp {

class DedispersionModule : public AsyncronousModule {
  public:
    // Options: timeBinsPerBufferPow2/value (15), dedispersionSamples/value
    // (1984), dedispersionStepSize/value (0), dedispersionMinimum/value (0,
    // clamped at 0), frequencyChannel1/MHz (required), channelBandwidth/MHz
    // (1; must be negative: channels in descending frequency),
    // numberOfBuffers/value (2); this module's own inputPrecision/value
    // (f32 | f16 | u8) and quantiseScale/value (127/10, u8 steps per unit).
    explicit DedispersionModule(const ConfigNode& config);
    ~DedispersionModule() override;

    void dedisperse(DataBlob* incoming);
    void dedisperse(WeightedSpectrumDataSet* incoming);
    void waitForJobCompletion();
    void resize(const SpectrumDataSet<float>* streamData);
    unsigned numberOfSamples() const { return _numSamplesBuffer; }
    int maxshift() const { return _maxshift; }
    int remainingSamples() const { return _remainingSamples; }

  protected:
    void exportComplete(DataBlob* data) override;

  private:
    struct Segment {
        SpectrumDataSetStokes* blob;
        unsigned start, count;
    };
    void _launch();
    void _cleanBuffers();
    DedispersionSpectra* _nextOutput();
    static void _gpuDone(void* user, int status);

    unsigned _tdms, _numSamplesBuffer, _maxBuffers;
    float _dmStep, _dmLow;
    double _fch1, _foff, _tsamp = 0.0;
    int _dtype;
    float _invScale;
    int _maxshift = 0, _remainingSamples = 0, _nOut = 0;
    unsigned _sampleSize = 0, _nSub = 0, _nCh = 0;
    ppl_dedisp_plan* _plan = nullptr;
    ppl_ddbuf* _buffer = nullptr;
    void* _sums = nullptr;  // int32 sums of the u8 mode
    std::vector<std::unique_ptr<DedispersionSpectra>> _outputs;
    std::deque<DedispersionSpectra*> _free;
    std::mutex _freeMutex;
    std::condition_variable _freeCv;
    std::vector<Segment> _segments;  // samples of the current buffer, in order
};

// DedispersionAnalyser (src/lib/DedispersionAnalyser.h:20-44,
// src/lib/src/DedispersionAnalyser.cpp:16-203): dyadic boxcar threshold
// detection on the device.
class DedispersionAnalyser : public AbstractModule {
  public:
    // Options: detectionThreshold/in_sigma (6.0), power2ForBinning/value (6),
    // useStokesStats/0_or_1 (accepted, unused as in the reference).
    explicit DedispersionAnalyser(const ConfigNode& config);
    // Fills `result` with the events of `data` (the dummy event first) and
    // returns how many there are.
    int analyse(DedispersionSpectra* data, DedispersionDataAnalysis* result);

  private:
    float _detectionThreshold;
    unsigned _binPow2;
    std::vector<int32_t> _ints;
    std::vector<float> _vals;
};

PELICAN_DECLARE_MODULE(DedispersionModule)
PELICAN_DECLARE_MODULE(DedispersionAnalyser)

}  // namespace ampp
}  // namespace pelican
