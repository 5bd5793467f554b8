// DataSets.h -- the hot-path data blobs of pelican-lofar with their layouts:
// TimeSeriesDataSet (src/lib/TimeSeriesDataSet.h:145-160, [sub][pol][block][t]),
// SpectrumDataSet (src/lib/SpectrumDataSet.h:229-242, [block][sub][pol][ch]),
// WeightedSpectrumDataSet (src/lib/WeightedSpectrumDataSet.h:26-47),
// BlobStatistics, BinMap (src/lib/src/BinMap.cpp), DedispersionSpectra
// (src/lib/DedispersionSpectra.h:30-85), DedispersionEvent and
// DedispersionDataAnalysis (src/lib/src/DedispersionDataAnalysis.cpp).
// Each array blob carries a DeviceMirror.
#pragma once

#include <complex>
#include <cstdint>
#include <list>
#include <numeric>
#include <vector>

#include "../pelican/DataBlob.h"
#include "DeviceMirror.h"

namespace pelican {
namespace ampp {

template <typename T>
class TimeSeriesDataSet : public DataBlob {
  public:
    explicit TimeSeriesDataSet(const std::string& type = "TimeSeriesDataSet") : DataBlob(type) {}
    void resize(unsigned nTimeBlocks, unsigned nSubbands, unsigned nPols, unsigned nTimes)
    {
        if (_nSubbands != nSubbands || _nPolarisations != nPols || _nTimeBlocks != nTimeBlocks ||
            _nTimesPerBlock != nTimes) {
            _nSubbands = nSubbands;
            _nPolarisations = nPols;
            _nTimeBlocks = nTimeBlocks;
            _nTimesPerBlock = nTimes;
            _data.resize((size_t)nSubbands * nPols * nTimeBlocks * nTimes);
            _dev.invalidate();
        }
    }
    unsigned nSubbands() const { return _nSubbands; }
    unsigned nPolarisations() const { return _nPolarisations; }
    unsigned nTimeBlocks() const { return _nTimeBlocks; }
    unsigned nTimesPerBlock() const { return _nTimesPerBlock; }
    size_t size() const { return _data.size(); }
    // subband (outer) -> pol -> block -> time (inner)
    static long index(unsigned subband, unsigned numTimesPerBlock, unsigned polarisation,
                      unsigned numPolarisations, unsigned block, unsigned numTimeBlocks)
    {
        return (long)numTimesPerBlock * ((long)numTimeBlocks * (subband * numPolarisations + polarisation) + block);
    }
    T* data()
    {
        _dev.hostAccess(_data.data(), bytes(), true);
        return _data.data();
    }
    const T* constData() const
    {
        _dev.hostAccess(const_cast<T*>(_data.data()), bytes(), false);
        return _data.data();
    }
    T* timeSeriesData(unsigned b, unsigned s, unsigned p)
    {
        return data() + index(s, _nTimesPerBlock, p, _nPolarisations, b, _nTimeBlocks);
    }
    double getBlockRate() const { return _blockRate; }
    void setBlockRate(double r) { _blockRate = r; }
    double getLofarTimestamp() const { return _lofarTimestamp; }
    void setLofarTimestamp(double t) { _lofarTimestamp = t; }
    size_t bytes() const { return _data.size() * sizeof(T); }
    const void* deviceData(void* stream) const
    {
        return const_cast<DeviceMirror&>(_dev).deviceRead(_data.data(), bytes(), stream);
    }

  protected:
    std::vector<T> _data;
    mutable DeviceMirror _dev;
    unsigned _nSubbands = 0, _nPolarisations = 0, _nTimeBlocks = 0, _nTimesPerBlock = 0;
    double _blockRate = 0.0, _lofarTimestamp = 0.0;
};

class TimeSeriesDataSetC32 : public TimeSeriesDataSet<std::complex<float>> {
  public:
    TimeSeriesDataSetC32() : TimeSeriesDataSet<std::complex<float>>("TimeSeriesDataSetC32") {}
};

class SpectrumDataSetBase : public DataBlob {
  public:
    explicit SpectrumDataSetBase(const std::string& type) : DataBlob(type) {}
    unsigned nTimeBlocks() const { return _nTimeBlocks; }
    unsigned nSubbands() const { return _nSubbands; }
    unsigned nPolarisations() const { return _nPolarisations; }
    unsigned nChannels() const { return _nChannels; }
    double getTime(unsigned sampleNumber) const { return _startTimestamp + sampleNumber * _blockRate; }
    double getBlockRate() const { return _blockRate; }
    void setBlockRate(double r) { _blockRate = r; }
    double getLofarTimestamp() const { return _startTimestamp; }
    void setLofarTimestamp(double t) { _startTimestamp = t; }
    static long index(unsigned subband, unsigned numSubbands, unsigned polarisation, unsigned numPolarisations,
                      unsigned block, unsigned numChannels)
    {
        return (long)numChannels * ((long)numPolarisations * ((long)numSubbands * block + subband) + polarisation);
    }

  protected:
    unsigned _nSubbands = 0, _nPolarisations = 0, _nTimeBlocks = 0, _nChannels = 0;
    double _blockRate = 0.0, _startTimestamp = 0.0;
};

template <typename T>
class SpectrumDataSet : public SpectrumDataSetBase {
  public:
    explicit SpectrumDataSet(const std::string& type = "SpectrumDataSet") : SpectrumDataSetBase(type) {}
    void resize(unsigned nTimeBlocks, unsigned nSubbands, unsigned nPolarisations, unsigned nChannels)
    {
        const size_t n = (size_t)nSubbands * nPolarisations * nTimeBlocks * nChannels;
        _nSubbands = nSubbands;
        _nPolarisations = nPolarisations;
        _nTimeBlocks = nTimeBlocks;
        _nChannels = nChannels;
        if (n != _data.size()) {
            _data.resize(n);
            _dev.invalidate();
        }
    }
    void resize(const SpectrumDataSet<T>& o)
    {
        resize(o.nTimeBlocks(), o.nSubbands(), o.nPolarisations(), o.nChannels());
    }
    void init(const T& v)
    {
        T* d = data();
        std::fill(d, d + _data.size(), v);
    }
    int size() const { return (int)_data.size(); }
    T* data()
    {
        _dev.hostAccess(_data.data(), bytes(), true);
        return _data.data();
    }
    const T* data() const
    {
        _dev.hostAccess(const_cast<T*>(_data.data()), bytes(), false);
        return _data.data();
    }
    T* spectrumData(unsigned b, unsigned s, unsigned p)
    {
        return data() + index(s, _nSubbands, p, _nPolarisations, b, _nChannels);
    }
    const T* spectrumData(unsigned b, unsigned s, unsigned p) const
    {
        return data() + index(s, _nSubbands, p, _nPolarisations, b, _nChannels);
    }
    typename std::vector<T>::const_iterator begin() const
    {
        data();
        return _data.begin();
    }
    typename std::vector<T>::const_iterator end() const { return _data.end(); }
    size_t bytes() const { return _data.size() * sizeof(T); }
    // device side (module use)
    const void* deviceData(void* stream) const
    {
        return const_cast<DeviceMirror&>(_dev).deviceRead(_data.data(), bytes(), stream);
    }
    void* deviceDataInPlace(void* stream) { return _dev.deviceReadWrite(_data.data(), bytes(), stream); }
    void* deviceDataForWrite() { return _dev.deviceWrite(bytes()); }

  protected:
    std::vector<T> _data;
    mutable DeviceMirror _dev;
};

class SpectrumDataSetC32 : public SpectrumDataSet<std::complex<float>> {
  public:
    SpectrumDataSetC32() : SpectrumDataSet<std::complex<float>>("SpectrumDataSetC32") {}
};

class SpectrumDataSetStokes : public SpectrumDataSet<float> {
  public:
    SpectrumDataSetStokes() : SpectrumDataSet<float>("SpectrumDataSetStokes") {}
    void setRawData(const SpectrumDataSetC32* raw) { _raw = raw; }
    const SpectrumDataSetC32* getRawData() const { return _raw; }

  private:
    const SpectrumDataSetC32* _raw = nullptr;
};

class BlobStatistics {
  public:
    void reset() { _rms = _mean = _median = 0.0f; }
    float rms() const { return _rms; }
    float mean() const { return _mean; }
    float median() const { return _median; }
    void setRMS(float v) { _rms = v; }
    void setMean(float v) { _mean = v; }
    void setMedian(float v) { _median = v; }

  private:
    float _rms = 0.0f, _mean = 0.0f, _median = 0.0f;
};

// WeightedSpectrumDataSet (src/lib/src/WeightedSpectrumDataSet.cpp:13-78):
// wraps a Stokes blob plus a same-shape weight blob initialised to 1.
class WeightedSpectrumDataSet : public DataBlob {
  public:
    explicit WeightedSpectrumDataSet(SpectrumDataSet<float>* data = nullptr) : DataBlob("WeightedSpectrumDataSet")
    {
        if (data) reset(data);
    }
    void reset(SpectrumDataSet<float>* data)
    {
        _dataSet = data;
        _weights.resize(*data);
        _weights.init(1.0f);
        _stats.reset();
    }
    SpectrumDataSet<float>* dataSet() const { return _dataSet; }
    SpectrumDataSet<float>* weights() { return &_weights; }
    float rms() const { return _stats.rms(); }
    float median() const { return _stats.median(); }
    float mean() const
    {
        const float m = _stats.mean();
        if (m) return m;
        return std::accumulate(_dataSet->begin(), _dataSet->end(), 0.0) /
               std::accumulate(_weights.begin(), _weights.end(), 0.0);
    }
    void setRMS(float v) { _stats.setRMS(v); }
    void setMedian(float v) { _stats.setMedian(v); }
    void setMean(float v) { _stats.setMean(v); }
    const BlobStatistics& stats() const { return _stats; }

  private:
    SpectrumDataSet<float>* _dataSet = nullptr;
    SpectrumDataSet<float> _weights;
    BlobStatistics _stats;
};

// BinMap (src/lib/src/BinMap.cpp:49-70), the parts DedispersionSpectra uses.
class BinMap {
  public:
    BinMap() = default;
    explicit BinMap(unsigned n) : _nBins(n) {}
    void reset(unsigned n)
    {
        _nBins = n;
        _lower = 0.0;
        _width = 1.0;
    }
    void setStart(double s) { _lower = s; }
    void setBinWidth(double w)
    {
        _width = w;
        _halfwidth = w / 2.0;
    }
    void setEnd(double end) { setBinWidth((end - _lower) / _nBins); }
    int binIndex(double value) const { return (int)(0.5 + ((value - _lower) / _width)); }
    double width() const { return _width; }
    unsigned numberBins() const { return _nBins; }
    double startValue() const { return _lower; }
    double endValue() const { return _lower + _width * _nBins; }
    double lastBinValue() const { return _lower + _width * (_nBins - 1); }
    double binAssignmentNumber(int index) const { return _lower + (_width * index); }
    double binStart(unsigned i) const { return binAssignmentNumber(i) - _halfwidth; }
    double binEnd(unsigned i) const { return binAssignmentNumber(i) + _halfwidth; }

  private:
    unsigned _nBins = 0;
    double _lower = 0.0, _width = 1.0, _halfwidth = 0.5;
};

// DedispersionSpectra (src/lib/src/DedispersionSpectra.cpp:25-66): [dm][t].
class DedispersionSpectra : public DataBlob {
  public:
    DedispersionSpectra() : DataBlob("DedispersionSpectra") {}
    void resize(unsigned timebins, unsigned dedispersionBins, float dedispersionBinStart, float dedispersionBinWidth)
    {
        _timeBins = timebins;
        _dmBin.reset(dedispersionBins);
        _dmBin.setStart(dedispersionBinStart);
        _dmBin.setBinWidth(dedispersionBinWidth);
        const size_t n = (size_t)timebins * dedispersionBins;
        if (n != _data.size()) {
            _data.resize(n);
            _dev.invalidate();
        }
    }
    std::vector<float>& data()
    {
        _dev.hostAccess(_data.data(), _data.size() * sizeof(float), true);
        return _data;
    }
    float dmAmplitude(unsigned timeSlice, float dm) const { return dmAmplitude(timeSlice, dmIndex(dm)); }
    float dmAmplitude(unsigned timeSlice, int dm) const
    {
        _dev.hostAccess(const_cast<float*>(_data.data()), _data.size() * sizeof(float), false);
        return _data[timeSlice + (size_t)_timeBins * dm];
    }
    int dmIndex(float dm) const { return _dmBin.binIndex(dm); }
    float dm(unsigned i) const { return (float)_dmBin.binAssignmentNumber(i); }
    float dmMax() const { return (float)_dmBin.lastBinValue(); }
    int dmBins() const { return (int)_dmBin.numberBins(); }
    int timeSamples() const { return (int)_timeBins; }
    const std::vector<SpectrumDataSetStokes*>& inputDataBlobs() const { return _inputBlobs; }
    void setInputDataBlobs(const std::vector<SpectrumDataSetStokes*>& b) { _inputBlobs = b; }
    void setFirstSample(unsigned n) { _firstSampleNumber = n; }
    double getTime(unsigned sampleNumber) const
    {
        return _inputBlobs.empty() ? 0.0 : _inputBlobs[0]->getTime(sampleNumber + _firstSampleNumber);
    }
    float getRMS() const { return _rms; }
    void setRMS(float r) { _rms = r; }
    void setLost(unsigned l) { _lost = l; }
    unsigned getLost() const { return _lost; }
    // device side
    void* deviceDataForWrite() { return _dev.deviceWrite(_data.size() * sizeof(float)); }
    const void* deviceData(void* stream) const
    {
        return const_cast<DeviceMirror&>(_dev).deviceRead(_data.data(), _data.size() * sizeof(float), stream);
    }

  private:
    BinMap _dmBin;
    unsigned _timeBins = 0, _firstSampleNumber = 0, _lost = 0;
    float _rms = 0.0f;
    std::vector<float> _data;
    mutable DeviceMirror _dev;
    std::vector<SpectrumDataSetStokes*> _inputBlobs;
};

// DedispersionEvent (src/lib/src/DedispersionEvent.cpp:13-55).
class DedispersionEvent {
  public:
    DedispersionEvent(int dmIndex, unsigned timeIndex, const DedispersionSpectra* d, float mfBinFactor,
                      float mfBinValue)
        : _dm(dmIndex), _time(timeIndex), _data(d), _mfBinFactor(mfBinFactor), _mfBinValue(mfBinValue) {}
    double getTime() const { return _data->getTime(_time); }
    unsigned timeBin() const { return _time; }
    int dmIndex() const { return _dm; }
    float mfBinning() const { return _mfBinFactor; }
    float mfValue() const { return _mfBinValue; }
    float dm() const { return _data->dm(_dm); }
    float amplitude() const { return _data->dmAmplitude(_time, _dm); }

  private:
    int _dm;
    unsigned _time;
    const DedispersionSpectra* _data;
    float _mfBinFactor, _mfBinValue;
};

// DedispersionDataAnalysis (src/lib/src/DedispersionDataAnalysis.cpp:24-39).
class DedispersionDataAnalysis : public DataBlob {
  public:
    DedispersionDataAnalysis() : DataBlob("DedispersionDataAnalysis") {}
    void reset(const DedispersionSpectra* data)
    {
        _data = data;
        _events.clear();
    }
    int eventsFound() const { return (int)_events.size(); }
    const std::vector<DedispersionEvent>& events() const { return _events; }
    void addEvent(unsigned dmIndex, unsigned timeIndex, float mfBinFactor, float mfBinValue)
    {
        _events.emplace_back((int)dmIndex, timeIndex, _data, mfBinFactor, mfBinValue);
    }
    const DedispersionSpectra* data() const { return _data; }
    float getRMS() const { return _rms; }
    void setRMS(float r) { _rms = r; }

  private:
    const DedispersionSpectra* _data = nullptr;
    std::vector<DedispersionEvent> _events;
    float _rms = 0.0f;
};

}  // namespace ampp
}  // namespace pelican
