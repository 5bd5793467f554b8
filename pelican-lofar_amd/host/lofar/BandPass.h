// BandPass.h -- the bandpass model blob (src/lib/BandPass.h:37-96,
// src/lib/src/BandPass.cpp:17-166): a polynomial in frequency tied to a
// primary BinMap, its values on any map it is rebinned to (scaled by the
// width ratio), and per map a mean, median and RMS.  The clipper builds one
// from its .bp file as BandPassAdapter does (fromFileValues) and reports it
// through RFI_Clipper::bandPass(); BandPassRecorder fills one and
// BandPassOutput writes it.  (The clipper's running model lives on the
// device; this object holds the file's model.)
#pragma once

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>
#include <tuple>
#include <utility>
#include <vector>

#include "../pelican/DataBlob.h"
#include "DataSets.h"

namespace pelican {
namespace ampp {

class BandPass : public DataBlob {
  public:
    BandPass() : DataBlob("BandPass"), _params(1, 0.0f) {}

    // setData (BandPass.cpp:26-42): the primary map and polynomial; the
    // model is built on the map (if not built before), and the primary
    // mean (float, bin order) and median (element n/2 of the sorted values)
    // are taken from it.
    void setData(const BinMap& map, const std::vector<float>& params)
    {
        _primaryMap = map;
        _primaryId = key(map);
        _params = params;
        reBin(map);
        std::vector<float> copy;
        float mean = 0.0f;
        const std::vector<float>& d = _data[_primaryId];
        for (unsigned i = 0; i < map.numberBins(); ++i) {
            mean += d[i];
            copy.push_back(d[i]);
        }
        mean /= map.numberBins();
        _mean[_primaryId] = mean;
        if (!copy.empty()) {
            std::nth_element(copy.begin(), copy.begin() + copy.size() / 2, copy.end());
            _median[_primaryId] = copy[copy.size() / 2];
        }
    }
    // setRMS (:44-46): the current map's RMS.
    void setRMS(float rms) { _rms[_currentId] = rms; }
    // setMedian (:48-61): moves the polynomial's constant term so that the
    // current map's median becomes `median`.
    void setMedian(float median)
    {
        const float delta = median - _median[_currentId];
        if (std::fabs(delta) > 0.0f) {
            const float scale = (float)(_currentMap.width() / _primaryMap.width());
            _median[_currentId] = median;
            _median[_primaryId] = median / scale;
            _params[0] += delta / scale;
            _mean[_currentId] += delta;
            _mean[_primaryId] += delta / scale;
            _data.clear();
            buildData(_currentMap, scale);
        }
    }
    // setMean (:63-74): the same for the mean.
    void setMean(float mean)
    {
        const float delta = mean - _mean[_currentId];
        if (std::fabs(delta) > 0.0f) {
            const float scale = (float)(_currentMap.width() / _primaryMap.width());
            _params[0] += delta / scale;
            _mean[_currentId] = mean;
            _mean[_primaryId] = mean / scale;
            _data.clear();
            buildData(_currentMap, scale);
        }
    }
    void resetMap() { reBin(_primaryMap); }
    // reBin (:81-99): makes `map` current; a map seen for the first time
    // gets the primary values scaled by the width ratio (RMS by its inverse
    // square root).
    void reBin(const BinMap& map)
    {
        _currentMap = map;
        const Key id = key(map);
        _currentId = id;
        if (!_data.count(id)) {
            const double scale = map.width() / _primaryMap.width();
            _rms[id] = (float)(_rms[_primaryId] * std::sqrt(1.0 / scale));
            _median[id] = (float)(_median[_primaryId] * scale);
            _mean[id] = (float)(_mean[_primaryId] * scale);
            buildData(_currentMap, (float)scale);
        }
    }

    const std::vector<float>& params() const { return _params; }
    const BinMap& primaryMap() const { return _primaryMap; }
    float startFrequency() const { return (float)_currentMap.startValue(); }
    float endFrequency() const { return (float)_currentMap.endValue(); }
    float intensity(float frequency) const { return intensityOfBin((unsigned)_currentMap.binIndex(frequency)); }
    float intensityOfBin(unsigned index) const
    {
        const std::vector<float>& d = _data.at(_currentId);
        if (index >= d.size()) throw std::out_of_range("BandPass: bin outside the current map");
        return d[index];
    }
    const std::vector<float>& currentSet() const { return _data.at(_currentId); }
    float median() const { return at(_median, _currentId); }
    float mean() const { return at(_mean, _currentId); }
    float rms() const { return at(_rms, _currentId); }
    float primaryMedian() const { return at(_median, _primaryId); }
    float primaryRms() const { return at(_rms, _primaryId); }
    // The primary map's model values (as setData built them).
    const std::vector<float>& data() const { return _data.at(_primaryId); }

    // killChannel / killBand (:141-157): zero the model over a frequency
    // range on every map built so far and every later one; filterBin
    // (:146-149): the bin's model is (near) zero.
    void killChannel(unsigned index) { killBand((float)_currentMap.binStart(index), (float)_currentMap.binEnd(index)); }
    void killBand(float start, float end)
    {
        _killed.emplace_back(start, end);
        for (auto& kv : _data) zeroChannels(kv.first, kv.second);
    }
    bool filterBin(unsigned index) const { return intensityOfBin(index) < 0.0000001f; }

    // _evaluate (:158-166): tot += p[i] * pow(v, i), the power and product in
    // double, the running total in float.
    float evaluate(float v) const
    {
        float tot = 0.0f;
        for (size_t i = 0; i < _params.size(); ++i) tot += _params[i] * std::pow((double)v, (double)i);
        return tot;
    }

    // BandPassAdapter::deserialise (src/lib/src/BandPassAdapter.cpp:76-100)
    // on the numbers of a .bp file: nChan, start, end, width, rms, median,
    // coefficients.  The adapter drops the first five numbers only, so the
    // median becomes the constant coefficient and the file's coefficients
    // move up one power; then setMedian(median) and setRMS(rms).
    // Kill lines (:58-71,101-106): the unique indices in ascending order
    // (a QMap), each checked against the file's map (the line's first number
    // is also its frequency) and killed on it.
    static BandPass fromFileValues(const std::vector<double>& v, const std::vector<unsigned>& kills = {})
    {
        if (v.size() < 6) throw std::runtime_error("BandPassAdapter: data format wrong");
        BandPass bp;
        BinMap map((unsigned)(float)v[0]);
        map.setStart((float)v[1]);
        map.setBinWidth((float)v[3]);
        std::vector<float> params;
        for (size_t i = 5; i < v.size(); ++i) params.push_back((float)v[i]);
        bp.setData(map, params);
        bp.setMedian((float)v[5]);
        bp.setRMS((float)v[4]);
        std::vector<unsigned> k(kills);
        std::sort(k.begin(), k.end());
        k.erase(std::unique(k.begin(), k.end()), k.end());
        for (unsigned index : k) {
            const float freq = (float)(double)index;
            if (map.binIndex(freq) != (int)index)
                throw std::runtime_error("BandPassAdapter:  Frequency kill mismatch: index " + std::to_string(index) +
                                         " does not match frequency " + std::to_string(freq));
            bp.killChannel(index);
        }
        return bp;
    }

    // The clipper's model after a run (RFI_Clipper.cpp:184-187,562-563):
    // current on the run map, with the constant coefficient, mean and RMS
    // that setMean / setRMS left; the values are rebuilt as setMean
    // rebuilds them (:63-74).
    void setRunState(const BinMap& runMap, float params0, float mean, float rms)
    {
        reBin(runMap);
        const float scale = (float)(_currentMap.width() / _primaryMap.width());
        if (params0 != _params[0] || mean != _mean[_currentId]) {
            _params[0] = params0;
            _mean[_currentId] = mean;
            _mean[_primaryId] = mean / scale;
            _data.clear();
            buildData(_currentMap, scale);
        }
        setRMS(rms);
    }

  private:
    using Key = std::tuple<unsigned, double, double>;  // BinMap::hash: (bins, start, width)
    static Key key(const BinMap& m) { return Key(m.numberBins(), m.startValue(), m.width()); }
    template <typename M>
    static float at(const M& m, const Key& k)
    {
        auto it = m.find(k);
        return it == m.end() ? 0.0f : it->second;
    }
    // _buildData (:101-109): scale * model at every bin centre, then the
    // killed ranges zeroed
    void buildData(const BinMap& map, float scale)
    {
        const Key id = key(map);
        std::vector<float>& d = _data[id];
        d.assign(map.numberBins(), 0.0f);
        for (unsigned i = 0; i < map.numberBins(); ++i) d[i] = scale * evaluate((float)map.binAssignmentNumber(i));
        zeroChannels(id, d);
    }
    // _zeroChannelsMap (:111-139), with the bin range clipped to the map
    void zeroChannels(const Key& id, std::vector<float>& d)
    {
        for (const auto& r : _killed) {
            BinMap map(std::get<0>(id));
            map.setStart(std::get<1>(id));
            map.setBinWidth(std::get<2>(id));
            int lo = map.binIndex(r.first), hi = map.binIndex(r.second);
            if (hi < lo) std::swap(lo, hi);
            lo = std::max(lo, 0);
            hi = std::min(hi, (int)d.size() - 1);
            for (int i = lo; i <= hi; ++i) d[(size_t)i] = 0.0f;
        }
    }

    BinMap _primaryMap, _currentMap;
    Key _primaryId{0, 0.0, 1.0}, _currentId{0, 0.0, 1.0};
    std::vector<float> _params;
    std::map<Key, std::vector<float>> _data;
    std::map<Key, float> _rms, _median, _mean;
    std::vector<std::pair<float, float>> _killed;
};

}  // namespace ampp
}  // namespace pelican
