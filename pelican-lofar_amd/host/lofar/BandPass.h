// BandPass.h -- the bandpass model blob (src/lib/BandPass.h,
// src/lib/src/BandPass.cpp:17-45,158-166), the part a recorder fills and an
// output stream writes: a polynomial in frequency over a primary BinMap,
// the model evaluated at every bin of that map, its mean and median, and an
// RMS.  (The clipper reads .bp files into its own device model instead.)
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "../pelican/DataBlob.h"
#include "DataSets.h"

namespace pelican {
namespace ampp {

class BandPass : public DataBlob {
  public:
    BandPass() : DataBlob("BandPass"), _params(1, 0.0f) {}

    // setData (BandPass.cpp:26-42): the map and polynomial; evaluates the
    // model at every bin and takes its mean (float, in bin order) and the
    // element n/2 of the sorted values as the median.
    void setData(const BinMap& map, const std::vector<float>& params)
    {
        _primaryMap = map;
        _params = params;
        const unsigned n = map.numberBins();
        _data.assign(n, 0.0f);
        for (unsigned i = 0; i < n; ++i) _data[i] = evaluate((float)map.binAssignmentNumber(i));
        _mean = 0.0f;
        for (unsigned i = 0; i < n; ++i) _mean += _data[i];
        _mean /= n;
        std::vector<float> copy(_data);
        std::nth_element(copy.begin(), copy.begin() + copy.size() / 2, copy.end());
        _median = copy.empty() ? 0.0f : copy[copy.size() / 2];
    }
    void setRMS(float rms) { _rms = rms; }

    const BinMap& primaryMap() const { return _primaryMap; }
    const std::vector<float>& params() const { return _params; }
    float primaryRms() const { return _rms; }
    float primaryMedian() const { return _median; }
    float mean() const { return _mean; }
    // the model at every bin of the primary map
    const std::vector<float>& data() const { return _data; }

    // _evaluate (BandPass.cpp:158-166): tot += p[i] * pow(v, i), the power
    // and product in double, the running total in float.
    float evaluate(float v) const
    {
        float tot = 0.0f;
        for (size_t i = 0; i < _params.size(); ++i) tot += _params[i] * std::pow((double)v, (double)i);
        return tot;
    }

  private:
    BinMap _primaryMap;
    std::vector<float> _params;
    std::vector<float> _data;
    float _mean = 0.0f, _median = 0.0f, _rms = 0.0f;
};

}  // namespace ampp
}  // namespace pelican
