// oracle/clipper.cpp -- CPU restatement of RFI_Clipper with its BandPass
// model.  TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Follows src/lib/src/RFI_Clipper.cpp:22-573, src/lib/src/BandPass.cpp:26-166,
// src/lib/src/BandPassAdapter.cpp:30-107 and src/lib/src/BinMap.cpp:49-70
// statement by statement, including the float/double promotions of each
// expression (C++11 semantics: std::pow(float,int) and sqrt(unsigned) are
// evaluated in double).  Compiled with -ffp-contract=off.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <deque>
#include <map>
#include <numeric>
#include <tuple>
#include <vector>

namespace {

// BinMap, src/lib/BinMap.h:24-78, src/lib/src/BinMap.cpp
struct BinMapR {
    unsigned nBins = 0;
    double lower = 0.0, width = 1.0, halfwidth = 0.5;
    void reset(unsigned n) { nBins = n; lower = 0.0; width = 1.0; }
    void setStart(double s) { lower = s; }
    void setBinWidth(double w) { width = w; halfwidth = w / 2.0; }
    void setEnd(double end) { setBinWidth((end - lower) / nBins); }  // BinMap.cpp:67-70
    double binAssignmentNumber(int i) const { return lower + (width * i); }
    double endValue() const { return lower + width * nBins; }
    int binIndex(double v) const { return (int)(0.5 + ((v - lower) / width)); }     // BinMap.cpp:49-52
    double binStart(unsigned i) const { return binAssignmentNumber(i) - halfwidth; }  // :72-75
    double binEnd(unsigned i) const { return binAssignmentNumber(i) + halfwidth; }    // :77-80
};

// boost::circular_buffer<float> semantics used by the clipper.
struct Ring {
    size_t cap = 0;
    std::deque<float> d;
    bool full() const { return d.size() == cap; }
    size_t size() const { return d.size(); }
    void push_back(float v) { if (cap == 0) return; if (d.size() == cap) d.pop_front(); d.push_back(v); }
    void pop_back() { d.pop_back(); }
    float front() const { return d.front(); }
    float back() const { return d.back(); }
    double accumulate0() const { double s = 0.0; for (float v : d) s = s + v; return s; }
    void clear() { d.clear(); }
};

struct BandPassR {
    std::map<std::tuple<unsigned, double, double>, int> ids;
    int idOf(const BinMapR& m) {
        auto key = std::make_tuple(m.nBins, m.lower, m.width);
        auto it = ids.find(key);
        if (it != ids.end()) return it->second;
        int id = (int)ids.size() + 1;
        ids[key] = id;
        return id;
    }
    BinMapR primary, current;
    int primaryId = 0, currentId = 0;
    std::vector<float> params{0.0f};
    std::map<int, std::vector<float>> dataSets;
    std::map<int, BinMapR> maps;  // the map of each data set
    std::map<int, float> rms_, median_, mean_;
    // Range<float>::subranges of the killed bands (src/lib/Range.h:40-92): each
    // (min, max) with min <= max
    std::vector<std::pair<float, float>> killed;
    bool killOutside = false;  // a kill reached outside a map (the reference writes out of bounds)

    // BandPass::_evaluate :159-166 -- tot is float, each term in double.
    float evaluate(float v) const {
        float tot = 0.0f;
        for (size_t i = 0; i < params.size(); ++i)
            tot += (double)params[i] * std::pow((double)v, (double)(int)i);
        return tot;
    }
    // BandPass::_buildData :101-109, then _zeroChannelsMap
    void buildData(const BinMapR& map, float scale) {
        int id = idOf(map);
        std::vector<float> d(map.nBins);
        for (unsigned i = 0; i < map.nBins; ++i) d[i] = scale * evaluate((float)map.binAssignmentNumber(i));
        dataSets[id] = d;
        maps[id] = map;
        zeroChannelsMap(id);
    }
    // BandPass::_zeroChannelsMap :111-125: every killed range's bins, from
    // binIndex(min) to binIndex(max) inclusive, are set to 0.0
    void zeroChannelsMap(int id) {
        const BinMapR& map = maps[id];
        std::vector<float>& d = dataSets[id];
        for (const auto& r : killed) {
            int lo = map.binIndex(r.first), hi = map.binIndex(r.second);
            if (hi < lo) std::swap(lo, hi);
            if ((int)d.size() < hi) d.resize(hi + 1);
            if (lo < 0 || hi >= (int)d.size()) { killOutside = true; continue; }
            do { d[lo] = 0.0f; } while (++lo <= hi);
        }
    }
    // BandPass::killBand :151-157 and killChannel :141-144 (on the current map)
    void killBand(float start, float end) {
        killed.emplace_back(std::min(start, end), std::max(start, end));
        for (auto& kv : dataSets) zeroChannelsMap(kv.first);
    }
    void killChannel(unsigned index) { killBand((float)current.binStart(index), (float)current.binEnd(index)); }
    // BandPass::reBin :81-99
    void reBin(const BinMapR& map) {
        current = map;
        int id = idOf(map);
        currentId = id;
        if (!dataSets.count(id)) {
            double scale = map.width / primary.width;
            rms_[id] = rms_[primaryId] * std::sqrt(1.0 / scale);
            median_[id] = median_[primaryId] * scale;
            mean_[id] = mean_[primaryId] * scale;
            buildData(current, (float)scale);
        }
    }
    // BandPass::setData :26-42
    void setData(const BinMapR& map, const std::vector<float>& p) {
        primary = map;
        primaryId = idOf(map);
        params = p;
        reBin(map);
        std::vector<float> copy;
        mean_[primaryId] = 0;
        for (unsigned i = 0; i < map.nBins; ++i) {
            mean_[primaryId] += dataSets[primaryId][i];
            copy.push_back(dataSets[primaryId][i]);
        }
        mean_[primaryId] /= map.nBins;
        std::nth_element(copy.begin(), copy.begin() + copy.size() / 2, copy.end());
        median_[primaryId] = copy[copy.size() / 2];
    }
    // BandPass::setMedian :48-61
    void setMedian(float median) {
        float delta = median - median_[currentId];
        if (std::fabs(delta) > 0.0f) {
            float scale = current.width / primary.width;
            median_[currentId] = median;
            median_[primaryId] = median / scale;
            params[0] += delta / scale;
            mean_[currentId] += delta;
            mean_[primaryId] += delta / scale;
            dataSets.clear();
            buildData(current, scale);
        }
    }
    // BandPass::setMean :63-74
    void setMean(float mean) {
        float delta = mean - mean_[currentId];
        if (std::fabs(delta) > 0.0f) {
            float scale = current.width / primary.width;
            params[0] += delta / scale;
            mean_[currentId] = mean;
            mean_[primaryId] = mean / scale;
            dataSets.clear();
            buildData(current, scale);
        }
    }
    void setRMS(float r) { rms_[currentId] = r; }
    float rms() { return rms_[currentId]; }
    const std::vector<float>& currentSet() { return dataSets[currentId]; }
};

}  // namespace

struct orc_clipper {
    BandPassR bp;
    BinMapR map;
    bool active = true;
    float startFrequency = 0, endFrequency = 0;
    float crFactor = 10.0f, srFactor = 4.0f;
    Ring meanBuffer, rmsBuffer;
    float lastGoodMean = 0, lastGoodRMS = 0;
    float rmsRunAve = 0, meanRunAve = 0, meanOverRMS = 0;
    float fractionBadChannels = 0;
    int badSpectra = 0, num = 0, maxHistory = 10;
    int zeroDMing = 0, useMeanOverRMS = 0, remainingZeros = 0;
    std::vector<float> lastGoodSpectrum;
    std::vector<uint8_t> lastGood;
    std::vector<int32_t> lastNclip;
};

extern "C" orc_clipper* orc_clipper_create(const double* v, int n, float crFactor, float srFactor,
                                           int maxHistory, int zeroDM, int bandMatching,
                                           int useMeanOverRMS)
{
    return orc_clipper_create_kill(v, n, nullptr, 0, crFactor, srFactor, maxHistory, zeroDM, bandMatching,
                                   useMeanOverRMS);
}

// The same with the indices of the file's K lines.  BandPassAdapter takes
// the kill line's first number for both the index and the frequency
// (BandPassAdapter.cpp:58-71), keeps them in a QMap (unique, ascending), and
// after setRMS checks binIndex(frequency) == index on the file's map before
// killChannel(index) (:101-106).  Returns nullptr when a check fails or a
// kill reaches outside the map (where the reference writes out of bounds).
extern "C" orc_clipper* orc_clipper_create_kill(const double* v, int n, const unsigned* kills, int nKill,
                                                float crFactor, float srFactor, int maxHistory, int zeroDM,
                                                int bandMatching, int useMeanOverRMS)
{
    if (n < 6) return nullptr;
    orc_clipper* c = new orc_clipper;
    // BandPassAdapter::deserialise :76-100.  Values pass through QVector<float>.
    std::vector<float> params;
    for (int i = 0; i < n; ++i) params.push_back((float)v[i]);
    unsigned nChan = (unsigned)params[0];
    float startFreq = params[1];
    float deltaF = params[3];
    float rms = params[4];
    float median = params[5];
    params.erase(params.begin(), params.begin() + 5);  // params.remove(0,5)
    BinMapR map;
    map.reset(nChan);
    map.setStart(startFreq);
    map.setBinWidth(deltaF);
    c->bp.setData(map, params);
    c->bp.setMedian(median);
    c->bp.setRMS(rms);
    std::map<unsigned, float> killChannels;
    for (int k = 0; k < nKill; ++k) killChannels[kills[k]] = (float)(double)kills[k];
    for (const auto& kv : killChannels) {
        const unsigned index = (unsigned)kv.second;  // foreach over the QMap's values
        if (map.binIndex(killChannels[index]) != (int)index) { delete c; return nullptr; }
        c->bp.killChannel(index);
    }
    if (c->bp.killOutside) { delete c; return nullptr; }
    // RFI_Clipper::RFI_Clipper :54-79
    c->crFactor = crFactor;
    c->srFactor = srFactor;
    c->maxHistory = maxHistory;
    c->meanBuffer.cap = maxHistory;
    c->rmsBuffer.cap = maxHistory;
    c->zeroDMing = zeroDM ? 1 : 0;
    c->useMeanOverRMS = useMeanOverRMS ? 1 : 0;
    if (bandMatching) {
        c->startFrequency = (float)c->bp.current.lower;
        c->endFrequency = (float)c->bp.current.endValue();
    }
    return c;
}

extern "C" void orc_clipper_destroy(orc_clipper* c) { delete c; }

extern "C" int orc_clipper_run(orc_clipper* C, float* I, float* W, unsigned nSamples, unsigned nSubbands,
                               unsigned nPolarisations, unsigned nChannels)
{
    unsigned nBins = nChannels * nSubbands;
    if (nBins % 8) return -1;
    auto index = [&](unsigned s, unsigned pol, unsigned t) -> long {
        return (long)nChannels * (nPolarisations * ((long)nSubbands * t + s) + pol);
    };
    float k = 4;
    float meanMinusMinimum = k / std::sqrt(2.0 * k);
    float spectrumRMS = 0.0f;
    float dataModel;
    C->lastGood.assign(nSamples, 0);
    C->lastNclip.assign(nSamples, 0);

    if (C->lastGoodSpectrum.size() != nBins) {
        C->lastGoodSpectrum.resize(nBins, 0.0f);
        C->remainingZeros = nBins;
    }
    C->map.reset(nBins);
    C->map.setStart(C->startFrequency);
    C->map.setEnd(C->endFrequency);
    C->bp.reBin(C->map);
    if (C->bp.killOutside) return -2;  // a killed range outside the run map

    for (unsigned t = 0; t < nSamples; ++t) {
        const std::vector<float> bandPass = C->bp.currentSet();
        float spectrumSum = 0.0;
        bool goodSpectrum = true;
        float spectrumSumSq = 0.0;
        std::vector<float> goodChannels(8, 0.0);
        unsigned channelsPerBand = nBins / 8;
        std::vector<float> miniData(8, 1e6), miniModel(8, 1e6), dataMinusModel(8), bandSigma(8);
        std::vector<float> bandMean(8, 0.0), bandMeanSquare(8, 0.0);

        if (!C->rmsBuffer.full()) {
            for (unsigned s = 0; s < nSubbands; ++s) {
                long idx = index(s, 0, t);
                for (unsigned c = 0; c < nChannels; ++c) {
                    int binLocal = s * nChannels + c;
                    unsigned band = (int)binLocal / channelsPerBand;
                    if (I[idx + c] < miniData[band]) miniData[band] = I[idx + c];
                    if (bandPass[binLocal] < miniModel[band]) miniModel[band] = bandPass[binLocal];
                    bandMean[band] += I[idx + c];
                    bandMeanSquare[band] += (I[idx + c] * I[idx + c]);
                }
            }
            for (unsigned b = 0; b < 8; ++b) {
                dataMinusModel[b] = miniData[b] - miniModel[b];
                bandSigma[b] = std::sqrt(bandMeanSquare[b] / channelsPerBand -
                                         std::pow((double)(bandMean[b] / channelsPerBand), 2.0));
            }
            spectrumRMS = *std::min_element(bandSigma.begin(), bandSigma.end());
            std::nth_element(dataMinusModel.begin(), dataMinusModel.begin() + 4, dataMinusModel.end());
            dataModel = dataMinusModel[4];
            C->rmsBuffer.push_back(spectrumRMS);
            C->rmsRunAve = C->rmsBuffer.accumulate0() / C->rmsBuffer.size();
            if (C->rmsBuffer.full()) C->meanOverRMS = C->meanRunAve / C->rmsRunAve;
            dataModel = dataModel + meanMinusMinimum * C->rmsRunAve;
        } else {
            C->rmsRunAve -= (C->rmsBuffer.front() / (float)C->rmsBuffer.size());
            C->rmsRunAve += (C->rmsBuffer.back() / (float)C->rmsBuffer.cap);
            if (C->useMeanOverRMS) spectrumRMS = std::fabs(C->meanRunAve / C->meanOverRMS);
            C->rmsBuffer.push_back(spectrumRMS);
            dataModel = 0.0;
        }

        float margin = C->crFactor * C->rmsRunAve;
        int nclip = 0;
        for (unsigned s = 0; s < nSubbands; ++s) {
            long idx = index(s, 0, t);
            for (unsigned c = 0; c < nChannels; ++c) {
                int binLocal = s * nChannels + c;
                if (I[idx + c] - dataModel - bandPass[binLocal] > margin) {
                    ++nclip;
                    for (unsigned pol = 0; pol < nPolarisations; ++pol) {
                        long ip = index(s, pol, t);
                        I[ip + c] = C->lastGoodSpectrum[binLocal];
                        if (W) W[ip + c] = 1.0;
                    }
                } else {
                    unsigned band = (int)binLocal / channelsPerBand;
                    ++goodChannels[band];
                    spectrumSum += I[idx + c];
                }
            }
        }
        unsigned totalGoodChannels = 0;
        {
            int acc = 0;
            for (float g : goodChannels) acc = acc + g;  // std::accumulate(..., 0)
            totalGoodChannels = acc;
        }
        C->fractionBadChannels += (float)(nBins - totalGoodChannels) / nBins;
        spectrumSum /= totalGoodChannels;

        unsigned badBands = 0;
        for (unsigned b = 0; b < 8; ++b) {
            if (goodChannels[b] < 0.8 * channelsPerBand) ++badBands;
            if (goodChannels[b] == 0) badBands += 4;
        }
        if (totalGoodChannels < 0.8 * nBins) badBands += 4;

        if (!C->meanBuffer.full()) {
            C->meanBuffer.push_back(spectrumSum);
            C->meanRunAve = C->meanBuffer.accumulate0() / C->meanBuffer.size();
        } else {
            C->meanRunAve -= C->meanBuffer.front() / (float)C->meanBuffer.size();
            C->meanRunAve += C->meanBuffer.back() / (float)C->meanBuffer.size();
            C->meanBuffer.push_back(spectrumSum);
        }

        float spectrumRMStolerance = C->srFactor * C->bp.rms() / std::sqrt((double)nBins);

        auto clipSample = [&]() {
            for (unsigned s = 0; s < nSubbands; ++s)
                for (unsigned pol = 0; pol < nPolarisations; ++pol) {
                    long ip = index(s, pol, t);
                    for (unsigned c = 0; c < nChannels; ++c) {
                        if (W) W[ip + c] = 1.0;
                        I[ip + c] = -0.2;
                    }
                }
        };

        if (C->meanBuffer.size() < 1000) {
            clipSample();
            goodSpectrum = false;
        } else if (spectrumSum - C->meanRunAve > spectrumRMStolerance || badBands >= 4) {
            clipSample();
            goodSpectrum = false;
            ++C->badSpectra;
            C->meanBuffer.pop_back();
            C->rmsBuffer.pop_back();
            C->meanBuffer.push_back(C->meanBuffer.back());
            C->rmsBuffer.push_back(C->rmsBuffer.back());
        } else {
            spectrumSum = 0.0;
            spectrumSumSq = 0.0;
            if (C->remainingZeros != 0) {
                for (unsigned s = 0; s < nSubbands; ++s) {
                    long idx = index(s, 0, t);
                    for (unsigned c = 0; c < nChannels; ++c) {
                        int binLocal = s * nChannels + c;
                        if (C->lastGoodSpectrum[binLocal] == 0.0 && I[idx + c] != 0.0) {
                            C->lastGoodSpectrum[binLocal] = I[idx + c];
                            --C->remainingZeros;
                            spectrumSum += C->lastGoodSpectrum[binLocal];
                            spectrumSumSq += C->lastGoodSpectrum[binLocal] * C->lastGoodSpectrum[binLocal];
                        }
                    }
                }
                C->lastGoodMean = spectrumSum / nBins;
                C->lastGoodRMS = std::sqrt(spectrumSumSq / nBins - std::pow((double)C->lastGoodMean, 2.0));
            }
        }

        spectrumSum = 0.0;
        spectrumSumSq = 0.0;
        for (unsigned s = 0; s < nSubbands; ++s) {
            long idx = index(s, 0, t);
            for (unsigned c = 0; c < nChannels; ++c) {
                int binLocal = s * nChannels + c;
                I[idx + c] -= (bandPass[binLocal] + dataModel);
                spectrumSum += I[idx + c];
                spectrumSumSq += I[idx + c] * I[idx + c];
            }
        }
        spectrumSum /= nBins;
        spectrumRMS = std::sqrt(spectrumSumSq / nBins - std::pow((double)spectrumSum, 2.0));
        if (spectrumRMS == 0.0) spectrumRMS = 1.0;

        if (goodSpectrum) {
            for (unsigned s = 0; s < nSubbands; ++s) {
                long idx = index(s, 0, t);
                for (unsigned c = 0; c < nChannels; ++c) {
                    if (C->zeroDMing == 1)
                        I[idx + c] -= C->zeroDMing * spectrumSum;
                    else
                        I[idx + c] -= C->meanRunAve;
                    I[idx + c] /= spectrumRMS;
                    if (I[idx + c] > C->crFactor) I[idx + c] = 0.0;
                }
            }
        }
        unsigned reportStatsEvery = 10 * C->maxHistory;
        if (C->num == 0) {
            float fractionBadSpectra = 100.0 * (float)C->badSpectra / (float)reportStatsEvery;
            if (fractionBadSpectra > 99.0) {
                C->rmsBuffer.clear();
                C->meanBuffer.clear();
            }
            C->badSpectra = 0;
            C->fractionBadChannels = 0.0;
        }
        C->bp.setMean(C->meanRunAve);
        C->bp.setRMS(C->rmsRunAve);
        ++C->num;
        C->num = C->num % reportStatsEvery;
        C->lastGood[t] = goodSpectrum ? 1 : 0;
        C->lastNclip[t] = nclip;
    }
    return 0;
}

extern "C" int orc_clipper_last_flags(orc_clipper* c, uint8_t* good, int32_t* nclip, unsigned nT)
{
    if (nT > c->lastGood.size()) return -1;
    for (unsigned t = 0; t < nT; ++t) { good[t] = c->lastGood[t]; nclip[t] = c->lastNclip[t]; }
    return 0;
}

extern "C" void orc_clipper_state(orc_clipper* c, float* o, int* sizes)
{
    o[0] = c->meanRunAve;
    o[1] = c->rmsRunAve;
    o[2] = c->bp.params[0];
    o[3] = c->bp.mean_[c->bp.currentId];
    o[4] = c->meanOverRMS;
    sizes[0] = (int)c->meanBuffer.size();
    sizes[1] = (int)c->rmsBuffer.size();
}

extern "C" int orc_bandpass_probe(unsigned nBins, double start, double width, const float* params,
                                  int nParams, float rms, float median, unsigned nBins2,
                                  double start2, double width2, float* out4)
{
    if (nParams < 1 || nBins == 0) return -1;
    BandPassR bp;
    BinMapR map;
    map.reset(nBins);
    map.setStart(start);
    map.setBinWidth(width);
    bp.setData(map, std::vector<float>(params, params + nParams));
    bp.setRMS(rms);
    if (!std::isnan(median)) bp.setMedian(median);
    if (nBins2 > 0) {
        BinMapR m2;
        m2.reset(nBins2);
        m2.setStart(start2);
        m2.setBinWidth(width2);
        bp.reBin(m2);
    }
    out4[0] = bp.median_[bp.currentId];
    out4[1] = bp.rms();
    out4[2] = bp.currentSet()[0];
    out4[3] = bp.params[0];
    return 0;
}

extern "C" int orc_clipper_bandpass(orc_clipper* c, float* out, unsigned nBins)
{
    const std::vector<float>& d = c->bp.currentSet();
    if (d.size() < nBins) return -1;  // a kill past the map's end lengthens it (BandPass.cpp:118)
    for (unsigned i = 0; i < nBins; ++i) out[i] = d[i];
    return 0;
}
