// PPFChanneliser.h -- polyphase channeliser module (src/lib/PPFChanneliser.h:
// 36-139, src/lib/src/PPFChanneliser.cpp:29-364) over the device channeliser.
#pragma once

#include <vector>

#include "../pelican/DataBlob.h"
#include "DataSets.h"

struct ppl_ppf_plan;

namespace pelican {
namespace ampp {

class PPFChanneliser : public AbstractModule {
  public:
    // Options: outputChannelsPerSubband/value (512), filter/nTaps (8),
    // filter/filterWindow (kaiser | gaussian | blackman | hamming | hann),
    // processingThreads/value (accepted, unused), and this module's own
    // dftPrecision/value (fp32 | bf16x3 | fp16x2).
    explicit PPFChanneliser(const ConfigNode& config);
    ~PPFChanneliser() override;
    // PPFChanneliser::run (:81-187): [sub][pol][block][nCh] complex samples
    // -> [block][sub][pol][nCh] spectra.  The FIR history of every
    // (subband, polarisation) stays on the device between calls.
    void run(const TimeSeriesDataSetC32* timeSeries, SpectrumDataSetC32* spectra);
    unsigned nChannels() const { return _nChannels; }
    const std::vector<float>& coefficients() const { return _coeffs; }

  private:
    void _checkData(const TimeSeriesDataSetC32* timeData) const;
    unsigned _nChannels, _nTaps;
    int _precision;
    std::vector<float> _coeffs;
    ppl_ppf_plan* _plan = nullptr;
    unsigned _planSub = 0, _planPol = 0;
};

PELICAN_DECLARE_MODULE(PPFChanneliser)

}  // namespace ampp
}  // namespace pelican
