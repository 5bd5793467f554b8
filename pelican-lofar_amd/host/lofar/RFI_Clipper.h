// RFI_Clipper.h -- bandpass-model RFI clipper (src/lib/RFI_Clipper.h:25-74,
// src/lib/src/RFI_Clipper.cpp:22-573) over the device clipper.  The .bp
// file is read as BandPassAdapter::deserialise does
// (src/lib/src/BandPassAdapter.cpp:30-107).
#pragma once

#include <string>
#include <vector>

#include "../pelican/DataBlob.h"
#include "BandPass.h"
#include "DataSets.h"

struct ppl_clipper;

namespace pelican {
namespace ampp {

// The numbers of a .bp file in order: nChan, start, end, width, rms,
// median, polynomial coefficients...; the indices of its K (kill) lines go
// to *kills when given.  Throws with the reference's messages.
std::vector<double> readBandPassFile(const std::string& fileName, std::vector<unsigned>* kills = nullptr);

class RFI_Clipper : public AbstractModule {
  public:
    // Options: attribute active (default true); BandPassData/file (required
    // when active); attributes channelRejectionRMS (10) and
    // spectrumRejectionRMS (4); History/maximum (10); zeroDMing/active;
    // computeRMSfromMean/active (true); Band/matching (must be true).
    explicit RFI_Clipper(const ConfigNode& config);
    ~RFI_Clipper() override;
    // In place on the Stokes blob of `weightedStokes` ([t][sub][pol][ch]);
    // every weight becomes 1; the blob's mean and rms become the running
    // mean and rms after the last spectrum.
    void run(WeightedSpectrumDataSet* weightedStokes);
    // Per-spectrum flags of the last run (1 = kept).
    std::vector<unsigned char> lastFlags() const;
    const std::vector<double>& bandPassValues() const { return _bp; }
    const std::vector<unsigned>& killedChannels() const { return _kill; }
    // The bandpass model (RFI_Clipper.h:38): BandPassAdapter's model from
    // the file until the first run, then the live model on the run map as
    // setMean / setRMS leave it after the last spectrum (its constant
    // coefficient, mean and RMS read back from the device).
    const BandPass& bandPass() const { return _bandPass; }

  private:
    bool _active = true;
    std::vector<double> _bp;
    std::vector<unsigned> _kill;
    BandPass _bandPass;
    float _startFrequency = 0.0f, _endFrequency = 0.0f;
    ppl_clipper* _clip = nullptr;
    unsigned _lastNT = 0;
};

PELICAN_DECLARE_MODULE(RFI_Clipper)

}  // namespace ampp
}  // namespace pelican
