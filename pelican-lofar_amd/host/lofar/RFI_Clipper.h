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
// median, polynomial coefficients...  Throws with the reference's messages.
std::vector<double> readBandPassFile(const std::string& fileName);

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
    // The bandpass read from the file (RFI_Clipper.h:38): BandPassAdapter's
    // model on the file's map, its median and RMS.  The running model the
    // clipper updates per spectrum stays on the device.
    const BandPass& bandPass() const { return _bandPass; }

  private:
    bool _active = true;
    std::vector<double> _bp;
    BandPass _bandPass;
    ppl_clipper* _clip = nullptr;
    unsigned _lastNT = 0;
};

PELICAN_DECLARE_MODULE(RFI_Clipper)

}  // namespace ampp
}  // namespace pelican
