// DedispersionDataAnalysisOutput.h -- the event text file writer
// (src/lib/DedispersionDataAnalysisOutput.h:27-45,
//  src/lib/src/DedispersionDataAnalysisOutput.cpp:18-112).
//
// Configuration (as the reference's):  <file name="PREFIX"/>.  The
// constructor opens PREFIX_DYYYYMMDDTHHMMSS.dat (local time) and writes the
// five-line header; addFile opens further files.  Every
// "DedispersionDataAnalysis" blob sent appends, to each file, one line per
// event after the first (the analyser's time-stamp event)
//
//     <mjd>,   <dm>, <snr>, <binFactor>
//
// with mjd = event time / 86400, snr = value / (rms * sqrt(binFactor)), the
// reals streamed with 14 significant digits, then the line
//
//     # Written buffer :<k> | MJDstart: <mjd of event 0> | Best DM: <dm> | Max SNR: <snr>  Done
//
// where Best DM / Max SNR belong to the first event of largest snr above 0
// (0 and 0 when none is), and k counts the summary lines this writer has
// written over all its files.
#pragma once

#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "../pelican/AbstractOutputStream.h"

namespace pelican {
namespace ampp {

class DedispersionDataAnalysis;
class DedispersionEvent;

// value / (rms * sqrt(binFactor)), the event signal-to-noise both event
// writers report.
float eventSnr(const DedispersionEvent& e, float rms);

class DedispersionDataAnalysisOutput : public AbstractOutputStream {
  public:
    explicit DedispersionDataAnalysisOutput(const ConfigNode& configNode);
    ~DedispersionDataAnalysisOutput() override;
    // Opens `filename` for writing (truncating) and writes the header; a file
    // that cannot be opened is reported on stderr and skipped.
    void addFile(const std::string& filename);
    // The name the constructor chose (empty if it could not open it).
    const std::string& fileName() const { return _firstFile; }

    // The text one blob adds to every file (also what sendStream writes).
    std::string format(const DedispersionDataAnalysis& data, int dumpIndex) const;
    static const char* header();

  protected:
    void sendStream(const std::string& streamName, const DataBlob* dataBlob) override;

  private:
    std::vector<std::unique_ptr<std::ofstream>> _streams;
    std::string _firstFile;
    int _indexOfDump = 0;
};

PELICAN_DECLARE(AbstractOutputStream, DedispersionDataAnalysisOutput)

}  // namespace ampp
}  // namespace pelican
