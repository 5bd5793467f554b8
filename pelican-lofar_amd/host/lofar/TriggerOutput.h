// TriggerOutput.h -- the FRATS trigger message for the best event of a
// buffer (src/lib/TriggerOutput.h:38-75, src/lib/src/TriggerOutput.cpp:32-203).
//
// Configuration (the reference's tags):
//   <Receiver host="IP" port="N"/>        UDP receiver (numeric address)
//   <Trigger format="FRATS"/>             the only format there is
//   <Logfile name="PATH"/>                messages appended, one per line
//   <Telescope stationID=".." beamID=".."/>
//   <RADec rajd=".." decjd=".."/>  <AltAz az=".." alt=".."/>
//   <frequencyChannel1 MHz=".."/>
//   <Threshold snr="S" events="E"/>
//
// For a "DedispersionDataAnalysis" blob with at least E events (the
// time-stamp event included), the event of largest
// snr = value / (rms * sqrt(binFactor)) -- the first one if several tie, and
// only if above 0 -- is reported when snr >= S as
//
//   artemis:<station>##ID:<yyyyMMddhhmm>-<beamID:2><n:4>##cFreqMHz:<MHz:10>
//   ##beamRA:<ra:10>##beamDec:<dec:9>##eTime:<time 'f',9 :20>
//   ##eDM:<dm 'f',3 :9>##eSNR:<snr 'f',2 :6>
//
// (one line; ":w" = left-padded with '0' to w characters; n counts messages
// from 1; the ID date is local time at construction).  The message goes to
// every receiver as one datagram and, with a newline, to every log file.
#pragma once

#include <memory>
#include <string>
#include <vector>
#include <fstream>

#include "../pelican/AbstractOutputStream.h"

namespace pelican {
namespace ampp {

class DedispersionDataAnalysis;

class TriggerOutput : public AbstractOutputStream {
  public:
    explicit TriggerOutput(const ConfigNode& configNode);
    ~TriggerOutput() override;
    // A UDP receiver; an address that is not a numeric IPv4/IPv6 literal is
    // reported on stderr and skipped.
    void addReceiver(const std::string& host, unsigned short port);
    // A log file the messages are appended to.
    void addFile(const std::string& filename);

    // The message for this blob, or "" when it raises no trigger (advances
    // the message counter when it builds one).
    std::string convertToTriggerFRATS(const DedispersionDataAnalysis& data);
    int messagesSent() const { return _messageCounter; }

  protected:
    void sendStream(const std::string& streamName, const DataBlob* dataBlob) override;

  private:
    void sendMessage(const std::string& message);

    struct Receiver {
        int fd;
        std::vector<unsigned char> addr;  // sockaddr storage
        unsigned addrLen;
    };
    std::vector<Receiver> _receivers;
    std::vector<std::unique_ptr<std::ofstream>> _devices;
    std::string _idroot, _format, _station, _beamRA, _beamDec, _beamAz, _beamAlt, _cfreqMHz;
    int _minEvents = 0;
    float _snrThreshold = 0.0f;
    int _messageCounter = 0;
};

PELICAN_DECLARE(AbstractOutputStream, TriggerOutput)

}  // namespace ampp
}  // namespace pelican
