// TriggerOutput.cpp -- FRATS trigger messages (src/lib/src/TriggerOutput.cpp:32-203).
#include "TriggerOutput.h"

#include <netdb.h>
#include <sys/socket.h>
#include <unistd.h>

#include <cerrno>
#include <cstring>
#include <ctime>
#include <iostream>

#include "DataSets.h"
#include "DedispersionDataAnalysisOutput.h"
#include "TextFormat.h"

namespace pelican {
namespace ampp {

TriggerOutput::TriggerOutput(const ConfigNode& configNode) : AbstractOutputStream(configNode)
{
    // receiver (:35-42)
    const int port = configNode.getOption("Receiver", "port").toInt();
    const std::string host = configNode.getOption("Receiver", "host").str();
    if (!host.empty() && port != 0) addReceiver(host, (unsigned short)port);

    _format = configNode.getOption("Trigger", "format").str();

    // log file (:48-52)
    const std::string logfile = configNode.getOption("Logfile", "name").str();
    if (!logfile.empty()) addFile(logfile);

    // identifier root "<yyyyMMddhhmm>-<beamID:2>%3" (:55-58)
    _station = configNode.getOption("Telescope", "stationID").str();
    char date[16];
    std::time_t now = std::time(nullptr);
    std::tm t = *std::localtime(&now);
    std::strftime(date, sizeof date, "%Y%m%d%H%M", &t);
    _idroot = text::arg("%1-%2%3", date);
    _idroot = text::arg(_idroot, configNode.getOption("Telescope", "beamID").str(), 2, '0');

    // beam direction, frequency and thresholds (:61-68)
    _beamRA = configNode.getOption("RADec", "rajd").str();
    _beamDec = configNode.getOption("RADec", "decjd").str();
    _beamAz = configNode.getOption("AltAz", "az").str();
    _beamAlt = configNode.getOption("AltAz", "alt").str();
    _cfreqMHz = configNode.getOption("frequencyChannel1", "MHz").str();
    _snrThreshold = configNode.getOption("Threshold", "snr").toFloat();
    _minEvents = configNode.getOption("Threshold", "events").toInt();
}

TriggerOutput::~TriggerOutput()
{
    for (auto& r : _receivers) ::close(r.fd);
}

void TriggerOutput::addFile(const std::string& logfile)
{
    std::unique_ptr<std::ofstream> f(new std::ofstream(logfile, std::ios::out | std::ios::app));
    if (!*f) {
        std::cerr << "unable to open file for writing: " << logfile << std::endl;
        return;
    }
    _devices.push_back(std::move(f));
}

void TriggerOutput::addReceiver(const std::string& host, unsigned short port)
{
    addrinfo hints;
    std::memset(&hints, 0, sizeof hints);
    hints.ai_family = AF_UNSPEC;
    hints.ai_socktype = SOCK_DGRAM;
    hints.ai_flags = AI_NUMERICHOST | AI_NUMERICSERV;
    addrinfo* res = nullptr;
    const std::string service = std::to_string(port);
    if (getaddrinfo(host.c_str(), service.c_str(), &hints, &res) != 0 || !res) {
        std::cerr << "TriggerOutput: not a numeric address: " << host << std::endl;
        return;
    }
    const int fd = ::socket(res->ai_family, res->ai_socktype, res->ai_protocol);
    if (fd < 0) {
        std::cerr << "TriggerOutput: cannot open a UDP socket: " << std::strerror(errno) << std::endl;
        freeaddrinfo(res);
        return;
    }
    Receiver r;
    r.fd = fd;
    r.addrLen = (unsigned)res->ai_addrlen;
    r.addr.assign((const unsigned char*)res->ai_addr, (const unsigned char*)res->ai_addr + res->ai_addrlen);
    freeaddrinfo(res);
    _receivers.push_back(std::move(r));
}

void TriggerOutput::sendStream(const std::string& /*streamName*/, const DataBlob* dataBlob)
{
    if (_format == "FRATS") {
        const std::string message = convertToTriggerFRATS(*static_cast<const DedispersionDataAnalysis*>(dataBlob));
        if (!message.empty()) sendMessage(message);
    } else {
        std::cerr << "Trigger format not recognized / implemented";
    }
}

std::string TriggerOutput::convertToTriggerFRATS(const DedispersionDataAnalysis& data)
{
    // best event over all events, the time-stamp event included (:148-157)
    if (data.eventsFound() < _minEvents) return std::string();
    const float rms = data.getRMS();
    const DedispersionEvent* emax = nullptr;
    float snrMax = 0.0f;
    for (const DedispersionEvent& e : data.events()) {
        const float snr = eventSnr(e, rms);
        if (snr > snrMax) {
            snrMax = snr;
            emax = &e;
        }
    }
    if (!emax || !(snrMax >= _snrThreshold)) return std::string();
    if (_beamRA.empty()) std::cerr << "No RA/Dec found, conversion not yet implemented";

    // the message (:161-169)
    std::string m = "artemis:%1##ID:%2##cFreqMHz:%3##beamRA:%4##beamDec:%5##eTime:%6##eDM:%7##eSNR:%8";
    m = text::arg(m, _station);
    m = text::arg(m, text::arg(_idroot, std::to_string(++_messageCounter), 4, '0'));
    m = text::arg(m, _cfreqMHz, 10, '0');
    m = text::arg(m, _beamRA, 10, '0');
    m = text::arg(m, _beamDec, 9, '0');
    m = text::arg(m, text::fixedReal(emax->getTime(), 9), 20, '0');
    m = text::arg(m, text::fixedReal(emax->dm(), 3), 9, '0');
    m = text::arg(m, text::fixedReal(snrMax, 2), 6, '0');
    return m;
}

void TriggerOutput::sendMessage(const std::string& message)
{
    // one datagram per receiver, then the log line (:180-203)
    for (const Receiver& r : _receivers) {
        if (::sendto(r.fd, message.data(), message.size(), 0, (const sockaddr*)r.addr.data(), r.addrLen) < 0)
            std::cerr << "TriggerOutput: datagram not sent: " << std::strerror(errno) << std::endl;
    }
    std::cout << "Datagram written" << std::endl;
    for (auto& d : _devices) {
        *d << message << "\n";
        d->flush();
    }
}

}  // namespace ampp
}  // namespace pelican
