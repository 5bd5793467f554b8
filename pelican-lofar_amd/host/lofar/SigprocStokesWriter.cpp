// SigprocStokesWriter.cpp -- Sigproc filterbank writer
// (src/lib/src/SigprocStokesWriter.cpp:15-290).
#include "SigprocStokesWriter.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <iostream>
#include <stdexcept>

#include "../../../include/ppl_amd.h"
#include "DataSets.h"
#include "DeviceMirror.h"

namespace pelican {
namespace ampp {

namespace {

// QString::toFloat(&ok): the whole (trimmed) text must be a number.
bool strictFloat(const std::string& text, float* out)
{
    const size_t a = text.find_first_not_of(" \t\r\n"), b = text.find_last_not_of(" \t\r\n");
    if (a == std::string::npos) return false;
    const std::string t = text.substr(a, b - a + 1);
    char* end = nullptr;
    const float v = std::strtof(t.c_str(), &end);
    if (end != t.c_str() + t.size()) return false;
    *out = v;
    return true;
}

}  // namespace

SigprocStokesWriter::SigprocStokesWriter(const ConfigNode& configNode) : AbstractOutputStream(configNode)
{
    // (:15-121), in the reference's types
    _nSubbands = configNode.getOption("subbandsPerPacket", "value", "1").toUInt();
    _nTotalSubbands = (int)configNode.getOption("totalComplexSubbands", "value", "1").toUInt();
    _clock = configNode.getOption("clock", "value", "200").toFloat();
    _integration = configNode.getOption("integrateTimeBins", "value", "1").toUInt();
    _integrationFreq = configNode.getOption("integrateFrequencyChannels", "value", "1").toUInt();
    _nChannels = configNode.getOption("outputChannelsPerSubband", "value", "128").toUInt();
    if (_integrationFreq == 0) throw std::runtime_error("SigprocStokesWriter: integrateFrequencyChannels is 0");
    _nChannels = _nChannels / _integrationFreq;
    _nRawPols = configNode.getOption("nRawPolarisations", "value", "2").toUInt();
    _nBits = configNode.getOption("dataBits", "value", "32").toUInt();
    // (int)pow(2, nBits) - 1: the byte range for 8 bits (unused at 32)
    _nRange = _nBits < 31 ? (float)((int)std::pow(2.0, (double)_nBits) - 1.0) : 0.0f;
    _cropMin = configNode.getOption("scale", "min", "0.0").toFloat();
    if (!strictFloat(configNode.getOption("scale", "max", "X").str(), &_cropMax)) _cropMax = _nRange;
    _scaleDelta = _cropMax - _cropMin;

    _filepath = configNode.getOption("file", "filepath").str();
    _topsubband = (unsigned)configNode.getOption("topSubbandIndex", "value", "150").toFloat();
    _lbahba = (unsigned)configNode.getOption("LBA_0_or_HBA_1", "value", "1").toFloat();
    const float subbandwidth = _clock / (float)(_nRawPols * (unsigned)_nTotalSubbands);  // MHz
    const float channelwidth = subbandwidth / (float)_nChannels;
    if (configNode.getOption("fch1", "value").isEmpty()) {
        // the centre frequency of the top channel (:44-68)
        if (_lbahba == 0) {
            _fch1 = (float)((double)(subbandwidth * (float)_topsubband) + 0.5 * subbandwidth - 0.5 * channelwidth);
        } else {
            // HBA: 100 MHz above the clock's first Nyquist zone at 200 MHz,
            // 160 MHz at 160; other clocks leave it at 0
            const int base = _clock == 200.0f ? 100 : _clock == 160.0f ? 160 : -1;
            if (base >= 0)
                _fch1 = (float)((double)((float)base + subbandwidth * (float)_topsubband) + 0.5 * subbandwidth -
                                0.5 * channelwidth);
        }
    } else {
        _fch1 = configNode.getOption("fch1", "value", "1400.0").toFloat();
    }
    _foff = configNode.getOption("foff", "value").isEmpty() ? -channelwidth
                                                            : configNode.getOption("foff", "value", "1.0").toFloat();
    if (configNode.getOption("tsamp", "value").isEmpty())
        _tsamp = (float)((double)((float)(_nRawPols * (unsigned)_nTotalSubbands * _nChannels * _integrationFreq *
                                          _integration) /
                                  _clock) /
                         1e6);
    else
        _tsamp = configNode.getOption("tsamp", "value", "1.0").toFloat();

    _nPols = configNode.getOption("params", "nPolsToWrite", "1").toUInt();
    _nchans = (int)(_nChannels * _nSubbands);
    _buffSize = (int)configNode.getOption("params", "bufferSize", "5120").toUInt();
    _first = configNode.hasAttribute("writeHeader") && configNode.getAttribute("writeHeader").toLower() == "true";
    _site = configNode.getOption("TelescopeID", "value", "0").toUInt();
    _machine = configNode.getOption("MachineID", "value", "9").toUInt();
    _raString = configNode.getOption("RAJ", "value", "000000.0").str();
    _decString = configNode.getOption("DecJ", "value", "000000.0").str();
    _ra = ConfigValue(_raString).toFloat();
    _dec = ConfigValue(_decString).toFloat();
    _sourceName = _raString.substr(0, 4);
    if (_dec > 0.0f) _sourceName += "+" + _decString.substr(0, 4);
    else _sourceName += _decString.substr(0, 5);

    char timestr[22];
    std::time_t now = std::time(nullptr);
    std::tm tstruct = *std::localtime(&now);
    std::strftime(timestr, sizeof timestr, "D%Y%m%dT%H%M%S", &tstruct);
    _fileName = _filepath + "_" + timestr + ".fil";
    _file.open(_fileName, std::ios::out | std::ios::binary);
    if (!_file) std::cerr << "SigprocStokesWriter: unable to open file for writing: " << _fileName << std::endl;
}

SigprocStokesWriter::~SigprocStokesWriter()
{
    _file.close();
    ppl_device_free(_dPack);
    ppl_host_free(_hPack);
}

void SigprocStokesWriter::writeString(const std::string& s)
{
    const int len = (int)s.size();
    _file.write(reinterpret_cast<const char*>(&len), sizeof(int));
    _file.write(s.data(), len);
}

void SigprocStokesWriter::writeInt(const std::string& name, int value)
{
    writeString(name);
    _file.write(reinterpret_cast<const char*>(&value), sizeof(int));
}

void SigprocStokesWriter::writeDouble(const std::string& name, double value)
{
    writeString(name);
    _file.write(reinterpret_cast<const char*>(&value), sizeof(double));
}

void SigprocStokesWriter::writeHeader(const SpectrumDataSetStokes* stokes)
{
    // (:123-158): tstart is the MJD of the blob's time stamp, the epoch
    // 2011-01-01 00:00:00 local time being MJD 55562
    const double timeStamp = stokes->getLofarTimestamp();
    std::tm tm;
    std::memset(&tm, 0, sizeof tm);
    tm.tm_year = 111;
    tm.tm_mon = 0;
    tm.tm_mday = 1;
    tm.tm_isdst = -1;
    const std::time_t epoch = std::mktime(&tm);
    if (epoch == (std::time_t)-1) throw std::runtime_error("SigprocStokesWriter: unable to set epoch.");
    const double mjdStamp = (timeStamp - (double)epoch) / 86400 + 55562.0;

    writeString("HEADER_START");
    writeInt("machine_id", (int)_machine);
    writeInt("telescope_id", (int)_site);
    writeInt("data_type", 1);
    writeString("source_name");
    writeString(_sourceName);
    writeDouble("src_raj", _ra);
    writeDouble("src_dej", _dec);
    writeDouble("fch1", _fch1);
    writeDouble("foff", _foff);
    writeInt("nchans", _nchans);
    writeDouble("tsamp", _tsamp);
    writeInt("nbits", (int)_nBits);
    writeDouble("tstart", mjdStamp);
    writeInt("nifs", (int)_nPols);
    writeString("HEADER_END");
    _file.flush();
}

void SigprocStokesWriter::sendStream(const std::string& /*streamName*/, const DataBlob* incoming)
{
    const auto* stokes = dynamic_cast<const SpectrumDataSetStokes*>(incoming);
    if (!stokes) {
        std::cerr << "SigprocStokesWriter::send(): "
                     "Only SpectrumDataSetStokes data can be written by the SigprocWriter"
                  << std::endl;
        return;
    }
    if (_first) {
        _first = false;
        writeHeader(stokes);
    }
    // (:184-240): the data, formed on the device
    const unsigned nSamples = stokes->nTimeBlocks(), nSubbands = stokes->nSubbands();
    const unsigned nChannels = stokes->nChannels(), nPolarisations = stokes->nPolarisations();
    if (_nBits != 32 && _nBits != 8)
        throw std::runtime_error("SigprocStokesWriter: " + std::to_string(_nBits) +
                                 " bit datafiles not yet supported");
    const size_t bytes = (size_t)nSamples * _nPols * nSubbands * nChannels * (_nBits / 8);
    if (bytes == 0) return;
    if (bytes > _packBytes) {
        ppl_device_free(_dPack);
        ppl_host_free(_hPack);
        _dPack = _hPack = nullptr;
        _packBytes = 0;
        ppl_check(ppl_device_alloc(&_dPack, bytes), "SigprocStokesWriter");
        ppl_check(ppl_host_alloc(&_hPack, bytes), "SigprocStokesWriter");
        _packBytes = bytes;
    }
    const float* d_I = static_cast<const float*>(stokes->deviceData(nullptr));
    ppl_check(ppl_sigproc_pack(d_I, nSamples, nSubbands, nPolarisations, nChannels, _nPols, (int)_nBits, _cropMin,
                               _cropMax, _dPack, nullptr),
              "SigprocStokesWriter");
    ppl_check(ppl_memcpy_d2h(_hPack, _dPack, bytes, nullptr), "SigprocStokesWriter");
    ppl_check(ppl_stream_sync(nullptr), "SigprocStokesWriter");
    _file.write(static_cast<const char*>(_hPack), (std::streamsize)bytes);
    _file.flush();
}

}  // namespace ampp
}  // namespace pelican
