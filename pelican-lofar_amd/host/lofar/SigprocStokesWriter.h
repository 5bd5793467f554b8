// SigprocStokesWriter.h -- Stokes blobs to a Sigproc filterbank file
// (src/lib/SigprocStokesWriter.h:26-77, src/lib/src/SigprocStokesWriter.cpp:15-290).
//
// Configuration: the reference's tags and defaults (subbandsPerPacket,
// totalComplexSubbands, clock, integrateTimeBins, integrateFrequencyChannels,
// outputChannelsPerSubband, nRawPolarisations, dataBits 32|8, scale min/max,
// file filepath, topSubbandIndex, LBA_0_or_HBA_1, fch1, foff, tsamp,
// params nPolsToWrite/bufferSize, TelescopeID, MachineID, RAJ, DecJ) and the
// element's own writeHeader="true" attribute.  The file is
// <filepath>_DYYYYMMDDTHHMMSS.fil (local time at construction).  With writeHeader, the first blob is preceded by the
// Sigproc header (HEADER_START ... HEADER_END: machine_id, telescope_id,
// data_type 1, source_name, src_raj, src_dej, fch1, foff, nchans, tsamp,
// nbits, tstart as MJD of the blob's time stamp, nifs).  Every blob then
// appends its spectra, formed on the device by ppl_sigproc_pack and copied
// back in one transfer.
#pragma once

#include <fstream>
#include <string>

#include "../pelican/AbstractOutputStream.h"

namespace pelican {
namespace ampp {

class SpectrumDataSetStokes;

class SigprocStokesWriter : public AbstractOutputStream {
  public:
    explicit SigprocStokesWriter(const ConfigNode& config);
    ~SigprocStokesWriter() override;
    const std::string& filepath() const { return _filepath; }
    const std::string& fileName() const { return _fileName; }
    // header values as configured (for callers that check them)
    float fch1() const { return _fch1; }
    float foff() const { return _foff; }
    float tsamp() const { return _tsamp; }
    int nchans() const { return _nchans; }
    const std::string& sourceName() const { return _sourceName; }

  protected:
    void sendStream(const std::string& streamName, const DataBlob* dataBlob) override;

  private:
    void writeHeader(const SpectrumDataSetStokes* stokes);
    void writeString(const std::string& s);
    void writeInt(const std::string& name, int value);
    void writeDouble(const std::string& name, double value);

    bool _first = false;
    std::string _filepath, _fileName, _sourceName, _raString, _decString;
    std::ofstream _file;
    float _fch1 = 0.0f, _foff = 0.0f, _tsamp = 0.0f, _clock = 0.0f, _ra = 0.0f, _dec = 0.0f;
    float _cropMin = 0.0f, _cropMax = 0.0f, _scaleDelta = 0.0f, _nRange = 0.0f;
    int _nchans = 0, _nTotalSubbands = 0, _buffSize = 0;
    unsigned _nRawPols = 0, _nChannels = 0, _nSubbands = 0, _integration = 0, _nPols = 0;
    unsigned _topsubband = 0, _lbahba = 0, _site = 0, _machine = 0, _nBits = 0, _integrationFreq = 0;
    // device and pinned host staging for one blob's bytes
    void* _dPack = nullptr;
    void* _hPack = nullptr;
    size_t _packBytes = 0;
};

PELICAN_DECLARE(AbstractOutputStream, SigprocStokesWriter)

}  // namespace ampp
}  // namespace pelican
