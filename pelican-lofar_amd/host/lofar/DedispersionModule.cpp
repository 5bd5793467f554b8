#include "DedispersionModule.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "../../../include/ppl_amd.h"

namespace pelican {
namespace ampp {

DedispersionModule::DedispersionModule(const ConfigNode& config) : AsyncronousModule(config)
{
    const float timeSamplesPow2 = config.getOption("timeBinsPerBufferPow2", "value", "15").toFloat();
    _numSamplesBuffer = (unsigned)(int)std::pow(2.0, timeSamplesPow2);
    _tdms = config.getOption("dedispersionSamples", "value", "1984").toUInt();
    _dmStep = config.getOption("dedispersionStepSize", "value", "0.0").toFloat();
    _dmLow = config.getOption("dedispersionMinimum", "value", "0.0").toFloat();
    if (_dmLow < 0.0f) _dmLow = 0.0f;
    _fch1 = config.getOption("frequencyChannel1", "MHz", "0.0").toDouble();
    _foff = config.getOption("channelBandwidth", "MHz", "1.0").toDouble();
    if (_fch1 == 0) throw std::runtime_error("DedispersionModule: frequencyChannel1 must be a positve number");
    if (_foff >= 0)
        throw std::runtime_error(
            "DedispersionModule: channelBandwidth must be negative (channels in descending frequency)");
    _maxBuffers = config.getOption("numberOfBuffers", "value", "2").toUInt();
    if (_maxBuffers < 1) throw std::runtime_error("DedispersionModule: Must have at least one buffer");
    const std::string prec = config.getOption("inputPrecision", "value", "f32").toLower();
    if (prec == "f32") _dtype = PPL_F32;
    else if (prec == "f16") _dtype = PPL_F16;
    else if (prec == "u8") _dtype = PPL_U8;
    else throw std::runtime_error("DedispersionModule: unknown inputPrecision '" + prec + "'");
    _invScale = config.getOption("quantiseScale", "value", "12.7").toFloat();
}

DedispersionModule::~DedispersionModule()
{
    waitForJobCompletion();
    _cleanBuffers();
}

void DedispersionModule::_cleanBuffers()
{
    ppl_dedisp_plan_destroy(_plan);
    ppl_ddbuf_destroy(_buffer);
    ppl_device_free(_sums);
    _plan = nullptr;
    _buffer = nullptr;
    _sums = nullptr;
    _outputs.clear();
    _free.clear();
    _segments.clear();
}

void DedispersionModule::waitForJobCompletion()
{
    // outputs come back to the free list in exportComplete
    ppl_check(ppl_stream_sync(nullptr), "DedispersionModule");
    AsyncronousModule::waitForJobCompletion();
}

void DedispersionModule::resize(const SpectrumDataSet<float>* streamData)
{
    const unsigned nCh = streamData->nChannels(), nSub = streamData->nSubbands();
    const unsigned sampleSize = nSub * nCh;
    if (sampleSize == _sampleSize && _plan) return;
    waitForJobCompletion();
    _cleanBuffers();
    _tsamp = streamData->getBlockRate();
    ppl_dedisp_config cfg{};
    cfg.nChan = sampleSize;
    cfg.nsampBuf = _numSamplesBuffer;
    cfg.tdms = _tdms;
    cfg.dmLow = _dmLow;
    cfg.dmStep = _dmStep;
    cfg.tsamp = _tsamp;
    cfg.fch1 = _fch1;
    cfg.foff = _foff;
    cfg.dtype = _dtype;
    const int rc = ppl_dedisp_plan_create(&_plan, &cfg);
    if (rc != PPL_OK) throw std::runtime_error(std::string("DedispersionModule: ") + ppl_last_error());
    ppl_dedisp_info info{};
    ppl_check(ppl_dedisp_plan_info(_plan, &info), "DedispersionModule");
    _maxshift = info.maxshift;
    _remainingSamples = info.remainingSamples;
    _nOut = info.nOut;
    ppl_check(ppl_ddbuf_create(&_buffer, sampleSize, _numSamplesBuffer, (unsigned)(_maxshift + _remainingSamples),
                               _dtype, _invScale),
              "DedispersionModule");
    if (_dtype == PPL_U8)
        ppl_check(ppl_device_alloc(&_sums, sizeof(int32_t) * (size_t)_tdms * (size_t)_nOut), "DedispersionModule");
    for (unsigned i = 0; i < _maxBuffers; ++i) {
        _outputs.emplace_back(new DedispersionSpectra());
        _free.push_back(_outputs.back().get());
    }
    _sampleSize = sampleSize;
    _nSub = nSub;
    _nCh = nCh;
}

void DedispersionModule::dedisperse(DataBlob* incoming) { dedisperse(dynamic_cast<WeightedSpectrumDataSet*>(incoming)); }

void DedispersionModule::dedisperse(WeightedSpectrumDataSet* weightedData)
{
    if (!weightedData) throw std::runtime_error("DedispersionModule: not a WeightedSpectrumDataSet");
    auto* streamData = static_cast<SpectrumDataSetStokes*>(weightedData->dataSet());
    resize(streamData);
    const unsigned maxSamples = streamData->nTimeBlocks();
    const float* d = static_cast<const float*>(streamData->deviceData(nullptr));
    unsigned sampleNumber = 0;
    bool lockedHere = false;
    while (sampleNumber != maxSamples) {
        unsigned consumed = 0;
        int full = 0;
        ppl_check(ppl_ddbuf_add_blob(_buffer, d, maxSamples, sampleNumber, streamData->nSubbands(),
                                     streamData->nPolarisations(), streamData->nChannels(), &consumed, &full, nullptr),
                  "DedispersionModule");
        if (consumed) {
            if (!lockedHere) {
                lock(streamData);  // once per buffer the blob enters
                lockedHere = true;
            }
            _segments.push_back({streamData, sampleNumber, consumed});
            sampleNumber += consumed;
        }
        if (full) {
            _launch();
            // a blob carried into (or continuing in) the next buffer was
            // locked for it by _launch, or is locked on its next add
            lockedHere = false;
            for (const auto& sgm : _segments)
                if (sgm.blob == streamData) lockedHere = true;
        }
    }
}

DedispersionSpectra* DedispersionModule::_nextOutput()
{
    std::unique_lock<std::mutex> l(_freeMutex);
    _freeCv.wait(l, [this] { return !_free.empty(); });
    DedispersionSpectra* out = _free.front();
    _free.pop_front();
    return out;
}

namespace {
struct GpuJob {
    DedispersionModule* module;
    DedispersionSpectra* out;
};
}  // namespace

void DedispersionModule::_launch()
{
    DedispersionSpectra* out = _nextOutput();
    out->resize((unsigned)_nOut, _tdms, _dmLow, _dmStep);
    std::vector<SpectrumDataSetStokes*> blobs;
    for (const auto& sgm : _segments)
        if (blobs.empty() || blobs.back() != sgm.blob) blobs.push_back(sgm.blob);
    out->setInputDataBlobs(blobs);
    out->setFirstSample(_segments.empty() ? 0u : _segments.front().start);
    void* d_out = out->deviceDataForWrite();
    if (_dtype == PPL_U8) {
        ppl_check(ppl_dedisperse(_plan, ppl_ddbuf_data(_buffer), _sums, nullptr), "DedispersionModule");
        // the float series with the buffer's exceptional rows restored, before
        // the rotate below moves its row offsets
        ppl_check(ppl_dedisp_u8_series(_plan, static_cast<const int32_t*>(_sums), ppl_ddbuf_row_offsets(_buffer),
                                       1.0f / _invScale, static_cast<float*>(d_out), nullptr),
                  "DedispersionModule");
    } else {
        ppl_check(ppl_dedisperse(_plan, ppl_ddbuf_data(_buffer), d_out, nullptr), "DedispersionModule");
    }
    // DedispersionBuffer::copy: the last maxshift + remaining samples start
    // the next buffer, with the blobs that hold them
    unsigned keep = (unsigned)(_maxshift + _remainingSamples);
    std::vector<Segment> next;
    for (auto it = _segments.rbegin(); it != _segments.rend() && keep > 0; ++it) {
        const unsigned take = std::min(keep, it->count);
        next.insert(next.begin(), Segment{it->blob, it->start + it->count - take, take});
        keep -= take;
    }
    {
        std::lock_guard<std::mutex> l(lockerMutex);
        SpectrumDataSetStokes* prev = nullptr;
        for (const auto& sgm : next)
            if (sgm.blob != prev) {
                lockUnprotected(sgm.blob);
                prev = sgm.blob;
            }
    }
    _segments.swap(next);
    ppl_check(ppl_ddbuf_rotate(_buffer, nullptr), "DedispersionModule");
    ppl_check(ppl_stream_callback(nullptr, &DedispersionModule::_gpuDone, new GpuJob{this, out}),
              "DedispersionModule");
}

void DedispersionModule::_gpuDone(void* user, int status)
{
    // runs on the runtime's callback thread: no device calls here, only a
    // hand-over to the export worker
    GpuJob* job = static_cast<GpuJob*>(user);
    DedispersionModule* m = job->module;
    DedispersionSpectra* out = job->out;
    delete job;
    if (status == PPL_OK) {
        m->exportData(out);
    } else {
        m->enqueue([m, out] { m->exportCancel(out); });
    }
}

void DedispersionModule::exportComplete(DataBlob* datablob)
{
    auto* data = static_cast<DedispersionSpectra*>(datablob);
    for (SpectrumDataSetStokes* d : data->inputDataBlobs()) unlock(d);
    {
        std::lock_guard<std::mutex> l(_freeMutex);
        _free.push_back(data);
    }
    _freeCv.notify_one();
}

DedispersionAnalyser::DedispersionAnalyser(const ConfigNode& config) : AbstractModule(config)
{
    _detectionThreshold = config.getOption("detectionThreshold", "in_sigma", "6.0").toFloat();
    _binPow2 = config.getOption("power2ForBinning", "value", "6").toUInt();
}

int DedispersionAnalyser::analyse(DedispersionSpectra* data, DedispersionDataAnalysis* result)
{
    result->reset(data);
    const auto& blobs = data->inputDataBlobs();
    if (blobs.empty()) throw std::runtime_error("DedispersionAnalyser: dedispersion data without input blobs");
    const int nChannels = (int)blobs[0]->nChannels(), nSubbands = (int)blobs[0]->nSubbands();
    result->setRMS(std::sqrt((float)nChannels * (float)nSubbands));
    const int tdms = data->dmBins(), nsamp = data->timeSamples();
    const float* d = static_cast<const float*>(data->deviceData(nullptr));
    if (_ints.empty()) {
        _ints.resize(3 * 65536);
        _vals.resize(65536);
    }
    int count = 0;
    for (;;) {
        ppl_check(ppl_detect(d, tdms, nsamp, nChannels, nSubbands, _detectionThreshold, _binPow2, _ints.data(),
                             _vals.data(), (int)_vals.size(), &count, nullptr),
                  "DedispersionAnalyser");
        if (count <= (int)_vals.size()) break;
        _ints.resize(3 * (size_t)count);
        _vals.resize((size_t)count);
    }
    for (int e = 0; e < count; ++e)
        result->addEvent((unsigned)_ints[3 * e], (unsigned)_ints[3 * e + 1], (float)_ints[3 * e + 2], _vals[e]);
    return result->eventsFound();
}

}  // namespace ampp
}  // namespace pelican
