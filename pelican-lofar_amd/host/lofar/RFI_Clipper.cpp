#include "RFI_Clipper.h"

#include <fstream>
#include <sstream>
#include <stdexcept>

#include "../../../include/ppl_amd.h"

namespace pelican {
namespace ampp {

std::vector<double> readBandPassFile(const std::string& fileName, std::vector<unsigned>* kills)
{
    std::ifstream in(fileName);
    if (!in) throw std::runtime_error("RFI_Clipper: Cannot open File \"" + fileName + "\"");
    std::vector<double> params;
    std::string line;
    int lineNo = 0;
    if (kills) kills->clear();
    while (std::getline(in, line)) {
        ++lineNo;
        if (!line.empty() && line[0] == '#') continue;  // comment
        if (!line.empty() && line[0] == 'K') {
            // kill line (BandPassAdapter.cpp:58-71): "index frequency", at
            // least two whitespace-separated fields, the first an unsigned
            // integer (the reference takes it for the frequency too)
            std::istringstream fields(line.substr(1));
            std::vector<std::string> nums;
            for (std::string w; fields >> w;) nums.push_back(w);
            if (nums.size() < 2)
                throw std::runtime_error("BandPassAdapter: syntax error on line " + std::to_string(lineNo) +
                                         ": expecting two numbers on Kill line: index frequency");
            const std::string& n0 = nums[0];
            if (n0.empty() || n0.find_first_not_of("0123456789") != std::string::npos || n0.size() > 10 ||
                std::stoull(n0) > 0xFFFFFFFFull)
                throw std::runtime_error("BandPassAdapter: syntax error on line " + std::to_string(lineNo) +
                                         " - cannot convert index to unsigned integer");
            if (kills) kills->push_back((unsigned)std::stoul(n0));
            continue;
        }
        // QByteArray::trimmed().toDouble(): the whole trimmed line is a number
        size_t a = line.find_first_not_of(" \t\r\n\v\f");
        size_t b = line.find_last_not_of(" \t\r\n\v\f");
        const std::string s = a == std::string::npos ? std::string() : line.substr(a, b - a + 1);
        char* end = nullptr;
        const double v = s.empty() ? 0.0 : std::strtod(s.c_str(), &end);
        if (s.empty() || end != s.c_str() + s.size())
            throw std::runtime_error("BandPassAdapter: syntax error on line " + std::to_string(lineNo));
        params.push_back(v);
    }
    if (params.size() < 6)
        throw std::runtime_error("BandPassAdapter: data format wrong : got " + std::to_string(params.size()) +
                                 " parameters, expecting at least 6");
    return params;
}

RFI_Clipper::RFI_Clipper(const ConfigNode& config) : AbstractModule(config)
{
    if (config.hasAttribute("active") && config.getAttribute("active").toLower() == "false") _active = false;
    const std::string file = config.getOption("BandPassData", "file", "");
    if (file.empty()) {
        if (_active) throw std::runtime_error("RFI_Clipper: <BandPassData file=?> not defined");
        return;
    }
    if (!_active) return;
    std::string path;
    try {
        path = config.searchFile(file);
    } catch (const std::exception& e) {
        throw std::runtime_error(std::string("RFI_Clipper: ") + e.what());
    }
    _bp = readBandPassFile(path, &_kill);
    _bandPass = BandPass::fromFileValues(_bp, _kill);
    _startFrequency = _bandPass.startFrequency();
    _endFrequency = _bandPass.endFrequency();
    ppl_clipper_config cfg{};
    cfg.bp_values = _bp.data();
    cfg.n_values = (int)_bp.size();
    cfg.kill_index = _kill.empty() ? nullptr : _kill.data();
    cfg.n_kill = (int)_kill.size();
    cfg.crFactor = config.hasAttribute("channelRejectionRMS") ? config.getAttribute("channelRejectionRMS").toFloat()
                                                              : 10.0f;
    cfg.srFactor = config.hasAttribute("spectrumRejectionRMS")
                       ? config.getAttribute("spectrumRejectionRMS").toFloat()
                       : 4.0f;
    cfg.maxHistory = config.getOption("History", "maximum", "10").toInt();
    cfg.zeroDM = config.getOption("zeroDMing", "active") == "true" ? 1 : 0;
    cfg.useMeanOverRMS = config.getOption("computeRMSfromMean", "active", "true") == "true" ? 1 : 0;
    cfg.bandMatching = config.getOption("Band", "matching") == "true" ? 1 : 0;
    ppl_check(ppl_clipper_create(&_clip, &cfg), "RFI_Clipper");
}

RFI_Clipper::~RFI_Clipper() { ppl_clipper_destroy(_clip); }

void RFI_Clipper::run(WeightedSpectrumDataSet* weightedStokes)
{
    if (!_active) return;
    SpectrumDataSet<float>* stokes = weightedStokes->dataSet();
    SpectrumDataSet<float>* weights = weightedStokes->weights();
    const unsigned nT = stokes->nTimeBlocks(), nSub = stokes->nSubbands();
    const unsigned nPol = stokes->nPolarisations(), nCh = stokes->nChannels();
    float* d_I = static_cast<float*>(stokes->deviceDataInPlace(nullptr));
    float* d_W = static_cast<float*>(weights->deviceDataForWrite());
    ppl_check(ppl_clipper_run(_clip, d_I, d_W, nT, nSub, nPol, nCh, nullptr), "RFI_Clipper");
    _lastNT = nT;
    float mean = 0.0f, rms = 0.0f;
    ppl_check(ppl_clipper_stats(_clip, &mean, &rms), "RFI_Clipper");
    weightedStokes->setMean(mean);
    weightedStokes->setRMS(rms);
    // the live model: the run map (:184-187) with the device's constant
    // coefficient, mean and RMS after the last spectrum (:562-563)
    float p0 = 0.0f, bpMean = 0.0f, bpRms = 0.0f;
    ppl_check(ppl_clipper_model(_clip, &p0, &bpMean, &bpRms, nullptr, 0), "RFI_Clipper");
    BinMap map(nSub * nCh);
    map.setStart(_startFrequency);
    map.setEnd(_endFrequency);
    _bandPass.setRunState(map, p0, bpMean, bpRms);
}

std::vector<unsigned char> RFI_Clipper::lastFlags() const
{
    std::vector<unsigned char> f(_lastNT);
    if (_clip && _lastNT) ppl_check(ppl_clipper_flags(_clip, f.data(), _lastNT), "RFI_Clipper");
    return f;
}

}  // namespace ampp
}  // namespace pelican
