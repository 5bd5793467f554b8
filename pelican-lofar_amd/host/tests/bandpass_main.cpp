// bandpass_main.cpp -- runs BandPassRecorder (and BandPassOutput) on a
// Stokes stream, or its fitting stage alone on an averaged spectrum.
//
//   bandpass_main <config.xml> stream <stokes.f32> <nT> <nSub> <nPol> <nCh> <chunk>
//   bandpass_main <config.xml> spectrum <spectrum.f32> <nBins>
//   bandpass_main - bp <file.bp> [<nBins2> <start2> <width2>]
//   bandpass_main - clip <file.bp> <history> <stokes.f32> <nT> <nSub> <nCh> <chunk>
//
// config.xml: a root element holding <BandPassRecorder> and, optionally,
// <BandPassOutput> sections.  "stream" feeds the float32 [t][sub][pol][ch]
// file in blobs of `chunk` spectra (the last one shorter) to run() until it
// reports a bandpass (the GPU integrates); "spectrum" hands the float32
// averaged spectrum to fitSpectrum (host only).  stdout: one JSON object
// with the blobs used, spectra integrated, fits made, the bandpass RMS,
// median, mean and parameters (%.9g) and, for "spectrum", the cleaned
// spectrum's file (<spectrum.f32>.clean).  "bp" reads a .bp file as the
// clipper does (BandPass::fromFileValues), optionally rebins it, and prints
// the current map's median, RMS, mean, first and last bin, the constant
// coefficient and the map's start and end frequencies (kill lines applied).
// "clip" runs the RFI_Clipper module (zero-DMing, 10 / 4 rms) on the float32
// [t][sub][1][ch] file in blobs of `chunk` spectra, in place, and prints its
// bandPass() afterwards: constant coefficient, mean, RMS, bins, and the
// values to <stokes.f32>.bp; the clipped spectra go to <stokes.f32>.out.
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../lofar/BandPass.h"
#include "../lofar/BandPassOutput.h"
#include "../lofar/BandPassRecorder.h"
#include "../lofar/RFI_Clipper.h"

using namespace pelican;
using namespace pelican::ampp;

static std::string g9(double v)
{
    if (v != v) return "NaN";
    char b[48];
    std::snprintf(b, sizeof b, "%.9g", v);
    return b;
}

static std::vector<float> readFloats(const std::string& path, size_t n)
{
    std::vector<float> v(n);
    std::ifstream f(path, std::ios::binary);
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(float)));
    if (!f) throw std::runtime_error("short file " + path);
    return v;
}

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::cerr << "usage: bandpass_main <config.xml> stream <stokes.f32> <nT> <nSub> <nPol> <nCh> <chunk>\n"
                     "       bandpass_main <config.xml> spectrum <spectrum.f32> <nBins>\n";
        return 2;
    }
    try {
        if (std::string(argv[2]) == "bp") {
            if (argc != 4 && argc != 7) throw std::runtime_error("bp needs a file and optionally a map");
            std::vector<unsigned> kills;
            const std::vector<double> v = readBandPassFile(argv[3], &kills);
            BandPass b = BandPass::fromFileValues(v, kills);
            if (argc == 7) {
                BinMap m((unsigned)std::atoi(argv[4]));
                m.setStart(std::atof(argv[5]));
                m.setBinWidth(std::atof(argv[6]));
                b.reBin(m);
            }
            const size_t n = b.currentSet().size();
            std::cout << "{\"median\": " << g9(b.median()) << ", \"rms\": " << g9(b.rms())
                      << ", \"mean\": " << g9(b.mean()) << ", \"first\": " << g9(b.intensityOfBin(0))
                      << ", \"last\": " << g9(b.intensityOfBin((unsigned)n - 1)) << ", \"p0\": "
                      << g9(b.params()[0]) << ", \"start\": " << g9(b.startFrequency())
                      << ", \"end\": " << g9(b.endFrequency()) << ", \"bins\": " << n << "}" << std::endl;
            return 0;
        }
        if (std::string(argv[2]) == "clip") {
            if (argc != 10) throw std::runtime_error("clip needs a file, history, stokes file, nT, nSub, nCh, chunk");
            const std::string xml = std::string("<RFI_Clipper active=\"true\" channelRejectionRMS=\"10\" "
                                                "spectrumRejectionRMS=\"4\"><BandPassData file=\"") +
                                    argv[3] + "\"/><History maximum=\"" + argv[4] +
                                    "\"/><zeroDMing active=\"true\"/><Band matching=\"true\"/></RFI_Clipper>";
            RFI_Clipper clipper{ConfigNode(xml)};
            const unsigned nT = (unsigned)std::atoi(argv[6]), nSub = (unsigned)std::atoi(argv[7]);
            const unsigned nCh = (unsigned)std::atoi(argv[8]), chunk = (unsigned)std::atoi(argv[9]);
            if (!chunk) throw std::runtime_error("chunk must be positive");
            const size_t row = (size_t)nSub * nCh;
            std::vector<float> all = readFloats(argv[5], row * nT);
            SpectrumDataSetStokes stokes;
            for (unsigned t0 = 0; t0 < nT; t0 += chunk) {
                const unsigned n = std::min(chunk, nT - t0);
                stokes.resize(n, nSub, 1, nCh);
                std::copy(all.begin() + (long)(row * t0), all.begin() + (long)(row * (t0 + n)), stokes.data());
                WeightedSpectrumDataSet ws(&stokes);
                clipper.run(&ws);
                const float* out = stokes.data();
                std::copy(out, out + row * n, all.begin() + (long)(row * t0));
            }
            const BandPass& b = clipper.bandPass();
            const std::vector<float>& vals = b.currentSet();
            std::ofstream(std::string(argv[5]) + ".bp", std::ios::binary)
                .write(reinterpret_cast<const char*>(vals.data()), (std::streamsize)(vals.size() * sizeof(float)));
            std::ofstream(std::string(argv[5]) + ".out", std::ios::binary)
                .write(reinterpret_cast<const char*>(all.data()), (std::streamsize)(all.size() * sizeof(float)));
            std::cout << "{\"p0\": " << g9(b.params()[0]) << ", \"mean\": " << g9(b.mean())
                      << ", \"rms\": " << g9(b.rms()) << ", \"bins\": " << vals.size() << "}" << std::endl;
            return 0;
        }
        std::ifstream xf(argv[1]);
        std::stringstream xs;
        xs << xf.rdbuf();
        ConfigNode config(xs.str());
        BandPassRecorder recorder(config.getNode("BandPassRecorder"));
        std::unique_ptr<BandPassOutput> output;
        if (!config.getNode("BandPassOutput").tag().empty())
            output.reset(new BandPassOutput(config.getNode("BandPassOutput")));
        BandPass bp;
        const std::string mode = argv[2];
        int blobs = 0, fits = 0;
        unsigned long spectra = 0;
        bool done = false;
        if (mode == "stream" && argc == 9) {
            const unsigned nT = (unsigned)std::atoi(argv[4]), nSub = (unsigned)std::atoi(argv[5]);
            const unsigned nPol = (unsigned)std::atoi(argv[6]), nCh = (unsigned)std::atoi(argv[7]);
            const unsigned chunk = (unsigned)std::atoi(argv[8]);
            if (!chunk) throw std::runtime_error("chunk must be positive");
            const size_t row = (size_t)nSub * nPol * nCh;
            const std::vector<float> all = readFloats(argv[3], row * nT);
            SpectrumDataSetStokes stokes;
            for (unsigned t0 = 0; t0 < nT && !done; t0 += chunk) {
                const unsigned n = std::min(chunk, nT - t0);
                stokes.resize(n, nSub, nPol, nCh);
                std::copy(all.begin() + (long)(row * t0), all.begin() + (long)(row * (t0 + n)), stokes.data());
                ++blobs;
                spectra += n;
                done = recorder.run(&stokes, &bp);
            }
            fits = recorder.fitsMade();
        } else if (mode == "spectrum" && argc == 5) {
            const unsigned nBins = (unsigned)std::atoi(argv[4]);
            std::vector<float> s = readFloats(argv[3], nBins);
            fits = recorder.fitSpectrum(s, recorder.binMap(nBins), &bp);
            done = true;
            std::ofstream clean(std::string(argv[3]) + ".clean", std::ios::binary);
            clean.write(reinterpret_cast<const char*>(s.data()), (std::streamsize)(s.size() * sizeof(float)));
        } else {
            throw std::runtime_error("bad arguments for mode " + mode);
        }
        if (done && output) output->send("BandPass", &bp);
        std::cout << "{\"done\": " << (done ? "true" : "false") << ", \"blobs\": " << blobs
                  << ", \"spectra\": " << spectra << ", \"fits\": " << fits << ", \"rms\": " << g9(bp.primaryRms())
                  << ", \"median\": " << g9(bp.primaryMedian()) << ", \"mean\": " << g9(bp.mean())
                  << ", \"params\": [";
        for (size_t i = 0; i < bp.params().size(); ++i) std::cout << (i ? ", " : "") << g9(bp.params()[i]);
        std::cout << "]}" << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "bandpass_main: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
