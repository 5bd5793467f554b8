// chain_main.cpp -- runs voltages through the drop-in module chain
// (DedispersionPipeline over the Pelican-style modules) and writes every
// analysed buffer's events.
//
//   chain_main <pipeline.xml> <voltages.i16> <nSub> <nSamples> <chunk> <sampleTime> <events.txt> [outputs.xml]
//
// voltages: complex int16 [sub][pol=2][t][re,im], nSamples per (sub, pol);
// each chunk of `chunk` samples per (sub, pol) becomes one
// TimeSeriesDataSetC32 of chunk/nCh blocks.  Output: one line per buffer
// "buffer <k> <n>" followed by n lines "dm t binFactor value" (the dummy
// event first), then "pool <free> <history>".  Chunk k carries the time
// stamp k * chunk * sampleTime.  outputs.xml, if given, is a Pelican <output>
// element (dataStreams + streamers): the pipeline publishes its
// "TriggerInput", "DedispersionDataAnalysis" and "SignalFoundSpectrum"
// streams to the writers it names.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../lofar/BandPassOutput.h"
#include "../lofar/DedispersionDataAnalysisOutput.h"
#include "../lofar/DedispersionPipeline.h"
#include "../lofar/SigprocStokesWriter.h"
#include "../lofar/TriggerOutput.h"

using namespace pelican;
using namespace pelican::ampp;

int main(int argc, char** argv)
{
    if (argc != 8 && argc != 9) {
        std::cerr << "usage: chain_main <pipeline.xml> <voltages.i16> <nSub> <nSamples> <chunk> <sampleTime> "
                     "<events.txt> [outputs.xml]\n";
        return 2;
    }
    try {
        std::ifstream xf(argv[1]);
        std::stringstream xs;
        xs << xf.rdbuf();
        ConfigNode config(xs.str());
        const unsigned nSub = (unsigned)std::atoi(argv[3]), nSamp = (unsigned)std::atoi(argv[4]);
        const unsigned chunk = (unsigned)std::atoi(argv[5]);
        const double sampleTime = std::atof(argv[6]);
        const unsigned nCh = config.getNode("PPFChanneliser").getOption("outputChannelsPerSubband", "value", "512").toUInt();
        const unsigned history = config.getNode("DedispersionPipeline").getOption("history", "value", "10").toUInt();
        std::vector<int16_t> v((size_t)nSub * 2 * nSamp * 2);
        std::ifstream vf(argv[2], std::ios::binary);
        vf.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(int16_t)));
        if (!vf) throw std::runtime_error("short voltage file");

        std::unique_ptr<OutputStreamManager> outputs;
        if (argc == 9) {
            std::ifstream of(argv[8]);
            std::stringstream os;
            os << of.rdbuf();
            outputs.reset(new OutputStreamManager(ConfigNode(os.str())));
        }

        std::ofstream out(argv[7]);
        int buffers = 0;
        DedispersionPipeline pipe(config);
        pipe.setOutputStreams(outputs.get());
        pipe.setEventCallback([&](const DedispersionDataAnalysis& r) {
            out << "buffer " << buffers++ << " " << r.eventsFound() << "\n";
            for (const auto& e : r.events()) {
                char v[32];
                std::snprintf(v, sizeof v, "%.9g", e.mfValue());
                out << e.dmIndex() << " " << e.timeBin() << " " << (int)e.mfBinning() << " " << v << "\n";
            }
        });
        TimeSeriesDataSetC32 ts;
        const unsigned nBlk = chunk / nCh;
        for (unsigned c0 = 0; c0 + chunk <= nSamp; c0 += chunk) {
            ts.resize(nBlk, nSub, 2, nCh);
            ts.setBlockRate(sampleTime);
            ts.setLofarTimestamp((double)c0 * sampleTime);
            std::complex<float>* d = ts.data();
            for (unsigned s = 0; s < nSub; ++s)
                for (unsigned p = 0; p < 2; ++p) {
                    const int16_t* src = &v[(((size_t)s * 2 + p) * nSamp + c0) * 2];
                    std::complex<float>* dst = d + TimeSeriesDataSetC32::index(s, nCh, p, 2, 0, nBlk);
                    for (unsigned t = 0; t < chunk; ++t) dst[t] = std::complex<float>(src[2 * t], src[2 * t + 1]);
                }
            pipe.run(&ts);
        }
        pipe.waitForCompletion();
        out << "pool " << pipe.freeStokesBlobs() << " " << history << "\n";
        std::cout << "{\"buffers\": " << pipe.buffersAnalysed() << "}" << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "chain_main: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
