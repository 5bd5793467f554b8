// output_main.cpp -- drives the output streams (DedispersionDataAnalysisOutput,
// TriggerOutput, SigprocStokesWriter, ...) through an OutputStreamManager from
// a small script.
//
//   output_main <output.xml>  < script
//
// output.xml: a Pelican <output> element (dataStreams + streamers).
// Script lines (whitespace separated):
//   spectra <timeBins> <dmBins> <dmStart> <dmStep> <startTime> <blockRate> <firstSample>
//   rms <rms>
//   event <dmIndex> <timeIndex> <binFactor> <value>
//   send <stream>        publish the analysis built so far on <stream>
//   clear                start a new analysis (time-stamp event first, as
//                        the analyser does)
//   stokes <file.f32> <nT> <nSub> <nPol> <nCh> <timeStamp> <blockRate>
//                        load a Stokes blob [t][sub][pol][ch] (nT may be 0:
//                        no file is read)
//   sendstokes <stream>  publish the Stokes blob on <stream>
// The Stokes blob's data reach the device only when a writer asks for them.
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>

#include "../lofar/DataSets.h"
#include "../lofar/DedispersionDataAnalysisOutput.h"
#include "../lofar/SigprocStokesWriter.h"
#include "../lofar/TriggerOutput.h"
#include "../pelican/OutputStreamManager.h"

using namespace pelican;
using namespace pelican::ampp;

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::cerr << "usage: output_main <output.xml> < script\n";
        return 2;
    }
    try {
        std::ifstream xf(argv[1]);
        std::stringstream xs;
        xs << xf.rdbuf();
        OutputStreamManager outputs{ConfigNode(xs.str())};
        if (auto* dat = dynamic_cast<DedispersionDataAnalysisOutput*>(outputs.streamer("DedispersionDataAnalysisOutput")))
            std::cout << "dat " << dat->fileName() << std::endl;
        if (auto* fil = dynamic_cast<SigprocStokesWriter*>(outputs.streamer("SigprocStokesWriter")))
            std::cout << "fil " << fil->fileName() << std::endl;

        SpectrumDataSetStokes stokes, timing;
        DedispersionSpectra spectra;
        spectra.setInputDataBlobs({&timing});
        DedispersionDataAnalysis analysis;
        analysis.reset(&spectra);
        analysis.addEvent(0, 0, 1, 0.0f);

        std::string line;
        while (std::getline(std::cin, line)) {
            std::istringstream ls(line);
            std::string op;
            if (!(ls >> op)) continue;
            if (op == "spectra") {
                unsigned tb, nd, first;
                double start, width, t0, rate;
                ls >> tb >> nd >> start >> width >> t0 >> rate >> first;
                if (!ls) throw std::runtime_error("bad spectra line: " + line);
                spectra.resize(tb, nd, (float)start, (float)width);
                spectra.setFirstSample(first);
                timing.setLofarTimestamp(t0);
                timing.setBlockRate(rate);
            } else if (op == "rms") {
                double r;
                ls >> r;
                analysis.setRMS((float)r);
            } else if (op == "event") {
                unsigned dm, t;
                double bf, v;
                ls >> dm >> t >> bf >> v;
                if (!ls) throw std::runtime_error("bad event line: " + line);
                analysis.addEvent(dm, t, (float)bf, (float)v);
            } else if (op == "send") {
                std::string stream;
                if (!(ls >> stream)) throw std::runtime_error("send needs a stream name");
                outputs.send(&analysis, stream);
            } else if (op == "clear") {
                const float rms = analysis.getRMS();
                analysis.reset(&spectra);
                analysis.setRMS(rms);
                analysis.addEvent(0, 0, 1, 0.0f);
            } else if (op == "stokes") {
                std::string file;
                unsigned nT, nSub, nPol, nCh;
                double t0, rate;
                ls >> file >> nT >> nSub >> nPol >> nCh >> t0 >> rate;
                if (!ls) throw std::runtime_error("bad stokes line: " + line);
                stokes.resize(nT, nSub, nPol, nCh);
                stokes.setLofarTimestamp(t0);
                stokes.setBlockRate(rate);
                if (nT) {
                    std::ifstream f(file, std::ios::binary);
                    f.read(reinterpret_cast<char*>(stokes.data()), (std::streamsize)stokes.bytes());
                    if (!f) throw std::runtime_error("short Stokes file " + file);
                }
            } else if (op == "sendstokes") {
                std::string stream;
                if (!(ls >> stream)) throw std::runtime_error("sendstokes needs a stream name");
                outputs.send(&stokes, stream);
            } else {
                throw std::runtime_error("unknown script line: " + line);
            }
        }
    } catch (const std::exception& e) {
        std::cerr << "output_main: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
