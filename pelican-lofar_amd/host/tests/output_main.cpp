// output_main.cpp -- drives the two event writers (DedispersionDataAnalysisOutput,
// TriggerOutput) from a small script, on the host only.
//
//   output_main <outputs.xml>  < script
//
// outputs.xml: a root element holding <DedispersionDataAnalysisOutput> and/or
// <TriggerOutput> sections in the writers' own configuration format.
// Script lines (whitespace separated):
//   spectra <timeBins> <dmBins> <dmStart> <dmStep> <startTime> <blockRate> <firstSample>
//   rms <rms>
//   event <dmIndex> <timeIndex> <binFactor> <value>
//   send                 hand the analysis built so far to every writer, then
//                        start a new one (time-stamp event first, as the
//                        analyser does)
// stdout: "dat <file>" for the analysis writer's file, and the trigger
// writer's "Datagram written" for each message it sends.
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>

#include "../lofar/DataSets.h"
#include "../lofar/DedispersionDataAnalysisOutput.h"
#include "../lofar/TriggerOutput.h"

using namespace pelican;
using namespace pelican::ampp;

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::cerr << "usage: output_main <outputs.xml> < script\n";
        return 2;
    }
    try {
        std::ifstream xf(argv[1]);
        std::stringstream xs;
        xs << xf.rdbuf();
        ConfigNode config(xs.str());
        std::unique_ptr<DedispersionDataAnalysisOutput> dat;
        std::unique_ptr<TriggerOutput> trig;
        if (!config.getNode("DedispersionDataAnalysisOutput").tag().empty()) {
            dat.reset(new DedispersionDataAnalysisOutput(config.getNode("DedispersionDataAnalysisOutput")));
            std::cout << "dat " << dat->fileName() << std::endl;
        }
        if (!config.getNode("TriggerOutput").tag().empty())
            trig.reset(new TriggerOutput(config.getNode("TriggerOutput")));

        SpectrumDataSetStokes stokes;
        DedispersionSpectra spectra;
        spectra.setInputDataBlobs({&stokes});
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
                stokes.setLofarTimestamp(t0);
                stokes.setBlockRate(rate);
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
                if (dat) dat->send("DedispersionDataAnalysis", &analysis);
                if (trig) trig->send("DedispersionDataAnalysis", &analysis);
                const float rms = analysis.getRMS();
                analysis.reset(&spectra);
                analysis.setRMS(rms);
                analysis.addEvent(0, 0, 1, 0.0f);
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
