// blob_main.cpp -- drives the host blobs' accessors from a small script, so
// the tests can check DedispersionSpectra and BinMap against the reference's
// known answers on a host without a GPU (no device copy is ever made).
//
//   blob_main < script
//
// Script lines (whitespace separated); every query prints one line:
//   resize <timeBins> <dmBins> <dmStart> <dmWidth>   DedispersionSpectra::resize
//   set <index> <value>                               data()[index] = value
//   size | dmMax | dmBins | timeSamples               accessors
//   dmIndex <dm> | dm <i> | ampf <t> <dm> | ampi <t> <i>
//   map <nBins> <start> <width>                       a BinMap
//   mapIndex <v> | mapNumber <i> | mapStart <i> | mapEnd <i>
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>

#include "../lofar/DataSets.h"

using namespace pelican::ampp;

int main()
{
    DedispersionSpectra spectra;
    BinMap map;
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string cmd;
        if (!(in >> cmd)) continue;
        if (cmd == "resize") {
            unsigned t, d;
            float start, width;
            in >> t >> d >> start >> width;
            spectra.resize(t, d, start, width);
        } else if (cmd == "set") {
            size_t i;
            float v;
            in >> i >> v;
            spectra.data().at(i) = v;
        } else if (cmd == "size") {
            std::printf("%zu\n", spectra.data().size());
        } else if (cmd == "dmMax") {
            std::printf("%.9g\n", (double)spectra.dmMax());
        } else if (cmd == "dmBins") {
            std::printf("%d\n", spectra.dmBins());
        } else if (cmd == "timeSamples") {
            std::printf("%d\n", spectra.timeSamples());
        } else if (cmd == "dmIndex") {
            float dm;
            in >> dm;
            std::printf("%d\n", spectra.dmIndex(dm));
        } else if (cmd == "dm") {
            unsigned i;
            in >> i;
            std::printf("%.9g\n", (double)spectra.dm(i));
        } else if (cmd == "ampf") {
            unsigned t;
            float dm;
            in >> t >> dm;
            std::printf("%.9g\n", (double)spectra.dmAmplitude(t, dm));
        } else if (cmd == "ampi") {
            unsigned t;
            int i;
            in >> t >> i;
            std::printf("%.9g\n", (double)spectra.dmAmplitude(t, i));
        } else if (cmd == "map") {
            unsigned n;
            double start, width;
            in >> n >> start >> width;
            map.reset(n);
            map.setStart(start);
            map.setBinWidth(width);
        } else if (cmd == "mapIndex") {
            double v;
            in >> v;
            std::printf("%d\n", map.binIndex(v));
        } else if (cmd == "mapNumber" || cmd == "mapStart" || cmd == "mapEnd") {
            unsigned i;
            in >> i;
            const double v = cmd == "mapNumber" ? map.binAssignmentNumber(i)
                           : cmd == "mapStart"  ? map.binStart(i)
                                                : map.binEnd(i);
            std::printf("%.17g\n", v);
        } else {
            std::cerr << "unknown command: " << cmd << "\n";
            return 2;
        }
    }
    return 0;
}
