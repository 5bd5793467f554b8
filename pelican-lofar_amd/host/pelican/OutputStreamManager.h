// OutputStreamManager.h -- routes a pipeline's published blobs to output
// streams by stream name (Pelican's OutputStreamManager, configured by the
// <output> element of a pipeline configuration, e.g.
// src/pipelines/udpstream/data/stream1.xml:69-107):
//
//   <output>
//     <dataStreams>
//       <stream name="DedispersionDataAnalysis" listeners="DedispersionDataAnalysisOutput"/>
//       <stream name="SignalFoundSpectrum" listeners="SigprocStokesWriter"/>
//     </dataStreams>
//     <streamers>
//       <DedispersionDataAnalysisOutput> ... </DedispersionDataAnalysisOutput>
//       <SigprocStokesWriter writeHeader="true"> ... </SigprocStokesWriter>
//     </streamers>
//   </output>
//
// Each listener is built once, by its registered class name, from its
// <streamers> section (an empty one if absent); listeners are comma
// separated; a stream named "all" receives every published blob.  A
// streamer section with active="false" is not built and its streams get
// nothing.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "AbstractOutputStream.h"

namespace pelican {

class OutputStreamManager {
  public:
    explicit OutputStreamManager(const ConfigNode& output);
    // Sends the blob to every listener of streamName (and of "all").
    void send(const DataBlob* blob, const std::string& streamName);
    // The listener built for a class name, or null.
    AbstractOutputStream* streamer(const std::string& name) const;
    std::vector<std::string> listeners(const std::string& streamName) const;

  private:
    std::map<std::string, std::unique_ptr<AbstractOutputStream>> _streamers;
    std::map<std::string, std::vector<AbstractOutputStream*>> _routes;
};

}  // namespace pelican
