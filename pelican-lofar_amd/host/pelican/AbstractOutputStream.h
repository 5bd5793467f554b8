// AbstractOutputStream.h -- the Pelican output-stream base class
// (pelican/output/AbstractOutputStream.h as used by
// src/lib/DedispersionDataAnalysisOutput.h:27 and src/lib/TriggerOutput.h):
// a stream is built from its configuration node, and send(streamName, blob)
// hands one data blob to the stream's sendStream.  A pipeline's output
// manager calls send for every blob it publishes under a stream name
// (DedispersionPipeline's "DedispersionDataAnalysis" output).
#pragma once

#include <string>

#include "ConfigNode.h"
#include "DataBlob.h"

namespace pelican {

class AbstractOutputStream {
  public:
    explicit AbstractOutputStream(const ConfigNode& config) : _config(config) {}
    virtual ~AbstractOutputStream() = default;
    AbstractOutputStream(const AbstractOutputStream&) = delete;
    AbstractOutputStream& operator=(const AbstractOutputStream&) = delete;

    void send(const std::string& streamName, const DataBlob* dataBlob) { sendStream(streamName, dataBlob); }
    const ConfigNode& config() const { return _config; }

  protected:
    virtual void sendStream(const std::string& streamName, const DataBlob* dataBlob) = 0;

  private:
    ConfigNode _config;
};

}  // namespace pelican
