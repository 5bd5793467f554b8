// AbstractOutputStream.h -- the Pelican output-stream base class
// (pelican/output/AbstractOutputStream.h as used by
// src/lib/DedispersionDataAnalysisOutput.h:27 and src/lib/TriggerOutput.h):
// a stream is built from its configuration node, and send(streamName, blob)
// hands one data blob to the stream's sendStream.  A pipeline's output
// manager calls send for every blob it publishes under a stream name
// (DedispersionPipeline's "DedispersionDataAnalysis" output).
#pragma once

#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <type_traits>

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

// Name -> constructor registry of output streams (PELICAN_DECLARE).
class OutputStreamFactory {
  public:
    using Maker = std::function<AbstractOutputStream*(const ConfigNode&)>;
    static OutputStreamFactory& instance()
    {
        static OutputStreamFactory f;
        return f;
    }
    bool add(const std::string& name, Maker m)
    {
        _makers[name] = std::move(m);
        return true;
    }
    AbstractOutputStream* create(const std::string& name, const ConfigNode& config) const
    {
        auto it = _makers.find(name);
        if (it == _makers.end()) throw std::runtime_error("OutputStreamFactory: unknown output stream " + name);
        return it->second(config);
    }

  private:
    std::map<std::string, Maker> _makers;
};

}  // namespace pelican

// PELICAN_DECLARE(AbstractOutputStream, TYPE): registers an output stream by
// its class name, as the reference's writers declare themselves.
#define PELICAN_DECLARE(BASE, TYPE)                                                                        \
    static_assert(std::is_same<BASE, ::pelican::AbstractOutputStream>::value,                               \
                  "only output streams are declared this way");                                              \
    static const bool pelican_output_registered_##TYPE = ::pelican::OutputStreamFactory::instance().add(      \
        #TYPE, [](const ::pelican::ConfigNode& c) -> ::pelican::AbstractOutputStream* { return new TYPE(c); });
