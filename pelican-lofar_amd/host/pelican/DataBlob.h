// DataBlob.h / AbstractModule.h equivalents: the Pelican base classes the
// pulsar-search modules derive from, and the declaration macros that register
// them by name (a pipeline creates modules and blobs from their type names).
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>

#include "ConfigNode.h"

namespace pelican {

class DataBlob {
  public:
    explicit DataBlob(const std::string& type) : _type(type) {}
    virtual ~DataBlob() = default;
    const std::string& type() const { return _type; }

  private:
    std::string _type;
};

class AbstractModule {
  public:
    explicit AbstractModule(const ConfigNode& config) : _config(config) {}
    virtual ~AbstractModule() = default;
    const ConfigNode& config() const { return _config; }

  private:
    ConfigNode _config;
};

// Name -> constructor registries (PELICAN_DECLARE_MODULE / _DATABLOB).
class ModuleFactory {
  public:
    using Maker = std::function<AbstractModule*(const ConfigNode&)>;
    static ModuleFactory& instance()
    {
        static ModuleFactory f;
        return f;
    }
    bool add(const std::string& name, Maker m)
    {
        _makers[name] = std::move(m);
        return true;
    }
    AbstractModule* create(const std::string& name, const ConfigNode& config) const
    {
        auto it = _makers.find(name);
        if (it == _makers.end()) throw std::runtime_error("ModuleFactory: unknown module " + name);
        return it->second(config);
    }

  private:
    std::map<std::string, Maker> _makers;
};

class DataBlobFactory {
  public:
    using Maker = std::function<DataBlob*()>;
    static DataBlobFactory& instance()
    {
        static DataBlobFactory f;
        return f;
    }
    bool add(const std::string& name, Maker m)
    {
        _makers[name] = std::move(m);
        return true;
    }
    DataBlob* create(const std::string& name) const
    {
        auto it = _makers.find(name);
        if (it == _makers.end()) throw std::runtime_error("DataBlobFactory: unknown blob " + name);
        return it->second();
    }

  private:
    std::map<std::string, Maker> _makers;
};

}  // namespace pelican

#define PELICAN_DECLARE_MODULE(TYPE)                                                                        \
    static const bool pelican_module_registered_##TYPE = ::pelican::ModuleFactory::instance().add(          \
        #TYPE, [](const ::pelican::ConfigNode& c) -> ::pelican::AbstractModule* { return new TYPE(c); });

#define PELICAN_DECLARE_DATABLOB(TYPE)                                                                      \
    static const bool pelican_blob_registered_##TYPE = ::pelican::DataBlobFactory::instance().add(          \
        #TYPE, []() -> ::pelican::DataBlob* { return new TYPE(); });
