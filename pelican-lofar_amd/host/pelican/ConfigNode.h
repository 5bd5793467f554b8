// ConfigNode.h -- the part of Pelican's configuration interface the
// pulsar-search modules use (ConfigNode::getOption / hasAttribute /
// getAttribute / setFromString / searchFile), over a small XML reader.
// Values are strings with the Qt-style conversions the modules call
// (toUInt, toULong, toInt, toFloat, toDouble, toLower).
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace pelican {

// A configuration value (the QString the reference's getOption returns).
class ConfigValue {
  public:
    ConfigValue() = default;
    ConfigValue(std::string s) : _s(std::move(s)) {}
    const std::string& str() const { return _s; }
    unsigned toUInt() const;
    unsigned long toULong() const;
    int toInt() const;
    float toFloat() const;
    double toDouble() const;
    ConfigValue toLower() const;
    bool isEmpty() const { return _s.empty(); }
    bool operator==(const char* o) const { return _s == o; }
    bool operator==(const std::string& o) const { return _s == o; }
    bool operator!=(const char* o) const { return _s != o; }
    operator const std::string&() const { return _s; }

  private:
    std::string _s;
};

struct XmlElement {
    std::string tag;
    std::map<std::string, std::string> attrs;
    std::vector<std::shared_ptr<XmlElement>> children;
};

class ConfigNode {
  public:
    ConfigNode() = default;
    explicit ConfigNode(const std::string& xml) { setFromString(xml); }

    // Parses one element (with children) from XML text; throws std::runtime_error.
    void setFromString(const std::string& xml);

    // Attribute of the first child element `tag` (or defaultValue).
    ConfigValue getOption(const std::string& tag, const std::string& attribute,
                          const std::string& defaultValue = "") const;
    bool hasAttribute(const std::string& attribute) const;
    ConfigValue getAttribute(const std::string& attribute) const;
    // A child element as its own node (e.g. a module's section of a pipeline file).
    ConfigNode getNode(const std::string& tag) const;
    // Every child element `tag` in document order (every child if tag is empty).
    std::vector<ConfigNode> getNodes(const std::string& tag = "") const;
    std::string tag() const { return _root ? _root->tag : std::string(); }
    // Resolves a file name: absolute, or relative to the directories listed
    // in the PPL_CONFIG_PATH environment variable; throws if missing.
    std::string searchFile(const std::string& file) const;

  private:
    std::shared_ptr<XmlElement> _root;
};

}  // namespace pelican
