#include "ConfigNode.h"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace pelican {

namespace {

std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// Minimal XML reader: elements, attributes (single or double quotes),
// comments, declarations; text content is ignored.
class Parser {
  public:
    explicit Parser(const std::string& s) : _s(s) {}
    std::shared_ptr<XmlElement> parseDocument()
    {
        skipMisc();
        auto e = parseElement();
        return e;
    }

  private:
    const std::string& _s;
    size_t _i = 0;

    [[noreturn]] void fail(const char* what)
    {
        throw std::runtime_error(std::string("ConfigNode: XML ") + what + " at offset " + std::to_string(_i));
    }
    void skipWs()
    {
        while (_i < _s.size() && std::isspace((unsigned char)_s[_i])) ++_i;
    }
    void skipMisc()
    {
        for (;;) {
            skipWs();
            if (_s.compare(_i, 4, "<!--") == 0) {
                size_t e = _s.find("-->", _i + 4);
                if (e == std::string::npos) fail("unterminated comment");
                _i = e + 3;
            } else if (_s.compare(_i, 2, "<?") == 0 || _s.compare(_i, 2, "<!") == 0) {
                size_t e = _s.find('>', _i);
                if (e == std::string::npos) fail("unterminated declaration");
                _i = e + 1;
            } else {
                return;
            }
        }
    }
    std::string name()
    {
        size_t a = _i;
        while (_i < _s.size() && (std::isalnum((unsigned char)_s[_i]) || _s[_i] == '_' || _s[_i] == '-' ||
                                  _s[_i] == ':' || _s[_i] == '.'))
            ++_i;
        if (a == _i) fail("expected a name");
        return _s.substr(a, _i - a);
    }
    std::shared_ptr<XmlElement> parseElement()
    {
        if (_i >= _s.size() || _s[_i] != '<') fail("expected an element");
        ++_i;
        auto e = std::make_shared<XmlElement>();
        e->tag = name();
        for (;;) {
            skipWs();
            if (_i >= _s.size()) fail("unterminated tag");
            if (_s[_i] == '/') {
                if (_s.compare(_i, 2, "/>") != 0) fail("bad empty-element end");
                _i += 2;
                return e;
            }
            if (_s[_i] == '>') {
                ++_i;
                break;
            }
            std::string key = name();
            skipWs();
            if (_i >= _s.size() || _s[_i] != '=') fail("expected '='");
            ++_i;
            skipWs();
            if (_i >= _s.size() || (_s[_i] != '"' && _s[_i] != '\'')) fail("expected a quoted value");
            const char q = _s[_i++];
            size_t end = _s.find(q, _i);
            if (end == std::string::npos) fail("unterminated value");
            e->attrs[key] = _s.substr(_i, end - _i);
            _i = end + 1;
        }
        // content
        for (;;) {
            size_t lt = _s.find('<', _i);
            if (lt == std::string::npos) fail("unterminated element");
            _i = lt;
            if (_s.compare(_i, 4, "<!--") == 0) {
                skipMisc();
                continue;
            }
            if (_s.compare(_i, 2, "</") == 0) {
                _i += 2;
                std::string n = name();
                if (n != e->tag) fail("mismatched closing tag");
                skipWs();
                if (_i >= _s.size() || _s[_i] != '>') fail("expected '>'");
                ++_i;
                return e;
            }
            e->children.push_back(parseElement());
        }
    }
};

}  // namespace

unsigned ConfigValue::toUInt() const { return (unsigned)std::strtoul(trim(_s).c_str(), nullptr, 10); }
unsigned long ConfigValue::toULong() const { return std::strtoul(trim(_s).c_str(), nullptr, 10); }
int ConfigValue::toInt() const { return (int)std::strtol(trim(_s).c_str(), nullptr, 10); }
float ConfigValue::toFloat() const { return std::strtof(trim(_s).c_str(), nullptr); }
double ConfigValue::toDouble() const { return std::strtod(trim(_s).c_str(), nullptr); }
ConfigValue ConfigValue::toLower() const
{
    std::string r = _s;
    std::transform(r.begin(), r.end(), r.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return ConfigValue(r);
}

void ConfigNode::setFromString(const std::string& xml)
{
    Parser p(xml);
    _root = p.parseDocument();
}

ConfigValue ConfigNode::getOption(const std::string& tag, const std::string& attribute,
                                  const std::string& defaultValue) const
{
    if (_root)
        for (const auto& c : _root->children)
            if (c->tag == tag) {
                auto it = c->attrs.find(attribute);
                if (it != c->attrs.end()) return ConfigValue(it->second);
                return ConfigValue(defaultValue);
            }
    return ConfigValue(defaultValue);
}

bool ConfigNode::hasAttribute(const std::string& attribute) const
{
    return _root && _root->attrs.count(attribute);
}

ConfigValue ConfigNode::getAttribute(const std::string& attribute) const
{
    if (!_root) return ConfigValue();
    auto it = _root->attrs.find(attribute);
    return it == _root->attrs.end() ? ConfigValue() : ConfigValue(it->second);
}

ConfigNode ConfigNode::getNode(const std::string& tag) const
{
    ConfigNode n;
    if (_root)
        for (const auto& c : _root->children)
            if (c->tag == tag) {
                n._root = c;
                break;
            }
    return n;
}

std::vector<ConfigNode> ConfigNode::getNodes(const std::string& tag) const
{
    std::vector<ConfigNode> out;
    if (_root)
        for (const auto& c : _root->children)
            if (tag.empty() || c->tag == tag) {
                ConfigNode n;
                n._root = c;
                out.push_back(n);
            }
    return out;
}

std::string ConfigNode::searchFile(const std::string& file) const
{
    auto exists = [](const std::string& f) { return std::ifstream(f).good(); };
    if (!file.empty() && file[0] == '/') {
        if (exists(file)) return file;
        throw std::runtime_error("file not found: " + file);
    }
    if (exists(file)) return file;
    const char* path = std::getenv("PPL_CONFIG_PATH");
    if (path) {
        std::stringstream ss(path);
        std::string dir;
        while (std::getline(ss, dir, ':'))
            if (!dir.empty() && exists(dir + "/" + file)) return dir + "/" + file;
    }
    throw std::runtime_error("file not found: " + file);
}

}  // namespace pelican
