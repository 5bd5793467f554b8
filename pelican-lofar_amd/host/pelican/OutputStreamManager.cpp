#include "OutputStreamManager.h"

#include <sstream>

namespace pelican {

namespace {
std::string trim(const std::string& s)
{
    const size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}
}  // namespace

OutputStreamManager::OutputStreamManager(const ConfigNode& output)
{
    const ConfigNode streamers = output.getNode("streamers");
    for (const ConfigNode& stream : output.getNode("dataStreams").getNodes("stream")) {
        const std::string name = stream.getAttribute("name").str();
        std::stringstream ls(stream.getAttribute("listeners").str());
        std::string listener;
        while (std::getline(ls, listener, ',')) {
            listener = trim(listener);
            if (listener.empty()) continue;
            const ConfigNode section = streamers.getNode(listener);
            if (section.getAttribute("active").toLower() == "false") continue;
            auto it = _streamers.find(listener);
            if (it == _streamers.end())
                it = _streamers
                         .emplace(listener, std::unique_ptr<AbstractOutputStream>(
                                                OutputStreamFactory::instance().create(listener, section)))
                         .first;
            _routes[name].push_back(it->second.get());
        }
    }
}

void OutputStreamManager::send(const DataBlob* blob, const std::string& streamName)
{
    for (const std::string& key : {streamName, std::string("all")}) {
        auto it = _routes.find(key);
        if (it == _routes.end()) continue;
        for (AbstractOutputStream* s : it->second) s->send(streamName, blob);
        if (streamName == "all") break;
    }
}

AbstractOutputStream* OutputStreamManager::streamer(const std::string& name) const
{
    auto it = _streamers.find(name);
    return it == _streamers.end() ? nullptr : it->second.get();
}

std::vector<std::string> OutputStreamManager::listeners(const std::string& streamName) const
{
    std::vector<std::string> out;
    auto it = _routes.find(streamName);
    if (it == _routes.end()) return out;
    for (AbstractOutputStream* s : it->second)
        for (const auto& kv : _streamers)
            if (kv.second.get() == s) out.push_back(kv.first);
    return out;
}

}  // namespace pelican
