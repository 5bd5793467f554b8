#include "AsyncronousModule.h"

#include "../../../include/ppl_amd.h"
#include "DeviceMirror.h"

namespace pelican {
namespace ampp {

AsyncronousModule::AsyncronousModule(const ConfigNode& config) : AbstractModule(config)
{
    // GPU_NVidia::initialiseResources (AsyncronousModule.cpp:31-33): device 0
    ppl_check(ppl_init(0), "AsyncronousModule: GPU initialisation");
    _thread = std::thread([this] { worker(); });
}

AsyncronousModule::~AsyncronousModule()
{
    {
        std::lock_guard<std::mutex> l(_qMutex);
        _stop = true;
    }
    _qCv.notify_all();
    if (_thread.joinable()) _thread.join();
}

void AsyncronousModule::connect(const CallBackT& functor) { _linkedFunctors.push_back(functor); }

void AsyncronousModule::unlockCallback(const UnlockCallBackT& callback) { _unlockTriggers.push_back(callback); }

void AsyncronousModule::onChainCompletion(const std::function<void()>& fn) { _callbacks.push_back(fn); }

int AsyncronousModule::lockNumber(const DataBlob* data) const
{
    std::lock_guard<std::mutex> l(lockerMutex);
    auto it = _dataLocker.find(data);
    return it == _dataLocker.end() ? 0 : it->second;
}

void AsyncronousModule::lock(const DataBlob* data)
{
    std::lock_guard<std::mutex> l(lockerMutex);
    ++_dataLocker[data];
}

int AsyncronousModule::unlock(DataBlob* data)
{
    auto it = _dataLocker.find(data);
    if (it == _dataLocker.end() || it->second <= 0) return 0;
    if (--it->second == 0) _recentUnlocked.push_back(data);
    return it->second;
}

void AsyncronousModule::enqueue(std::function<void()> task)
{
    {
        std::lock_guard<std::mutex> l(_qMutex);
        _queue.push_back(std::move(task));
    }
    _qCv.notify_one();
}

void AsyncronousModule::exportData(DataBlob* data)
{
    enqueue([this, data] {
        for (auto& f : _linkedFunctors) f(data);
        _exportComplete(data);
        for (auto& c : _callbacks) c();
    });
}

void AsyncronousModule::exportCancel(DataBlob* data) { _exportComplete(data); }

void AsyncronousModule::_exportComplete(DataBlob* blob)
{
    std::vector<DataBlob*> unlocked;
    {
        std::lock_guard<std::mutex> l(lockerMutex);
        exportComplete(blob);
        unlocked.swap(_recentUnlocked);
    }
    for (auto& f : _unlockTriggers) f(unlocked);
}

void AsyncronousModule::waitForJobCompletion() const
{
    std::unique_lock<std::mutex> l(_qMutex);
    _idleCv.wait(l, [this] { return _queue.empty() && _busy == 0; });
}

void AsyncronousModule::worker()
{
    for (;;) {
        std::function<void()> task;
        {
            std::unique_lock<std::mutex> l(_qMutex);
            _qCv.wait(l, [this] { return _stop || !_queue.empty(); });
            if (_queue.empty()) return;  // stopping with nothing left
            task = std::move(_queue.front());
            _queue.pop_front();
            ++_busy;
        }
        task();
        {
            std::lock_guard<std::mutex> l(_qMutex);
            --_busy;
        }
        _idleCv.notify_all();
    }
}

}  // namespace ampp
}  // namespace pelican
