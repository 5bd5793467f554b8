// AsyncronousModule.h -- base of modules whose results are exported from a
// worker thread (src/lib/AsyncronousModule.h:29-96): connected functors run
// on every exported blob, then the module's exportComplete, the unlock
// callbacks with the blobs whose lock count reached zero, and the
// end-of-chain callbacks.  Exports run on one worker thread in submission
// order.
#pragma once

#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../pelican/DataBlob.h"

namespace pelican {
namespace ampp {

class AsyncronousModule : public AbstractModule {
  public:
    using CallBackT = std::function<void(DataBlob*)>;
    using UnlockCallBackT = std::function<void(const std::vector<DataBlob*>&)>;

    explicit AsyncronousModule(const ConfigNode& config);
    ~AsyncronousModule() override;

    void connect(const CallBackT& functor);
    void unlockCallback(const UnlockCallBackT& callback);
    void onChainCompletion(const std::function<void()>& fn);
    int lockNumber(const DataBlob* data) const;
    // Blocks until every export queued so far has run.
    void waitForJobCompletion() const;

  protected:
    // Queue `data` for export (functors, then completion).
    void exportData(DataBlob* data);
    // Completion without running the functors (a failed job).
    void exportCancel(DataBlob* data);
    virtual void exportComplete(DataBlob* data) = 0;
    void lock(const DataBlob* data);
    void lockUnprotected(const DataBlob* data) { ++_dataLocker[data]; }
    template <class Container>
    void lockAllUnprotected(const Container& data)
    {
        for (const auto* d : data) lockUnprotected(d);
    }
    // Requires lockerMutex held; returns the remaining count.
    int unlock(DataBlob* data);
    // Called from the worker when the queue drains a task; used by
    // subclasses that wait on GPU completion before exporting.
    void enqueue(std::function<void()> task);

    mutable std::mutex lockerMutex;

  private:
    void _exportComplete(DataBlob* data);
    void worker();

    std::map<const DataBlob*, int> _dataLocker;
    std::vector<CallBackT> _linkedFunctors;
    std::vector<UnlockCallBackT> _unlockTriggers;
    std::vector<std::function<void()>> _callbacks;
    std::vector<DataBlob*> _recentUnlocked;

    mutable std::mutex _qMutex;
    mutable std::condition_variable _qCv, _idleCv;
    std::deque<std::function<void()>> _queue;
    int _busy = 0;
    bool _stop = false;
    std::thread _thread;
};

}  // namespace ampp
}  // namespace pelican
