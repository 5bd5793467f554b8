// DeviceMirror.h -- the device-resident copy of a data blob.
//
// Blobs keep their host vector (the reference's storage) plus a device
// buffer.  A module reads its input with deviceRead (uploads only when the
// host copy is newer) and writes its output with deviceWrite (the device copy
// becomes the newer one); host accessors call hostAccess, which downloads
// only when the device copy is newer.  Chained modules therefore hand blobs
// over in HBM without round trips through host memory.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>

#include "../../../include/ppl_amd.h"

namespace pelican {
namespace ampp {

inline void ppl_check(int rc, const char* what)
{
    if (rc != PPL_OK) throw std::runtime_error(std::string(what) + ": " + ppl_last_error());
}

class DeviceMirror {
  public:
    DeviceMirror() = default;
    DeviceMirror(const DeviceMirror&) = delete;
    DeviceMirror& operator=(const DeviceMirror&) = delete;
    ~DeviceMirror() { ppl_device_free(_d); }

    // Device pointer with the blob's current contents.
    void* deviceRead(const void* host, size_t bytes, void* stream)
    {
        reserve(bytes);
        if (!_deviceValid) {
            if (bytes) ppl_check(ppl_memcpy_h2d(_d, host, bytes, stream), "upload");
            _deviceValid = true;
        }
        return _d;
    }
    // Device pointer a kernel will overwrite; afterwards the device copy is newest.
    void* deviceWrite(size_t bytes)
    {
        reserve(bytes);
        _deviceValid = true;
        _deviceNewer = true;
        return _d;
    }
    // Device pointer whose contents a kernel updates in place.
    void* deviceReadWrite(const void* host, size_t bytes, void* stream)
    {
        void* p = deviceRead(host, bytes, stream);
        _deviceNewer = true;
        return p;
    }
    // Host access: download if the device copy is newer.  A writable access
    // makes the host copy the newest.
    void hostAccess(void* host, size_t bytes, bool writable) const
    {
        if (_deviceNewer) {
            if (bytes) {
                ppl_check(ppl_memcpy_d2h(host, _d, bytes, nullptr), "download");
                ppl_check(ppl_stream_sync(nullptr), "download");
            }
            _deviceNewer = false;
        }
        if (writable) _deviceValid = false;
    }
    void invalidate()
    {
        _deviceValid = false;
        _deviceNewer = false;
    }
    bool deviceNewer() const { return _deviceNewer; }

  private:
    void reserve(size_t bytes)
    {
        if (bytes > _cap) {
            ppl_device_free(_d);
            _d = nullptr;
            ppl_check(ppl_device_alloc(&_d, bytes), "device allocation");
            _cap = bytes;
            _deviceValid = false;
        }
    }
    void* _d = nullptr;
    size_t _cap = 0;
    mutable bool _deviceValid = false;
    mutable bool _deviceNewer = false;
};

}  // namespace ampp
}  // namespace pelican
