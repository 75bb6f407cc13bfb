// sicc_common.hpp -- shared host-side plumbing for the MI355X receive-and-decode
// library: HIP error checking, the library-owned last-error string and the
// opaque-handle registry behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <vector>
#include <stdexcept>
#include <string>
#include <unordered_map>

namespace sicc {

// A failed HIP call is reported through the C ABI's status code and message;
// the runtime's per-thread last error is cleared here so that the failure does
// not surface again in the caller's own later hipGetLastError checks (torch
// checks it after every kernel launch).
[[noreturn]] inline void throw_hip(hipError_t e, const char* what)
{
    (void)hipGetLastError();
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

#define SICC_HIP(call)                                      \
    do {                                                    \
        hipError_t sicc_e_ = (call);                        \
        if (sicc_e_ != hipSuccess) ::sicc::throw_hip(sicc_e_, #call); \
    } while (0)

// Last error message, as cuda_osd_get_last_error() reports it
// (source/cpp_cuda/portable/usrp_video_link_portable_20260513/third_party/cuda_osd/include/matlab_decoder_c_api.h:134).
// The string is owned by the library and stays valid until the next failing call
// on the same thread.
void set_last_error(const std::string& msg);
const char* last_error();

// Registry of live objects addressed by opaque 64-bit handles; 0 is never a
// valid handle (the reference treats handle 0 as "not created").
template <class T>
class HandleTable {
public:
    uint64_t add(std::unique_ptr<T> obj)
    {
        std::lock_guard<std::mutex> lk(mu_);
        const uint64_t h = next_++;
        map_.emplace(h, std::shared_ptr<T>(std::move(obj)));
        return h;
    }
    std::shared_ptr<T> get(uint64_t h)
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = map_.find(h);
        return it == map_.end() ? nullptr : it->second;
    }
    // Calls f on every live object, outside the table's lock.
    template <class F>
    void for_each(F f)
    {
        std::vector<std::shared_ptr<T>> all;
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (auto& kv : map_) all.push_back(kv.second);
        }
        for (auto& o : all) f(*o);
    }
    bool remove(uint64_t h)
    {
        std::shared_ptr<T> victim;
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = map_.find(h);
            if (it == map_.end()) return false;
            victim = std::move(it->second);
            map_.erase(it);
        }
        return true;
    }

private:
    std::mutex mu_;
    uint64_t next_ = 1;
    std::unordered_map<uint64_t, std::shared_ptr<T>> map_;
};

// Device buffer owned by one object; freed on destruction.
template <class T>
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    void alloc(size_t count)
    {
        release();
        if (count == 0) return;
        SICC_HIP(hipMalloc(reinterpret_cast<void**>(&p_), count * sizeof(T)));
        n_ = count;
    }
    void release()
    {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

}  // namespace sicc
