// decoder_handles.hpp -- the objects behind decoder handles of the C ABI
// (stream BP-OSD, OSD-only and BP-only CSR decoders), shared by the ABI
// implementation and the serving loop.
#pragma once

#include "../../include/matlab_decoder_c_api.h"
#include "engine.hpp"

#include <chrono>
#include <deque>
#include <memory>
#include <string>
#include <vector>

namespace sicc {

inline uint64_t now_us()
{
    return static_cast<uint64_t>(std::chrono::duration_cast<std::chrono::microseconds>(
                                     std::chrono::steady_clock::now().time_since_epoch())
                                     .count());
}

enum class Kind { Stream, OsdOnly, BpCsr };

// Pinned host buffer.
template <class T>
class PinBuf {
public:
    PinBuf() = default;
    ~PinBuf() { release(); }
    PinBuf(const PinBuf&) = delete;
    PinBuf& operator=(const PinBuf&) = delete;
    void alloc(size_t count)
    {
        release();
        if (count == 0) return;
        SICC_HIP(hipHostMalloc(reinterpret_cast<void**>(&p_), count * sizeof(T), hipHostMallocDefault));
        n_ = count;
    }
    void release()
    {
        if (p_) (void)hipHostFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// One launch of the stream decoder: device inputs/outputs and pinned results.
struct Batch {
    DevBuf<float> llr;
    DevBuf<uint8_t> info;
    DevBuf<uint8_t> used;
    DevBuf<uint8_t> ok;
    PinBuf<uint8_t> h_info;
    PinBuf<uint8_t> h_used;
    PinBuf<uint8_t> h_ok;
    std::vector<unsigned long long> ts, ids;
    std::vector<unsigned int> blk;
    int count = 0;
    uint64_t first_us = 0;
    bool inflight = false;
    hipEvent_t start = nullptr;
    hipEvent_t done = nullptr;
};

struct ReadyItem {
    unsigned long long ts, id;
    unsigned int blk;
    uint8_t used;
    std::vector<uint8_t> info;
};

struct Decoder {
    Kind kind;
    int device = 0;
    std::unique_ptr<Engine> eng;
    BpParams bp;
    int osd_order = 2;
    OsdParams osd;
    int schedule = 0;
    hipStream_t stream = nullptr;
    // stream decoder state
    cuda_osd_stream_config_t cfg{};
    int max_batch = 1;
    int min_batch = 1;
    std::vector<std::unique_ptr<Batch>> slots;
    int filling = -1;
    std::deque<int> inflight;
    std::deque<ReadyItem> ready;
    bool flushed = false;
    cuda_osd_stream_stats_t st{};
    uint64_t polled = 0;
    uint64_t last_arrival_us = 0;
    // scratch for synchronous host calls
    DevBuf<float> s_llr;
    DevBuf<uint8_t> s_a, s_b, s_c, s_d;
    DevBuf<float> s_app, s_dist;
    DevBuf<int> s_it;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    ~Decoder()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& b : slots) {
            if (b->start) (void)hipEventDestroy(b->start);
            if (b->done) (void)hipEventDestroy(b->done);
        }
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

HandleTable<Decoder>& table();

template <class F>
int guarded(F&& f)
{
    try {
        return f();
    } catch (const std::invalid_argument& e) {
        set_last_error(e.what());
        return CUDA_OSD_STATUS_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return CUDA_OSD_STATUS_INTERNAL_ERROR;
    } catch (...) {
        set_last_error("unknown error");
        return CUDA_OSD_STATUS_INTERNAL_ERROR;
    }
}

std::shared_ptr<Decoder> lookup(unsigned long long h, const char* who);

}  // namespace sicc
