// decode_noalloc.cpp -- sicc_decode_device_noalloc: sicc_decode_device on a
// fixed device scratch array, for callers that cannot allocate (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The scratch of a decode -- the BP -> OSD hand-off (rejected frames' LLRs and
// indices) and the hand-off / work-stealing counters -- lives in the
// file-scope device array `scratch` below, one copy per GPU, held until the
// process ends.  Its device address comes from hipGetSymbolAddress, once per
// device on that device's first call, and only that address is ever handed to
// kernels: the array's name in host code is a host-side stand-in.  Every call
// on a GPU shares that GPU's array (the counters are cleared by a kernel at the
// start of each decode), so calls that overlap on one GPU race.
#include "../../include/sicc_decode_ext.h"

#include "decoder_handles.hpp"

#include <map>
#include <mutex>
#include <string>

// Layout: [0, 256) counters, then osd_index [cap] ints (rounded up to 256 B),
// then osd_llr [cap][n] floats.
static __device__ __align__(256) unsigned char scratch[SICC_NOALLOC_SCRATCH_BYTES];

namespace sicc {

namespace {

constexpr size_t kScratchBytes = static_cast<size_t>(SICC_NOALLOC_SCRATCH_BYTES);
constexpr size_t kHeaderBytes = 256;

std::mutex g_mu;
std::map<int, unsigned char*>& addresses()
{
    // never destroyed: no static destructor touches the runtime at exit
    static auto& m = *new std::map<int, unsigned char*>;
    return m;
}

// Device address of this device's copy of `scratch` (device must be current).
unsigned char* device_scratch()
{
    int dev = 0;
    SICC_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = addresses().find(dev);
    if (it != addresses().end()) return it->second;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) throw_hip(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    auto* base = static_cast<unsigned char*>(p);
    addresses()[dev] = base;
    return base;
}

size_t round256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

// Codewords of length n whose OSD hand-off fits the array.
int capacity(int n)
{
    size_t lo = 0, hi = kScratchBytes / (4 * static_cast<size_t>(n) + 4) + 1;
    while (lo + 1 < hi) {  // largest cap with header + round256(4 cap) + 4 n cap <= bytes
        const size_t mid = (lo + hi) / 2;
        if (kHeaderBytes + round256(4 * mid) + 4 * mid * static_cast<size_t>(n) <= kScratchBytes)
            lo = mid;
        else
            hi = mid;
    }
    return static_cast<int>(lo);
}

}  // namespace
}  // namespace sicc

using namespace sicc;

extern "C" {

int sicc_decode_device_noalloc(unsigned long long handle, const float* device_llrs, int codewords,
                               unsigned char* d_info_bits, unsigned char* d_code_bits, float* d_app_llrs,
                               unsigned char* d_bp_success, unsigned char* d_used_osd, int* d_iterations,
                               float* d_distances, void* hip_stream)
{
    return guarded([&] {
        auto d = lookup(handle, "decode_device_noalloc");
        if (!device_llrs || codewords <= 0) throw std::invalid_argument("decode_device_noalloc: empty batch");
        hipStream_t s = hip_stream ? static_cast<hipStream_t>(hip_stream) : d->stream;
        DecodeOut out;
        out.info = d_info_bits;
        out.code = d_code_bits;
        out.app = d_app_llrs;
        out.bp_ok = d_bp_success;
        out.used_osd = d_used_osd;
        out.iterations = d_iterations;
        out.distance = d_distances;
        if (d->kind == Kind::OsdOnly) {
            // OSD on every codeword reads the caller's LLRs directly: no scratch
            OsdParams op = d->osd;
            op.order = d->osd_order;
            d->eng->osd_only(device_llrs, codewords, op, out, s);
            return CUDA_OSD_STATUS_OK;
        }
        const int n = d->eng->code().n;
        const int osd_order = d->kind == Kind::Stream ? d->osd_order : -1;
        const int cap = capacity(n);
        if (osd_order >= 0 && codewords > cap)
            throw std::invalid_argument("decode_device_noalloc: " + std::to_string(codewords) +
                                        " codewords exceed the scratch limit of " + std::to_string(cap) +
                                        " for n = " + std::to_string(n) + " (" +
                                        std::to_string(kScratchBytes >> 20) + " MiB array)");
        unsigned char* base = device_scratch();
        ScratchView v;
        v.counters = reinterpret_cast<int*>(base);
        v.osd_index = reinterpret_cast<int*>(base + kHeaderBytes);
        v.osd_llr = reinterpret_cast<float*>(base + kHeaderBytes + round256(4 * static_cast<size_t>(cap)));
        v.cap = cap;
        d->eng->decode_into(device_llrs, codewords, d->bp, osd_order, d->osd, out, s, v);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_decode_noalloc_capacity(unsigned long long handle, int* max_codewords, unsigned long long* scratch_bytes)
{
    return guarded([&] {
        auto d = lookup(handle, "decode_noalloc_capacity");
        if (max_codewords) *max_codewords = capacity(d->eng->code().n);
        if (scratch_bytes) *scratch_bytes = kScratchBytes;
        return CUDA_OSD_STATUS_OK;
    });
}

}  // extern "C"
