// occupancy.cpp -- cached residency queries for the persistent launches.
//
// Persistent kernels size their grid to what the GPU holds at once; the HIP
// occupancy and attribute queries behind that cost microseconds per call, which
// small launches (one server batch, one plain 256-codeword decode) would pay on
// every launch.  Results are cached per device, kernel, block size and LDS.
#include <map>
#include <mutex>
#include <stdexcept>
#include <tuple>

#include "kernels.hpp"

namespace sicc {

namespace {

std::mutex& cache_mutex()
{
    static std::mutex m;
    return m;
}

}  // namespace

int device_cu_count()
{
    static std::map<int, int> cus;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("hipGetDevice failed");
    std::lock_guard<std::mutex> lk(cache_mutex());
    auto it = cus.find(dev);
    if (it != cus.end()) return it->second;
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        throw std::runtime_error("multiprocessor count query failed");
    n = n > 0 ? n : 1;
    cus[dev] = n;
    return n;
}

int resident_blocks_per_cu(const void* kernel, int threads, size_t lds)
{
    static std::map<std::tuple<int, const void*, int, size_t>, int> blocks;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("hipGetDevice failed");
    const auto key = std::make_tuple(dev, kernel, threads, lds);
    std::lock_guard<std::mutex> lk(cache_mutex());
    auto it = blocks.find(key);
    if (it != blocks.end()) return it->second;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds) != hipSuccess)
        throw std::runtime_error("occupancy query failed");
    per_cu = per_cu > 0 ? per_cu : 1;
    blocks[key] = per_cu;
    return per_cu;
}

}  // namespace sicc
