// frame_kernels.hip -- per-frame aggregation of decoded blocks into the
// reference's FrameDecodeResult, on the GPU.
//
// Semantics (LINK = source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp):
//   SNR gate                 LINK:5482-5501
//   CPU FEC branch           LINK:5529-5553 (parity_ok = every block's parity,
//                            iterations summed)
//   result                   LINK:5555-5591 (lround mean iterations, frame id
//                            = first 32 decoded info bits MSB first, bit and
//                            pre-FEC error counts against the reference
//                            payload and its encoding, frame_ok)
//   SNR / EVM                LINK:5456-5467 from the demodulator's per-symbol
//                            sums (OfdmDemodArgs::stats)
//
// Mapping: persistent 256-thread workgroups stride over frames.  The pre-FEC
// count re-encodes the reference payload block by block: the block's k
// information bits are packed into 64-bit words in LDS by ballots, and each
// thread forms encoded bits as parities of (information AND generator column)
// words.  The generator's non-identity columns sit in LDS once per workgroup
// when they fit (48 KiB; the n = 1024 rate-1/2 QC code needs 32 KiB), else
// they are read through L2.  Information-bit and flag traffic is read once.
#include "frame.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace sicc {

namespace {

constexpr int kThreads = 256;
constexpr int kGenLdsBytes = 48 * 1024;

// Sums of three ints over the workgroup (all threads get the totals).
__device__ __forceinline__ void block_sum3(int& a, int& b, int& c, int* red)
{
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
        c += __shfl_xor(c, o);
    }
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[3 * w] = a;
        red[3 * w + 1] = b;
        red[3 * w + 2] = c;
    }
    __syncthreads();
    a = b = c = 0;
    for (int i = 0; i < kThreads / 64; ++i) {
        a += red[3 * i];
        b += red[3 * i + 1];
        c += red[3 * i + 2];
    }
}

__global__ __launch_bounds__(kThreads) void frame_finish_kernel(FrameFinishArgs a)
{
    extern __shared__ unsigned long long lds_words[];  // [kw] info words, then the LDS generator copy
    __shared__ int red[3 * (kThreads / 64)];
    __shared__ int s_flag[2];
    unsigned long long* info_w = lds_words;
    const unsigned long long* gen = a.gen_cols;  // [(n - c0)][kw]
    const int ncols = a.n - a.c0;
    const bool want_pre = a.ref_info && a.gen_cols;
    if (want_pre && a.gen_in_lds) {
        unsigned long long* g = lds_words + a.kw;
        const int words = ncols * a.kw;
        for (int i = threadIdx.x; i < words; i += blockDim.x) g[i] = a.gen_cols[i];
        gen = g;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const size_t ninfo = static_cast<size_t>(a.blocks) * a.k;
    for (int f = blockIdx.x; f < a.frames; f += gridDim.x) {
        sicc_frame_result_t r;
        r.detected = 1;
        r.peak = 0.0;
        r.cfo_hz = 0.0;
        r.snr_db = 0.0;
        r.evm_rms = NAN;
        r.snr_gated = 0;
        r.reserved = 0;
        r.blocks = a.blocks;
        if (a.sync) {
            const OfdmSyncResult sr = a.sync[f];
            r.detected = sr.start >= 0 && static_cast<long long>(sr.start) + a.frame_samples <= a.span;
            r.peak = sr.start >= 0 ? sr.metric : sr.coarse_metric;
            r.cfo_hz = sr.cfo_hz_d;
        }
        if (a.stats && r.detected) {
            // per-symbol sums, added in symbol order (LINK:5431-5467)
            double sh = 0.0, se = 0.0, sref = 0.0;
            const double* st = a.stats + static_cast<size_t>(f) * a.n_data * 4;
            for (int d = 0; d < a.n_data; ++d) {
                sh += st[4 * d];
                se += st[4 * d + 2];
                sref += st[4 * d + 3];
            }
            const double ratio = (sh / a.n_re) / a.noise_var;
            r.snr_db = 10.0 * log10(ratio > 1e-12 ? ratio : 1e-12);
            r.evm_rms = sqrt((se > 0.0 ? se : 0.0) / (sref > 1e-12 ? sref : 1e-12));
            r.snr_gated = r.snr_db < a.snr_gate_db;
        }
        // blocks: parity (BP accepted, or decoded by OSD), iteration sum, counts
        const size_t b0 = static_cast<size_t>(f) * a.blocks;
        int all_ok = 0, iters = 0, nbp = 0;  // all_ok counts failing blocks
        int nosd = 0;
        for (int b = threadIdx.x; b < a.blocks; b += blockDim.x) {
            const int bp = a.bp_ok[b0 + b] != 0;
            const int os = a.used_osd ? a.used_osd[b0 + b] != 0 : 0;
            const int ok = bp || os;
            if (a.block_ok) a.block_ok[b0 + b] = static_cast<uint8_t>(ok);
            all_ok += !ok;
            iters += a.iterations[b0 + b];
            nbp += bp;
            nosd += os;
        }
        block_sum3(all_ok, iters, nbp, red);
        {
            int z0 = nosd, z1 = 0, z2 = 0;
            block_sum3(z0, z1, z2, red);
            nosd = z0;
        }
        r.bp_ok_blocks = nbp;
        r.osd_blocks = nosd;
        uint8_t* info = a.info + static_cast<size_t>(f) * ninfo;
        if (!r.detected || r.snr_gated) {
            // not decoded (LINK:5482-5501; undetected frames are never decoded, LINK:5618-5631)
            for (size_t i = threadIdx.x; i < ninfo; i += blockDim.x) info[i] = 0;
            r.parity_ok = 0;
            r.iterations = 0;
            r.frame_id = 0;
            r.bit_errors = -1;
            r.pre_fec_bit_errors = -1;
            r.frame_ok = 0;
        } else {
            r.parity_ok = all_ok == 0;
            // lround(iters / blocks) for iters >= 0: floor((2 iters + blocks) / (2 blocks))
            r.iterations = a.blocks > 0 ? (2 * iters + a.blocks) / (2 * a.blocks) : 0;
            if (a.ref_info) {
                const uint8_t* ref = a.ref_info + static_cast<size_t>(f) * ninfo;
                int be = 0;
                for (size_t i = threadIdx.x; i < ninfo; i += blockDim.x) be += info[i] != ref[i];
                int pe = 0;
                if (want_pre) {
                    const float* L = a.llr + static_cast<size_t>(f) * a.blocks * a.n;
                    for (int b = 0; b < a.blocks; ++b) {
                        const uint8_t* rb = ref + static_cast<size_t>(b) * a.k;
                        const float* lb = L + static_cast<size_t>(b) * a.n;
                        // pack the block's reference information bits, 64 per word
                        for (int w = wave; w < a.kw; w += kThreads / 64) {
                            const int i = w * 64 + lane;
                            const unsigned long long m = __ballot(i < a.k && rb[i] != 0);
                            if (lane == 0) info_w[w] = m;
                        }
                        __syncthreads();
                        // systematic columns first (the information bits themselves)
                        for (int c = threadIdx.x; c < a.c0; c += blockDim.x) pe += (lb[c] < 0.0f) != (rb[c] != 0);
                        for (int j = threadIdx.x; j < ncols; j += blockDim.x) {
                            const unsigned long long* gc = gen + static_cast<size_t>(j) * a.kw;
                            int par = 0;
                            for (int w = 0; w < a.kw; ++w) par ^= __popcll(info_w[w] & gc[w]);
                            pe += (lb[a.c0 + j] < 0.0f) != ((par & 1) != 0);
                        }
                        __syncthreads();
                    }
                }
                int z = 0;
                block_sum3(be, pe, z, red);
                if (threadIdx.x == 0) {
                    unsigned v = 0;
                    for (int i = 0; i < 32; ++i) v = (v << 1) | (static_cast<size_t>(i) < ninfo ? (info[i] & 1u) : 0u);
                    s_flag[0] = static_cast<int>(v);
                }
                __syncthreads();
                r.frame_id = static_cast<unsigned>(s_flag[0]);
                r.bit_errors = be;
                r.pre_fec_bit_errors = want_pre ? pe : -1;
                r.frame_ok = r.parity_ok && be == 0;
                __syncthreads();
            } else {
                r.frame_id = 0;
                r.bit_errors = -1;
                r.pre_fec_bit_errors = -1;
                r.frame_ok = 0;
            }
        }
        if (threadIdx.x == 0) a.out[f] = r;
        __syncthreads();
    }
}

}  // namespace

bool frame_gen_fits_lds(int ncols, int kw)
{
    return static_cast<size_t>(ncols) * kw * 8 + static_cast<size_t>(kw) * 8 <= kGenLdsBytes;
}

void launch_frame_finish(const FrameFinishArgs& a, hipStream_t s)
{
    if (a.frames <= 0) return;
    if (a.blocks <= 0 || a.n <= 0 || a.k <= 0 || !a.bp_ok || !a.iterations || !a.info || !a.out)
        throw std::invalid_argument("frame_finish: bad arguments");
    if (a.ref_info && a.gen_cols && (a.kw != (a.k + 63) / 64 || a.c0 < 0 || a.c0 > a.n || !a.llr))
        throw std::invalid_argument("frame_finish: bad generator table");
    if (a.stats && (a.n_data <= 0 || !(a.n_re > 0.0) || !(a.noise_var > 0.0)))
        throw std::invalid_argument("frame_finish: bad statistics arguments");
    size_t lds = sizeof(unsigned long long) * static_cast<size_t>(a.kw > 0 ? a.kw : 1);
    if (a.ref_info && a.gen_cols && a.gen_in_lds) {
        if (!frame_gen_fits_lds(a.n - a.c0, a.kw)) throw std::invalid_argument("frame_finish: generator exceeds LDS");
        lds += sizeof(unsigned long long) * static_cast<size_t>(a.n - a.c0) * a.kw;
    }
    // persistent: as many workgroups as the GPU holds at once (each stages the generator)
    const int per_cu = resident_blocks_per_cu(reinterpret_cast<const void*>(frame_finish_kernel), kThreads, lds);
    const int blocks = std::min(a.frames, device_cu_count() * std::max(per_cu, 1));
    hipLaunchKernelGGL(frame_finish_kernel, dim3(blocks), dim3(kThreads), lds, s, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("frame_finish_kernel: ") + hipGetErrorString(e));
}

}  // namespace sicc
