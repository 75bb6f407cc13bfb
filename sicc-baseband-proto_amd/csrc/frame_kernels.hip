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
// count re-encodes the reference payload: each wavefront takes whole blocks of
// the frame (block b on wave b mod 4), packs the block's k information bits
// into 64-bit words by ballots into its own LDS words, and each lane forms
// encoded bits as parities of (information AND generator column) words, over
// columns lane, lane + 64, ...  The generator is stored word-major
// (gen[w][column]), so a wavefront's reads of one word of 64 consecutive
// columns are consecutive; its non-identity columns sit in LDS once per
// workgroup when they fit (48 KiB; the n = 1024 rate-1/2 QC code needs
// 32 KiB), else they are read through L2.  Blocks need no workgroup barrier;
// the LLRs, information bits and flags are each read once.
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

// Bytes that differ between two words (any byte values).
__device__ __forceinline__ int diff_bytes(unsigned x, unsigned y)
{
    unsigned d = x ^ y;
    d |= d >> 4;
    d |= d >> 2;
    d |= d >> 1;
    return __popc(d & 0x01010101u);
}

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// GEN_LDS: the generator is staged in LDS (its reads compile to LDS
// instructions; a pointer that may address either space would compile to FLAT
// loads, several times slower here), else it is read through L2.
template <bool GEN_LDS>
__global__ __launch_bounds__(kThreads) void frame_finish_kernel(FrameFinishArgs a)
{
    // [waves][kw] packed information words, then the LDS generator copy [kw][ncols]
    extern __shared__ unsigned long long lds_words[];
    __shared__ int red[3 * (kThreads / 64)];
    __shared__ int s_flag[2];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    constexpr int kWaves = kThreads / 64;
    unsigned long long* info_w = lds_words + wave * a.kw;
    const int ncols = a.n - a.c0;
    const bool want_pre = a.ref_info && a.gen_cols;
    const unsigned long long* gen;  // [kw][ncols]
    if (GEN_LDS) {
        unsigned long long* g = lds_words + kWaves * a.kw;
        const int words = want_pre ? ncols * a.kw : 0;
        constexpr int U = 8;  // loads in flight per thread
        for (int i0 = threadIdx.x; i0 < words; i0 += U * kThreads) {
            unsigned long long x[U];
#pragma unroll
            for (int t = 0; t < U; ++t) x[t] = i0 + t * kThreads < words ? a.gen_cols[i0 + t * kThreads] : 0ull;
#pragma unroll
            for (int t = 0; t < U; ++t)
                if (i0 + t * kThreads < words) g[i0 + t * kThreads] = x[t];
        }
        gen = g;
    } else {
        gen = a.gen_cols;
    }
    __syncthreads();
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
            // per-symbol sums (LINK:5431-5467), the symbols spread over the lanes of
            // every wave (all waves compute the same totals)
            double sh = 0.0, se = 0.0, sref = 0.0;
            const double* st = a.stats + static_cast<size_t>(f) * a.n_data * 4;
            for (int d = lane; d < a.n_data; d += 64) {
                sh += st[4 * d];
                se += st[4 * d + 2];
                sref += st[4 * d + 3];
            }
            for (int o = 32; o > 0; o >>= 1) {
                sh += __shfl_xor(sh, o);
                se += __shfl_xor(se, o);
                sref += __shfl_xor(sref, o);
            }
            const double ratio = (sh / a.n_re) / a.noise_var;
            r.snr_db = 10.0 * log10(ratio > 1e-12 ? ratio : 1e-12);
            r.evm_rms = sqrt((se > 0.0 ? se : 0.0) / (sref > 1e-12 ? sref : 1e-12));
            r.snr_gated = r.snr_db < a.snr_gate_db;
        }
        // blocks: parity (BP accepted, or decoded by OSD), iteration sum, counts
        const size_t b0 = static_cast<size_t>(f) * a.blocks;
        int nfail = 0, iters = 0, nbp = 0, nosd = 0;
        for (int b = threadIdx.x; b < a.blocks; b += blockDim.x) {
            const int bp = a.bp_ok[b0 + b] != 0;
            const int os = a.used_osd ? a.used_osd[b0 + b] != 0 : 0;
            const int ok = bp || os;
            if (a.block_ok) a.block_ok[b0 + b] = static_cast<uint8_t>(ok);
            nfail += !ok;
            iters += a.iterations[b0 + b];
            nbp += bp | (os << 16);  // both counts fit 16 bits (blocks <= 65535)
        }
        block_sum3(nfail, iters, nbp, red);
        nosd = nbp >> 16;
        nbp &= 0xffff;
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
            r.parity_ok = nfail == 0;
            // lround(iters / blocks) for iters >= 0: floor((2 iters + blocks) / (2 blocks))
            r.iterations = a.blocks > 0 ? (2 * iters + a.blocks) / (2 * a.blocks) : 0;
            if (a.ref_info) {
                const uint8_t* ref = a.ref_info + static_cast<size_t>(f) * ninfo;
                int be = 0;
                if (((reinterpret_cast<uintptr_t>(info) | reinterpret_cast<uintptr_t>(ref)) & 3u) == 0) {
                    const unsigned* iw = reinterpret_cast<const unsigned*>(info);
                    const unsigned* rw = reinterpret_cast<const unsigned*>(ref);
                    const size_t nw = ninfo / 4;
                    constexpr int U = 4;  // loads in flight per thread
                    for (size_t i0 = threadIdx.x; i0 < nw; i0 += U * blockDim.x) {
                        unsigned x[U], y[U];
#pragma unroll
                        for (int t = 0; t < U; ++t) {
                            const size_t i = i0 + static_cast<size_t>(t) * blockDim.x;
                            x[t] = i < nw ? iw[i] : 0u;
                            y[t] = i < nw ? rw[i] : 0u;
                        }
#pragma unroll
                        for (int t = 0; t < U; ++t) be += diff_bytes(x[t], y[t]);
                    }
                    for (size_t i = nw * 4 + threadIdx.x; i < ninfo; i += blockDim.x) be += info[i] != ref[i];
                } else {
                    for (size_t i = threadIdx.x; i < ninfo; i += blockDim.x) be += info[i] != ref[i];
                }
                int pe = 0;
                if (want_pre) {
                    const float* L = a.llr + static_cast<size_t>(f) * a.blocks * a.n;
                    for (int b = wave; b < a.blocks; b += kWaves) {
                        const uint8_t* rb = ref + static_cast<size_t>(b) * a.k;
                        const float* lb = L + static_cast<size_t>(b) * a.n;
                        // pack the block's reference information bits, 64 per word (PW
                        // words' loads in flight before their ballots)
                        constexpr int PW = 8;
                        for (int w0 = 0; w0 < a.kw; w0 += PW) {
                            uint8_t v[PW];
#pragma unroll
                            for (int t = 0; t < PW; ++t) {
                                const int i = (w0 + t) * 64 + lane;
                                v[t] = (w0 + t < a.kw && i < a.k) ? rb[i] : 0;
                            }
#pragma unroll
                            for (int t = 0; t < PW; ++t) {
                                const unsigned long long m = __ballot(v[t] != 0);
                                if (lane == 0 && w0 + t < a.kw) info_w[w0 + t] = m;
                            }
                        }
                        wave_lds_sync();
                        // systematic columns (the information bits themselves), then the
                        // parity columns; PC columns' loads in flight at a time
                        constexpr int PC = 8;
                        for (int c0 = lane; c0 < a.c0; c0 += 64 * PC) {
                            float l[PC];
                            uint8_t x[PC];
#pragma unroll
                            for (int t = 0; t < PC; ++t) {
                                const int c = min(c0 + 64 * t, a.c0 - 1);
                                l[t] = lb[c];
                                x[t] = rb[c];
                            }
#pragma unroll
                            for (int t = 0; t < PC; ++t)
                                pe += (c0 + 64 * t < a.c0) & ((l[t] < 0.0f) != (x[t] != 0));
                        }
                        for (int j0 = lane; j0 < ncols; j0 += 64 * PC) {
                            // an encoded bit is the parity of the XOR of its words (info AND
                            // column): XOR the words, one popcount at the end; lanes past the
                            // last column read the last one and are masked out (branch-free)
                            float l[PC];
                            unsigned long long acc[PC];
                            int jc[PC];
#pragma unroll
                            for (int t = 0; t < PC; ++t) {
                                jc[t] = min(j0 + 64 * t, ncols - 1);
                                l[t] = lb[a.c0 + jc[t]];
                                acc[t] = 0ull;
                            }
                            for (int w = 0; w < a.kw; ++w) {
                                const unsigned long long iw = info_w[w];
                                const unsigned long long* gw = gen + static_cast<size_t>(w) * ncols;
#pragma unroll
                                for (int t = 0; t < PC; ++t) acc[t] ^= iw & gw[jc[t]];
                            }
#pragma unroll
                            for (int t = 0; t < PC; ++t)
                                pe += (j0 + 64 * t < ncols) & ((l[t] < 0.0f) != ((__popcll(acc[t]) & 1) != 0));
                        }
                        wave_lds_sync();  // the next block's packing overwrites info_w
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
    return (static_cast<size_t>(ncols) + kThreads / 64) * kw * 8 <= kGenLdsBytes;
}

void launch_frame_finish(const FrameFinishArgs& a, hipStream_t s)
{
    if (a.frames <= 0) return;
    if (a.blocks <= 0 || a.blocks > 65535 || a.n <= 0 || a.k <= 0 || !a.bp_ok || !a.iterations || !a.info || !a.out)
        throw std::invalid_argument("frame_finish: bad arguments");
    if (a.ref_info && a.gen_cols && (a.kw != (a.k + 63) / 64 || a.c0 < 0 || a.c0 > a.n || !a.llr))
        throw std::invalid_argument("frame_finish: bad generator table");
    if (a.stats && (a.n_data <= 0 || !(a.n_re > 0.0) || !(a.noise_var > 0.0)))
        throw std::invalid_argument("frame_finish: bad statistics arguments");
    size_t lds = sizeof(unsigned long long) * (kThreads / 64) * static_cast<size_t>(a.kw > 0 ? a.kw : 1);
    if (a.ref_info && a.gen_cols && a.gen_in_lds) {
        if (!frame_gen_fits_lds(a.n - a.c0, a.kw)) throw std::invalid_argument("frame_finish: generator exceeds LDS");
        lds += sizeof(unsigned long long) * static_cast<size_t>(a.n - a.c0) * a.kw;
    }
    // persistent: as many workgroups as the GPU holds at once (each stages the generator)
    const bool gl = a.ref_info && a.gen_cols && a.gen_in_lds;
    const void* kfn = gl ? reinterpret_cast<const void*>(frame_finish_kernel<true>)
                         : reinterpret_cast<const void*>(frame_finish_kernel<false>);
    const int per_cu = resident_blocks_per_cu(kfn, kThreads, lds);
    const int blocks = std::min(a.frames, device_cu_count() * std::max(per_cu, 1));
    if (gl)
        hipLaunchKernelGGL(frame_finish_kernel<true>, dim3(blocks), dim3(kThreads), lds, s, a);
    else
        hipLaunchKernelGGL(frame_finish_kernel<false>, dim3(blocks), dim3(kThreads), lds, s, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("frame_finish_kernel: ") + hipGetErrorString(e));
}

}  // namespace sicc
