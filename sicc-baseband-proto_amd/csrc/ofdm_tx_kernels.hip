// ofdm_tx_kernels.hip -- batched OFDM transmitter for gfx950 (device-resident
// test IQ and the transmit half of the C++ pipeline front).
//
// Behaviour follows the reference GPU transmitter (gpu_ofdm_pipeline.cu):
//   encode   generator parity per coded bit (PIPE:129-150) or the DVB-S2
//            systematic copy + row sums + accumulator (PIPE:152-214);
//   grid     pilots / Gray QAM of the coded bit stream at the natural FFT bin
//            of each used subcarrier, optional pre-equalization
//            value * conj(h) / max(|h|^2, 1e-10) (PIPE:216-257);
//   channel  optional value * actual_h on every resource element, AWGN of
//            variance noise_var on data resource elements (PIPE:259-314);
//   compose  unnormalized inverse FFT, preamble * amp, then per symbol the
//            cyclic prefix and body scaled by amp / nfft (PIPE:316-342).
// The AWGN restates the reference's counter-based generator (the lowbias32
// integer hash of seed ^ (index * 0x9e3779b9), Box-Muller on two uniforms,
// imaginary part from seed ^ 0xa5a5a5a5) so that a seed gives the same noise
// realization; frame f of a batch uses seed + f.
//
// Mapping: one workgroup per codeword encodes (information words or row sums
// staged per workgroup); one workgroup per (frame, symbol) builds the grid in
// LDS, runs the FP64 radix-4 FFT on the conjugate (= inverse transform) and
// writes cyclic prefix and body in one pass.
#include "fft_lds.hpp"
#include "ofdm.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace sicc {

namespace {

constexpr int kThreads = 256;
constexpr float kPiF = 3.14159265358979323846f;

__device__ __forceinline__ unsigned int lowbias32(unsigned int x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ float uniform_open(unsigned int seed, unsigned int idx)
{
    const unsigned int h = lowbias32(seed ^ (idx * 0x9e3779b9u));
    return (static_cast<float>(h) + 1.0f) * 2.3283064365386963e-10f;
}

__device__ __forceinline__ float normal01(unsigned int seed, unsigned int idx)
{
    const float u1 = fmaxf(uniform_open(seed, idx * 2u), 1.0e-7f);
    const float u2 = uniform_open(seed, idx * 2u + 1u);
    return sqrtf(-2.0f * logf(u1)) * cosf(2.0f * kPiF * u2);
}

// Codewords encoded per workgroup by the generator kernel: each generator
// column word is read once from L2 for all of them.
constexpr int kTxCwPerGroup = 8;

__global__ __launch_bounds__(kThreads) void tx_encode_generator_kernel(OfdmTxArgs a, int total_cw, int cwg)
{
    extern __shared__ unsigned long long words[];  // [cwg][gen_words]
    const int cw0 = blockIdx.x * cwg;
    const int gw = a.gen_words;
    const int lane = threadIdx.x & 63;
    // information bits packed 64 to a word by wavefront ballots
    for (int idx = threadIdx.x >> 6; idx < cwg * gw; idx += blockDim.x >> 6) {
        const int j = idx / gw, w = idx - j * gw;
        const int cw = cw0 + j;
        const int i = w * 64 + lane;
        const bool bit = cw < total_cw && i < a.code_k &&
                         (a.info[static_cast<size_t>(cw) * a.code_k + i] & 1u) != 0u;
        const unsigned long long v = __ballot(bit);
        if (lane == 0) words[idx] = v;
    }
    __syncthreads();
    const int nj = min(cwg, total_cw - cw0);
    for (int c = threadIdx.x; c < a.code_n; c += blockDim.x) {
        const unsigned long long* g = a.gen_cols + static_cast<size_t>(c) * gw;
        int par[kTxCwPerGroup];
#pragma unroll
        for (int j = 0; j < kTxCwPerGroup; ++j) par[j] = 0;
        for (int w = 0; w < gw; ++w) {
            const unsigned long long gv = g[w];
#pragma unroll
            for (int j = 0; j < kTxCwPerGroup; ++j)
                if (j < nj) par[j] ^= __popcll(words[j * gw + w] & gv);
        }
#pragma unroll
        for (int j = 0; j < kTxCwPerGroup; ++j) {
            if (j >= nj) break;
            const int cw = cw0 + j;
            const int f = cw / a.blocks, b = cw - f * a.blocks;
            a.coded[static_cast<size_t>(f) * a.coded_bits + static_cast<size_t>(b) * a.code_n + c] =
                static_cast<uint8_t>(par[j] & 1);
        }
    }
    // pad bits after the last codeword of each frame
    for (int j = 0; j < nj; ++j) {
        const int cw = cw0 + j;
        const int f = cw / a.blocks, b = cw - f * a.blocks;
        if (b != 0) continue;
        uint8_t* fr = a.coded + static_cast<size_t>(f) * a.coded_bits;
        for (int i = a.blocks * a.code_n + threadIdx.x; i < a.coded_bits; i += blockDim.x) fr[i] = 0;
    }
}

__global__ __launch_bounds__(kThreads) void tx_encode_dvbs2_kernel(OfdmTxArgs a)
{
    __shared__ unsigned char tot[kThreads];
    __shared__ unsigned char pre[kThreads];
    const int cw = blockIdx.x;
    const int f = cw / a.blocks;
    const int b = cw - f * a.blocks;
    const int k = a.code_k;
    const int m = a.code_n - k;
    const uint8_t* info = a.info + static_cast<size_t>(cw) * k;
    uint8_t* out = a.coded + static_cast<size_t>(f) * a.coded_bits + static_cast<size_t>(b) * a.code_n;
    for (int i = threadIdx.x; i < k; i += blockDim.x) out[i] = info[i] & 1u;
    // row sums, staged in the parity positions
    for (int r = threadIdx.x; r < m; r += blockDim.x) {
        unsigned int s = 0;
        for (int j = a.row_info_offsets[r]; j < a.row_info_offsets[r + 1]; ++j) s ^= info[a.row_info_cols[j]];
        out[k + r] = static_cast<uint8_t>(s & 1u);
    }
    __syncthreads();
    // accumulator = prefix XOR over the row sums: per-thread contiguous chunks
    const int chunk = (m + kThreads - 1) / kThreads;
    const int lo = min(m, static_cast<int>(threadIdx.x) * chunk);
    const int hi = min(m, lo + chunk);
    unsigned char t = 0;
    for (int r = lo; r < hi; ++r) t ^= out[k + r];
    tot[threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned char p = 0;
        for (int i = 0; i < kThreads; ++i) {
            pre[i] = p;
            p ^= tot[i];
        }
    }
    __syncthreads();
    unsigned char p = pre[threadIdx.x];
    for (int r = lo; r < hi; ++r) {
        p ^= out[k + r];
        out[k + r] = p;
    }
    if (b == 0) {
        uint8_t* fr = a.coded + static_cast<size_t>(f) * a.coded_bits;
        for (int i = a.blocks * a.code_n + threadIdx.x; i < a.coded_bits; i += blockDim.x) fr[i] = 0;
    }
}

__global__ __launch_bounds__(kThreads) void tx_symbol_kernel(OfdmTxArgs a)
{
    // single-precision grid and inverse FFT (the frame is float IQ; the error of
    // an FP32 Stockham FFT, ~1e-7 of the frame rms, is far inside the frame tolerance)
    extern __shared__ float2 shf[];
    Cf* buf = reinterpret_cast<Cf*>(shf);
    const int f = blockIdx.x / a.num_symbols;
    const int sym = blockIdx.x - f * a.num_symbols;
    const int N = a.nfft;
    const int A = a.active_sc;
    const int B = a.bits_per_symbol;
    const int frame_len = a.preamble_len + a.num_symbols * (N + a.cp);
    float2* out = a.out + static_cast<size_t>(f) * frame_len;
    if (sym == 0)
        for (int i = threadIdx.x; i < a.preamble_len; i += blockDim.x)
            out[i] = make_float2(a.preamble[i].x * a.amplitude, a.preamble[i].y * a.amplitude);
    for (int i = threadIdx.x; i < N; i += blockDim.x) buf[i] = {0.0f, 0.0f};
    __syncthreads();
    const int pi = a.pilot_index[sym];
    const int di = a.data_index[sym];
    const uint8_t* bits = a.coded + static_cast<size_t>(f) * a.coded_bits;
    const unsigned long long s64 = a.noise_seed + static_cast<unsigned long long>(f);
    const unsigned int seed = static_cast<unsigned int>(s64 ^ (s64 >> 32));
    const float sigma = sqrtf(fmaxf(a.noise_var, 0.0f) * 0.5f);
    for (int sc = threadIdx.x; sc < A; sc += blockDim.x) {
        const int re = sym * A + sc;
        float2 v = make_float2(0.0f, 0.0f);
        if (pi >= 0) {
            v = a.pilot_grid[pi * A + sc];
        } else if (di >= 0) {
            const int base = (di * A + sc) * B;
            int idx = 0;
            for (int i = 0; i < B; ++i) idx = (idx << 1) | (base + i < a.coded_bits ? (bits[base + i] & 1) : 0);
            v = a.constellation[idx];
        }
        if (a.precomp_h) {
            const float2 h = a.precomp_h[re];
            const float inv = 1.0f / fmaxf(h.x * h.x + h.y * h.y, 1.0e-10f);
            const float2 c = make_float2(h.x * inv, -h.y * inv);
            v = make_float2(v.x * c.x - v.y * c.y, v.x * c.y + v.y * c.x);
        }
        if (a.actual_h) {
            const float2 h = a.actual_h[re];
            v = make_float2(v.x * h.x - v.y * h.y, v.x * h.y + v.y * h.x);
        }
        if (a.noise_var > 0.0f && di >= 0) {
            v.x += sigma * normal01(seed, static_cast<unsigned int>(re));
            v.y += sigma * normal01(seed ^ 0xa5a5a5a5u, static_cast<unsigned int>(re));
        }
        // conjugate in, forward FFT, conjugate out = unnormalized inverse FFT
        buf[a.used_natural[sc]] = {v.x, -v.y};
    }
    __syncthreads();
    fft_stockham<true>(buf, a.twiddle, N);
    const float scale = a.amplitude / static_cast<float>(N);
    float2* o = out + a.preamble_len + static_cast<size_t>(sym) * (N + a.cp);
    for (int i = threadIdx.x; i < N + a.cp; i += blockDim.x) {
        const Cf v = buf[i < a.cp ? N - a.cp + i : i - a.cp];
        o[i] = make_float2(v.x * scale, -v.y * scale);
    }
}

}  // namespace

void launch_ofdm_tx(const OfdmTxArgs& a, int frames, hipStream_t s)
{
    if (frames <= 0) return;
    if (!lds_fft_size_ok(a.nfft)) throw std::invalid_argument("TX FFT needs nfft a power of two in [16, 2048]");
    if (a.cp < 0 || a.cp > a.nfft) throw std::invalid_argument("TX cp out of range");
    if (a.code_k <= 0 || a.code_n <= a.code_k || a.blocks <= 0 ||
        static_cast<long long>(a.blocks) * a.code_n > a.coded_bits)
        throw std::invalid_argument("TX encoder shape");
    if (a.dvb_s2) {
        if (!a.row_info_offsets) throw std::invalid_argument("TX DVB-S2 encoder needs row_info_offsets");
        hipLaunchKernelGGL(tx_encode_dvbs2_kernel, dim3(frames * a.blocks), dim3(kThreads), 0, s, a);
    } else {
        if (!a.gen_cols || a.gen_words != (a.code_k + 63) / 64 || a.gen_words > 8192)
            throw std::invalid_argument("TX generator shape");
        const int total = frames * a.blocks;
        const int cwg = std::max(1, std::min(kTxCwPerGroup, 8192 / a.gen_words));  // <= 64 KiB of words
        hipLaunchKernelGGL(tx_encode_generator_kernel, dim3((total + cwg - 1) / cwg), dim3(kThreads),
                           sizeof(unsigned long long) * a.gen_words * cwg, s, a, total, cwg);
    }
    hipLaunchKernelGGL(tx_symbol_kernel, dim3(frames * a.num_symbols), dim3(kThreads), sizeof(float2) * a.nfft, s,
                       a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("ofdm tx kernels: ") + hipGetErrorString(e));
}

}  // namespace sicc
