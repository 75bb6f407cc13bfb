// bp_kernels.hip -- batched normalized min-sum LDPC BP for gfx950.
//
// Semantics follow the reference CPU decoder ldpc_decode_min_sum
// (source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp:3013-3094): flooding
// schedule, min1/min2/first-min-edge check update, r = alpha * sign * mag,
// APP = LLR + sum of check messages in increasing row order, hard = APP < 0,
// syndrome stop.  The MATLAB reference additions
// (source/matlab/uhd_cpp/matlab/cpu_bp_osd_decode.m:115-177) are carried too:
// best-syndrome APP tracking for the OSD hand-off and optional VN damping.
// Every float operation is the same operation, in the same order, as the CPU
// decoder's, so outputs are bit-identical (built with -ffp-contract=off).
//
// Mapping: one wavefront decodes one codeword.  Lane l owns check rows
// l, l+64, ... and variable nodes l, l+64, ...  Registers hold each lane's row
// edge lists, its rows' previous check messages (so q = APP - r_old needs no
// stored q) and its variables' LLR / APP / best-syndrome APP.  Per wave, LDS
// holds only what crosses lanes: APP[n] (gathered along rows) and the new
// check messages r[E] in column-major order (summed along columns).  The check
// update is branch-free: padded edges point at a FLT_MAX sentinel APP, min1/min2
// use min/max, and the first-minimum test is |q| == min1 (equal values imply
// min2 == min1).  HBM traffic per codeword is the LLR read and the bit/flag
// writes.  Waves are persistent and stride over the batch.
#include "kernels.hpp"

#include <cfloat>
#include <stdexcept>

namespace sicc {

namespace {

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void zero_i32_kernel(int* p, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

constexpr int kWavesPerBlock = 4;

template <int MPL, int NPL, int MAXDC, bool DAMP>
__global__ __launch_bounds__(256) void bp_flood_kernel(BpCodeDev code, BpParams p, BpIo io, int e_pad)
{
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    constexpr int np = 64 * NPL;
    // per wave: APP[np] + one sentinel slot (FLT_MAX) for padded edges, then r[e_pad]
    float* app = lds + wv * (np + 4 + e_pad);
    float* rcol = app + np + 4;

    const int n = code.n;
    const int m = code.m;
    // Padded edges read the sentinel APP and write the spare message slot e_pad-1,
    // so every row runs the same MAXDC-edge, branch-free update.
    const int pad_edge = (np << 16) | (e_pad - 1);
    int em[MPL][MAXDC];
#pragma unroll
    for (int s = 0; s < MPL; ++s) {
        const int r = lane + 64 * s;
#pragma unroll
        for (int j = 0; j < MAXDC; ++j) {
            const int e = (r < m && j < code.max_dc) ? code.row_meta[r * code.max_dc + j] : -1;
            em[s][j] = e >= 0 ? e : pad_edge;
        }
    }
    int cm[NPL];
#pragma unroll
    for (int t = 0; t < NPL; ++t) {
        const int v = lane + 64 * t;
        cm[t] = v < n ? code.col_meta[v] : 0;
    }
    if (lane == 0) app[np] = FLT_MAX;

    const int gw = blockIdx.x * (blockDim.x >> 6) + wv;
    const int nw = gridDim.x * (blockDim.x >> 6);
    const float alpha = p.alpha;
    const float off = p.offset;
    const float damp = p.damping;

    for (int cw = gw; cw < io.num_cw; cw += nw) {
        const float* L = io.llr + static_cast<size_t>(cw) * n;
        float lr[NPL], appr[NPL], bestr[NPL];
#pragma unroll
        for (int t = 0; t < NPL; ++t) {
            const int v = lane + 64 * t;
            lr[t] = v < n ? L[v] : 0.0f;
            appr[t] = lr[t];
            bestr[t] = lr[t];
            app[v] = lr[t];
        }
        float rold[MPL][MAXDC];
        float qold[MPL][MAXDC];
#pragma unroll
        for (int s = 0; s < MPL; ++s)
#pragma unroll
            for (int j = 0; j < MAXDC; ++j) {
                rold[s][j] = 0.0f;
                qold[s][j] = 0.0f;
            }
        wave_lds_sync();

        int best_w = 0;
        bool ok = false;
        int it = 0;
        for (;;) {
            // ---- check-node update; its APP gather also yields the syndrome ----
            int synw = 0;
#pragma unroll
            for (int s = 0; s < MPL; ++s) {
                unsigned hp = 0u, sp = 0u;
                float m1 = FLT_MAX, m2 = FLT_MAX;
                float qv[MAXDC];
#pragma unroll
                for (int j = 0; j < MAXDC; ++j) {
                    const float a = app[em[s][j] >> 16];
                    hp ^= __float_as_uint(a + 0.0f);  // +0.0f maps -0 to +0: sign bit == (a < 0)
                    float q = a - rold[s][j];
                    if (DAMP) {
                        if (it > 0) q = (1.0f - damp) * q + damp * qold[s][j];
                        qold[s][j] = q;
                    }
                    q = q + 0.0f;
                    qv[j] = q;
                    sp ^= __float_as_uint(q);
                    const float av = fabsf(q);
                    m2 = fminf(m2, fmaxf(m1, av));
                    m1 = fminf(m1, av);
                }
                synw += __popcll(__ballot((hp >> 31) != 0u));
                // The first minimum edge takes min2; any other edge equal to min1
                // implies min2 == min1, so |q| == min1 selects the same magnitude.
                float a1 = m1, a2 = m2;
                if (off != 0.0f) {
                    a1 = fmaxf(a1 - off, 0.0f);
                    a2 = fmaxf(a2 - off, 0.0f);
                }
                a1 = alpha * a1;
                a2 = alpha * a2;
#pragma unroll
                for (int j = 0; j < MAXDC; ++j) {
                    const float mag = (fabsf(qv[j]) == m1) ? a2 : a1;
                    const float r = __uint_as_float(__float_as_uint(mag) | ((sp ^ __float_as_uint(qv[j])) & 0x80000000u));
                    rold[s][j] = r;
                    rcol[em[s][j] & 0xffff] = r;
                }
            }
            if (it > 0) {
                if (synw < best_w || synw == 0) {
                    best_w = synw;
#pragma unroll
                    for (int t = 0; t < NPL; ++t) bestr[t] = appr[t];
                }
                if (synw == 0) {
                    ok = true;
                    break;
                }
            } else {
                best_w = synw;
            }
            if (it == p.max_iter) break;
            wave_lds_sync();
            // ---- variable-node update: APP = LLR + sum r, increasing row order ----
#pragma unroll
            for (int t = 0; t < NPL; ++t) {
                const int v = lane + 64 * t;
                if (v < n) {
                    float a = lr[t];
                    const int st = cm[t] >> 8;
                    const int dg = cm[t] & 255;
                    for (int j = 0; j < dg; ++j) a += rcol[st + j];
                    appr[t] = a;
                    app[v] = a;
                }
            }
            wave_lds_sync();
            ++it;
        }

        // ---- outputs ----
        bool accept = ok;
        if (ok && p.min_abs_accept > 0.0f) {
            float mn = FLT_MAX;
#pragma unroll
            for (int t = 0; t < NPL; ++t)
                if (lane + 64 * t < n) mn = fminf(mn, fabsf(appr[t]));
            for (int o = 32; o > 0; o >>= 1) mn = fminf(mn, __shfl_xor(mn, o));
            accept = mn >= p.min_abs_accept;
        }
        if (io.code_bits || io.app) {
#pragma unroll
            for (int t = 0; t < NPL; ++t) {
                const int v = lane + 64 * t;
                if (v < n) {
                    if (io.code_bits) io.code_bits[static_cast<size_t>(cw) * n + v] = appr[t] < 0.0f ? 1 : 0;
                    if (io.app) io.app[static_cast<size_t>(cw) * n + v] = appr[t];
                }
            }
        }
        if (io.info_bits) {
            const int k = code.k;
            for (int i = lane; i < k; i += 64)
                io.info_bits[static_cast<size_t>(cw) * k + i] = app[code.info_cols[i]] < 0.0f ? 1 : 0;
        }
        if (lane == 0) {
            if (io.success) io.success[cw] = accept ? 1 : 0;
            if (io.iterations) io.iterations[cw] = it;
            if (io.used_osd) io.used_osd[cw] = (!accept && io.osd_llr) ? 1 : 0;
            if (io.distance && accept) io.distance[cw] = 0.0f;
        }
        if (!accept && io.osd_llr) {
            int slot = 0;
            if (lane == 0) slot = atomicAdd(io.osd_count, 1);
            slot = __shfl(slot, 0);
            float* dst = io.osd_llr + static_cast<size_t>(slot) * n;
            // a converged-but-rejected frame hands over its final APP (== best)
#pragma unroll
            for (int t = 0; t < NPL; ++t) {
                const int v = lane + 64 * t;
                if (v < n) dst[v] = bestr[t];
            }
            if (lane == 0) io.osd_index[slot] = cw;
        }
        wave_lds_sync();
    }
}

template <int MPL, int NPL, int MAXDC>
void launch_variant(const BpCodeDev& code, const BpParams& p, const BpIo& io, int max_waves, hipStream_t s)
{
    const int e_pad = (code.E + 1 + 63) & ~63;  // one spare slot for padded edges
    const bool damp = p.damping != 0.0f;
    const size_t wave_bytes = sizeof(float) * static_cast<size_t>(64 * NPL + 4 + e_pad);
    // At most 64 KiB of LDS per workgroup; fewer waves per group for big codes.
    int wpb = static_cast<int>((64 * 1024) / wave_bytes);
    if (wpb > kWavesPerBlock) wpb = kWavesPerBlock;
    if (wpb < 1) throw std::invalid_argument("code too large for the LDS-resident BP kernel");
    const size_t lds = wave_bytes * wpb;
    int waves = io.num_cw < max_waves ? io.num_cw : max_waves;
    if (waves < 1) waves = 1;
    const int blocks = (waves + wpb - 1) / wpb;
    if (damp)
        hipLaunchKernelGGL((bp_flood_kernel<MPL, NPL, MAXDC, true>), dim3(blocks), dim3(64 * wpb), lds, s, code, p,
                           io, e_pad);
    else
        hipLaunchKernelGGL((bp_flood_kernel<MPL, NPL, MAXDC, false>), dim3(blocks), dim3(64 * wpb), lds, s, code, p,
                           io, e_pad);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("bp_flood_kernel launch: ") + hipGetErrorString(e));
}

}  // namespace

void launch_zero_i32(int* p, int count, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(zero_i32_kernel, dim3((count + 255) / 256), dim3(256), 0, s, p, count);
}

bool bp_shape_supported(const BpCodeDev& c)
{
    if (c.n <= 0 || c.m <= 0 || c.E <= 0 || c.E > 0xffff || c.n > 0x7fff) return false;
    const int mpl = (c.m + 63) / 64;
    const int npl = (c.n + 63) / 64;
    if (c.max_dc > 16) return false;
    if (c.max_dc > 8) return mpl <= 4 && npl <= 8;
    return mpl <= 8 && npl <= 16;
}

void launch_bp_flood(const BpCodeDev& code, const BpParams& p, const BpIo& io, int max_waves, hipStream_t s)
{
    if (!bp_shape_supported(code)) throw std::invalid_argument("code shape not supported by the BP kernel");
    if (io.num_cw <= 0) return;
    if (p.max_iter < 1) throw std::invalid_argument("max_iter must be >= 1");
    const int mpl = (code.m + 63) / 64;
    const int npl = (code.n + 63) / 64;
    const int dc = code.max_dc <= 8 ? 8 : 16;
    // Round the shape up to an instantiated variant.
    const int mp = mpl <= 1 ? 1 : mpl <= 2 ? 2 : mpl <= 4 ? 4 : 8;
    const int nq = npl <= 2 ? 2 : npl <= 4 ? 4 : npl <= 8 ? 8 : 16;
    if (dc == 8) {
        if (mp == 1 && nq == 2) return launch_variant<1, 2, 8>(code, p, io, max_waves, s);
        if (mp <= 2 && nq <= 4) return launch_variant<2, 4, 8>(code, p, io, max_waves, s);
        if (mp <= 4 && nq <= 8) return launch_variant<4, 8, 8>(code, p, io, max_waves, s);
        return launch_variant<8, 16, 8>(code, p, io, max_waves, s);
    }
    if (mp == 1 && nq == 2) return launch_variant<1, 2, 16>(code, p, io, max_waves, s);
    if (mp <= 2 && nq <= 4) return launch_variant<2, 4, 16>(code, p, io, max_waves, s);
    if (mp <= 4 && nq <= 8) return launch_variant<4, 8, 16>(code, p, io, max_waves, s);
    throw std::invalid_argument("code shape not supported by the BP kernel");
}

}  // namespace sicc
