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
// l, l+64, ... (registers hold their edge lists) and variable nodes l, l+64, ...
// Per wave, LDS holds APP[n], the best-syndrome APP[n] and the check-to-variable
// messages r[E] in column-major order, so both the check update (row gather of
// APP) and the variable update (contiguous column runs of r) read LDS only.
// The variable-to-check message q = APP - r_old is recomputed from LDS instead
// of being stored.  HBM traffic per codeword is the LLR read and the bit/flag
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
__global__ __launch_bounds__(256) void bp_flood_kernel(BpCodeDev code, BpParams p, BpIo io,
                                                       int e_pad)
{
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int np = 64 * NPL;
    const int wave_floats = 2 * np + e_pad * (DAMP ? 2 : 1);
    float* app = lds + wv * wave_floats;
    float* best = app + np;
    float* rcol = best + np;
    float* qst = rcol + e_pad;

    const int n = code.n;
    const int m = code.m;
    const int E = code.E;

    int em[MPL][MAXDC];
#pragma unroll
    for (int s = 0; s < MPL; ++s) {
        const int r = lane + 64 * s;
#pragma unroll
        for (int j = 0; j < MAXDC; ++j) em[s][j] = (r < m && j < code.max_dc) ? code.row_meta[r * code.max_dc + j] : -1;
    }
    int cm[NPL];
#pragma unroll
    for (int t = 0; t < NPL; ++t) {
        const int v = lane + 64 * t;
        cm[t] = v < n ? code.col_meta[v] : 0;
    }

    const int gw = blockIdx.x * (blockDim.x >> 6) + wv;
    const int nw = gridDim.x * (blockDim.x >> 6);
    const float alpha = p.alpha;

    for (int cw = gw; cw < io.num_cw; cw += nw) {
        const float* L = io.llr + static_cast<size_t>(cw) * n;
        float lr[NPL];
#pragma unroll
        for (int t = 0; t < NPL; ++t) {
            const int v = lane + 64 * t;
            lr[t] = v < n ? L[v] : 0.0f;
            app[v] = lr[t];
        }
        for (int i = lane; i < E; i += 64) rcol[i] = 0.0f;
        wave_lds_sync();

        int best_w = 0;
        bool best_is_channel = true;
        bool ok = false;
        int it = 0;
        for (;;) {
            // ---- check-node update + syndrome of the current hard decision ----
            int synw = 0;
#pragma unroll
            for (int s = 0; s < MPL; ++s) {
                int hp = 0;
                int sp = 0;
                float m1 = FLT_MAX;
                float m2 = FLT_MAX;
                int i1 = -1;
                float qv[MAXDC];
#pragma unroll
                for (int j = 0; j < MAXDC; ++j) {
                    const int e = em[s][j];
                    qv[j] = 0.0f;
                    if (e >= 0) {
                        const int col = e >> 16;
                        const int slot = e & 0xffff;
                        const float a = app[col];
                        hp ^= (a < 0.0f) ? 1 : 0;
                        float q = a - rcol[slot];
                        if (DAMP) {
                            if (it > 0) q = (1.0f - p.damping) * q + p.damping * qst[slot];
                            qst[slot] = q;
                        }
                        qv[j] = q;
                        sp ^= (q < 0.0f) ? 1 : 0;
                        const float av = fabsf(q);
                        if (av < m1) {
                            m2 = m1;
                            m1 = av;
                            i1 = j;
                        } else if (av < m2) {
                            m2 = av;
                        }
                    }
                }
                const bool row_valid = em[s][0] >= 0;
                synw += __popcll(__ballot(row_valid && hp));
                // Writing r_new right away is safe: each slot belongs to one row,
                // and nothing reads r before the variable update below.
#pragma unroll
                for (int j = 0; j < MAXDC; ++j) {
                    const int e = em[s][j];
                    if (e >= 0) {
                        const bool neg = (sp ^ (qv[j] < 0.0f ? 1 : 0)) != 0;
                        float mag = (j == i1) ? m2 : m1;
                        if (p.offset != 0.0f) mag = fmaxf(mag - p.offset, 0.0f);
                        const float sf = neg ? -1.0f : 1.0f;
                        rcol[e & 0xffff] = alpha * sf * mag;
                    }
                }
            }
            if (it > 0) {
                if (synw < best_w || synw == 0) {
                    best_w = synw;
                    best_is_channel = false;
#pragma unroll
                    for (int t = 0; t < NPL; ++t) {
                        const int v = lane + 64 * t;
                        best[v] = app[v];
                    }
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
            // ---- variable-node update: APP = LLR + sum r (increasing row) ----
#pragma unroll
            for (int t = 0; t < NPL; ++t) {
                const int v = lane + 64 * t;
                if (v < n) {
                    float a = lr[t];
                    const int st = cm[t] >> 8;
                    const int dg = cm[t] & 255;
                    for (int j = 0; j < dg; ++j) a += rcol[st + j];
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
            for (int t = 0; t < NPL; ++t) {
                const int v = lane + 64 * t;
                if (v < n) mn = fminf(mn, fabsf(app[v]));
            }
            for (int off = 32; off > 0; off >>= 1) mn = fminf(mn, __shfl_xor(mn, off));
            accept = mn >= p.min_abs_accept;
        }
        if (io.code_bits || io.app) {
#pragma unroll
            for (int t = 0; t < NPL; ++t) {
                const int v = lane + 64 * t;
                if (v < n) {
                    const float a = app[v];
                    if (io.code_bits) io.code_bits[static_cast<size_t>(cw) * n + v] = a < 0.0f ? 1 : 0;
                    if (io.app) io.app[static_cast<size_t>(cw) * n + v] = a;
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
#pragma unroll
            for (int t = 0; t < NPL; ++t) {
                const int v = lane + 64 * t;
                if (v < n) dst[v] = best_is_channel ? lr[t] : best[v];
            }
            if (lane == 0) io.osd_index[slot] = cw;
        }
        wave_lds_sync();
    }
}

template <int MPL, int NPL, int MAXDC>
void launch_variant(const BpCodeDev& code, const BpParams& p, const BpIo& io, int max_waves, hipStream_t s)
{
    const int e_pad = (code.E + 63) & ~63;
    const bool damp = p.damping != 0.0f;
    const size_t wave_bytes = sizeof(float) * static_cast<size_t>(2 * 64 * NPL + e_pad * (damp ? 2 : 1));
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
