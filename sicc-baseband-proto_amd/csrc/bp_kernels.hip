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
// Mapping: a team of TEAM wavefronts decodes one codeword (TEAM = 1 for short
// codes, up to 16 for n = 2048).  Team thread t owns check rows t, t + T, ...
// and variable nodes t, t + T, ... (T = 64 * TEAM).  Registers hold each
// thread's row edge lists, its rows' previous check messages (so
// q = APP - r_old needs no stored q) and its variables' LLR / APP /
// best-syndrome APP.  LDS holds only what crosses threads: APP[n] (gathered
// along rows) and the new check messages r[E] in column-major order (summed
// along columns).  The check update is branch-free: padded edges point at a
// FLT_MAX sentinel APP, min1/min2 use min/max, and the first-minimum test is
// |q| == min1 (equal values imply min2 == min1).  HBM traffic per codeword is
// the LLR read and the bit/flag writes.  Teams are persistent and stride over
// the batch; a single-wave team synchronizes with wave barriers only.
#include "kernels.hpp"

#include <cfloat>
#include <type_traits>
#include <stdexcept>
#include <string>

namespace sicc {

namespace {

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroup barrier ordering LDS only.  A multi-wave team exchanges nothing
// through global memory inside the decode loop, so its barriers need not wait
// for the next codeword's LLR loads or this codeword's output stores, as a
// full __syncthreads() would.
__device__ __forceinline__ void group_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int TEAM>
__device__ __forceinline__ void team_sync()
{
    if (TEAM == 1)
        wave_lds_sync();
    else
        group_lds_sync();
}

__global__ void zero_i32_kernel(int* p, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

// One workgroup per table entry; float4 moves when both ends are 16-byte aligned.
__global__ __launch_bounds__(256) void gather_rows_kernel(const GatherEntry* table)
{
    const GatherEntry e = table[blockIdx.x];
    const size_t len = static_cast<size_t>(e.count) * e.row_len;
    float* d = e.dst;
    const bool vec = ((reinterpret_cast<uintptr_t>(e.src) | reinterpret_cast<uintptr_t>(d)) & 15u) == 0 && (len & 3u) == 0;
    if (vec) {
        const float4* s4 = reinterpret_cast<const float4*>(e.src);
        float4* d4 = reinterpret_cast<float4*>(d);
        for (size_t i = threadIdx.x; i < len / 4; i += blockDim.x) d4[i] = s4[i];
    } else {
        for (size_t i = threadIdx.x; i < len; i += blockDim.x) d[i] = e.src[i];
    }
}

// Eight information bits (bytes of 0 / 1) to one byte, bit i from byte i.
// The 8-byte load is aligned; the two 4-byte halves fold to bits 0 and 4 of
// each byte, and one multiply gathers the four bytes' bits 0 / 4 into bits
// 24..31 (every partial product lands on its own bit: no carries).
__device__ __forceinline__ unsigned pack8_aligned(const uint8_t* p)
{
    const uint2 x = *reinterpret_cast<const uint2*>(p);
    const unsigned t = (x.x & 0x01010101u) | ((x.y & 0x01010101u) << 4);
    return (t * 0x01020408u) >> 24;
}

__device__ __forceinline__ unsigned pack_bits(const uint8_t* p, int nbits)
{
    unsigned b = 0;
    for (int i = 0; i < nbits; ++i) b |= (p[i] & 1u) << i;
    return b;
}

// Result packing for the server's copy back, all code groups of a launch in one
// kernel.  A group's output is one byte array: its rows' information bits,
// bit i of a row at byte i / 8, bit i % 8 (kb = ceil(k / 8) bytes a row), then
// its flag bytes.  Each thread builds four consecutive output bytes and writes
// them with one 4-byte store (group outputs start 4-byte aligned), so every
// wavefront writes 256 contiguous bytes of the pinned host results instead of
// byte-wide scattered stores across the bus.
__global__ __launch_bounds__(256) void pack_groups_kernel(const PackGroup* __restrict__ groups, int count)
{
    // the table may sit in host memory: fetch the group starts in one parallel
    // read into LDS instead of a chain of dependent reads over the bus
    constexpr int kLdsGroups = 64;
    __shared__ int first[kLdsGroups];
    int gi = 0;
    if (count <= kLdsGroups) {
        for (int i = threadIdx.x; i < count; i += blockDim.x) first[i] = groups[i].block0;
        __syncthreads();
        while (gi + 1 < count && static_cast<int>(blockIdx.x) >= first[gi + 1]) ++gi;
    } else {
        while (gi + 1 < count && static_cast<int>(blockIdx.x) >= groups[gi + 1].block0) ++gi;
    }
    const PackGroup g = groups[gi];
    const int kb = (g.k + 7) >> 3;
    const size_t packed = static_cast<size_t>(g.rows) * kb;
    const size_t total = packed + static_cast<size_t>(g.flag_bytes);
    const size_t q0 = (static_cast<size_t>(static_cast<int>(blockIdx.x) - g.block0) * 256 + threadIdx.x) * 4;
    if (q0 >= total) return;
    const bool vec = (g.k & 7) == 0 && (reinterpret_cast<uintptr_t>(g.bits) & 7u) == 0;
    size_t r = 0;
    int j = 0;
    if (q0 < packed) {
        r = q0 / static_cast<size_t>(kb);
        j = static_cast<int>(q0 - r * kb);
    }
    unsigned word = 0;
    const int nb = total - q0 < 4 ? static_cast<int>(total - q0) : 4;
    for (int b = 0; b < nb; ++b) {
        const size_t q = q0 + b;
        unsigned v;
        if (q < packed) {
            const uint8_t* p = g.bits + r * g.k + 8 * j;
            v = vec ? pack8_aligned(p) : pack_bits(p, min(8, g.k - 8 * j));
            if (++j == kb) {
                j = 0;
                ++r;
            }
        } else {
            v = g.flags[q - packed];
        }
        word |= v << (8 * b);
    }
    uint8_t* dst = g.out + q0;
    if (nb == 4) {
        *reinterpret_cast<unsigned*>(dst) = word;
    } else {
        for (int b = 0; b < nb; ++b) dst[b] = static_cast<uint8_t>(word >> (8 * b));
    }
}

// XOR of N words with three-input XORs (v_bitop3_b32, truth table 0x96):
// N - 1 two-input XORs become ceil((N - 1) / 2) instructions.
template <int N>
__device__ __forceinline__ unsigned xor_reduce(const unsigned (&v)[N])
{
    unsigned x = v[0];
    int j = 1;
#pragma unroll
    for (; j + 1 < N; j += 2) x = __builtin_amdgcn_bitop3_b32(x, v[j], v[j + 1], 0x96);
    if (j < N) x ^= v[j];
    return x;
}

// Smallest and second smallest (multiset order statistics: a repeated minimum
// is also the second) of N floats, none NaN.  Groups of three give
// (min3, med3), a pair (min, max), a lone value (v, FLT_MAX); the groups'
// minima recurse the same way, and the second smallest overall is
// min(second smallest of the minima, smallest of the seconds): for the group
// i holding the overall minimum, every other element is either another
// group's minimum or at least it, or a value of group i at least its second.
// For 8 values this is 10 v_min3 / v_med3 / v_min / v_max instructions in two
// dependent levels, against 18 in four for a pairwise tree.
template <int N>
__device__ __forceinline__ void min_two(const float* v, float& lo, float& se)
{
    constexpr int G = (N + 2) / 3;
    float gl[G], gs[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int b = 3 * g;
        const int c = N - b < 3 ? N - b : 3;
        if (c == 3) {
            gl[g] = fminf(fminf(v[b], v[b + 1]), v[b + 2]);
            gs[g] = __builtin_amdgcn_fmed3f(v[b], v[b + 1], v[b + 2]);
        } else if (c == 2) {
            gl[g] = fminf(v[b], v[b + 1]);
            gs[g] = fmaxf(v[b], v[b + 1]);
        } else {
            gl[g] = v[b];
            gs[g] = FLT_MAX;
        }
    }
    if constexpr (G == 1) {
        lo = gl[0];
        se = gs[0];
    } else {
        float ll, ls;
        min_two<G>(gl, ll, ls);
        float sm = gs[0];
#pragma unroll
        for (int g = 1; g < G; ++g) sm = fminf(sm, gs[g]);
        lo = ll;
        se = fminf(ls, sm);
    }
}

// a + r[0] + r[1] + ... + r[D-1], summed in that order, with every LDS read
// issued before the first add (immediate offsets: no per-term wait or address
// arithmetic).
template <int D>
__device__ __forceinline__ float col_sum(float a, const float* r)
{
    float x[D];
#pragma unroll
    for (int j = 0; j < D; ++j) x[j] = r[j];
#pragma unroll
    for (int j = 0; j < D; ++j) a += x[j];
    return a;
}

// One team's decode loop: team `gteam` of `nteams` takes codewords gteam,
// gteam + nteams, ... (or chunks from io.work_ctr).
// OFF = false: a build for launches whose offset is zero, where
// alpha * max(m - 0, 0) is alpha * m for every m >= +0.
template <int MPL, int NPL, int MAXDC, bool DAMP, int TEAM, bool OFF = true>
__device__ __forceinline__ void bp_flood_body(BpCodeDev code, BpParams p, BpIo io, int e_pad, int gteam, int nteams)
{
    extern __shared__ float lds[];
    constexpr int T = 64 * TEAM;     // threads per team
    constexpr int np = T * NPL;      // padded variable count
    const int tid = threadIdx.x % T;
    const int team = threadIdx.x / T;
    const int lane = threadIdx.x & 63;
    // per team: [8 reduction words] APP[np] + sentinel pad (4) + r[e_pad]
    float* tbase = lds + team * (8 + np + 4 + e_pad);
    int* red = reinterpret_cast<int*>(tbase);  // [0..1] syndrome counters, [2] OSD slot, [3] min|APP| bits
    float* app = tbase + 8;
    float* rcol = app + np + 4;

    const int n = code.n;
    const int m = code.m;
    // Padded edges read the sentinel APP and write the spare message slot e_pad-1,
    // so every row runs the same MAXDC-edge, branch-free update.
    const int pad_edge = (np << 16) | (e_pad - 1);
    int em[MPL][MAXDC];
#pragma unroll
    for (int s = 0; s < MPL; ++s) {
        const int r = tid + T * s;
#pragma unroll
        for (int j = 0; j < MAXDC; ++j) {
            const int e = (r < m && j < code.max_dc) ? code.row_meta[r * code.max_dc + j] : -1;
            em[s][j] = e >= 0 ? e : pad_edge;
        }
    }
    int cm[NPL];
    unsigned uni = 0u;  // bit t: all lanes of this wavefront hold a variable of one degree
#pragma unroll
    for (int t = 0; t < NPL; ++t) {
        const int v = tid + T * t;
        cm[t] = v < n ? code.col_meta[v] : 0;
        const int dg0 = __builtin_amdgcn_readfirstlane(cm[t] & 255);
        if (__ballot(v >= n || (cm[t] & 255) != dg0) == 0ull) uni |= 1u << t;
    }
    if (tid == 0) {
        app[np] = FLT_MAX;
        red[0] = red[1] = 0;
    }

    const float alpha = p.alpha;
    const float off = p.offset;
    const float damp = p.damping;

    // Codeword assignment: static striding, or chunks taken from a global
    // counter (work stealing) so teams that finish early take more work.
    // A single-wave team claims its next chunk when the current chunk's last
    // codeword starts (nb), so that chunk's first LLRs load while it decodes,
    // as under static striding.
    const bool steal = io.work_ctr != nullptr;
    const int chunk = io.chunk > 0 ? io.chunk : 1;
    int cw = gteam, lim = 0, nb = 0;
    if (steal) {
        int b = 0;
        if (TEAM == 1) {
            if (lane == 0) b = atomicAdd(io.work_ctr, chunk);
            b = __builtin_amdgcn_readlane(b, 0);
        } else {
            if (tid == 0) red[4] = atomicAdd(io.work_ctr, chunk);
            group_lds_sync();
            b = red[4];
        }
        cw = b;
        lim = min(b + chunk, io.num_cw);
    }
    // information column of bit tid + T * t, and the LLRs of the next codeword,
    // loaded while the current one decodes
    int icol[NPL];
    float lrn[NPL];
#pragma unroll
    for (int t = 0; t < NPL; ++t) {
        const int i = tid + T * t;
        icol[t] = i < code.k ? code.info_cols[i] : 0;
        lrn[t] = 0.0f;
    }
    if (cw < io.num_cw) {
        const float* L = io.llr + static_cast<size_t>(cw) * n;
#pragma unroll
        for (int t = 0; t < NPL; ++t) {
            const int v = tid + T * t;
            lrn[t] = v < n ? L[v] : 0.0f;
        }
    }
    while (cw < io.num_cw) {
        float lr[NPL], appr[NPL], bestr[NPL];
#pragma unroll
        for (int t = 0; t < NPL; ++t) lr[t] = lrn[t];
        {
            int nx = cw + nteams;
            if (steal) {
                nx = cw + 1 < lim ? cw + 1 : io.num_cw;
                if (TEAM == 1 && cw + 1 >= lim) {
                    int b = 0;
                    if (lane == 0) b = atomicAdd(io.work_ctr, chunk);
                    nb = __builtin_amdgcn_readlane(b, 0);
                    nx = nb;
                }
            }
            if (nx < io.num_cw) {
                const float* Ln = io.llr + static_cast<size_t>(nx) * n;
#pragma unroll
                for (int t = 0; t < NPL; ++t) {
                    const int v = tid + T * t;
                    lrn[t] = v < n ? Ln[v] : 0.0f;
                }
            }
        }
#pragma unroll
        for (int t = 0; t < NPL; ++t) {
            const int v = tid + T * t;
            appr[t] = lr[t];
            bestr[t] = lr[t];
            app[v] = lr[t] + 0.0f;  // LDS copy canonical (-0 -> +0): its sign bit is (APP < 0)
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
        team_sync<TEAM>();

        int best_w = 0;
        bool ok = false;
        int it = 0;
        for (;;) {
            // ---- check-node gather and the syndrome of the current APP ----
            // The syndrome decides first: a pass that ends the decode (zero syndrome,
            // or the last iteration) never needs q or the row minima, which are most
            // of a row's check-node instructions.
            int synw = 0;
            float ga[MPL][MAXDC];
#pragma unroll
            for (int s = 0; s < MPL; ++s) {
                unsigned hb[MAXDC];
#pragma unroll
                for (int j = 0; j < MAXDC; ++j) {
                    ga[s][j] = app[em[s][j] >> 16];
                    hb[j] = __float_as_uint(ga[s][j]);  // canonical APP: sign bit == (APP < 0)
                }
                synw += __popcll(__ballot((xor_reduce(hb) >> 31) != 0u));
            }
            if (TEAM > 1) {
                // team-wide syndrome weight through an LDS counter, double-buffered by iteration
                int* ctr = red + (it & 1);
                if (lane == 0 && synw) atomicAdd(ctr, synw);
                group_lds_sync();
                synw = *ctr;
                if (tid == 0) red[(it + 1) & 1] = 0;
            }
            if (it > 0) {
                if (synw < best_w || synw == 0) {
                    best_w = synw;
#pragma unroll
                    for (int t = 0; t < NPL; ++t) bestr[t] = appr[t];
                }
                if (synw == 0) {
                    ok = true;
                    break;  // the messages of this pass would never be used
                }
            } else {
                best_w = synw;
            }
            if (it == p.max_iter) break;
            // ---- q, min1 / min2 and the sign product of every row ----
            float qv[MPL][MAXDC];
            float m1s[MPL], m2s[MPL];
            unsigned sps[MPL];
#pragma unroll
            for (int s = 0; s < MPL; ++s) {
                unsigned qb[MAXDC];
                float av[MAXDC];
#pragma unroll
                for (int j = 0; j < MAXDC; ++j) {
                    // The LDS APP is never -0, so a - r_old is never -0 either (x - y is
                    // -0 only for x = -0, y = +0): q's sign bit is (q < 0) as in the
                    // reference, whose q can differ only in the sign of a zero.  Damped
                    // q can underflow to -0 and is canonicalized.
                    float q = ga[s][j] - rold[s][j];
                    if (DAMP) {
                        if (it > 0) q = (1.0f - damp) * q + damp * qold[s][j];
                        qold[s][j] = q;
                        q = q + 0.0f;
                    }
                    qv[s][j] = q;
                    qb[j] = __float_as_uint(q);
                    av[j] = fabsf(q);
                }
                // smallest and second smallest |q|: exact order statistics, capped at
                // FLT_MAX as a sequential scan from (FLT_MAX, FLT_MAX) leaves them (a
                // padded edge's |q| can overflow to +inf)
                float lo, se;
                min_two<MAXDC>(av, lo, se);
                m1s[s] = fminf(lo, FLT_MAX);
                m2s[s] = fminf(se, FLT_MAX);
                sps[s] = xor_reduce(qb);
            }
            // ---- check-node messages ----
#pragma unroll
            for (int s = 0; s < MPL; ++s) {
                // The first minimum edge takes min2; any other edge equal to min1
                // implies min2 == min1, so |q| == min1 selects the same magnitude.
                const float m1 = m1s[s];
                // with off == 0, m - 0 == m and max(m, 0) == m for every m >= +0, so
                // the offset needs no test
                const float a1 = OFF ? alpha * fmaxf(m1 - off, 0.0f) : alpha * m1;
                const float a2 = OFF ? alpha * fmaxf(m2s[s] - off, 0.0f) : alpha * m2s[s];
                // keep the two products per row (not one per edge after the select)
                float b1 = a1, b2 = a2;
                asm volatile("" : "+v"(b1), "+v"(b2));
                // the row's sign product folded into both magnitudes once (they are
                // >= +0, so XOR of a sign bit is OR); per edge only q's own sign is
                // XORed in: one three-input op (v_bitop3) instead of XOR + AND-OR
                const unsigned sg = sps[s] & 0x80000000u;
                unsigned b1s = __float_as_uint(b1) ^ sg;
                unsigned b2s = __float_as_uint(b2) ^ sg;
                asm volatile("" : "+v"(b1s), "+v"(b2s));  // keep the fold (not re-split per edge)
#pragma unroll
                for (int j = 0; j < MAXDC; ++j) {
                    const unsigned mag = (fabsf(qv[s][j]) == m1) ? b2s : b1s;
                    const float r = __uint_as_float(mag ^ (__float_as_uint(qv[s][j]) & 0x80000000u));
                    rold[s][j] = r;
                    rcol[em[s][j] & 0xffff] = r;
                }
            }
            team_sync<TEAM>();
            // ---- variable-node update: APP = LLR + sum r, increasing row order ----
#pragma unroll
            for (int t = 0; t < NPL; ++t) {
                const int v = tid + T * t;
                const int st = cm[t] >> 8;
                if (uni & (1u << t)) {
                    // every lane of the wavefront has this degree: scalar trip count
                    const int dgs = __builtin_amdgcn_readfirstlane(cm[t] & 255);
                    float a = lr[t];
                    const float* rc = rcol + st;
                    // common degrees fully unrolled (14 % fewer VALU instructions per
                    // (128,64) codeword than the generic loop, 1.5 % less time)
                    switch (dgs) {
                    case 1: a = col_sum<1>(a, rc); break;
                    case 2: a = col_sum<2>(a, rc); break;
                    case 3: a = col_sum<3>(a, rc); break;
                    case 4: a = col_sum<4>(a, rc); break;
                    case 5: a = col_sum<5>(a, rc); break;
                    case 6: a = col_sum<6>(a, rc); break;
                    default:
#pragma unroll 4
                        for (int j = 0; j < dgs; ++j) a += rc[j];
                    }
                    appr[t] = a;
                    app[v] = a + 0.0f;
                } else if (v < n) {
                    float a = lr[t];
                    const int dg = cm[t] & 255;
                    for (int j = 0; j < dg; ++j) a += rcol[st + j];
                    appr[t] = a;
                    app[v] = a + 0.0f;
                }
            }
            team_sync<TEAM>();
            ++it;
        }
        if (TEAM > 1) {
            // the counter the last check pass used must read zero for the next codeword
            group_lds_sync();
            if (tid == 0) red[0] = red[1] = 0;
        }

        // ---- outputs ----
        bool accept = ok;
        if (ok && p.min_abs_accept > 0.0f) {
            float mn = FLT_MAX;
#pragma unroll
            for (int t = 0; t < NPL; ++t)
                if (tid + T * t < n) mn = fminf(mn, fabsf(appr[t]));
            // Through an LDS minimum (non-negative floats order as their bit patterns):
            // a lane-shuffle tree would keep its lane addresses in registers for the
            // whole kernel.
            if (tid == 0) red[3] = __float_as_int(FLT_MAX);
            team_sync<TEAM>();
            atomicMin(red + 3, __float_as_int(mn));
            team_sync<TEAM>();
            mn = __int_as_float(red[3]);
            accept = mn >= p.min_abs_accept;
        }
        if (io.code_bits || io.app) {
#pragma unroll
            for (int t = 0; t < NPL; ++t) {
                const int v = tid + T * t;
                if (v < n) {
                    if (io.code_bits) io.code_bits[static_cast<size_t>(cw) * n + v] = appr[t] < 0.0f ? 1 : 0;
                    if (io.app) io.app[static_cast<size_t>(cw) * n + v] = appr[t];
                }
            }
        }
        if (io.info_bits) {
            const int k = code.k;
#pragma unroll
            for (int t = 0; t < NPL; ++t) {
                const int i = tid + T * t;
                if (i < k) io.info_bits[static_cast<size_t>(cw) * k + i] = app[icol[t]] < 0.0f ? 1 : 0;
            }
        }
        if (tid == 0) {
            if (io.success) io.success[cw] = accept ? 1 : 0;
            if (io.iterations) io.iterations[cw] = it;
            if (io.used_osd) io.used_osd[cw] = (!accept && io.osd_llr) ? 1 : 0;
            if (io.distance && accept) io.distance[cw] = 0.0f;
        }
        if (!accept && io.osd_llr) {
            int slot = 0;
            if (TEAM == 1) {
                if (lane == 0) slot = atomicAdd(io.osd_count, 1);
                slot = __builtin_amdgcn_readlane(slot, 0);
            } else {
                if (tid == 0) red[2] = atomicAdd(io.osd_count, 1);
                group_lds_sync();
                slot = red[2];
            }
            float* dst = io.osd_llr + static_cast<size_t>(slot) * n;
            // a converged-but-rejected frame hands over its final APP (== best)
#pragma unroll
            for (int t = 0; t < NPL; ++t) {
                const int v = tid + T * t;
                if (v < n) dst[v] = bestr[t];
            }
            if (tid == 0) io.osd_index[slot] = cw;
        }
        team_sync<TEAM>();
        if (!steal) {
            cw += nteams;
        } else if (TEAM == 1 && ++cw >= lim) {
            // the chunk claimed ahead; its first LLRs are already in lrn
            cw = nb;
            lim = min(nb + chunk, io.num_cw);
        } else if (TEAM > 1 && ++cw >= lim) {
            int b = 0;
            {
                if (tid == 0) red[4] = atomicAdd(io.work_ctr, chunk);
                group_lds_sync();
                b = red[4];
                group_lds_sync();
            }
            cw = b;
            lim = min(b + chunk, io.num_cw);
            if (cw < io.num_cw) {
                const float* L = io.llr + static_cast<size_t>(cw) * n;
#pragma unroll
                for (int t = 0; t < NPL; ++t) {
                    const int v = tid + T * t;
                    lrn[t] = v < n ? L[v] : 0.0f;
                }
            }
        }
    }
}

// Team index and team count of a one-batch launch (teams of T threads).
template <int T>
__device__ __forceinline__ int grid_team() { return static_cast<int>(blockIdx.x * (blockDim.x / T) + threadIdx.x / T); }
template <int T>
__device__ __forceinline__ int grid_teams() { return static_cast<int>(gridDim.x * (blockDim.x / T)); }

// Teams of 8 and 16 waves with one row a thread are held to 64 VGPRs (eight
// waves per SIMD, two workgroups per CU) up to the row-degree classes where the
// few spilled registers cost less than the second workgroup gains: 16-wave
// teams through 14 edges, 8-wave teams through 12 (32 768 codewords, n = 2048
// rate 11/16: 2.29 -> 1.65 ms, rate 3/4: 2.22 -> 2.00 ms; n = 1024 rate 1/2:
// 0.94 -> 0.85 ms; the 8-wave, 14-edge instance was slower capped).
template <int MPL, int MAXDC, int TEAM>
constexpr int team_waves_per_eu()
{
    return MPL == 1 && ((TEAM == 16 && MAXDC <= 14) || (TEAM == 8 && MAXDC <= 12)) ? 8 : 1;
}

template <int MPL, int NPL, int MAXDC, bool DAMP, int TEAM, bool OFF>
__global__ __launch_bounds__(TEAM == 1 ? 256 : 64 * TEAM) __attribute__((amdgpu_waves_per_eu(
    team_waves_per_eu<MPL, MAXDC, TEAM>()))) void bp_flood_kernel(BpCodeDev code, BpParams p, BpIo io, int e_pad)
{
    bp_flood_body<MPL, NPL, MAXDC, DAMP, TEAM, OFF>(code, p, io, e_pad, grid_team<64 * TEAM>(),
                                                    grid_teams<64 * TEAM>());
}

// Several batches, of codes that share one instance of the decode loop, in one
// launch: segment i owns blocks [block0, block0 of segment i + 1) and its
// `teams` teams stride over its own batch with its own code, parameters and
// outputs.  The per-codeword arithmetic is the one-batch kernel's.
struct BpSegment {
    BpCodeDev code;
    BpParams p;
    BpIo io;
    int e_pad;
    int block0;
    int teams;
};
constexpr int kBpSegmentsPerLaunch = 8;
struct BpSegmentTable {
    BpSegment seg[kBpSegmentsPerLaunch];
    int count;
};

template <int MPL, int NPL, int MAXDC, bool DAMP, int TEAM, bool OFF>
__global__ __launch_bounds__(TEAM == 1 ? 256 : 64 * TEAM) __attribute__((amdgpu_waves_per_eu(
    team_waves_per_eu<MPL, MAXDC, TEAM>()))) void bp_flood_multi_kernel(BpSegmentTable t)
{
    constexpr int T = 64 * TEAM;
    int si = 0;
    while (si + 1 < t.count && static_cast<int>(blockIdx.x) >= t.seg[si + 1].block0) ++si;
    const BpSegment& g = t.seg[si];
    const int gteam = static_cast<int>((blockIdx.x - g.block0) * (blockDim.x / T) + threadIdx.x / T);
    bp_flood_body<MPL, NPL, MAXDC, DAMP, TEAM, OFF>(g.code, g.p, g.io, g.e_pad, gteam, g.teams);
}

// The single-wave shape of codes up to 128 variables and row degree 8 without
// damping (the (128,64) class): its vector registers fit eight waves per SIMD,
// but the kernel arguments held in scalar registers cap residency at seven.
// Two ways to the eighth wave:
//  * bp_flood_kernel_lean takes only the launch's hard-decision outputs (and
//    the work-stealing cursor); with the soft outputs' code gone its arguments
//    fit the scalar budget outright;
//  * bp_flood_kernel_w8 (any outputs) caps the scalar budget, spilling some
//    arguments to vector lanes.
struct BpIoLean {
    const float* llr;
    int num_cw;
    int chunk;
    uint8_t* info_bits;
    uint8_t* success;
    int* iterations;
    uint8_t* used_osd;
    float* osd_llr;
    int* osd_index;
    int* osd_count;
    int* work_ctr;
};

// STEAL: the build with the work-stealing cursor needs a few more registers;
// it is held to eight waves per SIMD (a 12-byte spill), the static build is not
// constrained (62 registers, no spill).
template <bool STEAL, bool OFF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(STEAL ? 8 : 1))) void bp_flood_kernel_lean(
    BpCodeDev code, BpParams p, BpIoLean l, int e_pad)
{
    BpIo io;
    io.llr = l.llr;
    io.num_cw = l.num_cw;
    io.info_bits = l.info_bits;
    io.success = l.success;
    io.iterations = l.iterations;
    io.used_osd = l.used_osd;
    io.osd_llr = l.osd_llr;
    io.osd_index = l.osd_index;
    io.osd_count = l.osd_count;
    io.work_ctr = STEAL ? l.work_ctr : nullptr;
    io.chunk = l.chunk;
    bp_flood_body<1, 2, 8, false, 1, OFF>(code, p, io, e_pad, grid_team<64>(), grid_teams<64>());
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void bp_flood_kernel_w8(BpCodeDev code,
                                                                                                  BpParams p,
                                                                                                  BpIo io, int e_pad)
{
    bp_flood_body<1, 2, 8, false, 1>(code, p, io, e_pad, grid_team<64>(), grid_teams<64>());
}

template <int MPL, int NPL, int MAXDC, int TEAM>
void launch_variant(const BpCodeDev& code, const BpParams& p, const BpIo& io, int max_waves, int tpb_req,
                    hipStream_t s)
{
    const bool damp = p.damping != 0.0f;
    constexpr int T = 64 * TEAM;
    const int e_pad = (code.E + 1 + 63) & ~63;  // one spare slot for padded edges
    const size_t team_bytes = sizeof(float) * static_cast<size_t>(8 + T * NPL + 4 + e_pad);
    // Up to 4 single-wave teams per group; a multi-wave team synchronizes with
    // group barriers, so it must be alone in its group.  Within 64 KiB of LDS.
    int tpb = TEAM == 1 ? (tpb_req >= 4 ? 4 : (tpb_req >= 2 ? 2 : 1)) : 1;
    while (tpb > 1 && team_bytes * tpb > 64 * 1024) --tpb;
    if (team_bytes * tpb > 64 * 1024) throw std::invalid_argument("code too large for the LDS-resident BP kernel");
    const size_t lds = team_bytes * tpb;
    // damped launches take the general build (offset applied as given); undamped
    // ones a build with or without the offset clamp
    const bool off = p.offset != 0.0f;
    auto kern = damp  ? bp_flood_kernel<MPL, NPL, MAXDC, true, TEAM, true>
                : off ? bp_flood_kernel<MPL, NPL, MAXDC, false, TEAM, true>
                      : bp_flood_kernel<MPL, NPL, MAXDC, false, TEAM, false>;
    bool lean = false;
    if constexpr (MPL == 1 && NPL == 2 && MAXDC == 8 && TEAM == 1) {
        if (!damp) {
            lean = !io.code_bits && !io.app && !io.distance;
            kern = bp_flood_kernel_w8;
        }
    }
    auto lean_kern = io.work_ctr ? (off ? bp_flood_kernel_lean<true, true> : bp_flood_kernel_lean<true, false>)
                                 : (off ? bp_flood_kernel_lean<false, true> : bp_flood_kernel_lean<false, false>);
    const void* kfn = lean ? reinterpret_cast<const void*>(lean_kern) : reinterpret_cast<const void*>(kern);
    // A persistent grid larger than what the GPU holds at once would leave its
    // last teams to start after the first ones finished all their work.
    const long resident_teams = static_cast<long>(resident_blocks_per_cu(kfn, T * tpb, lds)) * device_cu_count() * tpb;
    long teams = io.num_cw < max_waves / TEAM ? io.num_cw : max_waves / TEAM;
    if (teams > resident_teams) teams = resident_teams;
    if (teams < 1) teams = 1;
    const int blocks = static_cast<int>((teams + tpb - 1) / tpb);
    if (lean) {
        const BpIoLean l{io.llr,     io.num_cw,   io.chunk,     io.info_bits, io.success, io.iterations,
                         io.used_osd, io.osd_llr, io.osd_index, io.osd_count, io.work_ctr};
        hipLaunchKernelGGL(lean_kern, dim3(blocks), dim3(T * tpb), lds, s, code, p, l, e_pad);
    } else {
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(T * tpb), lds, s, code, p, io, e_pad);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("bp_flood_kernel launch: ") + hipGetErrorString(e));
}

// Up to kBpSegmentsPerLaunch batches of one instance in one launch.  Each
// batch gets the blocks its codewords (capped by its max_waves) need; when
// they exceed what the GPU holds at once, the resident blocks are shared in
// proportion, so the batches' teams all start together and each strides over
// its own batch.
template <int MPL, int NPL, int MAXDC, int TEAM>
void launch_multi_variant(const BpBatch* b, int count, hipStream_t s)
{
    constexpr int T = 64 * TEAM;
    const bool damp = b[0].p.damping != 0.0f;
    int e_pad[kBpSegmentsPerLaunch];
    int e_max = 0;
    for (int i = 0; i < count; ++i) {
        e_pad[i] = (b[i].code.E + 1 + 63) & ~63;
        e_max = e_pad[i] > e_max ? e_pad[i] : e_max;
    }
    const size_t team_bytes = sizeof(float) * static_cast<size_t>(8 + T * NPL + 4 + e_max);
    const int tpb_req = b[0].tune.teams_per_block;
    int tpb = TEAM == 1 ? (tpb_req >= 4 ? 4 : (tpb_req >= 2 ? 2 : 1)) : 1;
    while (tpb > 1 && team_bytes * tpb > 64 * 1024) --tpb;
    if (team_bytes * tpb > 64 * 1024) throw std::invalid_argument("code too large for the LDS-resident BP kernel");
    const size_t lds = team_bytes * tpb;
    const bool off = b[0].p.offset != 0.0f;
    auto kern = damp  ? bp_flood_multi_kernel<MPL, NPL, MAXDC, true, TEAM, true>
                : off ? bp_flood_multi_kernel<MPL, NPL, MAXDC, false, TEAM, true>
                      : bp_flood_multi_kernel<MPL, NPL, MAXDC, false, TEAM, false>;
    const long resident =
        static_cast<long>(resident_blocks_per_cu(reinterpret_cast<const void*>(kern), T * tpb, lds)) * device_cu_count();
    long need[kBpSegmentsPerLaunch];
    long total_need = 0;
    for (int i = 0; i < count; ++i) {
        const int mw = b[i].max_waves / TEAM;
        long teams = b[i].io.num_cw < mw ? b[i].io.num_cw : mw;
        if (teams < 1) teams = 1;
        need[i] = (teams + tpb - 1) / tpb;
        total_need += need[i];
    }
    BpSegmentTable t{};
    t.count = count;
    long blocks = 0;
    for (int i = 0; i < count; ++i) {
        long nb = need[i];
        if (total_need > resident && resident > 0) nb = need[i] * resident / total_need;
        if (nb < 1) nb = 1;
        BpSegment& g = t.seg[i];
        g.code = b[i].code;
        g.p = b[i].p;
        g.io = b[i].io;
        g.io.work_ctr = nullptr;
        g.e_pad = e_pad[i];
        g.block0 = static_cast<int>(blocks);
        g.teams = static_cast<int>(nb * tpb);
        blocks += nb;
    }
    if (blocks > 0x7fffffffL) throw std::invalid_argument("fused BP launch too large");
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(blocks)), dim3(T * tpb), lds, s, t);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("bp_flood_multi_kernel launch: ") + hipGetErrorString(e));
}

template <int MPL, int NPL, int MAXDC>
void launch_multi_team(int team, const BpBatch* b, int count, hipStream_t s)
{
    switch (team) {
        case 1: return launch_multi_variant<MPL, NPL, MAXDC, 1>(b, count, s);
        case 2: return launch_multi_variant<MPL, NPL, MAXDC, 2>(b, count, s);
        case 4: return launch_multi_variant<MPL, NPL, MAXDC, 4>(b, count, s);
        case 8: return launch_multi_variant<MPL, NPL, MAXDC, 8>(b, count, s);
        case 16: return launch_multi_variant<MPL, NPL, MAXDC, 16>(b, count, s);
        default: throw std::invalid_argument("BP team size must be 1, 2, 4, 8 or 16 waves");
    }
}

template <int MPL, int NPL, int MAXDC>
void launch_team(int team, int tpb, const BpCodeDev& code, const BpParams& p, const BpIo& io, int max_waves,
                 hipStream_t s)
{
    switch (team) {
        case 1: return launch_variant<MPL, NPL, MAXDC, 1>(code, p, io, max_waves, tpb, s);
        case 2: return launch_variant<MPL, NPL, MAXDC, 2>(code, p, io, max_waves, tpb, s);
        case 4: return launch_variant<MPL, NPL, MAXDC, 4>(code, p, io, max_waves, tpb, s);
        case 8: return launch_variant<MPL, NPL, MAXDC, 8>(code, p, io, max_waves, tpb, s);
        case 16: return launch_variant<MPL, NPL, MAXDC, 16>(code, p, io, max_waves, tpb, s);
        default: throw std::invalid_argument("BP team size must be 1, 2, 4, 8 or 16 waves");
    }
}

// Row-degree bound of the kernel instance for a code: its largest row degree
// rounded up to 8, 10, 12, 14 or 16 (the check-node loop runs that many edge
// slots per row, padding included, so a code whose rows reach 10 edges would do
// 60 % more check-node work on the 16-slot instance).
int dc_class(int max_dc) { return max_dc <= 8 ? 8 : max_dc <= 10 ? 10 : max_dc <= 12 ? 12 : max_dc <= 14 ? 14 : 16; }

// Calls f(rows per thread, variables per thread, row-degree bound) with the three
// as compile-time constants: (1, 2) for one row and two variables a thread, else
// (2, 4).
template <class F>
void with_dc_class(int dc, bool small, F&& f)
{
    using std::integral_constant;
    auto go = [&](auto maxdc) {
        if (small)
            f(integral_constant<int, 1>{}, integral_constant<int, 2>{}, maxdc);
        else
            f(integral_constant<int, 2>{}, integral_constant<int, 4>{}, maxdc);
    };
    switch (dc) {
        case 8: return go(integral_constant<int, 8>{});
        case 10: return go(integral_constant<int, 10>{});
        case 12: return go(integral_constant<int, 12>{});
        case 14: return go(integral_constant<int, 14>{});
        case 16: return go(integral_constant<int, 16>{});
        default: throw std::invalid_argument("BP row-degree class must be 8, 10, 12, 14 or 16");
    }
}

// Rows and variables per thread for a team of `team` waves.
void per_thread(const BpCodeDev& c, int team, int* mpl, int* npl)
{
    const int T = 64 * team;
    *mpl = (c.m + T - 1) / T;
    *npl = (c.n + T - 1) / T;
}

}  // namespace

void launch_zero_i32(int* p, int count, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(zero_i32_kernel, dim3((count + 255) / 256), dim3(256), 0, s, p, count);
}

void launch_gather_rows(const GatherEntry* table, int entries, hipStream_t s)
{
    if (entries <= 0) return;
    hipLaunchKernelGGL(gather_rows_kernel, dim3(entries), dim3(256), 0, s, table);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("gather_rows_kernel launch: ") + hipGetErrorString(e));
}

void launch_pack_groups(PackGroup* groups, int count, hipStream_t s)
{
    if (count <= 0) return;
    long blocks = 0;
    for (int i = 0; i < count; ++i) {
        PackGroup& g = groups[i];
        if (g.rows <= 0 || g.k <= 0 || g.flag_bytes < 0) throw std::invalid_argument("pack_groups: bad shape");
        if (reinterpret_cast<uintptr_t>(g.out) & 3u) throw std::invalid_argument("pack_groups: output not 4-byte aligned");
        const long bytes = static_cast<long>(g.rows) * ((g.k + 7) / 8) + g.flag_bytes;
        g.block0 = static_cast<int>(blocks);
        g.pack_blocks = static_cast<int>((bytes + 1023) / 1024);  // 4 bytes a thread
        blocks += g.pack_blocks;
        if (blocks > 0x7fffffffL) throw std::invalid_argument("pack_groups: launch too large");
    }
    hipLaunchKernelGGL(pack_groups_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, groups, count);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("pack_groups_kernel launch: ") + hipGetErrorString(e));
}

bool bp_team_supported(const BpCodeDev& c, int team)
{
    if (c.n <= 0 || c.m <= 0 || c.E <= 0 || c.E > 0xfffe || c.n > 0x7fff || c.max_dc > 16) return false;
    if (team != 1 && team != 2 && team != 4 && team != 8 && team != 16) return false;
    int mpl = 0, npl = 0;
    per_thread(c, team, &mpl, &npl);
    if (mpl > 2 || npl > 4) return false;
    const bool small = mpl <= 1 && npl <= 2;
    const int e_pad = (c.E + 1 + 63) & ~63;
    const size_t np = static_cast<size_t>(64) * team * (small ? 2 : 4);
    return sizeof(float) * (8 + np + 4 + static_cast<size_t>(e_pad)) <= 64 * 1024;
}

int bp_default_team(const BpCodeDev& c)
{
    for (int team : {1, 2, 4, 8, 16}) {
        int mpl = 0, npl = 0;
        per_thread(c, team, &mpl, &npl);
        if (mpl <= 1 && npl <= 2 && bp_team_supported(c, team)) return team;
    }
    for (int team : {1, 2, 4, 8, 16})
        if (bp_team_supported(c, team)) return team;
    return 0;
}

bool bp_shape_supported(const BpCodeDev& c) { return bp_default_team(c) != 0; }

int bp_flood_instance(const BpCodeDev& c, const BpParams& p, const BpTune& tune)
{
    const int team = tune.team > 0 ? tune.team : bp_default_team(c);
    if (!bp_team_supported(c, team) || p.max_iter < 1) return -1;
    int mpl = 0, npl = 0;
    per_thread(c, team, &mpl, &npl);
    const bool small = mpl <= 1 && npl <= 2;
    const int tpb = team == 1 ? (tune.teams_per_block >= 4 ? 4 : (tune.teams_per_block >= 2 ? 2 : 1)) : 1;
    const bool damp = p.damping != 0.0f;
    const bool off = !damp && p.offset != 0.0f;  // undamped launches have builds with and without the offset
    return team << 12 | tpb << 8 | dc_class(c.max_dc) << 3 | (small ? 4 : 0) | (off ? 2 : 0) | (damp ? 1 : 0);
}

void launch_bp_flood_multi(const BpBatch* batches, int count, hipStream_t s)
{
    if (count <= 0) return;
    const int key = bp_flood_instance(batches[0].code, batches[0].p, batches[0].tune);
    if (key < 0) throw std::invalid_argument("code shape not supported by the BP kernel");
    for (int i = 0; i < count; ++i) {
        if (bp_flood_instance(batches[i].code, batches[i].p, batches[i].tune) != key)
            throw std::invalid_argument("fused BP batches must share one kernel instance");
        if (batches[i].io.num_cw <= 0 || !batches[i].io.llr) throw std::invalid_argument("fused BP batch is empty");
        if (batches[i].io.work_ctr) throw std::invalid_argument("fused BP batches use static striding");
    }
    const int team = key >> 12;
    const bool small = (key & 4) != 0;
    const int dc = (key >> 3) & 31;
    for (int c0 = 0; c0 < count; c0 += kBpSegmentsPerLaunch) {
        const int c = count - c0 < kBpSegmentsPerLaunch ? count - c0 : kBpSegmentsPerLaunch;
        const BpBatch* b = batches + c0;
        with_dc_class(dc, small, [&](auto mpl, auto npl, auto maxdc) {
            launch_multi_team<decltype(mpl)::value, decltype(npl)::value, decltype(maxdc)::value>(team, b, c, s);
        });
    }
}

void launch_bp_flood(const BpCodeDev& code, const BpParams& p, const BpIo& io, int max_waves, hipStream_t s,
                     const BpTune& tune)
{
    const int team = tune.team > 0 ? tune.team : bp_default_team(code);
    if (!bp_team_supported(code, team)) throw std::invalid_argument("code shape not supported by the BP kernel");
    if (io.num_cw <= 0) return;
    if (p.max_iter < 1) throw std::invalid_argument("max_iter must be >= 1");
    if (tune.max_waves > 0) max_waves = tune.max_waves;
    const int tpb = tune.teams_per_block;
    int mpl = 0, npl = 0;
    per_thread(code, team, &mpl, &npl);
    const bool small = mpl <= 1 && npl <= 2;
    with_dc_class(dc_class(code.max_dc), small, [&](auto mpl_c, auto npl_c, auto maxdc) {
        launch_team<decltype(mpl_c)::value, decltype(npl_c)::value, decltype(maxdc)::value>(team, tpb, code, p, io,
                                                                                           max_waves, s);
    });
}

}  // namespace sicc
