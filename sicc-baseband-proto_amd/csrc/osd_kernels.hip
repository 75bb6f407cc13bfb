// osd_kernels.hip -- bit-exact ordered-statistics decoding for gfx950.
//
// Semantics follow the MATLAB reference osd_decode / generator_to_systematic
// (source/matlab/uhd_cpp/matlab/cpu_bp_osd_decode.m:187-305):
//   1. stable sort of |L| descending, columns of G permuted accordingly;
//   2. Gaussian elimination to [I | P] where step t takes the first column
//      c >= t with a one in rows >= t, the first such row, swaps rows and
//      columns, then clears column t in every other row;
//   3. base message = hard decisions of the k most reliable independent
//      positions; test error patterns (TEPs) of order <= p over the
//      `max_flip` weakest of those positions (stable ascending sort), in
//      nchoosek order, capped at `max_tests`;
//   4. metric = sum of |L| over positions where the re-encoded word differs
//      from the hard decision, accumulated in double in ascending systematic
//      position order; strict-less keeps the first best TEP.
//
// Mapping: one wavefront per frame.  The generator is held column-major: lane l
// owns columns l, l+64, ... as one or two 64-bit words over the k <= 128 rows, so the
// pivot search is one ballot per column slot and a column swap is a lane
// exchange.  After elimination the parity part is transposed into LDS rows
// R[i] (ballots), and each lane evaluates whole TEPs: the disagreement
// vector of a TEP is D0 ^ R[f1] ^ R[f2] ..., summed in ascending position
// order so every lane reproduces the sequential CPU sum exactly.  The winner
// is the lexicographic minimum of (metric, test index) over the wave.
#include "kernels.hpp"

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

__device__ __forceinline__ long long nck(int n, int r)
{
    if (r < 0 || n < r) return 0;
    if (r == 0) return 1;
    if (r == 1) return n;
    if (r == 2) return static_cast<long long>(n) * (n - 1) / 2;
    return static_cast<long long>(n) * (n - 1) * (n - 2) / 6;
}

// Unrank test index t into flip-pool indices (lexicographic combinations,
// order blocks 0, 1, 2, 3 in sequence).  Returns the order.
__device__ __forceinline__ int unrank_tep(long long t, int pool, int order, int* idx)
{
    int o = 0;
    for (; o <= order; ++o) {
        const long long c = nck(pool, o);
        if (t < c) break;
        t -= c;
    }
    if (o == 1) {
        idx[0] = static_cast<int>(t);
        return 1;
    }
    if (o == 2) {
        // pairs (a, b), a < b, in lexicographic order: S(a) = a (2 pool - a - 1) / 2
        // pairs precede first index a; take the largest a with S(a) <= t
        const long long P = pool;
        auto S = [P](long long x) { return x * (2 * P - x - 1) / 2; };
        const double bq = static_cast<double>(2 * P - 1);
        long long a = static_cast<long long>((bq - sqrt(bq * bq - 8.0 * static_cast<double>(t))) * 0.5);
        if (a < 0) a = 0;
        if (a > P - 2) a = P - 2;
        while (a > 0 && S(a) > t) --a;
        while (a + 1 <= P - 2 && S(a + 1) <= t) ++a;
        idx[0] = static_cast<int>(a);
        idx[1] = static_cast<int>(a + 1 + (t - S(a)));
        return 2;
    }
    int prev = -1;
    for (int pos = 0; pos < o; ++pos) {
        int a = prev + 1;
        for (;; ++a) {
            const long long cnt = nck(pool - a - 1, o - pos - 1);
            if (t < cnt) break;
            t -= cnt;
        }
        idx[pos] = a;
        prev = a;
    }
    return o;
}

// Value of x in lane l, for a wave-uniform l: v_readlane into scalar registers,
// no LDS round trip (the elimination broadcasts a pivot column at every step).
__device__ __forceinline__ unsigned long long lane_u64(unsigned long long x, int l)
{
    const unsigned lo = __builtin_amdgcn_readlane(static_cast<unsigned>(x), l);
    const unsigned hi = __builtin_amdgcn_readlane(static_cast<unsigned>(x >> 32), l);
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// 64 x 64 bit-matrix transpose across a wavefront: on entry bit j of lane i's
// word is M[i][j]; on return it is M[j][i].  Six stages swap the off-diagonal
// blocks of halving size between lanes i and i ^ j.
// Lane i ^ j's value for j < 32: ds_swizzle in bit mode (and 0x1f, xor j), an
// immediate pattern, so no lane-address registers stay live.
template <int J>
__device__ __forceinline__ unsigned long long swz_xor(unsigned long long x)
{
    constexpr int pat = 0x1f | (J << 10);
    const int lo = __builtin_amdgcn_ds_swizzle(static_cast<int>(static_cast<unsigned>(x)), pat);
    const int hi = __builtin_amdgcn_ds_swizzle(static_cast<int>(static_cast<unsigned>(x >> 32)), pat);
    return (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) | static_cast<unsigned>(lo);
}

template <int J>
__device__ __forceinline__ unsigned long long transpose_stage(unsigned long long a, int lane, unsigned long long m)
{
    const unsigned long long b = J == 32 ? __shfl_xor(a, 32) : swz_xor<J & 31>(a);
    return (lane & J) == 0 ? (a & m) | ((b & m) << J) : (a & ~m) | ((b & ~m) >> J);
}

__device__ __forceinline__ unsigned long long transpose64(unsigned long long a, int lane)
{
    a = transpose_stage<32>(a, lane, 0x00000000FFFFFFFFull);
    a = transpose_stage<16>(a, lane, 0x0000FFFF0000FFFFull);
    a = transpose_stage<8>(a, lane, 0x00FF00FF00FF00FFull);
    a = transpose_stage<4>(a, lane, 0x0F0F0F0F0F0F0F0Full);
    a = transpose_stage<2>(a, lane, 0x3333333333333333ull);
    a = transpose_stage<1>(a, lane, 0x5555555555555555ull);
    return a;
}

constexpr int kMaxOrder = 3;
constexpr int kOsdWavesPerBlock = 4;

template <int CPL, int KW>
struct OsdLds {
    float llr[64 * CPL];
    double rel[64 * CPL];
    unsigned long long rows[64 * KW * CPL];  // R[i][s], i < k
    int rperm[64 * CPL];
    int pool[64 * KW];
    unsigned char code[64 * CPL];
};

// Bit r (< 64 KW) of a column held as KW words.
template <int KW>
__device__ __forceinline__ bool colbit(const unsigned long long* w, int r)
{
    return ((w[KW == 1 ? 0 : r >> 6] >> (r & 63)) & 1ull) != 0ull;
}

// Any one in rows >= t.
template <int KW>
__device__ __forceinline__ bool rows_from(const unsigned long long* w, int t)
{
    const int tw = KW == 1 ? 0 : t >> 6;  // t < k <= 64 KW
    bool any = false;
#pragma unroll
    for (int q = 0; q < KW; ++q)
        if (q > tw)
            any |= w[q] != 0ull;
        else if (q == tw)
            any |= (w[q] >> (t & 63)) != 0ull;
    return any;
}

template <int CPL, int KW>
__global__ __launch_bounds__(256) void osd_kernel(OsdCodeDev code, OsdParams p, OsdIo io)
{
    extern __shared__ unsigned char lds_raw[];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    OsdLds<CPL, KW>& sh = reinterpret_cast<OsdLds<CPL, KW>*>(lds_raw)[wv];
    const int n = code.n;
    const int k = code.k;
    const int total = io.count ? *io.count : io.host_count;
    const int gw = blockIdx.x * (blockDim.x >> 6) + wv;
    const int nw = gridDim.x * (blockDim.x >> 6);
    const int pool = k < p.max_flip ? k : p.max_flip;
    long long tests_total = 0;
    for (int o = 0; o <= p.order; ++o) tests_total += nck(pool, o);
    if (tests_total > p.max_tests) tests_total = p.max_tests;

    for (int f = gw; f < total; f += nw) {
        const int frame = io.index ? io.index[f] : f;
        const float* L = io.llr + static_cast<size_t>(f) * n;
        // 1. load and rank by |L| descending, stable
#pragma unroll
        for (int s = 0; s < CPL; ++s) {
            const int c = lane + 64 * s;
            sh.llr[c] = c < n ? L[c] : 0.0f;
        }
        wave_lds_sync();
#pragma unroll
        for (int s = 0; s < CPL; ++s) {
            const int c = lane + 64 * s;
            if (c < n) {
                const float a = fabsf(sh.llr[c]);
                int rank = 0;
                for (int j = 0; j < n; ++j) {
                    const float b = fabsf(sh.llr[j]);
                    rank += (b > a || (b == a && j < c)) ? 1 : 0;
                }
                sh.rperm[rank] = c;
            }
        }
        wave_lds_sync();
        // 2. column words of the permuted generator
        unsigned long long col[CPL][KW];
        int sp[CPL];
#pragma unroll
        for (int s = 0; s < CPL; ++s) {
            const int c = lane + 64 * s;
#pragma unroll
            for (int q = 0; q < KW; ++q) col[s][q] = c < n ? code.gen_cols[static_cast<size_t>(sh.rperm[c]) * KW + q] : 0ull;
            sp[s] = c;
        }
        int status = 0;
        for (int t = 0; t < k; ++t) {
            // t < k <= 64 KW: with one row word the word (and the column slot of
            // column t) is 0, which the compiler cannot infer on its own
            const int tw = KW == 1 ? 0 : t >> 6;
            const unsigned long long tbit = 1ull << (t & 63);
            int pc = -1;
#pragma unroll
            for (int s = 0; s < CPL; ++s) {
                const int c = lane + 64 * s;
                const bool cand = c >= t && c < n && rows_from<KW>(col[s], t);
                const unsigned long long b = __ballot(cand);
                if (pc < 0 && b != 0ull) pc = 64 * s + __ffsll(static_cast<long long>(b)) - 1;
            }
            if (pc < 0) {
                status = 1;
                break;
            }
            // pivot column words (uniform)
            unsigned long long wpc[KW];
#pragma unroll
            for (int q = 0; q < KW; ++q) wpc[q] = 0ull;
#pragma unroll
            for (int s = 0; s < CPL; ++s)
#pragma unroll
                for (int q = 0; q < KW; ++q)
                    if (s == (pc >> 6)) wpc[q] = lane_u64(col[s][q], pc & 63);
            // first row >= t with a one in the pivot column
            int prow = -1;
#pragma unroll
            for (int q = 0; q < KW; ++q) {
                if (prow >= 0 || q < tw) continue;
                const unsigned long long m = q == tw ? (wpc[q] >> (t & 63)) << (t & 63) : wpc[q];
                if (m != 0ull) prow = 64 * q + __ffsll(static_cast<long long>(m)) - 1;
            }
            if (prow != t) {
                const int pw = KW == 1 ? 0 : prow >> 6;
                const unsigned long long pbit = 1ull << (prow & 63);
#pragma unroll
                for (int s = 0; s < CPL; ++s) {
                    const bool b1 = (col[s][tw] & tbit) != 0ull, b2 = (col[s][pw] & pbit) != 0ull;
                    if (b1 != b2) {
                        col[s][tw] ^= tbit;
                        col[s][pw] ^= pbit;
                    }
                }
            }
            if (pc != t) {
                unsigned long long wt[KW], wp[KW];
                int st = 0, spc = 0;
#pragma unroll
                for (int q = 0; q < KW; ++q) wt[q] = wp[q] = 0ull;
#pragma unroll
                for (int s = 0; s < CPL; ++s) {
                    if (s == tw) {
#pragma unroll
                        for (int q = 0; q < KW; ++q) wt[q] = lane_u64(col[s][q], t & 63);
                        st = __builtin_amdgcn_readlane(sp[s], t & 63);
                    }
                    if (s == (pc >> 6)) {
#pragma unroll
                        for (int q = 0; q < KW; ++q) wp[q] = lane_u64(col[s][q], pc & 63);
                        spc = __builtin_amdgcn_readlane(sp[s], pc & 63);
                    }
                }
#pragma unroll
                for (int s = 0; s < CPL; ++s) {
                    const int c = lane + 64 * s;
                    if (c == t) {
#pragma unroll
                        for (int q = 0; q < KW; ++q) col[s][q] = wp[q];
                        sp[s] = spc;
                    } else if (c == pc) {
#pragma unroll
                        for (int q = 0; q < KW; ++q) col[s][q] = wt[q];
                        sp[s] = st;
                    }
                }
            }
            unsigned long long clr[KW];
#pragma unroll
            for (int q = 0; q < KW; ++q) clr[q] = 0ull;
#pragma unroll
            for (int s = 0; s < CPL; ++s)
#pragma unroll
                for (int q = 0; q < KW; ++q)
                    if (s == tw) clr[q] = lane_u64(col[s][q], t & 63);
            clr[tw] &= ~tbit;
#pragma unroll
            for (int s = 0; s < CPL; ++s)
                if (col[s][tw] & tbit)
#pragma unroll
                    for (int q = 0; q < KW; ++q) col[s][q] ^= clr[q];
        }
        if (status) {
            if (lane == 0 && io.status) io.status[frame] = 1;
            continue;
        }
        // 3. systematic hard decisions and reliabilities
        bool hs[CPL];
#pragma unroll
        for (int s = 0; s < CPL; ++s) {
            const int c = lane + 64 * s;
            if (c < n) {
                const float x = sh.llr[sh.rperm[sp[s]]];
                hs[s] = x < 0.0f;
                sh.rel[c] = static_cast<double>(fabsf(x));
            } else {
                hs[s] = false;
            }
        }
        // base message: the hard decisions of the k systematic positions (column
        // slots 0 .. KW-1 hold positions 0 .. 64 KW - 1)
        unsigned long long base[KW];
#pragma unroll
        for (int q = 0; q < KW; ++q) {
            const int lo = 64 * q;
            const unsigned long long m = k - lo >= 64 ? ~0ull : (k > lo ? ((1ull << (k - lo)) - 1ull) : 0ull);
            base[q] = __ballot(hs[q]) & m;
        }
        unsigned long long d0[CPL];
#pragma unroll
        for (int s = 0; s < CPL; ++s) {
            const int c = lane + 64 * s;
            const bool par = c >= k && c < n;
            int pb = 0;
#pragma unroll
            for (int q = 0; q < KW; ++q) pb ^= __popcll(base[q] & col[s][q]) & 1;
            d0[s] = __ballot(par && ((pb != 0) != hs[s]));
        }
        // parity rows by position: each 64 x 64 block (lane = position, bit = row)
        // transposed across the wave so that lane = row, bit = position
#pragma unroll
        for (int s = 0; s < CPL; ++s) {
            const int c = lane + 64 * s;
#pragma unroll
            for (int q = 0; q < KW; ++q) {
                const unsigned long long tr = transpose64(c >= k && c < n ? col[s][q] : 0ull, lane);
                const int r = 64 * q + lane;
                if (r < k) sh.rows[r * CPL + s] = tr;
            }
        }
        wave_lds_sync();
        // weak-order flip pool: stable ascending sort of rel[0..k)
        for (int c = lane; c < k; c += 64) {
            const double a = sh.rel[c];
            int rank = 0;
            for (int j = 0; j < k; ++j) {
                const double b = sh.rel[j];
                rank += (b < a || (b == a && j < c)) ? 1 : 0;
            }
            if (rank < pool) sh.pool[rank] = c;
        }
        wave_lds_sync();
        // 4. evaluate TEPs, one per lane at a time
        double best_m = 0.0;
        long long best_t = -1;
        for (long long t = lane; t < tests_total; t += 64) {
            int ix[kMaxOrder];
            const int o = unrank_tep(t, pool, p.order, ix);
            int pos[kMaxOrder];
            for (int i = 0; i < o; ++i) pos[i] = sh.pool[ix[i]];
            // ascending position order for the sum
            for (int i = 1; i < o; ++i)
                for (int j = i; j > 0 && pos[j - 1] > pos[j]; --j) {
                    const int x = pos[j];
                    pos[j] = pos[j - 1];
                    pos[j - 1] = x;
                }
            double metric = 0.0;
            unsigned long long d[CPL];
#pragma unroll
            for (int s = 0; s < CPL; ++s) d[s] = d0[s];
            for (int i = 0; i < o; ++i) {
                metric += sh.rel[pos[i]];
#pragma unroll
                for (int s = 0; s < CPL; ++s) d[s] ^= sh.rows[pos[i] * CPL + s];
            }
#pragma unroll
            for (int s = 0; s < CPL; ++s) {
                unsigned long long w = d[s];
                while (w) {
                    const int b = __ffsll(static_cast<long long>(w)) - 1;
                    metric += sh.rel[64 * s + b];
                    w &= w - 1ull;
                }
            }
            if (best_t < 0 || metric < best_m) {
                best_m = metric;
                best_t = t;
            }
        }
        // lexicographic (metric, index) minimum over the wave
        for (int off = 32; off > 0; off >>= 1) {
            const double om = __shfl_xor(best_m, off);
            const long long ot = __shfl_xor(best_t, off);
            const bool take = ot >= 0 && (best_t < 0 || om < best_m || (om == best_m && ot < best_t));
            if (take) {
                best_m = om;
                best_t = ot;
            }
        }
        // 5. rebuild the winner and un-permute
        {
            int ix[kMaxOrder];
            const int o = unrank_tep(best_t, pool, p.order, ix);
            unsigned long long fm[KW];
#pragma unroll
            for (int q = 0; q < KW; ++q) fm[q] = 0ull;
            unsigned long long d[CPL];
#pragma unroll
            for (int s = 0; s < CPL; ++s) d[s] = d0[s];
            for (int i = 0; i < o; ++i) {
                const int ps = sh.pool[ix[i]];
#pragma unroll
                for (int q = 0; q < KW; ++q)
                    if (q == (ps >> 6)) fm[q] |= 1ull << (ps & 63);
#pragma unroll
                for (int s = 0; s < CPL; ++s) d[s] ^= sh.rows[ps * CPL + s];
            }
#pragma unroll
            for (int s = 0; s < CPL; ++s) {
                const int c = lane + 64 * s;
                if (c < n) {
                    unsigned char bit;
                    if (c < k) {
                        unsigned long long w = 0ull;
#pragma unroll
                        for (int q = 0; q < KW; ++q)
                            if (q == s) w = base[q] ^ fm[q];
                        bit = static_cast<unsigned char>((w >> lane) & 1ull);
                    } else {
                        bit = static_cast<unsigned char>((hs[s] ? 1 : 0) ^ ((d[s] >> lane) & 1ull));
                    }
                    sh.code[sh.rperm[sp[s]]] = bit;
                }
            }
        }
        wave_lds_sync();
        if (io.code_bits)
            for (int c = lane; c < n; c += 64) io.code_bits[static_cast<size_t>(frame) * n + c] = sh.code[c];
        if (io.info_bits)
            for (int i = lane; i < k; i += 64)
                io.info_bits[static_cast<size_t>(frame) * k + i] = sh.code[code.info_cols[i]];
        if (lane == 0) {
            if (io.distance) io.distance[frame] = static_cast<float>(best_m);
            if (io.used_osd) io.used_osd[frame] = 1;
            if (io.status) io.status[frame] = 0;
        }
        wave_lds_sync();
    }
}

template <int CPL, int KW>
void launch_osd_variant(const OsdCodeDev& code, const OsdParams& p, const OsdIo& io, int max_frames, hipStream_t s)
{
    const size_t lds = sizeof(OsdLds<CPL, KW>) * kOsdWavesPerBlock;
    // One wave per frame up to what the GPU holds at once (the frame count may
    // live on the device; surplus waves find no frame and exit).
    const long resident = static_cast<long>(resident_blocks_per_cu(reinterpret_cast<const void*>(osd_kernel<CPL, KW>),
                                                                   64 * kOsdWavesPerBlock, lds)) *
                          device_cu_count() * kOsdWavesPerBlock;
    long waves = max_frames < resident ? max_frames : resident;
    if (waves < 1) waves = 1;
    const int blocks = static_cast<int>((waves + kOsdWavesPerBlock - 1) / kOsdWavesPerBlock);
    hipLaunchKernelGGL((osd_kernel<CPL, KW>), dim3(blocks), dim3(64 * kOsdWavesPerBlock), lds, s, code, p, io);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("osd_kernel launch: ") + hipGetErrorString(e));
}

}  // namespace

bool osd_shape_supported(int n, int k) { return n > 0 && k > 0 && k <= 128 && k < n && n <= 256; }

void launch_osd(const OsdCodeDev& code, const OsdParams& p, const OsdIo& io, int max_frames, hipStream_t s)
{
    if (!osd_shape_supported(code.n, code.k)) throw std::invalid_argument("OSD supports k <= 128, n <= 256");
    if (p.order < 0 || p.order > kMaxOrder) throw std::invalid_argument("OSD order must be 0..3");
    if (p.max_flip < 1 || p.max_tests < 1) throw std::invalid_argument("OSD max_flip and max_tests must be >= 1");
    if (code.kw != (code.k + 63) / 64) throw std::invalid_argument("OSD generator word count does not match k");
    if (max_frames <= 0) return;
    if (code.k <= 64) {
        if (code.n <= 128) return launch_osd_variant<2, 1>(code, p, io, max_frames, s);
        return launch_osd_variant<4, 1>(code, p, io, max_frames, s);
    }
    if (code.n <= 128) return launch_osd_variant<2, 2>(code, p, io, max_frames, s);
    return launch_osd_variant<4, 2>(code, p, io, max_frames, s);
}

}  // namespace sicc
