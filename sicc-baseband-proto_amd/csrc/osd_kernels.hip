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
// Mapping: one wavefront per frame, built for a short dependency chain (a BP
// failure waits on this kernel).  The |L| order and the flip pool order are
// bitonic sorts of (key, position) pairs across the wave.  The generator is
// held column-major: lane l owns columns l, l+64, ... as one or two 64-bit
// words over the k <= 128 rows.  The elimination never swaps rows: step t reads
// column t's word (one v_readlane), takes its lowest unused row as the pivot
// and clears column t elsewhere with one masked XOR per column; the ballot
// search and the column swap run only when column t is dependent.  The reduced
// form with the reference's pivot columns is unique, so logical row t is the
// physical row chosen at step t.  The parity part is transposed into LDS rows
// R[t] (by logical row), and each lane evaluates two TEPs at a time: the
// disagreement vector of a TEP is D0 ^ R[f1] ^ R[f2] ..., summed in ascending
// position order so every lane reproduces the sequential CPU sum exactly.  The
// winner is the lexicographic minimum of (metric, test index) over the wave.
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

// Branch-free: both halves of the exchange are formed and one is kept by mask.
template <int J>
__device__ __forceinline__ unsigned long long transpose_stage(unsigned long long a, int lane, unsigned long long m)
{
    const unsigned long long b = J == 32 ? __shfl_xor(a, 32) : swz_xor<J & 31>(a);
    const unsigned long long lo = (a & m) | ((b & m) << J);
    const unsigned long long hi = (a & ~m) | ((b & ~m) >> J);
    const unsigned long long upper = 0ull - static_cast<unsigned long long>((lane >> __builtin_ctz(J)) & 1);
    return (hi & upper) | (lo & ~upper);
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

// Value of x in lane lane ^ j (j a power of two below 64, a constant once the
// caller's loops are unrolled).
__device__ __forceinline__ unsigned long long xor_lane(unsigned long long x, int j)
{
    switch (j) {
    case 1: return swz_xor<1>(x);
    case 2: return swz_xor<2>(x);
    case 4: return swz_xor<4>(x);
    case 8: return swz_xor<8>(x);
    case 16: return swz_xor<16>(x);
    default: return __shfl_xor(x, 32);
    }
}

// Ascending bitonic sort of 64 S keys across a wavefront: element e = lane + 64 s
// is key[s] of lane `lane`; on return element e holds the e-th smallest key.
// Stages with partner distance j < 64 exchange lanes, the others exchange
// registers; every key is distinct where it is used, so the result is the
// unique sorted order (no stability question).
template <int S>
__device__ __forceinline__ void bitonic_sort(unsigned long long (&key)[S], int lane)
{
    constexpr int N = 64 * S;
#pragma unroll
    for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
        for (int j = size >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
                const int sj = j >> 6;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (s & sj) continue;
                    const int s2 = s | sj;
                    const bool asc = ((64 * s) & size) == 0;
                    const unsigned long long a = key[s], b = key[s2];
                    const unsigned long long lo = a < b ? a : b;
                    const unsigned long long hi = a < b ? b : a;
                    key[s] = asc ? lo : hi;
                    key[s2] = asc ? hi : lo;
                }
            } else {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const unsigned long long y = xor_lane(key[s], j);
                    const bool asc = ((lane + 64 * s) & size) == 0;
                    const bool lower = (lane & j) == 0;
                    const unsigned long long a = key[s];
                    const unsigned long long lo = a < y ? a : y;
                    const unsigned long long hi = a < y ? y : a;
                    key[s] = asc == lower ? lo : hi;
                }
            }
        }
    }
}

// One elimination step on column-major words: column t (words w, independent of
// the pivots so far) takes physical row p as its pivot; every column with a one
// in row p takes w without that bit (the row operation "every other row with a
// one in column t adds the pivot row", seen column by column).
template <int CPL, int KW>
__device__ __forceinline__ void elim_step(unsigned long long (&col)[CPL][KW], unsigned long long (&used)[KW],
                                          int (&lrow)[KW], const unsigned long long (&w)[KW], int p, int t, int lane)
{
    const int pq = KW == 1 ? 0 : p >> 6;
    const unsigned long long pbit = 1ull << (p & 63);
    unsigned long long clr[KW];
#pragma unroll
    for (int q = 0; q < KW; ++q) {
        clr[q] = w[q];
        if (q == pq) {
            used[q] |= pbit;
            clr[q] &= ~pbit;
            if (lane == (p & 63)) lrow[q] = t;
        }
    }
    // bit p of a column word as an all-ones / zero mask: the 32-bit half holding it
    // (p is the same in every lane), then a sign-extending one-bit field extract;
    // the masked XOR is then one three-input op per 32-bit half
    const bool phi = (p & 32) != 0;
    const int pb = p & 31;
#pragma unroll
    for (int s = 0; s < CPL; ++s) {
        unsigned long long pw = col[s][0];
#pragma unroll
        for (int q = 1; q < KW; ++q)
            if (q == pq) pw = col[s][q];
        const unsigned half = phi ? static_cast<unsigned>(pw >> 32) : static_cast<unsigned>(pw);
        const unsigned m = static_cast<unsigned>(__builtin_amdgcn_sbfe(static_cast<int>(half), pb, 1));
#pragma unroll
        for (int q = 0; q < KW; ++q) {
            // w ^ (c & m) per half: v_bitop3 truth table 0x78
            const unsigned lo = __builtin_amdgcn_bitop3_b32(static_cast<unsigned>(col[s][q]),
                                                            static_cast<unsigned>(clr[q]), m, 0x78);
            const unsigned hi = __builtin_amdgcn_bitop3_b32(static_cast<unsigned>(col[s][q] >> 32),
                                                            static_cast<unsigned>(clr[q] >> 32), m, 0x78);
            col[s][q] = (static_cast<unsigned long long>(hi) << 32) | lo;
        }
    }
}

constexpr int kMaxOrder = 3;
constexpr int kOsdWavesPerBlock = 4;

template <int CPL, int KW>
struct OsdLds {
    float llr[64 * CPL];
    double rel[64 * CPL + 2];                // [64 CPL] = +0.0, read for "no further bit"
    unsigned long long rows[64 * KW * CPL];  // R[i][s], i < k
    int rperm[64 * CPL];
    int pool[64 * KW];
    unsigned char code[64 * CPL];
};

// TPL: test patterns per lane per round -- 2 when the whole test set fits one
// round of the wave (order 1: 65 tests), 1 otherwise (fewer registers, more
// resident waves for the throughput-bound order-2 batches).
template <int CPL, int KW, int TPL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TPL == 1 ? 5 : 1))) void osd_kernel(
    OsdCodeDev code, OsdParams p, OsdIo io)
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
    constexpr int kZero = 64 * CPL;  // index of the +0.0 reliability
    if (lane == 0) sh.rel[kZero] = 0.0;

    for (int f = gw; f < total; f += nw) {
        const int frame = io.index ? io.index[f] : f;
        const float* L = io.llr + static_cast<size_t>(f) * n;
        // 1. order by |L| descending, ties by position (a stable sort): ascending
        //    sort of (0x7fffffff - bits of |L|, position) -- the bits of a
        //    non-negative float order as its value
        unsigned long long key[CPL];
#pragma unroll
        for (int s = 0; s < CPL; ++s) {
            const int c = lane + 64 * s;
            const float x = c < n ? L[c] : 0.0f;
            sh.llr[c] = x;
            key[s] = c < n ? (static_cast<unsigned long long>(0x7fffffffu - __float_as_uint(fabsf(x))) << 32) |
                                 static_cast<unsigned>(c)
                           : ~0ull;
        }
        bitonic_sort<CPL>(key, lane);
        // 2. column words of the permuted generator: lane c holds sorted column c
        unsigned long long col[CPL][KW];
        int sp[CPL];
#pragma unroll
        for (int s = 0; s < CPL; ++s) {
            const int c = lane + 64 * s;
            const int src = static_cast<int>(key[s] & 0xffffu);
            if (c < n) sh.rperm[c] = src;
#pragma unroll
            for (int q = 0; q < KW; ++q) col[s][q] = c < n ? code.gen_cols[static_cast<size_t>(src) * KW + q] : 0ull;
            sp[s] = c;
        }
        wave_lds_sync();
        // Elimination to [I | P] with the reference's column choice: step t takes
        // column t if it is independent of the pivot columns so far, else swaps in
        // the first later column that is.  Rows are never swapped: the pivot of
        // step t is the lowest physical row not yet used, recorded as the logical
        // row t of that physical row.  The reduced form with these pivot columns
        // is unique, so logical row t equals the reference's row t, and a column's
        // bits in the unused rows are nonzero exactly when the column is
        // independent -- the reference's "a one in rows >= t" test.  A step reads
        // column t's word (v_readlane), picks the pivot on the scalar unit and
        // clears column t elsewhere with one masked XOR per column; the ballot
        // search and the column swap run only for a dependent column.
        unsigned long long used[KW];
        int lrow[KW];  // logical index of physical row 64 q + lane
#pragma unroll
        for (int q = 0; q < KW; ++q) {
            used[q] = 0ull;
            lrow[q] = 0;
        }
        int status = 0;
        int t = 0;
        for (;;) {
            unsigned long long w[KW];
            for (; t < k; ++t) {
                const int ts = KW == 1 ? 0 : t >> 6;  // column slot of column t (t < k <= 64 KW)
#pragma unroll
                for (int q = 0; q < KW; ++q) w[q] = 0ull;
#pragma unroll
                for (int s = 0; s < (KW < CPL ? KW : CPL); ++s)
                    if (s == ts)
#pragma unroll
                        for (int q = 0; q < KW; ++q) w[q] = lane_u64(col[s][q], t & 63);
                int p = -1;
#pragma unroll
                for (int q = 0; q < KW; ++q) {
                    const unsigned long long m = w[q] & ~used[q];
                    if (p < 0 && m != 0ull) p = 64 * q + __ffsll(static_cast<long long>(m)) - 1;
                }
                if (p < 0) break;  // column t is dependent
                elim_step<CPL, KW>(col, used, lrow, w, p, t, lane);
            }
            if (t >= k) break;
            // swap in the first later column with a one in an unused row
            int pc = -1;
#pragma unroll
            for (int s = 0; s < CPL; ++s) {
                const int c = lane + 64 * s;
                bool ind = false;
#pragma unroll
                for (int q = 0; q < KW; ++q) ind = ind || (col[s][q] & ~used[q]) != 0ull;
                const unsigned long long b = __ballot(c > t && c < n && ind);
                if (pc < 0 && b != 0ull) pc = 64 * s + __ffsll(static_cast<long long>(b)) - 1;
            }
            if (pc < 0) {
                status = 1;
                break;
            }
            const int ts = KW == 1 ? 0 : t >> 6;
            unsigned long long wp[KW];
            int st = 0, spc = 0;
#pragma unroll
            for (int q = 0; q < KW; ++q) wp[q] = 0ull;
#pragma unroll
            for (int s = 0; s < CPL; ++s) {
                if (s == (pc >> 6)) {
#pragma unroll
                    for (int q = 0; q < KW; ++q) wp[q] = lane_u64(col[s][q], pc & 63);
                    spc = __builtin_amdgcn_readlane(sp[s], pc & 63);
                }
                if (s == ts) st = __builtin_amdgcn_readlane(sp[s], t & 63);
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
                    for (int q = 0; q < KW; ++q) col[s][q] = w[q];
                    sp[s] = st;
                }
            }
            int p = -1;
#pragma unroll
            for (int q = 0; q < KW; ++q) {
                const unsigned long long m = wp[q] & ~used[q];
                if (p < 0 && m != 0ull) p = 64 * q + __ffsll(static_cast<long long>(m)) - 1;
            }
            elim_step<CPL, KW>(col, used, lrow, wp, p, t, lane);
            ++t;
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
        // slots 0 .. KW-1 hold positions 0 .. 64 KW - 1), by logical row; and the
        // same bits by physical row, for products with the column words
        unsigned long long base[KW], basep[KW];
#pragma unroll
        for (int q = 0; q < KW; ++q) {
            const int lo = 64 * q;
            const unsigned long long m = k - lo >= 64 ? ~0ull : (k > lo ? ((1ull << (k - lo)) - 1ull) : 0ull);
            base[q] = __ballot(hs[q]) & m;
        }
#pragma unroll
        for (int q = 0; q < KW; ++q) {
            const int tl = lrow[q];
            unsigned long long bw = base[0];
#pragma unroll
            for (int u = 1; u < KW; ++u)
                if (u == (tl >> 6)) bw = base[u];
            basep[q] = __ballot(64 * q + lane < k && ((bw >> (tl & 63)) & 1ull) != 0ull);
        }
        unsigned long long d0[CPL];
#pragma unroll
        for (int s = 0; s < CPL; ++s) {
            const int c = lane + 64 * s;
            const bool par = c >= k && c < n;
            int pb = 0;
#pragma unroll
            for (int q = 0; q < KW; ++q) pb ^= __popcll(basep[q] & col[s][q]) & 1;
            d0[s] = __ballot(par && ((pb != 0) != hs[s]));
        }
        // parity rows by position: each 64 x 64 block (lane = position, bit =
        // physical row) transposed across the wave so that lane = physical row,
        // bit = position, stored at the row's logical index
#pragma unroll
        for (int s = 0; s < CPL; ++s) {
            const int c = lane + 64 * s;
#pragma unroll
            for (int q = 0; q < KW; ++q) {
                const unsigned long long tr = transpose64(c >= k && c < n ? col[s][q] : 0ull, lane);
                if (64 * q + lane < k) sh.rows[lrow[q] * CPL + s] = tr;
            }
        }
        // weak-order flip pool: stable ascending sort of rel[0..k), as an ascending
        // sort of (bits of |L|, position)
        {
            unsigned long long pk[KW];
#pragma unroll
            for (int q = 0; q < KW; ++q) {
                const int c = lane + 64 * q;
                pk[q] = ~0ull;
#pragma unroll
                for (int s = 0; s < CPL; ++s)
                    if (s == q && c < k)
                        pk[q] = (static_cast<unsigned long long>(__float_as_uint(fabsf(sh.llr[sh.rperm[sp[s]]]))) << 32) |
                                static_cast<unsigned>(c);
            }
            bitonic_sort<KW>(pk, lane);
#pragma unroll
            for (int q = 0; q < KW; ++q) {
                const int e = lane + 64 * q;
                if (e < pool) sh.pool[e] = static_cast<int>(pk[q] & 0xffffu);
            }
        }
        wave_lds_sync();
        // 4. evaluate TEPs, TPL per lane at a time (t, t + 64, ...): the metric of
        //    a TEP sums |L| over its flips (systematic positions, ascending) and
        //    then over the parity positions where the re-encoded word disagrees
        //    with the hard decisions, ascending; four set bits of each are read
        //    from LDS before their ordered adds (+0.0 past the last bit: the sum is
        //    >= +0.0, so adding +0.0 leaves it unchanged)
        double best_m = 0.0;
        long long best_t = -1;
        for (long long t0 = lane; t0 < tests_total; t0 += 64 * TPL) {
            double met[TPL];
            unsigned long long d[TPL][CPL];
#pragma unroll
            for (int u = 0; u < TPL; ++u) {
                const long long t = t0 + 64 * u;
                met[u] = 0.0;
#pragma unroll
                for (int s = 0; s < CPL; ++s) d[u][s] = t < tests_total ? d0[s] : 0ull;
                if (t >= tests_total) continue;
                int ix[kMaxOrder];
                const int o = unrank_tep(t, pool, p.order, ix);
                int pos[kMaxOrder];
                for (int i = 0; i < o; ++i) pos[i] = sh.pool[ix[i]];
                for (int i = 1; i < o; ++i)
                    for (int j = i; j > 0 && pos[j - 1] > pos[j]; --j) {
                        const int x = pos[j];
                        pos[j] = pos[j - 1];
                        pos[j - 1] = x;
                    }
                for (int i = 0; i < o; ++i) {
                    met[u] += sh.rel[pos[i]];
#pragma unroll
                    for (int s = 0; s < CPL; ++s) d[u][s] ^= sh.rows[pos[i] * CPL + s];
                }
            }
#pragma unroll
            for (int s = 0; s < CPL; ++s) {
                unsigned long long w[TPL];
                unsigned long long any = 0ull;
#pragma unroll
                for (int u = 0; u < TPL; ++u) {
                    w[u] = d[u][s];
                    any |= w[u];
                }
                while (any != 0ull) {
                    double x[TPL][4];
#pragma unroll
                    for (int u = 0; u < TPL; ++u) {
                        int ib[4];
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            ib[r] = w[u] ? 64 * s + __ffsll(static_cast<long long>(w[u])) - 1 : kZero;
                            w[u] &= w[u] - 1ull;
                        }
#pragma unroll
                        for (int r = 0; r < 4; ++r) x[u][r] = sh.rel[ib[r]];
                    }
                    any = 0ull;
#pragma unroll
                    for (int u = 0; u < TPL; ++u) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) met[u] += x[u][r];
                        any |= w[u];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < TPL; ++u) {
                const long long t = t0 + 64 * u;
                if (t < tests_total && (best_t < 0 || met[u] < best_m)) {
                    best_m = met[u];
                    best_t = t;
                }
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

template <int CPL, int KW, int TPL>
void launch_osd_tpl(const OsdCodeDev& code, const OsdParams& p, const OsdIo& io, int max_frames, hipStream_t s)
{
    const size_t lds = sizeof(OsdLds<CPL, KW>) * kOsdWavesPerBlock;
    // One wave per frame up to what the GPU holds at once (the frame count may
    // live on the device; surplus waves find no frame and exit).
    const long resident = static_cast<long>(resident_blocks_per_cu(
                              reinterpret_cast<const void*>(osd_kernel<CPL, KW, TPL>), 64 * kOsdWavesPerBlock, lds)) *
                          device_cu_count() * kOsdWavesPerBlock;
    long waves = max_frames < resident ? max_frames : resident;
    if (waves < 1) waves = 1;
    const int blocks = static_cast<int>((waves + kOsdWavesPerBlock - 1) / kOsdWavesPerBlock);
    hipLaunchKernelGGL((osd_kernel<CPL, KW, TPL>), dim3(blocks), dim3(64 * kOsdWavesPerBlock), lds, s, code, p, io);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("osd_kernel launch: ") + hipGetErrorString(e));
}

template <int CPL, int KW>
void launch_osd_variant(const OsdCodeDev& code, const OsdParams& p, const OsdIo& io, int max_frames, hipStream_t s)
{
    const int pool = code.k < p.max_flip ? code.k : p.max_flip;
    long long tests = 0;
    for (int o = 0; o <= p.order; ++o) {
        long long c = 1;
        for (int i = 0; i < o; ++i) c = c * (pool - i) / (i + 1);
        tests += pool >= o ? c : 0;
    }
    if (tests > p.max_tests) tests = p.max_tests;
    if (tests <= 128) return launch_osd_tpl<CPL, KW, 2>(code, p, io, max_frames, s);
    return launch_osd_tpl<CPL, KW, 1>(code, p, io, max_frames, s);
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
