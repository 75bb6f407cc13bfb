// ofdm_rx_kernels.hip -- batched OFDM receive for gfx950: Schmidl-Cox frame sync
// and a fused demodulator (CFO derotation, CP strip, FFT, LS pilot channel
// estimate, interpolation, equalization, max-log LLR).
//
// Semantics follow the reference CPU receiver (paths relative to the
// reference root, LINK = source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp):
//   find_frame LINK:5123-5212 -- global argmax of |P|^2 / (E1 E2) over the
//     coarse window, refinement by full-preamble normalized correlation
//     within +-radius, CFO = arg(sum conj(x) x[+L]) * fs / (2 pi L);
//   decode_frame LINK:5354-5521 -- phase origin at the frame start, FP64 FFT
//     per symbol (cast to float), y conj(p) / max(|p|^2, 1e-9), linear
//     interpolation, eq = y conj(h) / (|h|^2 + s2), s2_eq = max(s2/|h|^2, 1e-8),
//     LLR_b = (min_{b=1} |eq-c|^2 - min_{b=0} |eq-c|^2) / s2_eq (LINK:2552).
//
// Layout: frames are [frames][span] complex64 rows; each frame starts its
// search at sample 0 of its row.  LLRs are written per frame as
// [blocks * code_n] floats in (data symbol, subcarrier, bit) order; resource
// elements past blocks * code_n carry pad bits and produce no LLR.
//
// Mapping.  Sync: one 256-thread workgroup per frame; the search window sits
// in LDS and the CPU search's sliding recurrence runs as a block scan of
// per-candidate increments (FP64), then block argmax / refine (one wavefront
// per candidate, window energies from the scan) / CFO.  Demod: one 256-thread
// workgroup per (frame, pilot symbol) writes LS estimates (A floats2 per pilot,
// L2-resident), then one per (frame, data symbol) derotates + strips the CP
// while loading the FFT input, runs the Stockham FFT in FP64 through LDS
// (radix 4, 5 stages for 1024; a last radix-2 stage for 512 / 2048) and equalizes and turns the symbol into LLRs straight
// from LDS: the frequency grid never goes to HBM, each symbol window is read
// once and each LLR written once, and a batch is tens of thousands of
// independent workgroups.
#include "fft_lds.hpp"
#include "ofdm.hpp"

#include <cfloat>
#include <cmath>
#include <stdexcept>
#include <string>

namespace sicc {

namespace {

constexpr int kThreads = 256;
constexpr double kPiD = 3.141592653589793238462643383279502884;

__device__ __forceinline__ double block_sum(double v, double* red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
    return t;
}

// ---------------------------------------------------------------- sync ----
__global__ __launch_bounds__(kThreads) void ofdm_sync_kernel(OfdmSyncArgs a)
{
    extern __shared__ float2 win[];  // [wlen] samples of this frame
    __shared__ double red[3 * (kThreads / 64)];
    __shared__ double bm[kThreads];
    __shared__ int bi[kThreads];
    __shared__ int s_best;

    const int f = blockIdx.x;
    const float2* x = a.rx + static_cast<size_t>(f) * a.span;
    const int L = a.half_len;
    const int pre = 2 * L;
    const int last = a.span - pre;
    const int limit = (a.max_search > 0 && a.max_search < last) ? a.max_search : last;
    const int hi_all = min(last, limit + a.radius);
    const int wlen = hi_all + pre;  // samples [0, wlen) cover coarse and refine windows
    // Windows that fit are staged in LDS; longer rows are read through L2.
    const float2* w = x;
    if (a.window_in_lds) {
        for (int i = threadIdx.x; i < wlen; i += blockDim.x) win[i] = x[i];
        w = win;
    }
    __syncthreads();

    // coarse: M(d) = |sum conj(x[d+i]) x[d+i+L]|^2 / max(E1 E2, 1e-18)
    double best_m = 0.0;
    int best_d = 0x7fffffff;
    if (a.prefix) {
        // Windows staged in LDS: the sliding recurrence of the CPU search
        // (P(d+1) = P(d) + conj(x[d+L]) x[d+2L] - conj(x[d]) x[d+L], E1, E2 by
        // float differences of float norms, find_frame LINK:5123-5212), evaluated
        // in parallel: the sums at d = 0 by a block reduction, the per-candidate
        // increments by a block scan.  The window energies E1 + E2 of every
        // candidate up to the refinement range stay in LDS for the refinement.
        const int nd = hi_all + 1;  // candidates d = 0 .. hi_all (the coarse search uses d <= limit)
        double* Ed = reinterpret_cast<double*>(win + wlen);  // [nd]
        const int t = threadIdx.x;
        const int ln = t & 63, wv = t >> 6;
        __shared__ double base[4];
        __shared__ double wsum[4 * (kThreads / 64)];
        {
            double p0r = 0.0, p0i = 0.0, e10 = 0.0, e20 = 0.0;
            for (int i = t; i < L; i += kThreads) {
                const float2 u = w[i], v = w[i + L];
                const double ur = u.x, ui = u.y, vr = v.x, vi = v.y;
                p0r += ur * vr + ui * vi;
                p0i += ur * vi - ui * vr;
                e10 += static_cast<double>(u.x * u.x + u.y * u.y);
                e20 += static_cast<double>(v.x * v.x + v.y * v.y);
            }
            for (int o = 32; o > 0; o >>= 1) {
                p0r += __shfl_xor(p0r, o);
                p0i += __shfl_xor(p0i, o);
                e10 += __shfl_xor(e10, o);
                e20 += __shfl_xor(e20, o);
            }
            if (ln == 0) {
                wsum[4 * wv] = p0r;
                wsum[4 * wv + 1] = p0i;
                wsum[4 * wv + 2] = e10;
                wsum[4 * wv + 3] = e20;
            }
            __syncthreads();
            if (t < 4) {
                double b = 0.0;
                for (int u = 0; u < kThreads / 64; ++u) b += wsum[4 * u + t];
                base[t] = b;
            }
            __syncthreads();
        }
        // thread t owns candidates [lc, hc); the increment of d moves d -> d + 1
        const int per_c = (nd + kThreads - 1) / kThreads;
        const int lc = min(nd, t * per_c), hc = min(nd, lc + per_c);
        auto incr = [&](int d, double& dr, double& di, double& d1, double& d2) {
            const float2 x0 = w[d], x1 = w[d + L], x2 = w[d + pre];
            const double x0r = x0.x, x0i = x0.y, x1r = x1.x, x1i = x1.y, x2r = x2.x, x2i = x2.y;
            dr = (x1r * x2r + x1i * x2i) - (x0r * x1r + x0i * x1i);
            di = (x1r * x2i - x1i * x2r) - (x0r * x1i - x0i * x1r);
            const float n0 = x0.x * x0.x + x0.y * x0.y, n1 = x1.x * x1.x + x1.y * x1.y;
            const float n2 = x2.x * x2.x + x2.y * x2.y;
            d1 = static_cast<double>(n1 - n0);
            d2 = static_cast<double>(n2 - n1);
        };
        double tr = 0.0, ti = 0.0, t1 = 0.0, t2 = 0.0;  // this thread's increments (all but the last candidate's)
        for (int d = lc; d < hc && d < nd - 1; ++d) {
            double dr, di, d1, d2;
            incr(d, dr, di, d1, d2);
            tr += dr;
            ti += di;
            t1 += d1;
            t2 += d2;
        }
        // exclusive scan of the thread totals: shuffles within each wavefront, then the
        // totals of the earlier wavefronts
        double ir = tr, ii = ti, i1 = t1, i2 = t2;
        for (int o = 1; o < 64; o <<= 1) {
            const double ur = __shfl_up(ir, o), ui = __shfl_up(ii, o), u1 = __shfl_up(i1, o), u2 = __shfl_up(i2, o);
            if (ln >= o) {
                ir += ur;
                ii += ui;
                i1 += u1;
                i2 += u2;
            }
        }
        __syncthreads();  // wsum is reused
        if (ln == 63) {
            wsum[4 * wv] = ir;
            wsum[4 * wv + 1] = ii;
            wsum[4 * wv + 2] = i1;
            wsum[4 * wv + 3] = i2;
        }
        __syncthreads();
        double pr = base[0], pi = base[1], e1 = base[2], e2 = base[3];
        for (int u = 0; u < wv; ++u) {
            pr += wsum[4 * u];
            pi += wsum[4 * u + 1];
            e1 += wsum[4 * u + 2];
            e2 += wsum[4 * u + 3];
        }
        pr += ir - tr;
        pi += ii - ti;
        e1 += i1 - t1;
        e2 += i2 - t2;
        for (int d = lc; d < hc; ++d) {
            if (d <= limit) {
                const double den = e1 * e2 > 1e-18 ? e1 * e2 : 1e-18;
                const double m = (pr * pr + pi * pi) / den;
                if (m > best_m) {
                    best_m = m;
                    best_d = d;
                }
            }
            Ed[d] = e1 + e2;
            if (d < nd - 1) {
                double dr, di, d1, d2;
                incr(d, dr, di, d1, d2);
                pr += dr;
                pi += di;
                e1 += d1;
                e2 += d2;
            }
        }
        __syncthreads();  // bm is reused by the argmax below; Ed complete
    } else {
        for (int d = threadIdx.x; d <= limit; d += blockDim.x) {
            double pr = 0.0, pi = 0.0, e1 = 0.0, e2 = 0.0;
            for (int i = 0; i < L; ++i) {
                const float2 u = w[d + i];
                const float2 v = w[d + i + L];
                const double ur = u.x, ui = u.y, vr = v.x, vi = v.y;
                pr += ur * vr + ui * vi;
                pi += ur * vi - ui * vr;
                e1 += static_cast<double>(u.x * u.x + u.y * u.y);
                e2 += static_cast<double>(v.x * v.x + v.y * v.y);
            }
            const double den = e1 * e2 > 1e-18 ? e1 * e2 : 1e-18;
            const double m = (pr * pr + pi * pi) / den;
            if (m > best_m) {
                best_m = m;
                best_d = d;
            }
        }
    }
    // argmax, first index on ties (strict > in index order)
    bm[threadIdx.x] = best_m;
    bi[threadIdx.x] = best_d;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double om = bm[threadIdx.x + s];
            const int od = bi[threadIdx.x + s];
            if (om > bm[threadIdx.x] || (om == bm[threadIdx.x] && od < bi[threadIdx.x])) {
                bm[threadIdx.x] = om;
                bi[threadIdx.x] = od;
            }
        }
        __syncthreads();
    }
    const double coarse = bm[0];
    const int best = bi[0] == 0x7fffffff ? 0 : bi[0];
    __syncthreads();
    OfdmSyncResult res;
    res.start = -1;
    res.metric = 0.0f;
    res.cfo_hz = 0.0f;
    res.cfo_hz_d = 0.0;
    res.coarse_metric = static_cast<float>(coarse);
    if (coarse < a.threshold) {
        if (threadIdx.x == 0) a.out[f] = res;
        return;
    }
    // refine: full-preamble normalized correlation, one wave per candidate
    const int lo = best > a.radius ? best - a.radius : 0;
    const int hi = min(last, best + a.radius);
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int nwv = blockDim.x >> 6;
    double wbest = 0.0;
    int wbest_d = 0x7fffffff;
    // window energies of the candidates when the coarse search left them in LDS
    const double* Ep = a.prefix ? reinterpret_cast<const double*>(win + wlen) : nullptr;
    // Preambles of up to 64 * kRefineRegs samples: each lane keeps its preamble
    // samples in registers for all candidates and issues a candidate's window
    // reads together; longer ones loop through L2.
    constexpr int kRefineRegs = 16;
    const bool regs = pre <= 64 * kRefineRegs;
    double2 qv[kRefineRegs];
#pragma unroll
    for (int j = 0; j < kRefineRegs; ++j) {
        const int i = lane + 64 * j;
        qv[j] = (regs && i < pre) ? a.preamble_d[i] : make_double2(0.0, 0.0);
    }
    for (int d = lo + wv; d <= hi; d += nwv) {
        double cr = 0.0, ci = 0.0, eng = 0.0;
        // y conj(q) accumulated with fused multiply-adds (the refined metric needs
        // a few parts in 1e16, not the CPU's separately rounded products)
        if (regs) {
            double cr1 = 0.0, ci1 = 0.0;
#pragma unroll
            for (int j = 0; j < kRefineRegs; ++j) {
                const int i = lane + 64 * j;
                if (i < pre) {
                    const float2 y = w[d + i];
                    const double yr = y.x, yi = y.y, qr = qv[j].x, qi = qv[j].y;
                    if (j & 1) {
                        cr1 = fma(yi, qi, fma(yr, qr, cr1));
                        ci1 = fma(-yr, qi, fma(yi, qr, ci1));
                    } else {
                        cr = fma(yi, qi, fma(yr, qr, cr));
                        ci = fma(-yr, qi, fma(yi, qr, ci));
                    }
                    if (!Ep) eng += static_cast<double>(y.x * y.x + y.y * y.y);
                }
            }
            cr += cr1;
            ci += ci1;
        } else {
            for (int i = lane; i < pre; i += 64) {
                const float2 y = w[d + i];
                const double2 q = a.preamble_d[i];
                const double yr = y.x, yi = y.y;
                cr = fma(yi, q.y, fma(yr, q.x, cr));
                ci = fma(-yr, q.y, fma(yi, q.x, ci));
                if (!Ep) eng += static_cast<double>(y.x * y.x + y.y * y.y);
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            cr += __shfl_xor(cr, o);
            ci += __shfl_xor(ci, o);
            if (!Ep) eng += __shfl_xor(eng, o);
        }
        if (Ep) eng = Ep[d];
        const double den = eng * a.pre_energy > 1e-18 ? eng * a.pre_energy : 1e-18;
        const double m = (cr * cr + ci * ci) / den;
        if (m > wbest) {
            wbest = m;
            wbest_d = d;
        }
    }
    if (lane == 0) {
        bm[wv] = wbest;
        bi[wv] = wbest_d;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        int d = best;
        for (int wi = 0; wi < nwv; ++wi)
            if (bm[wi] > m || (bm[wi] == m && bm[wi] > 0.0 && bi[wi] < d)) {
                m = bm[wi];
                d = bi[wi];
            }
        bm[0] = m;
        s_best = d;
    }
    __syncthreads();
    const double peak = bm[0];
    const int start = s_best;
    res.metric = static_cast<float>(peak);
    if (peak < a.threshold) {
        if (threadIdx.x == 0) a.out[f] = res;
        return;
    }
    double ar = 0.0, ai = 0.0;
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
        const float2 u = w[start + i];
        const float2 v = w[start + i + L];
        const double ur = u.x, ui = u.y, vr = v.x, vi = v.y;
        ar += ur * vr + ui * vi;
        ai += ur * vi - ui * vr;
    }
    ar = block_sum(ar, red);
    ai = block_sum(ai, red + (kThreads / 64));
    if (threadIdx.x == 0) {
        res.start = start;
        res.cfo_hz = static_cast<float>(atan2(ai, ar) * a.fs / (2.0 * kPiD * static_cast<double>(L)));
        res.cfo_hz_d = atan2(ai, ar) * a.fs / (2.0 * kPiD * static_cast<double>(L));
        a.out[f] = res;
    }
}

// --------------------------------------------------------------- demod ----
__device__ __forceinline__ bool frame_usable(const OfdmDemodArgs& a, const OfdmSyncResult& sr)
{
    return sr.start >= 0 && static_cast<long long>(sr.start) + a.frame_samples <= a.span;
}

// Derotate (phase origin at the frame start), strip the CP and FFT symbol `sym`
// of frame x into buf (natural bin order, FP64).
__device__ void symbol_fft(const OfdmDemodArgs& a, const float2* x, double cfo, int sym, Cd* buf)
{
    const int N = a.nfft;
    const int base = a.preamble_len + sym * (N + a.cp) + a.cp;
    // Derotation factor e^{-2 pi i cfo n / fs}, rounded to float as the reference
    // applies it.  Each thread evaluates it in double at its first sample and steps
    // by the thread stride with double complex products (N / blockDim - 1 of them):
    // a relative error near 1e-15, far below the float rounding of the factor.
    const double hc = -2.0 * cfo / a.fs;  // phase per sample in units of pi
    double s0, c0, sd, cd;
    sincospi(hc * static_cast<double>(base + static_cast<int>(threadIdx.x)), &s0, &c0);
    sincospi(hc * static_cast<double>(blockDim.x), &sd, &cd);
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
        const int n = base + k;
        const float c = static_cast<float>(c0);
        const float s = static_cast<float>(s0);
        const float2 v = x[n];
        const float re = v.x * c - v.y * s;
        const float im = v.x * s + v.y * c;
        buf[k] = {static_cast<double>(re), static_cast<double>(im)};
        const double cn = fma(c0, cd, -(s0 * sd));
        s0 = fma(s0, cd, c0 * sd);
        c0 = cn;
    }
    __syncthreads();
    fft_stockham<false>(buf, a.twiddle, N);
}

// One workgroup per (frame, pilot symbol): LS estimate y conj(p) / max(|p|^2, 1e-9)
// of every active subcarrier into h_est[frame][pilot][active_sc].
__global__ __launch_bounds__(kThreads) void ofdm_pilot_kernel(OfdmDemodArgs a)
{
    extern __shared__ double2 sh[];
    Cd* buf = reinterpret_cast<Cd*>(sh);
    const int f = blockIdx.x / a.n_pilot;
    const int pi = blockIdx.x - f * a.n_pilot;
    const OfdmSyncResult sr = a.sync[f];
    if (!frame_usable(a, sr)) return;
    const float2* x = a.rx + static_cast<size_t>(f) * a.span + sr.start;
    symbol_fft(a, x, sr.cfo_hz_d, a.pilot_symbols[pi], buf);
    const int A = a.active_sc;
    float2* h = a.h_est + (static_cast<size_t>(f) * a.n_pilot + pi) * A;
    for (int sc = threadIdx.x; sc < A; sc += blockDim.x) {
        const Cd yv = buf[a.used_natural[sc]];
        const float yr = static_cast<float>(yv.x), yi = static_cast<float>(yv.y);
        const float2 p = a.pilot_grid[pi * A + sc];
        float den = p.x * p.x + p.y * p.y;
        if (den < 1e-9f) den = 1e-9f;
        const float tr = yr * p.x - yi * (-p.y);
        const float ti = yr * (-p.y) + yi * p.x;
        h[sc] = make_float2(tr / den, ti / den);
    }
}

// Sum of four doubles over the workgroup; the total is valid in thread 0.
__device__ __forceinline__ void block_sum4(double (&v)[4], double* red)
{
#pragma unroll
    for (int j = 0; j < 4; ++j)
        for (int o = 32; o > 0; o >>= 1) v[j] += __shfl_xor(v[j], o);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int j = 0; j < 4; ++j) red[4 * w + j] = v[j];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 1; i < static_cast<int>(blockDim.x >> 6); ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] += red[4 * i + j];
}

// One workgroup per (frame, data symbol): FFT, interpolated channel, equalize,
// max-log LLRs of the symbol's resource elements.
template <int B>
__global__ __launch_bounds__(kThreads) void ofdm_data_kernel(OfdmDemodArgs a)
{
    extern __shared__ double2 sh[];
    __shared__ double red4[4 * (kThreads / 64)];
    Cd* buf = reinterpret_cast<Cd*>(sh);
    const int f = blockIdx.x / a.n_data;
    const int di = blockIdx.x - f * a.n_data;
    const int A = a.active_sc;
    const OfdmSyncResult sr = a.sync[f];
    float* out = a.llr + static_cast<size_t>(f) * a.llr_per_frame;
    double* st = a.stats ? a.stats + (static_cast<size_t>(f) * a.n_data + di) * 4 : nullptr;
    if (!frame_usable(a, sr)) {
        // no frame detected, or the frame runs past the row: zero LLRs (erasures)
        const int lo = di * A * B;
        const int hi = min(lo + A * B, a.llr_per_frame);
        for (int i = lo + static_cast<int>(threadIdx.x); i < hi; i += blockDim.x) out[i] = 0.0f;
        if (st && threadIdx.x < 4) st[threadIdx.x] = 0.0;
        return;
    }
    if (!st && di * A * B >= a.llr_per_frame) return;  // a symbol of pad resource elements only
    const float2* x = a.rx + static_cast<size_t>(f) * a.span + sr.start;
    symbol_fft(a, x, sr.cfo_hz_d, a.data_symbols[di], buf);
    constexpr int AX = B == 1 ? 1 : B / 2;  // leading (in-phase) bits of a symbol index
    constexpr int QB = B - AX;               // trailing (quadrature) bits
    constexpr int LI = 1 << AX, LQ = 1 << QB;
    const float nv = a.noise_var;
    const int p0 = a.interp_p0[di], p1 = a.interp_p1[di];
    const float w0 = a.interp_w0[di], w1 = a.interp_w1[di];
    const float2* h0p = a.use_pilots ? a.h_est + (static_cast<size_t>(f) * a.n_pilot + p0) * A : nullptr;
    const float2* h1p = a.use_pilots ? a.h_est + (static_cast<size_t>(f) * a.n_pilot + p1) * A : nullptr;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};  // |h|^2, |y - h hard|^2, |eq0 - hard|^2, |hard|^2
    const bool vec_ok = (reinterpret_cast<uintptr_t>(out) & (B == 4 ? 15u : 7u)) == 0;
    for (int sc = threadIdx.x; sc < A; sc += blockDim.x) {
        const int obit = (di * A + sc) * B;
        if (!st && obit >= a.llr_per_frame) continue;  // pad resource elements
        float hr = 1.0f, hi = 0.0f;
        if (a.use_pilots) {
            const float2 h0 = h0p[sc], h1 = h1p[sc];
            hr = h0.x * w0 + h1.x * w1;
            hi = h0.y * w0 + h1.y * w1;
        }
        const float hn = hr * hr + hi * hi;
        const float h2 = hn < 1e-9f ? 1e-9f : hn;
        const Cd yv = buf[a.used_natural[sc]];
        const float yr = static_cast<float>(yv.x), yi = static_cast<float>(yv.y);
        const float tr = yr * hr - yi * (-hi);
        const float ti = yr * (-hi) + yi * hr;
        if (st) {
            // decode_frame's SNR / EVM terms (uhd_ldpc_ofdm_link.cpp:5431-5439, 5457-5459):
            // the nearest point to eq0 is the first of least squared distance
            const float qr = tr / h2, qi = ti / h2;
            float bd = INFINITY;
            float cr = 0.0f, ci = 0.0f;
            for (int i = 0; i < (1 << B); ++i) {
                const float2 c = a.constellation[i];
                const float dx = qr - c.x, dy = qi - c.y;
                const float d2 = dx * dx + dy * dy;
                if (d2 < bd) {
                    bd = d2;
                    cr = c.x;
                    ci = c.y;
                }
            }
            const float rr = yr - (hr * cr - hi * ci);
            const float ri = yi - (hr * ci + hi * cr);
            const float er0 = qr - cr, ei0 = qi - ci;
            acc[0] += static_cast<double>(hn);
            acc[1] += static_cast<double>(rr * rr + ri * ri);
            acc[2] += static_cast<double>(er0 * er0 + ei0 * ei0);
            acc[3] += static_cast<double>(cr * cr + ci * ci);
            if (obit >= a.llr_per_frame) continue;  // pad resource elements carry no LLR
        }
        // eq = y conj(h) / (|h|^2 + s2) through the hardware reciprocal (1 ulp) and
        // 1 / s2_eq = 1 / max(s2 / |h|^2, 1e-8) = min(|h|^2 / s2, 1e8) through the
        // launch's 1 / s2: no division per resource element.  The equalized point
        // is within 2 float ulps and the scale within 2 double ulps of the
        // divided values, so an LLR differs from the divided form by a few float
        // ulps of its own magnitude (the sign only where d1 == d0 to that precision).
        const float rden = __builtin_amdgcn_rcpf(h2 + nv);
        const float er = tr * rden, ei = ti * rden;
        const double inv_s2 = fmin(static_cast<double>(h2) * a.inv_noise_var_d, 1e8);
        // The Gray QAM table is a product of an in-phase and a quadrature level set:
        // point (gi << QB) | gq sits at (X[gi], Y[gq]), so its distance is
        // fl(DX[gi] + DY[gq]).  fl(a + b) is monotone in each operand, hence the
        // minimum over the points whose bit b is 0 (or 1) equals fl(min DX + min DY)
        // over the matching level subsets: the same floats as the 2^B-point scan.
        float dxs[LI], dys[LQ];
#pragma unroll
        for (int g = 0; g < LI; ++g) {
            const float dx = er - a.constellation[g << QB].x;
            dxs[g] = dx * dx;
        }
#pragma unroll
        for (int g = 0; g < LQ; ++g) {
            const float dy = ei - a.constellation[g].y;
            dys[g] = dy * dy;
        }
        float mx = INFINITY, my = INFINITY;
#pragma unroll
        for (int g = 0; g < LI; ++g) mx = dxs[g] < mx ? dxs[g] : mx;
#pragma unroll
        for (int g = 0; g < LQ; ++g) my = dys[g] < my ? dys[g] : my;
        float lv[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            float d0, d1;
            if (b < AX) {
                float m0 = INFINITY, m1 = INFINITY;
#pragma unroll
                for (int g = 0; g < LI; ++g) {
                    if (((g >> (AX - 1 - b)) & 1) == 0)
                        m0 = dxs[g] < m0 ? dxs[g] : m0;
                    else
                        m1 = dxs[g] < m1 ? dxs[g] : m1;
                }
                d0 = m0 + my;
                d1 = m1 + my;
            } else {
                float m0 = INFINITY, m1 = INFINITY;
#pragma unroll
                for (int g = 0; g < LQ; ++g) {
                    if (((g >> (B - 1 - b)) & 1) == 0)
                        m0 = dys[g] < m0 ? dys[g] : m0;
                    else
                        m1 = dys[g] < m1 ? dys[g] : m1;
                }
                d0 = mx + m0;
                d1 = mx + m1;
            }
            lv[b] = static_cast<float>(static_cast<double>(d1 - d0) * inv_s2);
        }
        // a resource element's B LLRs leave as one vector store (B floats at a
        // multiple of B: 16-byte aligned for B = 4, 8-byte for B = 2 and 6, when
        // the frame's LLR row starts so aligned)
        if (obit + B <= a.llr_per_frame && vec_ok) {
            if (B == 4) {
                *reinterpret_cast<float4*>(out + obit) = make_float4(lv[0], lv[1 % B], lv[2 % B], lv[3 % B]);
            } else if (B == 2 || B == 6) {
#pragma unroll
                for (int b = 0; b < B; b += 2)
                    *reinterpret_cast<float2*>(out + obit + b) = make_float2(lv[b], lv[(b + 1) % B]);
            } else {
                out[obit] = lv[0];
            }
        } else {
#pragma unroll
            for (int b = 0; b < B; ++b)
                if (obit + b < a.llr_per_frame) out[obit + b] = lv[b];
        }
    }
    if (st) {
        block_sum4(acc, red4);
        if (threadIdx.x == 0)
#pragma unroll
            for (int j = 0; j < 4; ++j) st[j] = acc[j];
    }
}

}  // namespace

void launch_ofdm_sync(const OfdmSyncArgs& a, int frames, hipStream_t s)
{
    if (frames <= 0) return;
    if (!a.preamble_d) throw std::invalid_argument("sync: missing double-precision preamble");
    const int L = a.half_len;
    const int last = a.span - 2 * L;
    if (last < 0) throw std::invalid_argument("frame span shorter than the preamble");
    const int limit = (a.max_search > 0 && a.max_search < last) ? a.max_search : last;
    const int wlen = std::min(last, limit + a.radius) + 2 * L;
    OfdmSyncArgs b = a;
    size_t lds = sizeof(float2) * static_cast<size_t>(wlen);
    // the sliding search keeps the window energies of candidates 0 .. wlen - 2L
    const size_t prefix_lds = lds + sizeof(double) * static_cast<size_t>(wlen - 2 * L + 1);
    b.window_in_lds = lds <= 56 * 1024 ? 1 : 0;
    b.prefix = prefix_lds <= 56 * 1024 ? 1 : 0;
    if (b.prefix)
        lds = prefix_lds;
    else if (!b.window_in_lds)
        lds = 0;
    hipLaunchKernelGGL(ofdm_sync_kernel, dim3(frames), dim3(kThreads), lds, s, b);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("ofdm_sync_kernel: ") + hipGetErrorString(e));
}

void launch_ofdm_demod(const OfdmDemodArgs& a, int frames, hipStream_t s)
{
    if (frames <= 0) return;
    if (!lds_fft_size_ok(a.nfft)) throw std::invalid_argument("demod FFT needs nfft a power of two in [16, 2048]");
    if (a.n_pilot < 1 && a.use_pilots) throw std::invalid_argument("pilot channel estimation needs a pilot symbol");
    if (a.use_pilots && !a.h_est) throw std::invalid_argument("demod: missing channel-estimate scratch");
    if (a.n_data < 1) throw std::invalid_argument("demod: no data symbol");
    const size_t lds = sizeof(double2) * a.nfft;
    if (a.use_pilots)
        hipLaunchKernelGGL(ofdm_pilot_kernel, dim3(frames * a.n_pilot), dim3(kThreads), lds, s, a);
    const dim3 grid(frames * a.n_data);
    switch (a.bits_per_symbol) {
        case 1: hipLaunchKernelGGL(ofdm_data_kernel<1>, grid, dim3(kThreads), lds, s, a); break;
        case 2: hipLaunchKernelGGL(ofdm_data_kernel<2>, grid, dim3(kThreads), lds, s, a); break;
        case 4: hipLaunchKernelGGL(ofdm_data_kernel<4>, grid, dim3(kThreads), lds, s, a); break;
        case 6: hipLaunchKernelGGL(ofdm_data_kernel<6>, grid, dim3(kThreads), lds, s, a); break;
        default: throw std::invalid_argument("bits_per_symbol must be 1, 2, 4 or 6");
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("ofdm demod kernels: ") + hipGetErrorString(e));
}

}  // namespace sicc
