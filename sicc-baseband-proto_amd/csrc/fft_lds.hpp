// fft_lds.hpp -- mixed-radix (8 / 4 / 2) Stockham FFT through LDS for one
// workgroup (device code shared by the receive kernels, FP64, radix 4 / 2, and
// the transmitter, FP32, radix 8 / 4 / 2).
#pragma once

#include <hip/hip_runtime.h>

namespace sicc {

struct Cd {
    double x, y;
};
struct Cf {
    float x, y;
};

// Complex products with one fused multiply-add per component (2 multiplies and
// 2 FMAs instead of 4 multiplies and 2 adds; one rounding fewer per component).
// The spectra are compared with the reference's within a relative tolerance,
// not bit for bit.
static __device__ __forceinline__ Cd cmul(Cd a, Cd b) { return {fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x)}; }
static __device__ __forceinline__ Cf cmul(Cf a, Cf b)
{
    return {fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x)};
}

// Radix-4 DFT of (a0, a1, a2, a3) in place: the forward 4-point transform
// (W4 = -i).
template <class C>
static __device__ __forceinline__ void dft4(C& a0, C& a1, C& a2, C& a3)
{
    const C b0 = {a0.x + a2.x, a0.y + a2.y};
    const C b1 = {a0.x - a2.x, a0.y - a2.y};
    const C b2 = {a1.x + a3.x, a1.y + a3.y};
    const C b3 = {a1.y - a3.y, -(a1.x - a3.x)};  // (a1 - a3) * (-i)
    a0 = {b0.x + b2.x, b0.y + b2.y};
    a2 = {b0.x - b2.x, b0.y - b2.y};
    a1 = {b1.x + b3.x, b1.y + b3.y};
    a3 = {b1.x - b3.x, b1.y - b3.y};
}

// One radix-8 Stockham stage over span ns (N / 8 <= blockDim butterflies, at
// most one per thread): the 8-point DFT as two 4-point DFTs of the even and odd
// inputs, the odd one turned by W8^m (W8 = (1 - i) / sqrt 2), then one
// butterfly layer.
template <class C, class TW>
static __device__ __forceinline__ void fft_stage8(C* buf, const TW* tw, int N, int ns)
{
    using R = decltype(C{}.x);
    const R h = static_cast<R>(0.70710678118654752440);
    const int q = N >> 3;
    const int j = threadIdx.x;
    C y[8];
    int o = -1;
    if (j < q) {
        const int k = j & (ns - 1);
        C a[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) a[r] = buf[j + r * q];
        if (ns > 1) {
            const int step = N / (8 * ns);
#pragma unroll
            for (int r = 1; r < 8; ++r) {
                const TW t = tw[(r * k * step) & (N - 1)];
                a[r] = cmul(a[r], {t.x, t.y});
            }
        }
        dft4(a[0], a[2], a[4], a[6]);
        dft4(a[1], a[3], a[5], a[7]);
        // a[1 + 2m] is the odd half's bin m
        const C o0 = a[1];
        const C o1 = {(a[3].x + a[3].y) * h, (a[3].y - a[3].x) * h};    // * (1 - i) / sqrt 2
        const C o2 = {a[5].y, -a[5].x};                                  // * (-i)
        const C o3 = {(a[7].y - a[7].x) * h, -(a[7].x + a[7].y) * h};   // * (-1 - i) / sqrt 2
        const C e[4] = {a[0], a[2], a[4], a[6]};
        const C od[4] = {o0, o1, o2, o3};
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            y[m] = {e[m].x + od[m].x, e[m].y + od[m].y};
            y[m + 4] = {e[m].x - od[m].x, e[m].y - od[m].y};
        }
        o = (j / ns) * 8 * ns + k;
    }
    __syncthreads();
    if (o >= 0)
#pragma unroll
        for (int r = 0; r < 8; ++r) buf[o + r * ns] = y[r];
    __syncthreads();
}

// Forward FFT of any power of two N in [16, 2048], Stockham autosort in LDS,
// in the precision of C (Cd with double2 twiddles, Cf with float2 twiddles),
// with 256 threads.  Radix-4 stages while the span ns has four more to go
// (after radix-8 stages when R8), then one radix-2 stage when a factor 2 is
// left (N = 2 * 4^k: 32, 128, 512, 2048).
// Stage of radix R over span ns: butterfly j < N / R takes buf[j + r N / R]
// times w^(r k), k = j mod ns, w = exp(-2 pi i / (R ns)), and writes its R-point
// DFT to (j / ns) R ns + k + r ns.  tw[m] = exp(-2 pi i m / N).
// buf holds the input in natural order and receives the output in natural order.
// R8: radix-8 stages first (then radix 4 / 2 for the remaining factor).  For
// N = 1024 with 256 threads a radix-8 stage keeps half the threads idle; in
// FP64 (the receiver) that made the receive step 5 % slower than radix 4
// (0.710 against 0.677 ms per 4 096 frames), in FP32 (the transmitter) the
// fewer LDS passes won (0.662 against 0.695 ms).
template <bool R8, class C, class TW>
static __device__ void fft_stockham(C* buf, const TW* tw, int N)
{
    const int q = N >> 2;
    int ns = 1;
    if (R8)
        for (; 8 * ns <= N; ns <<= 3) fft_stage8(buf, tw, N, ns);
    for (; 4 * ns <= N; ns <<= 2) {
        // up to two butterflies per thread (N <= 2048 with 256 threads); fixed
        // indices keep the staged outputs in registers
        C v[2][4];
        int outi[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int j = threadIdx.x + c * blockDim.x;
            outi[c] = -1;
            if (j < q) {
                const int k = j & (ns - 1);
                C a0 = buf[j], a1 = buf[j + q], a2 = buf[j + 2 * q], a3 = buf[j + 3 * q];
                if (ns > 1) {
                    // twiddle w^(r k), w = exp(-2 pi i / (4 ns)); table entry m = exp(-2 pi i m / N)
                    const int step = N / (4 * ns);
                    const TW t1 = tw[(k * step) & (N - 1)];
                    const TW t2 = tw[(2 * k * step) & (N - 1)];
                    const TW t3 = tw[(3 * k * step) & (N - 1)];
                    a1 = cmul(a1, {t1.x, t1.y});
                    a2 = cmul(a2, {t2.x, t2.y});
                    a3 = cmul(a3, {t3.x, t3.y});
                }
                const C b0 = {a0.x + a2.x, a0.y + a2.y};
                const C b1 = {a0.x - a2.x, a0.y - a2.y};
                const C b2 = {a1.x + a3.x, a1.y + a3.y};
                const C b3 = {a1.y - a3.y, -(a1.x - a3.x)};  // (a1 - a3) * (-i)
                v[c][0] = {b0.x + b2.x, b0.y + b2.y};
                v[c][2] = {b0.x - b2.x, b0.y - b2.y};
                v[c][1] = {b1.x + b3.x, b1.y + b3.y};
                v[c][3] = {b1.x - b3.x, b1.y - b3.y};
                outi[c] = (j / ns) * 4 * ns + k;
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int o = outi[c];
            if (o >= 0) {
                buf[o] = v[c][0];
                buf[o + ns] = v[c][1];
                buf[o + 2 * ns] = v[c][2];
                buf[o + 3 * ns] = v[c][3];
            }
        }
        __syncthreads();
    }
    if (ns < N) {
        // radix 2 over span ns = N / 2: butterfly j < ns reads and writes j and
        // j + ns only (k = j, w^k = tw[j]), so it runs in place
        for (int j = threadIdx.x; j < ns; j += blockDim.x) {
            const C a0 = buf[j];
            const TW t = tw[j];
            const C a1 = cmul(buf[j + ns], {t.x, t.y});
            buf[j] = {a0.x + a1.x, a0.y + a1.y};
            buf[j + ns] = {a0.x - a1.x, a0.y - a1.y};
        }
        __syncthreads();
    }
}

// N is a power of two the LDS FFT takes (with 256 threads).
static inline bool lds_fft_size_ok(int N) { return N >= 16 && N <= 2048 && (N & (N - 1)) == 0; }

}  // namespace sicc
