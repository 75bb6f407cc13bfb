// fft_lds.hpp -- radix-4 Stockham FFT through LDS for one workgroup (device
// code shared by the receive kernels, FP64, and the transmitter, FP32).
#pragma once

#include <hip/hip_runtime.h>

namespace sicc {

struct Cd {
    double x, y;
};
struct Cf {
    float x, y;
};

static __device__ __forceinline__ Cd cmul(Cd a, Cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
static __device__ __forceinline__ Cf cmul(Cf a, Cf b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

// 1024-point (or any 4^k) forward FFT, Stockham radix-4, in LDS, in the
// precision of C (Cd with double2 twiddles, Cf with float2 twiddles).
// buf holds the input in natural order and receives the output in natural order.
template <class C, class TW>
static __device__ void fft_radix4(C* buf, const TW* tw, int N)
{
    const int q = N >> 2;
    for (int ns = 1; ns < N; ns <<= 2) {
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
}

}  // namespace sicc
