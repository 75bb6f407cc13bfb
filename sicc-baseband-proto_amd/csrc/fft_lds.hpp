// fft_lds.hpp -- FP64 radix-4 Stockham FFT through LDS for one workgroup
// (device code shared by the receive and transmit kernels).
#pragma once

#include <hip/hip_runtime.h>

namespace sicc {

struct Cd {
    double x, y;
};

static __device__ __forceinline__ Cd cmul(Cd a, Cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

// 1024-point (or any 4^k) forward FFT, Stockham radix-4, in LDS, FP64.
// buf holds the input in natural order and receives the output in natural order.
static __device__ void fft_radix4(Cd* buf, const double2* tw, int N)
{
    const int q = N >> 2;
    for (int ns = 1; ns < N; ns <<= 2) {
        Cd v[4][2];
        int outi[2];
        int cnt = 0;
        for (int j = threadIdx.x; j < q; j += blockDim.x, ++cnt) {
            const int k = j & (ns - 1);
            Cd a0 = buf[j], a1 = buf[j + q], a2 = buf[j + 2 * q], a3 = buf[j + 3 * q];
            if (ns > 1) {
                // twiddle w^(r k), w = exp(-2 pi i / (4 ns)); table entry m = exp(-2 pi i m / N)
                const int step = N / (4 * ns);
                const double2 t1 = tw[(k * step) & (N - 1)];
                const double2 t2 = tw[(2 * k * step) & (N - 1)];
                const double2 t3 = tw[(3 * k * step) & (N - 1)];
                a1 = cmul(a1, {t1.x, t1.y});
                a2 = cmul(a2, {t2.x, t2.y});
                a3 = cmul(a3, {t3.x, t3.y});
            }
            const Cd b0 = {a0.x + a2.x, a0.y + a2.y};
            const Cd b1 = {a0.x - a2.x, a0.y - a2.y};
            const Cd b2 = {a1.x + a3.x, a1.y + a3.y};
            const Cd b3 = {a1.y - a3.y, -(a1.x - a3.x)};  // (a1 - a3) * (-i)
            v[0][cnt] = {b0.x + b2.x, b0.y + b2.y};
            v[2][cnt] = {b0.x - b2.x, b0.y - b2.y};
            v[1][cnt] = {b1.x + b3.x, b1.y + b3.y};
            v[3][cnt] = {b1.x - b3.x, b1.y - b3.y};
            outi[cnt] = (j / ns) * 4 * ns + k;
        }
        __syncthreads();
        for (int c = 0; c < cnt; ++c) {
            const int o = outi[c];
            buf[o] = v[0][c];
            buf[o + ns] = v[1][c];
            buf[o + 2 * ns] = v[2][c];
            buf[o + 3 * ns] = v[3][c];
        }
        __syncthreads();
    }
}

}  // namespace sicc
