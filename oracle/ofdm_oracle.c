/*
 * ofdm_oracle.c -- CPU restatement of the reference OFDM receive path.
 * TEST INFRASTRUCTURE ONLY (checker for tests/, smoke and bench.py's CPU leg).
 *
 * Sources restated (paths relative to the reference repository root):
 *   LINK = source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp
 *     fft_inplace            LINK:3788-3827
 *     fft_to_shifted_grid    LINK:3845-3859
 *     find_frame             LINK:5123-5212
 *     decode_frame (demod)   LINK:5331-5521, known-noise-variance branch
 *     symbol_to_llr_maxlog   LINK:2552-2568
 *
 * Complex arithmetic is written out as std::complex performs it in the
 * reference build: products (ac - bd, ad + bc), division by a real scalar per
 * component, std::norm(z) = re*re + im*im in the element type.
 * IQ buffers are interleaved float pairs (re, im).
 */
#include "sicc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double kPi = 3.141592653589793238462643383279502884;

/* std::norm of a complex<float> in libstdc++: re*re + im*im in float. */
static float normf(float re, float im) { return re * re + im * im; }

/* LINK:3788-3827: in-place radix-2, bit reversal then butterflies with the
 * twiddle advanced by repeated multiplication w *= wlen; inverse scales 1/n. */
void oracle_fft(double* re, double* im, int n, int inverse)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i];
            re[i] = re[j];
            re[j] = t;
            t = im[i];
            im[i] = im[j];
            im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const double ang = 2.0 * kPi / (double)len * (inverse ? 1.0 : -1.0);
        const double wr = cos(ang), wi = sin(ang);
        for (int i = 0; i < n; i += len) {
            double cr = 1.0, ci = 0.0;
            for (int j = 0; j < len / 2; ++j) {
                const double ur = re[i + j], ui = im[i + j];
                const double xr = re[i + j + len / 2], xi = im[i + j + len / 2];
                const double vr = xr * cr - xi * ci;
                const double vi = xr * ci + xi * cr;
                re[i + j] = ur + vr;
                im[i + j] = ui + vi;
                re[i + j + len / 2] = ur - vr;
                im[i + j + len / 2] = ui - vi;
                const double nr = cr * wr - ci * wi;
                const double ni = cr * wi + ci * wr;
                cr = nr;
                ci = ni;
            }
        }
    }
    if (inverse) {
        const double s = 1.0 / (double)n;
        for (int i = 0; i < n; ++i) {
            re[i] *= s;
            im[i] *= s;
        }
    }
}

/* LINK:5123-5212 with an optional bound on the coarse search (as the
 * reference's max-search-samples option, LINK:5214-5246): candidates d in
 * [0, min(count - pre_len, max_search)] (max_search <= 0: all of them),
 * refinement over [best - radius, best + radius] clamped to
 * [0, count - pre_len].
 * Returns 1 when a frame passes both thresholds. */
int oracle_find_frame(const float* iq, int count, int half_len, const float* preamble, int pre_len,
                      int refine_radius, int max_search, double threshold, double fs, int* start_out,
                      double* peak_out, double* cfo_out, double* coarse_out)
{
    const int L = half_len;
    *start_out = -1;
    *peak_out = 0.0;
    *cfo_out = 0.0;
    if (coarse_out) *coarse_out = 0.0;
    if (count < pre_len || pre_len != 2 * L) return 0;
    const int last = count - pre_len; /* highest candidate */
    const int limit = (max_search > 0 && max_search < last) ? max_search : last;
    if (refine_radius < 0) refine_radius = 0;
    double pr = 0.0, pi = 0.0, e1 = 0.0, e2 = 0.0;
    for (int i = 0; i < L; ++i) {
        const double ar = iq[2 * i], ai = iq[2 * i + 1];
        const double br = iq[2 * (i + L)], bi = iq[2 * (i + L) + 1];
        pr += ar * br + ai * bi; /* conj(a) * b */
        pi += ar * bi - ai * br;
        e1 += (double)normf(iq[2 * i], iq[2 * i + 1]);
        e2 += (double)normf(iq[2 * (i + L)], iq[2 * (i + L) + 1]);
    }
    int best = 0;
    double best_metric = 0.0;
    for (int d = 0; d <= limit; ++d) {
        const double den = e1 * e2 > 1e-18 ? e1 * e2 : 1e-18;
        const double metric = (pr * pr + pi * pi) / den;
        if (metric > best_metric) {
            best_metric = metric;
            best = d;
        }
        if (d + pre_len >= count) break;
        const double x0r = iq[2 * d], x0i = iq[2 * d + 1];
        const double x1r = iq[2 * (d + L)], x1i = iq[2 * (d + L) + 1];
        const double x2r = iq[2 * (d + pre_len)], x2i = iq[2 * (d + pre_len) + 1];
        /* p += conj(x1) x2 - conj(x0) x1 */
        pr += (x1r * x2r + x1i * x2i) - (x0r * x1r + x0i * x1i);
        pi += (x1r * x2i - x1i * x2r) - (x0r * x1i - x0i * x1r);
        /* float difference of float norms, then accumulated in double */
        e1 += (double)(normf(iq[2 * (d + L)], iq[2 * (d + L) + 1]) - normf(iq[2 * d], iq[2 * d + 1]));
        e2 += (double)(normf(iq[2 * (d + pre_len)], iq[2 * (d + pre_len) + 1]) -
                       normf(iq[2 * (d + L)], iq[2 * (d + L) + 1]));
    }
    if (coarse_out) *coarse_out = best_metric;
    if (best_metric < threshold) return 0;
    const int lo = best > refine_radius ? best - refine_radius : 0;
    const int hi = (best + refine_radius < last) ? best + refine_radius : last;
    double pre_energy = 0.0;
    for (int i = 0; i < pre_len; ++i) pre_energy += (double)normf(preamble[2 * i], preamble[2 * i + 1]);
    double best_corr_metric = 0.0;
    int best_corr = best;
    for (int d = lo; d <= hi; ++d) {
        double cr = 0.0, ci = 0.0, eng = 0.0;
        for (int i = 0; i < pre_len; ++i) {
            const double yr = iq[2 * (d + i)], yi = iq[2 * (d + i) + 1];
            const double qr = preamble[2 * i], qi = preamble[2 * i + 1];
            cr += yr * qr + yi * qi; /* y * conj(p) */
            ci += yi * qr - yr * qi;
            eng += (double)normf(iq[2 * (d + i)], iq[2 * (d + i) + 1]);
        }
        const double den = eng * pre_energy > 1e-18 ? eng * pre_energy : 1e-18;
        const double m = (cr * cr + ci * ci) / den;
        if (m > best_corr_metric) {
            best_corr_metric = m;
            best_corr = d;
        }
    }
    *peak_out = best_corr_metric;
    if (best_corr_metric < threshold) return 0;
    double ar = 0.0, ai = 0.0;
    for (int i = 0; i < L; ++i) {
        const double xr = iq[2 * (best_corr + i)], xi = iq[2 * (best_corr + i) + 1];
        const double yr = iq[2 * (best_corr + i + L)], yi = iq[2 * (best_corr + i + L) + 1];
        ar += xr * yr + xi * yi;
        ai += xr * yi - xi * yr;
    }
    *start_out = best_corr;
    *cfo_out = atan2(ai, ar) * fs / (2.0 * kPi * (double)L);
    return 1;
}

/* LINK:5331-5521 demodulation with a known noise variance: CFO derotation with
 * the phase origin at the frame start, per-symbol FP64 FFT to the shifted grid,
 * LS pilot estimate, linear interpolation between pilot symbols, equalization
 * eq = y conj(h) / (|h|^2 + s2), s2_eq = s2 / |h|^2, max-log LLR
 * (d1 - d0) / s2_eq over the constellation (LINK:2552-2568).  LLR order:
 * ((data symbol * active_sc) + subcarrier) * bits + bit. */
void oracle_demod_frame(const float* frame, int nfft, int cp, int pre_len, int num_symbols, int active_sc,
                        const int* used_shifted, int n_pilot, const int* pilot_sym, const float* pilot_grid,
                        int n_data, const int* data_sym, int bits_per_symbol, const float* constellation,
                        double cfo_hz, double fs, double noise_var, float* llr_out)
{
    oracle_demod_frame_ex(frame, nfft, cp, pre_len, num_symbols, active_sc, used_shifted, n_pilot, pilot_sym,
                          pilot_grid, n_data, data_sym, bits_per_symbol, constellation, cfo_hz, fs, noise_var,
                          llr_out, NULL, NULL);
}

/* LINK:2429-2441 nearest_constellation_symbol: the first point of least
 * squared distance (strict < in table order). */
static int nearest_point(float er, float ei, const float* constellation, int M)
{
    int best = 0;
    float best_d2 = INFINITY;
    for (int i = 0; i < M; ++i) {
        const float d2 = normf(er - constellation[2 * i], ei - constellation[2 * i + 1]);
        if (d2 < best_d2) {
            best_d2 = d2;
            best = i;
        }
    }
    return best;
}

/* As oracle_demod_frame, plus the frame's SNR and EVM estimates of
 * LINK:5431-5467 when snr_db / evm_rms are non-null: over every data resource
 * element (y, h), h2 = max(|h|^2, 1e-9), eq0 = y conj(h) / h2, hard = the
 * nearest constellation point to eq0; sums of |y - h hard|^2, |eq0 - hard|^2,
 * |hard|^2 and |h|^2 in double in (data symbol, subcarrier) order; with the
 * known noise variance nv = max(noise_var, 1e-8):
 * snr_db = 10 log10(max(mean |h|^2 / nv, 1e-12)),
 * evm_rms = sqrt(max(sum |eq0 - hard|^2, 0) / max(sum |hard|^2, 1e-12)). */
void oracle_demod_frame_ex(const float* frame, int nfft, int cp, int pre_len, int num_symbols, int active_sc,
                           const int* used_shifted, int n_pilot, const int* pilot_sym, const float* pilot_grid,
                           int n_data, const int* data_sym, int bits_per_symbol, const float* constellation,
                           double cfo_hz, double fs, double noise_var, float* llr_out, double* snr_db,
                           double* evm_rms)
{
    const int sym_len = nfft + cp;
    const int frame_len = pre_len + num_symbols * sym_len;
    float* y = (float*)malloc(sizeof(float) * 2 * (size_t)num_symbols * active_sc);
    double* re = (double*)malloc(sizeof(double) * (size_t)nfft);
    double* im = (double*)malloc(sizeof(double) * (size_t)nfft);
    float* rot = (float*)malloc(sizeof(float) * 2 * (size_t)frame_len);
    for (int n = 0; n < frame_len; ++n) {
        const double a = -2.0 * kPi * cfo_hz * (double)n / fs;
        const float c = (float)cos(a), s = (float)sin(a);
        const float xr = frame[2 * n], xi = frame[2 * n + 1];
        rot[2 * n] = xr * c - xi * s;
        rot[2 * n + 1] = xr * s + xi * c;
    }
    for (int sym = 0; sym < num_symbols; ++sym) {
        const float* td = rot + 2 * (size_t)(pre_len + sym * sym_len + cp);
        for (int i = 0; i < nfft; ++i) {
            re[i] = td[2 * i];
            im[i] = td[2 * i + 1];
        }
        oracle_fft(re, im, nfft, 0);
        for (int sc = 0; sc < active_sc; ++sc) {
            const int nat = (used_shifted[sc] + nfft / 2) % nfft;
            y[2 * ((size_t)sym * active_sc + sc)] = (float)re[nat];
            y[2 * ((size_t)sym * active_sc + sc) + 1] = (float)im[nat];
        }
    }
    float* hp = (float*)malloc(sizeof(float) * 2 * (size_t)(n_pilot > 0 ? n_pilot : 1) * active_sc);
    for (int p = 0; p < n_pilot; ++p) {
        const int sym = pilot_sym[p];
        for (int sc = 0; sc < active_sc; ++sc) {
            const float pr = pilot_grid[2 * ((size_t)p * active_sc + sc)];
            const float pi = pilot_grid[2 * ((size_t)p * active_sc + sc) + 1];
            float den = normf(pr, pi);
            if (den < 1e-9f) den = 1e-9f;
            const float yr = y[2 * ((size_t)sym * active_sc + sc)], yi = y[2 * ((size_t)sym * active_sc + sc) + 1];
            /* y * conj(p) / den */
            const float tr = yr * pr - yi * (-pi);
            const float ti = yr * (-pi) + yi * pr;
            hp[2 * ((size_t)p * active_sc + sc)] = tr / den;
            hp[2 * ((size_t)p * active_sc + sc) + 1] = ti / den;
        }
    }
    const int M = 1 << bits_per_symbol;
    if (noise_var < 1e-8) noise_var = 1e-8;
    const float nv = (float)noise_var;
    double s_resid = 0.0, s_err = 0.0, s_ref = 0.0, s_h = 0.0;
    for (int di = 0; di < n_data; ++di) {
        const int sym = data_sym[di];
        int hi = 0;
        while (hi + 1 < n_pilot && sym > pilot_sym[hi + 1]) ++hi;
        int p0 = hi, p1 = (hi + 1 < n_pilot - 1) ? hi + 1 : n_pilot - 1;
        if (sym <= pilot_sym[0]) p0 = p1 = 0;
        else if (sym >= pilot_sym[n_pilot - 1]) p0 = p1 = n_pilot - 1;
        double alpha = 0.0;
        if (p0 != p1) alpha = (double)(sym - pilot_sym[p0]) / (double)(pilot_sym[p1] - pilot_sym[p0]);
        const float w0 = (float)(1.0 - alpha), w1 = (float)alpha;
        for (int sc = 0; sc < active_sc; ++sc) {
            const float h0r = hp[2 * ((size_t)p0 * active_sc + sc)], h0i = hp[2 * ((size_t)p0 * active_sc + sc) + 1];
            const float h1r = hp[2 * ((size_t)p1 * active_sc + sc)], h1i = hp[2 * ((size_t)p1 * active_sc + sc) + 1];
            const float hr = h0r * w0 + h1r * w1;
            const float hi_ = h0i * w0 + h1i * w1;
            float h2 = normf(hr, hi_);
            if (h2 < 1e-9f) h2 = 1e-9f;
            const float yr = y[2 * ((size_t)sym * active_sc + sc)], yi = y[2 * ((size_t)sym * active_sc + sc) + 1];
            const float tr = yr * hr - yi * (-hi_);
            const float ti = yr * (-hi_) + yi * hr;
            if (snr_db || evm_rms) {
                /* LINK:5431-5439 */
                const float qr = tr / h2, qi = ti / h2;
                const int hp_i = nearest_point(qr, qi, constellation, M);
                const float cr = constellation[2 * hp_i], ci = constellation[2 * hp_i + 1];
                const float rr = yr - (hr * cr - hi_ * ci);
                const float ri = yi - (hr * ci + hi_ * cr);
                s_resid += (double)normf(rr, ri);
                s_err += (double)normf(qr - cr, qi - ci);
                s_ref += (double)normf(cr, ci);
                s_h += (double)normf(hr, hi_);
            }
            const float den = h2 + nv;
            const float er = tr / den, ei = ti / den;
            double s2 = noise_var / (double)h2;
            if (s2 < 1e-8) s2 = 1e-8;
            float* dst = llr_out + ((size_t)di * active_sc + sc) * bits_per_symbol;
            for (int b = 0; b < bits_per_symbol; ++b) {
                float d0 = INFINITY, d1 = INFINITY;
                for (int i = 0; i < M; ++i) {
                    const float d2 = normf(er - constellation[2 * i], ei - constellation[2 * i + 1]);
                    if (((i >> (bits_per_symbol - 1 - b)) & 1) == 0) {
                        if (d2 < d0) d0 = d2;
                    } else if (d2 < d1) {
                        d1 = d2;
                    }
                }
                dst[b] = (float)((double)(d1 - d0) / s2);
            }
        }
    }
    (void)s_resid;  /* the decision-directed noise estimate: unused with a known variance */
    if (snr_db) {
        const double cnt = (double)((size_t)n_data * active_sc > 0 ? (size_t)n_data * active_sc : 1);
        const double r = (s_h / cnt) / noise_var;
        *snr_db = 10.0 * log10(r > 1e-12 ? r : 1e-12);
    }
    if (evm_rms) *evm_rms = sqrt((s_err > 0.0 ? s_err : 0.0) / (s_ref > 1e-12 ? s_ref : 1e-12));
    free(y);
    free(re);
    free(im);
    free(rot);
    free(hp);
}

/* ------------------------------------------------------------------ TX ----
 * Restatement of the reference GPU transmitter,
 * PIPE = source/cpp_cuda/uhd_cpp/src/gpu_ofdm_pipeline.cu:
 *   ldpc_encode_generator_kernel            PIPE:129-150
 *   DVB-S2 systematic copy / row sums / acc  PIPE:152-214
 *   fill_resource_grid_kernel (+ precomp)    PIPE:216-257, qam_from_bits_dev
 *   hash_u32 / uniform_01 / normal01         PIPE:259-280
 *   apply_actual_channel_and_frequency_awgn  PIPE:282-314
 *   cuFFT inverse (unnormalized) + compose   PIPE:316-342, PIPE:1032-1053
 * Float operations in the reference's order; the inverse FFT in double. */

static uint32_t tx_hash(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

static float tx_uniform(uint32_t seed, uint32_t idx)
{
    const uint32_t h = tx_hash(seed ^ (idx * 0x9e3779b9u));
    return ((float)h + 1.0f) * 2.3283064365386963e-10f;
}

static float tx_normal(uint32_t seed, uint32_t idx)
{
    float u1 = tx_uniform(seed, idx * 2u);
    if (u1 < 1.0e-7f) u1 = 1.0e-7f;
    const float u2 = tx_uniform(seed, idx * 2u + 1u);
    return sqrtf(-2.0f * logf(u1)) * cosf(2.0f * 3.14159265358979323846f * u2);
}

int oracle_tx_frame(const uint8_t* info, int n, int k, int blocks, const uint8_t* gen_kn, int dvb_s2,
                    const int* row_info_offsets, const int* row_info_cols, int coded_bits, int nfft, int cp,
                    int pre_len, int num_symbols, int active_sc, int bits_per_symbol, const int* pilot_index,
                    const int* data_index, const int* used_shifted, const float* pilot_grid,
                    const float* preamble, const float* constellation, const float* precomp_h,
                    const float* actual_h, float amplitude, float noise_var, uint64_t seed,
                    uint8_t* coded_out, float* frame_out)
{
    if (k <= 0 || n <= k || blocks <= 0 || blocks * n > coded_bits || nfft <= 0 || (nfft & (nfft - 1)) != 0)
        return -1;
    if (!dvb_s2 && !gen_kn) return -1;
    memset(coded_out, 0, (size_t)coded_bits);
    const int m = n - k;
    for (int b = 0; b < blocks; ++b) {
        const uint8_t* in = info + (size_t)b * k;
        uint8_t* out = coded_out + (size_t)b * n;
        if (dvb_s2) {
            for (int i = 0; i < k; ++i) out[i] = in[i] & 1u;
            uint8_t prev = 0;
            for (int r = 0; r < m; ++r) {
                uint8_t s = 0;
                for (int j = row_info_offsets[r]; j < row_info_offsets[r + 1]; ++j) s ^= in[row_info_cols[j]] & 1u;
                const uint8_t p = (uint8_t)(s ^ (r == 0 ? 0u : prev));
                out[k + r] = p;
                prev = p;
            }
        } else {
            for (int c = 0; c < n; ++c) {
                uint8_t acc = 0;
                for (int r = 0; r < k; ++r) acc ^= (uint8_t)((in[r] & 1u) & (gen_kn[(size_t)r * n + c] & 1u));
                out[c] = acc;
            }
        }
    }
    const uint32_t seed32 = (uint32_t)(seed ^ (seed >> 32));
    const float sigma = sqrtf((noise_var > 0.0f ? noise_var : 0.0f) * 0.5f);
    const float scale = amplitude / (float)nfft;
    double* re = (double*)malloc(sizeof(double) * (size_t)nfft);
    double* im = (double*)malloc(sizeof(double) * (size_t)nfft);
    float* g = (float*)malloc(sizeof(float) * 2 * (size_t)nfft);
    for (int i = 0; i < pre_len; ++i) {
        frame_out[2 * i] = preamble[2 * i] * amplitude;
        frame_out[2 * i + 1] = preamble[2 * i + 1] * amplitude;
    }
    for (int sym = 0; sym < num_symbols; ++sym) {
        for (int i = 0; i < 2 * nfft; ++i) g[i] = 0.0f;
        const int pi = pilot_index[sym], di = data_index[sym];
        for (int sc = 0; sc < active_sc; ++sc) {
            const int re_idx = sym * active_sc + sc;
            float vr = 0.0f, vi = 0.0f;
            if (pi >= 0) {
                vr = pilot_grid[2 * (pi * active_sc + sc)];
                vi = pilot_grid[2 * (pi * active_sc + sc) + 1];
            } else if (di >= 0) {
                const int base = (di * active_sc + sc) * bits_per_symbol;
                int idx = 0;
                for (int i = 0; i < bits_per_symbol; ++i)
                    idx = (idx << 1) | (base + i < coded_bits ? (coded_out[base + i] & 1) : 0);
                vr = constellation[2 * idx];
                vi = constellation[2 * idx + 1];
            }
            if (precomp_h) {
                const float hr = precomp_h[2 * re_idx], hi = precomp_h[2 * re_idx + 1];
                float h2 = normf(hr, hi);
                if (h2 < 1.0e-10f) h2 = 1.0e-10f;
                const float inv = 1.0f / h2;
                const float cr = hr * inv, ci = -hi * inv;
                const float tr = vr * cr - vi * ci, ti = vr * ci + vi * cr;
                vr = tr;
                vi = ti;
            }
            if (actual_h) {
                const float hr = actual_h[2 * re_idx], hi = actual_h[2 * re_idx + 1];
                const float tr = vr * hr - vi * hi, ti = vr * hi + vi * hr;
                vr = tr;
                vi = ti;
            }
            if (noise_var > 0.0f && di >= 0) {
                vr += sigma * tx_normal(seed32, (uint32_t)re_idx);
                vi += sigma * tx_normal(seed32 ^ 0xa5a5a5a5u, (uint32_t)re_idx);
            }
            const int bin = (used_shifted[sc] + nfft / 2) % nfft;
            g[2 * bin] = vr;
            g[2 * bin + 1] = vi;
        }
        for (int i = 0; i < nfft; ++i) {
            re[i] = g[2 * i];
            im[i] = g[2 * i + 1];
        }
        oracle_fft(re, im, nfft, 1); /* 1/n normalized; undo for cuFFT's unnormalized inverse */
        float* o = frame_out + 2 * ((size_t)pre_len + (size_t)sym * (nfft + cp));
        for (int i = 0; i < nfft + cp; ++i) {
            const int src = i < cp ? nfft - cp + i : i - cp;
            o[2 * i] = (float)(re[src] * nfft) * scale;
            o[2 * i + 1] = (float)(im[src] * nfft) * scale;
        }
    }
    free(re);
    free(im);
    free(g);
    return 0;
}
