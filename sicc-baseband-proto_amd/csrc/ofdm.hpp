// ofdm.hpp -- device argument blocks and launchers of the OFDM receive kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sicc {

// Per-frame synchronization result (device).  start < 0: no frame passed the
// threshold.  cfo_hz_d keeps the double estimate the demodulator uses.
struct OfdmSyncResult {
    int start;
    float metric;         // refined peak (full-preamble normalized correlation)
    float cfo_hz;
    float coarse_metric;  // Schmidl-Cox metric at the coarse argmax
    double cfo_hz_d;
};

struct OfdmSyncArgs {
    const float2* rx = nullptr;  // [frames][span]
    int span = 0;
    int half_len = 0;
    int max_search = 0;          // coarse window [0, max_search]; <= 0: whole row
    int radius = 16;             // refinement half-width
    double threshold = 0.45;
    double fs = 15.36e6;
    const float2* preamble = nullptr;  // [2 * half_len]
    const double2* preamble_d = nullptr;  // the same samples in double
    double pre_energy = 0.0;           // sum of |preamble|^2 (float norms, double sum)
    OfdmSyncResult* out = nullptr;     // [frames]
    int window_in_lds = 1;             // set by the launcher
    int prefix = 0;                    // set by the launcher: prefix-sum coarse search
};

struct OfdmDemodArgs {
    const float2* rx = nullptr;  // [frames][span], frame f starts at sync[f].start
    int span = 0;
    int frame_samples = 0;  // preamble_len + num_symbols * (nfft + cp)
    const OfdmSyncResult* sync = nullptr;
    int nfft = 1024;
    int cp = 72;
    int preamble_len = 1024;
    int active_sc = 600;
    int bits_per_symbol = 4;
    int n_pilot = 1;
    int n_data = 13;
    const int* pilot_symbols = nullptr;  // [n_pilot]
    const int* data_symbols = nullptr;   // [n_data]
    const int* used_natural = nullptr;   // [active_sc] natural FFT bin of each subcarrier
    const float2* pilot_grid = nullptr;  // [n_pilot][active_sc]
    const int* interp_p0 = nullptr;      // [n_data]
    const int* interp_p1 = nullptr;
    const float* interp_w0 = nullptr;    // (float)(1 - alpha)
    const float* interp_w1 = nullptr;    // (float)alpha
    const float2* constellation = nullptr;  // [2^bits]
    const double2* twiddle = nullptr;       // [nfft] exp(-2 pi i m / nfft)
    float noise_var = 1e-3f;
    double noise_var_d = 1e-3;
    double inv_noise_var_d = 1e3;  // 1 / noise_var_d
    double fs = 15.36e6;
    int use_pilots = 1;
    float* llr = nullptr;  // [frames][llr_per_frame]
    int llr_per_frame = 0;
    float2* h_est = nullptr;  // scratch [frames][n_pilot][active_sc] LS estimates
    // Optional per-(frame, data symbol) sums over the symbol's active
    // subcarriers, pad resource elements included, of |h|^2, |y - h hard|^2,
    // |eq0 - hard|^2 and |hard|^2 (eq0 = y conj(h) / max(|h|^2, 1e-9), hard =
    // the nearest constellation point): [frames][n_data][4] doubles, zeros for a
    // frame sync did not deliver.  The SNR / EVM estimates of decode_frame
    // (uhd_ldpc_ofdm_link.cpp:5431-5467) are ratios of their per-frame totals.
    double* stats = nullptr;
};

void launch_ofdm_sync(const OfdmSyncArgs& a, int frames, hipStream_t s);
void launch_ofdm_demod(const OfdmDemodArgs& a, int frames, hipStream_t s);

// Transmit side (the reference GPU transmitter, gpu_ofdm_pipeline.cu:129-342).
struct OfdmTxArgs {
    // encoder: info [frames][blocks][k] bytes -> coded [frames][coded_bits] bytes
    const uint8_t* info = nullptr;
    int code_n = 0;
    int code_k = 0;
    int blocks = 0;
    // generator encoder: bit r of gen_cols[c * gen_words + r / 64] is G[r][c]
    const unsigned long long* gen_cols = nullptr;
    int gen_words = 0;
    // DVB-S2 accumulator encoder (used when dvb_s2 != 0): systematic info in
    // bits [0, k), parity p_r = rowsum_r ^ p_{r-1} in bits k + r, where rowsum_r
    // is the XOR of info[row_info_cols[j]] for j in [row_info_offsets[r], row_info_offsets[r + 1]).
    int dvb_s2 = 0;
    const int* row_info_offsets = nullptr;  // [n - k + 1]
    const int* row_info_cols = nullptr;
    uint8_t* coded = nullptr;               // [frames][coded_bits]
    int coded_bits = 0;                     // data capacity in bits (pad bits are 0)
    // grid + IFFT + compose
    int nfft = 1024;
    int cp = 72;
    int preamble_len = 1024;
    int num_symbols = 14;
    int active_sc = 600;
    int bits_per_symbol = 4;
    const int* pilot_index = nullptr;     // [num_symbols] pilot ordinal or -1
    const int* data_index = nullptr;      // [num_symbols] data ordinal or -1
    const int* used_natural = nullptr;    // [active_sc]
    const float2* pilot_grid = nullptr;   // [n_pilot][active_sc]
    const float2* preamble = nullptr;     // [preamble_len]
    const float2* constellation = nullptr;
    const float2* precomp_h = nullptr;    // optional [num_symbols][active_sc]
    const float2* actual_h = nullptr;     // optional [num_symbols][active_sc]
    const float2* twiddle = nullptr;      // [nfft] exp(-2 pi i m / nfft), single precision
    float amplitude = 0.2f;
    float noise_var = 0.0f;               // frequency-domain AWGN variance on data REs (0: none)
    unsigned long long noise_seed = 0;    // frame f of a batch uses noise_seed + f
    float2* out = nullptr;                // [frames][frame_samples]
};
void launch_ofdm_tx(const OfdmTxArgs& a, int frames, hipStream_t s);

}  // namespace sicc
