// frame.hpp -- device argument block and launcher of the per-frame result
// kernel (the reference's FrameDecodeResult aggregation).
#pragma once

#include "../../include/sicc_frame_c_api.h"
#include "kernels.hpp"
#include "ofdm.hpp"

#include <cstdint>

namespace sicc {

// Frame f's blocks are codewords f * blocks .. f * blocks + blocks - 1 of the
// decode; its information bits are info[f][blocks * k], its LLRs
// llr[f][blocks * n].  Null optional inputs: no sync (LLR path, every frame
// detected), no statistics (SNR 0, EVM NaN), no reference payload (no error
// counts), no generator (no pre-FEC count).
struct FrameFinishArgs {
    int frames = 0;
    int blocks = 0;
    int n = 0;
    int k = 0;
    const uint8_t* bp_ok = nullptr;       // [frames * blocks]
    const uint8_t* used_osd = nullptr;    // [frames * blocks] or null (BP-only decoder)
    const int* iterations = nullptr;      // [frames * blocks]
    uint8_t* info = nullptr;              // [frames][blocks * k]; zeroed for frames not decoded
    const uint8_t* ref_info = nullptr;    // [frames][blocks * k] reference payload
    const float* llr = nullptr;           // [frames][blocks * n]
    // generator columns c0 .. n - 1 (c0 = k for a systematic-front code, whose
    // first k columns are the identity), word-major: bit r % 64 of
    // gen_cols[(r / 64) * (n - c0) + (c - c0)] is G[r][c]
    const unsigned long long* gen_cols = nullptr;
    int kw = 0;
    int c0 = 0;
    int gen_in_lds = 0;
    const OfdmSyncResult* sync = nullptr;  // [frames]
    int span = 0;
    int frame_samples = 0;
    const double* stats = nullptr;         // [frames][n_data][4] (OfdmDemodArgs::stats)
    int n_data = 0;
    double n_re = 0.0;                     // data resource elements per frame
    double noise_var = 0.0;
    double snr_gate_db = -1e300;
    uint8_t* block_ok = nullptr;           // optional out [frames * blocks]
    sicc_frame_result_t* out = nullptr;    // [frames]
};

bool frame_gen_fits_lds(int ncols, int kw);
void launch_frame_finish(const FrameFinishArgs& a, hipStream_t s);

}  // namespace sicc
