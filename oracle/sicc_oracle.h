/*
 * sicc_oracle.h -- CPU restatement of the reference receive-and-decode path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the parity tests,
 * the smoke test and the CPU-baseline leg of bench.py compare the HIP product
 * against.  Nothing in the product (the sicc-baseband-proto_amd package or its
 * shared library) may link, load or call it.
 *
 * Every function names the reference file:line it restates.  Reference paths
 * are relative to the reference repository root:
 *   LINK  = source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp
 *   MOSD  = source/matlab/uhd_cpp/matlab/cpu_bp_osd_decode.m
 *   PIPE  = source/cpp_cuda/uhd_cpp/src/gpu_ofdm_pipeline.cu
 *
 * Pinning status of each function (how it is checked independently of the
 * product, see tests/test_oracle_*.py):
 *   oracle_alist_parse      pinned: CCSDS header/degree counts, H*G^T = 0.
 *   oracle_build_ldpc       pinned: H*G^T = 0, rank, front-info layout.
 *   oracle_encode           pinned: syndrome of every output is zero, linearity.
 *   oracle_nms_decode       pinned: single-check closed form, tree code max-log
 *                           MAP by brute force, 1-iteration decode of clean codewords,
 *                           VN damping by the linear degree-2 closed form (two passes,
 *                           damping-free fixed point).
 *   oracle_osd_decode       pinned: output in code, order-k OSD == brute-force ML,
 *                           argmin over explicit TEP set, basis independence, ties.
 *   oracle_ofdm_*           pinned: numpy FFT, closed-form QPSK LLR, brute force
 *                           sliding metric, noise-free known-answer loopback.
 *   oracle_demod_frame_ex   pinned: SNR / EVM equal to the reference decode_frame
 *                           (golden fixture written by its own CPU path).
 *   oracle_frame_finish     pinned: every FrameDecodeResult field equal to the
 *                           reference decode_frame in its reference-test mode
 *                           (golden fixture); SNR-gate branch by construction.
 */
#ifndef SICC_ORACLE_H
#define SICC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- code model (LINK:2722-3160) ---- */
int oracle_alist_parse(const char* path, int* n_out, int* m_out, int* e_out,
                       int* row_ptr, int* col_ind, int edge_cap);
int oracle_build_ldpc(int n, int m, const int* row_ptr, const int* col_ind, int front_info,
                      int* k_out, int* info_cols, int* pivot_cols, uint64_t* rref_rows,
                      int* accumulator_out);
void oracle_encode(int n, int m, int k, int accumulator, const int* row_ptr, const int* col_ind,
                   const int* info_cols, const int* pivot_cols, const uint64_t* rref_rows,
                   const uint8_t* info, uint8_t* cw);
void oracle_encode_batch(int count, int n, int m, int k, int accumulator, const int* row_ptr,
                         const int* col_ind, const int* info_cols, const int* pivot_cols,
                         const uint64_t* rref_rows, const uint8_t* info, uint8_t* cw);

/* ---- normalized min-sum BP (LINK:3013-3094, MOSD:101-185) ---- */
int oracle_nms_decode(int n, int m, const int* row_ptr, const int* col_ind, const float* llr,
                      int max_iter, float alpha, float offset, float damping, int early_stop,
                      uint8_t* hard_out, float* app_out, float* best_app_out,
                      int* iters_out, int* best_synw_out);

/* ---- OSD (MOSD:187-305) ---- */
int oracle_osd_decode(int n, int k, const uint8_t* gen, const float* llr, int order,
                      int max_flip, int max_tests, uint8_t* code_out, double* metric_out,
                      int* tests_out);

/* ---- layered normalized min-sum (GPU "colored layered" schedule) ----
 * oracle_layer_order: greedy first-fit row colouring in row order; order_out
 * [m] = rows sorted by (colour, row), color_out [m] optional; returns the
 * number of colours.  oracle_nms_layered_decode processes rows in `order`. */
int oracle_layer_order(int n, int m, const int* row_ptr, const int* col_ind, int* order_out, int* color_out);
int oracle_nms_layered_decode(int n, int m, const int* row_ptr, const int* col_ind, const int* order,
                              const float* llr, int max_iter, float alpha, float offset, float damping,
                              int early_stop, uint8_t* hard_out, float* app_out, int* iters_out);

/* ---- BP followed by OSD on failure, over a batch (MOSD:55-77) ---- */
int oracle_bposd_batch(int count, int n, int m, int k, const int* row_ptr, const int* col_ind,
                       const uint8_t* gen, const int* info_cols, const float* llr, int max_iter,
                       float alpha, float offset, float damping, int osd_order, int max_flip,
                       int max_tests, int threads, uint8_t* code_out, uint8_t* info_out,
                       uint8_t* bp_ok_out, uint8_t* used_osd_out, int* iters_out);

/* ---- OFDM receive (LINK:3788-3860, LINK:5123-5212, LINK:5331-5521) ---- */
void oracle_fft(double* re, double* im, int n, int inverse);
int oracle_find_frame(const float* iq, int count, int half_len, const float* preamble,
                      int pre_len, int refine_radius, int max_search, double threshold, double fs,
                      int* start_out, double* peak_out, double* cfo_out, double* coarse_out);
void oracle_demod_frame(const float* frame, int nfft, int cp, int pre_len, int num_symbols,
                        int active_sc, const int* used_shifted, int n_pilot, const int* pilot_sym,
                        const float* pilot_grid, int n_data, const int* data_sym,
                        int bits_per_symbol, const float* constellation, double cfo_hz, double fs,
                        double noise_var, float* llr_out);
/* oracle_demod_frame plus the SNR / EVM estimates of LINK:5431-5467 (known noise
 * variance); snr_db / evm_rms may be NULL. */
void oracle_demod_frame_ex(const float* frame, int nfft, int cp, int pre_len, int num_symbols,
                           int active_sc, const int* used_shifted, int n_pilot, const int* pilot_sym,
                           const float* pilot_grid, int n_data, const int* data_sym,
                           int bits_per_symbol, const float* constellation, double cfo_hz, double fs,
                           double noise_var, float* llr_out, double* snr_db, double* evm_rms);

/* ---- per-frame FEC aggregation (LINK:5482-5591) ----
 * block_parity / block_iters [blocks]; info [blocks*k] decoded; ref_info
 * [blocks*k] and ref_coded [blocks*n] (its encoding) or NULL for no reference
 * test; llr [blocks*n] (hard decisions for the pre-FEC count). */
void oracle_frame_finish(int blocks, int n, int k, const uint8_t* block_parity, const int* block_iters,
                         const uint8_t* info, const uint8_t* ref_info, const uint8_t* ref_coded,
                         const float* llr, int gated, int* parity_ok, int* iterations, uint32_t* frame_id,
                         int* bit_errors, int* pre_fec_bit_errors, int* frame_ok);

/* ---- OFDM transmit, GPU transmitter semantics (PIPE:129-342) ----
 * info [blocks][k] bytes; gen_kn [k][n] bytes (generator path) or, when
 * dvb_s2, row_info_offsets [n-k+1] / row_info_cols (accumulator path).
 * coded_out [coded_bits] (pad bits beyond blocks * n are 0); frame_out
 * [pre_len + num_symbols * (nfft + cp)] interleaved IQ.  precomp_h / actual_h
 * are [num_symbols][active_sc] IQ or NULL; noise_var > 0 adds the counter-based
 * AWGN of seed on data resource elements.  Returns 0, or -1 on a bad shape. */
int oracle_tx_frame(const uint8_t* info, int n, int k, int blocks, const uint8_t* gen_kn, int dvb_s2,
                    const int* row_info_offsets, const int* row_info_cols, int coded_bits, int nfft, int cp,
                    int pre_len, int num_symbols, int active_sc, int bits_per_symbol, const int* pilot_index,
                    const int* data_index, const int* used_shifted, const float* pilot_grid,
                    const float* preamble, const float* constellation, const float* precomp_h,
                    const float* actual_h, float amplitude, float noise_var, uint64_t seed,
                    uint8_t* coded_out, float* frame_out);

#ifdef __cplusplus
}
#endif
#endif
