/*
 * sicc_ofdm_c_api.h -- C ABI of the batched OFDM receive path in libsicc_amd.so.
 *
 * The reference exposes its GPU receive chain as the C++ class
 * usrp_gpu_pipeline_bridge::Pipeline (source/cpp_cuda/uhd_cpp/src/
 * gpu_ofdm_pipeline.hpp:65-93, mirrored in gpu_ofdm_pipeline.hpp of this
 * library).  These entry points are the plain-C layer under that class, and
 * the batched form an FFI caller binds: many frames per call, device or host
 * buffers, no C++ types.  Status codes and cuda_osd_get_last_error() as in
 * matlab_decoder_c_api.h.
 *
 * Receive semantics (reference CPU receiver, uhd_ldpc_ofdm_link.cpp):
 *   sync   find_frame LINK:5123-5212 (coarse window [0, max_rx_search_samples]);
 *   demod  decode_frame LINK:5354-5521 with the configured noise variance.
 * IQ is interleaved float32 (re, im).  Frame f of a batch is row f of a
 * [frames][span] sample array and is searched from the row's first sample.
 * LLRs: [frames][blocks_per_frame * code_n] float32, positive = bit 0, bit
 * order (data symbol, subcarrier, bit); resource elements beyond
 * blocks_per_frame * code_n carry pad bits and yield no LLR.
 */
#ifndef SICC_OFDM_C_API_H
#define SICC_OFDM_C_API_H

#ifdef __cplusplus
extern "C" {
#endif

/* Field meaning as usrp_gpu_pipeline_bridge::Config (gpu_ofdm_pipeline.hpp:12-37). */
typedef struct sicc_ofdm_config_t {
    int nfft;
    int cp;
    int preamble_len;
    int pre_half_len;
    int num_symbols;
    int active_sc;
    int bits_per_symbol;
    int code_n;
    int blocks_per_frame;
    int max_rx_search_samples; /* coarse search candidates 0..this; <= 0: whole row */
    int refine_radius;         /* < 0: min(16, cp) as the reference CPU receiver */
    int use_pilot_channel_est;
    float sample_rate;
    float sync_threshold;
    float demod_noise_var;
} sicc_ofdm_config_t;

/* Per-frame sync result; start < 0 when no frame passed the threshold. */
typedef struct sicc_ofdm_sync_t {
    int start;
    float metric;
    float cfo_hz;
    float coarse_metric;
    double cfo_hz_d;
} sicc_ofdm_sync_t;

/* Create a receiver.  preamble_iq [preamble_len], used_shifted [active_sc]
 * (shifted-grid subcarrier indices, LINK:4011-4018), pilot_symbols [n_pilot],
 * pilot_grid_iq [n_pilot][active_sc]; the data symbols are the others in order
 * and the constellation is Gray QAM of bits_per_symbol (LINK:2362-2414). */
int sicc_ofdm_create(const sicc_ofdm_config_t* config, const float* preamble_iq, const int* used_shifted,
                     int n_pilot, const int* pilot_symbols, const float* pilot_grid_iq,
                     unsigned long long* out_handle);
int sicc_ofdm_destroy(unsigned long long handle);
int sicc_ofdm_llr_per_frame(unsigned long long handle);

/* Device-resident batches on `hip_stream` (null: the receiver's own stream), asynchronous.
 * Demodulation keeps per-frame channel estimates in receiver-owned scratch:
 * calls on one handle must be issued on one stream (or must not overlap). */
int sicc_ofdm_sync_device(unsigned long long handle, const float* d_rx_iq, int frames, int span,
                          sicc_ofdm_sync_t* d_sync, void* hip_stream);
int sicc_ofdm_demod_device(unsigned long long handle, const float* d_rx_iq, int frames, int span,
                           const sicc_ofdm_sync_t* d_sync, float* d_llr, void* hip_stream);
/* sync then demod */
int sicc_ofdm_receive_device(unsigned long long handle, const float* d_rx_iq, int frames, int span,
                             sicc_ofdm_sync_t* d_sync, float* d_llr, void* hip_stream);
/* Synchronous host-buffer variant. */
int sicc_ofdm_receive_host(unsigned long long handle, const float* rx_iq, int frames, int span,
                           sicc_ofdm_sync_t* sync_out, float* llr_out);

#ifdef __cplusplus
}
#endif

#endif /* SICC_OFDM_C_API_H */
