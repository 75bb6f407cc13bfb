/*
 * sicc_pipeline_c_api.h -- plain-C layer of usrp_gpu_pipeline_bridge::Pipeline
 * (gpu_ofdm_pipeline.hpp of this library; reference class
 * source/cpp_cuda/uhd_cpp/src/gpu_ofdm_pipeline.hpp:65-93).
 *
 * One function per Pipeline method, with the same arguments and meaning, for
 * callers that bind C (ctypes, cgo, MATLAB loadlibrary).  The C++ exceptions of
 * the class become status codes (matlab_decoder_c_api.h) with the message in
 * cuda_osd_get_last_error().  IQ is interleaved float32 (re, im).
 */
#ifndef SICC_PIPELINE_C_API_H
#define SICC_PIPELINE_C_API_H

#ifdef __cplusplus
extern "C" {
#endif

/* Fields and meaning of usrp_gpu_pipeline_bridge::Config (gpu_ofdm_pipeline.hpp:12-37). */
typedef struct sicc_pipeline_config_t {
    int nfft;
    int cp;
    int preamble_len;
    int pre_half_len;
    int num_symbols;
    int active_sc;
    int bits_per_symbol;
    int code_n;
    int code_k;
    int blocks_per_frame;
    int max_rx_search_samples;
    float sample_rate;
    float sync_threshold;
    float tx_amplitude;
    float demod_noise_var;
    int use_pilot_channel_est;
} sicc_pipeline_config_t;

/* usrp_gpu_pipeline_bridge::StaticTables (gpu_ofdm_pipeline.hpp:39-57) as
 * pointer + element count pairs; complex tables count complex elements. */
typedef struct sicc_pipeline_tables_t {
    const unsigned char* generator_kn; int generator_kn_count;
    int dvb_s2_accumulator;
    const int* row_info_offsets; int row_info_offsets_count;
    const int* row_info_cols; int row_info_cols_count;
    const int* used_indices; int used_indices_count;
    const int* pilot_symbol_index; int pilot_symbol_index_count;
    const int* data_symbol_index; int data_symbol_index_count;
    const int* pilot_symbols; int pilot_symbols_count;
    const int* data_symbols; int data_symbols_count;
    const int* interp_p0; int interp_p0_count;
    const int* interp_p1; int interp_p1_count;
    const float* interp_alpha; int interp_alpha_count;
    const float* pilot_grid_iq; int pilot_grid_count;
    const float* preamble_iq; int preamble_count;
    const float* precomp_h_iq; int precomp_h_count;
    const float* actual_h_iq; int actual_h_count;
} sicc_pipeline_tables_t;

/* usrp_gpu_pipeline_bridge::SyncResult */
typedef struct sicc_pipeline_sync_t {
    int start;
    float metric;
    float cfo_hz;
} sicc_pipeline_sync_t;

int sicc_pipeline_create(const sicc_pipeline_config_t* config, int slots, unsigned long long* out_handle);
int sicc_pipeline_destroy(unsigned long long handle);
int sicc_pipeline_upload_static_tables(unsigned long long handle, const sicc_pipeline_tables_t* tables);

/* Pipeline::submit_tx_frame_async / submit_tx_frame_device_async / synchronize_tx */
int sicc_pipeline_submit_tx_frame(unsigned long long handle, int slot, const unsigned char* info_bits,
                                  float* tx_frame_iq);
int sicc_pipeline_submit_tx_frame_device(unsigned long long handle, int slot, const unsigned char* info_bits,
                                         float frequency_noise_var, unsigned long long noise_seed,
                                         int add_frequency_noise, const float** d_tx_frame_iq);
int sicc_pipeline_synchronize_tx(unsigned long long handle, int slot);

/* Pipeline::find_frame / find_frame_device */
int sicc_pipeline_find_frame(unsigned long long handle, int slot, const float* samples_iq, int sample_count,
                             sicc_pipeline_sync_t* out);
int sicc_pipeline_find_frame_device(unsigned long long handle, int slot, const float* d_samples_iq,
                                    int sample_count, sicc_pipeline_sync_t* out);

/* Pipeline::demod_frame_to_device_llr / demod_frame_device_to_device_llr / demod_frame_to_host_llr */
int sicc_pipeline_demod_frame_to_device_llr(unsigned long long handle, int slot, const float* frame_iq,
                                            sicc_pipeline_sync_t* sync_out, const float** d_llr);
int sicc_pipeline_demod_frame_device_to_device_llr(unsigned long long handle, int slot, const float* d_frame_iq,
                                                   const sicc_pipeline_sync_t* sync, const float** d_llr);
int sicc_pipeline_demod_frame_to_host_llr(unsigned long long handle, int slot, const float* frame_iq,
                                          float* llr, sicc_pipeline_sync_t* sync_out);

/* Batched form of submit_tx_frame_device_async for device-resident test IQ:
 * d_info_bits [frames][blocks_per_frame * code_k] bytes in device memory ->
 * d_out_iq [frames][frame_samples] complex; frame f uses noise_seed + f.
 * Asynchronous on hip_stream (null: the pipeline's TX stream).  Calls on one
 * handle share its coded-bit staging buffer and must not overlap. */
int sicc_pipeline_tx_batch_device(unsigned long long handle, const unsigned char* d_info_bits, int frames,
                                  float frequency_noise_var, unsigned long long noise_seed, float* d_out_iq,
                                  void* hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* SICC_PIPELINE_C_API_H */
