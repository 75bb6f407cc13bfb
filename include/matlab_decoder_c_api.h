/*
 * matlab_decoder_c_api.h -- BP-OSD decoder C ABI of the MI355X library
 * (libsicc_amd.so).  Drop-in for the reference's cuda_osd decoder interface:
 * same names, argument meaning, struct layouts and status codes, so a caller
 * built against the reference header links against this library unchanged.
 *
 * Reference interface replaced (paths relative to the reference root):
 *   OSDH = source/cpp_cuda/portable/usrp_video_link_portable_20260513/third_party/
 *          cuda_osd/include/matlab_decoder_c_api.h
 *   LINK = source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp (callers, and the
 *          three entry points the canonical tree uses beyond OSDH)
 *
 * Conventions (OSDH, LINK:3162-3786):
 *   - handles are opaque non-zero 64-bit values; 0 is never valid;
 *   - every entry point returns CUDA_OSD_STATUS_* and leaves a message for
 *     cuda_osd_get_last_error() on failure; outputs are caller-allocated;
 *   - LLRs are row-major [codewords][n] float32, positive means bit 0;
 *   - bits are one byte each; information bits are the first k code bits
 *     (systematic-front layout, LINK:3096-3118);
 *   - "device" pointers are HIP device pointers on the current device.
 * Decoder semantics: normalized min-sum BP (LINK:3013-3094) with the MATLAB
 * best-syndrome APP hand-off, then ordered-statistics decoding
 * (source/matlab/uhd_cpp/matlab/cpu_bp_osd_decode.m:187-305) on BP failure.
 */
#ifndef SICC_MATLAB_DECODER_C_API_H
#define SICC_MATLAB_DECODER_C_API_H

#ifdef __cplusplus
extern "C" {
#endif

/* OSDH:17-32 */
typedef struct cuda_osd_stream_config_t {
    int n;
    int k;
    int m;
    int max_batch_size;
    int min_batch_size;
    unsigned long long max_latency_us;
    double arrival_ewma_alpha;
    unsigned long long default_interarrival_us;
    int bp_max_iterations;
    float bp_normalization;
    float bp_offset;
    float bp_damping;
    float bp_min_abs_llr_accept;
    int osd_threads;
} cuda_osd_stream_config_t;

/* OSDH:34-39 */
typedef struct cuda_osd_osd_config_t {
    int n;
    int k;
    int max_batch_size;
    int osd_threads;
} cuda_osd_osd_config_t;

/* OSDH:41-51, plus the trailing `schedule` field the canonical tree sets
 * (LINK:3658): 0 flooding, 1 node-parallel, 2 colored-layered, 3 flat-layered,
 * negative = automatic (layered when n >= 1000). */
typedef struct cuda_osd_bp_csr_config_t {
    int n;
    int m;
    int max_batch_size;
    int max_iterations;
    float normalization;
    float offset;
    float damping;
    float min_abs_llr_accept;
    int nnz;
    int schedule;
} cuda_osd_bp_csr_config_t;

/* OSDH:53-65 */
typedef struct cuda_osd_stream_stats_t {
    unsigned long long submitted_codewords;
    unsigned long long decoded_codewords;
    unsigned long long bp_success_codewords;
    unsigned long long osd_fallback_codewords;
    unsigned long long decoded_batches;
    unsigned long long current_pending_codewords;
    unsigned long long max_pending_codewords;
    unsigned long long last_submit_timestamp_us;
    double ewma_interarrival_us;
    double average_batch_size;
    double total_decode_ms;
} cuda_osd_stream_stats_t;

/* OSDH:67-71 */
enum {
    CUDA_OSD_STATUS_OK = 0,
    CUDA_OSD_STATUS_INVALID_ARGUMENT = 1,
    CUDA_OSD_STATUS_INTERNAL_ERROR = 2
};

/* OSDH:73 -- stream BP-OSD decoder on the bundled CCSDS (128,64) code. */
int cuda_osd_create_stream_decoder(const cuda_osd_stream_config_t* config,
                                   unsigned long long* out_handle);
/* LINK:3251-3256 (declared nowhere in the reference tree) -- stream decoder on
 * a caller code: dense H [m][n] bytes and the systematic parity rows
 * [k][parity_words_per_row] (bit p of row i = parity bit p of encode(e_i)). */
int cuda_osd_create_stream_decoder_with_matrices(const cuda_osd_stream_config_t* config,
                                                 const unsigned char* dense_h,
                                                 const unsigned long long* systematic_parity,
                                                 int parity_words_per_row,
                                                 unsigned long long* out_handle);
/* OSDH:74 */
int cuda_osd_destroy_stream_decoder(unsigned long long handle);
/* OSDH:75-76 -- OSD-only decoder on the bundled CCSDS (128,64) code. */
int cuda_osd_create_osd_decoder(const cuda_osd_osd_config_t* config,
                                unsigned long long* out_handle);
/* OSDH:77 */
int cuda_osd_destroy_osd_decoder(unsigned long long handle);
/* OSDH:78-81 -- BP-only decoder on a CSR parity-check matrix. */
int cuda_osd_create_bp_csr_decoder(const cuda_osd_bp_csr_config_t* config,
                                   const int* row_ptr,
                                   const int* col_ind,
                                   unsigned long long* out_handle);
/* OSDH:82 */
int cuda_osd_destroy_bp_csr_decoder(unsigned long long handle);
/* OSDH:83-88 -- host LLRs in, host code bits and BP success out. */
int cuda_osd_decode_bp_csr_batch(unsigned long long handle,
                                 const float* llrs,
                                 int codewords,
                                 unsigned char* out_code_bits,
                                 unsigned char* out_success,
                                 double* out_total_decode_ms);
/* LINK:3753-3760 (declared nowhere in the reference tree) -- device LLRs in. */
int cuda_osd_decode_bp_csr_device_batch(unsigned long long handle,
                                        const float* device_llrs,
                                        int codewords,
                                        unsigned char* out_code_bits,
                                        unsigned char* out_success,
                                        double* out_total_decode_ms);
/* OSDH:89-95 */
int cuda_osd_decode_osd_batch(unsigned long long handle,
                              const float* llrs,
                              int codewords,
                              unsigned char* out_info_bits,
                              float* out_distances,
                              double* out_total_decode_ms,
                              double* out_throughput_mbps);
/* OSDH:96-101 */
int cuda_osd_submit_codewords(unsigned long long handle,
                              const float* llrs,
                              int codewords,
                              const unsigned long long* timestamps_us,
                              const unsigned long long* frame_ids,
                              const unsigned int* block_indices);
/* OSDH:102-107 */
int cuda_osd_submit_device_codewords(unsigned long long handle,
                                     const float* device_llrs,
                                     int codewords,
                                     const unsigned long long* timestamps_us,
                                     const unsigned long long* frame_ids,
                                     const unsigned int* block_indices);
/* OSDH:108-115 */
int cuda_osd_poll_ready(unsigned long long handle,
                        int max_codewords,
                        unsigned char* out_info_bits,
                        unsigned long long* out_timestamps_us,
                        unsigned long long* out_frame_ids,
                        unsigned int* out_block_indices,
                        unsigned char* out_used_osd,
                        int* out_codewords);
/* OSDH:116 */
int cuda_osd_flush(unsigned long long handle);
/* OSDH:117-118 */
int cuda_osd_get_stats(unsigned long long handle, cuda_osd_stream_stats_t* out_stats);
/* OSDH:119 */
unsigned long long cuda_osd_get_decoded_codewords(unsigned long long handle);
/* OSDH:120 */
double cuda_osd_get_total_decode_ms(unsigned long long handle);
/* OSDH:121-132 -- Schmidl-Cox frame search on double IQ (GPU). */
int cuda_osd_find_frame(const double* signal_real,
                        const double* signal_imag,
                        int signal_length,
                        const double* preamble_real,
                        const double* preamble_imag,
                        int preamble_length,
                        int pre_half_len,
                        double sync_peak_thresh,
                        double fs,
                        int* out_start_idx_1based,
                        double* out_peak,
                        double* out_cfo_hz);
/* OSDH:134 */
const char* cuda_osd_get_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SICC_MATLAB_DECODER_C_API_H */
