/*
 * sicc_decode_ext.h -- entry points of libsicc_amd.so beyond the reference's
 * cuda_osd ABI (matlab_decoder_c_api.h).  Each one states the reference
 * behaviour it exposes, or that it has no counterpart in the reference.
 * Status codes are CUDA_OSD_STATUS_*; messages go to cuda_osd_get_last_error().
 */
#ifndef SICC_DECODE_EXT_H
#define SICC_DECODE_EXT_H

#include "matlab_decoder_c_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Stream BP-OSD decoder on an alist file (LINK:2722-2945 read_alist + build_ldpc
 * with systematic-front info, as the reference runs with --systematic-front-info). */
int sicc_create_stream_decoder_alist(const cuda_osd_stream_config_t* config, const char* alist_path,
                                     unsigned long long* out_handle);

/* Stream BP-OSD decoder on a caller CSR parity-check matrix (rows in the caller's
 * column order), systematic-front layout built by GF(2) RREF (LINK:2811-2945). */
int sicc_create_stream_decoder_csr(const cuda_osd_stream_config_t* config, int n, int m,
                                   const int* row_ptr, const int* col_ind,
                                   unsigned long long* out_handle);

/* Seeded quasi-cyclic LDPC code: 16 block columns of (n/16)x(n/16) circulant
 * permutations, mb block rows, column weight col_weight.  No reference
 * counterpart (the reference only loads alists); it builds the mixed-code
 * workloads.  The CSR is written to caller buffers when they are non-null:
 * row_ptr[mb*n/16 + 1], col_ind[col_weight*n]. */
int sicc_qc_code_csr(int n, int mb, int col_weight, unsigned long long seed, int* out_m,
                     int* out_nnz, int* row_ptr, int* col_ind);

/* OSD parameters of a stream or OSD-only decoder: order 0..3, flip pool size,
 * test cap (the MATLAB reference options osdOrder/osdMaxFlipPositions/osdMaxTests,
 * source/matlab/uhd_cpp/matlab/cpu_bp_osd_decode.m:83-99).  order < 0 disables
 * OSD.  Defaults: stream decoders order 2 / 32 / 4096 (the MATLAB defaults);
 * OSD-only decoders order 2 / k / 4096. */
int sicc_set_osd_params(unsigned long long handle, int order, int max_flip, int max_tests);

/* Code dimensions of a decoder handle. */
int sicc_get_code_info(unsigned long long handle, int* n, int* k, int* m, int* nnz);
/* CSR of the decoder's parity-check matrix (row_ptr[m+1], col_ind[nnz]). */
int sicc_get_code_csr(unsigned long long handle, int* row_ptr, int* col_ind);
/* Dense generator [k][n] of the decoder's code (row i = encode(e_i)). */
int sicc_get_generator(unsigned long long handle, unsigned char* out_gen);
/* Systematic encoder (LINK:2947-2997): info [count][k] -> codewords [count][n]. */
int sicc_encode(unsigned long long handle, const unsigned char* info, int count, unsigned char* out_cw);

/* Device-resident decode of `codewords` LLR rows already in HBM, on
 * `hip_stream` (null = the decoder's own stream), asynchronous: BP
 * (LINK:3013-3094 semantics) then OSD on frames BP does not accept.  Every
 * output is a device pointer and may be null.  The stream/poll ABI of the
 * reference hands results back through host memory; this entry point keeps them
 * on the device for callers that chain further GPU work.  No counterpart in the
 * reference. */
int sicc_decode_device(unsigned long long handle, const float* device_llrs, int codewords,
                       unsigned char* d_info_bits, unsigned char* d_code_bits, float* d_app_llrs,
                       unsigned char* d_bp_success, unsigned char* d_used_osd, int* d_iterations,
                       float* d_distances, void* hip_stream);

/* Allocation-free variant of sicc_decode_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, outputs and semantics; the decode's scratch (BP -> OSD
 * hand-off and counters) lives in a fixed device array of
 * SICC_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, instead of
 * buffers the library allocates.  Limits: decoders that run OSD take at most
 * (SICC_NOALLOC_SCRATCH_BYTES - 512) / (4 n + 4) codewords per call (16 256 for
 * n = 128; sicc_decode_noalloc_capacity reports it), larger batches fail with
 * CUDA_OSD_STATUS_INVALID_ARGUMENT; only the short-code flooding kernel runs
 * here (a layered schedule or a long code fails the same way).  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU race. */
#define SICC_NOALLOC_SCRATCH_BYTES (8u << 20)
int sicc_decode_device_noalloc(unsigned long long handle, const float* device_llrs, int codewords,
                               unsigned char* d_info_bits, unsigned char* d_code_bits, float* d_app_llrs,
                               unsigned char* d_bp_success, unsigned char* d_used_osd, int* d_iterations,
                               float* d_distances, void* hip_stream);
int sicc_decode_noalloc_capacity(unsigned long long handle, int* max_codewords,
                                 unsigned long long* scratch_bytes);

/* Kernel timing: with timing on, HIP events recorded on the launch stream
 * bracket every BP and OSD kernel launch of the handle; the getter waits for
 * them and returns summed durations (ms) and launch counts since the last
 * enable.  No counterpart in the reference (its events are created with
 * timing disabled, gpu_ofdm_pipeline.cu:1434-1437). */
int sicc_set_kernel_timing(unsigned long long handle, int enable);
int sicc_get_kernel_timing(unsigned long long handle, double* bp_ms, double* osd_ms, int* bp_launches,
                           int* osd_launches);

/* BP launch shape of a decoder: wavefronts per codeword (1, 2, 4, 8, 16; 0 =
 * default), codewords per workgroup for single-wave teams (1, 2, 4) and the
 * persistent grid size in wavefronts (0 = 32 per CU), and work stealing
 * (steal_chunk > 0: teams take that many codewords at a time from a shared
 * cursor; 0: static striding).  The autotuner times
 * every supported shape on `codewords` device LLR rows and keeps the fastest;
 * it reports the choice and its average BP kernel time.  No counterpart in the
 * reference (its decoder is a prebuilt library with fixed launch shapes). */
int sicc_set_bp_tune(unsigned long long handle, int team, int teams_per_block, int max_waves,
                     int steal_chunk);
int sicc_get_bp_tune(unsigned long long handle, int* team, int* teams_per_block, int* max_waves,
                     int* steal_chunk);
int sicc_autotune(unsigned long long handle, const float* device_llrs, int codewords, int* team,
                  int* teams_per_block, int* max_waves, int* steal_chunk, double* best_ms);

/* Mixed-code batches.  A multi-decoder groups existing decoder handles (one per
 * code; it keeps them alive).  One call decodes any number of groups, each a
 * contiguous device batch of one member's code, with every group's persistent
 * BP kernel (and OSD for stream members) launched concurrently on forked
 * streams; with use_graph the whole fork/join sequence is captured once into a
 * hipGraph and replayed while the group descriptors stay the same.  Outputs may
 * be null.  No counterpart in the reference (it decodes one code per handle). */
typedef struct sicc_group_t {
    int member;                   /* index into the decoders given at creation */
    int count;                    /* codewords in this group */
    const float* d_llr;           /* [count][n] */
    unsigned char* d_info;        /* [count][k] */
    unsigned char* d_bp_ok;       /* [count] */
    int* d_iterations;            /* [count] */
    unsigned char* d_used_osd;    /* [count] */
} sicc_group_t;
int sicc_multi_create(const unsigned long long* decoders, int count, unsigned long long* out_handle);
int sicc_multi_destroy(unsigned long long handle);
int sicc_multi_decode_device(unsigned long long handle, const sicc_group_t* groups, int n_groups,
                             int use_graph, void* hip_stream);

/* Synchronous host-buffer variant of sicc_decode_device (hard bits, soft LLRs,
 * per-frame success) for CPU callers; outputs may be null. */
int sicc_decode_host(unsigned long long handle, const float* llrs, int codewords,
                     unsigned char* info_bits, unsigned char* code_bits, float* app_llrs,
                     unsigned char* bp_success, unsigned char* used_osd, int* iterations,
                     float* distances);

#ifdef __cplusplus
}
#endif

#endif /* SICC_DECODE_EXT_H */
