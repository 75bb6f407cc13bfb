/*
 * sicc_frame_c_api.h -- frame-level receive and decode in libsicc_amd.so:
 * IQ samples (or LLRs) in, per-frame information bits and the reference's
 * FrameDecodeResult out, with the whole chain on the GPU.
 *
 * Reference contract (paths relative to the reference root,
 * LINK = source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp):
 *   FrameDecodeResult        LINK:273-286
 *   decode_frame             LINK:5331-5603 (demodulate one synchronized frame,
 *                            decode its blocks, aggregate, compare with the
 *                            reference payload)
 *   finish_decode_from_llr   LINK:6615-6705 (the same from frame LLRs)
 * A frame carries blocks = coded bits / n codewords.  Aggregation follows the
 * reference's CPU FEC branch (LINK:5543-5553): parity_ok holds when every
 * block's decoded word satisfies its parity checks, iterations is the rounded
 * mean (std::lround) of the blocks' BP iteration counts.  With a reference
 * payload (the reference's reference_test mode): frame_id = the first 32
 * decoded information bits MSB first (bits_to_u32_msb, LINK:2570), bit_errors
 * = decoded vs reference information bits, pre_fec_bit_errors = hard channel
 * decisions (LLR < 0) vs the re-encoded reference payload, frame_ok =
 * parity_ok && bit_errors == 0.  Without one: frame_id 0, bit_errors and
 * pre_fec_bit_errors -1, frame_ok 0 (LINK:5584-5588).
 *
 * Differences from the reference, by design:
 *  - the reference payload is an input (the reference regenerates it from
 *    the decoded frame id with std::mt19937, LINK:4048-4056, whose output is
 *    not portable across standard libraries);
 *  - with a BP-OSD decoder a block BP rejects is decoded by OSD, whose output
 *    always satisfies the parity checks, so it counts as parity-satisfied;
 *    `iterations` stays the mean BP iteration count and osd_blocks counts the
 *    OSD blocks (the reference's CUDA BP-OSD branch, LINK:5538-5541, leaves
 *    parity_ok true and reports the OSD count as iterations);
 *  - the noise variance is the receiver's configured one (decode_frame's
 *    known_frequency_noise_var branch, LINK:5456-5461);
 *  - a frame that sync did not find, or whose samples run past its row, is
 *    reported with detected = 0 and zero information bits (the reference's
 *    receive loop never decodes such a frame, LINK:5618-5631).
 *
 * Status codes and cuda_osd_get_last_error() as in matlab_decoder_c_api.h.
 */
#ifndef SICC_FRAME_C_API_H
#define SICC_FRAME_C_API_H

#include "sicc_ofdm_c_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* FrameDecodeResult (LINK:273-286) as a fixed-size record (48 bytes of
 * integers, then four doubles); the information bits go to a separate
 * [frames][blocks * k] byte array. */
typedef struct sicc_frame_result_t {
    int detected;            /* frame found by sync and wholly inside its row (1 on the LLR path) */
    int parity_ok;           /* every block's decoded word satisfies H c = 0 */
    int frame_ok;            /* reference payload given: parity_ok && bit_errors == 0; else 0 */
    unsigned int frame_id;   /* reference payload given: first 32 decoded info bits, MSB first; else 0 */
    int bit_errors;          /* decoded vs reference information bits; -1 without a reference */
    int pre_fec_bit_errors;  /* hard LLR decisions vs the re-encoded reference payload; -1 without */
    int iterations;          /* lround(mean BP iterations over the frame's blocks) */
    int blocks;              /* codewords in the frame */
    int bp_ok_blocks;        /* blocks BP accepted */
    int osd_blocks;          /* blocks decoded by OSD (BP-OSD decoders) */
    int snr_gated;           /* 1 when snr_db < the configured gate: not decoded (LINK:5482-5501) */
    int reserved;
    double peak;             /* refined sync metric (IQ path), 0 on the LLR path */
    double cfo_hz;           /* CFO estimate (IQ path), 0 on the LLR path */
    double snr_db;           /* 10 log10(mean |h|^2 / noise variance) over data REs (LINK:5456-5463) */
    double evm_rms;          /* sqrt(sum |eq - hard|^2 / sum |hard|^2), NaN on the LLR path (LINK:5464-5466) */
} sicc_frame_result_t;

typedef struct sicc_frame_config_t {
    int blocks_per_frame;    /* codewords per frame; the IQ path takes the receiver's */
    double snr_gate_db;      /* rx_snr_gate_db (LINK:5482): IQ-path frames whose estimated SNR
                                is below it are not decoded; <= -1e9 disables the gate */
} sicc_frame_config_t;

/* A frame decoder binds an OFDM receiver (sicc_ofdm_create; 0 for an
 * LLR-only decoder) to an LDPC decoder handle (cuda_osd stream BP-OSD decoder
 * or BP CSR decoder) and owns the per-batch device scratch.  The receiver's
 * code_n must equal the decoder's n.  It keeps both handles' objects alive.
 * Calls on one frame decoder must be issued on one stream (or not overlap). */
int sicc_frame_decoder_create(unsigned long long ofdm_receiver, unsigned long long ldpc_decoder,
                              const sicc_frame_config_t* config, unsigned long long* out_handle);
int sicc_frame_decoder_destroy(unsigned long long handle);
/* Size the device scratch for batches of up to `frames` frames (entry points
 * grow it on demand, which allocates). */
int sicc_frame_decoder_reserve(unsigned long long handle, int frames);
int sicc_frame_decoder_info(unsigned long long handle, int* blocks_per_frame, int* code_n, int* code_k,
                            int* info_bits_per_frame);

/* Device batches, asynchronous on hip_stream (null: the decoder's own stream).
 * Row f of d_rx_iq ([frames][span] complex64) holds frame f, searched from
 * its first sample.  d_ref_info [frames][blocks * k] (null: no reference test).
 * Outputs: d_info_bits [frames][blocks * k], d_results [frames]; optional
 * (null = kept in scratch): d_sync [frames], d_llr [frames][blocks * n],
 * d_block_ok [frames * blocks] (1 = the block's output satisfies its checks),
 * d_block_iterations [frames * blocks]. */
int sicc_frame_receive_decode_device(unsigned long long handle, const float* d_rx_iq, int frames, int span,
                                     const unsigned char* d_ref_info, unsigned char* d_info_bits,
                                     sicc_frame_result_t* d_results, sicc_ofdm_sync_t* d_sync, float* d_llr,
                                     unsigned char* d_block_ok, int* d_block_iterations, void* hip_stream);
/* decode_frame proper: the caller's sync results (start, cfo_hz_d, metric)
 * select and derotate each frame, as decode_frame takes cfo_hz and peak. */
int sicc_frame_demod_decode_device(unsigned long long handle, const float* d_rx_iq, int frames, int span,
                                   const sicc_ofdm_sync_t* d_sync, const unsigned char* d_ref_info,
                                   unsigned char* d_info_bits, sicc_frame_result_t* d_results, float* d_llr,
                                   unsigned char* d_block_ok, int* d_block_iterations, void* hip_stream);
/* finish_decode_from_llr: d_llr [frames][blocks * n]. */
int sicc_frame_decode_llr_device(unsigned long long handle, const float* d_llr, int frames,
                                 const unsigned char* d_ref_info, unsigned char* d_info_bits,
                                 sicc_frame_result_t* d_results, unsigned char* d_block_ok,
                                 int* d_block_iterations, void* hip_stream);

/* Synchronous host-buffer variants (ref_info may be null). */
int sicc_frame_receive_decode_host(unsigned long long handle, const float* rx_iq, int frames, int span,
                                   const unsigned char* ref_info, unsigned char* info_bits,
                                   sicc_frame_result_t* results);
int sicc_frame_decode_llr_host(unsigned long long handle, const float* llr, int frames,
                               const unsigned char* ref_info, unsigned char* info_bits,
                               sicc_frame_result_t* results);

#ifdef __cplusplus
}
#endif

#endif /* SICC_FRAME_C_API_H */
