/*
 * frame_oracle.c -- CPU restatement of the reference's per-frame FEC
 * aggregation.  TEST INFRASTRUCTURE ONLY (checker for tests/ and bench.py's
 * CPU leg; the product never links it).
 *
 * Source restated (path relative to the reference repository root):
 *   LINK = source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp
 *     decode_frame, SNR gate         LINK:5482-5501
 *     decode_frame, CPU FEC branch   LINK:5529-5553 (blocks = coded bits / n,
 *                                    parity_ok = AND of the blocks' parity,
 *                                    iterations summed)
 *     decode_frame, result           LINK:5555-5591 (lround mean iterations,
 *                                    frame_id = bits_to_u32_msb, bit and
 *                                    pre-FEC error counts, frame_ok)
 *     bits_to_u32_msb                LINK:2570-2577
 *   finish_decode_from_llr (LINK:6615-6705) aggregates the same way.
 * The reference payload is an input here (the reference regenerates it from
 * the frame id, LINK:4048-4056).
 */
#include "sicc_oracle.h"

#include <math.h>
#include <stddef.h>

void oracle_frame_finish(int blocks, int n, int k, const uint8_t* block_parity, const int* block_iters,
                         const uint8_t* info, const uint8_t* ref_info, const uint8_t* ref_coded, const float* llr,
                         int gated, int* parity_ok, int* iterations, uint32_t* frame_id, int* bit_errors,
                         int* pre_fec_bit_errors, int* frame_ok)
{
    if (gated) {
        /* LINK:5482-5501: not decoded */
        *parity_ok = 0;
        *iterations = 0;
        *frame_id = 0;
        *bit_errors = -1;
        *pre_fec_bit_errors = -1;
        *frame_ok = 0;
        return;
    }
    int ok = 1;
    long iter_sum = 0;
    for (int b = 0; b < blocks; ++b) {
        ok = ok && block_parity[b];
        iter_sum += block_iters[b];
    }
    *parity_ok = ok;
    *iterations = blocks > 0 ? (int)lround((double)iter_sum / (double)blocks) : 0;
    const size_t ninfo = (size_t)blocks * (size_t)k;
    if (ref_info) {
        uint32_t v = 0;
        for (int i = 0; i < 32; ++i) v = (v << 1) | ((size_t)i < ninfo ? (uint32_t)(info[i] & 1u) : 0u);
        *frame_id = v;
        int be = 0;
        for (size_t i = 0; i < ninfo; ++i) be += info[i] != ref_info[i];
        *bit_errors = be;
        int pe = 0;
        const size_t ncoded = (size_t)blocks * (size_t)n;
        for (size_t i = 0; i < ncoded; ++i) pe += (llr[i] < 0.0f ? 1 : 0) != ref_coded[i];
        *pre_fec_bit_errors = pe;
        *frame_ok = ok && be == 0;
    } else {
        *frame_id = 0;
        *bit_errors = -1;
        *pre_fec_bit_errors = -1;
        *frame_ok = 0;
    }
}
