// code_model.hpp -- LDPC code model on the host: alist/CSR/dense-H loaders, the
// GF(2) RREF systematic encoder, generator tables for OSD and a seeded
// quasi-cyclic LDPC constructor.
//
// Reference behaviour (paths relative to the reference repository root,
// LINK = source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp):
//   read_alist                   LINK:2722-2770
//   build_ldpc (RREF, front info) LINK:2811-2945
//   ldpc_encode                  LINK:2947-2997
//   ldpc_has_front_systematic_info LINK:3096-3118
//   build_dense_h_for_cuda_osd   LINK:3120-3131
//   build_systematic_parity_for_cuda_osd LINK:3133-3160
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace sicc {

struct LdpcCode {
    int n = 0;
    int m = 0;  // parity checks (rows of H)
    int k = 0;  // information bits = n - rank(H)
    int words = 0;  // 64-bit words per packed row of length n
    bool accumulator = false;  // DVB-S2 style dual-diagonal parity (LINK:2834-2866)
    std::vector<int> row_ptr;  // CSR of H, rows keep the source's column order
    std::vector<int> col_ind;
    std::vector<int> info_cols;
    std::vector<int> pivot_cols;
    std::vector<uint64_t> rref;  // m x words, reduced row echelon form of H

    int edges() const { return row_ptr.empty() ? 0 : row_ptr.back(); }
    bool front_systematic() const;
    void encode(const uint8_t* info, uint8_t* cw) const;
    // Dense generator, k x n bytes, row i = encode(e_i).
    std::vector<uint8_t> dense_generator() const;
    // Dense parity-check matrix, m x n bytes.
    std::vector<uint8_t> dense_h() const;
};

// Build from a CSR parity-check matrix (row order of columns kept).
LdpcCode build_code(int n, int m, const std::vector<int>& row_ptr, const std::vector<int>& col_ind,
                    bool front_info);
LdpcCode load_alist(const std::string& path, bool front_info);
LdpcCode parse_alist_text(const std::string& text, bool front_info);
// Dense m x n H, columns taken in ascending order per row.
LdpcCode code_from_dense_h(int m, int n, const uint8_t* h, bool front_info);

// Seeded quasi-cyclic LDPC: 16 block columns of Z x Z circulant permutations,
// mb block rows, column weight `col_weight`; shifts drawn from `seed` with
// 4-cycles avoided, re-drawn deterministically until the parity part is full rank.
LdpcCode make_qc_code(int n, int mb, int col_weight, uint64_t seed);

// Directory holding the bundled code matrices (the library's ../data).
std::string bundled_data_dir();

}  // namespace sicc
