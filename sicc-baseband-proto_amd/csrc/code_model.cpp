// code_model.cpp -- host-side LDPC code construction (see code_model.hpp).
#include "code_model.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace sicc {

namespace {

inline bool bit_at(const uint64_t* row, int c) { return ((row[c >> 6] >> (c & 63)) & 1ull) != 0; }
inline void toggle(uint64_t* row, int c) { row[c >> 6] ^= 1ull << (c & 63); }

// Dual-diagonal parity part with front info: row 0 touches parity column 0,
// row r>0 touches parity columns r-1 and r (LINK:2834-2852).
bool accumulator_shape(int n, int m, const std::vector<int>& rp, const std::vector<int>& ci)
{
    if (n <= m) return false;
    const int k = n - m;
    std::vector<int> pc;
    for (int r = 0; r < m; ++r) {
        pc.clear();
        for (int e = rp[r]; e < rp[r + 1]; ++e)
            if (ci[e] >= k) pc.push_back(ci[e] - k);
        std::sort(pc.begin(), pc.end());
        if (r == 0) {
            if (pc.size() != 1 || pc[0] != 0) return false;
        } else if (pc.size() != 2 || pc[0] != r - 1 || pc[1] != r) {
            return false;
        }
    }
    return true;
}

uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Unbiased draw in [0, bound) by rejection.
uint32_t draw(uint64_t& s, uint32_t bound)
{
    const uint64_t lim = (~0ull) - ((~0ull) % bound);
    uint64_t x;
    do {
        x = splitmix64(s);
    } while (x >= lim);
    return static_cast<uint32_t>(x % bound);
}

}  // namespace

LdpcCode build_code(int n, int m, const std::vector<int>& row_ptr, const std::vector<int>& col_ind,
                    bool front_info)
{
    if (n <= 0 || m <= 0 || static_cast<int>(row_ptr.size()) != m + 1 || row_ptr[0] != 0 ||
        row_ptr[m] != static_cast<int>(col_ind.size()))
        throw std::invalid_argument("invalid CSR parity-check matrix");
    for (int r = 0; r < m; ++r)
        if (row_ptr[r + 1] < row_ptr[r]) throw std::invalid_argument("CSR row_ptr is not monotone");
    for (int c : col_ind)
        if (c < 0 || c >= n) throw std::invalid_argument("CSR column index out of range");

    LdpcCode code;
    code.n = n;
    code.m = m;
    code.words = (n + 63) / 64;
    code.row_ptr = row_ptr;
    code.col_ind = col_ind;

    if (front_info && accumulator_shape(n, m, row_ptr, col_ind)) {
        code.accumulator = true;
        code.k = n - m;
        for (int i = 0; i < code.k; ++i) code.info_cols.push_back(i);
        for (int r = 0; r < m; ++r) code.pivot_cols.push_back(code.k + r);
        return code;
    }

    const int W = code.words;
    std::vector<uint64_t> rows(static_cast<size_t>(m) * W, 0);
    for (int r = 0; r < m; ++r)
        for (int e = row_ptr[r]; e < row_ptr[r + 1]; ++e) toggle(&rows[static_cast<size_t>(r) * W], col_ind[e]);

    std::vector<int> order;
    order.reserve(n);
    if (front_info) {
        if (n <= m) throw std::invalid_argument("front-systematic layout requires n > m");
        for (int c = n - m; c < n; ++c) order.push_back(c);
        for (int c = 0; c < n - m; ++c) order.push_back(c);
    } else {
        for (int c = 0; c < n; ++c) order.push_back(c);
    }
    int rank = 0;
    for (int col : order) {
        if (rank >= m) break;
        int piv = -1;
        for (int r = rank; r < m; ++r)
            if (bit_at(&rows[static_cast<size_t>(r) * W], col)) {
                piv = r;
                break;
            }
        if (piv < 0) continue;
        if (piv != rank)
            std::swap_ranges(rows.begin() + static_cast<long>(piv) * W, rows.begin() + static_cast<long>(piv + 1) * W,
                             rows.begin() + static_cast<long>(rank) * W);
        const uint64_t* prow = &rows[static_cast<size_t>(rank) * W];
        for (int r = 0; r < m; ++r) {
            uint64_t* row = &rows[static_cast<size_t>(r) * W];
            if (r != rank && bit_at(row, col))
                for (int w = 0; w < W; ++w) row[w] ^= prow[w];
        }
        code.pivot_cols.push_back(col);
        ++rank;
    }
    if (rank != m) throw std::runtime_error("LDPC parity-check matrix is rank deficient for this encoder");
    std::vector<char> isp(n, 0);
    for (int c : code.pivot_cols) isp[c] = 1;
    for (int c = 0; c < n; ++c)
        if (!isp[c]) code.info_cols.push_back(c);
    code.k = static_cast<int>(code.info_cols.size());
    code.rref = std::move(rows);
    return code;
}

bool LdpcCode::front_systematic() const
{
    if (k <= 0 || m <= 0 || k + m != n) return false;
    if (static_cast<int>(info_cols.size()) != k || static_cast<int>(pivot_cols.size()) != m) return false;
    for (int i = 0; i < k; ++i)
        if (info_cols[i] != i) return false;
    std::vector<char> seen(m, 0);
    for (int c : pivot_cols) {
        if (c < k || c >= n) return false;
        seen[c - k] = 1;
    }
    return std::all_of(seen.begin(), seen.end(), [](char v) { return v != 0; });
}

void LdpcCode::encode(const uint8_t* info, uint8_t* cw) const
{
    std::fill(cw, cw + n, 0);
    if (accumulator) {
        for (int i = 0; i < k; ++i) cw[i] = info[i] & 1u;
        uint8_t prev = 0;
        for (int r = 0; r < m; ++r) {
            uint8_t s = 0;
            for (int e = row_ptr[r]; e < row_ptr[r + 1]; ++e)
                if (col_ind[e] < k) s ^= cw[col_ind[e]];
            const uint8_t p = static_cast<uint8_t>(s ^ (r == 0 ? 0u : prev));
            cw[k + r] = p;
            prev = p;
        }
        return;
    }
    std::vector<uint64_t> w(words, 0);
    for (int i = 0; i < k; ++i)
        if (info[i] & 1u) {
            cw[info_cols[i]] = 1;
            toggle(w.data(), info_cols[i]);
        }
    for (int r = 0; r < m; ++r) {
        int par = 0;
        const uint64_t* row = &rref[static_cast<size_t>(r) * words];
        for (int x = 0; x < words; ++x) par ^= __builtin_popcountll(row[x] & w[x]) & 1;
        if (par) {
            cw[pivot_cols[r]] = 1;
            toggle(w.data(), pivot_cols[r]);
        }
    }
}

std::vector<uint8_t> LdpcCode::dense_generator() const
{
    std::vector<uint8_t> g(static_cast<size_t>(k) * n, 0);
    std::vector<uint8_t> unit(k, 0);
    for (int i = 0; i < k; ++i) {
        std::fill(unit.begin(), unit.end(), 0);
        unit[i] = 1;
        encode(unit.data(), &g[static_cast<size_t>(i) * n]);
    }
    return g;
}

std::vector<uint8_t> LdpcCode::dense_h() const
{
    std::vector<uint8_t> h(static_cast<size_t>(m) * n, 0);
    for (int r = 0; r < m; ++r)
        for (int e = row_ptr[r]; e < row_ptr[r + 1]; ++e) h[static_cast<size_t>(r) * n + col_ind[e]] = 1;
    return h;
}

LdpcCode parse_alist_text(const std::string& text, bool front_info)
{
    std::istringstream in(text);
    int n = 0, m = 0, maxc = 0, maxr = 0;
    in >> n >> m >> maxc >> maxr;
    if (!in || n <= 0 || m <= 0 || maxc < 0 || maxr <= 0) throw std::invalid_argument("invalid alist header");
    std::vector<int> cw(n), rw(m);
    for (auto& v : cw) in >> v;
    for (auto& v : rw) in >> v;
    for (int c = 0; c < n; ++c)
        for (int j = 0; j < maxc; ++j) {
            int r = 0;
            in >> r;
        }
    std::vector<int> rp(1, 0), ci;
    for (int r = 0; r < m; ++r) {
        for (int j = 0; j < maxr; ++j) {
            int c = 0;
            in >> c;
            if (j < rw[r] && c > 0) ci.push_back(c - 1);
        }
        rp.push_back(static_cast<int>(ci.size()));
    }
    if (in.fail()) throw std::invalid_argument("truncated alist");
    return build_code(n, m, rp, ci, front_info);
}

LdpcCode load_alist(const std::string& path, bool front_info)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open alist: " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_alist_text(ss.str(), front_info);
}

LdpcCode code_from_dense_h(int m, int n, const uint8_t* h, bool front_info)
{
    if (!h || m <= 0 || n <= 0) throw std::invalid_argument("dense H is empty");
    std::vector<int> rp(1, 0), ci;
    for (int r = 0; r < m; ++r) {
        for (int c = 0; c < n; ++c)
            if (h[static_cast<size_t>(r) * n + c] & 1u) ci.push_back(c);
        rp.push_back(static_cast<int>(ci.size()));
    }
    return build_code(n, m, rp, ci, front_info);
}

LdpcCode make_qc_code(int n, int mb, int col_weight, uint64_t seed)
{
    constexpr int nb = 16;
    if (n <= 0 || n % nb != 0 || mb <= 0 || mb >= nb || col_weight <= 0 || col_weight > mb)
        throw std::invalid_argument("invalid QC-LDPC shape");
    const int Z = n / nb;
    for (int attempt = 0; attempt < 64; ++attempt) {
        uint64_t s = seed * 0x100000001B3ull + static_cast<uint64_t>(attempt) + 1;
        std::vector<int> shift(static_cast<size_t>(mb) * nb, -1);  // -1 = zero block
        for (int j = 0; j < nb; ++j) {
            // choose col_weight distinct block rows (partial Fisher-Yates)
            std::vector<int> rows(mb);
            for (int i = 0; i < mb; ++i) rows[i] = i;
            for (int i = 0; i < col_weight; ++i) {
                const int pick = i + static_cast<int>(draw(s, static_cast<uint32_t>(mb - i)));
                std::swap(rows[i], rows[pick]);
            }
            for (int i = 0; i < col_weight; ++i) {
                const int bi = rows[i];
                int best = static_cast<int>(draw(s, static_cast<uint32_t>(Z)));
                for (int tries = 0; tries < 64; ++tries) {
                    // 4-cycle test against every earlier block column
                    bool cyc = false;
                    for (int j2 = 0; j2 < j && !cyc; ++j2) {
                        if (shift[bi * nb + j2] < 0) continue;
                        for (int bi2 = 0; bi2 < mb && !cyc; ++bi2) {
                            if (bi2 == bi || shift[bi2 * nb + j] < 0 || shift[bi2 * nb + j2] < 0) continue;
                            const int d = shift[bi * nb + j2] - shift[bi2 * nb + j2] + shift[bi2 * nb + j] - best;
                            if (((d % Z) + Z) % Z == 0) cyc = true;
                        }
                    }
                    if (!cyc) break;
                    best = static_cast<int>(draw(s, static_cast<uint32_t>(Z)));
                }
                shift[bi * nb + j] = best;
            }
        }
        const int m = mb * Z;
        std::vector<int> rp(1, 0), ci;
        for (int bi = 0; bi < mb; ++bi)
            for (int z = 0; z < Z; ++z) {
                for (int j = 0; j < nb; ++j) {
                    const int sh = shift[bi * nb + j];
                    if (sh >= 0) ci.push_back(j * Z + (z + sh) % Z);
                }
                rp.push_back(static_cast<int>(ci.size()));
            }
        try {
            LdpcCode c = build_code(n, m, rp, ci, true);
            if (c.front_systematic()) return c;
        } catch (const std::runtime_error&) {
            // rank deficient draw: try the next attempt
        }
    }
    throw std::runtime_error("could not draw a full-rank QC-LDPC code");
}

std::string bundled_data_dir()
{
    if (const char* env = std::getenv("SICC_DATA_DIR")) return env;
    Dl_info info{};
    if (dladdr(reinterpret_cast<void*>(&bundled_data_dir), &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        const auto slash = p.find_last_of('/');
        std::string dir = slash == std::string::npos ? "." : p.substr(0, slash);
        return dir + "/../data";
    }
    return "data";
}

}  // namespace sicc
