// ref_harness.cpp -- C entry points into the reference's own C++ CPU path.
//
// TEST INFRASTRUCTURE ONLY.  build_ref.py compiles this file together with the
// reference sources where they lie (source/cpp_cuda/uhd_cpp/src, read from the
// reference checkout) into oracle/_ref/libref.so.  The reference translation
// unit is included textually because its decode functions have internal
// linkage; its `main` is renamed to an unused static function, so the radio
// and network code that only `main` reaches is never emitted and nothing needs
// the UHD library.  The harness adds no algorithm of its own: each entry point
// forwards to one reference function.
#define main static reference_main_unused
#include SICC_REF_LINK_CPP
#undef main

#include <cstring>

namespace {
LdpcCode g_code;
bool g_have_code = false;
PhyConfig g_phy;
bool g_have_phy = false;
thread_local std::string g_err;
}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// read_alist + build_ldpc (LINK:2722, LINK:2811).
int ref_code_load(const char* alist, int front_info, int* n, int* m, int* k)
{
    try {
        g_code = build_ldpc(read_alist(alist), front_info != 0);
        g_have_code = true;
        *n = g_code.n;
        *m = g_code.m;
        *k = g_code.k;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int ref_code_info_cols(int* info_cols, int* pivot_cols)
{
    if (!g_have_code) return -1;
    std::copy(g_code.info_cols.begin(), g_code.info_cols.end(), info_cols);
    std::copy(g_code.pivot_cols.begin(), g_code.pivot_cols.end(), pivot_cols);
    return 0;
}

// ldpc_encode (LINK:2947).
int ref_code_encode(const uint8_t* info, int count, uint8_t* cw)
{
    if (!g_have_code) return -1;
    for (int i = 0; i < count; ++i) {
        const auto c = ldpc_encode(g_code, info + static_cast<size_t>(i) * g_code.k);
        std::memcpy(cw + static_cast<size_t>(i) * g_code.n, c.data(), c.size());
    }
    return 0;
}

// ldpc_decode_min_sum (LINK:3013), one call per codeword.
int ref_min_sum(const float* llr, int count, int max_iter, double norm, uint8_t* code_bits, uint8_t* info_bits,
                uint8_t* ok, int* iters)
{
    if (!g_have_code) return -1;
    for (int i = 0; i < count; ++i) {
        const DecodeResult r = ldpc_decode_min_sum(g_code, llr + static_cast<size_t>(i) * g_code.n, max_iter, norm);
        if (code_bits) std::memcpy(code_bits + static_cast<size_t>(i) * g_code.n, r.code_bits.data(), g_code.n);
        if (info_bits) std::memcpy(info_bits + static_cast<size_t>(i) * g_code.k, r.info_bits.data(), g_code.k);
        if (ok) ok[i] = r.parity_ok ? 1 : 0;
        if (iters) iters[i] = r.iterations;
    }
    return 0;
}

// build_phy_config (LINK:3971) from the option subset the receive path uses.
int ref_phy_setup(int nfft, int cp, int num_symbols, int active_sc, int pre_half_len, const char* modulation,
                  int pilot_period, double rate, double sync_threshold, double amplitude, int test_seed)
{
    try {
        Options opt;
        opt.nfft = nfft;
        opt.cp = cp;
        opt.num_symbols = num_symbols;
        opt.active_sc = active_sc;
        opt.pre_half_len = pre_half_len;
        opt.modulation = modulation;
        opt.pilot_period = pilot_period;
        opt.rate = rate;
        opt.sync_threshold = sync_threshold;
        opt.amplitude = amplitude;
        opt.test_seed = test_seed;
        g_phy = build_phy_config(opt);
        g_have_phy = true;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int ref_phy_dims(int* frame_len, int* coded_bits, int* n_pilot, int* n_data, int* bits_per_symbol, int* pre_len)
{
    if (!g_have_phy) return -1;
    *frame_len = g_phy.frame_len();
    *coded_bits = g_phy.coded_bits_per_frame();
    *n_pilot = static_cast<int>(g_phy.pilot_symbols.size());
    *n_data = static_cast<int>(g_phy.data_symbols.size());
    *bits_per_symbol = g_phy.bits_per_symbol;
    *pre_len = static_cast<int>(g_phy.preamble.size());
    return 0;
}

// Interleaved re/im float tables of the configured PHY.
int ref_phy_tables(float* preamble, int* used_indices, int* pilot_syms, int* data_syms, float* pilot_grid,
                   float* constellation)
{
    if (!g_have_phy) return -1;
    for (size_t i = 0; i < g_phy.preamble.size(); ++i) {
        preamble[2 * i] = g_phy.preamble[i].real();
        preamble[2 * i + 1] = g_phy.preamble[i].imag();
    }
    std::copy(g_phy.used_indices.begin(), g_phy.used_indices.end(), used_indices);
    std::copy(g_phy.pilot_symbols.begin(), g_phy.pilot_symbols.end(), pilot_syms);
    std::copy(g_phy.data_symbols.begin(), g_phy.data_symbols.end(), data_syms);
    size_t o = 0;
    for (const auto& row : g_phy.pilot_grid)
        for (const auto& x : row) {
            pilot_grid[o++] = x.real();
            pilot_grid[o++] = x.imag();
        }
    for (size_t i = 0; i < g_phy.constellation.size(); ++i) {
        constellation[2 * i] = g_phy.constellation[i].real();
        constellation[2 * i + 1] = g_phy.constellation[i].imag();
    }
    return 0;
}

// build_tx_frame_from_info (LINK:4590).
int ref_build_tx_frame(const uint8_t* info_bits, float* out_iq)
{
    if (!g_have_phy || !g_have_code) return -1;
    try {
        const int blocks = g_phy.coded_bits_per_frame() / g_code.n;
        std::vector<uint8_t> info(info_bits, info_bits + static_cast<size_t>(blocks) * g_code.k);
        const auto f = build_tx_frame_from_info(g_phy, g_code, info);
        for (size_t i = 0; i < f.size(); ++i) {
            out_iq[2 * i] = f[i].real();
            out_iq[2 * i + 1] = f[i].imag();
        }
        return static_cast<int>(f.size());
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// find_frame (LINK:5123) on an interleaved IQ buffer.
int ref_find_frame(const float* iq, int count, int* start, double* peak, double* cfo_hz)
{
    if (!g_have_phy) return -1;
    std::deque<cf32> buf;
    for (int i = 0; i < count; ++i) buf.emplace_back(iq[2 * i], iq[2 * i + 1]);
    size_t s = 0;
    double p = 0.0, c = 0.0;
    const bool found = find_frame(buf, g_phy, s, p, c);
    *start = found ? static_cast<int>(s) : -1;
    *peak = p;
    *cfo_hz = c;
    return found ? 1 : 0;
}

// fft_to_shifted_grid (LINK:3845) of one nfft-sample symbol.
int ref_fft_shifted(const float* iq, int nfft, float* out)
{
    std::vector<cf32> t(static_cast<size_t>(nfft));
    for (int i = 0; i < nfft; ++i) t[i] = cf32(iq[2 * i], iq[2 * i + 1]);
    const auto g = fft_to_shifted_grid(t.data(), nfft);
    for (int i = 0; i < nfft; ++i) {
        out[2 * i] = g[i].real();
        out[2 * i + 1] = g[i].imag();
    }
    return 0;
}

// symbol_to_llr_maxlog (LINK:2552) under the configured modulation.
int ref_symbol_llr(const float* iq, int count, const double* sigma2, float* out)
{
    if (!g_have_phy) return -1;
    for (int i = 0; i < count; ++i)
        symbol_to_llr_maxlog(g_phy, cf32(iq[2 * i], iq[2 * i + 1]), sigma2[i],
                             out + static_cast<size_t>(i) * g_phy.bits_per_symbol);
    return 0;
}

// decode_frame (LINK:5331) with the CPU min-sum FEC; known_noise_var <= 0 uses
// the decision-directed estimate.
int ref_decode_frame(const float* frame_iq, double cfo_hz, double known_noise_var, int max_iter, double norm,
                     uint8_t* info_out, int* parity_ok, double* snr_db)
{
    if (!g_have_phy || !g_have_code) return -1;
    try {
        std::vector<cf32> f(static_cast<size_t>(g_phy.frame_len()));
        for (size_t i = 0; i < f.size(); ++i) f[i] = cf32(frame_iq[2 * i], frame_iq[2 * i + 1]);
        std::optional<double> nv;
        if (known_noise_var > 0.0) nv = known_noise_var;
        const FrameDecodeResult r = decode_frame(f, g_phy, g_code, nullptr, nullptr, cfo_hz, 1.0, max_iter, norm,
                                                 -120.0, false, nullptr, std::nullopt, nv);
        std::memcpy(info_out, r.info_bits.data(), r.info_bits.size());
        *parity_ok = r.parity_ok ? 1 : 0;
        *snr_db = r.snr_db;
        return static_cast<int>(r.info_bits.size());
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// build_info_bits (LINK:4048-4056): the reference payload of a frame id.
int ref_build_info_bits(unsigned frame_id, uint8_t* out)
{
    if (!g_have_phy || !g_have_code) return -1;
    const auto b = build_info_bits(g_phy, g_code, frame_id);
    std::memcpy(out, b.data(), b.size());
    return static_cast<int>(b.size());
}

// decode_frame (LINK:5331) with the CPU min-sum FEC, every FrameDecodeResult
// field: ires = {detected, parity_ok, frame_ok, frame_id, bit_errors,
// pre_fec_bit_errors, iterations}, dres = {peak, cfo_hz, snr_db, evm_rms}.
// expected_id < 0: the reference id is the decoded one.
int ref_decode_frame_full(const float* frame_iq, double cfo_hz, double peak, double known_noise_var, int max_iter,
                          double norm, double snr_gate_db, int reference_test, long long expected_id,
                          uint8_t* info_out, long long* ires, double* dres)
{
    if (!g_have_phy || !g_have_code) return -1;
    try {
        std::vector<cf32> f(static_cast<size_t>(g_phy.frame_len()));
        for (size_t i = 0; i < f.size(); ++i) f[i] = cf32(frame_iq[2 * i], frame_iq[2 * i + 1]);
        std::optional<double> nv;
        if (known_noise_var > 0.0) nv = known_noise_var;
        std::optional<uint32_t> eid;
        if (expected_id >= 0) eid = static_cast<uint32_t>(expected_id);
        const FrameDecodeResult r = decode_frame(f, g_phy, g_code, nullptr, nullptr, cfo_hz, peak, max_iter, norm,
                                                 snr_gate_db, reference_test != 0, nullptr, eid, nv);
        if (!r.info_bits.empty()) std::memcpy(info_out, r.info_bits.data(), r.info_bits.size());
        ires[0] = r.detected;
        ires[1] = r.parity_ok;
        ires[2] = r.frame_ok;
        ires[3] = r.frame_id;
        ires[4] = r.bit_errors;
        ires[5] = r.pre_fec_bit_errors;
        ires[6] = r.iterations;
        dres[0] = r.peak;
        dres[1] = r.cfo_hz;
        dres[2] = r.snr_db;
        dres[3] = r.evm_rms;
        return static_cast<int>(r.info_bits.size());
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

}  // extern "C"
