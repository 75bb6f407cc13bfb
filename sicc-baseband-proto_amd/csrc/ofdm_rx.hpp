// ofdm_rx.hpp -- host object owning the device tables of one OFDM receiver.
#pragma once

#include "../../include/sicc_ofdm_c_api.h"
#include "ofdm.hpp"
#include "sicc_common.hpp"

#include <memory>
#include <vector>

namespace sicc {

class OfdmRx {
public:
    // Tables: preamble [preamble_len] complex, used_shifted [active_sc],
    // pilot_symbols [n_pilot], pilot_grid [n_pilot][active_sc] complex.
    OfdmRx(const sicc_ofdm_config_t& cfg, const float2* preamble, const int* used_shifted, int n_pilot,
           const int* pilot_symbols, const float2* pilot_grid);
    // Caller-supplied symbol and interpolation tables (the pipeline front's
    // StaticTables): data symbol d is equalized with
    // (1 - alpha[d]) * h[p0[d]] + alpha[d] * h[p1[d]] (float weights).
    OfdmRx(const sicc_ofdm_config_t& cfg, const float2* preamble, const int* used_shifted,
           const std::vector<int>& pilot_symbols, const float2* pilot_grid, const std::vector<int>& data_symbols,
           const std::vector<int>& interp_p0, const std::vector<int>& interp_p1,
           const std::vector<float>& interp_alpha);

    const sicc_ofdm_config_t& config() const { return cfg_; }
    int frame_samples() const { return cfg_.preamble_len + cfg_.num_symbols * (cfg_.nfft + cfg_.cp); }
    int llr_per_frame() const { return cfg_.blocks_per_frame * cfg_.code_n; }

    void sync(const float2* d_rx, int frames, int span, OfdmSyncResult* d_out, int max_search, hipStream_t s) const;
    // d_stats (optional): [frames][n_data][4] per-symbol SNR / EVM sums (OfdmDemodArgs::stats).
    void demod(const float2* d_rx, int frames, int span, const OfdmSyncResult* d_sync, float* d_llr,
               hipStream_t s, double* d_stats = nullptr) const;
    int n_data() const { return n_data_; }
    int n_pilot() const { return n_pilot_; }

    // device tables shared with a transmitter of the same numerology
    const float2* d_preamble() const { return preamble_.get(); }
    const float2* d_pilot_grid() const { return pilot_grid_.get(); }
    const float2* d_constellation() const { return constellation_.get(); }
    const int* d_used_natural() const { return used_natural_.get(); }
    const double2* d_twiddle() const { return twiddle_.get(); }
    const float2* d_twiddle_f() const { return twiddle_f_.get(); }  // the transmitter's FP32 FFT

private:
    void init(const float2* preamble, const std::vector<int>& nat, const std::vector<int>& ps,
              const float2* pilot_grid, const std::vector<int>& ds, const std::vector<int>& p0,
              const std::vector<int>& p1, const std::vector<float>& w0, const std::vector<float>& w1);

    sicc_ofdm_config_t cfg_;
    int n_pilot_ = 0;
    int n_data_ = 0;
    double pre_energy_ = 0.0;
    DevBuf<float2> preamble_, pilot_grid_, constellation_;
    DevBuf<double2> preamble_d_;  // the preamble's float samples widened to double (sync refinement)
    DevBuf<int> pilot_symbols_, data_symbols_, used_natural_, p0_, p1_;
    DevBuf<float> w0_, w1_;
    DevBuf<double2> twiddle_;
    DevBuf<float2> twiddle_f_;
    // LS channel estimates of the last demod batch; demods on one receiver
    // must be issued on one stream (or not overlap)
    mutable DevBuf<float2> h_est_;
};

// The receiver behind a sicc_ofdm_* handle (kept alive by the returned
// pointer); throws std::invalid_argument for an unknown handle.
std::shared_ptr<OfdmRx> ofdm_receiver(unsigned long long handle);

// Gray QAM constellation indexed by the symbol's bits MSB first, unit average
// power (configure_modulation, uhd_ldpc_ofdm_link.cpp:2362-2414).
std::vector<float2> qam_constellation(int bits_per_symbol);

}  // namespace sicc
