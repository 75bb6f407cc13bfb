// ofdm_rx.cpp -- OFDM receiver tables, the sicc_ofdm_* C ABI and the
// cuda_osd_find_frame entry point of the decoder ABI.
#include "ofdm_rx.hpp"

#include "../../include/matlab_decoder_c_api.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace sicc {

std::vector<float2> qam_constellation(int bits)
{
    std::vector<float2> c;
    if (bits == 1) return {make_float2(1.0f, 0.0f), make_float2(-1.0f, 0.0f)};
    if (bits == 2) {
        const float s = static_cast<float>(1.0 / std::sqrt(2.0));
        for (int idx = 0; idx < 4; ++idx) c.push_back(make_float2((idx & 2) ? -s : s, (idx & 1) ? -s : s));
        return c;
    }
    if (bits != 4 && bits != 6) throw std::invalid_argument("bits_per_symbol must be 1, 2, 4 or 6");
    const int axis = bits / 2, levels = 1 << axis, m = 1 << bits;
    auto g2b = [](int g) {
        int b = 0;
        for (; g != 0; g >>= 1) b ^= g;
        return b;
    };
    double avg = 0.0;
    std::vector<float> ii(m), qq(m);
    for (int idx = 0; idx < m; ++idx) {
        const int gi = idx >> axis, gq = idx & (levels - 1);
        ii[idx] = static_cast<float>((levels - 1) - 2 * g2b(gi));
        qq[idx] = static_cast<float>((levels - 1) - 2 * g2b(gq));
        avg += static_cast<double>(ii[idx] * ii[idx] + qq[idx] * qq[idx]);
    }
    const float scale = static_cast<float>(1.0 / std::sqrt(avg / static_cast<double>(m)));
    for (int idx = 0; idx < m; ++idx) c.push_back(make_float2(ii[idx] * scale, qq[idx] * scale));
    return c;
}

namespace {

void check_common(const sicc_ofdm_config_t& cfg)
{
    if (cfg.nfft < 16 || cfg.nfft > 2048 || (cfg.nfft & (cfg.nfft - 1)) != 0)
        throw std::invalid_argument("nfft must be a power of two in [16, 2048]");
    if (cfg.cp < 0 || cfg.cp >= cfg.nfft) throw std::invalid_argument("cp must be in [0, nfft)");
    if (cfg.pre_half_len <= 0 || cfg.preamble_len != 2 * cfg.pre_half_len)
        throw std::invalid_argument("preamble_len must be 2 * pre_half_len (two identical halves)");
    if (cfg.active_sc <= 0 || cfg.active_sc >= cfg.nfft) throw std::invalid_argument("active_sc must be in (0, nfft)");
    if (cfg.num_symbols <= 0) throw std::invalid_argument("num_symbols must be positive");
}

std::vector<int> natural_bins(const sicc_ofdm_config_t& cfg, const int* used_shifted)
{
    std::vector<int> nat(cfg.active_sc);
    for (int sc = 0; sc < cfg.active_sc; ++sc) {
        if (used_shifted[sc] < 0 || used_shifted[sc] >= cfg.nfft) throw std::invalid_argument("subcarrier out of range");
        nat[sc] = (used_shifted[sc] + cfg.nfft / 2) % cfg.nfft;
    }
    return nat;
}

}  // namespace

OfdmRx::OfdmRx(const sicc_ofdm_config_t& cfg, const float2* preamble, const int* used_shifted, int n_pilot,
               const int* pilot_symbols, const float2* pilot_grid)
    : cfg_(cfg)
{
    check_common(cfg);
    if (n_pilot < 0 || n_pilot >= cfg.num_symbols) throw std::invalid_argument("need 0 <= pilots < num_symbols");
    if (cfg.use_pilot_channel_est && n_pilot == 0) throw std::invalid_argument("pilot estimation needs a pilot symbol");
    if (!preamble || !used_shifted || (n_pilot > 0 && (!pilot_symbols || !pilot_grid)))
        throw std::invalid_argument("null table");
    std::vector<int> ps(pilot_symbols, pilot_symbols + n_pilot);
    for (size_t i = 0; i < ps.size(); ++i) {
        if (ps[i] < 0 || ps[i] >= cfg.num_symbols) throw std::invalid_argument("pilot symbol out of range");
        if (i > 0 && ps[i] <= ps[i - 1]) throw std::invalid_argument("pilot symbols must be strictly increasing");
    }
    std::vector<int> ds;
    for (int s = 0; s < cfg.num_symbols; ++s)
        if (std::find(ps.begin(), ps.end(), s) == ps.end()) ds.push_back(s);
    const int n_data = static_cast<int>(ds.size());
    // interpolation tables (decode_frame, uhd_ldpc_ofdm_link.cpp:5402-5420)
    std::vector<int> p0(n_data, 0), p1(n_data, 0);
    std::vector<float> w0(n_data, 1.0f), w1(n_data, 0.0f);
    for (int di = 0; di < n_data && n_pilot > 0; ++di) {
        const int sym = ds[di];
        int hi = 0;
        while (hi + 1 < n_pilot && sym > ps[hi + 1]) ++hi;
        int a = hi, b = std::min(hi + 1, n_pilot - 1);
        if (sym <= ps.front()) a = b = 0;
        else if (sym >= ps.back()) a = b = n_pilot - 1;
        double alpha = 0.0;
        if (a != b) alpha = static_cast<double>(sym - ps[a]) / static_cast<double>(ps[b] - ps[a]);
        p0[di] = a;
        p1[di] = b;
        w0[di] = static_cast<float>(1.0 - alpha);
        w1[di] = static_cast<float>(alpha);
    }
    init(preamble, natural_bins(cfg, used_shifted), ps, pilot_grid, ds, p0, p1, w0, w1);
}

OfdmRx::OfdmRx(const sicc_ofdm_config_t& cfg, const float2* preamble, const int* used_shifted,
               const std::vector<int>& pilot_symbols, const float2* pilot_grid, const std::vector<int>& data_symbols,
               const std::vector<int>& interp_p0, const std::vector<int>& interp_p1,
               const std::vector<float>& interp_alpha)
    : cfg_(cfg)
{
    check_common(cfg);
    const int n_pilot = static_cast<int>(pilot_symbols.size());
    const size_t nd = data_symbols.size();
    if (!preamble || !used_shifted || (n_pilot > 0 && !pilot_grid)) throw std::invalid_argument("null table");
    if (interp_p0.size() != nd || interp_p1.size() != nd || interp_alpha.size() != nd)
        throw std::invalid_argument("interpolation tables must have one entry per data symbol");
    if (cfg.use_pilot_channel_est && n_pilot == 0) throw std::invalid_argument("pilot estimation needs a pilot symbol");
    for (int s : pilot_symbols)
        if (s < 0 || s >= cfg.num_symbols) throw std::invalid_argument("pilot symbol out of range");
    for (int s : data_symbols)
        if (s < 0 || s >= cfg.num_symbols) throw std::invalid_argument("data symbol out of range");
    std::vector<float> w0(nd), w1(nd);
    for (size_t i = 0; i < nd; ++i) {
        if (cfg.use_pilot_channel_est && (interp_p0[i] < 0 || interp_p0[i] >= n_pilot || interp_p1[i] < 0 ||
                                          interp_p1[i] >= n_pilot))
            throw std::invalid_argument("interpolation pilot index out of range");
        w0[i] = 1.0f - interp_alpha[i];
        w1[i] = interp_alpha[i];
    }
    init(preamble, natural_bins(cfg, used_shifted), pilot_symbols, pilot_grid, data_symbols, interp_p0, interp_p1, w0,
         w1);
}

void OfdmRx::init(const float2* preamble, const std::vector<int>& nat, const std::vector<int>& ps,
                  const float2* pilot_grid, const std::vector<int>& ds, const std::vector<int>& p0,
                  const std::vector<int>& p1, const std::vector<float>& w0, const std::vector<float>& w1)
{
    const auto& cfg = cfg_;
    if (cfg_.refine_radius < 0) cfg_.refine_radius = std::min(16, cfg.cp);
    n_pilot_ = static_cast<int>(ps.size());
    n_data_ = static_cast<int>(ds.size());
    const long capacity = static_cast<long>(n_data_) * cfg.active_sc * cfg.bits_per_symbol;
    if (cfg.blocks_per_frame <= 0 || cfg.code_n <= 0 || static_cast<long>(cfg.blocks_per_frame) * cfg.code_n > capacity)
        throw std::invalid_argument("blocks_per_frame * code_n must be positive and fit the data resource elements");
    pre_energy_ = 0.0;
    for (int i = 0; i < cfg.preamble_len; ++i)
        pre_energy_ += static_cast<double>(preamble[i].x * preamble[i].x + preamble[i].y * preamble[i].y);
    std::vector<double2> tw(cfg.nfft);
    for (int m = 0; m < cfg.nfft; ++m) {
        const double a = -2.0 * 3.141592653589793238462643383279502884 * m / cfg.nfft;
        tw[m] = make_double2(std::cos(a), std::sin(a));
    }
    const auto con = qam_constellation(cfg.bits_per_symbol);
    {
        // The demapper evaluates distances per axis: point (gi << qb) | gq must sit
        // at (x of point gi << qb, y of point gq).
        const int B = cfg.bits_per_symbol, ax = B == 1 ? 1 : B / 2, qb = B - ax;
        for (int idx = 0; idx < static_cast<int>(con.size()); ++idx)
            if (con[idx].x != con[(idx >> qb) << qb].x || con[idx].y != con[idx & ((1 << qb) - 1)].y)
                throw std::logic_error("constellation is not a product of per-axis levels");
    }

    auto up = [](auto& buf, const auto* src, size_t count) {
        buf.alloc(std::max<size_t>(count, 1));
        if (count) SICC_HIP(hipMemcpy(buf.get(), src, count * sizeof(*src), hipMemcpyHostToDevice));
    };
    up(preamble_, preamble, cfg.preamble_len);
    {
        std::vector<double2> pd(cfg.preamble_len);
        for (int i = 0; i < cfg.preamble_len; ++i) pd[i] = make_double2(preamble[i].x, preamble[i].y);
        up(preamble_d_, pd.data(), pd.size());
    }
    up(pilot_grid_, pilot_grid, static_cast<size_t>(n_pilot_) * cfg.active_sc);
    up(constellation_, con.data(), con.size());
    up(pilot_symbols_, ps.data(), ps.size());
    up(data_symbols_, ds.data(), ds.size());
    up(used_natural_, nat.data(), nat.size());
    up(p0_, p0.data(), p0.size());
    up(p1_, p1.data(), p1.size());
    up(w0_, w0.data(), w0.size());
    up(w1_, w1.data(), w1.size());
    up(twiddle_, tw.data(), tw.size());
    std::vector<float2> twf(tw.size());
    for (size_t m = 0; m < tw.size(); ++m) twf[m] = make_float2(static_cast<float>(tw[m].x), static_cast<float>(tw[m].y));
    up(twiddle_f_, twf.data(), twf.size());
}

void OfdmRx::sync(const float2* d_rx, int frames, int span, OfdmSyncResult* d_out, int max_search, hipStream_t s) const
{
    if (!d_rx || !d_out || frames <= 0) throw std::invalid_argument("sync: empty batch");
    if (span < cfg_.preamble_len) throw std::invalid_argument("sync: rows shorter than the preamble");
    OfdmSyncArgs a;
    a.rx = d_rx;
    a.span = span;
    a.half_len = cfg_.pre_half_len;
    a.max_search = max_search;
    a.radius = cfg_.refine_radius;
    a.threshold = cfg_.sync_threshold;
    a.fs = cfg_.sample_rate;
    a.preamble = preamble_.get();
    a.preamble_d = preamble_d_.get();
    a.pre_energy = pre_energy_;
    a.out = d_out;
    launch_ofdm_sync(a, frames, s);
}

void OfdmRx::demod(const float2* d_rx, int frames, int span, const OfdmSyncResult* d_sync, float* d_llr,
                   hipStream_t s, double* d_stats) const
{
    if (!d_rx || !d_sync || !d_llr || frames <= 0) throw std::invalid_argument("demod: empty batch");
    if (span < frame_samples()) throw std::invalid_argument("demod: rows shorter than one frame");
    OfdmDemodArgs a;
    a.rx = d_rx;
    a.span = span;
    a.frame_samples = frame_samples();
    a.sync = d_sync;
    a.nfft = cfg_.nfft;
    a.cp = cfg_.cp;
    a.preamble_len = cfg_.preamble_len;
    a.active_sc = cfg_.active_sc;
    a.bits_per_symbol = cfg_.bits_per_symbol;
    a.n_pilot = n_pilot_;
    a.n_data = n_data_;
    a.pilot_symbols = pilot_symbols_.get();
    a.data_symbols = data_symbols_.get();
    a.used_natural = used_natural_.get();
    a.pilot_grid = pilot_grid_.get();
    a.interp_p0 = p0_.get();
    a.interp_p1 = p1_.get();
    a.interp_w0 = w0_.get();
    a.interp_w1 = w1_.get();
    a.constellation = constellation_.get();
    a.twiddle = twiddle_.get();
    const double nv = std::max(static_cast<double>(cfg_.demod_noise_var), 1e-8);
    a.noise_var = static_cast<float>(nv);
    a.noise_var_d = nv;
    a.inv_noise_var_d = 1.0 / nv;
    a.fs = cfg_.sample_rate;
    a.use_pilots = cfg_.use_pilot_channel_est;
    a.llr = d_llr;
    a.llr_per_frame = llr_per_frame();
    a.stats = d_stats;
    if (cfg_.use_pilot_channel_est) {
        const size_t need = static_cast<size_t>(frames) * n_pilot_ * cfg_.active_sc;
        if (h_est_.size() < need) {
            SICC_HIP(hipStreamSynchronize(s));  // an earlier demod on this stream may still read it
            h_est_.alloc(need);
        }
        a.h_est = h_est_.get();
    }
    launch_ofdm_demod(a, frames, s);
}

namespace {

struct OfdmHandle {
    int device = 0;
    hipStream_t stream = nullptr;
    std::shared_ptr<OfdmRx> rx;
    DevBuf<float2> h_rx;
    DevBuf<OfdmSyncResult> h_sync;
    DevBuf<float> h_llr;
    ~OfdmHandle()
    {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }
};

HandleTable<OfdmHandle>& otable()
{
    // never destroyed: handles still open at exit must not reach the HIP
    // runtime from static destructors
    static HandleTable<OfdmHandle>& t = *new HandleTable<OfdmHandle>;
    return t;
}

template <class F>
int oguard(F&& f)
{
    try {
        return f();
    } catch (const std::invalid_argument& e) {
        set_last_error(e.what());
        return CUDA_OSD_STATUS_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return CUDA_OSD_STATUS_INTERNAL_ERROR;
    } catch (...) {
        set_last_error("unknown error");
        return CUDA_OSD_STATUS_INTERNAL_ERROR;
    }
}

std::shared_ptr<OfdmHandle> olookup(unsigned long long h)
{
    auto p = otable().get(h);
    if (!p) throw std::invalid_argument("ofdm: unknown handle");
    SICC_HIP(hipSetDevice(p->device));
    return p;
}

static_assert(sizeof(sicc_ofdm_sync_t) == sizeof(OfdmSyncResult), "sync result layout");

}  // namespace

std::shared_ptr<OfdmRx> ofdm_receiver(unsigned long long handle) { return olookup(handle)->rx; }
}  // namespace sicc

using namespace sicc;

extern "C" {

int sicc_ofdm_create(const sicc_ofdm_config_t* config, const float* preamble_iq, const int* used_shifted, int n_pilot,
                     const int* pilot_symbols, const float* pilot_grid_iq, unsigned long long* out_handle)
{
    return oguard([&] {
        if (!out_handle || !config) throw std::invalid_argument("null argument");
        *out_handle = 0;
        auto h = std::make_unique<OfdmHandle>();
        SICC_HIP(hipGetDevice(&h->device));
        h->rx = std::make_unique<OfdmRx>(*config, reinterpret_cast<const float2*>(preamble_iq), used_shifted, n_pilot,
                                         pilot_symbols, reinterpret_cast<const float2*>(pilot_grid_iq));
        SICC_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        *out_handle = otable().add(std::move(h));
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_ofdm_destroy(unsigned long long handle)
{
    return oguard([&] {
        if (!otable().remove(handle)) throw std::invalid_argument("ofdm: unknown handle");
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_ofdm_llr_per_frame(unsigned long long handle)
{
    auto p = otable().get(handle);
    return p ? p->rx->llr_per_frame() : -1;
}

int sicc_ofdm_sync_device(unsigned long long handle, const float* d_rx_iq, int frames, int span, sicc_ofdm_sync_t* d_sync,
                          void* hip_stream)
{
    return oguard([&] {
        auto h = olookup(handle);
        hipStream_t s = hip_stream ? static_cast<hipStream_t>(hip_stream) : h->stream;
        h->rx->sync(reinterpret_cast<const float2*>(d_rx_iq), frames, span, reinterpret_cast<OfdmSyncResult*>(d_sync),
                    h->rx->config().max_rx_search_samples, s);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_ofdm_demod_device(unsigned long long handle, const float* d_rx_iq, int frames, int span,
                           const sicc_ofdm_sync_t* d_sync, float* d_llr, void* hip_stream)
{
    return oguard([&] {
        auto h = olookup(handle);
        hipStream_t s = hip_stream ? static_cast<hipStream_t>(hip_stream) : h->stream;
        h->rx->demod(reinterpret_cast<const float2*>(d_rx_iq), frames, span,
                     reinterpret_cast<const OfdmSyncResult*>(d_sync), d_llr, s);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_ofdm_receive_device(unsigned long long handle, const float* d_rx_iq, int frames, int span,
                             sicc_ofdm_sync_t* d_sync, float* d_llr, void* hip_stream)
{
    return oguard([&] {
        auto h = olookup(handle);
        hipStream_t s = hip_stream ? static_cast<hipStream_t>(hip_stream) : h->stream;
        const auto* rx = reinterpret_cast<const float2*>(d_rx_iq);
        auto* sy = reinterpret_cast<OfdmSyncResult*>(d_sync);
        h->rx->sync(rx, frames, span, sy, h->rx->config().max_rx_search_samples, s);
        h->rx->demod(rx, frames, span, sy, d_llr, s);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_ofdm_receive_host(unsigned long long handle, const float* rx_iq, int frames, int span,
                           sicc_ofdm_sync_t* sync_out, float* llr_out)
{
    return oguard([&] {
        auto h = olookup(handle);
        if (!rx_iq || frames <= 0 || span <= 0) throw std::invalid_argument("receive_host: empty batch");
        const size_t ns = static_cast<size_t>(frames) * span;
        const size_t nl = static_cast<size_t>(frames) * h->rx->llr_per_frame();
        if (h->h_rx.size() < ns) h->h_rx.alloc(ns);
        if (h->h_sync.size() < static_cast<size_t>(frames)) h->h_sync.alloc(frames);
        if (h->h_llr.size() < nl) h->h_llr.alloc(nl);
        SICC_HIP(hipMemcpyAsync(h->h_rx.get(), rx_iq, ns * sizeof(float2), hipMemcpyHostToDevice, h->stream));
        h->rx->sync(h->h_rx.get(), frames, span, h->h_sync.get(), h->rx->config().max_rx_search_samples, h->stream);
        h->rx->demod(h->h_rx.get(), frames, span, h->h_sync.get(), h->h_llr.get(), h->stream);
        if (sync_out)
            SICC_HIP(hipMemcpyAsync(sync_out, h->h_sync.get(), frames * sizeof(OfdmSyncResult), hipMemcpyDeviceToHost,
                                    h->stream));
        if (llr_out)
            SICC_HIP(hipMemcpyAsync(llr_out, h->h_llr.get(), nl * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        SICC_HIP(hipStreamSynchronize(h->stream));
        return CUDA_OSD_STATUS_OK;
    });
}

// cuda_osd_find_frame (matlab_decoder_c_api.h, reference OSDH:121-132): Schmidl-Cox
// search over the whole signal on the GPU; the start index is 1-based, and 0
// when no frame passes sync_peak_thresh.
int cuda_osd_find_frame(const double* signal_real, const double* signal_imag, int signal_length,
                        const double* preamble_real, const double* preamble_imag, int preamble_length,
                        int pre_half_len, double sync_peak_thresh, double fs, int* out_start_idx_1based,
                        double* out_peak, double* out_cfo_hz)
{
    return oguard([&] {
        if (!signal_real || !signal_imag || !preamble_real || !preamble_imag || !out_start_idx_1based)
            throw std::invalid_argument("find_frame: null argument");
        if (pre_half_len <= 0 || preamble_length != 2 * pre_half_len || signal_length < preamble_length)
            throw std::invalid_argument("find_frame: need preamble_length = 2 * pre_half_len <= signal_length");
        std::vector<float2> sig(signal_length), pre(preamble_length);
        double pe = 0.0;
        for (int i = 0; i < signal_length; ++i)
            sig[i] = make_float2(static_cast<float>(signal_real[i]), static_cast<float>(signal_imag[i]));
        for (int i = 0; i < preamble_length; ++i) {
            pre[i] = make_float2(static_cast<float>(preamble_real[i]), static_cast<float>(preamble_imag[i]));
            pe += static_cast<double>(pre[i].x * pre[i].x + pre[i].y * pre[i].y);
        }
        hipStream_t s = nullptr;
        SICC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        std::vector<double2> pred(preamble_length);
        for (int i = 0; i < preamble_length; ++i) pred[i] = make_double2(pre[i].x, pre[i].y);
        DevBuf<float2> d_sig(signal_length), d_pre(preamble_length);
        DevBuf<double2> d_pred(preamble_length);
        DevBuf<OfdmSyncResult> d_out(1);
        SICC_HIP(hipMemcpyAsync(d_sig.get(), sig.data(), sig.size() * sizeof(float2), hipMemcpyHostToDevice, s));
        SICC_HIP(hipMemcpyAsync(d_pre.get(), pre.data(), pre.size() * sizeof(float2), hipMemcpyHostToDevice, s));
        SICC_HIP(hipMemcpyAsync(d_pred.get(), pred.data(), pred.size() * sizeof(double2), hipMemcpyHostToDevice, s));
        OfdmSyncArgs a;
        a.rx = d_sig.get();
        a.span = signal_length;
        a.half_len = pre_half_len;
        a.max_search = 0;
        a.radius = 16;
        a.threshold = sync_peak_thresh;
        a.fs = fs;
        a.preamble = d_pre.get();
        a.preamble_d = d_pred.get();
        a.pre_energy = pe;
        a.out = d_out.get();
        launch_ofdm_sync(a, 1, s);
        OfdmSyncResult r{};
        SICC_HIP(hipMemcpyAsync(&r, d_out.get(), sizeof(r), hipMemcpyDeviceToHost, s));
        SICC_HIP(hipStreamSynchronize(s));
        SICC_HIP(hipStreamDestroy(s));
        *out_start_idx_1based = r.start >= 0 ? r.start + 1 : 0;
        if (out_peak) *out_peak = r.start >= 0 ? r.metric : r.coarse_metric;
        if (out_cfo_hz) *out_cfo_hz = r.cfo_hz_d;
        return CUDA_OSD_STATUS_OK;
    });
}

}  // extern "C"
