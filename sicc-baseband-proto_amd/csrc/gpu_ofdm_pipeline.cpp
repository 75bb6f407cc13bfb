// gpu_ofdm_pipeline.cpp -- usrp_gpu_pipeline_bridge::Pipeline on the gfx950
// OFDM kernels, and its plain-C layer (sicc_pipeline_*).
//
// Reference: source/cpp_cuda/uhd_cpp/src/gpu_ofdm_pipeline.cu:791-1668
// (GpuOfdmPipeline and the bridge methods at 1552-1667).  Per slot the object
// owns the info / coded bits, the transmitted frame, a receive sample buffer
// of max(max_rx_search_samples, frame_samples) samples, the sync result and
// the frame's LLRs, so that up to `slots` frames are in flight.  TX work runs
// on one stream, receive work on another.
#include "../../include/gpu_ofdm_pipeline.hpp"
#include "../../include/matlab_decoder_c_api.h"
#include "../../include/sicc_pipeline_c_api.h"
#include "ofdm_rx.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <unordered_map>

namespace usrp_gpu_pipeline_bridge {

namespace {

using sicc::DevBuf;
using sicc::OfdmSyncResult;

// Argument errors: a std::runtime_error, as every error of the reference class,
// that the C layer reports as CUDA_OSD_STATUS_INVALID_ARGUMENT.
struct ArgumentError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check_hip(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void require_size(size_t actual, size_t expected, const char* name)
{
    if (actual != expected)
        throw ArgumentError(std::string(name) + " has wrong size; expected " + std::to_string(expected) + ", got " +
                            std::to_string(actual));
}

const float2* as_f2(const Complex32* p) { return reinterpret_cast<const float2*>(p); }

template <class T>
void upload(DevBuf<T>& dst, const T* src, size_t count)
{
    dst.alloc(std::max<size_t>(count, 1));
    if (count) check_hip(hipMemcpy(dst.get(), src, count * sizeof(T), hipMemcpyHostToDevice), "upload table");
}

}  // namespace

static_assert(sizeof(Complex32) == sizeof(float2), "Complex32 must be two packed floats");

namespace {

// The pipeline state; Pipeline::Impl holds one, and the C layer reaches it
// through core_of() for the batched transmitter.
struct PipelineCore {
    struct Slot {
        DevBuf<uint8_t> info, coded;
        DevBuf<float2> tx_frame, rx_samples;
        DevBuf<OfdmSyncResult> sync;
        DevBuf<float> llr;
        OfdmSyncResult* host_sync = nullptr;  // pinned
        hipEvent_t tx_done = nullptr;
    };

    Config cfg;
    int device = 0;
    int max_rx = 0;
    hipStream_t tx_stream = nullptr;
    hipStream_t rx_stream = nullptr;
    std::vector<Slot> slots;
    // tables
    std::unique_ptr<sicc::OfdmRx> rx;
    bool dvb_s2 = false;
    int gen_words = 0;
    DevBuf<unsigned long long> gen_cols;
    DevBuf<int> row_offsets, row_cols, pilot_index, data_index;
    DevBuf<float2> precomp_h, actual_h;
    bool have_precomp = false, have_actual = false;
    DevBuf<uint8_t> batch_coded;

    PipelineCore(const Config& c, int n_slots) : cfg(c)
    {
        if (cfg.blocks_per_frame <= 0)
            throw ArgumentError("blocks_per_frame must be set before constructing GpuOfdmPipeline");
        if (cfg.active_sc <= 0 || cfg.bits_per_symbol <= 0 ||
            cfg.coded_bits_per_frame() != cfg.active_sc * cfg.data_symbols_per_frame() * cfg.bits_per_symbol)
            throw ArgumentError("coded bits per frame must match active data resource elements");
        if (cfg.bits_per_symbol != 1 && cfg.bits_per_symbol != 2 && cfg.bits_per_symbol != 4 &&
            cfg.bits_per_symbol != 6)
            throw ArgumentError("bits_per_symbol must be 1, 2, 4, or 6");
        if (cfg.nfft < 16 || cfg.nfft > 2048 || (cfg.nfft & (cfg.nfft - 1)) != 0)
            throw ArgumentError("nfft must be a power of two in [16, 2048]");
        if (cfg.cp < 0 || cfg.cp >= cfg.nfft) throw ArgumentError("cp must be in [0, nfft)");
        if (cfg.pre_half_len <= 0 || cfg.preamble_len != 2 * cfg.pre_half_len)
            throw ArgumentError("preamble_len must be 2 * pre_half_len");
        if (cfg.code_k <= 0 || cfg.code_n <= cfg.code_k) throw ArgumentError("need 0 < code_k < code_n");
        check_hip(hipGetDevice(&device), "hipGetDevice");
        check_hip(hipStreamCreateWithFlags(&tx_stream, hipStreamNonBlocking), "create TX stream");
        check_hip(hipStreamCreateWithFlags(&rx_stream, hipStreamNonBlocking), "create RX stream");
        max_rx = cfg.max_rx_search_samples > 0 ? cfg.max_rx_search_samples : cfg.frame_samples();
        slots.resize(static_cast<size_t>(std::max(n_slots, 1)));
        for (auto& s : slots) {
            s.info.alloc(cfg.info_bits_per_frame());
            s.coded.alloc(cfg.coded_bits_per_frame());
            s.tx_frame.alloc(cfg.frame_samples());
            s.rx_samples.alloc(std::max(max_rx, cfg.frame_samples()));
            s.sync.alloc(1);
            s.llr.alloc(cfg.coded_bits_per_frame());
            check_hip(hipHostMalloc(reinterpret_cast<void**>(&s.host_sync), sizeof(OfdmSyncResult)), "pinned sync");
            check_hip(hipEventCreateWithFlags(&s.tx_done, hipEventDisableTiming), "create tx_done");
        }
    }

    ~PipelineCore()
    {
        (void)hipSetDevice(device);
        if (tx_stream) (void)hipStreamSynchronize(tx_stream);
        if (rx_stream) (void)hipStreamSynchronize(rx_stream);
        for (auto& s : slots) {
            if (s.host_sync) (void)hipHostFree(s.host_sync);
            if (s.tx_done) (void)hipEventDestroy(s.tx_done);
        }
        if (tx_stream) (void)hipStreamDestroy(tx_stream);
        if (rx_stream) (void)hipStreamDestroy(rx_stream);
    }

    Slot& slot(int i)
    {
        if (i < 0 || i >= static_cast<int>(slots.size())) throw ArgumentError("invalid GPU pipeline slot");
        check_hip(hipSetDevice(device), "hipSetDevice");
        return slots[static_cast<size_t>(i)];
    }

    sicc::OfdmRx& receiver()
    {
        if (!rx) throw ArgumentError("GPU RX requested before uploading static tables");
        return *rx;
    }

    void upload_tables(const StaticTables& t)
    {
        check_hip(hipSetDevice(device), "hipSetDevice");
        const size_t A = static_cast<size_t>(cfg.active_sc);
        const size_t S = static_cast<size_t>(cfg.num_symbols);
        const int k = cfg.code_k, n = cfg.code_n, m = n - k;
        if (!t.generator_kn.empty()) require_size(t.generator_kn.size(), static_cast<size_t>(k) * n, "generator_kn");
        require_size(t.used_indices.size(), A, "used_indices");
        require_size(t.pilot_symbol_index.size(), S, "pilot_symbol_index");
        require_size(t.data_symbol_index.size(), S, "data_symbol_index");
        require_size(t.data_symbols.size(), static_cast<size_t>(cfg.data_symbols_per_frame()), "data_symbols");
        require_size(t.interp_p0.size(), t.data_symbols.size(), "interp_p0");
        require_size(t.interp_p1.size(), t.data_symbols.size(), "interp_p1");
        require_size(t.interp_alpha.size(), t.data_symbols.size(), "interp_alpha");
        require_size(t.preamble.size(), static_cast<size_t>(cfg.preamble_len), "preamble");
        require_size(t.pilot_grid.size(), t.pilot_symbols.size() * A, "pilot_grid");
        if (t.dvb_s2_accumulator) {
            require_size(t.row_info_offsets.size(), static_cast<size_t>(m) + 1, "row_info_offsets");
            const int total = static_cast<int>(t.row_info_cols.size());
            for (int r = 0; r < m; ++r)
                if (t.row_info_offsets[r] < 0 || t.row_info_offsets[r] > t.row_info_offsets[r + 1] ||
                    t.row_info_offsets[r + 1] > total)
                    throw ArgumentError("row_info_offsets must be non-decreasing within row_info_cols");
            for (int c : t.row_info_cols)
                if (c < 0 || c >= k) throw ArgumentError("row_info_cols entry out of range");
        }
        if (!t.precomp_h.empty()) require_size(t.precomp_h.size(), S * A, "precomp_h");
        if (!t.actual_h.empty()) require_size(t.actual_h.size(), S * A, "actual_h");
        const int n_pilot = static_cast<int>(t.pilot_symbols.size());
        const int n_data = static_cast<int>(t.data_symbols.size());
        for (size_t s = 0; s < S; ++s) {
            if (t.pilot_symbol_index[s] < -1 || t.pilot_symbol_index[s] >= n_pilot)
                throw ArgumentError("pilot_symbol_index entry out of range");
            if (t.data_symbol_index[s] < -1 || t.data_symbol_index[s] >= n_data)
                throw ArgumentError("data_symbol_index entry out of range");
        }

        sicc_ofdm_config_t rc{};
        rc.nfft = cfg.nfft;
        rc.cp = cfg.cp;
        rc.preamble_len = cfg.preamble_len;
        rc.pre_half_len = cfg.pre_half_len;
        rc.num_symbols = cfg.num_symbols;
        rc.active_sc = cfg.active_sc;
        rc.bits_per_symbol = cfg.bits_per_symbol;
        rc.code_n = cfg.code_n;
        rc.blocks_per_frame = cfg.blocks_per_frame;
        rc.max_rx_search_samples = 0;  // the sample count bounds the search
        rc.refine_radius = -1;
        rc.use_pilot_channel_est = cfg.use_pilot_channel_est ? 1 : 0;
        rc.sample_rate = cfg.sample_rate;
        rc.sync_threshold = cfg.sync_threshold;
        rc.demod_noise_var = cfg.demod_noise_var;
        std::unique_ptr<sicc::OfdmRx> nrx;
        try {
            nrx = std::make_unique<sicc::OfdmRx>(rc, as_f2(t.preamble.data()), t.used_indices.data(), t.pilot_symbols,
                                                 as_f2(t.pilot_grid.data()), t.data_symbols, t.interp_p0, t.interp_p1,
                                                 t.interp_alpha);
        } catch (const std::invalid_argument& e) {
            throw ArgumentError(e.what());
        }

        if (!t.generator_kn.empty()) {
            gen_words = (k + 63) / 64;
            std::vector<unsigned long long> cols(static_cast<size_t>(n) * gen_words, 0ull);
            for (int r = 0; r < k; ++r)
                for (int c = 0; c < n; ++c)
                    if (t.generator_kn[static_cast<size_t>(r) * n + c] & 1u)
                        cols[static_cast<size_t>(c) * gen_words + r / 64] |= 1ull << (r % 64);
            upload(gen_cols, cols.data(), cols.size());
        } else {
            gen_words = 0;
            gen_cols.release();
        }
        dvb_s2 = t.dvb_s2_accumulator;
        if (dvb_s2) {
            upload(row_offsets, t.row_info_offsets.data(), t.row_info_offsets.size());
            upload(row_cols, t.row_info_cols.data(), t.row_info_cols.size());
        }
        upload(pilot_index, t.pilot_symbol_index.data(), S);
        upload(data_index, t.data_symbol_index.data(), S);
        have_precomp = !t.precomp_h.empty();
        have_actual = !t.actual_h.empty();
        if (have_precomp) upload(precomp_h, as_f2(t.precomp_h.data()), t.precomp_h.size());
        if (have_actual) upload(actual_h, as_f2(t.actual_h.data()), t.actual_h.size());
        rx = std::move(nrx);
    }

    sicc::OfdmTxArgs tx_args(float noise_var, uint64_t seed)
    {
        if (!rx) throw ArgumentError("GPU TX requested before uploading static tables");
        if (!dvb_s2 && gen_words == 0) throw ArgumentError("GPU TX requested before uploading a generator matrix");
        sicc::OfdmTxArgs a;
        a.code_n = cfg.code_n;
        a.code_k = cfg.code_k;
        a.blocks = cfg.blocks_per_frame;
        a.gen_cols = gen_cols.get();
        a.gen_words = gen_words;
        a.dvb_s2 = dvb_s2 ? 1 : 0;
        a.row_info_offsets = dvb_s2 ? row_offsets.get() : nullptr;
        a.row_info_cols = dvb_s2 ? row_cols.get() : nullptr;
        a.coded_bits = cfg.coded_bits_per_frame();
        a.nfft = cfg.nfft;
        a.cp = cfg.cp;
        a.preamble_len = cfg.preamble_len;
        a.num_symbols = cfg.num_symbols;
        a.active_sc = cfg.active_sc;
        a.bits_per_symbol = cfg.bits_per_symbol;
        a.pilot_index = pilot_index.get();
        a.data_index = data_index.get();
        a.used_natural = rx->d_used_natural();
        a.pilot_grid = rx->d_pilot_grid();
        a.preamble = rx->d_preamble();
        a.constellation = rx->d_constellation();
        a.precomp_h = have_precomp ? precomp_h.get() : nullptr;
        a.actual_h = have_actual ? actual_h.get() : nullptr;
        a.twiddle = rx->d_twiddle_f();
        a.amplitude = cfg.tx_amplitude;
        a.noise_var = noise_var;
        a.noise_seed = seed;
        return a;
    }

    const float2* tx_device(int si, const uint8_t* h_info, float noise_var, uint64_t seed, bool add_noise)
    {
        Slot& s = slot(si);
        if (!h_info) throw ArgumentError("TX info bits pointer is null");
        sicc::OfdmTxArgs a = tx_args(add_noise ? noise_var : 0.0f, seed);
        check_hip(hipMemcpyAsync(s.info.get(), h_info, static_cast<size_t>(cfg.info_bits_per_frame()),
                                 hipMemcpyHostToDevice, tx_stream),
                  "copy TX info bits");
        a.info = s.info.get();
        a.coded = s.coded.get();
        a.out = s.tx_frame.get();
        sicc::launch_ofdm_tx(a, 1, tx_stream);
        check_hip(hipEventRecord(s.tx_done, tx_stream), "record tx_done");
        return s.tx_frame.get();
    }

    SyncResult sync_on_device(Slot& s, const float2* d_samples, int count)
    {
        if (count > max_rx) throw ArgumentError("RX sample_count exceeds slot allocation");
        if (count - cfg.preamble_len + 1 <= 0) throw ArgumentError("RX sample_count is shorter than the preamble");
        receiver().sync(d_samples, 1, count, s.sync.get(), 0, rx_stream);
        check_hip(hipMemcpyAsync(s.host_sync, s.sync.get(), sizeof(OfdmSyncResult), hipMemcpyDeviceToHost, rx_stream),
                  "copy sync result");
        check_hip(hipStreamSynchronize(rx_stream), "synchronize RX sync");
        SyncResult r;
        r.start = s.host_sync->start;
        r.metric = s.host_sync->metric;
        r.cfo_hz = s.host_sync->cfo_hz;
        return r;
    }

    SyncResult sync_host(int si, const Complex32* h_samples, int count)
    {
        Slot& s = slot(si);
        if (!h_samples) throw ArgumentError("RX host sample pointer is null");
        if (count > max_rx) throw ArgumentError("RX sample_count exceeds slot allocation");
        if (count > 0)
            check_hip(hipMemcpyAsync(s.rx_samples.get(), h_samples, static_cast<size_t>(count) * sizeof(float2),
                                     hipMemcpyHostToDevice, rx_stream),
                      "copy RX samples");
        return sync_on_device(s, s.rx_samples.get(), count);
    }

    const float* demod_host(int si, const Complex32* h_frame, SyncResult* out)
    {
        const SyncResult r = sync_host(si, h_frame, cfg.frame_samples());
        Slot& s = slot(si);
        receiver().demod(s.rx_samples.get(), 1, cfg.frame_samples(), s.sync.get(), s.llr.get(), rx_stream);
        check_hip(hipStreamSynchronize(rx_stream), "synchronize RX demod device LLR");
        if (out) *out = r;
        return s.llr.get();
    }

    const float* demod_device(int si, const Complex32* d_frame, const SyncResult& sr)
    {
        Slot& s = slot(si);
        if (!d_frame) throw ArgumentError("RX device frame pointer is null");
        s.host_sync->start = 0;
        s.host_sync->metric = sr.metric;
        s.host_sync->cfo_hz = sr.cfo_hz;
        s.host_sync->coarse_metric = sr.metric;
        s.host_sync->cfo_hz_d = static_cast<double>(sr.cfo_hz);
        check_hip(hipMemcpyAsync(s.sync.get(), s.host_sync, sizeof(OfdmSyncResult), hipMemcpyHostToDevice, rx_stream),
                  "copy device-frame sync result");
        receiver().demod(as_f2(d_frame), 1, cfg.frame_samples(), s.sync.get(), s.llr.get(), rx_stream);
        check_hip(hipStreamSynchronize(rx_stream), "synchronize RX device-frame demod LLR");
        return s.llr.get();
    }

    void tx_batch(const uint8_t* d_info, int frames, float noise_var, uint64_t seed, float2* d_out, hipStream_t st)
    {
        check_hip(hipSetDevice(device), "hipSetDevice");
        if (!d_info || !d_out || frames <= 0) throw ArgumentError("tx_batch: empty batch");
        sicc::OfdmTxArgs a = tx_args(noise_var, seed);
        const size_t need = static_cast<size_t>(frames) * cfg.coded_bits_per_frame();
        if (batch_coded.size() < need) {
            check_hip(hipStreamSynchronize(st ? st : tx_stream), "synchronize before staging growth");
            batch_coded.alloc(need);
        }
        a.info = d_info;
        a.coded = batch_coded.get();
        a.out = d_out;
        sicc::launch_ofdm_tx(a, frames, st ? st : tx_stream);
    }
};

std::mutex& registry_mutex()
{
    static std::mutex m;
    return m;
}

std::unordered_map<const Pipeline*, PipelineCore*>& registry()
{
    static std::unordered_map<const Pipeline*, PipelineCore*> r;
    return r;
}

PipelineCore& core_of(const Pipeline* p)
{
    std::lock_guard<std::mutex> lk(registry_mutex());
    auto it = registry().find(p);
    if (it == registry().end()) throw std::runtime_error("pipeline: unregistered object");
    return *it->second;
}

}  // namespace

struct Pipeline::Impl {
    Impl(const Config& cfg, int slots) : core(cfg, slots) {}
    PipelineCore core;
};

Pipeline::Pipeline(const Config& cfg, int slots) : impl_(std::make_unique<Impl>(cfg, slots))
{
    std::lock_guard<std::mutex> lk(registry_mutex());
    registry()[this] = &impl_->core;
}

Pipeline::~Pipeline()
{
    std::lock_guard<std::mutex> lk(registry_mutex());
    registry().erase(this);
}

void Pipeline::upload_static_tables(const StaticTables& tables) { impl_->core.upload_tables(tables); }

void Pipeline::submit_tx_frame_async(int slot, const uint8_t* h_info_bits, Complex32* h_tx_frame)
{
    if (!h_tx_frame) throw ArgumentError("TX host frame pointer is null");
    const float2* d = impl_->core.tx_device(slot, h_info_bits, 0.0f, 0, false);
    auto& s = impl_->core.slots[static_cast<size_t>(slot)];
    check_hip(hipMemcpyAsync(h_tx_frame, d, static_cast<size_t>(impl_->core.cfg.frame_samples()) * sizeof(float2),
                             hipMemcpyDeviceToHost, impl_->core.tx_stream),
              "copy TX frame to host");
    check_hip(hipEventRecord(s.tx_done, impl_->core.tx_stream), "record tx_done");
}

const Complex32* Pipeline::submit_tx_frame_device_async(int slot, const uint8_t* h_info_bits,
                                                        float frequency_noise_var, uint64_t noise_seed,
                                                        bool add_frequency_noise)
{
    return reinterpret_cast<const Complex32*>(
        impl_->core.tx_device(slot, h_info_bits, frequency_noise_var, noise_seed, add_frequency_noise));
}

void Pipeline::synchronize_tx(int slot)
{
    auto& s = impl_->core.slot(slot);
    check_hip(hipEventSynchronize(s.tx_done), "synchronize tx_done");
}

SyncResult Pipeline::find_frame(int slot, const Complex32* h_samples, int sample_count)
{
    return impl_->core.sync_host(slot, h_samples, sample_count);
}

SyncResult Pipeline::find_frame_device(int slot, const Complex32* d_samples, int sample_count)
{
    auto& s = impl_->core.slot(slot);
    if (!d_samples) throw ArgumentError("RX device sample pointer is null");
    return impl_->core.sync_on_device(s, as_f2(d_samples), sample_count);
}

const float* Pipeline::demod_frame_to_device_llr(int slot, const Complex32* h_frame, SyncResult* sync_result)
{
    return impl_->core.demod_host(slot, h_frame, sync_result);
}

const float* Pipeline::demod_frame_device_to_device_llr(int slot, const Complex32* d_frame,
                                                        const SyncResult& sync_result)
{
    return impl_->core.demod_device(slot, d_frame, sync_result);
}

SyncResult Pipeline::demod_frame_to_host_llr(int slot, const Complex32* h_frame, float* h_llr)
{
    if (!h_llr) throw ArgumentError("RX host LLR pointer is null");
    SyncResult r;
    const float* d = impl_->core.demod_host(slot, h_frame, &r);
    check_hip(hipMemcpyAsync(h_llr, d, static_cast<size_t>(impl_->core.cfg.coded_bits_per_frame()) * sizeof(float),
                             hipMemcpyDeviceToHost, impl_->core.rx_stream),
              "copy RX LLRs to host");
    check_hip(hipStreamSynchronize(impl_->core.rx_stream), "synchronize RX demod LLR copy");
    return r;
}

}  // namespace usrp_gpu_pipeline_bridge

// ------------------------------------------------------------- C layer ----
namespace {

using usrp_gpu_pipeline_bridge::Complex32;
using usrp_gpu_pipeline_bridge::Pipeline;

struct PipelineHandle {
    std::unique_ptr<Pipeline> p;
    usrp_gpu_pipeline_bridge::Config cfg;
};

sicc::HandleTable<PipelineHandle>& ptable()
{
    // never destroyed: handles still open at exit must not reach the HIP
    // runtime from static destructors
    static sicc::HandleTable<PipelineHandle>& t = *new sicc::HandleTable<PipelineHandle>;
    return t;
}

template <class F>
int pguard(F&& f)
{
    try {
        return f();
    } catch (const usrp_gpu_pipeline_bridge::ArgumentError& e) {
        sicc::set_last_error(e.what());
        return CUDA_OSD_STATUS_INVALID_ARGUMENT;
    } catch (const std::invalid_argument& e) {
        sicc::set_last_error(e.what());
        return CUDA_OSD_STATUS_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        sicc::set_last_error(e.what());
        return CUDA_OSD_STATUS_INTERNAL_ERROR;
    } catch (...) {
        sicc::set_last_error("unknown error");
        return CUDA_OSD_STATUS_INTERNAL_ERROR;
    }
}

std::shared_ptr<PipelineHandle> plookup(unsigned long long h)
{
    auto p = ptable().get(h);
    if (!p) throw std::invalid_argument("pipeline: unknown handle");
    return p;
}

template <class T>
std::vector<T> vec(const T* p, int count)
{
    if (count < 0 || (count > 0 && !p)) throw std::invalid_argument("pipeline tables: null table with a count");
    return count > 0 ? std::vector<T>(p, p + count) : std::vector<T>();
}

std::vector<Complex32> cvec(const float* iq, int count)
{
    if (count < 0 || (count > 0 && !iq)) throw std::invalid_argument("pipeline tables: null table with a count");
    std::vector<Complex32> v(static_cast<size_t>(std::max(count, 0)));
    for (int i = 0; i < count; ++i) v[i] = Complex32(iq[2 * i], iq[2 * i + 1]);
    return v;
}

void put(sicc_pipeline_sync_t* out, const usrp_gpu_pipeline_bridge::SyncResult& r)
{
    if (!out) return;
    out->start = r.start;
    out->metric = r.metric;
    out->cfo_hz = r.cfo_hz;
}

}  // namespace

extern "C" {

int sicc_pipeline_create(const sicc_pipeline_config_t* c, int slots, unsigned long long* out_handle)
{
    return pguard([&] {
        if (!c || !out_handle) throw std::invalid_argument("null argument");
        *out_handle = 0;
        usrp_gpu_pipeline_bridge::Config cfg;
        cfg.nfft = c->nfft;
        cfg.cp = c->cp;
        cfg.preamble_len = c->preamble_len;
        cfg.pre_half_len = c->pre_half_len;
        cfg.num_symbols = c->num_symbols;
        cfg.active_sc = c->active_sc;
        cfg.bits_per_symbol = c->bits_per_symbol;
        cfg.code_n = c->code_n;
        cfg.code_k = c->code_k;
        cfg.blocks_per_frame = c->blocks_per_frame;
        cfg.max_rx_search_samples = c->max_rx_search_samples;
        cfg.sample_rate = c->sample_rate;
        cfg.sync_threshold = c->sync_threshold;
        cfg.tx_amplitude = c->tx_amplitude;
        cfg.demod_noise_var = c->demod_noise_var;
        cfg.use_pilot_channel_est = c->use_pilot_channel_est != 0;
        auto h = std::make_unique<PipelineHandle>();
        h->cfg = cfg;
        h->p = std::make_unique<Pipeline>(cfg, slots);
        *out_handle = ptable().add(std::move(h));
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_pipeline_destroy(unsigned long long handle)
{
    return pguard([&] {
        if (!ptable().remove(handle)) throw std::invalid_argument("pipeline: unknown handle");
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_pipeline_upload_static_tables(unsigned long long handle, const sicc_pipeline_tables_t* t)
{
    return pguard([&] {
        auto h = plookup(handle);
        if (!t) throw std::invalid_argument("null tables");
        usrp_gpu_pipeline_bridge::StaticTables st;
        st.generator_kn = vec(t->generator_kn, t->generator_kn_count);
        st.dvb_s2_accumulator = t->dvb_s2_accumulator != 0;
        st.row_info_offsets = vec(t->row_info_offsets, t->row_info_offsets_count);
        st.row_info_cols = vec(t->row_info_cols, t->row_info_cols_count);
        st.used_indices = vec(t->used_indices, t->used_indices_count);
        st.pilot_symbol_index = vec(t->pilot_symbol_index, t->pilot_symbol_index_count);
        st.data_symbol_index = vec(t->data_symbol_index, t->data_symbol_index_count);
        st.pilot_symbols = vec(t->pilot_symbols, t->pilot_symbols_count);
        st.data_symbols = vec(t->data_symbols, t->data_symbols_count);
        st.interp_p0 = vec(t->interp_p0, t->interp_p0_count);
        st.interp_p1 = vec(t->interp_p1, t->interp_p1_count);
        st.interp_alpha = vec(t->interp_alpha, t->interp_alpha_count);
        st.pilot_grid = cvec(t->pilot_grid_iq, t->pilot_grid_count);
        st.preamble = cvec(t->preamble_iq, t->preamble_count);
        st.precomp_h = cvec(t->precomp_h_iq, t->precomp_h_count);
        st.actual_h = cvec(t->actual_h_iq, t->actual_h_count);
        h->p->upload_static_tables(st);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_pipeline_submit_tx_frame(unsigned long long handle, int slot, const unsigned char* info_bits,
                                  float* tx_frame_iq)
{
    return pguard([&] {
        auto h = plookup(handle);
        h->p->submit_tx_frame_async(slot, info_bits, reinterpret_cast<Complex32*>(tx_frame_iq));
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_pipeline_submit_tx_frame_device(unsigned long long handle, int slot, const unsigned char* info_bits,
                                         float frequency_noise_var, unsigned long long noise_seed,
                                         int add_frequency_noise, const float** d_tx_frame_iq)
{
    return pguard([&] {
        auto h = plookup(handle);
        const Complex32* d = h->p->submit_tx_frame_device_async(slot, info_bits, frequency_noise_var, noise_seed,
                                                                add_frequency_noise != 0);
        if (d_tx_frame_iq) *d_tx_frame_iq = reinterpret_cast<const float*>(d);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_pipeline_synchronize_tx(unsigned long long handle, int slot)
{
    return pguard([&] {
        plookup(handle)->p->synchronize_tx(slot);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_pipeline_find_frame(unsigned long long handle, int slot, const float* samples_iq, int sample_count,
                             sicc_pipeline_sync_t* out)
{
    return pguard([&] {
        auto h = plookup(handle);
        put(out, h->p->find_frame(slot, reinterpret_cast<const Complex32*>(samples_iq), sample_count));
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_pipeline_find_frame_device(unsigned long long handle, int slot, const float* d_samples_iq,
                                    int sample_count, sicc_pipeline_sync_t* out)
{
    return pguard([&] {
        auto h = plookup(handle);
        put(out, h->p->find_frame_device(slot, reinterpret_cast<const Complex32*>(d_samples_iq), sample_count));
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_pipeline_demod_frame_to_device_llr(unsigned long long handle, int slot, const float* frame_iq,
                                            sicc_pipeline_sync_t* sync_out, const float** d_llr)
{
    return pguard([&] {
        auto h = plookup(handle);
        usrp_gpu_pipeline_bridge::SyncResult r;
        const float* d = h->p->demod_frame_to_device_llr(slot, reinterpret_cast<const Complex32*>(frame_iq), &r);
        put(sync_out, r);
        if (d_llr) *d_llr = d;
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_pipeline_demod_frame_device_to_device_llr(unsigned long long handle, int slot, const float* d_frame_iq,
                                                   const sicc_pipeline_sync_t* sync, const float** d_llr)
{
    return pguard([&] {
        auto h = plookup(handle);
        if (!sync) throw std::invalid_argument("null sync result");
        usrp_gpu_pipeline_bridge::SyncResult r;
        r.start = sync->start;
        r.metric = sync->metric;
        r.cfo_hz = sync->cfo_hz;
        const float* d =
            h->p->demod_frame_device_to_device_llr(slot, reinterpret_cast<const Complex32*>(d_frame_iq), r);
        if (d_llr) *d_llr = d;
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_pipeline_demod_frame_to_host_llr(unsigned long long handle, int slot, const float* frame_iq, float* llr,
                                          sicc_pipeline_sync_t* sync_out)
{
    return pguard([&] {
        auto h = plookup(handle);
        put(sync_out, h->p->demod_frame_to_host_llr(slot, reinterpret_cast<const Complex32*>(frame_iq), llr));
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_pipeline_tx_batch_device(unsigned long long handle, const unsigned char* d_info_bits, int frames,
                                  float frequency_noise_var, unsigned long long noise_seed, float* d_out_iq,
                                  void* hip_stream)
{
    return pguard([&] {
        auto h = plookup(handle);
        usrp_gpu_pipeline_bridge::core_of(h->p.get())
            .tx_batch(d_info_bits, frames, frequency_noise_var, noise_seed, reinterpret_cast<float2*>(d_out_iq),
                      static_cast<hipStream_t>(hip_stream));
        return CUDA_OSD_STATUS_OK;
    });
}

}  // extern "C"
