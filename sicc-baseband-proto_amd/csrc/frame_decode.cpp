// frame_decode.cpp -- frame-level receive and decode (include/sicc_frame_c_api.h):
// sync, demodulation, BP(-OSD) over every block of a batch of frames and the
// per-frame FrameDecodeResult, all stream-ordered on the GPU.
//
// Reference (LINK = source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp):
// decode_frame LINK:5331-5603 and finish_decode_from_llr LINK:6615-6705 for
// one frame at a time; here one call covers a batch: one sync launch, one
// demodulator launch pair, one BP launch (+ one OSD launch) over every block
// of every frame, and one aggregation launch.
#include "../../include/sicc_frame_c_api.h"

#include "decoder_handles.hpp"
#include "frame.hpp"
#include "ofdm_rx.hpp"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <memory>
#include <string>

namespace sicc {

namespace {

static_assert(sizeof(sicc_frame_result_t) == 80, "frame result layout");
static_assert(offsetof(sicc_frame_result_t, peak) == 48, "frame result layout");

struct FrameDec {
    int device = 0;
    std::shared_ptr<OfdmRx> rx;    // null: LLR input only
    std::shared_ptr<Decoder> dec;  // stream BP-OSD or BP CSR decoder
    int blocks = 0;
    int n = 0;
    int k = 0;
    double gate = -1e300;
    hipStream_t stream = nullptr;
    // generator columns for the pre-FEC count, word-major (built on first use)
    DevBuf<unsigned long long> gen;
    int kw = 0;
    int c0 = 0;
    bool gen_ready = false;
    bool gen_lds = false;
    // per-batch scratch
    int cap = 0;
    DevBuf<OfdmSyncResult> sync;
    DevBuf<float> llr;
    DevBuf<uint8_t> bp_ok, used, blk_ok;
    DevBuf<int> iters;
    DevBuf<double> stats;
    DecodeScratch dscratch;
    // host-call staging
    DevBuf<float2> h_rx;
    DevBuf<float> h_llr;
    DevBuf<uint8_t> h_ref, h_info;
    DevBuf<sicc_frame_result_t> h_res;

    ~FrameDec()
    {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }

    bool osd() const { return dec->kind == Kind::Stream && dec->osd_order >= 0; }

    void reserve(int frames, bool with_llr)
    {
        if (frames <= 0) throw std::invalid_argument("frame decode: empty batch");
        const size_t cw = static_cast<size_t>(frames) * blocks;
        if (cw > 0x7fffffffull) throw std::invalid_argument("frame decode: batch too large");
        if (frames > cap) {
            // a kernel of an earlier call on this stream may still read the old buffers
            SICC_HIP(hipStreamSynchronize(stream));
            const int c = std::max(frames, cap + cap / 2);
            const size_t ccw = static_cast<size_t>(c) * blocks;
            bp_ok.alloc(ccw);
            used.alloc(ccw);
            blk_ok.alloc(ccw);
            iters.alloc(ccw);
            if (rx) {
                sync.alloc(c);
                stats.alloc(static_cast<size_t>(c) * rx->n_data() * 4);
            }
            if (llr.size() > 0) llr.alloc(ccw * n);
            cap = c;
        }
        if (with_llr && llr.size() < static_cast<size_t>(cap) * blocks * n) {
            SICC_HIP(hipStreamSynchronize(stream));
            llr.alloc(static_cast<size_t>(cap) * blocks * n);
        }
        // the BP -> OSD hand-off is sized only for decoders that run OSD
        dec->eng->reserve(osd() ? static_cast<int>(cw) : 0, &dec->bp, &dscratch);
    }

    void ensure_generator()
    {
        if (gen_ready) return;
        const LdpcCode& c = dec->eng->code();
        kw = (c.k + 63) / 64;
        c0 = c.front_systematic() ? c.k : 0;
        const size_t words = static_cast<size_t>(c.n - c0) * kw;
        if (words * 8 > (256ull << 20)) throw std::invalid_argument("frame decode: generator too large for the pre-FEC count");
        const auto g = c.dense_generator();
        std::vector<unsigned long long> cols(std::max<size_t>(words, 1), 0ull);
        for (int r = 0; r < c.k; ++r)
            for (int col = c0; col < c.n; ++col)
                if (g[static_cast<size_t>(r) * c.n + col])
                    cols[static_cast<size_t>(r / 64) * (c.n - c0) + (col - c0)] |= 1ull << (r % 64);
        gen.alloc(cols.size());
        SICC_HIP(hipMemcpy(gen.get(), cols.data(), cols.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
        gen_lds = frame_gen_fits_lds(c.n - c0, kw);
        gen_ready = true;
    }

    // BP(-OSD) over frames * blocks codewords, then the per-frame results.
    void decode_and_finish(const float* d_llr, int frames, const OfdmSyncResult* d_sync, int span,
                           const double* d_stats, const uint8_t* d_ref, uint8_t* d_info, sicc_frame_result_t* d_res,
                           uint8_t* d_block_ok, int* d_block_it, hipStream_t s)
    {
        if (!d_info || !d_res) throw std::invalid_argument("frame decode: null information-bit or result output");
        if (d_ref) ensure_generator();
        const int cw = frames * blocks;
        DecodeOut out;
        out.info = d_info;
        out.bp_ok = bp_ok.get();
        out.iterations = d_block_it ? d_block_it : iters.get();
        out.used_osd = osd() ? used.get() : nullptr;
        dec->eng->decode(d_llr, cw, dec->bp, osd() ? dec->osd_order : -1, dec->osd, out, s, &dscratch);
        FrameFinishArgs a;
        a.frames = frames;
        a.blocks = blocks;
        a.n = n;
        a.k = k;
        a.bp_ok = bp_ok.get();
        a.used_osd = out.used_osd;
        a.iterations = out.iterations;
        a.info = d_info;
        a.ref_info = d_ref;
        a.llr = d_llr;
        if (d_ref) {
            a.gen_cols = gen.get();
            a.kw = kw;
            a.c0 = c0;
            a.gen_in_lds = gen_lds ? 1 : 0;
        }
        a.sync = d_sync;
        a.span = span;
        if (rx) {
            a.frame_samples = rx->frame_samples();
            a.n_data = rx->n_data();
            a.n_re = static_cast<double>(rx->n_data()) * rx->config().active_sc;
            a.noise_var = std::max(static_cast<double>(rx->config().demod_noise_var), 1e-8);
        }
        a.stats = d_stats;
        a.snr_gate_db = gate;
        a.block_ok = d_block_ok;
        a.out = d_res;
        launch_frame_finish(a, s);
    }

    void receive(const float2* d_rx, int frames, int span, OfdmSyncResult* d_sync, bool run_sync, const uint8_t* d_ref,
                 uint8_t* d_info, sicc_frame_result_t* d_res, float* d_llr, uint8_t* d_block_ok, int* d_block_it,
                 hipStream_t s)
    {
        if (!rx) throw std::invalid_argument("frame decode: this frame decoder has no OFDM receiver");
        if (!d_rx || span <= 0) throw std::invalid_argument("frame decode: empty batch");
        reserve(frames, d_llr == nullptr);
        OfdmSyncResult* sy = d_sync ? d_sync : sync.get();
        float* ll = d_llr ? d_llr : llr.get();
        if (run_sync) rx->sync(d_rx, frames, span, sy, rx->config().max_rx_search_samples, s);
        rx->demod(d_rx, frames, span, sy, ll, s, stats.get());
        decode_and_finish(ll, frames, sy, span, stats.get(), d_ref, d_info, d_res, d_block_ok, d_block_it, s);
    }
};

HandleTable<FrameDec>& ftable()
{
    // never destroyed: handles still open at exit must not reach the HIP
    // runtime from static destructors
    static HandleTable<FrameDec>& t = *new HandleTable<FrameDec>;
    return t;
}

std::shared_ptr<FrameDec> flookup(unsigned long long h)
{
    auto p = ftable().get(h);
    if (!p) throw std::invalid_argument("frame decoder: unknown handle");
    SICC_HIP(hipSetDevice(p->device));
    return p;
}

hipStream_t pick(const FrameDec& f, void* s) { return s ? static_cast<hipStream_t>(s) : f.stream; }

}  // namespace
}  // namespace sicc

using namespace sicc;

extern "C" {

int sicc_frame_decoder_create(unsigned long long ofdm_receiver_handle, unsigned long long ldpc_decoder,
                              const sicc_frame_config_t* config, unsigned long long* out_handle)
{
    return guarded([&] {
        if (!out_handle || !config) throw std::invalid_argument("null argument");
        *out_handle = 0;
        auto f = std::make_unique<FrameDec>();
        f->dec = lookup(ldpc_decoder, "frame_decoder_create");
        if (f->dec->kind == Kind::OsdOnly) throw std::invalid_argument("frame decoder: needs a BP or BP-OSD decoder");
        SICC_HIP(hipGetDevice(&f->device));
        f->n = f->dec->eng->code().n;
        f->k = f->dec->eng->code().k;
        if (ofdm_receiver_handle) {
            f->rx = ofdm_receiver(ofdm_receiver_handle);
            const auto& rc = f->rx->config();
            if (rc.code_n != f->n)
                throw std::invalid_argument("frame decoder: receiver code_n " + std::to_string(rc.code_n) +
                                            " differs from the decoder's n " + std::to_string(f->n));
            if (config->blocks_per_frame > 0 && config->blocks_per_frame != rc.blocks_per_frame)
                throw std::invalid_argument("frame decoder: blocks_per_frame differs from the receiver's");
            f->blocks = rc.blocks_per_frame;
        } else {
            if (config->blocks_per_frame <= 0) throw std::invalid_argument("frame decoder: blocks_per_frame must be > 0");
            f->blocks = config->blocks_per_frame;
        }
        f->gate = config->snr_gate_db > -1e9 ? config->snr_gate_db : -1e300;
        SICC_HIP(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
        *out_handle = ftable().add(std::move(f));
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_frame_decoder_destroy(unsigned long long handle)
{
    return guarded([&] {
        if (!ftable().remove(handle)) throw std::invalid_argument("frame decoder: unknown handle");
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_frame_decoder_reserve(unsigned long long handle, int frames)
{
    return guarded([&] {
        auto f = flookup(handle);
        f->reserve(frames, f->rx != nullptr);
        SICC_HIP(hipStreamSynchronize(f->stream));
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_frame_decoder_info(unsigned long long handle, int* blocks_per_frame, int* code_n, int* code_k,
                            int* info_bits_per_frame)
{
    return guarded([&] {
        auto f = flookup(handle);
        if (blocks_per_frame) *blocks_per_frame = f->blocks;
        if (code_n) *code_n = f->n;
        if (code_k) *code_k = f->k;
        if (info_bits_per_frame) *info_bits_per_frame = f->blocks * f->k;
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_frame_receive_decode_device(unsigned long long handle, const float* d_rx_iq, int frames, int span,
                                     const unsigned char* d_ref_info, unsigned char* d_info_bits,
                                     sicc_frame_result_t* d_results, sicc_ofdm_sync_t* d_sync, float* d_llr,
                                     unsigned char* d_block_ok, int* d_block_iterations, void* hip_stream)
{
    return guarded([&] {
        auto f = flookup(handle);
        f->receive(reinterpret_cast<const float2*>(d_rx_iq), frames, span, reinterpret_cast<OfdmSyncResult*>(d_sync),
                   true, d_ref_info, d_info_bits, d_results, d_llr, d_block_ok, d_block_iterations,
                   pick(*f, hip_stream));
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_frame_demod_decode_device(unsigned long long handle, const float* d_rx_iq, int frames, int span,
                                   const sicc_ofdm_sync_t* d_sync, const unsigned char* d_ref_info,
                                   unsigned char* d_info_bits, sicc_frame_result_t* d_results, float* d_llr,
                                   unsigned char* d_block_ok, int* d_block_iterations, void* hip_stream)
{
    return guarded([&] {
        auto f = flookup(handle);
        if (!d_sync) throw std::invalid_argument("frame decode: null sync results");
        f->receive(reinterpret_cast<const float2*>(d_rx_iq), frames, span,
                   const_cast<OfdmSyncResult*>(reinterpret_cast<const OfdmSyncResult*>(d_sync)), false, d_ref_info,
                   d_info_bits, d_results, d_llr, d_block_ok, d_block_iterations, pick(*f, hip_stream));
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_frame_decode_llr_device(unsigned long long handle, const float* d_llr, int frames,
                                 const unsigned char* d_ref_info, unsigned char* d_info_bits,
                                 sicc_frame_result_t* d_results, unsigned char* d_block_ok, int* d_block_iterations,
                                 void* hip_stream)
{
    return guarded([&] {
        auto f = flookup(handle);
        if (!d_llr) throw std::invalid_argument("frame decode: null LLRs");
        f->reserve(frames, false);
        f->decode_and_finish(d_llr, frames, nullptr, 0, nullptr, d_ref_info, d_info_bits, d_results, d_block_ok,
                             d_block_iterations, pick(*f, hip_stream));
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_frame_receive_decode_host(unsigned long long handle, const float* rx_iq, int frames, int span,
                                   const unsigned char* ref_info, unsigned char* info_bits,
                                   sicc_frame_result_t* results)
{
    return guarded([&] {
        auto f = flookup(handle);
        if (!rx_iq || frames <= 0 || span <= 0 || !info_bits || !results)
            throw std::invalid_argument("frame decode: empty batch or null output");
        const size_t ns = static_cast<size_t>(frames) * span;
        const size_t ni = static_cast<size_t>(frames) * f->blocks * f->k;
        SICC_HIP(hipStreamSynchronize(f->stream));
        if (f->h_rx.size() < ns) f->h_rx.alloc(ns);
        if (f->h_info.size() < ni) f->h_info.alloc(ni);
        if (ref_info && f->h_ref.size() < ni) f->h_ref.alloc(ni);
        if (f->h_res.size() < static_cast<size_t>(frames)) f->h_res.alloc(frames);
        SICC_HIP(hipMemcpyAsync(f->h_rx.get(), rx_iq, ns * sizeof(float2), hipMemcpyHostToDevice, f->stream));
        if (ref_info) SICC_HIP(hipMemcpyAsync(f->h_ref.get(), ref_info, ni, hipMemcpyHostToDevice, f->stream));
        f->receive(f->h_rx.get(), frames, span, nullptr, true, ref_info ? f->h_ref.get() : nullptr, f->h_info.get(),
                   f->h_res.get(), nullptr, nullptr, nullptr, f->stream);
        SICC_HIP(hipMemcpyAsync(info_bits, f->h_info.get(), ni, hipMemcpyDeviceToHost, f->stream));
        SICC_HIP(hipMemcpyAsync(results, f->h_res.get(), frames * sizeof(sicc_frame_result_t), hipMemcpyDeviceToHost,
                                f->stream));
        SICC_HIP(hipStreamSynchronize(f->stream));
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_frame_decode_llr_host(unsigned long long handle, const float* llr, int frames, const unsigned char* ref_info,
                               unsigned char* info_bits, sicc_frame_result_t* results)
{
    return guarded([&] {
        auto f = flookup(handle);
        if (!llr || frames <= 0 || !info_bits || !results)
            throw std::invalid_argument("frame decode: empty batch or null output");
        const size_t nl = static_cast<size_t>(frames) * f->blocks * f->n;
        const size_t ni = static_cast<size_t>(frames) * f->blocks * f->k;
        SICC_HIP(hipStreamSynchronize(f->stream));
        if (f->h_llr.size() < nl) f->h_llr.alloc(nl);
        if (f->h_info.size() < ni) f->h_info.alloc(ni);
        if (ref_info && f->h_ref.size() < ni) f->h_ref.alloc(ni);
        if (f->h_res.size() < static_cast<size_t>(frames)) f->h_res.alloc(frames);
        SICC_HIP(hipMemcpyAsync(f->h_llr.get(), llr, nl * sizeof(float), hipMemcpyHostToDevice, f->stream));
        if (ref_info) SICC_HIP(hipMemcpyAsync(f->h_ref.get(), ref_info, ni, hipMemcpyHostToDevice, f->stream));
        f->reserve(frames, false);
        f->decode_and_finish(f->h_llr.get(), frames, nullptr, 0, nullptr, ref_info ? f->h_ref.get() : nullptr,
                             f->h_info.get(), f->h_res.get(), nullptr, nullptr, f->stream);
        SICC_HIP(hipMemcpyAsync(info_bits, f->h_info.get(), ni, hipMemcpyDeviceToHost, f->stream));
        SICC_HIP(hipMemcpyAsync(results, f->h_res.get(), frames * sizeof(sicc_frame_result_t), hipMemcpyDeviceToHost,
                                f->stream));
        SICC_HIP(hipStreamSynchronize(f->stream));
        return CUDA_OSD_STATUS_OK;
    });
}

}  // extern "C"
