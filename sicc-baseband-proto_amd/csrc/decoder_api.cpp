// cuda_osd_api.cpp -- the BP-OSD C ABI (include/matlab_decoder_c_api.h and
// include/sicc_decode_ext.h) over the HIP decode engine.
//
// Handle kinds: stream BP-OSD decoders (dynamic batching, submit / flush /
// poll, OSDH:17-32, OSDH:96-121), OSD-only decoders (OSDH:75-95) and BP-only
// CSR decoders (OSDH:78-88).  All share one handle table.
#include "../../include/matlab_decoder_c_api.h"
#include "../../include/sicc_decode_ext.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <vector>

#include "decoder_handles.hpp"

namespace sicc {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

namespace {

}  // namespace

HandleTable<Decoder>& table()
{
    // never destroyed: handles still open at exit must not reach the HIP
    // runtime from static destructors
    static HandleTable<Decoder>& t = *new HandleTable<Decoder>;
    return t;
}

std::shared_ptr<Decoder> lookup(unsigned long long h, const char* who)
{
    auto d = table().get(h);
    if (!d) throw std::invalid_argument(std::string(who) + ": unknown handle");
    SICC_HIP(hipSetDevice(d->device));
    return d;
}

namespace {

void init_common(Decoder& d)
{
    SICC_HIP(hipGetDevice(&d.device));
    SICC_HIP(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
    SICC_HIP(hipEventCreate(&d.ev0));
    SICC_HIP(hipEventCreate(&d.ev1));
}

LdpcCode bundled_ccsds(int n, int k)
{
    if (n == 128 && k == 64) return load_alist(bundled_data_dir() + "/CCSDS_ldpc_n128_k64.alist", true);
    if (n == 256 && k == 128) return load_alist(bundled_data_dir() + "/CCSDS_ldpc_n256_k128.alist", true);
    if (n == 512 && k == 256) return load_alist(bundled_data_dir() + "/CCSDS_ldpc_n512_k256.alist", true);
    throw std::invalid_argument("no bundled code for n=" + std::to_string(n) + " k=" + std::to_string(k));
}

void validate_stream_cfg(const cuda_osd_stream_config_t* c)
{
    if (!c) throw std::invalid_argument("null config");
    if (c->max_batch_size <= 0 || c->min_batch_size < 0 || c->min_batch_size > c->max_batch_size)
        throw std::invalid_argument("invalid min/max batch size");
    if (c->bp_max_iterations < 1) throw std::invalid_argument("bp_max_iterations must be >= 1");
    if (!(c->bp_normalization > 0.0f)) throw std::invalid_argument("bp_normalization must be > 0");
    if (c->bp_offset < 0.0f || c->bp_damping < 0.0f || c->bp_damping >= 1.0f)
        throw std::invalid_argument("bp_offset must be >= 0 and bp_damping in [0, 1)");
}

unsigned long long make_stream(const cuda_osd_stream_config_t* cfg, LdpcCode code)
{
    validate_stream_cfg(cfg);
    if (code.n != cfg->n || code.k != cfg->k)
        throw std::invalid_argument("config n/k do not match the code (n=" + std::to_string(code.n) +
                                    " k=" + std::to_string(code.k) + ")");
    if (!code.front_systematic())
        throw std::invalid_argument("BP-OSD requires a systematic-front code layout (information columns first)");
    auto d = std::make_unique<Decoder>();
    d->kind = Kind::Stream;
    init_common(*d);
    d->eng = std::make_unique<Engine>(std::move(code));
    if (!d->eng->osd_supported()) throw std::invalid_argument("OSD supports k <= 128 and n <= 256");
    d->cfg = *cfg;
    d->bp.max_iter = cfg->bp_max_iterations;
    d->bp.alpha = cfg->bp_normalization;
    d->bp.offset = cfg->bp_offset;
    d->bp.damping = cfg->bp_damping;
    d->bp.min_abs_accept = cfg->bp_min_abs_llr_accept;
    d->osd_order = 2;
    d->osd.max_flip = 32;
    d->osd.max_tests = 4096;
    d->max_batch = std::max(cfg->max_batch_size, 1);
    d->min_batch = std::max(cfg->min_batch_size, 1);
    const int n = d->eng->code().n, k = d->eng->code().k;
    for (int i = 0; i < 4; ++i) {
        auto b = std::make_unique<Batch>();
        b->llr.alloc(static_cast<size_t>(d->max_batch) * n);
        b->info.alloc(static_cast<size_t>(d->max_batch) * k);
        b->used.alloc(d->max_batch);
        b->ok.alloc(d->max_batch);
        b->h_info.alloc(static_cast<size_t>(d->max_batch) * k);
        b->h_used.alloc(d->max_batch);
        b->h_ok.alloc(d->max_batch);
        SICC_HIP(hipEventCreate(&b->start));
        SICC_HIP(hipEventCreate(&b->done));
        d->slots.push_back(std::move(b));
    }
    d->st.ewma_interarrival_us = static_cast<double>(cfg->default_interarrival_us);
    return table().add(std::move(d));
}

// ---- stream decoder internals ----

void harvest(Decoder& d, int si)
{
    Batch& b = *d.slots[si];
    const int k = d.eng->code().k;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, b.start, b.done) == hipSuccess) d.st.total_decode_ms += ms;
    for (int i = 0; i < b.count; ++i) {
        ReadyItem it;
        it.ts = b.ts[i];
        it.id = b.ids[i];
        it.blk = b.blk[i];
        it.used = b.h_used.get()[i];
        it.info.assign(b.h_info.get() + static_cast<size_t>(i) * k, b.h_info.get() + static_cast<size_t>(i + 1) * k);
        d.st.bp_success_codewords += b.h_ok.get()[i] ? 1 : 0;
        d.st.osd_fallback_codewords += it.used ? 1 : 0;
        d.ready.push_back(std::move(it));
    }
    d.st.decoded_codewords += static_cast<unsigned long long>(b.count);
    d.st.decoded_batches += 1;
    d.st.average_batch_size = static_cast<double>(d.st.decoded_codewords) / static_cast<double>(d.st.decoded_batches);
    b.count = 0;
    b.ts.clear();
    b.ids.clear();
    b.blk.clear();
    b.inflight = false;
}

void harvest_completed(Decoder& d, bool wait_oldest)
{
    while (!d.inflight.empty()) {
        const int si = d.inflight.front();
        hipError_t q = hipEventQuery(d.slots[si]->done);
        if (q == hipErrorNotReady) {
            if (!wait_oldest) return;
            SICC_HIP(hipEventSynchronize(d.slots[si]->done));
            wait_oldest = false;
        } else if (q != hipSuccess) {
            throw_hip(q, "decode batch");
        }
        d.inflight.pop_front();
        harvest(d, si);
    }
}

void launch_filling(Decoder& d)
{
    if (d.filling < 0) return;
    Batch& b = *d.slots[d.filling];
    const int n = d.eng->code().n, k = d.eng->code().k;
    if (b.count > 0) {
        DecodeOut out;
        out.info = b.info.get();
        out.used_osd = b.used.get();
        out.bp_ok = b.ok.get();
        SICC_HIP(hipEventRecord(b.start, d.stream));
        d.eng->decode(b.llr.get(), b.count, d.bp, d.osd_order, d.osd, out, d.stream);
        SICC_HIP(hipMemcpyAsync(b.h_info.get(), b.info.get(), static_cast<size_t>(b.count) * k, hipMemcpyDeviceToHost,
                                d.stream));
        SICC_HIP(hipMemcpyAsync(b.h_used.get(), b.used.get(), b.count, hipMemcpyDeviceToHost, d.stream));
        SICC_HIP(hipMemcpyAsync(b.h_ok.get(), b.ok.get(), b.count, hipMemcpyDeviceToHost, d.stream));
        SICC_HIP(hipEventRecord(b.done, d.stream));
        b.inflight = true;
        d.inflight.push_back(d.filling);
    }
    (void)n;
    d.filling = -1;
}

int acquire_slot(Decoder& d)
{
    for (;;) {
        for (int i = 0; i < static_cast<int>(d.slots.size()); ++i)
            if (!d.slots[i]->inflight && i != d.filling) return i;
        harvest_completed(d, true);
    }
}

void maybe_launch_on_latency(Decoder& d)
{
    if (d.filling < 0) return;
    Batch& b = *d.slots[d.filling];
    if (b.count >= d.min_batch && now_us() - b.first_us >= d.cfg.max_latency_us) launch_filling(d);
}

void submit(Decoder& d, const float* llrs, int count, bool device_ptr, const unsigned long long* ts,
            const unsigned long long* ids, const unsigned int* blk)
{
    if (d.kind != Kind::Stream) throw std::invalid_argument("submit needs a stream decoder handle");
    if (!llrs || count <= 0) throw std::invalid_argument("submit received no codewords");
    const int n = d.eng->code().n;
    const uint64_t t = now_us();
    if (d.last_arrival_us != 0) {
        const double dt = static_cast<double>(t - d.last_arrival_us);
        const double a = d.cfg.arrival_ewma_alpha > 0.0 ? d.cfg.arrival_ewma_alpha : 0.15;
        d.st.ewma_interarrival_us = (1.0 - a) * d.st.ewma_interarrival_us + a * dt;
    }
    d.last_arrival_us = t;
    d.st.last_submit_timestamp_us = ts ? ts[count - 1] : t;
    int done = 0;
    while (done < count) {
        if (d.filling < 0) {
            d.filling = acquire_slot(d);
            d.slots[d.filling]->first_us = t;
        }
        Batch& b = *d.slots[d.filling];
        const int take = std::min(count - done, d.max_batch - b.count);
        SICC_HIP(hipMemcpyAsync(b.llr.get() + static_cast<size_t>(b.count) * n, llrs + static_cast<size_t>(done) * n,
                                static_cast<size_t>(take) * n * sizeof(float),
                                device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, d.stream));
        for (int i = 0; i < take; ++i) {
            b.ts.push_back(ts ? ts[done + i] : 0ull);
            b.ids.push_back(ids ? ids[done + i] : 0ull);
            b.blk.push_back(blk ? blk[done + i] : static_cast<unsigned int>(done + i));
        }
        b.count += take;
        done += take;
        if (b.count >= d.max_batch) launch_filling(d);
    }
    if (!device_ptr) SICC_HIP(hipStreamSynchronize(d.stream));  // caller may reuse its host buffer
    d.st.submitted_codewords += static_cast<unsigned long long>(count);
    d.flushed = false;
    maybe_launch_on_latency(d);
    const unsigned long long pending = d.st.submitted_codewords - d.polled;
    d.st.current_pending_codewords = pending;
    d.st.max_pending_codewords = std::max(d.st.max_pending_codewords, pending);
}

void ensure_scratch(Decoder& d, int count, bool app)
{
    const int n = d.eng->code().n, k = d.eng->code().k;
    const size_t need = static_cast<size_t>(count);
    if (d.s_llr.size() < need * n) {
        d.s_llr.alloc(need * n);
        d.s_a.alloc(need * std::max(n, k));
        d.s_b.alloc(need * n);
        d.s_c.alloc(need);
        d.s_d.alloc(need);
        d.s_it.alloc(need);
        d.s_dist.alloc(need);
        d.s_app.release();
    }
    if (app && d.s_app.size() < need * n) d.s_app.alloc(need * n);
}

// ---- mixed-code batches ----

struct MultiDec {
    int device = 0;
    std::vector<std::shared_ptr<Decoder>> members;
    hipStream_t main = nullptr;
    std::vector<hipStream_t> lanes;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<sicc_group_t> key;
    std::vector<unsigned char> key_params;  // every group's member parameters the graph bakes in
    // Per-group scratch (BP -> OSD hand-off, counters, long-code messages): groups
    // run concurrently, also several groups of one member, so none shares a
    // member's own scratch.  The graph's set is sized before capture and only
    // replaced together with the graph; eager calls use their own set.
    std::vector<std::unique_ptr<DecodeScratch>> graph_scratch, eager_scratch;

    ~MultiDec()
    {
        if (main) (void)hipStreamSynchronize(main);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto s : lanes) (void)hipStreamDestroy(s);
        for (auto e : joins) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (main) (void)hipStreamDestroy(main);
    }
};

HandleTable<MultiDec>& mtable()
{
    // never destroyed: handles still open at exit must not reach the HIP
    // runtime from static destructors
    static HandleTable<MultiDec>& t = *new HandleTable<MultiDec>;
    return t;
}

bool same_groups(const std::vector<sicc_group_t>& a, const sicc_group_t* b, int n)
{
    if (static_cast<int>(a.size()) != n) return false;
    for (int i = 0; i < n; ++i)
        if (std::memcmp(&a[i], &b[i], sizeof(sicc_group_t)) != 0) return false;
    return true;
}

// Bytes of everything a captured decode of these groups depends on besides the
// group descriptors: each member's BP parameters, OSD order and parameters and
// BP launch shape.
std::vector<unsigned char> group_params(const MultiDec& md, const sicc_group_t* groups, int n)
{
    std::vector<unsigned char> k;
    auto put = [&k](const void* p, size_t b) {
        const auto* c = static_cast<const unsigned char*>(p);
        k.insert(k.end(), c, c + b);
    };
    for (int i = 0; i < n; ++i) {
        const Decoder& d = *md.members[groups[i].member];
        const BpTune& t = d.eng->tune();
        put(&d.bp, sizeof(d.bp));
        put(&d.osd_order, sizeof(d.osd_order));
        put(&d.osd, sizeof(d.osd));
        put(&t, sizeof(t));
    }
    return k;
}

// Enqueue the groups on the lane streams, forked from and joined back to `s`;
// group i decodes with scratch[i].  BP-only groups whose batches share one
// instance of the flooding kernel decode in one fused launch (their slowest
// codewords overlap instead of queueing behind each other); every other group
// is a launch of its own.  Launches are dealt round-robin over the lanes in
// the order of their first group.
void enqueue_groups(MultiDec& md, const sicc_group_t* groups, int n, hipStream_t s,
                    std::vector<std::unique_ptr<DecodeScratch>>& scratch)
{
    while (static_cast<int>(scratch.size()) < n) scratch.push_back(std::make_unique<DecodeScratch>());
    std::vector<DecodeOut> outs(n);
    std::vector<BpBatch> batch(n);
    std::vector<int> key(n, -1);
    for (int i = 0; i < n; ++i) {
        const sicc_group_t& g = groups[i];
        Decoder& d = *md.members[g.member];
        DecodeOut& out = outs[i];
        out.info = g.d_info;
        out.bp_ok = g.d_bp_ok;
        out.iterations = g.d_iterations;
        out.used_osd = g.d_used_osd;
        const bool bp_only = d.kind != Kind::Stream || d.osd_order < 0;
        if (bp_only && d.eng->bp_batch(g.d_llr, g.count, d.bp, out, &batch[i]))
            key[i] = bp_flood_instance(batch[i].code, batch[i].p, batch[i].tune);
    }
    // launches: lists of group indices (a fused list shares one key)
    std::vector<std::vector<int>> launches;
    std::vector<int> launch_of_key;  // parallel to keys_seen
    std::vector<int> keys_seen;
    for (int i = 0; i < n; ++i) {
        if (key[i] >= 0) {
            const auto it = std::find(keys_seen.begin(), keys_seen.end(), key[i]);
            if (it != keys_seen.end()) {
                launches[launch_of_key[it - keys_seen.begin()]].push_back(i);
                continue;
            }
            keys_seen.push_back(key[i]);
            launch_of_key.push_back(static_cast<int>(launches.size()));
        }
        launches.push_back({i});
    }
    SICC_HIP(hipEventRecord(md.fork, s));
    const int nl = static_cast<int>(launches.size());
    for (int li = 0; li < nl; ++li) {
        const std::vector<int>& L = launches[li];
        hipStream_t ls = md.lanes[li % md.lanes.size()];
        SICC_HIP(hipStreamWaitEvent(ls, md.fork, 0));
        if (L.size() > 1) {
            std::vector<BpBatch> fb;
            for (int i : L) fb.push_back(batch[i]);
            launch_bp_flood_multi(fb.data(), static_cast<int>(fb.size()), ls);
        } else {
            const int i = L[0];
            const sicc_group_t& g = groups[i];
            Decoder& d = *md.members[g.member];
            d.eng->decode(g.d_llr, g.count, d.bp, d.kind == Kind::Stream ? d.osd_order : -1, d.osd, outs[i], ls,
                          scratch[i].get());
        }
        SICC_HIP(hipEventRecord(md.joins[li % md.joins.size()], ls));
    }
    const int used = std::min<int>(nl, static_cast<int>(md.lanes.size()));
    for (int i = 0; i < used; ++i) SICC_HIP(hipStreamWaitEvent(s, md.joins[i], 0));
}

}  // namespace
}  // namespace sicc

using namespace sicc;

extern "C" {

int cuda_osd_create_stream_decoder(const cuda_osd_stream_config_t* config, unsigned long long* out_handle)
{
    return guarded([&] {
        if (!out_handle) throw std::invalid_argument("null out_handle");
        *out_handle = 0;
        if (!config) throw std::invalid_argument("null config");
        if (config->m != 0 && config->m != config->n - config->k)
            throw std::invalid_argument("config m must equal n - k");
        *out_handle = make_stream(config, bundled_ccsds(config->n, config->k));
        return CUDA_OSD_STATUS_OK;
    });
}

int cuda_osd_create_stream_decoder_with_matrices(const cuda_osd_stream_config_t* config,
                                                 const unsigned char* dense_h,
                                                 const unsigned long long* systematic_parity,
                                                 int parity_words_per_row, unsigned long long* out_handle)
{
    return guarded([&] {
        if (!out_handle) throw std::invalid_argument("null out_handle");
        *out_handle = 0;
        if (!config || !dense_h) throw std::invalid_argument("null config or dense_h");
        const int n = config->n, k = config->k, m = config->m;
        if (n <= 0 || k <= 0 || m <= 0 || k + m != n) throw std::invalid_argument("need n = k + m > 0");
        LdpcCode code = code_from_dense_h(m, n, dense_h, true);
        if (systematic_parity) {
            // The caller's systematic parity must agree with the RREF encoder on H.
            if (parity_words_per_row != (n - k + 63) / 64) throw std::invalid_argument("parity_words_per_row mismatch");
            std::vector<uint8_t> unit(k, 0), cw(n);
            for (int i = 0; i < k; ++i) {
                std::fill(unit.begin(), unit.end(), 0);
                unit[i] = 1;
                code.encode(unit.data(), cw.data());
                for (int p = 0; p < n - k; ++p) {
                    const int bit = static_cast<int>(
                        (systematic_parity[static_cast<size_t>(i) * parity_words_per_row + p / 64] >> (p % 64)) & 1ull);
                    if (bit != cw[k + p]) throw std::invalid_argument("systematic parity does not match H");
                }
            }
        }
        *out_handle = make_stream(config, std::move(code));
        return CUDA_OSD_STATUS_OK;
    });
}

int cuda_osd_destroy_stream_decoder(unsigned long long handle)
{
    return guarded([&] {
        auto d = table().get(handle);
        if (!d || d->kind != Kind::Stream) throw std::invalid_argument("destroy_stream_decoder: unknown handle");
        table().remove(handle);
        return CUDA_OSD_STATUS_OK;
    });
}

int cuda_osd_create_osd_decoder(const cuda_osd_osd_config_t* config, unsigned long long* out_handle)
{
    return guarded([&] {
        if (!out_handle) throw std::invalid_argument("null out_handle");
        *out_handle = 0;
        if (!config || config->max_batch_size <= 0) throw std::invalid_argument("invalid OSD config");
        auto d = std::make_unique<Decoder>();
        d->kind = Kind::OsdOnly;
        init_common(*d);
        d->eng = std::make_unique<Engine>(bundled_ccsds(config->n, config->k));
        if (!d->eng->osd_supported()) throw std::invalid_argument("OSD supports k <= 128 and n <= 256");
        d->osd.order = 2;
        d->osd_order = 2;
        d->osd.max_flip = config->k;
        d->osd.max_tests = 4096;
        d->max_batch = config->max_batch_size;
        *out_handle = table().add(std::move(d));
        return CUDA_OSD_STATUS_OK;
    });
}

int cuda_osd_destroy_osd_decoder(unsigned long long handle)
{
    return guarded([&] {
        auto d = table().get(handle);
        if (!d || d->kind != Kind::OsdOnly) throw std::invalid_argument("destroy_osd_decoder: unknown handle");
        table().remove(handle);
        return CUDA_OSD_STATUS_OK;
    });
}

int cuda_osd_create_bp_csr_decoder(const cuda_osd_bp_csr_config_t* config, const int* row_ptr, const int* col_ind,
                                   unsigned long long* out_handle)
{
    return guarded([&] {
        if (!out_handle) throw std::invalid_argument("null out_handle");
        *out_handle = 0;
        if (!config || !row_ptr || !col_ind) throw std::invalid_argument("null config or CSR");
        const int n = config->n, m = config->m;
        if (n <= 0 || m <= 0 || config->max_iterations < 1 || !(config->normalization > 0.0f))
            throw std::invalid_argument("invalid BP config");
        if (row_ptr[0] != 0 || row_ptr[m] != config->nnz) throw std::invalid_argument("row_ptr does not match nnz");
        std::vector<int> rp(row_ptr, row_ptr + m + 1), ci(col_ind, col_ind + config->nnz);
        LdpcCode code;
        try {
            code = build_code(n, m, rp, ci, true);
        } catch (const std::runtime_error&) {
            code = build_code(n, m, rp, ci, false);
        }
        auto d = std::make_unique<Decoder>();
        d->kind = Kind::BpCsr;
        init_common(*d);
        d->eng = std::make_unique<Engine>(std::move(code));
        d->bp.max_iter = config->max_iterations;
        d->bp.alpha = config->normalization;
        d->bp.offset = config->offset;
        d->bp.damping = config->damping;
        d->bp.min_abs_accept = config->min_abs_llr_accept;
        d->schedule = config->schedule;
        d->bp.schedule = config->schedule;
        d->osd_order = -1;
        d->max_batch = std::max(config->max_batch_size, 1);
        *out_handle = table().add(std::move(d));
        return CUDA_OSD_STATUS_OK;
    });
}

int cuda_osd_destroy_bp_csr_decoder(unsigned long long handle)
{
    return guarded([&] {
        auto d = table().get(handle);
        if (!d || d->kind != Kind::BpCsr) throw std::invalid_argument("destroy_bp_csr_decoder: unknown handle");
        table().remove(handle);
        return CUDA_OSD_STATUS_OK;
    });
}

static int bp_csr_batch(unsigned long long handle, const float* llrs, int codewords, unsigned char* out_code_bits,
                        unsigned char* out_success, double* out_ms, bool device_in)
{
    return guarded([&] {
        auto d = lookup(handle, "decode_bp_csr_batch");
        if (d->kind != Kind::BpCsr) throw std::invalid_argument("decode_bp_csr_batch needs a BP CSR handle");
        if (!llrs || codewords <= 0 || !out_code_bits || !out_success)
            throw std::invalid_argument("decode_bp_csr_batch: empty batch or null output");
        const int n = d->eng->code().n;
        ensure_scratch(*d, codewords, false);
        const float* src = llrs;
        SICC_HIP(hipEventRecord(d->ev0, d->stream));
        if (!device_in) {
            SICC_HIP(hipMemcpyAsync(d->s_llr.get(), llrs, static_cast<size_t>(codewords) * n * sizeof(float),
                                    hipMemcpyHostToDevice, d->stream));
            src = d->s_llr.get();
        }
        DecodeOut out;
        out.code = d->s_a.get();
        out.bp_ok = d->s_c.get();
        d->eng->decode(src, codewords, d->bp, -1, d->osd, out, d->stream);
        SICC_HIP(hipMemcpyAsync(out_code_bits, d->s_a.get(), static_cast<size_t>(codewords) * n, hipMemcpyDeviceToHost,
                                d->stream));
        SICC_HIP(hipMemcpyAsync(out_success, d->s_c.get(), codewords, hipMemcpyDeviceToHost, d->stream));
        SICC_HIP(hipEventRecord(d->ev1, d->stream));
        SICC_HIP(hipStreamSynchronize(d->stream));
        float ms = 0.0f;
        SICC_HIP(hipEventElapsedTime(&ms, d->ev0, d->ev1));
        if (out_ms) *out_ms = ms;
        return CUDA_OSD_STATUS_OK;
    });
}

int cuda_osd_decode_bp_csr_batch(unsigned long long handle, const float* llrs, int codewords,
                                 unsigned char* out_code_bits, unsigned char* out_success, double* out_total_decode_ms)
{
    return bp_csr_batch(handle, llrs, codewords, out_code_bits, out_success, out_total_decode_ms, false);
}

int cuda_osd_decode_bp_csr_device_batch(unsigned long long handle, const float* device_llrs, int codewords,
                                        unsigned char* out_code_bits, unsigned char* out_success,
                                        double* out_total_decode_ms)
{
    return bp_csr_batch(handle, device_llrs, codewords, out_code_bits, out_success, out_total_decode_ms, true);
}

int cuda_osd_decode_osd_batch(unsigned long long handle, const float* llrs, int codewords,
                              unsigned char* out_info_bits, float* out_distances, double* out_total_decode_ms,
                              double* out_throughput_mbps)
{
    return guarded([&] {
        auto d = lookup(handle, "decode_osd_batch");
        if (d->kind != Kind::OsdOnly) throw std::invalid_argument("decode_osd_batch needs an OSD decoder handle");
        if (!llrs || codewords <= 0 || !out_info_bits) throw std::invalid_argument("decode_osd_batch: empty batch");
        const int n = d->eng->code().n, k = d->eng->code().k;
        ensure_scratch(*d, codewords, false);
        const auto t0 = std::chrono::steady_clock::now();
        SICC_HIP(hipMemcpyAsync(d->s_llr.get(), llrs, static_cast<size_t>(codewords) * n * sizeof(float),
                                hipMemcpyHostToDevice, d->stream));
        DecodeOut out;
        out.info = d->s_a.get();
        out.distance = d->s_dist.get();
        OsdParams op = d->osd;
        op.order = d->osd_order;
        d->eng->osd_only(d->s_llr.get(), codewords, op, out, d->stream);
        SICC_HIP(hipMemcpyAsync(out_info_bits, d->s_a.get(), static_cast<size_t>(codewords) * k, hipMemcpyDeviceToHost,
                                d->stream));
        if (out_distances)
            SICC_HIP(hipMemcpyAsync(out_distances, d->s_dist.get(), codewords * sizeof(float), hipMemcpyDeviceToHost,
                                    d->stream));
        SICC_HIP(hipStreamSynchronize(d->stream));
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (out_total_decode_ms) *out_total_decode_ms = ms;
        if (out_throughput_mbps) *out_throughput_mbps = ms > 0.0 ? codewords * static_cast<double>(k) / (ms * 1e3) : 0.0;
        return CUDA_OSD_STATUS_OK;
    });
}

int cuda_osd_submit_codewords(unsigned long long handle, const float* llrs, int codewords,
                              const unsigned long long* timestamps_us, const unsigned long long* frame_ids,
                              const unsigned int* block_indices)
{
    return guarded([&] {
        auto d = lookup(handle, "submit_codewords");
        submit(*d, llrs, codewords, false, timestamps_us, frame_ids, block_indices);
        return CUDA_OSD_STATUS_OK;
    });
}

int cuda_osd_submit_device_codewords(unsigned long long handle, const float* device_llrs, int codewords,
                                     const unsigned long long* timestamps_us, const unsigned long long* frame_ids,
                                     const unsigned int* block_indices)
{
    return guarded([&] {
        auto d = lookup(handle, "submit_device_codewords");
        submit(*d, device_llrs, codewords, true, timestamps_us, frame_ids, block_indices);
        return CUDA_OSD_STATUS_OK;
    });
}

int cuda_osd_poll_ready(unsigned long long handle, int max_codewords, unsigned char* out_info_bits,
                        unsigned long long* out_timestamps_us, unsigned long long* out_frame_ids,
                        unsigned int* out_block_indices, unsigned char* out_used_osd, int* out_codewords)
{
    return guarded([&] {
        auto d = lookup(handle, "poll_ready");
        if (d->kind != Kind::Stream) throw std::invalid_argument("poll_ready needs a stream decoder handle");
        if (!out_codewords) throw std::invalid_argument("null out_codewords");
        *out_codewords = 0;
        if (max_codewords <= 0) return CUDA_OSD_STATUS_OK;
        if (!out_info_bits) throw std::invalid_argument("null out_info_bits");
        maybe_launch_on_latency(*d);
        harvest_completed(*d, false);
        // After flush, poll blocks until the oldest batch is back, so a drained
        // flush never reports zero while work is in flight (LINK:3499-3530).
        if (d->ready.empty() && !d->inflight.empty() && (d->flushed || d->filling < 0)) harvest_completed(*d, true);
        const int k = d->eng->code().k;
        int got = 0;
        while (got < max_codewords && !d->ready.empty()) {
            ReadyItem& it = d->ready.front();
            std::memcpy(out_info_bits + static_cast<size_t>(got) * k, it.info.data(), k);
            if (out_timestamps_us) out_timestamps_us[got] = it.ts;
            if (out_frame_ids) out_frame_ids[got] = it.id;
            if (out_block_indices) out_block_indices[got] = it.blk;
            if (out_used_osd) out_used_osd[got] = it.used;
            d->ready.pop_front();
            ++got;
        }
        d->polled += static_cast<uint64_t>(got);
        d->st.current_pending_codewords = d->st.submitted_codewords - d->polled;
        *out_codewords = got;
        return CUDA_OSD_STATUS_OK;
    });
}

int cuda_osd_flush(unsigned long long handle)
{
    return guarded([&] {
        auto d = lookup(handle, "flush");
        if (d->kind != Kind::Stream) throw std::invalid_argument("flush needs a stream decoder handle");
        launch_filling(*d);
        d->flushed = true;
        return CUDA_OSD_STATUS_OK;
    });
}

int cuda_osd_get_stats(unsigned long long handle, cuda_osd_stream_stats_t* out_stats)
{
    return guarded([&] {
        auto d = lookup(handle, "get_stats");
        if (!out_stats) throw std::invalid_argument("null out_stats");
        harvest_completed(*d, false);
        *out_stats = d->st;
        return CUDA_OSD_STATUS_OK;
    });
}

unsigned long long cuda_osd_get_decoded_codewords(unsigned long long handle)
{
    auto d = table().get(handle);
    return d ? d->st.decoded_codewords : 0ull;
}

double cuda_osd_get_total_decode_ms(unsigned long long handle)
{
    auto d = table().get(handle);
    return d ? d->st.total_decode_ms : 0.0;
}

const char* cuda_osd_get_last_error(void) { return last_error(); }

// ---- extensions ----

int sicc_create_stream_decoder_alist(const cuda_osd_stream_config_t* config, const char* alist_path,
                                     unsigned long long* out_handle)
{
    return guarded([&] {
        if (!out_handle || !alist_path) throw std::invalid_argument("null argument");
        *out_handle = 0;
        LdpcCode code;
        try {
            code = load_alist(alist_path, true);
        } catch (const std::runtime_error& e) {
            throw std::invalid_argument(e.what());
        }
        *out_handle = make_stream(config, std::move(code));
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_create_stream_decoder_csr(const cuda_osd_stream_config_t* config, int n, int m, const int* row_ptr,
                                   const int* col_ind, unsigned long long* out_handle)
{
    return guarded([&] {
        if (!out_handle || !row_ptr || !col_ind) throw std::invalid_argument("null argument");
        *out_handle = 0;
        if (m <= 0 || n <= 0) throw std::invalid_argument("invalid n/m");
        std::vector<int> rp(row_ptr, row_ptr + m + 1), ci(col_ind, col_ind + row_ptr[m]);
        *out_handle = make_stream(config, build_code(n, m, rp, ci, true));
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_qc_code_csr(int n, int mb, int col_weight, unsigned long long seed, int* out_m, int* out_nnz, int* row_ptr,
                     int* col_ind)
{
    return guarded([&] {
        LdpcCode c = make_qc_code(n, mb, col_weight, seed);
        if (out_m) *out_m = c.m;
        if (out_nnz) *out_nnz = c.edges();
        if (row_ptr) std::copy(c.row_ptr.begin(), c.row_ptr.end(), row_ptr);
        if (col_ind) std::copy(c.col_ind.begin(), c.col_ind.end(), col_ind);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_set_osd_params(unsigned long long handle, int order, int max_flip, int max_tests)
{
    return guarded([&] {
        auto d = lookup(handle, "set_osd_params");
        if (d->kind == Kind::BpCsr) throw std::invalid_argument("BP CSR decoders have no OSD stage");
        if (order > 3 || (order >= 0 && (max_flip < 1 || max_tests < 1)))
            throw std::invalid_argument("OSD order must be <= 3, max_flip and max_tests >= 1");
        if (d->kind == Kind::OsdOnly && order < 0) throw std::invalid_argument("OSD-only decoders need order >= 0");
        d->osd_order = order;
        d->osd.order = order;
        if (order >= 0) {
            d->osd.max_flip = max_flip;
            d->osd.max_tests = max_tests;
        }
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_get_code_info(unsigned long long handle, int* n, int* k, int* m, int* nnz)
{
    return guarded([&] {
        auto d = lookup(handle, "get_code_info");
        const LdpcCode& c = d->eng->code();
        if (n) *n = c.n;
        if (k) *k = c.k;
        if (m) *m = c.m;
        if (nnz) *nnz = c.edges();
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_get_code_csr(unsigned long long handle, int* row_ptr, int* col_ind)
{
    return guarded([&] {
        auto d = lookup(handle, "get_code_csr");
        const LdpcCode& c = d->eng->code();
        if (!row_ptr || !col_ind) throw std::invalid_argument("null output");
        std::copy(c.row_ptr.begin(), c.row_ptr.end(), row_ptr);
        std::copy(c.col_ind.begin(), c.col_ind.end(), col_ind);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_get_generator(unsigned long long handle, unsigned char* out_gen)
{
    return guarded([&] {
        auto d = lookup(handle, "get_generator");
        if (!out_gen) throw std::invalid_argument("null output");
        const auto g = d->eng->code().dense_generator();
        std::memcpy(out_gen, g.data(), g.size());
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_encode(unsigned long long handle, const unsigned char* info, int count, unsigned char* out_cw)
{
    return guarded([&] {
        auto d = lookup(handle, "encode");
        if (!info || !out_cw || count < 0) throw std::invalid_argument("null argument");
        const LdpcCode& c = d->eng->code();
        for (int i = 0; i < count; ++i)
            c.encode(info + static_cast<size_t>(i) * c.k, out_cw + static_cast<size_t>(i) * c.n);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_decode_device(unsigned long long handle, const float* device_llrs, int codewords, unsigned char* d_info_bits,
                       unsigned char* d_code_bits, float* d_app_llrs, unsigned char* d_bp_success,
                       unsigned char* d_used_osd, int* d_iterations, float* d_distances, void* hip_stream)
{
    return guarded([&] {
        auto d = lookup(handle, "decode_device");
        if (!device_llrs || codewords <= 0) throw std::invalid_argument("decode_device: empty batch");
        hipStream_t s = hip_stream ? static_cast<hipStream_t>(hip_stream) : d->stream;
        DecodeOut out;
        out.info = d_info_bits;
        out.code = d_code_bits;
        out.app = d_app_llrs;
        out.bp_ok = d_bp_success;
        out.used_osd = d_used_osd;
        out.iterations = d_iterations;
        out.distance = d_distances;
        if (d->kind == Kind::OsdOnly) {
            OsdParams op = d->osd;
            op.order = d->osd_order;
            d->eng->osd_only(device_llrs, codewords, op, out, s);
        } else {
            d->eng->decode(device_llrs, codewords, d->bp, d->kind == Kind::Stream ? d->osd_order : -1, d->osd, out, s);
        }
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_set_bp_tune(unsigned long long handle, int team, int teams_per_block, int max_waves, int steal_chunk)
{
    return guarded([&] {
        auto d = lookup(handle, "set_bp_tune");
        BpTune t;
        t.team = team;
        t.teams_per_block = teams_per_block;
        t.max_waves = max_waves;
        t.steal_chunk = steal_chunk;
        d->eng->set_tune(t);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_get_bp_tune(unsigned long long handle, int* team, int* teams_per_block, int* max_waves, int* steal_chunk)
{
    return guarded([&] {
        auto d = lookup(handle, "get_bp_tune");
        const BpTune& t = d->eng->tune();
        if (team) *team = t.team;
        if (teams_per_block) *teams_per_block = t.teams_per_block;
        if (max_waves) *max_waves = t.max_waves > 0 ? t.max_waves : d->eng->max_waves();
        if (steal_chunk) *steal_chunk = t.steal_chunk;
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_autotune(unsigned long long handle, const float* device_llrs, int codewords, int* team, int* teams_per_block,
                  int* max_waves, int* steal_chunk, double* best_ms)
{
    return guarded([&] {
        auto d = lookup(handle, "autotune");
        const double ms = d->eng->autotune(device_llrs, codewords, d->bp, d->stream);
        SICC_HIP(hipStreamSynchronize(d->stream));
        const BpTune& t = d->eng->tune();
        if (team) *team = t.team;
        if (teams_per_block) *teams_per_block = t.teams_per_block;
        if (max_waves) *max_waves = t.max_waves;
        if (steal_chunk) *steal_chunk = t.steal_chunk;
        if (best_ms) *best_ms = ms;
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_multi_create(const unsigned long long* decoders, int count, unsigned long long* out_handle)
{
    return guarded([&] {
        if (!out_handle || !decoders || count <= 0) throw std::invalid_argument("multi_create: no decoders");
        *out_handle = 0;
        auto md = std::make_unique<MultiDec>();
        SICC_HIP(hipGetDevice(&md->device));
        for (int i = 0; i < count; ++i) {
            auto d = table().get(decoders[i]);
            if (!d || d->kind == Kind::OsdOnly) throw std::invalid_argument("multi_create: members must be BP decoders");
            if (d->device != md->device) throw std::invalid_argument("multi_create: members on another device");
            md->members.push_back(d);
        }
        SICC_HIP(hipStreamCreateWithFlags(&md->main, hipStreamNonBlocking));
        SICC_HIP(hipEventCreateWithFlags(&md->fork, hipEventDisableTiming));
        for (int i = 0; i < 8; ++i) {
            hipStream_t s = nullptr;
            hipEvent_t e = nullptr;
            SICC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            SICC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            md->lanes.push_back(s);
            md->joins.push_back(e);
        }
        *out_handle = mtable().add(std::move(md));
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_multi_destroy(unsigned long long handle)
{
    return guarded([&] {
        if (!mtable().remove(handle)) throw std::invalid_argument("multi_destroy: unknown handle");
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_multi_decode_device(unsigned long long handle, const sicc_group_t* groups, int n_groups, int use_graph,
                             void* hip_stream)
{
    return guarded([&] {
        auto md = mtable().get(handle);
        if (!md) throw std::invalid_argument("multi_decode: unknown handle");
        SICC_HIP(hipSetDevice(md->device));
        if (!groups || n_groups <= 0) throw std::invalid_argument("multi_decode: no groups");
        for (int i = 0; i < n_groups; ++i) {
            const sicc_group_t& g = groups[i];
            if (g.member < 0 || g.member >= static_cast<int>(md->members.size()) || g.count <= 0 || !g.d_llr)
                throw std::invalid_argument("multi_decode: bad group " + std::to_string(i));
        }
        hipStream_t s = hip_stream ? static_cast<hipStream_t>(hip_stream) : md->main;
        if (!use_graph) {
            enqueue_groups(*md, groups, n_groups, s, md->eager_scratch);
            return CUDA_OSD_STATUS_OK;
        }
        const std::vector<unsigned char> params = group_params(*md, groups, n_groups);
        if (!md->exec || !same_groups(md->key, groups, n_groups) || params != md->key_params) {
            if (md->exec) SICC_HIP(hipGraphExecDestroy(md->exec));
            if (md->graph) SICC_HIP(hipGraphDestroy(md->graph));
            md->exec = nullptr;
            md->graph = nullptr;
            // an earlier replay may still read the old graph's scratch
            SICC_HIP(hipStreamSynchronize(s));
            SICC_HIP(hipStreamSynchronize(md->main));
            md->graph_scratch.clear();
            for (int i = 0; i < n_groups; ++i) {
                Decoder& d = *md->members[groups[i].member];
                d.eng->set_timing(false);
                md->graph_scratch.push_back(std::make_unique<DecodeScratch>());
                const bool osd = d.kind == Kind::Stream && d.osd_order >= 0;
                d.eng->reserve(osd ? groups[i].count : 0, &d.bp, md->graph_scratch.back().get());
            }
            SICC_HIP(hipStreamBeginCapture(md->main, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_groups(*md, groups, n_groups, md->main, md->graph_scratch);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(md->main, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            SICC_HIP(hipStreamEndCapture(md->main, &md->graph));
            SICC_HIP(hipGraphInstantiate(&md->exec, md->graph, nullptr, nullptr, 0));
            md->key.assign(groups, groups + n_groups);
            md->key_params = params;
        }
        SICC_HIP(hipGraphLaunch(md->exec, s));
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_set_kernel_timing(unsigned long long handle, int enable)
{
    return guarded([&] {
        auto d = lookup(handle, "set_kernel_timing");
        d->eng->set_timing(enable != 0);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_get_kernel_timing(unsigned long long handle, double* bp_ms, double* osd_ms, int* bp_launches,
                           int* osd_launches)
{
    return guarded([&] {
        auto d = lookup(handle, "get_kernel_timing");
        d->eng->collect_timing(bp_ms, osd_ms, bp_launches, osd_launches);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_decode_host(unsigned long long handle, const float* llrs, int codewords, unsigned char* info_bits,
                     unsigned char* code_bits, float* app_llrs, unsigned char* bp_success, unsigned char* used_osd,
                     int* iterations, float* distances)
{
    return guarded([&] {
        auto d = lookup(handle, "decode_host");
        if (!llrs || codewords <= 0) throw std::invalid_argument("decode_host: empty batch");
        const int n = d->eng->code().n, k = d->eng->code().k;
        ensure_scratch(*d, codewords, app_llrs != nullptr);
        SICC_HIP(hipMemcpyAsync(d->s_llr.get(), llrs, static_cast<size_t>(codewords) * n * sizeof(float),
                                hipMemcpyHostToDevice, d->stream));
        DecodeOut out;
        out.info = d->s_a.get();
        out.code = d->s_b.get();
        out.app = app_llrs ? d->s_app.get() : nullptr;
        out.bp_ok = d->s_c.get();
        out.used_osd = d->s_d.get();
        out.iterations = d->s_it.get();
        out.distance = d->s_dist.get();
        if (d->kind == Kind::OsdOnly) {
            OsdParams op = d->osd;
            op.order = d->osd_order;
            out.bp_ok = nullptr;
            out.iterations = nullptr;
            d->eng->osd_only(d->s_llr.get(), codewords, op, out, d->stream);
        } else {
            d->eng->decode(d->s_llr.get(), codewords, d->bp, d->kind == Kind::Stream ? d->osd_order : -1, d->osd, out,
                           d->stream);
        }
        const size_t N = static_cast<size_t>(codewords);
        if (info_bits) SICC_HIP(hipMemcpyAsync(info_bits, d->s_a.get(), N * k, hipMemcpyDeviceToHost, d->stream));
        if (code_bits) SICC_HIP(hipMemcpyAsync(code_bits, d->s_b.get(), N * n, hipMemcpyDeviceToHost, d->stream));
        if (app_llrs)
            SICC_HIP(hipMemcpyAsync(app_llrs, d->s_app.get(), N * n * sizeof(float), hipMemcpyDeviceToHost, d->stream));
        if (bp_success && d->kind != Kind::OsdOnly)
            SICC_HIP(hipMemcpyAsync(bp_success, d->s_c.get(), N, hipMemcpyDeviceToHost, d->stream));
        if (used_osd) SICC_HIP(hipMemcpyAsync(used_osd, d->s_d.get(), N, hipMemcpyDeviceToHost, d->stream));
        if (iterations && d->kind != Kind::OsdOnly)
            SICC_HIP(hipMemcpyAsync(iterations, d->s_it.get(), N * sizeof(int), hipMemcpyDeviceToHost, d->stream));
        if (distances) SICC_HIP(hipMemcpyAsync(distances, d->s_dist.get(), N * sizeof(float), hipMemcpyDeviceToHost, d->stream));
        SICC_HIP(hipStreamSynchronize(d->stream));
        return CUDA_OSD_STATUS_OK;
    });
}

}  // extern "C"
