// engine.cpp -- device tables and BP -> OSD orchestration for one code.
#include "engine.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

namespace sicc {

Engine::Engine(LdpcCode code) : code_(std::move(code))
{
    const int n = code_.n, m = code_.m, E = code_.edges();
    int max_dc = 0;
    for (int r = 0; r < m; ++r) max_dc = std::max(max_dc, code_.row_ptr[r + 1] - code_.row_ptr[r]);
    info_cols_.alloc(std::max(code_.k, 1));
    if (code_.k > 0)
        SICC_HIP(hipMemcpy(info_cols_.get(), code_.info_cols.data(), code_.k * sizeof(int), hipMemcpyHostToDevice));
    int dev = 0;
    SICC_HIP(hipGetDevice(&dev));
    SICC_HIP(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev));
    cus_ = std::max(cus_, 1);
    max_waves_ = cus_ * 32;
    own_.ensure(1, n);

    // Short-code tables: (col << 16) | slot row entries and (slot << 8) | degree
    // column entries bound n, E and the variable degree.
    short_ok_ = n <= 0x7fff && E <= 0xffff;
    if (short_ok_) {
        std::vector<int> col_cnt(n + 1, 0);
        for (int e = 0; e < E; ++e) col_cnt[code_.col_ind[e] + 1]++;
        for (int v = 0; v < n; ++v) col_cnt[v + 1] += col_cnt[v];
        for (int v = 0; v < n && short_ok_; ++v) short_ok_ = col_cnt[v + 1] - col_cnt[v] <= 255;
        if (short_ok_) {
            // Column-major slots: a column's edges in increasing row order.
            std::vector<int> fill(n, 0), slot(E);
            for (int r = 0; r < m; ++r)
                for (int e = code_.row_ptr[r]; e < code_.row_ptr[r + 1]; ++e) {
                    const int v = code_.col_ind[e];
                    slot[e] = col_cnt[v] + fill[v]++;
                }
            std::vector<int> col_meta(n);
            for (int v = 0; v < n; ++v) col_meta[v] = (col_cnt[v] << 8) | (col_cnt[v + 1] - col_cnt[v]);
            std::vector<int> row_meta(static_cast<size_t>(m) * std::max(max_dc, 1), -1);
            for (int r = 0; r < m; ++r) {
                int j = 0;
                for (int e = code_.row_ptr[r]; e < code_.row_ptr[r + 1]; ++e, ++j)
                    row_meta[static_cast<size_t>(r) * max_dc + j] = (code_.col_ind[e] << 16) | slot[e];
            }
            bpdev_.n = n;
            bpdev_.m = m;
            bpdev_.k = code_.k;
            bpdev_.E = E;
            bpdev_.max_dc = max_dc;
            short_ok_ = bp_shape_supported(bpdev_);
            if (short_ok_) {
                row_meta_.alloc(row_meta.size());
                col_meta_.alloc(n);
                SICC_HIP(hipMemcpy(row_meta_.get(), row_meta.data(), row_meta.size() * sizeof(int),
                                   hipMemcpyHostToDevice));
                SICC_HIP(hipMemcpy(col_meta_.get(), col_meta.data(), n * sizeof(int), hipMemcpyHostToDevice));
                bpdev_.row_meta = row_meta_.get();
                bpdev_.col_meta = col_meta_.get();
                bpdev_.info_cols = info_cols_.get();
                tune_.team = bp_default_team(bpdev_);
            }
        }
    }

    osd_ok_ = short_ok_ && osd_shape_supported(n, code_.k);
    if (osd_ok_) {
        const auto g = code_.dense_generator();
        const int kw = (code_.k + 63) / 64;
        std::vector<unsigned long long> cols(static_cast<size_t>(n) * kw, 0);
        for (int r = 0; r < code_.k; ++r)
            for (int c = 0; c < n; ++c)
                if (g[static_cast<size_t>(r) * n + c]) cols[static_cast<size_t>(c) * kw + r / 64] |= 1ull << (r % 64);
        gen_cols_.alloc(cols.size());
        SICC_HIP(hipMemcpy(gen_cols_.get(), cols.data(), cols.size() * sizeof(unsigned long long),
                           hipMemcpyHostToDevice));
        osddev_.n = n;
        osddev_.k = code_.k;
        osddev_.kw = kw;
        osddev_.gen_cols = gen_cols_.get();
        osddev_.info_cols = info_cols_.get();
    }
}

bool Engine::uses_long_kernel(const BpParams& bp) const
{
    const bool layered = bp.schedule >= 2 || (bp.schedule < 0 && code_.n >= 1000);
    return layered || !short_ok_;
}

// Greedy first-fit colouring of the rows in row order (row r takes the least
// colour no earlier row sharing a column has); layered groups are the colours,
// rows sorted by (colour, row).  Flooding: one group of all rows in row order.
const LongTables& Engine::long_tables(bool layered)
{
    auto& slot = long_[layered ? 1 : 0];
    if (slot) return *slot;
    const int n = code_.n, m = code_.m, E = code_.edges();
    const auto& rp = code_.row_ptr;
    const auto& ci = code_.col_ind;
    std::vector<int> order(m), grp_row{0};
    if (layered) {
        std::vector<int> cptr(n + 1, 0), crow(E), fill(n, 0);
        for (int e = 0; e < E; ++e) cptr[ci[e] + 1]++;
        for (int v = 0; v < n; ++v) cptr[v + 1] += cptr[v];
        for (int r = 0; r < m; ++r)
            for (int e = rp[r]; e < rp[r + 1]; ++e) crow[cptr[ci[e]] + fill[ci[e]]++] = r;
        std::vector<int> color(m, 0), mark(m + 1, 0);
        int ncol = 0;
        for (int r = 0; r < m; ++r) {
            for (int e = rp[r]; e < rp[r + 1]; ++e) {
                const int v = ci[e];
                for (int t = cptr[v]; t < cptr[v + 1]; ++t) {
                    const int r2 = crow[t];
                    if (r2 < r) mark[color[r2]] = r + 1;
                    if (r2 == r && t > cptr[v] && crow[t - 1] == r)
                        throw std::invalid_argument("layered BP needs rows without repeated columns");
                }
            }
            int c = 0;
            while (mark[c] == r + 1) ++c;
            color[r] = c;
            ncol = std::max(ncol, c + 1);
        }
        std::vector<int> cnt(ncol + 1, 0);
        for (int r = 0; r < m; ++r) cnt[color[r] + 1]++;
        for (int c = 0; c < ncol; ++c) cnt[c + 1] += cnt[c];
        std::vector<int> pos(cnt.begin(), cnt.end() - 1);
        for (int r = 0; r < m; ++r) order[pos[color[r]]++] = r;
        grp_row.assign(cnt.begin(), cnt.end());
    } else {
        for (int r = 0; r < m; ++r) order[r] = r;
        grp_row.push_back(m);
    }
    const int groups = static_cast<int>(grp_row.size()) - 1;
    std::vector<int> grp_edge(groups), grp_width(groups);
    long long slots = 0;
    int max_width = 0;
    for (int g = 0; g < groups; ++g) {
        int w = 0;
        for (int i = grp_row[g]; i < grp_row[g + 1]; ++i) w = std::max(w, rp[order[i] + 1] - rp[order[i]]);
        grp_edge[g] = static_cast<int>(slots);
        grp_width[g] = w;
        max_width = std::max(max_width, w);
        slots += static_cast<long long>(w) * (grp_row[g + 1] - grp_row[g]);
        if (slots > (1ll << 30)) throw std::invalid_argument("code too large for the long-code BP tables");
    }
    if (max_width > 32) throw std::invalid_argument("long-code BP supports row degrees up to 32");
    std::vector<int> edge_col(static_cast<size_t>(std::max<long long>(slots, 1)), -1);
    std::vector<int> slot_of_edge(E);
    for (int g = 0; g < groups; ++g) {
        const int rows = grp_row[g + 1] - grp_row[g];
        for (int p = 0; p < rows; ++p) {
            const int r = order[grp_row[g] + p];
            for (int e = rp[r]; e < rp[r + 1]; ++e) {
                const int sl = grp_edge[g] + (e - rp[r]) * rows + p;
                edge_col[sl] = ci[e];
                slot_of_edge[e] = sl;
            }
        }
    }
    auto t = std::make_unique<LongTables>();
    auto up = [](DevBuf<int>& b, const std::vector<int>& v) {
        b.alloc(std::max<size_t>(v.size(), 1));
        if (!v.empty()) SICC_HIP(hipMemcpy(b.get(), v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
    };
    up(t->grp_row, grp_row);
    up(t->grp_edge, grp_edge);
    up(t->grp_width, grp_width);
    up(t->edge_col, edge_col);
    if (!layered) {
        // column v's slots in increasing row order (edges of a row appear once per row)
        std::vector<int> cptr(n + 1, 0), cslot(E), fill(n, 0);
        for (int e = 0; e < E; ++e) cptr[ci[e] + 1]++;
        for (int v = 0; v < n; ++v) cptr[v + 1] += cptr[v];
        for (int r = 0; r < m; ++r)
            for (int e = rp[r]; e < rp[r + 1]; ++e) cslot[cptr[ci[e]] + fill[ci[e]]++] = slot_of_edge[e];
        up(t->col_ptr, cptr);
        up(t->col_slot, cslot);
    }
    auto& d = t->dev;
    d.n = n;
    d.m = m;
    d.k = code_.k;
    d.slots = static_cast<int>(slots);
    d.groups = groups;
    d.max_width = max_width;
    d.layered = layered ? 1 : 0;
    d.grp_row = t->grp_row.get();
    d.grp_edge = t->grp_edge.get();
    d.grp_width = t->grp_width.get();
    d.edge_col = t->edge_col.get();
    d.col_ptr = layered ? nullptr : t->col_ptr.get();
    d.col_slot = layered ? nullptr : t->col_slot.get();
    d.info_cols = info_cols_.get();
    slot = std::move(t);
    return *slot;
}

BpLongScratch Engine::long_scratch(const LongTables& t, const BpParams& bp, bool layered, DecodeScratch& sc)
{
    BpLongScratch ls;
    ls.app_in_lds = bp_long_app_fits_lds(code_.n);
    // Workgroups of a launch, each with its own message scratch slice; after its
    // first codeword each claims the next through a counter.  With the APP in
    // LDS (about 64 KiB a workgroup) the occupancy query counts one workgroup per
    // CU; up to four times that many let the dispatcher start the others as LDS
    // frees up, and about four codewords a workgroup balance best (4096 N16200
    // frames: 9.8 ms on 1x, 9.3 on 2x, 8.5 on 4x, 9.4 on 8x the count; 1024
    // frames: 4.24 ms on 1x, 4.48 on 4x).  With the APP in global memory more
    // workgroups only add scratch traffic (1024 N64800 frames: 32.8 ms on 1x,
    // 34.4 on 2x).
    ls.resident = bp_long_resident_blocks(t.dev, ls.app_in_lds);
    ls.blocks = ls.resident * (ls.app_in_lds ? 4 : 1);
    const bool need_q = !layered && bp.damping != 0.0f;
    sc.ensure(0, code_.n);  // the counters
    sc.ensure_long(ls.blocks, t.dev.slots, code_.n, need_q, !ls.app_in_lds);
    ls.work_ctr = sc.counters.get() + 1;
    ls.r = sc.long_r.get();
    ls.q = need_q ? sc.long_q.get() : nullptr;
    ls.app = ls.app_in_lds ? nullptr : sc.long_app.get();
    return ls;
}

void Engine::reserve(int num_cw, const BpParams* bp, DecodeScratch* sc)
{
    DecodeScratch& s = sc ? *sc : own_;
    s.ensure(num_cw, code_.n);
    if (bp && uses_long_kernel(*bp)) {
        const bool layered = bp->schedule >= 2 || (bp->schedule < 0 && code_.n >= 1000);
        long_scratch(long_tables(layered), *bp, layered, s);
    }
}

void Engine::decode_long(const float* d_llr, int num_cw, const BpParams& bp, bool layered, const DecodeOut& out,
                         hipStream_t s, DecodeScratch& sc)
{
    const LongTables& t = long_tables(layered);
    const BpLongScratch ls = long_scratch(t, bp, layered, sc);
    BpIo io;
    io.llr = d_llr;
    io.num_cw = num_cw;
    io.code_bits = out.code;
    io.info_bits = out.info;
    io.app = out.app;
    io.success = out.bp_ok;
    io.iterations = out.iterations;
    io.used_osd = out.used_osd;
    io.distance = out.distance;
    timed_begin(0, s);
    launch_bp_long(t.dev, bp, io, ls, s);
    timed_end(s);
}

void DecodeScratch::ensure_long(int blocks, int slots, int n, bool need_q, bool need_app)
{
    const size_t rs = static_cast<size_t>(blocks) * slots;
    if (long_r.size() < rs) long_r.alloc(rs);
    if (need_q && long_q.size() < rs) long_q.alloc(rs);
    const size_t as = static_cast<size_t>(blocks) * n;
    if (need_app && long_app.size() < as) long_app.alloc(as);
}

void DecodeScratch::ensure(int num_cw, int n)
{
    if (counters.size() < 2) counters.alloc(2);
    if (num_cw <= cap) return;
    // Grow geometrically so a sweep of batch sizes does not reallocate each call.
    const int c = std::max(num_cw, cap * 2);
    osd_llr.alloc(static_cast<size_t>(c) * n);
    osd_index.alloc(c);
    cap = c;
}

void Engine::decode(const float* d_llr, int num_cw, const BpParams& bp, int osd_order, const OsdParams& osd,
                    const DecodeOut& out, hipStream_t s, DecodeScratch* scratch)
{
    DecodeScratch& sc = scratch ? *scratch : own_;
    if (num_cw <= 0) return;
    if (!d_llr) throw std::invalid_argument("null LLR pointer");
    const bool use_osd = osd_order >= 0;
    if (uses_long_kernel(bp)) {
        if (use_osd) throw std::invalid_argument("BP-OSD runs the flooding schedule of codes with n <= 256");
        decode_long(d_llr, num_cw, bp, bp.schedule >= 2 || (bp.schedule < 0 && code_.n >= 1000), out, s, sc);
        return;
    }
    if (use_osd && !osd_ok_) throw std::invalid_argument("OSD is not supported for this code shape (k <= 128, n <= 256)");
    sc.ensure(use_osd ? num_cw : 0, code_.n);
    ScratchView v;
    v.osd_llr = sc.osd_llr.get();
    v.osd_index = sc.osd_index.get();
    v.counters = sc.counters.get();
    v.cap = sc.cap;
    decode_short(d_llr, num_cw, bp, osd_order, osd, out, s, v);
}

void Engine::decode_into(const float* d_llr, int num_cw, const BpParams& bp, int osd_order, const OsdParams& osd,
                         const DecodeOut& out, hipStream_t s, const ScratchView& v)
{
    if (num_cw <= 0) return;
    if (!d_llr) throw std::invalid_argument("null LLR pointer");
    if (uses_long_kernel(bp))
        throw std::invalid_argument("the allocation-free decode runs the short-code flooding kernel only");
    if (!v.counters) throw std::invalid_argument("missing scratch counters");
    if (osd_order >= 0 && (num_cw > v.cap || !v.osd_llr || !v.osd_index))
        throw std::invalid_argument("batch of " + std::to_string(num_cw) + " codewords exceeds the " +
                                    std::to_string(v.cap) + "-codeword OSD scratch");
    decode_short(d_llr, num_cw, bp, osd_order, osd, out, s, v);
}

void Engine::decode_short(const float* d_llr, int num_cw, const BpParams& bp, int osd_order, const OsdParams& osd,
                          const DecodeOut& out, hipStream_t s, const ScratchView& v)
{
    const bool use_osd = osd_order >= 0;
    if (use_osd && !osd_ok_) throw std::invalid_argument("OSD is not supported for this code shape (k <= 128, n <= 256)");
    BpIo io;
    io.llr = d_llr;
    io.num_cw = num_cw;
    io.code_bits = out.code;
    io.info_bits = out.info;
    io.app = out.app;
    io.success = out.bp_ok;
    io.iterations = out.iterations;
    io.used_osd = out.used_osd;
    io.distance = out.distance;
    // counters: [0] OSD hand-off slots, [1] work-stealing cursor; cleared by a
    // kernel (no memset node when this sequence is captured in a graph)
    const bool steal = tune_.steal_chunk > 0;
    if (use_osd || steal) launch_zero_i32(v.counters, 2, s);
    if (use_osd) {
        io.osd_llr = v.osd_llr;
        io.osd_index = v.osd_index;
        io.osd_count = v.counters;
    }
    if (steal) {
        io.work_ctr = v.counters + 1;
        io.chunk = tune_.steal_chunk;
    }
    timed_begin(0, s);
    launch_bp_flood(bpdev_, bp, io, max_waves_, s, tune_);
    timed_end(s);
    if (use_osd) {
        OsdParams op = osd;
        op.order = osd_order;
        OsdIo oio;
        oio.llr = v.osd_llr;
        oio.index = v.osd_index;
        oio.count = v.counters;
        oio.code_bits = out.code;
        oio.info_bits = out.info;
        oio.distance = out.distance;
        timed_begin(1, s);
        launch_osd(osddev_, op, oio, num_cw, s);
        timed_end(s);
    }
}

bool Engine::bp_batch(const float* d_llr, int num_cw, const BpParams& bp, const DecodeOut& out, BpBatch* b) const
{
    if (!short_ok_ || uses_long_kernel(bp) || tune_.steal_chunk > 0 || num_cw <= 0 || !d_llr) return false;
    if (bp_flood_instance(bpdev_, bp, tune_) < 0) return false;
    b->code = bpdev_;
    b->p = bp;
    b->tune = tune_;
    b->max_waves = tune_.max_waves > 0 ? tune_.max_waves : max_waves_;
    BpIo io;
    io.llr = d_llr;
    io.num_cw = num_cw;
    io.code_bits = out.code;
    io.info_bits = out.info;
    io.app = out.app;
    io.success = out.bp_ok;
    io.iterations = out.iterations;
    io.used_osd = out.used_osd;
    io.distance = out.distance;
    b->io = io;
    return true;
}

void Engine::osd_only(const float* d_llr, int num_cw, const OsdParams& osd, const DecodeOut& out, hipStream_t s)
{
    if (num_cw <= 0) return;
    if (!osd_ok_) throw std::invalid_argument("OSD is not supported for this code shape (k <= 128, n <= 256)");
    OsdIo oio;
    oio.llr = d_llr;
    oio.host_count = num_cw;
    oio.code_bits = out.code;
    oio.info_bits = out.info;
    oio.distance = out.distance;
    oio.used_osd = out.used_osd;
    timed_begin(1, s);
    launch_osd(osddev_, osd, oio, num_cw, s);
    timed_end(s);
}

void Engine::set_tune(const BpTune& t)
{
    if (!short_ok_) throw std::invalid_argument("launch shapes apply to the short-code BP kernel");
    const int team = t.team > 0 ? t.team : bp_default_team(bpdev_);
    if (!bp_team_supported(bpdev_, team)) throw std::invalid_argument("BP launch shape not supported for this code");
    if (t.teams_per_block != 1 && t.teams_per_block != 2 && t.teams_per_block != 4)
        throw std::invalid_argument("teams_per_block must be 1, 2 or 4");
    if (t.max_waves < 0 || t.steal_chunk < 0) throw std::invalid_argument("max_waves and steal_chunk must be >= 0");
    tune_ = t;
    tune_.team = team;
}

double Engine::autotune(const float* d_llr, int num_cw, const BpParams& bp, hipStream_t s)
{
    if (!d_llr || num_cw <= 0) throw std::invalid_argument("autotune needs a sample batch");
    if (uses_long_kernel(bp)) throw std::invalid_argument("autotune applies to the flooding short-code BP kernel");
    DevBuf<uint8_t> bits(static_cast<size_t>(num_cw) * code_.k);
    DevBuf<uint8_t> ok(num_cw);
    BpIo io;
    io.llr = d_llr;
    io.num_cw = num_cw;
    io.info_bits = bits.get();
    io.success = ok.get();
    hipEvent_t e0, e1;
    SICC_HIP(hipEventCreate(&e0));
    SICC_HIP(hipEventCreate(&e1));
    int dev = 0, cus = 256;
    SICC_HIP(hipGetDevice(&dev));
    SICC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    BpTune best = tune_;
    double best_ms = 1e30;
    try {
        for (int team : {1, 2, 4, 8, 16}) {
            if (!bp_team_supported(bpdev_, team)) continue;
            for (int tpb : {1, 2, 4}) {
                if (team > 1 && tpb > 1) continue;
                for (int wpc : {8, 16, 32})
                    for (int sc : {0, 4, 32}) {
                        BpTune t;
                        t.team = team;
                        t.teams_per_block = tpb;
                        t.max_waves = cus * wpc;
                        t.steal_chunk = sc;
                        BpIo tio = io;
                        if (sc > 0) {
                            tio.work_ctr = own_.counters.get() + 1;
                            tio.chunk = sc;
                        }
                        auto run = [&] {
                            if (sc > 0) launch_zero_i32(own_.counters.get() + 1, 1, s);
                            launch_bp_flood(bpdev_, bp, tio, max_waves_, s, t);
                        };
                        run();  // warm-up
                        SICC_HIP(hipEventRecord(e0, s));
                        for (int r = 0; r < 3; ++r) run();
                        SICC_HIP(hipEventRecord(e1, s));
                        SICC_HIP(hipEventSynchronize(e1));
                        float ms = 0.0f;
                        SICC_HIP(hipEventElapsedTime(&ms, e0, e1));
                        if (ms / 3.0 < best_ms) {
                            best_ms = ms / 3.0;
                            best = t;
                        }
                    }
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tune_ = best;
    return best_ms;
}

void Engine::set_timing(bool on)
{
    drain_timing(true);
    timing_ = on;
    t_ms_[0] = t_ms_[1] = 0.0;
    t_cnt_[0] = t_cnt_[1] = 0;
    if (on && tpairs_.empty()) {
        tpairs_.resize(256);
        for (auto& p : tpairs_) {
            SICC_HIP(hipEventCreate(&p.a));
            SICC_HIP(hipEventCreate(&p.b));
        }
    }
}

void Engine::timed_begin(int kind, hipStream_t s)
{
    if (!timing_) return;
    TimedPair& p = tpairs_[tnext_];
    if (p.pending) {
        SICC_HIP(hipEventSynchronize(p.b));
        float ms = 0.0f;
        SICC_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        t_ms_[p.kind] += ms;
        t_cnt_[p.kind] += 1;
        p.pending = false;
    }
    p.kind = kind;
    SICC_HIP(hipEventRecord(p.a, s));
    tcur_ = tnext_;
    tnext_ = (tnext_ + 1) % static_cast<int>(tpairs_.size());
}

void Engine::timed_end(hipStream_t s)
{
    if (!timing_ || tcur_ < 0) return;
    TimedPair& p = tpairs_[tcur_];
    SICC_HIP(hipEventRecord(p.b, s));
    p.pending = true;
    tcur_ = -1;
}

void Engine::drain_timing(bool)
{
    for (auto& p : tpairs_) {
        if (!p.pending) continue;
        SICC_HIP(hipEventSynchronize(p.b));
        float ms = 0.0f;
        SICC_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        t_ms_[p.kind] += ms;
        t_cnt_[p.kind] += 1;
        p.pending = false;
    }
}

void Engine::collect_timing(double* bp_ms, double* osd_ms, int* bp_launches, int* osd_launches)
{
    drain_timing(true);
    if (bp_ms) *bp_ms = t_ms_[0];
    if (osd_ms) *osd_ms = t_ms_[1];
    if (bp_launches) *bp_launches = t_cnt_[0];
    if (osd_launches) *osd_launches = t_cnt_[1];
}

Engine::~Engine()
{
    for (auto& p : tpairs_) {
        if (p.a) (void)hipEventDestroy(p.a);
        if (p.b) (void)hipEventDestroy(p.b);
    }
}

}  // namespace sicc
