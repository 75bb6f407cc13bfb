// engine.hpp -- one LDPC code resident on the GPU: graph tables for BP,
// the systematic generator for OSD, and the device-side BP -> OSD hand-off.
// Everything here is stream-ordered and asynchronous; nothing synchronizes
// the device.
#pragma once

#include "code_model.hpp"
#include "kernels.hpp"
#include "sicc_common.hpp"

#include <memory>
#include <vector>

namespace sicc {

struct DecodeOut {
    uint8_t* info = nullptr;      // [cw][k]
    uint8_t* code = nullptr;      // [cw][n]
    float* app = nullptr;         // [cw][n] BP APP LLR (final iteration)
    uint8_t* bp_ok = nullptr;     // [cw]
    uint8_t* used_osd = nullptr;  // [cw]
    int* iterations = nullptr;    // [cw]
    float* distance = nullptr;    // [cw] OSD metric, 0 for BP-accepted frames
};

// Per-caller device scratch of a decode: the BP -> OSD hand-off buffers and
// the two counters ([0] hand-off slots, [1] work-stealing cursor).  Decodes
// that may run concurrently on different streams need separate scratch.
struct DecodeScratch {
    DevBuf<float> osd_llr;
    DevBuf<int> osd_index;
    DevBuf<int> counters;
    int cap = 0;
    void ensure(int num_cw, int n);
    // long-code BP: per-workgroup messages (and q for damped flooding, APP when
    // it does not fit LDS)
    DevBuf<float> long_r, long_q, long_app;
    void ensure_long(int blocks, int slots, int n, bool need_q, bool need_app);
};

// Device tables of the long-code BP kernel for one schedule.
struct LongTables {
    DevBuf<int> grp_row, grp_edge, grp_width, edge_col, col_ptr, col_slot;
    BpLongCodeDev dev;
};

// Scratch of the short-code kernel path held by the caller (no allocation):
// the BP -> OSD hand-off for up to `cap` codewords and two counters.
struct ScratchView {
    float* osd_llr = nullptr;   // [cap][n]
    int* osd_index = nullptr;   // [cap]
    int* counters = nullptr;    // [2]
    int cap = 0;
};

class Engine {
public:
    explicit Engine(LdpcCode code);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
    const LdpcCode& code() const { return code_; }
    bool osd_supported() const { return osd_ok_; }

    // BP over num_cw device-resident codewords, then OSD (order >= 0) on the
    // frames BP did not accept.  osd_order < 0 disables OSD.  The flooding
    // schedule of a code the short-code tables hold runs the short-code kernel;
    // layered schedules and longer codes run the long-code kernel (no OSD).
    void decode(const float* d_llr, int num_cw, const BpParams& bp, int osd_order, const OsdParams& osd,
                const DecodeOut& out, hipStream_t s, DecodeScratch* scratch = nullptr);
    // decode() on caller-held scratch: allocates nothing.  Only the short-code
    // (flooding) kernel path; throws std::invalid_argument when `bp` selects the
    // long-code kernel or OSD needs more than v.cap hand-off slots.
    void decode_into(const float* d_llr, int num_cw, const BpParams& bp, int osd_order, const OsdParams& osd,
                     const DecodeOut& out, hipStream_t s, const ScratchView& v);
    // The flooding-kernel batch decode(d_llr, num_cw, bp, -1, ...) runs, for a
    // launch that decodes it together with other codes' batches
    // (launch_bp_flood_multi; no kernel timing).  False when that decode cannot
    // share a launch: the long-code kernel, or work stealing.
    bool bp_batch(const float* d_llr, int num_cw, const BpParams& bp, const DecodeOut& out, BpBatch* b) const;
    // OSD on every codeword (no BP).
    void osd_only(const float* d_llr, int num_cw, const OsdParams& osd, const DecodeOut& out, hipStream_t s);

    int max_waves() const { return max_waves_; }
    // Pre-size the OSD hand-off scratch for batches of up to num_cw codewords,
    // and the long-code scratch when `bp` selects that kernel (decode() grows
    // them on demand, which a graph capture cannot do and which costs a
    // device-wide hipFree in a serving loop).  `sc` names a caller-owned
    // scratch (as passed to decode()); null means the engine's own.
    void reserve(int num_cw, const BpParams* bp = nullptr, DecodeScratch* sc = nullptr);
    // True when `bp` runs the long-code kernel (layered schedule or a code the
    // short-code tables cannot hold).
    bool uses_long_kernel(const BpParams& bp) const;
    bool short_kernel_ok() const { return short_ok_; }

    // BP launch shape.  autotune() times every supported shape (wavefronts per
    // codeword, codewords per workgroup, persistent grid size) on a sample batch
    // already in HBM and keeps the fastest; it writes only internal scratch.
    const BpTune& tune() const { return tune_; }
    void set_tune(const BpTune& t);
    double autotune(const float* d_llr, int num_cw, const BpParams& bp, hipStream_t s);

    // Kernel timing with HIP events recorded on the launch stream around each
    // BP and OSD launch.  collect_timing() waits for the recorded launches and
    // returns their summed durations since the last reset.
    void set_timing(bool on);
    void collect_timing(double* bp_ms, double* osd_ms, int* bp_launches, int* osd_launches);

private:
    struct TimedPair {
        hipEvent_t a = nullptr, b = nullptr;
        int kind = 0;  // 0 BP, 1 OSD
        bool pending = false;
    };
    void timed_begin(int kind, hipStream_t s);
    void timed_end(hipStream_t s);
    void drain_timing(bool all);
    BpTune tune_;
    bool timing_ = false;
    std::vector<TimedPair> tpairs_;
    int tnext_ = 0;
    int tcur_ = -1;
    double t_ms_[2] = {0.0, 0.0};
    int t_cnt_[2] = {0, 0};

private:

    void decode_long(const float* d_llr, int num_cw, const BpParams& bp, bool layered, const DecodeOut& out,
                     hipStream_t s, DecodeScratch& sc);
    void decode_short(const float* d_llr, int num_cw, const BpParams& bp, int osd_order, const OsdParams& osd,
                      const DecodeOut& out, hipStream_t s, const ScratchView& v);
    const LongTables& long_tables(bool layered);
    BpLongScratch long_scratch(const LongTables& t, const BpParams& bp, bool layered, DecodeScratch& sc);

    LdpcCode code_;
    bool short_ok_ = false;
    bool osd_ok_ = false;
    int cus_ = 256;
    std::unique_ptr<LongTables> long_[2];  // [0] flooding, [1] layered
    int max_waves_ = 8192;
    DevBuf<int> row_meta_;
    DevBuf<int> col_meta_;
    DevBuf<int> info_cols_;
    DevBuf<unsigned long long> gen_cols_;
    DecodeScratch own_;
    BpCodeDev bpdev_;
    OsdCodeDev osddev_;
};

}  // namespace sicc
