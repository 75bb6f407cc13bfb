// kernels.hpp -- host-visible launchers for the gfx950 decode kernels, and the
// device-side descriptors they take.  All pointers are device pointers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sicc {

// Parity-check graph of one code, laid out for the BP kernels.
//   row_meta[r * max_dc + j] = (col << 16) | slot, or -1 past the row degree;
//     j follows the row's CSR order, slot is the edge's position in the
//     column-major message array.
//   col_meta[v] = (first_slot << 8) | degree; a column's slots hold its edges
//     in increasing row order.
struct BpCodeDev {
    int n = 0;
    int m = 0;
    int k = 0;
    int E = 0;
    int max_dc = 0;
    const int* row_meta = nullptr;
    const int* col_meta = nullptr;
    const int* info_cols = nullptr;  // [k]
};

struct BpParams {
    int max_iter = 50;
    float alpha = 0.75f;     // normalization
    float offset = 0.0f;     // magnitude offset, mag = max(mag - offset, 0)
    float damping = 0.0f;    // VN message damping (MATLAB reference option)
    float min_abs_accept = 0.0f;  // a converged frame needs min|APP| >= this
    // 0 flooding, 1 node-parallel (= flooding), 2 colored layered, 3 flat layered
    // (= colored layered), < 0 automatic: layered when n >= 1000.
    int schedule = 0;
};

// Per-launch inputs/outputs of the BP kernel.  Null outputs are skipped.
struct BpIo {
    const float* llr = nullptr;   // [num_cw][n]
    int num_cw = 0;
    uint8_t* code_bits = nullptr; // [num_cw][n]
    uint8_t* info_bits = nullptr; // [num_cw][k]
    float* app = nullptr;         // [num_cw][n] final APP LLR
    uint8_t* success = nullptr;   // [num_cw] BP syndrome satisfied (and accepted)
    int* iterations = nullptr;    // [num_cw]
    uint8_t* used_osd = nullptr;  // [num_cw] 1 when handed to OSD
    float* distance = nullptr;    // [num_cw] 0 for accepted frames
    // Hand-off of rejected frames to OSD: best-syndrome APP and frame index.
    float* osd_llr = nullptr;     // [num_cw][n]
    int* osd_index = nullptr;     // [num_cw]
    int* osd_count = nullptr;     // one counter, zeroed before the launch
    // Work stealing: teams take `chunk` codewords at a time from this counter
    // (zeroed before the launch).  Null: static striding over the batch.
    int* work_ctr = nullptr;
    int chunk = 1;
};

// OSD inputs: the generator as column words (bit r % 64 of gen_cols[c * kw + r / 64]
// is G[r][c], k <= 128, kw = ceil(k / 64)) and the frames BP handed off.
struct OsdCodeDev {
    int n = 0;
    int k = 0;
    int kw = 1;
    const unsigned long long* gen_cols = nullptr;  // [n][kw]
    const int* info_cols = nullptr;                // [k]
};

struct OsdParams {
    int order = 1;
    int max_flip = 64;
    int max_tests = 4096;
};

struct OsdIo {
    const float* llr = nullptr;     // [count][n]
    const int* index = nullptr;     // [count] frame index, or null = identity
    const int* count = nullptr;     // device counter, or null = host_count
    int host_count = 0;
    uint8_t* code_bits = nullptr;   // [frames][n]
    uint8_t* info_bits = nullptr;   // [frames][k]
    float* distance = nullptr;      // [frames] best correlation metric
    uint8_t* used_osd = nullptr;    // [frames] set to 1 for processed frames
    int* status = nullptr;          // [frames] 0 ok, 1 rank-deficient generator
};

// Row gather for batch packing: entry i copies `count` rows of `row_len` floats
// from `src` to `dst`.  One launch serves every code of a mixed launch.  The
// table may live in pinned host memory (read by the kernel over the bus).
struct GatherEntry {
    const float* src;
    float* dst;
    int count;
    int row_len;
};
void launch_gather_rows(const GatherEntry* table, int entries, hipStream_t s);
// Result packing of one code group: `rows` rows of k 0/1 bytes at `bits` become
// ceil(k/8) bytes each at `out` (bit i at byte i/8, bit i%8), followed by
// `flag_bytes` bytes copied from `flags`; `out` must be 4-byte aligned.
// Blocks [block0, block0 + pack_blocks) of the launch serve the group, 1 KiB of
// its output each (pack_blocks = ceil((rows ceil(k/8) + flag_bytes) / 1024)).
struct PackGroup {
    const uint8_t* bits;
    const uint8_t* flags;
    uint8_t* out;
    int rows;
    int k;
    int flag_bytes;
    int block0;
    int pack_blocks;
};
// Fills block0 / pack_blocks of groups[0 .. count) and launches one kernel for
// all of them; the table may live in pinned host memory.
void launch_pack_groups(PackGroup* groups, int count, hipStream_t s);

// Kernel launchers.  They validate shapes on the host and throw
// std::invalid_argument when a code does not fit an instantiated variant.
void launch_zero_i32(int* p, int count, hipStream_t s);
// Cached residency queries (per device; kernel, block size and dynamic LDS).
int device_cu_count();
int resident_blocks_per_cu(const void* kernel, int threads, size_t lds);

bool bp_shape_supported(const BpCodeDev& c);
// Launch shape of the BP kernel (the autotuner's search space):
//   team            wavefronts cooperating on one codeword (1, 2, 4, 8, 16); 0 = default
//   teams_per_block codewords in flight per workgroup for single-wave teams (1, 2, 4)
//   max_waves       persistent-grid size in wavefronts (0 = 32 per CU)
//   steal_chunk     0: static striding; > 0: work stealing in chunks of this many codewords
struct BpTune {
    int team = 0;
    int teams_per_block = 4;
    int max_waves = 0;
    int steal_chunk = 0;
};
bool bp_team_supported(const BpCodeDev& c, int team);
int bp_default_team(const BpCodeDev& c);
void launch_bp_flood(const BpCodeDev& code, const BpParams& p, const BpIo& io, int max_waves,
                     hipStream_t s, const BpTune& tune = BpTune());
// One batch of the flooding kernel, for launches that decode the batches of
// several codes together (work stealing off: io.work_ctr must be null).
struct BpBatch {
    BpCodeDev code;
    BpParams p;
    BpIo io;
    BpTune tune;
    int max_waves = 0;
};
// Key of the kernel instance (launch shape, row-degree class, damping, and for
// undamped batches whether the offset is nonzero) a batch runs on; batches with
// equal keys can share a launch.  -1: no instance fits.
int bp_flood_instance(const BpCodeDev& code, const BpParams& p, const BpTune& tune);
// Decodes batches of one instance key, up to eight per launch, each with its
// own code, parameters and outputs; results equal launch_bp_flood's per batch.
void launch_bp_flood_multi(const BpBatch* batches, int count, hipStream_t s);
// Long-code BP (flooding or layered), one workgroup of 16 wavefronts per
// codeword (8 for row degrees above 16).
// Rows are processed in groups (layered: colours of the greedy row colouring;
// flooding: one group of all rows in row order).  Group g covers positions
// [grp_row[g], grp_row[g + 1]) of the processing order; edge j of the row at
// position p of the group is slot grp_edge[g] + j * rows_g + p, with
// edge_col[slot] = column, or -1 past the row's degree.
struct BpLongCodeDev {
    int n = 0;
    int m = 0;
    int k = 0;
    int slots = 0;  // padded edge slots
    int groups = 0;
    int max_width = 0;
    int layered = 0;
    const int* grp_row = nullptr;    // [groups + 1]
    const int* grp_edge = nullptr;   // [groups]
    const int* grp_width = nullptr;  // [groups]
    const int* edge_col = nullptr;   // [slots]
    const int* col_ptr = nullptr;    // [n + 1] flooding: column v's slots, increasing row
    const int* col_slot = nullptr;   // [edges]
    const int* info_cols = nullptr;  // [k]
};
// Per-workgroup scratch: r (and q for damped flooding) [blocks][slots]; APP
// [blocks][n] unless it is kept in LDS.
struct BpLongScratch {
    float* r = nullptr;       // [blocks][slots]
    float* q = nullptr;       // [blocks][slots], flooding with damping
    float* app = nullptr;     // [blocks][n] when the APP does not fit LDS
    int* work_ctr = nullptr;  // one counter: codewords claimed past the first grid
    int blocks = 0;           // scratch slices: the most workgroups of a launch
    int resident = 0;         // workgroups the occupancy query counts the GPU holding
    bool app_in_lds = false;
};
bool bp_long_app_fits_lds(int n);
// Workgroups of the long-code kernel for `c` the GPU holds at once.
int bp_long_resident_blocks(const BpLongCodeDev& c, bool app_in_lds);
void launch_bp_long(const BpLongCodeDev& c, const BpParams& p, const BpIo& io, const BpLongScratch& sc,
                    hipStream_t s);

bool osd_shape_supported(int n, int k);
void launch_osd(const OsdCodeDev& code, const OsdParams& p, const OsdIo& io, int max_frames,
                hipStream_t s);

}  // namespace sicc
