// bp_long_kernels.hip -- normalized min-sum BP for long codes (DVB-S2 N16200 /
// N64800 and any CSR code the short-code kernel cannot hold), flooding or
// layered schedule.
//
// One workgroup of 16 wavefronts (8 for row degrees above 16) per codeword, persistent over
// the batch: after its first codeword (its block index) each workgroup takes the
// next undecoded one from a global counter, so the batch's tail balances over
// the workgroups that finish first, whatever the grid size.  APP values live in LDS when n * 4 bytes fit in 64 KiB (n <= 16 368),
// otherwise in a per-workgroup slice of global scratch (L2 / MALL resident).
// Check-to-variable messages r live in global scratch in a group-major,
// position-minor layout: group g's rows are processed by consecutive threads,
// and edge j of the row at position p is slot grp_edge[g] + j * rows_g + p, so a
// wavefront's loads of one edge index are contiguous.
//
// Layered (schedule 2 / 3): the groups are the colours of a greedy first-fit
// row colouring; rows of a colour share no column and run in parallel, colours
// run in order.  Each row: q = APP - r_old, the min-sum check update of the
// flooding decoder (LINK:3035-3066), r' = alpha * s * mag, optional damping
// r' = (1 - d) r' + d r_old, APP = q + r'.
// Flooding (schedule 0 / 1): one group of all rows in row order; check phase
// from q = APP - r_old (damped against the stored q when damping != 0, as
// MOSD:174-177), then each column sums llr + r over its edges in increasing row
// order (LINK:3068-3074) -- bit-identical to the short-code kernel's result.
// Both: syndrome stop after every iteration; a frame is accepted when the
// syndrome is zero and min |APP| >= min_abs_accept.
#include "kernels.hpp"

#include <cfloat>
#include <stdexcept>
#include <string>

namespace sicc {

namespace {

// Workgroup barrier that orders LDS only: global stores still in flight (this
// thread's own check messages, which only this thread reads back) and global
// loads issued ahead (the next colour's messages) are not waited for, as a full
// __syncthreads() would.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Barrier after a step that changed the APP: LDS-only when the APP lives in LDS.
template <bool APP_LDS>
__device__ __forceinline__ void app_barrier()
{
    if (APP_LDS)
        lds_barrier();
    else
        __syncthreads();
}

// 16 wavefronts per codeword; rows of degree above 16 keep 32 messages in
// registers, which needs the larger register budget of 8 wavefronts.
template <int MAXW>
constexpr int long_threads()
{
    return MAXW > 16 ? 512 : 1024;
}

// Rows of up to 8 edges with the APP in LDS: held to 64 VGPRs (a few registers
// spill), so that two 1024-thread workgroups (eight waves per SIMD) share a CU:
// N16200 layered 8.43 -> 6.33 ms per 4 096 frames with the next-colour loads
// below.  With the APP in global memory the cap and those loads are slower
// (N64800 32.7 -> 59.3 ms): every barrier there waits for all memory operations.
template <int MAXW, bool LAYERED, bool APP_LDS>
__global__ __launch_bounds__(long_threads<MAXW>()) __attribute__((amdgpu_waves_per_eu(MAXW <= 8 && APP_LDS ? 8 : 1))) void
bp_long_kernel(BpLongCodeDev c, BpParams p, BpIo io, BpLongScratch sc)
{
    constexpr int kLongThreads = long_threads<MAXW>();
    extern __shared__ float lds_app[];
    __shared__ int s_flag[2];
    __shared__ float s_min[16];
    __shared__ int s_next;
    const int tid = threadIdx.x;
    float* r = sc.r + static_cast<size_t>(blockIdx.x) * c.slots;
    float* qst = sc.q ? sc.q + static_cast<size_t>(blockIdx.x) * c.slots : nullptr;
    float* app = APP_LDS ? lds_app : sc.app + static_cast<size_t>(blockIdx.x) * c.n;
    const float alpha = p.alpha, offset = p.offset, damping = p.damping;

    for (int cw = blockIdx.x; cw < io.num_cw;) {
        const float* llr = io.llr + static_cast<size_t>(cw) * c.n;
        for (int v = tid; v < c.n; v += kLongThreads) app[v] = llr[v];
        for (int e = tid; e < c.slots; e += kLongThreads) r[e] = 0.0f;
        if (tid < 2) s_flag[tid] = 0;
        __syncthreads();
        int it = 0;
        bool ok = false;
        // Column indices and old messages of one row (every slot of the group's
        // width exists, padding included: the loads are issued together, not
        // chained).
        auto load_row = [&](int g, int pos, int* cv, float* rv) {
            const int rows = c.grp_row[g + 1] - c.grp_row[g];
            const int base = c.grp_edge[g];
            const int width = c.grp_width[g];
#pragma unroll
            for (int j = 0; j < MAXW; ++j) {
                if (j < width) {
                    const int slot = base + j * rows + pos;
                    cv[j] = c.edge_col[slot];
                    rv[j] = r[slot];
                } else {
                    cv[j] = -1;
                    rv[j] = 0.0f;
                }
            }
        };
        while (it < p.max_iter) {
            // Layered: a thread's first row of the next colour is loaded before the
            // barrier that ends the current colour, so the loads overlap the wait.
            // Its messages were last written by this thread (rows keep their thread
            // from iteration to iteration) and the column indices are constant, so
            // only the APP reads need the barrier.  (Rows of more than 16 edges keep
            // one row in registers; with the APP in global memory the barrier waits
            // for the loads, so they are issued after it.)
            constexpr bool PF = LAYERED && APP_LDS && MAXW <= 16;
            int pcv[MAXW];
            float prv[MAXW];
            if (PF && tid < c.grp_row[1] - c.grp_row[0]) load_row(0, tid, pcv, prv);
            for (int g = 0; g < c.groups; ++g) {
                const int rows = c.grp_row[g + 1] - c.grp_row[g];
                const int base = c.grp_edge[g];
                const int width = c.grp_width[g];
                for (int pos = tid; pos < rows; pos += kLongThreads) {
                    float qv[MAXW], rv[MAXW];
                    int cv[MAXW];
                    float m1 = FLT_MAX, m2 = FLT_MAX;
                    int mi = -1, neg = 0, deg = 0;
                    if (PF && pos == tid) {
#pragma unroll
                        for (int j = 0; j < MAXW; ++j) {
                            cv[j] = pcv[j];
                            rv[j] = prv[j];
                        }
                    } else {
                        load_row(g, pos, cv, rv);
                    }
#pragma unroll
                    for (int j = 0; j < MAXW; ++j) {
                        if (j < width) {
                            const int slot = base + j * rows + pos;
                            const int col = cv[j];
                            if (col >= 0) {
                                const float rold = rv[j];
                                float q = app[col] - rold;
                                if (!LAYERED && damping != 0.0f) {
                                    if (it > 0) q = (1.0f - damping) * q + damping * qst[slot];
                                    qst[slot] = q;
                                }
                                qv[j] = q;
                                deg = j + 1;
                                neg ^= q < 0.0f ? 1 : 0;
                                const float av = fabsf(q);
                                if (av < m1) {
                                    m2 = m1;
                                    m1 = av;
                                    mi = j;
                                } else if (av < m2) {
                                    m2 = av;
                                }
                            }
                        }
                    }
#pragma unroll
                    for (int j = 0; j < MAXW; ++j) {
                        if (j < deg) {
                            const int s = ((qv[j] < 0.0f ? 1 : 0) ^ neg) ? -1 : 1;
                            float mag = (j == mi) ? m2 : m1;
                            if (offset != 0.0f) mag = fmaxf(mag - offset, 0.0f);
                            float rn = alpha * static_cast<float>(s) * mag;
                            if (LAYERED) {
                                if (damping != 0.0f) rn = (1.0f - damping) * rn + damping * rv[j];
                                app[cv[j]] = qv[j] + rn;
                            }
                            r[base + j * rows + pos] = rn;
                        }
                    }
                }
                if (PF && g + 1 < c.groups && tid < c.grp_row[g + 2] - c.grp_row[g + 1]) load_row(g + 1, tid, pcv, prv);
                // layered: the next colour reads this one's APP updates; flooding:
                // the column sums read other threads' messages from global memory
                if (LAYERED)
                    app_barrier<APP_LDS>();
                else
                    __syncthreads();
            }
            if (!LAYERED) {
                for (int v = tid; v < c.n; v += kLongThreads) {
                    float s = llr[v];
                    for (int t = c.col_ptr[v]; t < c.col_ptr[v + 1]; ++t) s += r[c.col_slot[t]];
                    app[v] = s;
                }
                __syncthreads();
            }
            // syndrome of the hard decision APP < 0; a row's column indices are
            // loaded four at a time before its APP reads, and once a violated row has
            // been flagged the remaining groups are skipped (a stale read of the flag
            // only costs time)
            const int fl = it & 1;
            for (int g = 0; g < c.groups; ++g) {
                const int rows = c.grp_row[g + 1] - c.grp_row[g];
                const int base = c.grp_edge[g];
                const int width = c.grp_width[g];
                const int lim = *static_cast<volatile int*>(&s_flag[fl]) ? 0 : rows;
                for (int pos = tid; pos < lim; pos += kLongThreads) {
                    const int* ec = c.edge_col + base + pos;
                    int par = 0;
                    for (int j0 = 0; j0 < width; j0 += 4) {
                        int cols[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) cols[u] = j0 + u < width ? ec[(j0 + u) * rows] : -1;
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (cols[u] >= 0) par ^= app[cols[u]] < 0.0f ? 1 : 0;
                    }
                    if (par) s_flag[fl] = 1;
                }
            }
            if (tid == 0) s_flag[fl ^ 1] = 0;
            app_barrier<APP_LDS>();
            ok = s_flag[fl] == 0;
            ++it;
            if (ok) break;
        }
        bool accept = ok;
        if (ok && p.min_abs_accept > 0.0f) {
            float mn = FLT_MAX;
            for (int v = tid; v < c.n; v += kLongThreads) mn = fminf(mn, fabsf(app[v]));
            for (int o = 32; o > 0; o >>= 1) mn = fminf(mn, __shfl_xor(mn, o));
            if ((tid & 63) == 0) s_min[tid >> 6] = mn;
            __syncthreads();
            mn = FLT_MAX;
            for (int w = 0; w < kLongThreads / 64; ++w) mn = fminf(mn, s_min[w]);
            accept = mn >= p.min_abs_accept;
        }
        const size_t ob = static_cast<size_t>(cw) * c.n;
        for (int v = tid; v < c.n; v += kLongThreads) {
            const float a = app[v];
            if (io.code_bits) io.code_bits[ob + v] = a < 0.0f ? 1 : 0;
            if (io.app) io.app[ob + v] = a;
        }
        if (io.info_bits)
            for (int i = tid; i < c.k; i += kLongThreads)
                io.info_bits[static_cast<size_t>(cw) * c.k + i] = app[c.info_cols[i]] < 0.0f ? 1 : 0;
        if (tid == 0) {
            if (io.success) io.success[cw] = accept ? 1 : 0;
            if (io.iterations) io.iterations[cw] = it;
            if (io.used_osd) io.used_osd[cw] = 0;
            if (io.distance) io.distance[cw] = 0.0f;
        }
        __syncthreads();  // APP and flags are reused by the next codeword
        if (tid == 0) s_next = static_cast<int>(gridDim.x) + atomicAdd(sc.work_ctr, 1);
        __syncthreads();
        cw = s_next;
    }
}

template <int MAXW>
void launch_width(const BpLongCodeDev& c, const BpParams& p, const BpIo& io, const BpLongScratch& sc, int grid,
                  hipStream_t s)
{
    const size_t lds = sc.app_in_lds ? sizeof(float) * static_cast<size_t>(c.n) : 0;
    constexpr int kLongThreads = long_threads<MAXW>();
    if (c.layered) {
        if (sc.app_in_lds)
            hipLaunchKernelGGL((bp_long_kernel<MAXW, true, true>), dim3(grid), dim3(kLongThreads), lds, s, c, p, io, sc);
        else
            hipLaunchKernelGGL((bp_long_kernel<MAXW, true, false>), dim3(grid), dim3(kLongThreads), 0, s, c, p, io, sc);
    } else {
        if (sc.app_in_lds)
            hipLaunchKernelGGL((bp_long_kernel<MAXW, false, true>), dim3(grid), dim3(kLongThreads), lds, s, c, p, io,
                               sc);
        else
            hipLaunchKernelGGL((bp_long_kernel<MAXW, false, false>), dim3(grid), dim3(kLongThreads), 0, s, c, p, io,
                               sc);
    }
}

}  // namespace

bool bp_long_app_fits_lds(int n) { return n > 0 && static_cast<size_t>(n) * sizeof(float) <= 65536 - 256; }

namespace {
template <int MAXW>
int resident_width(const BpLongCodeDev& c, bool app_in_lds)
{
    const size_t lds = app_in_lds ? sizeof(float) * static_cast<size_t>(c.n) : 0;
    constexpr int t = long_threads<MAXW>();
    const void* k = c.layered ? (app_in_lds ? reinterpret_cast<const void*>(bp_long_kernel<MAXW, true, true>)
                                            : reinterpret_cast<const void*>(bp_long_kernel<MAXW, true, false>))
                              : (app_in_lds ? reinterpret_cast<const void*>(bp_long_kernel<MAXW, false, true>)
                                            : reinterpret_cast<const void*>(bp_long_kernel<MAXW, false, false>));
    return resident_blocks_per_cu(k, t, lds) * device_cu_count();
}
}  // namespace

int bp_long_resident_blocks(const BpLongCodeDev& c, bool app_in_lds)
{
    int b = 0;
    if (c.max_width <= 8)
        b = resident_width<8>(c, app_in_lds);
    else if (c.max_width <= 16)
        b = resident_width<16>(c, app_in_lds);
    else if (c.max_width <= 32)
        b = resident_width<32>(c, app_in_lds);
    else
        throw std::invalid_argument("long BP: row degree above 32");
    return b > 0 ? b : 1;
}

void launch_bp_long(const BpLongCodeDev& c, const BpParams& p, const BpIo& io, const BpLongScratch& sc,
                    hipStream_t s)
{
    if (io.num_cw <= 0) return;
    if (!io.llr || !sc.r || !sc.work_ctr || sc.blocks <= 0)
        throw std::invalid_argument("long BP: missing input or scratch");
    if (!sc.app_in_lds && !sc.app) throw std::invalid_argument("long BP: missing APP scratch");
    if (sc.app_in_lds && !bp_long_app_fits_lds(c.n)) throw std::invalid_argument("long BP: APP does not fit LDS");
    if (!c.layered && p.damping != 0.0f && !sc.q) throw std::invalid_argument("long BP: damping needs q scratch");
    if (p.max_iter < 1) throw std::invalid_argument("long BP: max_iter must be positive");
    // about four codewords a workgroup, between the resident count and the
    // scratch slices (4096 N16200 frames: 1024 workgroups, 1024 frames: 256)
    int grid = (io.num_cw + 3) / 4;
    if (grid < sc.resident) grid = sc.resident;
    if (grid > sc.blocks) grid = sc.blocks;
    if (grid > io.num_cw) grid = io.num_cw;
    launch_zero_i32(sc.work_ctr, 1, s);  // a kernel, not a memset node, under graph capture
    if (c.max_width <= 8)
        launch_width<8>(c, p, io, sc, grid, s);
    else if (c.max_width <= 16)
        launch_width<16>(c, p, io, sc, grid, s);
    else if (c.max_width <= 32)
        launch_width<32>(c, p, io, sc, grid, s);
    else
        throw std::invalid_argument("long BP: row degree above 32");
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("bp_long_kernel: ") + hipGetErrorString(e));
}

}  // namespace sicc
