/*
 * sicc_serve_c_api.h -- continuous-batching decode server in libsicc_amd.so.
 *
 * No counterpart in the reference ABI; it generalizes the reference's
 * cross-frame batching loop (staging of frames into one decoder submit and
 * out-of-order completion by (frame id, block index),
 * source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp:3309-3432 and 11482-11633)
 * to jobs of any size and mixed codes.
 *
 * A server owns a device-memory pool of `budget_bytes`.  A submitted job
 * (codewords of one member code) is admitted, in arrival order, as soon as its
 * LLR, result and scratch space fits the pool; its LLRs then stay in HBM until
 * the job completes.  A scheduler thread keeps up to `slots` launches in
 * flight: whenever a slot is free it packs every admitted, not yet launched
 * job (up to `max_launch_codewords`) into per-code groups and launches them, so
 * jobs that arrive while others run overlap with them on the device.  Code
 * groups whose BP batches share one flooding-kernel instance decode in one
 * fused launch; the launch's decode units are spread over the streams of slots
 * that are idle when it is issued.
 * While launches are in flight the scheduler thread polls their completion
 * events (yielding between polls: one host core), so completions are seen
 * within microseconds; an idle server sleeps on a condition variable.  Two
 * copier threads copy each completed launch's results out of its pinned
 * buffer into the jobs' own buffers (recycled from fetched jobs) and complete
 * the jobs, so the scheduler keeps launching meanwhile; a slot is free again
 * once its copy is done, and a job's latency runs from its submission to the
 * end of that copy.  Submission does not take the scheduler's lock: jobs wait
 * in an inbox that the scheduler empties when it admits.
 * slots = 1 and max_jobs_per_launch = 1 give the plain one-job-per-launch server.
 * Results are identical to decoding each job alone (per-codeword decoding is
 * independent of batch composition).
 * Status codes and cuda_osd_get_last_error() as in matlab_decoder_c_api.h.
 */
#ifndef SICC_SERVE_C_API_H
#define SICC_SERVE_C_API_H

#ifdef __cplusplus
extern "C" {
#endif

/* Members are decoder handles (stream BP-OSD or BP CSR decoders); the server
 * keeps them alive.  max_jobs_per_launch <= 0: unlimited.  Besides the pool,
 * each slot allocates its launch staging at creation, at its largest size:
 * per member max_launch_codewords x (4n + k + 2) bytes of device memory (+ 4n
 * for the OSD hand-off of BP-OSD members; + 4 x edge slots x W for a member on
 * the long-code kernel, W = its workgroups per launch (1 024 for N16200, 256
 * for N64800 on 256 CUs), times 2 with flooding damping, + 4n x W when
 * n > 16 320) and max_launch_codewords x
 * (ceil(k_max / 8) + 2) + 4 x members bytes of pinned results. */
int sicc_server_create(const unsigned long long* decoders, int count, unsigned long long budget_bytes, int slots,
                       int max_jobs_per_launch, int max_launch_codewords, unsigned long long* out_handle);
int sicc_server_destroy(unsigned long long handle);

/* Queue a job.  llrs [count][n] of member `member`, host (device_ptr = 0, the
 * data is copied before the call returns) or device memory (must stay valid
 * until the job completes).  A job larger than the whole budget is rejected
 * with CUDA_OSD_STATUS_INVALID_ARGUMENT. */
int sicc_server_submit(unsigned long long handle, unsigned long long job_id, int member, const float* llrs,
                       int count, int device_ptr);

/* Ids of completed jobs (at most max_jobs), oldest completion first. */
int sicc_server_poll(unsigned long long handle, int max_jobs, unsigned long long* job_ids, int* out_count);

/* Results of a completed job (then forgotten): info [count][k], BP success and
 * OSD flags [count] (any may be null), latency from submit to completion. */
int sicc_server_fetch(unsigned long long handle, unsigned long long job_id, unsigned char* info_bits,
                      unsigned char* bp_success, unsigned char* used_osd, double* latency_us);

/* Block until every submitted job has completed (an error while held with
 * jobs pending). */
int sicc_server_drain(unsigned long long handle);

/* Batching policy.  While another launch is in flight, a new launch waits until
 * min_launch_codewords are resident or the oldest resident job has waited
 * max_wait_us; an idle GPU starts work at once.  0, 0 (the default): launch
 * whenever a slot is free. */
int sicc_server_set_batching(unsigned long long handle, int min_launch_codewords, double max_wait_us);

/* hold != 0: queue arrivals without admitting or launching them (launches in
 * flight complete); hold == 0: resume, so the queued burst is packed. */
int sicc_server_hold(unsigned long long handle, int hold);

/* Native load driver: submits job j at arrival_us[j] (microseconds after the
 * call starts, busy-waiting on a steady clock) from device LLR sources, waits
 * for all, and returns each job's latency and the wall time.  Jobs already due
 * when the driver gets to them are queued together (up to 256 under one lock),
 * as a client holding a backlog would.  Results are discarded.  For
 * benchmarks. */
int sicc_server_drive(unsigned long long handle, int n_jobs, const int* member, const int* count,
                      const unsigned long long* d_llr_ptrs, const double* arrival_us, double* latency_us,
                      double* wall_s);

typedef struct sicc_server_stats_t {
    unsigned long long jobs_submitted;
    unsigned long long jobs_completed;
    unsigned long long codewords_completed;
    unsigned long long launches;
    unsigned long long max_inflight_launches;
    unsigned long long pool_bytes;
    unsigned long long pool_high_water;
    double mean_jobs_per_launch;
} sicc_server_stats_t;
int sicc_server_stats(unsigned long long handle, sicc_server_stats_t* out);

/* Scheduler log of recent launches, for latency breakdowns: steady-clock
 * microseconds (CLOCK_MONOTONIC) at which the launch's oldest job was
 * submitted, the scheduler took the launch's jobs, it finished enqueueing the
 * launch's kernels, and it saw the launch complete; jobs, codewords and slot of
 * the launch, and the number of streams its code groups ran on (the slot's own
 * plus those of slots idle when it was issued).  The last 65 536 launches are
 * kept; reading returns up to `max` of them, oldest first, and forgets what it
 * returned. */
typedef struct sicc_server_launch_record_t {
    unsigned long long t_first_submit_us;
    unsigned long long t_taken_us;
    unsigned long long t_issued_us;
    unsigned long long t_done_us;
    int jobs;
    int codewords;
    int slot;
    int lanes;
} sicc_server_launch_record_t;
int sicc_server_launch_log(unsigned long long handle, sicc_server_launch_record_t* out, int max_records,
                           int* out_count);

#ifdef __cplusplus
}
#endif

#endif /* SICC_SERVE_C_API_H */
