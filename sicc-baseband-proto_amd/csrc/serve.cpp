// server.cpp -- continuous-batching decode server (include/sicc_serve_c_api.h).
//
// Admission: jobs wait in arrival order until their device footprint (LLRs,
// results, per-job scratch share) fits the pool; admitted jobs' LLRs are copied
// into HBM and stay there until the job completes.  Scheduling: a thread keeps
// up to `slots` launches in flight; each launch takes every admitted job not
// yet launched (bounded by jobs and codewords per launch), packs jobs of the
// same code contiguously (device-to-device), decodes each code group with its
// persistent BP(+OSD) kernels on the slot's stream, scatters results back to
// the jobs' buffers and copies them to pinned host memory.  Completion is seen
// by polling the slot's event, so jobs arriving during a launch are admitted
// and launched on the next free slot while earlier launches still run.
#include "../../include/sicc_serve_c_api.h"

#include <algorithm>
#include <cstdlib>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "decoder_handles.hpp"

namespace sicc {
namespace {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over [0, cap) with coalescing, plus reservations that
// count against the same capacity without taking address space (jobs whose
// LLRs already live in the caller's device memory).
class Pool {
public:
    explicit Pool(size_t cap) : cap_(cap) { free_[0] = cap; }
    size_t alloc(size_t bytes)
    {
        bytes = align_up(bytes);
        if (used_ + bytes > cap_) return SIZE_MAX;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const size_t off = it->first;
                const size_t rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                used_ += bytes;
                high_ = std::max(high_, used_);
                return off;
            }
        }
        return SIZE_MAX;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = align_up(bytes);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    bool reserve(size_t bytes)
    {
        bytes = align_up(bytes);
        if (used_ + bytes > cap_) return false;
        used_ += bytes;
        high_ = std::max(high_, used_);
        return true;
    }
    void unreserve(size_t bytes) { used_ -= align_up(bytes); }
    size_t capacity() const { return cap_; }
    size_t high_water() const { return high_; }

private:
    size_t cap_;
    size_t used_ = 0;
    size_t high_ = 0;
    std::map<size_t, size_t> free_;
};

struct Job {
    uint64_t id = 0;
    int member = 0;
    int count = 0;
    const float* src = nullptr;
    bool src_dev = false;
    std::vector<float> host_copy;
    uint64_t t_submit = 0, t_done = 0;
    size_t off = SIZE_MAX, bytes = 0;
    const float* d_llr = nullptr;  // HBM-resident LLRs: the caller's device rows or a pool copy
    size_t grp_base = 0;           // this job's group block in the slot's pinned results
    int row = 0;                   // first row of the job inside its group
    int grp_cw = 0;                // rows in the group
    std::vector<uint8_t> res;  // [info bit-packed, ceil(k/8) bytes per codeword | ok | used]
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    hipEvent_t gathered = nullptr;        // the launch's packed LLRs are in place
    std::vector<hipEvent_t> lane_done;    // per lane stream: the launch's last group there
    bool busy = false;     // a launch holds the slot (on the GPU, or its results being copied out)
    bool copying = false;  // the launch completed; a copier thread owns its results
    std::vector<std::shared_ptr<Job>> jobs;
    std::vector<DevBuf<float>> pack_llr;    // per member, max_cw rows
    std::vector<DevBuf<uint8_t>> pack_out;  // per member, [info | ok | used] of one launch
    std::vector<DecodeScratch> scratch;
    uint8_t* h_res = nullptr;  // pinned results of one launch (bit-packed info), written by the GPU
    size_t h_res_bytes = 0;
    GatherEntry* table = nullptr;  // pinned gather table of one launch
    PackGroup* groups = nullptr;   // pinned result-packing table, one group per member
    int table_cap = 0;
    sicc_server_launch_record_t rec{};  // the launch in flight
};

class Server {
public:
    Server(std::vector<std::shared_ptr<Decoder>> members, size_t budget, int slots, int max_jobs, int max_cw)
        : members_(std::move(members)), pool_(budget), max_jobs_(max_jobs), max_cw_(max_cw)
    {
        try {
            setup(slots, budget);
        } catch (...) {
            release();  // the destructor does not run for a constructor that throws
            throw;
        }
        thread_ = std::thread([this] { loop(); });
        for (int i = 0; i < kCopierThreads; ++i) copiers_.emplace_back([this] { copy_loop(); });
    }

    ~Server()
    {
        stop_thread();
        release();
    }

    // Stops and joins the scheduler and copier threads (idempotent).  Also run
    // at process exit for servers still open, so that no scheduler thread is
    // inside a HIP call while the runtime tears down.
    void stop_thread()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        wake_.store(true, std::memory_order_release);
        cv_.notify_all();
        cv_copy_.notify_all();
        std::lock_guard<std::mutex> jl(join_mu_);
        if (thread_.joinable()) thread_.join();
        for (auto& c : copiers_)
            if (c.joinable()) c.join();
    }

private:
    void setup(int slots, size_t budget)
    {
        known_.reserve(size_t(1) << 16);  // no rehash while tens of thousands of jobs are open
        SICC_HIP(hipGetDevice(&device_));
        SICC_HIP(hipMalloc(reinterpret_cast<void**>(&base_), budget));
        SICC_HIP(hipStreamCreateWithFlags(&admit_stream_, hipStreamNonBlocking));
        SICC_HIP(hipEventCreateWithFlags(&admit_event_, hipEventDisableTiming));
        size_t kbmax = 0;
        for (auto& m : members_) kbmax = std::max<size_t>(kbmax, (m->eng->code().k + 7) / 8);
        slots_.resize(std::max(slots, 1));
        for (auto& s : slots_) {
            SICC_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
            SICC_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
            SICC_HIP(hipEventCreateWithFlags(&s.gathered, hipEventDisableTiming));
            s.pack_llr.resize(members_.size());
            s.pack_out.resize(members_.size());
            s.scratch.resize(members_.size());
            // every group's block starts 4-byte aligned (the packing kernel's stores)
            s.h_res_bytes = static_cast<size_t>(max_cw_) * (kbmax + 2) + 4 * members_.size();
            SICC_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_res), s.h_res_bytes, hipHostMallocDefault));
            s.table_cap = max_cw_;  // at most one entry per codeword
            SICC_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.table), sizeof(GatherEntry) * s.table_cap,
                                   hipHostMallocDefault));
            SICC_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.groups), sizeof(PackGroup) * members_.size(),
                                   hipHostMallocDefault));
            // Launch staging at its largest size up front: a buffer grown in the
            // serving loop would hipFree (a device-wide wait behind every launch in
            // flight) and stall the pipeline just when launches get large.
            for (size_t m = 0; m < members_.size(); ++m) {
                Decoder& d = *members_[m];
                const LdpcCode& c = d.eng->code();
                const int osd = d.kind == Kind::Stream ? d.osd_order : -1;
                s.pack_llr[m].alloc(static_cast<size_t>(max_cw_) * c.n);
                s.pack_out[m].alloc(static_cast<size_t>(max_cw_) * (c.k + 2));
                d.eng->reserve(osd >= 0 ? max_cw_ : 0, &d.bp, &s.scratch[m]);
            }
        }
        // every slot's stream is also a lane for the code groups of any launch
        for (auto& s : slots_) {
            s.lane_done.resize(slots_.size(), nullptr);
            for (auto& e : s.lane_done) SICC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
    }

    // HIP handles and pinned buffers (device buffers free themselves).
    void release()
    {
        for (auto& s : slots_) {
            if (s.stream) (void)hipStreamSynchronize(s.stream);
            if (s.h_res) (void)hipHostFree(s.h_res);
            if (s.table) (void)hipHostFree(s.table);
            if (s.groups) (void)hipHostFree(s.groups);
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.gathered) (void)hipEventDestroy(s.gathered);
            for (auto e : s.lane_done)
                if (e) (void)hipEventDestroy(e);
            if (s.stream) (void)hipStreamDestroy(s.stream);
        }
        if (admit_stream_) {
            (void)hipStreamSynchronize(admit_stream_);
            (void)hipStreamDestroy(admit_stream_);
        }
        if (admit_event_) (void)hipEventDestroy(admit_event_);
        if (base_) (void)hipFree(base_);
        slots_.clear();
        admit_stream_ = nullptr;
        admit_event_ = nullptr;
        base_ = nullptr;
    }

public:

    // Device memory a job holds from admission to completion: its LLR rows and
    // its share of the packed launch (LLR rows, results, OSD hand-off rows).
    size_t footprint(int member, int count) const
    {
        const LdpcCode& c = members_[member]->eng->code();
        const size_t n = static_cast<size_t>(count);
        size_t b = 2 * align_up(n * c.n * sizeof(float)) + align_up(n * (c.k + 2));
        if (members_[member]->kind == Kind::Stream && members_[member]->osd_order >= 0) b += align_up(n * c.n * 4);
        return b;
    }

    // A validated job, not yet visible to the scheduler (no lock taken).
    std::shared_ptr<Job> make_job(uint64_t id, int member, const float* llr, int count, bool dev)
    {
        if (member < 0 || member >= static_cast<int>(members_.size())) throw std::invalid_argument("submit: bad member");
        if (!llr || count <= 0) throw std::invalid_argument("submit: empty job");
        if (count > max_cw_) throw std::invalid_argument("submit: job larger than max_launch_codewords");
        if (footprint(member, count) > pool_.capacity())
            throw std::invalid_argument("submit: job larger than the device-memory budget");
        auto j = std::make_shared<Job>();
        j->id = id;
        j->member = member;
        j->count = count;
        j->src_dev = dev;
        if (dev) {
            j->src = llr;
        } else {
            const size_t n = static_cast<size_t>(count) * members_[member]->eng->code().n;
            j->host_copy.assign(llr, llr + n);
            j->src = j->host_copy.data();
        }
        j->t_submit = now_us();
        return j;
    }

    // Queue jobs under one lock acquisition (all or none: a duplicate id rejects
    // the group).
    // Submission never takes the scheduler's lock (mu_): a client submitting in
    // a tight loop re-took mu_ faster than the woken scheduler could, which kept
    // the scheduler off the lock, and the GPU idle, for as long as the client
    // kept submitting.  Ids go into known_ under ids_mu_, the jobs into the
    // inbox under inbox_mu_ (held for the pointer pushes only), and the
    // scheduler empties the inbox in admit().
    void enqueue(const std::shared_ptr<Job>* jobs, size_t count)
    {
        if (count == 0) return;
        {
            std::lock_guard<std::mutex> lk(ids_mu_);
            // one hash insert per job; a duplicate (already known, or twice in the
            // group) takes the group's earlier inserts back out
            for (size_t i = 0; i < count; ++i) {
                if (!known_.emplace(jobs[i]->id, jobs[i]).second) {
                    for (size_t q = 0; q < i; ++q) known_.erase(jobs[q]->id);
                    throw std::invalid_argument("submit: job id already in use");
                }
            }
        }
        {
            std::lock_guard<std::mutex> lk(inbox_mu_);
            for (size_t i = 0; i < count; ++i) inbox_.push_back(jobs[i]);
            arrived_.fetch_add(count, std::memory_order_seq_cst);
        }
        // Only a scheduler asleep on cv_ needs the notify, and then under mu_:
        // it sets sleeping_ before it tests for arrivals, so either that test
        // sees this group or this load sees sleeping_ and the notify follows its
        // wait (which holds mu_ until it blocks).
        if (sleeping_.load(std::memory_order_seq_cst)) {
            std::lock_guard<std::mutex> lk(mu_);
            cv_.notify_all();
        }
    }

    void submit(uint64_t id, int member, const float* llr, int count, bool dev)
    {
        const auto j = make_job(id, member, llr, count, dev);
        enqueue(&j, 1);
    }

    int poll(int max, unsigned long long* ids)
    {
        std::lock_guard<std::mutex> lk(mu_);
        int got = 0;
        while (got < max && !completed_.empty()) {
            ids[got++] = completed_.front();
            completed_.pop_front();
        }
        return got;
    }

    void fetch(uint64_t id, uint8_t* info, uint8_t* ok, uint8_t* used, double* lat)
    {
        std::shared_ptr<Job> j;
        {
            std::lock_guard<std::mutex> lk(mu_);  // t_done is written under mu_
            std::lock_guard<std::mutex> il(ids_mu_);
            auto it = known_.find(id);
            if (it == known_.end() || it->second->t_done == 0) throw std::invalid_argument("fetch: job not completed");
            j = it->second;
            known_.erase(it);
        }
        const int k = members_[j->member]->eng->code().k;
        const size_t kb = (static_cast<size_t>(k) + 7) / 8;
        const size_t cnt = static_cast<size_t>(j->count);
        if (info) {
            // unpack: bit i of a codeword sits at byte i/8, bit i%8
            for (size_t r = 0; r < cnt; ++r) {
                const uint8_t* src = j->res.data() + r * kb;
                uint8_t* dst = info + r * k;
                for (int i = 0; i < k; ++i) dst[i] = (src[i >> 3] >> (i & 7)) & 1u;
            }
        }
        if (ok) std::memcpy(ok, j->res.data() + cnt * kb, cnt);
        if (used) std::memcpy(used, j->res.data() + cnt * (kb + 1), cnt);
        if (lat) *lat = static_cast<double>(j->t_done - j->t_submit);
        std::lock_guard<std::mutex> lk(mu_);
        recycle(*j);
    }

    // While held, arrivals queue and nothing new is admitted or launched
    // (launches in flight still complete); releasing lets the scheduler pack
    // the queued burst.
    void hold(bool on)
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            hold_ = on;
        }
        wake_.store(true, std::memory_order_release);
        cv_.notify_all();
    }

    void set_batching(int min_cw, double max_wait_us)
    {
        if (min_cw < 0 || max_wait_us < 0.0) throw std::invalid_argument("set_batching: negative limit");
        {
            std::lock_guard<std::mutex> lk(mu_);
            min_cw_ = min_cw;
            max_wait_us_ = max_wait_us;
        }
        wake_.store(true, std::memory_order_release);
        cv_.notify_all();
    }

    void drain()
    {
        std::unique_lock<std::mutex> lk(mu_);
        if (hold_ && st_.jobs_completed != arrived_.load())
            throw std::invalid_argument("drain: the server is held with jobs pending");
        cv_done_.wait(lk, [&] { return st_.jobs_completed == arrived_.load() || failed_; });
        if (failed_) throw std::runtime_error("server: " + fail_msg_);
    }

    double job_latency_and_forget(uint64_t id)
    {
        std::lock_guard<std::mutex> lk(mu_);
        std::lock_guard<std::mutex> il(ids_mu_);
        auto it = known_.find(id);
        if (it == known_.end()) return -1.0;
        const double l = static_cast<double>(it->second->t_done - it->second->t_submit);
        recycle(*it->second);
        known_.erase(it);
        return l;
    }

    void clear_completed()
    {
        std::lock_guard<std::mutex> lk(mu_);
        completed_.clear();
    }

    sicc_server_stats_t stats()
    {
        std::lock_guard<std::mutex> lk(mu_);
        sicc_server_stats_t s = st_;
        s.jobs_submitted = arrived_.load();
        s.pool_bytes = pool_.capacity();
        s.pool_high_water = pool_.high_water();
        s.mean_jobs_per_launch = st_.launches ? static_cast<double>(st_.jobs_completed) / st_.launches : 0.0;
        return s;
    }

    int device() const { return device_; }

    // Keep a released job's result buffer for a later job (lock held), up to
    // kSpareBytes in total.
    void recycle(Job& j)
    {
        const size_t cap = j.res.capacity();
        if (cap == 0 || spare_bytes_ + cap > kSpareBytes) return;
        int c = kMinClass;
        while (c + 1 < kMinClass + kClasses && (size_t(1) << (c + 1)) <= cap) ++c;
        if ((size_t(1) << c) > cap) return;  // smaller than the smallest class
        spare_bytes_ += cap;
        spare_[c - kMinClass].push_back(std::move(j.res));
        j.res = std::vector<uint8_t>();
    }

    int launch_log(sicc_server_launch_record_t* out, int max)
    {
        std::lock_guard<std::mutex> lk(mu_);
        int got = 0;
        while (got < max && !log_.empty()) {
            out[got++] = log_.front();
            log_.pop_front();
        }
        return got;
    }

private:
    void loop()
    {
        try {
            SICC_HIP(hipSetDevice(device_));
            std::unique_lock<std::mutex> lk(mu_);
            while (!stop_) {
                bool progress = harvest(lk);
                if (!hold_) {
                    progress |= admit();
                    progress |= launch(lk);
                }
                const bool inflight = std::any_of(slots_.begin(), slots_.end(), [](const Slot& s) { return s.busy; });
                if (!progress) {
                    if (inflight || (!hold_ && (pending_arrivals() || !resident_.empty()))) {
                        // Wait for a launch to finish, a new job, or the oldest
                        // resident job's batching deadline -- without the lock, so
                        // submitters are not starved by this thread's polling.
                        // jobs may arrive after admit() emptied the inbox: wake for
                        // any job not yet taken (while held, for new ones only)
                        const uint64_t seen = hold_ ? arrived_.load(std::memory_order_acquire) : moved_;
                        const uint64_t copies = copies_done_.load(std::memory_order_acquire);
                        polling_.clear();
                        for (auto& s : slots_)
                            if (s.busy && !s.copying) polling_.push_back(s.done);
                        const bool slot_free =
                            std::any_of(slots_.begin(), slots_.end(), [](const Slot& s) { return !s.busy; });
                        const uint64_t deadline = (resident_.empty() || !slot_free)
                                                      ? UINT64_MAX
                                                      : resident_.front()->t_submit +
                                                            static_cast<uint64_t>(max_wait_us_);
                        lk.unlock();
                        wait_event(seen, copies, deadline);
                        lk.lock();
                    } else {
                        sleeping_.store(true, std::memory_order_seq_cst);
                        cv_.wait(lk, [&] { return stop_ || (!hold_ && (pending_arrivals() || !resident_.empty())); });
                        sleeping_.store(false, std::memory_order_relaxed);
                    }
                }
            }
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> g(mu_);
            failed_ = true;
            fail_msg_ = e.what();
            cv_done_.notify_all();
        }
    }

    // Lock not held: polls the completion events of the launches still on the
    // GPU (polling_, taken under the lock) until one completes, a job arrives,
    // a copier frees a slot, the batching deadline passes or wake_ is set.
    void wait_event(uint64_t seen, uint64_t copies, uint64_t deadline)
    {
        for (;;) {
            if (arrived_.load(std::memory_order_acquire) != seen) return;
            if (copies_done_.load(std::memory_order_acquire) != copies) return;
            for (hipEvent_t e : polling_)
                if (hipEventQuery(e) != hipErrorNotReady) return;
            if (now_us() >= deadline) return;
            if (wake_.exchange(false, std::memory_order_acq_rel)) return;  // stop, hold or policy change
            std::this_thread::yield();
        }
    }

    // Completed launches go to the copier threads (lock held): the scheduler goes
    // straight on to admission and the next launches.
    bool harvest(std::unique_lock<std::mutex>&)
    {
        bool any = false;
        for (auto& s : slots_) {
            if (!s.busy || s.copying) continue;
            const hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) throw_hip(q, "server launch");
            s.copying = true;
            s.rec.t_done_us = now_us();
            to_copy_.push_back(&s);
            any = true;
        }
        if (any) cv_copy_.notify_one();
        return any;
    }

    // Copier thread: each completed launch's results leave the slot's pinned
    // buffer (written by the packing kernel) into the jobs' own buffers with the
    // lock released, then the jobs complete and the slot is free again.  Done
    // here, the copy (100-300 us for two 65 536-codeword launches) no longer
    // delays the scheduler's polling and launches.
    void copy_loop()
    {
        try {
            std::unique_lock<std::mutex> lk(mu_);
            for (;;) {
                cv_copy_.wait(lk, [&] { return stop_ || !to_copy_.empty(); });
                if (to_copy_.empty()) break;  // stopped with nothing left to copy
                Slot* s = to_copy_.front();
                to_copy_.pop_front();
                for (auto& j : s->jobs) take_spare(*j, result_bytes(*j));
                lk.unlock();
                for (auto& j : s->jobs) {
                    const size_t kb = (static_cast<size_t>(members_[j->member]->eng->code().k) + 7) / 8;
                    const size_t n = static_cast<size_t>(j->count);
                    const size_t row = static_cast<size_t>(j->row);
                    const size_t gcw = static_cast<size_t>(j->grp_cw);
                    const uint8_t* g = s->h_res + j->grp_base;
                    const size_t need = n * (kb + 2);
                    if (j->res.capacity() < need) j->res.reserve(size_t(1) << size_class(need));
                    j->res.resize(need);
                    uint8_t* o = j->res.data();
                    std::memcpy(o, g + row * kb, n * kb);
                    std::memcpy(o + n * kb, g + gcw * kb + row, n);
                    std::memcpy(o + n * (kb + 1), g + gcw * (kb + 1) + row, n);
                }
                const uint64_t t = now_us();
                lk.lock();
                if (log_.size() >= kLogRecords) log_.pop_front();
                log_.push_back(s->rec);
                for (auto& j : s->jobs) {
                    j->t_done = t;
                    if (j->off == SIZE_MAX)
                        pool_.unreserve(j->bytes);
                    else
                        pool_.release(j->off, j->bytes);
                    j->host_copy = std::vector<float>();
                    completed_.push_back(j->id);
                    ++st_.jobs_completed;
                    st_.codewords_completed += static_cast<unsigned long long>(j->count);
                }
                s->jobs.clear();
                s->copying = false;
                s->busy = false;
                copies_done_.fetch_add(1, std::memory_order_acq_rel);
                cv_done_.notify_all();
            }
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> g(mu_);
            failed_ = true;
            fail_msg_ = e.what();
            cv_done_.notify_all();
        }
    }

    // Bytes of a job's results: bit-packed information, then BP-success and OSD flags.
    size_t result_bytes(const Job& j) const
    {
        const size_t kb = (static_cast<size_t>(members_[j.member]->eng->code().k) + 7) / 8;
        return static_cast<size_t>(j.count) * (kb + 2);
    }

    // Result buffers are recycled by power-of-two capacity: a fresh heap buffer
    // per job costs page faults and zero-filling (about 1 us a job, 0.5-0.8 ms for
    // two 65 536-codeword launches), and a buffer smaller than the job would be
    // reallocated.
    static int size_class(size_t bytes)
    {
        int c = kMinClass;
        while (c < kMinClass + kClasses - 1 && (size_t(1) << c) < bytes) ++c;
        return c;
    }

    // Lock held: a recycled buffer of the job's class (or one or two classes up).
    void take_spare(Job& j, size_t need)
    {
        if (j.res.capacity() >= need) return;
        const int c = size_class(need);
        for (int cc = c; cc < std::min(c + 3, kMinClass + kClasses); ++cc) {
            auto& v = spare_[cc - kMinClass];
            if (v.empty() || v.back().capacity() < need) continue;
            spare_bytes_ -= v.back().capacity();
            j.res = std::move(v.back());
            v.pop_back();
            return;
        }
    }

    // Jobs submitted and not yet taken from the inbox, or taken and not admitted
    // (lock held).
    bool pending_arrivals() const
    {
        return !arrivals_.empty() || arrived_.load(std::memory_order_seq_cst) != moved_;
    }

    bool admit()
    {
        bool any = false, copied = false;
        if (arrived_.load(std::memory_order_acquire) != moved_) {
            std::lock_guard<std::mutex> il(inbox_mu_);
            moved_ += inbox_.size();
            for (auto& j : inbox_) arrivals_.push_back(std::move(j));
            inbox_.clear();
        }
        while (!arrivals_.empty()) {
            auto j = arrivals_.front();
            const size_t bytes = footprint(j->member, j->count);
            // in order: later jobs wait behind the head.  A device-sourced job only
            // reserves its share of the budget; a host job gets pool memory for its
            // LLR copy.
            const size_t off = j->src_dev ? SIZE_MAX : pool_.alloc(bytes);
            if (j->src_dev ? !pool_.reserve(bytes) : off == SIZE_MAX) break;
            arrivals_.pop_front();
            j->off = off;
            j->bytes = bytes;
            if (j->src_dev) {
                j->d_llr = j->src;  // already resident in HBM
            } else {
                const LdpcCode& c = members_[j->member]->eng->code();
                float* dst = reinterpret_cast<float*>(base_ + off);
                SICC_HIP(hipMemcpyAsync(dst, j->src, static_cast<size_t>(j->count) * c.n * sizeof(float),
                                        hipMemcpyHostToDevice, admit_stream_));
                j->d_llr = dst;
                copied = true;
            }
            resident_.push_back(j);
            any = true;
        }
        if (copied) {
            SICC_HIP(hipEventRecord(admit_event_, admit_stream_));
            admit_pending_ = true;
        }
        return any;
    }

    // Batching policy: with another launch in flight, a launch waits until
    // min_cw_ codewords are resident or the oldest resident job has waited
    // max_wait_us_; with the GPU idle it starts at once.
    bool launch_ready() const
    {
        if (resident_.empty()) return false;
        if (min_cw_ <= 0) return true;
        if (std::none_of(slots_.begin(), slots_.end(), [](const Slot& s) { return s.busy; })) return true;
        long long cw = 0;
        for (const auto& j : resident_) {
            cw += j->count;
            if (cw >= min_cw_) return true;
        }
        return static_cast<double>(now_us() - resident_.front()->t_submit) >= max_wait_us_;
    }

    // Called with the lock held; the slot's jobs are taken under the lock, the
    // packing and kernel launches run with it released (only this thread touches
    // slots and the jobs' packing fields), so submit/poll/fetch callers do not
    // wait behind the launch work.
    bool launch(std::unique_lock<std::mutex>& lk)
    {
        bool any = false;
        for (auto& s : slots_) {
            if (s.busy || resident_.empty()) continue;
            if (!launch_ready()) break;
            // take jobs in admission order within the launch limits
            int total = 0;
            while (!resident_.empty()) {
                auto& j = resident_.front();
                if (max_jobs_ > 0 && static_cast<int>(s.jobs.size()) >= max_jobs_) break;
                if (total + j->count > max_cw_) break;
                total += j->count;
                s.jobs.push_back(j);
                resident_.pop_front();
            }
            if (s.jobs.empty()) continue;
            const bool wait_admit = admit_pending_;
            s.rec = sicc_server_launch_record_t{};
            s.rec.t_first_submit_us = s.jobs.front()->t_submit;
            s.rec.t_taken_us = now_us();
            s.rec.jobs = static_cast<int>(s.jobs.size());
            s.rec.codewords = total;
            s.rec.slot = static_cast<int>(&s - slots_.data());
            lk.unlock();
            try {
                issue(s, wait_admit);
            } catch (...) {
                lk.lock();
                throw;
            }
            s.rec.t_issued_us = now_us();
            lk.lock();
            s.busy = true;
            ++st_.launches;
            const auto inflight = static_cast<unsigned long long>(
                std::count_if(slots_.begin(), slots_.end(), [](const Slot& x) { return x.busy; }));
            st_.max_inflight_launches = std::max(st_.max_inflight_launches, inflight);
            any = true;
        }
        return any;
    }

    // Pack, decode and copy back one launch on the slot's stream (lock not held).
    void issue(Slot& s, bool wait_admit)
    {
        {
            if (wait_admit) SICC_HIP(hipStreamWaitEvent(s.stream, admit_event_, 0));
            auto& by = by_;  // per member, reused across launches
            by.resize(members_.size());
            for (auto& v : by) v.clear();
            for (auto& j : s.jobs) by[j->member].push_back(j.get());
            // rows of every job in its code group, and one gather launch for every
            // group of more than one job (a single job is decoded in place)
            int tbl = 0;
            for (size_t m = 0; m < by.size(); ++m) {
                if (by[m].empty()) continue;
                const int n = members_[m]->eng->code().n;
                int cw = 0;
                for (Job* j : by[m]) {
                    j->row = cw;
                    cw += j->count;
                }
                if (by[m].size() > 1) {
                    // one gather workgroup per chunk of about 4096 LLRs, so a few
                    // large jobs still spread over the whole GPU
                    const int ch = std::max(1, 4096 / n);
                    float* dst = s.pack_llr[m].get();
                    for (Job* j : by[m])
                        for (int r = 0; r < j->count; r += ch)
                            s.table[tbl++] = {j->d_llr + static_cast<size_t>(r) * n,
                                              dst + static_cast<size_t>(j->row + r) * n, std::min(ch, j->count - r), n};
                }
            }
            launch_gather_rows(s.table, tbl, s.stream);
            // Decode units: the code groups whose BP batches share one instance of
            // the flooding kernel form one fused launch (their slowest codewords
            // then overlap instead of running one kernel after another); every
            // other group (OSD, long-code kernel, work stealing) is a unit of its
            // own.  With several units they are dealt over the slot's own stream
            // and the streams of the slots that are idle (one hardware queue
            // each), largest first onto the least loaded lane (edges x
            // codewords); with every slot busy the launch stays on its own stream,
            // the queues being full anyway.  Each lane waits for the gather; the
            // slot's stream joins every lane before packing the results.
            const int nslots = static_cast<int>(slots_.size());
            const int self = static_cast<int>(&s - slots_.data());
            const size_t nm = members_.size();
            g_out_.resize(nm);
            g_llr_.resize(nm);
            g_cw_.resize(nm);
            g_batch_.resize(nm);
            size_t res = 0;
            int ng = 0;
            int nu = 0;
            for (size_t m = 0; m < nm; ++m) {
                if (by[m].empty()) continue;
                Decoder& d = *members_[m];
                const LdpcCode& c = d.eng->code();
                const int osd = d.kind == Kind::Stream ? d.osd_order : -1;
                const int cw = by[m].back()->row + by[m].back()->count;
                const float* llr = by[m].size() == 1 ? by[m][0]->d_llr : s.pack_llr[m].get();
                uint8_t* o = s.pack_out[m].get();
                DecodeOut out;
                out.info = o;
                out.bp_ok = o + static_cast<size_t>(cw) * c.k;
                out.used_osd = o + static_cast<size_t>(cw) * (c.k + 1);
                g_out_[m] = out;
                g_llr_[m] = llr;
                g_cw_[m] = cw;
                int key = -1;
                if (osd < 0 && d.eng->bp_batch(llr, cw, d.bp, out, &g_batch_[m]))
                    key = bp_flood_instance(g_batch_[m].code, g_batch_[m].p, g_batch_[m].tune);
                int u = 0;
                if (key >= 0)
                    while (u < nu && units_[u].key != key) ++u;
                else
                    u = nu;
                if (u == nu) {
                    if (static_cast<int>(units_.size()) <= nu) units_.emplace_back();
                    units_[nu].key = key;
                    units_[nu].cost = 0;
                    units_[nu].members.clear();
                    ++nu;
                }
                units_[u].members.push_back(static_cast<int>(m));
                units_[u].cost += std::max(1LL, static_cast<long long>(cw) * c.edges());
                // results leave bit-packed (8x less to move and to read on the host),
                // ok / used flags behind them, written by the packing kernel straight
                // into the slot's pinned results: no copy-engine transfer, whose D2H
                // submissions stalled the host for milliseconds at times
                s.groups[ng++] = {o, out.bp_ok, s.h_res + res, cw, c.k, 2 * cw, 0, 0};
                for (Job* j : by[m]) {
                    j->grp_base = res;
                    j->grp_cw = cw;
                }
                res = (res + static_cast<size_t>(cw) * ((c.k + 7) / 8 + 2) + 3) & ~static_cast<size_t>(3);
            }
            lanes_.assign(1, self);
            for (int l = 0; l < nslots; ++l)
                if (l != self && !slots_[l].busy) lanes_.push_back(l);
            order_.resize(nu);
            for (int u = 0; u < nu; ++u) order_[u] = u;
            const bool spread = nu > 1 && lanes_.size() > 1;
            if (spread) {
                std::stable_sort(order_.begin(), order_.end(),
                                 [&](int x, int y) { return units_[x].cost > units_[y].cost; });
                SICC_HIP(hipEventRecord(s.gathered, s.stream));
            }
            lane_load_.assign(lanes_.size(), 0);
            for (int ui : order_) {
                const Unit& un = units_[ui];
                size_t li = 0;
                if (spread)
                    for (size_t j = 1; j < lanes_.size(); ++j)
                        if (lane_load_[j] < lane_load_[li]) li = j;
                const int lane = lanes_[li];
                hipStream_t ls = slots_[lane].stream;
                if (lane != self && lane_load_[li] == 0) SICC_HIP(hipStreamWaitEvent(ls, s.gathered, 0));
                lane_load_[li] += un.cost;
                if (un.key >= 0 && un.members.size() > 1) {
                    fused_.clear();
                    for (int m : un.members) fused_.push_back(g_batch_[m]);
                    launch_bp_flood_multi(fused_.data(), static_cast<int>(fused_.size()), ls);
                } else {
                    for (int m : un.members) {
                        Decoder& d = *members_[m];
                        const int osd = d.kind == Kind::Stream ? d.osd_order : -1;
                        d.eng->decode(g_llr_[m], g_cw_[m], d.bp, osd, d.osd, g_out_[m], ls, &s.scratch[m]);
                    }
                }
            }
            for (size_t j = 1; j < lanes_.size(); ++j)
                if (lane_load_[j]) {
                    const int l = lanes_[j];
                    SICC_HIP(hipEventRecord(s.lane_done[l], slots_[l].stream));
                    SICC_HIP(hipStreamWaitEvent(s.stream, s.lane_done[l], 0));
                }
            s.rec.lanes = spread ? static_cast<int>(std::count_if(lane_load_.begin(), lane_load_.end(),
                                                                  [](long long x) { return x > 0; }))
                                 : 1;
            launch_pack_groups(s.groups, ng, s.stream);
            SICC_HIP(hipEventRecord(s.done, s.stream));
        }
    }

    std::vector<std::shared_ptr<Decoder>> members_;
    Pool pool_;
    int max_jobs_;
    int max_cw_;
    int device_ = 0;
    char* base_ = nullptr;
    hipStream_t admit_stream_ = nullptr;
    hipEvent_t admit_event_ = nullptr;
    bool admit_pending_ = false;
    std::vector<Slot> slots_;
    std::vector<std::vector<Job*>> by_;       // issue's per-member job lists
    // issue's decode units: fused BP launches (key = kernel instance) or single groups
    struct Unit {
        int key = -1;
        long long cost = 0;
        std::vector<int> members;
    };
    std::vector<Unit> units_;
    std::vector<int> order_;                  // issue's units, largest first
    std::vector<int> lanes_;                  // issue's lanes: its own slot, then idle slots
    std::vector<long long> lane_load_;        // issue's work dealt to each lane
    std::vector<DecodeOut> g_out_;            // issue's per-member outputs
    std::vector<const float*> g_llr_;         // issue's per-member LLR rows
    std::vector<int> g_cw_;                   // issue's per-member codeword counts
    std::vector<BpBatch> g_batch_;            // issue's per-member BP batches
    std::vector<BpBatch> fused_;              // issue's batches of one fused launch
    std::mutex mu_;
    std::condition_variable cv_, cv_done_;
    std::deque<std::shared_ptr<Job>> arrivals_, resident_;  // admission order (under mu_)
    std::mutex inbox_mu_;                                   // guards inbox_ (taken after mu_)
    std::deque<std::shared_ptr<Job>> inbox_;                // submitted, not yet taken by the scheduler
    uint64_t moved_ = 0;                                    // jobs taken from the inbox (under mu_)
    std::atomic<bool> sleeping_{false};                     // the scheduler waits on cv_ for arrivals
    std::mutex ids_mu_;                                     // guards known_ (taken after mu_)
    std::deque<unsigned long long> completed_;
    static constexpr size_t kLogRecords = 65536;
    std::deque<sicc_server_launch_record_t> log_;  // completed launches, oldest first
    std::unordered_map<uint64_t, std::shared_ptr<Job>> known_;
    static constexpr size_t kSpareBytes = size_t(256) << 20;
    static constexpr int kMinClass = 8, kClasses = 24;          // 256 B .. 2 GiB
    std::vector<std::vector<uint8_t>> spare_[kClasses];         // recycled result buffers by capacity class
    size_t spare_bytes_ = 0;
    std::deque<Slot*> to_copy_;                                 // completed launches for the copiers
    std::condition_variable cv_copy_;
    std::atomic<uint64_t> copies_done_{0};                      // launches the copiers have completed
    std::vector<hipEvent_t> polling_;                           // scheduler: events of launches on the GPU
    // A launch's slot stays busy until its results are copied out (about 50 us of
    // host memcpy per 65 536-codeword launch); two copiers take completed
    // launches in parallel so that the copies do not queue behind each other.
    static constexpr int kCopierThreads = 2;
    std::vector<std::thread> copiers_;
    sicc_server_stats_t st_{};
    bool stop_ = false;
    bool hold_ = false;
    std::atomic<uint64_t> arrived_{0};  // jobs submitted so far (read without the lock)
    std::atomic<bool> wake_{false};      // stop / hold / batching change for the unlocked wait
    int min_cw_ = 0;
    double max_wait_us_ = 0.0;
    bool failed_ = false;
    std::string fail_msg_;
    std::mutex join_mu_;
    std::thread thread_;
};

HandleTable<Server>& stable()
{
    // never destroyed: handles still open at exit must not reach the HIP
    // runtime from static destructors
    static HandleTable<Server>& t = *new HandleTable<Server>;
    return t;
}

// Registered after the first server exists, hence after the HIP runtime set up
// its own exit handlers: runs before them and stops every open server's
// scheduler thread (their device resources are left to the process exit).
void stop_open_servers() { stable().for_each([](Server& s) { s.stop_thread(); }); }

void register_exit_stop()
{
    static std::once_flag once;
    std::call_once(once, [] { std::atexit(stop_open_servers); });
}

std::shared_ptr<Server> slookup(unsigned long long h)
{
    auto s = stable().get(h);
    if (!s) throw std::invalid_argument("server: unknown handle");
    SICC_HIP(hipSetDevice(s->device()));
    return s;
}

}  // namespace
}  // namespace sicc

using namespace sicc;

extern "C" {

int sicc_server_create(const unsigned long long* decoders, int count, unsigned long long budget_bytes, int slots,
                       int max_jobs_per_launch, int max_launch_codewords, unsigned long long* out_handle)
{
    return guarded([&] {
        if (!out_handle || !decoders || count <= 0) throw std::invalid_argument("server_create: no decoders");
        *out_handle = 0;
        if (budget_bytes < (1u << 20)) throw std::invalid_argument("server_create: budget below 1 MiB");
        if (slots <= 0 || max_launch_codewords <= 0) throw std::invalid_argument("server_create: bad limits");
        std::vector<std::shared_ptr<Decoder>> members;
        for (int i = 0; i < count; ++i) {
            auto d = table().get(decoders[i]);
            if (!d || d->kind == Kind::OsdOnly) throw std::invalid_argument("server_create: members must be BP decoders");
            members.push_back(d);
        }
        *out_handle = stable().add(std::make_unique<Server>(std::move(members), static_cast<size_t>(budget_bytes), slots,
                                                            max_jobs_per_launch, max_launch_codewords));
        register_exit_stop();
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_server_destroy(unsigned long long handle)
{
    return guarded([&] {
        if (!stable().remove(handle)) throw std::invalid_argument("server: unknown handle");
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_server_submit(unsigned long long handle, unsigned long long job_id, int member, const float* llrs, int count,
                       int device_ptr)
{
    return guarded([&] {
        slookup(handle)->submit(job_id, member, llrs, count, device_ptr != 0);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_server_poll(unsigned long long handle, int max_jobs, unsigned long long* job_ids, int* out_count)
{
    return guarded([&] {
        if (!out_count || (max_jobs > 0 && !job_ids)) throw std::invalid_argument("server_poll: null output");
        *out_count = max_jobs > 0 ? slookup(handle)->poll(max_jobs, job_ids) : 0;
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_server_fetch(unsigned long long handle, unsigned long long job_id, unsigned char* info_bits,
                      unsigned char* bp_success, unsigned char* used_osd, double* latency_us)
{
    return guarded([&] {
        slookup(handle)->fetch(job_id, info_bits, bp_success, used_osd, latency_us);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_server_drain(unsigned long long handle)
{
    return guarded([&] {
        slookup(handle)->drain();
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_server_set_batching(unsigned long long handle, int min_launch_codewords, double max_wait_us)
{
    return guarded([&] {
        slookup(handle)->set_batching(min_launch_codewords, max_wait_us);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_server_hold(unsigned long long handle, int hold)
{
    return guarded([&] {
        slookup(handle)->hold(hold != 0);
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_server_drive(unsigned long long handle, int n_jobs, const int* member, const int* count,
                      const unsigned long long* d_llr_ptrs, const double* arrival_us, double* latency_us,
                      double* wall_s)
{
    return guarded([&] {
        auto s = slookup(handle);
        if (n_jobs <= 0 || !member || !count || !d_llr_ptrs || !arrival_us)
            throw std::invalid_argument("server_drive: empty schedule");
        const uint64_t base_id = 0x8000000000000000ull + (now_us() << 20);
        const auto t0 = std::chrono::steady_clock::now();
        auto due_at = [&](int j) { return t0 + std::chrono::nanoseconds(static_cast<long long>(arrival_us[j] * 1000.0)); };
        // every job already due goes to the server in one group (one lock
        // acquisition), as a client thread holding a backlog would submit it
        std::vector<std::shared_ptr<Job>> group;
        for (int j = 0; j < n_jobs;) {
            while (std::chrono::steady_clock::now() < due_at(j)) {
            }
            const auto now = std::chrono::steady_clock::now();
            group.clear();
            while (j < n_jobs && group.size() < 256 && due_at(j) <= now) {
                group.push_back(s->make_job(base_id + j, member[j], reinterpret_cast<const float*>(d_llr_ptrs[j]),
                                            count[j], true));
                ++j;
            }
            s->enqueue(group.data(), group.size());
        }
        s->drain();
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (int j = 0; j < n_jobs; ++j) {
            const double l = s->job_latency_and_forget(base_id + j);
            if (latency_us) latency_us[j] = l;
        }
        s->clear_completed();
        if (wall_s) *wall_s = wall;
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_server_launch_log(unsigned long long handle, sicc_server_launch_record_t* out, int max_records,
                           int* out_count)
{
    return guarded([&] {
        if (!out_count || (max_records > 0 && !out)) throw std::invalid_argument("server_launch_log: null output");
        *out_count = max_records > 0 ? slookup(handle)->launch_log(out, max_records) : 0;
        return CUDA_OSD_STATUS_OK;
    });
}

int sicc_server_stats(unsigned long long handle, sicc_server_stats_t* out)
{
    return guarded([&] {
        if (!out) throw std::invalid_argument("server_stats: null output");
        *out = slookup(handle)->stats();
        return CUDA_OSD_STATUS_OK;
    });
}

}  // extern "C"
