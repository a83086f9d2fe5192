// ctd_serve.hpp -- serving form: continuous batching of jobs as they arrive.
//
// Included by ctd_engine.hip inside its internal namespace.  A server owns a
// scheduler thread, S streams with one resident prime slot each, and the
// engine's resident sieving base.  Jobs are submitted with caller-owned host
// factors buffers and come back through poll/wait.
//
// Scheduler loop (one host thread, no device-side waits):
//   admit   pending jobs while their buffers fit in the memory budget and the
//           streams need work.  A job is charged the power-of-two size classes
//           of its device factors buffer and, separately, of its pinned
//           staging buffer; an idle server always admits the next job.  Then:
//           allocate the device buffers, copy the host buffers in, run
//           the mod-2 kernel, and cut the prime window into tiles of about
//           2^35 VALU lane-instructions, into at most about 32 tiles (at
//           most tile_length integers each);
//   issue   each stream keeps up to SERVE_STREAM_WORK of estimated work in
//           flight (at least one tile, at most SERVE_DEPTH): enough that it
//           never runs dry between scheduler passes, little enough that a
//           new job does not queue behind much committed work.  Free places
//           alternate between the oldest admitted job that still has tiles
//           and the one with the least remaining work (pick_job; with
//           CTD_SERVE_FIFO always the oldest); one job can spread over every
//           stream;
//   retire  a tile is done when its event has completed; a job whose tiles are
//           all done is copied back to its host buffers on the copy stream and
//           completes when that copy has, which frees its device memory.
// Results travel device -> pinned staging buffer (asynchronous) -> the
// caller's buffers.  The last copy is done by the thread that calls
// ctd_hip_server_poll / ctd_hip_server_wait (or close): a copy into pageable
// memory would block the scheduler thread until the DMA finishes, and no tile
// could be issued meanwhile.  Only when the staging buffers of uncollected
// results exceed the budget does the scheduler copy a result out itself, so
// pinned memory stays bounded whatever the caller does.  Latency is
// measured from submit to the moment the results are back in host memory (the
// staging copy has completed), whenever the caller collects them.
// Residues, hits and factor buffers are those of the plain form: the same
// kernels, the same tiles, only the interleaving differs.

constexpr uint64_t SERVE_TILE_DEFAULT = 1ull << 26;
// work = estimated updates x the family's VALU instructions per update
constexpr double SERVE_TILE_WORK = 34359738368.0; // 2^35 lane-instructions: ~0.9 ms of the whole device
// a job's window is cut into at most about this many tiles: a job with much
// more work than SERVE_TILE_WORK per tile (a 2^30-integer window with many
// indices) gets longer tiles instead of more of them, since a tile with few
// primes splits its index range into chunks that each pay the per-prime
// set-up again
constexpr double SERVE_JOB_TILES = 32.0;
constexpr uint64_t SERVE_BUDGET_DEFAULT = 8ull << 30;
constexpr uint32_t SERVE_QUEUE_DEFAULT = 65536;
constexpr uint32_t SERVE_DEPTH = 8; // at most this many tiles in flight per stream
constexpr uint32_t SERVE_STREAMS_DEFAULT = 3;
constexpr double SERVE_STREAM_WORK = 2.0 * SERVE_TILE_WORK; // estimated work in flight per stream

// Buffers of finished jobs are kept for later ones, by power-of-two size
// class: hipFree can wait for the whole device, which would serialize the
// streams at every job completion (and hipHostMalloc is slow and
// synchronizing).  held counts every byte the pool has allocated and not yet
// freed, in use or cached; before a new allocation would take held past the
// limit, cached buffers of other classes are freed, largest first.
template <bool PINNED>
struct BufferPool {
    std::multimap<uint64_t, void *> free_bufs;
    std::unordered_set<void *> live;
    uint64_t limit = UINT64_MAX;
    uint64_t held = 0, cached = 0, peak = 0;
    static uint64_t size_class(uint64_t bytes)
    {
        uint64_t c = 4096;
        while (c < bytes) c <<= 1;
        return c;
    }
    uint64_t in_use() const { return held - cached; }
    void *get(uint64_t bytes)
    {
        const uint64_t c = size_class(bytes);
        auto it = free_bufs.find(c);
        if (it != free_bufs.end()) {
            void *p = it->second;
            free_bufs.erase(it);
            cached -= c;
            return p;
        }
        trim(c);
        void *p = nullptr;
        if constexpr (PINNED) HIPCHK(hipHostMalloc(&p, c, hipHostMallocDefault));
        else HIPCHK(hipMalloc(&p, c));
        live.insert(p);
        held += c;
        peak = std::max(peak, held);
        return p;
    }
    void put(void *p, uint64_t bytes)
    {
        const uint64_t c = size_class(bytes);
        free_bufs.emplace(c, p);
        cached += c;
    }
    // Frees cached buffers until `need` more bytes fit under the limit (or none is left).
    void trim(uint64_t need)
    {
        while (held + need > limit && !free_bufs.empty()) {
            auto it = std::prev(free_bufs.end());
            if constexpr (PINNED) HIPCHK(hipHostFree(it->second));
            else HIPCHK(hipFree(it->second));
            live.erase(it->second);
            held -= it->first;
            cached -= it->first;
            free_bufs.erase(it);
        }
    }
    void release()
    {
        for (void *p : live) {
            if constexpr (PINNED) (void)hipHostFree(p);
            else (void)hipFree(p);
        }
        live.clear();
        free_bufs.clear();
        held = cached = 0;
    }
};
using DevicePool = BufferPool<false>;
using PinnedPool = BufferPool<true>; // shared by the scheduler and the delivering threads: used under the server mutex

// Staging layout of one job: the two accumulators, then counts, then factors.
inline uint64_t stage_counts_off() { return 2 * sizeof(unsigned long long); }
inline uint64_t stage_factors_off(uint64_t N) { return (stage_counts_off() + N * sizeof(uint32_t) + 7) & ~7ull; }
inline uint64_t stage_bytes(uint64_t N, uint32_t maxf) { return stage_factors_off(N) + N * (uint64_t)maxf * 8; }
uint64_t job_pinned_bytes(uint64_t N, uint32_t maxf) { return PinnedPool::size_class(stage_bytes(N, maxf)); }

struct ServeJob {
    uint64_t ticket = 0;
    ctd_job job{};
    uint32_t *h_counts = nullptr;
    uint64_t *h_factors = nullptr;
    uint32_t *d_counts = nullptr;
    uint64_t *d_factors = nullptr;
    unsigned long long *d_acc = nullptr; // odd primes tested, summed over the tiles
    unsigned char *stage = nullptr;      // pinned staging (stage_bytes)
    bool delivering = false;
    uint64_t bytes = 0;
    JobPlan plan;
    std::deque<Tile> tiles;  // not yet issued
    uint32_t outstanding = 0; // issued, not yet retired
    hipEvent_t ready = nullptr;
    hipEvent_t done = nullptr;
    bool copying = false;
    std::chrono::steady_clock::time_point t_submit;
    double work_left = 0; // estimated work of the tiles not yet issued
};

struct Ticket {
    int state = 0; // 0 queued, 1 running, 3 staged (results in pinned memory), 2 done, < 0 error code
    double latency_ms = 0;
    ServeJob *staged = nullptr;
};

struct InFlight {
    hipEvent_t ev;
    ServeJob *job;
    double work; // estimated work of the tile (tile_work)
};

struct ctd_hip_server_impl {
    ctd_hip_engine *E = nullptr;
    ctd_server_opts opts{};
    std::vector<BatchSlot> slots;
    std::vector<std::deque<InFlight>> inflight;
    std::vector<double> inflight_work; // per stream, sum of InFlight::work
    bool pick_least = false;           // pick_job: the next place goes to the least remaining work
    double stream_work = SERVE_STREAM_WORK; // estimated work in flight per stream (opts.depth tiles)
    hipStream_t copy_st = nullptr;
    std::vector<hipEvent_t> event_pool;
    DevicePool pool;
    PinnedPool pinned; // under mu

    std::mutex mu;
    std::condition_variable cv_done;
    // Wake-ups of the scheduler: submit, close, a freed staging buffer, and a
    // host function the device runs after each tile and each result copy.
    // This mutex is never held across a HIP call (host functions take it).
    std::mutex wmu;
    std::condition_variable wcv;
    bool wake = false;
    std::deque<ServeJob *> pending;
    std::unordered_map<uint64_t, Ticket> tickets;
    uint64_t next_ticket = 1;
    bool stop = false;
    int fatal = CTD_OK;
    std::string fatal_msg;
    ctd_server_stats stats{};

    std::deque<ServeJob *> active; // scheduler thread only
    uint64_t used_dev = 0;         // device bytes charged to admitted jobs (size classes)
    std::atomic<double> sched_cpu_ms{0.0};
    std::atomic<uint64_t> passes{0};
    std::atomic<uint64_t> dev_peak{0}; // pool.peak, readable by ctd_hip_server_stats
    std::thread th;
};

void wake_scheduler(ctd_hip_server_impl *S)
{
    std::lock_guard<std::mutex> lk(S->wmu);
    S->wake = true;
    S->wcv.notify_one();
}

// Host function enqueued behind device work: the work before it has completed.
void wake_from_device(void *user) { wake_scheduler(static_cast<ctd_hip_server_impl *>(user)); }

// Pool bytes one job holds: its device factors buffer and accumulators, and
// its pinned staging buffer (each rounded to its size class).
uint64_t job_device_bytes(uint64_t N, uint32_t maxf)
{
    return DevicePool::size_class(N * sizeof(uint32_t)) + DevicePool::size_class(N * (uint64_t)maxf * 8) +
           DevicePool::size_class(2 * sizeof(unsigned long long));
}
uint64_t job_pinned_bytes(uint64_t N, uint32_t maxf);

hipEvent_t pool_event(ctd_hip_server_impl *S)
{
    if (!S->event_pool.empty()) {
        hipEvent_t e = S->event_pool.back();
        S->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
}

void finish_ticket(ctd_hip_server_impl *S, ServeJob *J, int state)
{
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - J->t_submit).count();
    std::lock_guard<std::mutex> lk(S->mu);
    Ticket &t = S->tickets[J->ticket];
    t.state = state;
    t.latency_ms = ms;
    if (state == 2) S->stats.completed++;
    else S->stats.failed++;
    S->cv_done.notify_all();
}

void deliver(ctd_hip_server_impl *S, Ticket &t, std::unique_lock<std::mutex> &lk);

// The job's results are in its pinned staging buffer (scheduler thread):
// count it and hand it to the delivering side.
void stage_ticket(ctd_hip_server_impl *S, ServeJob *J)
{
    const unsigned long long *acc = (const unsigned long long *)J->stage;
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - J->t_submit).count();
    std::unique_lock<std::mutex> lk(S->mu);
    Ticket &t = S->tickets[J->ticket];
    t.state = 3;
    t.staged = J;
    t.latency_ms = ms;
    S->stats.completed++;
    S->stats.primes += acc[0];
    S->stats.updates += acc[0] * (J->job.n1 > J->job.n0 ? J->job.n1 - J->job.n0 : 0);
    S->cv_done.notify_all();
    // results nobody has collected keep their staging buffers; past the
    // memory budget the scheduler copies them into the caller's buffers itself
    if (S->pinned.in_use() > S->opts.memory_budget) deliver(S, t, lk);
}

// Copies a staged job's results into the caller's buffers (the calling
// thread does the copy, outside the lock) and completes its ticket.  lk holds
// S->mu on entry and on return.
void deliver(ctd_hip_server_impl *S, Ticket &t, std::unique_lock<std::mutex> &lk)
{
    ServeJob *J = t.staged;
    if (!J || J->delivering) return;
    J->delivering = true;
    lk.unlock();
    const uint64_t N = J->job.n1 > J->job.n0 ? J->job.n1 - J->job.n0 : 0;
    std::memcpy(J->h_counts, J->stage + stage_counts_off(), N * sizeof(uint32_t));
    std::memcpy(J->h_factors, J->stage + stage_factors_off(N), N * (uint64_t)J->job.max_factors * 8);
    lk.lock();
    S->pinned.put(J->stage, stage_bytes(N, J->job.max_factors));
    t.staged = nullptr;
    t.state = 2;
    delete J;
    S->cv_done.notify_all();
    wake_scheduler(S); // admission may be waiting for staging memory
}

// Admission of one job (scheduler thread).  Returns false when the job has to
// wait for memory.
// Estimated work of a tile: primes in [blo, bhi) x the job's indices x the
// family's VALU instructions per update.
double tile_work(const JobPlan &P, uint64_t blo, uint64_t bhi)
{
    return (double)(bhi - blo) / std::log((double)std::max<uint64_t>(bhi, 16)) * (double)P.N * P.ops;
}

bool admit(ctd_hip_server_impl *S, ServeJob *J)
{
    ctd_hip_engine *E = S->E;
    const ctd_job &q = J->job;
    const uint64_t N = q.n1 > q.n0 ? q.n1 - q.n0 : 0;
    if (!N || q.f1 <= q.f0) { // trial_factoring_impl.hpp:380-387: nothing to do
        finish_ticket(S, J, 2);
        delete J;
        return true;
    }
    // the job's device and pinned pool bytes must fit in the budget next to
    // those of the admitted jobs (and of results not yet collected); a job
    // that arrives at an idle server is always admitted
    J->bytes = job_device_bytes(N, q.max_factors);
    const uint64_t pin = job_pinned_bytes(N, q.max_factors);
    if (!S->active.empty()) {
        if (S->used_dev + J->bytes > S->opts.memory_budget) return false;
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->pinned.in_use() + pin > S->opts.memory_budget) return false;
    }
    const uint64_t f0 = std::max<uint64_t>(q.f0, 3);
    if (f0 < q.f1 && isqrt_u64(q.f1 - 1) > E->base_lim && E->d_sp) {
        // the sieving base grows: no tile may be reading the old one
        for (auto &sl : S->slots) HIPCHK(hipStreamSynchronize(sl.st));
    }
    J->d_counts = (uint32_t *)S->pool.get(N * sizeof(uint32_t));
    J->d_factors = (uint64_t *)S->pool.get(N * (uint64_t)q.max_factors * sizeof(uint64_t));
    J->d_acc = (unsigned long long *)S->pool.get(2 * sizeof(unsigned long long));
    S->dev_peak.store(S->pool.peak, std::memory_order_relaxed);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        J->stage = (unsigned char *)S->pinned.get(stage_bytes(N, q.max_factors));
    }
    S->used_dev += J->bytes;
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, S->copy_st, J->d_acc, 2u);
    if (S->opts.flags & CTD_SERVE_FRESH_BUFFERS) {
        // the job starts from an empty factors buffer: clear it on the device
        const uint64_t wf = 2 * N * (uint64_t)q.max_factors;
        hipLaunchKernelGGL(zero_words_kernel, dim3((uint32_t)std::min<uint64_t>(div_up(N, 256), 4096)), dim3(256), 0,
                           S->copy_st, (uint32_t *)J->d_counts, N);
        hipLaunchKernelGGL(zero_words_kernel, dim3((uint32_t)std::min<uint64_t>(div_up(wf, 1024), 4096)), dim3(256),
                           0, S->copy_st, (uint32_t *)J->d_factors, wf);
    } else {
        // the caller's buffers go through the pinned staging buffer, so the
        // copies to the device are asynchronous
        std::memcpy(J->stage + stage_counts_off(), J->h_counts, N * sizeof(uint32_t));
        std::memcpy(J->stage + stage_factors_off(N), J->h_factors, N * (uint64_t)q.max_factors * 8);
        HIPCHK(hipMemcpyAsync(J->d_counts, J->stage + stage_counts_off(), N * sizeof(uint32_t),
                              hipMemcpyHostToDevice, S->copy_st));
        HIPCHK(hipMemcpyAsync(J->d_factors, J->stage + stage_factors_off(N), N * (uint64_t)q.max_factors * 8,
                              hipMemcpyHostToDevice, S->copy_st));
    }
    std::vector<JobPlan> plans;
    uint32_t tuned = 0;
    int rc = plan_jobs(E, &q, 1, &J->d_counts, &J->d_factors, false, plans, &tuned);
    if (rc != CTD_OK) throw HipError{hipErrorInvalidValue, "job plan"};
    J->plan = plans[0];
    if (J->plan.f0 < J->plan.f1) {
        const bool fast = Fast<0>::fits(J->plan.f1 - 1);
        const bool narrow = q.seq_id == FAM_SMARANDACHE && sm_narrow_ok(q.param, q.n1, J->plan.f1 - 1);
        auto it = E->shapes.find(shape_key(q.seq_id, fast, narrow, N));
        if (it != E->shapes.end()) J->plan.shape = it->second;
    }
    if (J->plan.has2) {
        uint32_t g = (uint32_t)std::min<uint64_t>(div_up(N, 256), 4096);
        hipLaunchKernelGGL(mod2_kernel, dim3(g), dim3(256), 0, S->copy_st, J->plan.J);
        HIPCHK(hipGetLastError());
    }
    J->ready = pool_event(S);
    HIPCHK(hipEventRecord(J->ready, S->copy_st));
    // tiles of about SERVE_TILE_WORK, or 1/SERVE_JOB_TILES of the window when
    // that is longer (at most tile_length integers, at least one segment): a
    // job with many indices or a costly family is cut finer, so its tiles
    // spread over every stream instead of holding one
    const double primes_per_int = 1.0 / std::log((double)std::max<uint64_t>(q.f1, 16));
    const double window = J->plan.f1 > J->plan.f0 ? (double)(J->plan.f1 - J->plan.f0) : 0.0;
    const double want = std::max(SERVE_TILE_WORK / ((double)N * J->plan.ops * primes_per_int), // integers
                                 window / SERVE_JOB_TILES);
    const uint64_t segs = std::max<uint64_t>(1, (uint64_t)(want / (double)SEG_SPAN));
    const uint64_t tl = std::min<uint64_t>(S->opts.tile_length, segs * SEG_SPAN);
    if (J->plan.f0 < J->plan.f1) {
        const uint64_t lo0 = J->plan.f0 & ~1ull;
        for (uint64_t blo = lo0; blo < J->plan.f1; blo += tl) {
            const uint64_t bhi = block_end(blo, tl, J->plan.f1);
            J->tiles.push_back(Tile{0, blo, bhi, std::max(blo, J->plan.f0), 0.0, 0u, false});
            J->work_left += tile_work(J->plan, blo, bhi);
            if (bhi == J->plan.f1) break;
        }
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->tickets[J->ticket].state = 1;
        S->stats.admitted++;
        S->stats.active_max = std::max<uint32_t>(S->stats.active_max, (uint32_t)S->active.size() + 1);
    }
    S->active.push_back(J);
    return true;
}

// The admitted job that gets the next free place in a stream.  Default:
// places alternate between the oldest job that still has tiles and the job
// with the least remaining work, so a short job that arrives behind long
// ones starts at the next free place, while the long ones keep every other
// place (no job waits for ever, and the streams keep a mix of long tiles
// that fill the device and short ones).  CTD_SERVE_FIFO: always the oldest.
ServeJob *pick_job(ctd_hip_server_impl *S)
{
    ServeJob *oldest = nullptr, *least = nullptr;
    for (ServeJob *a : S->active) {
        if (a->tiles.empty()) continue;
        if (!oldest) oldest = a;
        if (!least || a->work_left < least->work_left) least = a;
    }
    if (S->opts.flags & CTD_SERVE_FIFO) return oldest;
    S->pick_least = !S->pick_least;
    return S->pick_least ? least : oldest;
}

// One pass of the scheduler; returns true when it made progress.
bool schedule_once(ctd_hip_server_impl *S)
{
    ctd_hip_engine *E = S->E;
    bool progress = false;
    // admit, while the admitted jobs hold fewer unissued tiles than twice what
    // the streams keep in flight: enough to keep every stream fed, and jobs
    // that could not start yet hold no device memory
    uint64_t ready_tiles = 0;
    for (ServeJob *a : S->active) ready_tiles += a->tiles.size();
    const uint64_t want = 2ull * S->slots.size() * SERVE_DEPTH;
    for (;;) {
        if (ready_tiles >= want) break;
        ServeJob *J = nullptr;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (S->pending.empty()) break;
            J = S->pending.front();
        }
        const size_t before = S->active.size();
        if (!admit(S, J)) break; // an empty job is finished (and freed) by admit
        if (S->active.size() > before) ready_tiles += S->active.back()->tiles.size();
        std::lock_guard<std::mutex> lk(S->mu);
        S->pending.pop_front();
        progress = true;
    }
    // retire
    for (size_t s = 0; s < S->slots.size(); ++s) {
        auto &q = S->inflight[s];
        while (!q.empty()) {
            hipError_t e = hipEventQuery(q.front().ev);
            if (e == hipErrorNotReady) break;
            if (e != hipSuccess) throw HipError{e, "tile event"};
            S->event_pool.push_back(q.front().ev);
            q.front().job->outstanding--;
            S->inflight_work[s] -= q.front().work;
            q.pop_front();
            progress = true;
        }
        if (q.empty()) S->inflight_work[s] = 0.0; // no rounding drift
    }
    // issue
    for (size_t s = 0; s < S->slots.size(); ++s) {
        while (S->inflight[s].empty() ||
               (S->inflight[s].size() < SERVE_DEPTH && S->inflight_work[s] < S->stream_work)) {
            ServeJob *J = pick_job(S);
            if (!J) break;
            Tile T = J->tiles.front();
            J->tiles.pop_front();
            BatchSlot &sl = S->slots[s];
            HIPCHK(hipStreamWaitEvent(sl.st, J->ready, 0));
            issue_tile(E, J->plan, sl, T.blo, T.bhi, T.start, J->d_acc, true, SERVE_LANES_PER_CU);
            hipEvent_t ev = pool_event(S);
            HIPCHK(hipEventRecord(ev, sl.st));
            HIPCHK(hipLaunchHostFunc(sl.st, wake_from_device, S));
            const double work = tile_work(J->plan, T.blo, T.bhi);
            J->work_left = J->tiles.empty() ? 0.0 : J->work_left - work;
            S->inflight[s].push_back(InFlight{ev, J, work});
            S->inflight_work[s] += work;
            J->outstanding++;
            {
                std::lock_guard<std::mutex> lk(S->mu);
                S->stats.tiles++;
            }
            progress = true;
        }
    }
    // complete
    for (auto it = S->active.begin(); it != S->active.end();) {
        ServeJob *J = *it;
        if (!J->tiles.empty() || J->outstanding) {
            ++it;
            continue;
        }
        const uint64_t N = J->job.n1 - J->job.n0;
        if (!J->copying) {
            // every tile event has completed, so the device buffers are final
            HIPCHK(hipMemcpyAsync(J->stage + stage_counts_off(), J->d_counts, N * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, S->copy_st));
            HIPCHK(hipMemcpyAsync(J->stage + stage_factors_off(N), J->d_factors,
                                  N * (uint64_t)J->job.max_factors * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  S->copy_st));
            HIPCHK(hipMemcpyAsync(J->stage, J->d_acc, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                  S->copy_st));
            J->done = pool_event(S);
            HIPCHK(hipEventRecord(J->done, S->copy_st));
            HIPCHK(hipLaunchHostFunc(S->copy_st, wake_from_device, S));
            J->copying = true;
            progress = true;
        }
        hipError_t e = hipEventQuery(J->done);
        if (e == hipErrorNotReady) {
            ++it;
            continue;
        }
        if (e != hipSuccess) throw HipError{e, "copy event"};
        S->pool.put(J->d_counts, N * sizeof(uint32_t));
        S->pool.put(J->d_factors, N * (uint64_t)J->job.max_factors * sizeof(uint64_t));
        S->pool.put(J->d_acc, 2 * sizeof(unsigned long long));
        S->used_dev -= J->bytes;
        S->event_pool.push_back(J->ready);
        S->event_pool.push_back(J->done);
        it = S->active.erase(it);
        stage_ticket(S, J); // J now belongs to the delivering side
        progress = true;
    }
    (void)E;
    return progress;
}

void fail_all(ctd_hip_server_impl *S, int code, const std::string &msg)
{
    std::lock_guard<std::mutex> lk(S->mu);
    S->fatal = code;
    S->fatal_msg = msg;
    for (ServeJob *J : S->active) {
        S->tickets[J->ticket].state = code;
        S->stats.failed++;
        if (J->stage) S->pinned.put(J->stage, stage_bytes(J->job.n1 - J->job.n0, J->job.max_factors));
        delete J;
    }
    S->active.clear();
    for (ServeJob *J : S->pending) {
        S->tickets[J->ticket].state = code;
        S->stats.failed++;
        delete J;
    }
    S->pending.clear();
    S->cv_done.notify_all();
}

double thread_cpu_ms()
{
    timespec ts{};
    clock_gettime(CLOCK_THREAD_CPUTIME_ID, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

// The scheduler thread.  After a pass that changed nothing it sleeps until
// woken: by a submit, by close, by a delivered result (staging memory freed),
// or by the device itself through the host function behind every tile and
// every result copy.  It does not poll: while jobs are on the device a 2 ms
// time-out only bounds the wait should a wake-up ever be missed.
void serve_loop(ctd_hip_server_impl *S)
{
    try {
        HIPCHK(hipSetDevice(S->E->device));
        for (;;) {
            const bool progress = schedule_once(S);
            S->passes.fetch_add(1, std::memory_order_relaxed);
            S->sched_cpu_ms.store(thread_cpu_ms(), std::memory_order_relaxed);
            {
                std::lock_guard<std::mutex> lk(S->mu);
                if (S->stop && S->pending.empty() && S->active.empty()) break;
            }
            if (progress) continue;
            std::unique_lock<std::mutex> wl(S->wmu);
            if (S->active.empty()) {
                S->wcv.wait(wl, [&] { return S->wake; });
            } else {
                // work is on the device: its host functions wake the thread;
                // the time-out only bounds the wait should one ever be missed
                S->wcv.wait_for(wl, std::chrono::milliseconds(2), [&] { return S->wake; });
            }
            S->wake = false;
        }
    } catch (const HipError &h) {
        fail_all(S, h.code == hipErrorOutOfMemory ? CTD_E_NOMEM : CTD_E_DEVICE,
                 std::string(h.what) + ": " + hipGetErrorString(h.code));
    }
}

void release_server(ctd_hip_server_impl *S)
{
    for (auto &sl : S->slots) {
        if (sl.st) (void)hipStreamSynchronize(sl.st);
        (void)hipFree(sl.primes);
        (void)hipFree(sl.segcnt);
        (void)hipFree(sl.sc);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.st) (void)hipStreamDestroy(sl.st);
    }
    for (auto &q : S->inflight)
        for (auto &f : q) (void)hipEventDestroy(f.ev);
    if (S->copy_st) {
        (void)hipStreamSynchronize(S->copy_st);
        (void)hipStreamDestroy(S->copy_st);
    }
    for (auto e : S->event_pool) (void)hipEventDestroy(e);
    S->pool.release();
    for (auto &kv : S->tickets) { // staged results nobody collected
        if (kv.second.staged) delete kv.second.staged;
        kv.second.staged = nullptr;
    }
    S->pinned.release();
}
