// ctd_batch.hpp -- batched form of the engine: many jobs in one hipGraph.
//
// Included by ctd_engine.hip inside its internal namespace (it uses the
// engine's helpers).  A job's prime window [f0, f1) is cut into tiles of
// tile_length integers.  Per tile, on one of S streams that each own a
// resident prime slot:
//   zero the slot's counters -> sieve the tile into the slot -> keep its count
//   -> (first tile of a job whose window holds 2: mod-2 kernel)
//   -> the family's residue kernel over (slot primes) x (job indices),
//      reading the prime count from the slot (no host round trip).
// Tiles are ordered by estimated cost and dealt to the least-loaded stream,
// so the sieve of one tile overlaps the residue kernels of the others and the
// long tiles start first.  The whole schedule is captured once into a
// hipGraph and launched with one hipGraphLaunch per 1024 tiles.
//
// Autotuner: for each (family, kernel kind, index-count class) the first
// batch that needs it times a probe tile under every launch shape of a small
// grid -- workgroup size (64, 256 lanes), primes per lane (1, 2, 4: each lane
// takes that many work items in turn) and index-chunk split (1, 2, 4) --
// with one warm-up and three timed runs each, and keeps the shape of the
// lowest median.  Shapes change the schedule only: every shape computes the
// same residues and records the same hits.

constexpr uint32_t BATCH_GRAPH_TILES = 1024;
constexpr uint64_t BATCH_TILE_DEFAULT = 1ull << 28;
constexpr uint64_t TUNE_WINDOW = 1ull << 25;
constexpr uint64_t TUNE_MAX_N = 16384;
constexpr int TUNE_REPS = 3;

// Montgomery-Vaughan bound pi(x + y) - pi(x) <= 2y / log y, plus slack.
uint64_t prime_cap(uint64_t len)
{
    double ly = std::log((double)std::max<uint64_t>(len, 16));
    return (uint64_t)(2.0 * (double)len / ly) + 4096;
}

double est_primes(uint64_t lo, uint64_t hi)
{
    if (hi <= lo) return 0.0;
    double x = (double)std::max<uint64_t>(hi, 16);
    return (double)(hi - lo) / std::log(x);
}

void release_batch(ctd_hip_engine *E)
{
    for (auto &s : E->slots) {
        if (s.st) (void)hipStreamSynchronize(s.st);
        (void)hipFree(s.primes);
        (void)hipFree(s.segcnt);
        (void)hipFree(s.sc);
        if (s.ev) (void)hipEventDestroy(s.ev);
        if (s.st) (void)hipStreamDestroy(s.st);
    }
    E->slots.clear();
    if (E->cap_stream) (void)hipStreamDestroy(E->cap_stream);
    E->cap_stream = nullptr;
    if (E->ev_fork) (void)hipEventDestroy(E->ev_fork);
    E->ev_fork = nullptr;
    (void)hipFree(E->d_err);
    E->d_err = nullptr;
    (void)hipFree(E->d_tile);
    E->d_tile = nullptr;
    E->tile_cap = 0;
    (void)hipFree(E->d_scr_counts);
    (void)hipFree(E->d_scr_factors);
    E->d_scr_counts = nullptr;
    E->d_scr_factors = nullptr;
    E->scr_cap = 0;
}

// Slots able to hold the primes of a tile of tile_len integers.
void ensure_slots(ctd_hip_engine *E, uint32_t nslots, uint64_t tile_len)
{
    const uint64_t cap = prime_cap(tile_len);
    const uint64_t nseg = div_up(tile_len, SEG_SPAN) + 1;
    if (!E->cap_stream) HIPCHK(hipStreamCreateWithFlags(&E->cap_stream, hipStreamNonBlocking));
    if (!E->ev_fork) HIPCHK(hipEventCreateWithFlags(&E->ev_fork, hipEventDisableTiming));
    if (!E->d_err) HIPCHK(hipMalloc((void **)&E->d_err, sizeof(unsigned long long)));
    while (E->slots.size() < nslots) {
        BatchSlot s;
        HIPCHK(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
        HIPCHK(hipMalloc((void **)&s.sc, 2 * sizeof(unsigned long long)));
        E->slots.push_back(s);
    }
    for (auto &s : E->slots) {
        ensure_buf(s.primes, s.cap, cap);
        ensure_buf(s.segcnt, s.segcap, nseg);
    }
}

struct JobPlan {
    JobArgs J{};       // base arguments (chunking and flags are set per tile)
    uint64_t N = 0;
    uint64_t f0 = 0;   // first odd candidate (the mod-2 path covers 2)
    uint64_t f1 = 0;
    bool has2 = false;
    LaunchShape shape;
    double ops = 7.0;  // relative cost per update, for ordering tiles
    bool fp64 = true;  // the FP64 forms are allowed
    uint32_t stepped = 0; // JOB_STEPPED when the fixed-n families step every index
};

struct Tile {
    uint32_t job;
    uint64_t blo, bhi, start;
    double cost;
    uint32_t stream;
    bool first;
};

// Lanes a tile's residue launch should offer (the index range is split into
// chunks until it does): the batched graph's concurrent tiles keep the device
// full with 512 per CU, and fewer chunks mean less per-chunk set-up (2000
// jobs of 2^20-integer windows: 6.90e12 updates/s against 6.47e12 with 2048;
// tiles of 2^24 integers and more hold enough primes either way); the serving
// form, whose streams run dry between scheduler passes, measured best at 4096.
constexpr uint32_t BATCH_LANES_PER_CU = 512;
constexpr uint32_t SERVE_LANES_PER_CU = 4096;

// VALU instructions per update of the family's loop (FP64 forms below 2^48,
// integer forms above), for ordering tiles and balancing the streams.
double family_cost(uint32_t fam, bool fast, bool fp64, uint64_t N, uint32_t stepped = 0)
{
    if (!fast) return 50.0;
    // one power of about log2(n) multiplies per prime, whatever the index count
    if (closed_form(fam, stepped)) return (fp64 ? 300.0 : 2000.0) / (double)std::max<uint64_t>(N, 1);
    switch (fam) {
    case FAM_SMARANDACHE: return fp64 ? 7.0 : 12.25;
    case FAM_TRIBONACCI: return fp64 ? 5.5 : 11.25;
    case FAM_FIBONACCI:
    case FAM_LUCAS: return fp64 ? (N >= THETA_MIN_STEPS ? 2.1 : 5.0) : 6.0;
    // doubling and fixed-n strides: 1.1-1.2 in the loop, but their short work
    // items pay 64-target set-ups; 1.5 sized the serving tiles best
    default: return fp64 ? 1.5 : 7.0;
    }
}

// Tile arguments: chunking from the estimated prime count and the shape's
// split, the Smarandache narrow flag from the tile's largest candidate.
JobArgs tile_args(const ctd_hip_engine *E, const JobPlan &P, uint64_t est_np, uint64_t pmax,
                  uint32_t lanes_per_cu = BATCH_LANES_PER_CU)
{
    JobArgs J = P.J;
    uint64_t nch = 1;
    const uint64_t target = (uint64_t)E->cu * lanes_per_cu;
    if (est_np < target && P.N > 2048) nch = std::min<uint64_t>(div_up(target, std::max<uint64_t>(est_np, 1)),
                                                                 div_up(P.N, 1024));
    nch = std::min<uint64_t>(std::max<uint64_t>(nch * P.shape.split, div_up(P.N, CHUNK_MAX)), P.N);
    if (closed_form(J.fam, P.stepped)) nch = 1; // one solve per prime covers every index
    const uint64_t chunk = div_up(P.N, nch);
    J.chunk = chunk;
    J.nchunks = (uint32_t)div_up(P.N, chunk);
    J.flags = (J.fam == FAM_SMARANDACHE && sm_narrow_ok(J.param, J.n1, pmax)) ? JOB_SM_NARROW : 0u;
    if (pmax < FP_PMAX && P.fp64) J.flags |= JOB_FP64;
    J.flags |= P.stepped;
    return J;
}

dim3 tile_grid(const ctd_hip_engine *E, const LaunchShape &sh, uint64_t est_np, uint32_t nchunks)
{
    (void)E;
    const uint64_t items = (((uint64_t)(1.1 * (double)est_np) + 64 + 63) & ~63ull) * nchunks;
    const uint64_t g = div_up(items, (uint64_t)sh.block * sh.ppl);
    return dim3((uint32_t)std::max<uint64_t>(std::min<uint64_t>(g, 1ull << 30), 1));
}

uint64_t shape_key(uint32_t fam, bool fast, bool narrow, uint64_t N, uint32_t stepped = 0)
{
    uint32_t ncls = N < 1024 ? 0 : N < 16384 ? 1 : 2;
    return ((uint64_t)fam << 8) | ((uint64_t)(stepped != 0) << 5) | ((uint64_t)fast << 4) | ((uint64_t)narrow << 3) |
           ncls;
}

void ensure_scratch(ctd_hip_engine *E, uint64_t n)
{
    if (E->scr_cap >= n && E->d_scr_counts) return;
    (void)hipFree(E->d_scr_counts);
    (void)hipFree(E->d_scr_factors);
    E->d_scr_counts = nullptr;
    E->d_scr_factors = nullptr;
    HIPCHK(hipMalloc((void **)&E->d_scr_counts, n * sizeof(uint32_t)));
    HIPCHK(hipMalloc((void **)&E->d_scr_factors, n * sizeof(uint64_t)));
    E->scr_cap = n;
}

// The operations of one tile on its slot's stream: zero the slot counters,
// sieve, keep the count (tile_out[0..1]), propagate.
void issue_tile(ctd_hip_engine *E, const JobPlan &P, BatchSlot &S, uint64_t blo, uint64_t bhi, uint64_t start,
                unsigned long long *tile_out, bool accumulate = false, uint32_t lanes_per_cu = BATCH_LANES_PER_CU)
{
    hipStream_t st = S.st;
    SieveArgs A{};
    A.lo0 = blo;
    A.start = start;
    A.end = bhi;
    A.sp = E->d_sp;
    A.sp_mu = E->d_mu;
    A.nsp = E->nsp;
    A.nsmall = E->nsmall;
    A.nwheel = E->nwheel;
    A.mode = SIEVE_DENSE;
    A.seg_counts = S.segcnt;
    A.out64 = S.primes;
    A.total = S.sc;
    A.maxp = S.sc + 1;
    A.cap = S.cap;
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, S.sc, 2u);
    hipLaunchKernelGGL(sieve_kernel, dim3((uint32_t)div_up(bhi - blo, SEG_SPAN)), dim3(SIEVE_THREADS), 0, st, A);
    hipLaunchKernelGGL(tile_done_kernel, dim3(1), dim3(64), 0, st, S.sc, S.cap, E->d_err, tile_out,
                       (uint32_t)accumulate);
    const uint64_t est = (uint64_t)est_primes(start, bhi);
    JobArgs J = tile_args(E, P, est, bhi - 1, lanes_per_cu);
    const bool fast = Fast<0>::fits(bhi - 1);
    launch_tf_any(J, S.primes, S.cap, S.sc, fast, tile_grid(E, P.shape, est, J.nchunks), dim3(P.shape.block), st);
    HIPCHK(hipGetLastError());
}

// Slot buffers able to hold a tile of tile_len integers.
void ensure_slot_buffers(BatchSlot &s, uint64_t tile_len)
{
    ensure_buf(s.primes, s.cap, prime_cap(tile_len));
    ensure_buf(s.segcnt, s.segcap, div_up(tile_len, SEG_SPAN) + 1);
}

// Times the candidate shapes on a probe tile of the job and returns the best.
LaunchShape tune_shape(ctd_hip_engine *E, const JobPlan &P, bool fast)
{
    const uint64_t lo = P.f0, hi = P.f0 < P.f1 ? block_end(P.f0, TUNE_WINDOW, P.f1) : P.f0;
    LaunchShape best;
    if (lo >= hi) return best;
    const uint64_t blo = lo & ~1ull;
    ensure_slots(E, 1, std::max<uint64_t>(hi - blo, SEG_SPAN));
    BatchSlot &S = E->slots[0];
    hipStream_t st = S.st;
    SieveArgs A{};
    A.lo0 = blo;
    A.start = lo;
    A.end = hi;
    A.sp = E->d_sp;
    A.sp_mu = E->d_mu;
    A.nsp = E->nsp;
    A.nsmall = E->nsmall;
    A.nwheel = E->nwheel;
    A.mode = SIEVE_DENSE;
    A.seg_counts = S.segcnt;
    A.out64 = S.primes;
    A.total = S.sc;
    A.maxp = S.sc + 1;
    A.cap = S.cap;
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, S.sc, 2u);
    hipLaunchKernelGGL(sieve_kernel, dim3((uint32_t)div_up(hi - blo, SEG_SPAN)), dim3(SIEVE_THREADS), 0, st, A);
    HIPCHK(hipGetLastError());
    unsigned long long sc[2];
    HIPCHK(hipMemcpyAsync(sc, S.sc, sizeof sc, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const uint64_t np = std::min<uint64_t>(sc[0], S.cap);
    if (!np) return best;
    // probe job: at most TUNE_MAX_N indices, hits into scratch buffers
    JobPlan Q = P;
    Q.N = std::min<uint64_t>(P.N, TUNE_MAX_N);
    Q.J.n1 = Q.J.n0 + Q.N;
    Q.J.maxf = 1;
    ensure_scratch(E, Q.N);
    Q.J.counts = E->d_scr_counts;
    Q.J.factors = E->d_scr_factors;
    const uint32_t blocks[] = {64u, 256u};
    const uint32_t ppls[] = {1u, 2u, 4u};
    const uint32_t splits[] = {1u, 2u, 4u};
    float best_ms = 1e30f;
    for (uint32_t b : blocks)
        for (uint32_t pl : ppls)
            for (uint32_t sp : splits) {
                if (sp > 1 && (Q.N < 1024 * sp || closed_form(Q.J.fam, Q.stepped))) continue;
                Q.shape = LaunchShape{b, pl, sp};
                JobArgs J = tile_args(E, Q, np, hi - 1);
                dim3 grid = tile_grid(E, Q.shape, np, J.nchunks);
                float ms[TUNE_REPS + 1];
                for (int rep = 0; rep <= TUNE_REPS; ++rep) { // the first run warms the code object
                    HIPCHK(hipEventRecord(E->ev0, st));
                    launch_tf_any(J, S.primes, np, nullptr, fast, grid, dim3(b), st);
                    HIPCHK(hipGetLastError());
                    HIPCHK(hipEventRecord(E->ev1, st));
                    HIPCHK(hipStreamSynchronize(st));
                    HIPCHK(hipEventElapsedTime(&ms[rep], E->ev0, E->ev1));
                }
                std::sort(ms + 1, ms + TUNE_REPS + 1);
                const float med = ms[1 + TUNE_REPS / 2];
                if (med < best_ms) {
                    best_ms = med;
                    best = Q.shape;
                }
            }
    return best;
}

int plan_jobs(ctd_hip_engine *E, const ctd_job *jobs, uint32_t njobs, uint32_t *const *counts,
              uint64_t *const *factors, bool autotune, std::vector<JobPlan> &plans, uint32_t *tuned,
              uint32_t batch_flags = 0)
{
    plans.resize(njobs);
    uint64_t fmax = 0;
    for (uint32_t j = 0; j < njobs; ++j) {
        int code;
        if (!valid_job(E, &jobs[j], &code)) return fail(E, code, "job " + std::to_string(j) + ": " + E->err);
        if (!counts || !factors || !counts[j] || !factors[j])
            return fail(E, CTD_E_INVALID, "job " + std::to_string(j) + ": counts/factors are NULL");
        const ctd_job &q = jobs[j];
        JobPlan &P = plans[j];
        P.N = q.n1 > q.n0 ? q.n1 - q.n0 : 0;
        P.J.fam = q.seq_id;
        P.J.maxf = q.max_factors;
        P.J.param = q.param;
        P.J.n0 = q.n0;
        P.J.n1 = q.n1;
        P.J.chunk = P.N;
        P.J.nchunks = 1;
        P.J.counts = counts[j];
        P.J.factors = factors[j];
        P.has2 = P.N && q.f0 <= 2 && q.f1 > 2;
        P.f0 = std::max<uint64_t>(q.f0, 3);
        P.f1 = P.N ? q.f1 : 0;
        P.fp64 = !(batch_flags & CTD_BATCH_NO_FP64);
        P.stepped = (batch_flags & CTD_BATCH_STEPPED) ? JOB_STEPPED : 0u;
        if (P.f0 < P.f1) fmax = std::max(fmax, P.f1 - 1);
    }
    if (fmax) build_base(E, isqrt_u64(fmax));
    for (auto &P : plans) {
        if (P.f0 >= P.f1) continue;
        const bool fast = Fast<0>::fits(P.f1 - 1);
        P.ops = family_cost(P.J.fam, fast, P.fp64 && P.f1 - 1 < FP_PMAX, P.N, P.stepped);
        if (!autotune) continue;
        const bool narrow = P.J.fam == FAM_SMARANDACHE && sm_narrow_ok(P.J.param, P.J.n1, P.f1 - 1);
        const uint64_t key = shape_key(P.J.fam, fast, narrow, P.N, P.stepped);
        auto it = E->shapes.find(key);
        if (it == E->shapes.end()) {
            it = E->shapes.emplace(key, tune_shape(E, P, fast)).first;
            ++*tuned;
        }
        P.shape = it->second;
    }
    return CTD_OK;
}

// Issues the operations of tiles [a, b) on the slot streams, forked from and
// joined back to origin (a capturing stream, or the caller's stream).
void issue_tiles(ctd_hip_engine *E, const std::vector<JobPlan> &plans, const std::vector<Tile> &tiles, size_t a,
                 size_t b, uint32_t nst, hipStream_t origin)
{
    HIPCHK(hipEventRecord(E->ev_fork, origin));
    for (uint32_t s = 0; s < nst; ++s) HIPCHK(hipStreamWaitEvent(E->slots[s].st, E->ev_fork, 0));
    for (size_t t = a; t < b; ++t) {
        const Tile &T = tiles[t];
        const JobPlan &P = plans[T.job];
        BatchSlot &S = E->slots[T.stream];
        if (T.first && P.has2) {
            uint32_t g = (uint32_t)std::min<uint64_t>(div_up(P.N, 256), 4096);
            hipLaunchKernelGGL(mod2_kernel, dim3(g), dim3(256), 0, S.st, P.J);
        }
        issue_tile(E, P, S, T.blo, T.bhi, T.start, E->d_tile + 2 * t);
    }
    for (uint32_t s = 0; s < nst; ++s) {
        HIPCHK(hipEventRecord(E->slots[s].ev, E->slots[s].st));
        HIPCHK(hipStreamWaitEvent(origin, E->slots[s].ev, 0));
    }
}

// Tiles whose prime-2 job has no odd candidate still need the mod-2 kernel.
void mod2_only(const std::vector<JobPlan> &plans, hipStream_t st)
{
    for (const auto &P : plans) {
        if (!P.has2 || P.f0 < P.f1) continue;
        uint32_t g = (uint32_t)std::min<uint64_t>(div_up(P.N, 256), 4096);
        hipLaunchKernelGGL(mod2_kernel, dim3(g), dim3(256), 0, st, P.J);
    }
}

int run_batch(ctd_hip_engine *E, const ctd_job *jobs, uint32_t njobs, uint32_t *const *counts,
              uint64_t *const *factors, const ctd_batch_opts *opts, ctd_batch_stats *stats, hipStream_t st)
{
    using clk = std::chrono::steady_clock;
    ctd_batch_stats bs{};
    const uint32_t flags = opts ? opts->flags : 0u;
    uint64_t tile_len = (opts && opts->tile_length) ? opts->tile_length : BATCH_TILE_DEFAULT;
    tile_len = whole_spans(tile_len, SEG_SPAN, BLOCK_SEGS_MAX);
    uint32_t nst = (opts && opts->streams) ? opts->streams : 4u;
    nst = std::max<uint32_t>(1, std::min<uint32_t>(nst, 16));

    // device buffers (the caller's, or copies of host arrays)
    std::vector<uint32_t *> dc(njobs, nullptr);
    std::vector<uint64_t *> df(njobs, nullptr);
    const bool host = flags & CTD_BATCH_HOST_BUFFERS;
    struct Free {
        std::vector<uint32_t *> &c;
        std::vector<uint64_t *> &f;
        bool on;
        ~Free()
        {
            if (!on) return;
            for (auto p : c) (void)hipFree(p);
            for (auto p : f) (void)hipFree(p);
        }
    } guard{dc, df, host};
    for (uint32_t j = 0; j < njobs; ++j) {
        if (!counts || !factors || !counts[j] || !factors[j])
            return fail(E, CTD_E_INVALID, "job " + std::to_string(j) + ": counts/factors are NULL");
        if (!host) {
            dc[j] = counts[j];
            df[j] = factors[j];
            continue;
        }
        const uint64_t N = jobs[j].n1 > jobs[j].n0 ? jobs[j].n1 - jobs[j].n0 : 0;
        const uint64_t nf = N * (uint64_t)jobs[j].max_factors;
        HIPCHK(hipMalloc((void **)&dc[j], std::max<uint64_t>(N, 1) * sizeof(uint32_t)));
        HIPCHK(hipMalloc((void **)&df[j], std::max<uint64_t>(nf, 1) * sizeof(uint64_t)));
        if (N) {
            HIPCHK(hipMemcpyAsync(dc[j], counts[j], N * sizeof(uint32_t), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(df[j], factors[j], nf * sizeof(uint64_t), hipMemcpyHostToDevice, st));
        }
    }

    std::vector<JobPlan> plans;
    auto t_tune = clk::now();
    int rc = plan_jobs(E, jobs, njobs, dc.data(), df.data(), !(flags & CTD_BATCH_NO_AUTOTUNE), plans, &bs.tuned,
                       flags);
    if (rc != CTD_OK) return rc;
    bs.autotune_ms = std::chrono::duration<double, std::milli>(clk::now() - t_tune).count();

    // tiles, longest first, each to the least-loaded stream
    std::vector<Tile> tiles;
    for (uint32_t j = 0; j < njobs; ++j) {
        const JobPlan &P = plans[j];
        if (P.f0 >= P.f1) continue;
        const uint64_t lo0 = P.f0 & ~1ull;
        bool first = true;
        for (uint64_t blo = lo0; blo < P.f1; blo += tile_len) {
            const uint64_t bhi = block_end(blo, tile_len, P.f1);
            const uint64_t start = std::max(blo, P.f0);
            const double cost = est_primes(start, bhi) * ((double)P.N * P.ops + 300.0) + 2e6;
            tiles.push_back(Tile{j, blo, bhi, start, cost, 0u, first});
            first = false;
            if (bhi == P.f1) break;
        }
    }
    std::stable_sort(tiles.begin(), tiles.end(), [](const Tile &x, const Tile &y) { return x.cost > y.cost; });
    bs.tiles = (uint32_t)tiles.size();
    ensure_slots(E, nst, tile_len);
    if (E->tile_cap < 2 * tiles.size() || !E->d_tile) {
        (void)hipFree(E->d_tile);
        E->d_tile = nullptr;
        E->tile_cap = std::max<uint64_t>(2 * tiles.size(), 2);
        HIPCHK(hipMalloc((void **)&E->d_tile, E->tile_cap * sizeof(unsigned long long)));
    }
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, E->d_err, 1u);

    HIPCHK(hipEventRecord(E->ev0, st));
    mod2_only(plans, st);
    const bool graph = !(flags & CTD_BATCH_NO_GRAPH);
    struct Graphs {
        std::vector<hipGraphExec_t> x;
        std::vector<hipGraph_t> g;
        ~Graphs()
        {
            for (auto e : x) (void)hipGraphExecDestroy(e);
            for (auto e : g) (void)hipGraphDestroy(e);
        }
    } keep; // alive until the launches have completed
    for (size_t a = 0; a < tiles.size(); a += BATCH_GRAPH_TILES) {
        const size_t b = std::min(tiles.size(), a + BATCH_GRAPH_TILES);
        std::vector<double> load(nst, 0.0);
        for (size_t t = a; t < b; ++t) {
            uint32_t s = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
            tiles[t].stream = s;
            load[s] += tiles[t].cost;
        }
        if (!graph) {
            issue_tiles(E, plans, tiles, a, b, nst, st);
            continue;
        }
        auto t_build = clk::now();
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        HIPCHK(hipStreamBeginCapture(E->cap_stream, hipStreamCaptureModeThreadLocal));
        try {
            issue_tiles(E, plans, tiles, a, b, nst, E->cap_stream);
        } catch (...) {
            hipGraph_t dead = nullptr;
            (void)hipStreamEndCapture(E->cap_stream, &dead);
            if (dead) (void)hipGraphDestroy(dead);
            throw;
        }
        HIPCHK(hipStreamEndCapture(E->cap_stream, &g));
        keep.g.push_back(g);
        HIPCHK(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
        keep.x.push_back(x);
        bs.build_ms += std::chrono::duration<double, std::milli>(clk::now() - t_build).count();
        HIPCHK(hipGraphLaunch(x, st));
        bs.graphs++;
    }
    HIPCHK(hipEventRecord(E->ev1, st));
    HIPCHK(hipStreamSynchronize(st));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, E->ev0, E->ev1));
    bs.total_ms = ms;

    unsigned long long err = 0;
    HIPCHK(hipMemcpy(&err, E->d_err, sizeof err, hipMemcpyDeviceToHost));
    if (err) return fail(E, CTD_E_DEVICE, "prime slot overflow in a batch tile");
    std::vector<unsigned long long> tc(2 * tiles.size());
    if (!tiles.empty())
        HIPCHK(hipMemcpy(tc.data(), E->d_tile, tc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (size_t t = 0; t < tiles.size(); ++t) {
        bs.primes += tc[2 * t];
        bs.updates += tc[2 * t] * plans[tiles[t].job].N;
    }
    if (host) {
        for (uint32_t j = 0; j < njobs; ++j) {
            const uint64_t N = plans[j].N;
            if (!N) continue;
            HIPCHK(hipMemcpyAsync(counts[j], dc[j], N * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(factors[j], df[j], N * (uint64_t)jobs[j].max_factors * sizeof(uint64_t),
                                  hipMemcpyDeviceToHost, st));
        }
        HIPCHK(hipStreamSynchronize(st));
    }
    if (stats) *stats = bs;
    return CTD_OK;
}
