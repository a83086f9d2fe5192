// ctd_engine.hip -- MI355X trial-factoring engine behind the C ABI of ctd_hip.h.
//
// Replaces device_trial_factor<Seq> (trial_factoring.cuh:336-513): the segment
// loop now runs entirely on the device.  Per batch of the prime range:
//   1. sieve_kernel sieves 2^20-integer segments in LDS and streams the
//      primes into HBM (no host sieve, no PCIe prime copy);
//   2. tf kernel: one prime per lane, one index chunk per work item, the
//      family recurrence in registers, zero residues recorded with one atomic
//      per wave (ballot + popcount) into the factors_buffer layout.
// Prime 2 takes the mod-2 path (trial_factoring.cuh:139-154,415-418) as an
// index-parallel kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <atomic>
#include <ctime>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ctd_hip.h"
#include "ctd_families.hpp"
#include "ctd_sieve.hpp"
#include "ctd_tf_fast.hpp"
#include "ctd_probe.hpp"

using namespace ctd;

// When np_dev is given, the prime count is read from device memory (the
// sieve of the same stream wrote it) and nprimes is the capacity bounding it.
__device__ __forceinline__ uint64_t prime_count(uint64_t nprimes, const unsigned long long *np_dev)
{
    if (!np_dev) return nprimes;
    const uint64_t c = *np_dev;
    return c < nprimes ? c : nprimes;
}

// Generic kernel: any odd prime < 2^64, 64-bit Montgomery set-up.
template <uint32_t FAM>
__global__ void __launch_bounds__(256) tf_generic_kernel(JobArgs J, const uint64_t *__restrict__ primes,
                                                          uint64_t nprimes, const unsigned long long *np_dev)
{
    nprimes = prime_count(nprimes, np_dev);
    const uint64_t tr = (nprimes + 63) & ~63ull; // wave-aligned prime slots per chunk
    const uint64_t work = tr * J.nchunks;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < work;
         w += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t c = w / tr, i = w - c * tr;
        if (i >= nprimes) continue;
        const uint64_t p = primes[i];
        Mont M;
        M.init(p);
        const uint64_t ns = J.n0 + c * J.chunk;
        uint64_t ne = ns + J.chunk;
        if (ne > J.n1 || ne < ns) ne = J.n1;
        GenState s;
        Gen<FAM>::init(s, M, J, ns);
        record_hits(J, ns - J.n0, p, Gen<FAM>::zero(s, M));
        for (uint64_t n = ns; n + 1 < ne; ++n) {
            Gen<FAM>::step(s, M, n);
            record_hits(J, n + 1 - J.n0, p, Gen<FAM>::zero(s, M));
        }
        if (J.final_res) J.final_res[i] = Gen<FAM>::plain(s, M);
    }
}

// Fast kernel: primes below 2^63 with the family's narrow step (ctd_tf_fast.hpp).
template <uint32_t FAM>
__global__ void __launch_bounds__(256) tf_fast_kernel(JobArgs J, const uint64_t *__restrict__ primes,
                                                       uint64_t nprimes, const unsigned long long *np_dev)
{
    nprimes = prime_count(nprimes, np_dev);
    const uint64_t tr = (nprimes + 63) & ~63ull;
    const uint64_t work = tr * J.nchunks;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < work;
         w += (uint64_t)gridDim.x * blockDim.x) {
        // the chunk is wave-uniform (wave-aligned slots): keep it scalar
        const uint64_t c = __builtin_amdgcn_readfirstlane((uint32_t)(w / tr));
        const uint64_t i = w - c * tr;
        if (i >= nprimes) continue;
        const uint64_t p = primes[i];
        const uint64_t ns = J.n0 + c * J.chunk;
        uint64_t ne = ns + J.chunk;
        if (ne > J.n1 || ne < ns) ne = J.n1;
        Fast<FAM>::run(J, p, ns, ne, i);
    }
}

// Smarandache with two primes per lane: the lanes of a wave take the primes
// base + l and base + 64 + l of a 128-prime group and run the FP64 steps of
// both in one loop (run_fp_sm<2>), so the scalar stepping of the appended
// numbers, the loop control and the segment choices serve two updates per
// step.  A group that is not full, or whose primes do not all take the FP64
// form, runs its primes one after the other.
__global__ void __launch_bounds__(256) tf_sm_pair_kernel(JobArgs J, const uint64_t *__restrict__ primes,
                                                          uint64_t nprimes, const unsigned long long *np_dev)
{
    nprimes = prime_count(nprimes, np_dev);
    const uint64_t half = ((nprimes + 127) & ~127ull) / 2; // lanes of 128-prime groups per chunk
    const uint64_t work = half * J.nchunks;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < work;
         w += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t c = __builtin_amdgcn_readfirstlane((uint32_t)(w / half));
        const uint64_t j = w - c * half;
        const uint64_t i0 = ((j & ~63ull) << 1) + (j & 63), i1 = i0 + 64;
        if (i0 >= nprimes) continue;
        const uint64_t ns = J.n0 + c * J.chunk;
        uint64_t ne = ns + J.chunk;
        if (ne > J.n1 || ne < ns) ne = J.n1;
        const bool has1 = i1 < nprimes;
        const uint64_t p0 = primes[i0], p1 = has1 ? primes[i1] : p0;
        if (__ballot(!has1) == 0 && use_fp<FAM_SMARANDACHE>(J, p0, ne - ns - 1) &&
            use_fp<FAM_SMARANDACHE>(J, p1, ne - ns - 1)) {
            const uint64_t pp[2] = {p0, p1};
            uint64_t fin[2];
            run_fp_sm<2>(J, pp, ns, ne, ns - J.n0, fin);
            if (J.final_res) {
                J.final_res[i0] = fin[0];
                J.final_res[i1] = fin[1];
            }
        } else {
            Fast<FAM_SMARANDACHE>::run(J, p0, ns, ne, i0);
            if (has1) Fast<FAM_SMARANDACHE>::run(J, p1, ns, ne, i1);
        }
    }
}

// Prime 2: S(n) mod 2 per index.
__global__ void mod2_kernel(JobArgs J)
{
    const uint64_t N = J.n1 - J.n0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if (value_mod_2(J.fam, J.param, J.n0 + i) == 0) {
            uint32_t pos = atomicAdd(&J.counts[i], 1u);
            if (pos < J.maxf) J.factors[i * J.maxf + pos] = 2;
        }
    }
}

__global__ void zero_u64_kernel(unsigned long long *p, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void zero_words_kernel(uint32_t *p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// After a tile's sieve: keep its prime count and largest prime (or add the
// count to a running total), and flag a count above the slot's capacity (the
// sieve never writes past it).
__global__ void tile_done_kernel(const unsigned long long *sc, uint64_t cap, unsigned long long *err,
                                 unsigned long long *out, uint32_t accumulate)
{
    if (threadIdx.x == 0) {
        const unsigned long long n = sc[0];
        if (accumulate) {
            atomicAdd(out, n); // tiles of one job run on several streams
        } else {
            out[0] = n;
            out[1] = sc[1];
        }
        if (n > cap) atomicOr(err, 1ull);
    }
}

// Launch shape of a residue kernel: lanes per workgroup, work items (primes)
// each lane takes in turn (the grid holds items / (block * ppl) lanes; the
// sequential counterpart of the reference's B primes per thread,
// trial_factoring.cuh:472-481) and the extra index-chunk split.
struct LaunchShape {
    uint32_t block = 256;
    uint32_t ppl = 1;
    uint32_t split = 1;
};

// One resident prime slot of the batched form, with its own stream.
struct BatchSlot {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    uint64_t *primes = nullptr;
    uint64_t cap = 0;
    uint32_t *segcnt = nullptr;
    uint64_t segcap = 0;
    unsigned long long *sc = nullptr; // [0] count, [1] largest prime
};

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct ctd_hip_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::string name;
    uint64_t total_mem = 0;
    int cu = 0;
    // sieving base: odd primes <= base_lim
    uint64_t base_lim = 0;
    uint32_t *d_sp = nullptr;
    uint64_t *d_mu = nullptr;
    uint32_t nsp = 0, nsmall = 0, nwheel = 0;
    // prime stream and scratch
    uint64_t *d_primes = nullptr;
    uint64_t primes_cap = 0;
    uint32_t *d_segcnt = nullptr;
    uint64_t segcnt_cap = 0;
    uint64_t *d_segoff = nullptr;
    uint64_t segoff_cap = 0;
    unsigned long long *d_scalars = nullptr; // [0] total, [1] max prime
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_all0 = nullptr, ev_all1 = nullptr; // around a whole plain run
    void *scratch = nullptr; // this device's allocation-free workspace (device address)
    ctd_run_stats stats{};
    // batched form
    std::vector<BatchSlot> slots;
    hipStream_t cap_stream = nullptr;
    hipEvent_t ev_fork = nullptr;
    unsigned long long *d_err = nullptr;
    unsigned long long *d_tile = nullptr;
    uint64_t tile_cap = 0;
    uint32_t *d_scr_counts = nullptr;
    uint64_t *d_scr_factors = nullptr;
    uint64_t scr_cap = 0;
    std::map<uint64_t, LaunchShape> shapes;
};

static std::string g_open_error;

namespace {

struct HipError {
    hipError_t code;
    const char *what;
};

#define HIPCHK(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) throw HipError{e_, #call};                                        \
    } while (0)

int fail(ctd_hip_engine *E, int code, const std::string &msg)
{
    if (E) E->err = msg;
    return code;
}

int fail_hip(ctd_hip_engine *E, const HipError &h)
{
    std::string m = std::string(h.what) + ": " + hipGetErrorString(h.code);
    return fail(E, h.code == hipErrorOutOfMemory ? CTD_E_NOMEM : CTD_E_DEVICE, m);
}

template <typename T>
void ensure_buf(T *&ptr, uint64_t &cap, uint64_t need)
{
    if (need <= cap && ptr) return;
    if (ptr) HIPCHK(hipFree(ptr));
    ptr = nullptr;
    uint64_t n = std::max<uint64_t>(need, 1);
    HIPCHK(hipMalloc((void **)&ptr, n * sizeof(T)));
    cap = n;
}

uint64_t div_up(uint64_t a, uint64_t b) { return a / b + (a % b != 0); }

// Indices per work item at most: the index loops keep 32-bit counters.
constexpr uint64_t CHUNK_MAX = 1ull << 30;
// Segments per device batch or tile at most (2^40 integers): a requested
// length is rounded to whole segments and clamped to this.
constexpr uint64_t BLOCK_SEGS_MAX = 1ull << 20;

// The sieving numbers of one sieve pass (ascending odd numbers; nwheel of
// them <= WHEEL_MAX come from the wheel patterns, nsmall < SMALL_LIMIT).
struct SieveBase {
    const uint32_t *sp = nullptr;
    const uint64_t *mu = nullptr;
    uint32_t nsp = 0, nsmall = 0, nwheel = 0;
};

SieveArgs sieve_args(const SieveBase &B, uint64_t lo0, uint64_t start, uint64_t end, uint32_t mode)
{
    SieveArgs A{};
    A.lo0 = lo0;
    A.start = start;
    A.end = end;
    A.sp = B.sp;
    A.sp_mu = B.mu;
    A.nsp = B.nsp;
    A.nsmall = B.nsmall;
    A.nwheel = B.nwheel;
    A.mode = mode;
    return A;
}

// Count pass of an ordered sieve of [start, end): the per-segment prime
// counts go to the host (h_cnt), their exclusive prefix sums to h_off and to
// segoff on the device.  segcnt/segoff/h_cnt/h_off hold nseg entries, nseg =
// the segments of [start & ~1, end).  Returns the total.
uint64_t sieve_count_pass(hipStream_t st, const SieveBase &B, uint64_t start, uint64_t end, uint32_t *segcnt,
                          uint64_t *segoff, uint32_t *h_cnt, uint64_t *h_off, uint64_t nseg)
{
    SieveArgs A = sieve_args(B, start & ~1ull, start, end, SIEVE_COUNT);
    A.seg_counts = segcnt;
    hipLaunchKernelGGL(sieve_kernel, dim3((uint32_t)nseg), dim3(SIEVE_THREADS), 0, st, A);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h_cnt, segcnt, nseg * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    uint64_t total = 0;
    for (uint64_t i = 0; i < nseg; ++i) {
        h_off[i] = total;
        total += h_cnt[i];
    }
    HIPCHK(hipMemcpyAsync(segoff, h_off, nseg * sizeof(uint64_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st)); // h_off may be reused once this returns
    return total;
}

// Ordered pass after sieve_count_pass: the primes of [start, end) in
// ascending order into out32 or out64 (the count pass's total entries).
void sieve_order_pass(hipStream_t st, const SieveBase &B, uint64_t start, uint64_t end, uint32_t *segcnt,
                      const uint64_t *segoff, uint64_t nseg, uint32_t *out32, uint64_t *out64, uint64_t total)
{
    SieveArgs A = sieve_args(B, start & ~1ull, start, end, out32 ? SIEVE_ORDERED32 : SIEVE_ORDERED64);
    A.seg_counts = segcnt;
    A.seg_off = segoff;
    A.out32 = out32;
    A.out64 = out64;
    A.cap = total;
    hipLaunchKernelGGL(sieve_kernel, dim3((uint32_t)nseg), dim3(SIEVE_THREADS), 0, st, A);
    HIPCHK(hipGetLastError());
}

uint64_t sieve_segments(uint64_t start, uint64_t end) { return start < end ? div_up(end - (start & ~1ull), SEG_SPAN) : 0; }

// Sieves [start, end) into an ordered list (u32 or u64) allocated here.
uint64_t sieve_ordered(ctd_hip_engine *E, const SieveBase &B, uint64_t start, uint64_t end, uint32_t **out32,
                       uint64_t **out64)
{
    if (start >= end) return 0;
    const uint64_t nseg = sieve_segments(start, end);
    ensure_buf(E->d_segcnt, E->segcnt_cap, nseg);
    ensure_buf(E->d_segoff, E->segoff_cap, nseg);
    std::vector<uint32_t> cnt(nseg);
    std::vector<uint64_t> off(nseg);
    const uint64_t total = sieve_count_pass(E->stream, B, start, end, E->d_segcnt, E->d_segoff, cnt.data(),
                                            off.data(), nseg);
    if (out32) HIPCHK(hipMalloc((void **)out32, std::max<uint64_t>(total, 1) * sizeof(uint32_t)));
    else HIPCHK(hipMalloc((void **)out64, std::max<uint64_t>(total, 1) * sizeof(uint64_t)));
    sieve_order_pass(E->stream, B, start, end, E->d_segcnt, E->d_segoff, nseg, out32 ? *out32 : nullptr,
                     out32 ? nullptr : *out64, total);
    HIPCHK(hipStreamSynchronize(E->stream));
    return total;
}

// The small and wheel counts of an ascending sieving list of n numbers;
// h holds the first min(n, 4096) of them on the host.
void count_small(const uint32_t *h, uint32_t n, uint32_t &nsmall, uint32_t &nwheel)
{
    const uint32_t k = std::min<uint32_t>(n, 4096);
    nsmall = nwheel = 0;
    while (nsmall < k && h[nsmall] < SMALL_LIMIT) ++nsmall;
    while (nwheel < k && h[nwheel] <= WHEEL_MAX) ++nwheel;
    if (nsmall - nwheel > SMALL_MAX_COUNT) throw HipError{hipErrorInvalidValue, "sieving list: too many small numbers"};
}

void list_counts(hipStream_t st, const uint32_t *d_list, uint32_t n, uint32_t *h4096, SieveBase &B)
{
    const uint32_t k = std::min<uint32_t>(n, 4096);
    if (k) HIPCHK(hipMemcpyAsync(h4096, d_list, k * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    B.sp = d_list;
    B.nsp = n;
    count_small(h4096, n, B.nsmall, B.nwheel);
}

void make_mu(hipStream_t st, const uint32_t *sp, uint64_t *mu, uint32_t n)
{
    if (!n) return;
    hipLaunchKernelGGL(sieve_mu_kernel, dim3(div_up(n, 256)), dim3(256), 0, st, sp, mu, n);
    HIPCHK(hipGetLastError());
}

// Odd primes below 65536 sieve every window up to 2^32: they are the
// second level of every sieving base.
constexpr uint32_t L0_COUNT = 127;   // odd numbers 3..255 (level 0)
constexpr uint32_t L1_CAP = 6542;    // odd primes below 65536: 6541
constexpr uint64_t L1_END = 65537;

// Levels 0 and 1 of a sieving base into l0/m0 [L0_COUNT] and l1/m1 [L1_CAP]:
// the odd numbers 3..255 sieve [3, 65537) into the odd primes below 65536.
// segcnt/segoff/h_cnt/h_off need one entry.
SieveBase build_level1(hipStream_t st, uint32_t *l0, uint64_t *m0, uint32_t *l1, uint64_t *m1, uint32_t *segcnt,
                       uint64_t *segoff, uint32_t *h_cnt, uint64_t *h_off, uint32_t *h4096)
{
    hipLaunchKernelGGL(odd_numbers_kernel, dim3(1), dim3(128), 0, st, l0, L0_COUNT);
    make_mu(st, l0, m0, L0_COUNT);
    SieveBase B0{l0, m0, L0_COUNT, L0_COUNT, (WHEEL_MAX - 1) / 2};
    const uint64_t nseg = sieve_segments(3, L1_END);
    const uint64_t n1 = sieve_count_pass(st, B0, 3, L1_END, segcnt, segoff, h_cnt, h_off, nseg);
    if (n1 > L1_CAP) throw HipError{hipErrorInvalidValue, "sieving base: level-1 list overflow"};
    sieve_order_pass(st, B0, 3, L1_END, segcnt, segoff, nseg, l1, nullptr, n1);
    make_mu(st, l1, m1, (uint32_t)n1);
    SieveBase B1;
    B1.mu = m1;
    list_counts(st, l1, (uint32_t)n1, h4096, B1);
    return B1;
}

// Odd primes <= lim (lim < 2^32) as the engine's sieving base.
void build_base(ctd_hip_engine *E, uint64_t lim)
{
    if (lim < 65536) lim = 65536;
    if (E->d_sp && lim <= E->base_lim) return;
    // grow geometrically: consecutive batches of a window raise sqrt(f1) a
    // little each time, and sieving numbers beyond sqrt(end) cost a segment
    // nothing (the per-lane loop stops at the first one past it)
    if (E->d_sp) lim = std::min<uint64_t>(std::max<uint64_t>(lim, 2 * E->base_lim), 0xffffffffull);
    uint32_t *l0 = nullptr, *l1 = nullptr, *l2 = nullptr;
    uint64_t *m0 = nullptr, *m1 = nullptr, *m2 = nullptr;
    HIPCHK(hipMalloc((void **)&l0, L0_COUNT * sizeof(uint32_t)));
    HIPCHK(hipMalloc((void **)&m0, L0_COUNT * sizeof(uint64_t)));
    HIPCHK(hipMalloc((void **)&l1, L1_CAP * sizeof(uint32_t)));
    HIPCHK(hipMalloc((void **)&m1, L1_CAP * sizeof(uint64_t)));
    ensure_buf(E->d_segcnt, E->segcnt_cap, 1);
    ensure_buf(E->d_segoff, E->segoff_cap, 1);
    std::vector<uint32_t> h(4096);
    uint32_t hc;
    uint64_t ho;
    const SieveBase B1 = build_level1(E->stream, l0, m0, l1, m1, E->d_segcnt, E->d_segoff, &hc, &ho, h.data());
    // level 2: primes below 65536 sieve [3, lim + 1)
    const uint64_t n2 = sieve_ordered(E, B1, 3, lim + 1, &l2, nullptr);
    HIPCHK(hipMalloc((void **)&m2, std::max<uint64_t>(n2, 1) * sizeof(uint64_t)));
    make_mu(E->stream, l2, m2, (uint32_t)n2);
    HIPCHK(hipStreamSynchronize(E->stream));
    HIPCHK(hipFree(l0));
    HIPCHK(hipFree(m0));
    HIPCHK(hipFree(l1));
    HIPCHK(hipFree(m1));
    if (E->d_sp) HIPCHK(hipFree(E->d_sp));
    if (E->d_mu) HIPCHK(hipFree(E->d_mu));
    E->d_sp = l2;
    E->d_mu = m2;
    SieveBase B;
    B.mu = m2;
    list_counts(E->stream, l2, (uint32_t)n2, h.data(), B);
    E->nsp = B.nsp;
    E->nsmall = B.nsmall;
    E->nwheel = B.nwheel;
    E->base_lim = lim;
}

SieveBase engine_base(const ctd_hip_engine *E) { return SieveBase{E->d_sp, E->d_mu, E->nsp, E->nsmall, E->nwheel}; }

unsigned bitlen64(uint64_t x) { return x ? 64u - (unsigned)__builtin_clzll(x) : 0u; }

// Conditions of the smarandache 32-bit-quotient step (ctd_tf_fast.hpp):
// m = b^digits(n) < 2^21 for every appended n < n1, bitlen(p) <= 54,
// bitlen(m) + bitlen(p) + 1 <= 64, and n1 <= 2^32 <= p (n+1 < p).
// The fixed-n families solve for their hits once per prime (ctd_tf_fast.hpp)
// unless told to step every index; their index range is then never split.
bool closed_form(uint32_t fam, uint32_t job_flags)
{
    return (fam == FAM_PROTH_FIXED_N || fam == FAM_RIESEL_FIXED_N) && !(job_flags & JOB_STEPPED);
}

bool sm_narrow_ok(uint64_t b, uint64_t n1, uint64_t pmax)
{
    if (n1 > (1ull << 32)) return false;
    uint64_t m = b;
    while (m <= n1 && m < (1ull << 22)) m *= b; // b^digits(n1)
    if (m >= (1ull << 21)) return false;
    unsigned k = bitlen64(pmax);
    return k <= 54 && bitlen64(m) + k + 1 <= 64;
}

bool valid_job(ctd_hip_engine *E, const ctd_job *j, int *code)
{
    *code = CTD_E_INVALID;
    if (!j) return fail(E, CTD_E_INVALID, "job is NULL"), false;
    switch (j->seq_id) {
    case CTD_SEQ_MERSENNE: case CTD_SEQ_PROTH: case CTD_SEQ_RIESEL: case CTD_SEQ_SMARANDACHE:
    case CTD_SEQ_FIBONACCI: case CTD_SEQ_LUCAS: case CTD_SEQ_PROTH_FIXED_N: case CTD_SEQ_RIESEL_FIXED_N:
    case CTD_SEQ_TRIBONACCI: case CTD_SEQ_PEANO: break;
    default: return fail(E, CTD_E_INVALID, "Unexpected number sequence id: " + std::to_string(j->seq_id)), false;
    }
    if (j->max_factors == 0) return fail(E, CTD_E_INVALID, "max_factors must be >= 1"), false;
    if ((j->seq_id == CTD_SEQ_PROTH || j->seq_id == CTD_SEQ_RIESEL) && j->param == 0)
        return fail(E, CTD_E_INVALID, std::string(j->seq_id == CTD_SEQ_PROTH ? "Proth" : "Riesel") +
                                          " parameter k must be non-zero"), false;
    if ((j->seq_id == CTD_SEQ_PROTH_FIXED_N || j->seq_id == CTD_SEQ_RIESEL_FIXED_N) && j->param == 0)
        return fail(E, CTD_E_INVALID, "Invalid n value zero for Riesel/Proth fixed n"), false;
    if (j->seq_id == CTD_SEQ_SMARANDACHE && (j->param < 2 || j->param > 64))
        return fail(E, CTD_E_INVALID, "Base `" + std::to_string(j->param) +
                                          "' is not supported for Smarandache numbers"), false;
    *code = CTD_OK;
    return true;
}

template <uint32_t FAM>
void launch_tf(const JobArgs &J, const uint64_t *primes, uint64_t nprimes, const unsigned long long *np_dev,
               bool fast, dim3 grid, dim3 block, hipStream_t st)
{
    if (fast && FAM == FAM_SMARANDACHE && !(J.flags & JOB_SM_NO_PAIR))
        hipLaunchKernelGGL(tf_sm_pair_kernel, grid, block, 0, st, J, primes, nprimes, np_dev);
    else if (fast)
        hipLaunchKernelGGL(tf_fast_kernel<FAM>, grid, block, 0, st, J, primes, nprimes, np_dev);
    else
        hipLaunchKernelGGL(tf_generic_kernel<FAM>, grid, block, 0, st, J, primes, nprimes, np_dev);
}

// nprimes is the count, or the capacity when np_dev holds the count on the device.
void launch_tf_any(const JobArgs &J, const uint64_t *primes, uint64_t nprimes, const unsigned long long *np_dev,
                   bool fast, dim3 grid, dim3 block, hipStream_t st)
{
    switch (J.fam) {
    case FAM_MERSENNE: launch_tf<FAM_MERSENNE>(J, primes, nprimes, np_dev, fast, grid, block, st); break;
    case FAM_PROTH: launch_tf<FAM_PROTH>(J, primes, nprimes, np_dev, fast, grid, block, st); break;
    case FAM_RIESEL: launch_tf<FAM_RIESEL>(J, primes, nprimes, np_dev, fast, grid, block, st); break;
    case FAM_SMARANDACHE: launch_tf<FAM_SMARANDACHE>(J, primes, nprimes, np_dev, fast, grid, block, st); break;
    case FAM_FIBONACCI: launch_tf<FAM_FIBONACCI>(J, primes, nprimes, np_dev, fast, grid, block, st); break;
    case FAM_LUCAS: launch_tf<FAM_LUCAS>(J, primes, nprimes, np_dev, fast, grid, block, st); break;
    case FAM_PROTH_FIXED_N: launch_tf<FAM_PROTH_FIXED_N>(J, primes, nprimes, np_dev, fast, grid, block, st); break;
    case FAM_RIESEL_FIXED_N: launch_tf<FAM_RIESEL_FIXED_N>(J, primes, nprimes, np_dev, fast, grid, block, st); break;
    case FAM_TRIBONACCI: launch_tf<FAM_TRIBONACCI>(J, primes, nprimes, np_dev, fast, grid, block, st); break;
    default: launch_tf<FAM_PEANO>(J, primes, nprimes, np_dev, fast, grid, block, st); break;
    }
}

struct HostSync {
    uint32_t *h_counts;
    uint64_t *h_factors;
};

// Device buffers of the segment loop: the sieving base, one batch's prime
// stream, its per-segment counts and the batch's two scalars (count, largest
// prime).  The engine's own buffers grow on demand; a fixed workspace (the
// allocation-free entry point's) is used as it is, with batches of at most
// batch_max integers.
struct Workspace {
    SieveBase base;
    uint64_t *primes = nullptr;
    uint64_t primes_cap = 0;
    uint32_t *segcnt = nullptr;
    uint64_t segcnt_cap = 0;
    unsigned long long *scalars = nullptr;
    uint64_t batch_max = ~0ull;
};

// The segment loop.  d_counts/d_factors hold the job's factors buffer.
// fixed: the workspace to use (its base covers sqrt(f1)), or null for the
// engine's own buffers.
int run_job(ctd_hip_engine *E, const ctd_job *job, const ctd_run_opts *opts, const ctd_callbacks *cb,
            uint32_t *d_counts, uint64_t *d_factors, uint64_t *last_prime, hipStream_t st, const HostSync *hs,
            const Workspace *fixed = nullptr)
{
    int code;
    if (!valid_job(E, job, &code)) return code;
    // --block-size goes to the residue launch as given (trial_factoring.cuh:280-286);
    // the kernels run whole 64-lane waves, at most 256 lanes per workgroup
    if (opts && opts->block_size && (opts->block_size < 64 || opts->block_size > 256 || opts->block_size % 64))
        return fail(E, CTD_E_UNSUPPORTED, "block size " + std::to_string(opts->block_size) +
                                              " is not supported: the residue kernels take 64, 128, 192 or 256 "
                                              "threads per workgroup");
    if (opts && opts->grid_size < 0)
        return fail(E, CTD_E_INVALID, "grid size " + std::to_string(opts->grid_size) + " is not positive");
    E->stats = ctd_run_stats{};
    if (last_prime) *last_prime = 0;
    if (job->n1 <= job->n0 || job->f1 <= job->f0) return CTD_OK; // trial_factoring_impl.hpp:380-387
    const uint64_t N = job->n1 - job->n0;
    JobArgs J{};
    J.fam = job->seq_id;
    J.maxf = job->max_factors;
    J.param = job->param;
    J.n0 = job->n0;
    J.n1 = job->n1;
    J.counts = d_counts;
    J.factors = d_factors;

    HIPCHK(hipEventRecord(E->ev_all0, st));

    uint64_t f0 = job->f0;
    const uint64_t f1 = job->f1;
    if (f0 <= 2 && f1 > 2) {
        J.chunk = N;
        J.nchunks = 1;
        uint32_t g = (uint32_t)std::min<uint64_t>(div_up(N, 256), 4096);
        hipLaunchKernelGGL(mod2_kernel, dim3(g), dim3(256), 0, st, J);
        HIPCHK(hipGetLastError());
        f0 = 3;
        if (last_prime) *last_prime = 2;
    }
    uint64_t lastp = (job->f0 <= 2 && f1 > 2) ? 2 : 0;
    if (f0 < f1) {
        Workspace ws;
        if (fixed) {
            ws = *fixed;
        } else {
            build_base(E, isqrt_u64(f1 - 1));
            ws.base = engine_base(E);
        }
        const bool generic_only = opts && (opts->flags & CTD_RUN_GENERIC_ONLY);
        const bool fast = !generic_only && Fast<0>::fits(f1 - 1);
        const uint64_t blen = std::min(whole_spans((opts && opts->batch_length) ? opts->batch_length : (1ull << 31),
                                                   SEG_SPAN, BLOCK_SEGS_MAX),
                                       ws.batch_max);
        const uint64_t lo0 = f0 & ~1ull;
        for (uint64_t blo = lo0; blo < f1;) {
            const uint64_t bhi = block_end(blo, blen, f1);
            uint64_t bstart = std::max(blo, f0);
            uint64_t len = bhi - blo;
            uint64_t nseg = div_up(len, SEG_SPAN);
            // Montgomery-Vaughan: pi(x+y) - pi(x) <= 2y / log y
            double ly = std::log((double)std::max<uint64_t>(len, 16));
            uint64_t cap = (uint64_t)(2.0 * (double)len / ly) + 4096;
            if (!fixed) {
                ensure_buf(E->d_primes, E->primes_cap, cap);
                ensure_buf(E->d_segcnt, E->segcnt_cap, nseg);
                if (!E->d_scalars) HIPCHK(hipMalloc((void **)&E->d_scalars, 2 * sizeof(unsigned long long)));
                ws.primes = E->d_primes;
                ws.primes_cap = E->primes_cap;
                ws.segcnt = E->d_segcnt;
                ws.segcnt_cap = E->segcnt_cap;
                ws.scalars = E->d_scalars;
            } else if (nseg > ws.segcnt_cap) {
                return fail(E, CTD_E_INVALID, "batch of " + std::to_string(len) + " integers exceeds the workspace");
            }
            hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, ws.scalars, 2u);
            SieveArgs A = sieve_args(ws.base, blo, bstart, bhi, SIEVE_DENSE);
            A.seg_counts = ws.segcnt;
            A.out64 = ws.primes;
            A.total = ws.scalars;
            A.maxp = ws.scalars + 1;
            A.cap = ws.primes_cap;
            HIPCHK(hipEventRecord(E->ev0, st));
            hipLaunchKernelGGL(sieve_kernel, dim3((uint32_t)nseg), dim3(SIEVE_THREADS), 0, st, A);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(E->ev1, st));
            unsigned long long sc[2];
            HIPCHK(hipMemcpyAsync(sc, ws.scalars, sizeof sc, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, E->ev0, E->ev1));
            E->stats.sieve_ms += ms;
            E->stats.sieve_launches++;
            uint64_t np = sc[0];
            if (np > ws.primes_cap)
                return fail(E, CTD_E_DEVICE, "prime buffer overflow in batch starting at " + std::to_string(blo));
            if (np) {
                // index chunking so that small prime batches still fill the device
                const uint32_t stepped = (opts && (opts->flags & CTD_RUN_STEPPED)) ? JOB_STEPPED : 0u;
                uint64_t chunk = N, nch = 1;
                if (closed_form(J.fam, stepped)) {
                    // one solve per prime covers every index
                } else if (opts && opts->chunk_length) {
                    chunk = opts->chunk_length;
                    nch = div_up(N, chunk);
                } else {
                    const uint64_t target = (uint64_t)E->cu * 4096;
                    if (np < target && N > 2048) {
                        nch = std::min<uint64_t>(div_up(target, np), div_up(N, 1024));
                        chunk = div_up(N, nch);
                        nch = div_up(N, chunk);
                    }
                }
                if (chunk > CHUNK_MAX) { // the index loops count in 32 bits
                    nch = div_up(N, CHUNK_MAX);
                    chunk = div_up(N, nch);
                }
                J.chunk = chunk;
                J.nchunks = (uint32_t)div_up(N, chunk);
                J.flags = (J.fam == FAM_SMARANDACHE && sm_narrow_ok(J.param, J.n1, bhi - 1)) ? JOB_SM_NARROW : 0u;
                if (bhi - 1 < FP_PMAX && !(opts && (opts->flags & CTD_RUN_NO_FP64))) J.flags |= JOB_FP64;
                if (opts && (opts->flags & CTD_RUN_SM_FULL_STEP)) J.flags |= JOB_SM_FULL_STEP;
                if (opts && (opts->flags & CTD_RUN_SM_NO_SHORT)) J.flags |= JOB_SM_NO_SHORT;
                if (opts && (opts->flags & CTD_RUN_SM_NO_SHIFT)) J.flags |= JOB_SM_NO_SHIFT;
                if (opts && (opts->flags & CTD_RUN_SM_NO_PAIR)) J.flags |= JOB_SM_NO_PAIR;
                J.flags |= stepped;
                const uint32_t block = (opts && opts->block_size > 0) ? (uint32_t)opts->block_size : 256u;
                uint64_t items = ((np + 63) & ~63ull) * nch;
                uint64_t grid = (opts && opts->grid_size > 0) ? (uint64_t)opts->grid_size : div_up(items, block);
                grid = std::min<uint64_t>(grid, 1ull << 30);
                HIPCHK(hipEventRecord(E->ev0, st));
                launch_tf_any(J, ws.primes, np, nullptr, fast, dim3((uint32_t)grid), dim3(block), st);
                HIPCHK(hipGetLastError());
                HIPCHK(hipEventRecord(E->ev1, st));
                HIPCHK(hipStreamSynchronize(st));
                HIPCHK(hipEventElapsedTime(&ms, E->ev0, E->ev1));
                E->stats.tf_ms += ms;
                E->stats.tf_launches++;
                E->stats.primes += np;
                E->stats.updates += np * N;
                lastp = sc[1];
            }
            if (last_prime) *last_prime = lastp;
            if (cb && cb->progress) cb->progress(cb->user, bhi, lastp);
            if (cb && cb->checkpoint && hs && cb->checkpoint_due && cb->checkpoint_due(cb->user)) {
                HIPCHK(hipMemcpyAsync(hs->h_counts, d_counts, N * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
                HIPCHK(hipMemcpyAsync(hs->h_factors, d_factors, N * (uint64_t)job->max_factors * sizeof(uint64_t),
                                      hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                cb->checkpoint(cb->user, lastp);
            }
            blo = bhi;
        }
    }
    HIPCHK(hipEventRecord(E->ev_all1, st));
    HIPCHK(hipStreamSynchronize(st));
    float tot = 0;
    HIPCHK(hipEventElapsedTime(&tot, E->ev_all0, E->ev_all1));
    E->stats.total_ms = tot;
    if (last_prime) *last_prime = lastp;
    return CTD_OK;
}

#include "ctd_batch.hpp"
#include "ctd_serve.hpp"

// ---- the allocation-free workspace -----------------------------------------
// One fixed device array per GPU holds everything the plain form allocates
// for itself: the sieving base (odd primes below 2^24, enough for prime
// windows below 2^48), one batch's prime stream (batches of 2^25 integers:
// at most 2 * 2^25 / ln 2^25 + 4096 < 2^22 primes, Montgomery-Vaughan), the
// per-segment counts and the batch scalars.
constexpr uint64_t NA_F1_MAX = 1ull << 48;
constexpr uint32_t NA_BASE_CAP = (1u << 20) + (1u << 15); // >= 1,077,870 odd primes below 2^24
constexpr uint64_t NA_BATCH = 1ull << 25;
constexpr uint64_t NA_PRIME_CAP = 1ull << 22;
constexpr uint64_t NA_SEGS = 64; // segments of one pass at most: 32 per batch, 16 for the base
constexpr uint64_t na_align(uint64_t x) { return (x + 255) & ~255ull; }
constexpr uint64_t NA_L0 = 0;
constexpr uint64_t NA_M0 = na_align(NA_L0 + L0_COUNT * 4);
constexpr uint64_t NA_L1 = na_align(NA_M0 + L0_COUNT * 8);
constexpr uint64_t NA_M1 = na_align(NA_L1 + L1_CAP * 4);
constexpr uint64_t NA_SP = na_align(NA_M1 + L1_CAP * 8);
constexpr uint64_t NA_MU = na_align(NA_SP + NA_BASE_CAP * 4ull);
constexpr uint64_t NA_SEGCNT = na_align(NA_MU + NA_BASE_CAP * 8ull);
constexpr uint64_t NA_SEGOFF = na_align(NA_SEGCNT + NA_SEGS * 4);
constexpr uint64_t NA_SCALARS = na_align(NA_SEGOFF + NA_SEGS * 8);
constexpr uint64_t NA_PRIMES = na_align(NA_SCALARS + 2 * 8);
constexpr uint64_t NA_BYTES = na_align(NA_PRIMES + NA_PRIME_CAP * 8);
static_assert(NA_BYTES == CTD_NOALLOC_WORKSPACE_BYTES, "workspace size differs from the header's");
static_assert(NA_BATCH / SEG_SPAN + 1 <= NA_SEGS && (1ull << 24) / SEG_SPAN + 1 <= NA_SEGS, "segment counts");

// The array itself.  In host code its name is only a stand-in for the
// symbol: kernels get the device address hipGetSymbolAddress returns for it
// (g_scratch, per device), never the name.
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

std::mutex g_scratch_mu;
void *g_scratch[CTD_MAX_DEVICES] = {};

// This device's workspace address (the device is current), looked up once.
void *scratch_address(int device)
{
    if (device < 0 || device >= CTD_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[device] = p;
    }
    return g_scratch[device];
}

// The workspace of the allocation-free form at device address W, with the
// odd primes <= max(lim, 65536) (lim < 2^24) built into it.  Allocates
// nothing: the host side uses arrays on the stack.
Workspace scratch_workspace(hipStream_t st, unsigned char *W, uint64_t lim)
{
    if (lim < 65536) lim = 65536;
    uint32_t h_cnt[NA_SEGS];
    uint64_t h_off[NA_SEGS];
    uint32_t h4096[4096];
    Workspace ws;
    ws.segcnt = (uint32_t *)(W + NA_SEGCNT);
    ws.segcnt_cap = NA_SEGS;
    uint64_t *segoff = (uint64_t *)(W + NA_SEGOFF);
    const SieveBase B1 = build_level1(st, (uint32_t *)(W + NA_L0), (uint64_t *)(W + NA_M0), (uint32_t *)(W + NA_L1),
                                      (uint64_t *)(W + NA_M1), ws.segcnt, segoff, h_cnt, h_off, h4096);
    const uint64_t nseg = sieve_segments(3, lim + 1);
    if (nseg > NA_SEGS) throw HipError{hipErrorInvalidValue, "sieving base beyond the workspace"};
    uint32_t *sp = (uint32_t *)(W + NA_SP);
    uint64_t *mu = (uint64_t *)(W + NA_MU);
    const uint64_t n = sieve_count_pass(st, B1, 3, lim + 1, ws.segcnt, segoff, h_cnt, h_off, nseg);
    if (n > NA_BASE_CAP) throw HipError{hipErrorInvalidValue, "sieving base beyond the workspace"};
    sieve_order_pass(st, B1, 3, lim + 1, ws.segcnt, segoff, nseg, sp, nullptr, n);
    make_mu(st, sp, mu, (uint32_t)n);
    ws.base.mu = mu;
    list_counts(st, sp, (uint32_t)n, h4096, ws.base);
    ws.primes = (uint64_t *)(W + NA_PRIMES);
    ws.primes_cap = NA_PRIME_CAP;
    ws.scalars = (unsigned long long *)(W + NA_SCALARS);
    ws.batch_max = NA_BATCH;
    return ws;
}

} // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int ctd_hip_abi_version(void) { return CTD_HIP_ABI_VERSION; }

int ctd_hip_open(int device, ctd_hip_engine **out)
{
    if (!out) return CTD_E_INVALID;
    *out = nullptr;
    ctd_hip_engine *E = new ctd_hip_engine;
    try {
        int n = 0;
        HIPCHK(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) {
            g_open_error = "Error selecting device " + std::to_string(device) + ": no such device";
            delete E;
            return CTD_E_INVALID;
        }
        E->device = device;
        HIPCHK(hipSetDevice(device));
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        E->name = prop.name[0] ? prop.name : prop.gcnArchName;
        E->total_mem = prop.totalGlobalMem;
        E->cu = prop.multiProcessorCount;
        HIPCHK(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&E->ev0));
        HIPCHK(hipEventCreate(&E->ev1));
        HIPCHK(hipEventCreate(&E->ev_all0));
        HIPCHK(hipEventCreate(&E->ev_all1));
        E->scratch = scratch_address(device);
    } catch (const HipError &h) {
        g_open_error = std::string(h.what) + ": " + hipGetErrorString(h.code);
        delete E;
        return CTD_E_DEVICE;
    }
    *out = E;
    return CTD_OK;
}

void ctd_hip_close(ctd_hip_engine *E)
{
    if (!E) return;
    (void)hipSetDevice(E->device);
    if (E->stream) (void)hipStreamSynchronize(E->stream);
    (void)hipFree(E->d_sp);
    (void)hipFree(E->d_mu);
    (void)hipFree(E->d_primes);
    (void)hipFree(E->d_segcnt);
    (void)hipFree(E->d_segoff);
    (void)hipFree(E->d_scalars);
    release_batch(E);
    if (E->ev0) (void)hipEventDestroy(E->ev0);
    if (E->ev1) (void)hipEventDestroy(E->ev1);
    if (E->ev_all0) (void)hipEventDestroy(E->ev_all0);
    if (E->ev_all1) (void)hipEventDestroy(E->ev_all1);
    if (E->stream) (void)hipStreamDestroy(E->stream);
    delete E;
}

const char *ctd_hip_last_error(const ctd_hip_engine *E) { return E ? E->err.c_str() : g_open_error.c_str(); }

int ctd_hip_device_info(ctd_hip_engine *E, char *name, size_t name_len, uint64_t *total_mem, int32_t *cu_count)
{
    if (!E) return CTD_E_INVALID;
    if (name && name_len) {
        std::snprintf(name, name_len, "%s", E->name.c_str());
    }
    if (total_mem) *total_mem = E->total_mem;
    if (cu_count) *cu_count = E->cu;
    return CTD_OK;
}

int ctd_hip_last_stats(const ctd_hip_engine *E, ctd_run_stats *out)
{
    if (!E || !out) return CTD_E_INVALID;
    *out = E->stats;
    return CTD_OK;
}

int ctd_hip_sieve(ctd_hip_engine *E, uint64_t start, uint64_t end, uint64_t *out, uint64_t cap, uint64_t *count)
{
    if (!E || !count) return fail(E, CTD_E_INVALID, "invalid argument");
    *count = 0;
    if (start >= end) return CTD_OK;
    try {
        HIPCHK(hipSetDevice(E->device));
        uint64_t n2 = (start <= 2 && end > 2) ? 1 : 0;
        uint64_t s = std::max<uint64_t>(start, 3);
        uint64_t n = 0;
        uint64_t *d = nullptr;
        if (s < end) {
            build_base(E, isqrt_u64(end - 1));
            n = sieve_ordered(E, engine_base(E), s, end, nullptr, &d);
        }
        *count = n + n2;
        if (out && cap) {
            if (n2) out[0] = 2;
            uint64_t k = std::min<uint64_t>(n, cap - n2 > cap ? 0 : cap - n2);
            if (k) HIPCHK(hipMemcpy(out + n2, d, k * sizeof(uint64_t), hipMemcpyDeviceToHost));
        }
        if (d) HIPCHK(hipFree(d));
    } catch (const HipError &h) {
        return fail_hip(E, h);
    }
    return CTD_OK;
}

int ctd_hip_run_batch(ctd_hip_engine *E, const ctd_job *jobs, uint32_t njobs, uint32_t *const *counts,
                      uint64_t *const *factors, const ctd_batch_opts *opts, ctd_batch_stats *stats, void *stream)
{
    if (!E) return CTD_E_INVALID;
    if (stats) *stats = ctd_batch_stats{};
    if (!njobs) return CTD_OK;
    if (!jobs) return fail(E, CTD_E_INVALID, "jobs is NULL");
    try {
        HIPCHK(hipSetDevice(E->device));
        return run_batch(E, jobs, njobs, counts, factors, opts, stats, stream ? (hipStream_t)stream : E->stream);
    } catch (const HipError &h) {
        return fail_hip(E, h);
    }
}

int ctd_hip_batch_shape(ctd_hip_engine *E, const ctd_job *job, uint32_t *block, uint32_t *primes_per_lane,
                        uint32_t *chunk_split)
{
    if (!E || !job) return CTD_E_INVALID;
    try {
        HIPCHK(hipSetDevice(E->device));
        // a plan needs factors buffers: the probe writes to scratch only
        const uint64_t N = job->n1 > job->n0 ? job->n1 - job->n0 : 0;
        ensure_scratch(E, std::max<uint64_t>(N, 1));
        uint32_t *c = E->d_scr_counts;
        uint64_t *f = E->d_scr_factors;
        std::vector<JobPlan> plans;
        uint32_t tuned = 0;
        int rc = plan_jobs(E, job, 1, &c, &f, true, plans, &tuned);
        if (rc != CTD_OK) return rc;
        if (block) *block = plans[0].shape.block;
        if (primes_per_lane) *primes_per_lane = plans[0].shape.ppl;
        if (chunk_split) *chunk_split = plans[0].shape.split;
        return CTD_OK;
    } catch (const HipError &h) {
        return fail_hip(E, h);
    }
}

struct ctd_hip_server {
    ctd_hip_server_impl impl;
};

int ctd_hip_server_open(ctd_hip_engine *E, const ctd_server_opts *opts, ctd_hip_server **out)
{
    if (!E || !out) return fail(E, CTD_E_INVALID, "invalid argument");
    *out = nullptr;
    ctd_hip_server *srv = new ctd_hip_server;
    ctd_hip_server_impl *S = &srv->impl;
    S->E = E;
    if (opts) S->opts = *opts;
    if (!S->opts.memory_budget) S->opts.memory_budget = SERVE_BUDGET_DEFAULT;
    if (!S->opts.tile_length) S->opts.tile_length = SERVE_TILE_DEFAULT;
    S->opts.tile_length = whole_spans(S->opts.tile_length, SEG_SPAN, BLOCK_SEGS_MAX);
    // 3 streams: measured on the 10^4-job Poisson stream, 2 give the most
    // capacity, 3 nearly as much with the lowest p99; from 4 on, the sieve
    // workgroups of one stream wait for CUs held by the others' residue kernels
    if (!S->opts.streams) S->opts.streams = SERVE_STREAMS_DEFAULT;
    if (S->opts.depth) S->stream_work = std::min<uint32_t>(S->opts.depth, 16) * SERVE_TILE_WORK;
    S->opts.streams = std::min<uint32_t>(S->opts.streams, 16);
    if (!S->opts.queue_limit) S->opts.queue_limit = SERVE_QUEUE_DEFAULT;
    S->pool.limit = S->opts.memory_budget;
    S->pinned.limit = S->opts.memory_budget;
    try {
        HIPCHK(hipSetDevice(E->device));
        if (!E->d_err) HIPCHK(hipMalloc((void **)&E->d_err, sizeof(unsigned long long)));
        hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, E->stream, E->d_err, 1u);
        HIPCHK(hipStreamSynchronize(E->stream));
        // CTD_SERVE_OWN_QUEUES: a stream with a CU mask (here every CU) gets
        // a hardware queue of its own instead of sharing one of the
        // GPU_MAX_HW_QUEUES queues with the process's other streams, where
        // two busy streams run in turn
        const std::vector<uint32_t> all_cus((E->cu + 31) / 32, 0xffffffffu);
        auto make_stream = [&](hipStream_t *st) {
            if (S->opts.flags & CTD_SERVE_OWN_QUEUES)
                HIPCHK(hipExtStreamCreateWithCUMask(st, (uint32_t)all_cus.size(), all_cus.data()));
            else
                HIPCHK(hipStreamCreateWithFlags(st, hipStreamNonBlocking));
        };
        make_stream(&S->copy_st);
        S->slots.resize(S->opts.streams);
        S->inflight.resize(S->opts.streams);
        S->inflight_work.assign(S->opts.streams, 0.0);
        for (auto &sl : S->slots) {
            make_stream(&sl.st);
            HIPCHK(hipMalloc((void **)&sl.sc, 2 * sizeof(unsigned long long)));
            ensure_slot_buffers(sl, S->opts.tile_length);
        }
        S->th = std::thread(serve_loop, S);
    } catch (const HipError &h) {
        release_server(S);
        delete srv;
        return fail_hip(E, h);
    }
    *out = srv;
    return CTD_OK;
}

int ctd_hip_server_submit(ctd_hip_server *srv, const ctd_job *job, uint32_t *counts, uint64_t *factors,
                          uint64_t *ticket)
{
    if (!srv || !job || !ticket) return CTD_E_INVALID;
    ctd_hip_server_impl *S = &srv->impl;
    int code;
    if (!valid_job(S->E, job, &code)) return code;
    const uint64_t N = job->n1 > job->n0 ? job->n1 - job->n0 : 0;
    if (N && (!counts || !factors)) return fail(S->E, CTD_E_INVALID, "counts/factors are NULL");
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->fatal != CTD_OK) return fail(S->E, S->fatal, S->fatal_msg);
    if (S->pending.size() >= S->opts.queue_limit) return fail(S->E, CTD_E_BUSY, "serving queue is full");
    ServeJob *J = new ServeJob;
    J->ticket = S->next_ticket++;
    J->job = *job;
    J->h_counts = counts;
    J->h_factors = factors;
    J->t_submit = std::chrono::steady_clock::now();
    S->tickets[J->ticket] = Ticket{};
    S->pending.push_back(J);
    S->stats.submitted++;
    S->stats.pending_max = std::max<uint32_t>(S->stats.pending_max, (uint32_t)S->pending.size());
    *ticket = J->ticket;
    wake_scheduler(S);
    return CTD_OK;
}

int ctd_hip_server_poll(ctd_hip_server *srv, uint64_t ticket, int *state, double *latency_ms)
{
    if (!srv || !state) return CTD_E_INVALID;
    ctd_hip_server_impl *S = &srv->impl;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->tickets.find(ticket);
    if (it == S->tickets.end()) return fail(S->E, CTD_E_INVALID, "unknown ticket");
    Ticket &t = it->second; // stays valid while deliver() drops the lock
    if (t.state == 3) {
        // results staged: copy them into the caller's buffers now (another
        // thread may be doing it; then report the job as still running)
        deliver(S, t, lk);
    }
    *state = t.state == 3 ? 1 : t.state;
    if (latency_ms) *latency_ms = t.latency_ms;
    return CTD_OK;
}

int ctd_hip_server_wait(ctd_hip_server *srv, uint64_t ticket, double *latency_ms)
{
    if (!srv) return CTD_E_INVALID;
    ctd_hip_server_impl *S = &srv->impl;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->tickets.find(ticket);
    if (it == S->tickets.end()) return fail(S->E, CTD_E_INVALID, "unknown ticket");
    // a reference, not the iterator: submits on other threads may rehash the
    // map while this one waits, which keeps references valid but not iterators
    Ticket &t = it->second;
    for (;;) {
        S->cv_done.wait(lk, [&] { return t.state >= 2 || t.state < 0; });
        if (t.state != 3) break;
        deliver(S, t, lk); // returns at once when another thread delivers it
        if (t.state == 3) S->cv_done.wait(lk, [&] { return t.state != 3; });
    }
    if (latency_ms) *latency_ms = t.latency_ms;
    if (t.state < 0) return fail(S->E, t.state, S->fatal_msg);
    return CTD_OK;
}

int ctd_hip_server_stats(ctd_hip_server *srv, ctd_server_stats *out)
{
    if (!srv || !out) return CTD_E_INVALID;
    ctd_hip_server_impl *S = &srv->impl;
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->stats;
    out->pending = (uint32_t)S->pending.size();
    out->pinned_bytes_peak = S->pinned.peak;
    out->scheduler_cpu_ms = S->sched_cpu_ms.load(std::memory_order_relaxed);
    out->scheduler_passes = S->passes.load(std::memory_order_relaxed);
    out->device_bytes_peak = S->dev_peak.load(std::memory_order_relaxed);
    return CTD_OK;
}

int ctd_hip_server_close(ctd_hip_server *srv)
{
    if (!srv) return CTD_E_INVALID;
    ctd_hip_server_impl *S = &srv->impl;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    wake_scheduler(S);
    if (S->th.joinable()) S->th.join();
    {
        // results nobody collected still reach the caller's buffers
        std::unique_lock<std::mutex> lk(S->mu);
        std::vector<Ticket *> staged;
        for (auto &kv : S->tickets)
            if (kv.second.state == 3) staged.push_back(&kv.second);
        for (Ticket *t : staged) deliver(S, *t, lk);
    }
    int rc = S->fatal;
    if (rc == CTD_OK) {
        unsigned long long err = 0;
        if (hipMemcpy(&err, S->E->d_err, sizeof err, hipMemcpyDeviceToHost) != hipSuccess || err)
            rc = fail(S->E, CTD_E_DEVICE, "prime slot overflow in a served tile");
    } else {
        S->E->err = S->fatal_msg;
    }
    release_server(S);
    delete srv;
    return rc;
}

int ctd_hip_trial_factor_device(ctd_hip_engine *E, const ctd_job *job, const ctd_run_opts *opts,
                                uint32_t *d_counts, uint64_t *d_factors, uint64_t *last_prime, void *stream)
{
    if (!E) return CTD_E_INVALID;
    try {
        HIPCHK(hipSetDevice(E->device));
        hipStream_t st = stream ? (hipStream_t)stream : E->stream;
        return run_job(E, job, opts, nullptr, d_counts, d_factors, last_prime, st, nullptr);
    } catch (const HipError &h) {
        return fail_hip(E, h);
    }
}

int ctd_hip_trial_factor_device_noalloc(ctd_hip_engine *E, const ctd_job *job, const ctd_run_opts *opts,
                                        uint32_t *d_counts, uint64_t *d_factors, uint64_t *last_prime, void *stream)
{
    if (!E) return CTD_E_INVALID;
    int code;
    if (!valid_job(E, job, &code)) return code;
    if (job->f1 > CTD_NOALLOC_F1_MAX)
        return fail(E, CTD_E_UNSUPPORTED, "the allocation-free form takes prime windows up to 2^48 (its fixed "
                                          "workspace holds the sieving primes below 2^24); f1 = " +
                                              std::to_string(job->f1));
    if (!E->scratch) return fail(E, CTD_E_DEVICE, "no allocation-free workspace on this device");
    try {
        HIPCHK(hipSetDevice(E->device));
        hipStream_t st = stream ? (hipStream_t)stream : E->stream;
        Workspace ws;
        const uint64_t f0 = std::max<uint64_t>(job->f0, 3);
        if (job->n1 > job->n0 && f0 < job->f1)
            ws = scratch_workspace(st, (unsigned char *)E->scratch, isqrt_u64(job->f1 - 1));
        return run_job(E, job, opts, nullptr, d_counts, d_factors, last_prime, st, nullptr, &ws);
    } catch (const HipError &h) {
        return fail_hip(E, h);
    }
}

int ctd_hip_trial_factor(ctd_hip_engine *E, const ctd_job *job, const ctd_run_opts *opts, const ctd_callbacks *cb,
                         uint32_t *counts, uint64_t *factors, uint64_t *last_prime)
{
    if (!E) return CTD_E_INVALID;
    int code;
    if (!valid_job(E, job, &code)) return code;
    if (!counts || !factors) return fail(E, CTD_E_INVALID, "counts/factors are NULL");
    uint32_t *dc = nullptr;
    uint64_t *df = nullptr;
    try {
        HIPCHK(hipSetDevice(E->device));
        if (job->n1 <= job->n0) return CTD_OK;
        const uint64_t N = job->n1 - job->n0;
        const uint64_t nf = N * (uint64_t)job->max_factors;
        HIPCHK(hipMalloc((void **)&dc, N * sizeof(uint32_t)));
        HIPCHK(hipMalloc((void **)&df, nf * sizeof(uint64_t)));
        HIPCHK(hipMemcpyAsync(dc, counts, N * sizeof(uint32_t), hipMemcpyHostToDevice, E->stream));
        HIPCHK(hipMemcpyAsync(df, factors, nf * sizeof(uint64_t), hipMemcpyHostToDevice, E->stream));
        HostSync hs{counts, factors};
        int rc = run_job(E, job, opts, cb, dc, df, last_prime, E->stream, &hs);
        if (rc == CTD_OK) {
            HIPCHK(hipMemcpyAsync(counts, dc, N * sizeof(uint32_t), hipMemcpyDeviceToHost, E->stream));
            HIPCHK(hipMemcpyAsync(factors, df, nf * sizeof(uint64_t), hipMemcpyDeviceToHost, E->stream));
            HIPCHK(hipStreamSynchronize(E->stream));
        }
        HIPCHK(hipFree(dc));
        HIPCHK(hipFree(df));
        return rc;
    } catch (const HipError &h) {
        if (dc) (void)hipFree(dc);
        if (df) (void)hipFree(df);
        return fail_hip(E, h);
    }
}

int ctd_hip_residues(ctd_hip_engine *E, const ctd_job *job, const uint64_t *primes, uint64_t nprimes,
                     uint64_t *out, uint32_t flags)
{
    if (!E) return CTD_E_INVALID;
    int code;
    if (!valid_job(E, job, &code)) return code;
    if (!primes || !out || job->n1 <= job->n0) return fail(E, CTD_E_INVALID, "invalid residues request");
    uint64_t pmax = 0;
    for (uint64_t i = 0; i < nprimes; ++i) {
        if (!(primes[i] & 1) || primes[i] < 3) return fail(E, CTD_E_INVALID, "moduli must be odd and >= 3");
        pmax = std::max(pmax, primes[i]);
    }
    if (!nprimes) return CTD_OK;
    uint32_t *dc = nullptr;
    uint64_t *df = nullptr, *dr = nullptr;
    try {
        HIPCHK(hipSetDevice(E->device));
        const uint64_t N = job->n1 - job->n0;
        HIPCHK(hipMalloc((void **)&dc, N * sizeof(uint32_t)));
        HIPCHK(hipMalloc((void **)&df, N * (uint64_t)job->max_factors * sizeof(uint64_t)));
        HIPCHK(hipMalloc((void **)&dr, nprimes * sizeof(uint64_t)));
        HIPCHK(hipMemsetAsync(dc, 0, N * sizeof(uint32_t), E->stream));
        ensure_buf(E->d_primes, E->primes_cap, nprimes);
        HIPCHK(hipMemcpyAsync(E->d_primes, primes, nprimes * sizeof(uint64_t), hipMemcpyHostToDevice, E->stream));
        JobArgs J{};
        J.fam = job->seq_id;
        J.maxf = job->max_factors;
        J.param = job->param;
        J.n0 = job->n0;
        J.n1 = job->n1;
        J.chunk = std::min<uint64_t>(N, CHUNK_MAX);
        J.nchunks = (uint32_t)div_up(N, J.chunk);
        J.counts = dc;
        J.factors = df;
        J.final_res = dr;
        const bool fast = !(flags & CTD_RUN_GENERIC_ONLY) && Fast<0>::fits(pmax);
        J.flags = (J.fam == FAM_SMARANDACHE && sm_narrow_ok(J.param, J.n1, pmax)) ? JOB_SM_NARROW : 0u;
        if (pmax < FP_PMAX && !(flags & CTD_RUN_NO_FP64)) J.flags |= JOB_FP64;
        if (flags & CTD_RUN_SM_FULL_STEP) J.flags |= JOB_SM_FULL_STEP;
        if (flags & CTD_RUN_SM_NO_SHORT) J.flags |= JOB_SM_NO_SHORT;
        if (flags & CTD_RUN_SM_NO_SHIFT) J.flags |= JOB_SM_NO_SHIFT;
        if (flags & CTD_RUN_SM_NO_PAIR) J.flags |= JOB_SM_NO_PAIR;
        if (flags & CTD_RUN_STEPPED) J.flags |= JOB_STEPPED;
        const uint64_t items = (nprimes + 63) & ~63ull;
        launch_tf_any(J, E->d_primes, nprimes, nullptr, fast, dim3((uint32_t)div_up(items, 256)), dim3(256),
                      E->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, dr, nprimes * sizeof(uint64_t), hipMemcpyDeviceToHost, E->stream));
        HIPCHK(hipStreamSynchronize(E->stream));
        HIPCHK(hipFree(dc));
        HIPCHK(hipFree(df));
        HIPCHK(hipFree(dr));
    } catch (const HipError &h) {
        if (dc) (void)hipFree(dc);
        if (df) (void)hipFree(df);
        if (dr) (void)hipFree(dr);
        return fail_hip(E, h);
    }
    return CTD_OK;
}

int ctd_hip_valu_probe(ctd_hip_engine *E, int op, uint32_t iters, double *rate)
{
    if (!E || !rate || op < 0 || op >= PROBE_COUNT || iters == 0) return fail(E, CTD_E_INVALID, "invalid probe");
    try {
        HIPCHK(hipSetDevice(E->device));
        uint32_t *d = nullptr;
        HIPCHK(hipMalloc((void **)&d, sizeof(uint32_t)));
        const uint32_t blocks = (uint32_t)E->cu * 16, threads = 256;
        auto launch = [&](uint32_t it) {
            switch (op) {
#define CTD_PROBE_CASE(K)                                                                                     \
    case K: hipLaunchKernelGGL(valu_probe_kernel<K>, dim3(blocks), dim3(threads), 0, E->stream, d, it); break;
                CTD_PROBE_CASE(0) CTD_PROBE_CASE(1) CTD_PROBE_CASE(2) CTD_PROBE_CASE(3) CTD_PROBE_CASE(4)
                CTD_PROBE_CASE(5) CTD_PROBE_CASE(6) CTD_PROBE_CASE(7) CTD_PROBE_CASE(8) CTD_PROBE_CASE(9)
                CTD_PROBE_CASE(10) CTD_PROBE_CASE(11) CTD_PROBE_CASE(12) CTD_PROBE_CASE(13) CTD_PROBE_CASE(14)
                CTD_PROBE_CASE(15) CTD_PROBE_CASE(16) CTD_PROBE_CASE(17) CTD_PROBE_CASE(18) CTD_PROBE_CASE(19)
#undef CTD_PROBE_CASE
            }
        };
        launch(1); // warm-up
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(E->ev0, E->stream));
        launch(iters);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(E->ev1, E->stream));
        HIPCHK(hipEventSynchronize(E->ev1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, E->ev0, E->ev1));
        HIPCHK(hipFree(d));
        *rate = (double)blocks * threads * iters * PROBE_CHAINS * PROBE_DEPTH / (ms * 1e-3);
    } catch (const HipError &h) {
        return fail_hip(E, h);
    }
    return CTD_OK;
}

} // extern "C"
