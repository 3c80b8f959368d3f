// Host engine and C ABI of libwar_engine.so (declared in include/war_engine.h).
//
// One engine per process, bound to one GPU. GPU work runs on the engine's stream, guarded by one
// mutex; the serving loop is a host thread that takes the same mutex per step and alternates its
// steps between that stream and a second one, keeping one step in flight while it plans the next.
//
// A persistent batch is one hipGraph: k_setup, the FRESH stage, the queue stages and (if any
// request of the batch needs them) the cycle walks. Graphs are captured once per launch shape and
// replayed; the job list they read lives in device memory and is copied in before each replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <type_traits>
#include <vector>

#include "war_engine.h"
#include "war_kernels.h"
#include "war_tools.h"

using namespace war;

namespace {

// The reason of the last failed call on this thread (empty after a call that succeeded). A serving
// request's failure travels in its ReqHost and is raised again on the thread that waits for it.
thread_local std::string t_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_err = buf;
    return code;
}

#define HIPCHK(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) return fail(100 + (int)e_, "%s: %s (%s:%d)", #call,          \
                                          hipGetErrorString(e_), __FILE__, __LINE__);        \
    } while (0)

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

u64 fact(int k) { u64 f = 1; for (int i = 2; i <= k; i++) f *= (u64)i; return f; }
u64 range_total(int n, int mode) {
    if (n < 1) return 0;
    if (mode == WAR_MODE_CANON) return n < 2 ? 0 : fact(n - 1) * (u64)(n - 1);
    return fact(n);
}

constexpr int MAXSLOTS = 256;
constexpr int MAXJOBS = 1024;
constexpr size_t JOBS_OFF = 256;   // byte offset of the Job array after the JobHdr

// Per request: device counters and histograms (resident until the result is copied out).
struct ReqHost {
    war_request rq{};
    war_result *res = nullptr;
    int slot = -1;
    u64 *dmem = nullptr;            // one allocation: counters | hist | lenhist | dhist | htab
    ReqDev dev{};
    u64 next = 0;                   // next position to schedule
    uint64_t id = 0;
    double t_submit = 0, t_done = 0, dev_ms = 0;
    bool finished = false;
    int status = 0;
    std::string err;
    size_t bytes = 0;
    int layout = 4;                 // pile layout the request runs in (4, 5 or 6)
    u64 share = 0;                  // serving: positions given to the request in the current step
    int inflight = 0;               // serving: launched steps with ranges of this request, not yet waited for
    bool pooled = false;            // serving: dmem came from (and returns to) the engine's pool
    bool reach = false;             // war_reach: search each distinct cycle's predecessors for a deal
    u64 reach_cap = 0;
    war_reach_result *rres = nullptr;
};

// Device buffers of one pile layout: two region buffers (ping-pong between stages) of `cap` slots
// with their per-tile counts, and the cycle-record regions with their per-tile counts.
struct QueueBufs {
    void *qa[2] = {nullptr, nullptr}, *qb[2] = {nullptr, nullptr};
    u32 *qreq[2] = {nullptr, nullptr};
    u64 *qpos[2] = {nullptr, nullptr};
    u32 *qhome[2] = {nullptr, nullptr};
    u32 *qcnt[2] = {nullptr, nullptr};
    void *ca = nullptr, *cb = nullptr;
    u32 *cmeta = nullptr, *creq = nullptr, *clam = nullptr, *ccnt = nullptr;
    u64 *cpos = nullptr;
    void *qta[2] = {nullptr, nullptr}, *qtb[2] = {nullptr, nullptr};   // alternation detector: tortoise
    u32 *qmeta[2] = {nullptr, nullptr};
    u64 cap = 0, tmax = 0;
    bool have_c = false, have_t = false;
};

constexpr u64 MIN_TILE = 64, MAX_TILE = 8192;
constexpr int MAXB = 32;                 // launches one batch graph holds (one branch each at most)
constexpr size_t SLOT_BYTES = JOBS_OFF + MAXJOBS * 32;   // one launch's job header + jobs
constexpr u64 FLOW_BATCH = (1ull << 32) - (1ull << 20);  // flow kernel: positions per batch (below 2^32: the
                                                        // kernel's per-wave counts are 32-bit)
constexpr u32 PLAIN_CYC_TILE = 1024;    // home-tile size of the plain kernel's cycle records

// ---- the entry point that allocates nothing (war_run_range_noalloc): its device memory is this
// file-scope array, one copy per GPU, held until the process ends. Host code never passes the
// array's name to a kernel or a copy (on the host it is a stand-in, not a device address): war_init
// resolves the device address once per device with hipGetSymbolAddress and keeps it (Engine::na).
// Layout, in u64 words: the pipe counters, the job header and its one job, the request descriptor,
// then the request's counters and its two histograms (at most NA_BINS bins each).
constexpr u32 NA_BINS = 4096;
constexpr size_t NA_PIPE = 0, NA_HDR = 64, NA_JOB = NA_HDR + JOBS_OFF / 8, NA_REQ = 160, NA_CNT = 192,
                 NA_HIST = NA_CNT + C_N, NA_LEN = NA_HIST + NA_BINS, NA_WORDS = NA_LEN + NA_BINS;
static_assert(sizeof(Pipe) <= 8 * NA_HDR && sizeof(JobHdr) <= JOBS_OFF && NA_JOB + 4 <= NA_REQ &&
              sizeof(ReqDev) <= 8 * (NA_CNT - NA_REQ), "scratch layout");
static __device__ __align__(256) u64 scratch[NA_WORDS];

// Fills n words with v (a kernel, not hipMemsetAsync, so a caller may capture the calls).
__global__ void k_fill(u64 *p, u64 v, u32 n) {
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

struct GraphKey {
    int W, rule, kernel, wpc, per, cap0, walk, multi, hb, lb, kind, det;   // det: WAR_DET_* of the persistent path
    int blk;                                                                // flow kernel over whole blocks of 24 deals
    int fw = 0;                                                             // flow kernel, Brent, rule 0: first window (fw_of)
    bool operator<(const GraphKey &o) const {
        return std::tie(W, rule, kernel, wpc, per, cap0, walk, multi, hb, lb, kind, det, blk, fw) <
               std::tie(o.W, o.rule, o.kernel, o.wpc, o.per, o.cap0, o.walk, o.multi, o.hb, o.lb, o.kind, o.det, o.blk, o.fw);
    }
};

// One step context: MAXB slots of pipe counters and job lists on the device, their pinned host
// staging, the pipe counters read back, a stream, timing events and the batch graphs captured on
// these buffers (keyed by the shape of their branches). While a step is in flight, `stepped` holds
// the requests with ranges in it.
struct StepCtx {
    Pipe *pipe = nullptr;
    char *hdr = nullptr;
    void *stage = nullptr;
    Pipe *pipe_host = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::map<std::vector<int>, hipGraphExec_t> graphs;
    size_t nl = 0;              // launches of the step in flight (0: none)
    bool walk = false;          // ... and whether any of them has cycle walks
    std::vector<std::shared_ptr<ReqHost>> stepped;
};

struct Engine {
    int device = 0;
    bool verbose = false;
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int cus = 0, clock_mhz = 0;
    std::string arch;
    war_tuning tun{0, 0, 0, 0, 0};            // engine-wide override (war_set_tuning); zeros = per-workload default
    std::map<std::tuple<int, int, int, int>, war_tuning> tuned;   // (n, mode, rule, detector) -> war_autotune's choice
    u64 batch = 1ull << 28;          // positions per persistent batch (= queue capacity)
    Pipe *pipe = nullptr;
    JobHdr *hdr = nullptr;           // device: JobHdr followed by MAXJOBS Jobs
    Job *jobs = nullptr;
    void *stage_host = nullptr;      // pinned staging for the job list
    ReqDev *reqs = nullptr;          // device table of MAXSLOTS
    QueueBufs qb4, qb5;
    u64 *reach_stack = nullptr;     // reach searches: REACH_DEPTH entries per thread of a walk grid
    std::map<GraphKey, hipGraphExec_t> graphs;
    // batch graphs: two step contexts (the serving loop launches a step on one while the step on
    // the other runs; everything else uses context 0, whose stream is `st`), and the streams and
    // events that capture a graph's branches (used only while capturing)
    StepCtx sc[2];
    hipStream_t bst[MAXB] = {};
    hipEvent_t bfork = nullptr, bjoin[MAXB] = {};
    // the serving loop's request memory, kept for reuse by size class (hipFree waits for the whole
    // device, which would stall the step in flight); released when the loop goes idle
    std::multimap<size_t, void *> pool;
    // war_run_range_noalloc: the device address of `scratch` on this device, and pinned staging
    // (allocated at init) for its job header, request descriptor and results
    u64 *na = nullptr;
    u64 *na_host = nullptr;
    std::vector<int> free_slots;
    uint32_t slot_cap[MAXSLOTS] = {};  // each slot's explicit step cap (ReqDev::cap), for the kernel variant
    uint32_t slot_n[MAXSLOTS] = {};    // each slot's n and mode (the flow kernel's first-window variant)
    uint32_t slot_mode[MAXSLOTS] = {};
    std::mutex mu;                   // serialises all GPU work of the engine

    // serving
    std::thread worker;
    std::mutex smu;
    std::condition_variable scv;
    std::deque<std::shared_ptr<ReqHost>> pending;
    std::map<uint64_t, std::shared_ptr<ReqHost>> all;
    std::atomic<bool> serving{false};
    int serve_slots = 64;
    u64 step_positions = 1ull << 26;
    uint64_t next_id = 1;
};

Engine *g_eng = nullptr;

int check_request(const war_request *r) {
    if (!r) return fail(1, "null request");
    if (r->n < 1 || r->n > 20) return fail(2, "n must be in [1, 20] (got %d)", r->n);
    if (r->mode < 0 || r->mode > 2) return fail(3, "mode must be FULL, HALF or CANON");
    if (r->rule != 0 && r->rule != 1) return fail(4, "rule must be 0 or 1");
    if (r->kernel != WAR_KERNEL_PLAIN && r->kernel != WAR_KERNEL_PERSISTENT) return fail(5, "unknown kernel");
    const u64 total = range_total(r->n, r->mode);
    if (r->lo > r->hi || r->hi > total) return fail(6, "range [%llu, %llu) outside [0, %llu)", r->lo, r->hi, total);
    if (r->mode == WAR_MODE_HALF && (r->n & 1)) return fail(7, "half mode needs even n");
    if (r->hbins < 0 || r->hbins > 4096 || r->lbins < 0 || r->lbins > 4096 || r->dbins < 0 || r->dbins > 4096)
        return fail(8, "bins must be in [0, 4096]");
    if (r->hbins + r->lbins > 8192) return fail(8, "hbins + lbins must be <= 8192");
    if (r->distinct && (r->mode != WAR_MODE_CANON || r->n > 15))
        return fail(9, "distinct needs canon mode and n <= 15 (its keys are 4-bit cards under a 4-bit length)");
    if (r->dbins && (r->mode != WAR_MODE_CANON || r->n > 16))
        return fail(9, "dbins needs canon mode and n <= 16 (4-bit piles)");
    if (r->step_cap > WAR_SAFETY_CAP) return fail(13, "step_cap must be <= 2^24 (got %u)", r->step_cap);
    if (r->flags & ~(WAR_FLAG_UNDET_ERROR | WAR_FLAG_CYCLE_TABLE | WAR_FLAG_NO_CYCLE_TABLE))
        return fail(13, "unknown flags 0x%x", r->flags);
    if (r->shape_blk && r->mode != WAR_MODE_CANON) return fail(10, "shape needs canon mode");
    if (r->shape_blk < 0 || r->shape_blk > 512) return fail(10, "shape_blk must be in [0, 512]");
    if (r->detector != WAR_DET_BRENT && r->detector != WAR_DET_STREAK && r->detector != WAR_DET_ALT &&
        r->detector != WAR_DET_PAIR)
        return fail(12, "unknown detector");
    if (r->detector == WAR_DET_PAIR) {
        if (r->n % 4 != 2 || r->rule != 0)
            return fail(12, "the pair-run detector needs n = 2 mod 4 and rule 0 (runs of two: the loop-shape test)");
        if (r->kernel != WAR_KERNEL_PERSISTENT) return fail(12, "the pair-run detector runs in the persistent kernel");
    }
    if (r->detector == WAR_DET_ALT && ((r->n & 1) == 0 || r->rule != 0))
        return fail(12, "the alternation detector needs odd n and rule 0 (odd-n cycles alternate: Spivey Thm 6)");
    if (r->detector == WAR_DET_STREAK) {
        if ((r->n & 1) == 0) return fail(12, "the streak test is valid for odd n only (Spivey Thm 6)");
        if (r->rule != 0) return fail(12, "the streak test needs rule 0");
        if (r->kernel != WAR_KERNEL_PLAIN) return fail(12, "the streak test runs in the plain kernel");
        if (r->hbins || r->dbins || r->distinct || r->shape_blk)
            return fail(12, "the streak test gives no cycle lengths: hbins, dbins, distinct and shape must be 0");
    }
    return 0;
}

// Layout 6 applies to n = 17 persistent Brent requests without cycle walks (its records do not
// keep the players' deal order).
bool held_ok(const war_request &r) {
    return r.n == (int)HELD_N && r.kernel == WAR_KERNEL_PERSISTENT &&
           (r.detector == WAR_DET_BRENT || r.detector == WAR_DET_ALT) && !r.dbins && !r.distinct && !r.shape_blk;
}

// Defaults measured on MI355X: deals (long games, ~110 turns at n = 13) want a late first stage
// boundary; canonical positions (most end within ~30 turns) an early one and larger tiles.
war_tuning default_tuning(int n, int mode, int det) {
    war_tuning t{28, 1024, 15, 32, 3, 0};     // stage_kind 3: the flow kernel where it applies
    if (mode != WAR_MODE_CANON && n >= 12) t.brent_cap = 63;
    if (det == WAR_DET_ALT) t.brent_cap = mode == WAR_MODE_CANON ? 31 : 63;   // stage windows of 32 / 64 turns (when staged)
    if (mode == WAR_MODE_CANON) t.ranks_per_refill = 2048;
    return t;
}

// A single canonical range of at least BIG_CANON_RANGE positions, with no tile set by the request, the
// engine or the autotuner: lane runs of six whole orders (6 (n - 1) positions from split 0) instead of
// the default's two, so fewer tile grabs and first-order unranks per position; smaller ranges, the
// serving loop's steps and batches keep two, so each wave still gets many tiles. Measured against
// two: n = 14 whole 1453 -> 1423 ms, n = 16 (2^35 positions) 599 -> 585 ms, n = 13 239 -> 237 ms,
// n = 15 54.6 -> 54.3 s; five and seven orders were slower than six at n = 14; on ranges of 5.5e7
// positions (n = 14), six orders were 19 % slower than two, and on the 2^28-position batches of
// cycle walks (n = 14 with distinct + shape) 5 % slower.
constexpr u64 BIG_CANON_RANGE = 1ull << 31;
bool walks_of(const war_request &r);
void big_canon_tile(const Engine &E, const war_request &r, war_tuning &t) {
    if (r.mode != WAR_MODE_CANON || r.n < 3 || r.hi - r.lo < BIG_CANON_RANGE) return;
    if (walks_of(r)) return;         // walk batches hold 2^28 positions: two orders keep their tiles many
    if (r.tuning.ranks_per_refill > 0 || E.tun.ranks_per_refill > 0) return;
    if (E.tuned.count(std::make_tuple(r.n, r.mode, r.rule, r.detector))) return;
    t.ranks_per_refill = (int)std::min<u64>(64ull * 6u * (u64)(r.n - 1), MAX_TILE);
}

// Precedence: the request's own fields, then the engine-wide override, then war_autotune's choice
// for this (n, mode, rule), then the default.
void resolve_tuning(const Engine &E, const war_request &r, war_tuning &t) {
    t = default_tuning(r.n, r.mode, r.detector);
    auto it = E.tuned.find(std::make_tuple(r.n, r.mode, r.rule, r.detector));
    if (it != E.tuned.end()) {
        t = it->second;
        if (t.pile_layout == 6 && !held_ok(r)) t.pile_layout = 0;   // tuned on requests without walks
    }
    for (const war_tuning *src : {&E.tun, &r.tuning}) {
        if (src->waves_per_cu > 0) t.waves_per_cu = src->waves_per_cu;
        if (src->ranks_per_refill > 0) t.ranks_per_refill = src->ranks_per_refill;
        if (src->brent_cap > 0) t.brent_cap = src->brent_cap;
        if (src->plain_per > 0) t.plain_per = src->plain_per;
        if (src->stage_kind > 0) t.stage_kind = src->stage_kind;
        if (src->pile_layout > 0) t.pile_layout = src->pile_layout;
    }
    // the pair-run detector exists on the flow kernel only
    if (r.detector == WAR_DET_PAIR && r.tuning.stage_kind == 0 && E.tun.stage_kind == 0) t.stage_kind = 3;
}

// An explicit step cap needs a kernel that tests it exactly: the plain kernel (any cap), or the flow
// kernel (a multiple of its spans of 16 and 32 turns, tested at every span end).
int check_cap(const war_request &r, const war_tuning &t) {
    if (!r.step_cap || r.kernel == WAR_KERNEL_PLAIN) return 0;
    if (r.step_cap % 32u) return fail(13, "step_cap on the persistent kernel must be a multiple of 32 (got %u)", r.step_cap);
    if (t.stage_kind != 3) return fail(13, "step_cap on the persistent kernel needs the flow schedule (stage_kind 3)");
    return 0;
}

int check_tuning(const war_tuning &t) {
    if (t.waves_per_cu < 1 || t.waves_per_cu > 32) return fail(11, "waves_per_cu must be in [1, 32]");
    if (t.ranks_per_refill < 64 || t.ranks_per_refill > 8192 || (t.ranks_per_refill & 63))
        return fail(11, "ranks_per_refill must be a multiple of 64 in [64, 8192]");
    const int c = t.brent_cap;
    if (c < 1 || c > 65535 || ((c + 1) & c)) return fail(11, "brent_cap must be 2^k - 1 in [1, 65535]");
    if (t.plain_per < 1 || t.plain_per > 4096) return fail(11, "plain_per must be in [1, 4096]");
    if (t.stage_kind < 1 || t.stage_kind > 3) return fail(11, "stage_kind must be 1 (refill), 2 (lockstep) or 3 (flow)");
    if (t.pile_layout != 0 && t.pile_layout != 4 && t.pile_layout != 5 && t.pile_layout != 6)
        return fail(11, "pile_layout must be 0, 4, 5 or 6");
    return 0;
}

// The pile layout a request runs in. Layout 6 (n = 17, card 16 held apart, 4-bit piles) plays a
// turn in about half the instructions of the 5-bit layout; it runs in the persistent pipeline
// with Brent and without cycle walks (its records do not keep the players' deal order).
int layout_of(const war_request &r, const war_tuning &t, int *out) {
    int L = t.pile_layout;
    if (L == 0) L = held_ok(r) ? 6 : (r.n <= 16 ? 4 : 5);
    if (L == 4 && r.n > 16) return fail(11, "pile_layout 4 needs n <= 16");
    if (L == 6 && !held_ok(r))
        return fail(11, "pile_layout 6 needs n = 17, the persistent kernel, Brent and no dbins / distinct / shape");
    if (L == 5 && (r.distinct || r.dbins)) return fail(11, "distinct / dbins need 4-bit piles (n <= 15)");
    if (r.detector == WAR_DET_ALT && (u32)t.brent_cap + 1u < ALT_K)
        return fail(11, "the alternation detector needs brent_cap >= %u (stage windows of brent_cap + 1 turns, a "
                        "multiple of its %u-turn checkpoints)", ALT_K - 1u, ALT_K);
    *out = L;
    return 0;
}

// ---- per-request device memory
// Pool size class of a request's memory: a power of two of at least 64 KiB.
size_t pool_class(size_t bytes) {
    size_t c = (size_t)1 << 16;
    while (c < bytes) c <<= 1;
    return c;
}

void pool_release(Engine &E) {
    for (auto &kv : E.pool) (void)hipFree(kv.second);
    E.pool.clear();
}

u32 first_window(int n);

// Cycle walks (distinct / dbins / shape on canonical positions), except where no position cycles
// (rule 0, n = 12 and 16: first_window 0): the walk variants and their smaller batches would only add
// time, every walk output being zero there.
bool walks_of(const war_request &r) {
    const bool nocyc = r.rule == 0 && r.n >= 5 && r.n <= 16 && first_window(r.n) == 0u;
    return r.mode == WAR_MODE_CANON && (r.distinct || r.dbins || r.shape_blk) && !nocyc;
}

// The cycle-state table of a canonical rule-0 request with cycle walks, as log2 of its slots (0: none):
// at least twice the states on the cycles of n, from the reference's published distinct cycles and
// their lengths (data/checks.txt, data/long_runs.txt: n = 5: 2 of 6; 7: 52 of 8; 9: 144 of 20 and 192
// of 30; 10: 48 of 60; 11: 42560 of 12 and 13032 of 24; 13: 86400 of 42, 129600 of 56 and 103680 of
// 70; 14: 720 of 96 and 25920 of 112; none for n = 6, 8, 12, 16), each cycle in both player orders.
// n = 15 (48 million cycles) keeps the walks. Ranges of at least 2^28 positions by default
// (WAR_FLAG_CYCLE_TABLE / WAR_FLAG_NO_CYCLE_TABLE force it either way).
constexpr u64 STAB_MIN_RANGE = 1ull << 28;
u32 stab_log2(const war_request &r) {
    const bool walk = r.mode == WAR_MODE_CANON && (r.distinct || r.dbins || r.shape_blk);
    if (!walk || r.rule != 0 || (r.flags & WAR_FLAG_NO_CYCLE_TABLE)) return 0;
    if (!(r.flags & WAR_FLAG_CYCLE_TABLE) && r.hi - r.lo < STAB_MIN_RANGE) return 0;
    switch (r.n) {
        case 5: return 8;      // 24 states
        case 7: return 12;     // 832
        case 9: return 16;     // 17 280
        case 10: return 14;    // 5 760
        case 11: return 22;    // 1 648 512
        case 13: return 27;    // 36 288 000
        case 14: return 24;    // 5 944 320
        default: return 0;
    }
}

// The request's device memory (counters | hist | lenhist | dhist | htab | nolen | stab), zeroed, and its descriptor
// in the device table, both on stream st. pooled: memory from the serving loop's pool (a block of the
// request's size class when one is free), returned there by req_free.
int req_alloc(Engine &E, ReqHost &h, hipStream_t st = nullptr, bool pooled = false) {
    if (!st) st = E.st;
    const war_request &r = h.rq;
    const size_t nc = C_N, nh = (size_t)r.hbins, nl = (size_t)r.lbins, nd = (size_t)r.dbins;
    u32 hlog2 = 0;
    size_t nt = 0;
    if (r.distinct) {
        // table sized for the cycles a range can touch: at most the on-cycle positions of n
        hlog2 = r.n <= 13 ? 21 : (r.n == 14 ? 17 : 27);
        nt = (size_t)1 << hlog2;
    }
    const size_t nr = h.reach ? 4096 : 0;
    const u32 slog2 = stab_log2(r);
    const size_t ns = slog2 ? (size_t)1 << slog2 : 0;
    h.bytes = (nc + nh + nl + nd + nt + nr + ns) * sizeof(u64);
    void *p = nullptr;
    h.pooled = pooled;
    if (pooled) {
        const size_t c = pool_class(h.bytes);
        auto it = E.pool.find(c);
        if (it != E.pool.end()) { p = it->second; E.pool.erase(it); }
        else HIPCHK(hipMalloc(&p, c));
    } else {
        HIPCHK(hipMalloc(&p, h.bytes));
    }
    h.dmem = (u64 *)p;
    HIPCHK(hipMemsetAsync(p, 0, h.bytes, st));
    ReqDev &d = h.dev;
    d = ReqDev{};
    d.n = (u32)r.n; d.mode = (u32)r.mode; d.rule = (u32)r.rule;
    d.hbins = (u32)r.hbins; d.lbins = (u32)r.lbins; d.dbins = (u32)r.dbins;
    d.distinct = (u32)r.distinct; d.shape_blk = (u32)r.shape_blk;
    d.walk = walks_of(r) ? 1u : 0u;
    d.hlog2 = hlog2;
    d.cnt = h.dmem;
    d.hist = d.cnt + nc;
    d.lenhist = d.hist + nh;
    d.dhist = d.lenhist + nl;
    d.htab = nt ? d.dhist + nd : nullptr;
    d.reach = h.reach ? 1u : 0u;
    d.reach_cap = h.reach_cap;
    d.nolen = nr ? d.dhist + nd + nt : nullptr;
    d.stab = ns ? d.dhist + nd + nt + nr : nullptr;
    d.slog2 = slog2;
    d.cap = r.step_cap;
    E.slot_cap[h.slot] = r.step_cap;
    E.slot_n[h.slot] = (u32)r.n;
    E.slot_mode[h.slot] = (u32)r.mode;
    HIPCHK(hipMemcpyAsync(&E.reqs[h.slot], &d, sizeof(ReqDev), hipMemcpyHostToDevice, st));
    return 0;
}

// Copy a request's counters and histograms into buf on stream st (the caller synchronises).
int req_fetch(ReqHost &h, hipStream_t st, std::vector<u64> &buf) {
    const war_request &r = h.rq;
    buf.assign(C_N + (size_t)r.hbins + (size_t)r.lbins + (size_t)r.dbins, 0);
    HIPCHK(hipMemcpyAsync(buf.data(), h.dmem, buf.size() * sizeof(u64), hipMemcpyDeviceToHost, st));
    return 0;
}

// The result of a request from its fetched counters, with the checks of a finished request.
int req_finish(ReqHost &h, const std::vector<u64> &buf) {
    const war_request &r = h.rq;
    war_result *o = h.res;
    if (o) {
        o->cyc = buf[C_CYC]; o->done = buf[C_DONE]; o->maxfin = buf[C_MAXFIN]; o->turns = buf[C_TURNS];
        o->oncycle = buf[C_ONCYCLE]; o->distinct_minkey = buf[C_MINKEY]; o->distinct_hashset = buf[C_HASH];
        o->shape_nonstd = buf[C_NONSTD]; o->shape_unchk = buf[C_UNCHK]; o->undetermined = buf[C_UNDET];
        if (h.rres) {
            war_reach_result *q = h.rres;
            q->classes = buf[C_MINKEY]; q->reached = buf[C_RYES]; q->not_reached = buf[C_RNO];
            q->undetermined = buf[C_RUND]; q->device_ms = h.dev_ms; q->max_depth = buf[C_RDEPTH];
            if (q->nolen && q->nbins > 0) {
                std::vector<u64> nl(4096);
                HIPCHK(hipMemcpy(nl.data(), h.dev.nolen, 4096 * sizeof(u64), hipMemcpyDeviceToHost));
                for (int i = 0; i < q->nbins; i++) q->nolen[i] = 0;
                for (int i = 0; i < 4096; i++) q->nolen[std::min(i, q->nbins - 1)] += nl[(size_t)i];
            }
        }
        const u64 *hp = buf.data() + C_N, *lp = hp + r.hbins, *dp = lp + r.lbins;
        if (o->hist) for (int i = 0; i < std::min(o->hbins, r.hbins); i++) o->hist[i] = hp[i];
        if (o->lenhist) for (int i = 0; i < std::min(o->lbins, r.lbins); i++) o->lenhist[i] = lp[i];
        if (o->dhist) for (int i = 0; i < std::min(o->dbins, r.dbins); i++) o->dhist[i] = dp[i];
        o->device_ms = h.dev_ms;
    }
    if (buf[C_HOVF]) return fail(20, "distinct hash set overflowed (%llu keys dropped)", buf[C_HOVF]);
    // undetermined games are a count; an error where the caller asks for it, and for the default cap
    // of the exact detectors (a game that runs 2^24 turns means a corrupted state)
    if (buf[C_UNDET] && ((r.flags & WAR_FLAG_UNDET_ERROR) || (!r.step_cap && r.detector != WAR_DET_STREAK)))
        return fail(21, "%llu games not resolved within %u turns", buf[C_UNDET],
                    r.step_cap ? r.step_cap : (r.detector == WAR_DET_STREAK ? WAR_STREAK_CAP : WAR_SAFETY_CAP));
    const u64 want = r.hi - r.lo;
    if (r.mode != WAR_MODE_HALF && buf[C_DONE] != want)
        return fail(22, "played %llu positions, expected %llu", buf[C_DONE], want);
    return 0;
}

int req_collect(Engine &E, ReqHost &h) {
    std::vector<u64> buf;
    int rc = req_fetch(h, E.st, buf);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(E.st));
    return req_finish(h, buf);
}

void req_free(Engine &E, ReqHost &h) {
    if (h.dmem) {
        if (h.pooled) E.pool.insert({pool_class(h.bytes), (void *)h.dmem});
        else (void)hipFree(h.dmem);
        h.dmem = nullptr;
    }
    if (h.slot >= 0) { E.slot_cap[h.slot] = 0; E.free_slots.push_back(h.slot); h.slot = -1; }
}

// ---- queue buffers of one layout, allocated on first use
// layout 6 records are 16-byte u64 pairs like layout 4's, so the two share buffers
template <int W> QueueBufs &qbufs(Engine &E) { return W == 5 ? E.qb5 : E.qb4; }

template <int W> int ensure_queues(Engine &E, bool walk, bool alt = false) {
    typedef typename Lay<W>::T T;
    QueueBufs &Q = qbufs<W>(E);
    if (!Q.cap) {
        // T * tile_pos <= batch + jobs * tile_pos; a refill stage leaves at most one partly filled
        // region per wave (<= 32 waves/CU x CUs) beyond its records
        const u64 cap = E.batch + std::max<u64>((u64)MAXJOBS, (u64)E.cus * 32) * MAX_TILE;
        const u64 tmax = cap / MIN_TILE + MAXJOBS;
        for (int i = 0; i < 2; i++) {
            HIPCHK(hipMalloc(&Q.qa[i], cap * sizeof(T)));
            HIPCHK(hipMalloc(&Q.qb[i], cap * sizeof(T)));
            HIPCHK(hipMalloc((void **)&Q.qreq[i], cap * sizeof(u32)));
            HIPCHK(hipMalloc((void **)&Q.qpos[i], cap * sizeof(u64)));
            HIPCHK(hipMalloc((void **)&Q.qhome[i], cap * sizeof(u32)));
            HIPCHK(hipMalloc((void **)&Q.qcnt[i], tmax * sizeof(u32)));
        }
        Q.cap = cap;
        Q.tmax = tmax;
    }
    if (alt && !Q.have_t) {
        for (int i = 0; i < 2; i++) {
            HIPCHK(hipMalloc(&Q.qta[i], Q.cap * sizeof(T)));
            HIPCHK(hipMalloc(&Q.qtb[i], Q.cap * sizeof(T)));
            HIPCHK(hipMalloc((void **)&Q.qmeta[i], Q.cap * sizeof(u32)));
        }
        Q.have_t = true;
    }
    if (walk && !E.reach_stack) {
        // the walk grids have at most CUs x 8 blocks of 256 threads
        HIPCHK(hipMalloc((void **)&E.reach_stack, (size_t)E.cus * 8 * 256 * REACH_DEPTH * 2 * sizeof(u64)));
    }
    if (walk && !Q.have_c) {
        HIPCHK(hipMalloc(&Q.ca, Q.cap * sizeof(T)));
        HIPCHK(hipMalloc(&Q.cb, Q.cap * sizeof(T)));
        HIPCHK(hipMalloc((void **)&Q.cmeta, Q.cap * sizeof(u32)));
        HIPCHK(hipMalloc((void **)&Q.creq, Q.cap * sizeof(u32)));
        HIPCHK(hipMalloc((void **)&Q.clam, Q.cap * sizeof(u32)));
        HIPCHK(hipMalloc((void **)&Q.cpos, Q.cap * sizeof(u64)));
        HIPCHK(hipMalloc((void **)&Q.ccnt, Q.tmax * sizeof(u32)));
        Q.have_c = true;
    }
    return 0;
}

template <int W> CycQueue<W> cyc_queue(Engine &E) {
    typedef typename Lay<W>::T T;
    QueueBufs &Q = qbufs<W>(E);
    CycQueue<W> c;
    c.a = (T *)Q.ca; c.b = (T *)Q.cb; c.meta = Q.cmeta; c.req = Q.creq; c.lam = Q.clam; c.pos = Q.cpos;
    c.cnt = Q.ccnt;
    return c;
}

int round_pow2(int x) { int p = 0; if (x <= 0) return 0; p = 1; while (p < x) p <<= 1; return p; }

// Brent's first useful window under rule 0, as its power: a window of power p sees only cycles of
// length <= p, so the windows below the shortest cycle of n can be left out without changing any
// detection (or `turns`). The shortest cycles follow the reference's published cycle-length
// histograms over every deal and every canonical position (data/checks.txt, data/long_runs.txt:
// n = 9: 20, 30; 10: 60; 11: 12, 24; 13: 42, 56, 70; 14: 96, 112; no cycle at all for n = 12 and 16,
// where 0 means no window is needed); from n = 17 on a cycle is at least ceil(n/2) >= 9 long (each
// player plays every card it holds over a cycle). n <= 8 and 15 keep the window at age 7.
u32 first_window(int n) {
    switch (n) {
        case 9: return 32;
        case 10: return 64;
        case 11: return 16;
        case 12: return 0;
        case 13: return 64;
        case 14: return 128;
        case 16: return 0;
        default: return n >= 17 ? 16 : 8;
    }
}

// The flow kernel's first-window variant (GraphKey::fw, templates FWM and OFF) for a single request
// of n (4-bit piles, Brent, rule 0): 1 = the first window at age 127 (n = 14), 2 = none (no cycles),
// 0 = every window from age 7 (a window earlier than n needs is Brent's own, so still exact); 3 and
// 4 = 1 and 2 for canonical positions dealt after their first turn (without an explicit step cap).
int fw_of(int n, int mode, bool capped) {
    const u32 w = first_window(n);
    const int fw = w >= 128 ? 1 : (w == 0 ? 2 : 0);
    return (fw && mode == WAR_MODE_CANON && !capped) ? fw + 2 : fw;
}

// Stage boundaries: cap0, 2 cap0 + 1, ... while below 1024, then the unbounded last stage.
std::vector<u32> stage_caps(int cap0) {
    std::vector<u32> c;
    u32 a = (u32)cap0;
    while (a < 1024u && (int)c.size() < MAXSTAGE - 1) { c.push_back(a); a = 2 * a + 1; }
    c.push_back(UNBOUNDED);
    return c;
}

// Stage boundaries of the alternation detector: multiples of ALT_K, so every lane meets the same
// checkpoints: w = cap0 + 1, 2w, ..., 8w, then doubling while below 1024, then unbounded.
std::vector<u32> stage_caps_alt(int cap0) {
    std::vector<u32> c;
    const u32 w = (u32)cap0 + 1u;
    for (u32 i = 1; i <= 8 && (int)c.size() < MAXSTAGE - 1; i++) c.push_back(i * w);
    u32 a = c.back();
    while (a < 1024u && (int)c.size() < MAXSTAGE - 1) { a *= 2; c.push_back(a); }
    c.push_back(UNBOUNDED);
    return c;
}

template <int W, int RULE, bool FRESH, int AF>
void launch_stage_a(int flags, int blocks, size_t lds, hipStream_t st, const StageArgs<W> &A) {
    if constexpr (W == 6) {            // no cycle walks in layout 6
        if (flags & F_MULTI) hipLaunchKernelGGL((k_stage<W, RULE, FRESH, F_MULTI | AF>), dim3(blocks), dim3(256), lds, st, A);
        else hipLaunchKernelGGL((k_stage<W, RULE, FRESH, AF>), dim3(blocks), dim3(256), lds, st, A);
        return;
    }
    switch (flags & (F_MULTI | F_WALK)) {
        case 0: hipLaunchKernelGGL((k_stage<W, RULE, FRESH, AF>), dim3(blocks), dim3(256), lds, st, A); break;
        case F_MULTI: hipLaunchKernelGGL((k_stage<W, RULE, FRESH, F_MULTI | AF>), dim3(blocks), dim3(256), lds, st, A); break;
        case F_WALK: hipLaunchKernelGGL((k_stage<W, RULE, FRESH, F_WALK | AF>), dim3(blocks), dim3(256), lds, st, A); break;
        default: hipLaunchKernelGGL((k_stage<W, RULE, FRESH, F_MULTI | F_WALK | AF>), dim3(blocks), dim3(256), lds, st, A); break;
    }
}

template <int W, int RULE, bool FRESH>
void launch_stage_f(int flags, int blocks, size_t lds, hipStream_t st, const StageArgs<W> &A) {
    // the alternation detector exists for rule 0 only
    if (RULE == 0 && (flags & F_ALT)) launch_stage_a<W, RULE, FRESH, (RULE == 0 ? F_ALT : 0)>(flags, blocks, lds, st, A);
    else launch_stage_a<W, RULE, FRESH, 0>(flags, blocks, lds, st, A);
}

template <int W, int RULE>
void launch_stage(bool fresh, int flags, int blocks, size_t lds, hipStream_t st, const StageArgs<W> &A) {
    if (fresh) launch_stage_f<W, RULE, true>(flags, blocks, lds, st, A);
    else       launch_stage_f<W, RULE, false>(flags, blocks, lds, st, A);
}

template <int W, int RULE>
void launch_refill(int flags, int blocks, size_t lds, hipStream_t st, const StageArgs<W> &A) {
    if constexpr (W == 6) {            // layout 6 runs lockstep stages only
        launch_stage<W, RULE>(false, flags, blocks, lds, st, A);
        return;
    }
    switch (flags) {
        case 0: hipLaunchKernelGGL((k_refill<W, RULE, 0>), dim3(blocks), dim3(256), lds, st, A); break;
        case F_MULTI: hipLaunchKernelGGL((k_refill<W, RULE, F_MULTI>), dim3(blocks), dim3(256), lds, st, A); break;
        case F_WALK: hipLaunchKernelGGL((k_refill<W, RULE, F_WALK>), dim3(blocks), dim3(256), lds, st, A); break;
        default: hipLaunchKernelGGL((k_refill<W, RULE, F_MULTI | F_WALK>), dim3(blocks), dim3(256), lds, st, A); break;
    }
}

// The flow kernel for a batch's detector (WAR_DET_*): Brent for either rule, the alternation and
// pair-run detectors for rule 0 (pair-run: single requests, layouts 4 and 5). BLK: the variant for
// 4-bit deals whose jobs and lane runs are whole blocks of 24 deals; CAP: with the step-cap test.
template <int W, int RULE, bool WALK, bool MULTI, bool BLK, bool CAP>
void launch_flow_c(int det, int fw, int blocks, size_t lds, hipStream_t st, const StageArgs<W> &A) {
    if (det == WAR_DET_BRENT) {
        if constexpr (W == 4 && RULE == 0 && !MULTI) {
            if (fw == 1) {
                hipLaunchKernelGGL((k_flow<4, 0, FLOW_BRENT, WALK, false, BLK, CAP, 127u>), dim3(blocks), dim3(256), lds, st, A);
                return;
            }
            if (fw == 2) {
                hipLaunchKernelGGL((k_flow<4, 0, FLOW_BRENT, WALK, false, BLK, CAP, 0xFFFFFFFFu>), dim3(blocks), dim3(256), lds, st, A);
                return;
            }
            if constexpr (!BLK && !CAP) {
                if (fw == 3) {
                    hipLaunchKernelGGL((k_flow<4, 0, FLOW_BRENT, WALK, false, false, false, 127u, true>), dim3(blocks), dim3(256), lds, st, A);
                    return;
                }
                if (fw == 4) {
                    hipLaunchKernelGGL((k_flow<4, 0, FLOW_BRENT, WALK, false, false, false, 0xFFFFFFFFu, true>), dim3(blocks), dim3(256), lds, st, A);
                    return;
                }
            }
        }
        hipLaunchKernelGGL((k_flow<W, RULE, FLOW_BRENT, WALK, MULTI, BLK, CAP>), dim3(blocks), dim3(256), lds, st, A);
    } else if constexpr (RULE == 0) {
        if (det == WAR_DET_ALT)
            hipLaunchKernelGGL((k_flow<W, 0, FLOW_ALT, WALK, MULTI, BLK, CAP>), dim3(blocks), dim3(256), lds, st, A);
        else if constexpr (!MULTI && W != 6)
            hipLaunchKernelGGL((k_flow<W, 0, FLOW_PAIR, WALK, false, BLK, CAP>), dim3(blocks), dim3(256), lds, st, A);
    }
}
template <int W, int RULE, bool WALK, bool MULTI, bool BLK = false>
void launch_flow(int det, int fw, bool cap, int blocks, size_t lds, hipStream_t st, const StageArgs<W> &A) {
    if (cap) launch_flow_c<W, RULE, WALK, MULTI, BLK, true>(det, fw, blocks, lds, st, A);
    else launch_flow_c<W, RULE, WALK, MULTI, BLK, false>(det, fw, blocks, lds, st, A);
}

// Enqueue one batch's kernels on stream st (inside a capture): k_setup, then the flow kernel or the
// staged pipeline, then the cycle walks. pipe / hdr / jobs: the batch's counters and job list.
template <int W, int RULE>
void record_launch(Engine &E, const GraphKey &key, hipStream_t st, Pipe *pipe, JobHdr *hdr, Job *jobsd,
                   const ReqDev *reqs = nullptr) {
    typedef typename Lay<W>::T T;
    if (!reqs) reqs = E.reqs;
    QueueBufs &Q = qbufs<W>(E);
    const int blocks = std::max(1, E.cus * key.wpc / 4);
    const size_t lds = (size_t)4 * (key.hb + key.lb) * sizeof(u32);
    const std::vector<u32> caps = key.kind == 3 ? std::vector<u32>{}
                                  : key.det == WAR_DET_ALT ? stage_caps_alt(key.cap0) : stage_caps(key.cap0);
    hipLaunchKernelGGL(k_setup, dim3(256), dim3(256), 0, st, pipe, (const JobHdr *)hdr,
                       key.walk ? Q.ccnt : (u32 *)nullptr);
    if (key.kind == 3) {                // the flow kernel: one launch plays the whole batch
        const size_t flds = lds + flow_pool_bytes<W>(key.multi != 0, key.walk != 0);   // pools, then histograms
        StageArgs<W> A{};
        A.pipe = pipe; A.hdr = hdr; A.jobs = jobsd; A.reqs = reqs;
        A.per = (u32)key.per; A.tile_pos = 64u * (u32)key.per;
        A.hb = (u32)key.hb; A.lb = (u32)key.lb;
        constexpr bool CW = W != 6;     // cycle walks exist for layouts 4 and 5
        if (key.walk) {
            A.cq = cyc_queue<W>(E);
            if constexpr (CW) {
                if (key.multi) launch_flow<W, RULE, true, true>(key.det, key.fw, key.cap0 != 0, blocks, flds, st, A);
                else launch_flow<W, RULE, true, false>(key.det, key.fw, key.cap0 != 0, blocks, flds, st, A);
            }
        } else if (key.blk) {
            if constexpr (W == 4) {
                if (key.multi) launch_flow<4, RULE, false, true, true>(key.det, key.fw, key.cap0 != 0, blocks, flds, st, A);
                else launch_flow<4, RULE, false, false, true>(key.det, key.fw, key.cap0 != 0, blocks, flds, st, A);
            }
        } else {
            if (key.multi) launch_flow<W, RULE, false, true>(key.det, key.fw, key.cap0 != 0, blocks, flds, st, A);
            else launch_flow<W, RULE, false, false>(key.det, key.fw, key.cap0 != 0, blocks, flds, st, A);
        }
    }
    u32 a0 = 0;
    for (size_t s = 0; s < caps.size(); s++) {
        StageArgs<W> A{};
        A.pipe = pipe; A.hdr = hdr; A.jobs = jobsd; A.reqs = reqs;
        const int in = (int)((s + 1) & 1), out = (int)(s & 1);
        A.qin.a = (T *)Q.qa[in]; A.qin.b = (T *)Q.qb[in]; A.qin.req = Q.qreq[in]; A.qin.pos = Q.qpos[in];
        A.qin.home = Q.qhome[in]; A.qin.cnt = Q.qcnt[in];
        A.qout.a = (T *)Q.qa[out]; A.qout.b = (T *)Q.qb[out]; A.qout.req = Q.qreq[out]; A.qout.pos = Q.qpos[out];
        A.qout.home = Q.qhome[out]; A.qout.cnt = Q.qcnt[out];
        A.qin.ta = (T *)Q.qta[in]; A.qin.tb = (T *)Q.qtb[in]; A.qin.meta = Q.qmeta[in];
        A.qout.ta = (T *)Q.qta[out]; A.qout.tb = (T *)Q.qtb[out]; A.qout.meta = Q.qmeta[out];
        A.cq = cyc_queue<W>(E);
        A.per = (u32)key.per; A.tile_pos = 64u * (u32)key.per;
        A.a0 = a0; A.a1 = caps[s]; A.stage = (int)s;
        A.hb = (u32)key.hb; A.lb = (u32)key.lb;
        A.max_regions = Q.cap / A.tile_pos;
        const int flags = (key.multi ? F_MULTI : 0) | (key.walk ? F_WALK : 0) | (key.det == WAR_DET_ALT ? F_ALT : 0);
        // refill stages need a bounded window (no tortoise move inside); the unbounded last stage
        // and the first (enumerating) stage are lockstep
        const bool refill = W != 6 && key.det != WAR_DET_ALT && key.kind == 1 && s > 0 && caps[s] != UNBOUNDED;
        A.in_refill = (W != 6 && key.det != WAR_DET_ALT && key.kind == 1 && s > 1) ? 1u : 0u;
        if (refill) launch_refill<W, RULE>(flags, blocks, lds, st, A);
        else launch_stage<W, RULE>(s == 0, flags, blocks, lds, st, A);
        a0 = caps[s];
    }
    if constexpr (W != 6) {
        if (key.walk) {
            WalkArgs<W> WA{};
            WA.pipe = pipe; WA.hdr = hdr; WA.reqs = reqs; WA.cq = cyc_queue<W>(E); WA.tile_pos = 64u * (u32)key.per;
            WA.stack = E.reach_stack;
            for (u32 ph = 0; ph < 2; ph++) {     // the cycle-state tables' inserts, then the walks
                WA.phase = ph;
                hipLaunchKernelGGL((k_walk<W, RULE>), dim3(blocks), dim3(256), 0, st, WA);
            }
        }
    }
}

void record_launch_any(Engine &E, const GraphKey &key, hipStream_t st, Pipe *pipe, JobHdr *hdr, Job *jobsd,
                       const ReqDev *reqs = nullptr) {
    switch (key.W * 2 + key.rule) {
        case 8: record_launch<4, 0>(E, key, st, pipe, hdr, jobsd, reqs); break;
        case 9: record_launch<4, 1>(E, key, st, pipe, hdr, jobsd, reqs); break;
        case 10: record_launch<5, 0>(E, key, st, pipe, hdr, jobsd, reqs); break;
        case 11: record_launch<5, 1>(E, key, st, pipe, hdr, jobsd, reqs); break;
        case 12: record_launch<6, 0>(E, key, st, pipe, hdr, jobsd, reqs); break;
        default: record_launch<6, 1>(E, key, st, pipe, hdr, jobsd, reqs); break;
    }
}

int end_capture(Engine &E, hipGraphExec_t *out, hipStream_t st = nullptr) {
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(st ? st : E.st, &g);
    if (ce != hipSuccess) return fail(30, "graph capture failed: %s", hipGetErrorString(ce));
    HIPCHK(hipGraphInstantiate(out, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    return 0;
}

template <int W, int RULE>
int build_graph(Engine &E, const GraphKey &key, hipGraphExec_t *out) {
    HIPCHK(hipStreamBeginCapture(E.st, hipStreamCaptureModeThreadLocal));
    record_launch<W, RULE>(E, key, E.st, E.pipe, E.hdr, E.jobs);
    return end_capture(E, out);
}

// One persistent launch planned for a batch: the graph key of its kernels and its job list.
struct Launch {
    GraphKey key{};
    std::vector<Job> jobs;
    bool flow = false;
};

// Plan one launch of jobs (all of layout W and rule RULE): the checks, tiles and graph key.
template <int W, int RULE>
int plan_launch(Engine &E, std::vector<Job> &jobs, const war_tuning &t, int hb, int lb, bool walk, bool multi,
                int det, bool blk, Launch *out) {
    if ((int)jobs.size() > MAXJOBS) return fail(31, "too many jobs in one batch");
    const bool alt = det == WAR_DET_ALT;
    // the flow kernel (stage_kind 3, the default) plays every batch, of one request or several;
    // stage_kind 1 / 2 select the staged pipeline. Without cycle walks it needs no queue buffers.
    const bool flow = t.stage_kind == 3 && (RULE == 0 || det == WAR_DET_BRENT);
    if (det == WAR_DET_PAIR && (!flow || multi))
        return fail(12, "the pair-run detector runs on the flow kernel (stage_kind 3, single requests)");
    int rc = (flow && !walk) ? 0 : ensure_queues<W>(E, walk, alt);
    if (rc) return rc;
    const u64 tile_pos = (u64)t.ranks_per_refill;
    u64 tiles = 0;
    for (Job &j : jobs) { j.tile_base = tiles; tiles += (j.hi - j.lo + tile_pos - 1) / tile_pos; }
    const int kind = flow ? 3 : (t.stage_kind == 3 ? 2 : t.stage_kind);
    // flow: cap0 = 1 when some request of the launch has an explicit step cap (the kernel variant
    // with the cap test); brent_cap's stage windows are unused
    bool capped = false;
    for (const Job &j : jobs) capped |= E.slot_cap[j.req] != 0u;
    const int cap0 = flow ? (capped ? 1 : 0) : t.brent_cap;
    out->key = GraphKey{W, RULE, 1, t.waves_per_cu, t.ranks_per_refill / 64, cap0, walk ? 1 : 0, multi ? 1 : 0,
                        round_pow2(hb), round_pow2(lb), kind, det, (blk && flow && W == 4 && !walk) ? 1 : 0};
    if ((!flow || walk) && (tiles > qbufs<W>(E).tmax || tiles * tile_pos > qbufs<W>(E).cap))
        return fail(32, "batch exceeds the queue capacity");
    if (flow && W == 4 && RULE == 0 && det == WAR_DET_BRENT && !multi && !jobs.empty())
        out->key.fw = fw_of((int)E.slot_n[jobs[0].req], (int)E.slot_mode[jobs[0].req], capped);
    out->jobs = jobs;
    out->flow = flow;
    return 0;
}

// Write a launch's job list (header, then the jobs) into pinned staging at sp.
void stage_jobs(Engine &E, const Launch &L, char *sp) {
    JobHdr h{};
    u64 tiles = 0;
    const u64 tp = 64u * (u64)L.key.per;
    for (const Job &j : L.jobs) tiles += (j.hi - j.lo + tp - 1) / tp;
    h.njobs = (u32)L.jobs.size(); h.req0 = L.jobs[0].req; h.ntiles = tiles;
    memcpy(sp, &h, sizeof h);
    memcpy(sp + JOBS_OFF, L.jobs.data(), L.jobs.size() * sizeof(Job));
}

int get_graph(Engine &E, const GraphKey &key, hipGraphExec_t *ge) {
    auto it = E.graphs.find(key);
    if (it != E.graphs.end()) { *ge = it->second; return 0; }
    int rc;
    switch (key.W * 2 + key.rule) {
        case 8: rc = build_graph<4, 0>(E, key, ge); break;
        case 9: rc = build_graph<4, 1>(E, key, ge); break;
        case 10: rc = build_graph<5, 0>(E, key, ge); break;
        case 11: rc = build_graph<5, 1>(E, key, ge); break;
        case 12: rc = build_graph<6, 0>(E, key, ge); break;
        default: rc = build_graph<6, 1>(E, key, ge); break;
    }
    if (!rc) E.graphs[key] = *ge;
    return rc;
}

// One planned launch as its own graph (the engine's slot-0 pipe and job list), synchronously.
int exec_single(Engine &E, const Launch &L, float *ms) {
    if (L.jobs.empty()) return 0;
    // copy the job list in (the staging buffer is free: the previous batch was synchronised)
    char *sp = (char *)E.stage_host;
    stage_jobs(E, L, sp);
    HIPCHK(hipMemcpyAsync(E.hdr, sp, JOBS_OFF + L.jobs.size() * sizeof(Job), hipMemcpyHostToDevice, E.st));
    hipGraphExec_t ge = nullptr;
    int rc = get_graph(E, L.key, &ge);
    if (rc) return rc;
    HIPCHK(hipEventRecord(E.ev0, E.st));
    HIPCHK(hipGraphLaunch(ge, E.st));
    HIPCHK(hipEventRecord(E.ev1, E.st));
    u64 ovf = 0;
    HIPCHK(hipMemcpyAsync(&ovf, &E.pipe->ovf, sizeof ovf, hipMemcpyDeviceToHost, E.st));
    HIPCHK(hipStreamSynchronize(E.st));
    float m = 0;
    HIPCHK(hipEventElapsedTime(&m, E.ev0, E.ev1));
    if (ms) *ms = m;
    if (ovf) return fail(33, "persistent pipeline ran out of queue regions (%llu refused)", ovf);
    return 0;
}

// Whether a plan runs as ONE graph: every launch a flow launch, at most MAXB of them. *nl: launches.
bool one_graph(const std::vector<Launch> &plan, size_t *nl) {
    size_t n = 0;
    bool all_flow = true;
    for (const Launch &L : plan) {
        if (L.jobs.empty()) continue;
        n++;
        all_flow = all_flow && L.flow;
    }
    *nl = n;
    return n > 0 && all_flow && n <= (size_t)MAXB;
}

// Launch a plan as ONE graph on step context c, without waiting: each launch on a branch of its own
// with its own pipe counters and job list, forked from and joined to the context's stream; launches
// with cycle walks of one pile layout share that layout's cycle records and follow each other on one
// branch. The job lists are copied in first and the pipe counters read back after (into pinned
// memory), all on the context's stream, between its timing events. The caller has checked
// one_graph(plan) and that the context has no step in flight.
int launch_plan(Engine &E, StepCtx &c, std::vector<Launch> &plan) {
    // branches: each launch without walks alone; the launches with walks, one branch per layout
    std::vector<const Launch *> ls;
    for (const Launch &L : plan) if (!L.jobs.empty()) ls.push_back(&L);
    std::vector<std::vector<int>> br;
    int walk_branch[7] = {-1, -1, -1, -1, -1, -1, -1};
    bool walk = false;
    for (int i = 0; i < (int)ls.size(); i++) {
        const GraphKey &k = ls[(size_t)i]->key;
        if (k.walk) {
            walk = true;
            int &b = walk_branch[k.W];
            if (b < 0) { b = (int)br.size(); br.emplace_back(); }
            br[(size_t)b].push_back(i);
        } else {
            br.push_back({i});
        }
    }
    // the graph's shape: branch by branch, the keys of its launches (slot = position in ls)
    std::vector<int> bkey;
    for (const auto &b : br) {
        bkey.push_back(-1);
        for (int i : b) {
            const GraphKey &k = ls[(size_t)i]->key;
            for (int v : {i, k.W, k.rule, k.kernel, k.wpc, k.per, k.cap0, k.walk, k.multi, k.hb, k.lb, k.kind, k.det, k.blk, k.fw})
                bkey.push_back(v);
        }
    }
    // job lists of every slot, one copy
    char *sp = (char *)c.stage;
    for (size_t i = 0; i < ls.size(); i++) stage_jobs(E, *ls[i], sp + i * SLOT_BYTES);
    HIPCHK(hipMemcpyAsync(c.hdr, sp, ls.size() * SLOT_BYTES, hipMemcpyHostToDevice, c.st));
    hipGraphExec_t ge = nullptr;
    auto it = c.graphs.find(bkey);
    if (it != c.graphs.end()) {
        ge = it->second;
    } else {
        HIPCHK(hipStreamBeginCapture(c.st, hipStreamCaptureModeThreadLocal));
        HIPCHK(hipEventRecord(E.bfork, c.st));
        for (size_t b = 0; b < br.size(); b++) {
            hipStream_t st = E.bst[b];
            HIPCHK(hipStreamWaitEvent(st, E.bfork, 0));
            for (int i : br[b]) {
                char *hd = c.hdr + (size_t)i * SLOT_BYTES;
                record_launch_any(E, ls[(size_t)i]->key, st, c.pipe + i, (JobHdr *)hd, (Job *)(hd + JOBS_OFF));
            }
            HIPCHK(hipEventRecord(E.bjoin[b], st));
            HIPCHK(hipStreamWaitEvent(c.st, E.bjoin[b], 0));
        }
        const int rc = end_capture(E, &ge, c.st);
        if (rc) return rc;
        c.graphs[bkey] = ge;
    }
    HIPCHK(hipEventRecord(c.ev0, c.st));
    HIPCHK(hipGraphLaunch(ge, c.st));
    HIPCHK(hipEventRecord(c.ev1, c.st));
    HIPCHK(hipMemcpyAsync(c.pipe_host, c.pipe, ls.size() * sizeof(Pipe), hipMemcpyDeviceToHost, c.st));
    c.nl = ls.size();
    c.walk = walk;
    return 0;
}

// Wait for the step in flight on context c: its device time into *ms (added), and its pipe counters
// checked. The context is free again afterwards.
int wait_plan(StepCtx &c, double *ms) {
    const size_t nl = c.nl;
    c.nl = 0;
    c.walk = false;
    if (!nl) return 0;
    HIPCHK(hipStreamSynchronize(c.st));
    float m = 0;
    HIPCHK(hipEventElapsedTime(&m, c.ev0, c.ev1));
    if (ms) *ms += m;
    for (size_t i = 0; i < nl; i++)
        if (c.pipe_host[i].ovf)
            return fail(33, "persistent pipeline ran out of queue regions (%llu refused)", c.pipe_host[i].ovf);
    return 0;
}

// Several planned launches, synchronously, on context 0. Flow launches (up to MAXB of them) run as
// ONE graph (launch_plan); anything else (staged launches, more than MAXB) runs launch by launch.
// *ms: device time; *graphs: graph launches made.
int exec_plan(Engine &E, std::vector<Launch> &plan, double *ms, int *graphs) {
    size_t nl = 0;
    const bool one = one_graph(plan, &nl);
    if (nl == 0) return 0;
    if (nl == 1 || !one) {
        for (const Launch &L : plan) {
            if (L.jobs.empty()) continue;
            float m = 0;
            const int rc = exec_single(E, L, &m);
            if (ms) *ms += m;
            if (graphs) ++*graphs;
            if (rc) return rc;
        }
        return 0;
    }
    int rc = launch_plan(E, E.sc[0], plan);
    if (rc) { E.sc[0].nl = 0; (void)hipStreamSynchronize(E.sc[0].st); return rc; }
    if (graphs) ++*graphs;
    return wait_plan(E.sc[0], ms);
}

// One batch of jobs (all of layout W and rule RULE) through the persistent pipeline, synchronously.
template <int W, int RULE>
int run_persistent_batch(Engine &E, std::vector<Job> &jobs, const war_tuning &t, int hb, int lb, bool walk,
                         bool multi, float *ms, int det, bool blk = false) {
    if (jobs.empty()) return 0;
    Launch L;
    const int rc = plan_launch<W, RULE>(E, jobs, t, hb, lb, walk, multi, det, blk, &L);
    if (rc) return rc;
    return exec_single(E, L, ms);
}

template <int W>
int plan_dispatch(Engine &E, int rule, std::vector<Job> &jobs, const war_tuning &t, int hb, int lb, bool walk,
                  bool multi, int det, bool blk, Launch *out) {
    return rule ? plan_launch<W, 1>(E, jobs, t, hb, lb, walk, multi, WAR_DET_BRENT, blk, out)
                : plan_launch<W, 0>(E, jobs, t, hb, lb, walk, multi, det, blk, out);
}

// The plain kernel for a request's detector, with the explicit-cap variant where the request has one.
template <int W, int RULE>
void launch_plain(int det, bool capped, unsigned blocks, size_t lds, hipStream_t st, const PlainArgs<W> &P) {
    if (RULE == 0 && det == WAR_DET_STREAK) {
        if (capped) hipLaunchKernelGGL((k_plain<W, 0, 1, true>), dim3(blocks), dim3(256), lds, st, P);
        else hipLaunchKernelGGL((k_plain<W, 0, 1>), dim3(blocks), dim3(256), lds, st, P);
    } else if (RULE == 0 && det == WAR_DET_ALT) {
        if (capped) hipLaunchKernelGGL((k_plain<W, 0, 2, true>), dim3(blocks), dim3(256), lds, st, P);
        else hipLaunchKernelGGL((k_plain<W, 0, 2>), dim3(blocks), dim3(256), lds, st, P);
    } else {
        if (capped) hipLaunchKernelGGL((k_plain<W, RULE, 0, true>), dim3(blocks), dim3(256), lds, st, P);
        else hipLaunchKernelGGL((k_plain<W, RULE, 0>), dim3(blocks), dim3(256), lds, st, P);
    }
}

template <int W, int RULE>
int run_plain_batch(Engine &E, ReqHost &h, u64 lo, u64 hi, const war_tuning &t, float *ms) {
    const bool walk = h.dev.walk != 0;
    int rc = ensure_queues<W>(E, walk);
    if (rc) return rc;
    PlainArgs<W> P{};
    P.lo = lo; P.hi = hi; P.per = (u32)t.plain_per;
    P.hb = (u32)h.rq.hbins; P.lb = (u32)h.rq.lbins; P.tile_pos = PLAIN_CYC_TILE;
    P.R = h.dev; P.req = (u32)h.slot; P.pipe = E.pipe;
    P.cq = cyc_queue<W>(E);
    // the walks read the cycle records by home tile: (pos - lo) / PLAIN_CYC_TILE
    const u64 tiles = (hi - lo + PLAIN_CYC_TILE - 1) / PLAIN_CYC_TILE;
    JobHdr jh{};
    jh.njobs = 0; jh.req0 = (u32)h.slot; jh.ntiles = tiles;
    memcpy(E.stage_host, &jh, sizeof jh);
    HIPCHK(hipMemcpyAsync(E.hdr, E.stage_host, sizeof jh, hipMemcpyHostToDevice, E.st));
    const u64 threads = (hi - lo + P.per - 1) / P.per;
    const u64 blocks = (threads + 255) / 256;
    const size_t lds = (size_t)(P.hb + P.lb) * sizeof(u32);
    HIPCHK(hipEventRecord(E.ev0, E.st));
    hipLaunchKernelGGL(k_setup, dim3(256), dim3(256), 0, E.st, E.pipe, (const JobHdr *)E.hdr,
                       walk ? qbufs<W>(E).ccnt : (u32 *)nullptr);
    launch_plain<W, RULE>(h.rq.detector, h.rq.step_cap != 0, (unsigned)blocks, lds, E.st, P);
    HIPCHK(hipGetLastError());
    if (walk) {
        WalkArgs<W> WA{};
        WA.pipe = E.pipe; WA.hdr = E.hdr; WA.reqs = E.reqs; WA.cq = cyc_queue<W>(E); WA.tile_pos = PLAIN_CYC_TILE;
        WA.stack = E.reach_stack;
        for (u32 ph = 0; ph < 2; ph++) {         // the cycle-state tables' inserts, then the walks
            WA.phase = ph;
            hipLaunchKernelGGL((k_walk<W, RULE>), dim3(std::max(1, E.cus * 8)), dim3(256), 0, E.st, WA);
        }
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(E.ev1, E.st));
    HIPCHK(hipStreamSynchronize(E.st));
    float m = 0;
    HIPCHK(hipEventElapsedTime(&m, E.ev0, E.ev1));
    if (ms) *ms = m;
    return 0;
}

template <int W>
int run_batch_dispatch(Engine &E, int rule, std::vector<Job> &jobs, const war_tuning &t, int hb, int lb, bool walk,
                       bool multi, float *ms, int det = WAR_DET_BRENT, bool blk = false) {
    return rule ? run_persistent_batch<W, 1>(E, jobs, t, hb, lb, walk, multi, ms, WAR_DET_BRENT, blk)
                : run_persistent_batch<W, 0>(E, jobs, t, hb, lb, walk, multi, ms, det, blk);
}

int alloc_slot(Engine &E) {
    if (E.free_slots.empty()) return -1;
    int s = E.free_slots.back();
    E.free_slots.pop_back();
    return s;
}

// Deals in whole blocks of 24, lane runs too: a batch for the flow kernel's block-stepping variant.
bool deal_blocks(const war_request &r, int W, u64 lo, u64 hi, const war_tuning &t) {
    return W == 4 && r.mode == WAR_MODE_FULL && r.n >= 5 && lo % 24 == 0 && hi % 24 == 0 &&
           (t.ranks_per_refill / 64) % 24 == 0;
}

// One persistent batch [lo, hi) of a single request.
int run_range_batch(Engine &E, ReqHost &h, int W, u64 lo, u64 hi, const war_tuning &t, double *dev) {
    const war_request &r = h.rq;
    std::vector<Job> jobs(1);
    jobs[0].lo = lo; jobs[0].hi = hi; jobs[0].req = (u32)h.slot;
    float ms = 0;
    const int det = r.detector;
    const bool blk = deal_blocks(r, W, lo, hi, t);
    int rc;
    if (W == 4)      rc = run_batch_dispatch<4>(E, r.rule, jobs, t, r.hbins, r.lbins, h.dev.walk != 0, false, &ms, det, blk);
    else if (W == 6) rc = run_batch_dispatch<6>(E, r.rule, jobs, t, r.hbins, r.lbins, false, false, &ms, det);
    else             rc = run_batch_dispatch<5>(E, r.rule, jobs, t, r.hbins, r.lbins, h.dev.walk != 0, false, &ms, det);
    *dev += ms;
    return rc;
}

// The persistent batches of a request's range [r.lo, r.hi), in order: run_one(lo, hi, tuning) for
// each. On the flow kernel, batches of FLOW_BATCH positions (the cycle records bound them to `batch`
// with walks); canonical lane runs of whole orders starting at split 0; deals in whole blocks of 24
// with the parts before the first and after the last block boundary apart.
template <class F>
int for_each_persistent_batch(const war_request &r, const war_tuning &t, int W, bool walk, u64 batch, F &&run_one) {
    int rc = 0;
    const bool flow = t.stage_kind == 3;
    war_tuning bt = t;
    u64 bstep = batch;
    u64 lo = r.lo, hi = r.hi;    // batches cover [lo, hi); a tail [hi, r.hi) runs after them
    if (flow) {
        // the flow kernel needs no queues, so its batches are larger (fewer batch ends); with
        // cycle walks the cycle records bound them as in the staged pipeline
        if (!walk) bstep = FLOW_BATCH;
        if (r.mode == WAR_MODE_CANON && r.n >= 3) {
            // canonical positions: lane runs of whole orders (a multiple of n-1 positions) that
            // start at split 0, so every lane of a wave steps to its next order (next_perm) in
            // the same pool step instead of a few lanes in almost every step
            const u64 m = (u64)r.n - 1u;
            const u64 per = std::max<u64>(1, ((u64)bt.ranks_per_refill / 64 + m / 2) / m) * m;
            if (64 * per <= MAX_TILE) bt.ranks_per_refill = (int)(64 * per);
            const u64 tp = (u64)bt.ranks_per_refill;
            bstep = bstep / tp * tp;
            const u64 a = (lo + m - 1) / m * m;          // the first order boundary
            if (a > lo && a < r.hi) {                    // the part of an order before it
                rc = run_one(lo, a, bt);
                lo = a;
            }
        } else if (r.mode == WAR_MODE_FULL && W == 4 && r.n >= 5 && !walk) {
            // deals: lane runs of whole blocks of 24 (the orders that share all but the last
            // four cards), which the kernel steps through with one swap of two dealt cards
            // each; the parts before the first and after the last block boundary run apart
            const u64 per = std::max<u64>(1, ((u64)bt.ranks_per_refill / 64 + 12) / 24) * 24;
            if (64 * per <= MAX_TILE) bt.ranks_per_refill = (int)(64 * per);
            const u64 tp = (u64)bt.ranks_per_refill;
            bstep = bstep / tp * tp;
            const u64 a = (lo + 23) / 24 * 24, b = r.hi / 24 * 24;
            if (a < b) {
                if (a > lo) rc = run_one(lo, a, bt);
                lo = a;
                hi = b;
            }
        }
    }
    for (; lo < hi && !rc; lo += bstep) rc = run_one(lo, std::min<u64>(hi, lo + bstep), bt);
    if (hi < r.hi && !rc) rc = run_one(hi, r.hi, bt);
    return rc;
}

// Whole range of one request, synchronously (engine mutex held by the caller).
int run_request_locked(Engine &E, ReqHost &h) {
    const war_request &r = h.rq;
    war_tuning t;
    resolve_tuning(E, r, t);
    big_canon_tile(E, r, t);
    int rc = check_tuning(t);
    if (!rc) rc = check_cap(r, t);
    if (rc) return rc;
    int W = 4;
    rc = layout_of(r, t, &W);
    if (rc) return rc;
    h.layout = W;
    h.slot = alloc_slot(E);
    if (h.slot < 0) return fail(40, "no free request slot");
    rc = req_alloc(E, h);
    if (rc) { req_free(E, h); return rc; }
    double dev = 0;
    if (r.kernel == WAR_KERNEL_PLAIN) {
        const u64 chunk = 1ull << 26;   // contiguous chunks of 2^26 positions per launch
        for (u64 lo = r.lo; lo < r.hi && !rc; lo += chunk) {
            const u64 hi = std::min<u64>(r.hi, lo + chunk);
            float ms = 0;
            if (W == 4) rc = r.rule ? run_plain_batch<4, 1>(E, h, lo, hi, t, &ms) : run_plain_batch<4, 0>(E, h, lo, hi, t, &ms);
            else        rc = r.rule ? run_plain_batch<5, 1>(E, h, lo, hi, t, &ms) : run_plain_batch<5, 0>(E, h, lo, hi, t, &ms);
            dev += ms;
        }
    } else {
        rc = for_each_persistent_batch(r, t, W, h.dev.walk != 0, E.batch,
                                       [&](u64 a, u64 b, const war_tuning &bt) { return run_range_batch(E, h, W, a, b, bt, &dev); });
    }
    h.dev_ms = dev;
    if (!rc) rc = req_collect(E, h);
    req_free(E, h);
    return rc;
}

// One group's ranges (one pile layout, rule, detector and walk flag) planned as multi-request
// persistent launches, all later branches of one graph (exec_plan). With `blocks_apart`, 4-bit deal
// ranges without walks are split at blocks of 24 orders and their whole blocks run on the kernel's
// deal-block variant (lane runs of 24 deals), a launch of their own. On the flow kernel, canonical
// ranges of one n (when they hold 2^26 positions or more) take a launch of their own whose lane runs
// are whole orders from split 0, as a single request's do. Each set of ranges is packed, in order,
// into as few launches as the batch capacity allows (FLOW_BATCH positions on the flow kernel, the
// queue capacity with cycle walks or the staged pipeline; MAXJOBS ranges).
struct GroupRange { u32 slot, n, mode; u64 lo, hi; };
int plan_group(Engine &E, int W, int rule, int det, bool walk, const war_tuning &gt, int hb, int lb,
               const std::vector<GroupRange> &rs, bool blocks_apart, std::vector<Launch> *plan) {
    std::vector<GroupRange> rest, blocks;
    std::map<u32, std::vector<GroupRange>> canon;   // canonical ranges by n, from an order boundary on
    war_tuning bt = gt;
    {
        const u64 per = std::max<u64>(1, ((u64)gt.ranks_per_refill / 64 + 12) / 24) * 24;
        if (64 * per <= MAX_TILE) bt.ranks_per_refill = (int)(64 * per);
    }
    const bool blk_ok =
        blocks_apart && W == 4 && !walk && gt.stage_kind == 3 && (bt.ranks_per_refill / 64) % 24 == 0;
    for (const GroupRange &x : rs) {
        if (gt.stage_kind == 3 && x.mode == WAR_MODE_CANON && x.n >= 3) {
            // canonical positions as a single request runs them: a launch per n whose lane runs are
            // whole orders starting at split 0, so the lanes of a wave step to their next order
            // together; the part of an order before the first boundary runs with the rest
            const u64 m = x.n - 1u, a = (x.lo + m - 1) / m * m;
            if (a < x.hi) {
                if (x.lo < a) rest.push_back({x.slot, x.n, x.mode, x.lo, a});
                canon[x.n].push_back({x.slot, x.n, x.mode, a, x.hi});
                continue;
            }
        }
        if (blk_ok && x.mode == WAR_MODE_FULL && x.n >= 5) {
            const u64 a = (x.lo + 23) / 24 * 24, b = x.hi / 24 * 24;
            if (a < b) {
                if (x.lo < a) rest.push_back({x.slot, x.n, x.mode, x.lo, a});
                blocks.push_back({x.slot, x.n, x.mode, a, b});
                if (b < x.hi) rest.push_back({x.slot, x.n, x.mode, b, x.hi});
                continue;
            }
        }
        if (x.lo < x.hi) rest.push_back(x);
    }
    auto pack = [&](const std::vector<GroupRange> &v, const war_tuning &tt, bool blk) -> int {
        const u64 tp = (u64)tt.ranks_per_refill;
        const u64 cap = (walk || tt.stage_kind != 3) ? E.batch : FLOW_BATCH;
        std::vector<Job> jobs;
        u64 used = 0;          // tile-rounded positions of the launch being packed
        int rc = 0;
        auto flush = [&]() -> int {
            if (jobs.empty()) return 0;
            plan->emplace_back();
            Launch *L = &plan->back();
            const int r2 = (W == 4) ? plan_dispatch<4>(E, rule, jobs, tt, hb, lb, walk, true, det, blk, L)
                         : (W == 6) ? plan_dispatch<6>(E, rule, jobs, tt, hb, lb, false, true, det, false, L)
                                    : plan_dispatch<5>(E, rule, jobs, tt, hb, lb, walk, true, det, false, L);
            jobs.clear();
            used = 0;
            return r2;
        };
        for (const GroupRange &x : v) {
            for (u64 lo = x.lo; lo < x.hi && !rc;) {
                if ((int)jobs.size() == MAXJOBS || cap - used < tp) { rc = flush(); continue; }
                const u64 len = std::min<u64>(x.hi - lo, (cap - used) / tp * tp);
                Job j{};
                j.lo = lo; j.hi = lo + len; j.req = x.slot;
                jobs.push_back(j);
                used += (len + tp - 1) / tp * tp;
                lo += len;
            }
        }
        if (!rc) rc = flush();
        return rc;
    };
    // a bucket too small to fill the GPU a few times over joins the rest (its own launch would cost
    // more than the aligned runs save)
    for (auto it = canon.begin(); it != canon.end();) {
        u64 tot = 0;
        for (const GroupRange &x : it->second) tot += x.hi - x.lo;
        if (tot < (1ull << 26)) {
            rest.insert(rest.end(), it->second.begin(), it->second.end());
            it = canon.erase(it);
        } else {
            ++it;
        }
    }
    int rc = pack(blocks, bt, true);
    if (!rc) rc = pack(rest, gt, false);
    for (auto it = canon.begin(); it != canon.end() && !rc; ++it) {
        const u64 m = it->first - 1u;
        war_tuning ct = gt;
        const u64 per = std::max<u64>(1, ((u64)gt.ranks_per_refill / 64 + m / 2) / m) * m;
        if (64 * per <= MAX_TILE) ct.ranks_per_refill = (int)(64 * per);
        rc = pack(it->second, ct, false);
    }
    return rc;
}

// Several requests at once, synchronously (engine mutex held by the caller): the serving loop's
// multi-request launches without its thread. Requests are grouped by pile layout, rule, detector
// and whether they need cycle walks (so requests without walks keep the larger batches of the
// kernel without them); each group's whole ranges are planned as launches (one while they fit, or
// two with deal blocks), and every launch of the batch runs as branches of one graph (exec_plan).
// *launches counts the graph launches, *dev_total the device time.
int run_batch_locked(Engine &E, std::vector<ReqHost> &hs, int *launches, double *dev_total = nullptr) {
    int rc = 0;
    for (auto &h : hs) {
        war_tuning tt;
        resolve_tuning(E, h.rq, tt);
        rc = check_tuning(tt);
        if (!rc) rc = check_cap(h.rq, tt);
        if (!rc) rc = layout_of(h.rq, tt, &h.layout);
        if (!rc) {
            h.slot = alloc_slot(E);
            if (h.slot < 0) rc = fail(40, "more requests than free request slots");
        }
        if (!rc) rc = req_alloc(E, h);
        if (rc) break;
    }
    std::vector<Launch> plan;
    for (int W : {4, 5, 6}) {
        for (int rule : {0, 1}) {
            for (int alt : {0, 1}) {
              for (int wk : {0, 1}) {
                if (rc) break;
                std::vector<ReqHost *> grp;
                for (auto &h : hs)
                    if (h.layout == W && h.rq.rule == rule && (h.rq.detector == WAR_DET_ALT) == (alt != 0) &&
                        (h.dev.walk != 0) == (wk != 0) && h.rq.lo < h.rq.hi)
                        grp.push_back(&h);
                if (grp.empty()) continue;
                war_tuning gt;
                resolve_tuning(E, grp[0]->rq, gt);
                int hb = 0, lb = 0;
                bool walk = false;
                for (ReqHost *h : grp) {
                    hb = std::max(hb, h->rq.hbins); lb = std::max(lb, h->rq.lbins);
                    walk = walk || h->dev.walk;
                }
                const int det = alt ? WAR_DET_ALT : WAR_DET_BRENT;
                std::vector<GroupRange> rs;
                for (ReqHost *h : grp) rs.push_back({(u32)h->slot, (u32)h->rq.n, (u32)h->rq.mode, h->rq.lo, h->rq.hi});
                rc = plan_group(E, W, rule, det, walk, gt, hb, lb, rs, true, &plan);
              }
            }
        }
    }
    double dev = 0;
    if (!rc) rc = exec_plan(E, plan, &dev, launches);
    for (auto &h : hs) h.dev_ms = dev;     // device time of the batch's graph
    if (dev_total) *dev_total = dev;
    for (auto &h : hs) {
        if (!rc && h.slot >= 0) rc = req_collect(E, h);
        req_free(E, h);
    }
    return rc;
}

// ---------------------------------------------------------------------------------------------
// Serving loop: continuous batching. Each step takes, from every admitted request, its next share
// of positions, plans each group of one layout, rule, detector and walk flag as launches, and runs
// every group's launches as one graph. Steps alternate between the two step contexts and are
// pipelined: the worker launches step k while step k - 1 still runs, then waits for step k - 1 and
// retires the requests it completed, so the GPU does not wait for the host between steps (and the
// two steps' kernels may overlap on the device: a request's counters only take atomic adds, and the
// ranges of two steps are disjoint). Requests that arrive during a step are admitted at the next
// one, while slots and device memory allow.

// Retire the active requests with no positions left and no step in flight: their counters are
// fetched on stream st (one synchronisation), their memory returns to the pool, waiters are woken.
void retire(Engine &E, std::vector<std::shared_ptr<ReqHost>> &active, hipStream_t st) {
    std::vector<std::shared_ptr<ReqHost>> done;
    for (size_t i = 0; i < active.size();) {
        if (active[i]->next >= active[i]->rq.hi && active[i]->inflight == 0) {
            done.push_back(active[i]);
            active.erase(active.begin() + (long)i);
        } else {
            i++;
        }
    }
    if (done.empty()) return;
    std::vector<std::vector<u64>> bufs(done.size());
    for (size_t i = 0; i < done.size(); i++) {
        ReqHost &h = *done[i];
        if (!h.status) {
            const int c = req_fetch(h, st, bufs[i]);
            if (c) { h.status = c; h.err = t_err; }
        }
    }
    const hipError_t se = hipStreamSynchronize(st);
    for (size_t i = 0; i < done.size(); i++) {
        ReqHost &h = *done[i];
        if (!h.status && se != hipSuccess) {
            h.status = fail(100 + (int)se, "hipStreamSynchronize: %s", hipGetErrorString(se));
            h.err = t_err;
        }
        if (!h.status) {
            const int c = req_finish(h, bufs[i]);
            if (c) { h.status = c; h.err = t_err; }
        }
        req_free(E, h);
    }
    std::lock_guard<std::mutex> lk(E.smu);
    const double t = now_ms();
    for (auto &h : done) {
        h->finished = true;
        h->t_done = t;
        if (h->res) h->res->latency_ms = h->t_done - h->t_submit;
    }
    E.scv.notify_all();
}

// Wait for the step in flight on context c; the requests it held take its device time (or its
// failure).
void finish_step(StepCtx &c) {
    double dms = 0;
    const int rc = wait_plan(c, &dms);
    const std::string err = rc ? t_err : std::string();
    for (auto &h : c.stepped) {
        h->dev_ms += dms;        // device time of the steps the request took part in
        h->inflight--;
        if (rc && !h->status) { h->status = rc; h->err = err; h->next = h->rq.hi; }
    }
    c.stepped.clear();
}

void serve_loop(Engine *Ep) {
    Engine &E = *Ep;
    (void)hipSetDevice(E.device);     // this thread's current device: the engine's
    std::vector<std::shared_ptr<ReqHost>> active;
    int cur = 0;          // the context of the next step
    int prev = -1;        // the context of the step in flight (-1: none)
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(E.smu);
            E.scv.wait(lk, [&] { return !E.serving || !E.pending.empty() || !active.empty(); });
            if (!E.serving && E.pending.empty() && active.empty()) break;
        }
        std::lock_guard<std::mutex> g(E.mu);
        StepCtx &C = E.sc[cur];
        // admit: a new request's counters are cleared and its descriptor written on the stream of
        // the step that first runs it
        for (;;) {
            std::shared_ptr<ReqHost> h;
            {
                std::lock_guard<std::mutex> lk(E.smu);
                if (E.pending.empty() || (int)active.size() >= E.serve_slots) break;
                h = E.pending.front();
            }
            size_t fr = 0, tot = 0;
            (void)hipMemGetInfo(&fr, &tot);
            const u32 sl = stab_log2(h->rq);
            const size_t need = (size_t)(C_N + h->rq.hbins + h->rq.lbins + h->rq.dbins) * 8 +
                                (h->rq.distinct ? ((size_t)8 << 27) : 0) + (sl ? (size_t)8 << sl : 0) +
                                ((size_t)1 << 30);
            if (fr < need && !active.empty()) break;   // wait for memory to free up
            {
                std::lock_guard<std::mutex> lk(E.smu);
                E.pending.pop_front();
            }
            int rc = 0;
            {
                war_tuning tt;
                resolve_tuning(E, h->rq, tt);
                rc = check_tuning(tt);
                if (!rc) rc = check_cap(h->rq, tt);
                if (!rc) rc = layout_of(h->rq, tt, &h->layout);
            }
            if (!rc) {
                h->slot = alloc_slot(E);
                if (h->slot < 0) rc = fail(40, "no free request slot");
            }
            if (!rc) rc = req_alloc(E, *h, C.st, true);
            if (rc) { h->status = rc; h->err = t_err; }
            h->next = rc ? h->rq.hi : h->rq.lo;     // a refused request retires at the end of this pass
            active.push_back(h);
        }
        // the step's positions: a quarter to the oldest request with work left (none starves), the
        // rest shortest remaining first, so a request whose rest fits finishes in this step and a
        // step's launches hold a few large ranges rather than a sliver of every request
        {
            std::vector<std::pair<u64, ReqHost *>> order;
            ReqHost *oldest = nullptr;
            for (auto &h : active) {
                h->share = 0;
                if (h->next >= h->rq.hi) continue;
                order.push_back({h->rq.hi - h->next, h.get()});
                if (!oldest || h->id < oldest->id) oldest = h.get();
            }
            u64 budget = E.step_positions;
            if (oldest) {
                oldest->share = std::min<u64>(oldest->rq.hi - oldest->next, budget / 4);
                budget -= oldest->share;
            }
            std::sort(order.begin(), order.end(), [](const std::pair<u64, ReqHost *> &a, const std::pair<u64, ReqHost *> &b) {
                return a.first != b.first ? a.first < b.first : a.second->id < b.second->id;
            });
            for (auto &o : order) {
                if (!budget) break;
                const u64 more = std::min<u64>(o.first - o.second->share, budget);
                o.second->share += more;
                budget -= more;
            }
        }
        // one launch per (layout, rule, detector, cycle walks) group: requests without walks keep
        // the kernel without them
        std::vector<Launch> plan;
        std::vector<std::shared_ptr<ReqHost>> stepped;
        for (int W : {4, 5, 6}) {
            for (int rule : {0, 1})
            for (int alt : {0, 1})
            for (int wk : {0, 1}) {
                std::vector<std::shared_ptr<ReqHost>> grp;
                for (auto &h : active)
                    if (h->layout == W && h->rq.rule == rule && (h->rq.detector == WAR_DET_ALT) == (alt != 0) &&
                        (h->dev.walk != 0) == (wk != 0) && h->next < h->rq.hi && h->share)
                        grp.push_back(h);
                if (grp.empty()) continue;
                war_tuning gt;
                resolve_tuning(E, grp[0]->rq, gt);
                std::vector<GroupRange> rs;
                int hb = 0, lb = 0;
                bool walk = false;
                for (auto &h : grp) {
                    const u64 hi = std::min<u64>(h->rq.hi, h->next + h->share);
                    rs.push_back({(u32)h->slot, (u32)h->rq.n, (u32)h->rq.mode, h->next, hi});
                    h->next = hi;
                    hb = std::max(hb, h->rq.hbins); lb = std::max(lb, h->rq.lbins);
                    walk = walk || h->dev.walk;
                }
                const int det = alt ? WAR_DET_ALT : WAR_DET_BRENT;
                const size_t np = plan.size();
                const int rc = plan_group(E, W, rule, det, walk, gt, hb, lb, rs, true, &plan);
                if (rc) {
                    plan.resize(np);
                    for (auto &h : grp) { h->status = rc; h->err = t_err; h->next = h->rq.hi; }
                } else {
                    stepped.insert(stepped.end(), grp.begin(), grp.end());
                }
            }
        }
        // launch the step: one graph on this context, after the step in flight; staged launches (or
        // more than one graph holds) run synchronously once the step in flight is done
        bool launched = false;
        size_t nl = 0;
        if (one_graph(plan, &nl)) {
            bool walk = false;
            for (const Launch &L : plan) walk = walk || (!L.jobs.empty() && L.key.walk);
            int rc = 0;
            // the cycle records of a layout are shared: a step with cycle walks waits for the one in flight
            if (walk && prev >= 0 && E.sc[prev].walk) {
                const hipError_t we = hipStreamWaitEvent(C.st, E.sc[prev].ev1, 0);
                if (we != hipSuccess) rc = fail(100 + (int)we, "hipStreamWaitEvent: %s", hipGetErrorString(we));
            }
            if (!rc) rc = launch_plan(E, C, plan);
            if (rc) {
                (void)hipStreamSynchronize(C.st);
                for (auto &h : stepped) { h->status = rc; h->err = t_err; h->next = h->rq.hi; }
            } else {
                launched = true;
                C.stepped = stepped;
                for (auto &h : stepped) h->inflight++;
            }
        } else if (nl) {
            if (prev >= 0) { finish_step(E.sc[prev]); prev = -1; }
            double dms = 0;
            const int rc = exec_plan(E, plan, &dms, nullptr);
            for (auto &h : stepped) {
                h->dev_ms += dms;
                if (rc) { h->status = rc; h->err = t_err; h->next = h->rq.hi; }
            }
        }
        // the step before this one: wait for it, then retire what is complete (fetching on its now
        // idle stream)
        if (prev >= 0) { finish_step(E.sc[prev]); prev = -1; }
        retire(E, active, E.sc[cur ^ 1].st);
        if (launched) { prev = cur; cur ^= 1; }
        if (active.empty()) pool_release(E);   // nothing in flight either (a step holds active requests)
    }
}

// Every GPU entry point other than the serving worker waits for the serving loop's step in flight
// (if any) before it uses the device, which it then has to itself while it holds the engine's mutex.
// Before GPU work on the calling thread (the engine mutex held): make the engine's device current
// (HIP's current device is per host thread, and a call may come from any thread), then wait for the
// serving loop's steps in flight.
void drain(Engine &E) {
    (void)hipSetDevice(E.device);
    for (StepCtx &c : E.sc)
        if (c.nl) (void)hipStreamSynchronize(c.st);
}

}  // namespace

// =============================================================================================
extern "C" {

int war_abi_version(void) { return WAR_ABI_VERSION; }

const char *war_last_error(void) { return t_err.c_str(); }

int war_init(int32_t device, uint32_t flags) {
    t_err.clear();
    if (g_eng) return 0;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) return fail(50, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= ndev) return fail(51, "device %d out of range (%d devices)", device, ndev);
    std::unique_ptr<Engine> E(new Engine());
    E->device = device;
    E->verbose = flags & 1u;
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, device));
    E->cus = p.multiProcessorCount;
    E->clock_mhz = p.clockRate / 1000;
    E->arch = p.gcnArchName;
    if (E->arch.rfind("gfx950", 0) != 0)
        return fail(52, "libwar_engine is built for gfx950; device %d is %s", device, E->arch.c_str());
    HIPCHK(hipStreamCreateWithFlags(&E->st, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&E->ev0));
    HIPCHK(hipEventCreate(&E->ev1));
    HIPCHK(hipMalloc((void **)&E->pipe, sizeof(Pipe)));
    void *jb = nullptr;
    static_assert(sizeof(JobHdr) <= JOBS_OFF, "job header too large");
    HIPCHK(hipMalloc(&jb, JOBS_OFF + MAXJOBS * sizeof(Job)));
    E->hdr = (JobHdr *)jb;
    E->jobs = (Job *)((char *)jb + JOBS_OFF);
    HIPCHK(hipHostMalloc(&E->stage_host, JOBS_OFF + MAXJOBS * sizeof(Job), hipHostMallocDefault));
    HIPCHK(hipMalloc((void **)&E->reqs, MAXSLOTS * sizeof(ReqDev)));
    static_assert(sizeof(Job) == 32, "Job is 32 bytes (SLOT_BYTES)");
    for (int k = 0; k < 2; k++) {
        StepCtx &c = E->sc[k];
        HIPCHK(hipMalloc((void **)&c.pipe, MAXB * sizeof(Pipe)));
        HIPCHK(hipMalloc((void **)&c.hdr, MAXB * SLOT_BYTES));
        HIPCHK(hipHostMalloc(&c.stage, MAXB * SLOT_BYTES, hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void **)&c.pipe_host, MAXB * sizeof(Pipe), hipHostMallocDefault));
        if (k == 0) c.st = E->st;
        else HIPCHK(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking));   // default priority
        HIPCHK(hipEventCreate(&c.ev0));
        HIPCHK(hipEventCreate(&c.ev1));
    }
    for (int b = 0; b < MAXB; b++) {
        HIPCHK(hipStreamCreateWithFlags(&E->bst[b], hipStreamNonBlocking));   // default priority
        HIPCHK(hipEventCreateWithFlags(&E->bjoin[b], hipEventDisableTiming));
    }
    HIPCHK(hipEventCreateWithFlags(&E->bfork, hipEventDisableTiming));
    HIPCHK(hipGetSymbolAddress((void **)&E->na, HIP_SYMBOL(scratch)));
    HIPCHK(hipHostMalloc((void **)&E->na_host, NA_WORDS * sizeof(u64), hipHostMallocDefault));
    HIPCHK(hipMemsetAsync(E->reqs, 0, MAXSLOTS * sizeof(ReqDev), E->st));
    HIPCHK(hipMemsetAsync(E->pipe, 0, sizeof(Pipe), E->st));
    HIPCHK(hipStreamSynchronize(E->st));
    for (int i = MAXSLOTS - 1; i >= 0; i--) E->free_slots.push_back(i);
    if (E->verbose) fprintf(stderr, "war_engine: device %d %s, %d CUs, %d MHz\n", device, E->arch.c_str(), E->cus, E->clock_mhz);
    g_eng = E.release();
    return 0;
}

int war_shutdown(void) {
    t_err.clear();
    if (!g_eng) return 0;
    war_serve_stop();
    Engine *E = g_eng;
    {
        std::lock_guard<std::mutex> g(E->mu);
        (void)hipSetDevice(E->device);
        (void)hipStreamSynchronize(E->st);
        for (auto &kv : E->graphs) (void)hipGraphExecDestroy(kv.second);
        pool_release(*E);
        for (int k = 0; k < 2; k++) {
            StepCtx &c = E->sc[k];
            (void)hipStreamSynchronize(c.st);
            for (auto &kv : c.graphs) (void)hipGraphExecDestroy(kv.second);
            (void)hipFree(c.pipe);
            (void)hipFree(c.hdr);
            (void)hipHostFree(c.stage);
            (void)hipHostFree(c.pipe_host);
            (void)hipEventDestroy(c.ev0);
            (void)hipEventDestroy(c.ev1);
            if (k) (void)hipStreamDestroy(c.st);
        }
        for (int b = 0; b < MAXB; b++) { (void)hipStreamDestroy(E->bst[b]); (void)hipEventDestroy(E->bjoin[b]); }
        (void)hipEventDestroy(E->bfork);
        (void)hipHostFree(E->na_host);
        for (QueueBufs *Q : {&E->qb4, &E->qb5}) {
            for (int i = 0; i < 2; i++) {
                (void)hipFree(Q->qa[i]); (void)hipFree(Q->qb[i]); (void)hipFree(Q->qreq[i]); (void)hipFree(Q->qpos[i]);
                (void)hipFree(Q->qhome[i]); (void)hipFree(Q->qcnt[i]);
            }
            (void)hipFree(Q->ca); (void)hipFree(Q->cb); (void)hipFree(Q->cmeta); (void)hipFree(Q->creq); (void)hipFree(Q->clam);
            (void)hipFree(Q->cpos); (void)hipFree(Q->ccnt);
        }
        (void)hipFree(E->pipe);
        (void)hipFree(E->hdr);
        (void)hipFree(E->reqs);
        (void)hipHostFree(E->stage_host);
        (void)hipEventDestroy(E->ev0);
        (void)hipEventDestroy(E->ev1);
        (void)hipStreamDestroy(E->st);
    }
    delete E;
    g_eng = nullptr;
    return 0;
}

int war_device_info(war_devinfo *out) {
    t_err.clear();
    if (!g_eng) return fail(60, "war_init has not been called");
    if (!out) return fail(1, "null output");
    memset(out, 0, sizeof *out);
    out->cus = g_eng->cus; out->max_clock_mhz = g_eng->clock_mhz; out->wave_size = 64;
    size_t fr = 0, tot = 0;
    HIPCHK(hipSetDevice(g_eng->device));      // the free memory of the engine's device
    HIPCHK(hipMemGetInfo(&fr, &tot));
    out->free_bytes = fr; out->total_bytes = tot;
    snprintf(out->arch, sizeof out->arch, "%s", g_eng->arch.c_str());
    return 0;
}

int war_set_tuning(const war_tuning *t) {
    t_err.clear();
    if (!g_eng) return fail(60, "war_init has not been called");
    if (!t) return fail(1, "null tuning");
    war_tuning probe = default_tuning(13, WAR_MODE_FULL, WAR_DET_BRENT);
    if (t->waves_per_cu > 0) probe.waves_per_cu = t->waves_per_cu;
    if (t->ranks_per_refill > 0) probe.ranks_per_refill = t->ranks_per_refill;
    if (t->brent_cap > 0) probe.brent_cap = t->brent_cap;
    if (t->plain_per > 0) probe.plain_per = t->plain_per;
    if (t->stage_kind > 0) probe.stage_kind = t->stage_kind;
    if (t->pile_layout > 0) probe.pile_layout = t->pile_layout;
    int rc = check_tuning(probe);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(g_eng->mu);
    g_eng->tun = *t;
    return 0;
}

int war_get_tuning(war_tuning *t) {
    t_err.clear();
    if (!g_eng) return fail(60, "war_init has not been called");
    if (!t) return fail(1, "null tuning");
    *t = g_eng->tun;
    return 0;
}

int war_tuning_for(int32_t n, int32_t mode, int32_t rule, war_tuning *out) {
    t_err.clear();
    if (!g_eng) return fail(60, "war_init has not been called");
    if (!out) return fail(1, "null tuning");
    war_request r{};
    r.n = n; r.mode = mode; r.rule = rule;
    std::lock_guard<std::mutex> g(g_eng->mu);
    resolve_tuning(*g_eng, r, *out);
    return 0;
}

int war_run_range(const war_request *req, war_result *res) {
    t_err.clear();
    if (!g_eng) return fail(60, "war_init has not been called");
    int rc = check_request(req);
    if (rc) return rc;
    if (!res) return fail(1, "null result");
    ReqHost h;
    h.rq = *req;
    h.res = res;
    const double t0 = now_ms();
    {
        std::lock_guard<std::mutex> g(g_eng->mu);
        drain(*g_eng);
        rc = run_request_locked(*g_eng, h);
    }
    res->latency_ms = now_ms() - t0;
    return rc;
}

// The same query as war_run_range with no device allocation and no graph: the counters, histograms,
// pipe counters, job list and request descriptor live in `scratch` (resolved at war_init), cleared
// by a fill kernel at every call; kernels are launched directly on the engine's stream; the job list
// and the results pass through pinned staging allocated at init; nothing on the host heap. Requests
// without cycle walks, on the plain kernel or the flow schedule, histograms of at most NA_BINS bins.
// Every call on a GPU uses that GPU's one array, and keeps its contents until the next call: the
// engine mutex serialises the calls of a process, so overlapping calls on one GPU would race only
// if something else wrote the array.
int war_run_range_noalloc(const war_request *req, war_result *res) {
    t_err.clear();
    if (!g_eng) return fail(60, "war_init has not been called");
    int rc = check_request(req);
    if (rc) return rc;
    if (!res) return fail(1, "null result");
    const war_request &r = *req;
    if (r.distinct || r.dbins || r.shape_blk)
        return fail(90, "war_run_range_noalloc: no cycle walks (distinct, dbins, shape need war_run_range)");
    if ((u32)r.hbins > NA_BINS || (u32)r.lbins > NA_BINS)
        return fail(90, "war_run_range_noalloc: at most %u histogram bins (hbins %d, lbins %d)", NA_BINS, r.hbins, r.lbins);
    Engine &E = *g_eng;
    const double t0 = now_ms();
    std::lock_guard<std::mutex> g(E.mu);
    drain(E);
    war_tuning t;
    resolve_tuning(E, r, t);
    big_canon_tile(E, r, t);
    rc = check_tuning(t);
    if (!rc) rc = check_cap(r, t);
    int W = 4;
    if (!rc) rc = layout_of(r, t, &W);
    if (rc) return rc;
    const bool flow = t.stage_kind == 3 && (r.rule == 0 || r.detector == WAR_DET_BRENT);
    if (r.kernel == WAR_KERNEL_PERSISTENT && !flow)
        return fail(90, "war_run_range_noalloc: the persistent kernel runs the flow schedule only (stage_kind 3)");
    u64 *na = E.na;
    Pipe *pipe = (Pipe *)(na + NA_PIPE);
    JobHdr *hdr = (JobHdr *)(na + NA_HDR);
    Job *job = (Job *)(na + NA_JOB);
    ReqDev *rd = (ReqDev *)(na + NA_REQ);
    // counters and histograms start at zero (the array keeps the last call's contents)
    const u32 ncnt = C_N + (u32)r.hbins + (u32)r.lbins;
    hipLaunchKernelGGL(k_fill, dim3(std::min(64u, (ncnt + 255u) / 256u)), dim3(256), 0, E.st, na + NA_CNT, 0ull, ncnt);
    ReqDev d{};
    d.n = (u32)r.n; d.mode = (u32)r.mode; d.rule = (u32)r.rule;
    d.hbins = (u32)r.hbins; d.lbins = (u32)r.lbins;
    d.cap = r.step_cap;
    d.cnt = na + NA_CNT; d.hist = na + NA_HIST; d.lenhist = na + NA_HIST + r.hbins;
    u64 *hs = E.na_host;                                  // pinned staging: [0, 16) request, [16, ..) job list
    memcpy(hs, &d, sizeof d);
    HIPCHK(hipMemcpyAsync(rd, hs, sizeof d, hipMemcpyHostToDevice, E.st));
    double dev = 0;
    int batches = 0;
    auto one = [&](u64 lo, u64 hi, const war_tuning &bt) -> int {
        // the previous batch's staging copy is complete: every batch ends with a synchronisation
        char *js = (char *)(hs + 16);
        JobHdr h{};
        const u64 tp = (u64)bt.ranks_per_refill;
        h.njobs = 1; h.req0 = 0; h.ntiles = (hi - lo + tp - 1) / tp;
        Job j{};
        j.lo = lo; j.hi = hi; j.tile_base = 0; j.req = 0;
        memcpy(js, &h, sizeof h);
        memcpy(js + JOBS_OFF, &j, sizeof j);
        HIPCHK(hipMemcpyAsync(hdr, js, JOBS_OFF + sizeof j, hipMemcpyHostToDevice, E.st));
        const bool blk = deal_blocks(r, W, lo, hi, bt);
        GraphKey key{W, r.rule, 1, bt.waves_per_cu, bt.ranks_per_refill / 64, r.step_cap ? 1 : 0, 0, 0, round_pow2(r.hbins),
                     round_pow2(r.lbins), 3, r.rule ? WAR_DET_BRENT : r.detector, (blk && W == 4) ? 1 : 0};
        if (W == 4 && r.rule == 0 && key.det == WAR_DET_BRENT) key.fw = fw_of(r.n, r.mode, r.step_cap != 0);
        HIPCHK(hipEventRecord(E.ev0, E.st));
        record_launch_any(E, key, E.st, pipe, hdr, job, rd);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(E.ev1, E.st));
        HIPCHK(hipStreamSynchronize(E.st));
        float m = 0;
        HIPCHK(hipEventElapsedTime(&m, E.ev0, E.ev1));
        dev += m;
        batches++;
        return 0;
    };
    if (r.kernel == WAR_KERNEL_PERSISTENT) {
        rc = for_each_persistent_batch(r, t, W, false, E.batch, one);
    } else {
        // the plain kernel, chunks of 2^26 positions as war_run_range runs them
        const u32 per = (u32)t.plain_per;
        for (u64 lo = r.lo; lo < r.hi && !rc; lo += 1ull << 26) {
            const u64 hi = std::min<u64>(r.hi, lo + (1ull << 26));
            HIPCHK(hipEventRecord(E.ev0, E.st));
            const u64 threads = (hi - lo + per - 1) / per, blocks = (threads + 255) / 256;
            const size_t lds = (size_t)(r.hbins + r.lbins) * sizeof(u32);
            auto launch = [&](auto W_, auto R_) {
                constexpr int WW = decltype(W_)::value, RR = decltype(R_)::value;
                PlainArgs<WW> P{};
                P.lo = lo; P.hi = hi; P.per = per; P.hb = (u32)r.hbins; P.lb = (u32)r.lbins; P.tile_pos = PLAIN_CYC_TILE;
                P.R = d; P.req = 0; P.pipe = pipe;
                launch_plain<WW, RR>(r.detector, r.step_cap != 0, (unsigned)blocks, lds, E.st, P);
            };
            using I4 = std::integral_constant<int, 4>;
            using I5 = std::integral_constant<int, 5>;
            using R0 = std::integral_constant<int, 0>;
            using R1 = std::integral_constant<int, 1>;
            if (W == 4) { if (r.rule) launch(I4{}, R1{}); else launch(I4{}, R0{}); }
            else        { if (r.rule) launch(I5{}, R1{}); else launch(I5{}, R0{}); }
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(E.ev1, E.st));
            HIPCHK(hipStreamSynchronize(E.st));
            float m = 0;
            HIPCHK(hipEventElapsedTime(&m, E.ev0, E.ev1));
            dev += m;
        }
    }
    if (rc) return rc;
    // results: counters and histograms through the pinned staging
    HIPCHK(hipMemcpyAsync(hs, na + NA_CNT, (size_t)ncnt * sizeof(u64), hipMemcpyDeviceToHost, E.st));
    HIPCHK(hipStreamSynchronize(E.st));
    const u64 *b = hs;
    res->cyc = b[C_CYC]; res->done = b[C_DONE]; res->maxfin = b[C_MAXFIN]; res->turns = b[C_TURNS];
    res->oncycle = 0; res->distinct_minkey = 0; res->distinct_hashset = 0;
    res->shape_nonstd = 0; res->shape_unchk = 0; res->undetermined = b[C_UNDET];
    if (res->hist) for (int i = 0; i < std::min(res->hbins, r.hbins); i++) res->hist[i] = b[C_N + i];
    if (res->lenhist) for (int i = 0; i < std::min(res->lbins, r.lbins); i++) res->lenhist[i] = b[C_N + r.hbins + i];
    res->device_ms = dev;
    res->latency_ms = now_ms() - t0;
    if (b[C_UNDET] && ((r.flags & WAR_FLAG_UNDET_ERROR) || (!r.step_cap && r.detector != WAR_DET_STREAK)))
        return fail(21, "%llu games not resolved within %u turns", b[C_UNDET],
                    r.step_cap ? r.step_cap : (r.detector == WAR_DET_STREAK ? WAR_STREAK_CAP : WAR_SAFETY_CAP));
    if (r.mode != WAR_MODE_HALF && b[C_DONE] != r.hi - r.lo)
        return fail(22, "played %llu positions, expected %llu", b[C_DONE], r.hi - r.lo);
    (void)batches;
    return 0;
}

int war_run_batch(const war_request *reqs, int32_t count, war_result *res, int32_t *launches) {
    t_err.clear();
    if (!g_eng) return fail(60, "war_init has not been called");
    if (count < 0 || (count > 0 && (!reqs || !res))) return fail(1, "null requests or results");
    int rc = 0;
    for (int i = 0; i < count && !rc; i++) {
        rc = check_request(&reqs[i]);
        if (!rc && reqs[i].kernel != WAR_KERNEL_PERSISTENT)
            rc = fail(82, "request %d: a batch runs the persistent kernel only", i);
        if (!rc && reqs[i].detector != WAR_DET_BRENT && reqs[i].detector != WAR_DET_ALT)
            rc = fail(82, "request %d: a batch runs Brent or the alternation detector", i);
    }
    if (rc) return rc;
    std::vector<ReqHost> hs((size_t)count);
    for (int i = 0; i < count; i++) { hs[(size_t)i].rq = reqs[i]; hs[(size_t)i].res = &res[i]; }
    int nl = 0;
    const double t0 = now_ms();
    {
        std::lock_guard<std::mutex> g(g_eng->mu);
        drain(*g_eng);
        if ((size_t)count > g_eng->free_slots.size())
            return fail(40, "%d requests, %zu free request slots", count, g_eng->free_slots.size());
        rc = run_batch_locked(*g_eng, hs, &nl);
    }
    const double wall = now_ms() - t0;
    for (int i = 0; i < count; i++) res[i].latency_ms = wall;
    if (launches) *launches = nl;
    return rc;
}

int war_reach(int32_t n, uint64_t lo, uint64_t hi, uint64_t cap, war_reach_result *out) {
    t_err.clear();
    if (!g_eng) return fail(60, "war_init has not been called");
    if (!out) return fail(1, "null result");
    if (n < 2 || n > 15) return fail(2, "reach needs 2 <= n <= 15 (got %d)", n);
    if (cap < 1) return fail(2, "cap must be >= 1");
    war_request r{};
    r.n = n; r.mode = WAR_MODE_CANON; r.rule = 0; r.kernel = WAR_KERNEL_PERSISTENT;
    r.lo = lo; r.hi = hi; r.distinct = 1;
    int rc = check_request(&r);
    if (rc) return rc;
    war_result res{};
    ReqHost h;
    h.rq = r; h.res = &res; h.reach = true; h.reach_cap = cap; h.rres = out;
    {
        std::lock_guard<std::mutex> g(g_eng->mu);
        drain(*g_eng);
        rc = run_request_locked(*g_eng, h);
    }
    return rc;
}

int war_cycle_info_batch(const war_pile_position *pos, int32_t count, int32_t rule, uint64_t cap,
                         war_cycle_info *out) {
    t_err.clear();
    if (!g_eng) return fail(60, "war_init has not been called");
    if (count < 0 || count > 65536) return fail(2, "count must be in [0, 65536]");
    if (count == 0) return 0;
    if (!pos || !out) return fail(1, "null positions or results");
    if (rule != 0 && rule != 1) return fail(4, "rule must be 0 or 1");
    if (cap < 1) return fail(2, "cap must be >= 1");
    std::vector<CInfoDev> H((size_t)count);
    for (int q = 0; q < count; q++) {
        const war_pile_position &P = pos[q];
        if (P.n1 < 1 || P.n2 < 1 || P.n1 + P.n2 > 255 || !P.p1 || !P.p2)
            return fail(2, "position %d: piles must be non-empty with at most 255 cards in all", q);
        bool seen[256] = {false};
        CInfoDev &d = H[(size_t)q];
        memset(&d, 0, sizeof d);
        d.l1 = (u32)P.n1; d.l2 = (u32)P.n2;
        for (int p = 0; p < 2; p++) {
            const int L = p ? P.n2 : P.n1;
            const uint8_t *c = p ? P.p2 : P.p1;
            u64 sm = 0, ws = 0;
            for (int i = 0; i < L; i++) {
                if (seen[c[i]]) return fail(2, "position %d: card %d appears twice", q, (int)c[i]);
                seen[c[i]] = true;
                d.start[p][i] = c[i];
                d.ring[p][i] = c[i]; d.ring[p + 2][i] = c[i];
                sm += c[i]; ws += (u64)c[i] * (u64)(i + 1);
            }
            d.len[p] = d.len[p + 2] = (u32)L;
            d.sum[p] = d.sum[p + 2] = sm;
            d.wsum[p] = d.wsum[p + 2] = ws;
        }
        d.power = 1; d.phase = CI_BRENT;
    }
    Engine &E = *g_eng;
    std::lock_guard<std::mutex> g(E.mu);
    drain(E);
    CInfoDev *D = nullptr;
    HIPCHK(hipMalloc((void **)&D, H.size() * sizeof(CInfoDev)));
    int rc = 0;
    do {
        if (hipMemcpy(D, H.data(), H.size() * sizeof(CInfoDev), hipMemcpyHostToDevice) != hipSuccess) {
            rc = fail(50, "copy to the device failed"); break;
        }
        const unsigned blocks = (unsigned)((count + CI_LANES - 1) / CI_LANES);
        for (;;) {
            hipLaunchKernelGGL(k_cycinfo, dim3(blocks), dim3(CI_LANES), 0, E.st, D, (u32)count, (u32)rule, (u64)cap,
                               (u64)1 << 22);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(E.st) != hipSuccess) {
                rc = fail(50, "cycle-info kernel failed"); break;
            }
            if (hipMemcpy(H.data(), D, H.size() * sizeof(CInfoDev), hipMemcpyDeviceToHost) != hipSuccess) {
                rc = fail(50, "copy from the device failed"); break;
            }
            bool all = true;
            for (const CInfoDev &d : H) all = all && d.phase == CI_DONE;
            if (all) break;
        }
    } while (false);
    (void)hipFree(D);
    if (rc) return rc;
    for (int q = 0; q < count; q++) {
        const CInfoDev &d = H[(size_t)q];
        war_cycle_info &o = out[q];
        o.status = (int32_t)d.status; o.pad = 0;
        o.turns = d.rounds;
        o.lam = d.status == 1 ? d.lam : 0;
        o.mu = d.status == 1 ? d.mu : 0;
    }
    return 0;
}

int war_autotune(int32_t n, int32_t mode, int32_t rule, double budget_ms, war_tuning *best) {
    t_err.clear();
    if (!g_eng) return fail(60, "war_init has not been called");
    war_request r{};
    r.n = n; r.mode = mode; r.rule = rule; r.kernel = WAR_KERNEL_PERSISTENT;
    if (n < 1 || n > 20 || mode < 0 || mode > 2) return fail(2, "autotune: n must be in [1, 20] and mode FULL, HALF or CANON");
    const u64 total = range_total(n, mode);
    // sample: a slice from the middle of the range, big enough to fill the GPU many times over
    const u64 want = std::min<u64>(total, 1ull << 26);
    r.lo = total / 2 - std::min(total / 2, want / 2);
    r.hi = r.lo + want;
    return war_autotune_batch(&r, 1, budget_ms, nullptr, best);
}

int war_autotune_batch(const war_request *reqs, int32_t count, double budget_ms, const war_result *expect,
                       war_tuning *best) {
    t_err.clear();
    if (!g_eng) return fail(60, "war_init has not been called");
    if (count < 1 || count > MAXSLOTS || !reqs) return fail(1, "autotune: 1 to %d requests", MAXSLOTS);
    Engine &E = *g_eng;
    for (int i = 0; i < count; i++) {
        int rc = check_request(&reqs[i]);
        if (!rc && reqs[i].kernel != WAR_KERNEL_PERSISTENT) rc = fail(82, "autotune: request %d is not for the persistent kernel", i);
        if (!rc && reqs[i].detector != WAR_DET_BRENT && reqs[i].detector != WAR_DET_ALT && count > 1)
            rc = fail(82, "autotune: a batch runs Brent or the alternation detector (request %d)", i);
        if (!rc && reqs[i].detector == WAR_DET_STREAK) rc = fail(82, "autotune: the streak test runs on the plain kernel");
        if (rc) return rc;
    }
    // result buffers: one set per request for the reference and one for the candidate
    struct Buf {
        war_result r{};
        std::vector<uint64_t> h, l, d;
        void init(const war_request &q) {
            h.assign((size_t)std::max(1, q.hbins), 0); l.assign((size_t)std::max(1, q.lbins), 0);
            d.assign((size_t)std::max(1, q.dbins), 0);
            r = war_result{};
            r.hist = h.data(); r.hbins = q.hbins; r.lenhist = l.data(); r.lbins = q.lbins; r.dhist = d.data(); r.dbins = q.dbins;
        }
    };
    std::vector<Buf> ref((size_t)count), cand((size_t)count);
    for (int i = 0; i < count; i++) { ref[(size_t)i].init(reqs[i]); cand[(size_t)i].init(reqs[i]); }
    // the counts every schedule must reproduce: the caller's (expect), else the plain kernel's, an
    // independent code path (one position per lane, the reference's structure)
    if (!expect) {
        std::lock_guard<std::mutex> g(E.mu);
        drain(E);
        for (int i = 0; i < count; i++) {
            ReqHost h;
            h.rq = reqs[i]; h.rq.kernel = WAR_KERNEL_PLAIN;
            if (h.rq.detector == WAR_DET_PAIR) h.rq.detector = WAR_DET_BRENT;   // no plain form; turns not compared
            h.rq.tuning = war_tuning{};
            h.res = &ref[(size_t)i].r;
            const int c = run_request_locked(E, h);
            if (c) return c;
        }
    }
    auto same_counts = [&](int i, const war_result &a) -> bool {
        const war_result &b = expect ? expect[i] : ref[(size_t)i].r;
        const war_request &q = reqs[i];
        if (a.cyc != b.cyc || a.done != b.done || a.maxfin != b.maxfin || a.undetermined != b.undetermined) return false;
        if (q.detector != WAR_DET_PAIR && a.turns != b.turns) return false;
        if (q.dbins || q.distinct) { if (a.oncycle != b.oncycle) return false; }
        if (q.distinct && a.distinct_minkey != b.distinct_minkey) return false;
        if (q.shape_blk && (a.shape_nonstd != b.shape_nonstd || a.shape_unchk != b.shape_unchk)) return false;
        for (int k = 0; k < q.hbins && b.hist; k++) if (a.hist[k] != b.hist[k]) return false;
        for (int k = 0; k < q.lbins && b.lenhist; k++) if (a.lenhist[k] != b.lenhist[k]) return false;
        for (int k = 0; k < q.dbins && b.dhist; k++) if (a.dhist[k] != b.dhist[k]) return false;
        return true;
    };
    // one schedule: every request with tuning t, as one batch (or one range); timed on its second
    // run (the first captures its graphs); *ms = device time
    auto trial = [&](const war_tuning &t, double *ms) -> int {
        double dev = 0;
        for (int rep = 0; rep < 2; rep++) {
            std::lock_guard<std::mutex> g(E.mu);
            drain(E);
            int c = 0;
            if (count == 1) {
                ReqHost h;
                h.rq = reqs[0]; h.rq.tuning = t; cand[0].init(reqs[0]); h.res = &cand[0].r;
                c = run_request_locked(E, h);
                dev = cand[0].r.device_ms;
            } else {
                std::vector<ReqHost> hs((size_t)count);
                for (int i = 0; i < count; i++) {
                    hs[(size_t)i].rq = reqs[i]; hs[(size_t)i].rq.tuning = t;
                    cand[(size_t)i].init(reqs[i]); hs[(size_t)i].res = &cand[(size_t)i].r;
                }
                if ((size_t)count > E.free_slots.size()) return fail(40, "autotune: %d requests, %zu free slots", count, E.free_slots.size());
                int nl = 0;
                c = run_batch_locked(E, hs, &nl, &dev);
            }
            if (c) return c;
        }
        for (int i = 0; i < count; i++)
            if (!same_counts(i, cand[(size_t)i].r))
                return fail(70, "autotune: schedule (waves %d, tile %d, cap %d, kind %d, layout %d) changed the counts of "
                                "request %d (exactness violated)", t.waves_per_cu, t.ranks_per_refill, t.brent_cap,
                            t.stage_kind, t.pile_layout, i);
        *ms = dev;
        return 0;
    };
    war_tuning cur;
    {
        std::lock_guard<std::mutex> g(E.mu);
        resolve_tuning(E, reqs[0], cur);
    }
    const double t_start = now_ms();
    double best_ms = 1e30;
    int rc = trial(cur, &best_ms);
    if (rc) return rc;
    // coordinate descent, two passes, over: the schedule (staged refill / lockstep / flow), the tile
    // (positions per grab: a multiple of the 64 lanes of a wave, so 64 is the smallest; 32 would
    // leave half a wave without a position), waves per CU (1..8 as the occupancy sweep, and the
    // larger counts the flow kernel prefers), the first stage boundary (staged schedules only: the
    // flow kernel has no stages) and, for n = 17, the pile layout (5-bit or card 16 held apart)
    constexpr double AUTOTUNE_MARGIN = 0.02;
    const int kinds[] = {3, 1, 2};
    const int tiles[] = {64, 128, 256, 512, 1024, 2048, 4096, 8192};
    const int wpcs[] = {1, 2, 3, 4, 5, 6, 7, 8, 12, 16, 20, 24, 28, 32};
    const int caps[] = {7, 15, 31, 63, 127};
    const int lays[] = {5, 6};
    bool held = count == 1 && held_ok(reqs[0]);
    bool any_cap = false, any_pair = false;
    for (int i = 0; i < count; i++) { any_cap |= reqs[i].step_cap != 0; any_pair |= reqs[i].detector == WAR_DET_PAIR; }
    int trials = 1;
    for (int pass = 0; pass < 2; pass++) {
        for (int dim = 0; dim < 5; dim++) {
            const int *vals = dim == 0 ? kinds : dim == 1 ? tiles : dim == 2 ? wpcs : dim == 3 ? caps : lays;
            const int nv = dim == 0 ? 3 : dim == 1 ? 8 : dim == 2 ? 14 : dim == 3 ? 5 : 2;
            if (dim == 0 && (any_cap || any_pair)) continue;    // explicit caps and the pair-run detector: flow only
            if (dim == 3 && cur.stage_kind == 3) continue;      // no stage boundaries in the flow kernel
            if (dim == 4 && !held) continue;
            war_tuning bestt = cur;
            for (int i = 0; i < nv; i++) {
                if (now_ms() - t_start > budget_ms) break;
                war_tuning t = cur;
                if (dim == 0) t.stage_kind = vals[i];
                else if (dim == 1) t.ranks_per_refill = vals[i];
                else if (dim == 2) t.waves_per_cu = vals[i];
                else if (dim == 3) t.brent_cap = vals[i];
                else t.pile_layout = vals[i];
                if (memcmp(&t, &cur, sizeof t) == 0) continue;
                double ms = 0;
                rc = trial(t, &ms);
                if (rc) return rc;
                trials++;
                if (E.verbose)
                    fprintf(stderr, "autotune n=%d mode=%d: kind=%d tile=%d wpc=%d cap=%d layout=%d: %.3f ms\n", reqs[0].n,
                            reqs[0].mode, t.stage_kind, t.ranks_per_refill, t.waves_per_cu, t.brent_cap, t.pile_layout, ms);
                // a schedule replaces the best so far only if at least AUTOTUNE_MARGIN faster on the
                // sample: one timing per candidate, so noise alone does not move the choice away from
                // the default the descent starts from
                if (ms < best_ms * (1.0 - AUTOTUNE_MARGIN)) { best_ms = ms; bestt = t; }
            }
            cur = bestt;
        }
    }
    if (count == 1) {
        std::lock_guard<std::mutex> g(E.mu);
        drain(E);
        const war_request &q = reqs[0];
        E.tuned[std::make_tuple(q.n, q.mode, q.rule, q.detector)] = cur;
    }
    if (E.verbose) fprintf(stderr, "autotune: %d schedules in %.0f ms, best %.3f ms\n", trials, now_ms() - t_start, best_ms);
    if (best) *best = cur;
    return 0;
}

int war_serve_start(int32_t max_slots, uint64_t step_positions) {
    t_err.clear();
    if (!g_eng) return fail(60, "war_init has not been called");
    Engine &E = *g_eng;
    if (E.serving) return 0;
    if (max_slots < 1 || max_slots > MAXSLOTS - 8) return fail(80, "max_slots must be in [1, %d]", MAXSLOTS - 8);
    E.serve_slots = max_slots;
    if (step_positions) E.step_positions = std::min<u64>(step_positions, E.batch);
    E.serving = true;
    E.worker = std::thread(serve_loop, &E);
    return 0;
}

int war_serve_stop(void) {
    t_err.clear();
    if (!g_eng) return 0;
    Engine &E = *g_eng;
    if (!E.serving) return 0;
    {
        std::lock_guard<std::mutex> lk(E.smu);
        E.serving = false;
        E.scv.notify_all();
    }
    if (E.worker.joinable()) E.worker.join();
    return 0;
}

int war_submit(const war_request *req, war_result *res, uint64_t *id) {
    t_err.clear();
    if (!g_eng) return fail(60, "war_init has not been called");
    Engine &E = *g_eng;
    if (!E.serving) return fail(81, "serving loop not started (war_serve_start)");
    int rc = check_request(req);
    if (rc) return rc;
    if (req->kernel != WAR_KERNEL_PERSISTENT) return fail(82, "the serving loop runs the persistent kernel only");
    if (req->detector == WAR_DET_PAIR)
        return fail(82, "the pair-run detector runs on single requests (war_run_range), not in the serving loop");
    if (!res || !id) return fail(1, "null result or id");
    auto h = std::make_shared<ReqHost>();
    h->rq = *req;
    h->res = res;
    h->t_submit = now_ms();
    std::lock_guard<std::mutex> lk(E.smu);
    h->id = E.next_id++;
    *id = h->id;
    E.all[h->id] = h;
    E.pending.push_back(h);
    E.scv.notify_all();
    return 0;
}

int war_poll(uint64_t id, int32_t *done) {
    t_err.clear();
    if (!g_eng) return fail(60, "war_init has not been called");
    Engine &E = *g_eng;
    std::lock_guard<std::mutex> lk(E.smu);
    auto it = E.all.find(id);
    if (it == E.all.end()) return fail(83, "unknown request id %llu", (unsigned long long)id);
    if (done) *done = it->second->finished ? 1 : 0;
    return 0;
}

int war_wait(uint64_t id) {
    t_err.clear();
    if (!g_eng) return fail(60, "war_init has not been called");
    Engine &E = *g_eng;
    std::shared_ptr<ReqHost> h;
    {
        std::unique_lock<std::mutex> lk(E.smu);
        auto it = E.all.find(id);
        if (it == E.all.end()) return fail(83, "unknown request id %llu", (unsigned long long)id);
        h = it->second;
        E.scv.wait(lk, [&] { return h->finished; });
        E.all.erase(id);
    }
    if (h->status) return fail(h->status, "%s", h->err.c_str());
    return 0;
}

}  // extern "C"
