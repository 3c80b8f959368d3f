/* war_engine.h -- C ABI of the MI355X (gfx950) War enumeration engine, libwar_engine.so.
 *
 * The engine counts, over a range of deals or canonical positions of single-suit War, the games
 * that never end, exactly as the reference's GPU counter does (gpu/war_gpu.py), and adds a
 * persistent compacting kernel, an autotuner and a continuous-batching request loop.
 *
 * Plain pointers and fixed-width integers only; no framework types cross this boundary. Every
 * entry point returns 0 on success and a non-zero status otherwise (war_last_error() then gives
 * the reason). The library never aborts the process.
 *
 * The reference has no FFI of its own; the calls below replace the host driver of
 * gpu/war_gpu.py and its kernel launches:
 *   war_run_range   <- run() (gpu/war_gpu.py:272-352): counters, batching, launches, sync, and the
 *                      war_count / war_canon kernels (gpu/war_gpu.py:121-176, :183-259)
 *   war_result      <- the 7-tuple run() returns (gpu/war_gpu.py:346-352), plus lenhist/turns
 *   war_request.mode   FULL / HALF / CANON <- run(half=, canon=) (gpu/war_gpu.py:272-290)
 *   war_request.hbins  <- --hist (gpu/war_gpu.py:364), dbins <- --distinct (:369-370,379)
 *   war_request.shape_blk <- --shape / WAR_BLK_OVERRIDE (gpu/war_gpu.py:290-292)
 *   war_request.rule   <- --rule (gpu/war_gpu.py:367-368)
 *   5-bit piles for n >= 17 <- gpu/war24.py:31-47 (layout only)
 */
#ifndef WAR_ENGINE_H
#define WAR_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WAR_ABI_VERSION 3

/* mode */
#define WAR_MODE_FULL  0   /* deal ranks r in [0, n!), deck positions 0,2,4.. to player 1     */
#define WAR_MODE_HALF  1   /* as FULL, but only deals giving card n-1 to player 1 are played
                              (even n); counts are NOT doubled here (run() doubles them)       */
#define WAR_MODE_CANON 2   /* positions p = r*(n-1)+k: player 1 = [n-1, q0..q(k-1)],
                              player 2 = [qk..q(n-2)], r = lexicographic rank of q            */

/* kernel */
#define WAR_KERNEL_PLAIN      0  /* one position per lane, per-lane Brent (reference structure) */
#define WAR_KERNEL_PERSISTENT 1  /* persistent, age-synchronous stages with survivor compaction */

/* detector (cycle test) */
#define WAR_DET_BRENT  0   /* Brent (gpu/war_gpu.py:77-96): every output, any n and rule         */
#define WAR_DET_STREAK 1   /* the odd-n streak test of gpu/war24.py:98-125 (Spivey Thm 6):
                              odd n, rule 0, plain kernel; counts, game lengths and turns, but
                              no cycle lengths (hbins, dbins, distinct, shape must be 0)        */
#define WAR_DET_ALT    2   /* Brent with alternation checkpoints (this engine's; odd n, rule 0,
                              either kernel): the tortoise also jumps, at ages that are multiples
                              of 8, after a repeated winner (odd-n cycles alternate, Spivey
                              Thm 6), and after waiting 128, 256, ... turns. Every output is
                              exact and equals Brent's; only `turns` (its own) is smaller.     */
#define WAR_DET_PAIR   3   /* Brent with pair-run checkpoints (this engine's; n = 2 mod 4, rule 0,
                              the persistent kernel on single requests): as WAR_DET_ALT, but the
                              pattern jump follows a winner run of one turn or of three or more
                              (the cycles of such n are, all but a few, runs of exactly two: the
                              loop-shape test of gpu/war_gpu.py:51-73), and only while the game is
                              at most 128 turns old. Every output is exact and equals Brent's;
                              only `turns` (its own) is smaller.                               */

typedef struct war_tuning {
    int32_t waves_per_cu;      /* persistent grid = CUs x waves_per_cu wave64s (1..32)          */
    int32_t ranks_per_refill;  /* tile: positions per grab / slots per queue region (64..8192)  */
    int32_t brent_cap;         /* first stage boundary 2^k-1 (1..65535); later stages double    */
    int32_t plain_per;         /* plain kernel: consecutive positions per lane (1..4096)        */
    int32_t stage_kind;        /* persistent stages after the first: 1 = refill (a lane that
                                  finishes takes the next record at once), 2 = lockstep;
                                  3 = flow: no stages, one kernel per batch in which a lane whose
                                  game is over takes the next position at the start of the next
                                  span of 16 turns (32 for Brent over blocks of 24 deals; single
                                  requests and mixed batches, with or without cycle walks; rule 1
                                  with a detector other than Brent runs as 2); 0 = default (3) */
    int32_t pile_layout;       /* 0 = default; 4 = 4-bit piles (n <= 16); 5 = 5-bit piles (any n);
                                  6 = n = 17 with card 16 held apart and 4-bit piles (the layout
                                  of gpu/war17.py; persistent kernel, Brent, no cycle walks).
                                  The default is 6 where it applies, else 4 for n <= 16, else 5 */
} war_tuning;

/* war_request.flags */
#define WAR_FLAG_UNDET_ERROR 1u   /* fail the request (status 21) if any game is undetermined, instead
                                     of only counting it in war_result.undetermined               */
#define WAR_FLAG_CYCLE_TABLE 2u   /* CANON, rule 0, n <= 14, with cycle walks (distinct / dbins / shape):
                                     keep the cycle-state table whatever the range (by default only
                                     ranges of at least 2^28 positions keep it). The table holds every
                                     state of the cycles met, so the walks look a record up instead of
                                     walking its cycle; the outputs are the same                   */
#define WAR_FLAG_NO_CYCLE_TABLE 4u  /* walk every cycle record (no cycle-state table)              */

/* war_request.step_cap = 0: the default cap. The streak test stops a game after 2^22 turns, as the
 * reference's CAP (gpu/war24.py:57,103); Brent and the checkpoint detectors are exact and end every
 * valid game, and a game still running after 2^24 turns (a corrupted state) is an error (status 21). */
#define WAR_STREAK_CAP (1u << 22)
#define WAR_SAFETY_CAP (1u << 24)

typedef struct war_request {
    int32_t  n;            /* cards: 1..16 (4-bit piles), 17..20 (5-bit piles)                 */
    int32_t  mode;         /* WAR_MODE_*                                                         */
    int32_t  rule;         /* 0: winner's card first (Spivey); 1: loser's card first            */
    int32_t  kernel;       /* WAR_KERNEL_*                                                       */
    uint64_t lo, hi;       /* half-open range: deal ranks (FULL/HALF) or positions (CANON)      */
    int32_t  hbins;        /* cycle-length histogram bins, last = overflow (0 = off, <= 4096)   */
    int32_t  lbins;        /* ending-game-length histogram bins, last = overflow (0 = off)      */
    int32_t  dbins;        /* distinct weight bins (CANON, n <= 16; 0 = off, <= 4096)           */
    int32_t  distinct;     /* CANON, n <= 15: min-key count + device hash set of cycle keys    */
    int32_t  shape_blk;    /* CANON: Theorem-3 run length to test (0 = off)                      */
    int32_t  detector;     /* WAR_DET_* (0 = Brent)                                              */
    war_tuning tuning;     /* zeros = engine defaults                                            */
    uint32_t step_cap;     /* 0 = the default cap (above). Otherwise a game that its detector has
                              not resolved (ended or seen repeating) within step_cap turns counts
                              as undetermined, with step_cap turns in `turns`; the streak test keeps
                              the reference's form (still running with more than step_cap turns
                              played at the top of its loop: gpu/war24.py:103), so step_cap + 1
                              turns. <= 2^24; the persistent kernel needs a multiple of 32 and the
                              flow schedule (stage_kind 3)                                       */
    uint32_t flags;        /* WAR_FLAG_*                                                         */
} war_request;

typedef struct war_result {
    uint64_t cyc;              /* endless deals / cycling positions                              */
    uint64_t done;             /* positions played (must equal the range length; HALF: taken)   */
    uint64_t maxfin;           /* longest ending game, in turns                                  */
    uint64_t turns;            /* sum of the detector's rounds (game length, or turns until the
                                  repeat is seen)                                                */
    uint64_t oncycle;          /* distinct/dbins: positions lying on their own cycle             */
    uint64_t distinct_minkey;  /* distinct: positions whose key is their cycle's minimum key    */
    uint64_t distinct_hashset; /* distinct: different cycle-minimum keys met (device hash set)  */
    uint64_t shape_nonstd;     /* shape: cycles not made of winner runs of exactly shape_blk    */
    uint64_t shape_unchk;      /* shape: cycles longer than 512 turns (not checked)             */
    uint64_t undetermined;     /* games not resolved within the step cap (war_request.step_cap):
                                  counted in done, not in cyc, maxfin or the histograms; the
                                  reference's war24 counter prints it (gpu/war24.py:150-152,203-205) */
    uint64_t *hist;    int32_t hbins;   /* caller-owned, hbins entries (or NULL)               */
    uint64_t *lenhist; int32_t lbins;   /* caller-owned, lbins entries (or NULL)               */
    uint64_t *dhist;   int32_t dbins;   /* caller-owned, dbins entries (or NULL)               */
    double   device_ms;        /* device time of the request's kernels (HIP events)             */
    double   latency_ms;       /* serving: submit -> completion, host clock                     */
} war_result;

/* war_reach: the distinct cycles (up to swapping the players) met by the canonical positions of a
 * range, each taken once at its position of minimum key, and whether an ordinary deal enters it
 * (cpu/reach_cycles.c): a backward search from the cycle for a position of ceil(n/2) and
 * floor(n/2) cards. */
typedef struct war_reach_result {
    uint64_t classes;          /* distinct cycles (positions at their cycle's minimum key)       */
    uint64_t reached;          /* entered by a deal                                               */
    uint64_t not_reached;
    uint64_t undetermined;     /* the search visited more than `cap` positions (or its stack of
                                  64 entries per search was full)                                */
    uint64_t *nolen; int32_t nbins;  /* caller-owned: cycle lengths of the cycles not reached
                                        (last bin = overflow), or NULL                           */
    double   device_ms;
    uint64_t max_depth;        /* the most entries any search's stack held (at most 64; a full
                                  stack is counted in undetermined)                              */
} war_reach_result;

/* war_cycle_info: one position of up to 255 cards (cpu/cycle_length.c). */
typedef struct war_pile_position {
    int32_t n1, n2;            /* cards of player 1 and player 2: >= 1 each, n1 + n2 <= 255     */
    const uint8_t *p1, *p2;    /* card values, top card first; distinct                          */
} war_pile_position;

typedef struct war_cycle_info {
    int32_t  status;           /* 0: the game ends; 1: it cycles; 2: not settled within cap turns */
    int32_t  pad;
    uint64_t turns;            /* 0: the game's length; 1: the turn Brent saw the repeat; 2: cap */
    uint64_t lam;              /* 1: cycle length                                                */
    uint64_t mu;               /* 1: turns played before the game enters its cycle               */
} war_cycle_info;

typedef struct war_devinfo {
    int32_t  cus;              /* compute units                                                  */
    int32_t  max_clock_mhz;
    int32_t  wave_size;
    int32_t  reserved;
    uint64_t free_bytes, total_bytes;
    char     arch[64];
} war_devinfo;

/* Lifecycle. war_init selects the device, creates the engine's stream and device buffers.
 * flags: bit 0 = verbose to stderr. */
int war_init(int32_t device, uint32_t flags);
int war_shutdown(void);
const char *war_last_error(void);
int war_abi_version(void);
int war_device_info(war_devinfo *out);

/* Synchronous range query: plays every position of [lo, hi) on the GPU and fills *res. */
int war_run_range(const war_request *req, war_result *res);

/* war_run_range with no device allocation: no counterpart in the reference. For callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * same counts as war_run_range, from a static device array of the library (about 67 KB per GPU,
 * held until the process ends; its address is resolved at war_init) instead of the per-request
 * allocation, with kernels launched directly (no graph) and results through pinned staging set up at
 * war_init. Limits: no cycle walks (distinct, dbins, shape), at most 4096 bins per histogram, the
 * plain kernel or the flow schedule. Calls are serialised per process; calls that overlapped on one
 * GPU would share the array and race. */
int war_run_range_noalloc(const war_request *req, war_result *res);

/* Several requests of mixed n, mode and rule in shared persistent launches, synchronously: the
 * reference launches war_count / war_canon once per range (gpu/war_gpu.py:309-328); here requests
 * of one pile layout, rule, detector and cycle-walk need share each kernel launch (one, when their
 * positions fit a batch: 2^32 positions on the flow kernel, 2^28 with cycle walks), and all the
 * kernel launches of the call (up to 16) are branches of ONE captured hipGraph, launched once;
 * every result equals the request run alone. Persistent kernel, Brent or the alternation detector;
 * count <= the free request slots (248 less the serving loop's). res[i] receives request i's
 * result, device_ms the device time of the call's graph and latency_ms the whole call; *launches
 * (may be null) the number of graph launches (1 unless the call needs more than 16 kernel
 * launches or a staged schedule, which run one graph per kernel launch). */
int war_run_batch(const war_request *reqs, int32_t count, war_result *res, int32_t *launches);

/* Reachability of the distinct cycles of canonical positions [lo, hi) of n cards (2 <= n <= 15,
 * rule 0) by ordinary deals: replaces cpu/reach_cycles.c's main loop (:73-107) and search
 * (:48-69). cap: positions one cycle's backward search may visit (the reference's default is
 * 200000000). Synchronous. */
int war_reach(int32_t n, uint64_t lo, uint64_t hi, uint64_t cap, war_reach_result *out);

/* Cycle length and pre-period of `count` positions (one GPU lane each): replaces
 * cpu/cycle_length.c:114-126 (Brent, then two copies lam apart stepped together). rule 0 or 1.
 * cap bounds Brent's turns per position (the reference's default is 10^12). Synchronous; runs as
 * a series of launches of at most 2^22 turns per lane. count <= 65536. */
int war_cycle_info_batch(const war_pile_position *pos, int32_t count, int32_t rule, uint64_t cap,
                         war_cycle_info *out);

/* Engine-wide tuning override. A request's non-zero tuning fields win; then this override's
 * non-zero fields; then war_autotune's choice for the request's (n, mode, rule); then a built-in
 * default per (n, mode). war_get_tuning returns the override (zeros = none). */
int war_set_tuning(const war_tuning *t);
int war_get_tuning(war_tuning *t);
/* The tuning a request of (n, mode, rule) with zero tuning fields would run with. */
int war_tuning_for(int32_t n, int32_t mode, int32_t rule, war_tuning *out);

/* Autotuner (replaces the reference's fixed launch shapes, gpu/war_gpu.py:272-280,310-317): times
 * candidate schedules of the persistent kernel on a 2^26-position slice of (n, mode, rule) with
 * Brent, and keeps the fastest as the default for that workload. Coordinate descent, two passes,
 * over the schedule (flow / staged refill / staged lockstep), the tile (64..8192 positions per grab;
 * 64 = one position per lane of a wave is the smallest), waves per CU (1..8, 12..32), the first
 * stage boundary (staged schedules) and, for n = 17, the pile layout. Every candidate must
 * reproduce the plain kernel's counts on the slice exactly (an independent code path), else the
 * call fails. budget_ms bounds the search; *best receives the choice. */
int war_autotune(int32_t n, int32_t mode, int32_t rule, double budget_ms, war_tuning *best);

/* The same search over a given set of requests (count >= 1), run as one war_run_batch (count > 1)
 * or one war_run_range (count = 1) per candidate with the candidate as every request's tuning.
 * expect (may be null): the counts each request must reproduce (cyc, done, maxfin, undetermined,
 * turns, oncycle / minkey / shape where asked, and every histogram the request asks for); null =
 * the plain kernel's counts. A single request's choice is kept as the default for its
 * (n, mode, rule, detector); a batch's is returned only (pass it as the requests' tuning). */
int war_autotune_batch(const war_request *reqs, int32_t count, double budget_ms, const war_result *expect,
                       war_tuning *best);

/* Continuous-batching request loop. Requests submitted while others run are admitted into free
 * slots (bounded by slots and by free device memory) and share the persistent launches.
 * war_submit copies *req; results are retrieved once with war_wait (which blocks) after war_poll
 * reports done. The result's hist/lenhist/dhist pointers must stay valid until war_wait returns. */
int war_serve_start(int32_t max_slots, uint64_t step_positions);
int war_serve_stop(void);
int war_submit(const war_request *req, war_result *res, uint64_t *id);
int war_poll(uint64_t id, int32_t *done);
int war_wait(uint64_t id);

#ifdef __cplusplus
}
#endif
#endif /* WAR_ENGINE_H */
