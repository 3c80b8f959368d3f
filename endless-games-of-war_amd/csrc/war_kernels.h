// Kernels of the War engine (gfx950, wave64).
//
//  k_setup          zeroes the pipeline counters of one batch (a kernel, not a memset node, so the
//                   batch graph can be replayed).
//  k_plain          one position per lane at a time, per-lane Brent, `per` consecutive positions
//                   per lane; the structure of the reference's war_count / war_canon
//                   (gpu/war_gpu.py:121-176, :183-259).
//  k_stage<FRESH>   the persistent, age-synchronous form. Every lane of a wave starts its game at
//                   the same age, so Brent's tortoise resets (ages 2^k - 1) and every game length
//                   or cycle length seen at a turn are wave-uniform: the per-turn work is the turn
//                   plus two compares, and all accounting is scalar. A stage plays ages
//                   (a0, a1] with a1 = 2^k - 1; lanes still running at a1 are compacted into the
//                   next stage's queue (the tortoise equals the hare there, so a record is just the
//                   state). The FRESH stage enumerates positions from the job list (a global tile
//                   counter refills waves); later stages drain queues (a global record counter).
//                   The last stage has no cap (Brent's exactness), only a 2^24-turn safety stop
//                   that counts as "undetermined" (an error for the default cap).
//  k_walk           per cycling position of a --distinct / --dbins / --shape request: the
//                   Theorem-3 run test from the meeting state (gpu/war_gpu.py:51-73), and the walk
//                   of lam turns from the start that decides whether it lies on its cycle, then the
//                   weight (gpu/war_gpu.py:224-234), the minimum key (cpu/distinct_cycles.c:57-80)
//                   and the device hash-set insert of that key.
#pragma once
#include "war_device.h"

namespace war {

enum {
    C_CYC = 0, C_DONE, C_MAXFIN, C_TURNS, C_ONCYCLE, C_MINKEY, C_NONSTD, C_UNCHK, C_UNDET,
    C_HASH, C_HOVF, C_QOVF, C_RYES, C_RNO, C_RUND, C_RDEPTH, C_SOVF, C_N = 17
};

constexpr u32 AGE_CAP = 1u << 24;      // safety stop of the last stage (never reached by a valid game)
constexpr u32 UNBOUNDED = 0xFFFFFFFFu;
constexpr int MAXSTAGE = 16;

struct ReqDev {
    u32 n, mode, rule, hbins, lbins, dbins, distinct, shape_blk;
    u32 walk, hlog2, reach;
    u32 cap;            // the request's step cap (war_request.step_cap); 0 = the default
    u64 *cnt, *hist, *lenhist, *dhist, *htab;
    u64 reach_cap;      // reach: positions a cycle's backward search may visit
    u64 *nolen;         // reach: cycle lengths of the cycles no deal enters (4096 bins)
    u64 *stab;          // cycle walks: the cycle-state table (2^slog2 slots), or null
    u32 slog2, pad2;
};

struct Job { u64 lo, hi, tile_base; u32 req, pad; };
// Per batch: the job list and the number of tiles T.
struct JobHdr {
    u32 njobs, req0;        // req0: the request of job 0
    u64 ntiles;
    u64 pad;
};

struct Pipe {
    u64 grab[MAXSTAGE];     // grab counter of stage s
    u64 stotal[MAXSTAGE];   // records written by stage s (sizes the next stage's grabs)
    u64 nreg[MAXSTAGE];     // output regions allocated by refill stage s
    u64 wgrab, ctotal;      // walk grab counter (second pass); cycle records written
    u64 ovf;                // output regions refused for lack of room (an error; never with the sizing)
    u64 wgrab0;             // walk grab counter, first pass (the cycle-state table's inserts)
    u64 pad[4];
};

// Survivor records live in per-tile regions of tile_pos slots: tile t's records are
// a[t*tile_pos + 0 .. cnt[t]). A wave that takes a group of tiles writes its survivors densely into
// the same group's span of the other buffer, so no append needs an atomic. The pile lengths and n
// ride in the unused top bits of the packed piles (4-bit: an alive pile holds <= 15 cards = 60 bits;
// 5-bit: <= 19 cards = 95 bits): a record is 16 B (4-bit) or 32 B (5-bit). req / pos / home are
// written only when the batch needs them.
template <int W> struct Queue {
    typedef typename Lay<W>::T T;
    T *a, *b;
    u32 *req;
    u64 *pos;
    u32 *home;   // the tile the position was enumerated in (cycle records go to its region)
    u32 *cnt;    // records per tile
    T *ta, *tb;  // alternation detector: the tortoise (packed like the state)
    u32 *meta;   // alternation detector: age - tortoise age (bits 0-23), log2(wait / ALT_L0) (24-28),
                 // player 1 won the last turn (29)
};

// Layout 6 (n = 17 only, so n is not stored): an alive holder's word has <= 15 cards, so its top
// nibble takes the word's length; when that length is >= 1 the other player holds <= 15 cards and
// the other word's top nibble takes the depth of card n-1 (length 0 means depth 0).
template <int W> __device__ __forceinline__ void pack(const St<W> &s, u32 n, typename Lay<W>::T &qa, typename Lay<W>::T &qb) {
    typedef typename Lay<W>::T T;
    if constexpr (W == 6) {
        (void)n;
        qa = s.a | ((u64)(s.sa >> 2) << 60);
        qb = s.b | ((u64)(s.pb >> 2) << 60);
    } else if constexpr (W == 4) {
        qa = s.a | ((u64)(s.sa >> 2) << 60);
        qb = s.b | ((u64)(n - 1) << 60);
    } else {
        qa = s.a | ((T)(s.sa | (n << 8)) << 96);
        qb = s.b;
    }
}
template <int W> __device__ __forceinline__ void unpack(typename Lay<W>::T qa, typename Lay<W>::T qb, St<W> &s, u32 &n) {
    typedef typename Lay<W>::T T;
    if constexpr (W == 6) {
        const u32 lh = (u32)(qa >> 60);
        s.a = qa & 0x0FFFFFFFFFFFFFFFull; s.sa = lh << 2;
        s.b = lh ? (qb & 0x0FFFFFFFFFFFFFFFull) : qb;
        s.pb = lh ? (u32)(qb >> 60) << 2 : 0u;
        n = HELD_N;
    } else if constexpr (W == 4) {
        s.a = qa & 0x0FFFFFFFFFFFFFFFull; s.sa = (u32)(qa >> 60) << 2;
        s.b = qb & 0x0FFFFFFFFFFFFFFFull; n = (u32)(qb >> 60) + 1u;
    } else {
        const u32 meta = (u32)(qa >> 96);
        s.a = qa & lowmask_<T>(96); s.sa = meta & 0xFFu; n = meta >> 8;
        s.b = qb;
    }
}

// Cycle records (requests that need walks), in the region of their home tile.
template <int W> struct CycQueue {
    typedef typename Lay<W>::T T;
    T *a, *b;    // meeting state
    u32 *meta;   // sa | n << 8
    u32 *req;
    u32 *lam;
    u64 *pos;    // position index (the start is rebuilt from it)
    u32 *cnt;    // records per home tile (zeroed per batch)
};

template <int W> struct StageArgs {
    Pipe *pipe;
    const JobHdr *hdr;
    const Job *jobs;
    const ReqDev *reqs;
    Queue<W> qin, qout;
    CycQueue<W> cq;
    u32 tile_pos, per;     // positions per tile (= slots per region), FRESH positions per lane
    u32 a0, a1;            // ages played: (a0, a1]
    int stage;
    u32 hb, lb;            // LDS histogram bins per wave
    u64 max_regions;       // regions of tile_pos slots each queue buffer holds
    u32 in_refill;         // the input regions were written by a refill stage (count: nreg[stage-1])
    u32 zero;              // always 0 (see Zh)
};

// ------------------------------------------------------------------------------------------
__device__ __forceinline__ u32 lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
__device__ __forceinline__ u64 ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ u32 rank_in(u64 mask) {   // set bits of mask below this lane
    return __builtin_amdgcn_mbcnt_hi((u32)(mask >> 32), __builtin_amdgcn_mbcnt_lo((u32)mask, 0u));
}
// this lane is in mask m (the mask itself is the condition: no vector work to test a lane's bit)
__device__ __forceinline__ bool in_mask(u64 m) { return __builtin_amdgcn_inverse_ballot_w64(m); }
__device__ __forceinline__ u32 uni(u32 x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ u64 uni64(u64 x) {
    return ((u64)__builtin_amdgcn_readfirstlane((u32)(x >> 32)) << 32) | (u64)__builtin_amdgcn_readfirstlane((u32)x);
}
__device__ __forceinline__ u32 lanev(u32 x, u32 l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ u64 lanev64(u64 x, u32 l) {
    return ((u64)__builtin_amdgcn_readlane((u32)(x >> 32), l) << 32) | (u64)__builtin_amdgcn_readlane((u32)x, l);
}
// one atomicAdd per wave; every lane gets the old value
__device__ __forceinline__ u64 wave_fetch_add(u64 *p, u64 v) {
    u64 old = 0;
    if (lane_id() == 0) old = atomicAdd((unsigned long long *)p, (unsigned long long)v);
    return uni64(old);
}

// Zero the pipeline counters and the per-tile cycle counts of one batch (grid-stride).
__global__ void k_setup(Pipe *pipe, const JobHdr *hdr, u32 *ccnt) {
    const u32 tid = blockIdx.x * blockDim.x + threadIdx.x, nth = gridDim.x * blockDim.x;
    u64 *p = reinterpret_cast<u64 *>(pipe);
    for (u32 i = tid; i < sizeof(Pipe) / 8; i += nth) p[i] = 0;
    if (ccnt) {
        const u64 T = hdr->ntiles;
        for (u64 i = tid; i < T; i += nth) ccnt[i] = 0;
    }
}

// Inclusive prefix sum over the 64 lanes.
__device__ __forceinline__ u32 wave_scan(u32 v) {
    const u32 lane = lane_id();
    for (int off = 1; off < 64; off <<= 1) {
        const u32 t = __shfl_up(v, off);
        if (lane >= (u32)off) v += t;
    }
    return v;
}

// Tiles per grab so that a grab holds about 256 records (a power of two, at most 64).
__device__ __forceinline__ u32 group_size(u64 records, u64 tiles, u32 tile_pos) {
    u32 g = 1;
    while (g < 64 && (u64)g * records < 256ull * tiles) g <<= 1;
    (void)tile_pos;
    return g;
}

// Position k of a group's record list -> (tile i within the group, offset in its region).
// excl holds the exclusive prefix of the counts, lane j for tile j (j < ng). Must be called with all
// lanes of the wave active (cross-lane reads).
__device__ __forceinline__ void group_locate(u32 excl, u32 ng, u32 k, u32 &tile, u32 &off) {
    u32 lo = 0, elo = __shfl(excl, 0);
    if (ng > 1) {
        for (u32 step = 32; step; step >>= 1) {
            const u32 m = lo + step;
            const u32 e = __shfl(excl, (int)(m < 63u ? m : 63u));
            if (m < ng && e <= k) { lo = m; elo = e; }
        }
    }
    tile = lo;
    off = k - elo;
}

// ------------------------------------------------------------------------------------------
// Scalar (wave-uniform) accounting of one request, with wave-private LDS histograms. The counts
// are 32-bit (one scalar register each: the flow kernel's span loop has none to spare), as no
// launch holds 2^32 positions (a staged batch at most 2^28, a flow launch less than 2^32).
struct WaveAcc {
    u32 req;                   // request slot these totals belong to (UNBOUNDED = none)
    u32 hb, lb;                // that request's bins
    u32 cyc, done, maxfin, undet;
    u64 turns;
    u64 pushed;                // cycle records this wave wrote (all requests)
    u32 *lh, *ll;              // LDS: hist[hbmax], lenhist[lbmax] of this wave
    bool dirty;

    __device__ void init(u32 *lds, u32 hbmax, u32 lbmax) {
        req = UNBOUNDED; hb = lb = 0; cyc = done = turns = maxfin = undet = 0; dirty = false; pushed = 0;
        lh = lds; ll = lds + hbmax;
        for (u32 i = lane_id(); i < hbmax + lbmax; i += 64) lds[i] = 0;
    }
    // Push totals to the request's global counters. Every lane of the wave must call this.
    __device__ void flush(const ReqDev *reqs) {
        if (req != UNBOUNDED && dirty) {
            const ReqDev &R = reqs[req];
            if (lane_id() == 0) {
                if (cyc) atomicAdd((unsigned long long *)&R.cnt[C_CYC], (unsigned long long)cyc);
                if (done) atomicAdd((unsigned long long *)&R.cnt[C_DONE], (unsigned long long)done);
                if (turns) atomicAdd((unsigned long long *)&R.cnt[C_TURNS], (unsigned long long)turns);
                if (undet) atomicAdd((unsigned long long *)&R.cnt[C_UNDET], (unsigned long long)undet);
                atomicMax((unsigned long long *)&R.cnt[C_MAXFIN], (unsigned long long)maxfin);
            }
            for (u32 i = lane_id(); i < hb; i += 64) {
                u32 v = lh[i];
                if (v) { atomicAdd((unsigned long long *)&R.hist[i], (unsigned long long)v); lh[i] = 0; }
            }
            for (u32 i = lane_id(); i < lb; i += 64) {
                u32 v = ll[i];
                if (v) { atomicAdd((unsigned long long *)&R.lenhist[i], (unsigned long long)v); ll[i] = 0; }
            }
        }
        cyc = done = turns = maxfin = undet = 0; dirty = false;
    }
    __device__ void select(u32 r, const ReqDev *reqs) {
        if (r != req) {
            flush(reqs);
            req = r; hb = reqs[r].hbins; lb = reqs[r].lbins;
        }
    }
    __device__ void ended(u32 c, u32 age) {
        done += c; turns += (u64)c * age; maxfin = age > maxfin ? age : maxfin; dirty = true;
        if (lb && lane_id() == 0) atomicAdd(&ll[age < lb - 1 ? age : lb - 1], c);
    }
    __device__ void cycled(u32 c, u32 age, u32 lam) {
        cyc += c; done += c; turns += (u64)c * age; dirty = true;
        if (hb && lane_id() == 0) atomicAdd(&lh[lam < hb - 1 ? lam : hb - 1], c);
    }
    // Lanes with per-lane values: em ended at age `fin`, cm repeated at age `fin` with cycle length
    // `lam`. Every lane of the wave calls this (wave reductions, then one LDS add per lane).
    __device__ void lanes(u64 em, u64 cm, u32 fin, u32 lam) {
        const u32 lane = lane_id();
        const bool ie = (em >> lane) & 1ull, ic = (cm >> lane) & 1ull;
        u32 sum = (ie || ic) ? fin : 0u, mx = ie ? fin : 0u;
        for (int off = 32; off > 0; off >>= 1) {
            sum += (u32)__shfl_xor((int)sum, off);
            const u32 o = (u32)__shfl_xor((int)mx, off);
            mx = o > mx ? o : mx;
        }
        const u32 ne = (u32)__popcll(em), nc = (u32)__popcll(cm);
        done += ne + nc; cyc += nc; turns += sum; dirty = true;
        maxfin = mx > maxfin ? mx : maxfin;
        if (lb && ie) atomicAdd(&ll[fin < lb - 1 ? fin : lb - 1], 1u);
        if (hb && ic) atomicAdd(&lh[lam < hb - 1 ? lam : hb - 1], 1u);
    }
};

// Append the lanes of `mask` to the cycle records of their home tiles (one atomic per home tile
// present in the mask; the counters are spread over the tiles, so they do not contend).
template <int W>
__device__ __forceinline__ u32 push_cycles(const CycQueue<W> &cq, u32 tile_pos, u64 mask, const St<W> &s, u32 n,
                                           u32 req, u32 lam, u64 pos, u32 home) {
    const u32 lane = lane_id();
    const u32 total = (u32)__popcll(mask);
    while (mask) {
        const u32 leader = (u32)(__ffsll((long long)mask) - 1);
        const u32 h = lanev(home, leader);
        const u64 m = ballot(home == h) & mask;
        mask &= ~m;
        u32 base = 0;
        if (lane == leader) base = atomicAdd(&cq.cnt[h], (u32)__popcll(m));
        base = lanev(base, leader);
        if ((m >> lane) & 1ull) {
            const u32 slot = base + rank_in(m);
            if (slot < tile_pos) {
                const u64 i = (u64)h * tile_pos + slot;
                cq.a[i] = s.a; cq.b[i] = s.b; cq.meta[i] = s.sa | (n << 8);
                cq.req[i] = req; cq.lam[i] = lam; cq.pos[i] = pos;
            }
        }
    }
    return total;
}

// ------------------------------------------------------------------------------------------
// The lockstep game of one wave for ages (a0, a1]: every alive lane starts at age a0 with its
// tortoise equal to its state (a0 = 0 or 2^k - 1, where Brent resets). Ended / repeated lanes are
// accounted (wave-uniform values: the age is the game length, age - R the cycle length); returns
// the lanes still alive at a1 (or, for the last stage, at the safety stop).
struct LaneCtx {
    u32 req, home;
    u64 pos;
    bool walk;
};

// FLAGS: bit 0 = the batch holds several requests (records carry req), bit 1 = some request needs
// cycle walks (records carry pos and home), bit 2 = the alternation detector (records carry the
// tortoise). Compile-time, so the record loads are unconditional.
enum { F_MULTI = 1, F_WALK = 2, F_ALT = 4 };

// The alternation detector (odd n, rule 0): Brent's tortoise also jumps at ages that are multiples
// of ALT_K when one of the last ALT_K turns of odd age repeated the previous turn's winner (in a
// cycle of odd-n War the players alternate: Spivey Thm 6, gpu/war24.py:9-17), and after waiting
// ALT_L0 turns (then twice as long, ...). Before the first turn, player 1 counts as the last
// winner. Only turns of odd age are checked: that halves the scalar work of the check, which is
// the scalar unit's busiest part of the turn loop, and costs 0.5 % more turns (n = 13). Repeats
// are only declared on exact state matches, and the first match after a jump is the cycle length,
// so every output is exact; games are seen to repeat in about 30 % fewer turns.
constexpr u32 ALT_K = 8, ALT_L0 = 128;
// flow kernel: checkpoints per span (finished games are accounted and replaced once per span)
constexpr u32 FLOW_BLOCKS = 2;

template <int W, int RULE, bool FRESH, int FLAGS>
__device__ __forceinline__ u64 play(const StageArgs<W> &A, WaveAcc &acc, St<W> &s, u32 C, u32 n, u64 alive,
                                    bool uniform_req, const LaneCtx &L) {
    const ReqDev *reqs = A.reqs;
    const bool last = (A.a1 == UNBOUNDED);
    open_sa<W>(s, A.zero);
    u32 sb = C - s.sa;          // player 2's append offset, carried from turn to turn
    St<W> t = s;
    u32 R = A.a0;
    u32 age = A.a0;
    const u32 a1 = last ? AGE_CAP : A.a1;
    Zh<W> z = make_zh<W>(A.zero);
    // Brent windows (R, 2R + 1]: the tortoise stays put inside one. The inner loop is the turn, the
    // end test and one compare of player 1's pile with the tortoise's (a candidate repeat); it leaves
    // on any candidate or end, and the other two fields are compared only then.
    for (;;) {
        const u32 wend = (2u * R + 1u) < a1 ? (2u * R + 1u) : a1;
        while (age < wend) {
            u64 em, dm;
            do {
                ++age;
                u64 over;
                turn_sb<W, RULE>(s, sb, C, z, over);
                em = over & alive;
                dm = ballot(s.a == t.a) & alive;
            } while (!(em | dm) && age < wend);
            dm &= ~em;
            if (dm) dm &= ballot(rest_same<W>(s, t));
            if (em | dm) {
                const u32 lam = age - R;
                if (uniform_req) {
                    if (em) acc.ended(__popcll(em), age);
                    if (dm) acc.cycled(__popcll(dm), age, lam);
                } else {
                    u64 rem = em | dm;
                    while (rem) {
                        const u32 r = lanev(L.req, (u32)(__ffsll((long long)rem) - 1));
                        const u64 mr = ballot(L.req == r) & rem;
                        rem &= ~mr;
                        acc.select(r, reqs);
                        if (mr & em) acc.ended(__popcll(mr & em), age);
                        if (mr & dm) acc.cycled(__popcll(mr & dm), age, lam);
                    }
                }
                if ((FLAGS & F_WALK) && dm) {
                    const u64 wm = ballot(L.walk) & dm;
                    if (wm) acc.pushed += push_cycles<W>(A.cq, A.tile_pos, wm, s, n, L.req, lam, L.pos, L.home);
                }
                alive &= ~(em | dm);
                if (!alive) break;
            }
        }
        if (!alive || age == a1) break;
        t = s; R = age;                                      // Brent: tortoise := hare at 2^k - 1
    }
    if (alive && last) {
        // safety stop reached: counted as undetermined (an error for the caller)
        u64 rem = alive;
        while (rem) {
            const u32 r = lanev(L.req, (u32)(__ffsll((long long)rem) - 1));
            const u64 mr = ballot(L.req == r) & rem;
            rem &= ~mr;
            acc.select(r, reqs);
            acc.undet += __popcll(mr); acc.done += __popcll(mr); acc.dirty = true;
        }
        alive = 0;
    }
    return alive;
}

// The lanes whose player 2 pile and length (and layout 6's depth of card n-1) equal y's: the rest
// of the repeat test, as one mask (rest_same per lane, without turning the result into a value).
template <int W> __device__ __forceinline__ u64 rest_mask(const St<W> &x, const St<W> &y) {
    u64 m = ballot(x.b == y.b) & ballot(x.sa == y.sa);
    if constexpr (W == 6) m &= ballot(x.pb == y.pb);
    return m;
}

// Per-lane state of the alternation detector beside the game state.
template <int W> struct AltLane {
    St<W> t;         // tortoise
    u32 R, L;        // tortoise age, wait before the next timed jump
};

// The lockstep game of one wave for ages (a0, a1] under the alternation detector. a0 and a1 are
// multiples of ALT_K, so the checkpoints fall at the same ages for every lane. prevm: lanes whose
// previous turn player 1 won (before the first turn, player 1 counts as the last winner). Cycle
// lengths differ per lane (age - R); the accounting groups the lanes of an event by (request,
// cycle length).
template <int W, int RULE, bool FRESH, int FLAGS>
__device__ __forceinline__ u64 play_alt(const StageArgs<W> &A, WaveAcc &acc, St<W> &s, AltLane<W> &T_, u64 &prevm,
                                        u32 C, u32 n, u64 alive, bool uniform_req, const LaneCtx &L) {
    const ReqDev *reqs = A.reqs;
    const bool last = (A.a1 == UNBOUNDED);
    open_sa<W>(s, A.zero);
    open_sa<W>(T_.t, A.zero);
    u32 sb = C - s.sa;                     // player 2's append offset, carried from turn to turn
    u32 age = A.a0;
    const u32 a1 = last ? AGE_CAP : A.a1;
    Zh<W> z = make_zh<W>(A.zero);
    u32 due = T_.R + T_.L;                 // per lane: the age of the next timed jump
    u64 streakm = 0;
    // Finished lanes are recorded as they finish (their age, and which masks they are in) and
    // accounted once, when the wave leaves the stage: detections are spread over many ages under
    // this detector, so the accounting works on per-lane values (wave sums, one LDS add per lane)
    // instead of one scalar update per distinct age.
    u64 endm = 0, cycm = 0;
    u32 fin = 0;                           // per lane: the age its game ended or repeated
    auto settle = [&]() {
        const u32 lam = fin - T_.R;
        if (uniform_req) {
            acc.lanes(endm, cycm, fin, lam);
        } else {
            u64 rem = endm | cycm;
            while (rem) {
                const u32 r = lanev(L.req, (u32)(__ffsll((long long)rem) - 1));
                const u64 mr = ballot(L.req == r) & rem;
                rem &= ~mr;
                acc.select(r, reqs);
                acc.lanes(endm & mr, cycm & mr, fin, lam);
            }
        }
        endm = cycm = 0;
    };
#define WAR_ALT_TURN()                                                                     \
    {                                                                                      \
        ++age;                                                                             \
        u64 over;                                                                          \
        const u64 pm = ballot(turn_sb<W, RULE>(s, sb, C, z, over));                        \
        const u64 em = over & alive;                                                       \
        u64 dm = ballot(s.a == T_.t.a) & alive;                                            \
        if ((q & 1u) == 0u) streakm |= ~(pm ^ prevm);      /* odd age: a0 is a multiple of ALT_K */ \
        prevm = pm;                                                                        \
        if (em | dm) {                                                                     \
            dm &= ~em;                                                                     \
            if (dm) dm &= ballot(rest_same<W>(s, T_.t));                                   \
            const u64 ev = em | dm;                                                        \
            if (((ev >> lane_id()) & 1ull) != 0ull) fin = age;                             \
            endm |= em; cycm |= dm; alive &= ~ev;                                          \
            if ((FLAGS & F_WALK) && dm) {                                                  \
                const u64 wm = ballot(L.walk) & dm;                                        \
                if (wm) acc.pushed += push_cycles<W>(A.cq, A.tile_pos, wm, s, n, L.req,    \
                                                     age - T_.R, L.pos, L.home);           \
            }                                                                              \
        }                                                                                  \
    }
    for (;;) {
        // one block of ALT_K turns up to the next checkpoint (ages a0, a1 are multiples of ALT_K)
#pragma unroll
        for (u32 q = 0; q < ALT_K; q++) WAR_ALT_TURN();
        if (!alive) break;
        // checkpoint: timed jumps (the wait doubles), then jumps after a repeated winner
        const u64 ex = ballot(age >= due) & alive;
        const u64 js = (ex | streakm) & alive;
        streakm = 0;
        if (js) {
            if ((js >> lane_id()) & 1ull) {
                T_.t = s; T_.R = age;
                if ((ex >> lane_id()) & 1ull) T_.L <<= 1;
                due = age + T_.L;
            }
        }
        if (age == a1) break;
    }
#undef WAR_ALT_TURN
    if (endm | cycm) settle();
    if (alive && last) {
        u64 rem = alive;
        while (rem) {
            const u32 r = lanev(L.req, (u32)(__ffsll((long long)rem) - 1));
            const u64 mr = ballot(L.req == r) & rem;
            rem &= ~mr;
            acc.select(r, reqs);
            acc.undet += __popcll(mr); acc.done += __popcll(mr); acc.dirty = true;
        }
        alive = 0;
    }
    return alive;
}

// Write the survivors (lanes of `alive`) densely at span + out_n.
template <int W, int FLAGS>
__device__ __forceinline__ void put_survivors(const StageArgs<W> &A, u64 span, u32 out_n, u64 alive, const St<W> &s,
                                              u32 n, const LaneCtx &L, const AltLane<W> *T_ = nullptr, u64 prevm = 0) {
    typedef typename Lay<W>::T T;
    if ((alive >> lane_id()) & 1ull) {
        const u64 i = span + out_n + rank_in(alive);
        T qa, qb;
        pack<W>(s, n, qa, qb);
        A.qout.a[i] = qa; A.qout.b[i] = qb;
        if (FLAGS & F_MULTI) A.qout.req[i] = L.req;
        if (FLAGS & F_WALK) { A.qout.pos[i] = L.pos; A.qout.home[i] = L.home; }
        if constexpr ((FLAGS & F_ALT) != 0) {
            T ta, tb;
            pack<W>(T_->t, n, ta, tb);
            A.qout.ta[i] = ta; A.qout.tb[i] = tb;
            const u32 e = (u32)(__builtin_ctz(T_->L) - __builtin_ctz(ALT_L0));
            A.qout.meta[i] = (A.a1 - T_->R) | (e << 24) | ((u32)((prevm >> lane_id()) & 1ull) << 29);
        }
    }
}

// ------------------------------------------------------------------------------------------
// The age-synchronous stage. FRESH: a wave takes one tile of the job list at a time (lane l plays
// positions l*per .. l*per+per-1 of the tile, one per round). Otherwise: a wave takes a group of
// tiles of the previous stage's regions and plays their records 64 at a time. Survivors of a
// tile / group go densely into the same tiles' regions of the other buffer.
template <int W, int RULE, bool FRESH, int FLAGS>
__global__ __launch_bounds__(256) void k_stage(StageArgs<W> A) {
    typedef typename Lay<W>::T T;
    extern __shared__ u32 lds[];
    const u32 wid = threadIdx.x >> 6;
    const u32 lane = lane_id();
    WaveAcc acc;
    acc.init(lds + wid * (A.hb + A.lb), A.hb, A.lb);
    const ReqDev *reqs = A.reqs;
    const bool last = (A.a1 == UNBOUNDED);
    const u64 T_ = (!FRESH && A.in_refill) ? A.pipe->nreg[A.stage - 1] : A.hdr->ntiles;
    const u32 req0 = A.hdr->req0;
    const u32 TP = A.tile_pos;
    // tiles per grab: enough for ~256 records, from what the previous stage wrote in this batch
    u32 G = 1;
    if constexpr (!FRESH) {
        const u64 recs = A.pipe->stotal[A.stage - 1];
        if (recs == 0) { acc.flush(reqs); return; }
        G = group_size(recs, T_, TP);
    }
    u64 *grab = &A.pipe->grab[A.stage];
    // the next grab is requested one round ahead, so its latency hides behind a round of play
    u64 nxt = 0;
    if (lane == 0) nxt = atomicAdd((unsigned long long *)grab, 1ull);

    for (;;) {
        const u64 g0 = uni64(nxt) * G;
        if (g0 >= T_) break;
        if (lane == 0) nxt = atomicAdd((unsigned long long *)grab, 1ull);
        const u64 span = g0 * TP;           // this tile / group's slots in the output buffer
        u32 out_n = 0;                      // survivors written so far (wave-uniform)

        if constexpr (FRESH) {
            // the job holding tile g0 (jobs are few; binary search on tile_base)
            u32 lo = 0, hi = A.hdr->njobs;
            while (hi - lo > 1) { u32 mid = (lo + hi) >> 1; if (A.jobs[mid].tile_base <= g0) lo = mid; else hi = mid; }
            const Job J = A.jobs[lo];
            const ReqDev &Rq = reqs[J.req];
            const u32 n = Rq.n, mode = Rq.mode;
            const u32 C = c_of<W>(n);
            constexpr int PW = Lay<W>::PW;         // the permutation's bits per card
            const u32 per = A.per;
            acc.select(J.req, reqs);
            LaneCtx L;
            L.req = J.req; L.home = (u32)g0; L.walk = Rq.walk != 0;
            const u64 first = J.lo + (g0 - J.tile_base) * (u64)TP + (u64)lane * per;
            const u32 cnt = first < J.hi ? (u32)((J.hi - first) < per ? (J.hi - first) : per) : 0u;
            L.pos = first;
            typename Lay<W>::P q = 0;
            int k = 0;
            if (cnt) {
                if (mode == 2) { const u32 m = n - 1; const u64 r = first / m; k = (int)(first - r * m); q = unrank<PW>(r, (int)m); }
                else q = unrank<PW>(first, (int)n);
            }
            for (u32 j = 0; j < per; j++) {
                if (ballot(j < cnt) == 0) break;
                bool act = j < cnt;
                St<W> s = {};
                if (act) {
                    s = (mode == 2) ? canon<W>(q, (int)n, k) : deal<W>(q, (int)n);
                    if (mode == 1) act = top_card_to_p1<PW>(q, (int)n);
                }
                const u64 act_m = ballot(act);
                // a position that is over before it starts: rounds 0 (gpu/war_gpu.py:80)
                const u64 e0 = ballot(act && (n < 2 || ended<W>(s, C))) & act_m;
                if (e0) acc.ended(__popcll(e0), 0);
                u64 alive = act_m & ~e0;
                if constexpr ((FLAGS & F_ALT) != 0) {
                    AltLane<W> al;
                    al.t = s; al.R = 0; al.L = ALT_L0;
                    // before the first turn, the position's player 1 counts as the last winner
                    u64 prevm = ballot(p1_is_dealt_p1<W>(s, mode));
                    if (alive) alive = play_alt<W, RULE, FRESH, FLAGS>(A, acc, s, al, prevm, C, n, alive, true, L);
                    if (alive) {
                        put_survivors<W, FLAGS>(A, span, out_n, alive, s, n, L, &al, prevm);
                        out_n += __popcll(alive);
                    }
                } else {
                    if (alive) alive = play<W, RULE, FRESH, FLAGS>(A, acc, s, C, n, alive, true, L);
                    if (alive) {
                        put_survivors<W, FLAGS>(A, span, out_n, alive, s, n, L);
                        out_n += __popcll(alive);
                    }
                }
                if (j + 1 < cnt) {
                    if (mode == 2) { if (++k == (int)n - 1) { k = 0; q = next_perm<PW>(q, (int)n - 1); } }
                    else q = next_perm<PW>(q, (int)n);
                }
                L.pos++;
            }
            if (!last && lane == 0) A.qout.cnt[g0] = out_n;
        } else {
            // the records of tiles g0 .. g0+ng-1
            const u32 ng = (u32)((T_ - g0) < G ? (T_ - g0) : G);
            const u32 c = lane < ng ? A.qin.cnt[g0 + lane] : 0u;
            const u32 incl = wave_scan(c);
            const u32 excl = incl - c;
            const u32 total = lanev(incl, 63);
            // records of the next batch are loaded one batch ahead (their latency hides behind play)
            T na = 0, nb = 0, nta = 0, ntb = 0;
            u32 nreq = req0, nhome = 0, nmeta = 0;
            u64 npos = 0;
            // called with every lane active: group_locate reads other lanes (ds_bpermute), and a
            // read from an inactive lane does not return its value
            auto fetch = [&](u32 kk) {
                u32 ti, off;
                group_locate(excl, ng, kk < total ? kk : total - 1u, ti, off);
                if (kk < total && off < TP) {
                    const u64 i = (g0 + ti) * (u64)TP + off;
                    na = A.qin.a[i]; nb = A.qin.b[i];
                    if constexpr ((FLAGS & F_MULTI) != 0) nreq = A.qin.req[i];
                    if constexpr ((FLAGS & F_WALK) != 0) { npos = A.qin.pos[i]; nhome = A.qin.home[i]; }
                    if constexpr ((FLAGS & F_ALT) != 0) { nta = A.qin.ta[i]; ntb = A.qin.tb[i]; nmeta = A.qin.meta[i]; }
                }
            };
            fetch(lane);
            for (u32 b = 0; b < total; b += 64) {
                const u32 kk = b + lane;
                const bool act = kk < total;
                St<W> s = {};
                u32 n = 2;
                LaneCtx L;
                L.req = req0; L.pos = 0; L.home = 0; L.walk = false;
                AltLane<W> al;
                al.t = St<W>{}; al.R = 0; al.L = ALT_L0;
                u32 prevbit = 0;
                if (act) {
                    unpack<W>(na, nb, s, n);
                    L.req = nreq; L.pos = npos; L.home = nhome;
                    if constexpr ((FLAGS & F_ALT) != 0) {
                        u32 tn;
                        unpack<W>(nta, ntb, al.t, tn);
                        al.R = A.a0 - (nmeta & 0xFFFFFFu);
                        al.L = ALT_L0 << ((nmeta >> 24) & 31u);
                        prevbit = (nmeta >> 29) & 1u;
                    }
                }
                if (b + 64 < total) fetch(b + 64 + lane);
                const u32 C = c_of<W>(n);
                u64 alive = ballot(act);
                bool uniform_req = true;
                if constexpr ((FLAGS & F_MULTI) != 0) {
                    const u32 r0 = lanev(L.req, 0);
                    uniform_req = (ballot(L.req == r0) & alive) == alive;
                    if (uniform_req) acc.select(r0, reqs);
                } else {
                    acc.select(req0, reqs);
                }
                if constexpr ((FLAGS & F_WALK) != 0) L.walk = act && reqs[L.req].walk != 0;
                if constexpr ((FLAGS & F_ALT) != 0) {
                    u64 prevm = ballot(prevbit != 0u) & alive;
                    alive = play_alt<W, RULE, FRESH, FLAGS>(A, acc, s, al, prevm, C, n, alive, uniform_req, L);
                    if (alive) {
                        put_survivors<W, FLAGS>(A, span, out_n, alive, s, n, L, &al, prevm);
                        out_n += __popcll(alive);
                    }
                } else {
                    alive = play<W, RULE, FRESH, FLAGS>(A, acc, s, C, n, alive, uniform_req, L);
                    if (alive) {
                        put_survivors<W, FLAGS>(A, span, out_n, alive, s, n, L);
                        out_n += __popcll(alive);
                    }
                }
            }
            if (!last && lane < ng) {
                const u32 before = lane * TP;
                A.qout.cnt[g0 + lane] = out_n > before ? (out_n - before < TP ? out_n - before : TP) : 0u;
            }
        }
        if (!last && out_n && lane == 0) atomicAdd((unsigned long long *)&A.pipe->stotal[A.stage], (unsigned long long)out_n);
    }
    acc.flush(reqs);
    if (acc.pushed && lane == 0) atomicAdd((unsigned long long *)&A.pipe->ctotal, (unsigned long long)acc.pushed);
}

// ------------------------------------------------------------------------------------------
// Refill stage: ages (a0, a1] with a0 = 2^k - 1 and a1 = 2^(k+1) - 1. Brent's tortoise does not move
// inside such a window (its next reset is exactly at a1), so a lane that finishes can take another
// record of the same stage at once: the lane keeps its own turn count j (age = a0 + j, cycle length
// = j), and nothing else about Brent's schedule changes. Records come from the previous stage's
// regions through a 64-record buffer (loaded one buffer ahead, handed to idle lanes by cross-lane
// reads); survivors (j reaches a1 - a0) are appended to output regions the wave allocates one at a
// time (one atomic per region of tile_pos slots). In a multi-request batch a wave holds records of
// one request at a time (it drains before switching), so all accounting stays per wave.
template <int W>
__device__ __forceinline__ typename Lay<W>::T shfl_T(typename Lay<W>::T v, u32 src) {
    typedef typename Lay<W>::T T;
    if constexpr (W == 4) {
        return ((u64)(u32)__shfl((int)(u32)(v >> 32), (int)src) << 32) | (u64)(u32)__shfl((int)(u32)v, (int)src);
    } else {
        T r = 0;
        for (int q = 3; q >= 0; q--) r = (r << 32) | (T)(u32)__shfl((int)(u32)(v >> (32 * q)), (int)src);
        return r;
    }
}

constexpr u32 REFILL_IDLE = 24;    // refill once this many lanes of 64 are idle

template <int W, int RULE, int FLAGS>
__global__ __launch_bounds__(256) void k_refill(StageArgs<W> A) {
    typedef typename Lay<W>::T T;
    extern __shared__ u32 lds[];
    const u32 wid = threadIdx.x >> 6;
    const u32 lane = lane_id();
    WaveAcc acc;
    acc.init(lds + wid * (A.hb + A.lb), A.hb, A.lb);
    const ReqDev *reqs = A.reqs;
    const bool last = (A.a1 == UNBOUNDED);
    const u32 TP = A.tile_pos;
    const u32 req0 = A.hdr->req0;
    const u64 T_in = A.stage == 1 ? A.hdr->ntiles : A.pipe->nreg[A.stage - 1];
    const u64 recs = A.pipe->stotal[A.stage - 1];
    if (recs == 0) return;
    const u32 G = group_size(recs, T_in, TP);
    const u32 LEN = last ? AGE_CAP - A.a0 : A.a1 - A.a0;
    const u32 a0 = A.a0;
    u64 *grab = &A.pipe->grab[A.stage];
    u64 nxt = 0;
    if (lane == 0) nxt = atomicAdd((unsigned long long *)grab, 1ull);

    // ---- input stream (wave-uniform cursor over groups of regions)
    u64 g0 = 0;
    u32 ng = 0, excl = 0, total = 0, cursor = 0;
    bool stream_done = false;
    // next buffer (being loaded) and current buffer (being handed out)
    T xa = 0, xb = 0; u32 xreq = req0, xhome = 0; u64 xpos = 0; u32 x_n = 0;
    T ca = 0, cb = 0; u32 creq = req0, chome = 0; u64 cpos = 0; u32 c_n = 0, c_i = 0;
    auto load_next = [&]() {        // fill the next buffer with up to 64 records (all lanes)
        x_n = 0;
        while (cursor >= total) {
            if (stream_done) return;
            g0 = uni64(nxt) * G;
            if (g0 >= T_in) { stream_done = true; return; }
            if (lane == 0) nxt = atomicAdd((unsigned long long *)grab, 1ull);
            ng = (u32)((T_in - g0) < G ? (T_in - g0) : G);
            u32 c = lane < ng ? A.qin.cnt[g0 + lane] : 0u;
            c = c < TP ? c : TP;
            const u32 incl = wave_scan(c);
            excl = incl - c;
            total = lanev(incl, 63);
            cursor = 0;
        }
        const u32 kk = cursor + lane;
        u32 ti, off;
        group_locate(excl, ng, kk < total ? kk : total - 1u, ti, off);
        if (kk < total && off < TP) {
            const u64 i = (g0 + ti) * (u64)TP + off;
            xa = A.qin.a[i]; xb = A.qin.b[i];
            if constexpr ((FLAGS & F_MULTI) != 0) xreq = A.qin.req[i];
            if constexpr ((FLAGS & F_WALK) != 0) { xpos = A.qin.pos[i]; xhome = A.qin.home[i]; }
        }
        x_n = (total - cursor) < 64u ? (total - cursor) : 64u;
        cursor += x_n;
    };
    load_next();

    // ---- lanes (a finished lane keeps its final state, frozen, until the next refill accounts it)
    St<W> s, t;
    s.a = 0; s.b = 0; s.sa = 0; t = s;
    u32 n = 2, C = 0, j = 0;
    bool live = false;
    LaneCtx L;
    L.req = req0; L.pos = 0; L.home = 0; L.walk = false;
    u64 pending = 0;    // lanes that finished and are not accounted yet
    u32 cur_req = req0;
    acc.select(req0, reqs);
    u64 lturns = 0;     // per lane: sum of game lengths / detection rounds
    u32 lmax = 0;       // per lane: longest ending game
    // ---- output region
    u32 och = UNBOUNDED, ofill = 0;

    auto close_region = [&]() {
        if (och != UNBOUNDED && lane == 0) {
            A.qout.cnt[och] = ofill;
            if (ofill) atomicAdd((unsigned long long *)&A.pipe->stotal[A.stage], (unsigned long long)ofill);
        }
    };
    auto reduce_lanes = [&]() {     // per-lane totals into the wave accumulator
        u64 tsum = lturns;
        u32 tmax = lmax;
        for (int off = 32; off > 0; off >>= 1) {
            tsum += __shfl_xor(tsum, off);
            const u32 o = __shfl_xor(tmax, off);
            tmax = o > tmax ? o : tmax;
        }
        acc.turns += tsum;
        acc.maxfin = tmax > acc.maxfin ? tmax : acc.maxfin;
        acc.dirty = true;
        lturns = 0; lmax = 0;
    };
    // Account the finished lanes of `pending` from their frozen states: ended / repeated / reached
    // the stage end (survivor: appended to the output regions). All lanes call this.
    auto settle = [&]() {
        const bool mine = (pending >> lane) & 1ull;
        const bool e = mine && ended<W>(s, C);
        const bool d = mine && !e && same<W>(s, t);
        const u32 age = a0 + j;
        if (e) {
            lturns += age; lmax = age > lmax ? age : lmax;
            if (acc.lb) atomicAdd(&acc.ll[age < acc.lb - 1 ? age : acc.lb - 1], 1u);
        }
        if (d) {
            lturns += age;
            if (acc.hb) atomicAdd(&acc.lh[j < acc.hb - 1 ? j : acc.hb - 1], 1u);
        }
        const u64 em = ballot(e), dm = ballot(d);
        const u64 sm = pending & ~em & ~dm;
        acc.done += __popcll(em | dm); acc.cyc += __popcll(dm); acc.dirty = true;
        if constexpr ((FLAGS & F_WALK) != 0) {
            const u64 wm = ballot(L.walk) & dm;
            if (wm) acc.pushed += push_cycles<W>(A.cq, TP, wm, s, n, L.req, j, L.pos, L.home);
        }
        if (sm) {
            const u32 c = (u32)__popcll(sm);
            const u32 r = rank_in(sm);
            u32 placed = 0;
            while (placed < c) {
                if (och == UNBOUNDED || ofill == TP) {
                    close_region();
                    och = (u32)wave_fetch_add(&A.pipe->nreg[A.stage], 1);
                    ofill = 0;
                    if (och >= A.max_regions) {       // no room: an error, reported by the host
                        if (lane == 0) atomicAdd((unsigned long long *)&A.pipe->ovf, 1ull);
                        och = UNBOUNDED;
                        break;
                    }
                }
                const u32 kk = (TP - ofill) < (c - placed) ? (TP - ofill) : (c - placed);
                if (((sm >> lane) & 1ull) && r >= placed && r < placed + kk) {
                    const u64 i = (u64)och * TP + ofill + (r - placed);
                    T qa, qb;
                    pack<W>(s, n, qa, qb);
                    A.qout.a[i] = qa; A.qout.b[i] = qb;
                    if constexpr ((FLAGS & F_MULTI) != 0) A.qout.req[i] = L.req;
                    if constexpr ((FLAGS & F_WALK) != 0) { A.qout.pos[i] = L.pos; A.qout.home[i] = L.home; }
                }
                ofill += kk; placed += kk;
            }
        }
        pending = 0;
    };

    Zh<W> z = make_zh<W>(A.zero);
    for (;;) {
        // ---- refill idle lanes (when enough are idle, or all)
        u64 idle = ~ballot(live);
        if ((u32)__popcll(idle) >= REFILL_IDLE || idle == ~0ull) {
            if (pending) settle();
            for (;;) {
                if (c_i >= c_n) {
                    if (x_n == 0) break;                       // stream exhausted
                    ca = xa; cb = xb; creq = xreq; chome = xhome; cpos = xpos; c_n = x_n; c_i = 0;
                    load_next();                               // the following buffer, loaded ahead
                }
                if (!idle) break;
                u32 k = (u32)__popcll(idle);
                k = k < c_n - c_i ? k : c_n - c_i;
                if constexpr ((FLAGS & F_MULTI) != 0) {
                    // only records of the wave's current request; switch request once drained
                    const u64 other = ballot(lane >= c_i && lane < c_n && creq != cur_req);
                    const u32 first = other ? (u32)(__ffsll((long long)other) - 1) : c_n;
                    if (first - c_i < k) k = first - c_i;
                    if (k == 0) {
                        if (idle != ~0ull) break;
                        reduce_lanes();
                        cur_req = lanev(creq, c_i);
                        acc.select(cur_req, reqs);
                        continue;
                    }
                }
                const u32 r = rank_in(idle);
                const u32 src = c_i + (r < k ? r : 0u);
                const T na = shfl_T<W>(ca, src), nb = shfl_T<W>(cb, src);
                u32 nreq = req0, nhome = 0;
                u64 npos = 0;
                if constexpr ((FLAGS & F_MULTI) != 0) nreq = (u32)__shfl((int)creq, (int)src);
                if constexpr ((FLAGS & F_WALK) != 0) {
                    nhome = (u32)__shfl((int)chome, (int)src);
                    npos = ((u64)(u32)__shfl((int)(u32)(cpos >> 32), (int)src) << 32) | (u64)(u32)__shfl((int)(u32)cpos, (int)src);
                }
                const bool take = ((idle >> lane) & 1ull) && r < k;
                if (take) {
                    unpack<W>(na, nb, s, n);
                    open_sa<W>(s, A.zero);
                    t = s; j = 0; C = W * n - 2u * W; live = true;
                    L.req = nreq; L.pos = npos; L.home = nhome;
                    if constexpr ((FLAGS & F_WALK) != 0) L.walk = reqs[nreq].walk != 0;
                }
                idle &= ~ballot(take);
                c_i += k;
            }
            if (idle == ~0ull) break;                          // nothing left to play
        }

        // ---- two turns for the playing lanes (finished lanes stay frozen); the refill check and its
        // mask bookkeeping run once per two turns
#pragma unroll
        for (int u = 0; u < 2; u++) {
            if (live) {
                turn<W, RULE>(s, C, z);
                ++j;
                // bitwise, not short-circuit: no branch inside the turn
                if (ended<W>(s, C) | same<W>(s, t) | (j == LEN)) live = false;
            }
        }
        pending |= ballot(!live) & ~idle;
    }
    if (pending) settle();
    close_region();
    reduce_lanes();
    acc.flush(reqs);
    if (acc.pushed && lane == 0) atomicAdd((unsigned long long *)&A.pipe->ctotal, (unsigned long long)acc.pushed);
}

// ------------------------------------------------------------------------------------------
// Flow kernel: no stages. Every lane plays its own game, and a lane whose game is over takes the
// next position at the start of the next span of SPAN turns (FLOW_SPAN; twice that for Brent over
// deal blocks), so a wave never waits for the slowest game of a group and no survivor records
// are written. Births fall on span starts (the clock a multiple of SPAN); the age is clock -
// birth and the cycle length clock - R, per lane. Because every age is a multiple of SPAN at a
// span start, each detector's tortoise moves fall on fixed turns of the span:
// * Brent (DET 0, any n and rule): the tortoise takes the state at ages 2^k - 1 (after that turn's
//   repeat test), exactly as the reference's war_sim (gpu/war_gpu.py:77-96), so `turns` is the
//   reference's count. Age 2^k - 1 is turn 2^k - 2 of the lane's first span, and turn SPAN - 2
//   of a later span (when age + 1 is a power of 2).
// * the alternation detector (DET 2, odd n, rule 0): checkpoints every ALT_K turns.
// Positions come from tiles of the job list as in the first stage: lane l of a tile owns positions
// l*per .. l*per+per-1 (stepped with next_perm); the 64 lanes' current positions, dealt, are the
// pool that refills hand out in lane order (cross-lane reads), and the pool steps every lane to its
// next position once all 64 are handed out. With cycle walks (WALK), a repeat's meeting state (the
// tortoise, which stays put once the game is over), cycle length and position index go to the
// cycle records of the position's tile when the span is accounted, as in the staged pipeline, for
// k_walk.
template <typename P>
__device__ __forceinline__ P shfl_word(P v, u32 src) {
    if constexpr (sizeof(P) == 8) {
        return ((u64)(u32)__shfl((int)(u32)(v >> 32), (int)src) << 32) | (u64)(u32)__shfl((int)(u32)v, (int)src);
    } else {
        P r = 0;
        for (int q = 3; q >= 0; q--) r = (r << 32) | (P)(u32)__shfl((int)(u32)(v >> (32 * q)), (int)src);
        return r;
    }
}

// The turn of the span each lane's game ended on, from the bit planes of the end turns: bit b of the
// turn is this lane's bit of fb[b]. Horner's rule in add-with-carry steps (e = 2e + bit, the bit as
// the carry in), one VALU instruction per plane instead of a select and an add.
template <int FB>
__device__ __forceinline__ u32 end_turn(const u64 (&fb)[FB]) {
    u32 e = in_mask(fb[FB - 1]) ? 1u : 0u;
    for (int b = FB - 2; b >= 0; b--) {
        u32 r;
        u64 co;                // the carry out, unused
        asm volatile("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(r), "=s"(co) : "v"(e), "s"(fb[b]));
        (void)co;
        e = r;
    }
    return e;
}

constexpr int FLOW_BRENT = 0, FLOW_ALT = 2, FLOW_PAIR = 3;
// The pair-run detector (n = 2 mod 4, rule 0): the reference's loop-shape test
// (gpu/war_gpu.py:51-73, Theorem 3) finds the cycles of such n made of winner runs of exactly two
// turns, all but a few (n = 14 canonical positions: all but the cycles of length 96). So a game whose
// last ALT_K turns held a run of one turn or of three or more is, for all but those, not on its
// cycle yet, and its tortoise jumps at the checkpoint, as after a repeated winner in the alternation
// detector; such pattern jumps stop once the game is PAIR_P turns old, after which only the timed
// jumps (waits of ALT_L0, then twice as long, ...) move the tortoise, so a cycle of any shape is
// seen. Every output is exact and equals Brent's; only `turns` (its own) differs.
constexpr u32 PAIR_P = 128;
constexpr u32 FLOW_SPAN = ALT_K * FLOW_BLOCKS;
// The flow kernel's pools: dealt positions in LDS, ahead of the waves' histograms (blocks of 4
// waves); an entry's sa is FLOW_POOL_EMPTY when the run that deals it holds no more positions (sa of a
// real position is a bit offset below 256). A batch of one request without cycle walks keeps a ring
// of two generations of 64 (every entry shares the request and needs no position index or home
// tile), so one refill round serves every free lane; other batches keep one generation.
constexpr u32 FLOW_POOL_EMPTY = 0xFFFFFFFFu;
__host__ __device__ constexpr u32 flow_pool_entries(bool multi, bool walk) { return (multi || walk) ? 64u : 128u; }
template <int W> __host__ __device__ constexpr u32 flow_pool_bytes(bool multi, bool walk) {
    return 4u * flow_pool_entries(multi, walk) * (u32)sizeof(St<W>);
}

// MULTI: the batch holds several requests (the serving loop's steps, war_run_batch). A tile belongs
// to one job, so the pool's positions share one request (its n, mode and C change at a tile grab);
// a game keeps its request and C per lane, and accounting groups the finished lanes by request.
// BLK: every job is 4-bit deals in whole blocks of 24 orders, and so is every lane run.
// CAP: some request of the batch has an explicit step cap (war_request.step_cap), tested at span
// ends (a variant, not a run-time flag: the flag and the cap held across the loop cost the kernel
// scalar registers it spills into vector lanes, 1 % of its time at n = 14 canonical).
// FWM: Brent under rule 0, Brent's first window as its power minus 1 (0: every window, the tortoise
// first at age 7; 127: first at age 127; 0xFFFFFFFF: none), a variant for the same reason.
// OFF: canonical positions dealt after their first turn (canon_top1), which player 1's top card n - 1
// always wins: a game is born one turn old (its birth clock one before the span start), so ages at a
// span start are 1 mod SPAN and the old games' tortoise moves fall on turn SPAN - 3; the played turn
// count, the ages and every output are unchanged. With a first window of 16 or more only (no move in
// a game's first span).
template <int W, int RULE, int DET, bool WALK, bool MULTI, bool BLK = false, bool CAP = false, u32 FWM = 0,
          bool OFF = false>
__global__ __launch_bounds__(256) void k_flow(StageArgs<W> A) {
    static_assert(!OFF || (W == 4 && RULE == 0 && DET == 0 && !MULTI && !BLK && !CAP && FWM >= 15u),
                  "positions dealt after their first turn: 4-bit canonical single requests, Brent, rule 0, first window >= 16");
    static_assert(!BLK || (W == 4 && !WALK), "deal blocks: 4-bit deals without walks");
    static_assert(DET == FLOW_BRENT || ((DET == FLOW_ALT || DET == FLOW_PAIR) && RULE == 0),
                  "the alternation and pair-run detectors are for rule 0");
    static_assert(!WALK || W != 6, "no cycle walks in layout 6");
    static_assert(!MULTI || DET != FLOW_PAIR, "the pair-run detector runs on single requests");
    typedef typename Lay<W>::P P;
    constexpr int PW = Lay<W>::PW;
    extern __shared__ u32 lds[];
    const u32 wid = threadIdx.x >> 6;
    const u32 lane = lane_id();
    // LDS: the four waves' pools (PE dealt positions each), then their histograms
    constexpr bool RING = !MULTI && !WALK;
    constexpr u32 PE = flow_pool_entries(MULTI, WALK);
    St<W> *pool = reinterpret_cast<St<W> *>(lds) + wid * PE;
    WaveAcc acc;
    acc.init(lds + flow_pool_bytes<W>(MULTI, WALK) / 4u + wid * (A.hb + A.lb), A.hb, A.lb);
    const ReqDev *reqs = A.reqs;
    const u32 req = A.hdr->req0;
    acc.select(req, reqs);
    // the pool's request (all of a batch with one request; the current tile's with MULTI)
    u32 preq = req, n = reqs[req].n, mode = BLK ? 0u : reqs[req].mode;
    bool walk = WALK && reqs[req].walk != 0;   // cycles go to the cycle records for k_walk
    u32 C = c_of<W>(n);
    const u32 per = A.per, TP = A.tile_pos;
    const u64 ntiles = A.hdr->ntiles;
    const u32 rcap = CAP ? reqs[req].cap : 0u;
    u64 *grab = &A.pipe->grab[0];
    u64 nxt = 0;
    if (lane == 0) nxt = atomicAdd((unsigned long long *)grab, 1ull);

    // ---- the pool: lane l deals the current position of its run (a whole generation of 64 runs as
    // one wave step) into entry l of the wave's pool in LDS, from where free lanes take them
    P q = 0;
    u64 qr = 0;                // 4-bit canonical runs: the order reversed (canon_top builds from it)
    u32 k = 0, left = 0;       // this lane's run: current order / split; positions left, it included
    St<W> ps = {};             // the current position, dealt
    u64 pp = 0;                // its index (cycle walks rebuild the start from it)
    u32 tile = 0;              // the pool's tile: the home of its positions' cycle records
    u32 ptr = 64;              // next pool entry to hand out (64: all handed out)
    u32 cons = 0, avail = 0;   // RING: the first entry not handed out, and how many follow it
    bool more = true;          // positions remain (in the pool or in tiles not yet taken)
    // A canonical position with split n - 2 (player 2 holds one card, below player 1's top card
    // n - 1) ends on its first turn: a game of length 1. Single-request batches (SKIP) count such
    // positions per lane (ninst, added to the totals at the end) and step over them, so no lane idles
    // a span for one; the run's position count `left` includes them. step is called with left >= 2
    // (the next position is in the run); after an instant position it moves on to the next order
    // only if that is in the run too (left >= 2 again): next_perm is never asked for the successor
    // of the last order.
    constexpr bool SKIP = !MULTI && !BLK;
    u32 ninst = 0;
    auto step = [&]() {
        if (mode == 2) {
            ++k;
            if constexpr (SKIP) {
                if (k == n - 2u) {
                    ninst++;
                    left--;
                    if (left > 1u) {
                        ++k;
                        if constexpr (WALK) pp++;
                    }
                }
            }
            if (k == n - 1u) {
                k = 0;
                q = next_perm<PW>(q, (int)n - 1);
                if constexpr (W == 4) qr = rev_fields4(q);
            }
        } else {
            q = next_perm<PW>(q, (int)n);
        }
        if constexpr (WALK) pp++;
    };
    // --half (gpu/war_gpu.py:150-154) keeps the deals whose card n-1 goes to player 1
    auto skip_filtered = [&]() {
        if (mode == 1) while (left && !top_card_to_p1<PW>(q, (int)n)) { if (--left) step(); }
    };
    // BLK: every lane run is whole blocks of 24 deals (the orders that share all but the last four
    // cards; the host aligns jobs and runs to them): a run steps through each block in Heap's order,
    // swapping two cards of the dealt position in place (heap_swap_top), and deals from the order
    // only at a block's start (next_block24)
    u32 hj = 0;                // the step within the block (wave-uniform)
    // every lane to the next position of its run; a new tile once every run is done
    auto advance = [&]() -> bool {
        bool fresh = true;     // deal ps from q (else heap_swap_top moved it)
        if constexpr (BLK) {
            if (hj < 23u) { heap_swap_top(ps, hj, n); hj++; fresh = false; }
            else { if (left > 1u) q = next_block24(q, (int)n); hj = 0; }
        } else {
            if (left > 1u) step();
        }
        left = left ? left - 1u : 0u;
        skip_filtered();
        while (ballot(left != 0u) == 0ull) {
            const u64 g0 = uni64(nxt);
            if (g0 >= ntiles) return false;
            if (lane == 0) nxt = atomicAdd((unsigned long long *)grab, 1ull);
            u32 lo = 0, hi = A.hdr->njobs;
            while (hi - lo > 1) { u32 mid = (lo + hi) >> 1; if (A.jobs[mid].tile_base <= g0) lo = mid; else hi = mid; }
            const Job J = A.jobs[lo];
            if constexpr (MULTI) {
                preq = J.req; n = reqs[preq].n; mode = BLK ? 0u : reqs[preq].mode; C = c_of<W>(n);
            }
            const u64 first = J.lo + (g0 - J.tile_base) * (u64)TP + (u64)lane * per;
            left = first < J.hi ? (u32)((J.hi - first) < per ? (J.hi - first) : per) : 0u;
            u64 f = first;     // the run's first position to play
            if (left) {
                if (mode == 2) {
                    const u32 m = n - 1u;
                    u64 r = first / m;
                    k = (u32)(first - r * m);
                    if constexpr (SKIP) {
                        if (k == n - 2u) {                      // an instant position first
                            ninst++;
                            left--;
                            if (left) { k = 0; r++; f++; }      // the next order is in the run
                        }
                    }
                    q = unrank<PW>(r, (int)m);
                    if constexpr (W == 4) qr = rev_fields4(q);
                } else {
                    k = 0;
                    q = unrank<PW>(first, (int)n);
                }
            }
            if constexpr (WALK) { pp = f; tile = (u32)g0; }
            hj = 0;
            fresh = true;
            skip_filtered();
        }
        if (left && fresh) {
            if constexpr (W == 4) {                          // 4-bit games play in turn_top's form
                ps = (mode == 2) ? (OFF ? canon_top1(qr, (int)n, (int)k) : canon_top(qr, (int)n, (int)k))
                                 : flip_piles(deal<W>(q, (int)n));
            } else {
                ps = (mode == 2) ? canon<W>(q, (int)n, (int)k) : deal<W>(q, (int)n);
            }
        }
        // entry `lane` of the pool: this lane's position, or empty once its run is done (every lane
        // writes its own entry; the wave's later reads of other lanes' entries follow in program
        // order, and LDS operations of one wave complete in order)
        St<W> e = ps;
        if (!left) e.sa = FLOW_POOL_EMPTY;
        if constexpr (RING) {      // the generation after the entries not yet handed out
            pool[((cons + avail) & (PE - 1u)) + lane] = e;
            avail += 64u;
        } else {
            pool[lane] = e;
            ptr = 0;
        }
        __builtin_amdgcn_wave_barrier();
        return true;
    };

    // ---- the games
    St<W> s = {};
    u32 sb = C;                // C - s.sa: player 2's append offset, carried from turn to turn
    u32 gC = C, greq = req;    // MULTI: the game's C and request
    AltLane<W> T_;             // tortoise, its clock R; L: the alternation detector's timed wait
    T_.t = s; T_.R = 0; T_.L = ALT_L0;
    u32 birth = 0, clk = 0;
    u64 gpos = 0;              // cycle walks: the game's position index and home tile
    u32 ghome = 0;
    u64 alive = 0, newm = 0, prevm = 0, streakm = 0, endm = 0, cycm = 0;
    u64 preveqm = 0, youngm = 0;   // pair-run detector: last turn's same-winner mask; games <= PAIR_P old
    // the span: FLOW_SPAN turns (two alternation checkpoints); Brent over deal blocks runs spans
    // twice as long (long games: half the span ends, 4 % faster at n = 13; canonical positions,
    // most of which end within 30 turns, lose more to lanes idle until the span end, and the
    // alternation detector measured slower with 32)
    constexpr u32 SPAN = (BLK && DET == FLOW_BRENT) ? 2u * FLOW_SPAN : FLOW_SPAN;
    // Brent's turn takes turn_top's written compare and selects (the compiled ones put a VOP2 select
    // right after the compare): 2 % faster on canonical n = 14, 3 % on n = 13 deals; the
    // alternation detector's compiled turn already has the VOP3 selects and runs 4 % slower with it
    constexpr bool SELASM = DET == FLOW_BRENT;
    constexpr int FB = __builtin_ctz(SPAN);
    static_assert(FLOW_SPAN == 2u * ALT_K && PAIR_P % FLOW_SPAN == 0u, "checkpoints at turns 7 and 15 of a span");
    static_assert((SPAN & (SPAN - 1u)) == 0u && (AGE_CAP % SPAN) == 0u && SPAN <= 1024u,
                  "span: a power of 2 dividing AGE_CAP");
    u64 fb[FB];                // the turn of the span each finished game ended in, bit by bit
    for (int b = 0; b < FB; b++) fb[b] = 0;
    u64 lturns = 0;            // per lane: sum of its finished games' lengths / detection ages
    u32 lmax = 0;              // per lane: longest ending game
    // (MULTI: both belong to the request the wave's accumulator follows, acc.req)
    // Fold the per-lane sums into the wave's accumulator (wave sums; every lane calls this).
    auto fold_lanes = [&]() {
        u64 tsum = lturns;
        u32 tmax = lmax;
        for (int off = 32; off > 0; off >>= 1) {
            tsum += __shfl_xor(tsum, off);
            const u32 o = __shfl_xor(tmax, off);
            tmax = o > tmax ? o : tmax;
        }
        acc.turns += tsum;
        acc.maxfin = tmax > acc.maxfin ? tmax : acc.maxfin;
        acc.dirty = true;
        lturns = 0;
        lmax = 0;
    };
    // MULTI: make the accumulator follow request r (folding the per-lane sums of the previous one
    // first); every lane calls this
    auto switch_to = [&](u32 r) {
        if (r != acc.req) {
            fold_lanes();
            acc.select(r, reqs);
        }
    };
    Zh<W> z = make_zh<W>(A.zero);
    Zt zt = make_zt(A.zero);

    // free lanes take pool entries in lane order until every lane plays or no position is left
    // (every lane calls this)
    auto refill = [&]() {
        for (;;) {
            u64 dead = ~alive;
            while (dead) {
                const u32 i = rank_in(dead);
                u64 tm;                // the first free lanes, one per entry
                u32 src;
                if constexpr (RING) {
                    // fewer entries left than free lanes (at most 64): the other generation's slots
                    // are free, so deal one there
                    if (avail < (u32)__popcll(dead) && more && !advance()) more = false;
                    if (avail == 0u) break;
                    tm = ballot(i < avail) & dead;
                    src = (cons + i) & (PE - 1u);
                    const u32 took = (u32)__popcll(tm);
                    cons = (cons + took) & (PE - 1u);
                    avail -= took;
                } else {
                    if (ptr == 64u && !advance()) { more = false; break; }
                    tm = ballot(i < 64u - ptr) & dead;
                    src = (ptr + i) & 63u;
                    ptr += (u32)__popcll(tm);
                }
                // a free lane's state is not read (its game is over), so the lanes served take their
                // entry straight into it, empty entries too (those lanes stay free)
                if (in_mask(tm)) s = pool[src];
                u64 npos = 0;
                if constexpr (WALK) npos = shfl_word(pp, src);
                dead &= ~tm;
                const u64 gm = ballot(s.sa != FLOW_POOL_EMPTY) & tm;    // entries that hold a position
                if (in_mask(gm)) {
                    open_sa<W>(s, A.zero);
                    sb = C - s.sa;
                    birth = clk - (OFF ? 1u : 0u);
                    if constexpr (MULTI) { gC = C; greq = preq; }
                    // Brent's first tortoise (the start) needs no copy: no state repeats after one
                    // turn (player 1's card count changes every turn), so the repeat test of a new
                    // game's first turn is masked off, and the tortoise takes the state after it
                    if constexpr (DET != FLOW_BRENT) { T_.t = s; T_.R = clk; T_.L = ALT_L0; }
                    // rule 0: a tortoise no state can equal (every field the top card value) until
                    // its first move, so the repeat tests before it need no mask
                    else if constexpr (RULE == 0) { T_.t.a = ~decltype(T_.t.a)(0); }
                    if constexpr (WALK) { gpos = npos; ghome = tile; }
                }
                u64 e0 = 0;
                // a position that is over before it starts (only with fewer than 2 cards): length 0
                if (n < 2u) {
                    e0 = gm;
                    if constexpr (MULTI) switch_to(preq);
                    acc.ended((u32)__popcll(e0), 0);
                }
                alive |= gm & ~e0;
                newm |= gm & ~e0;
                // before the first turn, the position's player 1 counts as the last winner (layout
                // 6 names players by who holds card n-1)
                if constexpr (W == 6) prevm = (prevm & ~gm) | (ballot(p1_is_dealt_p1<W>(s, mode)) & gm);
                else prevm |= gm;
                preveqm &= ~gm;   // ... and as having changed hands (pair-run detector)
            }
            if (alive || !more) return;
        }
    };
    // account the games that finished in the span that began at clock clk0 (every lane calls this)
    auto settle = [&](u32 clk0) {
        const u32 fin = clk0 + 1u + end_turn<FB>(fb);
        for (int b = 0; b < FB; b++) fb[b] = 0;
        const u32 age = fin - birth, lam = fin - T_.R;
        if constexpr (MULTI) {
            // the games of the request the accumulator follows go to the per-lane sums as in a
            // single-request batch; only when some game belongs to another request are the sums
            // folded and the other requests accounted group by group (wave sums, one LDS add per
            // lane), the accumulator then following the last of them
            const u64 ended = endm | cycm;
            const u64 own = ballot(greq == acc.req) & ended;
            const u64 rest = ended & ~own;
            if (own) {
                if (in_mask(own)) lturns += age;
                if (in_mask(endm & own)) {
                    lmax = age > lmax ? age : lmax;
                    if (acc.lb) atomicAdd(&acc.ll[age < acc.lb - 1 ? age : acc.lb - 1], 1u);
                }
                if (acc.hb && in_mask(cycm & own)) atomicAdd(&acc.lh[lam < acc.hb - 1 ? lam : acc.hb - 1], 1u);
                acc.done += __popcll(own); acc.cyc += __popcll(cycm & own); acc.dirty = true;
                if constexpr (WALK) {
                    if (reqs[acc.req].walk && (cycm & own)) {
                        St<W> m = T_.t;
                        if constexpr (W == 4) m = flip_piles(m);
                        acc.pushed += push_cycles<W>(A.cq, TP, cycm & own, m, reqs[acc.req].n, acc.req, lam, gpos, ghome);
                    }
                }
            }
            for (u64 rem = rest; rem;) {
                const u32 r = lanev(greq, (u32)(__ffsll((long long)rem) - 1));
                const u64 mr = ballot(greq == r) & rem;
                rem &= ~mr;
                switch_to(r);
                acc.lanes(endm & mr, cycm & mr, age, lam);
                if constexpr (WALK) {
                    if (reqs[r].walk && (cycm & mr)) {
                        St<W> m = T_.t;
                        if constexpr (W == 4) m = flip_piles(m);
                        acc.pushed += push_cycles<W>(A.cq, TP, cycm & mr, m, reqs[r].n, r, lam, gpos, ghome);
                    }
                }
            }
            endm = cycm = 0;
            return;
        }
        if (in_mask(endm | cycm)) lturns += age;
        if (in_mask(endm)) {
            lmax = age > lmax ? age : lmax;
            if (acc.lb) atomicAdd(&acc.ll[age < acc.lb - 1 ? age : acc.lb - 1], 1u);
        }
        if (acc.hb && in_mask(cycm)) atomicAdd(&acc.lh[lam < acc.hb - 1 ? lam : acc.hb - 1], 1u);
        // cycle walks: the meeting state is the tortoise (a finished lane's tortoise stays put),
        // recorded in the usual pile form
        if constexpr (WALK) {
            if (walk && cycm) {
                St<W> m = T_.t;
                if constexpr (W == 4) m = flip_piles(m);
                acc.pushed += push_cycles<W>(A.cq, TP, cycm, m, n, req, lam, gpos, ghome);
            }
        }
        acc.done += __popcll(endm | cycm); acc.cyc += __popcll(cycm); acc.dirty = true;
        endm = cycm = 0;
    };

    refill();
    while (alive) {
        const u32 clk0 = clk;
        const u64 alive0 = alive;
        if constexpr (DET == FLOW_PAIR) {
            // the span's checkpoints are at ages clk0 - birth + 8 and + 16 (births fall on span
            // starts, and PAIR_P is a multiple of the span): both are <= PAIR_P iff this holds
            youngm = ballot(clk0 - birth <= PAIR_P - FLOW_SPAN);
        }
#pragma unroll
        for (u32 tix = 0; tix < SPAN; tix++) {       // turn tix of the span
            u64 over, pm;          // pm: the lanes whose player 1 won the turn
            if constexpr (W == 4) pm = turn_top<RULE, !MULTI, SELASM>(s, sb, MULTI ? gC : C, zt, over);
            else pm = ballot(turn_sb<W, RULE, !MULTI>(s, sb, MULTI ? gC : C, z, over));
            // the common case (no lane ended or matched the tortoise's player-1 pile) costs two
            // scalar operations and the branch; the scalar unit is busy about 70 % of the time here.
            // Every cycle is of even length (over one cycle player 1's card count comes back, so it
            // wins as many of the cycle's turns as it loses), so a game meets its tortoise only at
            // an even number of turns after the tortoise was taken. Tortoises are taken at odd ages
            // (Brent: 2^k - 1) or even ones (the checkpoint detectors: multiples of ALT_K, and a new
            // game's start), and births fall on span starts (OFF: one turn before), so the age
            // parity after turn tix is the same in every lane: the compare runs on every other turn
            // only, and each detection falls on the same turn as with a compare on every turn.
            constexpr u32 CMP_PAR = DET == FLOW_BRENT ? (OFF ? 1u : 0u) : 1u;
            u64 cand = 0;
            if ((tix & 1u) == CMP_PAR) cand = ballot(s.a == T_.t.a);
            if constexpr (DET == FLOW_BRENT) {
                // Brent's windows of the tortoise at ages 1 and 3 see only cycles of length <= 2
                // and <= 4, and no rule-0 War cycle is that short: over one cycle each player plays
                // every card it holds (its pile would otherwise come back with another card on top),
                // so a cycle is at least max(cards of player 1, of player 2) >= ceil(n/2) long,
                // which is 5 or more from n = 9 on, and the rule-0 cycles of n <= 8 are 6 and 8
                // long (every canonical position). So under rule 0 a new game's tortoise first
                // moves at age 7 (or later: the first window FWM) and no state
                // equals it before: every detection, and so `turns`, is Brent's. Rule 1 has cycles
                // of 2 and 4 turns (n = 3, 5) and keeps every window.
                if (RULE != 0 && tix < 1u) cand &= ~newm;
            }
            if constexpr (DET == FLOW_ALT) {
                // a repeated winner on a turn of odd age (tix even: births fall on span starts)
                if ((tix & 1u) == 0u) streakm |= ~(pm ^ prevm);
                prevm = pm;
            } else if constexpr (DET == FLOW_PAIR) {
                // a run violation: the winner repeated twice in a row (a run of three or more) or
                // changed twice in a row (a run of one)
                const u64 eqm = ~(pm ^ prevm);
                streakm |= ~(eqm ^ preveqm);
                prevm = pm;
                preveqm = eqm;
            } else {
                (void)pm;
            }
            if (const u64 x = (over | cand) & alive) {
                // ended games: x & over; the ended set of the span is rebuilt at its end (live at
                // the start, not now, not repeated), so only the rare repeats are collected here
                u64 ev = x & over;
                if (const u64 dc = x & ~over) {
                    const u64 dm = dc & rest_mask<W>(s, T_.t);
                    cycm |= dm;
                    ev |= dm;
                }
                for (int b = 0; b < FB; b++) if ((tix >> b) & 1u) fb[b] |= ev;
                alive &= ~ev;
            }
            if constexpr (DET == FLOW_BRENT) {
                // Brent's tortoise moves at ages 2^k - 1 (from 7 on under rule 0, above): turn 2^k - 2 of a
                // lane's first span; turn SPAN - 2 of a later span when age + 1 is a power of 2. Under
                // rule 0 only from the first window FWM on: a window whose power is below every cycle
                // length of the batch's n cannot see a repeat (the host derives the first window from
                // the published cycle lengths), so leaving it out changes no detection; age + 1 (a
                // power of 2) must then also have no bit of FWM set, for the lanes born
                // at this span start (age + 1 = tix + 2) as for the older ones. (A move left out must
                // not be replaced by an earlier one either: the tortoise would stay there, and its
                // longer window could see a repeat Brent sees later.)
                // (OFF: no move in a game's first span; the old games' at turn SPAN - 3, where age + 1
                // is a multiple of SPAN)
                const bool newb = !OFF && ((tix + 2u) & (tix + 1u)) == 0u && tix >= (RULE == 0 ? 6u : 0u) &&
                                  (RULE != 0 || ((tix + 2u) & FWM) == 0u);
                const bool oldw = tix == SPAN - 2u - (OFF ? 1u : 0u);
                if (newb || oldw) {   // unrolled: folded
                    u64 jm = 0;
                    if (newb) jm = newm & alive;
                    if (oldw) {
                        const u32 age1 = clk0 + tix + 2u - birth;    // age + 1
                        jm |= ballot((age1 & ((age1 - 1u) | (RULE == 0 ? FWM : 0u))) == 0u) & alive;
                    }
                    if (jm) {
                        if (in_mask(jm)) { T_.t = s; T_.R = clk0 + tix + 1u; }
                    }
                }
            } else {
                if ((tix + 1u) % ALT_K == 0u) {
                    // checkpoint: timed jumps (the wait doubles), then jumps after a repeated winner
                    // (alternation) or a run violation while the game is young (pair-run)
                    const u32 now = clk0 + tix + 1u;
                    const u64 ex = ballot(now - T_.R >= T_.L) & alive;
                    const u64 pat = DET == FLOW_PAIR ? (streakm & youngm) : streakm;
                    const u64 js = (ex | pat) & alive;
                    streakm = 0;
                    if (js) {
                        if (in_mask(js)) { T_.t = s; T_.R = now; }
                        // the timed jumps double their wait; they are rare (waits of ALT_L0 turns
                        // and longer, against pattern jumps at most checkpoints), so behind a
                        // branch the common checkpoint skips the wait's select and shift
                        if (__builtin_expect(ex != 0ull, 0)) {
                            if (in_mask(ex)) T_.L <<= 1;
                        }
                    }
                }
            }
        }
        clk += SPAN;
        newm = 0;
        endm = alive0 & ~alive & ~cycm;
        if (endm | cycm) settle(clk0);
        // the safety stop, tested every 1024 turns: a game still running after AGE_CAP turns is
        // undetermined (never for a valid game; its turns are not counted)
        if ((clk & 1023u) == 0u) {
            const u64 um = ballot(clk - birth >= AGE_CAP) & alive;
            if (um) {
                for (u64 rem = um; rem;) {
                    u64 mr = rem;
                    if constexpr (MULTI) {
                        const u32 r = lanev(greq, (u32)(__ffsll((long long)rem) - 1));
                        mr = ballot(greq == r) & rem;
                        switch_to(r);
                    }
                    rem &= ~mr;
                    acc.undet += __popcll(mr); acc.done += __popcll(mr); acc.dirty = true;
                }
                alive &= ~um;
            }
        }
        // an explicit step cap (war_request.step_cap), tested at every span end when a request of
        // the batch has one: ages at span ends are multiples of the span, which divides the cap, so
        // a game stops exactly when it has played the cap's turns unresolved, charged those turns
        if constexpr (CAP) {
            u32 gcap = rcap;
            if constexpr (MULTI) gcap = reqs[greq].cap;
            const u64 um = ballot(gcap != 0u && clk - birth >= gcap) & alive;
            if (um) {
                for (u64 rem = um; rem;) {
                    u64 mr = rem;
                    u32 cp = rcap;
                    if constexpr (MULTI) {
                        const u32 r = lanev(greq, (u32)(__ffsll((long long)rem) - 1));
                        mr = ballot(greq == r) & rem;
                        switch_to(r);
                        cp = reqs[r].cap;
                    }
                    rem &= ~mr;
                    const u32 c = (u32)__popcll(mr);
                    acc.undet += c; acc.done += c; acc.turns += (u64)c * cp; acc.dirty = true;
                }
                alive &= ~um;
            }
        }
        // refill once a lane is free (a higher threshold measured within 0.2 %: almost every span
        // frees several lanes)
        if (more && ~alive) refill();
    }
    fold_lanes();
    if constexpr (SKIP) {            // the length-1 games the pool stepped over
        u32 t = ninst;
        for (int off = 32; off > 0; off >>= 1) t += (u32)__shfl_xor((int)t, off);
        if (t) acc.ended(t, 1u);
    }
    acc.flush(reqs);
    if (WALK && acc.pushed && lane == 0) atomicAdd((unsigned long long *)&A.pipe->ctotal, (unsigned long long)acc.pushed);
}

// ------------------------------------------------------------------------------------------
// Plain kernel: the reference's structure. Lane g plays positions [lo + g*per, lo + (g+1)*per),
// one at a time, each with its own Brent loop (per-lane tortoise, power and lam).
template <int W> struct PlainArgs {
    u64 lo, hi;
    u32 per, hb, lb, tile_pos;   // tile_pos: slots per home tile of the cycle records
    u32 zero, pad;               // zero: always 0 (see Zh)
    ReqDev R;
    u32 req;
    CycQueue<W> cq;
    Pipe *pipe;
};

// The streak test (gpu/war24.py:98-125; odd n, rule 0): anchor at a turn whose winner-to-be holds an
// even number ws of cards and the other ls >= 2; if the players then alternate for
// lcm(n+1, ws, ls-1) turns and the state equals the anchor, the game never ends (Spivey Thm 6). A
// repeated winner drops the anchor. Still running with more than the cap's turns played at the top of
// its loop (2^22 by default, the reference's CAP): undetermined, as gpu/war24.py:103.
constexpr u32 STREAK_CAP = 1u << 22;
__device__ __forceinline__ u32 gcd_u(u32 a, u32 b) { while (b) { const u32 t = a % b; a = b; b = t; } return a; }
__device__ __forceinline__ u32 lcm_u(u32 a, u32 b) { return a / gcd_u(a, b) * b; }

// DET: 0 = Brent, 1 = the streak test (RULE 0 only; no cycle lengths, so no histogram or walks),
// 2 = the alternation detector (RULE 0 only; every output). CAPPED: the request has an explicit step
// cap (P.R.cap); without one the cap is the compile-time safety stop (a run-time cap in the loop
// measured 10 % slower).
template <int W, int RULE, int DET, bool CAPPED = false>
__global__ __launch_bounds__(256) void k_plain(PlainArgs<W> P) {
    typedef typename Lay<W>::T T;
    extern __shared__ u32 lds[];
    u32 *lh = lds, *ll = lds + P.hb;
    for (u32 i = threadIdx.x; i < P.hb + P.lb; i += blockDim.x) lds[i] = 0;
    __syncthreads();
    const u32 n = P.R.n, mode = P.R.mode, C = W * n - 2u * W;
    const u64 g = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    const u64 first = P.lo + g * (u64)P.per;
    u32 cnt = first < P.hi ? (u32)((P.hi - first) < P.per ? (P.hi - first) : P.per) : 0u;
    T q = 0; int k = 0;
    if (cnt) {
        if (mode == 2) { const u32 m = n - 1; const u64 r = first / m; k = (int)(first - r * m); q = unrank<W>(r, (int)m); }
        else q = unrank<W>(first, (int)n);
    }
    u64 cyc = 0, done = 0, maxfin = 0, turns = 0, undet = 0, pushed = 0;
    // step caps: Brent and the alternation detector stop a game unresolved after `cap` turns (the
    // request's, else the 2^24-turn safety stop); the streak test after more than `scap`
    const u32 cap = CAPPED ? P.R.cap : AGE_CAP, scap = CAPPED ? P.R.cap : STREAK_CAP;
    for (u32 j = 0; j < cnt; j++) {
        St<W> s = (mode == 2) ? canon<W>(q, (int)n, k) : deal<W>(q, (int)n);
        const bool take = (mode == 1) ? top_card_to_p1<W>(q, (int)n) : true;
        bool cy = false, und = false;
        u32 lam = 0;
        u64 rounds = 0;
        if (take && n >= 2 && !ended<W>(s, C)) {
            open_sa<W>(s, P.zero);
            St<W> t = s;
            Zh<W> z = make_zh<W>(P.zero);
            if constexpr (DET == 0) {
                // the generic turn with its own end compare: measured faster here than turn_top, the
                // flow kernel's form (n = 13: 4.64e9 against 3.81e9 deals/s), so the plain kernel,
                // the baseline of the persistent/plain ratio, keeps it
                u32 power = 1, l = 0;
                for (;;) {
                    turn<W, RULE>(s, C, z);
                    rounds++;
                    if (ended<W>(s, C)) break;
                    l++;
                    if (same<W>(s, t)) { cy = true; lam = l; break; }
                    if (l == power) { t = s; power <<= 1; l = 0; }
                    if (rounds >= cap) { und = true; break; }
                }
            } else if constexpr (DET == 2) {
                // the alternation detector, one lane at a time (same schedule as the persistent form)
                u32 R = 0, Lw = ALT_L0, prevw = 1u;
                bool streak = false;
                for (;;) {
                    const u32 w = turn<W, RULE>(s, C, z) ? 1u : 0u;
                    rounds++;
                    if (ended<W>(s, C)) break;
                    if (same<W>(s, t)) { cy = true; lam = (u32)rounds - R; break; }
                    streak = streak || (w == prevw && (rounds & 1u));   // turns of odd age
                    prevw = w;
                    if ((rounds & (ALT_K - 1u)) == 0) {
                        if ((u32)rounds - R >= Lw) { t = s; R = (u32)rounds; Lw <<= 1; }
                        else if (streak) { t = s; R = (u32)rounds; }
                        streak = false;
                    }
                    if (rounds >= cap) { und = true; break; }
                }
            } else {
                constexpr u32 M = (1u << W) - 1u;
                bool anch = false;
                u32 lam_s = 0, since = 0, prevw = 2u;     // prevw 2: no previous turn since the anchor
                for (;;) {
                    if (rounds > scap) { und = true; break; }
                    if (!anch) {
                        const u32 ta = (u32)s.a & M, tb = (u32)s.b & M;
                        const u32 l1 = s.sa / (u32)W + 1u;
                        const u32 ws = ta > tb ? l1 : n - l1, ls = n - ws;
                        if ((ws & 1u) == 0u && ls >= 2u) {
                            anch = true; since = 0; prevw = 2u;
                            lam_s = lcm_u(n + 1u, lcm_u(ws, ls - 1u));
                            t = s;
                        }
                    }
                    const u32 w = turn<W, 0>(s, C, z) ? 1u : 0u;
                    rounds++;
                    if (ended<W>(s, C)) break;
                    if (anch) {
                        if (w == prevw) anch = false;
                        else {
                            prevw = w;
                            if (++since == lam_s) {
                                if (same<W>(s, t)) { cy = true; break; }
                                anch = false;
                            }
                        }
                    }
                }
            }
        }
        if (und) undet++;
        if (take) {
            done++;
            turns += rounds;
            if (cy) {
                cyc++;
                if (P.hb) atomicAdd(&lh[lam < P.hb - 1 ? lam : P.hb - 1], 1u);
            } else if (!und) {
                if (rounds > maxfin) maxfin = rounds;
                if (P.lb) atomicAdd(&ll[rounds < P.lb - 1 ? (u32)rounds : P.lb - 1], 1u);
            }
        }
        // cycle records for the walks (reconverged here: aggregated per home tile)
        if (P.R.walk) {
            const u64 wm = ballot(cy);
            if (wm) {
                const u64 pos = first + j;
                pushed += push_cycles<W>(P.cq, P.tile_pos, wm, s, n, P.req, lam, pos, (u32)((pos - P.lo) / P.tile_pos));
            }
        }
        if (j + 1 < cnt) {
            if (mode == 2) { if (++k == (int)n - 1) { k = 0; q = next_perm<W>(q, (int)n - 1); } }
            else q = next_perm<W>(q, (int)n);
        }
    }
    // wave reductions, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        cyc += __shfl_xor(cyc, off);
        done += __shfl_xor(done, off);
        turns += __shfl_xor(turns, off);
        undet += __shfl_xor(undet, off);
        const u64 o = __shfl_xor(maxfin, off);
        maxfin = o > maxfin ? o : maxfin;
        const u64 pu = __shfl_xor(pushed, off);      // every lane active in a push saw its whole total
        pushed = pu > pushed ? pu : pushed;
    }
    if (lane_id() == 0 && done) {
        atomicAdd((unsigned long long *)&P.R.cnt[C_CYC], (unsigned long long)cyc);
        atomicAdd((unsigned long long *)&P.R.cnt[C_DONE], (unsigned long long)done);
        atomicAdd((unsigned long long *)&P.R.cnt[C_TURNS], (unsigned long long)turns);
        if (undet) atomicAdd((unsigned long long *)&P.R.cnt[C_UNDET], (unsigned long long)undet);
        atomicMax((unsigned long long *)&P.R.cnt[C_MAXFIN], (unsigned long long)maxfin);
    }
    if (lane_id() == 0 && pushed) atomicAdd((unsigned long long *)&P.pipe->ctotal, (unsigned long long)pushed);
    __syncthreads();
    for (u32 i = threadIdx.x; i < P.hb; i += blockDim.x)
        if (lh[i]) atomicAdd((unsigned long long *)&P.R.hist[i], (unsigned long long)lh[i]);
    for (u32 i = threadIdx.x; i < P.lb; i += blockDim.x)
        if (ll[i]) atomicAdd((unsigned long long *)&P.R.lenhist[i], (unsigned long long)ll[i]);
}

// ------------------------------------------------------------------------------------------
// Cycle walks. One record per lane; lockstep over the longest lam of the wave.
template <int W> struct WalkArgs {
    Pipe *pipe;
    const JobHdr *hdr;
    const ReqDev *reqs;
    CycQueue<W> cq;
    u32 tile_pos;
    u32 zero;      // always 0 (see Zh)
    u64 *stack;    // reach: REACH_DEPTH 16-byte entries per thread of the grid
    u32 phase;     // 0: insert the cycles not yet in the requests' cycle-state tables; 1: the walks
};

// ------------------------------------------------------------------------------------------
// Is a distinct cycle entered by an ordinary deal? (cpu/reach_cycles.c:32-69) Positions of deal
// size hold ceil(n/2) and floor(n/2) cards. A position has at most two predecessors: undo a win of
// player w, whose bottom two cards must be (winner, loser) with winner > loser (rule 0): they go
// back on top of the two piles. Off the loop the predecessors form a finite tree, searched depth
// first from each loop position (skipping the loop's own predecessor) with a bounded per-thread
// stack in device memory. 1 = reached, 0 = not, -1 = undetermined (more than `cap` positions
// visited, or the stack is full).
constexpr u32 REACH_DEPTH = 64;

__device__ __forceinline__ bool deal_sized4(const St<4> &s, u32 n) {
    const u32 la = (s.sa >> 2) + 1u;
    return la == (n + 1u) / 2u || la == n / 2u;
}
__device__ __forceinline__ bool undo_win4(const St<4> &s, u32 w, u32 n, St<4> &o) {
    const u32 la = (s.sa >> 2) + 1u;
    const u32 L = w ? n - la : la;
    if (L < 2u) return false;
    const u64 word = w ? s.b : s.a;
    const u64 cw = (word >> (4u * (L - 2u))) & 15ull, cl = (word >> (4u * (L - 1u))) & 15ull;
    if (cw <= cl) return false;
    const u64 back = ((word & lowmask_<u64>(4u * (L - 2u))) << 4) | cw;
    if (w == 0u) { o.a = back; o.b = (s.b << 4) | cl; o.sa = s.sa - 4u; }
    else         { o.b = back; o.a = (s.a << 4) | cl; o.sa = s.sa + 4u; }
    return true;
}
template <int RULE>
__device__ int reach_search(const St<4> &s0, u32 lam, u32 n, u32 C, u64 cap, u64 *stk, Zh<4> &z, u32 &depth) {
    St<4> prev = s0;
    for (u32 i = 0; i + 1u < lam; i++) turn<4, RULE>(prev, C, z);
    St<4> t = s0;
    u64 visited = 0;
    for (u32 i = 0; i < lam; i++) {
        if (deal_sized4(t, n)) return 1;
        for (u32 w = 0; w < 2u; w++) {
            St<4> q;
            if (!undo_win4(t, w, n, q) || same<4>(q, prev)) continue;
            u32 sp = 0;
            stk[0] = q.a | ((u64)(q.sa >> 2) << 60); stk[1] = q.b; sp = 1;
            depth = depth > 1u ? depth : 1u;
            while (sp) {
                sp--;
                St<4> u;
                u.a = stk[2 * sp] & 0x0FFFFFFFFFFFFFFFull; u.sa = (u32)(stk[2 * sp] >> 60) << 2; u.b = stk[2 * sp + 1];
                if (deal_sized4(u, n)) return 1;
                if (++visited > cap) return -1;
                for (u32 w2 = 0; w2 < 2u; w2++) {
                    St<4> r;
                    if (!undo_win4(u, w2, n, r)) continue;
                    if (sp == REACH_DEPTH) return -1;
                    stk[2 * sp] = r.a | ((u64)(r.sa >> 2) << 60); stk[2 * sp + 1] = r.b; sp++;
                    depth = sp > depth ? sp : depth;
                }
            }
        }
        prev = t;
        turn<4, RULE>(t, C, z);
    }
    return 0;
}

// Insert key K into the open-addressing set; returns 1 if it was new. The probe length is bounded;
// a full table raises C_HOVF instead of looping.
__device__ __forceinline__ u32 hash_insert(const ReqDev &R, u64 K) {
    const u64 mask = (1ull << R.hlog2) - 1ull;
    u64 h = (K * 0x9E3779B97F4A7C15ull) >> (64 - R.hlog2);
    for (u32 probe = 0; probe < 4096u; probe++) {
        const u64 prev = atomicCAS((unsigned long long *)&R.htab[h], 0ull, (unsigned long long)K);
        if (prev == 0ull) return 1u;
        if (prev == K) return 0u;
        h = (h + 1ull) & mask;
    }
    atomicAdd((unsigned long long *)&R.cnt[C_HOVF], 1ull);
    return 0u;
}

// The cycle-state table of a canonical request with cycle walks (4-bit piles, n <= 14): every state of
// every cycle the request's positions met, as its key4 (exact for n <= 15; bits 0..3 are zero for
// n <= 14) with the cycle's shape verdict in bit 0 (STAB_NONSTD); 0 marks an empty slot. The walks'
// first pass walks each cycle not yet in the table once per record that finds it missing and inserts
// every state; the second pass, after the whole first pass, reads a record's shape verdict from its
// meeting state's entry and decides whether its start lies on its cycle by looking the start up (a
// state in the table lies on a cycle, and the start's cycle, if it is on one, is the meeting state's,
// which the first pass inserted whole). An insert that finds no free slot within STAB_PROBES counts
// C_SOVF, after which the second pass walks every record as without a table.
constexpr u64 STAB_NONSTD = 1ull, STAB_FLAGS = 15ull;
constexpr u32 STAB_PROBES = 256;
__device__ __forceinline__ u64 stab_slot(u64 K, u32 lg) { return (K * 0x9E3779B97F4A7C15ull) >> (64 - lg); }
// insert entry E (key | flags); false if no free slot was found
__device__ __forceinline__ bool stab_insert(const ReqDev &R, u64 E) {
    const u64 mask = (1ull << R.slog2) - 1ull, K = E & ~STAB_FLAGS;
    u64 h = stab_slot(K, R.slog2);
    for (u32 probe = 0; probe < STAB_PROBES; probe++) {
        const u64 prev = atomicCAS((unsigned long long *)&R.stab[h], 0ull, (unsigned long long)E);
        if (prev == 0ull || (prev & ~STAB_FLAGS) == K) return true;
        h = (h + 1ull) & mask;
    }
    return false;
}
// the entry of key K, or 0. Plain loads: the second pass runs after the first pass's kernel has
// completed, whose inserts it then sees; in the first pass a load that misses an insert made
// meanwhile elsewhere only makes the record walk and insert its cycle again.
__device__ __forceinline__ u64 stab_find(const ReqDev &R, u64 K) {
    const u64 mask = (1ull << R.slog2) - 1ull;
    u64 h = stab_slot(K, R.slog2);
    for (u32 probe = 0; probe < STAB_PROBES; probe++) {
        const u64 e = R.stab[h];
        if (e == 0ull) return 0ull;
        if ((e & ~STAB_FLAGS) == K) return e;
        h = (h + 1ull) & mask;
    }
    return 0ull;
}
__device__ __forceinline__ u64 key_of4(const St<4> &v) { return key4(v.a, v.b, bits1<4>(v)); }

// Add the number of lanes where pred holds to counter ctr of each such lane's request, one atomic per
// request among them (a wave's records may belong to several requests); every lane calls this. Per-
// lane atomics on one counter serialise: the n = 14 canonical shape test counts 4.4e7 positions.
__device__ __forceinline__ void wave_count(bool pred, u32 req, const ReqDev &R, u32 ctr) {
    u64 m = ballot(pred);
    while (m) {
        const u32 l = (u32)__ffsll((long long)m) - 1u;
        const u32 r = lanev(req, l);
        const u64 mr = ballot(pred && req == r);
        m &= ~mr;
        if (lane_id() == l) atomicAdd((unsigned long long *)&R.cnt[ctr], (unsigned long long)__popcll(mr));
    }
}

template <int W, int RULE>
__global__ __launch_bounds__(256) void k_walk(WalkArgs<W> A) {
    const u32 lane = lane_id();
    Zh<W> z = make_zh<W>(A.zero);
    const u64 T_ = A.hdr->ntiles;
    const u64 recs = A.pipe->ctotal;
    if (recs == 0) return;
    const u32 G = group_size(recs, T_, A.tile_pos);
    const u32 TP = A.tile_pos;
    const bool ins = A.phase == 0u;     // the first pass: the cycle-state tables' inserts
    for (;;) {
      const u64 g0 = wave_fetch_add(ins ? &A.pipe->wgrab0 : &A.pipe->wgrab, 1) * G;
      if (g0 >= T_) break;
      const u32 ng = (u32)((T_ - g0) < G ? (T_ - g0) : G);
      const u32 cc = lane < ng ? A.cq.cnt[g0 + lane] : 0u;
      const u32 c = cc < TP ? cc : TP;
      const u32 incl = wave_scan(c);
      const u32 excl = incl - c;
      const u32 total = lanev(incl, 63);
      for (u32 bb = 0; bb < total; bb += 64) {
        const u32 kk = bb + lane;
        const bool act = kk < total;
        u32 ti, off;
        group_locate(excl, ng, act ? kk : 0u, ti, off);
        const u64 i = (g0 + ti) * (u64)TP + off;
        St<W> m;
        u32 n = 0, req = 0, lam = 0;
        u64 pos = 0;
        if (act && off < TP) {
            m.a = A.cq.a[i]; m.b = A.cq.b[i];
            const u32 meta = A.cq.meta[i];
            m.sa = meta & 0xFFu; n = meta >> 8;
            req = A.cq.req[i]; lam = A.cq.lam[i]; pos = A.cq.pos[i];
        } else {
            m.a = 0; m.b = 0; m.sa = 0; n = 2;
        }
        const u32 B = W * n, C = B - 2u * W;
        ReqDev R;
        u32 blk = 0, mode = 0;
        bool dist = false;
        if (act) { R = A.reqs[req]; blk = R.shape_blk; mode = R.mode; dist = (R.dbins || R.distinct) && mode == 2; }
        // the record's request keeps a cycle-state table (and it has not overflowed: read after the
        // first pass, in the second)
        bool tab = false;
        if constexpr (W == 4) tab = act && R.stab != nullptr && lam <= 512u && (ins || R.cnt[C_SOVF] == 0ull);

        if (ins) {
            // first pass: a record whose meeting state is not in the table walks its cycle for the
            // shape verdict (if asked), then walks it again inserting every state with the verdict
            if constexpr (W == 4) {
                const bool miss = tab && stab_find(R, key_of4(m)) == 0ull;
                if (ballot(miss)) {
                    const u32 L = miss ? lam : 0u;
                    u32 maxL = L;
                    for (int o = 32; o > 0; o >>= 1) { const u32 x = __shfl_xor(maxL, o); maxL = x > maxL ? x : maxL; }
                    u64 fl = 0;
                    if (ballot(miss && blk != 0)) {
                        u32 first = 0, cur = 0, run = 0, first_run = 0;
                        bool changed = false, bad = false;
                        St<W> v = m;
                        open_sa<W>(v, A.zero);
                        for (u32 t = 0; t < maxL; t++) {
                            const u32 w = turn<W, RULE>(v, C, z) ? 1u : 0u;
                            if (t < L) {
                                if (t == 0) { first = w; cur = w; run = 1; }
                                else if (w == cur) run++;
                                else {
                                    if (!changed) { first_run = run; changed = true; }
                                    else if (run != blk) bad = true;
                                    cur = w; run = 1;
                                }
                            }
                        }
                        u32 shp;
                        if (!changed || bad) shp = 1;
                        else if (cur == first) shp = (run + first_run == blk) ? 0u : 1u;
                        else shp = (first_run == blk && run == blk) ? 0u : 1u;
                        if (miss && blk != 0 && shp) fl = STAB_NONSTD;
                    }
                    St<W> v = m;
                    open_sa<W>(v, A.zero);
                    bool full = false;
                    for (u32 t = 0; t < maxL; t++) {
                        if (t < L) {
                            St<W> u = v;
                            u.sa &= ~3u;
                            full |= !stab_insert(R, key_of4(u) | fl);
                            turn<W, RULE>(v, C, z);
                        }
                    }
                    if (full) atomicAdd((unsigned long long *)&R.cnt[C_SOVF], 1ull);
                }
            }
            continue;
        }
        u64 ment = 0;                  // the meeting state's entry: its cycle is in the table whole
        if constexpr (W == 4) if (tab) ment = stab_find(R, key_of4(m));
        const bool known = ment != 0ull;
        wave_count(known && blk != 0 && (ment & STAB_NONSTD) != 0ull, req, R, C_NONSTD);

        // (1) loop shape from the meeting state (gpu/war_gpu.py:51-73), online over the winners
        if (ballot(act && blk != 0 && !known)) {
            const bool doit = act && blk != 0 && !known && lam <= 512u;
            u32 first = 0, cur = 0, run = 0, first_run = 0;
            bool changed = false, bad = false;
            u32 L = doit ? lam : 0u;
            u32 maxL = L;
            for (int off = 32; off > 0; off >>= 1) { const u32 o = __shfl_xor(maxL, off); maxL = o > maxL ? o : maxL; }
            St<W> v = m;
            open_sa<W>(v, A.zero);
            for (u32 t = 0; t < maxL; t++) {
                const u32 w = turn<W, RULE>(v, C, z) ? 1u : 0u;
                if (t < L) {
                    if (t == 0) { first = w; cur = w; run = 1; }
                    else if (w == cur) run++;
                    else {
                        if (!changed) { first_run = run; changed = true; }
                        else if (run != blk) bad = true;
                        cur = w; run = 1;
                    }
                }
            }
            u32 shp = 0;
            if (doit) {
                if (!changed) shp = 1;                         // one player wins every turn
                else if (bad) shp = 1;
                else if (cur == first) shp = (run + first_run == blk) ? 0u : 1u;
                else shp = (first_run == blk && run == blk) ? 0u : 1u;
            }
            wave_count(doit && shp != 0u, req, R, C_NONSTD);
            wave_count(act && blk != 0 && !known && !doit, req, R, C_UNCHK);
        }

        // (2) distinct: does the start lie on its cycle? (lam turns from the start return to it)
        if constexpr (W == 4) {
            if (ballot(dist)) {
                St<W> s0;
                s0.a = 0; s0.b = 0; s0.sa = 0;
                if (dist) {
                    const u32 mm = n - 1;
                    const u64 r = pos / mm;
                    const int kk = (int)(pos - r * mm);
                    s0 = canon<W>(unrank<W>(r, (int)mm), (int)n, kk);
                }
                // with the cycle in the table the start lies on its cycle iff the table holds it;
                // otherwise lam turns from the start come back to it
                const bool look = dist && known;
                const bool onk = look && stab_find(R, key_of4(s0)) != 0ull;
                const u32 L = (dist && !known) ? lam : 0u;
                u32 maxL = L;
                for (int off = 32; off > 0; off >>= 1) { const u32 o = __shfl_xor(maxL, off); maxL = o > maxL ? o : maxL; }
                open_sa<W>(s0, A.zero);
                St<W> v = s0;
                for (u32 t = 0; t < maxL; t++) if (t < L) turn<W, RULE>(v, C, z);
                const bool onc = look ? onk : (dist && same<W>(v, s0));
                if (ballot(onc)) {
                    // (3) weight, minimum key and hash insert for positions on their cycle
                    u64 c1n = 0, c2n = 0, K = ~0ull;
                    bool sym = false;
                    const u32 L2 = onc ? lam : 0u;
                    u32 maxL2 = L2;
                    for (int off = 32; off > 0; off >>= 1) { const u32 o = __shfl_xor(maxL2, off); maxL2 = o > maxL2 ? o : maxL2; }
                    v = s0;
                    const u64 top = (u64)(n - 1);
                    for (u32 t = 0; t < maxL2; t++) {
                        if (t < L2) {
                            const u32 b1 = bits1<W>(v);
                            if (b1 > 0 && (v.a & 15ull) == top) { c1n++; const u64 kk = key4(v.a, v.b, b1); K = kk < K ? kk : K; }
                            if (b1 < B && (v.b & 15ull) == top) { c2n++; const u64 kk = key4(v.b, v.a, B - b1); K = kk < K ? kk : K; }
                            if (v.a == s0.b && v.b == s0.a && b1 == B - bits1<W>(s0)) sym = true;
                            turn<W, RULE>(v, C, z);
                        }
                    }
                    if (onc) {
                        atomicAdd((unsigned long long *)&R.cnt[C_ONCYCLE], 1ull);
                        if (R.dbins) {
                            const u64 d = sym ? c1n : c1n + c2n;
                            atomicAdd((unsigned long long *)&R.dhist[d < (u64)(R.dbins - 1) ? d : (u64)(R.dbins - 1)], 1ull);
                        }
                        if (R.distinct) {
                            const bool rep = key4(s0.a, s0.b, bits1<W>(s0)) == K;
                            if (rep) atomicAdd((unsigned long long *)&R.cnt[C_MINKEY], 1ull);
                            if (hash_insert(R, K)) atomicAdd((unsigned long long *)&R.cnt[C_HASH], 1ull);
                            if (rep && R.reach && RULE == 0) {
                                const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x;
                                u32 depth = 0;
                                const int res = reach_search<RULE>(s0, lam, n, C, R.reach_cap,
                                                                   A.stack + gtid * (2u * REACH_DEPTH), z, depth);
                                atomicMax((unsigned long long *)&R.cnt[C_RDEPTH], (unsigned long long)depth);
                                atomicAdd((unsigned long long *)&R.cnt[res == 1 ? C_RYES : res == 0 ? C_RNO : C_RUND], 1ull);
                                if (res == 0) atomicAdd((unsigned long long *)&R.nolen[lam < 4095u ? lam : 4095u], 1ull);
                            }
                        }
                    }
                }
            }
        }
    }
    }
}

}  // namespace war
