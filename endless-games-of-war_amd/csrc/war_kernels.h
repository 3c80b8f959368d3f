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
//                   plus three compares, and all accounting is scalar. A stage plays ages
//                   (a0, a1] with a1 = 2^k - 1; lanes still running at a1 are compacted into the
//                   next stage's queue (the tortoise equals the hare there, so a record is just the
//                   state). The FRESH stage enumerates positions from the job list (a global tile
//                   counter refills waves); later stages drain queues (a global record counter).
//                   The last stage has no cap (Brent's exactness), only a 2^24-turn safety stop
//                   that counts as "undetermined" and is reported as an error.
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
    C_HASH, C_HOVF, C_QOVF, C_N = 16
};

constexpr u32 AGE_CAP = 1u << 24;      // safety stop of the last stage (never reached by a valid game)
constexpr u32 UNBOUNDED = 0xFFFFFFFFu;
constexpr int MAXSTAGE = 16;

struct ReqDev {
    u32 n, mode, rule, hbins, lbins, dbins, distinct, shape_blk;
    u32 walk, hlog2, pad0, pad1;
    u64 *cnt, *hist, *lenhist, *dhist, *htab;
};

struct Job { u64 lo, hi, tile_base; u32 req, pad; };
struct JobHdr { u32 njobs, req0; u64 ntiles; };   // req0: the request of job 0

struct Pipe {
    u64 tile_ctr;
    u64 qcount[MAXSTAGE];   // records written by stage s
    u64 qread[MAXSTAGE];    // read cursor of stage s over the records of stage s-1
    u64 ccount, cread;      // cycle-record queue (walk requests)
    u64 pad[6];
};

// Survivor records. The pile lengths and n ride in the unused top bits of the packed piles
// (4-bit: an alive pile holds <= 15 cards = 60 bits; 5-bit: <= 19 cards = 95 bits), so a record is
// 16 B (4-bit) or 32 B (5-bit). req / pos are written only when the batch needs them.
template <int W> struct Queue {
    typedef typename Lay<W>::T T;
    T *a, *b;
    u32 *req;
    u64 *pos;
};

template <int W> __device__ __forceinline__ void pack(const St<W> &s, u32 n, typename Lay<W>::T &qa, typename Lay<W>::T &qb) {
    typedef typename Lay<W>::T T;
    if constexpr (W == 4) {
        qa = s.a | ((u64)(s.sa >> 2) << 60);
        qb = s.b | ((u64)(n - 1) << 60);
    } else {
        qa = s.a | ((T)(s.sa | (n << 8)) << 96);
        qb = s.b;
    }
}
template <int W> __device__ __forceinline__ void unpack(typename Lay<W>::T qa, typename Lay<W>::T qb, St<W> &s, u32 &n) {
    typedef typename Lay<W>::T T;
    if constexpr (W == 4) {
        s.a = qa & 0x0FFFFFFFFFFFFFFFull; s.sa = (u32)(qa >> 60) << 2;
        s.b = qb & 0x0FFFFFFFFFFFFFFFull; n = (u32)(qb >> 60) + 1u;
    } else {
        const u32 meta = (u32)(qa >> 96);
        s.a = qa & lowmask_<T>(96); s.sa = meta & 0xFFu; n = meta >> 8;
        s.b = qb;
    }
}

template <int W> struct CycQueue {
    typedef typename Lay<W>::T T;
    T *a, *b;    // meeting state
    u32 *meta;   // sa | n << 8
    u32 *req;
    u32 *lam;
    u64 *pos;    // position index (the start is rebuilt from it)
};

template <int W> struct StageArgs {
    Pipe *pipe;
    const JobHdr *hdr;
    const Job *jobs;
    const ReqDev *reqs;
    Queue<W> qin, qout;
    CycQueue<W> cq;
    u64 qcap, cqcap;
    u32 tile_pos, per;     // FRESH: positions per tile, positions per lane (tile_pos = 64 * per)
    u32 a0, a1;            // ages played: (a0, a1]
    int stage;             // index of the counters this stage writes (qcount[stage])
    u32 hb, lb;            // LDS histogram bins per wave
    u32 need_req, need_pos;  // batch holds several requests / a request that needs cycle walks
};

// ------------------------------------------------------------------------------------------
__device__ __forceinline__ u32 lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
__device__ __forceinline__ u64 ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ u32 rank_in(u64 mask) {   // set bits of mask below this lane
    return __builtin_amdgcn_mbcnt_hi((u32)(mask >> 32), __builtin_amdgcn_mbcnt_lo((u32)mask, 0u));
}
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

__global__ void k_setup(Pipe *pipe) {
    u64 *p = reinterpret_cast<u64 *>(pipe);
    for (u32 i = threadIdx.x; i < sizeof(Pipe) / 8; i += blockDim.x) p[i] = 0;
}

// ------------------------------------------------------------------------------------------
// Scalar (wave-uniform) accounting of one request, with wave-private LDS histograms.
struct WaveAcc {
    u32 req;                   // request slot these totals belong to (UNBOUNDED = none)
    u32 hb, lb;                // that request's bins
    u64 cyc, done, turns, maxfin, undet;
    u32 *lh, *ll;              // LDS: hist[hbmax], lenhist[lbmax] of this wave
    bool dirty;

    __device__ void init(u32 *lds, u32 hbmax, u32 lbmax) {
        req = UNBOUNDED; hb = lb = 0; cyc = done = turns = maxfin = undet = 0; dirty = false;
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
};

// Append the lanes of `mask` to the cycle-record queue.
template <int W>
__device__ __forceinline__ void push_cycles(const StageArgs<W> &A, u64 mask, const St<W> &s, u32 n, u32 req,
                                            u32 lam, u64 pos) {
    const u32 c = __popcll(mask);
    const u64 base = wave_fetch_add(&A.pipe->ccount, c);
    if ((mask >> lane_id()) & 1ull) {
        const u64 i = base + rank_in(mask);
        if (i < A.cqcap) {
            A.cq.a[i] = s.a; A.cq.b[i] = s.b; A.cq.meta[i] = s.sa | (n << 8);
            A.cq.req[i] = req; A.cq.lam[i] = lam; A.cq.pos[i] = pos;
        }
    }
}

// ------------------------------------------------------------------------------------------
// The age-synchronous stage. FRESH: positions come from tiles of the job list. Otherwise: records
// written by the previous stage.
template <int W, int RULE, bool FRESH>
__global__ __launch_bounds__(256) void k_stage(StageArgs<W> A) {
    typedef typename Lay<W>::T T;
    extern __shared__ u32 lds[];
    const u32 wid = threadIdx.x >> 6;
    const u32 lane = lane_id();
    WaveAcc acc;
    acc.init(lds + wid * (A.hb + A.lb), A.hb, A.lb);
    const ReqDev *reqs = A.reqs;
    const bool last = (A.a1 == UNBOUNDED);
    const u64 *qcount_in = FRESH ? nullptr : &A.pipe->qcount[A.stage - 1];

    for (;;) {
        // ---- obtain work: per lane a start state, its request, its position index
        St<W> s;
        u32 C = 0, n = 0, req = 0;
        u64 pos = 0;
        u64 alive = 0;
        // FRESH bookkeeping
        u32 per = 1, cnt = 0, mode = 0;
        T q = 0;
        int k = 0;
        bool uniform_req = true, walk = false;
        if constexpr (FRESH) {
            const u64 t = wave_fetch_add(&A.pipe->tile_ctr, 1);
            if (t >= A.hdr->ntiles) break;
            // the job holding tile t (jobs are few; binary search on tile_base)
            u32 lo = 0, hi = A.hdr->njobs;
            while (hi - lo > 1) { u32 mid = (lo + hi) >> 1; if (A.jobs[mid].tile_base <= t) lo = mid; else hi = mid; }
            const Job J = A.jobs[lo];
            req = J.req;
            const ReqDev &R = reqs[req];
            n = R.n; mode = R.mode; walk = R.walk != 0;
            C = W * n - 2u * W;
            per = A.per;
            const u64 first = J.lo + (t - J.tile_base) * (u64)A.tile_pos + (u64)lane * per;
            cnt = first < J.hi ? (u32)((J.hi - first) < per ? (J.hi - first) : per) : 0u;
            pos = first;
            if (cnt) {
                if (mode == 2) { const u32 m = n - 1; const u64 r = first / m; k = (int)(first - r * m); q = unrank<W>(r, (int)m); }
                else q = unrank<W>(first, (int)n);
            }
            acc.select(req, reqs);
        } else {
            const u64 base = wave_fetch_add(&A.pipe->qread[A.stage], 64);
            const u64 total = *qcount_in < A.qcap ? *qcount_in : A.qcap;
            if (base >= total) break;
            const u64 i = base + lane;
            const bool act = i < total;
            if (act) {
                unpack<W>(A.qin.a[i], A.qin.b[i], s, n);
                req = A.need_req ? A.qin.req[i] : A.hdr->req0;
                pos = A.need_pos ? A.qin.pos[i] : 0ull;
            } else {
                s.a = 0; s.b = 0; s.sa = 0; n = 2; req = 0;
            }
            C = W * n - 2u * W;
            alive = ballot(act);
            const u32 r0 = lanev(req, (u32)(__ffsll((long long)alive) - 1));
            uniform_req = (ballot(req == r0) & alive) == alive;
            if (uniform_req) acc.select(r0, reqs);
        }

        // ---- FRESH: loop over each lane's `per` positions; otherwise a single round
        for (u32 j = 0;; j++) {
            if constexpr (FRESH) {
                if (j >= per || ballot(j < cnt) == 0) break;
                bool act = j < cnt;
                if (act) {
                    s = (mode == 2) ? canon<W>(q, (int)n, k) : deal<W>(q, (int)n);
                    if (mode == 1) act = top_card_to_p1<W>(q, (int)n);
                }
                u64 act_m = ballot(act);
                // a position that is over before it starts: rounds 0 (gpu/war_gpu.py:80)
                const u64 e0 = ballot(act && (n < 2 || ended<W>(s, C))) & act_m;
                if (e0) acc.ended(__popcll(e0), 0);
                alive = act_m & ~e0;
            }

            // ---- the lockstep game
            if (alive) {
                St<W> t = s;
                u32 R = A.a0;
                u32 age = A.a0;
                const u32 a1 = last ? AGE_CAP : A.a1;
                for (;;) {
                    ++age;
                    turn<W, RULE>(s, C);
                    const u64 em = ballot(ended<W>(s, C)) & alive;
                    // three ballots of plain compares: each is one v_cmp into an SGPR pair
                    const u64 dm = ballot(s.a == t.a) & ballot(s.b == t.b) & ballot(s.sa == t.sa) & alive & ~em;
                    if (em | dm) {
                        const u32 lam = age - R;
                        u64 rem = em | dm;
                        if (uniform_req) {
                            if (em) acc.ended(__popcll(em), age);
                            if (dm) acc.cycled(__popcll(dm), age, lam);
                        } else {
                            while (rem) {
                                const u32 r = lanev(req, (u32)(__ffsll((long long)rem) - 1));
                                const u64 mr = ballot(req == r) & rem;
                                rem &= ~mr;
                                acc.select(r, reqs);
                                if (mr & em) acc.ended(__popcll(mr & em), age);
                                if (mr & dm) acc.cycled(__popcll(mr & dm), age, lam);
                            }
                        }
                        if (dm) {
                            bool need = false;
                            if constexpr (FRESH) need = walk;
                            else need = reqs[req].walk != 0;
                            const u64 wm = ballot(need) & dm;
                            if (wm) push_cycles<W>(A, wm, s, n, req, lam, pos);
                        }
                        alive &= ~(em | dm);
                        if (!alive) break;
                    }
                    if (age == a1) break;
                    if ((age & (age + 1u)) == 0u) { t = s; R = age; }   // Brent: tortoise := hare
                }
                if (alive) {
                    if (last) {
                        // safety stop reached: count as undetermined (an error for the caller)
                        if (uniform_req) { acc.undet += __popcll(alive); acc.done += __popcll(alive); acc.dirty = true; }
                        else {
                            u64 rem = alive;
                            while (rem) {
                                const u32 r = lanev(req, (u32)(__ffsll((long long)rem) - 1));
                                const u64 mr = ballot(req == r) & rem;
                                rem &= ~mr;
                                acc.select(r, reqs);
                                acc.undet += __popcll(mr); acc.done += __popcll(mr); acc.dirty = true;
                            }
                        }
                    } else {
                        // compact the survivors into the next stage's queue
                        const u64 base = wave_fetch_add(&A.pipe->qcount[A.stage], __popcll(alive));
                        if ((alive >> lane) & 1ull) {
                            const u64 i = base + rank_in(alive);
                            if (i < A.qcap) {
                                T qa, qb;
                                pack<W>(s, n, qa, qb);
                                A.qout.a[i] = qa; A.qout.b[i] = qb;
                                if (A.need_req) A.qout.req[i] = req;
                                if (A.need_pos) A.qout.pos[i] = pos;
                            }
                        }
                    }
                }
            }

            if constexpr (FRESH) {
                if (j + 1 < cnt) {
                    if (mode == 2) { if (++k == (int)n - 1) { k = 0; q = next_perm<W>(q, (int)n - 1); } }
                    else q = next_perm<W>(q, (int)n);
                }
                pos++;
            } else {
                break;
            }
        }
    }
    acc.flush(reqs);
}

// ------------------------------------------------------------------------------------------
// Plain kernel: the reference's structure. Lane g plays positions [lo + g*per, lo + (g+1)*per),
// one at a time, each with its own Brent loop (per-lane tortoise, power and lam).
template <int W> struct PlainArgs {
    u64 lo, hi;
    u32 per, hb, lb, pad;
    ReqDev R;
    u32 req;
    Pipe *pipe;
    CycQueue<W> cq;
    u64 cqcap;
};

template <int W, int RULE>
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
    u64 cyc = 0, done = 0, maxfin = 0, turns = 0, undet = 0;
    for (u32 j = 0; j < cnt; j++) {
        St<W> s = (mode == 2) ? canon<W>(q, (int)n, k) : deal<W>(q, (int)n);
        const bool take = (mode == 1) ? top_card_to_p1<W>(q, (int)n) : true;
        bool cy = false;
        u32 lam = 0;
        u64 rounds = 0;
        if (take && n >= 2 && !ended<W>(s, C)) {
            St<W> t = s;
            u32 power = 1, l = 0;
            for (;;) {
                turn<W, RULE>(s, C);
                rounds++;
                if (ended<W>(s, C)) break;
                l++;
                if (same<W>(s, t)) { cy = true; lam = l; break; }
                if (l == power) { t = s; power <<= 1; l = 0; }
                if (rounds >= AGE_CAP) { undet++; break; }
            }
        }
        if (take) {
            done++;
            turns += rounds;
            if (cy) {
                cyc++;
                if (P.hb) atomicAdd(&lh[lam < P.hb - 1 ? lam : P.hb - 1], 1u);
            } else if (rounds < AGE_CAP) {
                if (rounds > maxfin) maxfin = rounds;
                if (P.lb) atomicAdd(&ll[rounds < P.lb - 1 ? (u32)rounds : P.lb - 1], 1u);
            }
        }
        // cycle records for the walks (reconverged here: wave-aggregated append)
        if (P.R.walk) {
            const u64 wm = ballot(cy);
            if (wm) {
                const u64 base = wave_fetch_add(&P.pipe->ccount, __popcll(wm));
                if (cy) {
                    const u64 i = base + rank_in(wm);
                    if (i < P.cqcap) {
                        P.cq.a[i] = s.a; P.cq.b[i] = s.b; P.cq.meta[i] = s.sa | (n << 8);
                        P.cq.req[i] = P.req; P.cq.lam[i] = lam; P.cq.pos[i] = first + j;
                    }
                }
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
    }
    if (lane_id() == 0 && done) {
        atomicAdd((unsigned long long *)&P.R.cnt[C_CYC], (unsigned long long)cyc);
        atomicAdd((unsigned long long *)&P.R.cnt[C_DONE], (unsigned long long)done);
        atomicAdd((unsigned long long *)&P.R.cnt[C_TURNS], (unsigned long long)turns);
        if (undet) atomicAdd((unsigned long long *)&P.R.cnt[C_UNDET], (unsigned long long)undet);
        atomicMax((unsigned long long *)&P.R.cnt[C_MAXFIN], (unsigned long long)maxfin);
    }
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
    const ReqDev *reqs;
    CycQueue<W> cq;
    u64 cqcap;
};

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

template <int W, int RULE>
__global__ __launch_bounds__(256) void k_walk(WalkArgs<W> A) {
    typedef typename Lay<W>::T T;
    const u32 lane = lane_id();
    for (;;) {
        const u64 base = wave_fetch_add(&A.pipe->cread, 64);
        const u64 total = A.pipe->ccount < A.cqcap ? A.pipe->ccount : A.cqcap;
        if (base >= total) break;
        const u64 i = base + lane;
        const bool act = i < total;
        St<W> m;
        u32 n = 0, req = 0, lam = 0;
        u64 pos = 0;
        if (act) {
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

        // (1) loop shape from the meeting state (gpu/war_gpu.py:51-73), online over the winners
        if (ballot(act && blk != 0)) {
            const bool doit = act && blk != 0 && lam <= 512u;
            u32 first = 0, cur = 0, run = 0, first_run = 0;
            bool changed = false, bad = false;
            u32 L = doit ? lam : 0u;
            u32 maxL = L;
            for (int off = 32; off > 0; off >>= 1) { const u32 o = __shfl_xor(maxL, off); maxL = o > maxL ? o : maxL; }
            St<W> v = m;
            for (u32 t = 0; t < maxL; t++) {
                const u32 w = turn<W, RULE>(v, C) ? 1u : 0u;
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
            if (doit) {
                u32 shp;
                if (!changed) shp = 1;                         // one player wins every turn
                else if (bad) shp = 1;
                else if (cur == first) shp = (run + first_run == blk) ? 0u : 1u;
                else shp = (first_run == blk && run == blk) ? 0u : 1u;
                if (shp) atomicAdd((unsigned long long *)&R.cnt[C_NONSTD], 1ull);
            } else if (act && blk != 0) {
                atomicAdd((unsigned long long *)&R.cnt[C_UNCHK], 1ull);
            }
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
                const u32 L = dist ? lam : 0u;
                u32 maxL = L;
                for (int off = 32; off > 0; off >>= 1) { const u32 o = __shfl_xor(maxL, off); maxL = o > maxL ? o : maxL; }
                St<W> v = s0;
                for (u32 t = 0; t < maxL; t++) if (t < L) turn<W, RULE>(v, C);
                const bool onc = dist && same<W>(v, s0);
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
                            turn<W, RULE>(v, C);
                        }
                    }
                    if (onc) {
                        atomicAdd((unsigned long long *)&R.cnt[C_ONCYCLE], 1ull);
                        if (R.dbins) {
                            const u64 d = sym ? c1n : c1n + c2n;
                            atomicAdd((unsigned long long *)&R.dhist[d < (u64)(R.dbins - 1) ? d : (u64)(R.dbins - 1)], 1ull);
                        }
                        if (R.distinct) {
                            if (key4(s0.a, s0.b, bits1<W>(s0)) == K) atomicAdd((unsigned long long *)&R.cnt[C_MINKEY], 1ull);
                            if (hash_insert(R, K)) atomicAdd((unsigned long long *)&R.cnt[C_HASH], 1ull);
                        }
                    }
                }
            }
        }
    }
}

}  // namespace war
