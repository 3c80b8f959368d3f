// War enumeration oracle: a CPU restatement (C++17 + OpenMP) of the reference's hot path.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product (the HIP engine, its C ABI or the Python
// drop-in) links, loads or calls this file. It is used by the test suite as the checker, by
// the smoke check, and by the benchmark's CPU-baseline leg.
//
// What it restates (reference file:line):
//   turn ........................ gpu/war_gpu.py:38-45 (WAR_STEP), cpu/war.cpp:17-23 (step)
//   Brent cycle test ............ gpu/war_gpu.py:77-96 (war_sim), cpu/war.cpp:26-37 (run)
//   deal -> piles ............... gpu/war_gpu.py:98-107, cpu/war.cpp:48-49
//   lexicographic unrank ........ gpu/war_gpu.py:135-146 and :196-206
//   next permutation ............ gpu/war_gpu.py:109-119 (std::next_permutation in war.cpp:53)
//   deal counter (+ --half) ..... gpu/war_gpu.py:121-176, cpu/war.cpp:38-60
//   canonical positions ......... gpu/war_gpu.py:183-259, cpu/war_all.cpp:47-51
//   distinct weights (dhist) .... gpu/war_gpu.py:217-235, :387-392
//   distinct by minimum key ..... cpu/distinct_cycles.c:24-31 (key), :57-80
//   loop shape (Theorem-3 test) . gpu/war_gpu.py:51-73, :238-242
//   5-bit piles (n >= 17) ....... gpu/war24.py:31-47 (only the layout; the game is the same)
//   streak test (odd n) ......... gpu/war24.py:98-125
//   cycle length / pre-period ... cpu/cycle_length.c:16-30,114-126
//   reach by a deal ............. cpu/reach_cycles.c:32-69,73-107
// and two detectors of this engine's own (alt_brent, odd n; pair_brent, n = 2 mod 4), pinned
// against Brent.
//
// Extensions the reference does not have (pinned by invariants in the tests, not by goldens):
// arbitrary rank / position ranges, the histogram of ending-game lengths (lenhist), the total
// number of Brent turns (turns), and the count of distinct cycle keys touched by a range.
//
// Range units: full / half mode use deal ranks r in [0, n!); canon mode uses positions
// p = r*(n-1) + k with r in [0, (n-1)!) and split k in [0, n-2] (k = cards below P1's top).
//
// Build: see oracle/Makefile (g++ -O3 -march=native -fopenmp -shared -fPIC).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <vector>
#include <omp.h>

typedef unsigned long long u64;
typedef unsigned __int128 u128;

enum { ORC_FULL = 0, ORC_HALF = 1, ORC_CANON = 2 };

extern "C" {
struct orc_res {
    u64 cyc;              // endless deals / cycling positions
    u64 done;             // positions played (half mode: the deals taken)
    u64 maxfin;           // longest ending game, in turns
    u64 turns;            // sum over positions of Brent's rounds (game length, or turns to detection)
    u64 oncycle;          // canon + distinct: positions that lie on their own cycle
    u64 distinct_minkey;  // canon + distinct: positions whose key is their cycle's minimum key
    u64 distinct_keys;    // canon + distinct: number of different cycle-minimum keys met in the range
    u64 shape_nonstd;     // canon + shape: cycles whose winner runs are not all of length blk
    u64 shape_unchk;      // canon + shape: cycles longer than 512 turns (not checked)
    u64* hist;  int hbins;    // lambda histogram of endless positions (last bin = overflow)
    u64* lenhist; int lbins;  // histogram of ending-game lengths (last bin = overflow)
    u64* dhist; int dbins;    // distinct weights (canon + distinct)
    u64 undet;                // games not resolved within the step cap (orc_run_cap)
};
}

// ---------------------------------------------------------------------------------------------
// Pile layouts. Card i of a pile (0 = top) sits at bit W*i. W = 4 for n <= 16 (u64), 5 for
// n <= 25 (u128). The game rules do not depend on the layout.
template <int W> struct Lay;
template <> struct Lay<4> { typedef u64 T; };
template <> struct Lay<5> { typedef u128 T; };

template <int W> struct St {
    typedef typename Lay<W>::T T;
    T a, b;   // P1, P2
    int la, lb;
};

template <int W> static inline bool same(const St<W>& x, const St<W>& y) {
    // (w1, w2, l1) as in war_gpu.py:91 / war.cpp:24; l1 is needed because card 0 and an empty
    // slot look alike.
    return x.a == y.a && x.b == y.b && x.la == y.la;
}

// One turn (war_gpu.py:38-45). Returns 1 if player 1 won the turn.
template <int W> static inline int step(St<W>& s, int rule) {
    typedef typename Lay<W>::T T;
    const T M = (T)((1u << W) - 1u);
    T c1 = s.a & M, c2 = s.b & M;
    s.a >>= W; s.b >>= W; s.la--; s.lb--;
    if (c1 > c2) {
        T x = rule ? c2 : c1, y = rule ? c1 : c2;
        s.a |= x << (W * s.la); s.la++;
        s.a |= y << (W * s.la); s.la++;
        return 1;
    }
    T x = rule ? c1 : c2, y = rule ? c2 : c1;
    s.b |= x << (W * s.lb); s.lb++;
    s.b |= y << (W * s.lb); s.lb++;
    return 0;
}

// Brent (war_gpu.py:77-96). Returns 1 if the game never ends, 0 if it ends; rounds = turns played
// until the end or until the repeat is seen; lam = cycle length; meet = state at detection. With a
// step cap (not in the reference, whose Brent has none): 2 = still unresolved after cap turns
// (rounds = cap).
template <int W> static inline int brent(St<W> s, int rule, u64& rounds, u64& lam_out, St<W>* meet, u64 cap = 0) {
    if (s.la == 0 || s.lb == 0) { rounds = 0; return 0; }
    St<W> t = s;
    u64 power = 1, lam = 0, r = 0;
    for (;;) {
        step<W>(s, rule);
        r++;
        if (s.la == 0 || s.lb == 0) { rounds = r; return 0; }
        lam++;
        if (same<W>(s, t)) { rounds = r; lam_out = lam; if (meet) *meet = s; return 1; }
        if (lam == power) { t = s; power <<= 1; lam = 0; }
        if (cap && r >= cap) { rounds = r; return 2; }
    }
}

// Streak test for odd n (gpu/war24.py:98-125, Spivey Thm 6): anchor the state at a turn whose
// winner-to-be holds an even number ws of cards while the other holds ls >= 2; if the players then
// alternate for lcm(n+1, ws, ls-1) turns and the state equals the anchor, the game never ends.
// A repeated winner drops the anchor. Returns 0 = ends, 1 = endless, 2 = undetermined (still running
// with more than cap turns played at the top of the loop: the reference's CAP = 2^22,
// gpu/war24.py:57,103); steps = turns played (the game length for a game that ends).
static inline int gcd_i(int a, int b) { while (b) { int t = a % b; a = b; b = t; } return a; }
static inline int lcm_i(int a, int b) { return a / gcd_i(a, b) * b; }
template <int W> static inline int streak(St<W> s, int n, u64& steps, u64 cap = 0) {
    typedef typename Lay<W>::T T;
    const T M = (T)((1u << W) - 1u);
    const u64 CAP = cap ? cap : (1ULL << 22);
    steps = 0;
    bool anch = false; int lam = 0, since = 0, prevw = -1;
    St<W> a = s;
    for (;;) {
        if (s.la == 0 || s.lb == 0) return 0;
        if (steps > CAP) return 2;
        if (!anch) {
            int ta = (int)(s.a & M), tb = (int)(s.b & M);
            int ws = ta > tb ? s.la : s.lb, ls = n - ws;
            if ((ws & 1) == 0 && ls >= 2) {
                anch = true; since = 0; prevw = -1;
                lam = lcm_i(n + 1, lcm_i(ws, ls - 1));
                a = s;
            }
        }
        int w = step<W>(s, 0);
        steps++;
        if (anch) {
            if (w == prevw) anch = false;
            else {
                prevw = w;
                if (++since == lam) {
                    if (same<W>(s, a)) return 1;
                    anch = false;
                }
            }
        }
    }
}

// Brent with alternation checkpoints (this engine's detector for odd n, rule 0; not in the
// reference). In a cycle of odd-n War the players alternate winning (Spivey Thm 6, the fact the
// reference's streak test rests on, gpu/war24.py:9-17). Every ALT_K turns (ages multiple of
// ALT_K), a game one of whose last ALT_K turns of odd age repeated the previous turn's winner
// cannot have been on its cycle for those turns, so its tortoise jumps to the present state; a
// tortoise that has waited ALT_L0 turns (then twice as long, and so on) jumps too, as in Brent.
// A repeat is only ever declared on an exact state match, and the first match after a jump is the
// cycle length, so every output is exact; the turns to the repeat are fewer than Brent's (about
// 30 % at n = 13).
constexpr u64 ALT_K = 8, ALT_L0 = 128;
template <int W> static inline int alt_brent(St<W> s, u64& rounds, u64& lam_out, St<W>* meet, u64 cap = 0) {
    if (s.la == 0 || s.lb == 0) { rounds = 0; return 0; }
    St<W> t = s;
    u64 R = 0, age = 0, L = ALT_L0;
    bool streak = false;
    int prevw = 1;          // before the first turn, player 1 counts as the last winner
    for (;;) {
        const int w = step<W>(s, 0);
        age++;
        if (s.la == 0 || s.lb == 0) { rounds = age; return 0; }
        if (same<W>(s, t)) { rounds = age; lam_out = age - R; if (meet) *meet = s; return 1; }
        if (w == prevw && (age & 1)) streak = true;
        prevw = w;
        if (age % ALT_K == 0) {
            if (age - R >= L) { t = s; R = age; L *= 2; }
            else if (streak) { t = s; R = age; }
            streak = false;
        }
        if (cap && age >= cap) { rounds = age; return 2; }
    }
}

// Brent with pair-run checkpoints (this engine's detector for n = 2 mod 4, rule 0; not in the
// reference). The reference's loop-shape test (gpu/war_gpu.py:51-73, Theorem 3) finds the cycles of
// such n made of winner runs of exactly n & -n = 2 turns, all but a few (at n = 14 the canonical
// positions of every cycle but those of length 96, data/long_runs.txt:20-22). Every ALT_K turns, a
// game whose last ALT_K turns held a run of one turn or of three or more (the winner changed twice
// in a row, or repeated twice in a row) jumps its tortoise to the present state, as long as the game
// is at most PAIR_P turns old; a tortoise that has waited ALT_L0 turns (then twice as long, and so
// on) jumps too, so the pattern jumps stop and the timed ones find a cycle of any shape. Before the
// first turn, player 1 counts as the last winner and the hands as having changed. A repeat is only
// declared on an exact state match and the first match after a jump is the cycle length, so every
// output is exact; the turns to the repeat are fewer than Brent's (about 8 % at n = 14 canonical).
constexpr u64 PAIR_P = 128;
template <int W> static inline int pair_brent(St<W> s, u64& rounds, u64& lam_out, St<W>* meet, u64 cap = 0) {
    if (s.la == 0 || s.lb == 0) { rounds = 0; return 0; }
    St<W> t = s;
    u64 R = 0, age = 0, L = ALT_L0;
    bool viol = false;
    int prevw = 1, preveq = 0;
    for (;;) {
        const int w = step<W>(s, 0);
        age++;
        if (s.la == 0 || s.lb == 0) { rounds = age; return 0; }
        if (same<W>(s, t)) { rounds = age; lam_out = age - R; if (meet) *meet = s; return 1; }
        const int eq = w == prevw;
        if (eq == preveq) viol = true;
        prevw = w;
        preveq = eq;
        if (age % ALT_K == 0) {
            if (age - R >= L) { t = s; R = age; L *= 2; }
            else if (viol && age <= PAIR_P) { t = s; R = age; }
            viol = false;
        }
        if (cap && age >= cap) { rounds = age; return 2; }
    }
}

// ---------------------------------------------------------------------------------------------
// Permutations (lexicographic, p[0] most significant), war_gpu.py:135-146 and :109-119.
static u64 fact_u64(int k) { u64 f = 1; for (int i = 2; i <= k; i++) f *= (u64)i; return f; }

static void unrank(u64 r, int m, int* p) {
    int avail[32]; int na = m;
    for (int i = 0; i < m; i++) avail[i] = i;
    for (int i = 0; i < m; i++) {
        u64 f = fact_u64(m - 1 - i);
        int d = (int)(r / f); r -= (u64)d * f;
        p[i] = avail[d];
        for (int t = d; t < na - 1; t++) avail[t] = avail[t + 1];
        na--;
    }
}

// ---------------------------------------------------------------------------------------------
// Position builders.
template <int W> static inline St<W> deal(const int* p, int n) {   // war_gpu.py:98-107
    typedef typename Lay<W>::T T;
    St<W> s; s.a = 0; s.b = 0; s.la = 0; s.lb = 0;
    for (int i = 0; i < n; i++) {
        if ((i & 1) == 0) { s.a |= ((T)p[i]) << (W * s.la); s.la++; }
        else              { s.b |= ((T)p[i]) << (W * s.lb); s.lb++; }
    }
    return s;
}

template <int W> static inline St<W> canon_pos(const int* q, int n, int k) {   // war_gpu.py:210-213
    typedef typename Lay<W>::T T;
    int m = n - 1;
    St<W> s; s.a = (T)(n - 1); s.la = 1; s.b = 0; s.lb = 0;
    for (int i = 0; i < k; i++) { s.a |= ((T)q[i]) << (W * s.la); s.la++; }
    for (int i = k; i < m; i++) { s.b |= ((T)q[i]) << (W * s.lb); s.lb++; }
    return s;
}

// ---------------------------------------------------------------------------------------------
// Distinct cycles. key() follows cpu/distinct_cycles.c:24-31 (4-bit cards, n <= 15).
static inline u64 key4(const St<4>& s) {
    u64 k = (u64)s.la << 60; int pos = 56;
    u64 v = s.a;
    for (int i = 0; i < s.la; i++) { k |= (v & 15ULL) << pos; v >>= 4; pos -= 4; }
    v = s.b;
    for (int i = 0; i < s.lb; i++) { k |= (v & 15ULL) << pos; v >>= 4; pos -= 4; }
    return k;
}
static inline St<4> swapst(const St<4>& s) { St<4> t; t.a = s.b; t.b = s.a; t.la = s.lb; t.lb = s.la; return t; }

// loop_shape (war_gpu.py:51-73): 0 = runs of exactly blk, 1 = other shape, 2 = lam > 512.
template <int W> static int loop_shape(St<W> s, u64 lam, int blk, int rule) {
    if (lam > 512) return 2;
    std::vector<int> wb((size_t)lam);
    for (u64 i = 0; i < lam; i++) wb[i] = step<W>(s, rule);
    int L = (int)lam;
    auto WB = [&](int i) { return wb[(size_t)(i % L)]; };
    int st = -1;
    for (int i = 0; i < L; i++) if (WB(i) != WB(i + L - 1)) { st = i; break; }
    if (st < 0) return 1;
    int run = 1;
    for (int j = 1; j <= L; j++) {
        if (j < L && WB(st + j) == WB(st + j - 1)) { run++; continue; }
        if (run != blk) return 1;
        run = 1;
    }
    return 0;
}

// ---------------------------------------------------------------------------------------------
struct Acc {
    u64 cyc = 0, done = 0, maxfin = 0, turns = 0, oncycle = 0, minkey = 0, nonstd = 0, unchk = 0, undet = 0;
    std::vector<u64> hist, lenhist, dhist;
    std::vector<u64> keys;   // cycle-minimum keys of on-cycle positions (for the distinct-key count)
    Acc(int hb, int lb, int db) : hist(hb > 0 ? hb : 0, 0), lenhist(lb > 0 ? lb : 0, 0), dhist(db > 0 ? db : 0, 0) {}
};

struct Cfg { int n, mode, rule, hbins, lbins, dbins, distinct, blk, det; u64 cap; };

template <int W> static void play_cycle(const St<W>& s, const Cfg& c, Acc& A, u64 lam, const St<W>& meet);

template <int W> static void play(const St<W>& s, const Cfg& c, Acc& A) {
    u64 rounds = 0, lam = 0; St<W> meet;
    if (c.det == 2 || c.det == 3) {   // alternation / pair-run checkpoints: every output, own turn count
        const int r = c.det == 2 ? alt_brent<W>(s, rounds, lam, &meet, c.cap) : pair_brent<W>(s, rounds, lam, &meet, c.cap);
        A.done++;
        A.turns += rounds;
        if (r == 2) { A.undet++; return; }
        if (r == 0) {
            if (rounds > A.maxfin) A.maxfin = rounds;
            if (c.lbins) A.lenhist[rounds < (u64)(c.lbins - 1) ? rounds : (u64)(c.lbins - 1)]++;
            return;
        }
        play_cycle<W>(s, c, A, lam, meet);
        return;
    }
    if (c.det == 1) {      // streak detector: counts, game lengths and turns; no cycle lengths
        int r = streak<W>(s, c.n, rounds, c.cap);
        A.done++;
        A.turns += rounds;
        if (r == 1) A.cyc++;
        else if (r == 2) A.undet++;
        else {
            if (rounds > A.maxfin) A.maxfin = rounds;
            if (c.lbins) A.lenhist[rounds < (u64)(c.lbins - 1) ? rounds : (u64)(c.lbins - 1)]++;
        }
        return;
    }
    const int cyc = brent<W>(s, c.rule, rounds, lam, &meet, c.cap);
    A.done++;
    A.turns += rounds;
    if (cyc == 2) { A.undet++; return; }
    if (!cyc) {
        if (rounds > A.maxfin) A.maxfin = rounds;
        if (c.lbins) A.lenhist[rounds < (u64)(c.lbins - 1) ? rounds : (u64)(c.lbins - 1)]++;
        return;
    }
    play_cycle<W>(s, c, A, lam, meet);
}

// An endless position s with cycle length lam, seen at state meet (on the cycle).
template <int W> static void play_cycle(const St<W>& s, const Cfg& c, Acc& A, u64 lam, const St<W>& meet) {
    A.cyc++;
    if (c.hbins) A.hist[lam < (u64)(c.hbins - 1) ? lam : (u64)(c.hbins - 1)]++;
    if (c.mode == ORC_CANON && (c.dbins || c.distinct)) {
        if constexpr (W == 4) {
            // war_gpu.py:224-234 (weights) and distinct_cycles.c:184-197 (minimum key)
            St<4> v = s; u64 c1n = 0, c2n = 0; int sym = 0;
            u64 K = ~0ULL;
            for (u64 t = 0; t < lam; t++) {
                if (v.la > 0 && (int)(v.a & 15ULL) == c.n - 1) { c1n++; u64 k = key4(v); if (k < K) K = k; }
                if (v.lb > 0 && (int)(v.b & 15ULL) == c.n - 1) { c2n++; u64 k = key4(swapst(v)); if (k < K) K = k; }
                if (v.a == s.b && v.b == s.a && v.la == s.lb) sym = 1;
                step<4>(v, c.rule);
            }
            if (same<4>(v, s)) {
                A.oncycle++;
                if (c.dbins) {
                    u64 d = sym ? c1n : c1n + c2n;
                    A.dhist[d < (u64)(c.dbins - 1) ? d : (u64)(c.dbins - 1)]++;
                }
                if (c.distinct) {
                    if (key4(s) == K) A.minkey++;
                    A.keys.push_back(K);
                }
            }
        }
    }
    if (c.blk) {
        int shp = loop_shape<W>(meet, lam, c.blk, c.rule);
        if (shp == 1) A.nonstd++; else if (shp == 2) A.unchk++;
    }
}

template <int W> static void run_chunk(const Cfg& c, u64 lo, u64 hi, Acc& A) {
    int n = c.n, p[32];
    if (c.mode == ORC_CANON) {
        int m = n - 1;
        u64 r = lo / (u64)m; int k = (int)(lo % (u64)m);
        unrank(r, m, p);
        for (u64 pos = lo; pos < hi; pos++) {
            play<W>(canon_pos<W>(p, n, k), c, A);
            if (++k == m) { k = 0; std::next_permutation(p, p + m); }
        }
        return;
    }
    unrank(lo, n, p);
    for (u64 r = lo; r < hi; r++) {
        bool take = true;
        if (c.mode == ORC_HALF) {   // war_gpu.py:151-154: keep deals where card n-1 goes to player 1
            int pos = 0; while (p[pos] != n - 1) pos++;
            take = (pos & 1) == 0;
        }
        if (take) play<W>(deal<W>(p, n), c, A);
        std::next_permutation(p, p + n);
    }
}

static u64 range_total(int n, int mode) {
    if (n < 1) return 0;
    if (mode == ORC_CANON) return n < 2 ? 0 : fact_u64(n - 1) * (u64)(n - 1);
    return fact_u64(n);
}

static std::string g_err;

extern "C" const char* orc_last_error(void) { return g_err.c_str(); }

extern "C" u64 orc_total(int n, int mode) { return range_total(n, mode); }

// Play every position in [lo, hi) (deal ranks, or canon positions). 0 on success.
// det: 0 = Brent (every output), 1 = the streak test (odd n, rule 0; no cycle lengths, so no
// hist / dhist / distinct / shape), 2 = Brent with alternation checkpoints (odd n, rule 0; every
// output, its own turn count), 3 = Brent with pair-run checkpoints (n = 2 mod 4, rule 0; every
// output, its own turn count).
// cap: the step cap (0 = none for Brent and the checkpoint detectors, the reference's 2^22 for the
// streak test): a game unresolved after cap turns (the streak test: with more than cap turns played
// at the top of its loop) counts as undetermined, its turns in `turns`.
extern "C" int orc_run_cap(int n, int mode, int rule, u64 lo, u64 hi, int threads,
                           int distinct, int shape_blk, int det, u64 cap, orc_res* res) {
    if (!res) { g_err = "null result"; return 1; }
    if (n < 1 || n > 20) { g_err = "n must be in [1, 20]"; return 2; }
    if (mode < 0 || mode > 2) { g_err = "bad mode"; return 3; }
    if (rule != 0 && rule != 1) { g_err = "rule must be 0 or 1"; return 4; }
    u64 total = range_total(n, mode);
    if (lo > hi || hi > total) { g_err = "range outside [0, total]"; return 5; }
    if (mode == ORC_HALF && (n & 1)) { g_err = "half mode needs even n"; return 6; }
    if (distinct && (mode != ORC_CANON || n > 15)) { g_err = "distinct needs canon mode and n <= 15"; return 7; }
    if (res->dbins && (mode != ORC_CANON || n > 16)) { g_err = "dbins needs canon mode and n <= 16"; return 7; }
    if (shape_blk && mode != ORC_CANON) { g_err = "shape needs canon mode"; return 8; }
    if (det < 0 || det > 3) {
        g_err = "det must be 0 (Brent), 1 (streak), 2 (alternation checkpoints) or 3 (pair-run checkpoints)"; return 9;
    }
    if (det == 2 && ((n % 2) == 0 || rule != 0)) {
        g_err = "alternation checkpoints need odd n and rule 0"; return 10;
    }
    if (det == 3 && ((n % 4) != 2 || rule != 0)) {
        g_err = "pair-run checkpoints need n = 2 mod 4 and rule 0"; return 10;
    }
    if (det == 1 && ((n % 2) == 0 || rule != 0 || distinct || shape_blk || (res->hist && res->hbins) ||
                     (res->dhist && res->dbins))) {
        g_err = "the streak test needs odd n, rule 0 and no cycle-length outputs"; return 10;
    }
    Cfg c{n, mode, rule, res->hist ? res->hbins : 0, res->lenhist ? res->lbins : 0,
          res->dhist ? res->dbins : 0, distinct, shape_blk, det, cap};
    if (threads > 0) omp_set_num_threads(threads);
    const u64 CH = (mode == ORC_CANON) ? (u64)std::max(1, n - 1) * 512 : 4096;
    u64 nchunks = (hi - lo + CH - 1) / CH;   // 0 for an empty range
    Acc G(c.hbins, c.lbins, c.dbins);
    #pragma omp parallel
    {
        Acc A(c.hbins, c.lbins, c.dbins);
        #pragma omp for schedule(dynamic, 1)
        for (long long ci = 0; ci < (long long)nchunks; ci++) {
            u64 a = lo + (u64)ci * CH, b = std::min(hi, a + CH);
            if (n <= 16) run_chunk<4>(c, a, b, A); else run_chunk<5>(c, a, b, A);
        }
        #pragma omp critical
        {
            G.cyc += A.cyc; G.done += A.done; G.turns += A.turns; G.oncycle += A.oncycle;
            G.minkey += A.minkey; G.nonstd += A.nonstd; G.unchk += A.unchk; G.undet += A.undet;
            G.maxfin = std::max(G.maxfin, A.maxfin);
            for (size_t i = 0; i < G.hist.size(); i++) G.hist[i] += A.hist[i];
            for (size_t i = 0; i < G.lenhist.size(); i++) G.lenhist[i] += A.lenhist[i];
            for (size_t i = 0; i < G.dhist.size(); i++) G.dhist[i] += A.dhist[i];
            G.keys.insert(G.keys.end(), A.keys.begin(), A.keys.end());
        }
    }
    std::sort(G.keys.begin(), G.keys.end());
    u64 nk = (u64)(std::unique(G.keys.begin(), G.keys.end()) - G.keys.begin());
    res->cyc = G.cyc; res->done = G.done; res->maxfin = G.maxfin; res->turns = G.turns;
    res->oncycle = G.oncycle; res->distinct_minkey = G.minkey; res->distinct_keys = nk;
    res->shape_nonstd = G.nonstd; res->shape_unchk = G.unchk; res->undet = G.undet;
    for (int i = 0; i < c.hbins; i++) res->hist[i] = G.hist[(size_t)i];
    for (int i = 0; i < c.lbins; i++) res->lenhist[i] = G.lenhist[(size_t)i];
    for (int i = 0; i < c.dbins; i++) res->dhist[i] = G.dhist[(size_t)i];
    return 0;
}

extern "C" int orc_run_det(int n, int mode, int rule, u64 lo, u64 hi, int threads,
                           int distinct, int shape_blk, int det, orc_res* res) {
    return orc_run_cap(n, mode, rule, lo, hi, threads, distinct, shape_blk, det, 0, res);
}

extern "C" int orc_run(int n, int mode, int rule, u64 lo, u64 hi, int threads,
                       int distinct, int shape_blk, orc_res* res) {
    return orc_run_det(n, mode, rule, lo, hi, threads, distinct, shape_blk, 0, res);
}

// One position given as two piles (top card first). Returns 1 if endless, 0 if it ends.
// rounds / lam as in orc_run. Used for known-answer tests on single published positions.
extern "C" int orc_play(int n, const int* p1, int l1, const int* p2, int l2, int rule,
                        u64* rounds, u64* lam) {
    u64 r = 0, L = 0;
    bool cyc;
    if (n <= 16) {
        St<4> s; s.a = 0; s.b = 0; s.la = l1; s.lb = l2;
        for (int i = 0; i < l1; i++) s.a |= (u64)p1[i] << (4 * i);
        for (int i = 0; i < l2; i++) s.b |= (u64)p2[i] << (4 * i);
        cyc = brent<4>(s, rule, r, L, nullptr);
    } else {
        St<5> s; s.a = 0; s.b = 0; s.la = l1; s.lb = l2;
        for (int i = 0; i < l1; i++) s.a |= (u128)p1[i] << (5 * i);
        for (int i = 0; i < l2; i++) s.b |= (u128)p2[i] << (5 * i);
        cyc = brent<5>(s, rule, r, L, nullptr);
    }
    if (rounds) *rounds = r;
    if (lam) *lam = cyc ? L : 0;
    return cyc ? 1 : 0;
}

// ---------------------------------------------------------------------------------------------
// One position of up to 255 cards (cpu/cycle_length.c:16-30,114-126): byte cards in two rings.
// Brent finds the cycle length lam; then one copy is advanced lam turns and both are stepped
// together until they meet: the number of turns before the cycle (mu). Card values are any
// distinct bytes. Returns 0 = the game ends (turns = its length), 1 = it cycles (turns = the
// round Brent saw the repeat), 2 = not settled within cap turns.
struct Ring {
    unsigned char c[2][256];
    int h[2], len[2];
};
static inline int ring_step(Ring& s, int rule) {
    int x = s.c[0][s.h[0]], y = s.c[1][s.h[1]];
    s.h[0] = (s.h[0] + 1) & 255; s.h[1] = (s.h[1] + 1) & 255;
    s.len[0]--; s.len[1]--;
    int w = x > y ? 0 : 1;
    int hi = x > y ? x : y, lo = x > y ? y : x;
    int first = rule ? lo : hi, second = rule ? hi : lo;
    s.c[w][(s.h[w] + s.len[w]) & 255] = (unsigned char)first; s.len[w]++;
    s.c[w][(s.h[w] + s.len[w]) & 255] = (unsigned char)second; s.len[w]++;
    return s.len[0] == 0 || s.len[1] == 0;
}
static inline bool ring_same(const Ring& x, const Ring& y) {
    for (int p = 0; p < 2; p++) {
        if (x.len[p] != y.len[p]) return false;
        for (int i = 0; i < x.len[p]; i++)
            if (x.c[p][(x.h[p] + i) & 255] != y.c[p][(y.h[p] + i) & 255]) return false;
    }
    return true;
}
extern "C" int orc_cycle_info(const int* p1, int l1, const int* p2, int l2, int rule, u64 cap,
                              u64* turns, u64* lam_out, u64* mu_out) {
    if (l1 < 1 || l2 < 1 || l1 + l2 > 255) { g_err = "piles must be non-empty, at most 255 cards"; return -1; }
    Ring s0; memset(&s0, 0, sizeof s0);
    for (int i = 0; i < l1; i++) s0.c[0][i] = (unsigned char)p1[i];
    for (int i = 0; i < l2; i++) s0.c[1][i] = (unsigned char)p2[i];
    s0.len[0] = l1; s0.len[1] = l2;
    Ring s = s0, t = s0;
    u64 power = 1, lam = 0, r = 0;
    for (;;) {
        if (ring_step(s, rule)) { *turns = r + 1; return 0; }
        r++; lam++;
        if (ring_same(s, t)) break;
        if (lam == power) { t = s; power <<= 1; lam = 0; }
        if (r >= cap) { *turns = r; return 2; }
    }
    Ring x = s0, y = s0;
    for (u64 i = 0; i < lam; i++) ring_step(y, rule);
    u64 mu = 0;
    while (!ring_same(x, y)) { ring_step(x, rule); ring_step(y, rule); mu++; }
    *turns = r; *lam_out = lam; *mu_out = mu;
    return 1;
}

// ---------------------------------------------------------------------------------------------
// Are the distinct cycles entered by ordinary deals? (cpu/reach_cycles.c:2-11,32-69,73-107)
// Each distinct cycle (up to swapping the players) is taken once, at its canonical position of
// minimum key (as in distinct_cycles.c); then the positions that lead into it are searched
// backwards for one of deal size (ceil(n/2) and floor(n/2) cards, either way round). A position
// has at most two predecessors: undo a win of player w, whose bottom two cards must be (winner,
// loser) with winner > loser (rule 0). Off the loop the predecessors form a finite tree, searched
// depth first. More than cap positions visited for one cycle: undetermined.
static inline bool undo_win(const St<4>& s, int w, St<4>& out) {
    const int L = w ? s.lb : s.la;
    if (L < 2) return false;
    const u64 word = w ? s.b : s.a;
    const u64 cw = (word >> (4 * (L - 2))) & 15ULL, cl = (word >> (4 * (L - 1))) & 15ULL;
    if (cw <= cl) return false;
    const u64 rest = word & ((1ULL << (4 * (L - 2))) - 1ULL);
    out = s;
    if (w == 0) { out.a = (rest << 4) | cw; out.la = L - 1; out.b = (s.b << 4) | cl; out.lb = s.lb + 1; }
    else        { out.b = (rest << 4) | cw; out.lb = L - 1; out.a = (s.a << 4) | cl; out.la = s.la + 1; }
    return true;
}
static inline bool deal_sized(const St<4>& s, int n) {
    const int hi = (n + 1) / 2, lo = n / 2;
    return (s.la == hi && s.lb == lo) || (s.la == lo && s.lb == hi);
}
// 1 = some deal-size position leads into the cycle through s0, 0 = none, -1 = capped. *depth: the
// most positions the search's stack held at once (no limit here, as the reference's 2^18 is never
// approached).
static int reached_by_deal(const St<4>& s0, u64 lam, u64 cap, int n, u64* depth) {
    St<4> prev = s0;                              // the loop predecessor of t
    for (u64 i = 0; i + 1 < lam; i++) step<4>(prev, 0);
    St<4> t = s0;
    std::vector<St<4>> todo;
    u64 visited = 0;
    for (u64 i = 0; i < lam; i++) {
        if (deal_sized(t, n)) return 1;
        for (int w = 0; w < 2; w++) {
            St<4> q;
            if (!undo_win(t, w, q) || same<4>(q, prev)) continue;
            todo.clear();
            todo.push_back(q);
            if (*depth < 1) *depth = 1;
            while (!todo.empty()) {
                St<4> u = todo.back(); todo.pop_back();
                if (deal_sized(u, n)) return 1;
                if (++visited > cap) return -1;
                for (int w2 = 0; w2 < 2; w2++) { St<4> r; if (undo_win(u, w2, r)) todo.push_back(r); }
                if (todo.size() > *depth) *depth = todo.size();
            }
        }
        prev = t;
        step<4>(t, 0);
    }
    return 0;
}
// Reach over canonical positions [lo, hi) (classes taken at their minimum-key position, as the
// engine's war_reach does); *maxdepth: the deepest search stack.
extern "C" int orc_reach_range(int n, u64 lo, u64 hi, int threads, u64 cap, u64* classes, u64* yes, u64* no,
                               u64* und, u64* nolen, u64* maxdepth) {
    if (n < 2 || n > 15) { g_err = "reach needs 2 <= n <= 15"; return 1; }
    if (lo > hi || hi > range_total(n, ORC_CANON)) { g_err = "range outside the canonical positions"; return 2; }
    if (threads > 0) omp_set_num_threads(threads);
    const int m = n - 1;
    u64 C = 0, Y = 0, N = 0, U = 0, D = 0;
    std::vector<u64> NL(4096, 0);
    const u64 r0 = lo / (u64)m, r1 = (hi + (u64)m - 1) / (u64)m;
    #pragma omp parallel
    {
        std::vector<u64> nl(4096, 0);
        u64 d = 0;
        #pragma omp for schedule(dynamic, 64) reduction(+ : C, Y, N, U)
        for (long long r = (long long)r0; r < (long long)r1; r++) {
            int p[16];
            unrank((u64)r, m, p);
            for (int k = 0; k < m; k++) {
                const u64 pos = (u64)r * (u64)m + (u64)k;
                if (pos < lo || pos >= hi) continue;
                St<4> s = canon_pos<4>(p, n, k);
                u64 rounds = 0, lam = 0;
                if (!brent<4>(s, 0, rounds, lam, nullptr)) continue;
                St<4> v = s;
                for (u64 t = 0; t < lam; t++) step<4>(v, 0);
                if (!same<4>(v, s)) continue;                     // not on its own cycle
                u64 K = ~0ULL;
                v = s;
                for (u64 t = 0; t < lam; t++) {
                    if (v.la > 0 && (int)(v.a & 15ULL) == n - 1) { u64 kk = key4(v); if (kk < K) K = kk; }
                    if (v.lb > 0 && (int)(v.b & 15ULL) == n - 1) { u64 kk = key4(swapst(v)); if (kk < K) K = kk; }
                    step<4>(v, 0);
                }
                if (key4(s) != K) continue;                       // not the class's representative
                C++;
                const int res = reached_by_deal(s, lam, cap, n, &d);
                if (res == 1) Y++;
                else if (res == 0) { N++; nl[lam < 4095 ? lam : 4095]++; }
                else U++;
            }
        }
        #pragma omp critical
        {
            for (int i = 0; i < 4096; i++) NL[(size_t)i] += nl[(size_t)i];
            if (d > D) D = d;
        }
    }
    *classes = C; *yes = Y; *no = N; *und = U;
    if (nolen) for (int i = 0; i < 4096; i++) nolen[i] = NL[(size_t)i];
    if (maxdepth) *maxdepth = D;
    return 0;
}

extern "C" int orc_reach(int n, int threads, u64 cap, u64* classes, u64* yes, u64* no, u64* und, u64* nolen) {
    if (n < 2 || n > 15) { g_err = "reach needs 2 <= n <= 15"; return 1; }
    return orc_reach_range(n, 0, range_total(n, ORC_CANON), threads, cap, classes, yes, no, und, nolen, nullptr);
}

// Lexicographic unrank, exposed for the permutation tests.
extern "C" void orc_unrank(u64 r, int m, int* p) { unrank(r, m, p); }

#ifdef ORC_MAIN
static void usage() {
    fprintf(stderr, "usage: war_oracle n [--mode full|half|canon] [--rule 0|1] [--lo L] [--hi H]\n"
                    "                  [--hist B] [--lbins B] [--dbins B] [--distinct] [--shape] [--threads T]\n"
                    "                  [--det brent|streak|alt|pair]\n");
}
int main(int argc, char** argv) {
    if (argc < 2) { usage(); return 2; }
    int n = atoi(argv[1]), mode = ORC_FULL, rule = 0, threads = 0, hb = 0, lb = 0, db = 0, distinct = 0, shape = 0;
    int det = 0;
    long long lo = 0, hi = -1;
    for (int i = 2; i < argc; i++) {
        std::string a = argv[i];
        auto nxt = [&]() { if (i + 1 >= argc) { usage(); exit(2); } return std::string(argv[++i]); };
        if (a == "--mode") { std::string m = nxt(); mode = m == "canon" ? ORC_CANON : m == "half" ? ORC_HALF : ORC_FULL; }
        else if (a == "--rule") rule = atoi(nxt().c_str());
        else if (a == "--lo") lo = atoll(nxt().c_str());
        else if (a == "--hi") hi = atoll(nxt().c_str());
        else if (a == "--hist") hb = atoi(nxt().c_str());
        else if (a == "--lbins") lb = atoi(nxt().c_str());
        else if (a == "--dbins") db = atoi(nxt().c_str());
        else if (a == "--distinct") distinct = 1;
        else if (a == "--shape") shape = 1;
        else if (a == "--threads") threads = atoi(nxt().c_str());
        else if (a == "--det") { std::string d = nxt(); det = d == "streak" ? 1 : d == "alt" ? 2 : d == "pair" ? 3 : 0; }
        else { usage(); return 2; }
    }
    u64 total = range_total(n, mode);
    if (hi < 0) hi = (long long)total;
    std::vector<u64> H(hb), L(lb), D(db);
    orc_res r{}; r.hist = H.data(); r.hbins = hb; r.lenhist = L.data(); r.lbins = lb; r.dhist = D.data(); r.dbins = db;
    int blk = shape ? (n & -n) : 0;
    double t0 = omp_get_wtime();
    int rc = orc_run_det(n, mode, rule, (u64)lo, (u64)hi, threads, distinct, blk, det, &r);
    double el = omp_get_wtime() - t0;
    if (rc) { fprintf(stderr, "error: %s\n", orc_last_error()); return 1; }
    printf("n=%d mode=%d rule=%d range=[%lld,%lld) done=%llu cycling=%llu longest_finite=%llu turns=%llu "
           "undetermined=%llu threads=%d time=%.3fs\n", n, mode, rule, lo, hi, r.done, r.cyc, r.maxfin, r.turns,
           r.undet, threads > 0 ? threads : omp_get_max_threads(), el);
    if (distinct || db)
        printf("   oncycle=%llu distinct_minkey=%llu distinct_keys=%llu\n", r.oncycle, r.distinct_minkey, r.distinct_keys);
    if (shape) printf("   shape blk=%d nonstd=%llu unchecked=%llu\n", blk, r.shape_nonstd, r.shape_unchk);
    if (hb) { printf("   hist:"); for (int i = 0; i < hb; i++) if (H[i]) printf(" %d:%llu", i, H[i]); printf("\n"); }
    if (lb) { printf("   lenhist:"); for (int i = 0; i < lb; i++) if (L[i]) printf(" %d:%llu", i, L[i]); printf("\n"); }
    if (db) { printf("   dhist:"); for (int i = 0; i < db; i++) if (D[i]) printf(" %d:%llu", i, D[i]); printf("\n"); }
    return 0;
}
#endif
