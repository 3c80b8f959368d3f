// Device-side War arithmetic for gfx950: packed piles, one turn, permutations, position builders.
//
// Piles are packed W bits per card, top card in the lowest W bits (gpu/war_gpu.py:38-45,
// cpu/war.cpp:17-23):
//   W = 4: one u64 per pile, n <= 16 (a full pile of 16 cards is 64 bits);
//   W = 5: one u128 per pile, 17 <= n <= 20 (gpu/war24.py:31-47 uses 5-bit cards in two words).
// The state keeps one length (player 1's, as a bit offset); player 2's follows from n. A turn is
// branch-free: each pile ORs in a two-card value that is zero for the turn's loser.
//
// Permutations of m <= 16 cards are nibble-packed in a u64 (card i of the order in nibble i);
// for m <= 20 they are 5-bit packed in a u128. No per-lane arrays are used anywhere, so nothing
// spills to scratch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef unsigned long long u64;
typedef unsigned int u32;
typedef unsigned __int128 u128;

namespace war {

// 0! .. 20! (20! < 2^63, so every deal rank of n <= 20 fits a u64).
__constant__ u64 c_fact[21] = {
    1ull, 1ull, 2ull, 6ull, 24ull, 120ull, 720ull, 5040ull, 40320ull, 362880ull, 3628800ull,
    39916800ull, 479001600ull, 6227020800ull, 87178291200ull, 1307674368000ull,
    20922789888000ull, 355687428096000ull, 6402373705728000ull, 121645100408832000ull,
    2432902008176640000ull};

// Pile layouts, by id W:
//   4: 4-bit cards, one u64 per pile (n <= 16);
//   5: 5-bit cards, one u128 per pile (n <= 20);
//   6: n = 17 with the top card held apart (the layout of gpu/war17.py:7-16,33-72): card n-1 never
//      loses a turn, so it never changes hands and its holder never runs out. Both players' other
//      cards (0..15) are 4-bit in one u64 each, and the holder's pile is its word with card n-1
//      at depth p. `a` is the holder's word, `b` the other player's; the game is the same, with the
//      players named by who holds card n-1 instead of by deal order (the counts do not depend on
//      the names).
// T: pile word; P / PW: permutation word and its bits per card; CB: bits per card in a pile.
template <int W> struct Lay;
template <> struct Lay<4> { typedef u64 T; typedef u64 P; static constexpr int PW = 4, CB = 4; };
template <> struct Lay<5> { typedef u128 T; typedef u128 P; static constexpr int PW = 5, CB = 5; };
template <> struct Lay<6> { typedef u64 T; typedef u128 P; static constexpr int PW = 5, CB = 4; };
constexpr u32 HELD_N = 17;     // the one n of layout 6

// sa = CB * (player 1's cards - 1): the bit index of player 1's bottom card. Player 1 wins a turn:
// the two cards go to bit sa (after the pop), then sa += CB. Player 2 wins: they go to bit
// C - sa with C = CB * (n - 2), then sa -= CB. The game is over iff sa > C as an unsigned number
// (player 1 empty: sa = -CB; player 2 empty: sa = C + CB). Layout 6: "player 1" is the holder of
// card n-1 and sa counts only its word's cards (sa = 4 * cards in `a`, the holder's full pile
// minus one), so the same formulas hold; pb = 4 * the depth of card n-1 in the holder's pile.
template <int W> struct St {
    typename Lay<W>::T a, b;   // player 1, player 2
    u32 sa;
};
template <> struct St<6> {
    u64 a, b;                  // the holder's other cards, the other player's cards
    u32 sa, pb;
};
template <int W> __host__ __device__ constexpr u32 c_of(u32 n) { return (u32)Lay<W>::CB * (n - 2u); }

// x >> s and x << s that give 0 for s >= bit width (the hardware masks the shift count).
template <class T> __device__ __forceinline__ T shr_(T x, u32 s) { return s >= 8 * sizeof(T) ? (T)0 : x >> s; }
template <class T> __device__ __forceinline__ T shl_(T x, u32 s) { return s >= 8 * sizeof(T) ? (T)0 : x << s; }
template <class T> __device__ __forceinline__ T lowmask_(u32 s) { return s >= 8 * sizeof(T) ? ~(T)0 : (((T)1 << s) - (T)1); }

template <int W> __device__ __forceinline__ u32 bits1(const St<W> &s) { return s.sa + W; }   // W * cards of player 1

// Two 64-bit words for the two-card values of a turn, whose high halves stay zero for the whole game
// loop. Built from a zero the compiler cannot see (a kernel argument), they keep one register per
// high half instead of a shared zero register that would be copied before every turn.
// The same zero, OR-ed into sa's low bits when a game loop starts (open_sa), hides from the
// compiler that sa is a multiple of W; otherwise it narrows the shift-count masks below to a
// constant the hardware's own count masking does not cover, and emits an AND per shift.
template <int W> struct Zh {};
template <> struct Zh<4> { u64 h1, h2; };
template <> struct Zh<6> { u64 h1, h2; };
template <int W> __device__ __forceinline__ void open_sa(St<W> &s, u32 zero) { s.sa |= zero & 3u; }

template <int W> __device__ __forceinline__ Zh<W> make_zh(u32 zero) {
    Zh<W> z;
    if constexpr (W == 4 || W == 6) { z.h1 = (u64)zero << 32; z.h2 = z.h1; }
    (void)zero;
    return z;
}

// One turn (gpu/war_gpu.py:38-45). Rule 0: the winner puts the winning card, then the losing
// card, under its pile; rule 1 reverses the two. C = W * (n - 2). Returns true if player 1 won.
// Branch-free: each pile ORs in its own two-card value (zero for the loser) at its own offset.
template <int RULE>
__device__ __forceinline__ bool turn_held(St<6> &s, u32 C, Zh<6> &z);

template <int W, int RULE>
__device__ __forceinline__ bool turn(St<W> &s, u32 C, Zh<W> &z) {
    if constexpr (W == 6) { return turn_held<RULE>(s, C, z); } else {
    typedef typename Lay<W>::T T;
    constexpr u32 M = (1u << W) - 1u;
    const u32 c1 = (u32)s.a & M, c2 = (u32)s.b & M;
    s.a >>= W;
    s.b >>= W;
    const bool p1 = c1 > c2;
    const u32 v1 = RULE ? (c2 | (c1 << W)) : (c1 | (c2 << W));   // player 1 won: its card pair
    const u32 v2 = RULE ? (c1 | (c2 << W)) : (c2 | (c1 << W));   // player 2 won
    const u32 x1 = p1 ? v1 : 0u, x2 = p1 ? 0u : v2;
    // the masks only matter for a finished lane (whose state is no longer read); for u64 they are
    // free, the hardware shift already takes the count modulo 64
    constexpr u32 SM = 8 * sizeof(T) - 1;
    if constexpr (W == 4) {
        z.h1 = (z.h1 & 0xFFFFFFFF00000000ull) | x1;
        z.h2 = (z.h2 & 0xFFFFFFFF00000000ull) | x2;
        s.a |= z.h1 << (s.sa & SM);
        s.b |= z.h2 << ((C - s.sa) & SM);
    } else {
        s.a |= (T)x1 << (s.sa & SM);
        s.b |= (T)x2 << ((C - s.sa) & SM);
    }
    s.sa += p1 ? (u32)W : (u32)(-W);
    return p1;
    }
}

// Layout 6 (held top card). The holder plays card n-1 when it is on top (pb == 0): it wins, the
// other player's card goes under its word, and card n-1 goes under its pile (gpu/war17.py:36-41).
// Otherwise both play their words' top cards and the holder's card n-1 moves up one. The holder
// never runs out, so the game is over iff the other player is out: sa > C. Returns true if the
// holder won.
// sbo = C - s.sa, the other player's append offset.
template <int RULE>
__device__ __forceinline__ bool turn_held_at(St<6> &s, u32 sbo, Zh<6> &z) {
    const u32 ch = (u32)s.a & 15u, co = (u32)s.b & 15u;
    const bool top = s.pb == 0u;
    const bool hw = top | (ch > co);
    const u32 sh = top ? 0u : 4u;
    s.a >>= sh;
    s.b >>= 4;
    const u32 hoff = s.sa - sh;                                     // the holder's word, after the pop
    const u32 vp = RULE ? (co | (ch << 4)) : (ch | (co << 4));      // holder won with a word card
    const u32 vh = top ? co : vp;                                   // with card n-1: only the other card
    const u32 vo = RULE ? (ch | (co << 4)) : (co | (ch << 4));      // the other player won
    const u32 xh = hw ? vh : 0u, xo = hw ? 0u : vo;
    z.h1 = (z.h1 & 0xFFFFFFFF00000000ull) | xh;
    z.h2 = (z.h2 & 0xFFFFFFFF00000000ull) | xo;
    s.a |= z.h1 << (hoff & 63u);
    s.b |= z.h2 << (sbo & 63u);
    // card n-1 went under: rule 0 puts it before the captured card (depth = cards in the word
    // before the capture), rule 1 after it
    s.pb = top ? (RULE ? s.sa + 4u : s.sa) : s.pb - 4u;
    s.sa += hw ? 4u : (u32)-4;
    return hw;
}
template <int RULE>
__device__ __forceinline__ bool turn_held(St<6> &s, u32 C, Zh<6> &z) { return turn_held_at<RULE>(s, C - s.sa, z); }

template <int W> __device__ __forceinline__ bool ended(const St<W> &s, u32 C) { return s.sa > C; }

// Whether the state's player 1 is the position's player 1, for a position as dealt (before its first
// turn); detectors that name a turn's winner by player need it (the alternation detector counts the
// position's player 1 as the winner of the turn before the first). Always, except in layout 6, whose
// player 1 is the holder of card n-1: that is the position's player 1 for canonical positions (mode 2;
// player 1 leads with card n-1) and, for a deal, iff it holds the deal's (n+1)/2 cards, that is
// (n-1)/2 besides card n-1 in its word (sa = 4 * (n-1)/2; the other player's would be one fewer).
template <int W> __device__ __forceinline__ bool p1_is_dealt_p1(const St<W> &s, u32 mode) {
    if constexpr (W == 6) return mode == 2u || s.sa == 4u * ((HELD_N - 1u) >> 1);
    else { (void)s; (void)mode; return true; }
}

// d = C - x and the wave's mask of lanes where it borrows (x > C), in one instruction (bits of
// inactive lanes are 0). UC: C is wave-uniform and sits in a scalar register.
template <bool UC>
__device__ __forceinline__ void sub_borrow(u32 C, u32 x, u32 &d, u64 &borrow) {
    if constexpr (UC) asm("v_sub_co_u32 %0, %1, %2, %3" : "=v"(d), "=s"(borrow) : "s"(C), "v"(x));
    else asm("v_sub_co_u32 %0, %1, %2, %3" : "=v"(d), "=s"(borrow) : "v"(C), "v"(x));
}

// turn() for loops that carry sb = C - sa from turn to turn (player 2's append offset): the
// subtraction that makes the next turn's sb borrows exactly when sa > C, so its borrow is the end
// test and the turn needs no compare of its own for it. `over` receives the wave's lane mask of
// games that ended (bits of inactive lanes are 0). The subtraction is written as one instruction so
// that its borrow lands in a scalar register pair as the mask itself (the compiler would otherwise
// copy the borrow through a vector register). UC: C is the same in every lane of the wave (one
// request: a scalar register); without it C is taken in a vector register, as where a wave mixes
// requests of several n. The two cards go in with a 64-bit add (they land on
// zero bits, so it equals the OR), one v_lshl_add_u64 instead of two 32-bit ORs: 17 VALU per turn
// for 4-bit piles.
template <int W, int RULE, bool UC = false>
__device__ __forceinline__ bool turn_sb(St<W> &s, u32 &sb, u32 C, Zh<W> &z, u64 &over) {
    if constexpr (W == 4) {
        const u32 c1 = (u32)s.a & 15u, c2 = (u32)s.b & 15u;
        s.a >>= 4;
        s.b >>= 4;
        const bool p1 = c1 > c2;
        const u32 v1 = RULE ? (c2 | (c1 << 4)) : (c1 | (c2 << 4));
        const u32 v2 = RULE ? (c1 | (c2 << 4)) : (c2 | (c1 << 4));
        const u32 x1 = p1 ? v1 : 0u, x2 = p1 ? 0u : v2;
        z.h1 = (z.h1 & 0xFFFFFFFF00000000ull) | x1;
        z.h2 = (z.h2 & 0xFFFFFFFF00000000ull) | x2;
        s.a += z.h1 << (s.sa & 63u);
        s.b += z.h2 << (sb & 63u);
        s.sa += p1 ? 4u : (u32)-4;
        sub_borrow<UC>(C, s.sa, sb, over);
        return p1;
    } else if constexpr (W == 6) {
        const bool hw = turn_held_at<RULE>(s, sb, z);
        sub_borrow<UC>(C, s.sa, sb, over);
        return hw;
    } else {
        const bool p1 = turn<W, RULE>(s, C, z);
        over = __builtin_amdgcn_ballot_w64(ended<W>(s, C));
        sb = C - s.sa;
        return p1;
    }
}

// The same turn for 4-bit piles kept top card first from the high end (the "top" form: a pile of L
// cards fills bits 64-4L..63, its top card in bits 60..63; flip_piles converts). The pop is then a
// shift left by 4, which v_lshl_add_u64 does together with the add of the appended pair: the pair
// is built in bits 24..31 of the high word (first card above) and shifted right by sa (player 1)
// or sb (player 2) to sit just below the remaining cards, so both piles take one 64-bit shift and
// one v_lshl_add_u64 each, against two 64-bit shifts and an add in turn_sb. The offsets, the
// lengths and the end test are turn_sb's (sa = 4 * (player 1's cards - 1)). On gfx950 the turn
// issues in about 70 cycles per wave against about 80 for turn_sb (measured, 7 waves per SIMD).
// Zt: the two 64-bit words the pairs are built in; their low halves stay an opaque zero. UC: C is
// wave-uniform (a batch of one request), else it is taken per lane.
struct Zt { u64 h1, h2; };
__device__ __forceinline__ Zt make_zt(u32 zero) {
    // two vector-register zeros the compiler can neither rebuild from a scalar one inside the loop
    // nor merge into one register (each word keeps its own low half)
    u32 z1, z2;
    asm volatile("v_mov_b32 %0, %1" : "=v"(z1) : "s"(zero));
    asm volatile("v_mov_b32 %0, %1" : "=v"(z2) : "s"(zero));
    Zt z; z.h1 = z1; z.h2 = z2; return z;
}
template <int RULE, bool UC = true, bool SELASM = false>
__device__ __forceinline__ u64 turn_top(St<4> &s, u32 &sb, u32 C, Zt &z, u64 &over) {
    const u32 ah = (u32)(s.a >> 32) & 0xF0000000u, bh = (u32)(s.b >> 32) & 0xF0000000u;   // top cards << 28
    u64 pm;                    // the wave's mask of lanes whose player 1 won the turn
    u32 x1, x2, d;
    if constexpr (SELASM) {
        // The compare and the three selects as written: compiled from C, the first select can come
        // out as the VOP2 form reading vcc right after the compare wrote it, which gfx950 issues
        // several times slower than the VOP3 form (measured). Two wait states between the compare
        // and the first read of its mask.
        const u32 wa = ah | (bh >> 4), wb = bh | (ah >> 4);
        asm("v_cmp_gt_u32_e64 %0, %4, %5\n\t"
            "s_nop 1\n\t"
            "v_cndmask_b32_e64 %1, 0, %6, %0\n\t"
            "v_cndmask_b32_e64 %2, %7, 0, %0\n\t"
            "v_cndmask_b32_e64 %3, -4, 4, %0"
            : "=&s"(pm), "=&v"(x1), "=&v"(x2), "=&v"(d)
            : "v"(ah), "v"(bh), "v"(RULE ? wb : wa), "v"(RULE ? wa : wb));
    } else {
        const bool p1 = ah > bh;
        const u32 wa = ah | (bh >> 4), wb = bh | (ah >> 4);     // player 1's card above / player 2's above
        x1 = p1 ? (RULE ? wb : wa) : 0u;
        x2 = p1 ? 0u : (RULE ? wa : wb);
        d = p1 ? 4u : (u32)-4;
        pm = __builtin_amdgcn_ballot_w64(p1);
    }
    z.h1 = (z.h1 & 0xFFFFFFFFull) | ((u64)x1 << 32);
    z.h2 = (z.h2 & 0xFFFFFFFFull) | ((u64)x2 << 32);
    s.a = (s.a << 4) + (z.h1 >> (s.sa & 63u));
    s.b = (s.b << 4) + (z.h2 >> (sb & 63u));
    s.sa += d;
    sub_borrow<UC>(C, s.sa, sb, over);
    return pm;
}

// The repeat test compares (w1, w2, l1): card 0 and an empty slot look alike (gpu/war_gpu.py:91).
// Layout 6 adds the depth of card n-1.
template <int W> __device__ __forceinline__ bool rest_same(const St<W> &x, const St<W> &y) {
    if constexpr (W == 6) return (x.b == y.b) & (x.sa == y.sa) & (x.pb == y.pb);
    else return (x.b == y.b) & (x.sa == y.sa);
}
template <int W> __device__ __forceinline__ bool same(const St<W> &x, const St<W> &y) {
    return (x.a == y.a) & rest_same<W>(x, y);
}

template <int W> __device__ __forceinline__ u32 card(typename Lay<W>::T q, int i) {
    return (u32)(q >> (W * i)) & ((1u << W) - 1u);
}

// Lexicographic unrank (gpu/war_gpu.py:135-146): digit i = r / (m-1-i)! picks the digit-th lowest
// unused card. The unused cards are kept as an ascending packed list; picking one removes its
// field. The first digits of a large rank need a 64-bit division; once r < 2^32 and the factorial
// fits 32 bits the division is 32-bit.
template <int W>
__device__ __forceinline__ typename Lay<W>::T unrank(u64 r, int m) {
    typedef typename Lay<W>::T T;
    T avail = 0;
    for (int i = 0; i < m; i++) avail |= (T)i << (W * i);
    T q = 0;
    for (int i = 0; i < m; i++) {
        const u64 f = c_fact[m - 1 - i];
        u32 d;
        if ((r >> 32) == 0 && (f >> 32) == 0) d = (u32)r / (u32)f;
        else d = (u32)(r / f);
        r -= (u64)d * f;
        const u32 sd = W * d;
        const u32 c = (u32)(avail >> sd) & ((1u << W) - 1u);
        avail = (avail & lowmask_<T>(sd)) | shl_<T>(shr_<T>(avail, sd + W), sd);
        q |= (T)c << (W * i);
    }
    return q;
}

// Reverse the order of the W-bit fields of a whole word.
__device__ __forceinline__ u64 rev_fields4(u64 x) {
    x = __builtin_bswap64(x);
    return ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
}

// 4-bit piles between the usual form (top card in the low nibble) and turn_top's (top card in the
// high nibble): reversing the 16 fields maps one to the other, both ways.
__device__ __forceinline__ St<4> flip_piles(St<4> s) {
    s.a = rev_fields4(s.a);
    s.b = rev_fields4(s.b);
    return s;
}

// Next lexicographic permutation of the m packed cards (gpu/war_gpu.py:109-119). The caller never
// asks for the successor of the last permutation.
template <int W>
__device__ __forceinline__ typename Lay<W>::T next_perm(typename Lay<W>::T q, int m) {
    typedef typename Lay<W>::T T;
    constexpr u32 M = (1u << W) - 1u;
    int i = m - 2;
    u32 right = card<W>(q, m - 1);
    u32 ai = card<W>(q, i);
    while (i > 0 && ai >= right) { right = ai; i--; ai = card<W>(q, i); }
    // here a[i] < a[i+1] (a[0] >= a[1] only for the last permutation, never passed in)
    int j = m - 1;
    u32 aj = card<W>(q, j);
    while (aj <= ai) { j--; aj = card<W>(q, j); }
    // swap a[i], a[j]
    q &= ~(((T)M << (W * i)) | ((T)M << (W * j)));
    q |= ((T)aj << (W * i)) | ((T)ai << (W * j));
    // reverse a[i+1 .. m-1]
    const int L = m - 1 - i;
    if (L > 1) {
        T rs = 0;
        if constexpr (W == 4) {
            rs = rev_fields4(q) >> (4 * (16 - m));      // a[m-1] now in field 0
        } else {
            for (int t = 0; t < L; t++) rs |= (T)card<W>(q, m - 1 - t) << (W * t);
        }
        rs &= lowmask_<T>(W * L);
        q = (q & lowmask_<T>(W * (i + 1))) | (rs << (W * (i + 1)));
    }
    return q;
}

// Deal (gpu/war_gpu.py:98-107): deck positions 0,2,4,.. to player 1, 1,3,5,.. to player 2,
// first dealt on top.
__device__ __forceinline__ u64 even_fields4(u64 x) {
    x &= 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
    x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
    x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
    return x;
}

// Layout 6 from a 5-bit order of the 17 cards: each player's cards in deal order, card n-1 left out
// and remembered as (holder, depth) (gpu/war17.py:50-58).
__device__ __forceinline__ St<6> deal_held(u128 q) {
    u64 w0 = 0, w1 = 0;
    u32 l0 = 0, l1 = 0, holder = 0, p = 0;
#pragma unroll
    for (int i = 0; i < (int)HELD_N; i++) {
        const u32 c = (u32)(q >> (5 * i)) & 31u;
        const bool top = c == HELD_N - 1u;
        if ((i & 1) == 0) {
            holder = top ? 0u : holder; p = top ? l0 : p;
            w0 |= top ? 0ull : (u64)c << (4u * l0); l0 += top ? 0u : 1u;
        } else {
            holder = top ? 1u : holder; p = top ? l1 : p;
            w1 |= top ? 0ull : (u64)c << (4u * l1); l1 += top ? 0u : 1u;
        }
    }
    St<6> s;
    s.a = holder ? w1 : w0;
    s.b = holder ? w0 : w1;
    s.sa = 4u * (holder ? l1 : l0);
    s.pb = 4u * p;
    return s;
}

// 5-bit packed cards 0..15 (16 fields) to 4-bit packed.
__device__ __forceinline__ u64 fields5to4(u128 q) {
    u64 r = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) r |= (u64)((u32)(q >> (5 * i)) & 15u) << (4 * i);
    return r;
}

template <int W>
__device__ __forceinline__ St<W> deal(typename Lay<W>::P q, int n) {
    typedef typename Lay<W>::T T;
    St<W> s;
    if constexpr (W == 6) {
        (void)n;
        return deal_held(q);
    } else if constexpr (W == 4) {
        s.a = even_fields4(q);
        s.b = even_fields4(q >> 4);
    } else {
        s.a = 0; s.b = 0;
        for (int i = 0; i < n; i++) {
            const T c = (T)card<W>(q, i);
            if ((i & 1) == 0) s.a |= c << (W * (i >> 1));
            else              s.b |= c << (W * (i >> 1));
        }
    }
    s.sa = W * (u32)((n + 1) >> 1) - W;
    return s;
}

// Canonical position (gpu/war_gpu.py:210-213): player 1 = [n-1, q0..q(k-1)], player 2 = [qk..q(n-2)].
// Layout 6: player 1 holds card n-1 on top (depth 0), its word is q0..q(k-1).
template <int W>
__device__ __forceinline__ St<W> canon(typename Lay<W>::P q, int n, int k) {
    typedef typename Lay<W>::T T;
    St<W> s;
    if constexpr (W == 6) {
        (void)n;
        const u64 q4 = fields5to4(q);
        const u32 sk = 4u * (u32)k;                 // k <= 15
        s.a = q4 & lowmask_<u64>(sk);
        s.b = q4 >> sk;
        s.sa = sk;
        s.pb = 0;
    } else {
        const u32 sk = W * (u32)k;
        s.a = (T)(u32)(n - 1) | ((q & lowmask_<T>(sk)) << W);
        s.b = shr_<T>(q, sk);
        s.sa = W * (u32)k;
    }
    return s;
}

// The canonical position of canon<4> directly in turn_top's form (top card in the high nibble),
// from the order reversed (qr = rev_fields4(q): q0 in the top nibble): player 1 = [n-1, q0..q(k-1)]
// from the top, player 2 = [qk..q(n-2)]. k <= n - 2 <= 14, so no shift reaches 64.
__device__ __forceinline__ St<4> canon_top(u64 qr, int n, int k) {
    St<4> s;
    const u32 sk = 4u * (u32)k;
    s.a = ((qr >> 4) | ((u64)(u32)(n - 1) << 60)) & (~0ull << (60u - sk));
    s.b = qr << sk;
    s.sa = sk;
    return s;
}

// The same position after its first turn, which player 1's top card n - 1 always wins (rule 0: it goes
// under player 1's pile first, then player 2's q(k)): player 1 = [q0..q(k-1), n-1, qk], player 2 =
// [q(k+1)..q(n-2)]. k <= n - 3 (split n - 2 ends on that turn and is never dealt), so every shift
// is below 64; the bits below each pile's bottom card are zero, as turn_top needs.
__device__ __forceinline__ St<4> canon_top1(u64 qr, int n, int k) {
    St<4> s;
    const u32 sk = 4u * (u32)k;
    s.a = (qr & ~(~0ull >> sk)) | ((u64)(u32)(n - 1) << (60u - sk)) | ((qr >> 4) & (0xFull << (56u - sk)));
    s.b = qr << (sk + 4u);
    s.sa = sk + 4u;
    return s;
}

// The 24 deals whose orders share the first n-4 cards (a block of 24 consecutive ranks that starts
// at a multiple of 24) differ in the last four deck positions only: Heap's algorithm visits them
// with one swap of two of those positions per step. Step t (0..22) swaps positions i, j of the
// four: (0,1) at even t, (t/6, 3) at t = 5 mod 6, (0,2) else. Deck position p sits in player 1's
// pile (p even) or player 2's (p odd) at depth p/2, which in turn_top's form is bits 28-4(p/2) of
// the pile's high word (n <= 16: depth <= 7). n >= 5; t and n are wave-uniform.
__device__ __forceinline__ void heap_swap_top(St<4> &s, u32 t, u32 n) {
    const u32 i = (t & 1u) == 0u ? 0u : (t % 6u == 5u ? t / 6u : 0u);
    const u32 j = (t & 1u) == 0u ? 1u : (t % 6u == 5u ? 3u : 2u);
    const u32 pi = n - 4u + i, pj = n - 4u + j;
    const u32 si = 28u - 4u * (pi >> 1), sj = 28u - 4u * (pj >> 1);
    u32 ah = (u32)(s.a >> 32), bh = (u32)(s.b >> 32);
    if (((pi ^ pj) & 1u) == 0u) {              // both in one pile
        u32 w = (pi & 1u) ? bh : ah;
        const u32 d = ((w >> si) ^ (w >> sj)) & 15u;
        w ^= (d << si) | (d << sj);
        if (pi & 1u) bh = w; else ah = w;
    } else {                                   // one in each
        const u32 s1 = (pi & 1u) ? sj : si, s2 = (pi & 1u) ? si : sj;
        const u32 d = ((ah >> s1) ^ (bh >> s2)) & 15u;
        ah ^= d << s1;
        bh ^= d << s2;
    }
    s.a = (s.a & 0xFFFFFFFFull) | ((u64)ah << 32);
    s.b = (s.b & 0xFFFFFFFFull) | ((u64)bh << 32);
}

// The first order of the next block of 24 from the first order of this one (its last four cards
// ascending): the block's last order has them descending, and its successor is next_perm's.
__device__ __forceinline__ u64 next_block24(u64 q, int n) {
    const u32 sh = 4u * (u32)(n - 4);
    const u64 tail = (rev_fields4(q) >> (4u * (16u - (u32)n))) & 0xFFFFull;   // cards n-1..n-4
    return next_perm<4>((q & ((1ull << sh) - 1ull)) | (tail << sh), n);
}

// --half filter (gpu/war_gpu.py:151-154): true if card n-1 sits at an even deck position.
template <int W>
__device__ __forceinline__ bool top_card_to_p1(typename Lay<W>::T q, int n) {
    int pos = 0;
    for (int i = 0; i < n; i++) if (card<W>(q, i) == (u32)(n - 1)) pos = i;
    return (pos & 1) == 0;
}

// Key of a 4-bit state for distinct cycles (cpu/distinct_cycles.c:24-31): player 1's length in
// bits 60..63, then player 1's cards and player 2's cards, top first, from bit 56 downwards.
// bl = 4 * (player 1's cards).
__device__ __forceinline__ u64 key4(u64 a, u64 b, u32 bl) {
    const u64 c = a | shl_<u64>(b, bl);           // n <= 15 cards, top of player 1 in nibble 0
    return ((u64)(bl >> 2) << 60) | (rev_fields4(c) >> 4);
}

}  // namespace war
