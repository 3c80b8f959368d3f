// Single-position tool of the War engine: cycle length and pre-period of given positions of up to
// 255 cards (the job of the reference's cpu/cycle_length.c, here for a batch of positions, one
// per lane).
//
// Piles of byte cards live in per-lane rings in LDS (hare and tortoise, 2 x 256 bytes each); the
// start position and the rest of a lane's state live in device memory between launches, so a
// launch plays at most `budget` turns per lane and the host launches again until every lane is
// settled. The repeat test compares two cheap exact functions of each pile first (the card sum
// and the sum of card x (depth + 1)), and the bytes only when they agree.
#pragma once
#include "war_device.h"

namespace war {

struct CInfoDev {
    unsigned char start[2][256];   // the position, player 1 then player 2, top card first
    u32 h[4], len[4];              // rings 0, 1: hare (players 1, 2); 2, 3: tortoise
    u64 sum[4], wsum[4];           // card sum and sum of card x (depth + 1), per ring
    unsigned char ring[4][256];    // the rings between launches
    u64 rounds, power, lam, mu, steps;
    u32 phase, status, l1, l2;     // phase: 0 Brent, 1 advance one copy lam turns, 2 meet, 3 done
};

enum { CI_BRENT = 0, CI_ADVANCE = 1, CI_MEET = 2, CI_DONE = 3 };
constexpr u32 CI_LANES = 64;      // one wave per block; LDS: 4 rings x 256 bytes per lane

struct CInfoRings {
    unsigned char *r;              // this lane's 4 x 256 bytes of LDS
    u32 h[4], len[4];
    u64 sum[4], wsum[4];

    __device__ unsigned char at(int k, u32 i) const { return r[k * 256 + ((h[k] + i) & 255u)]; }
    __device__ u32 pop(int k) {
        const u32 c = r[k * 256 + h[k]];
        h[k] = (h[k] + 1u) & 255u;
        len[k]--;
        wsum[k] -= sum[k];
        sum[k] -= c;
        return c;
    }
    __device__ void push(int k, u32 c) {
        r[k * 256 + ((h[k] + len[k]) & 255u)] = (unsigned char)c;
        len[k]++;
        sum[k] += c;
        wsum[k] += (u64)c * len[k];
    }
    // one turn of rings (k, k+1); true if the game is over
    __device__ bool turn(int k, int rule) {
        const u32 x = pop(k), y = pop(k + 1);
        const int w = x > y ? k : k + 1;
        const u32 hi = x > y ? x : y, lo = x > y ? y : x;
        push(w, rule ? lo : hi);
        push(w, rule ? hi : lo);
        return len[k] == 0 || len[k + 1] == 0;
    }
    __device__ bool same(int k, int j) const {      // position (k, k+1) equals position (j, j+1)
        for (int p = 0; p < 2; p++)
            if (len[k + p] != len[j + p] || sum[k + p] != sum[j + p] || wsum[k + p] != wsum[j + p]) return false;
        for (int p = 0; p < 2; p++)
            for (u32 i = 0; i < len[k + p]; i++)
                if (at(k + p, i) != at(j + p, i)) return false;
        return true;
    }
    __device__ void copy(int dst, int src) {         // position src -> position dst
        for (int p = 0; p < 2; p++) {
            for (u32 i = 0; i < len[src + p]; i++) r[(dst + p) * 256 + i] = at(src + p, i);
            h[dst + p] = 0; len[dst + p] = len[src + p]; sum[dst + p] = sum[src + p]; wsum[dst + p] = wsum[src + p];
        }
    }
    __device__ void load_start(int k, const CInfoDev &D) {
        for (int p = 0; p < 2; p++) {
            const u32 L = p ? D.l2 : D.l1;
            u64 s = 0, ws = 0;
            for (u32 i = 0; i < L; i++) {
                const u32 c = D.start[p][i];
                r[(k + p) * 256 + i] = (unsigned char)c;
                s += c; ws += (u64)c * (i + 1u);
            }
            h[k + p] = 0; len[k + p] = L; sum[k + p] = s; wsum[k + p] = ws;
        }
    }
};

// One launch: every lane plays its position for at most `budget` turns. Cycle-length.c's order
// (cpu/cycle_length.c:114-126): Brent until the game ends, repeats or reaches `cap` turns; then one
// copy of the start is advanced lam turns and both copies step together until they meet.
__global__ __launch_bounds__(CI_LANES) void k_cycinfo(CInfoDev *D, u32 count, u32 rule, u64 cap, u64 budget) {
    __shared__ unsigned char lds[CI_LANES * 4 * 256];
    const u32 g = blockIdx.x * CI_LANES + threadIdx.x;
    if (g >= count) return;
    CInfoDev &d = D[g];
    if (d.phase == CI_DONE) return;
    CInfoRings R;
    R.r = lds + threadIdx.x * 4 * 256;
    for (int k = 0; k < 4; k++) {
        R.h[k] = 0; R.len[k] = d.len[k]; R.sum[k] = d.sum[k]; R.wsum[k] = d.wsum[k];
        for (u32 i = 0; i < d.len[k]; i++) R.r[k * 256 + i] = d.ring[k][(d.h[k] + i) & 255u];
    }
    u64 left = budget;
    while (left && d.phase != CI_DONE) {
        if (d.phase == CI_BRENT) {
            left--;
            if (R.turn(0, (int)rule)) { d.status = 0; d.rounds++; d.phase = CI_DONE; break; }
            d.rounds++; d.lam++;
            if (R.same(0, 2)) {
                d.phase = CI_ADVANCE; d.steps = 0;
                R.load_start(0, d);
                continue;
            }
            if (d.lam == d.power) { R.copy(2, 0); d.power <<= 1; d.lam = 0; }
            if (d.rounds >= cap) { d.status = 2; d.phase = CI_DONE; }
        } else if (d.phase == CI_ADVANCE) {
            if (d.steps == d.lam) {
                R.load_start(2, d);
                d.mu = 0;
                d.phase = CI_MEET;
                continue;
            }
            left--;
            R.turn(0, (int)rule);
            d.steps++;
        } else {
            if (R.same(0, 2)) { d.status = 1; d.phase = CI_DONE; break; }
            left -= left > 1 ? 2 : 1;
            R.turn(0, (int)rule);
            R.turn(2, (int)rule);
            d.mu++;
        }
    }
    for (int k = 0; k < 4; k++) {
        d.h[k] = 0; d.len[k] = R.len[k]; d.sum[k] = R.sum[k]; d.wsum[k] = R.wsum[k];
        for (u32 i = 0; i < R.len[k]; i++) d.ring[k][i] = R.at(k, i);
    }
}

}  // namespace war
