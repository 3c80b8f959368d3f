// VALU issue-rate probe for gfx950: the measurements behind the flow kernel's roofline.
//
//   valu_probe [waves_per_cu]      (default 28, the flow kernel's default occupancy)
//
// Part 1: cycles per wave64 instruction per SIMD for the instruction forms the War turn is made
// of, each timed alone: every wave runs 8 independent chains of the form (no dependence between
// consecutive instructions), so the rate is the issue rate, not a latency. Two instructions per
// SIMD per 4 clocks (about 2 cycles each) is dual issue; one per 4 clocks is single issue.
// Part 2: cycles per wave-turn of the two 4-bit turn forms (turn_sb, top card in the low nibble;
// turn_top, top card in the high nibble), every lane playing an n = 13 game from its own deal with
// the repeat and end tests of the flow kernel, restarted from its start when it ends.
//
// Grid: CUs x waves_per_cu / 4 blocks of 256 threads; each thread writes one u64 of `out`, sized
// for exactly that grid. The loops have fixed trip counts.
#include "war_device.h"
#include <cstdio>
#include <cstdlib>

using namespace war;

#define CHK(x)                                                                                   \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) {                                                                  \
            fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);       \
            return 1;                                                                            \
        }                                                                                        \
    } while (0)

// eight copies of one instruction on operands %0..%7 (the chains), with extra operands from %8
#define X8(I) I(0) I(1) I(2) I(3) I(4) I(5) I(6) I(7)
#define C32 "+v"(c[0]), "+v"(c[1]), "+v"(c[2]), "+v"(c[3]), "+v"(c[4]), "+v"(c[5]), "+v"(c[6]), "+v"(c[7])
#define C64 "+v"(w[0]), "+v"(w[1]), "+v"(w[2]), "+v"(w[3]), "+v"(w[4]), "+v"(w[5]), "+v"(w[6]), "+v"(w[7])
#define I_OR(i) "v_or_b32 %" #i ", %" #i ", %8\n"
#define I_AND(i) "v_and_b32 %" #i ", 15, %" #i "\n"
#define I_ADD(i) "v_add_u32 %" #i ", %" #i ", %8\n"
#define I_LSHLOR(i) "v_lshl_or_b32 %" #i ", %" #i ", 4, %8\n"
#define I_CND(i) "v_cndmask_b32_e64 %" #i ", %" #i ", %8, %9\n"
#define I_SHL32(i) "v_lshlrev_b32 %" #i ", %8, %" #i "\n"
#define I_SHR64(i) "v_lshrrev_b64 %" #i ", %8, %" #i "\n"
#define I_LSHLADD(i) "v_lshl_add_u64 %" #i ", %" #i ", 4, %8\n"
// further forms (K >= 10): candidates for rebuilding the turn around dual-issuable instructions
#define I_XOR(i) "v_xor_b32 %" #i ", %" #i ", %8\n"
#define I_SUB(i) "v_sub_u32 %" #i ", %" #i ", %8\n"
#define I_MOV(i) "v_mov_b32 %" #i ", %8\n"
#define I_MAX(i) "v_max_u32 %" #i ", %" #i ", %8\n"
#define I_MIN(i) "v_min_u32 %" #i ", %" #i ", %8\n"
#define I_CNDVCC(i) "v_cndmask_b32_e32 %" #i ", %" #i ", %8, vcc\n"
#define I_SHR32(i) "v_lshrrev_b32 %" #i ", %8, %" #i "\n"
#define I_SHR16(i) "v_lshrrev_b16 %" #i ", %8, %" #i "\n"
#define I_PKSHR16(i) "v_pk_lshrrev_b16 %" #i ", %8, %" #i "\n"
#define I_PKADD16(i) "v_pk_add_u16 %" #i ", %" #i ", %8\n"
#define I_BFI(i) "v_bfi_b32 %" #i ", %8, %" #i ", %8\n"
#define I_PERM(i) "v_perm_b32 %" #i ", %" #i ", %8, %9\n"
#define I_MOV64(i) "v_mov_b64 %" #i ", %8\n"
#define I_ADDS(i) "v_add_u32 %" #i ", %9, %" #i "\n"
#define I_ANDLIT(i) "v_and_b32 %" #i ", 0xf0000000, %" #i "\n"
#define I_CMPVCC(i) "v_cmp_gt_u32 vcc, %" #i ", %8\n"
#define I_SUBCOVCC(i) "v_sub_co_u32 %" #i ", vcc, %8, %" #i "\n"
#define I_ADDC(i) "v_addc_co_u32 %" #i ", vcc, %" #i ", %8, vcc\n"
#define I_CND0(i) "v_cndmask_b32_e64 %" #i ", 0, %" #i ", %9\n"
#define I_ADD3(i) "v_add3_u32 %" #i ", %" #i ", %8, %8\n"
#define I_OR3(i) "v_or3_b32 %" #i ", %" #i ", %8, %8\n"
#define I_MULU24(i) "v_mul_u32_u24 %" #i ", %" #i ", %8\n"
#define I_SHL64(i) "v_lshlrev_b64 %" #i ", %8, %" #i "\n"
#define I_ALIGN(i) "v_alignbit_b32 %" #i ", %" #i ", %8, 4\n"
#define I_CMP64VCC(i) "v_cmp_eq_u64 vcc, %" #i ", %8\n"

constexpr int NFORM = 40;
const char *form_name[NFORM] = {"v_or_b32", "v_and_b32 (inline constant)", "v_add_u32", "v_lshl_or_b32",
                                "v_cndmask_b32_e64 (scalar mask)", "v_lshlrev_b32", "v_lshrrev_b64",
                                "v_lshl_add_u64", "v_cmp_gt_u32 (to a scalar pair)",
                                "v_sub_co_u32 (borrow to a scalar pair)",
                                "v_xor_b32", "v_sub_u32", "v_mov_b32", "v_max_u32", "v_min_u32",
                                "v_cndmask_b32_e32 (vcc)", "v_lshrrev_b32", "v_lshrrev_b16", "v_pk_lshrrev_b16",
                                "v_pk_add_u16", "v_bfi_b32", "v_perm_b32", "v_mov_b64",
                                "v_add_u32 (scalar operand)", "v_and_b32 (literal 0xf0000000)",
                                "v_cmp_gt_u32 (to vcc)", "v_sub_co_u32 (borrow to vcc)",
                                "v_addc_co_u32 (carry in and out in vcc)", "v_cndmask_b32_e64 (0, scalar mask)",
                                "v_add3_u32", "v_or3_b32", "v_mul_u32_u24", "v_lshlrev_b64", "v_alignbit_b32",
                                "v_cmp_eq_u64 (to vcc)",
                                "v_cndmask_b32_e32 after a VALU compare to vcc (1 compare + 8 selects, per select)",
                                "v_cndmask_b32_e64 after a VALU compare to a scalar pair (1 + 8, per select)",
                                "v_cndmask_b32_e64 reading vcc after a VALU compare to vcc (1 + 8, per select)",
                                "v_cndmask_b32_e32 after a VALU compare to vcc, 8 ORs between (1 + 8 + 8, per select)",
                                "v_or_b32 x 8 after a VALU compare to vcc (1 + 8, per OR; reference for the row above)"};

template <int K>
__global__ __launch_bounds__(256) void k_rate(u64 *out, int iters, u32 seed) {
    u32 c[8];
    u64 w[8];
    for (int i = 0; i < 8; i++) {
        c[i] = seed * (i + 3) + threadIdx.x;
        w[i] = (u64)c[i] * 0x9E3779B97F4A7C15ull;
    }
    const u32 x = seed ^ threadIdx.x, sh = seed & 7u;
    const u64 y = w[0] | 1ull;
    u64 m = ~0ull;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int r = 0; r < 8; r++) {
            if constexpr (K == 0) asm volatile(X8(I_OR) : C32 : "v"(x));
            if constexpr (K == 1) asm volatile(X8(I_AND) : C32);
            if constexpr (K == 2) asm volatile(X8(I_ADD) : C32 : "v"(x));
            if constexpr (K == 3) asm volatile(X8(I_LSHLOR) : C32 : "v"(x));
            if constexpr (K == 4) asm volatile(X8(I_CND) : C32 : "v"(x), "s"(m));
            if constexpr (K == 5) asm volatile(X8(I_SHL32) : C32 : "v"(sh));
            if constexpr (K == 6) asm volatile(X8(I_SHR64) : C64 : "v"(sh));
            if constexpr (K == 7) asm volatile(X8(I_LSHLADD) : C64 : "v"(y));
            if constexpr (K == 8) {
                u64 s0, s1, s2, s3;
                asm volatile("v_cmp_gt_u32 %0, %4, %8\n v_cmp_gt_u32 %1, %5, %8\n v_cmp_gt_u32 %2, %6, %8\n"
                             "v_cmp_gt_u32 %3, %7, %8\n v_cmp_gt_u32 %0, %4, %9\n v_cmp_gt_u32 %1, %5, %9\n"
                             "v_cmp_gt_u32 %2, %6, %9\n v_cmp_gt_u32 %3, %7, %9"
                             : "=s"(s0), "=s"(s1), "=s"(s2), "=s"(s3)
                             : "v"(c[0]), "v"(c[1]), "v"(c[2]), "v"(c[3]), "v"(x), "v"(sh));
                m ^= s0 ^ s1 ^ s2 ^ s3;
            }
            if constexpr (K == 10) asm volatile(X8(I_XOR) : C32 : "v"(x));
            if constexpr (K == 11) asm volatile(X8(I_SUB) : C32 : "v"(x));
            if constexpr (K == 12) asm volatile(X8(I_MOV) : C32 : "v"(x));
            if constexpr (K == 13) asm volatile(X8(I_MAX) : C32 : "v"(x));
            if constexpr (K == 14) asm volatile(X8(I_MIN) : C32 : "v"(x));
            if constexpr (K == 15) asm volatile("s_mov_b64 vcc, %9\n" X8(I_CNDVCC) : C32 : "v"(x), "s"(m) : "vcc");
            if constexpr (K == 16) asm volatile(X8(I_SHR32) : C32 : "v"(sh));
            if constexpr (K == 17) asm volatile(X8(I_SHR16) : C32 : "v"(sh));
            if constexpr (K == 18) asm volatile(X8(I_PKSHR16) : C32 : "v"(sh));
            if constexpr (K == 19) asm volatile(X8(I_PKADD16) : C32 : "v"(x));
            if constexpr (K == 20) asm volatile(X8(I_BFI) : C32 : "v"(x));
            if constexpr (K == 21) asm volatile(X8(I_PERM) : C32 : "v"(x), "s"(0x05010400u));
            if constexpr (K == 22) asm volatile(X8(I_MOV64) : C64 : "v"(y));
            if constexpr (K == 23) asm volatile(X8(I_ADDS) : C32 : "v"(x), "s"(seed));
            if constexpr (K == 24) asm volatile(X8(I_ANDLIT) : C32);
            if constexpr (K == 25) asm volatile(X8(I_CMPVCC) : C32 : "v"(x) : "vcc");
            if constexpr (K == 26) asm volatile(X8(I_SUBCOVCC) : C32 : "v"(x) : "vcc");
            if constexpr (K == 27) asm volatile("s_mov_b64 vcc, 0\n" X8(I_ADDC) : C32 : "v"(x) : "vcc");
            if constexpr (K == 28) asm volatile(X8(I_CND0) : C32 : "v"(x), "s"(m));
            if constexpr (K == 29) asm volatile(X8(I_ADD3) : C32 : "v"(x));
            if constexpr (K == 30) asm volatile(X8(I_OR3) : C32 : "v"(x));
            if constexpr (K == 31) asm volatile(X8(I_MULU24) : C32 : "v"(x));
            if constexpr (K == 32) asm volatile(X8(I_SHL64) : C64 : "v"(sh));
            if constexpr (K == 33) asm volatile(X8(I_ALIGN) : C32 : "v"(x));
            if constexpr (K == 34) asm volatile(X8(I_CMP64VCC) : C64 : "v"(y) : "vcc");
            if constexpr (K == 35)
                asm volatile("v_cmp_gt_u32 vcc, %8, %9\n s_nop 1\n" X8(I_CNDVCC) : C32 : "v"(x), "v"(c[0]) : "vcc");
            if constexpr (K == 36) {
                u64 s0;
                asm volatile("v_cmp_gt_u32 %8, %9, %10\n s_nop 1\n"
                             "v_cndmask_b32_e64 %0, %0, %9, %8\n v_cndmask_b32_e64 %1, %1, %9, %8\n"
                             "v_cndmask_b32_e64 %2, %2, %9, %8\n v_cndmask_b32_e64 %3, %3, %9, %8\n"
                             "v_cndmask_b32_e64 %4, %4, %9, %8\n v_cndmask_b32_e64 %5, %5, %9, %8\n"
                             "v_cndmask_b32_e64 %6, %6, %9, %8\n v_cndmask_b32_e64 %7, %7, %9, %8"
                             : C32, "=&s"(s0)
                             : "v"(x), "v"(c[0]));
                m ^= s0;
            }
            if constexpr (K == 37)
                asm volatile("v_cmp_gt_u32 vcc, %8, %9\n s_nop 1\n"
                             "v_cndmask_b32_e64 %0, %0, %8, vcc\n v_cndmask_b32_e64 %1, %1, %8, vcc\n"
                             "v_cndmask_b32_e64 %2, %2, %8, vcc\n v_cndmask_b32_e64 %3, %3, %8, vcc\n"
                             "v_cndmask_b32_e64 %4, %4, %8, vcc\n v_cndmask_b32_e64 %5, %5, %8, vcc\n"
                             "v_cndmask_b32_e64 %6, %6, %8, vcc\n v_cndmask_b32_e64 %7, %7, %8, vcc"
                             : C32 : "v"(x), "v"(c[0]) : "vcc");
            if constexpr (K == 38)
                asm volatile("v_cmp_gt_u32 vcc, %8, %9\n" X8(I_OR) X8(I_CNDVCC) : C32 : "v"(x), "v"(c[0]) : "vcc");
            if constexpr (K == 39)
                asm volatile("v_cmp_gt_u32 vcc, %8, %9\n s_nop 1\n" X8(I_OR) : C32 : "v"(x), "v"(c[0]) : "vcc");
            if constexpr (K == 9) {
                u64 s0;
                asm volatile("v_sub_co_u32 %0, %8, %9, %0\n v_sub_co_u32 %1, %8, %9, %1\n v_sub_co_u32 %2, %8, %9, %2\n"
                             "v_sub_co_u32 %3, %8, %9, %3\n v_sub_co_u32 %4, %8, %9, %4\n v_sub_co_u32 %5, %8, %9, %5\n"
                             "v_sub_co_u32 %6, %8, %9, %6\n v_sub_co_u32 %7, %8, %9, %7"
                             : C32, "=s"(s0)
                             : "s"(seed));
                m ^= s0;
            }
        }
    }
    u64 r = m;
    for (int i = 0; i < 8; i++) r ^= c[i] ^ w[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

// Part 2. FORM 0: turn_sb; FORM 1: turn_top. 16 turns per span as in the flow kernel; a lane whose
// game ended or repeated restarts from its start state at the span end.
template <int FORM>
__global__ __launch_bounds__(256) void k_turns(u64 *out, int spans, u32 zero, u32 n) {
    const u32 C = c_of<4>(n);
    const u64 r = ((u64)(blockIdx.x * blockDim.x + threadIdx.x) * 2654435761ull) % c_fact[n];
    St<4> s = deal<4>(unrank<4>(r, (int)n), (int)n);
    if constexpr (FORM == 1) s = flip_piles(s);
    open_sa<4>(s, zero);
    u32 sb = C - s.sa;
    const St<4> t = s;
    Zh<4> z = make_zh<4>(zero);
    Zt zt = make_zt(zero);
    u64 acc = 0, alive = ~0ull;
    for (int sp = 0; sp < spans; sp++) {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            u64 over;
            if constexpr (FORM == 0) turn_sb<4, 0>(s, sb, C, z, over);
            else turn_top<0>(s, sb, C, zt, over);
            const u64 em = over & alive;
            const u64 dm = __builtin_amdgcn_ballot_w64(s.a == t.a) & alive;
            if (em | dm) {
                acc += em ^ dm;
                alive &= ~(em | dm);
            }
        }
        if (alive != ~0ull) {
            if (!((alive >> __lane_id()) & 1ull)) {
                s = t;
                sb = C - s.sa;
            }
            alive = ~0ull;
        }
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = acc ^ s.a ^ s.b;
}

template <int K> void launch_rate(int blocks, u64 *out, int iters) {
    hipLaunchKernelGGL(k_rate<K>, dim3(blocks), dim3(256), 0, 0, out, iters, 7u);
}
template <int F> void launch_turns(int blocks, u64 *out, int spans) {
    hipLaunchKernelGGL(k_turns<F>, dim3(blocks), dim3(256), 0, 0, out, spans, 0u, 13u);
}

int main(int argc, char **argv) {
    hipDeviceProp_t p;
    CHK(hipGetDeviceProperties(&p, 0));
    const int wpc = argc > 1 ? atoi(argv[1]) : 28;
    if (wpc < 4 || wpc > 32 || wpc % 4) {
        fprintf(stderr, "waves per CU: a multiple of 4 in 4..32\n");
        return 2;
    }
    const int blocks = p.multiProcessorCount * wpc / 4;
    u64 *out;
    CHK(hipMalloc(&out, (size_t)blocks * 256 * sizeof(u64)));
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    const double waves_per_simd = (double)wpc / 4.0;
    const double clk = p.clockRate * 1e3;
    printf("%d CUs, %.0f MHz, %d waves per CU (%.0f per SIMD)\n", p.multiProcessorCount, clk / 1e6, wpc,
           waves_per_simd);
    auto time_it = [&](auto fn, int arg, float &best) -> int {
        best = 1e30f;
        for (int rep = 0; rep < 3; rep++) {     // the first run warms up and is not counted
            CHK(hipEventRecord(e0));
            fn(blocks, out, arg);
            CHK(hipGetLastError());
            CHK(hipEventRecord(e1));
            CHK(hipEventSynchronize(e1));
            float ms;
            CHK(hipEventElapsedTime(&ms, e0, e1));
            if (rep && ms < best) best = ms;
        }
        return 0;
    };
    void (*rates[NFORM])(int, u64 *, int) = {
        launch_rate<0>,  launch_rate<1>,  launch_rate<2>,  launch_rate<3>,  launch_rate<4>,  launch_rate<5>,
        launch_rate<6>,  launch_rate<7>,  launch_rate<8>,  launch_rate<9>,  launch_rate<10>, launch_rate<11>,
        launch_rate<12>, launch_rate<13>, launch_rate<14>, launch_rate<15>, launch_rate<16>, launch_rate<17>,
        launch_rate<18>, launch_rate<19>, launch_rate<20>, launch_rate<21>, launch_rate<22>, launch_rate<23>,
        launch_rate<24>, launch_rate<25>, launch_rate<26>, launch_rate<27>, launch_rate<28>, launch_rate<29>,
        launch_rate<30>, launch_rate<31>, launch_rate<32>, launch_rate<33>, launch_rate<34>, launch_rate<35>,
        launch_rate<36>, launch_rate<37>, launch_rate<38>, launch_rate<39>};
    const int iters = 10000;
    printf("\n| instruction form | cycles per wave-instruction per SIMD |\n|---|---:|\n");
    for (int k = 0; k < NFORM; k++) {
        float ms;
        if (time_it(rates[k], iters, ms)) return 1;
        const double inst = waves_per_simd * iters * 64.0;
        printf("| `%s` | %.2f |\n", form_name[k], ms * 1e-3 * clk / inst);
    }
    void (*turns[2])(int, u64 *, int) = {launch_turns<0>, launch_turns<1>};
    const char *turn_name[2] = {"turn_sb (top card in the low nibble)", "turn_top (top card in the high nibble)"};
    const int spans = 20000;
    printf("\n| turn form, n = 13 | cycles per wave-turn per SIMD |\n|---|---:|\n");
    for (int f = 0; f < 2; f++) {
        float ms;
        if (time_it(turns[f], spans, ms)) return 1;
        const double wave_turns = waves_per_simd * spans * 16.0;
        printf("| %s | %.1f |\n", turn_name[f], ms * 1e-3 * clk / wave_turns);
    }
    CHK(hipFree(out));
    return 0;
}
