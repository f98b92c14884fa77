// Device 1-D transform kernels for the intra candidate chain (gfx950, integer VALU).
//
// Forward: the encoder-defined Q16 butterfly DCT-4/8/16/32 and the Q12 matrix ADST-4/8/16
// of the maroontree encoder (reference src/dct.rs:60-633, :304, :445, :3178).  Inverse: the
// AV1-normative dav1d-exact DCT-4..64 / ADST-4/8/16 / identity kernels (reference
// src/idct.rs:1432-2675, :2837).  Every value here is held in registers of one lane; the
// 2-D drivers in mtgpu.hip stage lines through LDS.  All arithmetic wraps like a Rust
// release build (the library is compiled with -fwrapv).
#pragma once
#include <stdint.h>

#define MT_DEV __device__ __forceinline__

namespace mt {

// ------------------------------------------------------------------ Q16 forward helpers
MT_DEV int32_t mulq16(int32_t d, int32_t c) {
    return (int32_t)(((int64_t)d * (int64_t)c + 32768) >> 16);
}

// Butterfly constants of the forward DCT: round(65536 / (2 cos((2k+1) pi / 2N))) and
// round(65536 sqrt 2) (reference src/dct.rs:48-54, :231-238, :589, pinned by the closed
// form at src/dct.rs:3818-3857).
constexpr int32_t kWC4[2] = {35468, 85627};
constexpr int32_t kWC8[4] = {33410, 39410, 58981, 167963};
constexpr int32_t kWC16[8] = {32927, 34242, 37155, 42390, 51653, 69513, 112882, 334309};
constexpr int32_t kWC32[16] = {32808, 33127, 33780, 34802, 36248, 38203, 40796, 44224,
                               48794, 55008, 63738, 76640, 97266, 134859, 223321, 667812};
constexpr int32_t kSqrt2Q16 = 92682;

MT_DEV int32_t fmla_sqrt2(int32_t a, int32_t b) { return mulq16(a, kSqrt2Q16) + b; }

// Forward DCT of length N in place (AAN-style split: even half by recursion, odd half
// scaled by the WC row then recursed, then the running-sum recombination).
template <int N> struct FDct;

template <> struct FDct<4> {
    static MT_DEV void run(int32_t* b) {
        const int32_t s0 = b[0] + b[3], s1 = b[1] + b[2];
        const int32_t d0 = mulq16(b[0] - b[3], kWC4[0]);
        const int32_t d1 = mulq16(b[1] - b[2], kWC4[1]);
        const int32_t odd_hi = d0 - d1;
        b[0] = s0 + s1;
        b[2] = s0 - s1;
        b[1] = fmla_sqrt2(d0 + d1, odd_hi);
        b[3] = odd_hi;
    }
};

template <int N> MT_DEV const int32_t* wc_row();
template <> MT_DEV const int32_t* wc_row<8>() { return kWC8; }
template <> MT_DEV const int32_t* wc_row<16>() { return kWC16; }
template <> MT_DEV const int32_t* wc_row<32>() { return kWC32; }

template <int N> struct FDct {
    static MT_DEV void run(int32_t* b) {
        constexpr int H = N / 2;
        int32_t e[H], o[H];
#pragma unroll
        for (int i = 0; i < H; i++) {
            e[i] = b[i] + b[N - 1 - i];
            o[i] = b[i] - b[N - 1 - i];
        }
        FDct<H>::run(e);
        const int32_t* wc = wc_row<N>();
#pragma unroll
        for (int i = 0; i < H; i++) o[i] = mulq16(o[i], wc[i]);
        FDct<H>::run(o);
        // o[0] = o[0]*sqrt2 + o[1]; o[i] += o[i+1] for 1 <= i < H-1 (ascending: each sum
        // reads the not-yet-updated successor).
        o[0] = fmla_sqrt2(o[0], o[1]);
#pragma unroll
        for (int i = 1; i < H - 1; i++) o[i] += o[i + 1];
#pragma unroll
        for (int i = 0; i < H; i++) {
            b[2 * i] = e[i];
            b[2 * i + 1] = o[i];
        }
    }
};

// Forward ADST of length N (Q12 matrix in constant memory, i64 accumulation, round >> 12).
template <int N> MT_DEV void fadst(const int32_t* __restrict__ m, int32_t* b) {
    int32_t out[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        int64_t acc = 0;
#pragma unroll
        for (int j = 0; j < N; j++) acc += (int64_t)m[i * N + j] * (int64_t)b[j];
        out[i] = (int32_t)((acc + 2048) >> 12);
    }
#pragma unroll
    for (int i = 0; i < N; i++) b[i] = out[i];
}

// ------------------------------------------------------------------ inverse (dav1d exact)
struct Clip {
    int32_t lo, hi;
    MT_DEV int32_t operator()(int32_t v) const { return min(max(v, lo), hi); }  // lo <= hi
};

// Product of the inverse transforms.  Every operand there is a clipped value, an input
// bounded by the coefficient limit, or a sum of two such (all below 2^21 in magnitude), so
// the 24 x 24-bit multiply (full rate) gives the same low 32 bits as the 32-bit one.
MT_DEV int32_t m24(int32_t a, int32_t b) { return __mul24(a, b); }

// Rounded products in the normative forms: r12 = (x + 2048) >> 12, r11 = (x + 1024) >> 11,
// and the 181/256 rotation.
MT_DEV int32_t r12(int32_t x) { return (x + 2048) >> 12; }
MT_DEV int32_t r11(int32_t x) { return (x + 1024) >> 11; }
MT_DEV int32_t r181(int32_t x) { return (m24(x, 181) + 128) >> 8; }

// Inverse DCT-4 over the strided values v[0], v[S], v[2S], v[3S] of a register array.
template <int S> MT_DEV void idct4(int32_t* v, Clip c) {
    const int32_t i0 = v[0], i1 = v[S], i2 = v[2 * S], i3 = v[3 * S];
    const int32_t t0 = r181(i0 + i2);
    const int32_t t1 = r181(i0 - i2);
    const int32_t t2 = r12(m24(i1, 1567) - m24(i3, (3784 - 4096))) - i3;
    const int32_t t3 = r12(m24(i1, (3784 - 4096)) + m24(i3, 1567)) + i1;
    v[0] = c(t0 + t3);
    v[S] = c(t1 + t2);
    v[2 * S] = c(t1 - t2);
    v[3 * S] = c(t0 - t3);
}

template <int S> MT_DEV void idct8(int32_t* v, Clip c) {
    idct4<2 * S>(v, c);
    const int32_t i1 = v[S], i3 = v[3 * S], i5 = v[5 * S], i7 = v[7 * S];
    const int32_t t4a = r12(m24(i1, 799) - m24(i7, (4017 - 4096))) - i7;
    int32_t t5a = r11(m24(i5, 1703) - m24(i3, 1138));
    int32_t t6a = r11(m24(i5, 1138) + m24(i3, 1703));
    const int32_t t7a = r12(m24(i1, (4017 - 4096)) + m24(i7, 799)) + i1;
    const int32_t t4 = c(t4a + t5a);
    t5a = c(t4a - t5a);
    const int32_t t7 = c(t7a + t6a);
    t6a = c(t7a - t6a);
    const int32_t t5 = r181(t6a - t5a);
    const int32_t t6 = r181(t6a + t5a);
    const int32_t e0 = v[0], e1 = v[2 * S], e2 = v[4 * S], e3 = v[6 * S];
    v[0] = c(e0 + t7);
    v[S] = c(e1 + t6);
    v[2 * S] = c(e2 + t5);
    v[3 * S] = c(e3 + t4);
    v[4 * S] = c(e3 - t4);
    v[5 * S] = c(e2 - t5);
    v[6 * S] = c(e1 - t6);
    v[7 * S] = c(e0 - t7);
}

template <int S> MT_DEV void idct16(int32_t* v, Clip c) {
    idct8<2 * S>(v, c);
    const int32_t i1 = v[S], i3 = v[3 * S], i5 = v[5 * S], i7 = v[7 * S];
    const int32_t i9 = v[9 * S], i11 = v[11 * S], i13 = v[13 * S], i15 = v[15 * S];
    int32_t t8a = r12(m24(i1, 401) - m24(i15, (4076 - 4096))) - i15;
    int32_t t9a = r11(m24(i9, 1583) - m24(i7, 1299));
    int32_t t10a = r12(m24(i5, 1931) - m24(i11, (3612 - 4096))) - i11;
    int32_t t11a = r12(m24(i13, (3920 - 4096)) - m24(i3, 1189)) + i13;
    int32_t t12a = r12(m24(i13, 1189) + m24(i3, (3920 - 4096))) + i3;
    int32_t t13a = r12(m24(i5, (3612 - 4096)) + m24(i11, 1931)) + i5;
    int32_t t14a = r11(m24(i9, 1299) + m24(i7, 1583));
    int32_t t15a = r12(m24(i1, (4076 - 4096)) + m24(i15, 401)) + i1;
    int32_t t8 = c(t8a + t9a), t9 = c(t8a - t9a), t10 = c(t11a - t10a), t11 = c(t11a + t10a);
    int32_t t12 = c(t12a + t13a), t13 = c(t12a - t13a), t14 = c(t15a - t14a), t15 = c(t15a + t14a);
    t9a = r12(m24(t14, 1567) - m24(t9, (3784 - 4096))) - t9;
    t14a = r12(m24(t14, (3784 - 4096)) + m24(t9, 1567)) + t14;
    t10a = r12(-(m24(t13, (3784 - 4096)) + m24(t10, 1567))) - t13;
    t13a = r12(m24(t13, 1567) - m24(t10, (3784 - 4096))) - t10;
    t8a = c(t8 + t11);
    t9 = c(t9a + t10a);
    t10 = c(t9a - t10a);
    t11a = c(t8 - t11);
    t12a = c(t15 - t12);
    t13 = c(t14a - t13a);
    t14 = c(t14a + t13a);
    t15a = c(t15 + t12);
    t10a = r181(t13 - t10);
    t13a = r181(t13 + t10);
    t11 = r181(t12a - t11a);
    t12 = r181(t12a + t11a);
    const int32_t o[8] = {t15a, t14, t13a, t12, t11, t10a, t9, t8a};
    int32_t e[8];
#pragma unroll
    for (int i = 0; i < 8; i++) e[i] = v[2 * i * S];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        v[i * S] = c(e[i] + o[i]);
        v[(15 - i) * S] = c(e[i] - o[i]);
    }
}

template <int S> MT_DEV void idct32(int32_t* v, Clip c) {
    idct16<2 * S>(v, c);
    int32_t in[32];
#pragma unroll
    for (int i = 1; i < 32; i += 2) in[i] = v[i * S];
    int32_t t16a = r12(m24(in[1], 201) - m24(in[31], (4091 - 4096))) - in[31];
    int32_t t17a = r12(m24(in[17], (3035 - 4096)) - m24(in[15], 2751)) + in[17];
    int32_t t18a = r12(m24(in[9], 1751) - m24(in[23], (3703 - 4096))) - in[23];
    int32_t t19a = r12(m24(in[25], (3857 - 4096)) - m24(in[7], 1380)) + in[25];
    int32_t t20a = r12(m24(in[5], 995) - m24(in[27], (3973 - 4096))) - in[27];
    int32_t t21a = r12(m24(in[21], (3513 - 4096)) - m24(in[11], 2106)) + in[21];
    int32_t t22a = r11(m24(in[13], 1220) - m24(in[19], 1645));
    int32_t t23a = r12(m24(in[29], (4052 - 4096)) - m24(in[3], 601)) + in[29];
    int32_t t24a = r12(m24(in[29], 601) + m24(in[3], (4052 - 4096))) + in[3];
    int32_t t25a = r11(m24(in[13], 1645) + m24(in[19], 1220));
    int32_t t26a = r12(m24(in[21], 2106) + m24(in[11], (3513 - 4096))) + in[11];
    int32_t t27a = r12(m24(in[5], (3973 - 4096)) + m24(in[27], 995)) + in[5];
    int32_t t28a = r12(m24(in[25], 1380) + m24(in[7], (3857 - 4096))) + in[7];
    int32_t t29a = r12(m24(in[9], (3703 - 4096)) + m24(in[23], 1751)) + in[9];
    int32_t t30a = r12(m24(in[17], 2751) + m24(in[15], (3035 - 4096))) + in[15];
    int32_t t31a = r12(m24(in[1], (4091 - 4096)) + m24(in[31], 201)) + in[1];

    int32_t t16 = c(t16a + t17a), t17 = c(t16a - t17a), t18 = c(t19a - t18a), t19 = c(t19a + t18a);
    int32_t t20 = c(t20a + t21a), t21 = c(t20a - t21a), t22 = c(t23a - t22a), t23 = c(t23a + t22a);
    int32_t t24 = c(t24a + t25a), t25 = c(t24a - t25a), t26 = c(t27a - t26a), t27 = c(t27a + t26a);
    int32_t t28 = c(t28a + t29a), t29 = c(t28a - t29a), t30 = c(t31a - t30a), t31 = c(t31a + t30a);

    t17a = r12(m24(t30, 799) - m24(t17, (4017 - 4096))) - t17;
    t30a = r12(m24(t30, (4017 - 4096)) + m24(t17, 799)) + t30;
    t18a = r12(-(m24(t29, (4017 - 4096)) + m24(t18, 799))) - t29;
    t29a = r12(m24(t29, 799) - m24(t18, (4017 - 4096))) - t18;
    t21a = r11(m24(t26, 1703) - m24(t21, 1138));
    t26a = r11(m24(t26, 1138) + m24(t21, 1703));
    t22a = r11(-(m24(t25, 1138) + m24(t22, 1703)));
    t25a = r11(m24(t25, 1703) - m24(t22, 1138));

    t16a = c(t16 + t19); t17 = c(t17a + t18a); t18 = c(t17a - t18a); t19a = c(t16 - t19);
    t20a = c(t23 - t20); t21 = c(t22a - t21a); t22 = c(t22a + t21a); t23a = c(t23 + t20);
    t24a = c(t24 + t27); t25 = c(t25a + t26a); t26 = c(t25a - t26a); t27a = c(t24 - t27);
    t28a = c(t31 - t28); t29 = c(t30a - t29a); t30 = c(t30a + t29a); t31a = c(t31 + t28);

    t18a = r12(m24(t29, 1567) - m24(t18, (3784 - 4096))) - t18;
    t29a = r12(m24(t29, (3784 - 4096)) + m24(t18, 1567)) + t29;
    t19 = r12(m24(t28a, 1567) - m24(t19a, (3784 - 4096))) - t19a;
    t28 = r12(m24(t28a, (3784 - 4096)) + m24(t19a, 1567)) + t28a;
    t20 = r12(-(m24(t27a, (3784 - 4096)) + m24(t20a, 1567))) - t27a;
    t27 = r12(m24(t27a, 1567) - m24(t20a, (3784 - 4096))) - t20a;
    t21a = r12(-(m24(t26, (3784 - 4096)) + m24(t21, 1567))) - t26;
    t26a = r12(m24(t26, 1567) - m24(t21, (3784 - 4096))) - t21;

    t16 = c(t16a + t23a); t17a = c(t17 + t22); t18 = c(t18a + t21a); t19a = c(t19 + t20);
    t20a = c(t19 - t20); t21 = c(t18a - t21a); t22a = c(t17 - t22); t23 = c(t16a - t23a);
    t24 = c(t31a - t24a); t25a = c(t30 - t25); t26 = c(t29a - t26a); t27a = c(t28 - t27);
    t28a = c(t28 + t27); t29 = c(t29a + t26a); t30a = c(t30 + t25); t31 = c(t31a + t24a);

    t20 = r181(t27a - t20a); t27 = r181(t27a + t20a);
    t21a = r181(t26 - t21); t26a = r181(t26 + t21);
    t22 = r181(t25a - t22a); t25 = r181(t25a + t22a);
    t23a = r181(t24 - t23); t24a = r181(t24 + t23);

    const int32_t o[16] = {t31, t30a, t29, t28a, t27, t26a, t25, t24a,
                           t23a, t22, t21a, t20, t19a, t18, t17a, t16};
    int32_t e[16];
#pragma unroll
    for (int i = 0; i < 16; i++) e[i] = v[2 * i * S];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        v[i * S] = c(e[i] + o[i]);
        v[(31 - i) * S] = c(e[i] - o[i]);
    }
}

// Inverse DCT-64 (dav1d inv_dct64_1d_c with the tx64 half-zero inputs): the 32 coded inputs
// sit at the even... all of v[0..31*S]; v[32*S..63*S] are zero on entry.
template <int S> MT_DEV void idct64(int32_t* v, Clip c) {
    idct32<2 * S>(v, c);
    int32_t in[32];
#pragma unroll
    for (int i = 1; i < 32; i += 2) in[i] = v[i * S];
    int32_t t32a = r12(m24(in[1], 101)), t33a = r12(m24(in[31], -2824)), t34a = r12(m24(in[17], 1660));
    int32_t t35a = r12(m24(in[15], -1474)), t36a = r12(m24(in[9], 897)), t37a = r12(m24(in[23], -2191));
    int32_t t38a = r12(m24(in[25], 2359)), t39a = r12(m24(in[7], -700)), t40a = r12(m24(in[5], 501));
    int32_t t41a = r12(m24(in[27], -2520)), t42a = r12(m24(in[21], 2019)), t43a = r12(m24(in[11], -1092));
    int32_t t44a = r12(m24(in[13], 1285)), t45a = r12(m24(in[19], -1842)), t46a = r12(m24(in[29], 2675));
    int32_t t47a = r12(m24(in[3], -301)), t48a = r12(m24(in[3], 4085)), t49a = r12(m24(in[29], 3102));
    int32_t t50a = r12(m24(in[19], 3659)), t51a = r12(m24(in[13], 3889)), t52a = r12(m24(in[11], 3948));
    int32_t t53a = r12(m24(in[21], 3564)), t54a = r12(m24(in[27], 3229)), t55a = r12(m24(in[5], 4065));
    int32_t t56a = r12(m24(in[7], 4036)), t57a = r12(m24(in[25], 3349)), t58a = r12(m24(in[23], 3461));
    int32_t t59a = r12(m24(in[9], 3996)), t60a = r12(m24(in[15], 3822)), t61a = r12(m24(in[17], 3745));
    int32_t t62a = r12(m24(in[31], 2967)), t63a = r12(m24(in[1], 4095));
    int32_t t32 = c(t32a + t33a), t33 = c(t32a - t33a), t34 = c(t35a - t34a), t35 = c(t35a + t34a);
    int32_t t36 = c(t36a + t37a), t37 = c(t36a - t37a), t38 = c(t39a - t38a), t39 = c(t39a + t38a);
    int32_t t40 = c(t40a + t41a), t41 = c(t40a - t41a), t42 = c(t43a - t42a), t43 = c(t43a + t42a);
    int32_t t44 = c(t44a + t45a), t45 = c(t44a - t45a), t46 = c(t47a - t46a), t47 = c(t47a + t46a);
    int32_t t48 = c(t48a + t49a), t49 = c(t48a - t49a), t50 = c(t51a - t50a), t51 = c(t51a + t50a);
    int32_t t52 = c(t52a + t53a), t53 = c(t52a - t53a), t54 = c(t55a - t54a), t55 = c(t55a + t54a);
    int32_t t56 = c(t56a + t57a), t57 = c(t56a - t57a), t58 = c(t59a - t58a), t59 = c(t59a + t58a);
    int32_t t60 = c(t60a + t61a), t61 = c(t60a - t61a), t62 = c(t63a - t62a), t63 = c(t63a + t62a);
    t33a = r12(m24(t33, (4096 - 4076)) + m24(t62, 401)) - t33;
    t34a = r12(m24(t34, -401) + m24(t61, (4096 - 4076))) - t61;
    t37a = r11(m24(t37, -1299) + m24(t58, 1583));
    t38a = r11(m24(t38, -1583) + m24(t57, -1299));
    t41a = r12(m24(t41, (4096 - 3612)) + m24(t54, 1931)) - t41;
    t42a = r12(m24(t42, -1931) + m24(t53, (4096 - 3612))) - t53;
    t45a = r12(m24(t45, -1189) + m24(t50, (3920 - 4096))) + t50;
    t46a = r12(m24(t46, (4096 - 3920)) + m24(t49, -1189)) - t46;
    t49a = r12(m24(t46, -1189) + m24(t49, (3920 - 4096))) + t49;
    t50a = r12(m24(t45, (3920 - 4096)) + m24(t50, 1189)) + t45;
    t53a = r12(m24(t42, (4096 - 3612)) + m24(t53, 1931)) - t42;
    t54a = r12(m24(t41, 1931) + m24(t54, (3612 - 4096))) + t54;
    t57a = r11(m24(t38, -1299) + m24(t57, 1583));
    t58a = r11(m24(t37, 1583) + m24(t58, 1299));
    t61a = r12(m24(t34, (4096 - 4076)) + m24(t61, 401)) - t34;
    t62a = r12(m24(t33, 401) + m24(t62, (4076 - 4096))) + t62;
    t32a = c(t32 + t35); t33 = c(t33a + t34a); t34 = c(t33a - t34a); t35a = c(t32 - t35);
    t36a = c(t39 - t36); t37 = c(t38a - t37a); t38 = c(t38a + t37a); t39a = c(t39 + t36);
    t40a = c(t40 + t43); t41 = c(t41a + t42a); t42 = c(t41a - t42a); t43a = c(t40 - t43);
    t44a = c(t47 - t44); t45 = c(t46a - t45a); t46 = c(t46a + t45a); t47a = c(t47 + t44);
    t48a = c(t48 + t51); t49 = c(t49a + t50a); t50 = c(t49a - t50a); t51a = c(t48 - t51);
    t52a = c(t55 - t52); t53 = c(t54a - t53a); t54 = c(t54a + t53a); t55a = c(t55 + t52);
    t56a = c(t56 + t59); t57 = c(t57a + t58a); t58 = c(t57a - t58a); t59a = c(t56 - t59);
    t60a = c(t63 - t60); t61 = c(t62a - t61a); t62 = c(t62a + t61a); t63a = c(t63 + t60);
    t34a = r12(m24(t34, (4096 - 4017)) + m24(t61, 799)) - t34;
    t35 = r12(m24(t35a, (4096 - 4017)) + m24(t60a, 799)) - t35a;
    t36 = r12(m24(t36a, -799) + m24(t59a, (4096 - 4017))) - t59a;
    t37a = r12(m24(t37, -799) + m24(t58, (4096 - 4017))) - t58;
    t42a = r11(m24(t42, -1138) + m24(t53, 1703));
    t43 = r11(m24(t43a, -1138) + m24(t52a, 1703));
    t44 = r11(m24(t44a, -1703) + m24(t51a, -1138));
    t45a = r11(m24(t45, -1703) + m24(t50, -1138));
    t50a = r11(m24(t45, -1138) + m24(t50, 1703));
    t51 = r11(m24(t44a, -1138) + m24(t51a, 1703));
    t52 = r11(m24(t43a, 1703) + m24(t52a, 1138));
    t53a = r11(m24(t42, 1703) + m24(t53, 1138));
    t58a = r12(m24(t37, (4096 - 4017)) + m24(t58, 799)) - t37;
    t59 = r12(m24(t36a, (4096 - 4017)) + m24(t59a, 799)) - t36a;
    t60 = r12(m24(t35a, 799) + m24(t60a, (4017 - 4096))) + t60a;
    t61a = r12(m24(t34, 799) + m24(t61, (4017 - 4096))) + t61;
    t32 = c(t32a + t39a); t33a = c(t33 + t38); t34 = c(t34a + t37a); t35a = c(t35 + t36);
    t36a = c(t35 - t36); t37 = c(t34a - t37a); t38a = c(t33 - t38); t39 = c(t32a - t39a);
    t40 = c(t47a - t40a); t41a = c(t46 - t41); t42 = c(t45a - t42a); t43a = c(t44 - t43);
    t44a = c(t44 + t43); t45 = c(t45a + t42a); t46a = c(t46 + t41); t47 = c(t47a + t40a);
    t48 = c(t48a + t55a); t49a = c(t49 + t54); t50 = c(t50a + t53a); t51a = c(t51 + t52);
    t52a = c(t51 - t52); t53 = c(t50a - t53a); t54a = c(t49 - t54); t55 = c(t48a - t55a);
    t56 = c(t63a - t56a); t57a = c(t62 - t57); t58 = c(t61a - t58a); t59a = c(t60 - t59);
    t60a = c(t60 + t59); t61 = c(t61a + t58a); t62a = c(t62 + t57); t63 = c(t63a + t56a);
    t36 = r12(m24(t36a, (4096 - 3784)) + m24(t59a, 1567)) - t36a;
    t37a = r12(m24(t37, (4096 - 3784)) + m24(t58, 1567)) - t37;
    t38 = r12(m24(t38a, (4096 - 3784)) + m24(t57a, 1567)) - t38a;
    t39a = r12(m24(t39, (4096 - 3784)) + m24(t56, 1567)) - t39;
    t40a = r12(m24(t40, -1567) + m24(t55, (4096 - 3784))) - t55;
    t41 = r12(m24(t41a, -1567) + m24(t54a, (4096 - 3784))) - t54a;
    t42a = r12(m24(t42, -1567) + m24(t53, (4096 - 3784))) - t53;
    t43 = r12(m24(t43a, -1567) + m24(t52a, (4096 - 3784))) - t52a;
    t52 = r12(m24(t43a, (4096 - 3784)) + m24(t52a, 1567)) - t43a;
    t53a = r12(m24(t42, (4096 - 3784)) + m24(t53, 1567)) - t42;
    t54 = r12(m24(t41a, (4096 - 3784)) + m24(t54a, 1567)) - t41a;
    t55a = r12(m24(t40, (4096 - 3784)) + m24(t55, 1567)) - t40;
    t56a = r12(m24(t39, 1567) + m24(t56, (3784 - 4096))) + t56;
    t57 = r12(m24(t38a, 1567) + m24(t57a, (3784 - 4096))) + t57a;
    t58a = r12(m24(t37, 1567) + m24(t58, (3784 - 4096))) + t58;
    t59 = r12(m24(t36a, 1567) + m24(t59a, (3784 - 4096))) + t59a;
    t32a = c(t32 + t47); t33 = c(t33a + t46a); t34a = c(t34 + t45); t35 = c(t35a + t44a);
    t36a = c(t36 + t43); t37 = c(t37a + t42a); t38a = c(t38 + t41); t39 = c(t39a + t40a);
    t40 = c(t39a - t40a); t41a = c(t38 - t41); t42 = c(t37a - t42a); t43a = c(t36 - t43);
    t44 = c(t35a - t44a); t45a = c(t34 - t45); t46 = c(t33a - t46a); t47a = c(t32 - t47);
    t48a = c(t63 - t48); t49 = c(t62a - t49a); t50a = c(t61 - t50); t51 = c(t60a - t51a);
    t52a = c(t59 - t52); t53 = c(t58a - t53a); t54a = c(t57 - t54); t55 = c(t56a - t55a);
    t56 = c(t56a + t55a); t57a = c(t57 + t54); t58 = c(t58a + t53a); t59a = c(t59 + t52);
    t60 = c(t60a + t51a); t61a = c(t61 + t50); t62 = c(t62a + t49a); t63a = c(t63 + t48);
    t40a = r181(t55 - t40); t41 = r181(t54a - t41a); t42a = r181(t53 - t42); t43 = r181(t52a - t43a);
    t44a = r181(t51 - t44); t45 = r181(t50a - t45a); t46a = r181(t49 - t46); t47 = r181(t48a - t47a);
    t48 = r181(t47a + t48a); t49a = r181(t46 + t49); t50 = r181(t45a + t50a); t51a = r181(t44 + t51);
    t52 = r181(t43a + t52a); t53a = r181(t42 + t53); t54 = r181(t41a + t54a); t55a = r181(t40 + t55);
    const int32_t o[32] = {t63a, t62, t61a, t60, t59a, t58, t57a, t56, t55a, t54, t53a,
                           t52, t51a, t50, t49a, t48, t47, t46a, t45, t44a, t43, t42a,
                           t41, t40a, t39, t38a, t37, t36a, t35, t34a, t33, t32a};
    int32_t e[32];
#pragma unroll
    for (int i = 0; i < 32; i++) e[i] = v[2 * i * S];
#pragma unroll
    for (int i = 0; i < 32; i++) {
        v[i * S] = c(e[i] + o[i]);
        v[(63 - i) * S] = c(e[i] - o[i]);
    }
}

// Inverse ADST-4 (no internal clip).
MT_DEV void iadst4(int32_t* v) {
    const int32_t i0 = v[0], i1 = v[1], i2 = v[2], i3 = v[3];
    const int32_t o0 = r12(m24(1321, i0) + m24((3803 - 4096), i2) + m24((2482 - 4096), i3) + m24((3344 - 4096), i1)) + i2 + i3 + i1;
    const int32_t o1 = r12(m24((2482 - 4096), i0) - m24(1321, i2) - m24((3803 - 4096), i3) + m24((3344 - 4096), i1)) + i0 - i3 + i1;
    const int32_t o2 = (m24(209, (i0 - i2 + i3)) + 128) >> 8;
    const int32_t o3 = r12(m24((3803 - 4096), i0) + m24((2482 - 4096), i2) - m24(1321, i3) - m24((3344 - 4096), i1)) + i0 + i2 - i1;
    v[0] = o0;
    v[1] = o1;
    v[2] = o2;
    v[3] = o3;
}

MT_DEV void iadst8(int32_t* v, Clip c) {
    const int32_t i0 = v[0], i1 = v[1], i2 = v[2], i3 = v[3], i4 = v[4], i5 = v[5], i6 = v[6], i7 = v[7];
    int32_t t0a = r12(m24((4076 - 4096), i7) + m24(401, i0)) + i7;
    int32_t t1a = r12(m24(401, i7) - m24((4076 - 4096), i0)) - i0;
    int32_t t2a = r12(m24((3612 - 4096), i5) + m24(1931, i2)) + i5;
    int32_t t3a = r12(m24(1931, i5) - m24((3612 - 4096), i2)) - i2;
    int32_t t4a = r11(m24(1299, i3) + m24(1583, i4));
    int32_t t5a = r11(m24(1583, i3) - m24(1299, i4));
    int32_t t6a = r12(m24(1189, i1) + m24((3920 - 4096), i6)) + i6;
    int32_t t7a = r12(m24((3920 - 4096), i1) - m24(1189, i6)) + i1;
    int32_t t0 = c(t0a + t4a), t1 = c(t1a + t5a), t2 = c(t2a + t6a), t3 = c(t3a + t7a);
    int32_t t4 = c(t0a - t4a), t5 = c(t1a - t5a), t6 = c(t2a - t6a), t7 = c(t3a - t7a);
    t4a = r12(m24((3784 - 4096), t4) + m24(1567, t5)) + t4;
    t5a = r12(m24(1567, t4) - m24((3784 - 4096), t5)) - t5;
    t6a = r12(m24((3784 - 4096), t7) - m24(1567, t6)) + t7;
    t7a = r12(m24(1567, t7) + m24((3784 - 4096), t6)) + t6;
    v[0] = c(t0 + t2);
    v[7] = -c(t1 + t3);
    t2 = c(t0 - t2);
    t3 = c(t1 - t3);
    v[1] = -c(t4a + t6a);
    v[6] = c(t5a + t7a);
    t6 = c(t4a - t6a);
    t7 = c(t5a - t7a);
    v[3] = -r181(t2 + t3);
    v[4] = r181(t2 - t3);
    v[2] = r181(t6 + t7);
    v[5] = -r181(t6 - t7);
}

MT_DEV void iadst16(int32_t* v, Clip c) {
    int32_t in[16];
#pragma unroll
    for (int i = 0; i < 16; i++) in[i] = v[i];
    int32_t t0 = r12(m24(in[15], (4091 - 4096)) + m24(in[0], 201)) + in[15];
    int32_t t1 = r12(m24(in[15], 201) - m24(in[0], (4091 - 4096))) - in[0];
    int32_t t2 = r12(m24(in[13], (3973 - 4096)) + m24(in[2], 995)) + in[13];
    int32_t t3 = r12(m24(in[13], 995) - m24(in[2], (3973 - 4096))) - in[2];
    int32_t t4 = r12(m24(in[11], (3703 - 4096)) + m24(in[4], 1751)) + in[11];
    int32_t t5 = r12(m24(in[11], 1751) - m24(in[4], (3703 - 4096))) - in[4];
    int32_t t6 = r11(m24(in[9], 1645) + m24(in[6], 1220));
    int32_t t7 = r11(m24(in[9], 1220) - m24(in[6], 1645));
    int32_t t8 = r12(m24(in[7], 2751) + m24(in[8], (3035 - 4096))) + in[8];
    int32_t t9 = r12(m24(in[7], (3035 - 4096)) - m24(in[8], 2751)) + in[7];
    int32_t t10 = r12(m24(in[5], 2106) + m24(in[10], (3513 - 4096))) + in[10];
    int32_t t11 = r12(m24(in[5], (3513 - 4096)) - m24(in[10], 2106)) + in[5];
    int32_t t12 = r12(m24(in[3], 1380) + m24(in[12], (3857 - 4096))) + in[12];
    int32_t t13 = r12(m24(in[3], (3857 - 4096)) - m24(in[12], 1380)) + in[3];
    int32_t t14 = r12(m24(in[1], 601) + m24(in[14], (4052 - 4096))) + in[14];
    int32_t t15 = r12(m24(in[1], (4052 - 4096)) - m24(in[14], 601)) + in[1];
    int32_t t0a = c(t0 + t8), t1a = c(t1 + t9), t2a = c(t2 + t10), t3a = c(t3 + t11);
    int32_t t4a = c(t4 + t12), t5a = c(t5 + t13), t6a = c(t6 + t14), t7a = c(t7 + t15);
    int32_t t8a = c(t0 - t8), t9a = c(t1 - t9), t10a = c(t2 - t10), t11a = c(t3 - t11);
    int32_t t12a = c(t4 - t12), t13a = c(t5 - t13), t14a = c(t6 - t14), t15a = c(t7 - t15);
    t8 = r12(m24(t8a, (4017 - 4096)) + m24(t9a, 799)) + t8a;
    t9 = r12(m24(t8a, 799) - m24(t9a, (4017 - 4096))) - t9a;
    t10 = r12(m24(t10a, 2276) + m24(t11a, (3406 - 4096))) + t11a;
    t11 = r12(m24(t10a, (3406 - 4096)) - m24(t11a, 2276)) + t10a;
    t12 = r12(m24(t13a, (4017 - 4096)) - m24(t12a, 799)) + t13a;
    t13 = r12(m24(t13a, 799) + m24(t12a, (4017 - 4096))) + t12a;
    t14 = r12(m24(t15a, 2276) - m24(t14a, (3406 - 4096))) - t14a;
    t15 = r12(m24(t15a, (3406 - 4096)) + m24(t14a, 2276)) + t15a;
    t0 = c(t0a + t4a); t1 = c(t1a + t5a); t2 = c(t2a + t6a); t3 = c(t3a + t7a);
    t4 = c(t0a - t4a); t5 = c(t1a - t5a); t6 = c(t2a - t6a); t7 = c(t3a - t7a);
    t8a = c(t8 + t12); t9a = c(t9 + t13); t10a = c(t10 + t14); t11a = c(t11 + t15);
    t12a = c(t8 - t12); t13a = c(t9 - t13); t14a = c(t10 - t14); t15a = c(t11 - t15);
    t4a = r12(m24(t4, (3784 - 4096)) + m24(t5, 1567)) + t4;
    t5a = r12(m24(t4, 1567) - m24(t5, (3784 - 4096))) - t5;
    t6a = r12(m24(t7, (3784 - 4096)) - m24(t6, 1567)) + t7;
    t7a = r12(m24(t7, 1567) + m24(t6, (3784 - 4096))) + t6;
    t12 = r12(m24(t12a, (3784 - 4096)) + m24(t13a, 1567)) + t12a;
    t13 = r12(m24(t12a, 1567) - m24(t13a, (3784 - 4096))) - t13a;
    t14 = r12(m24(t15a, (3784 - 4096)) - m24(t14a, 1567)) + t15a;
    t15 = r12(m24(t15a, 1567) + m24(t14a, (3784 - 4096))) + t14a;
    v[0] = c(t0 + t2);
    v[15] = -c(t1 + t3);
    t2a = c(t0 - t2);
    t3a = c(t1 - t3);
    v[3] = -c(t4a + t6a);
    v[12] = c(t5a + t7a);
    t6 = c(t4a - t6a);
    t7 = c(t5a - t7a);
    v[1] = -c(t8a + t10a);
    v[14] = c(t9a + t11a);
    t10 = c(t8a - t10a);
    t11 = c(t9a - t11a);
    v[2] = c(t12 + t14);
    v[13] = -c(t13 + t15);
    t14a = c(t12 - t14);
    t15a = c(t13 - t15);
    v[7] = -r181(t2a + t3a);
    v[8] = r181(t2a - t3a);
    v[4] = r181(t6 + t7);
    v[11] = -r181(t6 - t7);
    v[6] = r181(t10 + t11);
    v[9] = -r181(t10 - t11);
    v[5] = -r181(t14a + t15a);
    v[10] = r181(t14a - t15a);
}

// Identity scalings (no internal clamp): x sqrt2, x 2, x 2 sqrt2 in the normative roundings.
MT_DEV int32_t iid4(int32_t v) { return v + r12(m24(v, 1697)); }
MT_DEV int32_t iid8(int32_t v) { return v * 2; }
MT_DEV int32_t iid16(int32_t v) { return 2 * v + r11(m24(v, 1697)); }

}  // namespace mt
