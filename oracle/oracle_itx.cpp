// CPU oracle: dequantization + AV1 inverse transforms, SSE, microbench chain.
//
// TEST INFRASTRUCTURE ONLY -- the checker for the HIP library, never part of the product.
// Restates the reference's dav1d-exact inverse (src/idct.rs) function by function.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <thread>
#include <vector>

#include "oracle.h"

namespace {

enum { DCT_DCT = 0, ADST_DCT = 1, DCT_ADST = 2, ADST_ADST = 3, IDTX = 9, V_DCT = 10, H_DCT = 11 };
enum {
    TX_4X4 = 0, TX_8X8 = 1, TX_16X16 = 2, TX_32X32 = 3, TX_64X64 = 4, TX_4X8 = 5, TX_8X4 = 6,
    TX_8X16 = 7, TX_16X8 = 8, TX_16X32 = 9, TX_32X16 = 10, TX_4X16 = 13, TX_16X4 = 14
};

inline int32_t clip(int32_t x, int32_t mn, int32_t mx) { return x < mn ? mn : x > mx ? mx : x; }

// inv_dct4_1d (src/idct.rs:1432)
void inv_dct4(int32_t* c, int s, int mn, int mx) {
    int32_t in0 = c[0], in1 = c[s], in2 = c[2 * s], in3 = c[3 * s];
    int32_t t0 = ((in0 + in2) * 181 + 128) >> 8;
    int32_t t1 = ((in0 - in2) * 181 + 128) >> 8;
    int32_t t2 = ((in1 * 1567 - in3 * (3784 - 4096) + 2048) >> 12) - in3;
    int32_t t3 = ((in1 * (3784 - 4096) + in3 * 1567 + 2048) >> 12) + in1;
    c[0] = clip(t0 + t3, mn, mx);
    c[s] = clip(t1 + t2, mn, mx);
    c[2 * s] = clip(t1 - t2, mn, mx);
    c[3 * s] = clip(t0 - t3, mn, mx);
}

// inv_dct8_1d (src/idct.rs:1446)
void inv_dct8(int32_t* c, int s, int mn, int mx) {
    inv_dct4(c, 2 * s, mn, mx);
    int32_t in1 = c[s], in3 = c[3 * s], in5 = c[5 * s], in7 = c[7 * s];
    int32_t t4a = ((in1 * 799 - in7 * (4017 - 4096) + 2048) >> 12) - in7;
    int32_t t5a = (in5 * 1703 - in3 * 1138 + 1024) >> 11;
    int32_t t6a = (in5 * 1138 + in3 * 1703 + 1024) >> 11;
    int32_t t7a = ((in1 * (4017 - 4096) + in7 * 799 + 2048) >> 12) + in1;
    int32_t t4 = clip(t4a + t5a, mn, mx);
    t5a = clip(t4a - t5a, mn, mx);
    int32_t t7 = clip(t7a + t6a, mn, mx);
    t6a = clip(t7a - t6a, mn, mx);
    int32_t t5 = ((t6a - t5a) * 181 + 128) >> 8;
    int32_t t6 = ((t6a + t5a) * 181 + 128) >> 8;
    int32_t t0 = c[0], t1 = c[2 * s], t2 = c[4 * s], t3 = c[6 * s];
    c[0] = clip(t0 + t7, mn, mx);
    c[s] = clip(t1 + t6, mn, mx);
    c[2 * s] = clip(t2 + t5, mn, mx);
    c[3 * s] = clip(t3 + t4, mn, mx);
    c[4 * s] = clip(t3 - t4, mn, mx);
    c[5 * s] = clip(t2 - t5, mn, mx);
    c[6 * s] = clip(t1 - t6, mn, mx);
    c[7 * s] = clip(t0 - t7, mn, mx);
}

// inv_dct16_1d (src/idct.rs:1863)
void inv_dct16(int32_t* c, int s, int mn, int mx) {
    inv_dct8(c, 2 * s, mn, mx);
    int32_t in1 = c[s], in3 = c[3 * s], in5 = c[5 * s], in7 = c[7 * s];
    int32_t in9 = c[9 * s], in11 = c[11 * s], in13 = c[13 * s], in15 = c[15 * s];
    int32_t t8a = ((in1 * 401 - in15 * (4076 - 4096) + 2048) >> 12) - in15;
    int32_t t9a = (in9 * 1583 - in7 * 1299 + 1024) >> 11;
    int32_t t10a = ((in5 * 1931 - in11 * (3612 - 4096) + 2048) >> 12) - in11;
    int32_t t11a = ((in13 * (3920 - 4096) - in3 * 1189 + 2048) >> 12) + in13;
    int32_t t12a = ((in13 * 1189 + in3 * (3920 - 4096) + 2048) >> 12) + in3;
    int32_t t13a = ((in5 * (3612 - 4096) + in11 * 1931 + 2048) >> 12) + in5;
    int32_t t14a = (in9 * 1299 + in7 * 1583 + 1024) >> 11;
    int32_t t15a = ((in1 * (4076 - 4096) + in15 * 401 + 2048) >> 12) + in1;
    int32_t t8 = clip(t8a + t9a, mn, mx);
    int32_t t9 = clip(t8a - t9a, mn, mx);
    int32_t t10 = clip(t11a - t10a, mn, mx);
    int32_t t11 = clip(t11a + t10a, mn, mx);
    int32_t t12 = clip(t12a + t13a, mn, mx);
    int32_t t13 = clip(t12a - t13a, mn, mx);
    int32_t t14 = clip(t15a - t14a, mn, mx);
    int32_t t15 = clip(t15a + t14a, mn, mx);
    t9a = ((t14 * 1567 - t9 * (3784 - 4096) + 2048) >> 12) - t9;
    t14a = ((t14 * (3784 - 4096) + t9 * 1567 + 2048) >> 12) + t14;
    t10a = ((-(t13 * (3784 - 4096) + t10 * 1567) + 2048) >> 12) - t13;
    t13a = ((t13 * 1567 - t10 * (3784 - 4096) + 2048) >> 12) - t10;
    t8a = clip(t8 + t11, mn, mx);
    t9 = clip(t9a + t10a, mn, mx);
    t10 = clip(t9a - t10a, mn, mx);
    t11a = clip(t8 - t11, mn, mx);
    t12a = clip(t15 - t12, mn, mx);
    t13 = clip(t14a - t13a, mn, mx);
    t14 = clip(t14a + t13a, mn, mx);
    t15a = clip(t15 + t12, mn, mx);
    t10a = ((t13 - t10) * 181 + 128) >> 8;
    t13a = ((t13 + t10) * 181 + 128) >> 8;
    t11 = ((t12a - t11a) * 181 + 128) >> 8;
    t12 = ((t12a + t11a) * 181 + 128) >> 8;
    int32_t t0 = c[0], t1 = c[2 * s], t2 = c[4 * s], t3 = c[6 * s];
    int32_t t4 = c[8 * s], t5 = c[10 * s], t6 = c[12 * s], t7 = c[14 * s];
    c[0] = clip(t0 + t15a, mn, mx);
    c[s] = clip(t1 + t14, mn, mx);
    c[2 * s] = clip(t2 + t13a, mn, mx);
    c[3 * s] = clip(t3 + t12, mn, mx);
    c[4 * s] = clip(t4 + t11, mn, mx);
    c[5 * s] = clip(t5 + t10a, mn, mx);
    c[6 * s] = clip(t6 + t9, mn, mx);
    c[7 * s] = clip(t7 + t8a, mn, mx);
    c[8 * s] = clip(t7 - t8a, mn, mx);
    c[9 * s] = clip(t6 - t9, mn, mx);
    c[10 * s] = clip(t5 - t10a, mn, mx);
    c[11 * s] = clip(t4 - t11, mn, mx);
    c[12 * s] = clip(t3 - t12, mn, mx);
    c[13 * s] = clip(t2 - t13a, mn, mx);
    c[14 * s] = clip(t1 - t14, mn, mx);
    c[15 * s] = clip(t0 - t15a, mn, mx);
}

// inv_dct32_1d (src/idct.rs:2130)
void inv_dct32(int32_t* c, int s, int mn, int mx) {
    inv_dct16(c, 2 * s, mn, mx);
    int32_t in1 = c[s], in3 = c[3 * s], in5 = c[5 * s], in7 = c[7 * s];
    int32_t in9 = c[9 * s], in11 = c[11 * s], in13 = c[13 * s], in15 = c[15 * s];
    int32_t in17 = c[17 * s], in19 = c[19 * s], in21 = c[21 * s], in23 = c[23 * s];
    int32_t in25 = c[25 * s], in27 = c[27 * s], in29 = c[29 * s], in31 = c[31 * s];
    int32_t t16a = ((in1 * 201 - in31 * (4091 - 4096) + 2048) >> 12) - in31;
    int32_t t17a = ((in17 * (3035 - 4096) - in15 * 2751 + 2048) >> 12) + in17;
    int32_t t18a = ((in9 * 1751 - in23 * (3703 - 4096) + 2048) >> 12) - in23;
    int32_t t19a = ((in25 * (3857 - 4096) - in7 * 1380 + 2048) >> 12) + in25;
    int32_t t20a = ((in5 * 995 - in27 * (3973 - 4096) + 2048) >> 12) - in27;
    int32_t t21a = ((in21 * (3513 - 4096) - in11 * 2106 + 2048) >> 12) + in21;
    int32_t t22a = (in13 * 1220 - in19 * 1645 + 1024) >> 11;
    int32_t t23a = ((in29 * (4052 - 4096) - in3 * 601 + 2048) >> 12) + in29;
    int32_t t24a = ((in29 * 601 + in3 * (4052 - 4096) + 2048) >> 12) + in3;
    int32_t t25a = (in13 * 1645 + in19 * 1220 + 1024) >> 11;
    int32_t t26a = ((in21 * 2106 + in11 * (3513 - 4096) + 2048) >> 12) + in11;
    int32_t t27a = ((in5 * (3973 - 4096) + in27 * 995 + 2048) >> 12) + in5;
    int32_t t28a = ((in25 * 1380 + in7 * (3857 - 4096) + 2048) >> 12) + in7;
    int32_t t29a = ((in9 * (3703 - 4096) + in23 * 1751 + 2048) >> 12) + in9;
    int32_t t30a = ((in17 * 2751 + in15 * (3035 - 4096) + 2048) >> 12) + in15;
    int32_t t31a = ((in1 * (4091 - 4096) + in31 * 201 + 2048) >> 12) + in1;

    int32_t t16 = clip(t16a + t17a, mn, mx);
    int32_t t17 = clip(t16a - t17a, mn, mx);
    int32_t t18 = clip(t19a - t18a, mn, mx);
    int32_t t19 = clip(t19a + t18a, mn, mx);
    int32_t t20 = clip(t20a + t21a, mn, mx);
    int32_t t21 = clip(t20a - t21a, mn, mx);
    int32_t t22 = clip(t23a - t22a, mn, mx);
    int32_t t23 = clip(t23a + t22a, mn, mx);
    int32_t t24 = clip(t24a + t25a, mn, mx);
    int32_t t25 = clip(t24a - t25a, mn, mx);
    int32_t t26 = clip(t27a - t26a, mn, mx);
    int32_t t27 = clip(t27a + t26a, mn, mx);
    int32_t t28 = clip(t28a + t29a, mn, mx);
    int32_t t29 = clip(t28a - t29a, mn, mx);
    int32_t t30 = clip(t31a - t30a, mn, mx);
    int32_t t31 = clip(t31a + t30a, mn, mx);

    t17a = ((t30 * 799 - t17 * (4017 - 4096) + 2048) >> 12) - t17;
    t30a = ((t30 * (4017 - 4096) + t17 * 799 + 2048) >> 12) + t30;
    t18a = ((-(t29 * (4017 - 4096) + t18 * 799) + 2048) >> 12) - t29;
    t29a = ((t29 * 799 - t18 * (4017 - 4096) + 2048) >> 12) - t18;
    t21a = (t26 * 1703 - t21 * 1138 + 1024) >> 11;
    t26a = (t26 * 1138 + t21 * 1703 + 1024) >> 11;
    t22a = (-(t25 * 1138 + t22 * 1703) + 1024) >> 11;
    t25a = (t25 * 1703 - t22 * 1138 + 1024) >> 11;

    t16a = clip(t16 + t19, mn, mx);
    t17 = clip(t17a + t18a, mn, mx);
    t18 = clip(t17a - t18a, mn, mx);
    t19a = clip(t16 - t19, mn, mx);
    t20a = clip(t23 - t20, mn, mx);
    t21 = clip(t22a - t21a, mn, mx);
    t22 = clip(t22a + t21a, mn, mx);
    t23a = clip(t23 + t20, mn, mx);
    t24a = clip(t24 + t27, mn, mx);
    t25 = clip(t25a + t26a, mn, mx);
    t26 = clip(t25a - t26a, mn, mx);
    t27a = clip(t24 - t27, mn, mx);
    t28a = clip(t31 - t28, mn, mx);
    t29 = clip(t30a - t29a, mn, mx);
    t30 = clip(t30a + t29a, mn, mx);
    t31a = clip(t31 + t28, mn, mx);

    t18a = ((t29 * 1567 - t18 * (3784 - 4096) + 2048) >> 12) - t18;
    t29a = ((t29 * (3784 - 4096) + t18 * 1567 + 2048) >> 12) + t29;
    t19 = ((t28a * 1567 - t19a * (3784 - 4096) + 2048) >> 12) - t19a;
    t28 = ((t28a * (3784 - 4096) + t19a * 1567 + 2048) >> 12) + t28a;
    t20 = ((-(t27a * (3784 - 4096) + t20a * 1567) + 2048) >> 12) - t27a;
    t27 = ((t27a * 1567 - t20a * (3784 - 4096) + 2048) >> 12) - t20a;
    t21a = ((-(t26 * (3784 - 4096) + t21 * 1567) + 2048) >> 12) - t26;
    t26a = ((t26 * 1567 - t21 * (3784 - 4096) + 2048) >> 12) - t21;

    t16 = clip(t16a + t23a, mn, mx);
    t17a = clip(t17 + t22, mn, mx);
    t18 = clip(t18a + t21a, mn, mx);
    t19a = clip(t19 + t20, mn, mx);
    t20a = clip(t19 - t20, mn, mx);
    t21 = clip(t18a - t21a, mn, mx);
    t22a = clip(t17 - t22, mn, mx);
    t23 = clip(t16a - t23a, mn, mx);
    t24 = clip(t31a - t24a, mn, mx);
    t25a = clip(t30 - t25, mn, mx);
    t26 = clip(t29a - t26a, mn, mx);
    t27a = clip(t28 - t27, mn, mx);
    t28a = clip(t28 + t27, mn, mx);
    t29 = clip(t29a + t26a, mn, mx);
    t30a = clip(t30 + t25, mn, mx);
    t31 = clip(t31a + t24a, mn, mx);

    t20 = ((t27a - t20a) * 181 + 128) >> 8;
    t27 = ((t27a + t20a) * 181 + 128) >> 8;
    t21a = ((t26 - t21) * 181 + 128) >> 8;
    t26a = ((t26 + t21) * 181 + 128) >> 8;
    t22 = ((t25a - t22a) * 181 + 128) >> 8;
    t25 = ((t25a + t22a) * 181 + 128) >> 8;
    t23a = ((t24 - t23) * 181 + 128) >> 8;
    t24a = ((t24 + t23) * 181 + 128) >> 8;

    int32_t e[16];
    for (int i = 0; i < 16; i++) e[i] = c[2 * i * s];
    const int32_t o[16] = {t31, t30a, t29, t28a, t27, t26a, t25, t24a,
                           t23a, t22, t21a, t20, t19a, t18, t17a, t16};
    for (int i = 0; i < 16; i++) {
        c[i * s] = clip(e[i] + o[i], mn, mx);
        c[(31 - i) * s] = clip(e[i] - o[i], mn, mx);
    }
}

// inv_dct64_1d (src/idct.rs:2331): only the first 32 inputs may be nonzero.
void inv_dct64(int32_t* c, int s, int mn, int mx) {
    inv_dct32(c, 2 * s, mn, mx);
    int32_t in[32];
    for (int i = 1; i < 32; i += 2) in[i] = c[i * s];
    auto r12 = [](int32_t v, int32_t k) { return (v * k + 2048) >> 12; };
    int32_t t32a = r12(in[1], 101), t33a = r12(in[31], -2824), t34a = r12(in[17], 1660);
    int32_t t35a = r12(in[15], -1474), t36a = r12(in[9], 897), t37a = r12(in[23], -2191);
    int32_t t38a = r12(in[25], 2359), t39a = r12(in[7], -700), t40a = r12(in[5], 501);
    int32_t t41a = r12(in[27], -2520), t42a = r12(in[21], 2019), t43a = r12(in[11], -1092);
    int32_t t44a = r12(in[13], 1285), t45a = r12(in[19], -1842), t46a = r12(in[29], 2675);
    int32_t t47a = r12(in[3], -301), t48a = r12(in[3], 4085), t49a = r12(in[29], 3102);
    int32_t t50a = r12(in[19], 3659), t51a = r12(in[13], 3889), t52a = r12(in[11], 3948);
    int32_t t53a = r12(in[21], 3564), t54a = r12(in[27], 3229), t55a = r12(in[5], 4065);
    int32_t t56a = r12(in[7], 4036), t57a = r12(in[25], 3349), t58a = r12(in[23], 3461);
    int32_t t59a = r12(in[9], 3996), t60a = r12(in[15], 3822), t61a = r12(in[17], 3745);
    int32_t t62a = r12(in[31], 2967), t63a = r12(in[1], 4095);
    auto C = [&](int32_t v) { return clip(v, mn, mx); };
    int32_t t32 = C(t32a + t33a), t33 = C(t32a - t33a), t34 = C(t35a - t34a), t35 = C(t35a + t34a);
    int32_t t36 = C(t36a + t37a), t37 = C(t36a - t37a), t38 = C(t39a - t38a), t39 = C(t39a + t38a);
    int32_t t40 = C(t40a + t41a), t41 = C(t40a - t41a), t42 = C(t43a - t42a), t43 = C(t43a + t42a);
    int32_t t44 = C(t44a + t45a), t45 = C(t44a - t45a), t46 = C(t47a - t46a), t47 = C(t47a + t46a);
    int32_t t48 = C(t48a + t49a), t49 = C(t48a - t49a), t50 = C(t51a - t50a), t51 = C(t51a + t50a);
    int32_t t52 = C(t52a + t53a), t53 = C(t52a - t53a), t54 = C(t55a - t54a), t55 = C(t55a + t54a);
    int32_t t56 = C(t56a + t57a), t57 = C(t56a - t57a), t58 = C(t59a - t58a), t59 = C(t59a + t58a);
    int32_t t60 = C(t60a + t61a), t61 = C(t60a - t61a), t62 = C(t63a - t62a), t63 = C(t63a + t62a);
    t33a = ((t33 * (4096 - 4076) + t62 * 401 + 2048) >> 12) - t33;
    t34a = ((t34 * -401 + t61 * (4096 - 4076) + 2048) >> 12) - t61;
    t37a = (t37 * -1299 + t58 * 1583 + 1024) >> 11;
    t38a = (t38 * -1583 + t57 * -1299 + 1024) >> 11;
    t41a = ((t41 * (4096 - 3612) + t54 * 1931 + 2048) >> 12) - t41;
    t42a = ((t42 * -1931 + t53 * (4096 - 3612) + 2048) >> 12) - t53;
    t45a = ((t45 * -1189 + t50 * (3920 - 4096) + 2048) >> 12) + t50;
    t46a = ((t46 * (4096 - 3920) + t49 * -1189 + 2048) >> 12) - t46;
    t49a = ((t46 * -1189 + t49 * (3920 - 4096) + 2048) >> 12) + t49;
    t50a = ((t45 * (3920 - 4096) + t50 * 1189 + 2048) >> 12) + t45;
    t53a = ((t42 * (4096 - 3612) + t53 * 1931 + 2048) >> 12) - t42;
    t54a = ((t41 * 1931 + t54 * (3612 - 4096) + 2048) >> 12) + t54;
    t57a = (t38 * -1299 + t57 * 1583 + 1024) >> 11;
    t58a = (t37 * 1583 + t58 * 1299 + 1024) >> 11;
    t61a = ((t34 * (4096 - 4076) + t61 * 401 + 2048) >> 12) - t34;
    t62a = ((t33 * 401 + t62 * (4076 - 4096) + 2048) >> 12) + t62;
    t32a = C(t32 + t35); t33 = C(t33a + t34a); t34 = C(t33a - t34a); t35a = C(t32 - t35);
    t36a = C(t39 - t36); t37 = C(t38a - t37a); t38 = C(t38a + t37a); t39a = C(t39 + t36);
    t40a = C(t40 + t43); t41 = C(t41a + t42a); t42 = C(t41a - t42a); t43a = C(t40 - t43);
    t44a = C(t47 - t44); t45 = C(t46a - t45a); t46 = C(t46a + t45a); t47a = C(t47 + t44);
    t48a = C(t48 + t51); t49 = C(t49a + t50a); t50 = C(t49a - t50a); t51a = C(t48 - t51);
    t52a = C(t55 - t52); t53 = C(t54a - t53a); t54 = C(t54a + t53a); t55a = C(t55 + t52);
    t56a = C(t56 + t59); t57 = C(t57a + t58a); t58 = C(t57a - t58a); t59a = C(t56 - t59);
    t60a = C(t63 - t60); t61 = C(t62a - t61a); t62 = C(t62a + t61a); t63a = C(t63 + t60);
    t34a = ((t34 * (4096 - 4017) + t61 * 799 + 2048) >> 12) - t34;
    t35 = ((t35a * (4096 - 4017) + t60a * 799 + 2048) >> 12) - t35a;
    t36 = ((t36a * -799 + t59a * (4096 - 4017) + 2048) >> 12) - t59a;
    t37a = ((t37 * -799 + t58 * (4096 - 4017) + 2048) >> 12) - t58;
    t42a = (t42 * -1138 + t53 * 1703 + 1024) >> 11;
    t43 = (t43a * -1138 + t52a * 1703 + 1024) >> 11;
    t44 = (t44a * -1703 + t51a * -1138 + 1024) >> 11;
    t45a = (t45 * -1703 + t50 * -1138 + 1024) >> 11;
    t50a = (t45 * -1138 + t50 * 1703 + 1024) >> 11;
    t51 = (t44a * -1138 + t51a * 1703 + 1024) >> 11;
    t52 = (t43a * 1703 + t52a * 1138 + 1024) >> 11;
    t53a = (t42 * 1703 + t53 * 1138 + 1024) >> 11;
    t58a = ((t37 * (4096 - 4017) + t58 * 799 + 2048) >> 12) - t37;
    t59 = ((t36a * (4096 - 4017) + t59a * 799 + 2048) >> 12) - t36a;
    t60 = ((t35a * 799 + t60a * (4017 - 4096) + 2048) >> 12) + t60a;
    t61a = ((t34 * 799 + t61 * (4017 - 4096) + 2048) >> 12) + t61;
    t32 = C(t32a + t39a); t33a = C(t33 + t38); t34 = C(t34a + t37a); t35a = C(t35 + t36);
    t36a = C(t35 - t36); t37 = C(t34a - t37a); t38a = C(t33 - t38); t39 = C(t32a - t39a);
    t40 = C(t47a - t40a); t41a = C(t46 - t41); t42 = C(t45a - t42a); t43a = C(t44 - t43);
    t44a = C(t44 + t43); t45 = C(t45a + t42a); t46a = C(t46 + t41); t47 = C(t47a + t40a);
    t48 = C(t48a + t55a); t49a = C(t49 + t54); t50 = C(t50a + t53a); t51a = C(t51 + t52);
    t52a = C(t51 - t52); t53 = C(t50a - t53a); t54a = C(t49 - t54); t55 = C(t48a - t55a);
    t56 = C(t63a - t56a); t57a = C(t62 - t57); t58 = C(t61a - t58a); t59a = C(t60 - t59);
    t60a = C(t60 + t59); t61 = C(t61a + t58a); t62a = C(t62 + t57); t63 = C(t63a + t56a);
    t36 = ((t36a * (4096 - 3784) + t59a * 1567 + 2048) >> 12) - t36a;
    t37a = ((t37 * (4096 - 3784) + t58 * 1567 + 2048) >> 12) - t37;
    t38 = ((t38a * (4096 - 3784) + t57a * 1567 + 2048) >> 12) - t38a;
    t39a = ((t39 * (4096 - 3784) + t56 * 1567 + 2048) >> 12) - t39;
    t40a = ((t40 * -1567 + t55 * (4096 - 3784) + 2048) >> 12) - t55;
    t41 = ((t41a * -1567 + t54a * (4096 - 3784) + 2048) >> 12) - t54a;
    t42a = ((t42 * -1567 + t53 * (4096 - 3784) + 2048) >> 12) - t53;
    t43 = ((t43a * -1567 + t52a * (4096 - 3784) + 2048) >> 12) - t52a;
    t52 = ((t43a * (4096 - 3784) + t52a * 1567 + 2048) >> 12) - t43a;
    t53a = ((t42 * (4096 - 3784) + t53 * 1567 + 2048) >> 12) - t42;
    t54 = ((t41a * (4096 - 3784) + t54a * 1567 + 2048) >> 12) - t41a;
    t55a = ((t40 * (4096 - 3784) + t55 * 1567 + 2048) >> 12) - t40;
    t56a = ((t39 * 1567 + t56 * (3784 - 4096) + 2048) >> 12) + t56;
    t57 = ((t38a * 1567 + t57a * (3784 - 4096) + 2048) >> 12) + t57a;
    t58a = ((t37 * 1567 + t58 * (3784 - 4096) + 2048) >> 12) + t58;
    t59 = ((t36a * 1567 + t59a * (3784 - 4096) + 2048) >> 12) + t59a;
    t32a = C(t32 + t47); t33 = C(t33a + t46a); t34a = C(t34 + t45); t35 = C(t35a + t44a);
    t36a = C(t36 + t43); t37 = C(t37a + t42a); t38a = C(t38 + t41); t39 = C(t39a + t40a);
    t40 = C(t39a - t40a); t41a = C(t38 - t41); t42 = C(t37a - t42a); t43a = C(t36 - t43);
    t44 = C(t35a - t44a); t45a = C(t34 - t45); t46 = C(t33a - t46a); t47a = C(t32 - t47);
    t48a = C(t63 - t48); t49 = C(t62a - t49a); t50a = C(t61 - t50); t51 = C(t60a - t51a);
    t52a = C(t59 - t52); t53 = C(t58a - t53a); t54a = C(t57 - t54); t55 = C(t56a - t55a);
    t56 = C(t56a + t55a); t57a = C(t57 + t54); t58 = C(t58a + t53a); t59a = C(t59 + t52);
    t60 = C(t60a + t51a); t61a = C(t61 + t50); t62 = C(t62a + t49a); t63a = C(t63 + t48);
    auto r8 = [](int32_t v) { return (v * 181 + 128) >> 8; };
    t40a = r8(t55 - t40); t41 = r8(t54a - t41a); t42a = r8(t53 - t42); t43 = r8(t52a - t43a);
    t44a = r8(t51 - t44); t45 = r8(t50a - t45a); t46a = r8(t49 - t46); t47 = r8(t48a - t47a);
    t48 = r8(t47a + t48a); t49a = r8(t46 + t49); t50 = r8(t45a + t50a); t51a = r8(t44 + t51);
    t52 = r8(t43a + t52a); t53a = r8(t42 + t53); t54 = r8(t41a + t54a); t55a = r8(t40 + t55);
    int32_t e[32];
    for (int i = 0; i < 32; i++) e[i] = c[2 * i * s];
    const int32_t o[32] = {t63a, t62, t61a, t60, t59a, t58, t57a, t56, t55a, t54, t53a,
                           t52, t51a, t50, t49a, t48, t47, t46a, t45, t44a, t43, t42a,
                           t41, t40a, t39, t38a, t37, t36a, t35, t34a, t33, t32a};
    for (int i = 0; i < 32; i++) {
        c[i * s] = C(e[i] + o[i]);
        c[(63 - i) * s] = C(e[i] - o[i]);
    }
}

// inv_adst4_1d (src/idct.rs:2837): no internal clip.
void inv_adst4(int32_t* c, int s, int, int) {
    int32_t in0 = c[0], in1 = c[s], in2 = c[2 * s], in3 = c[3 * s];
    int32_t o0 = ((1321 * in0 + (3803 - 4096) * in2 + (2482 - 4096) * in3 + (3344 - 4096) * in1 + 2048) >> 12) +
                 in2 + in3 + in1;
    int32_t o1 = (((2482 - 4096) * in0 - 1321 * in2 - (3803 - 4096) * in3 + (3344 - 4096) * in1 + 2048) >> 12) +
                 in0 - in3 + in1;
    int32_t o2 = (209 * (in0 - in2 + in3) + 128) >> 8;
    int32_t o3 = (((3803 - 4096) * in0 + (2482 - 4096) * in2 - 1321 * in3 - (3344 - 4096) * in1 + 2048) >> 12) +
                 in0 + in2 - in1;
    c[0] = o0;
    c[s] = o1;
    c[2 * s] = o2;
    c[3 * s] = o3;
}

// inv_adst8_1d (src/idct.rs:1473)
void inv_adst8(int32_t* c, int s, int mn, int mx) {
    int32_t in0 = c[0], in1 = c[s], in2 = c[2 * s], in3 = c[3 * s];
    int32_t in4 = c[4 * s], in5 = c[5 * s], in6 = c[6 * s], in7 = c[7 * s];
    int32_t t0a = (((4076 - 4096) * in7 + 401 * in0 + 2048) >> 12) + in7;
    int32_t t1a = ((401 * in7 - (4076 - 4096) * in0 + 2048) >> 12) - in0;
    int32_t t2a = (((3612 - 4096) * in5 + 1931 * in2 + 2048) >> 12) + in5;
    int32_t t3a = ((1931 * in5 - (3612 - 4096) * in2 + 2048) >> 12) - in2;
    int32_t t4a = (1299 * in3 + 1583 * in4 + 1024) >> 11;
    int32_t t5a = (1583 * in3 - 1299 * in4 + 1024) >> 11;
    int32_t t6a = ((1189 * in1 + (3920 - 4096) * in6 + 2048) >> 12) + in6;
    int32_t t7a = (((3920 - 4096) * in1 - 1189 * in6 + 2048) >> 12) + in1;
    int32_t t0 = clip(t0a + t4a, mn, mx);
    int32_t t1 = clip(t1a + t5a, mn, mx);
    int32_t t2 = clip(t2a + t6a, mn, mx);
    int32_t t3 = clip(t3a + t7a, mn, mx);
    int32_t t4 = clip(t0a - t4a, mn, mx);
    int32_t t5 = clip(t1a - t5a, mn, mx);
    int32_t t6 = clip(t2a - t6a, mn, mx);
    int32_t t7 = clip(t3a - t7a, mn, mx);
    t4a = (((3784 - 4096) * t4 + 1567 * t5 + 2048) >> 12) + t4;
    t5a = ((1567 * t4 - (3784 - 4096) * t5 + 2048) >> 12) - t5;
    t6a = (((3784 - 4096) * t7 - 1567 * t6 + 2048) >> 12) + t7;
    t7a = ((1567 * t7 + (3784 - 4096) * t6 + 2048) >> 12) + t6;
    c[0] = clip(t0 + t2, mn, mx);
    c[7 * s] = -clip(t1 + t3, mn, mx);
    t2 = clip(t0 - t2, mn, mx);
    t3 = clip(t1 - t3, mn, mx);
    c[s] = -clip(t4a + t6a, mn, mx);
    c[6 * s] = clip(t5a + t7a, mn, mx);
    t6 = clip(t4a - t6a, mn, mx);
    t7 = clip(t5a - t7a, mn, mx);
    c[3 * s] = -(((t2 + t3) * 181 + 128) >> 8);
    c[4 * s] = ((t2 - t3) * 181 + 128) >> 8;
    c[2 * s] = ((t6 + t7) * 181 + 128) >> 8;
    c[5 * s] = -(((t6 - t7) * 181 + 128) >> 8);
}

// inv_adst16_1d (src/idct.rs:1968)
void inv_adst16(int32_t* c, int s, int mn, int mx) {
    int32_t in[16];
    for (int i = 0; i < 16; i++) in[i] = c[i * s];
    auto C = [&](int32_t v) { return clip(v, mn, mx); };
    int32_t t0 = ((in[15] * (4091 - 4096) + in[0] * 201 + 2048) >> 12) + in[15];
    int32_t t1 = ((in[15] * 201 - in[0] * (4091 - 4096) + 2048) >> 12) - in[0];
    int32_t t2 = ((in[13] * (3973 - 4096) + in[2] * 995 + 2048) >> 12) + in[13];
    int32_t t3 = ((in[13] * 995 - in[2] * (3973 - 4096) + 2048) >> 12) - in[2];
    int32_t t4 = ((in[11] * (3703 - 4096) + in[4] * 1751 + 2048) >> 12) + in[11];
    int32_t t5 = ((in[11] * 1751 - in[4] * (3703 - 4096) + 2048) >> 12) - in[4];
    int32_t t6 = (in[9] * 1645 + in[6] * 1220 + 1024) >> 11;
    int32_t t7 = (in[9] * 1220 - in[6] * 1645 + 1024) >> 11;
    int32_t t8 = ((in[7] * 2751 + in[8] * (3035 - 4096) + 2048) >> 12) + in[8];
    int32_t t9 = ((in[7] * (3035 - 4096) - in[8] * 2751 + 2048) >> 12) + in[7];
    int32_t t10 = ((in[5] * 2106 + in[10] * (3513 - 4096) + 2048) >> 12) + in[10];
    int32_t t11 = ((in[5] * (3513 - 4096) - in[10] * 2106 + 2048) >> 12) + in[5];
    int32_t t12 = ((in[3] * 1380 + in[12] * (3857 - 4096) + 2048) >> 12) + in[12];
    int32_t t13 = ((in[3] * (3857 - 4096) - in[12] * 1380 + 2048) >> 12) + in[3];
    int32_t t14 = ((in[1] * 601 + in[14] * (4052 - 4096) + 2048) >> 12) + in[14];
    int32_t t15 = ((in[1] * (4052 - 4096) - in[14] * 601 + 2048) >> 12) + in[1];
    int32_t t0a = C(t0 + t8), t1a = C(t1 + t9), t2a = C(t2 + t10), t3a = C(t3 + t11);
    int32_t t4a = C(t4 + t12), t5a = C(t5 + t13), t6a = C(t6 + t14), t7a = C(t7 + t15);
    int32_t t8a = C(t0 - t8), t9a = C(t1 - t9), t10a = C(t2 - t10), t11a = C(t3 - t11);
    int32_t t12a = C(t4 - t12), t13a = C(t5 - t13), t14a = C(t6 - t14), t15a = C(t7 - t15);
    t8 = ((t8a * (4017 - 4096) + t9a * 799 + 2048) >> 12) + t8a;
    t9 = ((t8a * 799 - t9a * (4017 - 4096) + 2048) >> 12) - t9a;
    t10 = ((t10a * 2276 + t11a * (3406 - 4096) + 2048) >> 12) + t11a;
    t11 = ((t10a * (3406 - 4096) - t11a * 2276 + 2048) >> 12) + t10a;
    t12 = ((t13a * (4017 - 4096) - t12a * 799 + 2048) >> 12) + t13a;
    t13 = ((t13a * 799 + t12a * (4017 - 4096) + 2048) >> 12) + t12a;
    t14 = ((t15a * 2276 - t14a * (3406 - 4096) + 2048) >> 12) - t14a;
    t15 = ((t15a * (3406 - 4096) + t14a * 2276 + 2048) >> 12) + t15a;
    t0 = C(t0a + t4a); t1 = C(t1a + t5a); t2 = C(t2a + t6a); t3 = C(t3a + t7a);
    t4 = C(t0a - t4a); t5 = C(t1a - t5a); t6 = C(t2a - t6a); t7 = C(t3a - t7a);
    t8a = C(t8 + t12); t9a = C(t9 + t13); t10a = C(t10 + t14); t11a = C(t11 + t15);
    t12a = C(t8 - t12); t13a = C(t9 - t13); t14a = C(t10 - t14); t15a = C(t11 - t15);
    t4a = ((t4 * (3784 - 4096) + t5 * 1567 + 2048) >> 12) + t4;
    t5a = ((t4 * 1567 - t5 * (3784 - 4096) + 2048) >> 12) - t5;
    t6a = ((t7 * (3784 - 4096) - t6 * 1567 + 2048) >> 12) + t7;
    t7a = ((t7 * 1567 + t6 * (3784 - 4096) + 2048) >> 12) + t6;
    t12 = ((t12a * (3784 - 4096) + t13a * 1567 + 2048) >> 12) + t12a;
    t13 = ((t12a * 1567 - t13a * (3784 - 4096) + 2048) >> 12) - t13a;
    t14 = ((t15a * (3784 - 4096) - t14a * 1567 + 2048) >> 12) + t15a;
    t15 = ((t15a * 1567 + t14a * (3784 - 4096) + 2048) >> 12) + t14a;
    c[0] = C(t0 + t2);
    c[15 * s] = -C(t1 + t3);
    t2a = C(t0 - t2);
    t3a = C(t1 - t3);
    c[3 * s] = -C(t4a + t6a);
    c[12 * s] = C(t5a + t7a);
    t6 = C(t4a - t6a);
    t7 = C(t5a - t7a);
    c[s] = -C(t8a + t10a);
    c[14 * s] = C(t9a + t11a);
    t10 = C(t8a - t10a);
    t11 = C(t9a - t11a);
    c[2 * s] = C(t12 + t14);
    c[13 * s] = -C(t13 + t15);
    t14a = C(t12 - t14);
    t15a = C(t13 - t15);
    c[7 * s] = -(((t2a + t3a) * 181 + 128) >> 8);
    c[8 * s] = ((t2a - t3a) * 181 + 128) >> 8;
    c[4 * s] = ((t6 + t7) * 181 + 128) >> 8;
    c[11 * s] = -(((t6 - t7) * 181 + 128) >> 8);
    c[6 * s] = ((t10 + t11) * 181 + 128) >> 8;
    c[9 * s] = -(((t10 - t11) * 181 + 128) >> 8);
    c[5 * s] = -(((t14a + t15a) * 181 + 128) >> 8);
    c[10 * s] = ((t14a - t15a) * 181 + 128) >> 8;
}

// Identity 1-D passes (src/idct.rs:1094, :2733, :1154): no internal clamp.
inline int32_t id4(int32_t v) { return v + ((v * 1697 + 2048) >> 12); }
inline int32_t id8(int32_t v) { return v * 2; }
inline int32_t id16(int32_t v) { return 2 * v + ((v * 1697 + 1024) >> 11); }

enum Kind { K_DCT, K_ADST, K_ID };

void pass_1d(Kind k, int n, int32_t* c, int s, int mn, int mx) {
    if (k == K_ID) {
        for (int i = 0; i < n; i++) {
            int32_t v = c[i * s];
            c[i * s] = n == 4 ? id4(v) : n == 8 ? id8(v) : id16(v);
        }
        return;
    }
    if (k == K_ADST) {
        if (n == 4) inv_adst4(c, s, mn, mx);
        else if (n == 8) inv_adst8(c, s, mn, mx);
        else inv_adst16(c, s, mn, mx);
        return;
    }
    switch (n) {
        case 4: inv_dct4(c, s, mn, mx); break;
        case 8: inv_dct8(c, s, mn, mx); break;
        case 16: inv_dct16(c, s, mn, mx); break;
        case 32: inv_dct32(c, s, mn, mx); break;
        case 64: inv_dct64(c, s, mn, mx); break;
    }
}

// Dequantize one level (src/idct.rs:1576-1590 and every sibling): the 24-bit mask, the
// dq_shift of the 32/64-class sizes, then the cf_max clamp.
inline int32_t dequant(int32_t lvl, int32_t step, int dq_shift, int32_t cf_max) {
    if (lvl == 0) return 0;
    uint32_t a = lvl < 0 ? (uint32_t)0 - (uint32_t)lvl : (uint32_t)lvl;
    int32_t mag = (int32_t)((((uint64_t)a * (uint64_t)(uint32_t)step) & 0xffffff) >> dq_shift);
    mag = std::min(mag, cf_max + (lvl < 0 ? 1 : 0));
    return lvl < 0 ? -mag : mag;
}

// The common 2-D inverse of every reference size (src/idct.rs:31, :858-1430, :1572-1862,
// :1926-2129, :2281, :2719-3136): dequant, optional rect2 prescale while transposing into
// row-major, row pass with the row clip, the per-size inter-pass shift and column clip,
// column pass with the column clip, then (t + 8) >> 4.
void inv2d(const int32_t* levels, const orc_quant* q, int W, int H, Kind rowk, Kind colk,
           bool use_qm, int dq_shift, bool rect2, int mid_shift, int32_t* out) {
    int32_t coeff[1024];
    for (int rc = 0; rc < W * H; rc++) {
        int32_t step;
        if (use_qm) step = orc_dequant_step(q, rc, W, H);
        else step = rc == 0 ? q->dc : q->ac;
        coeff[rc] = dequant(levels[rc], step, dq_shift, q->cf_max);
    }
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int32_t v = coeff[y + x * H];
            out[y * W + x] = rect2 ? (v * 181 + 128) >> 8 : v;
        }
    for (int y = 0; y < H; y++) pass_1d(rowk, W, out + y * W, 1, q->rmin, q->rmax);
    for (int i = 0; i < W * H; i++) {
        int32_t t = out[i];
        if (mid_shift) t = (t + (1 << (mid_shift - 1))) >> mid_shift;
        out[i] = clip(t, q->cmin, q->cmax);
    }
    for (int x = 0; x < W; x++) pass_1d(colk, H, out + x, W, q->cmin, q->cmax);
    for (int i = 0; i < W * H; i++) out[i] = (out[i] + 8) >> 4;
}

// idct_dequant_64x64 (src/idct.rs:2677): 32x32 coded levels, 32x32 QM, dq_shift 2.
void inv64(const int32_t* levels, const orc_quant* q, int32_t* out) {
    std::fill(out, out + 4096, 0);
    for (int y = 0; y < 32; y++) {
        int32_t* row = out + y * 64;
        bool any = false;
        for (int x = 0; x < 32; x++) {
            int32_t lvl = levels[x * 32 + y];
            if (lvl == 0) continue;
            any = true;
            row[x] = dequant(lvl, orc_dequant_step(q, x * 32 + y, 32, 32), 2, q->cf_max);
        }
        if (any) inv_dct64(row, 1, q->rmin, q->rmax);
    }
    for (int i = 0; i < 32 * 64; i++) out[i] = clip((out[i] + 2) >> 2, q->cmin, q->cmax);
    for (int x = 0; x < 64; x++) inv_dct64(out + x, 64, q->cmin, q->cmax);
    for (int i = 0; i < 4096; i++) out[i] = (out[i] + 8) >> 4;
}

}  // namespace

extern "C" int orc_inv_txfm(int s, int t, const int32_t* levels, const orc_quant* q, int32_t* out) {
    if (!orc_tx_supported(s, t)) return -1;
    const int W = orc_tx_w(s), H = orc_tx_h(s);
    if (s == TX_64X64) {
        inv64(levels, q, out);
        return 0;
    }
    // Row kind (first pass, horizontal) and column kind (second pass, vertical).  ADST_DCT
    // is vertical ADST, so its column pass is the ADST (src/idct.rs:1683, :2893).
    Kind rowk = K_DCT, colk = K_DCT;
    switch (t) {
        case ADST_ADST: rowk = K_ADST; colk = K_ADST; break;
        case ADST_DCT: rowk = K_DCT; colk = K_ADST; break;
        case DCT_ADST: rowk = K_ADST; colk = K_DCT; break;
        case V_DCT: rowk = K_ID; colk = K_DCT; break;
        case H_DCT: rowk = K_DCT; colk = K_ID; break;
        case IDTX: rowk = K_ID; colk = K_ID; break;
        default: break;
    }
    // The identity family dequantizes flat (FlatDct, src/idct.rs:1614).
    const bool use_qm = !(t == IDTX || t == V_DCT || t == H_DCT);
    int dq_shift = 0, mid = 0;
    bool rect2 = false;
    switch (s) {
        case TX_4X4: mid = 0; break;                         // src/idct.rs:2808
        case TX_4X8: case TX_8X4: rect2 = true; mid = 0; break;  // :1394, :1299, :1338
        case TX_4X16: case TX_16X4: mid = 1; break;          // :934, :899
        case TX_8X8: mid = 1; break;                         // :1572
        case TX_8X16: case TX_16X8: rect2 = true; mid = 1; break;  // :1260, :968, :1098
        case TX_16X16: mid = 2; break;                       // :1926, :1826
        case TX_16X32: case TX_32X16: rect2 = true; mid = 1; dq_shift = 1; break;  // :858, :31
        case TX_32X32: mid = 2; dq_shift = 1; break;         // :2281
        default: return -1;
    }
    inv2d(levels, q, W, H, rowk, colk, use_qm, dq_shift, rect2, mid, out);
    return 0;
}

extern "C" void orc_inv_dct1d(int n, int32_t* c, int mn, int mx) { pass_1d(K_DCT, n, c, 1, mn, mx); }
extern "C" void orc_inv_adst1d(int n, int32_t* c, int mn, int mx) { pass_1d(K_ADST, n, c, 1, mn, mx); }

extern "C" int64_t orc_sse_recon(const int32_t* pred, const int32_t* resid, const uint16_t* src,
                                 int stride, int w, int h, int maxv) {
    // sse_recon_scalar (src/rd_sse.rs:1262)
    int64_t sse = 0;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int32_t r = clip(pred[y * w + x] + resid[y * w + x], 0, maxv);
            int64_t d = (int64_t)((int32_t)src[y * stride + x] - r);
            sse += d * d;
        }
    return sse;
}

namespace {
// The library's block checksum: sum of fold(value * odd 24-bit position weight) with the
// low 32 bits of 24 x 24-bit products.
inline uint32_t u24mul(uint32_t a, uint32_t b) { return (uint32_t)((uint64_t)(a & 0xFFFFFFu) * (b & 0xFFFFFFu)); }
inline uint32_t hw24(uint32_t p) { return (u24mul(p, 0x9E3779u) & 0xFFFFFFu) | 1u; }
inline uint32_t fold13(uint32_t t) { return t ^ (t >> 13); }
}  // namespace

extern "C" uint32_t orc_block_hash(int ncoef, const int32_t* levels, const float* targets, int npx,
                                   const int32_t* recon) {
    uint32_t h = 0;
    for (int i = 0; i < ncoef; i++) {
        uint32_t tb;
        memcpy(&tb, &targets[i], 4);
        h += fold13(u24mul((uint32_t)levels[i], hw24((uint32_t)i)));
        h += fold13(u24mul(tb, hw24((uint32_t)i + 4096u)) ^ (tb >> 24));
    }
    for (int j = 0; j < npx; j++) h += fold13(u24mul((uint32_t)recon[j], hw24((uint32_t)j + 8192u)));
    return h;
}

extern "C" int64_t orc_chain_block(int s, int t, const int16_t* resid, const orc_quant* q,
                                   int32_t* levels_out, float* targets_out, int32_t* recon_out) {
    if (!orc_tx_supported(s, t)) return -1;
    const int W = orc_tx_w(s), H = orc_tx_h(s), N = W * H, NC = orc_tx_ncoef(s);
    int32_t in[4096], lv[1024], rec[4096], pred[4096];
    float tg[1024];
    uint16_t src[4096];
    const int P = 1 << (q->bd - 1), maxv = (1 << q->bd) - 1;
    for (int i = 0; i < N; i++) {
        in[i] = resid[i];
        pred[i] = P;
        src[i] = (uint16_t)clip(P + resid[i], 0, maxv);
    }
    orc_fwd_txfm(s, t, in, q, lv, tg);
    orc_inv_txfm(s, t, lv, q, rec);
    if (levels_out) memcpy(levels_out, lv, sizeof(int32_t) * NC);
    if (targets_out) memcpy(targets_out, tg, sizeof(float) * NC);
    if (recon_out) memcpy(recon_out, rec, sizeof(int32_t) * N);
    return orc_sse_recon(pred, rec, src, W, W, H, maxv);
}

extern "C" int orc_chain_batch(int n, const uint8_t* tx_size, const uint8_t* tx_type,
                               const uint8_t* qidx, const uint8_t* bd, const uint8_t* qm_level,
                               const uint64_t* offset, const int16_t* pool, int64_t* sse_out,
                               uint32_t* hash_out, int nthreads) {
    auto work = [&](int begin, int end) {
        std::vector<int32_t> lv(1024), rec(4096);
        std::vector<float> tg(1024);
        for (int i = begin; i < end; i++) {
            orc_quant q;
            // qm_level bit 6: the chroma quantizer with zero qindex deltas
            // (Quant::new_chroma_with_delta_qm, src/quant.rs:394)
            if (qm_level[i] & 0x40)
                orc_quant_new_chroma(qidx[i], 0, 0, bd[i], qm_level[i] & 15, &q);
            else
                orc_quant_new(qidx[i], bd[i], qm_level[i] & 15, &q);
            int64_t sse = orc_chain_block(tx_size[i], tx_type[i], pool + offset[i], &q, lv.data(),
                                          tg.data(), rec.data());
            sse_out[i] = sse;
            if (hash_out) {
                int s = tx_size[i];
                hash_out[i] = sse < 0 ? 0u
                                      : orc_block_hash(orc_tx_ncoef(s), lv.data(), tg.data(),
                                                       orc_tx_w(s) * orc_tx_h(s), rec.data());
            }
        }
    };
    if (nthreads <= 1) {
        work(0, n);
        return 0;
    }
    std::vector<std::thread> th;
    int chunk = (n + nthreads - 1) / nthreads;
    for (int k = 0; k < nthreads; k++) {
        int b = k * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back(work, b, e);
    }
    for (auto& x : th) x.join();
    return 0;
}

// The chain with the trellis (src/coder/partition_search.rs:509-513 order: forward +
// quantization, trellis_optimize[_ctx], dequantization + inverse, SSE).
extern "C" int orc_chain_batch_trellis(int n, const uint8_t* tx_size, const uint8_t* tx_type,
                                       const uint8_t* qidx, const uint8_t* bd, const uint8_t* qm_level,
                                       const uint64_t* offset, const int16_t* pool, const orc_trellisblk* tr,
                                       int64_t* sse_out, uint32_t* hash_out, int nthreads) {
    auto work = [&](int begin, int end) {
        int32_t in[4096], lv[1024], rec[4096], pred[4096];
        float tg[1024];
        uint16_t src[4096];
        for (int i = begin; i < end; i++) {
            orc_quant q;
            if (qm_level[i] & 0x40)
                orc_quant_new_chroma(qidx[i], 0, 0, bd[i], qm_level[i] & 15, &q);
            else
                orc_quant_new(qidx[i], bd[i], qm_level[i] & 15, &q);
            const int s = tx_size[i], t = tx_type[i];
            if (!orc_tx_supported(s, t)) {
                sse_out[i] = -1;
                if (hash_out) hash_out[i] = 0;
                continue;
            }
            const int W = orc_tx_w(s), H = orc_tx_h(s), N = W * H, NC = orc_tx_ncoef(s);
            const int P = 1 << (q.bd - 1), maxv = (1 << q.bd) - 1;
            const int16_t* r = pool + offset[i];
            for (int k = 0; k < N; k++) {
                in[k] = r[k];
                pred[k] = P;
                src[k] = (uint16_t)clip(P + r[k], 0, maxv);
            }
            orc_fwd_txfm(s, t, in, &q, lv, tg);
            const orc_trellisblk& d = tr[i];
            const int cw = W < 32 ? W : 32, ch = H < 32 ? H : 32;
            if (d.flags & 1) {
                orc_rate_ctx c{};
                c.qctx = d.qctx;
                c.cls = ((31 - __builtin_clz(W)) - 2 + (31 - __builtin_clz(H)) - 2 + 1) >> 1;
                c.plane = d.plane;
                c.w = cw;
                c.h = ch;
                c.eob_bin = d.eob_bin;
                c.dcs_ctx = d.dcs_ctx;
                orc_trellis_ctx(lv, tg, (float)q.dc, (float)q.ac, orc_scan(s), d.lambda0, &c, q.qm_level,
                                d.band_tilt);
            } else {
                orc_trellis(lv, tg, (float)q.dc, (float)q.ac, orc_scan(s), NC, d.lambda0);
            }
            orc_inv_txfm(s, t, lv, &q, rec);
            sse_out[i] = orc_sse_recon(pred, rec, src, W, W, H, maxv);
            if (hash_out) hash_out[i] = orc_block_hash(NC, lv, tg, N, rec);
        }
    };
    if (nthreads <= 1) {
        work(0, n);
        return 0;
    }
    std::vector<std::thread> th;
    const int chunk = (n + nthreads - 1) / nthreads;
    for (int k = 0; k < nthreads; k++) {
        const int b = k * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back(work, b, e);
    }
    for (auto& x : th) x.join();
    return 0;
}
