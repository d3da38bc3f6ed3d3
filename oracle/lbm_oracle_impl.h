/*
 * Type-generic body of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 * Included twice by lbm_oracle.c with T = float / S = f and T = double / S = d.
 *
 * Every expression below keeps the operand order and parenthesisation of the
 * reference so that, compiled without FMA contraction, each floating-point
 * operation rounds exactly as the reference's does.
 */

#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, S)

/* Equilibrium macros, common.h:132-160 (same operand order). */
static inline T FN(eq_a0)(T vela, T vela2, T rho_alpha)
{
    return ((T)1 / (T)18) * ((rho_alpha) + ((T)3) * (vela) + ((T)9 / (T)2) * (vela2));
}
static inline T FN(eq_a1)(T vela, T vela2, T rho_alpha)
{
    return ((T)1 / (T)18) * ((rho_alpha) + ((T)-3) * (vela) + ((T)9 / (T)2) * (vela2));
}
static inline T FN(eq_4)(T vab, T vab2, T rho_alpha)
{
    return ((T)1 / (T)36) * ((rho_alpha) + ((T)3) * (vab) + ((T)9 / (T)2) * (vab2));
}
static inline T FN(eq_5)(T vab, T vab2, T rho_alpha)
{
    return ((T)1 / (T)36) * ((rho_alpha) + ((T)-3) * (vab) + ((T)9 / (T)2) * (vab2));
}
static inline T FN(eq_18)(T rho_alpha)
{
    return ((T)1 / (T)3) * (rho_alpha);
}

/* Moment accumulation order shared by CLbmAlphaCPU.cpp:87-182 and
 * CLbmBetaCPU.cpp:93-217: rho in index order, j without division by rho. */
static inline void FN(moments)(const T d[19], T* rho, T* ux, T* uy, T* uz)
{
    T r, x, y, z;
    r = d[0];  x = d[0];
    r += d[1]; x -= d[1];
    r += d[2]; y = d[2];
    r += d[3]; y -= d[3];
    r += d[4]; x += d[4]; y += d[4];
    r += d[5]; x -= d[5]; y -= d[5];
    r += d[6]; x += d[6]; y -= d[6];
    r += d[7]; x -= d[7]; y += d[7];
    r += d[8]; x += d[8]; z = d[8];
    r += d[9]; x -= d[9]; z -= d[9];
    r += d[10]; x += d[10]; z -= d[10];
    r += d[11]; x -= d[11]; z += d[11];
    r += d[12]; y += d[12]; z += d[12];
    r += d[13]; y -= d[13]; z -= d[13];
    r += d[14]; y += d[14]; z -= d[14];
    r += d[15]; y -= d[15]; z += d[15];
    r += d[16]; z += d[16];
    r += d[17]; z -= d[17];
    r += d[18];
    *rho = r; *ux = x; *uy = y; *uz = z;
}

/*
 * One cell of the alpha step, CLbmAlphaCPU.cpp:75-494.
 * Reads the 19 own slots, collides, and stores the post-collision value of
 * direction l into slot opp(l) of the same cell. The variable `tmp` of the
 * reference is a #define for `rho` (CLbmAlphaCPU.cpp:199), so each gravity
 * term overwrites rho; the stored density is the last gravity term.
 */
static void FN(alpha_cell)(T* dd, long n, long id, int32_t flag, T* rho_out, T* u_out,
                           T inv_tau, T lid, const T g[3], int store_rho, int store_u)
{
    T d[19];
    T rho, ux, uy, uz, vel2, vela2, vab, dd_param;
    int l;

    if (flag == ORACLE_GHOST_LAYER)
        return;
    for (l = 0; l < 19; l++)
        d[l] = dd[l * n + id];
    FN(moments)(d, &rho, &ux, &uy, &uz);

    switch (flag) {
    case ORACLE_FLUID:
        vel2 = ux * ux + uy * uy + uz * uz;
        dd_param = rho - ((T)3 / (T)2) * vel2;

        rho = g[0] * ((T)1 / (T)18) * rho;
        vela2 = ux * ux;
        d[1] += inv_tau * (FN(eq_a1)(ux, vela2, dd_param) - d[1]);
        d[1] -= rho;
        dd[0 * n + id] = d[1];
        d[0] += inv_tau * (FN(eq_a0)(ux, vela2, dd_param) - d[0]);
        d[0] += rho;
        dd[1 * n + id] = d[0];

        rho = g[1] * ((T)-1 / (T)18) * rho;
        vela2 = uy * uy;
        d[3] += inv_tau * (FN(eq_a1)(uy, vela2, dd_param) - d[3]);
        d[3] -= rho;
        dd[2 * n + id] = d[3];
        d[2] += inv_tau * (FN(eq_a0)(uy, vela2, dd_param) - d[2]);
        d[2] += rho;
        dd[3 * n + id] = d[2];

        vab = ux + uy;
        vela2 = vab * vab;
        rho = (g[0] - g[1]) * ((T)1 / (T)36) * rho;
        d[5] += inv_tau * (FN(eq_5)(vab, vela2, dd_param) - d[5]);
        d[5] -= rho;
        dd[4 * n + id] = d[5];
        d[4] += inv_tau * (FN(eq_4)(vab, vela2, dd_param) - d[4]);
        d[4] += rho;
        dd[5 * n + id] = d[4];

        vab = ux - uy;
        vela2 = vab * vab;
        rho = (g[0] + g[1]) * ((T)1 / (T)36) * rho;
        d[7] += inv_tau * (FN(eq_5)(vab, vela2, dd_param) - d[7]);
        d[7] -= rho;
        dd[6 * n + id] = d[7];
        d[6] += inv_tau * (FN(eq_4)(vab, vela2, dd_param) - d[6]);
        d[6] += rho;
        dd[7 * n + id] = d[6];

        vab = ux + uz;
        vela2 = vab * vab;
        rho = (g[0] + g[2]) * ((T)1 / (T)36) * rho;
        d[9] += inv_tau * (FN(eq_5)(vab, vela2, dd_param) - d[9]);
        d[9] -= rho;
        dd[8 * n + id] = d[9];
        d[8] += inv_tau * (FN(eq_4)(vab, vela2, dd_param) - d[8]);
        d[8] += rho;
        dd[9 * n + id] = d[8];

        rho = (g[0] - g[2]) * ((T)1 / (T)36) * rho;
        vab = ux - uz;
        vela2 = vab * vab;
        d[11] += inv_tau * (FN(eq_5)(vab, vela2, dd_param) - d[11]);
        d[11] -= rho;
        dd[10 * n + id] = d[11];
        d[10] += inv_tau * (FN(eq_4)(vab, vela2, dd_param) - d[10]);
        d[10] += rho;
        dd[11 * n + id] = d[10];

        vab = uy + uz;
        vela2 = vab * vab;
        rho = (g[2] - g[1]) * ((T)1 / (T)36) * rho;
        d[13] += inv_tau * (FN(eq_5)(vab, vela2, dd_param) - d[13]);
        d[13] -= rho;
        dd[12 * n + id] = d[13];
        d[12] += inv_tau * (FN(eq_4)(vab, vela2, dd_param) - d[12]);
        d[12] += rho;
        dd[13 * n + id] = d[12];

        vab = uy - uz;
        vela2 = vab * vab;
        rho = (g[2] + g[1]) * ((T)-1 / (T)36) * rho;
        d[15] += inv_tau * (FN(eq_5)(vab, vela2, dd_param) - d[15]);
        d[15] -= rho;
        dd[14 * n + id] = d[15];
        d[14] += inv_tau * (FN(eq_4)(vab, vela2, dd_param) - d[14]);
        d[14] += rho;
        dd[15 * n + id] = d[14];

        vela2 = uz * uz;
        rho = g[2] * ((T)1 / (T)18) * rho;
        d[17] += inv_tau * (FN(eq_a1)(uz, vela2, dd_param) - d[17]);
        d[17] -= rho;
        dd[16 * n + id] = d[17];
        d[16] += inv_tau * (FN(eq_a0)(uz, vela2, dd_param) - d[16]);
        d[16] += rho;
        dd[17 * n + id] = d[16];

        d[18] += inv_tau * (FN(eq_18)(dd_param) - d[18]);
        dd[18 * n + id] = d[18];
        break;

    case ORACLE_OBSTACLE:
        /* no stores: bounce-back is implicit in the A-A layout (:330-343) */
        if (store_u) {
            ux = (T)0;
            uy = (T)0;
            uz = (T)0;
        }
        break;

    case ORACLE_VELOCITY_INJECTION:
        ux = lid;
        uy = (T)0;
        uz = (T)0;
        rho = (T)1;
        vel2 = ux * ux + uy * uy + uz * uz;
        dd_param = rho - ((T)3 / (T)2) * (vel2);

        vela2 = ux * ux;
        rho = g[0] * ((T)1 / (T)18) * rho;
        d[1] = FN(eq_a1)(ux, vela2, dd_param);
        d[1] -= rho;
        dd[0 * n + id] = d[1];
        d[0] = FN(eq_a0)(ux, vela2, dd_param);
        d[0] += rho;
        dd[1 * n + id] = d[0];

        vela2 = uy * uy;
        rho = g[1] * ((T)-1 / (T)18) * rho;
        d[3] = FN(eq_a1)(uy, vela2, dd_param);
        d[3] -= rho;
        dd[2 * n + id] = d[3];
        d[2] = FN(eq_a0)(uy, vela2, dd_param);
        d[2] += rho;
        dd[3 * n + id] = d[2];

        vab = ux + uy;
        vela2 = vab * vab;
        rho = (g[0] - g[1]) * ((T)1 / (T)36) * rho;
        d[5] = FN(eq_5)(vab, vela2, dd_param);
        d[5] -= rho;
        dd[4 * n + id] = d[5];
        d[4] = FN(eq_4)(vab, vela2, dd_param);
        d[4] += rho;
        dd[5 * n + id] = d[4];

        vab = ux - uy;
        vela2 = vab * vab;
        rho = (g[0] + g[1]) * ((T)1 / (T)36) * rho;
        d[7] = FN(eq_5)(vab, vela2, dd_param);
        d[7] -= rho;
        dd[6 * n + id] = d[7];
        d[6] = FN(eq_4)(vab, vela2, dd_param);
        d[6] += rho;
        dd[7 * n + id] = d[6];

        vab = ux + uz;
        vela2 = vab * vab;
        rho = (g[0] + g[2]) * ((T)1 / (T)36) * rho;
        d[9] = FN(eq_5)(vab, vela2, dd_param);
        d[9] -= rho;
        dd[8 * n + id] = d[9];
        d[8] = FN(eq_4)(vab, vela2, dd_param);
        d[8] += rho;
        dd[9 * n + id] = d[8];

        vab = ux - uz;
        vela2 = vab * vab;
        rho = (g[0] - g[2]) * ((T)1 / (T)36) * rho;
        d[11] = FN(eq_5)(vab, vela2, dd_param);
        d[11] -= rho;
        dd[10 * n + id] = d[11];
        d[10] = FN(eq_4)(vab, vela2, dd_param);
        d[10] += rho;
        dd[11 * n + id] = d[10];

        vab = uy + uz;
        vela2 = vab * vab;
        rho = (g[2] - g[1]) * ((T)1 / (T)36) * rho;
        d[13] = FN(eq_5)(vab, vela2, dd_param);
        d[13] -= rho;
        dd[12 * n + id] = d[13];
        d[12] = FN(eq_4)(vab, vela2, dd_param);
        d[12] += rho;
        dd[13 * n + id] = d[12];

        vab = uy - uz;
        vela2 = vab * vab;
        rho = (g[2] + g[1]) * ((T)-1 / (T)36) * rho;
        d[15] = FN(eq_5)(vab, vela2, dd_param);
        d[15] -= rho;
        dd[14 * n + id] = d[15];
        d[14] = FN(eq_4)(vab, vela2, dd_param);
        d[14] += rho;
        dd[15 * n + id] = d[14];

        vela2 = uz * uz;
        rho = g[2] * ((T)1 / (T)18) * rho;
        d[17] = FN(eq_a1)(uz, vela2, dd_param);
        d[17] -= rho;
        dd[16 * n + id] = d[17];
        d[16] = FN(eq_a0)(uz, vela2, dd_param);
        d[16] += rho;
        dd[17 * n + id] = d[16];

        d[18] = FN(eq_18)(dd_param);
        dd[18 * n + id] = d[18];
        break;

    default:
        break;
    }

    if (store_u) {
        u_out[0 * n + id] = ux;
        u_out[1 * n + id] = uy;
        u_out[2 * n + id] = uz;
    }
    if (store_rho)
        rho_out[id] = rho;
}

/*
 * One cell of the beta step, CLbmBetaCPU.cpp:83-451.
 * Gathers direction l from slot opp(l) of neighbour id - c_l, collides, and
 * scatters direction l to slot l of neighbour id + c_l. Neighbour indices wrap
 * modulo the cell count with halo (domainWrap, CLbmBetaCPU.cpp:44-50).
 * `dd_param` is a #define for `rho` (CLbmBetaCPU.cpp:224), so the stored
 * FLUID density is rho - 1.5 |j|^2. GHOST cells are not skipped: they pass
 * values through unchanged.
 */
static void FN(beta_cell)(T* dd, long n, long id, const long delta[19], int32_t flag,
                          T* rho_out, T* u_out, T inv_tau, T lid, int store_rho,
                          int store_u)
{
    /* opp[l]: slot holding direction l at the neighbour, the reverse direction */
    static const int opp[19] = {1, 0, 3, 2, 5, 4, 7, 6, 9, 8, 11, 10, 13, 12, 15, 14, 17, 16, 18};
    T d[19];
    T rho, ux, uy, uz, vel2, vela2, swap;
    long pos[19];
    int l;

    for (l = 0; l < 19; l++) {
        long j = id + delta[l];
        if (j >= n)
            j -= n;
        if (j < 0)
            j += n;
        pos[l] = j; /* slot l of neighbour id + c_l */
    }
    /* direction l arrives from id - c_l, stored there in slot opp(l); that is
     * plane opp(l) at position pos[opp(l)] */
    for (l = 0; l < 19; l++)
        d[l] = dd[(long)opp[l] * n + pos[opp[l]]];

    FN(moments)(d, &rho, &ux, &uy, &uz);

    switch (flag) {
    case ORACLE_FLUID:
        vel2 = ux * ux + uy * uy + uz * uz;
        rho = rho - ((T)3 / (T)2) * (vel2);

        vela2 = ux * ux;
        d[0] += inv_tau * (FN(eq_a0)(ux, vela2, rho) - d[0]);
        d[1] += inv_tau * (FN(eq_a1)(ux, vela2, rho) - d[1]);

        vela2 = uy * uy;
        d[2] += inv_tau * (FN(eq_a0)(uy, vela2, rho) - d[2]);
        d[3] += inv_tau * (FN(eq_a1)(uy, vela2, rho) - d[3]);

        vel2 = ux + uy;
        vela2 = vel2 * vel2;
        d[4] += inv_tau * (FN(eq_4)(vel2, vela2, rho) - d[4]);
        d[5] += inv_tau * (FN(eq_5)(vel2, vela2, rho) - d[5]);

        vel2 = ux - uy;
        vela2 = vel2 * vel2;
        d[6] += inv_tau * (FN(eq_4)(vel2, vela2, rho) - d[6]);
        d[7] += inv_tau * (FN(eq_5)(vel2, vela2, rho) - d[7]);

        vel2 = ux + uz;
        vela2 = vel2 * vel2;
        d[8] += inv_tau * (FN(eq_4)(vel2, vela2, rho) - d[8]);
        d[9] += inv_tau * (FN(eq_5)(vel2, vela2, rho) - d[9]);

        vel2 = ux - uz;
        vela2 = vel2 * vel2;
        d[10] += inv_tau * (FN(eq_4)(vel2, vela2, rho) - d[10]);
        d[11] += inv_tau * (FN(eq_5)(vel2, vela2, rho) - d[11]);

        vel2 = uy + uz;
        vela2 = vel2 * vel2;
        d[12] += inv_tau * (FN(eq_4)(vel2, vela2, rho) - d[12]);
        d[13] += inv_tau * (FN(eq_5)(vel2, vela2, rho) - d[13]);

        vel2 = uy - uz;
        vela2 = vel2 * vel2;
        d[14] += inv_tau * (FN(eq_4)(vel2, vela2, rho) - d[14]);
        d[15] += inv_tau * (FN(eq_5)(vel2, vela2, rho) - d[15]);

        vela2 = uz * uz;
        d[16] += inv_tau * (FN(eq_a0)(uz, vela2, rho) - d[16]);
        d[17] += inv_tau * (FN(eq_a1)(uz, vela2, rho) - d[17]);

        d[18] += inv_tau * (FN(eq_18)(rho) - d[18]);
        break;

    case ORACLE_OBSTACLE:
        if (store_u) {
            ux = (T)0;
            uy = (T)0;
            uz = (T)0;
        }
        /* simple bounce-back: swap opposite pairs (:302-311) */
        for (l = 0; l < 18; l += 2) {
            swap = d[l + 1];
            d[l + 1] = d[l];
            d[l] = swap;
        }
        break;

    case ORACLE_VELOCITY_INJECTION:
        ux = lid;
        uy = 0;
        uz = 0;
        rho = (T)1;
        vel2 = ux * ux + uy * uy + uz * uz;
        rho = rho - ((T)3 / (T)2) * vel2;

        vela2 = ux * ux;
        d[0] = FN(eq_a0)(ux, vela2, rho);
        d[1] = FN(eq_a1)(ux, vela2, rho);
        vela2 = uy * uy;
        d[2] = FN(eq_a0)(uy, vela2, rho);
        d[3] = FN(eq_a1)(uy, vela2, rho);

        vel2 = ux + uy;
        vela2 = vel2 * vel2;
        d[4] = FN(eq_4)(vel2, vela2, rho);
        d[5] = FN(eq_5)(vel2, vela2, rho);
        vel2 = ux - uy;
        vela2 = vel2 * vel2;
        d[6] = FN(eq_4)(vel2, vela2, rho);
        d[7] = FN(eq_5)(vel2, vela2, rho);

        vel2 = ux + uz;
        vela2 = vel2 * vel2;
        d[8] = FN(eq_4)(vel2, vela2, rho);
        d[9] = FN(eq_5)(vel2, vela2, rho);
        vel2 = ux - uz;
        vela2 = vel2 * vel2;
        d[10] = FN(eq_4)(vel2, vela2, rho);
        d[11] = FN(eq_5)(vel2, vela2, rho);

        vel2 = uy + uz;
        vela2 = vel2 * vel2;
        d[12] = FN(eq_4)(vel2, vela2, rho);
        d[13] = FN(eq_5)(vel2, vela2, rho);
        vel2 = uy - uz;
        vela2 = vel2 * vel2;
        d[14] = FN(eq_4)(vel2, vela2, rho);
        d[15] = FN(eq_5)(vel2, vela2, rho);

        vela2 = uz * uz;
        d[16] = FN(eq_a0)(uz, vela2, rho);
        d[17] = FN(eq_a1)(uz, vela2, rho);
        d[18] = FN(eq_18)(rho);
        break;

    default: /* GHOST_LAYER: pass-through */
        break;
    }

    for (l = 0; l < 19; l++)
        dd[(long)l * n + pos[l]] = d[l];

    if (store_u) {
        u_out[0 * n + id] = ux;
        u_out[1 * n + id] = uy;
        u_out[2 * n + id] = uz;
    }
    if (store_rho)
        rho_out[id] = rho;
}

/* CLbmInitCPU::initLbm (CLbmInitCPU.cpp:77-171): rest equilibrium via the
 * equilibrium macros with rho = 1, u = 0. */
void FN(oracle_init)(T* dd, T* rho_out, T* u_out, const int size_h[3], int store_rho,
                     int store_u)
{
    long n = (long)size_h[0] * size_h[1] * size_h[2];
    long id;
    for (id = 0; id < n; id++) {
        T ux = (T)0, uy = (T)0, uz = (T)0, rho = (T)1;
        T vela2, dd_param, vab, vab2;

        vela2 = ux * ux;
        dd_param = rho - ((T)3 / (T)2) * (vela2);
        dd[id + 0 * n] = FN(eq_a0)(ux, vela2, dd_param);
        dd[id + 1 * n] = FN(eq_a1)(ux, vela2, dd_param);
        vela2 = uy * uy;
        dd[id + 2 * n] = FN(eq_a0)(uy, vela2, dd_param);
        dd[id + 3 * n] = FN(eq_a1)(uy, vela2, dd_param);

        vab = ux + uy;
        vab2 = vab * vab;
        dd[id + 4 * n] = FN(eq_4)(vab, vab2, dd_param);
        dd[id + 5 * n] = FN(eq_5)(vab, vab2, dd_param);
        vab = ux - uy;
        vab2 = vab * vab;
        dd[id + 6 * n] = FN(eq_4)(vab, vab2, dd_param);
        dd[id + 7 * n] = FN(eq_5)(vab, vab2, dd_param);

        vab = ux + uz;
        vab2 = vab * vab;
        dd[id + 8 * n] = FN(eq_4)(vab, vab2, dd_param);
        dd[id + 9 * n] = FN(eq_5)(vab, vab2, dd_param);
        vab = ux - uz;
        vab2 = vab * vab;
        dd[id + 10 * n] = FN(eq_4)(vab, vab2, dd_param);
        dd[id + 11 * n] = FN(eq_5)(vab, vab2, dd_param);

        vab = uy + uz;
        vab2 = vab * vab;
        dd[id + 12 * n] = FN(eq_4)(vab, vab2, dd_param);
        dd[id + 13 * n] = FN(eq_5)(vab, vab2, dd_param);
        vab = uy - uz;
        vab2 = vab * vab;
        dd[id + 14 * n] = FN(eq_4)(vab, vab2, dd_param);
        dd[id + 15 * n] = FN(eq_5)(vab, vab2, dd_param);

        vela2 = uz * uz;
        dd[id + 16 * n] = FN(eq_a0)(uz, vela2, dd_param);
        dd[id + 17 * n] = FN(eq_a1)(uz, vela2, dd_param);
        dd[id + 18 * n] = FN(eq_18)(dd_param);

        if (store_u) {
            u_out[id] = ux;
            u_out[id + n] = uy;
            u_out[id + 2 * n] = uz;
        }
        if (store_rho)
            rho_out[id] = rho;
    }
}

/* CLbmSolverCPU::simulationStepAlpha(origin, size) -> alphaKernelCPU over the
 * box (CLbmSolverCPU.cpp:305-345). Cells are independent, so the reference's
 * cache blocking (CLbmAlphaCPU.cpp:54-68) does not change any result. */
void FN(oracle_alpha)(T* dd, const int32_t* flags, T* rho_out, T* u_out, T inv_tau,
                      T lid_velocity, const T g[3], const int size_h[3], const int origin[3],
                      const int box[3], int store_rho, int store_u, int nthreads)
{
    long nx = size_h[0], nxy = (long)size_h[0] * size_h[1];
    long n = nxy * size_h[2];
    int z;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (z = 0; z < box[2]; z++) {
        int y, x;
        for (y = 0; y < box[1]; y++)
            for (x = 0; x < box[0]; x++) {
                long id = (long)(origin[2] + z) * nxy + (long)(origin[1] + y) * nx + (origin[0] + x);
                FN(alpha_cell)(dd, n, id, flags[id], rho_out, u_out, inv_tau, lid_velocity, g,
                               store_rho, store_u);
            }
    }
}

/* CLbmSolverCPU::simulationStepBeta(origin, size) -> betaKernelCPU
 * (CLbmSolverCPU.cpp:430-471). Each (cell, slot) pair is read and written by
 * exactly one cell, so any traversal order gives the same result. */
void FN(oracle_beta)(T* dd, const int32_t* flags, T* rho_out, T* u_out, T inv_tau,
                     T lid_velocity, const int size_h[3], const int origin[3], const int box[3],
                     int store_rho, int store_u, int nthreads)
{
    long nx = size_h[0], nxy = (long)size_h[0] * size_h[1];
    long n = nxy * size_h[2];
    long delta[19];
    int l, z;
    for (l = 0; l < 19; l++)
        delta[l] = oracle_c[l][0] + oracle_c[l][1] * nx + oracle_c[l][2] * nxy;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (z = 0; z < box[2]; z++) {
        int y, x;
        for (y = 0; y < box[1]; y++)
            for (x = 0; x < box[0]; x++) {
                long id = (long)(origin[2] + z) * nxy + (long)(origin[1] + y) * nx + (origin[0] + x);
                FN(beta_cell)(dd, n, id, delta, flags[id], rho_out, u_out, inv_tau, lid_velocity,
                              store_rho, store_u);
            }
    }
}

/* getVariable / getDensityDistributions box copy (CLbmSolverCPU.cpp:142-166,
 * 497-530): strided field -> dense dst[c][z][y][x] over the box. */
void FN(oracle_get_box)(const T* field, int ncomp, const int size_h[3], const int origin[3],
                        const int box[3], T* dst)
{
    long nx = size_h[0], nxy = (long)size_h[0] * size_h[1], n = nxy * size_h[2];
    long cells = (long)box[0] * box[1] * box[2];
    int c, z, y, x;
    for (c = 0; c < ncomp; c++)
        for (z = 0; z < box[2]; z++)
            for (y = 0; y < box[1]; y++)
                for (x = 0; x < box[0]; x++)
                    dst[c * cells + ((long)z * box[1] + y) * box[0] + x] =
                        field[c * n + (long)(origin[2] + z) * nxy + (long)(origin[1] + y) * nx +
                              (origin[0] + x)];
}

/* setVariable / setDensityDistributions[(direction)] (CLbmSolverCPU.cpp:194-218,
 * 562-676): dense src[c][z][y][x] -> strided field, for components in mask. */
void FN(oracle_set_box)(T* field, int ncomp, const int size_h[3], const int origin[3],
                        const int box[3], const T* src, uint32_t comp_mask)
{
    long nx = size_h[0], nxy = (long)size_h[0] * size_h[1], n = nxy * size_h[2];
    long cells = (long)box[0] * box[1] * box[2];
    int c, z, y, x;
    for (c = 0; c < ncomp; c++) {
        if (!((comp_mask >> c) & 1u))
            continue;
        for (z = 0; z < box[2]; z++)
            for (y = 0; y < box[1]; y++)
                for (x = 0; x < box[0]; x++)
                    field[c * n + (long)(origin[2] + z) * nxy + (long)(origin[1] + y) * nx +
                          (origin[0] + x)] = src[c * cells + ((long)z * box[1] + y) * box[0] + x];
    }
}

/* Density and momentum of every cell of a box from its populations, with the
 * accumulation of CLbmAlphaCPU.cpp:87-182 (the `moments` above, which the
 * alpha and beta cells use): dst[4][z][y][x] = rho, jx, jy, jz. layout 0: slot
 * l holds direction l (after beta or init); 1: slot l holds direction opp(l)
 * (after alpha). */
void FN(oracle_moments)(const T* dd, const int size_h[3], const int origin[3], const int box[3],
                        int layout, T* dst)
{
    long nx = size_h[0], nxy = (long)size_h[0] * size_h[1], n = nxy * size_h[2];
    long cells = (long)box[0] * box[1] * box[2];
    int z, y, x, l;
    for (z = 0; z < box[2]; z++)
        for (y = 0; y < box[1]; y++)
            for (x = 0; x < box[0]; x++) {
                long id = (long)(origin[2] + z) * nxy + (long)(origin[1] + y) * nx + (origin[0] + x);
                long g = ((long)z * box[1] + y) * box[0] + x;
                T d[19], r, jx, jy, jz;
                for (l = 0; l < 19; l++)
                    d[layout ? (l < 18 ? (l ^ 1) : 18) : l] = dd[l * n + id];
                FN(moments)(d, &r, &jx, &jy, &jz);
                dst[g] = r;
                dst[cells + g] = jx;
                dst[2 * cells + g] = jy;
                dst[3 * cells + g] = jz;
            }
}

static T FN(length3)(const T v[3])
{
    return FN(oracle_sqrt)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

/*
 * Time-step adaptation of the CLbmSolver constructor (CLbmSolver.cpp:43-281),
 * all in T except where the reference mixes in its double constants
 * REYNOLDS_DEFAULT and TAU_DEFAULT (constants.h:26-27).
 * Returns 0, or -1 where the reference prints "No valid timestep size" and
 * exits (CLbmSolver.cpp:263-272). The reference loop has no iteration cap;
 * this one stops after 64 rounds and returns -2.
 */
int FN(oracle_solver_params)(T length_x, int size_x, T length_max, const T velocity[3],
                             const T acceleration[3], T viscosity, T timestep, T max_velocity,
                             T max_acceleration, T out_vel_dimless[3], T out_acc_dimless[3],
                             T* out_visc_dimless, T* out_tau, T* out_inv_tau, T* out_timestep)
{
    const T eps = ORACLE_EPS;
    int limited_by_velocity = 0, limited_by_acceleration = 0, iter;
    T cell_length = length_x / (T)size_x;
    T old_timestep, tau = (T)0, visc_dimless = (T)0;
    T vel_dimless[3], acc_dimless[3];

    if (viscosity <= (T)0)
        viscosity = length_max * velocity[0] / ORACLE_REYNOLDS_DEFAULT;

    for (iter = 0;; iter++) {
        T s;
        int i;
        if (iter >= 64)
            return -2;
        s = timestep / cell_length;
        for (i = 0; i < 3; i++)
            vel_dimless[i] = velocity[i] * s;
        s = (timestep * timestep) / cell_length;
        for (i = 0; i < 3; i++)
            acc_dimless[i] = acceleration[i] * s;
        visc_dimless = viscosity * (timestep / (cell_length * cell_length));

        if (FN(length3)(vel_dimless) > max_velocity + eps) {
            limited_by_velocity = 1;
            old_timestep = timestep;
            timestep = (max_velocity * cell_length) / FN(length3)(velocity);
            continue;
        }
        if (FN(length3)(acc_dimless) > max_acceleration + eps) {
            limited_by_acceleration = 1;
            old_timestep = timestep;
            timestep = FN(oracle_sqrt)((max_acceleration * cell_length) / FN(length3)(acceleration));
            continue;
        }
        tau = (T)0.5 * ((T)6 * visc_dimless + (T)1);
        if (tau - eps < (T)ORACLE_TAU_LOWER_LIMIT || tau + eps > (T)ORACLE_TAU_UPPER_LIMIT) {
            old_timestep = timestep;
            timestep = (cell_length * cell_length) *
                       (((T)2 * ORACLE_TAU_DEFAULT - (T)1) / ((T)6 * viscosity));
            if ((limited_by_velocity || limited_by_acceleration) && timestep > old_timestep)
                return -1;
            continue;
        }
        break;
    }
    (void)old_timestep;
    out_vel_dimless[0] = vel_dimless[0];
    out_vel_dimless[1] = vel_dimless[1];
    out_vel_dimless[2] = vel_dimless[2];
    out_acc_dimless[0] = acc_dimless[0];
    out_acc_dimless[1] = acc_dimless[1];
    out_acc_dimless[2] = acc_dimless[2];
    *out_visc_dimless = visc_dimless;
    *out_tau = tau;
    *out_inv_tau = (T)1 / tau;
    *out_timestep = timestep;
    return 0;
}

#undef FN
#undef CAT
#undef CAT_
