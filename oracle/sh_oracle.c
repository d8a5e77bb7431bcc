/*
 * sh_oracle.c -- CPU restatement of DENISE-SH's SH time-stepping and
 * adjoint-gradient path.  TEST INFRASTRUCTURE ONLY: this file checks the
 * HIP product path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load the libraries built from it.
 *
 * The file is compiled twice (see oracle/Makefile):
 *   -DORC_REAL=float  -DORC_SFX=_f32   the reference's own arithmetic: float
 *                                      storage, float operators, and the
 *                                      reference's double-literal promotions
 *                                      reproduced by keeping its literals;
 *   -DORC_REAL=double -DORC_SFX=_f64   the same algorithm in FP64 (the
 *                                      "FP64 oracle" every result is compared
 *                                      against for its error statement).
 * Both builds use -ffp-contract=off so no multiply-add is fused, as in the
 * reference's x86 builds (icc -xAVX / gcc x86-64 have no FMA contraction).
 *
 * Arrays follow the reference's index space: wavefields and model fields are
 * NR matrices [-nd+1..NY+nd][-nd+1..NX+nd] with nd = FDORDER/2+1, flattened
 * row-major (x fastest) with width W = NX+2nd.  Memory variables r, q and the
 * coefficient fields dip, d are stored [l][j][i] (one padded plane per
 * mechanism) instead of the reference's f3tensor [j][i][l]; only the storage
 * order differs.  CPML vectors are 1-based [1..2FW].
 *
 * Pinning of each function (what checks it beyond itself) is listed in the
 * header block before it.  Functions without an independent pin say
 * "parity unpinned".
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#ifndef ORC_REAL
#define ORC_REAL float
#endif
#ifndef ORC_SFX
#define ORC_SFX _f32
#endif

typedef ORC_REAL real;

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define F(name) CAT(name, ORC_SFX)

#ifndef PI
#define PI (3.141592653589793)
#endif
/* fd.h:18 */
#define iround(x) ((int)(floor)(x + 0.5))

/* ------------------------------------------------------------------------
 * Configuration shared by all entry points.  Scalar parameters arrive as
 * doubles holding float32 values (the reference reads them as float).
 * ---------------------------------------------------------------------- */
typedef struct {
    int nx, ny;          /* interior grid                                   */
    int fdorder;         /* 2..12                                           */
    int maxrelerror;     /* 0..4 (Holberg table row)                        */
    int fw;              /* CPML width in grid points                       */
    int free_surf;       /* 1: implicit free surface on top                 */
    int L;               /* relaxation mechanisms, 0 = elastic              */
    int nt;              /* time steps                                      */
    int dtinv;           /* snapshot stride                                 */
    double dt, dh;
    double damping, fpml, npower, k_max_pml;
    double fl[8];        /* relaxation frequencies FL[1..L] at fl[0..L-1]   */
    int eprecond;        /* EPRECOND: 1 = eprecond1 energy scaling sqrt(Ws Wr), 3 = the receiver
                            approximation (eprecond1.c:34-50), 0 = none */
    double c_vs, c_rho;  /* > 0: the normalisation means C_vs, C_rho (fixed at
                            the first FWI iteration, denise.c:745-816);
                            0: the means of the model passed in            */
    const unsigned char *skip_res, *skip_en;   /* [nshots][ntr] data
                            conditioning masks of orc_residual_cond, or NULL */
    int periodic;        /* BOUNDARY 1: no CPML at the left and right edges
                            (update_v_PML.c:63,71; update_s_*_PML.c: the same
                            guards) and the wavefields wrapped across them
                            after each half step (exchange_v.c:99-161 for vy,
                            exchange_s.c:95-157 for sxy, FDORDER/2 columns) */
    int snap_bf16;       /* 1: the gradient's forward snapshots are held in
                            BF16 with a power-of-two scale per snapshot
                            (orc_snap_bf16) before the adjoint reads them  */
} orc_cfg;

#define ND(c) ((c)->fdorder / 2 + 1)
#define WID(c) ((c)->nx + 2 * ND(c))
#define HEI(c) ((c)->ny + 2 * ND(c))
/* element (j,i) of a padded field, j,i in the reference's index space */
#define AT(a, c, j, i) ((a)[((size_t)((j) + ND(c) - 1)) * WID(c) + (size_t)((i) + ND(c) - 1)])

size_t F(orc_padded_size)(const orc_cfg *c) { return (size_t)WID(c) * (size_t)HEI(c); }

/* ------------------------------------------------------------------------
 * holbergcoeff.c:19-90.  Pinned by tests: Taylor rows equal the closed-form
 * staggered-grid coefficients; Holberg rows reproduce the published table
 * values typed independently in the test.
 * ---------------------------------------------------------------------- */
static const float HCALL[5][6][7] = {
    {{23.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0},
     {8.0, 9.0 / 8.0, -1.0 / 24.0, 0.0, 0.0, 0.0, 0.0},
     {6.0, 75.0 / 64.0, -25.0 / 384.0, 3.0 / 640.0, 0.0, 0.0, 0.0},
     {5.0, 1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0, 0.0, 0.0},
     {5.0, 19845.0 / 16384.0, -735.0 / 8192.0, 567.0 / 40960.0, -405.0 / 229376.0, 35.0 / 294912.0, 0.0},
     {4.0, 160083.0 / 131072.0, -12705.0 / 131072.0, 22869.0 / 1310720.0, -5445.0 / 1835008.0, 847.0 / 2359296.0, -63.0 / 2883584.0}},
    {{49.7, 1.0010, 0.0, 0.0, 0.0, 0.0, 0.0},
     {8.32, 1.1382, -0.046414, 0.0, 0.0, 0.0, 0.0},
     {4.77, 1.1965, -0.078804, 0.0081781, 0.0, 0.0, 0.0},
     {3.69, 1.2257, -0.099537, 0.018063, -0.0026274, 0.0, 0.0},
     {3.19, 1.2415, -0.11231, 0.026191, -0.0064682, 0.001191, 0.0},
     {2.91, 1.2508, -0.12034, 0.032131, -0.010142, 0.0029857, -0.00066667}},
    {{22.2, 1.0050, 0.0, 0.0, 0.0, 0.0, 0.0},
     {5.65, 1.1534, -0.052806, 0.0, 0.0, 0.0, 0.0},
     {3.74, 1.2111, -0.088313, 0.011768, 0.0, 0.0, 0.0},
     {3.11, 1.2367, -0.10815, 0.023113, -0.0046905, 0.0, 0.0},
     {2.80, 1.2496, -0.11921, 0.031130, -0.0093272, 0.0025161, 0.0},
     {2.62, 1.2568, -0.12573, 0.036423, -0.013132, 0.0047484, -0.0015979}},
    {{15.8, 1.0100, 0.0, 0.0, 0.0, 0.0, 0.0},
     {4.80, 1.1640, -0.057991, 0.0, 0.0, 0.0, 0.0},
     {3.39, 1.2192, -0.094070, 0.014608, 0.0, 0.0, 0.0},
     {2.90, 1.2422, -0.11269, 0.026140, -0.0064054, 0.0, 0.0},
     {2.65, 1.2534, -0.12257, 0.033755, -0.011081, 0.0036784, 0.0},
     {2.51, 1.2596, -0.12825, 0.038550, -0.014763, 0.0058619, -0.0024538}},
    {{9.16, 1.0300, 0.0, 0.0, 0.0, 0.0, 0.0},
     {3.47, 1.1876, -0.072518, 0.0, 0.0, 0.0, 0.0},
     {2.91, 1.2341, -0.10569, 0.022589, 0.0, 0.0, 0.0},
     {2.61, 1.2516, -0.12085, 0.032236, -0.011459, 0.0, 0.0},
     {2.45, 1.2596, -0.12829, 0.038533, -0.014681, 0.0072580, 0.0},
     {2.36, 1.2640, -0.13239, 0.042217, -0.017803, 0.0081959, -0.0051848}},
};

int F(orc_holberg)(int fdorder, int maxrelerror, real *hc)
{
    int i;
    if (fdorder < 2 || fdorder > 12 || (fdorder & 1)) return -1;
    if (maxrelerror < 0 || maxrelerror > 4) return -2;
    /* the reference table is float; in FP64 mode it is promoted unchanged */
    for (i = 0; i <= 6; i++) hc[i] = (real)HCALL[maxrelerror][fdorder / 2 - 1][i];
    return 0;
}

/* ------------------------------------------------------------------------
 * PML_pro.c:65-297 (CPML profiles, Komatitsch & Martin 2007).  All twenty
 * vectors are 1-based [1..2FW] and must be zero-filled by the caller, as the
 * reference's vector() does (util.c:77-87).  Reproduces Q1: the `a`
 * coefficients are set only where the *integer* abs() of d is nonzero
 * (PML_pro.c:153-154, 200-201, 250-251, 292-293).  Pinned by a numpy FP64
 * brute-force evaluation of the closed forms in the tests.
 * ---------------------------------------------------------------------- */
void F(orc_pml_pro)(const orc_cfg *c, real *d_x, real *K_x, real *alpha_prime_x, real *a_x, real *b_x,
                    real *d_x_half, real *K_x_half, real *alpha_prime_x_half, real *a_x_half, real *b_x_half,
                    real *d_y, real *K_y, real *alpha_prime_y, real *a_y, real *b_y,
                    real *d_y_half, real *K_y_half, real *alpha_prime_y_half, real *a_y_half, real *b_y_half)
{
    const int FW = c->fw, NXG = c->nx, NYG = c->ny;
    const real DH = (real)c->dh, DT = (real)c->dt, DAMPING = (real)c->damping, FPML = (real)c->fpml;
    const real npower = (real)c->npower, k_max_PML = (real)c->k_max_pml;
    const real alpha_max_PML = 2.0 * PI * (FPML / 2.0);
    real thickness_PML_x, thickness_PML_y, xoriginleft, xoriginright, yoriginbottom, yorigintop;
    real Rcoef, d0_x, d0_y, xval, yval, abscissa_in_PML, abscissa_normalized;
    int i, h;

    thickness_PML_x = FW * DH;
    thickness_PML_y = FW * DH;
    Rcoef = 0.001;
    d0_x = -(npower + 1) * DAMPING * log(Rcoef) / (2.0 * thickness_PML_x);
    d0_y = -(npower + 1) * DAMPING * log(Rcoef) / (2.0 * thickness_PML_y);

    xoriginleft = thickness_PML_x;
    xoriginright = (NXG - 1) * DH - thickness_PML_x;

    for (i = 1; i <= FW; i++) {
        K_x[i] = 1.0;
        K_x_half[i] = 1.0;
        xval = DH * (i - 1);
        abscissa_in_PML = xoriginleft - xval;
        if (abscissa_in_PML >= 0.0) {
            abscissa_normalized = abscissa_in_PML / thickness_PML_x;
            d_x[i] = d0_x * pow(abscissa_normalized, npower);
            K_x[i] = 1.0 + (k_max_PML - 1.0) * pow(abscissa_normalized, npower);
            alpha_prime_x[i] = alpha_max_PML * (1.0 - abscissa_normalized) + 0.1 * alpha_max_PML;
        }
        abscissa_in_PML = xoriginleft - (xval + DH / 2.0);
        if (abscissa_in_PML >= 0.0) {
            abscissa_normalized = abscissa_in_PML / thickness_PML_x;
            d_x_half[i] = d0_x * pow(abscissa_normalized, npower);
            K_x_half[i] = 1.0 + (k_max_PML - 1.0) * pow(abscissa_normalized, npower);
            alpha_prime_x_half[i] = alpha_max_PML * (1.0 - abscissa_normalized) + 0.1 * alpha_max_PML;
        }
        if (alpha_prime_x[i] < 0.0) alpha_prime_x[i] = 0.0;
        if (alpha_prime_x_half[i] < 0.0) alpha_prime_x_half[i] = 0.0;
        b_x[i] = exp(-(d_x[i] / K_x[i] + alpha_prime_x[i]) * DT);
        b_x_half[i] = exp(-(d_x_half[i] / K_x_half[i] + alpha_prime_x_half[i]) * DT);
        if (abs((int)d_x[i]) > 1.0e-6)
            a_x[i] = d_x[i] * (b_x[i] - 1.0) / (K_x[i] * (d_x[i] + K_x[i] * alpha_prime_x[i]));
        if (abs((int)d_x_half[i]) > 1.0e-6)
            a_x_half[i] = d_x_half[i] * (b_x_half[i] - 1.0) /
                          (K_x_half[i] * (d_x_half[i] + K_x_half[i] * alpha_prime_x_half[i]));
    }

    for (i = NXG - FW + 1; i <= NXG; i++) {
        h = i - NXG + 2 * FW;
        K_x[h] = 1.0;
        K_x_half[h] = 1.0;
        xval = DH * (i - 1);
        abscissa_in_PML = xval - xoriginright;
        if (abscissa_in_PML >= 0.0) {
            abscissa_normalized = abscissa_in_PML / thickness_PML_x;
            d_x[h] = d0_x * pow(abscissa_normalized, npower);
            K_x[h] = 1.0 + (k_max_PML - 1.0) * pow(abscissa_normalized, npower);
            alpha_prime_x[h] = alpha_max_PML * (1.0 - abscissa_normalized) + 0.1 * alpha_max_PML;
        }
        abscissa_in_PML = xval + DH / 2.0 - xoriginright;
        if (abscissa_in_PML >= 0.0) {
            abscissa_normalized = abscissa_in_PML / thickness_PML_x;
            d_x_half[h] = d0_x * pow(abscissa_normalized, npower);
            K_x_half[h] = 1.0 + (k_max_PML - 1.0) * pow(abscissa_normalized, npower);
            alpha_prime_x_half[h] = alpha_max_PML * (1.0 - abscissa_normalized) + 0.1 * alpha_max_PML;
        }
        if (alpha_prime_x[h] < 0.0) alpha_prime_x[h] = 0.0;
        if (alpha_prime_x_half[h] < 0.0) alpha_prime_x_half[h] = 0.0;
        b_x[h] = exp(-(d_x[h] / K_x[h] + alpha_prime_x[h]) * DT);
        b_x_half[h] = exp(-(d_x_half[h] / K_x_half[h] + alpha_prime_x_half[h]) * DT);
        if (abs((int)d_x[h]) > 1.0e-6)
            a_x[h] = d_x[h] * (b_x[h] - 1.0) / (K_x[h] * (d_x[h] + K_x[h] * alpha_prime_x[h]));
        if (abs((int)d_x_half[h]) > 1.0e-6)
            a_x_half[h] = d_x_half[h] * (b_x_half[h] - 1.0) /
                          (K_x_half[h] * (d_x_half[h] + K_x_half[h] * alpha_prime_x_half[h]));
    }

    yoriginbottom = thickness_PML_y;
    yorigintop = (NYG - 1) * DH - thickness_PML_y;

    for (i = 1; i <= FW; i++) {
        K_y[i] = 1.0;
        K_y_half[i] = 1.0;
        yval = DH * (i - 1);
        abscissa_in_PML = yoriginbottom - yval;
        if (abscissa_in_PML >= 0.0) {
            abscissa_normalized = abscissa_in_PML / thickness_PML_y;
            d_y[i] = d0_y * pow(abscissa_normalized, npower);
            K_y[i] = 1.0 + (k_max_PML - 1.0) * pow(abscissa_normalized, npower);
            alpha_prime_y[i] = alpha_max_PML * (1.0 - abscissa_normalized) + 0.1 * alpha_max_PML;
        }
        abscissa_in_PML = yoriginbottom - (yval + DH / 2.0);
        if (abscissa_in_PML >= 0.0) {
            abscissa_normalized = abscissa_in_PML / thickness_PML_y;
            d_y_half[i] = d0_y * pow(abscissa_normalized, npower);
            K_y_half[i] = 1.0 + (k_max_PML - 1.0) * pow(abscissa_normalized, npower);
            alpha_prime_y_half[i] = alpha_max_PML * (1.0 - abscissa_normalized) + 0.1 * alpha_max_PML;
        }
        b_y[i] = exp(-(d_y[i] / K_y[i] + alpha_prime_y[i]) * DT);
        b_y_half[i] = exp(-(d_y_half[i] / K_y_half[i] + alpha_prime_y_half[i]) * DT);
        if (abs((int)d_y[i]) > 1.0e-6)
            a_y[i] = d_y[i] * (b_y[i] - 1.0) / (K_y[i] * (d_y[i] + K_y[i] * alpha_prime_y[i]));
        if (abs((int)d_y_half[i]) > 1.0e-6)
            a_y_half[i] = d_y_half[i] * (b_y_half[i] - 1.0) /
                          (K_y_half[i] * (d_y_half[i] + K_y_half[i] * alpha_prime_y_half[i]));
    }

    for (i = NYG - FW + 1; i <= NYG; i++) {
        h = i - NYG + 2 * FW;
        K_y[h] = 1.0;
        K_y_half[h] = 1.0;
        yval = DH * (i - 1);
        abscissa_in_PML = yval - yorigintop;
        if (abscissa_in_PML >= 0.0) {
            abscissa_normalized = abscissa_in_PML / thickness_PML_y;
            d_y[h] = d0_y * pow(abscissa_normalized, npower);
            K_y[h] = 1.0 + (k_max_PML - 1.0) * pow(abscissa_normalized, npower);
            alpha_prime_y[h] = alpha_max_PML * (1.0 - abscissa_normalized) + 0.1 * alpha_max_PML;
        }
        abscissa_in_PML = yval + DH / 2.0 - yorigintop;
        if (abscissa_in_PML >= 0.0) {
            abscissa_normalized = abscissa_in_PML / thickness_PML_y;
            /* PML_pro.c:281 uses d0_x here; d0_x == d0_y because both strips
               have the same thickness, so the value is the same */
            d_y_half[h] = d0_x * pow(abscissa_normalized, npower);
            K_y_half[h] = 1.0 + (k_max_PML - 1.0) * pow(abscissa_normalized, npower);
            alpha_prime_y_half[h] = alpha_max_PML * (1.0 - abscissa_normalized) + 0.1 * alpha_max_PML;
        }
        b_y[h] = exp(-(d_y[h] / K_y[h] + alpha_prime_y[h]) * DT);
        b_y_half[h] = exp(-(d_y_half[h] / K_y_half[h] + alpha_prime_y_half[h]) * DT);
        if (abs((int)d_y[h]) > 1.0e-6)
            a_y[h] = d_y[h] * (b_y[h] - 1.0) / (K_y[h] * (d_y[h] + K_y[h] * alpha_prime_y[h]));
        if (abs((int)d_y_half[h]) > 1.0e-6)
            a_y_half[h] = d_y_half[h] * (b_y_half[h] - 1.0) /
                          (K_y_half[h] * (d_y_half[h] + K_y_half[h] * alpha_prime_y_half[h]));
    }
}

/* ------------------------------------------------------------------------
 * matcopy.c:12-137 / matcopy_elastic.c with one rank: every neighbour index
 * wraps to the rank itself (initproc.c:52-55), so the global-edge model halo
 * becomes a periodic copy, rows first and then columns (Q3).  Pinned by the
 * test that compares against numpy.pad(mode="wrap") on the one-cell ring.
 * ---------------------------------------------------------------------- */
void F(orc_matcopy_wrap)(const orc_cfg *c, real *x)
{
    const int NX = c->nx, NY = c->ny;
    int i, j;
    for (i = 0; i <= NX + 1; i++) AT(x, c, NY + 1, i) = AT(x, c, 1, i);
    for (i = 0; i <= NX + 1; i++) AT(x, c, 0, i) = AT(x, c, NY, i);
    for (j = 0; j <= NY + 1; j++) AT(x, c, j, NX + 1) = AT(x, c, j, 1);
    for (j = 0; j <= NY + 1; j++) AT(x, c, j, 0) = AT(x, c, j, NX);
}

/* ------------------------------------------------------------------------
 * av_mue.c:28-51 (INVMAT1 == 1): mu = rho*vs*vs and harmonic averages to the
 * x- and y-half points.  Pinned by a numpy brute force in the tests.
 * ---------------------------------------------------------------------- */
void F(orc_av_mue)(const orc_cfg *c, const real *u, const real *rho, real *uip, real *ujp)
{
    const int NX = c->nx, NY = c->ny;
    int i, j;
    real u1, u2, u3;
    for (j = 1; j <= NY; j++) {
        for (i = 1; i <= NX; i++) {
            u1 = AT(rho, c, j, i) * AT(u, c, j, i) * AT(u, c, j, i);
            u2 = AT(rho, c, j, i + 1) * AT(u, c, j, i + 1) * AT(u, c, j, i + 1);
            u3 = AT(rho, c, j + 1, i) * AT(u, c, j + 1, i) * AT(u, c, j + 1, i);
            AT(uip, c, j, i) = 2.0 / ((1.0 / u1) + (1.0 / u2));
            AT(ujp, c, j, i) = 2.0 / ((1.0 / u1) + (1.0 / u3));
            if ((u1 == 0.0) || (u2 == 0.0)) AT(uip, c, j, i) = 0.0;
            if ((u1 == 0.0) || (u3 == 0.0)) AT(ujp, c, j, i) = 0.0;
        }
    }
}

/* av_tau.c:5-15.  Pinned by a numpy brute force. */
void F(orc_av_tau)(const orc_cfg *c, const real *taus, real *tausipjp)
{
    int i, j;
    for (j = 1; j <= c->ny; j++)
        for (i = 1; i <= c->nx; i++)
            AT(tausipjp, c, j, i) = 0.25 * (AT(taus, c, j, i) + AT(taus, c, j, i + 1) + AT(taus, c, j + 1, i) +
                                            AT(taus, c, j + 1, i + 1));
}

/* ------------------------------------------------------------------------
 * readmod.c:30-34 (eta[l] = DT/pts[l], pts[l] = 1/(2 pi FL[l])) and
 * prepare_update_s.c:20-62 (INVMAT1 == 1).  Outputs per-cell fipjp, f, and
 * per-mechanism planes dip[l], d[l] ([l][j][i]); per-mechanism scalars
 * bip, bjm, cip, cjm (0-based l).  Pinned: with taus == 0 the outputs reduce
 * to fipjp = uip*DT, f = ujp*DT, dip = d = 0 (checked in the tests), and the
 * visco update with these reproduces the elastic update (test); the r/q
 * recursion is checked against its closed-form single-cell decay.
 * ---------------------------------------------------------------------- */
void F(orc_prepare_update_s)(const orc_cfg *c, const real *pujp, const real *puip, const real *ptaus,
                             const real *ptausipjp, real *fipjp, real *f, real *dip, real *d, real *bip, real *bjm,
                             real *cip, real *cjm)
{
    const int NX = c->nx, NY = c->ny, L = c->L;
    const real DT = (real)c->dt;
    const size_t plane = F(orc_padded_size)(c);
    real pts[9], eta[9], etaip[9], etajm[9], FL[9];
    real sumu, ws, mujp, muip;
    int i, j, l;

    for (l = 1; l <= L; l++) FL[l] = (real)c->fl[l - 1];
    for (l = 1; l <= L; l++) {
        pts[l] = 1.0 / (2.0 * PI * FL[l]);
        eta[l] = DT / pts[l];
    }
    ws = 2.0 * PI * FL[1];
    sumu = 0.0;
    for (l = 1; l <= L; l++) sumu = sumu + ((ws * ws * pts[l] * pts[l]) / (1.0 + ws * ws * pts[l] * pts[l]));
    for (l = 1; l <= L; l++) {
        etajm[l] = eta[l];
        etaip[l] = eta[l];
    }
    for (l = 1; l <= L; l++) {
        bip[l - 1] = 1.0 / (1.0 + (etaip[l] * 0.5));
        bjm[l - 1] = 1.0 / (1.0 + (etajm[l] * 0.5));
        cip[l - 1] = 1.0 - (etaip[l] * 0.5);
        cjm[l - 1] = 1.0 - (etajm[l] * 0.5);
    }
    for (j = 1; j <= NY; j++) {
        for (i = 1; i <= NX; i++) {
            mujp = AT(pujp, c, j, i) / (1.0 + sumu * AT(ptaus, c, j, i));
            muip = AT(puip, c, j, i) / (1.0 + sumu * AT(ptausipjp, c, j, i));
            AT(fipjp, c, j, i) = muip * DT * (1.0 + L * AT(ptausipjp, c, j, i));
            AT(f, c, j, i) = mujp * DT * (1.0 + L * AT(ptaus, c, j, i));
            for (l = 1; l <= L; l++) {
                AT(dip + (size_t)(l - 1) * plane, c, j, i) = muip * etaip[l] * AT(ptausipjp, c, j, i);
                AT(d + (size_t)(l - 1) * plane, c, j, i) = mujp * etajm[l] * AT(ptaus, c, j, i);
            }
        }
    }
}

/* ------------------------------------------------------------------------
 * checkfd_ssg_visc.c:14-126 / checkfd_ssg_elastic.c: stability limit
 * dtstab = DH/(sqrt(2)*gamma*cmax).  ts = TS (source duration).  Pinned by
 * the closed form for a homogeneous elastic model (cmax = vs).
 * ---------------------------------------------------------------------- */
double F(orc_dtstab)(const orc_cfg *c, const float *vs, const float *rho, const float *qs, double TS)
{
    const int NX = c->nx, NY = c->ny, L = c->L;
    const real DT = (real)c->dt, DH = (real)c->dh;
    real hc[7], pts, eta[9], ws, sumu, ts, c1, cmax = 0.0, gamma, dtstab, taus, pu_ref;
    int i, j, l;
    F(orc_holberg)(c->fdorder, c->maxrelerror, hc);
    ws = 2.0 * PI / (real)TS;
    for (l = 1; l <= L; l++) {
        pts = 1.0 / (2.0 * PI * (real)c->fl[l - 1]);
        eta[l] = DT / pts;
    }
    for (i = 1; i <= NX; i++) {
        for (j = 1; j <= NY; j++) {
            size_t k = (size_t)(j - 1) * NX + (size_t)(i - 1);
            real rv = rho[k], uv = vs[k];
            sumu = 0.0;
            taus = (L > 0) ? (real)(2.0 / qs[k]) : (real)0.0;
            for (l = 1; l <= L; l++) {
                ts = DT / eta[l];
                sumu = sumu + ((ws * ws * ts * ts * taus) / (1.0 + ws * ws * ts * ts));
            }
            pu_ref = rv * uv * uv;
            if (L > 0)
                c1 = sqrt(pu_ref * (1.0 + L * taus) / (rv * (1.0 + sumu)));
            else
                c1 = uv; /* checkfd_ssg_elastic.c: c = pu for INVMAT1 == 1 */
            if (cmax < c1) cmax = c1;
        }
    }
    gamma = fabs(hc[1]) + fabs(hc[2]) + fabs(hc[3]) + fabs(hc[4]) + fabs(hc[5]) + fabs(hc[6]);
    dtstab = DH / (sqrt(2) * gamma * cmax);
    return (double)dtstab;
}

/* ------------------------------------------------------------------------
 * wavelet.c:12-85 for QUELLART 1 (DENISE's Ricker, Q9), 2 (Fuchs-Mueller),
 * 4 (sin^3), 5 (Gaussian derivative) and 6 (spike).  signals[nt-1] holds
 * sample nt = 1..NT.  Pinned by the closed forms in the tests.
 * ---------------------------------------------------------------------- */
int F(orc_wavelet)(int quellart, int NT, double dt_in, double tshift_in, double fc_in, double amp_in, real *out)
{
    const real DT = (real)dt_in, tshift = (real)tshift_in, fc = (real)fc_in, a = (real)amp_in;
    real t, ts, ag, tau, amp = 0.0;
    int nt;
    for (nt = 1; nt <= NT; nt++) {
        t = (real)nt * DT;
        ts = 1.0 / fc;
        ag = PI * PI * fc * fc;
        switch (quellart) {
        case 1:
            tau = PI * (t - 1.5 * ts - tshift) / (1.5 * ts);
            amp = (((1.0 - 4.0 * tau * tau) * exp(-2.0 * tau * tau)));
            break;
        case 2:
            if ((t < tshift) || (t > (tshift + ts)))
                amp = 0.0;
            else
                amp = ((sin(2.0 * PI * (t - tshift) * fc) - 0.5 * sin(4.0 * PI * (t - tshift) * fc)));
            break;
        case 4:
            if ((t < tshift) || (t > (tshift + ts)))
                amp = 0.0;
            else
                amp = pow(sin(PI * (t + tshift) / ts), 3.0);
            break;
        case 5:
            ts = 1.2 / fc;
            ag = PI * PI * fc * fc;
            amp = -2.0 * ag * (t - ts) * exp(-ag * (t - ts) * (t - ts));
            break;
        case 6:
            amp = 0.0;
            if (nt == 1 + iround(tshift / DT)) amp = 1.0;
            break;
        default:
            return -1;
        }
        out[nt - 1] = amp * a;
    }
    return 0;
}

/* ------------------------------------------------------------------------
 * Butterworth filter: libcseife cseife.c:147-177 (framework), cseife_rfk.c
 * (coefficients), cseife_rekfl.c (recursion), FP64.  orc_seife_butterworth
 * filters x[0..n-1] in place (lowpass: method 1, highpass: method 2).
 * Pinned bit-for-bit against libcseife compiled from the reference sources
 * into oracle/_ref, and against scipy.signal.bilinear + lfilter.
 * ---------------------------------------------------------------------- */
static void orc_rekfl(double *x, int n, double f0, double f1, double f2, double g1, double g2)
{
    double xa = 0.0, xaa = 0.0, ya = 0.0, yaa = 0.0, xn, y;
    int j;
    for (j = 0; j < n; j++) {
        xn = x[j];
        y = f0 * xn + f1 * xa + f2 * xaa + g1 * ya + g2 * yaa;
        x[j] = y;
        xaa = xa;
        xa = xn;
        yaa = ya;
        ya = y;
    }
}

void F(orc_seife_butterworth)(double *x, int n, double dt, double t0, int m, int method)
{
    double t0dt = t0 / dt, zpi = 8.0 * atan(1.0), eps = zpi / t0dt, pih = 2.0 * atan(1.0);
    int mm = m / 2, j, i;
    if (method == 2) {
        double x0 = x[0];
        for (i = 0; i < n; ++i) x[i] -= x0;
    }
    if (m > 2 * mm) {
        double g1 = (2.0 - eps) / (2.0 + eps), f0, f1;
        if (method == 1) {
            f0 = eps / (2.0 + eps);
            f1 = f0;
        } else {
            f0 = 2.0 / (2.0 + eps);
            f1 = -f0;
        }
        orc_rekfl(x, n, f0, f1, 0.0, g1, 0.0);
    }
    for (j = 1; j <= mm; ++j) {
        double h = sin(pih * (2 * j - 1) / m);
        double epsq = eps * eps;
        double a = 1.0 - eps * h + epsq / 4.0;
        double b = -2.0 + epsq / 2.0;
        double cc = 1.0 + eps * h + epsq / 4.0;
        double g1 = -b / cc, g2 = -a / cc, f0, f1, f2;
        if (method == 1) {
            f0 = epsq / 4.0 / cc;
            f1 = f0 + f0;
            f2 = f0;
        } else {
            f0 = 1.0 / cc;
            f1 = -f0 - f0;
            f2 = f0;
        }
        orc_rekfl(x, n, f0, f1, f2, g1, g2);
    }
}

/* timedomain_filt.c:10-53 including Q8: the 1-based trace is handed to the
 * 0-based filter, so the filter sees [0, s1, ..., s_{ns-1}] (the zeroed
 * allocation cell util.c:143-146 first) and s_ns stays unfiltered.
 * data is [ntr][ns] float, in place. */
void F(orc_timedomain_filt)(float *data, int ntr, int ns, double dt_in, double fc_in, int order, int method)
{
    const float DT = (float)dt_in, fc = (float)fc_in;
    double *seis = (double *)calloc((size_t)ns + 1, sizeof(double));
    double T0 = 1.0 / fc;
    int itr, j;
    for (itr = 0; itr < ntr; itr++) {
        float *row = data + (size_t)itr * ns;
        seis[0] = 0.0;
        for (j = 1; j <= ns; j++) seis[j] = (double)row[j - 1];
        F(orc_seife_butterworth)(seis, ns, DT, T0, order, method);
        for (j = 1; j <= ns; j++) row[j - 1] = (float)seis[j];
    }
    free(seis);
}

/* ------------------------------------------------------------------------
 * update_v_PML.c:12-583 (FDORDER cases 2..12), one sweep over the interior.
 * sw = 0 forward (vyp1 = acceleration), sw = 1 adjoint (vyp1 = old vy,
 * GRAD_FORM 2).  Q6: for FDORDER 12 the last syy_y term uses hc[5].
 * Source injection is separate (orc_inject).  psi_syy_y is [1..2FW][1..NX]
 * (width NX), psi_sxy_x is [1..NY][1..2FW] (width 2FW); both 1-based and
 * flattened with (row-1)*width + (col-1).
 * Each row is processed in three passes that keep every cell's operations
 * and their order: (1) the stencil sums for the whole row, (2) the CPML
 * corrections on the strip cells, in the reference's left, right, top,
 * bottom order, (3) the update.  Passes 1 and 3 vectorise.
 * Pinned by: FP64 energy conservation of the discrete leapfrog scheme (no
 * CPML, no source) and the travel time in a homogeneous medium (tests).
 * ---------------------------------------------------------------------- */
#define PSI_Y(p, c, row, col) ((p)[(size_t)((row)-1) * (c)->nx + (size_t)((col)-1)])
#define PSI_X(p, c, row, col) ((p)[(size_t)((row)-1) * 2 * (c)->fw + (size_t)((col)-1)])

/* stencil sums of update_v for row j, columns 1..NX, into sx[1..NX], sy[1..NX] */
static inline __attribute__((always_inline)) void v_sums(const orc_cfg *c, const int k, int j,
                                                         const real *restrict syz, const real *restrict sxy,
                                                         const real *restrict hc, real *restrict sx,
                                                         real *restrict sy)
{
    const int NX = c->nx, W = WID(c);
    const real *sr = &AT(sxy, c, j, 0);
    int i, m;
    for (i = 1; i <= NX; i++) {
        real a = hc[1] * (sr[i] - sr[i - 1]);
        real b = hc[1] * (AT(syz, c, j, i) - AT(syz, c, j - 1, i));
        for (m = 2; m <= k; m++) {
            a = a + hc[m] * (sr[i + m - 1] - sr[i - m]);
            b = b + hc[(k == 6 && m == 6) ? 5 : m] * (AT(syz, c, j + m - 1, i) - AT(syz, c, j - m, i));
        }
        sx[i] = a;
        sy[i] = b;
    }
    (void)W;
}

void F(orc_update_v)(const orc_cfg *c, int sw, real *vy, real *vyp1, const real *syz, const real *sxy,
                     const real *rho, const real *hc, const real *K_x, const real *a_x, const real *b_x,
                     const real *K_y, const real *a_y, const real *b_y, real *psi_syy_y, real *psi_sxy_x)
{
    const int NX = c->nx, NY = c->ny, FW = c->fw, k = c->fdorder / 2;
    const real DT = (real)c->dt, DH = (real)c->dh;
    int j;
#pragma omp parallel for schedule(static)
    for (j = 1; j <= NY; j++) {
        real sx[NX + 1], sy[NX + 1];
        int i, h1;
        switch (k) {
        case 1: v_sums(c, 1, j, syz, sxy, hc, sx, sy); break;
        case 2: v_sums(c, 2, j, syz, sxy, hc, sx, sy); break;
        case 3: v_sums(c, 3, j, syz, sxy, hc, sx, sy); break;
        case 4: v_sums(c, 4, j, syz, sxy, hc, sx, sy); break;
        case 5: v_sums(c, 5, j, syz, sxy, hc, sx, sy); break;
        default: v_sums(c, 6, j, syz, sxy, hc, sx, sy); break;
        }
        if (FW > 0) {
            for (i = 1; i <= NX; i++) {
                const int xl = !c->periodic && i <= FW, xr = !c->periodic && i >= NX - FW + 1;
                if (!(xl || xr || ((!c->free_surf) && (j <= FW)) || j >= NY - FW + 1)) continue;
                real sxy_x = sx[i], syy_y = sy[i];
                if (xl) {
                    PSI_X(psi_sxy_x, c, j, i) = b_x[i] * PSI_X(psi_sxy_x, c, j, i) + a_x[i] * sxy_x;
                    sxy_x = sxy_x / K_x[i] + PSI_X(psi_sxy_x, c, j, i);
                }
                if (xr) {
                    h1 = i - NX + 2 * FW;
                    PSI_X(psi_sxy_x, c, j, h1) = b_x[h1] * PSI_X(psi_sxy_x, c, j, h1) + a_x[h1] * sxy_x;
                    sxy_x = sxy_x / K_x[h1] + PSI_X(psi_sxy_x, c, j, h1);
                }
                if ((!c->free_surf) && (j <= FW)) {
                    PSI_Y(psi_syy_y, c, j, i) = b_y[j] * PSI_Y(psi_syy_y, c, j, i) + a_y[j] * syy_y;
                    syy_y = syy_y / K_y[j] + PSI_Y(psi_syy_y, c, j, i);
                }
                if (j >= NY - FW + 1) {
                    h1 = j - NY + 2 * FW;
                    PSI_Y(psi_syy_y, c, h1, i) = b_y[h1] * PSI_Y(psi_syy_y, c, h1, i) + a_y[h1] * syy_y;
                    syy_y = syy_y / K_y[h1] + PSI_Y(psi_syy_y, c, h1, i);
                }
                sx[i] = sxy_x;
                sy[i] = syy_y;
            }
        }
        {
            real *restrict v = &AT(vy, c, j, 0);
            real *restrict v1 = &AT(vyp1, c, j, 0);
            const real *restrict rr = &AT(rho, c, j, 0);
            for (i = 1; i <= NX; i++) {
                real b = 1.0 / rr[i];
                if (rr[i] < 1e-20) b = 0.0;
                if (sw == 0) v1[i] = b * (sx[i] + sy[i]) / DH;
                if (sw == 1) v1[i] = v[i];
                v[i] += b * DT * (sx[i] + sy[i]) / DH;
            }
        }
    }
}

/* update_v_PML.c:563-583: raw addition after the sweep (Q10). */
void F(orc_inject)(const orc_cfg *c, real *vy, int n, const int *ii, const int *jj, const real *val)
{
    int l;
    for (l = 0; l < n; l++) AT(vy, c, jj[l], ii[l]) += val[l];
}

/* stencil sums of the stress updates for row j: vyx, vyy (before *dhi) */
static inline __attribute__((always_inline)) void s_sums(const orc_cfg *c, const int k, int j,
                                                         const real *restrict vy, const real *restrict hc,
                                                         real *restrict gx, real *restrict gy)
{
    const int NX = c->nx;
    const real *vr = &AT(vy, c, j, 0);
    int i, m;
    for (i = 1; i <= NX; i++) {
        real a = hc[1] * (vr[i + 1] - vr[i]);
        real b = hc[1] * (AT(vy, c, j + 1, i) - AT(vy, c, j, i));
        for (m = 2; m <= k; m++) {
            a = a + hc[m] * (vr[i + m] - vr[i - m + 1]);
            b = b + hc[m] * (AT(vy, c, j + m, i) - AT(vy, c, j - m + 1, i));
        }
        gx[i] = a;
        gy[i] = b;
    }
}

#define S_SUMS_DISPATCH                                                                                         \
    switch (k) {                                                                                                \
    case 1: s_sums(c, 1, j, vy, hc, gx, gy); break;                                                             \
    case 2: s_sums(c, 2, j, vy, hc, gx, gy); break;                                                             \
    case 3: s_sums(c, 3, j, vy, hc, gx, gy); break;                                                             \
    case 4: s_sums(c, 4, j, vy, hc, gx, gy); break;                                                             \
    case 5: s_sums(c, 5, j, vy, hc, gx, gy); break;                                                             \
    default: s_sums(c, 6, j, vy, hc, gx, gy); break;                                                            \
    }

/* CPML of the stress updates on the strip cells of row j (reference order).
 * top_full selects the full-grid y profile on top (Q2, viscoelastic). */
static inline void s_cpml_row(const orc_cfg *c, int j, real *gx, real *gy, const real *K_x_half,
                              const real *a_x_half, const real *b_x_half, const real *Kt, const real *at,
                              const real *bt, const real *K_y_half, const real *a_y_half, const real *b_y_half,
                              real *psi_vyy, real *psi_vyx)
{
    const int NX = c->nx, NY = c->ny, FW = c->fw;
    int i, h1;
    for (i = 1; i <= NX; i++) {
        const int xl = !c->periodic && i <= FW, xr = !c->periodic && i >= NX - FW + 1;
        if (!(xl || xr || ((!c->free_surf) && (j <= FW)) || j >= NY - FW + 1)) continue;
        real vyx = gx[i], vyy = gy[i];
        if (xl) {
            PSI_X(psi_vyx, c, j, i) = b_x_half[i] * PSI_X(psi_vyx, c, j, i) + a_x_half[i] * vyx;
            vyx = vyx / K_x_half[i] + PSI_X(psi_vyx, c, j, i);
        }
        if (xr) {
            h1 = i - NX + 2 * FW;
            PSI_X(psi_vyx, c, j, h1) = b_x_half[h1] * PSI_X(psi_vyx, c, j, h1) + a_x_half[h1] * vyx;
            vyx = vyx / K_x_half[h1] + PSI_X(psi_vyx, c, j, h1);
        }
        if ((!c->free_surf) && (j <= FW)) {
            PSI_Y(psi_vyy, c, j, i) = bt[j] * PSI_Y(psi_vyy, c, j, i) + at[j] * vyy;
            vyy = vyy / Kt[j] + PSI_Y(psi_vyy, c, j, i);
        }
        if (j >= NY - FW + 1) {
            h1 = j - NY + 2 * FW;
            PSI_Y(psi_vyy, c, h1, i) = b_y_half[h1] * PSI_Y(psi_vyy, c, h1, i) + a_y_half[h1] * vyy;
            vyy = vyy / K_y_half[h1] + PSI_Y(psi_vyy, c, h1, i);
        }
        gx[i] = vyx;
        gy[i] = vyy;
    }
}

/* ------------------------------------------------------------------------
 * update_s_elastic_PML.c:12-446 (cases 2..12): dhi = DT/DH, half-grid CPML
 * on all four strips, uy/uyx include the DT factor (Q5).  psi_vyy is
 * [1..2FW][1..NX], psi_vyx is [1..NY][1..2FW].  uy/uyx may be NULL.
 * Pinned by energy conservation (with orc_update_v) and by the visco kernel
 * with taus = 0 (test).
 * ---------------------------------------------------------------------- */
void F(orc_update_s_elastic)(const orc_cfg *c, const real *vy, real *uy, real *uyx, real *syz, real *sxy,
                             const real *uip, const real *ujp, const real *hc, const real *K_x_half,
                             const real *a_x_half, const real *b_x_half, const real *K_y_half, const real *a_y_half,
                             const real *b_y_half, real *psi_vyy, real *psi_vyx)
{
    const int NX = c->nx, NY = c->ny, FW = c->fw, k = c->fdorder / 2;
    const real DT = (real)c->dt, DH = (real)c->dh;
    const real dhi = DT / DH;
    int j;
#pragma omp parallel for schedule(static)
    for (j = 1; j <= NY; j++) {
        real gx[NX + 1], gy[NX + 1];
        int i;
        S_SUMS_DISPATCH
        for (i = 1; i <= NX; i++) {
            gx[i] = gx[i] * dhi;
            gy[i] = gy[i] * dhi;
        }
        if (FW > 0)
            s_cpml_row(c, j, gx, gy, K_x_half, a_x_half, b_x_half, K_y_half, a_y_half, b_y_half, K_y_half,
                       a_y_half, b_y_half, psi_vyy, psi_vyx);
        {
            real *restrict sx = &AT(sxy, c, j, 0), *restrict sz = &AT(syz, c, j, 0);
            const real *restrict ui = &AT(uip, c, j, 0), *restrict uj = &AT(ujp, c, j, 0);
            if (uy) {
                real *restrict u1 = &AT(uy, c, j, 0), *restrict u2 = &AT(uyx, c, j, 0);
                for (i = 1; i <= NX; i++) {
                    u1[i] = gy[i];
                    u2[i] = gx[i];
                }
            }
            for (i = 1; i <= NX; i++) {
                sx[i] += ui[i] * gx[i];
                sz[i] += uj[i] * gy[i];
            }
        }
    }
}

/* ------------------------------------------------------------------------
 * update_s_visc_PML.c:12-577 (cases 2..12): dhi = 1/DH, dthalbe = DT/2,
 * x CPML half-grid, top CPML full-grid (Q2), bottom half-grid; memory
 * variables r (x) and q (y) stored [l][j][i]; uy/uyx without DT (Q5).
 * dip, d are [l][j][i] planes; bip, bjm, cip, cjm are 0-based [l].
 * Pinned by: taus = 0 reduces to the elastic update (to rounding), and the
 * single-cell memory-variable recursion matches its closed form (tests).
 * ---------------------------------------------------------------------- */
void F(orc_update_s_visc)(const orc_cfg *c, const real *vy, real *uy, real *uyx, real *syz, real *sxy, real *r,
                          real *q, const real *fipjp, const real *f, const real *dip, const real *d,
                          const real *bip, const real *bjm, const real *cip, const real *cjm, const real *hc,
                          const real *K_x_half, const real *a_x_half, const real *b_x_half, const real *K_y,
                          const real *a_y, const real *b_y, const real *K_y_half, const real *a_y_half,
                          const real *b_y_half, real *psi_vyy, real *psi_vyx)
{
    const int NX = c->nx, NY = c->ny, FW = c->fw, k = c->fdorder / 2, L = c->L;
    const real DT = (real)c->dt, DH = (real)c->dh;
    const size_t plane = F(orc_padded_size)(c);
    const real dhi = 1.0 / DH, dthalbe = DT / 2.0;
    int j;
#pragma omp parallel for schedule(static)
    for (j = 1; j <= NY; j++) {
        real gx[NX + 1], gy[NX + 1];
        int i, l;
        S_SUMS_DISPATCH
        for (i = 1; i <= NX; i++) {
            gx[i] = gx[i] * dhi;
            gy[i] = gy[i] * dhi;
        }
        if (FW > 0)
            s_cpml_row(c, j, gx, gy, K_x_half, a_x_half, b_x_half, K_y, a_y, b_y, K_y_half, a_y_half, b_y_half,
                       psi_vyy, psi_vyx);
        {
            real *restrict sx = &AT(sxy, c, j, 0), *restrict sz = &AT(syz, c, j, 0);
            const real *restrict fi = &AT(fipjp, c, j, 0), *restrict ff = &AT(f, c, j, 0);
            if (uy) {
                real *restrict u1 = &AT(uy, c, j, 0), *restrict u2 = &AT(uyx, c, j, 0);
                for (i = 1; i <= NX; i++) {
                    u1[i] = gy[i];
                    u2[i] = gx[i];
                }
            }
            real sr[NX + 1], sq[NX + 1];
            for (i = 1; i <= NX; i++) sr[i] = sq[i] = 0.0;
            for (l = 0; l < L; l++) {
                const real *restrict rl = &AT(r + (size_t)l * plane, c, j, 0);
                const real *restrict ql = &AT(q + (size_t)l * plane, c, j, 0);
                for (i = 1; i <= NX; i++) {
                    sr[i] += rl[i];
                    sq[i] += ql[i];
                }
            }
            for (i = 1; i <= NX; i++) {
                sx[i] += (fi[i] * gx[i]) + (dthalbe * sr[i]);
                sz[i] += (ff[i] * gy[i]) + (dthalbe * sq[i]);
                sr[i] = sq[i] = 0.0;
            }
            for (l = 0; l < L; l++) {
                real *restrict rl = &AT(r + (size_t)l * plane, c, j, 0);
                real *restrict ql = &AT(q + (size_t)l * plane, c, j, 0);
                const real *restrict dl = &AT(dip + (size_t)l * plane, c, j, 0);
                const real *restrict el = &AT(d + (size_t)l * plane, c, j, 0);
                const real bi = bip[l], bj = bjm[l], ci = cip[l], cj = cjm[l];
                for (i = 1; i <= NX; i++) {
                    rl[i] = bi * (rl[i] * ci - (dl[i] * gx[i]));
                    ql[i] = bj * (ql[i] * cj - (el[i] * gy[i]));
                    sr[i] += rl[i];
                    sq[i] += ql[i];
                }
            }
            for (i = 1; i <= NX; i++) {
                sx[i] += (dthalbe * sr[i]);
                sz[i] += (dthalbe * sq[i]);
            }
        }
    }
}

/* ========================================================================
 * Composed shot simulation (denise.c:1079-1346 forward, 1452-1651 adjoint,
 * 1665-1694 per-shot preconditioning, 1743-1813 chain rule).
 * ====================================================================== */

typedef struct {
    const orc_cfg *c;
    real hc[7];
    real *pml[20]; /* d,K,alpha,a,b for x, x_half, y, y_half */
    real *vs, *rho, *taus, *uip, *ujp, *tausipjp, *fipjp, *f, *dip, *d;
    real bip[8], bjm[8], cip[8], cjm[8];
    /* wavefield state */
    real *vy, *sxy, *syz, *vyp1, *uy, *uyx, *r, *q;
    real *psi_syy_y, *psi_sxy_x, *psi_vyy, *psi_vyx;
} orc_sim;

#define PML_KX(s) ((s)->pml[1])
#define PML_AX(s) ((s)->pml[3])
#define PML_BX(s) ((s)->pml[4])
#define PML_KXH(s) ((s)->pml[6])
#define PML_AXH(s) ((s)->pml[8])
#define PML_BXH(s) ((s)->pml[9])
#define PML_KY(s) ((s)->pml[11])
#define PML_AY(s) ((s)->pml[13])
#define PML_BY(s) ((s)->pml[14])
#define PML_KYH(s) ((s)->pml[16])
#define PML_AYH(s) ((s)->pml[18])
#define PML_BYH(s) ((s)->pml[19])

static real *zalloc(size_t n) { return (real *)calloc(n ? n : 1, sizeof(real)); }

/* model setup: readmod.c:82-104 (taus = 2/qs), matcopy, av_mue, av_tau,
 * prepare_update_s, PML_pro, holbergcoeff.  vs/rho/qs are [NY][NX] row-major
 * float (qs may be NULL for L == 0). */
static int sim_init(orc_sim *s, const orc_cfg *c, const float *vs, const float *rho, const float *qs)
{
    const size_t P = F(orc_padded_size)(c);
    const int NX = c->nx, NY = c->ny, L = c->L, FW = c->fw;
    int i, j, t;
    memset(s, 0, sizeof(*s));
    s->c = c;
    if (F(orc_holberg)(c->fdorder, c->maxrelerror, s->hc)) return -1;
    for (t = 0; t < 20; t++) s->pml[t] = zalloc((size_t)2 * FW + 2);
    if (FW > 0)
        F(orc_pml_pro)(c, s->pml[0], s->pml[1], s->pml[2], s->pml[3], s->pml[4], s->pml[5], s->pml[6], s->pml[7],
                       s->pml[8], s->pml[9], s->pml[10], s->pml[11], s->pml[12], s->pml[13], s->pml[14],
                       s->pml[15], s->pml[16], s->pml[17], s->pml[18], s->pml[19]);
    s->vs = zalloc(P);
    s->rho = zalloc(P);
    s->taus = zalloc(P);
    s->uip = zalloc(P);
    s->ujp = zalloc(P);
    for (j = 1; j <= NY; j++)
        for (i = 1; i <= NX; i++) {
            size_t k = (size_t)(j - 1) * NX + (size_t)(i - 1);
            AT(s->vs, c, j, i) = vs[k];
            AT(s->rho, c, j, i) = rho[k];
            if (L > 0) AT(s->taus, c, j, i) = 2.0 / qs[k];
        }
    F(orc_matcopy_wrap)(c, s->rho);
    F(orc_matcopy_wrap)(c, s->vs);
    F(orc_matcopy_wrap)(c, s->taus);
    F(orc_av_mue)(c, s->vs, s->rho, s->uip, s->ujp);
    if (L > 0) {
        s->tausipjp = zalloc(P);
        s->fipjp = zalloc(P);
        s->f = zalloc(P);
        s->dip = zalloc(P * (size_t)L);
        s->d = zalloc(P * (size_t)L);
        F(orc_av_tau)(c, s->taus, s->tausipjp);
        F(orc_prepare_update_s)(c, s->ujp, s->uip, s->taus, s->tausipjp, s->fipjp, s->f, s->dip, s->d, s->bip,
                                s->bjm, s->cip, s->cjm);
    }
    s->vy = zalloc(P);
    s->sxy = zalloc(P);
    s->syz = zalloc(P);
    s->vyp1 = zalloc(P);
    s->uy = zalloc(P);
    s->uyx = zalloc(P);
    if (L > 0) {
        s->r = zalloc(P * (size_t)L);
        s->q = zalloc(P * (size_t)L);
    }
    s->psi_syy_y = zalloc((size_t)2 * FW * NX);
    s->psi_vyy = zalloc((size_t)2 * FW * NX);
    s->psi_sxy_x = zalloc((size_t)2 * FW * NY);
    s->psi_vyx = zalloc((size_t)2 * FW * NY);
    return 0;
}

static void sim_free(orc_sim *s)
{
    int t;
    for (t = 0; t < 20; t++) free(s->pml[t]);
    free(s->vs); free(s->rho); free(s->taus); free(s->uip); free(s->ujp);
    free(s->tausipjp); free(s->fipjp); free(s->f); free(s->dip); free(s->d);
    free(s->vy); free(s->sxy); free(s->syz); free(s->vyp1); free(s->uy); free(s->uyx);
    free(s->r); free(s->q);
    free(s->psi_syy_y); free(s->psi_vyy); free(s->psi_sxy_x); free(s->psi_vyx);
}

/* zero_fdveps.c:11-52 / zero_fdveps_visc.c:11-64 */
static void sim_zero(orc_sim *s)
{
    const orc_cfg *c = s->c;
    const size_t P = F(orc_padded_size)(c);
    const int FW = c->fw, L = c->L;
    memset(s->vy, 0, P * sizeof(real));
    memset(s->sxy, 0, P * sizeof(real));
    memset(s->syz, 0, P * sizeof(real));
    memset(s->vyp1, 0, P * sizeof(real));
    memset(s->uy, 0, P * sizeof(real));
    memset(s->uyx, 0, P * sizeof(real));
    if (L > 0) {
        memset(s->r, 0, P * (size_t)L * sizeof(real));
        memset(s->q, 0, P * (size_t)L * sizeof(real));
    }
    memset(s->psi_syy_y, 0, (size_t)2 * FW * c->nx * sizeof(real));
    memset(s->psi_vyy, 0, (size_t)2 * FW * c->nx * sizeof(real));
    memset(s->psi_sxy_x, 0, (size_t)2 * FW * c->ny * sizeof(real));
    memset(s->psi_vyx, 0, (size_t)2 * FW * c->ny * sizeof(real));
}

/* exchange_v.c:99-161 / exchange_s.c:95-157 with BOUNDARY 1 on one
 * domain: columns 1..FDORDER/2 are copied to NX+1..NX+FDORDER/2 and
 * NX-FDORDER/2+1..NX to 1-FDORDER/2..0 (fdo = FDORDER/2 + 1, l < fdo). */
void F(orc_wrap_x)(const orc_cfg *c, real *a)
{
    const int NX = c->nx, NY = c->ny, k = c->fdorder / 2;
    int j, l;
    for (j = 1; j <= NY; j++)
        for (l = 1; l <= k; l++) {
            AT(a, c, j, NX + l) = AT(a, c, j, l);
            AT(a, c, j, 1 - l) = AT(a, c, j, NX - l + 1);
        }
}

static void sim_step(orc_sim *s, int sw, int ninj, const int *ii, const int *jj, const real *val)
{
    const orc_cfg *c = s->c;
    F(orc_update_v)(c, sw, s->vy, s->vyp1, s->syz, s->sxy, s->rho, s->hc, PML_KX(s), PML_AX(s), PML_BX(s),
                    PML_KY(s), PML_AY(s), PML_BY(s), s->psi_syy_y, s->psi_sxy_x);
    F(orc_inject)(c, s->vy, ninj, ii, jj, val);
    if (c->periodic) F(orc_wrap_x)(c, s->vy);   /* exchange_v after update_v_PML (denise.c:975) */
    if (c->L > 0)
        F(orc_update_s_visc)(c, s->vy, s->uy, s->uyx, s->syz, s->sxy, s->r, s->q, s->fipjp, s->f, s->dip, s->d,
                             s->bip, s->bjm, s->cip, s->cjm, s->hc, PML_KXH(s), PML_AXH(s), PML_BXH(s), PML_KY(s),
                             PML_AY(s), PML_BY(s), PML_KYH(s), PML_AYH(s), PML_BYH(s), s->psi_vyy, s->psi_vyx);
    else
        F(orc_update_s_elastic)(c, s->vy, s->uy, s->uyx, s->syz, s->sxy, s->uip, s->ujp, s->hc, PML_KXH(s),
                                PML_AXH(s), PML_BXH(s), PML_KYH(s), PML_AYH(s), PML_BYH(s), s->psi_vyy, s->psi_vyx);
    if (c->periodic) F(orc_wrap_x)(c, s->sxy);  /* exchange_s (denise.c:1009) */
}

/* ------------------------------------------------------------------------
 * BF16 snapshot storage (the product's reduced-precision option; the
 * reference stores FP32 only).  Snapshot k of a field is replaced by
 * bf16(x * 2^-e) * 2^e, with e the frexp exponent of max |x| over snapshot
 * k - 1 of the same field (0 for k = 0 or a zero maximum; clamped to
 * [-100, 100]); bf16 rounds to nearest even on the float bits.  snaps
 * [nsnap][3][nxny] as written by orc_forward.
 * ---------------------------------------------------------------------- */
static float bf16_round(float v)
{
    uint32_t u;
    memcpy(&u, &v, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&v, &u, 4);
    return v;
}

void F(orc_snap_bf16)(int nsnap, size_t nxny, real *snaps)
{
    float prev[3] = {0.0f, 0.0f, 0.0f};
    int k, f;
    size_t t;
    for (k = 0; k < nsnap; k++)
        for (f = 0; f < 3; f++) {
            real *x = snaps + ((size_t)k * 3 + (size_t)f) * nxny;
            float m = 0.0f, enc, dec;
            int e = 0;
            for (t = 0; t < nxny; t++) {
                const float a = fabsf((float)x[t]);
                if (a > m) m = a;
            }
            if (prev[f] > 0.0f && isfinite(prev[f])) {
                (void)frexpf(prev[f], &e);
                e = e < -100 ? -100 : (e > 100 ? 100 : e);
            }
            enc = ldexpf(1.0f, -e);
            dec = ldexpf(1.0f, e);
            for (t = 0; t < nxny; t++) x[t] = (real)(bf16_round((float)x[t] * enc) * dec);
            prev[f] = m;
        }
}

/* number of stored steps, denise.c:287 */
int F(orc_ntdtinv)(int nt, int dtinv) { return (int)ceil((float)nt / (float)dtinv); }

/* ------------------------------------------------------------------------
 * Forward modelling of one shot: denise.c:1125-1346.  Sources are nsrc
 * points (grid indices, 1-based) with signals [nsrc][NT]; receivers are ntr
 * points.  seis_out [ntr][NT] (REAL).  If snaps != NULL it receives the
 * stored snapshots [NTDTINV][3][NY][NX] (order vyp1, uy, uyx) and Ws
 * (padded, accumulated) the source-side energy.  Returns 0, or 1 if the
 * centre-cell NaN probe (denise.c:1183-1185) fired.
 * ---------------------------------------------------------------------- */
int F(orc_forward)(const orc_cfg *c, const float *vs, const float *rho, const float *qs, int nsrc, const int *src_i,
                   const int *src_j, const float *signals, int ntr, const int *rec_i, const int *rec_j,
                   real *seis_out, real *snaps, real *Ws)
{
    orc_sim s;
    const int NT = c->nt, NX = c->nx, NY = c->ny;
    real *val = zalloc((size_t)nsrc);
    int nt, hin1 = 1, kstore = 0, l, i, j, rc = 0;
    if (sim_init(&s, c, vs, rho, qs)) { free(val); return -1; }
    sim_zero(&s);
    for (nt = 1; nt <= NT; nt++) {
        if (isnan(AT(s.vy, c, NY / 2, NX / 2))) { rc = 1; break; }
        for (l = 0; l < nsrc; l++) val[l] = (real)signals[(size_t)l * NT + (size_t)(nt - 1)];
        sim_step(&s, 0, nsrc, src_i, src_j, val);
        for (l = 0; l < ntr; l++) seis_out[(size_t)l * NT + (size_t)(nt - 1)] = AT(s.vy, c, rec_j[l], rec_i[l]);
        if (nt == hin1) {
            if (snaps) {
                real *base = snaps + (size_t)kstore * 3 * NX * NY;
                for (j = 1; j <= NY; j++)
                    for (i = 1; i <= NX; i++) {
                        size_t o = (size_t)(j - 1) * NX + (size_t)(i - 1);
                        base[o] = AT(s.vyp1, c, j, i);
                        base[(size_t)NX * NY + o] = AT(s.uy, c, j, i);
                        base[(size_t)2 * NX * NY + o] = AT(s.uyx, c, j, i);
                    }
            }
            if (Ws) /* eprecond.c:8-14 */
                for (j = 1; j <= NY; j++)
                    for (i = 1; i <= NX; i++) AT(Ws, c, j, i) += AT(s.vy, c, j, i) * AT(s.vy, c, j, i);
            kstore++;
            hin1 += c->dtinv;
        }
    }
    sim_free(&s);
    free(val);
    return rc;
}

/* ------------------------------------------------------------------------
 * calc_res.c:150-257 (LNORM 2, GRAD_FORM 2, RTM 0) and calc_energy.c:53-61.
 * obs is modified in place (obs[i][1] := syn[i][1], Q7) for the traces the
 * residual covers.  diff [ntr][ns] is time-reversed.  l2 and energy are
 * accumulated in double here; the reference accumulates in float (noted
 * where compared).
 * Data conditioning (calc_res.c:37-113, calc_energy.c:23-49): skip_res[i]
 * != 0 (trace killed by TRKILL or muted by OFFSET_MUTE) leaves the trace's
 * residual zero, out of L2 and its obs[1] unchanged; skip_en[i] != 0 (the
 * trace kill as calc_energy reads the kill file) leaves it out of the
 * energy.  NULL masks: no conditioning.
 * ---------------------------------------------------------------------- */
void F(orc_residual_cond)(int ntr, int ns, double dt_in, const real *syn, real *obs, real *diff, double *l2,
                          double *energy, const unsigned char *skip_res, const unsigned char *skip_en)
{
    const real DT = (real)dt_in;
    int i, j, invtime;
    double L2 = 0.0, en = 0.0;
    for (i = 0; i < ntr; i++) {
        const real *sy = syn + (size_t)i * ns;
        real *ob = obs + (size_t)i * ns, *df = diff + (size_t)i * ns;
        real intseis;
        for (j = 0; j < ns; j++) df[j] = 0.0;
        if (!(skip_res && skip_res[i])) {
            invtime = ns;
            for (j = 1; j <= ns; j++) {
                if (j == 1) ob[0] = sy[0];
                intseis = sy[j - 1] - ob[j - 1];
                df[invtime - 1] = intseis;
                invtime--;
            }
            for (j = 0; j < ns; j++) L2 += (double)df[j] * (double)df[j];
        }
        if (!(skip_en && skip_en[i])) {
            intseis = 0.0;
            for (j = 0; j < ns; j++) {
                intseis += ob[j];
                en += (double)(intseis * intseis * DT * DT);
            }
        }
    }
    *l2 += L2;
    *energy += en;
}

void F(orc_residual)(int ntr, int ns, double dt_in, const real *syn, real *obs, real *diff, double *l2,
                     double *energy)
{
    F(orc_residual_cond)(ntr, ns, dt_in, syn, obs, diff, l2, energy, NULL, NULL);
}

/* ------------------------------------------------------------------------
 * Adjoint (back-propagation) of one shot with the zero-lag correlation:
 * denise.c:1452-1651.  Sources are at the receivers with the time-reversed
 * residual diff [ntr][NT].  snaps as written by orc_forward.  grad_u,
 * grad_rho, Wr are padded fields accumulated in place.
 * ---------------------------------------------------------------------- */
int F(orc_adjoint)(const orc_cfg *c, const float *vs, const float *rho, const float *qs, int ntr, const int *rec_i,
                   const int *rec_j, const real *diff, const real *snaps, real *grad_u, real *grad_rho, real *Wr)
{
    orc_sim s;
    const int NT = c->nt, NX = c->nx, NY = c->ny, dtinv = c->dtinv;
    const int NTDTINV = F(orc_ntdtinv)(NT, dtinv);
    real *val = zalloc((size_t)ntr);
    int nt, hin = 1, l, i, j, rc = 0;
    if (sim_init(&s, c, vs, rho, qs)) { free(val); return -1; }
    sim_zero(&s);
    for (nt = 1; nt <= NT; nt++) {
        if (isnan(AT(s.vy, c, NY / 2, NX / 2))) { rc = 1; break; }
        for (l = 0; l < ntr; l++) val[l] = diff[(size_t)l * NT + (size_t)(nt - 1)];
        sim_step(&s, 1, ntr, rec_i, rec_j, val);
        /* DTINV_help[NT-nt+1]: forward step NT-nt+1 was stored */
        if (((NT - nt + 1) - 1) % dtinv == 0) {
            const real *base = snaps + (size_t)(NTDTINV - hin) * 3 * NX * NY;
            for (j = 1; j <= NY; j++)
                for (i = 1; i <= NX; i++) {
                    size_t o = (size_t)(j - 1) * NX + (size_t)(i - 1);
                    AT(grad_rho, c, j, i) += (AT(s.vyp1, c, j, i) * base[o]);
                    AT(grad_u, c, j, i) += (base[(size_t)2 * NX * NY + o] * AT(s.sxy, c, j, i)) +
                                           (base[(size_t)NX * NY + o] * AT(s.syz, c, j, i));
                }
            for (j = 1; j <= NY; j++)
                for (i = 1; i <= NX; i++) AT(Wr, c, j, i) += AT(s.vy, c, j, i) * AT(s.vy, c, j, i);
            hin++;
        }
    }
    sim_free(&s);
    free(val);
    return rc;
}

/* ------------------------------------------------------------------------
 * Pinned (with orc_adjoint) by central differences of the misfit on smooth
 * bumps with EPRECOND = 0: the density correlation equals c*dJ/drho|mu and
 * the shear correlation c*DTF*mu*dJ/dmu (DTF = DT elastic, 1 visco) with one
 * constant c, within 8% / 25%.  eprecond1 and the C_vs/C_rho means are
 * pinned by the ratio of the EPRECOND 0 and 1 gradients: (mean vs / mean
 * rho)^2 between the parameters, the relative floor, invariance under a
 * joint scaling of signals and data, and per-shot weights.
 * Full gradient over nshots (denise.c:746-816, 875-1694, 1743-1813):
 *   C_vs, C_rho = float means of the model (i outer, j inner, Q14);
 *   per shot: forward, residual, adjoint, eprecond1 (Q16), scaling;
 *   chain rule to vs and rho (INVMAT1 = 1).
 * Shots: src_i/src_j [nshots] (one source each), signals [nshots][NT];
 * obs [nshots][ntr][NT] float (copied, then modified as calc_res does).
 * Outputs (all [NY][NX] row-major REAL): grad_vs, grad_rho; optional
 * We_last (the last shot's energy weights).  l2/energy in double.
 * ---------------------------------------------------------------------- */
static int gradient_impl(const orc_cfg *c, const float *vs, const float *rho, const float *qs, int nshots,
                         const int *src_i, const int *src_j, const float *signals, int ntr, const int *rec_i,
                         const int *rec_j, const float *obs, const real *obs_real, real *grad_vs, real *grad_rho,
                         double *l2, double *energy, real *waveconv_u_out, real *waveconv_rho_out);

int F(orc_gradient)(const orc_cfg *c, const float *vs, const float *rho, const float *qs, int nshots,
                    const int *src_i, const int *src_j, const float *signals, int ntr, const int *rec_i,
                    const int *rec_j, const float *obs, real *grad_vs, real *grad_rho, double *l2, double *energy,
                    real *waveconv_u_out, real *waveconv_rho_out)
{
    return gradient_impl(c, vs, rho, qs, nshots, src_i, src_j, signals, ntr, rec_i, rec_j, obs, NULL, grad_vs,
                         grad_rho, l2, energy, waveconv_u_out, waveconv_rho_out);
}

/* The same with the observed data in the oracle's own precision (the FP64
 * oracle's whole chain in double, observed data from orc_forward_f64). */
int F(orc_gradient_obsreal)(const orc_cfg *c, const float *vs, const float *rho, const float *qs, int nshots,
                            const int *src_i, const int *src_j, const float *signals, int ntr, const int *rec_i,
                            const int *rec_j, const real *obs, real *grad_vs, real *grad_rho, double *l2,
                            double *energy, real *waveconv_u_out, real *waveconv_rho_out)
{
    return gradient_impl(c, vs, rho, qs, nshots, src_i, src_j, signals, ntr, rec_i, rec_j, NULL, obs, grad_vs,
                         grad_rho, l2, energy, waveconv_u_out, waveconv_rho_out);
}

static int gradient_impl(const orc_cfg *c, const float *vs, const float *rho, const float *qs, int nshots,
                         const int *src_i, const int *src_j, const float *signals, int ntr, const int *rec_i,
                         const int *rec_j, const float *obs, const real *obs_real, real *grad_vs, real *grad_rho,
                         double *l2, double *energy, real *waveconv_u_out, real *waveconv_rho_out)
{
    const int NT = c->nt, NX = c->nx, NY = c->ny;
    const size_t P = F(orc_padded_size)(c);
    const int NTDTINV = F(orc_ntdtinv)(NT, c->dtinv);
    real *snaps = zalloc((size_t)NTDTINV * 3 * NX * NY);
    real *seis = zalloc((size_t)ntr * NT), *obsr = zalloc((size_t)ntr * NT), *diff = zalloc((size_t)ntr * NT);
    real *Ws = zalloc(P), *Wr = zalloc(P), *We = zalloc(P);
    real *gu_shot = zalloc(P), *grho_shot = zalloc(P), *gu = zalloc(P), *grho = zalloc(P);
    real C_vs, C_rho, Vs_avg = 0.0, rho_avg = 0.0, maxWe, EPSILON_WE = 0.005, DT = (real)c->dt;
    int Lcount = 1, ishot, i, j, rc = 0;
    size_t t;

    for (i = 1; i <= NX; i++)
        for (j = 1; j <= NY; j++) {
            Vs_avg += (real)vs[(size_t)(j - 1) * NX + (i - 1)];
            rho_avg += (real)rho[(size_t)(j - 1) * NX + (i - 1)];
            Lcount++;
        }
    Lcount--;
    Vs_avg /= Lcount;
    rho_avg /= Lcount;
    C_vs = Vs_avg;
    C_rho = rho_avg;
    if (c->c_vs > 0.0) C_vs = (real)c->c_vs;
    if (c->c_rho > 0.0) C_rho = (real)c->c_rho;
    *l2 = 0.0;
    *energy = 0.0;

    for (ishot = 0; ishot < nshots; ishot++) {
        memset(Ws, 0, P * sizeof(real));
        memset(Wr, 0, P * sizeof(real));
        memset(gu_shot, 0, P * sizeof(real));
        memset(grho_shot, 0, P * sizeof(real));
        rc = F(orc_forward)(c, vs, rho, qs, 1, &src_i[ishot], &src_j[ishot], signals + (size_t)ishot * NT, ntr,
                            rec_i, rec_j, seis, snaps, Ws);
        if (rc) break;
        if (c->snap_bf16) F(orc_snap_bf16)(NTDTINV, (size_t)NX * NY, snaps);
        for (t = 0; t < (size_t)ntr * NT; t++)
            obsr[t] = obs ? (real)obs[(size_t)ishot * ntr * NT + t] : obs_real[(size_t)ishot * ntr * NT + t];
        F(orc_residual_cond)(ntr, NT, c->dt, seis, obsr, diff, l2, energy,
                             c->skip_res ? c->skip_res + (size_t)ishot * ntr : NULL,
                             c->skip_en ? c->skip_en + (size_t)ishot * ntr : NULL);
        rc = F(orc_adjoint)(c, vs, rho, qs, ntr, rec_i, rec_j, diff, snaps, gu_shot, grho_shot, Wr);
        if (rc) break;
        if (!c->eprecond) { /* EPRECOND 0: the shot gradients are summed unscaled (denise.c:1665) */
            for (i = 1; i <= NX; i++)
                for (j = 1; j <= NY; j++) {
                    AT(grho, c, j, i) += AT(grho_shot, c, j, i);
                    AT(gu, c, j, i) += AT(gu_shot, c, j, i);
                }
            continue;
        }
        /* eprecond1.c:15-70: EPRECOND 1 sqrt(Ws Wr); EPRECOND 3 sqrt(Ws)
         * times asinh((xmax-x)/y) - asinh((xmin-x)/y) (eprecond1.c:34-50,
         * float coordinates x = i DH, y = j DH, xmin = DH, xmax = NX DH) */
        maxWe = 0.0;
        for (i = 1; i <= NX; i++)
            for (j = 1; j <= NY; j++) {
                if (c->eprecond == 3) {
                    const float DHf = (float)c->dh, x = (float)i * DHf, y = (float)j * DHf;
                    const float xmin = DHf, xmax = (float)NX * DHf;
                    AT(We, c, j, i) = (real)(sqrt((double)AT(Ws, c, j, i)) *
                                             (asinh((double)((xmax - x) / y)) - asinh((double)((xmin - x) / y))));
                } else {
                    AT(We, c, j, i) = sqrt(AT(Ws, c, j, i) * AT(Wr, c, j, i));
                }
                if (AT(We, c, j, i) > maxWe) maxWe = AT(We, c, j, i);
            }
        for (i = 1; i <= NX; i++)
            for (j = 1; j <= NY; j++) AT(We, c, j, i) = AT(We, c, j, i) + (EPSILON_WE * maxWe);
        /* denise.c:1671-1694 */
        for (i = 1; i <= NX; i++)
            for (j = 1; j <= NY; j++) {
                AT(gu_shot, c, j, i) = AT(gu_shot, c, j, i) / (AT(We, c, j, i) * C_vs * C_vs);
                AT(grho_shot, c, j, i) = AT(grho_shot, c, j, i) / (AT(We, c, j, i) * C_rho * C_rho);
                AT(grho, c, j, i) += AT(grho_shot, c, j, i);
                AT(gu, c, j, i) += AT(gu_shot, c, j, i);
            }
    }
    /* denise.c:1756-1811, INVMAT 0, INVMAT1 1 */
    for (j = 1; j <= NY; j++)
        for (i = 1; i <= NX; i++) {
            size_t o = (size_t)(j - 1) * NX + (size_t)(i - 1);
            real pu = (real)vs[o], pr = (real)rho[o];
            real gmu = -DT * AT(gu, c, j, i);
            real grs = -DT * AT(grho, c, j, i);
            if (waveconv_u_out) waveconv_u_out[o] = AT(gu, c, j, i);
            if (waveconv_rho_out) waveconv_rho_out[o] = AT(grho, c, j, i);
            grad_vs[o] = 2.0 * pr * pu * gmu;
            grad_rho[o] = (pu * pu * gmu) + grs;
        }
    free(snaps); free(seis); free(obsr); free(diff); free(Ws); free(Wr); free(We);
    free(gu_shot); free(grho_shot); free(gu); free(grho);
    return rc;
}

/* ------------------------------------------------------------------------
 * Pseudo-Hessian diagonal, HESSIAN = 1: hess_save.c:2-17 stores uy,
 * 0.5*uyx and vyp1 at every stored forward step; hessian.c:91-94 sums their
 * squares over stored steps 1..NTDTINV-1 (the last stored step is left
 * out) into one accumulator that runs on across shots; hessian_out.c:25-36
 * applies the chain rule to vs and rho (INVMAT1 = 1, rho first).  Pinned
 * against sums of squares of orc_forward's snapshots in the tests.
 * Outputs [NY][NX].
 * ---------------------------------------------------------------------- */
int F(orc_pseudo_hessian)(const orc_cfg *c, const float *vs, const float *rho, const float *qs, int nshots,
                          const int *src_i, const int *src_j, const float *signals, real *hess_vs, real *hess_rho)
{
    const int NT = c->nt, NX = c->nx, NY = c->ny;
    const int NTDTINV = F(orc_ntdtinv)(NT, c->dtinv);
    const size_t nxny = (size_t)NX * NY;
    real *snaps = zalloc((size_t)NTDTINV * 3 * nxny);
    real *hu = zalloc(nxny), *hr = zalloc(nxny);
    int ishot, k, rc = 0;
    size_t o;
    for (ishot = 0; ishot < nshots && rc == 0; ishot++) {
        rc = F(orc_forward)(c, vs, rho, qs, 1, &src_i[ishot], &src_j[ishot], signals + (size_t)ishot * NT, 0, NULL,
                            NULL, NULL, snaps, NULL);
        for (k = 0; k < NTDTINV - 1; k++) {
            const real *base = snaps + (size_t)k * 3 * nxny;
            for (o = 0; o < nxny; o++) {
                const real eyy = base[nxny + o];
                const real eyx = 0.5 * base[2 * nxny + o];
                const real utty = base[o];
                hu[o] += (eyy * eyy) + (eyx * eyx);
                hr[o] += (utty * utty);
            }
        }
    }
    for (o = 0; o < nxny; o++) {
        const real u = (real)vs[o], r = (real)rho[o];
        hr[o] += u * u * u * u * hu[o];
        hu[o] = 4.0 * r * r * u * u * hu[o];
        hess_vs[o] = hu[o];
        hess_rho[o] = hr[o];
    }
    free(snaps); free(hu); free(hr);
    return rc;
}

/* ========================================================================
 * Optimiser pieces (host-side in the reference), restated for the tests.
 * ====================================================================== */

/* solvelin.c:12-61 (Gauss elimination without pivoting, n = 3) and
 * calc_opt_step.c:13-109.  L2t[1..3], epst[1..3] passed 0-based [0..2]. */
double F(orc_calc_opt_step)(const real *L2t, const real *epst)
{
    real A[4][4], b[4], x[4], a, cc, opteps;
    int i, k, n, m;
    const int N = 3;
    for (i = 1; i <= N; i++) {
        A[i][3] = (epst[i - 1] * epst[i - 1]);
        A[i][2] = (epst[i - 1]);
        A[i][1] = (1.0);
        b[i] = L2t[i - 1];
    }
    for (k = 1; k <= N - 1; k++)
        for (n = k; n <= N - 1; n++) {
            a = A[n + 1][k] / A[k][k];
            for (m = 1; m <= N; m++) A[n + 1][m] = A[n + 1][m] - a * A[k][m];
            b[n + 1] = b[n + 1] - a * b[k];
        }
    for (k = N; k >= 1; k--) {
        cc = b[k];
        for (m = N; m >= k + 1; m--) cc = cc - A[k][m] * x[m];
        x[k] = cc / A[k][k];
    }
    opteps = -x[2] / (2.0 * x[3]);
    if (((2.0 * x[3]) < 0.0) && (L2t[1] > L2t[2])) opteps = epst[2];
    if (opteps > epst[2]) opteps = epst[2];
    if (opteps < 0.0) opteps = epst[0];
    return (double)opteps;
}

/* LBFGS.c:26-95 two-loop recursion in slot order (Q18).  y, s are
 * [nslots][n]; q (in, destroyed) and r (out) are [n]; cur is the 1-based
 * slot holding the newest pair.  Dots in REAL (float in the reference). */
void F(orc_lbfgs)(int nslots, int n, const real *y, const real *s, real *q, real *r, int cur)
{
    real rho[64], alpha[64], gamma, sum_nom, sum_denom, beta;
    int k, i;
    sum_nom = sum_denom = 0.0;
    for (i = 0; i < n; i++) sum_nom += y[(size_t)(cur - 1) * n + i] * s[(size_t)(cur - 1) * n + i];
    for (i = 0; i < n; i++) sum_denom += y[(size_t)(cur - 1) * n + i] * y[(size_t)(cur - 1) * n + i];
    gamma = sum_nom / sum_denom;
    for (k = 1; k <= nslots; k++) {
        sum_nom = 0.0;
        for (i = 0; i < n; i++) sum_nom += y[(size_t)(k - 1) * n + i] * s[(size_t)(k - 1) * n + i];
        rho[k] = (fabs(sum_nom) > 0.0) ? (real)(1.0 / sum_nom) : (real)0.0;
    }
    for (k = nslots; k >= 1; k--) {
        sum_nom = 0.0;
        for (i = 0; i < n; i++) sum_nom += s[(size_t)(k - 1) * n + i] * q[i];
        alpha[k] = rho[k] * sum_nom;
        for (i = 0; i < n; i++) q[i] = q[i] - alpha[k] * y[(size_t)(k - 1) * n + i];
    }
    for (i = 0; i < n; i++) r[i] = gamma * q[i];
    for (k = 1; k <= nslots; k++) {
        sum_nom = 0.0;
        for (i = 0; i < n; i++) sum_nom += y[(size_t)(k - 1) * n + i] * r[i];
        beta = rho[k] * sum_nom;
        for (i = 0; i < n; i++) r[i] = r[i] + s[(size_t)(k - 1) * n + i] * (alpha[k] - beta);
    }
}

/* PCG.c:24-101 (Polak-Ribiere, PCG_BETA 2) for nclass classes of length
 * n each: dir = new + beta*dir with beta reset to 0 when negative. */
void F(orc_pcg)(int nclass, int n, const real *gnew, const real *gold, real *dir)
{
    int k, i;
    for (k = 0; k < nclass; k++) {
        real betaz = 0.0, betan = 0.0, beta = 0.0;
        const real *pn = gnew + (size_t)k * n, *po = gold + (size_t)k * n;
        real *pd = dir + (size_t)k * n;
        for (i = 0; i < n; i++) {
            betaz += pn[i] * (pn[i] - po[i]);
            betan += po[i] * po[i];
        }
        if (betan != 0.0f) beta = betaz / betan;
        if (beta < 0.0) beta = 0.0;
        for (i = 0; i < n; i++) pd[i] = pn[i] + beta * pd[i];
    }
}

/* calc_mat_change_test.c:41-191 for INVMAT1 == 1 (Q17): returns
 * EPSILON_u; writes unp1/rhonp1 ([NY][NX]).  iter/INV_*_ITER gates apply. */
double F(orc_calc_mat_change)(int nx, int ny, const real *gu, const real *grho, const real *u, const real *rho,
                              real *unp1, real *rhonp1, double eps_scale_in, int iter, int inv_vs_iter,
                              int inv_rho_iter, double scalerho_in, double vsl, double vsu, double rhol,
                              double rhou, real *eps_rho_out)
{
    const real eps_scale = (real)eps_scale_in, SCALERHO = (real)scalerho_in;
    const real VSLOWERLIM = (real)vsl, VSUPPERLIM = (real)vsu, RHOLOWERLIM = (real)rhol, RHOUPPERLIM = (real)rhou;
    real umax = 0.0, gradmax_u = 0.0, rhomax = 0.0, gradmax_rho = 0.0, EPSILON_u, EPSILON_rho;
    int i, j;
#define E(a, j, i) ((a)[(size_t)((j)-1) * nx + (size_t)((i)-1)])
    for (i = 1; i <= nx; i++)
        for (j = 1; j <= ny; j++) {
            if (E(u, j, i) > umax) umax = E(u, j, i);
            if ((i * j == 1) || (gradmax_u == 0.0))
                gradmax_u = fabs(E(gu, j, i));
            else if (fabs(E(gu, j, i)) > gradmax_u)
                gradmax_u = fabs(E(gu, j, i));
        }
    for (i = 1; i <= nx; i++)
        for (j = 1; j <= ny; j++) {
            if (E(rho, j, i) > rhomax) rhomax = E(rho, j, i);
            if ((i * j == 1) || (gradmax_rho == 0.0))
                gradmax_rho = fabs(E(grho, j, i));
            else if (fabs(E(grho, j, i)) > gradmax_rho)
                gradmax_rho = fabs(E(grho, j, i));
        }
    EPSILON_u = eps_scale * (umax / gradmax_u);
    if (iter < inv_vs_iter) EPSILON_u = 0.0;
    EPSILON_rho = eps_scale * (rhomax / gradmax_rho) * SCALERHO;
    if (iter < inv_rho_iter) EPSILON_rho = 0.0;
    for (i = 1; i <= nx; i++)
        for (j = 1; j <= ny; j++) {
            E(unp1, j, i) = E(u, j, i) - EPSILON_u * E(gu, j, i);
            E(rhonp1, j, i) = E(rho, j, i) - EPSILON_rho * E(grho, j, i);
            if ((E(unp1, j, i) < VSLOWERLIM) && (E(unp1, j, i) > 1e-6)) E(unp1, j, i) = E(u, j, i);
            if (E(unp1, j, i) > VSUPPERLIM) E(unp1, j, i) = E(u, j, i);
            if (E(rhonp1, j, i) < RHOLOWERLIM) E(rhonp1, j, i) = E(rho, j, i);
            if (E(rhonp1, j, i) > RHOUPPERLIM) E(rhonp1, j, i) = E(rho, j, i);
            if (inv_vs_iter > iter) E(unp1, j, i) = E(u, j, i);
            if (inv_rho_iter > iter) E(rhonp1, j, i) = E(rho, j, i);
            if (E(unp1, j, i) < 0.0) E(unp1, j, i) = E(u, j, i);
            if (E(rhonp1, j, i) < 0.0) E(rhonp1, j, i) = E(rho, j, i);
        }
#undef E
    if (eps_rho_out) *eps_rho_out = EPSILON_rho;
    return (double)EPSILON_u;
}

/* ------------------------------------------------------------------------
 * Gradient post-processing after the chain rule (denise.c:1818-1847), one
 * domain (NXG = NX, NYG = NY, POS = {0, 0}).  Arrays are [NY][NX] row-major
 * (x fastest), index (j-1)*NX + (i-1) for the reference's [j][i].
 * ---------------------------------------------------------------------- */

/* util.c:38-72 minimum_m / maximum_m: first element, then all others. */
static float orc_min_m(const float *m, size_t n)
{
    float v = m[0];
    for (size_t t = 1; t < n; ++t)
        if (m[t] < v) v = m[t];
    return v;
}
static float orc_max_m(const float *m, size_t n)
{
    float v = m[0];
    for (size_t t = 1; t < n; ++t)
        if (m[t] > v) v = m[t];
    return v;
}

/* gauss_filt_var.c:10-215 (SPATFILTER 4 of smooth_grad.c:31-33): every cell
 * is replaced by the normalised sum of a Gaussian of the gradient around it,
 * whose half widths follow the local wavelength lam = vel / FC_END:
 *   FILT_SIZE_X = round(WD_DAMP1 lam / DH), FILT_SIZE_Z = round(WD_DAMP lam / DH)
 *   (made odd, at least 6), hfs = FILT_SIZE / 2, sigma = hfs / 2 (integer
 *   division), s = 2 sigma^2, kernel(ii, jj) = exp(-ii^2/sx - jj^2/sz),
 *   ii in [-hfsx, hfsx] (x, outer loop), jj in [-hfsz, hfsz] (y, inner loop);
 * cells with FILT_SIZE_Z <= 0 are copied.  Outside the grid the gradient is
 * continued with its edge values (gauss_filt_var.c:126-157), i.e. indices
 * are clamped.  The float build keeps the reference's float operands, the
 * double exp() rounded to float, and the sequential float sums.
 * vel: the current vs model (pu at denise.c:1818).
 * Pins (tests/test_oracle_pins.py): a constant gradient is returned
 * unchanged; for uniform velocity the result equals scipy.ndimage.correlate
 * with the normalised kernel built from the closed form (mode "nearest", x
 * and y half widths different so a transposed kernel fails); for a two-layer
 * model, cells farther than their half width from the interface equal the
 * uniform result of their own layer; FILT_SIZE_Z <= 0 copies. */
void F(orc_gauss_filt_var)(int nx, int ny, double dh_in, double wd_damp_in, double wd_damp1_in, double fc_end_in,
                           const float *vel, real *grad)
{
    const real DH = (real)dh_in, WD_DAMP = (real)wd_damp_in, WD_DAMP1 = (real)wd_damp1_in, FC_END = (real)fc_end_in;
    const size_t n = (size_t)nx * (size_t)ny;
    real *tmp = (real *)malloc((n ? n : 1) * sizeof(real));
    memcpy(tmp, grad, n * sizeof(real));
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            const size_t t = (size_t)(j - 1) * nx + (size_t)(i - 1);
            real lam = (real)vel[t] / FC_END;
            int fx = (int)round((WD_DAMP1 * lam) / DH);
            int fz = (int)round((WD_DAMP * lam) / DH);
            if (fz > 0) {
                if (!(fx % 2)) fx += 1;
                if (!(fz % 2)) fz += 1;
                if (fx < 6) fx = 6;
                if (fz < 6) fz = 6;
                const int hx = abs(fx) / 2, hz = abs(fz) / 2;
                const real sigx = hx / 2, sigz = hz / 2;
                const real sx = 2.0 * sigx * sigx, sz = 2.0 * sigz * sigz;
                real conv = 0.0, sum = 0.0;
                for (int ii = -hx; ii <= hx; ii++)
                    for (int jj = -hz; jj <= hz; jj++) {
                        const real kern = exp(-((ii * ii) / sx) - ((jj * jj) / sz));
                        int y = j + jj, x = i + ii;
                        y = y < 1 ? 1 : (y > ny ? ny : y);
                        x = x < 1 ? 1 : (x > nx ? nx : x);
                        conv += tmp[(size_t)(y - 1) * nx + (size_t)(x - 1)] * kern;
                        sum += kern;
                    }
                grad[t] = conv / sum;
            } else {
                grad[t] = tmp[t];
            }
        }
    free(tmp);
}

/* tripd (util.c:638-671): symmetric tridiagonal solve, d diagonal, e
 * superdiagonal, b right-hand side overwritten with the solution. */
static void orc_tripd(float *d, float *e, float *b, int n)
{
    int k;
    float temp;
    for (k = 1; k < n; ++k) {
        temp = e[k - 1];
        e[k - 1] = temp / d[k - 1];
        d[k] -= temp * e[k - 1];
    }
    for (k = 1; k < n; ++k) b[k] -= e[k - 1] * b[k - 1];
    b[n - 1] /= d[n - 1];
    for (k = n - 1; k > 0; --k) b[k - 1] = b[k - 1] / d[k - 1] - e[k - 1] * b[k];
}

/* smooth2.c:28-165 (SPATFILTER 2) on a gradient [ny][nx] (always float,
 * as the reference), r1 = r2 = WD_DAMP^2 / 4.  The reference loads the
 * gradient into v[1..n2][1..n1] of a zeroed matrix v[0..n2][0..n1] (n2 = NX,
 * n1 = NY) and smooths over 0..n-1: the zero row and column 0 take part, and
 * the last row and column (x = NX, y = NY) keep their values.  The window
 * covers the whole grid, so its smoothing (rw = 0) is skipped.  Needs
 * nx >= 2, ny >= 3 (returns -1 otherwise).
 * Pins (tests/test_oracle_smooth2.py): the last row and column are kept; the
 * two tridiagonal operators built from the stated structure, applied to the
 * result in reverse order, give back the input (float rounding); the result
 * is linear in the gradient. */
int F(orc_smooth2)(int nx, int ny, double wd_damp_in, float *grad)
{
    const int n1 = ny, n2 = nx;
    if (n2 < 2 || n1 < 3) return -1;
    float r1 = (float)wd_damp_in, r2;
    r2 = r1;
    r1 = (float)(r1 * r1 * 0.25);
    r2 = (float)(r2 * r2 * 0.25);
    const int nmax = (n1 < n2) ? n2 : n1;
    const size_t ld = (size_t)n1 + 1;
    float *v = (float *)calloc(((size_t)n2 + 1) * ld, sizeof(float));
    float *w = (float *)calloc(((size_t)n2 + 1) * ld, sizeof(float));
    float *d = (float *)calloc((size_t)nmax + 1, sizeof(float));
    float *e = (float *)calloc((size_t)nmax + 1, sizeof(float));
    float *f = (float *)calloc((size_t)nmax + 1, sizeof(float));
#define V(ix, iz) v[(size_t)(ix) * ld + (size_t)(iz)]
#define W(ix, iz) w[(size_t)(ix) * ld + (size_t)(iz)]
    for (int i = 1; i <= n2; i++)
        for (int j = 1; j <= n1; j++) V(i, j) = grad[(size_t)(j - 1) * nx + (size_t)(i - 1)];
    for (int ix = 0; ix < n2; ++ix)
        for (int iz = 0; iz < n1; ++iz) W(ix, iz) = 1;
    for (int iz = 0; iz < n1; ++iz) {
        for (int ix = 0; ix < n2 - 1; ++ix) {
            d[ix] = (float)(1.0 + r2 * (W(ix, iz) + W(ix + 1, iz)));
            e[ix] = -r2 * W(ix + 1, iz);
            f[ix] = V(ix, iz);
        }
        d[0] -= r2 * W(0, iz);
        d[n2 - 1] = (float)(1.0 + r2 * W(n2 - 1, iz));
        f[n2 - 1] = V(n2 - 1, iz);
        orc_tripd(d, e, f, n2);
        for (int ix = 0; ix < n2; ++ix) V(ix, iz) = f[ix];
    }
    for (int ix = 0; ix < n2; ++ix) {
        for (int iz = 0; iz < n1 - 2; ++iz) {
            d[iz] = (float)(1.0 + r1 * (W(ix, iz + 1) + W(ix, iz + 2)));
            e[iz] = -r1 * W(ix, iz + 2);
            f[iz] = V(ix, iz + 1);
        }
        f[0] += r1 * W(ix, 1) * V(ix, 0);
        d[n1 - 2] = (float)(1.0 + r1 * W(ix, n1 - 1));
        f[n1 - 2] = V(ix, n1 - 1);
        orc_tripd(d, e, f, n1 - 1);
        for (int iz = 0; iz < n1 - 1; ++iz) V(ix, iz + 1) = f[iz];
    }
    for (int i = 1; i <= n2; i++)
        for (int j = 1; j <= n1; j++) grad[(size_t)(j - 1) * nx + (size_t)(i - 1)] = V(i, j);
#undef V
#undef W
    free(v);
    free(w);
    free(d);
    free(e);
    free(f);
    return 0;
}

/* gauss_filt.c:56-160 (SPATFILTER 3): one Gaussian for the whole grid.
 * lam = C_vs / FC_END; FILT_SIZE_GRAD = round(WD_DAMP lam / DH), made odd;
 * 0 leaves the gradient as it is; hfs = |FILT_SIZE_GRAD| / 2, sigma = hfs / 2
 * in integer arithmetic, s = 2 sigma^2; weights exp(-ii^2/s - jj^2/s) for
 * ii (x, outer) and jj (y, inner) in [-hfs, hfs], summed in that order from
 * zero (the reference leaves `sum` uninitialised: this is its result when it
 * starts at zero) and divided by the sum; the gradient, continued with its
 * edge values (clamped indices, gauss_filt.c:96-120), convolved with them in
 * the same order.  Returns 0, 1 when FILT_SIZE_GRAD is 0 (no filtering), -1
 * for a negative size (the reference then skips the normalisation) or
 * hfs < 2 (sigma 0: the reference computes 0/0).
 * Pins (tests/test_oracle_gauss_filt.py): a delta returns the weights, equal to the closed-form normalised Gaussian
 * of the integer sigma; a constant is kept; scipy.ndimage.correlate with
 * that kernel (mode "nearest") matches, and the kernel of a real-valued
 * sigma = hfs / 2 does not. */
int F(orc_gauss_filt)(int nx, int ny, double dh_in, double wd_damp_in, double fc_end_in, double c_vs_in, real *grad)
{
    const float DH = (float)dh_in, WD_DAMP = (float)wd_damp_in, FC_END = (float)fc_end_in, C_vs = (float)c_vs_in;
    const float lam = C_vs / FC_END;
    int fs = (int)round((WD_DAMP * lam) / DH);
    if (fs == 0) return 1;
    if (fs < 0) return -1;
    if (!(fs % 2)) fs += 1;
    const int hfs = abs(fs) / 2;
    if (hfs < 2) return -1;
    const float sigma = (float)(hfs / 2);
    const float s = (float)(2.0 * sigma * sigma);
    const int w = 2 * hfs + 1;
    float *kern = (float *)malloc((size_t)w * w * sizeof(float));
    float sum = 0.0f;
    for (int ii = -hfs; ii <= hfs; ii++)
        for (int jj = -hfs; jj <= hfs; jj++) {
            const float k = (float)exp((double)(-((float)(ii * ii) / s) - ((float)(jj * jj) / s)));
            kern[(ii + hfs) * w + (jj + hfs)] = k;
            sum += k;
        }
    for (int t = 0; t < w * w; t++) kern[t] /= sum;
    const size_t n = (size_t)nx * (size_t)ny;
    real *tmp = (real *)malloc((n ? n : 1) * sizeof(real));
    memcpy(tmp, grad, n * sizeof(real));
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            real conv = 0.0;
            for (int ii = -hfs; ii <= hfs; ii++)
                for (int jj = -hfs; jj <= hfs; jj++) {
                    int y = j + jj, x = i + ii;
                    y = y < 1 ? 1 : (y > ny ? ny : y);
                    x = x < 1 ? 1 : (x > nx ? nx : x);
                    conv += tmp[(size_t)(y - 1) * nx + (size_t)(x - 1)] * (real)kern[(ii + hfs) * w + (jj + hfs)];
                }
            grad[(size_t)(j - 1) * nx + (size_t)(i - 1)] = conv;
        }
    free(tmp);
    free(kern);
    return 0;
}

/* taper_grad.c:11-584 for sws 1 (vertical), 2 (horizontal) and 3 (taper
 * around the sources); the taper coefficients go to taper [NY][NX], the
 * gradient is multiplied by them and then zeroed where the reference zeroes
 * it.  Sources and receivers are grid indices (iround(srcpos / DH) and
 * recpos).  gradt = GRADT1..GRADT4.
 *   sws 1: window over columns GRADT1 < i < GRADT4, Gaussian flanks of
 *          lengths GRADT2-GRADT1 and GRADT4-GRADT3 (a = 3), zero elsewhere;
 *          source and receiver cells zeroed (taper_grad.c:50-160).
 *   sws 2: zero above row GRADT1, Gaussian ramp to row GRADT2, j*DH below
 *          (EXP_TAPER_GRAD_HOR = 1) (taper_grad.c:166-290).
 *   sws 3: erf (SRTSHAPE 1) or log (2) bowl of diameter 2*SRTRADIUS around
 *          every source, minimum over the sources, shifted and normalised to
 *          [0, 1]; the (2 FILTSIZE + 1)^2 box around each source zeroed
 *          (taper_grad.c:295-520).
 * Pins: the closed forms of each window and bowl in float64
 * (tests/test_oracle_pins.py). Returns 0, or -1 for an unknown sws. */
int F(orc_taper_grad)(int nx, int ny, double dh_in, int sws, const int *gradt, int srtshape, double srtradius_in,
                      int filtsize, int nsrc, const int *src_i, const int *src_j, int nrec, const int *rec_i,
                      const int *rec_j, real *grad, float *taper)
{
    const float DH = (float)dh_in, SRTRADIUS = (float)srtradius_in;
    const int GRADT1 = gradt[0], GRADT2 = gradt[1], GRADT3 = gradt[2], GRADT4 = gradt[3];
    const size_t n = (size_t)nx * (size_t)ny;
#define T(j, i) taper[(size_t)((j)-1) * nx + (size_t)((i)-1)]
#define G(j, i) grad[(size_t)((j)-1) * nx + (size_t)((i)-1)]
    if (sws == 1) {
        const int tl = GRADT2 - GRADT1, tl2 = GRADT4 - GRADT3, ifw = GRADT4 - GRADT1;
        float *win = (float *)calloc((size_t)(ifw > 0 ? ifw : 0) + 2, sizeof(float));
        const float a = 3;
        for (int i = 1; i <= ifw; i++) {
            win[i] = 1.0;
            if (i <= tl) win[i] = exp(-(1.0 / 2.0) * (a * (i - tl) / (tl / 2.0)) * (a * (i - tl) / (tl / 2.0)));
            if (i >= (ifw - tl2))
                win[i] = exp(-(1.0 / 2.0) * (a * (i - (ifw - tl2)) / (tl2 / 2.0)) * (a * (i - (ifw - tl2)) / (tl2 / 2.0)));
        }
        for (int j = 1; j <= ny; j++) {
            int h = 1;
            for (int i = 1; i <= nx; i++) {
                float tap = 0.0;
                if ((i > GRADT1) && (i < GRADT4)) tap = win[h++];
                T(j, i) = tap;
            }
        }
        free(win);
        for (size_t t = 0; t < n; t++) grad[t] *= taper[t];
        for (int s = 0; s < nsrc; s++)
            if (src_i[s] >= 1 && src_i[s] <= nx && src_j[s] >= 1 && src_j[s] <= ny) G(src_j[s], src_i[s]) = 0.0;
        for (int r = 0; r < nrec; r++)
            if (rec_i[r] >= 1 && rec_i[r] <= nx && rec_j[r] >= 1 && rec_j[r] <= ny) G(rec_j[r], rec_i[r]) = 0.0;
        return 0;
    }
    if (sws == 2) {
        const int tl = GRADT2 - GRADT1, ifw = GRADT2 - GRADT1 + 1;
        float *win = (float *)calloc((size_t)(ifw > 0 ? ifw : 0) + 2, sizeof(float));
        const float a = 3;
        for (int i = 1; i <= ifw; i++) {
            win[i] = 1.0;
            if (i <= tl) win[i] = exp(-(1.0 / 2.0) * (a * (i - tl) / (tl / 2.0)) * (a * (i - tl) / (tl / 2.0)));
        }
        int h = 1;
        for (int j = 1; j <= ny; j++) {
            for (int i = 1; i <= nx; i++) {
                float tap = 0.0;
                if ((j >= GRADT1) && (j <= GRADT2)) tap = win[h];
                if (j > GRADT2) tap = pow((float)(j * DH), 1.0);
                T(j, i) = tap;
            }
            if (j >= GRADT1) h++;
        }
        free(win);
        for (size_t t = 0; t < n; t++) grad[t] *= taper[t];
        return 0;
    }
    if (sws == 3) {
        int ng = (int)(ceil(2.0 * SRTRADIUS / DH));
        if (ng < 5) ng = 5;
        float *m = (float *)malloc((size_t)ng * ng * sizeof(float));
        float *msum = (float *)malloc(n * sizeof(float));
#define M(j, i) m[(size_t)((j)-1) * ng + (size_t)((i)-1)]
        for (size_t t = 0; t < n; t++) msum[t] = 1.0;
        const float a = 1.0;
        const float maxrad = sqrt(2.0 * SRTRADIUS * SRTRADIUS);
        for (int j = 1; j <= ng; j++)
            for (int i = 1; i <= ng; i++) {
                const float x = ((float)i - ((float)ng) / 2.0 - 0.5) * DH;
                const float y = ((float)j - ((float)ng) / 2.0 - 0.5) * DH;
                const float rad = sqrt(x * x + y * y);
                if (srtshape == 1) M(j, i) = erf(a * rad / maxrad);
                else if (srtshape == 2) M(j, i) = (rad > 0) ? log(rad) : 0.0;
                else M(j, i) = 0.0;
            }
        /* minimum_m(m, ng, ng) visits i outer, j inner; the first minimum
         * is the same value whatever the order */
        const float minm = orc_min_m(m, (size_t)ng * ng);
        for (size_t t = 0; t < (size_t)ng * ng; t++) m[t] -= minm;
        float edge[4] = {M(1, ng / 2), M(ng / 2, 1), M(ng / 2, ng), M(ng, ng / 2)};
        const float maxm = orc_max_m(edge, 4);
        for (size_t t = 0; t < (size_t)ng * ng; t++) {
            m[t] /= maxm;
            if (m[t] > 1.0) m[t] = 1.0;
        }
        const int ijc = (int)(ceil((float)ng / 2));
        for (int s = 0; s < nsrc; s++)
            for (int iy = 1; iy <= ng; iy++)
                for (int ix = 1; ix <= ng; ix++) {
                    const int xx = src_i[s] + ix - ijc, yy = src_j[s] + iy - ijc;
                    if ((xx < 1) || (xx > nx) || (yy < 1) || (yy > ny)) continue;
                    float *p = &msum[(size_t)(yy - 1) * nx + (size_t)(xx - 1)];
                    if (*p > M(iy, ix)) *p = M(iy, ix);
                }
#undef M
        const float mn = orc_min_m(msum, n);
        for (size_t t = 0; t < n; t++) msum[t] -= mn;
        const float mx = orc_max_m(msum, n);
        for (size_t t = 0; t < n; t++) taper[t] = msum[t] / mx;
        free(m);
        free(msum);
        for (size_t t = 0; t < n; t++) grad[t] *= taper[t];
        for (int s = 0; s < nsrc; s++)
            for (int i = src_i[s] - filtsize; i <= src_i[s] + filtsize; i++)
                for (int j = src_j[s] - filtsize; j <= src_j[s] + filtsize; j++)
                    if (i >= 1 && i <= nx && j >= 1 && j <= ny) {
                        G(j, i) = 0.0;
                        T(j, i) = 0.0;
                    }
        return 0;
    }
#undef T
#undef G
    return -1;
}

/* time_window.c:40-60 on data [ntr][ns]: samples with t = j DT >= pick are
 * multiplied by exp(-GAMMA (t - pick)^2); pick = DT (TIMEWIN 2) or
 * picks[i] + TWLENGTH_PLUS (TIMEWIN 1).  Pin: the closed form in float64
 * (tests/test_oracle_pins.py). */
void F(orc_time_window)(int ntr, int ns, double dt_in, int timewin, double gamma_in, double twplus_in,
                        const float *picks, real *data)
{
    const real DT = (real)dt_in, GAMMA = (real)gamma_in, TWLENGTH_PLUS = (real)twplus_in;
    for (int i = 0; i < ntr; i++) {
        const real pick = (timewin == 1) ? (real)picks[i] + TWLENGTH_PLUS : DT;
        for (int j = 1; j <= ns; j++) {
            const real time = (real)(j * DT);
            const real dump = time - pick;
            const real taper = exp(-GAMMA * dump * dump);
            if (time >= pick) data[(size_t)i * ns + (j - 1)] = data[(size_t)i * ns + (j - 1)] * taper;
        }
    }
}
