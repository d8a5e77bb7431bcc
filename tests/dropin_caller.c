/*
 * A C caller of the reference-named drop-ins, written against
 * include/denise_sh.h and linked to libdenise_sh.so the way DENISE-SH's own
 * denise.c would be: Numerical-Recipes offset arrays, the forward time loop
 * of denise.c:1179-1256 (update_v_PML, update_s_*_PML, seismo_ssg) after
 * zero_fdveps, with the drop-ins in copy mode (0), resident mode (1) or
 * resident mode with deferred seismograms (2).
 *
 * usage: dropin_caller <input.bin> <output.bin>
 *
 * input (little-endian, written by tests/test_gpu_dropin.py):
 *   int32  NX NY FDORDER FW FREE_SURF L NT ntr mode src_i src_j
 *   float  DT DH hc[0..6]
 *   float  12 CPML vectors of 2FW+1 entries (entry 0 unused): K_x a_x b_x
 *          K_x_half a_x_half b_x_half K_y a_y b_y K_y_half a_y_half b_y_half
 *   float  [NY][NX] planes: rho, then uip ujp (L = 0) or fipjp f dip[L] d[L]
 *   float  bip[L] bjm[L] cip[L] cjm[L] (L > 0)
 *   float  signal[NT] (a force in y, source type 3, at (src_i, src_j))
 *   int32  rec_i[ntr] rec_j[ntr]
 * output: float sectionvy [ntr][NT], then the final vy [NY][NX].
 * stdout: the time per step of the loop.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "denise_sh.h"

static float **mat(int nrl, int nrh, int ncl, int nch)
{
    int r, nrow = nrh - nrl + 1, ncol = nch - ncl + 1;
    float **m = (float **)malloc((size_t)nrow * sizeof(float *));
    float *data = (float *)calloc((size_t)nrow * ncol, sizeof(float));
    for (r = 0; r < nrow; ++r) m[r] = data + (size_t)r * ncol - ncl;
    return m - nrl;
}

static float ***tensor(int nrl, int nrh, int ncl, int nch, int ndl, int ndh)
{
    int r, c, nrow = nrh - nrl + 1, ncol = nch - ncl + 1, ndep = ndh - ndl + 1;
    float ***t = (float ***)malloc((size_t)nrow * sizeof(float **));
    float **rows = (float **)malloc((size_t)nrow * ncol * sizeof(float *));
    float *data = (float *)calloc((size_t)nrow * ncol * ndep, sizeof(float));
    for (r = 0; r < nrow; ++r) {
        t[r] = rows + (size_t)r * ncol - ncl;
        for (c = 0; c < ncol; ++c) rows[(size_t)r * ncol + c] = data + ((size_t)r * ncol + c) * ndep - ndl;
    }
    return t - nrl;
}

static void rd(FILE *fp, void *p, size_t n)
{
    if (fread(p, 1, n, fp) != n) {
        fprintf(stderr, "dropin_caller: short input\n");
        exit(2);
    }
}

static void plane(FILE *fp, float **m, int nx, int ny)
{
    int j;
    for (j = 1; j <= ny; ++j) rd(fp, &m[j][1], (size_t)nx * sizeof(float));
}

int main(int argc, char **argv)
{
    int hdr[11], NX, NY, FDORDER, FW, FREE_SURF, L, NT, ntr, mode, nd, l, v, nt, i, j;
    float head[9], *vecs[12];
    dn_globals G;
    FILE *fp;
    struct timespec t0, t1, t2;
    if (argc != 3) {
        fprintf(stderr, "usage: dropin_caller <input.bin> <output.bin>\n");
        return 2;
    }
    fp = fopen(argv[1], "rb");
    if (!fp) return 2;
    rd(fp, hdr, sizeof(hdr));
    NX = hdr[0]; NY = hdr[1]; FDORDER = hdr[2]; FW = hdr[3]; FREE_SURF = hdr[4]; L = hdr[5]; NT = hdr[6];
    ntr = hdr[7]; mode = hdr[8];
    rd(fp, head, sizeof(head));
    nd = FDORDER / 2 + 1;
    for (v = 0; v < 12; ++v) {
        vecs[v] = (float *)calloc((size_t)2 * FW + 1, sizeof(float));
        rd(fp, vecs[v], ((size_t)2 * FW + 1) * sizeof(float));
    }
    /* the reference's allocations (denise.c:320-470) */
    float **vy = mat(-nd + 1, NY + nd, -nd + 1, NX + nd), **vyp1 = mat(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **sxy = mat(-nd + 1, NY + nd, -nd + 1, NX + nd), **syz = mat(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **uy = mat(-nd + 1, NY + nd, -nd + 1, NX + nd), **uyx = mat(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **rho = mat(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **uip = mat(-nd + 1, NY + nd, -nd + 1, NX + nd), **ujp = mat(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **fipjp = mat(-nd + 1, NY + nd, -nd + 1, NX + nd), **f = mat(-nd + 1, NY + nd, -nd + 1, NX + nd);
    int Lm = L > 0 ? L : 1, fw2 = FW > 0 ? 2 * FW : 1;
    float ***r = tensor(-nd + 1, NY + nd, -nd + 1, NX + nd, 1, Lm), ***p = tensor(-nd + 1, NY + nd, -nd + 1, NX + nd, 1, Lm);
    float ***q = tensor(-nd + 1, NY + nd, -nd + 1, NX + nd, 1, Lm);
    float ***dip = tensor(-nd + 1, NY + nd, -nd + 1, NX + nd, 1, Lm), ***dd = tensor(-nd + 1, NY + nd, -nd + 1, NX + nd, 1, Lm);
    float **psi_sxy_x = mat(1, NY, 1, fw2), **psi_vyx = mat(1, NY, 1, fw2);
    float **psi_syy_y = mat(1, fw2, 1, NX), **psi_vyy = mat(1, fw2, 1, NX);
    float *bip = (float *)calloc(Lm + 1, sizeof(float)), *bjm = (float *)calloc(Lm + 1, sizeof(float));
    float *cip = (float *)calloc(Lm + 1, sizeof(float)), *cjm = (float *)calloc(Lm + 1, sizeof(float));
    float **srcpos = mat(1, 8, 1, 1), **signals = mat(1, 1, 1, NT), **section = mat(1, ntr, 1, NT);
    int **recpos = (int **)malloc(4 * sizeof(int *));
    for (i = 1; i <= 3; ++i) recpos[i] = (int *)calloc(ntr + 1, sizeof(int));
    plane(fp, rho, NX, NY);
    if (L == 0) {
        plane(fp, uip, NX, NY);
        plane(fp, ujp, NX, NY);
    } else {
        float **tmp = mat(1, NY, 1, NX);
        plane(fp, fipjp, NX, NY);
        plane(fp, f, NX, NY);
        for (l = 1; l <= L; ++l) {
            plane(fp, tmp, NX, NY);
            for (j = 1; j <= NY; ++j)
                for (i = 1; i <= NX; ++i) dip[j][i][l] = tmp[j][i];
        }
        for (l = 1; l <= L; ++l) {
            plane(fp, tmp, NX, NY);
            for (j = 1; j <= NY; ++j)
                for (i = 1; i <= NX; ++i) dd[j][i][l] = tmp[j][i];
        }
        rd(fp, &bip[1], L * sizeof(float));
        rd(fp, &bjm[1], L * sizeof(float));
        rd(fp, &cip[1], L * sizeof(float));
        rd(fp, &cjm[1], L * sizeof(float));
    }
    rd(fp, &signals[1][1], (size_t)NT * sizeof(float));
    {
        int *ri = (int *)malloc(ntr * sizeof(int)), *rj = (int *)malloc(ntr * sizeof(int));
        rd(fp, ri, ntr * sizeof(int));
        rd(fp, rj, ntr * sizeof(int));
        for (i = 1; i <= ntr; ++i) {
            recpos[1][i] = ri[i - 1];
            recpos[2][i] = rj[i - 1];
        }
    }
    fclose(fp);
    srcpos[1][1] = (float)hdr[9];
    srcpos[2][1] = (float)hdr[10];
    srcpos[8][1] = 3.0f;
    float hc[7];
    memcpy(hc, &head[2], sizeof(hc));

    memset(&G, 0, sizeof(G));
    G.NX = NX; G.NY = NY; G.FDORDER = FDORDER; G.FW = FW; G.FREE_SURF = FREE_SURF; G.BOUNDARY = 0; G.L = L;
    G.GRAD_FORM = 2; G.QUELLTYPB = 2; G.NT = NT; G.DT = head[0]; G.DH = head[1];
    if (dn_set_globals(&G) || dn_dropin_resident(mode)) {
        fprintf(stderr, "dropin_caller: %s\n", dn_last_error(NULL));
        return 1;
    }
    if (L)
        zero_fdveps_visc(-nd + 1, NY + nd, -nd + 1, NX + nd, vy, syz, sxy, uy, vyp1, psi_sxy_x, psi_vyx, psi_syy_y,
                         psi_vyy, r, p, q);
    else
        zero_fdveps(-nd + 1, NY + nd, -nd + 1, NX + nd, vy, syz, sxy, uy, vyp1, psi_sxy_x, psi_vyx, psi_syy_y, psi_vyy);
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (nt = 1; nt <= NT; ++nt) {
        update_v_PML(1, NX, 1, NY, nt, vy, vyp1, NULL, NULL, syz, sxy, rho, srcpos, signals, 1, NULL, hc, 0, 0,
                     vecs[0], vecs[1], vecs[2], vecs[3], vecs[4], vecs[5], vecs[6], vecs[7], vecs[8], vecs[9],
                     vecs[10], vecs[11], psi_syy_y, psi_sxy_x);
        if (L)
            update_s_visc_PML(1, NX, 1, NY, vy, uy, uyx, syz, sxy, ujp, uip, rho, hc, 0, r, p, q, fipjp, f, NULL, bip,
                              bjm, cip, cjm, dd, NULL, dip, vecs[0], vecs[1], vecs[2], vecs[3], vecs[4], vecs[5],
                              vecs[6], vecs[7], vecs[8], vecs[9], vecs[10], vecs[11], psi_vyy, psi_vyx);
        else
            update_s_elastic_PML(1, NX, 1, NY, vy, uy, uyx, syz, sxy, NULL, uip, ujp, NULL, rho, hc, 0, vecs[0],
                                 vecs[1], vecs[2], vecs[3], vecs[4], vecs[5], vecs[6], vecs[7], vecs[8], vecs[9],
                                 vecs[10], vecs[11], psi_vyy, psi_vyx);
        seismo_ssg(nt, ntr, recpos, section, vy, syz, NULL, hc);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (dn_dropin_sync(1)) {
        fprintf(stderr, "dropin_caller: %s\n", dn_last_error(NULL));
        return 1;
    }
    clock_gettime(CLOCK_MONOTONIC, &t2);
    {
        /* per step: the loop and the final download of every array the
         * steps wrote (in the deferred mode the loop only enqueues, so the
         * download also waits for the device) */
        double loop = (t1.tv_sec - t0.tv_sec) * 1e3 + (t1.tv_nsec - t0.tv_nsec) * 1e-6;
        double sync = (t2.tv_sec - t1.tv_sec) * 1e3 + (t2.tv_nsec - t1.tv_nsec) * 1e-6;
        printf("mode %d: %d steps of %dx%d; loop %.2f ms, final download %.2f ms; %.4f ms per step\n", mode, NT,
               NX, NY, loop, sync, (loop + sync) / NT);
    }
    if (dn_dropin_resident(0)) return 1;
    fp = fopen(argv[2], "wb");
    if (!fp) return 2;
    for (i = 1; i <= ntr; ++i) fwrite(&section[i][1], sizeof(float), NT, fp);
    for (j = 1; j <= NY; ++j) fwrite(&vy[j][1], sizeof(float), NX, fp);
    fclose(fp);
    return 0;
}
