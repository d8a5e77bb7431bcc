// Reference-named drop-ins (layer 1 of include/denise_sh.h).
//
// The reference passes Numerical-Recipes offset matrices (util.c:177-203):
// m[j][i] with row pointers, rows contiguous in i.  Each call gathers the
// rows it reads into the device layout, runs the split HIP kernel on the
// device and scatters what it writes back into the caller's arrays, so the
// caller sees exactly the reference's in-place semantics.  The caller's
// hc and CPML vectors are used as given.
#include "dn_core.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

using namespace dn;

namespace {

struct DropIn {
    bool have_globals = false;
    dn_globals G{};
    Geo g{};
    hipStream_t stream = nullptr;
    size_t cap_plane = 0;
    int cap_L = 0, cap_fw = 0, cap_nx = 0, cap_ny = 0;
    // device fields
    float *vy = nullptr, *vyp1 = nullptr, *sxy = nullptr, *syz = nullptr, *uy = nullptr, *uyx = nullptr;
    float *rho = nullptr, *uip = nullptr, *ujp = nullptr, *fipjp = nullptr, *fcoef = nullptr;
    float *dip = nullptr, *dcoef = nullptr, *r = nullptr, *q = nullptr;
    float *psi_a = nullptr, *psi_b = nullptr, *pml = nullptr;
    std::vector<float> h;  // host staging
};

DropIn &state()
{
    static DropIn s;
    return s;
}

bool fail(const char *msg)
{
    dn_fail(nullptr, "%s", msg);
    std::fprintf(stderr, "denise_sh drop-in error: %s\n", msg);
    return false;
}

void release(DropIn &d)
{
    float **all[] = {&d.vy, &d.vyp1, &d.sxy, &d.syz, &d.uy, &d.uyx, &d.rho, &d.uip, &d.ujp, &d.fipjp,
                     &d.fcoef, &d.dip, &d.dcoef, &d.r, &d.q, &d.psi_a, &d.psi_b, &d.pml};
    for (float **p : all)
        if (*p) {
            (void)hipFree(*p);
            *p = nullptr;
        }
    d.cap_plane = 0;
}

bool ensure(DropIn &d)
{
    if (!d.have_globals) return fail("dn_set_globals() must be called before the drop-ins");
    const dn_globals &G = d.G;
    if (d.stream == nullptr && hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking) != hipSuccess)
        return fail("no HIP device");
    Geo &g = d.g;
    g = make_geo(G.NX, G.NY, G.FW, G.FREE_SURF, G.L);
    if (d.cap_plane == (size_t)g.plane && d.cap_L == G.L && d.cap_fw == G.FW && d.cap_nx == G.NX &&
        d.cap_ny == G.NY)
        return true;
    release(d);
    const size_t P = (size_t)g.plane, Lm = (size_t)std::max(1, G.L), fw2 = (size_t)std::max(1, 2 * G.FW);
    const size_t psin = fw2 * (size_t)std::max(G.NX, G.NY);
    struct A {
        float **p;
        size_t n;
    } list[] = {{&d.vy, P},      {&d.vyp1, P},     {&d.sxy, P},     {&d.syz, P},      {&d.uy, P},
                {&d.uyx, P},     {&d.rho, P},      {&d.uip, P},     {&d.ujp, P},      {&d.fipjp, P},
                {&d.fcoef, P},   {&d.dip, P * Lm}, {&d.dcoef, P * Lm}, {&d.r, P * Lm}, {&d.q, P * Lm},
                {&d.psi_a, psin}, {&d.psi_b, psin}, {&d.pml, 12 * (fw2 + 1)}};
    for (auto &a : list) {
        if (hipMalloc((void **)a.p, a.n * sizeof(float)) != hipSuccess) {
            release(d);
            return fail("out of device memory");
        }
        (void)launch_fill(*a.p, a.n, 0.0f, d.stream);
    }
    d.cap_plane = P;
    d.cap_L = G.L;
    d.cap_fw = G.FW;
    d.cap_nx = G.NX;
    d.cap_ny = G.NY;
    d.h.assign(P * Lm, 0.0f);
    return hipStreamSynchronize(d.stream) == hipSuccess ? true : fail("device initialisation failed");
}

// NR matrix rows [jlo..jhi] x cols [ilo..ihi] <-> padded device field
bool put(DropIn &d, float *dev, float **m, int jlo, int jhi, int ilo, int ihi)
{
    const Geo &g = d.g;
    std::fill(d.h.begin(), d.h.begin() + g.plane, 0.0f);
    for (int j = jlo; j <= jhi; ++j) std::memcpy(&d.h[gidx(g, j, ilo)], &m[j][ilo], (size_t)(ihi - ilo + 1) * 4);
    return hipMemcpyAsync(dev, d.h.data(), (size_t)g.plane * 4, hipMemcpyHostToDevice, d.stream) == hipSuccess &&
           hipStreamSynchronize(d.stream) == hipSuccess;
}
bool get(DropIn &d, const float *dev, float **m, int jlo, int jhi, int ilo, int ihi)
{
    const Geo &g = d.g;
    if (hipMemcpyAsync(d.h.data(), dev, (size_t)g.plane * 4, hipMemcpyDeviceToHost, d.stream) != hipSuccess ||
        hipStreamSynchronize(d.stream) != hipSuccess)
        return false;
    for (int j = jlo; j <= jhi; ++j) std::memcpy(&m[j][ilo], &d.h[gidx(g, j, ilo)], (size_t)(ihi - ilo + 1) * 4);
    return true;
}
// f3tensor t[j][i][1..L] <-> L padded planes
bool put3(DropIn &d, float *dev, float ***t, int L)
{
    const Geo &g = d.g;
    std::fill(d.h.begin(), d.h.begin() + g.plane * L, 0.0f);
    for (int l = 0; l < L; ++l)
        for (int j = 1; j <= g.ny; ++j)
            for (int i = 1; i <= g.nx; ++i) d.h[(size_t)l * g.plane + gidx(g, j, i)] = t[j][i][l + 1];
    return hipMemcpyAsync(dev, d.h.data(), (size_t)g.plane * L * 4, hipMemcpyHostToDevice, d.stream) == hipSuccess &&
           hipStreamSynchronize(d.stream) == hipSuccess;
}
bool get3(DropIn &d, const float *dev, float ***t, int L)
{
    const Geo &g = d.g;
    if (hipMemcpyAsync(d.h.data(), dev, (size_t)g.plane * L * 4, hipMemcpyDeviceToHost, d.stream) != hipSuccess ||
        hipStreamSynchronize(d.stream) != hipSuccess)
        return false;
    for (int l = 0; l < L; ++l)
        for (int j = 1; j <= g.ny; ++j)
            for (int i = 1; i <= g.nx; ++i) t[j][i][l + 1] = d.h[(size_t)l * g.plane + gidx(g, j, i)];
    return true;
}
// psi matrices: [1..rows][1..cols] contiguous rows -> flat (row-1)*cols+(col-1)
bool put_psi(DropIn &d, float *dev, float **m, int rows, int cols)
{
    std::vector<float> t((size_t)rows * cols);
    for (int r = 1; r <= rows; ++r) std::memcpy(&t[(size_t)(r - 1) * cols], &m[r][1], (size_t)cols * 4);
    return hipMemcpy(dev, t.data(), t.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
}
bool get_psi(DropIn &d, const float *dev, float **m, int rows, int cols)
{
    std::vector<float> t((size_t)rows * cols);
    if (hipMemcpy(t.data(), dev, t.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return false;
    for (int r = 1; r <= rows; ++r) std::memcpy(&m[r][1], &t[(size_t)(r - 1) * cols], (size_t)cols * 4);
    return true;
}

bool put_pml(DropIn &d, PmlVec &pv, float *const vecs[12])
{
    const int n = 2 * d.G.FW + 1;
    std::vector<float> t((size_t)12 * n, 0.0f);
    for (int v = 0; v < 12; ++v)
        if (vecs[v] && d.G.FW > 0) std::memcpy(&t[(size_t)v * n + 1], vecs[v] + 1, (size_t)(n - 1) * 4);
    if (hipMemcpy(d.pml, t.data(), t.size() * 4, hipMemcpyHostToDevice) != hipSuccess) return false;
    const float *b = d.pml;
    pv.Kx = b + 0 * n; pv.ax = b + 1 * n; pv.bx = b + 2 * n;
    pv.Kxh = b + 3 * n; pv.axh = b + 4 * n; pv.bxh = b + 5 * n;
    pv.Ky = b + 6 * n; pv.ay = b + 7 * n; pv.by = b + 8 * n;
    pv.Kyh = b + 9 * n; pv.ayh = b + 10 * n; pv.byh = b + 11 * n;
    return true;
}

bool check_range(const DropIn &d, int nx1, int nx2, int ny1, int ny2)
{
    if (nx1 != 1 || ny1 != 1 || nx2 != d.G.NX || ny2 != d.G.NY)
        return fail("drop-ins update the whole single domain: call with 1, NX, 1, NY");
    return true;
}

}  // namespace

extern "C" int dn_set_globals(const dn_globals *G)
{
    if (!G) return dn_fail(nullptr, "dn_set_globals: NULL");
    if (G->FDORDER < 2 || G->FDORDER > 12 || (G->FDORDER % 2)) return dn_fail(nullptr, "wrong FDORDER %d", G->FDORDER);
    if (G->BOUNDARY != 0) return dn_fail(nullptr, "BOUNDARY=1 (periodic wavefield) is not provided");
    if (G->GRAD_FORM != 2) return dn_fail(nullptr, "only GRAD_FORM=2 is provided");
    if (G->L < 0 || G->L > kMaxL) return dn_fail(nullptr, "L=%d outside 0..%d", G->L, kMaxL);
    DropIn &d = state();
    d.G = *G;
    d.have_globals = true;
    return 0;
}

extern "C" void update_v_PML(int nx1, int nx2, int ny1, int ny2, int nt, float **vy, float **vyp1, float **vym1,
                             float **utty, float **syz, float **sxy, float **rho, float **srcpos_loc,
                             float **signals1, int nsrc, float **absorb_coeff, float *hc, int infoout, int sw,
                             float *K_x, float *a_x, float *b_x, float *K_x_half, float *a_x_half, float *b_x_half,
                             float *K_y, float *a_y, float *b_y, float *K_y_half, float *a_y_half, float *b_y_half,
                             float **psi_syy_y, float **psi_sxy_x)
{
    (void)vym1;
    (void)utty;
    (void)absorb_coeff;
    (void)infoout;
    DropIn &d = state();
    if (!ensure(d) || !check_range(d, nx1, nx2, ny1, ny2)) return;
    const dn_globals &G = d.G;
    const int K = G.FDORDER / 2;
    SplitArgs a{};
    a.g = d.g;
    a.DT = G.DT;
    a.DH = G.DH;
    for (int i = 0; i < 7; ++i) a.hc[i] = (i <= K) ? hc[i] : 0.0f;
    a.hcy6 = (K == 6) ? hc[5] : 0.0f;
    a.K = K;
    a.sw = sw;
    float *vecs[12] = {K_x, a_x, b_x, K_x_half, a_x_half, b_x_half, K_y, a_y, b_y, K_y_half, a_y_half, b_y_half};
    if (!put_pml(d, a.pml, vecs)) {
        fail("upload failed");
        return;
    }
    const int nd = K;
    bool ok = put(d, d.vy, vy, 1, G.NY, 1, G.NX) && put(d, d.sxy, sxy, 1 - nd, G.NY + nd, 1 - nd, G.NX + nd) &&
              put(d, d.syz, syz, 1 - nd, G.NY + nd, 1 - nd, G.NX + nd) && put(d, d.rho, rho, 1, G.NY, 1, G.NX) &&
              put(d, d.vyp1, vyp1, 1, G.NY, 1, G.NX);
    if (ok && G.FW > 0) ok = put_psi(d, d.psi_a, psi_syy_y, 2 * G.FW, G.NX) && put_psi(d, d.psi_b, psi_sxy_x, G.NY, 2 * G.FW);
    if (!ok) {
        fail("upload failed");
        return;
    }
    a.vy = d.vy;
    a.vyp1 = d.vyp1;
    a.sxy = d.sxy;
    a.syz = d.syz;
    a.rho = d.rho;
    a.psi_a = d.psi_a;
    a.psi_b = d.psi_b;
    if (launch_split_v(a, d.stream) != hipSuccess || hipStreamSynchronize(d.stream) != hipSuccess) {
        fail("update_v_PML kernel failed");
        return;
    }
    ok = get(d, d.vy, vy, 1, G.NY, 1, G.NX) && get(d, d.vyp1, vyp1, 1, G.NY, 1, G.NX);
    if (ok && G.FW > 0) ok = get_psi(d, d.psi_a, psi_syy_y, 2 * G.FW, G.NX) && get_psi(d, d.psi_b, psi_sxy_x, G.NY, 2 * G.FW);
    if (!ok) {
        fail("download failed");
        return;
    }
    // update_v_PML.c:563-583: raw addition after the sweep
    if (sw == 0) {
        for (int l = 1; l <= nsrc; ++l) {
            const int i = (int)srcpos_loc[1][l], j = (int)srcpos_loc[2][l];
            if ((int)srcpos_loc[8][l] == 3) vy[j][i] += signals1[l][nt];
        }
    } else if (sw == 1) {
        for (int l = 1; l <= nsrc; ++l) {
            const int i = (int)srcpos_loc[1][l], j = (int)srcpos_loc[2][l];
            if (G.QUELLTYPB == 2) vy[j][i] += signals1[l][nt];
        }
    }
}

static void update_s_common(bool visco, int nx1, int nx2, int ny1, int ny2, float **vy, float **uy, float **uyx,
                            float **syz, float **sxy, float **uip, float **ujp, float ***r, float ***q,
                            float **fipjp, float **f, float *bip, float *bjm, float *cip, float *cjm, float ***d3,
                            float ***dip, float *hc, float *const vecs[12], float **psi_vyy, float **psi_vyx)
{
    DropIn &d = state();
    if (!ensure(d) || !check_range(d, nx1, nx2, ny1, ny2)) return;
    const dn_globals &G = d.G;
    const int K = G.FDORDER / 2;
    if (visco && G.L < 1) {
        fail("update_s_visc_PML needs L >= 1");
        return;
    }
    SplitArgs a{};
    a.g = d.g;
    a.DT = G.DT;
    a.DH = G.DH;
    for (int i = 0; i < 7; ++i) a.hc[i] = (i <= K) ? hc[i] : 0.0f;
    a.K = K;
    if (!put_pml(d, a.pml, vecs)) {
        fail("upload failed");
        return;
    }
    const int nd = K;
    bool ok = put(d, d.vy, vy, 1 - nd, G.NY + nd, 1 - nd, G.NX + nd) && put(d, d.sxy, sxy, 1, G.NY, 1, G.NX) &&
              put(d, d.syz, syz, 1, G.NY, 1, G.NX);
    if (!visco)
        ok = ok && put(d, d.uip, uip, 1, G.NY, 1, G.NX) && put(d, d.ujp, ujp, 1, G.NY, 1, G.NX);
    else
        ok = ok && put(d, d.fipjp, fipjp, 1, G.NY, 1, G.NX) && put(d, d.fcoef, f, 1, G.NY, 1, G.NX) &&
             put3(d, d.r, r, G.L) && put3(d, d.q, q, G.L) && put3(d, d.dip, dip, G.L) && put3(d, d.dcoef, d3, G.L);
    if (ok && G.FW > 0) ok = put_psi(d, d.psi_a, psi_vyy, 2 * G.FW, G.NX) && put_psi(d, d.psi_b, psi_vyx, G.NY, 2 * G.FW);
    if (!ok) {
        fail("upload failed");
        return;
    }
    a.vy = d.vy;
    a.sxy = d.sxy;
    a.syz = d.syz;
    a.uy = d.uy;
    a.uyx = d.uyx;
    a.uip = d.uip;
    a.ujp = d.ujp;
    a.fipjp = d.fipjp;
    a.fcoef = d.fcoef;
    a.dip = d.dip;
    a.dcoef = d.dcoef;
    a.r = d.r;
    a.q = d.q;
    a.psi_a = d.psi_a;
    a.psi_b = d.psi_b;
    if (visco)
        for (int l = 0; l < G.L; ++l) {
            a.bip[l] = bip[l + 1];
            a.bjm[l] = bjm[l + 1];
            a.cip[l] = cip[l + 1];
            a.cjm[l] = cjm[l + 1];
        }
    if (launch_split_s(a, visco, d.stream) != hipSuccess || hipStreamSynchronize(d.stream) != hipSuccess) {
        fail("stress kernel failed");
        return;
    }
    ok = get(d, d.sxy, sxy, 1, G.NY, 1, G.NX) && get(d, d.syz, syz, 1, G.NY, 1, G.NX) &&
         get(d, d.uy, uy, 1, G.NY, 1, G.NX) && get(d, d.uyx, uyx, 1, G.NY, 1, G.NX);
    if (ok && visco) ok = get3(d, d.r, r, G.L) && get3(d, d.q, q, G.L);
    if (ok && G.FW > 0) ok = get_psi(d, d.psi_a, psi_vyy, 2 * G.FW, G.NX) && get_psi(d, d.psi_b, psi_vyx, G.NY, 2 * G.FW);
    if (!ok) fail("download failed");
}

extern "C" void update_s_elastic_PML(int nx1, int nx2, int ny1, int ny2, float **vy, float **uy, float **uyx,
                                     float **syz, float **sxy, float **u, float **uip, float **ujp,
                                     float **absorb_coeff, float **rho, float *hc, int infoout, float *K_x, float *a_x,
                                     float *b_x, float *K_x_half, float *a_x_half, float *b_x_half, float *K_y,
                                     float *a_y, float *b_y, float *K_y_half, float *a_y_half, float *b_y_half,
                                     float **psi_vyy, float **psi_vyx)
{
    (void)u;
    (void)absorb_coeff;
    (void)rho;
    (void)infoout;
    float *vecs[12] = {K_x, a_x, b_x, K_x_half, a_x_half, b_x_half, K_y, a_y, b_y, K_y_half, a_y_half, b_y_half};
    update_s_common(false, nx1, nx2, ny1, ny2, vy, uy, uyx, syz, sxy, uip, ujp, nullptr, nullptr, nullptr, nullptr,
                    nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, hc, vecs, psi_vyy, psi_vyx);
}

extern "C" void update_s_visc_PML(int nx1, int nx2, int ny1, int ny2, float **vy, float **uy, float **uyx,
                                  float **syz, float **sxy, float **ujp, float **uip, float **rho, float *hc,
                                  int infoout, float ***r, float ***p, float ***q, float **fipjp, float **f,
                                  float **g, float *bip, float *bjm, float *cip, float *cjm, float ***d, float ***e,
                                  float ***dip, float *K_x, float *a_x, float *b_x, float *K_x_half,
                                  float *a_x_half, float *b_x_half, float *K_y, float *a_y, float *b_y,
                                  float *K_y_half, float *a_y_half, float *b_y_half, float **psi_vyy,
                                  float **psi_vyx)
{
    (void)ujp;
    (void)uip;
    (void)rho;
    (void)infoout;
    (void)p;
    (void)g;
    (void)e;
    float *vecs[12] = {K_x, a_x, b_x, K_x_half, a_x_half, b_x_half, K_y, a_y, b_y, K_y_half, a_y_half, b_y_half};
    update_s_common(true, nx1, nx2, ny1, ny2, vy, uy, uyx, syz, sxy, nullptr, nullptr, r, q, fipjp, f, bip, bjm, cip,
                    cjm, d, dip, hc, vecs, psi_vyy, psi_vyx);
}

extern "C" void seismo_ssg(int lsamp, int ntr, int **recpos, float **sectionvy, float **vy, float **syy, float **u,
                           float *hc)
{
    (void)syy;
    (void)u;
    (void)hc;
    for (int itr = 1; itr <= ntr; ++itr) sectionvy[itr][lsamp] = vy[recpos[2][itr]][recpos[1][itr]];
}

extern "C" void psource(int nt, float **sxx, float **syy, float **srcpos_loc, float **signals, int nsrc)
{
    DropIn &d = state();
    if (!d.have_globals) {
        fail("dn_set_globals() must be called before psource");
        return;
    }
    const int NT = d.G.NT;
    const float DT = d.G.DT;
    for (int l = 1; l <= nsrc; ++l) {
        const int i = (int)srcpos_loc[1][l], j = (int)srcpos_loc[2][l];
        float amp = 0.0f;
        if (nt == 1) amp = signals[l][nt + 1] / DT;
        if ((nt > 1) && (nt < NT)) amp = (signals[l][nt + 1] - signals[l][nt - 1]) / DT;
        if (nt == NT) amp = -signals[l][nt - 1] / DT;
        sxx[j][i] += amp;
        syy[j][i] += amp;
    }
}
