// Host orchestration of the fast path (handle API of include/denise_sh.h):
// resident model and shot data, batched time loops, residual, energy
// preconditioning and chain rule.  The reference's shot loop
// (denise.c:875-1702) becomes: shots are grouped into batches that fit in
// HBM; each batch runs its forward loop (one fused launch per step for all
// shots), its residuals, its adjoint loop, then the per-shot preconditioning
// and accumulation in shot order.
#include "dn_core.h"
#include "dn_setup.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

using namespace dn;

static thread_local std::string g_create_error;

int dn_fail(dn_ctx *c, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c)
        c->err = buf;
    else
        g_create_error = buf;
    return -1;
}

static void dfree(void *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}
template <typename T> static void dfree_t(T *&p)
{
    void *q = (void *)p;
    dfree(q);
    p = nullptr;
}

template <typename T> static hipError_t dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void **)p, n * sizeof(T));
}

extern "C" const char *dn_last_error(const dn_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

extern "C" dn_ctx *dn_create(const dn_params *p)
{
    if (!p) {
        dn_fail(nullptr, "dn_create: params is NULL");
        return nullptr;
    }
    if (p->nx < 2 || p->ny < 2) {
        dn_fail(nullptr, "dn_create: grid %dx%d too small", p->nx, p->ny);
        return nullptr;
    }
    if (p->L < 0 || p->L > kMaxL) {
        dn_fail(nullptr, "dn_create: L=%d outside 0..%d", p->L, kMaxL);
        return nullptr;
    }
    if (p->fw < 0 || 2 * p->fw > std::min(p->nx, p->ny) + p->fw) {
        dn_fail(nullptr, "dn_create: FW=%d larger than the grid", p->fw);
        return nullptr;
    }
    // the fused time step takes the absorbing frames on opposite sides to be
    // disjoint (the reference would apply both to a cell where they overlap)
    if (p->fw > 0 && (p->nx < 2 * p->fw + 1 || (!p->free_surf && p->ny < 2 * p->fw + 1))) {
        dn_fail(nullptr, "dn_create: the CPML frames of width FW=%d overlap on a %dx%d grid", p->fw, p->nx, p->ny);
        return nullptr;
    }
    {
        // the fused step addresses one shot's fields with 32-bit byte offsets
        const Geo t = make_geo(p->nx, p->ny, p->fw, p->free_surf, p->L);
        if ((double)t.plane * (double)(16 + 8 * p->L) >= 2147483648.0) {
            dn_fail(nullptr, "dn_create: a %dx%d grid is too large for one device context", p->nx, p->ny);
            return nullptr;
        }
    }
    if (p->dtinv < 1) {
        dn_fail(nullptr, "dn_create: DTINV must be >= 1");
        return nullptr;
    }
    if (p->snap_precision < 0 || p->snap_precision > 2) {
        dn_fail(nullptr, "dn_create: snap_precision must be 0, 1 or 2");
        return nullptr;
    }
    if (p->eprecond != 0 && p->eprecond != 1 && p->eprecond != 3) {
        dn_fail(nullptr, "dn_create: eprecond must be 0, 1 or 3");
        return nullptr;
    }
    if (p->boundary != 0 && p->boundary != 1) {
        dn_fail(nullptr, "dn_create: BOUNDARY must be 0 or 1");
        return nullptr;
    }
    if (p->boundary == 1 && p->nx < kWave) {
        // a wave of the fused step spans 64 columns; wrapped once, they must
        // stay inside the grid
        dn_fail(nullptr, "dn_create: BOUNDARY=1 needs NX >= %d (got %d)", kWave, p->nx);
        return nullptr;
    }
    dn_ctx *c = new dn_ctx();
    c->p = *p;
    if (fd_coefficients(p->fdorder, p->maxrelerror, c->hc) != 0) {
        dn_fail(nullptr, "Error in selection of FD coefficients: FDORDER=%d MAXRELERROR=%d", p->fdorder,
                p->maxrelerror);
        delete c;
        return nullptr;
    }
    c->K = p->fdorder / 2;
    Geo &g = c->g;
    g = make_geo(p->nx, p->ny, p->fw, p->free_surf, p->L, p->boundary);

    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        dn_fail(nullptr, "dn_create: hipStreamCreate failed (no HIP device?)");
        delete c;
        return nullptr;
    }
    CpmlIn ci{p->nx, p->ny, p->fw, p->dh, p->dt, p->damping, p->fpml, p->npower, p->k_max_pml};
    CpmlOut co;
    cpml_profiles(ci, co);
    const size_t nv = (size_t)co.n;
    std::vector<float> packed(12 * nv);
    for (int v = 0; v < 12; ++v) std::copy(co.v[v].begin(), co.v[v].end(), packed.begin() + v * nv);
    if (dalloc(&c->d_pml, packed.size()) != hipSuccess ||
        hipMemcpy(c->d_pml, packed.data(), packed.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        dn_fail(nullptr, "dn_create: CPML upload failed");
        dn_destroy(c);
        return nullptr;
    }
    c->pml_host = packed;
    PmlVec &pv = c->pml;
    const float *b = c->d_pml;
    pv.Kx = b + 0 * nv; pv.ax = b + 1 * nv; pv.bx = b + 2 * nv;
    pv.Kxh = b + 3 * nv; pv.axh = b + 4 * nv; pv.bxh = b + 5 * nv;
    pv.Ky = b + 6 * nv; pv.ay = b + 7 * nv; pv.by = b + 8 * nv;
    pv.Kyh = b + 9 * nv; pv.ayh = b + 10 * nv; pv.byh = b + 11 * nv;

    if (p->L > 0)
        relaxation_constants(p->L, p->fl, p->dt, c->eta, &c->sumu, c->bip, c->bjm, c->cip, c->cjm);

    // model buffers (padded)
    const size_t P = (size_t)g.plane;
    bool ok = dalloc(&c->d_vs, P) == hipSuccess && dalloc(&c->d_rho, P) == hipSuccess &&
              dalloc(&c->d_taus, P) == hipSuccess && dalloc(&c->d_uip, P) == hipSuccess &&
              dalloc(&c->d_ujp, P) == hipSuccess && dalloc(&c->d_b, P) == hipSuccess &&
              dalloc(&c->d_tausipjp, P) == hipSuccess && dalloc(&c->d_fipjp, P) == hipSuccess &&
              dalloc(&c->d_f, P) == hipSuccess && dalloc(&c->d_dip, P * std::max(1, p->L)) == hipSuccess &&
              dalloc(&c->d_d, P * std::max(1, p->L)) == hipSuccess &&
              dalloc(&c->d_spk, spk_floats(P, p->L)) == hipSuccess &&
              dalloc(&c->d_gu, (size_t)p->nx * p->ny) == hipSuccess &&
              dalloc(&c->d_grho, (size_t)p->nx * p->ny) == hipSuccess &&
              dalloc(&c->d_gvs_out, (size_t)p->nx * p->ny) == hipSuccess &&
              dalloc(&c->d_grho_out, (size_t)p->nx * p->ny) == hipSuccess &&
              dalloc(&c->d_ws_sum, (size_t)p->nx * p->ny) == hipSuccess &&
              dalloc(&c->d_wr_sum, (size_t)p->nx * p->ny) == hipSuccess &&
              dalloc(&c->d_maxbits, 4096) == hipSuccess && dalloc(&c->d_status, 4) == hipSuccess;
    if (!ok) {
        dn_fail(nullptr, "dn_create: out of device memory for the model");
        dn_destroy(c);
        return nullptr;
    }
    float *zs[] = {c->d_vs, c->d_rho, c->d_taus, c->d_uip, c->d_ujp, c->d_b, c->d_tausipjp, c->d_fipjp, c->d_f};
    for (float *z : zs) (void)launch_fill(z, P, 0.0f, c->stream);
    (void)launch_fill(c->d_dip, P * std::max(1, p->L), 0.0f, c->stream);
    (void)launch_fill(c->d_d, P * std::max(1, p->L), 0.0f, c->stream);
    (void)launch_fill(c->d_spk, spk_floats(P, p->L), 0.0f, c->stream);
    (void)launch_fill(c->d_gu, (size_t)p->nx * p->ny, 0.0f, c->stream);
    (void)launch_fill(c->d_grho, (size_t)p->nx * p->ny, 0.0f, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        dn_fail(nullptr, "dn_create: device initialisation failed");
        dn_destroy(c);
        return nullptr;
    }
    (void)hipEventCreate(&c->ev0);
    (void)hipEventCreate(&c->ev1);
    return c;
}

void dn_drop_graphs(dn_ctx *c)
{
    for (auto &gr : c->graphs) (void)hipGraphExecDestroy(gr.exec);
    c->graphs.clear();
}

static void free_batch(dn_ctx *c)
{
    dn_drop_graphs(c);
    for (int k = 0; k < 2; ++k) {
        dfree_t(c->d_vy[k]);
        dfree_t(c->d_sxy[k]);
        dfree_t(c->d_syz[k]);
        dfree_t(c->d_psx[k]);
        dfree_t(c->d_psy[k]);
    }
    dfree_t(c->d_rq); dfree_t(c->d_pvx); dfree_t(c->d_pvy);
    dfree_t(c->d_seis); dfree_t(c->d_resid); dfree_t(c->d_snap); dfree_t(c->d_snap16); dfree_t(c->d_snap_scale);
    dfree_t(c->d_stage); dfree_t(c->d_stage_max);
    dfree_t(c->d_Ws); dfree_t(c->d_Wr); dfree_t(c->d_gu_shot); dfree_t(c->d_grho_shot); dfree_t(c->d_part);
    dfree_t(c->d_shot_map);
    c->cur_map.clear();   // the shot map must be uploaded again into the new workspace
    c->batch_cap = 0;
    c->batch_nt_max = 0;
    // allocated after the workspace, from what it leaves
    dfree_t(c->d_terms);
    c->terms_cap = 0;
    dfree_t(c->d_rank);
    c->rank_cap = 0;
}

extern "C" void dn_destroy(dn_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    dn_fwi_destroy(c);
    dn_gradproc_destroy(c);
    free_batch(c);
    dfree_t(c->d_pml);
    dfree_t(c->d_vs); dfree_t(c->d_rho); dfree_t(c->d_taus); dfree_t(c->d_uip); dfree_t(c->d_ujp);
    dfree_t(c->d_b); dfree_t(c->d_tausipjp); dfree_t(c->d_fipjp); dfree_t(c->d_f); dfree_t(c->d_dip);
    dfree_t(c->d_spk);
    dfree_t(c->d_d); dfree_t(c->d_gu); dfree_t(c->d_grho); dfree_t(c->d_gvs_out); dfree_t(c->d_grho_out);
    dfree_t(c->d_maxbits);
    dfree_t(c->d_status);
    dfree_t(c->d_trace_flags);
    dfree_t(c->d_rec_row); dfree_t(c->d_rec_tab);
    dfree_t(c->d_shot_nt); dfree_t(c->d_src_off); dfree_t(c->d_src_i); dfree_t(c->d_src_j); dfree_t(c->d_sig);
    dfree_t(c->d_obs);
    dfree_t(c->d_ntbase);
    dfree_t(c->d_hu);
    dfree_t(c->d_hr);
    dfree_t(c->d_ws_sum);
    dfree_t(c->d_wr_sum);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// host [ny][nx] -> padded device field (interior only; halo zero)
static int upload_padded(dn_ctx *c, float *dst, const float *src)
{
    const Geo &g = c->g;
    HIPCHK(c, hipMemcpy2DAsync(dst + gidx(g, 1, 1), (size_t)g.pitch * sizeof(float), src, (size_t)g.nx * sizeof(float),
                               (size_t)g.nx * sizeof(float), (size_t)g.ny, hipMemcpyHostToDevice, c->stream));
    return 0;
}
static int download_padded(dn_ctx *c, float *dst, const float *src)
{
    const Geo &g = c->g;
    HIPCHK(c, hipMemcpy2DAsync(dst, (size_t)g.nx * sizeof(float), src + gidx(g, 1, 1), (size_t)g.pitch * sizeof(float),
                               (size_t)g.nx * sizeof(float), (size_t)g.ny, hipMemcpyDeviceToHost, c->stream));
    return 0;
}

int dn_prepare_model(dn_ctx *c)
{
    ModelPrepArgs a{};
    a.g = c->g;
    a.DT = c->p.dt;
    a.L = c->p.L;
    a.vs = c->d_vs;
    a.rho = c->d_rho;
    a.taus = c->d_taus;
    a.uip = c->d_uip;
    a.ujp = c->d_ujp;
    a.tausipjp = c->d_tausipjp;
    a.bfield = c->d_b;
    a.fipjp = c->d_fipjp;
    a.fcoef = c->d_f;
    a.dip = c->d_dip;
    a.dcoef = c->d_d;
    a.spk = c->d_spk;
    a.sumu = c->sumu;
    for (int l = 0; l < kMaxL; ++l) a.eta[l] = (l < c->p.L) ? c->eta[l] : 0.0f;
    HIPCHK(c, launch_model_prep(a, c->stream));
    return 0;
}

// denise.c:778-816: the means of vs and rho as float sums, i outer, j
// inner, divided by the cell count
static void model_means(const float *vs, const float *rho, int nx, int ny, float &C_vs, float &C_rho)
{
    float vs_avg = 0.0f, rho_avg = 0.0f;
    int count = 1;
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j) {
            vs_avg += vs[(size_t)(j - 1) * nx + (i - 1)];
            rho_avg += rho[(size_t)(j - 1) * nx + (i - 1)];
            ++count;
        }
    --count;
    vs_avg /= (float)count;
    rho_avg /= (float)count;
    C_vs = vs_avg;
    C_rho = rho_avg;
}

extern "C" int dn_set_model(dn_ctx *c, const float *vs, const float *rho, const float *qs)
{
    if (!c || !vs || !rho) return dn_fail(c, "dn_set_model: NULL argument");
    const int nx = c->p.nx, ny = c->p.ny, L = c->p.L;
    if (L > 0 && !qs) return dn_fail(c, "dn_set_model: qs required for L=%d", L);
    std::vector<float> taus;
    if (L > 0) {
        taus.resize((size_t)nx * ny);
        for (size_t k = 0; k < taus.size(); ++k) taus[k] = (float)(2.0 / (double)qs[k]);  // readmod.c:99
    }
    if (upload_padded(c, c->d_vs, vs) || upload_padded(c, c->d_rho, rho)) return -1;
    if (L > 0 && upload_padded(c, c->d_taus, taus.data())) return -1;
    if (dn_prepare_model(c)) return -1;
    c->have_model = true;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// The means that normalise the gradient and the L-BFGS vectors.  The
// reference takes them once, from the model of its first iteration
// (denise.c:746-819, inside `if (iter_true == 1)`), and keeps them for the
// run; here they are taken from the device model at the first gradient of a
// context and at dn_fwi_start, and kept until then.
int dn_update_means(dn_ctx *c)
{
    const size_t n = (size_t)c->g.nx * c->g.ny;
    std::vector<float> vs(n), rho(n);
    if (dn_get_model(c, vs.data(), rho.data())) return -1;
    model_means(vs.data(), rho.data(), c->g.nx, c->g.ny, c->C_vs, c->C_rho);
    c->C_fixed = true;
    return 0;
}

extern "C" int dn_get_means(const dn_ctx *c, float *c_vs, float *c_rho)
{
    if (!c) return -1;
    if (c_vs) *c_vs = c->C_vs;
    if (c_rho) *c_rho = c->C_rho;
    return c->C_fixed ? 0 : 1;
}

extern "C" int dn_get_model(dn_ctx *c, float *vs, float *rho)
{
    if (!c) return -1;
    if (vs && download_padded(c, vs, c->d_vs)) return -1;
    if (rho && download_padded(c, rho, c->d_rho)) return -1;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int dn_set_receivers(dn_ctx *c, int ntr, const int *ri, const int *rj)
{
    if (!c || ntr < 0 || (ntr > 0 && (!ri || !rj))) return dn_fail(c, "dn_set_receivers: bad arguments");
    const int nx = c->p.nx, ny = c->p.ny;
    std::vector<int> row(ny + 1, -1);
    std::vector<int> tab;
    for (int r = 0; r < ntr; ++r) {
        if (ri[r] < 1 || ri[r] > nx || rj[r] < 1 || rj[r] > ny)
            return dn_fail(c, "receiver %d at (%d,%d) outside the grid", r + 1, ri[r], rj[r]);
        if (row[rj[r]] < 0) {
            row[rj[r]] = (int)tab.size();
            tab.insert(tab.end(), (size_t)nx + 1, -1);
        }
        int &slot = tab[(size_t)row[rj[r]] + ri[r]];
        if (slot >= 0) return dn_fail(c, "receivers %d and %d share a grid point", slot + 1, r + 1);
        slot = r;
    }
    if (tab.empty()) tab.push_back(-1);
    dfree_t(c->d_rec_row);
    dfree_t(c->d_rec_tab);
    HIPCHK(c, dalloc(&c->d_rec_row, row.size()));
    HIPCHK(c, dalloc(&c->d_rec_tab, tab.size()));
    HIPCHK(c, hipMemcpy(c->d_rec_row, row.data(), row.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_rec_tab, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice));
    c->ntr = ntr;
    c->rec_i.assign(ri, ri + ntr);
    c->rec_j.assign(rj, rj + ntr);
    c->dc = dn_ctx::DataCond{};   // conditioning refers to receivers: reset
    c->dc_dirty = true;
    dfree_t(c->d_obs);  // observed data are [shot][ntr][nt_max]: set them again
    c->h_obs.clear();
    free_batch(c);      // seismogram buffers depend on ntr
    return 0;
}

extern "C" int dn_set_shots_multi(dn_ctx *c, int nshots, int nt_max, const int *nt, const int *src_off,
                                  const int *si, const int *sj, const float *signals)
{
    if (!c || nshots < 1 || nt_max < 1 || !nt || !src_off || !si || !sj || !signals)
        return dn_fail(c, "dn_set_shots: bad arguments");
    if (src_off[0] != 0) return dn_fail(c, "dn_set_shots: src_off[0] must be 0");
    for (int s = 0; s < nshots; ++s) {
        if (nt[s] < 1 || nt[s] > nt_max) return dn_fail(c, "shot %d: nt=%d outside 1..%d", s + 1, nt[s], nt_max);
        if (src_off[s + 1] < src_off[s]) return dn_fail(c, "shot %d: bad source offsets", s + 1);
    }
    const int nsrc = src_off[nshots];
    for (int l = 0; l < nsrc; ++l)
        if (si[l] < 1 || si[l] > c->p.nx || sj[l] < 1 || sj[l] > c->p.ny)
            return dn_fail(c, "source %d at (%d,%d) outside the grid", l + 1, si[l], sj[l]);
    dfree_t(c->d_shot_nt); dfree_t(c->d_src_off); dfree_t(c->d_src_i); dfree_t(c->d_src_j); dfree_t(c->d_sig);
    dfree_t(c->d_obs);
    HIPCHK(c, dalloc(&c->d_shot_nt, nshots));
    HIPCHK(c, dalloc(&c->d_src_off, nshots + 1));
    HIPCHK(c, dalloc(&c->d_src_i, std::max(1, nsrc)));
    HIPCHK(c, dalloc(&c->d_src_j, std::max(1, nsrc)));
    HIPCHK(c, dalloc(&c->d_sig, (size_t)std::max(1, nsrc) * nt_max));
    HIPCHK(c, hipMemcpy(c->d_shot_nt, nt, nshots * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_src_off, src_off, (nshots + 1) * sizeof(int), hipMemcpyHostToDevice));
    if (nsrc > 0) {
        HIPCHK(c, hipMemcpy(c->d_src_i, si, nsrc * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_src_j, sj, nsrc * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_sig, signals, (size_t)nsrc * nt_max * sizeof(float), hipMemcpyHostToDevice));
    }
    c->nshots = nshots;
    c->nt_max = nt_max;
    c->shot_nt.assign(nt, nt + nshots);
    c->src_i.assign(si, si + nsrc);
    c->src_j.assign(sj, sj + nsrc);
    c->src_off.assign(src_off, src_off + nshots + 1);
    c->dc = dn_ctx::DataCond{};   // conditioning refers to shots: reset
    c->dc_dirty = true;
    c->shot_l2.assign(nshots, 0.0);
    c->shot_energy.assign(nshots, 0.0);
    c->cur_map.clear();
    c->h_sig.assign(signals, signals + (size_t)nsrc * nt_max);
    c->h_obs.clear();
    c->filt_sig = c->filt_obs = dn_ctx::TimeFilt{};
    free_batch(c);
    return 0;
}

extern "C" int dn_set_shots(dn_ctx *c, int nshots, int nt_max, const int *nt, const int *si, const int *sj,
                            const float *signals)
{
    if (!c || nshots < 1) return dn_fail(c, "dn_set_shots: bad arguments");
    std::vector<int> off(nshots + 1);
    for (int s = 0; s <= nshots; ++s) off[s] = s;
    return dn_set_shots_multi(c, nshots, nt_max, nt, off.data(), si, sj, signals);
}

extern "C" int dn_set_signals(dn_ctx *c, const float *signals)
{
    if (!c || !signals || c->nshots < 1) return dn_fail(c, "dn_set_signals: set shots first");
    const size_t n = c->src_i.size() * (size_t)c->nt_max;
    HIPCHK(c, hipMemcpy(c->d_sig, signals, n * sizeof(float), hipMemcpyHostToDevice));
    c->h_sig.assign(signals, signals + n);
    c->filt_sig = dn_ctx::TimeFilt{};
    return 0;
}

extern "C" int dn_set_observed(dn_ctx *c, const float *obs)
{
    if (!c || !obs) return dn_fail(c, "dn_set_observed: NULL argument");
    if (c->nshots < 1 || c->ntr < 1) return dn_fail(c, "dn_set_observed: set shots and receivers first");
    const size_t n = (size_t)c->nshots * c->ntr * c->nt_max;
    if (!c->d_obs) HIPCHK(c, dalloc(&c->d_obs, n));
    HIPCHK(c, hipMemcpy(c->d_obs, obs, n * sizeof(float), hipMemcpyHostToDevice));
    c->h_obs.assign(obs, obs + n);
    c->filt_obs = dn_ctx::TimeFilt{};
    return 0;
}

// Filters copies of [rows][nt_max] traces whose first ns(row) samples carry
// the record, as timedomain_filt.c does per trace (low-pass at fc, then for
// TIME_FILT 2 high-pass at fc_start), spread over host threads.
static void filter_rows(std::vector<float> &out, const std::vector<float> &in, int rows, int nt_max,
                        const std::vector<int> &row_ns, const dn_ctx::TimeFilt &f, float dt)
{
    out = in;
    if (f.time_filt == 0) return;
    const int nth = std::max(1, std::min<int>(16, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int t = 0; t < nth; ++t)
        pool.emplace_back([&, t]() {
            for (int r = t; r < rows; r += nth) {
                float *row = out.data() + (size_t)r * nt_max;
                dn::timedomain_filt(row, 1, row_ns[r], dt, f.fc, f.order, 1);
                if (f.time_filt == 2) dn::timedomain_filt(row, 1, row_ns[r], dt, f.fc_start, f.order, 2);
            }
        });
    for (auto &th : pool) th.join();
}

extern "C" int dn_set_time_filt(dn_ctx *c, int time_filt, float fc_start, float fc, int order)
{
    if (!c) return -1;
    if (time_filt < 0 || time_filt > 2) return dn_fail(c, "dn_set_time_filt: TIME_FILT must be 0, 1 or 2");
    if (time_filt && (order < 1 || !(fc > 0.0f) || (time_filt == 2 && !(fc_start > 0.0f))))
        return dn_fail(c, "dn_set_time_filt: need ORDER >= 1 and positive corner frequencies");
    if (c->nshots < 1) return dn_fail(c, "dn_set_time_filt: set shots first");
    dn_ctx::TimeFilt want;
    want.time_filt = time_filt;
    if (time_filt) {
        want.order = order;
        want.fc = fc;
        want.fc_start = time_filt == 2 ? fc_start : 0.0f;
    }
    std::vector<float> buf;
    if (!(c->filt_sig == want)) {
        // sources of shot s: src_off[s] .. src_off[s+1]-1, each nt(s) samples
        std::vector<int> ns(c->src_i.size());
        for (int s = 0; s < c->nshots; ++s)
            for (int l = c->src_off[s]; l < c->src_off[s + 1]; ++l) ns[l] = c->shot_nt[s];
        filter_rows(buf, c->h_sig, (int)ns.size(), c->nt_max, ns, want, c->p.dt);
        if (!buf.empty()) HIPCHK(c, hipMemcpy(c->d_sig, buf.data(), buf.size() * sizeof(float), hipMemcpyHostToDevice));
        c->filt_sig = want;
    }
    if (!c->h_obs.empty() && c->d_obs && !(c->filt_obs == want)) {
        std::vector<int> ns((size_t)c->nshots * c->ntr);
        for (int s = 0; s < c->nshots; ++s)
            for (int r = 0; r < c->ntr; ++r) ns[(size_t)s * c->ntr + r] = c->shot_nt[s];
        filter_rows(buf, c->h_obs, (int)ns.size(), c->nt_max, ns, want, c->p.dt);
        HIPCHK(c, hipMemcpy(c->d_obs, buf.data(), buf.size() * sizeof(float), hipMemcpyHostToDevice));
        c->filt_obs = want;
    }
    return 0;
}

extern "C" int dn_set_batch(dn_ctx *c, int max_shots, int rows_per_segment)
{
    if (!c) return -1;
    c->batch_req = std::max(0, max_shots);
    c->rows_req = std::max(0, rows_per_segment);
    free_batch(c);
    return 0;
}

extern "C" int dn_set_batch_order(dn_ctx *c, int by_record_length)
{
    if (!c) return -1;
    c->order_by_nt = by_record_length != 0;
    return 0;
}

extern "C" int dn_batch_shots(dn_ctx *c, int gradient, int *nshots)
{
    if (!c || !nshots) return dn_fail(c, "dn_batch_shots: NULL argument");
    if (dn_ensure_batch(c, c->nshots, gradient != 0)) return -1;
    *nshots = std::min(c->batch_cap, c->nshots);
    return 0;
}

extern "C" int dn_set_exec(dn_ctx *c, int graph_steps)
{
    if (!c) return -1;
    if (graph_steps < -1) return dn_fail(c, "dn_set_exec: graph_steps must be >= -1");
    c->graph_steps = graph_steps;
    dn_drop_graphs(c);
    return 0;
}

extern "C" int dn_set_wave_shots(dn_ctx *c, int n)
{
    if (!c) return -1;
    if (n != 1 && n != 2) return dn_fail(c, "dn_set_wave_shots: 1 or 2 shots per wave");
    c->two_shots = (n == 2);
    dn_drop_graphs(c);
    return 0;
}

extern "C" int dn_get_launch_shape(const dn_ctx *c, int *rows, int *graph, int *wave_shots, int *group_shots)
{
    if (!c) return dn_fail(nullptr, "dn_get_launch_shape: NULL context");
    if (rows) *rows = c->rows_req;
    if (graph) *graph = c->graph_steps;
    if (wave_shots) *wave_shots = c->two_shots ? 2 : 1;
    if (group_shots) *group_shots = c->sg;
    return 0;
}

extern "C" int dn_set_group_layout(dn_ctx *c, int mode)
{
    if (!c) return -1;
    if (mode != 0 && mode != 1 && mode != 2 && mode != 4 && mode != 8)
        return dn_fail(c, "dn_set_group_layout: 0 (adjacent strips of one shot), 2 or 4 (shots x strips), 1 or 8 "
                          "(one strip of eight shots)");
    c->sg = (mode == 0) ? 1 : (mode == 1 ? 8 : mode);
    dn_drop_graphs(c);
    return 0;
}

extern "C" int dn_autotune(dn_ctx *c, int nshots, int nsteps, int *rows_out, int *graph_out)
{
    if (!c) return -1;
    if (nshots < 1 || nshots > c->nshots || nsteps < 1) return dn_fail(c, "dn_autotune: bad arguments");
    if (!c->have_model) return dn_fail(c, "dn_autotune: model not set");
    // tune the launch the gradient runs: as many of the shots as share a
    // batch once the snapshots are stored (the forward-only batch of the
    // same shots can be much larger and wants other segment heights)
    if (dn_ensure_batch(c, nshots, true)) return -1;
    const int n = std::min(nshots, c->batch_cap);
    std::vector<int> shots(n);
    for (int s = 0; s < n; ++s) shots[s] = s;
    const int saved_rows = c->rows_req, saved_graph = c->graph_steps;
    // 0: the rows the cost model of choose_rows picks (whole rounds of
    // resident workgroups), then fixed heights
    const int cand[] = {0, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 1 << 20};
    auto time_it = [&](int rows, int graph, double *ms) -> int {
        c->rows_req = rows;
        c->graph_steps = graph;
        dn_drop_graphs(c);
        c->nt_limit = nsteps;
        dn_reset_timing(c);
        int rc = dn_run_pass(c, shots, false, false, false);   // first pass: warm-up and capture
        // the faster of two timed passes
        double t = 1e300;
        for (int k = 0; rc == 0 && k < 2; ++k) {
            dn_reset_timing(c);
            rc = dn_run_pass(c, shots, false, false, false);
            t = std::min(t, c->ms_steps);
        }
        c->nt_limit = 0;
        *ms = t;
        return rc;
    };
    double best = 1e300;
    int best_rows = 0;
    std::vector<int> tried;
    for (int R : cand) {
        const int Rr = (R == 0) ? 0 : std::min(R, c->g.ny);
        if ((Rr != 0 && Rr < std::min(2 * c->K, c->g.ny)) || std::find(tried.begin(), tried.end(), Rr) != tried.end())
            continue;
        tried.push_back(Rr);
        double ms = 0.0;
        if (time_it(Rr, 0, &ms)) {
            c->rows_req = saved_rows;
            c->graph_steps = saved_graph;
            return -1;
        }
        if (ms < best) {
            best = ms;
            best_rows = Rr;
        }
    }
    // one or two shots per wave on the plain steps
    if (n >= 2) {
        const bool saved_two = c->two_shots;
        c->two_shots = !saved_two;
        double ms_other = 0.0;
        if (time_it(best_rows, 0, &ms_other)) {
            c->two_shots = saved_two;
            c->rows_req = saved_rows;
            c->graph_steps = saved_graph;
            return -1;
        }
        if (ms_other < best)
            best = ms_other;
        else
            c->two_shots = saved_two;
    }
    // the other workgroup layouts (shots x strips per workgroup, each with
    // its own rows per segment, as the workgroup count differs)
    {
        const int saved_sg = c->sg;
        int best_sg = saved_sg;
        for (int sg : {1, 2, 4, 8}) {
            if (sg == saved_sg || sg > std::max(1, n)) continue;
            c->sg = sg;
            for (int R : {0, best_rows / 2, best_rows, best_rows * 2}) {
                const int Rr = std::min(R, c->g.ny);
                if (R != 0 && (best_rows == 0 || Rr < std::min(2 * c->K, c->g.ny))) continue;
                double ms = 0.0;
                if (time_it(Rr, 0, &ms)) {
                    c->sg = saved_sg;
                    c->rows_req = saved_rows;
                    c->graph_steps = saved_graph;
                    return -1;
                }
                if (ms < best) {
                    best = ms;
                    best_rows = Rr;
                    best_sg = sg;
                }
            }
        }
        c->sg = best_sg;
    }
    double ms_graph = 0.0;
    if (time_it(best_rows, -1, &ms_graph)) {
        c->rows_req = saved_rows;
        c->graph_steps = saved_graph;
        return -1;
    }
    c->rows_req = best_rows;
    // graph replay unless launching every step from the host is clearly
    // faster (the two differ by the launch gaps only)
    c->graph_steps = (ms_graph <= 1.02 * best) ? -1 : 0;
    dn_drop_graphs(c);
    dn_reset_timing(c);
    if (rows_out) *rows_out = best_rows;
    if (graph_out) *graph_out = c->graph_steps;
    return 0;
}

int dn_nsnap(const dn_ctx *c) { return (c->nt_max + c->p.dtinv - 1) / c->p.dtinv; }

// bytes of per-shot device state for a gradient batch
static size_t per_shot_bytes(const dn_ctx *c, bool gradient)
{
    const Geo &g = c->g;
    const size_t P = (size_t)g.plane, nxny = (size_t)g.nx * g.ny;
    size_t b = 6 * P * 4 + 2 * (size_t)g.L * P * 4;                 // vy,sxy,syz x2, r, q
    b += (size_t)(2 * 2 + 2) * (size_t)(2 * g.fw) * (g.nx + g.ny) * 4;  // psi
    b += 2 * (size_t)c->ntr * c->nt_max * 4;                        // seis + resid
    b += 4 * nxny * 4 + (size_t)c->ntr * 16;                        // Ws, Wr, gu, grho, partials
    if (gradient) {
        const size_t ns = (size_t)dn_nsnap(c);
        b += (c->p.snap_precision == 0 ? 4 : 2) * 3 * nxny * ns;
        if (c->p.snap_precision != 0) b += ns * 12 + 36;
        if (c->p.snap_precision == 2) b += 3 * nxny * 4;   // FP16: FP32 staging plane
    }
    return b;
}

int dn_ensure_batch(dn_ctx *c, int want, bool gradient)
{
    if (c->nshots < 1) return dn_fail(c, "no shots set");
    if (c->ntr < 1) return dn_fail(c, "no receivers set");
    int cap = want;
    if (c->batch_req > 0) cap = std::min(cap, c->batch_req);
    const size_t per = per_shot_bytes(c, gradient || c->batch_has_snap);
    // reuse the workspace when it holds the request, or as many shots as
    // fitted into HBM when it was sized
    if (c->batch_cap > 0 && c->batch_cap >= std::min(cap, c->batch_fit) && c->batch_nt_max == c->nt_max &&
        (!gradient || c->batch_has_snap))
        return 0;
    free_batch(c);
    size_t freeb = 0, total = 0;
    HIPCHK(c, hipMemGetInfo(&freeb, &total));
    const size_t budget = (size_t)((double)freeb * 0.85);
    const int fit = (int)std::max<size_t>(1, budget / std::max<size_t>(per, 1));
    cap = std::max(1, std::min(std::min(cap, fit), 4096));
    c->batch_fit = std::min(fit, 4096);
    if (per > budget) return dn_fail(c, "one shot needs %zu MB but only %zu MB of HBM are free", per >> 20, freeb >> 20);
    const Geo &g = c->g;
    const size_t P = (size_t)g.plane, nxny = (size_t)g.nx * g.ny;
    const size_t S = (size_t)cap;
    const size_t fw2 = (size_t)(2 * g.fw);
    for (int k = 0; k < 2; ++k) {
        HIPCHK(c, dalloc(&c->d_vy[k], S * P));
        HIPCHK(c, dalloc(&c->d_sxy[k], S * P));
        HIPCHK(c, dalloc(&c->d_syz[k], S * P));
        HIPCHK(c, dalloc(&c->d_psx[k], S * fw2 * g.ny));
        HIPCHK(c, dalloc(&c->d_psy[k], S * fw2 * g.nx));
    }
    HIPCHK(c, dalloc(&c->d_rq, S * P * 2 * std::max(1, g.L)));
    HIPCHK(c, dalloc(&c->d_pvx, S * fw2 * g.ny));
    HIPCHK(c, dalloc(&c->d_pvy, S * fw2 * g.nx));
    HIPCHK(c, dalloc(&c->d_seis, S * c->ntr * c->nt_max));
    HIPCHK(c, dalloc(&c->d_resid, S * c->ntr * c->nt_max));
    HIPCHK(c, dalloc(&c->d_Ws, S * nxny));
    HIPCHK(c, dalloc(&c->d_Wr, S * nxny));
    HIPCHK(c, dalloc(&c->d_gu_shot, S * nxny));
    HIPCHK(c, dalloc(&c->d_grho_shot, S * nxny));
    HIPCHK(c, dalloc(&c->d_part, S * c->ntr * 2));
    HIPCHK(c, dalloc(&c->d_shot_map, S));
    c->batch_has_snap = gradient;
    if (gradient) {
        const size_t ns = (size_t)dn_nsnap(c);
        if (c->p.snap_precision == 0) {
            HIPCHK(c, dalloc(&c->d_snap, S * ns * 3 * nxny));
        } else {
            HIPCHK(c, dalloc(&c->d_snap16, S * ns * 3 * nxny));
            HIPCHK(c, dalloc(&c->d_snap_scale, S * ns * 3));
            // BF16 is encoded by the storing step itself (a ring of three
            // maxima per field); FP16 needs the snapshot's own maximum
            // first, so it goes through an FP32 staging plane
            if (c->p.snap_precision == 2) HIPCHK(c, dalloc(&c->d_stage, S * 3 * nxny));
            HIPCHK(c, dalloc(&c->d_stage_max, S * 9));
        }
    }
    c->batch_cap = cap;
    c->batch_nt_max = c->nt_max;
    return 0;
}

// rows per segment: enough waves to fill 256 CUs several times over (about
// 36 per CU: the latency of the row loads is hidden by many resident waves
// rather than by long segments; 1200x300, 20 shots, FDORDER 8 -> 16 rows)
// Rows per segment of the fused step.  A segment of R rows costs its wave
// R + 2K - 1 velocity rows (the first 2K - 1 only fill the register
// windows), and the segments of all (shot, strip) pairs run as workgroups
// of kStripsPerGroup waves, two per CU at the step kernel's four waves per
// SIMD, each XCD taking an equal share of them (the kernel's block remap):
// the launch lasts about as many row times as its fullest CU walks,
// rounds x (R + 2K - 1) with rounds = the workgroups of an XCD over its
// slots rounded up.  The smallest such cost wins; a part-filled last round
// costs as much as a full one.
static int choose_rows(const dn_ctx *c, int nshots, int nstrips)
{
    if (c->rows_req > 0) return std::min(c->rows_req, c->g.ny);
    static int cus = 0;
    if (cus <= 0) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
            cus = n;
        else
            cus = 256;
    }
    const int ny = c->g.ny, K = c->K;
    const long long xcd_slots = std::max(1LL, 2LL * cus / 8);   // 8 XCDs, two workgroups per CU
    const int rmin = std::min(ny, 2 * K);
    int best = ny;
    long long best_cost = -1;
    for (int segs = 1; segs <= (ny + rmin - 1) / rmin; ++segs) {
        const int R = (ny + segs - 1) / segs;       // the fewest rows giving this many segments
        if (R < rmin || (ny + R - 1) / R != segs) continue;
        const long long wgs = ((long long)nshots * nstrips * segs + kStripsPerGroup - 1) / kStripsPerGroup;
        const long long per_xcd = (wgs + 7) / 8;
        const long long rounds = (per_xcd + xcd_slots - 1) / xcd_slots;
        const long long cost = rounds * (R + 2 * K - 1);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = R;
        }
    }
    return best;
}

void dn_fill_step_common(dn_ctx *c, StepArgs &a, int n)
{
    const Geo &g = c->g;
    a = StepArgs{};
    a.g = g;
    a.nshots = n;
    a.nshots_all = c->nshots;
    a.ow = kWave + 1 - 2 * c->K;
    a.nstrips = (g.nx + a.ow - 1) / a.ow;
    a.ngroups = (a.nstrips + kStripsPerGroup - 1) / kStripsPerGroup;
    a.sg = c->sg;
    const int nq = (n + a.sg - 1) / a.sg;                              // shot groups
    const int spg = kStripsPerGroup / a.sg;                            // strips per workgroup
    const int ngs = (a.nstrips + spg - 1) / spg;                       // strip groups
    a.R = (a.sg > 1) ? choose_rows(c, nq * kStripsPerGroup, ngs) : choose_rows(c, n, a.ngroups * kStripsPerGroup);
    a.nsegs = (g.ny + a.R - 1) / a.R;
    a.nblocks = n * a.ngroups * a.nsegs;
    a.nblocks8 = (a.nblocks + 7) / 8;
    a.nblocks_sg = nq * ngs * a.nsegs;
    a.nblocks_sg8 = (a.nblocks_sg + 7) / 8;
    a.nblocks2 = (n >= 2 && c->two_shots) ? ((n + 1) / 2) * a.ngroups * a.nsegs : 0;
    a.nblocks2_8 = (a.nblocks2 + 7) / 8;
    a.DT = c->p.dt;
    a.DH = c->p.dh;
    a.dhi = (c->p.L > 0) ? (float)(1.0 / (double)c->p.dh) : c->p.dt / c->p.dh;
    a.dthalbe = (float)((double)c->p.dt / 2.0);
    for (int i = 0; i < 7; ++i) a.hc[i] = c->hc[i];
    a.hcy6 = c->hc[5];  // update_v_PML.c:469 (Q6)
    a.pml = c->pml;
    a.bfield = c->d_b;
    a.uip = c->d_uip;
    a.ujp = c->d_ujp;
    a.fipjp = c->d_fipjp;
    a.fcoef = c->d_f;
    a.dip = c->d_dip;
    a.dcoef = c->d_d;
    a.spk = c->d_spk;
    for (int l = 0; l < kMaxL; ++l) {
        a.bip[l] = c->bip[l];
        a.bjm[l] = c->bjm[l];
        a.cip[l] = c->cip[l];
        a.cjm[l] = c->cjm[l];
    }
    a.rq = c->d_rq;
    a.pvx = c->d_pvx;
    a.pvy = c->d_pvy;
    a.shot_map = c->d_shot_map;
    a.status = c->d_status;
    a.shot_nt = c->d_shot_nt;
    a.src_off = c->d_src_off;
    a.src_i = c->d_src_i;
    a.src_j = c->d_src_j;
    a.src_sig = c->d_sig;
    a.nt_max = c->nt_max;
    a.rec_row = c->d_rec_row;
    a.rec_tab = c->d_rec_tab;
    a.ntr = c->ntr;
    a.dtinv = c->p.dtinv;
}

int dn_zero_state(dn_ctx *c, int n)
{
    const Geo &g = c->g;
    const size_t P = (size_t)g.plane, S = (size_t)n, fw2 = (size_t)(2 * g.fw);
    for (int k = 0; k < 2; ++k) {
        HIPCHK(c, launch_fill(c->d_vy[k], S * P, 0.0f, c->stream));
        HIPCHK(c, launch_fill(c->d_sxy[k], S * P, 0.0f, c->stream));
        HIPCHK(c, launch_fill(c->d_syz[k], S * P, 0.0f, c->stream));
        HIPCHK(c, launch_fill(c->d_psx[k], S * fw2 * g.ny, 0.0f, c->stream));
        HIPCHK(c, launch_fill(c->d_psy[k], S * fw2 * g.nx, 0.0f, c->stream));
    }
    if (g.L > 0) {
        HIPCHK(c, launch_fill(c->d_rq, S * P * 2 * g.L, 0.0f, c->stream));
    }
    HIPCHK(c, launch_fill(c->d_pvx, S * fw2 * g.ny, 0.0f, c->stream));
    HIPCHK(c, launch_fill(c->d_pvy, S * fw2 * g.nx, 0.0f, c->stream));
    return 0;
}

// Enqueue `count` time steps starting at step nt0 (absolute, or an offset to
// the device counter nt_base inside a captured chunk).  cur selects the
// ping-pong half holding the current fields and is advanced per step.
static int enqueue_steps(dn_ctx *c, StepArgs &a, int n, bool adjoint, bool pack_mode, int nsnap, int nt0, int count,
                         const int *nt_base, int &cur, const std::vector<char> &corr_res, bool fwd_extra)
{
    const int kind = c->p.snap_precision;
    for (int k = 0; k < count; ++k) {
        const int nt = nt0 + k;
        a.nt = nt;
        a.nt_base = nt_base;
        a.vy_in = c->d_vy[cur];
        a.sxy_in = c->d_sxy[cur];
        a.syz_in = c->d_syz[cur];
        a.vy_out = c->d_vy[cur ^ 1];
        a.sxy_out = c->d_sxy[cur ^ 1];
        a.syz_out = c->d_syz[cur ^ 1];
        a.psx_in = c->d_psx[cur];
        a.psx_out = c->d_psx[cur ^ 1];
        a.psy_in = c->d_psy[cur];
        a.psy_out = c->d_psy[cur ^ 1];
        // inside a chunk the counter is a multiple of dtinv, so the store
        // pattern depends on the offset only
        const bool pack = pack_mode && ((nt - 1) % c->p.dtinv) == 0;
        if (pack) HIPCHK(c, launch_fill_u32(c->d_stage_max, (size_t)n * 3, 0u, c->stream));
        // adjoint: shot s correlates at step nt when nt = NT_s (mod dtinv);
        // inside a chunk the counter is a multiple of dtinv
        const int dt = c->p.dtinv;
        const bool extra = adjoint ? corr_res[((nt % dt) + dt) % dt] != 0 : (fwd_extra && ((nt - 1) % dt) == 0);
        HIPCHK(c, launch_step(a, c->K, c->p.L > 0, adjoint, extra, c->stream));
        if (pack)
            HIPCHK(c, launch_snap_pack(n, c->g.nx, c->g.ny, c->d_stage, c->d_stage_max, c->d_snap16, c->d_snap_scale,
                                       a.snap_shot_stride, nsnap, nt_base, nt, c->p.dtinv, kind, c->stream));
        cur ^= 1;
    }
    return 0;
}

static int chunk_steps(const dn_ctx *c)
{
    if (c->graph_steps == 0) return 0;
    const int target = c->graph_steps > 0 ? c->graph_steps : 64;
    const int dt = c->p.dtinv;
    const int l = (dt % 2 == 0) ? dt : 2 * dt;   // lcm(2, dtinv): ping-pong parity and store pattern repeat
    return l * std::max(1, (target + l - 1) / l);
}

// Captured chunk for this batch shape (built on first use): `steps` time
// steps reading the device counter, followed by the counter advance.
static int get_graph(dn_ctx *c, StepArgs a, int n, bool adjoint, bool store, bool record, bool pack_mode,
                     int nsnap, int steps, const std::vector<char> &corr_res, bool fwd_extra, hipGraphExec_t *out)
{
    const int kind = c->p.snap_precision;
    const bool hess = a.hess != 0;
    for (auto &gr : c->graphs)
        if (gr.n == n && gr.R == a.R && gr.steps == steps && gr.kind == kind && gr.adjoint == adjoint &&
            gr.store == store && gr.record == record && gr.hess == hess && gr.corr_res == corr_res) {
            *out = gr.exec;
            return 0;
        }
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int cur = 0;
    int rc = enqueue_steps(c, a, n, adjoint, pack_mode, nsnap, 1, steps, c->d_ntbase, cur, corr_res, fwd_extra);
    if (rc == 0 && launch_advance(c->d_ntbase, steps, c->stream) != hipSuccess)
        rc = dn_fail(c, "graph capture: counter advance failed");
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return dn_fail(c, "graph capture failed: %s", hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return dn_fail(c, "graph instantiate failed: %s", hipGetErrorString(ei));
    c->graphs.push_back({n, a.R, steps, kind, adjoint, store, record, hess, corr_res, exec});
    *out = exec;
    return 0;
}

// One pass (forward or adjoint) over the given shots, resident in batch
// slots 0..n-1.  store: forward snapshots + Ws; record: forward seismograms.
// The time loop runs as replays of a captured chunk of steps (hipGraph) plus
// eager launches for the remainder.
int dn_run_pass(dn_ctx *c, const std::vector<int> &shots, bool adjoint, bool store, bool record)
{
    const int n = (int)shots.size();
    if (n < 1 || n > c->batch_cap) return dn_fail(c, "dn_run_pass: %d shots for a batch of %d", n, c->batch_cap);
    int ntmax = 0;
    for (int s = 0; s < n; ++s) ntmax = std::max(ntmax, c->shot_nt[shots[s]]);
    if (c->nt_limit > 0) ntmax = std::min(ntmax, c->nt_limit);
    if (c->cur_map != shots) {
        HIPCHK(c, hipMemcpyAsync(c->d_shot_map, shots.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
        c->cur_map = shots;
    }
    if (dn_zero_state(c, n)) return -1;
    StepArgs a;
    dn_fill_step_common(c, a, n);
    const size_t nxny = (size_t)c->g.nx * c->g.ny;
    const int nsnap = dn_nsnap(c);
    const int kind = c->p.snap_precision;
    if (!adjoint) {
        a.seis = record ? c->d_seis : nullptr;
        if (record) HIPCHK(c, launch_fill(c->d_seis, (size_t)n * c->ntr * c->nt_max, 0.0f, c->stream));
        a.store_mode = store ? (kind == 0 ? 1 : (kind == 1 ? 3 : 2)) : 0;
        if (store && kind == 1) HIPCHK(c, launch_fill_u32(c->d_stage_max, (size_t)n * 9, 0u, c->stream));
        a.wacc = c->d_Ws;
        if (store) HIPCHK(c, launch_fill(c->d_Ws, (size_t)n * nxny, 0.0f, c->stream));
        if (c->hess_mode) {
            a.hess = 1;
            a.grad_u = c->d_gu_shot;
            a.grad_rho = c->d_grho_shot;
            HIPCHK(c, launch_fill(c->d_gu_shot, (size_t)n * nxny, 0.0f, c->stream));
            HIPCHK(c, launch_fill(c->d_grho_shot, (size_t)n * nxny, 0.0f, c->stream));
        }
    } else {
        a.resid = c->d_resid;
        a.wacc = c->d_Wr;
        a.grad_u = c->d_gu_shot;
        a.grad_rho = c->d_grho_shot;
        HIPCHK(c, launch_fill(c->d_Wr, (size_t)n * nxny, 0.0f, c->stream));
        HIPCHK(c, launch_fill(c->d_gu_shot, (size_t)n * nxny, 0.0f, c->stream));
        HIPCHK(c, launch_fill(c->d_grho_shot, (size_t)n * nxny, 0.0f, c->stream));
    }
    a.snap = c->d_snap;
    a.snap16 = c->d_snap16;
    a.snap_scale = c->d_snap_scale;
    a.snap_kind = kind;
    a.snap_shot_stride = (long long)nsnap * 3 * (long long)nxny;
    a.stage = c->d_stage;
    a.stage_max = c->d_stage_max;
    const bool pack_mode = !adjoint && store && kind == 2;
    std::vector<char> corr_res(c->p.dtinv, 0);
    for (int sh : shots) corr_res[c->shot_nt[sh] % c->p.dtinv] = 1;
    const bool fwd_extra = store || c->hess_mode;
    const int CH = chunk_steps(c);
    const int nchunks = (CH > 0 && ntmax >= 2 * CH) ? ntmax / CH : 0;
    hipGraphExec_t exec = nullptr;
    if (nchunks > 0) {
        if (!c->d_ntbase) HIPCHK(c, hipMalloc((void **)&c->d_ntbase, 256));
        if (get_graph(c, a, n, adjoint, store, record, pack_mode, nsnap, CH, corr_res, fwd_extra, &exec)) return -1;
    }
    HIPCHK(c, launch_fill_u32((unsigned int *)c->d_status, 4, 0u, c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    int cur = 0;
    if (nchunks > 0) {
        HIPCHK(c, launch_fill_u32((unsigned int *)c->d_ntbase, 1, 0u, c->stream));
        for (int k = 0; k < nchunks; ++k) HIPCHK(c, hipGraphLaunch(exec, c->stream));
        c->graph_replays += nchunks;
    }
    const int done = nchunks * CH;   // even: the fields are back in half 0
    if (enqueue_steps(c, a, n, adjoint, pack_mode, nsnap, done + 1, ntmax - done, nullptr, cur, corr_res, fwd_extra))
        return -1;
    c->launches += ntmax;
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->ms_steps += ms;
    long long cells = 0;
    for (int s = 0; s < n; ++s) cells += (long long)c->g.nx * c->g.ny * std::min(c->shot_nt[shots[s]], ntmax);
    c->cell_updates += cells;
    c->ms_pass[adjoint ? 1 : 0] += ms;
    c->launches_pass[adjoint ? 1 : 0] += ntmax;
    c->cells_pass[adjoint ? 1 : 0] += cells;
    c->cur_final = cur;
    // the kernel's latch: NaN at the probe cell after some step
    // (denise.c:1183-1185 probes that cell before every step), or a shot map
    // the batch did not upload
    int status[4] = {0, 0, 0, 0};
    HIPCHK(c, hipMemcpy(status, c->d_status, sizeof(status), hipMemcpyDeviceToHost));
    if (status[2]) return dn_fail(c, "step kernel: a batch slot held no valid shot (shot map not uploaded)");
    if (status[0])
        return dn_fail(c, "Simulation is unstable ! (time step %d, shot %d)", status[0] + 1, status[1]);
    return 0;
}

extern "C" int dn_set_datacond(dn_ctx *c, const dn_datacond *d)
{
    if (!c) return dn_fail(nullptr, "dn_set_datacond: ctx is NULL");
    dn_ctx::DataCond n{};
    if (d) {
        if (c->nshots < 1 || c->ntr < 1) return dn_fail(c, "dn_set_datacond: set receivers and shots first");
        if (d->offset_mute < 0 || d->offset_mute > 2) return dn_fail(c, "OFFSET_MUTE must be 0, 1 or 2");
        n.on = d->trkill != 0 || d->offset_mute != 0;
        n.trkill = d->trkill;
        n.offset_mute = d->offset_mute;
        n.offsetc = d->offsetc;
        if (d->trkill) {
            if (!d->kill || d->kill_nsrc < c->nshots || d->kill_ntr < c->ntr)
                return dn_fail(c, "TRKILL: the kill list must cover %d shots x %d receivers", c->nshots, c->ntr);
            n.kill_nsrc = d->kill_nsrc;
            n.kill_ntr = d->kill_ntr;
            n.kill.assign(d->kill, d->kill + (size_t)d->kill_nsrc * d->kill_ntr);
        }
        if (d->offset_mute) {
            n.src_x.resize(c->nshots);
            n.src_y.resize(c->nshots);
            for (int s = 0; s < c->nshots; ++s) {
                const int l = c->src_off[s];
                n.src_x[s] = d->src_x ? d->src_x[s] : (float)c->src_i[l] * c->p.dh;
                n.src_y[s] = d->src_y ? d->src_y[s] : (float)c->src_j[l] * c->p.dh;
            }
        }
    }
    c->dc = std::move(n);
    c->dc_dirty = true;
    return 0;
}

// per (shot, receiver) flags of the residual kernel from the conditioning
static int trace_flags(dn_ctx *c, const unsigned char **out)
{
    *out = nullptr;
    if (!c->dc.on) return 0;
    const size_t n = (size_t)c->nshots * c->ntr;
    if (c->dc_dirty || !c->d_trace_flags) {
        const dn_ctx::DataCond &d = c->dc;
        std::vector<unsigned char> f(n, 0);
        const float DH = c->p.dh;
        for (int s = 0; s < c->nshots; ++s)
            for (int r = 0; r < c->ntr; ++r) {
                unsigned char v = 0;
                if (d.trkill) {
                    if (d.kill[(size_t)s * d.kill_ntr + r] == 1) v |= 1;   // calc_res.c:44-47 kill_tmp[ishot][rec]
                    if (d.kill[(size_t)r * d.kill_nsrc + s] == 1) v |= 2;  // calc_energy.c:30-33 kill_tmp[rec][ishot]
                }
                if (d.offset_mute) {
                    // calc_res.c:101-108 in float
                    const float xr = (float)c->rec_i[r] * DH, yr = (float)c->rec_j[r] * DH;
                    const float xs = d.src_x[s], ys = d.src_y[s];
                    const float offset = (float)std::sqrt((double)(((xs - xr) * (xs - xr)) + ((ys - yr) * (ys - yr))));
                    if ((d.offset_mute == 1) && (offset >= d.offsetc)) v |= 1;
                    if ((d.offset_mute == 2) && (offset <= d.offsetc)) v |= 1;
                }
                f[(size_t)s * c->ntr + r] = v;
            }
        dfree_t(c->d_trace_flags);
        HIPCHK(c, hipMalloc((void **)&c->d_trace_flags, n));
        HIPCHK(c, hipMemcpy(c->d_trace_flags, f.data(), n, hipMemcpyHostToDevice));
        c->dc_dirty = false;
    }
    *out = c->d_trace_flags;
    return 0;
}

int dn_batch_residual(dn_ctx *c, const std::vector<int> &shots)
{
    const int n = (int)shots.size();
    const unsigned char *flags = nullptr;
    if (trace_flags(c, &flags)) return -1;
    HIPCHK(c, launch_residual(n, c->nshots, c->ntr, c->nt_max, c->d_shot_map, c->d_shot_nt, c->d_seis, c->d_obs, c->d_resid,
                              c->d_part, c->p.dt, flags, c->stream));
    std::vector<double> part((size_t)n * c->ntr * 2);
    HIPCHK(c, hipMemcpyAsync(part.data(), c->d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int s = 0; s < n; ++s) {
        double l2 = 0.0, en = 0.0;
        for (int r = 0; r < c->ntr; ++r) {
            l2 += part[2 * ((size_t)s * c->ntr + r)];
            en += part[2 * ((size_t)s * c->ntr + r) + 1];
        }
        c->shot_l2[shots[s]] = l2;
        c->shot_energy[shots[s]] = en;
    }
    return 0;
}

void dn_reset_timing(dn_ctx *c)
{
    c->ms_steps = 0.0;
    c->launches = 0;
    c->cell_updates = 0;
    for (int k = 0; k < 2; ++k) {
        c->ms_pass[k] = 0.0;
        c->launches_pass[k] = 0;
        c->cells_pass[k] = 0;
    }
}

int dn_forward_shots(dn_ctx *c, const std::vector<int> &shots, float *seis)
{
    for (size_t b0 = 0; b0 < shots.size();) {
        if (dn_ensure_batch(c, (int)(shots.size() - b0), false)) return -1;
        const size_t n = std::min((size_t)c->batch_cap, shots.size() - b0);
        std::vector<int> part(shots.begin() + b0, shots.begin() + b0 + n);
        if (dn_run_pass(c, part, false, false, true)) return -1;
        if (seis)
            HIPCHK(c, hipMemcpy(seis + b0 * c->ntr * c->nt_max, c->d_seis, n * c->ntr * c->nt_max * sizeof(float),
                                hipMemcpyDeviceToHost));
        if (c->d_obs && dn_batch_residual(c, part)) return -1;
        b0 += n;
    }
    return 0;
}

extern "C" int dn_forward(dn_ctx *c, int shot0, int nshots, float *seis)
{
    if (!c) return -1;
    if (shot0 < 0 || nshots < 1 || shot0 + nshots > c->nshots) return dn_fail(c, "dn_forward: bad shot range");
    dn_reset_timing(c);
    std::vector<int> shots(nshots);
    for (int s = 0; s < nshots; ++s) shots[s] = shot0 + s;
    return dn_forward_shots(c, shots, seis);
}

int dn_gradient_shots(dn_ctx *c, const std::vector<int> &shots, bool accumulate)
{
    if (!c->d_obs) return dn_fail(c, "gradient: observed data not set");
    if (!c->have_model) return dn_fail(c, "gradient: model not set");
    if (!c->C_fixed && dn_update_means(c)) return -1;
    const size_t nxny = (size_t)c->g.nx * c->g.ny;
    if (!accumulate) {
        HIPCHK(c, launch_fill(c->d_gu, nxny, 0.0f, c->stream));
        HIPCHK(c, launch_fill(c->d_grho, nxny, 0.0f, c->stream));
    }
    if (dn_ensure_batch(c, (int)shots.size(), true)) return -1;
    // Uneven record lengths over several batches: the batches take the shots
    // by decreasing length, so that shots of similar length advance together
    // and few launches carry finished shots.  The sums over shots keep the
    // shots' own order: each batch stores its per-shot terms at the shots'
    // ranks and one kernel adds them up in rank order at the end, with the
    // arithmetic of the in-order accumulation (the same bits).
    std::vector<int> order = shots, rank_of;
    bool by_nt = false;
    if (c->order_by_nt && (int)shots.size() > c->batch_cap) {
        int lo = c->shot_nt[shots[0]], hi = lo;
        for (int s : shots) {
            lo = std::min(lo, c->shot_nt[s]);
            hi = std::max(hi, c->shot_nt[s]);
        }
        if (lo != hi) {
            const size_t need = shots.size() * 4 * nxny;
            if (c->terms_cap < need) {
                dfree_t(c->d_terms);
                c->terms_cap = 0;
                if (dalloc(&c->d_terms, need) == hipSuccess)
                    c->terms_cap = need;
                else
                    (void)hipGetLastError();   // no room: the batches keep the shots' order
            }
            if (c->rank_cap < (int)shots.size()) {
                dfree_t(c->d_rank);
                c->rank_cap = 0;
                if (dalloc(&c->d_rank, shots.size()) == hipSuccess)
                    c->rank_cap = (int)shots.size();
                else
                    (void)hipGetLastError();
            }
            by_nt = c->terms_cap >= need && c->rank_cap >= (int)shots.size();
        }
    }
    if (by_nt) {
        std::vector<int> idx(shots.size());
        for (size_t k = 0; k < idx.size(); ++k) idx[k] = (int)k;
        std::stable_sort(idx.begin(), idx.end(),
                         [&](int a, int b) { return c->shot_nt[shots[a]] > c->shot_nt[shots[b]]; });
        rank_of = idx;
        for (size_t k = 0; k < idx.size(); ++k) order[k] = shots[idx[k]];
        // the ranks of the whole order at once; batch b0.. reads d_rank + b0
        HIPCHK(c, hipMemcpyAsync(c->d_rank, rank_of.data(), rank_of.size() * sizeof(int), hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    for (size_t b0 = 0; b0 < order.size();) {
        if (dn_ensure_batch(c, (int)(order.size() - b0), true)) return -1;
        const size_t n = std::min((size_t)c->batch_cap, order.size() - b0);
        std::vector<int> part(order.begin() + b0, order.begin() + b0 + n);
        if (dn_run_pass(c, part, false, true, true)) return -1;
        if (dn_batch_residual(c, part)) return -1;
        if (dn_run_pass(c, part, true, false, false)) return -1;
        if (by_nt) {
            HIPCHK(c, launch_eprecond_terms((int)n, c->g.nx, c->g.ny, c->p.dh, c->d_Ws, c->d_Wr, c->d_gu_shot,
                                            c->d_grho_shot, c->C_vs, c->C_rho, c->d_maxbits, c->p.eprecond,
                                            c->d_rank + b0, c->d_terms, c->stream));
        } else {
            HIPCHK(c, launch_eprecond_accumulate((int)n, c->g.nx, c->g.ny, c->p.dh, c->d_Ws, c->d_Wr,
                                                 c->d_gu_shot, c->d_grho_shot, c->d_gu, c->d_grho, c->C_vs,
                                                 c->C_rho, c->d_maxbits, c->p.eprecond, c->d_ws_sum, c->d_wr_sum,
                                                 b0 == 0 && !accumulate, c->stream));
        }
        b0 += n;
    }
    if (by_nt)
        HIPCHK(c, launch_eprecond_sum((int)shots.size(), c->g.nx, c->g.ny, c->d_terms, c->d_gu, c->d_grho,
                                      c->d_ws_sum, c->d_wr_sum, !accumulate, c->stream));
    HIPCHK(c, launch_chain_rule(c->g, c->g.nx, c->g.ny, c->d_gu, c->d_grho, c->d_vs, c->d_rho, c->p.dt,
                                c->d_gvs_out, c->d_grho_out, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int dn_gradient(dn_ctx *c, int shot0, int nshots, int accumulate, double *l2, double *energy)
{
    if (!c) return -1;
    if (shot0 < 0 || nshots < 1 || shot0 + nshots > c->nshots) return dn_fail(c, "dn_gradient: bad shot range");
    dn_reset_timing(c);
    std::vector<int> shots(nshots);
    for (int s = 0; s < nshots; ++s) shots[s] = shot0 + s;
    if (dn_gradient_shots(c, shots, accumulate != 0)) return -1;
    double L2 = 0.0, EN = 0.0;
    for (int s : shots) {
        L2 += c->shot_l2[s];
        EN += c->shot_energy[s];
    }
    if (l2) *l2 = L2;
    if (energy) *energy = EN;
    return 0;
}

extern "C" int dn_get_energy_maps(dn_ctx *c, float *ws, float *wr)
{
    if (!c) return -1;
    const size_t n = (size_t)c->g.nx * c->g.ny * sizeof(float);
    if (ws) HIPCHK(c, hipMemcpy(ws, c->d_ws_sum, n, hipMemcpyDeviceToHost));
    if (wr) HIPCHK(c, hipMemcpy(wr, c->d_wr_sum, n, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int dn_hessian_probe(dn_ctx *c, int shot0, int nshots, int nprobes, float eps, unsigned int seed,
                                float *diag_vs)
{
    if (!c) return -1;
    if (shot0 < 0 || nshots < 1 || shot0 + nshots > c->nshots || nprobes < 1 || !(eps > 0.0f) || !diag_vs)
        return dn_fail(c, "dn_hessian_probe: bad arguments");
    if (!c->have_model || !c->d_obs) return dn_fail(c, "dn_hessian_probe: model and observed data must be set");
    const size_t P = (size_t)c->g.plane, nxny = (size_t)c->g.nx * c->g.ny;
    float *vs0 = nullptr, *g0 = nullptr, *diag = nullptr;
    auto cleanup = [&]() {
        dfree_t(vs0);
        dfree_t(g0);
        dfree_t(diag);
    };
    if (dalloc(&vs0, P) != hipSuccess || dalloc(&g0, nxny) != hipSuccess || dalloc(&diag, nxny) != hipSuccess) {
        cleanup();
        return dn_fail(c, "dn_hessian_probe: out of device memory");
    }
    std::vector<int> shots(nshots);
    for (int s = 0; s < nshots; ++s) shots[s] = shot0 + s;
    int rc = 0;
    auto step = [&](hipError_t e) {
        if (rc == 0 && e != hipSuccess) rc = dn_fail(c, "dn_hessian_probe: %s", hipGetErrorString(e));
    };
    step(hipMemcpyAsync(vs0, c->d_vs, P * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    if (rc == 0) rc = dn_gradient_shots(c, shots, false);
    step(hipMemcpyAsync(g0, c->d_gvs_out, nxny * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    step(launch_fill(diag, nxny, 0.0f, c->stream));
    for (int k = 0; k < nprobes && rc == 0; ++k) {
        step(launch_probe_perturb(c->g, vs0, c->d_vs, eps, seed, k, c->stream));
        if (rc == 0) rc = dn_prepare_model(c);
        if (rc == 0) rc = dn_gradient_shots(c, shots, false);
        step(launch_probe_accumulate(c->g, vs0, g0, c->d_gvs_out, eps, seed, k, diag, c->stream));
    }
    // restore the model (also after a failure)
    (void)hipMemcpyAsync(c->d_vs, vs0, P * sizeof(float), hipMemcpyDeviceToDevice, c->stream);
    if (dn_prepare_model(c) != 0 && rc == 0) rc = -1;
    std::vector<float> h(nxny);
    if (rc == 0) step(hipMemcpyAsync(h.data(), diag, nxny * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    step(hipStreamSynchronize(c->stream));
    cleanup();
    if (rc) return rc;
    for (size_t t = 0; t < nxny; ++t) diag_vs[t] = h[t] / (float)nprobes;
    return 0;
}

extern "C" int dn_pseudo_hessian(dn_ctx *c, int shot0, int nshots, float *h_vs, float *h_rho)
{
    if (!c) return -1;
    if (shot0 < 0 || nshots < 1 || shot0 + nshots > c->nshots) return dn_fail(c, "dn_pseudo_hessian: bad shot range");
    if (!c->have_model) return dn_fail(c, "dn_pseudo_hessian: model not set");
    const size_t nxny = (size_t)c->g.nx * c->g.ny;
    if (!c->d_hu) HIPCHK(c, dalloc(&c->d_hu, nxny));
    if (!c->d_hr) HIPCHK(c, dalloc(&c->d_hr, nxny));
    dn_reset_timing(c);
    std::vector<int> shots(nshots);
    for (int s = 0; s < nshots; ++s) shots[s] = shot0 + s;
    for (size_t b0 = 0; b0 < shots.size();) {
        if (dn_ensure_batch(c, (int)(shots.size() - b0), false)) return -1;
        const size_t n = std::min((size_t)c->batch_cap, shots.size() - b0);
        std::vector<int> part(shots.begin() + b0, shots.begin() + b0 + n);
        c->hess_mode = true;
        const int rc = dn_run_pass(c, part, false, false, false);
        c->hess_mode = false;
        if (rc) return rc;
        HIPCHK(c, launch_hessian_finish((int)n, c->g.nx, c->g.ny, c->g, c->d_gu_shot, c->d_grho_shot, c->d_vs,
                                        c->d_rho, c->d_hu, c->d_hr, b0 == 0, b0 + n == shots.size(), c->stream));
        b0 += n;
    }
    if (h_vs) HIPCHK(c, hipMemcpyAsync(h_vs, c->d_hu, nxny * 4, hipMemcpyDeviceToHost, c->stream));
    if (h_rho) HIPCHK(c, hipMemcpyAsync(h_rho, c->d_hr, nxny * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int dn_get_gradient(dn_ctx *c, float *gvs, float *grho, float *raw_u, float *raw_rho)
{
    if (!c) return -1;
    const size_t n = (size_t)c->g.nx * c->g.ny * sizeof(float);
    if (gvs) HIPCHK(c, hipMemcpy(gvs, c->d_gvs_out, n, hipMemcpyDeviceToHost));
    if (grho) HIPCHK(c, hipMemcpy(grho, c->d_grho_out, n, hipMemcpyDeviceToHost));
    if (raw_u) HIPCHK(c, hipMemcpy(raw_u, c->d_gu, n, hipMemcpyDeviceToHost));
    if (raw_rho) HIPCHK(c, hipMemcpy(raw_rho, c->d_grho, n, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int dn_forward_fields(dn_ctx *c, int shot, int nsteps, float *vy, float *sxy, float *syz)
{
    if (!c || shot < 0 || shot >= c->nshots || nsteps < 1) return dn_fail(c, "dn_forward_fields: bad arguments");
    if (dn_ensure_batch(c, 1, false)) return -1;
    const int saved = c->shot_nt[shot];
    c->shot_nt[shot] = std::min(nsteps, saved);
    HIPCHK(c, hipMemcpy(c->d_shot_nt + shot, &c->shot_nt[shot], sizeof(int), hipMemcpyHostToDevice));
    const int rc = dn_run_pass(c, std::vector<int>{shot}, false, false, false);
    c->shot_nt[shot] = saved;
    HIPCHK(c, hipMemcpy(c->d_shot_nt + shot, &saved, sizeof(int), hipMemcpyHostToDevice));
    if (rc) return rc;
    const int cur = c->cur_final;
    if (vy && download_padded(c, vy, c->d_vy[cur])) return -1;
    if (sxy && download_padded(c, sxy, c->d_sxy[cur])) return -1;
    if (syz && download_padded(c, syz, c->d_syz[cur])) return -1;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int dn_last_timing(dn_ctx *c, double *ms, long long *launches, long long *cells)
{
    if (!c) return -1;
    if (ms) *ms = c->ms_steps;
    if (launches) *launches = c->launches;
    if (cells) *cells = c->cell_updates;
    return 0;
}

extern "C" int dn_pass_timing(dn_ctx *c, int adjoint, double *ms, long long *launches, long long *cells)
{
    if (!c || adjoint < 0 || adjoint > 1) return -1;
    if (ms) *ms = c->ms_pass[adjoint];
    if (launches) *launches = c->launches_pass[adjoint];
    if (cells) *cells = c->cells_pass[adjoint];
    return 0;
}
