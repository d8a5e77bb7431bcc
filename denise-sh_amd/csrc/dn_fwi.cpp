// Full-waveform-inversion iteration on the device (denise.c:695-2018 with
// line_search.c, calc_mat_change_test.c, calc_opt_step.c, LBFGS.c,
// store_/extract_LBFGS_SH.c, PCG.c, store_/extract_PCG_SH.c, calc_mres.c).
//
// Differences from the reference, all deliberate:
//  * dot products of L-BFGS / PCG and the misfit sums accumulate in FP64
//    (the reference accumulates in float);
//  * the L-BFGS history stays in HBM instead of files, which also removes the
//    reference's store/extract file-name mismatch that aborts iteration 2;
//  * L2_hist entries before the first iteration read as "no history" (the
//    reference reads before its allocation when MIN_ITER < 2).
#include "dn_core.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#pragma clang fp contract(off)

using namespace dn;

namespace {

constexpr int kRedBlocks = 512;

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// FP64 dot product: per-block partial sums, then one block adds them in order
__global__ void k_dot_part(const float *a, const float *b, size_t n, double *part)
{
    __shared__ double sm[4];
    double acc = 0.0;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        acc += (double)a[t] * (double)b[t];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
}
__global__ void k_sum_part(const double *part, int n, double *out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += part[i];
        *out = s;
    }
}

// (a - b) . (a - b) and b . b for the model misfit (calc_mres.c)
__global__ void k_mres_part(const float *u, const float *u0, size_t n, double *part)
{
    __shared__ double s1[4], s2[4];
    double a = 0.0, b = 0.0;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const double d = (double)u[t] - (double)u0[t];
        a += d * d;
        b += (double)u0[t] * (double)u0[t];
    }
    a = wave_sum(a);
    b = wave_sum(b);
    if ((threadIdx.x & 63) == 0) {
        s1[threadIdx.x >> 6] = a;
        s2[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1[0] + s1[1] + s1[2] + s1[3];
        part[2 * blockIdx.x + 1] = s2[0] + s2[1] + s2[2] + s2[3];
    }
}

// y = x * a  (float scale, as the reference's normalisations)
__global__ void k_scale(float *y, const float *x, float a, size_t n)
{
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        y[t] = x[t] * a;
}
// y = a - b
__global__ void k_sub(float *y, const float *a, const float *b, size_t n)
{
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        y[t] = a[t] - b[t];
}
// y = y + x * c  with c in FP64 (two-loop recursion updates)
__global__ void k_axpy(float *y, const float *x, double c, size_t n)
{
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        y[t] = (float)((double)y[t] + c * (double)x[t]);
}
// y = d * x (learned diagonal scaling)
__global__ void k_mul(float *y, const float *d, const float *x, size_t n)
{
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        y[t] = d[t] * x[t];
}
// PCG: dir = new + beta * dir
__global__ void k_pcg_dir(float *dir, const float *nw, float beta, size_t n)
{
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        dir[t] = nw[t] + beta * dir[t];
}
// maximum of max(0, x) (absval = 0) or of |x| (absval = 1) as float bits
__global__ void k_max(const float *x, size_t n, int absval, unsigned int *out)
{
    float m = 0.0f;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const float v = absval ? fabsf(x[t]) : x[t];
        if (v > m) m = v;
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

// calc_mat_change_test.c:123-191 (INVMAT1 1): trial/updated model with the
// revert rules (Q17).  Writes the compact outputs and the padded model.
__global__ void k_mat_change(Geo g, const float *u, const float *rho, const float *gu, const float *grho, float eps_u,
                             float eps_rho, float vsl, float vsu, float rhol, float rhou, int keep_u, int keep_rho,
                             float *unp1, float *rhonp1, float *vs_pad, float *rho_pad)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nx * g.ny) return;
    const float u0 = u[t], r0 = rho[t];
    float un = u0 - eps_u * gu[t];
    float rn = r0 - eps_rho * grho[t];
    if ((un < vsl) && ((double)un > 1e-6)) un = u0;
    if (un > vsu) un = u0;
    if (rn < rhol) rn = r0;
    if (rn > rhou) rn = r0;
    if (keep_u) un = u0;
    if (keep_rho) rn = r0;
    if (un < 0.0f) un = u0;
    if (rn < 0.0f) rn = r0;
    unp1[t] = un;
    rhonp1[t] = rn;
    const int j = t / g.nx + 1, i = t % g.nx + 1;
    vs_pad[gidx(g, j, i)] = un;
    rho_pad[gidx(g, j, i)] = rn;
}

inline unsigned grid_for(size_t n) { return (unsigned)std::min<size_t>(kRedBlocks, (n + 255) / 256); }

}  // namespace

struct dn_fwi_state {
    dn_fwi_opts o{};
    int iter = 1;              // iteration inside the stage
    int iter_true = 1;         // iteration over all stages
    int stage = 1;
    int min_iter_help = 0;
    int lbfgs_ptr = 1;
    std::vector<float> L2_hist;
    size_t nxny = 0;
    float *m_vs = nullptr, *m_rho = nullptr, *m0_vs = nullptr, *m0_rho = nullptr;  // compact [ny][nx]
    float *t_vs = nullptr, *t_rho = nullptr;                                       // trial/updated model
    float *gu = nullptr, *grho = nullptr;                                          // search direction
    float *y = nullptr, *s = nullptr, *q = nullptr, *r = nullptr, *gold = nullptr, *mold = nullptr;  // [rho; u]
    float *pcg_new = nullptr, *pcg_old = nullptr, *pcg_dir = nullptr;              // [u; rho]
    double *part = nullptr, *scalar = nullptr;
    unsigned int *maxbits = nullptr;
    // learned preconditioner: diagonal D of the initial L-BFGS Hessian for
    // the vs part, the learned direction and a spare copy of the plain one
    float *D = nullptr, *r2 = nullptr, *gu_l = nullptr, *grho_l = nullptr, *gu_p = nullptr, *grho_p = nullptr;
    float *feat = nullptr, *pred = nullptr;
    double last_plain_drop = -1.0;
    bool D_valid = false;      // D holds this stage's learned diagonal (fixed mode)
    bool learned_off = false;  // the plain direction replaced the learned one in this stage
};

static void fwi_free(dn_fwi_state *f)
{
    if (!f) return;
    void *ps[] = {f->m_vs, f->m_rho, f->m0_vs, f->m0_rho, f->t_vs,   f->t_rho,  f->gu,   f->grho,   f->y,
                  f->s,    f->q,     f->r,    f->gold,   f->mold, f->pcg_new, f->pcg_old, f->pcg_dir, f->part,
                  f->scalar, f->maxbits, f->D, f->r2, f->gu_l, f->grho_l, f->gu_p, f->grho_p, f->feat, f->pred};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    delete f;
}

void dn_fwi_destroy(dn_ctx *c)
{
    if (c && c->fwi) {
        fwi_free(c->fwi);
        c->fwi = nullptr;
    }
}

static int dot(dn_ctx *c, const float *a, const float *b, size_t n, double *out)
{
    dn_fwi_state *f = c->fwi;
    hipLaunchKernelGGL(k_dot_part, dim3(kRedBlocks), dim3(256), 0, c->stream, a, b, n, f->part);
    hipLaunchKernelGGL(k_sum_part, dim3(1), dim3(64), 0, c->stream, f->part, kRedBlocks, f->scalar);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(out, f->scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

static int maxval(dn_ctx *c, const float *x, size_t n, int absval, float *out)
{
    dn_fwi_state *f = c->fwi;
    HIPCHK(c, launch_fill_u32(f->maxbits, 1, 0u, c->stream));
    hipLaunchKernelGGL(k_max, dim3(grid_for(n)), dim3(256), 0, c->stream, x, n, absval, f->maxbits);
    HIPCHK(c, hipGetLastError());
    unsigned int bits = 0;
    HIPCHK(c, hipMemcpyAsync(&bits, f->maxbits, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::memcpy(out, &bits, sizeof(float));
    return 0;
}

// calc_mat_change_test: returns EPSILON_u; writes trial model into f->t_* and
// the padded model, then prepares the model coefficients.
static int mat_change(dn_ctx *c, float eps_scale, float *eps_u_out, float *eps_rho_out)
{
    dn_fwi_state *f = c->fwi;
    const size_t n = f->nxny;
    float umax, gmax_u, rhomax, gmax_rho;
    if (maxval(c, f->m_vs, n, 0, &umax) || maxval(c, f->gu, n, 1, &gmax_u) || maxval(c, f->m_rho, n, 0, &rhomax) ||
        maxval(c, f->grho, n, 1, &gmax_rho))
        return -1;
    const dn_fwi_opts &o = f->o;
    float eps_u = eps_scale * (umax / gmax_u);
    if (f->iter < o.inv_vs_iter) eps_u = 0.0f;
    float eps_rho = eps_scale * (rhomax / gmax_rho) * o.scalerho;
    if (f->iter < o.inv_rho_iter) eps_rho = 0.0f;
    if (!std::isfinite(eps_u)) eps_u = 0.0f;
    if (!std::isfinite(eps_rho)) eps_rho = 0.0f;
    hipLaunchKernelGGL(k_mat_change, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->g, f->m_vs,
                       f->m_rho, f->gu, f->grho, eps_u, eps_rho, o.vs_lower, o.vs_upper, o.rho_lower, o.rho_upper,
                       o.inv_vs_iter > f->iter ? 1 : 0, o.inv_rho_iter > f->iter ? 1 : 0, f->t_vs, f->t_rho,
                       c->d_vs, c->d_rho);
    HIPCHK(c, hipGetLastError());
    if (dn_prepare_model(c)) return -1;
    if (eps_u_out) *eps_u_out = eps_u;
    if (eps_rho_out) *eps_rho_out = eps_rho;
    return 0;
}

static std::vector<int> test_shots(const dn_ctx *c)
{
    const dn_fwi_opts &o = c->fwi->o;
    std::vector<int> v;
    const int inc = std::max(1, o.testshot_incr);
    for (int s = o.testshot_start; s <= std::min(o.testshot_end, c->nshots); s += inc)
        if (s >= 1) v.push_back(s - 1);
    return v;
}

static int misfit_of(dn_ctx *c, const std::vector<int> &shots, double *l2, double *en)
{
    double a = 0.0, b = 0.0;
    for (int s : shots) {
        a += c->shot_l2[s];
        b += c->shot_energy[s];
    }
    *l2 = a;
    *en = b;
    return 0;
}

extern "C" int dn_fwi_start(dn_ctx *c, const dn_fwi_opts *o)
{
    if (!c || !o) return dn_fail(c, "dn_fwi_start: NULL argument");
    if (!c->have_model || c->nshots < 1 || !c->d_obs) return dn_fail(c, "dn_fwi_start: set model, shots and data first");
    if (o->grad_method != 1 && o->grad_method != 2) return dn_fail(c, "GRAD_METHOD must be 1 (PCG) or 2 (L-BFGS)");
    if (o->grad_method == 2 && (o->nlbfgs < 1 || o->nlbfgs > 64)) return dn_fail(c, "NLBFGS outside 1..64");
    // the stage's frequency filter on the wavelets and observed data
    if (dn_set_time_filt(c, o->time_filt, o->fc_start, o->fc_end, o->order)) return -1;
    const size_t nxny = (size_t)c->g.nx * c->g.ny;
    dn_fwi_state *f = c->fwi;
    const bool fresh = (f == nullptr);
    if (fresh) {
        // C_vs, C_rho of the starting model, kept for the whole run over all
        // stages (denise.c:746-819, at iter_true == 1)
        if (dn_update_means(c)) return -1;
        f = new dn_fwi_state();
        c->fwi = f;
        f->nxny = nxny;
        f->L2_hist.assign(100000, 0.0f);
        const size_t v2 = 2 * nxny;
        float **vs[] = {&f->m_vs, &f->m_rho, &f->m0_vs, &f->m0_rho, &f->t_vs, &f->t_rho, &f->gu, &f->grho};
        for (float **p : vs) HIPCHK(c, hipMalloc((void **)p, nxny * sizeof(float)));
        float **v2s[] = {&f->q, &f->r, &f->gold, &f->mold, &f->pcg_new, &f->pcg_old, &f->pcg_dir};
        for (float **p : v2s) HIPCHK(c, hipMalloc((void **)p, v2 * sizeof(float)));
        HIPCHK(c, hipMalloc((void **)&f->part, 2 * kRedBlocks * sizeof(double)));
        HIPCHK(c, hipMalloc((void **)&f->scalar, sizeof(double)));
        HIPCHK(c, hipMalloc((void **)&f->maxbits, sizeof(unsigned int)));
        // current and starting model from the context (compact copies)
        const Geo &g = c->g;
        HIPCHK(c, hipMemcpy2DAsync(f->m_vs, g.nx * sizeof(float), c->d_vs + gidx(g, 1, 1), g.pitch * sizeof(float),
                                   g.nx * sizeof(float), g.ny, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpy2DAsync(f->m_rho, g.nx * sizeof(float), c->d_rho + gidx(g, 1, 1), g.pitch * sizeof(float),
                                   g.nx * sizeof(float), g.ny, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(f->m0_vs, f->m_vs, nxny * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(f->m0_rho, f->m_rho, nxny * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        f->min_iter_help = 0;   // denise.c:138: MIN_ITER first applies after a stage change
    } else {
        // a new stage: the reference resets the optimiser (zero_LBFGS / zero_PCG) and iter
        f->stage++;
    }
    if (f->y) (void)hipFree(f->y);
    if (f->s) (void)hipFree(f->s);
    f->y = f->s = nullptr;
    f->o = *o;
    const size_t v2 = 2 * nxny;
    if (o->grad_method == 2) {
        HIPCHK(c, hipMalloc((void **)&f->y, (size_t)o->nlbfgs * v2 * sizeof(float)));
        HIPCHK(c, hipMalloc((void **)&f->s, (size_t)o->nlbfgs * v2 * sizeof(float)));
        HIPCHK(c, launch_fill(f->y, (size_t)o->nlbfgs * v2, 0.0f, c->stream));
        HIPCHK(c, launch_fill(f->s, (size_t)o->nlbfgs * v2, 0.0f, c->stream));
    }
    float *zs[] = {f->q, f->r, f->gold, f->mold, f->pcg_new, f->pcg_old, f->pcg_dir};
    for (float *z : zs) HIPCHK(c, launch_fill(z, v2, 0.0f, c->stream));
    f->iter = 1;
    f->lbfgs_ptr = 1;
    f->D_valid = false;
    f->learned_off = false;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// L-BFGS (store_LBFGS_SH.c, LBFGS.c, extract_LBFGS_SH.c) on [rho; u] vectors.
// With D (learned preconditioner) a second direction with the initial
// Hessian gamma * diag(1, D) is formed from the same history into
// f->grho_l / f->gu_l.
static int lbfgs_direction(dn_ctx *c, const float *D)
{
    dn_fwi_state *f = c->fwi;
    const size_t n = f->nxny, v2 = 2 * n;
    const int NL = f->o.nlbfgs, iter = f->iter, ptr = f->lbfgs_ptr;
    const unsigned gb = grid_for(v2);
    // normalisation (store_LBFGS_SH.c:39-61): waveconv *= C; gradp = waveconv
    hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, c->stream, f->gu, f->gu, c->C_vs, n);
    hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, c->stream, f->grho, f->grho, c->C_rho, n);
    HIPCHK(c, hipGetLastError());
    if (iter > 1) {
        float *yk = f->y + (size_t)(ptr - 1) * v2, *sk = f->s + (size_t)(ptr - 1) * v2;
        hipLaunchKernelGGL(k_sub, dim3(grid_for(n)), dim3(256), 0, c->stream, yk, f->grho, f->gold, n);
        hipLaunchKernelGGL(k_sub, dim3(grid_for(n)), dim3(256), 0, c->stream, yk + n, f->gu, f->gold + n, n);
        hipLaunchKernelGGL(k_sub, dim3(grid_for(n)), dim3(256), 0, c->stream, sk, f->m_rho, f->mold, n);
        hipLaunchKernelGGL(k_sub, dim3(grid_for(n)), dim3(256), 0, c->stream, sk + n, f->m_vs, f->mold + n, n);
        HIPCHK(c, hipMemcpyAsync(f->q, f->grho, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(f->q + n, f->gu, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        // LBFGS.c:26-95, slot order 20..1 / 1..20 (Q18), dots in FP64
        double ys, yy;
        if (dot(c, yk, sk, v2, &ys) || dot(c, yk, yk, v2, &yy)) return -1;
        const double gamma = ys / yy;
        std::vector<double> rho(NL + 1, 0.0), alpha(NL + 1, 0.0);
        for (int k = 1; k <= NL; ++k) {
            double d;
            if (dot(c, f->y + (size_t)(k - 1) * v2, f->s + (size_t)(k - 1) * v2, v2, &d)) return -1;
            rho[k] = (std::fabs(d) > 0.0) ? 1.0 / d : 0.0;
        }
        for (int k = NL; k >= 1; --k) {
            double d;
            if (dot(c, f->s + (size_t)(k - 1) * v2, f->q, v2, &d)) return -1;
            alpha[k] = rho[k] * d;
            hipLaunchKernelGGL(k_axpy, dim3(gb), dim3(256), 0, c->stream, f->q, f->y + (size_t)(k - 1) * v2,
                               -alpha[k], v2);
        }
        // r = gamma * q in FP64
        HIPCHK(c, launch_fill(f->r, v2, 0.0f, c->stream));
        hipLaunchKernelGGL(k_axpy, dim3(gb), dim3(256), 0, c->stream, f->r, f->q, gamma, v2);
        for (int k = 1; k <= NL; ++k) {
            double d;
            if (dot(c, f->y + (size_t)(k - 1) * v2, f->r, v2, &d)) return -1;
            const double beta = rho[k] * d;
            hipLaunchKernelGGL(k_axpy, dim3(gb), dim3(256), 0, c->stream, f->r, f->s + (size_t)(k - 1) * v2,
                               alpha[k] - beta, v2);
        }
        HIPCHK(c, hipGetLastError());
        if (D) {
            // r2 = gamma_D * [q_rho; D q_u], then the same second loop, with
            // the scaling of the initial Hessian diag(1, D) taken from the
            // newest pair: gamma_D = s'y / (y' diag(1, D) y)
            hipLaunchKernelGGL(k_mul, dim3(grid_for(n)), dim3(256), 0, c->stream, f->r2 + n, D, yk + n, n);
            double yDy_u, yy_rho;
            if (dot(c, yk + n, f->r2 + n, n, &yDy_u) || dot(c, yk, yk, n, &yy_rho)) return -1;
            const double yDy = yDy_u + yy_rho;
            const double gamma_d = (yDy > 0.0) ? ys / yDy : gamma;
            hipLaunchKernelGGL(k_mul, dim3(grid_for(n)), dim3(256), 0, c->stream, f->r2 + n, D, f->q + n, n);
            HIPCHK(c, hipMemcpyAsync(f->r2, f->q, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
            hipLaunchKernelGGL(k_scale, dim3(gb), dim3(256), 0, c->stream, f->r2, f->r2, (float)gamma_d, v2);
            for (int k = 1; k <= NL; ++k) {
                double d;
                if (dot(c, f->y + (size_t)(k - 1) * v2, f->r2, v2, &d)) return -1;
                const double beta = rho[k] * d;
                hipLaunchKernelGGL(k_axpy, dim3(gb), dim3(256), 0, c->stream, f->r2, f->s + (size_t)(k - 1) * v2,
                                   alpha[k] - beta, v2);
            }
            HIPCHK(c, hipGetLastError());
        }
    }
    // save old model and normalised gradient (extract_LBFGS_SH.c:68-178)
    HIPCHK(c, hipMemcpyAsync(f->gold, f->grho, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(f->gold + n, f->gu, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(f->mold, f->m_rho, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(f->mold + n, f->m_vs, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    if (D && iter == 1) {
        // first iteration: the (normalised) gradient, scaled by D for vs
        HIPCHK(c, hipMemcpyAsync(f->grho_l, f->grho, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        hipLaunchKernelGGL(k_mul, dim3(grid_for(n)), dim3(256), 0, c->stream, f->gu_l, D, f->gu, n);
        HIPCHK(c, hipGetLastError());
    }
    if (iter > 1) {
        // extract and denormalise (extract_LBFGS_SH.c:31-64)
        hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, c->stream, f->grho, f->r, c->C_rho, n);
        hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, c->stream, f->gu, f->r + n, c->C_vs, n);
        if (D) {
            hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, c->stream, f->grho_l, f->r2, c->C_rho, n);
            hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, c->stream, f->gu_l, f->r2 + n, c->C_vs, n);
        }
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// PCG (descent.c, store_PCG_SH.c, PCG.c, extract_PCG_SH.c) on [u; rho]
static int pcg_direction(dn_ctx *c)
{
    dn_fwi_state *f = c->fwi;
    const size_t n = f->nxny;
    hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, c->stream, f->pcg_new, f->gu, -1.0f, n);
    hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, c->stream, f->pcg_new + n, f->grho, -1.0f, n);
    HIPCHK(c, hipGetLastError());
    float beta[2];
    for (int k = 0; k < 2; ++k) {
        const float *nw = f->pcg_new + (size_t)k * n, *od = f->pcg_old + (size_t)k * n;
        double nn, no, oo;
        if (dot(c, nw, nw, n, &nn) || dot(c, nw, od, n, &no) || dot(c, od, od, n, &oo)) return -1;
        const double betaz = nn - no, betan = oo;  // Polak-Ribiere, PCG.c:43-44
        double b = (betan != 0.0) ? betaz / betan : 0.0;
        if (b < 0.0) b = 0.0;
        beta[k] = (float)b;
    }
    for (int k = 0; k < 2; ++k)
        hipLaunchKernelGGL(k_pcg_dir, dim3(grid_for(n)), dim3(256), 0, c->stream, f->pcg_dir + (size_t)k * n,
                           f->pcg_new + (size_t)k * n, beta[k], n);
    HIPCHK(c, hipMemcpyAsync(f->pcg_old, f->pcg_new, 2 * n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    // extract the direction and flip back to a gradient direction
    hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, c->stream, f->gu, f->pcg_dir, -1.0f, n);
    hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, c->stream, f->grho, f->pcg_dir + n, -1.0f, n);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// calc_opt_step.c:56-94 (Gauss elimination of solvelin.c), float as the reference
static float calc_opt_step(const float *L2t, const float *epst)
{
    float A[4][4], b[4], x[4] = {0, 0, 0, 0}, a, cc, opteps;
    for (int i = 1; i <= 3; i++) {
        A[i][3] = epst[i] * epst[i];
        A[i][2] = epst[i];
        A[i][1] = 1.0f;
        b[i] = L2t[i];
    }
    for (int k = 1; k <= 2; k++)
        for (int n = k; n <= 2; n++) {
            a = A[n + 1][k] / A[k][k];
            for (int m = 1; m <= 3; m++) A[n + 1][m] = A[n + 1][m] - a * A[k][m];
            b[n + 1] = b[n + 1] - a * b[k];
        }
    for (int k = 3; k >= 1; k--) {
        cc = b[k];
        for (int m = 3; m >= k + 1; m--) cc = cc - A[k][m] * x[m];
        x[k] = cc / A[k][k];
    }
    opteps = (float)(-(double)x[2] / (2.0 * (double)x[3]));
    if (((2.0 * x[3]) < 0.0) && (L2t[2] > L2t[3])) opteps = epst[3];
    if (opteps > epst[3]) opteps = epst[3];
    if (opteps < 0.0f) opteps = epst[1];
    return opteps;
}

// line_search.c:36-381
static int line_search(dn_ctx *c, double energy_sum, float *L2t, float *epst1, int *step1, int *step3,
                       float *eps_out, int *nforward)
{
    dn_fwi_state *f = c->fwi;
    const dn_fwi_opts &o = f->o;
    const std::vector<int> tshots = test_shots(c);
    float eps_scale = o.eps_scale, scalefac = o.scalefac;
    int step2 = 0, countstep = 0, itests = 2, iteste = 2;
    *step1 = 0;
    *step3 = 0;
    while ((step2 != 1) || (*step1 != 1)) {
        for (int itest = itests; itest <= iteste; itest++) {
            if (mat_change(c, eps_scale, nullptr, nullptr)) return -1;
            if (dn_forward_shots(c, tshots, nullptr)) return -1;
            *nforward += (int)tshots.size();
            double l2, en;
            misfit_of(c, tshots, &l2, &en);
            epst1[itest] = eps_scale;
            epst1[1] = 0.0f;
            L2t[itest] = (float)(l2 / energy_sum);
        }
        if ((*step1 == 0) && (L2t[1] <= L2t[2])) {
            eps_scale = eps_scale / scalefac;
            countstep++;
        }
        if ((*step1 == 1) && (L2t[2] < L2t[3])) {
            epst1[3] = eps_scale;
            step2 = 1;
        }
        if ((*step1 == 1) && (L2t[2] >= L2t[3])) {
            epst1[3] = eps_scale;
            eps_scale = eps_scale + (eps_scale / scalefac);
            countstep++;
        }
        if ((*step1 == 0) && (L2t[1] > L2t[2])) {
            epst1[2] = eps_scale;
            *step1 = 1;
            iteste = 3;
            itests = 3;
            countstep = 0;
            eps_scale = eps_scale + (eps_scale / scalefac);
        }
        *step3 = 0;
        if ((*step1 == 0) && (countstep > o.stepmax)) {
            *step3 = 1;
            break;
        }
        if ((*step1 == 1) && (countstep > o.stepmax)) {
            *step1 = 1;
            step2 = 1;
        }
    }
    if (*step1 == 1) eps_scale = calc_opt_step(L2t, epst1);
    *eps_out = eps_scale;
    return 0;
}

// ---------------------------------------------------------------------------
// Learned preconditioner: input maps and the diagonal D
// ---------------------------------------------------------------------------
static int features_host(dn_ctx *c, std::vector<float> &x, std::vector<float> &ws, std::vector<float> &wr, int &H8,
                         int &W8)
{
    const Geo &g = c->g;
    const int nx = g.nx, ny = g.ny;
    const size_t nxny = (size_t)nx * ny;
    H8 = (ny + 7) / 8 * 8;
    W8 = (nx + 7) / 8 * 8;
    ws.resize(nxny);
    wr.resize(nxny);
    std::vector<float> vs(nxny);
    HIPCHK(c, hipMemcpy(ws.data(), c->d_ws_sum, nxny * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(wr.data(), c->d_wr_sum, nxny * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy2D(vs.data(), nx * sizeof(float), c->d_vs + gidx(g, 1, 1), g.pitch * sizeof(float),
                          nx * sizeof(float), ny, hipMemcpyDeviceToHost));
    double mws = 0.0, mwr = 0.0, svs = 0.0;
    for (size_t t = 0; t < nxny; ++t) {
        mws = std::max(mws, (double)ws[t]);
        mwr = std::max(mwr, (double)wr[t]);
        svs += vs[t];
    }
    if (!(mws > 0.0) || !(mwr > 0.0)) return dn_fail(c, "learned preconditioner: energy maps are empty (run a gradient)");
    const double mvs = svs / (double)nxny;
    x.assign((size_t)3 * H8 * W8, 0.0f);
    for (int j = 0; j < H8; ++j)
        for (int i = 0; i < W8; ++i) {
            const size_t o = (size_t)std::min(j, ny - 1) * nx + std::min(i, nx - 1);
            const size_t d = (size_t)j * W8 + i;
            x[d] = (float)(std::log10(std::max((double)ws[o] / mws, 1e-12)) / 12.0 + 1.0);
            x[(size_t)H8 * W8 + d] = (float)(std::log10(std::max((double)wr[o] / mwr, 1e-12)) / 12.0 + 1.0);
            x[(size_t)2 * H8 * W8 + d] = (float)(vs[o] / mvs - 1.0);
        }
    return 0;
}

extern "C" int dn_precond_features(dn_ctx *c, float *x)
{
    if (!c || !x) return dn_fail(c, "dn_precond_features: NULL argument");
    std::vector<float> h, ws, wr;
    int H8, W8;
    if (features_host(c, h, ws, wr, H8, W8)) return -1;
    std::memcpy(x, h.data(), h.size() * sizeof(float));
    return 0;
}

extern "C" int dn_fwi_set_learned_fixed(dn_ctx *c, int fixed)
{
    if (!c) return -1;
    c->learned_fixed = fixed != 0;
    if (c->fwi) c->fwi->D_valid = false;
    return 0;
}

extern "C" int dn_fwi_set_learned_revert(dn_ctx *c, int per_stage)
{
    if (!c) return -1;
    c->learned_revert = per_stage != 0;
    if (c->fwi) c->fwi->learned_off = false;
    return 0;
}

extern "C" int dn_fwi_set_learned_power(dn_ctx *c, float power)
{
    if (!c) return -1;
    if (!(power > 0.0f && power <= 2.0f)) return dn_fail(c, "dn_fwi_set_learned_power: power must be in (0, 2]");
    c->learned_power = power;
    if (c->fwi) c->fwi->D_valid = false;
    return 0;
}

extern "C" int dn_fwi_set_learned(dn_ctx *c, dn_unet *net, int judge_every)
{
    if (!c) return -1;
    if (judge_every < -1000) return dn_fail(c, "dn_fwi_set_learned: judge_every must be >= -1000");
    c->unet = net;
    c->judge_every = judge_every;
    return 0;
}

// D = (We / H_pred)^power (We = sqrt(sum Ws * sum Wr) + 0.005 max, the
// energy weight of eprecond1.c; H_pred = 10^net; power 1 by default),
// normalised to mean 1 and kept in [1e-2, 1e2]: the learned diagonal
// relative to the plain preconditioning
static int learned_diagonal(dn_ctx *c)
{
    dn_fwi_state *f = c->fwi;
    const Geo &g = c->g;
    const size_t nxny = f->nxny;
    std::vector<float> x, ws, wr;
    int H8, W8;
    if (features_host(c, x, ws, wr, H8, W8)) return -1;
    const size_t P8 = (size_t)H8 * W8;
    if (!f->feat) HIPCHK(c, hipMalloc((void **)&f->feat, 3 * P8 * sizeof(float)));
    if (!f->pred) HIPCHK(c, hipMalloc((void **)&f->pred, P8 * sizeof(float)));
    HIPCHK(c, hipMemcpy(f->feat, x.data(), 3 * P8 * sizeof(float), hipMemcpyHostToDevice));
    if (dn_unet_forward_device(c->unet, 1, H8, W8, f->feat, f->pred)) return dn_fail(c, "%s", dn_last_error(nullptr));
    std::vector<float> pred(P8), D(nxny);
    HIPCHK(c, hipMemcpy(pred.data(), f->pred, P8 * sizeof(float), hipMemcpyDeviceToHost));
    double wmax = 0.0;
    for (size_t t = 0; t < nxny; ++t) wmax = std::max(wmax, std::sqrt((double)ws[t] * (double)wr[t]));
    double sum = 0.0;
    for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
            const size_t t = (size_t)j * g.nx + i;
            const double we = std::sqrt((double)ws[t] * (double)wr[t]) + 0.005 * wmax;
            const double h = std::pow(10.0, std::min(30.0, std::max(-30.0, (double)pred[(size_t)j * W8 + i])));
            D[t] = (float)(c->learned_power == 1.0 ? we / h : std::pow(we / h, c->learned_power));
            sum += D[t];
        }
    const double mean = sum / (double)nxny;
    for (float &d : D) d = (float)std::min(1e2, std::max(1e-2, (double)d / mean));
    if (!f->D) HIPCHK(c, hipMalloc((void **)&f->D, nxny * sizeof(float)));
    HIPCHK(c, hipMemcpy(f->D, D.data(), nxny * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int dn_fwi_iteration(dn_ctx *c, dn_fwi_record *rec)
{
    if (!c || !c->fwi) return dn_fail(c, "dn_fwi_iteration: call dn_fwi_start first");
    dn_fwi_state *f = c->fwi;
    dn_fwi_record R{};
    R.stage = f->stage;
    R.iter = f->iter;
    R.iter_true = f->iter_true;
    const size_t n = f->nxny;
    if (f->o.grad_method == 2) {
        if (f->iter > 2) f->lbfgs_ptr++;
        if (f->lbfgs_ptr > f->o.nlbfgs) f->lbfgs_ptr = 1;
    }
    // current model into the context (coefficients for the gradient pass)
    {
        const Geo &g = c->g;
        HIPCHK(c, hipMemcpy2DAsync(c->d_vs + gidx(g, 1, 1), g.pitch * sizeof(float), f->m_vs, g.nx * sizeof(float),
                                   g.nx * sizeof(float), g.ny, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpy2DAsync(c->d_rho + gidx(g, 1, 1), g.pitch * sizeof(float), f->m_rho, g.nx * sizeof(float),
                                   g.nx * sizeof(float), g.ny, hipMemcpyDeviceToDevice, c->stream));
        if (dn_prepare_model(c)) return -1;
    }
    std::vector<int> all(c->nshots);
    for (int s = 0; s < c->nshots; ++s) all[s] = s;
    dn_reset_timing(c);
    if (dn_gradient_shots(c, all, false)) return -1;
    const std::vector<int> tshots = test_shots(c);
    double l2, energy;
    misfit_of(c, tshots, &l2, &energy);
    float L2t[5] = {0, 0, 0, 0, 0}, epst1[4] = {0, 0, 0, 0};
    L2t[1] = L2t[4] = (float)(l2 / energy);
    R.l2 = l2;
    R.energy = energy;
    {  // model misfit (calc_mres.c) in FP64
        double m2[2];
        for (int k = 0; k < 2; ++k) {
            hipLaunchKernelGGL(k_mres_part, dim3(kRedBlocks), dim3(256), 0, c->stream, k ? f->m_rho : f->m_vs,
                               k ? f->m0_rho : f->m0_vs, n, f->part);
            std::vector<double> part(2 * kRedBlocks);
            HIPCHK(c, hipMemcpyAsync(part.data(), f->part, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                                     c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            double a = 0.0, b = 0.0;
            for (int t = 0; t < kRedBlocks; ++t) {
                a += part[2 * t];
                b += part[2 * t + 1];
            }
            m2[k] = (b > 0.0) ? a / b / (double)n : 0.0;
        }
        R.M2_vs = (float)m2[0];
        R.M2_rho = (float)m2[1];
    }
    // direction
    HIPCHK(c, hipMemcpyAsync(f->gu, c->d_gvs_out, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(f->grho, c->d_grho_out, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    // smooth_grad and the tapers (denise.c:1818-1847)
    if (f->o.spatfilter &&
        dn_smooth_grad_device(c, f->o.spatfilter, f->o.wd_damp, f->o.wd_damp1, f->o.fc_end, f->gu, f->grho))
        return -1;
    {
        const int flags[4] = {f->o.taper_vert, f->o.taper_hor, f->o.taper_sources, f->o.taper_file};
        const int sws[4] = {1, 2, 3, 5};
        for (int k = 0; k < 4; ++k) {
            if (!flags[k]) continue;
            dn_taper T = f->o.taper;
            T.sws = sws[k];
            if (dn_taper_grad_device(c, &T, f->gu, f->grho)) return -1;
        }
    }
    if (c->precond_hook && c->precond_hook(c, f->gu, f->grho)) return -1;
    // judge_every < 0: the learned direction only on the first -judge_every
    // iterations of a stage, plain L-BFGS (with the same history) after; the
    // same after the plain direction replaced the learned one (revert mode)
    const bool learned = (c->unet != nullptr) && (f->o.grad_method == 2) &&
                         !(c->judge_every < 0 && f->iter > -c->judge_every) && !f->learned_off;
    if (learned) {
        if (!f->r2) {
            HIPCHK(c, hipMalloc((void **)&f->r2, 2 * n * sizeof(float)));
            float **bs[] = {&f->gu_l, &f->grho_l, &f->gu_p, &f->grho_p};
            for (float **b : bs) HIPCHK(c, hipMalloc((void **)b, n * sizeof(float)));
        }
        if (!(c->learned_fixed && f->D_valid)) {
            if (learned_diagonal(c)) return -1;
            f->D_valid = true;
        }
    }
    if (f->o.grad_method == 1) {
        if (pcg_direction(c)) return -1;
    } else {
        if (lbfgs_direction(c, learned ? f->D : nullptr)) return -1;
    }
    // line search on the test shots
    int step1 = 0, step3 = 0, nfwd = 0;
    float eps_scale = 0.0f;
    R.precond = 0;
    R.judged = 0;
    R.drop_learned = R.drop_plain = -1.0f;
    if (!learned) {
        if (line_search(c, energy, L2t, epst1, &step1, &step3, &eps_scale, &nfwd)) return -1;
    } else {
        // learned direction first; the plain (eprecond) direction judges it
        // on the first iteration of a stage and every judge_every-th one
        // (the better misfit drop of the two is taken), and replaces it
        // whenever its line search finds no decrease
        auto drop_of = [](const float *L, int st1) {
            return st1 ? (double)(L[1] - std::min(L[2], L[3])) / (double)L[1] : 0.0;
        };
        HIPCHK(c, hipMemcpyAsync(f->gu_p, f->gu, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(f->grho_p, f->grho, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(f->gu, f->gu_l, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(f->grho, f->grho_l, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        float L2l[5], epl[4];
        std::memcpy(L2l, L2t, sizeof(L2l));
        std::memcpy(epl, epst1, sizeof(epl));
        int s1l = 0, s3l = 0;
        float epsl = 0.0f;
        if (line_search(c, energy, L2l, epl, &s1l, &s3l, &epsl, &nfwd)) return -1;
        const double dl = drop_of(L2l, s1l);
        R.drop_learned = (float)dl;
        const bool judge = (f->iter == 1) || (c->judge_every > 0 && (f->iter - 1) % c->judge_every == 0);
        const bool fail = (s1l == 0);
        bool take_learned = true;
        if (judge || fail) {
            HIPCHK(c, hipMemcpyAsync(f->gu, f->gu_p, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(f->grho, f->grho_p, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
            float L2p[5], epp[4];
            std::memcpy(L2p, L2t, sizeof(L2p));
            std::memcpy(epp, epst1, sizeof(epp));
            int s1p = 0, s3p = 0;
            float epsp = 0.0f;
            if (line_search(c, energy, L2p, epp, &s1p, &s3p, &epsp, &nfwd)) return -1;
            const double dp = drop_of(L2p, s1p);
            f->last_plain_drop = dp;
            R.judged = 1;
            R.drop_plain = (float)dp;
            take_learned = !fail && dl >= dp;
            if (!take_learned && c->learned_revert) f->learned_off = true;
            if (!take_learned) {
                std::memcpy(L2t, L2p, sizeof(L2p));
                std::memcpy(epst1, epp, sizeof(epp));
                step1 = s1p;
                step3 = s3p;
                eps_scale = epsp;
            } else {
                HIPCHK(c, hipMemcpyAsync(f->gu, f->gu_l, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
                HIPCHK(c, hipMemcpyAsync(f->grho, f->grho_l, n * sizeof(float), hipMemcpyDeviceToDevice,
                                         c->stream));
            }
        }
        if (take_learned) {
            std::memcpy(L2t, L2l, sizeof(L2l));
            std::memcpy(epst1, epl, sizeof(epl));
            step1 = s1l;
            step3 = s3l;
            eps_scale = epsl;
            R.precond = 1;
        }
    }
    float opteps_vs = eps_scale;
    if ((f->iter > f->min_iter_help) && (step1 == 0)) {
        eps_scale = 0.0f;
        opteps_vs = 0.0f;
    }
    // commit (calc_mat_change_test with itest 0)
    float eps_u = 0.0f, eps_rho = 0.0f;
    if (mat_change(c, eps_scale, &eps_u, &eps_rho)) return -1;
    HIPCHK(c, hipMemcpyAsync(f->m_vs, f->t_vs, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(f->m_rho, f->t_rho, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    R.opteps_vs = opteps_vs;
    R.eps_u = eps_u;
    R.eps_rho = eps_rho;
    for (int k = 0; k < 3; ++k) R.epst1[k] = epst1[k + 1];
    for (int k = 0; k < 4; ++k) R.L2t[k] = L2t[k + 1];
    R.step1 = step1;
    R.step3 = step3;
    R.test_forwards = nfwd;
    // abort criterion and stage change (denise.c:1937-2018)
    if (f->iter < (int)f->L2_hist.size()) f->L2_hist[f->iter] = L2t[4];
    R.next_stage = 0;
    if (f->iter > f->min_iter_help) {
        const int k = f->iter - 2;
        if (k >= 1 || step3 == 1) {
            const float prev = (k >= 1) ? f->L2_hist[k] : 0.0f;
            const float diff = (k >= 1) ? std::fabs((prev - f->L2_hist[f->iter]) / prev) : 0.0f;
            if (((k >= 1) && (diff <= f->o.pro)) || (step3 == 1)) {
                f->min_iter_help = f->iter + f->o.min_iter;
                f->iter = 0;
                R.next_stage = 1;
            }
        }
    }
    f->iter++;
    // denise.c:2014-2019: the stage change leaves the iteration loop with
    // break, before iter_true++ (the next stage's first iteration carries the
    // same iter_true)
    if (!R.next_stage) f->iter_true++;
    if (rec) *rec = R;
    return 0;
}

extern "C" int dn_fwi_write_files(dn_ctx *c, const dn_fwi_record *rec, const dn_fwi_files *F)
{
    if (!c || !c->fwi || !rec || !F) return dn_fail(c, "dn_fwi_write_files: NULL argument or no FWI state");
    const Geo &g = c->g;
    const size_t n = c->fwi->nxny;
    if (F->misfit_log) {
        FILE *fp = std::fopen(F->misfit_log, rec->iter_true == 1 ? "w" : "a");
        if (!fp) return dn_fail(c, "Could not open %s", F->misfit_log);
        const float *e = rec->epst1, *L = rec->L2t;
        if (F->time_filt == 0)
            std::fprintf(fp, "%e \t %e \t %e \t %e \t %e \t %e \t %e \t %e \t %e \t  %e \t %d \n", rec->opteps_vs,
                         e[0], e[1], e[2], L[0], L[1], L[2], L[3], rec->M2_vs, rec->M2_rho, rec->stage);
        else
            std::fprintf(fp, "%e \t %e \t %e \t %e \t %e \t %e \t %e \t %e \t %e \t %e \t %f \t %f \t %d \n",
                         rec->opteps_vs, e[0], e[1], e[2], L[0], L[1], L[2], L[3], rec->M2_vs, rec->M2_rho,
                         F->fc_start, F->fc, rec->stage);
        std::fclose(fp);
    }
    std::vector<float> a(n), b(n);
    auto write = [&](const std::string &path, const float *v) { return dn_write_model_file(path.c_str(), g.nx, g.ny, v); };
    if (F->jacobian) {
        HIPCHK(c, hipMemcpy(a.data(), c->fwi->gu, n * sizeof(float), hipMemcpyDeviceToHost));
        if (write(std::string(F->jacobian) + "_grad_vs.bin", a.data())) return -1;
    }
    if (F->jacobian && F->lbfgs_old && c->fwi->o.grad_method == 2) {
        // extract_LBFGS_SH.c:68-178: the model the iteration started from,
        // the normalised gradient there and the direction (H^-1 g), each as
        // the rank file <JACOBIAN>_<name>.old.<POS[1]>.<POS[2]> and the
        // merged <JACOBIAN>_<name>.old (mergemod.c; one domain: POS = 0, 0)
        const dn_fwi_state *f = c->fwi;
        const std::string jac = F->jacobian;
        const struct {
            const char *name;
            const float *v;
        } files[] = {{"_p_vs.old", f->mold + n}, {"_p_u.old", f->gold + n}, {"_c_u.old", f->gu},
                     {"_p_mrho.old", f->mold}, {"_p_rho.old", f->gold}, {"_c_rho.old", f->grho}};
        for (const auto &fl : files) {
            HIPCHK(c, hipMemcpy(a.data(), fl.v, n * sizeof(float), hipMemcpyDeviceToHost));
            if (write(jac + fl.name + ".0.0", a.data()) || write(jac + fl.name, a.data())) return -1;
        }
    }
    if (F->inv_modelfile) {
        const std::string base = F->inv_modelfile;
        HIPCHK(c, hipMemcpy(a.data(), c->fwi->m_vs, n * sizeof(float), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(b.data(), c->fwi->m_rho, n * sizeof(float), hipMemcpyDeviceToHost));
        if (write(base + "_vs.bin", a.data()) || write(base + "_rho.bin", b.data())) return -1;
        if (F->write_it) {
            const std::string it = std::to_string(rec->iter);
            if (write(base + "_vs_it_" + it + ".bin", a.data()) || write(base + "_rho_it_" + it + ".bin", b.data()))
                return -1;
        }
        if (rec->next_stage) {
            const std::string st = std::to_string(rec->stage);
            if (write(base + "_vs_stage_" + st + ".bin", a.data()) || write(base + "_rho_stage_" + st + ".bin", b.data()))
                return -1;
        }
    }
    return 0;
}

extern "C" int dn_fwi_get_model(dn_ctx *c, float *vs, float *rho)
{
    if (!c || !c->fwi) return dn_fail(c, "dn_fwi_get_model: no FWI state");
    const size_t n = c->fwi->nxny * sizeof(float);
    if (vs) HIPCHK(c, hipMemcpy(vs, c->fwi->m_vs, n, hipMemcpyDeviceToHost));
    if (rho) HIPCHK(c, hipMemcpy(rho, c->fwi->m_rho, n, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int dn_fwi_get_direction(dn_ctx *c, float *gvs, float *grho)
{
    if (!c || !c->fwi) return dn_fail(c, "dn_fwi_get_direction: no FWI state");
    const size_t n = c->fwi->nxny * sizeof(float);
    if (gvs) HIPCHK(c, hipMemcpy(gvs, c->fwi->gu, n, hipMemcpyDeviceToHost));
    if (grho) HIPCHK(c, hipMemcpy(grho, c->fwi->grho, n, hipMemcpyDeviceToHost));
    return 0;
}

// ---------------------------------------------------------------------------
// Optimiser state on disk.  The reference keeps the L-BFGS ring and the PCG
// vectors in memory only (a restarted denise.c begins with an empty ring) and
// reads the previous model and normalised gradient back from the files of
// extract_LBFGS_SH.c (store_LBFGS_SH.c:78-112).  dn_fwi_save_state writes
// everything an FWI run needs to continue; dn_fwi_load_state restores it and
// can take the previous model and gradient from those reference-named files.
// ---------------------------------------------------------------------------
namespace {
constexpr char kStateMagic[8] = {'D', 'N', 'F', 'W', 'I', 'S', '0', '2'};

struct StateHead {
    char magic[8];
    int32_t nx, ny, grad_method, nlbfgs, iter, iter_true, stage, min_iter_help, lbfgs_ptr, nhist;
    float C_vs, C_rho;
    double last_plain_drop;
    int32_t learned;   // learned preconditioner: bit 0 the stage's diagonal D follows the
                       // arrays, bit 1 the stage reverted to plain L-BFGS
    int32_t reserved;
};

// the device arrays of the state, in file order
std::vector<std::pair<float *, size_t>> state_arrays(dn_fwi_state *f)
{
    const size_t n = f->nxny, v2 = 2 * n;
    std::vector<std::pair<float *, size_t>> v = {{f->m_vs, n},     {f->m_rho, n},    {f->m0_vs, n},
                                                 {f->m0_rho, n},   {f->gold, v2},    {f->mold, v2},
                                                 {f->pcg_old, v2}, {f->pcg_dir, v2}};
    if (f->o.grad_method == 2) {
        v.push_back({f->y, (size_t)f->o.nlbfgs * v2});
        v.push_back({f->s, (size_t)f->o.nlbfgs * v2});
    }
    return v;
}
}  // namespace

extern "C" int dn_fwi_save_state(dn_ctx *c, const char *path)
{
    if (!c || !c->fwi || !path) return dn_fail(c, "dn_fwi_save_state: no FWI state or no path");
    dn_fwi_state *f = c->fwi;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    StateHead h{};
    std::memcpy(h.magic, kStateMagic, 8);
    h.nx = c->g.nx;
    h.ny = c->g.ny;
    h.grad_method = f->o.grad_method;
    h.nlbfgs = f->o.nlbfgs;
    h.iter = f->iter;
    h.iter_true = f->iter_true;
    h.stage = f->stage;
    h.min_iter_help = f->min_iter_help;
    h.lbfgs_ptr = f->lbfgs_ptr;
    h.nhist = (int32_t)std::min<size_t>(f->L2_hist.size(), (size_t)std::max(0, f->iter) + 1);
    h.C_vs = c->C_vs;
    h.C_rho = c->C_rho;
    h.last_plain_drop = f->last_plain_drop;
    const bool save_D = f->D_valid && f->D != nullptr;
    h.learned = (save_D ? 1 : 0) | (f->learned_off ? 2 : 0);
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return dn_fail(c, "Could not open %s", path);
    bool ok = std::fwrite(&h, sizeof(h), 1, fp) == 1 &&
              std::fwrite(f->L2_hist.data(), sizeof(float), (size_t)h.nhist, fp) == (size_t)h.nhist;
    std::vector<float> buf;
    for (const auto &a : state_arrays(f)) {
        if (!ok) break;
        buf.resize(a.second);
        if (hipMemcpy(buf.data(), a.first, a.second * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
            std::fclose(fp);
            return dn_fail(c, "dn_fwi_save_state: device copy failed");
        }
        ok = std::fwrite(buf.data(), sizeof(float), a.second, fp) == a.second;
    }
    if (ok && save_D) {
        buf.resize(f->nxny);
        if (hipMemcpy(buf.data(), f->D, f->nxny * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
            std::fclose(fp);
            return dn_fail(c, "dn_fwi_save_state: device copy failed");
        }
        ok = std::fwrite(buf.data(), sizeof(float), f->nxny, fp) == f->nxny;
    }
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? 0 : dn_fail(c, "Could not write %s", path);
}

extern "C" int dn_fwi_load_state(dn_ctx *c, const char *path, const char *jacobian, int name_fix)
{
    if (!c || !c->fwi || !path) return dn_fail(c, "dn_fwi_load_state: call dn_fwi_start first, and give a path");
    dn_fwi_state *f = c->fwi;
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return dn_fail(c, "Could not open %s", path);
    StateHead h{};
    if (std::fread(&h, sizeof(h), 1, fp) != 1 || std::memcmp(h.magic, kStateMagic, 8) != 0) {
        std::fclose(fp);
        return dn_fail(c, "%s is not an FWI state file", path);
    }
    if (h.nx != c->g.nx || h.ny != c->g.ny || h.grad_method != f->o.grad_method || h.nlbfgs != f->o.nlbfgs ||
        h.nhist < 0 || h.nhist > (int)f->L2_hist.size()) {
        std::fclose(fp);
        return dn_fail(c, "%s: grid %dx%d, GRAD_METHOD %d, NLBFGS %d do not match this run", path, h.nx, h.ny,
                       h.grad_method, h.nlbfgs);
    }
    bool ok = std::fread(f->L2_hist.data(), sizeof(float), (size_t)h.nhist, fp) == (size_t)h.nhist;
    std::vector<float> buf;
    for (const auto &a : state_arrays(f)) {
        if (!ok) break;
        buf.resize(a.second);
        ok = std::fread(buf.data(), sizeof(float), a.second, fp) == a.second &&
             hipMemcpy(a.first, buf.data(), a.second * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    }
    if (ok && (h.learned & 1)) {
        if (!f->D && hipMalloc((void **)&f->D, f->nxny * sizeof(float)) != hipSuccess) {
            std::fclose(fp);
            return dn_fail(c, "dn_fwi_load_state: device allocation failed");
        }
        buf.resize(f->nxny);
        ok = std::fread(buf.data(), sizeof(float), f->nxny, fp) == f->nxny &&
             hipMemcpy(f->D, buf.data(), f->nxny * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    }
    std::fclose(fp);
    if (!ok) return dn_fail(c, "%s: truncated FWI state file", path);
    if (jacobian) {
        // store_LBFGS_SH.c:78-112 reads the rank files as
        // <JACOBIAN>_<name>.old.%i%i, extract_LBFGS_SH.c writes them as
        // .old.%i.%i: name_fix reads what was written, 0 the reference's name
        const std::string sfx = name_fix ? ".0.0" : ".00";
        const size_t n = f->nxny;
        const struct {
            const char *name;
            float *v;
        } files[] = {{"_p_rho.old", f->gold}, {"_p_mrho.old", f->mold}, {"_p_u.old", f->gold + n},
                     {"_p_vs.old", f->mold + n}};
        buf.resize(n);
        for (const auto &fl : files) {
            const std::string p = std::string(jacobian) + fl.name + sfx;
            if (dn_read_model_file(p.c_str(), c->g.nx, c->g.ny, buf.data()))
                return dn_fail(c, "dn_fwi_load_state: cannot read %s", p.c_str());
            HIPCHK(c, hipMemcpy(fl.v, buf.data(), n * sizeof(float), hipMemcpyHostToDevice));
        }
    }
    f->iter = h.iter;
    f->iter_true = h.iter_true;
    f->stage = h.stage;
    f->min_iter_help = h.min_iter_help;
    f->lbfgs_ptr = h.lbfgs_ptr;
    f->last_plain_drop = h.last_plain_drop;
    f->D_valid = (h.learned & 1) != 0;
    f->learned_off = (h.learned & 2) != 0;
    c->C_vs = h.C_vs;
    c->C_rho = h.C_rho;
    c->C_fixed = true;
    return 0;
}
