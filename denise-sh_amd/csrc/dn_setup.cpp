// Host-side setup of the SH path: FD coefficients (holbergcoeff.c), CPML
// profiles (PML_pro.c), relaxation constants (readmod.c, prepare_update_s.c),
// source wavelets (wavelet.c) and the causal Butterworth filter
// (timedomain_filt.c + libcseife).  All of it runs once per model, stage or
// shot; the arithmetic keeps the reference's float/double mix so the
// coefficients handed to the kernels are the reference's values.
#include "dn_setup.h"

#include <cmath>
#include <cstdlib>
#include <cstring>

namespace dn {

static const double kPi = 3.141592653589793;  // fd.h:22

// Staggered-grid coefficient table, holbergcoeff.c:25-66 (float).  Row
// MAXRELERROR, column FDORDER/2-1; entry 0 is the points-per-wavelength
// figure used by the dispersion check, 1..6 the coefficients.
static const float kHolberg[5][6][7] = {
    {{23.0f, 1.0f, 0, 0, 0, 0, 0},
     {8.0f, (float)(9.0 / 8.0), (float)(-1.0 / 24.0), 0, 0, 0, 0},
     {6.0f, (float)(75.0 / 64.0), (float)(-25.0 / 384.0), (float)(3.0 / 640.0), 0, 0, 0},
     {5.0f, (float)(1225.0 / 1024.0), (float)(-245.0 / 3072.0), (float)(49.0 / 5120.0), (float)(-5.0 / 7168.0), 0, 0},
     {5.0f, (float)(19845.0 / 16384.0), (float)(-735.0 / 8192.0), (float)(567.0 / 40960.0),
      (float)(-405.0 / 229376.0), (float)(35.0 / 294912.0), 0},
     {4.0f, (float)(160083.0 / 131072.0), (float)(-12705.0 / 131072.0), (float)(22869.0 / 1310720.0),
      (float)(-5445.0 / 1835008.0), (float)(847.0 / 2359296.0), (float)(-63.0 / 2883584.0)}},
    {{49.7f, 1.0010f, 0, 0, 0, 0, 0},
     {8.32f, 1.1382f, -0.046414f, 0, 0, 0, 0},
     {4.77f, 1.1965f, -0.078804f, 0.0081781f, 0, 0, 0},
     {3.69f, 1.2257f, -0.099537f, 0.018063f, -0.0026274f, 0, 0},
     {3.19f, 1.2415f, -0.11231f, 0.026191f, -0.0064682f, 0.001191f, 0},
     {2.91f, 1.2508f, -0.12034f, 0.032131f, -0.010142f, 0.0029857f, -0.00066667f}},
    {{22.2f, 1.0050f, 0, 0, 0, 0, 0},
     {5.65f, 1.1534f, -0.052806f, 0, 0, 0, 0},
     {3.74f, 1.2111f, -0.088313f, 0.011768f, 0, 0, 0},
     {3.11f, 1.2367f, -0.10815f, 0.023113f, -0.0046905f, 0, 0},
     {2.80f, 1.2496f, -0.11921f, 0.031130f, -0.0093272f, 0.0025161f, 0},
     {2.62f, 1.2568f, -0.12573f, 0.036423f, -0.013132f, 0.0047484f, -0.0015979f}},
    {{15.8f, 1.0100f, 0, 0, 0, 0, 0},
     {4.80f, 1.1640f, -0.057991f, 0, 0, 0, 0},
     {3.39f, 1.2192f, -0.094070f, 0.014608f, 0, 0, 0},
     {2.90f, 1.2422f, -0.11269f, 0.026140f, -0.0064054f, 0, 0},
     {2.65f, 1.2534f, -0.12257f, 0.033755f, -0.011081f, 0.0036784f, 0},
     {2.51f, 1.2596f, -0.12825f, 0.038550f, -0.014763f, 0.0058619f, -0.0024538f}},
    {{9.16f, 1.0300f, 0, 0, 0, 0, 0},
     {3.47f, 1.1876f, -0.072518f, 0, 0, 0, 0},
     {2.91f, 1.2341f, -0.10569f, 0.022589f, 0, 0, 0},
     {2.61f, 1.2516f, -0.12085f, 0.032236f, -0.011459f, 0, 0},
     {2.45f, 1.2596f, -0.12829f, 0.038533f, -0.014681f, 0.0072580f, 0},
     {2.36f, 1.2640f, -0.13239f, 0.042217f, -0.017803f, 0.0081959f, -0.0051848f}},
};

int fd_coefficients(int fdorder, int maxrelerror, float hc[7])
{
    if (fdorder < 2 || fdorder > 12 || (fdorder % 2) != 0) return -1;
    if (maxrelerror < 0 || maxrelerror > 4) return -2;
    for (int i = 0; i < 7; ++i) hc[i] = kHolberg[maxrelerror][fdorder / 2 - 1][i];
    return 0;
}

// One CPML strip side.  `origin_dist(xval)` returns the distance into the
// layer at grid coordinate xval (float, as PML_pro.c); the half-grid point
// uses a double expression (PML_pro.c:134, 181) rounded to float.
namespace {
struct Prof {
    float d, K, alpha, a, b;
};

inline void finish(Prof &p, float DT, bool clip_alpha)
{
    if (clip_alpha && p.alpha < 0.0f) p.alpha = 0.0f;
    p.b = (float)std::exp((double)(-(p.d / p.K + p.alpha) * DT));
    // PML_pro.c:153: stdlib abs() on an int -- a is set only where (int)d != 0
    if (std::abs((int)p.d) > 1.0e-6)
        p.a = (float)((double)p.d * ((double)p.b - 1.0) / (double)(p.K * (p.d + p.K * p.alpha)));
}

inline void fill(Prof &p, float abscissa, float thickness, float d0, float npower, float kmax, float alpha_max)
{
    if (abscissa >= 0.0f) {
        const float an = abscissa / thickness;
        p.d = (float)((double)d0 * std::pow((double)an, (double)npower));
        p.K = (float)(1.0 + ((double)kmax - 1.0) * std::pow((double)an, (double)npower));
        p.alpha = (float)((double)alpha_max * (1.0 - (double)an) + 0.1 * (double)alpha_max);
    }
}
}  // namespace

void cpml_profiles(const CpmlIn &in, CpmlOut &out)
{
    const int FW = in.fw;
    out.n = 2 * FW + 1;
    for (int v = 0; v < 12; ++v) out.v[v].assign((size_t)out.n, 0.0f);
    if (FW <= 0) return;
    const float DH = in.dh, DT = in.dt;
    const float alpha_max = (float)(2.0 * kPi * ((double)in.fpml / 2.0));
    const float thick = (float)FW * DH;
    const float Rcoef = 0.001f;
    const float d0 = (float)((double)(-(in.npower + 1.0f) * in.damping) * std::log((double)Rcoef) / (2.0 * (double)thick));

    auto store = [&](int base, int idx, const Prof &p) {
        // base 0: full-grid x, 3: half-grid x, 6: full-grid y, 9: half-grid y  (K, a, b)
        out.v[base + 0][idx] = p.K;
        out.v[base + 1][idx] = p.a;
        out.v[base + 2][idx] = p.b;
    };

    for (int dir = 0; dir < 2; ++dir) {
        const int N = dir == 0 ? in.nx : in.ny;
        const bool clip = (dir == 0);  // PML_pro.c:146-147 clips alpha in x only
        const float origin_lo = thick;
        const float origin_hi = (float)(N - 1) * DH - thick;
        for (int i = 1; i <= FW; ++i) {
            Prof full{0, 1, 0, 0, 0}, half{0, 1, 0, 0, 0};
            const float xval = DH * (float)(i - 1);
            fill(full, origin_lo - xval, thick, d0, in.npower, in.k_max_pml, alpha_max);
            fill(half, (float)((double)origin_lo - ((double)xval + (double)DH / 2.0)), thick, d0, in.npower,
                 in.k_max_pml, alpha_max);
            finish(full, DT, clip);
            finish(half, DT, clip);
            store(dir == 0 ? 0 : 6, i, full);
            store(dir == 0 ? 3 : 9, i, half);
        }
        for (int i = N - FW + 1; i <= N; ++i) {
            const int h = i - N + 2 * FW;
            Prof full{0, 1, 0, 0, 0}, half{0, 1, 0, 0, 0};
            const float xval = DH * (float)(i - 1);
            fill(full, xval - origin_hi, thick, d0, in.npower, in.k_max_pml, alpha_max);
            fill(half, (float)((double)xval + (double)DH / 2.0 - (double)origin_hi), thick, d0, in.npower,
                 in.k_max_pml, alpha_max);
            finish(full, DT, clip);
            finish(half, DT, clip);
            store(dir == 0 ? 0 : 6, h, full);
            store(dir == 0 ? 3 : 9, h, half);
        }
    }
}

void relaxation_constants(int L, const float *FL, float DT, float *eta, float *sumu, float *bip, float *bjm,
                          float *cip, float *cjm)
{
    float pts[kSetupMaxL + 1];
    for (int l = 0; l < L; ++l) {
        pts[l] = (float)(1.0 / (2.0 * kPi * (double)FL[l]));  // readmod.c:31
        eta[l] = DT / pts[l];                                  // readmod.c:32
    }
    const float ws = (float)(2.0 * kPi * (double)FL[0]);       // prepare_update_s.c:27 (Q15)
    float su = 0.0f;
    for (int l = 0; l < L; ++l) {
        const float num = ws * ws * pts[l] * pts[l];
        su = (float)((double)su + (double)num / (1.0 + (double)num));
    }
    *sumu = su;
    for (int l = 0; l < L; ++l) {
        bip[l] = (float)(1.0 / (1.0 + ((double)eta[l] * 0.5)));
        bjm[l] = bip[l];
        cip[l] = (float)(1.0 - ((double)eta[l] * 0.5));
        cjm[l] = cip[l];
    }
}

int wavelet(int quellart, int nt_total, float DT, float tshift, float fc, float amp, float *out)
{
    for (int nt = 1; nt <= nt_total; ++nt) {
        const float t = (float)nt * DT;
        float ts = (float)(1.0 / (double)fc);
        float a = 0.0f;
        switch (quellart) {
        case 1: {  // wavelet.c:43-46 (Q9)
            const float tau = (float)(kPi * ((double)t - 1.5 * (double)ts - (double)tshift) / (1.5 * (double)ts));
            a = (float)((1.0 - 4.0 * (double)tau * (double)tau) * std::exp(-2.0 * (double)tau * (double)tau));
            break;
        }
        case 2:
            if ((t < tshift) || (t > (tshift + ts)))
                a = 0.0f;
            else
                a = (float)(std::sin(2.0 * kPi * (double)(t - tshift) * (double)fc) -
                            0.5 * std::sin(4.0 * kPi * (double)(t - tshift) * (double)fc));
            break;
        case 4:
            if ((t < tshift) || (t > (tshift + ts)))
                a = 0.0f;
            else
                a = (float)std::pow(std::sin(kPi * (double)(t + tshift) / (double)ts), 3.0);
            break;
        case 5: {
            ts = (float)(1.2 / (double)fc);
            const float ag = (float)(kPi * kPi * (double)fc * (double)fc);
            a = (float)(-2.0 * (double)ag * (double)(t - ts) * std::exp((double)(-ag * (t - ts) * (t - ts))));
            break;
        }
        case 6:
            a = (nt == 1 + (int)std::floor((double)(tshift / DT) + 0.5)) ? 1.0f : 0.0f;
            break;
        default:
            return -1;
        }
        out[nt - 1] = a * amp;
    }
    return 0;
}

// libcseife recursion (cseife_rekfl.c) on x[0..n-1] in place, FP64.
static void iir(double *x, int n, double f0, double f1, double f2, double g1, double g2)
{
    double xa = 0.0, xaa = 0.0, ya = 0.0, yaa = 0.0;
    for (int j = 0; j < n; ++j) {
        const double xn = x[j];
        const double y = f0 * xn + f1 * xa + f2 * xaa + g1 * ya + g2 * yaa;
        x[j] = y;
        xaa = xa;
        xa = xn;
        yaa = ya;
        ya = y;
    }
}

void butterworth(double *x, int n, double dt, double t0, int order, int highpass)
{
    const double eps = 8.0 * std::atan(1.0) / (t0 / dt);  // cseife_rfk.c: zpi/t0 with t0 in samples
    if (highpass) {
        const double x0 = x[0];  // seife_first
        for (int i = 0; i < n; ++i) x[i] -= x0;
    }
    const int mm = order / 2;
    if (order > 2 * mm) {
        const double g1 = (2.0 - eps) / (2.0 + eps);
        const double f0 = highpass ? 2.0 / (2.0 + eps) : eps / (2.0 + eps);
        iir(x, n, f0, highpass ? -f0 : f0, 0.0, g1, 0.0);
    }
    const double pih = 2.0 * std::atan(1.0);
    for (int j = 1; j <= mm; ++j) {
        const double h = std::sin(pih * (2 * j - 1) / order);
        const double epsq = eps * eps;
        const double a = 1.0 - eps * h + epsq / 4.0;
        const double b = -2.0 + epsq / 2.0;
        const double c = 1.0 + eps * h + epsq / 4.0;
        double f0, f1, f2;
        if (!highpass) {
            f0 = epsq / 4.0 / c;
            f1 = f0 + f0;
            f2 = f0;
        } else {
            f0 = 1.0 / c;
            f1 = -f0 - f0;
            f2 = f0;
        }
        iir(x, n, f0, f1, f2, -b / c, -a / c);
    }
}

void timedomain_filt(float *data, int ntr, int ns, float DT, float fc, int order, int method)
{
    // timedomain_filt.c:31-49 (Q8): the filter sees [0, s1 .. s_{ns-1}]
    double *buf = (double *)std::calloc((size_t)ns + 1, sizeof(double));
    const double T0 = 1.0 / (double)fc;
    for (int t = 0; t < ntr; ++t) {
        float *row = data + (size_t)t * ns;
        buf[0] = 0.0;
        for (int j = 1; j <= ns; ++j) buf[j] = (double)row[j - 1];
        butterworth(buf, ns, (double)DT, T0, order, method == 2);
        for (int j = 1; j <= ns; ++j) row[j - 1] = (float)buf[j];
    }
    std::free(buf);
}

}  // namespace dn

// ---------------------------------------------------------------------------
// C-ABI exports of the host setup (include/denise_sh.h)
// ---------------------------------------------------------------------------
#include "../../include/denise_sh.h"

extern "C" int dn_holbergcoeff(int fdorder, int maxrelerror, float *hc)
{
    return hc ? dn::fd_coefficients(fdorder, maxrelerror, hc) : -1;
}

extern "C" int dn_pml_pro(int nx, int ny, int fw, float dh, float dt, float damping, float fpml, float npower,
                          float k_max_pml, float *out)
{
    if (!out || fw < 0) return -1;
    dn::CpmlIn in{nx, ny, fw, dh, dt, damping, fpml, npower, k_max_pml};
    dn::CpmlOut co;
    dn::cpml_profiles(in, co);
    for (int v = 0; v < 12; ++v) std::memcpy(out + (size_t)v * co.n, co.v[v].data(), (size_t)co.n * sizeof(float));
    return 0;
}

extern "C" int dn_relaxation_constants(int L, const float *fl, float dt, float *eta, float *sumu, float *bip,
                                       float *bjm, float *cip, float *cjm)
{
    if (L < 1 || L > dn::kSetupMaxL || !fl) return -1;
    dn::relaxation_constants(L, fl, dt, eta, sumu, bip, bjm, cip, cjm);
    return 0;
}

extern "C" int dn_wavelet(int quellart, int nt, float dt, float tshift, float fc, float amp, float *out)
{
    return out ? dn::wavelet(quellart, nt, dt, tshift, fc, amp, out) : -1;
}

extern "C" int dn_timedomain_filt(float *data, int ntr, int ns, float dt, float fc, int order, int method)
{
    if (!data || ntr < 0 || ns < 1 || order < 1 || (method != 1 && method != 2)) return -1;
    dn::timedomain_filt(data, ntr, ns, dt, fc, order, method);
    return 0;
}

// checkfd_ssg_visc.c:14-142 / checkfd_ssg_elastic.c:14-130 with INVMAT1 = 1
// (the model holds vs, readmod.c:41-60): phase-velocity extremes, the
// dispersion and stability limits, and the absorbing-frame width check.
// Returns 1 where the reference calls err() (DT > dtstab, FW wider than the
// grid), 0 otherwise; warnings are reported through the flags.
extern "C" int dn_checkfd(int nx, int ny, int fdorder, int maxrelerror, int L, const float *fl, float dt, float dh,
                          float ts_src, int fw, const float *vs, const float *rho, const float *qs,
                          dn_checkfd_result *out)
{
    if (nx < 1 || ny < 1 || !vs || !rho || !out || L < 0 || L > dn::kSetupMaxL || (L > 0 && (!fl || !qs)))
        return -1;
    float hc[7];
    if (dn::fd_coefficients(fdorder, maxrelerror, hc) != 0) return -1;
    float eta[dn::kSetupMaxL + 1] = {0};
    for (int l = 1; l <= L; ++l) {
        const float pts = 1.0 / (2.0 * dn::kPi * fl[l - 1]);   // readmod.c:32-33
        eta[l] = dt / pts;
    }
    const float ws = 2.0 * dn::kPi / ts_src;
    float cmax = 0.0f, cmin = 1e9f;
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j) {
            const size_t k = (size_t)(j - 1) * nx + (size_t)(i - 1);
            float c;
            if (L > 0) {
                const float taus = 2.0 / qs[k];                  // readmod.c:99
                float sumu = 0.0f;
                for (int l = 1; l <= L; ++l) {
                    const float t = dt / eta[l];
                    sumu = sumu + ((ws * ws * t * t * taus) / (1.0 + ws * ws * t * t));
                }
                const float pu_ref = rho[k] * vs[k] * vs[k];
                c = std::sqrt(pu_ref / (rho[k] * (1.0 + sumu)));
                if (cmin > c) cmin = c;
                c = std::sqrt(pu_ref * (1.0 + L * taus) / (rho[k] * (1.0 + sumu)));
                if (cmax < c) cmax = c;
            } else {
                c = vs[k];
                if (cmax < c) cmax = c;
                if (cmin > c) cmin = c;
            }
        }
    const float fmax = 2.0 / ts_src;
    const float dhstab = cmin / (hc[0] * fmax);
    // C's fabs takes and returns double: the sum is formed in double
    const float gamma = std::fabs((double)hc[1]) + std::fabs((double)hc[2]) + std::fabs((double)hc[3]) +
                        std::fabs((double)hc[4]) + std::fabs((double)hc[5]) + std::fabs((double)hc[6]);
    const float dtstab = dh / (std::sqrt(2.0) * gamma * cmax);
    out->cmin = cmin;
    out->cmax = cmax;
    out->dtstab = dtstab;
    out->dhstab = dhstab;
    out->dh_dispersive = dh > dhstab;
    out->dt_unstable = dt > dtstab;
    out->fw_too_large = (fw > nx) || (fw > ny);
    out->fw_small = fw < 10;
    return (out->dt_unstable || out->fw_too_large) ? 1 : 0;
}


int dn_fail(dn_ctx *c, const char *fmt, ...);   // dn_core.cpp

// time_window.c:40-60: damping exp(-GAMMA (t - pick)^2) for t = j DT >= pick
extern "C" int dn_time_window(int ntr, int ns, float dt, int timewin, float gamma, float twlength_plus,
                              const float *picks, float *data)
{
    if (ntr < 0 || ns < 0 || !data || (timewin != 1 && timewin != 2) || (timewin == 1 && !picks)) {
        dn_fail(nullptr, "dn_time_window: TIMEWIN must be 1 (with picks) or 2");
        return -1;
    }
    const float DT = dt, GAMMA = gamma;
    for (int i = 0; i < ntr; ++i) {
        const float pick = (timewin == 1) ? picks[i] + twlength_plus : DT;
        float *row = data + (size_t)i * ns;
        for (int j = 1; j <= ns; ++j) {
            const float time = (float)(j * DT);
            const float dump = time - pick;
            const float taper = (float)std::exp((double)(-GAMMA * dump * dump));
            if (time >= pick) row[j - 1] = row[j - 1] * taper;
        }
    }
    return 0;
}
