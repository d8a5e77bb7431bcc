// Source-time-function estimate by Wiener deconvolution (stf.c:11-250) on
// the device, with the project's own FFT.
//
// For one shot with observed traces d_i, modelled traces m_i (modelled with
// the current source signal s) and Npad = 2^(ceil(log2 ns) + 2):
//   S_n[k] = sum_i ntr D_i[k] conj(M_i[k]),  S_d[k] = sum_i ntr |M_i[k]|^2
//   Ebar   = sum_i sum_k S_d^(i)[k] / (Npad ntr)   (S_d^(i): the running sum
//            after trace i, as stf.c:165-168 accumulates it)
//   STF    = Re IFFT( FFT(s) S_n / (S_d + EPS_STF ntr Ebar) )[0..ns-1]
// Killed and muted traces are zero (stf.c:57-112); NORMALIZE 1 scales each
// trace by its largest |sample| over samples 2..ns, NORMALIZE 2 the observed
// trace to the modelled trace's (util.c:489-579).  The reference transforms
// with FFTW's +1 sign forward and -1 inverse; for real input this gives the
// same real part as the usual convention used here.  Sums run in FP64 in a
// fixed order; the reference accumulates Ebar in float.
//
// FFT: radix-2 decimation in time, a bit-reversal gather then log2(Npad)
// butterfly passes over all traces at once, complex double in HBM.
#include "dn_core.h"

#include <cmath>
#include <vector>

using namespace dn;

namespace {

struct cd {
    double re, im;
};

__device__ inline cd cmul(cd a, cd b) { return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

// x [ntr][ns] float (row stride ld) -> X [ntr][N] complex, bit-reversed
// order, zero padding beyond ns; mute[i] != 0 gives a zero trace; samples
// 2..ns are divided by scale[i] (mul = 0) or multiplied by it (mul = 1),
// sample 1 is kept (util.c:502-520, 545-571)
__global__ void k_fft_load(int ntr, int ns, int ld, int N, int logN, const float *__restrict__ x,
                           const unsigned char *__restrict__ mute, const float *__restrict__ scale, int mul,
                           cd *__restrict__ X)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ntr * N) return;
    const int i = (int)(t / N), j = (int)(t % N);
    const unsigned r = __builtin_bitreverse32((unsigned)j) >> (32 - logN);
    double v = 0.0;
    if (j < ns && !(mute && mute[i])) {
        float f = x[(long long)i * ld + j];
        if (scale && j >= 1) f = mul ? f * scale[i] : f / scale[i];
        v = (double)f;
    }
    X[(long long)i * N + r] = cd{v, 0.0};
}

// one butterfly pass of half-length h for every trace; sign +1 or -1
__global__ void k_fft_pass(int ntr, int N, int h, int sign, cd *__restrict__ X)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nb = N / 2;
    if (t >= (long long)ntr * nb) return;
    const int i = (int)(t / nb), b = (int)(t % nb);
    const int grp = b / h, pos = b % h;
    const int i0 = grp * 2 * h + pos, i1 = i0 + h;
    double sn, cs;
    sincospi((double)pos / (double)h, &sn, &cs);   // angle pi pos / h = 2 pi pos / (2h)
    const cd w{cs, sign * sn};
    cd *row = X + (long long)i * N;
    const cd a = row[i0], b1 = cmul(row[i1], w);
    row[i0] = cd{a.re + b1.re, a.im + b1.im};
    row[i1] = cd{a.re - b1.re, a.im - b1.im};
}

// per-trace scales of NORMALIZE 1 / 2 (max |x| over samples 2..ns, 0 -> 1)
__global__ void k_trace_max(int ntr, int ns, int ld, const float *__restrict__ x, float *__restrict__ mx)
{
    const int i = blockIdx.x;
    float m = 0.0f, mn = 0.0f;
    for (int j = 1 + threadIdx.x; j < ns; j += blockDim.x) {
        const float v = x[(long long)i * ld + j];
        m = fmaxf(m, v);
        mn = fminf(mn, v);
    }
    __shared__ float sm[256], sn[256];
    sm[threadIdx.x] = m;
    sn[threadIdx.x] = mn;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sm[threadIdx.x] = fmaxf(sm[threadIdx.x], sm[threadIdx.x + o]);
            sn[threadIdx.x] = fminf(sn[threadIdx.x], sn[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float a = sm[0], b = fabsf(sn[0]);
        float t = (a > b) ? a : b;
        if (t == 0.0f) t = 1.0f;
        mx[i] = t;
    }
}

// NORMALIZE 2: the observed trace times max_syn / max_obs (util.c:570)
__global__ void k_norm_ratio(int ntr, const float *mx_syn, const float *mx_obs, float *ratio)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ntr) ratio[i] = mx_syn[i] / mx_obs[i];
}

// S_n, S_d and the per-frequency share of Ebar (sum over the running sums)
__global__ void k_wiener_sums(int ntr, int N, const cd *__restrict__ D, const cd *__restrict__ M, cd *__restrict__ sn,
                              double *__restrict__ sd, double *__restrict__ erow)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const double n = (double)ntr;
    cd a{0.0, 0.0};
    double d = 0.0, e = 0.0;
    for (int i = 0; i < ntr; ++i) {
        const cd dd = D[(long long)i * N + k], mm = M[(long long)i * N + k];
        const cd nd{n * dd.re, n * dd.im}, nm{n * mm.re, n * mm.im};
        const cd p = cmul(nd, cd{mm.re, -mm.im});
        a.re += p.re;
        a.im += p.im;
        d += nm.re * mm.re + nm.im * mm.im;
        e += d;
    }
    sn[k] = a;
    sd[k] = d;
    erow[k] = e;
}

// fixed-order sum of erow (one block)
__global__ void k_sum_fixed(int N, const double *__restrict__ v, double *__restrict__ out)
{
    __shared__ double sm[256];
    double a = 0.0;
    for (int k = threadIdx.x; k < N; k += blockDim.x) a += v[k];
    sm[threadIdx.x] = a;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sm[0];
}

// D_s = FFT(s) S_n / (S_d + eps), written bit-reversed for the inverse pass
__global__ void k_wiener_div(int N, int logN, const cd *__restrict__ S, const cd *__restrict__ sn,
                             const double *__restrict__ sd, const double *__restrict__ eps, cd *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const double den = sd[k] + *eps;
    const cd q{sn[k].re / den, sn[k].im / den};
    const unsigned r = __builtin_bitreverse32((unsigned)k) >> (32 - logN);
    out[r] = cmul(S[k], q);
}

__global__ void k_eps(const double *esum, int N, int ntr, float eps_stf, double *eps)
{
    // stf.c:170-174: Ebar = Ebar / (Npad ntr); the regulariser EPS_STF ntr Ebar
    const float ebar = (float)(*esum / ((double)N * (double)ntr));
    *eps = (double)((eps_stf * (float)ntr) * ebar);
}

__global__ void k_stf_out(int ns, int N, const cd *__restrict__ x, float *__restrict__ stf)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < ns) stf[j] = (float)(x[j].re / (double)N);
}

unsigned blocks(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

struct StfDev {
    hipStream_t st = nullptr;
    float *obs = nullptr, *syn = nullptr, *sig = nullptr, *stf = nullptr, *mx = nullptr;
    unsigned char *mute = nullptr;
    cd *D = nullptr, *M = nullptr, *S = nullptr, *sn = nullptr, *Y = nullptr;
    double *sd = nullptr, *erow = nullptr, *esum = nullptr;
    void release()
    {
        void *ps[] = {obs, syn, sig, stf, mx, mute, D, M, S, sn, Y, sd, erow, esum};
        for (void *p : ps)
            if (p) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
        *this = StfDev{};
    }
};

#define SCHK(x)                                                                                  \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) {                                                                  \
            R.release();                                                                         \
            return dn_fail(nullptr, "%s: %s", #x, hipGetErrorString(e_));                        \
        }                                                                                        \
    } while (0)

void fft(cd *X, int ntr, int N, int logN, int sign, hipStream_t st)
{
    for (int h = 1; h < N; h *= 2)
        hipLaunchKernelGGL(k_fft_pass, dim3(blocks((long long)ntr * (N / 2))), dim3(256), 0, st, ntr, N, h, sign, X);
}

}  // namespace

extern "C" int dn_stf(int ntr, int ns, const float *obs, const float *syn, const float *signal,
                      const unsigned char *mute, const dn_stf_opts *o, float *stf)
{
    if (ntr < 1 || ns < 2 || !obs || !syn || !signal || !o || !stf) return dn_fail(nullptr, "dn_stf: bad arguments");
    if (o->normalize < 0 || o->normalize > 2) return dn_fail(nullptr, "dn_stf: NORMALIZE must be 0, 1 or 2");
    // stf.c:38: Npad = 2^(ceil(log2 ns) + 2)
    const int logN = (int)std::ceil(std::log((double)ns) / std::log(2.0)) + 2;
    if (logN > 26) return dn_fail(nullptr, "dn_stf: record too long");
    const int N = 1 << logN;
    StfDev R;
    SCHK(hipStreamCreateWithFlags(&R.st, hipStreamNonBlocking));
    const size_t nd = (size_t)ntr * ns;
    SCHK(hipMalloc((void **)&R.obs, nd * sizeof(float)));
    SCHK(hipMalloc((void **)&R.syn, nd * sizeof(float)));
    SCHK(hipMalloc((void **)&R.sig, (size_t)ns * sizeof(float)));
    SCHK(hipMalloc((void **)&R.stf, (size_t)ns * sizeof(float)));
    SCHK(hipMalloc((void **)&R.mx, (size_t)4 * ntr * sizeof(float)));
    SCHK(hipMalloc((void **)&R.mute, (size_t)ntr));
    SCHK(hipMalloc((void **)&R.D, (size_t)ntr * N * sizeof(cd)));
    SCHK(hipMalloc((void **)&R.M, (size_t)ntr * N * sizeof(cd)));
    SCHK(hipMalloc((void **)&R.S, (size_t)N * sizeof(cd)));
    SCHK(hipMalloc((void **)&R.sn, (size_t)N * sizeof(cd)));
    SCHK(hipMalloc((void **)&R.Y, (size_t)N * sizeof(cd)));
    SCHK(hipMalloc((void **)&R.sd, (size_t)N * sizeof(double)));
    SCHK(hipMalloc((void **)&R.erow, (size_t)N * sizeof(double)));
    SCHK(hipMalloc((void **)&R.esum, 2 * sizeof(double)));
    hipStream_t st = R.st;
    SCHK(hipMemcpyAsync(R.obs, obs, nd * sizeof(float), hipMemcpyHostToDevice, st));
    SCHK(hipMemcpyAsync(R.syn, syn, nd * sizeof(float), hipMemcpyHostToDevice, st));
    SCHK(hipMemcpyAsync(R.sig, signal, (size_t)ns * sizeof(float), hipMemcpyHostToDevice, st));
    std::vector<unsigned char> mu(ntr, 0);
    if (mute)
        for (int i = 0; i < ntr; ++i) mu[i] = mute[i] ? 1 : 0;
    SCHK(hipMemcpyAsync(R.mute, mu.data(), (size_t)ntr, hipMemcpyHostToDevice, st));
    // normalisation of samples 2..ns (muted traces are zero whatever the scale)
    const float *sc_syn = nullptr, *sc_obs = nullptr;
    int mul_obs = 0;
    if (o->normalize) {
        float *mxs = R.mx, *mxo = R.mx + ntr;
        hipLaunchKernelGGL(k_trace_max, dim3(ntr), dim3(256), 0, st, ntr, ns, ns, R.syn, mxs);
        hipLaunchKernelGGL(k_trace_max, dim3(ntr), dim3(256), 0, st, ntr, ns, ns, R.obs, mxo);
        if (o->normalize == 1) {
            sc_syn = mxs;
            sc_obs = mxo;
        } else {
            hipLaunchKernelGGL(k_norm_ratio, dim3(blocks(ntr)), dim3(256), 0, st, ntr, mxs, mxo, R.mx + 2 * ntr);
            sc_obs = R.mx + 2 * ntr;
            mul_obs = 1;
        }
    }
    hipLaunchKernelGGL(k_fft_load, dim3(blocks((long long)ntr * N)), dim3(256), 0, st, ntr, ns, ns, N, logN, R.obs,
                       R.mute, sc_obs, mul_obs, R.D);
    hipLaunchKernelGGL(k_fft_load, dim3(blocks((long long)ntr * N)), dim3(256), 0, st, ntr, ns, ns, N, logN, R.syn,
                       R.mute, sc_syn, 0, R.M);
    hipLaunchKernelGGL(k_fft_load, dim3(blocks(N)), dim3(256), 0, st, 1, ns, ns, N, logN, R.sig,
                       (const unsigned char *)nullptr, (const float *)nullptr, 0, R.S);
    fft(R.D, ntr, N, logN, +1, st);
    fft(R.M, ntr, N, logN, +1, st);
    fft(R.S, 1, N, logN, +1, st);
    hipLaunchKernelGGL(k_wiener_sums, dim3(blocks(N)), dim3(256), 0, st, ntr, N, R.D, R.M, R.sn, R.sd, R.erow);
    hipLaunchKernelGGL(k_sum_fixed, dim3(1), dim3(256), 0, st, N, R.erow, R.esum);
    hipLaunchKernelGGL(k_eps, dim3(1), dim3(1), 0, st, R.esum, N, ntr, o->eps_stf, R.esum + 1);
    hipLaunchKernelGGL(k_wiener_div, dim3(blocks(N)), dim3(256), 0, st, N, logN, R.S, R.sn, R.sd, R.esum + 1, R.Y);
    fft(R.Y, 1, N, logN, -1, st);
    hipLaunchKernelGGL(k_stf_out, dim3(blocks(ns)), dim3(256), 0, st, ns, N, R.Y, R.stf);
    SCHK(hipGetLastError());
    SCHK(hipMemcpyAsync(stf, R.stf, (size_t)ns * sizeof(float), hipMemcpyDeviceToHost, st));
    SCHK(hipStreamSynchronize(st));
    R.release();
    return 0;
}

// stf.c on the context: the shot's observed data, its synthetics from the
// current signal, the context's trace kill / offset mute plus OFFSETC_STF
extern "C" int dn_stf_shot(dn_ctx *c, int shot, const dn_stf_opts *o, float *stf)
{
    if (!c || !o || !stf) return dn_fail(c, "dn_stf_shot: NULL argument");
    if (shot < 0 || shot >= c->nshots) return dn_fail(c, "dn_stf_shot: shot %d outside 0..%d", shot, c->nshots - 1);
    if (!c->d_obs) return dn_fail(c, "dn_stf_shot: set the observed data first");
    const int ntr = c->ntr, nt_max = c->nt_max, ns = c->shot_nt[shot];
    std::vector<float> syn((size_t)ntr * nt_max), obs((size_t)ntr * nt_max), sig(nt_max);
    if (dn_forward_shots(c, std::vector<int>{shot}, syn.data())) return -1;
    HIPCHK(c, hipMemcpy(obs.data(), c->d_obs + (size_t)shot * ntr * nt_max, obs.size() * sizeof(float),
                        hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(sig.data(), c->d_sig + (size_t)c->src_off[shot] * nt_max, nt_max * sizeof(float),
                        hipMemcpyDeviceToHost));
    // [ntr][ns] rows of the shot's record length
    std::vector<float> o2((size_t)ntr * ns), s2((size_t)ntr * ns);
    for (int r = 0; r < ntr; ++r)
        for (int j = 0; j < ns; ++j) {
            o2[(size_t)r * ns + j] = obs[(size_t)r * nt_max + j];
            s2[(size_t)r * ns + j] = syn[(size_t)r * nt_max + j];
        }
    std::vector<unsigned char> mute(ntr, 0);
    const dn_ctx::DataCond &d = c->dc;
    const float DH = c->p.dh;
    const int l0 = c->src_off[shot];
    const float xs = d.offset_mute ? d.src_x[shot] : (float)c->src_i[l0] * DH;
    const float ys = d.offset_mute ? d.src_y[shot] : (float)c->src_j[l0] * DH;
    for (int r = 0; r < ntr; ++r) {
        if (d.on && d.trkill && d.kill[(size_t)shot * d.kill_ntr + r] == 1) mute[r] = 1;   // stf.c:46-64
        const float xr = (float)c->rec_i[r] * DH, yr = (float)c->rec_j[r] * DH;
        const float offset = (float)std::sqrt((double)(((xs - xr) * (xs - xr)) + ((ys - yr) * (ys - yr))));
        if (d.on && d.offset_mute == 1 && offset >= d.offsetc) mute[r] = 1;                  // stf.c:67-90
        if (d.on && d.offset_mute == 2 && offset <= d.offsetc) mute[r] = 1;
        if (o->offsetc_stf > 0.0f && std::fabs(offset) >= o->offsetc_stf) mute[r] = 1;        // stf.c:93-110
    }
    if (dn_stf(ntr, ns, o2.data(), s2.data(), sig.data(), mute.data(), o, stf)) return dn_fail(c, "%s", dn_last_error(nullptr));
    return 0;
}
