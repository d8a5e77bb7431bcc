// Kernels of the learned diagonal-Hessian preconditioner: a 4-level U-Net of
// 3x3 convolutions in BF16 with FP32 accumulation on the CDNA matrix cores
// (v_mfma_f32_16x16x16_bf16), stored channels-last ([n][H][W][C], C a
// multiple of 16), plus the pooling, upsampling, loss and Adam kernels of
// its training.  Host side: dn_unet.cpp.
#include "dn_unet.h"

#include <algorithm>

namespace dn {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t u_f2bf(float x)
{
    // round to nearest even (finite inputs)
    unsigned int u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float u_bf2f(uint16_t h) { return __uint_as_float(((unsigned int)h) << 16); }

// ---------------------------------------------------------------------------
// 3x3 "same" convolution as an implicit GEMM on MFMA 16x16x16 (bf16 in,
// fp32 accumulate).  D[co][pixel] = sum_{ky,kx,ci} W[ky][kx][co][ci] *
// X[pixel + (ky-1, kx-1)][ci].  Per MFMA: A = 16 output channels x 16 input
// channels of one tap, B = 16 input channels x 16 pixels of one row.
// Operand layout (16x16x16): lane l supplies A[l%16][4(l/16)+j] and
// B[4(l/16)+j][l%16], j = 0..3, and receives D[4(l/16)+j][l%16].
// Workgroup: 4 waves, 64 consecutive pixels of one image row and one group
// of up to 64 output channels (blockIdx.z); each wave keeps the group's
// (<= 4) accumulator tiles of its 16 pixels.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_conv3x3(ConvArgs a)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = blockIdx.y;
    const int img = row / a.H, y = row - img * a.H;
    const int xb = blockIdx.x * 64 + wave * 16;
    const int nl = lane & 15, kq = lane >> 4;
    const int cg = blockIdx.z * 64;                 // first output channel of the group
    const int nmb = min(4, (a.Co - cg) >> 4);
    f4v acc[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[m] = f4v{0.0f, 0.0f, 0.0f, 0.0f};
    for (int ky = 0; ky < 3; ++ky) {
        const int yy = y + ky - 1;
        if (yy < 0 || yy >= a.H) continue;   // uniform in the workgroup
        const uint16_t *xrow = a.x + (long long)(img * a.H + yy) * a.W * a.Ci;
        for (int kx = 0; kx < 3; ++kx) {
            const int xx = xb + nl + kx - 1;
            const bool inb = (xx >= 0) && (xx < a.W);
            const uint16_t *wp = a.w + (long long)(ky * 3 + kx) * a.Co * a.Ci;
            for (int c0 = 0; c0 < a.Ci; c0 += 16) {
                s4v bv = s4v{0, 0, 0, 0};
                if (inb) bv = *(const s4v *)(xrow + (long long)xx * a.Ci + c0 + 4 * kq);
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    if (m < nmb) {
                        const s4v av = *(const s4v *)(wp + (long long)(cg + m * 16 + nl) * a.Ci + c0 + 4 * kq);
                        acc[m] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[m], 0, 0, 0);
                    }
            }
        }
    }
    const int xo = xb + nl;
    if (xo >= a.W) return;
    const long long pix = (long long)(img * a.H + y) * a.W + xo;
#pragma unroll
    for (int m = 0; m < 4; ++m)
        if (m < nmb)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int co = cg + m * 16 + 4 * kq + j;
                float v = acc[m][j];
                if (a.bias) v += a.bias[co];
                if (a.relu) v = fmaxf(v, 0.0f);
                if (a.mask && !(u_bf2f(a.mask[pix * a.Co + co]) > 0.0f)) v = 0.0f;
                if (a.y16) a.y16[pix * a.Co + co] = u_f2bf(v);
                if (a.y32) a.y32[pix * a.Co + co] = v;
            }
}

hipError_t launch_conv3x3(const ConvArgs &a, hipStream_t st)
{
    if (a.Ci % 16 || a.Co % 16 || a.Ci < 16 || a.Co < 16) return hipErrorInvalidValue;
    dim3 grid((unsigned)((a.W + 63) / 64), (unsigned)(a.n * a.H), (unsigned)((a.Co + 63) / 64));
    hipLaunchKernelGGL(k_conv3x3, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[ky][kx][co][ci] = sum_p dY[p][co] * X[p + (ky-1,kx-1)][ci]
// as a GEMM over pixels (K).  Stage 1: one wave per (pixel chunk, tap,
// 16x16 channel tile) writes a partial; stage 2 adds the chunks in order
// (deterministic).  Also the bias gradient (sum over pixels of dY).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_wgrad(WgradArgs a)
{
    const int lane = threadIdx.x, nl = lane & 15, kq = lane >> 4;
    const int chunk = blockIdx.x, tap = blockIdx.y;
    const int ky = tap / 3, kx = tap - 3 * (tap / 3);
    const int ncb = a.Ci >> 4;
    const int cob = blockIdx.z / ncb, cib = blockIdx.z - cob * ncb;
    const long long P = (long long)a.n * a.H * a.W;
    const long long p0 = (long long)chunk * a.chunk;
    const long long p1 = min(P, p0 + a.chunk);
    const long long shift = (long long)(ky - 1) * a.W + (kx - 1);
    f4v acc = f4v{0.0f, 0.0f, 0.0f, 0.0f};
    for (long long pb = p0; pb < p1; pb += 16) {
        s4v av, bv;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long long p = pb + 4 * kq + j;
            uint16_t dyv = 0, xv = 0;
            if (p < p1) {
                dyv = a.dy[p * a.Co + cob * 16 + nl];
                const int xp = (int)(p % a.W);
                const int yp = (int)((p / a.W) % a.H);
                const int yy = yp + ky - 1, xx = xp + kx - 1;
                if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) xv = a.x[(p + shift) * a.Ci + cib * 16 + nl];
            }
            av[j] = (short)dyv;
            bv[j] = (short)xv;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int co = cob * 16 + 4 * kq + j, ci = cib * 16 + nl;
        a.part[(((long long)chunk * 9 + tap) * a.Co + co) * a.Ci + ci] = acc[j];
    }
}

__global__ void k_bgrad(WgradArgs a)
{
    const int co = threadIdx.x;
    if (co >= a.Co) return;
    const long long P = (long long)a.n * a.H * a.W;
    const long long p0 = (long long)blockIdx.x * a.chunk;
    const long long p1 = min(P, p0 + a.chunk);
    float s = 0.0f;
    for (long long p = p0; p < p1; ++p) s += u_bf2f(a.dy[p * a.Co + co]);
    a.bpart[(long long)blockIdx.x * a.Co + co] = s;
}

__global__ void k_reduce_chunks(int nchunks, long long len, const float *part, float *out)
{
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < len;
         e += (long long)gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int c = 0; c < nchunks; ++c) s += part[(long long)c * len + e];
        out[e] = s;
    }
}

int wgrad_chunks(const WgradArgs &a)
{
    const long long P = (long long)a.n * a.H * a.W;
    return (int)((P + a.chunk - 1) / a.chunk);
}

hipError_t launch_wgrad(const WgradArgs &a, float *dw, float *db, hipStream_t st)
{
    if (a.Ci % 16 || a.Co % 16 || a.Co > 64 || a.chunk % 16) return hipErrorInvalidValue;
    const int nch = wgrad_chunks(a);
    hipLaunchKernelGGL(k_wgrad, dim3(nch, 9, (a.Co / 16) * (a.Ci / 16)), dim3(64), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const long long len = 9LL * a.Co * a.Ci;
    hipLaunchKernelGGL(k_reduce_chunks, dim3((unsigned)std::min<long long>(1024, (len + 255) / 256)), dim3(256), 0, st,
                       nch, len, a.part, dw);
    e = hipGetLastError();
    if (e != hipSuccess || !db) return e;
    hipLaunchKernelGGL(k_bgrad, dim3(nch), dim3(64), 0, st, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_chunks, dim3(1), dim3(64), 0, st, nch, (long long)a.Co, a.bpart, db);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// 2x2 average pooling, nearest upsampling with channel concatenation, and
// their adjoints with the ReLU mask of the activation they feed.
// ---------------------------------------------------------------------------
__global__ void k_pool(int n, int H, int W, int C, const uint16_t *x, uint16_t *y)
{
    const int h = H / 2, w = W / 2;
    const long long total = (long long)n * h * w * C;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(e % C);
        const long long q = e / C;
        const int xo = (int)(q % w), yo = (int)((q / w) % h), im = (int)(q / ((long long)w * h));
        const long long b = ((long long)(im * H + 2 * yo) * W + 2 * xo) * C + c;
        const float s = (u_bf2f(x[b]) + u_bf2f(x[b + C])) + (u_bf2f(x[b + (long long)W * C]) +
                                                             u_bf2f(x[b + (long long)W * C + C]));
        y[e] = u_f2bf(0.25f * s);
    }
}

// dact[p][c] = (act > 0) * (0.25 * dpool[p/2][c] + dcat[p][off + c])
__global__ void k_pool_bwd(int n, int H, int W, int C, const uint16_t *dpool, const uint16_t *dcat, int ccat, int off,
                           const uint16_t *act, uint16_t *dact)
{
    const int h = H / 2, w = W / 2;
    const long long total = (long long)n * H * W * C;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(e % C);
        const long long p = e / C;
        const int x = (int)(p % W), yq = (int)((p / W) % H), im = (int)(p / ((long long)W * H));
        const long long q = ((long long)(im * h + yq / 2) * w + x / 2) * C + c;
        float v = 0.25f * u_bf2f(dpool[q]);
        if (dcat) v += u_bf2f(dcat[p * ccat + off + c]);
        if (!(u_bf2f(act[e]) > 0.0f)) v = 0.0f;
        dact[e] = u_f2bf(v);
    }
}

// cat[p][0..Cu) = up[p/2][c]; cat[p][Cu..Cu+Cs) = skip[p][c]
__global__ void k_upcat(int n, int H, int W, int Cu, int Cs, const uint16_t *up, const uint16_t *skip, uint16_t *cat)
{
    const int Cc = Cu + Cs, h = H / 2, w = W / 2;
    const long long total = (long long)n * H * W * Cc;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(e % Cc);
        const long long p = e / Cc;
        if (c < Cu) {
            const int x = (int)(p % W), yq = (int)((p / W) % H), im = (int)(p / ((long long)W * H));
            cat[e] = up[((long long)(im * h + yq / 2) * w + x / 2) * Cu + c];
        } else {
            cat[e] = skip[p * Cs + (c - Cu)];
        }
    }
}

// dlow[q][c] = (act_low > 0) * sum over the 2x2 children of dcat[child][c], c < Cu
__global__ void k_up_bwd(int n, int H, int W, int Cu, int Ccat, const uint16_t *dcat, const uint16_t *act_low,
                         uint16_t *dlow)
{
    const int h = H / 2, w = W / 2;
    const long long total = (long long)n * h * w * Cu;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(e % Cu);
        const long long q = e / Cu;
        const int xo = (int)(q % w), yo = (int)((q / w) % h), im = (int)(q / ((long long)w * h));
        const long long b = ((long long)(im * H + 2 * yo) * W + 2 * xo);
        float s = (u_bf2f(dcat[b * Ccat + c]) + u_bf2f(dcat[(b + 1) * Ccat + c])) +
                  (u_bf2f(dcat[(b + W) * Ccat + c]) + u_bf2f(dcat[(b + W + 1) * Ccat + c]));
        if (!(u_bf2f(act_low[e]) > 0.0f)) s = 0.0f;
        dlow[e] = u_f2bf(s);
    }
}

static dim3 grid_for(long long total)
{
    return dim3((unsigned)std::max<long long>(1, std::min<long long>(8192, (total + 255) / 256)));
}

hipError_t launch_pool(int n, int H, int W, int C, const uint16_t *x, uint16_t *y, hipStream_t st)
{
    hipLaunchKernelGGL(k_pool, grid_for((long long)n * (H / 2) * (W / 2) * C), dim3(256), 0, st, n, H, W, C, x, y);
    return hipGetLastError();
}
hipError_t launch_pool_bwd(int n, int H, int W, int C, const uint16_t *dpool, const uint16_t *dcat, int ccat, int off,
                           const uint16_t *act, uint16_t *dact, hipStream_t st)
{
    hipLaunchKernelGGL(k_pool_bwd, grid_for((long long)n * H * W * C), dim3(256), 0, st, n, H, W, C, dpool, dcat, ccat,
                       off, act, dact);
    return hipGetLastError();
}
hipError_t launch_upcat(int n, int H, int W, int Cu, int Cs, const uint16_t *up, const uint16_t *skip, uint16_t *cat,
                        hipStream_t st)
{
    hipLaunchKernelGGL(k_upcat, grid_for((long long)n * H * W * (Cu + Cs)), dim3(256), 0, st, n, H, W, Cu, Cs, up, skip,
                       cat);
    return hipGetLastError();
}
hipError_t launch_up_bwd(int n, int H, int W, int Cu, int Ccat, const uint16_t *dcat, const uint16_t *act_low,
                         uint16_t *dlow, hipStream_t st)
{
    hipLaunchKernelGGL(k_up_bwd, grid_for((long long)n * (H / 2) * (W / 2) * Cu), dim3(256), 0, st, n, H, W, Cu, Ccat,
                       dcat, act_low, dlow);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Input packing, output extraction, loss, parameter casts and Adam
// ---------------------------------------------------------------------------
// x [n][3][H][W] fp32 (planar) -> [n][H][W][16] bf16 with channels 3..15 zero
__global__ void k_pack_input(int n, int H, int W, const float *x, uint16_t *xp)
{
    const long long total = (long long)n * H * W * 16;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(e % 16);
        const long long p = e / 16;
        const long long hw = (long long)H * W;
        const long long im = p / hw, r = p - im * hw;
        xp[e] = (c < 3) ? u_f2bf(x[(im * 3 + c) * hw + r]) : (uint16_t)0;
    }
}

// loss = mean (y[p][0] - t[p])^2 ; dy[p][c] = c == 0 ? 2 (y - t) / P : 0
__global__ void k_mse(long long P, const float *y, const float *t, uint16_t *dy, double *part)
{
    __shared__ double sh[256];
    double acc = 0.0;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < P; p += (long long)gridDim.x * blockDim.x) {
        const float d = y[p * 16] - t[p];
        acc += (double)d * (double)d;
        if (dy) {
            dy[p * 16] = u_f2bf(2.0f * d / (float)P);
#pragma unroll
            for (int c = 1; c < 16; ++c) dy[p * 16 + c] = 0;
        }
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_extract(long long P, const float *y, float *out)
{
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < P; p += (long long)gridDim.x * blockDim.x)
        out[p] = y[p * 16];
}

// fp32 master W [9][Co][Ci] -> bf16 W and bf16 W^T for the data gradient:
// wt[8 - tap][ci][co] = w[tap][co][ci]
__global__ void k_cast_weights(int Co, int Ci, const float *w32, uint16_t *w16, uint16_t *wt16)
{
    const long long total = 9LL * Co * Ci;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int ci = (int)(e % Ci);
        const int co = (int)((e / Ci) % Co);
        const int tap = (int)(e / ((long long)Ci * Co));
        const uint16_t v = u_f2bf(w32[e]);
        w16[e] = v;
        wt16[((long long)(8 - tap) * Ci + ci) * Co + co] = v;
    }
}

__global__ void k_adam(long long n, float *w, const float *g, float *m, float *v, float lr_t, float b1, float b2,
                       float eps)
{
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const float gi = g[e];
        const float mi = b1 * m[e] + (1.0f - b1) * gi;
        const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
        m[e] = mi;
        v[e] = vi;
        w[e] = w[e] - lr_t * mi / (sqrtf(vi) + eps);
    }
}

hipError_t launch_pack_input(int n, int H, int W, const float *x, uint16_t *xp, hipStream_t st)
{
    hipLaunchKernelGGL(k_pack_input, grid_for((long long)n * H * W * 16), dim3(256), 0, st, n, H, W, x, xp);
    return hipGetLastError();
}
hipError_t launch_mse(long long P, const float *y, const float *t, uint16_t *dy, double *part, int nblocks,
                      hipStream_t st)
{
    hipLaunchKernelGGL(k_mse, dim3(nblocks), dim3(256), 0, st, P, y, t, dy, part);
    return hipGetLastError();
}
hipError_t launch_extract(long long P, const float *y, float *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_extract, grid_for(P), dim3(256), 0, st, P, y, out);
    return hipGetLastError();
}
hipError_t launch_cast_weights(int Co, int Ci, const float *w32, uint16_t *w16, uint16_t *wt16, hipStream_t st)
{
    hipLaunchKernelGGL(k_cast_weights, grid_for(9LL * Co * Ci), dim3(256), 0, st, Co, Ci, w32, w16, wt16);
    return hipGetLastError();
}
hipError_t launch_adam(long long n, float *w, const float *g, float *m, float *v, float lr_t, float b1, float b2,
                       float eps, hipStream_t st)
{
    hipLaunchKernelGGL(k_adam, grid_for(n), dim3(256), 0, st, n, w, g, m, v, lr_t, b1, b2, eps);
    return hipGetLastError();
}

}  // namespace dn
