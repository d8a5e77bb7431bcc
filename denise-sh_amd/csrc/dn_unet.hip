// Kernels of the learned diagonal-Hessian preconditioner: a 4-level U-Net of
// 3x3 convolutions in BF16 with FP32 accumulation on the CDNA4 matrix cores
// (v_mfma_f32_16x16x32_bf16), stored channels-last ([n][H][W][C], C a
// multiple of 16), plus the pooling, upsampling, loss and Adam kernels of
// its training.  Host side: dn_unet.cpp.
#include "dn_unet.h"

#include <algorithm>

namespace dn {

typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t u_f2bf(float x)
{
    // round to nearest even (finite inputs)
    unsigned int u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float u_bf2f(uint16_t h) { return __uint_as_float(((unsigned int)h) << 16); }

// 8 BF16 values: one 16-byte MFMA operand per lane
typedef unsigned short u8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f4v mfma32(u8v a, u8v b, f4v c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, a), __builtin_bit_cast(bf8v, b), c, 0, 0,
                                                   0);
}

// ---------------------------------------------------------------------------
// 3x3 "same" convolution as an implicit GEMM on MFMA 16x16x32 (bf16 in, fp32
// accumulate): D[co][px] = sum over k = (tap, ci) of W[tap][co][ci] *
// X[px + shift(tap)][ci].  K is walked in groups of 8 input channels of one
// tap; one MFMA takes 4 groups: lane quarter kq = lane / 16 supplies group
// 4i + kq, A = W[tap][co = lane % 16][8 channels] and B = X[pixel lane % 16
// + shift][8 channels] (16 bytes each), and receives D[4kq + r][lane % 16],
// r = 0..3.  Groups past the last tap are zero.
// Workgroup: 4 waves over 256 consecutive pixels of one image row (64 per
// wave, 4 pixel tiles) and a group of up to 64 output channels (blockIdx.z,
// up to 4 channel tiles): 16 accumulator tiles per wave.  The weights of one
// kernel row (3 taps x 64 channels x CI) are staged in LDS once and read by
// all four waves; the input comes with one 16-byte load per lane and
// operand, the three taps of a row re-reading a pixel through L1.
// ---------------------------------------------------------------------------
template <int CI>
__global__ __launch_bounds__(256) void k_conv3x3(ConvArgs a)
{
    constexpr int CB = CI / 8;          // 8-channel groups per tap
    constexpr int NG = 3 * CB;          // groups of one kernel row
    constexpr int NKI = (NG + 3) / 4;   // MFMAs per kernel row
    constexpr int SW = 65;              // LDS entries per group (64 channels + 1 against bank conflicts)
    __shared__ u8v sw[NKI * 4 * SW];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nl = lane & 15, kq = lane >> 4;
    const int row = blockIdx.y;
    const int img = row / a.H, y = row - img * a.H;
    const int x0 = blockIdx.x * 256 + wave * 64;
    const int cg = blockIdx.z * 64;
    const int nmt = min(4, (a.Co - cg) >> 4);
    f4v acc[4][4];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[m][t] = f4v{0.0f, 0.0f, 0.0f, 0.0f};
    for (int ky = 0; ky < 3; ++ky) {
        const int yy = y + ky - 1;
        if (yy < 0 || yy >= a.H) continue;   // uniform in the workgroup
        __syncthreads();                     // the previous kernel row's weights are consumed
        for (int e = threadIdx.x; e < NKI * 4 * 64; e += 256) {
            const int g = e >> 6, co = e & 63;
            u8v v = u8v{0, 0, 0, 0, 0, 0, 0, 0};
            if (g < NG && cg + co < a.Co) {
                const int kx = g / CB, cb = g - kx * CB;
                v = *(const u8v *)(a.w + ((long long)(3 * ky + kx) * a.Co + cg + co) * CI + 8 * cb);
            }
            sw[g * SW + co] = v;
        }
        __syncthreads();
        const uint16_t *xrow = a.x + (long long)(img * a.H + yy) * a.W * CI;
#pragma unroll
        for (int i = 0; i < NKI; ++i) {
            const int g = 4 * i + kq;
            const bool live = g < NG;
            const int kx = live ? g / CB : 0, cb = live ? g - (g / CB) * CB : 0;
            u8v b[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int xx = x0 + 16 * t + nl + kx - 1;
                b[t] = u8v{0, 0, 0, 0, 0, 0, 0, 0};
                if (live && xx >= 0 && xx < a.W) b[t] = *(const u8v *)(xrow + (long long)xx * CI + 8 * cb);
            }
#pragma unroll
            for (int m = 0; m < 4; ++m)
                if (m < nmt) {
                    const u8v av = sw[g * SW + 16 * m + nl];
#pragma unroll
                    for (int t = 0; t < 4; ++t) acc[m][t] = mfma32(av, b[t], acc[m][t]);
                }
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int xo = x0 + 16 * t + nl;
        if (xo >= a.W) continue;
        const long long pix = (long long)(img * a.H + y) * a.W + xo;
#pragma unroll
        for (int m = 0; m < 4; ++m)
            if (m < nmt) {
                const int co0 = cg + m * 16 + 4 * kq;
                float v[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    v[r] = acc[m][t][r];
                    if (a.bias) v[r] += a.bias[co0 + r];
                    if (a.relu) v[r] = fmaxf(v[r], 0.0f);
                }
                if (a.mask) {
                    const ushort4 mk = *(const ushort4 *)(a.mask + pix * a.Co + co0);
                    if (!(u_bf2f(mk.x) > 0.0f)) v[0] = 0.0f;
                    if (!(u_bf2f(mk.y) > 0.0f)) v[1] = 0.0f;
                    if (!(u_bf2f(mk.z) > 0.0f)) v[2] = 0.0f;
                    if (!(u_bf2f(mk.w) > 0.0f)) v[3] = 0.0f;
                }
                if (a.y16)
                    *(ushort4 *)(a.y16 + pix * a.Co + co0) =
                        make_ushort4(u_f2bf(v[0]), u_f2bf(v[1]), u_f2bf(v[2]), u_f2bf(v[3]));
                if (a.y32) *(float4 *)(a.y32 + pix * a.Co + co0) = make_float4(v[0], v[1], v[2], v[3]);
            }
    }
}

hipError_t launch_conv3x3(const ConvArgs &a, hipStream_t st)
{
    if (a.Co % 16 || a.Co < 16 || a.n < 1 || a.H < 1 || a.W < 1) return hipErrorInvalidValue;
    dim3 grid((unsigned)((a.W + 255) / 256), (unsigned)(a.n * a.H), (unsigned)((a.Co + 63) / 64));
    switch (a.Ci) {
    case 16: hipLaunchKernelGGL(k_conv3x3<16>, grid, dim3(256), 0, st, a); break;
    case 32: hipLaunchKernelGGL(k_conv3x3<32>, grid, dim3(256), 0, st, a); break;
    case 48: hipLaunchKernelGGL(k_conv3x3<48>, grid, dim3(256), 0, st, a); break;
    case 64: hipLaunchKernelGGL(k_conv3x3<64>, grid, dim3(256), 0, st, a); break;
    case 80: hipLaunchKernelGGL(k_conv3x3<80>, grid, dim3(256), 0, st, a); break;
    case 112: hipLaunchKernelGGL(k_conv3x3<112>, grid, dim3(256), 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Weight gradient dW[tap][co][ci] = sum_p dY[p][co] * X[p + shift(tap)][ci]
// as a GEMM over pixels (K) on MFMA 16x16x32, with the bias gradient
// db[co] = sum_p dY[p][co].  Workgroup (4 waves): one chunk of image rows
// (blockIdx.x) and one kernel row ky (blockIdx.y), walked in segments of 64
// pixels of a row.  Per segment the output gradient dY[64 px][CO] and, for
// each kx, the shifted input row X[y + ky - 1][x + kx - 1][CI] are staged in
// LDS transposed ([channel][pixel], zero outside the image; one 16-byte load
// of 8 channels per thread and item), so that a lane reads 8 consecutive
// pixels of one channel as one 16-byte operand: A = dY^T (co = lane % 16),
// B = X_shift (ci = lane % 16).  The 3 x (CO/16) x (CI/16) tiles of the
// kernel row are dealt round-robin to the waves (at most 16 each) and
// accumulated over the chunk; each chunk writes a partial, added over the
// chunks in chunk order (deterministic).
// ---------------------------------------------------------------------------
constexpr int kSegPx = 64;     // pixels per segment
constexpr int kSegPad = 72;    // LDS pixels per channel row (16-byte aligned rows, staggered banks)

template <int CI, int CO>
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a)
{
    constexpr int NMT = CO / 16, NNT = CI / 16, NTILES = 3 * NMT * NNT;
    static_assert(NTILES <= 64, "at most 16 tiles per wave");
    __shared__ __attribute__((aligned(16))) uint16_t sdy[CO * kSegPad];
    __shared__ __attribute__((aligned(16))) uint16_t sx[3 * CI * kSegPad];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nl = lane & 15, kq = lane >> 4;
    const int ky = blockIdx.y;
    const int rows_total = a.n * a.H;
    const int r0 = blockIdx.x * a.chunk_rows, r1 = min(rows_total, r0 + a.chunk_rows);
    f4v acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = f4v{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    for (int r = r0; r < r1; ++r) {
        const int img = r / a.H, y = r - img * a.H;
        const int yy = y + ky - 1;
        const bool yin = yy >= 0 && yy < a.H;
        for (int x0 = 0; x0 < a.W; x0 += kSegPx) {
            __syncthreads();   // the previous segment is consumed
            // dY[y][x0 + j][8c .. 8c+7] -> sdy[8c + k][j]
            for (int e = threadIdx.x; e < kSegPx * (CO / 8); e += 256) {
                const int j = e / (CO / 8), c8 = e - j * (CO / 8);
                const int x = x0 + j;
                u8v v = u8v{0, 0, 0, 0, 0, 0, 0, 0};
                if (x < a.W) v = *(const u8v *)(a.dy + ((long long)r * a.W + x) * CO + 8 * c8);
#pragma unroll
                for (int k = 0; k < 8; ++k) sdy[(8 * c8 + k) * kSegPad + j] = v[k];
            }
            // X[yy][x0 + jj - 1][8c .. 8c+7] -> sx[kx][8c + k][jj - kx]
            for (int e = threadIdx.x; e < (kSegPx + 2) * (CI / 8); e += 256) {
                const int jj = e / (CI / 8), c8 = e - jj * (CI / 8);
                const int x = x0 + jj - 1;
                u8v v = u8v{0, 0, 0, 0, 0, 0, 0, 0};
                if (yin && x >= 0 && x < a.W) v = *(const u8v *)(a.x + ((long long)(img * a.H + yy) * a.W + x) * CI + 8 * c8);
#pragma unroll
                for (int kx = 0; kx < 3; ++kx) {
                    const int j = jj - kx;
                    if (j >= 0 && j < kSegPx)
#pragma unroll
                        for (int k = 0; k < 8; ++k) sx[(kx * CI + 8 * c8 + k) * kSegPad + j] = v[k];
                }
            }
            __syncthreads();
            if (ky == 0 && (int)threadIdx.x < CO)
#pragma unroll 8
                for (int j = 0; j < kSegPx; ++j) bsum += u_bf2f(sdy[threadIdx.x * kSegPad + j]);
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int tile = wave + 4 * q;
                if (tile < NTILES) {
                    const int kx = tile / (NMT * NNT), rem = tile - kx * (NMT * NNT);
                    const int mt = rem / NNT, nt = rem - mt * NNT;
                    const uint16_t *pa = sdy + (mt * 16 + nl) * kSegPad + 8 * kq;
                    const uint16_t *pb = sx + (kx * CI + nt * 16 + nl) * kSegPad + 8 * kq;
#pragma unroll
                    for (int s = 0; s < kSegPx; s += 32)
                        acc[q] = mfma32(*(const u8v *)(pa + s), *(const u8v *)(pb + s), acc[q]);
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int tile = wave + 4 * q;
        if (tile < NTILES) {
            const int kx = tile / (NMT * NNT), rem = tile - kx * (NMT * NNT);
            const int mt = rem / NNT, nt = rem - mt * NNT;
            const int tap = 3 * ky + kx;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int co = mt * 16 + 4 * kq + j, ci = nt * 16 + nl;
                a.part[(((long long)blockIdx.x * 9 + tap) * CO + co) * CI + ci] = acc[q][j];
            }
        }
    }
    if (ky == 0 && (int)threadIdx.x < CO) a.bpart[(long long)blockIdx.x * CO + threadIdx.x] = bsum;
}

// out[e] = sum over c of part[c][e], c in increasing order: each thread
// walks the chunks of one element (adjacent threads, adjacent elements)
__global__ void k_reduce_chunks(int nchunks, long long len, const float *part, float *out)
{
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < len;
         e += (long long)gridDim.x * blockDim.x) {
        float s = 0.0f;
        int c = 0;
        for (; c + 4 <= nchunks; c += 4) {
            const float p0 = part[(long long)c * len + e], p1 = part[(long long)(c + 1) * len + e];
            const float p2 = part[(long long)(c + 2) * len + e], p3 = part[(long long)(c + 3) * len + e];
            s += p0;
            s += p1;
            s += p2;
            s += p3;
        }
        for (; c < nchunks; ++c) s += part[(long long)c * len + e];
        out[e] = s;
    }
}

// The same sum for a short vector (the bias gradient): one workgroup per
// element; thread t adds chunks t, t + 256, ... in order, then a fixed
// pairwise tree over the threads
__global__ __launch_bounds__(256) void k_reduce_chunks_short(int nchunks, int len, const float *part, float *out)
{
    __shared__ float sh[256];
    const int e = blockIdx.x;
    float s = 0.0f;
    for (int c = threadIdx.x; c < nchunks; c += 256) s += part[(long long)c * len + e];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[e] = sh[0];
}

int wgrad_chunk_rows(int n, int H)
{
    return std::max(1, (n * H + 255) / 256);
}

int wgrad_chunks(int n, int H)
{
    const int cr = wgrad_chunk_rows(n, H);
    return (n * H + cr - 1) / cr;
}

template <int CI, int CO> static void wgrad_launch(const WgradArgs &a, int nch, hipStream_t st)
{
    hipLaunchKernelGGL((k_wgrad<CI, CO>), dim3(nch, 3), dim3(256), 0, st, a);
}

hipError_t launch_wgrad(const WgradArgs &a, float *dw, float *db, hipStream_t st)
{
    if (a.chunk_rows != wgrad_chunk_rows(a.n, a.H) || a.n < 1 || a.H < 1 || a.W < 1) return hipErrorInvalidValue;
    const int nch = wgrad_chunks(a.n, a.H);
    const int key = a.Ci * 1000 + a.Co;
    switch (key) {   // the (input, output) channel pairs of the network's layers
    case 16016: wgrad_launch<16, 16>(a, nch, st); break;
    case 16032: wgrad_launch<16, 32>(a, nch, st); break;
    case 32032: wgrad_launch<32, 32>(a, nch, st); break;
    case 32048: wgrad_launch<32, 48>(a, nch, st); break;
    case 48048: wgrad_launch<48, 48>(a, nch, st); break;
    case 48064: wgrad_launch<48, 64>(a, nch, st); break;
    case 64064: wgrad_launch<64, 64>(a, nch, st); break;
    case 112048: wgrad_launch<112, 48>(a, nch, st); break;
    case 80032: wgrad_launch<80, 32>(a, nch, st); break;
    case 48016: wgrad_launch<48, 16>(a, nch, st); break;
    default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const long long len = 9LL * a.Co * a.Ci;
    hipLaunchKernelGGL(k_reduce_chunks, dim3((unsigned)std::min<long long>(1024, (len + 255) / 256)), dim3(256), 0, st,
                       nch, len, a.part, dw);
    e = hipGetLastError();
    if (e != hipSuccess || !db) return e;
    hipLaunchKernelGGL(k_reduce_chunks_short, dim3(a.Co), dim3(256), 0, st, nch, a.Co, a.bpart, db);
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
