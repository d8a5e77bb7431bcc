// Learned diagonal-Hessian preconditioner: host side of the U-Net (network
// definition, parameters, forward, backward, Adam) on the kernels of
// dn_unet.hip.  Not in the reference; see include/denise_sh.h for the ABI.
//
// Network (channels-last BF16 activations, FP32 master weights):
//   enc1  3(->16 padded) -> 16 -> 16      full resolution
//   enc2  16 -> 32 -> 32                  1/2   (2x2 average pooling)
//   enc3  32 -> 48 -> 48                  1/4
//   mid   48 -> 64 -> 64                  1/8
//   dec3  [up(64) | enc3 48] 112 -> 48 -> 48
//   dec2  [up(48) | enc2 32]  80 -> 32 -> 32
//   dec1  [up(32) | enc1 16]  48 -> 16 -> 16
//   head  16 -> 1 (3x3, linear; computed as 16 outputs, channel 0 used)
// All convolutions 3x3 "same" with bias; ReLU after every one but the head.
#include "../../include/denise_sh.h"
#include "dn_core.h"
#include "dn_unet.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

using namespace dn;

namespace {

struct LSpec {
    int ci, co, ci_real, relu;
};
const LSpec kLayers[] = {
    {16, 16, 3, 1},  {16, 16, 16, 1}, {16, 32, 16, 1}, {32, 32, 32, 1}, {32, 48, 32, 1},
    {48, 48, 48, 1}, {48, 64, 48, 1}, {64, 64, 64, 1}, {112, 48, 112, 1}, {48, 48, 48, 1},
    {80, 32, 80, 1}, {32, 32, 32, 1}, {48, 16, 48, 1}, {16, 16, 16, 1}, {16, 16, 16, 0},
};
constexpr int kNL = sizeof(kLayers) / sizeof(kLayers[0]);

template <typename T> hipError_t dal(T **p, size_t n) { return hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)); }
template <typename T> void dfr(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

struct dn_unet {
    hipStream_t stream = nullptr;
    size_t woff[kNL], boff[kNL];
    size_t nparam = 0;
    float *w32 = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *w16 = nullptr, *wt16 = nullptr;
    int step = 0;
    // activation buffers for one (n, H, W)
    int bn = 0, bH = 0, bW = 0;
    uint16_t *x = nullptr, *a[kNL] = {nullptr}, *p1 = nullptr, *p2 = nullptr, *p3 = nullptr;
    uint16_t *c8 = nullptr, *c10 = nullptr, *c12 = nullptr;
    uint16_t *d[kNL + 1] = {nullptr}, *dc8 = nullptr, *dc10 = nullptr, *dc12 = nullptr, *dp1 = nullptr,
             *dp2 = nullptr, *dp3 = nullptr;
    float *out32 = nullptr, *xin = nullptr, *tgt = nullptr, *yout = nullptr, *part = nullptr, *bpart = nullptr;
    double *lpart = nullptr;
    size_t part_len = 0;
};

#define UCHK(expr)                                                                                           \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return dn_fail(nullptr, "unet: %s failed: %s", #expr, hipGetErrorString(e_));  \
    } while (0)

static void free_buffers(dn_unet *u)
{
    dfr(u->x);
    for (int l = 0; l < kNL; ++l) dfr(u->a[l]);
    for (int l = 0; l <= kNL; ++l) dfr(u->d[l]);
    dfr(u->p1); dfr(u->p2); dfr(u->p3); dfr(u->c8); dfr(u->c10); dfr(u->c12);
    dfr(u->dc8); dfr(u->dc10); dfr(u->dc12); dfr(u->dp1); dfr(u->dp2); dfr(u->dp3);
    dfr(u->out32); dfr(u->xin); dfr(u->tgt); dfr(u->yout); dfr(u->part); dfr(u->bpart); dfr(u->lpart);
    u->bn = u->bH = u->bW = 0;
}

// resolution level of layer l's input and output (1/2^level of the input)
static int level_of(int l)
{
    static const int lev[kNL] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0};
    return lev[l];
}

// pixels of the level holding layer l's input and output
static long long level_pixels(int l, long long P0) { return P0 >> (2 * level_of(l)); }

static int ensure_buffers(dn_unet *u, int n, int H, int W)
{
    if (u->bn == n && u->bH == H && u->bW == W) return 0;
    free_buffers(u);
    const long long P0 = (long long)n * H * W, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
    UCHK(dal(&u->x, P0 * 16));
    for (int l = 0; l < kNL; ++l) UCHK(dal(&u->a[l], level_pixels(l, P0) * kLayers[l].co));
    for (int l = 0; l <= kNL; ++l) {
        const int ll = std::min(l, kNL - 1);
        UCHK(dal(&u->d[l], level_pixels(ll, P0) * kLayers[ll].co));
    }
    UCHK(dal(&u->p1, P1 * 16));
    UCHK(dal(&u->p2, P2 * 32));
    UCHK(dal(&u->p3, P3 * 48));
    UCHK(dal(&u->c8, P2 * 112));
    UCHK(dal(&u->c10, P1 * 80));
    UCHK(dal(&u->c12, P0 * 48));
    UCHK(dal(&u->dc8, P2 * 112));
    UCHK(dal(&u->dc10, P1 * 80));
    UCHK(dal(&u->dc12, P0 * 48));
    UCHK(dal(&u->dp1, P1 * 16));
    UCHK(dal(&u->dp2, P2 * 32));
    UCHK(dal(&u->dp3, P3 * 48));
    UCHK(dal(&u->out32, P0 * 16));
    UCHK(dal(&u->xin, P0 * 3));
    UCHK(dal(&u->tgt, P0));
    UCHK(dal(&u->yout, P0));
    size_t plen = 0, blen = 0;
    for (int l = 0; l < kNL; ++l) {
        const size_t nch = (size_t)wgrad_chunks(n, H >> level_of(l));
        plen = std::max(plen, nch * 9 * kLayers[l].co * kLayers[l].ci);
        blen = std::max(blen, nch * kLayers[l].co);
    }
    UCHK(dal(&u->part, plen));
    UCHK(dal(&u->bpart, blen));
    UCHK(dal(&u->lpart, 256));
    u->part_len = plen;
    u->bn = n;
    u->bH = H;
    u->bW = W;
    return 0;
}

static int cast_all(dn_unet *u)
{
    for (int l = 0; l < kNL; ++l)
        UCHK(launch_cast_weights(kLayers[l].co, kLayers[l].ci, u->w32 + u->woff[l], u->w16 + u->woff[l],
                                 u->wt16 + u->woff[l], u->stream));
    return 0;
}

static int conv(dn_unet *u, int l, int n, int H, int W, const uint16_t *x, uint16_t *y16, float *y32)
{
    ConvArgs c{};
    c.n = n;
    c.H = H;
    c.W = W;
    c.Ci = kLayers[l].ci;
    c.Co = kLayers[l].co;
    c.x = x;
    c.w = u->w16 + u->woff[l];
    c.bias = u->w32 + u->boff[l];
    c.relu = kLayers[l].relu;
    c.y16 = y16;
    c.y32 = y32;
    UCHK(launch_conv3x3(c, u->stream));
    return 0;
}

// data gradient of layer l: dx = conv(dy, W^T) (optionally masked by the
// ReLU of the activation the layer read)
static int conv_bwd(dn_unet *u, int l, int n, int H, int W, const uint16_t *dy, const uint16_t *mask, uint16_t *dx)
{
    ConvArgs c{};
    c.n = n;
    c.H = H;
    c.W = W;
    c.Ci = kLayers[l].co;
    c.Co = kLayers[l].ci;
    c.x = dy;
    c.w = u->wt16 + u->woff[l];
    c.mask = mask;
    c.y16 = dx;
    UCHK(launch_conv3x3(c, u->stream));
    return 0;
}

static int wgrad(dn_unet *u, int l, int n, int H, int W, const uint16_t *x, const uint16_t *dy)
{
    WgradArgs w{};
    w.n = n;
    w.H = H;
    w.W = W;
    w.Ci = kLayers[l].ci;
    w.Co = kLayers[l].co;
    w.x = x;
    w.dy = dy;
    w.chunk_rows = wgrad_chunk_rows(n, H);
    w.part = u->part;
    w.bpart = u->bpart;
    UCHK(launch_wgrad(w, u->g + u->woff[l], u->g + u->boff[l], u->stream));
    return 0;
}

// forward of the packed input u->x; result in u->out32 (channel 0)
static int forward_dev(dn_unet *u, int n, int H, int W)
{
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
    if (conv(u, 0, n, H, W, u->x, u->a[0], nullptr) || conv(u, 1, n, H, W, u->a[0], u->a[1], nullptr)) return -1;
    UCHK(launch_pool(n, H, W, 16, u->a[1], u->p1, u->stream));
    if (conv(u, 2, n, H1, W1, u->p1, u->a[2], nullptr) || conv(u, 3, n, H1, W1, u->a[2], u->a[3], nullptr)) return -1;
    UCHK(launch_pool(n, H1, W1, 32, u->a[3], u->p2, u->stream));
    if (conv(u, 4, n, H2, W2, u->p2, u->a[4], nullptr) || conv(u, 5, n, H2, W2, u->a[4], u->a[5], nullptr)) return -1;
    UCHK(launch_pool(n, H2, W2, 48, u->a[5], u->p3, u->stream));
    if (conv(u, 6, n, H3, W3, u->p3, u->a[6], nullptr) || conv(u, 7, n, H3, W3, u->a[6], u->a[7], nullptr)) return -1;
    UCHK(launch_upcat(n, H2, W2, 64, 48, u->a[7], u->a[5], u->c8, u->stream));
    if (conv(u, 8, n, H2, W2, u->c8, u->a[8], nullptr) || conv(u, 9, n, H2, W2, u->a[8], u->a[9], nullptr)) return -1;
    UCHK(launch_upcat(n, H1, W1, 48, 32, u->a[9], u->a[3], u->c10, u->stream));
    if (conv(u, 10, n, H1, W1, u->c10, u->a[10], nullptr) || conv(u, 11, n, H1, W1, u->a[10], u->a[11], nullptr))
        return -1;
    UCHK(launch_upcat(n, H, W, 32, 16, u->a[11], u->a[1], u->c12, u->stream));
    if (conv(u, 12, n, H, W, u->c12, u->a[12], nullptr) || conv(u, 13, n, H, W, u->a[12], u->a[13], nullptr))
        return -1;
    if (conv(u, 14, n, H, W, u->a[13], nullptr, u->out32)) return -1;
    return 0;
}

// backward from d[kNL] (gradient of the head output) into u->g
static int backward_dev(dn_unet *u, int n, int H, int W)
{
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
    uint16_t **d = u->d;
    // head and dec1
    if (wgrad(u, 14, n, H, W, u->a[13], d[14]) || conv_bwd(u, 14, n, H, W, d[14], u->a[13], d[13])) return -1;
    if (wgrad(u, 13, n, H, W, u->a[12], d[13]) || conv_bwd(u, 13, n, H, W, d[13], u->a[12], d[12])) return -1;
    if (wgrad(u, 12, n, H, W, u->c12, d[12]) || conv_bwd(u, 12, n, H, W, d[12], nullptr, u->dc12)) return -1;
    UCHK(launch_up_bwd(n, H, W, 32, 48, u->dc12, u->a[11], d[11], u->stream));
    // dec2
    if (wgrad(u, 11, n, H1, W1, u->a[10], d[11]) || conv_bwd(u, 11, n, H1, W1, d[11], u->a[10], d[10])) return -1;
    if (wgrad(u, 10, n, H1, W1, u->c10, d[10]) || conv_bwd(u, 10, n, H1, W1, d[10], nullptr, u->dc10)) return -1;
    UCHK(launch_up_bwd(n, H1, W1, 48, 80, u->dc10, u->a[9], d[9], u->stream));
    // dec3
    if (wgrad(u, 9, n, H2, W2, u->a[8], d[9]) || conv_bwd(u, 9, n, H2, W2, d[9], u->a[8], d[8])) return -1;
    if (wgrad(u, 8, n, H2, W2, u->c8, d[8]) || conv_bwd(u, 8, n, H2, W2, d[8], nullptr, u->dc8)) return -1;
    UCHK(launch_up_bwd(n, H2, W2, 64, 112, u->dc8, u->a[7], d[7], u->stream));
    // mid
    if (wgrad(u, 7, n, H3, W3, u->a[6], d[7]) || conv_bwd(u, 7, n, H3, W3, d[7], u->a[6], d[6])) return -1;
    if (wgrad(u, 6, n, H3, W3, u->p3, d[6]) || conv_bwd(u, 6, n, H3, W3, d[6], nullptr, u->dp3)) return -1;
    UCHK(launch_pool_bwd(n, H2, W2, 48, u->dp3, u->dc8, 112, 64, u->a[5], d[5], u->stream));
    // enc3
    if (wgrad(u, 5, n, H2, W2, u->a[4], d[5]) || conv_bwd(u, 5, n, H2, W2, d[5], u->a[4], d[4])) return -1;
    if (wgrad(u, 4, n, H2, W2, u->p2, d[4]) || conv_bwd(u, 4, n, H2, W2, d[4], nullptr, u->dp2)) return -1;
    UCHK(launch_pool_bwd(n, H1, W1, 32, u->dp2, u->dc10, 80, 48, u->a[3], d[3], u->stream));
    // enc2
    if (wgrad(u, 3, n, H1, W1, u->a[2], d[3]) || conv_bwd(u, 3, n, H1, W1, d[3], u->a[2], d[2])) return -1;
    if (wgrad(u, 2, n, H1, W1, u->p1, d[2]) || conv_bwd(u, 2, n, H1, W1, d[2], nullptr, u->dp1)) return -1;
    UCHK(launch_pool_bwd(n, H, W, 16, u->dp1, u->dc12, 48, 32, u->a[1], d[1], u->stream));
    // enc1
    if (wgrad(u, 1, n, H, W, u->a[0], d[1]) || conv_bwd(u, 1, n, H, W, d[1], u->a[0], d[0])) return -1;
    if (wgrad(u, 0, n, H, W, u->x, d[0])) return -1;
    return 0;
}

static int check_shape(int n, int H, int W)
{
    if (n < 1 || H < 8 || W < 8 || H % 8 || W % 8)
        return dn_fail(nullptr, "unet: n >= 1 and H, W multiples of 8 (>= 8) required, got %d x %d x %d", n, H, W);
    return 0;
}

static int loss_dev(dn_unet *u, long long P, bool grad, double *loss)
{
    const int nb = 256;
    UCHK(launch_mse(P, u->out32, u->tgt, grad ? u->d[14] : nullptr, u->lpart, nb, u->stream));
    std::vector<double> h(nb);
    UCHK(hipMemcpyAsync(h.data(), u->lpart, nb * sizeof(double), hipMemcpyDeviceToHost, u->stream));
    UCHK(hipStreamSynchronize(u->stream));
    double s = 0.0;
    for (double v : h) s += v;
    if (loss) *loss = s / (double)P;
    return 0;
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" dn_unet *dn_unet_create(unsigned int seed)
{
    dn_unet *u = new dn_unet();
    size_t off = 0;
    for (int l = 0; l < kNL; ++l) {
        u->woff[l] = off;
        off += 9 * (size_t)kLayers[l].co * kLayers[l].ci;
        u->boff[l] = off;
        off += kLayers[l].co;
    }
    u->nparam = off;
    // He-normal weights (Box-Muller on a Mersenne twister), zero biases;
    // padded input channels and unused head outputs are zero
    std::mt19937 rng(seed);
    auto unif = [&]() { return ((double)rng() + 0.5) / 4294967296.0; };
    std::vector<float> h(off, 0.0f);
    for (int l = 0; l < kNL; ++l) {
        const LSpec &s = kLayers[l];
        const double sd = (l == kNL - 1) ? 0.1 * std::sqrt(1.0 / (9.0 * s.ci_real)) : std::sqrt(2.0 / (9.0 * s.ci_real));
        for (int tap = 0; tap < 9; ++tap)
            for (int co = 0; co < s.co; ++co)
                for (int ci = 0; ci < s.ci; ++ci) {
                    const double z = std::sqrt(-2.0 * std::log(unif())) * std::cos(6.283185307179586 * unif());
                    const bool live = ci < s.ci_real && (l != kNL - 1 || co == 0);
                    h[u->woff[l] + ((size_t)tap * s.co + co) * s.ci + ci] = live ? (float)(sd * z) : 0.0f;
                }
    }
    bool ok = hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking) == hipSuccess &&
              dal(&u->w32, off) == hipSuccess && dal(&u->g, off) == hipSuccess && dal(&u->m, off) == hipSuccess &&
              dal(&u->v, off) == hipSuccess && dal(&u->w16, off) == hipSuccess && dal(&u->wt16, off) == hipSuccess;
    ok = ok && hipMemcpy(u->w32, h.data(), off * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    std::fill(h.begin(), h.end(), 0.0f);
    ok = ok && hipMemcpy(u->m, h.data(), off * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(u->v, h.data(), off * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(u->g, h.data(), off * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && cast_all(u) == 0 && hipStreamSynchronize(u->stream) == hipSuccess;
    if (!ok) {
        dn_fail(nullptr, "dn_unet_create: device initialisation failed");
        dn_unet_destroy(u);
        return nullptr;
    }
    return u;
}

extern "C" void dn_unet_destroy(dn_unet *u)
{
    if (!u) return;
    if (u->stream) (void)hipStreamSynchronize(u->stream);
    free_buffers(u);
    dfr(u->w32); dfr(u->g); dfr(u->m); dfr(u->v); dfr(u->w16); dfr(u->wt16);
    if (u->stream) (void)hipStreamDestroy(u->stream);
    delete u;
}

extern "C" long long dn_unet_num_params(const dn_unet *u) { return u ? (long long)u->nparam : -1; }

extern "C" int dn_unet_get_params(dn_unet *u, float *out)
{
    if (!u || !out) return dn_fail(nullptr, "dn_unet_get_params: NULL");
    UCHK(hipMemcpy(out, u->w32, u->nparam * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int dn_unet_set_params(dn_unet *u, const float *in)
{
    if (!u || !in) return dn_fail(nullptr, "dn_unet_set_params: NULL");
    UCHK(hipMemcpy(u->w32, in, u->nparam * sizeof(float), hipMemcpyHostToDevice));
    if (cast_all(u)) return -1;
    UCHK(hipStreamSynchronize(u->stream));
    return 0;
}

// device-side entry used by the FWI preconditioner: x_dev [n][3][H][W] fp32
int dn_unet_forward_device(dn_unet *u, int n, int H, int W, const float *x_dev, float *y_dev)
{
    if (check_shape(n, H, W) || ensure_buffers(u, n, H, W)) return -1;
    UCHK(launch_pack_input(n, H, W, x_dev, u->x, u->stream));
    if (forward_dev(u, n, H, W)) return -1;
    UCHK(launch_extract((long long)n * H * W, u->out32, y_dev, u->stream));
    UCHK(hipStreamSynchronize(u->stream));
    return 0;
}

extern "C" int dn_unet_forward(dn_unet *u, int n, int H, int W, const float *x, float *y)
{
    if (!u || !x || !y) return dn_fail(nullptr, "dn_unet_forward: NULL");
    if (check_shape(n, H, W) || ensure_buffers(u, n, H, W)) return -1;
    const size_t P = (size_t)n * H * W;
    UCHK(hipMemcpyAsync(u->xin, x, P * 3 * sizeof(float), hipMemcpyHostToDevice, u->stream));
    if (dn_unet_forward_device(u, n, H, W, u->xin, u->yout)) return -1;
    UCHK(hipMemcpy(y, u->yout, P * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

static int loss_grad_common(dn_unet *u, int n, int H, int W, const float *x, const float *t, double *loss)
{
    if (check_shape(n, H, W) || ensure_buffers(u, n, H, W)) return -1;
    const size_t P = (size_t)n * H * W;
    UCHK(hipMemcpyAsync(u->xin, x, P * 3 * sizeof(float), hipMemcpyHostToDevice, u->stream));
    UCHK(hipMemcpyAsync(u->tgt, t, P * sizeof(float), hipMemcpyHostToDevice, u->stream));
    UCHK(launch_pack_input(n, H, W, u->xin, u->x, u->stream));
    if (forward_dev(u, n, H, W)) return -1;
    if (loss_dev(u, (long long)P, true, loss)) return -1;
    if (backward_dev(u, n, H, W)) return -1;
    return 0;
}

extern "C" int dn_unet_loss_grad(dn_unet *u, int n, int H, int W, const float *x, const float *t, double *loss,
                                 float *grad)
{
    if (!u || !x || !t) return dn_fail(nullptr, "dn_unet_loss_grad: NULL");
    if (loss_grad_common(u, n, H, W, x, t, loss)) return -1;
    if (grad) UCHK(hipMemcpyAsync(grad, u->g, u->nparam * sizeof(float), hipMemcpyDeviceToHost, u->stream));
    UCHK(hipStreamSynchronize(u->stream));
    return 0;
}

extern "C" int dn_unet_train_step(dn_unet *u, int n, int H, int W, const float *x, const float *t, float lr,
                                  double *loss)
{
    if (!u || !x || !t || !(lr > 0.0f)) return dn_fail(nullptr, "dn_unet_train_step: bad arguments");
    if (loss_grad_common(u, n, H, W, x, t, loss)) return -1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    u->step++;
    const double bc1 = 1.0 - std::pow((double)b1, u->step), bc2 = 1.0 - std::pow((double)b2, u->step);
    const float lr_t = (float)(lr * std::sqrt(bc2) / bc1);
    UCHK(launch_adam((long long)u->nparam, u->w32, u->g, u->m, u->v, lr_t, b1, b2, eps, u->stream));
    if (cast_all(u)) return -1;
    UCHK(hipStreamSynchronize(u->stream));
    return 0;
}
