// Gradient post-processing on the device: the spatially variable Gaussian
// smoothing of smooth_grad.c (SPATFILTER 4, gauss_filt_var.c:10-215), its
// fixed Gaussian (SPATFILTER 3, gauss_filt.c:56-160: one weight table for
// every cell, normalised before the sum), its tridiagonal smoothing
// (SPATFILTER 2, smooth2.c:28-165) and the gradient tapers of
// taper_grad.c:11-584 (sws 1, 2, 3 and a user map).
//
// smooth2.  Every x system (one per grid row) and every z system (one per
// column) has the same matrix, so the host factors the two matrices once
// with tripd's float operations (util.c:638-671) and the device runs only
// the substitutions: one thread per row for x, then one per column for z,
// in tripd's order.  The reference's index ranges are kept: its gradient
// sits at v[1..NX][1..NY] of a zeroed v[0..NX][0..NY] and it smooths over
// 0..n-1, so the zero row and column take part and the last row and
// column keep their values.
//
// Gaussian smoothing.  Each cell's half widths (hx, hz) follow its local
// wavelength vs / FC_END; its kernel weights exp(-ii^2/sx - jj^2/sz) depend
// only on (hx, hz).  The host turns the vs model into a per-cell index of
// the distinct (hx, hz) pairs and builds one weight table per pair with the
// same float operands, the double exp() rounded to float and the same
// sequential float sum as the reference, so the weights and their sums are
// bit-identical to it.  The device then sums gradient x weight per cell in
// the reference's order (x offset outer, y offset inner; unfused float
// multiply and add) with clamped indices, which is the reference's edge
// continuation of the gradient (gauss_filt_var.c:126-157), and divides by
// the sum: the result is bit-identical to the reference's loop.
//
// Memory: the gradient plane (nx*ny floats) is read (2hx+1)(2hz+1) times
// per cell, but a workgroup's neighbourhood of it stays in LDS: a 64 x 4
// tile of cells stages its (64 + 2H) x (4 + 2H) halo window once (H the
// largest half width in the tile), so the HBM traffic is one read of the
// gradient and the cell map plus one write.
//
// Tapers.  The coefficient maps are cheap one-off host computations of the
// reference's formulas; the device multiplies the gradient by the map and
// zeroes the cells the reference zeroes.
#include "dn_core.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <vector>

using namespace dn;

namespace {

constexpr int kTileX = 64, kTileY = 4;
constexpr int kMaxHalf = 1024;          // largest half width accepted
constexpr int kLdsFloats = 16 * 1024;   // 64 KiB staging window (half widths up to 48)

// hx, hz, first weight, bits of the weight sum
struct GPair {
    int hx, hz, woff;
    float sum;
};

__global__ __launch_bounds__(256) void k_gauss_filt_var(int nx, int ny, const float *__restrict__ in,
                                                         float *__restrict__ out, const int *__restrict__ cpair,
                                                         const GPair *__restrict__ pairs,
                                                         const float *__restrict__ w,
                                                         const int *__restrict__ tile_half)
{
    extern __shared__ float win[];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = blockIdx.x * kTileX, j0 = blockIdx.y * kTileY;
    const int i = i0 + tx, j = j0 + ty;
    // largest half width of the tile (from the host): window rows/columns
    const int H = tile_half[blockIdx.y * gridDim.x + blockIdx.x];
    const int wx = kTileX + 2 * H, wy = kTileY + 2 * H;
    const bool staged = (wx * wy <= kLdsFloats);
    if (staged) {
        // window cell (a, b) holds the gradient at the clamped (j0 - H + b, i0 - H + a)
        const int nthr = kTileX * kTileY, lt = ty * kTileX + tx;
        for (int e = lt; e < wx * wy; e += nthr) {
            const int b = e / wx, a = e - b * wx;
            const int y = min(max(j0 - H + b, 0), ny - 1), x = min(max(i0 - H + a, 0), nx - 1);
            win[e] = in[(size_t)y * nx + x];
        }
    }
    __syncthreads();
    if (i >= nx || j >= ny) return;
    const size_t t = (size_t)j * nx + i;
    const int p = cpair[t];
    if (p < 0) {
        out[t] = in[t];
        return;
    }
    const GPair P = pairs[p];
    const float *wp = w + P.woff;
    float conv = 0.0f;
    if (staged) {
        // (i + ii, j + jj) is window cell (tx + H + ii, ty + H + jj); the
        // window already holds the clamped values, so no clamp here
        for (int ii = -P.hx; ii <= P.hx; ++ii) {
            const float *col = win + (ty + H - P.hz) * wx + (tx + H + ii);
            for (int jj = 0; jj <= 2 * P.hz; ++jj) {
                conv = conv + col[jj * wx] * wp[0];
                ++wp;
            }
        }
    } else {
        for (int ii = -P.hx; ii <= P.hx; ++ii) {
            const int x = min(max(i + ii, 0), nx - 1);
            for (int jj = -P.hz; jj <= P.hz; ++jj) {
                const int y = min(max(j + jj, 0), ny - 1);
                conv = conv + in[(size_t)y * nx + x] * wp[0];
                ++wp;
            }
        }
    }
    out[t] = conv / P.sum;
}

// x substitutions of smooth2: thread j (1..ny-1) solves the system of grid
// row j over v[0..nx-1][j], v[0][j] = 0 (the zero column), in place on
// g[j-1][0..nx-2]; e, d: the factored x matrix (tripd's e[k-1] and d[k])
__global__ void k_smooth2_x(int nx, int ny, float *__restrict__ g, const float *__restrict__ e,
                            const float *__restrict__ d, float v0)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (j > ny - 1) return;
    float *row = g + (size_t)(j - 1) * nx;   // v[ix][j] = row[ix - 1] for ix >= 1
    float prev = v0;                         // b[0]
    for (int k = 1; k < nx; ++k) {
        const float b = row[k - 1] - e[k - 1] * prev;
        row[k - 1] = b;
        prev = b;
    }
    float bk = prev / d[nx - 1];             // b[n-1]
    row[nx - 2] = bk;
    for (int k = nx - 1; k >= 2; --k) {      // b[k-1] for k-1 >= 1; b[0] is not kept
        const float b = row[k - 2] / d[k - 1] - e[k - 1] * bk;
        row[k - 2] = b;
        bk = b;
    }
}

// z substitutions of smooth2: thread i (1..nx-1) solves the system of grid
// column i over v[i][1..ny-1] (n = ny - 1), its first right-hand side plus
// r1 w v[i][0] (v[i][0] = v0 = 0, the zero row after its x smoothing)
__global__ void k_smooth2_z(int nx, int ny, float *__restrict__ g, const float *__restrict__ e,
                            const float *__restrict__ d, float r1w, float v0)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (i > nx - 1) return;
    float *col = g + (i - 1);                // v[i][iz + 1] = col[iz * nx]
    const int n = ny - 1;
    float prev = col[0] + (r1w * v0);
    col[0] = prev;
    for (int k = 1; k < n; ++k) {
        const float b = col[(size_t)k * nx] - e[k - 1] * prev;
        col[(size_t)k * nx] = b;
        prev = b;
    }
    float bk = prev / d[n - 1];
    col[(size_t)(n - 1) * nx] = bk;
    for (int k = n - 1; k >= 1; --k) {
        const float b = col[(size_t)(k - 1) * nx] / d[k - 1] - e[k - 1] * bk;
        col[(size_t)(k - 1) * nx] = b;
        bk = b;
    }
}

__global__ void k_taper_apply(size_t n, float *__restrict__ g, const float *__restrict__ tap,
                              const unsigned char *__restrict__ zero)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float v = g[t] * tap[t];
    if (zero && zero[t]) v = 0.0f;
    g[t] = v;
}

// Host plan of the Gaussian smoothing for one vs model.
struct GaussPlan {
    std::vector<int> cell;       // [ny][nx] pair index, -1 copy
    std::vector<GPair> pairs;
    std::vector<float> w;        // weight tables, concatenated
    std::vector<int> tile_half;  // per 64x4 tile: largest half width
    int tiles_x = 0, tiles_y = 0;
};

// gauss_filt_var.c:160-200 per cell; float operands as in the reference
bool make_gauss_plan(int nx, int ny, float DH, float WD_DAMP, float WD_DAMP1, float FC_END, const float *vel,
                     GaussPlan &P, std::string &err)
{
    const size_t n = (size_t)nx * ny;
    P.cell.assign(n, -1);
    P.pairs.clear();
    P.w.clear();
    std::map<std::pair<int, int>, int> index;
    for (size_t t = 0; t < n; ++t) {
        const float lam = vel[t] / FC_END;
        const float rx = (WD_DAMP1 * lam) / DH, rz = (WD_DAMP * lam) / DH;
        if (!(std::fabs(rz) < 1e7f) || !(std::fabs(rx) < 1e7f)) {
            err = "smooth_grad: filter size out of range (check vs, WD_DAMP, WD_DAMP1, FC_END)";
            return false;
        }
        int fx = (int)std::round(rx), fz = (int)std::round(rz);
        if (fz <= 0) continue;
        if (!(fx % 2)) fx += 1;
        if (!(fz % 2)) fz += 1;
        if (fx < 6) fx = 6;
        if (fz < 6) fz = 6;
        const int hx = std::abs(fx) / 2, hz = std::abs(fz) / 2;
        if (hx > kMaxHalf || hz > kMaxHalf) {
            err = "smooth_grad: Gaussian half width above 1024 cells";
            return false;
        }
        auto key = std::make_pair(hx, hz);
        auto it = index.find(key);
        if (it == index.end()) {
            const int id = (int)P.pairs.size();
            const size_t woff = P.w.size();
            if (woff + (size_t)(2 * hx + 1) * (2 * hz + 1) > (size_t)1 << 28) {
                err = "smooth_grad: Gaussian weight tables above 2^28 entries";
                return false;
            }
            const float sigx = (float)(hx / 2), sigz = (float)(hz / 2);
            const float sx = (float)(2.0 * sigx * sigx), sz = (float)(2.0 * sigz * sigz);
            float sum = 0.0f;
            for (int ii = -hx; ii <= hx; ++ii)
                for (int jj = -hz; jj <= hz; ++jj) {
                    const float a = (float)(ii * ii) / sx, b = (float)(jj * jj) / sz;
                    const float k = (float)::exp((double)(-a - b));
                    P.w.push_back(k);
                    sum += k;
                }
            P.pairs.push_back(GPair{hx, hz, (int)woff, sum});
            it = index.emplace(key, id).first;
        }
        P.cell[t] = it->second;
    }
    P.tiles_x = (nx + kTileX - 1) / kTileX;
    P.tiles_y = (ny + kTileY - 1) / kTileY;
    P.tile_half.assign((size_t)P.tiles_x * P.tiles_y, 0);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const int p = P.cell[(size_t)j * nx + i];
            if (p < 0) continue;
            int &h = P.tile_half[(size_t)(j / kTileY) * P.tiles_x + i / kTileX];
            h = std::max(h, std::max(P.pairs[p].hx, P.pairs[p].hz));
        }
    return true;
}

// gauss_filt.c:56-160 (SPATFILTER 3): one Gaussian for every cell.
// FILT_SIZE_GRAD = round(WD_DAMP (C_vs / FC_END) / DH), made odd; sigma =
// hfs / 2 in integer arithmetic; the weights, summed in order from zero (the
// reference leaves its sum uninitialised: this is its result when the sum
// starts at zero), are divided by the sum before the convolution, so the
// pair's divisor is 1.  Returns 1 for FILT_SIZE_GRAD 0 (nothing to do), 0
// for a plan, -1 with err set for sizes the reference cannot filter (a
// negative size skips its normalisation; hfs < 2 gives sigma 0 and 0/0).
int make_gauss_fixed_plan(int nx, int ny, float DH, float WD_DAMP, float FC_END, float C_vs, GaussPlan &P,
                          std::string &err)
{
    const float lam = C_vs / FC_END;
    const float r = (WD_DAMP * lam) / DH;
    if (!(std::fabs(r) < 1e7f)) {
        err = "smooth_grad: filter size out of range (check C_vs, WD_DAMP, FC_END)";
        return -1;
    }
    int fs = (int)std::round(r);
    if (fs == 0) return 1;
    if (fs < 0) {
        err = "smooth_grad: SPATFILTER 3 needs WD_DAMP > 0 (gauss_filt.c normalises only positive sizes)";
        return -1;
    }
    if (!(fs % 2)) fs += 1;
    const int hfs = fs / 2;
    if (hfs < 2) {
        err = "smooth_grad: SPATFILTER 3 needs FILT_SIZE_GRAD >= 5 (below, sigma = hfs / 2 is 0)";
        return -1;
    }
    if (hfs > kMaxHalf) {
        err = "smooth_grad: Gaussian half width above 1024 cells";
        return -1;
    }
    const float sigma = (float)(hfs / 2);
    const float s = (float)(2.0 * sigma * sigma);
    P.cell.assign((size_t)nx * ny, 0);
    P.pairs.assign(1, GPair{hfs, hfs, 0, 1.0f});
    P.w.clear();
    float sum = 0.0f;
    for (int ii = -hfs; ii <= hfs; ++ii)
        for (int jj = -hfs; jj <= hfs; ++jj) {
            const float a = (float)(ii * ii) / s, b = (float)(jj * jj) / s;
            const float k = (float)::exp((double)(-a - b));
            P.w.push_back(k);
            sum += k;
        }
    for (float &w : P.w) w /= sum;
    P.tiles_x = (nx + kTileX - 1) / kTileX;
    P.tiles_y = (ny + kTileY - 1) / kTileY;
    P.tile_half.assign((size_t)P.tiles_x * P.tiles_y, hfs);
    return 0;
}

// smooth2.c:28-165: the x matrix (n = NX: diagonal 1 + 2 r2, 1 + r2 at both
// ends, superdiagonal -r2) and the z matrix (n = NY - 1: 1 + 2 r1, 1 + r1 at
// the end, -r1), built with the reference's float operands (w = 1: the window
// is the whole grid) and factored as tripd does.  r1 = r2 = WD_DAMP^2 / 4.
struct Smooth2Plan {
    std::vector<float> ex, dx, ez, dz;
    float r1 = 0.0f;
};

static void tripd_factor(std::vector<float> &d, std::vector<float> &e)
{
    const int n = (int)d.size();
    for (int k = 1; k < n; ++k) {
        const float temp = e[k - 1];
        e[k - 1] = temp / d[k - 1];
        d[k] -= temp * e[k - 1];
    }
}

bool make_smooth2_plan(int nx, int ny, float WD_DAMP, Smooth2Plan &P, std::string &err)
{
    if (nx < 2 || ny < 3) {
        err = "smooth_grad: SPATFILTER 2 needs NX >= 2 and NY >= 3";
        return false;
    }
    float r1 = WD_DAMP, r2 = r1;
    r1 = (float)(r1 * r1 * 0.25);
    r2 = (float)(r2 * r2 * 0.25);
    const float w = 1.0f;
    P.dx.assign(nx, 0.0f);
    P.ex.assign(nx, 0.0f);
    for (int ix = 0; ix < nx - 1; ++ix) {
        P.dx[ix] = (float)(1.0 + r2 * (w + w));
        P.ex[ix] = -r2 * w;
    }
    P.dx[0] -= r2 * w;
    P.dx[nx - 1] = (float)(1.0 + r2 * w);
    tripd_factor(P.dx, P.ex);
    const int n = ny - 1;
    P.dz.assign(n, 0.0f);
    P.ez.assign(n, 0.0f);
    for (int iz = 0; iz < ny - 2; ++iz) {
        P.dz[iz] = (float)(1.0 + r1 * (w + w));
        P.ez[iz] = -r1 * w;
    }
    P.dz[n - 1] = (float)(1.0 + r1 * w);
    tripd_factor(P.dz, P.ez);
    P.r1 = r1 * w;
    return true;
}

// device copies of a plan plus a scratch plane; grown on demand
struct GaussDev {
    int *cell = nullptr, *tile_half = nullptr;
    GPair *pairs = nullptr;
    float *w = nullptr, *tmp = nullptr;
    size_t cap_cell = 0, cap_tiles = 0, cap_pairs = 0, cap_w = 0, cap_tmp = 0;
    void release()
    {
        (void)hipFree(cell);
        (void)hipFree(tile_half);
        (void)hipFree(pairs);
        (void)hipFree(w);
        (void)hipFree(tmp);
        *this = GaussDev{};
    }
};

template <typename T>
hipError_t grow(T **p, size_t &cap, size_t n)
{
    if (n <= cap && *p) return hipSuccess;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) cap = n;
    return e;
}

hipError_t upload_plan(const GaussPlan &P, GaussDev &D, size_t n, hipStream_t st)
{
    hipError_t e;
    if ((e = grow(&D.cell, D.cap_cell, n)) != hipSuccess) return e;
    if ((e = grow(&D.tile_half, D.cap_tiles, P.tile_half.size())) != hipSuccess) return e;
    if ((e = grow(&D.pairs, D.cap_pairs, P.pairs.size())) != hipSuccess) return e;
    if ((e = grow(&D.w, D.cap_w, P.w.size())) != hipSuccess) return e;
    if ((e = grow(&D.tmp, D.cap_tmp, n)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(D.cell, P.cell.data(), n * sizeof(int), hipMemcpyHostToDevice, st)) != hipSuccess)
        return e;
    if ((e = hipMemcpyAsync(D.tile_half, P.tile_half.data(), P.tile_half.size() * sizeof(int),
                            hipMemcpyHostToDevice, st)) != hipSuccess)
        return e;
    if (!P.pairs.empty()) {
        if ((e = hipMemcpyAsync(D.pairs, P.pairs.data(), P.pairs.size() * sizeof(GPair), hipMemcpyHostToDevice,
                                st)) != hipSuccess)
            return e;
        if ((e = hipMemcpyAsync(D.w, P.w.data(), P.w.size() * sizeof(float), hipMemcpyHostToDevice, st)) !=
            hipSuccess)
            return e;
    }
    // the host vectors are released by the caller only after a stream sync
    return hipSuccess;
}

// filters the device plane g [ny][nx] in place (through D.tmp)
hipError_t run_gauss(int nx, int ny, const GaussPlan &P, GaussDev &D, float *g, hipStream_t st)
{
    const size_t n = (size_t)nx * ny;
    hipError_t e = hipMemcpyAsync(D.tmp, g, n * sizeof(float), hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
    int hmax = 0;
    for (int h : P.tile_half) hmax = std::max(hmax, h);
    const size_t want = (size_t)(kTileX + 2 * hmax) * (kTileY + 2 * hmax);
    const size_t lds = std::min<size_t>(want, kLdsFloats) * sizeof(float);
    hipLaunchKernelGGL(k_gauss_filt_var, dim3(P.tiles_x, P.tiles_y), dim3(kTileX, kTileY), lds, st, nx, ny,
                       D.tmp, g, D.cell, D.pairs, D.w, D.tile_half);
    return hipGetLastError();
}

// taper_grad.c coefficient maps ([ny][nx] row-major) and zeroed cells
struct TaperMap {
    std::vector<float> tap;
    std::vector<unsigned char> zero;
};

float min_first(const std::vector<float> &m)
{
    float v = m[0];
    for (size_t t = 1; t < m.size(); ++t)
        if (m[t] < v) v = m[t];
    return v;
}
float max_first(const float *m, size_t n)
{
    float v = m[0];
    for (size_t t = 1; t < n; ++t)
        if (m[t] > v) v = m[t];
    return v;
}

bool make_taper(int nx, int ny, const dn_taper *T, const std::vector<int> &si, const std::vector<int> &sj,
                const std::vector<int> &ri, const std::vector<int> &rj, TaperMap &M, std::string &err)
{
    const size_t n = (size_t)nx * ny;
    M.tap.assign(n, 0.0f);
    M.zero.assign(n, 0);
    const float DH = T->dh;
    const int G1 = T->gradt[0], G2 = T->gradt[1], G3 = T->gradt[2], G4 = T->gradt[3];
    auto at = [&](int j, int i) -> size_t { return (size_t)(j - 1) * nx + (size_t)(i - 1); };
    auto inside = [&](int j, int i) { return i >= 1 && i <= nx && j >= 1 && j <= ny; };
    const float a = 3;
    if (T->sws == 1) {  // taper_grad.c:50-160
        const int tl = G2 - G1, tl2 = G4 - G3, ifw = G4 - G1;
        std::vector<float> win((size_t)std::max(ifw, 0) + 2, 0.0f);
        for (int i = 1; i <= ifw; ++i) {
            win[i] = 1.0f;
            if (i <= tl) win[i] = (float)::exp(-(1.0 / 2.0) * (a * (i - tl) / (tl / 2.0)) * (a * (i - tl) / (tl / 2.0)));
            if (i >= (ifw - tl2))
                win[i] = (float)::exp(-(1.0 / 2.0) * (a * (i - (ifw - tl2)) / (tl2 / 2.0)) *
                                      (a * (i - (ifw - tl2)) / (tl2 / 2.0)));
        }
        for (int j = 1; j <= ny; ++j) {
            int h = 1;
            for (int i = 1; i <= nx; ++i) M.tap[at(j, i)] = ((i > G1) && (i < G4)) ? win[h++] : 0.0f;
        }
        for (size_t s = 0; s < si.size(); ++s)
            if (inside(sj[s], si[s])) M.zero[at(sj[s], si[s])] = 1;
        for (size_t r = 0; r < ri.size(); ++r)
            if (inside(rj[r], ri[r])) M.zero[at(rj[r], ri[r])] = 1;
        return true;
    }
    if (T->sws == 2) {  // taper_grad.c:166-290
        const int tl = G2 - G1, ifw = G2 - G1 + 1;
        std::vector<float> win((size_t)std::max(ifw, 0) + 2, 0.0f);
        for (int i = 1; i <= ifw; ++i) {
            win[i] = 1.0f;
            if (i <= tl) win[i] = (float)::exp(-(1.0 / 2.0) * (a * (i - tl) / (tl / 2.0)) * (a * (i - tl) / (tl / 2.0)));
        }
        int h = 1;
        for (int j = 1; j <= ny; ++j) {
            float tap = 0.0f;
            if ((j >= G1) && (j <= G2)) tap = win[h];
            if (j > G2) tap = (float)::pow((double)(float)(j * DH), 1.0);
            for (int i = 1; i <= nx; ++i) M.tap[at(j, i)] = tap;
            if (j >= G1) h++;
        }
        return true;
    }
    if (T->sws == 3) {  // taper_grad.c:295-520
        const float R = T->srtradius;
        int ng = (int)std::ceil(2.0 * R / DH);
        if (ng < 5) ng = 5;
        if (ng > 4096) {
            err = "taper_grad: SRTRADIUS / DH too large";
            return false;
        }
        std::vector<float> m((size_t)ng * ng), msum(n, 1.0f);
        const float one = 1.0f, maxrad = (float)std::sqrt(2.0 * R * R);
        for (int j = 1; j <= ng; ++j)
            for (int i = 1; i <= ng; ++i) {
                const float x = (float)(((float)i - ((float)ng) / 2.0 - 0.5) * DH);
                const float y = (float)(((float)j - ((float)ng) / 2.0 - 0.5) * DH);
                const float rad = (float)std::sqrt((double)(x * x + y * y));
                float v = 0.0f;
                if (T->srtshape == 1) v = (float)std::erf((double)(one * rad / maxrad));
                else if (T->srtshape == 2) v = (rad > 0) ? (float)std::log((double)rad) : 0.0f;
                m[(size_t)(j - 1) * ng + (i - 1)] = v;
            }
        const float minm = min_first(m);
        for (float &v : m) v -= minm;
        auto M1 = [&](int j, int i) { return m[(size_t)(j - 1) * ng + (i - 1)]; };
        const float edge[4] = {M1(1, ng / 2), M1(ng / 2, 1), M1(ng / 2, ng), M1(ng, ng / 2)};
        const float maxm = max_first(edge, 4);
        for (float &v : m) {
            v /= maxm;
            if (v > 1.0f) v = 1.0f;
        }
        const int ijc = (int)std::ceil((float)ng / 2);
        for (size_t s = 0; s < si.size(); ++s)
            for (int iy = 1; iy <= ng; ++iy)
                for (int ix = 1; ix <= ng; ++ix) {
                    const int xx = si[s] + ix - ijc, yy = sj[s] + iy - ijc;
                    if (!inside(yy, xx)) continue;
                    float &p = msum[at(yy, xx)];
                    if (p > M1(iy, ix)) p = M1(iy, ix);
                }
        const float mn = min_first(msum);
        for (float &v : msum) v -= mn;
        const float mx = max_first(msum.data(), n);
        for (size_t t = 0; t < n; ++t) M.tap[t] = msum[t] / mx;
        const int F = T->filtsize;
        for (size_t s = 0; s < si.size(); ++s)
            for (int i = si[s] - F; i <= si[s] + F; ++i)
                for (int j = sj[s] - F; j <= sj[s] + F; ++j)
                    if (inside(j, i)) {
                        M.zero[at(j, i)] = 1;
                        M.tap[at(j, i)] = 0.0f;
                    }
        return true;
    }
    if (T->sws == 5 || T->sws == 6) {  // taper_grad.c:525-580 (taper.bin as a host map)
        if (!T->map) {
            err = "taper_grad: sws 5/6 need the taper map";
            return false;
        }
        std::copy(T->map, T->map + n, M.tap.begin());
        return true;
    }
    err = "taper_grad: sws must be 1, 2, 3, 5 or 6";
    return false;
}

hipError_t run_taper(const TaperMap &M, float *g, size_t n, float *d_tap, unsigned char *d_zero, hipStream_t st)
{
    hipError_t e = hipMemcpyAsync(d_tap, M.tap.data(), n * sizeof(float), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    if ((e = hipMemcpyAsync(d_zero, M.zero.data(), n, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_taper_apply, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, g, d_tap, d_zero);
    return hipGetLastError();
}

// smooth2 on the device plane g [ny][nx] in place.  The factored matrices go
// to the device on the first call with an empty host vector (ex, dx: nx
// each; ez, dz: ny - 1 each); later calls with the same vector reuse them.
// host must stay alive until the stream is synchronised.
hipError_t run_smooth2(int nx, int ny, const Smooth2Plan &P, std::vector<float> &host, float **buf, size_t &cap,
                       float *g, hipStream_t st)
{
    hipError_t e;
    if (host.empty()) {
        host.insert(host.end(), P.ex.begin(), P.ex.end());
        host.insert(host.end(), P.dx.begin(), P.dx.end());
        host.insert(host.end(), P.ez.begin(), P.ez.end());
        host.insert(host.end(), P.dz.begin(), P.dz.end());
        if ((e = grow(buf, cap, host.size())) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(*buf, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice, st)) !=
            hipSuccess)
            return e;
    }
    const float *ex = *buf, *dx = ex + nx, *ez = dx + nx, *dz = ez + (ny - 1);
    hipLaunchKernelGGL(k_smooth2_x, dim3((unsigned)((ny - 1 + 63) / 64)), dim3(64), 0, st, nx, ny, g, ex, dx, 0.0f);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_smooth2_z, dim3((unsigned)((nx - 1 + 63) / 64)), dim3(64), 0, st, nx, ny, g, ez, dz, P.r1,
                       0.0f);
    return hipGetLastError();
}

// per-process device state of the context-free entry points
struct Standalone {
    hipStream_t st = nullptr;
    GaussDev gd;
    float *g = nullptr, *tap = nullptr, *s2 = nullptr;
    unsigned char *zero = nullptr;
    size_t cap_g = 0, cap_tap = 0, cap_zero = 0, cap_s2 = 0;
};
Standalone &standalone()
{
    static Standalone s;
    return s;
}

}  // namespace

// ---------------------------------------------------------------------------
// handle API
// ---------------------------------------------------------------------------
struct dn_gradproc_state {
    GaussDev gd;
    float *tap = nullptr, *s2 = nullptr;
    unsigned char *zero = nullptr;
    size_t cap_tap = 0, cap_zero = 0, cap_s2 = 0;
    float ms_smooth = 0.0f;
};

void dn_gradproc_destroy(dn_ctx *c)
{
    if (!c || !c->gradproc) return;
    c->gradproc->gd.release();
    (void)hipFree(c->gradproc->tap);
    (void)hipFree(c->gradproc->s2);
    (void)hipFree(c->gradproc->zero);
    delete c->gradproc;
    c->gradproc = nullptr;
}

static dn_gradproc_state *gp(dn_ctx *c)
{
    if (!c->gradproc) c->gradproc = new dn_gradproc_state();
    return c->gradproc;
}

static int vs_host(dn_ctx *c, std::vector<float> &vel)
{
    const Geo &g = c->g;
    vel.resize((size_t)g.nx * g.ny);
    HIPCHK(c, hipMemcpy2DAsync(vel.data(), g.nx * sizeof(float), c->d_vs + gidx(g, 1, 1), g.pitch * sizeof(float),
                               g.nx * sizeof(float), g.ny, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dn_smooth_grad_device(dn_ctx *c, int spatfilter, float wd_damp, float wd_damp1, float fc_end, float *g_vs,
                          float *g_rho)
{
    if (spatfilter == 0) return 0;
    if (spatfilter < 2 || spatfilter > 4)
        return dn_fail(c,
                       "smooth_grad: SPATFILTER %d is not available (0 none, 2 smooth2, 3 gauss_filt, "
                       "4 gauss_filt_var)",
                       spatfilter);
    const Geo &g = c->g;
    if (spatfilter == 2) {
        Smooth2Plan S2;
        std::string err2;
        if (!make_smooth2_plan(g.nx, g.ny, wd_damp, S2, err2)) return dn_fail(c, "%s", err2.c_str());
        dn_gradproc_state *s = gp(c);
        std::vector<float> host;
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        if (g_vs) HIPCHK(c, run_smooth2(g.nx, g.ny, S2, host, &s->s2, s->cap_s2, g_vs, c->stream));
        if (g_rho) HIPCHK(c, run_smooth2(g.nx, g.ny, S2, host, &s->s2, s->cap_s2, g_rho, c->stream));
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipEventElapsedTime(&s->ms_smooth, c->ev0, c->ev1));
        return 0;
    }
    if (!(fc_end > 0.0f)) return dn_fail(c, "smooth_grad: FC_END must be > 0");
    GaussPlan P;
    std::string err;
    if (spatfilter == 3) {
        // C_vs: the vs mean the gradient normalisation uses (denise.c:746-819)
        if (!(c->C_vs > 0.0f)) return dn_fail(c, "smooth_grad: SPATFILTER 3 needs C_vs (run a gradient first)");
        const int rc = make_gauss_fixed_plan(g.nx, g.ny, c->p.dh, wd_damp, fc_end, c->C_vs, P, err);
        if (rc < 0) return dn_fail(c, "%s", err.c_str());
        if (rc > 0) return 0;
    } else {
        std::vector<float> vel;
        if (vs_host(c, vel)) return -1;
        if (!make_gauss_plan(g.nx, g.ny, c->p.dh, wd_damp, wd_damp1, fc_end, vel.data(), P, err))
            return dn_fail(c, "%s", err.c_str());
    }
    dn_gradproc_state *s = gp(c);
    const size_t n = (size_t)g.nx * g.ny;
    HIPCHK(c, upload_plan(P, s->gd, n, c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (g_vs) HIPCHK(c, run_gauss(g.nx, g.ny, P, s->gd, g_vs, c->stream));
    if (g_rho) HIPCHK(c, run_gauss(g.nx, g.ny, P, s->gd, g_rho, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipEventElapsedTime(&s->ms_smooth, c->ev0, c->ev1));
    return 0;
}

int dn_taper_grad_device(dn_ctx *c, const dn_taper *T, float *g_vs, float *g_rho)
{
    if (!T) return dn_fail(c, "dn_taper_grad: taper is NULL");
    const Geo &g = c->g;
    dn_taper TT = *T;
    TT.dh = c->p.dh;
    TaperMap M;
    std::string err;
    if (!make_taper(g.nx, g.ny, &TT, c->src_i, c->src_j, c->rec_i, c->rec_j, M, err))
        return dn_fail(c, "%s", err.c_str());
    dn_gradproc_state *s = gp(c);
    const size_t n = (size_t)g.nx * g.ny;
    HIPCHK(c, grow(&s->tap, s->cap_tap, n));
    HIPCHK(c, grow(&s->zero, s->cap_zero, n));
    if (g_vs) HIPCHK(c, run_taper(M, g_vs, n, s->tap, s->zero, c->stream));
    if (g_rho) HIPCHK(c, run_taper(M, g_rho, n, s->tap, s->zero, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int dn_smooth_grad(dn_ctx *c, int spatfilter, float wd_damp, float wd_damp1, float fc_end, float *ms)
{
    if (!c) return dn_fail(nullptr, "dn_smooth_grad: ctx is NULL");
    if (dn_smooth_grad_device(c, spatfilter, wd_damp, wd_damp1, fc_end, c->d_gvs_out, c->d_grho_out)) return -1;
    if (ms) *ms = c->gradproc ? c->gradproc->ms_smooth : 0.0f;
    return 0;
}

extern "C" int dn_taper_grad(dn_ctx *c, int which, const dn_taper *T)
{
    if (!c) return dn_fail(nullptr, "dn_taper_grad: ctx is NULL");
    if (which < 1 || which > 3) return dn_fail(c, "dn_taper_grad: which must be 1 (vs), 2 (rho) or 3 (both)");
    return dn_taper_grad_device(c, T, (which & 1) ? c->d_gvs_out : nullptr, (which & 2) ? c->d_grho_out : nullptr);
}

// ---------------------------------------------------------------------------
// context-free entry points on host arrays [ny][nx]
// ---------------------------------------------------------------------------
static int sa_init(Standalone &S)
{
    if (S.st) return 0;
    if (hipStreamCreateWithFlags(&S.st, hipStreamNonBlocking) != hipSuccess) {
        S.st = nullptr;
        return dn_fail(nullptr, "no HIP device");
    }
    return 0;
}

#define SACHK(x)                                                                              \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess) return dn_fail(nullptr, "%s: %s", #x, hipGetErrorString(e_));   \
    } while (0)

extern "C" int dn_gauss_filt_var(int nx, int ny, float dh, float wd_damp, float wd_damp1, float fc_end,
                                 const float *vel, float *grad)
{
    if (nx < 1 || ny < 1 || !vel || !grad) return dn_fail(nullptr, "dn_gauss_filt_var: bad arguments");
    if (!(fc_end > 0.0f)) return dn_fail(nullptr, "dn_gauss_filt_var: FC_END must be > 0");
    Standalone &S = standalone();
    if (sa_init(S)) return -1;
    GaussPlan P;
    std::string err;
    if (!make_gauss_plan(nx, ny, dh, wd_damp, wd_damp1, fc_end, vel, P, err))
        return dn_fail(nullptr, "%s", err.c_str());
    const size_t n = (size_t)nx * ny;
    SACHK(grow(&S.g, S.cap_g, n));
    SACHK(upload_plan(P, S.gd, n, S.st));
    SACHK(hipMemcpyAsync(S.g, grad, n * sizeof(float), hipMemcpyHostToDevice, S.st));
    SACHK(run_gauss(nx, ny, P, S.gd, S.g, S.st));
    SACHK(hipMemcpyAsync(grad, S.g, n * sizeof(float), hipMemcpyDeviceToHost, S.st));
    SACHK(hipStreamSynchronize(S.st));
    return 0;
}

extern "C" int dn_smooth2(int nx, int ny, float wd_damp, float *grad)
{
    if (nx < 1 || ny < 1 || !grad) return dn_fail(nullptr, "dn_smooth2: bad arguments");
    Smooth2Plan P;
    std::string err;
    if (!make_smooth2_plan(nx, ny, wd_damp, P, err)) return dn_fail(nullptr, "%s", err.c_str());
    Standalone &S = standalone();
    if (sa_init(S)) return -1;
    const size_t n = (size_t)nx * ny;
    std::vector<float> host;
    SACHK(grow(&S.g, S.cap_g, n));
    SACHK(hipMemcpyAsync(S.g, grad, n * sizeof(float), hipMemcpyHostToDevice, S.st));
    SACHK(run_smooth2(nx, ny, P, host, &S.s2, S.cap_s2, S.g, S.st));
    SACHK(hipMemcpyAsync(grad, S.g, n * sizeof(float), hipMemcpyDeviceToHost, S.st));
    SACHK(hipStreamSynchronize(S.st));
    return 0;
}

extern "C" int dn_gauss_filt(int nx, int ny, float dh, float wd_damp, float fc_end, float c_vs, float *grad)
{
    if (nx < 1 || ny < 1 || !grad) return dn_fail(nullptr, "dn_gauss_filt: bad arguments");
    if (!(fc_end > 0.0f) || !(c_vs > 0.0f)) return dn_fail(nullptr, "dn_gauss_filt: FC_END and C_vs must be > 0");
    Standalone &S = standalone();
    if (sa_init(S)) return -1;
    GaussPlan P;
    std::string err;
    const int rc = make_gauss_fixed_plan(nx, ny, dh, wd_damp, fc_end, c_vs, P, err);
    if (rc < 0) return dn_fail(nullptr, "%s", err.c_str());
    if (rc > 0) return 0;
    const size_t n = (size_t)nx * ny;
    SACHK(grow(&S.g, S.cap_g, n));
    SACHK(upload_plan(P, S.gd, n, S.st));
    SACHK(hipMemcpyAsync(S.g, grad, n * sizeof(float), hipMemcpyHostToDevice, S.st));
    SACHK(run_gauss(nx, ny, P, S.gd, S.g, S.st));
    SACHK(hipMemcpyAsync(grad, S.g, n * sizeof(float), hipMemcpyDeviceToHost, S.st));
    SACHK(hipStreamSynchronize(S.st));
    return 0;
}

extern "C" int dn_taper_grad_host(int nx, int ny, const dn_taper *T, int nsrc, const int *src_i, const int *src_j,
                                  int nrec, const int *rec_i, const int *rec_j, float *grad, float *taper_out)
{
    if (nx < 1 || ny < 1 || !T || !grad || nsrc < 0 || nrec < 0)
        return dn_fail(nullptr, "dn_taper_grad_host: bad arguments");
    Standalone &S = standalone();
    if (sa_init(S)) return -1;
    std::vector<int> si(src_i, src_i + nsrc), sj(src_j, src_j + nsrc), ri(rec_i, rec_i + nrec),
        rj(rec_j, rec_j + nrec);
    TaperMap M;
    std::string err;
    if (!make_taper(nx, ny, T, si, sj, ri, rj, M, err)) return dn_fail(nullptr, "%s", err.c_str());
    const size_t n = (size_t)nx * ny;
    SACHK(grow(&S.g, S.cap_g, n));
    SACHK(grow(&S.tap, S.cap_tap, n));
    SACHK(grow(&S.zero, S.cap_zero, n));
    SACHK(hipMemcpyAsync(S.g, grad, n * sizeof(float), hipMemcpyHostToDevice, S.st));
    SACHK(run_taper(M, S.g, n, S.tap, S.zero, S.st));
    SACHK(hipMemcpyAsync(grad, S.g, n * sizeof(float), hipMemcpyDeviceToHost, S.st));
    SACHK(hipStreamSynchronize(S.st));
    if (taper_out) std::copy(M.tap.begin(), M.tap.end(), taper_out);
    return 0;
}
