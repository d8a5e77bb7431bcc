// Reference-named drop-ins (layer 1 of include/denise_sh.h).
//
// The reference passes Numerical-Recipes offset matrices (util.c:177-203):
// m[j][i] with row pointers, rows contiguous in i.  Two modes:
//
//  * copy (default): each call gathers the rows it reads into the device
//    layout, runs the split HIP kernel and scatters what it writes back into
//    the caller's arrays, so the caller sees the reference's in-place
//    semantics after every call.
//  * resident (dn_dropin_resident): every array a drop-in receives is copied
//    to HBM the first time its host pointer is seen and stays there; later
//    calls run on the device copies without host transfers.  The host
//    copies go stale until dn_dropin_sync(1).  seismo_ssg reads the
//    receiver samples (and the cell the reference's NaN probe reads) back
//    from the device, or, deferred, keeps the seismograms on the device
//    until dn_dropin_sync(1).  zero_fdveps / zero_fdveps_visc zero both
//    copies.
//
// The caller's hc and CPML vectors are used as given (resident: as first
// seen, re-read by dn_dropin_sync(0)).
#include "dn_core.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <unordered_map>
#include <vector>

using namespace dn;

namespace {

// roles of the host arrays: how much of them the device copy mirrors
enum Kind { kField, kInterior, kTensor, kPsi, kVec };

struct Resident {
    float *dev = nullptr;
    Kind kind = kInterior;
    int rows = 0, cols = 0;   // kPsi: [1..rows][1..cols]; kVec: [1..cols]
    bool written = false;     // the kernels write it (copied back by dn_dropin_sync(1))
};

// deferred seismograms of one sectionvy array: device [ntr][NT]
struct Section {
    float *dev = nullptr;
    int ntr = 0, nt = 0;
    std::vector<char> recorded;   // samples written on the device since the last sync
};

// one source or receiver point of a device injection: padded index, value
struct InjectPoint {
    long long o;
    float v;
};
constexpr int kInjectBatch = 64;
struct InjectArgs {
    int n;
    InjectPoint p[kInjectBatch];
};

__global__ void k_inject_points(float *vy, InjectArgs a)
{
    // one lane, in list order: points listed twice add up as in the
    // reference's loop
    if (threadIdx.x == 0 && blockIdx.x == 0)
        for (int k = 0; k < a.n; ++k) vy[a.p[k].o] += a.p[k].v;
}

__global__ void k_gather_points(const float *vy, const long long *o, int n, float *out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = vy[o[t]];
}

__global__ void k_gather_section(const float *vy, const long long *o, int n, float *sec, int nt, int lsamp)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) sec[(long long)t * nt + (lsamp - 1)] = vy[o[t]];
}

struct DropIn {
    bool have_globals = false;
    dn_globals G{};
    Geo g{};
    hipStream_t stream = nullptr;
    size_t cap_plane = 0;
    int cap_L = 0, cap_fw = 0, cap_nx = 0, cap_ny = 0;
    // copy mode: device fields
    float *vy = nullptr, *vyp1 = nullptr, *sxy = nullptr, *syz = nullptr, *uy = nullptr, *uyx = nullptr;
    float *rho = nullptr, *uip = nullptr, *ujp = nullptr, *fipjp = nullptr, *fcoef = nullptr;
    float *dip = nullptr, *dcoef = nullptr, *r = nullptr, *q = nullptr;
    float *psi_a = nullptr, *psi_b = nullptr, *pml = nullptr;
    std::vector<float> h;  // host staging
    // resident mode
    int mode = 0;                                   // 0 copy, 1 resident, 2 resident + deferred seismograms
    std::unordered_map<const void *, Resident> reg;
    std::unordered_map<const void *, Section> sections;
    std::vector<int> rec_host;                      // receiver (i, j) pairs of the gather list
    long long *d_rec = nullptr;                     // padded indices: receivers, then the probe cell
    int rec_cap = 0, rec_n = 0;
    float *d_gather = nullptr, *h_gather = nullptr; // pinned staging of the gathered samples
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

void release_resident(DropIn &d)
{
    for (auto &kv : d.reg) (void)hipFree(kv.second.dev);
    d.reg.clear();
    for (auto &kv : d.sections) (void)hipFree(kv.second.dev);
    d.sections.clear();
    if (d.d_rec) (void)hipFree(d.d_rec);
    if (d.d_gather) (void)hipFree(d.d_gather);
    if (d.h_gather) (void)hipHostFree(d.h_gather);
    d.d_rec = nullptr;
    d.d_gather = d.h_gather = nullptr;
    d.rec_cap = d.rec_n = 0;
    d.rec_host.clear();
}

bool ensure_stream(DropIn &d)
{
    if (d.stream == nullptr && hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking) != hipSuccess)
        return fail("no HIP device");
    return true;
}

bool ensure(DropIn &d)
{
    if (!d.have_globals) return fail("dn_set_globals() must be called before the drop-ins");
    const dn_globals &G = d.G;
    if (!ensure_stream(d)) return false;
    Geo &g = d.g;
    g = make_geo(G.NX, G.NY, G.FW, G.FREE_SURF, G.L, G.BOUNDARY);
    if (d.mode) return true;   // resident arrays are allocated as they are seen
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
    d.h.resize(std::max(d.h.size(), (size_t)g.plane));
    std::fill(d.h.begin(), d.h.begin() + g.plane, 0.0f);
    for (int j = jlo; j <= jhi; ++j) std::memcpy(&d.h[gidx(g, j, ilo)], &m[j][ilo], (size_t)(ihi - ilo + 1) * 4);
    return hipMemcpyAsync(dev, d.h.data(), (size_t)g.plane * 4, hipMemcpyHostToDevice, d.stream) == hipSuccess &&
           hipStreamSynchronize(d.stream) == hipSuccess;
}
bool get(DropIn &d, const float *dev, float **m, int jlo, int jhi, int ilo, int ihi)
{
    const Geo &g = d.g;
    d.h.resize(std::max(d.h.size(), (size_t)g.plane));
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
    d.h.resize(std::max(d.h.size(), (size_t)g.plane * L));
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
    d.h.resize(std::max(d.h.size(), (size_t)g.plane * L));
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

// ---- resident mode ------------------------------------------------------
size_t res_floats(const DropIn &d, const Resident &r)
{
    switch (r.kind) {
    case kTensor: return (size_t)d.g.plane * std::max(1, d.G.L);
    case kPsi: return (size_t)std::max(1, r.rows) * std::max(1, r.cols);
    case kVec: return (size_t)r.cols + 1;
    default: return (size_t)d.g.plane;
    }
}

// host -> device copy of a resident array over its role's extent: fields
// with the K-wide halo the stencils read, everything else its interior
bool res_upload(DropIn &d, const void *host, const Resident &r)
{
    const int K = d.G.FDORDER / 2, NX = d.G.NX, NY = d.G.NY;
    switch (r.kind) {
    case kField: return put(d, r.dev, (float **)host, 1 - K, NY + K, 1 - K, NX + K);
    case kInterior: return put(d, r.dev, (float **)host, 1, NY, 1, NX);
    case kTensor: return put3(d, r.dev, (float ***)host, d.G.L);
    case kPsi: return r.rows < 1 || r.cols < 1 || put_psi(d, r.dev, (float **)host, r.rows, r.cols);
    case kVec: {
        std::vector<float> t((size_t)r.cols + 1, 0.0f);
        std::memcpy(&t[1], (const float *)host + 1, (size_t)r.cols * 4);
        return hipMemcpy(r.dev, t.data(), t.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
    }
    }
    return false;
}

// device -> host copy of what the kernels write (interiors, psi, tensors)
bool res_download(DropIn &d, const void *host, const Resident &r)
{
    const int NX = d.G.NX, NY = d.G.NY;
    switch (r.kind) {
    case kField:
    case kInterior: return get(d, r.dev, (float **)host, 1, NY, 1, NX);
    case kTensor: return get3(d, r.dev, (float ***)host, d.G.L);
    case kPsi: return r.rows < 1 || r.cols < 1 || get_psi(d, r.dev, (float **)host, r.rows, r.cols);
    case kVec: return true;
    }
    return false;
}

// device copy of a host array, made at its first sight
float *bind(DropIn &d, const void *host, Kind kind, bool written, int rows = 0, int cols = 0)
{
    if (!host) return nullptr;
    auto it = d.reg.find(host);
    if (it != d.reg.end()) {
        it->second.written = it->second.written || written;
        return it->second.dev;
    }
    Resident r;
    r.kind = kind;
    r.rows = rows;
    r.cols = cols;
    r.written = written;
    const size_t n = res_floats(d, r);
    if (hipMalloc((void **)&r.dev, n * sizeof(float)) != hipSuccess) {
        fail("out of device memory (resident drop-in arrays)");
        return nullptr;
    }
    // the zero fill runs on the library's stream; the psi and vector uploads
    // are plain copies, so wait for it first
    if (launch_fill(r.dev, n, 0.0f, d.stream) != hipSuccess || hipStreamSynchronize(d.stream) != hipSuccess ||
        !res_upload(d, host, r)) {
        (void)hipFree(r.dev);
        fail("upload failed");
        return nullptr;
    }
    d.reg[host] = r;
    return r.dev;
}

bool bind_pml(DropIn &d, PmlVec &pv, float *const vecs[12])
{
    const float **slots[12] = {&pv.Kx, &pv.ax, &pv.bx, &pv.Kxh, &pv.axh, &pv.bxh,
                               &pv.Ky, &pv.ay, &pv.by, &pv.Kyh, &pv.ayh, &pv.byh};
    const int n2 = 2 * d.G.FW;
    for (int v = 0; v < 12; ++v) {
        if (n2 < 1 || !vecs[v]) {
            *slots[v] = nullptr;
            continue;
        }
        float *p = bind(d, vecs[v], kVec, false, 0, n2);
        if (!p) return false;
        *slots[v] = p;
    }
    return true;
}

// point additions on the device (update_v_PML.c:563-583 after the sweep)
bool inject(DropIn &d, float *vy_dev, const std::vector<InjectPoint> &pts)
{
    for (size_t k0 = 0; k0 < pts.size(); k0 += kInjectBatch) {
        InjectArgs a{};
        a.n = (int)std::min<size_t>(kInjectBatch, pts.size() - k0);
        for (int k = 0; k < a.n; ++k) a.p[k] = pts[k0 + k];
        hipLaunchKernelGGL(k_inject_points, dim3(1), dim3(64), 0, d.stream, vy_dev, a);
        if (hipGetLastError() != hipSuccess) return fail("source injection failed");
    }
    return true;
}

bool check_range(const DropIn &d, int nx1, int nx2, int ny1, int ny2)
{
    if (nx1 != 1 || ny1 != 1 || nx2 != d.G.NX || ny2 != d.G.NY)
        return fail("drop-ins update the whole single domain: call with 1, NX, 1, NY");
    return true;
}

bool sync_to_host(DropIn &d)
{
    if (!d.stream) return true;
    if (hipStreamSynchronize(d.stream) != hipSuccess) return fail("device work failed");
    for (auto &kv : d.reg)
        if (kv.second.written && !res_download(d, kv.first, kv.second)) return fail("download failed");
    for (auto &kv : d.sections) {
        Section &s = kv.second;
        std::vector<float> t((size_t)s.ntr * s.nt);
        if (hipMemcpy(t.data(), s.dev, t.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
            return fail("download failed");
        float **sec = (float **)kv.first;
        for (int l = 1; l <= s.nt; ++l)
            if (s.recorded[l])
                for (int r = 1; r <= s.ntr; ++r) sec[r][l] = t[(size_t)(r - 1) * s.nt + (l - 1)];
        std::fill(s.recorded.begin(), s.recorded.end(), 0);
    }
    return true;
}

}  // namespace

extern "C" int dn_set_globals(const dn_globals *G)
{
    if (!G) return dn_fail(nullptr, "dn_set_globals: NULL");
    if (G->FDORDER < 2 || G->FDORDER > 12 || (G->FDORDER % 2)) return dn_fail(nullptr, "wrong FDORDER %d", G->FDORDER);
    if (G->BOUNDARY != 0 && G->BOUNDARY != 1) return dn_fail(nullptr, "BOUNDARY must be 0 or 1");
    if (G->BOUNDARY == 1 && G->NX < G->FDORDER)
        return dn_fail(nullptr, "BOUNDARY=1 needs NX >= FDORDER (got NX=%d)", G->NX);
    if (G->GRAD_FORM != 2) return dn_fail(nullptr, "only GRAD_FORM=2 is provided");
    if (G->L < 0 || G->L > kMaxL) return dn_fail(nullptr, "L=%d outside 0..%d", G->L, kMaxL);
    DropIn &d = state();
    if (d.mode && d.have_globals &&
        (G->NX != d.G.NX || G->NY != d.G.NY || G->L != d.G.L || G->FW != d.G.FW || G->FDORDER != d.G.FDORDER ||
         G->BOUNDARY != d.G.BOUNDARY))
        return dn_fail(nullptr, "dn_set_globals: grid changed while drop-in arrays are resident; "
                                "call dn_dropin_resident(0) first");
    d.G = *G;
    d.have_globals = true;
    return 0;
}

extern "C" int dn_dropin_resident(int mode)
{
    if (mode < 0 || mode > 2) return dn_fail(nullptr, "dn_dropin_resident: mode 0, 1 or 2");
    DropIn &d = state();
    if (mode == d.mode) return 0;
    if (d.mode && mode == 0) {
        // leaving residency: the host arrays become current again
        if (!sync_to_host(d)) return -1;
        release_resident(d);
    }
    if (mode) {
        if (!d.have_globals) return dn_fail(nullptr, "dn_dropin_resident: call dn_set_globals() first");
        if (!ensure_stream(d)) return -1;
        d.g = make_geo(d.G.NX, d.G.NY, d.G.FW, d.G.FREE_SURF, d.G.L, d.G.BOUNDARY);
        if (d.mode == 2 && mode == 1 && !sync_to_host(d)) return -1;   // deliver deferred seismograms
    }
    d.mode = mode;
    return 0;
}

extern "C" int dn_dropin_sync(int to_host)
{
    DropIn &d = state();
    if (!d.mode) return 0;
    if (to_host) return sync_to_host(d) ? 0 : -1;
    if (hipStreamSynchronize(d.stream) != hipSuccess) return dn_fail(nullptr, "device work failed");
    for (auto &kv : d.reg)
        if (!res_upload(d, kv.first, kv.second)) return dn_fail(nullptr, "upload failed");
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
    // the points added after the sweep (update_v_PML.c:563-583)
    std::vector<InjectPoint> pts;
    for (int l = 1; l <= nsrc; ++l) {
        const int i = (int)srcpos_loc[1][l], j = (int)srcpos_loc[2][l];
        const bool add = (sw == 0) ? ((int)srcpos_loc[8][l] == 3) : (sw == 1 && G.QUELLTYPB == 2);
        if (add) pts.push_back({gidx(d.g, j, i), signals1[l][nt]});
    }
    if (d.mode) {
        bool ok = bind_pml(d, a.pml, vecs);
        a.vy = bind(d, vy, kField, true);
        a.vyp1 = bind(d, vyp1, kInterior, true);
        a.sxy = bind(d, sxy, kField, true);
        a.syz = bind(d, syz, kField, true);
        a.rho = bind(d, rho, kInterior, false);
        if (G.FW > 0) {
            a.psi_a = bind(d, psi_syy_y, kPsi, true, 2 * G.FW, G.NX);
            a.psi_b = bind(d, psi_sxy_x, kPsi, true, G.NY, 2 * G.FW);
            ok = ok && a.psi_a && a.psi_b;
        }
        if (!ok || !a.vy || !a.vyp1 || !a.sxy || !a.syz || !a.rho) return;
        // a few points ride in the sweep's launch, more take their own
        const bool fused = pts.size() <= (size_t)kSplitPoints;
        a.npts = fused ? (int)pts.size() : 0;
        for (int k = 0; k < a.npts; ++k) {
            a.pt_o[k] = pts[k].o;
            a.pt_v[k] = pts[k].v;
        }
        if (launch_split_v(a, d.stream) != hipSuccess) {
            fail("update_v_PML kernel failed");
            return;
        }
        if (!fused) inject(d, a.vy, pts);
        return;
    }
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
    for (int l = 1, k = 0; l <= nsrc; ++l) {
        const int i = (int)srcpos_loc[1][l], j = (int)srcpos_loc[2][l];
        const bool add = (sw == 0) ? ((int)srcpos_loc[8][l] == 3) : (sw == 1 && G.QUELLTYPB == 2);
        if (add) vy[j][i] += pts[k++].v;
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
    if (visco)
        for (int l = 0; l < G.L; ++l) {
            a.bip[l] = bip[l + 1];
            a.bjm[l] = bjm[l + 1];
            a.cip[l] = cip[l + 1];
            a.cjm[l] = cjm[l + 1];
        }
    if (d.mode) {
        bool ok = bind_pml(d, a.pml, vecs);
        a.vy = bind(d, vy, kField, true);
        a.sxy = bind(d, sxy, kField, true);
        a.syz = bind(d, syz, kField, true);
        a.uy = bind(d, uy, kInterior, true);
        a.uyx = bind(d, uyx, kInterior, true);
        ok = ok && a.vy && a.sxy && a.syz && a.uy && a.uyx;
        if (!visco) {
            a.uip = bind(d, uip, kInterior, false);
            a.ujp = bind(d, ujp, kInterior, false);
            ok = ok && a.uip && a.ujp;
        } else {
            a.fipjp = bind(d, fipjp, kInterior, false);
            a.fcoef = bind(d, f, kInterior, false);
            a.r = bind(d, r, kTensor, true);
            a.q = bind(d, q, kTensor, true);
            a.dip = bind(d, dip, kTensor, false);
            a.dcoef = bind(d, d3, kTensor, false);
            ok = ok && a.fipjp && a.fcoef && a.r && a.q && a.dip && a.dcoef;
        }
        if (G.FW > 0) {
            a.psi_a = bind(d, psi_vyy, kPsi, true, 2 * G.FW, G.NX);
            a.psi_b = bind(d, psi_vyx, kPsi, true, G.NY, 2 * G.FW);
            ok = ok && a.psi_a && a.psi_b;
        }
        if (ok && launch_split_s(a, visco, d.stream) != hipSuccess) fail("stress kernel failed");
        return;
    }
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
    DropIn &d = state();
    auto it = d.mode ? d.reg.find(vy) : d.reg.end();
    if (it == d.reg.end()) {
        for (int itr = 1; itr <= ntr; ++itr) sectionvy[itr][lsamp] = vy[recpos[2][itr]][recpos[1][itr]];
        return;
    }
    const float *vdev = it->second.dev;
    const dn_globals &G = d.G;
    // gather list: the receivers, then the centre cell denise.c:1183 probes
    std::vector<int> want(2 * (size_t)ntr + 2);
    for (int itr = 1; itr <= ntr; ++itr) {
        want[2 * (itr - 1)] = recpos[1][itr];
        want[2 * (itr - 1) + 1] = recpos[2][itr];
    }
    want[2 * ntr] = G.NX / 2;
    want[2 * ntr + 1] = G.NY / 2;
    const int n = ntr + 1;
    if (want != d.rec_host) {
        if (n > d.rec_cap) {
            if (d.d_rec) (void)hipFree(d.d_rec);
            if (d.d_gather) (void)hipFree(d.d_gather);
            if (d.h_gather) (void)hipHostFree(d.h_gather);
            d.d_rec = nullptr;
            d.d_gather = d.h_gather = nullptr;
            d.rec_cap = 0;
            if (hipMalloc((void **)&d.d_rec, n * sizeof(long long)) != hipSuccess ||
                hipMalloc((void **)&d.d_gather, n * sizeof(float)) != hipSuccess ||
                hipHostMalloc((void **)&d.h_gather, n * sizeof(float), hipHostMallocDefault) != hipSuccess) {
                fail("out of memory (receiver gather)");
                return;
            }
            d.rec_cap = n;
        }
        std::vector<long long> o(n);
        for (int k = 0; k < n; ++k) o[k] = gidx(d.g, want[2 * k + 1], want[2 * k]);
        if (hipMemcpyAsync(d.d_rec, o.data(), n * sizeof(long long), hipMemcpyHostToDevice, d.stream) != hipSuccess ||
            hipStreamSynchronize(d.stream) != hipSuccess) {
            fail("upload failed");
            return;
        }
        d.rec_host = want;
    }
    d.rec_n = n;
    if (d.mode == 2 && G.NT >= 1 && lsamp >= 1 && lsamp <= G.NT) {
        // deferred: the samples stay on the device until dn_dropin_sync(1)
        Section &s = d.sections[sectionvy];
        if (!s.dev || s.ntr != ntr || s.nt != G.NT) {
            if (s.dev) (void)hipFree(s.dev);
            s = Section{};
            if (hipMalloc((void **)&s.dev, (size_t)ntr * G.NT * sizeof(float)) != hipSuccess) {
                fail("out of device memory (deferred seismograms)");
                return;
            }
            s.ntr = ntr;
            s.nt = G.NT;
            s.recorded.assign((size_t)G.NT + 1, 0);
        }
        hipLaunchKernelGGL(k_gather_section, dim3((ntr + 63) / 64), dim3(64), 0, d.stream, vdev, d.d_rec, ntr, s.dev,
                           s.nt, lsamp);
        if (hipGetLastError() != hipSuccess) fail("receiver gather failed");
        s.recorded[lsamp] = 1;
        return;
    }
    hipLaunchKernelGGL(k_gather_points, dim3((n + 63) / 64), dim3(64), 0, d.stream, vdev, d.d_rec, n, d.d_gather);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(d.h_gather, d.d_gather, n * sizeof(float), hipMemcpyDeviceToHost, d.stream) != hipSuccess ||
        hipStreamSynchronize(d.stream) != hipSuccess) {
        fail("receiver readback failed");
        return;
    }
    for (int itr = 1; itr <= ntr; ++itr) sectionvy[itr][lsamp] = d.h_gather[itr - 1];
    vy[G.NY / 2][G.NX / 2] = d.h_gather[ntr];
}

// zero_fdveps.c:11-52 / zero_fdveps_visc.c:11-64 on the host arrays, and on
// their device copies when resident (whose halo frames are zero anyway)
static void zero_resident(DropIn &d, const void *host)
{
    auto it = d.reg.find(host);
    if (it == d.reg.end()) return;
    if (launch_fill(it->second.dev, res_floats(d, it->second), 0.0f, d.stream) != hipSuccess)
        fail("device zeroing failed");
}

static void zero_common(int ny1, int ny2, int nx1, int nx2, float **vy, float **syy, float **sxy, float **vym1,
                        float **vyp1, float **psi_sxy_x, float **psi_vyx, float **psi_syy_y, float **psi_vyy,
                        float ***pr, float ***pp, float ***pq)
{
    DropIn &d = state();
    if (!d.have_globals) {
        fail("dn_set_globals() must be called before zero_fdveps");
        return;
    }
    const dn_globals &G = d.G;
    for (int j = ny1; j <= ny2; ++j)
        for (int i = nx1; i <= nx2; ++i) {
            vy[j][i] = 0.0f;
            syy[j][i] = 0.0f;
            sxy[j][i] = 0.0f;
            vym1[j][i] = 0.0f;
            vyp1[j][i] = 0.0f;
        }
    for (int j = 1; j <= G.NY; ++j)
        for (int i = 1; i <= 2 * G.FW; ++i) psi_sxy_x[j][i] = psi_vyx[j][i] = 0.0f;
    for (int j = 1; j <= 2 * G.FW; ++j)
        for (int i = 1; i <= G.NX; ++i) psi_syy_y[j][i] = psi_vyy[j][i] = 0.0f;
    if (pr)
        for (int j = ny1; j <= ny2; ++j)
            for (int i = nx1; i <= nx2; ++i)
                for (int l = 1; l <= G.L; ++l) pr[j][i][l] = pp[j][i][l] = pq[j][i][l] = 0.0f;
    if (d.mode) {
        const void *all[] = {vy, syy, sxy, vym1, vyp1, psi_sxy_x, psi_vyx, psi_syy_y, psi_vyy, pr, pp, pq};
        for (const void *p : all)
            if (p) zero_resident(d, p);
    }
}

extern "C" void zero_fdveps(int ny1, int ny2, int nx1, int nx2, float **vy, float **syy, float **sxy, float **vym1,
                            float **vyp1, float **psi_sxy_x, float **psi_vyx, float **psi_syy_y, float **psi_vyy)
{
    zero_common(ny1, ny2, nx1, nx2, vy, syy, sxy, vym1, vyp1, psi_sxy_x, psi_vyx, psi_syy_y, psi_vyy, nullptr,
                nullptr, nullptr);
}

extern "C" void zero_fdveps_visc(int ny1, int ny2, int nx1, int nx2, float **vy, float **syy, float **sxy,
                                 float **vym1, float **vyp1, float **psi_sxy_x, float **psi_vyx, float **psi_syy_y,
                                 float **psi_vyy, float ***pr, float ***pp, float ***pq)
{
    zero_common(ny1, ny2, nx1, nx2, vy, syy, sxy, vym1, vyp1, psi_sxy_x, psi_vyx, psi_syy_y, psi_vyy, pr, pp, pq);
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
