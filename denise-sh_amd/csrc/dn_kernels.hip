// Hand-written CDNA4 (gfx950) kernels of the SH time-stepping and adjoint
// gradient path.  Every floating-point expression keeps the reference's
// operand order and rounding (FP32 operators, double where the reference
// promotes through a double literal); the file is compiled with
// -ffp-contract=off so no multiply-add is fused, which makes the results
// reproduce the reference's CPU FP32 arithmetic.
#include "dn_internal.h"
#include <hip/hip_fp16.h>

#pragma clang fp contract(off)

namespace dn {

hipError_t launch_fill_u32(unsigned int *p, size_t n, unsigned int v, hipStream_t st);

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// 32-bit offset of (j, i) inside one padded plane (plane < 2^31 elements)
__device__ __forceinline__ unsigned gi32(const Geo &g, int j, int i)
{
    return (unsigned)((j - 1 + g.yoff) * g.pitch + (i - 1 + g.xoff));
}

// loads and stores at a 32-bit byte offset from a uniform base: the
// addresses become base (scalar registers) + offset (one vector register)
__device__ __forceinline__ float ldf(const float *base, unsigned boff)
{
    return *(const float *)((const char *)base + boff);
}
__device__ __forceinline__ void stf(float *base, unsigned boff, float v) { *(float *)((char *)base + boff) = v; }
__device__ __forceinline__ void st2(float *base, unsigned boff, float2 v) { *(float2 *)((char *)base + boff) = v; }
__device__ __forceinline__ void st4(float *base, unsigned boff, float4 v) { *(float4 *)((char *)base + boff) = v; }
__device__ __forceinline__ float2 ld2(const float *base, unsigned boff)
{
    return *(const float2 *)((const char *)base + boff);
}
__device__ __forceinline__ float4 ld4(const float *base, unsigned boff)
{
    return *(const float4 *)((const char *)base + boff);
}

__device__ __forceinline__ uint16_t f2bf16(float x)
{
    // round-to-nearest-even truncation of the fp32 pattern (NaN not expected)
    unsigned int u = __float_as_uint(x);
    unsigned int lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf162f(uint16_t h) { return __uint_as_float(((unsigned int)h) << 16); }
__device__ __forceinline__ uint16_t f2h(float x) { return __half_as_ushort(__float2half_rn(x)); }
__device__ __forceinline__ float h2f(uint16_t h) { return __half2float(__ushort_as_half(h)); }

// ---------------------------------------------------------------------------
// Fused time step.
//
// One 64-lane wavefront (= one workgroup) owns `ow = 65 - 2K` consecutive
// columns of one shot and walks down a segment of R rows.  Its lanes cover
// ow + 2K - 1 = 64 columns: the velocity update (update_v_PML.c) is computed
// for all 64 (the 2K-1 outer ones redundantly, as their owners in the
// neighbouring strips also do) so that the stress update
// (update_s_{elastic,visc}_PML.c) of the owned columns finds every x
// neighbour of the new vy inside the wave.  In y the wave keeps register
// windows: 2K rows of old syz (velocity y-derivative) and 2K rows of new vy
// (stress y-derivative); the velocity runs K rows ahead of the stress.  The
// x-neighbours come through LDS.  Old and new vy, sxy, syz and the velocity
// CPML memory variables are separate buffers (ping-pong), so the redundant
// reads of neighbouring strips never race with their owners' writes; the
// stress CPML and the relaxation memory variables are touched by their owner
// only and are updated in place.
//
// Also fused: source / residual injection (update_v_PML.c:563-583), receiver
// sampling (seismo_ssg.c), snapshot store and Ws (denise.c:1259-1327),
// correlation and Wr (denise.c:1580-1625).
//
// Block ids are remapped so that each of the 8 XCDs receives a contiguous
// range of (tile, shot) pairs with the shot index fastest: the shots of one
// tile share the static model fields through that XCD's L2.
// ---------------------------------------------------------------------------
template <int K, int L, bool ADJ, int EX, int NSH>
__global__ __launch_bounds__(kWave * kStripsPerGroup) void k_step(const StepArgs a)
{
    // EX: 0 for the plain steps (most steps), which carry none of the
    // snapshot code and its pointers; 1 stores FP32 snapshots or stages
    // FP16 ones / sums the pseudo-Hessian (forward) or correlates with FP32
    // snapshots (adjoint); 2 stores BF16 snapshots (forward) or correlates
    // with 16-bit snapshots, decoded where used (adjoint).
    // NSH: shots per wave.  With two, a wave advances the same strip of two
    // shots of the batch: the static model fields and the receiver table are
    // loaded once for both, and the two shots' independent arithmetic hides
    // each other's latency.
    static_assert(NSH == 1 || (NSH == 2 && !EX), "two shots per wave only on steps without EX work");
    static_assert(EX >= 0 && EX <= 2, "EX variant");
    constexpr bool S16 = ADJ && EX == 2;
    constexpr bool B16W = !ADJ && EX == 2;   // BF16 snapshots written by the storing step
    constexpr bool VISCO = L > 0;
    const int w = blockIdx.x;
    const int nblk = (NSH == 2) ? a.nblocks2 : (a.sg > 1 ? a.nblocks_sg : a.nblocks);
    const int nblk8 = (NSH == 2) ? a.nblocks2_8 : (a.sg > 1 ? a.nblocks_sg8 : a.nblocks8);
    const int lin = (w & 7) * nblk8 + (w >> 3);
    if (lin >= nblk) return;                      // whole workgroup
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    // The waves of a workgroup step through the rows together (one
    // s_barrier per row).  Strip groups (default): they take adjacent strips
    // of one shot, so their loads of a row arrive at the memory as one
    // contiguous run per field; spare waves of the last group only keep the
    // row barriers.  Shot groups (sg > 1, one shot per wave only): wave wv
    // takes shot wv % sg of sg consecutive shots and strip wv / sg of
    // kStripsPerGroup / sg adjacent strips, so the sg waves of a strip load
    // the same static model rows in the same row iteration and those loads
    // meet in the CU's L1; a wave without an active shot fetches only those
    // rows and stores nothing.
    const bool shot_grp = NSH == 1 && a.sg > 1;
    int p, strip, seg, q0 = 0;
    bool live;
    if (shot_grp) {
        const int sg = a.sg, spg = kStripsPerGroup / sg;
        const int nq = (a.nshots + sg - 1) / sg;
        const int ngs = (a.nstrips + spg - 1) / spg;
        const int q = lin % nq;
        const int tile = lin / nq;
        const int grp = tile % ngs;
        seg = tile / ngs;
        q0 = q * sg;
        p = q0 + wv % sg;
        const int st = grp * spg + wv / sg;
        live = st < a.nstrips;
        strip = live ? st : a.nstrips - 1;
    } else {
        const int npair = (a.nshots + NSH - 1) / NSH;
        p = lin % npair;
        const int tile = lin / npair;
        const int grp = tile % a.ngroups;
        seg = tile / a.ngroups;
        live = grp * kStripsPerGroup + wv < a.nstrips;
        strip = live ? grp * kStripsPerGroup + wv : a.nstrips - 1;
    }
    const Geo &g = a.g;
    const int nt = a.nt_base ? (*a.nt_base + a.nt) : a.nt;

    // the wave's shots: batch slot, absolute shot, record length, active
    int sl[NSH], Sx[NSH], NTs[NSH];
    bool act[NSH];
    bool any = false;
#pragma unroll
    for (int u = 0; u < NSH; ++u) {
        sl[u] = p * NSH + u;
        Sx[u] = 0;
        NTs[u] = 0;
        act[u] = false;
        if (sl[u] < a.nshots) {
            const int S = a.shot_map[sl[u]];
            if (S >= 0 && S < a.nshots_all) {   // a stale or foreign shot map: touch nothing
                Sx[u] = S;
                NTs[u] = a.shot_nt[S];
                act[u] = nt <= NTs[u];          // finished shots (uneven record lengths) rest
            } else {
                a.status[2] = 1;                // reported by the host after the pass
            }
        }
        any = any || act[u];
    }
    if (shot_grp) {
        // the workgroup goes on while any of its shots does (barriers)
        for (int k = 0; k < a.sg && !any; ++k) {
            const int t = q0 + k;
            if (t < a.nshots) {
                const int S = a.shot_map[t];
                any = S >= 0 && S < a.nshots_all && nt <= a.shot_nt[S];
            }
        }
    }
    if (!any) return;                             // uniform over the workgroup

    const int lane = threadIdx.x % kWave;
    const int c0 = 1 + strip * a.ow;          // first owned column
    const int c = c0 - (K - 1) + lane;        // column of this lane
    // BOUNDARY 1: a lane left of column 1 or right of NX computes the
    // column it wraps to (the halo the reference fills by exchange_v /
    // exchange_s); every address of the lane uses cw.  NX >= 64, so one
    // wrap lands inside the grid.
    const int cw = !g.periodic ? c : (c < 1 ? c + g.nx : (c > g.nx ? c - g.nx : c));
    const bool in_x = g.periodic || ((c >= 1) && (c <= g.nx));
    const bool own_x = live && (lane >= K - 1) && (lane < K - 1 + a.ow) && (c >= 1) && (c <= g.nx);
    const int j0 = 1 + seg * a.R;
    const int j1 = min(j0 + a.R - 1, g.ny);
    const int nxny = g.nx * g.ny;                 // per-lane offsets are 32-bit
    const int plane = (int)g.plane;
    const int fw = g.fw, fw2 = 2 * g.fw;

    // base pointers at the wave's first shot (batch slot sl[0]) and the byte
    // offset of shot u from it (the shots of a wave are consecutive slots);
    // per-lane offsets stay 32-bit, as one shot's planes are below 2^31 bytes
    const long long s0 = sl[0];
    const float *__restrict__ vyi = a.vy_in + s0 * g.plane;
    const float *__restrict__ sxi = a.sxy_in + s0 * g.plane;
    const float *__restrict__ szi = a.syz_in + s0 * g.plane;
    float *__restrict__ vyo = a.vy_out + s0 * g.plane;
    float *__restrict__ sxo_ = a.sxy_out + s0 * g.plane;
    float *__restrict__ szo = a.syz_out + s0 * g.plane;
    // r, q interleaved: float4 planes of pairs (l, l+1), then a float2 plane
    // when L is odd; rqp(k) is the base of plane k
    float *__restrict__ rqb = VISCO ? a.rq + s0 * 2 * L * g.plane : nullptr;
    auto rqp = [&](int k) -> float * { return rqb + (long long)k * 4 * g.plane; };
    const float *__restrict__ pxi = a.psx_in + s0 * g.ny * fw2;
    float *__restrict__ pxo = a.psx_out + s0 * g.ny * fw2;
    const float *__restrict__ pyi = a.psy_in + s0 * fw2 * g.nx;
    float *__restrict__ pyo = a.psy_out + s0 * fw2 * g.nx;
    float *__restrict__ pvxb = a.pvx + s0 * g.ny * fw2;
    float *__restrict__ pvyb = a.pvy + s0 * fw2 * g.nx;
    unsigned sof[NSH], sor[NSH], spx[NSH], spy[NSH];
#pragma unroll
    for (int u = 0; u < NSH; ++u) {
        sof[u] = (unsigned)u * (unsigned)plane * 4u;
        sor[u] = (unsigned)u * (unsigned)(L > 0 ? 2 * L : 1) * (unsigned)plane * 4u;
        spx[u] = (unsigned)u * (unsigned)(g.ny * fw2) * 4u;
        spy[u] = (unsigned)u * (unsigned)(fw2 * g.nx) * 4u;
    }

    // EX work of the single shot (NSH == 1)
    const int s = sl[0];
    bool fstore = false, corr = false, hsum = false;
    int slot = 0;
    if (!EX || !act[0]) {
    } else if (!ADJ) {
        if (a.store_mode != 0 && ((nt - 1) % a.dtinv) == 0) {
            fstore = true;
            slot = (nt - 1) / a.dtinv;
        }
        // hessian.c:91: stored steps 1..NTDTINV-1 of this shot
        if (a.hess && ((nt - 1) % a.dtinv) == 0 && (nt - 1) / a.dtinv < (NTs[0] + a.dtinv - 1) / a.dtinv - 1)
            hsum = true;
    } else {
        const int sf = NTs[0] - nt + 1;  // forward step paired with this adjoint step
        if (((sf - 1) % a.dtinv) == 0) {
            corr = true;
            slot = (sf - 1) / a.dtinv;
        }
    }

    __shared__ float lsx_all[kStripsPerGroup][NSH][kWave + 2 * K];
    __shared__ float lvy_all[kStripsPerGroup][NSH][kWave];

    // The padded planes hold a zero frame of at least 2K-1 rows above and
    // below the interior, 2K-1 columns left and 64 right (dn_internal.h), so
    // every row and column this wave touches is inside the plane: loads are
    // unconditional and cells outside the interior read as 0 (wavefields)
    // or are not used (model fields).  Lanes 0..K-1 also fetch the K sxy
    // columns left of the wave, lanes K..2K-2 the K-1 columns right of it.
    const int hcol0 = (lane < K) ? (c0 - 2 * K + 1 + lane) : ((lane < 2 * K - 1) ? (c0 + 65 - 2 * K + lane) : c);
    const int hcol = !g.periodic ? hcol0 : (hcol0 < 1 ? hcol0 + g.nx : (hcol0 > g.nx ? hcol0 - g.nx : hcol0));
    const int hslot = (lane < K) ? lane : (kWave + lane);
    const long long shot_off = (long long)s * nxny;
    float *__restrict__ grb = (EX && a.grad_rho) ? a.grad_rho + shot_off : nullptr;
    float *__restrict__ gub = (EX && a.grad_u) ? a.grad_u + shot_off : nullptr;
    float *__restrict__ wab = (EX && a.wacc) ? a.wacc + shot_off : nullptr;
    // the twelve CPML profile vectors are packed with stride 2*FW+1 in the
    // order K_x, a_x, b_x, K_x_half, a_x_half, b_x_half, K_y, a_y, b_y,
    // K_y_half, a_y_half, b_y_half (one base keeps the scalar registers free)
    const float *__restrict__ pmlb = a.pml.Kx;
    const int pn = fw2 + 1;
    // CPML: the host guarantees that the left and right (top and bottom)
    // absorbing regions do not overlap, so a lane (row) lies in at most one.
    // x: the lane's profile index (1..2FW, 0 outside) and its six profile
    // values, constant over the rows; y: the profile index of a row
    const int xk = (fw > 0 && in_x && !g.periodic) ? ((c <= fw) ? c : ((c >= g.nx - fw + 1) ? c - g.nx + fw2 : 0)) : 0;
    const bool xin = xk > 0;
    const float vKx = pmlb[0 * pn + xk], vax = pmlb[1 * pn + xk], vbx = pmlb[2 * pn + xk];
    const float sKx = pmlb[3 * pn + xk], sax = pmlb[4 * pn + xk], sbx = pmlb[5 * pn + xk];
    auto ykof = [&](int r) -> int {
        if (fw <= 0 || r < 1 || r > g.ny) return 0;
        if (!g.free_surf && r <= fw) return r;
        return (r >= g.ny - fw + 1) ? r - g.ny + fw2 : 0;
    };

    float zw[NSH][2 * K];   // old syz, slot t <-> row jv-K+t (after the head load)
    float vw[NSH][2 * K];   // new vy, slot t <-> row jv-2K+1+t (after the push)
    float sq[NSH][K + 1];   // old sxy at the lane's column, slot t <-> row jv-K+t
    const int jstart = j0 - K + 1;
    const int jend = j1 + K;
    if (!live) {
        // a spare wave of the last strip group: it only keeps the workgroup's
        // per-row barriers (one per iteration of the row loop below)
        for (int jv = jstart; jv <= jend; ++jv) __builtin_amdgcn_s_barrier();
        return;
    }
#pragma unroll
    for (int u = 0; u < NSH; ++u) {
#pragma unroll
        for (int t = 0; t < 2 * K - 1; ++t) zw[u][t] = act[u] ? ldf(szi, sof[u] + gi32(g, jstart - K + t, cw) * 4u) : 0.0f;
#pragma unroll
        for (int t = 0; t < 2 * K; ++t) vw[u][t] = 0.0f;
#pragma unroll
        for (int t = 0; t <= K; ++t) sq[u][t] = 0.0f;
    }

    // one row of inputs, fetched one iteration ahead: velocity row jv and
    // the static fields / memory variables of stress row jv - K
    struct RowM {               // shared by the wave's shots
        float b, s1, s2;
        float dp[L > 0 ? L : 1], dd[L > 0 ? L : 1];
        int rrow;               // receiver row table entry of row jv (-1: none)
        float vKy, vay, vby;    // y-CPML profile of velocity row jv (K, a, b)
        float sKy, say, sby;    // y-CPML profile of stress row jv - K
    };
    struct RowS {               // per shot
        float syz, sc, halo, vy;
        float px, py;           // CPML memory of velocity row jv: psi_sxy_x, psi_syy_y
        float qx, qy;           // CPML memory of stress row jv - K: psi_vyx, psi_vyy
        float r[L > 0 ? L : 1], q[L > 0 ? L : 1];
        float gr, f0;           // velocity row: grad_rho / H_rho sum, stored vyp1 (fp32 snapshots)
        float wa, gu, fy, fu;   // stress row: Ws / Wr, grad_u / H_mu sum, stored uy, uyx (fp32)
        unsigned h0, hy, hu;    // the stored fields of 16-bit snapshots (decoded where used;
                                // full registers, so their loads never merge into one)
    };
    // per-shot snapshot slot of this step and its decode scales
    const long long sbase = (long long)s * a.snap_shot_stride + (long long)slot * 3 * nxny;
    const float *__restrict__ snb = (EX && a.snap) ? a.snap + sbase : nullptr;
    float *__restrict__ snw = (EX && a.snap) ? a.snap + sbase : nullptr;
    const uint16_t *__restrict__ snb16 = (EX && a.snap16) ? a.snap16 + sbase : nullptr;
    float *__restrict__ stg = (EX && a.stage) ? a.stage + (long long)s * 3 * nxny : nullptr;
    uint16_t *__restrict__ sn16w = (B16W && a.snap16) ? a.snap16 + sbase : nullptr;
    // BF16 snapshots stored directly (EX 2 forward): snapshot k of a field is
    // encoded as bf16(x * 2^-e) and decoded by 2^e, with e the exponent of
    // the largest |x| of snapshot k - 1 (0 for the first).  Scaling by a
    // power of two leaves the rounding of normal values unchanged, so the
    // scale only normalises the stored magnitudes.  The maxima go through a
    // ring of three per shot and field: read k - 1, accumulate k, clear k + 1
    // (read by the launch of snapshot k - 2, written by that of k + 1).
    float enc[3] = {1.0f, 1.0f, 1.0f};
    if (B16W && fstore) {
        unsigned int *ring = a.stage_max + (long long)s * 9;
        const int kp = (slot + 2) % 3, kc = (slot + 1) % 3;
        const bool first = live && strip == 0 && seg == 0 && lane == 0;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const float m = __uint_as_float(ring[kp * 3 + f]);
            int e = 0;
            if (m > 0.0f && isfinite(m)) {
                (void)frexpf(m, &e);
                e = min(max(e, -100), 100);
            }
            enc[f] = ldexpf(1.0f, -e);
            if (first) {
                ring[kc * 3 + f] = 0u;
                a.snap_scale[((long long)s * (a.snap_shot_stride / (3 * nxny)) + slot) * 3 + f] = ldexpf(1.0f, e);
            }
        }
    }
    float dsc[3] = {1.0f, 1.0f, 1.0f};
    if (S16 && corr) {
        const float *scp = a.snap_scale + ((long long)s * (a.snap_shot_stride / (3 * nxny)) + slot) * 3;
        dsc[0] = scp[0];
        dsc[1] = scp[1];
        dsc[2] = scp[2];
    }
    auto decode = [&](unsigned h, float scale) -> float {
        // h holds a 16-bit code in its low half
        return (a.snap_kind == 1 ? bf162f(h) : h2f(h)) * scale;
    };
    auto fetch = [&](int jv, RowM &m, RowS (&o)[NSH]) {
        const unsigned gv = gi32(g, jv, cw) * 4u;
        const unsigned gs = gi32(g, jv - K, cw) * 4u;
        // the stress-row inputs are fetched only where the iteration updates
        // the stress (row jv - K inside the segment): the first 2K-1
        // iterations of a segment advance the velocity alone (wave-uniform)
        const int jsr = jv - K;
        const bool sneed = jsr >= j0;
        m.b = ldf(a.bfield, gv);
        m.rrow = (jv >= 1 && jv <= g.ny) ? a.rec_row[jv] : -1;
        if (!sneed) {
        } else if (!VISCO) {
            const float2 t = ld2(a.spk, gs * 2u);
            m.s1 = t.x;
            m.s2 = t.y;
        } else {
            const float4 t = ld4(a.spk, gs * 4u);
            m.s1 = t.x;
            m.s2 = t.y;
            m.dp[0] = t.z;
            m.dd[0] = t.w;
#pragma unroll
            for (int l = 1; l < L; ++l) {
                const float2 u = ld2(a.spk + 4 * g.plane + 2 * (long long)(l - 1) * g.plane, gs * 2u);
                m.dp[l] = u.x;
                m.dd[l] = u.y;
            }
        }
        const int ykv = ykof(jv), yks = sneed ? ykof(jsr) : 0;
        if (ykv > 0) {
            m.vKy = pmlb[6 * pn + ykv];
            m.vay = pmlb[7 * pn + ykv];
            m.vby = pmlb[8 * pn + ykv];
        }
        if (yks > 0) {
            // Q2: the viscoelastic kernel uses the full-grid y profile on top
            const int b0 = (VISCO && !g.free_surf && jsr <= fw) ? 6 : 9;
            m.sKy = pmlb[b0 * pn + yks];
            m.say = pmlb[(b0 + 1) * pn + yks];
            m.sby = pmlb[(b0 + 2) * pn + yks];
        }
        const bool vrow = jv >= 1 && jv <= g.ny, srow = jsr >= 1 && jsr <= g.ny;
        const unsigned gz = gi32(g, jv + K - 1, cw) * 4u, gh = gi32(g, jv, hcol) * 4u;
#pragma unroll
        for (int u = 0; u < NSH; ++u) {
            if (!act[u]) continue;
            o[u].syz = ldf(szi, sof[u] + gz);
            o[u].sc = ldf(sxi, sof[u] + gv);
            o[u].halo = ldf(sxi, sof[u] + gh);
            o[u].vy = ldf(vyi, sof[u] + gv);
            if (xin && vrow) o[u].px = ldf(pxi, spx[u] + (unsigned)((jv - 1) * fw2 + (xk - 1)) * 4u);
            if (ykv > 0 && in_x) o[u].py = ldf(pyi, spy[u] + (unsigned)((ykv - 1) * g.nx + (cw - 1)) * 4u);
            if (xin && srow && sneed) o[u].qx = ldf(pvxb, spx[u] + (unsigned)((jsr - 1) * fw2 + (xk - 1)) * 4u);
            if (yks > 0 && in_x) o[u].qy = ldf(pvyb, spy[u] + (unsigned)((yks - 1) * g.nx + (cw - 1)) * 4u);
            if (VISCO && sneed) {
#pragma unroll
                for (int l = 0; l + 1 < L; l += 2) {
                    const float4 t = ld4(rqp(l / 2), sor[u] + gs * 4u);
                    o[u].r[l] = t.x;
                    o[u].q[l] = t.y;
                    o[u].r[l + 1] = t.z;
                    o[u].q[l + 1] = t.w;
                }
                if (L & 1) {
                    const float2 t = ld2(rqp(L / 2), sor[u] + gs * 2u);
                    o[u].r[L - 1] = t.x;
                    o[u].q[L - 1] = t.y;
                }
            }
        }
        if (EX) {
            // The accumulators and snapshots are read only where this wave
            // writes them back.  Every lane loads (others at element 0 of the
            // same arrays, unused), so no lane keeps an older value in a
            // register with a load in flight and nothing waits here; the
            // snapshot storage (FP32 or 16-bit) is fixed per variant.
            RowS &e = o[0];
            const bool vin = (jv >= j0 && jv <= j1) && own_x;
            const unsigned cv = vin ? (unsigned)((jv - 1) * g.nx + (cw - 1)) : 0u;
            if ((ADJ && corr) || (!ADJ && hsum)) {
                e.gr = grb[cv];
                if (ADJ && !S16) e.f0 = snb[cv];
                if (S16) e.h0 = snb16[cv];
            }
            const int jsn = jv - K;
            const bool sin_ = (jsn >= j0 && jsn <= j1) && own_x;
            const unsigned cs = sin_ ? (unsigned)((jsn - 1) * g.nx + (cw - 1)) : 0u;
            if ((ADJ && corr) || (!ADJ && fstore)) e.wa = wab[cs];
            if ((ADJ && corr) || (!ADJ && hsum)) e.gu = gub[cs];
            if (ADJ && corr && !S16) {
                e.fy = snb[nxny + cs];
                e.fu = snb[2 * nxny + cs];
            }
            if (S16 && corr) {
                e.hy = snb16[nxny + cs];
                e.hu = snb16[2 * nxny + cs];
            }
        }
    };
    // does a source of the shot fall into the wave's columns and rows?
    bool has_src[NSH];
    int sb[NSH], se[NSH];
#pragma unroll
    for (int u = 0; u < NSH; ++u) {
        has_src[u] = false;
        sb[u] = (ADJ || !act[u]) ? 0 : a.src_off[Sx[u]];
        se[u] = (ADJ || !act[u]) ? 0 : a.src_off[Sx[u] + 1];
        if (!ADJ)
            for (int l = sb[u]; l < se[u]; ++l) {
                const int si = a.src_i[l], sj = a.src_j[l];
                if (sj >= jstart && sj <= jend && (g.periodic || (si >= c0 - (K - 1) && si <= c0 - K + kWave)))
                    has_src[u] = true;
            }
    }
    // per-snapshot max |x| of the values this wave stages for the 16-bit
    // codec (store_mode 2), kept while they are written
    float smx0 = 0.0f, smx1 = 0.0f, smx2 = 0.0f;
    // one iteration: velocity row jv and stress row jv - K, from inputs
    // fetched earlier (cm, cr)
    auto row = [&](const int jv, const RowM &cm, const RowS (&cr)[NSH]) {
        float vnew[NSH], sc[NSH];
#pragma unroll
        for (int u = 0; u < NSH; ++u) {
            zw[u][2 * K - 1] = cr[u].syz;
            sc[u] = cr[u].sc;
            lsx_all[wv][u][K + lane] = sc[u];
            if (lane < 2 * K - 1) lsx_all[wv][u][hslot] = cr[u].halo;
        }
        __builtin_amdgcn_wave_barrier();

#pragma unroll
        for (int u = 0; u < NSH; ++u) {
            vnew[u] = 0.0f;
            if (!act[u] || !(jv >= 1 && jv <= g.ny && in_x)) continue;
            const float *lsx = lsx_all[wv][u];
            const bool own_v = own_x && jv >= j0 && jv <= j1;
            float sxy_x = a.hc[1] * (lsx[K + lane] - lsx[K + lane - 1]);
            float syy_y = a.hc[1] * (zw[u][K] - zw[u][K - 1]);
#pragma unroll
            for (int m = 2; m <= K; ++m) {
                sxy_x = sxy_x + a.hc[m] * (lsx[K + lane + m - 1] - lsx[K + lane - m]);
                const float hy = (K == 6 && m == 6) ? a.hcy6 : a.hc[m];
                syy_y = syy_y + hy * (zw[u][K + m - 1] - zw[u][K - m]);
            }
            if (xin) {   // update_v_PML.c: x-CPML, left (c <= FW) or right (c >= NX-FW+1)
                const float p = vbx * cr[u].px + vax * sxy_x;
                if (own_v) stf(pxo, spx[u] + (unsigned)((jv - 1) * fw2 + (xk - 1)) * 4u, p);
                sxy_x = sxy_x / vKx + p;
            }
            {
                const int ykv = ykof(jv);
                if (ykv > 0) {   // y-CPML, top (no free surface) or bottom
                    const float p = cm.vby * cr[u].py + cm.vay * syy_y;
                    if (own_v) stf(pyo, spy[u] + (unsigned)((ykv - 1) * g.nx + (cw - 1)) * 4u, p);
                    syy_y = syy_y / cm.vKy + p;
                }
            }
            const unsigned gi = gi32(g, jv, cw);
            const float vold = cr[u].vy;
            const float bb = cm.b;
            const float tsum = sxy_x + syy_y;
            const unsigned ci = (unsigned)((jv - 1) * g.nx + (cw - 1));
            if (EX && !ADJ && hsum && own_v) {
                const float acc = (bb * tsum) / a.DH;  // vyp1 (hess_utty)
                grb[ci] = cr[u].gr + (acc * acc);
            }
            if (EX && !ADJ && fstore && own_v) {
                const float acc = (bb * tsum) / a.DH;  // vyp1, update_v_PML.c:184
                if (B16W) {
                    sn16w[ci] = f2bf16(acc * enc[0]);
                    smx0 = fmaxf(smx0, fabsf(acc));
                } else if (a.store_mode == 1) {
                    snw[ci] = acc;
                } else {
                    stg[ci] = acc;
                    smx0 = fmaxf(smx0, fabsf(acc));
                }
            }
            if (EX && ADJ && corr && own_v) {
                // vyp1 = old vy (update_v_PML.c:189); denise.c:1589
                const float f0 = S16 ? decode(cr[u].h0, dsc[0]) : cr[u].f0;
                grb[ci] = cr[u].gr + (vold * f0);
            }
            float vn = vold + ((bb * a.DT) * tsum) / a.DH;
            if (!ADJ) {
                if (has_src[u])
                    for (int l = sb[u]; l < se[u]; ++l)
                        if (a.src_j[l] == jv && a.src_i[l] == cw)
                            vn = vn + a.src_sig[(long long)l * a.nt_max + (nt - 1)];
            } else {
                const int off = cm.rrow;
                if (off >= 0) {
                    const int r = a.rec_tab[off + cw];
                    if (r >= 0) vn = vn + a.resid[((long long)sl[u] * a.ntr + r) * a.nt_max + (nt - 1)];
                }
            }
            vnew[u] = vn;
            // instability latch at the reference's probe cell (row test first:
            // wave-uniform, true on one row of the grid)
            if (jv == (g.ny >> 1) && own_v && c == (g.nx >> 1) && !(vn == vn))
                if (atomicCAS(&a.status[0], 0, nt) == 0) a.status[1] = Sx[u] + 1;
            if (own_v) {
                stf(vyo, sof[u] + gi * 4u, vn);
                if (!ADJ && a.seis != nullptr) {
                    const int off = cm.rrow;
                    if (off >= 0) {
                        const int r = a.rec_tab[off + cw];
                        if (r >= 0) a.seis[((long long)sl[u] * a.ntr + r) * a.nt_max + (nt - 1)] = vn;
                    }
                }
            }
        }

#pragma unroll
        for (int u = 0; u < NSH; ++u) {
#pragma unroll
            for (int t = 0; t < 2 * K - 1; ++t) vw[u][t] = vw[u][t + 1];
            vw[u][2 * K - 1] = vnew[u];
#pragma unroll
            for (int t = 0; t < K; ++t) sq[u][t] = sq[u][t + 1];
            sq[u][K] = sc[u];
        }

        const int js = jv - K;  // stress row of this iteration
        if (js >= j0) {
#pragma unroll
            for (int u = 0; u < NSH; ++u) lvy_all[wv][u][lane] = vw[u][K - 1];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int u = 0; u < NSH; ++u) {
                if (!act[u] || !own_x) continue;
                const float *lvy = lvy_all[wv][u];
                float vyx = a.hc[1] * (lvy[lane + 1] - lvy[lane]);
                float vyy = a.hc[1] * (vw[u][K] - vw[u][K - 1]);
#pragma unroll
                for (int m = 2; m <= K; ++m) {
                    vyx = vyx + a.hc[m] * (lvy[lane + m] - lvy[lane - m + 1]);
                    vyy = vyy + a.hc[m] * (vw[u][K - 1 + m] - vw[u][K - m]);
                }
                vyx = vyx * a.dhi;
                vyy = vyy * a.dhi;
                if (xin) {   // update_s_*_PML.c: x-CPML of the stress
                    const float p = sbx * cr[u].qx + sax * vyx;
                    stf(pvxb, spx[u] + (unsigned)((js - 1) * fw2 + (xk - 1)) * 4u, p);
                    vyx = vyx / sKx + p;
                }
                {
                    const int yks = ykof(js);
                    if (yks > 0) {   // y-CPML of the stress (profile chosen in fetch, Q2)
                        const float p = cm.sby * cr[u].qy + cm.say * vyy;
                        stf(pvyb, spy[u] + (unsigned)((yks - 1) * g.nx + (cw - 1)) * 4u, p);
                        vyy = vyy / cm.sKy + p;
                    }
                }
                const unsigned gs = gi32(g, js, cw);
                const float sxo = sq[u][0];
                const float syo = zw[u][0];
                float sxn, syn;
                if (!VISCO) {
                    sxn = sxo + cm.s1 * vyx;
                    syn = syo + cm.s2 * vyy;
                } else {
                    float rr[L > 0 ? L : 1], qq[L > 0 ? L : 1];
                    float sumr = 0.0f, sumq = 0.0f;
#pragma unroll
                    for (int l = 0; l < L; ++l) {
                        rr[l] = cr[u].r[l];
                        qq[l] = cr[u].q[l];
                        sumr += rr[l];
                        sumq += qq[l];
                    }
                    sxn = sxo + ((cm.s1 * vyx) + (a.dthalbe * sumr));
                    syn = syo + ((cm.s2 * vyy) + (a.dthalbe * sumq));
                    sumr = 0.0f;
                    sumq = 0.0f;
#pragma unroll
                    for (int l = 0; l < L; ++l) {
                        rr[l] = a.bip[l] * (rr[l] * a.cip[l] - (cm.dp[l] * vyx));
                        qq[l] = a.bjm[l] * (qq[l] * a.cjm[l] - (cm.dd[l] * vyy));
                        sumr += rr[l];
                        sumq += qq[l];
                    }
#pragma unroll
                    for (int l = 0; l + 1 < L; l += 2)
                        st4(rqp(l / 2), sor[u] + gs * 16u, make_float4(rr[l], qq[l], rr[l + 1], qq[l + 1]));
                    if (L & 1) st2(rqp(L / 2), sor[u] + gs * 8u, make_float2(rr[L - 1], qq[L - 1]));
                    sxn = sxn + (a.dthalbe * sumr);
                    syn = syn + (a.dthalbe * sumq);
                }
                stf(sxo_, sof[u] + gs * 4u, sxn);
                stf(szo, sof[u] + gs * 4u, syn);
                if (EX) {
                    const unsigned ci = (unsigned)((js - 1) * g.nx + (cw - 1));
                    const float vc = vw[u][K - 1];
                    if (!ADJ && fstore) {
                        if (B16W) {
                            uint16_t *b0 = sn16w + ci;
                            b0[nxny] = f2bf16(vyy * enc[1]);   // uy
                            b0[2 * nxny] = f2bf16(vyx * enc[2]);   // uyx
                            smx1 = fmaxf(smx1, fabsf(vyy));
                            smx2 = fmaxf(smx2, fabsf(vyx));
                        } else if (a.store_mode == 1) {
                            float *b0 = snw + ci;
                            b0[nxny] = vyy;        // uy  (forward_prop_y)
                            b0[2 * nxny] = vyx;    // uyx (forward_prop_u)
                        } else {
                            float *b0 = stg + ci;
                            b0[nxny] = vyy;
                            b0[2 * nxny] = vyx;
                            smx1 = fmaxf(smx1, fabsf(vyy));
                            smx2 = fmaxf(smx2, fabsf(vyx));
                        }
                        wab[ci] = cr[u].wa + vc * vc;   // eprecond.c:11
                    }
                    if (!ADJ && hsum) {
                        const float eyx = 0.5f * vyx;     // hess_save.c:13 (exact in float)
                        gub[ci] = cr[u].gu + ((vyy * vyy) + (eyx * eyx));
                    }
                    if (ADJ && corr) {
                        const float fy = S16 ? decode(cr[u].hy, dsc[1]) : cr[u].fy;
                        const float fu = S16 ? decode(cr[u].hu, dsc[2]) : cr[u].fu;
                        gub[ci] = cr[u].gu + ((fu * sxn) + (fy * syn));   // denise.c:1609
                        wab[ci] = cr[u].wa + vc * vc;
                    }
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_s_barrier();   // keep the workgroup's strips on the same row
#pragma unroll
        for (int u = 0; u < NSH; ++u) {
#pragma unroll
            for (int t = 0; t < 2 * K - 1; ++t) zw[u][t] = zw[u][t + 1];
        }
    };

    // inputs are fetched one row ahead: the wait at the end of an
    // iteration is for the loads issued at its start
    RowM nm_;
    RowS ns_[NSH];
    fetch(jstart, nm_, ns_);
    for (int jv = jstart; jv <= jend; ++jv) {
        const RowM cm = nm_;
        RowS cr[NSH];
#pragma unroll
        for (int u = 0; u < NSH; ++u) cr[u] = ns_[u];
        // the next row's loads go out at raised wave priority, so a wave
        // that reaches its fetch issues it before the others' arithmetic
        __builtin_amdgcn_s_setprio(2);
        if (jv < jend) fetch(jv + 1, nm_, ns_);
        __builtin_amdgcn_s_setprio(0);
        row(jv, cm, cr);
    }

    // per-snapshot max |x| for the reduced-precision codecs (FP16 staging:
    // the scale of this snapshot; BF16: the scale of the next one)
    if (EX && !ADJ && fstore && (B16W || a.store_mode == 2) && live) {
        // the maxima over the cells this wave stored (owned cells only)
        const float m0 = wave_max(smx0);
        const float m1 = wave_max(smx1);
        const float m2 = wave_max(smx2);
        unsigned int *mx = B16W ? a.stage_max + s * 9 + (slot % 3) * 3 : a.stage_max + s * 3;
        if (lane == 0) {
            atomicMax(&mx[0], __float_as_uint(m0));
            atomicMax(&mx[1], __float_as_uint(m1));
            atomicMax(&mx[2], __float_as_uint(m2));
        }
    }
}

template <int K, int L>
static hipError_t launch_step_kl(const StepArgs &a, bool adjoint, bool extra, hipStream_t st)
{
    // plain steps of batches with two or more shots take two shots per wave
    const bool two = !extra && a.nshots >= 2 && a.nblocks2 > 0;
    const int n8 = two ? a.nblocks2_8 : (a.sg > 1 ? a.nblocks_sg8 : a.nblocks8);
    dim3 grid((unsigned)(n8 * 8)), block(kWave * kStripsPerGroup);
    if (adjoint) {
        if (extra && a.snap_kind != 0)
            hipLaunchKernelGGL((k_step<K, L, true, 2, 1>), grid, block, 0, st, a);
        else if (extra)
            hipLaunchKernelGGL((k_step<K, L, true, 1, 1>), grid, block, 0, st, a);
        else if (two)
            hipLaunchKernelGGL((k_step<K, L, true, 0, 2>), grid, block, 0, st, a);
        else
            hipLaunchKernelGGL((k_step<K, L, true, 0, 1>), grid, block, 0, st, a);
    } else {
        if (extra && a.store_mode == 3)
            hipLaunchKernelGGL((k_step<K, L, false, 2, 1>), grid, block, 0, st, a);
        else if (extra)
            hipLaunchKernelGGL((k_step<K, L, false, 1, 1>), grid, block, 0, st, a);
        else if (two)
            hipLaunchKernelGGL((k_step<K, L, false, 0, 2>), grid, block, 0, st, a);
        else
            hipLaunchKernelGGL((k_step<K, L, false, 0, 1>), grid, block, 0, st, a);
    }
    return hipGetLastError();
}

template <int K>
static hipError_t launch_step_k(const StepArgs &a, int L, bool adjoint, bool extra, hipStream_t st)
{
    switch (L) {
    case 0: return launch_step_kl<K, 0>(a, adjoint, extra, st);
    case 1: return launch_step_kl<K, 1>(a, adjoint, extra, st);
    case 2: return launch_step_kl<K, 2>(a, adjoint, extra, st);
    case 3: return launch_step_kl<K, 3>(a, adjoint, extra, st);
    case 4: return launch_step_kl<K, 4>(a, adjoint, extra, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_step(const StepArgs &a, int K, bool visco, bool adjoint, bool extra, hipStream_t st)
{
    const int L = visco ? a.g.L : 0;
    if (visco != (a.g.L > 0)) return hipErrorInvalidValue;
    switch (K) {
    case 1: return launch_step_k<1>(a, L, adjoint, extra, st);
    case 2: return launch_step_k<2>(a, L, adjoint, extra, st);
    case 3: return launch_step_k<3>(a, L, adjoint, extra, st);
    case 4: return launch_step_k<4>(a, L, adjoint, extra, st);
    case 5: return launch_step_k<5>(a, L, adjoint, extra, st);
    case 6: return launch_step_k<6>(a, L, adjoint, extra, st);
    default: return hipErrorInvalidValue;
    }
}

// Pseudo-Hessian: add the per-shot sums in shot order (the reference's
// accumulator runs on across shots) and, after the last batch, the chain
// rule of hessian_out.c:25-31 with the reference's double promotion.
__global__ void k_hessian_finish(int nshots, int nx, int ny, Geo g, const float *hu_shot, const float *hr_shot,
                                 const float *vs_pad, const float *rho_pad, float *hu, float *hr, int first,
                                 int last)
{
    const int nxny = nx * ny;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nxny; t += gridDim.x * blockDim.x) {
        float a = first ? 0.0f : hu[t];
        float b = first ? 0.0f : hr[t];
        for (int s = 0; s < nshots; ++s) {
            a = a + hu_shot[(long long)s * nxny + t];
            b = b + hr_shot[(long long)s * nxny + t];
        }
        if (last) {
            const int j = t / nx + 1, i = t % nx + 1;
            const float u = vs_pad[gidx(g, j, i)], r = rho_pad[gidx(g, j, i)];
            b = b + u * u * u * u * a;
            a = (float)(4.0 * r * r * u * u * a);
        }
        hu[t] = a;
        hr[t] = b;
    }
}

hipError_t launch_hessian_finish(int nshots, int nx, int ny, const Geo &g, const float *hu_shot,
                                 const float *hr_shot, const float *vs_pad, const float *rho_pad, float *hu,
                                 float *hr, int first, int last, hipStream_t st)
{
    const int nxny = nx * ny;
    hipLaunchKernelGGL(k_hessian_finish, dim3(std::min(2048, (nxny + 255) / 256)), dim3(256), 0, st, nshots, nx, ny,
                       g, hu_shot, hr_shot, vs_pad, rho_pad, hu, hr, first, last);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Model preparation (denise.c:726-743): periodic halo (matcopy.c with one
// rank), av_mue.c (INVMAT1 1), av_tau.c, 1/rho of update_v_PML.c:181-182 and
// prepare_update_s.c:41-62.  Double where the reference promotes.
// ---------------------------------------------------------------------------
__global__ void k_model_wrap(Geo g, float *vs, float *rho, float *taus)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // 0..nx+1
    const int j = blockIdx.y;                              // 0..ny+1
    if (i > g.nx + 1 || j > g.ny + 1) return;
    const bool edge = (i == 0 || i == g.nx + 1 || j == 0 || j == g.ny + 1);
    if (!edge) return;
    // rows first, then columns (matcopy.c:40-123): the final value of a halo
    // cell is the interior cell reached by wrapping each index once
    const int jj = (j == 0) ? g.ny : (j == g.ny + 1 ? 1 : j);
    const int ii = (i == 0) ? g.nx : (i == g.nx + 1 ? 1 : i);
    const long long d = gidx(g, j, i), sidx = gidx(g, jj, ii);
    vs[d] = vs[sidx];
    rho[d] = rho[sidx];
    if (taus) taus[d] = taus[sidx];
}

__global__ void k_model_coef(ModelPrepArgs a)
{
    const Geo &g = a.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y + 1;
    if (i > g.nx || j > g.ny) return;
    const long long o = gidx(g, j, i);
    const long long oi = gidx(g, j, i + 1), oj = gidx(g, j + 1, i), oij = gidx(g, j + 1, i + 1);
    // av_mue.c:33-46
    const float u1 = a.rho[o] * a.vs[o] * a.vs[o];
    const float u2 = a.rho[oi] * a.vs[oi] * a.vs[oi];
    const float u3 = a.rho[oj] * a.vs[oj] * a.vs[oj];
    float uip = (float)(2.0 / ((1.0 / (double)u1) + (1.0 / (double)u2)));
    float ujp = (float)(2.0 / ((1.0 / (double)u1) + (1.0 / (double)u3)));
    if ((u1 == 0.0f) || (u2 == 0.0f)) uip = 0.0f;
    if ((u1 == 0.0f) || (u3 == 0.0f)) ujp = 0.0f;
    a.uip[o] = uip;
    a.ujp[o] = ujp;
    if (a.L == 0) {
        a.spk[2 * o] = uip;
        a.spk[2 * o + 1] = ujp;
    }
    // update_v_PML.c:181-182
    const float rv = a.rho[o];
    float b = (float)(1.0 / (double)rv);
    if ((double)rv < 1e-20) b = 0.0f;
    a.bfield[o] = b;
    if (a.L > 0) {
        // av_tau.c:12
        const float tsum = a.taus[o] + a.taus[oi] + a.taus[oj] + a.taus[oij];
        const float tip = (float)(0.25 * (double)tsum);
        a.tausipjp[o] = tip;
        const float ptaus = a.taus[o];
        // prepare_update_s.c:45-57
        const float mujp = (float)((double)ujp / (1.0 + (double)(a.sumu * ptaus)));
        const float muip = (float)((double)uip / (1.0 + (double)(a.sumu * tip)));
        const float fipjp = (float)((double)(muip * a.DT) * (1.0 + (double)((float)a.L * tip)));
        const float f = (float)((double)(mujp * a.DT) * (1.0 + (double)((float)a.L * ptaus)));
        a.fipjp[o] = fipjp;
        a.fcoef[o] = f;
        a.spk[4 * o] = fipjp;
        a.spk[4 * o + 1] = f;
        for (int l = 0; l < a.L; ++l) {
            const float dp = muip * a.eta[l] * tip, dd = mujp * a.eta[l] * ptaus;
            a.dip[(long long)l * g.plane + o] = dp;
            a.dcoef[(long long)l * g.plane + o] = dd;
            float *pk = (l == 0) ? a.spk + 4 * o + 2 : a.spk + 4 * g.plane + 2 * ((long long)(l - 1) * g.plane + o);
            pk[0] = dp;
            pk[1] = dd;
        }
    }
}

hipError_t launch_model_prep(const ModelPrepArgs &a, hipStream_t st)
{
    dim3 b1(128), g1((a.g.nx + 2 + 127) / 128, a.g.ny + 2);
    hipLaunchKernelGGL(k_model_wrap, g1, b1, 0, st, a.g, a.vs, a.rho, a.L > 0 ? a.taus : (float *)nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    dim3 b2(128), g2((a.g.nx + 127) / 128, a.g.ny);
    hipLaunchKernelGGL(k_model_coef, g2, b2, 0, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Residual and misfit (calc_res.c:150-257 LNORM 2 / GRAD_FORM 2, Q7, and
// calc_energy.c:53-61).  One thread per (shot, trace); per-trace FP64 partial
// sums are written in place and reduced on the host in a fixed order.
// ---------------------------------------------------------------------------
// flags [nshots_all][ntr] (or null): bit 0 leaves the trace out of the
// residual and L2 (trace kill or offset mute, calc_res.c:99-113), bit 1 out
// of the energy (trace kill as calc_energy.c:23-49 reads the kill file)
__global__ void k_residual(int nshots, int nshots_all, int ntr, int nt_max, const int *shot_map, const int *shot_nt,
                           const float *syn, const float *obs, float *resid, double *part, float DT,
                           const unsigned char *flags)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nshots * ntr) return;
    const int s = t / ntr, r = t % ntr;
    const int S = shot_map[s];
    if (S < 0 || S >= nshots_all) {
        part[2 * (long long)t] = 0.0;
        part[2 * (long long)t + 1] = 0.0;
        return;
    }
    const int ns = shot_nt[S];
    const float *sy = syn + (long long)t * nt_max;
    const float *ob = obs + ((long long)S * ntr + r) * nt_max;
    float *df = resid + (long long)t * nt_max;
    const unsigned fl = flags ? flags[(long long)S * ntr + r] : 0u;
    const bool res = !(fl & 1u), nrg = !(fl & 2u);
    double l2 = 0.0, en = 0.0;
    float intseis = 0.0f;
    for (int j = 1; j <= ns; ++j) {
        // calc_res.c:152 sets obs[1] = syn[1] for the traces it covers
        const float o = (j == 1 && res) ? sy[0] : ob[j - 1];
        const float d = res ? sy[j - 1] - o : 0.0f;
        df[ns - j] = d;                                 // time reversal, calc_res.c:224
        l2 += (double)d * (double)d;
        if (nrg) {
            intseis = intseis + o;                      // calc_energy.c:59-60
            en += (double)(intseis * intseis * DT * DT);
        }
    }
    for (int j = ns; j < nt_max; ++j) df[j] = 0.0f;
    part[2 * (long long)t] = l2;
    part[2 * (long long)t + 1] = en;
}

hipError_t launch_residual(int nshots, int nshots_all, int ntr, int nt_max, const int *shot_map, const int *shot_nt,
                           const float *syn, const float *obs, float *resid, double *part, float DT,
                           const unsigned char *flags, hipStream_t st)
{
    const int n = nshots * ntr;
    hipLaunchKernelGGL(k_residual, dim3((n + 63) / 64), dim3(64), 0, st, nshots, nshots_all, ntr, nt_max, shot_map,
                       shot_nt, syn, obs, resid, part, DT, flags);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// eprecond1.c:15-70 and denise.c:1665-1694: We = sqrt(Ws*Wr) per shot,
// We += 0.005*max(We), scale the shot gradients and add them to the sums in
// shot order (deterministic).
// ---------------------------------------------------------------------------
// energy weight of one cell (eprecond1.c:15-50): EPRECOND 1 sqrt(Ws Wr);
// EPRECOND 3 sqrt(Ws) times the Plessix-Mulder (2004) receiver-side
// approximation asinh((xmax - x)/y) - asinh((xmin - x)/y) for receivers
// along the surface from xmin = DH to xmax = NX DH, at x = i DH, y = j DH
// (float coordinates, double asinh and sqrt as the reference's C calls)
__device__ __forceinline__ float energy_weight(int precond, float ws, float wr, int t, int nx, float DH)
{
    if (precond == 3) {
        const int i = t % nx + 1, j = t / nx + 1;
        const float x = (float)i * DH, y = (float)j * DH, xmin = DH, xmax = (float)nx * DH;
        return (float)(sqrt((double)ws) * (asinh((double)((xmax - x) / y)) - asinh((double)((xmin - x) / y))));
    }
    return (float)sqrt((double)(ws * wr));
}

__global__ void k_we_max(int nxny, int nx, float DH, int precond, const float *Ws, const float *Wr,
                         unsigned int *maxbits)
{
    const int s = blockIdx.y;
    float m = 0.0f;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nxny; t += gridDim.x * blockDim.x) {
        const long long o = (long long)s * nxny + t;
        const float we = energy_weight(precond, Ws[o], Wr[o], t, nx, DH);
        m = fmaxf(m, we);
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) atomicMax(&maxbits[s], __float_as_uint(m));
}

__global__ void k_eprecond_acc(int nshots, int nxny, int nx, float DH, const float *Ws, const float *Wr,
                               const float *gu_shot,
                               const float *grho_shot, float *gu, float *grho, float C_vs, float C_rho,
                               const unsigned int *maxbits, int precond, float *ws_sum, float *wr_sum, int first)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nxny) return;
    const float EPSILON_WE = 0.005f;
    float au = gu[t], ar = grho[t];
    float es = first ? 0.0f : ws_sum[t], er = first ? 0.0f : wr_sum[t];
    for (int s = 0; s < nshots; ++s) {
        const long long o = (long long)s * nxny + t;
        es = es + Ws[o];
        er = er + Wr[o];
        if (precond) {
            float we = energy_weight(precond, Ws[o], Wr[o], t, nx, DH);
            we = we + (EPSILON_WE * __uint_as_float(maxbits[s]));
            const float su = gu_shot[o] / (we * C_vs * C_vs);
            const float sr = grho_shot[o] / (we * C_rho * C_rho);
            ar = ar + sr;
            au = au + su;
        } else {
            ar = ar + grho_shot[o];
            au = au + gu_shot[o];
        }
    }
    gu[t] = au;
    grho[t] = ar;
    ws_sum[t] = es;
    wr_sum[t] = er;
}

hipError_t launch_eprecond_accumulate(int nshots, int nx, int ny, float DH, const float *Ws, const float *Wr,
                                      const float *gu_shot, const float *grho_shot, float *gu, float *grho,
                                      float C_vs, float C_rho, unsigned int *maxbits, int precond, float *ws_sum,
                                      float *wr_sum, int first, hipStream_t st)
{
    const int nxny = nx * ny;
    if (precond) {
        hipError_t e = launch_fill_u32(maxbits, (size_t)nshots, 0u, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_we_max, dim3(64, nshots), dim3(256), 0, st, nxny, nx, DH, precond, Ws, Wr, maxbits);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_eprecond_acc, dim3((nxny + 255) / 256), dim3(256), 0, st, nshots, nxny, nx, DH, Ws, Wr,
                       gu_shot, grho_shot, gu, grho, C_vs, C_rho, maxbits, precond, ws_sum, wr_sum, first);
    return hipGetLastError();
}

__global__ void k_eprecond_terms(int nxny, int nx, float DH, const float *Ws, const float *Wr, const float *gu_shot,
                                 const float *grho_shot, float C_vs, float C_rho, const unsigned int *maxbits,
                                 int precond, const int *rank, float *terms)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (t >= nxny) return;
    const float EPSILON_WE = 0.005f;
    const long long o = (long long)s * nxny + t;
    float su, sr;
    if (precond) {
        float we = energy_weight(precond, Ws[o], Wr[o], t, nx, DH);
        we = we + (EPSILON_WE * __uint_as_float(maxbits[s]));
        su = gu_shot[o] / (we * C_vs * C_vs);
        sr = grho_shot[o] / (we * C_rho * C_rho);
    } else {
        su = gu_shot[o];
        sr = grho_shot[o];
    }
    float *T = terms + (long long)rank[s] * 4 * nxny;
    T[t] = Ws[o];
    T[nxny + t] = Wr[o];
    T[2 * (long long)nxny + t] = su;
    T[3 * (long long)nxny + t] = sr;
}

__global__ void k_eprecond_sum(int nterms, int nxny, const float *terms, float *gu, float *grho, float *ws_sum,
                               float *wr_sum, int first)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nxny) return;
    float au = gu[t], ar = grho[t];
    float es = first ? 0.0f : ws_sum[t], er = first ? 0.0f : wr_sum[t];
    for (int r = 0; r < nterms; ++r) {
        const float *T = terms + (long long)r * 4 * nxny;
        es = es + T[t];
        er = er + T[nxny + t];
        au = au + T[2 * (long long)nxny + t];
        ar = ar + T[3 * (long long)nxny + t];
    }
    gu[t] = au;
    grho[t] = ar;
    ws_sum[t] = es;
    wr_sum[t] = er;
}

hipError_t launch_eprecond_terms(int nshots, int nx, int ny, float DH, const float *Ws, const float *Wr,
                                 const float *gu_shot, const float *grho_shot, float C_vs, float C_rho,
                                 unsigned int *maxbits, int precond, const int *rank, float *terms, hipStream_t st)
{
    const int nxny = nx * ny;
    if (precond) {
        hipError_t e = launch_fill_u32(maxbits, (size_t)nshots, 0u, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_we_max, dim3(64, nshots), dim3(256), 0, st, nxny, nx, DH, precond, Ws, Wr, maxbits);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_eprecond_terms, dim3((nxny + 255) / 256, nshots), dim3(256), 0, st, nxny, nx, DH, Ws, Wr,
                       gu_shot, grho_shot, C_vs, C_rho, maxbits, precond, rank, terms);
    return hipGetLastError();
}

hipError_t launch_eprecond_sum(int nterms, int nx, int ny, const float *terms, float *gu, float *grho,
                               float *ws_sum, float *wr_sum, int first, hipStream_t st)
{
    const int nxny = nx * ny;
    hipLaunchKernelGGL(k_eprecond_sum, dim3((nxny + 255) / 256), dim3(256), 0, st, nterms, nxny, terms, gu, grho,
                       ws_sum, wr_sum, first);
    return hipGetLastError();
}

// Rademacher sign of (seed, probe, cell): one round of a 32-bit integer
// hash (splitmix-style finaliser) of the combined key, top bit.
__device__ inline float rademacher(unsigned int seed, int probe, int cell)
{
    unsigned int x = seed * 0x9E3779B9u ^ ((unsigned int)probe * 0x85EBCA6Bu) ^ ((unsigned int)cell * 0xC2B2AE35u);
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return (x & 0x80000000u) ? -1.0f : 1.0f;
}

__global__ void k_probe_perturb(Geo g, const float *vs0, float *vs, float eps, unsigned int seed, int probe)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nx * g.ny) return;
    const int j = t / g.nx + 1, i = t % g.nx + 1;
    const long long o = gidx(g, j, i);
    vs[o] = vs0[o] * (1.0f + eps * rademacher(seed, probe, t));
}

__global__ void k_probe_acc(Geo g, const float *vs0, const float *g0, const float *g1, float eps, unsigned int seed,
                            int probe, float *diag)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nx * g.ny) return;
    const int j = t / g.nx + 1, i = t % g.nx + 1;
    const float v = rademacher(seed, probe, t);
    const float dm = eps * vs0[gidx(g, j, i)];
    diag[t] = diag[t] + v * (g1[t] - g0[t]) / dm;
}

hipError_t launch_probe_perturb(const Geo &g, const float *vs0_pad, float *vs_pad, float eps, unsigned int seed,
                                int probe, hipStream_t st)
{
    const int n = g.nx * g.ny;
    hipLaunchKernelGGL(k_probe_perturb, dim3((n + 255) / 256), dim3(256), 0, st, g, vs0_pad, vs_pad, eps, seed, probe);
    return hipGetLastError();
}

hipError_t launch_probe_accumulate(const Geo &g, const float *vs0_pad, const float *g0, const float *g1, float eps,
                                   unsigned int seed, int probe, float *diag, hipStream_t st)
{
    const int n = g.nx * g.ny;
    hipLaunchKernelGGL(k_probe_acc, dim3((n + 255) / 256), dim3(256), 0, st, g, vs0_pad, g0, g1, eps, seed, probe,
                       diag);
    return hipGetLastError();
}

// denise.c:1756-1811 with INVMAT 0, INVMAT1 1
__global__ void k_chain_rule(Geo g, const float *gu, const float *grho, const float *vs, const float *rho, float DT,
                             float *gvs, float *grh)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nx * g.ny) return;
    const int j = t / g.nx + 1, i = t % g.nx + 1;
    const long long o = gidx(g, j, i);
    const float pu = vs[o], pr = rho[o];
    const float gmu = -DT * gu[t];
    const float grs = -DT * grho[t];
    gvs[t] = (float)(2.0 * (double)pr * (double)pu * (double)gmu);
    grh[t] = (pu * pu * gmu) + grs;
}

hipError_t launch_chain_rule(const Geo &g, int nx, int ny, const float *gu, const float *grho, const float *vs_pad,
                             const float *rho_pad, float DT, float *grad_vs, float *grad_rho, hipStream_t st)
{
    (void)nx;
    (void)ny;
    const int n = g.nx * g.ny;
    hipLaunchKernelGGL(k_chain_rule, dim3((n + 255) / 256), dim3(256), 0, st, g, gu, grho, vs_pad, rho_pad, DT,
                       grad_vs, grad_rho);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Staged snapshot codec (FP16 snapshots, store_mode 2): x16 = round(x *
// T/max), decode scale max/T per snapshot and field (T = 2^15 for fp16, 1
// for bf16; BF16 snapshots are normally encoded by the storing step).
// ---------------------------------------------------------------------------
__global__ void k_snap_pack(int nxny, const float *stage, const unsigned int *stage_max, uint16_t *snap16,
                            float *snap_scale, long long shot_stride16, int nsnap, const int *nt_base, int nt,
                            int dtinv, int kind)
{
    const int s = blockIdx.z, f = blockIdx.y;
    const int slot = ((nt_base ? *nt_base : 0) + nt - 1) / dtinv;
    const float m = __uint_as_float(stage_max[s * 3 + f]);
    const float T = (kind == 1) ? 1.0f : 32768.0f;
    const float enc = (m > 0.0f) ? (T / m) : 1.0f;
    const float dec = (m > 0.0f) ? (m / T) : 1.0f;
    if (blockIdx.x == 0 && threadIdx.x == 0) snap_scale[((long long)s * nsnap + slot) * 3 + f] = dec;
    const float *src = stage + ((long long)s * 3 + f) * nxny;
    uint16_t *dst = snap16 + (long long)s * shot_stride16 + ((long long)slot * 3 + f) * nxny;
    auto code = [&](float v) -> unsigned { return (kind == 1) ? f2bf16(v * enc) : f2h(v * enc); };
    if ((nxny & 3) == 0) {
        // four values per lane: 16-byte loads and 8-byte stores (every plane
        // starts at a multiple of 4 elements of 16-byte aligned buffers)
        const float4 *s4 = reinterpret_cast<const float4 *>(src);
        uint2 *d4 = reinterpret_cast<uint2 *>(dst);
        for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < (nxny >> 2); t += gridDim.x * blockDim.x) {
            const float4 x = s4[t];
            d4[t] = make_uint2(code(x.x) | (code(x.y) << 16), code(x.z) | (code(x.w) << 16));
        }
    } else {
        for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nxny; t += gridDim.x * blockDim.x)
            dst[t] = (uint16_t)code(src[t]);
    }
}

hipError_t launch_snap_pack(int nshots, int nx, int ny, const float *stage, const unsigned int *stage_max,
                            uint16_t *snap16, float *snap_scale, long long shot_stride16, int nsnap,
                            const int *nt_base, int nt, int dtinv, int kind, hipStream_t st)
{
    const int nxny = nx * ny;
    const int per = (nxny & 3) == 0 ? nxny / 4 : nxny;   // work items per plane
    hipLaunchKernelGGL(k_snap_pack, dim3(std::min(1024, (per + 255) / 256), 3, nshots), dim3(256), 0, st, nxny,
                       stage, stage_max, snap16, snap_scale, shot_stride16, nsnap, nt_base, nt, dtinv, kind);
    return hipGetLastError();
}

__global__ void k_advance(int *p, int v)
{
    if (threadIdx.x == 0) p[0] += v;
}

hipError_t launch_advance(int *p, int v, hipStream_t st)
{
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, p, v);
    return hipGetLastError();
}

__global__ void k_fill(float *p, size_t n, float v)
{
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) p[t] = v;
}
__global__ void k_fill_u32(unsigned int *p, size_t n, unsigned int v)
{
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) p[t] = v;
}

hipError_t launch_fill(float *p, size_t n, float v, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}
hipError_t launch_fill_u32(unsigned int *p, size_t n, unsigned int v, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Split kernels behind the reference-named drop-ins: one thread per interior
// cell, in-place update exactly as the reference loop bodies.  With BOUNDARY 1
// the x neighbours beyond the edges are read from the columns they wrap to
// (what exchange_v / exchange_s copy into the reference's halo).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int xwrap(const Geo &g, int i)
{
    return !g.periodic ? i : (i < 1 ? i + g.nx : (i > g.nx ? i - g.nx : i));
}

// K = FDORDER / 2 is a template parameter so the stencil loops unroll; the
// arithmetic is the same in the same order for every K.  A workgroup takes
// 64 columns of kSplitRows rows.
constexpr int kSplitRows = 4;

template <int K>
__global__ __launch_bounds__(64 * kSplitRows) void k_split_v(SplitArgs a)
{
    const Geo &g = a.g;
    const int i = blockIdx.x * 64 + (int)(threadIdx.x % 64) + 1;
    const int j = blockIdx.y * kSplitRows + (int)(threadIdx.x / 64) + 1;
    if (i > g.nx || j > g.ny) return;
    const int fw = g.fw, fw2 = 2 * g.fw;
    float sxy_x = a.hc[1] * (a.sxy[gidx(g, j, i)] - a.sxy[gidx(g, j, xwrap(g, i - 1))]);
    float syy_y = a.hc[1] * (a.syz[gidx(g, j, i)] - a.syz[gidx(g, j - 1, i)]);
#pragma unroll
    for (int m = 2; m <= K; ++m) {
        sxy_x = sxy_x + a.hc[m] * (a.sxy[gidx(g, j, xwrap(g, i + m - 1))] - a.sxy[gidx(g, j, xwrap(g, i - m))]);
        const float hy = (K == 6 && m == 6) ? a.hcy6 : a.hc[m];
        syy_y = syy_y + hy * (a.syz[gidx(g, j + m - 1, i)] - a.syz[gidx(g, j - m, i)]);
    }
    if (fw > 0) {
        float *psx = a.psi_b, *psy = a.psi_a;
        if (!g.periodic && i <= fw) {
            float &p = psx[(long long)(j - 1) * fw2 + (i - 1)];
            p = a.pml.bx[i] * p + a.pml.ax[i] * sxy_x;
            sxy_x = sxy_x / a.pml.Kx[i] + p;
        }
        if (!g.periodic && i >= g.nx - fw + 1) {
            const int h1 = i - g.nx + fw2;
            float &p = psx[(long long)(j - 1) * fw2 + (h1 - 1)];
            p = a.pml.bx[h1] * p + a.pml.ax[h1] * sxy_x;
            sxy_x = sxy_x / a.pml.Kx[h1] + p;
        }
        if (!g.free_surf && j <= fw) {
            float &p = psy[(long long)(j - 1) * g.nx + (i - 1)];
            p = a.pml.by[j] * p + a.pml.ay[j] * syy_y;
            syy_y = syy_y / a.pml.Ky[j] + p;
        }
        if (j >= g.ny - fw + 1) {
            const int h1 = j - g.ny + fw2;
            float &p = psy[(long long)(h1 - 1) * g.nx + (i - 1)];
            p = a.pml.by[h1] * p + a.pml.ay[h1] * syy_y;
            syy_y = syy_y / a.pml.Ky[h1] + p;
        }
    }
    const long long o = gidx(g, j, i);
    const float rv = a.rho[o];
    float b = (float)(1.0 / (double)rv);
    if ((double)rv < 1e-20) b = 0.0f;
    if (a.sw == 0) a.vyp1[o] = (b * (sxy_x + syy_y)) / a.DH;
    if (a.sw == 1) a.vyp1[o] = a.vy[o];
    float vn = a.vy[o] + ((b * a.DT) * (sxy_x + syy_y)) / a.DH;
    for (int k = 0; k < a.npts; ++k)
        if (a.pt_o[k] == o) vn = vn + a.pt_v[k];
    a.vy[o] = vn;
}

template <bool VISCO, int K>
__global__ __launch_bounds__(64 * kSplitRows) void k_split_s(SplitArgs a)
{
    const Geo &g = a.g;
    const int i = blockIdx.x * 64 + (int)(threadIdx.x % 64) + 1;
    const int j = blockIdx.y * kSplitRows + (int)(threadIdx.x / 64) + 1;
    if (i > g.nx || j > g.ny) return;
    const int fw = g.fw, fw2 = 2 * g.fw;
    const float dhi = VISCO ? (float)(1.0 / (double)a.DH) : a.DT / a.DH;
    const float dthalbe = (float)((double)a.DT / 2.0);
    float vyx = a.hc[1] * (a.vy[gidx(g, j, xwrap(g, i + 1))] - a.vy[gidx(g, j, i)]);
    float vyy = a.hc[1] * (a.vy[gidx(g, j + 1, i)] - a.vy[gidx(g, j, i)]);
#pragma unroll
    for (int m = 2; m <= K; ++m) {
        vyx = vyx + a.hc[m] * (a.vy[gidx(g, j, xwrap(g, i + m))] - a.vy[gidx(g, j, xwrap(g, i - m + 1))]);
        vyy = vyy + a.hc[m] * (a.vy[gidx(g, j + m, i)] - a.vy[gidx(g, j - m + 1, i)]);
    }
    vyx = vyx * dhi;
    vyy = vyy * dhi;
    if (fw > 0) {
        float *pvy = a.psi_a, *pvx = a.psi_b;
        if (!g.periodic && i <= fw) {
            float &p = pvx[(long long)(j - 1) * fw2 + (i - 1)];
            p = a.pml.bxh[i] * p + a.pml.axh[i] * vyx;
            vyx = vyx / a.pml.Kxh[i] + p;
        }
        if (!g.periodic && i >= g.nx - fw + 1) {
            const int h1 = i - g.nx + fw2;
            float &p = pvx[(long long)(j - 1) * fw2 + (h1 - 1)];
            p = a.pml.bxh[h1] * p + a.pml.axh[h1] * vyx;
            vyx = vyx / a.pml.Kxh[h1] + p;
        }
        if (!g.free_surf && j <= fw) {
            float &p = pvy[(long long)(j - 1) * g.nx + (i - 1)];
            const float bY = VISCO ? a.pml.by[j] : a.pml.byh[j];
            const float aY = VISCO ? a.pml.ay[j] : a.pml.ayh[j];
            const float KY = VISCO ? a.pml.Ky[j] : a.pml.Kyh[j];
            p = bY * p + aY * vyy;
            vyy = vyy / KY + p;
        }
        if (j >= g.ny - fw + 1) {
            const int h1 = j - g.ny + fw2;
            float &p = pvy[(long long)(h1 - 1) * g.nx + (i - 1)];
            p = a.pml.byh[h1] * p + a.pml.ayh[h1] * vyy;
            vyy = vyy / a.pml.Kyh[h1] + p;
        }
    }
    const long long o = gidx(g, j, i);
    if (!VISCO) {
        a.uy[o] = vyy;
        a.uyx[o] = vyx;
        a.sxy[o] = a.sxy[o] + a.uip[o] * vyx;
        a.syz[o] = a.syz[o] + a.ujp[o] * vyy;
    } else {
        float sumr = 0.0f, sumq = 0.0f;
        for (int l = 0; l < g.L; ++l) {
            sumr += a.r[(long long)l * g.plane + o];
            sumq += a.q[(long long)l * g.plane + o];
        }
        a.sxy[o] = a.sxy[o] + ((a.fipjp[o] * vyx) + (dthalbe * sumr));
        a.syz[o] = a.syz[o] + ((a.fcoef[o] * vyy) + (dthalbe * sumq));
        sumr = 0.0f;
        sumq = 0.0f;
        for (int l = 0; l < g.L; ++l) {
            float &rl = a.r[(long long)l * g.plane + o];
            float &ql = a.q[(long long)l * g.plane + o];
            rl = a.bip[l] * (rl * a.cip[l] - (a.dip[(long long)l * g.plane + o] * vyx));
            ql = a.bjm[l] * (ql * a.cjm[l] - (a.dcoef[(long long)l * g.plane + o] * vyy));
            sumr += rl;
            sumq += ql;
        }
        a.uy[o] = vyy;
        a.uyx[o] = vyx;
        a.sxy[o] = a.sxy[o] + (dthalbe * sumr);
        a.syz[o] = a.syz[o] + (dthalbe * sumq);
    }
}

template <int K>
static void launch_split_k(const SplitArgs &a, int which, hipStream_t st)
{
    const dim3 grid((a.g.nx + 63) / 64, (a.g.ny + kSplitRows - 1) / kSplitRows), block(64 * kSplitRows);
    if (which == 0)
        hipLaunchKernelGGL((k_split_v<K>), grid, block, 0, st, a);
    else if (which == 1)
        hipLaunchKernelGGL((k_split_s<false, K>), grid, block, 0, st, a);
    else
        hipLaunchKernelGGL((k_split_s<true, K>), grid, block, 0, st, a);
}

static hipError_t launch_split(const SplitArgs &a, int which, hipStream_t st)
{
    switch (a.K) {
    case 1: launch_split_k<1>(a, which, st); break;
    case 2: launch_split_k<2>(a, which, st); break;
    case 3: launch_split_k<3>(a, which, st); break;
    case 4: launch_split_k<4>(a, which, st); break;
    case 5: launch_split_k<5>(a, which, st); break;
    case 6: launch_split_k<6>(a, which, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_split_v(const SplitArgs &a, hipStream_t st) { return launch_split(a, 0, st); }
hipError_t launch_split_s(const SplitArgs &a, bool visco, hipStream_t st) { return launch_split(a, visco ? 2 : 1, st); }

}  // namespace dn
