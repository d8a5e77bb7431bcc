// Internal declarations shared by the HIP kernels and the host side of
// libdenise_sh.so.  Not part of the public C-ABI (include/denise_sh.h).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace dn {

constexpr int kWave = 64;      // CDNA wavefront width: one wave per strip
constexpr int kStripsPerGroup = 8;   // waves (adjacent strips) per workgroup of the time step
constexpr int kMaxL = 4;       // relaxation mechanisms supported
// Zero frame around the interior of every padded field.  The fused step
// reads up to 2K-1 rows above/below the interior, 2K-1 columns left of it
// and 63 right of it (the last strip's wave), K <= 6.
constexpr int kPadX = 16;      // columns left of i = 1 (64-byte aligned rows)
constexpr int kPadR = 64;      // columns right of i = nx (at least)
constexpr int kPadY = 12;      // rows above j = 1 and below j = ny


// Padded, x-fastest layout of one 2-D field.  Element (j,i) of the
// reference's index space (1-based interior) lives at
// (j-1+yoff)*pitch + (i-1+xoff).  The halo is a permanent zero frame for
// wavefields (zero_fdveps.c) and holds the periodic copy for model fields
// (matcopy.c).
struct Geo {
    int nx, ny;
    int pitch;          // floats per padded row (multiple of 32: 128-byte rows)
    int xoff, yoff;     // position of i = 1 / j = 1
    int rows;           // padded rows
    long long plane;    // pitch * rows
    int fw, free_surf, L;
    int periodic;       // BOUNDARY 1: wavefields wrap across the left and right edges, no x-CPML
};

inline Geo make_geo(int nx, int ny, int fw, int free_surf, int L, int periodic = 0)
{
    Geo g{};
    g.periodic = periodic ? 1 : 0;
    g.nx = nx;
    g.ny = ny;
    g.xoff = kPadX;
    g.yoff = kPadY;
    g.pitch = (nx + kPadX + kPadR + 31) / 32 * 32;
    g.rows = ny + 2 * kPadY;
    g.plane = (long long)g.pitch * g.rows;
    g.fw = fw;
    g.free_surf = free_surf ? 1 : 0;
    g.L = L;
    return g;
}

// floats of the packed stress-coefficient planes (StepArgs::spk)
inline size_t spk_floats(size_t plane, int L) { return L > 0 ? plane * (size_t)(4 + 2 * (L - 1)) : plane * 2; }

__host__ __device__ inline long long gidx(const Geo &g, int j, int i)
{
    return (long long)(j - 1 + g.yoff) * g.pitch + (long long)(i - 1 + g.xoff);
}

// CPML profile vectors, 1-based [1..2FW] as in PML_pro.c, packed on device.
struct PmlVec {
    const float *Kx, *ax, *bx, *Kxh, *axh, *bxh, *Ky, *ay, *by, *Kyh, *ayh, *byh;
};

// Arguments of the fused time-step kernel (one launch advances every active
// shot of a batch by one time step).
struct StepArgs {
    Geo g;
    int nshots;                    // shots in this launch
    int nshots_all;                // shots of the context (bound of shot_map entries)
    int ow, nstrips, R, nsegs;     // strip width (owned columns), strips, rows per segment, segments
    int ngroups;                   // workgroups of kStripsPerGroup adjacent strips per segment row
    int nblocks, nblocks8;         // real workgroups and workgroups per XCD slice (one shot per wave)
    int nblocks2, nblocks2_8;      // the same with two shots per wave (0: not used)
    int sg;                        // shots per workgroup: 1 = kStripsPerGroup adjacent strips of one
                                   // shot; 2, 4, 8 = kStripsPerGroup / sg adjacent strips of sg
                                   // consecutive shots (one shot per wave)
    int nblocks_sg, nblocks_sg8;   // workgroups in a layout with sg > 1
    float DT, DH, dhi, dthalbe;
    float hc[7];
    float hcy6;                    // coefficient used for the 6th y-term of update_v (Q6: hc[5])
    PmlVec pml;
    // static model fields (shared by all shots), padded layout
    const float *__restrict__ bfield;   // 1/rho (0 where rho < 1e-20)
    const float *__restrict__ uip;      // elastic
    const float *__restrict__ ujp;
    const float *__restrict__ fipjp;    // visco
    const float *__restrict__ fcoef;
    const float *__restrict__ dip;      // [L][plane]
    const float *__restrict__ dcoef;    // [L][plane]
    // the same stress coefficients packed per cell, read by the fused step
    // with one wide load instead of one load per field: elastic one float2
    // plane {uip, ujp}; viscoelastic one float4 plane {fipjp, f, dip[0],
    // d[0]} followed by L-1 float2 planes {dip[l], d[l]}
    const float *__restrict__ spk;
    float bip[kMaxL], bjm[kMaxL], cip[kMaxL], cjm[kMaxL];
    // per-shot dynamic state; shot stride = plane (2*L*plane for rq)
    const float *vy_in, *sxy_in, *syz_in;
    float *vy_out, *sxy_out, *syz_out;
    // the relaxation memory variables r[L], q[L] of a shot, interleaved per
    // cell so that one wide access moves a cell's pairs: float4 planes
    // {r[l], q[l], r[l+1], q[l+1]} for l = 0, 2, ..., then a float2 plane
    // {r[L-1], q[L-1]} when L is odd
    float *rq;
    const float *psx_in;  float *psx_out;   // psi_sxy_x [s][ny][2fw] (ping-pong)
    const float *psy_in;  float *psy_out;   // psi_syy_y [s][2fw][nx] (ping-pong)
    float *pvx, *pvy;                       // psi_vyx [s][ny][2fw], psi_vyy [s][2fw][nx]
    // shot tables (absolute shot index of batch slot s: shot_map[s])
    const int *shot_map;
    const int *shot_nt;                     // [all shots] time steps of each shot
    const int *src_off, *src_i, *src_j;     // sources of shot s: src_off[s]..src_off[s+1]-1
    const float *src_sig;                   // [nsrc_total][nt_max]
    int nt_max;
    // receivers (shared by all shots)
    const int *rec_row;                     // [ny+1] -> offset into rec_tab, or -1
    const int *rec_tab;                     // rows of [nx+1]: receiver index or -1
    int ntr;
    float *seis;                            // forward: [s][ntr][nt_max]
    const float *resid;                     // adjoint: [s][ntr][nt_max] time-reversed residual
    // gradient storage
    int dtinv;
    int store_mode;                         // forward: 0 no snapshots, 1 fp32 direct, 2 fp32 staging
                                            // (fp16), 3 bf16 direct
    float *snap;                            // fp32 [s][nsnap][3][ny][nx] (store_mode 1 / adjoint fp32)
    long long snap_shot_stride;             // elements per shot
    uint16_t *snap16;                       // bf16/fp16 snapshots [s][nsnap][3][ny][nx]: written by
                                            // store_mode 3 and k_snap_pack, read by the adjoint
    float *snap_scale;                      // [s][nsnap][3] decode scale
    int snap_kind;                          // 0 fp32, 1 bf16, 2 fp16
    float *stage;                           // store_mode 2: [s][3][ny][nx]
    unsigned int *stage_max;                // max |x| as float bits; store_mode 2: [s][3];
                                            // store_mode 3: [s][3 snapshots, ring][3]
    float *wacc;                            // Ws (forward) or Wr (adjoint) [s][ny][nx]
    float *grad_u, *grad_rho;               // adjoint [s][ny][nx]; forward with hess: H_mu, H_rho sums
    int hess;                               // forward: accumulate the pseudo-Hessian (hess_save.c, hessian.c)
    int nt;                                 // current step, 1-based (offset within a graph chunk if nt_base)
    const int *nt_base;                     // graph replay: device step counter added to nt, or null
    // latched by the kernel: [0] the first step after which vy at the probe
    // cell (NY/2, NX/2) is NaN (denise.c:1183-1185 checks that cell before
    // every step; 0: none), [1] that shot + 1, [2] nonzero: a shot-map
    // entry was out of range (the batch slot was skipped)
    int *status;
};

// launchers (dn_kernels.hip)
// extra: the step may store snapshots / sum the pseudo-Hessian (forward) or
// correlate with the stored snapshots (adjoint) for some shot of the launch
hipError_t launch_step(const StepArgs &a, int K, bool visco, bool adjoint, bool extra, hipStream_t st);

struct ModelPrepArgs {
    Geo g;
    float DT;
    int L;
    float *vs, *rho, *taus;                 // padded, halo filled by prep
    float *uip, *ujp, *tausipjp, *bfield;
    float *fipjp, *fcoef, *dip, *dcoef;
    float *spk;                             // packed copy (StepArgs::spk)
    float sumu;                             // prepare_update_s.c:30-33 (host)
    float eta[kMaxL];                       // readmod.c:33 (host)
};
hipError_t launch_model_prep(const ModelPrepArgs &a, hipStream_t st);

// pseudo-Hessian: per-shot sums added in shot order, then the chain rule of
// hessian_out.c:25-31 (INVMAT1 1)
hipError_t launch_hessian_finish(int nshots, int nx, int ny, const Geo &g, const float *hu_shot,
                                 const float *hr_shot, const float *vs_pad, const float *rho_pad, float *hu,
                                 float *hr, int first, int last, hipStream_t st);

// per-shot residual (calc_res / calc_energy): one thread per trace
hipError_t launch_residual(int nshots, int nshots_all, int ntr, int nt_max, const int *shot_map, const int *shot_nt,
                           const float *syn, const float *obs, float *resid, double *part, float DT,
                           const unsigned char *flags, hipStream_t st);

// energy preconditioning + shot accumulation (eprecond1.c, denise.c:1665-1694)
// (precond 0: EPRECOND 0, plain sum; 1 and 3: EPRECOND 1 and 3).  ws_sum/wr_sum collect the energy maps
// over shots (reset when first).
hipError_t launch_eprecond_accumulate(int nshots, int nx, int ny, float DH, const float *Ws, const float *Wr,
                                      const float *gu_shot, const float *grho_shot, float *gu, float *grho,
                                      float C_vs, float C_rho, unsigned int *maxbits, int precond, float *ws_sum,
                                      float *wr_sum, int first, hipStream_t st);
// the same in two parts, for batches that do not follow the shots' order:
// the per-shot terms (Ws, Wr and the scaled gradients) of a batch stored at
// the shots' ranks, terms [rank][4][nx*ny], then their sum in rank order,
// with the arithmetic of launch_eprecond_accumulate
hipError_t launch_eprecond_terms(int nshots, int nx, int ny, float DH, const float *Ws, const float *Wr,
                                 const float *gu_shot, const float *grho_shot, float C_vs, float C_rho,
                                 unsigned int *maxbits, int precond, const int *rank, float *terms, hipStream_t st);
hipError_t launch_eprecond_sum(int nterms, int nx, int ny, const float *terms, float *gu, float *grho,
                               float *ws_sum, float *wr_sum, int first, hipStream_t st);

// finite-difference Hessian probes: Rademacher +-1 of (seed, probe, cell)
hipError_t launch_probe_perturb(const Geo &g, const float *vs0_pad, float *vs_pad, float eps, unsigned int seed,
                                int probe, hipStream_t st);
hipError_t launch_probe_accumulate(const Geo &g, const float *vs0_pad, const float *g0, const float *g1, float eps,
                                   unsigned int seed, int probe, float *diag, hipStream_t st);

// chain rule (denise.c:1756-1811)
hipError_t launch_chain_rule(const Geo &g, int nx, int ny, const float *gu, const float *grho,
                             const float *vs_pad, const float *rho_pad, float DT, float *grad_vs,
                             float *grad_rho, hipStream_t st);

// fp16 snapshot packing (store_mode 2)
// (snapshot slot = (nt_base ? *nt_base : 0) + nt - 1) / dtinv)
hipError_t launch_snap_pack(int nshots, int nx, int ny, const float *stage, const unsigned int *stage_max,
                            uint16_t *snap16, float *snap_scale, long long shot_stride16, int nsnap,
                            const int *nt_base, int nt, int dtinv, int kind, hipStream_t st);

// graph replay step counter: *p += v (one lane)
hipError_t launch_advance(int *p, int v, hipStream_t st);

hipError_t launch_fill(float *p, size_t n, float v, hipStream_t st);
hipError_t launch_fill_u32(unsigned int *p, size_t n, unsigned int v, hipStream_t st);

// split (one field update per launch) kernels used by the reference-named drop-ins
constexpr int kSplitPoints = 8;   // point additions carried by one k_split_v launch
struct SplitArgs {
    Geo g;
    float DT, DH;
    float hc[7];
    float hcy6;
    int K;
    PmlVec pml;
    int sw;                                  // update_v: 0 forward, 1 adjoint
    float *vy, *vyp1, *sxy, *syz, *uy, *uyx;
    const float *rho, *uip, *ujp, *fipjp, *fcoef, *dip, *dcoef;
    float *r, *q;
    float bip[kMaxL], bjm[kMaxL], cip[kMaxL], cjm[kMaxL];
    float *psi_a, *psi_b;                   // v: psi_syy_y [2fw][nx], psi_sxy_x [ny][2fw]; s: psi_vyy, psi_vyx
    // update_v: point additions after the sweep (update_v_PML.c:563-583),
    // made by the thread of the cell, in list order
    int npts;
    long long pt_o[kSplitPoints];           // padded cell index
    float pt_v[kSplitPoints];
};
hipError_t launch_split_v(const SplitArgs &a, hipStream_t st);
hipError_t launch_split_s(const SplitArgs &a, bool visco, hipStream_t st);

}  // namespace dn
