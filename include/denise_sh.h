/*
 * denise_sh.h -- C-ABI of the MI355X SH time-stepping and adjoint-gradient
 * library (libdenise_sh.so).
 *
 * Two layers:
 *
 *  1. Drop-in entry points with the reference's names and signatures
 *     (fd.h of DENISE-SH).  They take the reference's Numerical-Recipes
 *     offset-indexed float** arrays in host memory, run the hand-written HIP
 *     kernels on the device and leave the results in the same host arrays,
 *     so a caller written against the reference links against this library
 *     unchanged.  The reference's implicit inputs (globals DT, DH, FDORDER,
 *     FW, FREE_SURF, L, ...) are handed over once with dn_set_globals().
 *
 *  2. A handle API (dn_*) for the fast path: the model, acquisition and
 *     per-shot data stay resident in HBM, many shots advance together in one
 *     launch per time step, and only scalars and requested results cross the
 *     boundary.
 *
 * Conventions: grid indices are 1-based like the reference (i = x, j = y,
 * y pointing down).  Host 2-D arrays in the handle API are row-major
 * [ny][nx] float32 unless stated otherwise.  Functions returning int return
 * 0 on success and a negative code on error; dn_last_error() gives the text.
 */
#ifndef DENISE_SH_H
#define DENISE_SH_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * Layer 1: reference drop-ins
 * ---------------------------------------------------------------------- */

/* Snapshot of the reference globals the kernels read implicitly
 * (update_v_PML.c:23-29, update_s_visc_PML.c:24-28, globvar.h).  POS,
 * NPROCX and NPROCY are fixed to the single-domain values (0, 1, 1). */
typedef struct dn_globals {
    int NX, NY;          /* local = global grid size                       */
    int FDORDER;         /* 2..12                                          */
    int FW;              /* CPML width                                     */
    int FREE_SURF;       /* 1: no top CPML (implicit free surface)         */
    int BOUNDARY;        /* 0, or 1: periodic left and right edges (NX >= FDORDER) */
    int L;               /* relaxation mechanisms                          */
    int GRAD_FORM;       /* must be 2                                      */
    int QUELLTYPB;       /* adjoint source type (2: force in y)            */
    int NT;              /* number of time steps (psource)                 */
    float DT, DH;
} dn_globals;

/* Sets the implicit inputs of the drop-ins below; call once after the
 * parameter file is read (where the reference calls exchange_par()). */
int dn_set_globals(const dn_globals *g);

/* update_v_PML.c:12-16 (fd.h:406-411).  vym1, utty, absorb_coeff are
 * unused, as in the reference. */
void update_v_PML(int nx1, int nx2, int ny1, int ny2, int nt, float **vy, float **vyp1, float **vym1,
                  float **utty, float **syz, float **sxy, float **rho, float **srcpos_loc, float **signals1,
                  int nsrc, float **absorb_coeff, float *hc, int infoout, int sw, float *K_x, float *a_x,
                  float *b_x, float *K_x_half, float *a_x_half, float *b_x_half, float *K_y, float *a_y,
                  float *b_y, float *K_y_half, float *a_y_half, float *b_y_half, float **psi_syy_y,
                  float **psi_sxy_x);

/* update_s_elastic_PML.c:12-16 (fd.h:376-381).  u, absorb_coeff and rho
 * are unused, as in the reference. */
void update_s_elastic_PML(int nx1, int nx2, int ny1, int ny2, float **vy, float **uy, float **uyx, float **syz,
                          float **sxy, float **u, float **uip, float **ujp, float **absorb_coeff, float **rho,
                          float *hc, int infoout, float *K_x, float *a_x, float *b_x, float *K_x_half,
                          float *a_x_half, float *b_x_half, float *K_y, float *a_y, float *b_y, float *K_y_half,
                          float *a_y_half, float *b_y_half, float **psi_vyy, float **psi_vyx);

/* update_s_visc_PML.c:12-17 (fd.h:386-392).  r, q, dip, d are the
 * reference's f3tensors [j][i][1..L]; ujp, uip, rho, p, g, e are unused, as
 * in the reference. */
void update_s_visc_PML(int nx1, int nx2, int ny1, int ny2, float **vy, float **uy, float **uyx, float **syz,
                       float **sxy, float **ujp, float **uip, float **rho, float *hc, int infoout, float ***r,
                       float ***p, float ***q, float **fipjp, float **f, float **g, float *bip, float *bjm,
                       float *cip, float *cjm, float ***d, float ***e, float ***dip, float *K_x, float *a_x,
                       float *b_x, float *K_x_half, float *a_x_half, float *b_x_half, float *K_y, float *a_y,
                       float *b_y, float *K_y_half, float *a_y_half, float *b_y_half, float **psi_vyy,
                       float **psi_vyx);

/* seismo_ssg.c:9 (fd.h:296-297): sectionvy[itr][lsamp] = vy[recpos[2][itr]][recpos[1][itr]]. */
void seismo_ssg(int lsamp, int ntr, int **recpos, float **sectionvy, float **vy, float **syy, float **u,
                float *hc);

/* psource.c:9-10 (no prototype in fd.h): explosive source, sxx,syy +=
 * centred time derivative of the signal (one-sided at nt = 1 and NT). */
void psource(int nt, float **sxx, float **syy, float **srcpos_loc, float **signals, int nsrc);

/* zero_fdveps.c:11-13 / zero_fdveps_visc.c:11-14 (fd.h:447-454): zero the
 * wavefields, the CPML memory and (visco) the memory variables as the
 * reference does, and, in resident mode, their device copies too. */
void zero_fdveps(int ny1, int ny2, int nx1, int nx2, float **vy, float **syy, float **sxy, float **vym1, float **vyp1,
                 float **psi_sxy_x, float **psi_vyx, float **psi_syy_y, float **psi_vyy);
void zero_fdveps_visc(int ny1, int ny2, int nx1, int nx2, float **vy, float **syy, float **sxy, float **vym1,
                      float **vyp1, float **psi_sxy_x, float **psi_vyx, float **psi_syy_y, float **psi_vyy,
                      float ***pr, float ***pp, float ***pq);

/* Residency of the drop-ins' arrays (no counterpart in the reference; it
 * replaces the reference's per-step host arrays with device-resident ones).
 * mode 0 (default): every drop-in call copies what it reads to the device
 * and what it writes back, so the host arrays are current after each call.
 * mode 1: each array a drop-in receives is copied to HBM when its host
 * pointer is first seen and then stays there; the drop-ins run on the device
 * copies without host transfers, and the host arrays go stale.  seismo_ssg
 * then reads the receiver samples back from the device (the one sync per
 * step) and also refreshes vy[NY/2][NX/2], the cell the reference's NaN
 * probe reads (denise.c:1183).  mode 2: as 1, but seismo_ssg keeps the
 * samples on the device until dn_dropin_sync(1) writes them into sectionvy
 * (no sync per step).  Switching to 0 copies the written arrays back and
 * frees the device copies. */
int dn_dropin_resident(int mode);
/* to_host 1: wait for the device, copy every resident array the kernels
 * write (wavefields, memory variables, CPML memory) and the deferred
 * seismograms back into the host arrays, e.g. before the host reads fields
 * (the snapshot loop of denise.c:1262-1306).  to_host 0: re-read every
 * resident array from the host, e.g. after the host changed the model
 * (prepare_update_s after a model update) or a CPML vector. */
int dn_dropin_sync(int to_host);

/* ------------------------------------------------------------------------
 * Host-side setup routines of the path (no device needed).  The reference
 * versions read globals; these take the same quantities as arguments.
 * ---------------------------------------------------------------------- */

/* holbergcoeff.c:19-90: hc[0..6] for FDORDER and MAXRELERROR. */
int dn_holbergcoeff(int fdorder, int maxrelerror, float *hc);

/* PML_pro.c:65-297: twelve 1-based vectors of length 2*fw+1 written to
 * out[12][2*fw+1] in the order K_x, a_x, b_x, K_x_half, a_x_half, b_x_half,
 * K_y, a_y, b_y, K_y_half, a_y_half, b_y_half. */
int dn_pml_pro(int nx, int ny, int fw, float dh, float dt, float damping, float fpml, float npower,
               float k_max_pml, float *out);

/* readmod.c:30-34 + prepare_update_s.c:20-56: eta[L], sumu, bip, bjm,
 * cip, cjm [L] for relaxation frequencies fl[L]. */
int dn_relaxation_constants(int L, const float *fl, float dt, float *eta, float *sumu, float *bip, float *bjm,
                            float *cip, float *cjm);

/* wavelet.c:12-96 for QUELLART 1, 2, 4, 5, 6: out[nt-1] = sample nt. */
int dn_wavelet(int quellart, int nt, float dt, float tshift, float fc, float amp, float *out);

/* timedomain_filt.c:10-53 (Butterworth, libcseife) on data [ntr][ns] in
 * place; method 1 lowpass, 2 highpass. */
int dn_timedomain_filt(float *data, int ntr, int ns, float dt, float fc, int order, int method);

/* checkfd_ssg_visc.c:14-142 (L > 0) / checkfd_ssg_elastic.c:14-130 (L = 0),
 * INVMAT1 = 1: vs, rho, qs host [ny][nx]; ts_src = TS (source duration).
 * Returns 1 where the reference aborts (DT > dtstab or FW > NX, NY),
 * 0 otherwise, -1 on bad arguments. */
typedef struct {
    float cmin, cmax;     /* S-wave phase velocity at f = 0 / f = inf */
    float dtstab, dhstab; /* stability limit for DT, recommended DH */
    int dt_unstable, dh_dispersive, fw_too_large, fw_small;
} dn_checkfd_result;
int dn_checkfd(int nx, int ny, int fdorder, int maxrelerror, int L, const float *fl, float dt, float dh, float ts_src,
               int fw, const float *vs, const float *rho, const float *qs, dn_checkfd_result *out);

/* ------------------------------------------------------------------------
 * Layer 2: handle API (fast path).  Not in the reference: it replaces the
 * per-step call sequence of denise.c:1179-1346 / 1493-1651 and the shot
 * loop denise.c:875-1702 with resident, batched execution.
 * ---------------------------------------------------------------------- */

typedef struct dn_ctx dn_ctx;
typedef struct dn_unet dn_unet;   /* learned preconditioner network (below) */

typedef struct dn_params {
    int nx, ny;          /* NX, NY                                         */
    int fdorder;         /* FDORDER 2..12                                  */
    int maxrelerror;     /* MAXRELERROR 0..4                               */
    int fw;              /* FW                                             */
    int free_surf;       /* FREE_SURF                                      */
    int L;               /* relaxation mechanisms (0 = elastic), <= 4      */
    float fl[4];         /* FL[1..L]                                       */
    float dt, dh;        /* DT, DH                                         */
    float damping, fpml; /* DAMPING, FPML                                  */
    float npower, k_max_pml;
    int dtinv;           /* DTINV: snapshot every dtinv-th step            */
    int snap_precision;  /* 0: fp32; 1: bf16, scale 2^-e per snapshot and field with e
                            the exponent of the previous snapshot's max |x|;
                            2: fp16, scale 2^15/max |x| of the snapshot itself */
    int eprecond;        /* EPRECOND: 1 energy preconditioning sqrt(Ws Wr), 3 sqrt(Ws)
                            times the Plessix-Mulder receiver approximation
                            (eprecond1.c:15-50), 0 none */
    int boundary;        /* BOUNDARY: 0 absorbing (CPML) left and right; 1 periodic:
                            vy and sxy wrap across the left and right edges
                            after each half step (exchange_v.c:99-161,
                            exchange_s.c:95-157 on one domain) and the x-CPML
                            is off (the !BOUNDARY guards of update_v_PML.c:63,71
                            and update_s_*_PML.c); needs NX >= 64 */
} dn_params;

/* Context on the current HIP device; NULL on failure (dn_last_error(NULL)).
 * The fused step addresses one shot's fields with 32-bit byte offsets: one
 * padded plane ((NX+80 rounded up to 32) x (NY+24) floats) times 16 + 8L
 * bytes must stay below 2 GiB (about 7 500 x 7 500 cells at L = 2). */
dn_ctx *dn_create(const dn_params *p);
void dn_destroy(dn_ctx *ctx);
/* Error text of the last failing call on ctx (or of dn_create when NULL). */
const char *dn_last_error(const dn_ctx *ctx);

/* Model: vs, rho and (visco) qs, host [ny][nx].  Runs matcopy (periodic
 * halo), av_mue, av_tau and prepare_update_s on the device
 * (denise.c:726-743).  The gradient normalisation means C_vs and C_rho are
 * taken from the model of the first gradient of the context (and again at
 * dn_fwi_start) and kept: the reference computes them in its first FWI
 * iteration only (denise.c:746-819). */
int dn_set_model(dn_ctx *ctx, const float *vs, const float *rho, const float *qs);
int dn_get_model(dn_ctx *ctx, float *vs, float *rho);
/* The means C_vs, C_rho (the globals of denise.c:746-819 that the gradient
 * normalisation and gauss_filt.c read); returns 1 while they are not taken
 * yet (no gradient and no dn_fwi_start so far), 0 after. */
int dn_get_means(const dn_ctx *ctx, float *c_vs, float *c_rho);

/* Receivers: grid indices (1-based), duplicates must already be removed
 * (receiver.c:50-63 does this on the host). */
int dn_set_receivers(dn_ctx *ctx, int ntr, const int *rec_i, const int *rec_j);

/* Shots: nshots shots with one point force each (QUELLTYP 3) at grid
 * (src_i, src_j); signals host [nshots][nt_max] (already filtered); nt[s]
 * the number of time steps of shot s (<= nt_max). */
int dn_set_shots(dn_ctx *ctx, int nshots, int nt_max, const int *nt, const int *src_i, const int *src_j,
                 const float *signals);

/* Shots with several simultaneous point forces (RUN_MULTIPLE_SHOTS = 0 or
 * several sources per shot): the sources of shot s are src_off[s] ..
 * src_off[s+1]-1, signals host [src_off[nshots]][nt_max]. */
int dn_set_shots_multi(dn_ctx *ctx, int nshots, int nt_max, const int *nt, const int *src_off, const int *src_i,
                       const int *src_j, const float *signals);

/* Replaces the source signals (same layout as the last dn_set_shots*),
 * e.g. after the per-stage low-pass filter of the workflow. */
int dn_set_signals(dn_ctx *ctx, const float *signals);

/* Observed data host [nshots][ntr][nt_max] (samples 1..nt[s] used). */
int dn_set_observed(dn_ctx *ctx, const float *obs);

/* Frequency filter of a workflow stage (denise.c:1099-1111 on the wavelets,
 * 1381-1394 and line_search.c:138-145, 258-270 on the observed data):
 * TIME_FILT 1 low-passes every source signal and observed trace at FC with
 * timedomain_filt.c (Butterworth of ORDER, libcseife), TIME_FILT 2 then
 * high-passes at FC_START as well, 0 leaves them unfiltered.  The filter is
 * applied to the signals and data as last given to dn_set_shots /
 * dn_set_signals / dn_set_observed (the context keeps host copies), never
 * on top of an earlier stage's filter, each shot's traces over its own nt
 * samples.  dn_fwi_start calls it with the stage's values. */
int dn_set_time_filt(dn_ctx *ctx, int time_filt, float fc_start, float fc, int order);

/* Maximum number of shots advanced together (0 = automatic from free HBM)
 * and rows per workgroup segment of the time-step kernel (0 = automatic). */
int dn_set_batch(dn_ctx *ctx, int max_shots_per_batch, int rows_per_segment);
/* by_record_length 1 (default): a gradient whose shots have uneven record
 * lengths and take several batches forms the batches by decreasing length
 * (shots of similar length advance together, fewer launches carry finished
 * shots); the sums over shots keep the shots' order, so the result is the
 * same bits as with 0 (batches in shot order).  Not in the reference, whose
 * shot loop is sequential (denise.c:875). */
int dn_set_batch_order(dn_ctx *ctx, int by_record_length);

/* Shots one batch advances together (gradient != 0: with snapshot storage),
 * sized from free HBM and the dn_set_batch cap; the batch workspace is
 * allocated by this call.  The reference runs one shot at a time per rank
 * (denise.c:875); no counterpart. */
int dn_batch_shots(dn_ctx *ctx, int gradient, int *nshots);

/* Source- and receiver-side energy maps of the last gradient, summed over
 * its shots: sum_s Ws, sum_s Wr (eprecond.c), host [ny][nx]. */
int dn_get_energy_maps(dn_ctx *ctx, float *ws, float *wr);

/* Diagonal of the misfit Hessian w.r.t. vs estimated with nprobes
 * finite-difference Hessian-vector probes (the idea of the reference's
 * unused hessian_FD.c): diag ~ mean_k v_k * (g(m + eps*vs*v_k) - g(m)) /
 * (eps*vs), Rademacher v_k from seed, g = the vs gradient of shots
 * [shot0, shot0+nshots) with the context's EPRECOND.  The model is restored
 * afterwards.  Output host [ny][nx]. */
int dn_hessian_probe(dn_ctx *ctx, int shot0, int nshots, int nprobes, float eps, unsigned int seed, float *diag_vs);

/* Pseudo-Hessian diagonal (HESSIAN = 1, hess_save.c:2-17, hessian.c:91-94,
 * hessian_out.c:25-31): forward modelling of shots [shot0, shot0+nshots)
 * with the signals set (the reference uses the QUELLART 6 spike low-passed
 * at FC_HESSIAN), summing uy^2 + (uyx/2)^2 and vyp1^2 over stored steps
 * 1..NTDTINV-1 and applying the chain rule to vs and rho.  Outputs host
 * [ny][nx]; either may be NULL. */
int dn_pseudo_hessian(dn_ctx *ctx, int shot0, int nshots, float *h_vs, float *h_rho);

/* Time-loop execution: graph_steps = 0 launches every step from the host;
 * > 0 captures chunks of about that many steps once per batch shape
 * (hipGraph) and replays them; -1 (default) uses 64. */
int dn_set_exec(dn_ctx *ctx, int graph_steps);
/* Shots per wavefront on the steps without snapshot / correlation work:
 * 1 (default) one shot per wave; 2 advances two shots of a batch in one
 * wave, sharing the loads of the static model fields (more registers per
 * wave).  Results are identical; dn_autotune picks the faster. */
int dn_set_wave_shots(dn_ctx *ctx, int shots_per_wave);
/* Workgroup layout of the time-step kernel (one shot per wave): 0 (default)
 * eight adjacent strips of one shot; 2 or 4: four or two adjacent strips of
 * 2 or 4 consecutive shots; 1 or 8: one strip of eight consecutive shots.
 * With several shots per workgroup, their waves load the same static model
 * rows in the same row iteration, and those loads meet in the compute
 * unit's cache.  Results are identical; dn_autotune picks the fastest.  No
 * counterpart in the reference (a launch-shape choice). */
int dn_set_group_layout(dn_ctx *ctx, int mode);

/* Tile-shape autotuner: on the batch the gradient runs (as many of the
 * first nshots shots as share a batch with their snapshots stored), times
 * nsteps forward steps for each candidate rows-per-segment (0: the cost
 * model's choice, then fixed heights), then one or two shots per wave, the
 * workgroup layouts (at the cost model's, half, equal and double the rows)
 * and graph replay (kept unless eager launches are more than 2 % faster),
 * each the faster of two timed passes; keeps the fastest setting and
 * returns rows (0: the cost model's) and graph use. */
int dn_autotune(dn_ctx *ctx, int nshots, int nsteps, int *rows_per_segment, int *graph_steps);
/* The launch shape in use (set by dn_set_batch / dn_set_exec /
 * dn_set_wave_shots / dn_set_group_layout or chosen by dn_autotune): rows per
 * segment (0: the cost model's choice per batch), graph steps, shots per wave
 * and shots per workgroup (1: eight adjacent strips of one shot).  No
 * counterpart in the reference. */
int dn_get_launch_shape(const dn_ctx *ctx, int *rows_per_segment, int *graph_steps, int *wave_shots,
                        int *group_shots);

/* Forward modelling of shots [shot0, shot0+nshots): seismograms to seis
 * (host [nshots][ntr][nt_max]) when non-NULL. */
int dn_forward(dn_ctx *ctx, int shot0, int nshots, float *seis);

/* Misfit gradient over shots [shot0, shot0+nshots) (denise.c:875-1694 and
 * the chain rule 1743-1813): forward with snapshot storage, residual
 * (calc_res LNORM 2), adjoint with correlation, energy preconditioning.
 * Accumulates into the context's gradient when accumulate != 0, else
 * starts from zero.  l2/energy receive the FP64 misfit sums over these
 * shots. */
int dn_gradient(dn_ctx *ctx, int shot0, int nshots, int accumulate, double *l2, double *energy);

/* Gradient after the chain rule (waveconv_u -> vs, waveconv_rho -> rho),
 * host [ny][nx].  raw_u/raw_rho (optional) receive the preconditioned
 * shot sums before the chain rule. */
int dn_get_gradient(dn_ctx *ctx, float *grad_vs, float *grad_rho, float *raw_u, float *raw_rho);

/* Debug/parity access: wavefields of one shot after forward modelling of
 * nsteps steps (vy, sxy, syz host [ny][nx]). */
int dn_forward_fields(dn_ctx *ctx, int shot, int nsteps, float *vy, float *sxy, float *syz);

/* Timing of the last dn_forward/dn_gradient: device milliseconds of the
 * time-step kernels and number of kernel launches. */
int dn_last_timing(dn_ctx *ctx, double *ms_steps, long long *launches, long long *cell_updates);

/* Same, split by pass: adjoint = 0 forward passes, 1 adjoint passes.  ms is
 * measured with HIP events on the library's stream around each pass's
 * time-step launches; launches counts fused step launches. */
int dn_pass_timing(dn_ctx *ctx, int adjoint, double *ms, long long *launches, long long *cell_updates);

/* ------------------------------------------------------------------------
 * Drop-in file formats (host side)
 * ---------------------------------------------------------------------- */
typedef struct dn_par dn_par;
typedef struct dn_workflow dn_workflow;

/* read_par.c / read_par_json.c: a ".json" file name selects the JSON reader
 * (denise.c:186-191), anything else the positional .inp reader.  Values are
 * looked up by the reference's variable names (REFRECX/REFRECY for REFREC,
 * FL as a list of L floats). */
dn_par *dn_par_read(const char *path);
void dn_par_free(dn_par *par);
int dn_par_get_int(const dn_par *par, const char *key, int *out);
int dn_par_get_float(const dn_par *par, const char *key, float *out);
int dn_par_get_floats(const dn_par *par, const char *key, float *out, int max);   /* returns the count */
int dn_par_get_string(const dn_par *par, const char *key, char *buf, int n);

/* read_par_inv.c: workflow table; columns by the reference's variable names
 * (PRO, TIME_FILT, FC_START, FC_END, ORDER, ..., SCALERHO). */
dn_workflow *dn_workflow_read(const char *path);
void dn_workflow_free(dn_workflow *w);
int dn_workflow_nstages(const dn_workflow *w);
int dn_workflow_get(const dn_workflow *w, int stage, const char *column, float *out);

/* sources.c:36-77: srcpos [nsrc][8] = x, y, 0, tshift, fc, amp, azimuth, type
 * (file columns x z y tshift fc amp azimuth type). */
int dn_read_sources(const char *path, int max, float *srcpos, int *nsrc);
/* receiver.c:28-74: grid indices, duplicates removed. */
int dn_read_receivers(const char *path, float dh, float refrec_x, float refrec_y, int max, int *rec_i, int *rec_j,
                      int *ntr);
/* readmod.c / writemod.c: x-major float32 files <-> row-major [ny][nx]. */
int dn_read_model_file(const char *path, int nx, int ny, float *out);
int dn_write_model_file(const char *path, int nx, int ny, const float *in);
int dn_read_model(const char *mfile, int nx, int ny, int L, float *vs, float *rho, float *qs);
/* outseis_glob.c (SEIS_FORMAT 1) / inseis.c: SU traces, data [ntr][ns]. */
int dn_write_su(const char *path, int ntr, int ns, const float *data, const int *rec_i, const int *rec_j, float dh,
                float refrec_x, float refrec_y, float src_x, float src_y, float dt, int ndt);
int dn_read_su(const char *path, int ntr, int ns, float *out);

/* ------------------------------------------------------------------------
 * Data conditioning of the residual (calc_res.c:37-113, calc_energy.c:23-49)
 * and of observed data (time_window.c).
 * ---------------------------------------------------------------------- */
typedef struct dn_datacond {
    int trkill;            /* TRKILL: 1 = leave out the traces the kill list marks        */
    const int *kill;       /* TRKILL_FILE contents in file order (kill_nsrc x kill_ntr)   */
    int kill_nsrc, kill_ntr;   /* its dimensions: the reference's nsrc_glob, ntr_glob     */
    int offset_mute;       /* OFFSET_MUTE: 1 mute offsets >= OFFSETC, 2 offsets <= OFFSETC */
    float offsetc;         /* OFFSETC [m]                                                 */
    const float *src_x, *src_y;   /* per shot source position [m] (srcpos[1], srcpos[2]);
                                     NULL: the shot's first source grid index times DH    */
} dn_datacond;
/* Applies to every later residual of the context (gradients, misfits, line
 * search): a killed or muted trace has a zero residual and no L2 share
 * (calc_res.c); a killed trace has no energy share, where the kill list is
 * read as calc_energy.c reads it, with receiver and shot indices swapped
 * (calc_energy.c:23-33 vs calc_res.c:37-47).  The reference's offset uses
 * srcpos in metres and recpos * DH.  NULL switches conditioning off.
 * Call after dn_set_receivers / dn_set_shots (they reset it). */
int dn_set_datacond(dn_ctx *ctx, const dn_datacond *dc);
/* time_window.c on data [ntr][ns] in place: samples at t = j DT >= pick are
 * multiplied by exp(-GAMMA (t - pick)^2).  TIMEWIN 2: pick = DT for every
 * trace; TIMEWIN 1: pick = picks[i] + TWLENGTH_PLUS (the reference reads the
 * picks from a file whose name it never sets, time_window.c:17-33; here the
 * caller passes them). */
int dn_time_window(int ntr, int ns, float dt, int timewin, float gamma, float twlength_plus, const float *picks,
                   float *data);

/* calc_res.c:150-257 (LNORM 2, GRAD_FORM 2) and calc_energy.c:53-61 on one
 * shot's host seismograms syn, obs [ntr][ns]: diff [ntr][ns] receives the
 * time-reversed residual syn - obs, *l2 the sum of its squares and *energy
 * the energy of the time-integrated observed traces (sum over samples of
 * (DT * running sum of obs)^2), both summed in FP64 per trace and then in
 * trace order.  As in the reference, obs[i][0] is replaced by syn[i][0]
 * first (on return obs holds that substitution). */
int dn_calc_res(int ntr, int ns, float dt, const float *syn, float *obs, float *diff, double *l2, double *energy);
/* No counterpart in the reference: dn_calc_res for callers that cannot
 * allocate device memory (code under stream capture, serving loops, harnesses
 * that forbid hipMalloc).  syn, obs, diff [ntr][ns] and l2_energy[2] (l2,
 * energy) are device arrays; obs is read only (the obs[i][0] substitution is
 * applied in the sums, not written back).  Work is queued on `stream` (a
 * hipStream_t; NULL: the null stream) and the call returns without waiting.
 * The per-trace partial sums live in a fixed device array of the library,
 * one per GPU, kept until the process ends: ntr up to
 * DN_CALC_RES_NOALLOC_MAX_TRACES (1 MiB); larger ntr is rejected.  Calls
 * that overlap on one GPU (other streams or threads) share that array and
 * race. */
#define DN_CALC_RES_NOALLOC_MAX_TRACES 65536
int dn_calc_res_noalloc(int ntr, int ns, float dt, const float *syn, const float *obs, float *diff, double *l2_energy,
                        void *stream);

/* stf.c:11-250: source time function of one shot by Wiener deconvolution of
 * the observed by the modelled traces, applied to the current signal:
 *   STF = Re IFFT( S sum_i D_i conj(M_i) / (sum_i |M_i|^2 + EPS_STF ntr Ebar) )
 * on Npad = 2^(ceil(log2 ns)+2) points; killed / muted traces left out;
 * NORMALIZE 1 / 2 trace scaling (util.c:489-579).  FFT and sums on the
 * device (FP64; the reference sums Ebar in float). */
typedef struct dn_stf_opts {
    float eps_stf;        /* EPS_STF                                              */
    float offsetc_stf;    /* OFFSETC_STF [m]; <= 0: no STF offset mute            */
    int normalize;        /* NORMALIZE: 0, 1 (each trace to its max), 2 (observed
                             trace to the modelled trace's max)                   */
} dn_stf_opts;
/* On the context: the shot's observed data and its synthetics modelled with
 * its current signal, the context's trace kill and offset mute plus
 * OFFSETC_STF; stf receives the shot's nt samples.  The caller installs the
 * result as the shot's signal (dn_set_signals), as denise.c:1063-1066 does
 * by switching to QUELLART 3. */
int dn_stf_shot(dn_ctx *ctx, int shot, const dn_stf_opts *opts, float *stf);
/* The same on host arrays: obs, syn [ntr][ns], signal [ns], mute [ntr]
 * (nonzero: trace left out) or NULL. */
int dn_stf(int ntr, int ns, const float *obs, const float *syn, const float *signal, const unsigned char *mute,
           const dn_stf_opts *opts, float *stf);

/* ------------------------------------------------------------------------
 * Gradient post-processing after the chain rule (denise.c:1818-1847).
 * Arrays are [NY][NX] row-major.
 * ---------------------------------------------------------------------- */
/* taper_grad.c:11-584 parameters.  sws 1: vertical window over columns
 * GRADT1 < i < GRADT4 with Gaussian flanks, source and receiver cells zeroed;
 * 2: zero above row GRADT1, Gaussian ramp to GRADT2, j*DH below; 3: erf
 * (SRTSHAPE 1) or log (2) bowl of radius SRTRADIUS around every source, the
 * (2 FILTSIZE + 1)^2 box at each source zeroed; 5/6: the coefficients of
 * taper.bin given as map (5 for vs, 6 for rho in denise.c:1832, 1846). */
typedef struct dn_taper {
    int sws;
    int gradt[4];         /* GRADT1..GRADT4 (grid indices)                         */
    int srtshape;         /* SRTSHAPE                                              */
    float srtradius;      /* SRTRADIUS [m]                                         */
    int filtsize;         /* FILTSIZE [cells]                                      */
    float dh;             /* DH [m]; the handle API uses the context's DH           */
    const float *map;     /* sws 5/6: [NY][NX] taper coefficients                   */
} dn_taper;

/* smooth_grad.c:10-35 on the context's gradients (after dn_gradient):
 * SPATFILTER 4 = gauss_filt_var.c:10-215, the spatially variable Gaussian
 * whose half widths follow the local wavelength of the context's vs model
 * (WD_DAMP1 along x, WD_DAMP along y, times vs / FC_END); SPATFILTER 3 =
 * gauss_filt.c:56-160, one Gaussian of FILT_SIZE_GRAD = round(WD_DAMP C_vs /
 * FC_END / DH) cells (made odd; 0: nothing; sizes below 5 or negative are
 * rejected: the reference's sigma is 0 or its kernel not normalised; its
 * uninitialised weight sum is taken to start at zero); SPATFILTER 2 =
 * smooth2.c:28-165, tridiagonal smoothing with r = WD_DAMP^2 / 4 over the
 * reference's index ranges (its zero row and column 0 take part, the last
 * row and column are kept); 0 = none.  SPATFILTER 1 (FFTW wavenumber
 * damping) is not available (error).  Bit-identical to the reference's
 * float loops.  ms (may be NULL) receives the device time. */
int dn_smooth_grad(dn_ctx *ctx, int spatfilter, float wd_damp, float wd_damp1, float fc_end, float *ms);
/* taper_grad.c on the context's gradients: which 1 vs, 2 rho, 3 both;
 * sources and receivers are the context's. */
int dn_taper_grad(dn_ctx *ctx, int which, const dn_taper *taper);
/* The same two on host arrays, without a context (gauss_filt_var(waveconv,
 * vel_mod) and taper_grad(waveconv, taper_coeff, srcpos, nshots, recpos, ntr,
 * iter, sws) of fd.h with grid-index sources): grad is filtered in place;
 * taper_out (may be NULL) receives the coefficients. */
int dn_gauss_filt_var(int nx, int ny, float dh, float wd_damp, float wd_damp1, float fc_end, const float *vel,
                      float *grad);
/* gauss_filt(waveconv) of fd.h:133 (gauss_filt.c:56-160, SPATFILTER 3) on a
 * host array, with the globals DH, WD_DAMP, FC_END and C_vs as arguments. */
int dn_gauss_filt(int nx, int ny, float dh, float wd_damp, float fc_end, float c_vs, float *grad);
/* smooth2(grad) of fd.h:277 (smooth2.c:28-165, SPATFILTER 2) on a host
 * array, with the global WD_DAMP as an argument (nx >= 2, ny >= 3). */
int dn_smooth2(int nx, int ny, float wd_damp, float *grad);
int dn_taper_grad_host(int nx, int ny, const dn_taper *taper, int nsrc, const int *src_i, const int *src_j, int nrec,
                       const int *rec_i, const int *rec_j, float *grad, float *taper_out);

/* ------------------------------------------------------------------------
 * FWI iteration on the device (denise.c:695-2018, line_search.c,
 * calc_mat_change_test.c, calc_opt_step.c, LBFGS.c, PCG.c and the
 * store/extract helpers).  Dot products and misfit sums are FP64; the
 * L-BFGS history is kept in HBM.
 * ---------------------------------------------------------------------- */
typedef struct dn_fwi_opts {
    int grad_method;      /* GRAD_METHOD: 1 PCG (Polak-Ribiere), 2 L-BFGS          */
    int nlbfgs;           /* L-BFGS ring size (the reference fixes 20)             */
    float eps_scale;      /* EPS_SCALE: first relative step length                 */
    int stepmax;          /* STEPMAX                                               */
    float scalefac;       /* SCALEFAC                                              */
    int testshot_start, testshot_end, testshot_incr;   /* TESTSHOT_*, 1-based      */
    float vs_lower, vs_upper, rho_lower, rho_upper;    /* VS/RHO LOWER/UPPERLIM    */
    int inv_vs_iter, inv_rho_iter;                     /* INV_VS_ITER, INV_RHO_ITER */
    float scalerho;       /* SCALERHO                                              */
    int min_iter;         /* MIN_ITER                                              */
    float pro;            /* PRO: relative misfit change that ends a stage          */
    /* gradient post-processing (denise.c:1818-1847), applied to the summed
     * gradient before the optimiser */
    int spatfilter;       /* SPATFILTER of the stage: 0 none, 4 gauss_filt_var     */
    float wd_damp, wd_damp1, fc_end;   /* WD_DAMP (y), WD_DAMP1 (x), FC_END           */
    int taper_vert, taper_hor, taper_sources, taper_file;   /* SWS_TAPER_GRAD_*, SWS_TAPER_FILE */
    dn_taper taper;       /* GRADT1..4, SRTSHAPE, SRTRADIUS, FILTSIZE, map (sws ignored) */
    /* frequency filter of the stage (dn_set_time_filt): low-pass at fc_end */
    int time_filt;        /* TIME_FILT: 0 none, 1 low-pass, 2 band-pass            */
    float fc_start;       /* FC_START (FC_low of the workflow): high-pass corner    */
    int order;            /* ORDER of the Butterworth filters                      */
} dn_fwi_opts;

/* One line of the reference's misfit log (denise.c:1929-1933) plus extras. */
typedef struct dn_fwi_record {
    int stage, iter, iter_true;
    float opteps_vs;      /* applied relative step (0 when no update)              */
    float epst1[3];       /* trial steps epst1[1..3]                               */
    float L2t[4];         /* normalised misfits L2t[1..4]                          */
    float M2_vs, M2_rho;  /* model misfit against the starting model               */
    float eps_u, eps_rho; /* absolute steps EPSILON_u, EPSILON_rho of the update   */
    double l2, energy;    /* FP64 misfit and energy of the test shots              */
    int step1, step3;     /* line-search flags                                     */
    int test_forwards;    /* shot simulations spent in the line search             */
    int next_stage;       /* 1: the abort criterion ended the stage                */
    int precond;          /* direction taken: 0 eprecond (plain), 1 learned         */
    /* learned mode: 1 when the plain direction was line-searched as well (the
     * judge, or the learned line search found no decrease), and the relative
     * misfit drops (L2t[1] - min(L2t[2], L2t[3])) / L2t[1] that the two line
     * searches saw (0: no decrease found, -1: not run) */
    int judged;
    float drop_learned, drop_plain;
} dn_fwi_record;

/* Starts FWI on the current model, or a new stage (optimiser reset) when
 * called again.  Needs model, receivers, shots and observed data. */
int dn_fwi_start(dn_ctx *ctx, const dn_fwi_opts *opts);
/* One iteration: gradient over all shots, direction, line search on the
 * test shots, model update; fills rec. */
int dn_fwi_iteration(dn_ctx *ctx, dn_fwi_record *rec);
/* Files of an FWI iteration under the reference's names and formats:
 *   misfit_log: one line per iteration (denise.c:1929-1933), the file
 *     created at iteration 1 over all stages and appended to after;
 *   <jacobian>_grad_vs.bin: the vs search direction (denise.c:1889-1891);
 *   <inv_modelfile>_vs.bin / _rho.bin: the updated model
 *     (calc_mat_change_test.c:191-220), with write_it also
 *     <inv_modelfile>_vs_it_<iter>.bin / _rho_it_<iter>.bin (at iter ==
 *     nfstart, calc_mat_change_test.c:221-241);
 *   after the last iteration of a stage (rec->next_stage)
 *     <inv_modelfile>_vs_stage_<stage>.bin / _rho_stage_<stage>.bin
 *     (model_freq_out.c).
 * Model files are x-major float32 (writemod.c + mergemod.c on one domain).
 * NULL names skip a file. */
typedef struct dn_fwi_files {
    const char *misfit_log;   /* MISFIT_LOG_FILE                                    */
    int time_filt;            /* TIME_FILT: the log line carries FC_START and FC    */
    float fc_start, fc;       /* FC_START, FC (FC_END of the stage)                 */
    const char *jacobian;     /* JACOBIAN                                           */
    const char *inv_modelfile;/* INV_MODELFILE                                      */
    int write_it;             /* also the per-iteration model files                 */
    int lbfgs_old;            /* L-BFGS: also the files of extract_LBFGS_SH.c:68-178,
                                 <JACOBIAN>_p_vs.old, _p_u.old, _c_u.old, _p_mrho.old,
                                 _p_rho.old, _c_rho.old (model the iteration started
                                 from, normalised gradient there, direction), each as
                                 the rank file <name>.0.0 and the merged <name>        */
} dn_fwi_files;
int dn_fwi_write_files(dn_ctx *ctx, const dn_fwi_record *rec, const dn_fwi_files *files);

/* The FWI optimiser state on disk (no counterpart in the reference, which
 * keeps the L-BFGS ring and the PCG vectors in memory and so cannot resume):
 * dn_fwi_save_state writes iteration, stage and abort counters, the misfit
 * history, C_vs / C_rho, current and starting model, previous model and
 * normalised gradient, the PCG vectors and the whole L-BFGS ring, and for
 * the learned preconditioner the stage's fixed diagonal D and whether the
 * stage reverted to plain L-BFGS (the network itself is the caller's).
 * dn_fwi_load_state restores it into a context after dn_fwi_start with the
 * same GRAD_METHOD and NLBFGS (model, shots and data set as for the saved
 * run); the next dn_fwi_iteration continues the saved run.  With jacobian
 * non-NULL the previous model and normalised gradient are then read from the
 * reference-named files written with lbfgs_old (store_LBFGS_SH.c:78-112
 * reads them): name_fix 1 reads <JACOBIAN>_p_rho.old.0.0 etc., the names
 * extract_LBFGS_SH.c writes; name_fix 0 reads <JACOBIAN>_p_rho.old.00 etc.,
 * the names store_LBFGS_SH.c asks for (absent unless renamed: an error). */
int dn_fwi_save_state(dn_ctx *ctx, const char *path);
int dn_fwi_load_state(dn_ctx *ctx, const char *path, const char *jacobian, int name_fix);

/* Learned preconditioner for L-BFGS (NULL: off).  The network's
 * prediction H of the vs diagonal Hessian gives the initial L-BFGS Hessian
 * gamma_D * diag(1, D), D = We / H for vs (We the eprecond1.c energy weight;
 * D normalised to mean 1), gamma_D = s'y / (y' diag(1, D) y) from the newest
 * pair; the plain direction (the reference's) judges it on the first
 * iteration of a stage and every judge_every-th (0: only the first), the
 * larger misfit drop of the two being taken, and replaces it whenever its
 * line search finds no decrease.  judge_every = -k uses the learned
 * direction on the first k iterations of a stage only (judged on the
 * first) and the plain L-BFGS direction, with the same history, after.
 * The network stays owned by the caller.  Not in the reference. */
int dn_fwi_set_learned(dn_ctx *ctx, dn_unet *net, int judge_every);
/* fixed 1 (default): the learned diagonal D is predicted on the first
 * iteration of a stage and kept for the whole stage (a fixed initial Hessian,
 * so the L-BFGS pairs and the scaling stay consistent); 0: predicted again at
 * every iteration from the current energy maps and model.  Not in the
 * reference. */
int dn_fwi_set_learned_fixed(dn_ctx *ctx, int fixed);
/* per_stage 1 (default): once the plain direction replaces the learned one
 * (the judge prefers it, or the learned line search finds no decrease), the
 * rest of the stage runs plain L-BFGS with the same history, so plain and
 * learned steps never alternate; 0: the learned direction is formed again at
 * the next iteration.  Not in the reference. */
int dn_fwi_set_learned_revert(dn_ctx *ctx, int per_stage);
/* The learned diagonal as D = (We / H_pred)^power, power in (0, 2] (default
 * 1): between the plain energy scaling (power -> 0) and the predicted Hessian
 * diagonal (1).  Not in the reference. */
int dn_fwi_set_learned_power(dn_ctx *ctx, float power);
/* The network's input maps from the last gradient and the current model,
 * host [3][H8][W8] with H8, W8 = ny, nx rounded up to multiples of 8 (edge
 * rows/columns repeated): log10 of sum Ws and sum Wr relative to their
 * maxima (floored at 1e-12) / 12 + 1, and vs / mean(vs) - 1. */
int dn_precond_features(dn_ctx *ctx, float *x);

/* Current FWI model / last search direction, host [ny][nx]. */
int dn_fwi_get_model(dn_ctx *ctx, float *vs, float *rho);
int dn_fwi_get_direction(dn_ctx *ctx, float *grad_vs, float *grad_rho);

/* ------------------------------------------------------------------------
 * Learned diagonal-Hessian preconditioner.  Not in the reference: a 4-level
 * U-Net (3x3 convolutions, 16/32/48/64 channels, BF16 weights and
 * activations, FP32 accumulation on the matrix cores) mapping
 * [source energy, receiver energy, vs] maps to the log of a per-cell
 * diagonal-Hessian scaling.  Images are host [n][3][h][w] (inputs) and
 * [n][h][w] (outputs, targets) float32, h and w multiples of 8.  Errors:
 * dn_last_error(NULL).
 * ---------------------------------------------------------------------- */
dn_unet *dn_unet_create(unsigned int seed);
void dn_unet_destroy(dn_unet *u);
long long dn_unet_num_params(const dn_unet *u);
/* FP32 master parameters, per layer W[3][3][co][ci] then b[co]. */
int dn_unet_get_params(dn_unet *u, float *out);
int dn_unet_set_params(dn_unet *u, const float *in);
int dn_unet_forward(dn_unet *u, int n, int h, int w, const float *x, float *y);
/* Mean squared error against t and its gradient w.r.t. the parameters
 * (grad may be NULL); no update. */
int dn_unet_loss_grad(dn_unet *u, int n, int h, int w, const float *x, const float *t, double *loss, float *grad);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) with learning rate lr; loss
 * is the value before the step. */
int dn_unet_train_step(dn_unet *u, int n, int h, int w, const float *x, const float *t, float lr, double *loss);

#ifdef __cplusplus
}
#endif
#endif /* DENISE_SH_H */
