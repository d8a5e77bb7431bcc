// Context of the handle API (include/denise_sh.h) and helpers shared by the
// host-side translation units of libdenise_sh.so.
#pragma once
#include "../../include/denise_sh.h"
#include "dn_internal.h"
#include "dn_setup.h"

#include <string>
#include <vector>

struct dn_fwi_state;
struct dn_gradproc_state;
struct dn_unet;
// U-Net forward on device buffers (dn_unet.cpp): x [n][3][H][W] -> y [n][H][W]
int dn_unet_forward_device(dn_unet *u, int n, int H, int W, const float *x_dev, float *y_dev);

struct dn_ctx {
    dn_params p{};
    int K = 1;
    dn::Geo g{};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    float hc[7] = {0};
    std::vector<float> pml_host;   // 12 x (2FW+1), same packing as d_pml
    float *d_pml = nullptr;
    dn::PmlVec pml{};
    // relaxation constants
    float eta[dn::kMaxL] = {0}, sumu = 0.0f;
    float bip[dn::kMaxL] = {0}, bjm[dn::kMaxL] = {0}, cip[dn::kMaxL] = {0}, cjm[dn::kMaxL] = {0};
    // model (padded)
    float *d_vs = nullptr, *d_rho = nullptr, *d_taus = nullptr, *d_uip = nullptr, *d_ujp = nullptr, *d_b = nullptr;
    float *d_tausipjp = nullptr, *d_fipjp = nullptr, *d_f = nullptr, *d_dip = nullptr, *d_d = nullptr;
    float *d_spk = nullptr;   // the stress coefficients packed per cell (see StepArgs::spk)
    float C_vs = 0.0f, C_rho = 0.0f;
    bool have_model = false;
    bool C_fixed = false;            // C_vs, C_rho taken (first gradient or dn_fwi_start)
    // gradient sums [ny][nx] and outputs after the chain rule
    float *d_gu = nullptr, *d_grho = nullptr, *d_gvs_out = nullptr, *d_grho_out = nullptr;
    unsigned int *d_maxbits = nullptr;
    int *d_status = nullptr;         // instability / shot-map latch of the step kernel (StepArgs::status)
    // receivers
    int ntr = 0;
    std::vector<int> rec_i, rec_j;
    int *d_rec_row = nullptr, *d_rec_tab = nullptr;
    // shots
    int nshots = 0, nt_max = 0;
    std::vector<int> shot_nt, src_i, src_j;
    int *d_shot_nt = nullptr, *d_src_off = nullptr, *d_src_i = nullptr, *d_src_j = nullptr;
    float *d_sig = nullptr, *d_obs = nullptr;
    // batch workspace
    int batch_req = 0, rows_req = 0, batch_cap = 0, batch_nt_max = 0, batch_fit = 0;
    // gradients over several batches of uneven record lengths: batches by
    // decreasing length, per-shot terms [rank][4][nx*ny] summed in shot order
    bool order_by_nt = true;
    float *d_terms = nullptr;
    size_t terms_cap = 0;
    int *d_rank = nullptr;            // [shots of the gradient] rank of each position of the order
    int rank_cap = 0;
    bool batch_has_snap = false;
    float *d_vy[2] = {nullptr, nullptr}, *d_sxy[2] = {nullptr, nullptr}, *d_syz[2] = {nullptr, nullptr};
    float *d_psx[2] = {nullptr, nullptr}, *d_psy[2] = {nullptr, nullptr};
    float *d_rq = nullptr, *d_pvx = nullptr, *d_pvy = nullptr;
    float *d_seis = nullptr, *d_resid = nullptr, *d_snap = nullptr;
    uint16_t *d_snap16 = nullptr;
    float *d_snap_scale = nullptr, *d_stage = nullptr;
    unsigned int *d_stage_max = nullptr;
    float *d_Ws = nullptr, *d_Wr = nullptr, *d_gu_shot = nullptr, *d_grho_shot = nullptr;
    double *d_part = nullptr;
    int *d_shot_map = nullptr;
    std::vector<int> cur_map;        // absolute shot of each batch slot in the last pass
    std::vector<double> shot_l2, shot_energy;   // per-shot misfit sums of the last residual pass
    int cur_final = 0;
    // timing of the last pass set
    double ms_steps = 0.0;
    long long launches = 0, cell_updates = 0;
    double ms_pass[2] = {0.0, 0.0};          // [forward, adjoint]
    long long launches_pass[2] = {0, 0}, cells_pass[2] = {0, 0};
    // time-loop replay: captured chunks of `steps` time steps per batch shape
    struct Graph {
        int n, R, steps, kind;
        bool adjoint, store, record, hess;
        std::vector<char> corr_res;   // record-length residues that select the correlating launches
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    int graph_steps = -1;            // -1 default chunk, 0 eager launches, > 0 requested chunk
    bool two_shots = false;          // plain steps advance two shots per wave (dn_set_wave_shots)
    int sg = 1;                      // shots per workgroup of the step kernel (dn_set_group_layout)
    int *d_ntbase = nullptr;         // device step counter read by replayed chunks
    int nt_limit = 0;                // > 0: passes stop after this step (autotuner)
    long long graph_replays = 0;
    // pseudo-Hessian (forward passes accumulate it when hess_mode is set)
    bool hess_mode = false;
    float *d_hu = nullptr, *d_hr = nullptr;
    float *d_ws_sum = nullptr, *d_wr_sum = nullptr;   // energy maps of the last gradient
    // FWI optimiser state (dn_fwi.cpp) and an optional gradient preconditioner
    dn_fwi_state *fwi = nullptr;
    int (*precond_hook)(dn_ctx *, float *grad_vs, float *grad_rho) = nullptr;
    // learned preconditioner (owned by the caller) and the judge interval
    dn_unet *unet = nullptr;
    int judge_every = 5;
    bool learned_fixed = true;       // D from the first iteration of a stage, kept for the stage
    bool learned_revert = true;      // once the plain direction replaces the learned one, plain
                                     // L-BFGS for the rest of the stage
    double learned_power = 1.0;      // D = (We / H_pred)^power
    // gradient smoothing / taper workspace (dn_gradproc.hip)
    dn_gradproc_state *gradproc = nullptr;
    // data conditioning of the residual (dn_set_datacond): parameters and the
    // per (shot, receiver) flags derived from them for the residual kernel
    struct DataCond {
        bool on = false;
        int trkill = 0, kill_nsrc = 0, kill_ntr = 0, offset_mute = 0;
        float offsetc = 0.0f;
        std::vector<int> kill;
        std::vector<float> src_x, src_y;
    } dc;
    bool dc_dirty = false;
    unsigned char *d_trace_flags = nullptr;
    std::vector<int> src_off;        // host copy of the per-shot source ranges
    // the signals and observed data as the caller set them (host copies):
    // a stage's frequency filter is applied to these afresh, never on top of
    // an earlier stage's (denise.c:1099-1111, 1381-1394 filter the freshly
    // made wavelet and the data read from disk each time)
    struct TimeFilt {
        int time_filt = 0, order = 0;
        float fc_start = 0.0f, fc = 0.0f;
        bool operator==(const TimeFilt &o) const
        {
            return time_filt == o.time_filt && (time_filt == 0 || (order == o.order && fc == o.fc &&
                                                                 (time_filt != 2 || fc_start == o.fc_start)));
        }
    };
    std::vector<float> h_sig, h_obs;
    TimeFilt filt_sig, filt_obs;     // the filter the device signals / observed data carry
};

int dn_fail(dn_ctx *c, const char *fmt, ...);
// returns dn_fail(ctx, ...) from the enclosing function when a HIP call fails
#define HIPCHK(ctx, expr)                                                                                  \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return dn_fail((ctx), "%s failed: %s", #expr, hipGetErrorString(e_));        \
    } while (0)
int dn_prepare_model(dn_ctx *c);
int dn_update_means(dn_ctx *c);   // C_vs, C_rho from the current device model, then fixed
int dn_ensure_batch(dn_ctx *c, int want, bool gradient);
int dn_run_pass(dn_ctx *c, const std::vector<int> &shots, bool adjoint, bool store, bool record);
// residual of the batch slots of the last forward pass: resid + per-shot l2/energy
int dn_batch_residual(dn_ctx *c, const std::vector<int> &shots);
int dn_gradient_shots(dn_ctx *c, const std::vector<int> &shots, bool accumulate);
int dn_forward_shots(dn_ctx *c, const std::vector<int> &shots, float *seis_host);
int dn_nsnap(const dn_ctx *c);
void dn_reset_timing(dn_ctx *c);
void dn_fwi_destroy(dn_ctx *c);
void dn_gradproc_destroy(dn_ctx *c);
// smooth_grad (SPATFILTER 0 or 4) and taper_grad on device planes [ny][nx]
// (null planes are skipped), with the context's vs model and geometry
int dn_smooth_grad_device(dn_ctx *c, int spatfilter, float wd_damp, float wd_damp1, float fc_end, float *g_vs,
                          float *g_rho);
int dn_taper_grad_device(dn_ctx *c, const dn_taper *T, float *g_vs, float *g_rho);
void dn_drop_graphs(dn_ctx *c);
