// Host-side setup routines of the SH path (see dn_setup.cpp).
#pragma once
#include <vector>

namespace dn {

constexpr int kSetupMaxL = 8;

// holbergcoeff.c:19-90; returns <0 for an unsupported FDORDER/MAXRELERROR.
int fd_coefficients(int fdorder, int maxrelerror, float hc[7]);

struct CpmlIn {
    int nx, ny, fw;
    float dh, dt, damping, fpml, npower, k_max_pml;
};
// PML_pro.c:65-297.  v[0..2] = K_x, a_x, b_x; v[3..5] = K_x_half, a_x_half,
// b_x_half; v[6..8] = K_y, a_y, b_y; v[9..11] = K_y_half, a_y_half, b_y_half;
// each 1-based with n = 2FW+1 entries.
struct CpmlOut {
    int n = 0;
    std::vector<float> v[12];
};
void cpml_profiles(const CpmlIn &in, CpmlOut &out);

// readmod.c:30-34 and prepare_update_s.c:20-56 (per-mechanism constants).
void relaxation_constants(int L, const float *FL, float DT, float *eta, float *sumu, float *bip, float *bjm,
                          float *cip, float *cjm);

// wavelet.c:32-84 for QUELLART 1, 2, 4, 5, 6; out[nt-1] = sample nt.
int wavelet(int quellart, int nt_total, float DT, float tshift, float fc, float amp, float *out);

// libcseife Butterworth (cseife.c:147-177) on x[0..n-1], FP64, in place.
void butterworth(double *x, int n, double dt, double t0, int order, int highpass);

// timedomain_filt.c:10-53 on data [ntr][ns] (method 1 lowpass, 2 highpass).
void timedomain_filt(float *data, int ntr, int ns, float DT, float fc, int order, int method);

}  // namespace dn
