// Misfit and energy of one shot's traces outside a context: calc_res.c
// (LNORM 2) and calc_energy.c on host [ntr][ns] seismograms, through the
// residual kernel the gradient passes use (k_residual), with FP64 sums.
#include "dn_core.h"

#include <vector>

using namespace dn;

extern "C" int dn_calc_res(int ntr, int ns, float dt, const float *syn, float *obs, float *diff, double *l2,
                           double *energy)
{
    if (ntr < 1 || ns < 1 || !syn || !obs || !diff || !l2 || !energy)
        return dn_fail(nullptr, "dn_calc_res: need ntr >= 1, ns >= 1 and five arrays");
    const size_t n = (size_t)ntr * ns;
    float *d_syn = nullptr, *d_obs = nullptr, *d_diff = nullptr;
    double *d_part = nullptr;
    int *d_tab = nullptr;
    hipStream_t st = nullptr;
    std::vector<double> part(2 * (size_t)ntr);
    const int tab[2] = {0, ns};
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&d_syn, n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&d_obs, n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&d_diff, n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&d_part, part.size() * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&d_tab, sizeof(tab));
    if (e == hipSuccess) e = hipMemcpyAsync(d_syn, syn, n * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_obs, obs, n * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_tab, tab, sizeof(tab), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = launch_residual(1, 1, ntr, ns, d_tab, d_tab + 1, d_syn, d_obs, d_diff, d_part, dt, nullptr, st);
    if (e == hipSuccess) e = hipMemcpyAsync(diff, d_diff, n * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(part.data(), d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_syn);
    (void)hipFree(d_obs);
    (void)hipFree(d_diff);
    (void)hipFree(d_part);
    (void)hipFree(d_tab);
    if (st) (void)hipStreamDestroy(st);
    if (e != hipSuccess) return dn_fail(nullptr, "dn_calc_res: %s", hipGetErrorString(e));
    double a = 0.0, b = 0.0;
    for (int r = 0; r < ntr; ++r) {
        a += part[2 * (size_t)r];
        b += part[2 * (size_t)r + 1];
    }
    *l2 = a;
    *energy = b;
    // calc_res.c:152: the first observed sample of each trace becomes the
    // synthetic one (the kernel used it that way)
    for (int r = 0; r < ntr; ++r) obs[(size_t)r * ns] = syn[(size_t)r * ns];
    return 0;
}
