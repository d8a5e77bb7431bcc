// dn_calc_res_noalloc: dn_calc_res on device arrays without allocating.
// Its per-trace partial sums and shot tables live in `scratch`, a
// file-scope device array of fixed size, so callers that must not allocate
// (code under stream capture, serving loops) can use it.
#include "dn_core.h"

#include <mutex>

using namespace dn;

namespace {

// Layout of the scratch: two ints of shot tables (shot map, record length)
// in the first 256 bytes, then 2 doubles per trace.
constexpr size_t kScratchHead = 256;
constexpr size_t kScratchBytes = kScratchHead + 2 * sizeof(double) * (size_t)DN_CALC_RES_NOALLOC_MAX_TRACES;

__global__ void k_one_shot_tables(int *t, int ns)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        t[0] = 0;     // shot map: slot 0 is shot 0
        t[1] = ns;    // its record length
    }
}

// l2 and energy over the traces, added in trace order (the order
// dn_calc_res and the context's residual pass sum them on the host)
__global__ void k_sum_trace_pairs(const double *part, int ntr, double *out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double l2 = 0.0, en = 0.0;
        for (int r = 0; r < ntr; ++r) {
            l2 += part[2 * r];
            en += part[2 * r + 1];
        }
        out[0] = l2;
        out[1] = en;
    }
}

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

extern "C" int dn_calc_res_noalloc(int ntr, int ns, float dt, const float *syn, const float *obs, float *diff,
                                   double *l2_energy, void *stream)
{
    if (ntr < 1 || ns < 1 || !syn || !obs || !diff || !l2_energy)
        return dn_fail(nullptr, "dn_calc_res_noalloc: need ntr >= 1, ns >= 1 and four device arrays");
    if (ntr > DN_CALC_RES_NOALLOC_MAX_TRACES)
        return dn_fail(nullptr, "dn_calc_res_noalloc: %d traces exceed the maximum of %d", ntr,
                       DN_CALC_RES_NOALLOC_MAX_TRACES);
    // the scratch array's device address, looked up once per device with
    // that device current (each device holds its own copy of the array)
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static unsigned char *dev_scratch[kMaxDevices] = {nullptr};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return dn_fail(nullptr, "dn_calc_res_noalloc: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return dn_fail(nullptr, "dn_calc_res_noalloc: device %d not supported", dev);
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(mu);
        if (!dev_scratch[dev]) {
            void *p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return dn_fail(nullptr, "dn_calc_res_noalloc: scratch address: %s", hipGetErrorString(e));
            dev_scratch[dev] = (unsigned char *)p;
        }
        base = dev_scratch[dev];
    }
    hipStream_t st = (hipStream_t)stream;
    int *tab = (int *)base;
    double *part = (double *)(base + kScratchHead);
    hipLaunchKernelGGL(k_one_shot_tables, dim3(1), dim3(64), 0, st, tab, ns);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_residual(1, 1, ntr, ns, tab, tab + 1, syn, obs, diff, part, dt, nullptr, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_sum_trace_pairs, dim3(1), dim3(64), 0, st, part, ntr, l2_energy);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return dn_fail(nullptr, "dn_calc_res_noalloc: %s", hipGetErrorString(e));
    return 0;
}
