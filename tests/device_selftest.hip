// device_selftest.hip -- GPU checks of device helpers from
// chllmodel_amd/csrc/chll_device.h whose results depend on hardware
// instructions (v_exp_f32), built into tests/libdevice_selftest.so by
// __graft_entry__.build() and loaded by tests/test_gpu_selftest.py.
#include "chll_device.h"

using namespace chll;

// For x_i = -90 (i + 0.5) / n: ratio[i] = |e / exp(x) - 1| / r (the measured
// error of the fp32 estimate over the margin the bracket assumes; 0 below
// x = -80, where only the upper bound is used), bad[i] = number of failed
// checks: the bracket must hold exp(x), and accept_exp(x, u) must equal
// u <= exp(x) for u placed on, inside and outside the bracket edges.
__global__ void k_exp_check(int n, double *ratio, int *bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = -90.0 * ((double)i + 0.5) / (double)n;
    double lo, hi;
    exp_bracket(x, lo, hi);
    const double ex = exp(x);
    int nb = 0;
    if (!(lo <= ex && ex <= hi)) ++nb;
    if (x >= -80.0) {
        const float xf = (float)x;
        const double e = (double)__builtin_amdgcn_exp2f(xf * 1.44269504f);
        const double r = fma(-(double)xf, 0x1.0p-20, 0x1.0p-17);
        ratio[i] = fabs(e / ex - 1.0) / r;
    } else {
        ratio[i] = 0.0;
    }
    const double us[] = {0.0, lo * 0.5, lo, nextafter(lo, 0.0), nextafter(lo, 1.0), nextafter(ex, 0.0), ex,
                         nextafter(ex, 1.0), 0.5 * (ex + hi), hi, nextafter(hi, 1.0), 2.0 * hi, 0.999999999};
    for (double u : us)
        if (accept_exp(x, u) != (u <= ex)) ++nb;
    bad[i] = nb;
}

extern "C" int chll_selftest_exp(int n, double *ratio_out, int *bad_out)
{
    double *ratio = nullptr;
    int *bad = nullptr;
    if (hipMalloc(&ratio, sizeof(double) * n) != hipSuccess) return -1;
    if (hipMalloc(&bad, sizeof(int) * n) != hipSuccess) { (void)hipFree(ratio); return -1; }
    hipLaunchKernelGGL(k_exp_check, dim3((n + 255) / 256), dim3(256), 0, 0, n, ratio, bad);
    int rc = hipGetLastError() == hipSuccess ? 0 : -1;
    if (rc == 0 && hipMemcpy(ratio_out, ratio, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    if (rc == 0 && hipMemcpy(bad_out, bad, sizeof(int) * n, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    (void)hipFree(ratio);
    (void)hipFree(bad);
    return rc;
}
