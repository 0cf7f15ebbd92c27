// fx_noalloc.hip -- fx_mass_balance_noalloc: the water ledger of
// fx_mass_balance without any device or pinned-host allocation, for callers
// that cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// fx_mass_balance allocates the volume kernel's block partials on every call
// and sums them on the host.  This variant writes the same partials, with the
// same kernels, into a file-scope device array and sums them on the device in
// the same order with the same Kahan steps, so its results are bit-identical.
// The array is each device's own copy, found once per device with
// hipGetSymbolAddress; the host never passes the array's name to a kernel.
#include <mutex>

#include "fx_kern32.h"

// block partials of the volume kernel, then three result slots
#define FX_NOALLOC_PARTS ((int64_t)1 << 20)
#define FX_NOALLOC_RES 8
#define FX_NOALLOC_MAXDEV 64

static __device__ __align__(256) double scratch[FX_NOALLOC_PARTS + FX_NOALLOC_RES];

namespace {

// the device address of `scratch` on each device (nullptr until looked up), and
// one lock per device: calls on one device share its array
double *g_scratch[FX_NOALLOC_MAXDEV];
std::mutex g_lock[FX_NOALLOC_MAXDEV];

constexpr int FIN_CHUNK = 2048;   // partials staged in LDS per round (16 KiB)

// Kahan sum of the volume partials in index order (as fx_mass_balance sums
// them on the host), the plain sum of the pending weir-face partials of an
// FP32-mixed domain (stride FX32_NPV, entry 2), and the three ledger values.
__global__ __launch_bounds__(256) void k_mass_finish(const double *vpart, int64_t nv, const double *wpart,
                                                     int64_t nw, const fx_ctl *ctl, double arbase, double dxy,
                                                     double *res)
{
    __shared__ double buf[FIN_CHUNK];
    double s = 0.0, c = 0.0, w = 0.0;
    for (int64_t base = 0; base < nv; base += FIN_CHUNK) {
        const int n = (int)min((int64_t)FIN_CHUNK, nv - base);
        for (int i = threadIdx.x; i < n; i += blockDim.x) buf[i] = vpart[base + i];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int i = 0; i < n; i++) {
                const double y = buf[i] - c;
                const double t = s + y;
                c = (t - s) - y;
                s = t;
            }
        __syncthreads();
    }
    for (int64_t base = 0; base < nw; base += FIN_CHUNK) {
        const int n = (int)min((int64_t)FIN_CHUNK, nw - base);
        for (int i = threadIdx.x; i < n; i += blockDim.x) buf[i] = wpart[FX32_NPV * (base + i) + 2];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int i = 0; i < n; i++) w += buf[i];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        res[0] = s * arbase;
        res[1] = ctl->led_melt;
        res[2] = ctl->led_weir + w * dxy * dxy;
    }
}

// this device's copy of the array, looked up on the first call on the device
// (the caller holds g_lock[dev] and has made dev current)
int scratch_of(int dev, double **out)
{
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fx_hip_check(e != hipSuccess ? e : hipErrorInvalidSymbol,
                                                       "hipGetSymbolAddress(scratch)");
        g_scratch[dev] = (double *)p;
    }
    *out = g_scratch[dev];
    return FX_OK;
}

} // namespace

extern "C" int fx_mass_balance_noalloc(fx_domain d, double out[3])
{
    if (!d) return fx_fail(FX_ERR_ARG, "domain is NULL");
    if (!out) return fx_fail(FX_ERR_ARG, "out is NULL");
    const int dev = d->desc.device;
    if (dev < 0 || dev >= FX_NOALLOC_MAXDEV) return fx_fail(FX_ERR_ARG, "device ordinal above the no-alloc limit");
    FX_HIP(hipSetDevice(dev));
    const bool fp64 = d->desc.precision == FX_FP64_PARITY;
    const int64_t nv = fp64 ? fx64_volume_nparts(d) : fx32_volume_nparts(d);
    if (nv > FX_NOALLOC_PARTS)
        return fx_fail(FX_ERR_ARG, "fx_mass_balance_noalloc: the domain's volume kernel has " + std::to_string(nv) +
                                       " blocks, above the " + std::to_string(FX_NOALLOC_PARTS) +
                                       " partials its fixed device buffer holds (use fx_mass_balance)");
    std::lock_guard<std::mutex> hold(g_lock[dev]);
    double *base = nullptr;
    int rc = scratch_of(dev, &base);
    if (rc) return rc;
    if ((rc = fp64 ? fx64_volume_partials(d, base) : fx32_volume_partials(d, base))) return rc;
    double *res = base + FX_NOALLOC_PARTS;
    const fx_params P = fx_make_params(d);
    hipLaunchKernelGGL(k_mass_finish, dim3(1), dim3(256), 0, d->stream, base, nv,
                       fp64 ? (const double *)nullptr : (const double *)d->s32.partials,
                       fp64 ? (int64_t)0 : (int64_t)d->s32.npart, d->ctl, P.arbase, (double)d->desc.dxy, res);
    FX_HIP(hipGetLastError());
    FX_HIP(hipMemcpyAsync(out, res, 3 * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    FX_HIP(hipStreamSynchronize(d->stream));
    return FX_OK;
}
