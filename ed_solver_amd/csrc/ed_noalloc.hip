// ed_dynamic_noalloc: compute_dynamic (ed_dynamic) for callers that cannot allocate
// device memory.  Every buffer beyond the caller's inputs and outputs lives in one
// file-scope device array, `scratch`, of kScratchBytes (about 34 MB), which each GPU the
// library runs on holds until the process ends.  Its address on a device comes from
// hipGetSymbolAddress, taken once per device with that device current (at ed_init, or
// at the first call on a device); host code uses that address, never the array's name.
//
// Layout of the array (offsets 256-byte aligned):
//   LzState | partial sums (kPartA .. kPartC + 2*8192 doubles) | 3 Krylov slots of
//   kNoAllocMaxDim doubles | start vector [kNoAllocMaxDim] | a, b, scale [2*kMaxIterG+1] |
//   w_n [kNoAllocMaxNw] | dyn [2*kNoAllocMaxNw] | error flag
//
// Every call on one GPU shares that GPU's array and leaves its contents behind: the
// recurrence state is re-initialised by the start kernel and the continued-fraction
// error flag is cleared by a fill kernel at each call.  Calls that overlap on one GPU
// (two host threads, or two streams) race on the array.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "ed_kernels.h"

namespace ed {

int cfrac_launch_dev(int n, const double *d_a, const double *d_b, int lda, int njobs, int nw, const double *d_wn,
                     const double *d_scale, double *d_out, int *d_err, hipStream_t st);

constexpr size_t align256(size_t x) { return (x + 255) / 256 * 256; }

constexpr size_t kOffState = 0;
constexpr size_t kOffPart = align256(kOffState + sizeof(LzState));
constexpr size_t kNPart = kPartC + 2 * 8192;   // partials without reorthogonalisation
constexpr size_t kOffBasis = align256(kOffPart + sizeof(double) * kNPart);
constexpr size_t kOffStart = align256(kOffBasis + sizeof(double) * 3 * (size_t)kNoAllocMaxDim);
constexpr size_t kOffAB = align256(kOffStart + sizeof(double) * (size_t)kNoAllocMaxDim);
constexpr size_t kOffWn = align256(kOffAB + sizeof(double) * (2 * (size_t)kNoAllocMaxIter + 1));
constexpr size_t kOffOut = align256(kOffWn + sizeof(double) * (size_t)kNoAllocMaxNw);
constexpr size_t kOffErr = align256(kOffOut + sizeof(double) * 2 * (size_t)kNoAllocMaxNw);
constexpr size_t kScratchBytes = align256(kOffErr + sizeof(int));

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static unsigned char *g_scratch[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Address of this device's copy of `scratch` (the current device must be `device`).
int noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(ED_ERR_ARG, "no-allocation scratch: device index above 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return ED_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(ED_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = (unsigned char *)p;
    return ED_OK;
}

__global__ void k_scale_inplace(double *x, int64_t n, double s)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = s * x[i];
}

__global__ void k_clear_int(int *p)
{
    if (threadIdx.x == 0) *p = 0;
}

}  // namespace ed

using namespace ed;

extern "C" int ed_dynamic_noalloc(ed_sector *sec, int iph, double energy, const double *phi, int nitergreen,
                                  double cutoff_dynamic, int precision, int nw, const double *wn, double *dyn,
                                  double *a, double *b)
{
    if (!sec || !phi || !wn || !dyn || nw < 1) return fail(ED_ERR_ARG, "ed_dynamic_noalloc: bad argument");
    if (precision != ED_PREC_F64 && precision != ED_PREC_F32)
        return fail(ED_ERR_ARG, "ed_dynamic_noalloc: bad precision");
    if (iph != 1 && iph != 2) return fail(ED_ERR_ARG, "ed_dynamic_noalloc: iph must be 1 (particle) or 2 (hole)");
    if (nitergreen < 1) return fail(ED_ERR_ARG, "ed_dynamic_noalloc: nitergreen < 1");
    if (sec->dim > kNoAllocMaxDim)
        return fail(ED_ERR_ARG, "ed_dynamic_noalloc: sector vector length " + std::to_string(sec->dim) +
                                    " above the scratch maximum " + std::to_string(kNoAllocMaxDim));
    if (nw > kNoAllocMaxNw)
        return fail(ED_ERR_ARG, "ed_dynamic_noalloc: nw above the scratch maximum " + std::to_string(kNoAllocMaxNw));
    const int niter = (int)std::min<int64_t>(nitergreen, sec->cplx ? sec->cdim : sec->dim);
    if (niter > kNoAllocMaxIter)
        return fail(ED_ERR_ARG, "ed_dynamic_noalloc: nitergreen above the scratch maximum " +
                                    std::to_string(kNoAllocMaxIter));
    ED_HIP_CHECK(hipSetDevice(sec->device));
    ED_TRY(noalloc_setup(sec->device));
    unsigned char *base = g_scratch[sec->device];
    const double nrm2 = host_norm2(phi, sec->dim);
    double ah[kNoAllocMaxIter], bh[kNoAllocMaxIter];
    if (nrm2 < cutoff_dynamic) {   // green_class_compute_dynamic.f90:263-265
        std::memset(dyn, 0, sizeof(double) * 2 * (size_t)nw);
        if (a) std::memset(a, 0, sizeof(double) * niter);
        if (b) std::memset(b, 0, sizeof(double) * niter);
        return ED_OK;
    }
    hipStream_t st = sec->stream;
    double *d_start = (double *)(base + kOffStart);
    ED_HIP_CHECK(hipMemcpyAsync(d_start, phi, sizeof(double) * sec->dim, hipMemcpyHostToDevice, st));
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>((sec->dim + kBlock * 4 - 1) / (kBlock * 4), 2048));
    hipLaunchKernelGGL(k_scale_inplace, dim3(g), dim3(kBlock), 0, st, d_start, sec->dim, 1.0 / std::sqrt(nrm2));
    ED_HIP_CHECK(hipGetLastError());
    const LzWork w{base + kOffBasis, (sec->dim + 63) / 64 * 64, 3, (LzState *)(base + kOffState),
                   (double *)(base + kOffPart)};
    ED_TRY(lanczos_run_ws(sec, w, niter, d_start, precision, ah, bh));
    if (a) std::memcpy(a, ah, sizeof(double) * niter);
    if (b) std::memcpy(b, bh, sizeof(double) * niter);
    const double ph = iph == 1 ? 1.0 : -1.0;   // tri%diag = ph_sign * (diag - E), :310-311
    for (int i = 0; i < niter; ++i) ah[i] = ph * (ah[i] - energy);
    double *d_ab = (double *)(base + kOffAB), *d_wn = (double *)(base + kOffWn), *d_out = (double *)(base + kOffOut);
    int *d_err = (int *)(base + kOffErr);
    ED_HIP_CHECK(hipMemcpyAsync(d_ab, ah, sizeof(double) * niter, hipMemcpyHostToDevice, st));
    ED_HIP_CHECK(hipMemcpyAsync(d_ab + niter, bh, sizeof(double) * niter, hipMemcpyHostToDevice, st));
    // dyn = -fermion_sign * CF * ||phi||^2 with fermion_sign = -1 (:224-230, :324)
    ED_HIP_CHECK(hipMemcpyAsync(d_ab + 2 * niter, &nrm2, sizeof(double), hipMemcpyHostToDevice, st));
    ED_HIP_CHECK(hipMemcpyAsync(d_wn, wn, sizeof(double) * nw, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, st, d_err);
    ED_TRY(cfrac_launch_dev(niter, d_ab, d_ab + niter, niter, 1, nw, d_wn, d_ab + 2 * niter, d_out, d_err, st));
    int h_err = 0;
    ED_HIP_CHECK(hipMemcpyAsync(&h_err, d_err, sizeof(int), hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipMemcpyAsync(dyn, d_out, sizeof(double) * 2 * nw, hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipStreamSynchronize(st));
    if (h_err) return fail(ED_ERR_CF, "continued fraction: denominator below epsilon (matrices_tridiag.f90:174-186)");
    return ED_OK;
}
