// Green's-function path: excitation vectors, excitation Lanczos and the continued
// fraction.
//
//   k_cmap / k_apply_c   apply_Cupdo (apply_C.f90:497-598).  c+/c on one impurity
//                        orbital of one spin maps each spin state to at most one state
//                        of the N+-1 spin sector, so phi is a signed gather of psi:
//                        one thread per phi element, coalesced writes.  An S_z
//                        sector (apply_Csz, apply_C.f90:277-327) is one string of
//                        2 Ns Nambu orbitals: the same map over that string.
//   k_cfrac              invert_zmtridiag (matrices_tridiag.f90:163-198), one thread
//                        per Matsubara frequency (and per job when batched).
//   ed_dynamic           compute_dynamic_cpu (green_class_compute_dynamic.f90:191-407).
//   (the batched form of ed_dynamic lives in ed_batch.hip)
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ed_kernels.h"

namespace ed {

int lanczos_run_dev(ed_sector *s, int niter, const double *d_start, int prec, double *diag, double *subdiag,
                    int *breakdown_at);

// For every destination spin state t: src[t] = source rank (or -1), sg[t] = operator sign.
// c+ (pm > 0): t must hold `bit`, source = t without it; c (pm < 0): t must not hold it,
// source = t with it.  The sign counts occupied orbitals above `bit` in t
// (Fermion_sign, quantum_fermion_ket_class.f90:605-618).  With cbit >= 0 the operator is
// conditioned on orbital cbit holding cval (0 or 1) in the source state, which is its
// value in t because cbit != bit: the correlated excitation of apply_Csz_N_updo
// (apply_C.f90:404-483).
__global__ void k_cmap(const int32_t *dst_states, int d, int bit, int pm, const int32_t *rank, int32_t *src,
                       int32_t *sg, int cbit, int cval)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d) return;
    const uint32_t s = (uint32_t)dst_states[t];
    const bool has = (s >> bit) & 1u;
    if ((pm > 0) != has || (cbit >= 0 && (int)((s >> cbit) & 1u) != cval)) {
        src[t] = -1;
        sg[t] = 0;
        return;
    }
    const uint32_t from = pm > 0 ? (s & ~(1u << bit)) : (s | (1u << bit));
    src[t] = rank[from];
    sg[t] = (__popc(s >> (bit + 1)) & 1) ? -1 : 1;
}

// phi(iup', ido') = sg * psi(src(iup'), ido')   (spin up)
// phi(iup', ido') = sg * psi(iup', src(ido'))   (spin down)
__global__ void k_apply_c(const double *psi, int dup_src, double *phi, int dup_dst, int64_t n_dst, int spin,
                          const int32_t *src, const int32_t *sg)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_dst; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i % dup_dst);
        const int64_t c = i / dup_dst;
        double v = 0.0;
        if (spin == 0) {
            const int s = src[r];
            if (s >= 0) v = sg[r] * psi[s + c * (int64_t)dup_src];
        } else {
            const int s = src[c];
            if (s >= 0) v = sg[c] * psi[r + (int64_t)s * dup_src];
        }
        phi[i] = v;
    }
}

// out[job][k] = scale[job] * CF_job(i wn[k]) of the tridiagonal (a_job, b_job), a_job
// already shifted.  Bottom-up evaluation as invert_zmtridiag; err set when a
// denominator falls below machine epsilon.
__global__ void k_cfrac(int n, const double *a, const double *b, int lda, int nw, const double *wn,
                        const double *scale, double *out, int *err)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int job = blockIdx.y;
    if (k >= nw) return;
    const double *aj = a + (size_t)job * lda, *bj = b + (size_t)job * lda;
    const double sc = scale ? scale[job] : 1.0;
    double *o = out + ((size_t)job * nw + k) * 2;
    if (sc == 0.0) {
        o[0] = 0.0;
        o[1] = 0.0;
        return;
    }
    const double eps = 2.220446049250313e-16;
    const double zi = wn[k];
    // den = z - a[n-1]; r = 1/den
    double dr = -aj[n - 1], di = zi;
    if (sqrt(dr * dr + di * di) < eps) { atomicOr(err, 1); return; }
    double m2 = dr * dr + di * di;
    double rr = dr / m2, ri = -di / m2;
    for (int it = n - 2; it >= 0; --it) {
        const double b2 = bj[it + 1] * bj[it + 1];
        dr = -aj[it] - rr * b2;
        di = zi - ri * b2;
        if (sqrt(dr * dr + di * di) < eps) { atomicOr(err, 1); return; }
        m2 = dr * dr + di * di;
        rr = dr / m2;
        ri = -di / m2;
    }
    o[0] = sc * rr;
    o[1] = sc * ri;
}

__global__ void k_scale_copy(const double *x, double *y, int64_t n, double s)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = s * x[i];
}

static int grid_for(int64_t n)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + kBlock * 4 - 1) / (kBlock * 4), 2048));
}

static int check_pair(const ed_sector *src, const ed_sector *dst, int site, int spin, int pm)
{
    if (!src || !dst) return fail(ED_ERR_ARG, "ed_apply_c: NULL sector");
    if (src->sz != dst->sz) return fail(ED_ERR_ARG, "ed_apply_c: one S_z and one (N_up, N_dn) sector");
    if (src->cplx != dst->cplx) return fail(ED_ERR_ARG, "ed_apply_c: one complex and one real sector");
    if (src->nc != dst->nc || src->nb != dst->nb) return fail(ED_ERR_ARG, "ed_apply_c: sectors of different models");
    if (site < 0 || site >= src->nc) return fail(ED_ERR_ARG, "ed_apply_c: site outside [0, nc)");
    if ((spin != 0 && spin != 1) || (pm != 1 && pm != -1)) return fail(ED_ERR_ARG, "ed_apply_c: spin 0/1, pm +1/-1");
    if (src->sz) {   // Nambu orbital (site, spin): npart -> npart + pm (Csz_sector_updo, apply_C.f90:214-228)
        if (dst->nup != src->nup + pm)
            return fail(ED_ERR_ARG, "ed_apply_c: destination S_z sector is not npart + pm");
        return ED_OK;
    }
    const int nu = src->nup + (spin == 0 ? pm : 0), nd = src->ndn + (spin == 1 ? pm : 0);
    if (dst->nup != nu || dst->ndn != nd)
        return fail(ED_ERR_ARG, "ed_apply_c: destination sector is not (nup, ndn) shifted by the operator (Cupdo_sector)");
    return ED_OK;
}

int apply_c_dev(const ed_sector *src, const ed_sector *dst, int site, int spin, int pm, const double *d_psi,
                double *d_phi, hipStream_t st, int ncond)
{
    ED_TRY(check_pair(src, dst, site, spin, pm));
    if (ncond && !src->sz)   // apply_N_Cup / apply_N_Cdo STOP there (apply_C.f90:134-145,168-179)
        return fail(ED_ERR_ARG, "ed_apply_c_n: the correlated excitation exists for S_z sectors only");
    // An S_z sector is one string of 2 Ns Nambu orbitals held as the "up" table (ddn = 1):
    // the operator creates (pm > 0) or destroys the Nambu orbital of (site, spin), bit
    // spin Ns + nb + site (apply_Csz, apply_C.f90:277-327).
    const int bit = src->sz ? spin * src->ns + src->nb + site : src->nb + site;
    const int kspin = src->sz ? 0 : spin;
    const SpinTables &T = kspin == 0 ? dst->up : dst->dn;
    int32_t *d_map = nullptr;
    ED_HIP_CHECK(hipMallocAsync((void **)&d_map, sizeof(int32_t) * 2 * std::max(1, T.d), st));
    // correlated excitation (apply_Csz_N_updo): the Nambu up orbital of the site acts only
    // where the Nambu down orbital is empty (NDN == 0: a physical down electron), the
    // Nambu down orbital only where the Nambu up orbital is occupied (NUP == 1)
    const int cbit = ncond ? (spin == 0 ? src->ns + src->nb + site : src->nb + site) : -1;
    const int cval = spin == 0 ? 0 : 1;
    hipLaunchKernelGGL(k_cmap, dim3((T.d + kBlock - 1) / kBlock), dim3(kBlock), 0, st, T.states, T.d,
                       bit, pm, src->d_rank, d_map, d_map + T.d, cbit, cval);
    // a complex sector's planar vectors: the same real map on the Re and the Im plane
    const int planes = dst->cplx ? 2 : 1;
    const int64_t ns = dst->cplx ? src->cdim : src->dim, nd = dst->cplx ? dst->cdim : dst->dim;
    for (int p = 0; p < planes; ++p)
        hipLaunchKernelGGL(k_apply_c, dim3(grid_for(nd)), dim3(kBlock), 0, st, d_psi + p * ns, src->up.d,
                           d_phi + p * nd, dst->up.d, nd, kspin, (const int32_t *)d_map, (const int32_t *)(d_map + T.d));
    ED_HIP_CHECK(hipGetLastError());
    ED_HIP_CHECK(hipFreeAsync(d_map, st));
    return ED_OK;
}

// k_cfrac over njobs tridiagonals on stream st; *d_err (cleared by the caller) is set
// when a denominator falls below epsilon.
int cfrac_launch_dev(int n, const double *d_a, const double *d_b, int lda, int njobs, int nw, const double *d_wn,
                     const double *d_scale, double *d_out, int *d_err, hipStream_t st)
{
    hipLaunchKernelGGL(k_cfrac, dim3((nw + kBlock - 1) / kBlock, njobs), dim3(kBlock), 0, st, n, d_a, d_b, lda, nw,
                       d_wn, d_scale, d_out, d_err);
    ED_HIP_CHECK(hipGetLastError());
    return ED_OK;
}

int cfrac_batch_dev(int n, const double *d_a, const double *d_b, int lda, int njobs, int nw, const double *d_wn,
                    const double *d_scale, double *d_out, hipStream_t st)
{
    int *d_err = nullptr;
    ED_HIP_CHECK(hipMallocAsync((void **)&d_err, sizeof(int), st));
    ED_HIP_CHECK(hipMemsetAsync(d_err, 0, sizeof(int), st));
    ED_TRY(cfrac_launch_dev(n, d_a, d_b, lda, njobs, nw, d_wn, d_scale, d_out, d_err, st));
    int h_err = 0;
    ED_HIP_CHECK(hipMemcpyAsync(&h_err, d_err, sizeof(int), hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipStreamSynchronize(st));
    ED_HIP_CHECK(hipFreeAsync(d_err, st));
    if (h_err) return fail(ED_ERR_CF, "continued fraction: denominator below epsilon (matrices_tridiag.f90:174-186)");
    return ED_OK;
}

// compute_dynamic_cpu for one excitation vector already on the device.
int dynamic_dev(ed_sector *sec, int iph, double energy, const double *d_phi, double nrm2, int nitergreen,
                       double cutoff_dynamic, int prec, int nw, const double *d_wn, double *dyn, double *a_out,
                       double *b_out)
{
    if (iph != 1 && iph != 2) return fail(ED_ERR_ARG, "ed_dynamic: iph must be 1 (particle) or 2 (hole)");
    if (nitergreen < 1) return fail(ED_ERR_ARG, "ed_dynamic: nitergreen < 1");
    const int niter = (int)std::min<int64_t>(nitergreen, sec->cplx ? sec->cdim : sec->dim);
    if (nrm2 < cutoff_dynamic) {   // green_class_compute_dynamic.f90:263-265
        std::fill(dyn, dyn + 2 * (size_t)nw, 0.0);
        if (a_out) std::fill(a_out, a_out + niter, 0.0);
        if (b_out) std::fill(b_out, b_out + niter, 0.0);
        return ED_OK;
    }
    hipStream_t st = sec->stream;
    double *d_start = nullptr;
    ED_HIP_CHECK(hipMallocAsync((void **)&d_start, sizeof(double) * sec->dim, st));
    hipLaunchKernelGGL(k_scale_copy, dim3(grid_for(sec->dim)), dim3(kBlock), 0, st, d_phi, d_start, sec->dim,
                       1.0 / std::sqrt(nrm2));
    std::vector<double> a(niter), b(niter);
    int rc = lanczos_run_dev(sec, niter, d_start, prec, a.data(), b.data(), nullptr);
    (void)hipFreeAsync(d_start, st);
    if (rc != ED_OK) return rc;
    if (a_out) std::copy(a.begin(), a.end(), a_out);
    if (b_out) std::copy(b.begin(), b.end(), b_out);
    const double ph = iph == 1 ? 1.0 : -1.0;   // tri%diag = ph_sign * (diag - E), :310-311
    for (int i = 0; i < niter; ++i) a[i] = ph * (a[i] - energy);
    double *d_ab = nullptr, *d_out = nullptr, *d_sc = nullptr;
    ED_HIP_CHECK(hipMallocAsync((void **)&d_ab, sizeof(double) * (2 * niter + 1), st));
    ED_HIP_CHECK(hipMallocAsync((void **)&d_out, sizeof(double) * 2 * nw, st));
    d_sc = d_ab + 2 * niter;
    ED_HIP_CHECK(hipMemcpyAsync(d_ab, a.data(), sizeof(double) * niter, hipMemcpyHostToDevice, st));
    ED_HIP_CHECK(hipMemcpyAsync(d_ab + niter, b.data(), sizeof(double) * niter, hipMemcpyHostToDevice, st));
    // dyn = -fermion_sign * CF * ||phi||^2 with fermion_sign = -1 (:224-230, :324)
    ED_HIP_CHECK(hipMemcpyAsync(d_sc, &nrm2, sizeof(double), hipMemcpyHostToDevice, st));
    rc = cfrac_batch_dev(niter, d_ab, d_ab + niter, niter, 1, nw, d_wn, d_sc, d_out, st);
    if (rc == ED_OK) {
        ED_HIP_CHECK(hipMemcpyAsync(dyn, d_out, sizeof(double) * 2 * nw, hipMemcpyDeviceToHost, st));
        ED_HIP_CHECK(hipStreamSynchronize(st));
    }
    (void)hipFreeAsync(d_ab, st);
    (void)hipFreeAsync(d_out, st);
    (void)hipStreamSynchronize(st);
    return rc;
}

// |<n|phi>|^2 for the nst eigenvectors (rows of vec, dim each) of a small sector: one
// workgroup per eigenvector, fixed-order block sum.
__global__ __launch_bounds__(kBlock) void k_overlap2(int dim, const double *vec, const double *phi, double *w)
{
    __shared__ double red[kRed];
    const double *v = vec + (size_t)blockIdx.x * dim;
    double s = 0.0;
    for (int i = threadIdx.x; i < dim; i += kBlock) s += v[i] * phi[i];
    s = block_sum_all(s, red);
    if (threadIdx.x == 0) w[blockIdx.x] = s * s;
}

// |<n|phi>|^2 for planar complex vectors (Re[dim], Im[dim]): the real and imaginary
// parts of sum conj(v) phi by one fixed-order block sum each.
__global__ __launch_bounds__(kBlock) void k_overlap2_c(int dim, const double *vec, const double *phi, double *w)
{
    __shared__ double red[kRed];
    const double *v = vec + (size_t)blockIdx.x * 2 * dim;
    double re = 0.0, im = 0.0;
    for (int i = threadIdx.x; i < dim; i += kBlock) {
        re += v[i] * phi[i] + v[dim + i] * phi[dim + i];
        im += v[i] * phi[dim + i] - v[dim + i] * phi[i];
    }
    re = block_sum_all(re, red);
    im = block_sum_all(im, red);
    if (threadIdx.x == 0) w[blockIdx.x] = re * re + im * im;
}

// Lehmann sum of compute_dynamic_full_ed: dyn(i w_k) = sum_n w_n / (i w_k - ph (E_n - E)),
// poles in index order, one thread per frequency.
__global__ void k_lehmann(int nst, const double *e, const double *w, double e0, double ph, int nw, const double *wn,
                          double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nw) return;
    double re = 0.0, im = 0.0;
    for (int n = 0; n < nst; ++n) {
        // w / (x + i y) with x = -ph (E_n - E), y = w_k
        const double x = -ph * (e[n] - e0), y = wn[k], m2 = x * x + y * y;
        re += w[n] * x / m2;
        im -= w[n] * y / m2;
    }
    out[2 * k] = re;
    out[2 * k + 1] = im;
}

}  // namespace ed

using namespace ed;

extern "C" {

static int apply_c_host(const ed_sector *src, const ed_sector *dst, int site, int spin, int pm, const double *psi,
                        double *phi, int ncond)
{
    ED_TRY(check_pair(src, dst, site, spin, pm));
    if (!psi || !phi) return fail(ED_ERR_ARG, "ed_apply_c: NULL vector");
    ED_HIP_CHECK(hipSetDevice(src->device));
    hipStream_t st = dst->stream;
    double *d_psi = nullptr, *d_phi = nullptr;
    ED_HIP_CHECK(hipMallocAsync((void **)&d_psi, sizeof(double) * src->dim, st));
    ED_HIP_CHECK(hipMallocAsync((void **)&d_phi, sizeof(double) * dst->dim, st));
    ED_HIP_CHECK(hipMemcpyAsync(d_psi, psi, sizeof(double) * src->dim, hipMemcpyHostToDevice, st));
    int rc = apply_c_dev(src, dst, site, spin, pm, d_psi, d_phi, st, ncond);
    if (rc == ED_OK) ED_HIP_CHECK(hipMemcpyAsync(phi, d_phi, sizeof(double) * dst->dim, hipMemcpyDeviceToHost, st));
    (void)hipFreeAsync(d_psi, st);
    (void)hipFreeAsync(d_phi, st);
    ED_HIP_CHECK(hipStreamSynchronize(st));
    return rc;
}

int ed_apply_c(const ed_sector *src, const ed_sector *dst, int site, int spin, int pm, const double *psi, double *phi)
{
    return apply_c_host(src, dst, site, spin, pm, psi, phi, 0);
}

int ed_apply_c_n(const ed_sector *src, const ed_sector *dst, int site, int spin, int pm, const double *psi,
                 double *phi)
{
    return apply_c_host(src, dst, site, spin, pm, psi, phi, 1);
}

}  // extern "C"

// compute_dynamic_full_ed for real (planes = 1) or planar complex (planes = 2) vectors of
// dim amplitudes.
static int dynamic_full_ed(int dim, int planes, int iph, double energy, const double *phi, int nst,
                           const double *evals, const double *evecs, double cutoff_dynamic, int nw, const double *wn,
                           double *dyn)
{
    if (dim < 1 || nst < 0 || !phi || (nst > 0 && (!evals || !evecs)) || nw < 1 || !wn || !dyn)
        return fail(ED_ERR_ARG, "ed_dynamic_full_ed: bad argument");
    if (iph != 1 && iph != 2) return fail(ED_ERR_ARG, "ed_dynamic_full_ed: iph must be 1 (particle) or 2 (hole)");
    std::fill(dyn, dyn + 2 * (size_t)nw, 0.0);
    double nrm2 = 0.0;
    for (int i = 0; i < planes * dim; ++i) nrm2 += phi[i] * phi[i];
    // green_class_compute_dynamic.f90:117-119: below the cutoff, or no stored state of the
    // target sector (iisector = 0, :135), the contribution is zero
    if (nrm2 < cutoff_dynamic || nst == 0) return ED_OK;
    hipStream_t st = nullptr;
    double *d = nullptr;
    const size_t len = (size_t)planes * dim;
    const size_t tot = (size_t)nst * len + len + 2 * (size_t)nst + nw + 2 * (size_t)nw;
    ED_HIP_CHECK(hipMalloc((void **)&d, sizeof(double) * tot));
    double *d_vec = d, *d_phi = d_vec + (size_t)nst * len, *d_e = d_phi + len, *d_w = d_e + nst, *d_wn = d_w + nst,
           *d_out = d_wn + nw;
    int rc = ED_OK;
    if (hipMemcpy(d_vec, evecs, sizeof(double) * nst * len, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_phi, phi, sizeof(double) * len, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_e, evals, sizeof(double) * nst, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_wn, wn, sizeof(double) * nw, hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(ED_ERR_HIP, "ed_dynamic_full_ed: copy to the device failed");
    } else {
        if (planes == 1)
            hipLaunchKernelGGL(k_overlap2, dim3(nst), dim3(kBlock), 0, st, dim, (const double *)d_vec,
                               (const double *)d_phi, d_w);
        else
            hipLaunchKernelGGL(k_overlap2_c, dim3(nst), dim3(kBlock), 0, st, dim, (const double *)d_vec,
                               (const double *)d_phi, d_w);
        hipLaunchKernelGGL(k_lehmann, dim3((nw + kBlock - 1) / kBlock), dim3(kBlock), 0, st, nst, (const double *)d_e,
                           (const double *)d_w, energy, iph == 1 ? 1.0 : -1.0, nw, (const double *)d_wn, d_out);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpy(dyn, d_out, sizeof(double) * 2 * nw, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(ED_ERR_HIP, "ed_dynamic_full_ed: kernel or copy back failed");
    }
    (void)hipFree(d);
    return rc;
}

extern "C" {

int ed_dynamic_full_ed(int dim, int iph, double energy, const double *phi, int nst, const double *evals,
                       const double *evecs, double cutoff_dynamic, int nw, const double *wn, double *dyn)
{
    return dynamic_full_ed(dim, 1, iph, energy, phi, nst, evals, evecs, cutoff_dynamic, nw, wn, dyn);
}

int ed_dynamic_full_ed_c(int dim, int iph, double energy, const double *phi, int nst, const double *evals,
                         const double *evecs, double cutoff_dynamic, int nw, const double *wn, double *dyn)
{
    return dynamic_full_ed(dim, 2, iph, energy, phi, nst, evals, evecs, cutoff_dynamic, nw, wn, dyn);
}

int ed_continued_fraction(int n, const double *a, const double *b, int nw, const double *wn, double *out)
{
    if (n < 1 || nw < 1 || !a || !b || !wn || !out) return fail(ED_ERR_ARG, "ed_continued_fraction: bad argument");
    hipStream_t st = nullptr;
    double *d = nullptr;
    const size_t tot = 2 * (size_t)n + (size_t)nw + 2 * (size_t)nw;
    ED_HIP_CHECK(hipMalloc((void **)&d, sizeof(double) * tot));
    double *d_a = d, *d_b = d + n, *d_wn = d + 2 * n, *d_out = d_wn + nw;
    ED_HIP_CHECK(hipMemcpy(d_a, a, sizeof(double) * n, hipMemcpyHostToDevice));
    ED_HIP_CHECK(hipMemcpy(d_b, b, sizeof(double) * n, hipMemcpyHostToDevice));
    ED_HIP_CHECK(hipMemcpy(d_wn, wn, sizeof(double) * nw, hipMemcpyHostToDevice));
    int rc = cfrac_batch_dev(n, d_a, d_b, n, 1, nw, d_wn, nullptr, d_out, st);
    if (rc == ED_OK) ED_HIP_CHECK(hipMemcpy(out, d_out, sizeof(double) * 2 * nw, hipMemcpyDeviceToHost));
    (void)hipFree(d);
    return rc;
}

int ed_dynamic(ed_sector *sec, int iph, double energy, const double *phi, int nitergreen, double cutoff_dynamic,
               int precision, int nw, const double *wn, double *dyn, double *a, double *b)
{
    if (!sec || !phi || !wn || !dyn || nw < 1) return fail(ED_ERR_ARG, "ed_dynamic: bad argument");
    if (precision != ED_PREC_F64 && precision != ED_PREC_F32) return fail(ED_ERR_ARG, "ed_dynamic: bad precision");
    ED_HIP_CHECK(hipSetDevice(sec->device));
    double nrm2 = 0.0;
    nrm2 = host_norm2(phi, sec->dim);
    hipStream_t st = sec->stream;
    double *d_phi = nullptr, *d_wn = nullptr;
    ED_HIP_CHECK(hipMallocAsync((void **)&d_phi, sizeof(double) * sec->dim, st));
    ED_HIP_CHECK(hipMallocAsync((void **)&d_wn, sizeof(double) * nw, st));
    ED_HIP_CHECK(hipMemcpyAsync(d_phi, phi, sizeof(double) * sec->dim, hipMemcpyHostToDevice, st));
    ED_HIP_CHECK(hipMemcpyAsync(d_wn, wn, sizeof(double) * nw, hipMemcpyHostToDevice, st));
    int rc = dynamic_dev(sec, iph, energy, d_phi, nrm2, nitergreen, cutoff_dynamic, precision, nw, d_wn, dyn, a, b);
    (void)hipFreeAsync(d_phi, st);
    (void)hipFreeAsync(d_wn, st);
    (void)hipStreamSynchronize(st);
    return rc;
}

}  // extern "C"
