// H.v kernels and the device-resident Lanczos engine.
//
// H.v (HAIMupdo_multr, HAIMupdo.f90:677-1195), gather form, on the column-major
// d_up x d_dn coefficient matrix V (istate = iup + ido*d_up):
//
//   W(:, ido) = [diag_up + diag_dn(ido) + Uimp(imp_up, imp_dn(ido))] o V(:, ido)   (i)-(iii)
//             + T_up V(:, ido)                                                   (iv) up hops
//             + sum_k off_dn(ido, k) V(:, jdo_k)                                 (v)  down hops
//             + sum_t J_t V(swap_up(:, cu_t), swap_dn(ido, cd_t))                (vii)-(viii) Hund
//
// A workgroup owns CPB whole columns: they are staged in LDS once; each thread owns
// rows of those columns, loads the row's up-hop entries (ELL, width unrolled at compile
// time) once and applies them to all CPB columns from LDS; the down hops stream whole
// neighbour columns (uniform per column, coalesced over rows); the Hund terms gather
// single elements.  No atomics: every output element is written by exactly one thread.
// CPB is picked per sector by a small autotuner (1, 2 or 4 columns per workgroup).
//
// Lanczos (lanc_opt, Lanczos_plain.h:66-173) runs as two kernels per step on
// device-resident scalars (LzState), so a chunk of steps runs from a captured graph
// without the host:
//   k_hv<MODE=1>:   beta_j = ||u_j||; w = H(u_j)/beta_j - (beta_j/beta_{j-1}) u_{j-1};
//                   partial sums of w.v_j
//   k_lz_update:    alpha_j = w.v_j; w -= alpha_j v_j; partial sums of ||w||^2
// with u_j the un-normalised Lanczos vectors (v_j = u_j / beta_j) kept in a Krylov basis
// on the device.  Each kernel writes one partial sum per workgroup; the NEXT kernel
// finishes that reduction in every workgroup, in the same fixed order (so every
// workgroup obtains the same bits).  This avoids cross-XCD atomics and release fences,
// which cost more than the reduction itself on MI355X.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <type_traits>
#include <vector>

#include "ed_kernels.h"

namespace ed {

__device__ inline bool block0() { return blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0; }

// ------------------------------------------------------------------ H.v kernels

// Row-owner kernel: CPB columns per workgroup, ELL width KU unrolled.
// MODE 0: y = H x for grid.y vectors (leading dimension ld).
// MODE 1: Lanczos phase 1 on the Krylov slots of `lz` (x = u_j, y = u_{j+1}).
template <typename T, int MODE, int CPB, int KU>
__global__ __launch_bounds__(kBlock) void k_hv_rows(HvArgs<T> a, const T *__restrict__ xin, T *__restrict__ yout,
                                                     int64_t ld, LzArgs<T> lz)
{
    extern __shared__ double smem_d[];
    __shared__ double red[kRed];
    __shared__ double grp[kSums];
    const int64_t off = MODE ? 0 : (int64_t)blockIdx.y * ld;
    hv_rows_body<T, MODE, CPB, KU, false>(a, MODE ? xin : xin + off, MODE ? yout : yout + off, lz, blockIdx.x,
                                          reinterpret_cast<T *>(smem_d), red, grp);
}

// Fast kernel (one column per workgroup, blockDim.x up to 1024), see hv_fast_body.
// Largest workgroup of the fast kernel with R rows per thread: 1024 / R threads (R >= 2
// keeps the register budget of 1024 / R threads, so R rows of loads stay in registers).
__host__ __device__ constexpr int fast_max_threads(int R) { return R > 1 ? (1024 / R + 63) / 64 * 64 : 1024; }

template <typename T, int MODE, int K, int R>
__global__ __launch_bounds__(fast_max_threads(R)) void k_hv_fast(HvArgs<T> a, const T *__restrict__ xin, T *__restrict__ yout,
                                                   int64_t ld, LzArgs<T> lz)
{
    extern __shared__ double smem_d[];
    __shared__ double red[kRed];
    __shared__ double grp[kSums];
    const int64_t off = MODE ? 0 : (int64_t)blockIdx.y * ld;
    hv_fast_body<T, MODE, K, R>(a, MODE ? xin : xin + off, MODE ? yout : yout + off, lz, blockIdx.x,
                                  reinterpret_cast<unsigned char *>(smem_d), red, grp);
}

// Generic kernel for short columns (d_up < 128: several columns per workgroup, element-
// strided) and for ELL widths above the compiled ones.
template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_hv_generic(HvArgs<T> a, const T *__restrict__ xin, T *__restrict__ yout,
                                                        int64_t ld, LzArgs<T> lz)
{
    extern __shared__ double smem_d[];
    T *col = reinterpret_cast<T *>(smem_d);
    __shared__ double red[kRed];
    __shared__ double grp[kSums];
    const int dup = a.dup, ncfg = a.ncfg;
    const int c0 = blockIdx.x * a.cpb;
    const int ncol = min(a.cpb, a.ddn - c0);
    const T *__restrict__ x;
    T *__restrict__ y;
    int j = 0;
    if (MODE == 1) {
        if (!lz_hv_begin(lz, block0(), j)) return;
        x = lz.slot(j);
        y = lz.slot(j + 1);
    } else {
        x = xin + (int64_t)blockIdx.y * ld;
        y = yout + (int64_t)blockIdx.y * ld;
    }
    const int base = c0 * dup;
    const int nel = ncol * dup;
    for (int e = threadIdx.x; e < nel; e += kBlock) col[e] = x[base + e];
    __syncthreads();
    double dot = 0.0;
    for (int e = threadIdx.x; e < nel; e += kBlock) {
        const int cl = e / dup, r = e - cl * dup, ido = c0 + cl;
        const T *cv = col + cl * dup;
        const int cu = a.imp_up[r], cd = a.imp_dn[ido];
        const T xv = cv[r];
        T acc = (a.diag_up[r] + a.diag_dn[ido] + a.uimp[cu * ncfg + cd]) * xv;
        for (int k = 0; k < a.kup; ++k) acc += a.ell_val_up[k * dup + r] * cv[a.ell_rk_up[k * dup + r]];
        for (int k = a.ptr_dn[ido]; k < a.ptr_dn[ido + 1]; ++k) acc += a.off_dn[k] * x[r + a.rk_dn[k] * dup];
        if (a.hund) {
            const int h0 = a.hund_ptr[cu * ncfg + cd], h1 = a.hund_ptr[cu * ncfg + cd + 1];
            for (int t = h0; t < h1; ++t)
                acc += a.hund_val[t] * x[a.swap_up[r * ncfg + a.hund_cu[t]] + a.swap_dn[ido * ncfg + a.hund_cd[t]] * dup];
        }
        y[base + e] = acc;
        if (MODE == 1) dot += (double)acc * (double)xv;
    }
    if (MODE == 1) block_partial(dot, red, lz.partA + blockIdx.x);
}

// Interleaved complex (re, im pairs: Fortran COMPLEX(8)) <-> planar (Re[n], Im[n]).
__global__ void k_planar(const double *in, double *out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double re = in[2 * i], im = in[2 * i + 1];
        out[i] = re;
        out[n + i] = im;
    }
}

__global__ void k_interleaved(const double *in, double *out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double re = in[i], im = in[n + i];
        out[2 * i] = re;
        out[2 * i + 1] = im;
    }
}

// Complex sectors: HAIMupdo_multc, one thread per amplitude (FP64).
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_hv_cplx(CplxArgs a, const double *__restrict__ xin,
                                                     double *__restrict__ yout, int64_t ld, LzArgs<double> lz)
{
    __shared__ double red[kRed];
    const int64_t off = MODE ? 0 : (int64_t)blockIdx.y * ld;
    hv_cplx_body<MODE>(a, MODE ? xin : xin + off, MODE ? yout : yout + off, lz, blockIdx.x, red);
}

// S_z sectors, on-the-fly entries (HAIMsz_mult_fly), one thread per row.
template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_hv_sz_fly(SzArgs<T> a, SpinModel m, const int32_t *states,
                                                       const int32_t *rank, const T *__restrict__ xin,
                                                       T *__restrict__ yout, int64_t ld, LzArgs<T> lz)
{
    __shared__ double red[kRed];
    const int64_t off = MODE ? 0 : (int64_t)blockIdx.y * ld;
    hv_sz_fly_body<T, MODE>(a, m, states, rank, MODE ? xin : xin + off, MODE ? yout : yout + off, lz, blockIdx.x, red);
}

// Complex S_z sectors: HAIMsz_multc, one thread per amplitude (FP64).
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_hv_sz_cplx(SzArgs<double> a, const double *__restrict__ xin,
                                                        double *__restrict__ yout, int64_t ld, LzArgs<double> lz)
{
    __shared__ double red[kRed];
    const int64_t off = MODE ? 0 : (int64_t)blockIdx.y * ld;
    hv_sz_cplx_body<MODE>(a, MODE ? xin : xin + off, MODE ? yout : yout + off, lz, blockIdx.x, red);
}

// Complex S_z sectors, on-the-fly entries (HAIMsz_mult_fly_c).
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_hv_sz_cplx_fly(SzArgs<double> a, SpinModel m, const int32_t *states,
                                                            const int32_t *rank, const double *__restrict__ xin,
                                                            double *__restrict__ yout, int64_t ld, LzArgs<double> lz)
{
    __shared__ double red[kRed];
    const int64_t off = MODE ? 0 : (int64_t)blockIdx.y * ld;
    hv_sz_cplx_fly_body<MODE>(a, m, states, rank, MODE ? xin : xin + off, MODE ? yout : yout + off, lz, blockIdx.x,
                              red);
}

// S_z sectors: HAIMsz_multr, one thread per row.
template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_hv_sz(SzArgs<T> a, const T *__restrict__ xin, T *__restrict__ yout,
                                                   int64_t ld, LzArgs<T> lz)
{
    __shared__ double red[kRed];
    const int64_t off = MODE ? 0 : (int64_t)blockIdx.y * ld;
    hv_sz_body<T, MODE>(a, MODE ? xin : xin + off, MODE ? yout : yout + off, lz, blockIdx.x, red);
}

// ------------------------------------------------------------- Lanczos kernels

template <typename T>
__global__ __launch_bounds__(kBlock) void k_lz_init(LzArgs<T> lz, const double *start, int64_t n)
{
    __shared__ double red[kRed];
    __shared__ double grp[kSums];
    lz_init_body(lz, start, n, blockIdx.x, gridDim.x, red);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_lz_update(LzArgs<T> lz, int64_t n, int advance)
{
    __shared__ double red[kRed];
    __shared__ double grp[kSums];
    lz_update_body(lz, n, advance, blockIdx.x, gridDim.x, red, grp);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_lz_finish(LzArgs<T> lz)
{
    __shared__ double red[kRed];
    __shared__ double grp[kSums];
    lz_finish_body(lz, grp);
}

// Persistent Lanczos: nsteps steps of the fast H.v and the update in one cooperative
// launch (see lz_persist_body).  256 threads, at most 4 workgroups per CU resident.
// MON: the monitored form (LzMonitor), its own instance so that the plain one carries no
// monitor registers.
template <typename T, int K, int R, bool MON>
__global__ __launch_bounds__(kBlock, 4) void k_lz_persist(HvArgs<T> a, LzArgs<T> lz, int nsteps, int spin_limit,
                                                          LzMonitor mon)
{
    extern __shared__ double smem_d[];
    __shared__ double red[kRed];
    __shared__ double grp[kSums];
    __shared__ double pub[2];
    __shared__ int flag;
    lz_persist_body<T, K, R, MON>(a, lz, nsteps, spin_limit, mon, reinterpret_cast<unsigned char *>(smem_d), red,
                                  grp, pub, &flag);
}

// Persistent Lanczos with pm.cpw columns per workgroup of 256 pm.cpw threads (see
// lz_persist_cols_body); one workgroup per CU at cpw = 4.
template <typename T, int K, int R, bool MON>
__global__ __launch_bounds__(1024, 1) void k_lz_persist_cols(HvArgs<T> a, LzArgs<T> lz, PersistMap pm, int nsteps,
                                                             int spin_limit, LzMonitor mon)
{
    extern __shared__ double smem_d[];
    __shared__ double red[kRed];
    __shared__ double grp[kSums];
    __shared__ double pub[2];
    __shared__ int flag;
    lz_persist_cols_body<T, K, R, MON>(a, lz, pm, nsteps, spin_limit, mon, reinterpret_cast<unsigned char *>(smem_d),
                                       red, grp, pub, &flag);
}

// Persistent Lanczos with one grid barrier per step (see lz_persist_1b_body): the
// multi-column kernel's map and LDS layout, two exchange vectors at qx (leading dimension
// lz.ld) and the partial sums at p1 (kP1Arrays x 2 x kP1Stride doubles).
template <typename T, int K, int R, bool MON>
__global__ __launch_bounds__(1024, 1) void k_lz_persist_1b(HvArgs<T> a, LzArgs<T> lz, PersistMap pm, int nsteps,
                                                           int spin_limit, LzMonitor mon, T *qx, double *p1,
                                                           int exact_beta)
{
    extern __shared__ double smem_d[];
    __shared__ double red[kP1Arrays * 16 + 2];
    __shared__ double sh[8];
    __shared__ int flag;
    lz_persist_1b_body<T, K, R, MON>(a, lz, pm, nsteps, spin_limit, mon, qx, p1, exact_beta,
                                     reinterpret_cast<unsigned char *>(smem_d), red, sh, &flag);
}

__global__ void k_bar_reset(LzState *st)
{
    for (int g = threadIdx.x; g < kBarGroups; g += blockDim.x) {
        st->bar_grp[16 * g] = 0;
        st->bar_rel[16 * g] = 0;
    }
    if (threadIdx.x == 0) {
        st->bar_count = 0;
        st->bar_err = 0;
        st->stop_at = 0;
        st->exact_beta_steps = 0;
    }
}

// Full reorthogonalisation (classical Gram-Schmidt, applied twice): partial dots of
// w = u_{j+1} with every kept u_k, k <= j.  One workgroup per slice of the vectors.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_reorth_dots(LzArgs<T> lz, int64_t n, double *partR)
{
    __shared__ double red[kRed];
    __shared__ double grp[kSums];
    if (lz.st->status) return;
    const int j = lz.st->j_k2;
    const T *w = lz.slot(j + 1);
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * per, hi = min(n, lo + per);
    for (int k = 1; k <= j; ++k) {
        const T *v = lz.slot(k);
        double acc = 0.0;
        for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) acc += (double)w[i] * (double)v[i];
        acc = block_sum_all(acc, red);
        if (threadIdx.x == 0) partR[(int64_t)(k - 1) * gridDim.x + blockIdx.x] = acc;
    }
}

// coef[k-1] = (w . u_k) / beta_k^2, one workgroup, fixed summation order.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_reorth_coef(LzArgs<T> lz, const double *partR, int nr, double *coef)
{
    if (lz.st->status) return;
    const int j = lz.st->j_k2;
    for (int k = 1 + threadIdx.x; k <= j; k += blockDim.x) {
        double s = 0.0;
        for (int b = 0; b < nr; ++b) s += partR[(int64_t)(k - 1) * nr + b];
        const double bk = lz.st->beta[k - 1];
        coef[k - 1] = s / (bk * bk);
    }
}

// w -= sum_k coef[k] u_k; partial sums of ||w||^2 and <w, u_j> for step j+1; advances
// j when `advance`.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_reorth_update(LzArgs<T> lz, int64_t n, const double *coef, int advance)
{
    __shared__ double red[kRed];
    __shared__ int cnt;
    if (lz.st->status) return;
    const int j = lz.st->j_k2;
    T *w = lz.slot(j + 1);
    const T *uj = lz.slot(j);
    int64_t i0, i1, st;
    elem_range(lz, n, blockIdx.x, gridDim.x, i0, i1, st);
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    double nn = 0.0, cr = 0.0;
    for (int64_t i = i0; i < i1; i += st) {
        double v = (double)w[i];
        for (int k = 1; k <= j; ++k) v -= coef[k - 1] * (double)lz.slot(k)[i];
        const T vt = (T)v;
        w[i] = vt;
        nn += (double)vt * (double)vt;
        cr += (double)vt * (double)uj[i];
    }
    block_partial_last(nn, cr, red, &cnt, lz.nrm(j + 1) + blockIdx.x, lz.crs(j + 1) + blockIdx.x);
    if (advance && blockIdx.x == 0 && threadIdx.x == 0) lz.st->j_k1 = j + 1;
}

// psi = sum_{k<m} coef[k] u_{k+1} / beta_{k+1} (FP64 output).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_combine(LzArgs<T> lz, int m, const double *coef, double *psi, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        double s = 0.0;
        for (int k = 0; k < m; ++k) s += coef[k] / lz.st->beta[k] * (double)lz.slot(k + 1)[i];
        psi[i] = s;
    }
}

// Two-kernel deterministic dot product: partial sums, then one workgroup sums them.
__global__ __launch_bounds__(kBlock) void k_dot_part(const double *x, const double *y, int64_t n, double *partials)
{
    __shared__ double red[kRed];
    __shared__ double grp[kSums];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        acc += x[i] * y[i];
    const double p = block_sum_all(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = p;
}

__global__ __launch_bounds__(kBlock) void k_sum_part(const double *partials, int n, double *out)
{
    __shared__ double red[kRed];
    __shared__ double grp[kSums];
    const double s = sum_partials(partials, n, red);
    if (threadIdx.x == 0) *out = s;
}

__global__ void k_scale(double *x, int64_t n, const double *nrm2)
{
    const double s = 1.0 / sqrt(*nrm2);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] *= s;
}

__global__ void k_start_vector(int64_t n, double *out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t x = ((uint64_t)(i % 10000) + 1ull) * 0x9E3779B97F4A7C15ull;
        x ^= x >> 29;
        x *= 0xBF58476D1CE4E5B9ull;
        x ^= x >> 32;
        const uint64_t m = x >> 11;
        const uint64_t k = (m * 1000ull) >> 53;
        out[i] = (double)(k + 1ull) / 1000.0;
    }
}

// ------------------------------------------------------------------ host side

template <typename T>
static HvArgs<T> hv_args(const ed_sector *s, int cpb_rows)
{
    const bool f64 = sizeof(T) == sizeof(double);
    HvArgs<T> a;
    a.dup = s->up.d; a.ddn = s->dn.d; a.ncfg = s->ncfg; a.kup = s->up.kpad;
    a.cpb = cpb_rows;
    a.hund = s->hund;
    a.ell_rk_up = s->up.ell_rk; a.ell_val_up = (const T *)(f64 ? (void *)s->up.ell_val : (void *)s->up.ell_val32);
    a.ptr_dn = s->dn.ptr; a.rk_dn = s->dn.rk; a.off_dn = (const T *)(f64 ? (void *)s->dn.off : (void *)s->dn.off32);
    a.diag_up = (const T *)(f64 ? (void *)s->up.diag : (void *)s->up.diag32);
    a.diag_dn = (const T *)(f64 ? (void *)s->dn.diag : (void *)s->dn.diag32);
    a.imp_up = s->up.imp; a.imp_dn = s->dn.imp;
    a.uimp = (const T *)(f64 ? (void *)s->d_uimp : (void *)s->d_uimp32);
    a.hund_ptr = s->d_hund_ptr; a.hund_cu = s->d_hund_cu; a.hund_cd = s->d_hund_cd;
    a.hund_val = (const T *)(f64 ? (void *)s->d_hund_val : (void *)s->d_hund_val32);
    a.swap_up = s->up.swap; a.swap_dn = s->dn.swap;
    a.ell_pk_up = s->up.ell_pk;
    a.coef_up = (const T *)(f64 ? (void *)s->up.coef : (void *)s->up.coef32);
    a.ncoef_up = s->up.ncoef;
    a.kdn = s->dn.kpad;
    a.nhund = s->nhund;
    a.hund_tr = s->d_hund_tr;
    a.colord = nullptr;
    return a;
}

static CplxArgs cplx_args(const ed_sector *s)
{
    CplxArgs a;
    a.dup = s->up.d; a.ddn = s->dn.d; a.ncfg = s->ncfg; a.kup = s->up.kpad; a.kdn = s->dn.kpad;
    a.hund = s->hund;
    a.D = s->cdim;
    a.rk_up = s->up.ell_rk; a.rk_dn = s->dn.ell_rk;
    a.vr_up = s->up.ell_val; a.vi_up = s->up.ell_vim; a.vr_dn = s->dn.ell_val; a.vi_dn = s->dn.ell_vim;
    a.diag_up = s->up.diag; a.diag_dn = s->dn.diag; a.uimp = s->d_uimp;
    a.imp_up = s->up.imp; a.imp_dn = s->dn.imp;
    a.hund_ptr = s->d_hund_ptr; a.hund_tr = s->d_hund_tr; a.hund_val = s->d_hund_val;
    a.swap_up = s->up.swap; a.swap_dn = s->dn.swap;
    return a;
}

HvArgs<double> hv_args_f64(const ed_sector *s) { return hv_args<double>(s, 1); }
HvArgs<float> hv_args_f32(const ed_sector *s) { return hv_args<float>(s, 1); }

// Launch shape of one H.v: which kernel, columns per workgroup, threads, grid.
struct HvShape {
    int kind;    // 0 generic, 1 row kernel, 2 fast kernel, 3 S_z sector, 4 complex sector
    int cpb;
    int nblk;
    int nt;      // threads per workgroup
    int k;       // compiled table width
    size_t lds;
    int xcd;     // fast kernel over the XCD-grouped column order (s->d_colord)
    int rows;    // fast kernel: rows per thread with their loads in flight together (0 = loop)
};

static int width_of(int kmax)
{
    const int widths[] = {4, 8, 12, 16, 24, 32};
    for (int w : widths)
        if (kmax <= w) return w;
    return 0;
}

static bool fast_ok(const ed_sector *s)
{
    if (s->sz || s->cplx) return false;
    const int k = width_of(std::max(s->up.kpad, s->dn.kpad));
    // widths above 16 exceed the 128-VGPR budget of a 1024-thread workgroup in FP64
    return s->up.ncoef > 0 && k > 0 && k <= 16 && s->up.d >= 64 && s->up.d <= 65535 && s->dn.d <= 8184 &&
           hv_fast_lds(s->up.d, s->up.ncoef, s->ncfg, s->nhund, sizeof(double)) <= 56 * 1024;   // + 4.2 KB static
}

// sched: 1, 2, 4 = row kernel with that many columns per workgroup; 8 = fast kernel;
// 16 = fast kernel with the XCD-grouped column order (needs s->d_colord); 18, 19, 20 =
// that kernel with 2, 3, 4 rows per thread (smaller workgroups, all rows' loads in flight).
static bool sched_valid(int sched)
{
    return sched == 1 || sched == 2 || sched == 4 || sched == 8 || sched == 16 || (sched >= 18 && sched <= 20);
}

static HvShape hv_shape(const ed_sector *s, int sched)
{
    HvShape h;
    const int dup = s->up.d, ddn = s->dn.d;
    h.nt = kBlock;
    h.k = s->up.kpad;
    h.xcd = 0;
    h.rows = 0;
    if (s->cplx && !s->sz) {   // complex sector: one amplitude per thread, grid-stride
        h.kind = 4;
        h.cpb = 1;
        h.nblk = (int)std::min<int64_t>(kMaxRowBlocks, (s->cdim + kBlock - 1) / kBlock);
        h.lds = 0;
        return h;
    }
    if (s->sz) {   // S_z sector: one row (complex: amplitude) per thread, grid-stride beyond kMaxRowBlocks
        h.kind = 3;
        h.cpb = 1;
        h.nblk = (int)std::min<int64_t>(kMaxRowBlocks, ((s->cplx ? s->cdim : s->dim) + kBlock - 1) / kBlock);
        h.lds = 0;
        return h;
    }
    const int rows = sched >= 18 ? sched - 16 : 1;
    const int nt_rows = std::max(64, ((dup + rows - 1) / rows + 63) / 64 * 64);
    if ((sched == 8 || sched >= 16) && fast_ok(s) && (rows == 1 || nt_rows <= fast_max_threads(rows))) {
        h.kind = 2;
        h.cpb = 1;
        h.rows = rows == 1 ? (dup <= 1024 ? 1 : 0) : rows;
        h.nt = rows == 1 ? std::min(1024, std::max(64, (dup + 63) / 64 * 64)) : nt_rows;
        h.k = width_of(std::max(s->up.kpad, s->dn.kpad));
        h.xcd = sched >= 16 && s->d_colord;
        h.nblk = h.xcd ? s->colord_blocks : ddn;
        h.lds = 0;   // set per precision at launch
        return h;
    }
    h.kind = (dup >= 128 && s->up.kpad <= 32) ? 1 : 0;
    if (h.kind == 1) {
        h.cpb = (sched == 2 || sched == 4) ? sched : 1;
        // LDS holds CPB FP64 columns and the packed coefficients; stay within 64 KB
        const size_t nco = (s->up.ell_pk && s->up.ncoef <= kMaxRowsCoef) ? (size_t)s->up.ncoef : 0;
        while (h.cpb > 1 && ((size_t)h.cpb * dup + nco) * sizeof(double) > 56 * 1024) h.cpb /= 2;
    } else {
        h.cpb = dup >= kBlock ? 1 : std::max(1, kBlock / std::max(1, dup));
    }
    h.nblk = (ddn + h.cpb - 1) / h.cpb;
    h.lds = 0;
    return h;
}

template <typename T, int MODE, int R>
static void launch_fast_rows(const HvArgs<T> &a, int k, dim3 grid, int nt, size_t lds, hipStream_t st, const T *x,
                             T *y, int64_t ld, const LzArgs<T> &lz)
{
    switch (k) {
    case 4: hipLaunchKernelGGL((k_hv_fast<T, MODE, 4, R>), grid, dim3(nt), lds, st, a, x, y, ld, lz); break;
    case 8: hipLaunchKernelGGL((k_hv_fast<T, MODE, 8, R>), grid, dim3(nt), lds, st, a, x, y, ld, lz); break;
    case 12: hipLaunchKernelGGL((k_hv_fast<T, MODE, 12, R>), grid, dim3(nt), lds, st, a, x, y, ld, lz); break;
    default: hipLaunchKernelGGL((k_hv_fast<T, MODE, 16, R>), grid, dim3(nt), lds, st, a, x, y, ld, lz); break;
    }
}

template <typename T, int MODE>
static void launch_fast(const HvArgs<T> &a, int k, int rows, dim3 grid, int nt, size_t lds, hipStream_t st,
                        const T *x, T *y, int64_t ld, const LzArgs<T> &lz)
{
    switch (rows) {
    case 1: launch_fast_rows<T, MODE, 1>(a, k, grid, nt, lds, st, x, y, ld, lz); break;
    case 2: launch_fast_rows<T, MODE, 2>(a, k, grid, nt, lds, st, x, y, ld, lz); break;
    case 3: launch_fast_rows<T, MODE, 3>(a, k, grid, nt, lds, st, x, y, ld, lz); break;
    case 4: launch_fast_rows<T, MODE, 4>(a, k, grid, nt, lds, st, x, y, ld, lz); break;
    default: launch_fast_rows<T, MODE, 0>(a, k, grid, nt, lds, st, x, y, ld, lz); break;
    }
}

template <typename T, int MODE, int CPB>
static void launch_rows_ku(const HvArgs<T> &a, int ku, dim3 grid, size_t lds, hipStream_t st, const T *x, T *y,
                           int64_t ld, const LzArgs<T> &lz)
{
    switch (ku) {
    case 4: hipLaunchKernelGGL((k_hv_rows<T, MODE, CPB, 4>), grid, dim3(kBlock), lds, st, a, x, y, ld, lz); break;
    case 8: hipLaunchKernelGGL((k_hv_rows<T, MODE, CPB, 8>), grid, dim3(kBlock), lds, st, a, x, y, ld, lz); break;
    case 12: hipLaunchKernelGGL((k_hv_rows<T, MODE, CPB, 12>), grid, dim3(kBlock), lds, st, a, x, y, ld, lz); break;
    case 16: hipLaunchKernelGGL((k_hv_rows<T, MODE, CPB, 16>), grid, dim3(kBlock), lds, st, a, x, y, ld, lz); break;
    case 24: hipLaunchKernelGGL((k_hv_rows<T, MODE, CPB, 24>), grid, dim3(kBlock), lds, st, a, x, y, ld, lz); break;
    default: hipLaunchKernelGGL((k_hv_rows<T, MODE, CPB, 32>), grid, dim3(kBlock), lds, st, a, x, y, ld, lz); break;
    }
}

template <typename T, int MODE>
static void hv_launch_mode(const ed_sector *s, const HvShape &h, const T *x, T *y, int64_t ld, int nvec,
                           const LzArgs<T> &lz, hipStream_t st)
{
    if (h.kind == 4) {   // complex sectors run FP64 only (check_common)
        if constexpr (std::is_same<T, double>::value)
            hipLaunchKernelGGL((k_hv_cplx<MODE>), dim3(h.nblk, MODE ? 1 : nvec), dim3(kBlock), 0, st, cplx_args(s), x,
                               y, ld, lz);
        return;
    }
    if (h.kind == 3) {
        const bool f64 = std::is_same<T, double>::value;
        SzArgs<T> za;
        za.d = s->dim;
        za.kw = s->up.kpad;
        za.rk = s->up.ell_rk;
        za.val = (const T *)(f64 ? (const void *)s->up.ell_val : (const void *)s->up.ell_val32);
        za.diag = (const T *)(f64 ? (const void *)s->up.diag : (const void *)s->up.diag32);
        za.quart = (const T *)(f64 ? (const void *)s->d_quart : (const void *)s->d_quart32);
        za.kh = s->sz_kh;
        za.hrk = s->d_sz_hrk;
        za.hval = (const T *)(f64 ? (const void *)s->d_sz_hval : (const void *)s->d_sz_hval32);
        za.vim = nullptr;
        if (s->cplx) {   // complex S_z sectors run FP64 only (check_common)
            if constexpr (std::is_same<T, double>::value) {
                za.d = s->cdim;
                za.vim = s->up.ell_vim;
                if (s->sz_fly)
                    hipLaunchKernelGGL((k_hv_sz_cplx_fly<MODE>), dim3(h.nblk, MODE ? 1 : nvec), dim3(kBlock), 0, st,
                                       za, s->sz_model, (const int32_t *)s->up.states, (const int32_t *)s->d_rank, x,
                                       y, ld, lz);
                else
                    hipLaunchKernelGGL((k_hv_sz_cplx<MODE>), dim3(h.nblk, MODE ? 1 : nvec), dim3(kBlock), 0, st, za,
                                       x, y, ld, lz);
            }
            return;
        }
        if (s->sz_fly) {
            hipLaunchKernelGGL((k_hv_sz_fly<T, MODE>), dim3(h.nblk, MODE ? 1 : nvec), dim3(kBlock), 0, st, za,
                               s->sz_model, (const int32_t *)s->up.states, (const int32_t *)s->d_rank, x, y, ld, lz);
            return;
        }
        hipLaunchKernelGGL((k_hv_sz<T, MODE>), dim3(h.nblk, MODE ? 1 : nvec), dim3(kBlock), 0, st, za, x, y, ld, lz);
        return;
    }
    HvArgs<T> a = hv_args<T>(s, h.cpb);
    if (h.xcd) a.colord = s->d_colord;
    dim3 grid(h.nblk, MODE ? 1 : nvec);
    if (h.kind == 2) {
        const size_t lds = hv_fast_lds(a.dup, a.ncoef_up, a.ncfg, a.nhund, sizeof(T));
        launch_fast<T, MODE>(a, h.k, h.rows, grid, h.nt, lds, st, x, y, ld, lz);
        return;
    }
    // row kernels: packed up hops when the coefficient table is small enough for LDS
    if (a.ncoef_up > kMaxRowsCoef) a.ell_pk_up = nullptr;
    const size_t lds = ((size_t)h.cpb * a.dup + (a.ell_pk_up ? a.ncoef_up : 0)) * sizeof(T);
    if (h.kind == 0) {
        hipLaunchKernelGGL((k_hv_generic<T, MODE>), grid, dim3(kBlock), lds, st, a, x, y, ld, lz);
        return;
    }
    switch (h.cpb) {
    case 2: launch_rows_ku<T, MODE, 2>(a, a.kup, grid, lds, st, x, y, ld, lz); break;
    case 4: launch_rows_ku<T, MODE, 4>(a, a.kup, grid, lds, st, x, y, ld, lz); break;
    default: launch_rows_ku<T, MODE, 1>(a, a.kup, grid, lds, st, x, y, ld, lz); break;
    }
}

static int elem_grid(int64_t n)
{
    int64_t g = (n + kBlock * 4 - 1) / (kBlock * 4);
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 2048));
}

// Schedule of the kernels of precision T: FP32 has its own when the autotuner set one.
template <typename T>
static int sched_of(const ed_sector *s)
{
    return (sizeof(T) == sizeof(float) && s->cpb32) ? s->cpb32 : s->cpb;
}

// Schedule of the Lanczos recurrence (two-kernel step and persistent kernel) in precision T:
// the H.v schedule unless the tuner chose another one for it (lanczos_sched).
template <typename T>
static int sched_lz(const ed_sector *s)
{
    const int l = sizeof(T) == sizeof(float) ? s->lzs32 : s->lzs;
    return l ? l : sched_of<T>(s);
}

int launch_hv(ed_sector *s, int prec, const void *x, void *y, int64_t ld, int nvec, hipStream_t st)
{
    const HvShape h = hv_shape(s, prec == ED_PREC_F32 ? sched_of<float>(s) : sched_of<double>(s));
    if (prec == ED_PREC_F64) {
        LzArgs<double> lz{};
        hv_launch_mode<double, 0>(s, h, (const double *)x, (double *)y, ld, nvec, lz, st);
    } else {
        LzArgs<float> lz{};
        hv_launch_mode<float, 0>(s, h, (const float *)x, (float *)y, ld, nvec, lz, st);
    }
    ED_HIP_CHECK(hipGetLastError());
    return ED_OK;
}

int start_vector_dev(int64_t n, double *d_out, hipStream_t st)
{
    hipLaunchKernelGGL(k_start_vector, dim3(elem_grid(n)), dim3(kBlock), 0, st, n, d_out);
    ED_HIP_CHECK(hipGetLastError());
    return ED_OK;
}

static int ensure_state(ed_sector *s)
{
    if (!s->d_state) {
        // Cleared on the sector's own stream: that stream is non-blocking, so a clear on
        // the null stream (hipMemset) is not ordered before the kernels queued on it and
        // could land after the first Lanczos initialisation has written its partials.
        ED_HIP_CHECK(hipMalloc((void **)&s->d_state, sizeof(LzState)));
        ED_HIP_CHECK(hipMemsetAsync(s->d_state, 0, sizeof(LzState), s->stream));
        ED_HIP_CHECK(hipMalloc((void **)&s->d_partials, sizeof(double) * (size_t)kScratchDoubles));
        ED_HIP_CHECK(hipMemsetAsync(s->d_partials, 0, sizeof(double) * (size_t)kScratchDoubles, s->stream));
    }
    return ED_OK;
}

static int ensure_io(ed_sector *s)
{
    if (!s->d_io_a) {
        ED_HIP_CHECK(hipMalloc((void **)&s->d_io_a, sizeof(double) * std::max<int64_t>(1, s->dim)));
        ED_HIP_CHECK(hipMalloc((void **)&s->d_io_b, sizeof(double) * std::max<int64_t>(1, s->dim)));
    }
    return ED_OK;
}

static void drop_graphs(ed_sector *s)
{
    for (GraphCache &G : s->graphs)
        if (G.exec) {
            (void)hipGraphExecDestroy(G.exec);
            (void)hipGraphDestroy(G.graph);
            G = GraphCache();
        }
}

// Krylov basis with nslots vectors (leading dimension padded to 64 elements), sized for
// FP64 so that FP32 and FP64 runs share one allocation.
static int ensure_basis(ed_sector *s, int prec, int nslots)
{
    const int64_t ld = (s->dim + 63) / 64 * 64;
    const size_t need = sizeof(double) * (size_t)ld * nslots;
    if (s->d_basis && s->basis_bytes >= need) {
        s->basis_prec = prec;
        s->basis_nslots = nslots;
        s->ld = ld;
        return ED_OK;
    }
    if (s->d_basis) (void)hipFree(s->d_basis);
    s->d_basis = nullptr;
    s->basis_bytes = 0;
    drop_graphs(s);
    hipError_t e = hipMalloc(&s->d_basis, need);
    if (e != hipSuccess)
        return fail(ED_ERR_NOMEM, "Krylov basis of " + std::to_string(need) + " bytes: " + hipGetErrorString(e));
    s->basis_bytes = need;
    s->basis_prec = prec;
    s->basis_nslots = nslots;
    s->ld = ld;
    return ED_OK;
}

// Krylov slots, leading dimension and slot count, recurrence scalars and partial sums
// of a recurrence in work space given by the caller (lz_args_at), or the sector's own.
template <typename T>
static LzArgs<T> lz_args_at(const ed_sector *s, const LzWork &w, int sched)
{
    LzArgs<T> lz;
    lz.base = (T *)w.basis;
    lz.ld = w.ld;
    lz.nslots = w.nslots;
    lz.st = w.st;
    lz.partA = w.partials + kPartA;
    lz.partB = w.partials + kPartB;
    lz.partC = w.partials + kPartC;
    lz.pstride = 8192;
    const HvShape h = hv_shape(s, sched);
    lz.nA = h.nblk;
    lz.dup = s->up.d;
    lz.colmap = h.kind == 2 ? 1 : 0;
    lz.colord = h.xcd ? s->d_colord : nullptr;
    lz.nB = lz.colmap ? h.nblk : elem_grid(s->dim);
    return lz;
}

template <typename T>
static LzArgs<T> lz_args(ed_sector *s)
{
    return lz_args_at<T>(s, LzWork{s->d_basis, s->ld, s->basis_nslots, s->d_state, s->d_partials}, sched_lz<T>(s));
}

// nsteps two-kernel Lanczos steps (H.v phase, update phase, optional CGS2 with its dots
// in `partials`), then the finishing kernel, as plain launches on stream st.
template <typename T>
static void launch_steps_at(ed_sector *s, const LzArgs<T> &lz, const HvShape &h, double *partials, int nsteps,
                            int reorth, hipStream_t st)
{
    const int64_t n = s->dim;
    const int g = lz.nB;   // element-wise kernels: one workgroup per partial of partB
    const int gr = (int)std::max<int64_t>(1, std::min<int64_t>(kReorthGrid, (n + kBlock * 16 - 1) / (kBlock * 16)));
    double *partR = partials + kPartR, *coef = partials + kCoef;
    for (int i = 0; i < nsteps; ++i) {
        hv_launch_mode<T, 1>(s, h, (const T *)nullptr, (T *)nullptr, 0, 1, lz, st);
        hipLaunchKernelGGL(k_lz_update<T>, dim3(g), dim3(kBlock), 0, st, lz, n, reorth ? 0 : 1);
        if (reorth) {
            for (int pass = 0; pass < 2; ++pass) {
                hipLaunchKernelGGL(k_reorth_dots<T>, dim3(gr), dim3(kBlock), 0, st, lz, n, partR);
                hipLaunchKernelGGL(k_reorth_coef<T>, dim3(1), dim3(kBlock), 0, st, lz, (const double *)partR, gr, coef);
                hipLaunchKernelGGL(k_reorth_update<T>, dim3(g), dim3(kBlock), 0, st, lz, n, (const double *)coef,
                                   pass == 1 ? 1 : 0);
            }
        }
    }
    hipLaunchKernelGGL(k_lz_finish<T>, dim3(1), dim3(kBlock), 0, st, lz);
}

template <typename T>
static void launch_steps_impl(ed_sector *s, int nsteps, int reorth, hipStream_t st)
{
    launch_steps_at<T>(s, lz_args<T>(s), hv_shape(s, sched_lz<T>(s)), s->d_partials, nsteps, reorth, st);
}

// ---------------------------------------------------------- persistent Lanczos

// Wait rounds of a grid-barrier poll before the kernel gives up (a few seconds).
constexpr int kBarrierSpins = 1 << 21;
// Shortest run given to the persistent kernel: below it the launch and barrier set-up
// cost more than the captured graph of two-kernel steps.
constexpr int kPersistMinSteps = 16;

template <typename T, int K, int R>
static const void *persist_fn(bool mon)
{
    return mon ? reinterpret_cast<const void *>(&k_lz_persist<T, K, R, true>)
               : reinterpret_cast<const void *>(&k_lz_persist<T, K, R, false>);
}

// The instances compiled: those whose rows fit the 128-register budget of a co-resident
// grid (4 workgroups of 256 threads per CU); width 12 with 4 rows per thread (the excited
// sectors of the G(i w_n) runs at Ns = 12) reloads its packed entries every step and
// keeps its down-hop entries in LDS to get there.  Wider tables keep the two-kernel step.
template <typename T>
static const void *persist_kernel(int k, int rows, bool mon)
{
    constexpr bool f32 = std::is_same<T, float>::value;
    switch (k * 8 + rows) {
    case 4 * 8 + 1: return persist_fn<T, 4, 1>(mon);
    case 4 * 8 + 2: return persist_fn<T, 4, 2>(mon);
    case 4 * 8 + 4: return persist_fn<T, 4, 4>(mon);
    case 8 * 8 + 1: return persist_fn<T, 8, 1>(mon);
    case 8 * 8 + 2: return persist_fn<T, 8, 2>(mon);
    case 8 * 8 + 4: return persist_fn<T, 8, 4>(mon);
    case 12 * 8 + 1: return persist_fn<T, 12, 1>(mon);
    case 12 * 8 + 2: return f32 ? persist_fn<float, 12, 2>(mon) : nullptr;
    case 12 * 8 + 4: return persist_fn<T, 12, 4>(mon);
    case 16 * 8 + 1: return persist_fn<T, 16, 1>(mon);
    default: return nullptr;
    }
}

template <typename T, int K, int R>
static const void *persist_cols_fn(bool mon)
{
    return mon ? reinterpret_cast<const void *>(&k_lz_persist_cols<T, K, R, true>)
               : reinterpret_cast<const void *>(&k_lz_persist_cols<T, K, R, false>);
}

// The multi-column instances: table widths 4, 8 and 12, 1, 2 or 4 rows per thread.
template <typename T>
static const void *persist_cols_kernel(int k, int rows, bool mon)
{
    switch (k * 8 + rows) {
    case 4 * 8 + 1: return persist_cols_fn<T, 4, 1>(mon);
    case 4 * 8 + 2: return persist_cols_fn<T, 4, 2>(mon);
    case 4 * 8 + 4: return persist_cols_fn<T, 4, 4>(mon);
    case 8 * 8 + 1: return persist_cols_fn<T, 8, 1>(mon);
    case 8 * 8 + 2: return persist_cols_fn<T, 8, 2>(mon);
    case 8 * 8 + 4: return persist_cols_fn<T, 8, 4>(mon);
    case 12 * 8 + 1: return persist_cols_fn<T, 12, 1>(mon);
    case 12 * 8 + 2: return persist_cols_fn<T, 12, 2>(mon);
    case 12 * 8 + 4: return persist_cols_fn<T, 12, 4>(mon);
    default: return nullptr;
    }
}

template <typename T, int K, int R>
static const void *persist_1b_fn(bool mon)
{
    return mon ? reinterpret_cast<const void *>(&k_lz_persist_1b<T, K, R, true>)
               : reinterpret_cast<const void *>(&k_lz_persist_1b<T, K, R, false>);
}

// The one-barrier instances: one row per thread (d_up <= 256).  With more rows a thread
// cannot hold q and u of all its gathered elements at once (128 registers at 4 waves per
// SIMD): its rows then take twice the batches of loads of the two-barrier kernel, which
// costs more than the barrier saves (measured at d_up = 495 and 924).
template <typename T>
static const void *persist_1b_kernel(int k, int rows, bool mon)
{
    if (rows != 1) return nullptr;
    switch (k) {
    case 4: return persist_1b_fn<T, 4, 1>(mon);
    case 8: return persist_1b_fn<T, 8, 1>(mon);
    case 12: return persist_1b_fn<T, 12, 1>(mon);
    default: return nullptr;
    }
}

// the one-barrier kernel's partial sums live in the reorthogonalisation dots' space, which
// a persistent run (never reorthogonalised) does not use
static_assert((size_t)kP1Arrays * 2 * kP1Stride <= (size_t)kMaxIter * kReorthGrid, "partials of the one-barrier kernel");
static_assert(kBarGroups * kBarGroupSize <= kP1Stride, "one partial slot per workgroup");

struct PersistShape {
    const void *fn = nullptr, *fn_mon = nullptr;   // plain and monitored instance
    int nblk = 0, rows = 0;
    size_t lds = 0;
    bool onebar = false;   // the one-barrier form of the multi-column kernel
    // the multi-column kernel: cpw columns per workgroup of 256 cpw threads (0: the one-column kernel)
    int cpw = 0, nthreads = kBlock, nA = 0;
    PersistMap map{};
};

// The map of the multi-column kernel for `cpw` columns per workgroup over the column
// order of the sector's H.v schedule (XCD-grouped: the breadth-first order of the down-hop
// graph cut into 8 chunks, workgroup b taking cpw consecutive columns of chunk b mod 8, so
// they sit on one XCD; natural: the same over 0 .. d_dn - 1), with each column's partial
// slot in the two-kernel step's layout.
static int ensure_pmap(ed_sector *s, int cpw, bool xcd)
{
    if (s->d_pmap && s->pmap_cpw == cpw && s->pmap_xcd == (int)xcd) return ED_OK;
    const int d = s->dn.d;
    std::vector<int32_t> order(d), pslot(d), idle;
    if (xcd) {
        if (!s->d_colord || (int)s->col_bfs.size() != d) return fail(ED_ERR_STATE, "persistent map: no column order");
        order = s->col_bfs;
        std::vector<int32_t> cmap(s->colord_blocks);
        ED_HIP_CHECK(hipMemcpy(cmap.data(), s->d_colord, sizeof(int32_t) * cmap.size(), hipMemcpyDeviceToHost));
        for (int b = 0; b < (int)cmap.size(); ++b) {
            if (cmap[b] >= 0) pslot[cmap[b]] = b;
            else idle.push_back(b);
        }
    } else {
        for (int c = 0; c < d; ++c) order[c] = pslot[c] = c;
    }
    const int q = (d + 7) / 8, per = (q + cpw - 1) / cpw, nwg = 8 * per;
    std::vector<int32_t> map((size_t)nwg * cpw + d + idle.size() + 1);
    for (int b = 0; b < nwg; ++b)
        for (int c = 0; c < cpw; ++c) {
            const int pos = (b / 8) * cpw + c, at = (b % 8) * q + pos;
            map[(size_t)b * cpw + c] = (pos < q && at < d) ? order[at] : -1;
        }
    std::copy(pslot.begin(), pslot.end(), map.begin() + (size_t)nwg * cpw);
    std::copy(idle.begin(), idle.end(), map.begin() + (size_t)nwg * cpw + d);
    if (s->d_pmap) (void)hipFree(s->d_pmap);
    s->d_pmap = nullptr;
    ED_HIP_CHECK(hipMalloc((void **)&s->d_pmap, sizeof(int32_t) * map.size()));
    ED_HIP_CHECK(hipMemcpy(s->d_pmap, map.data(), sizeof(int32_t) * map.size(), hipMemcpyHostToDevice));
    s->pmap_cpw = cpw;
    s->pmap_xcd = xcd ? 1 : 0;
    s->pmap_nwg = nwg;
    s->pmap_nidle = (int)idle.size();
    return ED_OK;
}

// The persistent kernel applies when the sector's schedule is a fast-kernel one (one
// column per workgroup, so its partial sums have the fast kernel's layout), d_up fits 4
// rows per thread of 256, and every workgroup of the grid can be resident at once.
// The multi-column kernel applies to fast-kernel schedules with tables of width <= 12, at
// most kPersistHund Hund transitions per row, when its LDS fits and the grid is resident.
template <typename T>
static bool persist_cols_shape(ed_sector *s, const HvShape &h, PersistShape &p, int ncu)
{
    if (!s->persist_cols || h.k > 12 || (s->hund && s->hund_maxrow > kPersistHund)) return false;
    const int rows = s->up.d <= kBlock ? 1 : s->up.d <= 2 * kBlock ? 2 : 4;
    const bool onebar = s->persist_1b != 0 && rows == 1;
    const void *fn = onebar ? persist_1b_kernel<T>(h.k, rows, false) : persist_cols_kernel<T>(h.k, rows, false);
    const void *fn_mon = onebar ? persist_1b_kernel<T>(h.k, rows, true) : persist_cols_kernel<T>(h.k, rows, true);
    if (!fn || !fn_mon) return false;
    int cpw = std::min(4, std::max(1, (s->dn.d + ncu - 1) / ncu));
    for (; cpw >= 1; cpw /= 2) {
        const size_t lds =
            persist_cols_lds(s->up.d, h.k, s->up.ncoef, s->ncfg, s->nhund, s->hund, cpw, sizeof(T), onebar ? 3 : 2)
                .total;
        if (lds > 160 * 1024) continue;
        if (ensure_pmap(s, cpw, h.xcd != 0) != ED_OK) return false;
        int per_cu = 0, per_cu_mon = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock * cpw, lds) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_mon, fn_mon, kBlock * cpw, lds) != hipSuccess)
            return false;
        const int nwg = s->pmap_nwg;
        if ((int64_t)std::min(per_cu, per_cu_mon) * ncu < nwg || nwg > kBarGroups * kBarGroupSize) continue;
        p.fn = fn;
        p.fn_mon = fn_mon;
        p.onebar = onebar;
        p.rows = rows;
        p.cpw = cpw;
        p.nthreads = kBlock * cpw;
        p.nblk = nwg;
        p.lds = lds;
        p.nA = h.nblk;
        const int32_t *m = s->d_pmap;
        p.map = PersistMap{m, m + (size_t)nwg * cpw, m + (size_t)nwg * cpw + s->dn.d, s->pmap_nidle, cpw};
        return true;
    }
    return false;
}

template <typename T>
static bool persist_shape(ed_sector *s, PersistShape &p)
{
    const int pi = std::is_same<T, double>::value ? 0 : 1;
    if (!s->persist || s->persist_ok[pi] == 0) return false;
    const HvShape h = hv_shape(s, sched_lz<T>(s));
    if (h.kind != 2 || s->up.d > 4 * kBlock) return false;
    {
        int dev = 0, ncu = 0, coop = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) == hipSuccess && coop &&
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            persist_cols_shape<T>(s, h, p, ncu)) {
            (void)hipGetLastError();
            if (s->persist_ok[pi] < 0) s->persist_ok[pi] = 1;
            return s->persist_ok[pi] == 1;
        }
        (void)hipGetLastError();
        p = PersistShape();
    }
    p.rows = s->up.d <= kBlock ? 1 : s->up.d <= 2 * kBlock ? 2 : 4;
    p.fn = persist_kernel<T>(h.k, p.rows, false);
    p.fn_mon = persist_kernel<T>(h.k, p.rows, true);
    p.nblk = h.nblk;
    p.nA = h.nblk;
    p.lds = lz_persist_lds(s->up.d, s->up.ncoef, s->ncfg, s->nhund, sizeof(T));
    if (!p.fn) return false;
    if (s->persist_ok[pi] < 0) {
        int dev = 0, per_cu = 0, per_cu_mon = 0, ncu = 0, coop = 0;
        bool ok = hipGetDevice(&dev) == hipSuccess &&
                  hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) == hipSuccess && coop &&
                  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
                  hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, p.fn, kBlock, p.lds) == hipSuccess &&
                  hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_mon, p.fn_mon, kBlock, p.lds) == hipSuccess &&
                  (int64_t)std::min(per_cu, per_cu_mon) * ncu >= p.nblk && p.nblk <= kBarGroups * kBarGroupSize &&
                  p.lds <= 64 * 1024;
        (void)hipGetLastError();
        s->persist_ok[pi] = ok ? 1 : 0;
    }
    return s->persist_ok[pi] == 1;
}

// launch_persist's return when the runtime refuses the cooperative launch (not an ED_ code)
constexpr int kPersistRefused = 1;

template <typename T>
static int launch_persist(ed_sector *s, const PersistShape &p, int nsteps, hipStream_t st,
                          LzMonitor mon = LzMonitor{nullptr, nullptr})
{
    HvArgs<T> a = hv_args<T>(s, 1);
    LzArgs<T> lz = lz_args<T>(s);
    const HvShape h = hv_shape(s, sched_lz<T>(s));
    if (h.xcd) a.colord = s->d_colord;
    lz.nA = lz.nB = p.nA;
    int spins = kBarrierSpins;
    PersistMap pm = p.map;
    if (p.onebar && !s->d_qx) {
        const hipError_t ea = hipMalloc(&s->d_qx, sizeof(double) * 2 * (size_t)s->ld);
        if (ea != hipSuccess) {
            s->d_qx = nullptr;
            return fail(ED_ERR_NOMEM, std::string("exchange vectors of the persistent kernel: ") + hipGetErrorString(ea));
        }
    }
    T *qx = (T *)s->d_qx;
    double *p1 = s->d_partials + kPartR;
    int exact = s->persist_exact_beta;
    void *args[] = {&a, &lz, &nsteps, &spins, &mon};
    void *args_cols[] = {&a, &lz, &pm, &nsteps, &spins, &mon};
    void *args_1b[] = {&a, &lz, &pm, &nsteps, &spins, &mon, &qx, &p1, &exact};
    hipLaunchKernelGGL(k_bar_reset, dim3(1), dim3(256), 0, st, lz.st);
    // persist == 2 (a test hook) takes the refusal path without launching
    const hipError_t e = s->persist == 2 ? hipErrorCooperativeLaunchTooLarge
                                         : hipLaunchCooperativeKernel(mon.stop ? p.fn_mon : p.fn, dim3(p.nblk),
                                                                      dim3(p.nthreads),
                                                                      p.onebar ? args_1b : p.cpw ? args_cols : args,
                                                                      (unsigned)p.lds, st);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        // A refusal of the cooperative launch itself (nothing ran): the grid does not fit at
        // once (hipErrorCooperativeLaunchTooLarge), or the runtime reported it cannot launch
        // one in its present state (hipErrorIllegalState, seen once in a fixed-length run
        // at dim 63,504 between runs that launched).  The sector then leaves the persistent
        // kernel for this precision, the refusal is counted (ed_lanczos_persistent_refusals)
        // and the caller runs the two-kernel step.  Every other error is an error.
        if (e != hipErrorCooperativeLaunchTooLarge && e != hipErrorIllegalState)
            return fail(ED_ERR_HIP, std::string("hipLaunchCooperativeKernel: ") + hipGetErrorString(e));
        s->persist_ok[sizeof(T) == sizeof(double) ? 0 : 1] = 0;
        ++s->persist_refused;
        return kPersistRefused;
    }
    hipLaunchKernelGGL(k_lz_finish<T>, dim3(1), dim3(kBlock), 0, st, lz);
    return ED_OK;
}

// Runs nsteps steps with the persistent kernel when it applies; *done = false otherwise.
static int run_persist(ed_sector *s, int prec, int nsteps, hipStream_t st, bool &done)
{
    PersistShape p;
    done = false;
    int rc;
    if (prec == ED_PREC_F64) {
        if (!persist_shape<double>(s, p)) return ED_OK;
        rc = launch_persist<double>(s, p, nsteps, st);
    } else {
        if (!persist_shape<float>(s, p)) return ED_OK;
        rc = launch_persist<float>(s, p, nsteps, st);
    }
    if (rc == kPersistRefused) return ED_OK;
    ED_TRY(rc);
    done = true;
    return ED_OK;
}

static void run_steps(ed_sector *s, int prec, int nsteps, int reorth, hipStream_t st)
{
    if (prec == ED_PREC_F64) launch_steps_impl<double>(s, nsteps, reorth, st);
    else launch_steps_impl<float>(s, nsteps, reorth, st);
}

static int tune(ed_sector *s);

// Chunks of two-kernel steps replay captured graphs unless the environment sets
// ED_LANCZOS_GRAPHS=0: then every chunk is launched kernel by kernel (the same kernels, so
// the same results), for tools that trace kernel launches and do not follow graph capture.
static bool graphs_enabled()
{
    static const bool on = [] {
        const char *e = std::getenv("ED_LANCZOS_GRAPHS");
        return !(e && e[0] == '0');
    }();
    return on;
}

// Runs nsteps Lanczos steps; chunks of exactly `chunk` steps replay a captured graph.
static int steps(ed_sector *s, int prec, int nsteps, int chunk, int reorth)
{
    hipStream_t st = s->stream;
    if (!reorth && nsteps >= kPersistMinSteps) {
        bool done = false;
        ED_TRY(run_persist(s, prec, nsteps, st, done));
        if (done) {
            ED_HIP_CHECK(hipGetLastError());
            return ED_OK;
        }
    }
    if (nsteps == chunk && chunk > 1 && graphs_enabled()) {
        const int key = prec * 2 + (reorth ? 1 : 0);
        GraphCache &G = s->graphs[key];
        if (!(G.exec && G.chunk == chunk && G.nslots == s->basis_nslots && G.base == s->d_basis)) {
            if (G.exec) {
                (void)hipGraphExecDestroy(G.exec);
                (void)hipGraphDestroy(G.graph);
                G = GraphCache();
            }
            ED_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            run_steps(s, prec, nsteps, reorth, st);
            ED_HIP_CHECK(hipStreamEndCapture(st, &G.graph));
            ED_HIP_CHECK(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
            G.prec = key;
            G.chunk = chunk;
            G.nslots = s->basis_nslots;
            G.base = s->d_basis;
        }
        ED_HIP_CHECK(hipGraphLaunch(G.exec, st));
    } else {
        run_steps(s, prec, nsteps, reorth, st);
    }
    ED_HIP_CHECK(hipGetLastError());
    return ED_OK;
}

static int lz_start(ed_sector *s, int prec, const double *h_start, const double *d_start)
{
    hipStream_t st = s->stream;
    if (!s->cpb) ED_TRY(tune(s));
    const double *src = d_start;
    if (!src) {
        ED_TRY(ensure_io(s));
        if (h_start)
            ED_HIP_CHECK(hipMemcpyAsync(s->d_io_a, h_start, sizeof(double) * s->dim, hipMemcpyHostToDevice, st));
        else
            ED_TRY(start_vector_dev(s->dim, s->d_io_a, st));
        src = s->d_io_a;
    }
    if (prec == ED_PREC_F64)
        hipLaunchKernelGGL(k_lz_init<double>, dim3(lz_args<double>(s).nB), dim3(kBlock), 0, st, lz_args<double>(s), src,
                           s->dim);
    else
        hipLaunchKernelGGL(k_lz_init<float>, dim3(lz_args<float>(s).nB), dim3(kBlock), 0, st, lz_args<float>(s), src,
                           s->dim);
    ED_HIP_CHECK(hipGetLastError());
    return ED_OK;
}

struct HostLz {
    int j, status;
    std::vector<double> alpha, beta;
};

// Copies j (next step), status, alpha[0..upto), beta[0..upto] to the host.
static int read_state(ed_sector *s, HostLz &h, int upto)
{
    LzState *d = s->d_state;
    int hdr[3];
    h.alpha.resize(std::max(upto, 1));
    h.beta.resize(upto + 1);
    ED_HIP_CHECK(hipMemcpyAsync(hdr, d, 3 * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    if (upto > 0)
        ED_HIP_CHECK(hipMemcpyAsync(h.alpha.data(), (char *)d + offsetof(LzState, alpha), sizeof(double) * upto,
                                    hipMemcpyDeviceToHost, s->stream));
    ED_HIP_CHECK(hipMemcpyAsync(h.beta.data(), (char *)d + offsetof(LzState, beta), sizeof(double) * (upto + 1),
                                hipMemcpyDeviceToHost, s->stream));
    ED_HIP_CHECK(hipStreamSynchronize(s->stream));
    h.j = hdr[0];
    h.status = hdr[2];
    if (h.status == 3) {   // the persistent kernel's barrier timed out: never use it here again
        hipLaunchKernelGGL(k_bar_reset, dim3(1), dim3(256), 0, s->stream, d);
        (void)hipStreamSynchronize(s->stream);
        s->persist = 0;
        return fail(ED_ERR_HIP, "persistent Lanczos kernel: grid barrier wait timed out (run stopped)");
    }
    return ED_OK;
}

static int check_common(ed_sector *s, int prec)
{
    if (!s) return fail(ED_ERR_ARG, "NULL sector");
    if (prec != ED_PREC_F64 && prec != ED_PREC_F32) return fail(ED_ERR_ARG, "precision must be ED_PREC_F64 or ED_PREC_F32");
    if (s->cplx && prec == ED_PREC_F32) return fail(ED_ERR_ARG, "complex sectors run in FP64 only");
    ED_HIP_CHECK(hipSetDevice(s->device));
    return ED_OK;
}

// XCD-grouped column order of the fast H.v.  Workgroups are dispatched to the 8 XCDs
// round-robin (workgroup b runs on XCD b mod 8) and every XCD has an L2 of its own.  The
// down hops of column ido read whole neighbour columns, so the columns one XCD owns
// should be neighbours of each other: the down-hop graph is ordered breadth-first,
// component by component, each component from a column of least degree; the order is
// cut into 8 contiguous chunks and workgroup b takes entry b / 8 of chunk b mod 8.
static int ensure_colord(ed_sector *s)
{
    if (s->d_colord) return ED_OK;
    const int d = s->dn.d;
    std::vector<int32_t> ptr(d + 1), rk(std::max<int64_t>(1, s->dn.nnz));
    ED_HIP_CHECK(hipMemcpy(ptr.data(), s->dn.ptr, sizeof(int32_t) * (d + 1), hipMemcpyDeviceToHost));
    if (s->dn.nnz > 0)
        ED_HIP_CHECK(hipMemcpy(rk.data(), s->dn.rk, sizeof(int32_t) * s->dn.nnz, hipMemcpyDeviceToHost));
    std::vector<int> by_deg(d), order;
    order.reserve(d);
    for (int i = 0; i < d; ++i) by_deg[i] = i;
    std::stable_sort(by_deg.begin(), by_deg.end(),
                     [&](int a, int b) { return ptr[a + 1] - ptr[a] < ptr[b + 1] - ptr[b]; });
    std::vector<char> seen(d, 0);
    for (int root : by_deg) {
        if (seen[root]) continue;
        size_t head = order.size();
        seen[root] = 1;
        order.push_back(root);
        while (head < order.size()) {
            const int u = order[head++];
            for (int k = ptr[u]; k < ptr[u + 1]; ++k) {
                const int v = rk[k];
                if (v >= 0 && v < d && !seen[v]) {
                    seen[v] = 1;
                    order.push_back(v);
                }
            }
        }
    }
    s->col_bfs.assign(order.begin(), order.end());
    const int q = (d + 7) / 8, nblk = 8 * q;
    std::vector<int32_t> map(nblk);
    for (int b = 0; b < nblk; ++b) {
        const int pos = (b % 8) * q + b / 8;
        map[b] = pos < d ? order[pos] : -1;
    }
    ED_HIP_CHECK(hipMalloc((void **)&s->d_colord, sizeof(int32_t) * nblk));
    ED_HIP_CHECK(hipMemcpy(s->d_colord, map.data(), sizeof(int32_t) * nblk, hipMemcpyHostToDevice));
    s->colord_blocks = nblk;
    return ED_OK;
}

int ensure_col_bfs(ed_sector *s)
{
    return ensure_colord(s);
}

// Autotuner: times the FP64 H.v under every applicable schedule (row kernel with 1, 2
// or 4 columns per workgroup; fast kernel in natural and in XCD-grouped column order),
// 3 launches each after one warm-up, HIP events on the sector's stream, and keeps the
// fastest.
template <typename T>
static int tune_prec(ed_sector *s, double *ms_out, int &choice)
{
    hipStream_t st = s->stream;
    const T *x = (const T *)s->d_io_a;   // timing only: the contents do not matter
    T *y = (T *)s->d_io_b;
    float best = 1e30f;
    int best_c = 1;
    const int cands[kSchedules] = {1, 2, 4, 8, 16, 18, 19, 20};
    for (int i = 0; i < kSchedules; ++i) {
        const HvShape h = hv_shape(s, cands[i]);
        const bool applies = (cands[i] >= 16) ? h.kind == 2 && h.xcd && (cands[i] == 16 || h.rows == cands[i] - 16)
                             : (cands[i] == 8) ? h.kind == 2
                                               : (h.kind == 1 && h.cpb == cands[i]) || (h.kind == 0 && i == 0);
        if (!applies) continue;
        LzArgs<T> lz{};
        hv_launch_mode<T, 0>(s, h, x, y, 0, 1, lz, st);
        ED_HIP_CHECK(hipEventRecord(s->ev0, st));
        for (int r = 0; r < 3; ++r) hv_launch_mode<T, 0>(s, h, x, y, 0, 1, lz, st);
        ED_HIP_CHECK(hipEventRecord(s->ev1, st));
        ED_HIP_CHECK(hipEventSynchronize(s->ev1));
        float ms = 0.f;
        ED_HIP_CHECK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        ms_out[i] = ms / 3;
        if (ms < best) {
            best = ms;
            best_c = cands[i];
        }
    }
    ED_HIP_CHECK(hipGetLastError());
    choice = best_c;
    return ED_OK;
}

// Whether the persistent Lanczos kernel has an instance for the sector under schedule
// `sched` (the fast kernel, d_up <= 1024, a compiled width and rows per thread); the
// co-residency test of persist_shape still applies at launch.
template <typename T>
static bool persist_may_apply(ed_sector *s, int sched)
{
    const HvShape h = hv_shape(s, sched);
    if (h.kind != 2 || s->up.d > 4 * kBlock || !s->persist) return false;
    const int rows = s->up.d <= kBlock ? 1 : s->up.d <= 2 * kBlock ? 2 : 4;
    const bool cols = s->persist_cols && h.k <= 12 && !(s->hund && s->hund_maxrow > kPersistHund) &&
                      persist_cols_kernel<T>(h.k, rows, false);
    return cols || persist_kernel<T>(h.k, rows, false);
}

// Schedule of the Lanczos recurrence (0: the H.v schedule `hv_choice`).  Where the fastest
// H.v is the row kernel, which keeps the persistent kernel out, the fastest fast-kernel
// schedule under which the persistent kernel applies is taken instead if its H.v is within
// 1.5x: a persistent step (H.v and update in one launch, no launch gaps, one grid barrier
// where a thread owns one row) is 1.7x the two-kernel step there (6.3 against 10.6 us at
// (5,5) of Ns = 10 with H.v times 4.91 against 4.85 us; 8.5 against 14.5 us at (3,6) of
// Ns = 12 with 6.96 against 6.23 us).
template <typename T>
static int lanczos_sched(ed_sector *s, const double *ms, int hv_choice)
{
    const int cands[kSchedules] = {1, 2, 4, 8, 16, 18, 19, 20};
    if (hv_shape(s, hv_choice).kind == 2) return 0;
    double thv = 0.0, tbest = 1e30;
    int best = 0;
    for (int i = 0; i < kSchedules; ++i) {
        if (cands[i] == hv_choice) thv = ms[i];
        if (cands[i] >= 8 && ms[i] > 0.0 && ms[i] < tbest && persist_may_apply<T>(s, cands[i])) {
            tbest = ms[i];
            best = cands[i];
        }
    }
    return (best && thv > 0.0 && tbest <= 1.5 * thv) ? best : 0;
}

// Autotuner: times the H.v under every applicable schedule (row kernel with 1, 2 or 4
// columns per workgroup; fast kernel in natural and XCD-grouped column order, with 1-4
// rows per thread), 3 launches each after one warm-up, HIP events on the sector's
// stream, separately for the FP64 and the FP32 kernels, and keeps the fastest of each.
static int tune(ed_sector *s)
{
    ED_TRY(ensure_io(s));
    hipStream_t st = s->stream;
    for (double &t : s->tune_ms) t = 0.0;
    for (double &t : s->tune_ms32) t = 0.0;
    if (s->sz || s->cplx || s->dim < 4096) {
        s->cpb = fast_ok(s) ? 8 : 1;
        s->cpb32 = 0;
        s->lzs = s->lzs32 = 0;
        return ED_OK;
    }
    if (fast_ok(s)) ED_TRY(ensure_colord(s));
    ED_TRY(start_vector_dev(s->dim, s->d_io_a, st));
    int c64 = 1, c32 = 1;
    ED_TRY(tune_prec<double>(s, s->tune_ms, c64));
    ED_TRY(tune_prec<float>(s, s->tune_ms32, c32));
    s->cpb = c64;
    s->cpb32 = c32;
    s->lzs = lanczos_sched<double>(s, s->tune_ms, c64);
    s->lzs32 = lanczos_sched<float>(s, s->tune_ms32, c32);
    s->persist_ok[0] = s->persist_ok[1] = -1;
    drop_graphs(s);
    return ED_OK;
}

struct PhaseOut {
    int iters = 0;
    std::vector<double> alpha, beta;
    // Rayleigh-form FP32 phase: the target was not met (gave up, or the run ended first);
    // second Ritz value and its residual estimate at the last target test
    bool rq_gave_up = true;
    double rq_theta2 = 0.0, rq_rho2 = 0.0;
};

// The convergence test of Lanczos_fast_diagonalize (Lanczos_plain.f90:33-134), one
// iteration at a time: iteration `it` ends the loop when the neig lowest eigenvalues of
// T_it all moved by less than tol since T_{it-1} (it > 2), at it == niter, or when the
// first step already gives beta_2 = 0.
struct ConvTest {
    int neig, niter;
    double tol;
    // > 0 (the FP32 phase of the Rayleigh form): once the eigenvalue test is met, also wait
    // until the Lanczos estimate of the lowest Ritz pair's residual, rho_1 = beta_{n+1} |z_n1|
    // of T_n (n = it - 1), meets the Kato-Temple target rho_1^2 <= rq_tol (l_2 - theta_1) / 2
    // the FP64 check applies, with l_2 = theta_2 - beta_{n+1} |z_n2| the lower estimate of
    // lambda_2 from the second Ritz pair.  The wait ends without the target (rq_gave_up)
    // when the target is below what vectors of unit roundoff rq_eps can reach
    // ((4 rq_eps ||T||)^2 > rq_tol gap / 2), when theta_2 collapses onto theta_1 (a Lanczos
    // copy of theta_1, or a splitting the phase cannot resolve), or after
    // max(16, n_conv) further steps (n_conv: the step that met the eigenvalue test).
    double rq_tol = 0.0, rq_eps = 0.0;
    bool rq_gave_up = false;
    int rq_conv_at = 0;             // iteration at which the eigenvalue test was first met
    double rq_theta2 = 0.0, rq_rho2 = 0.0;   // second Ritz pair of the last target test
    std::vector<double> last, w;
    ConvTest(int neig_, int niter_, double tol_) : neig(neig_), niter(niter_), tol(tol_), last(neig_, 1e300), w(neig_) {}
    // false: keep iterating for the Rayleigh-form target
    bool rq_target(int it, const double *alpha, const double *beta)
    {
        const int n = it - 1;
        if (!rq_conv_at) rq_conv_at = it;
        if (it - rq_conv_at > std::max(16, rq_conv_at)) {
            rq_gave_up = true;
            return true;
        }
        double ev[2];
        std::vector<double> z(2 * (size_t)std::max(n, 2));
        if (n < 2 || tridiag_lowest_pairs(n, alpha, beta, 2, ev, z.data()) != ED_OK) return false;
        const double rho1 = std::fabs(beta[n] * z[n - 1]), rho2 = std::fabs(beta[n] * z[2 * (size_t)n - 1]);
        double tn = 0.0;   // Gershgorin bound of ||T_n||
        for (int i = 0; i < n; ++i)
            tn = std::max(tn, std::fabs(alpha[i]) + (i > 0 ? std::fabs(beta[i]) : 0.0) +
                                  (i + 1 < n ? std::fabs(beta[i + 1]) : 0.0));
        rq_theta2 = ev[1];
        rq_rho2 = rho2;
        const double gap = (ev[1] - rho2) - ev[0];
        const double fl = 4.0 * rq_eps * tn;
        if (!(ev[1] - ev[0] > 1e-7 * std::max(1.0, std::fabs(ev[0]))) || (gap > 0.0 && fl * fl > rq_tol * 0.5 * gap)) {
            rq_gave_up = true;
            return true;
        }
        return gap > 0.0 && rho1 * rho1 <= rq_tol * 0.5 * gap;
    }
    bool ends(int it, const double *alpha, const double *beta)
    {
        tridiag_lowest_warm(it, alpha, beta, neig, last.data(), w.data());
        const int m = std::min(neig, it);
        bool allc = true;
        for (int k = 0; k < neig; ++k)
            if (!((k < m) && std::fabs(w[k] - last[k]) < tol && it > 2)) allc = false;
        if (allc && rq_tol > 0.0 && !rq_gave_up) allc = rq_target(it, alpha, beta);
        const bool one_step = (it == 1) && std::fabs(beta[1]) < 1e-24;
        if (allc || it == niter || one_step) return true;
        for (int k = 0; k < m; ++k) last[k] = w[k];
        return false;
    }
};

// The pinned host channel of monitored persistent runs (LzMonitor): fine-grained, so the
// kernel's system-scope stores reach the host while it runs and its loads see the flag.
static int ensure_monitor(ed_sector *s)
{
    if (s->h_mon) return ED_OK;
    void *h = nullptr, *d = nullptr;
    ED_HIP_CHECK(hipHostMalloc(&h, sizeof(double) * (16 + 2 * (size_t)kMaxIter),
                               hipHostMallocMapped | hipHostMallocCoherent));
    const hipError_t e = hipHostGetDevicePointer(&d, h, 0);
    if (e != hipSuccess) {
        (void)hipHostFree(h);
        return fail(ED_ERR_HIP, std::string("hipHostGetDevicePointer: ") + hipGetErrorString(e));
    }
    s->h_mon = (double *)h;
    s->d_mon = (double *)d;
    return ED_OK;
}

// Steps done+1 .. done+nsteps in one persistent launch while the host runs the
// convergence test on the scalars the kernel publishes; once an iteration ends the loop
// the kernel is told to stop (it does a few steps later; those steps are never used).
// alpha/beta: host copies, filled from index `done` on as the scalars arrive; `tested`:
// the last iteration tested (in order, as the chunked loop does).  Returns with the
// stream drained; the scalars of the steps not yet tested are then read from the device.
template <typename T>
static int run_monitored(ed_sector *s, const PersistShape &p, int done, int nsteps, ConvTest &ct,
                         std::vector<double> &alpha, std::vector<double> &beta, int &tested, int &final_iter)
{
    ED_TRY(ensure_monitor(s));
    hipStream_t st = s->stream;
    volatile int *stop = reinterpret_cast<volatile int *>(s->h_mon);
    volatile double *ab = s->h_mon + 16;
    *stop = 0;
    for (int i = 2 * done; i < 2 * (done + nsteps); ++i) ab[i] = std::numeric_limits<double>::quiet_NaN();
    std::atomic_thread_fence(std::memory_order_seq_cst);
    ED_TRY(launch_persist<T>(s, p, nsteps, st, LzMonitor{reinterpret_cast<const int *>(s->d_mon), s->d_mon + 16}));
    ED_HIP_CHECK(hipGetLastError());
    int avail = done;   // steps whose alpha and beta are on the host
    while (!final_iter) {
        while (avail < done + nsteps) {
            const double b = ab[2 * avail], a = ab[2 * avail + 1];
            if (std::isnan(a) || std::isnan(b)) break;
            alpha[avail] = a;
            beta[avail] = b;
            ++avail;
        }
        bool progressed = false;
        // iteration 1 also needs beta_2 (published with step 2)
        while (!final_iter && tested < avail && (tested > 0 || avail >= 2)) {
            ++tested;
            progressed = true;
            if (ct.ends(tested, alpha.data(), beta.data())) {
                final_iter = tested;
                *stop = 1;
                std::atomic_thread_fence(std::memory_order_seq_cst);
            }
        }
        if (!progressed && hipStreamQuery(st) != hipErrorNotReady) break;   // ended (errors: the sync reports them)
    }
    ED_HIP_CHECK(hipStreamSynchronize(st));
    return ED_OK;
}

// One Lanczos_fast_diagonalize loop (Lanczos_plain.f90:33-134) in precision `prec`
// from a host or device start vector, with the reference's convergence test on the host
// for every step.  Where the persistent kernel applies, one launch runs up to the whole
// loop while the host tests the steps as they come (run_monitored); otherwise chunks of
// `chunk` steps alternate with the tests.  Both give the same scalars, so the same
// iteration ends the loop.
static int lanczos_phase(ed_sector *s, int prec, double tol, int niter, int neig, int chunk, int reorth,
                         const double *h_start, const double *d_start, PhaseOut &out, double rq_tol = 0.0)
{
    out.rq_gave_up = rq_tol > 0.0;
    ED_TRY(ensure_state(s));
    ED_TRY(ensure_basis(s, prec, niter + 1));
    s->basis_valid_iters = 0;
    ED_TRY(lz_start(s, prec, h_start, d_start));
    ConvTest ct(neig, niter, tol);
    ct.rq_tol = rq_tol;
    ct.rq_eps = prec == ED_PREC_F32 ? 5.9604644775390625e-8 : 1.1102230246251565e-16;   // 2^-24, 2^-53
    std::vector<double> alpha(niter + 1), beta(niter + 2);
    HostLz h;
    int done = 0, final_iter = 0, tested = 0;
    while (done < niter && !final_iter) {
        int c = std::min(chunk, niter - done);
        PersistShape ps;
        const bool mon = !reorth && niter - done >= kPersistMinSteps &&
                         (prec == ED_PREC_F64 ? persist_shape<double>(s, ps) : persist_shape<float>(s, ps));
        if (mon) {
            c = niter - done;
            const int rc = prec == ED_PREC_F64
                               ? run_monitored<double>(s, ps, done, c, ct, alpha, beta, tested, final_iter)
                               : run_monitored<float>(s, ps, done, c, ct, alpha, beta, tested, final_iter);
            if (rc == kPersistRefused) continue;   // nothing ran; the two-kernel step takes over
            ED_TRY(rc);
        } else {
            ED_TRY(steps(s, prec, c, chunk, reorth));
        }
        ED_TRY(read_state(s, h, done + c));
        if (h.status == 2) return fail(ED_ERR_BREAKDOWN, "Lanczos start vector has zero norm (Lanczos_plain.h:70-76)");
        s->last_steps_run += std::min(h.j - 1, done + c) - done;
        if (final_iter) break;
        const int completed = h.j - 1;   // steps whose beta_{j+1} is known
        for (int i = tested; i < done + c; ++i) {
            alpha[i] = h.alpha[i];
            beta[i] = h.beta[i];
        }
        beta[done + c] = h.beta[done + c];
        for (int it = tested + 1; it <= done + c; ++it) {
            if (it > completed)   // the reference STOPs at the start of this step (Lanczos_plain.h:85-92)
                return fail(ED_ERR_BREAKDOWN, "zero-norm Lanczos vector at iteration " + std::to_string(it));
            tested = it;
            if (ct.ends(it, alpha.data(), beta.data())) {
                final_iter = it;
                break;
            }
        }
        done += c;
    }
    out.iters = final_iter;
    out.rq_gave_up = ct.rq_gave_up || !ct.rq_conv_at;
    out.rq_theta2 = ct.rq_theta2;
    out.rq_rho2 = ct.rq_rho2;
    out.alpha.assign(h.alpha.begin(), h.alpha.begin() + final_iter);
    out.beta.assign(h.beta.begin(), h.beta.begin() + final_iter);
    s->basis_valid_iters = final_iter;
    s->basis_prec = prec;
    return ED_OK;
}

static int gs_vector_dev(ed_sector *s, int iters, const double *vecp, double *d_psi);

// Per-workgroup partial sums of <x, x>, <x, y> and ||y - c x||^2 (grid-stride, fixed order).
__global__ __launch_bounds__(kBlock) void k_dot3_part(const double *x, const double *y, int64_t n, double c,
                                                      double *part)
{
    __shared__ double red[kRed];
    double a = 0.0, b = 0.0, r = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double xi = x[i], yi = y[i], di = yi - c * xi;
        a += xi * xi;
        b += xi * yi;
        r += di * di;
    }
    a = block_sum_all(a, red);
    b = block_sum_all(b, red);
    r = block_sum_all(r, red);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = a;
        part[3 * blockIdx.x + 1] = b;
        part[3 * blockIdx.x + 2] = r;
    }
}

// FP64 Rayleigh quotient E = <psi, H psi> / <psi, psi> of the device vector d_psi and the
// squared residual ||H psi - E psi||^2 / <psi, psi> (one H.v into d_w, one pass of sums).
// The residual is summed about the shift c (an estimate of E) and corrected by
// ||Hpsi - E psi||^2 = ||H psi - c psi||^2 - (E - c)^2 ||psi||^2, which cancels only
// (E - c)^2 instead of E^2.
static int rayleigh_residual(ed_sector *s, const double *d_psi, double *d_w, double c, double &e, double &res2)
{
    hipStream_t st = s->stream;
    ED_TRY(launch_hv(s, ED_PREC_F64, d_psi, d_w, 0, 1, st));
    constexpr int kParts = 256;
    double *d_part = nullptr;
    ED_HIP_CHECK(hipMallocAsync((void **)&d_part, sizeof(double) * 3 * kParts, st));
    hipLaunchKernelGGL(k_dot3_part, dim3(kParts), dim3(kBlock), 0, st, d_psi, (const double *)d_w, s->dim, c, d_part);
    ED_HIP_CHECK(hipGetLastError());
    std::vector<double> part(3 * kParts);
    ED_HIP_CHECK(hipMemcpyAsync(part.data(), d_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipStreamSynchronize(st));
    (void)hipFreeAsync(d_part, st);
    double xx = 0.0, xy = 0.0, rr = 0.0;
    for (int q = 0; q < kParts; ++q) {
        xx += part[3 * q];
        xy += part[3 * q + 1];
        rr += part[3 * q + 2];
    }
    if (!(xx > 0.0)) return fail(ED_ERR_BREAKDOWN, "Rayleigh quotient of a zero vector");
    e = xy / xx;
    res2 = std::max(0.0, rr / xx - (e - c) * (e - c));
    return ED_OK;
}

// Lanczos_fast_diagonalize (Lanczos_plain.f90:33-134).  ED_PREC_F64 is the reference
// algorithm.  ED_PREC_F32 is the mixed-precision form: the loop runs on FP32 vectors
// and FP32 H.v (FP64 dots and scalars) to |dlambda| < max(tol, 1e-6), then an FP64
// Lanczos restarted on the device from that Ritz vector finishes to `tol`; the
// returned T, eigenpairs and kept basis are those of the FP64 phase.  ED_PREC_MIXED_RQ
// (one eigenvalue): after the FP32 phase, the FP64 Rayleigh quotient of its Ritz vector
// is E0 when ||r||^2 <= tol * (lambda_2 - E) / 2; the outputs are then those of the FP32
// phase with evals[0] = E.  Otherwise (and for neigen > 1) it continues as ED_PREC_F32.
static int lanczos_diagonalize(ed_sector *s, const ed_lanczos_opts *o, const double *start, int *iters_out,
                               int *neig_out, double *evals, double *vecp, double *diag, double *subdiag)
{
    if (!o || !iters_out) return fail(ED_ERR_ARG, "ed_lanczos_diagonalize: NULL argument");
    const bool rq = o->precision == ED_PREC_MIXED_RQ;
    ED_TRY(check_common(s, rq ? ED_PREC_F32 : o->precision));
    const int64_t D = s->cplx ? s->cdim : s->dim;   // Niter = min(dimen_H - 1, Nitermax)
    const int niter = (int)std::min<int64_t>(D - 1, o->nitermax);
    if (niter < 1) return fail(ED_ERR_ARG, "ed_lanczos_diagonalize: sector dimension 1 has no Lanczos step (full ED case)");
    if (niter > kMaxIter - 2) return fail(ED_ERR_ARG, "ed_lanczos_diagonalize: nitermax above " + std::to_string(kMaxIter - 2));
    int neig = (int)std::min<int64_t>(o->neigen, D);
    neig = std::min(neig, niter);
    if (neig < 1) return fail(ED_ERR_ARG, "ed_lanczos_diagonalize: neigen < 1");
    const int chunk = std::max(1, o->check_every);
    PhaseOut ph;
    s->last_f32_iters = 0;
    s->last_steps_run = 0;
    s->last_rq_accepted = 0;
    s->last_rq_res2 = s->last_rq_gap = 0.0;
    if (o->precision == ED_PREC_F32 || rq) {
        PhaseOut p32;
        ED_TRY(lanczos_phase(s, ED_PREC_F32, std::max(o->tolerance, 1e-6), niter, 1, chunk, o->reorth, start, nullptr,
                             p32, rq && neig == 1 ? o->tolerance : 0.0));
        // the RQ form also takes the second Ritz value of the final T for the gap of its bound
        const int k32 = std::min(rq ? 2 : 1, p32.iters);
        std::vector<double> ev(k32), z((size_t)p32.iters * k32);
        ED_TRY(tridiag_lowest_pairs(p32.iters, p32.alpha.data(), p32.beta.data(), k32, ev.data(), z.data()));
        ED_TRY(ensure_io(s));
        ED_TRY(gs_vector_dev(s, p32.iters, z.data(), s->d_io_b));
        s->last_f32_iters = p32.iters;
        if (rq && neig == 1 && k32 == 2 && !p32.rq_gave_up) {
            // E0 from the FP64 Rayleigh quotient E of the FP32 Ritz vector, with the
            // Kato-Temple bound E - lambda_1 <= ||r||^2 / (l_2 - E) for a lower bound l_2 of
            // lambda_2.  l_2 = theta_2 - rho_2 (the second Ritz value less its residual
            // estimate) stands in for it: Lanczos gives no guaranteed lower bound (a missed
            // eigenvalue), so the test keeps a factor 2 of margin on the gap.
            double e = 0.0, r2 = 0.0;
            ED_TRY(rayleigh_residual(s, s->d_io_b, s->d_io_a, ev[0], e, r2));
            const double gap = (p32.rq_theta2 - p32.rq_rho2) - e;
            s->last_rq_res2 = r2;
            s->last_rq_gap = gap;
            if (gap > 0.0 && r2 <= o->tolerance * 0.5 * gap) {
                s->last_rq_accepted = 1;
                if (evals) evals[0] = e;
                if (vecp)
                    for (int r = 0; r < p32.iters; ++r) vecp[r] = z[r];
                if (diag) std::memcpy(diag, p32.alpha.data(), sizeof(double) * p32.iters);
                if (subdiag) std::memcpy(subdiag, p32.beta.data(), sizeof(double) * p32.iters);
                *iters_out = p32.iters;
                if (neig_out) *neig_out = 1;
                return ED_OK;
            }
        }
        ED_TRY(lanczos_phase(s, ED_PREC_F64, o->tolerance, niter, neig, chunk, o->reorth, nullptr, s->d_io_b, ph));
    } else {
        ED_TRY(lanczos_phase(s, ED_PREC_F64, o->tolerance, niter, neig, chunk, o->reorth, start, nullptr, ph));
    }
    const int it = ph.iters;
    neig = std::min(neig, it);
    std::vector<double> ev(neig), z((size_t)it * neig);
    ED_TRY(tridiag_lowest_pairs(it, ph.alpha.data(), ph.beta.data(), neig, ev.data(), z.data()));
    for (int k = 0; k < neig; ++k) {
        if (evals) evals[k] = ev[k];
        if (vecp)
            for (int r = 0; r < it; ++r) vecp[(size_t)k * it + r] = z[(size_t)k * it + r];
    }
    if (diag) std::memcpy(diag, ph.alpha.data(), sizeof(double) * it);
    if (subdiag) std::memcpy(subdiag, ph.beta.data(), sizeof(double) * it);
    *iters_out = it;
    if (neig_out) *neig_out = neig;
    return ED_OK;
}

static int lanczos_run(ed_sector *s, int niter, const double *h_start, const double *d_start, int prec, int chunk,
                       double *diag, double *subdiag, int *breakdown_at)
{
    ED_TRY(check_common(s, prec));
    if (niter < 1 || niter > kMaxIter - 2) return fail(ED_ERR_ARG, "ed_lanczos_run: niter outside [1, 2046]");
    ED_TRY(ensure_state(s));
    ED_TRY(ensure_basis(s, prec, niter + 1));
    s->basis_valid_iters = 0;
    ED_TRY(lz_start(s, prec, h_start, d_start));
    // no test between chunks here: the persistent kernel runs all steps in one launch
    PersistShape ps;
    const bool one = niter >= kPersistMinSteps &&
                     (prec == ED_PREC_F64 ? persist_shape<double>(s, ps) : persist_shape<float>(s, ps));
    int done = 0;
    while (done < niter) {
        const int c = one ? niter - done : std::min(chunk, niter - done);
        ED_TRY(steps(s, prec, c, chunk, 0));
        done += c;
    }
    HostLz h;
    ED_TRY(read_state(s, h, niter));
    if (h.status == 2) return fail(ED_ERR_BREAKDOWN, "Lanczos start vector has zero norm (Lanczos_plain.h:70-76)");
    const int reached = h.j - 1;   // completed steps
    if (breakdown_at) *breakdown_at = (h.status == 1) ? reached : 0;
    for (int i = 0; i < niter; ++i) {
        if (diag) diag[i] = i < reached ? h.alpha[i] : 0.0;
        if (subdiag) subdiag[i] = i <= reached ? h.beta[i] : 0.0;
    }
    s->basis_valid_iters = reached;
    s->basis_prec = prec;
    if (h.status == 1 && reached < niter)
        return fail(ED_ERR_BREAKDOWN, "zero-norm Lanczos vector at iteration " + std::to_string(reached + 1) +
                                          " (Lanczos_plain.h:85-92)");
    return ED_OK;
}

// x.y into *out (device scalar), deterministic.
static void dot_dev(ed_sector *s, const double *x, const double *y, int64_t n, double *out, hipStream_t st)
{
    const int g = elem_grid(n);
    double *part = s->d_partials + kPartR;   // free outside a reorthogonalised step
    hipLaunchKernelGGL(k_dot_part, dim3(g), dim3(kBlock), 0, st, x, y, n, part);
    hipLaunchKernelGGL(k_sum_part, dim3(1), dim3(kBlock), 0, st, (const double *)part, g, out);
}

static int gs_vector_dev(ed_sector *s, int iters, const double *vecp, double *d_psi)
{
    ED_TRY(check_common(s, 0));
    if (!vecp || !d_psi || iters < 1) return fail(ED_ERR_ARG, "ed_lanczos_gs_vector: bad argument");
    if (iters > kMaxIter - 1) return fail(ED_ERR_ARG, "ed_lanczos_gs_vector: iters above 2047");
    if (iters > s->basis_valid_iters || !s->d_basis) {
        // No kept basis (a fresh handle, as after the reference's second new_H, solver.h:43):
        // the second pass of Lanczos_get_GS_sector (Lanczos_plain.f90:156-229), the FP64
        // recurrence again from the deterministic start vector (create_fix_initial_vector,
        // :196), here with its iters vectors kept for the combination below.
        const int n = std::max(1, iters - 1);
        ED_TRY(lanczos_run(s, n, nullptr, nullptr, ED_PREC_F64, 8, nullptr, nullptr, nullptr));
        s->basis_valid_iters = std::max(s->basis_valid_iters, iters);   // u_1 .. u_{n+1}, beta_1 .. beta_{n+1}
    }
    hipStream_t st = s->stream;
    double *d_coef = s->d_partials + kMisc;
    ED_HIP_CHECK(hipMemcpyAsync(d_coef, vecp, sizeof(double) * iters, hipMemcpyHostToDevice, st));
    const int g = elem_grid(s->dim);
    if (s->basis_prec == ED_PREC_F64)
        hipLaunchKernelGGL(k_combine<double>, dim3(g), dim3(kBlock), 0, st, lz_args<double>(s), iters,
                           (const double *)d_coef, d_psi, s->dim);
    else
        hipLaunchKernelGGL(k_combine<float>, dim3(g), dim3(kBlock), 0, st, lz_args<float>(s), iters,
                           (const double *)d_coef, d_psi, s->dim);
    double *d_nrm = d_coef + iters;
    dot_dev(s, d_psi, d_psi, s->dim, d_nrm, st);
    hipLaunchKernelGGL(k_scale, dim3(g), dim3(kBlock), 0, st, d_psi, s->dim, (const double *)d_nrm);
    ED_HIP_CHECK(hipGetLastError());
    return ED_OK;
}

}  // namespace ed

using namespace ed;

extern "C" {

int ed_hmult_f64_dev(ed_sector *sec, int64_t n, double *d_out, const double *d_in, void *stream)
{
    ED_TRY(check_common(sec, 0));
    if (n != sec->dim || !d_out || !d_in) return fail(ED_ERR_ARG, "ed_hmult_f64_dev: size mismatch or NULL vector");
    return launch_hv(sec, ED_PREC_F64, d_in, d_out, 0, 1, stream ? (hipStream_t)stream : sec->stream);
}

int ed_hmult_f32_dev(ed_sector *sec, int64_t n, float *d_out, const float *d_in, void *stream)
{
    ED_TRY(check_common(sec, 0));
    if (n != sec->dim || !d_out || !d_in) return fail(ED_ERR_ARG, "ed_hmult_f32_dev: size mismatch or NULL vector");
    return launch_hv(sec, ED_PREC_F32, d_in, d_out, 0, 1, stream ? (hipStream_t)stream : sec->stream);
}

int ed_hmult_block_f64_dev(ed_sector *sec, int64_t n, int nvec, int64_t ld, double *d_out, const double *d_in,
                           void *stream)
{
    ED_TRY(check_common(sec, 0));
    if (n != sec->dim || nvec < 1 || nvec > 65535 || ld < n || !d_out || !d_in)
        return fail(ED_ERR_ARG, "ed_hmult_block_f64_dev: bad sizes");
    return launch_hv(sec, ED_PREC_F64, d_in, d_out, ld, nvec, stream ? (hipStream_t)stream : sec->stream);
}

int ed_sector_synchronize(ed_sector *sec)
{
    ED_TRY(check_common(sec, 0));
    ED_HIP_CHECK(hipStreamSynchronize(sec->stream));
    return ED_OK;
}

int ed_hmult_c128_dev(ed_sector *sec, int64_t n, double *d_out, const double *d_in, void *stream)
{
    ED_TRY(check_common(sec, 0));
    if (!sec->cplx) return fail(ED_ERR_ARG, "ed_hmult_c128: not a complex sector (ed_sector_create_c)");
    if (n != sec->cdim || !d_out || !d_in) return fail(ED_ERR_ARG, "ed_hmult_c128: n differs from dimen_H()");
    ED_TRY(ensure_io(sec));
    hipStream_t st = stream ? (hipStream_t)stream : sec->stream;
    const int g = elem_grid(n);
    hipLaunchKernelGGL(k_planar, dim3(g), dim3(kBlock), 0, st, d_in, sec->d_io_a, n);
    ED_TRY(launch_hv(sec, ED_PREC_F64, sec->d_io_a, sec->d_io_b, 0, 1, st));
    hipLaunchKernelGGL(k_interleaved, dim3(g), dim3(kBlock), 0, st, (const double *)sec->d_io_b, d_out, n);
    ED_HIP_CHECK(hipGetLastError());
    return ED_OK;
}

int ed_hmult_c128(ed_sector *sec, int64_t n, double *vec_out, const double *vec_in)
{
    ED_TRY(check_common(sec, 0));
    if (!sec->cplx) return fail(ED_ERR_ARG, "ed_hmult_c128: not a complex sector (ed_sector_create_c)");
    if (n != sec->cdim || !vec_out || !vec_in) return fail(ED_ERR_ARG, "ed_hmult_c128: n differs from dimen_H()");
    ED_TRY(ensure_io(sec));   // two staging vectors of 2 n doubles
    hipStream_t st = sec->stream;
    const int g = elem_grid(n);
    ED_HIP_CHECK(hipMemcpyAsync(sec->d_io_b, vec_in, sizeof(double) * 2 * n, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_planar, dim3(g), dim3(kBlock), 0, st, (const double *)sec->d_io_b, sec->d_io_a, n);
    ED_TRY(launch_hv(sec, ED_PREC_F64, sec->d_io_a, sec->d_io_b, 0, 1, st));
    hipLaunchKernelGGL(k_interleaved, dim3(g), dim3(kBlock), 0, st, (const double *)sec->d_io_b, sec->d_io_a, n);
    ED_HIP_CHECK(hipGetLastError());
    ED_HIP_CHECK(hipMemcpyAsync(vec_out, sec->d_io_a, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipStreamSynchronize(st));
    return ED_OK;
}

int ed_hmult_f64(ed_sector *sec, int64_t n, double *vec_out, const double *vec_in)
{
    ED_TRY(check_common(sec, 0));
    if (n != sec->dim || !vec_out || !vec_in) return fail(ED_ERR_ARG, "ed_hmult_f64: n differs from dimen_H()");
    ED_TRY(ensure_io(sec));
    hipStream_t st = sec->stream;
    ED_HIP_CHECK(hipMemcpyAsync(sec->d_io_a, vec_in, sizeof(double) * n, hipMemcpyHostToDevice, st));
    ED_TRY(launch_hv(sec, ED_PREC_F64, sec->d_io_a, sec->d_io_b, 0, 1, st));
    ED_HIP_CHECK(hipMemcpyAsync(vec_out, sec->d_io_b, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipStreamSynchronize(st));
    return ED_OK;
}

int ed_hmult_tune(ed_sector *sec, int *cpb, double *ms_per_choice)
{
    ED_TRY(check_common(sec, 0));
    ED_TRY(tune(sec));
    if (cpb) *cpb = sec->cpb;
    if (ms_per_choice)
        for (int i = 0; i < kSchedules; ++i) ms_per_choice[i] = sec->tune_ms[i];
    return ED_OK;
}

int ed_hmult_set_schedule(ed_sector *sec, int choice)
{
    ED_TRY(check_common(sec, 0));
    if (sec->sz || sec->cplx) return fail(ED_ERR_ARG, "ed_hmult_set_schedule: S_z and complex sectors have one H.v schedule");
    if (!sched_valid(choice))
        return fail(ED_ERR_ARG, "ed_hmult_set_schedule: choice must be 1, 2, 4, 8, 16, 18, 19 or 20");
    if (choice >= 8 && !fast_ok(sec)) return fail(ED_ERR_ARG, "ed_hmult_set_schedule: fast kernel does not apply here");
    if (choice >= 16) ED_TRY(ensure_colord(sec));
    if (choice >= 18 && hv_shape(sec, choice).rows != choice - 16)
        return fail(ED_ERR_ARG, "ed_hmult_set_schedule: that many rows per thread do not apply here");
    sec->cpb = choice;
    sec->cpb32 = 0;   // FP32 and the Lanczos recurrence follow the forced schedule too
    sec->lzs = sec->lzs32 = 0;
    sec->persist_ok[0] = sec->persist_ok[1] = -1;
    drop_graphs(sec);
    return ED_OK;
}

int ed_sector_set_fly(ed_sector *sec, int on)
{
    ED_TRY(check_common(sec, 0));
    if (!sec->sz) return fail(ED_ERR_ARG, "ed_sector_set_fly: S_z sectors only (HAIMsz_mult_fly)");
    sec->sz_fly = on ? 1 : 0;
    drop_graphs(sec);
    return ED_OK;
}

int ed_lanczos_benchmark(ed_sector *sec, int precision, int niter, int reps, double *avg_ms, double *bytes_per_run,
                         int *persistent)
{
    ED_TRY(check_common(sec, precision));
    if (niter < 1 || niter > kMaxIter - 2 || reps < 1 || !avg_ms) return fail(ED_ERR_ARG, "ed_lanczos_benchmark: bad argument");
    if (sec->dim < 2) return fail(ED_ERR_ARG, "ed_lanczos_benchmark: sector of dimension 1");
    hipStream_t st = sec->stream;
    ED_TRY(ensure_state(sec));
    ED_TRY(ensure_basis(sec, precision, niter + 1));
    sec->basis_valid_iters = 0;
    PersistShape ps;
    const bool one = niter >= kPersistMinSteps &&
                     (precision == ED_PREC_F64 ? persist_shape<double>(sec, ps) : persist_shape<float>(sec, ps));
    double tot = 0.0;
    for (int r = 0; r <= reps; ++r) {   // r = 0: warm-up (graph capture, first launch)
        ED_TRY(lz_start(sec, precision, nullptr, nullptr));
        ED_HIP_CHECK(hipEventRecord(sec->ev0, st));
        if (one) {
            ED_TRY(steps(sec, precision, niter, 8, 0));
        } else {
            for (int done = 0; done < niter; done += 8) ED_TRY(steps(sec, precision, std::min(8, niter - done), 8, 0));
        }
        ED_HIP_CHECK(hipEventRecord(sec->ev1, st));
        ED_HIP_CHECK(hipEventSynchronize(sec->ev1));
        float ms = 0.f;
        ED_HIP_CHECK(hipEventElapsedTime(&ms, sec->ev0, sec->ev1));
        if (r > 0) tot += ms;
    }
    HostLz h;
    ED_TRY(read_state(sec, h, niter));
    if (h.status != 0) return fail(ED_ERR_BREAKDOWN, "ed_lanczos_benchmark: the recurrence broke down");
    const double s = precision == ED_PREC_F64 ? 8.0 : 4.0;
    *avg_ms = tot / reps;
    if (bytes_per_run) *bytes_per_run = 5.0 * s * (double)sec->dim * niter;
    if (persistent) *persistent = one ? 1 : 0;
    return ED_OK;
}

int ed_lanczos_set_persistent(ed_sector *sec, int enable)
{
    ED_TRY(check_common(sec, 0));
    if (enable < 0 || enable > 5) return fail(ED_ERR_ARG, "ed_lanczos_set_persistent: enable must be 0 .. 5");
    sec->persist = enable >= 3 ? 1 : enable;
    sec->persist_cols = enable == 3 ? 0 : 1;
    sec->persist_1b = (enable == 1 || enable == 2 || enable == 5) ? 1 : 0;
    sec->persist_exact_beta = enable == 5 ? 1 : 0;
    sec->persist_ok[0] = sec->persist_ok[1] = -1;
    return ED_OK;
}

int ed_lanczos_persistent_active(ed_sector *sec, int precision, int *active)
{
    ED_TRY(check_common(sec, precision));
    if (!active) return fail(ED_ERR_ARG, "ed_lanczos_persistent_active: NULL output");
    if (!sec->cpb) ED_TRY(tune(sec));
    PersistShape p;
    *active = precision == ED_PREC_F64 ? persist_shape<double>(sec, p) : persist_shape<float>(sec, p);
    return ED_OK;
}

int ed_lanczos_persistent_shape(ed_sector *sec, int precision, int *cols_per_wg, int *workgroups, int *rows,
                                int *width)
{
    ED_TRY(check_common(sec, precision));
    if (!sec->cpb) ED_TRY(tune(sec));
    PersistShape p;
    const bool on = precision == ED_PREC_F64 ? persist_shape<double>(sec, p) : persist_shape<float>(sec, p);
    const HvShape h = hv_shape(sec, precision == ED_PREC_F64 ? sched_lz<double>(sec) : sched_lz<float>(sec));
    if (cols_per_wg) *cols_per_wg = on ? std::max(1, p.cpw) : 0;
    if (workgroups) *workgroups = on ? p.nblk : 0;
    if (rows) *rows = on ? p.rows : 0;
    if (width) *width = on ? h.k : 0;
    return ED_OK;
}

int ed_lanczos_persistent_refusals(ed_sector *sec, int *refused)
{
    ED_TRY(check_common(sec, 0));
    if (!refused) return fail(ED_ERR_ARG, "ed_lanczos_persistent_refusals: NULL output");
    *refused = sec->persist_refused;
    return ED_OK;
}

int ed_lanczos_exact_beta_steps(ed_sector *sec, int *steps)
{
    ED_TRY(check_common(sec, 0));
    if (!steps) return fail(ED_ERR_ARG, "ed_lanczos_exact_beta_steps: NULL output");
    *steps = 0;
    if (!sec->d_state) return ED_OK;
    ED_HIP_CHECK(hipStreamSynchronize(sec->stream));
    ED_HIP_CHECK(hipMemcpy(steps, (const char *)sec->d_state + offsetof(LzState, exact_beta_steps), sizeof(int),
                           hipMemcpyDeviceToHost));
    return ED_OK;
}

int ed_start_vector(int64_t n, double *out)
{
    if (n < 1 || !out) return fail(ED_ERR_ARG, "ed_start_vector: bad argument");
    for (int64_t i = 0; i < n; ++i) {
        uint64_t x = ((uint64_t)(i % 10000) + 1ull) * 0x9E3779B97F4A7C15ull;
        x ^= x >> 29;
        x *= 0xBF58476D1CE4E5B9ull;
        x ^= x >> 32;
        const uint64_t m = x >> 11;
        out[i] = (double)(((m * 1000ull) >> 53) + 1ull) / 1000.0;
    }
    return ED_OK;
}

int ed_lanczos_diagonalize(ed_sector *sec, const ed_lanczos_opts *opts, const double *start, int *iters,
                           int *neig_out, double *evals, double *vecp, double *diag, double *subdiag)
{
    return lanczos_diagonalize(sec, opts, start, iters, neig_out, evals, vecp, diag, subdiag);
}

int ed_lanczos_info(ed_sector *sec, int *iters_f32_phase)
{
    ED_TRY(check_common(sec, 0));
    if (iters_f32_phase) *iters_f32_phase = sec->last_f32_iters;
    return ED_OK;
}

int ed_lanczos_rq_info(ed_sector *sec, int *accepted, double *res2, double *gap)
{
    ED_TRY(check_common(sec, 0));
    if (accepted) *accepted = sec->last_rq_accepted;
    if (res2) *res2 = sec->last_rq_res2;
    if (gap) *gap = sec->last_rq_gap;
    return ED_OK;
}

int ed_lanczos_steps_run(ed_sector *sec, int *steps)
{
    ED_TRY(check_common(sec, 0));
    if (!steps) return fail(ED_ERR_ARG, "ed_lanczos_steps_run: NULL argument");
    *steps = sec->last_steps_run;
    return ED_OK;
}

int ed_lanczos_run(ed_sector *sec, int niter, const double *start, int precision, double *diag, double *subdiag)
{
    return lanczos_run(sec, niter, start, nullptr, precision, 8, diag, subdiag, nullptr);
}

int ed_lanczos_converged(int n, const double *alpha, const double *beta, int neigen, double tolerance, int nitermax,
                         int *iters, double *evals)
{
    if (n < 1 || !alpha || !beta || neigen < 1 || nitermax < 1 || !iters || !evals)
        return fail(ED_ERR_ARG, "ed_lanczos_converged: bad argument");
    // Niter = min(dimen_H - 1, Nitermax) is the caller's n when the run was that long; the
    // test reads beta one past the current step, so the copies end with a zero
    ConvTest ct(neigen, std::min(n, nitermax), tolerance);
    std::vector<double> a(alpha, alpha + n), b(beta, beta + n);
    a.push_back(0.0);
    b.push_back(0.0);
    int it = 1;
    for (; it <= std::min(n, nitermax); ++it)
        if (ct.ends(it, a.data(), b.data())) break;
    it = std::min(it, std::min(n, nitermax));
    *iters = it;
    const int m = std::min(neigen, it);
    for (int k = 0; k < neigen; ++k) evals[k] = k < m ? ct.w[k] : 0.0;
    return ED_OK;
}

int ed_lanczos_gs_vector_dev(ed_sector *sec, int iters, const double *vecp, double *d_psi)
{
    return gs_vector_dev(sec, iters, vecp, d_psi);
}

int ed_lanczos_gs_vector(ed_sector *sec, int iters, const double *vecp, double *psi)
{
    ED_TRY(check_common(sec, 0));
    if (!psi) return fail(ED_ERR_ARG, "ed_lanczos_gs_vector: NULL psi");
    ED_TRY(ensure_io(sec));
    ED_TRY(gs_vector_dev(sec, iters, vecp, sec->d_io_b));
    ED_HIP_CHECK(hipMemcpyAsync(psi, sec->d_io_b, sizeof(double) * sec->dim, hipMemcpyDeviceToHost, sec->stream));
    ED_HIP_CHECK(hipStreamSynchronize(sec->stream));
    return ED_OK;
}

int ed_rayleigh_f64(ed_sector *sec, int64_t n, const double *psi, double *energy)
{
    ED_TRY(check_common(sec, 0));
    if (n != sec->dim || !psi || !energy) return fail(ED_ERR_ARG, "ed_rayleigh_f64: bad argument");
    ED_TRY(ensure_io(sec));
    ED_TRY(ensure_state(sec));
    hipStream_t st = sec->stream;
    ED_HIP_CHECK(hipMemcpyAsync(sec->d_io_a, psi, sizeof(double) * n, hipMemcpyHostToDevice, st));
    ED_TRY(launch_hv(sec, ED_PREC_F64, sec->d_io_a, sec->d_io_b, 0, 1, st));
    double *d_r = sec->d_partials + kMisc + kMaxIter;
    dot_dev(sec, sec->d_io_a, sec->d_io_b, n, d_r, st);
    dot_dev(sec, sec->d_io_a, sec->d_io_a, n, d_r + 1, st);
    double h[2];
    ED_HIP_CHECK(hipMemcpyAsync(h, d_r, sizeof(h), hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipStreamSynchronize(st));
    *energy = h[0] / h[1];
    return ED_OK;
}

int ed_hmult_benchmark(ed_sector *sec, int precision, int reps, double *avg_ms, double *bytes_per_launch)
{
    ED_TRY(check_common(sec, precision));
    if (reps < 1 || reps > kMaxIter - 2 || !avg_ms) return fail(ED_ERR_ARG, "ed_hmult_benchmark: bad argument");
    if (sec->dim < 2) return fail(ED_ERR_ARG, "ed_hmult_benchmark: sector of dimension 1");
    hipStream_t st = sec->stream;
    ED_TRY(ensure_state(sec));
    ED_TRY(ensure_basis(sec, precision, reps + 1));
    sec->basis_valid_iters = 0;
    ED_TRY(lz_start(sec, precision, nullptr, nullptr));
    const int64_t n = sec->dim;
    const bool f64 = precision == ED_PREC_F64;
    const HvShape h = hv_shape(sec, f64 ? sched_lz<double>(sec) : sched_lz<float>(sec));
    // Phase 1 of step 1 (z = H u_1 and the partials of <z, u_1>), launched reps times back
    // to back between one event pair: every launch reads u_1 and writes z and the
    // partials again, the same traffic as a step of the recurrence, so the average is the
    // kernel's duration plus the dispatch gap (no per-launch event overhead).
    auto launch = [&]() {
        if (f64) {
            const LzArgs<double> lz = lz_args<double>(sec);
            hv_launch_mode<double, 1>(sec, h, (const double *)nullptr, (double *)nullptr, 0, 1, lz, st);
        } else {
            const LzArgs<float> lz = lz_args<float>(sec);
            hv_launch_mode<float, 1>(sec, h, (const float *)nullptr, (float *)nullptr, 0, 1, lz, st);
        }
    };
    launch();   // warm-up
    ED_HIP_CHECK(hipEventRecord(sec->ev0, st));
    for (int r = 0; r < reps; ++r) launch();
    ED_HIP_CHECK(hipEventRecord(sec->ev1, st));
    ED_HIP_CHECK(hipGetLastError());
    ED_HIP_CHECK(hipEventSynchronize(sec->ev1));
    float ms = 0.f;
    ED_HIP_CHECK(hipEventElapsedTime(&ms, sec->ev0, sec->ev1));
    const size_t esz = precision == ED_PREC_F64 ? sizeof(double) : sizeof(float);
    *avg_ms = (double)ms / reps;
    if (bytes_per_launch) *bytes_per_launch = 2.0 * esz * (double)n;
    return ED_OK;
}

}  // extern "C"

// Internal entry points used by the Green's-function translation unit.
namespace ed {
// ED_diago (ED_ARPACK.f90:590-687): the dense H of a small sector, column i = H e_i
// from the block H.v kernel, then all eigenpairs on the host by cyclic Jacobi rotations
// (Golub and Van Loan, Matrix Computations, sec. 8.5), ascending.
static int jacobi_eig(int n, std::vector<double> &A, std::vector<double> &w, std::vector<double> &V)
{
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = 0; q < n; ++q) {
                const double a = A[(size_t)q * n + p];
                tot += a * a;
                if (p != q) off += a * a;
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[(size_t)q * n + p];
                if (apq == 0.0) continue;
                const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {   // columns p, q
                    const double akp = A[(size_t)p * n + k], akq = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * akp - s * akq;
                    A[(size_t)q * n + k] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {   // rows p, q
                    const double apk = A[(size_t)k * n + p], aqk = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * apk - s * aqk;
                    A[(size_t)k * n + q] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[(size_t)p * n + k], vkq = V[(size_t)q * n + k];
                    V[(size_t)p * n + k] = c * vkp - s * vkq;
                    V[(size_t)q * n + k] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return A[(size_t)a * n + a] < A[(size_t)b * n + b]; });
    w.resize(n);
    std::vector<double> Vs((size_t)n * n);
    for (int k = 0; k < n; ++k) {
        w[k] = A[(size_t)order[k] * n + order[k]];
        const double *src = V.data() + (size_t)order[k] * n;
        int imax = 0;
        for (int r = 1; r < n; ++r)
            if (std::fabs(src[r]) > std::fabs(src[imax])) imax = r;
        const double sg = src[imax] < 0 ? -1.0 : 1.0;
        for (int r = 0; r < n; ++r) Vs[(size_t)k * n + r] = sg * src[r];
    }
    V.swap(Vs);
    return ED_OK;
}

}  // namespace ed

namespace ed {
int hermitian_from_embedding(int m, const std::vector<double> &w, const std::vector<double> &V, int neig,
                             double *evals, double *evecs);
}

extern "C" int ed_full_ed(ed_sector *sec, int neig, double *evals, double *evecs)
{
    using namespace ed;
    ED_TRY(check_common(sec, 0));
    const int64_t n = sec->dim;   // a complex sector: the planar length 2 dimen_H
    if (n < 1 || n > 1024) return fail(ED_ERR_ARG, "ed_full_ed: sector dimension must be in [1, 1024] (512 complex)");
    const int64_t nphys = sec->cplx ? n / 2 : n;
    if (neig < 1 || neig > nphys || !evals) return fail(ED_ERR_ARG, "ed_full_ed: neig outside [1, dim] or NULL evals");
    hipStream_t st = sec->stream;
    std::vector<double> eye((size_t)n * n, 0.0), H((size_t)n * n);
    for (int64_t i = 0; i < n; ++i) eye[(size_t)i * n + i] = 1.0;
    double *d = nullptr;
    ED_HIP_CHECK(hipMallocAsync((void **)&d, sizeof(double) * 2 * n * n, st));
    ED_HIP_CHECK(hipMemcpyAsync(d, eye.data(), sizeof(double) * n * n, hipMemcpyHostToDevice, st));
    int rc = launch_hv(sec, ED_PREC_F64, d, d + n * n, n, (int)n, st);
    if (rc == ED_OK) ED_HIP_CHECK(hipMemcpyAsync(H.data(), d + n * n, sizeof(double) * n * n, hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipStreamSynchronize(st));
    (void)hipFreeAsync(d, st);
    if (rc != ED_OK) return rc;
    std::vector<double> w, V;
    ED_TRY(jacobi_eig((int)n, H, w, V));
    if (!sec->cplx) {
        for (int k = 0; k < neig; ++k) {
            evals[k] = w[k];
            if (evecs) std::memcpy(evecs + (size_t)k * n, V.data() + (size_t)k * n, sizeof(double) * n);
        }
        return ED_OK;
    }
    return ed::hermitian_from_embedding((int)nphys, w, V, neig, evals, evecs);
}

namespace ed {

// Eigenpairs of a Hermitian H = A + iB of dimension m from those of its real embedding
// [[A, -B], [B, A]] (2m x 2m, eigenvalues w ascending, eigenvectors rows of V): every
// eigenvalue of H appears twice there, with the vectors (x; y) and (-y; x) of the one
// complex vector x + iy.  Within each cluster of equal eigenvalues (2q of them) the
// complex vectors x + iy of the cluster's real vectors span H's q-dimensional
// eigenspace; q orthonormal ones are kept by complex Gram-Schmidt.  evecs rows are
// planar (Re[m], Im[m]).
int hermitian_from_embedding(int m, const std::vector<double> &w, const std::vector<double> &V, int neig,
                             double *evals, double *evecs)
{
    const int n = 2 * m;
    double scale = 1.0;
    for (double x : w) scale = std::max(scale, std::fabs(x));
    const double tol = 1e-10 * scale;
    int got = 0;
    for (int k = 0; k < n && got < neig;) {
        int k2 = k + 1;
        while (k2 < n && w[k2] - w[k] <= tol) ++k2;
        const int q = (k2 - k + 1) / 2;
        double lam = 0.0;
        for (int t = k; t < k2; ++t) lam += w[t];
        lam /= (k2 - k);
        std::vector<std::vector<double>> acc;   // accepted, planar
        for (int t = k; t < k2 && (int)acc.size() < q; ++t) {
            std::vector<double> z(V.begin() + (size_t)t * n, V.begin() + (size_t)(t + 1) * n);
            for (const auto &a : acc) {   // z -= <a, z> a  (complex inner product on planar vectors)
                double re = 0.0, im = 0.0;
                for (int i = 0; i < m; ++i) {
                    re += a[i] * z[i] + a[m + i] * z[m + i];
                    im += a[i] * z[m + i] - a[m + i] * z[i];
                }
                for (int i = 0; i < m; ++i) {
                    const double ar = a[i], ai = a[m + i];
                    z[i] -= re * ar - im * ai;
                    z[m + i] -= re * ai + im * ar;
                }
            }
            double nrm = 0.0;
            for (double v : z) nrm += v * v;
            nrm = std::sqrt(nrm);
            if (nrm < 0.5) continue;   // i times a vector already kept
            for (double &v : z) v /= nrm;
            acc.push_back(std::move(z));
        }
        for (const auto &z : acc) {
            if (got >= neig) break;
            evals[got] = lam;
            if (evecs) std::memcpy(evecs + (size_t)got * n, z.data(), sizeof(double) * n);
            ++got;
        }
        k = k2;
    }
    return got == neig ? ED_OK : fail(ED_ERR_HIP, "ed_full_ed: fewer complex eigenvectors than requested");
}

int lanczos_run_dev(ed_sector *s, int niter, const double *d_start, int prec, double *diag, double *subdiag,
                    int *breakdown_at)
{
    return lanczos_run(s, niter, nullptr, d_start, prec, 8, diag, subdiag, breakdown_at);
}

template <typename T>
static void run_ws_impl(ed_sector *s, const LzWork &w, int sched, int niter, const double *d_start)
{
    const LzArgs<T> lz = lz_args_at<T>(s, w, sched);
    hipLaunchKernelGGL(k_lz_init<T>, dim3(lz.nB), dim3(kBlock), 0, s->stream, lz, d_start, s->dim);
    launch_steps_at<T>(s, lz, hv_shape(s, sched), w.partials, niter, 0, s->stream);
}

int lanczos_run_ws(ed_sector *s, const LzWork &w, int niter, const double *d_start, int prec, double *diag,
                   double *subdiag, int *breakdown_at)
{
    if (breakdown_at) *breakdown_at = 0;
    ED_TRY(check_common(s, prec));
    if (niter < 1 || niter > kMaxIter - 2) return fail(ED_ERR_ARG, "lanczos run: niter outside [1, 2046]");
    if (w.nslots < 3 || w.ld < s->dim) return fail(ED_ERR_ARG, "lanczos run: work space below 3 vectors of dimen_H");
    // the tuned schedule of that precision, or the default of an untuned sector (the fast
    // kernel in natural column order where it applies): neither needs an allocation
    const int tuned = prec == ED_PREC_F32 ? sched_of<float>(s) : sched_of<double>(s);
    const int sched = tuned ? tuned : (fast_ok(s) ? 8 : 1);
    if (prec == ED_PREC_F64) run_ws_impl<double>(s, w, sched, niter, d_start);
    else run_ws_impl<float>(s, w, sched, niter, d_start);
    ED_HIP_CHECK(hipGetLastError());
    int hdr[3];
    ED_HIP_CHECK(hipMemcpyAsync(hdr, w.st, 3 * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    ED_HIP_CHECK(hipMemcpyAsync(diag, (char *)w.st + offsetof(LzState, alpha), sizeof(double) * niter,
                                hipMemcpyDeviceToHost, s->stream));
    ED_HIP_CHECK(hipMemcpyAsync(subdiag, (char *)w.st + offsetof(LzState, beta), sizeof(double) * niter,
                                hipMemcpyDeviceToHost, s->stream));
    ED_HIP_CHECK(hipStreamSynchronize(s->stream));
    if (hdr[2] == 2) return fail(ED_ERR_BREAKDOWN, "Lanczos start vector has zero norm (Lanczos_plain.h:70-76)");
    const int reached = hdr[0] - 1;   // completed steps
    for (int i = 0; i < niter; ++i) {
        if (i >= reached) diag[i] = 0.0;
        if (i > reached) subdiag[i] = 0.0;
    }
    if (breakdown_at && hdr[2] == 1) *breakdown_at = reached;
    if (hdr[2] == 1 && reached < niter)
        return fail(ED_ERR_BREAKDOWN, "zero-norm Lanczos vector at iteration " + std::to_string(reached + 1) +
                                          " (Lanczos_plain.h:85-92)");
    return ED_OK;
}
}  // namespace ed
