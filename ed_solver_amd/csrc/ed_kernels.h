// Device-side bodies of the H.v and Lanczos kernels, shared by the single-sector
// kernels (ed_engine.hip) and the batched ones (ed_batch.hip).  A body gets the
// index of its workgroup inside its own job (`blk`) and the number of workgroups the
// job owns (`nblk`); the single-sector wrappers pass blockIdx.x / gridDim.x.
#pragma once

#include "ed_internal.h"

namespace ed {

__device__ inline double wave_sum(double v)
{
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block sum in a fixed order, returned to every thread.  `red` holds blockDim/64 + 1.
__device__ inline double block_sum_all(double v, double *red)
{
    v = wave_sum(v);
    const int nw = blockDim.x >> 6;
    __syncthreads();   // red may still be read from a previous call
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < nw; ++i) t += red[i];
        red[nw] = t;
    }
    __syncthreads();
    return red[nw];
}

// Sum of n partials in a fixed order (thread-strided, then the block tree); every
// workgroup that calls it on the same data obtains the same value.
__device__ inline double sum_partials(const double *p, int n, double *red)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += p[i];
    return block_sum_all(v, red);
}

// Fixed-order sum of n <= kGroup*kMaxGroups partial sums, in two steps so that the
// loads are issued together and the order does not depend on the workgroup size:
//   grp_issue:  group g = entries [16g, 16g+16) summed left to right into grp[g]
//               (threads take groups g = tid, tid + blockDim.x, ...);
//   (barrier)
//   grp_total:  the groups summed left to right (every thread, LDS broadcast reads).
// Every kernel that reduces the same partials therefore obtains the same bits.
__device__ inline void grp_issue(const double *p, int n, double *grp)
{
    const int ng = (n + kGroup - 1) / kGroup;
    for (int g = threadIdx.x; g < ng; g += blockDim.x) {
        double v[kGroup];
#pragma unroll
        for (int q = 0; q < kGroup; ++q) {
            const int i = g * kGroup + q;
            v[q] = i < n ? p[i] : 0.0;
        }
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < kGroup; ++q) s += v[q];
        grp[g] = s;
    }
}

__device__ inline double grp_total(const double *grp, int n)
{
    const int ng = (n + kGroup - 1) / kGroup;
    double s = 0.0;
    for (int g = 0; g < ng; ++g) s += grp[g];
    return s;
}

// Element range of workgroup blk of an element-wise Lanczos kernel (see LzArgs::colmap).
template <typename T>
__device__ inline void elem_range(const LzArgs<T> &lz, int64_t n, int blk, int nblk, int64_t &i0, int64_t &i1,
                                  int64_t &step)
{
    if (lz.colmap) {
        i0 = (int64_t)blk * lz.dup + threadIdx.x;
        i1 = min(n, (int64_t)(blk + 1) * lz.dup);
        step = blockDim.x;
    } else {
        i0 = (int64_t)blk * blockDim.x + threadIdx.x;
        i1 = n;
        step = (int64_t)nblk * blockDim.x;
    }
}

// Prologue of a Lanczos H.v phase: beta_j from the update phase's partials, breakdown
// test (Lanczos_plain.h:70-92), scale factors.  Returns false when the step is skipped.
// `lead` is true in exactly one thread of the job (it publishes beta_j and j).
template <typename T>
__device__ inline bool lz_prologue(const LzArgs<T> &lz, double *red, double *grp, bool lead, int &j, double &sj,
                                   double &cprev)
{
    if (lz.st->status) return false;
    j = lz.st->j_k1;
    grp_issue(lz.partB, lz.nB, grp);
    __syncthreads();
    const double bj = sqrt(grp_total(grp, lz.nB));
    if (bj < kMinLanczosNorm) {
        if (lead) lz.st->status = (j == 1) ? 2 : 1;
        return false;
    }
    if (lead) {
        lz.st->beta[j - 1] = bj;
        lz.st->j_k2 = j;
    }
    sj = 1.0 / bj;
    cprev = (j > 1) ? bj / lz.st->beta[j - 2] : 0.0;
    return true;
}

// Deterministic block sum of one value per thread into `out` (thread 0 only): one
// butterfly per wave, one barrier, waves added in order by thread 0.
__device__ inline void block_partial(double v, double *red, double *out)
{
    v = wave_sum(v);
    __syncthreads();   // red may still be in use
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
        *out = t;
    }
}

// H.v for CPB columns starting at column blk*CPB, rows owned by threads.
// MODE 0: y = H x.  MODE 1: Lanczos phase 1 (x, y from the Krylov slots).
// CLAMP: the ELL width of this job may be below KU (batched kernels); entries k >=
// a.kup reread entry a.kup-1 with a zero coefficient, so no load leaves the table.
template <typename T, int MODE, int CPB, int KU, bool CLAMP>
__device__ inline void hv_rows_body(const HvArgs<T> &a, const T *__restrict__ xin, T *__restrict__ yout,
                                    const LzArgs<T> &lz, int blk, T *col, double *red, double *grp)
{
    const int dup = a.dup, ncfg = a.ncfg;
    const int c0 = blk * CPB;
    const T *__restrict__ x;
    T *__restrict__ y;
    const T *__restrict__ prev = nullptr;
    double sj = 1.0, cprev = 0.0;
    int j = 0;
    if (MODE == 1) {
        if (!lz_prologue(lz, red, grp, blk == 0 && threadIdx.x == 0, j, sj, cprev)) return;
        x = lz.slot(j);
        y = lz.slot(j + 1);
        if (j > 1) prev = lz.slot(j - 1);
    } else {
        x = xin;
        y = yout;
    }
    int ido[CPB], cd[CPB], p0[CPB], p1[CPB];
    T ddg[CPB];
    bool ok[CPB];
#pragma unroll
    for (int c = 0; c < CPB; ++c) {
        ok[c] = c0 + c < a.ddn;
        ido[c] = ok[c] ? c0 + c : c0;
        cd[c] = a.imp_dn[ido[c]];
        p0[c] = a.ptr_dn[ido[c]];
        p1[c] = ok[c] ? a.ptr_dn[ido[c] + 1] : p0[c];
        ddg[c] = a.diag_dn[ido[c]];
    }
    const int ncol = min(CPB, a.ddn - c0);
    const T *xc = x + (int64_t)c0 * dup;
    for (int e = threadIdx.x; e < ncol * dup; e += blockDim.x) col[e] = xc[e];
    __syncthreads();

    double dot = 0.0;
    for (int r = threadIdx.x; r < dup; r += blockDim.x) {
        int rk[KU];
        T v[KU];
#pragma unroll
        for (int k = 0; k < KU; ++k) {
            const int kk = CLAMP ? min(k, a.kup - 1) : k;
            rk[k] = a.ell_rk_up[kk * dup + r];
            const T vv = a.ell_val_up[kk * dup + r];
            v[k] = (CLAMP && k >= a.kup) ? (T)0 : vv;
        }
        const T du = a.diag_up[r];
        const int cu = a.imp_up[r];
#pragma unroll
        for (int c = 0; c < CPB; ++c) {
            if (!ok[c]) continue;
            const T *cv = col + c * dup;
            const T xv = cv[r];
            T acc = (du + ddg[c] + a.uimp[cu * ncfg + cd[c]]) * xv;
#pragma unroll
            for (int k = 0; k < KU; ++k) acc += v[k] * cv[rk[k]];
#pragma unroll 4
            for (int k = p0[c]; k < p1[c]; ++k) acc += a.off_dn[k] * x[r + a.rk_dn[k] * dup];
            if (a.hund) {
                const int h0 = a.hund_ptr[cu * ncfg + cd[c]], h1 = a.hund_ptr[cu * ncfg + cd[c] + 1];
                for (int t = h0; t < h1; ++t)
                    acc += a.hund_val[t] *
                           x[a.swap_up[r * ncfg + a.hund_cu[t]] + a.swap_dn[ido[c] * ncfg + a.hund_cd[t]] * dup];
            }
            const int idx = (c0 + c) * dup + r;
            if (MODE == 1) {
                double yv = (double)acc * sj;
                if (j > 1) yv -= cprev * (double)prev[idx];
                const T yt = (T)yv;
                y[idx] = yt;
                dot += (double)yt * ((double)xv * sj);
            } else {
                y[idx] = acc;
            }
        }
    }
    if (MODE == 1) block_partial(dot, red, lz.partA + blk);
}

// Bytes of dynamic LDS the fast H.v needs (column + coefficient/interaction tables).
__host__ __device__ inline size_t hv_fast_lds(int dup, int ncoef, int ncfg, int nhund, size_t tsize)
{
    const size_t nq = (size_t)ncfg * ncfg;
    return tsize * ((size_t)dup + ncoef + nq + nhund) + sizeof(int) * (nq + 1 + nhund + ncfg);
}

// Fast H.v: one column per workgroup, blockDim.x threads own rows r = tid + i*blockDim.x
// (ONE: at most one row per thread).  All loads that do not depend on the staged column
// (packed up-hop entries, down-hop neighbour values, diagonal, previous Lanczos vector)
// are issued before the one LDS barrier, together with the column staging and, in
// Lanczos mode, the reduction of beta_j by wave 0.  Up-hop coefficients, interaction
// diagonal and Hund transitions are read from LDS.  K bounds the entries per row of both
// spin tables.  Dynamic LDS: col[dup] coef[ncoef] uimp[ncfg^2] hval[nhund] (T), then
// hptr[ncfg^2+1] htr[nhund] swdn[ncfg] (int).
template <typename T, int MODE, int K, bool ONE>
__device__ inline void hv_fast_body(const HvArgs<T> &a, const T *__restrict__ xin, T *__restrict__ yout,
                                    const LzArgs<T> &lz, int blk, unsigned char *smem, double *red, double *grp)
{
    const int dup = a.dup, ncfg = a.ncfg, nq = ncfg * ncfg, nh = a.nhund;
    T *col = reinterpret_cast<T *>(smem);
    T *coef = col + dup;
    T *uimp = coef + a.ncoef_up;
    T *hval = uimp + nq;
    int *hptr = reinterpret_cast<int *>(hval + nh);
    int *htr = hptr + nq + 1;
    int *swdn = htr + nh;
    const int ido = blk, NT = blockDim.x, tid = threadIdx.x;
    const T *__restrict__ x;
    T *__restrict__ y;
    const T *__restrict__ prev = nullptr;
    int j = 0;
    if (MODE == 1) {
        if (lz.st->status) return;
        j = lz.st->j_k1;
        x = lz.slot(j);
        y = lz.slot(j + 1);
        if (j > 1) prev = lz.slot(j - 1);
    } else {
        x = xin;
        y = yout;
    }
    const int cd = a.imp_dn[ido];
    const int p0 = a.ptr_dn[ido], nd = a.ptr_dn[ido + 1] - p0;
    const T ddg = a.diag_dn[ido];
    T od[K];
    int jd[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        od[k] = k < nd ? a.off_dn[p0 + k] : (T)0;
        jd[k] = k < nd ? a.rk_dn[p0 + k] * dup : ido * dup;
    }
    const int base = ido * dup;

    // ---- loads of the first row, issued before the column is staged
    int r = tid;
    uint32_t pk[K];
    T xd[K];
    T du = 0, pv = 0;
    int cu = 0;
    if (r < dup) {
#pragma unroll
        for (int k = 0; k < K; ++k) pk[k] = a.ell_pk_up[k * dup + r];
#pragma unroll
        for (int k = 0; k < K; ++k) xd[k] = k < nd ? x[r + jd[k]] : (T)0;
        du = a.diag_up[r];
        cu = a.imp_up[r];
        if (MODE == 1 && j > 1) pv = prev[base + r];
    }
    // ---- stage the column and the small tables in LDS; wave 0 sums beta_j^2
    for (int e = tid; e < dup; e += NT) col[e] = x[base + e];
    for (int e = tid; e < a.ncoef_up; e += NT) coef[e] = a.coef_up[e];
    for (int e = tid; e < nq; e += NT) uimp[e] = a.uimp[e];
    if (a.hund) {
        for (int e = tid; e <= nq; e += NT) hptr[e] = a.hund_ptr[e];
        for (int e = tid; e < nh; e += NT) {
            htr[e] = a.hund_tr[e];
            hval[e] = a.hund_val[e];
        }
        for (int e = tid; e < ncfg; e += NT) swdn[e] = a.swap_dn[ido * ncfg + e];
    }
    if (MODE == 1) grp_issue(lz.partB, lz.nB, grp);
    __syncthreads();
    double sj = 1.0, cprev = 0.0;
    if (MODE == 1) {
        const double bj = sqrt(grp_total(grp, lz.nB));
        if (bj < kMinLanczosNorm) {   // the reference STOPs here (Lanczos_plain.h:70-92)
            if (blk == 0 && tid == 0) lz.st->status = (j == 1) ? 2 : 1;
            return;
        }
        if (blk == 0 && tid == 0) {
            lz.st->beta[j - 1] = bj;
            lz.st->j_k2 = j;
        }
        sj = 1.0 / bj;
        cprev = (j > 1) ? bj / lz.st->beta[j - 2] : 0.0;
    }

    double dot = 0.0;
    while (r < dup) {
        const T xv = col[r];
        T acc = (du + ddg + uimp[cu * ncfg + cd]) * xv;
#pragma unroll
        for (int k = 0; k < K; ++k) acc += coef[pk[k] >> 16] * col[pk[k] & 0xFFFFu];
#pragma unroll
        for (int k = 0; k < K; ++k) acc += od[k] * xd[k];
        if (a.hund) {
            const int h0 = hptr[cu * ncfg + cd], h1 = hptr[cu * ncfg + cd + 1];
            for (int t = h0; t < h1; ++t) {
                const int tr = htr[t];
                acc += hval[t] * x[a.swap_up[r * ncfg + (tr & 0xFF)] + swdn[tr >> 8] * dup];
            }
        }
        const int idx = base + r;
        if (MODE == 1) {
            double yv = (double)acc * sj;
            if (j > 1) yv -= cprev * (double)pv;
            const T yt = (T)yv;
            y[idx] = yt;
            dot += (double)yt * ((double)xv * sj);
        } else {
            y[idx] = acc;
        }
        if (ONE) break;
        // next row of this thread
        r += NT;
        if (r < dup) {
#pragma unroll
            for (int k = 0; k < K; ++k) pk[k] = a.ell_pk_up[k * dup + r];
#pragma unroll
            for (int k = 0; k < K; ++k) xd[k] = k < nd ? x[r + jd[k]] : (T)0;
            du = a.diag_up[r];
            cu = a.imp_up[r];
            if (MODE == 1 && j > 1) pv = prev[base + r];
        }
    }
    if (MODE == 1) block_partial(dot, red, lz.partA + blk);
}

// u_1 = start (converted to T); partial sums of ||u_1||^2; j = 1.
template <typename T>
__device__ inline void lz_init_body(const LzArgs<T> &lz, const double *start, int64_t n, int blk, int nblk,
                                    double *red)
{
    T *u = lz.slot(1);
    int64_t i0, i1, st;
    elem_range(lz, n, blk, nblk, i0, i1, st);
    double acc = 0.0;
    for (int64_t i = i0; i < i1; i += st) {
        const T v = (T)start[i];
        u[i] = v;
        acc += (double)v * (double)v;
    }
    block_partial(acc, red, lz.partB + blk);
    if (blk == 0 && threadIdx.x == 0) {
        lz.st->j_k1 = 1;
        lz.st->j_k2 = 0;
        lz.st->status = 0;
    }
}

// Lanczos phase 2: alpha_j from the H.v partials, w -= alpha_j v_j, partial ||w||^2.
// The first kPre element pairs of every thread are loaded before alpha_j is known.
template <typename T>
__device__ inline void lz_update_body(const LzArgs<T> &lz, int64_t n, int advance, int blk, int nblk, double *red,
                                      double *grp)
{
    constexpr int kPre = 4;
    if (lz.st->status) return;
    const int j = lz.st->j_k2;
    T *__restrict__ y = lz.slot(j + 1);
    const T *__restrict__ x = lz.slot(j);
    int64_t i0, i1, st;
    elem_range(lz, n, blk, nblk, i0, i1, st);
    T yv[kPre], xv[kPre];
#pragma unroll
    for (int q = 0; q < kPre; ++q) {
        const int64_t i = i0 + q * st;
        yv[q] = i < i1 ? y[i] : (T)0;
        xv[q] = i < i1 ? x[i] : (T)0;
    }
    grp_issue(lz.partA, lz.nA, grp);
    __syncthreads();
    const double alpha = grp_total(grp, lz.nA);
    const bool lead = blk == 0 && threadIdx.x == 0;
    if (lead) lz.st->alpha[j - 1] = alpha;
    const double c = alpha / lz.st->beta[j - 1];
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < kPre; ++q) {
        const int64_t i = i0 + q * st;
        if (i < i1) {
            const T v = (T)((double)yv[q] - c * (double)xv[q]);
            y[i] = v;
            acc += (double)v * (double)v;
        }
    }
    for (int64_t i = i0 + kPre * st; i < i1; i += st) {
        const T v = (T)((double)y[i] - c * (double)x[i]);
        y[i] = v;
        acc += (double)v * (double)v;
    }
    block_partial(acc, red, lz.partB + blk);
    if (advance && lead) lz.st->j_k1 = j + 1;
}

// End of a chunk: beta_j of the next step from the update partials (one workgroup), the
// same reduction the next H.v phase performs.
template <typename T>
__device__ inline void lz_finish_body(const LzArgs<T> &lz, double *grp)
{
    if (lz.st->status) return;
    const int j = lz.st->j_k1;
    grp_issue(lz.partB, lz.nB, grp);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double b = sqrt(grp_total(grp, lz.nB));
        lz.st->beta[j - 1] = b;
        if (b < kMinLanczosNorm) lz.st->status = (j == 1) ? 2 : 1;
    }
}

}  // namespace ed
