// Device-side bodies of the H.v and Lanczos kernels, shared by the single-sector
// kernels (ed_engine.hip) and the batched ones (ed_batch.hip).  A body gets the
// index of its workgroup inside its own job (`blk`) and the number of workgroups the
// job owns (`nblk`); the single-sector wrappers pass blockIdx.x / gridDim.x.
#pragma once

#include "ed_internal.h"

namespace ed {

// x from lane (lane - k) of the same 16-lane row (row_shr:k), or 0 when that lane is
// outside the row; dpp_ctrl as the DPP encoding of the instruction.
template <int CTRL, int ROW_MASK = 0xF, int BANK_MASK = 0xF>
__device__ inline double dpp_shift(double x)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROW_MASK, BANK_MASK, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROW_MASK, BANK_MASK, true);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Sum over the 64 lanes of a wave in a fixed order, returned in every lane with the same
// bits: within each 16-lane row, lane i adds the partial sums of lanes i-1, i-2, i-4, i-8
// (DPP row shifts), so lane 15 of a row holds the row's sum; row_bcast15 adds row 0 into
// row 1 and row 2 into row 3, row_bcast31 adds lane 31 into rows 2 and 3, leaving the wave's
// sum in lane 63, which is broadcast.  DPP moves stay in the vector ALU, off the LDS
// crossbar a shuffle would use.
__device__ inline double wave_sum(double v)
{
    v += dpp_shift<0x111>(v);               // row_shr:1
    v += dpp_shift<0x112>(v);               // row_shr:2
    v += dpp_shift<0x114, 0xF, 0xE>(v);     // row_shr:4, banks 1-3
    v += dpp_shift<0x118, 0xF, 0xC>(v);     // row_shr:8, banks 2-3
    v += dpp_shift<0x142, 0xA>(v);          // row_bcast:15 into rows 1, 3
    v += dpp_shift<0x143, 0xC>(v);          // row_bcast:31 into rows 2, 3
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), 63);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
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

// Fixed-order sum of n partial sums by ONE whole wave (64 lanes): lane l adds entries
// l, l+64, l+128, ... left to right (sixteen coalesced loads in flight per batch), then
// the xor butterfly.  Addition is commutative, so the butterfly leaves the same bits in
// every lane, and the order depends only on n: every kernel that reduces the same
// partials with this function obtains the same value.
// Device-scope coherent access (agent-scope relaxed atomics: no cache fences needed) for
// data another workgroup of the same kernel writes.
template <typename T>
__device__ inline T ld_agent(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ inline void st_agent(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// COH: the partials come from other workgroups of the running kernel (coherent loads).
template <bool COH = false>
__device__ inline double wave_part_sum(const double *p, int n)
{
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    for (int b0 = 0; b0 < n; b0 += 1024) {
        double v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int i = b0 + q * 64 + lane;
            v[q] = i < n ? (COH ? ld_agent(p + i) : p[i]) : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) s += v[q];
    }
    return wave_sum(s);
}

// Element range of workgroup blk of an element-wise Lanczos kernel (see LzArgs::colmap).
template <typename T>
__device__ inline void elem_range(const LzArgs<T> &lz, int64_t n, int blk, int nblk, int64_t &i0, int64_t &i1,
                                  int64_t &step)
{
    if (lz.colmap) {
        const int c = lz.colord ? lz.colord[blk] : blk;
        step = blockDim.x;
        if (c < 0) {   // idle workgroup of the XCD-grouped order
            i0 = 0;
            i1 = 0;
            return;
        }
        i0 = (int64_t)c * lz.dup + threadIdx.x;
        i1 = min(n, (int64_t)(c + 1) * lz.dup);
    } else {
        i0 = (int64_t)blk * blockDim.x + threadIdx.x;
        i1 = n;
        step = (int64_t)nblk * blockDim.x;
    }
}

// Start of a Lanczos H.v phase: skipped once the run has stopped; j = step index.
// `lead` (one thread of the job) hands j on to the update phase.
template <typename T>
__device__ inline bool lz_hv_begin(const LzArgs<T> &lz, bool lead, int &j)
{
    if (lz.st->status) return false;
    j = lz.st->j_k1;
    if (lead) lz.st->j_k2 = j;
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

// Two deterministic block sums in one pass, each in the order of block_partial (out1 may
// be null).  COH: stored with device-scope coherent stores.
template <bool COH = false>
__device__ inline void block_partial2(double v0, double v1, double *red, double *out0, double *out1)
{
    v0 = wave_sum(v0);
    v1 = wave_sum(v1);
    __syncthreads();   // red may still be in use
    if ((threadIdx.x & 63) == 0) {
        red[2 * (threadIdx.x >> 6)] = v0;
        red[2 * (threadIdx.x >> 6) + 1] = v1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0.0, t1 = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
            t0 += red[2 * i];
            t1 += red[2 * i + 1];
        }
        if (COH) {
            st_agent(out0, t0);
            if (out1) st_agent(out1, t1);
        } else {
            *out0 = t0;
            if (out1) *out1 = t1;
        }
    }
}

// The same fixed-order block sum without a workgroup barrier: every wave stores its
// butterfly sum and bumps an LDS counter (`cnt`, zeroed before an earlier barrier); the
// wave that arrives last adds the wave sums in index order and stores the result.  Two
// values (v0 -> out0, v1 -> out1, out1 may be null) share one pass.
__device__ inline void block_partial_last(double v0, double v1, double *red, int *cnt, double *out0, double *out1)
{
    v0 = wave_sum(v0);
    if (out1) v1 = wave_sum(v1);
    if ((threadIdx.x & 63) == 0) {
        const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
        red[2 * w] = v0;
        red[2 * w + 1] = v1;
        __threadfence_block();
        if (atomicAdd(cnt, 1) == nw - 1) {
            __threadfence_block();
            const volatile double *rv = red;
            double t0 = 0.0, t1 = 0.0;
            for (int i = 0; i < nw; ++i) {
                t0 += rv[2 * i];
                t1 += rv[2 * i + 1];
            }
            *out0 = t0;
            if (out1) *out1 = t1;
        }
    }
}

// H.v for CPB columns starting at column blk*CPB, rows owned by threads.
// MODE 0: y = H x.  MODE 1: Lanczos phase 1 (x, y from the Krylov slots).
// CLAMP: the ELL width of this job may be below KU (batched kernels); entries k >=
// a.kup reread entry a.kup-1 with a zero coefficient, so no load leaves the table.
// With a.ell_pk_up set, the up hops come from the packed table (rank | coefficient
// index << 16, 4 bytes an entry instead of 12) and the coefficients from an LDS copy
// placed after the CPB columns (dynamic LDS: col[CPB dup], coef[ncoef_up]); same entries
// in the same order, so the same sums.
template <typename T, int MODE, int CPB, int KU, bool CLAMP>
__device__ inline void hv_rows_body(const HvArgs<T> &a, const T *__restrict__ xin, T *__restrict__ yout,
                                    const LzArgs<T> &lz, int blk, T *col, double *red, double *grp)
{
    const int dup = a.dup, ncfg = a.ncfg;
    const int c0 = blk * CPB;
    const T *__restrict__ x;
    T *__restrict__ y;
    int j = 0;
    if (MODE == 1) {
        if (!lz_hv_begin(lz, blk == 0 && threadIdx.x == 0, j)) return;
        x = lz.slot(j);
        y = lz.slot(j + 1);
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
    T *cf = col + CPB * dup;
    const bool pkd = a.ell_pk_up != nullptr;
    if (pkd)
        for (int e = threadIdx.x; e < a.ncoef_up; e += blockDim.x) cf[e] = a.coef_up[e];
    __syncthreads();

    double dot = 0.0;
    for (int r = threadIdx.x; r < dup; r += blockDim.x) {
        int rk[KU];
        T v[KU];
        if (pkd) {
            uint32_t pw[KU];
#pragma unroll
            for (int k = 0; k < KU; ++k) pw[k] = a.ell_pk_up[(CLAMP ? min(k, a.kup - 1) : k) * dup + r];
#pragma unroll
            for (int k = 0; k < KU; ++k) {
                rk[k] = (int)(pw[k] & 0xFFFFu);
                v[k] = (CLAMP && k >= a.kup) ? (T)0 : cf[pw[k] >> 16];
            }
        } else {
#pragma unroll
            for (int k = 0; k < KU; ++k) {
                const int kk = CLAMP ? min(k, a.kup - 1) : k;
                rk[k] = a.ell_rk_up[kk * dup + r];
                const T vv = a.ell_val_up[kk * dup + r];
                v[k] = (CLAMP && k >= a.kup) ? (T)0 : vv;
            }
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
            y[(c0 + c) * dup + r] = acc;
            if (MODE == 1) dot += (double)acc * (double)xv;
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
// (R >= 1: exactly the R rows i < R, all their loads in flight together; R == 0: a loop
// over the rows, each row's loads issued after the previous row is done).  All loads that do not depend on the staged column
// (packed up-hop entries, down-hop neighbour values, diagonal, previous Lanczos vector)
// are issued before the one LDS barrier, together with the column staging and, in
// Lanczos mode, the reduction of beta_j by wave 0.  Up-hop coefficients, interaction
// diagonal and Hund transitions are read from LDS.  K bounds the entries per row of both
// spin tables; the packed up table has K rows (padded with the zero coefficient when the
// down table is the wider one).  Dynamic LDS: col[dup] coef[ncoef] uimp[ncfg^2] hval[nhund] (T), then
// hptr[ncfg^2+1] htr[nhund] swdn[ncfg] (int).
template <typename T, int MODE, int K, int R>
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
    const int ido = a.colord ? a.colord[blk] : blk, NT = blockDim.x, tid = threadIdx.x;
    const T *__restrict__ x;
    T *__restrict__ y;
    int j = 0;
    if (MODE == 1 && !lz_hv_begin(lz, blk == 0 && tid == 0, j)) return;
    if (ido < 0) {   // idle workgroup of the XCD-grouped order: contributes a zero partial
        if (MODE == 1 && tid == 0) lz.partA[blk] = 0.0;
        return;
    }
    if (MODE == 1) {
        x = lz.slot(j);
        y = lz.slot(j + 1);
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

    // ---- per-row loads issued before the column is staged: all R rows of the thread
    // (R >= 1), or the first of a loop over rows tid, tid + NT, ... (R == 0)
    constexpr int RR = R > 0 ? R : 1;
    uint32_t pk[RR][K];
    T xd[RR][K];
    T du[RR];
    int cu[RR];
#pragma unroll
    for (int i = 0; i < RR; ++i) {
        const int r = tid + i * NT;
        du[i] = (T)0;
        cu[i] = 0;
        if (r < dup) {
#pragma unroll
            for (int k = 0; k < K; ++k) pk[i][k] = a.ell_pk_up[k * dup + r];
#pragma unroll
            for (int k = 0; k < K; ++k) xd[i][k] = k < nd ? x[r + jd[k]] : (T)0;
            du[i] = a.diag_up[r];
            cu[i] = a.imp_up[r];
        }
    }
    // ---- stage the column and the small tables in LDS
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
    __shared__ int cnt;
    if (MODE == 1 && tid == 0) cnt = 0;
    __syncthreads();

    double dot = 0.0;
    auto row = [&](int r, const uint32_t *pr, const T *xr, T dur, int cur) {
        const T xv = col[r];
        T acc = (dur + ddg + uimp[cur * ncfg + cd]) * xv;
#pragma unroll
        for (int k = 0; k < K; ++k) acc += coef[pr[k] >> 16] * col[pr[k] & 0xFFFFu];
#pragma unroll
        for (int k = 0; k < K; ++k) acc += od[k] * xr[k];
        if (a.hund) {
            const int h0 = hptr[cur * ncfg + cd], h1 = hptr[cur * ncfg + cd + 1];
            for (int t = h0; t < h1; ++t) {
                const int tr = htr[t];
                acc += hval[t] * x[a.swap_up[r * ncfg + (tr & 0xFF)] + swdn[tr >> 8] * dup];
            }
        }
        y[base + r] = acc;
        if (MODE == 1) dot += (double)acc * (double)xv;
    };
    if constexpr (R > 0) {
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int r = tid + i * NT;
            if (r < dup) row(r, pk[i], xd[i], du[i], cu[i]);
        }
    } else {
        for (int r = tid; r < dup;) {
            row(r, pk[0], xd[0], du[0], cu[0]);
            // next row of this thread
            r += NT;
            if (r < dup) {
#pragma unroll
                for (int k = 0; k < K; ++k) pk[0][k] = a.ell_pk_up[k * dup + r];
#pragma unroll
                for (int k = 0; k < K; ++k) xd[0][k] = k < nd ? x[r + jd[k]] : (T)0;
                du[0] = a.diag_up[r];
                cu[0] = a.imp_up[r];
            }
        }
    }
    if (MODE == 1) block_partial_last(dot, 0.0, red, &cnt, lz.partA + blk, nullptr);
}

// u_1 = start (converted to T); partial sums of ||u_1||^2 (and zero cross sums); j = 1.
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
    block_partial(acc, red, lz.nrm(1) + blk);
    if (threadIdx.x == 0) lz.crs(1)[blk] = 0.0;
    if (blk == 0 && threadIdx.x == 0) {
        lz.st->j_k1 = 1;
        lz.st->j_k2 = 0;
        lz.st->status = 0;
    }
}

// Recurrence scalars of step j, identical in every workgroup of the update phase.  From
// the fixed-order sums (one wave each, in parallel) of
//   N = ||u_j||^2,  P = <u_j, u_{j-1}>   (update phase of step j-1, or initialisation)
//   Q = <z, u_j>,   z = H u_j            (H.v phase of step j)
// with v_j = u_j / beta_j:
//   beta_j  = sqrt(N)                     (breakdown below kMinLanczosNorm, Lanczos_plain.h:70-92)
//   alpha_j = <H v_j - beta_j v_{j-1}, v_j> = Q / beta_j^2 - P / beta_{j-1}
// which is lanc_opt's alpha (taken after the beta-subtraction, Lanczos_plain.h:116-140)
// rearranged so that the H.v phase needs no scalar of step j.  sh[] receives N, P, Q.
template <typename T, bool COH = false>
__device__ inline void lz_scalars(const LzArgs<T> &lz, int j, double *sh)
{
    const int w = threadIdx.x >> 6;
    if (w == 0) {
        const double t = wave_part_sum<COH>(lz.nrm(j), lz.nB);
        if (threadIdx.x == 0) sh[0] = t;
    } else if (w == 1) {
        const double t = j > 1 ? wave_part_sum<COH>(lz.crs(j), lz.nB) : 0.0;
        if (threadIdx.x == 64) sh[1] = t;
    } else if (w == 2) {
        const double t = wave_part_sum<COH>(lz.partA, lz.nA);
        if (threadIdx.x == 128) sh[2] = t;
    }
}

// Lanczos phase 2 of step j (lanc_opt, Lanczos_plain.h:66-173): with z = H u_j in slot
// j+1,  u_{j+1} = z / beta_j - (beta_j / beta_{j-1}) u_{j-1} - (alpha_j / beta_j) u_j,
// written over z; partial sums of ||u_{j+1}||^2 and <u_{j+1}, u_j> for step j+1.  The
// first kPre elements of every thread are loaded before the scalars are known.
template <typename T>
__device__ inline void lz_update_body(const LzArgs<T> &lz, int64_t n, int advance, int blk, int nblk, double *red,
                                      double *sh)
{
    constexpr int kPre = 4;
    if (lz.st->status) return;
    const int j = lz.st->j_k2;
    T *__restrict__ y = lz.slot(j + 1);
    const T *__restrict__ x = lz.slot(j);
    const T *__restrict__ pr = lz.slot(j > 1 ? j - 1 : j);
    int64_t i0, i1, st;
    elem_range(lz, n, blk, nblk, i0, i1, st);
    T yv[kPre], xv[kPre], pv[kPre];
#pragma unroll
    for (int q = 0; q < kPre; ++q) {
        const int64_t i = i0 + q * st;
        yv[q] = i < i1 ? y[i] : (T)0;
        xv[q] = i < i1 ? x[i] : (T)0;
        pv[q] = (i < i1 && j > 1) ? pr[i] : (T)0;
    }
    const double bprev = j > 1 ? lz.st->beta[j - 2] : 1.0;
    lz_scalars(lz, j, sh);
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const double bj = sqrt(sh[0]);
    const bool lead = blk == 0 && threadIdx.x == 0;
    if (bj < kMinLanczosNorm) {   // the reference STOPs here (Lanczos_plain.h:70-92)
        if (lead) {
            lz.st->beta[j - 1] = bj;   // the test of iteration j - 1 reads it
            lz.st->status = (j == 1) ? 2 : 1;
        }
        return;
    }
    const double alpha = sh[2] / (bj * bj) - (j > 1 ? sh[1] / bprev : 0.0);
    if (lead) {
        lz.st->beta[j - 1] = bj;
        lz.st->alpha[j - 1] = alpha;
    }
    const double cz = 1.0 / bj, cp = j > 1 ? bj / bprev : 0.0, cx = alpha / bj;
    double nn = 0.0, cr = 0.0;
#pragma unroll
    for (int q = 0; q < kPre; ++q) {
        const int64_t i = i0 + q * st;
        if (i < i1) {
            const T v = (T)(cz * (double)yv[q] - cp * (double)pv[q] - cx * (double)xv[q]);
            y[i] = v;
            nn += (double)v * (double)v;
            cr += (double)v * (double)xv[q];
        }
    }
    for (int64_t i = i0 + kPre * st; i < i1; i += st) {
        const double xi = (double)x[i];
        const T v = (T)(cz * (double)y[i] - cp * (j > 1 ? (double)pr[i] : 0.0) - cx * xi);
        y[i] = v;
        nn += (double)v * (double)v;
        cr += (double)v * xi;
    }
    block_partial_last(nn, cr, red, &cnt, lz.nrm(j + 1) + blk, lz.crs(j + 1) + blk);
    if (advance && lead) lz.st->j_k1 = j + 1;
}

// End of a chunk: beta_j of the next step j from the update partials (one workgroup),
// the same reduction the next update phase performs, so the host sees those bits.
template <typename T>
__device__ inline void lz_finish_body(const LzArgs<T> &lz, double *sh)
{
    if (lz.st->status) return;
    const int j = lz.st->j_k1;
    if (threadIdx.x < 64) {
        const double t = wave_part_sum(lz.nrm(j), lz.nB);
        if (threadIdx.x == 0) {
            const double b = sqrt(t);
            lz.st->beta[j - 1] = b;
            if (b < kMinLanczosNorm) lz.st->status = (j == 1) ? 2 : 1;
        }
    }
}

// ------------------------------------------------------------- Fock-space hops

// Sign of c+/c on `bit` = (-1)^{# occupied orbitals above bit} of the resulting state
// (Fermion_sign, quantum_fermion_ket_class.f90:605-618).  Bits above `bit` do not
// change under the operator, so the input state can be used.
__device__ inline int sign_above(uint32_t s, int bit)
{
    return (__popc(s >> (bit + 1)) & 1) ? -1 : 1;
}

// hop(out, to, from, in) = c+_to c_from: destroy then create (ket_class.f90:483-492).
__device__ inline bool hop(uint32_t s, int to, int from, uint32_t *out, int *sg)
{
    if (!((s >> from) & 1u)) return false;
    uint32_t t = s & ~(1u << from);
    int g = sign_above(t, from);
    if ((t >> to) & 1u) return false;
    uint32_t o = t | (1u << to);
    g *= sign_above(o, to);
    *out = o;
    *sg = g;
    return true;
}

// The mask and cutoff of tab_HAIM2 for one coefficient: kept when nonzero and
// |t| > cutoff (the modulus for a complex t).
__device__ inline bool keep_coef(double re, double im, double cut)
{
    return !(re == 0.0 && im == 0.0) && hypot(re, im) > cut;
}


// ------------------------------------------------------------- S_z (Nambu) sectors

// Tables of an S_z sector: ELL [kw][d] of tab_HAIM2 over the Nambu orbitals, the
// quadratic diagonal and QUART_INT_SZ.
template <typename T>
struct SzArgs {
    int64_t d;
    int kw;
    const int32_t *rk;
    const T *val;
    const T *diag;
    const T *quart;
    int kh;                 // Hund transitions per row (0: none)
    const int32_t *hrk;     // [kh][d]
    const T *hval;
    const T *vim;           // complex sector: imaginary parts of val ([kw][d]); null when real
};

// HAIMsz_multr (HAIMsz.f90:453-560, USE_CC gather form): one thread per row,
//   w_i = QUART_i v_i + diag_i v_i + sum_k off_ik v_{rank_ik}   in this order,
// then (|Jhund| > 1e-5) the spin-flip and pair-hopping terms of HAIMsz_multc
// (HAIMsz.f90:300-406) from their ELL, for real vectors.
// MODE 1: Lanczos phase 1 on the Krylov slots with the partial <z, u_j> per workgroup.
template <typename T, int MODE>
__device__ inline void hv_sz_body(const SzArgs<T> &a, const T *__restrict__ xin, T *__restrict__ yout,
                                  const LzArgs<T> &lz, int blk, double *red)
{
    const T *__restrict__ x;
    T *__restrict__ y;
    int j = 0;
    if (MODE == 1) {
        if (!lz_hv_begin(lz, blk == 0 && threadIdx.x == 0, j)) return;
        x = lz.slot(j);
        y = lz.slot(j + 1);
    } else {
        x = xin;
        y = yout;
    }
    // grid-stride over the rows: the grid stays within the kPartA partials of a Lanczos step
    const int64_t d = a.d, stride = (int64_t)gridDim.x * blockDim.x;
    double dot = 0.0;
    for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < d; i += stride) {
        const T xv = x[i];
        T acc = a.quart[i] * xv;
        acc += a.diag[i] * xv;
        for (int k = 0; k < a.kw; ++k) acc += a.val[k * d + i] * x[a.rk[k * d + i]];
        for (int k = 0; k < a.kh; ++k) acc += a.hval[k * d + i] * x[a.hrk[k * d + i]];
        y[i] = acc;
        dot += (double)acc * (double)xv;
    }
    if (MODE == 1) block_partial(dot, red, lz.partA + blk);
}

// HAIMsz_multc (HAIMsz.f90:188-452, USE_CC gather form) of a complex S_z sector on planar
// vectors (Re[d], Im[d]), one thread per amplitude (grid-stride):
//   w_i = (QUART_i + diag_i) v_i + sum_k conj(h_ik) v_{rank_ik} + sum_Hund (-J) v_target
// with conj(h) v = (hr vr + hi vi) + i (hr vi - hi vr) (HAIMsz.f90:279-289; QUART, the
// diagonal and the Hund elements are real).  MODE 1: Lanczos phase 1 of the real
// embedding, partial Re <u_j, z> per workgroup.
template <int MODE>
__device__ inline void hv_sz_cplx_body(const SzArgs<double> &a, const double *__restrict__ xin,
                                       double *__restrict__ yout, const LzArgs<double> &lz, int blk, double *red)
{
    const double *__restrict__ x;
    double *__restrict__ y;
    int j = 0;
    if (MODE == 1) {
        if (!lz_hv_begin(lz, blk == 0 && threadIdx.x == 0, j)) return;
        x = lz.slot(j);
        y = lz.slot(j + 1);
    } else {
        x = xin;
        y = yout;
    }
    const int64_t d = a.d, stride = (int64_t)gridDim.x * blockDim.x;
    double dot = 0.0;
    for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < d; i += stride) {
        const double xr = x[i], xi = x[d + i];
        const double dg = a.quart[i] + a.diag[i];
        double yr = dg * xr, yi = dg * xi;
        for (int k = 0; k < a.kw; ++k) {
            const int64_t e = (int64_t)k * d + i, t = a.rk[e];
            const double hr = a.val[e], hi = a.vim[e], vr = x[t], vi = x[d + t];
            yr += hr * vr + hi * vi;
            yi += hr * vi - hi * vr;
        }
        for (int k = 0; k < a.kh; ++k) {
            const int64_t e = (int64_t)k * d + i, t = a.hrk[e];
            yr += a.hval[e] * x[t];
            yi += a.hval[e] * x[d + t];
        }
        y[i] = yr;
        y[d + i] = yi;
        dot += yr * xr + yi * xi;
    }
    if (MODE == 1) block_partial(dot, red, lz.partA + blk);
}

// ------------------------------------------------------------- complex sectors

// Tables of a complex (N_up, N_dn) sector: the ELL of both spins with the real and
// imaginary parts of <target|H|state>, the diagonals, the interaction per impurity pair
// and the Hund transitions (real).
struct CplxArgs {
    int dup, ddn, ncfg, kup, kdn, hund;
    int64_t D;                        // complex dimension; vectors are (Re[D], Im[D])
    const int32_t *rk_up, *rk_dn;     // ELL [k][d]
    const double *vr_up, *vi_up, *vr_dn, *vi_dn;
    const double *diag_up, *diag_dn, *uimp;
    const int32_t *imp_up, *imp_dn;
    const int32_t *hund_ptr, *hund_tr;
    const double *hund_val;
    const int32_t *swap_up, *swap_dn;
};

// HAIMupdo_multc (HAIMupdo.f90:150-656, USE_CC gather form) on planar complex vectors,
// one thread per amplitude (grid-stride beyond the grid), i = iup + ido d_up:
//   w_i = (U terms + diag_up + diag_dn) v_i + sum_up conj(h) v(jup, ido)
//         + sum_dn conj(h) v(iup, jdo) + sum_Hund J v(target)
// with conj(h) v = (hr vr + hi vi) + i (hr vi - hi vr).  MODE 1: Lanczos phase 1 on the
// Krylov slots with the partial Re <u_j, z> = sum zr ur + zi ui per workgroup (the real
// embedding: Lanczos on (Re, Im) with the symmetric [[A, -B], [B, A]] of H = A + iB has
// the complex Lanczos' alpha and beta, which are real).
template <int MODE>
__device__ inline void hv_cplx_body(const CplxArgs &a, const double *__restrict__ xin, double *__restrict__ yout,
                                    const LzArgs<double> &lz, int blk, double *red)
{
    const double *__restrict__ x;
    double *__restrict__ y;
    int j = 0;
    if (MODE == 1) {
        if (!lz_hv_begin(lz, blk == 0 && threadIdx.x == 0, j)) return;
        x = lz.slot(j);
        y = lz.slot(j + 1);
    } else {
        x = xin;
        y = yout;
    }
    const int64_t D = a.D, stride = (int64_t)gridDim.x * blockDim.x;
    const int dup = a.dup, ddn = a.ddn, ncfg = a.ncfg;
    double dot = 0.0;
    for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < D; i += stride) {
        const int iup = (int)(i % dup), ido = (int)(i / dup);
        const int64_t cb = (int64_t)ido * dup;
        const double xr = x[i], xi = x[D + i];
        const int cu = a.imp_up[iup], cd = a.imp_dn[ido];
        const double dg = a.uimp[cu * ncfg + cd] + (a.diag_up[iup] + a.diag_dn[ido]);
        double yr = dg * xr, yi = dg * xi;
        for (int k = 0; k < a.kup; ++k) {
            const int64_t e = (int64_t)k * dup + iup, t = a.rk_up[e] + cb;
            const double hr = a.vr_up[e], hi = a.vi_up[e], vr = x[t], vi = x[D + t];
            yr += hr * vr + hi * vi;
            yi += hr * vi - hi * vr;
        }
        for (int k = 0; k < a.kdn; ++k) {
            const int64_t e = (int64_t)k * ddn + ido, t = iup + (int64_t)a.rk_dn[e] * dup;
            const double hr = a.vr_dn[e], hi = a.vi_dn[e], vr = x[t], vi = x[D + t];
            yr += hr * vr + hi * vi;
            yi += hr * vi - hi * vr;
        }
        if (a.hund) {
            const int h0 = a.hund_ptr[cu * ncfg + cd], h1 = a.hund_ptr[cu * ncfg + cd + 1];
            for (int t = h0; t < h1; ++t) {
                const int tr = a.hund_tr[t];
                const int64_t g = a.swap_up[iup * ncfg + (tr & 0xFF)] + (int64_t)a.swap_dn[ido * ncfg + (tr >> 8)] * dup;
                yr += a.hund_val[t] * x[g];
                yi += a.hund_val[t] * x[D + g];
            }
        }
        y[i] = yr;
        y[D + i] = yi;
        dot += yr * xr + yi * xi;
    }
    if (MODE == 1) block_partial(dot, red, lz.partA + blk);
}

// HAIMsz_mult_fly (HAIMsz.f90:615-1027): the S_z H.v with each row's tab_HAIM2 entries
// generated on the fly from the Nambu matrices (k_tab's loops: impurity hops, bath hops,
// the hybridisation both ways, each kept when nonzero and above the cutoff) instead of
// read from the ELL.  Per row it reads the state word and a rank lookup per target
// instead of kw (rank, value) pairs.  The entries come in the table's order with the
// table's values (t sign, rounded to T), so the result has the table path's bits.
template <typename T, int MODE>
__device__ inline void hv_sz_fly_body(const SzArgs<T> &a, const SpinModel &m, const int32_t *__restrict__ states,
                                      const int32_t *__restrict__ rank, const T *__restrict__ xin,
                                      T *__restrict__ yout, const LzArgs<T> &lz, int blk, double *red)
{
    const T *__restrict__ x;
    T *__restrict__ y;
    int j = 0;
    if (MODE == 1) {
        if (!lz_hv_begin(lz, blk == 0 && threadIdx.x == 0, j)) return;
        x = lz.slot(j);
        y = lz.slot(j + 1);
    } else {
        x = xin;
        y = yout;
    }
    const int64_t d = a.d, stride = (int64_t)gridDim.x * blockDim.x;
    const int ni = m.nimp, nbt = m.nbath;
    double dot = 0.0;
    for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < d; i += stride) {
        const uint32_t s = (uint32_t)states[i];
        const T xv = x[i];
        T acc = a.quart[i] * xv;
        acc += a.diag[i] * xv;
        uint32_t o;
        int sg;
        for (int p = 0; p < ni; ++p)
            for (int q = 0; q < ni; ++q) {
                const double t = m.ec[p * ni + q];
                if (p == q || t == 0.0 || !(fabs(t) > m.cut)) continue;
                if (hop(s, m.ibit[p], m.ibit[q], &o, &sg)) acc += (T)(t * sg) * x[rank[o]];
            }
        for (int p = 0; p < nbt; ++p)
            for (int q = 0; q < nbt; ++q) {
                const double t = m.eb[p * nbt + q];
                if (p == q || t == 0.0 || !(fabs(t) > m.cut)) continue;
                if (hop(s, m.bbit[p], m.bbit[q], &o, &sg)) acc += (T)(t * sg) * x[rank[o]];
            }
        for (int b = 0; b < nbt; ++b)
            for (int c = 0; c < ni; ++c) {
                const double t = m.vbc[b * ni + c];
                if (t == 0.0 || !(fabs(t) > m.cut)) continue;
                if (hop(s, m.bbit[b], m.ibit[c], &o, &sg)) acc += (T)(t * sg) * x[rank[o]];
                if (hop(s, m.ibit[c], m.bbit[b], &o, &sg)) acc += (T)(t * sg) * x[rank[o]];
            }
        for (int k = 0; k < a.kh; ++k) acc += a.hval[k * d + i] * x[a.hrk[k * d + i]];
        y[i] = acc;
        dot += (double)acc * (double)xv;
    }
    if (MODE == 1) block_partial(dot, red, lz.partA + blk);
}

// HAIMsz_mult_fly_c (HAIMsz.f90:833-1027): the complex S_z H.v with each row's entries
// generated on the fly from the complex Nambu matrices in the table's order (k_tab's
// loops and keep rule; the hybridisation's reverse hop carries the conjugate), then the
// conj(h) gather of hv_sz_cplx_body on planar vectors.  Same values as the table, so the
// same bits.
template <int MODE>
__device__ inline void hv_sz_cplx_fly_body(const SzArgs<double> &a, const SpinModel &m,
                                           const int32_t *__restrict__ states, const int32_t *__restrict__ rank,
                                           const double *__restrict__ xin, double *__restrict__ yout,
                                           const LzArgs<double> &lz, int blk, double *red)
{
    const double *__restrict__ x;
    double *__restrict__ y;
    int j = 0;
    if (MODE == 1) {
        if (!lz_hv_begin(lz, blk == 0 && threadIdx.x == 0, j)) return;
        x = lz.slot(j);
        y = lz.slot(j + 1);
    } else {
        x = xin;
        y = yout;
    }
    const int64_t d = a.d, stride = (int64_t)gridDim.x * blockDim.x;
    const int ni = m.nimp, nbt = m.nbath;
    double dot = 0.0;
    for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < d; i += stride) {
        const uint32_t s = (uint32_t)states[i];
        const double xr = x[i], xi = x[d + i];
        const double dg = a.quart[i] + a.diag[i];
        double yr = dg * xr, yi = dg * xi;
        uint32_t o;
        int sg;
        auto add = [&](double tr, double ti, int to, int frm) {
            if (!hop(s, to, frm, &o, &sg)) return;
            const int64_t t = rank[o];
            const double hr = tr * sg, hi = ti * sg, vr = x[t], vi = x[d + t];
            yr += hr * vr + hi * vi;
            yi += hr * vi - hi * vr;
        };
        for (int p = 0; p < ni; ++p)
            for (int q = 0; q < ni; ++q) {
                const double t = m.ec[p * ni + q], ti = m.ec_im[p * ni + q];
                if (p != q && keep_coef(t, ti, m.cut)) add(t, ti, m.ibit[p], m.ibit[q]);
            }
        for (int p = 0; p < nbt; ++p)
            for (int q = 0; q < nbt; ++q) {
                const double t = m.eb[p * nbt + q], ti = m.eb_im[p * nbt + q];
                if (p != q && keep_coef(t, ti, m.cut)) add(t, ti, m.bbit[p], m.bbit[q]);
            }
        for (int b = 0; b < nbt; ++b)
            for (int c = 0; c < ni; ++c) {
                const double t = m.vbc[b * ni + c], ti = m.vbc_im[b * ni + c];
                if (!keep_coef(t, ti, m.cut)) continue;
                add(t, ti, m.bbit[b], m.ibit[c]);
                add(t, -ti, m.ibit[c], m.bbit[b]);
            }
        for (int k = 0; k < a.kh; ++k) {
            const int64_t e = (int64_t)k * d + i, t = a.hrk[e];
            yr += a.hval[e] * x[t];
            yi += a.hval[e] * x[d + t];
        }
        y[i] = yr;
        y[d + i] = yi;
        dot += yr * xr + yi * xi;
    }
    if (MODE == 1) block_partial(dot, red, lz.partA + blk);
}

// ------------------------------------------------ persistent Lanczos (cooperative launch)

// Grid-wide barrier of a cooperative launch (every workgroup resident at once).  The
// data the workgroups exchange inside the kernel (Krylov vector, partial sums) is written
// and read with device-scope coherent accesses (ld_agent / st_agent), so the barrier
// needs no cache write-back or invalidation: every wave waits for its own accesses
// (s_waitcnt 0) and the workgroup meets.  Then a two-level arrival on counters that only
// grow during the launch (zeroed before it): thread 0 adds one to its group's counter
// (kBarGroupSize workgroups per group); the group's last arrival adds one to
// st->bar_count, and the last of those stores the barrier number `b` (1, 2, ... in the
// launch) in every group's release word; every other workgroup polls its group's word.
// So no word has more than kBarGroupSize writers or readers, and a release reaches the
// pollers one store after the last arrival.  A poll that exceeds `spin_limit` rounds (or sees
// another workgroup's timeout) sets st->bar_err and status 3 and leaves, so the kernel
// always ends; returns 1 then.  With `stop_word`, thread 0 reads it after the release and
// the barrier returns 2 when it holds a step <= `jdone` (the word is written before the
// writer's arrival, so every workgroup reads the same value), else 0.
__device__ inline bool bar_wait(unsigned *word, unsigned b, LzState *st, int spin_limit)
{
    for (int spins = 0;; ++spins) {
        if (__hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= b) return false;
        if (spins >= spin_limit ||
            ((spins & 63) == 63 && __hip_atomic_load(&st->bar_err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
            return true;
    }
}

__device__ inline int grid_barrier(LzState *st, unsigned b, int spin_limit, int *flag,
                                   const int *stop_word = nullptr, int jdone = 0)
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned nblk = gridDim.x, g = blockIdx.x / kBarGroupSize;
        const unsigned ng = (nblk + kBarGroupSize - 1) / kBarGroupSize;
        const unsigned sz = min((unsigned)kBarGroupSize, nblk - g * kBarGroupSize);
        bool err = false;
        const unsigned c = __hip_atomic_fetch_add(&st->bar_grp[16 * g], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bool released = false;
        if (c == sz * b - 1) {   // last of the group: arrive at the top
            const unsigned t = __hip_atomic_fetch_add(&st->bar_count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (t == ng * b - 1) {   // last of all: release every group directly
                for (unsigned h = 0; h < ng; ++h)
                    __hip_atomic_store(&st->bar_rel[16 * h], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                released = true;
            }
        }
        if (!released) err = bar_wait(&st->bar_rel[16 * g], b, st, spin_limit);
        if (err) {
            __hip_atomic_store(&st->bar_err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            st->status = 3;
        }
        int stop = 0;
        if (!err && stop_word) {
            const int sa = ld_agent(stop_word);
            stop = sa != 0 && sa <= jdone;
        }
        *flag = err ? 1 : stop ? 2 : 0;
    }
    __syncthreads();
    return *flag;
}

// grid_barrier in two halves, for work that needs no other workgroup between arriving and
// leaving: grid_arrive (every wave waits for its memory accesses, the workgroup meets,
// thread 0 arrives; true when it was the last arrival and has released barrier b), then
// grid_wait (thread 0 waits for the release unless it made it; the rest as grid_barrier).
__device__ inline bool grid_arrive(LzState *st, unsigned b)
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    bool released = false;
    if (threadIdx.x == 0) {
        const unsigned nblk = gridDim.x, g = blockIdx.x / kBarGroupSize;
        const unsigned ng = (nblk + kBarGroupSize - 1) / kBarGroupSize;
        const unsigned sz = min((unsigned)kBarGroupSize, nblk - g * kBarGroupSize);
        const unsigned c = __hip_atomic_fetch_add(&st->bar_grp[16 * g], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (c == sz * b - 1) {
            const unsigned t = __hip_atomic_fetch_add(&st->bar_count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (t == ng * b - 1) {
                for (unsigned h = 0; h < ng; ++h)
                    __hip_atomic_store(&st->bar_rel[16 * h], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                released = true;
            }
        }
    }
    return released;
}

__device__ inline int grid_wait(LzState *st, unsigned b, bool released, int spin_limit, int *flag,
                                const int *stop_word = nullptr, int jdone = 0)
{
    if (threadIdx.x == 0) {
        const unsigned g = blockIdx.x / kBarGroupSize;
        const bool err = released ? false : bar_wait(&st->bar_rel[16 * g], b, st, spin_limit);
        if (err) {
            __hip_atomic_store(&st->bar_err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            st->status = 3;
        }
        int stop = 0;
        if (!err && stop_word) {
            const int sa = ld_agent(stop_word);
            stop = sa != 0 && sa <= jdone;
        }
        *flag = err ? 1 : stop ? 2 : 0;
    }
    __syncthreads();
    return *flag;
}

// Bytes of dynamic LDS of the persistent Lanczos kernel: the fast H.v's layout, then
// u_{j-1} and the diagonal of the workgroup's rows (T[dup] each, 16-byte aligned).
__host__ __device__ inline size_t lz_persist_lds(int dup, int ncoef, int ncfg, int nhund, size_t tsize)
{
    return (hv_fast_lds(dup, ncoef, ncfg, nhund, tsize) + 15) / 16 * 16 + 2 * tsize * (size_t)dup;
}

// `nsteps` Lanczos steps in one kernel: the fast H.v (one column per workgroup, 256
// threads, rows r = tid + i*256, i < R) and the update of the same column, separated by
// grid barriers.  The static data of a workgroup (its column's down-hop entries, the
// packed up-hop entries of its rows in registers, the diagonal and the LDS tables) are
// loaded once per launch; u_j and u_{j-1} of the workgroup's column stay in LDS and
// z = H u_j in registers across the steps, so a step reads from memory only the
// neighbour columns (and Hund targets) of u_j and the partial sums.  Per step:
//   H.v:     z = H u_j, partial <z, u_j> -> partA
//   barrier
//   update:  N, P, Q as in lz_scalars; beta_j, alpha_j; u_{j+1} = z/beta_j -
//            (beta_j/beta_{j-1}) u_{j-1} - (alpha_j/beta_j) u_j -> slot(j+1) and LDS;
//            partials of ||u_{j+1}||^2, <u_{j+1}, u_j>
//   barrier
// Every sum is formed in the order of the two-kernel step with the same thread layout
// (the fast kernel at 256 threads and the update kernel), so alpha and beta carry the
// same bits.  Breakdown (beta_j below kMinLanczosNorm) is seen by every workgroup alike
// and ends the loop everywhere.  The caller launches k_lz_finish afterwards.
// With a monitor (MON), thread 0 of workgroup 0 publishes beta and alpha of each step
// (kept in LDS, pub[2]) to the host at the start of the next one (the stores complete
// during its H.v), loads the host's stop flag at the start of every H.v phase and, once it is set,
// writes the current step to st->stop_at in the update phase; every workgroup reads that
// word after the second barrier of the step and leaves the loop together.  The word is
// written after the first barrier of a step and read before the first barrier of the
// next, so no workgroup can see a value written in a later step.
template <typename T, int K, int R, bool MON>
__device__ inline void lz_persist_body(const HvArgs<T> &a, const LzArgs<T> &lz, int nsteps, int spin_limit,
                                       LzMonitor mon, unsigned char *smem, double *red, double *sh, double *pub,
                                       int *flag)
{
    constexpr int NT = kBlock;
    LzState *st = lz.st;
    if (st->status) return;   // written by an earlier kernel: the same for every workgroup
    const int dup = a.dup, ncfg = a.ncfg, nq = ncfg * ncfg, nh = a.nhund;
    T *col = reinterpret_cast<T *>(smem);
    T *coef = col + dup;
    T *uimp = coef + a.ncoef_up;
    T *hval = uimp + nq;
    int *hptr = reinterpret_cast<int *>(hval + nh);
    int *htr = hptr + nq + 1;
    int *swdn = htr + nh;
    T *colm = reinterpret_cast<T *>(smem + (hv_fast_lds(dup, a.ncoef_up, ncfg, nh, sizeof(T)) + 15) / 16 * 16);
    T *dgl = colm + dup;
    const int blk = blockIdx.x, tid = threadIdx.x;
    const bool lead = blk == 0 && tid == 0;
    unsigned nbar = 0;   // barriers passed in this launch
    int j = st->j_k1;
    const int ido = lz.colord ? lz.colord[blk] : blk;
    const bool act = ido >= 0;
    // ---- static data of the column and of the rows
    int cd = 0, nd = 0;
    T ddg = (T)0;
    T od[K];
    int jd[K];
    if (act) {
        cd = a.imp_dn[ido];
        const int p0 = a.ptr_dn[ido];
        nd = a.ptr_dn[ido + 1] - p0;
        ddg = a.diag_dn[ido];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        od[k] = (act && k < nd) ? a.off_dn[a.ptr_dn[ido] + k] : (T)0;
        jd[k] = (act && k < nd) ? a.rk_dn[a.ptr_dn[ido] + k] * dup : 0;
    }
    const int base = act ? ido * dup : 0;
    // Wide tables with many rows per thread (K R > 40) do not keep the packed up-hop
    // entries in registers across the steps: each step reloads them (read-only, L2
    // resident) with the neighbour loads of its row batch; the column's down-hop entries
    // sit in LDS instead of scalar registers.
    constexpr bool PKG = K * R > 40;
    __shared__ T s_od[PKG ? K : 1];
    __shared__ int s_jd[PKG ? K : 1];
    if constexpr (PKG) {
        if (tid < K) {
            s_od[tid] = (act && tid < nd) ? a.off_dn[a.ptr_dn[ido] + tid] : (T)0;
            s_jd[tid] = (act && tid < nd) ? a.rk_dn[a.ptr_dn[ido] + tid] * dup : 0;
        }
    }
    constexpr int RP = PKG ? 1 : R;
    uint32_t pk[RP][K];
    int cu[R];
    {
        const T *x0 = lz.slot(j);
        const T *xm = lz.slot(j > 1 ? j - 1 : j);
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int r = tid + i * NT;
            const bool ok = act && r < dup;
            cu[i] = ok ? a.imp_up[r] : 0;
            if constexpr (!PKG) {
#pragma unroll
                for (int k = 0; k < K; ++k) pk[i][k] = ok ? a.ell_pk_up[k * dup + r] : 0u;
            }
            if (ok) {
                dgl[r] = (a.diag_up[r] + ddg) + a.uimp[cu[i] * ncfg + cd];
                col[r] = x0[base + r];
                colm[r] = j > 1 ? xm[base + r] : (T)0;
            }
        }
    }
    for (int e = tid; e < a.ncoef_up; e += NT) coef[e] = a.coef_up[e];
    for (int e = tid; e < nq; e += NT) uimp[e] = a.uimp[e];
    if (a.hund && act) {
        for (int e = tid; e <= nq; e += NT) hptr[e] = a.hund_ptr[e];
        for (int e = tid; e < nh; e += NT) {
            htr[e] = a.hund_tr[e];
            hval[e] = a.hund_val[e];
        }
        for (int e = tid; e < ncfg; e += NT) swdn[e] = a.swap_dn[ido * ncfg + e];
    }
    double bprev = j > 1 ? st->beta[j - 2] : 1.0;
    int hflag = 0;   // lead: the host's stop flag as read in this step
    __syncthreads();

    for (int s = 0; s < nsteps; ++s) {
        if (MON && lead) {
            if (s > 0) {
                __hip_atomic_store(mon.ab + 2 * (j - 2), pub[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(mon.ab + 2 * (j - 2) + 1, pub[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            hflag = __hip_atomic_load(mon.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        // ---- H.v phase of step j (u_j of the own column is in col)
        const T *__restrict__ x = lz.slot(j);
        T *__restrict__ y = lz.slot(j + 1);
        T z[R];
        double dot = 0.0;
        {
            // neighbour columns through uniform per-hop pointers: one 32-bit row offset
            // per row serves every hop.  Rows go in batches of RB (the loads of a batch
            // in flight together) to stay within the register budget of a co-resident grid.
            const T *xk[PKG ? 1 : K];
            if constexpr (!PKG) {
#pragma unroll
                for (int k = 0; k < K; ++k) xk[k] = x + jd[k];
            }
            constexpr int RB = (R < 2 || PKG) ? 1 : 2;
#pragma unroll
            for (int i0 = 0; i0 < R; i0 += RB) {
                T xd[RB][K];
                uint32_t pkb[PKG ? RB : 1][K];
#pragma unroll
                for (int b = 0; b < RB; ++b) {
                    const unsigned r = (unsigned)(tid + (i0 + b) * NT);
                    const bool ok = act && (int)r < dup;
#pragma unroll
                    for (int k = 0; k < K; ++k)
                        xd[b][k] = (ok && k < nd) ? ld_agent((PKG ? x + s_jd[k] : xk[PKG ? 0 : k]) + r) : (T)0;
                    if constexpr (PKG) {
#pragma unroll
                        for (int k = 0; k < K; ++k) pkb[b][k] = ok ? a.ell_pk_up[k * dup + r] : 0u;
                    }
                }
#pragma unroll
                for (int b = 0; b < RB; ++b) {
                    const int i = i0 + b, r = tid + i * NT;
                    const uint32_t *pr = PKG ? pkb[PKG ? b : 0] : pk[PKG ? 0 : i];
                    z[i] = (T)0;
                    if (act && r < dup) {
                        const T xv = col[r];
                        T acc = dgl[r] * xv;
#pragma unroll
                        for (int k = 0; k < K; ++k) acc += coef[pr[k] >> 16] * col[pr[k] & 0xFFFFu];
#pragma unroll
                        for (int k = 0; k < K; ++k) acc += (PKG ? s_od[k] : od[k]) * xd[b][k];
                        if (a.hund) {
                            const int h0 = hptr[cu[i] * ncfg + cd], h1 = hptr[cu[i] * ncfg + cd + 1];
                            for (int t = h0; t < h1; ++t) {
                                const int tr = htr[t];
                                acc += hval[t] * ld_agent(x + a.swap_up[r * ncfg + (tr & 0xFF)] + swdn[tr >> 8] * dup);
                            }
                        }
                        z[i] = acc;
                        dot += (double)acc * (double)xv;
                    }
                }
                // keep the scheduler from hoisting the next batch's loads into this one
                // (it would hold every batch's loads at once and spill)
                if constexpr (PKG) __builtin_amdgcn_sched_barrier(0);
            }
        }
        block_partial2<true>(dot, 0.0, red, lz.partA + blk, nullptr);
        if (grid_barrier(st, ++nbar, spin_limit, flag)) return;

        // ---- update phase of step j
        lz_scalars<T, true>(lz, j, sh);
        __syncthreads();
        const double bj = sqrt(sh[0]);
        if (bj < kMinLanczosNorm) {   // the reference STOPs here (Lanczos_plain.h:70-92)
            if (lead) {
                st->beta[j - 1] = bj;   // the test of iteration j - 1 reads it
                st->status = (j == 1) ? 2 : 1;
            }
            return;
        }
        const double alpha = sh[2] / (bj * bj) - (j > 1 ? sh[1] / bprev : 0.0);
        if (lead) {
            st->beta[j - 1] = bj;
            st->alpha[j - 1] = alpha;
            if (MON) {
                pub[0] = bj;
                pub[1] = alpha;
                if (hflag) st_agent(&st->stop_at, j);
            }
        }
        const double cz = 1.0 / bj, cp = j > 1 ? bj / bprev : 0.0, cx = alpha / bj;
        double nn = 0.0, cr = 0.0;
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int r = tid + i * NT;
            if (act && r < dup) {
                const T xv = col[r];
                const T v = (T)(cz * (double)z[i] - cp * (double)colm[r] - cx * (double)xv);
                st_agent(y + base + r, v);
                nn += (double)v * (double)v;
                cr += (double)v * (double)xv;
                colm[r] = xv;
                col[r] = v;
            }
        }
        block_partial2<true>(nn, cr, red, lz.nrm(j + 1) + blk, lz.crs(j + 1) + blk);
        bprev = bj;
        ++j;
        if (lead) {
            st->j_k1 = j;
            st->j_k2 = j - 1;
        }
        if (grid_barrier(st, ++nbar, spin_limit, flag, MON ? &st->stop_at : nullptr, j - 1)) return;
    }
}


// ------------------------------------ persistent Lanczos, several columns per workgroup

// Where the columns of the multi-column persistent kernel live and where their partial
// sums go: workgroup b owns the columns cols[b * cpw + c] (-1 = idle), column ido writes
// its partial sums at index pslot[ido] of partA / partB / partC, the index the two-kernel
// step's workgroup for that column uses, so both sum the same partials in the same order;
// the nidle indices idle[] of that layout hold no column and are zeroed at launch.
struct PersistMap {
    const int32_t *cols, *pslot, *idle;
    int nidle, cpw;
};

// Hund transitions a row of this kernel gathers in the same batch as its down-hop
// neighbours; sectors with more per row keep the one-column kernel.
constexpr int kPersistHund = 2;

// Bytes of dynamic LDS of the multi-column persistent kernel (offsets 16-byte aligned):
//   shared by the workgroup's columns: pk[K][dup] (u32), du[dup] (T), cu[dup] (i32),
//   swap[dup][ncfg] (u16, Hund only), coef[ncoef], uimp[ncfg^2], hval[nhund] (T),
//   hptr[ncfg^2 + 1], htr[nhund] (i32);
//   per column: nvec vectors of dup (T: col, colm; the one-barrier kernel has a third),
//   swdn[ncfg] (i32), then (read by tables of width >= 12, whose down-hop entries stay
//   out of scalar registers, and by the one-barrier kernel) cod[K] (T), cjd[K] (i32).
struct PersistLds {
    size_t pk, du, cu, sw, coef, uimp, hval, hptr, htr, cols, col_stride, total;
};
__host__ __device__ inline size_t al16(size_t x) { return (x + 15) / 16 * 16; }
__host__ __device__ inline PersistLds persist_cols_lds(int dup, int K, int ncoef, int ncfg, int nhund, int hund,
                                                       int cpw, size_t tsize, int nvec = 2)
{
    PersistLds L;
    const size_t nq = (size_t)ncfg * ncfg;
    L.pk = 0;
    L.du = al16(L.pk + 4 * (size_t)K * dup);
    L.cu = al16(L.du + tsize * dup);
    L.sw = al16(L.cu + 4 * (size_t)dup);
    L.coef = al16(L.sw + (hund ? 2 * (size_t)dup * ncfg : 0));
    L.uimp = al16(L.coef + tsize * ncoef);
    L.hval = al16(L.uimp + tsize * nq);
    L.hptr = al16(L.hval + tsize * nhund);
    L.htr = al16(L.hptr + 4 * (nq + 1));
    L.cols = al16(L.htr + 4 * (size_t)nhund);
    L.col_stride = al16(al16(nvec * tsize * dup + 4 * (size_t)ncfg) + (tsize + 4) * (size_t)K);
    L.total = L.cols + L.col_stride * cpw;
    return L;
}

// `nsteps` Lanczos steps in one cooperative launch, as lz_persist_body, with map.cpw
// columns per workgroup of 256 * cpw threads: threads 256 c .. 256 c + 255 own column c of
// the workgroup, their rows r = t + 256 i (i < R) exactly as one workgroup of the two-kernel
// step's fast H.v owns its column, so every partial sum is formed in the same order.  What
// the columns of a workgroup share sits in LDS once: the rows' packed up-hop entries, their
// diagonal and impurity bits, their Hund swap table.  The Hund targets of a row are found in
// LDS, so its down-hop neighbours and its Hund targets (at most kPersistHund) are gathered
// in one batch of device-coherent loads, RB rows at a time.  The grid has a quarter as many
// workgroups as the one-column kernel at cpw = 4, so each grid barrier has fewer arrivals
// and each workgroup's reading of the partial sums is shared by four columns.
// The second grid barrier of a step (after u_{j+1} is stored) is split: a workgroup
// arrives, computes the part of the next H.v that needs only its own columns (diagonal and
// up hops, from LDS, the first terms of every row's sum), then waits for the release
// before it gathers its neighbours.

template <typename T, int K, int R, bool MON>
__device__ inline void lz_persist_cols_body(const HvArgs<T> &a, const LzArgs<T> &lz, const PersistMap &pm,
                                            int nsteps, int spin_limit, LzMonitor mon, unsigned char *smem,
                                            double *red, double *sh, double *pub, int *flag)
{
    constexpr int NT = kBlock;                           // threads per column
    // rows per batch of loads: all of them, or 2 (1 for the widest FP64 rows) where more
    // would spill registers or measured slower
    constexpr int LB = K * R * (int)sizeof(T);
    constexpr int RB = (K * R <= 16) ? R : (LB >= 384 ? 1 : 2);
    LzState *st = lz.st;
    if (st->status) return;   // written by an earlier kernel: the same for every workgroup
    const int dup = a.dup, ncfg = a.ncfg, nq = ncfg * ncfg, nh = a.nhund, cpw = pm.cpw;
    const PersistLds L = persist_cols_lds(dup, K, a.ncoef_up, ncfg, nh, a.hund, cpw, sizeof(T));
    uint32_t *pk = reinterpret_cast<uint32_t *>(smem + L.pk);
    T *du = reinterpret_cast<T *>(smem + L.du);
    int *cup = reinterpret_cast<int *>(smem + L.cu);
    uint16_t *swp = reinterpret_cast<uint16_t *>(smem + L.sw);
    T *coef = reinterpret_cast<T *>(smem + L.coef);
    T *uimp = reinterpret_cast<T *>(smem + L.uimp);
    T *hval = reinterpret_cast<T *>(smem + L.hval);
    int *hptr = reinterpret_cast<int *>(smem + L.hptr);
    int *htr = reinterpret_cast<int *>(smem + L.htr);
    const int tid = threadIdx.x, nt = blockDim.x;
    // the column group of this wave (uniform per wave)
    const int cg = __builtin_amdgcn_readfirstlane(tid >> 8), lt = tid & (NT - 1);
    T *col = reinterpret_cast<T *>(smem + L.cols + L.col_stride * cg);
    T *colm = col + dup;
    int *swdn = reinterpret_cast<int *>(colm + dup);
    T *cod = reinterpret_cast<T *>(reinterpret_cast<unsigned char *>(col) + al16(2 * sizeof(T) * dup + 4 * ncfg));
    int *cjd = reinterpret_cast<int *>(cod + K);
    constexpr bool DL = K >= 12;   // down-hop entries of the column read from LDS
    const bool lead = blockIdx.x == 0 && tid == 0;
    unsigned nbar = 0;
    int j = st->j_k1;
    const int ido = __builtin_amdgcn_readfirstlane(pm.cols[blockIdx.x * cpw + cg]);
    const bool act = ido >= 0;
    const int ps = act ? pm.pslot[ido] : 0;
    int cd = 0, nd = 0;
    T ddg = (T)0;
    T od[K];
    int jd[K];
    if (act) {
        cd = a.imp_dn[ido];
        nd = a.ptr_dn[ido + 1] - a.ptr_dn[ido];
        ddg = a.diag_dn[ido];
    }
    if constexpr (DL) {
        if (lt < K) {
            cod[lt] = (act && lt < nd) ? a.off_dn[a.ptr_dn[ido] + lt] : (T)0;
            cjd[lt] = (act && lt < nd) ? a.rk_dn[a.ptr_dn[ido] + lt] * dup : 0;
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            od[k] = (act && k < nd) ? a.off_dn[a.ptr_dn[ido] + k] : (T)0;
            jd[k] = (act && k < nd) ? a.rk_dn[a.ptr_dn[ido] + k] * dup : 0;
        }
    }
    const int base = act ? ido * dup : 0;
    // ---- tables shared by the workgroup's columns, and this column's vectors
    for (int e = tid; e < K * dup; e += nt) pk[e] = a.ell_pk_up[e];
    for (int e = tid; e < dup; e += nt) {
        du[e] = a.diag_up[e];
        cup[e] = a.imp_up[e];
    }
    for (int e = tid; e < a.ncoef_up; e += nt) coef[e] = a.coef_up[e];
    for (int e = tid; e < nq; e += nt) uimp[e] = a.uimp[e];
    if (a.hund) {
        for (int e = tid; e < dup * ncfg; e += nt) swp[e] = (uint16_t)a.swap_up[e];
        for (int e = tid; e <= nq; e += nt) hptr[e] = a.hund_ptr[e];
        for (int e = tid; e < nh; e += nt) {
            htr[e] = a.hund_tr[e];
            hval[e] = a.hund_val[e];
        }
        if (act)
            for (int e = lt; e < ncfg; e += NT) swdn[e] = a.swap_dn[ido * ncfg + e];
    }
    {
        const T *x0 = lz.slot(j);
        const T *xm = lz.slot(j > 1 ? j - 1 : j);
        for (int r = lt; r < dup; r += NT) {
            col[r] = act ? x0[base + r] : (T)0;
            colm[r] = (act && j > 1) ? xm[base + r] : (T)0;
        }
    }
    // partial-sum slots no column writes hold zeros (read by every sum of a step): the H.v
    // partials and both parities of the update partials
    if (blockIdx.x == 0)
        for (int e = tid; e < pm.nidle; e += nt) {
            const int b = pm.idle[e];
            st_agent(lz.partA + b, 0.0);
            st_agent(lz.nrm(0) + b, 0.0);
            st_agent(lz.nrm(1) + b, 0.0);
            st_agent(lz.crs(0) + b, 0.0);
            st_agent(lz.crs(1) + b, 0.0);
        }
    double bprev = j > 1 ? st->beta[j - 2] : 1.0;
    int hflag = 0;
    T zl[R];   // diagonal and up-hop part of the next H.v of each row
    __syncthreads();
    for (int s = 0; s < nsteps; ++s) {
        if (MON && lead) {
            if (s > 0) {
                __hip_atomic_store(mon.ab + 2 * (j - 2), pub[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(mon.ab + 2 * (j - 2) + 1, pub[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            hflag = __hip_atomic_load(mon.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        // ---- H.v phase of step j (u_j of the own column is in col)
        const T *__restrict__ x = lz.slot(j);
        T *__restrict__ y = lz.slot(j + 1);
        double *partA = lz.partA;
        T z[R];
        double dot = 0.0;
        if (s == 0) {   // the local part of the first step (later steps form it during the barrier)
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int r = lt + i * NT;
                zl[i] = (T)0;
                if (act && r < dup) {
                    const T xv = col[r];
                    T acc = ((du[r] + ddg) + uimp[cup[r] * ncfg + cd]) * xv;
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const uint32_t w = pk[k * dup + r];
                        acc += coef[w >> 16] * col[w & 0xFFFFu];
                    }
                    zl[i] = acc;
                }
            }
        }
#pragma unroll
        for (int i0 = 0; i0 < R; i0 += RB) {
            T xd[RB][K], xh[RB][kPersistHund];
            int h0[RB], nht[RB];
#pragma unroll
            for (int b = 0; b < RB; ++b) {
                const int r = lt + (i0 + b) * NT;
                const bool ok = act && r < dup;
#pragma unroll
                for (int k = 0; k < K; ++k) xd[b][k] = (ok && k < nd) ? ld_agent(x + (DL ? cjd[k] : jd[k]) + r) : (T)0;
                h0[b] = 0;
                nht[b] = 0;
                if (a.hund && ok) {
                    const int q = cup[r] * ncfg + cd;
                    h0[b] = hptr[q];
                    nht[b] = hptr[q + 1] - h0[b];
                }
#pragma unroll
                for (int h = 0; h < kPersistHund; ++h) {
                    xh[b][h] = (T)0;
                    if (h < nht[b]) {
                        const int tr = htr[h0[b] + h];
                        xh[b][h] = ld_agent(x + (int)swp[r * ncfg + (tr & 0xFF)] + swdn[tr >> 8] * dup);
                    }
                }
            }
#pragma unroll
            for (int b = 0; b < RB; ++b) {
                const int i = i0 + b, r = lt + i * NT;
                z[i] = (T)0;
                if (act && r < dup) {
                    const T xv = col[r];
                    T acc = zl[i];
#pragma unroll
                    for (int k = 0; k < K; ++k) acc += (DL ? cod[k] : od[k]) * xd[b][k];
#pragma unroll
                    for (int h = 0; h < kPersistHund; ++h)
                        if (h < nht[b]) acc += hval[h0[b] + h] * xh[b][h];
                    z[i] = acc;
                    dot += (double)acc * (double)xv;
                }
            }
            if constexpr (RB < R) __builtin_amdgcn_sched_barrier(0);
        }
        // column partial <z, u_j>: the four wave sums of the column in wave order
        {
            const double v = wave_sum(dot);
            if ((tid & 63) == 0) red[2 * (tid >> 6)] = v;
            __syncthreads();
            if (lt == 0 && act) {
                double t = 0.0;
                for (int w = 0; w < NT / 64; ++w) t += red[2 * (cg * (NT / 64) + w)];
                st_agent(partA + ps, t);
            }
        }
        if (grid_barrier(st, ++nbar, spin_limit, flag)) return;

        // ---- update phase of step j: N, P, Q as lz_scalars
        {
            const int w = tid >> 6;
            if (w == 0) {
                const double t = wave_part_sum<true>(lz.nrm(j), lz.nB);
                if (tid == 0) sh[0] = t;
            } else if (w == 1) {
                const double t = j > 1 ? wave_part_sum<true>(lz.crs(j), lz.nB) : 0.0;
                if (tid == 64) sh[1] = t;
            } else if (w == 2) {
                const double t = wave_part_sum<true>(partA, lz.nA);
                if (tid == 128) sh[2] = t;
            }
        }
        __syncthreads();
        const double bj = sqrt(sh[0]);
        if (bj < kMinLanczosNorm) {   // the reference STOPs here (Lanczos_plain.h:70-92)
            if (lead) {
                st->beta[j - 1] = bj;
                st->status = (j == 1) ? 2 : 1;
            }
            return;
        }
        const double alpha = sh[2] / (bj * bj) - (j > 1 ? sh[1] / bprev : 0.0);
        if (lead) {
            st->beta[j - 1] = bj;
            st->alpha[j - 1] = alpha;
            if (MON) {
                pub[0] = bj;
                pub[1] = alpha;
                if (hflag) st_agent(&st->stop_at, j);
            }
        }
        const double cz = 1.0 / bj, cp = j > 1 ? bj / bprev : 0.0, cx = alpha / bj;
        double nn = 0.0, cr = 0.0;
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int r = lt + i * NT;
            if (act && r < dup) {
                const T xv = col[r];
                const T v = (T)(cz * (double)z[i] - cp * (double)colm[r] - cx * (double)xv);
                st_agent(y + base + r, v);
                nn += (double)v * (double)v;
                cr += (double)v * (double)xv;
                colm[r] = xv;
                col[r] = v;
            }
        }
        {
            const double v0 = wave_sum(nn), v1 = wave_sum(cr);
            __syncthreads();   // red of the H.v phase is no longer read
            if ((tid & 63) == 0) {
                red[2 * (tid >> 6)] = v0;
                red[2 * (tid >> 6) + 1] = v1;
            }
            __syncthreads();
            if (lt == 0 && act) {
                double t0 = 0.0, t1 = 0.0;
                for (int w = 0; w < NT / 64; ++w) {
                    t0 += red[2 * (cg * (NT / 64) + w)];
                    t1 += red[2 * (cg * (NT / 64) + w) + 1];
                }
                st_agent(lz.nrm(j + 1) + ps, t0);
                st_agent(lz.crs(j + 1) + ps, t1);
            }
        }
        bprev = bj;
        ++j;
        if (lead) {
            st->j_k1 = j;
            st->j_k2 = j - 1;
        }
        // arrive at the exchange barrier, then the local part of the next H.v while it fills
        const bool last = s + 1 == nsteps;
        bool rel = false;
        if (!last) rel = grid_arrive(st, ++nbar);
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int r = lt + i * NT;
            zl[i] = (T)0;
            if (!last && act && r < dup) {
                const T xv = col[r];
                T acc = ((du[r] + ddg) + uimp[cup[r] * ncfg + cd]) * xv;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const uint32_t w = pk[k * dup + r];
                    acc += coef[w >> 16] * col[w & 0xFFFFu];
                }
                zl[i] = acc;
            }
        }
        if (!last && grid_wait(st, nbar, rel, spin_limit, flag, MON ? &st->stop_at : nullptr, j - 1)) return;
    }
}

// ------------------------- persistent Lanczos with one grid barrier per step

// Partial sums of the one-barrier kernel: array k (0 A = <z_j, u_j>, 1 X = <u_j, u_{j-1}>,
// 2 E = ||u_j||^2, 3 N = ||q_{j+1}||^2, 4 C = <q_{j+1}, u_j>) of the step j that wrote it,
// by parity of j, one slot per workgroup (the sum over its columns), kP1Stride slots each.
constexpr int kP1Arrays = 5;
constexpr int kP1Stride = 8192;
__host__ __device__ inline double *p1_arr(double *p1, int k, int j)
{
    return p1 + (size_t)(2 * k + (j & 1)) * kP1Stride;
}

// u_j = q_j - c u_{j-1}: one fused multiply-add in FP64, rounded once to T, so the
// workgroup that owns an element and every workgroup that gathers it form the same bits.
template <typename T>
__device__ inline T lz_unfold(T q, T um, double c)
{
    return (T)__builtin_fma(-c, (double)um, (double)q);
}

// n <= 256 partial sums, 4 per lane, with coherent loads, then their fixed-order wave sum
// (lane l adds entries l, l + 64, ...): split in two so that other loads can be issued
// between the loads and the sum.
__device__ inline void part_load4(const double *p, int n, double *v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = q * 64 + lane;
        v[q] = i < n ? ld_agent(p + i) : 0.0;
    }
}
__device__ inline double part_sum4(const double *v)
{
    return wave_sum(((v[0] + v[1]) + v[2]) + v[3]);
}

// Squared-norm ratio below which beta_j^2 = N - 2 c C + c^2 E (the expansion of
// ||q_j - c u_{j-1}||^2) has lost too many digits to cancellation: beta_j is then formed
// from ||u_j||^2 itself, at the cost of one more grid barrier.
constexpr double kP1Cancel = 1e-3;

// Gathered elements (q and u of each) a thread of the one-barrier kernel keeps in flight:
// the row's whole batch of down hops and Hund targets at table width 8 (more spill).
constexpr int kP1LoadElems = 20;

// `nsteps` Lanczos steps of the recurrence of lz_persist_cols_body (same map, same LDS
// layout, same scalars up to rounding) with ONE grid barrier per step instead of two.
// The barrier of the two-barrier kernel that only serves alpha_j is removed by not
// subtracting alpha_j u_j before the exchange: step j stores
//     q_{j+1} = z_j / beta_j - (beta_j / beta_{j-1}) u_{j-1},   z_j = H u_j,
// for the other workgroups (two exchange vectors, by parity), together with one partial
// sum per workgroup of each of <z_j, u_j>, <u_j, u_{j-1}>, ||u_j||^2, ||q_{j+1}||^2 and
// <q_{j+1}, u_j>.  After the barrier every workgroup forms
//     alpha_j = <z_j, u_j> / beta_j^2 - <u_j, u_{j-1}> / beta_{j-1},   c = alpha_j / beta_j,
//     beta_{j+1}^2 = ||q_{j+1}||^2 - 2 c <q_{j+1}, u_j> + c^2 ||u_j||^2,
// then u_{j+1} = q_{j+1} - c u_j for its own rows (LDS and the Krylov slot), and every
// neighbour element it gathers as the same fused multiply-add of q_{j+1} and u_j loaded
// from memory, so H u_{j+1} is H applied to the stored u_{j+1}: nothing is propagated
// from one step to the next but the vectors themselves.  beta_{j+1} differs from the
// norm of the stored vector by the cancellation of the expansion (N / beta^2 is about
// (alpha^2 + beta^2) / beta^2, a few hundred eps); where that ratio exceeds 1 / kP1Cancel
// (or always, with exact_beta) the norm is reduced over the grid instead.  The last
// step (or the one a monitored run stops after) ends with the two-barrier kernel's state:
// u_{j+1} in its slot, its norm and <u_{j+1}, u_j> partials in nrm / crs, j_k1 = j + 1.
template <typename T, int K, int R, bool MON>
__device__ inline void lz_persist_1b_body(const HvArgs<T> &a, const LzArgs<T> &lz, const PersistMap &pm,
                                          int nsteps, int spin_limit, LzMonitor mon, T *qx, double *p1,
                                          int exact_beta, unsigned char *smem, double *red, double *sh,
                                          int *flag)
{
    constexpr int NT = kBlock;                 // threads per column
    constexpr int NL = K + kPersistHund;       // gathered elements per row
    // rows per batch of gathers (q and u per element)
    constexpr int RB = (2 * NL * R <= kP1LoadElems) ? R : (2 * NL * 2 <= kP1LoadElems) ? 2 : 1;
    LzState *st = lz.st;
    if (st->status) return;
    const int dup = a.dup, ncfg = a.ncfg, nq = ncfg * ncfg, nh = a.nhund, cpw = pm.cpw;
    const PersistLds L = persist_cols_lds(dup, K, a.ncoef_up, ncfg, nh, a.hund, cpw, sizeof(T), 3);
    uint32_t *pk = reinterpret_cast<uint32_t *>(smem + L.pk);
    T *du = reinterpret_cast<T *>(smem + L.du);
    int *cup = reinterpret_cast<int *>(smem + L.cu);
    uint16_t *swp = reinterpret_cast<uint16_t *>(smem + L.sw);
    T *coef = reinterpret_cast<T *>(smem + L.coef);
    T *uimp = reinterpret_cast<T *>(smem + L.uimp);
    T *hval = reinterpret_cast<T *>(smem + L.hval);
    int *hptr = reinterpret_cast<int *>(smem + L.hptr);
    int *htr = reinterpret_cast<int *>(smem + L.htr);
    const int tid = threadIdx.x, nt = blockDim.x;
    const int cg = __builtin_amdgcn_readfirstlane(tid >> 8), lt = tid & (NT - 1);
    // three vectors of the column, by step j: vb[j % 3] holds q_j (u_j once formed),
    // vb[(j + 2) % 3] u_{j-1}, and q_{j+1} goes to vb[(j + 1) % 3]
    T *vb = reinterpret_cast<T *>(smem + L.cols + L.col_stride * cg);
    int *swdn = reinterpret_cast<int *>(vb + 3 * dup);
    T *cod = reinterpret_cast<T *>(reinterpret_cast<unsigned char *>(vb) + al16(3 * sizeof(T) * dup + 4 * ncfg));
    int *cjd = reinterpret_cast<int *>(cod + K);
    constexpr bool DL = true;   // down-hop entries from LDS: their registers go to the gathers
    const bool lead = blockIdx.x == 0 && tid == 0;
    unsigned nbar = 0;
    int j = st->j_k1;
    const int ido = __builtin_amdgcn_readfirstlane(pm.cols[blockIdx.x * cpw + cg]);
    const bool act = ido >= 0;
    const int ps = act ? pm.pslot[ido] : 0;
    int cd = 0, nd = 0;
    T ddg = (T)0;
    T od[K];
    int jd[K];
    if (act) {
        cd = a.imp_dn[ido];
        nd = a.ptr_dn[ido + 1] - a.ptr_dn[ido];
        ddg = a.diag_dn[ido];
    }
    if constexpr (DL) {
        if (lt < K) {
            cod[lt] = (act && lt < nd) ? a.off_dn[a.ptr_dn[ido] + lt] : (T)0;
            cjd[lt] = (act && lt < nd) ? a.rk_dn[a.ptr_dn[ido] + lt] * dup : 0;
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            od[k] = (act && k < nd) ? a.off_dn[a.ptr_dn[ido] + k] : (T)0;
            jd[k] = (act && k < nd) ? a.rk_dn[a.ptr_dn[ido] + k] * dup : 0;
        }
    }
    const int base = act ? ido * dup : 0;
    for (int e = tid; e < K * dup; e += nt) pk[e] = a.ell_pk_up[e];
    for (int e = tid; e < dup; e += nt) {
        du[e] = a.diag_up[e];
        cup[e] = a.imp_up[e];
    }
    for (int e = tid; e < a.ncoef_up; e += nt) coef[e] = a.coef_up[e];
    for (int e = tid; e < nq; e += nt) uimp[e] = a.uimp[e];
    if (a.hund) {
        for (int e = tid; e < dup * ncfg; e += nt) swp[e] = (uint16_t)a.swap_up[e];
        for (int e = tid; e <= nq; e += nt) hptr[e] = a.hund_ptr[e];
        for (int e = tid; e < nh; e += nt) {
            htr[e] = a.hund_tr[e];
            hval[e] = a.hund_val[e];
        }
        if (act)
            for (int e = lt; e < ncfg; e += NT) swdn[e] = a.swap_dn[ido * ncfg + e];
    }
    {
        const T *x0 = lz.slot(j);
        const T *xm = lz.slot(j > 1 ? j - 1 : j);
        T *c0 = vb + (j % 3) * dup, *cm = vb + ((j + 2) % 3) * dup;
        for (int r = lt; r < dup; r += NT) {
            c0[r] = act ? x0[base + r] : (T)0;
            cm[r] = (act && j > 1) ? xm[base + r] : (T)0;
        }
    }
    double bj1 = j > 1 ? st->beta[j - 2] : 1.0;   // beta_{j-1}
    double bj2 = j > 2 ? st->beta[j - 3] : 1.0;   // beta_{j-2}
    const int nwg = gridDim.x;
    int hflag = 0;
    __syncthreads();
    for (int s = 0;; ++s) {
        const int w = tid >> 6;
        T *col = vb + (j % 3) * dup, *colm = vb + ((j + 2) % 3) * dup, *colq = vb + ((j + 1) % 3) * dup;
        // ---- gather sources of step j: q_j (u_j itself in the first step) and u_{j-1}
        const T *G = s == 0 ? lz.slot(j) : qx + (int64_t)(j & 1) * lz.ld;
        const T *U = s == 0 ? G : lz.slot(j - 1);
        // ---- scalars: beta_j (and alpha_{j-1}, c with u_j = q_j - c u_{j-1})
        double c = 0.0, bj;
        if (s == 0) {
            // u_j is whole in its slot: beta_j from the update partials of the step that made it
            if (w == 0) {
                const double t = wave_part_sum<true>(lz.nrm(j), lz.nB);
                if (tid == 0) sh[0] = t;
            }
            if (MON && lead) hflag = __hip_atomic_load(mon.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __syncthreads();
            bj = sqrt(sh[0]);
        } else {
            // wave w reduces arrays w, w + (waves), ... (a one-column workgroup has 4 waves)
            for (int k = w; k < kP1Arrays; k += nt >> 6) {
                double t;
                if (nwg <= 256) {
                    double pv[4];
                    part_load4(p1_arr(p1, k, j - 1), nwg, pv);
                    t = part_sum4(pv);
                } else {
                    t = wave_part_sum<true>(p1_arr(p1, k, j - 1), nwg);
                }
                if ((tid & 63) == 0) sh[k] = t;
            }
            __syncthreads();
            const double alpha = sh[0] / (bj1 * bj1) - (j > 2 ? sh[1] / bj2 : 0.0);
            c = alpha / bj1;
            double b2 = sh[3] - 2.0 * c * sh[4] + c * c * sh[2];
            if (exact_beta || !(b2 >= kP1Cancel * sh[3])) {
                // ||u_j||^2 over the grid: one more barrier (the same branch in every workgroup)
                double nn = 0.0;
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    const int r = lt + i * NT;
                    if (act && r < dup) {
                        const double u = (double)lz_unfold(col[r], colm[r], c);
                        nn += u * u;
                    }
                }
                const double v = wave_sum(nn);
                if ((tid & 63) == 0) red[w] = v;
                __syncthreads();
                if (tid == 0) {
                    double t = 0.0;
                    for (int q = 0; q < (nt >> 6); ++q) t += red[q];
                    st_agent(p1_arr(p1, 2, j) + blockIdx.x, t);
                }
                if (grid_barrier(st, ++nbar, spin_limit, flag)) return;
                if (w == 0) {
                    const double t = wave_part_sum<true>(p1_arr(p1, 2, j), nwg);
                    if (tid == 0) sh[5] = t;
                }
                __syncthreads();
                b2 = sh[5];
                if (lead) atomicAdd(&st->exact_beta_steps, 1);
            }
            bj = sqrt(fmax(b2, 0.0));
            if (lead) {
                st->alpha[j - 2] = alpha;
                if (MON) {
                    __hip_atomic_store(mon.ab + 2 * (j - 2), bj1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    __hip_atomic_store(mon.ab + 2 * (j - 2) + 1, alpha, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    hflag = __hip_atomic_load(mon.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                }
            }
        }
        if (bj < kMinLanczosNorm) {   // the reference STOPs here (Lanczos_plain.h:70-92)
            if (lead) {
                st->beta[j - 1] = bj;
                st->status = (j == 1) ? 2 : 1;
            }
            return;
        }
        if (lead) st->beta[j - 1] = bj;
        // ---- u_j of the own rows (LDS and its Krylov slot)
        if (s > 0) {
            T *uj = lz.slot(j);
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int r = lt + i * NT;
                if (act && r < dup) {
                    const T u = lz_unfold(col[r], colm[r], c);
                    col[r] = u;
                    st_agent(uj + base + r, u);
                }
            }
            __syncthreads();   // the up hops read other rows of col
        }
        // ---- z = H u_j (diagonal and up hops from LDS, down hops and Hund terms gathered),
        // q_{j+1} = z / beta_j - (beta_j / beta_{j-1}) u_{j-1} for the other workgroups
        const double cz = 1.0 / bj, cp = j > 1 ? bj / bj1 : 0.0;
        T *Q = qx + (int64_t)((j + 1) & 1) * lz.ld;
        double dA = 0.0, dX = 0.0, dE = 0.0, dN = 0.0, dC = 0.0;
#pragma unroll
        for (int i0 = 0; i0 < R; i0 += RB) {
            // diagonal and up hops of the batch's rows first (LDS), then its gathers in flight
            T zl[RB];
#pragma unroll
            for (int b = 0; b < RB; ++b) {
                const int r = lt + (i0 + b) * NT;
                zl[b] = (T)0;
                if (act && r < dup) {
                    T acc = ((du[r] + ddg) + uimp[cup[r] * ncfg + cd]) * col[r];
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const uint32_t e = pk[k * dup + r];
                        acc += coef[e >> 16] * col[e & 0xFFFFu];
                    }
                    zl[b] = acc;
                }
            }
            T xq[RB][NL], xu[RB][NL];
            int h0[RB], nht[RB];
#pragma unroll
            for (int b = 0; b < RB; ++b) {
                const int r = lt + (i0 + b) * NT;
                const bool ok = act && r < dup;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int o = (DL ? cjd[k] : jd[k]) + r;
                    xq[b][k] = (ok && k < nd) ? ld_agent(G + o) : (T)0;
                    xu[b][k] = (ok && k < nd) ? ld_agent(U + o) : (T)0;
                }
                h0[b] = 0;
                nht[b] = 0;
                if (a.hund && ok) {
                    const int qq = cup[r] * ncfg + cd;
                    h0[b] = hptr[qq];
                    nht[b] = hptr[qq + 1] - h0[b];
                }
#pragma unroll
                for (int h = 0; h < kPersistHund; ++h) {
                    xq[b][K + h] = (T)0;
                    xu[b][K + h] = (T)0;
                    if (h < nht[b]) {
                        const int tr = htr[h0[b] + h];
                        const int o = (int)swp[r * ncfg + (tr & 0xFF)] + swdn[tr >> 8] * dup;
                        xq[b][K + h] = ld_agent(G + o);
                        xu[b][K + h] = ld_agent(U + o);
                    }
                }
            }
#pragma unroll
            for (int b = 0; b < RB; ++b) {
                const int i = i0 + b, r = lt + i * NT;
                if (act && r < dup) {
                    const T xv = col[r];
                    T acc = zl[b];
#pragma unroll
                    for (int k = 0; k < K; ++k) acc += (DL ? cod[k] : od[k]) * lz_unfold(xq[b][k], xu[b][k], c);
#pragma unroll
                    for (int h = 0; h < kPersistHund; ++h)
                        if (h < nht[b]) acc += hval[h0[b] + h] * lz_unfold(xq[b][K + h], xu[b][K + h], c);
                    const double um = (double)colm[r];
                    dA += (double)acc * (double)xv;
                    dX += (double)xv * um;
                    dE += (double)xv * (double)xv;
                    const T q = (T)(cz * (double)acc - cp * um);
                    st_agent(Q + base + r, q);
                    colq[r] = q;
                    dN += (double)q * (double)q;
                    dC += (double)q * (double)xv;
                }
            }
            if constexpr (RB < R) __builtin_amdgcn_sched_barrier(0);
        }
        {
            const double v[kP1Arrays] = {wave_sum(dA), wave_sum(dX), wave_sum(dE), wave_sum(dN), wave_sum(dC)};
            if ((tid & 63) == 0)
#pragma unroll
                for (int k = 0; k < kP1Arrays; ++k) red[kP1Arrays * w + k] = v[k];
            __syncthreads();
            if (tid < kP1Arrays) {   // the workgroup's sum: its waves in order
                double t = 0.0;
                for (int q = 0; q < (nt >> 6); ++q) t += red[kP1Arrays * q + tid];
                st_agent(p1_arr(p1, tid, j) + blockIdx.x, t);
            }
        }
        if (MON && lead && hflag) st_agent(&st->stop_at, j);
        bj2 = bj1;
        bj1 = bj;
        ++j;
        if (lead) {
            st->j_k1 = j;
            st->j_k2 = j - 1;
        }
        const int fl = grid_barrier(st, ++nbar, spin_limit, flag, MON ? &st->stop_at : nullptr, j - 1);
        if (fl == 1) return;
        if (fl == 2 || s + 1 == nsteps) break;
    }
    // ---- end of the launch: alpha_{j-1}, u_j = q_j - c u_{j-1} into its slot, and the
    // partials of ||u_j||^2 and <u_j, u_{j-1}> where the two-barrier kernel leaves them
    {
        const int w = tid >> 6;
        const T *col = vb + (j % 3) * dup, *colm = vb + ((j + 2) % 3) * dup;
        if (w < 2) {
            const double t = wave_part_sum<true>(p1_arr(p1, w, j - 1), nwg);
            if ((tid & 63) == 0) sh[w] = t;
        }
        __syncthreads();
        const double alpha = sh[0] / (bj1 * bj1) - (j > 2 ? sh[1] / bj2 : 0.0);
        const double c = alpha / bj1;
        if (lead) {
            st->alpha[j - 2] = alpha;
            if (MON) {
                __hip_atomic_store(mon.ab + 2 * (j - 2), bj1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(mon.ab + 2 * (j - 2) + 1, alpha, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
        T *uj = lz.slot(j);
        double nn = 0.0, cr = 0.0;
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int r = lt + i * NT;
            if (act && r < dup) {
                const T u = lz_unfold(col[r], colm[r], c);
                st_agent(uj + base + r, u);
                nn += (double)u * (double)u;
                cr += (double)u * (double)colm[r];
            }
        }
        const double v0 = wave_sum(nn), v1 = wave_sum(cr);
        if ((tid & 63) == 0) {
            red[2 * w] = v0;
            red[2 * w + 1] = v1;
        }
        __syncthreads();
        if (lt == 0 && act) {
            double t0 = 0.0, t1 = 0.0;
            for (int q = 0; q < NT / 64; ++q) {
                t0 += red[2 * (cg * (NT / 64) + q)];
                t1 += red[2 * (cg * (NT / 64) + q) + 1];
            }
            st_agent(lz.nrm(j) + ps, t0);
            st_agent(lz.crs(j) + ps, t1);
        }
        // canonical slots no column writes hold zeros (the first step of this launch read
        // nrm of its parity before any barrier, so they are cleared only here)
        if (blockIdx.x == 0)
            for (int e = tid; e < pm.nidle; e += nt) {
                st_agent(lz.nrm(j) + pm.idle[e], 0.0);
                st_agent(lz.crs(j) + pm.idle[e], 0.0);
            }
    }
}

// ------------------------------------ persistent Lanczos, several runs in one launch

// One Lanczos run of the multi-run persistent kernel: its sector's tables, its recurrence
// (Krylov slots, scalars, one partial sum per column of the sector: partA[d_dn] and, by
// parity of j, nrm / crs with pstride >= d_dn), its start vector (FP64, unit norm), the
// workgroups [wg0, wg0 + nwg) of the grid that serve it, their column map
// cmap[(w * nround + round) * 4 + group] (-1 = idle) and the sector's table width kk.
template <typename T>
struct PersistRun {
    HvArgs<T> a;
    LzArgs<T> lz;
    const double *start;
    const int32_t *cmap;
    int wg0, nwg, nround, kk;
};

// Column groups (of 256 threads) per workgroup of the multi-run kernel, and the most
// rounds of columns a workgroup takes.
constexpr int kRunCpw = 4;
constexpr int kRunMaxRounds = 16;

// Dynamic LDS of the multi-run kernel for one run's sector (offsets 16-byte aligned):
//   the shared tables of persist_cols_lds for kk table rows;
//   per (round, group) column record: cod[K] (T), cjd[K], swdn[ncfg], ci, cd, nd (i32),
//   ddg (T) -- the column's down-hop entries and what the diagonal and Hund terms need;
//   per group two column buffers col[2][dup] (T): u_j of the round being computed and of
//   the next one, staged while this one computes;
//   wave sums sums[round][16][2] (double).
struct RunsLds {
    PersistLds t;
    size_t rec, rec_stride, col, col_stride, sums, total;
};
__host__ __device__ inline RunsLds persist_runs_lds(int dup, int kk, int K, int ncoef, int ncfg, int nhund, int hund,
                                                    int nround, size_t tsize)
{
    RunsLds L;
    L.t = persist_cols_lds(dup, kk, ncoef, ncfg, nhund, hund, 0, tsize);
    L.rec = L.t.cols;
    L.rec_stride = al16(tsize * (K + 1) + 4 * (size_t)(K + ncfg + 3));
    L.col = al16(L.rec + L.rec_stride * (size_t)nround * kRunCpw);
    L.col_stride = al16(2 * tsize * dup);
    L.sums = L.col + L.col_stride * kRunCpw;
    L.total = L.sums + sizeof(double) * 2 * 16 * (size_t)nround;
    return L;
}

// One column record of the multi-run kernel (views into LDS).
template <typename T, int K>
struct RunCol {
    T *od;        // [K] down-hop coefficients (0 beyond nd)
    int *jd;      // [K] neighbour column offsets rk * dup
    int *swdn;    // [ncfg] down state with its impurity bits replaced (Hund)
    int *hdr;     // ci, cd, nd
    T *ddg;       // diagonal of the down state
    __device__ RunCol(unsigned char *p, int ncfg)
    {
        od = reinterpret_cast<T *>(p);
        ddg = od + K;
        jd = reinterpret_cast<int *>(ddg + 1);
        swdn = jd + K;
        hdr = swdn + ncfg;
    }
};

// `nsteps` Lanczos steps of every run of `runs` in ONE cooperative launch (the excitation
// runs of a Green's function, green_class_computeAA.h:101-302, each a compute_dynamic
// recurrence, green_class_compute_dynamic.f90:279-334, in sectors of uneven dimension).
// Workgroup b serves the run whose range holds it; its four column groups take the columns
// of the run's column map round by round.  What a column needs besides u_j (down-hop
// entries, diagonal, impurity bits, Hund swap row) is read once, at launch, into LDS.
// H.v phase, per round: a group gathers the down-hop and Hund neighbours of its column from
// the other columns with device-coherent loads while the next round's column of u_j is
// loaded, computes z = H u_j with the up hops and the diagonal from the staged column in
// LDS, stores z in the slot of u_{j+1} and keeps the wave sums of <z, u_j>; then it stages
// the next column (one workgroup barrier per round).  The per-column partial sums are
// formed after the last round, in the order lz_persist_cols_body uses.  One grid barrier
// for all runs; then every workgroup forms its run's scalars from the fixed-order sums
// and updates its columns (u_{j+1} over z; wave sums of ||u_{j+1}||^2 and <u_{j+1}, u_j>,
// added per column after the last round); a second grid barrier ends the step.  The
// partial sums are indexed by column, so the sums do not depend on the map.  A run whose
// beta falls below kMinLanczosNorm records it (status 1, or 2 at j = 1) and idles; the
// others go on.  Every workgroup reaches every barrier, so the launch always drains; a
// barrier wait past spin_limit ends it with bar->status = 3.  u_1 = start is written in
// the prologue.
template <typename T, int K, int R>
__device__ inline void lz_persist_runs_body(const PersistRun<T> *__restrict__ runs, int nruns, LzState *bar,
                                            int nsteps, int spin_limit, unsigned char *smem, double *sh, int *flag)
{
    constexpr int NT = kBlock;
    constexpr int LB = K * R * (int)sizeof(T);
    constexpr int RB = (K * R <= 16) ? R : (LB >= 384 ? 1 : 2);
    int jr = 0;
    while (jr + 1 < nruns && (int)blockIdx.x >= runs[jr + 1].wg0) ++jr;
    const PersistRun<T> &P = runs[jr];
    const HvArgs<T> &a = P.a;
    const LzArgs<T> lz = P.lz;
    LzState *st = lz.st;
    const int lw = blockIdx.x - P.wg0, nround = P.nround, kk = P.kk;
    const int dup = a.dup, ddn = a.ddn, ncfg = a.ncfg, nq = ncfg * ncfg, nh = a.nhund, hund = a.hund;
    const RunsLds L = persist_runs_lds(dup, kk, K, a.ncoef_up, ncfg, nh, hund, nround, sizeof(T));
    uint32_t *pk = reinterpret_cast<uint32_t *>(smem + L.t.pk);
    T *du = reinterpret_cast<T *>(smem + L.t.du);
    int *cup = reinterpret_cast<int *>(smem + L.t.cu);
    uint16_t *swp = reinterpret_cast<uint16_t *>(smem + L.t.sw);
    T *coef = reinterpret_cast<T *>(smem + L.t.coef);
    T *uimp = reinterpret_cast<T *>(smem + L.t.uimp);
    T *hval = reinterpret_cast<T *>(smem + L.t.hval);
    int *hptr = reinterpret_cast<int *>(smem + L.t.hptr);
    int *htr = reinterpret_cast<int *>(smem + L.t.htr);
    double *sums = reinterpret_cast<double *>(smem + L.sums);
    const int tid = threadIdx.x, nt = blockDim.x;
    const int cg = __builtin_amdgcn_readfirstlane(tid >> 8), lt = tid & (NT - 1), wv = tid >> 6;
    T *colb = reinterpret_cast<T *>(smem + L.col + L.col_stride * cg);   // [2][dup]
    auto rec = [&](int rd) { return RunCol<T, K>(smem + L.rec + L.rec_stride * ((size_t)rd * kRunCpw + cg), ncfg); };
    const bool lead = lw == 0 && tid == 0;   // the run's first workgroup records its scalars
    const int32_t *cmap = P.cmap + (size_t)lw * nround * kRunCpw + cg;
    // ---- tables of the run's sector, shared by the column groups
    for (int e = tid; e < kk * dup; e += nt) pk[e] = a.ell_pk_up[e];
    for (int e = tid; e < dup; e += nt) {
        du[e] = a.diag_up[e];
        cup[e] = a.imp_up[e];
    }
    for (int e = tid; e < a.ncoef_up; e += nt) coef[e] = a.coef_up[e];
    for (int e = tid; e < nq; e += nt) uimp[e] = a.uimp[e];
    if (hund) {
        for (int e = tid; e < dup * ncfg; e += nt) swp[e] = (uint16_t)a.swap_up[e];
        for (int e = tid; e <= nq; e += nt) hptr[e] = a.hund_ptr[e];
        for (int e = tid; e < nh; e += nt) {
            htr[e] = a.hund_tr[e];
            hval[e] = a.hund_val[e];
        }
    }
    // ---- the group's column records, every round
    for (int rd = 0; rd < nround; ++rd) {
        RunCol<T, K> c = rec(rd);
        const int ci = cmap[rd * kRunCpw];
        const int p0 = ci >= 0 ? a.ptr_dn[ci] : 0, nd = ci >= 0 ? a.ptr_dn[ci + 1] - p0 : 0;
        if (lt < K) {
            c.od[lt] = lt < nd ? a.off_dn[p0 + lt] : (T)0;
            c.jd[lt] = lt < nd ? a.rk_dn[p0 + lt] * dup : 0;
        }
        if (hund && ci >= 0)
            for (int e = lt; e < ncfg; e += NT) c.swdn[e] = a.swap_dn[ci * ncfg + e];
        if (lt == 0) {
            c.hdr[0] = ci;
            c.hdr[1] = ci >= 0 ? a.imp_dn[ci] : 0;
            c.hdr[2] = nd;
            *c.ddg = ci >= 0 ? a.diag_dn[ci] : (T)0;
        }
    }
    // per-column sums of the wave sums sums[rd][16][slot] of the last phase, in wave order
    auto column_sums = [&](double *out0, double *out1) {
        __syncthreads();
        if (lt == 0)
            for (int rd = 0; rd < nround; ++rd) {
                const int ci = cmap[rd * kRunCpw];
                if (ci < 0) continue;
                double t0 = 0.0, t1 = 0.0;
                for (int q = 0; q < NT / 64; ++q) {
                    t0 += sums[(rd * 16 + cg * (NT / 64) + q) * 2];
                    t1 += sums[(rd * 16 + cg * (NT / 64) + q) * 2 + 1];
                }
                st_agent(out0 + ci, t0);
                if (out1) st_agent(out1 + ci, t1);
            }
    };
    // ---- u_1 = start (lz_init_body per column): ||u_1||^2 partials, <u_1, u_0> = 0
    {
        T *u1 = lz.slot(1);
        for (int rd = 0; rd < nround; ++rd) {
            const int ci = __builtin_amdgcn_readfirstlane(cmap[rd * kRunCpw]);
            double nn = 0.0;
            if (ci >= 0)
                for (int r = lt; r < dup; r += NT) {
                    const int64_t i = (int64_t)ci * dup + r;
                    const T v = (T)P.start[i];
                    st_agent(u1 + i, v);
                    nn += (double)v * (double)v;
                }
            const double w0 = wave_sum(nn);
            if ((tid & 63) == 0) {
                sums[(rd * 16 + wv) * 2] = w0;
                sums[(rd * 16 + wv) * 2 + 1] = 0.0;
            }
        }
        column_sums(lz.nrm(1), lz.crs(1));
        if (lead) {
            st->j_k1 = 1;
            st->j_k2 = 0;
            st->status = 0;
        }
    }
    unsigned nbar = 0;
    if (grid_barrier(bar, ++nbar, spin_limit, flag)) return;
    int j = 1;
    bool live = true;
    double bprev = 1.0;
    for (int s = 0; s < nsteps; ++s) {
        const T *__restrict__ x = lz.slot(j);
        T *__restrict__ y = lz.slot(j + 1);
        // ---- H.v of step j, round by round
        if (live) {
            {   // stage round 0
                const RunCol<T, K> c = rec(0);
                const int ci = __builtin_amdgcn_readfirstlane(c.hdr[0]);
                for (int r = lt; r < dup; r += NT) colb[r] = ci >= 0 ? ld_agent(x + (size_t)ci * dup + r) : (T)0;
                __syncthreads();
            }
            for (int rd = 0; rd < nround; ++rd) {
                const RunCol<T, K> c = rec(rd);
                const int ci = __builtin_amdgcn_readfirstlane(c.hdr[0]);
                const bool act = ci >= 0;
                const int base = act ? ci * dup : 0, cd = c.hdr[1], nd = __builtin_amdgcn_readfirstlane(c.hdr[2]);
                const T ddg = *c.ddg;
                const T *col = colb + (rd & 1) * dup;
                // the next round's column of u_j, in flight during this round's gathers
                T nx[R];
                int nbase = -1;
                if (rd + 1 < nround) nbase = __builtin_amdgcn_readfirstlane(rec(rd + 1).hdr[0]);
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    const int r = lt + i * NT;
                    nx[i] = (nbase >= 0 && r < dup) ? ld_agent(x + (size_t)nbase * dup + r) : (T)0;
                }
                double dot = 0.0;
#pragma unroll
                for (int i0 = 0; i0 < R; i0 += RB) {
                    T xd[RB][K], xh[RB][kPersistHund];
                    int h0[RB], nht[RB];
#pragma unroll
                    for (int b = 0; b < RB; ++b) {
                        const int r = lt + (i0 + b) * NT;
                        const bool ok = act && r < dup;
#pragma unroll
                        for (int k = 0; k < K; ++k) xd[b][k] = (ok && k < nd) ? ld_agent(x + c.jd[k] + r) : (T)0;
                        h0[b] = 0;
                        nht[b] = 0;
                        if (hund && ok) {
                            const int q = cup[r] * ncfg + cd;
                            h0[b] = hptr[q];
                            nht[b] = hptr[q + 1] - h0[b];
                        }
#pragma unroll
                        for (int h = 0; h < kPersistHund; ++h) {
                            xh[b][h] = (T)0;
                            if (h < nht[b]) {
                                const int tr = htr[h0[b] + h];
                                xh[b][h] = ld_agent(x + (int)swp[r * ncfg + (tr & 0xFF)] + c.swdn[tr >> 8] * dup);
                            }
                        }
                    }
#pragma unroll
                    for (int b = 0; b < RB; ++b) {
                        const int r = lt + (i0 + b) * NT;
                        if (act && r < dup) {
                            const T xv = col[r];
                            T acc = ((du[r] + ddg) + uimp[cup[r] * ncfg + cd]) * xv;
#pragma unroll
                            for (int k = 0; k < K; ++k)
                                if (k < kk) {
                                    const uint32_t w = pk[k * dup + r];
                                    acc += coef[w >> 16] * col[w & 0xFFFFu];
                                }
#pragma unroll
                            for (int k = 0; k < K; ++k) acc += c.od[k] * xd[b][k];
#pragma unroll
                            for (int h = 0; h < kPersistHund; ++h)
                                if (h < nht[b]) acc += hval[h0[b] + h] * xh[b][h];
                            st_agent(y + base + r, acc);
                            dot += (double)acc * (double)xv;
                        }
                    }
                    if constexpr (RB < R) __builtin_amdgcn_sched_barrier(0);
                }
                const double w0 = wave_sum(dot);
                if ((tid & 63) == 0) sums[(rd * 16 + wv) * 2] = w0;
                if (rd + 1 < nround) {
                    T *nb = colb + ((rd + 1) & 1) * dup;
#pragma unroll
                    for (int i = 0; i < R; ++i) {
                        const int r = lt + i * NT;
                        if (r < dup) nb[r] = nx[i];
                    }
                }
                __syncthreads();   // the next column staged; this one no longer read
            }
            column_sums(lz.partA, nullptr);
        }
        if (grid_barrier(bar, ++nbar, spin_limit, flag)) return;
        // ---- scalars and update of step j (N, P, Q as lz_scalars over the run's columns)
        if (live) {
            if (wv == 0) {
                const double t = wave_part_sum<true>(lz.nrm(j), ddn);
                if (tid == 0) sh[0] = t;
            } else if (wv == 1) {
                const double t = j > 1 ? wave_part_sum<true>(lz.crs(j), ddn) : 0.0;
                if (tid == 64) sh[1] = t;
            } else if (wv == 2) {
                const double t = wave_part_sum<true>(lz.partA, ddn);
                if (tid == 128) sh[2] = t;
            }
            __syncthreads();
            const double bj = sqrt(sh[0]);
            const double q = sh[2], pp = sh[1];
            if (bj < kMinLanczosNorm) {   // the reference STOPs here (Lanczos_plain.h:70-92)
                if (lead) {
                    st->beta[j - 1] = bj;
                    st->status = (j == 1) ? 2 : 1;
                }
                live = false;
            } else {
                const double alpha = q / (bj * bj) - (j > 1 ? pp / bprev : 0.0);
                if (lead) {
                    st->beta[j - 1] = bj;
                    st->alpha[j - 1] = alpha;
                }
                const double cz = 1.0 / bj, cp = j > 1 ? bj / bprev : 0.0, cx = alpha / bj;
                const T *__restrict__ xm = lz.slot(j > 1 ? j - 1 : j);
                for (int rd = 0; rd < nround; ++rd) {
                    const int ci = __builtin_amdgcn_readfirstlane(rec(rd).hdr[0]);
                    const bool act = ci >= 0;
                    const int base = act ? ci * dup : 0;
                    T zv[R], xv[R], pv[R];
#pragma unroll
                    for (int i = 0; i < R; ++i) {
                        const int r = lt + i * NT;
                        const bool ok = act && r < dup;
                        zv[i] = ok ? ld_agent(y + base + r) : (T)0;
                        xv[i] = ok ? ld_agent(x + base + r) : (T)0;
                        pv[i] = (ok && j > 1) ? ld_agent(xm + base + r) : (T)0;
                    }
                    double nn = 0.0, cr = 0.0;
#pragma unroll
                    for (int i = 0; i < R; ++i) {
                        const int r = lt + i * NT;
                        if (act && r < dup) {
                            const T v = (T)(cz * (double)zv[i] - cp * (double)pv[i] - cx * (double)xv[i]);
                            st_agent(y + base + r, v);
                            nn += (double)v * (double)v;
                            cr += (double)v * (double)xv[i];
                        }
                    }
                    const double w0 = wave_sum(nn), w1 = wave_sum(cr);
                    if ((tid & 63) == 0) {
                        sums[(rd * 16 + wv) * 2] = w0;
                        sums[(rd * 16 + wv) * 2 + 1] = w1;
                    }
                }
                column_sums(lz.nrm(j + 1), lz.crs(j + 1));
                bprev = bj;
            }
        }
        ++j;
        if (lead && live) {
            st->j_k1 = j;
            st->j_k2 = j - 1;
        }
        if (s + 1 < nsteps && grid_barrier(bar, ++nbar, spin_limit, flag)) return;
    }
}

}  // namespace ed
