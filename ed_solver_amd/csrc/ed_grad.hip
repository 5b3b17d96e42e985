// Derivatives on the device: Hellmann-Feynman gradients of an eigen-energy and the
// bath-fit distance with its analytic gradient.
//
// Hellmann-Feynman: for a normalised eigenvector psi of H(theta), dE/dtheta =
// <psi| dH/dtheta |psi>.  The parameters covered are the one-body levels of every site
// (Ec(c,c) / Eb(k,k), both spins), the real hybridisations Vbc(k,c) (both spins, the
// terms Vbc b+_k c_c + h.c. of HAIM_offdiag_arrays.f90:483-527) and the interaction
// scale U of the Kanamori form U_ii = U, U_ij = U - 2J at fixed J (HAIMupdo.f90:785-829).
// With psi the column-major d_up x d_dn matrix (istate = iup + ido * d_up):
//   <n_a,up>  = sum_r P_up[r] bit_a(state_up[r]),  P_up[r] = sum_c psi[r,c]^2
//   <n_a,dn>  = sum_c P_dn[c] bit_a(state_dn[c]),  P_dn[c] = sum_r psi[r,c]^2
//   every hop r -> r' between bath site k and impurity orbital c (either direction) of
//   the up states contributes sign * sum_c psi[r,c] psi[r',c] to d/dVbc(k,c); down hops
//   likewise with sum_r psi[r,c] psi[r,c'].  The hops come from a table built here for
//   all (k, c) pairs, not from the H.v tables, which omit couplings that are zero;
//   <dH/dU> = sum F[imp_up(r)][imp_dn(c)] psi[r,c]^2 with F the interaction operator
//   differentiated in U.
// Two passes over psi: a column kernel (workgroup per column: P_dn, the U term and the
// down-hop pair sums) and a row kernel (thread per row over a slice of the columns: P_up
// and the up-hop pair sums, then a fixed-order sum over the slices).  Deterministic.
//
// Bath-fit distance (distance_func / diff_correl, bath_class_hybrid.f90:1151-1204,
// correl_class.f90:1061-1160), for a diagonal bath shared by both spins:
//   Delta_ij(i w_n) = sum_k Vbc(k,i) Vbc(k,j) / (i w_n - eps_k)        (bath2hybrid)
//   S = sum_n sum_{sigma,i,j} |Delta_ij(i w_n) - T^sigma_ij(i w_n)|^p / (|w_n - w_1| + s)^q
//   dist = |S| / (N_w N^2), N = 2 nc (the spin-resolved matrix size of the reference)
// with the analytic gradient with respect to eps_k and Vbc(k,c).  One thread per
// (frequency, parameter); the per-frequency terms are summed in a fixed order.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "ed_kernels.h"

namespace ed {

// ------------------------------------------------------------ Hellmann-Feynman

// Column c: P_dn[c], the U term and, for every down hop k of column c (target column
// tgt_dn[k][c]; padding entries point to c itself), sum_r psi[r,c] psi[r,c'].
template <int K>
__global__ __launch_bounds__(256) void k_hf_cols(int dup, int ddn, const double *__restrict__ psi,
                                                 const int32_t *tgt_dn, const int32_t *imp_up, const int32_t *imp_dn,
                                                 const double *F, int ncfg, double *out_pdn, double *out_fu,
                                                 double *out_dn)
{
    __shared__ double red[4][K + 2];
    const int c = blockIdx.x, tid = threadIdx.x;
    const int cd = imp_dn[c];
    int jd[K];
#pragma unroll
    for (int k = 0; k < K; ++k) jd[k] = tgt_dn[(int64_t)k * ddn + c];
    double acc[K + 2];
#pragma unroll
    for (int q = 0; q < K + 2; ++q) acc[q] = 0.0;
    const double *col = psi + (int64_t)c * dup;
    for (int r = tid; r < dup; r += 256) {
        const double v = col[r];
        acc[0] += v * v;
        acc[1] += F[imp_up[r] * ncfg + cd] * v * v;
#pragma unroll
        for (int k = 0; k < K; ++k) acc[2 + k] += v * psi[(int64_t)jd[k] * dup + r];
    }
#pragma unroll
    for (int q = 0; q < K + 2; ++q) {
        const double t = wave_sum(acc[q]);
        if ((tid & 63) == 0) red[tid >> 6][q] = t;
    }
    __syncthreads();
    if (tid < K + 2) {
        const double t = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
        if (tid == 0) out_pdn[c] = t;
        else if (tid == 1) out_fu[c] = t;
        else out_dn[(int64_t)(tid - 2) * ddn + c] = t;
    }
}

// part[g][q][r]: q = 0 -> sum psi[r,c]^2, q = 1 + k -> sum psi[r,c] psi[tgt_up[k][r], c],
// over the columns of slice g.
template <int K>
__global__ __launch_bounds__(256) void k_hf_rows(int dup, int ddn, int cols_per_slice, const double *__restrict__ psi,
                                                 const int32_t *tgt_up, double *part)
{
    const int r = blockIdx.x * 256 + threadIdx.x, g = blockIdx.y;
    if (r >= dup) return;
    int rk[K];
#pragma unroll
    for (int k = 0; k < K; ++k) rk[k] = tgt_up[(int64_t)k * dup + r];
    double acc[K + 1];
#pragma unroll
    for (int q = 0; q < K + 1; ++q) acc[q] = 0.0;
    const int c0 = g * cols_per_slice, c1 = min(ddn, c0 + cols_per_slice);
    for (int c = c0; c < c1; ++c) {
        const double *col = psi + (int64_t)c * dup;
        const double v = col[r];
        acc[0] += v * v;
#pragma unroll
        for (int k = 0; k < K; ++k) acc[1 + k] += v * col[rk[k]];
    }
#pragma unroll
    for (int q = 0; q < K + 1; ++q) part[((int64_t)g * (K + 1) + q) * dup + r] = acc[q];
}

// out[q][r] = sum over slices g (in order) of part[g][q][r].
__global__ void k_hf_rows_sum(int n_q_r, int nslice, const double *part, double *out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_q_r) return;
    double s = 0.0;
    for (int g = 0; g < nslice; ++g) s += part[(int64_t)g * n_q_r + e];
    out[e] = s;
}

template <int K>
static void launch_hf(const ed_sector *s, const double *psi, const int32_t *tgt_dn, const double *F, double *pdn,
                      double *fu, double *odn, hipStream_t st)
{
    hipLaunchKernelGGL(k_hf_cols<K>, dim3(s->dn.d), dim3(256), 0, st, s->up.d, s->dn.d, psi, tgt_dn, s->up.imp,
                       s->dn.imp, F, s->ncfg, pdn, fu, odn);
}

template <int K>
static void launch_hf_rows(const ed_sector *s, const double *psi, const int32_t *tgt_up, double *part, int nslice,
                           int cps, hipStream_t st)
{
    hipLaunchKernelGGL(k_hf_rows<K>, dim3((s->up.d + 255) / 256, nslice), dim3(256), 0, st, s->up.d, s->dn.d, cps, psi,
                       tgt_up, part);
}

static int hop_sign(uint32_t s, int a, int b);

// Hops between bath sites and impurity orbitals of one spin's states, in ELL form of
// width K: tgt[k][i] (target rank, i itself for padding), par[k][i] = k_bath * nc + c
// (or -1), sgn[k][i] = fermion sign of the hop.
struct HopTable {
    int K = 0;
    std::vector<int32_t> tgt, par, sgn;
};

static int pick_width(int kmax);

static HopTable hyb_hops(const std::vector<int32_t> &states, int nb, int nc)
{
    const int d = (int)states.size();
    std::vector<int32_t> rank(1u << (nb + nc), -1);
    for (int i = 0; i < d; ++i) rank[(uint32_t)states[i]] = i;
    std::vector<std::vector<int>> lst(d);
    int kmax = 1;
    for (int i = 0; i < d; ++i) {
        const uint32_t st = (uint32_t)states[i];
        for (int k = 0; k < nb; ++k)
            for (int c = 0; c < nc; ++c) {
                const int bk = k, bc = nb + c;
                const int nk = (st >> bk) & 1, ncc = (st >> bc) & 1;
                if (nk == ncc) continue;   // no hop between k and c in this state
                const uint32_t t = st ^ (1u << bk) ^ (1u << bc);
                lst[i].push_back(rank[t]);
                lst[i].push_back(k * nc + c);
                // created = the empty one of (k, c), destroyed = the occupied one
                lst[i].push_back(hop_sign(st, nk ? bc : bk, nk ? bk : bc));
            }
        kmax = std::max(kmax, (int)lst[i].size() / 3);
    }
    HopTable T;
    T.K = pick_width(kmax);
    T.tgt.assign((size_t)T.K * d, 0);
    T.par.assign((size_t)T.K * d, -1);
    T.sgn.assign((size_t)T.K * d, 0);
    for (int i = 0; i < d; ++i)
        for (int k = 0; k < T.K; ++k) {
            const bool ok = k < (int)lst[i].size() / 3;
            T.tgt[(size_t)k * d + i] = ok ? lst[i][3 * k] : i;
            if (ok) {
                T.par[(size_t)k * d + i] = lst[i][3 * k + 1];
                T.sgn[(size_t)k * d + i] = lst[i][3 * k + 2];
            }
        }
    return T;
}

static int pick_width(int kmax)
{
    const int widths[] = {4, 8, 12, 16, 24, 32};
    for (int w : widths)
        if (kmax <= w) return w;
    return 0;
}

// Parity of the occupied bits strictly between a and b (the sign of c+_a c_b).
static int hop_sign(uint32_t s, int a, int b)
{
    const int lo = std::min(a, b), hi = std::max(a, b);
    const uint32_t mask = (hi - lo > 1) ? (((1u << (hi - lo - 1)) - 1u) << (lo + 1)) : 0u;
    return (__builtin_popcount(s & mask) & 1) ? -1 : 1;
}

static int hf_gradient(ed_sector *s, const double *d_psi, double *d_level, double *d_v, double *d_u)
{
    const int dup = s->up.d, ddn = s->dn.d, nc = s->nc, nb = s->nb, ns = s->ns, ncfg = s->ncfg;
    hipStream_t st = s->stream;
    // F[cu][cd]: the interaction operator differentiated in U (U_ii = U, U_ij = U - 2J)
    std::vector<double> F((size_t)ncfg * ncfg, 0.0);
    for (int cu = 0; cu < ncfg; ++cu)
        for (int cd = 0; cd < ncfg; ++cd) {
            double f = 0.0;
            for (int c = 0; c < nc; ++c)
                if (s->u[c * nc + c] != 0.0 && ((cu >> c) & 1) && ((cd >> c) & 1)) f += 1.0;
            for (int c1 = 0; c1 < nc; ++c1)
                for (int c2 = c1 + 1; c2 < nc; ++c2) {
                    if (s->u[c1 * nc + c2] == 0.0) continue;
                    f += (double)((((cu >> c1) & 1) + ((cd >> c1) & 1)) * (((cu >> c2) & 1) + ((cd >> c2) & 1)));
                }
            F[(size_t)cu * ncfg + cd] = f;
        }
    std::vector<int32_t> st_up(dup), st_dn(ddn);
    ED_HIP_CHECK(hipMemcpyAsync(st_up.data(), s->up.states, sizeof(int32_t) * dup, hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipMemcpyAsync(st_dn.data(), s->dn.states, sizeof(int32_t) * ddn, hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipStreamSynchronize(st));
    const HopTable Hu = hyb_hops(st_up, nb, nc), Hd = hyb_hops(st_dn, nb, nc);
    if (!Hu.K || !Hd.K) return fail(ED_ERR_ARG, "hf_gradient: more than 32 hybridisation hops per state");
    const int nslice = std::max(1, std::min(64, ddn / 8));
    const int cps = (ddn + nslice - 1) / nslice;
    const size_t n_part = (size_t)nslice * (Hu.K + 1) * dup, n_rows = (size_t)(Hu.K + 1) * dup;
    const size_t n_odn = (size_t)Hd.K * ddn;
    const size_t n_dbl = ncfg * ncfg + 2 * (size_t)ddn + n_odn + n_part + n_rows;
    const size_t n_int = (size_t)Hu.K * dup + (size_t)Hd.K * ddn;
    char *w = nullptr;
    ED_HIP_CHECK(hipMalloc((void **)&w, sizeof(double) * n_dbl + sizeof(int32_t) * n_int));
    double *dF = (double *)w, *pdn = dF + ncfg * ncfg, *fu = pdn + ddn, *odn = fu + ddn, *part = odn + n_odn,
           *rows = part + n_part;
    int32_t *tgt_up = (int32_t *)(rows + n_rows), *tgt_dn = tgt_up + (size_t)Hu.K * dup;
    std::vector<double> h_pdn(ddn), h_fu(ddn), h_odn(n_odn), h_rows(n_rows);
    auto body = [&]() -> int {
        ED_HIP_CHECK(hipMemcpyAsync(dF, F.data(), sizeof(double) * F.size(), hipMemcpyHostToDevice, st));
        ED_HIP_CHECK(hipMemcpyAsync(tgt_up, Hu.tgt.data(), sizeof(int32_t) * Hu.tgt.size(), hipMemcpyHostToDevice, st));
        ED_HIP_CHECK(hipMemcpyAsync(tgt_dn, Hd.tgt.data(), sizeof(int32_t) * Hd.tgt.size(), hipMemcpyHostToDevice, st));
        switch (Hd.K) {
        case 4: launch_hf<4>(s, d_psi, tgt_dn, dF, pdn, fu, odn, st); break;
        case 8: launch_hf<8>(s, d_psi, tgt_dn, dF, pdn, fu, odn, st); break;
        case 12: launch_hf<12>(s, d_psi, tgt_dn, dF, pdn, fu, odn, st); break;
        case 16: launch_hf<16>(s, d_psi, tgt_dn, dF, pdn, fu, odn, st); break;
        case 24: launch_hf<24>(s, d_psi, tgt_dn, dF, pdn, fu, odn, st); break;
        default: launch_hf<32>(s, d_psi, tgt_dn, dF, pdn, fu, odn, st); break;
        }
        switch (Hu.K) {
        case 4: launch_hf_rows<4>(s, d_psi, tgt_up, part, nslice, cps, st); break;
        case 8: launch_hf_rows<8>(s, d_psi, tgt_up, part, nslice, cps, st); break;
        case 12: launch_hf_rows<12>(s, d_psi, tgt_up, part, nslice, cps, st); break;
        case 16: launch_hf_rows<16>(s, d_psi, tgt_up, part, nslice, cps, st); break;
        case 24: launch_hf_rows<24>(s, d_psi, tgt_up, part, nslice, cps, st); break;
        default: launch_hf_rows<32>(s, d_psi, tgt_up, part, nslice, cps, st); break;
        }
        hipLaunchKernelGGL(k_hf_rows_sum, dim3((unsigned)((n_rows + 255) / 256)), dim3(256), 0, st, (int)n_rows, nslice,
                           (const double *)part, rows);
        ED_HIP_CHECK(hipGetLastError());
        ED_HIP_CHECK(hipMemcpyAsync(h_pdn.data(), pdn, sizeof(double) * ddn, hipMemcpyDeviceToHost, st));
        ED_HIP_CHECK(hipMemcpyAsync(h_fu.data(), fu, sizeof(double) * ddn, hipMemcpyDeviceToHost, st));
        ED_HIP_CHECK(hipMemcpyAsync(h_odn.data(), odn, sizeof(double) * n_odn, hipMemcpyDeviceToHost, st));
        ED_HIP_CHECK(hipMemcpyAsync(h_rows.data(), rows, sizeof(double) * n_rows, hipMemcpyDeviceToHost, st));
        ED_HIP_CHECK(hipStreamSynchronize(st));
        return ED_OK;
    };
    const int rc = body();
    (void)hipFree(w);
    if (rc != ED_OK) return rc;
    double nrm = 0.0, fuu = 0.0;
    for (int c = 0; c < ddn; ++c) {
        nrm += h_pdn[c];
        fuu += h_fu[c];
    }
    if (!(nrm > 0.0)) return fail(ED_ERR_ARG, "hf_gradient: zero vector");
    std::vector<double> lev(ns, 0.0), dv((size_t)nb * nc, 0.0);
    for (int r = 0; r < dup; ++r)
        for (int a = 0; a < ns; ++a)
            if ((st_up[r] >> a) & 1) lev[a] += h_rows[r];   // q = 0: P_up[r]
    for (int c = 0; c < ddn; ++c)
        for (int a = 0; a < ns; ++a)
            if ((st_dn[c] >> a) & 1) lev[a] += h_pdn[c];
    for (int k = 0; k < Hu.K; ++k)
        for (int r = 0; r < dup; ++r) {
            const int p = Hu.par[(size_t)k * dup + r];
            if (p >= 0) dv[p] += Hu.sgn[(size_t)k * dup + r] * h_rows[(size_t)(1 + k) * dup + r];
        }
    for (int k = 0; k < Hd.K; ++k)
        for (int c = 0; c < ddn; ++c) {
            const int p = Hd.par[(size_t)k * ddn + c];
            if (p >= 0) dv[p] += Hd.sgn[(size_t)k * ddn + c] * h_odn[(size_t)k * ddn + c];
        }
    if (d_level)
        for (int a = 0; a < ns; ++a) d_level[a] = lev[a] / nrm;
    if (d_v)
        for (size_t i = 0; i < dv.size(); ++i) d_v[i] = dv[i] / nrm;
    if (d_u) *d_u = fuu / nrm;
    return ED_OK;
}

// ------------------------------------------------------------ bath-fit distance

struct BathArgs {
    int nc, nb, nw;
    const double *eps, *vbc, *wn, *tgt;
    double p, qpow, shift;
};

// part[q][n]: q = 0 the distance term of frequency n, q = 1 + k its derivative in eps_k,
// q = 1 + nb + k * nc + c its derivative in Vbc(k, c).
__global__ void k_bath_terms(BathArgs A, double *part)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
    if (n >= A.nw) return;
    const int nc = A.nc, nb = A.nb;
    const double w = A.wn[n];
    const double weight = 1.0 / pow(fabs(w - A.wn[0]) + A.shift, A.qpow);
    int qk = -1, qc = -1;
    bool d_eps = false;
    if (q >= 1 && q <= nb) {
        qk = q - 1;
        d_eps = true;
    } else if (q > nb) {
        qk = (q - 1 - nb) / nc;
        qc = (q - 1 - nb) - qk * nc;
    }
    double acc = 0.0;
    for (int i = 0; i < nc; ++i)
        for (int j = 0; j < nc; ++j) {
            double dre = 0.0, dim = 0.0, gre = 0.0, gim = 0.0;
            for (int k = 0; k < nb; ++k) {
                const double vv = A.vbc[k * nc + i] * A.vbc[k * nc + j];
                // 1 / (i w - e) = (-e - i w) / (e^2 + w^2)
                const double e = A.eps[k], den = e * e + w * w;
                const double ire = -e / den, iim = -w / den;
                dre += vv * ire;
                dim += vv * iim;
                if (k == qk) {
                    if (d_eps) {   // d/d eps_k: vv / (i w - e)^2
                        gre += vv * (ire * ire - iim * iim);
                        gim += vv * (2.0 * ire * iim);
                    } else {       // d/d Vbc(k, qc): (delta_i,qc V_kj + V_ki delta_j,qc) / (i w - e)
                        const double dv = (i == qc ? A.vbc[k * nc + j] : 0.0) + (j == qc ? A.vbc[k * nc + i] : 0.0);
                        gre += dv * ire;
                        gim += dv * iim;
                    }
                }
            }
            for (int sg = 0; sg < 2; ++sg) {
                const double *t = A.tgt + ((((int64_t)n * 2 + sg) * nc + i) * nc + j) * 2;
                const double xr = dre - t[0], xi = dim - t[1];
                const double ad = sqrt(xr * xr + xi * xi);
                if (q == 0) {
                    acc += pow(ad, A.p);
                } else if (ad > 0.0) {
                    acc += A.p * pow(ad, A.p - 2.0) * (xr * gre + xi * gim);
                }
            }
        }
    part[(int64_t)q * A.nw + n] = acc * weight;
}

__global__ void k_bath_sum(int nq, int nw, const double *part, double *out)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    double s = 0.0;
    for (int n = 0; n < nw; ++n) s += part[(int64_t)q * nw + n];
    out[q] = s;
}

static int bath_distance(int nc, int nb, const double *eps, const double *vbc, int nw, const double *wn,
                         const double *target, double weight_expo, double fit_weight_power, double fit_shift,
                         double *dist, double *grad_eps, double *grad_v)
{
    const int nq = 1 + nb + nb * nc;
    const size_t n_tgt = (size_t)nw * 2 * nc * nc * 2;
    const size_t n_tot = nb + (size_t)nb * nc + nw + n_tgt + (size_t)nq * nw + nq;
    double *w = nullptr;
    ED_HIP_CHECK(hipMalloc((void **)&w, sizeof(double) * n_tot));
    double *d_eps = w, *d_vbc = d_eps + nb, *d_wn = d_vbc + (size_t)nb * nc, *d_tgt = d_wn + nw, *part = d_tgt + n_tgt,
           *out = part + (size_t)nq * nw;
    std::vector<double> h(nq);
    auto body = [&]() -> int {
        ED_HIP_CHECK(hipMemcpy(d_eps, eps, sizeof(double) * nb, hipMemcpyHostToDevice));
        ED_HIP_CHECK(hipMemcpy(d_vbc, vbc, sizeof(double) * nb * nc, hipMemcpyHostToDevice));
        ED_HIP_CHECK(hipMemcpy(d_wn, wn, sizeof(double) * nw, hipMemcpyHostToDevice));
        ED_HIP_CHECK(hipMemcpy(d_tgt, target, sizeof(double) * n_tgt, hipMemcpyHostToDevice));
        BathArgs A{nc, nb, nw, d_eps, d_vbc, d_wn, d_tgt, weight_expo, fit_weight_power,
                   std::fabs(fit_shift) < 1e-10 ? std::fabs(wn[0]) : fit_shift};
        const int nq_launch = (grad_eps || grad_v) ? nq : 1;
        hipLaunchKernelGGL(k_bath_terms, dim3((nw + 255) / 256, nq_launch), dim3(256), 0, 0, A, part);
        hipLaunchKernelGGL(k_bath_sum, dim3((nq_launch + 63) / 64), dim3(64), 0, 0, nq_launch, nw, (const double *)part,
                           out);
        ED_HIP_CHECK(hipGetLastError());
        ED_HIP_CHECK(hipMemcpy(h.data(), out, sizeof(double) * nq_launch, hipMemcpyDeviceToHost));
        return ED_OK;
    };
    const int rc = body();
    (void)hipFree(w);
    if (rc != ED_OK) return rc;
    const double norm = 1.0 / ((double)nw * (2.0 * nc) * (2.0 * nc));
    const double sgn = h[0] < 0 ? -1.0 : 1.0;
    *dist = std::fabs(h[0]) * norm;
    if (grad_eps)
        for (int k = 0; k < nb; ++k) grad_eps[k] = sgn * h[1 + k] * norm;
    if (grad_v)
        for (int i = 0; i < nb * nc; ++i) grad_v[i] = sgn * h[1 + nb + i] * norm;
    return ED_OK;
}

}  // namespace ed

extern "C" {

int ed_hf_gradient_dev(ed_sector *sec, const double *d_psi, double *d_level, double *d_v, double *d_u)
{
    if (!sec || !d_psi) return ed::fail(ED_ERR_ARG, "ed_hf_gradient_dev: NULL argument");
    if (sec->sz || sec->cplx) return ed::fail(ED_ERR_ARG, "ed_hf_gradient_dev: real (N_up, N_dn) sectors only");
    if (!sec->slater.empty())
        return ed::fail(ED_ERR_ARG, "ed_hf_gradient_dev: the U derivative is of the Kanamori form, not a Slater tensor");
    ED_HIP_CHECK(hipSetDevice(sec->device));
    return ed::hf_gradient(sec, d_psi, d_level, d_v, d_u);
}

int ed_hf_gradient(ed_sector *sec, int64_t n, const double *psi, double *d_level, double *d_v, double *d_u)
{
    if (!sec || !psi) return ed::fail(ED_ERR_ARG, "ed_hf_gradient: NULL argument");
    if (sec->sz || sec->cplx) return ed::fail(ED_ERR_ARG, "ed_hf_gradient: real (N_up, N_dn) sectors only");
    if (!sec->slater.empty())
        return ed::fail(ED_ERR_ARG, "ed_hf_gradient: the U derivative is of the Kanamori form, not a Slater tensor");
    if (n != sec->dim) return ed::fail(ED_ERR_ARG, "ed_hf_gradient: vector length differs from the sector dimension");
    ED_HIP_CHECK(hipSetDevice(sec->device));
    double *d = nullptr;
    ED_HIP_CHECK(hipMalloc((void **)&d, sizeof(double) * n));
    int rc = ED_OK;
    hipError_t e = hipMemcpy(d, psi, sizeof(double) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = ed::fail(ED_ERR_HIP, std::string("ed_hf_gradient: copy-in: ") + hipGetErrorString(e));
    if (rc == ED_OK) rc = ed::hf_gradient(sec, d, d_level, d_v, d_u);
    (void)hipFree(d);
    return rc;
}

int ed_bath_distance(int nc, int nb, const double *eps, const double *vbc, int nw, const double *wn,
                     const double *target, double weight_expo, double fit_weight_power, double fit_shift, double *dist,
                     double *grad_eps, double *grad_v)
{
    if (nc < 1 || nb < 1 || nw < 1 || !eps || !vbc || !wn || !target || !dist)
        return ed::fail(ED_ERR_ARG, "ed_bath_distance: bad argument");
    if (!(weight_expo > 0.0)) return ed::fail(ED_ERR_ARG, "ed_bath_distance: weight_expo must be positive");
    return ed::bath_distance(nc, nb, eps, vbc, nw, wn, target, weight_expo, fit_weight_power, fit_shift, dist, grad_eps,
                             grad_v);
}

}  // extern "C"
