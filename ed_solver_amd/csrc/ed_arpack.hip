// ARPACK_diago (ED_ARPACK.f90:475-570) on the device: the implicitly restarted Lanczos
// method that arpack_eigenvector_sym_matrix (ED_ARPACK.f90:35-250) drives through
// ARPACK's dsaupd / dseupd with bmat = 'I', mode 1 and exact shifts, restated from the
// published method (Sorensen, SIAM J. Matrix Anal. Appl. 13 (1992) 357; Lehoucq and
// Sorensen, SIAM J. Matrix Anal. Appl. 17 (1996) 789; the ARPACK Users' Guide, sections
// 4.2-4.5), with every vector on the device:
//
//   * A V_m = V_m T_m + f e_m^T with V the (ncv + 1)-slot basis (slot m holds f / ||f||);
//     each step is one H.v (the sector's tuned kernel) and a full classical Gram-Schmidt
//     pass repeated once (DGKS-style, always), both passes as FP64 MFMA contractions;
//   * the Ritz values of T_m (host, implicit QL) and their estimates |beta_m s_(m,i)|;
//     the wanted set by `which` ('BE': floor(nev/2) from the low end and the rest from the
//     high end, as dsgets; 'SA', 'LA'); converged when the estimate is below
//     tol max(eps^(2/3), |theta|) (dsconv);
//   * otherwise nev grows by min(nconv, np/2) (dsaup2), the other Ritz values are applied
//     as exact shifts by implicit QR on T_m (host, Givens bulge chases, split at negligible
//     subdiagonals as dsapps), and V_k = V_m Q(:, 1:k), f_k = v_(k+1) T(k+1,k) +
//     f_m Q(m,k) on the device;
//   * at the end the converged wanted Ritz pairs, ascending, X = V_m S on the device, and
//     the residuals ||H x - theta x|| / |theta| that arpack_eigenvector_sym_matrix prints.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ed_kernels.h"

namespace ed {

// y = s x
__global__ void k_arp_scale(const double *x, double s, double *y, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = s * x[i];
}

// z = a x + b y (z may be x or y)
__global__ void k_arp_lin2(double a, const double *x, double b, const double *y, double *z, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        z[i] = a * x[i] + b * y[i];
}

// a fresh direction in (-1, 1) for a basis that broke down: a hash of (row, attempt)
__global__ void k_arp_random(double *x, int64_t n, int attempt)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t h = ((uint64_t)i + 1ull + 0x51ED27ull * (uint64_t)(attempt + 1)) * 0x9E3779B97F4A7C15ull;
        h ^= h >> 31;
        h *= 0xD6E8FEB86659FD93ull;
        h ^= h >> 32;
        x[i] = (double)(h >> 11) / 4503599627370496.0 - 1.0;
    }
}

static int arp_grid(int64_t n)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + kBlock * 4 - 1) / (kBlock * 4), 2048));
}

namespace {

struct Irl {
    ed_sector *s = nullptr;
    hipStream_t st = nullptr;
    int dev = 0;
    int64_t n = 0, ld = 0;
    int m = 0;                 // ncv
    double *V = nullptr;       // (m + 1) slots of ld
    double *W = nullptr;       // (m + 1) slots: the rotated basis of a restart
    double *d_small = nullptr; // 3 (m + 1) + 16 doubles: projections and norms
    std::vector<double> alpha, beta;   // T: alpha[j] = T(j, j), beta[j] = T(j, j - 1) (j >= 1); beta[m] = ||f||
    int64_t matvecs = 0;

    double *slot(int j) const { return V + (int64_t)j * ld; }

    // f (in slot m) -= V(:, 0:p) (V(:, 0:p)^T f) twice; h1 + h2 -> host coefficients,
    // then ||f||^2.  One host wait.
    int orthogonalise(int p, double *coef, double *nrm2)
    {
        double *f = slot(m);
        double *h1 = d_small, *h2 = d_small + (m + 1), *nn = d_small + 2 * (m + 1);
        if (p > 0) {
            ED_TRY(block_gram_dev(dev, n, p, 1, V, ld, f, ld, h1, st));
            block_xc_dev(n, p, 1, V, ld, h1, -1.0, 1.0, f, ld, st);
            ED_TRY(block_gram_dev(dev, n, p, 1, V, ld, f, ld, h2, st));
            block_xc_dev(n, p, 1, V, ld, h2, -1.0, 1.0, f, ld, st);
        }
        ED_TRY(block_gram_dev(dev, n, 1, 1, f, ld, f, ld, nn, st));
        ED_HIP_CHECK(hipGetLastError());
        std::vector<double> h(2 * (m + 1) + 1);
        ED_HIP_CHECK(hipMemcpyAsync(h.data(), d_small, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
        ED_HIP_CHECK(hipStreamSynchronize(st));
        for (int i = 0; i < p; ++i) coef[i] = h[i] + h[(m + 1) + i];
        *nrm2 = h[2 * (m + 1)];
        return ED_OK;
    }

    // Slot j <- f / ||f|| (f in slot m; slot j may be m).  A residual below the
    // breakdown threshold is replaced by a fresh direction orthogonal to V(:, 0:j)
    // (dgetv0's role in dsaitr) and beta_j = 0.
    int set_next(int j, double fnorm, double tnorm, double *beta_out)
    {
        const double eps = 2.220446049250313e-16;
        double *f = slot(m);
        if (fnorm > eps * std::max(1.0, tnorm) * 1e2) {
            hipLaunchKernelGGL(k_arp_scale, dim3(arp_grid(n)), dim3(kBlock), 0, st, (const double *)f, 1.0 / fnorm,
                               slot(j), n);
            ED_HIP_CHECK(hipGetLastError());
            *beta_out = fnorm;
            return ED_OK;
        }
        for (int attempt = 0; attempt < 3; ++attempt) {
            hipLaunchKernelGGL(k_arp_random, dim3(arp_grid(n)), dim3(kBlock), 0, st, f, n, attempt);
            std::vector<double> c(std::max(1, j));
            double r2 = 0.0;
            ED_TRY(orthogonalise(j, c.data(), &r2));
            if (r2 > 1e-20 * (double)n) {
                hipLaunchKernelGGL(k_arp_scale, dim3(arp_grid(n)), dim3(kBlock), 0, st, (const double *)f,
                                   1.0 / std::sqrt(r2), slot(j), n);
                ED_HIP_CHECK(hipGetLastError());
                *beta_out = 0.0;
                return ED_OK;
            }
        }
        return fail(ED_ERR_BREAKDOWN, "ed_arpack: no direction orthogonal to the basis (the sector's dimension is "
                                      "exhausted)");
    }

    double tnorm(int k) const
    {
        double t = 0.0;
        for (int i = 0; i < k; ++i) t = std::max(t, std::fabs(alpha[i]) + std::fabs(beta[i]) + std::fabs(beta[i + 1]));
        return t;
    }

    // Extend A V_k = V_k T_k + f e_k^T to length m (dsaitr).  Slot k holds v_(k+1).
    int extend(int k)
    {
        for (int j = k; j < m; ++j) {
            ED_TRY(launch_hv(s, ED_PREC_F64, slot(j), slot(m), ld, 1, st));
            ++matvecs;
            std::vector<double> c(j + 1);
            double r2 = 0.0;
            ED_TRY(orthogonalise(j + 1, c.data(), &r2));
            alpha[j] = c[j];
            double b = 0.0;
            ED_TRY(set_next(j + 1, std::sqrt(std::max(r2, 0.0)), tnorm(j + 1), &b));
            beta[j + 1] = b;
        }
        return ED_OK;
    }
};

// One implicitly shifted QR step with shift mu on the dense symmetric tridiagonal T
// (m x m, row-major) restricted to its unreduced blocks; Q (m x m, row-major) <- Q G.
void qr_shift(int m, std::vector<double> &T, std::vector<double> &Q, double mu)
{
    const double eps = 2.220446049250313e-16;
    auto at = [&](int i, int j) -> double & { return T[(size_t)i * m + j]; };
    int l = 0;
    while (l < m) {
        int r = l;   // block [l, r]
        while (r + 1 < m) {
            const double tst = std::fabs(at(r, r)) + std::fabs(at(r + 1, r + 1));
            if (std::fabs(at(r + 1, r)) <= eps * tst) {
                at(r + 1, r) = at(r, r + 1) = 0.0;
                break;
            }
            ++r;
        }
        if (r > l) {
            double x = at(l, l) - mu, z = at(l + 1, l);
            for (int i = l; i < r; ++i) {
                const double rr = std::hypot(x, z);
                const double c = rr == 0.0 ? 1.0 : x / rr, sn = rr == 0.0 ? 0.0 : z / rr;
                // T <- G^T T G in the plane (i, i + 1), G = [[c, -s], [s, c]]
                for (int k = 0; k < m; ++k) {   // rows i, i+1
                    const double a = at(i, k), b = at(i + 1, k);
                    at(i, k) = c * a + sn * b;
                    at(i + 1, k) = -sn * a + c * b;
                }
                for (int k = 0; k < m; ++k) {   // columns i, i+1
                    const double a = at(k, i), b = at(k, i + 1);
                    at(k, i) = c * a + sn * b;
                    at(k, i + 1) = -sn * a + c * b;
                }
                for (int k = 0; k < m; ++k) {
                    const double a = Q[(size_t)k * m + i], b = Q[(size_t)k * m + i + 1];
                    Q[(size_t)k * m + i] = c * a + sn * b;
                    Q[(size_t)k * m + i + 1] = -sn * a + c * b;
                }
                if (i + 1 < r) {
                    x = at(i + 1, i);
                    z = at(i + 2, i);
                }
            }
            // clean the fill-in outside the tridiagonal band of the block
            for (int i = l; i <= r; ++i)
                for (int k = l; k <= r; ++k)
                    if (std::abs(i - k) > 1) at(i, k) = 0.0;
        }
        l = r + 1;
    }
}

// Indices (into the ascending Ritz values) of the k wanted ones: 'BE' takes k / 2 from the
// low end and the rest from the high end (dsgets: k = 1 wants the largest), 'SA' the
// smallest, 'LA' the largest.
std::vector<int> wanted_set(int m, int k, int which)
{
    std::vector<int> w;
    if (which == ED_ARPACK_SA) {
        for (int i = 0; i < k; ++i) w.push_back(i);
    } else if (which == ED_ARPACK_LA) {
        for (int i = m - k; i < m; ++i) w.push_back(i);
    } else {
        const int lo = k / 2;
        for (int i = 0; i < lo; ++i) w.push_back(i);
        for (int i = m - (k - lo); i < m; ++i) w.push_back(i);
    }
    return w;
}

}  // namespace

static int arpack_run(ed_sector *sec, const ed_arpack_opts &o, const double *start, double *evals, double *resid,
                      double *d_x, int64_t ldx, int *nconv_out, int *iters_out, int *matvecs_out)
{
    Irl r;
    r.s = sec;
    r.st = sec->stream;
    r.dev = sec->device;
    r.n = sec->dim;
    r.ld = (r.n + 63) / 64 * 64;
    r.m = o.ncv;
    const int m = r.m, nev = o.nev;
    r.alpha.assign(m + 1, 0.0);
    r.beta.assign(m + 2, 0.0);
    ED_TRY(block_ws_reserve(r.dev, r.n, m + 1));
    const size_t bytes = sizeof(double) * (size_t)(m + 1) * r.ld;
    ED_HIP_CHECK(hipMalloc((void **)&r.V, bytes));
    if (hipMalloc((void **)&r.W, bytes) != hipSuccess) {
        (void)hipFree(r.V);
        return fail(ED_ERR_NOMEM, "ed_arpack: basis");
    }
    if (hipMalloc((void **)&r.d_small, sizeof(double) * (3 * (m + 1) + 16)) != hipSuccess) {
        (void)hipFree(r.V);
        (void)hipFree(r.W);
        return fail(ED_ERR_NOMEM, "ed_arpack: work space");
    }
    // zero padding rows once: the contractions read whole 16-row groups
    (void)hipMemsetAsync(r.V, 0, bytes, r.st);
    (void)hipMemsetAsync(r.W, 0, bytes, r.st);
    int rc = ED_OK;
    const double eps = 2.220446049250313e-16, eps23 = std::pow(eps, 2.0 / 3.0);
    const double tol = o.tol > 0.0 ? o.tol : eps;
    int iter = 0, nconv = 0;
    std::vector<double> theta(m), S((size_t)m * m);
    std::vector<int> want;
    do {
        // v_1 = start / ||start|| (the caller's vector, or create_fix_initial_vector's)
        if (start) {
            if (hipMemcpyAsync(r.slot(m), start, sizeof(double) * r.n, hipMemcpyHostToDevice, r.st) != hipSuccess) {
                rc = fail(ED_ERR_HIP, "ed_arpack: start vector");
                break;
            }
        } else if ((rc = start_vector_dev(r.n, r.slot(m), r.st)) != ED_OK) {
            break;
        }
        double r2 = 0.0, dummy = 0.0;
        if ((rc = r.orthogonalise(0, &dummy, &r2)) != ED_OK) break;
        if (!(r2 > 0.0)) {
            rc = fail(ED_ERR_BREAKDOWN, "ed_arpack: zero start vector");
            break;
        }
        double b0 = 0.0;
        if ((rc = r.set_next(0, std::sqrt(r2), 1.0, &b0)) != ED_OK) break;
        int k = 0;
        for (;;) {
            if ((rc = r.extend(k)) != ED_OK) break;
            ++iter;
            // Ritz values and estimates of T_m
            std::vector<double> off(m, 0.0);
            for (int j = 1; j < m; ++j) off[j] = r.beta[j];
            if ((rc = tridiag_eig(m, r.alpha.data(), off.data(), theta.data(), S.data())) != ED_OK) break;
            // S column-major: eigenvector i in S[i*m .. i*m + m - 1]
            const double bm = r.beta[m];
            auto bound = [&](int i) { return std::fabs(bm * S[(size_t)i * m + (m - 1)]); };
            want = wanted_set(m, nev, o.which);
            nconv = 0;
            for (int i : want)
                if (bound(i) <= tol * std::max(eps23, std::fabs(theta[i]))) ++nconv;
            if (nconv >= nev || iter >= o.maxitr) break;
            // grow the kept set as dsaup2 does, the rest are the exact shifts
            int np = m - nev;
            int kev = nev + std::min(nconv, np / 2);
            if (kev == 1 && m >= 6)
                kev = m / 2;
            else if (kev == 1 && m > 2)
                kev = 2;
            kev = std::min(kev, m - 1);
            np = m - kev;
            const std::vector<int> keep = wanted_set(m, kev, o.which);
            std::vector<char> is_keep(m, 0);
            for (int i : keep) is_keep[i] = 1;
            std::vector<int> shifts;
            for (int i = 0; i < m; ++i)
                if (!is_keep[i]) shifts.push_back(i);
            // apply the shifts with the largest Ritz estimates first
            std::stable_sort(shifts.begin(), shifts.end(), [&](int a, int b) { return bound(a) > bound(b); });
            std::vector<double> T((size_t)m * m, 0.0), Q((size_t)m * m, 0.0);
            for (int i = 0; i < m; ++i) {
                T[(size_t)i * m + i] = r.alpha[i];
                Q[(size_t)i * m + i] = 1.0;
                if (i > 0) T[(size_t)i * m + i - 1] = T[(size_t)(i - 1) * m + i] = r.beta[i];
            }
            for (int i : shifts) qr_shift(m, T, Q, theta[i]);
            const double bhat = T[(size_t)kev * m + kev - 1], sigma = Q[(size_t)(m - 1) * m + kev - 1];
            // W(:, 0:kev+1) = V(:, 0:m) Q(:, 0:kev+1), in column chunks of at most 16
            std::vector<double> C((size_t)m * 16);
            double *d_C = nullptr, *d_Cbuf = nullptr;
            if (hipMallocAsync((void **)&d_Cbuf, sizeof(double) * (size_t)m * 16, r.st) != hipSuccess) {
                rc = fail(ED_ERR_NOMEM, "ed_arpack: restart coefficients");
                break;
            }
            d_C = d_Cbuf;
            for (int c0 = 0; c0 <= kev && rc == ED_OK; c0 += 16) {
                const int q = std::min(16, kev + 1 - c0);
                for (int i = 0; i < m; ++i)
                    for (int j = 0; j < q; ++j) C[(size_t)i * q + j] = Q[(size_t)i * m + c0 + j];
                if (hipMemcpyAsync(d_C, C.data(), sizeof(double) * (size_t)m * q, hipMemcpyHostToDevice, r.st) !=
                    hipSuccess) {
                    rc = fail(ED_ERR_HIP, "ed_arpack: restart coefficients");
                    break;
                }
                block_xc_dev(r.n, m, q, r.V, r.ld, d_C, 1.0, 0.0, r.W + (int64_t)c0 * r.ld, r.ld, r.st);
                // the host buffer is rewritten by the next chunk
                if (hipStreamSynchronize(r.st) != hipSuccess) rc = fail(ED_ERR_HIP, "ed_arpack: restart rotation");
            }
            (void)hipFreeAsync(d_Cbuf, r.st);
            if (rc != ED_OK) break;
            // f_k = W(:, kev) bhat + f_m sigma, with f_m = beta_m v_(m+1) (slot m of V)
            hipLaunchKernelGGL(k_arp_lin2, dim3(arp_grid(r.n)), dim3(kBlock), 0, r.st, bhat,
                               (const double *)(r.W + (int64_t)kev * r.ld), bm * sigma, (const double *)r.slot(m),
                               r.W + (int64_t)m * r.ld, r.n);
            std::swap(r.V, r.W);
            for (int i = 0; i < kev; ++i) {
                r.alpha[i] = T[(size_t)i * m + i];
                r.beta[i] = i > 0 ? T[(size_t)i * m + i - 1] : 0.0;
            }
            double f2 = 0.0;
            std::vector<double> cf(kev);
            if ((rc = r.orthogonalise(kev, cf.data(), &f2)) != ED_OK) break;
            double bk = 0.0;
            if ((rc = r.set_next(kev, std::sqrt(std::max(f2, 0.0)), r.tnorm(kev), &bk)) != ED_OK) break;
            r.beta[kev] = bk;
            k = kev;
        }
    } while (false);
    if (rc == ED_OK) {
        // the converged wanted pairs, ascending (dseupd with rvec, HowMny = 'All')
        std::vector<int> sel;
        const double bm = r.beta[m];
        for (int i : want)
            if (std::fabs(bm * S[(size_t)i * m + (m - 1)]) <= tol * std::max(eps23, std::fabs(theta[i])))
                sel.push_back(i);
        std::sort(sel.begin(), sel.end());
        const int nc = (int)sel.size();
        for (int t = 0; t < nc; ++t) evals[t] = theta[sel[t]];
        // X = V(:, 0:m) S(:, sel), in chunks of 16, into W; then residuals and the copy out
        std::vector<double> C((size_t)m * 16);
        double *d_C = nullptr;
        if (nc > 0 && hipMalloc((void **)&d_C, sizeof(double) * (size_t)m * 16) != hipSuccess)
            rc = fail(ED_ERR_NOMEM, "ed_arpack: Ritz vectors");
        for (int c0 = 0; c0 < nc && rc == ED_OK; c0 += 16) {
            const int q = std::min(16, nc - c0);
            for (int i = 0; i < m; ++i)
                for (int j = 0; j < q; ++j) C[(size_t)i * q + j] = S[(size_t)sel[c0 + j] * m + i];
            if (hipMemcpyAsync(d_C, C.data(), sizeof(double) * (size_t)m * q, hipMemcpyHostToDevice, r.st) !=
                    hipSuccess ||
                (block_xc_dev(r.n, m, q, r.V, r.ld, d_C, 1.0, 0.0, r.W + (int64_t)c0 * r.ld, r.ld, r.st),
                 hipStreamSynchronize(r.st) != hipSuccess))
                rc = fail(ED_ERR_HIP, "ed_arpack: Ritz vectors");
        }
        if (d_C) (void)hipFree(d_C);
        // ||H x - theta x|| / |theta| (arpack_eigenvector_sym_matrix, ED_ARPACK.f90:213-219),
        // with W's slot m as the work vector
        for (int t = 0; t < nc && rc == ED_OK; ++t) {
            double *x = r.W + (int64_t)t * r.ld, *hx = r.W + (int64_t)m * r.ld;
            if ((rc = launch_hv(sec, ED_PREC_F64, x, hx, r.ld, 1, r.st)) != ED_OK) break;
            ++r.matvecs;
            hipLaunchKernelGGL(k_arp_lin2, dim3(arp_grid(r.n)), dim3(kBlock), 0, r.st, 1.0, (const double *)hx,
                               -evals[t], (const double *)x, hx, r.n);
            double rr = 0.0;
            if ((rc = block_gram_dev(r.dev, r.n, 1, 1, hx, r.ld, hx, r.ld, r.d_small, r.st)) != ED_OK) break;
            if (hipMemcpyAsync(&rr, r.d_small, sizeof(double), hipMemcpyDeviceToHost, r.st) != hipSuccess ||
                hipStreamSynchronize(r.st) != hipSuccess) {
                rc = fail(ED_ERR_HIP, "ed_arpack: residual");
                break;
            }
            if (resid) resid[t] = std::sqrt(std::max(rr, 0.0)) / std::fabs(evals[t]);
        }
        if (rc == ED_OK && d_x) {   // host eigenvectors [nconv][ldx]
            for (int t = 0; t < nc && rc == ED_OK; ++t)
                if (hipMemcpyAsync(d_x + (int64_t)t * ldx, r.W + (int64_t)t * r.ld, sizeof(double) * (size_t)r.n,
                                   hipMemcpyDeviceToHost, r.st) != hipSuccess)
                    rc = fail(ED_ERR_HIP, "ed_arpack: eigenvector copy");
            if (rc == ED_OK && hipStreamSynchronize(r.st) != hipSuccess) rc = fail(ED_ERR_HIP, "ed_arpack: copy");
        }
        *nconv_out = nc;
    }
    if (iters_out) *iters_out = iter;
    if (matvecs_out) *matvecs_out = (int)std::min<int64_t>(r.matvecs, INT32_MAX);
    (void)hipStreamSynchronize(r.st);
    (void)hipFree(r.V);
    (void)hipFree(r.W);
    (void)hipFree(r.d_small);
    return rc;
}

}  // namespace ed

extern "C" {

int ed_arpack_default_opts(ed_arpack_opts *o)
{
    if (!o) return ed::fail(ED_ERR_ARG, "ed_arpack_default_opts: NULL");
    o->nev = 8;
    o->ncv = 9;
    o->which = ED_ARPACK_BE;
    o->tol = 1e-12;
    o->maxitr = 300;
    return ED_OK;
}

int ed_arpack(ed_sector *sec, const ed_arpack_opts *o, const double *start, double *evals, double *resid,
              double *evecs, int64_t ldv, int *nconv, int *iters, int *matvecs)
{
    using namespace ed;
    if (!sec || !o || !evals || !nconv) return fail(ED_ERR_ARG, "ed_arpack: NULL argument");
    if (sec->cplx) return fail(ED_ERR_ARG, "ed_arpack: real sectors only (a complex H doubles every eigenvalue "
                                           "in the real embedding)");
    if (o->which != ED_ARPACK_BE && o->which != ED_ARPACK_SA && o->which != ED_ARPACK_LA)
        return fail(ED_ERR_ARG, "ed_arpack: which must be ED_ARPACK_BE, _SA or _LA");
    if (o->nev < 1 || o->ncv <= o->nev || (int64_t)o->ncv > sec->dim || o->ncv > 1024 || o->maxitr < 1)
        return fail(ED_ERR_ARG, "ed_arpack: need 1 <= nev < ncv <= min(dimension, 1024) and maxitr >= 1 "
                                "(the _SDRV1 checks of ED_ARPACK.f90:88-100)");
    if (evecs && ldv < sec->dim) return fail(ED_ERR_ARG, "ed_arpack: ldv below the dimension");
    ED_HIP_CHECK(hipSetDevice(sec->device));
    *nconv = 0;
    return arpack_run(sec, *o, start, evals, resid, evecs, ldv, nconv, iters, matvecs);
}

}  // extern "C"
