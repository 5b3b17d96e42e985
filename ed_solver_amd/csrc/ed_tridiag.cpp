// Host FP64 eigen-solver for the symmetric tridiagonal Lanczos matrix T.
//
// Takes the role of LAPACK DSTEDC in diagonalize_tridiag (matrices_tridiag.f90:210-259);
// the method is the implicit QL iteration with Wilkinson shifts (Wilkinson and
// Reinsch, Handbook for Automatic Computation II, "tql1/tql2"; Golub and Van Loan,
// Matrix Computations, sec. 8.3), written here from that description.  T is at most
// a few hundred rows, so this is negligible host work.
//
// Storage follows the Lanczos arrays: diag[0..n-1]; off[i] couples rows i-1 and i for
// i = 1..n-1 (off[0] = ||v_1|| is not referenced), i.e. subdiag of matrices_tridiag.f90.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "ed_internal.h"

namespace ed {

// Implicit QL on (d, e) with e[i] coupling i and i+1.  z (column-major, n x n) is
// rotated along when non-null; on exit column k of z is the eigenvector of d[k].
static int ql_implicit(int n, double *d, double *e, double *z)
{
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= DBL_EPSILON * dd) break;
            }
            if (m != l) {
                if (++iter > 100) return -1;
                // Wilkinson shift from the leading 2x2 block
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i];
                    const double b = c * e[i];
                    r = std::hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {   // deflation inside the sweep
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    if (z) {
                        double *zi = z + (size_t)i * n, *zi1 = z + (size_t)(i + 1) * n;
                        for (int k = 0; k < n; ++k) {
                            f = zi1[k];
                            zi1[k] = s * zi[k] + c * f;
                            zi[k] = c * zi[k] - s * f;
                        }
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    return 0;
}

// All eigenpairs of T, ascending; z[k*n + r] = component r of eigenvector k, with the
// sign chosen so that the largest-magnitude component is positive.
int tridiag_eig(int n, const double *diag, const double *off, double *w, double *z)
{
    if (n < 1) return fail(ED_ERR_ARG, "tridiag_eig: n < 1");
    std::vector<double> d(diag, diag + n), e(n, 0.0), q((size_t)n * n, 0.0);
    for (int i = 0; i + 1 < n; ++i) e[i] = off[i + 1];
    for (int i = 0; i < n; ++i) q[(size_t)i * n + i] = 1.0;
    if (ql_implicit(n, d.data(), e.data(), q.data()) != 0)
        return fail(ED_ERR_ARG, "tridiag_eig: QL iteration did not converge");
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return d[a] < d[b]; });
    for (int k = 0; k < n; ++k) {
        const int src = order[k];
        w[k] = d[src];
        const double *col = q.data() + (size_t)src * n;
        int imax = 0;
        for (int r = 1; r < n; ++r)
            if (std::fabs(col[r]) > std::fabs(col[imax])) imax = r;
        const double sg = col[imax] < 0 ? -1.0 : 1.0;
        for (int r = 0; r < n; ++r) z[(size_t)k * n + r] = sg * col[r];
    }
    return ED_OK;
}

// The k lowest eigenvalues of T (ascending).
void tridiag_lowest(int n, const double *diag, const double *off, int k, double *w)
{
    std::vector<double> d(diag, diag + n), e(n, 0.0);
    for (int i = 0; i + 1 < n; ++i) e[i] = off[i + 1];
    ql_implicit(n, d.data(), e.data(), nullptr);
    std::sort(d.begin(), d.end());
    for (int i = 0; i < k && i < n; ++i) w[i] = d[i];
    for (int i = n; i < k; ++i) w[i] = 1e300;
}

}  // namespace ed

extern "C" {

// Exposed for the parity tests of the host eigen-solver (no reference counterpart as a
// separate symbol: it is the DSTEDC call inside diagonalize_tridiag).
int ed_tridiag_eig(int n, const double *diag, const double *subdiag, double *w, double *z)
{
    return ed::tridiag_eig(n, diag, subdiag, w, z);
}

}  // extern "C"
