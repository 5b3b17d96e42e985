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
#include <cstdint>
#include <string>
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
                if (std::fabs(e[m]) <= DBL_EPSILON * dd || std::fabs(e[m]) <= DBL_MIN) break;
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

static void tri_vectors(int n, const double *d, const double *e, int nev, const double *w, double *z);

// Number of eigenvalues of (d, e) below x (Sturm count of the LDL^T pivots).
static int sturm_count(int n, const double *d, const double *e, double x)
{
    int cnt = 0;
    double q = 1.0;
    for (int i = 0; i < n; ++i) {
        q = d[i] - x - (i ? e[i - 1] * e[i - 1] / q : 0.0);
        if (q == 0.0) q = -DBL_EPSILON * (std::fabs(x) + 1e-300);
        if (q < 0.0) ++cnt;
    }
    return cnt;
}

// All eigenvalues of (d, e) by bisection on the Sturm count (Golub and Van Loan, sec. 8.4):
// the fallback when the QL iteration does not converge.
static void sturm_all(int n, const double *d, const double *e, double *w)
{
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; ++i) {
        const double r = (i ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
        lo = std::min(lo, d[i] - r);
        hi = std::max(hi, d[i] + r);
    }
    const double span = std::max(hi - lo, 1e-300);
    lo -= 1e-12 * span;
    hi += 1e-12 * span;
    for (int k = 0; k < n; ++k) {
        double a = lo, b = hi;
        for (int it = 0; it < 200 && b - a > 2.0 * DBL_EPSILON * std::max(std::fabs(a), std::fabs(b)); ++it) {
            const double mid = 0.5 * (a + b);
            if (sturm_count(n, d, e, mid) > k) b = mid;
            else a = mid;
        }
        w[k] = 0.5 * (a + b);
    }
}

// All eigenpairs of T, ascending; z[k*n + r] = component r of eigenvector k, with the
// sign chosen so that the largest-magnitude component is positive.
int tridiag_eig(int n, const double *diag, const double *off, double *w, double *z)
{
    if (n < 1) return fail(ED_ERR_ARG, "tridiag_eig: n < 1");
    double amax = 0.0;
    for (int i = 0; i < n; ++i) {
        if (!std::isfinite(diag[i]) || (i > 0 && !std::isfinite(off[i])))
            return fail(ED_ERR_ARG, "tridiag_eig: non-finite entry in row " + std::to_string(i) + " of " +
                                        std::to_string(n));
        amax = std::max(amax, std::fabs(diag[i]) + (i > 0 ? std::fabs(off[i]) : 0.0));
    }
    std::vector<double> d(diag, diag + n), e(n, 0.0), q((size_t)n * n, 0.0);
    for (int i = 0; i + 1 < n; ++i) e[i] = off[i + 1];
    for (int i = 0; i < n; ++i) q[(size_t)i * n + i] = 1.0;
    if (ql_implicit(n, d.data(), e.data(), q.data()) != 0) {
        // bisection + inverse iteration (the dstebz / dstein pair) on the original matrix
        std::vector<double> d0(diag, diag + n), e0(n, 0.0);
        for (int i = 0; i + 1 < n; ++i) e0[i] = off[i + 1];
        sturm_all(n, d0.data(), e0.data(), w);
        tri_vectors(n, d0.data(), e0.data(), n, w, z);
        for (int k = 0; k < n; ++k) {   // sign: largest-magnitude component positive
            double *col = z + (size_t)k * n;
            int imax = 0;
            for (int r = 1; r < n; ++r)
                if (std::fabs(col[r]) > std::fabs(col[imax])) imax = r;
            if (col[imax] < 0)
                for (int r = 0; r < n; ++r) col[r] = -col[r];
        }
        return ED_OK;
    }
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

// The k lowest eigenpairs of T: eigenvalues by QL (all, sorted), eigenvectors of the k
// lowest by inverse iteration, z[j*n + r] with the largest component positive.  O(n^2)
// instead of the O(n^3) of accumulating every rotation.
int tridiag_lowest_pairs(int n, const double *diag, const double *off, int k, double *w, double *z)
{
    if (n < 1 || k < 1 || k > n) return fail(ED_ERR_ARG, "tridiag_lowest_pairs: bad argument");
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(diag[i]) || (i > 0 && !std::isfinite(off[i])))
            return fail(ED_ERR_ARG, "tridiag_eig: non-finite entry in row " + std::to_string(i) + " of " +
                                        std::to_string(n));
    tridiag_lowest(n, diag, off, k, w);
    std::vector<double> e(n, 0.0);
    for (int i = 0; i + 1 < n; ++i) e[i] = off[i + 1];
    tri_vectors(n, diag, e.data(), k, w, z);
    for (int j = 0; j < k; ++j) {
        double *col = z + (size_t)j * n;
        int imax = 0;
        for (int r = 1; r < n; ++r)
            if (std::fabs(col[r]) > std::fabs(col[imax])) imax = r;
        if (col[imax] < 0)
            for (int r = 0; r < n; ++r) col[r] = -col[r];
    }
    return ED_OK;
}

// The k lowest eigenvalues of (d, e) by bisection on the Sturm count, each to
// 2 eps |lambda|: O(k n) per call against the O(n^2) of a QL sweep over all of them.
static void sturm_lowest(int n, const double *d, const double *e, int k, double *w)
{
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; ++i) {
        const double r = (i ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
        lo = std::min(lo, d[i] - r);
        hi = std::max(hi, d[i] + r);
    }
    const double span = std::max(hi - lo, 1e-300);
    lo -= 1e-12 * span;
    hi += 1e-12 * span;
    double left = lo;   // eigenvalue j >= eigenvalue j-1: reuse the previous one as the lower end
    for (int j = 0; j < k; ++j) {
        double a = left, b = hi;
        for (int it = 0; it < 200 && b - a > 2.0 * DBL_EPSILON * std::max(std::fabs(a), std::fabs(b)) + 1e-300;
             ++it) {
            const double mid = 0.5 * (a + b);
            if (sturm_count(n, d, e, mid) > j) b = mid;
            else a = mid;
        }
        w[j] = 0.5 * (a + b);
        left = a;
    }
}

// The k lowest eigenvalues of T (ascending).
void tridiag_lowest(int n, const double *diag, const double *off, int k, double *w)
{
    std::vector<double> d(diag, diag + n), e(n, 0.0);
    for (int i = 0; i + 1 < n; ++i) e[i] = off[i + 1];
    if (k <= 32 && k < n) {
        bool finite = true;
        for (int i = 0; i < n; ++i) finite = finite && std::isfinite(d[i]) && std::isfinite(e[i]);
        if (finite) {
            sturm_lowest(n, d.data(), e.data(), k, w);
            return;
        }
    }
    if (ql_implicit(n, d.data(), e.data(), nullptr) != 0) {
        std::vector<double> e0(n, 0.0);
        for (int i = 0; i + 1 < n; ++i) e0[i] = off[i + 1];
        sturm_all(n, diag, e0.data(), d.data());
    }
    std::sort(d.begin(), d.end());
    for (int i = 0; i < k && i < n; ++i) w[i] = d[i];
    for (int i = n; i < k; ++i) w[i] = 1e300;
}

// Pivots q_i of T - x = L D L^T for (d[n], e2[i] = e[i]^2): *cnt = the number of negative
// ones (eigenvalues below x) and *s = d/dx log |det(T - x)| = sum_i q_i'/q_i.  False when a
// pivot vanishes or a value is not finite.
static bool log_det_slope(int n, const double *d, const double *e2, double x, int *cnt, double *s)
{
    double q = d[0] - x, dq = -1.0;
    if (q == 0.0) return false;
    int c = q < 0.0;
    double r = 1.0 / q, acc = dq * r;
    for (int i = 1; i < n; ++i) {
        const double t = e2[i - 1] * r;   // e^2 / q_{i-1}
        dq = -1.0 + t * dq * r;           // q_i' = -1 + e^2 q_{i-1}' / q_{i-1}^2
        q = (d[i] - x) - t;               // q_i = d_i - x - e^2 / q_{i-1}
        if (q == 0.0) return false;
        c += q < 0.0;
        r = 1.0 / q;
        acc += dq * r;
    }
    *cnt = c;
    *s = acc;
    return std::isfinite(acc);
}

static int count_below(int n, const double *d, const double *e2, double x)
{
    int cnt = 0;
    double q = 1.0;
    for (int i = 0; i < n; ++i) {
        q = d[i] - x - (i ? e2[i - 1] / q : 0.0);
        if (q == 0.0) q = -DBL_EPSILON * (std::fabs(x) + 1e-300);
        if (q < 0.0) ++cnt;
    }
    return cnt;
}

// Eigenvalue number j in [a, b] (count_below(a) <= j < count_below(b)) by bisection to
// 2 eps |lambda|.
static double bisect_eig(int n, const double *d, const double *e2, int j, double a, double b)
{
    for (int it = 0; it < 200 && b - a > 2.0 * DBL_EPSILON * std::max(std::fabs(a), std::fabs(b)) + 1e-300; ++it) {
        const double mid = 0.5 * (a + b);
        if (count_below(n, d, e2, mid) > j) b = mid;
        else a = mid;
    }
    return 0.5 * (a + b);
}

// Eigenvalue number j by Newton's method on det(T - x) / prod_{i<j} (x - below[i])
// (Maehly's deflation of the j eigenvalues already found) from an x with exactly j
// eigenvalues below it: the deflated polynomial has real roots, all at or above lambda_j,
// so the iterates increase monotonically to lambda_j (quadratically near a simple root).
// A plain Lanczos run whose Ritz value has converged carries copies of it (lost
// orthogonality), and at a root of multiplicity m Newton's steps shrink only by (m-1)/m:
// when two steps shrink by more than 0.3, the step is multiplied by the m that ratio gives
// (kept only while the point stays below lambda_j).  Stops when the step falls below
// 2 eps |x|; an iterate that passes the root by rounding leaves a bracket of one step,
// which bisection finishes.  False when x does not have j eigenvalues below it, a value
// is not finite or 100 steps do not settle.
static bool newton_eig(int n, const double *d, const double *e2, int j, const double *below, double x, double &w)
{
    int c;
    double s, last = 0.0;
    auto slope = [&](double y, double &g) {
        if (!log_det_slope(n, d, e2, y, &c, &g) || c != j) return false;
        for (int i = 0; i < j; ++i) g -= 1.0 / (y - below[i]);
        return std::isfinite(g) && g < 0.0;
    };
    if (!slope(x, s)) return false;
    for (int it = 0; it < 100; ++it) {
        const double step = -1.0 / s;
        if (!(step > 2.0 * DBL_EPSILON * std::fabs(x) + 1e-300)) {
            w = x + step;
            return std::isfinite(w);
        }
        const double ratio = last > 0.0 ? step / last : 0.0;
        last = step;
        double sn, xn;
        if (ratio > 0.3 && ratio < 1.0) {
            const double m = std::min(8.0, std::floor(1.0 / (1.0 - ratio) + 0.5));
            xn = x + m * step;
            if (m > 1.0 && slope(xn, sn)) {
                x = xn;
                s = sn;
                last = 0.0;   // the next ratio compares plain steps again
                continue;
            }
        }
        xn = x + step;
        if (!slope(xn, sn)) {
            if (count_below(n, d, e2, xn) <= j) return false;   // not finite on the way
            w = bisect_eig(n, d, e2, j, x, xn);
            return true;
        }
        x = xn;
        s = sn;
    }
    return false;
}

// Eigenvalue number j in [a, b] (count_below(a) <= j < count_below(b)), lambda_{j-1} =
// below[j-1] for j > 0: bisection until a has exactly j eigenvalues below it and the
// bracket is small against its distance to lambda_{j-1} (or to the lower end of the
// spectrum `lo`), then Newton's method from a (newton_eig, quadratic from there); bisection
// to the end if that fails.
static double bracketed_eig(int n, const double *d, const double *e2, int j, const double *below, double lo, double a,
                            double b)
{
    const double floor_j = j ? below[j - 1] : lo;
    bool tried = false;
    int ca = count_below(n, d, e2, a);
    for (int it = 0; it < 200 && b - a > 2.0 * DBL_EPSILON * std::max(std::fabs(a), std::fabs(b)) + 1e-300; ++it) {
        if (!tried && ca == j && a > floor_j && b - a < 0.25 * (a - floor_j)) {
            tried = true;
            double w;
            if (newton_eig(n, d, e2, j, below, a, w) && w >= a && w <= b) return w;
        }
        const double mid = 0.5 * (a + b);
        const int cm = count_below(n, d, e2, mid);
        if (cm > j) {
            b = mid;
        } else {
            a = mid;
            ca = cm;
        }
    }
    return 0.5 * (a + b);
}

// The k lowest eigenvalues of T_n (diag[n], off[1..n-1]) when those of T_{n-1} are known
// (prev[j]; 1e300 or more where not), as the Lanczos convergence test needs them every
// step.  Cauchy interlacing gives lambda_j(T_n) <= lambda_j(T_{n-1}): each is found by
// Newton's method (newton_eig) from just below prev[j], the offset grown fourfold until
// the Sturm count confirms the start lies between lambda_{j-1} and lambda_j, else in a
// bracket below prev[j] (bracketed_eig).  A non-finite entry falls back to tridiag_lowest.
void tridiag_lowest_warm(int n, const double *diag, const double *off, int k, const double *prev, double *w)
{
    if (!prev || k > 32 || k >= n || n < 2) {
        tridiag_lowest(n, diag, off, k, w);
        return;
    }
    std::vector<double> d(diag, diag + n), e2(n, 0.0);
    bool finite = true;
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; ++i) {
        const double el = i ? std::fabs(off[i]) : 0.0, er = i + 1 < n ? std::fabs(off[i + 1]) : 0.0;
        finite = finite && std::isfinite(d[i]) && std::isfinite(er);
        lo = std::min(lo, d[i] - el - er);
        hi = std::max(hi, d[i] + el + er);
        if (i + 1 < n) e2[i] = off[i + 1] * off[i + 1];
    }
    if (!finite) {
        tridiag_lowest(n, diag, off, k, w);
        return;
    }
    const double span = std::max(hi - lo, 1e-300);
    lo -= 1e-12 * span;
    hi += 1e-12 * span;
    for (int j = 0; j < k; ++j) {
        const double floor_j = j ? w[j - 1] : lo;   // the start must lie above lambda_{j-1}
        bool ok = false;
        double a = j ? w[j - 1] - 4.0 * DBL_EPSILON * std::fabs(w[j - 1]) - 1e-300 : lo, b = hi;
        if (prev[j] < 1e299) {
            double dl = 1e-7 * std::max(1.0, std::fabs(prev[j]));
            for (int t = 0; t < 24 && !ok; ++t, dl *= 4.0) {
                const double x0 = prev[j] - dl;
                if (x0 <= floor_j) break;
                int c;
                double g;
                if (!log_det_slope(n, d.data(), e2.data(), x0, &c, &g)) continue;
                if (c > j) {   // lambda_j moved below the start: a tighter upper end
                    b = std::min(b, x0);
                    continue;
                }
                if (c < j) break;   // below lambda_{j-1}: bracket it instead
                a = std::max(a, x0);
                ok = newton_eig(n, d.data(), e2.data(), j, w, x0, w[j]);
            }
        }
        if (!ok) {
            if (count_below(n, d.data(), e2.data(), a) > j) a = lo;
            if (b == hi && prev[j] < 1e299) {
                const double bb = prev[j] + 8.0 * DBL_EPSILON * std::fabs(prev[j]) + 1e-300;
                if (bb < hi && count_below(n, d.data(), e2.data(), bb) > j) b = bb;
            }
            w[j] = bracketed_eig(n, d.data(), e2.data(), j, w, lo, a, b);
        }
    }
}

// ---------------------------------------------------------------- band matrices
// Lowest eigenpairs of the symmetric band matrix of a block Lanczos run (half-bandwidth
// b = block size), which takes the role of BLZPACK's Ritz computation inside BLZDRD
// (Block_Lanczos.f90:132,195; the library is not part of the reference sources).
//  1. Reduction to tridiagonal form by Givens rotations that annihilate the band column
//     by column and chase each bulge down the band (Schwarz's method; Golub and Van
//     Loan, Matrix Computations, sec. 8.3): O(n^2 b) operations, rotations logged.
//  2. Eigenvalues of the tridiagonal matrix by implicit QL.
//  3. Eigenvectors of the tridiagonal matrix by inverse iteration with partial pivoting
//     (Wilkinson's method; Gram-Schmidt inside clusters of close eigenvalues), carried
//     back through the logged rotations.

namespace {

using Rot = BandRot;

// Symmetric band storage with room for the bulge: L[c * (b + 2) + (r - c)] = A[r][c] for
// 0 <= r - c <= b + 1 (the band plus the one diagonal a bulge occupies while it is chased).
struct Band {
    int n, b, w;
    std::vector<double> L;
    Band(int n_, int b_) : n(n_), b(b_), w(b_ + 2), L((size_t)n_ * (b_ + 2), 0.0) {}
    bool in(int r, int c) const { return std::abs(r - c) <= b + 1; }
    double get(int r, int c) const
    {
        if (r < c) std::swap(r, c);
        return r - c <= b + 1 ? L[(size_t)c * w + (r - c)] : 0.0;
    }
    void set(int r, int c, double v)
    {
        if (r < c) std::swap(r, c);
        L[(size_t)c * w + (r - c)] = v;
    }
};

// A (symmetric band, half-bandwidth b) -> tridiagonal by similarity rotations G on the
// planes (p, p+1), G = [[c, s], [-s, c]]; the rotations are appended to `log` in order when
// it is not null (only the eigenvectors need them).  A rotation mixes rows (and columns) p
// and q = p + 1 at the columns x within b + 1 of either; the 2 x 2 block transforms as
// G B G^T.  Storage: L[c w + (r - c)] = A[r][c], 0 <= r - c <= b + 1.
void band_to_tridiag(Band &A, std::vector<Rot> *log)
{
    const int n = A.n, b = A.b;
    double *__restrict__ L = A.L.data();
    const int w = A.w;
    auto at = [&](int r, int c) -> double & { return L[(size_t)c * w + (r - c)]; };   // r >= c
    for (int k = 0; k + 2 < n; ++k) {
        for (int i = std::min(n - 1, k + b); i >= k + 2; --i) {
            int q = i, col = k;
            while (true) {
                const int p = q - 1;
                const double x = at(p, col), y = at(q, col);
                if (y == 0.0) break;
                const double r = std::sqrt(x * x + y * y), c = x / r, s = y / r;
                // rows p, q at the columns left of p (column x holds A[p][x], A[q][x] adjacent)
                const int lo = std::max(0, q - b - 1);
                for (int xx = lo; xx < p; ++xx) {
                    double *colx = L + (size_t)xx * w;
                    const double t1 = colx[p - xx], t2 = colx[q - xx];
                    colx[p - xx] = c * t1 + s * t2;
                    colx[q - xx] = -s * t1 + c * t2;
                }
                // columns p, q below q (two contiguous runs)
                const int hi = std::min(n - 1, p + b + 1);
                double *colp = L + (size_t)p * w + 1, *colq = L + (size_t)q * w;   // colp[x - q] = A[x][p]
                for (int d = 1; d <= hi - q; ++d) {
                    const double t1 = colp[d], t2 = colq[d];
                    colp[d] = c * t1 + s * t2;
                    colq[d] = -s * t1 + c * t2;
                }
                const double app = at(p, p), apq = at(q, p), aqq = at(q, q);
                at(p, p) = c * c * app + 2.0 * c * s * apq + s * s * aqq;
                at(q, q) = s * s * app - 2.0 * c * s * apq + c * c * aqq;
                at(q, p) = -c * s * app + (c * c - s * s) * apq + c * s * aqq;
                at(q, col) = 0.0;
                if (log) log->push_back({p, c, s});
                // the rotation filled (q + b, p): chase it one band further down
                if (q + b >= n) break;
                col = p;
                q = q + b;
            }
        }
    }
}

// Solves (T - lam) x = rhs in place for the tridiagonal (d, e), e[i] coupling i and i+1,
// by Gaussian elimination with partial pivoting; zero pivots are replaced by `tiny`.
void tri_shift_solve(int n, const double *d, const double *e, double lam, double tiny, double *x)
{
    std::vector<double> dl(n, 0.0), dd(n), du(n, 0.0), du2(n, 0.0);
    std::vector<char> swp(n, 0);
    for (int i = 0; i < n; ++i) dd[i] = d[i] - lam;
    for (int i = 0; i + 1 < n; ++i) {
        dl[i] = e[i];
        du[i] = e[i];
    }
    for (int i = 0; i + 1 < n; ++i) {
        if (std::fabs(dd[i]) >= std::fabs(dl[i])) {
            if (dd[i] == 0.0) dd[i] = tiny;
            const double f = dl[i] / dd[i];
            dl[i] = f;
            dd[i + 1] -= f * du[i];
        } else {   // interchange rows i and i+1
            const double f = dd[i] / dl[i];
            dd[i] = dl[i];
            dl[i] = f;
            const double t = du[i];
            du[i] = dd[i + 1];
            dd[i + 1] = t - f * dd[i + 1];
            if (i + 2 < n) {
                du2[i] = du[i + 1];
                du[i + 1] = -f * du[i + 1];
            }
            swp[i] = 1;
        }
    }
    if (dd[n - 1] == 0.0) dd[n - 1] = tiny;
    for (int i = 0; i + 1 < n; ++i) {
        if (!swp[i]) {
            x[i + 1] -= dl[i] * x[i];
        } else {
            const double t = x[i];
            x[i] = x[i + 1];
            x[i + 1] = t - dl[i] * x[i];
        }
    }
    x[n - 1] /= dd[n - 1];
    if (n > 1) x[n - 2] = (x[n - 2] - du[n - 2] * x[n - 1]) / dd[n - 2];
    for (int i = n - 3; i >= 0; --i) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / dd[i];
}

}  // namespace

// Eigenvectors z (n x nev, column-major) of the tridiagonal (d, e), e[i] coupling i and
// i+1, for the ascending eigenvalues w[0..nev): inverse iteration with partial pivoting,
// Gram-Schmidt inside clusters of close eigenvalues (Wilkinson; LAPACK dstein's scheme).
static void tri_vectors(int n, const double *d, const double *e, int nev, const double *w, double *z)
{
    double tnorm = 0.0;
    for (int i = 0; i < n; ++i)
        tnorm = std::max(tnorm, std::fabs(d[i]) + (i + 1 < n ? std::fabs(e[i]) : 0.0) + (i ? std::fabs(e[i - 1]) : 0.0));
    const double eps = DBL_EPSILON, tiny = eps * std::max(tnorm, 1e-300), ortol = 1e-3 * tnorm;
    int cluster0 = 0;
    for (int k = 0; k < nev; ++k) {
        double *x = z + (size_t)k * n;
        if (k > 0 && w[k] - w[k - 1] > ortol) cluster0 = k;
        // a perturbed copy of a cluster's eigenvalue keeps the factorisation regular
        const double lam = w[k] + (k - cluster0) * 10.0 * eps * std::max(tnorm, 1.0);
        uint64_t h = 0x2545F4914F6CDD1Dull + (uint64_t)k * 0x9E3779B97F4A7C15ull;
        for (int i = 0; i < n; ++i) {   // deterministic start vector in (-1, 1)
            h ^= h << 13;
            h ^= h >> 7;
            h ^= h << 17;
            x[i] = (double)(h >> 11) / 4503599627370496.0 - 1.0;
        }
        for (int it = 0; it < 4; ++it) {
            tri_shift_solve(n, d, e, lam, tiny, x);
            for (int c = cluster0; c < k; ++c) {   // Gram-Schmidt against the cluster
                const double *v = z + (size_t)c * n;
                double dot = 0.0;
                for (int i = 0; i < n; ++i) dot += v[i] * x[i];
                for (int i = 0; i < n; ++i) x[i] -= dot * v[i];
            }
            double nrm = 0.0;
            for (int i = 0; i < n; ++i) nrm += x[i] * x[i];
            nrm = std::sqrt(nrm);
            for (int i = 0; i < n; ++i) x[i] /= nrm;
        }
    }
}

int band_reduce(int n, int b, const double *a, bool keep_log, BandTri &t)
{
    if (n < 1 || b < 0 || !a) return fail(ED_ERR_ARG, "band_reduce: bad argument");
    Band A(n, std::max(b, 1));
    for (int c = 0; c < n; ++c)
        for (int r = c; r < std::min(n, c + std::max(b, 1) + 1); ++r) A.set(r, c, a[(size_t)r * n + c]);
    t.n = n;
    t.log.clear();
    if (keep_log && b > 1) t.log.reserve((size_t)n * n / 2);
    if (b > 1) band_to_tridiag(A, keep_log ? &t.log : nullptr);
    t.d.assign(n, 0.0);
    t.e.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
        t.d[i] = A.get(i, i);
        if (i + 1 < n) t.e[i] = A.get(i + 1, i);
        if (!std::isfinite(t.d[i]) || !std::isfinite(t.e[i])) return fail(ED_ERR_ARG, "band_lowest: non-finite entry");
    }
    return ED_OK;
}

int band_tri_values(const BandTri &t, int nev, double *w)
{
    const int n = t.n;
    if (nev < 1 || nev > n || !w) return fail(ED_ERR_ARG, "band_lowest: bad argument");
    if (nev <= 32 && nev < n) {
        sturm_lowest(n, t.d.data(), t.e.data(), nev, w);
    } else {
        std::vector<double> dq(t.d), eq(t.e), ds(t.d), es(t.e);
        if (ql_implicit(n, dq.data(), eq.data(), nullptr) != 0) sturm_all(n, ds.data(), es.data(), dq.data());
        std::sort(dq.begin(), dq.end());
        for (int k = 0; k < nev; ++k) w[k] = dq[k];
    }
    return ED_OK;
}

int band_tri_vectors(const BandTri &t, int nev, const double *w, double *z)
{
    const int n = t.n;
    if (nev < 1 || nev > n || !w || !z) return fail(ED_ERR_ARG, "band_lowest: bad argument");
    tri_vectors(n, t.d.data(), t.e.data(), nev, w, z);
    // back through the rotations: x = G_1^T ... G_N^T y
    for (size_t r = t.log.size(); r-- > 0;) {
        const Rot &g = t.log[r];
        for (int k = 0; k < nev; ++k) {
            double *x = z + (size_t)k * n;
            const double yp = x[g.p], yq = x[g.p + 1];
            x[g.p] = g.c * yp - g.s * yq;
            x[g.p + 1] = g.s * yp + g.c * yq;
        }
    }
    for (int k = 0; k < nev; ++k) {   // sign: largest-magnitude component positive
        double *x = z + (size_t)k * n;
        int imax = 0;
        for (int i = 1; i < n; ++i)
            if (std::fabs(x[i]) > std::fabs(x[imax])) imax = i;
        if (x[imax] < 0)
            for (int i = 0; i < n; ++i) x[i] = -x[i];
    }
    return ED_OK;
}

int band_lowest(int n, int b, const double *a, int nev, double *w, double *z)
{
    if (n < 1 || b < 0 || nev < 1 || nev > n || !a || !w) return fail(ED_ERR_ARG, "band_lowest: bad argument");
    BandTri t;
    ED_TRY(band_reduce(n, b, a, z != nullptr, t));
    ED_TRY(band_tri_values(t, nev, w));
    return z ? band_tri_vectors(t, nev, w, z) : ED_OK;
}

}  // namespace ed

extern "C" {

// Lowest eigenpairs of a symmetric band matrix (the Ritz step of the block Lanczos
// driver, exposed for its host-side tests).
int ed_band_eig(int n, int b, const double *a, int nev, double *w, double *z)
{
    return ed::band_lowest(n, b, a, nev, w, z);
}

// Exposed for the parity tests of the host eigen-solver (no reference counterpart as a
// separate symbol: it is the DSTEDC call inside diagonalize_tridiag).
int ed_tridiag_eig(int n, const double *diag, const double *subdiag, double *w, double *z)
{
    return ed::tridiag_eig(n, diag, subdiag, w, z);
}

int ed_tridiag_lowest(int n, const double *diag, const double *subdiag, int k, const double *prev, double *w)
{
    if (n < 1 || k < 1 || !diag || !subdiag || !w) return ed::fail(ED_ERR_ARG, "ed_tridiag_lowest: bad argument");
    ed::tridiag_lowest_warm(n, diag, subdiag, k, prev, w);
    return ED_OK;
}

}  // extern "C"
