// Block Lanczos (Block_Lanczos_diagonalize, Block_Lanczos.f90:37-355) on the device, and
// the FP64 MFMA tall-skinny contractions its orthogonalisation is built from.
//
// The reference hands H.U products to BLZPACK (BLZDRD, reverse communication: LFLAG = 1
// asks for V = H U on NVOPU <= Block_size vectors, LFLAG = 4 for start vectors, LFLAG = 0
// returns converged eigenpairs VALP/VECP, Block_Lanczos.f90:189-259).  BLZPACK is not
// part of the reference sources; this driver is a block Lanczos of its own, written from
// the textbook method (Golub and Underwood 1977; Golub and Van Loan, Matrix
// Computations, sec. 10.3.6) with full reorthogonalisation:
//
//   Q_1 R_0 = start block (n x b, Cholesky-QR2)
//   for j = 1, 2, ...:
//     W  = H Q_j                                   block H.v (b vectors per launch)
//     W -= Q_{1..j} (Q_{1..j}^T W), twice          classical Gram-Schmidt, CGS2
//     A_j = the Q_j rows of the two projections     diagonal block of T
//     Q_{j+1} B_{j+1} = W                           Cholesky-QR2; B_{j+1} upper triangular
//     T = block tridiagonal (A_j, B_j), half-bandwidth b; its lowest Ritz pairs from the
//     band eigen-solver on the host; stop when every wanted Ritz value has moved by less
//     than `tolerance` since the previous test and its residual bound
//     ||B_{j+1} s_k(last b rows)|| is below sqrt(tolerance) * max(1, |theta_k|), or when
//     the block steps or the Krylov space run out.
//   Ritz vectors X = Q_{1..m} S, residuals ||H x_k - theta_k x_k|| on the device.
//
// Contractions (FP64):
//   Gram    G = X^T Y   (X n x p, Y n x q, q <= 16) on the matrix cores
//           (v_mfma_f64_16x16x4f64, one wave per 16 x 16 output tile): the reduction over
//           n is split in row blocks; each workgroup writes its partial tile sums, a second
//           kernel adds them in row-block order (deterministic);
//   update  Y = beta Y + alpha X C  (C p x q): one thread per row on the vector unit, so
//           that every load and store of a wave is 512 contiguous bytes of one column.
// Both are HBM-bound (about q/4 flop per byte).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ed_kernels.h"

namespace ed {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// D(16x16) += A(16x4) B(4x16) in FP64 on the matrix cores.  Lane l (0..63) supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; it holds D[4 * r + l / 16][l % 16] in
// element r of the accumulator (layout checked against NumPy by the GPU tests).
__device__ inline dbl4 mfma_f64_16x16x4(double a, double b, dbl4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int kGramTiles = 4;    // 16-column tiles of X per Gram workgroup
// Row blocks of a Gram product: with one tile group (p <= 16 kGramTiles) the grid is the
// row blocks alone, and 1024 of them put four workgroups on every CU (one wave per SIMD is
// too few to keep a streaming reduction at HBM speed); wider products have tile groups
// enough with 256.
constexpr int kGramRowBlocks = 256, kGramRowBlocksNarrow = 1024;

// 16 consecutive rows [r, r + 16) of column `col` of a column-major matrix, zero
// outside [0, n) or when !ok.
__device__ inline void load_rows16(const double *m, int64_t ld, bool ok, int col, int64_t r, int64_t n, double *v)
{
    if (ok && r + 16 <= n) {
        const double2 *p = reinterpret_cast<const double2 *>(m + (int64_t)col * ld + r);
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const double2 t = p[s];
            v[2 * s] = t.x;
            v[2 * s + 1] = t.y;
        }
    } else {
#pragma unroll
        for (int s = 0; s < 16; ++s) v[s] = (ok && r + s < n) ? m[(int64_t)col * ld + r + s] : 0.0;
    }
}

// Partial Gram sums: workgroup (tile group, row block); part[rb][p16][16].  Row block rb
// covers rows [rb * rows, (rb + 1) * rows), rows a multiple of 256; wave w takes the
// 64-row chunks w, w + 4, ...; lane group g (l / 16) takes rows 16 g .. 16 g + 15 of a
// chunk, one row per MFMA step, so each lane reads 128 contiguous bytes of a column.
__global__ __launch_bounds__(256) void k_gram_part(int64_t n, int p, int q, const double *X, int64_t ldx,
                                                   const double *Y, int64_t ldy, int64_t rows, int p16, double *part)
{
    __shared__ double red[4][kGramTiles][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, i16 = lane & 15, g = lane >> 4;
    const int ntile = p16 / 16, t0 = blockIdx.x * kGramTiles;
    const int64_t r_begin = (int64_t)blockIdx.y * rows, r_end = min(n, r_begin + rows);
    dbl4 acc[kGramTiles];
#pragma unroll
    for (int t = 0; t < kGramTiles; ++t) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    // X^T X (the Cholesky-QR Gram): the X tile is the Y tile, loaded once
    const bool same = X == Y && ldx == ldy && p == q;
    for (int64_t c0 = r_begin + 64 * w; c0 < r_end; c0 += 256) {
        const int64_t r = c0 + 16 * g;
        double yv[16];
        load_rows16(Y, ldy, i16 < q, i16, r, n, yv);
#pragma unroll
        for (int t = 0; t < kGramTiles; ++t) {
            const int tile = t0 + t;
            if (tile >= ntile) break;
            const int xc = tile * 16 + i16;
            double xv[16];
            if (same) {
#pragma unroll
                for (int s = 0; s < 16; ++s) xv[s] = yv[s];
            } else {
                load_rows16(X, ldx, xc < p, xc, r, n, xv);
            }
#pragma unroll
            for (int s = 0; s < 16; ++s) acc[t] = mfma_f64_16x16x4(xv[s], yv[s], acc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < kGramTiles; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) red[w][t][lane * 4 + r] = acc[t][r];
    __syncthreads();
    const int e = threadIdx.x, el = e >> 2, er = e & 3;
    const int i = 4 * er + (el >> 4), jj = el & 15;
    for (int t = 0; t < kGramTiles; ++t) {
        const int tile = t0 + t;
        if (tile >= ntile) break;
        const double s = ((red[0][t][e] + red[1][t][e]) + red[2][t][e]) + red[3][t][e];
        part[((int64_t)blockIdx.y * p16 + tile * 16 + i) * 16 + jj] = s;
    }
}

// G[i][j] (row-major p x q) = sum over the nrb row blocks of the partial tiles.  One
// workgroup per row i of G: thread t takes column j = t % 16 and the row blocks
// b = t / 16, t / 16 + 16, ... (each load reads one contiguous 128-byte row of a tile
// per 16 threads), then the 16 partial sums of a column are added in group order: a
// fixed order, so every call on the same data gives the same bits.
__global__ __launch_bounds__(256) void k_gram_sum(int nrb, int p, int q, int p16, const double *part, double *G)
{
    __shared__ double red[256];
    const int i = blockIdx.x, t = threadIdx.x, j = t & 15, grp = t >> 4;
    // eight independent partial sums per thread (row blocks grp + 16 (8 u + v), v = 0..7),
    // so that eight loads are in flight, added in a fixed order
    double acc[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) acc[v] = 0.0;
    for (int b0 = grp; b0 < nrb; b0 += 128) {
        double x[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) {   // unconditional loads (clamped index), so none waits on another
            const int b = min(b0 + 16 * v, nrb - 1);
            x[v] = part[((int64_t)b * p16 + i) * 16 + j];
        }
#pragma unroll
        for (int v = 0; v < 8; ++v) acc[v] += b0 + 16 * v < nrb ? x[v] : 0.0;
    }
    const double s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    red[t] = s;
    __syncthreads();
    if (t < q) {
        double acc = 0.0;
        for (int g = 0; g < 16; ++g) acc += red[g * 16 + t];
        G[(int64_t)i * q + t] = acc;
    }
}

// Y[:, 0:q] = beta Y + alpha X[:, 0:p] C (C row-major p x q, q <= Q), one thread per row
// on the vector unit: the product is HBM-bound (2q flop per 8-byte element of X, well below
// the FP64 FMA rate), so what matters is that every load and store of a wave covers 64
// consecutive rows of one column (512 contiguous bytes).  C passes through LDS in chunks
// of 64 of its rows, zero-padded to Q columns; every thread reads the same LDS word
// (broadcast).  Y may alias X when p == q <= 64: a thread reads its row of X before it
// writes that row of Y.
template <int Q>
__global__ __launch_bounds__(256) void k_xc_rows(int64_t n, int p, int q, const double *X, int64_t ldx,
                                                 const double *C, double alpha, double beta, double *Y, int64_t ldy)
{
    __shared__ double sc[64 * Q];
    const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool ok = row < n;
    double acc[Q];
#pragma unroll
    for (int c = 0; c < Q; ++c) acc[c] = 0.0;
    for (int k0 = 0; k0 < p; k0 += 64) {
        const int kn = min(64, p - k0);
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * Q; e += 256) {
            const int k = e / Q, c = e % Q;
            sc[e] = (k < kn && c < q) ? C[(int64_t)(k0 + k) * q + c] : 0.0;
        }
        __syncthreads();
        if (!ok) continue;
        const double *x = X + (int64_t)k0 * ldx + row;
        int k = 0;
        for (; k + 4 <= kn; k += 4) {
            const double x0 = x[(int64_t)k * ldx], x1 = x[(int64_t)(k + 1) * ldx];
            const double x2 = x[(int64_t)(k + 2) * ldx], x3 = x[(int64_t)(k + 3) * ldx];
#pragma unroll
            for (int c = 0; c < Q; ++c) {
                double t = acc[c];
                t = __builtin_fma(x0, sc[k * Q + c], t);
                t = __builtin_fma(x1, sc[(k + 1) * Q + c], t);
                t = __builtin_fma(x2, sc[(k + 2) * Q + c], t);
                acc[c] = __builtin_fma(x3, sc[(k + 3) * Q + c], t);
            }
        }
        for (; k < kn; ++k) {
            const double x0 = x[(int64_t)k * ldx];
#pragma unroll
            for (int c = 0; c < Q; ++c) acc[c] = __builtin_fma(x0, sc[k * Q + c], acc[c]);
        }
    }
    if (!ok) return;
#pragma unroll
    for (int c = 0; c < Q; ++c) {
        if (c >= q) break;
        double *y = Y + (int64_t)c * ldy + row;
        const double v = alpha * acc[c];
        *y = beta == 0.0 ? v : __builtin_fma(beta, *y, v);
    }
}

// Deterministic start block in (-1, 1): the hash of (row mod 10000, column).
__global__ void k_block_start(int64_t n, int b, double *Q, int64_t ld)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < b; ++c) {
            uint64_t x = ((uint64_t)(i % 10000) + 1ull + 10007ull * (uint64_t)c) * 0x9E3779B97F4A7C15ull;
            x ^= x >> 29;
            x *= 0xBF58476D1CE4E5B9ull;
            x ^= x >> 32;
            Q[(int64_t)c * ld + i] = (double)(x >> 11) / 4503599627370496.0 - 1.0;
        }
}

// Cholesky factor of a b x b Gram matrix G (row-major, b <= 16) on the device: one
// workgroup, thread t owning entry (t / 16, t % 16).  Right-looking elimination writes
// R (upper triangular, G = R^T R) and Ri = R^{-1} by back substitution, and *ok = 1; when
// G is not numerically positive definite relative to its largest diagonal entry (a pivot
// <= 1e-26 max_i G_ii, or none positive) it writes R = 0, Ri = I and *ok = 0, so that
// W Ri leaves W as it is.  With a_out set it also copies the b x b blocks a1 and a2 to
// a_out[0, b^2) and a_out[b^2, 2 b^2), for the host to read with the factors in one copy.
__global__ __launch_bounds__(256) void k_chol_upper(int b, const double *G, double *R, double *Ri, double *ok,
                                                    const double *a1, const double *a2, double *a_out)
{
    __shared__ double A[16][17], Rs[16][17], Is[16][17];
    const int t = threadIdx.x, i = t >> 4, j = t & 15;
    const bool in = i < b && j < b;
    A[i][j] = in ? G[i * b + j] : 0.0;
    Rs[i][j] = 0.0;
    Is[i][j] = 0.0;
    if (a_out && t < b * b) {
        a_out[t] = a1[t];
        a_out[b * b + t] = a2[t];
    }
    __syncthreads();
    double gmax = 0.0;
    for (int k = 0; k < b; ++k) gmax = fmax(gmax, A[k][k]);
    bool good = gmax > 0.0;
    for (int k = 0; k < b && good; ++k) {   // every thread reads the same pivot: uniform exit
        const double piv = A[k][k];
        if (!(piv > 1e-26 * gmax)) {
            good = false;
            break;
        }
        const double rkk = sqrt(piv);
        if (i == k && j >= k && j < b) Rs[k][j] = j == k ? rkk : A[k][j] / rkk;
        __syncthreads();
        if (i > k && j >= i && j < b) A[i][j] -= Rs[k][i] * Rs[k][j];
        __syncthreads();
    }
    if (good)
        for (int r = b - 1; r >= 0; --r) {
            if (i == r && j >= r && j < b) {
                double v = j == r ? 1.0 : 0.0;
                for (int k = r + 1; k <= j; ++k) v -= Rs[r][k] * Is[k][j];
                Is[r][j] = v / Rs[r][r];
            }
            __syncthreads();
        }
    if (in) {
        R[i * b + j] = good ? Rs[i][j] : 0.0;
        Ri[i * b + j] = good ? Is[i][j] : (i == j ? 1.0 : 0.0);
    }
    if (t == 0) *ok = good ? 1.0 : 0.0;
}

// ---------------------------------------------------------------- host side

// Work space of the contractions: Gram partials and the small matrices, per device,
// grown on demand (the block driver and the C-ABI contraction calls share it).
struct BlockWS {
    double *part = nullptr, *small = nullptr;
    size_t part_n = 0, small_n = 0;
    int device = -1;
};
static BlockWS g_bws[16];

static int bws_get(int dev, size_t part_n, size_t small_n, BlockWS *&w)
{
    if (dev < 0 || dev >= 16) return fail(ED_ERR_ARG, "device index out of range");
    w = &g_bws[dev];
    if (w->part_n < part_n) {
        if (w->part) (void)hipFree(w->part);
        w->part = nullptr;
        w->part_n = 0;
        ED_HIP_CHECK(hipMalloc((void **)&w->part, sizeof(double) * part_n));
        w->part_n = part_n;
    }
    if (w->small_n < small_n) {
        if (w->small) (void)hipFree(w->small);
        w->small = nullptr;
        w->small_n = 0;
        ED_HIP_CHECK(hipMalloc((void **)&w->small, sizeof(double) * small_n));
        w->small_n = small_n;
    }
    return ED_OK;
}

static int gram_rows(int64_t n, int p)
{
    const int nrb = p <= 16 * kGramTiles ? kGramRowBlocksNarrow : kGramRowBlocks;
    const int64_t per = (n + nrb - 1) / nrb;
    return (int)std::max<int64_t>(256, (per + 255) / 256 * 256);
}

// Partial-sum workspace of the Gram products with up to p columns of X over n rows.
static size_t gram_part_size(int64_t n, int p)
{
    auto size = [n](int pc) {
        const int p16 = (pc + 15) / 16 * 16;
        const int64_t rows = gram_rows(n, pc);
        return (size_t)((n + rows - 1) / rows) * p16 * 16;
    };
    return std::max(size(p), size(std::min(p, 16 * kGramTiles)));
}

// d_G (device, row-major p x q) = X^T Y, queued on st (no host wait).  The partial-sum
// workspace must not be reallocated while a queued Gram still uses it: callers that
// queue several size it first (bws_get with the largest p).
static int gram_dev(int dev, int64_t n, int p, int q, const double *X, int64_t ldx, const double *Y, int64_t ldy,
                    double *d_G, hipStream_t st)
{
    const int p16 = (p + 15) / 16 * 16;
    const int64_t rows = gram_rows(n, p);
    const int nrb = (int)((n + rows - 1) / rows);
    BlockWS *w;
    ED_TRY(bws_get(dev, gram_part_size(n, p), 256, w));
    const dim3 grid((p16 / 16 + kGramTiles - 1) / kGramTiles, nrb);
    hipLaunchKernelGGL(k_gram_part, grid, dim3(256), 0, st, n, p, q, X, ldx, Y, ldy, rows, p16, w->part);
    hipLaunchKernelGGL(k_gram_sum, dim3(p), dim3(256), 0, st, nrb, p, q, p16,
                       (const double *)w->part, d_G);
    ED_HIP_CHECK(hipGetLastError());
    return ED_OK;
}

// G (host, row-major p x q) = X^T Y.
static int gram(int dev, int64_t n, int p, int q, const double *X, int64_t ldx, const double *Y, int64_t ldy,
                double *G, hipStream_t st)
{
    BlockWS *w;
    ED_TRY(bws_get(dev, gram_part_size(n, p), (size_t)p * q + 256, w));
    ED_TRY(gram_dev(dev, n, p, q, X, ldx, Y, ldy, w->small, st));
    ED_HIP_CHECK(hipMemcpyAsync(G, w->small, sizeof(double) * p * q, hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipStreamSynchronize(st));
    return ED_OK;
}

// Y = beta Y + alpha X C with C on the device (row-major p x q), queued on st.
static void xc_dev(int64_t n, int p, int q, const double *X, int64_t ldx, const double *d_C, double alpha,
                   double beta, double *Y, int64_t ldy, hipStream_t st)
{
    const dim3 grid((unsigned)((n + 255) / 256));
    if (q <= 1)
        hipLaunchKernelGGL(k_xc_rows<1>, grid, dim3(256), 0, st, n, p, q, X, ldx, d_C, alpha, beta, Y, ldy);
    else if (q <= 2)
        hipLaunchKernelGGL(k_xc_rows<2>, grid, dim3(256), 0, st, n, p, q, X, ldx, d_C, alpha, beta, Y, ldy);
    else if (q <= 4)
        hipLaunchKernelGGL(k_xc_rows<4>, grid, dim3(256), 0, st, n, p, q, X, ldx, d_C, alpha, beta, Y, ldy);
    else if (q <= 8)
        hipLaunchKernelGGL(k_xc_rows<8>, grid, dim3(256), 0, st, n, p, q, X, ldx, d_C, alpha, beta, Y, ldy);
    else
        hipLaunchKernelGGL(k_xc_rows<16>, grid, dim3(256), 0, st, n, p, q, X, ldx, d_C, alpha, beta, Y, ldy);
}

// Y = beta Y + alpha X C with C given on the host (row-major p x q).
static int xc(int dev, int64_t n, int p, int q, const double *X, int64_t ldx, const double *C, double alpha,
              double beta, double *Y, int64_t ldy, hipStream_t st)
{
    BlockWS *w;
    ED_TRY(bws_get(dev, 16, (size_t)p * q + 256, w));
    ED_HIP_CHECK(hipMemcpyAsync(w->small, C, sizeof(double) * p * q, hipMemcpyHostToDevice, st));
    xc_dev(n, p, q, X, ldx, w->small, alpha, beta, Y, ldy, st);
    ED_HIP_CHECK(hipGetLastError());
    // the small buffer is reused by the next call: finish this one first
    ED_HIP_CHECK(hipStreamSynchronize(st));
    return ED_OK;
}

// The contractions for other drivers (the implicitly restarted Lanczos of ed_arpack):
// reserve the partial-sum work space for Gram products of up to pmax columns over n rows
// before queueing several of them, then queue them without host waits.
int block_ws_reserve(int dev, int64_t n, int pmax)
{
    BlockWS *w;
    return bws_get(dev, gram_part_size(n, pmax), 256, w);
}

int block_gram_dev(int dev, int64_t n, int p, int q, const double *X, int64_t ldx, const double *Y, int64_t ldy,
                   double *d_G, hipStream_t st)
{
    return gram_dev(dev, n, p, q, X, ldx, Y, ldy, d_G, st);
}

void block_xc_dev(int64_t n, int p, int q, const double *X, int64_t ldx, const double *d_C, double alpha, double beta,
                  double *Y, int64_t ldy, hipStream_t st)
{
    xc_dev(n, p, q, X, ldx, d_C, alpha, beta, Y, ldy, st);
}

// Small device buffers of the Cholesky-QR2 below: the Gram matrix, the inverse factor, and
// the read-back block [R1 | R2 | A1 | A2 | ok1 | ok2] (b x b each, two flags).
struct CholBufs {
    double *G, *Ri, *out;
    static size_t out_size(int b) { return 4 * (size_t)b * b + 2; }
};

// In place W <- W R1^{-1} R2^{-1} with W^T W = R1^T R1, then (W R1^{-1})^T (W R1^{-1}) =
// R2^T R2 (Cholesky-QR, twice), queued on st with no host wait.  The factors and flags
// land in c.out (R1, R2, ok1, ok2; W Rtot with Rtot = R2 R1 is the original block); with
// a1, a2 set the last factor kernel copies those b x b projection blocks to
// c.out + 2 b^2.  A block that is numerically rank deficient gives ok1 or ok2 = 0 (W is
// then left as the identity factor found it).
static int chol_qr2_dev(int dev, int64_t n, int b, double *W, int64_t ld, const CholBufs &c, const double *a1,
                        const double *a2, hipStream_t st)
{
    const size_t bb = (size_t)b * b;
    for (int pass = 0; pass < 2; ++pass) {
        ED_TRY(gram_dev(dev, n, b, b, W, ld, W, ld, c.G, st));
        const bool last = pass == 1 && a1;
        hipLaunchKernelGGL(k_chol_upper, dim3(1), dim3(256), 0, st, b, (const double *)c.G, c.out + pass * bb, c.Ri,
                           c.out + 4 * bb + pass, last ? a1 : nullptr, last ? a2 : nullptr,
                           last ? c.out + 2 * bb : nullptr);
        xc_dev(n, b, b, W, ld, c.Ri, 1.0, 0.0, W, ld, st);
        ED_HIP_CHECK(hipGetLastError());
    }
    return ED_OK;
}

// Rtot = R2 R1 (upper triangular, row-major) from a read-back block of chol_qr2_dev;
// false when either factorisation failed.
static bool chol_result(int b, const double *h, std::vector<double> &Rtot)
{
    const size_t bb = (size_t)b * b;
    if (!(h[4 * bb] == 1.0 && h[4 * bb + 1] == 1.0)) return false;
    const double *R1 = h, *R2 = h + bb;
    Rtot.assign(bb, 0.0);
    for (int i = 0; i < b; ++i)
        for (int j = i; j < b; ++j) {
            double s = 0.0;
            for (int k = i; k <= j; ++k) s += R2[(size_t)i * b + k] * R1[(size_t)k * b + j];
            Rtot[(size_t)i * b + j] = s;
        }
    return true;
}

int launch_hv(ed_sector *s, int prec, const void *x, void *y, int64_t ld, int nvec, hipStream_t st);

struct BlockResult {
    std::vector<double> evals, resid;
    int steps = 0, nconv = 0, full_passes = 0;
};

// The block omega-recurrence of partial reorthogonalisation (Simon's recurrence for the
// loss of orthogonality of Lanczos vectors, in the block form used by block Lanczos codes
// such as BLZPACK).  With H Q_j = Q_{j+1} B_{j+1} + Q_j A_j + Q_{j-1} B_j^T and
// Om_{k,i} = Q_k^T Q_i (b x b), the symmetry of H gives for k < j
//   Om_{k,j+1} B_{j+1} = B_{k+1}^T Om_{k+1,j} + A_k Om_{k,j} + B_k Om_{k-1,j}
//                        - Om_{k,j} A_j - Om_{k,j-1} B_j^T,
// with Om_{j,j} = I, Om_{-1,.} = 0.  Rounding seeds the recurrence: every new estimate gets
// eps (||T|| estimate) ||B_{j+1}^{-1}|| in every entry, with alternating signs, and a block
// just projected against everything restarts at that level.  cur holds Om_{k,j} (k < j),
// old Om_{k,j-1}; A_k, B_k are read from T (B_k = T block (k, k-1), upper triangular).
struct BlockOmega {
    int b;
    std::vector<std::vector<double>> cur, old;
    double tnorm = 0.0;
    explicit BlockOmega(int b_) : b(b_) {}
    static double fro(const std::vector<double> &m)
    {
        double t = 0.0;
        for (double x : m) t += x * x;
        return std::sqrt(t);
    }
    double seed = 1e-16;
    // Om_{k,jn}, k < jn, back to the rounding level: the block was just projected against
    // every block (Om_{k,jn-1} is kept)
    void reset(int jn)
    {
        cur.assign(jn, std::vector<double>((size_t)b * b, 0.0));
        for (auto &m : cur)
            for (int e = 0; e < b * b; ++e) m[e] = (e & 1) ? -seed : seed;
    }
    // Om_{k,j+1} for k < j from the recurrence; Om_{j,j+1} at the seed level (the step
    // projected the block against Q_j).  Returns whether some |Om_{k,j+1}| > sqrt(eps).
    bool advance(int j, int bs, int nmax, const std::vector<double> &T, const std::vector<double> &Bn, bool was_full)
    {
        const double eps = 2.220446049250313e-16;
        auto blk = [&](int bi, int bj, std::vector<double> &out) {   // T block (bi, bj)
            out.resize((size_t)bs * bs);
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < bs; ++c) out[(size_t)r * bs + c] = T[(size_t)(bi * bs + r) * nmax + bj * bs + c];
        };
        auto mul = [&](const std::vector<double> &x, const std::vector<double> &y, bool tx) {
            std::vector<double> z((size_t)bs * bs, 0.0);   // z = op(x) y, op = transpose when tx
            for (int r = 0; r < bs; ++r)
                for (int k = 0; k < bs; ++k) {
                    const double a = tx ? x[(size_t)k * bs + r] : x[(size_t)r * bs + k];
                    if (a == 0.0) continue;
                    for (int c = 0; c < bs; ++c) z[(size_t)r * bs + c] += a * y[(size_t)k * bs + c];
                }
            return z;
        };
        auto mulT = [&](const std::vector<double> &x, const std::vector<double> &y) {   // x y^T
            std::vector<double> z((size_t)bs * bs, 0.0);
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < bs; ++c) {
                    double t = 0.0;
                    for (int k = 0; k < bs; ++k) t += x[(size_t)r * bs + k] * y[(size_t)c * bs + k];
                    z[(size_t)r * bs + c] = t;
                }
            return z;
        };
        std::vector<double> Aj, Bj, Ak, Bk, Bk1, I((size_t)bs * bs, 0.0);
        for (int r = 0; r < bs; ++r) I[(size_t)r * bs + r] = 1.0;
        blk(j, j, Aj);
        if (j > 0) blk(j, j - 1, Bj);
        // ||T|| estimate and ||B_{j+1}^{-1}|| (B upper triangular, back substitution on I)
        tnorm = std::max(tnorm, fro(Aj) + (j > 0 ? fro(Bj) : 0.0) + fro(Bn));
        std::vector<double> Binv((size_t)bs * bs, 0.0);
        for (int c = 0; c < bs; ++c)
            for (int r = bs - 1; r >= 0; --r) {
                double t = (r == c) ? 1.0 : 0.0;
                for (int k = r + 1; k < bs; ++k) t -= Bn[(size_t)r * bs + k] * Binv[(size_t)k * bs + c];
                Binv[(size_t)r * bs + c] = t / Bn[(size_t)r * bs + r];
            }
        seed = eps * tnorm * fro(Binv);
        std::vector<std::vector<double>> nxt(j + 1);
        bool over = false;
        if (!was_full)
            for (int k = 0; k < j; ++k) {
                blk(k, k, Ak);
                blk(k + 1, k, Bk1);                   // B_{k+1}
                const std::vector<double> &Okj = cur[k];
                const std::vector<double> &Ok1j = (k + 1 == j) ? I : cur[k + 1];
                std::vector<double> M = mul(Bk1, Ok1j, true);   // B_{k+1}^T Om_{k+1,j}
                std::vector<double> t = mul(Ak, Okj, false);
                for (size_t e = 0; e < M.size(); ++e) M[e] += t[e];
                if (k > 0) {
                    blk(k, k - 1, Bk);                // B_k
                    t = mul(Bk, cur[k - 1], false);
                    for (size_t e = 0; e < M.size(); ++e) M[e] += t[e];
                }
                t = mul(Okj, Aj, false);
                for (size_t e = 0; e < M.size(); ++e) M[e] -= t[e];
                if (j > 0) {
                    const std::vector<double> &Okjm = (k == j - 1) ? I : old[k];
                    t = mulT(Okjm, Bj);              // Om_{k,j-1} B_j^T
                    for (size_t e = 0; e < M.size(); ++e) M[e] -= t[e];
                }
                std::vector<double> O = mul(M, Binv, false);
                for (int e = 0; e < bs * bs; ++e) {
                    O[e] += (e & 1) ? -seed : seed;
                    if (std::fabs(O[e]) > std::sqrt(eps)) over = true;
                }
                nxt[k] = std::move(O);
            }
        else
            for (int k = 0; k < j; ++k) {
                nxt[k].assign((size_t)bs * bs, 0.0);
                for (int e = 0; e < bs * bs; ++e) nxt[k][e] = (e & 1) ? -seed : seed;
            }
        nxt[j].assign((size_t)bs * bs, 0.0);
        for (int e = 0; e < bs * bs; ++e) nxt[j][e] = (e & 1) ? -seed : seed;
        old = cur;
        cur = std::move(nxt);
        return over;
    }
};

// Ritz tests of the block driver on worker threads.  A test after `step` block steps
// reduces the leading m x m block of T (m = step * b, row-major with leading dimension nmax)
// to tridiagonal form, logging the rotations, and finds its nev lowest eigenvalues.  Later
// block steps write T only outside that block, so a test may run while the driver goes on.
constexpr int kRitzWorkers = 4;

// Worker threads of the Ritz tests: kRitzWorkers, or ED_BLOCK_RITZ_THREADS (0..16) from the
// environment; with 0 every test runs on the driving thread when the schedule reaches it
// (the same tests in the same order, so the same results).
static int ritz_threads()
{
    const char *e = std::getenv("ED_BLOCK_RITZ_THREADS");
    if (!e || !*e) return kRitzWorkers;
    return std::max(0, std::min(16, std::atoi(e)));
}

struct RitzJob {
    int step = 0, rc = ED_OK;
    bool ready = false;
    std::string err;
    std::vector<double> w;
    BandTri bt;
    double ms = 0.0;
};

class RitzPool {
  public:
    RitzPool(const double *T, int nmax, int b, int nev, int nthreads) : T_(T), nmax_(nmax), b_(b), nev_(nev)
    {
        for (int i = 0; i < nthreads; ++i) {
            try {
                th_.emplace_back([this] { run(); });
            } catch (...) {
                break;   // fewer threads; with none, get() runs the tests itself
            }
        }
    }
    ~RitzPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            queue_.clear();
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    // queue the test after `step` block steps, unless it is queued already
    void submit(int step)
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (jobs_.count(step)) return;
        auto jb = std::make_shared<RitzJob>();
        jb->step = step;
        jobs_[step] = jb;
        queue_.push_back(jb);
        cv_.notify_one();
    }
    // tests queued or running
    int busy()
    {
        std::lock_guard<std::mutex> lk(mu_);
        int k = 0;
        for (auto &e : jobs_) k += e.second->ready ? 0 : 1;
        return k;
    }
    // the finished test after `step` steps, or null (not queued, or with wait == false not
    // finished); with wait, blocks until it has finished
    std::shared_ptr<RitzJob> get(int step, bool wait)
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = jobs_.find(step);
        if (it == jobs_.end()) return nullptr;
        auto jb = it->second;
        if (wait && !jb->ready && th_.empty()) {   // no workers: run it here
            for (auto q = queue_.begin(); q != queue_.end(); ++q)
                if (*q == jb) {
                    queue_.erase(q);
                    break;
                }
            lk.unlock();
            const int rc = compute(*jb);
            lk.lock();
            jb->rc = rc;
            jb->ready = true;
        }
        if (wait) done_.wait(lk, [&] { return jb->ready; });
        return jb->ready ? jb : nullptr;
    }
    // forget the tests at or below `step`, and those queued below `keep` that have not
    // started (off the schedule now); a running one finishes on its own
    void forget(int step, int keep)
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto it = jobs_.begin(); it != jobs_.end() && it->first <= step;) it = jobs_.erase(it);
        for (auto q = queue_.begin(); q != queue_.end();) {
            if ((*q)->step < keep) {
                jobs_.erase((*q)->step);
                q = queue_.erase(q);
            } else {
                ++q;
            }
        }
    }

  private:
    int compute(RitzJob &jb)
    {
        const auto t0 = std::chrono::steady_clock::now();
        const int m = jb.step * b_;
        std::vector<double> Tm((size_t)m * m);
        for (int r = 0; r < m; ++r) std::memcpy(&Tm[(size_t)r * m], T_ + (size_t)r * nmax_, sizeof(double) * m);
        jb.w.assign(nev_, 0.0);
        int rc = band_reduce(m, b_, Tm.data(), true, jb.bt);
        if (rc == ED_OK) rc = band_tri_values(jb.bt, nev_, jb.w.data());
        if (rc != ED_OK) jb.err = ed_last_error();
        jb.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return rc;
    }
    void run()
    {
        for (;;) {
            std::shared_ptr<RitzJob> jb;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
                if (stop_) return;
                jb = queue_.front();
                queue_.pop_front();
            }
            const int rc = compute(*jb);
            {
                std::lock_guard<std::mutex> lk(mu_);
                jb->rc = rc;
                jb->ready = true;
            }
            done_.notify_all();
        }
    }
    const double *T_;
    const int nmax_, b_, nev_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    std::deque<std::shared_ptr<RitzJob>> queue_;
    std::map<int, std::shared_ptr<RitzJob>> jobs_;
    std::vector<std::thread> th_;
    bool stop_ = false;
};

// The block Lanczos driver; Ritz vectors go to d_out (column-major n x nev, leading
// dimension ldo) when d_out is not null.
// the last block Lanczos call (ed_block_last_stats): steps, full projections, and host
// wall time of the whole call and of its Ritz tests (band solver)
static int g_last_steps = 0, g_last_full = 0;
static double g_last_ms = 0.0, g_last_ritz_ms = 0.0;

static int block_lanczos(ed_sector *s, const ed_block_opts &o, double *d_out, int64_t ldo, BlockResult &res)
{
    const auto t_start = std::chrono::steady_clock::now();
    double ritz_ms = 0.0;
    const int64_t n = s->dim;
    const int b = o.block_size, nev = o.nev, dev = s->device;
    hipStream_t st = s->stream;
    if (!s->cpb) {   // H.v schedule: the autotuner runs with the first Lanczos call
        int choice;
        double ms[kSchedules];
        ED_TRY(ed_hmult_tune(s, &choice, ms));
    }
    const int64_t ld = (n + 63) / 64 * 64;
    const int maxsteps = (int)std::max<int64_t>(1, std::min<int64_t>(o.max_steps, (n + b - 1) / b));
    if ((int64_t)maxsteps * b < nev) return fail(ED_ERR_ARG, "block Lanczos: max_steps * block_size < nev");
    const size_t nvec = (size_t)(maxsteps + 1) * b;
    const size_t bb = (size_t)b * b, nout = CholBufs::out_size(b);
    double *Q = nullptr, *HX = nullptr, *dG = nullptr, *hb = nullptr;
    std::string err;
    {
        hipError_t e = hipMalloc((void **)&Q, sizeof(double) * ld * nvec);
        if (e != hipSuccess)
            return fail(ED_ERR_NOMEM, "block Lanczos basis of " + std::to_string(nvec) + " vectors: " + hipGetErrorString(e));
        // projections of the two CGS passes, then the Cholesky-QR2 buffers: the Gram matrix,
        // the inverse factor, and two read-back blocks (block steps; start block and
        // re-projections)
        e = hipMalloc((void **)&dG, sizeof(double) * (2 * nvec * b + 2 * bb + 2 * nout));
        if (e == hipSuccess) e = hipHostMalloc((void **)&hb, sizeof(double) * 2 * nout, hipHostMallocDefault);
        if (e != hipSuccess) {
            (void)hipFree(Q);
            if (dG) (void)hipFree(dG);
            return fail(ED_ERR_NOMEM, std::string("block Lanczos projections: ") + hipGetErrorString(e));
        }
    }
    struct Guard {
        double *&a, *&b, *&c, *&h;
        hipStream_t st;
        ~Guard()
        {
            (void)hipStreamSynchronize(st);   // queued steps may still use the buffers
            if (a) (void)hipFree(a);
            if (b) (void)hipFree(b);
            if (c) (void)hipFree(c);
            if (h) (void)hipHostFree(h);
        }
    } guard{Q, HX, dG, hb, st};
    {   // size the Gram workspace once for the widest product, so queued Grams never move it
        BlockWS *w;
        ED_TRY(bws_get(dev, gram_part_size(n, (int)nvec), 256, w));
    }
    double *dG1 = dG, *dG2 = dG + nvec * b;
    const CholBufs cb_step{dG + 2 * nvec * b, dG + 2 * nvec * b + bb, dG + 2 * nvec * b + 2 * bb};
    const CholBufs cb_aux{cb_step.G, cb_step.Ri, cb_step.out + nout};
    double *hb_step = hb, *hb_aux = hb + nout;
    auto read_back = [&](const CholBufs &c, double *h) -> int {
        ED_HIP_CHECK(hipMemcpyAsync(h, c.out, sizeof(double) * nout, hipMemcpyDeviceToHost, st));
        ED_HIP_CHECK(hipStreamSynchronize(st));
        return ED_OK;
    };
    hipLaunchKernelGGL(k_block_start, dim3(1024), dim3(256), 0, st, n, b, Q, ld);
    ED_HIP_CHECK(hipGetLastError());
    std::vector<double> R0;
    ED_TRY(chol_qr2_dev(dev, n, b, Q, ld, cb_aux, nullptr, nullptr, st));
    ED_TRY(read_back(cb_aux, hb_aux));
    if (!chol_result(b, hb_aux, R0)) return fail(ED_ERR_BREAKDOWN, "block Lanczos: start block is rank deficient");

    const int nmax = maxsteps * b;
    std::vector<double> T((size_t)nmax * nmax, 0.0);
    auto Tat = [&](int i, int j) -> double & { return T[(size_t)i * nmax + j]; };
    std::vector<double> Bnext, prev(nev, 1e300), w(nev), Z;
    int steps = 0;
    bool exhausted = false;
    // per completed step j: B_{j+1} (for the residual bounds of a Ritz test after j + 1
    // steps) and the full projections so far
    std::vector<std::vector<double>> Bhist;
    std::vector<int> full_hist;
    // Partial reorthogonalisation (o.reorth == 1): the estimates Om[k] ~ Q_k^T Q_j (b x b,
    // k < j) of the block omega-recurrence; a step projects its new block against Q_{j-1}
    // and Q_j only, unless an estimate exceeds sqrt(eps) (then, and at the step after, against
    // every block).  o.reorth == 0: against every block at every step (full CGS2).
    const bool partial = o.reorth == 1;
    BlockOmega om(b);
    int full_due = 0;
    // One block step on the device, queued with no host wait: W = H Q_j, CGS2 against
    // Q_{c0..j}, Cholesky-QR2 of W (not at the last step), and the read-back of A_j's
    // projections and the factors into hb_step.  The host takes the step's results after
    // one stream wait, and issues the next step before its own work on them (the omega
    // recurrence, the Ritz test), so that the device runs step j + 1 meanwhile.  When the
    // recurrence then asks for a full projection of Q_{j+1}, that projection is queued
    // behind the step already issued (which read Q_{j+1} first) and the step is issued
    // again: the results are those of the sequential order.
    struct Step {
        int j = 0, c0 = 0;
        bool full = false, last = false;
    };
    auto issue = [&](int j, Step &sp) -> int {
        const int p = (j + 1) * b;   // columns of Q_{1..j+1}
        sp.j = j;
        sp.full = !partial || full_due > 0 || j < 2;
        sp.c0 = sp.full ? 0 : (j - 1) * b;   // first column the new block is projected against
        sp.last = (j + 1 == maxsteps) || (int64_t)p + b > n;
        const int pw = p - sp.c0;
        double *Qj = Q + (size_t)j * b * ld, *W = Q + (size_t)(j + 1) * b * ld, *Qc = Q + (size_t)sp.c0 * ld;
        ED_TRY(launch_hv(s, ED_PREC_F64, Qj, W, ld, b, st));
        ED_TRY(gram_dev(dev, n, pw, b, Qc, ld, W, ld, dG1, st));
        xc_dev(n, pw, b, Qc, ld, dG1, -1.0, 1.0, W, ld, st);
        ED_TRY(gram_dev(dev, n, pw, b, Qc, ld, W, ld, dG2, st));
        xc_dev(n, pw, b, Qc, ld, dG2, -1.0, 1.0, W, ld, st);
        ED_HIP_CHECK(hipGetLastError());
        const double *a1 = dG1 + (size_t)(j * b - sp.c0) * b, *a2 = dG2 + (size_t)(j * b - sp.c0) * b;
        if (sp.last) {
            ED_HIP_CHECK(hipMemcpyAsync(hb_step + 2 * bb, a1, sizeof(double) * bb, hipMemcpyDeviceToHost, st));
            ED_HIP_CHECK(hipMemcpyAsync(hb_step + 3 * bb, a2, sizeof(double) * bb, hipMemcpyDeviceToHost, st));
        } else {
            ED_TRY(chol_qr2_dev(dev, n, b, W, ld, cb_step, a1, a2, st));
            ED_HIP_CHECK(hipMemcpyAsync(hb_step, cb_step.out, sizeof(double) * nout, hipMemcpyDeviceToHost, st));
        }
        return ED_OK;
    };
    // Ritz tests.  Their schedule is that of a sequential loop: the first after
    // max(check_every, nev / b) steps; then after max(check_every, steps / 4) more steps while
    // the wanted values still move by more than sqrt(tolerance) between tests (the band solver
    // costs O(m^2 b) per test), after check_every steps once they move less (so convergence is
    // seen within a few steps of the test that would see it); and at the last step.  The
    // values come from one band reduction with its rotations logged, the vectors (residual
    // bounds, Ritz vectors) from the same reduction once every wanted value has moved by less
    // than `tolerance`, or at the last step.  Converged: that, and every residual bound
    // ||B_{j+1} s_k(last b rows)|| below sqrt(tolerance) max(1, |theta_k|).
    //
    // A test reads only the leading block of T, so it runs on a worker thread as soon as its
    // steps are done while the device runs further steps; once the values move by less than
    // sqrt(tolerance), the next tests of the check_every cadence are started as well.  The
    // results are taken in schedule order (a test off the schedule is never used), so the
    // solve stops at the same step with the same values and vectors as the sequential loop;
    // the device runs at most `lookahead` steps past the next scheduled test, into blocks
    // outside the basis that is used.
    const int every = std::max(1, o.check_every);
    int sched = std::max(every, (nev + b - 1) / b);   // steps of the next scheduled test
    bool cadence = false;   // the last test moved by < sqrt(tolerance): tests every `every` steps
    int L = -1;             // steps at the last step, once reached
    int avail = 0;          // completed steps
    bool finished = false;
    RitzPool pool(T.data(), nmax, b, nev, ritz_threads());
    auto eff = [&]() { return (L >= 0 && sched >= L) ? L : sched; };
    auto allow = [&](int jn) {   // may step jn (0-based) be issued now
        const int lookahead = std::max(every * kRitzWorkers, sched / 4) + every;
        return jn + 1 <= sched + lookahead;
    };
    auto dispatch = [&]() {
        const int e = eff();
        if (e > avail || e * b < nev) return;
        pool.submit(e);
        if (cadence)
            for (int k = 1; k < kRitzWorkers; ++k) {
                const int s2 = e + k * every;
                if (s2 > avail || (L >= 0 && s2 > L) || pool.busy() >= kRitzWorkers) break;
                pool.submit(s2);
            }
    };
    // take the finished tests in schedule order (wait: block on the next one)
    auto process = [&](bool wait) -> int {
        for (;;) {
            const int e = eff();
            if (e > avail) return ED_OK;
            if (e * b < nev) {   // the last step came before the first test: no Ritz pairs
                finished = true;
                steps = e;
                return ED_OK;
            }
            dispatch();
            auto jb = pool.get(e, wait);
            if (!jb) return ED_OK;
            wait = false;
            if (jb->rc != ED_OK) return fail(jb->rc, jb->err);
            const bool is_last = e == L;
            bool settled = e > 1;
            double moved = 0.0;
            for (int k = 0; k < nev; ++k) {
                const double dk = std::fabs(jb->w[k] - prev[k]);
                if (!(dk < o.tolerance)) settled = false;
                moved = std::max(moved, std::isfinite(dk) ? dk : 1e300);
            }
            w = jb->w;
            ritz_ms += jb->ms;
            const int m = e * b;
            if (is_last || settled) {
                const auto t_v = std::chrono::steady_clock::now();
                Z.assign((size_t)m * nev, 0.0);
                ED_TRY(band_tri_vectors(jb->bt, nev, w.data(), Z.data()));
                ritz_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_v).count();
            }
            if (is_last) {
                finished = true;
                steps = e;
                return ED_OK;
            }
            bool conv = settled;
            const std::vector<double> &lastB = Bhist[e - 1];
            for (int k = 0; k < nev && conv; ++k) {
                // ||H x_k - theta_k x_k|| = ||B_{j+1} s_k(last b rows)||
                double rr = 0.0;
                for (int r = 0; r < b; ++r) {
                    double t = 0.0;
                    for (int c = r; c < b; ++c) t += lastB[(size_t)r * b + c] * Z[(size_t)k * m + (m - b + c)];
                    rr += t * t;
                }
                const double rnorm = std::sqrt(rr);
                if (!(std::fabs(w[k] - prev[k]) < o.tolerance &&
                      rnorm < std::sqrt(o.tolerance) * std::max(1.0, std::fabs(w[k]))))
                    conv = false;
            }
            for (int k = 0; k < nev; ++k) prev[k] = w[k];
            if (conv) {
                finished = true;
                steps = e;
                return ED_OK;
            }
            sched = e + (moved < std::sqrt(o.tolerance) ? every : std::max(every, e / 4));
            cadence = moved < std::sqrt(o.tolerance);
            pool.forget(e, sched);
        }
    };
    Step cur, nxt;
    ED_TRY(issue(0, cur));
    for (;;) {
        ED_HIP_CHECK(hipStreamSynchronize(st));
        const int j = cur.j, p = (j + 1) * b;
        if (cur.full) {
            ++res.full_passes;
            if (full_due > 0) --full_due;
        }
        {   // A_j, symmetrised: the Q_j rows of the two projections
            const double *a1 = hb_step + 2 * bb, *a2 = hb_step + 3 * bb;
            for (int r = 0; r < b; ++r)
                for (int c = 0; c < b; ++c) {
                    const double x = a1[(size_t)r * b + c] + a2[(size_t)r * b + c];
                    const double y = a1[(size_t)c * b + r] + a2[(size_t)c * b + r];
                    Tat(j * b + r, j * b + c) = 0.5 * (x + y);
                }
        }
        bool last = cur.last;
        if (!last && !chol_result(b, hb_step, Bnext)) {
            exhausted = true;   // the Krylov space is invariant: T is exact on it
            last = true;
        }
        const bool ahead = !last && allow(j + 1);
        if (ahead) ED_TRY(issue(j + 1, nxt));
        if (!last && partial) {
            // omega-recurrence for Q_k^T Q_{j+1}, k < j; above sqrt(eps): project the new
            // block against every block now (and the next block as well)
            const bool over = om.advance(j, b, nmax, T, Bnext, cur.full);
            if (over && !cur.full) {
                double *W = Q + (size_t)(j + 1) * b * ld;
                ED_TRY(gram_dev(dev, n, p, b, Q, ld, W, ld, dG1, st));
                xc_dev(n, p, b, Q, ld, dG1, -1.0, 1.0, W, ld, st);
                ED_TRY(gram_dev(dev, n, p, b, Q, ld, W, ld, dG2, st));
                xc_dev(n, p, b, Q, ld, dG2, -1.0, 1.0, W, ld, st);
                ED_TRY(chol_qr2_dev(dev, n, b, W, ld, cb_aux, nullptr, nullptr, st));
                ED_TRY(read_back(cb_aux, hb_aux));
                std::vector<double> R2;
                if (!chol_result(b, hb_aux, R2)) {
                    exhausted = true;
                    last = true;
                } else {
                    // Q_{j+1} B_{j+1} = (Q' R2) B_{j+1}: the new B is R2 B (upper triangular)
                    std::vector<double> Bn(bb, 0.0);
                    for (int r = 0; r < b; ++r)
                        for (int c = r; c < b; ++c) {
                            double t = 0.0;
                            for (int k = r; k <= c; ++k) t += R2[(size_t)r * b + k] * Bnext[(size_t)k * b + c];
                            Bn[(size_t)r * b + c] = t;
                        }
                    Bnext = Bn;
                    om.reset(j + 1);
                }
                ++res.full_passes;
                full_due = 1;
                // the step issued above used the unprojected block
                if (!last && ahead) ED_TRY(issue(j + 1, nxt));
            }
        }
        if (!last)
            for (int r = 0; r < b; ++r)
                for (int c = 0; c < b; ++c) {
                    Tat((j + 1) * b + r, j * b + c) = Bnext[(size_t)r * b + c];
                    Tat(j * b + c, (j + 1) * b + r) = Bnext[(size_t)r * b + c];
                }
        Bhist.push_back(last ? std::vector<double>() : Bnext);
        full_hist.push_back(res.full_passes);
        avail = j + 1;
        if (last) L = avail;
        ED_TRY(process(false));
        // with no step in flight: wait for tests until one may be issued or the solve ends
        // (a step issued before the last one turned out last is dropped)
        bool have_next = ahead && !last;
        while (!finished && !have_next) {
            if (L < 0 && allow(j + 1)) {
                ED_TRY(issue(j + 1, nxt));
                have_next = true;
            } else {
                ED_TRY(process(true));
            }
        }
        if (finished) break;
        cur = nxt;
    }
    res.full_passes = full_hist[steps - 1];
    const int m = steps * b;
    (void)exhausted;
    if (Z.size() < (size_t)m * nev)
        return fail(ED_ERR_BREAKDOWN, "block Lanczos: the Krylov space (dimension " + std::to_string(m) +
                                          ") is smaller than the number of states asked for");
    // Ritz vectors X = Q_{1..m} S and the true residuals
    res.evals.assign(w.begin(), w.end());
    res.steps = steps;
    res.resid.assign(nev, 0.0);
    {
        hipError_t e = hipMalloc((void **)&HX, sizeof(double) * ld * 2 * nev);
        if (e != hipSuccess) return fail(ED_ERR_NOMEM, std::string("block Lanczos Ritz vectors: ") + hipGetErrorString(e));
    }
    double *X = HX + (size_t)nev * ld;
    std::vector<double> S((size_t)m * nev);   // row-major m x nev
    for (int k = 0; k < nev; ++k)
        for (int r = 0; r < m; ++r) S[(size_t)r * nev + k] = Z[(size_t)k * m + r];
    ED_TRY(xc(dev, n, m, nev, Q, ld, S.data(), 1.0, 0.0, X, ld, st));
    ED_TRY(launch_hv(s, ED_PREC_F64, X, HX, ld, nev, st));
    std::vector<double> D((size_t)nev * nev, 0.0), RtR((size_t)nev * nev);
    for (int k = 0; k < nev; ++k) D[(size_t)k * nev + k] = w[k];
    ED_TRY(xc(dev, n, nev, nev, X, ld, D.data(), -1.0, 1.0, HX, ld, st));
    ED_TRY(gram(dev, n, nev, nev, HX, ld, HX, ld, RtR.data(), st));
    int nconv = 0;
    for (int k = 0; k < nev; ++k) {
        res.resid[k] = std::sqrt(std::max(0.0, RtR[(size_t)k * nev + k]));
        if (res.resid[k] < std::sqrt(o.tolerance) * std::max(1.0, std::fabs(w[k]))) ++nconv;
    }
    res.nconv = nconv;
    g_last_steps = res.steps;
    g_last_full = res.full_passes;
    g_last_ritz_ms = ritz_ms;
    g_last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    if (d_out)
        ED_HIP_CHECK(hipMemcpy2DAsync(d_out, sizeof(double) * ldo, X, sizeof(double) * ld, sizeof(double) * n, nev,
                                      hipMemcpyDeviceToDevice, st));
    ED_HIP_CHECK(hipStreamSynchronize(st));
    return ED_OK;
}

}  // namespace ed

extern "C" {

int ed_block_default_opts(ed_block_opts *o)
{
    if (!o) return ed::fail(ED_ERR_ARG, "ed_block_default_opts: NULL");
    o->block_size = 8;
    o->nev = 8;
    o->max_steps = 128;
    o->check_every = 2;
    o->tolerance = 1e-12;
    o->reorth = 1;
    return ED_OK;
}

static int block_check(ed_sector *sec, const ed_block_opts *o)
{
    if (!sec || !o) return ed::fail(ED_ERR_ARG, "ed_block_lanczos: NULL argument");
    // the real embedding of a complex H doubles every eigenvalue: a block method would
    // return each state twice (psi and i psi)
    if (sec->cplx) return ed::fail(ED_ERR_ARG, "ed_block_lanczos: real sectors only");
    if (o->block_size < 1 || o->block_size > 16) return ed::fail(ED_ERR_ARG, "ed_block_lanczos: block_size must be 1..16");
    if (o->nev < 1 || o->nev > 16 || o->nev > sec->dim)
        return ed::fail(ED_ERR_ARG, "ed_block_lanczos: nev must be 1..16 and at most the sector dimension");
    if (o->max_steps < 1) return ed::fail(ED_ERR_ARG, "ed_block_lanczos: max_steps < 1");
    if (!(o->tolerance > 0.0)) return ed::fail(ED_ERR_ARG, "ed_block_lanczos: tolerance must be positive");
    if (o->reorth != 0 && o->reorth != 1) return ed::fail(ED_ERR_ARG, "ed_block_lanczos: reorth must be 0 or 1");
    if ((int64_t)o->block_size > sec->dim) return ed::fail(ED_ERR_ARG, "ed_block_lanczos: block_size exceeds the dimension");
    ED_HIP_CHECK(hipSetDevice(sec->device));
    return ED_OK;
}

// the host-vector forms: every eigenvector (window = +inf), or those within `window` of
// the lowest eigenvalue (a leading run, the values being ascending)
static int block_lanczos_host(ed_sector *sec, const ed_block_opts *opts, double window, double *evals, double *resid,
                              double *evecs, int64_t ldv, int *nconv, int *steps, int *nkept)
{
    ED_TRY(block_check(sec, opts));
    if (!evals) return ed::fail(ED_ERR_ARG, "ed_block_lanczos: NULL evals");
    if (evecs && ldv < sec->dim) return ed::fail(ED_ERR_ARG, "ed_block_lanczos: ldv < dimension");
    if (!(window >= 0.0)) return ed::fail(ED_ERR_ARG, "ed_block_lanczos_window: dEmax must be >= 0");
    ed::BlockResult r;
    double *d_out = nullptr;
    if (evecs) {
        hipError_t e = hipMalloc((void **)&d_out, sizeof(double) * sec->dim * opts->nev);
        if (e != hipSuccess) return ed::fail(ED_ERR_NOMEM, "ed_block_lanczos: eigenvector staging");
    }
    int rc = ed::block_lanczos(sec, *opts, d_out, sec->dim, r);
    int keep = 0;
    if (rc == ED_OK)
        while (keep < opts->nev && r.evals[keep] - r.evals[0] <= window) ++keep;
    if (rc == ED_OK && evecs && keep > 0) {
        hipError_t e = hipMemcpy2D(evecs, sizeof(double) * ldv, d_out, sizeof(double) * sec->dim,
                                   sizeof(double) * sec->dim, keep, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = ed::fail(ED_ERR_HIP, std::string("ed_block_lanczos: copy-out: ") + hipGetErrorString(e));
    }
    if (d_out) (void)hipFree(d_out);
    if (rc != ED_OK) return rc;
    for (int k = 0; k < opts->nev; ++k) {
        evals[k] = r.evals[k];
        if (resid) resid[k] = r.resid[k];
    }
    if (nconv) *nconv = r.nconv;
    if (steps) *steps = r.steps;
    if (nkept) *nkept = keep;
    return ED_OK;
}

int ed_block_lanczos(ed_sector *sec, const ed_block_opts *opts, double *evals, double *resid, double *evecs,
                     int64_t ldv, int *nconv, int *steps)
{
    return block_lanczos_host(sec, opts, HUGE_VAL, evals, resid, evecs, ldv, nconv, steps, nullptr);
}

int ed_block_lanczos_window(ed_sector *sec, const ed_block_opts *opts, double dEmax, double *evals, double *resid,
                            double *evecs, int64_t ldv, int *nconv, int *steps, int *nkept)
{
    return block_lanczos_host(sec, opts, dEmax, evals, resid, evecs, ldv, nconv, steps, nkept);
}

int ed_block_lanczos_dev(ed_sector *sec, const ed_block_opts *opts, double *evals, double *resid, double *d_evecs,
                         int64_t ldv, int *nconv, int *steps)
{
    ED_TRY(block_check(sec, opts));
    if (!evals || !d_evecs || ldv < sec->dim) return ed::fail(ED_ERR_ARG, "ed_block_lanczos_dev: bad argument");
    ed::BlockResult r;
    ED_TRY(ed::block_lanczos(sec, *opts, d_evecs, ldv, r));
    for (int k = 0; k < opts->nev; ++k) {
        evals[k] = r.evals[k];
        if (resid) resid[k] = r.resid[k];
    }
    if (nconv) *nconv = r.nconv;
    if (steps) *steps = r.steps;
    return ED_OK;
}

int ed_block_last_stats(int *steps, int *full_passes, double *ms, double *ritz_ms)
{
    if (steps) *steps = ed::g_last_steps;
    if (full_passes) *full_passes = ed::g_last_full;
    if (ms) *ms = ed::g_last_ms;
    if (ritz_ms) *ritz_ms = ed::g_last_ritz_ms;
    return ED_OK;
}

int ed_gram_f64_dev(int64_t n, int p, int q, const double *d_x, int64_t ldx, const double *d_y, int64_t ldy, double *g,
                    void *stream)
{
    if (n < 1 || p < 1 || q < 1 || q > 16 || !d_x || !d_y || !g || ldx < n || ldy < n)
        return ed::fail(ED_ERR_ARG, "ed_gram_f64_dev: bad argument");
    if ((ldx & 1) || (ldy & 1) || ((uintptr_t)d_x & 15) || ((uintptr_t)d_y & 15))
        return ed::fail(ED_ERR_ARG, "ed_gram_f64_dev: pointers must be 16-byte aligned and leading dimensions even");
    int dev = 0;
    ED_HIP_CHECK(hipGetDevice(&dev));
    return ed::gram(dev, n, p, q, d_x, ldx, d_y, ldy, g, (hipStream_t)stream);
}

int ed_xc_f64_dev(int64_t n, int p, int q, const double *d_x, int64_t ldx, const double *c, double alpha, double beta,
                  double *d_y, int64_t ldy, void *stream)
{
    if (n < 1 || p < 1 || q < 1 || q > 16 || !d_x || !d_y || !c || ldx < n || ldy < n)
        return ed::fail(ED_ERR_ARG, "ed_xc_f64_dev: bad argument");
    if (d_x == d_y && (p != q || ldx != ldy)) return ed::fail(ED_ERR_ARG, "ed_xc_f64_dev: in place needs p == q");
    int dev = 0;
    ED_HIP_CHECK(hipGetDevice(&dev));
    return ed::xc(dev, n, p, q, d_x, ldx, c, alpha, beta, d_y, ldy, (hipStream_t)stream);
}

}  // extern "C"
