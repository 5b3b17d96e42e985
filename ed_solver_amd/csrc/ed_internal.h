// Internal definitions shared by the ed_hip translation units (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/ed_hip.h"

namespace ed {

// ------------------------------------------------------------------ errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define ED_HIP_CHECK(expr)                                                            \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return ::ed::fail(ED_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define ED_TRY(x)                     \
    do {                              \
        int rc_ = (x);                \
        if (rc_ != ED_OK) return rc_; \
    } while (0)

constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kBarGroups = 64;        // grid-barrier groups of the persistent Lanczos kernel
constexpr int kMaxRowsCoef = 2048;   // packed up-hop coefficients the row kernels keep in LDS
constexpr int kBarGroupSize = 32;     // workgroups per group
constexpr int kRed = 34;             // reduction scratch: two slots per wave of a 1024-thread workgroup + 2
constexpr int kSums = 4;             // LDS slots for the reduced recurrence sums
constexpr int kSchedules = 8;        // H.v schedules the autotuner times
constexpr int kMaxIter = 2048;       // capacity of the on-device alpha/beta arrays
constexpr double kMinLanczosNorm = 1e-25;  // cutoff_min_lanczos_vec (dmft_ed_solver_variables.h)
// limits of ed_dynamic_noalloc's static scratch (ed_noalloc.hip)
constexpr int64_t kNoAllocMaxDim = ED_NOALLOC_MAX_DIM;
constexpr int kNoAllocMaxIter = ED_NOALLOC_MAX_ITER;
constexpr int kNoAllocMaxNw = ED_NOALLOC_MAX_NW;

// Layout of the per-sector scratch array of doubles (offsets in doubles).
constexpr int kPartA = 0;                          // partial sums written by the H.v phase
constexpr int kPartB = kPartA + 8192;              // update phase, ||u_j||^2, [parity of j][8192]
constexpr int kMaxRowBlocks = 8192;                // workgroups of a row-per-thread H.v (one kPartA partial each)
constexpr int kPartC = kPartB + 2 * 8192;          // update phase, <u_j, u_{j-1}>, [parity of j][8192]
constexpr int kPartR = kPartC + 2 * 8192;          // reorthogonalisation dots [k][workgroup]
constexpr int kReorthGrid = 256;                   // workgroups of the reorthogonalisation dots
constexpr int kCoef = kPartR + kMaxIter * kReorthGrid;   // reorthogonalisation coefficients
constexpr int kMisc = kCoef + kMaxIter;            // Ritz coefficients, norms, Rayleigh quotient
constexpr int kScratchDoubles = kMisc + kMaxIter + 64;

// ------------------------------------------------------------ device state
// Recurrence scalars of one Lanczos run, resident on the device so that a chunk of
// iterations runs (and can be replayed from a captured graph) without the host.
// The iteration index is double-buffered between the two phases of a step so that no
// kernel ever writes a word that its own workgroups read: the H.v phase reads j_k1
// and writes j_k2, the update phase reads j_k2 and writes j_k1.
struct LzState {
    int j_k1;          // step index seen by the H.v phase (1-based)
    int j_k2;          // step index seen by the update phase
    int status;        // 0 ok, 1 zero-norm vector appeared, 2 zero-norm start vector,
                       // 3 grid barrier of the persistent Lanczos kernel timed out
    int stop_at;       // monitored persistent run: the step after which every workgroup
                       // stops (0 = not asked to stop); zeroed before each launch
    unsigned bar_count;   // persistent Lanczos kernel: group leaders' arrivals since its launch
    int bar_err;          // a barrier wait exceeded its spin limit
    int exact_beta_steps; // one-barrier persistent kernel: steps whose beta took the exact
                          // norm (one more grid barrier) since its launch
    double alpha[kMaxIter];      // alpha_j        -> alpha[j-1]
    double beta[kMaxIter + 1];   // beta_1 = ||u_1|| -> beta[0]; beta_{j+1} -> beta[j]
    // grid barrier groups of the persistent Lanczos kernel, one 64-byte line per word
    unsigned bar_grp[kBarGroups * 16];   // arrivals in group g: bar_grp[16 g]
    unsigned bar_rel[kBarGroups * 16];   // barriers released in group g: bar_rel[16 g]
};

// Quadratic model tabulated by tab_HAIM2: nimp impurity and nbath bath orbitals, their
// bit positions in the state word, and device copies of the row-major matrices
// ec[nimp][nimp], eb[nbath][nbath], vbc[nbath][nimp].  One spin of the (N_up, N_dn)
// basis: nimp = Nc at bits nb + c, nbath = Nb at bits b.  The Nambu basis of the S_z
// sectors: 2 Nc and 2 Nb orbitals over 2 Ns bits (AIM2%IMPiorb / BATHiorb,
// quantum_algebra_Anderson_Impurity_class2.f90:83-89).
constexpr int kMaxImpOrb = 12, kMaxBathOrb = 28;
struct SpinModel {
    int nimp, nbath;
    int ibit[kMaxImpOrb], bbit[kMaxBathOrb];
    const double *ec, *eb, *vbc;
    const double *ec_im, *eb_im, *vbc_im;   // imaginary parts of a complex model (null: real)
    double cut;
};

// One spin's tables on the device (tab_HAIM2 output plus derived layouts).
struct SpinTables {
    int d = 0;              // spin-sector dimension
    int64_t nnz = 0;        // tabulated off-diagonal entries
    int kmax = 0;           // max entries per state
    int kpad = 0;           // ELL width (kmax rounded up to a compiled width)
    int32_t *states = nullptr;   // [d]
    double *diag = nullptr;      // [d]
    int32_t *noff = nullptr;     // [d]
    int32_t *ptr = nullptr;      // [d+1] exclusive prefix sum of noff
    int32_t *rk = nullptr;       // [nnz] 0-based target rank (CSR order = reference order)
    double *off = nullptr;       // [nnz]
    int32_t *ell_rk = nullptr;   // [kpad][d], padded with the row itself
    double *ell_val = nullptr;   // [kpad][d], padded with 0
    int32_t *imp = nullptr;      // [d] impurity bits (state >> nb)
    int32_t *swap = nullptr;     // [d][ncfg] rank of the state with its impurity bits replaced
    // FP32 copies for the mixed-precision kernels
    float *diag32 = nullptr, *off32 = nullptr, *ell_val32 = nullptr;
    // complex sector: imaginary parts of off (CSR) and ell_val (ELL)
    double *off_im = nullptr, *ell_vim = nullptr;
    // packed ELL: entry = target rank (low 16 bits) | coefficient index (high 16 bits)
    // into the table of distinct coefficients coef[ncoef] (coef[0] = 0 pads short rows)
    uint32_t *ell_pk = nullptr;  // [max(kpad of both spins)][d]
    double *coef = nullptr;
    float *coef32 = nullptr;
    int ncoef = 0;
};

// Kernel-side view of a sector's Hamiltonian for precision T.
template <typename T>
struct HvArgs {
    int dup, ddn, ncfg, kup, cpb, hund;
    const int32_t *ell_rk_up;
    const T *ell_val_up;
    const int32_t *ptr_dn, *rk_dn;
    const T *off_dn;
    const T *diag_up, *diag_dn;
    const int32_t *imp_up, *imp_dn;
    const T *uimp;                 // [ncfg][ncfg] interaction diagonal per impurity pair
    const int32_t *hund_ptr;       // [ncfg*ncfg + 1]
    const int32_t *hund_cu, *hund_cd;
    const T *hund_val;
    const int32_t *swap_up, *swap_dn;
    // packed layouts of the fast kernel
    const uint32_t *ell_pk_up;     // [max(kup, kdn)][dup]: as many rows as the fast kernel unrolls
    const T *coef_up;              // [ncoef_up]
    int ncoef_up, kdn, nhund;
    const int32_t *hund_tr;        // [nhund] cu' | cd' << 8
    const int32_t *colord;         // fast kernel: workgroup -> column (XCD-grouped order, -1 = idle), or null
};

// Host-visible channel of a monitored persistent Lanczos run, in fine-grained pinned host
// memory: the kernel publishes beta_j and alpha_j of every step (ab[2 (j-1)] and
// ab[2 (j-1) + 1], NaN until written) while the host tests convergence, and stops after
// the first step in which it reads *stop != 0.  Null pointers: no monitor.
struct LzMonitor {
    const int *stop;
    double *ab;
};

// Where the Lanczos vectors live: slot(j) = base + ((j-1) % nslots) * ld, j >= 1, and
// where the per-workgroup partial sums of the two phases go.
template <typename T>
struct LzArgs {
    T *base;
    int64_t ld;
    int nslots;
    LzState *st;
    double *partA;     // [nA] H.v phase: <H u_j, u_j>
    double *partB;     // [2][pstride] update phase, by parity of j: ||u_j||^2
    double *partC;     // [2][pstride] update phase, by parity of j: <u_j, u_{j-1}>
    int pstride;
    int nA, nB;        // workgroups that wrote partA (H.v phase) / partB, partC (update phase)
    int colmap, dup;   // colmap: element-wise kernels give workgroup b the column b (d_up
                       // elements), the column the H.v workgroup b reads and writes, so it
                       // stays in the same XCD's L2; else a grid-stride split
    const int32_t *colord;   // with colmap: workgroup -> column as in the H.v (null = identity)
    __host__ __device__ T *slot(int j) const { return base + (int64_t)((j - 1) % nslots) * ld; }
    __host__ __device__ double *nrm(int j) const { return partB + (j & 1) * pstride; }
    __host__ __device__ double *crs(int j) const { return partC + (j & 1) * pstride; }
};

// Device work space of one Lanczos recurrence: nslots Krylov slots of ld elements
// (sized for FP64) at `basis`, the scalars `st` and kScratchDoubles-layout partial sums
// (only kPartA .. kPartC + 2 * 8192 are used without reorthogonalisation).
struct LzWork {
    void *basis;
    int64_t ld;
    int nslots;
    LzState *st;
    double *partials;
};

struct GraphCache {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int prec = -1, chunk = 0, nslots = 0;
    void *base = nullptr;
};

// The HIP stream of a sector, taken from the per-device stream pool when first used (the
// conversion below): a sector that only joins batched work, which runs on another sector's
// stream, never needs one of its own.  The sector build runs on a pooled stream that goes
// back to the pool when the build is done.
struct SectorStream {
    mutable hipStream_t s = nullptr;
    int device = 0;
    operator hipStream_t() const;
};

}  // namespace ed

struct ed_sector {
    int device = 0;
    ed::SectorStream stream;
    int nc = 0, nb = 0, ns = 0, nup = 0, ndn = 0, ncfg = 0;
    int64_t dim = 0;
    std::vector<double> ec, eb, vbc, u, jj;
    // Slater 4-index impurity interaction U(s1, s2, s3, s4) [nc^4] (flag_slater_int); empty = off
    std::vector<double> slater;
    double jhund = 0.0, cutoff = 0.0;
    int32_t *d_rank = nullptr;       // [2^ns] 0-based rank inside the popcount bucket
    ed::SpinTables up, dn;
    // keys of up / dn in the process's cache of spin tables (empty: the sector owns them)
    std::string up_key, dn_key;
    double *d_uimp = nullptr;
    float *d_uimp32 = nullptr;
    int hund = 0, nhund = 0;
    int32_t *d_hund_ptr = nullptr, *d_hund_cu = nullptr, *d_hund_cd = nullptr, *d_hund_tr = nullptr;
    double *d_hund_val = nullptr;
    float *d_hund_val32 = nullptr;
    // Lanczos work space
    void *d_basis = nullptr;
    size_t basis_bytes = 0;
    int basis_prec = -1, basis_nslots = 0, basis_valid_iters = 0;
    int64_t ld = 0;
    ed::LzState *d_state = nullptr;
    double *d_partials = nullptr;
    double *d_io_a = nullptr, *d_io_b = nullptr;   // host-API staging vectors [dim]
    ed::GraphCache graphs[4];      // captured Lanczos chunks, key = precision * 2 + reorth
    int last_f32_iters = 0;
    // last ED_PREC_MIXED_RQ solve: accepted without the FP64 phase, squared residual of the
    // Ritz vector, Ritz gap lambda_2 - E
    int last_rq_accepted = 0;
    double last_rq_res2 = 0.0, last_rq_gap = 0.0;
    int cpb = 0;                   // H.v schedule chosen by the autotuner: 1/2/4 = row kernel with that many
                                   // columns per workgroup, 8 = fast kernel, 16 = fast kernel with the
                                   // XCD-grouped column order; 0 = not tuned
    double tune_ms[ed::kSchedules] = {};   // measured FP64 H.v time per schedule (1, 2, 4, 8, 16, 18, 19, 20)
    int cpb32 = 0;                 // schedule of the FP32 kernels (tuned on their own); 0 = as cpb
    int lzs = 0, lzs32 = 0;        // schedule of the Lanczos recurrence (FP64, FP32); 0 = as the H.v's
    double tune_ms32[ed::kSchedules] = {};
    int32_t *d_colord = nullptr;   // [colord_blocks] workgroup -> down column, -1 = idle
    int colord_blocks = 0;
    std::vector<int32_t> col_bfs;  // down columns in the breadth-first order of the down-hop graph
    int hund_maxrow = 0;           // most Hund transitions of one (impurity up, impurity down) pair
    // map of the multi-column persistent kernel (PersistMap): [cols][pslot][idle] for the
    // column order of pmap_xcd (1: XCD-grouped, 0: natural) and pmap_cpw columns per workgroup
    int32_t *d_pmap = nullptr;
    int pmap_cpw = 0, pmap_xcd = -1, pmap_nwg = 0, pmap_nidle = 0;
    int persist_cols = 1;          // the multi-column persistent kernel where it applies (1) or never (0)
    int persist_1b = 1;            // its one-grid-barrier-per-step form (1) or the two-barrier form (0)
    int persist_exact_beta = 0;    // one-barrier form: every beta from the reduced norm (test hook)
    void *d_qx = nullptr;          // one-barrier form: two exchange vectors of ld FP64 elements
    // S_z (Nambu) sector: one table over 2 Ns bits in `up` (dn.d = 1), the quartic
    // diagonal QUART_INT_SZ in d_quart, and the constant tr(Ec_dn) + tr(Eb_dn) that the
    // particle-hole transformation moves out of H
    int sz = 0;
    // complex sector (Hmultc): Ec, Eb, Vbc with imaginary parts; vectors are planar,
    // (Re[cdim], Im[cdim]), so dim = 2 cdim is the length of every real-typed vector
    int cplx = 0;
    int64_t cdim = 0;
    std::vector<double> ec_im, eb_im, vbc_im;
    double *d_quart = nullptr;
    float *d_quart32 = nullptr;
    // S_z sector with Hund terms (HAIMsz_multc): ELL [sz_kh][dim] of the spin-flip and
    // pair-hopping transitions (target rank, -J sign), padded with the row itself and 0
    int sz_kh = 0;
    int32_t *d_sz_hrk = nullptr;
    double *d_sz_hval = nullptr;
    float *d_sz_hval32 = nullptr;
    // HAIMsz_mult_fly: the Nambu matrices on the device (d_sz_mat) and the sector's
    // tab_HAIM2 model over them; sz_fly = 1 generates the entries in the H.v
    ed::SpinModel sz_model{};
    double *d_sz_mat = nullptr;
    int sz_fly = 0;
    double energy_shift = 0.0;
    int persist = 1;               // use the persistent Lanczos kernel where it applies (1) or not (0);
                                   // 2: as 1, but every cooperative launch takes the refusal path (test hook)
    int persist_refused = 0;       // cooperative launches the runtime refused on this sector
    double *h_mon = nullptr;       // monitored persistent run: pinned fine-grained host channel
    double *d_mon = nullptr;       //   (int stop flag at [0], published scalars from [16]), device view
    int last_steps_run = 0;        // Lanczos steps the device ran in the last diagonalisation
    int persist_ok[2] = {-1, -1};  // per precision (F64, F32): -1 not checked, 0 no, 1 co-resident grid
    // timing of the H.v kernel inside Lanczos chunks
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double hv_ms_sum = 0.0;
    int hv_count = 0;
    int hv_prec = 0;
};

namespace ed {
// tridiagonal eigen-solver (host, FP64): implicit QL with Wilkinson shifts
int tridiag_eig(int n, const double *diag, const double *off, double *w, double *z);
// ||v||^2 of a host vector in eight interleaved partial sums (added in a fixed order), so
// the loop vectorises: one pass at memory speed instead of a chain of dependent adds.
inline double host_norm2(const double *v, int64_t n)
{
    double p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t i = 0;
    for (; i + 8 <= n; i += 8)
        for (int q = 0; q < 8; ++q) p[q] += v[i + q] * v[i + q];
    for (int q = 0; i < n; ++i, ++q) p[q] += v[i] * v[i];
    return ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
}

void tridiag_lowest(int n, const double *diag, const double *off, int k, double *w);
void tridiag_lowest_warm(int n, const double *diag, const double *off, int k, const double *prev, double *w);
int tridiag_lowest_pairs(int n, const double *diag, const double *off, int k, double *w, double *z);
// nev lowest eigenpairs of a symmetric band matrix (n x n row-major, half-bandwidth b);
// z (n x nev, column-major) may be null
int band_lowest(int n, int b, const double *a, int nev, double *w, double *z);
// The same in two parts: band_reduce brings the band matrix to the tridiagonal (d, e) by
// plane rotations, logged when `keep_log` (the vectors need them); band_tri_values gives
// the nev lowest eigenvalues and band_tri_vectors their vectors from one reduction.
struct BandRot {
    int p;
    double c, s;
};
struct BandTri {
    int n = 0;
    std::vector<double> d, e;
    std::vector<BandRot> log;
};
int band_reduce(int n, int b, const double *a, bool keep_log, BandTri &t);
int band_tri_values(const BandTri &t, int nev, double *w);
int band_tri_vectors(const BandTri &t, int nev, const double *w, double *z);
// launch helpers implemented in ed_engine.hip
int launch_hv(ed_sector *s, int prec, const void *x, void *y, int64_t ld, int nvec, hipStream_t st);
HvArgs<double> hv_args_f64(const ed_sector *s);
HvArgs<float> hv_args_f32(const ed_sector *s);
int start_vector_dev(int64_t n, double *d_out, hipStream_t st);
// the no-allocation scratch of this device (ed_noalloc.hip); the device must be current
int noalloc_setup(int device);
// Fixed-length Lanczos run (as ed_lanczos_run) entirely in the caller's work space w
// (w.nslots >= 3) with plain two-kernel steps on the sector's stream: no device memory is
// allocated and the sector's own Krylov basis is left alone.  The H.v schedule is the
// tuned one, or the untuned default when the sector was never tuned.  diag/subdiag
// (capacity niter, host) as ed_lanczos_run; ED_ERR_BREAKDOWN as there, with the completed
// steps in *breakdown_at when given (0 otherwise).
int lanczos_run_ws(ed_sector *s, const LzWork &w, int niter, const double *d_start, int prec, double *diag,
                   double *subdiag, int *breakdown_at = nullptr);
// s->col_bfs: the sector's down columns in the breadth-first order of its down-hop graph
int ensure_col_bfs(ed_sector *s);
// FP64 MFMA contractions of ed_block.hip on device matrices (column-major, 16-byte aligned,
// even leading dimensions), queued on st: d_G (row-major p x q) = X^T Y and
// Y = beta Y + alpha X C (d_C row-major p x q), q <= 16; block_ws_reserve sizes the Gram
// work space for up to pmax columns so that queued products never reallocate it.
int block_ws_reserve(int dev, int64_t n, int pmax);
int block_gram_dev(int dev, int64_t n, int p, int q, const double *X, int64_t ldx, const double *Y, int64_t ldy,
                   double *d_G, hipStream_t st);
void block_xc_dev(int64_t n, int p, int q, const double *X, int64_t ldx, const double *d_C, double alpha, double beta,
                  double *Y, int64_t ldy, hipStream_t st);
}  // namespace ed
