// Batched excitation Lanczos: the per-(state, orbital, iph) compute_dynamic runs of a
// Green's function (green_class_computeAA.h:101-302 -> green_class_compute_dynamic.f90:
// 191-407) advanced together.  Every job (one excitation vector in its own N+-1 sector,
// sectors of uneven dimension) owns a contiguous range of workgroups in ONE launch per
// recurrence phase:
//
//   k_hv_batch      H.v + beta-subtraction + alpha partials of every job
//   k_update_batch  alpha-subtraction + beta partials of every job
//
// so a Green's function with J excitation runs costs 2 launches per Lanczos step instead
// of 2J, and small sectors fill the GPU together.  A workgroup finds its job by a scan
// of the (uniform) job table.  The continued fractions of all jobs and frequencies run
// in one more launch (k_cfrac).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ed_kernels.h"

namespace ed {

int lanczos_run_dev(ed_sector *s, int niter, const double *d_start, int prec, double *diag, double *subdiag,
                    int *breakdown_at);
int cfrac_batch_dev(int n, const double *d_a, const double *d_b, int lda, int njobs, int nw, const double *d_wn,
                    const double *d_scale, double *d_out, hipStream_t st);

template <typename T>
struct JobDesc {
    HvArgs<T> a;
    LzArgs<T> lz;
    const double *start;   // normalised excitation vector (device, FP64)
    int64_t n;
    int blk0_hv, nblk_hv;  // workgroups of this job in the batched H.v grid
    int blk0_el, nblk_el;  // workgroups of this job in the batched element-wise grid
};

template <typename T>
__device__ inline int find_job_hv(const JobDesc<T> *jobs, int njobs)
{
    int j = 0;
    while (j + 1 < njobs && (int)blockIdx.x >= jobs[j + 1].blk0_hv) ++j;
    return j;
}

template <typename T>
__device__ inline int find_job_el(const JobDesc<T> *jobs, int njobs)
{
    int j = 0;
    while (j + 1 < njobs && (int)blockIdx.x >= jobs[j + 1].blk0_el) ++j;
    return j;
}

template <typename T, int KU>
__global__ __launch_bounds__(kBlock) void k_hv_batch(const JobDesc<T> *jobs, int njobs)
{
    extern __shared__ double smem_d[];
    __shared__ double red[kRed];
    __shared__ double grp[kSums];
    const int j = find_job_hv(jobs, njobs);
    const HvArgs<T> a = jobs[j].a;
    const LzArgs<T> lz = jobs[j].lz;
    hv_rows_body<T, 1, 1, KU, true>(a, (const T *)nullptr, (T *)nullptr, lz, blockIdx.x - jobs[j].blk0_hv,
                                    reinterpret_cast<T *>(smem_d), red, grp);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_update_batch(const JobDesc<T> *jobs, int njobs)
{
    __shared__ double red[kRed];
    __shared__ double grp[kSums];
    const int j = find_job_el(jobs, njobs);
    const LzArgs<T> lz = jobs[j].lz;
    lz_update_body(lz, jobs[j].n, 1, blockIdx.x - jobs[j].blk0_el, jobs[j].nblk_el, red, grp);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_init_batch(const JobDesc<T> *jobs, int njobs)
{
    __shared__ double red[kRed];
    __shared__ double grp[kSums];
    const int j = find_job_el(jobs, njobs);
    const LzArgs<T> lz = jobs[j].lz;
    lz_init_body(lz, jobs[j].start, jobs[j].n, blockIdx.x - jobs[j].blk0_el, jobs[j].nblk_el, red);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_finish_batch(const JobDesc<T> *jobs)
{
    __shared__ double red[kRed];
    __shared__ double grp[kSums];
    const LzArgs<T> lz = jobs[blockIdx.x].lz;
    lz_finish_body(lz, grp);
}

// phi_k / ||phi_k|| for every job into its start area.
// Partial sums of ||x||^2 per workgroup (grid-stride, fixed order); the host adds them.
__global__ __launch_bounds__(kBlock) void k_norm2_part(const double *x, int64_t n, double *part)
{
    __shared__ double red[kRed];
    double v = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v += x[i] * x[i];
    v = block_sum_all(v, red);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void k_scale_into(const double *x, double *y, int64_t n, double s)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = s * x[i];
}

static int egrid(int64_t n)
{
    int64_t g = (n + kBlock * 4 - 1) / (kBlock * 4);
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 2048));
}

// Workgroups of a job in the batched element-wise kernels: 8 elements per thread, so the
// per-step scalar reductions (every workgroup sums the partials of its job) read half as
// many partials as at 4.
static int egrid_el(int64_t n)
{
    int64_t g = (n + kBlock * 8 - 1) / (kBlock * 8);
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 2048));
}

// Grow-only device work space of the batched runs (one per process; calls serialise).
struct BatchWS {
    std::mutex mu;
    void *ptr = nullptr;
    size_t bytes = 0;
    int device = -1;
};
static BatchWS g_ws;
// work space of the sectors ed_lanczos_batched runs one by one (recurrence state, partial
// sums, 3 Krylov slots, start vector), its own lock: those runs also happen while g_ws is held
static BatchWS g_alone;
// launches of the multi-run persistent kernel, and launches the runtime refused (then the
// two-launch batch ran), since the library was loaded
static int g_runs_launched = 0, g_runs_refused = 0;
// the last launch: its duration (HIP event pair on the stream), job-steps and algorithmic
// bytes (5 s D per job-step: H.v reads u_j and writes H u_j, the recurrence reads u_{j-1},
// u_j and writes u_{j+1})
static double g_runs_ms = 0.0, g_runs_job_steps = 0.0, g_runs_bytes = 0.0;
// wait rounds of a grid-barrier poll of the multi-run kernel before it gives up (seconds)
constexpr int kRunsBarrierSpins = 1 << 21;

static constexpr size_t al256(size_t x) { return (x + 255) / 256 * 256; }

static int ws_reserve_in(BatchWS &w, size_t bytes, int device)
{
    if (w.ptr && w.bytes >= bytes && w.device == device) return ED_OK;
    if (w.ptr) {
        (void)hipDeviceSynchronize();
        (void)hipFree(w.ptr);
        w.ptr = nullptr;
        w.bytes = 0;
    }
    hipError_t e = hipMalloc(&w.ptr, bytes);
    if (e != hipSuccess) return fail(ED_ERR_NOMEM, "batched Lanczos work space: " + std::string(hipGetErrorString(e)));
    w.bytes = bytes;
    w.device = device;
    return ED_OK;
}

static int ws_reserve(size_t bytes, int device) { return ws_reserve_in(g_ws, bytes, device); }

static bool rows_capable(const ed_sector *s) { return !s->sz && !s->cplx && s->up.d >= 128 && s->up.kpad <= 32; }

template <typename T>
static HvArgs<T> hv_args_of(const ed_sector *s)
{
    const bool f64 = sizeof(T) == sizeof(double);
    HvArgs<T> a;
    a.dup = s->up.d; a.ddn = s->dn.d; a.ncfg = s->ncfg; a.kup = s->up.kpad; a.cpb = 1; a.hund = s->hund;
    a.ell_rk_up = s->up.ell_rk; a.ell_val_up = (const T *)(f64 ? (void *)s->up.ell_val : (void *)s->up.ell_val32);
    a.ptr_dn = s->dn.ptr; a.rk_dn = s->dn.rk; a.off_dn = (const T *)(f64 ? (void *)s->dn.off : (void *)s->dn.off32);
    a.diag_up = (const T *)(f64 ? (void *)s->up.diag : (void *)s->up.diag32);
    a.diag_dn = (const T *)(f64 ? (void *)s->dn.diag : (void *)s->dn.diag32);
    a.imp_up = s->up.imp; a.imp_dn = s->dn.imp;
    a.uimp = (const T *)(f64 ? (void *)s->d_uimp : (void *)s->d_uimp32);
    a.hund_ptr = s->d_hund_ptr; a.hund_cu = s->d_hund_cu; a.hund_cd = s->d_hund_cd;
    a.hund_val = (const T *)(f64 ? (void *)s->d_hund_val : (void *)s->d_hund_val32);
    a.swap_up = s->up.swap; a.swap_dn = s->dn.swap;
    // packed up hops (see hv_rows_body) when the coefficient table fits the LDS budget
    a.ncoef_up = s->up.ncoef;
    a.coef_up = (const T *)(f64 ? (void *)s->up.coef : (void *)s->up.coef32);
    a.ell_pk_up = (s->up.ell_pk && s->up.ncoef <= kMaxRowsCoef) ? s->up.ell_pk : nullptr;
    a.kdn = s->dn.kpad; a.nhund = s->nhund; a.hund_tr = s->d_hund_tr; a.colord = nullptr;
    return a;
}

template <typename T, int KU>
static void launch_hv_batch(int total, size_t lds, hipStream_t st, const JobDesc<T> *d_jobs, int nj)
{
    hipLaunchKernelGGL((k_hv_batch<T, KU>), dim3(total), dim3(kBlock), lds, st, d_jobs, nj);
}

// Runs `niter` Lanczos steps for every job together; fills a/b [nj][niter].
template <typename T>
static int run_batch(const std::vector<ed_sector *> &secs, const std::vector<const double *> &h_phi,
                     const std::vector<double> &nrm2, int niter, hipStream_t st, std::vector<double> &a_out,
                     std::vector<double> &b_out, bool phi_on_device)
{
    const int nj = (int)secs.size();
    // ---- layout of the work space
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) / 256 * 256; return o; };
    std::vector<size_t> o_basis(nj), o_start(nj), o_part(nj);
    std::vector<int64_t> ld(nj);
    const size_t o_state = take(sizeof(LzState) * nj);
    const size_t o_jobs = take(sizeof(JobDesc<T>) * nj);
    for (int k = 0; k < nj; ++k) {
        ld[k] = (secs[k]->dim + 63) / 64 * 64;
        o_basis[k] = take(sizeof(T) * (size_t)ld[k] * (niter + 1));
        o_start[k] = take(sizeof(double) * (size_t)secs[k]->dim);
        o_part[k] = take(sizeof(double) * (8192 + 4 * 2048));
    }
    ED_TRY(ws_reserve(off, secs[0]->device));
    char *base = (char *)g_ws.ptr;
    // ---- job table
    std::vector<JobDesc<T>> jobs(nj);
    int blk_hv = 0, blk_el = 0, ku = 4;
    size_t lds = 0;
    for (int k = 0; k < nj; ++k) {
        ed_sector *s = secs[k];
        JobDesc<T> &J = jobs[k];
        J.a = hv_args_of<T>(s);
        J.n = s->dim;
        J.nblk_hv = s->dn.d;
        J.nblk_el = egrid_el(s->dim);
        J.blk0_hv = blk_hv;
        J.blk0_el = blk_el;
        blk_hv += J.nblk_hv;
        blk_el += J.nblk_el;
        J.start = (const double *)(base + o_start[k]);
        J.lz.base = (T *)(base + o_basis[k]);
        J.lz.ld = ld[k];
        J.lz.nslots = niter + 1;
        J.lz.st = (LzState *)(base + o_state) + k;
        J.lz.partA = (double *)(base + o_part[k]);
        J.lz.partB = J.lz.partA + 8192;
        J.lz.partC = J.lz.partB + 2 * 2048;
        J.lz.pstride = 2048;
        J.lz.nA = J.nblk_hv;
        J.lz.nB = J.nblk_el;
        J.lz.colmap = 0;
        J.lz.colord = nullptr;
        J.lz.dup = s->up.d;
        ku = std::max(ku, s->up.kpad);
        lds = std::max(lds, ((size_t)s->up.d + (J.a.ell_pk_up ? (size_t)J.a.ncoef_up : 0)) * sizeof(T));
        ED_HIP_CHECK(hipMemcpyAsync(base + o_start[k] + 0, h_phi[k], sizeof(double) * s->dim,
                                    phi_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_scale_into, dim3(egrid(s->dim)), dim3(kBlock), 0, st, (const double *)(base + o_start[k]),
                           (double *)(base + o_start[k]), s->dim, 1.0 / std::sqrt(nrm2[k]));
    }
    const JobDesc<T> *d_jobs = (const JobDesc<T> *)(base + o_jobs);
    ED_HIP_CHECK(hipMemcpyAsync(base + o_jobs, jobs.data(), sizeof(JobDesc<T>) * nj, hipMemcpyHostToDevice, st));
    ED_HIP_CHECK(hipMemsetAsync(base + o_state, 0, sizeof(LzState) * nj, st));
    hipLaunchKernelGGL(k_init_batch<T>, dim3(blk_el), dim3(kBlock), 0, st, d_jobs, nj);
    for (int it = 0; it < niter; ++it) {
        switch (ku) {
        case 4: launch_hv_batch<T, 4>(blk_hv, lds, st, d_jobs, nj); break;
        case 8: launch_hv_batch<T, 8>(blk_hv, lds, st, d_jobs, nj); break;
        case 12: launch_hv_batch<T, 12>(blk_hv, lds, st, d_jobs, nj); break;
        case 16: launch_hv_batch<T, 16>(blk_hv, lds, st, d_jobs, nj); break;
        case 24: launch_hv_batch<T, 24>(blk_hv, lds, st, d_jobs, nj); break;
        default: launch_hv_batch<T, 32>(blk_hv, lds, st, d_jobs, nj); break;
        }
        hipLaunchKernelGGL(k_update_batch<T>, dim3(blk_el), dim3(kBlock), 0, st, d_jobs, nj);
    }
    hipLaunchKernelGGL(k_finish_batch<T>, dim3(nj), dim3(kBlock), 0, st, d_jobs);
    ED_HIP_CHECK(hipGetLastError());
    std::vector<LzState> hs(nj);
    ED_HIP_CHECK(hipMemcpyAsync(hs.data(), base + o_state, sizeof(LzState) * nj, hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipStreamSynchronize(st));
    a_out.assign((size_t)nj * niter, 0.0);
    b_out.assign((size_t)nj * niter, 0.0);
    for (int k = 0; k < nj; ++k) {
        const LzState &h = hs[k];
        const int reached = h.j_k1 - 1;
        if (h.status == 2) return fail(ED_ERR_BREAKDOWN, "batched Lanczos: zero-norm start vector in job " + std::to_string(k));
        if (h.status == 1 && reached < niter)
            return fail(ED_ERR_BREAKDOWN, "batched Lanczos: zero-norm vector at iteration " + std::to_string(reached + 1) +
                                              " of job " + std::to_string(k) + " (Lanczos_plain.h:85-92)");
        for (int i = 0; i < niter; ++i) {
            a_out[(size_t)k * niter + i] = h.alpha[i];
            b_out[(size_t)k * niter + i] = h.beta[i];
        }
    }
    return ED_OK;
}

// ------------------------------------------------- every run in one persistent launch

template <typename T, int K, int R>
__global__ __launch_bounds__(1024, 1) void k_lz_persist_runs(const PersistRun<T> *runs, int nruns, LzState *bar,
                                                             int nsteps, int spin_limit)
{
    extern __shared__ double smem_d[];
    __shared__ double sh[kSums];
    __shared__ int flag;
    lz_persist_runs_body<T, K, R>(runs, nruns, bar, nsteps, spin_limit, reinterpret_cast<unsigned char *>(smem_d), sh,
                                  &flag);
}

// beta_{n+1} of every run after the launch (one workgroup per run)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_finish_runs(const PersistRun<T> *runs)
{
    __shared__ double grp[kSums];
    lz_finish_body(runs[blockIdx.x].lz, grp);
}

// Zeroes n words (a barrier state before a launch; a kernel, so a captured graph holds no
// memset node).
__global__ void k_zero_words(unsigned *p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

template <typename T, int K, int R>
static const void *runs_fn()
{
    return reinterpret_cast<const void *>(&k_lz_persist_runs<T, K, R>);
}

// The instances: table widths 4, 8 and 12, 1, 2 or 4 rows per thread (d_up <= 1024).
template <typename T>
static const void *runs_kernel(int k, int rows)
{
    switch (k * 8 + rows) {
    case 4 * 8 + 1: return runs_fn<T, 4, 1>();
    case 4 * 8 + 2: return runs_fn<T, 4, 2>();
    case 4 * 8 + 4: return runs_fn<T, 4, 4>();
    case 8 * 8 + 1: return runs_fn<T, 8, 1>();
    case 8 * 8 + 2: return runs_fn<T, 8, 2>();
    case 8 * 8 + 4: return runs_fn<T, 8, 4>();
    case 12 * 8 + 1: return runs_fn<T, 12, 1>();
    case 12 * 8 + 2: return runs_fn<T, 12, 2>();
    case 12 * 8 + 4: return runs_fn<T, 12, 4>();
    default: return nullptr;
    }
}

static int runs_width(const ed_sector *s) { return std::max(s->up.kpad, s->dn.kpad); }
static int runs_rows(const ed_sector *s) { return s->up.d <= kBlock ? 1 : s->up.d <= 2 * kBlock ? 2 : 4; }

// Sectors the multi-run kernel takes: real, not S_z, with the packed up-hop table,
// 64 <= d_up <= 1024, table width <= 12, at most kPersistHund Hund transitions per row,
// and the persistent kernel not switched off (ed_lanczos_set_persistent).
static bool runs_capable(const ed_sector *s)
{
    return !s->sz && !s->cplx && s->persist != 0 && s->up.ell_pk && s->up.ncoef > 0 && s->up.d >= 64 &&
           s->up.d <= 4 * kBlock && runs_width(s) <= 12 && (!s->hund || s->hund_maxrow <= kPersistHund);
}

// Return of run_batch_persist when the kernel does not take the batch (not an ED_ code):
// the caller runs the two-launch batch instead.
constexpr int kRunsNotUsed = 1;
// Return of run_fixed when a refused launch held a run that only the multi-run kernel takes
// (d_up below the two-launch batch's 128): the caller runs those sectors one by one.
constexpr int kRunsNeedAlone = 2;

// Workgroups per run.  A run's columns go round by round to its workgroups' 4 column
// groups and a step lasts about as long as the most rounds any run needs, so run J gets
// ceil(d_dn,J / (4 nr)) workgroups for the least round count nr whose total fits the grid
// (gmax), each then raised to a multiple of 8 (the XCD count, so every XCD serves every
// run) while the grid has room.  False when not even one workgroup per run fits.
static bool runs_table(const std::vector<int> &ddn, int gmax, std::vector<int> &nwg)
{
    const int n = (int)ddn.size();
    const int maxd = *std::max_element(ddn.begin(), ddn.end());
    nwg.assign(n, 0);
    int tot = 0;
    for (int nr = 1;; ++nr) {
        tot = 0;
        for (int k = 0; k < n; ++k) {
            nwg[k] = (ddn[k] + kRunCpw * nr - 1) / (kRunCpw * nr);
            tot += nwg[k];
        }
        if (tot <= gmax) break;
        if (nr >= maxd) return false;
    }
    for (int k = 0; k < n; ++k) {
        const int up = std::min((nwg[k] + 7) / 8 * 8, (ddn[k] + kRunCpw - 1) / kRunCpw);
        if (tot - nwg[k] + up <= gmax) {
            tot += up - nwg[k];
            nwg[k] = up;
        }
    }
    return true;
}

// Column map of one run: workgroup w of the run (grid index wg0 + w, on XCD (wg0 + w) mod 8)
// takes in round rd, group g the column cmap[(w * nround + rd) * 4 + g].  The breadth-first
// column order of the down-hop graph is cut into one contiguous share per XCD, in
// proportion to the XCD's workgroups, so the columns an XCD holds are neighbours of each
// other; within a share, each round gives the XCD's column groups consecutive columns.
static int runs_cmap(const std::vector<int32_t> &order, int ddn, int wg0, int nwg, int nround,
                     std::vector<int32_t> &cmap)
{
    cmap.assign((size_t)nwg * nround * kRunCpw, -1);
    std::vector<int> members[8];
    for (int w = 0; w < nwg; ++w) members[(wg0 + w) % 8].push_back(w);
    int off = 0, seen = 0;
    for (int x = 0; x < 8; ++x) {
        const int nx = (int)members[x].size();
        if (!nx) continue;
        seen += nx;
        const int end = (int)((int64_t)ddn * seen / nwg);
        for (int p = 0; p < end - off; ++p) {
            const int rd = p / (kRunCpw * nx), q = p % (kRunCpw * nx);
            if (rd >= nround) return fail(ED_ERR_STATE, "multi-run Lanczos: column map overflow");
            cmap[((size_t)members[x][q / kRunCpw] * nround + rd) * kRunCpw + q % kRunCpw] = order[off + p];
        }
        off = end;
    }
    std::vector<char> hit(ddn, 0);
    for (int32_t c : cmap)
        if (c >= 0) {
            if (c >= ddn || hit[c]) return fail(ED_ERR_STATE, "multi-run Lanczos: column map is not a permutation");
            hit[c] = 1;
        }
    if (off != ddn) return fail(ED_ERR_STATE, "multi-run Lanczos: column map misses columns");
    return ED_OK;
}

// Runs `niter` Lanczos steps of every job in one cooperative launch of k_lz_persist_runs
// (same contract as run_batch); kRunsNotUsed when the kernel does not take the batch or the
// runtime refuses the launch.
template <typename T>
static int run_batch_persist(const std::vector<ed_sector *> &secs, const std::vector<const double *> &h_phi,
                             const std::vector<double> &nrm2, int niter, hipStream_t st, std::vector<double> &a_out,
                             std::vector<double> &b_out, bool phi_on_device)
{
    const int nj = (int)secs.size();
    int K = 4, R = 1;
    for (ed_sector *s : secs) {
        if (!runs_capable(s)) return kRunsNotUsed;
        K = std::max(K, runs_width(s));
        R = std::max(R, runs_rows(s));
    }
    const void *fn = runs_kernel<T>(K, R);
    if (!fn) return kRunsNotUsed;
    // one workgroup per CU (1024 threads, LDS-heavy): the grid is the CU count
    int dev = 0, ncu = 0, coop = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        return kRunsNotUsed;
    }
    std::vector<int> ddn(nj), nwg;
    for (int k = 0; k < nj; ++k) ddn[k] = secs[k]->dn.d;
    if (!runs_table(ddn, std::min(ncu, kBarGroups * kBarGroupSize), nwg)) return kRunsNotUsed;
    size_t lds = 0;
    for (int k = 0; k < nj; ++k) {
        const ed_sector *s = secs[k];
        const int nr = (ddn[k] + kRunCpw * nwg[k] - 1) / (kRunCpw * nwg[k]);
        if (nr > kRunMaxRounds) return kRunsNotUsed;
        lds = std::max(lds, persist_runs_lds(s->up.d, runs_width(s), K, s->up.ncoef, s->ncfg, s->nhund, s->hund, nr,
                                             sizeof(T)).total);
    }
    int per_cu = 0;
    if (lds > 159 * 1024 ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock * kRunCpw, lds) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        return kRunsNotUsed;
    }
    // ---- work space: barrier state, run table, then per run its state, 3 Krylov slots,
    //      start vector, partial sums and column map
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) / 256 * 256; return o; };
    const size_t o_bar = take(sizeof(LzState));
    const size_t o_runs = take(sizeof(PersistRun<T>) * nj);
    std::vector<size_t> o_state(nj), o_basis(nj), o_start(nj), o_part(nj), o_map(nj);
    std::vector<int64_t> ld(nj);
    std::vector<int> ps(nj), nround(nj), wg0(nj);
    std::vector<std::vector<int32_t>> cmaps(nj);
    int g = 0;
    for (int k = 0; k < nj; ++k) {
        ed_sector *s = secs[k];
        ED_TRY(ensure_col_bfs(s));
        wg0[k] = g;
        g += nwg[k];
        nround[k] = (ddn[k] + kRunCpw * nwg[k] - 1) / (kRunCpw * nwg[k]);
        ED_TRY(runs_cmap(s->col_bfs, ddn[k], wg0[k], nwg[k], nround[k], cmaps[k]));
        ld[k] = (s->dim + 63) / 64 * 64;
        ps[k] = (ddn[k] + 63) / 64 * 64;
        o_state[k] = take(sizeof(LzState));
        o_basis[k] = take(sizeof(T) * (size_t)ld[k] * 3);
        o_start[k] = take(sizeof(double) * (size_t)s->dim);
        o_part[k] = take(sizeof(double) * 5 * (size_t)ps[k]);
        o_map[k] = take(sizeof(int32_t) * cmaps[k].size());
    }
    ED_TRY(ws_reserve(off, secs[0]->device));
    char *base = (char *)g_ws.ptr;
    std::vector<PersistRun<T>> runs(nj);
    for (int k = 0; k < nj; ++k) {
        ed_sector *s = secs[k];
        PersistRun<T> &P = runs[k];
        P.a = hv_args_of<T>(s);
        P.a.ell_pk_up = s->up.ell_pk;   // every row of the packed table (hv_args_of drops it for wide coefficient tables)
        P.lz.base = (T *)(base + o_basis[k]);
        P.lz.ld = ld[k];
        P.lz.nslots = 3;
        P.lz.st = (LzState *)(base + o_state[k]);
        P.lz.partA = (double *)(base + o_part[k]);
        P.lz.partB = P.lz.partA + ps[k];
        P.lz.partC = P.lz.partB + 2 * ps[k];
        P.lz.pstride = ps[k];
        P.lz.nA = P.lz.nB = ddn[k];
        P.lz.colmap = 0;
        P.lz.dup = s->up.d;
        P.lz.colord = nullptr;
        P.start = (const double *)(base + o_start[k]);
        P.cmap = (const int32_t *)(base + o_map[k]);
        P.wg0 = wg0[k];
        P.nwg = nwg[k];
        P.nround = nround[k];
        P.kk = runs_width(s);
        ED_HIP_CHECK(hipMemcpyAsync(base + o_start[k], h_phi[k], sizeof(double) * s->dim,
                                    phi_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_scale_into, dim3(egrid(s->dim)), dim3(kBlock), 0, st, (const double *)(base + o_start[k]),
                           (double *)(base + o_start[k]), s->dim, 1.0 / std::sqrt(nrm2[k]));
        ED_HIP_CHECK(hipMemcpyAsync(base + o_map[k], cmaps[k].data(), sizeof(int32_t) * cmaps[k].size(),
                                    hipMemcpyHostToDevice, st));
    }
    const PersistRun<T> *d_runs = (const PersistRun<T> *)(base + o_runs);
    LzState *bar = (LzState *)(base + o_bar);
    ED_HIP_CHECK(hipMemcpyAsync(base + o_runs, runs.data(), sizeof(PersistRun<T>) * nj, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_zero_words, dim3(16), dim3(kBlock), 0, st, (unsigned *)bar, (int)(sizeof(LzState) / 4));
    int nruns = nj, nsteps = niter, spins = kRunsBarrierSpins;
    void *args[] = {&d_runs, &nruns, &bar, &nsteps, &spins};
    hipEvent_t ev[2] = {nullptr, nullptr};
    ED_HIP_CHECK(hipEventCreate(&ev[0]));
    ED_HIP_CHECK(hipEventCreate(&ev[1]));
    struct EvGuard {
        hipEvent_t *e;
        ~EvGuard() { (void)hipEventDestroy(e[0]); (void)hipEventDestroy(e[1]); }
    } ev_guard{ev};
    ED_HIP_CHECK(hipEventRecord(ev[0], st));
    const hipError_t e = secs[0]->persist == 2 ? hipErrorCooperativeLaunchTooLarge
                                               : hipLaunchCooperativeKernel(fn, dim3(g), dim3(kBlock * kRunCpw), args,
                                                                            (unsigned)lds, st);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        // the refusals launch_persist also takes (nothing ran); every other code is an error
        if (e != hipErrorCooperativeLaunchTooLarge && e != hipErrorIllegalState)
            return fail(ED_ERR_HIP, std::string("multi-run Lanczos: hipLaunchCooperativeKernel: ") + hipGetErrorString(e));
        ++g_runs_refused;
        return kRunsNotUsed;
    }
    ED_HIP_CHECK(hipEventRecord(ev[1], st));
    hipLaunchKernelGGL(k_finish_runs<T>, dim3(nj), dim3(kBlock), 0, st, d_runs);
    ED_HIP_CHECK(hipGetLastError());
    ++g_runs_launched;
    std::vector<LzState> hs(nj + 1);
    ED_HIP_CHECK(hipMemcpyAsync(&hs[nj], bar, sizeof(LzState), hipMemcpyDeviceToHost, st));
    for (int k = 0; k < nj; ++k)
        ED_HIP_CHECK(hipMemcpyAsync(&hs[k], base + o_state[k], sizeof(LzState), hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipStreamSynchronize(st));
    if (hs[nj].status == 3) return fail(ED_ERR_HIP, "multi-run Lanczos: grid barrier wait timed out (run stopped)");
    float ms = 0.0f;
    ED_HIP_CHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    g_runs_ms = ms;
    g_runs_job_steps = (double)nj * niter;
    g_runs_bytes = 0.0;
    for (ed_sector *s : secs) g_runs_bytes += 5.0 * sizeof(T) * (double)s->dim * niter;
    a_out.assign((size_t)nj * niter, 0.0);
    b_out.assign((size_t)nj * niter, 0.0);
    for (int k = 0; k < nj; ++k) {
        const LzState &h = hs[k];
        const int reached = h.j_k1 - 1;
        if (h.status == 2) return fail(ED_ERR_BREAKDOWN, "batched Lanczos: zero-norm start vector in job " + std::to_string(k));
        if (h.status == 1 && reached < niter)
            return fail(ED_ERR_BREAKDOWN, "batched Lanczos: zero-norm vector at iteration " + std::to_string(reached + 1) +
                                              " of job " + std::to_string(k) + " (Lanczos_plain.h:85-92)");
        for (int i = 0; i < niter; ++i) {
            a_out[(size_t)k * niter + i] = h.alpha[i];
            b_out[(size_t)k * niter + i] = h.beta[i];
        }
    }
    return ED_OK;
}

}  // namespace ed

using namespace ed;

extern "C" int ed_dynamic(ed_sector *sec, int iph, double energy, const double *phi, int nitergreen,
                          double cutoff_dynamic, int precision, int nw, const double *wn, double *dyn, double *a,
                          double *b);

namespace ed {
int dynamic_dev(ed_sector *sec, int iph, double energy, const double *d_phi, double nrm2, int nitergreen,
                double cutoff_dynamic, int prec, int nw, const double *d_wn, double *dyn, double *a_out, double *b_out);
}

// ||phi_k||^2 of device vectors: partial sums on the device, added on the host in order.
static int device_norms(int njobs, ed_sector *const *secs, const double *const *phis, std::vector<double> &nrm2)
{
    constexpr int kParts = 256;
    hipStream_t st = secs[0]->stream;
    double *d_part = nullptr;
    ED_HIP_CHECK(hipMallocAsync((void **)&d_part, sizeof(double) * kParts * njobs, st));
    for (int k = 0; k < njobs; ++k)
        hipLaunchKernelGGL(k_norm2_part, dim3(kParts), dim3(kBlock), 0, st, phis[k], secs[k]->dim, d_part + k * kParts);
    std::vector<double> part((size_t)kParts * njobs);
    ED_HIP_CHECK(hipMemcpyAsync(part.data(), d_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipStreamSynchronize(st));
    (void)hipFreeAsync(d_part, st);
    for (int k = 0; k < njobs; ++k) {
        double v = 0.0;
        for (int q = 0; q < kParts; ++q) v += part[(size_t)k * kParts + q];
        nrm2[k] = v;
    }
    return ED_OK;
}

// Fixed-length recurrences of niter steps from the start vectors bp (device when
// phi_on_device, norms^2 bn) in the sectors bs, all advancing together: the runs the
// multi-run persistent kernel takes in cooperative launches cut (in order) so that each
// fills one workgroup per CU at kRunMaxRounds rounds, then the rest (or a group the
// runtime refuses) in the two-launch batch.  ab, bb: [run][niter] alpha, beta.
static int run_fixed(const std::vector<ed_sector *> &bs, const std::vector<const double *> &bp,
                     const std::vector<double> &bn, int niter, int precision, hipStream_t st, std::vector<double> &ab,
                     std::vector<double> &bb, bool phi_on_device)
{
    const int nb = (int)bs.size();
    ab.assign((size_t)nb * niter, 0.0);
    bb.assign((size_t)nb * niter, 0.0);
    // groups: the runs the multi-run kernel takes, cut (in order) into launches that each fit
    // one workgroup per CU at kRunMaxRounds rounds of kRunCpw columns; then the rest
    std::vector<std::vector<int>> grp;
    std::vector<int> rest;
    {
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            ncu = 0;
        (void)hipGetLastError();
        const int gmax = std::min(ncu, kBarGroups * kBarGroupSize);
        int used = 0;
        for (int i = 0; i < nb; ++i) {
            const int need = (bs[i]->dn.d + kRunCpw * kRunMaxRounds - 1) / (kRunCpw * kRunMaxRounds);
            if (!runs_capable(bs[i]) || need > gmax) {
                rest.push_back(i);
                continue;
            }
            if (grp.empty() || used + need > gmax) {
                grp.emplace_back();
                used = 0;
            }
            grp.back().push_back(i);
            used += need;
        }
    }
    const int npersist = (int)grp.size();
    if (!rest.empty()) grp.push_back(rest);
    int rc = ED_OK;
    for (int gi = 0; gi < (int)grp.size(); ++gi) {
        if (grp[gi].empty()) continue;
        std::vector<ed_sector *> gs;
        std::vector<const double *> gp;
        std::vector<double> gn, ga, gb;
        for (int i : grp[gi]) {
            gs.push_back(bs[i]);
            gp.push_back(bp[i]);
            gn.push_back(bn[i]);
        }
        rc = kRunsNotUsed;
        if (gi < npersist)
            rc = precision == ED_PREC_F64 ? run_batch_persist<double>(gs, gp, gn, niter, st, ga, gb, phi_on_device)
                                          : run_batch_persist<float>(gs, gp, gn, niter, st, ga, gb, phi_on_device);
        if (rc == kRunsNotUsed) {
            for (ed_sector *x : gs)
                if (!rows_capable(x)) return kRunsNeedAlone;
            rc = precision == ED_PREC_F64 ? run_batch<double>(gs, gp, gn, niter, st, ga, gb, phi_on_device)
                                          : run_batch<float>(gs, gp, gn, niter, st, ga, gb, phi_on_device);
        }
        if (rc != ED_OK) return rc;
        for (size_t q = 0; q < grp[gi].size(); ++q) {
            std::copy(ga.begin() + q * niter, ga.begin() + (q + 1) * niter, ab.begin() + (size_t)grp[gi][q] * niter);
            std::copy(gb.begin() + q * niter, gb.begin() + (q + 1) * niter, bb.begin() + (size_t)grp[gi][q] * niter);
        }
    }
    return rc;
}

static int dynamic_batched(int njobs, ed_sector *const *secs, const double *const *phis, const int *iph,
                           const double *energy, int nitergreen, double cutoff_dynamic, int precision, int nw,
                           const double *wn, double *dyn, double *a, double *b, bool phi_on_device)
{
    if (njobs < 1 || !secs || !phis || !iph || !energy || !wn || !dyn || nw < 1 || nitergreen < 1)
        return fail(ED_ERR_ARG, "ed_dynamic_batched: bad argument");
    if (precision != ED_PREC_F64 && precision != ED_PREC_F32) return fail(ED_ERR_ARG, "ed_dynamic_batched: bad precision");
    std::lock_guard<std::mutex> lock(g_ws.mu);
    for (int k = 0; k < njobs; ++k) {
        if (!secs[k] || !phis[k]) return fail(ED_ERR_ARG, "ed_dynamic_batched: NULL sector or vector");
        if (iph[k] != 1 && iph[k] != 2) return fail(ED_ERR_ARG, "ed_dynamic_batched: iph must be 1 or 2");
        if (secs[k]->device != secs[0]->device) return fail(ED_ERR_ARG, "ed_dynamic_batched: sectors on different devices");
    }
    ED_HIP_CHECK(hipSetDevice(secs[0]->device));
    std::fill(dyn, dyn + 2 * (size_t)njobs * nw, 0.0);
    if (a) std::fill(a, a + (size_t)njobs * nitergreen, 0.0);
    if (b) std::fill(b, b + (size_t)njobs * nitergreen, 0.0);
    // Group: all jobs with the same Lanczos length that the batched kernel can take run
    // together; the others (short columns, very wide tables) run one by one.
    std::vector<double> nrm2(njobs, 0.0);
    if (phi_on_device) {
        ED_TRY(device_norms(njobs, secs, phis, nrm2));
    } else {
        for (int k = 0; k < njobs; ++k) nrm2[k] = host_norm2(phis[k], secs[k]->dim);
    }
    std::vector<int> batch;
    double *d_wn1 = nullptr;   // frequencies on the device for the one-by-one runs of device vectors
    for (int k = 0; k < njobs; ++k) {
        if (nrm2[k] < cutoff_dynamic) continue;   // dyn = 0 (green_class_compute_dynamic.f90:263-265)
        const int niter = (int)std::min<int64_t>(nitergreen, secs[k]->cplx ? secs[k]->cdim : secs[k]->dim);
        if (rows_capable(secs[k]) && niter == nitergreen) {
            batch.push_back(k);
        } else if (phi_on_device) {
            hipStream_t sk = secs[k]->stream;
            if (!d_wn1) {
                ED_HIP_CHECK(hipMalloc((void **)&d_wn1, sizeof(double) * nw));
                ED_HIP_CHECK(hipMemcpy(d_wn1, wn, sizeof(double) * nw, hipMemcpyHostToDevice));
            }
            ED_HIP_CHECK(hipStreamSynchronize(sk));
            int rc = dynamic_dev(secs[k], iph[k], energy[k], phis[k], nrm2[k], nitergreen, cutoff_dynamic, precision, nw,
                                 d_wn1, dyn + (size_t)k * 2 * nw, a ? a + (size_t)k * nitergreen : nullptr,
                                 b ? b + (size_t)k * nitergreen : nullptr);
            if (rc != ED_OK) {
                (void)hipFree(d_wn1);
                return rc;
            }
        } else {
            int rc = ed_dynamic(secs[k], iph[k], energy[k], phis[k], nitergreen, cutoff_dynamic, precision, nw, wn,
                                dyn + (size_t)k * 2 * nw, a ? a + (size_t)k * nitergreen : nullptr,
                                b ? b + (size_t)k * nitergreen : nullptr);
            if (rc != ED_OK) return rc;
        }
    }
    if (d_wn1) (void)hipFree(d_wn1);
    if (batch.empty()) return ED_OK;
    const int nb = (int)batch.size();
    std::vector<ed_sector *> bs(nb);
    std::vector<const double *> bp(nb);
    std::vector<double> bn(nb);
    for (int i = 0; i < nb; ++i) {
        bs[i] = secs[batch[i]];
        bp[i] = phis[batch[i]];
        bn[i] = nrm2[batch[i]];
    }
    hipStream_t st = bs[0]->stream;
    const int niter = nitergreen;
    std::vector<double> ab, bb;
    int rc = run_fixed(bs, bp, bn, niter, precision, st, ab, bb, phi_on_device);
    if (rc != ED_OK) return rc;
    // shifted tridiagonals: ph (alpha - E) (green_class_compute_dynamic.f90:310-311)
    std::vector<double> sa((size_t)nb * niter), scale(nb);
    for (int i = 0; i < nb; ++i) {
        const int k = batch[i];
        const double ph = iph[k] == 1 ? 1.0 : -1.0;
        for (int t = 0; t < niter; ++t) sa[(size_t)i * niter + t] = ph * (ab[(size_t)i * niter + t] - energy[k]);
        scale[i] = nrm2[k];
        if (a) std::copy(ab.begin() + (size_t)i * niter, ab.begin() + (size_t)(i + 1) * niter, a + (size_t)k * nitergreen);
        if (b) std::copy(bb.begin() + (size_t)i * niter, bb.begin() + (size_t)(i + 1) * niter, b + (size_t)k * nitergreen);
    }
    double *d = nullptr;
    const size_t tot = 2 * (size_t)nb * niter + nb + nw + 2 * (size_t)nb * nw;
    ED_HIP_CHECK(hipMallocAsync((void **)&d, sizeof(double) * tot, st));
    double *d_a = d, *d_b = d + (size_t)nb * niter, *d_sc = d_b + (size_t)nb * niter, *d_wn = d_sc + nb,
           *d_out = d_wn + nw;
    ED_HIP_CHECK(hipMemcpyAsync(d_a, sa.data(), sizeof(double) * nb * niter, hipMemcpyHostToDevice, st));
    ED_HIP_CHECK(hipMemcpyAsync(d_b, bb.data(), sizeof(double) * nb * niter, hipMemcpyHostToDevice, st));
    ED_HIP_CHECK(hipMemcpyAsync(d_sc, scale.data(), sizeof(double) * nb, hipMemcpyHostToDevice, st));
    ED_HIP_CHECK(hipMemcpyAsync(d_wn, wn, sizeof(double) * nw, hipMemcpyHostToDevice, st));
    rc = cfrac_batch_dev(niter, d_a, d_b, niter, nb, nw, d_wn, d_sc, d_out, st);
    std::vector<double> out(2 * (size_t)nb * nw);
    if (rc == ED_OK) {
        ED_HIP_CHECK(hipMemcpyAsync(out.data(), d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost, st));
        ED_HIP_CHECK(hipStreamSynchronize(st));
        for (int i = 0; i < nb; ++i)
            std::copy(out.begin() + (size_t)i * 2 * nw, out.begin() + (size_t)(i + 1) * 2 * nw,
                      dyn + (size_t)batch[i] * 2 * nw);
    }
    (void)hipFreeAsync(d, st);
    (void)hipStreamSynchronize(st);
    return rc;
}

// Lanczos_GPU_updo for many sectors at once (ed_lanczos_batched).
extern "C" int ed_lanczos_batched(int nsec, ed_sector *const *secs, int niter, int precision, double *alpha,
                                  double *beta, int *steps)
{
    if (nsec < 1 || !secs || niter < 1 || niter > kMaxIter - 2 || !alpha || !beta || !steps)
        return fail(ED_ERR_ARG, "ed_lanczos_batched: bad argument");
    if (precision != ED_PREC_F64 && precision != ED_PREC_F32)
        return fail(ED_ERR_ARG, "ed_lanczos_batched: precision must be ED_PREC_F64 or ED_PREC_F32");
    for (int k = 0; k < nsec; ++k) {
        if (!secs[k]) return fail(ED_ERR_ARG, "ed_lanczos_batched: NULL sector");
        if (secs[k]->device != secs[0]->device) return fail(ED_ERR_ARG, "ed_lanczos_batched: sectors on different devices");
    }
    ED_HIP_CHECK(hipSetDevice(secs[0]->device));
    std::fill(alpha, alpha + (size_t)nsec * niter, 0.0);
    std::fill(beta, beta + (size_t)nsec * niter, 0.0);
    // one sector alone: a fixed-length run from create_fix_initial_vector's vector in a
    // shared work space (no Krylov basis, scalars or partials of the sector's own, no
    // autotuning), its breakdown step reported
    auto alone = [&](int k) -> int {
        ed_sector *x = secs[k];
        std::lock_guard<std::mutex> lk(g_alone.mu);
        const int64_t ld = (x->dim + 63) / 64 * 64;
        const size_t o_part = al256(sizeof(LzState)), n_part = (size_t)kPartC + 2 * 8192;
        const size_t o_basis = al256(o_part + sizeof(double) * n_part);
        const size_t o_start = al256(o_basis + sizeof(double) * 3 * (size_t)ld);
        ED_TRY(ws_reserve_in(g_alone, o_start + sizeof(double) * (size_t)x->dim, x->device));
        char *base = (char *)g_alone.ptr;
        double *d_start = (double *)(base + o_start);
        ED_TRY(start_vector_dev(x->dim, d_start, x->stream));
        const LzWork w{base + o_basis, ld, 3, (LzState *)base, (double *)(base + o_part)};
        int at = 0;
        int rc = lanczos_run_ws(x, w, niter, d_start, precision, alpha + (size_t)k * niter, beta + (size_t)k * niter,
                                &at);
        if (rc == ED_ERR_BREAKDOWN && at > 0) {
            steps[k] = at;
            return ED_OK;
        }
        steps[k] = niter;
        return rc;
    };
    std::vector<int> batch;
    for (int k = 0; k < nsec; ++k) {
        const int64_t d = secs[k]->cplx ? secs[k]->cdim : secs[k]->dim;
        if ((rows_capable(secs[k]) || runs_capable(secs[k])) && d > niter) batch.push_back(k);
        else ED_TRY(alone(k));
    }
    if (batch.empty()) return ED_OK;
    std::lock_guard<std::mutex> lock(g_ws.mu);
    const int nb = (int)batch.size();
    std::vector<ed_sector *> bs(nb);
    std::vector<size_t> off(nb + 1, 0);
    for (int i = 0; i < nb; ++i) {
        bs[i] = secs[batch[i]];
        off[i + 1] = off[i] + (size_t)bs[i]->dim;
    }
    hipStream_t st = bs[0]->stream;
    // every start vector (create_fix_initial_vector's) on the device, then one batched run
    double *d = nullptr;
    ED_HIP_CHECK(hipMalloc((void **)&d, sizeof(double) * off[nb]));
    std::vector<const double *> bp(nb);
    int rc = ED_OK;
    for (int i = 0; i < nb && rc == ED_OK; ++i) {
        bp[i] = d + off[i];
        rc = start_vector_dev(bs[i]->dim, d + off[i], st);
    }
    std::vector<double> bn(nb, 0.0);
    if (rc == ED_OK) rc = device_norms(nb, bs.data(), bp.data(), bn);
    std::vector<double> ab, bb;
    if (rc == ED_OK) rc = run_fixed(bs, bp, bn, niter, precision, st, ab, bb, true);
    (void)hipStreamSynchronize(st);
    (void)hipFree(d);
    if (rc == ED_ERR_BREAKDOWN || rc == kRunsNeedAlone) {
        // a Krylov space closed before niter, or a refused launch held a run only the
        // multi-run kernel takes: each sector alone
        for (int i = 0; i < nb; ++i) ED_TRY(alone(batch[i]));
        return ED_OK;
    }
    if (rc != ED_OK) return rc;
    for (int i = 0; i < nb; ++i) {
        const int k = batch[i];
        std::copy(ab.begin() + (size_t)i * niter, ab.begin() + (size_t)(i + 1) * niter, alpha + (size_t)k * niter);
        std::copy(bb.begin() + (size_t)i * niter, bb.begin() + (size_t)(i + 1) * niter, beta + (size_t)k * niter);
        beta[(size_t)k * niter] = std::sqrt(bn[i]);   // subdiag(1) = ||v_1||, as ed_lanczos_run
        steps[k] = niter;
    }
    return ED_OK;
}

extern "C" int ed_dynamic_batched_stats(int *persistent_launches, int *refusals, double *last_ms,
                                        double *last_job_steps, double *last_bytes)
{
    std::lock_guard<std::mutex> lock(g_ws.mu);
    if (persistent_launches) *persistent_launches = g_runs_launched;
    if (refusals) *refusals = g_runs_refused;
    if (last_ms) *last_ms = g_runs_ms;
    if (last_job_steps) *last_job_steps = g_runs_job_steps;
    if (last_bytes) *last_bytes = g_runs_bytes;
    return ED_OK;
}

extern "C" int ed_dynamic_batched(int njobs, ed_sector *const *secs, const double *const *phis, const int *iph,
                                  const double *energy, int nitergreen, double cutoff_dynamic, int precision, int nw,
                                  const double *wn, double *dyn, double *a, double *b)
{
    return dynamic_batched(njobs, secs, phis, iph, energy, nitergreen, cutoff_dynamic, precision, nw, wn, dyn, a, b,
                           false);
}

namespace ed {
int apply_c_dev(const ed_sector *src, const ed_sector *dst, int site, int spin, int pm, const double *d_psi,
                double *d_phi, hipStream_t st, int ncond);
}

// y += s x
__global__ void k_axpy(double s, const double *x, double *y, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] += s * x[i];
}

// The excitation runs of several eigenstates at once: every state's vector copied to the
// device once, job k's vector built there from state state_of[k], every run advanced in
// one batched call (dynamic_batched: one launch of the multi-run persistent kernel for all
// the runs it takes).
static int dynamic_from_states(int nst, ed_sector *const *gs, const double *const *psi, const double *energy,
                               int njobs, const int *state_of, ed_sector *const *dst, const int *iph, const int *nterm,
                               const int *sites, const double *coefs, int spin, int nitergreen, double cutoff_dynamic,
                               int precision, int nw, const double *wn, double *dyn)
{
    if (nst < 1 || !gs || !psi || !energy || njobs < 1 || !state_of || !dst || !iph || !nterm || !sites || !coefs ||
        !dyn)
        return fail(ED_ERR_ARG, "ed_dynamic_from_states: bad argument");
    for (int q = 0; q < nst; ++q) {
        if (!gs[q] || !psi[q]) return fail(ED_ERR_ARG, "ed_dynamic_from_states: NULL state sector or vector");
        if (gs[q]->cplx) return fail(ED_ERR_ARG, "ed_dynamic_from_state: real sectors only");
        if (gs[q]->device != gs[0]->device) return fail(ED_ERR_ARG, "ed_dynamic_from_states: sectors on different devices");
    }
    ED_HIP_CHECK(hipSetDevice(gs[0]->device));
    hipStream_t st = gs[0]->stream;
    // every psi once on the device, then every excitation vector built there
    size_t tot = 0, maxd = 1;
    std::vector<size_t> o_psi(nst);
    for (int q = 0; q < nst; ++q) {
        o_psi[q] = tot;
        tot += (size_t)gs[q]->dim;
    }
    for (int k = 0; k < njobs; ++k) {
        if (!dst[k] || nterm[k] < 1 || state_of[k] < 0 || state_of[k] >= nst)
            return fail(ED_ERR_ARG, "ed_dynamic_from_states: bad job");
        tot += (size_t)dst[k]->dim;
        maxd = std::max<size_t>(maxd, (size_t)dst[k]->dim);
    }
    double *d = nullptr;
    ED_HIP_CHECK(hipMalloc((void **)&d, sizeof(double) * (tot + maxd)));
    double *d_tmp = d + tot;
    std::vector<const double *> phis(njobs);
    std::vector<double> en(njobs);
    int rc = ED_OK;
    for (int q = 0; q < nst && rc == ED_OK; ++q)
        if (hipMemcpyAsync(d + o_psi[q], psi[q], sizeof(double) * gs[q]->dim, hipMemcpyHostToDevice, st) != hipSuccess)
            rc = fail(ED_ERR_HIP, "ed_dynamic_from_states: copy of psi");
    size_t off = o_psi[nst - 1] + (size_t)gs[nst - 1]->dim;
    int t0 = 0;
    for (int k = 0; k < njobs && rc == ED_OK; ++k) {
        const int q = state_of[k];
        const ed_sector *g = gs[q];
        const double *d_psi = d + o_psi[q];
        double *phi = d + off;
        off += dst[k]->dim;
        phis[k] = phi;
        en[k] = energy[q];
        const int pm = iph[k] == 1 ? +1 : -1;
        if (iph[k] != 1 && iph[k] != 2) rc = fail(ED_ERR_ARG, "ed_dynamic_from_state: iph must be 1 or 2");
        // sum_t coef_t c^(pm)_{site_t} psi (the vectors of symmetric_combineAA,
        // green_class_compute_symmetric_combination.f90:39-137)
        for (int t = 0; t < nterm[k] && rc == ED_OK; ++t) {
            // S_z sectors: sites[] are Nambu orbitals spin * nc + site (apply_Csz), `spin`
            // unused; ED_SITE_NCOND on a Nambu orbital asks for the correlated excitation of
            // apply_Csz_N_updo (apply_C.f90:404-483)
            int site = sites[t0 + t], tspin = spin;
            const int ncond = (site & ED_SITE_NCOND) ? 1 : 0;
            site &= ~ED_SITE_NCOND;
            if (ncond && !g->sz) {
                rc = fail(ED_ERR_ARG, "ed_dynamic_from_state: the correlated excitation exists for S_z sectors only");
                break;
            }
            if (g->sz) {
                if (site < 0 || site >= 2 * g->nc) {
                    rc = fail(ED_ERR_ARG, "ed_dynamic_from_state: Nambu orbital outside [0, 2 nc)");
                    break;
                }
                tspin = site / g->nc;
                site %= g->nc;
            }
            const double c = coefs[t0 + t];
            if (t == 0 && c == 1.0) {
                rc = apply_c_dev(g, dst[k], site, tspin, pm, d_psi, phi, st, ncond);
            } else {
                rc = apply_c_dev(g, dst[k], site, tspin, pm, d_psi, d_tmp, st, ncond);
                if (rc == ED_OK) {
                    if (t == 0) ED_HIP_CHECK(hipMemsetAsync(phi, 0, sizeof(double) * dst[k]->dim, st));
                    hipLaunchKernelGGL(k_axpy, dim3(egrid(dst[k]->dim)), dim3(kBlock), 0, st, c, (const double *)d_tmp,
                                       phi, dst[k]->dim);
                }
            }
        }
        t0 += nterm[k];
    }
    if (rc == ED_OK && hipStreamSynchronize(st) != hipSuccess) rc = fail(ED_ERR_HIP, "ed_dynamic_from_state: apply_C");
    if (rc == ED_OK) {
        std::vector<ed_sector *> ds(dst, dst + njobs);
        rc = dynamic_batched(njobs, ds.data(), phis.data(), iph, en.data(), nitergreen, cutoff_dynamic, precision, nw,
                             wn, dyn, nullptr, nullptr, true);
    }
    (void)hipDeviceSynchronize();
    (void)hipFree(d);
    return rc;
}

extern "C" int ed_dynamic_from_state(ed_sector *gs, const double *psi, double energy, int njobs, ed_sector *const *dst,
                                     const int *iph, const int *nterm, const int *sites, const double *coefs, int spin,
                                     int nitergreen, double cutoff_dynamic, int precision, int nw, const double *wn,
                                     double *dyn)
{
    if (!gs || !psi || njobs < 1) return fail(ED_ERR_ARG, "ed_dynamic_from_state: bad argument");
    std::vector<int> state_of(njobs, 0);
    return dynamic_from_states(1, &gs, &psi, &energy, njobs, state_of.data(), dst, iph, nterm, sites, coefs, spin,
                               nitergreen, cutoff_dynamic, precision, nw, wn, dyn);
}

extern "C" int ed_dynamic_from_states(int nstates, ed_sector *const *gs, const double *const *psi,
                                      const double *energy, int njobs, const int *state_of, ed_sector *const *dst,
                                      const int *iph, const int *nterm, const int *sites, const double *coefs,
                                      int spin, int nitergreen, double cutoff_dynamic, int precision, int nw,
                                      const double *wn, double *dyn)
{
    return dynamic_from_states(nstates, gs, psi, energy, njobs, state_of, dst, iph, nterm, sites, coefs, spin,
                               nitergreen, cutoff_dynamic, precision, nw, wn, dyn);
}

