// Sector construction on the device: Fock-sector enumeration, rank tables and the
// quadratic-part tables of tab_HAIM2, all as integer kernels, bit-exact with the
// reference.
//
//   k_enum_states   quantum_fermion_hilbert_class.f90:346-421 (states of popcount N in
//                   increasing integer order) -- here by combinatorial unranking: the
//                   k-th N-subset in colexicographic order, which is increasing order.
//   k_rank_table    rank(state) -- here the combinatorial number system
//                   sum_i C(c_i, i) over the set bits c_1 < c_2 < ...
//   k_tab_count /   tab_HAIM2 count and fill passes (HAIM_offdiag_arrays.f90:89-148,
//   k_tab_fill      210-606), same entry order: impurity hops, bath hops, V b+c, V* c+b.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "ed_kernels.h"

namespace ed {

__host__ __device__ inline int binom(int n, int k)
{
    if (k < 0 || n < k) return 0;
    if (k > n - k) k = n - k;
    int r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

__global__ void k_enum_states(int ns, int n, int d, int32_t *states)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d) return;
    int r = t;
    uint32_t s = 0;
    int c = ns - 1;
    for (int i = n; i >= 1; --i) {
        while (binom(c, i) > r) --c;
        s |= 1u << c;
        r -= binom(c, i);
        --c;
    }
    states[t] = (int32_t)s;
}

__global__ void k_rank_table(int ns, int32_t *rank)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (1 << ns)) return;
    int r = 0, i = 0;
    for (int c = 0; c < ns; ++c)
        if ((s >> c) & 1) r += binom(c, ++i);
    rank[s] = r;
}

static SpinModel spin_model(int nc, int nb, const double *ec, const double *eb, const double *vbc, double cut)
{
    SpinModel m{};
    m.nimp = nc;
    m.nbath = nb;
    for (int c = 0; c < nc; ++c) m.ibit[c] = nb + c;
    for (int b = 0; b < nb; ++b) m.bbit[b] = b;
    m.ec = ec;
    m.eb = eb;
    m.vbc = vbc;
    m.ec_im = m.eb_im = m.vbc_im = nullptr;
    m.cut = cut;
    return m;
}

// tab_HAIM2 (HAIM_offdiag_arrays.f90:419-527) for one spin, one thread per state: the
// diagonal and the entries (target rank, <target|H|state>) in the reference order.  For a
// complex model off_im receives the imaginary parts: Ec(a,b) sign, Eb(a,b) sign,
// Vbc(b,c) sign (impurity -> bath) and conj(Vbc(b,c)) sign (bath -> impurity).
template <bool FILL>
__global__ void k_tab(SpinModel m, const int32_t *states, int d, const int32_t *rank,
                      const int32_t *ptr, int32_t *noff, double *diag, int32_t *rk, double *off, double *off_im)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d) return;
    const int nc = m.nimp, nb = m.nbath;
    const bool cx = m.ec_im != nullptr;
    const uint32_t s = (uint32_t)states[i];
    int k = FILL ? ptr[i] : 0;
    const int k0 = k;
    uint32_t o;
    int sg;
    if (FILL) {
        double dg = 0.0;
        for (int c = 0; c < nc; ++c)
            if ((s >> m.ibit[c]) & 1u) dg += m.ec[c * nc + c];
        for (int b = 0; b < nb; ++b)
            if ((s >> m.bbit[b]) & 1u) dg += m.eb[b * nb + b];
        diag[i] = dg;
    }
    auto put = [&](double tr, double ti) {
        if (FILL) {
            rk[k] = rank[o];
            off[k] = tr * sg;
            if (cx) off_im[k] = ti * sg;
        }
        ++k;
    };
    for (int a = 0; a < nc; ++a)
        for (int b = 0; b < nc; ++b) {
            const double t = m.ec[a * nc + b], ti = cx ? m.ec_im[a * nc + b] : 0.0;
            if (a == b || !(cx ? keep_coef(t, ti, m.cut) : (t != 0.0 && fabs(t) > m.cut))) continue;
            if (hop(s, m.ibit[a], m.ibit[b], &o, &sg)) put(t, ti);
        }
    for (int a = 0; a < nb; ++a)
        for (int b = 0; b < nb; ++b) {
            const double t = m.eb[a * nb + b], ti = cx ? m.eb_im[a * nb + b] : 0.0;
            if (a == b || !(cx ? keep_coef(t, ti, m.cut) : (t != 0.0 && fabs(t) > m.cut))) continue;
            if (hop(s, m.bbit[a], m.bbit[b], &o, &sg)) put(t, ti);
        }
    for (int b = 0; b < nb; ++b)
        for (int c = 0; c < nc; ++c) {
            const double t = m.vbc[b * nc + c], ti = cx ? m.vbc_im[b * nc + c] : 0.0;
            if (!(cx ? keep_coef(t, ti, m.cut) : (t != 0.0 && fabs(t) > m.cut))) continue;
            if (hop(s, m.bbit[b], m.ibit[c], &o, &sg)) put(t, ti);
            if (hop(s, m.ibit[c], m.bbit[b], &o, &sg)) put(t, -ti);
        }
    if (!FILL) noff[i] = k - k0;
}

// Exclusive prefix sum of noff[0..d) into ptr[0..d], one workgroup; also the row max.
__global__ void k_scan(const int32_t *noff, int d, int32_t *ptr, int32_t *kmax)
{
    __shared__ int32_t part[kBlock], mx[kBlock];
    const int t = threadIdx.x;
    const int per = (d + kBlock - 1) / kBlock;
    const int lo = min(d, t * per), hi = min(d, lo + per);
    int32_t s = 0, m = 0;
    for (int i = lo; i < hi; ++i) { s += noff[i]; m = max(m, noff[i]); }
    part[t] = s;
    mx[t] = m;
    __syncthreads();
    if (t == 0) {
        int32_t acc = 0, mm = 0;
        for (int i = 0; i < kBlock; ++i) { int32_t v = part[i]; part[i] = acc; acc += v; mm = max(mm, mx[i]); }
        ptr[d] = acc;
        *kmax = mm;
    }
    __syncthreads();
    int32_t acc = part[t];
    for (int i = lo; i < hi; ++i) { ptr[i] = acc; acc += noff[i]; }
}

__global__ void k_ell(int d, int kmax, const int32_t *ptr, const int32_t *rk, const double *off,
                      int32_t *ell_rk, double *ell_val, const double *off_im, double *ell_vim)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d) return;
    int p = ptr[i], n = ptr[i + 1] - p;
    for (int k = 0; k < kmax; ++k) {
        ell_rk[(int64_t)k * d + i] = k < n ? rk[p + k] : i;
        ell_val[(int64_t)k * d + i] = k < n ? off[p + k] : 0.0;
        if (off_im) ell_vim[(int64_t)k * d + i] = k < n ? off_im[p + k] : 0.0;
    }
}

// imp[i] = impurity bits of state i; swap[i][c] = rank of the state whose bath bits
// are those of state i and whose impurity bits are c (-1 if the popcount differs).
__global__ void k_imp_swap(const int32_t *states, int d, int nb, int ncfg, const int32_t *rank,
                           int32_t *imp, int32_t *swp)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d * ncfg) return;
    int i = t / ncfg, c = t - i * ncfg;
    uint32_t s = (uint32_t)states[i];
    uint32_t bath = s & ((1u << nb) - 1u);
    uint32_t s2 = bath | ((uint32_t)c << nb);
    swp[t] = (__popc(s2) == __popc(s)) ? rank[s2] : -1;
    if (c == 0) imp[i] = (int32_t)(s >> nb);
}

__global__ void k_to_f32(const double *in, float *out, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

static inline int nblocks(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

template <typename T>
static int dalloc(T **p, size_t n)
{
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) return fail(ED_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return ED_OK;
}

static void free_spin(SpinTables &t)
{
    void *ps[] = {t.states, t.diag, t.noff, t.ptr, t.rk, t.off, t.ell_rk, t.ell_val, t.imp, t.swap,
                  t.diag32, t.off32, t.ell_val32, t.ell_pk, t.coef, t.coef32, t.off_im, t.ell_vim};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    t = SpinTables();
}

// Packed ELL of one spin: entry (k, i) = rank | cidx << 16, where cidx indexes the
// table of the distinct coefficient values of the spin (index 0 holds 0.0 and pads rows
// shorter than kpad).  The values are copies of tab_HAIM2's, so the products are the
// same bits as with the plain ELL.  Built on the host from the device tables.
// Packed ELL of one spin with `rows` >= kpad rows (the unrolled width of the fast H.v,
// the wider of the two spins' widths; rows beyond kpad hold the zero coefficient).
static int build_packed(SpinTables &T, int rows, hipStream_t st)
{
    const int d = T.d, kw = rows;
    std::vector<int32_t> ptr(d + 1), rk(std::max<int64_t>(1, T.nnz));
    std::vector<double> off(std::max<int64_t>(1, T.nnz));
    ED_HIP_CHECK(hipMemcpy(ptr.data(), T.ptr, sizeof(int32_t) * (d + 1), hipMemcpyDeviceToHost));
    if (T.nnz) {
        ED_HIP_CHECK(hipMemcpy(rk.data(), T.rk, sizeof(int32_t) * T.nnz, hipMemcpyDeviceToHost));
        ED_HIP_CHECK(hipMemcpy(off.data(), T.off, sizeof(double) * T.nnz, hipMemcpyDeviceToHost));
    }
    std::vector<double> coef(1, 0.0);
    std::unordered_map<uint64_t, uint32_t> index;
    index[0] = 0;
    std::vector<uint32_t> pk((size_t)kw * d);
    T.ncoef = 0;
    if (d > 65535) return ED_OK;   // ranks do not fit 16 bits: the packed kernel is not used
    for (int i = 0; i < d; ++i) {
        const int n = ptr[i + 1] - ptr[i];
        for (int k = 0; k < kw; ++k) {
            uint32_t r = (uint32_t)i, c = 0;
            if (k < n) {
                const double v = off[ptr[i] + k];
                uint64_t b;
                std::memcpy(&b, &v, sizeof b);
                auto it = index.find(b);
                if (it == index.end()) {
                    if (coef.size() > 65535) return ED_OK;
                    it = index.emplace(b, (uint32_t)coef.size()).first;
                    coef.push_back(v);
                }
                c = it->second;
                r = (uint32_t)rk[ptr[i] + k];
            }
            pk[(size_t)k * d + i] = r | (c << 16);
        }
    }
    std::vector<float> coef32(coef.begin(), coef.end());
    ED_TRY(dalloc(&T.ell_pk, pk.size()));
    ED_TRY(dalloc(&T.coef, coef.size()));
    ED_TRY(dalloc(&T.coef32, coef.size()));
    ED_HIP_CHECK(hipMemcpyAsync(T.ell_pk, pk.data(), sizeof(uint32_t) * pk.size(), hipMemcpyHostToDevice, st));
    ED_HIP_CHECK(hipMemcpyAsync(T.coef, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice, st));
    ED_HIP_CHECK(hipMemcpyAsync(T.coef32, coef32.data(), sizeof(float) * coef32.size(), hipMemcpyHostToDevice, st));
    ED_HIP_CHECK(hipStreamSynchronize(st));
    T.ncoef = (int)coef.size();
    return ED_OK;
}

static int build_spin(ed_sector *S, int spin, int n, SpinTables &T)
{
    const int ns = S->ns, nc = S->nc, nb = S->nb;
    hipStream_t st = S->stream;
    T.d = binom(ns, n);
    const int d = T.d;
    ED_TRY(dalloc(&T.states, d));
    ED_TRY(dalloc(&T.diag, d));
    ED_TRY(dalloc(&T.noff, d));
    ED_TRY(dalloc(&T.ptr, d + 1));
    hipLaunchKernelGGL(k_enum_states, dim3(nblocks(d)), dim3(kBlock), 0, st, ns, n, d, T.states);
    ED_HIP_CHECK(hipGetLastError());

    // model matrices of this spin on the device: re parts, then (complex model) im parts
    const int nmat = S->cplx ? 2 : 1;
    double *d_ec = nullptr, *d_eb = nullptr, *d_vbc = nullptr;
    ED_TRY(dalloc(&d_ec, (size_t)nmat * nc * nc));
    ED_TRY(dalloc(&d_eb, (size_t)nmat * nb * nb));
    ED_TRY(dalloc(&d_vbc, (size_t)nmat * nb * nc));
    for (int part = 0; part < nmat; ++part) {
        const std::vector<double> &ec = part ? S->ec_im : S->ec, &eb = part ? S->eb_im : S->eb,
                                  &vbc = part ? S->vbc_im : S->vbc;
        ED_HIP_CHECK(hipMemcpyAsync(d_ec + (size_t)part * nc * nc, ec.data() + (size_t)spin * nc * nc,
                                    sizeof(double) * nc * nc, hipMemcpyHostToDevice, st));
        if (nb) {
            ED_HIP_CHECK(hipMemcpyAsync(d_eb + (size_t)part * nb * nb, eb.data() + (size_t)spin * nb * nb,
                                        sizeof(double) * nb * nb, hipMemcpyHostToDevice, st));
            ED_HIP_CHECK(hipMemcpyAsync(d_vbc + (size_t)part * nb * nc, vbc.data() + (size_t)spin * nb * nc,
                                        sizeof(double) * nb * nc, hipMemcpyHostToDevice, st));
        }
    }
    SpinModel m = spin_model(nc, nb, d_ec, d_eb, d_vbc, S->cutoff);
    if (S->cplx) {
        m.ec_im = d_ec + (size_t)nc * nc;
        m.eb_im = d_eb + (size_t)nb * nb;
        m.vbc_im = d_vbc + (size_t)nb * nc;
    }
    hipLaunchKernelGGL(k_tab<false>, dim3(nblocks(d)), dim3(kBlock), 0, st, m, T.states, d, S->d_rank,
                       (const int32_t *)nullptr, T.noff, (double *)nullptr, (int32_t *)nullptr, (double *)nullptr,
                       (double *)nullptr);
    int32_t *d_kmax = nullptr;
    ED_TRY(dalloc(&d_kmax, 1));
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(kBlock), 0, st, T.noff, d, T.ptr, d_kmax);
    ED_HIP_CHECK(hipGetLastError());
    int32_t h[2];
    ED_HIP_CHECK(hipMemcpyAsync(&h[0], T.ptr + d, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipMemcpyAsync(&h[1], d_kmax, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipStreamSynchronize(st));
    T.nnz = h[0];
    T.kmax = h[1];
    ED_TRY(dalloc(&T.rk, (size_t)T.nnz));
    ED_TRY(dalloc(&T.off, (size_t)T.nnz));
    if (S->cplx) ED_TRY(dalloc(&T.off_im, (size_t)T.nnz));
    hipLaunchKernelGGL(k_tab<true>, dim3(nblocks(d)), dim3(kBlock), 0, st, m, T.states, d, S->d_rank,
                       (const int32_t *)T.ptr, T.noff, T.diag, T.rk, T.off, T.off_im);
    ED_HIP_CHECK(hipGetLastError());
    // ELL width: the next compiled width of the unrolled H.v kernel, or kmax itself
    // above the largest one (the generic kernel then loops over it)
    const int widths[] = {4, 8, 12, 16, 24, 32};
    T.kpad = std::max(T.kmax, 1);
    for (int w : widths)
        if (T.kmax <= w) { T.kpad = w; break; }
    const int kw = T.kpad;
    ED_TRY(dalloc(&T.ell_rk, (size_t)kw * d));
    ED_TRY(dalloc(&T.ell_val, (size_t)kw * d));
    if (S->cplx) ED_TRY(dalloc(&T.ell_vim, (size_t)kw * d));
    hipLaunchKernelGGL(k_ell, dim3(nblocks(d)), dim3(kBlock), 0, st, d, kw, T.ptr, T.rk, T.off, T.ell_rk,
                       T.ell_val, (const double *)T.off_im, T.ell_vim);
    ED_TRY(dalloc(&T.imp, d));
    ED_TRY(dalloc(&T.swap, (size_t)d * S->ncfg));
    hipLaunchKernelGGL(k_imp_swap, dim3(nblocks((int64_t)d * S->ncfg)), dim3(kBlock), 0, st, T.states, d, nb,
                       S->ncfg, S->d_rank, T.imp, T.swap);
    ED_TRY(dalloc(&T.diag32, d));
    ED_TRY(dalloc(&T.off32, (size_t)T.nnz));
    ED_TRY(dalloc(&T.ell_val32, (size_t)kw * d));
    hipLaunchKernelGGL(k_to_f32, dim3(nblocks(d)), dim3(kBlock), 0, st, T.diag, T.diag32, (int64_t)d);
    if (T.nnz)
        hipLaunchKernelGGL(k_to_f32, dim3(nblocks(T.nnz)), dim3(kBlock), 0, st, T.off, T.off32, (int64_t)T.nnz);
    hipLaunchKernelGGL(k_to_f32, dim3(nblocks((int64_t)kw * d)), dim3(kBlock), 0, st, T.ell_val, T.ell_val32,
                       (int64_t)kw * d);
    ED_HIP_CHECK(hipGetLastError());
    ED_HIP_CHECK(hipStreamSynchronize(st));
    (void)hipFree(d_ec);
    (void)hipFree(d_eb);
    (void)hipFree(d_vbc);
    (void)hipFree(d_kmax);
    return ED_OK;
}

// The Slater 4-index term of impurity configuration (cu, cd) (flag_slater_int,
// HAIMupdo.f90:905-999): destroy site4 in spin is2, destroy site3 in spin is1, create
// site2 in spin is1, create site1 in spin is2; the result adds U(s1, s2, s3, s4) * sign
// times v(result) to the row.  Signs: orbitals occupied below site4 and below site1 in
// the is2 string, strictly between site2 and site3 in the is1 string.  The bath orbitals
// lie below every impurity orbital and are counted twice in the is2 string, so the sign
// is a function of the impurity bits alone.  Returns (target cu | cd << 8, value), one
// entry per target.
static std::vector<std::pair<int32_t, double>> slater_row(const double *u4, int nc, uint32_t cu, uint32_t cd)
{
    auto below = [](uint32_t x, int c) { return (__builtin_popcount(x & ((1u << c) - 1u)) & 1) ? -1 : 1; };
    auto between = [](uint32_t x, int a, int b) {
        const int lo = std::min(a, b), hi = std::max(a, b);
        const uint32_t mask = (hi - lo > 1) ? (((1u << (hi - lo - 1)) - 1u) << (lo + 1)) : 0u;
        return (__builtin_popcount(x & mask) & 1) ? -1 : 1;
    };
    std::vector<std::pair<int32_t, double>> out;
    const uint32_t in[2] = {cu, cd};
    for (int s4 = 0; s4 < nc; ++s4)
        for (int is2 = 0; is2 < 2; ++is2) {
            if (!((in[is2] >> s4) & 1u)) continue;
            uint32_t a[2] = {in[0], in[1]};
            const int g1 = below(a[is2], s4);
            a[is2] &= ~(1u << s4);
            for (int s3 = 0; s3 < nc; ++s3)
                for (int is1 = 0; is1 < 2; ++is1) {
                    if (!((a[is1] >> s3) & 1u)) continue;
                    uint32_t b[2] = {a[0], a[1]};
                    b[is1] &= ~(1u << s3);
                    for (int s2 = 0; s2 < nc; ++s2) {
                        if ((b[is1] >> s2) & 1u) continue;
                        const int g2 = g1 * between(b[is1], s2, s3);
                        uint32_t c[2] = {b[0], b[1]};
                        c[is1] |= 1u << s2;
                        for (int s1 = 0; s1 < nc; ++s1) {
                            if ((c[is2] >> s1) & 1u) continue;
                            const double u = u4[((s1 * nc + s2) * nc + s3) * nc + s4];
                            if (u == 0.0) continue;
                            const int g = g2 * below(c[is2], s1);
                            uint32_t d[2] = {c[0], c[1]};
                            d[is2] |= 1u << s1;
                            const int32_t key = (int32_t)(d[0] | (d[1] << 8));
                            size_t q = 0;
                            while (q < out.size() && out[q].first != key) ++q;
                            if (q == out.size()) out.push_back({key, 0.0});
                            out[q].second += u * g;
                        }
                    }
                }
        }
    return out;
}

// Interaction diagonal and Hund transitions per impurity configuration pair.  Both act
// on impurity bits only, and the Hund signs are parities of impurity bits strictly
// between two impurity orbitals, so they tabulate over (cu, cd) in [0, 2^nc)^2
// (HAIMupdo.f90:785-829 density terms, :1017-1129 spin flip and pair hopping).  A Slater
// tensor adds its diagonal to the first and its other transitions to the second list.
static int build_interaction(ed_sector *S)
{
    const int nc = S->nc, ncfg = S->ncfg;
    const double *U = S->u.data(), *JJ = S->jj.data();
    bool offc = false;
    for (int a = 0; a < nc; ++a)
        for (int b = 0; b < nc; ++b)
            if (a != b && U[a * nc + b] != 0.0) offc = true;
    const bool hund_j = std::fabs(S->jhund) > 1e-5 && offc;
    std::vector<double> uimp((size_t)ncfg * ncfg, 0.0);
    std::vector<int32_t> hptr((size_t)ncfg * ncfg + 1, 0), hcu, hcd, htr;
    std::vector<double> hval;
    auto par = [](uint32_t x, int a, int b) {   // parity of bits strictly between a and b
        int lo = std::min(a, b), hi = std::max(a, b);
        uint32_t mask = (hi - lo > 1) ? (((1u << (hi - lo - 1)) - 1u) << (lo + 1)) : 0u;
        return (__builtin_popcount(x & mask) & 1) ? -1 : 1;
    };
    for (int cu = 0; cu < ncfg; ++cu)
        for (int cd = 0; cd < ncfg; ++cd) {
            double e = 0.0;
            for (int c = 0; c < nc; ++c)
                if (U[c * nc + c] != 0.0 && ((cu >> c) & 1) && ((cd >> c) & 1)) e += U[c * nc + c];
            if (offc)
                for (int c1 = 0; c1 < nc - 1; ++c1) {
                    int u1 = (cu >> c1) & 1, d1 = (cd >> c1) & 1;
                    if (!(u1 + d1)) continue;
                    for (int c2 = c1 + 1; c2 < nc; ++c2) {
                        if (U[c1 * nc + c2] == 0.0) continue;
                        int u2 = (cu >> c2) & 1, d2 = (cd >> c2) & 1;
                        if (!(u2 + d2)) continue;
                        e += U[c1 * nc + c2] * (u1 + d1) * (u2 + d2);
                        if (u1 && u2) e -= JJ[c1 * nc + c2];
                        if (d1 && d2) e -= JJ[c1 * nc + c2];
                    }
                }
            std::vector<std::pair<int32_t, double>> sl;
            if (!S->slater.empty()) {
                sl = slater_row(S->slater.data(), nc, (uint32_t)cu, (uint32_t)cd);
                for (const auto &t : sl)
                    if (t.first == (int32_t)(cu | (cd << 8))) e += t.second;
            }
            uimp[(size_t)cu * ncfg + cd] = e;
            if (hund_j) {
                for (int c1 = 0; c1 < nc; ++c1) {
                    int u1 = (cu >> c1) & 1, d1 = (cd >> c1) & 1;
                    if (u1 + d1 != 1) continue;
                    for (int c2 = c1 + 1; c2 < nc; ++c2) {
                        if (U[c1 * nc + c2] == 0.0) continue;
                        int u2 = (cu >> c2) & 1, d2 = (cd >> c2) & 1;
                        if (u2 + d2 != 1) continue;
                        if (u1 && d2) {   // up c1 -> c2, down c2 -> c1
                            uint32_t ou = (cu & ~(1u << c1)) | (1u << c2), od = (cd & ~(1u << c2)) | (1u << c1);
                            hcu.push_back(ou); hcd.push_back(od);
                            hval.push_back(JJ[c1 * nc + c2] * par(ou, c1, c2) * par(od, c1, c2));
                        }
                        if (d1 && u2) {   // up c2 -> c1, down c1 -> c2
                            uint32_t ou = (cu & ~(1u << c2)) | (1u << c1), od = (cd & ~(1u << c1)) | (1u << c2);
                            hcu.push_back(ou); hcd.push_back(od);
                            hval.push_back(JJ[c1 * nc + c2] * par(ou, c1, c2) * par(od, c1, c2));
                        }
                    }
                }
                for (int c1 = 0; c1 < nc; ++c1) {
                    if (!(((cu >> c1) & 1) && ((cd >> c1) & 1))) continue;
                    for (int c2 = 0; c2 < nc; ++c2) {
                        if (c2 == c1 || U[c1 * nc + c2] == 0.0) continue;
                        if (((cu >> c2) & 1) || ((cd >> c2) & 1)) continue;
                        uint32_t ou = (cu & ~(1u << c1)) | (1u << c2), od = (cd & ~(1u << c1)) | (1u << c2);
                        hcu.push_back(ou); hcd.push_back(od);
                        hval.push_back(JJ[c1 * nc + c2] * par(ou, c1, c2) * par(od, c1, c2));
                    }
                }
            }
            for (const auto &t : sl)
                if (t.first != (int32_t)(cu | (cd << 8))) {
                    hcu.push_back(t.first & 0xFF); hcd.push_back(t.first >> 8);
                    hval.push_back(t.second);
                }
            hptr[(size_t)cu * ncfg + cd + 1] = (int32_t)hcu.size();
        }
    S->nhund = (int)hcu.size();
    S->hund_maxrow = 0;
    for (size_t q = 0; q + 1 < hptr.size(); ++q) S->hund_maxrow = std::max(S->hund_maxrow, hptr[q + 1] - hptr[q]);
    S->hund = (hund_j || S->nhund > 0) ? 1 : 0;
    for (size_t t = 0; t < hcu.size(); ++t) htr.push_back(hcu[t] | (hcd[t] << 8));
    std::vector<float> uimp32(uimp.begin(), uimp.end()), hval32(hval.begin(), hval.end());
    ED_TRY(dalloc(&S->d_uimp, uimp.size()));
    ED_TRY(dalloc(&S->d_uimp32, uimp.size()));
    ED_TRY(dalloc(&S->d_hund_ptr, hptr.size()));
    ED_TRY(dalloc(&S->d_hund_cu, hcu.size()));
    ED_TRY(dalloc(&S->d_hund_cd, hcd.size()));
    ED_TRY(dalloc(&S->d_hund_val, hval.size()));
    ED_TRY(dalloc(&S->d_hund_tr, htr.size()));
    ED_TRY(dalloc(&S->d_hund_val32, hval.size()));
    ED_HIP_CHECK(hipMemcpy(S->d_uimp, uimp.data(), uimp.size() * sizeof(double), hipMemcpyHostToDevice));
    ED_HIP_CHECK(hipMemcpy(S->d_uimp32, uimp32.data(), uimp32.size() * sizeof(float), hipMemcpyHostToDevice));
    ED_HIP_CHECK(hipMemcpy(S->d_hund_ptr, hptr.data(), hptr.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!hcu.empty()) {
        ED_HIP_CHECK(hipMemcpy(S->d_hund_cu, hcu.data(), hcu.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        ED_HIP_CHECK(hipMemcpy(S->d_hund_cd, hcd.data(), hcd.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        ED_HIP_CHECK(hipMemcpy(S->d_hund_val, hval.data(), hval.size() * sizeof(double), hipMemcpyHostToDevice));
        ED_HIP_CHECK(hipMemcpy(S->d_hund_tr, htr.data(), htr.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        ED_HIP_CHECK(hipMemcpy(S->d_hund_val32, hval32.data(), hval32.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    return ED_OK;
}

// ------------------------------------------------------------ spin-table cache
// The tables of one spin depend only on that spin's model matrices, the cutoff, the orbital
// counts, the electron count and the device, not on the other spin: the 157 Lanczos sectors
// of an Ns = 12 scan have 13 distinct tables per spin.  Sectors of the same model share
// them: a table is built by the first sector that needs it, kept while any sector holds it
// (reference count) and freed with the last one.  The tables are never written after they
// are built, and the builder synchronises its stream before it publishes them.  The packed
// up-hop table has kPkRows rows (at least the spin's own width), enough for every
// compiled width of the fast kernels, so it does not depend on the other spin either.
constexpr int kPkRows = 16;
struct SpinCacheEntry {
    SpinTables T;
    int refs = 0;
};
static std::mutex g_spin_mu;
static std::map<std::string, SpinCacheEntry> g_spin_cache;

static std::string spin_key(const ed_sector *S, int spin, int n)
{
    std::string k;
    auto put = [&](const void *p, size_t bytes) { k.append(reinterpret_cast<const char *>(p), bytes); };
    const int hdr[7] = {S->device, S->ns, S->nc, S->nb, n, S->cplx, S->ncfg};
    put(hdr, sizeof hdr);
    put(&S->cutoff, sizeof(double));
    const size_t nc2 = (size_t)S->nc * S->nc, nb2 = (size_t)S->nb * S->nb, nbc = (size_t)S->nb * S->nc;
    put(S->ec.data() + spin * nc2, sizeof(double) * nc2);
    put(S->eb.data() + spin * nb2, sizeof(double) * nb2);
    put(S->vbc.data() + spin * nbc, sizeof(double) * nbc);
    if (S->cplx) {
        put(S->ec_im.data() + spin * nc2, sizeof(double) * nc2);
        put(S->eb_im.data() + spin * nb2, sizeof(double) * nb2);
        put(S->vbc_im.data() + spin * nbc, sizeof(double) * nbc);
    }
    return k;
}

// The tables of `spin` with n electrons for sector S: from the cache, or built (with the
// packed table of a real sector) and entered in it.  `key` is set once S holds a reference.
static int spin_acquire(ed_sector *S, int spin, int n, SpinTables &T, std::string &key)
{
    const std::string k = spin_key(S, spin, n);
    {
        std::lock_guard<std::mutex> lk(g_spin_mu);
        auto it = g_spin_cache.find(k);
        if (it != g_spin_cache.end()) {
            T = it->second.T;
            ++it->second.refs;
            key = k;
            return ED_OK;
        }
    }
    ED_TRY(build_spin(S, spin, n, T));
    if (!S->cplx) ED_TRY(build_packed(T, std::max(kPkRows, T.kpad), S->stream));
    std::lock_guard<std::mutex> lk(g_spin_mu);
    auto ins = g_spin_cache.emplace(k, SpinCacheEntry{T, 1});
    if (!ins.second) {   // another thread built the same tables meanwhile: use those
        free_spin(T);
        T = ins.first->second.T;
        ++ins.first->second.refs;
    }
    key = k;
    return ED_OK;
}

static void spin_release(SpinTables &T, std::string &key)
{
    if (key.empty()) {   // owned by the sector (S_z tables, or a build that failed half-way)
        free_spin(T);
        return;
    }
    std::lock_guard<std::mutex> lk(g_spin_mu);
    auto it = g_spin_cache.find(key);
    if (it != g_spin_cache.end() && --it->second.refs == 0) {
        free_spin(it->second.T);
        g_spin_cache.erase(it);
    }
    T = SpinTables();
    key.clear();
}

// Streams of released sectors, kept for the next sectors created on the same device:
// creating and destroying a HIP stream costs about a millisecond each (844 and 615 us on
// average over the 157 Lanczos sectors of an Ns = 12 scan), more than building a small
// sector's tables.  A stream goes back to the pool only once it is idle; the pool holds at
// most kStreamPool streams (default priority, non-blocking, as created).
constexpr size_t kStreamPool = 256;
static std::mutex g_stream_mu;
static std::vector<std::pair<int, hipStream_t>> g_stream_pool;

static hipError_t stream_acquire(hipStream_t *s)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    {
        std::lock_guard<std::mutex> lk(g_stream_mu);
        for (size_t i = g_stream_pool.size(); i-- > 0;)
            if (g_stream_pool[i].first == dev) {
                *s = g_stream_pool[i].second;
                g_stream_pool.erase(g_stream_pool.begin() + (std::ptrdiff_t)i);
                return hipSuccess;
            }
    }
    return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}

static void stream_release(int device, hipStream_t s)
{
    if (!s) return;
    if (hipStreamSynchronize(s) != hipSuccess) {   // not idle or broken: not reused
        (void)hipGetLastError();
        (void)hipStreamDestroy(s);
        return;
    }
    std::lock_guard<std::mutex> lk(g_stream_mu);
    if (g_stream_pool.size() < kStreamPool) g_stream_pool.emplace_back(device, s);
    else (void)hipStreamDestroy(s);
}

SectorStream::operator hipStream_t() const
{
    if (!s) {
        int cur = -1;
        const bool other = hipGetDevice(&cur) == hipSuccess && cur != device;
        if (other) (void)hipSetDevice(device);
        if (stream_acquire(&s) != hipSuccess) {
            (void)hipGetLastError();
            s = nullptr;
        }
        if (other) (void)hipSetDevice(cur);
    }
    return s;
}

static void free_sector(ed_sector *S)
{
    spin_release(S->up, S->up_key);
    spin_release(S->dn, S->dn_key);
    void *ps[] = {S->d_rank, S->d_uimp, S->d_uimp32, S->d_hund_ptr, S->d_hund_cu, S->d_hund_cd, S->d_hund_val, S->d_hund_tr,
                  S->d_hund_val32, S->d_basis, S->d_state, S->d_partials, S->d_io_a, S->d_io_b, S->d_colord, S->d_pmap, S->d_qx,
                  S->d_quart, S->d_quart32, S->d_sz_hrk, S->d_sz_hval, S->d_sz_hval32, S->d_sz_mat};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    for (ed::GraphCache &G : S->graphs) {
        if (G.exec) (void)hipGraphExecDestroy(G.exec);
        if (G.graph) (void)hipGraphDestroy(G.graph);
    }
    if (S->h_mon) (void)hipHostFree(S->h_mon);
    if (S->ev0) (void)hipEventDestroy(S->ev0);
    if (S->ev1) (void)hipEventDestroy(S->ev1);
    stream_release(S->device, S->stream.s);
    delete S;
}

// (N_up, N_dn) sector; with ec_im (and eb_im, vbc_im when nb > 0) a complex model, whose
// tables carry complex entries and whose vectors are planar complex (dim = 2 D).
static int create_sector(const ed_model *m, int nup, int ndn, ed_sector **out, const double *ec_im = nullptr,
                         const double *eb_im = nullptr, const double *vbc_im = nullptr, const double *u4 = nullptr)
{
    if (!m || !out) return fail(ED_ERR_ARG, "ed_sector_create: NULL argument");
    *out = nullptr;
    const int nc = m->nc, nb = m->nb, ns = nc + nb;
    if (nc < 1 || nc > 6 || nb < 0 || ns > 15)
        return fail(ED_ERR_ARG, "ed_sector_create: need 1 <= nc <= 6, nb >= 0, nc + nb <= 15");
    if (nup < 0 || nup > ns || ndn < 0 || ndn > ns)
        return fail(ED_ERR_ARG, "ed_sector_create: particle numbers outside [0, Ns]");
    if (!m->ec || !m->u || !m->jj || (nb && (!m->eb || !m->vbc)))
        return fail(ED_ERR_ARG, "ed_sector_create: NULL model array");
    ed_sector *S = new ed_sector();
    ED_HIP_CHECK(hipGetDevice(&S->device));
    S->nc = nc; S->nb = nb; S->ns = ns; S->nup = nup; S->ndn = ndn; S->ncfg = 1 << nc;
    S->ec.assign(m->ec, m->ec + 2 * nc * nc);
    if (nb) {
        S->eb.assign(m->eb, m->eb + 2 * nb * nb);
        S->vbc.assign(m->vbc, m->vbc + 2 * nb * nc);
    }
    S->u.assign(m->u, m->u + nc * nc);
    S->jj.assign(m->jj, m->jj + nc * nc);
    S->jhund = m->jhund;
    S->cutoff = m->cutoff;
    if (u4) S->slater.assign(u4, u4 + (size_t)nc * nc * nc * nc);
    if (ec_im) {
        S->cplx = 1;
        S->ec_im.assign(ec_im, ec_im + 2 * nc * nc);
        S->eb_im.assign((size_t)2 * nb * nb, 0.0);
        S->vbc_im.assign((size_t)2 * nb * nc, 0.0);
        if (nb && eb_im) S->eb_im.assign(eb_im, eb_im + 2 * nb * nb);
        if (nb && vbc_im) S->vbc_im.assign(vbc_im, vbc_im + 2 * nb * nc);
    }
    int rc = ED_OK;
    auto bail = [&](int code) { free_sector(S); return code; };
    S->stream.device = S->device;
    if (stream_acquire(&S->stream.s) != hipSuccess) return bail(fail(ED_ERR_HIP, "hipStreamCreate failed"));
    if ((rc = dalloc(&S->d_rank, (size_t)1 << ns)) != ED_OK) return bail(rc);
    hipLaunchKernelGGL(k_rank_table, dim3(nblocks((int64_t)1 << ns)), dim3(kBlock), 0, S->stream, ns, S->d_rank);
    // the spin tables, shared with the other sectors of the model (with the packed
    // coefficients of the fast real kernels)
    if ((rc = spin_acquire(S, 0, nup, S->up, S->up_key)) != ED_OK) return bail(rc);
    if ((rc = spin_acquire(S, 1, ndn, S->dn, S->dn_key)) != ED_OK) return bail(rc);
    if ((rc = build_interaction(S)) != ED_OK) return bail(rc);
    S->dim = (int64_t)S->up.d * S->dn.d;
    if (S->cplx) {
        S->cdim = S->dim;
        S->dim = 2 * S->cdim;
    }
    if (hipEventCreate(&S->ev0) != hipSuccess || hipEventCreate(&S->ev1) != hipSuccess)
        return bail(fail(ED_ERR_HIP, "hipEventCreate failed"));
    if (hipStreamSynchronize(S->stream) != hipSuccess) return bail(fail(ED_ERR_HIP, "sector build failed"));
    stream_release(S->device, S->stream.s);   // the sector takes a stream again when it first needs one
    S->stream.s = nullptr;
    *out = S;
    return ED_OK;
}

// QUART_INT_SZ (HAIMsz.f90:129-177): the impurity interaction of every Nambu state,
// with the actual occupations n_up = bit of IMPiorb(c), n_dn = NOT bit of IMPiorb(c + Nc):
//   sum_c U_cc [n_c = 2]  +  sum_{c < c', U_cc' != 0} U_cc' n_c n_c'.
// (The reference gates the second sum with a flag computed from an uninitialised
// diagonal mask, HAIMsz.f90:75-79; every term it can add has U_cc' != 0, so the sum over
// the nonzero off-diagonal U is the same result.)  J = 0 here: the real-valued S_z
// H.v stops for |Jhund| > 1e-5 (HAIMsz.f90:473-476).
// jj (non-null when |Jhund| > 1e-5, Jhund_Slater_type): - J_c1c2 for every pair c1 < c2
// whose physical up orbitals are both occupied, and again for the down orbitals
// (HAIMsz.f90:161-170).
__global__ void k_quart_sz(const int32_t *states, int d, int nc, int nb, int ns, const double *u, const double *jj,
                           double *quart)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d) return;
    const uint32_t s = (uint32_t)states[i];
    double q = 0.0;
    for (int c1 = 0; c1 < nc; ++c1) {
        const int u1 = (int)((s >> (nb + c1)) & 1u), p1 = (int)(((s >> (ns + nb + c1)) & 1u) ^ 1u);
        const int n1 = u1 + p1;
        if (n1 == 2 && u[c1 * nc + c1] != 0.0) q += u[c1 * nc + c1];
        if (n1 != 0)
            for (int c2 = c1 + 1; c2 < nc; ++c2) {
                if (u[c1 * nc + c2] == 0.0) continue;
                const int n2 = (int)((s >> (nb + c2)) & 1u) + (int)(((s >> (ns + nb + c2)) & 1u) ^ 1u);
                q += u[c1 * nc + c2] * n1 * n2;
            }
        if (jj)
            for (int c2 = c1 + 1; c2 < nc; ++c2) {
                const int u2 = (int)((s >> (nb + c2)) & 1u), p2 = (int)(((s >> (ns + nb + c2)) & 1u) ^ 1u);
                if (u1 && u2) q -= jj[c1 * nc + c2];
                if (p1 && p2) q -= jj[c1 * nc + c2];
            }
    }
    quart[i] = q;
}

// The Hund terms of HAIMsz_multc (HAIMsz.f90:300-406) for every S_z state, in the
// reference's order, as an ELL [kh][d] (padded with the row itself and 0).  Raw Nambu
// bits: u_c = bit nb + c (up), p_c = bit ns + nb + c (particle-hole down orbital).
//   spin flip, c1 < c2: (u, p) = (1, 1) at c1 and (0, 0) at c2 swap places, then the
//     reverse; pair hopping, c1 != c2: (1, 0) at c1 and (0, 1) at c2 become (0, 1), (1, 0);
// element -J_c1c2 times the parities of the target's bits strictly between u_c1 and u_c2
// and between p_c1 and p_c2.  FILL = false counts the entries per state.
template <bool FILL>
__global__ void k_hund_sz(const int32_t *states, int d, int nc, int nb, int ns, const double *jj, const int32_t *rank,
                          int kh, int32_t *cnt, int32_t *hrk, double *hval)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d) return;
    const uint32_t s = (uint32_t)states[i];
    auto between = [](uint32_t o, int a, int b) {
        const int lo = min(a, b), hi = max(a, b);
        const uint32_t mask = hi - lo > 1 ? ((1u << (hi - lo - 1)) - 1u) << (lo + 1) : 0u;
        return (__popc(o & mask) & 1) ? -1 : 1;
    };
    int k = 0;
    auto put = [&](uint32_t o, int c1, int c2) {
        if (FILL) {
            const int sg = between(o, nb + c1, nb + c2) * between(o, ns + nb + c1, ns + nb + c2);
            hrk[(int64_t)k * d + i] = rank[o];
            hval[(int64_t)k * d + i] = -jj[c1 * nc + c2] * sg;
        }
        ++k;
    };
    auto ub = [&](int c) { return 1u << (nb + c); };
    auto pb = [&](int c) { return 1u << (ns + nb + c); };
    for (int c1 = 0; c1 < nc; ++c1) {
        const bool u1 = s & ub(c1), p1 = s & pb(c1);
        for (int c2 = c1 + 1; c2 < nc; ++c2) {
            const bool u2 = s & ub(c2), p2 = s & pb(c2);
            if (u1 && p1 && !u2 && !p2) put((s & ~ub(c1) & ~pb(c1)) | pb(c2) | ub(c2), c1, c2);
            if (!u1 && !p1 && u2 && p2) put((s & ~ub(c2) & ~pb(c2)) | pb(c1) | ub(c1), c1, c2);
        }
    }
    for (int c1 = 0; c1 < nc; ++c1) {
        if (!((s & ub(c1)) && !(s & pb(c1)))) continue;
        for (int c2 = 0; c2 < nc; ++c2)
            if (c2 != c1 && (s & pb(c2)) && !(s & ub(c2))) put((s & ~ub(c1) & ~pb(c2)) | pb(c1) | ub(c2), c1, c2);
    }
    if (FILL) {
        for (; k < kh; ++k) {
            hrk[(int64_t)k * d + i] = i;
            hval[(int64_t)k * d + i] = 0.0;
        }
    } else {
        cnt[i] = k;
    }
}

// Nambu matrices (Nambu_Ec, impurity_class.f90:1018-1037; Nambu_Eb / Nambu_Vbc,
// bath_class.f90:956-1035), row-major:
//   Ec_N  = [[Ec_up, 0], [0, -Ec_dn^T]]                       (2 nc x 2 nc)
//   Eb_N  = [[Eb_up, Pb], [Pb^T, -Eb_dn^T]]                   (2 nb x 2 nb)
//   Vbc_N = [[Vbc_up, PVbc_dn], [PVbc_up, -Vbc_dn]]           (2 nb x 2 nc; the PVbc
//           blocks only with pairing_imp_to_bath, PAIRING_IMP_TO_BATH)
static void nambu_matrices(int nc, int nb, const double *ec, const double *eb, const double *vbc, const double *pb,
                           const double *pvbc, int pair_ib, std::vector<double> &ecn, std::vector<double> &ebn,
                           std::vector<double> &vbn)
{
    const int n2 = 2 * nc, b2 = 2 * nb;
    ecn.assign((size_t)n2 * n2, 0.0);
    ebn.assign((size_t)b2 * b2, 0.0);
    vbn.assign((size_t)b2 * n2, 0.0);
    for (int a = 0; a < nc; ++a)
        for (int b = 0; b < nc; ++b) {
            ecn[(size_t)a * n2 + b] = ec[a * nc + b];
            ecn[(size_t)(nc + a) * n2 + nc + b] = -ec[nc * nc + b * nc + a];
        }
    for (int k = 0; k < nb; ++k)
        for (int l = 0; l < nb; ++l) {
            ebn[(size_t)k * b2 + l] = eb[k * nb + l];
            ebn[(size_t)(nb + k) * b2 + nb + l] = -eb[nb * nb + l * nb + k];
            ebn[(size_t)k * b2 + nb + l] = pb ? pb[k * nb + l] : 0.0;
            ebn[(size_t)(nb + l) * b2 + k] = pb ? pb[k * nb + l] : 0.0;
        }
    for (int k = 0; k < nb; ++k)
        for (int c = 0; c < nc; ++c) {
            vbn[(size_t)k * n2 + c] = vbc[k * nc + c];
            vbn[(size_t)(nb + k) * n2 + nc + c] = -vbc[nb * nc + k * nc + c];
            if (pair_ib && pvbc) {
                vbn[(size_t)(nb + k) * n2 + c] = pvbc[k * nc + c];
                vbn[(size_t)k * n2 + nc + c] = pvbc[nb * nc + k * nc + c];
            }
        }
}

// Imaginary parts of the Nambu matrices of a complex model (the _complex branches of
// Nambu_Ec / Nambu_Eb / Nambu_Vbc): Im Ec_N = [[Im Ec_up, 0], [0, -Im Ec_dn^T]],
// Im Eb_N = [[Im Eb_up, Im Pb], [-Im Pb^T, -Im Eb_dn^T]] (conj(Pb)^T below the diagonal),
// Im Vbc_N = [[Im Vbc_up, Im PVbc_dn], [Im PVbc_up, +Im Vbc_dn]] (-conj(Vbc_dn)).  Any
// input may be null (zero imaginary part).
static void nambu_matrices_im(int nc, int nb, const double *ec_im, const double *eb_im, const double *vbc_im,
                              const double *pb_im, const double *pvbc_im, int pair_ib, std::vector<double> &eci,
                              std::vector<double> &ebi, std::vector<double> &vbi)
{
    const int n2 = 2 * nc, b2 = 2 * nb;
    eci.assign((size_t)n2 * n2, 0.0);
    ebi.assign((size_t)b2 * b2, 0.0);
    vbi.assign((size_t)b2 * n2, 0.0);
    for (int a = 0; a < nc; ++a)
        for (int b = 0; b < nc; ++b)
            if (ec_im) {
                eci[(size_t)a * n2 + b] = ec_im[a * nc + b];
                eci[(size_t)(nc + a) * n2 + nc + b] = -ec_im[nc * nc + b * nc + a];
            }
    for (int k = 0; k < nb; ++k)
        for (int l = 0; l < nb; ++l) {
            if (eb_im) {
                ebi[(size_t)k * b2 + l] = eb_im[k * nb + l];
                ebi[(size_t)(nb + k) * b2 + nb + l] = -eb_im[nb * nb + l * nb + k];
            }
            if (pb_im) {
                ebi[(size_t)k * b2 + nb + l] = pb_im[k * nb + l];
                ebi[(size_t)(nb + l) * b2 + k] = -pb_im[k * nb + l];
            }
        }
    for (int k = 0; k < nb; ++k)
        for (int c = 0; c < nc; ++c) {
            if (vbc_im) {
                vbi[(size_t)k * n2 + c] = vbc_im[k * nc + c];
                vbi[(size_t)(nb + k) * n2 + nc + c] = vbc_im[nb * nc + k * nc + c];
            }
            if (pair_ib && pvbc_im) {
                vbi[(size_t)(nb + k) * n2 + c] = pvbc_im[k * nc + c];
                vbi[(size_t)k * n2 + nc + c] = pvbc_im[nb * nc + k * nc + c];
            }
        }
}

// Imaginary parts of a complex S_z model (each may be null).
struct SzImag {
    const double *ec = nullptr, *eb = nullptr, *vbc = nullptr, *pb = nullptr, *pvbc = nullptr;
};

// The S_z sector of Nambu particle number npart = Ns + 2 S_z (new_HAIMsz + tab_HAIMsz,
// HAIMsz.f90:51-177; the sector is the popcount-npart states of 2 Ns orbitals in
// increasing order, quantum_fermion_hilbert_class.f90:346-421).  With `im` the model is
// complex (the COMPLEX(8) tables of HAIMsz_multc): vectors are planar as in every
// complex sector.
static int create_sector_sz(const ed_model *m, const double *pb, const double *pvbc, int pair_ib, int npart,
                            ed_sector **out, const SzImag *im = nullptr)
{
    if (!m || !out) return fail(ED_ERR_ARG, "ed_sector_create_sz: NULL argument");
    *out = nullptr;
    const int nc = m->nc, nb = m->nb, ns = nc + nb;
    if (nc < 1 || nc > 6 || nb < 0 || ns > 13)
        return fail(ED_ERR_ARG, "ed_sector_create_sz: need 1 <= nc <= 6, nb >= 0, nc + nb <= 13 (2 Ns bits)");
    if (npart < 0 || npart > 2 * ns) return fail(ED_ERR_ARG, "ed_sector_create_sz: npart outside [0, 2 Ns]");
    if (!m->ec || !m->u || !m->jj || (nb && (!m->eb || !m->vbc)))
        return fail(ED_ERR_ARG, "ed_sector_create_sz: NULL model array");
    if (pair_ib && !pvbc) return fail(ED_ERR_ARG, "ed_sector_create_sz: pairing_imp_to_bath needs pvbc");
    ed_sector *S = new ed_sector();
    ED_HIP_CHECK(hipGetDevice(&S->device));
    S->nc = nc; S->nb = nb; S->ns = ns; S->nup = npart; S->ndn = -1; S->ncfg = 1 << nc;
    S->sz = 1;
    S->ec.assign(m->ec, m->ec + 2 * nc * nc);
    if (nb) {
        S->eb.assign(m->eb, m->eb + 2 * nb * nb);
        S->vbc.assign(m->vbc, m->vbc + 2 * nb * nc);
    }
    S->u.assign(m->u, m->u + nc * nc);
    S->jj.assign(m->jj, m->jj + nc * nc);
    S->jhund = m->jhund;
    S->cutoff = m->cutoff;
    // energy_global_shift (impurity_class.f90:1037) + energy_global_shift2 (bath_class.f90:976)
    {
        double s1 = 0.0, s2 = 0.0;
        for (int c = 0; c < nc; ++c) s1 += m->ec[nc * nc + c * nc + c];
        for (int k = 0; k < nb; ++k) s2 += m->eb[nb * nb + k * nb + k];
        S->energy_shift = s1 + s2;
    }
    int rc = ED_OK;
    auto bail = [&](int code) { free_sector(S); return code; };
    S->stream.device = S->device;
    if (stream_acquire(&S->stream.s) != hipSuccess) return bail(fail(ED_ERR_HIP, "hipStreamCreate failed"));
    hipStream_t st = S->stream;
    const int nbits = 2 * ns;
    if ((rc = dalloc(&S->d_rank, (size_t)1 << nbits)) != ED_OK) return bail(rc);
    hipLaunchKernelGGL(k_rank_table, dim3(nblocks((int64_t)1 << nbits)), dim3(kBlock), 0, st, nbits, S->d_rank);
    SpinTables &T = S->up;
    T.d = binom(nbits, npart);
    const int d = T.d;
    if ((rc = dalloc(&T.states, d)) != ED_OK || (rc = dalloc(&T.diag, d)) != ED_OK ||
        (rc = dalloc(&T.noff, d)) != ED_OK || (rc = dalloc(&T.ptr, d + 1)) != ED_OK)
        return bail(rc);
    hipLaunchKernelGGL(k_enum_states, dim3(nblocks(d)), dim3(kBlock), 0, st, nbits, npart, d, T.states);
    std::vector<double> ecn, ebn, vbn, eci, ebi, vbi;
    nambu_matrices(nc, nb, m->ec, m->eb, m->vbc, pb, pvbc, pair_ib, ecn, ebn, vbn);
    const bool cx = im != nullptr;
    if (cx) nambu_matrices_im(nc, nb, im->ec, im->eb, im->vbc, im->pb, im->pvbc, pair_ib, eci, ebi, vbi);
    double *d_mat = nullptr, *d_u = nullptr;   // d_u: U[nc][nc], then JJ[nc][nc]
    const size_t nre = ecn.size() + std::max<size_t>(1, ebn.size()) + std::max<size_t>(1, vbn.size());
    const size_t nmat = cx ? 2 * nre : nre;   // real parts, then (complex) the imaginary parts
    if ((rc = dalloc(&d_mat, nmat)) != ED_OK || (rc = dalloc(&d_u, (size_t)2 * nc * nc)) != ED_OK) {
        (void)hipFree(d_mat);
        return bail(rc);
    }
    struct Tmp {
        double *&a, *&b;
        ~Tmp()
        {
            (void)hipFree(a);
            (void)hipFree(b);
        }
    } tmp{d_mat, d_u};
    double *d_ec = d_mat, *d_eb = d_ec + ecn.size(), *d_vbc = d_eb + std::max<size_t>(1, ebn.size());
    ED_HIP_CHECK(hipMemcpyAsync(d_ec, ecn.data(), sizeof(double) * ecn.size(), hipMemcpyHostToDevice, st));
    if (nb) {
        ED_HIP_CHECK(hipMemcpyAsync(d_eb, ebn.data(), sizeof(double) * ebn.size(), hipMemcpyHostToDevice, st));
        ED_HIP_CHECK(hipMemcpyAsync(d_vbc, vbn.data(), sizeof(double) * vbn.size(), hipMemcpyHostToDevice, st));
    }
    double *d_eci = cx ? d_mat + nre : nullptr, *d_ebi = cx ? d_eci + ecn.size() : nullptr,
           *d_vbi = cx ? d_ebi + std::max<size_t>(1, ebn.size()) : nullptr;
    if (cx) {
        ED_HIP_CHECK(hipMemcpyAsync(d_eci, eci.data(), sizeof(double) * eci.size(), hipMemcpyHostToDevice, st));
        if (nb) {
            ED_HIP_CHECK(hipMemcpyAsync(d_ebi, ebi.data(), sizeof(double) * ebi.size(), hipMemcpyHostToDevice, st));
            ED_HIP_CHECK(hipMemcpyAsync(d_vbi, vbi.data(), sizeof(double) * vbi.size(), hipMemcpyHostToDevice, st));
        }
    }
    ED_HIP_CHECK(hipMemcpyAsync(d_u, m->u, sizeof(double) * nc * nc, hipMemcpyHostToDevice, st));
    ED_HIP_CHECK(hipMemcpyAsync(d_u + nc * nc, m->jj, sizeof(double) * nc * nc, hipMemcpyHostToDevice, st));
    SpinModel sm{};
    sm.nimp = 2 * nc;
    sm.nbath = 2 * nb;
    for (int c = 0; c < nc; ++c) {
        sm.ibit[c] = nb + c;
        sm.ibit[nc + c] = ns + nb + c;
    }
    for (int k = 0; k < nb; ++k) {
        sm.bbit[k] = k;
        sm.bbit[nb + k] = ns + k;
    }
    sm.ec = d_ec;
    sm.eb = d_eb;
    sm.vbc = d_vbc;
    sm.ec_im = d_eci;   // null for a real model
    sm.eb_im = d_ebi;
    sm.vbc_im = d_vbi;
    sm.cut = m->cutoff;
    // the sector keeps the Nambu matrices for the on-the-fly H.v (HAIMsz_mult_fly)
    S->sz_model = sm;
    S->d_sz_mat = d_mat;
    d_mat = nullptr;
    hipLaunchKernelGGL(k_tab<false>, dim3(nblocks(d)), dim3(kBlock), 0, st, sm, T.states, d, S->d_rank,
                       (const int32_t *)nullptr, T.noff, (double *)nullptr, (int32_t *)nullptr, (double *)nullptr,
                       (double *)nullptr);
    int32_t *d_kmax = nullptr;
    if ((rc = dalloc(&d_kmax, 1)) != ED_OK) return bail(rc);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(kBlock), 0, st, T.noff, d, T.ptr, d_kmax);
    ED_HIP_CHECK(hipGetLastError());
    int32_t h[2];
    ED_HIP_CHECK(hipMemcpyAsync(&h[0], T.ptr + d, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipMemcpyAsync(&h[1], d_kmax, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    ED_HIP_CHECK(hipStreamSynchronize(st));
    (void)hipFree(d_kmax);
    T.nnz = h[0];
    T.kmax = h[1];
    T.kpad = std::max(T.kmax, 1);
    if ((rc = dalloc(&T.rk, (size_t)T.nnz)) != ED_OK || (rc = dalloc(&T.off, (size_t)T.nnz)) != ED_OK) return bail(rc);
    if (cx && (rc = dalloc(&T.off_im, (size_t)std::max<int64_t>(1, T.nnz))) != ED_OK) return bail(rc);
    hipLaunchKernelGGL(k_tab<true>, dim3(nblocks(d)), dim3(kBlock), 0, st, sm, T.states, d, S->d_rank,
                       (const int32_t *)T.ptr, T.noff, T.diag, T.rk, T.off, T.off_im);
    const int kw = T.kpad;
    if ((rc = dalloc(&T.ell_rk, (size_t)kw * d)) != ED_OK || (rc = dalloc(&T.ell_val, (size_t)kw * d)) != ED_OK)
        return bail(rc);
    if (cx && (rc = dalloc(&T.ell_vim, (size_t)kw * d)) != ED_OK) return bail(rc);
    hipLaunchKernelGGL(k_ell, dim3(nblocks(d)), dim3(kBlock), 0, st, d, kw, T.ptr, T.rk, T.off, T.ell_rk, T.ell_val,
                       (const double *)T.off_im, T.ell_vim);
    if ((rc = dalloc(&S->d_quart, d)) != ED_OK || (rc = dalloc(&S->d_quart32, d)) != ED_OK ||
        (rc = dalloc(&T.diag32, d)) != ED_OK || (rc = dalloc(&T.off32, (size_t)T.nnz)) != ED_OK ||
        (rc = dalloc(&T.ell_val32, (size_t)kw * d)) != ED_OK)
        return bail(rc);
    // Hund terms of HAIMsz_multc when |Jhund| > 1e-5 and some U_cc' != 0 (offdiag_coulomb)
    bool offc = false;
    for (int a = 0; a < nc; ++a)
        for (int b = 0; b < nc; ++b)
            if (a != b && m->u[a * nc + b] != 0.0) offc = true;
    const bool jterm = std::fabs(m->jhund) > 1e-5;
    const double *d_jj = jterm ? (const double *)(d_u + nc * nc) : nullptr;
    hipLaunchKernelGGL(k_quart_sz, dim3(nblocks(d)), dim3(kBlock), 0, st, T.states, d, nc, nb, ns, (const double *)d_u,
                       d_jj, S->d_quart);
    hipLaunchKernelGGL(k_to_f32, dim3(nblocks(d)), dim3(kBlock), 0, st, S->d_quart, S->d_quart32, (int64_t)d);
    S->sz_kh = 0;
    if (jterm && offc) {
        int32_t *d_cnt = nullptr, *d_ptr = nullptr, *d_kh = nullptr;
        if ((rc = dalloc(&d_cnt, d)) != ED_OK || (rc = dalloc(&d_ptr, (size_t)d + 1)) != ED_OK ||
            (rc = dalloc(&d_kh, 1)) != ED_OK)
            return bail(rc);
        hipLaunchKernelGGL(k_hund_sz<false>, dim3(nblocks(d)), dim3(kBlock), 0, st, T.states, d, nc, nb, ns, d_jj,
                           (const int32_t *)S->d_rank, 0, d_cnt, (int32_t *)nullptr, (double *)nullptr);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(kBlock), 0, st, (const int32_t *)d_cnt, d, d_ptr, d_kh);
        int32_t kh = 0;
        ED_HIP_CHECK(hipMemcpyAsync(&kh, d_kh, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        ED_HIP_CHECK(hipStreamSynchronize(st));
        (void)hipFree(d_cnt);
        (void)hipFree(d_ptr);
        (void)hipFree(d_kh);
        if (kh > 0) {
            if ((rc = dalloc(&S->d_sz_hrk, (size_t)kh * d)) != ED_OK ||
                (rc = dalloc(&S->d_sz_hval, (size_t)kh * d)) != ED_OK ||
                (rc = dalloc(&S->d_sz_hval32, (size_t)kh * d)) != ED_OK)
                return bail(rc);
            hipLaunchKernelGGL(k_hund_sz<true>, dim3(nblocks(d)), dim3(kBlock), 0, st, T.states, d, nc, nb, ns, d_jj,
                               (const int32_t *)S->d_rank, (int)kh, (int32_t *)nullptr, S->d_sz_hrk, S->d_sz_hval);
            hipLaunchKernelGGL(k_to_f32, dim3(nblocks((int64_t)kh * d)), dim3(kBlock), 0, st, S->d_sz_hval,
                               S->d_sz_hval32, (int64_t)kh * d);
            S->sz_kh = kh;
        }
    }
    hipLaunchKernelGGL(k_to_f32, dim3(nblocks(d)), dim3(kBlock), 0, st, T.diag, T.diag32, (int64_t)d);
    if (T.nnz)
        hipLaunchKernelGGL(k_to_f32, dim3(nblocks(T.nnz)), dim3(kBlock), 0, st, T.off, T.off32, (int64_t)T.nnz);
    hipLaunchKernelGGL(k_to_f32, dim3(nblocks((int64_t)kw * d)), dim3(kBlock), 0, st, T.ell_val, T.ell_val32,
                       (int64_t)kw * d);
    ED_HIP_CHECK(hipGetLastError());
    S->dn.d = 1;
    S->dim = d;
    if (cx) {   // planar vectors (Re[d], Im[d]), as every complex sector
        S->cplx = 1;
        S->cdim = d;
        S->dim = 2 * (int64_t)d;
    }
    if (hipEventCreate(&S->ev0) != hipSuccess || hipEventCreate(&S->ev1) != hipSuccess)
        return bail(fail(ED_ERR_HIP, "hipEventCreate failed"));
    if (hipStreamSynchronize(st) != hipSuccess) return bail(fail(ED_ERR_HIP, "S_z sector build failed"));
    stream_release(S->device, S->stream.s);   // the sector takes a stream again when it first needs one
    S->stream.s = nullptr;
    *out = S;
    return ED_OK;
}

}  // namespace ed

extern "C" {

int ed_sector_create_sz(const ed_model *model, const double *pb, const double *pvbc, int pairing_imp_to_bath,
                        int npart, ed_sector **out)
{
    return ed::create_sector_sz(model, pb, pvbc, pairing_imp_to_bath, npart, out);
}

int ed_sector_create_sz_c(const ed_model *model, const double *ec_im, const double *eb_im, const double *vbc_im,
                          const double *pb, const double *pb_im, const double *pvbc, const double *pvbc_im,
                          int pairing_imp_to_bath, int npart, ed_sector **out)
{
    if (!ec_im) return ed::fail(ED_ERR_ARG, "ed_sector_create_sz_c: NULL ec_im");
    if (pb_im && !pb) return ed::fail(ED_ERR_ARG, "ed_sector_create_sz_c: pb_im without pb");
    if (pvbc_im && !pvbc) return ed::fail(ED_ERR_ARG, "ed_sector_create_sz_c: pvbc_im without pvbc");
    ed::SzImag im;
    im.ec = ec_im;
    im.eb = eb_im;
    im.vbc = vbc_im;
    im.pb = pb_im;
    im.pvbc = pvbc_im;
    return ed::create_sector_sz(model, pb, pvbc, pairing_imp_to_bath, npart, out, &im);
}

int ed_sector_quart_sz(const ed_sector *sec, double *quart)
{
    if (!sec || !quart) return ed::fail(ED_ERR_ARG, "ed_sector_quart_sz: NULL argument");
    if (!sec->sz) return ed::fail(ED_ERR_ARG, "ed_sector_quart_sz: not an S_z sector");
    ED_HIP_CHECK(hipMemcpy(quart, sec->d_quart, sizeof(double) * sec->up.d, hipMemcpyDeviceToHost));
    return ED_OK;
}

int ed_sector_hund_sz(const ed_sector *sec, int *kh, int32_t *rk, double *val)
{
    if (!sec || !kh) return ed::fail(ED_ERR_ARG, "ed_sector_hund_sz: NULL argument");
    if (!sec->sz) return ed::fail(ED_ERR_ARG, "ed_sector_hund_sz: not an S_z sector");
    *kh = sec->sz_kh;
    const size_t n = (size_t)sec->sz_kh * sec->up.d;
    if (rk && n) ED_HIP_CHECK(hipMemcpy(rk, sec->d_sz_hrk, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    if (val && n) ED_HIP_CHECK(hipMemcpy(val, sec->d_sz_hval, sizeof(double) * n, hipMemcpyDeviceToHost));
    return ED_OK;
}

int ed_sector_energy_shift(const ed_sector *sec, double *shift)
{
    if (!sec || !shift) return ed::fail(ED_ERR_ARG, "ed_sector_energy_shift: NULL argument");
    *shift = sec->sz ? sec->energy_shift : 0.0;
    return ED_OK;
}

int ed_sector_create_c(const ed_model *model, const double *ec_im, const double *eb_im, const double *vbc_im,
                       int nup, int ndn, ed_sector **out)
{
    if (!ec_im) return ed::fail(ED_ERR_ARG, "ed_sector_create_c: NULL ec_im");
    return ed::create_sector(model, nup, ndn, out, ec_im, eb_im, vbc_im);
}

int ed_sector_create(const ed_model *model, int nup, int ndn, ed_sector **out)
{
    return ed::create_sector(model, nup, ndn, out);
}

int ed_sector_create_slater(const ed_model *model, const double *u4, int nup, int ndn, ed_sector **out)
{
    if (!u4) return ed::fail(ED_ERR_ARG, "ed_sector_create_slater: NULL u4");
    return ed::create_sector(model, nup, ndn, out, nullptr, nullptr, nullptr, u4);
}

int ed_sector_destroy(ed_sector *sec)
{
    if (!sec) return ED_OK;
    if (sec->stream.s) (void)hipStreamSynchronize(sec->stream.s);   // a sector that never took a stream has none to wait on
    ed::free_sector(sec);
    return ED_OK;
}

int64_t ed_sector_dim(const ed_sector *sec) { return sec ? (sec->cplx ? sec->cdim : sec->dim) : 0; }

int ed_sector_spin_dims(const ed_sector *sec, int32_t *d_up, int32_t *d_dn)
{
    if (!sec) return ed::fail(ED_ERR_ARG, "ed_sector_spin_dims: NULL sector");
    if (d_up) *d_up = sec->up.d;
    if (d_dn) *d_dn = sec->dn.d;
    return ED_OK;
}

int ed_sector_nnz(const ed_sector *sec, int spin, int64_t *nnz)
{
    if (!sec || !nnz || (spin != 0 && spin != 1)) return ed::fail(ED_ERR_ARG, "ed_sector_nnz: bad argument");
    *nnz = spin ? sec->dn.nnz : sec->up.nnz;
    return ED_OK;
}

int ed_sector_tables(const ed_sector *sec, int spin, int32_t *states, double *diag, int32_t *noff,
                     int32_t *rankoff, double *offdiag)
{
    if (!sec || (spin != 0 && spin != 1)) return ed::fail(ED_ERR_ARG, "ed_sector_tables: bad argument");
    if (sec->sz && spin != 0) return ed::fail(ED_ERR_ARG, "ed_sector_tables: an S_z sector has one table (spin 0)");
    const ed::SpinTables &T = spin ? sec->dn : sec->up;
    if (states) ED_HIP_CHECK(hipMemcpy(states, T.states, sizeof(int32_t) * T.d, hipMemcpyDeviceToHost));
    if (diag) ED_HIP_CHECK(hipMemcpy(diag, T.diag, sizeof(double) * T.d, hipMemcpyDeviceToHost));
    if (noff) ED_HIP_CHECK(hipMemcpy(noff, T.noff, sizeof(int32_t) * T.d, hipMemcpyDeviceToHost));
    if (rankoff && T.nnz) ED_HIP_CHECK(hipMemcpy(rankoff, T.rk, sizeof(int32_t) * T.nnz, hipMemcpyDeviceToHost));
    if (offdiag && T.nnz) ED_HIP_CHECK(hipMemcpy(offdiag, T.off, sizeof(double) * T.nnz, hipMemcpyDeviceToHost));
    return ED_OK;
}

int ed_sector_tables_im(const ed_sector *sec, int spin, double *offdiag_im)
{
    if (!sec || (spin != 0 && spin != 1) || !offdiag_im) return ed::fail(ED_ERR_ARG, "ed_sector_tables_im: bad argument");
    if (!sec->cplx) return ed::fail(ED_ERR_ARG, "ed_sector_tables_im: not a complex sector");
    const ed::SpinTables &T = spin ? sec->dn : sec->up;
    if (T.nnz) ED_HIP_CHECK(hipMemcpy(offdiag_im, T.off_im, sizeof(double) * T.nnz, hipMemcpyDeviceToHost));
    return ED_OK;
}

}  // extern "C"
