/*
 * ed_oracle.c -- CPU restatement of the reference's exact-diagonalisation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is compiled into
 * oracle/libed_oracle.so and may be loaded only by tests/, by the smoke check
 * and by the cpu_baseline leg of bench.py, as the checker or as the timed CPU
 * baseline.  The product library (ed_solver_amd) never links or calls it.
 *
 * It restates, in plain C99 + OpenMP, the algorithms of the Fortran reference
 * francoisjamet/ed_solver (file:line citations are relative to the reference
 * repository, ed_solver/ directory):
 *
 *   orc_sector_states / orc_rank_table
 *       quantum_fermion_hilbert_class.f90:346-421 (enumerate 0..2^Ns-1, bucket by
 *       popcount, rank = 1-based position inside the bucket)
 *   orc_destroy / orc_create / orc_hop
 *       quantum_fermion_ket_class.f90:483-618 (sign = (-1)^{# occupied above})
 *   orc_tab_spin
 *       HAIM_offdiag_arrays.f90:210-606 (tab_HAIM2: diag, noff, rankoff, offdiag,
 *       entry order impurity hops, bath hops, V b+c, V* c+b)
 *   orc_hmult
 *       HAIMupdo.f90:677-1195 (HAIMupdo_multr, gather form, OpenMP split of the
 *       up-spin range as openmpmod.f90:135-175)
 *   orc_lanczos_step
 *       Lanczos_plain.f90:247-302 + Lanczos_plain.h:66-173 (one_step_Lanczos_fast,
 *       lanc_opt)
 *   orc_lanczos_diagonalize
 *       Lanczos_plain.f90:33-134 (Lanczos_fast_diagonalize convergence loop)
 *   orc_lanczos_gs
 *       Lanczos_plain.f90:156-229 (Lanczos_get_GS_sector second pass)
 *   orc_tridiag_eigvals / orc_tridiag_eigvec
 *       the role of DSTEDC in matrices_tridiag.f90:210-259; restated as Sturm
 *       bisection (eigenvalues) + inverse iteration (vectors), a textbook method,
 *       not LAPACK's divide and conquer.
 *   orc_cfrac
 *       matrices_tridiag.f90:163-198 (invert_zmtridiag)
 *   orc_apply_c
 *       apply_C.f90:497-598 (apply_Cupdo)
 *   orc_dynamic
 *       green_class_compute_dynamic.f90:191-407 (compute_dynamic_cpu, real build)
 *
 * Conventions (all indices 0-based here; the reference is 1-based):
 *   per-spin bit layout: bath site k -> bit k, impurity orbital c -> bit nb+c
 *   (quantum_algebra_Anderson_Impurity_class.f90:73-86);
 *   sector vector index istate = iup + ido*d_up (column-major d_up x d_dn,
 *   quantum_fermion_sector_class.f90:195-200).
 *   rankoff entries are 0-based ranks (reference value minus one).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int32_t nc, nb;
    const double *ec;   /* [2][nc][nc] row-major, Ec(sigma)(i,j) */
    const double *eb;   /* [2][nb][nb] */
    const double *vbc;  /* [2][nb][nc] */
    const double *u;    /* [nc][nc] U_ii and U'_ij (already U-2J off the diagonal) */
    const double *jj;   /* [nc][nc] Hund matrix JJmatrix */
    double jhund;       /* scalar Jhund gating the spin-flip / pair-hop terms */
    double cutoff;      /* cutoff_hamilt_param */
} orc_model;

/* ---------------------------------------------------------------- Fock basis */

/* quantum_fermion_hilbert_class.f90:385-404: scan every state in increasing order,
 * append it to its popcount bucket. */
int orc_sector_states(int ns, int n, int32_t *states)
{
    int cnt = 0;
    for (int32_t s = 0; s < (1 << ns); ++s) {
        int pc = __builtin_popcount((unsigned)s);
        if (pc == n) states[cnt++] = s;
    }
    return cnt;
}

/* rank(state) = 1-based position of state inside its popcount bucket (:403). */
void orc_rank_table(int ns, int32_t *rank)
{
    int cnt[33] = {0};
    for (int32_t s = 0; s < (1 << ns); ++s) {
        int pc = __builtin_popcount((unsigned)s);
        rank[s] = ++cnt[pc];
    }
}

/* --------------------------------------------------------- fermion operators */

/* Fermion_sign (quantum_fermion_ket_class.f90:605-618): multiply by (-1) for every
 * occupied orbital above iorb in the ket *after* the operator acted. */
static int fermion_sign_above(uint32_t s, int bit, int norbs)
{
    int sg = 1;
    for (int j = bit + 1; j < norbs; ++j)
        if ((s >> j) & 1u) sg = -sg;
    return sg;
}

/* destroy (:560-580) */
static int orc_destroy(uint32_t in, int bit, int norbs, uint32_t *out, int *sign)
{
    if (!((in >> bit) & 1u)) return 0;
    uint32_t t = in & ~(1u << bit);
    *sign *= fermion_sign_above(t, bit, norbs);
    *out = t;
    return 1;
}

/* create (:535-555) */
static int orc_create(uint32_t in, int bit, int norbs, uint32_t *out, int *sign)
{
    if ((in >> bit) & 1u) return 0;
    uint32_t t = in | (1u << bit);
    *sign *= fermion_sign_above(t, bit, norbs);
    *out = t;
    return 1;
}

/* hop(out, jorb, iorb, in) = c+_jorb c_iorb (:483-492): destroy iorb, then create jorb. */
static int orc_hop(uint32_t in, int bit_to, int bit_from, int norbs, uint32_t *out, int *sign)
{
    uint32_t t;
    *sign = 1;
    if (!orc_destroy(in, bit_from, norbs, &t, sign)) return 0;
    if (!orc_create(t, bit_to, norbs, out, sign)) return 0;
    return 1;
}

/* ------------------------------------------------------ quadratic tables */

/* tab_HAIM2 for one spin (HAIM_offdiag_arrays.f90:210-606).  If diag/rankoff/offdiag
 * are NULL only noff is filled (the NO_STORAGE count pass, :89-148).  Returns nnz. */
int64_t orc_tab_spin(const orc_model *m, int spin, int n, const int32_t *states, int d,
                     const int32_t *rank, double *diag, int32_t *noff, int32_t *rankoff,
                     double *offdiag)
{
    const int nc = m->nc, nb = m->nb, ns = nc + nb;
    const double *ec = m->ec + (size_t)spin * nc * nc;
    const double *eb = m->eb + (size_t)spin * nb * nb;
    const double *vbc = m->vbc + (size_t)spin * nb * nc;
    const double cut = m->cutoff;
    int64_t k = 0;
    (void)n;
    for (int i = 0; i < d; ++i) {
        uint32_t s = (uint32_t)states[i];
        int64_t k0 = k;
        uint32_t o;
        int sg;
        if (diag) {
            double dg = 0.0;
            for (int c = 0; c < nc; ++c)
                if ((s >> (nb + c)) & 1u) dg += ec[c * nc + c];
            for (int b = 0; b < nb; ++b)
                if ((s >> b) & 1u) dg += eb[b * nb + b];
            diag[i] = dg;
        }
        /* impurity hops Ec(i,j) c+_i c_j, i != j (:437-452) */
        for (int a = 0; a < nc; ++a)
            for (int b = 0; b < nc; ++b) {
                double t = ec[a * nc + b];
                if (a == b || t == 0.0 || !(fabs(t) > cut)) continue;
                if (orc_hop(s, nb + a, nb + b, ns, &o, &sg)) {
                    if (rankoff) { rankoff[k] = rank[o] - 1; offdiag[k] = t * sg; }
                    ++k;
                }
            }
        /* bath hops Eb(k,l) (:457-474) */
        for (int a = 0; a < nb; ++a)
            for (int b = 0; b < nb; ++b) {
                double t = eb[a * nb + b];
                if (a == b || t == 0.0 || !(fabs(t) > cut)) continue;
                if (orc_hop(s, a, b, ns, &o, &sg)) {
                    if (rankoff) { rankoff[k] = rank[o] - 1; offdiag[k] = t * sg; }
                    ++k;
                }
            }
        /* hybridisation, bath outer / impurity inner (:479-521) */
        for (int b = 0; b < nb; ++b)
            for (int c = 0; c < nc; ++c) {
                double t = vbc[b * nc + c];
                if (t == 0.0 || !(fabs(t) > cut)) continue;
                if (orc_hop(s, b, nb + c, ns, &o, &sg)) {  /* b+_k c_c */
                    if (rankoff) { rankoff[k] = rank[o] - 1; offdiag[k] = t * sg; }
                    ++k;
                }
                if (orc_hop(s, nb + c, b, ns, &o, &sg)) {  /* c+_c b_k, conj(V) = V (real) */
                    if (rankoff) { rankoff[k] = rank[o] - 1; offdiag[k] = t * sg; }
                    ++k;
                }
            }
        noff[i] = (int32_t)(k - k0);
    }
    return k;
}

/* --------------------------------------------------------------- sector */

typedef struct {
    orc_model m;
    double *ec, *eb, *vbc, *u, *jj;   /* owned copies of the model arrays */
    int ns, nup, ndn, dup, ddn;
    int64_t dim;
    int32_t *st_up, *st_dn, *rank;
    double *diag_up, *diag_dn, *off_up, *off_dn;
    int32_t *noff_up, *noff_dn, *rk_up, *rk_dn;
    int64_t *ptr_up, *ptr_dn;         /* CSR offsets (prefix sums of noff) */
    int64_t nnz_up, nnz_dn;
    int offdiag_coulomb;
    int nthreads;
} orc_sector;

static double *dupd(const double *p, size_t n)
{
    double *q = (double *)calloc(n ? n : 1, sizeof(double));
    if (n) memcpy(q, p, n * sizeof(double));
    return q;
}

/* new_HAIMupdo (HAIMupdo.f90:53-121): both spin tables + U mask bookkeeping. */
orc_sector *orc_sector_create(const orc_model *m, int nup, int ndn, int nthreads)
{
    orc_sector *S = (orc_sector *)calloc(1, sizeof(orc_sector));
    int nc = m->nc, nb = m->nb, ns = nc + nb;
    S->ec = dupd(m->ec, 2 * (size_t)nc * nc);
    S->eb = dupd(m->eb, 2 * (size_t)nb * nb);
    S->vbc = dupd(m->vbc, 2 * (size_t)nb * nc);
    S->u = dupd(m->u, (size_t)nc * nc);
    S->jj = dupd(m->jj, (size_t)nc * nc);
    S->m = *m;
    S->m.ec = S->ec; S->m.eb = S->eb; S->m.vbc = S->vbc; S->m.u = S->u; S->m.jj = S->jj;
    S->ns = ns; S->nup = nup; S->ndn = ndn;
    S->nthreads = nthreads > 0 ? nthreads : 1;
    S->rank = (int32_t *)malloc(sizeof(int32_t) << ns);
    orc_rank_table(ns, S->rank);
    S->st_up = (int32_t *)malloc(sizeof(int32_t) << ns);
    S->st_dn = (int32_t *)malloc(sizeof(int32_t) << ns);
    S->dup = orc_sector_states(ns, nup, S->st_up);
    S->ddn = orc_sector_states(ns, ndn, S->st_dn);
    S->dim = (int64_t)S->dup * S->ddn;
    S->diag_up = (double *)malloc(sizeof(double) * (S->dup + 1));
    S->diag_dn = (double *)malloc(sizeof(double) * (S->ddn + 1));
    S->noff_up = (int32_t *)malloc(sizeof(int32_t) * (S->dup + 1));
    S->noff_dn = (int32_t *)malloc(sizeof(int32_t) * (S->ddn + 1));
    S->nnz_up = orc_tab_spin(m, 0, nup, S->st_up, S->dup, S->rank, NULL, S->noff_up, NULL, NULL);
    S->nnz_dn = orc_tab_spin(m, 1, ndn, S->st_dn, S->ddn, S->rank, NULL, S->noff_dn, NULL, NULL);
    S->rk_up = (int32_t *)malloc(sizeof(int32_t) * (S->nnz_up + 1));
    S->rk_dn = (int32_t *)malloc(sizeof(int32_t) * (S->nnz_dn + 1));
    S->off_up = (double *)malloc(sizeof(double) * (S->nnz_up + 1));
    S->off_dn = (double *)malloc(sizeof(double) * (S->nnz_dn + 1));
    orc_tab_spin(m, 0, nup, S->st_up, S->dup, S->rank, S->diag_up, S->noff_up, S->rk_up, S->off_up);
    orc_tab_spin(m, 1, ndn, S->st_dn, S->ddn, S->rank, S->diag_dn, S->noff_dn, S->rk_dn, S->off_dn);
    S->ptr_up = (int64_t *)malloc(sizeof(int64_t) * (S->dup + 1));
    S->ptr_dn = (int64_t *)malloc(sizeof(int64_t) * (S->ddn + 1));
    S->ptr_up[0] = 0;
    for (int i = 0; i < S->dup; ++i) S->ptr_up[i + 1] = S->ptr_up[i] + S->noff_up[i];
    S->ptr_dn[0] = 0;
    for (int i = 0; i < S->ddn; ++i) S->ptr_dn[i + 1] = S->ptr_dn[i] + S->noff_dn[i];
    /* offdiag_coulomb = ANY(UMASK .AND. .NOT. diag) (HAIMupdo.f90:117-118) */
    S->offdiag_coulomb = 0;
    for (int a = 0; a < nc; ++a)
        for (int b = 0; b < nc; ++b)
            if (a != b && S->u[a * nc + b] != 0.0) S->offdiag_coulomb = 1;
    return S;
}

void orc_sector_destroy(orc_sector *S)
{
    if (!S) return;
    free(S->ec); free(S->eb); free(S->vbc); free(S->u); free(S->jj);
    free(S->rank); free(S->st_up); free(S->st_dn);
    free(S->diag_up); free(S->diag_dn); free(S->noff_up); free(S->noff_dn);
    free(S->rk_up); free(S->rk_dn); free(S->off_up); free(S->off_dn);
    free(S->ptr_up); free(S->ptr_dn);
    free(S);
}

int64_t orc_sector_dim(const orc_sector *S) { return S->dim; }
int orc_sector_dup(const orc_sector *S) { return S->dup; }
int orc_sector_ddn(const orc_sector *S) { return S->ddn; }
int64_t orc_sector_nnz(const orc_sector *S, int spin) { return spin ? S->nnz_dn : S->nnz_up; }
void orc_set_threads(orc_sector *S, int nthreads) { S->nthreads = nthreads > 0 ? nthreads : 1; }

/* Copy one spin's tables out (for bit-exact comparison with the device tables). */
void orc_sector_tables(const orc_sector *S, int spin, int32_t *states, double *diag,
                       int32_t *noff, int32_t *rankoff, double *offdiag)
{
    int d = spin ? S->ddn : S->dup;
    int64_t nnz = spin ? S->nnz_dn : S->nnz_up;
    if (states) memcpy(states, spin ? S->st_dn : S->st_up, sizeof(int32_t) * d);
    if (diag) memcpy(diag, spin ? S->diag_dn : S->diag_up, sizeof(double) * d);
    if (noff) memcpy(noff, spin ? S->noff_dn : S->noff_up, sizeof(int32_t) * d);
    if (rankoff) memcpy(rankoff, spin ? S->rk_dn : S->rk_up, sizeof(int32_t) * nnz);
    if (offdiag) memcpy(offdiag, spin ? S->off_dn : S->off_up, sizeof(double) * nnz);
}

/* parity of the occupied bits strictly between bits a and b (the reference's
 * "DO jj=i1+1,i2-1; IF(BTEST(...)) sign=-sign" loops, HAIMupdo.f90:1044-1047). */
static int between_sign(uint32_t s, int a, int b)
{
    int lo = a < b ? a : b, hi = a < b ? b : a, sg = 1;
    for (int j = lo + 1; j < hi; ++j)
        if ((s >> j) & 1u) sg = -sg;
    return sg;
}

/* openmp_split_array (openmpmod.f90:135-175): first MOD(n,np) chunks get one extra. */
static void split_range(int n, int np, int t, int *lo, int *hi)
{
    int pp = n % np, base = n / np;
    int start = 0;
    for (int i = 0; i < t; ++i) start += base + (i < pp ? 1 : 0);
    *lo = start;
    *hi = start + base + (t < pp ? 1 : 0);
}

/* HAIMupdo_multr (HAIMupdo.f90:677-1195) for the real build, one MPI rank, gather
 * form (USE_CC), Slater 4-index term off (flag_slater_int=.false., the default). */
void orc_hmult(const orc_sector *S, const double *vin, double *vout)
{
    const int nc = S->m.nc, nb = S->m.nb;
    const int dup = S->dup, ddn = S->ddn;
    const double *U = S->u, *JJ = S->jj;
    const int hund = fabs(S->m.jhund) > 1e-5 && S->offdiag_coulomb;
    const int np = S->nthreads;

    memset(vout, 0, sizeof(double) * (size_t)S->dim);
#pragma omp parallel num_threads(np)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        int i0, i1;
        split_range(dup, np, tid, &i0, &i1);

        /* (i) onsite U_ii n_i,up n_i,dn (:785-795) */
        for (int iup = i0; iup < i1; ++iup)
            for (int c = 0; c < nc; ++c) {
                if (U[c * nc + c] == 0.0 || !((S->st_up[iup] >> (nb + c)) & 1)) continue;
                for (int ido = 0; ido < ddn; ++ido)
                    if ((S->st_dn[ido] >> (nb + c)) & 1) {
                        int64_t is = iup + (int64_t)ido * dup;
                        vout[is] += U[c * nc + c] * vin[is];
                    }
            }
        /* (ii) inter-orbital U and Slater-type -J same-spin term (:798-829) */
        if (S->offdiag_coulomb)
            for (int iup = i0; iup < i1; ++iup)
                for (int ido = 0; ido < ddn; ++ido)
                    for (int c1 = 0; c1 < nc - 1; ++c1) {
                        int u1 = (S->st_up[iup] >> (nb + c1)) & 1, d1 = (S->st_dn[ido] >> (nb + c1)) & 1;
                        int n1 = u1 + d1;
                        if (!n1) continue;
                        for (int c2 = c1 + 1; c2 < nc; ++c2) {
                            if (U[c1 * nc + c2] == 0.0) continue;
                            int u2 = (S->st_up[iup] >> (nb + c2)) & 1, d2 = (S->st_dn[ido] >> (nb + c2)) & 1;
                            int n2 = u2 + d2;
                            if (!n2) continue;
                            int64_t is = iup + (int64_t)ido * dup;
                            vout[is] += U[c1 * nc + c2] * n1 * n2 * vin[is];
                            if (u1 && u2) vout[is] -= JJ[c1 * nc + c2] * vin[is];
                            if (d1 && d2) vout[is] -= JJ[c1 * nc + c2] * vin[is];
                        }
                    }
        /* (iii) quadratic diagonal (:841-848) */
        for (int iup = i0; iup < i1; ++iup)
            for (int ido = 0; ido < ddn; ++ido) {
                int64_t is = iup + (int64_t)ido * dup;
                vout[is] += (S->diag_up[iup] + S->diag_dn[ido]) * vin[is];
            }
        /* (iv) up hops: stride-d_up walk over ido (:865-881) */
        for (int iup = i0; iup < i1; ++iup)
            for (int64_t k = S->ptr_up[iup]; k < S->ptr_up[iup + 1]; ++k) {
                double h = S->off_up[k];
                int jup = S->rk_up[k];
                for (int ido = 0; ido < ddn; ++ido)
                    vout[iup + (int64_t)ido * dup] += h * vin[jup + (int64_t)ido * dup];
            }
        /* (v) down hops restricted to this thread's up range (:883-900) */
        for (int ido = 0; ido < ddn; ++ido)
            for (int64_t k = S->ptr_dn[ido]; k < S->ptr_dn[ido + 1]; ++k) {
                double h = S->off_dn[k];
                int jdo = S->rk_dn[k];
                for (int iup = i0; iup < i1; ++iup)
                    vout[iup + (int64_t)ido * dup] += h * vin[iup + (int64_t)jdo * dup];
            }
        if (hund) {
            /* (vii) Hund spin flip, element +JJ * s_up * s_dn (:1017-1085) */
            for (int iup = i0; iup < i1; ++iup)
                for (int ido = 0; ido < ddn; ++ido) {
                    int64_t is = iup + (int64_t)ido * dup;
                    uint32_t su = (uint32_t)S->st_up[iup], sd = (uint32_t)S->st_dn[ido];
                    for (int c1 = 0; c1 < nc; ++c1) {
                        int u1 = (su >> (nb + c1)) & 1, d1 = (sd >> (nb + c1)) & 1;
                        if (u1 + d1 != 1) continue;
                        for (int c2 = c1 + 1; c2 < nc; ++c2) {
                            if (U[c1 * nc + c2] == 0.0) continue;
                            int u2 = (su >> (nb + c2)) & 1, d2 = (sd >> (nb + c2)) & 1;
                            if (u2 + d2 != 1) continue;
                            int b1 = nb + c1, b2 = nb + c2;
                            if (u1 && d2) {
                                uint32_t uo = (su & ~(1u << b1)) | (1u << b2);
                                uint32_t dq = (sd & ~(1u << b2)) | (1u << b1);
                                int sg = between_sign(uo, b1, b2) * between_sign(dq, b1, b2);
                                int64_t js = (S->rank[uo] - 1) + (int64_t)(S->rank[dq] - 1) * dup;
                                vout[is] += vin[js] * JJ[c1 * nc + c2] * (double)sg;
                            }
                            if (d1 && u2) {
                                uint32_t uo = (su & ~(1u << b2)) | (1u << b1);
                                uint32_t dq = (sd & ~(1u << b1)) | (1u << b2);
                                int sg = between_sign(uo, b1, b2) * between_sign(dq, b1, b2);
                                int64_t js = (S->rank[uo] - 1) + (int64_t)(S->rank[dq] - 1) * dup;
                                vout[is] += vin[js] * JJ[c1 * nc + c2] * (double)sg;
                            }
                        }
                    }
                }
            /* (viii) pair hopping i (doubly occupied) -> j (empty) (:1087-1129) */
            for (int iup = i0; iup < i1; ++iup)
                for (int ido = 0; ido < ddn; ++ido) {
                    int64_t is = iup + (int64_t)ido * dup;
                    uint32_t su = (uint32_t)S->st_up[iup], sd = (uint32_t)S->st_dn[ido];
                    for (int c1 = 0; c1 < nc; ++c1) {
                        int b1 = nb + c1;
                        if (!(((su >> b1) & 1) && ((sd >> b1) & 1))) continue;
                        for (int c2 = 0; c2 < nc; ++c2) {
                            if (c2 == c1 || U[c1 * nc + c2] == 0.0) continue;
                            int b2 = nb + c2;
                            if (((su >> b2) & 1) || ((sd >> b2) & 1)) continue;
                            uint32_t uo = (su & ~(1u << b1)) | (1u << b2);
                            uint32_t dq = (sd & ~(1u << b1)) | (1u << b2);
                            int sg = between_sign(uo, b1, b2) * between_sign(dq, b1, b2);
                            int64_t js = (S->rank[uo] - 1) + (int64_t)(S->rank[dq] - 1) * dup;
                            vout[is] += vin[js] * JJ[c1 * nc + c2] * (double)sg;
                        }
                    }
                }
        }
    }
}

/* ------------------------------------------------------- Lanczos recurrence */

typedef struct {
    double normv_in, normv_out;  /* the SAVEd norms of one_step_Lanczos_fast (:255) */
    int n;                       /* tri%N */
    double *diag, *subdiag;      /* subdiag[0] = ||v1|| (Dagotto storage) */
    int np;
} orc_lanczos_ctx;

/* one_step_Lanczos_fast + lanc_opt (Lanczos_plain.f90:247-302, Lanczos_plain.h:66-173).
 * iter is 1-based as in the reference.  vec_in, tmp, vec_out have length dim.
 * Returns 0 on success, 1 on the 'converged in one step' flag, -1 when the
 * reference would STOP (null start vector / zero norm). */
int orc_lanczos_step(const orc_sector *S, int iter, double *vec_in, double *tmp, double *vec_out,
                     double *diag, double *subdiag, int ntri, double *norms)
{
    const int64_t D = S->dim;
    const int np = S->nthreads;
    double normv_in = norms[0], normv_out = norms[1];
    if (iter == 1) {
        double s = 0.0;
#pragma omp parallel for reduction(+ : s) num_threads(np)
        for (int64_t i = 0; i < D; ++i) s += vec_in[i] * vec_in[i];
        normv_in = sqrt(fabs(s));
        normv_out = 0.0;
        for (int i = 0; i < ntri; ++i) { diag[i] = 0.0; subdiag[i] = 0.0; }
        subdiag[0] = normv_in;
    } else if (iter > ntri) {
        return -2;
    }
    /* normalise the incoming vector (lanc_opt :69-101) */
    double aa;
    if (iter == 1) {
        if (fabs(normv_in) < 1e-25) return -1;
        aa = 1.0 / normv_in;
    } else {
        if (fabs(normv_out) < 1e-25) return -1;
        aa = 1.0 / normv_out;
    }
#pragma omp parallel for num_threads(np)
    for (int64_t i = 0; i < D; ++i) vec_in[i] *= aa;

    orc_hmult(S, vec_in, vec_out);

    double dd = 0.0;
    if (iter > 1) {
        double b = subdiag[iter - 1];
#pragma omp parallel for reduction(+ : dd) num_threads(np)
        for (int64_t i = 0; i < D; ++i) {
            vec_out[i] = vec_out[i] - b * tmp[i];
            dd += vec_out[i] * vec_in[i];
        }
    } else {
#pragma omp parallel for reduction(+ : dd) num_threads(np)
        for (int64_t i = 0; i < D; ++i) dd += vec_out[i] * vec_in[i];
    }
    diag[iter - 1] = dd;
    double nn = 0.0;
#pragma omp parallel for reduction(+ : nn) num_threads(np)
    for (int64_t i = 0; i < D; ++i) {
        vec_out[i] = vec_out[i] - dd * vec_in[i];
        nn += vec_out[i] * vec_out[i];
    }
    normv_out = sqrt(nn);
    if (iter < ntri) subdiag[iter] = normv_out;
#pragma omp parallel for num_threads(np)
    for (int64_t i = 0; i < D; ++i) {
        tmp[i] = vec_in[i];
        vec_in[i] = vec_out[i];
    }
    norms[0] = normv_in;
    norms[1] = normv_out;
    return (fabs(normv_out) < 1e-24 && iter == 1) ? 1 : 0;
}

/* -------------------------------------------- symmetric tridiagonal (oracle) */

/* Sturm count: number of eigenvalues of T (diag a[0..n-1], off-diagonal b[1..n-1])
 * strictly below x. */
static int sturm_count(int n, const double *a, const double *b, double x)
{
    int cnt = 0;
    double q = a[0] - x;
    if (q < 0) ++cnt;
    for (int i = 1; i < n; ++i) {
        if (q == 0.0) q = 1e-300;
        q = a[i] - x - b[i] * b[i] / q;
        if (q < 0) ++cnt;
    }
    return cnt;
}

/* The kmax lowest eigenvalues of the tridiagonal T, ascending, by bisection on the
 * Sturm count (Wilkinson, The Algebraic Eigenvalue Problem, ch. 5).  b[0] is not
 * referenced. */
void orc_tridiag_eigvals(int n, const double *a, const double *b, int kmax, double *w)
{
    double lo = a[0], hi = a[0];
    for (int i = 0; i < n; ++i) {
        double r = (i > 0 ? fabs(b[i]) : 0.0) + (i + 1 < n ? fabs(b[i + 1]) : 0.0);
        if (a[i] - r < lo) lo = a[i] - r;
        if (a[i] + r > hi) hi = a[i] + r;
    }
    double span = hi - lo;
    lo -= 1e-12 * (fabs(span) + 1.0);
    hi += 1e-12 * (fabs(span) + 1.0);
    if (kmax > n) kmax = n;
    for (int k = 0; k < kmax; ++k) {
        double l = lo, h = hi;
        for (int it = 0; it < 200; ++it) {
            double mid = 0.5 * (l + h);
            if (mid == l || mid == h) break;
            if (sturm_count(n, a, b, mid) > k) h = mid;
            else l = mid;
        }
        w[k] = 0.5 * (l + h);
    }
}

/* Eigenvector of T for eigenvalue lam by inverse iteration (two sweeps, tridiagonal
 * Gaussian elimination with partial pivoting).  Normalised, sign fixed so that
 * the largest-magnitude component is positive. */
void orc_tridiag_eigvec(int n, const double *a, const double *b, double lam, double *x)
{
    double *dl = (double *)malloc(sizeof(double) * n), *dd = (double *)malloc(sizeof(double) * n);
    double *du = (double *)malloc(sizeof(double) * n), *du2 = (double *)malloc(sizeof(double) * n);
    int *piv = (int *)malloc(sizeof(int) * n);
    double scale = 0.0;
    for (int i = 0; i < n; ++i) scale += fabs(a[i]) + (i ? fabs(b[i]) : 0.0);
    double shift = lam + 1e-14 * (scale + 1.0);
    for (int i = 0; i < n; ++i) {
        dd[i] = a[i] - shift;
        dl[i] = (i + 1 < n) ? b[i + 1] : 0.0;
        du[i] = (i + 1 < n) ? b[i + 1] : 0.0;
        du2[i] = 0.0;
    }
    /* LU with partial pivoting of the tridiagonal (T - shift) */
    for (int i = 0; i < n - 1; ++i) {
        if (fabs(dd[i]) >= fabs(dl[i])) {
            piv[i] = 0;
            if (dd[i] == 0.0) dd[i] = 1e-300;
            double f = dl[i] / dd[i];
            dl[i] = f;
            dd[i + 1] -= f * du[i];
        } else {
            piv[i] = 1;
            double f = dd[i] / dl[i];
            dd[i] = dl[i];
            dl[i] = f;
            double t = du[i];
            du[i] = dd[i + 1];
            dd[i + 1] = t - f * dd[i + 1];
            if (i < n - 2) {
                du2[i] = du[i + 1];
                du[i + 1] = -f * du[i + 1];
            }
        }
    }
    if (dd[n - 1] == 0.0) dd[n - 1] = 1e-300;
    for (int i = 0; i < n; ++i) x[i] = 1.0;
    for (int sweep = 0; sweep < 3; ++sweep) {
        for (int i = 0; i < n - 1; ++i) {
            if (piv[i] == 0) x[i + 1] -= dl[i] * x[i];
            else { double t = x[i]; x[i] = x[i + 1]; x[i + 1] = t - dl[i] * x[i + 1]; }
        }
        x[n - 1] /= dd[n - 1];
        if (n > 1) x[n - 2] = (x[n - 2] - du[n - 2] * x[n - 1]) / dd[n - 2];
        for (int i = n - 3; i >= 0; --i) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / dd[i];
        double nrm = 0.0;
        for (int i = 0; i < n; ++i) nrm += x[i] * x[i];
        nrm = sqrt(nrm);
        for (int i = 0; i < n; ++i) x[i] /= nrm;
    }
    int imax = 0;
    for (int i = 1; i < n; ++i) if (fabs(x[i]) > fabs(x[imax])) imax = i;
    if (x[imax] < 0) for (int i = 0; i < n; ++i) x[i] = -x[i];
    free(dl); free(dd); free(du); free(du2); free(piv);
}

/* ------------------------------------------------ Lanczos_fast_diagonalize */

/* Lanczos_fast_diagonalize (Lanczos_plain.f90:33-134) for Neigen lowest states,
 * start vector start[dim] (copied).  Outputs: alpha/beta[0..iters-1] (beta[0] =
 * ||v1||), evals[0..neig-1] of T_iters, vecp[neig][iters] Lanczos coefficients.
 * Returns the iteration count at exit (<= min(dim-1, nitermax)); negative on a
 * reference STOP. */
int orc_lanczos_diagonalize(const orc_sector *S, int nitermax, int neigen, double tol,
                            const double *start, double *alpha, double *beta, double *evals,
                            double *vecp, int *conv_flags)
{
    const int64_t D = S->dim;
    int niter = (int)((D - 1) < nitermax ? (D - 1) : nitermax);
    int neig = neigen < D ? neigen : (int)D;
    if (neig > niter) neig = niter;
    if (niter < 1) return -3;
    double *v = (double *)malloc(sizeof(double) * D), *tmp = (double *)calloc(D, sizeof(double));
    double *last = (double *)calloc(D, sizeof(double));
    double *w = (double *)malloc(sizeof(double) * niter), *lastev = (double *)malloc(sizeof(double) * (neig + 1));
    memcpy(v, start, sizeof(double) * D);
    for (int k = 0; k < neig; ++k) lastev[k] = 1e300;
    double norms[2] = {0, 0};
    int iter, ret = 0;
    for (iter = 1; iter <= niter; ++iter) {
        int st = orc_lanczos_step(S, iter, v, tmp, last, alpha, beta, niter, norms);
        if (st < 0) { ret = st; break; }
        orc_tridiag_eigvals(iter, alpha, beta, neig, w);
        int m = neig < iter ? neig : iter, allc = 1;
        for (int k = 0; k < neig; ++k) {
            int c = 0;
            if (k < m) c = fabs(w[k] - lastev[k]) < tol && iter > 2;
            if (conv_flags) conv_flags[k] = c;
            if (!c) allc = 0;
        }
        if (allc || iter == niter || st == 1) {
            for (int k = 0; k < neig; ++k) {
                evals[k] = w[k];
                orc_tridiag_eigvec(iter, alpha, beta, w[k], vecp + (size_t)k * iter);
            }
            ret = iter;
            break;
        }
        for (int k = 0; k < m; ++k) lastev[k] = w[k];
    }
    free(v); free(tmp); free(last); free(w); free(lastev);
    return ret;
}

/* Lanczos_get_GS_sector (Lanczos_plain.f90:156-229): rebuild psi = sum_j vecp(j) v_j
 * by re-running the recurrence from the same start vector. */
int orc_lanczos_gs(const orc_sector *S, int iters, const double *vecp, const double *start, double *psi)
{
    const int64_t D = S->dim;
    double *v = (double *)malloc(sizeof(double) * D), *tmp = (double *)calloc(D, sizeof(double));
    double *last = (double *)calloc(D, sizeof(double));
    double *dg = (double *)calloc(iters + 1, sizeof(double)), *sd = (double *)calloc(iters + 1, sizeof(double));
    double norms[2] = {0, 0};
    const double rerror = 1e-13;
    memcpy(v, start, sizeof(double) * D);
    double aa = 0.0;
    for (int64_t i = 0; i < D; ++i) aa += v[i] * v[i];
    aa = sqrt(aa);
    if (fabs(aa) < rerror) aa = rerror;
    aa = 1.0 / aa;
    for (int64_t i = 0; i < D; ++i) psi[i] = vecp[0] * v[i] * aa;
    int st = orc_lanczos_step(S, 1, v, tmp, last, dg, sd, iters, norms);
    if (st == 0) {
        for (int it2 = 2; it2 <= iters - 1; ++it2) {
            double n2 = 0.0;
            for (int64_t i = 0; i < D; ++i) n2 += last[i] * last[i];
            n2 = sqrt(n2);
            if (fabs(n2) < rerror) n2 = rerror;
            n2 = 1.0 / n2;
            for (int64_t i = 0; i < D; ++i) { last[i] *= n2; psi[i] += vecp[it2 - 1] * last[i]; }
            orc_lanczos_step(S, it2, v, tmp, last, dg, sd, iters, norms);
        }
        if (iters >= 2) {
            double n2 = 0.0;
            for (int64_t i = 0; i < D; ++i) n2 += last[i] * last[i];
            n2 = sqrt(n2);
            if (fabs(n2) < rerror) n2 = rerror;
            n2 = 1.0 / n2;
            for (int64_t i = 0; i < D; ++i) { last[i] *= n2; psi[i] += vecp[iters - 1] * last[i]; }
        }
        double nn = 0.0;
        for (int64_t i = 0; i < D; ++i) nn += psi[i] * psi[i];
        nn = sqrt(nn);
        if (fabs(nn) < rerror) nn = rerror;
        for (int64_t i = 0; i < D; ++i) psi[i] /= nn;
    }
    free(v); free(tmp); free(last); free(dg); free(sd);
    return st;
}

/* Fixed-length recurrence (no convergence test): the loop of compute_dynamic_cpu
 * (green_class_compute_dynamic.f90:279-299).  alpha/beta length niter. */
int orc_lanczos_run(const orc_sector *S, int niter, const double *start, double *alpha, double *beta)
{
    const int64_t D = S->dim;
    double *v = (double *)malloc(sizeof(double) * D), *tmp = (double *)calloc(D, sizeof(double));
    double *last = (double *)calloc(D, sizeof(double));
    double norms[2] = {0, 0};
    memcpy(v, start, sizeof(double) * D);
    int ret = 0;
    for (int iter = 1; iter <= niter; ++iter) {
        int st = orc_lanczos_step(S, iter, v, tmp, last, alpha, beta, niter, norms);
        if (st < 0) { ret = st; break; }
        if (st == 1) {
            /* compute_dynamic_cpu zeroes lastvec after a one-step convergence (:295-298) */
            memset(last, 0, sizeof(double) * D);
        }
    }
    free(v); free(tmp); free(last);
    return ret;
}

/* ----------------------------------------------------- continued fraction */

/* invert_zmtridiag (matrices_tridiag.f90:163-198): <0|(z-T)^-1|0>, evaluated from
 * the bottom of T upward.  Returns 0, or -1 where the reference would STOP on a
 * denominator below machine epsilon. */
int orc_cfrac(int n, const double *a, const double *b, double zr, double zi, double *out_re, double *out_im)
{
    double complex z = zr + zi * I;
    double complex den = z - a[n - 1];
    if (cabs(den) < 2.220446049250313e-16) return -1;
    double complex r = 1.0 / den;
    for (int it = n - 2; it >= 0; --it) {
        den = z - a[it] - r * b[it + 1] * b[it + 1];
        if (cabs(den) < 2.220446049250313e-16) return -1;
        r = 1.0 / den;
    }
    *out_re = creal(r);
    *out_im = cimag(r);
    return 0;
}

/* ------------------------------------------------------------- apply_C */

/* apply_Cupdo (apply_C.f90:497-598) for one impurity orbital `site`, spin (0 up,
 * 1 down), pm = +1 (c+) or -1 (c).  psi lives in sector (nup, ndn) with tables
 * from `src`; phi in `dst` = (nup +- 1, ndn) or (nup, ndn +- 1).  phi is zeroed
 * first.  Returns the number of spin states with a non-nil image. */
int orc_apply_c(const orc_sector *src, const orc_sector *dst, int site, int spin, int pm,
                const double *psi, double *phi)
{
    const int nb = src->m.nb, ns = src->ns;
    const int bit = nb + site;
    memset(phi, 0, sizeof(double) * (size_t)dst->dim);
    int d_sec = spin ? src->ddn : src->dup;
    const int32_t *st = spin ? src->st_dn : src->st_up;
    int hits = 0;
    for (int i = 0; i < d_sec; ++i) {
        uint32_t o;
        int sg = 1;
        int ok = pm > 0 ? orc_create((uint32_t)st[i], bit, ns, &o, &sg)
                        : orc_destroy((uint32_t)st[i], bit, ns, &o, &sg);
        if (!ok) continue;
        ++hits;
        int j = dst->rank[o] - 1;
        if (spin == 0) {
            for (int ido = 0; ido < src->ddn; ++ido)
                phi[j + (int64_t)ido * dst->dup] += sg * psi[i + (int64_t)ido * src->dup];
        } else {
            for (int iup = 0; iup < src->dup; ++iup)
                phi[iup + (int64_t)j * dst->dup] += sg * psi[iup + (int64_t)i * src->dup];
        }
    }
    return hits;
}

/* ------------------------------------------------------ compute_dynamic */

/* compute_dynamic_cpu (green_class_compute_dynamic.f90:191-407), fermionic, real
 * build: niter = min(nitergreen, dim) Lanczos steps from phi/||phi||, then
 * dyn(w_n) = ||phi||^2 * CF(i w_n) of T' = ph_sign (T - E).  iph = 1 particle,
 * 2 hole.  wn[nw] are the Matsubara frequencies (imaginary parts).  Returns the
 * number of Lanczos steps, 0 when ||phi||^2 < cutoff_dynamic (dyn = 0), negative
 * on a reference STOP.  alpha/beta (length >= niter) receive the raw T. */
int orc_dynamic(const orc_sector *S, int iph, double energy, const double *phi, int nitergreen,
                double cutoff_dynamic, int nw, const double *wn, double *dyn_re, double *dyn_im,
                double *alpha, double *beta)
{
    const int64_t D = S->dim;
    int niter = nitergreen < D ? nitergreen : (int)D;
    double n2 = 0.0;
    for (int64_t i = 0; i < D; ++i) n2 += phi[i] * phi[i];
    for (int k = 0; k < nw; ++k) { dyn_re[k] = 0.0; dyn_im[k] = 0.0; }
    if (n2 < cutoff_dynamic) return 0;
    double *start = (double *)malloc(sizeof(double) * D);
    double inv = 1.0 / sqrt(n2);
    for (int64_t i = 0; i < D; ++i) start[i] = phi[i] * inv;
    int st = orc_lanczos_run(S, niter, start, alpha, beta);
    free(start);
    if (st < 0) return st;
    double ph = iph == 1 ? 1.0 : -1.0;
    double *a2 = (double *)malloc(sizeof(double) * niter);
    for (int i = 0; i < niter; ++i) a2[i] = ph * (alpha[i] - energy);
    for (int k = 0; k < nw; ++k) {
        double re, im;
        if (orc_cfrac(niter, a2, beta, 0.0, wn[k], &re, &im) < 0) { free(a2); return -4; }
        /* dyn = -fermion_sign * CF * ||phi||^2, fermion_sign = -1 (:224-230, :324) */
        dyn_re[k] = re * n2;
        dyn_im[k] = im * n2;
    }
    free(a2);
    return niter;
}

int orc_num_procs(void)
{
#ifdef _OPENMP
    return omp_get_num_procs();
#else
    return 1;
#endif
}
