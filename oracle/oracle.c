/* CPU oracle for the AnsatzEvolve statevector path -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates, in plain C and operator by operator, what the reference computes.
 * The reference groups mutually commuting operators into fused blocks and
 * permutes them (src/fusedevolve.cpp:1797-1967); that only reorders commuting
 * rotations, so the unitary, the adjoint brackets and the Hessian it defines
 * are the ones restated here in plain operator order.  Loaded only by tests,
 * the smoke check and the benchmark's CPU-baseline leg; never by the product.
 *
 * Parity pins (tests/test_oracle.py): the reference's own test.F90 energies
 * and gradient norms; golden vectors written by the reference's CPU build
 * (oracle/_ref); exact diagonalisation; finite differences of E and g;
 * unitarity; E(0) = <HF|H|HF> in closed form.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct or_engine {
    int nq, mode, nel, nup, ndown;
    long N;
    uint32_t *states; /* ascending full-space bit strings of the sector */
    int nops;
    int8_t *ops;      /* nops x 4 */
    int *order;
    double *scale;
    int nparams;
    double *start;    /* N */
    double *start_im; /* N, complex mode (NULL: real start) */
    /* Hamiltonian: explicit rows (y[j] = sum val*x[col]) or an operator list */
    int hkind;        /* 0 none, 1 rows, 2 operator list */
    long *rowptr;
    int32_t *col;
    double *val;
    long nhops;
    uint32_t *hcre, *hdes, *hmask;
    double *hval;
};

static int popc(uint32_t x) { return __builtin_popcount(x); }

/* ---- sector (src/sparsematrix.h:65-128, src/operatorpool.h:49-77,
 *              src/operatorpool.cpp:207-242) ---- */
static int in_sector(uint32_t s, int nq, int mode, int nel, int nup, int ndown)
{
    if (mode == OR_SECTOR_FULL) return 1;
    if (mode == OR_SECTOR_NUMBER) return popc(s) == nel;
    uint32_t lo = (1u << (nq / 2)) - 1u;
    return popc(s >> (nq / 2)) == nup && popc(s & lo) == ndown;
}

long or_sector_size(int nq, int mode, int nel, int nup, int ndown)
{
    return or_sector_states(nq, mode, nel, nup, ndown, NULL);
}

long or_sector_states(int nq, int mode, int nel, int nup, int ndown, uint32_t *out)
{
    if (nq < 1 || nq > 30) return -1;
    long cnt = 0;
    uint32_t full = 1u << nq;
    for (uint32_t s = 0; s < full; s++)
        if (in_sector(s, nq, mode, nel, nup, ndown)) {
            if (out) out[cnt] = s;
            cnt++;
        }
    return cnt;
}

static long rank_of(const or_engine *e, uint32_t s)
{
    long lo = 0, hi = e->N - 1;
    while (lo <= hi) {
        long mid = (lo + hi) >> 1;
        uint32_t v = e->states[mid];
        if (v == s) return mid;
        if (v < s) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/* ---- applyExcToBasisState_ (src/fusedevolve.cpp:27-145), real mode ---- */
int or_apply_exc(uint32_t state, const int8_t a[4], uint32_t *out)
{
    int phase = 1;
    uint32_t create, annihilate;
    *out = state;
    if (a[0] < 0 && a[1] < 0) return 1;
    if (a[2] > -1 && a[3] > -1) {
        if (a[0] == a[1] || a[2] == a[3]) return 1; /* reference: "Wrong order", identity */
        create = (1u << a[0]) | (1u << a[1]);
        annihilate = (1u << a[2]) | (1u << a[3]);
        uint32_t mask = ((1u << a[0]) - 1) ^ ((1u << a[1]) - 1) ^ ((1u << a[2]) - 1) ^ ((1u << a[3]) - 1);
        mask &= ~((1u << a[0]) | (1u << a[1]) | (1u << a[2]) | (1u << a[3]));
        if (popc(state & mask) & 1) phase = -phase;
        if (a[0] > a[1]) phase = -phase;
        if (a[2] > a[3]) phase = -phase;
    } else {
        create = 1u << a[0];
        annihilate = 1u << a[1];
        uint32_t mask = ((1u << a[0]) - 1) ^ ((1u << a[1]) - 1);
        mask &= ~((1u << a[0]) | (1u << a[1]));
        if (popc(state & mask) & 1) phase = -phase;
    }
    uint32_t active = create | annihilate;
    uint32_t m = state & active;
    if (create == annihilate) return 0; /* number operator: complex builds only */
    if (((m & annihilate) ^ annihilate) == 0 && ((m ^ annihilate) & create) == 0) {
        *out = (state ^ annihilate) ^ create;
        return phase;
    }
    if (((m & create) ^ create) == 0 && ((m ^ create) & annihilate) == 0) {
        *out = (state ^ create) ^ annihilate;
        return -phase;
    }
    *out = state;
    return -1; /* inactive: the reference returns (state, false) */
}

/* ---- engine set-up: symmetry detection as in src/AnsatzManager.cpp:271-374,
 *      202-237 and the compressor choice at :88-100 ---- */
or_engine *or_engine_create(int nq, int nstart, const int *start_idx, const double *start_coeff,
                            int nops, const int *ops, const int *order, const double *scale)
{
    if (nq < 1 || nq > 30 || nstart < 1 || nops < 1) return NULL;
    int nel = -1, up = -1, down = -1, same = 1, sz = (nq % 2 == 0);
    for (int i = 0; i < nstart; i++) {
        if (start_coeff[i] == 0.0) continue;
        uint32_t s = (uint32_t)start_idx[i];
        int c = popc(s);
        if (nel == -1) nel = c;
        if (c != nel) { same = 0; break; }
        int cu = popc(s >> (nq / 2));
        int cd = popc(s & ((1u << (nq / 2)) - 1));
        if (up == -1) up = cu;
        if (down == -1) down = cd;
        if (cu != up || cd != down) sz = 0;
    }
    int opsz = 1;
    uint32_t dmask = (1u << (nq / 2)) - 1, umask = ((1u << nq) - 1) ^ dmask;
    for (int k = 0; k < nops; k++) {
        const int *o = ops + 4 * k;
        uint32_t cr, de;
        if (o[0] >= 0 && o[1] >= 0 && o[2] < 0 && o[3] < 0) { cr = 1u << o[0]; de = 1u << o[1]; }
        else if (o[0] >= 0 && o[1] >= 0 && o[2] >= 0 && o[3] >= 0) {
            cr = (1u << o[0]) | (1u << o[1]);
            de = (1u << o[2]) | (1u << o[3]);
        } else return NULL; /* neither single nor double: the reference asserts */
        opsz = opsz && popc(cr & umask) == popc(de & umask) && popc(cr & dmask) == popc(de & dmask);
    }
    int mode = OR_SECTOR_FULL;
    if (same && nel != -1) mode = (sz && up != -1 && down != -1 && opsz) ? OR_SECTOR_SZ : OR_SECTOR_NUMBER;

    or_engine *e = calloc(1, sizeof(*e));
    e->nq = nq; e->mode = mode; e->nel = nel; e->nup = up; e->ndown = down;
    e->N = or_sector_size(nq, mode, nel, up, down);
    e->states = malloc(sizeof(uint32_t) * e->N);
    or_sector_states(nq, mode, nel, up, down, e->states);
    e->nops = nops;
    e->ops = malloc(4 * nops);
    e->order = malloc(sizeof(int) * nops);
    e->scale = malloc(sizeof(double) * nops);
    e->nparams = 0;
    for (int k = 0; k < nops; k++) {
        for (int j = 0; j < 4; j++) e->ops[4 * k + j] = (int8_t)ops[4 * k + j];
        e->order[k] = order[k];
        e->scale[k] = scale ? scale[k] : 1.0;
        if (order[k] + 1 > e->nparams) e->nparams = order[k] + 1;
    }
    /* compressVector: out-of-sector amplitudes of the start are dropped */
    e->start = calloc(e->N, sizeof(double));
    for (int i = 0; i < nstart; i++) {
        long r = rank_of(e, (uint32_t)start_idx[i]);
        if (r >= 0) e->start[r] = start_coeff[i];
    }
    return e;
}

void or_engine_free(or_engine *e)
{
    if (!e) return;
    free(e->start_im);
    free(e->states); free(e->ops); free(e->order); free(e->scale); free(e->start);
    free(e->rowptr); free(e->col); free(e->val);
    free(e->hcre); free(e->hdes); free(e->hmask); free(e->hval);
    free(e);
}

int or_engine_mode(const or_engine *e) { return e->mode; }
long or_engine_dim(const or_engine *e) { return e->N; }
int or_engine_nparams(const or_engine *e) { return e->nparams; }
const uint32_t *or_engine_states(const or_engine *e) { return e->states; }
long or_engine_nnz(const or_engine *e) { return e->hkind == 1 ? e->rowptr[e->N] : 0; }

/* ---- explicit rows from (output, input, value) entries, duplicates summed ---- */
typedef struct { long row; long col; double v; } trip_t;

static int trip_cmp(const void *a, const void *b)
{
    const trip_t *x = a, *y = b;
    if (x->row != y->row) return x->row < y->row ? -1 : 1;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return 0;
}

static void rows_from_trips(or_engine *e, trip_t *t, long n)
{
    qsort(t, n, sizeof(trip_t), trip_cmp);
    free(e->rowptr); free(e->col); free(e->val);
    e->rowptr = calloc(e->N + 1, sizeof(long));
    e->col = malloc(sizeof(int32_t) * (n ? n : 1));
    e->val = malloc(sizeof(double) * (n ? n : 1));
    long w = 0;
    for (long i = 0; i < n; i++) {
        if (w > 0 && i > 0 && t[i].row == t[i - 1].row && t[i].col == t[i - 1].col) {
            e->val[w - 1] += t[i].v;
            continue;
        }
        e->col[w] = (int32_t)t[i].col;
        e->val[w] = t[i].v;
        e->rowptr[t[i].row + 1]++;
        w++;
    }
    for (long r = 0; r < e->N; r++) e->rowptr[r + 1] += e->rowptr[r];
    e->hkind = 1;
}

/* Triplet constructor (src/hamiltonianmatrix.cpp:628-659): M(ci,cj) += v for
 * in-sector i, j; applied as y = M^T x, so output row cj collects (ci, v). */
int or_engine_set_triplets(or_engine *e, long ntrip, const int *i0, const int *j0, const double *v)
{
    trip_t *t = malloc(sizeof(trip_t) * (ntrip ? ntrip : 1));
    long n = 0;
    for (long k = 0; k < ntrip; k++) {
        long ci = rank_of(e, (uint32_t)i0[k]);
        long cj = rank_of(e, (uint32_t)j0[k]);
        if (ci < 0 || cj < 0) continue;
        t[n].row = cj; t[n].col = ci; t[n].v = v[k];
        n++;
    }
    rows_from_trips(e, t, n);
    free(t);
    return 0;
}

/* Integral operator list (src/hamiltonianmatrix.cpp:402-554). */
int or_engine_set_integrals(or_engine *e, int n, const double *h, const double *g, int explicit_matrix)
{
    int nq = e->nq;
    if (nq != 2 * n) return 1;
    long cap = (long)nq * nq * nq * nq + (long)nq * nq;
    uint32_t *cr = malloc(4 * cap), *de = malloc(4 * cap), *mk = malloc(4 * cap);
    double *vv = malloc(8 * cap);
    long cnt = 0;
    const double tol = 1e-13;
#define GIDX(i, j, k, l) ((((size_t)(i) * n + (l)) * n + (j)) * n + (k))
    for (int a = 0; a < nq; a++)
        for (int b = a + 1; b < nq; b++)
            for (int c = 0; c < nq; c++)
                for (int d = c + 1; d < nq; d++) {
                    int sa = a >= n, sb = b >= n, sc = c >= n, sd = d >= n;
                    int pa = a % n, pb = b % n, pc = c % n, pd = d % n;
                    double E = 0;
                    if (sa == sc && sb == sd) E += g[GIDX(pb, pa, pc, pd)];
                    if (sa == sd && sb == sc) E -= g[GIDX(pb, pa, pd, pc)];
                    if (fabs(E) > tol) {
                        cr[cnt] = (1u << a) | (1u << b);
                        de[cnt] = (1u << c) | (1u << d);
                        uint32_t m = ((1u << a) - 1) ^ ((1u << b) - 1) ^ ((1u << c) - 1) ^ ((1u << d) - 1);
                        mk[cnt] = m & ~((1u << a) | (1u << b) | (1u << c) | (1u << d));
                        vv[cnt] = E;
                        cnt++;
                    }
                }
    for (int a = 0; a < nq; a++)
        for (int c = 0; c < nq; c++) {
            double E = h[(size_t)(a % n) * n + (c % n)];
            if (fabs(E) > tol) {
                cr[cnt] = 1u << a;
                de[cnt] = 1u << c;
                mk[cnt] = (((1u << a) - 1) ^ ((1u << c) - 1)) & ~((1u << a) | (1u << c));
                vv[cnt] = E;
                cnt++;
            }
        }
#undef GIDX
    /* postProcessOperators (:603-626) for the Sz compressor (opDoesSomething,
     * src/operatorpool.cpp:244-253); other sectors keep every operator. */
    if (e->mode == OR_SECTOR_SZ) {
        uint32_t dm = (1u << n) - 1, um = ((1u << nq) - 1) ^ dm;
        long w = 0;
        for (long k = 0; k < cnt; k++) {
            if (popc(cr[k] & um) == popc(de[k] & um) && popc(cr[k] & dm) == popc(de[k] & dm)) {
                cr[w] = cr[k]; de[w] = de[k]; mk[w] = mk[k]; vv[w] = vv[k]; w++;
            }
        }
        cnt = w;
    }
    free(e->hcre); free(e->hdes); free(e->hmask); free(e->hval);
    e->hcre = cr; e->hdes = de; e->hmask = mk; e->hval = vv; e->nhops = cnt;
    e->hkind = 2;
    if (explicit_matrix) {
        /* constructFromSecQuant (:556-601), enumerated per output row j with
         * the matrix-free rule (:915-1009); both define the same matrix. */
        long *rowcnt = calloc(e->N + 1, sizeof(long));
#pragma omp parallel for schedule(dynamic, 64)
        for (long j = 0; j < e->N; j++) {
            uint32_t jb = e->states[j];
            long c = 0;
            for (long k = 0; k < cnt; k++) {
                uint32_t dst = jb ^ cr[k];
                if ((dst & cr[k]) != 0 || (dst & de[k]) != 0) continue;
                if (rank_of(e, dst | de[k]) < 0) continue;
                c++;
            }
            rowcnt[j + 1] = c;
        }
        for (long j = 0; j < e->N; j++) rowcnt[j + 1] += rowcnt[j];
        long tot = rowcnt[e->N];
        trip_t *t = malloc(sizeof(trip_t) * (tot ? tot : 1));
#pragma omp parallel for schedule(dynamic, 64)
        for (long j = 0; j < e->N; j++) {
            uint32_t jb = e->states[j];
            long w = rowcnt[j];
            for (long k = 0; k < cnt; k++) {
                uint32_t dst = jb ^ cr[k];
                if ((dst & cr[k]) != 0 || (dst & de[k]) != 0) continue;
                long ci = rank_of(e, dst | de[k]);
                if (ci < 0) continue;
                t[w].row = j; t[w].col = ci;
                t[w].v = (popc(jb & mk[k]) & 1) ? -vv[k] : vv[k];
                w++;
            }
        }
        free(rowcnt);
        rows_from_trips(e, t, tot);
        free(t);
    }
    return 0;
}

int or_ham_apply(const or_engine *e, const double *x, double *y)
{
    if (e->hkind == 1) {
#pragma omp parallel for schedule(static)
        for (long j = 0; j < e->N; j++) {
            double acc = 0;
            for (long p = e->rowptr[j]; p < e->rowptr[j + 1]; p++) acc += x[e->col[p]] * e->val[p];
            y[j] = acc;
        }
        return 0;
    }
    if (e->hkind == 2) {
#pragma omp parallel for schedule(dynamic, 64)
        for (long j = 0; j < e->N; j++) {
            uint32_t jb = e->states[j];
            double acc = 0;
            for (long k = 0; k < e->nhops; k++) {
                uint32_t dst = jb ^ e->hcre[k];
                if ((dst & e->hcre[k]) != 0 || (dst & e->hdes[k]) != 0) continue;
                long ci = rank_of(e, dst | e->hdes[k]);
                if (ci < 0) continue;
                double v = (popc(jb & e->hmask[k]) & 1) ? -e->hval[k] : e->hval[k];
                acc += x[ci] * v;
            }
            y[j] = acc;
        }
        return 0;
    }
    return 1;
}

/* ---- rotations (src/fusedevolve.cpp:213-491) ----
 * One operator applied to nvec vectors: pair (b, b') with b' = T b > b and
 * sign s = applyExc(b).  x_b <- C x_b + s S x_b',  x_b' <- -s S x_b + C x_b'.
 * If phi/lam are given, first accumulate s*(lam_b phi_b' - lam_b' phi_b)
 * from the pre-rotation values (the braket rule at :224-233, :474-483). */
static int rotate_op(const or_engine *e, int k, double theta, double **vecs, int nvec,
                     const double *phi, const double *lam, double *bracket)
{
    const int8_t *a = e->ops + 4 * k;
    double S = sin(theta), C = cos(theta);
    double acc = 0;
    int bad = 0;
    for (long i = 0; i < e->N; i++) {
        uint32_t b = e->states[i], b2;
        int s = or_apply_exc(b, a, &b2);
        if (s == 0) { bad = 1; break; }
        if (b2 == b || b2 < b) continue;
        long j = rank_of(e, b2);
        if (j < 0) continue;
        if (phi) acc += s * (lam[i] * phi[j] - lam[j] * phi[i]);
        for (int v = 0; v < nvec; v++) {
            double x = vecs[v][i], y = vecs[v][j];
            vecs[v][i] = s * y * S + x * C;
            vecs[v][j] = -s * x * S + y * C;
        }
    }
    if (bracket) *bracket = acc;
    return bad;
}

/* G_k x into t (storeTangents, src/fusedevolve.cpp:492-518): t_b = s x_b',
 * t_b' = -s x_b on the pairs of operator k, zero elsewhere. */
static void generator_apply(const or_engine *e, int k, const double *x, double *t)
{
    const int8_t *a = e->ops + 4 * k;
    memset(t, 0, sizeof(double) * e->N);
    for (long i = 0; i < e->N; i++) {
        uint32_t b = e->states[i], b2;
        int s = or_apply_exc(b, a, &b2);
        if (s == 0 || b2 == b || b2 < b) continue;
        long j = rank_of(e, b2);
        if (j < 0) continue;
        t[i] = s * x[j];
        t[j] = -s * x[i];
    }
}

int or_decompress_angles(const or_engine *e, const double *theta, double *angles_ops)
{
    for (int k = 0; k < e->nops; k++) angles_ops[k] = e->scale[k] * theta[e->order[k]];
    return 0;
}

/* FusedEvolve::evolve (src/fusedevolve.cpp:2105-2159) */
int or_evolve(const or_engine *e, const double *angles, double *psi)
{
    memcpy(psi, e->start, sizeof(double) * e->N);
    double *v[1] = {psi};
    for (int k = 0; k < e->nops; k++)
        if (rotate_op(e, k, angles[k], v, 1, NULL, NULL, NULL)) return 2;
    return 0;
}

static double dotp(const double *a, const double *b, long n)
{
    double s = 0;
    for (long i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* FusedEvolve::getEnergy (src/fusedevolve.cpp:2782-2795) */
int or_energy(const or_engine *e, const double *angles, double *E)
{
    if (!e->hkind) return 1;
    double *psi = malloc(sizeof(double) * e->N), *hp = malloc(sizeof(double) * e->N);
    int rc = or_evolve(e, angles, psi);
    if (!rc) rc = or_ham_apply(e, psi, hp);
    if (!rc) *E = dotp(hp, psi, e->N);
    free(psi); free(hp);
    return rc;
}

/* FusedEvolve::evolveDerivative (src/fusedevolve.cpp:2263-2346): reverse
 * sweep with negated angles over (psi, H psi), brackets x2, then C = D^T. */
int or_energy_gradient(const or_engine *e, const double *angles, double *E, double *grad_comp)
{
    if (!e->hkind) return 1;
    long N = e->N;
    double *phi = malloc(sizeof(double) * N), *lam = malloc(sizeof(double) * N);
    double *deriv = calloc(e->nops, sizeof(double));
    int rc = or_evolve(e, angles, phi);
    if (!rc) rc = or_ham_apply(e, phi, lam);
    if (!rc) {
        *E = dotp(lam, phi, N);
        double *v[2] = {phi, lam};
        for (int k = e->nops - 1; k >= 0 && !rc; k--)
            rc = rotate_op(e, k, -angles[k], v, 2, phi, lam, &deriv[k]);
    }
    if (!rc) {
        for (int p = 0; p < e->nparams; p++) grad_comp[p] = 0;
        for (int k = 0; k < e->nops; k++) grad_comp[e->order[k]] += e->scale[k] * 2 * deriv[k];
    }
    free(phi); free(lam); free(deriv);
    return rc;
}

/* FusedEvolve::evolveHessian (src/fusedevolve.cpp:2437-2780). */
int or_hessian(const or_engine *e, const double *angles, double *hess, double *grad_comp, double *E)
{
    if (!e->hkind) return 1;
    long N = e->N;
    int M = e->nops, P = e->nparams;
    double *Ts = calloc((size_t)M * N, sizeof(double));
    double *psi = malloc(sizeof(double) * N);
    memcpy(psi, e->start, sizeof(double) * N);
    int rc = 0;
    /* 1. forward sweep storing G_k psi before operator k */
    for (int k = 0; k < M && !rc; k++) {
        generator_apply(e, k, psi, Ts + (size_t)k * N);
        double *v[1] = {psi};
        rc = rotate_op(e, k, angles[k], v, 1, NULL, NULL, NULL);
    }
    /* 2. propagate each tangent through operators k..M-1 */
    for (int k = 0; k < M && !rc; k++) {
        double *v[1] = {Ts + (size_t)k * N};
        for (int m = k; m < M && !rc; m++) rc = rotate_op(e, m, angles[m], v, 1, NULL, NULL, NULL);
    }
    if (rc) { free(Ts); free(psi); return rc; }
    /* 3. T_C = C_psi Ts (P rows + psi) */
    double *TC = calloc((size_t)(P + 1) * N, sizeof(double));
    double *HTC = malloc(sizeof(double) * (size_t)(P + 1) * N);
    for (int k = 0; k < M; k++) {
        double *dst = TC + (size_t)e->order[k] * N, *src = Ts + (size_t)k * N;
        for (long i = 0; i < N; i++) dst[i] += e->scale[k] * src[i];
    }
    memcpy(TC + (size_t)P * N, psi, sizeof(double) * N);
    /* 4. HT_C = H T_C */
    for (int r = 0; r <= P; r++) or_ham_apply(e, TC + (size_t)r * N, HTC + (size_t)r * N);
    /* 5. temp = 2 T_C HT_C^T : THT (upper triangle mirrored), gradient, E */
    double *THT = malloc(sizeof(double) * (size_t)P * P);
    for (int i = 0; i <= P; i++)
        for (int j = i; j <= P; j++) {
            double t = 2 * dotp(TC + (size_t)i * N, HTC + (size_t)j * N, N);
            if (i < P && j < P) { THT[(size_t)i * P + j] = t; THT[(size_t)j * P + i] = t; }
            else if (i < P && j == P) grad_comp[i] = t;
            else if (i == P && j == P) *E = t / 2;
        }
    /* 6. per-angle reverse sweeps of (T_a, H psi) with brackets for k >= a */
    double *Hunc = calloc((size_t)M * M, sizeof(double));
    const double *hpsi = HTC + (size_t)P * N;
    double *X = malloc(sizeof(double) * N), *lam = malloc(sizeof(double) * N);
    for (int a = 0; a < M && !rc; a++) {
        memcpy(X, Ts + (size_t)a * N, sizeof(double) * N);
        memcpy(lam, hpsi, sizeof(double) * N);
        double *v[2] = {X, lam};
        for (int k = M - 1; k >= a && !rc; k--) {
            double br = 0;
            rc = rotate_op(e, k, -angles[k], v, 2, X, lam, &br);
            Hunc[(size_t)a * M + k] += 2 * br;
            if (k != a) Hunc[(size_t)k * M + a] += 2 * br;
        }
    }
    /* 7. Hessian = C Hunc C^T + THT, written column-major */
    if (!rc) {
        double *Hc = calloc((size_t)P * P, sizeof(double));
        for (int a = 0; a < M; a++)
            for (int k = 0; k < M; k++)
                Hc[(size_t)e->order[a] * P + e->order[k]] += e->scale[a] * e->scale[k] * Hunc[(size_t)a * M + k];
        for (int i = 0; i < P; i++)
            for (int j = 0; j < P; j++) hess[(size_t)P * j + i] = Hc[(size_t)i * P + j] + THT[(size_t)i * P + j];
        free(Hc);
    }
    free(Ts); free(psi); free(TC); free(HTC); free(THT); free(Hunc); free(X); free(lam);
    return rc;
}

/* FusedEvolve::evolveDerivativeProj (src/fusedevolve.cpp:2814-2889): the
 * reverse sweep of (psi, proj) with negated angles; brackets without the
 * factor 2; *ov = proj . psi. */
int or_overlap_gradient(const or_engine *e, const double *angles, const double *proj, double *ov, double *grad_comp)
{
    long N = e->N;
    double *phi = malloc(sizeof(double) * N), *lam = malloc(sizeof(double) * N);
    double *deriv = calloc(e->nops, sizeof(double));
    int rc = or_evolve(e, angles, phi);
    if (!rc) {
        memcpy(lam, proj, sizeof(double) * N);
        *ov = dotp(lam, phi, N);
        double *v[2] = {phi, lam};
        for (int k = e->nops - 1; k >= 0 && !rc; k--)
            rc = rotate_op(e, k, -angles[k], v, 2, phi, lam, &deriv[k]);
    }
    if (!rc) {
        for (int p = 0; p < e->nparams; p++) grad_comp[p] = 0;
        for (int k = 0; k < e->nops; k++) grad_comp[e->order[k]] += e->scale[k] * deriv[k];
    }
    free(phi); free(lam); free(deriv);
    return rc;
}

/* FusedEvolve::evolveHessianProj (src/fusedevolve.cpp:2892-3183): the
 * tangents of evolveHessian; the gradient T_C . proj and the value
 * psi . proj (:3023-3037); per-angle reverse sweeps of (T_a, proj) with
 * brackets for k >= a, factor 1 (:3120-3125); Hessian = C Hunc C^T, no
 * <dpsi|H|dpsi> term (:3135).  Column-major output. */
int or_hessian_proj(const or_engine *e, const double *angles, const double *proj, double *hess, double *grad_comp,
                    double *ov)
{
    long N = e->N;
    int M = e->nops, P = e->nparams;
    double *Ts = calloc((size_t)M * N, sizeof(double));
    double *psi = malloc(sizeof(double) * N);
    memcpy(psi, e->start, sizeof(double) * N);
    int rc = 0;
    for (int k = 0; k < M && !rc; k++) {
        generator_apply(e, k, psi, Ts + (size_t)k * N);
        double *v[1] = {psi};
        rc = rotate_op(e, k, angles[k], v, 1, NULL, NULL, NULL);
    }
    for (int k = 0; k < M && !rc; k++) {
        double *v[1] = {Ts + (size_t)k * N};
        for (int m = k; m < M && !rc; m++) rc = rotate_op(e, m, angles[m], v, 1, NULL, NULL, NULL);
    }
    if (rc) { free(Ts); free(psi); return rc; }
    for (int p = 0; p < P; p++) grad_comp[p] = 0;
    for (int k = 0; k < M; k++) grad_comp[e->order[k]] += e->scale[k] * dotp(proj, Ts + (size_t)k * N, N);
    *ov = dotp(proj, psi, N);
    double *Hunc = calloc((size_t)M * M, sizeof(double));
    double *X = malloc(sizeof(double) * N), *lam = malloc(sizeof(double) * N);
    for (int a = 0; a < M && !rc; a++) {
        memcpy(X, Ts + (size_t)a * N, sizeof(double) * N);
        memcpy(lam, proj, sizeof(double) * N);
        double *v[2] = {X, lam};
        for (int k = M - 1; k >= a && !rc; k--) {
            double br = 0;
            rc = rotate_op(e, k, -angles[k], v, 2, X, lam, &br);
            Hunc[(size_t)a * M + k] += br;
            if (k != a) Hunc[(size_t)k * M + a] += br;
        }
    }
    if (!rc) {
        double *Hc = calloc((size_t)P * P, sizeof(double));
        for (int a = 0; a < M; a++)
            for (int k = 0; k < M; k++)
                Hc[(size_t)e->order[a] * P + e->order[k]] += e->scale[a] * e->scale[k] * Hunc[(size_t)a * M + k];
        for (int i = 0; i < P; i++)
            for (int j = 0; j < P; j++) hess[(size_t)P * j + i] = Hc[(size_t)i * P + j];
        free(Hc);
    }
    free(Ts); free(psi); free(Hunc); free(X); free(lam);
    return rc;
}

/* ---- complex mode (the reference's useComplex build) ----
 * Amplitudes interleaved: entry 2i is Re psi_i, 2i + 1 is Im psi_i, i the
 * sector index; a doubled basis state is (bits << 1) | c.  The pair rule
 * (src/fusedevolve.cpp:27-145 with isComplex): excitations map
 * (b, c) -> (b', c) with the real rule's phase; a number operator
 * (creation set = annihilation set) acts only where every annihilated
 * orbital is occupied and the phase (-1)^popc(b & mask) * flips equals +i:
 * (b, 0) -> (b, 1) and (b, 1) -> (b, 0); any other phase leaves the state
 * alone.  Number operators are evolved by their own routine
 * (RunFuseNDiagonal, src/fusedevolve.cpp:627-717) as
 * Re' = C Re - S Im, Im' = S Re + C Im (z -> e^{+i theta} z), with the
 * bracket Im(lambda) Re(phi) - Re(lambda) Im(phi): in this file's rotation
 * rule that is the pair sign -1 from the real entry, +1 from the imaginary
 * one. */
int or_apply_exc_complex(uint32_t state2, const int8_t a[4], uint32_t *out)
{
    *out = state2;
    if (a[0] < 0 && a[1] < 0) return 1;
    const uint32_t c = state2 & 1u, b = state2 >> 1;
    int phase = 1;
    uint32_t create, annihilate;
    if (a[2] > -1 && a[3] > -1) {
        if (a[0] == a[1] || a[2] == a[3]) return 1;
        create = (1u << a[0]) | (1u << a[1]);
        annihilate = (1u << a[2]) | (1u << a[3]);
        uint32_t mask = ((1u << a[0]) - 1) ^ ((1u << a[1]) - 1) ^ ((1u << a[2]) - 1) ^ ((1u << a[3]) - 1);
        mask &= ~((1u << a[0]) | (1u << a[1]) | (1u << a[2]) | (1u << a[3]));
        if (popc(b & mask) & 1) phase = -phase;
        if (a[0] > a[1]) phase = -phase;
        if (a[2] > a[3]) phase = -phase;
    } else {
        create = 1u << a[0];
        annihilate = 1u << a[1];
        uint32_t mask = ((1u << a[0]) - 1) ^ ((1u << a[1]) - 1);
        mask &= ~((1u << a[0]) | (1u << a[1]));
        if (popc(b & mask) & 1) phase = -phase;
    }
    const uint32_t m = b & (create | annihilate);
    if (create == annihilate) {
        if ((m & annihilate) != annihilate || phase != 1) return -1; /* phase 0 or -i: no partner */
        *out = (b << 1) | (c ^ 1u);
        return c ? 1 : -1;
    }
    if (((m & annihilate) ^ annihilate) == 0 && ((m ^ annihilate) & create) == 0) {
        *out = (((b ^ annihilate) ^ create) << 1) | c;
        return phase;
    }
    if (((m & create) ^ create) == 0 && ((m ^ create) & annihilate) == 0) {
        *out = (((b ^ create) ^ annihilate) << 1) | c;
        return -phase;
    }
    return -1;
}

/* one operator on interleaved vectors (2N entries), optional brackets as in
 * rotate_op */
static int rotate_op_complex(const or_engine *e, int k, double theta, double **vecs, int nvec, const double *phi,
                             const double *lam, double *bracket)
{
    const int8_t *a = e->ops + 4 * k;
    double S = sin(theta), C = cos(theta), acc = 0;
    for (long i = 0; i < 2 * e->N; i++) {
        const uint32_t s2 = (e->states[i >> 1] << 1) | (uint32_t)(i & 1);
        uint32_t t2;
        const int s = or_apply_exc_complex(s2, a, &t2);
        if (t2 == s2 || t2 < s2) continue;
        const long r = rank_of(e, t2 >> 1);
        if (r < 0) continue;
        const long j = 2 * r + (long)(t2 & 1);
        if (phi) acc += s * (lam[i] * phi[j] - lam[j] * phi[i]);
        for (int v = 0; v < nvec; v++) {
            const double x = vecs[v][i], y = vecs[v][j];
            vecs[v][i] = s * y * S + x * C;
            vecs[v][j] = -s * x * S + y * C;
        }
    }
    if (bracket) *bracket = acc;
    return 0;
}

int or_engine_set_start_complex(or_engine *e, int nstart, const int *idx, const double *re, const double *im)
{
    free(e->start_im);
    e->start_im = calloc(e->N, sizeof(double));
    memset(e->start, 0, sizeof(double) * e->N);
    for (int i = 0; i < nstart; i++) {
        const long r = rank_of(e, (uint32_t)idx[i]);
        if (r >= 0) {
            e->start[r] = re[i];
            e->start_im[r] = im[i];
        }
    }
    return 0;
}

/* psi interleaved (2N) */
int or_evolve_complex(const or_engine *e, const double *angles, double *psi)
{
    for (long i = 0; i < e->N; i++) {
        psi[2 * i] = e->start[i];
        psi[2 * i + 1] = e->start_im ? e->start_im[i] : 0.0;
    }
    double *v[1] = {psi};
    for (int k = 0; k < e->nops; k++) rotate_op_complex(e, k, angles[k], v, 1, NULL, NULL, NULL);
    return 0;
}

/* H on an interleaved vector: the real and imaginary parts separately */
static int ham_apply_complex(const or_engine *e, const double *x, double *y)
{
    long N = e->N;
    double *a = malloc(sizeof(double) * N), *b = malloc(sizeof(double) * N);
    int rc = 0;
    for (int part = 0; part < 2 && !rc; part++) {
        for (long i = 0; i < N; i++) a[i] = x[2 * i + part];
        rc = or_ham_apply(e, a, b);
        for (long i = 0; i < N; i++) y[2 * i + part] = b[i];
    }
    free(a); free(b);
    return rc;
}

/* E = Re <psi|H|psi> and the gradient: the real build's evolveDerivative
 * on the interleaved vectors (src/fusedevolve.cpp:2263-2346) */
int or_energy_gradient_complex(const or_engine *e, const double *angles, double *E, double *grad_comp)
{
    if (!e->hkind) return 1;
    long N2 = 2 * e->N;
    double *phi = malloc(sizeof(double) * N2), *lam = malloc(sizeof(double) * N2);
    double *deriv = calloc(e->nops, sizeof(double));
    or_evolve_complex(e, angles, phi);
    int rc = ham_apply_complex(e, phi, lam);
    if (!rc) {
        *E = dotp(lam, phi, N2);
        double *v[2] = {phi, lam};
        for (int k = e->nops - 1; k >= 0; k--) rotate_op_complex(e, k, -angles[k], v, 2, phi, lam, &deriv[k]);
        for (int p = 0; p < e->nparams; p++) grad_comp[p] = 0;
        for (int k = 0; k < e->nops; k++) grad_comp[e->order[k]] += e->scale[k] * 2 * deriv[k];
    }
    free(phi); free(lam); free(deriv);
    return rc;
}
