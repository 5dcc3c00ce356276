/*
 * gnx_oracle.c -- CPU restatement of the Gross-Neveu fermion hot path.
 *
 * TEST INFRASTRUCTURE ONLY (correctness oracle and serial CPU baseline).
 * The product library never links this file.
 *
 * Build with gcc -O3 -ffp-contract=off (no -ffast-math): the expressions
 * below keep the reference's C99 complex arithmetic and summation order, so
 * results are bit-identical to the reference built with gcc -O3 on x86-64.
 * Reference files are cited as file:line of alefrajoe/3D-Gross-Neveu-model.
 */
#include "gnx_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* ======================================================================= */
/* geometry                                                                 */
/* ======================================================================= */

/* lex n = (x0*L + x1)*LT + x2, time (x2) fastest: coordinate.c:25-65 */
static inline int lex_of(int L, int LT, int x0, int x1, int x2)
{
    return (x0 * L + x1) * LT + x2;
}

/* f = n/2 + parity*V/2 : coordinate.c:301-313 */
static inline int f_of(int L, int LT, int half, int x0, int x1, int x2)
{
    int n = lex_of(L, LT, x0, x1, x2);
    int par = (x0 + x1 + x2) % 2;
    return n / 2 + par * half;
}

int orc_geom_init(orc_geom *g, int L, int LT)
{
    memset(g, 0, sizeof(*g));
    if (L < 4 || LT < 4 || (LT % 2) != 0 || (L % 2) != 0) return -1;
    g->L = L; g->LT = LT; g->V = L * L * LT; g->half = g->V / 2;
    int V = g->V;
    g->lex2f = (int *)malloc(sizeof(int) * V);
    g->f2lex = (int *)malloc(sizeof(int) * V);
    g->coords = (int *)malloc(sizeof(int) * V * 3);
    g->fpp = (int *)malloc(sizeof(int) * V * 3);
    g->fmm = (int *)malloc(sizeof(int) * V * 3);
    g->eta_pp = (double *)malloc(sizeof(double) * V * 3);
    g->eta_mm = (double *)malloc(sizeof(double) * V * 3);
    g->eta_pbc = (double *)malloc(sizeof(double) * V * 3);
    if (!g->lex2f || !g->f2lex || !g->coords || !g->fpp || !g->fmm || !g->eta_pp ||
        !g->eta_mm || !g->eta_pbc) { orc_geom_free(g); return -2; }

    for (int x0 = 0; x0 < L; x0++)
    for (int x1 = 0; x1 < L; x1++)
    for (int x2 = 0; x2 < LT; x2++) {
        int n = lex_of(L, LT, x0, x1, x2);
        int f = f_of(L, LT, g->half, x0, x1, x2);
        int c[3] = {x0, x1, x2};
        int ext[3] = {L, L, LT};
        g->lex2f[n] = f;
        g->f2lex[f] = n;
        g->coords[3 * n + 0] = x0; g->coords[3 * n + 1] = x1; g->coords[3 * n + 2] = x2;
        for (int d = 0; d < 3; d++) {
            int cp[3] = {x0, x1, x2}, cm[3] = {x0, x1, x2};
            cp[d] = (c[d] + 1) % ext[d];
            cm[d] = (c[d] + ext[d] - 1) % ext[d];
            g->fpp[3 * f + d] = f_of(L, LT, g->half, cp[0], cp[1], cp[2]);
            g->fmm[3 * f + d] = f_of(L, LT, g->half, cm[0], cm[1], cm[2]);
            /* eta_mu(x) = (-1)^{x_0+..+x_{mu-1}}; APBC in time is folded into
             * the forward hop from x2=LT-1 and the backward hop from x2=0
             * (coordinate.c:133-224, macro.h:100). */
            int sum = 0;
            for (int t = 0; t < d; t++) sum += c[t];
            double eta = (sum % 2) ? -1.0 : 1.0;
            g->eta_pbc[3 * f + d] = eta;
            g->eta_pp[3 * f + d] = (d == 2 && x2 == LT - 1) ? -eta : eta;
            g->eta_mm[3 * f + d] = (d == 2 && x2 == 0) ? -eta : eta;
        }
    }
    return 0;
}

void orc_geom_free(orc_geom *g)
{
    free(g->lex2f); free(g->f2lex); free(g->coords); free(g->fpp); free(g->fmm);
    free(g->eta_pp); free(g->eta_mm); free(g->eta_pbc);
    memset(g, 0, sizeof(*g));
}

/* coordinate.c:253-267: the lex index of f-site fermion, by testing every
 * lex site (no early exit, as there) */
static int lex_by_search(const orc_geom *g, int fermion)
{
    int stereo = 0;
    for (int x = 0; x < g->V; x++) {
        int x2 = x % g->LT, q = x / g->LT, x1 = q % g->L, x0 = q / g->L;
        if (f_of(g->L, g->LT, g->half, x0, x1, x2) == fermion) stereo = x;
    }
    return stereo;
}

/* coordinate.c:322-374 for every site (lattice.c:32-33): one search per table */
void orc_setup_tables_by_search(const orc_geom *g, int *fpp, int *fmm)
{
    int ext[3] = {g->L, g->L, g->LT};
    for (int f = 0; f < g->V; f++) {
        for (int dir = 0; dir < 2; dir++) {
            int n = lex_by_search(g, f);
            int c[3] = {n / g->LT / g->L, (n / g->LT) % g->L, n % g->LT};
            int *dst = dir == 0 ? fpp : fmm;
            for (int d = 0; d < 3; d++) {
                int t[3] = {c[0], c[1], c[2]};
                t[d] = dir == 0 ? (c[d] + 1) % ext[d] : (c[d] + ext[d] - 1) % ext[d];
                dst[3 * f + d] = f_of(g->L, g->LT, g->half, t[0], t[1], t[2]);
            }
        }
    }
}

/* ======================================================================= */
/* dSFMT-132049 (Saito & Matsumoto, "SIMD-oriented Fast Mersenne Twister   */
/* for double precision"), restated from the published algorithm; the     */
/* parameter set is the MEXP=132049 one the reference selects              */
/* (myrngparam.h:4, dSFMT-params132049.h:6-16).                             */
/* ======================================================================= */
#define DS_N ORC_DSFMT_N
#define DS_POS1 371
#define DS_SL1 23
#define DS_SR 12
#define DS_MSK1 UINT64_C(0x000fb9f4eff4bf77)
#define DS_MSK2 UINT64_C(0x000fffffbfefff37)
#define DS_FIX1 UINT64_C(0x4ce24c0e4e234f3b)
#define DS_FIX2 UINT64_C(0x62612409b5665c2d)
#define DS_PCV1 UINT64_C(0x181232889145d000)
#define DS_PCV2 UINT64_C(0x0000000000000001)
#define DS_LOW_MASK UINT64_C(0x000FFFFFFFFFFFFF)
#define DS_HIGH_CONST UINT64_C(0x3FF0000000000000)

static void ds_recursion(uint64_t *r, const uint64_t *a, const uint64_t *b, uint64_t *lung)
{
    uint64_t t0 = a[0], t1 = a[1], L0 = lung[0], L1 = lung[1];
    lung[0] = (t0 << DS_SL1) ^ (L1 >> 32) ^ (L1 << 32) ^ b[0];
    lung[1] = (t1 << DS_SL1) ^ (L0 >> 32) ^ (L0 << 32) ^ b[1];
    r[0] = (lung[0] >> DS_SR) ^ (lung[0] & DS_MSK1) ^ t0;
    r[1] = (lung[1] >> DS_SR) ^ (lung[1] & DS_MSK2) ^ t1;
}

static void ds_gen_all(orc_rng *r)
{
    uint64_t *s = r->st;
    uint64_t lung[2] = {s[2 * DS_N], s[2 * DS_N + 1]};
    int i;
    for (i = 0; i < DS_N - DS_POS1; i++) ds_recursion(&s[2 * i], &s[2 * i], &s[2 * (i + DS_POS1)], lung);
    for (; i < DS_N; i++) ds_recursion(&s[2 * i], &s[2 * i], &s[2 * (i + DS_POS1 - DS_N)], lung);
    s[2 * DS_N] = lung[0];
    s[2 * DS_N + 1] = lung[1];
}

void orc_rng_init(orc_rng *r, uint32_t seed)
{
    /* 32-bit Knuth-style seeding over the whole (N+1)*128-bit state */
    uint32_t w[(DS_N + 1) * 4];
    w[0] = seed;
    for (uint32_t i = 1; i < (uint32_t)((DS_N + 1) * 4); i++)
        w[i] = 1812433253UL * (w[i - 1] ^ (w[i - 1] >> 30)) + i;
    for (int i = 0; i < (DS_N + 1) * 2; i++)
        r->st[i] = (uint64_t)w[2 * i] | ((uint64_t)w[2 * i + 1] << 32);
    /* force the N*2 payload words into [1,2) */
    for (int i = 0; i < DS_N * 2; i++) r->st[i] = (r->st[i] & DS_LOW_MASK) | DS_HIGH_CONST;
    /* period certification on the lung word */
    uint64_t t0 = r->st[2 * DS_N] ^ DS_FIX1, t1 = r->st[2 * DS_N + 1] ^ DS_FIX2;
    uint64_t inner = (t0 & DS_PCV1) ^ (t1 & DS_PCV2);
    for (int i = 32; i > 0; i >>= 1) inner ^= inner >> i;
    if ((inner & 1) == 0) r->st[2 * DS_N + 1] ^= 1; /* PCV2 has its low bit set */
    r->idx = DS_N * 2;
}

/* dsfmt_genrand_open_open: (0,1) */
double orc_casuale(orc_rng *r)
{
    if (r->idx >= DS_N * 2) { ds_gen_all(r); r->idx = 0; }
    union { double d; uint64_t u; } v;
    v.u = r->st[r->idx++] | 1;
    return v.d - 1.0;
}

void orc_fill_casuale(orc_rng *r, double *out, long n)
{
    for (long i = 0; i < n; i++) out[i] = orc_casuale(r);
}

/* random.c:88-102 (polar method, returns the first variate only) */
double orc_gauss1(orc_rng *r)
{
    double v1, v2, s, ris;
    do {
        v1 = 1.0 - 2.0 * orc_casuale(r);
        v2 = 1.0 - 2.0 * orc_casuale(r);
        s = v1 * v1 + v2 * v2;
    } while (s >= 1);
    ris = v1 * sqrt(-2 * log(s) / s);
    return ris;
}

/* random.c:129-143 */
void orc_gauss2mv(orc_rng *r, double *ris1, double *ris2, double mean, double variance)
{
    double v1, v2, s;
    do {
        v1 = -1.0 + 2.0 * orc_casuale(r);
        v2 = -1.0 + 2.0 * orc_casuale(r);
        s = v1 * v1 + v2 * v2;
    } while (s >= 1);
    *ris1 = mean + sqrt(variance) * v1 * sqrt(-2 * log(s) / s);
    *ris2 = mean + sqrt(variance) * v2 * sqrt(-2 * log(s) / s);
}

void orc_draw_sigma(orc_rng *r, int V, double var, double *sigma)
{
    int half = V / 2;
    for (int x = 0; x < half; x++) orc_gauss2mv(r, &sigma[x], &sigma[x + half], 0.0, var);
}

void orc_draw_z2_complex(orc_rng *r, int V, double *v)
{
    for (int x = 0; x < V; x++) {
        v[2 * x] = (orc_casuale(r) <= 0.5) ? 1.0 : -1.0;
        v[2 * x + 1] = (orc_casuale(r) <= 0.5) ? 1.0 : -1.0;
    }
}

/* condensate.c:52-60 */
void orc_draw_z2_real(orc_rng *r, int V, double *v)
{
    cplx *c = (cplx *)v;
    for (int x = 0; x < V; x++) {
        if (orc_casuale(r) <= 0.5) c[x] = 1.0;
        else c[x] = -1.0;
    }
}

/* ======================================================================= */
/* sigma mass and Dirac operator                                           */
/* ======================================================================= */

/* sigma.c:75-94 */
void orc_sigma_mass(const orc_geom *g, const double *sigma, double *sm)
{
    const int *mm = g->fmm;
    for (int x = 0; x < g->V; x++) {
        sm[x] = sigma[x];
        sm[x] += sigma[mm[3 * x + 0]];
        sm[x] += sigma[mm[3 * x + 1]];
        sm[x] += sigma[mm[3 * x + 2]];
        sm[x] += sigma[mm[3 * mm[3 * x + 0] + 1]];
        sm[x] += sigma[mm[3 * mm[3 * x + 0] + 2]];
        sm[x] += sigma[mm[3 * mm[3 * x + 1] + 2]];
        sm[x] += sigma[mm[3 * mm[3 * mm[3 * x + 0] + 1] + 2]];
        sm[x] = sm[x] / 8.0;
    }
}

/* dirac.c:46-82 */
void orc_dirac(const orc_geom *g, const double *sm, const double *in, double *out)
{
    const cplx *v = (const cplx *)in;
    cplx *t = (cplx *)out;
    int V = g->V;
    for (int x = 0; x < V; x++) t[x] = (sm[x]) * v[x];
    for (int x = 0; x < V; x++)
    for (int d = 0; d < 3; d++) {
        t[x] += (0.5) * g->eta_pp[3 * x + d] * v[g->fpp[3 * x + d]];
        t[x] -= (0.5) * g->eta_mm[3 * x + d] * v[g->fmm[3 * x + d]];
    }
}

/* dirac.c:88-123 (scatter form: accumulation order at a site follows the
 * ascending source index) */
void orc_adj_dirac(const orc_geom *g, const double *sm, const double *in, double *out)
{
    const cplx *v = (const cplx *)in;
    cplx *t = (cplx *)out;
    int V = g->V;
    for (int x = 0; x < V; x++) t[x] = (sm[x]) * v[x];
    for (int x = 0; x < V; x++)
    for (int d = 0; d < 3; d++) {
        t[g->fpp[3 * x + d]] += (0.5) * g->eta_pp[3 * x + d] * v[x];
        t[g->fmm[3 * x + d]] -= (0.5) * g->eta_mm[3 * x + d] * v[x];
    }
}

/* dirac.c:213-247: out -= M in */
void orc_sub_dirac(const orc_geom *g, const double *sm, const double *in, double *out)
{
    const cplx *v = (const cplx *)in;
    cplx *t = (cplx *)out;
    int V = g->V;
    for (int x = 0; x < V; x++) t[x] -= (sm[x]) * v[x];
    for (int x = 0; x < V; x++)
    for (int d = 0; d < 3; d++) {
        t[x] -= (0.5) * g->eta_pp[3 * x + d] * v[g->fpp[3 * x + d]];
        t[x] += (0.5) * g->eta_mm[3 * x + d] * v[g->fmm[3 * x + d]];
    }
}

/* dirac.c:253-287: out -= M^dag in */
void orc_sub_adj_dirac(const orc_geom *g, const double *sm, const double *in, double *out)
{
    const cplx *v = (const cplx *)in;
    cplx *t = (cplx *)out;
    int V = g->V;
    for (int x = 0; x < V; x++) t[x] -= (sm[x]) * v[x];
    for (int x = 0; x < V; x++)
    for (int d = 0; d < 3; d++) {
        t[g->fpp[3 * x + d]] -= (0.5) * g->eta_pp[3 * x + d] * v[x];
        t[g->fmm[3 * x + d]] += (0.5) * g->eta_mm[3 * x + d] * v[x];
    }
}

/* ----------------------------------------------------------------------- */
/* link-coupled operator: the commented-out copy_dirac.c:46-123 (the live    */
/* operator, dirac.c, ignores the links).  U is the reference's dlink layout: */
/* lex order [V][3] complex (lattice.h:15); the mm link of f-site x is the    */
/* link of its mm neighbour, lat->mm[COOR(site[x], d)].                       */
/* ----------------------------------------------------------------------- */
static inline int lex_of_f(const orc_geom *g, int f) { return g->f2lex[f]; }

/* copy_dirac.c:46-82 */
void orc_link_dirac(const orc_geom *g, const double *sm, const double *Ud, const double *in, double *out)
{
    const cplx *v = (const cplx *)in;
    const cplx *U = (const cplx *)Ud;
    cplx *t = (cplx *)out;
    int V = g->V;
    for (int x = 0; x < V; x++) t[x] = (sm[x]) * v[x];
    for (int x = 0; x < V; x++)
    for (int d = 0; d < 3; d++) {
        int fp = g->fpp[3 * x + d], fm = g->fmm[3 * x + d];
        t[x] += (0.5) * g->eta_pp[3 * x + d] * U[3 * lex_of_f(g, x) + d] * v[fp];
        t[x] -= (0.5) * g->eta_mm[3 * x + d] * conj(U[3 * lex_of_f(g, fm) + d]) * v[fm];
    }
}

/* copy_dirac.c:88-123 (gather form, D^dag = -D) */
void orc_link_adj_dirac(const orc_geom *g, const double *sm, const double *Ud, const double *in, double *out)
{
    const cplx *v = (const cplx *)in;
    const cplx *U = (const cplx *)Ud;
    cplx *t = (cplx *)out;
    int V = g->V;
    for (int x = 0; x < V; x++) t[x] = (sm[x]) * v[x];
    for (int x = 0; x < V; x++)
    for (int d = 0; d < 3; d++) {
        int fp = g->fpp[3 * x + d], fm = g->fmm[3 * x + d];
        t[x] -= (0.5) * g->eta_pp[3 * x + d] * U[3 * lex_of_f(g, x) + d] * v[fp];
        t[x] += (0.5) * g->eta_mm[3 * x + d] * conj(U[3 * lex_of_f(g, fm) + d]) * v[fm];
    }
}

/* out -= M_U in: the terms of copy_dirac.c:46-82 with flipped signs (the CG
 * prologue, conjugategradient.c:21, on the link-coupled operator) */
void orc_link_sub_dirac(const orc_geom *g, const double *sm, const double *Ud, const double *in, double *out)
{
    const cplx *v = (const cplx *)in;
    const cplx *U = (const cplx *)Ud;
    cplx *t = (cplx *)out;
    int V = g->V;
    for (int x = 0; x < V; x++) t[x] -= (sm[x]) * v[x];
    for (int x = 0; x < V; x++)
    for (int d = 0; d < 3; d++) {
        int fp = g->fpp[3 * x + d], fm = g->fmm[3 * x + d];
        t[x] -= (0.5) * g->eta_pp[3 * x + d] * U[3 * lex_of_f(g, x) + d] * v[fp];
        t[x] += (0.5) * g->eta_mm[3 * x + d] * conj(U[3 * lex_of_f(g, fm) + d]) * v[fm];
    }
}

/* fermion.h:67 */
double orc_norm(int V, const double *in)
{
    const cplx *v = (const cplx *)in;
    double total = 0.0;
    for (int x = 0; x < V; x++) total += creal(conj(v[x]) * v[x]);
    return pow(total, 0.5);
}

/* fermion.h:62 with the complex coefficient of the reference signature */
static inline void axpy_c(int V, const cplx *v, cplx coeff, cplx *target)
{
    for (int x = 0; x < V; x++) target[x] += (coeff * v[x]);
}

/* ======================================================================= */
/* CG on A = M M^dag (conjugategradient.c)                                  */
/* ======================================================================= */
/* the operator a solve runs on: the live sigma-only M (U == NULL) or the
 * link-coupled M_U of copy_dirac.c */
typedef struct {
    const orc_geom *g;
    const double *sm;
    const double *U;
} orc_op;

static void op_M(const orc_op *o, const double *in, double *out)
{
    if (o->U) orc_link_dirac(o->g, o->sm, o->U, in, out);
    else orc_dirac(o->g, o->sm, in, out);
}
static void op_Mdag(const orc_op *o, const double *in, double *out)
{
    if (o->U) orc_link_adj_dirac(o->g, o->sm, o->U, in, out);
    else orc_adj_dirac(o->g, o->sm, in, out);
}
static void op_subM(const orc_op *o, const double *in, double *out)
{
    if (o->U) orc_link_sub_dirac(o->g, o->sm, o->U, in, out);
    else orc_sub_dirac(o->g, o->sm, in, out);
}

static bool cg_loop(const orc_op *o, cplx *X, cplx *r, cplx *p,
                    cplx *t1, cplx *t2, double tol, long max_iter, orc_cgstat *st)
{
    int V = o->g->V;
    for (long iteration = 0; iteration < max_iter; iteration++) {
        double alpha_k = 0.0, norm_r_k = 0.0, norm_p_k = 0.0;
        for (int x = 0; x < V; x++) norm_r_k += pow(cabs(r[x]), 2.0);
        op_Mdag(o, (const double *)p, (double *)t1);
        for (int x = 0; x < V; x++) norm_p_k += pow(cabs(t1[x]), 2.0);
        alpha_k = norm_r_k / norm_p_k;
        axpy_c(V, p, alpha_k, X);
        op_M(o, (const double *)t1, (double *)t2);
        axpy_c(V, t2, -alpha_k, r);
        double intermediate_norm = orc_norm(V, (const double *)r);
        if (intermediate_norm < tol) {
            if (st) { st->iterations += (iteration + 1); st->cg_call++; st->last_iters = (int)(iteration + 1); }
            return true;
        }
        double beta_k = 0.0, norm_r_kplus1 = 0.0;
        for (int x = 0; x < V; x++) norm_r_kplus1 += pow(cabs(r[x]), 2.0);
        beta_k = norm_r_kplus1 / norm_r_k;
        for (int x = 0; x < V; x++) p[x] = r[x] + beta_k * p[x];
    }
    if (st) st->last_iters = (int)max_iter;
    return false;
}

static bool cg_null_op(const orc_op *o, const double *b, double *x, double tol, long max_iter, double *work,
                       orc_cgstat *st)
{
    int V = o->g->V;
    const cplx *B = (const cplx *)b;
    cplx *X = (cplx *)x;
    cplx *r = (cplx *)work, *p = r + V, *t1 = p + V, *t2 = t1 + V;
    memcpy(r, B, sizeof(cplx) * V);
    op_Mdag(o, x, (double *)t1);
    op_subM(o, (const double *)t1, (double *)r);
    if (orc_norm(V, (const double *)r) < tol) { if (st) st->last_iters = -1; return true; }
    for (int i = 0; i < V; i++) X[i] = 0.0;
    memcpy(r, B, sizeof(cplx) * V);
    memcpy(p, B, sizeof(cplx) * V);
    return cg_loop(o, X, r, p, t1, t2, tol, max_iter, st);
}

static bool cg_imposed_op(const orc_op *o, const double *b, double *x, double tol, long max_iter, double *work,
                          orc_cgstat *st)
{
    int V = o->g->V;
    const cplx *B = (const cplx *)b;
    cplx *X = (cplx *)x;
    cplx *r = (cplx *)work, *p = r + V, *t1 = p + V, *t2 = t1 + V;
    for (int i = 0; i < V; i++) X[i] = 0.0;
    memcpy(r, B, sizeof(cplx) * V);
    if (orc_norm(V, (const double *)r) < tol) { if (st) st->last_iters = -1; return true; }
    memcpy(p, B, sizeof(cplx) * V);
    return cg_loop(o, X, r, p, t1, t2, tol, max_iter, st);
}

/* conjugategradient.c:13-107 (START_X0_NULL) */
bool orc_cg_null(const orc_geom *g, const double *sm, const double *b, double *x,
                 double tol, long max_iter, double *work, orc_cgstat *st)
{
    orc_op o = {g, sm, NULL};
    return cg_null_op(&o, b, x, tol, max_iter, work, st);
}

/* conjugategradient.c:115-182 (StartX0Imposed) */
bool orc_cg_imposed(const orc_geom *g, const double *sm, const double *b, double *x,
                    double tol, long max_iter, double *work, orc_cgstat *st)
{
    orc_op o = {g, sm, NULL};
    return cg_imposed_op(&o, b, x, tol, max_iter, work, st);
}

/* the same solvers on the link-coupled operator */
bool orc_link_cg_null(const orc_geom *g, const double *sm, const double *U, const double *b, double *x,
                      double tol, long max_iter, double *work, orc_cgstat *st)
{
    orc_op o = {g, sm, U};
    return cg_null_op(&o, b, x, tol, max_iter, work, st);
}

bool orc_link_cg_imposed(const orc_geom *g, const double *sm, const double *U, const double *b, double *x,
                         double tol, long max_iter, double *work, orc_cgstat *st)
{
    orc_op o = {g, sm, U};
    return cg_imposed_op(&o, b, x, tol, max_iter, work, st);
}

/* ======================================================================= */
/* sigma HMC                                                                */
/* ======================================================================= */
int orc_kit_alloc(orc_kit *k, int V)
{
    memset(k, 0, sizeof(*k));
    k->V = V;
    k->work = (double *)calloc((size_t)8 * V, sizeof(double));
    for (int f = 0; f < 2; f++) {
        k->chi[f] = (double *)calloc((size_t)2 * V, sizeof(double));
        k->phi[f] = (double *)calloc((size_t)2 * V, sizeof(double));
        k->inv_phi[f] = (double *)calloc((size_t)2 * V, sizeof(double));
        k->inv_phi_obs[f] = (double *)calloc((size_t)2 * V, sizeof(double));
    }
    k->ddag = (double *)calloc((size_t)2 * V, sizeof(double));
    k->vec = (double *)calloc((size_t)2 * V, sizeof(double));
    return (k->work && k->ddag && k->vec) ? 0 : -1;
}

void orc_kit_free(orc_kit *k)
{
    free(k->work);
    for (int f = 0; f < 2; f++) { free(k->chi[f]); free(k->phi[f]); free(k->inv_phi[f]); free(k->inv_phi_obs[f]); }
    free(k->ddag); free(k->vec);
    memset(k, 0, sizeof(*k));
}

/* moleculardynamics.c:606-620 then :576-581 (force starts at zero, :634) */
void orc_sigma_force_from_solutions(const orc_geom *g, const double *sigma, double beta_s,
                                    double *const Xd[2], double *const Yd[2], double *force)
{
    int V = g->V;
    const int *pp = g->fpp;
    for (int x = 0; x < V; x++) force[x] = 0.0;
    for (int f = 0; f < ORC_HALFNFLAV; f++) {
        const cplx *X = (const cplx *)Xd[f];
        const cplx *Y = (const cplx *)Yd[f];
        for (int x = 0; x < V; x++) {
            int c1 = pp[3 * pp[3 * x + 0] + 1];
            int c2 = pp[3 * pp[3 * x + 0] + 2];
            int c3 = pp[3 * pp[3 * x + 1] + 2];
            int c4 = pp[3 * x + 0];
            int c5 = pp[3 * x + 1];
            int c6 = pp[3 * x + 2];
            int c7 = pp[3 * pp[3 * pp[3 * x + 0] + 1] + 2];
            force[x] += (0.125) * creal(conj(X[x]) * Y[x]);
            force[x] += (0.125) * creal(conj(X[c1]) * Y[c1]);
            force[x] += (0.125) * creal(conj(X[c2]) * Y[c2]);
            force[x] += (0.125) * creal(conj(X[c3]) * Y[c3]);
            force[x] += (0.125) * creal(conj(X[c4]) * Y[c4]);
            force[x] += (0.125) * creal(conj(X[c5]) * Y[c5]);
            force[x] += (0.125) * creal(conj(X[c6]) * Y[c6]);
            force[x] += (0.125) * creal(conj(X[c7]) * Y[c7]);
        }
    }
    for (int x = 0; x < V; x++) force[x] -= (ORC_NFLAV * beta_s / 2.0) * sigma[x];
}

/* moleculardynamics.c:590-647 */
bool orc_sigma_force(const orc_geom *g, const double *sigma, const double *sm, double beta_s,
                     double tol, long max_iter, orc_kit *k, double *force)
{
    for (int f = 0; f < ORC_HALFNFLAV; f++)
        if (!orc_cg_null(g, sm, k->phi[f], k->inv_phi[f], tol, max_iter, k->work, &k->st)) return false;
    for (int f = 0; f < ORC_HALFNFLAV; f++) orc_adj_dirac(g, sm, k->inv_phi[f], k->inv_phi_obs[f]);
    orc_sigma_force_from_solutions(g, sigma, beta_s, k->inv_phi, k->inv_phi_obs, force);
    return true;
}

/* moleculardynamics.c:299-319 */
double orc_energy(const orc_geom *g, const double *mom, const double *sigma, double beta_s,
                  double *const phid[2], double *const Xd[2])
{
    int V = g->V;
    double total_energy = 0.0;
    for (int r = 0; r < V; r++) total_energy += (pow(mom[r], 2) / 2.0);
    for (int f = 0; f < ORC_HALFNFLAV; f++) {
        const cplx *phi = (const cplx *)phid[f];
        const cplx *X = (const cplx *)Xd[f];
        for (int x = 0; x < V; x++) total_energy += ((0.5) * creal(conj(phi[x]) * X[x]));
    }
    double total = 0.0;
    for (int x = 0; x < V; x++) total += pow(sigma[x], 2);
    total_energy += ((ORC_NFLAV * beta_s / 4.0) * total);
    return total_energy;
}

/* moleculardynamics.c:824-836 */
static void evolve_coords(const orc_geom *g, double *sigma, double *sm, const double *mom, double eps)
{
    for (int x = 0; x < g->V; x++) sigma[x] += (eps * mom[x]);
    orc_sigma_mass(g, sigma, sm);
}

/* moleculardynamics.c:843-848 */
static void evolve_momenta(int V, double *mom, const double *force, double eps)
{
    for (int x = 0; x < V; x++) mom[x] += (force[x] * eps);
}

/* moleculardynamics.c:1234-1434 */
void orc_lfmn2vf(const orc_geom *g, orc_rng *rng, double *sigma, double *sm, double *mom,
                 double *force, double beta_s, double epsilon, int step, double tol_high,
                 double tol_medium, long max_iter, orc_kit *k, orc_traj *out)
{
    int V = g->V;
    const double lam = ORC_MAGIC_LAMBDA;
    memset(out, 0, sizeof(*out));
    for (int x = 0; x < V; x++) mom[x] = orc_gauss1(rng);                       /* :781-786 */
    for (int f = 0; f < ORC_HALFNFLAV; f++) {
        cplx *chi = (cplx *)k->chi[f];
        for (int r = 0; r < V; r++) chi[r] = orc_gauss1(rng);                   /* fermion.c:337-345 */
        orc_dirac(g, sm, k->chi[f], k->phi[f]);
    }
    double *sigma_copy = (double *)malloc(sizeof(double) * V);
    memcpy(sigma_copy, sigma, sizeof(double) * V);

#define ORC_FAIL() do { memcpy(sigma, sigma_copy, sizeof(double) * V); orc_sigma_mass(g, sigma, sm); \
                        free(sigma_copy); out->ok = 0; return; } while (0)

    for (int f = 0; f < ORC_HALFNFLAV; f++)
        if (!orc_cg_imposed(g, sm, k->phi[f], k->inv_phi[f], tol_high, max_iter, k->work, &k->st)) ORC_FAIL();
    double starting_energy = orc_energy(g, mom, sigma, beta_s, k->phi, k->inv_phi);

    if (!orc_sigma_force(g, sigma, sm, beta_s, tol_medium, max_iter, k, force)) ORC_FAIL();
    evolve_momenta(V, mom, force, (lam * epsilon));
    evolve_coords(g, sigma, sm, mom, (epsilon / 2.0));
    if (!orc_sigma_force(g, sigma, sm, beta_s, tol_medium, max_iter, k, force)) ORC_FAIL();
    evolve_momenta(V, mom, force, ((1.0 - (2.0 * lam)) * epsilon));
    evolve_coords(g, sigma, sm, mom, (epsilon / 2.0));

    for (int s = 1; s < step - 1; s++) {
        if (!orc_sigma_force(g, sigma, sm, beta_s, tol_medium, max_iter, k, force)) ORC_FAIL();
        evolve_momenta(V, mom, force, (2.0 * lam * epsilon));
        evolve_coords(g, sigma, sm, mom, (epsilon / 2.0));
        if (!orc_sigma_force(g, sigma, sm, beta_s, tol_medium, max_iter, k, force)) ORC_FAIL();
        evolve_momenta(V, mom, force, ((1.0 - (2.0 * lam)) * epsilon));
        evolve_coords(g, sigma, sm, mom, (epsilon / 2.0));
    }

    if (!orc_sigma_force(g, sigma, sm, beta_s, tol_medium, max_iter, k, force)) ORC_FAIL();
    evolve_momenta(V, mom, force, (2.0 * lam * epsilon));
    evolve_coords(g, sigma, sm, mom, (epsilon / 2.0));
    if (!orc_sigma_force(g, sigma, sm, beta_s, tol_medium, max_iter, k, force)) ORC_FAIL();
    evolve_momenta(V, mom, force, ((1.0 - (2.0 * lam)) * epsilon));
    evolve_coords(g, sigma, sm, mom, (epsilon / 2.0));
    if (!orc_sigma_force(g, sigma, sm, beta_s, tol_high, max_iter, k, force)) ORC_FAIL();
    evolve_momenta(V, mom, force, (lam * epsilon));

    double final_energy = orc_energy(g, mom, sigma, beta_s, k->phi, k->inv_phi);
    out->H0 = starting_energy;
    out->H1 = final_energy;
    out->dH = final_energy - starting_energy;
    out->ok = 1;
    if (orc_casuale(rng) <= exp(-(final_energy - starting_energy))) {
        out->accepted = 1;
    } else {
        memcpy(sigma, sigma_copy, sizeof(double) * V);
        orc_sigma_mass(g, sigma, sm);
        out->accepted = 0;
    }
    free(sigma_copy);
#undef ORC_FAIL
}

/* ======================================================================= */
/* condensate estimators (condensate.c:114-244)                             */
/* ======================================================================= */
#define ORC_PI (3.141592653589793238462643383279502884197169399375105820974944)

void orc_condensate(const orc_geom *g, orc_rng *rng, const double *sm, double tol, long max_iter,
                    orc_kit *k, double *x_inout, double *out)
{
    int V = g->V;
    cplx *vec = (cplx *)k->vec;
    cplx *ddag = (cplx *)k->ddag;
    double tr1 = 0.0, tr1_abs, tr2, tr4, mu2 = 0.0, mu4, chi, chi_abs, chip, nos = 0.0, withs = 0.0;
    double temp[4] = {0.0, 0.0, 0.0, 0.0};
    for (int p = 0; p < 4; p++) {
        orc_draw_z2_real(rng, V, k->vec);
        orc_cg_null(g, sm, k->vec, x_inout, tol, max_iter, k->work, &k->st);
        orc_adj_dirac(g, sm, x_inout, k->ddag);
        for (int x = 0; x < V; x++) temp[p] += creal(conj(vec[x]) * ddag[x]) / V;
    }
    /* Mean over CONDENSATE_PRECISION=1 element (inputoutput.c:307-315) */
    double t0 = 0.0 + temp[0] / 1, t1 = 0.0 + temp[1] / 1, t2 = 0.0 + temp[2] / 1, t3 = 0.0 + temp[3] / 1;
    tr1 = (t0 + t1 + t2 + t3) / 4.0;
    tr1_abs = fabs(tr1);
    tr2 = ((t0 * t1) + (t0 * t2) + (t0 * t3) + (t1 * t2) + (t1 * t3) + (t2 * t3)) / 6.0;
    tr4 = (t0 * t1 * t2 * t3);

    for (int x = 0; x < V; x++) vec[x] = 1.0 / V;
    orc_cg_null(g, sm, k->vec, x_inout, tol, max_iter, k->work, &k->st);
    orc_adj_dirac(g, sm, x_inout, k->ddag);
    for (int x = 0; x < V; x++) mu2 += creal(conj(vec[x]) * ddag[x]);
    mu4 = pow(mu2, 2);
    chi = V * mu2;
    chi_abs = fabs(chi);

    for (int x = 0; x < V; x++)
        vec[g->lex2f[x]] = cexp(I * ((ORC_PI) / g->L) * g->coords[3 * x + 2]) / V;
    orc_cg_null(g, sm, k->vec, x_inout, tol, max_iter, k->work, &k->st);
    orc_adj_dirac(g, sm, x_inout, k->ddag);
    double temp_chip = 0.0;
    for (int x = 0; x < V; x++) temp_chip += cimag(conj(vec[x]) * ddag[x]);
    chip = (temp_chip * V);

    double *copy_z2 = (double *)malloc(sizeof(double) * V);
    orc_draw_z2_real(rng, V, k->vec);
    for (int x = 0; x < V; x++) copy_z2[x] = creal(vec[x]);
    orc_cg_null(g, sm, k->vec, x_inout, tol, max_iter, k->work, &k->st);
    orc_adj_dirac(g, sm, x_inout, k->ddag);
    orc_cg_null(g, sm, k->ddag, x_inout, tol, max_iter, k->work, &k->st);
    orc_adj_dirac(g, sm, x_inout, k->ddag);
    for (int x = 0; x < V; x++) nos += ddag[x] * copy_z2[x] / V;

    orc_draw_z2_real(rng, V, k->vec);
    for (int x = 0; x < V; x++) copy_z2[x] = creal(vec[x]);
    for (int x = 0; x < V; x++) vec[x] = vec[x] * sm[x];
    orc_cg_null(g, sm, k->vec, x_inout, tol, max_iter, k->work, &k->st);
    orc_adj_dirac(g, sm, x_inout, k->ddag);
    for (int y = 0; y < V; y++) ddag[y] = ddag[y] * sm[y];
    orc_cg_null(g, sm, k->ddag, x_inout, tol, max_iter, k->work, &k->st);
    orc_adj_dirac(g, sm, x_inout, k->ddag);
    for (int x = 0; x < V; x++) withs += ddag[x] * copy_z2[x] / V;
    free(copy_z2);

    out[0] = tr1; out[1] = tr1_abs; out[2] = tr2; out[3] = tr4; out[4] = mu2; out[5] = mu4;
    out[6] = chi; out[7] = chi_abs; out[8] = chip; out[9] = nos; out[10] = withs;
}

/* observables.c:323-369 */
void orc_condensate_sigma(const orc_geom *g, const double *sigma, double *out)
{
    int V = g->V;
    double total = 0.0;
    cplx box[3] = {0.0, 0.0, 0.0};
    for (int x = 0; x < V; x++) {
        double temp = sigma[g->lex2f[x]] / V;
        total += temp;
        for (int d = 0; d < 3; d++)
            box[d] += cexp(I * ((2 * ORC_PI) / g->L) * g->coords[3 * x + d]) * temp;
    }
    double s = total;
    double mu2 = pow(s, 2);
    out[0] = s;
    out[1] = fabs(s);
    out[2] = mu2;
    out[3] = pow(mu2, 2);
    out[4] = V * mu2;
    out[5] = (V * pow(cabs(box[2]), 2));
    out[6] = (V * pow(cabs(box[0]), 2));
    out[7] = (V * pow(cabs(box[1]), 2));
}

/* ======================================================================= */
/* link sector (moleculardynamics.c:127-222, 268-294, 381-542;             */
/* linklattice.c:16-172).  U is lex order [V][3] complex (lattice.h:15);    */
/* momenta and force are f-order [V][3] (moleculardynamics.c:389).          */
/* No live reference path reaches this code (observables.c:127,183 are      */
/* commented out), so it is pinned by invariants only.                      */
/* ======================================================================= */
static inline int lex_pp(const orc_geom *g, int lex, int d) { return g->f2lex[g->fpp[3 * g->lex2f[lex] + d]]; }
static inline int lex_mm(const orc_geom *g, int lex, int d) { return g->f2lex[g->fmm[3 * g->lex2f[lex] + d]]; }

/* linklattice.c:16-33 (LinkLink, LinkAdjLink of link.h:20-21) */
static double plaq_re(const orc_geom *g, const cplx *U, int site, int mu, int nu)
{
    cplx t1 = U[3 * site + mu] * U[3 * lex_pp(g, site, mu) + nu];
    cplx t2 = t1 * conj(U[3 * lex_pp(g, site, nu) + mu]);
    t1 = t2 * conj(U[3 * site + nu]);
    return creal(t1);
}

/* linklattice.c:62-80 */
double orc_plaquette_sum(const orc_geom *g, const double *Ud)
{
    const cplx *U = (const cplx *)Ud;
    double temp = 0.0;
    for (int x = 0; x < g->V; x++)
        for (int mu = 0; mu < 3; mu++)
            for (int nu = 0; nu < mu; nu++) temp += plaq_re(g, U, x, mu, nu);
    return temp;
}

/* linklattice.c:146-172 */
static double imag_all_plaquettes(const orc_geom *g, const cplx *U, int site, int mu)
{
    double total = 0.0;
    for (int nu = 0; nu < 3; nu++)
        if (nu != mu) {
            cplx t1 = U[3 * site + mu] * U[3 * lex_pp(g, site, mu) + nu];
            cplx t2 = t1 * conj(U[3 * lex_pp(g, site, nu) + mu]);
            t1 = t2 * conj(U[3 * site + nu]);
            total -= cimag(t1);
            int smn = lex_mm(g, site, nu);
            t1 = U[3 * smn + mu] * U[3 * lex_pp(g, smn, mu) + nu];
            t2 = t1 * conj(U[3 * site + mu]);
            t1 = t2 * conj(U[3 * smn + nu]);
            total += cimag(t1);
        }
    return total;
}

/* moleculardynamics.c:127-138 (overwrites its elements) */
void orc_gauge_force(const orc_geom *g, const double *Ud, double beta, double *force)
{
    const cplx *U = (const cplx *)Ud;
    for (int x = 0; x < g->V; x++)
        for (int d = 0; d < 3; d++) force[3 * g->lex2f[x] + d] = beta * imag_all_plaquettes(g, U, x, d);
}

/* moleculardynamics.c:181-190, after the solves (accumulates) */
void orc_link_force_from_solutions(const orc_geom *g, const double *Ud, double *const Xd[2], double *const Yd[2],
                                   double *force)
{
    const cplx *U = (const cplx *)Ud;
    for (int f = 0; f < ORC_HALFNFLAV; f++) {
        const cplx *X = (const cplx *)Xd[f];
        const cplx *Y = (const cplx *)Yd[f];
        for (int x = 0; x < g->V; x++)
            for (int mu = 0; mu < 3; mu++) {
                int site = g->f2lex[x];
                int fp = g->fpp[3 * x + mu];
                force[3 * x + mu] += creal(conj(X[x]) * I * g->eta_pp[3 * x + mu] * U[3 * site + mu] * Y[fp]);
                force[3 * x + mu] += creal(conj(X[fp]) * I * g->eta_mm[3 * fp + mu] * conj(U[3 * site + mu]) * Y[x]);
            }
    }
}

/* ComputeForce (moleculardynamics.c:203-222) with
 * AddForceFromFermionSectorRespectToLink (:164-193).  linked: the solves and
 * M^dag run on M_U (copy_dirac.c); otherwise on the live M, as the reference. */
bool orc_link_force(const orc_geom *g, const double *sm, const double *U, double beta, int pure_gauge, int linked,
                    double tol, long max_iter, orc_kit *k, double *force)
{
    orc_op o = {g, sm, linked ? U : NULL};
    for (int r = 0; r < 3 * g->V; r++) force[r] = 0.0;
    if (pure_gauge) orc_gauge_force(g, U, beta, force);
    for (int f = 0; f < ORC_HALFNFLAV; f++)
        if (!cg_null_op(&o, k->phi[f], k->inv_phi[f], tol, max_iter, k->work, &k->st)) return false;
    for (int f = 0; f < ORC_HALFNFLAV; f++) op_Mdag(&o, k->inv_phi[f], k->inv_phi_obs[f]);
    orc_link_force_from_solutions(g, U, k->inv_phi, k->inv_phi_obs, force);
    return true;
}

/* ComputeFullEnergyWithPhiWithoutSigma (moleculardynamics.c:268-294) */
double orc_link_energy(const orc_geom *g, const double *U, const double *mom, double beta, int pure_gauge,
                       double *const phid[2], double *const Xd[2])
{
    double total_energy = 0.0;
    for (int r = 0; r < 3 * g->V; r++) total_energy += pow(mom[r], 2.0) / 2.0;
    if (pure_gauge) total_energy -= (beta * orc_plaquette_sum(g, U));
    for (int f = 0; f < ORC_HALFNFLAV; f++) {
        const cplx *phi = (const cplx *)phid[f];
        const cplx *X = (const cplx *)Xd[f];
        for (int x = 0; x < g->V; x++) total_energy += creal(conj(phi[x]) * X[x]);
    }
    return total_energy;
}

/* moleculardynamics.c:381-391, Q_CHARGE = 1 (macro.h:26) */
static void link_evolve_coords(const orc_geom *g, double *Ud, const double *mom, double eps)
{
    cplx *U = (cplx *)Ud;
    for (int x = 0; x < g->V; x++)
        for (int d = 0; d < 3; d++)
            U[3 * x + d] = cexp(I * (carg(U[3 * x + d]) + 1 * eps * mom[3 * g->lex2f[x] + d]));
}

/* moleculardynamics.c:400-407 */
static void link_evolve_momenta(int n, double *mom, const double *force, double eps)
{
    for (int r = 0; r < n; r++) mom[r] += 1 * (force[r] * eps);
}

/* LeapFrog2Link (moleculardynamics.c:419-542), RNG draws in the reference
 * order: 3V/2 Gauss2MeanVariance momentum pairs (:38-46), chi_0, chi_1
 * (fermion.c:337-345), the Metropolis uniform. */
void orc_leapfrog2link(const orc_geom *g, orc_rng *rng, const double *sm, double *U, double *mom, double *force,
                       double beta, double epsilon, int step, double tol_high, long max_iter, int pure_gauge,
                       int linked, orc_kit *k, orc_traj *out)
{
    int V = g->V;
    orc_op o = {g, sm, linked ? U : NULL};
    memset(out, 0, sizeof(*out));
    for (int r = 0; r < 3 * V / 2; r++) orc_gauss2mv(rng, &mom[2 * r], &mom[2 * r + 1], 0.0, 1.0);
    for (int f = 0; f < ORC_HALFNFLAV; f++) {
        cplx *chi = (cplx *)k->chi[f];
        for (int r = 0; r < V; r++) chi[r] = orc_gauss1(rng);
        op_M(&o, k->chi[f], k->phi[f]);
    }
    double *copy = (double *)malloc(sizeof(double) * 6 * V);
    memcpy(copy, U, sizeof(double) * 6 * V);
#define ORC_LFAIL() do { memcpy(U, copy, sizeof(double) * 6 * V); free(copy); out->ok = 0; return; } while (0)
    for (int f = 0; f < ORC_HALFNFLAV; f++)
        if (!cg_null_op(&o, k->phi[f], k->inv_phi[f], tol_high, max_iter, k->work, &k->st)) ORC_LFAIL();
    double starting_energy = orc_link_energy(g, U, mom, beta, pure_gauge, k->phi, k->inv_phi);
    if (!orc_link_force(g, sm, U, beta, pure_gauge, linked, tol_high, max_iter, k, force)) ORC_LFAIL();
    link_evolve_momenta(3 * V, mom, force, epsilon / 2.0);
    for (int s = 0; s < step - 1; s++) {
        link_evolve_coords(g, U, mom, epsilon);
        if (!orc_link_force(g, sm, U, beta, pure_gauge, linked, tol_high, max_iter, k, force)) ORC_LFAIL();
        link_evolve_momenta(3 * V, mom, force, epsilon);
    }
    link_evolve_coords(g, U, mom, epsilon);
    if (!orc_link_force(g, sm, U, beta, pure_gauge, linked, tol_high, max_iter, k, force)) ORC_LFAIL();
    link_evolve_momenta(3 * V, mom, force, epsilon / 2.0);
    double final_energy = orc_link_energy(g, U, mom, beta, pure_gauge, k->phi, k->inv_phi);
    out->H0 = starting_energy;
    out->H1 = final_energy;
    out->dH = final_energy - starting_energy;
    out->ok = 1;
    if (orc_casuale(rng) <= exp(-(final_energy - starting_energy))) {
        out->accepted = 1;
    } else {
        memcpy(U, copy, sizeof(double) * 6 * V);
        out->accepted = 0;
    }
    free(copy);
#undef ORC_LFAIL
}
