/*
 * oracle.c -- CPU FP64 restatement of IMSRG++'s scalar IMSRG(2) commutator,
 * BCH and Magnus-Euler path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for the HIP product.  Only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the reference repository root).  The restatement keeps the reference's
 * factorization (per-channel matrix products, Pandya transform, inverse
 * Pandya gather) and its conventions (normalized TBMEs, ket phases, the
 * |6j|<1e-8 skips, the Hermitian folds), so that it reproduces the reference
 * to rounding.  It is pinned independently by tests/ (m-scheme brute force,
 * Fock-space commutators, exact 6j values, invariants).
 *
 * Parity status: every function here is pinned by at least one test in
 * tests/test_oracle_*.py; see the test names cited next to each function.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OCC_CUT 1e-6 /* src/ModelSpace.cc:14 */

typedef struct {
  int J, par, tz, n;
  int *kets;   /* local -> global ket index */
  int *map;    /* global ket index -> local, -1 if absent */
  int nhh, nph, npp;
  int *hh, *ph, *pp; /* local indices, ascending */
  double *occ_hh, *unocc_hh, *occ_ph, *unocc_ph;
  int ncore, nnoncore; /* generator lists: core-core and (non-core)-(non-core) kets */
  int *core, *noncore;
  long off;
} chan_t;

struct oms {
  int emax, norb, nkets, nch, jmax;
  int *n, *l, *j2, *tz2, *cvq;
  double *occ;
  int *kp, *kq;
  int *obn;
  int **obl; /* one-body channel of each orbit: orbits with equal (l, j2, tz2), ascending */
  int nholes;
  int *holes;
  int ncore;
  int *core;
  chan_t *ch, *cc;
  long tbsize;
  int d1, d2;
  double *sixj;
};

static inline int phase_of(int x) { return (x % 2 == 0) ? 1 : -1; }

/* Planted-mistake switch for the checker tests (0 = off).  1: drop the exchange
 * term of the inverse Pandya gather; 2: wrong sign of the swapped-ket phase;
 * 3: transpose of the one-body factor in comm122ss; 4: drop the (2J'+1) weight
 * in the Pandya transform; 5: two-body norm weighted by (2J+1) instead of
 * (2J+1)^2; 6: one-body norm without the (2j+1) weight; 7: Omega not cleared at
 * a split; 8: BCH_Transform of H(s) in place of H0 / H_saved; 9: BCH_Product
 * without its 1/12 [[Y,X],X] term; 10: normal-ordering one-body sum without the
 * 1/(2j_a+1) factor; 11: basis transformation without the 1/sqrt2 for new-basis
 * kets with p' == q'. */
static int g_fault = 0;
void orc_set_fault(int f) { g_fault = f; }

static double SQRT2;

/* ------------------------------------------------------------------ */
/* double-double arithmetic for the exact Racah 6j                     */
/* ------------------------------------------------------------------ */
typedef struct {
  double hi, lo;
} dd_t;

static inline dd_t dd_quick(double a, double b) {
  double s = a + b;
  dd_t r = {s, b - (s - a)};
  return r;
}
static inline dd_t dd_two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  dd_t r = {s, (a - (s - bb)) + (b - bb)};
  return r;
}
static inline dd_t dd_add(dd_t x, dd_t y) {
  dd_t s = dd_two_sum(x.hi, y.hi);
  dd_t t = dd_two_sum(x.lo, y.lo);
  s.lo += t.hi;
  s = dd_quick(s.hi, s.lo);
  s.lo += t.lo;
  return dd_quick(s.hi, s.lo);
}
static inline dd_t dd_neg(dd_t x) {
  dd_t r = {-x.hi, -x.lo};
  return r;
}
static inline dd_t dd_mul(dd_t x, dd_t y) {
  double p = x.hi * y.hi;
  double e = fma(x.hi, y.hi, -p);
  e += x.hi * y.lo + x.lo * y.hi;
  return dd_quick(p, e);
}
static inline dd_t dd_mul_d(dd_t x, double y) {
  double p = x.hi * y;
  double e = fma(x.hi, y, -p);
  e += x.lo * y;
  return dd_quick(p, e);
}
static inline dd_t dd_div(dd_t x, dd_t y) {
  double q1 = x.hi / y.hi;
  dd_t r = dd_add(x, dd_neg(dd_mul_d(y, q1)));
  double q2 = r.hi / y.hi;
  r = dd_add(r, dd_neg(dd_mul_d(y, q2)));
  double q3 = r.hi / y.hi;
  dd_t q = dd_quick(q1, q2);
  dd_t q3d = {q3, 0.0};
  return dd_add(q, q3d);
}
static inline dd_t dd_sqrt(dd_t a) {
  if (a.hi <= 0) {
    dd_t z = {0, 0};
    return z;
  }
  double x = sqrt(a.hi);
  dd_t xx = dd_mul(dd_quick(x, 0), dd_quick(x, 0));
  dd_t r = dd_add(a, dd_neg(xx));
  double corr = r.hi / (2.0 * x);
  return dd_two_sum(x, corr);
}

#define NFACT 400
static dd_t g_fact[NFACT];
static int g_fact_init = 0;
static void init_fact(void) {
  if (g_fact_init) return;
  g_fact[0].hi = 1.0;
  g_fact[0].lo = 0.0;
  for (int i = 1; i < NFACT; ++i) g_fact[i] = dd_mul_d(g_fact[i - 1], (double)i);
  SQRT2 = sqrt(2.0);
  g_fact_init = 1;
}

static inline int tri_ok(int a, int b, int c) { /* doubled values */
  if (((a + b + c) & 1) != 0) return 0;
  if (c < abs(a - b) || c > a + b) return 0;
  return 1;
}

/* Wigner 6j {ja jb jc; jd je jf} from the Racah single-sum formula (Racah 1942;
 * Edmonds, "Angular Momentum in Quantum Mechanics", eq. 6.3.7), evaluated in
 * double-double.  Arguments are doubled angular momenta.  Replaces GSL's
 * gsl_sf_coupling_6j called at src/AngMom.cc:94-97. */
double orc_sixj_exact(int a, int b, int c, int d, int e, int f) {
  init_fact();
  if (a < 0 || b < 0 || c < 0 || d < 0 || e < 0 || f < 0) return 0.0;
  if (!tri_ok(a, b, c) || !tri_ok(a, e, f) || !tri_ok(d, b, f) || !tri_ok(d, e, c)) return 0.0;
  int tri[4][3] = {{a, b, c}, {a, e, f}, {d, b, f}, {d, e, c}};
  /* product of the four triangle coefficients, each formed as its own ratio
   * (at most 1): the four denominators multiplied first overflow a double
   * for j = 29/2, J = 29 (emax = 14) */
  dd_t d2 = {1.0, 0.0};
  for (int k = 0; k < 4; ++k) {
    int x = tri[k][0], y = tri[k][1], z = tri[k][2];
    dd_t num = dd_mul(dd_mul(g_fact[(x + y - z) / 2], g_fact[(x - y + z) / 2]), g_fact[(-x + y + z) / 2]);
    d2 = dd_mul(d2, dd_div(num, g_fact[(x + y + z) / 2 + 1]));
  }
  dd_t pref = dd_sqrt(d2);
  int al[4] = {(a + b + c) / 2, (a + e + f) / 2, (d + b + f) / 2, (d + e + c) / 2};
  int be[3] = {(a + b + d + e) / 2, (b + c + e + f) / 2, (a + c + d + f) / 2};
  int tmin = al[0], tmax = be[0];
  for (int k = 1; k < 4; ++k)
    if (al[k] > tmin) tmin = al[k];
  for (int k = 1; k < 3; ++k)
    if (be[k] < tmax) tmax = be[k];
  dd_t sum = {0.0, 0.0};
  for (int t = tmin; t <= tmax; ++t) {
    dd_t dnm = g_fact[t - al[0]];
    for (int k = 1; k < 4; ++k) dnm = dd_mul(dnm, g_fact[t - al[k]]);
    for (int k = 0; k < 3; ++k) dnm = dd_mul(dnm, g_fact[be[k] - t]);
    dd_t term = dd_div(g_fact[t + 1], dnm);
    if (t & 1) term = dd_neg(term);
    sum = dd_add(sum, term);
  }
  dd_t r = dd_mul(pref, sum);
  return r.hi + r.lo;
}

/* Dense 6j table {jj1 jj2 J3; jj4 jj5 J6}: index layout of
 * src/AngMomCache.hh:15-22, dims from src/AngMomCache.cc:8-14 with
 * jj_1b_max = 2*emax+1 (src/ModelSpace.cc:176,192). */
static inline long sixj_index(const oms_t *ms, int jj1, int jj2, int J3, int jj4, int jj5, int J6) {
  long d1 = ms->d1, d2 = ms->d2;
  return (((((long)((jj1 - 1) / 2) * d1 + (jj2 - 1) / 2) * d2 + J3) * d1 + (jj4 - 1) / 2) * d1 +
          (jj5 - 1) / 2) * d2 + J6;
}

double orc_sixj_cached(const oms_t *ms, int jj1, int jj2, int J3, int jj4, int jj5, int J6) {
  return ms->sixj[sixj_index(ms, jj1, jj2, J3, jj4, jj5, J6)];
}

static void build_sixj(oms_t *ms) {
  int jjmax = 2 * ms->emax + 1;
  ms->d1 = (jjmax - 1) / 2 + 1;
  ms->d2 = jjmax + 1;
  long tot = (long)ms->d1 * ms->d1 * ms->d2 * ms->d1 * ms->d1 * ms->d2;
  ms->sixj = (double *)calloc((size_t)tot, sizeof(double));
  int d1 = ms->d1, d2 = ms->d2;
#pragma omp parallel for collapse(3) schedule(dynamic, 4)
  for (int i1 = 0; i1 < d1; ++i1)
    for (int i2 = 0; i2 < d1; ++i2)
      for (int J3 = 0; J3 < d2; ++J3) {
        int jj1 = 2 * i1 + 1, jj2 = 2 * i2 + 1;
        if (!tri_ok(jj1, jj2, 2 * J3)) continue;
        for (int i4 = 0; i4 < d1; ++i4)
          for (int i5 = 0; i5 < d1; ++i5) {
            int jj4 = 2 * i4 + 1, jj5 = 2 * i5 + 1;
            if (!tri_ok(jj4, jj5, 2 * J3)) continue;
            for (int J6 = 0; J6 < d2; ++J6) {
              double v = orc_sixj_exact(jj1, jj2, 2 * J3, jj4, jj5, 2 * J6);
              ms->sixj[sixj_index(ms, jj1, jj2, J3, jj4, jj5, J6)] = v;
            }
          }
      }
}

/* ------------------------------------------------------------------ */
/* basis (src/ModelSpace.cc:353-417, 604-645, 923-980, 1108-1201;       */
/*        src/TwoBodyChannel.cc:19-188)                                 */
/* ------------------------------------------------------------------ */
static inline int ket_index(const oms_t *ms, int p, int q) { /* Index2, p<=q */
  return p * (2 * ms->norb - 1 - p) / 2 + q;
}

static int check_std(const oms_t *ms, int J, int par, int tz, int p, int q) {
  if (p == q && (J % 2) != 0) return 0;
  if ((ms->l[p] + ms->l[q]) % 2 != par) return 0;
  if (ms->tz2[p] + ms->tz2[q] != 2 * tz) return 0;
  if (ms->j2[p] + ms->j2[q] < 2 * J) return 0;
  if (abs(ms->j2[p] - ms->j2[q]) > 2 * J) return 0;
  return 1;
}
static int check_cc(const oms_t *ms, int J, int par, int tz, int p, int q) {
  if ((ms->l[p] + ms->l[q]) % 2 != par) return 0;
  if (ms->j2[p] + ms->j2[q] < 2 * J) return 0;
  if (abs(ms->j2[p] - ms->j2[q]) > 2 * J) return 0;
  if (abs(ms->tz2[p] - ms->tz2[q]) != 2 * tz) return 0;
  return 1;
}

static void setup_chan(oms_t *ms, chan_t *c, int is_cc) {
  int nk = ms->nkets;
  c->map = (int *)malloc(sizeof(int) * nk);
  c->kets = (int *)malloc(sizeof(int) * (nk > 0 ? nk : 1));
  c->n = 0;
  for (int k = 0; k < nk; ++k) {
    int p = ms->kp[k], q = ms->kq[k];
    int ok = is_cc ? check_cc(ms, c->J, c->par, c->tz, p, q) : check_std(ms, c->J, c->par, c->tz, p, q);
    c->map[k] = -1;
    if (ok) {
      c->map[k] = c->n;
      c->kets[c->n++] = k;
    }
  }
  int n = c->n;
  c->hh = (int *)malloc(sizeof(int) * (n + 1));
  c->ph = (int *)malloc(sizeof(int) * (n + 1));
  c->pp = (int *)malloc(sizeof(int) * (n + 1));
  c->occ_hh = (double *)malloc(sizeof(double) * (n + 1));
  c->unocc_hh = (double *)malloc(sizeof(double) * (n + 1));
  c->occ_ph = (double *)malloc(sizeof(double) * (n + 1));
  c->unocc_ph = (double *)malloc(sizeof(double) * (n + 1));
  c->core = (int *)malloc(sizeof(int) * (n + 1));
  c->noncore = (int *)malloc(sizeof(int) * (n + 1));
  c->nhh = c->nph = c->npp = c->ncore = c->nnoncore = 0;
  for (int i = 0; i < n; ++i) {
    int k = c->kets[i];
    int p = ms->kp[k], q = ms->kq[k];
    double op = ms->occ[p], oq = ms->occ[q];
    if (op < OCC_CUT && oq < OCC_CUT) c->pp[c->npp++] = i;
    if ((op > OCC_CUT) != (oq > OCC_CUT)) {
      c->occ_ph[c->nph] = op * oq;
      c->unocc_ph[c->nph] = (1 - op) * (1 - oq);
      c->ph[c->nph++] = i;
    }
    if (op > OCC_CUT && oq > OCC_CUT) {
      c->occ_hh[c->nhh] = op * oq;
      c->unocc_hh[c->nhh] = (1 - op) * (1 - oq);
      c->hh[c->nhh++] = i;
    }
    int cp = ms->cvq[p], cq = ms->cvq[q];
    if (cp + cq == 0) c->core[c->ncore++] = i;
    if (cp != 0 && cq != 0) c->noncore[c->nnoncore++] = i;
  }
}

oms_t *orc_ms_new(int norb, const int *n, const int *l, const int *j2, const int *tz2, const double *occ,
                  const int *cvq, int emax) {
  init_fact();
  oms_t *ms = (oms_t *)calloc(1, sizeof(oms_t));
  ms->emax = emax;
  ms->norb = norb;
  ms->n = (int *)malloc(sizeof(int) * norb);
  ms->l = (int *)malloc(sizeof(int) * norb);
  ms->j2 = (int *)malloc(sizeof(int) * norb);
  ms->tz2 = (int *)malloc(sizeof(int) * norb);
  ms->cvq = (int *)malloc(sizeof(int) * norb);
  ms->occ = (double *)malloc(sizeof(double) * norb);
  int jmax = 0;
  for (int i = 0; i < norb; ++i) {
    ms->n[i] = n[i];
    ms->l[i] = l[i];
    ms->j2[i] = j2[i];
    ms->tz2[i] = tz2[i];
    ms->occ[i] = occ[i];
    ms->cvq[i] = cvq[i];
    if (j2[i] > jmax) jmax = j2[i];
  }
  ms->jmax = jmax; /* TwoBodyJmax = OneBodyJmax (src/ModelSpace.cc:939-945) */
  ms->nch = 6 * (jmax + 1);
  ms->nkets = norb * (norb + 1) / 2;
  ms->kp = (int *)malloc(sizeof(int) * ms->nkets);
  ms->kq = (int *)malloc(sizeof(int) * ms->nkets);
  for (int p = 0; p < norb; ++p)
    for (int q = p; q < norb; ++q) {
      int k = ket_index(ms, p, q);
      ms->kp[k] = p;
      ms->kq[k] = q;
    }
  ms->obn = (int *)calloc(norb, sizeof(int));
  ms->obl = (int **)malloc(sizeof(int *) * norb);
  for (int i = 0; i < norb; ++i) {
    ms->obl[i] = (int *)malloc(sizeof(int) * norb);
    for (int j = 0; j < norb; ++j)
      if (l[j] == l[i] && j2[j] == j2[i] && tz2[j] == tz2[i]) ms->obl[i][ms->obn[i]++] = j;
  }
  ms->holes = (int *)malloc(sizeof(int) * norb);
  ms->core = (int *)malloc(sizeof(int) * norb);
  for (int i = 0; i < norb; ++i) {
    if (occ[i] > OCC_CUT) ms->holes[ms->nholes++] = i; /* src/ModelSpace.cc:956-959 */
    if (cvq[i] == 0) ms->core[ms->ncore++] = i;
  }
  ms->ch = (chan_t *)calloc(ms->nch, sizeof(chan_t));
  ms->cc = (chan_t *)calloc(ms->nch, sizeof(chan_t));
  long off = 0;
  for (int c = 0; c < ms->nch; ++c) {
    /* src/ModelSpace.cc:1069-1081: J = ch/6, Tz = (ch%6)/2-1, parity = ch%2 */
    chan_t *s = &ms->ch[c];
    s->J = c / 6;
    s->tz = (c % 6) / 2 - 1;
    s->par = c % 2;
    setup_chan(ms, s, 0);
    s->off = off;
    off += (long)s->n * s->n;
    chan_t *x = &ms->cc[c];
    x->J = c / 6;
    x->tz = (c % 6) / 2 - 1;
    x->par = c % 2;
    setup_chan(ms, x, 1);
  }
  ms->tbsize = off;
  build_sixj(ms);
  return ms;
}

/* Single-reference HO basis: orbit order of src/ModelSpace.cc:376-399 and
 * occupations of GetOrbitsAZ, src/ModelSpace.cc:611-645.  Core = reference
 * (single-reference case, src/ModelSpace.cc:274-281). */
oms_t *orc_ms_new_ho(int emax, double A, double Z) {
  int cap = (emax + 1) * (emax + 2) * 2 + 4;
  int *n = malloc(sizeof(int) * cap), *l = malloc(sizeof(int) * cap), *j2 = malloc(sizeof(int) * cap),
      *tz = malloc(sizeof(int) * cap), *cvq = malloc(sizeof(int) * cap);
  double *occ = calloc(cap, sizeof(double));
  int norb = 0;
  for (int N = 0; N <= emax; ++N)
    for (int ll = N; ll >= 0; ll -= 2) {
      int nn = (N - ll) / 2;
      for (int jj = 2 * ll + 1; jj >= 2 * ll - 1 && jj > 0; jj -= 2)
        for (int t = -1; t <= 1; t += 2) {
          n[norb] = nn;
          l[norb] = ll;
          j2[norb] = jj;
          tz[norb] = t;
          occ[norb] = 0.0;
          cvq[norb] = 2;
          norb++;
        }
    }
  double zz = 0, nnn = 0;
  int done = 0;
  for (int N = 0; N <= emax && !done; ++N)
    for (int g = 2 * N + 1; g >= -2 * N && !done; g -= 4) {
      int jj = abs(g);
      int ll = g < 0 ? (jj + 1) / 2 : (jj - 1) / 2;
      int nn = (N - ll) / 2;
      for (int i = 0; i < norb; ++i) {
        if (n[i] != nn || l[i] != ll || j2[i] != jj) continue;
        if (tz[i] == -1 && zz < Z) {
          double dz = fmin(Z - zz, jj + 1.0);
          occ[i] = dz / (jj + 1.0);
          cvq[i] = 0;
          zz += dz;
        }
      }
      for (int i = 0; i < norb; ++i) {
        if (n[i] != nn || l[i] != ll || j2[i] != jj) continue;
        if (tz[i] == 1 && nnn < A - Z) {
          double dn = fmin(A - Z - nnn, jj + 1.0);
          occ[i] = dn / (jj + 1.0);
          cvq[i] = 0;
          nnn += dn;
        }
      }
      if (zz == Z && nnn == A - Z) done = 1;
    }
  oms_t *ms = orc_ms_new(norb, n, l, j2, tz, occ, cvq, emax);
  free(n);
  free(l);
  free(j2);
  free(tz);
  free(cvq);
  free(occ);
  return ms;
}

static void free_chan(chan_t *c) {
  free(c->map);
  free(c->kets);
  free(c->hh);
  free(c->ph);
  free(c->pp);
  free(c->occ_hh);
  free(c->unocc_hh);
  free(c->occ_ph);
  free(c->unocc_ph);
  free(c->core);
  free(c->noncore);
}

void orc_ms_free(oms_t *ms) {
  if (!ms) return;
  for (int c = 0; c < ms->nch; ++c) {
    free_chan(&ms->ch[c]);
    free_chan(&ms->cc[c]);
  }
  free(ms->ch);
  free(ms->cc);
  for (int i = 0; i < ms->norb; ++i) free(ms->obl[i]);
  free(ms->obl);
  free(ms->obn);
  free(ms->holes);
  free(ms->core);
  free(ms->n);
  free(ms->l);
  free(ms->j2);
  free(ms->tz2);
  free(ms->cvq);
  free(ms->occ);
  free(ms->kp);
  free(ms->kq);
  free(ms->sixj);
  free(ms);
}

int orc_ms_norb(const oms_t *ms) { return ms->norb; }
int orc_ms_nkets(const oms_t *ms) { return ms->nkets; }
int orc_ms_nch(const oms_t *ms) { return ms->nch; }
long orc_ms_tbsize(const oms_t *ms) { return ms->tbsize; }
void orc_ms_orbit(const oms_t *ms, int i, int *n, int *l, int *j2, int *tz2, double *occ, int *cvq) {
  *n = ms->n[i];
  *l = ms->l[i];
  *j2 = ms->j2[i];
  *tz2 = ms->tz2[i];
  *occ = ms->occ[i];
  *cvq = ms->cvq[i];
}
void orc_ms_channel(const oms_t *ms, int ch, int *J, int *par, int *tz, int *nkets, long *off) {
  const chan_t *c = &ms->ch[ch];
  *J = c->J;
  *par = c->par;
  *tz = c->tz;
  *nkets = c->n;
  *off = c->off;
}
void orc_ms_channel_ket(const oms_t *ms, int ch, int i, int *p, int *q) {
  int k = ms->ch[ch].kets[i];
  *p = ms->kp[k];
  *q = ms->kq[k];
}
void orc_ms_cc_channel(const oms_t *ms, int ch, int *J, int *par, int *tz, int *nkets, int *nph) {
  const chan_t *c = &ms->cc[ch];
  *J = c->J;
  *par = c->par;
  *tz = c->tz;
  *nkets = c->n;
  *nph = c->nhh + c->nph;
}
void orc_ms_cc_channel_ket(const oms_t *ms, int ch, int i, int *p, int *q) {
  int k = ms->cc[ch].kets[i];
  *p = ms->kp[k];
  *q = ms->kq[k];
}

/* ------------------------------------------------------------------ */
/* TBME accessors                                                        */
/* ------------------------------------------------------------------ */
/* Ket::Phase, src/ModelSpace.hh:77-80: |pq J> = (-1)^{jp+jq+1+J} |qp J> */
static inline int ket_phase(const oms_t *ms, int p, int q, int J) {
  int ph = phase_of((ms->j2[p] + ms->j2[q]) / 2 + 1 + J);
  return g_fault == 2 ? -ph : ph;
}

/* GetTBME_J_norm for a scalar operator, src/TwoBodyME.cc:333-352 + 128-147 */
static double tbme_norm_J(const oms_t *ms, const double *two, int J, int a, int b, int c, int d) {
  int pab = (ms->l[a] + ms->l[b]) % 2, pcd = (ms->l[c] + ms->l[d]) % 2;
  if (pab != pcd) return 0.0;
  int tab = (ms->tz2[a] + ms->tz2[b]) / 2, tcd = (ms->tz2[c] + ms->tz2[d]) / 2;
  if (tab != tcd) return 0.0;
  if (J < 0 || J > ms->jmax) return 0.0;
  const chan_t *ch = &ms->ch[6 * J + 2 * (tab + 1) + pab];
  int i = ch->map[ket_index(ms, a < b ? a : b, a < b ? b : a)];
  int k = ch->map[ket_index(ms, c < d ? c : d, c < d ? d : c)];
  if (i < 0 || k < 0) return 0.0;
  double ph = 1.0;
  if (a > b) ph *= ket_phase(ms, a, b, J);
  if (c > d) ph *= ket_phase(ms, c, d, J);
  return ph * two[ch->off + (long)k * ch->n + i];
}

/* GetTBME_J (tilde normalization), src/TwoBodyME.cc:122-129, 305-325 */
double orc_tbme_J(const oms_t *ms, const double *two, int J, int a, int b, int c, int d) {
  double nrm = 1.0;
  if (a == b) nrm *= SQRT2;
  if (c == d) nrm *= SQRT2;
  return nrm * tbme_norm_J(ms, two, J, a, b, c, d);
}

/* GetTBMEmonopole, src/TwoBodyME.cc:615-682 (scalar operator) */
double orc_tbme_monopole(const oms_t *ms, const double *two, int a, int b, int c, int d) {
  double nrm = 1.0;
  if (a == b) nrm *= SQRT2;
  if (c == d) nrm *= SQRT2;
  int pab = (ms->l[a] + ms->l[b]) % 2, pcd = (ms->l[c] + ms->l[d]) % 2;
  if (pab != pcd) return 0.0;
  int tab = (ms->tz2[a] + ms->tz2[b]) / 2, tcd = (ms->tz2[c] + ms->tz2[d]) / 2;
  if (tab != tcd) return 0.0;
  int jmin = abs(ms->j2[a] - ms->j2[b]) / 2, jmax = (ms->j2[a] + ms->j2[b]) / 2;
  double mon = 0.0;
  for (int J = jmin; J <= jmax; ++J) mon += (2 * J + 1) * tbme_norm_J(ms, two, J, a, b, c, d);
  mon /= (ms->j2[a] + 1) * (ms->j2[b] + 1);
  return nrm * mon;
}

#define ONE(op, p, q) ((op)->one[(long)(q) * norb + (p)])

/* ------------------------------------------------------------------ */
/* commutator terms (src/Commutator.cc)                                 */
/* ------------------------------------------------------------------ */

/* comm110ss, src/Commutator.cc:564-581 */
static void comm110ss(const oms_t *ms, const orc_op *X, const orc_op *Y, orc_op *Z) {
  if (X->herm == 1 && Y->herm == 1) return;
  if (X->herm == -1 && Y->herm == -1) return;
  int norb = ms->norb;
  for (int ih = 0; ih < ms->nholes; ++ih) {
    int a = ms->holes[ih];
    double xy = 0, yx = 0;
    for (int k = 0; k < norb; ++k) {
      xy += ONE(X, a, k) * ONE(Y, k, a);
      yx += ONE(Y, a, k) * ONE(X, k, a);
    }
    *Z->zero += (ms->j2[a] + 1) * ms->occ[a] * (xy - yx);
  }
}

/* comm220ss, src/Commutator.cc:601-631 */
static void comm220ss(const oms_t *ms, const orc_op *X, const orc_op *Y, orc_op *Z) {
  if (X->herm == 1 && Y->herm == 1) return;
  if (X->herm == -1 && Y->herm == -1) return;
  double z0 = 0.0;
  for (int c = 0; c < ms->nch; ++c) {
    const chan_t *ch = &ms->ch[c];
    if (ch->nhh == 0) continue;
    int n = ch->n;
    const double *X2 = X->two + ch->off, *Y2 = Y->two + ch->off;
    double acc = 0.0;
    for (int ih = 0; ih < ch->nhh; ++ih) {
      int h = ch->hh[ih];
      double d = 0.0;
      /* nbar_indices = [hh, ph, pp]; rows of hh and ph weighted by nbar nbar */
      for (int m = 0; m < ch->nhh; ++m) d += X2[(long)ch->hh[m] * n + h] * (Y2[(long)h * n + ch->hh[m]] * ch->unocc_hh[m]);
      for (int m = 0; m < ch->nph; ++m) d += X2[(long)ch->ph[m] * n + h] * (Y2[(long)h * n + ch->ph[m]] * ch->unocc_ph[m]);
      for (int m = 0; m < ch->npp; ++m) d += X2[(long)ch->pp[m] * n + h] * Y2[(long)h * n + ch->pp[m]];
      acc += d * ch->occ_hh[ih];
    }
    z0 += 2 * (2 * ch->J + 1) * acc;
  }
  *Z->zero += z0;
}

/* comm111ss, src/Commutator.cc:644-649 */
static void comm111ss(const oms_t *ms, const orc_op *X, const orc_op *Y, orc_op *Z) {
  int norb = ms->norb;
#pragma omp parallel for
  for (int j = 0; j < norb; ++j)
    for (int i = 0; i < norb; ++i) {
      double xy = 0, yx = 0;
      for (int k = 0; k < norb; ++k) {
        xy += ONE(X, i, k) * ONE(Y, k, j);
        yx += ONE(Y, i, k) * ONE(X, k, j);
      }
      ONE(Z, i, j) += xy - yx;
    }
}

/* comm121ss, src/Commutator.cc:670-728 */
static void comm121ss(const oms_t *ms, const orc_op *X, const orc_op *Y, orc_op *Z) {
  int norb = ms->norb;
  int hZ = (Z->herm == 1) ? 1 : -1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < norb; ++i) {
    int jmin = (Z->herm == 0) ? 0 : i;
    for (int jj = 0; jj < ms->obn[i]; ++jj) {
      int j = ms->obl[i][jj];
      if (j < jmin) continue;
      double zij = 0.0;
      for (int ia = 0; ia < ms->nholes; ++ia) {
        int a = ms->holes[ia];
        for (int ib = 0; ib < ms->obn[a]; ++ib) {
          int b = ms->obl[a][ib];
          double nanb = ms->occ[a] * (1 - ms->occ[b]);
          if (fabs(nanb) < OCC_CUT) continue;
          double ybiaj = orc_tbme_monopole(ms, Y->two, b, i, a, j);
          double yaibj = orc_tbme_monopole(ms, Y->two, a, i, b, j);
          zij += (ms->j2[b] + 1) * nanb * ONE(X, a, b) * ybiaj;
          zij -= (ms->j2[a] + 1) * nanb * ONE(X, b, a) * yaibj;
        }
        for (int ib = 0; ib < ms->obn[a]; ++ib) {
          int b = ms->obl[a][ib];
          double nanb = ms->occ[a] * (1 - ms->occ[b]);
          if (fabs(nanb) < OCC_CUT) continue;
          zij -= (ms->j2[b] + 1) * nanb * ONE(Y, a, b) * orc_tbme_monopole(ms, X->two, b, i, a, j);
          zij += (ms->j2[a] + 1) * nanb * ONE(Y, b, a) * orc_tbme_monopole(ms, X->two, a, i, b, j);
        }
      }
      ONE(Z, i, j) += zij;
      if (jmin == i && j != i) ONE(Z, j, i) += hZ * zij;
    }
  }
}

/* comm122ss fast path, src/Commutator.cc:837-932 */
static void comm122ss(const oms_t *ms, const orc_op *X, const orc_op *Y, orc_op *Z) {
  int norb = ms->norb;
  int hZ = (Z->herm == 1) ? 1 : -1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < ms->nch; ++c) {
    const chan_t *ch = &ms->ch[c];
    int n = ch->n;
    if (n == 0) continue;
    const double *X2 = X->two + ch->off, *Y2 = Y->two + ch->off;
    double *Z2 = Z->two + ch->off;
    double *W = (double *)calloc((size_t)n * n, sizeof(double));
    int *ind = (int *)malloc(sizeof(int) * 2 * norb);
    double *cx = (double *)malloc(sizeof(double) * 2 * norb);
    double *cy = (double *)malloc(sizeof(double) * 2 * norb);
    for (int ij = 0; ij < n; ++ij) {
      int k = ch->kets[ij];
      int i = ms->kp[k], j = ms->kq[k];
      int flip = -phase_of((ms->j2[i] + ms->j2[j]) / 2 - ch->J);
      int m = 0;
      for (int ia = 0; ia < ms->obn[i]; ++ia) {
        int a = ms->obl[i][ia];
        int loc = ch->map[ket_index(ms, a < j ? a : j, a < j ? j : a)];
        if (loc < 0) continue;
        double f = (a > j) ? (double)flip : (a == j ? SQRT2 : 1.0);
        if (i == j) f /= SQRT2;
        ind[m] = loc;
        cx[m] = (g_fault == 3 ? ONE(X, i, a) : ONE(X, a, i)) * f;
        cy[m] = (g_fault == 3 ? ONE(Y, i, a) : ONE(Y, a, i)) * f;
        m++;
      }
      if (i != j) {
        for (int ia = 0; ia < ms->obn[j]; ++ia) {
          int a = ms->obl[j][ia];
          int loc = ch->map[ket_index(ms, a < i ? a : i, a < i ? i : a)];
          if (loc < 0) continue;
          double f = (i > a) ? (double)flip : (i == a ? SQRT2 : 1.0);
          ind[m] = loc;
          cx[m] = ONE(X, a, j) * f;
          cy[m] = ONE(Y, a, j) * f;
          m++;
        }
      }
      double *w = W + (long)ij * n;
      for (int t = 0; t < m; ++t) {
        const double *yc = Y2 + (long)ind[t] * n;
        for (int r = 0; r < n; ++r) w[r] += yc[r] * cx[t];
      }
      for (int t = 0; t < m; ++t) {
        const double *xc = X2 + (long)ind[t] * n;
        for (int r = 0; r < n; ++r) w[r] += xc[r] * (-cy[t]);
      }
      if (i == j)
        for (int r = 0; r < n; ++r) w[r] *= 2;
    }
    for (int s = 0; s < n; ++s)
      for (int r = 0; r < n; ++r) Z2[(long)s * n + r] -= W[(long)s * n + r] + hZ * W[(long)r * n + s];
    free(W);
    free(ind);
    free(cx);
    free(cy);
  }
}

/* C(:, :) += A(:, K) diag(w) B(K, :) for n x n column-major A, B, C */
/* Optional BLAS for the channel GEMMs (the reference runs them through
 * Armadillo -> dgemm, src/Commutator.cc:1176-1183,1789).  The host sets a
 * Fortran-convention dgemm (orc_set_dgemm); without one the loops below run. */
typedef void (*dgemm_fn)(const char *ta, const char *tb, const int *m, const int *n, const int *k,
                         const double *alpha, const double *a, const int *lda, const double *b, const int *ldb,
                         const double *beta, double *c, const int *ldc);
static dgemm_fn g_dgemm = NULL;
void orc_set_dgemm(void *fn) { g_dgemm = (dgemm_fn)fn; }
int orc_has_dgemm(void) { return g_dgemm != NULL; }

/* C(n x n) += sign * A(n x n)[:, K] diag(w) B[K, :] */
static void gemm_sub(int n, const double *A, const double *B, double *C, const int *K, const double *w, int nk,
                     double sign) {
  if (g_dgemm && nk > 0) {
    double *Ag = (double *)malloc(sizeof(double) * (size_t)n * nk);
    double *Bg = (double *)malloc(sizeof(double) * (size_t)n * nk);
    for (int t = 0; t < nk; ++t) {
      const double wt = w ? w[t] : 1.0;
      const double *ak = A + (long)K[t] * n;
      for (int r = 0; r < n; ++r) Ag[(long)t * n + r] = ak[r] * wt;
    }
    for (int s = 0; s < n; ++s)
      for (int t = 0; t < nk; ++t) Bg[(long)s * nk + t] = B[(long)s * n + K[t]];
    const double one = 1.0;
    g_dgemm("N", "N", &n, &n, &nk, &sign, Ag, &n, Bg, &nk, &one, C, &n);
    free(Ag);
    free(Bg);
    return;
  }
  for (int s = 0; s < n; ++s) {
    double *cs = C + (long)s * n;
    for (int t = 0; t < nk; ++t) {
      int k = K[t];
      double coef = B[(long)s * n + k] * (w ? w[t] : 1.0) * sign;
      const double *ak = A + (long)k * n;
      for (int r = 0; r < n; ++r) cs[r] += ak[r] * coef;
    }
  }
}

static int all_zero(const double *a, long len) {
  for (long i = 0; i < len; ++i)
    if (a[i] != 0.0) return 0;
  return 1;
}

/* ConstructScalarMpp_Mhh, src/Commutator.cc:1082-1210 (scalar, parity 0 case) */
static void construct_mpp_mhh(const oms_t *ms, const orc_op *X, const orc_op *Y, int zherm, double *Mpp,
                              double *Mhh) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < ms->nch; ++c) {
    const chan_t *ch = &ms->ch[c];
    int n = ch->n;
    if (n == 0) continue;
    long nn = (long)n * n;
    const double *X2 = X->two + ch->off, *Y2 = Y->two + ch->off;
    if (all_zero(X2, nn) || all_zero(Y2, nn)) continue;
    double *P = Mpp + ch->off, *H = Mhh + ch->off;
    /* each product is formed separately, then added (arma: M += A*B) */
    double *T = (double *)calloc((size_t)nn, sizeof(double));
#define ADD_PROD(dst, AA, BB, KK, WW, NK, SG)         \
  do {                                                \
    if ((NK) > 0) {                                   \
      memset(T, 0, sizeof(double) * nn);              \
      gemm_sub(n, AA, BB, T, KK, WW, NK, 1.0);        \
      for (long e = 0; e < nn; ++e) dst[e] += (SG) * T[e]; \
    }                                                 \
  } while (0)
    ADD_PROD(P, X2, Y2, ch->pp, NULL, ch->npp, 1.0);
    ADD_PROD(H, X2, Y2, ch->hh, ch->occ_hh, ch->nhh, 1.0);
    ADD_PROD(P, X2, Y2, ch->hh, ch->unocc_hh, ch->nhh, 1.0);
    ADD_PROD(P, X2, Y2, ch->ph, ch->unocc_ph, ch->nph, 1.0);
    if (zherm == 1 || zherm == -1) {
      double sg = (double)zherm;
      for (int s = 0; s < n; ++s)
        for (int r = 0; r < n; ++r) T[(long)s * n + r] = P[(long)r * n + s];
      for (long e = 0; e < nn; ++e) P[e] += sg * T[e];
      for (int s = 0; s < n; ++s)
        for (int r = 0; r < n; ++r) T[(long)s * n + r] = H[(long)r * n + s];
      for (long e = 0; e < nn; ++e) H[e] += sg * T[e];
    } else {
      ADD_PROD(P, Y2, X2, ch->pp, NULL, ch->npp, -1.0);
      ADD_PROD(H, Y2, X2, ch->hh, ch->occ_hh, ch->nhh, -1.0);
      ADD_PROD(P, Y2, X2, ch->hh, ch->unocc_hh, ch->nhh, -1.0);
      ADD_PROD(P, Y2, X2, ch->ph, ch->unocc_ph, ch->nph, -1.0);
    }
#undef ADD_PROD
    free(T);
  }
}

/* one-body part of comm222_pp_hh_221ss, src/Commutator.cc:1246-1294 (= comm221ss 760-819) */
static void comm221_onebody(const oms_t *ms, const double *Mpp, const double *Mhh, orc_op *Z) {
  int norb = ms->norb;
  int hZ = (Z->herm == 1) ? 1 : -1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < norb; ++i) {
    int jmin = (Z->herm == 0) ? 0 : i;
    for (int jj = 0; jj < ms->obn[i]; ++jj) {
      int j = ms->obl[i][jj];
      if (j < jmin) continue;
      double zij = 0.0;
      for (int c = 0; c < norb; ++c) {
        double nc = ms->occ[c], nbarc = 1.0 - nc;
        int Jmin = abs(ms->j2[c] - ms->j2[i]);
        int t2 = abs(ms->j2[c] - ms->j2[j]);
        if (t2 > Jmin) Jmin = t2;
        Jmin /= 2;
        int Jmax = (ms->j2[c] + (ms->j2[i] < ms->j2[j] ? ms->j2[i] : ms->j2[j])) / 2;
        if (fabs(nc) > 1e-9)
          for (int J = Jmin; J <= Jmax; ++J) zij += (2 * J + 1) * nc * orc_tbme_J(ms, Mpp, J, c, i, c, j);
        if (fabs(nbarc) > 1e-9)
          for (int J = Jmin; J <= Jmax; ++J) zij += (2 * J + 1) * nbarc * orc_tbme_J(ms, Mhh, J, c, i, c, j);
      }
      ONE(Z, i, j) += zij / (ms->j2[i] + 1.0);
      if (jmin == i && i != j) ONE(Z, j, i) += hZ * zij / (ms->j2[i] + 1.0);
    }
  }
}

/* comm222_pp_hh_221ss, comm222_pp_hhss, comm221ss: src/Commutator.cc:1062-1076, 1215-1301, 760-819 */
static void comm222_pphh_221(const oms_t *ms, const orc_op *X, const orc_op *Y, orc_op *Z, int do2b, int do1b) {
  long ts = ms->tbsize;
  double *Mpp = (double *)calloc((size_t)(ts > 0 ? ts : 1), sizeof(double));
  double *Mhh = (double *)calloc((size_t)(ts > 0 ? ts : 1), sizeof(double));
  construct_mpp_mhh(ms, X, Y, Z->herm, Mpp, Mhh);
  if (do2b) {
    for (long e = 0; e < ts; ++e) Z->two[e] += Mpp[e];
    for (long e = 0; e < ts; ++e) Z->two[e] -= Mhh[e];
  }
  if (do1b) comm221_onebody(ms, Mpp, Mhh, Z);
  free(Mpp);
  free(Mhh);
}

/* DoPandyaTransformation_SingleChannel_XandY, src/Commutator.cc:1392-1485 */
static void pandya_xy(const oms_t *ms, const orc_op *X, const orc_op *Y, int chcc, double *Xt, double *Yb) {
  const chan_t *cc = &ms->cc[chcc];
  int nk = cc->n, nph = cc->nhh + cc->nph;
  int hY = (Y->herm == 1) ? 1 : -1;
  int Jcc = cc->J;
  for (int ibra = 0; ibra < nph; ++ibra) {
    int loc = ibra < cc->nhh ? cc->hh[ibra] : cc->ph[ibra - cc->nhh];
    int kb = cc->kets[loc];
    int pq[2] = {ms->kp[kb], ms->kq[kb]};
    for (int abc = 0; abc <= 1; ++abc) {
      int a = pq[abc], b = pq[1 - abc];
      int shift = abc * nph;
      double nanb = ms->occ[a] - ms->occ[b];
      for (int ik = 0; ik < nk; ++ik) {
        int kk = cc->kets[ik];
        int c = ms->kp[kk], d = ms->kq[kk];
        double xbar = 0, ybar = 0, normf = 1.0;
        if (ms->tz2[a] + ms->tz2[d] - ms->tz2[b] - ms->tz2[c] == 0) {
          if (a == d) normf *= SQRT2;
          if (c == b) normf *= SQRT2;
          int jmin = abs(ms->j2[a] - ms->j2[d]);
          int t = abs(ms->j2[c] - ms->j2[b]);
          if (t > jmin) jmin = t;
          jmin /= 2;
          int jmax = ms->j2[a] + ms->j2[d];
          t = ms->j2[c] + ms->j2[b];
          if (t < jmax) jmax = t;
          jmax /= 2;
          int dJ = 1;
          if (a == d || c == b) {
            dJ = 2;
            jmin += jmin % 2;
          }
          for (int J = jmin; J <= jmax; J += dJ) {
            double sj = orc_sixj_cached(ms, ms->j2[a], ms->j2[b], Jcc, ms->j2[c], ms->j2[d], J);
            if (fabs(sj) < 1e-8) continue;
            /* GetTBME_J_norm_twoOps(c,b,a,d): same element of X and Y */
            double xcbad = tbme_norm_J(ms, X->two, J, c, b, a, d);
            double ycbad = tbme_norm_J(ms, Y->two, J, c, b, a, d);
            double wj = (g_fault == 4) ? 1.0 : (double)(2 * J + 1);
            xbar -= wj * sj * xcbad;
            ybar -= wj * sj * ycbad * hY;
          }
        }
        Xt[(long)(ibra + shift) * nk + ik] = xbar * normf * nanb;
        Yb[(long)ik * (2 * nph) + (ibra + shift)] = ybar * normf;
      }
    }
  }
}

/* comm222_phss, src/Commutator.cc:1681-1823 and AddInversePandyaTransformation 1501-1616 */
static double g_timer[8];
static int g_ncalls;

static void comm222_phss(const oms_t *ms, const orc_op *X, const orc_op *Y, orc_op *Z) {
  double t_ph = omp_get_wtime();
  int nch = ms->nch;
  int hy = (Y->herm == 1) ? 1 : -1;
  double **Zbar = (double **)calloc(nch, sizeof(double *));
  for (int c = 0; c < nch; ++c) {
    int nk = ms->cc[c].n;
    Zbar[c] = (double *)calloc((size_t)2 * nk * nk + 1, sizeof(double));
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < nch; ++c) {
    const chan_t *cc = &ms->cc[c];
    int nk = cc->n, nph = cc->nhh + cc->nph;
    if (nk == 0 || nph == 0) continue;
    int K = 2 * nph;
    double *Xt = (double *)calloc((size_t)nk * K, sizeof(double));  /* nk x K */
    double *Yb = (double *)calloc((size_t)K * nk, sizeof(double));  /* K x nk */
    double *Yf = (double *)calloc((size_t)K * nk, sizeof(double));  /* K x nk */
    pandya_xy(ms, X, Y, c, Xt, Yb);
    int *phi = (int *)malloc(sizeof(int) * nk);
    for (int s = 0; s < nk; ++s) {
      int k = cc->kets[s];
      phi[s] = phase_of((ms->j2[ms->kp[k]] + ms->j2[ms->kq[k]]) / 2);
    }
    for (int s = 0; s < nk; ++s)
      for (int m = 0; m < nph; ++m) {
        int loc = m < cc->nhh ? cc->hh[m] : cc->ph[m - cc->nhh];
        double pm = (double)(phi[loc] * phi[s] * hy);
        Yf[(long)s * K + m] = Yb[(long)s * K + nph + m] * pm;
        Yf[(long)s * K + nph + m] = Yb[(long)s * K + m] * pm;
      }
    double *Zb = Zbar[c]; /* nk x 2nk column-major */
    if (g_dgemm) {
      const double one = 1.0, zero = 0.0;
      g_dgemm("N", "N", &nk, &nk, &K, &one, Xt, &nk, Yb, &K, &zero, Zb, &nk);
      g_dgemm("N", "N", &nk, &nk, &K, &one, Xt, &nk, Yf, &K, &zero, Zb + (long)nk * nk, &nk);
    } else
    for (int s = 0; s < nk; ++s) {
      double *zh = Zb + (long)s * nk, *zt = Zb + (long)(nk + s) * nk;
      for (int m = 0; m < K; ++m) {
        double bh = Yb[(long)s * K + m], bt = Yf[(long)s * K + m];
        const double *xm = Xt + (long)m * nk;
        for (int r = 0; r < nk; ++r) {
          zh[r] += xm[r] * bh;
          zt[r] += xm[r] * bt;
        }
      }
    }
    /* Hermitian folds: head +-= head^T, tail += tail^T % PhaseMatZ */
    double sg = (Z->herm == 1) ? 1.0 : -1.0;
    for (int r = 0; r < nk; ++r)
      for (int s = r; s < nk; ++s) {
        double hrs = Zb[(long)s * nk + r], hsr = Zb[(long)r * nk + s];
        if (r == s) {
          Zb[(long)s * nk + r] = hrs + sg * hrs;
        } else {
          Zb[(long)s * nk + r] = hrs + sg * hsr;
          Zb[(long)r * nk + s] = hsr + sg * hrs;
        }
        double trs = Zb[(long)(nk + s) * nk + r], tsr = Zb[(long)(nk + r) * nk + s];
        double pz = (double)(phi[r] * phi[s]);
        if (r == s) {
          Zb[(long)(nk + s) * nk + r] = trs + trs * pz;
        } else {
          Zb[(long)(nk + s) * nk + r] = trs + tsr * pz;
          Zb[(long)(nk + r) * nk + s] = tsr + trs * pz;
        }
      }
    free(Xt);
    free(Yb);
    free(Yf);
    free(phi);
  }

  g_timer[6] += omp_get_wtime() - t_ph;
  t_ph = omp_get_wtime();
  /* inverse Pandya */
  int hZ = (Z->herm == 1) ? 1 : -1;
  long npairs = 0;
  for (int c = 0; c < nch; ++c) npairs += ms->ch[c].n;
  int *pch = (int *)malloc(sizeof(int) * (npairs + 1)), *pbr = (int *)malloc(sizeof(int) * (npairs + 1));
  long t = 0;
  for (int c = 0; c < nch; ++c)
    for (int b = 0; b < ms->ch[c].n; ++b) {
      pch[t] = c;
      pbr[t] = b;
      t++;
    }
#pragma omp parallel for schedule(dynamic, 1)
  for (long e = 0; e < npairs; ++e) {
    const chan_t *ch = &ms->ch[pch[e]];
    int ibra = pbr[e];
    int J = ch->J, n = ch->n;
    double *Zm = Z->two + ch->off;
    int kb = ch->kets[ibra];
    int i = ms->kp[kb], j = ms->kq[kb];
    int jji = ms->j2[i], jjj = ms->j2[j];
    int ketmin = (Z->herm == 1) ? ibra : ibra + 1;
    for (int iket = ketmin; iket < n; ++iket) {
      int kk = ch->kets[iket];
      int k = ms->kp[kk], l = ms->kq[kk];
      int jjk = ms->j2[k], jjl = ms->j2[l];
      double commij = 0, commji = 0;
      int pcc = (ms->l[i] + ms->l[l]) % 2;
      int tcc = abs(ms->tz2[i] - ms->tz2[l]) / 2;
      int a1 = abs(jji - jjl) / 2, a2 = abs(jjk - jjj) / 2;
      int Jpmin = a1 > a2 ? a1 : a2;
      a1 = (jji + jjl) / 2;
      a2 = (jjk + jjj) / 2;
      int Jpmax = a1 < a2 ? a1 : a2;
      for (int Jp = Jpmin; Jp <= Jpmax; ++Jp) {
        double sj = orc_sixj_cached(ms, jji, jjj, J, jjk, jjl, Jp);
        if (fabs(sj) < 1e-8) continue;
        const chan_t *cc = &ms->cc[6 * Jp + 2 * (tcc + 1) + pcc];
        int nkc = cc->n;
        int il = cc->map[ket_index(ms, i < l ? i : l, i < l ? l : i)] + (i > l ? nkc : 0);
        int kj = cc->map[ket_index(ms, j < k ? j : k, j < k ? k : j)] + (k > j ? nkc : 0);
        double me1 = Zbar[6 * Jp + 2 * (tcc + 1) + pcc][(long)kj * nkc + il];
        commij -= (2 * Jp + 1) * sj * me1;
      }
      if (k == l) {
        commji = commij;
      } else if (i == j) {
        commji = phase_of((jji + jjj + jjk + jjl) / 2) * commij;
      } else {
        pcc = (ms->l[i] + ms->l[k]) % 2;
        tcc = abs(ms->tz2[i] - ms->tz2[k]) / 2;
        a1 = abs(jjj - jjl) / 2;
        a2 = abs(jjk - jji) / 2;
        Jpmin = a1 > a2 ? a1 : a2;
        a1 = (jjj + jjl) / 2;
        a2 = (jjk + jji) / 2;
        Jpmax = a1 < a2 ? a1 : a2;
        for (int Jp = Jpmin; Jp <= Jpmax; ++Jp) {
          double sj = orc_sixj_cached(ms, jjj, jji, J, jjk, jjl, Jp);
          if (fabs(sj) < 1e-8) continue;
          const chan_t *cc = &ms->cc[6 * Jp + 2 * (tcc + 1) + pcc];
          int nkc = cc->n;
          int ik = cc->map[ket_index(ms, i < k ? i : k, i < k ? k : i)] + (i > k ? nkc : 0);
          int lj = cc->map[ket_index(ms, l < j ? l : j, l < j ? j : l)] + (l > j ? nkc : 0);
          double me1 = Zbar[6 * Jp + 2 * (tcc + 1) + pcc][(long)lj * nkc + ik];
          commji -= (2 * Jp + 1) * sj * me1;
        }
      }
      if (g_fault == 1) commji = 0.0;
      int dbra = (i == j), dket = (k == l);
      double nrm = (dbra == dket) ? 1.0 + dbra : SQRT2;
      double zijkl = -(commij - phase_of((jjk + jjl) / 2 - J) * commji) / nrm;
      Zm[(long)iket * n + ibra] += zijkl;
      if (ibra != iket) Zm[(long)ibra * n + iket] += hZ * zijkl;
    }
  }
  free(pch);
  free(pbr);
  for (int c = 0; c < nch; ++c) free(Zbar[c]);
  free(Zbar);
  g_timer[7] += omp_get_wtime() - t_ph;
}

/* CommutatorScalarScalar, src/Commutator.cc:280-428 (IMSRG(2) part) */
int orc_commutator(const oms_t *ms, const orc_op *X, const orc_op *Y, orc_op *Z, unsigned mask) {
  int hx = X->herm, hy = Y->herm;
  if ((hx == 1 && hy == 1) || (hx == -1 && hy == -1))
    Z->herm = -1;
  else if ((hx == 1 && hy == -1) || (hx == -1 && hy == 1))
    Z->herm = 1;
  else
    Z->herm = 0;
  *Z->zero = 0.0;
  memset(Z->one, 0, sizeof(double) * ms->norb * ms->norb);
  memset(Z->two, 0, sizeof(double) * ms->tbsize);
  double t = omp_get_wtime();
#define TIMED(slot, call)                  \
  do {                                     \
    call;                                  \
    double t2_ = omp_get_wtime();          \
    g_timer[slot] += t2_ - t;              \
    t = t2_;                               \
  } while (0)
  if (mask & ORC_COMM110) TIMED(0, comm110ss(ms, X, Y, Z));
  if (mask & ORC_COMM220) TIMED(1, comm220ss(ms, X, Y, Z));
  if (mask & ORC_COMM111) TIMED(2, comm111ss(ms, X, Y, Z));
  if (mask & ORC_COMM121) TIMED(3, comm121ss(ms, X, Y, Z));
  if (mask & ORC_COMM122) TIMED(4, comm122ss(ms, X, Y, Z));
  int pphh = (mask & ORC_COMM222_PPHH) != 0, c221 = (mask & ORC_COMM221) != 0;
  if (pphh || c221) TIMED(5, comm222_pphh_221(ms, X, Y, Z, pphh, c221));
  if (mask & ORC_COMM222_PH) comm222_phss(ms, X, Y, Z);
#undef TIMED
  g_ncalls++;
  return 0;
}

/* Accumulated seconds per phase since the last reset, named as the reference's
 * IMSRGProfiler timers: [0] comm110ss [1] comm220ss [2] comm111ss [3] comm121ss
 * [4] comm122ss [5] comm222_pp_hh_221ss (ConstructScalarMpp_Mhh + comm221ss)
 * [6] DoPandyaTransformation + ph GEMM ("Build Z_bar", src/Commutator.cc:1808)
 * [7] InversePandyaTransformation (:1821); returns the commutators timed. */
int orc_timers(double *out, int n, int reset) {
  for (int i = 0; i < n && i < 8; ++i) out[i] = g_timer[i];
  int c = g_ncalls;
  if (reset) {
    for (int i = 0; i < 8; ++i) g_timer[i] = 0.0;
    g_ncalls = 0;
  }
  return c;
}

/* ------------------------------------------------------------------ */
/* norms, BCH                                                           */
/* ------------------------------------------------------------------ */
/* Operator::OneBodyNorm, src/Operator.cc:1630-1646 */
double orc_onebody_norm(const oms_t *ms, const orc_op *op) {
  int norb = ms->norb;
  double nrm = 0.0;
  for (int p = 0; p < norb; ++p)
    for (int iq = 0; iq < ms->obn[p]; ++iq) {
      int q = ms->obl[p][iq];
      double v = ONE(op, p, q);
      nrm += v * v * (g_fault == 6 ? 1 : ms->j2[p] + 1);
    }
  return sqrt(nrm);
}

/* TwoBodyME::Norm, src/TwoBodyME.cc:703-724 (scalar: degeneracy 2J+1) */
double orc_twobody_norm(const oms_t *ms, const orc_op *op) {
  double nrm = 0.0;
  for (int c = 0; c < ms->nch; ++c) {
    const chan_t *ch = &ms->ch[c];
    long nn = (long)ch->n * ch->n;
    double f = 0.0;
    for (long e = 0; e < nn; ++e) f += op->two[ch->off + e] * op->two[ch->off + e];
    double n2 = sqrt(f) * (2 * ch->J + 1);
    if (g_fault == 5) n2 = sqrt(f * (2 * ch->J + 1));
    nrm += n2 * n2;
  }
  return sqrt(nrm);
}

/* Operator::DoNormalOrdering2 for a scalar operator (src/Operator.cc:461-579),
 * with sign +1 (DoNormalOrdering, :392-400) or -1 (UndoNormalOrdering, :402-415)
 * over the reference's holes with their occupations:
 *   E0 += sign sum_k (2j_k+1) n_k h_kk + sum_{k<=l} sum_J (2J+1) n_k n_l V^J_{kl,kl} (normalized)
 *   f_ab += sign (2J+1)/(2j_a+1) n_h V~^J_{ah,bh} over the channels (ascending), then the holes h,
 *           for b >= a in a's one-body block (only when the two-body norm exceeds 1e-7);
 * then the one-body part is made (anti-)Hermitian from its upper triangle (Symmetrize /
 * AntiSymmetrize, :1797-1830).  The two-body part is unchanged.  out may alias in. */
void orc_normal_order(const oms_t *ms, const orc_op *in, orc_op *out, int sign) {
  const int norb = ms->norb;
  double e0 = *in->zero;
  for (int ik = 0; ik < ms->nholes; ++ik) {  /* :469-490 */
    const int k = ms->holes[ik];
    e0 += (ms->j2[k] + 1) * sign * ms->occ[k] * ONE(in, k, k);
    for (int il = 0; il < ms->nholes; ++il) {
      const int l = ms->holes[il];
      if (l < k) continue;
      const int Jmin = abs(ms->j2[k] - ms->j2[l]) / 2, Jmax = (ms->j2[k] + ms->j2[l]) / 2;
      for (int J = Jmin; J <= Jmax; ++J)
        e0 += (2 * J + 1) * ms->occ[k] * ms->occ[l] * tbme_norm_J(ms, in->two, J, k, l, k, l);
    }
  }
  double *f = (double *)malloc(sizeof(double) * (size_t)norb * norb);
  memcpy(f, in->one, sizeof(double) * (size_t)norb * norb);
  if (orc_twobody_norm(ms, in) > 1e-7) {  /* :493 */
    for (int c = 0; c < ms->nch; ++c) {   /* :495-560, scalar operator: ch_bra = ch_ket */
      const chan_t *ch = &ms->ch[c];
      const int J = ch->J;
      for (int a = 0; a < norb; ++a) {
        const double ja = ms->j2[a] / 2.0;
        for (int ib = 0; ib < ms->obn[a]; ++ib) {
          const int b = ms->obl[a][ib];
          if (b < a) continue;
          for (int ih = 0; ih < ms->nholes; ++ih) {
            const int h = ms->holes[ih];
            /* GetTBME(ch, ch, a, h, b, h): zero unless both kets lie in channel c */
            const int par = (ms->l[a] + ms->l[h]) % 2, tz = (ms->tz2[a] + ms->tz2[h]) / 2;
            if (par != ch->par || tz != ch->tz) continue;
            double v = orc_tbme_J(ms, in->two, J, a, h, b, h);
            double w = (2 * J + 1.0) / (2 * ja + 1.0);
            if (g_fault == 10) w = (2 * J + 1.0);
            f[(long)b * norb + a] += w * sign * ms->occ[h] * v;
          }
        }
      }
    }
  }
  for (int a = 0; a < norb; ++a)  /* upper triangle -> lower */
    for (int b = a; b < norb; ++b) {
      const double u = f[(long)b * norb + a];
      if (in->herm == -1) {
        f[(long)b * norb + a] = (a == b) ? 0.0 : u;
        f[(long)a * norb + b] = (a == b) ? 0.0 : -u;
      } else {
        f[(long)a * norb + b] = u;
      }
    }
  memcpy(out->one, f, sizeof(double) * (size_t)norb * norb);
  free(f);
  if (out->two != in->two) memcpy(out->two, in->two, sizeof(double) * (size_t)ms->tbsize);
  *out->zero = e0;
  out->herm = in->herm;
}

/* HartreeFock::TransformToHFBasis of a scalar two-body operator,
 * src/HartreeFock.cc:905-1000: OneBody' = C^T OneBody C (:920-923) and, for each
 * channel, V' = Dbra V Dket with Dbra = Dket^T (:933-989), where for the kets
 * i = (p, q) of the old basis and a = (p', q') of the new (both p <= q)
 *   Dket(i, a) = C(p,p') C(q,q') + [p != q] Phase_pq(J) C(q,p') C(p,q')  (:948-953)
 * times sqrt2 if p == q and divided by sqrt2 if p' == q' (:954-955).  C is norb x
 * norb column-major, C(i, a) = <i|a>.  out may alias in. */
void orc_transform_basis(const oms_t *ms, const orc_op *in, const double *C, orc_op *out) {
  const int norb = ms->norb;
  const long n1 = (long)norb * norb;
  double *t1 = (double *)calloc((size_t)n1, sizeof(double));
  double *o1 = (double *)calloc((size_t)n1, sizeof(double));
  for (int b = 0; b < norb; ++b)  /* t1 = OneBody C */
    for (int k = 0; k < norb; ++k) {
      const double c = C[(long)b * norb + k];
      for (int a = 0; a < norb; ++a) t1[(long)b * norb + a] += in->one[(long)k * norb + a] * c;
    }
  for (int b = 0; b < norb; ++b)  /* o1 = C^T t1 */
    for (int a = 0; a < norb; ++a) {
      double acc = 0.0;
      for (int k = 0; k < norb; ++k) acc += C[(long)a * norb + k] * t1[(long)b * norb + k];
      o1[(long)b * norb + a] = acc;
    }
  memcpy(out->one, o1, sizeof(double) * (size_t)n1);
  free(t1);
  free(o1);
  for (int c = 0; c < ms->nch; ++c) {
    const chan_t *ch = &ms->ch[c];
    const int n = ch->n, J = ch->J;
    if (n == 0) continue;
    const long nn = (long)n * n;
    double *D = (double *)malloc(sizeof(double) * (size_t)nn);
    double *T = (double *)calloc((size_t)nn, sizeof(double));
    for (int a = 0; a < n; ++a) {
      const int pp = ms->kp[ch->kets[a]], qq = ms->kq[ch->kets[a]];
      for (int i = 0; i < n; ++i) {
        const int p = ms->kp[ch->kets[i]], q = ms->kq[ch->kets[i]];
        double d = C[(long)pp * norb + p] * C[(long)qq * norb + q];
        if (p != q) d += C[(long)pp * norb + q] * C[(long)qq * norb + p] * ket_phase(ms, p, q, J);
        if (p == q) d *= SQRT2;
        if (pp == qq && g_fault != 11) d /= SQRT2;
        D[(long)a * n + i] = d;
      }
    }
    const double *V = in->two + ch->off;
    double *W = out->two + ch->off;
    if (g_dgemm) {  /* T = V D; W = D^T T */
      const double one = 1.0, zero = 0.0;
      g_dgemm("N", "N", &n, &n, &n, &one, V, &n, D, &n, &zero, T, &n);
      g_dgemm("T", "N", &n, &n, &n, &one, D, &n, T, &n, &zero, W, &n);
    } else {
#pragma omp parallel for schedule(static)
      for (int b = 0; b < n; ++b)
        for (int k = 0; k < n; ++k) {
          const double dk = D[(long)b * n + k];
          for (int a = 0; a < n; ++a) T[(long)b * n + a] += V[(long)k * n + a] * dk;
        }
#pragma omp parallel for schedule(static)
      for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a) {
          double acc = 0.0;
          for (int k = 0; k < n; ++k) acc += D[(long)a * n + k] * T[(long)b * n + k];
          W[(long)b * n + a] = acc;
        }
    }
    free(D);
    free(T);
  }
  if (out != in) *out->zero = *in->zero;
  out->herm = in->herm;
}

/* Operator::Norm, src/Operator.cc:1610-1628 */
double orc_norm(const oms_t *ms, const orc_op *op) {
  double n1 = orc_onebody_norm(ms, op), n2 = orc_twobody_norm(ms, op);
  return sqrt(n1 * n1 + n2 * n2);
}

typedef struct {
  double zero;
  double *one, *two;
  orc_op v;
} ownop;

static void own_alloc(const oms_t *ms, ownop *o, int herm) {
  o->zero = 0;
  o->one = (double *)calloc((size_t)ms->norb * ms->norb, sizeof(double));
  o->two = (double *)calloc((size_t)(ms->tbsize > 0 ? ms->tbsize : 1), sizeof(double));
  o->v.zero = &o->zero;
  o->v.one = o->one;
  o->v.two = o->two;
  o->v.herm = herm;
}
static void own_free(ownop *o) {
  free(o->one);
  free(o->two);
}
static void op_copy(const oms_t *ms, orc_op *dst, const orc_op *src) {
  *dst->zero = *src->zero;
  memcpy(dst->one, src->one, sizeof(double) * ms->norb * ms->norb);
  memcpy(dst->two, src->two, sizeof(double) * ms->tbsize);
  dst->herm = src->herm;
}
/* dst += f * src, formed as (f*src) then added (Operator operator* then +=) */
static void op_axpy(const oms_t *ms, orc_op *dst, double f, const orc_op *src) {
  double t = f * *src->zero;
  *dst->zero += t;
  long n1 = (long)ms->norb * ms->norb;
  for (long e = 0; e < n1; ++e) {
    double v = f * src->one[e];
    dst->one[e] += v;
  }
  for (long e = 0; e < ms->tbsize; ++e) {
    double v = f * src->two[e];
    dst->two[e] += v;
  }
}

/* Standard_BCH_Transform, src/BCH.cc:87-201 */
int orc_bch_transform(const oms_t *ms, const orc_op *Op, const orc_op *Omega, orc_op *Out, int mode, double thr) {
  int max_iter = 40;
  double nx = orc_norm(ms, Op), ny = orc_norm(ms, Omega);
  op_copy(ms, Out, Op);
  int ncomm = 0;
  if (nx > thr) {
    ownop A, B;
    own_alloc(ms, &A, Op->herm);
    own_alloc(ms, &B, Op->herm);
    op_copy(ms, &A.v, Op);
    orc_op *nested = &A.v, *next = &B.v;
    double fact = 1.0;
    double eps = nx * exp(-2 * ny) * thr / (2 * ny);
    for (int i = 1; i <= max_iter; ++i) {
      orc_commutator(ms, Omega, nested, next, ORC_ALL_TERMS);
      ncomm++;
      orc_op *tmp = nested;
      nested = next;
      next = tmp;
      fact /= i;
      op_axpy(ms, Out, fact, nested);
      eps *= i + 1;
      double nn = orc_norm(ms, nested);
      if (mode == 0) {
        if (nn < eps) break;
      } else {
        if (nn < thr) break;
      }
    }
    own_free(&A);
    own_free(&B);
  }
  return ncomm;
}

/* BCH_Product, src/BCH.cc:267-350 */
int orc_bch_product(const oms_t *ms, const orc_op *X, const orc_op *Y, orc_op *Out, double thr) {
  static const double bern[9] = {1.0, -0.5, 1. / 6, 0.0, -1. / 30, 0.0, 1. / 42, 0, -1. / 30};
  static const double fact[9] = {1.0, 1.0, 2.0, 6.0, 24., 120., 720., 5040., 40320.};
  double nx = orc_norm(ms, X), ny = orc_norm(ms, Y);
  /* Z = X + Y (copy of X, then += Y) */
  op_copy(ms, Out, X);
  op_axpy(ms, Out, 1.0, Y);
  Out->herm = X->herm;
  int ncomm = 0;
  ownop A, B;
  own_alloc(ms, &A, 0);
  own_alloc(ms, &B, 0);
  orc_op *nested = &A.v, *next = &B.v;
  orc_commutator(ms, Y, X, nested, ORC_ALL_TERMS);
  ncomm++;
  double nxy = orc_norm(ms, nested);
  if (nxy * nx > thr) {
    orc_commutator(ms, nested, X, next, ORC_ALL_TERMS);
    ncomm++;
    if (g_fault != 9) op_axpy(ms, Out, 1. / 12, next);
  }
  int k = 1;
  while (nxy > thr) {
    if (k < 2 || k % 2 == 0) op_axpy(ms, Out, bern[k] / fact[k], nested);
    k++;
    if (k >= 9) break;
    if (2 * ny * nxy < thr) break;
    orc_commutator(ms, Y, nested, next, ORC_ALL_TERMS);
    ncomm++;
    orc_op *tmp = nested;
    nested = next;
    next = tmp;
    nxy = orc_norm(ms, nested);
  }
  own_free(&A);
  own_free(&B);
  return ncomm;
}

/* ------------------------------------------------------------------ */
/* White generator and Magnus-Euler                                     */
/* ------------------------------------------------------------------ */
/* Get1bDenominator / Get2bDenominator (Epstein-Nesbet), src/Generator.cc:110-188;
 * ConstructGenerator_SingleRef, src/Generator.cc:236-291 */
void orc_white_generator(const oms_t *ms, const orc_op *H, orc_op *Eta) {
  int norb = ms->norb;
  *Eta->zero = 0.0;
  memset(Eta->one, 0, sizeof(double) * norb * norb);
  memset(Eta->two, 0, sizeof(double) * ms->tbsize);
  Eta->herm = -1;
  const double cut = 1e-6;
  for (int ia = 0; ia < ms->ncore; ++ia) {
    int a = ms->core[ia];
    for (int i = 0; i < norb; ++i) {
      if (ms->cvq[i] == 0) continue;
      double d = ONE(H, i, i) - ONE(H, a, a);
      d += (ms->occ[i] - ms->occ[a]) * orc_tbme_monopole(ms, H->two, i, a, i, a);
      if (fabs(d) < cut) d = cut;
      double e = ONE(H, i, a) / d;
      ONE(Eta, i, a) = e;
      ONE(Eta, a, i) = -e;
    }
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < ms->nch; ++c) {
    const chan_t *ch = &ms->ch[c];
    int n = ch->n;
    const double *H2 = H->two + ch->off;
    double *E2 = Eta->two + ch->off;
    for (int ik = 0; ik < ch->ncore; ++ik) {
      int iket = ch->core[ik];
      int kk = ch->kets[iket];
      int k = ms->kp[kk], l = ms->kq[kk];
      for (int ib = 0; ib < ch->nnoncore; ++ib) {
        int ibra = ch->noncore[ib];
        int kb = ch->kets[ibra];
        int i = ms->kp[kb], j = ms->kq[kb];
        double ni = ms->occ[i], nj = ms->occ[j], nk = ms->occ[k], nl = ms->occ[l];
        double d = ONE(H, i, i) + ONE(H, j, j) - ONE(H, k, k) - ONE(H, l, l);
        d += (1 - ni - nj) * orc_tbme_monopole(ms, H->two, i, j, i, j);
        d -= (1 - nk - nl) * orc_tbme_monopole(ms, H->two, k, l, k, l);
        d += (ni - nk) * orc_tbme_monopole(ms, H->two, i, k, i, k);
        d += (ni - nl) * orc_tbme_monopole(ms, H->two, i, l, i, l);
        d += (nj - nk) * orc_tbme_monopole(ms, H->two, j, k, j, k);
        d += (nj - nl) * orc_tbme_monopole(ms, H->two, j, l, j, l);
        if (fabs(d) < cut) d = cut;
        double e = H2[(long)iket * n + ibra] / d;
        E2[(long)iket * n + ibra] = e;
        E2[(long)ibra * n + iket] = -e;
      }
    }
  }
}

/* Solve_magnus_euler, src/IMSRGSolver.cc:236-321, with magnus_adaptive=false
 * (fixed ds) and NewOmega splitting (src/IMSRGSolver.cc:44-84): a finished
 * Omega stays in the solver's list (Omega.emplace_back, then the new last
 * entry is erased), and the flow continues from H_saved = H(s).  When omegas
 * is not NULL, the list (finished segments, then the current one) is copied
 * into omegas[0 .. *n_omega-1], at most max_omega of them. */
int orc_magnus_flow(const oms_t *ms, const orc_op *H0, orc_op *Hs, double ds, double smax, double eta_crit,
                    double omega_norm_max, int bch_mode, double bch_thr, double product_thr, int max_steps,
                    double *e0_out, double *eta_norm_out, int *ncomm_out, orc_op *omegas, int max_omega,
                    int *n_omega) {
  ownop Eta, Omega, Omega2, Hsaved;
  own_alloc(ms, &Eta, -1);
  own_alloc(ms, &Omega, -1);
  own_alloc(ms, &Omega2, -1);
  own_alloc(ms, &Hsaved, 1);
  op_copy(ms, Hs, H0);
  op_copy(ms, &Hsaved.v, H0);
  int nsplit = 0;
  orc_white_generator(ms, Hs, &Eta.v);
  double s = 0.0;
  int step = 0;
  double Elast = *H0->zero;
  while (s < smax && step < max_steps) {
    double norm_eta = orc_norm(ms, &Eta.v);
    if (eta_norm_out) eta_norm_out[step] = norm_eta;
    if (norm_eta < eta_crit) break;
    if (norm_eta > 1e12 || fabs(Elast) > 1e9) break;
    double norm_omega = orc_norm(ms, &Omega.v);
    if (norm_omega > omega_norm_max) {
      op_copy(ms, &Hsaved.v, Hs);
      if (omegas && nsplit < max_omega) op_copy(ms, &omegas[nsplit], &Omega.v);
      if (g_fault != 7) {
        memset(Omega.one, 0, sizeof(double) * ms->norb * ms->norb);
        memset(Omega.two, 0, sizeof(double) * ms->tbsize);
        Omega.zero = 0;
      }
      nsplit++;
    }
    double dss = ds;
    if (dss > smax - s) dss = smax - s;
    s += dss;
    long n1 = (long)ms->norb * ms->norb;
    Eta.zero *= dss;
    for (long e = 0; e < n1; ++e) Eta.one[e] *= dss;
    for (long e = 0; e < ms->tbsize; ++e) Eta.two[e] *= dss;
    int nc = orc_bch_product(ms, &Eta.v, &Omega.v, &Omega2.v, product_thr);
    op_copy(ms, &Omega.v, &Omega2.v);
    Omega.v.herm = -1;
    if (g_fault == 8) op_copy(ms, &Hsaved.v, Hs);
    nc += orc_bch_transform(ms, (nsplit == 0 && g_fault != 8) ? H0 : &Hsaved.v, &Omega.v, Hs, bch_mode, bch_thr);
    orc_white_generator(ms, Hs, &Eta.v);
    if (e0_out) e0_out[step] = *Hs->zero;
    if (ncomm_out) ncomm_out[step] = nc;
    Elast = *Hs->zero;
    step++;
  }
  if (omegas && nsplit < max_omega) op_copy(ms, &omegas[nsplit], &Omega.v);
  if (n_omega) *n_omega = nsplit + 1;
  own_free(&Eta);
  own_free(&Omega);
  own_free(&Omega2);
  own_free(&Hsaved);
  return step;
}

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_num_threads(int n) { omp_set_num_threads(n); }
