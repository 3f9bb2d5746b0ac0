/*
 * oracle.h -- CPU FP64 restatement of IMSRG++'s scalar IMSRG(2) commutator
 * path (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the checker for the HIP product in imsrg_amd/.  It is
 * built from oracle.c by oracle/Makefile and loaded only by tests/, by
 * __graft_entry__.smoke() and by bench.py's CPU-baseline leg.  Product code
 * never links, loads or calls it.
 *
 * Operator layout (shared with the product's host-buffer API):
 *   zero : one double
 *   one  : norb*norb doubles, column-major, element (p,q) at one[q*norb+p]
 *   two  : sum_ch n_ch^2 doubles; channel ch (ascending ch = 6J+2(Tz+1)+parity)
 *          is an n_ch x n_ch column-major block starting at offset off[ch].
 *          Stored values are the normalized J-coupled antisymmetrized TBMEs
 *          for kets p<=q (reference: src/TwoBodyME.hh:36-41).
 *   herm : +1 Hermitian, -1 anti-Hermitian, 0 non-Hermitian.
 *
 * Term mask bits follow Commutator::comm_term_on (src/Commutator.cc:36-46).
 */
#ifndef IMSRG_ORACLE_H
#define IMSRG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_COMM110 (1u << 0)
#define ORC_COMM220 (1u << 1)
#define ORC_COMM111 (1u << 2)
#define ORC_COMM121 (1u << 3)
#define ORC_COMM221 (1u << 4)
#define ORC_COMM122 (1u << 5)
#define ORC_COMM222_PPHH (1u << 6)
#define ORC_COMM222_PH (1u << 7)
#define ORC_ALL_TERMS 0xFFu

typedef struct oms oms_t;

typedef struct {
  double *zero;
  double *one;
  double *two;
  int herm;
} orc_op;

/* basis construction */
oms_t *orc_ms_new_ho(int emax, double A_ref, double Z_ref);
oms_t *orc_ms_new(int norb, const int *n, const int *l, const int *j2, const int *tz2,
                  const double *occ, const int *cvq, int emax);
void orc_ms_free(oms_t *ms);

int orc_ms_norb(const oms_t *ms);
int orc_ms_nkets(const oms_t *ms);
int orc_ms_nch(const oms_t *ms);
long orc_ms_tbsize(const oms_t *ms);
void orc_ms_orbit(const oms_t *ms, int i, int *n, int *l, int *j2, int *tz2, double *occ, int *cvq);
void orc_ms_channel(const oms_t *ms, int ch, int *J, int *par, int *tz, int *nkets, long *off);
void orc_ms_channel_ket(const oms_t *ms, int ch, int i, int *p, int *q);
void orc_ms_cc_channel(const oms_t *ms, int ch, int *J, int *par, int *tz, int *nkets, int *nph);
void orc_ms_cc_channel_ket(const oms_t *ms, int ch, int i, int *p, int *q);

/* 6j: exact Racah value (double-double) and the cached dense table */
double orc_sixj_exact(int j2a, int j2b, int j2c, int j2d, int j2e, int j2f);
double orc_sixj_cached(const oms_t *ms, int jj1, int jj2, int J3, int jj4, int jj5, int J6);

/* TBME accessors (reference src/TwoBodyME.cc:305-409, 615-682) */
double orc_tbme_J(const oms_t *ms, const double *two, int J, int a, int b, int c, int d);
double orc_tbme_monopole(const oms_t *ms, const double *two, int a, int b, int c, int d);

/* commutator Z = [X,Y]; Z's herm is set by the rule of src/Commutator.cc:292-297 */
int orc_commutator(const oms_t *ms, const orc_op *X, const orc_op *Y, orc_op *Z, unsigned mask);

/* norms and arithmetic (reference src/Operator.cc:1610-1646, src/TwoBodyME.cc:703-724) */
double orc_norm(const oms_t *ms, const orc_op *op);
double orc_onebody_norm(const oms_t *ms, const orc_op *op);
double orc_twobody_norm(const oms_t *ms, const orc_op *op);

/* Operator::DoNormalOrdering (sign +1) / UndoNormalOrdering (sign -1) of a scalar
 * operator over the reference's holes (reference src/Operator.cc:392-415, 461-579);
 * out may alias in */
void orc_normal_order(const oms_t *ms, const orc_op *in, orc_op *out, int sign);

/* HartreeFock::TransformToHFBasis of a scalar operator (reference
 * src/HartreeFock.cc:905-1000) with the basis change C (norb x norb, column-major,
 * C(i,a) = <i|a>); out may alias in */
void orc_transform_basis(const oms_t *ms, const orc_op *in, const double *C, orc_op *out);

/* BCH (reference src/BCH.cc:87-201, 267-350).
 * mode 0: reference epsilon rule with threshold thr (default 1e-9);
 * mode 1: stop when ||nested|| < thr (absolute).
 * Return value: number of nested commutators evaluated. */
int orc_bch_transform(const oms_t *ms, const orc_op *Op, const orc_op *Omega, orc_op *Out,
                      int mode, double thr);
int orc_bch_product(const oms_t *ms, const orc_op *X, const orc_op *Y, orc_op *Out, double thr);

/* White generator, Epstein-Nesbet denominators (reference src/Generator.cc:110-291) */
void orc_white_generator(const oms_t *ms, const orc_op *H, orc_op *Eta);

/* Magnus-Euler flow (reference src/IMSRGSolver.cc:236-321), fixed ds.
 * Writes E0 after each step into e0_out[0..max_steps-1]; returns number of steps.
 * product_thr is the BCH_Product threshold (reference default 1e-4).  omegas
 * (may be NULL) receives the Omega list: the segments finished at each split
 * (src/IMSRGSolver.cc:44-84), then the current one; *n_omega = splits + 1. */
int orc_magnus_flow(const oms_t *ms, const orc_op *H0, orc_op *Hs, double ds, double smax,
                    double eta_crit, double omega_norm_max, int bch_mode, double bch_thr,
                    double product_thr, int max_steps, double *e0_out, double *eta_norm_out,
                    int *ncomm_out, orc_op *omegas, int max_omega, int *n_omega);

/* Channel GEMMs through a Fortran-convention dgemm (the reference calls BLAS
 * through Armadillo); NULL restores the plain loops.  orc_has_dgemm: 1 if set. */
void orc_set_dgemm(void *fn);
int orc_has_dgemm(void);

/* Seconds per phase accumulated by orc_commutator since the last reset, named
 * as the reference's profiler timers: [0] comm110ss [1] comm220ss [2] comm111ss
 * [3] comm121ss [4] comm122ss [5] comm222_pp_hh_221ss [6] Pandya transform and
 * ph GEMM ("Build Z_bar") [7] InversePandyaTransformation.  Returns the number
 * of commutators timed; reset != 0 clears the sums. */
int orc_timers(double *out, int n, int reset);

/* planted-mistake switch used by the checker tests (0 = off) */
void orc_set_fault(int f);

int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
