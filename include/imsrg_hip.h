/*
 * imsrg_hip.h -- C ABI of the MI355X-native scalar IMSRG(2) commutator and
 * Magnus/BCH path (libimsrg_hip.so).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, never
 * throws or exits, and returns an int status (IMSRG_OK or a negative code;
 * imsrg_hip_last_error() gives the message).  Each declaration names the
 * reference interface it replaces (paths relative to the IMSRG++ source).
 *
 * Operator layout (identical to the reference's Operator storage):
 *   zero_body  double                        Operator::ZeroBody   (src/Operator.hh:49)
 *   one_body   norb*norb, column-major       Operator::OneBody    (arma::mat)
 *   two_body   one n_ch x n_ch column-major  TwoBodyME::MatEl[{ch,ch}] (src/TwoBodyME.hh:55)
 *              block per channel ch = 6J + 2(Tz+1) + parity, J = 0..max j2,
 *              kets p<=q in ascending ket index, normalized TBMEs.
 * The "packed" form concatenates the channel blocks in ascending ch.
 *
 * Supported operators: scalar (rank_J = rank_T = 0), parity-conserving,
 * not reduced, particle rank <= 2, Hermitian (+1) or anti-Hermitian (-1)
 * and exactly (anti)symmetric as declared.  Anything else returns
 * IMSRG_E_UNSUPPORTED so that the caller keeps the reference path
 * (src/Commutator.cc:182-276 dispatch).
 */
#ifndef IMSRG_HIP_H
#define IMSRG_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define IMSRG_OK 0
#define IMSRG_E_INVALID (-1)
#define IMSRG_E_UNSUPPORTED (-2)
#define IMSRG_E_HIP (-3)
#define IMSRG_E_NOMEM (-4)
#define IMSRG_E_NONFINITE (-5)
#define IMSRG_E_NOT_CONVERGED (-6)

/* term switches, src/Commutator.cc:36-46 (comm_term_on) */
#define IMSRG_COMM110 (1u << 0)
#define IMSRG_COMM220 (1u << 1)
#define IMSRG_COMM111 (1u << 2)
#define IMSRG_COMM121 (1u << 3)
#define IMSRG_COMM221 (1u << 4)
#define IMSRG_COMM122 (1u << 5)
#define IMSRG_COMM222_PPHH (1u << 6)
#define IMSRG_COMM222_PH (1u << 7)
#define IMSRG_ALL_TERMS 0xFFu

/* arithmetic of the channel GEMMs */
#define IMSRG_PREC_FP64 0  /* every product and sum in FP64 */
#define IMSRG_PREC_MIXED 1 /* pp/hh and ph GEMMs on reduced-precision MFMA, everything else
                              FP64.  Default (tuning mixed_*_fp32 = 2): each FP64 operand
                              split into three bf16 values (relative error 2^-26), six
                              bf16 MFMA products per product, FP32 partial sums over 32 k
                              flushed into FP64 accumulators; channels of up to
                              pphh_small_tile_max kets on FP64 64x64 tiles.  Setting 1:
                              FP32-rounded operands on FP32 MFMA, FP32 partial sums over
                              32 k (16 k on 64x64 tiles) flushed into FP64; the ph GEMM on
                              this setting (its default) stores its output Zbar in FP32,
                              which the inverse Pandya gather sums in FP64 */
#define IMSRG_PREC_SPLIT 2 /* pp/hh and ph GEMMs on 128x128 tiles: each FP64 operand split into
                              two FP32 values (hi + lo), three FP32 MFMA products per
                              product (hi*hi + hi*lo + lo*hi), FP32 partial sums flushed
                              into FP64 every 4 k; 64x64-tile GEMMs and everything else FP64.
                              Not faster than FP64 on MI355X (three FP32 MFMAs cost 1.5x
                              one FP64 MFMA): an accuracy reference for the mixed mode */

typedef struct imsrg_ctx imsrg_ctx; /* basis tables, 6j table, stream, workspaces */
typedef struct imsrg_dop imsrg_dop; /* device-resident operator */

/* Single-particle basis in the reference's orbit order (src/ModelSpace.cc:376-399). */
typedef struct {
  int norb;
  const int *n, *l, *j2, *tz2; /* j2, tz2 doubled */
  const double *occ;           /* occupation in [0,1] */
  const int *cvq;              /* 0 core, 1 valence, 2 q-space */
  int emax;                    /* 6j table covers j <= emax + 1/2 (src/ModelSpace.cc:176) */
} imsrg_basis_desc;

/* Zero-copy view of a host Operator (src/Operator.hh:49-71). */
typedef struct {
  double zero_body;
  double *one_body;         /* norb*norb column-major */
  double *const *two_body;  /* [nch] block pointers; entries may be NULL where n_ch = 0 */
  int herm;                 /* +1 Hermitian, -1 anti-Hermitian, 0 non-Hermitian */
  int particle_rank;        /* 1 or 2 */
  int rank_J, rank_T, parity, reduced;
} imsrg_op_view;

typedef struct {
  unsigned term_mask; /* IMSRG_COMM* bits; 0 means all */
  int precision;      /* IMSRG_PREC_* */
  int use_imsrg3;     /* must be 0 (IMSRG(3) terms, src/Commutator.cc:337-420, are not provided) */
} imsrg_comm_opts;

typedef struct {
  int mode;            /* (0 or 1; others are IMSRG_E_INVALID)
                          0: reference rule eps = ||Op|| e^{-2||Omega||} thr / (2||Omega||) growing by (i+1)
                          (src/BCH.cc:132,189-191); 1: stop when ||nested|| < thr */
  double transform_thr;/* bch_transform_threshold, default 1e-9 (src/BCH.cc:16) */
  double product_thr;  /* bch_product_threshold, default 1e-4 (src/BCH.cc:17) */
  int max_terms;       /* default 40 (src/BCH.cc:91) */
  int precision;       /* IMSRG_PREC_* for the nested commutators */
} imsrg_bch_opts;

/* ---- context ---------------------------------------------------------- */
/* Builds kets, (J,parity,Tz) and cross-coupled channels (ModelSpace::SetupKets,
 * src/ModelSpace.cc:1120-1201; TwoBodyChannel.cc:19-188) and the FP64 6j table
 * on the device (replaces SixJCache_112112, src/AngMomCache.cc:8-54). */
int imsrg_hip_ctx_create(const imsrg_basis_desc *basis, int device, imsrg_ctx **out);
/* Convenience: HO basis to emax with the single-reference filling of
 * ModelSpace::GetOrbitsAZ (src/ModelSpace.cc:611-645), core = reference. */
int imsrg_hip_ctx_create_ho(int emax, double A, double Z, int device, imsrg_ctx **out);
int imsrg_hip_ctx_destroy(imsrg_ctx *ctx);
const char *imsrg_hip_last_error(const imsrg_ctx *ctx);
int imsrg_hip_ctx_info(const imsrg_ctx *ctx, int *norb, int *nch, long *two_body_size);
int imsrg_hip_ctx_orbit(const imsrg_ctx *ctx, int i, int *n, int *l, int *j2, int *tz2, double *occ, int *cvq);
int imsrg_hip_ctx_channel(const imsrg_ctx *ctx, int ch, int *J, int *parity, int *Tz, int *nkets, long *offset);
int imsrg_hip_ctx_channel_ket(const imsrg_ctx *ctx, int ch, int i, int *p, int *q);
int imsrg_hip_ctx_sixj(imsrg_ctx *ctx, int jj1, int jj2, int J3, int jj4, int jj5, int J6, double *value);
int imsrg_hip_synchronize(imsrg_ctx *ctx);

/* ---- host-buffer drop-ins (PCIe-inclusive) ----------------------------- */
/* Operator Commutator::CommutatorScalarScalar(const Operator& X, const Operator& Y)
 * (src/Commutator.hh:58, src/Commutator.cc:280-428).  Z->herm is written with the
 * rule of src/Commutator.cc:292-297; Z's buffers are overwritten. */
int imsrg_hip_commutator_ss(imsrg_ctx *ctx, const imsrg_op_view *X, const imsrg_op_view *Y, imsrg_op_view *Z,
                            const imsrg_comm_opts *opts);
/* Operator BCH::BCH_Transform(const Operator& Op, const Operator& Omega) (src/BCH.hh:23, src/BCH.cc:78-201) */
int imsrg_hip_bch_transform(imsrg_ctx *ctx, const imsrg_op_view *Op, const imsrg_op_view *Omega, imsrg_op_view *Out,
                            const imsrg_bch_opts *opts, int *n_commutators);
/* Operator BCH::BCH_Product(Operator& X, Operator& Y) (src/BCH.hh:22, src/BCH.cc:267-350) */
int imsrg_hip_bch_product(imsrg_ctx *ctx, const imsrg_op_view *X, const imsrg_op_view *Y, imsrg_op_view *Out,
                          const imsrg_bch_opts *opts, int *n_commutators);

/* ---- device-resident operators ----------------------------------------- */
int imsrg_hip_op_create(imsrg_ctx *ctx, int herm, imsrg_dop **out);
int imsrg_hip_op_destroy(imsrg_dop *op);
int imsrg_hip_op_get_herm(const imsrg_dop *op, int *herm);
int imsrg_hip_op_set_herm(imsrg_dop *op, int herm);
int imsrg_hip_op_upload(imsrg_dop *op, const imsrg_op_view *view);
int imsrg_hip_op_download(imsrg_dop *op, imsrg_op_view *view);
/* packed layout: two = concatenated channel blocks, length two_body_size */
int imsrg_hip_op_upload_packed(imsrg_dop *op, double zero, const double *one, const double *two, int herm);
int imsrg_hip_op_download_packed(imsrg_dop *op, double *zero, double *one, double *two, int *herm);
/* Seeded synthetic operator: one-body (R +/- R^T)/sqrt(2m) per (l,j,tz) block,
 * two-body (R +/- R^T)/sqrt(2n) per channel, R ~ N(0,1) from SplitMix64 +
 * Box-Muller; zero-body 0. */
int imsrg_hip_op_fill_random(imsrg_dop *op, int herm, unsigned long long seed);
/* Operator arithmetic used by BCH (src/Operator.cc:132-246) */
int imsrg_hip_op_copy(imsrg_dop *dst, const imsrg_dop *src);
int imsrg_hip_op_scale(imsrg_dop *op, double f);
int imsrg_hip_op_axpy(imsrg_dop *y, double a, const imsrg_dop *x); /* y += a*x */
int imsrg_hip_op_zero(imsrg_dop *op);
/* Operator::Norm / OneBodyNorm / TwoBodyNorm (src/Operator.cc:1610-1646, src/TwoBodyME.cc:703-724) */
int imsrg_hip_op_norm(imsrg_dop *op, double *norm1, double *norm2, double *norm);
/* Operator::DoNormalOrdering (sign = +1) and Operator::UndoNormalOrdering (sign = -1) of a
   scalar two-body operator, DoNormalOrdering2 over the reference's holes
   (src/Operator.cc:392-415, 461-579): E0 += sign sum_k (2j_k+1) n_k h_kk + sum_{k<=l} sum_J
   (2J+1) n_k n_l V^J_klkl, f_ab += sign sum_J (2J+1)/(2j_a+1) sum_h n_h V~^J_ahbh (when the
   two-body norm exceeds 1e-7), the one-body part completed from its upper triangle; the
   two-body part is unchanged.  out may be op (in place); both on the same context.
   Asynchronous on the context's stream. */
int imsrg_hip_op_normal_order(imsrg_dop *out, const imsrg_dop *op, int sign);
/* HartreeFock::TransformToHFBasis of a scalar two-body operator (src/HartreeFock.cc:905-1000):
   OneBody' = C^T OneBody C, and per (J, parity, Tz) channel V' = D^T V D with
   D(pq, p'q') = C(p,p') C(q,q') + [p != q] (-1)^{(j_p+j_q)/2+1+J} C(q,p') C(p,q'),
   times sqrt2 if p == q, divided by sqrt2 if p' == q' (kets p <= q of the old basis,
   p' <= q' of the new); the zero-body part and the Hermiticity are unchanged.
   C: host memory, norb x norb, column-major, C(i,a) = <i|a> (HartreeFock::C, block
   diagonal in (l, j, tz) for an HF basis).  The channel products run as DGEMMs through
   rocBLAS, loaded on the first call.  out may be op; both on the same context.
   Synchronous: returns when out holds the result. */
int imsrg_hip_op_transform_basis(imsrg_dop *out, const imsrg_dop *op, const double *C);

/* ReadWrite::ReadBareTBME_Darmstadt (src/ReadWrite.cc:664-728, 834-957): the two-body part
   of op (Hermitian) from a Darmstadt me2j file -- plain text (first line a header), gzip (.gz)
   or binary (.bin: 255-byte header, then floats) -- with the file's emax, E2max and lmax
   (emax < 0, E2max <= 0, lmax <= 0: the basis's emax, 2 emax, emax).  Records (a b c d J) in
   the file's order carry (T,Tz) = (0,0), (1,1), (1,0), (1,-1) un-normalized elements in float;
   pp and nn are set, proton-neutron kets get the isospin combination (Set_pn_TBME_from_iso).
   The rest of op is unchanged.  A file with fewer records than the loops need is
   IMSRG_E_INVALID (the reference reads zeros).  *nrecords (may be NULL): records read.
   Synchronous. */
int imsrg_hip_read_me2j(imsrg_dop *op, const char *path, int emax, int E2max, int lmax, long *nrecords);

/* HartreeFock::Solve options (src/HartreeFock.cc:55-167); zero fields take the defaults. */
typedef struct {
  double tolerance; /* convergence of the single-particle energies (default 1e-8, HartreeFock::tolerance) */
  int max_iter;     /* iterations (default and maximum 600: beyond 600 the reference mixes with
                       random factors, which is not provided) */
} imsrg_hf_opts;

/* HartreeFock::Solve + CalcEHF for a two-body Hermitian Hamiltonian H on the context's
   reference, occupations frozen (the reference's default): the monopoles
   sum_J (2J+1) V~^J_{ia,jb} of BuildMonopoleV / UpdateF (src/HartreeFock.cc:250-285, 750-822)
   come from the device; the iteration (Diagonalize per (l, j, tz) block, ReorderCoefficients'
   sign rule, UpdateDensityMatrix with density mixing 0.2, from iteration 101 the DIIS update
   UpdateDensityMatrix_DIIS while its last error matrix exceeds 0.01, from iteration 500 the
   mixing factors 0.7 / 0.5, UpdateF, CheckConvergence: src/HartreeFock.cc:55-167, 211-245,
   613-683, 834-903) runs on the host.  Outputs (host):
   C (norb x norb, column-major, C(i,a) = <i|a>), energies (norb, may be NULL), F (the Fock
   matrix, norb x norb, may be NULL), ehf[3] = {EHF, e1hf, e2hf} (may be NULL), iterations
   (may be NULL).  IMSRG_E_NOT_CONVERGED (outputs written) when max_iter is reached. */
int imsrg_hip_hf_solve(const imsrg_dop *H, const imsrg_hf_opts *opts, double *C, double *energies, double *F,
                       double *ehf, int *iterations);
/* HartreeFock::GetNormalOrderedH (src/HartreeFock.cc:1079-1172) without three-body forces:
   out = H in the basis C with ZeroBody = ehf, OneBody = C^T F C, two-body part as
   imsrg_hip_op_transform_basis; out is Hermitian.  Synchronous. */
int imsrg_hip_hf_normal_ordered_h(imsrg_dop *out, const imsrg_dop *H, const double *C, const double *F,
                                  double ehf);

/* device-resident versions of the drop-ins (no PCIe traffic) */
int imsrg_hip_commutator_ss_dev(imsrg_ctx *ctx, const imsrg_dop *X, const imsrg_dop *Y, imsrg_dop *Z,
                                const imsrg_comm_opts *opts);
/* Stream-ordered form of imsrg_hip_commutator_ss_dev for callers that manage
 * their own streams (one context per host thread, see the thread-safety notes
 * of the reference's single-caller design): `stream` is a hipStream_t of the
 * context's device (NULL: the null stream).  The commutator starts after the
 * work queued on `stream` so far, and work queued on `stream` later starts
 * after it; the call returns without a host synchronisation.  Z's Hermiticity
 * is set at once. */
int imsrg_hip_commutator_ss_dev_async(imsrg_ctx *ctx, const imsrg_dop *X, const imsrg_dop *Y, imsrg_dop *Z,
                                      const imsrg_comm_opts *opts, void *stream);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The same
 * commutator as imsrg_hip_commutator_ss_dev, with every per-call workspace
 * (Pandya matrices, partial sums, comm122 coefficients) carved from one static
 * device array of 56 MiB per GPU, held from module load until the process
 * ends; the call allocates nothing.  Bases whose workspaces exceed it (above
 * emax = 6 for the O16 reference) return IMSRG_E_INVALID.  All calls on one
 * GPU share the array, so calls that overlap on one GPU (other host threads,
 * other contexts' streams) race; calls ordered on one stream do not. */
int imsrg_hip_commutator_ss_dev_noalloc(imsrg_ctx *ctx, const imsrg_dop *X, const imsrg_dop *Y, imsrg_dop *Z,
                                        const imsrg_comm_opts *opts);
int imsrg_hip_bch_transform_dev(imsrg_ctx *ctx, const imsrg_dop *Op, const imsrg_dop *Omega, imsrg_dop *Out,
                                const imsrg_bch_opts *opts, int *n_commutators);
int imsrg_hip_bch_product_dev(imsrg_ctx *ctx, const imsrg_dop *X, const imsrg_dop *Y, imsrg_dop *Out,
                              const imsrg_bch_opts *opts, int *n_commutators);

/* Many independent commutators Z_i = [X_i, Y_i] (contexts may differ) in one
 * grouped launch sequence.  All contexts must live on the same device. */
int imsrg_hip_commutator_ss_batched(int npairs, imsrg_ctx *const *ctxs, const imsrg_dop *const *X,
                                    const imsrg_dop *const *Y, imsrg_dop *const *Z, const imsrg_comm_opts *opts);

/* White generator with Epstein-Nesbet denominators, single reference
 * (Generator::ConstructGenerator_SingleRef, src/Generator.cc:110-291). */
int imsrg_hip_white_generator_dev(imsrg_ctx *ctx, const imsrg_dop *H, imsrg_dop *Eta);

/* Magnus-Euler flow (IMSRGSolver::Solve_magnus_euler, src/IMSRGSolver.cc:236-321)
 * with fixed ds, Omega splitting at omega_norm_max (NewOmega, :44-84).
 * Hs receives H(s).  e0/eta_norm/ncomm (length max_steps, may be NULL) receive
 * E0(s), ||eta|| and commutators per step; *nsteps the number of steps.  A
 * BCH_Transform that reaches max_terms does not stop the flow (the reference
 * only warns, src/BCH.cc:192-195).  The Omega segments are dropped at the end;
 * imsrg_hip_solver_magnus_euler keeps them. */
int imsrg_hip_magnus_flow(imsrg_ctx *ctx, const imsrg_dop *H0, imsrg_dop *Hs, double ds, double smax,
                          double eta_criterion, double omega_norm_max, const imsrg_bch_opts *opts, int max_steps,
                          double *e0, double *eta_norm, int *ncomm, int *nsteps);

/* ---- Operator binary format ------------------------------------------- */
/* Operator::WriteBinary / ReadBinary (src/Operator.cc:341-388) with
 * TwoBodyME::WriteBinary / ReadBinary (src/TwoBodyME.cc:812-838): the byte
 * stream the reference writes to an std::ofstream (Omega checkpoints,
 * src/IMSRGSolver.cc:57-78).  Fields are packed with no padding: int rank_J,
 * rank_T, parity, particle_rank, legs, E2max, E3max; bool hermitian,
 * antihermitian; int nChannels; double ZeroBody; norb*norb one-body doubles
 * (column-major); for legs > 3: size_t nChannels, bool hermitian,
 * antihermitian, int rank_J, rank_T, parity, then every channel block in
 * ascending ch (column-major).  E2max/E3max are written as the reference
 * ModelSpace defaults 2*emax and min(14, 3*emax) and ignored when read.  Reading
 * accepts parity-conserving scalars of particle rank 1 or 2 of the same
 * channel count, and sets the operator's Hermiticity from the flags. */
int imsrg_hip_op_binary_size(const imsrg_dop *op, long *bytes);
int imsrg_hip_op_to_binary(imsrg_dop *op, void *buf, long size);
int imsrg_hip_op_from_binary(imsrg_dop *op, const void *buf, long size);
int imsrg_hip_op_write_binary(imsrg_dop *op, const char *path);
int imsrg_hip_op_read_binary(imsrg_dop *op, const char *path);

/* ---- IMSRGSolver: Magnus flow with its Omega list ----------------------- */
/* The state of IMSRGSolver (src/IMSRGSolver.hh) for the Magnus-Euler flow:
 * the list of Omega segments (the finished ones, then the current one). */
typedef struct imsrg_solver imsrg_solver;
int imsrg_hip_solver_create(imsrg_ctx *ctx, imsrg_solver **out);
int imsrg_hip_solver_destroy(imsrg_solver *s);
/* IMSRGSolver::SetScratchDir: with dir != NULL and not empty, each finished
 * segment is written to dir/OMEGA_<pid><n> in the binary format above and
 * freed from HBM (NewOmega, src/IMSRGSolver.cc:44-84); NULL keeps them in HBM. */
int imsrg_hip_solver_set_scratch(imsrg_solver *s, const char *dir);
/* IMSRGSolver::Solve_magnus_euler (src/IMSRGSolver.cc:236-321), arguments as
 * imsrg_hip_magnus_flow.  A previous flow's segments are dropped first. */
int imsrg_hip_solver_magnus_euler(imsrg_solver *s, const imsrg_dop *H0, imsrg_dop *Hs, double ds, double smax,
                                  double eta_criterion, double omega_norm_max, const imsrg_bch_opts *opts,
                                  int max_steps, double *e0, double *eta_norm, int *ncomm, int *nsteps);
/* Omega.size(); segment i copied into out (the last is the current one) */
int imsrg_hip_solver_num_omega(const imsrg_solver *s, int *n);
int imsrg_hip_solver_get_omega(imsrg_solver *s, int i, imsrg_dop *out);
/* IMSRGSolver::Transform_Partial(Op, first) (src/IMSRGSolver.cc:959-1004; Transform
 * is first = 0): Out = BCH_Transform over segments first .. n-1 in turn.
 * IMSRGSolver::InverseTransform (:939-951): over -Omega_i from the last to the
 * first.  *ncomm receives the nested commutators used (may be NULL). */
int imsrg_hip_solver_transform(imsrg_solver *s, const imsrg_dop *Op, imsrg_dop *Out, int first,
                               const imsrg_bch_opts *opts, int *ncomm);
int imsrg_hip_solver_inverse_transform(imsrg_solver *s, const imsrg_dop *Op, imsrg_dop *Out,
                                       const imsrg_bch_opts *opts, int *ncomm);

/* No counterpart in the reference.  on != 0: every device commutator on ctx
 * (single calls and the nested commutators of BCH_Transform, BCH_Product and
 * the Magnus flow) replays a hipGraph of its launch sequence, captured on the
 * first call for that set of operand buffers, Hermiticity flags, terms and
 * precision (up to 16 kept per context, oldest dropped).  on == 0 drops the
 * graphs.  Ignored while profiling is on. */
int imsrg_hip_set_graphs(imsrg_ctx *ctx, int on);

/* ---- launch tuning ------------------------------------------------------
 * No counterpart in the reference (its GEMMs go to BLAS).  The choices below
 * change only how work is tiled and ordered, and under IMSRG_PREC_MIXED which
 * GEMM family runs on FP32 MFMA; FP64 results agree to rounding. */
typedef struct {
  int pphh_small_tile_max;  /* channels with n <= this use 64x64 pp/hh tiles, larger ones 128x128
                               (default 64) */
  int pphh_order;           /* 0: pp/hh tiles by decreasing k-length; 1: row-block groups balanced over XCDs;
                               2: 8x8-tile super-tiles balanced over XCDs (shared operand blocks in one L2) */
  int mixed_pphh_fp32;      /* IMSRG_PREC_MIXED: pp/hh GEMM on 2 = bf16 MFMA with three-way split
                               operands (default; channels on 64x64 tiles stay FP64), 1 = FP32 MFMA
                               on FP32-rounded operands, 0 = FP64 MFMA (IMSRG_PREC_SPLIT ignores both
                               flags) */
  int mixed_ph_fp32;        /* IMSRG_PREC_MIXED: ph GEMM, same choices (default 1, which also keeps
                               Zbar in FP32: its time is set by the Zbar it writes and its per-tile
                               costs, and the bf16 kernel, which writes FP64 Zbar, is slower) */
  int ph_tile_128;          /* 1: ph GEMM on 128x128 tiles, 0: on 64x64 tiles */
  int finish_order;         /* tile pairs of the last kernel (fold, comm122, mirror): 0 in channel and
                               row-block order, 1 cut into 8 runs of consecutive tiles, one per XCD */
  int overlap_pphh;         /* pp/hh GEMM (then comm122) on a second stream of the context: 0 no (one
                               stream); 1 forked before the first kernel, beside the one-body, Pandya,
                               ph GEMM and inverse Pandya kernels; 2 forked after the ph GEMM, beside
                               the inverse Pandya kernel only */
  double us_per_commutator; /* autotune: measured time of the chosen setting (0 when set by hand) */
} imsrg_tuning;
int imsrg_hip_get_tuning(imsrg_ctx *ctx, imsrg_tuning *t);
int imsrg_hip_set_tuning(imsrg_ctx *ctx, const imsrg_tuning *t);
/* Times every distinct setting on seeded synthetic operators of ctx's basis
 * (one warm-up and eight timed commutators each, at `precision`) and applies
 * the fastest; the default setting is timed first and another replaces the
 * best so far only when more than 2% faster.  With cache_path != NULL, a setting stored there for the same
 * basis, precision and GPU architecture is applied without timing, and a new
 * result is appended (one text line per entry). */
int imsrg_hip_autotune(imsrg_ctx *ctx, int precision, const char *cache_path, imsrg_tuning *out);

/* Per-phase timing of commutators on ctx (microseconds, HIP events):
 * [0] zero-body, [1] one-body, [2] comm122 (coefficients, and with the pp/hh
 * overlap its upper-triangle pass), [3] pp/hh GEMM, [4] Pandya gather, [5] ph
 * GEMM, [6] inverse Pandya, [7] last kernel (fold of the inverse-Pandya terms,
 * comm122 when not overlapped, mirror).  Enabled by imsrg_hip_set_profiling:
 * each commutator records HIP events on the stream its phases run on (the
 * pp/hh GEMM's own stream when it overlaps the others) without a host
 * synchronisation; imsrg_hip_phase_times returns the average over the
 * commutators profiled since its previous call (0 if none) and resets it. */
int imsrg_hip_set_profiling(imsrg_ctx *ctx, int on);
int imsrg_hip_phase_times(imsrg_ctx *ctx, double *us, int n);

/* Diagnostic with no counterpart in the reference: out1024[0..255] = D of one
 * v_mfma_f64_16x16x4f64 with A = lane id, B = 1; out1024[256..511] with A = 1,
 * B = lane id (lane l, register q at index 4l+q); out1024[512..1023] the same
 * two products on v_mfma_f32_16x16x4f32.  Used by the tests to pin the MFMA
 * lane layouts the GEMM kernels assume. */
int imsrg_hip_selftest_mfma(imsrg_ctx *ctx, double *out1024);
/* Diagnostic with no counterpart in the reference: one v_mfma_f32_16x16x32_bf16
 * with C = 0 whose lane l takes operand slots j = 0..7 from a512[8l + j] and
 * b512[8l + j] (rounded to bf16); d256[4l + q] receives register q of lane l. */
int imsrg_hip_selftest_mfma_bf16(imsrg_ctx *ctx, const float *a512, const float *b512, float *d256);

#ifdef __cplusplus
}
#endif
#endif
