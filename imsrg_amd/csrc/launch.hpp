// launch.hpp -- host-callable launchers of the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace imsrg {

// dense 6j table {jj1 jj2 J3; jj4 jj5 J6} (double-double Racah): `table` in the layout of
// SixJCache_112112, `tableQ` in the sixjQ_index layout as the weights (2 J6 + 1) {6j},
// 0 in place of |6j| < 1e-8 (padding entries are left untouched)
hipError_t launch_sixj_table(double *table, double *tableQ, int d1, int d2, int d2q, hipStream_t s);

// commutator phases; items are device arrays of int4 work items
hipError_t launch_zerobody(const PairArg &pa, int npairs, int maxnch, hipStream_t s);
// one-body terms: comm111 + comm121 per pair (items), comm221 in partial sums
// (items221, none when that term is off), then the sums into Z1
hipError_t launch_onebody(const PairArg &pa, const int4 *items, int nitems, const int4 *items221, int n221,
                          int npairs, int maxn1pairs, hipStream_t s);
// pp/hh SYR2K: items on 128x128 tiles, items64 on 64x64 tiles (disjoint channels);
// under bf16x3 products (kGemmB3) the pack items first write the operand panels
// maxn: the largest channel dimension of the pairs (sizes the kernel's LDS index list)
hipError_t launch_pphh(const PairArg &pa, const int4 *items, int nitems, const int4 *items64, int nitems64,
                       const int4 *pack, int npack, int precision, int maxn, hipStream_t s);
// comm122 partner coefficients of this call (used by launch_finish with c122)
hipError_t launch_c122_coef(const PairArg &pa, int npairs, long maxn122, hipStream_t s);
// maxnch: the largest channel count of the pairs (sizes the kernel's LDS channel table)
hipError_t launch_pandya(const PairArg &pa, const int4 *items, int nitems, int maxnch, hipStream_t s);
// ph GEMM (head and tail of Zbar) on 128x128 tiles (big) or 64x64 tiles; precision 1
// (FP32 products) stores Zbar in FP32, the others in FP64
hipError_t launch_ph_gemm(const PairArg &pa, const int4 *items, int nitems, int precision, bool big, hipStream_t s);
// jmax, norb, nch, ncls: the largest of the pairs (they size the kernel's LDS tables);
// zbar_f32: Zbar holds FP32 elements (written by the precision-1 ph GEMM)
hipError_t launch_invpandya(const PairArg &pa, const int4 *items, int nitems, int jmax, int norb, int nch,
                            int ncls, bool zbar_f32, hipStream_t s);
// last phase: with fold, upper += the inverse-Pandya contributions stored at the
// transposed slots; with c122, Z2 -= W + hZ W^T (comm122ss); then lower triangle
// <- hZ * upper
hipError_t launch_finish(const PairArg &pa, const int4 *items, int nitems, bool fold, bool c122, hipStream_t s);
// comm122's W (two-body layout, P.c122w) from the partner coefficients of this
// call, by column bins; before launch_finish with c122
hipError_t launch_c122_w(const PairArg &pa, const int4 *items, int nitems, int maxbin, int maxent, hipStream_t s);
hipError_t launch_zero_two(const PairArg &pa, const int4 *items, int nitems, hipStream_t s);

// operator arithmetic
hipError_t launch_axpy(double *y, double a, const double *x, long n, hipStream_t s);
hipError_t launch_scale(double *y, double a, long n, hipStream_t s);
hipError_t launch_fill(double *y, double v, long n, hipStream_t s);  // y[i] = v
// Operator::DoNormalOrdering2 / UndoNormalOrdering of a scalar operator in place
// (sign +1 / -1); n2sq: the squared two-body norm on the device (launch_norm's second
// value): the one-body sum over the holes runs when the norm exceeds 1e-7, the
// (anti-)Hermitian completion of the one-body part always
hipError_t launch_normal_order(const BasisDev *bd, const BasisDev &bh, double *op, int sign, int herm,
                               const double *n2sq, hipStream_t s);
// y += a x over an operator buffer, with the norm sums of x into out[0..1] (as launch_norm)
hipError_t launch_axpy_norm(const BasisDev *bd, double *y, double a, const double *x, double *partials, int npart,
                            double *out, hipStream_t s);
// partial sums of squares; out[0] = one-body, out[1] = two-body (weighted)
hipError_t launch_norm(const BasisDev *bd, const BasisDev &bh, const double *op, double *partials, int npart,
                       double *out, hipStream_t s);
// the basis-change matrices D of every channel (two-body layout, tbsize doubles) from
// the one-body basis change C (norb x norb, column-major)
// the Hartree-Fock monopoles W[q] = sum_J (2J+1) V~^J_{ia,jb} of quads[q] = (i, a, j, b)
hipError_t launch_hf_mono(const BasisDev *bd, const double *op, const int4 *quads, int nq, double *W,
                          hipStream_t s);
// block-sparse basis transformation (C block diagonal in (l, j2, tz2)): the D blocks
// into Dg, T = V D (items (group, first row), R rows each), out two-body = D^T T (items
// (channel, first column), kc columns each), out one-body = C^T O C (through one_tmp)
hipError_t launch_xform_blocks(const BasisDev *bd, const BasisDev &bh, const XformDev &X, const double *C,
                               const double *op, double *out, double *T, double *Dg, double *one_tmp,
                               const int2 *right_items, int nright, int R, const int2 *left_items, int nleft, int kc,
                               hipStream_t s);
hipError_t launch_hf_D(const BasisDev *bd, const BasisDev &bh, const double *C, double *D, hipStream_t s);
hipError_t launch_fill_random(const BasisDev *bd, const BasisDev &bh, double *op, int herm,
                              unsigned long long seed, hipStream_t s);
hipError_t launch_white(const BasisDev *bd, const BasisDev &bh, const double *H, double *Eta, const int4 *gen_items,
                        int ngen, hipStream_t s);

// self-test of the MFMA f64 and f32 lane layouts: out[1024] (see imsrg_hip_selftest_mfma)
hipError_t launch_mfma_probe(double *out, hipStream_t s);
// one v_mfma_f32_16x16x32_bf16 on host-given operand slots (a, b: 512 floats, d: 256)
hipError_t launch_mfma_bf16_probe(const float *a, const float *b, float *d, hipStream_t s);

}  // namespace imsrg
