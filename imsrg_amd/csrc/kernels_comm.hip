// kernels_comm.hip -- HIP kernels of the scalar IMSRG(2) commutator Z = [X, Y]
// on CDNA4 (gfx950).  Every kernel reads work items (int4: pair, a, b, c)
// so that one launch can cover all channels of one or many operator pairs.
//
// Two-body accumulation scheme: for Hermitian / anti-Hermitian Z every term
// adds only to the upper triangle (row <= col) of each channel block; the
// mirror kernel writes the lower triangle as hZ * upper at the end.  This is
// exact because each term of the reference produces a contribution whose
// (s,r) element is hZ times its (r,s) element (src/Commutator.cc:929, 1185-1194,
// 1611-1612).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "internal.hpp"
#include "launch.hpp"

namespace imsrg {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

// Column (within a 16x16 block) held by register q of a lane in the D^T-block
// formulation of the GEMM kernels: f64 16x16x4 puts D[m][n] in lane n+16(m%4),
// register m/4; f32 16x16x4 in lane n+16(m/4), register m%4.
template <bool MIXED>
__device__ __forceinline__ int mfma_col(int lane, int q) {
  return MIXED ? 4 * (lane >> 4) + q : 4 * q + (lane >> 4);
}

#define SQRT2_D 1.4142135623730951

__device__ __forceinline__ double X1at(const double *one, int norb, int p, int q) { return one[(long)q * norb + p]; }

// Normalized scalar TBME <ab|V|cd>_J with the swapped-ket phases of
// TwoBodyME::GetTBME_norm (src/TwoBodyME.cc:128-147).
__device__ __forceinline__ double tbme_norm(const BasisDev &B, const double *two, int J, int a, int b, int c, int d) {
  int pab = (B.l[a] + B.l[b]) & 1, pcd = (B.l[c] + B.l[d]) & 1;
  if (pab != pcd) return 0.0;
  int tab = (B.tz2[a] + B.tz2[b]) / 2, tcd = (B.tz2[c] + B.tz2[d]) / 2;
  if (tab != tcd || J > B.jmax) return 0.0;
  int k1 = ket_index(B.norb, min(a, b), max(a, b));
  int k2 = ket_index(B.norb, min(c, d), max(c, d));
  int i = B.kmap[(long)k1 * (B.jmax + 1) + J];
  int k = B.kmap[(long)k2 * (B.jmax + 1) + J];
  if (i < 0 || k < 0) return 0.0;
  int ch = 6 * J + 2 * (tab + 1) + pab;
  double ph = 1.0;
  if (a > b) ph *= sgn_parity((B.j2[a] + B.j2[b]) / 2 + 1 + J);
  if (c > d) ph *= sgn_parity((B.j2[c] + B.j2[d]) / 2 + 1 + J);
  return ph * two[B.ch_off[ch] + (long)k * B.ch_n[ch] + i];
}

// TwoBodyME::GetTBMEmonopole for a scalar operator (src/TwoBodyME.cc:615-682)
__device__ double tbme_monopole(const BasisDev &B, const double *two, int a, int b, int c, int d) {
  double nrm = (a == b ? SQRT2_D : 1.0) * (c == d ? SQRT2_D : 1.0);
  int jmin = abs(B.j2[a] - B.j2[b]) / 2, jmax = (B.j2[a] + B.j2[b]) / 2;
  double mon = 0.0;
  for (int J = jmin; J <= jmax; ++J) mon += (2 * J + 1) * tbme_norm(B, two, J, a, b, c, d);
  mon /= (B.j2[a] + 1) * (B.j2[b] + 1);
  return nrm * mon;
}

template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV]) {
  __shared__ double red[NV][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double x = v[q];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if (lane == 0) red[q][w] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = (red[q][0] + red[q][1]) + (red[q][2] + red[q][3]);
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// comm110ss (src/Commutator.cc:564-581) and comm220ss (601-631) -> Z0.
// Grid (nch + 1, npairs): block 0 clears Z's one-body block and forms the
// comm110 sum, block 1+ch the comm220 contribution of channel ch; partial
// sums go to scratch and k_zerobody_final adds them in channel order.
__global__ __launch_bounds__(256) void k_zerobody(PairArg pa) {
  PairPtrs P = pa.get(blockIdx.y);
  const BasisDev &B = *P.basis;
  const int norb = B.norb;
  const int bid = blockIdx.x;
  if (bid > B.nch) return;
  double acc[1] = {0.0};
  const bool skip = (P.hX == P.hY);
  if (bid == 0) {
    const double *X1 = P.X + B.one_off, *Y1 = P.Y + B.one_off;
    double *Z1 = P.Z + B.one_off;
    for (int e = threadIdx.x; e < norb * norb; e += blockDim.x) Z1[e] = 0.0;
    if (!skip && (P.mask & M110)) {
      for (int t = threadIdx.x; t < B.nholes * norb; t += blockDim.x) {
        int a = B.holes[t / norb], k = t % norb;
        double xy = X1at(X1, norb, a, k) * X1at(Y1, norb, k, a);
        double yx = X1at(Y1, norb, a, k) * X1at(X1, norb, k, a);
        acc[0] += (B.j2[a] + 1) * B.occ[a] * (xy - yx);
      }
    }
  } else if (!skip && (P.mask & M220)) {
    const int ch = bid - 1;
    const int h0 = B.hh_start[ch], nhh = B.hh_start[ch + 1] - h0;
    if (nhh > 0) {
      const int n = B.ch_n[ch];
      const long off = B.ch_off[ch];
      const double *X2 = P.X + B.two_off + off, *Y2 = P.Y + B.two_off + off;
      const double *w = B.w220 + B.ch_kstart[ch];
      const double f = 2.0 * (2 * (ch / 6) + 1);
      // X(h,m) = hX X(m,h): both operands read down column h
      for (int t = threadIdx.x; t < nhh * n; t += blockDim.x) {
        int ih = t / n, m = t % n;
        int h = B.hh_loc[h0 + ih];
        acc[0] += f * B.hh_occ[h0 + ih] * ((P.hX * X2[(long)h * n + m]) * (Y2[(long)h * n + m] * w[m]));
      }
    }
  }
  block_reduce<1>(acc);
  if (threadIdx.x == 0) P.scratch[bid] = acc[0];
}

__global__ void k_zerobody_final(PairArg pa) {
  PairPtrs P = pa.get(blockIdx.x);
  if (threadIdx.x != 0) return;
  const BasisDev &B = *P.basis;
  const int nch = B.nch;
  double s220 = 0.0;
  // the channels' partial sums added in channel order (the order fixes the
  // result), loaded in groups of 16 issued before the group's adds: one thread
  // waiting on each load in turn took 30 us at emax=10
  for (int c0 = 0; c0 < nch; c0 += 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = P.scratch[1 + min(c0 + u, nch - 1)];
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (c0 + u < nch) s220 += v[u];
  }
  P.Z[0] = (0.0 + P.scratch[0]) + s220;
}

// comm122 coefficients of this call: x = X1(a,o) f, y = Y1(a,o) f for every
// partner entry (src/Commutator.cc:857-896).
__global__ __launch_bounds__(256) void k_c122_coef(PairArg pa, int npairs) {
  PairPtrs P = pa.get(blockIdx.y);
  const BasisDev &B = *P.basis;
  const int norb = B.norb;
  const double *X1 = P.X + B.one_off, *Y1 = P.Y + B.one_off;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < B.n122; e += (long)gridDim.x * blockDim.x) {
    const int2 ao = B.c122_ao[e];
    const double f = B.c122_f[e];
    P.c122[e] = make_double2(X1at(X1, norb, ao.x, ao.y) * f, X1at(Y1, norb, ao.x, ao.y) * f);
  }
}

// ---------------------------------------------------------------------------
// One-body terms for one (i, j) pair of the same (l, j2, tz2) block, i <= j:
// comm111ss (src/Commutator.cc:644-649), comm121ss (670-728) and the one-body
// part of comm222_pp_hh_221ss (1246-1294).  M_pp / M_hh elements are formed
// directly as folded dot products, M(r,s) + hZ M(s,r) with M = X W Y.
__device__ __forceinline__ double fold_dot(const double *xr, const double *xs, const double *yr, const double *ys,
                                           const int *loc, const double *w, int nk, int start, int stride,
                                           double hZ) {
  // the two products summed apart, as the reference's M and its transpose: for
  // r == s and hZ = -1 they cancel exactly (an anti-Hermitian diagonal is 0)
  double m1 = 0.0, m2 = 0.0;
  for (int q = start; q < nk; q += stride) {
    const int k = loc[q];
    m1 += w[q] * (xr[k] * ys[k]);
    m2 += w[q] * (xs[k] * yr[k]);
  }
  return m1 + hZ * m2;
}

__global__ __launch_bounds__(256) void k_onebody(PairArg pa, const int4 *items) {
  const int4 it = items[blockIdx.x];
  PairPtrs P = pa.get(it.x);
  const BasisDev &B = *P.basis;
  const int norb = B.norb, i = it.y, j = it.z;
  const double *X1 = P.X + B.one_off, *Y1 = P.Y + B.one_off;
  const double *X2 = P.X + B.two_off, *Y2 = P.Y + B.two_off;
  // (X1 Y1)(i,j) and (Y1 X1)(i,j) summed apart, as the reference's two matrix
  // products: commuting blocks then give exactly 0
  double acc[3] = {0.0, 0.0, 0.0};
  if (P.mask & M111) {
    for (int k = threadIdx.x; k < norb; k += blockDim.x) {
      acc[0] += X1at(X1, norb, i, k) * X1at(Y1, norb, k, j);
      acc[2] += X1at(Y1, norb, i, k) * X1at(X1, norb, k, j);
    }
  }
  if (P.mask & M121) {
    // one task per (hole a, partner b, monopole term)
    const int ntask = B.nholes * B.maxblk * 4;
    for (int t = threadIdx.x; t < ntask; t += blockDim.x) {
      const int term = t & 3, ib = (t >> 2) % B.maxblk, ia = (t >> 2) / B.maxblk;
      const int a = B.holes[ia];
      if (ib >= B.obstart[a + 1] - B.obstart[a]) continue;
      const int b = B.oblist[B.obstart[a] + ib];
      const double nanb = B.occ[a] * (1.0 - B.occ[b]);
      if (fabs(nanb) < 1e-6) continue;
      double v;
      if (term == 0) v = (B.j2[b] + 1) * nanb * X1at(X1, norb, a, b) * tbme_monopole(B, Y2, b, i, a, j);
      else if (term == 1) v = -(B.j2[a] + 1) * nanb * X1at(X1, norb, b, a) * tbme_monopole(B, Y2, a, i, b, j);
      else if (term == 2) v = -(B.j2[b] + 1) * nanb * X1at(Y1, norb, a, b) * tbme_monopole(B, X2, b, i, a, j);
      else v = (B.j2[a] + 1) * nanb * X1at(Y1, norb, b, a) * tbme_monopole(B, X2, a, i, b, j);
      acc[1] += v;
    }
  }
  block_reduce<3>(acc);
  if (threadIdx.x == 0) P.obpart[B.ob_start[it.w]] = (acc[0] - acc[2]) + acc[1];
}

// Lane's share of (2J+1) fac M(ci, cj) for the comm221 one-body part
// (src/Commutator.cc:1246-1294): M = M_pp (pp = true, ConstructScalarMpp_Mhh's
// pp list) or M_hh, formed as the folded dot product M(r,s) + hZ M(s,r) over
// the list entries start, start + stride, ...; 0 when a ket is missing.
__device__ __forceinline__ double m221_part(const BasisDev &B, const double *X2, const double *Y2, int i, int j,
                                            int c, int J, bool pp, int start, int stride, double hZ) {
  const int norb = B.norb, nJ = B.jmax + 1;
  const int kci = ket_index(norb, min(c, i), max(c, i)), kcj = ket_index(norb, min(c, j), max(c, j));
  const int r = B.kmap[(long)kci * nJ + J], s = B.kmap[(long)kcj * nJ + J];
  if (r < 0 || s < 0) return 0.0;
  const int ch = 6 * J + 2 * ((B.tz2[c] + B.tz2[i]) / 2 + 1) + ((B.l[c] + B.l[i]) & 1);
  const int n = B.ch_n[ch];
  const long off = B.ch_off[ch];
  double fac = (c == i ? SQRT2_D : 1.0) * (c == j ? SQRT2_D : 1.0);
  if (c > i) fac *= sgn_parity((B.j2[c] + B.j2[i]) / 2 + 1 + J);
  if (c > j) fac *= sgn_parity((B.j2[c] + B.j2[j]) / 2 + 1 + J);
  const double *xr = X2 + off + (long)r * n, *xs = X2 + off + (long)s * n;
  const double *yr = Y2 + off + (long)r * n, *ys = Y2 + off + (long)s * n;
  const int k0 = pp ? B.kpp_start[ch] : B.khh_start[ch], k1 = pp ? B.kpp_start[ch + 1] : B.khh_start[ch + 1];
  const double m = pp ? fold_dot(xr, xs, yr, ys, B.kpp_loc + k0, B.kpp_w + k0, k1 - k0, start, stride, hZ)
                      : fold_dot(xr, xs, yr, ys, B.khh_loc + k0, B.khh_w + k0, k1 - k0, start, stride, hZ);
  return (2 * J + 1) * fac * m;
}

// comm221 one-body part in partial sums, one slot per item (it.w):
//  kind 0 (n_c != 0): n_c M_pp for one (c, J), the block's lanes over the pp list;
//  kind 1 (1 - n_c != 0): (1 - n_c) M_hh for 4 orbits c of the list cpart
//  (one per wave), all J, the wave's lanes over the hh list.
__global__ __launch_bounds__(256) void k_onebody221(PairArg pa, const int4 *items) {
  const int4 it = items[blockIdx.x];
  PairPtrs P = pa.get(it.x);
  const BasisDev &B = *P.basis;
  const int2 pr = B.pairs1[it.y];
  const int i = pr.x, j = pr.y;
  const double *X2 = P.X + B.two_off, *Y2 = P.Y + B.two_off;
  const double hZ = P.hZ;
  double acc[1] = {0.0};
  if (!(it.z >> 30)) {
    const int c = (it.z >> 8) & 0xffff, J = it.z & 0xff;
    acc[0] = B.occ[c] * m221_part(B, X2, Y2, i, j, c, J, true, threadIdx.x, blockDim.x, hZ);
  } else {
    const int q = (it.z & 0x3fffffff) + (threadIdx.x >> 6);
    if (q < B.ncpart) {
      const int c = B.cpart[q], lane = threadIdx.x & 63;
      const int Jmin = max(abs(B.j2[c] - B.j2[i]), abs(B.j2[c] - B.j2[j])) / 2;
      const int Jmax = (B.j2[c] + min(B.j2[i], B.j2[j])) / 2;
      double m = 0.0;
      for (int J = Jmin; J <= Jmax; ++J) m += m221_part(B, X2, Y2, i, j, c, J, false, lane, 64, hZ);
      acc[0] = (1.0 - B.occ[c]) * m;
    }
  }
  block_reduce<1>(acc);
  if (threadIdx.x == 0) P.obpart[it.w] = P.hX * acc[0];
}

// Z1(i, j) = comm111 + comm121 + comm221 / (j_i + 1) from the partial sums of
// one-body pair p, one wave per pair (lane-strided sums and a fixed shuffle
// tree: the same order every call), and Z1(j, i) = hZ Z1(i, j).
__global__ __launch_bounds__(256) void k_onebody_final(PairArg pa) {
  PairPtrs P = pa.get(blockIdx.y);
  const BasisDev &B = *P.basis;
  const int p = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (p >= B.n1pairs) return;  // wave-uniform
  const int2 pr = B.pairs1[p];
  const int i = pr.x, j = pr.y, s0 = B.ob_start[p], s1 = B.ob_start[p + 1];
  double z221 = 0.0;
  if (P.mask & M221)
    for (int s = s0 + 1 + lane; s < s1; s += 64) z221 += P.obpart[s];
  for (int o = 32; o > 0; o >>= 1) z221 += __shfl_down(z221, o, 64);
  if (lane != 0) return;
  const double z = P.obpart[s0] + z221 / (B.j2[i] + 1.0);
  double *Z1 = P.Z + B.one_off;
  Z1[(long)j * B.norb + i] = z;
  if (i != j) Z1[(long)i * B.norb + j] = P.hZ * z;
}

// ---------------------------------------------------------------------------
// SYR2K tile on MFMA f64 16x16x4:
//   D(r,s) = alpha * cph(s) * sum_k w_k [ A(r,ka_k) B(s,kb_k) + h B(r,kb_k) A(s,ka_k) ],  r <= s
// with A, B column-major n x * matrices (A(r,k) = A[k*n + r]).
// MODE 0: pp/hh term of comm222_pp_hh_221ss (src/Commutator.cc:1171-1194, 1237-1238):
//         A = X2, B = Y2 of one channel, k over pp/hh kets with weight w_pp - w_hh,
//         alpha = hY, h = hZ.  (X W Y + hZ (X W Y)^T written with Y^T = hY Y.)
// MODE 1: head of Zbar in comm222_phss (src/Commutator.cc:1789-1799): A = Xbar,
//         B = Ybar^T, h = +1 (Z Hermitian) or -1.
// MODE 2: tail of Zbar (src/Commutator.cc:1788-1789, 1805): B columns permuted
//         by ab <-> ba with weights hY*phi(ph ket), cph(s) = phi(s), h = +1.
//
// MIXED: operands rounded to FP32 in LDS, products on f32 16x16x4 MFMA, FP32
// partial sums over kF32Flush k-steps added into the FP64 accumulators.
template <int MODE, bool MIXED>
__global__ __launch_bounds__(256) void k_syr2k(PairArg pa, const int4 *items, int accumulate) {
  constexpr int LD = kTile + 2;
  using ST = typename std::conditional<MIXED, float, double>::type;
  // the FP32-product ph GEMM keeps Zbar in FP32 (its inverse Pandya gather reads
  // 4-byte elements); the FP64 kernels keep it in FP64
  using ZT = typename std::conditional<MIXED && MODE != 0, float, double>::type;
  __shared__ ST As[2][kTileK][LD];
  __shared__ ST Bs[2][kTileK][LD];
  const int4 it = items[blockIdx.x];
  PairPtrs P = pa.get(it.x);
  const BasisDev &B = *P.basis;
  const double *Am, *Bm;
  int n, nk, nph = 0;
  const int *kl = nullptr;
  const double *kw = nullptr;
  const int *phl = nullptr;
  const signed char *phi = nullptr;
  double alpha, h;
  double *D = nullptr;
  ZT *Dz = nullptr;
  if (MODE == 0) {
    const int ch = it.y;
    n = B.ch_n[ch];
    const long off = B.ch_off[ch];
    Am = P.X + B.two_off + off;
    Bm = P.Y + B.two_off + off;
    const int k0 = B.k_start[ch];
    nk = B.k_start[ch + 1] - k0;
    kl = B.k_loc + k0;
    kw = B.k_w + k0;
    alpha = (double)P.hY;
    h = (double)P.hZ;
    D = P.Z + B.two_off + off;
  } else {
    const int cc = it.y;
    n = B.cc_n[cc];
    nph = B.cc_nph[cc];
    nk = 2 * nph;
    const long xo = B.cc_xoff[cc];
    Am = P.Xt + xo;
    Bm = P.Yt + xo;
    Dz = reinterpret_cast<ZT *>(P.Zbar) + B.cc_zoff[cc] + (MODE == 2 ? (long)n * n : 0);
    alpha = 1.0;
    h = (MODE == 1) ? (P.hZ == 1 ? 1.0 : -1.0) : 1.0;
    phi = B.cc_phi + B.cc_kstart[cc];
    phl = B.cc_ph + B.cc_phstart[cc];
  }
  const int R0 = it.z * kTile, S0 = it.w * kTile;
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wr = (wv & 1) * 32, wc = (wv >> 1) * 32;
  // 16x16 blocks with upper-triangle elements inside the matrix (the others issue no MFMA)
  unsigned bm = 0;
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni) {
      const int r0 = R0 + wr + mi * 16, c0 = S0 + wc + ni * 16;
      if (r0 < n && c0 < n && r0 <= c0 + 15) bm |= 1u << (mi * 2 + ni);
    }
  dbl4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};

  const int tk = threadIdx.x >> 4, tr = (threadIdx.x & 15) * 4;
  const int K2 = 2 * nk;
  // global -> register loads of one k-stage (16 values of k' per stage)
  auto load_stage = [&](int k0, double (&av)[4], double (&bv)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) av[q] = bv[q] = 0.0;
    const int kk = k0 + tk;
    if (kk < K2) {
      const int k = kk < nk ? kk : kk - nk;
      int ca, cb;
      double wk;
      if (MODE == 0) {
        ca = cb = kl[k];
        wk = kw[k];
      } else if (MODE == 1) {
        ca = cb = k;
        wk = 1.0;
      } else {
        ca = k;
        cb = k < nph ? k + nph : k - nph;
        wk = (double)P.hY * (double)phi[phl[k < nph ? k : k - nph]];
      }
      const double *sa, *sb;
      double fa;
      if (kk < nk) {
        sa = Am + (long)ca * n;
        sb = Bm + (long)cb * n;
        fa = wk;
      } else {
        sa = Bm + (long)cb * n;
        sb = Am + (long)ca * n;
        fa = wk * h;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = R0 + tr + q, c = S0 + tr + q;
        av[q] = (r < n) ? sa[r] * fa : 0.0;
        bv[q] = (c < n) ? sb[c] : 0.0;
      }
    }
  };
  double av[4], bv[4];
  load_stage(0, av, bv);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    As[0][tk][tr + q] = av[q];
    Bs[0][tk][tr + q] = bv[q];
  }
  __syncthreads();
  int cur = 0;
  for (int k0 = 0; k0 < K2; k0 += kTileK) {
    const bool more = k0 + kTileK < K2;
    if (more) load_stage(k0 + kTileK, av, bv);  // in flight during the MFMAs below
    if constexpr (MIXED) {
      flt4 t[2][2];
#pragma unroll
      for (int ks = 0; ks < kTileK / 4; ++ks) {
        const int kr = ks * 4 + (lane >> 4);
        const float a0 = As[cur][kr][wr + (lane & 15)], a1 = As[cur][kr][wr + 16 + (lane & 15)];
        const float b0 = Bs[cur][kr][wc + (lane & 15)], b1 = Bs[cur][kr][wc + 16 + (lane & 15)];
        const bool fresh = ks % kF32Flush == 0;
        const flt4 fz = {0.f, 0.f, 0.f, 0.f};
        if (bm & 1) t[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(b0, a0, fresh ? fz : t[0][0], 0, 0, 0);
        if (bm & 2) t[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(b1, a0, fresh ? fz : t[0][1], 0, 0, 0);
        if (bm & 4) t[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(b0, a1, fresh ? fz : t[1][0], 0, 0, 0);
        if (bm & 8) t[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(b1, a1, fresh ? fz : t[1][1], 0, 0, 0);
        if (ks % kF32Flush == kF32Flush - 1) {
#pragma unroll
          for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
#pragma unroll
              for (int q = 0; q < 4; ++q) acc[mi][ni][q] += (double)t[mi][ni][q];
        }
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < kTileK / 4; ++ks) {
        const int kr = ks * 4 + (lane >> 4);
        const double a0 = As[cur][kr][wr + (lane & 15)], a1 = As[cur][kr][wr + 16 + (lane & 15)];
        const double b0 = Bs[cur][kr][wc + (lane & 15)], b1 = Bs[cur][kr][wc + 16 + (lane & 15)];
        // D^T blocks: lane l, register q holds D row (lane&15), column mfma_col(l, q)
        if (bm & 1) acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a0, acc[0][0], 0, 0, 0);
        if (bm & 2) acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a0, acc[0][1], 0, 0, 0);
        if (bm & 4) acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a1, acc[1][0], 0, 0, 0);
        if (bm & 8) acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a1, acc[1][1], 0, 0, 0);
      }
    }
    if (more) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        As[cur ^ 1][tk][tr + q] = av[q];
        Bs[cur ^ 1][tk][tr + q] = bv[q];
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  if (MODE == 0) {
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int ni = 0; ni < 2; ++ni)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = R0 + wr + mi * 16 + (lane & 15);
          const int col = S0 + wc + ni * 16 + mfma_col<MIXED>(lane, q);
          if (row < n && col < n && row <= col) {
            const double v = alpha * acc[mi][ni][q];
            double *dp = D + (long)col * n + row;
            *dp = accumulate ? *dp + v : v;
          }
        }
  } else {
    // Zbar blocks are stored in full: the upper triangle as computed, the lower
    // one as (head) hZ-sign or (tail) phi_r phi_s times the upper element, as
    // the reference's folds make them (src/Commutator.cc:1792-1805).
    __shared__ double Tt[kTile][kTile + 1];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int ni = 0; ni < 2; ++ni)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int rr = wr + mi * 16 + (lane & 15), cc = wc + ni * 16 + mfma_col<MIXED>(lane, q);
          double v = alpha * acc[mi][ni][q];
          if (MODE == 2) {
            const int col = S0 + cc;
            v *= (col < n) ? (double)phi[col] : 0.0;
          }
          Tt[rr][cc] = v;
        }
    __syncthreads();
    for (int e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
      const int rr = e & (kTile - 1), cc = e / kTile;
      const int row = R0 + rr, col = S0 + cc;
      if (row < n && col < n && row <= col) Dz[(long)col * n + row] = (ZT)Tt[rr][cc];
      // mirrored element (S0+rr, R0+cc) from the upper element (R0+cc, S0+rr)
      const int row2 = S0 + rr, col2 = R0 + cc;
      if (row2 < n && col2 < n && row2 > col2) {
        const double sym = (MODE == 1) ? h : (double)(phi[row2] * phi[col2]);
        Dz[(long)col2 * n + row2] = (ZT)(sym * Tt[cc][rr]);
      }
    }
  }
}


// ---------------------------------------------------------------------------
// SYR2K on 128x128 output tiles: each of the 4 waves owns a 64x64 sub-tile
// (4x4 MFMA blocks), so one LDS read feeds four MFMAs; k-stages of 16 are
// double-buffered in LDS and filled by direct global->LDS copies issued one
// stage ahead.
//   D(r,s) = alpha * cph(s) * sum_k w_k [ A(r,ka_k) B(s,kb_k) + h B(r,kb_k) A(s,ka_k) ],  r <= s
// MODE 0: pp/hh term of comm222_pp_hh_221ss (src/Commutator.cc:1171-1194, 1237-1238):
//         A = X2, B = Y2 of one channel, k over pp/hh kets with weight w_pp - w_hh,
//         alpha = hY, h = hZ; the upper triangle of Z2 is written.
// MODE 1: head of Zbar in comm222_phss (src/Commutator.cc:1789-1799): A = Xbar,
//         B = Ybar^T, h = +1 (Z Hermitian) or -1.
// MODE 2: tail of Zbar (src/Commutator.cc:1788-1789, 1805): B columns permuted
//         by ab <-> ba with weights hY*phi(ph ket), cph(s) = phi(s), h = +1.
// MODE 1/2 store Zbar in full: the lower triangle as (head) hZ-sign or (tail)
// phi_r phi_s times the upper element, as the reference's folds make it
// (src/Commutator.cc:1792-1805), written through an LDS transpose.
// MIXED: operands rounded to FP32 when read from LDS, products on f32
// 16x16x4 MFMA, FP32 partial sums over two stages (32 k) added into the FP64
// accumulators.
// SPLIT (with MIXED): each operand is split into two FP32 values, a = ah + al
// with ah = fl32(a), al = fl32(a - ah), and a product takes three f32 MFMAs,
// ah*bh + ah*bl + al*bh (the dropped al*bl is 2^-48 of it); the FP32 partial
// sums are flushed into FP64 after every 4 k.
// Non-temporal stores of Zbar: the inverse Pandya gather reads it only after
// the whole ph GEMM, so keeping its lines out of the L2 spares that gather
// the write-back of dirty lines (inverse Pandya 880 -> 835 us, ph GEMM
// 398 -> 437 us: a net 1% on the overlapped commutator).  The last kernel's
// stores of Z use them too (Z is read again only by a later call).
__device__ __forceinline__ void st_nt(double *p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt2(double *p, double a, double b) {
  typedef double d2v __attribute__((ext_vector_type(2)));
  const d2v v = {a, b};
  __builtin_nontemporal_store(v, reinterpret_cast<d2v *>(p));
}
// FP32 Zbar (the FP32-product ph GEMM): the values rounded once, when stored
__device__ __forceinline__ void st_nt(float *p, double v) { __builtin_nontemporal_store((float)v, p); }
__device__ __forceinline__ void st_nt2(float *p, double a, double b) {
  typedef float f2s __attribute__((ext_vector_type(2), aligned(4)));
  const f2s v = {(float)a, (float)b};
  __builtin_nontemporal_store(v, reinterpret_cast<f2s *>(p));
}
__device__ __forceinline__ void st_nt4(float *p, double a, double b, double c, double d) {
  typedef float f4s __attribute__((ext_vector_type(4), aligned(4)));  // 4-byte aligned rows
  const f4s v = {(float)a, (float)b, (float)c, (float)d};
  __builtin_nontemporal_store(v, reinterpret_cast<f4s *>(p));
}

// A wave of a 128x128 tile owns the 64 x kGW sub-tile at rows (w & 1) * 64,
// columns (w >> 1) * kGW; it has work when that sub-tile holds an upper-
// triangle element inside the n x n matrix.
__device__ __forceinline__ bool tile_wave_live(int R0, int S0, int n, int w) {
  return R0 + (w & 1) * 64 < n && S0 + (w >> 1) * kGW < n && R0 + (w & 1) * 64 <= S0 + (w >> 1) * kGW + kGW - 1;
}

// Stores a 128x128 tile from the waves' FP64 accumulators (lane layout of the
// f32 MFMAs when F32L, else of the f64 16x16x4 MFMA).
// MODE 0: the upper triangle of D = alpha * acc.
// MODE 1/2: Zbar in full through LDS transposes of the 64 x kGW wave sub-tiles,
// four per round in the buffers tb0..tb3 (each >= 64 (kGW + 1) doubles): the
// lower triangle as (head) h or (tail) phi_r phi_s times the upper element, as
// the reference's folds make it (src/Commutator.cc:1792-1805); s_ph holds the
// tile's row (0) and column (1) phases.
template <int MODE, bool F32L, typename TD = double>
__device__ __forceinline__ void gemm_epilogue(const dbl4 (&acc)[4][kGN], TD *D, int n, int R0, int S0,
                                              double alpha, double h, const signed char (*s_ph)[kTileB],
                                              double *tb0, double *tb1, double *tb2, double *tb3) {
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wr = (wv & 1) * 64, wc = (wv >> 1) * kGW;
  const bool wave_live = tile_wave_live(R0, S0, n, wv);
  if constexpr (MODE == 0) {
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < kGN; ++ni)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = R0 + wr + mi * 16 + (lane & 15);
          const int col = S0 + wc + ni * 16 + mfma_col<F32L>(lane, q);
          if (row < n && col < n && row <= col) D[(long)col * n + row] = alpha * acc[mi][ni][q];
        }
  } else {
    static_assert(kGWaves == 8, "two rounds of four sub-tiles");
    auto Tb = [&](int j) {
      double *p = j == 0 ? tb0 : j == 1 ? tb1 : j == 2 ? tb2 : tb3;
      return reinterpret_cast<double(*)[kGW + 1]>(p);
    };
    for (int rd = 0; rd < 2; ++rd) {
      if ((wv >> 2) == rd && wave_live) {
        double(*Tt)[kGW + 1] = Tb(wv & 3);
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
#pragma unroll
          for (int ni = 0; ni < kGN; ++ni)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int rr = mi * 16 + (lane & 15), cc = ni * 16 + mfma_col<F32L>(lane, q);
              double v = alpha * acc[mi][ni][q];
              if (MODE == 2) {
                const int col = S0 + wc + cc;
                v *= (col < n) ? (double)s_ph[1][wc + cc] : 0.0;
              }
              Tt[rr][cc] = v;
            }
      }
      __syncthreads();
      for (int j = 0; j < 4; ++j) {
        const int w = 4 * rd + j;
        if (!tile_wave_live(R0, S0, n, w)) continue;  // block-uniform
        double(*Tt)[kGW + 1] = Tb(j);
        const int R1 = R0 + (w & 1) * 64, S1 = S0 + (w >> 1) * kGW;
        if constexpr (std::is_same<TD, float>::value) {
          // FP32 Zbar: four rows per thread and one 16-byte store where all four
          // are written; R1 and S1 are multiples of 4
          for (int e = threadIdx.x; e < 16 * kGW; e += blockDim.x) {
            const int rr = 4 * (e & 15), cc = e >> 4;
            const int row = R1 + rr, col = S1 + cc;
            if (col >= n || row >= n || row > col) continue;
            float *dp = D + (long)col * n + row;
            if (row + 3 < n && row + 3 <= col) {
              st_nt4(dp, Tt[rr][cc], Tt[rr + 1][cc], Tt[rr + 2][cc], Tt[rr + 3][cc]);
            } else {
#pragma unroll
              for (int q = 0; q < 4; ++q)
                if (row + q < n && row + q <= col) st_nt(dp + q, Tt[rr + q][cc]);
            }
          }
          for (int e = threadIdx.x; e < 16 * kGW; e += blockDim.x) {
            const int cc = 4 * (e & (kGW / 4 - 1)), rr = e / (kGW / 4);
            const int row2 = S1 + cc, col2 = R1 + rr;
            if (col2 >= n || row2 >= n) continue;
            double v[4];
            bool ok[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              ok[q] = row2 + q < n && row2 + q > col2;
              // phases from LDS: row2 + q - S0 and col2 - R0 lie in [0, 128)
              const double sq = (MODE == 1) ? h : (double)(s_ph[1][min(row2 + q - S0, kTileB - 1)] * s_ph[0][col2 - R0]);
              v[q] = sq * Tt[rr][cc + q];
            }
            float *dp = D + (long)col2 * n + row2;
            if (ok[0] && ok[1] && ok[2] && ok[3]) {
              st_nt4(dp, v[0], v[1], v[2], v[3]);
            } else {
#pragma unroll
              for (int q = 0; q < 4; ++q)
                if (ok[q]) st_nt(dp + q, v[q]);
            }
          }
        } else {
          // Two rows per thread and one 16-byte store where both are written (store
          // instructions, not bytes, bound this epilogue); R1 and S1 are even.
          // upper elements (R1 + rr, S1 + cc), consecutive threads on consecutive row pairs
          for (int e = threadIdx.x; e < 32 * kGW; e += blockDim.x) {
            const int rr = 2 * (e & 31), cc = e >> 5;
            const int row = R1 + rr, col = S1 + cc;
            if (col >= n || row >= n || row > col) continue;
            TD *dp = D + (long)col * n + row;
            if (row + 1 < n && row + 1 <= col)
              st_nt2(dp, Tt[rr][cc], Tt[rr + 1][cc]);
            else
              st_nt(dp, Tt[rr][cc]);
          }
          // mirrored elements (S1 + cc, R1 + rr) = sym * upper (R1 + rr, S1 + cc),
          // consecutive threads on consecutive row pairs S1 + cc
          for (int e = threadIdx.x; e < 32 * kGW; e += blockDim.x) {
            const int cc = 2 * (e & (kGW / 2 - 1)), rr = e / (kGW / 2);
            const int row2 = S1 + cc, col2 = R1 + rr;
            if (col2 >= n || row2 >= n) continue;
            const bool lo = row2 > col2, hi = row2 + 1 < n && row2 + 1 > col2;
            // phases from LDS: row2 - S0 and col2 - R0 lie in [0, 128)
            const double s0 = (MODE == 1) ? h : (double)(s_ph[1][row2 - S0] * s_ph[0][col2 - R0]);
            TD *dp = D + (long)col2 * n + row2;
            if (lo && hi) {
              const double s1 = (MODE == 1) ? h : (double)(s_ph[1][row2 + 1 - S0] * s_ph[0][col2 - R0]);
              st_nt2(dp, s0 * Tt[rr][cc], s1 * Tt[rr][cc + 1]);
            } else if (lo) {
              st_nt(dp, s0 * Tt[rr][cc]);
            } else if (hi) {
              const double s1 = (MODE == 1) ? h : (double)(s_ph[1][row2 + 1 - S0] * s_ph[0][col2 - R0]);
              st_nt(dp + 1, s1 * Tt[rr][cc + 1]);
            }
          }
        }
      }
      __syncthreads();
    }
  }
}

template <int MODE, bool MIXED, bool SPLIT = false>
__device__ __forceinline__ void gemm_tile(PairArg pa, const int4 *items) {
  static_assert(MIXED || !SPLIT, "split products are a mixed-precision mode");
  constexpr int T = kTileB, LD = kTileB + 4;
  // operands stay FP64 in LDS (they arrive by direct global->LDS copies); the
  // mixed kernel rounds them to FP32 when it reads them
  // the two stage buffers are separate LDS objects and every access names one
  // statically (the stage loop is unrolled by two), so the compiler can tell
  // the reads of one buffer from the direct copies into the other and does
  // not wait for the copies in flight before computing
  __shared__ __align__(16) double A0[kTileK][LD], A1[kTileK][LD];
  __shared__ __align__(16) double B0[kTileK][LD], B1[kTileK][LD];
  __shared__ double W0[kTileK], W1[kTileK];
  const int4 it = items[blockIdx.x];
  if (it.y < 0) return;  // padding item of the XCD-balanced order
  PairPtrs P = pa.get(it.x);
  const BasisDev &B = *P.basis;
  int n, nk, nph = 0;
  const double *Am, *Bm;
  const int *kl = nullptr;
  const double *kw = nullptr;
  const signed char *phi = nullptr, *phph = nullptr;
  double alpha, h;
  // Zbar in FP32 from the FP32-product ph GEMM, in FP64 otherwise
  using ZT = typename std::conditional<MODE != 0 && MIXED && !SPLIT, float, double>::type;
  ZT *D;
  if constexpr (MODE == 0) {
    const int ch = it.y;
    n = B.ch_n[ch];
    const long off = B.ch_off[ch];
    Am = P.X + B.two_off + off;
    Bm = P.Y + B.two_off + off;
    const int k0s = B.k_start[ch];
    nk = B.k_start[ch + 1] - k0s;
    kl = B.k_loc + k0s;
    kw = B.k_w + k0s;
    alpha = (double)P.hY;
    h = (double)P.hZ;
    D = P.Z + B.two_off + off;
  } else {
    const int cc = it.y;
    n = B.cc_n[cc];
    nph = B.cc_nph[cc];
    nk = 2 * nph;
    const long xo = B.cc_xoff[cc];
    Am = P.Xt + xo;
    Bm = P.Yt + xo;
    D = reinterpret_cast<ZT *>(P.Zbar) + B.cc_zoff[cc] + (MODE == 2 ? (long)n * n : 0);
    alpha = 1.0;
    h = (MODE == 1) ? (P.hZ == 1 ? 1.0 : -1.0) : 1.0;
    phi = B.cc_phi + B.cc_kstart[cc];
    phph = B.cc_phph + B.cc_phstart[cc];
  }
  const int R0 = it.z * T, S0 = it.w * T;
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // MODE 2: the phases of the tile's 128 rows and 128 columns, staged once for
  // the epilogue (the prologue's barrier below publishes them)
  __shared__ signed char s_ph[2][T];
  if constexpr (MODE == 2) {
    if (threadIdx.x < 2 * T) {
      const int sd = threadIdx.x / T, e = threadIdx.x - sd * T;
      s_ph[sd][e] = phi[min((sd ? S0 : R0) + e, n - 1)];
    }
  }
  // 8 waves, each a 64x32 sub-tile: rows wr .. wr+63, columns wc .. wc+31
  const int wr = (wv & 1) * 64, wc = (wv >> 1) * kGW;
  // a wave whose sub-tile holds no upper-triangle element inside the matrix
  // (padding past n, or below the diagonal) issues no MFMA
  const bool wave_live = tile_wave_live(R0, S0, n, wv);
  dbl4 acc[4][kGN];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < kGN; ++b) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
  const int K2 = 2 * nk;
  // Stage loader: wave w copies the k' columns 2w, 2w+1 of the stage, for
  // the A and the B row block, each a 128-row segment of one column, with one
  // 16-byte-per-lane global->LDS copy (lane l: rows 2l, 2l+1).  Rows past n
  // are never written back; their source is clamped so that a row pair reads
  // at most one element past the column (inside the arena, or the buffer's
  // slack).  Column indices (wave-uniform) are loaded one stage ahead.
  const int ra = min(R0 + 2 * lane, n - 1), rb = min(S0 + 2 * lane, n - 1);
  // MODE 0: the channel's column-index list kl[0..nk) is staged in LDS (dynamic
  // shared memory of at least nk ints), so index reads never wait behind the
  // global->LDS copies in flight
  extern __shared__ int KLs[];
  if constexpr (MODE == 0) {
    for (int e = threadIdx.x; e < nk; e += blockDim.x) KLs[e] = kl[e];
    __syncthreads();
  }
  auto load_idx = [&](int k0, int (&ki)[kGC]) {
#pragma unroll
    for (int j = 0; j < kGC; ++j) {
      const int kk = min(k0 + wv * kGC + j, K2 - 1);
      const int k = kk < nk ? kk : kk - nk;
      ki[j] = (MODE == 0) ? KLs[k] : k;
    }
  };
  auto issue_stage = [&](int k0, const int (&ki)[kGC], double (&As)[kTileK][LD], double (&Bs)[kTileK][LD]) {
#pragma unroll
    for (int j = 0; j < kGC; ++j) {
      const int kq = wv * kGC + j;
      const int kk = min(k0 + kq, K2 - 1);
      const int ca = ki[j];
      const int cb = (MODE == 2) ? (ca < nph ? ca + nph : ca - nph) : ca;
      const double *sa = kk < nk ? Am + (long)ca * n : Bm + (long)cb * n;
      const double *sb = kk < nk ? Bm + (long)cb * n : Am + (long)ca * n;
      __builtin_amdgcn_global_load_lds((const IMSRG_G void *)(sa + ra),
                                       (__attribute__((address_space(3))) void *)&As[kq][0], 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const IMSRG_G void *)(sb + rb),
                                       (__attribute__((address_space(3))) void *)&Bs[kq][0], 16, 0, 0);
    }
  };
  // the weight of k' = k0 + t (t < 16) is loaded into a register and only
  // scaled and stored to LDS at the end of the stage
  auto load_w = [&](int k0) -> double {
    const int kk = min(k0 + (int)(threadIdx.x & (kTileK - 1)), K2 - 1);
    const int k = kk < nk ? kk : kk - nk;
    if constexpr (MODE == 0) return kw[k];
    if constexpr (MODE == 1) return 1.0;
    return (double)P.hY * (double)phph[k < nph ? k : k - nph];
  };
  auto store_w = [&](int k0, double (&Ws)[kTileK], double w) {
    if (threadIdx.x < kTileK) {
      const int kk = k0 + threadIdx.x;
      Ws[threadIdx.x] = kk >= K2 ? 0.0 : kk < nk ? w : w * h;
    }
  };
  int ki[kGC];
  if (K2 > 0) {  // a channel without pp or hh kets only writes zeros
    load_idx(0, ki);
    issue_stage(0, ki, A0, B0);
    store_w(0, W0, load_w(0));
    load_idx(kTileK, ki);  // clamped: harmless past the end
  }
  __builtin_amdgcn_s_waitcnt(0);  // the copies into LDS have landed
  __syncthreads();
  // one stage: copies of the next stage into (An, Bn, Wn), MFMAs on (Ac, Bc, Wc)
  // mixed precision: FP32 partial sums of the wave's 4 x kGN blocks, flushed into
  // the FP64 accumulators every second stage (32 k)
  flt4 t[4][kGN];
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < kGN; ++ni) t[mi][ni] = flt4{0.f, 0.f, 0.f, 0.f};
  auto flush_t = [&]() {
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < kGN; ++ni) {
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[mi][ni][q] += (double)t[mi][ni][q];
        t[mi][ni] = flt4{0.f, 0.f, 0.f, 0.f};
      }
  };
  auto stage = [&](int k0, double (&Ac)[kTileK][LD], double (&Bc)[kTileK][LD], double (&Wc)[kTileK],
                   double (&An)[kTileK][LD], double (&Bn)[kTileK][LD], double (&Wn)[kTileK], bool flush) {
    const bool more = k0 + kTileK < K2;
    // loads with register results are unconditional (clamped indices): a
    // branch around them would make the compiler wait for them, and for the
    // copies issued with them, where the paths join
    const double w_next = load_w(k0 + kTileK);
    if (more) issue_stage(k0 + kTileK, ki, An, Bn);
    load_idx(k0 + 2 * kTileK, ki);
    if (!wave_live) {
    } else if constexpr (MIXED) {
#pragma unroll
      for (int ks = 0; ks < kTileK / 4; ++ks) {
        const int kr = ks * 4 + (lane >> 4);
        const double wk = Wc[kr];
        float am[4], bm[kGN], al[4], bl[kGN];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const double a = Ac[kr][wr + m * 16 + (lane & 15)] * wk;
          am[m] = (float)a;
          if constexpr (SPLIT) al[m] = (float)(a - (double)am[m]);
        }
#pragma unroll
        for (int m = 0; m < kGN; ++m) {
          const double b = Bc[kr][wc + m * 16 + (lane & 15)];
          bm[m] = (float)b;
          if constexpr (SPLIT) bl[m] = (float)(b - (double)bm[m]);
        }
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
#pragma unroll
          for (int ni = 0; ni < kGN; ++ni) {
            if constexpr (SPLIT) {
              t[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x4f32(bl[ni], am[mi], t[mi][ni], 0, 0, 0);
              t[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x4f32(bm[ni], al[mi], t[mi][ni], 0, 0, 0);
            }
            t[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x4f32(bm[ni], am[mi], t[mi][ni], 0, 0, 0);
          }
        // split products flush every 4 k: with the operand rounding gone, the
        // FP32 partial sums over 32 k would set the error
        if constexpr (SPLIT) flush_t();
      }
      if (!SPLIT && flush) flush_t();
    } else {
#pragma unroll
      for (int ks = 0; ks < kTileK / 4; ++ks) {
        const int kr = ks * 4 + (lane >> 4);
        const double wk = Wc[kr];
        // the weight of k scales the kGN B values rather than the four A
        // values (1.8% off the FP64 pp/hh GEMM; the mixed kernel keeps it on A,
        // where moving it costs spills)
        double a[4], b[kGN];
#pragma unroll
        for (int m = 0; m < 4; ++m) a[m] = Ac[kr][wr + m * 16 + (lane & 15)];
#pragma unroll
        for (int m = 0; m < kGN; ++m) b[m] = Bc[kr][wc + m * 16 + (lane & 15)] * wk;
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
#pragma unroll
          for (int ni = 0; ni < kGN; ++ni)
            acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[ni], a[mi], acc[mi][ni], 0, 0, 0);
      }
    }
    if (more) store_w(k0 + kTileK, Wn, w_next);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's copies of the next stage have landed
    __syncthreads();
  };
  for (int k0 = 0; k0 < K2; k0 += 2 * kTileK) {
    stage(k0, A0, B0, W0, A1, B1, W1, false);
    if (k0 + kTileK < K2) stage(k0 + kTileK, A1, B1, W1, A0, B0, W0, true);
  }
  if constexpr (MIXED) flush_t();  // a last stage without its pair
  static_assert(64 * (kGW + 1) <= kTileK * LD, "transpose tile exceeds the operand buffer");
  gemm_epilogue<MODE, MIXED, ZT>(acc, D, n, R0, S0, alpha, h, s_ph, &A0[0][0], &A1[0][0], &B0[0][0], &B1[0][0]);
}

// ---------------------------------------------------------------------------
// The same 128x128 SYR2K tiles on bf16 MFMA with three-way split operands
// (mixed precision, "bf16x3").  Each FP64 operand x (A values already scaled
// by their weight w_k) is held as three bf16 values x = p0 + p1 + p2:
// p0 = bf16(fl32(x)), p1 = bf16(fl32(x) - p0) and p2 = bf16(fl32(x - (p0 + p1)))
// from the FP64 residual, so |x - (p0 + p1 + p2)| <= 2^-26 |x| (the FP32
// mixed mode rounds each operand by 2^-24).  A product takes six
// v_mfma_f32_16x16x32_bf16 (p0q0, p0q1, p1q0, p1q1, p0q2, p2q0; the dropped
// p1q2 + p2q1 + p2q2 are <= 2^-26 of it), accumulated in FP32 over one stage
// of 32 k and flushed into the FP64 accumulators after every stage.  bf16
// MFMA runs at 16x the FP32 matrix rate on MI355X, so six products cost
// 3/8 of one FP32 product.
// The ph GEMM's kernel (gemm_tile_b3) splits operands once per tile when a
// stage is staged: each wave loads 8 consecutive k of two rows of one operand
// (16-byte global loads, lanes on consecutive row pairs), splits them in
// registers and stores the three bf16x8 vectors in the MFMA operand layout
// (lane l of a 16-row block reads row l % 16, k octet l / 16: one 16-byte LDS
// read per fragment).  The next stage's global loads are issued before the
// current stage's MFMAs.  The pp/hh GEMM splits its operands once per
// commutator instead (k_pack_pphh, below).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));

// LDS of one stage: [operand A/B][piece 0..2][k octet 0..3][row 0..127], 16 bytes each
typedef bf16x8 StageB3[2][3][4][kTileB];

// The FP32 intermediates pass through empty asm statements (no instructions):
// otherwise the compiler fuses the double -> float -> bf16 conversions into one
// correctly rounded double -> bf16 conversion (round-to-odd emulation, eight
// instructions), where any nearby p0 serves, the residuals being exact.
typedef float f2v __attribute__((ext_vector_type(2)));
typedef __bf16 b2v __attribute__((ext_vector_type(2)));
// The three pieces of two values (consecutive k of one row) at once: each
// rounding to bf16 is one v_cvt_pk_bf16_f32 for the pair.
__device__ __forceinline__ void split3x2(double x0, double x1, b2v &p0, b2v &p1, b2v &p2) {
  f2v f = {(float)x0, (float)x1};
  asm("" : "+v"(f));
  p0 = __builtin_convertvector(f, b2v);
  const f2v f0 = __builtin_convertvector(p0, f2v);
  p1 = __builtin_convertvector(f - f0, b2v);
  const f2v s01 = f0 + __builtin_convertvector(p1, f2v);  // exact: p0 and p1 span at most 18 bits
  f2v r2 = {(float)(x0 - (double)s01[0]), (float)(x1 - (double)s01[1])};
  asm("" : "+v"(r2));
  p2 = __builtin_convertvector(r2, b2v);
}

template <int MODE>
__device__ __forceinline__ void gemm_tile_b3(PairArg pa, const int4 *items) {
  constexpr int T = kTileB, KS = 32;
  __shared__ __align__(16) StageB3 L0, L1;
  const int4 it = items[blockIdx.x];
  if (it.y < 0) return;  // padding item of the XCD-balanced order
  PairPtrs P = pa.get(it.x);
  const BasisDev &B = *P.basis;
  int n, nk, nph = 0;
  const double *Am, *Bm;
  const signed char *phi = nullptr, *phph = nullptr;
  double alpha, h;
  double *D;
  static_assert(MODE == 1 || MODE == 2, "the ph GEMM's head (1) and tail (2); the pp/hh GEMM packs its panels");
  {
    const int cc = it.y;
    n = B.cc_n[cc];
    nph = B.cc_nph[cc];
    nk = 2 * nph;
    const long xo = B.cc_xoff[cc];
    Am = P.Xt + xo;
    Bm = P.Yt + xo;
    D = P.Zbar + B.cc_zoff[cc] + (MODE == 2 ? (long)n * n : 0);
    alpha = 1.0;
    h = (MODE == 1) ? (P.hZ == 1 ? 1.0 : -1.0) : 1.0;
    phi = B.cc_phi + B.cc_kstart[cc];
    phph = B.cc_phph + B.cc_phstart[cc];
  }
  const int R0 = it.z * T, S0 = it.w * T;
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  __shared__ signed char s_ph[2][T];
  if constexpr (MODE == 2) {
    if (threadIdx.x < 2 * T) {
      const int sd = threadIdx.x / T, e = threadIdx.x - sd * T;
      s_ph[sd][e] = phi[min((sd ? S0 : R0) + e, n - 1)];
    }
  }
  __syncthreads();  // s_ph
  const int K2 = 2 * nk;
  // loader role of this wave: operand (0: rows R0.. of A-side values, 1: rows
  // S0.. of B-side values), k octet, and the lane's two rows
  const int opnd = wv >> 2, oct = wv & 3;
  const int rbase = opnd ? S0 : R0;
  const int row = min(rbase + 2 * lane, n - 1);  // row, row + 1 (one past a column at most: slack)
  // two register sets: the global loads of a stage are issued two stages
  // before it is computed (one stage of compute between issue and use left the
  // load latency exposed at two waves per SIMD)
  d2u va[8], vb[8];
  // global loads of one stage and the factors of its 8 k slots (the weights on
  // the A side, 1 on the B side, 0 past the end).  Indices and weights are
  // wave-uniform; all their scalar loads are issued before any is used, with
  // no branch between them
  auto load_stage = [&](int k0, d2u (&v)[8], double (&wf)[8]) {
    const int kb = k0 + 8 * oct;
    int kk[8], k[8], ca[8];
    double wr[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      kk[j] = min(kb + j, K2 - 1);
      k[j] = kk[j] < nk ? kk[j] : kk[j] - nk;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      ca[j] = k[j];
      wr[j] = (MODE == 1) ? 1.0 : (double)P.hY * (double)phph[k[j] < nph ? k[j] : k[j] - nph];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int cb = (MODE == 2) ? (ca[j] < nph ? ca[j] + nph : ca[j] - nph) : ca[j];
      // A side: A(., ca) for kk < nk, else B(., cb); B side the other one
      const bool first = kk[j] < nk;
      const double *src = (first != (opnd != 0)) ? Am + (long)ca[j] * n : Bm + (long)cb * n;
      v[j] = *reinterpret_cast<const d2u *>(src + row);
      const double w = first ? wr[j] : wr[j] * h;
      wf[j] = kb + j >= K2 ? 0.0 : opnd ? 1.0 : w;
    }
  };
  auto store_stage = [&](const d2u (&v)[8], const double (&wf)[8], StageB3 &L) {
    union {
      bf16x8 v8;
      b2v h[4];
    } q[2][3];
#pragma unroll
    for (int j = 0; j < 8; j += 2) {
      b2v x0, x1, x2, y0, y1, y2;
      split3x2(v[j].x * wf[j], v[j + 1].x * wf[j + 1], x0, x1, x2);  // row 2 lane
      split3x2(v[j].y * wf[j], v[j + 1].y * wf[j + 1], y0, y1, y2);  // row 2 lane + 1
      q[0][0].h[j / 2] = x0; q[0][1].h[j / 2] = x1; q[0][2].h[j / 2] = x2;
      q[1][0].h[j / 2] = y0; q[1][1].h[j / 2] = y1; q[1][2].h[j / 2] = y2;
    }
    const int r = 2 * lane;
#pragma unroll
    for (int pc = 0; pc < 3; ++pc) {
      L[opnd][pc][oct][r] = q[0][pc].v8;
      L[opnd][pc][oct][r + 1] = q[1][pc].v8;
    }
  };
  const int wr = (wv & 1) * 64, wc = (wv >> 1) * kGW;
  const bool wave_live = tile_wave_live(R0, S0, n, wv);
  dbl4 acc[4][kGN];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < kGN; ++b) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
  auto compute = [&](const StageB3 &L) {
    if (!wave_live) return;
    const int oc = lane >> 4, rl = lane & 15;
    // the B-side fragments of all three pieces stay in registers; the A side
    // is read one piece at a time (fewer registers live at once)
    bf16x8 bfr[kGN][3];
#pragma unroll
    for (int pc = 0; pc < 3; ++pc)
#pragma unroll
      for (int ni = 0; ni < kGN; ++ni) bfr[ni][pc] = L[1][pc][oc][wc + ni * 16 + rl];
    flt4 t[4][kGN];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < kGN; ++ni) t[mi][ni] = flt4{0.f, 0.f, 0.f, 0.f};
    // products p_a q_b with a + b <= 2: (2,0), then (1,0), (1,1), then (0,2), (0,1), (0,0);
    // one product for all 8 blocks before the next, so that consecutive MFMAs
    // into one accumulator are 8 instructions apart
#pragma unroll
    for (int pa3 = 2; pa3 >= 0; --pa3) {
      bf16x8 af[4];
#pragma unroll
      for (int mi = 0; mi < 4; ++mi) af[mi] = L[0][pa3][oc][wr + mi * 16 + rl];
#pragma unroll
      for (int pb3 = 2 - pa3; pb3 >= 0; --pb3)
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
#pragma unroll
          for (int ni = 0; ni < kGN; ++ni)
            t[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[ni][pb3], af[mi], t[mi][ni], 0, 0, 0);
    }
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < kGN; ++ni)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[mi][ni][q] += (double)t[mi][ni][q];
  };
  if (K2 > 0) {  // a channel without pp or hh kets only writes zeros
    const int nst = (K2 + KS - 1) / KS;
    double wa[8], wb[8];
    load_stage(0, va, wa);
    if (nst > 1) load_stage(KS, vb, wb);
    store_stage(va, wa, L0);
    __syncthreads();
    for (int st = 0; st < nst; st += 2) {
      // stage st in L0, stage st+1 in flight in vb, va free
      if (st + 2 < nst) load_stage((st + 2) * KS, va, wa);
      compute(L0);
      if (st + 1 < nst) store_stage(vb, wb, L1);
      __syncthreads();
      if (st + 1 >= nst) break;
      // stage st+1 in L1, stage st+2 in flight in va, vb free
      if (st + 3 < nst) load_stage((st + 3) * KS, vb, wb);
      compute(L1);
      if (st + 2 < nst) store_stage(va, wa, L0);
      __syncthreads();
    }
  }
  static_assert(2 * 64 * (kGW + 1) * sizeof(double) <= sizeof(StageB3), "transpose tiles exceed the stage buffer");
  double *t0 = reinterpret_cast<double *>(&L0[0][0][0][0]), *t2 = reinterpret_cast<double *>(&L1[0][0][0][0]);
  gemm_epilogue<MODE, true>(acc, D, n, R0, S0, alpha, h, s_ph, t0, t0 + 64 * (kGW + 1), t2, t2 + 64 * (kGW + 1));
}

// ---------------------------------------------------------------------------
// Mixed pp/hh GEMM from packed panels.  k_pack_pphh splits each operand element
// once per commutator into its three bf16 pieces (split3x2, X already scaled by
// its weight) and stores them in the MFMA operand layout; the tile kernel then
// fills its LDS stages with direct global->LDS copies of those panels (no
// register staging, no conversion per tile) and runs six bf16 products per
// product as gemm_tile_b3 does.  The SYR2K's two halves are
//   sum_k (X w)(r,k) Y(s,k)  and  h * sum_k Y(r,k) (X w)(s,k),
// so the second half reads the same two panels with the sides swapped and its
// FP32 partial sums are flushed with the factor h.

// one work item: k octet it.z of channel it.y, rows 256 it.w .. + 255
__global__ __launch_bounds__(256) void k_pack_pphh(PairArg pa, const int4 *items) {
  const int4 it = items[blockIdx.x];
  PairPtrs P = pa.get(it.x);
  const BasisDev &B = *P.basis;
  const int ch = it.y, oct = it.z;
  const int n = B.ch_n[ch];
  const int k0s = B.k_start[ch], nk = B.k_start[ch + 1] - k0s;
  const int nrp = (n + kTileB - 1) / kTileB * kTileB;
  const int noct = (nk + kB3K - 1) / kB3K * (kB3K / 8);
  const int r = it.w * 256 + (int)threadIdx.x;
  if (r >= nrp) return;
  const long off = B.ch_off[ch];
  const double *X2 = P.X + B.two_off + off, *Y2 = P.Y + B.two_off + off;
  const int rr = min(r, n - 1);
  double xv[8], yv[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = 8 * oct + j;
    const int kc = min(k, nk - 1);
    const int col = B.k_loc[k0s + kc];  // wave-uniform
    const double w = B.k_w[k0s + kc];
    const bool ok = k < nk && r < n;
    const double x = X2[(long)col * n + rr], y = Y2[(long)col * n + rr];
    xv[j] = ok ? x * w : 0.0;
    yv[j] = ok ? y : 0.0;
  }
  union {
    bf16x8 v8;
    b2v h[4];
  } qx[3], qy[3];
#pragma unroll
  for (int j = 0; j < 8; j += 2) {
    split3x2(xv[j], xv[j + 1], qx[0].h[j / 2], qx[1].h[j / 2], qx[2].h[j / 2]);
    split3x2(yv[j], yv[j + 1], qy[0].h[j / 2], qy[1].h[j / 2], qy[2].h[j / 2]);
  }
  bf16x8 *px = reinterpret_cast<bf16x8 *>(P.panel) + B.pk_off[ch];
  bf16x8 *py = px + (long)noct * 3 * nrp;
#pragma unroll
  for (int pc = 0; pc < 3; ++pc) {
    px[((long)oct * 3 + pc) * nrp + r] = qx[pc].v8;
    py[((long)oct * 3 + pc) * nrp + r] = qy[pc].v8;
  }
}

// Measured alternatives at emax = 10 (pp/hh phase with the packing, 740 us
// for this kernel): register-staged copies two stages ahead 751 us; three LDS
// stages with the copies' waits written by hand 790 us; two wave groups that
// alternate between LDS reads and MFMA 800 us.  The kernel streams 4.0 GB of
// panels per commutator through L2 into LDS, and that stream sets its time:
// the copies alone take 510 us, 205 us when every tile re-reads data that
// stays in L2.
__global__ __launch_bounds__(kGWaves * 64, 2) void k_pphh_big_b3p(PairArg pa, const int4 *items) {
  constexpr int T = kTileB;
  __shared__ __align__(16) StageB3 L0, L1;
  __shared__ signed char s_ph[2][T];  // unused in this mode (epilogue signature)
  const int4 it = items[blockIdx.x];
  if (it.y < 0) return;  // padding item of the XCD-balanced order
  PairPtrs P = pa.get(it.x);
  const BasisDev &B = *P.basis;
  const int ch = it.y;
  const int n = B.ch_n[ch];
  const int nk = B.k_start[ch + 1] - B.k_start[ch];
  const int nrp = (n + kTileB - 1) / kTileB * kTileB;
  const int noct = (nk + kB3K - 1) / kB3K * (kB3K / 8);
  const bf16x8 *px = reinterpret_cast<const bf16x8 *>(P.panel) + B.pk_off[ch];
  const bf16x8 *py = px + (long)noct * 3 * nrp;
  const double alpha = (double)P.hY, h = (double)P.hZ;
  double *D = P.Z + B.two_off + B.ch_off[ch];
  const int R0 = it.z * T, S0 = it.w * T;
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nst1 = noct / 4, nst = 2 * nst1;  // stages of 32 k per half, both halves
  // stage copies: 2 sides x 3 pieces x 4 k octets x 2 halves of 64 rows = 48
  // wave copies of 1 KiB (16 bytes per lane), six per wave
  auto issue = [&](int st, StageB3 &L) {
    const bool second = st >= nst1;
    const int o0 = (second ? st - nst1 : st) * 4;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const int q = wv * 6 + i;
      const int side = q / 24, pc = (q / 8) % 3, oct = (q / 2) % 4, half = q & 1;
      const bf16x8 *src = ((side == 0) != second) ? px : py;
      const int row0 = (side ? S0 : R0) + half * 64;
      const bf16x8 *g = src + ((long)(o0 + oct) * 3 + pc) * nrp + row0 + lane;
      __builtin_amdgcn_global_load_lds((const IMSRG_G void *)g,
                                       (__attribute__((address_space(3))) void *)&L[side][pc][oct][half * 64], 16, 0,
                                       0);
    }
  };
  const int wr = (wv & 1) * 64, wc = (wv >> 1) * kGW;
  const bool wave_live = tile_wave_live(R0, S0, n, wv);
  dbl4 acc[4][kGN];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < kGN; ++b) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
  // six bf16 products of a stage summed in FP32 (smallest first), then
  // flushed into FP64 with the half's factor (1, or h for the second half)
  auto compute = [&](const StageB3 &L, double fl) {
    if (!wave_live) return;
    const int oc = lane >> 4, rl = lane & 15;
    bf16x8 bfr[kGN][3];
#pragma unroll
    for (int pc = 0; pc < 3; ++pc)
#pragma unroll
      for (int ni = 0; ni < kGN; ++ni) bfr[ni][pc] = L[1][pc][oc][wc + ni * 16 + rl];
    flt4 t[4][kGN];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < kGN; ++ni) t[mi][ni] = flt4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int pa3 = 2; pa3 >= 0; --pa3) {
      bf16x8 af[4];
#pragma unroll
      for (int mi = 0; mi < 4; ++mi) af[mi] = L[0][pa3][oc][wr + mi * 16 + rl];
#pragma unroll
      for (int pb3 = 2 - pa3; pb3 >= 0; --pb3)
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
#pragma unroll
          for (int ni = 0; ni < kGN; ++ni)
            t[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[ni][pb3], af[mi], t[mi][ni], 0, 0, 0);
    }
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < kGN; ++ni)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[mi][ni][q] += fl * (double)t[mi][ni][q];
  };
  if (nst > 0) {  // a channel without pp or hh kets only writes zeros
    issue(0, L0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int st = 0; st < nst; st += 2) {
      if (st + 1 < nst) issue(st + 1, L1);
      compute(L0, st >= nst1 ? h : 1.0);
      __builtin_amdgcn_s_waitcnt(0);  // this wave's copies of the next stage have landed
      __syncthreads();
      if (st + 1 >= nst) break;
      if (st + 2 < nst) issue(st + 2, L0);
      compute(L1, st + 1 >= nst1 ? h : 1.0);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
    }
  }
  double *t0 = reinterpret_cast<double *>(&L0[0][0][0][0]), *t2 = reinterpret_cast<double *>(&L1[0][0][0][0]);
  gemm_epilogue<0, true>(acc, D, n, R0, S0, alpha, h, s_ph, t0, t0 + 64 * (kGW + 1), t2, t2 + 64 * (kGW + 1));
}

template <int MODE>
__global__ __launch_bounds__(kGWaves * 64, 2) void k_ph_big_b3(PairArg pa, const int4 *items) {
  gemm_tile_b3<MODE>(pa, items);
}

__global__ __launch_bounds__(kGWaves * 64, 4) void k_pphh_big_f64(PairArg pa, const int4 *items) {
  gemm_tile<0, false>(pa, items);
}
__global__ __launch_bounds__(kGWaves * 64, 4) void k_pphh_big_mixed(PairArg pa, const int4 *items) {
  gemm_tile<0, true>(pa, items);
}
__global__ __launch_bounds__(kGWaves * 64, 4) void k_pphh_big_split(PairArg pa, const int4 *items) {
  gemm_tile<0, true, true>(pa, items);
}
template <int MODE, bool MIXED, bool SPLIT = false>
__global__ __launch_bounds__(kGWaves * 64, 4) void k_ph_big(PairArg pa, const int4 *items) {
  gemm_tile<MODE, MIXED, SPLIT>(pa, items);
}

// ---------------------------------------------------------------------------
// DoPandyaTransformation_SingleChannel_XandY (src/Commutator.cc:1392-1485):
// Xbar (nk x 2nph) and Ybar^T (nk x 2nph), one element each per thread.
__global__ __launch_bounds__(256) void k_pandya(PairArg pa, const int4 *items) {
  extern __shared__ __align__(8) unsigned char lds_raw[];
  const int4 it = items[blockIdx.x];
  PairPtrs P = pa.get(it.x);
  const BasisDev &B = *P.basis;
  // channel offsets and sizes in LDS: the gather of each J then waits on one
  // LDS read instead of two dependent global loads
  long *s_choff = (long *)lds_raw;
  int *s_chn = (int *)(s_choff + B.nch);
  for (int c = threadIdx.x; c < B.nch; c += blockDim.x) {
    s_choff[c] = B.ch_off[c];
    s_chn[c] = B.ch_n[c];
  }
  __syncthreads();
  const int cc = it.y;
  const int nk = B.cc_n[cc], nph = B.cc_nph[cc];
  const long tot = (long)nk * 2 * nph;
  const long e = (long)it.z + threadIdx.x;
  if (e >= tot) return;
  const int mp = (int)(e / nk), ik = (int)(e % nk);
  const int abc = mp / nph, ibra = mp - abc * nph;
  const int locb = B.cc_ph[B.cc_phstart[cc] + ibra];
  const int kb = B.cc_kets[B.cc_kstart[cc] + locb];
  const int p = B.kp[kb], q = B.kq[kb];
  const int a = abc ? q : p, b = abc ? p : q;
  const int kk = B.cc_kets[B.cc_kstart[cc] + ik];
  const int c = B.kp[kk], d = B.kq[kk];
  const int Jcc = cc / 6;
  const double *X2 = P.X + B.two_off, *Y2 = P.Y + B.two_off;
  double xbar = 0.0, ybar = 0.0, normf = 1.0;
  const double nanb = B.occ[a] - B.occ[b];
  if (B.tz2[a] + B.tz2[d] == B.tz2[b] + B.tz2[c]) {
    if (a == d) normf *= SQRT2_D;
    if (c == b) normf *= SQRT2_D;
    const int ja = B.j2[a], jb = B.j2[b], jc = B.j2[c], jd = B.j2[d];
    int jmin = max(abs(ja - jd), abs(jc - jb)) / 2;
    const int jmax = min(ja + jd, jc + jb) / 2;
    int dJ = 1;
    if (a == d || c == b) {
      dJ = 2;
      jmin += jmin & 1;
    }
    // index work common to X and Y (GetTBME_J_norm_twoOps, src/TwoBodyME.cc:368-409)
    const int pcb = (B.l[c] + B.l[b]) & 1, pad = (B.l[a] + B.l[d]) & 1;
    const int tcb = (B.tz2[c] + B.tz2[b]) / 2;
    const double hY = P.hY;
    if (pcb == pad) {
      // two J per pass, every load of both issued before use; an invalid term
      // reads a clamped in-bounds element and is weighted by 0.  Local indices of
      // the kets (c,b) and (a,d) come from the per-orbit rows of b and a, where
      // the lanes' consecutive c and d read neighbouring entries.
      const IMSRG_G double *sixt = B.sixj;
      const int nJ = B.jmax + 1, norb = B.norb;
      const IMSRG_G int *kb1 = B.kmorb + (long)b * nJ * norb + c;
      const IMSRG_G int *kb2 = B.kmorb + (long)a * nJ * norb + d;
      for (int J0 = jmin; J0 <= jmax; J0 += 2 * dJ) {
        double sj[2], xv[2], yv[2], wv2[2];
        int i1[2], i2[2], ch[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int J = min(J0 + u * dJ, jmax);
          sj[u] = sixt[sixj_index(B.d1, B.d2, ja, jb, Jcc, jc, jd, J)];
          i1[u] = kb1[J * norb];
          i2[u] = kb2[J * norb];
          ch[u] = 6 * J + 2 * (tcb + 1) + pcb;
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int J = J0 + u * dJ;
          const bool ok = J <= jmax && fabs(sj[u]) >= 1e-8 && i1[u] >= 0 && i2[u] >= 0;
          double ph = 1.0;
          if (c > b) ph *= sgn_parity((jc + jb) / 2 + 1 + J);
          if (a > d) ph *= sgn_parity((ja + jd) / 2 + 1 + J);
          const long idx = s_choff[ch[u]] + (long)max(i2[u], 0) * s_chn[ch[u]] + max(i1[u], 0);
          xv[u] = X2[idx];
          yv[u] = Y2[idx];
          wv2[u] = ok ? (2 * J + 1) * sj[u] * ph : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          xbar -= wv2[u] * xv[u];
          ybar -= wv2[u] * yv[u] * hY;
        }
      }
    }
  }
  const long o = B.cc_xoff[cc] + (long)mp * nk + ik;
  P.Xt[o] = xbar * normf * nanb;
  P.Yt[o] = ybar * normf;
}


// wave-wide min / max of an int, returned as a wave-uniform (scalar) value
__device__ __forceinline__ int wave_min(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int wave_max(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return __builtin_amdgcn_readfirstlane(v);
}

// AddInversePandyaTransformation (src/Commutator.cc:1501-1616), by orbit quadruple.
// One work-group per bra ket (ij) of a (parity, Tz) group; its lanes take the
// group's kets (kl) = e >= (ij) in (j_k, j_l) class order, so that the lanes of
// a wave mostly share their J and J' ranges (one lane per ket, in chunks of the
// work-group size).  The cross-coupled elements Zbar^{J'}[(il),(kj)] (direct) and
// Zbar^{J'}[(jl),(ki)] (exchange) do not depend on the standard-channel J, so
// each is loaded once and recoupled into every J of the quadruple (register
// chunks of kJC values):
//   commij(J) = -sum_J' (2J'+1) {ji jj J; jk jl J'} Zbar^{J'}[il, kj]
//   commji(J) = -sum_J' (2J'+1) {jj ji J; jk jl J'} Zbar^{J'}[jl, ki]
//   Z^J_{ij,kl} += -(commij - (-1)^{(jk+jl)/2 - J} commji) / norm
// The 6j copy sixjQ holds the weights (2J'+1) {...}, with zeros where the
// reference skips |6j| < 1e-8.  Every J of a chunk is accumulated unmasked: a J
// outside the lane's triangles has a zero 6j, and the accumulators of J values
// the lane does not own (odd J of a (p,p) ket, J past the range) are never
// stored.  CC channels without ph kets (Zbar = 0, not written by the ph GEMM)
// are skipped.  TZ is the element type of Zbar: float after the FP32-product ph
// GEMM, double otherwise; the sums are FP64.  The cross-coupled local indices of
// the kets (i,x) and (j,x), the CC channel sizes/offsets and the bra's
// standard-channel data are staged in LDS once per work-group.  Every upper
// element (ij <= kl, or ij < kl for anti-Hermitian Z) of every channel is
// produced exactly once and stored at its transposed slot (row ib of channel J,
// written by this work-group only, so the L2 merges the lanes' scattered
// stores); k_finish adds it to the upper triangle, the diagonal is kept apart.
__host__ __device__ inline size_t invp_lds_bytes(int jmax, int norb, int nch, int ncls) {
  const size_t nJ = (size_t)jmax + 1;
  return 8 * (size_t)nch + 8 * nJ + 4 * (size_t)nch + 4 * nJ * (2 * (size_t)norb + 3) + 4 * (2 * (size_t)ncls + 1);
}

template <typename TZ>
__global__ __launch_bounds__(256, 5) void k_invpandya(PairArg pa, const int4 *items) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int4 it = items[blockIdx.x];
  PairPtrs P = pa.get(it.x);
  const BasisDev &B = *P.basis;
  const int g = it.y, b = it.z;
  const int gs = B.grp_start[g];
  const int nk = B.nkets, norb = B.norb, nJ = B.jmax + 1, nch = B.nch;
  long *s_zoff = (long *)lds_raw;      // [nch] CC channel offsets in Zbar
  long *s_choff = s_zoff + nch;        // [nJ]  offsets of the standard channels 6J + g
  int *s_ccn = (int *)(s_choff + nJ);  // [nch] CC channel sizes
  int *s_chn = s_ccn + nch;            // [nJ]  sizes of the standard channels 6J + g
  int *s_ib = s_chn + nJ;              // [nJ]  local index of the bra in channel 6J + g
  int *s_chks = s_ib + nJ;             // [nJ]  first ket (all-channel index) of channel 6J + g
  int *Ti = s_chks + nJ;               // [nJ * norb] CC local index of ket (i,x) in channel J'
  int *Tj = Ti + nJ * norb;            // [nJ * norb] CC local index of ket (j,x) in channel J'
  const int ncls = B.ncls;
  int *s_us = Tj + nJ * norb;          // [ncls] first position of class c with e >= emin
  int *s_pre = s_us + ncls;            // [ncls + 1] positions of the classes in the bra's ket order
  const int kb = B.grp_kets[gs + b];
  const int emin = (P.hZ == 1) ? b : b + 1;
  const int i = B.kp[kb], j = B.kq[kb];
  // a CC channel without ph kets has Zbar = 0, which the ph GEMM does not write:
  // offset -1 marks it, and its terms are skipped
  for (int c = threadIdx.x; c < nch; c += blockDim.x) {
    s_zoff[c] = B.cc_nph[c] > 0 ? B.cc_zoff[c] : -1;
    s_ccn[c] = B.cc_n[c];
  }
  for (int J = threadIdx.x; J < nJ; J += blockDim.x) {
    s_choff[J] = B.ch_off[6 * J + g];
    s_chn[J] = B.ch_n[6 * J + g];
    s_ib[J] = B.kmapT[(long)J * nk + kb];
    s_chks[J] = B.ch_kstart[6 * J + g];
  }
  // the bra's rows of the per-orbit cross-coupled index table: contiguous reads
  const IMSRG_G int *ci = B.ccorb + (long)i * nJ * norb, *cj = B.ccorb + (long)j * nJ * norb;
  for (int e = threadIdx.x; e < nJ * norb; e += blockDim.x) {
    Ti[e] = ci[e];
    Tj[e] = cj[e];
  }
  // kets e >= emin taken class by class: the lanes of a wave then mostly share
  // j_k and j_l, hence their J and J' ranges
  for (int c = threadIdx.x; c < ncls; c += blockDim.x) {
    int lo = B.cls_start[g * ncls + c];
    const int end = B.cls_start[g * ncls + c + 1];
    int hi = end;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (B.cls_e[mid] < emin) lo = mid + 1; else hi = mid;
    }
    s_us[c] = lo;
    s_pre[c + 1] = end - lo;
  }
  __syncthreads();
  // inclusive prefix sum of s_pre[1..ncls] by the first wave, 64 classes per pass
  if (threadIdx.x < 64) {
    if (threadIdx.x == 0) s_pre[0] = 0;
    int carry = 0;
    for (int c0 = 0; c0 < ncls; c0 += 64) {
      const int c = c0 + (int)threadIdx.x;
      int v = c < ncls ? s_pre[c + 1] : 0;
      for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if ((int)threadIdx.x >= o) v += u;
      }
      if (c < ncls) s_pre[c + 1] = v + carry;
      carry += __shfl(v, 63, 64);
    }
  }
  __syncthreads();
  const int total = s_pre[ncls];
  const int jji = B.j2[i], jjj = B.j2[j];
  const int li = B.l[i], tzi = B.tz2[i];
  const int d2q = B.d2q;
  const double hf = (P.hZ == 1) ? 1.0 : -1.0;
  const TZ *Zb = reinterpret_cast<const TZ *>(P.Zbar);  // FP32 after the FP32-product ph GEMM
  double *Ztwo = P.Z + B.two_off;
  const IMSRG_G double *sixq = B.sixjQ;  // held in registers: the stores below could alias B
  const IMSRG_G int *kmapT = B.kmapT;
  for (int q0 = 0; q0 < total; q0 += blockDim.x) {
    const int q = q0 + threadIdx.x;
    const bool active = q < total;
    int clo = 0, chi = ncls;  // class of position q: s_pre[clo] <= q < s_pre[chi]
    while (chi - clo > 1) {
      const int mid = (clo + chi) >> 1;
      if (s_pre[mid] <= q) clo = mid; else chi = mid;
    }
    const int4 rc = B.cls_rec[active ? s_us[clo] + q - s_pre[clo] : 0];
    const int kk = active ? rc.x : kb;
    const int k = active ? rc.y >> 16 : i, l = active ? rc.y & 0xFFFF : j;  // i <= k <= l
    const int jjk = active ? rc.z >> 16 : jji, jjl = active ? rc.z & 0xFFFF : jjj;
    const int lk = active ? rc.w >> 24 : B.l[i], ll = active ? (rc.w >> 16) & 0xFF : B.l[j];
    const int tzk = active ? ((rc.w >> 8) & 0xFF) - 1 : B.tz2[i], tzl = active ? (rc.w & 0xFF) - 1 : B.tz2[j];
    const bool pauli = (i == j) || (k == l);  // kets (p,p) exist for even J only
    int Jlo = max(abs(jji - jjj), abs(jjk - jjl)) / 2, Jhi = min(jji + jjj, jjk + jjl) / 2;
    if (pauli) {
      Jlo += Jlo & 1;
      Jhi -= Jhi & 1;
    }
    const bool live = active && Jlo <= Jhi;
    if (!live) {
      Jlo = 1 << 20;
      Jhi = -1;
    }
    const int wlo = wave_min(Jlo), whi = wave_max(Jhi);
    if (wlo > whi) continue;  // wave-uniform: no lane of this wave has a channel
    // direct term: CC channels of (i,l) for J' in tri(ji,jl) ∩ tri(jk,jj)
    const int jpz = B.ccjmax_ph;  // Zbar is zero for J' above it
    int Dlo = max(abs(jji - jjl), abs(jjk - jjj)) / 2, Dhi = min(min(jji + jjl, jjk + jjj) / 2, jpz);
    if (!live) {
      Dlo = 1 << 20;
      Dhi = -1;
    }
    const int dlo = wave_min(Dlo), dhi = wave_max(Dhi);
    // exchange term: CC channels of (i,k) for J' in tri(jj,jl) ∩ tri(jk,ji)
    const bool exch = live && k != l && i != j;
    int Xlo = max(abs(jjj - jjl), abs(jjk - jji)) / 2, Xhi = min(min(jjj + jjl, jjk + jji) / 2, jpz);
    if (!exch) {
      Xlo = 1 << 20;
      Xhi = -1;
    }
    const int xlo = wave_min(Xlo), xhi = wave_max(Xhi);
    const int pd = (li + ll) & 1, td = abs(tzi - tzl) / 2;
    const int px = (li + lk) & 1, tx = abs(tzi - tzk) / 2;
    // 6j entries of J' and the J pair (c0, c0+1) sit at s6 + J' * s6J + (c0/2) * s6P; the next
    // pair (c0+2, c0+3) one pair stride further
    const long s6P = 2L * B.d1 * B.d1, s6J = s6P * (d2q >> 1);
    const long s6d = sixjQ_index(B.d1, d2q, jji, jjj, 0, jjk, jjl, 0);
    const long s6x = sixjQ_index(B.d1, d2q, jjj, jji, 0, jjk, jjl, 0);
    // tail element T(ik, lj) = phi_ik phi_lj T(lj, ik), phi_pq = (-1)^{(j_p+j_q)/2};
    // head H(ik, lj) = hZ-sign H(lj, ik)
    const double symx = (l > j) ? (double)sgn_parity((jji + jjk) / 2 + (jjl + jjj) / 2) : hf;
    const int dbra = (i == j), dket = (k == l);
    const double nrm = (dbra == dket) ? 1.0 + dbra : SQRT2_D;
    const double pij = sgn_parity((jji + jjj + jjk + jjl) / 2);
    // J chunks start at even J so that each pair of 6j values is one aligned 16-byte load
    for (int c0 = wlo & ~1; c0 <= whi; c0 += kJC) {
      unsigned vm = 0;  // lane's valid J of this chunk
#pragma unroll
      for (int t = 0; t < kJC; ++t) {
        const int J = c0 + t;
        if (J >= Jlo && J <= Jhi && !(pauli && (J & 1))) vm |= 1u << t;
      }
      // acc(J) = sum_J' w_d z_d - (-1)^{(jk+jl)/2 - J} sum_J' w_x z_x, i.e. -(commij - sign commji)
      // before the special cases k == l, i == j (no exchange sum; factor applied below)
      double acc[kJC];
#pragma unroll
      for (int t = 0; t < kJC; ++t) acc[t] = 0.0;
      // symx times -(-1)^{(jk+jl)/2 - J} at J = c0
      const double fx = symx * -(double)sgn_parity((jjk + jjl) / 2 - c0);
      // two J' per step, the loads of both (Zbar element, three 16-byte 6j pairs each)
      // issued before the first use; a J' outside the lane's range loads from a
      // valid address and adds nothing
      for (int Jp0 = dlo; Jp0 <= dhi; Jp0 += 2) {
        if (Jp0 + 1 < Dlo || Jp0 > Dhi) continue;
        bool v[2];
        double z[2];
        double2 w[2][kJC / 2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int Jp = min(Jp0 + u, dhi);
          const int cc = 6 * Jp + 2 * (td + 1) + pd;
          const int nkc = s_ccn[cc];
          const long zo = s_zoff[cc];
          v[u] = Jp0 + u <= dhi && Jp >= Dlo && Jp <= Dhi && zo >= 0;
          const int il = Ti[Jp * norb + l];
          const int kj = Tj[Jp * norb + k] + (k > j ? nkc : 0);
          z[u] = Zb[v[u] ? zo + (long)kj * nkc + il : 0];
          const double *s6 = sixq + s6d + (long)Jp * s6J + (long)(c0 >> 1) * s6P;
#pragma unroll
          for (int hh = 0; hh < kJC / 2; ++hh) w[u][hh] = *(const double2 *)(s6 + hh * s6P);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const double zz = v[u] ? z[u] : 0.0;
#pragma unroll
          for (int hh = 0; hh < kJC / 2; ++hh) {
            acc[2 * hh] += w[u][hh].x * zz;
            acc[2 * hh + 1] += w[u][hh].y * zz;
          }
        }
      }
      for (int Jp0 = xlo; Jp0 <= xhi; Jp0 += 2) {
        if (Jp0 + 1 < Xlo || Jp0 > Xhi) continue;
        bool v[2];
        double z[2];
        double2 w[2][kJC / 2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int Jp = min(Jp0 + u, xhi);
          const int cc = 6 * Jp + 2 * (tx + 1) + px;
          const int nkc = s_ccn[cc];
          const long zo = s_zoff[cc];
          v[u] = Jp0 + u <= xhi && Jp >= Xlo && Jp <= Xhi && zo >= 0;
          const int ik = Ti[Jp * norb + k];
          const int lj = Tj[Jp * norb + l];
          z[u] = Zb[v[u] ? zo + (long)((l > j ? nkc : 0) + ik) * nkc + lj : 0];
          const double *s6 = sixq + s6x + (long)Jp * s6J + (long)(c0 >> 1) * s6P;
#pragma unroll
          for (int hh = 0; hh < kJC / 2; ++hh) w[u][hh] = *(const double2 *)(s6 + hh * s6P);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const double zz = v[u] ? fx * z[u] : 0.0;
#pragma unroll
          for (int hh = 0; hh < kJC / 2; ++hh) {
            acc[2 * hh] += w[u][hh].x * zz;
            acc[2 * hh + 1] -= w[u][hh].y * zz;
          }
        }
      }
      // the ket's local indices in the chunk's channels, all loaded and waited for
      // before the first store: loads and stores share one counter, so a load
      // still in flight behind a store would cost a wait for that store
      int ikv[kJC];
#pragma unroll
      for (int t = 0; t < kJC; ++t) ikv[t] = ((vm >> t) & 1) ? kmapT[(long)(c0 + t) * nk + kk] : -1;
      __builtin_amdgcn_s_waitcnt(0);
#pragma unroll
      for (int t = 0; t < kJC; ++t) {
        if (!((vm >> t) & 1)) continue;
        const int J = c0 + t;
        // commji = commij for k == l, pij commij for i == j (no exchange sum then)
        const double s = sgn_parity((jjk + jjl) / 2 - J);
        const double f = (k == l) ? 1.0 - s : (i == j) ? 1.0 - s * pij : 1.0;
        const double z = acc[t] * f / nrm;
        const int n = s_chn[J];
        const int ib = s_ib[J], ik = ikv[t];
        if (ik == ib)  // diagonal: kept apart (the pp/hh GEMM may still be writing Z's diagonal)
          P.diag[s_chks[J] + max(ib, 0)] = z;
        else
          Ztwo[s_choff[J] + (long)ib * n + ik] = z;
      }
    }
  }
}

// Lower triangle <- hZ * upper triangle; anti-Hermitian diagonal <- 0.
__global__ __launch_bounds__(256) void k_mirror(PairArg pa, const int4 *items) {
  __shared__ double t[kTile][kTile + 1];
  const int4 it = items[blockIdx.x];
  if (it.y < 0) return;  // empty item (padding of an XCD run)
  PairPtrs P = pa.get(it.x);
  const BasisDev &B = *P.basis;
  const int ch = it.y, n = B.ch_n[ch];
  const int R0 = it.z * kTile, S0 = it.w * kTile;
  double *Z2 = P.Z + B.two_off + B.ch_off[ch];
  for (int e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
    const int rr = e & (kTile - 1), cc = e / kTile;
    const int r = R0 + rr, c = S0 + cc;
    if (r < n && c < n) t[cc][rr] = Z2[(long)c * n + r];
  }
  __syncthreads();
  const double hZ = P.hZ;
  for (int e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
    const int rr = e & (kTile - 1), cc = e / kTile;
    const int row = S0 + rr, col = R0 + cc;
    if (row >= n || col >= n) continue;
    if (row > col)
      Z2[(long)col * n + row] = hZ * t[rr][cc];
    else if (row == col && P.hZ == -1)
      Z2[(long)col * n + row] = 0.0;
  }
}

// Element idx of a channel block through a 32-bit byte offset (global loads and
// stores then take the block base in SGPRs); a block holds < 2^29 doubles.
template <class T>
__device__ __forceinline__ T &blk_at(T *base, unsigned idx) {
  return *(T *)((char *)base + (idx << 3));
}
template <class T>
__device__ __forceinline__ const T &blk_at(const T *base, unsigned idx) {
  return *(const T *)((const char *)base + (idx << 3));
}

// W of comm122ss (src/Commutator.cc:908-926) by column bins:
//   W(r, q) = sum over the partner entries e of ket q of Y2(r, ind_e) x_e - X2(r, ind_e) y_e
// (x_e, y_e from k_c122_coef).  Every partner of a bin's kets lies in the bin,
// so a work-group stages the bin's columns of X2 and Y2 for kC122Rows rows in
// LDS (each column segment read once, coalesced), together with the bin's
// partner entries (position in the bin, coefficients); each lane then forms W
// for one row of every column of the bin, reading only LDS, and W is stored in
// the two-body layout.
__host__ __device__ inline size_t c122w_lds_bytes(int maxbin, int maxent) {
  return 2 * sizeof(double) * kC122Rows * (size_t)maxbin + (sizeof(double2) + sizeof(int)) * (size_t)maxent +
         sizeof(int) * (size_t)maxbin;
}
__global__ __launch_bounds__(256) void k_c122_w(PairArg pa, const int4 *items, int maxbin, int maxent) {
  static_assert(kC122Rows == 64, "one row per lane of a wave");
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double(*sx)[kC122Rows] = reinterpret_cast<double(*)[kC122Rows]>(lds_raw);  // [maxbin][kC122Rows]
  double(*sy)[kC122Rows] = sx + maxbin;
  double2 *sc = reinterpret_cast<double2 *>(sy + maxbin);  // [maxent] coefficients (x_e, y_e)
  int *sp = reinterpret_cast<int *>(sc + maxent);          // [maxent] partner position in the bin
  int *se = sp + maxent;                                   // [maxbin] end of each column's entries
  const int4 it = items[blockIdx.x];
  PairPtrs P = pa.get(it.x);
  const BasisDev &B = *P.basis;
  const int ch = it.y, n = B.ch_n[ch];
  const int b0 = B.c122_boff[it.z], nc = B.c122_boff[it.z + 1] - b0;
  const int eo = B.c122_beoff[it.z], ne = B.c122_beoff[it.z + 1] - eo;
  const IMSRG_G int *cols = B.c122_bcols + b0;
  const int R0 = it.w * kC122Rows;
  const long off = B.ch_off[ch];
  const double *X2 = P.X + B.two_off + off, *Y2 = P.Y + B.two_off + off;
  double *W = P.c122w + off;
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned r = (unsigned)min(R0 + lane, n - 1);
  for (int t = threadIdx.x; t < ne; t += blockDim.x) {
    const int e = B.c122_bent[eo + t];
    sp[t] = B.c122_gpos[e];
    sc[t] = P.c122[e];
  }
  for (int t = threadIdx.x; t < nc; t += blockDim.x) se[t] = B.c122_bcend[b0 + t];
  // columns: up to 16 per wave and pass, every load issued before the LDS
  // writes (a loop that wrote each column before loading the next waited for
  // each load in turn)
  for (int cb = 0; cb < nc; cb += 64) {
    double xr[16], yr[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int c = min(cb + wv + 4 * m, nc - 1);
      const unsigned q = (unsigned)cols[c];
      xr[m] = blk_at(X2, q * n + r);
      yr[m] = blk_at(Y2, q * n + r);
    }
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int c = cb + wv + 4 * m;
      if (c < nc) {
        sx[c][lane] = xr[m];
        sy[c][lane] = yr[m];
      }
    }
  }
  __syncthreads();
  const bool row_ok = R0 + lane < n;
  static_assert(kC122Blk == 4, "partner blocks of four entries");
  for (int cp = wv; cp < nc; cp += 4) {
    const int q = cols[cp];
    const int ea = cp ? se[cp - 1] : 0, eb = se[cp];
    double w = 0.0;
    for (int t = ea; t < eb; t += 4) {  // a column's list is a multiple of four entries
      int pp[4];
      double2 cf[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        pp[m] = sp[t + m];
        cf[m] = sc[t + m];
      }
      double yv[4], xv[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        yv[m] = sy[pp[m]][lane];
        xv[m] = sx[pp[m]][lane];
      }
#pragma unroll
      for (int m = 0; m < 4; ++m) w += yv[m] * cf[m].x - xv[m] * cf[m].y;
    }
    if (row_ok) blk_at(W, (unsigned)q * n + r) = w;
  }
}

// 16-byte access to elements idx, idx + 1 of a channel block (8-byte aligned)
__device__ __forceinline__ double2 blk_ld2(const double *base, unsigned idx) {
  return *reinterpret_cast<const double2 *>(&blk_at(base, idx));
}
__device__ __forceinline__ void blk_st2(double *base, unsigned idx, double2 v) {
  *reinterpret_cast<double2 *>(&blk_at(base, idx)) = v;
}

// Last kernel of a commutator, over the tile pairs of every two-body block:
//   FOLD: upper triangle += the transposed-slot contributions of k_invpandya,
//         diagonal += its diagonal terms (kept in P.diag);
//   C122: comm122ss (src/Commutator.cc:837-932), Z2 -= W + hZ W^T, with W
//         formed beforehand by k_c122_w;
// then lower triangle <- hZ * upper triangle; anti-Hermitian diagonal <- 0.
// Tile pair: U = rows R0.., cols S0.. (R0 <= S0) and its transposed position.
// U stays in registers, two rows per thread in 8 of the wave's 16 columns.  The
// contributions held at the transposed position go through LDS, two rows per
// thread of 8 of the wave's 16 columns R0 + rr of Z2.  Loads and stores of
// Z2 then move 16 bytes per lane (address processing per instruction, not
// bytes, bounds these accesses).
template <bool FOLD, bool C122>
__global__ __launch_bounds__(256, 4) void k_finish(PairArg pa, const int4 *items) {
  __shared__ double t[kTile][kTile + 1];  // t[rr][cc]: contribution to upper (R0 + rr, S0 + cc)
  const int4 it = items[blockIdx.x];
  if (it.y < 0) return;  // empty item (padding of an XCD run)
  PairPtrs P = pa.get(it.x);
  const BasisDev &B = *P.basis;
  const int ch = it.y, n = B.ch_n[ch];
  const int R0 = it.z * kTile, S0 = it.w * kTile;
  const long off = B.ch_off[ch];
  double *Z2 = P.Z + B.two_off + off;
  const double *W2 = P.c122w + off;                    // comm122 W of this channel (k_c122_w)
  const IMSRG_G double *dg = P.diag + B.ch_kstart[ch];  // inverse-Pandya diagonal terms of this channel
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // lane (h, p): rows p2, p2 + 1 (p2 = 2p) of the tile, in the columns (U) or
  // rows (T) wv + 4 (2 j + h), j < kPairs: two elements per load and store
  const int h = lane >> 5, p2 = 2 * (lane & 31);
  const double hZ = P.hZ;
  constexpr int kPairs = kTile * kTile / 512;
  double2 u[kPairs];
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {  // upper elements (R0 + p2 (+1), S0 + cc)
    const int cc = wv + 4 * (2 * j + h), r = R0 + p2, c = S0 + cc;
    // clamped: no branch around the load (a pair may read one element past its column)
    const double2 z = blk_ld2(Z2, (unsigned)(min(c, n - 1) * n + min(r, n - 1)));
    u[j].x = (r < n && c < n) ? z.x : 0.0;
    u[j].y = (r + 1 < n && c < n) ? z.y : 0.0;
    if (C122) {  // Z2 -= W at the upper element
      const double2 w = blk_ld2(W2, (unsigned)(min(c, n - 1) * n + min(r, n - 1)));
      u[j].x -= w.x;
      u[j].y -= w.y;
    }
  }
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {  // transposed positions (S0 + p2 (+1), R0 + rr)
    const int rr = wv + 4 * (2 * j + h), r = R0 + rr, c = S0 + p2;
    double2 v = {0.0, 0.0};
    if (FOLD) {
      const double2 z = blk_ld2(Z2, (unsigned)(min(r, n - 1) * n + min(c, n - 1)));
      v.x = (r < n && c < n && c != r) ? z.x : 0.0;
      v.y = (r < n && c + 1 < n && c + 1 != r) ? z.y : 0.0;
    }
    if (C122) {  // and -= hZ W at the transposed position
      const double2 w = blk_ld2(W2, (unsigned)(min(r, n - 1) * n + min(c, n - 1)));
      v.x -= hZ * w.x;
      v.y -= hZ * w.y;
    }
    t[rr][p2] = v.x;
    t[rr][p2 + 1] = v.y;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    const int cc = wv + 4 * (2 * j + h), r = R0 + p2, c = S0 + cc;
    double2 v = {u[j].x + t[p2][cc], u[j].y + t[p2 + 1][cc]};
    const bool ok0 = r < n && c < n && r <= c, ok1 = r + 1 < n && c < n && r + 1 <= c;  // ok1 implies ok0
    // diagonal elements inside the block only: dg holds n entries of this channel
    if (ok0 && r == c) v.x = (P.hZ == -1) ? 0.0 : v.x + (FOLD ? dg[r] : 0.0);
    if (ok1 && r + 1 == c) v.y = (P.hZ == -1) ? 0.0 : v.y + (FOLD ? dg[r + 1] : 0.0);
    if (ok1)
      st_nt2(&blk_at(Z2, (unsigned)(c * n + r)), v.x, v.y);
    else if (ok0)
      st_nt(&blk_at(Z2, (unsigned)(c * n + r)), v.x);
    u[j] = v;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {  // new upper values to the threads writing their mirror
    const int cc = wv + 4 * (2 * j + h);
    t[p2][cc] = u[j].x;
    t[p2 + 1][cc] = u[j].y;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {  // lower elements (S0 + p2 (+1), R0 + rr) = hZ * upper
    const int rr = wv + 4 * (2 * j + h), r = R0 + rr, c = S0 + p2;
    const bool ok0 = r < n && c < n && r < c, ok1 = r < n && c + 1 < n && r < c + 1;
    const double2 v = {hZ * t[rr][p2], hZ * t[rr][p2 + 1]};
    if (ok0 && ok1)
      st_nt2(&blk_at(Z2, (unsigned)(r * n + c)), v.x, v.y);
    else if (ok0)
      st_nt(&blk_at(Z2, (unsigned)(r * n + c)), v.x);
    else if (ok1)
      st_nt(&blk_at(Z2, (unsigned)(r * n + c + 1)), v.y);
  }
}

// Z2 <- 0 over whole tiles (used when the pp/hh term is switched off)
__global__ __launch_bounds__(256) void k_zero_two(PairArg pa, const int4 *items) {
  const int4 it = items[blockIdx.x];
  if (it.y < 0) return;  // empty item (padding of an XCD run)
  PairPtrs P = pa.get(it.x);
  const BasisDev &B = *P.basis;
  const int ch = it.y, n = B.ch_n[ch];
  const int R0 = it.z * kTile, S0 = it.w * kTile;
  double *Z2 = P.Z + B.two_off + B.ch_off[ch];
  for (int e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
    const int r = R0 + (e & (kTile - 1)), c = S0 + e / kTile;
    if (r < n && c < n) {
      Z2[(long)c * n + r] = 0.0;
      Z2[(long)r * n + c] = 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
hipError_t launch_zerobody(const PairArg &pa, int npairs, int maxnch, hipStream_t s) {
  if (npairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zerobody, dim3(maxnch + 1, npairs), dim3(256), 0, s, pa);
  hipLaunchKernelGGL(k_zerobody_final, dim3(npairs), dim3(64), 0, s, pa);
  return hipGetLastError();
}
hipError_t launch_onebody(const PairArg &pa, const int4 *items, int nitems, const int4 *items221, int n221,
                          int npairs, int maxn1pairs, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_onebody, dim3(nitems), dim3(256), 0, s, pa, items);
  if (n221 > 0) hipLaunchKernelGGL(k_onebody221, dim3(n221), dim3(256), 0, s, pa, items221);
  hipLaunchKernelGGL(k_onebody_final, dim3((maxn1pairs + 3) / 4, npairs), dim3(256), 0, s, pa);
  return hipGetLastError();
}
hipError_t launch_pphh(const PairArg &pa, const int4 *items, int nitems, const int4 *items64, int nitems64,
                       const int4 *pack, int npack, int precision, int maxn, hipStream_t s) {
  if (nitems > 0) {
    const size_t lds = sizeof(int) * (size_t)std::max(maxn, 1);  // column-index list of the largest channel
    if (precision == kGemmB3) {
      if (npack > 0) hipLaunchKernelGGL(k_pack_pphh, dim3(npack), dim3(256), 0, s, pa, pack);
      hipLaunchKernelGGL(k_pphh_big_b3p, dim3(nitems), dim3(kGWaves * 64), 0, s, pa, items);
    }
    else if (precision == 1)
      hipLaunchKernelGGL(k_pphh_big_mixed, dim3(nitems), dim3(kGWaves * 64), lds, s, pa, items);
    else if (precision == 2)
      hipLaunchKernelGGL(k_pphh_big_split, dim3(nitems), dim3(kGWaves * 64), lds, s, pa, items);
    else
      hipLaunchKernelGGL(k_pphh_big_f64, dim3(nitems), dim3(kGWaves * 64), lds, s, pa, items);
  }
  if (nitems64 > 0) {  // split and bf16x3 products take FP64 64x64 tiles
    if (precision == 1)
      hipLaunchKernelGGL((k_syr2k<0, true>), dim3(nitems64), dim3(256), 0, s, pa, items64, 0);
    else
      hipLaunchKernelGGL((k_syr2k<0, false>), dim3(nitems64), dim3(256), 0, s, pa, items64, 0);
  }
  return hipGetLastError();
}
hipError_t launch_ph_gemm(const PairArg &pa, const int4 *items, int nitems, int precision, bool big, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  if (big) {
    if (precision == kGemmB3) {
      hipLaunchKernelGGL((k_ph_big_b3<1>), dim3(nitems), dim3(kGWaves * 64), 0, s, pa, items);
      hipLaunchKernelGGL((k_ph_big_b3<2>), dim3(nitems), dim3(kGWaves * 64), 0, s, pa, items);
    } else if (precision == 1) {
      hipLaunchKernelGGL((k_ph_big<1, true>), dim3(nitems), dim3(kGWaves * 64), 0, s, pa, items);
      hipLaunchKernelGGL((k_ph_big<2, true>), dim3(nitems), dim3(kGWaves * 64), 0, s, pa, items);
    } else if (precision == 2) {
      hipLaunchKernelGGL((k_ph_big<1, true, true>), dim3(nitems), dim3(kGWaves * 64), 0, s, pa, items);
      hipLaunchKernelGGL((k_ph_big<2, true, true>), dim3(nitems), dim3(kGWaves * 64), 0, s, pa, items);
    } else {
      hipLaunchKernelGGL((k_ph_big<1, false>), dim3(nitems), dim3(kGWaves * 64), 0, s, pa, items);
      hipLaunchKernelGGL((k_ph_big<2, false>), dim3(nitems), dim3(kGWaves * 64), 0, s, pa, items);
    }
    return hipGetLastError();
  }
  if (precision == 1) {
    hipLaunchKernelGGL((k_syr2k<1, true>), dim3(nitems), dim3(256), 0, s, pa, items, 0);
    hipLaunchKernelGGL((k_syr2k<2, true>), dim3(nitems), dim3(256), 0, s, pa, items, 0);
  } else {
    hipLaunchKernelGGL((k_syr2k<1, false>), dim3(nitems), dim3(256), 0, s, pa, items, 0);
    hipLaunchKernelGGL((k_syr2k<2, false>), dim3(nitems), dim3(256), 0, s, pa, items, 0);
  }
  return hipGetLastError();
}
hipError_t launch_c122_coef(const PairArg &pa, int npairs, long maxn122, hipStream_t s) {
  if (maxn122 <= 0) return hipSuccess;
  int cb = (int)std::max<long>(1, std::min<long>((maxn122 + 255) / 256, 4096));
  hipLaunchKernelGGL(k_c122_coef, dim3(cb, npairs), dim3(256), 0, s, pa, npairs);
  return hipGetLastError();
}
hipError_t launch_pandya(const PairArg &pa, const int4 *items, int nitems, int maxnch, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  const size_t lds = (sizeof(long) + sizeof(int)) * (size_t)std::max(maxnch, 1);
  hipLaunchKernelGGL(k_pandya, dim3(nitems), dim3(kChunk), lds, s, pa, items);
  return hipGetLastError();
}
hipError_t launch_invpandya(const PairArg &pa, const int4 *items, int nitems, int jmax, int norb, int nch,
                            int ncls, bool zbar_f32, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  const size_t lds = invp_lds_bytes(jmax, norb, nch, ncls);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  if (zbar_f32)
    hipLaunchKernelGGL(k_invpandya<float>, dim3(nitems), dim3(kChunk), lds, s, pa, items);
  else
    hipLaunchKernelGGL(k_invpandya<double>, dim3(nitems), dim3(kChunk), lds, s, pa, items);
  return hipGetLastError();
}
hipError_t launch_finish(const PairArg &pa, const int4 *items, int nitems, bool fold, bool c122, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  if (fold && c122)
    hipLaunchKernelGGL((k_finish<true, true>), dim3(nitems), dim3(256), 0, s, pa, items);
  else if (fold)
    hipLaunchKernelGGL((k_finish<true, false>), dim3(nitems), dim3(256), 0, s, pa, items);
  else if (c122)
    hipLaunchKernelGGL((k_finish<false, true>), dim3(nitems), dim3(256), 0, s, pa, items);
  else
    hipLaunchKernelGGL(k_mirror, dim3(nitems), dim3(256), 0, s, pa, items);
  return hipGetLastError();
}
hipError_t launch_c122_w(const PairArg &pa, const int4 *items, int nitems, int maxbin, int maxent, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  const size_t lds = c122w_lds_bytes(maxbin, maxent);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_c122_w, dim3(nitems), dim3(256), lds, s, pa, items, maxbin, maxent);
  return hipGetLastError();
}
hipError_t launch_zero_two(const PairArg &pa, const int4 *items, int nitems, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_two, dim3(nitems), dim3(256), 0, s, pa, items);
  return hipGetLastError();
}

}  // namespace imsrg
