// kernels_util.hip -- 6j table, operator arithmetic and norms, normal ordering,
// basis transformation, Hartree-Fock monopoles, synthetic operators, White
// generator, MFMA layout probe.
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "launch.hpp"

namespace imsrg {

// ---------------------------------------------------------------------------
// double-double arithmetic (Dekker / Knuth two-sum, fma two-product)
struct dd {
  double hi, lo;
};
__device__ __forceinline__ dd dd_quick(double a, double b) {
  double s = a + b;
  return {s, b - (s - a)};
}
__device__ __forceinline__ dd dd_two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ dd dd_add(dd x, dd y) {
  dd s = dd_two_sum(x.hi, y.hi);
  dd t = dd_two_sum(x.lo, y.lo);
  s.lo += t.hi;
  s = dd_quick(s.hi, s.lo);
  s.lo += t.lo;
  return dd_quick(s.hi, s.lo);
}
__device__ __forceinline__ dd dd_neg(dd x) { return {-x.hi, -x.lo}; }
__device__ __forceinline__ dd dd_mul(dd x, dd y) {
  double p = x.hi * y.hi;
  double e = __fma_rn(x.hi, y.hi, -p);
  e += x.hi * y.lo + x.lo * y.hi;
  return dd_quick(p, e);
}
__device__ __forceinline__ dd dd_mul_d(dd x, double y) {
  double p = x.hi * y;
  double e = __fma_rn(x.hi, y, -p);
  e += x.lo * y;
  return dd_quick(p, e);
}
__device__ __forceinline__ dd dd_div(dd x, dd y) {
  double q1 = x.hi / y.hi;
  dd r = dd_add(x, dd_neg(dd_mul_d(y, q1)));
  double q2 = r.hi / y.hi;
  r = dd_add(r, dd_neg(dd_mul_d(y, q2)));
  double q3 = r.hi / y.hi;
  return dd_add(dd_quick(q1, q2), dd{q3, 0.0});
}
__device__ __forceinline__ dd dd_sqrt(dd a) {
  if (a.hi <= 0.0) return {0.0, 0.0};
  double x = sqrt(a.hi);
  dd xx = dd_mul(dd{x, 0.0}, dd{x, 0.0});
  dd r = dd_add(a, dd_neg(xx));
  return dd_two_sum(x, r.hi / (2.0 * x));
}

constexpr int kNFact = 256;

__global__ void k_factorials(dd *f) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    dd v = {1.0, 0.0};
    f[0] = v;
    for (int i = 1; i < kNFact; ++i) {
      v = dd_mul_d(v, (double)i);
      f[i] = v;
    }
  }
}

__device__ __forceinline__ bool tri_ok(int a, int b, int c) {
  return ((a + b + c) & 1) == 0 && c >= abs(a - b) && c <= a + b;
}

// Wigner 6j {a b c; d e f} (doubled arguments), Racah single-sum formula
// (Edmonds eq. 6.3.7) in double-double; replaces gsl_sf_coupling_6j
// (src/AngMom.cc:94-97).
__device__ double sixj_dd(const dd *F, int a, int b, int c, int d, int e, int f) {
  if (!tri_ok(a, b, c) || !tri_ok(a, e, f) || !tri_ok(d, b, f) || !tri_ok(d, e, c)) return 0.0;
  const int tri[4][3] = {{a, b, c}, {a, e, f}, {d, b, f}, {d, e, c}};
  // product of the four triangle coefficients, each formed as its own ratio
  // (at most 1): the four denominators multiplied first overflow a double
  // for j = 29/2, J = 29 (emax = 14)
  dd d2 = {1.0, 0.0};
  for (int k = 0; k < 4; ++k) {
    int x = tri[k][0], y = tri[k][1], z = tri[k][2];
    const dd num = dd_mul(dd_mul(F[(x + y - z) / 2], F[(x - y + z) / 2]), F[(-x + y + z) / 2]);
    d2 = dd_mul(d2, dd_div(num, F[(x + y + z) / 2 + 1]));
  }
  dd pref = dd_sqrt(d2);
  const int al[4] = {(a + b + c) / 2, (a + e + f) / 2, (d + b + f) / 2, (d + e + c) / 2};
  const int be[3] = {(a + b + d + e) / 2, (b + c + e + f) / 2, (a + c + d + f) / 2};
  int tmin = max(max(al[0], al[1]), max(al[2], al[3]));
  int tmax = min(min(be[0], be[1]), be[2]);
  dd sum = {0.0, 0.0};
  for (int t = tmin; t <= tmax; ++t) {
    dd dn = F[t - al[0]];
    for (int k = 1; k < 4; ++k) dn = dd_mul(dn, F[t - al[k]]);
    for (int k = 0; k < 3; ++k) dn = dd_mul(dn, F[be[k] - t]);
    dd term = dd_div(F[t + 1], dn);
    if (t & 1) term = dd_neg(term);
    sum = dd_add(sum, term);
  }
  dd r = dd_mul(pref, sum);
  return r.hi + r.lo;
}

__global__ __launch_bounds__(256) void k_sixj_table(const dd *F, double *table, double *tableQ, int d1, int d2,
                                                   int d2q, long total) {
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    long r = idx;
    const int J6 = r % d2; r /= d2;
    const int i5 = r % d1; r /= d1;
    const int i4 = r % d1; r /= d1;
    const int J3 = r % d2; r /= d2;
    const int i2 = r % d1; r /= d1;
    const int i1 = (int)r;
    const double v = sixj_dd(F, 2 * i1 + 1, 2 * i2 + 1, 2 * J3, 2 * i4 + 1, 2 * i5 + 1, 2 * J6);
    table[idx] = v;
    // the inverse Pandya copy holds the weights (2 J6 + 1) {...}, and 0 where
    // |6j| < 1e-8, the values whose terms the reference skips (src/Commutator.cc:1564-1572,
    // 1593-1604)
    tableQ[sixjQ_index(d1, d2q, 2 * i1 + 1, 2 * i2 + 1, J3, 2 * i4 + 1, 2 * i5 + 1, J6)] =
        fabs(v) >= 1e-8 ? (2 * J6 + 1) * v : 0.0;
  }
}

hipError_t launch_sixj_table(double *table, double *tableQ, int d1, int d2, int d2q, hipStream_t s) {
  if (4 * (d1 - 1) + 4 > kNFact - 4) return hipErrorInvalidValue;
  dd *F = nullptr;
  hipError_t e = hipMalloc(&F, sizeof(dd) * kNFact);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_factorials, dim3(1), dim3(64), 0, s, F);
  long total = (long)d1 * d1 * d2 * d1 * d1 * d2;
  int blocks = (int)std::min<long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_sixj_table, dim3(blocks), dim3(256), 0, s, F, table, tableQ, d1, d2, d2q, total);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(s);
  (void)hipFree(F);
  return e;
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_axpy(double *y, double a, const double *x, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double t = a * x[i];  // (a*x) formed first, then added (Operator operator* then +=)
    y[i] += t;
  }
}
__global__ __launch_bounds__(256) void k_scale(double *y, double a, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] *= a;
}
// y[i] = v (clears memory that may hold NaN or Inf, which a scale by 0 keeps)
__global__ __launch_bounds__(256) void k_fill(double *y, double v, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = v;
}
static int grid_for(long n) { return (int)std::max<long>(1, std::min<long>((n + 255) / 256, 8192)); }
hipError_t launch_fill(double *y, double v, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, s, y, v, n);
  return hipGetLastError();
}
hipError_t launch_axpy(double *y, double a, const double *x, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(256), 0, s, y, a, x, n);
  return hipGetLastError();
}
hipError_t launch_scale(double *y, double a, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, s, y, a, n);
  return hipGetLastError();
}

// Norm partial sums.  Block b < npart-1 covers a slice of the two-body arena
// (weights (2J+1)^2 of the element's channel); the last block does the
// one-body part with weights (2 j_p + 1).
__global__ __launch_bounds__(256) void k_norm_partial(const BasisDev *bd, const double *op, double *part, int nblk2) {
  const BasisDev &B = *bd;
  __shared__ double red[4];
  double acc = 0.0;
  if ((int)blockIdx.x == nblk2) {
    const double *one = op + B.one_off;
    for (long e = threadIdx.x; e < (long)B.norb * B.norb; e += blockDim.x) acc += one[e] * one[e] * B.w1norm[e];
  } else {
    // block b takes the b-th contiguous chunk of the two-body arena and the
    // channels it overlaps, each with its weight (2J+1)^2
    const double *two = op + B.two_off;
    const long tb = B.tbsize, chunk = (tb + nblk2 - 1) / nblk2;
    const long c0 = (long)blockIdx.x * chunk, c1 = min(c0 + chunk, tb);
    int lo = 0, hi = B.nch - 1;  // last channel starting at or before c0
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (B.ch_off[mid] <= c0) lo = mid; else hi = mid - 1;
    }
    for (int ch = lo; ch < B.nch && B.ch_off[ch] < c1; ++ch) {
      const long off = B.ch_off[ch], nn = (long)B.ch_n[ch] * B.ch_n[ch];
      const long e0 = max(c0, off), e1 = min(c1, off + nn);
      const double w = (double)((2 * (ch / 6) + 1) * (2 * (ch / 6) + 1));
      double s = 0.0;
      for (long e = e0 + threadIdx.x; e < e1; e += blockDim.x) s += two[e] * two[e];
      acc += w * s;
    }
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// y += a x over the whole operator buffer, and the norm partial sums of x with the
// blocks of k_norm_partial (the same sums in the same order): one pass over x in
// place of an axpy and a norm (the BCH_Transform step, src/BCH.cc:176-191)
__global__ __launch_bounds__(256) void k_axpy_norm(const BasisDev *bd, double *y, double a, const double *x,
                                                   double *part, int nblk2) {
  const BasisDev &B = *bd;
  __shared__ double red[4];
  double acc = 0.0;
  if ((int)blockIdx.x == nblk2) {
    const long n1 = (long)B.norb * B.norb;
    // the zero-body and padding words around the one-body block and after the two-body arena
    const long rg[3][2] = {{0, B.one_off}, {B.one_off + n1, B.two_off}, {B.two_off + B.tbsize, B.opsize}};
    for (int g = 0; g < 3; ++g)
      for (long e = rg[g][0] + threadIdx.x; e < rg[g][1]; e += blockDim.x) {
        double t = a * x[e];
        y[e] += t;
      }
    const double *one = x + B.one_off;
    double *yo = y + B.one_off;
    for (long e = threadIdx.x; e < n1; e += blockDim.x) {
      const double v = one[e];
      acc += v * v * B.w1norm[e];
      double t = a * v;
      yo[e] += t;
    }
  } else {
    const double *two = x + B.two_off;
    double *y2 = y + B.two_off;
    const long tb = B.tbsize, chunk = (tb + nblk2 - 1) / nblk2;
    const long c0 = (long)blockIdx.x * chunk, c1 = min(c0 + chunk, tb);
    int lo = 0, hi = B.nch - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (B.ch_off[mid] <= c0) lo = mid; else hi = mid - 1;
    }
    for (int ch = lo; ch < B.nch && B.ch_off[ch] < c1; ++ch) {
      const long off = B.ch_off[ch], nn = (long)B.ch_n[ch] * B.ch_n[ch];
      const long e0 = max(c0, off), e1 = min(c1, off + nn);
      const double w = (double)((2 * (ch / 6) + 1) * (2 * (ch / 6) + 1));
      double s = 0.0;
      for (long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const double v = two[e];
        s += v * v;
        double t = a * v;
        y2[e] += t;
      }
      acc += w * s;
    }
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void k_norm_final(const double *part, int nblk2, double *out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nblk2; i += blockDim.x) acc += part[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = part[nblk2];
    out[1] = red[0];
  }
}

// Operator::DoNormalOrdering2 of a scalar operator (src/Operator.cc:461-579) in
// place, sign +1 (normal ordering) or -1 (undoing it), over the reference's holes.
// Zero-body part (:467-490): E0 += sign sum_k (2j_k+1) n_k h_kk + sum_{k<=l} sum_J
// (2J+1) n_k n_l V^J_{kl,kl} (normalized), one thread per hole pair and a
// fixed-order block sum; it reads the one-body diagonal before k_no_one changes it.
__global__ __launch_bounds__(256) void k_no_zero(const BasisDev *Bp, double *op, int sign) {
  __shared__ double red[4];
  const BasisDev &B = *Bp;
  const int norb = B.norb, nh = B.nholes;
  const double *one = op + B.one_off, *two = op + B.two_off;
  double acc = 0.0;
  for (int t = threadIdx.x; t < nh * nh; t += blockDim.x) {
    const int ik = t / nh, il = t - ik * nh;
    const int k = B.holes[ik], l = B.holes[il];
    if (ik == il) acc += (B.j2[k] + 1) * sign * B.occ[k] * one[(long)k * norb + k];
    if (l < k) continue;
    const int par = (B.l[k] + B.l[l]) & 1, tz = (B.tz2[k] + B.tz2[l]) / 2;
    const int kl = ket_index(norb, k, l);
    for (int J = abs(B.j2[k] - B.j2[l]) / 2; J <= (B.j2[k] + B.j2[l]) / 2; ++J) {
      const int i = B.kmapT[(long)J * B.nkets + kl];  // -1: (k,k) at odd J
      if (i < 0) continue;
      const int ch = 6 * J + 2 * (tz + 1) + par, n = B.ch_n[ch];
      acc += (2 * J + 1) * B.occ[k] * B.occ[l] * two[B.ch_off[ch] + (long)i * n + i];
    }
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) op[0] += (red[0] + red[1]) + (red[2] + red[3]);
}

// One-body part (:493-560) and the (anti-)Hermitian completion (Symmetrize /
// AntiSymmetrize, :1797-1830): one wave per pair a <= b of a one-body block, one
// lane per hole h,
//   f_ab += sign sum_J (2J+1)/(2j_a+1) n_h V~^J_{ah,bh}
// with the tilde element (sqrt 2 per ket with equal orbits, (-1)^{(j_p+j_q)/2+1+J}
// per swapped ket), summed over the lanes in a fixed order.  The sum is skipped
// when the two-body norm (n2sq = its square, on the device) is at most 1e-7.
__global__ __launch_bounds__(256) void k_no_one(const BasisDev *Bp, double *op, int sign, int herm,
                                                const double *n2sq) {
  const BasisDev &B = *Bp;
  const int p = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (p >= B.n1pairs) return;  // wave-uniform
  const int2 pr = B.pairs1[p];
  const int a = pr.x, b = pr.y, norb = B.norb;
  double *one = op + B.one_off;
  const double *two = op + B.two_off;
  double acc = 0.0;
  if (*n2sq > 1e-14) {
    const double w0 = 1.0 / (B.j2[a] + 1.0);  // 1 / (2 j_a + 1)
    for (int ih = lane; ih < B.nholes; ih += 64) {
      const int h = B.holes[ih];
      const int par = (B.l[a] + B.l[h]) & 1, tz = (B.tz2[a] + B.tz2[h]) / 2;
      const int kah = ket_index(norb, min(a, h), max(a, h)), kbh = ket_index(norb, min(b, h), max(b, h));
      const double nrm = (a == h ? 1.4142135623730951 : 1.0) * (b == h ? 1.4142135623730951 : 1.0);
      for (int J = abs(B.j2[a] - B.j2[h]) / 2; J <= (B.j2[a] + B.j2[h]) / 2; ++J) {
        const int i = B.kmapT[(long)J * B.nkets + kah], k = B.kmapT[(long)J * B.nkets + kbh];
        if (i < 0 || k < 0) continue;
        const int ch = 6 * J + 2 * (tz + 1) + par, n = B.ch_n[ch];
        double ph = 1.0;
        if (a > h) ph *= ((B.j2[a] + B.j2[h]) / 2 + 1 + J) & 1 ? -1.0 : 1.0;
        if (b > h) ph *= ((B.j2[b] + B.j2[h]) / 2 + 1 + J) & 1 ? -1.0 : 1.0;
        acc += (2 * J + 1) * w0 * sign * B.occ[h] * (nrm * ph * two[B.ch_off[ch] + (long)k * n + i]);
      }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  }
  if (lane != 0) return;
  double f = one[(long)b * norb + a] + acc;
  if (herm == -1 && a == b) f = 0.0;
  one[(long)b * norb + a] = f;
  if (a != b) one[(long)a * norb + b] = herm * f;
}

hipError_t launch_normal_order(const BasisDev *bd, const BasisDev &bh, double *op, int sign, int herm,
                               const double *n2sq, hipStream_t s) {
  hipLaunchKernelGGL(k_no_zero, dim3(1), dim3(256), 0, s, bd, op, sign);
  if (bh.n1pairs > 0)
    hipLaunchKernelGGL(k_no_one, dim3((bh.n1pairs + 3) / 4), dim3(256), 0, s, bd, op, sign, herm, n2sq);
  return hipGetLastError();
}

hipError_t launch_axpy_norm(const BasisDev *bd, double *y, double a, const double *x, double *partials, int npart,
                            double *out, hipStream_t s) {
  const int nblk2 = npart - 1;
  hipLaunchKernelGGL(k_axpy_norm, dim3(npart), dim3(256), 0, s, bd, y, a, x, partials, nblk2);
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(256), 0, s, partials, nblk2, out);
  return hipGetLastError();
}
hipError_t launch_norm(const BasisDev *bd, const BasisDev &bh, const double *op, double *partials, int npart,
                       double *out, hipStream_t s) {
  (void)bh;
  int nblk2 = npart - 1;
  hipLaunchKernelGGL(k_norm_partial, dim3(npart), dim3(256), 0, s, bd, op, partials, nblk2);
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(256), 0, s, partials, nblk2, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The two-body basis-change matrix of HartreeFock::TransformToHFBasis
// (src/HartreeFock.cc:942-960), for every channel at once, in the two-body layout:
// element (i, a) of channel ch's block is
//   D(i, a) = C(p, p') C(q, q') + [p != q] Phase_pq(J) C(q, p') C(p, q'),
// times sqrt2 if p == q, divided by sqrt2 if p' == q', for the channel's kets
// i = (p, q) (old basis, rows) and a = (p', q') (new basis, columns);
// Phase_pq(J) = (-1)^{(j_p + j_q)/2 + 1 + J} (src/ModelSpace.hh:72-78).
// C is norb x norb column-major, C(p, p') = <p|p'>.  One thread per element.
__global__ __launch_bounds__(256) void k_hf_D(const BasisDev *bd, const double *C, double *D) {
  const BasisDev &B = *bd;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= B.tbsize) return;
  int lo = 0, hi = B.nch - 1;  // the last channel starting at or before e holds it
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (B.ch_off[mid] <= e) lo = mid; else hi = mid - 1;
  }
  const int ch = lo, n = B.ch_n[ch], J = ch / 6;
  const long r = e - B.ch_off[ch];
  const int i = (int)(r % n), a = (int)(r / n);
  const int ki = B.ch_kets[B.ch_kstart[ch] + i], ka = B.ch_kets[B.ch_kstart[ch] + a];
  const int p = B.kp[ki], q = B.kq[ki], pp = B.kp[ka], qq = B.kq[ka];
  const long N = B.norb;
  double d = C[p + pp * N] * C[q + qq * N];
  if (p != q) {
    const double ph = (((B.j2[p] + B.j2[q]) / 2 + 1 + J) & 1) ? -1.0 : 1.0;
    double x = C[q + pp * N] * C[p + qq * N];
    d += ph * x;
  }
  if (p == q) d *= 1.4142135623730951;
  if (pp == qq) d /= 1.4142135623730951;
  D[e] = d;
}

// The two-body monopoles of the Hartree-Fock field (HartreeFock::BuildMonopoleV and
// the lookups of UpdateF, src/HartreeFock.cc:250-285, 771-800):
//   W(i a, j b) = sum_J (2J+1) V~^J_{ia,jb}
// with the tilde element (sqrt2 per ket with equal orbits, (-1)^{(j_p+j_q)/2+1+J}
// per swapped ket), J over the couplings of (i, a), summed in ascending J; equal to
// the reference's Vmon / Vmon_exch entry, GetTBMEmonopole times (2j_i+1)(2j_a+1)
// (src/TwoBodyME.cc:614-679).  quads[q] = (i, a, j, b); one thread per quad.
__global__ __launch_bounds__(256) void k_hf_mono(const BasisDev *Bp, const double *op, const int4 *quads, int nq,
                                                 double *W) {
  const BasisDev &B = *Bp;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const int4 t = quads[q];
  const int i = t.x, a = t.y, j = t.z, b = t.w, norb = B.norb;
  const double *two = op + B.two_off;
  const int par = (B.l[i] + B.l[a]) & 1, tz = (B.tz2[i] + B.tz2[a]) / 2;
  double acc = 0.0;
  if (par == ((B.l[j] + B.l[b]) & 1) && tz == (B.tz2[j] + B.tz2[b]) / 2) {
    const int kia = ket_index(norb, min(i, a), max(i, a)), kjb = ket_index(norb, min(j, b), max(j, b));
    const double nrm = (i == a ? 1.4142135623730951 : 1.0) * (j == b ? 1.4142135623730951 : 1.0);
    for (int J = abs(B.j2[i] - B.j2[a]) / 2; J <= (B.j2[i] + B.j2[a]) / 2; ++J) {
      const int r = B.kmapT[(long)J * B.nkets + kia], c = B.kmapT[(long)J * B.nkets + kjb];
      if (r < 0 || c < 0) continue;
      const int ch = 6 * J + 2 * (tz + 1) + par, n = B.ch_n[ch];
      double ph = 1.0;
      if (i > a) ph *= ((B.j2[i] + B.j2[a]) / 2 + 1 + J) & 1 ? -1.0 : 1.0;
      if (j > b) ph *= ((B.j2[j] + B.j2[b]) / 2 + 1 + J) & 1 ? -1.0 : 1.0;
      acc += (2 * J + 1) * (nrm * ph * two[B.ch_off[ch] + (long)c * n + r]);
    }
  }
  W[q] = acc;
}

hipError_t launch_hf_mono(const BasisDev *bd, const double *op, const int4 *quads, int nq, double *W,
                          hipStream_t s) {
  if (nq > 0) hipLaunchKernelGGL(k_hf_mono, dim3((nq + 255) / 256), dim3(256), 0, s, bd, op, quads, nq, W);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Basis transformation with a basis change C that is block diagonal in (l, j, tz),
// as every Hartree-Fock C is.  D(i, a) is then nonzero only when the kets i and a of
// a channel take their orbits from the same pair of one-body blocks, so after
// grouping each channel's kets by that pair, D is block diagonal with groups of at
// most (block size)^2 kets, and V' = D^T V D is two passes over the channel:
//   T  = V D     column a of T mixes the columns of V in a's group   (k_xf_right)
//   V' = D^T T   row a of V' mixes the rows of T in a's group        (k_xf_left)
// Every element of V, T and V' is read or written once from HBM.

// D blocks, one work-group per group g: Dg[doff[g] + r + c m] = D(kets r, c of g)
// (the formula of k_hf_D).
__global__ __launch_bounds__(256) void k_xf_D(const BasisDev *bd, XformDev X, const double *C, double *Dg) {
  const BasisDev &B = *bd;
  const int g = blockIdx.x, s = X.g_start[g], m = X.g_start[g + 1] - s, ch = X.g_ch[g], J = ch / 6;
  const long N = B.norb;
  const int *kets = B.ch_kets + B.ch_kstart[ch];
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
    const int r = e % m, c = e / m;
    const int ki = kets[X.g_kets[s + r]], ka = kets[X.g_kets[s + c]];
    const int p = B.kp[ki], q = B.kq[ki], pp = B.kp[ka], qq = B.kq[ka];
    double d = C[p + pp * N] * C[q + qq * N];
    if (p != q) {
      const double ph = (((B.j2[p] + B.j2[q]) / 2 + 1 + J) & 1) ? -1.0 : 1.0;
      double x = C[q + pp * N] * C[p + qq * N];
      d += ph * x;
    }
    if (p == q) d *= 1.4142135623730951;
    if (pp == qq) d /= 1.4142135623730951;
    Dg[X.g_doff[g] + e] = d;
  }
}

// T = V D on one item (group g, rows r0 .. r0+R-1 of its channel): V's rows in
// the group's columns and the group's D block staged in LDS, then each of the
// R x m outputs is a dot product of length m.
__global__ __launch_bounds__(256) void k_xf_right(const BasisDev *bd, XformDev X, const double *V, const double *Dg,
                                                  double *T, const int2 *items, int R) {
  extern __shared__ double lds[];
  const BasisDev &B = *bd;
  const int2 it = items[blockIdx.x];
  const int g = it.x, r0 = it.y, s = X.g_start[g], m = X.g_start[g + 1] - s, ch = X.g_ch[g];
  const int n = B.ch_n[ch];
  const long off = B.ch_off[ch];
  double *Vs = lds, *Ds = lds + (long)X.maxm * R;
  for (int e = threadIdx.x; e < m * R; e += blockDim.x) {
    const int k = e / R, row = e - k * R;
    Vs[e] = (r0 + row < n) ? V[off + (long)X.g_kets[s + k] * n + r0 + row] : 0.0;
  }
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) Ds[e] = Dg[X.g_doff[g] + e];
  __syncthreads();
  const int row = threadIdx.x % R, lanes = blockDim.x / R;
  if (r0 + row >= n) return;
  for (int c = threadIdx.x / R; c < m; c += lanes) {
    double acc = 0.0;
    for (int k = 0; k < m; ++k) acc += Vs[k * R + row] * Ds[c * m + k];
    T[off + (long)X.g_kets[s + c] * n + r0 + row] = acc;
  }
}

// V' = D^T T on one item (channel ch, columns b0 .. b0+kc-1): the columns of T
// staged in LDS (ket-major, kc values per ket), then row a of each column is the
// dot product of a's D column with the rows of T in a's group.
__global__ __launch_bounds__(256) void k_xf_left(const BasisDev *bd, XformDev X, const double *T, const double *Dg,
                                                 double *Vout, const int2 *items, int kc) {
  extern __shared__ double lds[];
  const BasisDev &B = *bd;
  const int2 it = items[blockIdx.x];
  const int ch = it.x, b0 = it.y, n = B.ch_n[ch];
  const int nb = min(kc, n - b0);
  const long off = B.ch_off[ch];
  for (int e = threadIdx.x; e < n * nb; e += blockDim.x) {
    const int bb = e / n, i = e - bb * n;
    lds[i * kc + bb] = T[off + (long)(b0 + bb) * n + i];
  }
  __syncthreads();
  const int kbase = X.k_start[ch];
  for (int a = threadIdx.x; a < n; a += blockDim.x) {
    const int g = X.k_grp[kbase + a], c = X.k_pos[kbase + a];
    const int s = X.g_start[g], m = X.g_start[g + 1] - s;
    const double *Dc = Dg + X.g_doff[g] + (long)c * m;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < m; ++k) {
      const double d = Dc[k];
      const double *t = lds + X.g_kets[s + k] * kc;
      for (int bb = 0; bb < 4; ++bb)
        if (bb < nb) acc[bb] += d * t[bb];
    }
    for (int bb = 0; bb < 4; ++bb)
      if (bb < nb) Vout[off + (long)(b0 + bb) * n + a] = acc[bb];
  }
}

// One-body part, C^T O C with C block diagonal: out(a, b) = sum over k in a's block
// and l in b's block of C(k, a) O(k, l) C(l, b); one thread per element.
__global__ __launch_bounds__(256) void k_xf_one(const BasisDev *bd, const double *C, const double *O, double *out) {
  const BasisDev &B = *bd;
  const int N = B.norb;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N * N) return;
  const int a = e % N, b = e / N;
  double acc = 0.0;
  for (int ik = B.obstart[a]; ik < B.obstart[a + 1]; ++ik) {
    const int k = B.oblist[ik];
    double t = 0.0;
    for (int il = B.obstart[b]; il < B.obstart[b + 1]; ++il) {
      const int l = B.oblist[il];
      t += O[k + (long)l * N] * C[l + (long)b * N];
    }
    acc += C[k + (long)a * N] * t;
  }
  out[e] = acc;
}

hipError_t launch_xform_blocks(const BasisDev *bd, const BasisDev &bh, const XformDev &X, const double *C,
                               const double *op, double *out, double *T, double *Dg, double *one_tmp,
                               const int2 *right_items, int nright, int R, const int2 *left_items, int nleft, int kc,
                               hipStream_t s) {
  if (X.ngroups > 0) {
    hipLaunchKernelGGL(k_xf_D, dim3(X.ngroups), dim3(256), 0, s, bd, X, C, Dg);
    const size_t lr = sizeof(double) * ((size_t)X.maxm * R + (size_t)X.maxm * X.maxm);
    hipLaunchKernelGGL(k_xf_right, dim3(nright), dim3(256), lr, s, bd, X, op + bh.two_off, Dg, T, right_items, R);
    const size_t ll = sizeof(double) * (size_t)bh.maxn * kc;
    hipLaunchKernelGGL(k_xf_left, dim3(nleft), dim3(256), ll, s, bd, X, T, Dg, out + bh.two_off, left_items, kc);
  }
  const int n1 = bh.norb * bh.norb;
  if (n1 > 0) {
    hipLaunchKernelGGL(k_xf_one, dim3((n1 + 255) / 256), dim3(256), 0, s, bd, C, op + bh.one_off, one_tmp);
    hipError_t e = hipMemcpyAsync(out + bh.one_off, one_tmp, sizeof(double) * n1, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

hipError_t launch_hf_D(const BasisDev *bd, const BasisDev &bh, const double *C, double *D, hipStream_t s) {
  if (bh.tbsize > 0) {
    const long nb = (bh.tbsize + 255) / 256;
    hipLaunchKernelGGL(k_hf_D, dim3((unsigned)nb), dim3(256), 0, s, bd, C, D);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Counter-based N(0,1): SplitMix64 of (seed, index) -> two uniforms -> Box-Muller.
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ double normal_at(unsigned long long seed, unsigned long long idx) {
  unsigned long long a = splitmix64(seed ^ splitmix64(2 * idx + 1));
  unsigned long long b = splitmix64(a ^ 0x5851F42D4C957F2Dull);
  double u1 = ((a >> 11) + 1.0) * (1.0 / 9007199254740993.0);  // (0, 1]
  double u2 = (b >> 11) * (1.0 / 9007199254740992.0);          // [0, 1)
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__global__ __launch_bounds__(256) void k_fill_random(const BasisDev *bd, double *op, int herm, unsigned long long seed) {
  const BasisDev &B = *bd;
  const double inv_sqrt2 = 0.7071067811865476;
  const long n2 = (long)B.norb * B.norb;
  const long total = n2 + B.tbsize;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    if (e < n2) {
      const int p = e % B.norb, q = e / B.norb;
      double v = 0.0;
      if (B.l[p] == B.l[q] && B.j2[p] == B.j2[q] && B.tz2[p] == B.tz2[q]) {
        // the value of the pair is formed once from the canonical (lo, hi) order,
        // so that (p,q) and (q,p) are exactly herm-related
        const int lo = min(p, q), hi = max(p, q);
        const int m = B.obstart[p + 1] - B.obstart[p];
        double a = normal_at(seed, (unsigned long long)((long)hi * B.norb + lo));
        double b = normal_at(seed, (unsigned long long)((long)lo * B.norb + hi));
        double vc = (lo == hi) ? (herm == 1 ? 2.0 * a : 0.0) : (a + herm * b);
        vc = vc * inv_sqrt2 / sqrt((double)m);
        v = (p <= q) ? vc : herm * vc;
      }
      op[B.one_off + e] = v;
    } else {
      const long t = e - n2;
      int clo = 0, chi = B.nch - 1;  // last channel starting at or before t (non-empty)
      while (clo < chi) {
        int mid = (clo + chi + 1) / 2;
        if (B.ch_off[mid] <= t) clo = mid; else chi = mid - 1;
      }
      const int ch = clo;
      const int n = B.ch_n[ch];
      const long loc = t - B.ch_off[ch];
      const int r = loc % n, c = loc / n;
      const int lo = min(r, c), hi = max(r, c);
      double a = normal_at(seed, (unsigned long long)(n2 + B.ch_off[ch] + (long)hi * n + lo));
      double b = normal_at(seed, (unsigned long long)(n2 + B.ch_off[ch] + (long)lo * n + hi));
      double vc = (lo == hi) ? (herm == 1 ? 2.0 * a : 0.0) : (a + herm * b);
      vc = vc * inv_sqrt2 / sqrt((double)n);
      op[B.two_off + t] = (r <= c) ? vc : herm * vc;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) op[0] = 0.0;
}

hipError_t launch_fill_random(const BasisDev *bd, const BasisDev &bh, double *op, int herm, unsigned long long seed,
                              hipStream_t s) {
  long total = (long)bh.norb * bh.norb + bh.tbsize;
  hipLaunchKernelGGL(k_fill_random, dim3(grid_for(total)), dim3(256), 0, s, bd, op, herm, seed);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// White generator, Epstein-Nesbet denominators (src/Generator.cc:110-188, 236-291).
__device__ double tbme_norm_u(const BasisDev &B, const double *two, int J, int a, int b, int c, int d) {
  int pab = (B.l[a] + B.l[b]) & 1, pcd = (B.l[c] + B.l[d]) & 1;
  if (pab != pcd) return 0.0;
  int tab = (B.tz2[a] + B.tz2[b]) / 2, tcd = (B.tz2[c] + B.tz2[d]) / 2;
  if (tab != tcd || J > B.jmax) return 0.0;
  int k1 = ket_index(B.norb, min(a, b), max(a, b)), k2 = ket_index(B.norb, min(c, d), max(c, d));
  int i = B.kmap[(long)k1 * (B.jmax + 1) + J], k = B.kmap[(long)k2 * (B.jmax + 1) + J];
  if (i < 0 || k < 0) return 0.0;
  int ch = 6 * J + 2 * (tab + 1) + pab;
  double ph = 1.0;
  if (a > b) ph *= sgn_parity((B.j2[a] + B.j2[b]) / 2 + 1 + J);
  if (c > d) ph *= sgn_parity((B.j2[c] + B.j2[d]) / 2 + 1 + J);
  return ph * two[B.ch_off[ch] + (long)k * B.ch_n[ch] + i];
}
__device__ double monopole_u(const BasisDev &B, const double *two, int a, int b, int c, int d) {
  double nrm = (a == b ? 1.4142135623730951 : 1.0) * (c == d ? 1.4142135623730951 : 1.0);
  int jmin = abs(B.j2[a] - B.j2[b]) / 2, jmax = (B.j2[a] + B.j2[b]) / 2;
  double mon = 0.0;
  for (int J = jmin; J <= jmax; ++J) mon += (2 * J + 1) * tbme_norm_u(B, two, J, a, b, c, d);
  mon /= (B.j2[a] + 1) * (B.j2[b] + 1);
  return nrm * mon;
}

// one-body part: one thread per (a core, i non-core)
__global__ __launch_bounds__(256) void k_white1(const BasisDev *bd, const double *H, double *Eta) {
  const BasisDev &B = *bd;
  const int norb = B.norb;
  const double *H1 = H + B.one_off, *H2 = H + B.two_off;
  double *E1 = Eta + B.one_off;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)norb * norb;
       e += (long)gridDim.x * blockDim.x) {
    const int i = e % norb, a = e / norb;
    if (B.cvq[a] != 0 || B.cvq[i] == 0) continue;
    double d = H1[(long)i * norb + i] - H1[(long)a * norb + a];
    d += (B.occ[i] - B.occ[a]) * monopole_u(B, H2, i, a, i, a);
    if (fabs(d) < 1e-6) d = 1e-6;
    const double v = H1[(long)a * norb + i] / d;
    E1[(long)a * norb + i] = v;
    E1[(long)i * norb + a] = -v;
  }
}

// two-body part: block per channel, threads over (core ket, non-core ket)
__global__ __launch_bounds__(256) void k_white2(const BasisDev *bd, const double *H, double *Eta,
                                                const int4 *items) {
  const BasisDev &B = *bd;
  const int ch = items[blockIdx.x].y;
  const int n = B.ch_n[ch];
  const long off = B.ch_off[ch];
  const int norb = B.norb;
  const double *H1 = H + B.one_off, *H2 = H + B.two_off;
  double *E2 = Eta + B.two_off + off;
  const int c0 = B.gcore_start[ch], nc = B.gcore_start[ch + 1] - c0;
  const int q0 = B.gnc_start[ch], nq = B.gnc_start[ch + 1] - q0;
  const int ks = B.ch_kstart[ch];
  for (int t = threadIdx.x; t < nc * nq; t += blockDim.x) {
    const int iket = B.gcore_loc[c0 + t / nq], ibra = B.gnc_loc[q0 + t % nq];
    const int kb = B.ch_kets[ks + ibra], kk = B.ch_kets[ks + iket];
    const int i = B.kp[kb], j = B.kq[kb], k = B.kp[kk], l = B.kq[kk];
    const double ni = B.occ[i], nj = B.occ[j], nk = B.occ[k], nl = B.occ[l];
    double d = H1[(long)i * norb + i] + H1[(long)j * norb + j] - H1[(long)k * norb + k] - H1[(long)l * norb + l];
    d += (1 - ni - nj) * monopole_u(B, H2, i, j, i, j);
    d -= (1 - nk - nl) * monopole_u(B, H2, k, l, k, l);
    d += (ni - nk) * monopole_u(B, H2, i, k, i, k);
    d += (ni - nl) * monopole_u(B, H2, i, l, i, l);
    d += (nj - nk) * monopole_u(B, H2, j, k, j, k);
    d += (nj - nl) * monopole_u(B, H2, j, l, j, l);
    if (fabs(d) < 1e-6) d = 1e-6;
    const double v = H2[off + (long)iket * n + ibra] / d;
    E2[(long)iket * n + ibra] = v;
    E2[(long)ibra * n + iket] = -v;
  }
}

hipError_t launch_white(const BasisDev *bd, const BasisDev &bh, const double *H, double *Eta, const int4 *gen_items,
                        int ngen, hipStream_t s) {
  long n2 = (long)bh.norb * bh.norb;
  hipLaunchKernelGGL(k_white1, dim3(grid_for(n2)), dim3(256), 0, s, bd, H, Eta);
  if (ngen > 0) hipLaunchKernelGGL(k_white2, dim3(ngen), dim3(256), 0, s, bd, H, Eta, gen_items);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
typedef double dbl4u __attribute__((ext_vector_type(4)));
// One wave: out[0..255] = D with A-operand = lane id, B-operand = 1;
//           out[256..511] = D with A-operand = 1, B-operand = lane id.
__global__ void k_mfma_probe(double *out) {
  const int lane = threadIdx.x;
  dbl4u z = {0.0, 0.0, 0.0, 0.0};
  dbl4u r1 = __builtin_amdgcn_mfma_f64_16x16x4f64((double)lane, 1.0, z, 0, 0, 0);
  dbl4u r2 = __builtin_amdgcn_mfma_f64_16x16x4f64(1.0, (double)lane, z, 0, 0, 0);
  typedef float flt4u __attribute__((ext_vector_type(4)));
  flt4u zf = {0.f, 0.f, 0.f, 0.f};
  flt4u f1 = __builtin_amdgcn_mfma_f32_16x16x4f32((float)lane, 1.0f, zf, 0, 0, 0);
  flt4u f2 = __builtin_amdgcn_mfma_f32_16x16x4f32(1.0f, (float)lane, zf, 0, 0, 0);
  for (int q = 0; q < 4; ++q) {
    out[lane * 4 + q] = r1[q];
    out[256 + lane * 4 + q] = r2[q];
    out[512 + lane * 4 + q] = f1[q];
    out[768 + lane * 4 + q] = f2[q];
  }
}
hipError_t launch_mfma_probe(double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, out);
  return hipGetLastError();
}

// One v_mfma_f32_16x16x32_bf16 with lane l's operand slots j = 0..7 taken from
// a[8l + j], b[8l + j] (rounded to bf16) and C = 0; d[4l + q] = register q of
// lane l.  One wave, 512 + 512 inputs, 256 outputs.
__global__ void k_mfma_bf16_probe(const float *a, const float *b, float *d) {
  typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
  typedef float flt4u __attribute__((ext_vector_type(4)));
  const int lane = threadIdx.x;
  bf16x8 va, vb;
  for (int j = 0; j < 8; ++j) {
    va[j] = (__bf16)a[lane * 8 + j];
    vb[j] = (__bf16)b[lane * 8 + j];
  }
  const flt4u z = {0.f, 0.f, 0.f, 0.f};
  const flt4u r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb, z, 0, 0, 0);
  for (int q = 0; q < 4; ++q) d[lane * 4 + q] = r[q];
}
hipError_t launch_mfma_bf16_probe(const float *a, const float *b, float *d, hipStream_t s) {
  hipLaunchKernelGGL(k_mfma_bf16_probe, dim3(1), dim3(64), 0, s, a, b, d);
  return hipGetLastError();
}

}  // namespace imsrg
