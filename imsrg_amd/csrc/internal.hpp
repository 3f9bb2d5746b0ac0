// internal.hpp -- device-side tables, operator addressing and work items shared
// by the host driver (api.cpp, basis.cpp) and the HIP kernels (*.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace imsrg {

// Pointer members of the structs below address global memory.  In the device
// pass they carry address space 1, so that every load and store through them
// (and through pointers derived from them) compiles to global_* instructions:
// flat_* ones would also count against the LDS wait counter and serialize LDS
// reads with outstanding global loads.  Host code fills the same layout.
#if defined(__HIP_DEVICE_COMPILE__)
#define IMSRG_G __attribute__((address_space(1)))
#else
#define IMSRG_G
#endif

constexpr int kTile = 64;         // SYR2K output tile edge
constexpr int kTileB = 128;       // output tile edge of the large pp/hh SYR2K
constexpr int kTileK = 16;        // k-depth of one LDS stage
constexpr int kGWaves = 8;        // waves of a 128x128 GEMM work-group (2 x 4 grid of 64 x kGW sub-tiles)
constexpr int kGW = 32;           // columns of a wave's sub-tile in the 128x128 GEMM
constexpr int kGN = kGW / 16;     // 16-column MFMA blocks per wave sub-tile
constexpr int kGC = kTileK / kGWaves;  // k columns each wave copies into LDS per stage
constexpr int kChunk = 256;       // elements per work item of the gather kernels
constexpr int kJC = 8;            // J values per register chunk of the inverse Pandya gather (even)
constexpr int kC122Blk = 4;       // comm122 partner lists are padded to multiples of this (one load block)
constexpr int kC122Cols = 32;     // kets (columns) a bin of the comm122 W kernel is filled to
constexpr int kC122MaxCols = 150; // largest bin (LDS: 16 * 64 bytes per column)
constexpr int kC122Rows = 64;     // rows of one W-kernel work item
constexpr int kF32Flush = 4;      // MFMA k-steps (of 4) summed in FP32 before the FP64 flush (mixed, 64x64 tiles)

// GEMM kernel families passed to launch_pphh / launch_ph_gemm: 0 FP64,
// 1 FP32 MFMA on FP32-rounded operands, 2 split FP32 products, kGemmB3 bf16
// MFMA on three-way split operands
constexpr int kGemmB3 = 3;
constexpr int kB3K = 32;          // k-depth of one stage of the bf16x3 GEMMs (one v_mfma_f32_16x16x32_bf16)

// Term bits (same values as IMSRG_COMM* in include/imsrg_hip.h)
constexpr unsigned M110 = 1u << 0, M220 = 1u << 1, M111 = 1u << 2, M121 = 1u << 3, M221 = 1u << 4,
                   M122 = 1u << 5, MPPHH = 1u << 6, MPH = 1u << 7;

// All basis tables of one model space, device pointers.  Index conventions:
//  * global ket k of orbits p<=q: k = p(2 norb - 1 - p)/2 + q   (ModelSpace::Index2)
//  * standard channel ch = 6J + 2(Tz+1) + parity, J = 0..jmax  (GetTwoBodyChannelIndex)
//  * cross-coupled channel uses the same formula with Tz = |tz_p - tz_q|/2
//  * kmap[k*(jmax+1)+J]: local index of ket k in the standard channel of
//    angular momentum J with the ket's parity and Tz, or -1
//  * cmap[k*(jmax+1)+J]: same for the cross-coupled channel
struct BasisDev {
  int norb, nkets, nch, jmax, d1, d2;
  int nholes, n1pairs;
  long tbsize;   // doubles in the two-body arena
  long one_off;  // offset of the one-body block inside an operator buffer
  long two_off;  // offset of the two-body arena inside an operator buffer
  long opsize;   // doubles per operator buffer (including 8 doubles of zero slack at the end)
  int maxn, maxcc, maxblk;
  int ccjmax_ph;  // largest J of a CC channel with ph kets (-1: none); Zbar is zero above it
  long n122;           // comm122 partner entries (including the padding to multiples of kC122Blk per ket)
  const IMSRG_G int *j2, *l, *tz2, *cvq;
  const IMSRG_G double *occ;
  const IMSRG_G int *holes;
  const IMSRG_G int *obstart, *oblist;  // one-body block of each orbit
  const IMSRG_G int *kp, *kq;
  const IMSRG_G int *kmap, *cmap;
  const IMSRG_G int *cmapT;     // cmapT[J*nkets + k] = cmap[k*(jmax+1) + J]
  const IMSRG_G int *kmapT;     // kmapT[J*nkets + k] = kmap[k*(jmax+1) + J]
  const IMSRG_G int *ccorb;     // ccorb[(i*(jmax+1) + J)*norb + x] = cmap of the ket (min(i,x), max(i,x)) at J
  const IMSRG_G int *kmorb;     // kmorb: the same layout over kmap (standard channels)
  // kets of each (parity, Tz) group g = 2(Tz+1) + parity (the channels 6J + g),
  // ascending global index: grp_kets[grp_start[g] .. grp_start[g+1])
  const IMSRG_G int *grp_start, *grp_kets;
  // standard channels
  const IMSRG_G long *ch_off;
  const IMSRG_G int *ch_n, *ch_kstart, *ch_kets;
  const IMSRG_G double *w220;  // per channel ket: (1-n_p)(1-n_q) for hh/ph kets, 1 for pp kets
  const IMSRG_G int *hh_start, *hh_loc;
  const IMSRG_G double *hh_occ;
  const IMSRG_G int *k_start, *k_loc;  // pp/hh product list with combined weight (w_pp - w_hh)
  const IMSRG_G double *k_w;
  // bf16x3 operand panels of the mixed pp/hh GEMM (16-byte units): channel ch
  // starts at pk_off[ch] with the panel of X W, then the panel of Y, each
  // [nkp/8 k octets][3 pieces][nrp rows] (nkp = nk rounded up to kB3K, nrp = n
  // rounded up to kTileB); pk_size units in all
  const IMSRG_G long *pk_off;
  long pk_size;
  const IMSRG_G int *kpp_start, *kpp_loc;  // M_pp list (weights w_pp)
  const IMSRG_G double *kpp_w;
  const IMSRG_G int *khh_start, *khh_loc;  // M_hh list (weights n_p n_q)
  const IMSRG_G double *khh_w;
  // comm122 partner lists (CSR): the partners of local ket r of channel ch are
  // the entries c122_start[ch_kstart[ch] + r] .. c122_start[ch_kstart[ch] + r + 1],
  // a multiple of kC122Blk (padding: the first partner again with factor 0)
  const IMSRG_G int *c122_start;
  const IMSRG_G int *c122_ind;     // local ket index ind2 of the partner
  const IMSRG_G int2 *c122_ao;     // (a, o): coefficient X1(a, o) * f
  const IMSRG_G double *c122_f;    // phase / sqrt2 factors, times 2 for kets with p == q
  // comm122 column bins (kets of whole one-body block pairs, closed under
  // partners): bin b holds the channel-local kets c122_bcols[c122_boff[b] ..
  // c122_boff[b + 1]); c122_gpos[e] is the position of entry e's partner in
  // its column's bin
  const IMSRG_G int *c122_bcols;
  const IMSRG_G int *c122_boff;
  const IMSRG_G int *c122_gpos;
  // the bin's partner entries, column by column: global entry indices
  // c122_bent[c122_beoff[b] ..]; column cp's entries end at c122_bcend[c122_boff[b] + cp]
  // within the bin
  const IMSRG_G int *c122_bent;
  const IMSRG_G int *c122_beoff;
  const IMSRG_G int *c122_bcend;
  int c122_maxbin;                 // kets of the largest bin
  int c122_maxent;                 // partner entries of the largest bin
  // cross-coupled channels
  const IMSRG_G int *cc_n, *cc_nph, *cc_kstart, *cc_kets, *cc_phstart, *cc_ph;
  const IMSRG_G signed char *cc_phi;  // (-1)^{(j_p+j_q)} per CC ket
  const IMSRG_G signed char *cc_phph; // cc_phi of the ket at each position of the cc_ph lists
  const IMSRG_G long *cc_xoff, *cc_zoff;
  long xsize, zsize;
  // kets of each (parity, Tz) group g by (j_k, j_l) class c = ((j2k-1)/2) d1 + (j2l-1)/2: the group
  // indices e, ascending, in cls_e[cls_start[g ncls + c] .. cls_start[g ncls + c + 1]), ncls = d1 * d1
  const IMSRG_G int *cls_start, *cls_e;
  // per cls_e entry, the ket and its orbits in one record: (ket, k << 16 | l,
  // j2k << 16 | j2l, l_k << 24 | l_l << 16 | (tz2k + 1) << 8 | (tz2l + 1))
  const IMSRG_G int4 *cls_rec;
  int ncls;
  const IMSRG_G int2 *pairs1;   // one-body pairs (i<=j, same (l,j2,tz2) block)
  // partial sums of the one-body part: slots ob_start[p] .. ob_start[p+1]-1 of pair p
  // (the first: comm111 + comm121, the others: comm221 items); cpart: orbits with 1 - n_c != 0
  const IMSRG_G int *ob_start, *cpart;
  int nobpart, ncpart;
  int ndiag;  // kets summed over all channels (ch_kstart[nch])
  const IMSRG_G double *w1norm; // norb*norb one-body norm weights
  const IMSRG_G double *sixj;   // dense 6j table, layout of SixJCache_112112
  const IMSRG_G double *sixjQ;  // (2 J6 + 1) times the values (0 where |6j| < 1e-8), layout
                                // [jj1][jj2][J6][J3/2][jj4][jj5][J3%2] (sixjQ_index):
                                // lanes of one (jj1, jj2, J6, J3 pair) that differ in jj4, jj5 read
                                // neighbouring 16-byte entries; d2q = d2 rounded up to even
  int d2q;
  // generator lists: per channel, core-core kets and (non-core)-(non-core) kets
  const IMSRG_G int *gcore_start, *gcore_loc, *gnc_start, *gnc_loc;
};

// Ket groups of the block-sparse basis transformation: the kets of a channel whose
// orbits come from the same unordered pair of one-body (l, j2, tz2) blocks.
struct XformDev {
  const IMSRG_G int *g_start;   // ngroups + 1: group g holds g_kets[g_start[g] .. g_start[g+1])
  const IMSRG_G int *g_kets;    // channel-local ket indices, ascending within a group
  const IMSRG_G int *g_ch;      // channel of each group
  const IMSRG_G long *g_doff;   // offset of each group's m x m block of D (column-major)
  const IMSRG_G int *k_start;   // per channel: where its kets start in k_grp / k_pos
  const IMSRG_G int *k_grp;     // group of each channel ket
  const IMSRG_G int *k_pos;     // its position inside the group
  int ngroups, maxm;
};

// Everything one commutator needs: operand buffers, workspaces, flags.
struct PairPtrs {
  const IMSRG_G BasisDev *basis;  // device pointer
  const IMSRG_G double *X, *Y;    // operator buffers (zero | one | two)
  IMSRG_G double *Z;
  IMSRG_G double *Xt, *Yt, *Zbar; // Pandya workspaces
  IMSRG_G double *scratch;        // >= nch + 8 doubles (zero-body partial sums)
  IMSRG_G double2 *c122;          // per-call comm122 coefficients (X1(a,o) f, Y1(a,o) f), n122 entries
  IMSRG_G double *obpart;         // per-call one-body partial sums, nobpart entries
  IMSRG_G double *diag;           // per-call inverse-Pandya diagonal terms, one per ket of every channel (ndiag)
  IMSRG_G double *panel;          // per-call bf16x3 panels of the mixed pp/hh GEMM (2 pk_size doubles)
  IMSRG_G double *c122w;          // per-call comm122 W, two-body layout (tbsize + 8 doubles)
  int hX, hY, hZ;
  unsigned mask;
};

struct PairArg {
  PairPtrs p0;           // used when tab == nullptr
  const IMSRG_G PairPtrs *tab;  // batched: device table
  __device__ __forceinline__ PairPtrs get(int i) const {
    if (tab) return tab[i];  // separate returns keep each load in its own address space
    return p0;
  }
};

__host__ __device__ inline int ket_index(int norb, int p, int q) { return p * (2 * norb - 1 - p) / 2 + q; }
__host__ __device__ inline int sgn_parity(int x) { return (x & 1) ? -1 : 1; }
__host__ __device__ inline long sixjQ_index(int d1, int d2q, int jj1, int jj2, int J3, int jj4, int jj5, int J6) {
  return ((((((long)((jj1 - 1) >> 1) * d1 + ((jj2 - 1) >> 1)) * d2q + J6) * (d2q >> 1) + (J3 >> 1)) * d1 +
           ((jj4 - 1) >> 1)) * d1 + ((jj5 - 1) >> 1)) * 2 + (J3 & 1);
}
__host__ __device__ inline long sixj_index(int d1, int d2, int jj1, int jj2, int J3, int jj4, int jj5, int J6) {
  return (((((long)((jj1 - 1) >> 1) * d1 + ((jj2 - 1) >> 1)) * d2 + J3) * d1 + ((jj4 - 1) >> 1)) * d1 +
          ((jj5 - 1) >> 1)) * d2 + J6;
}

}  // namespace imsrg
