// basis.hpp -- host-side model space (orbits, kets, channels) and its device
// tables.  Restates the construction of ModelSpace::Init / SetupKets
// (src/ModelSpace.cc:353-417, 1120-1201) and TwoBodyChannel::Initialize /
// CheckChannel_ket (src/TwoBodyChannel.cc:19-188).
#pragma once
#include <string>
#include <vector>

#include "internal.hpp"

namespace imsrg {

struct HostChannel {
  int J = 0, par = 0, tz = 0;
  std::vector<int> kets;  // global ket indices, ascending
};

struct HostBasis {
  int emax = 0, norb = 0, nkets = 0, nch = 0, jmax = 0;
  std::vector<int> n, l, j2, tz2, cvq;
  std::vector<double> occ;
  std::vector<int> kp, kq;
  std::vector<HostChannel> ch, cc;
  std::vector<long> ch_off;
  long tbsize = 0;
  std::vector<int> kmap, cmap;
  std::vector<int> holes;
  std::vector<std::vector<int>> ob;  // one-body block of each orbit

  // returns empty string on success, else an error message
  std::string build(int norb, const int *n, const int *l, const int *j2, const int *tz2, const double *occ,
                    const int *cvq, int emax);
  static void ho_orbits(int emax, double A, double Z, std::vector<int> &n, std::vector<int> &l, std::vector<int> &j2,
                        std::vector<int> &tz2, std::vector<double> &occ, std::vector<int> &cvq);
};

constexpr int kXcd = 8;    // XCDs of an MI355X; workgroup i runs on XCD i % 8
constexpr int kSuper = 8;  // edge (in tiles) of the pp/hh super-tiles of tile order 2

struct TileGroup {
  std::vector<int4> tiles;  // tiles sharing one 128-row block of one channel
  long cost = 0;
};
// Launch order of tile groups balanced over the XCDs (padding items have .y = -1).
std::vector<int4> xcd_order(std::vector<TileGroup> groups, int nxcd);

// Launch choices that the autotuner sets (defaults: measured best at emax=10).
struct Tuning {
  int pphh_small = 64;  // channels with n <= pphh_small take 64x64 pp/hh tiles, the others 128x128
  int pphh_order = 0;  // 0: tiles by decreasing k-length, 1: row-block groups balanced over the XCDs,
                       // 2: square super-tiles balanced over the XCDs
  int mixed_pphh = 2;  // mixed precision pp/hh kernel: 0 FP64, 1 FP32 MFMA, 2 bf16x3 MFMA
  int mixed_ph = 1;    // mixed precision ph kernel: 0 FP64, 1 FP32 MFMA, 2 bf16x3 MFMA
  int ph_big = 1;      // ph GEMM on 128x128 tiles (1) or 64x64 tiles (0)
  int finish_order = 1;  // tile pairs of the finishing kernel: 0 channel/row order, 1 XCD-contiguous runs
  int overlap_pphh = 1;  // pp/hh GEMM on a second stream: 0 no, 1 forked first, 2 forked after the ph GEMM
};

// Work-item lists (int4: pair, a, b, c) for every kernel of one commutator.
struct WorkLists {
  std::vector<int4> pphh;    // (pair, ch, rb, sb), rb <= sb, 128x128 tiles
  std::vector<int4> pphh64;  // (pair, ch, rb, sb), rb <= sb, 64x64 tiles (channels n <= pphh_small)
  std::vector<int4> pack;    // (pair, ch, k octet, 256-row block): bf16x3 panels of the 128x128-tile channels
  std::vector<int4> ph;      // (pair, cc, rb, sb), rb <= sb, tiles of 128 (ph_big) or 64
  std::vector<int4> pandya;  // (pair, cc, start, 0)
  std::vector<int4> invp;    // (pair, group, bra, first ket) of the (parity, Tz) ket groups
  std::vector<int4> one;     // (pair, i, j, index of the one-body pair)
  std::vector<int4> ob221;   // (pair, one-body pair, kind/c/J, partial-sum slot) of the comm221 one-body part
  std::vector<int4> mirror;  // (pair, ch, rb, sb), rb < sb  and diagonal tiles; ch = -1: empty item
  std::vector<int4> gen;     // (pair, ch, 0, 0) channels with work for the generator
  std::vector<int4> c122w;   // (pair, ch, comm122 column bin, block of kC122Rows rows) of the W kernel
};

// Device copy of the tables; owns the allocations.
struct DeviceBasis {
  BasisDev h{};             // host copy holding device pointers
  BasisDev *d = nullptr;    // device copy of h
  std::vector<void *> allocs;
  std::vector<void *> list_allocs;  // device work lists (rebuilt when the tuning changes)
  WorkLists wl;             // host lists with pair = 0
  Tuning tune;
  int4 *d_pphh = nullptr, *d_pphh64 = nullptr, *d_ph = nullptr, *d_pandya = nullptr, *d_pack = nullptr,
       *d_invp = nullptr, *d_one = nullptr, *d_mirror = nullptr, *d_gen = nullptr, *d_ob221 = nullptr,
       *d_c122w = nullptr;
  // host tables the work lists are built from
  std::vector<int> ch_n, k_start, cc_n, cc_nph, grp_start, grp_kets, gen_ok;
  std::vector<int> ch_kstart, c122_start;
  std::vector<int2> pairs1;
  std::vector<int4> c122w_items;  // the W kernel's items (basis only, not tuned)
  std::vector<int4> ob221;
  std::string upload(const HostBasis &hb, int emax6j);  // builds tables, lists, 6j table
  std::string build_lists(const Tuning &t);             // (re)builds and uploads the work lists
  void release();
};

}  // namespace imsrg
