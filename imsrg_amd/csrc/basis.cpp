// basis.cpp -- model-space construction and device tables.
#include "basis.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>

#include "launch.hpp"

namespace imsrg {

static const double kOccCut = 1e-6;  // ModelSpace::OCC_CUT, src/ModelSpace.cc:14

void HostBasis::ho_orbits(int emax, double A, double Z, std::vector<int> &n, std::vector<int> &l,
                          std::vector<int> &j2, std::vector<int> &tz2, std::vector<double> &occ,
                          std::vector<int> &cvq) {
  // orbit order of src/ModelSpace.cc:376-399
  n.clear(); l.clear(); j2.clear(); tz2.clear(); occ.clear(); cvq.clear();
  for (int N = 0; N <= emax; ++N)
    for (int ll = N; ll >= 0; ll -= 2) {
      int nn = (N - ll) / 2;
      for (int jj = 2 * ll + 1; jj >= 2 * ll - 1 && jj > 0; jj -= 2)
        for (int t : {-1, 1}) {
          n.push_back(nn); l.push_back(ll); j2.push_back(jj); tz2.push_back(t);
          occ.push_back(0.0); cvq.push_back(2);
        }
    }
  // filling of GetOrbitsAZ, src/ModelSpace.cc:611-645; core = reference
  double zz = 0, nn = 0;
  bool done = false;
  for (int N = 0; N <= emax && !done; ++N)
    for (int g = 2 * N + 1; g >= -2 * N && !done; g -= 4) {
      int jj = std::abs(g);
      int ll = g < 0 ? (jj + 1) / 2 : (jj - 1) / 2;
      int nr = (N - ll) / 2;
      for (size_t i = 0; i < n.size(); ++i)
        if (n[i] == nr && l[i] == ll && j2[i] == jj && tz2[i] == -1 && zz < Z) {
          double dz = std::min(Z - zz, jj + 1.0);
          occ[i] = dz / (jj + 1.0);
          cvq[i] = 0;
          zz += dz;
        }
      for (size_t i = 0; i < n.size(); ++i)
        if (n[i] == nr && l[i] == ll && j2[i] == jj && tz2[i] == 1 && nn < A - Z) {
          double dn = std::min(A - Z - nn, jj + 1.0);
          occ[i] = dn / (jj + 1.0);
          cvq[i] = 0;
          nn += dn;
        }
      if (zz == Z && nn == A - Z) done = true;
    }
}

std::string HostBasis::build(int norb_, const int *n_, const int *l_, const int *j2_, const int *tz2_,
                             const double *occ_, const int *cvq_, int emax_) {
  if (norb_ <= 0) return "norb must be positive";
  emax = emax_;
  norb = norb_;
  n.assign(n_, n_ + norb); l.assign(l_, l_ + norb); j2.assign(j2_, j2_ + norb); tz2.assign(tz2_, tz2_ + norb);
  occ.assign(occ_, occ_ + norb); cvq.assign(cvq_, cvq_ + norb);
  jmax = 0;
  for (int i = 0; i < norb; ++i) {
    if (j2[i] <= 0 || (j2[i] & 1) == 0) return "j2 must be odd and positive";
    if (tz2[i] != 1 && tz2[i] != -1) return "tz2 must be +-1";
    if (l[i] < 0 || std::abs(2 * l[i] - j2[i]) != 1) return "inconsistent l and j2";
    if (j2[i] > 2 * emax + 1) return "j2 exceeds 2*emax+1 (6j table range)";
    jmax = std::max(jmax, j2[i]);
  }
  nch = 6 * (jmax + 1);
  nkets = norb * (norb + 1) / 2;
  kp.assign(nkets, 0); kq.assign(nkets, 0);
  for (int p = 0; p < norb; ++p)
    for (int q = p; q < norb; ++q) {
      int k = ket_index(norb, p, q);
      kp[k] = p; kq[k] = q;
    }
  ob.assign(norb, {});
  holes.clear();
  for (int i = 0; i < norb; ++i) {
    for (int j = 0; j < norb; ++j)
      if (l[j] == l[i] && j2[j] == j2[i] && tz2[j] == tz2[i]) ob[i].push_back(j);
    if (occ[i] > kOccCut) holes.push_back(i);
  }
  ch.assign(nch, {});
  cc.assign(nch, {});
  kmap.assign((size_t)nkets * (jmax + 1), -1);
  cmap.assign((size_t)nkets * (jmax + 1), -1);
  ch_off.assign(nch, 0);
  long off = 0;
  for (int c = 0; c < nch; ++c) {
    int J = c / 6, tz = (c % 6) / 2 - 1, par = c % 2;
    ch[c].J = cc[c].J = J; ch[c].tz = cc[c].tz = tz; ch[c].par = cc[c].par = par;
    for (int k = 0; k < nkets; ++k) {
      int p = kp[k], q = kq[k];
      bool tri = (j2[p] + j2[q] >= 2 * J) && (std::abs(j2[p] - j2[q]) <= 2 * J);
      bool parok = ((l[p] + l[q]) % 2) == par;
      // TwoBodyChannel::CheckChannel_ket, src/TwoBodyChannel.cc:142-151
      if (tri && parok && !(p == q && (J % 2) != 0) && tz2[p] + tz2[q] == 2 * tz) {
        kmap[(size_t)k * (jmax + 1) + J] = (int)ch[c].kets.size();
        ch[c].kets.push_back(k);
      }
      // TwoBodyChannel_CC::CheckChannel_ket, src/TwoBodyChannel.cc:180-188
      if (tri && parok && std::abs(tz2[p] - tz2[q]) == 2 * tz) {
        cmap[(size_t)k * (jmax + 1) + J] = (int)cc[c].kets.size();
        cc[c].kets.push_back(k);
      }
    }
    ch_off[c] = off;
    off += (long)ch[c].kets.size() * ch[c].kets.size();
  }
  tbsize = off;
  return "";
}

// ---------------------------------------------------------------------------
template <typename T>
static T *dupload(DeviceBasis &db, const std::vector<T> &v, std::string &err) {
  size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    err = "hipMalloc failed for basis table";
    return nullptr;
  }
  db.allocs.push_back(p);
  if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    err = "hipMemcpy failed for basis table";
  return (T *)p;
}

std::vector<int4> xcd_order(std::vector<TileGroup> groups, int nxcd) {
  // longest-processing-time assignment of tile groups to XCDs, then an
  // interleaved order in which launch position p runs on XCD p % nxcd
  std::stable_sort(groups.begin(), groups.end(),
                   [](const TileGroup &a, const TileGroup &b) { return a.cost > b.cost; });
  std::vector<long> load(nxcd, 0);
  std::vector<std::vector<int4>> per(nxcd);
  for (const auto &g : groups) {
    int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[x] += g.cost;
    per[x].insert(per[x].end(), g.tiles.begin(), g.tiles.end());
  }
  size_t len = 0;
  for (auto &v : per) len = std::max(len, v.size());
  std::vector<int4> out;
  out.reserve(len * nxcd);
  for (size_t t = 0; t < len; ++t)
    for (int x = 0; x < nxcd; ++x) out.push_back(t < per[x].size() ? per[x][t] : make_int4(0, -1, 0, 0));
  return out;
}

std::string DeviceBasis::upload(const HostBasis &hb, int emax6j) {
  std::string err;
  BasisDev &B = h;
  std::memset(&B, 0, sizeof(B));
  B.norb = hb.norb; B.nkets = hb.nkets; B.nch = hb.nch; B.jmax = hb.jmax;
  int jjmax = 2 * emax6j + 1;
  B.d1 = (jjmax - 1) / 2 + 1;
  B.d2 = jjmax + 1;
  B.nholes = (int)hb.holes.size();
  B.tbsize = hb.tbsize;
  B.one_off = 8;
  long n2 = (long)hb.norb * hb.norb;
  B.two_off = 8 + ((n2 + 7) / 8) * 8;
  // 8 doubles of slack after the two-body arena: the pp/hh tile loader copies
  // row pairs and may read one element past the last column of a channel
  B.opsize = B.two_off + ((hb.tbsize + 7) / 8) * 8 + 8;
  const int norb = hb.norb, nch = hb.nch;

  B.j2 = dupload(*this, hb.j2, err); B.l = dupload(*this, hb.l, err); B.tz2 = dupload(*this, hb.tz2, err);
  B.cvq = dupload(*this, hb.cvq, err); B.occ = dupload(*this, hb.occ, err);
  B.holes = dupload(*this, hb.holes, err);
  std::vector<int> obstart(norb + 1, 0), oblist;
  for (int i = 0; i < norb; ++i) {
    obstart[i] = (int)oblist.size();
    oblist.insert(oblist.end(), hb.ob[i].begin(), hb.ob[i].end());
  }
  obstart[norb] = (int)oblist.size();
  B.obstart = dupload(*this, obstart, err); B.oblist = dupload(*this, oblist, err);
  B.kp = dupload(*this, hb.kp, err); B.kq = dupload(*this, hb.kq, err);
  B.kmap = dupload(*this, hb.kmap, err); B.cmap = dupload(*this, hb.cmap, err);
  {
    std::vector<int> cT((size_t)hb.nkets * (hb.jmax + 1));
    for (int k = 0; k < hb.nkets; ++k)
      for (int J = 0; J <= hb.jmax; ++J) cT[(size_t)J * hb.nkets + k] = hb.cmap[(size_t)k * (hb.jmax + 1) + J];
    B.cmapT = dupload(*this, cT, err);
    for (int k = 0; k < hb.nkets; ++k)
      for (int J = 0; J <= hb.jmax; ++J) cT[(size_t)J * hb.nkets + k] = hb.kmap[(size_t)k * (hb.jmax + 1) + J];
    B.kmapT = dupload(*this, cT, err);
    // ccorb[(i (jmax+1) + J) norb + x]: cross-coupled local index of the ket of
    // orbits i and x in channel J (of its parity and |tz_i - tz_x|/2), or -1
    std::vector<int> co((size_t)norb * (hb.jmax + 1) * norb);
    for (int i = 0; i < norb; ++i)
      for (int J = 0; J <= hb.jmax; ++J)
        for (int x = 0; x < norb; ++x) {
          const int k = ket_index(norb, std::min(i, x), std::max(i, x));
          co[((size_t)i * (hb.jmax + 1) + J) * norb + x] = hb.cmap[(size_t)k * (hb.jmax + 1) + J];
        }
    B.ccorb = dupload(*this, co, err);
    // kmorb: the same for the standard channels (kmap)
    for (int i = 0; i < norb; ++i)
      for (int J = 0; J <= hb.jmax; ++J)
        for (int x = 0; x < norb; ++x) {
          const int k = ket_index(norb, std::min(i, x), std::max(i, x));
          co[((size_t)i * (hb.jmax + 1) + J) * norb + x] = hb.kmap[(size_t)k * (hb.jmax + 1) + J];
        }
    B.kmorb = dupload(*this, co, err);
  }
  // (parity, Tz) groups of kets
  std::vector<int> grp_start(7, 0), grp_kets;
  for (int g = 0; g < 6; ++g) {
    grp_start[g] = (int)grp_kets.size();
    for (int k = 0; k < hb.nkets; ++k) {
      const int p = hb.kp[k], q = hb.kq[k];
      const int par = (hb.l[p] + hb.l[q]) & 1, tz = (hb.tz2[p] + hb.tz2[q]) / 2;
      if (2 * (tz + 1) + par == g) grp_kets.push_back(k);
    }
  }
  grp_start[6] = (int)grp_kets.size();
  B.grp_start = dupload(*this, grp_start, err); B.grp_kets = dupload(*this, grp_kets, err);

  // ---- standard channels ----
  std::vector<int> ch_n(nch), ch_kstart(nch + 1), ch_kets;
  std::vector<double> w220;
  std::vector<int> hh_start(nch + 1), hh_loc, k_start(nch + 1), k_loc, kpp_start(nch + 1), kpp_loc,
      khh_start(nch + 1), khh_loc, gcore_start(nch + 1), gcore_loc, gnc_start(nch + 1), gnc_loc;
  std::vector<double> hh_occ, k_w, kpp_w, khh_w;
  // comm122 lists per ket (pushed in ket order), concatenated after the channel loop
  std::vector<std::vector<int4>> p122_e(nch);   // (ind2, a, o, ket)
  std::vector<std::vector<double>> p122_f(nch);
  int maxn = 0;
  const double sqrt2 = std::sqrt(2.0);
  for (int c = 0; c < nch; ++c) {
    const HostChannel &C = hb.ch[c];
    int nk = (int)C.kets.size();
    maxn = std::max(maxn, nk);
    ch_n[c] = nk;
    ch_kstart[c] = (int)ch_kets.size();
    hh_start[c] = (int)hh_loc.size(); k_start[c] = (int)k_loc.size();
    kpp_start[c] = (int)kpp_loc.size(); khh_start[c] = (int)khh_loc.size();
    gcore_start[c] = (int)gcore_loc.size(); gnc_start[c] = (int)gnc_loc.size();
    for (int i = 0; i < nk; ++i) {
      int k = C.kets[i];
      int p = hb.kp[k], q = hb.kq[k];
      ch_kets.push_back(k);
      double op = hb.occ[p], oq = hb.occ[q];
      bool hp = op > kOccCut, hq = oq > kOccCut;
      bool is_pp = (op < kOccCut && oq < kOccCut), is_hh = hp && hq, is_ph = (hp != hq);
      // comm220ss weights, src/Commutator.cc:615-627
      w220.push_back(is_pp ? 1.0 : ((is_hh || is_ph) ? (1 - op) * (1 - oq) : 0.0));
      if (is_hh) { hh_loc.push_back(i); hh_occ.push_back(op * oq); }
      // ConstructScalarMpp_Mhh weights, src/Commutator.cc:1171-1183
      double wpp = is_pp ? 1.0 : ((is_hh || is_ph) ? (1 - op) * (1 - oq) : 0.0);
      double whh = is_hh ? op * oq : 0.0;
      if (wpp != 0.0) { kpp_loc.push_back(i); kpp_w.push_back(wpp); }
      if (whh != 0.0) { khh_loc.push_back(i); khh_w.push_back(whh); }
      if (wpp - whh != 0.0) { k_loc.push_back(i); k_w.push_back(wpp - whh); }
      // generator lists (src/Generator.cc:263-271): cc kets and (qq, vv, qv) kets
      if (hb.cvq[p] == 0 && hb.cvq[q] == 0) gcore_loc.push_back(i);
      if (hb.cvq[p] != 0 && hb.cvq[q] != 0) gnc_loc.push_back(i);
      // comm122ss partner list, src/Commutator.cc:857-896
      int ii = p, jj = q;
      int flip = -sgn_parity((hb.j2[ii] + hb.j2[jj]) / 2 - C.J);
      for (int a : hb.ob[ii]) {
        int lo = std::min(a, jj), hi = std::max(a, jj);
        int loc = hb.kmap[(size_t)ket_index(norb, lo, hi) * (hb.jmax + 1) + C.J];
        if (loc < 0) continue;
        double f = (a > jj) ? (double)flip : (a == jj ? sqrt2 : 1.0);
        if (ii == jj) f = f / sqrt2 * 2.0;  // 1/sqrt2 (src/Commutator.cc:898-902) and the column's 2 (925-926)
        p122_e[c].push_back(make_int4(loc, a, ii, i));
        p122_f[c].push_back(f);
      }
      if (ii != jj)
        for (int a : hb.ob[jj]) {
          int lo = std::min(a, ii), hi = std::max(a, ii);
          int loc = hb.kmap[(size_t)ket_index(norb, lo, hi) * (hb.jmax + 1) + C.J];
          if (loc < 0) continue;
          double f = (ii > a) ? (double)flip : (ii == a ? sqrt2 : 1.0);
          p122_e[c].push_back(make_int4(loc, a, jj, i));
          p122_f[c].push_back(f);
        }
    }
  }
  ch_kstart[nch] = (int)ch_kets.size();
  B.ndiag = ch_kstart[nch];
  hh_start[nch] = (int)hh_loc.size(); k_start[nch] = (int)k_loc.size();
  kpp_start[nch] = (int)kpp_loc.size(); khh_start[nch] = (int)khh_loc.size();
  gcore_start[nch] = (int)gcore_loc.size(); gnc_start[nch] = (int)gnc_loc.size();
  // comm122 lists in CSR form over the kets of all channels; each ket's list is
  // padded to a multiple of kC122Blk entries with its first partner and factor 0
  std::vector<int> c122_start(ch_kets.size() + 1, 0), c122_ind;
  std::vector<int2> c122_ao;
  std::vector<double> c122_f;
  for (int c = 0; c < nch; ++c) {
    const size_t ne = p122_e[c].size();
    for (size_t t = 0; t < ne;) {
      const int ket = p122_e[c][t].w;
      size_t u = t;
      for (; u < ne && p122_e[c][u].w == ket; ++u) {
        const int4 e = p122_e[c][u];
        c122_ind.push_back(e.x);
        c122_ao.push_back(make_int2(e.y, e.z));
        c122_f.push_back(p122_f[c][u]);
      }
      for (size_t v = u - t; v % kC122Blk != 0; ++v) {
        c122_ind.push_back(p122_e[c][t].x);
        c122_ao.push_back(make_int2(p122_e[c][t].y, p122_e[c][t].z));
        c122_f.push_back(0.0);
      }
      c122_start[ch_kstart[c] + ket + 1] = (int)c122_ind.size();
      t = u;
    }
  }
  for (size_t i = 0; i + 1 < c122_start.size(); ++i)  // kets without partners
    c122_start[i + 1] = std::max(c122_start[i + 1], c122_start[i]);
  B.n122 = (long)c122_ind.size();
  // comm122 column bins.  The partners of ket (p, q) are kets (a, q) and (p, a)
  // with a in the one-body block of p or q, so the kets of one unordered pair
  // of one-body blocks form a group closed under partners.  Groups are packed
  // into bins of at most kC122Cols kets (first fit, largest first); the W
  // kernel stages a bin's columns of X2 and Y2 in LDS, and c122_gpos gives each
  // partner entry's position within its bin.
  std::vector<int> c122_bcols, c122_boff(1, 0), c122_gpos(c122_ind.size(), 0);
  // bin entries: c122_bent[c122_beoff[b] ..] the global entry indices of bin b,
  // column by column; c122_bcend[c122_boff[b] + cp] the end of column cp's
  // entries within the bin
  std::vector<int> c122_bent, c122_beoff(1, 0), c122_bcend;
  c122w_items.clear();
  int c122_maxbin = 1, c122_maxent = 4;
  {
    std::vector<int> blk(norb);
    for (int o = 0; o < norb; ++o) blk[o] = *std::min_element(hb.ob[o].begin(), hb.ob[o].end());
    for (int c = 0; c < nch; ++c) {
      const int n = ch_n[c];
      if (n == 0) continue;
      std::map<std::pair<int, int>, std::vector<int>> groups;
      for (int i = 0; i < n; ++i) {
        const int k = ch_kets[ch_kstart[c] + i];
        const int b1 = blk[hb.kp[k]], b2 = blk[hb.kq[k]];
        groups[{std::min(b1, b2), std::max(b1, b2)}].push_back(i);
      }
      std::vector<std::vector<int>> gl;
      for (auto &g : groups) gl.push_back(g.second);
      std::stable_sort(gl.begin(), gl.end(),
                       [](const std::vector<int> &a, const std::vector<int> &b) { return a.size() > b.size(); });
      std::vector<std::vector<int>> bins;
      for (auto &g : gl) {
        // a group larger than the bin capacity (emax > 14) takes a bin of its own
        if ((int)g.size() > kC122MaxCols) return "comm122 ket group larger than the W kernel's LDS";
        size_t t = 0;
        while (t < bins.size() && bins[t].size() + g.size() > (size_t)kC122Cols) ++t;
        if (t == bins.size()) bins.emplace_back();
        bins[t].insert(bins[t].end(), g.begin(), g.end());
      }
      std::vector<int> pos(n, -1);
      for (auto &bn : bins) c122_maxbin = std::max(c122_maxbin, (int)bn.size());
      for (auto &bn : bins) {
        for (size_t t = 0; t < bn.size(); ++t) pos[bn[t]] = (int)t;
        const int ent0 = (int)c122_bent.size();
        for (int q : bn) {
          c122_bcols.push_back(q);
          for (int e = c122_start[ch_kstart[c] + q]; e < c122_start[ch_kstart[c] + q + 1]; ++e) {
            c122_gpos[e] = pos[c122_ind[e]];
            c122_bent.push_back(e);
          }
          c122_bcend.push_back((int)c122_bent.size() - ent0);
        }
        c122_beoff.push_back((int)c122_bent.size());
        c122_maxent = std::max(c122_maxent, (int)c122_bent.size() - ent0);
        const int bi = (int)c122_boff.size() - 1;
        c122_boff.push_back((int)c122_bcols.size());
        for (int r = 0; r < n; r += kC122Rows) c122w_items.push_back(make_int4(0, c, bi, r / kC122Rows));
      }
      for (auto &bn : bins)  // every partner of a bin's column lies in the bin
        for (int q : bn)
          for (int e = c122_start[ch_kstart[c] + q]; e < c122_start[ch_kstart[c] + q + 1]; ++e)
            if (pos[c122_ind[e]] < 0 || (int)bn.size() <= pos[c122_ind[e]] || bn[pos[c122_ind[e]]] != c122_ind[e])
              return "comm122 partner outside its ket group";
    }
  }
  B.maxn = maxn;
  if ((long)maxn * maxn >= (1L << 29)) return "two-body channel too large (kernels address a channel block with 32-bit byte offsets)";
  {
    int mb = 0;
    for (int i = 0; i < norb; ++i) mb = std::max(mb, (int)hb.ob[i].size());
    B.maxblk = mb;
  }
  B.ch_off = dupload(*this, hb.ch_off, err);
  B.ch_n = dupload(*this, ch_n, err); B.ch_kstart = dupload(*this, ch_kstart, err);
  B.ch_kets = dupload(*this, ch_kets, err); B.w220 = dupload(*this, w220, err);
  B.hh_start = dupload(*this, hh_start, err); B.hh_loc = dupload(*this, hh_loc, err);
  B.hh_occ = dupload(*this, hh_occ, err);
  B.k_start = dupload(*this, k_start, err); B.k_loc = dupload(*this, k_loc, err); B.k_w = dupload(*this, k_w, err);
  {
    // bf16x3 operand panels of the mixed pp/hh GEMM: per channel, two panels
    // (X W and Y over the pp/hh list) of nkp/8 k octets x 3 pieces x nrp rows,
    // 16 bytes each, nkp = nk rounded up to 32 and nrp = n rounded up to 128
    std::vector<long> pk(nch + 1, 0);
    long o = 0;
    for (int c = 0; c < nch; ++c) {
      pk[c] = o;
      const long nk = k_start[c + 1] - k_start[c], n = ch_n[c];
      if (n > 0 && nk > 0) o += 2 * ((nk + kB3K - 1) / kB3K) * (kB3K / 8) * 3 * ((n + kTileB - 1) / kTileB * kTileB);
    }
    pk[nch] = o;
    B.pk_off = dupload(*this, pk, err);
    B.pk_size = o;
  }
  B.kpp_start = dupload(*this, kpp_start, err); B.kpp_loc = dupload(*this, kpp_loc, err);
  B.kpp_w = dupload(*this, kpp_w, err);
  B.khh_start = dupload(*this, khh_start, err); B.khh_loc = dupload(*this, khh_loc, err);
  B.khh_w = dupload(*this, khh_w, err);
  B.c122_start = dupload(*this, c122_start, err); B.c122_ind = dupload(*this, c122_ind, err);
  B.c122_bcols = dupload(*this, c122_bcols, err); B.c122_boff = dupload(*this, c122_boff, err);
  B.c122_gpos = dupload(*this, c122_gpos, err);
  B.c122_bent = dupload(*this, c122_bent, err);
  B.c122_beoff = dupload(*this, c122_beoff, err);
  B.c122_bcend = dupload(*this, c122_bcend, err);
  B.c122_maxbin = c122_maxbin;
  B.c122_maxent = c122_maxent;
  B.c122_ao = dupload(*this, c122_ao, err); B.c122_f = dupload(*this, c122_f, err);
  B.gcore_start = dupload(*this, gcore_start, err); B.gcore_loc = dupload(*this, gcore_loc, err);
  B.gnc_start = dupload(*this, gnc_start, err); B.gnc_loc = dupload(*this, gnc_loc, err);

  // ---- cross-coupled channels ----
  std::vector<int> cc_n(nch), cc_nph(nch), cc_kstart(nch + 1), cc_kets, cc_phstart(nch + 1), cc_ph;
  std::vector<signed char> cc_phi, cc_phph;
  std::vector<long> cc_xoff(nch + 1), cc_zoff(nch + 1);
  long xo = 0, zo = 0;
  int maxcc = 0;
  for (int c = 0; c < nch; ++c) {
    const HostChannel &C = hb.cc[c];
    int nk = (int)C.kets.size();
    maxcc = std::max(maxcc, nk);
    cc_n[c] = nk;
    cc_kstart[c] = (int)cc_kets.size();
    cc_phstart[c] = (int)cc_ph.size();
    std::vector<int> hh, ph;
    for (int i = 0; i < nk; ++i) {
      int k = C.kets[i];
      int p = hb.kp[k], q = hb.kq[k];
      cc_kets.push_back(k);
      cc_phi.push_back((signed char)sgn_parity((hb.j2[p] + hb.j2[q]) / 2));
      bool hp = hb.occ[p] > kOccCut, hq = hb.occ[q] > kOccCut;
      if (hp && hq) hh.push_back(i);
      if (hp != hq) ph.push_back(i);
    }
    // kets_ph = join_cols(hh, ph), src/Commutator.cc:1398
    cc_ph.insert(cc_ph.end(), hh.begin(), hh.end());
    cc_ph.insert(cc_ph.end(), ph.begin(), ph.end());
    for (int i : hh) cc_phph.push_back(cc_phi[cc_kstart[c] + i]);
    for (int i : ph) cc_phph.push_back(cc_phi[cc_kstart[c] + i]);
    cc_nph[c] = (int)(hh.size() + ph.size());
    cc_xoff[c] = xo;
    cc_zoff[c] = zo;
    xo += (long)nk * 2 * cc_nph[c];
    zo += (long)nk * 2 * nk;
  }
  cc_kstart[nch] = (int)cc_kets.size();
  cc_phstart[nch] = (int)cc_ph.size();
  cc_xoff[nch] = xo;
  cc_zoff[nch] = zo;
  B.maxcc = maxcc;
  B.ccjmax_ph = -1;
  for (int c = 0; c < nch; ++c)
    if (cc_nph[c] > 0) B.ccjmax_ph = std::max(B.ccjmax_ph, c / 6);
  B.xsize = xo;
  B.zsize = zo;
  B.cc_n = dupload(*this, cc_n, err); B.cc_nph = dupload(*this, cc_nph, err);
  B.cc_kstart = dupload(*this, cc_kstart, err); B.cc_kets = dupload(*this, cc_kets, err);
  B.cc_phstart = dupload(*this, cc_phstart, err); B.cc_ph = dupload(*this, cc_ph, err);
  B.cc_phi = dupload(*this, cc_phi, err);
  B.cc_phph = dupload(*this, cc_phph, err);
  B.cc_xoff = dupload(*this, cc_xoff, err); B.cc_zoff = dupload(*this, cc_zoff, err);

  // ---- one-body pairs and norm weights ----
  std::vector<int2> pairs1;
  std::vector<double> w1(n2, 0.0);
  for (int i = 0; i < norb; ++i)
    for (int j : hb.ob[i]) {
      if (j >= i) pairs1.push_back(make_int2(i, j));
      w1[(long)j * norb + i] = hb.j2[i] + 1;  // Operator::OneBodyNorm degeneracy (scalar)
    }
  B.n1pairs = (int)pairs1.size();
  B.pairs1 = dupload(*this, pairs1, err);
  B.w1norm = dupload(*this, w1, err);
  // comm221 one-body work (src/Commutator.cc:1246-1294) in partial sums: pair p
  // owns the slots ob_start[p] .. ob_start[p + 1] - 1 of the per-call buffer;
  // the first takes its comm111 + comm121 part, every item one more slot.
  //  kind 0 (n_c != 0, M_pp over the long pp lists): one item per (c, J);
  //  kind 1 (1 - n_c != 0, M_hh over the short hh lists): one item per 4 entries
  //  of the list cpart of such orbits c (one wave each, all J)
  std::vector<int> ob_start(pairs1.size() + 1, 0), cpart;
  for (int c = 0; c < norb; ++c)
    if (std::fabs(1.0 - hb.occ[c]) > 1e-9) cpart.push_back(c);
  ob221.clear();
  {
    int slot = 0;
    for (size_t p = 0; p < pairs1.size(); ++p) {
      ob_start[p] = slot++;
      const int i = pairs1[p].x, j = pairs1[p].y;
      for (int c = 0; c < norb; ++c) {
        if (!(std::fabs(hb.occ[c]) > 1e-9)) continue;
        const int Jmin = std::max(std::abs(hb.j2[c] - hb.j2[i]), std::abs(hb.j2[c] - hb.j2[j])) / 2;
        const int Jmax = (hb.j2[c] + std::min(hb.j2[i], hb.j2[j])) / 2;
        for (int J = Jmin; J <= Jmax; ++J) ob221.push_back(make_int4(0, (int)p, (c << 8) | J, slot++));
      }
      for (size_t q = 0; q < cpart.size(); q += 4)
        ob221.push_back(make_int4(0, (int)p, (1 << 30) | (int)q, slot++));
    }
    ob_start[pairs1.size()] = slot;
    B.nobpart = slot;
  }
  B.ncpart = (int)cpart.size();
  B.ob_start = dupload(*this, ob_start, err);
  B.cpart = dupload(*this, cpart, err);
  if (!err.empty()) return err;

  // ---- 6j table on the device ----
  {
    long tot = (long)B.d1 * B.d1 * B.d2 * B.d1 * B.d1 * B.d2;
    B.d2q = (B.d2 + 1) & ~1;
    // the gather copy: padded J3 range, plus kJC / 2 (jj4, jj5) blocks of J3 pairs
    // so that a kJC-wide J chunk starting at the last pair stays inside the allocation
    long totq = (long)B.d1 * B.d1 * B.d1 * B.d1 * B.d2q * B.d2q + (long)(kJC / 2) * 2 * B.d1 * B.d1;
    void *p = nullptr, *pQ = nullptr;
    if (hipMalloc(&p, tot * sizeof(double)) != hipSuccess) return "hipMalloc failed for the 6j table";
    allocs.push_back(p);
    if (hipMalloc(&pQ, totq * sizeof(double)) != hipSuccess) return "hipMalloc failed for the 6j table";
    allocs.push_back(pQ);
    if (hipMemset(pQ, 0, totq * sizeof(double)) != hipSuccess) return "hipMemset failed for the 6j table";
    if (launch_sixj_table((double *)p, (double *)pQ, B.d1, B.d2, B.d2q, 0) != hipSuccess)
      return "6j table kernel failed";
    if (hipDeviceSynchronize() != hipSuccess) return "6j table kernel failed";
    B.sixj = (const double *)p;
    B.sixjQ = (const double *)pQ;
  }

  // ---- (j_k, j_l) class lists of the ket groups (inverse Pandya ket order) ----
  {
    const int ncls = B.d1 * B.d1;
    std::vector<int> cls_start(6 * ncls + 1, 0), cls_e;
    for (int g = 0; g < 6; ++g) {
      std::vector<std::vector<int>> per(ncls);
      for (int e = 0; e < grp_start[g + 1] - grp_start[g]; ++e) {
        const int kk = grp_kets[grp_start[g] + e];
        per[((hb.j2[hb.kp[kk]] - 1) / 2) * B.d1 + (hb.j2[hb.kq[kk]] - 1) / 2].push_back(e);
      }
      for (int c = 0; c < ncls; ++c) {
        cls_start[g * ncls + c] = (int)cls_e.size();
        cls_e.insert(cls_e.end(), per[c].begin(), per[c].end());
      }
    }
    cls_start[6 * ncls] = (int)cls_e.size();
    B.ncls = ncls;
    B.cls_start = dupload(*this, cls_start, err);
    B.cls_e = dupload(*this, cls_e, err);
    // per class entry: the ket and its orbits' quantum numbers in one record
    std::vector<int4> rec(cls_e.size());
    for (int g = 0; g < 6; ++g)
      for (int t = cls_start[g * ncls]; t < cls_start[(g + 1) * ncls]; ++t) {
        const int kk = grp_kets[grp_start[g] + cls_e[t]];
        const int k = hb.kp[kk], l = hb.kq[kk];
        rec[t] = make_int4(kk, (k << 16) | l, (hb.j2[k] << 16) | hb.j2[l],
                           (hb.l[k] << 24) | (hb.l[l] << 16) | ((hb.tz2[k] + 1) << 8) | (hb.tz2[l] + 1));
      }
    B.cls_rec = dupload(*this, rec, err);
  }

  // ---- work lists ----
  this->ch_n = ch_n;
  this->ch_kstart = ch_kstart;
  this->c122_start = c122_start;
  this->k_start = k_start;
  this->cc_n = cc_n;
  this->cc_nph = cc_nph;
  this->grp_start = grp_start;
  this->grp_kets = grp_kets;
  this->pairs1 = pairs1;
  gen_ok.assign(nch, 0);
  for (int c = 0; c < nch; ++c) gen_ok[c] = gcore_start[c + 1] > gcore_start[c] && gnc_start[c + 1] > gnc_start[c];
  if (!err.empty()) return err;
  std::string le = build_lists(tune);
  if (!le.empty()) return le;

  void *p = nullptr;
  if (hipMalloc(&p, sizeof(BasisDev)) != hipSuccess) return "hipMalloc failed for BasisDev";
  allocs.push_back(p);
  d = (BasisDev *)p;
  if (hipMemcpy(d, &B, sizeof(BasisDev), hipMemcpyHostToDevice) != hipSuccess) return "hipMemcpy BasisDev failed";
  return "";
}

std::string DeviceBasis::build_lists(const Tuning &t) {
  const int nch = (int)ch_n.size();
  wl = WorkLists();
  {
    // pp/hh tiles: 64x64 for channels with n <= t.pphh_small, 128x128 otherwise
    std::vector<TileGroup> groups;
    for (int c = 0; c < nch; ++c) {
      const int nk = ch_n[c];
      if (nk == 0) continue;
      if (nk <= t.pphh_small) {
        const int nt = (nk + kTile - 1) / kTile;
        for (int rb = 0; rb < nt; ++rb)
          for (int sb = rb; sb < nt; ++sb) wl.pphh64.push_back(make_int4(0, c, rb, sb));
        continue;
      }
      const int nt = (nk + kTileB - 1) / kTileB;
      const long kk = k_start[c + 1] - k_start[c];
      if (t.pphh_order == 2) {
        // square super-tiles of kSuper x kSuper tiles: their work-groups run
        // together on one XCD and share the A and B row blocks in its L2
        for (int rs = 0; rs < nt; rs += kSuper)
          for (int ss = rs; ss < nt; ss += kSuper) {
            TileGroup g;
            for (int rb = rs; rb < std::min(rs + kSuper, nt); ++rb)
              for (int sb = std::max(rb, ss); sb < std::min(ss + kSuper, nt); ++sb)
                g.tiles.push_back(make_int4(0, c, rb, sb));
            if (g.tiles.empty()) continue;
            g.cost = (kk + 1) * (long)g.tiles.size();
            groups.push_back(g);
          }
        continue;
      }
      for (int rb = 0; rb < nt; ++rb) {
        TileGroup g;
        for (int sb = rb; sb < nt; ++sb) g.tiles.push_back(make_int4(0, c, rb, sb));
        g.cost = (kk + 1) * (long)g.tiles.size();
        groups.push_back(g);
      }
    }
    // pack items of the mixed pp/hh GEMM: (pair, ch, k octet, block of 256 rows)
    // for the channels on 128x128 tiles
    for (int c = 0; c < nch; ++c) {
      const int n = ch_n[c], nk = k_start[c + 1] - k_start[c];
      if (n == 0 || n <= t.pphh_small || nk == 0) continue;
      const int noct = (nk + kB3K - 1) / kB3K * (kB3K / 8), nrp = (n + kTileB - 1) / kTileB * kTileB;
      for (int o = 0; o < noct; ++o)
        for (int r = 0; r < nrp; r += 256) wl.pack.push_back(make_int4(0, c, o, r / 256));
    }
    if (t.pphh_order >= 1) {
      wl.pphh = xcd_order(groups, kXcd);
    } else {
      std::vector<std::pair<long, int4>> ts;
      for (auto &g : groups)
        for (auto &w : g.tiles) ts.push_back({g.cost / (long)g.tiles.size(), w});
      std::stable_sort(ts.begin(), ts.end(), [](const std::pair<long, int4> &a, const std::pair<long, int4> &b) {
        return a.first > b.first;
      });
      for (auto &x : ts) wl.pphh.push_back(x.second);
    }
  }
  for (int c = 0; c < nch; ++c) {
    const int nk = ch_n[c];
    if (nk == 0) continue;
    const int nt = (nk + kTile - 1) / kTile;
    for (int rb = 0; rb < nt; ++rb)
      for (int sb = rb; sb < nt; ++sb) wl.mirror.push_back(make_int4(0, c, rb, sb));
    if (gen_ok[c]) wl.gen.push_back(make_int4(0, c, 0, 0));
  }
  wl.c122w = c122w_items;
  if (t.finish_order == 1) {
    // XCD-contiguous order: workgroup i runs on XCD i % 8, so the list is cut
    // into 8 runs of consecutive tiles (whole block rows, mostly) of about equal
    // cost, padded with empty items to equal length, and the runs are
    // interleaved; each XCD then sweeps its own block rows, whose partner-column
    // segments (comm122) stay in that XCD's L2.  Cost of a tile pair: partner
    // entries of its row- and column-block kets plus its columns.
    const std::vector<int4> &L = wl.mirror;
    std::vector<double> acc(L.size() + 1, 0.0);
    for (size_t i = 0; i < L.size(); ++i) {
      const int c = L[i].y, n = ch_n[c], ks = ch_kstart[c];
      double w = 0.0;
      for (int blk : {L[i].z, L[i].w}) {
        const int lo = blk * kTile, hi = std::min(n, lo + kTile);
        w += (c122_start[ks + hi] - c122_start[ks + lo]) + 0.5 * (hi - lo);
      }
      acc[i + 1] = acc[i] + w;
    }
    size_t bound[9];
    bound[0] = 0;
    bound[8] = L.size();
    for (int x = 1; x < 8; ++x) {
      const double target = acc[L.size()] * x / 8;
      size_t j = bound[x - 1];
      while (j < L.size() && acc[j] < target) ++j;
      bound[x] = j;
    }
    size_t len = 0;
    for (int x = 0; x < 8; ++x) len = std::max(len, bound[x + 1] - bound[x]);
    std::vector<int4> v(8 * len, make_int4(0, -1, 0, 0));
    for (int x = 0; x < 8; ++x)
      for (size_t j = bound[x]; j < bound[x + 1]; ++j) v[8 * (j - bound[x]) + x] = L[j];
    wl.mirror.swap(v);
  }
  // inverse Pandya by orbit quadruple: one item (pair, group g, bra b, 0) per
  // bra ket, whose work-group loops over the group kets e >= b; longest first
  {
    std::vector<std::pair<int, int4>> ts;
    for (int g = 0; g < 6; ++g) {
      const int gs = grp_start[g], ng = grp_start[g + 1] - gs;
      for (int b = 0; b < ng; ++b) ts.push_back({ng - b, make_int4(0, g, b, 0)});
    }
    std::stable_sort(ts.begin(), ts.end(),
                     [](const std::pair<int, int4> &x, const std::pair<int, int4> &y) { return x.first > y.first; });
    for (auto &x : ts) wl.invp.push_back(x.second);
  }
  for (int c = 0; c < nch; ++c) {
    const int nk = cc_n[c], nph = cc_nph[c];
    if (nk == 0 || nph == 0) continue;
    const int tile = t.ph_big ? kTileB : kTile;
    const int nt = (nk + tile - 1) / tile;
    for (int rb = 0; rb < nt; ++rb)
      for (int sb = rb; sb < nt; ++sb) wl.ph.push_back(make_int4(0, c, rb, sb));
    const long tot = (long)nk * 2 * nph;
    for (long s = 0; s < tot; s += kChunk) wl.pandya.push_back(make_int4(0, c, (int)s, 0));
  }
  for (size_t p = 0; p < pairs1.size(); ++p) wl.one.push_back(make_int4(0, pairs1[p].x, pairs1[p].y, (int)p));
  wl.ob221 = ob221;

  // upload, replacing the previous lists
  for (void *p : list_allocs) (void)hipFree(p);
  list_allocs.clear();
  std::string err;
  auto up = [&](const std::vector<int4> &v) -> int4 * {
    void *p = nullptr;
    const size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(int4);
    if (hipMalloc(&p, bytes) != hipSuccess) {
      err = "hipMalloc failed for a work list";
      return nullptr;
    }
    list_allocs.push_back(p);
    if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess)
      err = "hipMemcpy failed for a work list";
    return (int4 *)p;
  };
  d_pphh = up(wl.pphh); d_pphh64 = up(wl.pphh64); d_ph = up(wl.ph); d_pack = up(wl.pack);
  d_pandya = up(wl.pandya); d_invp = up(wl.invp);
  d_one = up(wl.one); d_mirror = up(wl.mirror); d_gen = up(wl.gen); d_ob221 = up(wl.ob221);
  d_c122w = up(wl.c122w);
  if (err.empty()) tune = t;
  return err;
}

void DeviceBasis::release() {
  for (void *p : list_allocs) (void)hipFree(p);
  list_allocs.clear();
  for (void *p : allocs) (void)hipFree(p);
  allocs.clear();
  d = nullptr;
}

}  // namespace imsrg
