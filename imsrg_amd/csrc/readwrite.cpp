// readwrite.cpp -- interaction files into the two-body layout of an operator
// (the reference's ReadWrite.cc): the Darmstadt "me2j" two-body format.
#include "readwrite.hpp"

#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <algorithm>
#include <atomic>
#include <map>
#include <sstream>
#include <thread>
#include <tuple>

namespace imsrg {

namespace {

constexpr double kSqrt2 = 1.4142135623730951;
constexpr int kBinHeader = 255;  // bytes before the floats of a .bin file (HEADERSIZE)

// The numbers of an me2j file in file order.  Text (plain or gzip): the first line
// is a header, then whitespace-separated numbers, each rounded to float as the
// reference reads them into float variables; .bin: a 255-byte header, then floats.
struct NumberSource {
  std::string text;
  std::vector<float> bin;
  size_t pos = 0;
  bool binary = false;
  // up to n numbers from the current position into out
  void take_all(std::vector<float> &out, long n) {
    if (binary) {
      const size_t k = std::min(bin.size() - std::min(pos, bin.size()), (size_t)std::max(n, 0L));
      out.assign(bin.begin() + pos, bin.begin() + pos + k);
      pos += k;
      return;
    }
    out.clear();
    out.reserve((size_t)std::max(n, 0L));
    float x;
    while ((long)out.size() < n && next(x)) out.push_back(x);
  }
  bool next(float &x) {
    if (binary) {
      if (pos >= bin.size()) return false;
      x = bin[pos++];
      return true;
    }
    const char *s = text.c_str() + pos;
    char *end = nullptr;
    const float v = std::strtof(s, &end);
    if (end == s) return false;
    x = v;
    pos = (size_t)(end - text.c_str());
    return true;
  }
};

bool ends_with(const std::string &s, const std::string &t) {
  return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

std::string open_source(const std::string &path, NumberSource &src) {
  if (ends_with(path, ".gz")) {
    gzFile f = gzopen(path.c_str(), "rb");
    if (!f) return "cannot open " + path;
    std::string out;
    char buf[1 << 16];
    int n;
    while ((n = gzread(f, buf, sizeof buf)) > 0) out.append(buf, (size_t)n);
    const bool bad = n < 0;
    gzclose(f);
    if (bad) return "gzip error in " + path;
    src.text.swap(out);
  } else if (ends_with(path, ".bin")) {
    std::ifstream in(path, std::ios::binary);
    if (!in.good()) return "cannot open " + path;
    in.seekg(0, std::ios::end);
    const long size = (long)in.tellg();
    if (size < kBinHeader) return "binary file shorter than its header: " + path;
    in.seekg(kBinHeader, std::ios::beg);
    src.bin.resize((size_t)(size - kBinHeader) / sizeof(float));
    in.read(reinterpret_cast<char *>(src.bin.data()), (std::streamsize)(src.bin.size() * sizeof(float)));
    src.binary = true;
    return "";
  } else {
    std::ifstream in(path);
    if (!in.good()) return "cannot open " + path;
    std::stringstream ss;
    ss << in.rdbuf();
    src.text = ss.str();
  }
  // skip the header line
  const size_t nl = src.text.find('\n');
  src.pos = nl == std::string::npos ? src.text.size() : nl + 1;
  return "";
}

struct TwoBodyWriter {
  const HostBasis &hb;
  std::vector<double> &two;
  int ket(int p, int q) const { return p * (2 * hb.norb - 1 - p) / 2 + q; }
  int phase(int p, int q, int J) const { return (((hb.j2[p] + hb.j2[q]) / 2 + 1 + J) % 2 == 0) ? 1 : -1; }
  // local indices and the phase of the kets (a,b), (c,d) in channel (J, parity, Tz)
  bool locate(int J, int par, int tz, int a, int b, int c, int d, long &ebra, long &eket, double &ph) const {
    const int ch = 6 * J + 2 * (tz + 1) + par;
    if (ch < 0 || ch >= hb.nch) return false;
    const int jm = hb.jmax + 1;
    const int ib = hb.kmap[(size_t)ket(std::min(a, b), std::max(a, b)) * jm + J];
    const int ik = hb.kmap[(size_t)ket(std::min(c, d), std::max(c, d)) * jm + J];
    if (ib < 0 || ik < 0) return false;
    const long n = (long)hb.ch[ch].kets.size(), off = hb.ch_off[ch];
    ph = 1.0;
    if (a > b) ph *= phase(b, a, J);
    if (c > d) ph *= phase(d, c, J);
    ebra = off + (long)ik * n + ib;  // element (bra, ket), column-major
    eket = off + (long)ib * n + ik;  // element (ket, bra)
    return true;
  }
  // TwoBodyME::SetTBME for a Hermitian operator (src/TwoBodyME.cc:155-170, 275-279)
  bool set(int J, int par, int tz, int a, int b, int c, int d, double v) {
    long e1, e2;
    double ph;
    if (!locate(J, par, tz, a, b, c, d, e1, e2, ph)) return false;
    two[e1] = ph * v;
    two[e2] = ph * v;
    return true;
  }
  // TwoBodyME::AddToTBME for a Hermitian operator (src/TwoBodyME.cc:172-198, 281-285)
  bool add(int J, int par, int tz, int a, int b, int c, int d, double v) {
    long e1, e2;
    double ph;
    if (!locate(J, par, tz, a, b, c, d, e1, e2, ph)) return false;
    two[e1] += ph * v;
    if (e1 != e2) two[e2] += ph * v;
    return true;
  }
  // TwoBodyME::Set_pn_TBME_from_iso (src/TwoBodyME.cc): the isospin-coupled element
  // spread over the proton-neutron kets; a, b, c, d are proton orbits, +1 the neutron
  bool pn_from_iso(int J, int T, int a, int b, int c, int d, double v) {
    const int par = (hb.l[a] + hb.l[b]) % 2;
    const double eta = 2 * T - 1;
    v *= 0.5;
    if (a == b) v *= kSqrt2;
    if (c == d) v *= kSqrt2;
    bool ok = add(J, par, 0, a, b + 1, c, d + 1, v);
    if (c != d) ok = ok && add(J, par, 0, a, b + 1, c + 1, d, v * eta);
    if (a != b && c != d) ok = ok && add(J, par, 0, a + 1, b, c + 1, d, v);
    if (a != b && (a != c || b != d)) ok = ok && add(J, par, 0, a + 1, b, c, d + 1, v * eta);
    return ok;
  }
};

}  // namespace

// ReadWrite::ReadBareTBME_Darmstadt(_from_stream), src/ReadWrite.cc:664-728, 834-957
std::string read_me2j(const HostBasis &hb, const std::string &path, int emax, int E2max_in, int lmax_in,
                      std::vector<double> &two, long *nrecords) {
  NumberSource src;
  std::string err = open_source(path, src);
  if (!err.empty()) return err;
  if (emax < 0) emax = hb.emax;
  const int lmax = lmax_in > 0 ? lmax_in : emax;
  const int E2max = E2max_in > 0 ? E2max_in : 2 * emax;
  std::map<std::tuple<int, int, int, int>, int> orbit;  // (n, l, j2, tz2) -> index
  for (int i = 0; i < hb.norb; ++i) orbit[std::make_tuple(hb.n[i], hb.l[i], hb.j2[i], hb.tz2[i])] = i;
  std::vector<int> remap, ev, lv, jv;  // the file's nlj orbits, in file order
  for (int e = 0; e <= std::min(emax, hb.emax); ++e)
    for (int l = e % 2; l <= std::min(e, lmax); l += 2) {
      const int n = (e - l) / 2;
      for (int tj = std::abs(2 * l - 1); tj <= 2 * l + 1; tj += 2) {
        auto it = orbit.find(std::make_tuple(n, l, tj, -1));
        if (it == orbit.end() || it->second + 1 >= hb.norb || hb.tz2[it->second + 1] != 1 ||
            hb.n[it->second + 1] != n || hb.l[it->second + 1] != l || hb.j2[it->second + 1] != tj)
          return "the model space lacks the proton-neutron orbit pair n=" + std::to_string(n) + " l=" +
                 std::to_string(l) + " j2=" + std::to_string(tj);
        remap.push_back(it->second);
        ev.push_back(2 * n + l);
        lv.push_back(l);
        jv.push_back(tj);
      }
    }
  const int nlj = (int)remap.size();
  // The records of one nlj1 (the file's outermost orbit) write only elements whose bra
  // holds that orbit (or, mirrored, whose ket does with nlj3 = nlj1), so the nlj1 blocks
  // write disjoint elements: they run in parallel, each in file order, and the result
  // equals the sequential read bit for bit.  First the records of each block.
  auto for_records = [&](int i1, auto &&fn) {
    for (int i2 = 0; i2 <= i1; ++i2) {
      if (ev[i1] + ev[i2] > E2max) break;
      for (int i3 = 0; i3 <= i1; ++i3)
        for (int i4 = 0; i4 <= (i3 == i1 ? i2 : i3); ++i4) {
          if (ev[i3] + ev[i4] > E2max) break;
          if ((lv[i1] + lv[i2] + lv[i3] + lv[i4]) % 2 != 0) continue;
          const int Jmin = std::max(std::abs(jv[i1] - jv[i2]), std::abs(jv[i3] - jv[i4])) / 2;
          const int Jmax = std::min(jv[i1] + jv[i2], jv[i3] + jv[i4]) / 2;
          for (int J = Jmin; J <= Jmax; ++J) fn(i2, i3, i4, J);
        }
    }
  };
  std::vector<long> first(nlj + 1, 0);
  for (int i1 = 0; i1 < nlj; ++i1) {
    long cnt = 0;
    if (ev[i1] <= hb.emax) for_records(i1, [&](int, int, int, int) { ++cnt; });
    first[i1 + 1] = first[i1] + cnt;
  }
  std::vector<float> vals;
  src.take_all(vals, 4 * first[nlj]);
  if ((long)vals.size() < 4 * first[nlj])
    return "file ends after " + std::to_string(vals.size() / 4) + " records of 4 elements: " + path;
  std::atomic<int> next{0};
  std::atomic<bool> bad{false};
  auto work = [&]() {
    TwoBodyWriter w{hb, two};
    for (int i1; (i1 = next.fetch_add(1)) < nlj;) {
      const int a = remap[i1];
      long rec = first[i1];
      for_records(i1, [&](int i2, int i3, int i4, int J) {
        const int b = remap[i2], c = remap[i3], d = remap[i4];
        const int par = (hb.l[a] + hb.l[b]) % 2;
        // (T, Tz) = (0,0), (1,1), (1,0), (1,-1): T=0, nn, T=1 pn, pp
        const float *v = vals.data() + 4 * rec++;
        const float v00 = v[0], vnn = v[1], v10 = v[2], vpp = v[3];
        double nf = 1.0;  // the file holds un-normalized elements
        if (a == b) nf /= kSqrt2;
        if (c == d) nf /= kSqrt2;
        bool ok = true;
        if (nf > 0.9 || J % 2 == 0) {
          ok = ok && w.set(J, par, -1, a, b, c, d, vpp * nf);
          ok = ok && w.set(J, par, 1, a + 1, b + 1, c + 1, d + 1, vnn * nf);
          ok = ok && w.pn_from_iso(J, 1, a, b, c, d, v10 * nf);
        }
        if (nf > 0.9 || J % 2 != 0) ok = ok && w.pn_from_iso(J, 0, a, b, c, d, v00 * nf);
        if (!ok) bad = true;
      });
    }
  };
  const int nthr = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (int t = 1; t < nthr; ++t) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  if (bad) return "a ket of the file is missing from its channel";
  const long nread = first[nlj];
  if (nrecords) *nrecords = nread;
  return "";
}

}  // namespace imsrg
