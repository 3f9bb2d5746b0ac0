// api.cpp -- C ABI of libimsrg_hip.so: contexts, device operators, the
// commutator launch sequence, BCH loops, Magnus-Euler flow, batching.
#include "../../include/imsrg_hip.h"

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include <dlfcn.h>
#include <unistd.h>

#include "basis.hpp"
#include "launch.hpp"
#include "noalloc.hpp"
#include "readwrite.hpp"

using namespace imsrg;

struct imsrg_dop {
  imsrg_ctx *ctx;
  double *buf;
  int herm;
};

struct BatchCache {
  std::vector<long long> key;
  PairPtrs *d_tab = nullptr;
  int4 *d_items = nullptr;
  size_t off[10] = {0};
  int cnt[10] = {0};
  double *work = nullptr;
  int maxn = 0, maxnch = 0, maxjmax = 0, maxnorb = 0, maxn1pairs = 0, maxncls = 0, maxbin = 0, maxent = 0;
  long maxn122 = 0;
  void release() {
    if (d_tab) (void)hipFree(d_tab);
    if (d_items) (void)hipFree(d_items);
    if (work) (void)hipFree(work);
    d_tab = nullptr; d_items = nullptr; work = nullptr;
    maxn = maxnch = maxjmax = maxnorb = maxn1pairs = maxncls = maxbin = maxent = 0;
    maxn122 = 0;
    key.clear();
  }
};

// A small pool of host threads for the memcpy side of view transfers:
// run(n, f) calls f(0) .. f(n-1) on the workers and the caller and returns
// when all are done.
class CopyPool {
 public:
  explicit CopyPool(int nworkers) {
    for (int i = 0; i < nworkers; ++i) th_.emplace_back([this] { work(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  void run(int n, const std::function<void(int)> &f) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &f;
      n_ = n;
      next_ = 0;
      busy_ = (int)th_.size();
      ++gen_;
    }
    cv_.notify_all();
    for (int i; (i = next_.fetch_add(1)) < n;) f(i);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [this] { return busy_ == 0; });
    job_ = nullptr;
  }

 private:
  void work() {
    long seen = 0;
    for (;;) {
      const std::function<void(int)> *job;
      int n;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        job = job_;
        n = n_;
      }
      for (int i; (i = next_.fetch_add(1)) < n;) (*job)(i);
      std::lock_guard<std::mutex> lk(mu_);
      if (--busy_ == 0) done_.notify_one();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int)> *job_ = nullptr;
  int n_ = 0, busy_ = 0;
  long gen_ = 0;
  bool stop_ = false;
  std::atomic<int> next_{0};
};

// One captured commutator launch sequence (operator buffers, Hermiticity,
// terms and precision are baked into its kernel arguments).
struct GraphEntry {
  const double *x, *y, *z;
  int hx, hy;
  unsigned mask;
  int prec;
  hipGraphExec_t exec;
};
constexpr size_t kMaxGraphs = 16;

struct imsrg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream for the pp/hh GEMM when it overlaps the one-body, Pandya and
  // ph phases (tuning field overlap_pphh); default priority, like `stream`
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  // orders a batched call against the streams of every context in the batch
  hipEvent_t batch_ev = nullptr;
  // order imsrg_hip_commutator_ss_dev_async against the caller's stream
  hipEvent_t async_in = nullptr, async_out = nullptr;
  long long serial = 0;  // unique per context of the process (batch cache keys)
  HostBasis hb;
  DeviceBasis db;
  double *Xt = nullptr, *Yt = nullptr, *Zbar = nullptr;
  double *scratch = nullptr;
  double2 *c122 = nullptr;
  double *obpart = nullptr;  // one-body partial sums (nobpart)
  double *diag = nullptr;    // inverse-Pandya diagonal terms (ndiag)
  double *panel = nullptr;   // bf16x3 operand panels of the mixed pp/hh GEMM (allocated on first use)
  double *c122w = nullptr;   // comm122 W in the two-body layout (tbsize + 8)
  double *partials = nullptr;
  // basis transformation: channel matrices D, one channel's product, C and a
  // one-body product (allocated on first use), and the rocBLAS handle
  double *xform = nullptr;
  long xform_n = 0;
  rocblas_handle blas = nullptr;
  // ket groups of the block-sparse transformation (built on first use)
  struct {
    bool built = false, usable = false;
    XformDev d{};
    std::vector<void *> allocs;
    int2 *right = nullptr, *left = nullptr;
    int nright = 0, nleft = 0, R = 64, kc = 4;
    long dg_size = 0;
    std::vector<int> blk;  // one-body block of each orbit
  } xf;
  double *norm_dev = nullptr;
  double *norm_host = nullptr;  // pinned
  int npart = 0;
  std::string err;
  bool profiling = false;
  // phase events of profiled commutators: a ring of sets recorded without host
  // synchronisation, read (and summed) by imsrg_hip_phase_times; marks 0..8 on
  // the main stream, 9, 10, 11 around the pp/hh GEMM and comm122 on the side stream
  static constexpr int kRing = 32, kMarks = 12;
  std::vector<std::array<hipEvent_t, kMarks>> ring;
  int ring_head = 0, ring_count = 0;
  bool ring_side[kRing] = {};  // set recorded with the pp/hh GEMM on the side stream
  double phase_sum[8] = {0};
  int phase_n = 0;
  imsrg_dop *stage[3] = {nullptr, nullptr, nullptr};
  imsrg_dop *tmp[4] = {nullptr, nullptr, nullptr, nullptr};
  // host staging of view transfers: two pinned buffers (the host copy of one
  // chunk overlaps the DMA of the other) and a pool of host copy threads
  char *stage_buf[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  std::unique_ptr<CopyPool> copy_pool;
  BatchCache batch;
  bool use_graphs = false;
  std::vector<GraphEntry> graphs;  // oldest first
  long lists_gen = 0;              // bumped whenever the work lists are rebuilt
};

static int fail(imsrg_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}
#define HIPCHK(ctx, call)                                                                      \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) return fail(ctx, IMSRG_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

static std::string g_err;

extern "C" const char *imsrg_hip_last_error(const imsrg_ctx *ctx) {
  return ctx ? ctx->err.c_str() : g_err.c_str();
}

// ---------------------------------------------------------------------------
// rocBLAS, opened on the first basis transformation (nothing else needs it): the
// channel products of imsrg_hip_op_transform_basis are plain DGEMMs.
namespace {
struct BlasLib {
  decltype(&rocblas_create_handle) create = nullptr;
  decltype(&rocblas_destroy_handle) destroy = nullptr;
  decltype(&rocblas_set_stream) set_stream = nullptr;
  decltype(&rocblas_dgemm) dgemm = nullptr;
  std::string err;
};
const BlasLib &blas_lib() {
  static const BlasLib lib = [] {
    BlasLib b;
    void *h = dlopen("librocblas.so.5", RTLD_NOW | RTLD_LOCAL);
    if (!h) h = dlopen("librocblas.so", RTLD_NOW | RTLD_LOCAL);
    if (!h) {
      const char *m = dlerror();
      b.err = std::string("cannot load rocBLAS: ") + (m ? m : "dlopen failed");
      return b;
    }
    b.create = reinterpret_cast<decltype(b.create)>(dlsym(h, "rocblas_create_handle"));
    b.destroy = reinterpret_cast<decltype(b.destroy)>(dlsym(h, "rocblas_destroy_handle"));
    b.set_stream = reinterpret_cast<decltype(b.set_stream)>(dlsym(h, "rocblas_set_stream"));
    b.dgemm = reinterpret_cast<decltype(b.dgemm)>(dlsym(h, "rocblas_dgemm"));
    if (!b.create || !b.destroy || !b.set_stream || !b.dgemm) b.err = "rocBLAS lacks a needed symbol";
    return b;
  }();
  return lib;
}
}  // namespace

// ---------------------------------------------------------------------------
static int ctx_init(imsrg_ctx *ctx, const imsrg_basis_desc *b, int device) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(ctx, IMSRG_E_HIP, "no HIP device");
  if (device < 0 || device >= ndev) return fail(ctx, IMSRG_E_INVALID, "device index out of range");
  ctx->device = device;
  HIPCHK(ctx, hipSetDevice(device));
  std::string e = ctx->hb.build(b->norb, b->n, b->l, b->j2, b->tz2, b->occ, b->cvq, b->emax);
  if (!e.empty()) return fail(ctx, IMSRG_E_INVALID, e);
  e = ctx->db.upload(ctx->hb, b->emax);
  if (!e.empty()) return fail(ctx, IMSRG_E_HIP, e);
  HIPCHK(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  HIPCHK(ctx, hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
  HIPCHK(ctx, hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming));
  HIPCHK(ctx, hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming));
  HIPCHK(ctx, hipEventCreateWithFlags(&ctx->batch_ev, hipEventDisableTiming));
  HIPCHK(ctx, hipEventCreateWithFlags(&ctx->async_in, hipEventDisableTiming));
  HIPCHK(ctx, hipEventCreateWithFlags(&ctx->async_out, hipEventDisableTiming));
  {  // this device's copy of the allocation-free commutator's static workspace
    std::string se;
    if (int rc = noalloc_setup(device, se)) return fail(ctx, rc, se);
  }
  static std::atomic<long long> next_serial{1};
  ctx->serial = next_serial++;
  const BasisDev &B = ctx->db.h;
  // 8 doubles of slack: the GEMM tile loader may read one element past a column
  HIPCHK(ctx, hipMalloc(&ctx->Xt, sizeof(double) * (std::max<long>(B.xsize, 1) + 8)));
  HIPCHK(ctx, hipMalloc(&ctx->Yt, sizeof(double) * (std::max<long>(B.xsize, 1) + 8)));
  HIPCHK(ctx, hipMalloc(&ctx->Zbar, sizeof(double) * std::max<long>(B.zsize, 1)));
  // cross-coupled channels without ph kets are never written and must read as 0
  HIPCHK(ctx, hipMemset(ctx->Zbar, 0, sizeof(double) * std::max<long>(B.zsize, 1)));
  HIPCHK(ctx, hipMalloc(&ctx->scratch, sizeof(double) * (B.nch + 8)));
  HIPCHK(ctx, hipMalloc(&ctx->c122, sizeof(double2) * std::max<long>(B.n122, 1)));
  HIPCHK(ctx, hipMalloc(&ctx->obpart, sizeof(double) * std::max(B.nobpart, 1)));
  HIPCHK(ctx, hipMalloc(&ctx->diag, sizeof(double) * std::max(B.ndiag, 1)));
  HIPCHK(ctx, hipMalloc(&ctx->c122w, sizeof(double) * (B.tbsize + 8)));
  ctx->npart = 2049;  // 2048 two-body chunks and the one-body block
  HIPCHK(ctx, hipMalloc(&ctx->partials, sizeof(double) * ctx->npart));
  HIPCHK(ctx, hipMalloc(&ctx->norm_dev, sizeof(double) * 2));
  HIPCHK(ctx, hipHostMalloc(&ctx->norm_host, sizeof(double) * 2, hipHostMallocDefault));
  ctx->ring.resize(imsrg_ctx::kRing);
  for (auto &set : ctx->ring)
    for (auto &ev : set) HIPCHK(ctx, hipEventCreate(&ev));
  return IMSRG_OK;
}

extern "C" int imsrg_hip_ctx_create(const imsrg_basis_desc *b, int device, imsrg_ctx **out) {
  if (!b || !out) {
    g_err = "null argument";
    return IMSRG_E_INVALID;
  }
  *out = nullptr;
  if (b->norb <= 0 || !b->n || !b->l || !b->j2 || !b->tz2 || !b->occ || !b->cvq || b->emax < 0) {
    g_err = "invalid basis description";
    return IMSRG_E_INVALID;
  }
  imsrg_ctx *ctx = new imsrg_ctx();
  int rc = ctx_init(ctx, b, device);
  if (rc != IMSRG_OK) {
    g_err = ctx->err;
    imsrg_hip_ctx_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return IMSRG_OK;
}

extern "C" int imsrg_hip_ctx_create_ho(int emax, double A, double Z, int device, imsrg_ctx **out) {
  if (emax < 0 || emax > 20 || A < 0 || Z < 0 || Z > A) {
    g_err = "invalid emax/A/Z";
    return IMSRG_E_INVALID;
  }
  std::vector<int> n, l, j2, tz2, cvq;
  std::vector<double> occ;
  HostBasis::ho_orbits(emax, A, Z, n, l, j2, tz2, occ, cvq);
  imsrg_basis_desc d{(int)n.size(), n.data(), l.data(), j2.data(), tz2.data(), occ.data(), cvq.data(), emax};
  return imsrg_hip_ctx_create(&d, device, out);
}

extern "C" int imsrg_hip_ctx_destroy(imsrg_ctx *ctx) {
  if (!ctx) return IMSRG_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->side) (void)hipStreamSynchronize(ctx->side);
  for (auto *&o : ctx->stage) if (o) { imsrg_hip_op_destroy(o); o = nullptr; }
  for (auto *&o : ctx->tmp) if (o) { imsrg_hip_op_destroy(o); o = nullptr; }
  ctx->batch.release();
  if (ctx->Xt) (void)hipFree(ctx->Xt);
  if (ctx->Yt) (void)hipFree(ctx->Yt);
  if (ctx->Zbar) (void)hipFree(ctx->Zbar);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  if (ctx->c122) (void)hipFree(ctx->c122);
  if (ctx->obpart) (void)hipFree(ctx->obpart);
  if (ctx->diag) (void)hipFree(ctx->diag);
  if (ctx->panel) (void)hipFree(ctx->panel);
  if (ctx->c122w) (void)hipFree(ctx->c122w);
  if (ctx->partials) (void)hipFree(ctx->partials);
  if (ctx->xform) (void)hipFree(ctx->xform);
  for (void *p : ctx->xf.allocs) (void)hipFree(p);
  if (ctx->blas) blas_lib().destroy(ctx->blas);
  if (ctx->norm_dev) (void)hipFree(ctx->norm_dev);
  if (ctx->norm_host) (void)hipHostFree(ctx->norm_host);
  for (int i = 0; i < 2; ++i) {
    if (ctx->stage_buf[i]) (void)hipHostFree(ctx->stage_buf[i]);
    if (ctx->stage_ev[i]) (void)hipEventDestroy(ctx->stage_ev[i]);
  }
  ctx->copy_pool.reset();
  for (auto &set : ctx->ring)
    for (auto &ev : set) if (ev) (void)hipEventDestroy(ev);
  for (auto &g : ctx->graphs) (void)hipGraphExecDestroy(g.exec);
  ctx->graphs.clear();
  ctx->db.release();
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  if (ctx->fork) (void)hipEventDestroy(ctx->fork);
  if (ctx->join) (void)hipEventDestroy(ctx->join);
  if (ctx->batch_ev) (void)hipEventDestroy(ctx->batch_ev);
  if (ctx->async_in) (void)hipEventDestroy(ctx->async_in);
  if (ctx->async_out) (void)hipEventDestroy(ctx->async_out);
  delete ctx;
  return IMSRG_OK;
}

extern "C" int imsrg_hip_ctx_info(const imsrg_ctx *ctx, int *norb, int *nch, long *tbsize) {
  if (!ctx) return IMSRG_E_INVALID;
  if (norb) *norb = ctx->hb.norb;
  if (nch) *nch = ctx->hb.nch;
  if (tbsize) *tbsize = ctx->hb.tbsize;
  return IMSRG_OK;
}
extern "C" int imsrg_hip_ctx_orbit(const imsrg_ctx *ctx, int i, int *n, int *l, int *j2, int *tz2, double *occ,
                                   int *cvq) {
  if (!ctx || i < 0 || i >= ctx->hb.norb) return IMSRG_E_INVALID;
  *n = ctx->hb.n[i]; *l = ctx->hb.l[i]; *j2 = ctx->hb.j2[i]; *tz2 = ctx->hb.tz2[i];
  *occ = ctx->hb.occ[i]; *cvq = ctx->hb.cvq[i];
  return IMSRG_OK;
}
extern "C" int imsrg_hip_ctx_channel(const imsrg_ctx *ctx, int ch, int *J, int *par, int *tz, int *nk, long *off) {
  if (!ctx || ch < 0 || ch >= ctx->hb.nch) return IMSRG_E_INVALID;
  const HostChannel &c = ctx->hb.ch[ch];
  *J = c.J; *par = c.par; *tz = c.tz; *nk = (int)c.kets.size(); *off = ctx->hb.ch_off[ch];
  return IMSRG_OK;
}
extern "C" int imsrg_hip_ctx_channel_ket(const imsrg_ctx *ctx, int ch, int i, int *p, int *q) {
  if (!ctx || ch < 0 || ch >= ctx->hb.nch) return IMSRG_E_INVALID;
  const HostChannel &c = ctx->hb.ch[ch];
  if (i < 0 || i >= (int)c.kets.size()) return IMSRG_E_INVALID;
  *p = ctx->hb.kp[c.kets[i]]; *q = ctx->hb.kq[c.kets[i]];
  return IMSRG_OK;
}
extern "C" int imsrg_hip_ctx_sixj(imsrg_ctx *ctx, int jj1, int jj2, int J3, int jj4, int jj5, int J6, double *v) {
  if (!ctx || !v) return IMSRG_E_INVALID;
  const BasisDev &B = ctx->db.h;
  int jjm = 2 * B.d1 - 1;
  for (int jj : {jj1, jj2, jj4, jj5})
    if (jj < 1 || jj > jjm || (jj & 1) == 0) return fail(ctx, IMSRG_E_INVALID, "6j: half-integer argument out of range");
  if (J3 < 0 || J3 >= B.d2 || J6 < 0 || J6 >= B.d2) return fail(ctx, IMSRG_E_INVALID, "6j: integer argument out of range");
  long idx = sixj_index(B.d1, B.d2, jj1, jj2, J3, jj4, jj5, J6);
  HIPCHK(ctx, hipMemcpy(v, B.sixj + idx, sizeof(double), hipMemcpyDeviceToHost));
  return IMSRG_OK;
}
extern "C" int imsrg_hip_synchronize(imsrg_ctx *ctx) {
  if (!ctx) return IMSRG_E_INVALID;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return IMSRG_OK;
}

// ---------------------------------------------------------------------------
extern "C" int imsrg_hip_op_create(imsrg_ctx *ctx, int herm, imsrg_dop **out) {
  if (!ctx || !out) return IMSRG_E_INVALID;
  *out = nullptr;
  if (herm != 1 && herm != -1 && herm != 0) return fail(ctx, IMSRG_E_INVALID, "herm must be +1, -1 or 0");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  imsrg_dop *o = new imsrg_dop{ctx, nullptr, herm};
  hipError_t e = hipMalloc(&o->buf, sizeof(double) * ctx->db.h.opsize);
  if (e != hipSuccess) {
    delete o;
    return fail(ctx, IMSRG_E_NOMEM, "hipMalloc of an operator failed");
  }
  e = hipMemsetAsync(o->buf, 0, sizeof(double) * ctx->db.h.opsize, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(o->buf);
    delete o;
    return fail(ctx, IMSRG_E_HIP, "operator clear failed");
  }
  *out = o;
  return IMSRG_OK;
}
extern "C" int imsrg_hip_op_destroy(imsrg_dop *op) {
  if (!op) return IMSRG_OK;
  if (op->buf) (void)hipFree(op->buf);
  delete op;
  return IMSRG_OK;
}
extern "C" int imsrg_hip_op_get_herm(const imsrg_dop *op, int *h) {
  if (!op || !h) return IMSRG_E_INVALID;
  *h = op->herm;
  return IMSRG_OK;
}
extern "C" int imsrg_hip_op_set_herm(imsrg_dop *op, int h) {
  if (!op || (h != 1 && h != -1 && h != 0)) return IMSRG_E_INVALID;
  op->herm = h;
  return IMSRG_OK;
}

static int view_ok(imsrg_ctx *ctx, const imsrg_op_view *v, bool need_two) {
  if (!v) return fail(ctx, IMSRG_E_INVALID, "null operator view");
  if (v->rank_J != 0 || v->rank_T != 0 || v->parity != 0 || v->reduced)
    return fail(ctx, IMSRG_E_UNSUPPORTED, "only scalar, parity-conserving, non-reduced operators are supported");
  if (v->herm != 1 && v->herm != -1)
    return fail(ctx, IMSRG_E_UNSUPPORTED, "only Hermitian or anti-Hermitian operators are supported");
  if (v->particle_rank < 1 || v->particle_rank > 2)
    return fail(ctx, IMSRG_E_UNSUPPORTED, "particle rank must be 1 or 2");
  if (!v->one_body) return fail(ctx, IMSRG_E_INVALID, "null one-body pointer");
  if (need_two && v->particle_rank == 2 && !v->two_body) return fail(ctx, IMSRG_E_INVALID, "null two-body table");
  return IMSRG_OK;
}

// View transfers.  The pieces of an operator view (ZeroBody, the one-body
// matrix, one block per channel) map to device ranges of the operator buffer;
// they are packed into chunks of at most kStageBytes that are contiguous on
// the device, and each chunk moves through one of two pinned buffers: while the
// DMA of one chunk runs, the host threads copy the next one.
constexpr size_t kStageBytes = size_t(32) << 20;

struct Piece {
  char *host;
  size_t dev;  // byte offset in the operator buffer
  size_t bytes;
};
struct Chunk {
  size_t dev = 0, bytes = 0;
  std::vector<Piece> pieces;  // host ranges, consecutive on the device
};

static std::vector<Chunk> make_chunks(const std::vector<Piece> &in) {
  std::vector<Chunk> out;
  for (Piece p : in) {
    while (p.bytes > 0) {
      if (out.empty() || out.back().bytes == kStageBytes || out.back().dev + out.back().bytes != p.dev) {
        out.emplace_back();
        out.back().dev = p.dev;
      }
      Chunk &c = out.back();
      const size_t take = std::min(p.bytes, kStageBytes - c.bytes);
      c.pieces.push_back(Piece{p.host, p.dev, take});
      c.bytes += take;
      p.host += take;
      p.dev += take;
      p.bytes -= take;
    }
  }
  return out;
}

static int ensure_staging(imsrg_ctx *ctx) {
  for (int i = 0; i < 2; ++i) {
    if (!ctx->stage_buf[i]) HIPCHK(ctx, hipHostMalloc((void **)&ctx->stage_buf[i], kStageBytes, hipHostMallocDefault));
    if (!ctx->stage_ev[i]) HIPCHK(ctx, hipEventCreateWithFlags(&ctx->stage_ev[i], hipEventDisableTiming));
  }
  if (!ctx->copy_pool) {
    int nt = 7;  // workers besides the calling thread
    if (const char *e = std::getenv("IMSRG_HIP_COPY_THREADS")) nt = std::max(0, std::atoi(e) - 1);
    ctx->copy_pool.reset(new CopyPool(nt));
  }
  return IMSRG_OK;
}

// host memcpy of one chunk's pieces (to_stage: host -> pinned, else pinned -> host),
// split into parts of about 1 MiB over the copy threads
static void copy_chunk(imsrg_ctx *ctx, const Chunk &c, char *stage, bool to_stage) {
  struct Part {
    char *host;
    char *pinned;
    size_t bytes;
  };
  std::vector<Part> parts;
  constexpr size_t kPart = size_t(1) << 20;
  for (const Piece &p : c.pieces)
    for (size_t o = 0; o < p.bytes; o += kPart)
      parts.push_back(Part{p.host + o, stage + (p.dev - c.dev) + o, std::min(kPart, p.bytes - o)});
  ctx->copy_pool->run((int)parts.size(), [&](int i) {
    const Part &q = parts[i];
    if (to_stage)
      std::memcpy(q.pinned, q.host, q.bytes);
    else
      std::memcpy(q.host, q.pinned, q.bytes);
  });
}

static std::vector<Piece> view_pieces(imsrg_ctx *ctx, const imsrg_op_view *v, double *zero, bool two) {
  const BasisDev &B = ctx->db.h;
  std::vector<Piece> pc;
  pc.push_back(Piece{(char *)zero, 0, sizeof(double)});
  pc.push_back(Piece{(char *)v->one_body, sizeof(double) * B.one_off, sizeof(double) * B.norb * B.norb});
  if (two)
    for (int c = 0; c < ctx->hb.nch; ++c) {
      const size_t n = ctx->hb.ch[c].kets.size();
      if (n == 0 || !v->two_body[c]) continue;
      pc.push_back(Piece{(char *)v->two_body[c], sizeof(double) * (B.two_off + ctx->hb.ch_off[c]),
                         sizeof(double) * n * n});
    }
  return pc;
}

extern "C" int imsrg_hip_op_upload(imsrg_dop *op, const imsrg_op_view *v) {
  if (!op) return IMSRG_E_INVALID;
  imsrg_ctx *ctx = op->ctx;
  int rc = view_ok(ctx, v, true);
  if (rc) return rc;
  const BasisDev &B = ctx->db.h;
  const bool two = v->particle_rank == 2;
  if (two)
    for (int c = 0; c < ctx->hb.nch; ++c)
      if (!ctx->hb.ch[c].kets.empty() && !v->two_body[c])
        return fail(ctx, IMSRG_E_INVALID, "null two-body block for a non-empty channel");
  if ((rc = ensure_staging(ctx))) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  double zero = v->zero_body;
  if (!two) HIPCHK(ctx, launch_fill(op->buf + B.two_off, 0.0, B.tbsize, ctx->stream));
  const std::vector<Chunk> ch = make_chunks(view_pieces(ctx, v, &zero, two));
  char *dev = reinterpret_cast<char *>(op->buf);
  for (size_t c = 0; c < ch.size(); ++c) {
    const int b = c & 1;
    if (c >= 2) HIPCHK(ctx, hipEventSynchronize(ctx->stage_ev[b]));  // its previous DMA is done
    copy_chunk(ctx, ch[c], ctx->stage_buf[b], true);
    HIPCHK(ctx, hipMemcpyAsync(dev + ch[c].dev, ctx->stage_buf[b], ch[c].bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->stage_ev[b], ctx->stream));
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  op->herm = v->herm;
  return IMSRG_OK;
}

extern "C" int imsrg_hip_op_download(imsrg_dop *op, imsrg_op_view *v) {
  if (!op || !v || !v->one_body) return IMSRG_E_INVALID;
  imsrg_ctx *ctx = op->ctx;
  if (int rc = ensure_staging(ctx)) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  double zero = 0.0;
  const std::vector<Chunk> ch = make_chunks(view_pieces(ctx, v, &zero, v->two_body != nullptr));
  const char *dev = reinterpret_cast<const char *>(op->buf);
  auto issue = [&](size_t c) -> hipError_t {
    const int b = c & 1;
    hipError_t e = hipMemcpyAsync(ctx->stage_buf[b], dev + ch[c].dev, ch[c].bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->stage_ev[b], ctx->stream);
    return e;
  };
  if (!ch.empty()) HIPCHK(ctx, issue(0));
  for (size_t c = 0; c < ch.size(); ++c) {
    // the other buffer's host copy (chunk c - 1) is done: its next DMA may start
    if (c + 1 < ch.size()) HIPCHK(ctx, issue(c + 1));
    HIPCHK(ctx, hipEventSynchronize(ctx->stage_ev[c & 1]));
    copy_chunk(ctx, ch[c], ctx->stage_buf[c & 1], false);
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  v->zero_body = zero;
  v->herm = op->herm;
  return IMSRG_OK;
}

extern "C" int imsrg_hip_op_upload_packed(imsrg_dop *op, double zero, const double *one, const double *two, int herm) {
  if (!op || !one || !two) return IMSRG_E_INVALID;
  imsrg_ctx *ctx = op->ctx;
  if (herm != 1 && herm != -1 && herm != 0) return fail(ctx, IMSRG_E_INVALID, "herm must be +1, -1 or 0");
  const BasisDev &B = ctx->db.h;
  HIPCHK(ctx, hipMemcpyAsync(op->buf, &zero, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(op->buf + B.one_off, one, sizeof(double) * B.norb * B.norb, hipMemcpyHostToDevice,
                             ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(op->buf + B.two_off, two, sizeof(double) * B.tbsize, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  op->herm = herm;
  return IMSRG_OK;
}

extern "C" int imsrg_hip_op_download_packed(imsrg_dop *op, double *zero, double *one, double *two, int *herm) {
  if (!op) return IMSRG_E_INVALID;
  imsrg_ctx *ctx = op->ctx;
  const BasisDev &B = ctx->db.h;
  if (zero) HIPCHK(ctx, hipMemcpyAsync(zero, op->buf, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (one)
    HIPCHK(ctx, hipMemcpyAsync(one, op->buf + B.one_off, sizeof(double) * B.norb * B.norb, hipMemcpyDeviceToHost,
                               ctx->stream));
  if (two)
    HIPCHK(ctx, hipMemcpyAsync(two, op->buf + B.two_off, sizeof(double) * B.tbsize, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (herm) *herm = op->herm;
  return IMSRG_OK;
}

extern "C" int imsrg_hip_op_fill_random(imsrg_dop *op, int herm, unsigned long long seed) {
  if (!op || (herm != 1 && herm != -1)) return IMSRG_E_INVALID;
  imsrg_ctx *ctx = op->ctx;
  HIPCHK(ctx, launch_fill_random(ctx->db.d, ctx->db.h, op->buf, herm, seed, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  op->herm = herm;
  return IMSRG_OK;
}

static int same_ctx(const imsrg_dop *a, const imsrg_dop *b) { return a && b && a->ctx == b->ctx; }

extern "C" int imsrg_hip_op_copy(imsrg_dop *dst, const imsrg_dop *src) {
  if (!same_ctx(dst, src)) return IMSRG_E_INVALID;
  imsrg_ctx *ctx = dst->ctx;
  if (dst != src)
    HIPCHK(ctx, hipMemcpyAsync(dst->buf, src->buf, sizeof(double) * ctx->db.h.opsize, hipMemcpyDeviceToDevice,
                               ctx->stream));
  dst->herm = src->herm;
  return IMSRG_OK;
}
extern "C" int imsrg_hip_op_scale(imsrg_dop *op, double f) {
  if (!op) return IMSRG_E_INVALID;
  HIPCHK(op->ctx, launch_scale(op->buf, f, op->ctx->db.h.opsize, op->ctx->stream));
  return IMSRG_OK;
}
extern "C" int imsrg_hip_op_axpy(imsrg_dop *y, double a, const imsrg_dop *x) {
  if (!same_ctx(y, x)) return IMSRG_E_INVALID;
  HIPCHK(y->ctx, launch_axpy(y->buf, a, x->buf, y->ctx->db.h.opsize, y->ctx->stream));
  return IMSRG_OK;
}
extern "C" int imsrg_hip_op_zero(imsrg_dop *op) {
  if (!op) return IMSRG_E_INVALID;
  HIPCHK(op->ctx, launch_fill(op->buf, 0.0, op->ctx->db.h.opsize, op->ctx->stream));
  return IMSRG_OK;
}

// norm of op; with axpy_to, axpy_to += axpy_a * op in the same pass over op
static int op_norm(imsrg_dop *op, double *n1, double *n2, double *n, imsrg_dop *axpy_to = nullptr,
                   double axpy_a = 0.0) {
  imsrg_ctx *ctx = op->ctx;
  if (axpy_to)
    HIPCHK(ctx, launch_axpy_norm(ctx->db.d, axpy_to->buf, axpy_a, op->buf, ctx->partials, ctx->npart,
                                 ctx->norm_dev, ctx->stream));
  else
    HIPCHK(ctx, launch_norm(ctx->db.d, ctx->db.h, op->buf, ctx->partials, ctx->npart, ctx->norm_dev, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->norm_host, ctx->norm_dev, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  double a = std::sqrt(ctx->norm_host[0]), b = std::sqrt(ctx->norm_host[1]);
  if (n1) *n1 = a;
  if (n2) *n2 = b;
  if (n) *n = std::sqrt(a * a + b * b);
  if (!std::isfinite(ctx->norm_host[0]) || !std::isfinite(ctx->norm_host[1]))
    return fail(ctx, IMSRG_E_NONFINITE, "operator holds non-finite values");
  return IMSRG_OK;
}
extern "C" int imsrg_hip_op_norm(imsrg_dop *op, double *n1, double *n2, double *n) {
  if (!op) return IMSRG_E_INVALID;
  return op_norm(op, n1, n2, n);
}

// Operator::DoNormalOrdering (sign +1) / UndoNormalOrdering (sign -1) of a scalar
// operator over the context's reference (src/Operator.cc:392-415, 461-579).
extern "C" int imsrg_hip_op_normal_order(imsrg_dop *out, const imsrg_dop *op, int sign) {
  if (!out || !op) return IMSRG_E_INVALID;
  if (!same_ctx(out, op)) return IMSRG_E_INVALID;
  imsrg_ctx *ctx = out->ctx;
  if (sign != 1 && sign != -1) return fail(ctx, IMSRG_E_INVALID, "normal-ordering sign must be +1 or -1");
  if (op->herm != 1 && op->herm != -1)
    return fail(ctx, IMSRG_E_UNSUPPORTED, "normal ordering needs a Hermitian or anti-Hermitian operator");
  // the one-body sum runs only when the two-body norm exceeds 1e-7 (src/Operator.cc:493):
  // the norm stays on the device, where the kernel tests it
  HIPCHK(ctx, launch_norm(ctx->db.d, ctx->db.h, op->buf, ctx->partials, ctx->npart, ctx->norm_dev, ctx->stream));
  if (out != op) HIPCHK(ctx, hipMemcpyAsync(out->buf, op->buf, sizeof(double) * ctx->db.h.opsize,
                                            hipMemcpyDeviceToDevice, ctx->stream));
  out->herm = op->herm;
  HIPCHK(ctx, launch_normal_order(ctx->db.d, ctx->db.h, out->buf, sign, op->herm, ctx->norm_dev, ctx->stream));
  return IMSRG_OK;
}

// Ket groups of the block-sparse transformation: per channel, the kets whose orbits
// lie in the same unordered pair of one-body blocks; work items of the two passes.
template <typename T>
static T *xf_upload(imsrg_ctx *ctx, const std::vector<T> &v, hipError_t &e) {
  T *p = nullptr;
  if (e != hipSuccess) return nullptr;
  e = hipMalloc(&p, sizeof(T) * std::max<size_t>(v.size(), 1));
  if (e != hipSuccess) return nullptr;
  ctx->xf.allocs.push_back(p);
  if (!v.empty()) e = hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
  return p;
}

static int xform_tables(imsrg_ctx *ctx) {
  auto &xf = ctx->xf;
  if (xf.built) return IMSRG_OK;
  const HostBasis &hb = ctx->hb;
  const int norb = hb.norb;
  xf.blk.assign(norb, -1);
  for (int o = 0; o < norb; ++o) xf.blk[o] = hb.ob[o].front();  // the block's first orbit names it
  std::vector<int> g_start{0}, g_kets, g_ch, k_start(hb.nch + 1, 0), k_grp, k_pos;
  std::vector<long> g_doff;
  std::vector<int2> right, left;
  long doff = 0;
  int maxm = 0, maxn = 0;
  for (int c = 0; c < hb.nch; ++c) {
    const auto &kets = hb.ch[c].kets;
    const int n = (int)kets.size();
    k_start[c] = (int)k_grp.size();
    maxn = std::max(maxn, n);
    std::map<std::pair<int, int>, std::vector<int>> groups;
    for (int i = 0; i < n; ++i) {
      const int bp = xf.blk[hb.kp[kets[i]]], bq = xf.blk[hb.kq[kets[i]]];
      groups[{std::min(bp, bq), std::max(bp, bq)}].push_back(i);
    }
    k_grp.resize(k_start[c] + n);
    k_pos.resize(k_start[c] + n);
    for (const auto &kv : groups) {
      const int G = (int)g_ch.size(), m = (int)kv.second.size();
      g_ch.push_back(c);
      g_doff.push_back(doff);
      doff += (long)m * m;
      maxm = std::max(maxm, m);
      for (int pos = 0; pos < m; ++pos) {
        g_kets.push_back(kv.second[pos]);
        k_grp[k_start[c] + kv.second[pos]] = G;
        k_pos[k_start[c] + kv.second[pos]] = pos;
      }
      g_start.push_back((int)g_kets.size());
    }
  }
  k_start[hb.nch] = (int)k_grp.size();
  xf.built = true;
  // LDS of one work-group stays within 64 KiB; beyond that the dense path runs
  xf.R = maxm <= 48 ? 64 : 32;
  xf.kc = maxn <= 2048 ? 4 : maxn <= 4096 ? 2 : 1;
  if (sizeof(double) * ((size_t)maxm * xf.R + (size_t)maxm * maxm) > 65536 || (size_t)maxn * 8 * xf.kc > 65536)
    return IMSRG_OK;
  for (int G = 0; G < (int)g_ch.size(); ++G)
    for (int r0 = 0; r0 < (int)hb.ch[g_ch[G]].kets.size(); r0 += xf.R) right.push_back(make_int2(G, r0));
  for (int c = 0; c < hb.nch; ++c)
    for (int b0 = 0; b0 < (int)hb.ch[c].kets.size(); b0 += xf.kc) left.push_back(make_int2(c, b0));
  hipError_t e = hipSuccess;
  XformDev &d = xf.d;
  d.g_start = xf_upload(ctx, g_start, e);
  d.g_kets = xf_upload(ctx, g_kets, e);
  d.g_ch = xf_upload(ctx, g_ch, e);
  d.g_doff = xf_upload(ctx, g_doff, e);
  d.k_start = xf_upload(ctx, k_start, e);
  d.k_grp = xf_upload(ctx, k_grp, e);
  d.k_pos = xf_upload(ctx, k_pos, e);
  xf.right = xf_upload(ctx, right, e);
  xf.left = xf_upload(ctx, left, e);
  if (e != hipSuccess) return fail(ctx, IMSRG_E_HIP, std::string("basis transformation tables: ") + hipGetErrorString(e));
  d.ngroups = (int)g_ch.size();
  d.maxm = maxm;
  xf.nright = (int)right.size();
  xf.nleft = (int)left.size();
  xf.dg_size = doff;
  xf.usable = true;
  return IMSRG_OK;
}

// HartreeFock::TransformToHFBasis of a scalar operator (src/HartreeFock.cc:905-1000):
// OneBody' = C^T OneBody C and, per channel, V' = D^T V D with the matrices D of
// k_hf_D; the zero-body part and Hermiticity carry over.  C: host, norb x norb,
// column-major.  Synchronous: returns when out holds the result.
extern "C" int imsrg_hip_op_transform_basis(imsrg_dop *out, const imsrg_dop *op, const double *C) {
  if (!out || !op || !C) return IMSRG_E_INVALID;
  if (!same_ctx(out, op)) return IMSRG_E_INVALID;
  imsrg_ctx *ctx = out->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const BasisDev &B = ctx->db.h;
  const long n1 = (long)B.norb * B.norb;
  if (int rc = xform_tables(ctx)) return rc;
  // a C that is block diagonal in (l, j2, tz2) takes the block-sparse passes
  bool blockdiag = ctx->xf.usable;
  for (int a = 0; a < B.norb && blockdiag; ++a)
    for (int i = 0; i < B.norb; ++i)
      if (ctx->xf.blk[i] != ctx->xf.blk[a] && C[(long)a * B.norb + i] != 0.0) {
        blockdiag = false;
        break;
      }
  const long need = B.tbsize + std::max((long)B.maxn * B.maxn, ctx->xf.dg_size) + 2 * n1;
  if (ctx->xform_n < need) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->xform) HIPCHK(ctx, hipFree(ctx->xform));
    ctx->xform = nullptr;
    ctx->xform_n = 0;
    HIPCHK(ctx, hipMalloc(&ctx->xform, sizeof(double) * need));
    ctx->xform_n = need;
  }
  if (blockdiag) {
    double *T = ctx->xform, *Dg = T + B.tbsize, *Cd = Dg + std::max((long)B.maxn * B.maxn, ctx->xf.dg_size);
    double *T1 = Cd + n1;
    HIPCHK(ctx, hipMemcpyAsync(Cd, C, sizeof(double) * n1, hipMemcpyHostToDevice, ctx->stream));
    if (out != op) {  // the zero-body part and the slack around the one- and two-body parts
      const long rg[3][2] = {{0, B.one_off}, {B.one_off + n1, B.two_off}, {B.two_off + B.tbsize, B.opsize}};
      for (const auto &r : rg)
        if (r[1] > r[0])
          HIPCHK(ctx, hipMemcpyAsync(out->buf + r[0], op->buf + r[0], sizeof(double) * (r[1] - r[0]),
                                     hipMemcpyDeviceToDevice, ctx->stream));
    }
    out->herm = op->herm;
    HIPCHK(ctx, launch_xform_blocks(ctx->db.d, B, ctx->xf.d, Cd, op->buf, out->buf, T, Dg, T1, ctx->xf.right,
                                    ctx->xf.nright, ctx->xf.R, ctx->xf.left, ctx->xf.nleft, ctx->xf.kc, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return IMSRG_OK;
  }
  const BlasLib &bl = blas_lib();
  if (!bl.err.empty()) return fail(ctx, IMSRG_E_HIP, bl.err);
  if (!ctx->blas) {
    rocblas_handle h = nullptr;
    if (bl.create(&h) != rocblas_status_success) return fail(ctx, IMSRG_E_HIP, "rocblas_create_handle failed");
    ctx->blas = h;
  }
  if (bl.set_stream(ctx->blas, ctx->stream) != rocblas_status_success)
    return fail(ctx, IMSRG_E_HIP, "rocblas_set_stream failed");
  double *D = ctx->xform, *T = D + B.tbsize, *Cd = T + std::max((long)B.maxn * B.maxn, ctx->xf.dg_size);
  double *T1 = Cd + n1;
  HIPCHK(ctx, hipMemcpyAsync(Cd, C, sizeof(double) * n1, hipMemcpyHostToDevice, ctx->stream));
  if (out != op) HIPCHK(ctx, hipMemcpyAsync(out->buf, op->buf, sizeof(double) * B.opsize,
                                            hipMemcpyDeviceToDevice, ctx->stream));
  out->herm = op->herm;
  HIPCHK(ctx, launch_hf_D(ctx->db.d, B, Cd, D, ctx->stream));
  const double one = 1.0, zero = 0.0;
  const auto N = rocblas_operation_none, Tr = rocblas_operation_transpose;
  // one-body: T1 = O C, then O' = C^T T1
  const int no = B.norb;
  if (no > 0) {
    if (bl.dgemm(ctx->blas, N, N, no, no, no, &one, op->buf + B.one_off, no, Cd, no, &zero, T1, no) !=
            rocblas_status_success ||
        bl.dgemm(ctx->blas, Tr, N, no, no, no, &one, Cd, no, T1, no, &zero, out->buf + B.one_off, no) !=
            rocblas_status_success)
      return fail(ctx, IMSRG_E_HIP, "rocblas_dgemm failed (one-body)");
  }
  // two-body, channel by channel: T = V D, then V' = D^T T
  for (int c = 0; c < ctx->hb.nch; ++c) {
    const int n = (int)ctx->hb.ch[c].kets.size();
    if (n == 0) continue;
    const long off = ctx->hb.ch_off[c];
    if (bl.dgemm(ctx->blas, N, N, n, n, n, &one, op->buf + B.two_off + off, n, D + off, n, &zero, T, n) !=
            rocblas_status_success ||
        bl.dgemm(ctx->blas, Tr, N, n, n, n, &one, D + off, n, T, n, &zero, out->buf + B.two_off + off, n) !=
            rocblas_status_success)
      return fail(ctx, IMSRG_E_HIP, "rocblas_dgemm failed (channel " + std::to_string(c) + ")");
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return IMSRG_OK;
}

// ReadWrite::ReadBareTBME_Darmstadt (src/ReadWrite.cc:664-728, 834-957): the two-body
// part of a Hermitian operator from a Darmstadt me2j file; the rest is unchanged.
extern "C" int imsrg_hip_read_me2j(imsrg_dop *op, const char *path, int emax, int E2max, int lmax, long *nrecords) {
  if (!op || !path) return IMSRG_E_INVALID;
  imsrg_ctx *ctx = op->ctx;
  if (op->herm != 1) return fail(ctx, IMSRG_E_UNSUPPORTED, "me2j files hold Hermitian interactions");
  const BasisDev &B = ctx->db.h;
  std::vector<double> two((size_t)B.tbsize, 0.0);
  const std::string e = read_me2j(ctx->hb, path, emax, E2max, lmax, two, nrecords);
  if (!e.empty()) return fail(ctx, IMSRG_E_INVALID, e);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipMemcpyAsync(op->buf + B.two_off, two.data(), sizeof(double) * B.tbsize, hipMemcpyHostToDevice,
                             ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return IMSRG_OK;
}

// ---------------------------------------------------------------------------
// Hartree-Fock (src/HartreeFock.cc) for a two-body Hamiltonian with the reference's
// frozen occupations.  The monopoles come from the device (k_hf_mono); the
// iteration works on norb x norb matrices (column-major) on the host.

// Eigenvalues (ascending) and eigenvectors (columns) of a symmetric m x m matrix A
// (column-major), by cyclic Jacobi rotations (Golub and Van Loan, Matrix
// Computations, algorithms 8.4.1-8.4.3): the Fock blocks have a few rows each.
static void sym_eig(int m, std::vector<double> A, std::vector<double> &w, std::vector<double> &V) {
  auto at = [m](std::vector<double> &M, int r, int c) -> double & { return M[(size_t)c * m + r]; };
  std::vector<double> U((size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i) at(U, i, i) = 1.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, all = 0.0;
    for (int c = 0; c < m; ++c)
      for (int r = 0; r < m; ++r) {
        all += at(A, r, c) * at(A, r, c);
        if (r != c) off += at(A, r, c) * at(A, r, c);
      }
    if (off <= 1e-34 * all || off == 0.0) break;
    for (int p = 0; p < m - 1; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = at(A, p, q);
        if (apq == 0.0) continue;
        const double tau = (at(A, q, q) - at(A, p, p)) / (2.0 * apq);
        const double t = tau >= 0.0 ? 1.0 / (tau + std::sqrt(1.0 + tau * tau)) : -1.0 / (-tau + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t), sn = t * c;
        for (int k = 0; k < m; ++k) {  // A J
          const double x = at(A, k, p), y = at(A, k, q);
          at(A, k, p) = c * x - sn * y;
          at(A, k, q) = sn * x + c * y;
        }
        for (int k = 0; k < m; ++k) {  // J^T (A J)
          const double x = at(A, p, k), y = at(A, q, k);
          at(A, p, k) = c * x - sn * y;
          at(A, q, k) = sn * x + c * y;
        }
        for (int k = 0; k < m; ++k) {  // U J
          const double x = at(U, k, p), y = at(U, k, q);
          at(U, k, p) = c * x - sn * y;
          at(U, k, q) = sn * x + c * y;
        }
      }
  }
  std::vector<int> ord(m);
  for (int i = 0; i < m; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return at(A, x, x) < at(A, y, y); });
  w.assign(m, 0.0);
  V.assign((size_t)m * m, 0.0);
  for (int k = 0; k < m; ++k) {
    w[k] = at(A, ord[k], ord[k]);
    for (int r = 0; r < m; ++r) V[(size_t)k * m + r] = at(U, r, ord[k]);
  }
}

namespace {
struct HFState {
  int norb = 0;
  std::vector<std::vector<int>> blocks;  // orbits of each (l, j2, tz2) block, ascending
  std::vector<int> blk;                  // block of each orbit
  std::vector<int> pos;                  // position of each orbit in its block
  std::vector<long> wbase;               // quad offset of block pair (bi, ba): wbase[bi * nb + ba]
  std::vector<double> W;                 // monopoles of the quads (i, a, j, b)
  std::vector<double> KE, C, rho, Vij, F, energies;
  std::vector<int> j2;
  std::vector<int> holes;
  std::vector<double> hole_occ;
  double e1 = 0, e2 = 0;
  double &m(std::vector<double> &M, int r, int c) { return M[(size_t)c * norb + r]; }
  double Wq(int i, int a, int j, int b) const {
    const int bi = blk[i], ba = blk[a], nb = (int)blocks.size();
    const int mi = (int)blocks[bi].size(), ma = (int)blocks[ba].size();
    return W[wbase[(size_t)bi * nb + ba] + (((long)pos[i] * mi + pos[j]) * ma + pos[a]) * ma + pos[b]];
  }
  // UpdateDensityMatrix, src/HartreeFock.cc:613-617: rho = C_h diag(occ) C_h^T
  void density() {
    std::fill(rho.begin(), rho.end(), 0.0);
    for (int c = 0; c < norb; ++c)
      for (int r = 0; r < norb; ++r) {
        double acc = 0.0;
        for (size_t h = 0; h < holes.size(); ++h) acc += m(C, r, holes[h]) * hole_occ[h] * m(C, c, holes[h]);
        m(rho, r, c) = acc;
      }
  }
  // UpdateF, src/HartreeFock.cc:750-822 (two-body part): Vij(i, j) = sum_{a, b in
  // block(a)} rho(a, b) W(ia, jb) / (2j_i+1) for i <= j of one block, mirrored
  void fock() {
    std::fill(Vij.begin(), Vij.end(), 0.0);
    for (const auto &B : blocks)
      for (int i : B)
        for (int j : B) {
          if (j < i) continue;
          double acc = 0.0;
          for (int a = 0; a < norb; ++a)
            for (int b : blocks[blk[a]]) acc += m(rho, a, b) * Wq(i, a, j, b);
          m(Vij, i, j) = acc / (j2[i] + 1);
        }
    for (int c = 0; c < norb; ++c)
      for (int r = c + 1; r < norb; ++r) m(Vij, r, c) = m(Vij, c, r);
    for (size_t e = 0; e < F.size(); ++e) F[e] = KE[e] + Vij[e];
  }
  // Diagonalize + ReorderCoefficients, src/HartreeFock.cc:211-245, 866-903: each
  // block's eigenvectors in ascending energy, then columns with C(i, i) < 0 negated
  void diagonalize() {
    for (const auto &B : blocks) {
      const int mb = (int)B.size();
      std::vector<double> Fb((size_t)mb * mb), w, V;
      for (int c = 0; c < mb; ++c)
        for (int r = 0; r < mb; ++r) Fb[(size_t)c * mb + r] = m(F, B[r], B[c]);
      sym_eig(mb, Fb, w, V);
      for (int c = 0; c < mb; ++c) {
        energies[B[c]] = w[c];
        for (int r = 0; r < mb; ++r) m(C, B[r], B[c]) = V[(size_t)c * mb + r];
      }
    }
    for (int c = 0; c < norb; ++c)
      if (m(C, c, c) < 0)
        for (int r = 0; r < norb; ++r) m(C, r, c) = -m(C, r, c);
  }
  // UpdateDensityMatrix_DIIS, src/HartreeFock.cc:623-683: the error F rho - rho F and
  // the new density join the last ones kept (up to 6); with 5 kept before this one, rho
  // is the combination of the kept densities whose coefficients solve the DIIS system
  // B c = (0, .., 0, 1), B(i, j) = ||E_i E_j^T||_F over the first nsave of them, taken
  // in absolute value and normalised to sum 1
  void diis(std::deque<std::vector<double>> &derr, std::deque<std::vector<double>> &drho) {
    const size_t n2 = (size_t)norb * norb;
    std::vector<double> err(n2, 0.0);
    for (int c = 0; c < norb; ++c)
      for (int r = 0; r < norb; ++r) {
        double x = 0.0, y = 0.0;
        for (int k = 0; k < norb; ++k) {
          x += m(F, r, k) * m(rho, k, c);
          y += m(rho, r, k) * m(F, k, c);
        }
        err[(size_t)c * norb + r] = x - y;
      }
    density();
    const size_t nsave = derr.size();
    if (nsave > 5) derr.pop_front();
    derr.push_back(err);
    if (drho.size() > 5) drho.pop_front();
    drho.push_back(rho);
    if (nsave < 5) return;
    const int nb = (int)nsave + 1;
    std::vector<double> Bm((size_t)nb * nb, 1.0), rhs(nb, 0.0);
    Bm[(size_t)nsave * nb + nsave] = 0.0;
    for (size_t i = 0; i < nsave; ++i)
      for (size_t j = 0; j < nsave; ++j) {
        double acc = 0.0;
        for (int c = 0; c < norb; ++c)
          for (int r = 0; r < norb; ++r) {
            double x = 0.0;
            for (int k = 0; k < norb; ++k) x += derr[i][(size_t)k * norb + r] * derr[j][(size_t)k * norb + c];
            acc += x * x;
          }
        Bm[i * nb + j] = std::sqrt(acc);  // row-major here
      }
    rhs[nsave] = 1.0;
    // Gaussian elimination with partial pivoting
    for (int col = 0; col < nb; ++col) {
      int piv = col;
      for (int r = col + 1; r < nb; ++r)
        if (std::fabs(Bm[(size_t)r * nb + col]) > std::fabs(Bm[(size_t)piv * nb + col])) piv = r;
      if (piv != col) {
        for (int k = 0; k < nb; ++k) std::swap(Bm[(size_t)col * nb + k], Bm[(size_t)piv * nb + k]);
        std::swap(rhs[col], rhs[piv]);
      }
      const double d = Bm[(size_t)col * nb + col];
      if (d == 0.0) continue;
      for (int r = col + 1; r < nb; ++r) {
        const double f = Bm[(size_t)r * nb + col] / d;
        for (int k = col; k < nb; ++k) Bm[(size_t)r * nb + k] -= f * Bm[(size_t)col * nb + k];
        rhs[r] -= f * rhs[col];
      }
    }
    std::vector<double> cv(nb, 0.0);
    for (int r = nb - 1; r >= 0; --r) {
      double x = rhs[r];
      for (int k = r + 1; k < nb; ++k) x -= Bm[(size_t)r * nb + k] * cv[k];
      const double d = Bm[(size_t)r * nb + r];
      cv[r] = d != 0.0 ? x / d : 0.0;
    }
    double sum = 0.0;
    for (size_t i = 0; i < nsave; ++i) sum += std::fabs(cv[i]);
    std::fill(rho.begin(), rho.end(), 0.0);
    for (size_t i = 0; i < nsave; ++i) {
      const double ci = std::fabs(cv[i]) / sum;
      for (size_t e = 0; e < n2; ++e) rho[e] += ci * drho[i][e];
    }
  }
  // CalcEHF, src/HartreeFock.cc:169-190 (no three-body part)
  double ehf() {
    e1 = e2 = 0.0;
    for (int i = 0; i < norb; ++i) {
      const int jf = j2[i] + 1;
      for (int j : blocks[blk[i]]) {
        e1 += m(rho, i, j) * jf * m(KE, i, j);
        e2 += m(rho, i, j) * jf * 0.5 * m(Vij, i, j);
      }
    }
    return e1 + e2;
  }
};
}  // namespace

static double frob(const std::vector<double> &M) {
  double acc = 0.0;
  for (double x : M) acc += x * x;
  return std::sqrt(acc);
}

extern "C" int imsrg_hip_hf_solve(const imsrg_dop *H, const imsrg_hf_opts *opts, double *C_out,
                                  double *energies_out, double *F_out, double *ehf_out, int *iterations) {
  if (!H || !C_out) return IMSRG_E_INVALID;
  imsrg_ctx *ctx = H->ctx;
  const double tol = (opts && opts->tolerance > 0) ? opts->tolerance : 1e-8;
  const int maxiter = (opts && opts->max_iter > 0) ? opts->max_iter : 600;
  if (maxiter > 600)
    return fail(ctx, IMSRG_E_UNSUPPORTED,
                "Hartree-Fock beyond 600 iterations (the reference's randomised mixing) is not provided");
  if (H->herm != 1) return fail(ctx, IMSRG_E_UNSUPPORTED, "Hartree-Fock needs a Hermitian Hamiltonian");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const HostBasis &hb = ctx->hb;
  const BasisDev &B = ctx->db.h;
  HFState st;
  const int norb = st.norb = hb.norb;
  st.blk.assign(norb, -1);
  st.pos.assign(norb, 0);
  for (int o = 0; o < norb; ++o)
    if (hb.ob[o].front() == o) {
      for (size_t k = 0; k < hb.ob[o].size(); ++k) {
        st.blk[hb.ob[o][k]] = (int)st.blocks.size();
        st.pos[hb.ob[o][k]] = (int)k;
      }
      st.blocks.push_back(hb.ob[o]);
    }
  const int nb = (int)st.blocks.size();
  std::vector<int4> quads;
  st.wbase.assign((size_t)nb * nb, 0);
  for (int bi = 0; bi < nb; ++bi)
    for (int ba = 0; ba < nb; ++ba) {
      st.wbase[(size_t)bi * nb + ba] = (long)quads.size();
      for (int i : st.blocks[bi])
        for (int j : st.blocks[bi])
          for (int a : st.blocks[ba])
            for (int b : st.blocks[ba]) quads.push_back(make_int4(i, a, j, b));
    }
  const int nq = (int)quads.size();
  st.W.assign(nq, 0.0);
  {  // monopoles on the device
    int4 *dq = nullptr;
    double *dw = nullptr;
    HIPCHK(ctx, hipMalloc(&dq, sizeof(int4) * std::max(nq, 1)));
    hipError_t e = hipMalloc(&dw, sizeof(double) * std::max(nq, 1));
    if (e == hipSuccess) e = hipMemcpyAsync(dq, quads.data(), sizeof(int4) * nq, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = launch_hf_mono(ctx->db.d, H->buf, dq, nq, dw, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(st.W.data(), dw, sizeof(double) * nq, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dq);
    if (dw) (void)hipFree(dw);
    if (e != hipSuccess) return fail(ctx, IMSRG_E_HIP, std::string("Hartree-Fock monopoles: ") + hipGetErrorString(e));
  }
  const size_t n2 = (size_t)norb * norb;
  st.KE.assign(n2, 0.0);
  HIPCHK(ctx, hipMemcpyAsync(st.KE.data(), H->buf + B.one_off, sizeof(double) * n2, hipMemcpyDeviceToHost,
                             ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  st.j2 = hb.j2;
  st.holes = hb.holes;
  for (int h : hb.holes) st.hole_occ.push_back(hb.occ[h]);
  st.C.assign(n2, 0.0);
  for (int i = 0; i < norb; ++i) st.m(st.C, i, i) = 1.0;
  st.rho.assign(n2, 0.0);
  st.Vij.assign(n2, 0.0);
  st.F.assign(n2, 0.0);
  st.energies.assign(norb, 0.0);
  for (int i = 0; i < norb; ++i) st.energies[i] = st.m(st.KE, i, i);
  // Solve, src/HartreeFock.cc:55-167: density mixing, the DIIS update from iteration 101
  // while its last error matrix exceeds 0.01, the mixing factors 0.7 / 0.5 from 500
  st.density();
  st.fock();
  double dmix = 0.2, fmix = 0.0;  // density_mixing_factor, field_mixing_factor
  std::deque<std::vector<double>> derr, drho;
  int it = 0;
  bool converged = false;
  for (it = 0; it < maxiter; ++it) {
    const std::vector<double> prev = st.energies;
    st.diagonalize();
    if (it == 500) {
      dmix = 0.7;
      fmix = 0.5;
    }
    if (it > 100 && (derr.empty() || frob(derr.back()) > 0.01)) {
      st.diis(derr, drho);
    } else {
      const std::vector<double> last_rho = st.rho;
      st.density();
      for (size_t e = 0; e < n2; ++e) st.rho[e] = (1.0 - dmix) * st.rho[e] + dmix * last_rho[e];
    }
    const std::vector<double> last_F = st.F;
    st.fock();
    for (size_t e = 0; e < n2; ++e) st.F[e] = (1.0 - fmix) * st.F[e] + fmix * last_F[e];
    st.ehf();  // CheckConvergence, :834-844
    double d2 = 0.0;
    for (int i = 0; i < norb; ++i) d2 += (st.energies[i] - prev[i]) * (st.energies[i] - prev[i]);
    if (std::sqrt(d2) / norb < tol) {
      converged = true;
      break;
    }
  }
  const double e = st.ehf();
  std::memcpy(C_out, st.C.data(), sizeof(double) * n2);
  if (energies_out) std::memcpy(energies_out, st.energies.data(), sizeof(double) * norb);
  if (F_out) std::memcpy(F_out, st.F.data(), sizeof(double) * n2);
  if (ehf_out) {
    ehf_out[0] = e;
    ehf_out[1] = st.e1;
    ehf_out[2] = st.e2;
  }
  if (iterations) *iterations = it;
  if (!converged) return fail(ctx, IMSRG_E_NOT_CONVERGED, "Hartree-Fock did not converge within max_iter iterations");
  return IMSRG_OK;
}

// HartreeFock::GetNormalOrderedH (src/HartreeFock.cc:1079-1172) for a two-body
// Hamiltonian: E0 = EHF, OneBody = C^T F C, and the two-body part D^T V D as in
// TransformToHFBasis; the result is Hermitian.
extern "C" int imsrg_hip_hf_normal_ordered_h(imsrg_dop *out, const imsrg_dop *H, const double *C, const double *F,
                                             double ehf) {
  if (!out || !H || !C || !F) return IMSRG_E_INVALID;
  if (!same_ctx(out, H)) return IMSRG_E_INVALID;
  imsrg_ctx *ctx = out->ctx;
  if (int rc = imsrg_hip_op_transform_basis(out, H, C)) return rc;
  const int norb = ctx->hb.norb;
  const size_t n2 = (size_t)norb * norb;
  std::vector<double> T(n2, 0.0), f(n2, 0.0);
  for (int b = 0; b < norb; ++b)  // T = F C
    for (int k = 0; k < norb; ++k) {
      const double c = C[(size_t)b * norb + k];
      if (c == 0.0) continue;
      for (int a = 0; a < norb; ++a) T[(size_t)b * norb + a] += F[(size_t)k * norb + a] * c;
    }
  for (int b = 0; b < norb; ++b)  // f = C^T T
    for (int a = 0; a < norb; ++a) {
      double acc = 0.0;
      for (int k = 0; k < norb; ++k) acc += C[(size_t)a * norb + k] * T[(size_t)b * norb + k];
      f[(size_t)b * norb + a] = acc;
    }
  HIPCHK(ctx, hipMemcpyAsync(out->buf + ctx->db.h.one_off, f.data(), sizeof(double) * n2, hipMemcpyHostToDevice,
                             ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(out->buf, &ehf, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  out->herm = 1;
  return IMSRG_OK;
}

// ---------------------------------------------------------------------------
static int comm_opts_ok(imsrg_ctx *ctx, const imsrg_comm_opts *o, unsigned *mask, int *prec) {
  *mask = IMSRG_ALL_TERMS;
  *prec = IMSRG_PREC_FP64;
  if (o) {
    if (o->use_imsrg3) return fail(ctx, IMSRG_E_UNSUPPORTED, "IMSRG(3) terms are not provided");
    if (o->term_mask) *mask = o->term_mask & IMSRG_ALL_TERMS;
    if (o->precision != IMSRG_PREC_FP64 && o->precision != IMSRG_PREC_MIXED && o->precision != IMSRG_PREC_SPLIT)
      return fail(ctx, IMSRG_E_INVALID, "unknown precision mode");
    *prec = o->precision;
  }
  return IMSRG_OK;
}

static long r8(long n) { return (n + 7) / 8 * 8; }  // 64-byte aligned segment length in doubles

static int herm_rule(int hx, int hy) { return (hx == hy) ? -1 : 1; }  // src/Commutator.cc:292-297

// Launch sequence of one commutator (or a batch when pa.tab is set).
// items / cnt: [0] pp/hh 128-tiles, [1] ph tiles, [2] comm221 one-body items,
// [3] Pandya chunks, [4] inverse-Pandya items,
// [5] one-body pairs, [6] tile pairs of the finishing kernel,
// [7] pp/hh 64-tiles, [8] bf16x3 panel packing, [9] comm122 W column bins.
// Under mixed precision the
// context's tuning picks the kernel of each GEMM family.
// Sums the phase durations of the oldest recorded event set (waits for it).
static int harvest_oldest(imsrg_ctx *ctx) {
  auto &set = ctx->ring[ctx->ring_head];
  HIPCHK(ctx, hipEventSynchronize(set[8]));
  // sequence order of the phases: zero-body, one-body, pp/hh, 122 coefficients,
  // Pandya, ph GEMM, inverse Pandya, last kernel; reported in the header's order
  const int map[8] = {0, 1, 3, 2, 4, 5, 6, 7};
  for (int i = 0; i < 8; ++i) {
    float ms = 0.f;
    if (i == 2 && ctx->ring_side[ctx->ring_head])
      HIPCHK(ctx, hipEventElapsedTime(&ms, set[9], set[10]));
    else if (i == 3 && ctx->ring_side[ctx->ring_head])  // comm122 on the side stream
      HIPCHK(ctx, hipEventElapsedTime(&ms, set[10], set[11]));
    else
      HIPCHK(ctx, hipEventElapsedTime(&ms, set[i], set[i + 1]));
    ctx->phase_sum[map[i]] += ms * 1000.0;
  }
  ++ctx->phase_n;
  ctx->ring_head = (ctx->ring_head + 1) % imsrg_ctx::kRing;
  --ctx->ring_count;
  return IMSRG_OK;
}

// GEMM kernel family of a phase: under mixed precision the tuning picks FP64
// (0), FP32 MFMA (1) or bf16x3 MFMA (2); other precisions keep theirs
static int gemm_family(int prec, int flag) {
  return prec != IMSRG_PREC_MIXED ? prec : flag == 0 ? 0 : flag == 1 ? 1 : kGemmB3;
}
static bool uses_panels(const imsrg_ctx *ctx, int prec) {
  return gemm_family(prec, ctx->db.tune.mixed_pphh) == kGemmB3 && !ctx->db.wl.pack.empty();
}

static int run_commutator(imsrg_ctx *ctx, const PairArg &pa, int npairs, const int4 *const *items, const int *cnt,
                          unsigned mask, int prec, int maxn, int maxnch, long maxn122, int maxjmax, int maxnorb,
                          int maxn1pairs, int maxncls, int maxbin, int maxent) {
  const Tuning &tu = ctx->db.tune;
  // split products take both GEMM families whatever the mixed-precision choices
  const int prec_pphh = gemm_family(prec, tu.mixed_pphh);
  const int prec_ph = gemm_family(prec, tu.mixed_ph);
  hipStream_t s = ctx->stream;
  // With overlap, the pp/hh GEMM and then comm122's coefficients and W run on
  // the side stream next to the zero-body, one-body, Pandya, ph GEMM and
  // inverse Pandya kernels (none of them touches the upper triangle of Z's
  // two-body blocks); the main stream waits for them before the last kernel,
  // which folds, applies W and mirrors.
  const bool overlap = tu.overlap_pphh && (mask & MPPHH);
  // overlap_pphh 1: fork before the first kernel; 2: fork after the ph GEMM, so
  // that the pp/hh GEMM shares the GPU with the inverse Pandya gather only
  const bool fork_late = overlap && tu.overlap_pphh == 2;
  // profiling: this commutator's phase events go to the next free set of the ring
  std::array<hipEvent_t, imsrg_ctx::kMarks> *set = nullptr;
  if (ctx->profiling) {
    if (ctx->ring_count == imsrg_ctx::kRing)
      if (int rc = harvest_oldest(ctx)) return rc;
    const int slot = (ctx->ring_head + ctx->ring_count) % imsrg_ctx::kRing;
    set = &ctx->ring[slot];
    ctx->ring_side[slot] = overlap;
  }
  auto mark = [&](int i) -> hipError_t { return set ? hipEventRecord((*set)[i], s) : hipSuccess; };
  auto fork_side = [&]() -> int {
    HIPCHK(ctx, hipEventRecord(ctx->fork, s));
    HIPCHK(ctx, hipStreamWaitEvent(ctx->side, ctx->fork, 0));
    if (set) HIPCHK(ctx, hipEventRecord((*set)[9], ctx->side));
    HIPCHK(ctx, launch_pphh(pa, items[0], cnt[0], items[7], cnt[7], items[8], cnt[8], prec_pphh, maxn, ctx->side));
    if (set) HIPCHK(ctx, hipEventRecord((*set)[10], ctx->side));
    if (mask & M122) {
      HIPCHK(ctx, launch_c122_coef(pa, npairs, maxn122, ctx->side));
      HIPCHK(ctx, launch_c122_w(pa, items[9], cnt[9], maxbin, maxent, ctx->side));
    }
    if (set) HIPCHK(ctx, hipEventRecord((*set)[11], ctx->side));
    HIPCHK(ctx, hipEventRecord(ctx->join, ctx->side));
    return IMSRG_OK;
  };
  if (overlap && !fork_late)
    if (int rc = fork_side()) return rc;
  HIPCHK(ctx, mark(0));
  HIPCHK(ctx, launch_zerobody(pa, npairs, maxnch, s));
  HIPCHK(ctx, mark(1));
  HIPCHK(ctx, launch_onebody(pa, items[5], cnt[5], (mask & M221) ? items[2] : nullptr, (mask & M221) ? cnt[2] : 0,
                             npairs, maxn1pairs, s));
  HIPCHK(ctx, mark(2));
  if (overlap) {
  } else if (mask & MPPHH) {
    HIPCHK(ctx, launch_pphh(pa, items[0], cnt[0], items[7], cnt[7], items[8], cnt[8], prec_pphh, maxn, s));
  } else {
    HIPCHK(ctx, launch_zero_two(pa, items[6], cnt[6], s));
  }
  HIPCHK(ctx, mark(3));
  if ((mask & M122) && !overlap) {
    HIPCHK(ctx, launch_c122_coef(pa, npairs, maxn122, s));
    HIPCHK(ctx, launch_c122_w(pa, items[9], cnt[9], maxbin, maxent, s));
  }
  HIPCHK(ctx, mark(4));
  if (mask & MPH) HIPCHK(ctx, launch_pandya(pa, items[3], cnt[3], maxnch, s));
  HIPCHK(ctx, mark(5));
  if (mask & MPH) HIPCHK(ctx, launch_ph_gemm(pa, items[1], cnt[1], prec_ph, tu.ph_big != 0, s));
  if (fork_late)
    if (int rc = fork_side()) return rc;
  HIPCHK(ctx, mark(6));
  if (mask & MPH)
    HIPCHK(ctx, launch_invpandya(pa, items[4], cnt[4], maxjmax, maxnorb, maxnch, maxncls, prec_ph == 1, s));
  if (overlap) HIPCHK(ctx, hipStreamWaitEvent(s, ctx->join, 0));
  HIPCHK(ctx, mark(7));
  HIPCHK(ctx, launch_finish(pa, items[6], cnt[6], (mask & MPH) != 0, (mask & M122) != 0, s));
  HIPCHK(ctx, mark(8));
  if (set) ++ctx->ring_count;
  return IMSRG_OK;
}

// Per-call workspaces of one commutator inside one buffer of `doubles` doubles
// (the layout of the batched path's per-pair workspaces).
struct Workspace {
  double *Xt, *Yt, *Zbar, *scratch;
  double2 *c122;
  double *obpart, *diag, *c122w, *panel;
};
// panels: room for the bf16x3 operand panels (last region, used by the mixed
// pp/hh GEMM only)
static long workspace_doubles(const BasisDev &B, bool panels) {
  return 2 * r8(B.xsize + 8) + r8(B.zsize) + r8(B.nch + 8) + 2 * r8(B.n122) + r8(B.nobpart) + r8(B.ndiag) +
         r8(B.tbsize + 8) + (panels ? r8(2 * B.pk_size) : 0) + 64;
}
static Workspace carve_workspace(const BasisDev &B, double *base) {
  Workspace w;
  // 8 doubles of slack after X and Y bar: the GEMM tile loader may read one element past a column
  w.Xt = base;
  w.Yt = w.Xt + r8(B.xsize + 8);
  w.Zbar = w.Yt + r8(B.xsize + 8);
  w.scratch = w.Zbar + r8(B.zsize);
  w.c122 = (double2 *)(w.scratch + r8(B.nch + 8));
  w.obpart = (double *)(w.c122) + 2 * r8(B.n122);
  w.diag = w.obpart + r8(B.nobpart);
  w.c122w = w.diag + r8(B.ndiag);
  w.panel = w.c122w + r8(B.tbsize + 8);
  return w;
}

static int commutator_dev(imsrg_ctx *ctx, const imsrg_dop *X, const imsrg_dop *Y, imsrg_dop *Z, unsigned mask,
                          int prec, const Workspace *ws = nullptr) {
  if (X->herm != 1 && X->herm != -1) return fail(ctx, IMSRG_E_UNSUPPORTED, "X must be Hermitian or anti-Hermitian");
  if (Y->herm != 1 && Y->herm != -1) return fail(ctx, IMSRG_E_UNSUPPORTED, "Y must be Hermitian or anti-Hermitian");
  if (Z == X || Z == Y) return fail(ctx, IMSRG_E_INVALID, "Z must not alias X or Y");
  const DeviceBasis &db = ctx->db;
  if (!ws && !ctx->panel && uses_panels(ctx, prec)) {
    const size_t bytes = 16 * std::max<size_t>(db.h.pk_size, 1);
    if (hipMalloc(&ctx->panel, bytes) != hipSuccess) {
      ctx->panel = nullptr;
      return fail(ctx, IMSRG_E_NOMEM, "hipMalloc of the bf16x3 operand panels failed");
    }
  }
  PairArg pa;
  pa.tab = nullptr;
  pa.p0.basis = db.d;
  pa.p0.X = X->buf;
  pa.p0.Y = Y->buf;
  pa.p0.Z = Z->buf;
  pa.p0.Xt = ws ? ws->Xt : ctx->Xt;
  pa.p0.Yt = ws ? ws->Yt : ctx->Yt;
  pa.p0.Zbar = ws ? ws->Zbar : ctx->Zbar;
  pa.p0.scratch = ws ? ws->scratch : ctx->scratch;
  pa.p0.c122 = ws ? ws->c122 : ctx->c122;
  pa.p0.obpart = ws ? ws->obpart : ctx->obpart;
  pa.p0.diag = ws ? ws->diag : ctx->diag;
  pa.p0.panel = ws ? ws->panel : ctx->panel;
  pa.p0.c122w = ws ? ws->c122w : ctx->c122w;
  pa.p0.hX = X->herm;
  pa.p0.hY = Y->herm;
  pa.p0.hZ = herm_rule(X->herm, Y->herm);
  pa.p0.mask = mask;
  Z->herm = pa.p0.hZ;
  const int4 *items[10] = {db.d_pphh, db.d_ph,     db.d_ob221,  db.d_pandya, db.d_invp,
                           db.d_one,  db.d_mirror, db.d_pphh64, db.d_pack,   db.d_c122w};
  int cnt[10] = {(int)db.wl.pphh.size(),   (int)db.wl.ph.size(),   (int)db.wl.ob221.size(),
                 (int)db.wl.pandya.size(), (int)db.wl.invp.size(),  (int)db.wl.one.size(),
                 (int)db.wl.mirror.size(), (int)db.wl.pphh64.size(), (int)db.wl.pack.size(),
                 (int)db.wl.c122w.size()};
  if (!ctx->use_graphs || ctx->profiling || ws)
    return run_commutator(ctx, pa, 1, items, cnt, mask, prec, db.h.maxn, db.h.nch, db.h.n122, db.h.jmax,
                          db.h.norb, db.h.n1pairs, db.h.ncls, db.h.c122_maxbin, db.h.c122_maxent);
  // replay a captured launch sequence for this operand set, capturing it on first use
  for (const GraphEntry &g : ctx->graphs)
    if (g.x == X->buf && g.y == Y->buf && g.z == Z->buf && g.hx == X->herm && g.hy == Y->herm && g.mask == mask &&
        g.prec == prec) {
      HIPCHK(ctx, hipGraphLaunch(g.exec, ctx->stream));
      return IMSRG_OK;
    }
  HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int rc = run_commutator(ctx, pa, 1, items, cnt, mask, prec, db.h.maxn, db.h.nch, db.h.n122, db.h.jmax,
                                db.h.norb, db.h.n1pairs, db.h.ncls, db.h.c122_maxbin, db.h.c122_maxent);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);  // always leave capture mode
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  HIPCHK(ctx, ec);
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIPCHK(ctx, ei);
  if (ctx->graphs.size() >= kMaxGraphs) {
    (void)hipGraphExecDestroy(ctx->graphs.front().exec);
    ctx->graphs.erase(ctx->graphs.begin());
  }
  ctx->graphs.push_back(GraphEntry{X->buf, Y->buf, Z->buf, X->herm, Y->herm, mask, prec, exec});
  HIPCHK(ctx, hipGraphLaunch(exec, ctx->stream));
  return IMSRG_OK;
}

int imsrg::ctx_device(const imsrg_ctx *ctx) { return ctx ? ctx->device : -1; }
int imsrg::ctx_fail(imsrg_ctx *ctx, int code, const char *msg) { return fail(ctx, code, msg); }

// The allocation-free commutator: the context's own workspaces are replaced by
// ones carved from ws.  Cross-coupled channels with no ph kets are never
// written by the ph GEMM and must read as 0, so that part (all of Zbar) is
// cleared by a fill kernel on every call: ws keeps the previous call's
// contents, possibly of another basis.  Every other workspace is written
// before it is read.
int imsrg::commutator_in_workspace(imsrg_ctx *ctx, const imsrg_dop *X, const imsrg_dop *Y, imsrg_dop *Z,
                                   const imsrg_comm_opts *opts, double *ws, long cap) {
  if (!X || !Y || !Z) return IMSRG_E_INVALID;
  if (X->ctx != ctx || Y->ctx != ctx || Z->ctx != ctx) return fail(ctx, IMSRG_E_INVALID, "operator from another context");
  unsigned mask;
  int prec;
  if (int rc = comm_opts_ok(ctx, opts, &mask, &prec)) return rc;
  const BasisDev &B = ctx->db.h;
  const long need = workspace_doubles(B, uses_panels(ctx, prec));
  if (need > cap) {
    char m[160];
    std::snprintf(m, sizeof(m), "basis needs %ld bytes of workspace; the static workspace holds %ld (emax <= 6 fits)",
                  8 * need, 8 * cap);
    return fail(ctx, IMSRG_E_INVALID, m);
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const Workspace w = carve_workspace(B, ws);
  HIPCHK(ctx, launch_fill(w.Zbar, 0.0, std::max<long>(B.zsize, 1), ctx->stream));
  return commutator_dev(ctx, X, Y, Z, mask, prec, &w);
}

extern "C" int imsrg_hip_set_graphs(imsrg_ctx *ctx, int on) {
  if (!ctx) return IMSRG_E_INVALID;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (!on) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (auto &g : ctx->graphs) (void)hipGraphExecDestroy(g.exec);
    ctx->graphs.clear();
  }
  ctx->use_graphs = on != 0;
  return IMSRG_OK;
}

extern "C" int imsrg_hip_commutator_ss_dev(imsrg_ctx *ctx, const imsrg_dop *X, const imsrg_dop *Y, imsrg_dop *Z,
                                           const imsrg_comm_opts *opts) {
  if (!ctx || !X || !Y || !Z) return IMSRG_E_INVALID;
  if (X->ctx != ctx || Y->ctx != ctx || Z->ctx != ctx) return fail(ctx, IMSRG_E_INVALID, "operator from another context");
  unsigned mask;
  int prec;
  int rc = comm_opts_ok(ctx, opts, &mask, &prec);
  if (rc) return rc;
  return commutator_dev(ctx, X, Y, Z, mask, prec);
}

extern "C" int imsrg_hip_commutator_ss_dev_async(imsrg_ctx *ctx, const imsrg_dop *X, const imsrg_dop *Y,
                                                 imsrg_dop *Z, const imsrg_comm_opts *opts, void *stream) {
  if (!ctx || !X || !Y || !Z) return IMSRG_E_INVALID;
  if (X->ctx != ctx || Y->ctx != ctx || Z->ctx != ctx) return fail(ctx, IMSRG_E_INVALID, "operator from another context");
  unsigned mask;
  int prec;
  int rc = comm_opts_ok(ctx, opts, &mask, &prec);
  if (rc) return rc;
  hipStream_t cs = static_cast<hipStream_t>(stream);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipEventRecord(ctx->async_in, cs));
  HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, ctx->async_in, 0));
  if ((rc = commutator_dev(ctx, X, Y, Z, mask, prec))) return rc;
  HIPCHK(ctx, hipEventRecord(ctx->async_out, ctx->stream));
  HIPCHK(ctx, hipStreamWaitEvent(cs, ctx->async_out, 0));
  return IMSRG_OK;
}

static int ensure_ops(imsrg_ctx *ctx, imsrg_dop **arr, int n) {
  for (int i = 0; i < n; ++i)
    if (!arr[i]) {
      int rc = imsrg_hip_op_create(ctx, 1, &arr[i]);
      if (rc) return rc;
    }
  return IMSRG_OK;
}

extern "C" int imsrg_hip_commutator_ss(imsrg_ctx *ctx, const imsrg_op_view *X, const imsrg_op_view *Y,
                                       imsrg_op_view *Z, const imsrg_comm_opts *opts) {
  if (!ctx || !X || !Y || !Z) return IMSRG_E_INVALID;
  unsigned mask;
  int prec;
  int rc = comm_opts_ok(ctx, opts, &mask, &prec);
  if (rc) return rc;
  if ((rc = view_ok(ctx, X, true)) || (rc = view_ok(ctx, Y, true))) return rc;
  if (!Z->one_body) return fail(ctx, IMSRG_E_INVALID, "null Z one-body pointer");
  if ((rc = ensure_ops(ctx, ctx->stage, 3))) return rc;
  if ((rc = imsrg_hip_op_upload(ctx->stage[0], X))) return rc;
  if ((rc = imsrg_hip_op_upload(ctx->stage[1], Y))) return rc;
  if ((rc = commutator_dev(ctx, ctx->stage[0], ctx->stage[1], ctx->stage[2], mask, prec))) return rc;
  return imsrg_hip_op_download(ctx->stage[2], Z);
}

// ---------------------------------------------------------------------------
// BCH on device operators
static int bch_defaults(imsrg_ctx *ctx, const imsrg_bch_opts *o, imsrg_bch_opts *d) {
  d->mode = 0;
  d->transform_thr = 1e-9;
  d->product_thr = 1e-4;
  d->max_terms = 40;
  d->precision = IMSRG_PREC_FP64;
  if (o) {
    if (o->mode != 0 && o->mode != 1) return fail(ctx, IMSRG_E_INVALID, "unknown BCH truncation mode (0 or 1)");
    d->mode = o->mode;
    if (o->transform_thr > 0) d->transform_thr = o->transform_thr;
    if (o->product_thr > 0) d->product_thr = o->product_thr;
    if (o->max_terms > 0) d->max_terms = o->max_terms;
    if (o->precision != IMSRG_PREC_FP64 && o->precision != IMSRG_PREC_MIXED && o->precision != IMSRG_PREC_SPLIT)
      return fail(ctx, IMSRG_E_INVALID, "unknown precision mode");
    d->precision = o->precision;
  }
  return IMSRG_OK;
}

// Standard_BCH_Transform, src/BCH.cc:87-201
static int bch_transform(imsrg_ctx *ctx, const imsrg_dop *Op, const imsrg_dop *Omega, imsrg_dop *Out,
                         const imsrg_bch_opts &o, int *ncomm) {
  int rc;
  if ((rc = ensure_ops(ctx, ctx->tmp, 2))) return rc;
  imsrg_dop *A = ctx->tmp[0], *Bn = ctx->tmp[1];
  double nx, ny;
  if ((rc = op_norm(const_cast<imsrg_dop *>(Op), nullptr, nullptr, &nx))) return rc;
  if ((rc = op_norm(const_cast<imsrg_dop *>(Omega), nullptr, nullptr, &ny))) return rc;
  if ((rc = imsrg_hip_op_copy(Out, Op))) return rc;
  int nc = 0;
  bool converged = true;
  if (nx > o.transform_thr) {
    converged = false;
    if ((rc = imsrg_hip_op_copy(A, Op))) return rc;
    imsrg_dop *nested = A, *next = Bn;
    double fact = 1.0;
    double eps = nx * std::exp(-2 * ny) * o.transform_thr / (2 * ny);
    for (int i = 1; i <= o.max_terms; ++i) {
      if ((rc = commutator_dev(ctx, Omega, nested, next, IMSRG_ALL_TERMS, o.precision))) return rc;
      nc++;
      std::swap(nested, next);
      fact /= i;
      eps *= i + 1;
      double nn;  // Out += fact * nested, and the norm of nested, in one pass
      if ((rc = op_norm(nested, nullptr, nullptr, &nn, Out, fact))) return rc;
      if ((o.mode == 0 && nn < eps) || (o.mode == 1 && nn < o.transform_thr)) {
        converged = true;
        break;
      }
    }
  }
  if (ncomm) *ncomm = nc;
  Out->herm = Op->herm;
  if (!converged) return fail(ctx, IMSRG_E_NOT_CONVERGED, "BCH_Transform did not converge within max_terms");
  return IMSRG_OK;
}

// BCH_Product, src/BCH.cc:267-350
static int bch_product(imsrg_ctx *ctx, const imsrg_dop *X, const imsrg_dop *Y, imsrg_dop *Out,
                       const imsrg_bch_opts &o, int *ncomm) {
  static const double bern[9] = {1.0, -0.5, 1. / 6, 0.0, -1. / 30, 0.0, 1. / 42, 0, -1. / 30};
  static const double fact[9] = {1.0, 1.0, 2.0, 6.0, 24., 120., 720., 5040., 40320.};
  int rc;
  if ((rc = ensure_ops(ctx, ctx->tmp, 3))) return rc;
  if (Out == X || Out == Y) return fail(ctx, IMSRG_E_INVALID, "Out must not alias X or Y");
  double nx, ny;
  if ((rc = op_norm(const_cast<imsrg_dop *>(X), nullptr, nullptr, &nx))) return rc;
  if ((rc = op_norm(const_cast<imsrg_dop *>(Y), nullptr, nullptr, &ny))) return rc;
  if ((rc = imsrg_hip_op_copy(Out, X))) return rc;
  if ((rc = imsrg_hip_op_axpy(Out, 1.0, Y))) return rc;
  Out->herm = X->herm;
  imsrg_dop *nested = ctx->tmp[0], *next = ctx->tmp[1];
  int nc = 0;
  if ((rc = commutator_dev(ctx, Y, X, nested, IMSRG_ALL_TERMS, o.precision))) return rc;
  nc++;
  double nxy;
  if ((rc = op_norm(nested, nullptr, nullptr, &nxy))) return rc;
  if (nxy * nx > o.product_thr) {
    if ((rc = commutator_dev(ctx, nested, X, next, IMSRG_ALL_TERMS, o.precision))) return rc;
    nc++;
    if ((rc = imsrg_hip_op_axpy(Out, 1. / 12, next))) return rc;
  }
  int k = 1;
  while (nxy > o.product_thr) {
    if (k < 2 || k % 2 == 0)
      if ((rc = imsrg_hip_op_axpy(Out, bern[k] / fact[k], nested))) return rc;
    k++;
    if (k >= 9) break;
    if (2 * ny * nxy < o.product_thr) break;
    if ((rc = commutator_dev(ctx, Y, nested, next, IMSRG_ALL_TERMS, o.precision))) return rc;
    nc++;
    std::swap(nested, next);
    if ((rc = op_norm(nested, nullptr, nullptr, &nxy))) return rc;
  }
  if (ncomm) *ncomm = nc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return IMSRG_OK;
}

extern "C" int imsrg_hip_bch_transform_dev(imsrg_ctx *ctx, const imsrg_dop *Op, const imsrg_dop *Omega,
                                           imsrg_dop *Out, const imsrg_bch_opts *opts, int *ncomm) {
  if (!ctx || !Op || !Omega || !Out) return IMSRG_E_INVALID;
  if (Op->ctx != ctx || Omega->ctx != ctx || Out->ctx != ctx) return fail(ctx, IMSRG_E_INVALID, "foreign operator");
  if (Out == Omega) return fail(ctx, IMSRG_E_INVALID, "Out must not alias Omega");
  imsrg_bch_opts o;
  if (int e = bch_defaults(ctx, opts, &o)) return e;
  return bch_transform(ctx, Op, Omega, Out, o, ncomm);
}
extern "C" int imsrg_hip_bch_product_dev(imsrg_ctx *ctx, const imsrg_dop *X, const imsrg_dop *Y, imsrg_dop *Out,
                                         const imsrg_bch_opts *opts, int *ncomm) {
  if (!ctx || !X || !Y || !Out) return IMSRG_E_INVALID;
  if (X->ctx != ctx || Y->ctx != ctx || Out->ctx != ctx) return fail(ctx, IMSRG_E_INVALID, "foreign operator");
  imsrg_bch_opts o;
  if (int e = bch_defaults(ctx, opts, &o)) return e;
  return bch_product(ctx, X, Y, Out, o, ncomm);
}

extern "C" int imsrg_hip_bch_transform(imsrg_ctx *ctx, const imsrg_op_view *Op, const imsrg_op_view *Omega,
                                       imsrg_op_view *Out, const imsrg_bch_opts *opts, int *ncomm) {
  if (!ctx || !Op || !Omega || !Out) return IMSRG_E_INVALID;
  int rc;
  if ((rc = view_ok(ctx, Op, true)) || (rc = view_ok(ctx, Omega, true))) return rc;
  if ((rc = ensure_ops(ctx, ctx->stage, 3))) return rc;
  if ((rc = imsrg_hip_op_upload(ctx->stage[0], Op))) return rc;
  if ((rc = imsrg_hip_op_upload(ctx->stage[1], Omega))) return rc;
  imsrg_bch_opts o;
  if (int e = bch_defaults(ctx, opts, &o)) return e;
  rc = bch_transform(ctx, ctx->stage[0], ctx->stage[1], ctx->stage[2], o, ncomm);
  if (rc && rc != IMSRG_E_NOT_CONVERGED) return rc;
  int rc2 = imsrg_hip_op_download(ctx->stage[2], Out);
  return rc2 ? rc2 : rc;
}
extern "C" int imsrg_hip_bch_product(imsrg_ctx *ctx, const imsrg_op_view *X, const imsrg_op_view *Y,
                                     imsrg_op_view *Out, const imsrg_bch_opts *opts, int *ncomm) {
  if (!ctx || !X || !Y || !Out) return IMSRG_E_INVALID;
  int rc;
  if ((rc = view_ok(ctx, X, true)) || (rc = view_ok(ctx, Y, true))) return rc;
  if ((rc = ensure_ops(ctx, ctx->stage, 3))) return rc;
  if ((rc = imsrg_hip_op_upload(ctx->stage[0], X))) return rc;
  if ((rc = imsrg_hip_op_upload(ctx->stage[1], Y))) return rc;
  imsrg_bch_opts o;
  if (int e = bch_defaults(ctx, opts, &o)) return e;
  if ((rc = bch_product(ctx, ctx->stage[0], ctx->stage[1], ctx->stage[2], o, ncomm))) return rc;
  return imsrg_hip_op_download(ctx->stage[2], Out);
}

// ---------------------------------------------------------------------------
extern "C" int imsrg_hip_white_generator_dev(imsrg_ctx *ctx, const imsrg_dop *H, imsrg_dop *Eta) {
  if (!ctx || !H || !Eta || H->ctx != ctx || Eta->ctx != ctx || H == Eta) return IMSRG_E_INVALID;
  HIPCHK(ctx, launch_fill(Eta->buf, 0.0, ctx->db.h.opsize, ctx->stream));
  HIPCHK(ctx, launch_white(ctx->db.d, ctx->db.h, H->buf, Eta->buf, ctx->db.d_gen, (int)ctx->db.wl.gen.size(),
                           ctx->stream));
  Eta->herm = -1;
  return IMSRG_OK;
}

// ---------------------------------------------------------------------------
// Operator binary format: Operator::WriteBinary / ReadBinary
// (src/Operator.cc:341-388) with TwoBodyME::WriteBinary / ReadBinary
// (src/TwoBodyME.cc:812-838).  Fields are written one after another, with no
// padding: int rank_J, rank_T, parity, particle_rank, legs, E2max, E3max;
// bool hermitian, antihermitian; int nChannels; double ZeroBody; the norb x
// norb one-body matrix (column-major); then, for legs > 3, size_t nChannels,
// bool hermitian, antihermitian, int rank_J, rank_T, parity and every
// (ch, ch) block of TwoBodyME::MatEl in map order (ascending ch, column-major).
namespace {
struct BinHeader {
  int rank_J = 0, rank_T = 0, parity = 0, particle_rank = 2, legs = 4, E2max = 0, E3max = 0;
  bool herm = false, antiherm = false;
  int nch = 0;
};
constexpr long kBinOpHead = 7 * 4 + 2 + 4;       // Operator fields before ZeroBody
constexpr long kBinTwoHead = 8 + 2 + 3 * 4;      // TwoBodyME fields before the blocks
template <class T>
void put(std::vector<char> &b, long &o, const T &v) {
  std::memcpy(b.data() + o, &v, sizeof(T));
  o += sizeof(T);
}
template <class T>
T get(const char *b, long &o) {
  T v;
  std::memcpy(&v, b + o, sizeof(T));
  o += sizeof(T);
  return v;
}
}  // namespace

static long binary_size(const imsrg_ctx *ctx) {
  const BasisDev &B = ctx->db.h;
  return kBinOpHead + 8 + 8L * B.norb * B.norb + kBinTwoHead + 8L * B.tbsize;
}

extern "C" int imsrg_hip_op_binary_size(const imsrg_dop *op, long *bytes) {
  if (!op || !bytes) return IMSRG_E_INVALID;
  *bytes = binary_size(op->ctx);
  return IMSRG_OK;
}

extern "C" int imsrg_hip_op_to_binary(imsrg_dop *op, void *buf, long size) {
  if (!op || !buf) return IMSRG_E_INVALID;
  imsrg_ctx *ctx = op->ctx;
  const BasisDev &B = ctx->db.h;
  const long need = binary_size(ctx);
  if (size < need) return fail(ctx, IMSRG_E_INVALID, "binary buffer too small");
  std::vector<char> b(kBinOpHead + kBinTwoHead + 8);
  long o = 0;
  const int emax = ctx->hb.emax;
  for (int v : {0, 0, 0, 2, 4, 2 * emax, std::min(14, 3 * emax)}) put(b, o, v);  // rank_J .. E3max
  put(b, o, (bool)(op->herm == 1));
  put(b, o, (bool)(op->herm == -1));
  put(b, o, (int)B.nch);
  char *out = static_cast<char *>(buf);
  std::memcpy(out, b.data(), o);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipMemcpyAsync(out + o, op->buf, 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(out + o + 8, op->buf + B.one_off, 8L * B.norb * B.norb, hipMemcpyDeviceToHost,
                             ctx->stream));
  long t = o + 8 + 8L * B.norb * B.norb;
  long o2 = 0;
  put(b, o2, (size_t)B.nch);
  put(b, o2, (bool)(op->herm == 1));
  put(b, o2, (bool)(op->herm == -1));
  for (int v : {0, 0, 0}) put(b, o2, v);  // rank_J, rank_T, parity
  std::memcpy(out + t, b.data(), o2);
  HIPCHK(ctx, hipMemcpyAsync(out + t + o2, op->buf + B.two_off, 8L * B.tbsize, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return IMSRG_OK;
}

extern "C" int imsrg_hip_op_from_binary(imsrg_dop *op, const void *buf, long size) {
  if (!op || !buf) return IMSRG_E_INVALID;
  imsrg_ctx *ctx = op->ctx;
  const BasisDev &B = ctx->db.h;
  const char *in = static_cast<const char *>(buf);
  if (size < kBinOpHead + 8 + 8L * B.norb * B.norb) return fail(ctx, IMSRG_E_INVALID, "binary operator truncated");
  long o = 0;
  BinHeader h;
  h.rank_J = get<int>(in, o); h.rank_T = get<int>(in, o); h.parity = get<int>(in, o);
  h.particle_rank = get<int>(in, o); h.legs = get<int>(in, o); h.E2max = get<int>(in, o); h.E3max = get<int>(in, o);
  h.herm = get<bool>(in, o); h.antiherm = get<bool>(in, o);
  h.nch = get<int>(in, o);
  if (h.rank_J != 0 || h.rank_T != 0 || h.parity != 0)
    return fail(ctx, IMSRG_E_UNSUPPORTED, "binary operator is not a parity-conserving scalar");
  if (h.legs != 2 && h.legs != 4) return fail(ctx, IMSRG_E_UNSUPPORTED, "binary operator has a particle rank above 2");
  if (h.nch != B.nch) return fail(ctx, IMSRG_E_INVALID, "binary operator has another channel count");
  const int herm = h.herm ? 1 : (h.antiherm ? -1 : 0);
  const long n1 = 8L * B.norb * B.norb;
  const long two_at = o + 8 + n1;
  if (h.legs > 3) {
    if (size < two_at + kBinTwoHead + 8L * B.tbsize) return fail(ctx, IMSRG_E_INVALID, "binary operator truncated");
    long t = two_at;
    const size_t nch2 = get<size_t>(in, t);
    get<bool>(in, t);
    get<bool>(in, t);
    const int rj = get<int>(in, t), rt = get<int>(in, t), par = get<int>(in, t);
    if (nch2 != (size_t)B.nch || rj != 0 || rt != 0 || par != 0)
      return fail(ctx, IMSRG_E_INVALID, "binary two-body header does not match the basis");
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, launch_fill(op->buf, 0.0, B.opsize, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(op->buf, in + o, 8, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(op->buf + B.one_off, in + o + 8, n1, hipMemcpyHostToDevice, ctx->stream));
  if (h.legs > 3)
    HIPCHK(ctx, hipMemcpyAsync(op->buf + B.two_off, in + two_at + kBinTwoHead, 8L * B.tbsize, hipMemcpyHostToDevice,
                               ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  op->herm = herm;
  return IMSRG_OK;
}

extern "C" int imsrg_hip_op_write_binary(imsrg_dop *op, const char *path) {
  if (!op || !path) return IMSRG_E_INVALID;
  imsrg_ctx *ctx = op->ctx;
  std::vector<char> b(binary_size(ctx));
  if (int rc = imsrg_hip_op_to_binary(op, b.data(), (long)b.size())) return rc;
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(ctx, IMSRG_E_INVALID, std::string("cannot open ") + path + " for writing");
  const size_t w = std::fwrite(b.data(), 1, b.size(), f);
  const int c = std::fclose(f);
  if (w != b.size() || c != 0) return fail(ctx, IMSRG_E_INVALID, std::string("short write to ") + path);
  return IMSRG_OK;
}

extern "C" int imsrg_hip_op_read_binary(imsrg_dop *op, const char *path) {
  if (!op || !path) return IMSRG_E_INVALID;
  imsrg_ctx *ctx = op->ctx;
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(ctx, IMSRG_E_INVALID, std::string("cannot open ") + path);
  std::vector<char> b;
  char chunk[1 << 16];
  size_t r;
  while ((r = std::fread(chunk, 1, sizeof(chunk), f)) > 0) b.insert(b.end(), chunk, chunk + r);
  std::fclose(f);
  return imsrg_hip_op_from_binary(op, b.data(), (long)b.size());
}

// ---------------------------------------------------------------------------
// IMSRGSolver: the Magnus flow with its list of Omega segments
// (src/IMSRGSolver.cc:44-84 NewOmega, :236-321 Solve_magnus_euler,
// :928-1004 Transform / Transform_Partial, :939-951 InverseTransform).
struct imsrg_solver {
  imsrg_ctx *ctx;
  std::vector<imsrg_dop *> omega;  // finished segments held in HBM (nullptr when written to scratch)
  std::vector<std::string> files;  // segment i written to files[i] (scratch mode), else ""
  imsrg_dop *current = nullptr;    // the segment being accumulated (Omega.back())
  std::string scratch;             // SetScratchDir: finished segments go to files
  int written = 0;
};

static void solver_clear(imsrg_solver *sv) {
  for (imsrg_dop *o : sv->omega) imsrg_hip_op_destroy(o);
  for (const std::string &f : sv->files)
    if (!f.empty()) std::remove(f.c_str());
  sv->omega.clear();
  sv->files.clear();
  imsrg_hip_op_destroy(sv->current);
  sv->current = nullptr;
}

extern "C" int imsrg_hip_solver_create(imsrg_ctx *ctx, imsrg_solver **out) {
  if (!ctx || !out) return IMSRG_E_INVALID;
  *out = new imsrg_solver{ctx, {}, {}, nullptr, "", 0};
  return IMSRG_OK;
}
extern "C" int imsrg_hip_solver_destroy(imsrg_solver *sv) {
  if (!sv) return IMSRG_OK;
  solver_clear(sv);
  delete sv;
  return IMSRG_OK;
}
extern "C" int imsrg_hip_solver_set_scratch(imsrg_solver *sv, const char *dir) {
  if (!sv) return IMSRG_E_INVALID;
  sv->scratch = dir ? dir : "";
  return IMSRG_OK;
}
extern "C" int imsrg_hip_solver_num_omega(const imsrg_solver *sv, int *n) {
  if (!sv || !n) return IMSRG_E_INVALID;
  *n = (int)sv->omega.size() + (sv->current ? 1 : 0);
  return IMSRG_OK;
}
// segment i into out (the last one is the current segment)
static int solver_omega(imsrg_solver *sv, int i, imsrg_dop *out) {
  imsrg_ctx *ctx = sv->ctx;
  const int nfin = (int)sv->omega.size();
  if (i == nfin && sv->current) return imsrg_hip_op_copy(out, sv->current);
  if (i < 0 || i >= nfin) return fail(ctx, IMSRG_E_INVALID, "Omega index out of range");
  if (sv->omega[i]) return imsrg_hip_op_copy(out, sv->omega[i]);
  return imsrg_hip_op_read_binary(out, sv->files[i].c_str());
}
extern "C" int imsrg_hip_solver_get_omega(imsrg_solver *sv, int i, imsrg_dop *out) {
  if (!sv || !out || out->ctx != sv->ctx) return IMSRG_E_INVALID;
  return solver_omega(sv, i, out);
}

// NewOmega: the current segment joins the list (in HBM, or written to the
// scratch directory as OMEGA_<pid><n>, the reference's file name) and a
// cleared one takes its place.
static int new_omega(imsrg_solver *sv) {
  imsrg_ctx *ctx = sv->ctx;
  if (sv->scratch.empty()) {
    sv->omega.push_back(sv->current);
    sv->files.emplace_back();
    sv->current = nullptr;
    if (int rc = imsrg_hip_op_create(ctx, -1, &sv->current)) return rc;
    return IMSRG_OK;
  }
  char name[64];
  std::snprintf(name, sizeof(name), "/OMEGA_%06d%03d", (int)getpid(), sv->written);
  std::string path = sv->scratch + name;
  if (int rc = imsrg_hip_op_write_binary(sv->current, path.c_str())) return rc;
  sv->omega.push_back(nullptr);
  sv->files.push_back(path);
  sv->written++;
  return imsrg_hip_op_zero(sv->current);
}

static int magnus_run(imsrg_ctx *ctx, imsrg_solver *sv, const imsrg_dop *H0, imsrg_dop *Hs, double ds, double smax,
                      double eta_crit, double omega_norm_max, const imsrg_bch_opts &o, int max_steps, double *e0,
                      double *eta_norm, int *ncomm, int *nsteps) {
  imsrg_dop *Eta = nullptr, *Om2 = nullptr, *Hsaved = nullptr;
  int rc = IMSRG_OK;
  auto cleanup = [&]() {
    imsrg_hip_op_destroy(Eta); imsrg_hip_op_destroy(Om2); imsrg_hip_op_destroy(Hsaved);
  };
  solver_clear(sv);
  if ((rc = imsrg_hip_op_create(ctx, -1, &Eta)) || (rc = imsrg_hip_op_create(ctx, -1, &sv->current)) ||
      (rc = imsrg_hip_op_create(ctx, -1, &Om2)) || (rc = imsrg_hip_op_create(ctx, 1, &Hsaved))) {
    cleanup();
    return rc;
  }
  rc = imsrg_hip_op_copy(Hs, H0);
  if (!rc) rc = imsrg_hip_op_copy(Hsaved, H0);
  if (!rc) rc = imsrg_hip_white_generator_dev(ctx, Hs, Eta);
  int nsplit = 0, step = 0;
  double s = 0.0, Elast = 0.0;
  if (!rc) {
    HIPCHK(ctx, hipMemcpyAsync(&Elast, H0->buf, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  }
  while (!rc && s < smax && step < max_steps) {
    double ne;
    if ((rc = op_norm(Eta, nullptr, nullptr, &ne))) break;
    if (eta_norm) eta_norm[step] = ne;
    if (ne < eta_crit) break;
    if (ne > 1e12 || std::fabs(Elast) > 1e9) {
      rc = fail(ctx, IMSRG_E_NONFINITE, "flow diverged (||eta|| > 1e12 or |E0| > 1e9)");
      break;
    }
    double no;
    if ((rc = op_norm(sv->current, nullptr, nullptr, &no))) break;
    if (no > omega_norm_max) {  // NewOmega, src/IMSRGSolver.cc:44-84
      if ((rc = imsrg_hip_op_copy(Hsaved, Hs))) break;
      if ((rc = new_omega(sv))) break;
      nsplit++;
    }
    double dss = std::min(ds, smax - s);
    s += dss;
    if ((rc = imsrg_hip_op_scale(Eta, dss))) break;
    int nc1 = 0, nc2 = 0;
    if ((rc = bch_product(ctx, Eta, sv->current, Om2, o, &nc1))) break;
    std::swap(sv->current, Om2);
    sv->current->herm = -1;
    rc = bch_transform(ctx, nsplit == 0 ? H0 : Hsaved, sv->current, Hs, o, &nc2);
    // a transform that reaches max_terms only warns in the reference
    // (src/BCH.cc:192-195), and the flow goes on
    if (rc == IMSRG_E_NOT_CONVERGED) rc = IMSRG_OK;
    if (rc) break;
    if ((rc = imsrg_hip_white_generator_dev(ctx, Hs, Eta))) break;
    double ez;
    HIPCHK(ctx, hipMemcpyAsync(&ez, Hs->buf, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (e0) e0[step] = ez;
    if (ncomm) ncomm[step] = nc1 + nc2;
    Elast = ez;
    step++;
  }
  if (nsteps) *nsteps = step;
  cleanup();
  return rc;
}

extern "C" int imsrg_hip_solver_magnus_euler(imsrg_solver *sv, const imsrg_dop *H0, imsrg_dop *Hs, double ds,
                                             double smax, double eta_crit, double omega_norm_max,
                                             const imsrg_bch_opts *opts, int max_steps, double *e0,
                                             double *eta_norm, int *ncomm, int *nsteps) {
  if (!sv) return IMSRG_E_INVALID;
  imsrg_ctx *ctx = sv->ctx;
  if (!H0 || !Hs || H0 == Hs || H0->ctx != ctx || Hs->ctx != ctx || ds <= 0 || max_steps < 0)
    return fail(ctx, IMSRG_E_INVALID, "invalid flow arguments");
  imsrg_bch_opts o;
  if (int e = bch_defaults(ctx, opts, &o)) return e;
  return magnus_run(ctx, sv, H0, Hs, ds, smax, eta_crit, omega_norm_max, o, max_steps, e0, eta_norm, ncomm, nsteps);
}

extern "C" int imsrg_hip_magnus_flow(imsrg_ctx *ctx, const imsrg_dop *H0, imsrg_dop *Hs, double ds, double smax,
                                     double eta_crit, double omega_norm_max, const imsrg_bch_opts *opts,
                                     int max_steps, double *e0, double *eta_norm, int *ncomm, int *nsteps) {
  if (!ctx) return IMSRG_E_INVALID;
  imsrg_solver sv{ctx, {}, {}, nullptr, "", 0};
  const int rc = imsrg_hip_solver_magnus_euler(&sv, H0, Hs, ds, smax, eta_crit, omega_norm_max, opts, max_steps, e0,
                                               eta_norm, ncomm, nsteps);
  solver_clear(&sv);
  return rc;
}

// Transform_Partial(OpIn, first): Out = BCH_Transform over the segments
// first .. n-1 in turn; InverseTransform: over -Omega_i from the last to the first.
static int solver_apply(imsrg_solver *sv, const imsrg_dop *Op, imsrg_dop *Out, int first, bool inverse,
                        const imsrg_bch_opts *opts, int *ncomm) {
  imsrg_ctx *ctx = sv->ctx;
  if (!Op || !Out || Op->ctx != ctx || Out->ctx != ctx) return fail(ctx, IMSRG_E_INVALID, "invalid operator");
  imsrg_bch_opts o;
  if (int e = bch_defaults(ctx, opts, &o)) return e;
  int n = 0;
  imsrg_hip_solver_num_omega(sv, &n);
  if (first < 0) return fail(ctx, IMSRG_E_INVALID, "first segment index is negative");
  imsrg_dop *om = nullptr, *cur = nullptr;
  int rc = imsrg_hip_op_create(ctx, -1, &om);
  if (!rc) rc = imsrg_hip_op_create(ctx, Op->herm, &cur);
  if (!rc) rc = imsrg_hip_op_copy(Out, Op);
  int total = 0;
  for (int k = 0; !rc && k < n - (inverse ? 0 : first); ++k) {
    const int i = inverse ? n - 1 - k : first + k;
    if ((rc = solver_omega(sv, i, om))) break;
    if (inverse && (rc = imsrg_hip_op_scale(om, -1.0))) break;
    if ((rc = imsrg_hip_op_copy(cur, Out))) break;
    int nc = 0;
    rc = bch_transform(ctx, cur, om, Out, o, &nc);
    if (rc == IMSRG_E_NOT_CONVERGED) rc = IMSRG_OK;  // warned only (src/BCH.cc:192-195)
    total += nc;
  }
  if (ncomm) *ncomm = total;
  imsrg_hip_op_destroy(om);
  imsrg_hip_op_destroy(cur);
  if (!rc) HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return rc;
}

extern "C" int imsrg_hip_solver_transform(imsrg_solver *sv, const imsrg_dop *Op, imsrg_dop *Out, int first,
                                          const imsrg_bch_opts *opts, int *ncomm) {
  if (!sv) return IMSRG_E_INVALID;
  return solver_apply(sv, Op, Out, first, false, opts, ncomm);
}
extern "C" int imsrg_hip_solver_inverse_transform(imsrg_solver *sv, const imsrg_dop *Op, imsrg_dop *Out,
                                                  const imsrg_bch_opts *opts, int *ncomm) {
  if (!sv) return IMSRG_E_INVALID;
  return solver_apply(sv, Op, Out, 0, true, opts, ncomm);
}

// ---------------------------------------------------------------------------
// Batched commutators: one launch per phase over the concatenated work lists.
extern "C" int imsrg_hip_commutator_ss_batched(int npairs, imsrg_ctx *const *ctxs, const imsrg_dop *const *X,
                                               const imsrg_dop *const *Y, imsrg_dop *const *Z,
                                               const imsrg_comm_opts *opts) {
  if (npairs <= 0 || !ctxs || !X || !Y || !Z) return IMSRG_E_INVALID;
  imsrg_ctx *c0 = ctxs[0];
  if (!c0) return IMSRG_E_INVALID;
  unsigned mask;
  int prec;
  int rc = comm_opts_ok(c0, opts, &mask, &prec);
  if (rc) return rc;
  for (int i = 0; i < npairs; ++i) {
    if (!ctxs[i] || !X[i] || !Y[i] || !Z[i]) return fail(c0, IMSRG_E_INVALID, "null entry in batch");
    if (ctxs[i]->device != c0->device) return fail(c0, IMSRG_E_INVALID, "batch spans devices");
    if (X[i]->ctx != ctxs[i] || Y[i]->ctx != ctxs[i] || Z[i]->ctx != ctxs[i])
      return fail(c0, IMSRG_E_INVALID, "operator from another context");
    if (X[i]->herm == 0 || Y[i]->herm == 0) return fail(c0, IMSRG_E_UNSUPPORTED, "non-Hermitian operator in batch");
    if (Z[i] == X[i] || Z[i] == Y[i]) return fail(c0, IMSRG_E_INVALID, "Z aliases X or Y");
  }
  BatchCache &bc = c0->batch;
  // The cached device table holds the operator buffers and each context's
  // basis tables, so the key names those (and each context by its serial
  // number, never reused), not the handle addresses a new object may take over.
  std::vector<long long> key;
  for (int i = 0; i < npairs; ++i) {
    key.push_back(ctxs[i]->serial);
    key.push_back((long long)(intptr_t)X[i]->buf);
    key.push_back((long long)(intptr_t)Y[i]->buf);
    key.push_back((long long)(intptr_t)Z[i]->buf);
    key.push_back(X[i]->herm * 4 + Y[i]->herm + 8 * (long long)mask);
    key.push_back(ctxs[i]->lists_gen);  // work lists rebuilt by a new tuning
  }
  if (key != bc.key) {
    bc.release();
    // per-pair workspaces (a context used by several pairs needs separate ones)
    long wtot = 0;
    std::vector<long> woff(npairs);
    for (int i = 0; i < npairs; ++i) {
      woff[i] = wtot;
      wtot += workspace_doubles(ctxs[i]->db.h, true);
    }
    HIPCHK(c0, hipMalloc(&bc.work, sizeof(double) * std::max<long>(wtot, 1)));
    HIPCHK(c0, hipMemset(bc.work, 0, sizeof(double) * std::max<long>(wtot, 1)));
    std::vector<PairPtrs> tab(npairs);
    for (int i = 0; i < npairs; ++i) {
      const BasisDev &B = ctxs[i]->db.h;
      PairPtrs &p = tab[i];
      p.basis = ctxs[i]->db.d;
      p.X = X[i]->buf; p.Y = Y[i]->buf; p.Z = Z[i]->buf;
      const Workspace w = carve_workspace(B, bc.work + woff[i]);
      p.Xt = w.Xt; p.Yt = w.Yt; p.Zbar = w.Zbar; p.scratch = w.scratch;
      p.c122 = w.c122; p.obpart = w.obpart; p.diag = w.diag; p.panel = w.panel; p.c122w = w.c122w;
      p.hX = X[i]->herm; p.hY = Y[i]->herm; p.hZ = herm_rule(X[i]->herm, Y[i]->herm);
      p.mask = mask;
    }
    HIPCHK(c0, hipMalloc(&bc.d_tab, sizeof(PairPtrs) * npairs));
    HIPCHK(c0, hipMemcpy(bc.d_tab, tab.data(), sizeof(PairPtrs) * npairs, hipMemcpyHostToDevice));
    // concatenated lists; pp/hh tiles of all pairs go through the XCD-balanced order
    std::vector<int4> all[10];
    std::vector<TileGroup> groups;
    for (int i = 0; i < npairs; ++i) {
      const WorkLists &wl = ctxs[i]->db.wl;
      const std::vector<int4> *src[10] = {&wl.pphh, &wl.ph,     &wl.ob221,  &wl.pandya, &wl.invp,
                                          &wl.one,  &wl.mirror, &wl.pphh64, &wl.pack,   &wl.c122w};
      {  // ph tiles in the tile size of the first context, whose kernels run the batch
        const int tile = c0->db.tune.ph_big ? kTileB : kTile;
        const DeviceBasis &dbi = ctxs[i]->db;
        for (int c = 0; c < (int)dbi.cc_n.size(); ++c) {
          const int nk = dbi.cc_n[c];
          if (nk == 0 || dbi.cc_nph[c] == 0) continue;
          const int nt = (nk + tile - 1) / tile;
          for (int rb = 0; rb < nt; ++rb)
            for (int sb = rb; sb < nt; ++sb) all[1].push_back(make_int4(i, c, rb, sb));
        }
      }
      for (int t = 2; t < 10; ++t)
        for (int4 w : *src[t]) {
          w.x = i;
          all[t].push_back(w);
        }
      // regroup this pair's pp/hh tiles by (channel, row block)
      std::map<std::pair<int, int>, TileGroup> gm;
      for (int4 w : wl.pphh) {
        if (w.y < 0) continue;
        w.x = i;
        TileGroup &g = gm[{w.y, w.z}];
        g.tiles.push_back(w);
        g.cost += (long)ctxs[i]->hb.ch[w.y].kets.size();
      }
      for (auto &kv : gm) groups.push_back(kv.second);
      bc.maxn = std::max(bc.maxn, ctxs[i]->db.h.maxn);
      bc.maxnch = std::max(bc.maxnch, ctxs[i]->db.h.nch);
      bc.maxjmax = std::max(bc.maxjmax, ctxs[i]->db.h.jmax);
      bc.maxnorb = std::max(bc.maxnorb, ctxs[i]->db.h.norb);
      bc.maxn1pairs = std::max(bc.maxn1pairs, ctxs[i]->db.h.n1pairs);
      bc.maxn122 = std::max(bc.maxn122, ctxs[i]->db.h.n122);
      bc.maxncls = std::max(bc.maxncls, ctxs[i]->db.h.ncls);
      bc.maxbin = std::max(bc.maxbin, ctxs[i]->db.h.c122_maxbin);
      bc.maxent = std::max(bc.maxent, ctxs[i]->db.h.c122_maxent);
    }
    if (c0->db.tune.pphh_order >= 1) {
      all[0] = xcd_order(groups, kXcd);
    } else {  // groups by decreasing k-length, as for a single commutator
      std::stable_sort(groups.begin(), groups.end(), [](const TileGroup &a, const TileGroup &b) {
        return a.cost / (long)a.tiles.size() > b.cost / (long)b.tiles.size();
      });
      for (auto &g : groups) all[0].insert(all[0].end(), g.tiles.begin(), g.tiles.end());
    }
    size_t tot = 0;
    for (int t = 0; t < 10; ++t) {
      bc.off[t] = tot;
      bc.cnt[t] = (int)all[t].size();
      tot += all[t].size();
    }
    HIPCHK(c0, hipMalloc(&bc.d_items, sizeof(int4) * std::max<size_t>(tot, 1)));
    for (int t = 0; t < 10; ++t)
      if (!all[t].empty())
        HIPCHK(c0, hipMemcpy(bc.d_items + bc.off[t], all[t].data(), sizeof(int4) * all[t].size(),
                             hipMemcpyHostToDevice));
    bc.key = key;
  }
  for (int i = 0; i < npairs; ++i) Z[i]->herm = herm_rule(X[i]->herm, Y[i]->herm);
  PairArg pa;
  std::memset(&pa.p0, 0, sizeof(pa.p0));
  pa.tab = bc.d_tab;
  const int4 *items[10];
  for (int t = 0; t < 10; ++t) items[t] = bc.d_items + bc.off[t];
  // The batch runs on c0's stream.  Work queued on the other contexts' streams
  // (uploads, copies, axpy, generators of X_i / Y_i) finishes before it starts,
  // and their later work (downloads, norms of Z_i) waits for it to end.
  std::vector<imsrg_ctx *> others;
  for (int i = 0; i < npairs; ++i)
    if (ctxs[i] != c0 && std::find(others.begin(), others.end(), ctxs[i]) == others.end()) others.push_back(ctxs[i]);
  for (imsrg_ctx *c : others) {
    HIPCHK(c0, hipEventRecord(c->batch_ev, c->stream));
    HIPCHK(c0, hipStreamWaitEvent(c0->stream, c->batch_ev, 0));
  }
  rc = run_commutator(c0, pa, npairs, items, bc.cnt, mask, prec, bc.maxn, bc.maxnch, bc.maxn122, bc.maxjmax,
                      bc.maxnorb, bc.maxn1pairs, bc.maxncls, bc.maxbin, bc.maxent);
  if (rc) return rc;
  if (!others.empty()) {
    HIPCHK(c0, hipEventRecord(c0->batch_ev, c0->stream));
    for (imsrg_ctx *c : others) HIPCHK(c0, hipStreamWaitEvent(c->stream, c0->batch_ev, 0));
  }
  return IMSRG_OK;
}

// ---------------------------------------------------------------------------
// launch tuning
static void tuning_out(const Tuning &t, double us, imsrg_tuning *o) {
  o->pphh_small_tile_max = t.pphh_small;
  o->pphh_order = t.pphh_order;
  o->mixed_pphh_fp32 = t.mixed_pphh;
  o->mixed_ph_fp32 = t.mixed_ph;
  o->ph_tile_128 = t.ph_big;
  o->finish_order = t.finish_order;
  o->overlap_pphh = t.overlap_pphh;
  o->us_per_commutator = us;
}

static int apply_tuning(imsrg_ctx *ctx, const Tuning &t) {
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (auto &g : ctx->graphs) (void)hipGraphExecDestroy(g.exec);  // their work lists are gone
  ctx->graphs.clear();
  ctx->batch.release();
  const std::string e = ctx->db.build_lists(t);
  if (!e.empty()) return fail(ctx, IMSRG_E_HIP, e);
  ++ctx->lists_gen;
  return IMSRG_OK;
}

extern "C" int imsrg_hip_get_tuning(imsrg_ctx *ctx, imsrg_tuning *t) {
  if (!ctx || !t) return IMSRG_E_INVALID;
  tuning_out(ctx->db.tune, 0.0, t);
  return IMSRG_OK;
}

extern "C" int imsrg_hip_set_tuning(imsrg_ctx *ctx, const imsrg_tuning *t) {
  if (!ctx || !t) return IMSRG_E_INVALID;
  if (t->pphh_small_tile_max < 0 || t->pphh_order < 0 || t->pphh_order > 2 ||
      t->mixed_pphh_fp32 < 0 || t->mixed_pphh_fp32 > 2 || t->mixed_ph_fp32 < 0 || t->mixed_ph_fp32 > 2 ||
      (t->ph_tile_128 != 0 && t->ph_tile_128 != 1) || (t->finish_order != 0 && t->finish_order != 1) ||
      t->overlap_pphh < 0 || t->overlap_pphh > 2)
    return fail(ctx, IMSRG_E_INVALID, "invalid tuning values");
  Tuning n;
  n.pphh_small = t->pphh_small_tile_max;
  n.pphh_order = t->pphh_order;
  n.mixed_pphh = t->mixed_pphh_fp32;
  n.mixed_ph = t->mixed_ph_fp32;
  n.ph_big = t->ph_tile_128;
  n.finish_order = t->finish_order;
  n.overlap_pphh = t->overlap_pphh;
  return apply_tuning(ctx, n);
}

// cache key: basis (orbits and occupations), precision, GPU architecture
static std::string tuning_key(imsrg_ctx *ctx, int precision) {
  unsigned long long h = 1469598103934665603ull;  // FNV-1a over the orbit table
  auto mix = [&](long long v) {
    for (int b = 0; b < 8; ++b) {
      h ^= (unsigned long long)((v >> (8 * b)) & 0xff);
      h *= 1099511628211ull;
    }
  };
  const HostBasis &hb = ctx->hb;
  for (int i = 0; i < hb.norb; ++i) {
    mix(hb.n[i]); mix(hb.l[i]); mix(hb.j2[i]); mix(hb.tz2[i]); mix(hb.cvq[i]);
    mix((long long)std::llround(hb.occ[i] * 1e9));
  }
  hipDeviceProp_t prop;
  std::string arch = "unknown";
  if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) arch = prop.gcnArchName;
  char buf[256];
  std::snprintf(buf, sizeof(buf), "imsrg_hip_tuning norb=%d emax=%d basis=%016llx prec=%d arch=%s", hb.norb,
                hb.emax, h, precision, arch.c_str());
  return buf;
}

extern "C" int imsrg_hip_autotune(imsrg_ctx *ctx, int precision, const char *cache_path, imsrg_tuning *out) {
  if (!ctx) return IMSRG_E_INVALID;
  if (precision != IMSRG_PREC_FP64 && precision != IMSRG_PREC_MIXED && precision != IMSRG_PREC_SPLIT)
    return fail(ctx, IMSRG_E_INVALID, "unknown precision mode");
  const std::string key = tuning_key(ctx, precision);
  if (cache_path) {
    if (FILE *f = std::fopen(cache_path, "r")) {
      char line[512];
      Tuning t;
      double us = 0.0;
      bool found = false;
      while (std::fgets(line, sizeof(line), f)) {
        if (std::strncmp(line, key.c_str(), key.size()) != 0 || line[key.size()] != ' ') continue;
        if (std::sscanf(line + key.size(),
                        " pphh_small=%d pphh_order=%d mixed_pphh=%d mixed_ph=%d ph_big=%d finish_order=%d"
                        " overlap_pphh=%d us=%lf",
                        &t.pphh_small, &t.pphh_order, &t.mixed_pphh, &t.mixed_ph, &t.ph_big, &t.finish_order,
                        &t.overlap_pphh, &us) == 8)
          found = true;  // the last matching line wins
      }
      std::fclose(f);
      if (found) {
        int rc = apply_tuning(ctx, t);
        if (rc) return rc;
        if (out) tuning_out(t, us, out);
        return IMSRG_OK;
      }
    }
  }
  // candidates: distinct 64-tile channel sets, both tile orders, and under
  // mixed precision each GEMM family on FP32 or FP64 MFMA
  std::vector<int> smalls = {0};
  {
    std::vector<int> ns(ctx->db.ch_n.begin(), ctx->db.ch_n.end());
    for (int th : {64, 128, 256}) {
      bool changes = false;
      for (int n : ns)
        if (n > smalls.back() && n <= th) changes = true;
      if (changes) smalls.push_back(th);
    }
  }
  // the default setting first; a later candidate is kept only if it is more than
  // 2% faster than the best so far, so that timing noise between near-equal
  // settings does not decide
  std::vector<Tuning> cand{Tuning{}};
  // under mixed precision each GEMM family runs FP32 (1) or bf16x3 (2) MFMA;
  // the FP64 kernels (0) stay a setting by hand
  const bool mixed = precision == IMSRG_PREC_MIXED;
  for (int sm : smalls)
    for (int order = 0; order < 3; ++order)
      for (int mp = mixed ? 1 : 2; mp <= 2; ++mp)
        for (int mh = mixed ? 1 : 2; mh <= 2; ++mh)
          for (int pb = 0; pb < 2; ++pb) {
            Tuning t;
            t.pphh_small = sm;
            t.pphh_order = order;
            t.mixed_pphh = mp;
            t.mixed_ph = mh;
            t.ph_big = pb;
            const Tuning d{};
            if (t.pphh_small == d.pphh_small && t.pphh_order == d.pphh_order && t.mixed_pphh == d.mixed_pphh &&
                t.mixed_ph == d.mixed_ph && t.ph_big == d.ph_big)
              continue;  // the default, already first
            cand.push_back(t);
          }
  imsrg_dop *X = nullptr, *Y = nullptr, *Z = nullptr;
  int rc = imsrg_hip_op_create(ctx, -1, &X);
  if (!rc) rc = imsrg_hip_op_create(ctx, 1, &Y);
  if (!rc) rc = imsrg_hip_op_create(ctx, 1, &Z);
  if (!rc) rc = imsrg_hip_op_fill_random(X, -1, 911);
  if (!rc) rc = imsrg_hip_op_fill_random(Y, 1, 912);
  const bool graphs = ctx->use_graphs, prof = ctx->profiling;
  ctx->use_graphs = false;
  ctx->profiling = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(ctx, IMSRG_E_HIP, "hipEventCreate failed");
  Tuning best = ctx->db.tune;
  double best_us = 1e300;
  constexpr int kTimed = 8;          // timed commutators per setting, after one warm-up
  constexpr double kMargin = 0.02;   // a setting must beat the best so far by this fraction
  auto trial = [&](const Tuning &t) -> int {  // time one setting; keep it if it is the fastest so far
    int r = apply_tuning(ctx, t);
    if (!r) r = commutator_dev(ctx, X, Y, Z, IMSRG_ALL_TERMS, precision);  // warm-up
    if (r) return r;
    if (hipEventRecord(e0, ctx->stream) != hipSuccess) return fail(ctx, IMSRG_E_HIP, "hipEventRecord");
    for (int i = 0; i < kTimed && !r; ++i) r = commutator_dev(ctx, X, Y, Z, IMSRG_ALL_TERMS, precision);
    if (r) return r;
    float ms = 0.f;
    if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      return fail(ctx, IMSRG_E_HIP, "timing failed");
    const double us = 1000.0 * ms / kTimed;
    if (us < (1.0 - kMargin) * best_us) {
      best_us = us;
      best = t;
    }
    return IMSRG_OK;
  };
  for (size_t c = 0; !rc && c < cand.size(); ++c) rc = trial(cand[c]);
  // the last kernel's tile order and the pp/hh overlap are independent of the
  // GEMM choices: one more trial each
  if (!rc) {
    Tuning t = best;
    t.finish_order = 1 - best.finish_order;
    rc = trial(t);
  }
  for (int ov = 0; ov <= 2 && !rc; ++ov) {
    if (ov == best.overlap_pphh) continue;
    Tuning t = best;
    t.overlap_pphh = ov;
    rc = trial(t);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (imsrg_dop *o : {X, Y, Z})
    if (o) imsrg_hip_op_destroy(o);
  ctx->use_graphs = graphs;
  ctx->profiling = prof;
  if (rc) return rc;
  if ((rc = apply_tuning(ctx, best))) return rc;
  if (cache_path) {
    if (FILE *f = std::fopen(cache_path, "a")) {
      std::fprintf(f,
                   "%s pphh_small=%d pphh_order=%d mixed_pphh=%d mixed_ph=%d ph_big=%d finish_order=%d"
                   " overlap_pphh=%d us=%.3f\n",
                   key.c_str(), best.pphh_small, best.pphh_order, best.mixed_pphh, best.mixed_ph, best.ph_big,
                   best.finish_order, best.overlap_pphh, best_us);
      std::fclose(f);
    }
  }
  if (out) tuning_out(best, best_us, out);
  return IMSRG_OK;
}

extern "C" int imsrg_hip_set_profiling(imsrg_ctx *ctx, int on) {
  if (!ctx) return IMSRG_E_INVALID;
  ctx->profiling = on != 0;
  return IMSRG_OK;
}
extern "C" int imsrg_hip_phase_times(imsrg_ctx *ctx, double *us, int n) {
  if (!ctx || !us) return IMSRG_E_INVALID;
  while (ctx->ring_count > 0)
    if (int rc = harvest_oldest(ctx)) return rc;
  for (int i = 0; i < n && i < 8; ++i) us[i] = ctx->phase_n ? ctx->phase_sum[i] / ctx->phase_n : 0.0;
  for (double &v : ctx->phase_sum) v = 0.0;
  ctx->phase_n = 0;
  return IMSRG_OK;
}

// Lane layout self-test of v_mfma_f64_16x16x4f64 (not part of the reference interface).
extern "C" int imsrg_hip_selftest_mfma(imsrg_ctx *ctx, double *out1024) {
  if (!ctx || !out1024) return IMSRG_E_INVALID;
  double *d = nullptr;
  HIPCHK(ctx, hipMalloc(&d, 1024 * sizeof(double)));
  hipError_t e = launch_mfma_probe(d, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out1024, d, 1024 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(ctx, IMSRG_E_HIP, hipGetErrorString(e));
  return IMSRG_OK;
}

// Lane layout self-test of v_mfma_f32_16x16x32_bf16 (not part of the reference interface).
extern "C" int imsrg_hip_selftest_mfma_bf16(imsrg_ctx *ctx, const float *a512, const float *b512, float *d256) {
  if (!ctx || !a512 || !b512 || !d256) return IMSRG_E_INVALID;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  float *d = nullptr;
  HIPCHK(ctx, hipMalloc(&d, 1280 * sizeof(float)));
  hipError_t e = hipMemcpyAsync(d, a512, 512 * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d + 512, b512, 512 * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = launch_mfma_bf16_probe(d, d + 512, d + 1024, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d256, d + 1024, 256 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(ctx, IMSRG_E_HIP, hipGetErrorString(e));
  return IMSRG_OK;
}
