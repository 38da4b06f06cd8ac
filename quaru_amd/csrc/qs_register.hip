// qs_register.hip -- the C ABI of include/quaru_hip.h: register lifetime,
// Operation validation and classification, dispatch to the kernels, and
// measurement.
//
// Reference: quaru 0.2.0, src/register.rs (Register) and src/operation.rs
// (Operation).  The reference multiplies every group by the full 2^k x 2^k
// matrix.  Here each incoming matrix is first classified (exact comparisons,
// no tolerance): a matrix bit j is a *control* when the matrix is the
// identity on the half where bit j is 0; the remaining *active* bits carry a
// smaller matrix, which may further be diagonal, or a diagonal that is 1
// except at a few entries.  Applying the classified form gives the same
// amplitudes (identity rows reproduce the input exactly), while touching only
// the amplitudes the gate can change.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/quaru_hip.h"
#include "qs_internal.h"
#include "qs_sweep.h"

using qs::cd;
using qs::Gate;

// ---------------------------------------------------------------------------
// errors

namespace {
struct LastError {
  size_t a = 0, b = 0;
  std::string msg;
};
thread_local LastError g_err;

int fail(int code, size_t a = 0, size_t b = 0, const std::string &msg = std::string()) {
  g_err.a = a;
  g_err.b = b;
  g_err.msg = msg;
  return code;
}

int hip_fail(hipError_t e, const char *where) {
  std::string m = std::string(where) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return fail(QS_ERR_OUT_OF_MEMORY, 0, 0, m);
  return fail(QS_ERR_HIP, (size_t)e, 0, m);
}
}  // namespace

#define QS_HIP(expr)                                    \
  do {                                                  \
    hipError_t e_ = (expr);                             \
    if (e_ != hipSuccess) return hip_fail(e_, #expr);   \
  } while (0)

// ---------------------------------------------------------------------------
// the register

struct qs_register {
  int n = 0;
  int dtype = QS_C128;
  uint64_t len = 1;
  int device = 0;
  hipStream_t stream = nullptr;
  void *psi = nullptr;
  void *scratch = nullptr;  // 2^n amplitudes, only for dense ops with > 5 active bits
  double *d_red = nullptr;  // reduction partials + results
  size_t d_red_bytes = 0;
  double *d_big = nullptr;  // measurement and readout work buffer, grown on demand
  size_t d_big_bytes = 0;
  void *d_upload = nullptr;  // tables of gates too large for the ring, grown on demand
  size_t d_upload_bytes = 0;
  double *h_red = nullptr;  // pinned host mirror of the results
  // matrix upload ring: device + pinned host, slots reused after a sync
  char *d_ring = nullptr;
  char *h_ring = nullptr;
  size_t ring_cap = 0, ring_off = 0;
  bool fusion = true;
  std::vector<Gate> pending;
  // the collapse of the last measure, not yet launched: the next measure
  // runs it fused with its own probability split; anything else launches it
  // first (settle)
  bool coll_pending = false;
  int coll_target = 0, coll_keep = 0;
  double coll_total = 1.0;
  uint64_t st_sweeps = 0, st_kernels = 0, st_gates = 0;
  uint64_t st_sweep_ops = 0, st_merged = 0;
  qs::SweepEngine *engine = nullptr;
  // kernel timing: event pairs recorded on `stream` around each launch
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  struct Mark {
    int kind;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;
  int open_kind = -1;
  hipEvent_t open_ev = nullptr;
  double t_ms[QS_K_COUNT] = {};
  uint64_t t_n[QS_K_COUNT] = {};

  size_t amp_bytes() const { return dtype == QS_C128 ? 16 : 8; }
  size_t state_bytes() const { return amp_bytes() * (size_t)len; }
};

namespace {

constexpr size_t kRingBytes = 8u << 20;
// Readouts (state, probabilities, amplitudes) go through the register's work
// buffer in pieces of at most 2^22 amplitudes (64 MiB widened to complex128),
// so they allocate nothing after the first call.
constexpr int kReadoutBits = 22;
constexpr uint64_t kReadoutAmps = 1ull << kReadoutBits;
constexpr size_t kRedDoubles = 4 * qs::kReduceBlocks + 64;

bool valid_dtype(int d) { return d == QS_C128 || d == QS_C64; }

int ensure_device() {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0)
    return fail(QS_ERR_NO_DEVICE, 0, 0, "no HIP device available");
  return QS_OK;
}

int alloc_register(size_t n, int dtype, qs_register **out) {
  if (!out) return fail(QS_ERR_ARGUMENT, 0, 0, "out is NULL");
  *out = nullptr;
  if (!valid_dtype(dtype)) return fail(QS_ERR_ARGUMENT, (size_t)dtype, 0, "bad dtype");
  if (n > (size_t)qs::kMaxQubits)
    return fail(QS_ERR_ARGUMENT, n, 0, "register too large for one device");
  int rc = ensure_device();
  if (rc) return rc;
  qs_register *r = new qs_register();
  r->n = (int)n;
  r->dtype = dtype;
  r->len = 1ull << n;
  hipError_t e = hipGetDevice(&r->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&r->psi, r->state_bytes());
  if (e == hipSuccess) {
    r->d_red_bytes = kRedDoubles * sizeof(double);
    e = hipMalloc((void **)&r->d_red, r->d_red_bytes);
  }
  if (e == hipSuccess) e = hipHostMalloc((void **)&r->h_red, 64 * sizeof(double));
  if (e == hipSuccess) {
    r->ring_cap = kRingBytes;
    e = hipMalloc((void **)&r->d_ring, r->ring_cap);
  }
  if (e == hipSuccess) e = hipHostMalloc((void **)&r->h_ring, r->ring_cap);
  if (e != hipSuccess) {
    qs_register_free(r);
    return hip_fail(e, "allocating the register");
  }
  r->engine = qs::sweep_engine_new();
  *out = r;
  return QS_OK;
}

// A device work buffer of at least `bytes`; the stream is drained before the
// old buffer is released.
hipError_t ensure_big(qs_register *r, size_t bytes, double **out) {
  if (r->d_big_bytes < bytes) {
    hipError_t e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return e;
    if (r->d_big) (void)hipFree(r->d_big);
    r->d_big = nullptr;
    r->d_big_bytes = 0;
    e = hipMalloc((void **)&r->d_big, bytes);
    if (e != hipSuccess) return e;
    r->d_big_bytes = bytes;
  }
  *out = r->d_big;
  return hipSuccess;
}

template <typename F>
hipError_t by_dtype(const qs_register *r, F &&f) {
  if (r->dtype == QS_C128) return f(double{});
  return f(float{});
}

// Copy `bytes` of host data into the device ring and return its device
// address.  A slot is reused only after the stream has drained.
hipError_t ring_upload(qs_register *r, const void *src, size_t bytes, void **dev) {
  const size_t need = (bytes + 255) & ~(size_t)255;
  if (need > r->ring_cap) return hipErrorInvalidValue;
  if (r->ring_off + need > r->ring_cap) {
    hipError_t e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return e;
    r->ring_off = 0;
  }
  char *h = r->h_ring + r->ring_off;
  char *d = r->d_ring + r->ring_off;
  std::memcpy(h, src, bytes);
  r->ring_off += need;
  *dev = d;
  return hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, r->stream);
}

// Upload a table too large for the ring (dense or permutation gates on many
// active qubits) into the register's upload buffer, grown on demand and
// kept: the stream is drained before a smaller buffer is replaced, and after
// the copy, so the host source may go away.
hipError_t big_upload(qs_register *r, const void *src, size_t bytes, void **dev) {
  hipError_t e = hipSuccess;
  if (r->d_upload_bytes < bytes) {
    e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return e;
    if (r->d_upload) (void)hipFree(r->d_upload);
    r->d_upload = nullptr;
    r->d_upload_bytes = 0;
    e = hipMalloc(&r->d_upload, bytes);
    if (e != hipSuccess) return e;
    r->d_upload_bytes = bytes;
  }
  *dev = r->d_upload;
  e = hipMemcpyAsync(*dev, src, bytes, hipMemcpyHostToDevice, r->stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(r->stream);
}

// ---------------------------------------------------------------------------
// scratch of the no-allocation readout (qs_get_amplitudes_noalloc)
//
// One fixed array per GPU, defined here as a device variable, so that the
// entry point needs no hipMalloc: up to kNoAllocMaxIndices uint64 indices,
// then as many complex128 amplitudes (1.5 MiB, held by every GPU the library
// runs on until the process ends).  On the host, `scratch` names a host-side
// stand-in, not device memory: the device address of each GPU's copy comes
// from hipGetSymbolAddress, called once per device with that device current,
// and only that address is used.  A per-device mutex serialises the calls of
// this process on one GPU; the array keeps the last call's contents.

constexpr size_t kNoAllocMaxIndices = 1u << 16;
constexpr size_t kNoAllocBytes = kNoAllocMaxIndices * (sizeof(uint64_t) + 2 * sizeof(double));
constexpr int kNoAllocDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_noalloc_init_mu;
std::mutex g_noalloc_mu[kNoAllocDevices];
char *g_noalloc_ptr[kNoAllocDevices] = {};

int noalloc_scratch(int device, char **out) {
  if (device < 0 || device >= kNoAllocDevices)
    return fail(QS_ERR_ARGUMENT, (size_t)device, kNoAllocDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(g_noalloc_init_mu);
  if (!g_noalloc_ptr[device]) {
    int prev = -1;
    QS_HIP(hipGetDevice(&prev));
    if (prev != device) QS_HIP(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != device) QS_HIP(hipSetDevice(prev));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(QS_ERR_HIP, 0, 0, "hipGetSymbolAddress(scratch) returned NULL");
    g_noalloc_ptr[device] = static_cast<char *>(p);
  }
  *out = g_noalloc_ptr[device];
  return QS_OK;
}

// ---------------------------------------------------------------------------
// validation (src/register.rs:141-153, :418-429; src/operation.rs:60-69)

int validate_targets(size_t n, const size_t *t, size_t k) {
  if (k >= 2) {
    std::vector<size_t> c(t, t + k);
    std::sort(c.begin(), c.end());
    for (size_t i = 1; i < k; ++i)
      if (c[i] == c[i - 1]) return fail(QS_INVALID_TARGET, c[i], 0, "duplicate target");
  }
  for (size_t j = 0; j < k; ++j)
    if (t[j] >= n) return fail(QS_INVALID_TARGET, t[j], 0, "target out of range");
  return QS_OK;
}

int validate_dims(size_t rows, size_t cols, size_t k) {
  if (k >= 63 || rows != (1ull << k) || cols != (1ull << k))
    return fail(QS_INVALID_DIMENSIONS, rows, cols, "matrix side must be 2^k");
  return QS_OK;
}

// ---------------------------------------------------------------------------
// classification

inline bool is_one(const cd &v) { return v.re == 1.0 && v.im == 0.0; }
inline bool is_zero(const cd &v) { return v.re == 0.0 && v.im == 0.0; }

// Classify a dense 2^k x 2^k matrix on `targets`; `cmask_in/cval_in` are
// controls already known (qs_apply_controlled).
Gate classify(const double *mat, const size_t *targets, size_t k, uint64_t cmask_in,
              uint64_t cval_in) {
  const size_t D = (size_t)1 << k;
  const cd *M = reinterpret_cast<const cd *>(mat);
  Gate g;
  g.cmask = cmask_in;
  g.cval = cval_in;
  // 1. control bits: identity wherever bit j of the row or the column is 0.
  std::vector<int> ctrl(k, 0);
  for (size_t j = 0; j < k; ++j) {
    bool ok = true;
    const size_t bj = (size_t)1 << j;
    for (size_t r = 0; r < D && ok; ++r)
      for (size_t c = 0; c < D; ++c) {
        if ((r & bj) && (c & bj)) continue;
        const cd v = M[r * D + c];
        if (r == c ? !is_one(v) : !is_zero(v)) {
          ok = false;
          break;
        }
      }
    ctrl[j] = ok;
  }
  size_t cbits = 0;
  for (size_t j = 0; j < k; ++j)
    if (ctrl[j]) {
      cbits |= (size_t)1 << j;
      g.cmask |= 1ull << targets[j];
      g.cval |= 1ull << targets[j];
    }
  std::vector<size_t> actj;
  for (size_t j = 0; j < k; ++j)
    if (!ctrl[j]) actj.push_back(j);
  g.na = (int)actj.size();
  for (int a = 0; a < g.na; ++a) g.act[a] = (int)targets[actj[a]];
  const size_t d = (size_t)1 << g.na;
  // reduced matrix: rows/cols with every control bit set
  auto full_index = [&](size_t ri) {
    size_t x = cbits;
    for (int a = 0; a < g.na; ++a)
      if ((ri >> a) & 1) x |= (size_t)1 << actj[a];
    return x;
  };
  std::vector<cd> red(d * d);
  for (size_t r = 0; r < d; ++r)
    for (size_t c = 0; c < d; ++c) red[r * d + c] = M[full_index(r) * D + full_index(c)];
  // 2. diagonal?
  bool diag = true;
  for (size_t r = 0; r < d && diag; ++r)
    for (size_t c = 0; c < d; ++c)
      if (r != c && !is_zero(red[r * d + c])) {
        diag = false;
        break;
      }
  if (!diag) {
    // a permutation with factors (one nonzero per row and column) on more
    // than 5 active qubits is a gather (2 passes over the state) instead of
    // 2^na complex products per amplitude
    if (g.na > 5) {
      std::vector<uint32_t> src(d);
      std::vector<char> used(d, 0);
      bool mono = true;
      for (size_t r = 0; r < d && mono; ++r) {
        size_t nz = 0;
        for (size_t c = 0; c < d; ++c)
          if (!is_zero(red[r * d + c])) {
            ++nz;
            src[r] = (uint32_t)c;
          }
        mono = nz == 1 && !used[src[r]];
        if (mono) used[src[r]] = 1;
      }
      if (mono) {
        g.kind = Gate::kPerm;
        g.perm = std::move(src);
        g.m.resize(d);
        for (size_t r = 0; r < d; ++r) g.m[r] = red[r * d + g.perm[r]];
        return g;
      }
    }
    g.kind = Gate::kDense;
    g.m = std::move(red);
    return g;
  }
  std::vector<cd> dv(d);
  size_t nonunit = 0;
  for (size_t r = 0; r < d; ++r) {
    dv[r] = red[r * d + r];
    if (!is_one(dv[r])) ++nonunit;
  }
  if (nonunit == 0) {
    g.kind = Gate::kIdentity;
    return g;
  }
  if (nonunit <= 16 && (nonunit < d || g.na == 0)) {
    g.kind = Gate::kSparse;
    uint64_t amask = 0;
    for (int a = 0; a < g.na; ++a) amask |= 1ull << g.act[a];
    g.smask = amask | g.cmask;
    for (size_t r = 0; r < d; ++r) {
      if (is_one(dv[r])) continue;
      uint64_t pat = g.cval;
      for (int a = 0; a < g.na; ++a)
        if ((r >> a) & 1) pat |= 1ull << g.act[a];
      g.spat.push_back(pat);
      g.m.push_back(dv[r]);
    }
    return g;
  }
  g.kind = Gate::kDiag;
  g.m = std::move(dv);
  return g;
}

// ---------------------------------------------------------------------------
// dispatch of one classified gate with per-gate kernels

template <typename R>
hipError_t launch_one(qs_register *r, const Gate &g) {
  using T = typename std::conditional<std::is_same<R, double>::value, double, float>::type;
  switch (g.kind) {
    case Gate::kIdentity:
      return hipSuccess;
    case Gate::kSparse:
      ++r->st_kernels;
      return qs::launch_gate_sparse<R>(r->psi, r->n, g, r->stream);
    case Gate::kDiag: {
      std::vector<T> buf(2 * g.m.size());
      for (size_t i = 0; i < g.m.size(); ++i) {
        buf[2 * i] = (T)g.m[i].re;
        buf[2 * i + 1] = (T)g.m[i].im;
      }
      void *dm = nullptr;
      const size_t bytes = buf.size() * sizeof(T);
      hipError_t e = bytes + 256 <= r->ring_cap ? ring_upload(r, buf.data(), bytes, &dm)
                                                : big_upload(r, buf.data(), bytes, &dm);
      if (e != hipSuccess) return e;
      ++r->st_kernels;
      e = qs::launch_gate_diag<R>(r->psi, r->n, g, dm, r->stream);
      return e;
    }
    case Gate::kPerm: {
      const size_t d = (size_t)1 << g.na;
      const size_t off = qs::perm_table_phase_offset(g.na);
      std::vector<char> buf(off + d * 2 * sizeof(T));
      std::memcpy(buf.data(), g.perm.data(), d * sizeof(uint32_t));
      T *ph = reinterpret_cast<T *>(buf.data() + off);
      for (size_t i = 0; i < d; ++i) {
        ph[2 * i] = (T)g.m[i].re;
        ph[2 * i + 1] = (T)g.m[i].im;
      }
      void *dm = nullptr;
      hipError_t e = buf.size() + 256 <= r->ring_cap ? ring_upload(r, buf.data(), buf.size(), &dm)
                                                     : big_upload(r, buf.data(), buf.size(), &dm);
      if (e != hipSuccess) return e;
      if (!r->scratch) {
        e = hipMalloc(&r->scratch, r->state_bytes());
        if (e != hipSuccess) return e;
      }
      ++r->st_kernels;
      e = qs::launch_gate_perm<R>(r->psi, r->n, g, dm, r->scratch, r->stream);
      return e;
    }
    case Gate::kDense: {
      const size_t d = (size_t)1 << g.na;
      void *dm = nullptr;
      hipError_t e = hipSuccess;
      if (g.na >= 3) {
        // row-major {re, im} in the register's precision
        std::vector<T> buf(2 * d * d);
        for (size_t e = 0; e < d * d; ++e) {
          buf[2 * e] = (T)g.m[e].re;
          buf[2 * e + 1] = (T)g.m[e].im;
        }
        const size_t bytes = buf.size() * sizeof(T);
        e = bytes + 256 <= r->ring_cap ? ring_upload(r, buf.data(), bytes, &dm)
                                       : big_upload(r, buf.data(), bytes, &dm);
        if (e != hipSuccess) return e;
        if (g.na > 5 && !r->scratch) {
          e = hipMalloc(&r->scratch, r->state_bytes());
          if (e != hipSuccess) return e;
        }
      }
      ++r->st_kernels;
      e = qs::launch_gate_dense<R>(r->psi, r->n, g, dm, r->scratch, r->stream);
      return e;
    }
  }
  return hipSuccess;
}

hipEvent_t take_event(qs_register *r) {
  if (!r->ev_pool.empty()) {
    hipEvent_t e = r->ev_pool.back();
    r->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Timing brackets: begin records an event before the launch, end one after.
void mark(qs_register *r, int kind, bool begin) {
  if (!r->timing) return;
  if (begin) {
    r->open_ev = take_event(r);
    r->open_kind = kind;
    if (r->open_ev) (void)hipEventRecord(r->open_ev, r->stream);
    return;
  }
  hipEvent_t b = take_event(r);
  if (!r->open_ev || !b) return;
  (void)hipEventRecord(b, r->stream);
  r->marks.push_back({r->open_kind, r->open_ev, b});
  r->open_ev = nullptr;
}

int gate_class(const Gate &g) {
  switch (g.kind) {
    case Gate::kDense:
      return g.na == 1 ? QS_K_GATE1 : g.na == 2 ? QS_K_GATE2 : g.na <= 5 ? QS_K_GATEK : QS_K_BIG;
    case Gate::kDiag: return QS_K_DIAG;
    case Gate::kSparse: return QS_K_SPARSE;
    case Gate::kPerm: return QS_K_PERM;
    default: return QS_K_COUNT - 1;
  }
}

hipError_t launch_gate(qs_register *r, const Gate &g) {
  if (g.kind == Gate::kIdentity) return hipSuccess;
  mark(r, gate_class(g), true);
  hipError_t e = by_dtype(r, [&](auto z) { return launch_one<decltype(z)>(r, g); });
  mark(r, gate_class(g), false);
  return e;
}

// Fold recorded event pairs into the per-class totals (synchronises).
hipError_t collect_marks(qs_register *r) {
  if (r->marks.empty()) return hipSuccess;
  hipError_t e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess) return e;
  for (const auto &m : r->marks) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess && m.kind >= 0 && m.kind < QS_K_COUNT) {
      r->t_ms[m.kind] += ms;
      r->t_n[m.kind] += 1;
    }
    r->ev_pool.push_back(m.a);
    r->ev_pool.push_back(m.b);
  }
  r->marks.clear();
  return hipSuccess;
}

// Launch the pending collapse of the last measure, if any.
int settle(qs_register *r) {
  if (!r->coll_pending) return QS_OK;
  r->coll_pending = false;
  const hipError_t e =
      r->dtype == QS_C128
          ? qs::launch_collapse<double>(r->psi, r->n, r->coll_target, r->coll_keep,
                                        r->coll_total, r->stream)
          : qs::launch_collapse<float>(r->psi, r->n, r->coll_target, r->coll_keep, r->coll_total,
                                       r->stream);
  if (e != hipSuccess) return hip_fail(e, "collapse");
  return QS_OK;
}

// Apply the pending collapse, then every pending gate: through the sweep
// engine when fusion is on, otherwise one kernel per gate.
int flush(qs_register *r) {
  int rc = settle(r);
  if (rc) return rc;
  if (r->pending.empty()) return QS_OK;
  std::vector<Gate> gates;
  gates.swap(r->pending);
  hipError_t e = hipSuccess;
  if (r->fusion) {
    qs::SweepRunner run;
    run.n = r->n;
    run.dtype_c64 = r->dtype == QS_C64;
    run.psi = r->psi;
    run.stream = r->stream;
    run.upload = [r](const void *src, size_t bytes, void **dev) {
      return ring_upload(r, src, bytes, dev);
    };
    run.single = [r](const Gate &g) { return launch_gate(r, g); };
    if (r->timing) run.mark = [r](bool begin) { mark(r, QS_K_SWEEP, begin); };
    run.sweep_ops = &r->st_sweep_ops;
    run.merged = &r->st_merged;
    e = qs::sweep_execute(r->engine, run, gates, &r->st_sweeps, &r->st_kernels);
  } else {
    for (const Gate &g : gates) {
      e = launch_gate(r, g);
      if (e != hipSuccess) break;
    }
  }
  if (e != hipSuccess) return hip_fail(e, "applying gates");
  return QS_OK;
}

constexpr size_t kMaxPending = 4096;

int enqueue(qs_register *r, Gate &&g) {
  r->st_gates += (uint64_t)g.logical;
  if (g.kind == Gate::kIdentity) return QS_OK;
  r->pending.push_back(std::move(g));
  if (!r->fusion || r->pending.size() >= kMaxPending) return flush(r);
  return QS_OK;
}

int sync(qs_register *r) {
  int rc = flush(r);
  if (rc) return rc;
  QS_HIP(hipStreamSynchronize(r->stream));
  return QS_OK;
}

template <typename R>
hipError_t reduce_norm(qs_register *r, double *out) {
  hipError_t e = qs::launch_norm<R>(r->psi, r->n, r->d_red, r->d_red + 4 * qs::kReduceBlocks,
                                    r->stream);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(r->h_red, r->d_red + 4 * qs::kReduceBlocks, sizeof(double),
                     hipMemcpyDeviceToHost, r->stream);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(r->stream);
  *out = r->h_red[0];
  return e;
}

// One measure (src/register.rs:297-342) on the device state.  The split
// reads the state (fused with the previous measure's collapse when that is
// still pending); the collapse itself is left pending, for the next measure
// to fuse or for the next other call to launch (settle).  Pending gates must
// have been flushed.
template <typename R>
int measure_one(qs_register *r, size_t target, double x, int *result) {
  double *res_dev = r->d_red + 4 * qs::kReduceBlocks;
  if (r->coll_pending) {
    r->coll_pending = false;
    QS_HIP(qs::launch_collapse_split<R>(r->psi, r->n, r->coll_target, r->coll_keep,
                                        r->coll_total, (int)target, r->d_red, res_dev,
                                        r->stream));
  } else {
    QS_HIP(qs::launch_prob_split<R>(r->psi, r->n, (int)target, r->d_red, res_dev, r->stream));
  }
  QS_HIP(hipMemcpyAsync(r->h_red, res_dev, 2 * sizeof(double), hipMemcpyDeviceToHost,
                        r->stream));
  QS_HIP(hipStreamSynchronize(r->stream));
  const double p0 = r->h_red[0], p1 = r->h_red[1];
  const int res = x > p0;
  r->coll_pending = true;
  r->coll_target = (int)target;
  r->coll_keep = res;
  r->coll_total = res ? p1 : p0;
  *result = res;
  return QS_OK;
}

// Device slots of deferred measurements: after the reduction results and the
// collapse parameters in the tail of d_red.
unsigned char *slot_base(qs_register *r) {
  return reinterpret_cast<unsigned char *>(r->d_red + 4 * qs::kReduceBlocks + 32);
}

int measure_dispatch(qs_register *r, size_t target, double x, int *result) {
  if (r->dtype == QS_C128) return measure_one<double>(r, target, x, result);
  return measure_one<float>(r, target, x, result);
}

// Mask/pattern collapse: keep amplitudes with (idx & mask) == pat, divided by
// sqrt(total); zero the rest.  Implemented with the sparse-free path: one
// diagonal-style pass.
template <typename R>
__global__ void k_collapse_mask(typename std::conditional<std::is_same<R, double>::value, double2,
                                                          float2>::type *psi,
                                uint64_t len, uint64_t mask, uint64_t pat, double sq) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * 256) {
    auto a = psi[i];
    if ((i & mask) != pat) {
      a.x = 0;
      a.y = 0;
    } else {
      a.x = (R)((double)a.x / sq);
      a.y = (R)((double)a.y / sq);
    }
    psi[i] = a;
  }
}

template <typename R>
hipError_t collapse_mask(qs_register *r, uint64_t mask, uint64_t pat, double total) {
  using T = typename std::conditional<std::is_same<R, double>::value, double2, float2>::type;
  // states above the last-level cache (256 MiB): the stream order of
  // launch_collapse_pattern
  if (((uint64_t)sizeof(T) << r->n) > (256ull << 20))
    return qs::launch_collapse_pattern<R>(r->psi, r->n, mask, pat, total, r->stream);
  uint64_t g = (r->len + 1023) / 1024;
  if (g > 262144) g = 262144;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_collapse_mask<R>, dim3((unsigned)g), dim3(256), 0, r->stream, (T *)r->psi,
                     r->len, mask, pat, std::sqrt(total));
  return hipGetLastError();
}

// The sequence of measure calls computed from one probability pass over the
// first m qubits of `order` plus a gather of the 2^(n-m) amplitudes that stay
// consistent with those outcomes.  Returns 1 if it could not decide exactly
// (a zero-probability branch), 0 on success.
template <typename R>
int measure_fast(qs_register *r, const size_t *order, size_t count, const double *u,
                 uint64_t *index, int *fallback) {
  *fallback = 0;
  const int n = r->n;
  const int m = (int)std::min<size_t>({count, (size_t)12, (size_t)n});
  // histogram over the pattern of the qubits order[0..m)
  uint64_t smask = 0;
  for (int j = 0; j < m; ++j) smask |= 1ull << order[j];
  const int H = 1 << m;
  std::vector<double> hist(H);
  {
    qs::Fixed fx = qs::make_fixed(smask, 0);
    const uint64_t nfree = 1ull << (n - m);
    int blocks = (int)std::min<uint64_t>(1024, nfree);
    const size_t rows = (size_t)blocks * (H >= 256 ? 1 : 256 / H);  // of k_set_hist
    double *d_work = nullptr;
    QS_HIP(ensure_big(r, (rows + 1) * H * sizeof(double), &d_work));
    double *d_part = d_work, *d_hist = d_work + rows * H;
    uint8_t pos_order[64];
    for (int j = 0; j < m; ++j) pos_order[j] = (uint8_t)order[j];
    hipError_t e = qs::launch_set_histogram<R>(r->psi, n, fx, pos_order, m, d_part, blocks,
                                               d_hist, r->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(hist.data(), d_hist, H * sizeof(double), hipMemcpyDeviceToHost,
                         r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return hip_fail(e, "measure histogram");
  }
  double total = 0.0;
  for (int h = 0; h < H; ++h) total += hist[h];
  // decide order[0..m) sequentially on bins consistent with the decisions
  uint64_t pat = 0;  // bin-space pattern of decided bits
  uint64_t gidx = 0;  // global-index pattern
  double branch = total;
  for (int j = 0; j < m; ++j) {
    double s0 = 0.0, s1 = 0.0;
    const uint64_t low = (1ull << j) - 1;
    for (int h = 0; h < H; ++h) {
      if (((uint64_t)h & low) != pat) continue;
      if ((h >> j) & 1)
        s1 += hist[h];
      else
        s0 += hist[h];
    }
    const double p0 = s0 / (s0 + s1);
    const int res = u[j] > p0;
    const double keep = res ? s1 : s0;
    if (!(keep > 0.0)) {
      *fallback = 1;
      return QS_OK;
    }
    if (res) {
      pat |= 1ull << j;
      gidx |= 1ull << order[j];
    }
    branch = keep;
  }
  uint64_t decided_mask = smask;
  if ((int)count > m) {
    // gather the 2^(n-m) probabilities consistent with the first m outcomes
    const uint64_t nfree = 1ull << (n - m);
    std::vector<double> pr(nfree);
    double *d_pr = nullptr;
    QS_HIP(ensure_big(r, nfree * sizeof(double), &d_pr));
    qs::Fixed fx = qs::make_fixed(smask, gidx);
    hipError_t e = qs::launch_gather_probs<R>(r->psi, fx, nfree, d_pr, r->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(pr.data(), d_pr, nfree * sizeof(double), hipMemcpyDeviceToHost,
                         r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return hip_fail(e, "measure gather");
    // free-counter bit of global qubit q
    auto free_bit = [&](size_t q) {
      int below = 0;
      for (int b = 0; b < (int)q; ++b)
        if ((smask >> b) & 1ull) ++below;
      return (int)q - below;
    };
    std::vector<uint32_t> live(nfree);
    for (uint64_t i = 0; i < nfree; ++i) live[i] = (uint32_t)i;
    for (size_t j = (size_t)m; j < count; ++j) {
      const int fb = free_bit(order[j]);
      double s0 = 0.0, s1 = 0.0;
      for (uint32_t i : live) {
        if ((i >> fb) & 1u)
          s1 += pr[i];
        else
          s0 += pr[i];
      }
      const double p0 = s0 / (s0 + s1);
      const int res = u[j] > p0;
      const double keep = res ? s1 : s0;
      if (!(keep > 0.0)) {
        *fallback = 1;
        return QS_OK;
      }
      if (res) gidx |= 1ull << order[j];
      decided_mask |= 1ull << order[j];
      std::vector<uint32_t> nl;
      nl.reserve(live.size() / 2 + 1);
      for (uint32_t i : live)
        if ((int)((i >> fb) & 1u) == res) nl.push_back(i);
      live.swap(nl);
      branch = keep;
    }
  }
  // The reference divides by sqrt(t_j) at every step; the product of the
  // t_j telescopes to branch / total.
  QS_HIP(collapse_mask<R>(r, decided_mask, gidx, branch / total));
  *index = gidx;
  return QS_OK;
}

// A dense gate on several active qubits that is a Kronecker product of
// 2x2 factors (Hadamard transforms, layers of one-qubit gates that a caller
// fused into one matrix) is applied as those factors: each is a one-qubit
// gate that the sweep engine fuses with its neighbours, one pass over the
// state for many of them, where the dense form costs a pass (and 2^k complex
// products per amplitude) for each block.  Controls carry over to every
// factor (C-(A (x) B) = C-A . C-B).  The product must reproduce the matrix
// to 1e-14 of its largest entry, far inside the engine's 1e-12 bar.
bool kron_factors(const Gate &g, std::vector<std::array<cd, 4>> &f) {
  if (g.kind != Gate::kDense || g.na < 2 || g.na > 12) return false;
  const int k = g.na;
  const size_t d = (size_t)1 << k;
  const cd *M = g.m.data();
  size_t piv = 0;
  double amax = -1.0;
  for (size_t i = 0; i < d * d; ++i) {
    const double a = std::hypot(M[i].re, M[i].im);
    if (a > amax) {
      amax = a;
      piv = i;
    }
  }
  if (!(amax > 0.0)) return false;
  const size_t r0 = piv / d, c0 = piv % d;
  const cd p = M[piv];
  const double pn = p.re * p.re + p.im * p.im;
  auto cdiv = [&](const cd &v) {  // v / p
    return cd{(v.re * p.re + v.im * p.im) / pn, (v.im * p.re - v.re * p.im) / pn};
  };
  auto cmul = [](const cd &a, const cd &b) {
    return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  };
  f.assign(k, {});
  for (int a = 0; a < k; ++a) {
    const size_t ba = (size_t)1 << a;
    for (int x = 0; x < 2; ++x)
      for (int y = 0; y < 2; ++y) {
        const size_t r = x ? (r0 | ba) : (r0 & ~ba), c = y ? (c0 | ba) : (c0 & ~ba);
        f[a][2 * x + y] = cdiv(M[r * d + c]);
      }
  }
  for (int q = 0; q < 4; ++q) f[0][q] = cmul(f[0][q], p);
  const double tol = 1e-14 * amax;
  for (size_t r = 0; r < d; ++r)
    for (size_t c = 0; c < d; ++c) {
      cd v{1.0, 0.0};
      for (int a = 0; a < k; ++a) v = cmul(v, f[a][2 * ((r >> a) & 1) + ((c >> a) & 1)]);
      if (std::hypot(v.re - M[r * d + c].re, v.im - M[r * d + c].im) > tol) return false;
    }
  return true;
}

// Enqueue a classified gate, as its one-qubit factors when it has them and
// fusion is on (with fusion off every factor would be a pass of its own).
int enqueue_split(qs_register *r, Gate &&g) {
  std::vector<std::array<cd, 4>> f;
  if (!r->fusion || !kron_factors(g, f)) return enqueue(r, std::move(g));
  int rc = QS_OK;
  for (int a = 0; a < g.na && rc == QS_OK; ++a) {
    const size_t t = (size_t)g.act[a];
    Gate one = classify(reinterpret_cast<const double *>(f[a].data()), &t, 1, g.cmask, g.cval);
    one.logical = a == 0 ? g.logical : 0;
    rc = enqueue(r, std::move(one));
  }
  return rc;
}

}  // namespace

qs::Gate qs::classify_matrix(const double *mat, const size_t *targets, size_t k, uint64_t cmask,
                             uint64_t cval) {
  return classify(mat, targets, k, cmask, cval);
}

int qs::set_error(int code, size_t a, size_t b, const std::string &msg) {
  return fail(code, a, b, msg);
}

int qs::set_hip_error(hipError_t e, const char *where) { return hip_fail(e, where); }

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

int qs_last_error(size_t *a, size_t *b, const char **message) {
  if (a) *a = g_err.a;
  if (b) *b = g_err.b;
  if (message) *message = g_err.msg.c_str();
  return QS_OK;
}

const char *qs_version(void) { return "quaru-hip 0.1.0 (gfx950)"; }

int qs_register_new(const unsigned char *bits, size_t n, int dtype, qs_register **out) {
  if (n > 0 && !bits) return fail(QS_ERR_ARGUMENT, 0, 0, "bits is NULL");
  int rc = alloc_register(n, dtype, out);
  if (rc) return rc;
  uint64_t k = 0;
  for (size_t i = 0; i < n; ++i)
    if (bits[i]) k |= 1ull << i;
  qs_register *r = *out;
  hipError_t e = by_dtype(r, [&](auto z) {
    return qs::launch_init_basis<decltype(z)>(r->psi, r->n, k, r->stream);
  });
  if (e != hipSuccess) {
    qs_register_free(r);
    *out = nullptr;
    return hip_fail(e, "qs_register_new");
  }
  return QS_OK;
}

int qs_register_from_int(size_t n, size_t k, int dtype, qs_register **out) {
  if (n < 64 && k >= (1ull << n))
    return fail(QS_ERR_ARGUMENT, k, n, "basis index out of range");
  int rc = alloc_register(n, dtype, out);
  if (rc) return rc;
  qs_register *r = *out;
  hipError_t e = by_dtype(r, [&](auto z) {
    return qs::launch_init_basis<decltype(z)>(r->psi, r->n, k, r->stream);
  });
  if (e != hipSuccess) {
    qs_register_free(r);
    *out = nullptr;
    return hip_fail(e, "qs_register_from_int");
  }
  return QS_OK;
}

int qs_register_from_qubits(const double *qubits, const size_t *lens, size_t count, int dtype,
                            qs_register **out) {
  if (count > 0 && (!qubits || !lens)) return fail(QS_ERR_ARGUMENT, 0, 0, "NULL input");
  if (!out) return fail(QS_ERR_ARGUMENT, 0, 0, "out is NULL");
  // is_qubit / is_one (src/register.rs:104-126), first failing index wins.
  size_t off = 0;
  for (size_t i = 0; i < count; ++i) {
    bool ok = false;
    if (lens[i] == 2) {
      double total = -1.0;
      for (size_t j = 0; j < 2; ++j) {
        const double re = qubits[2 * (off + j)], im = qubits[2 * (off + j) + 1];
        total += re * re + im * im;
      }
      ok = std::fabs(total) < 0.000001;
    }
    if (!ok) return fail(QS_INVALID_QUBIT, i, 0, "not a qubit");
    off += lens[i];
  }
  int rc = alloc_register(count, dtype, out);
  if (rc) return rc;
  qs_register *r = *out;
  double *d_q = nullptr;
  hipError_t e = hipMalloc((void **)&d_q, std::max<size_t>(1, count) * 4 * sizeof(double));
  if (e == hipSuccess && count)
    e = hipMemcpy(d_q, qubits, count * 4 * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = by_dtype(r, [&](auto z) {
      return qs::launch_init_product<decltype(z)>(r->psi, r->n, d_q, r->stream);
    });
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if (d_q) (void)hipFree(d_q);
  if (e != hipSuccess) {
    qs_register_free(r);
    *out = nullptr;
    return hip_fail(e, "qs_register_from_qubits");
  }
  return QS_OK;
}

int qs_register_clone(qs_register *src, qs_register **out) {
  if (!src) return fail(QS_ERR_ARGUMENT, 0, 0, "src is NULL");
  int rc = sync(src);
  if (rc) return rc;
  rc = alloc_register((size_t)src->n, src->dtype, out);
  if (rc) return rc;
  qs_register *r = *out;
  r->fusion = src->fusion;
  hipError_t e =
      hipMemcpyAsync(r->psi, src->psi, src->state_bytes(), hipMemcpyDeviceToDevice, r->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess) {
    qs_register_free(r);
    *out = nullptr;
    return hip_fail(e, "qs_register_clone");
  }
  return QS_OK;
}

void qs_register_free(qs_register *r) {
  if (!r) return;
  if (r->stream) (void)hipStreamSynchronize(r->stream);
  for (const auto &m : r->marks) {
    (void)hipEventDestroy(m.a);
    (void)hipEventDestroy(m.b);
  }
  for (hipEvent_t e : r->ev_pool) (void)hipEventDestroy(e);
  if (r->open_ev) (void)hipEventDestroy(r->open_ev);
  if (r->engine) qs::sweep_engine_free(r->engine);
  if (r->psi) (void)hipFree(r->psi);
  if (r->scratch) (void)hipFree(r->scratch);
  if (r->d_red) (void)hipFree(r->d_red);
  if (r->d_big) (void)hipFree(r->d_big);
  if (r->d_upload) (void)hipFree(r->d_upload);
  if (r->h_red) (void)hipHostFree(r->h_red);
  if (r->d_ring) (void)hipFree(r->d_ring);
  if (r->h_ring) (void)hipHostFree(r->h_ring);
  if (r->stream) (void)hipStreamDestroy(r->stream);
  delete r;
}

size_t qs_register_size(const qs_register *r) { return r ? (size_t)r->n : 0; }

int qs_register_dtype(const qs_register *r) { return r ? r->dtype : -1; }

int qs_apply(qs_register *r, const double *matrix, size_t rows, size_t cols,
             const size_t *targets, size_t k) {
  if (!r) return fail(QS_ERR_ARGUMENT, 0, 0, "register is NULL");
  if (!matrix || (k > 0 && !targets)) return fail(QS_ERR_ARGUMENT, 0, 0, "NULL input");
  int rc = validate_dims(rows, cols, k);
  if (rc) return rc;
  rc = validate_targets((size_t)r->n, targets, k);
  if (rc) return rc;
  return enqueue_split(r, classify(matrix, targets, k, 0, 0));
}

int qs_apply_controlled(qs_register *r, const size_t *controls, const unsigned char *cvals,
                        size_t c, const double *matrix, const size_t *targets, size_t k) {
  if (!r) return fail(QS_ERR_ARGUMENT, 0, 0, "register is NULL");
  if (!matrix || (k > 0 && !targets) || (c > 0 && !controls))
    return fail(QS_ERR_ARGUMENT, 0, 0, "NULL input");
  if (k >= 20) return fail(QS_INVALID_DIMENSIONS, k, k, "too many active targets");
  std::vector<size_t> all(targets, targets + k);
  all.insert(all.end(), controls, controls + c);
  int rc = validate_targets((size_t)r->n, all.data(), all.size());
  if (rc) return rc;
  uint64_t cmask = 0, cval = 0;
  for (size_t i = 0; i < c; ++i) {
    cmask |= 1ull << controls[i];
    if (!cvals || cvals[i]) cval |= 1ull << controls[i];
  }
  return enqueue_split(r, classify(matrix, targets, k, cmask, cval));
}

int qs_apply_diagonal(qs_register *r, const double *diag, const size_t *targets, size_t k) {
  if (!r) return fail(QS_ERR_ARGUMENT, 0, 0, "register is NULL");
  if (!diag || (k > 0 && !targets)) return fail(QS_ERR_ARGUMENT, 0, 0, "NULL input");
  if (k >= 30) return fail(QS_INVALID_DIMENSIONS, k, k, "diagonal too large");
  int rc = validate_targets((size_t)r->n, targets, k);
  if (rc) return rc;
  const size_t D = (size_t)1 << k;
  // Build the gate directly (a dense 2^k x 2^k matrix would be wasteful).
  Gate g;
  std::vector<cd> dv(D);
  std::memcpy(dv.data(), diag, D * sizeof(cd));
  size_t nonunit = 0;
  for (size_t i = 0; i < D; ++i)
    if (!is_one(dv[i])) ++nonunit;
  if (nonunit == 0) return enqueue(r, Gate());
  // controls: bits on which every entry with that bit 0 is 1
  std::vector<int> ctrl(k, 0);
  for (size_t j = 0; j < k; ++j) {
    bool ok = true;
    for (size_t i = 0; i < D && ok; ++i)
      if (!((i >> j) & 1) && !is_one(dv[i])) ok = false;
    ctrl[j] = ok;
  }
  size_t cbits = 0;
  for (size_t j = 0; j < k; ++j)
    if (ctrl[j]) {
      cbits |= (size_t)1 << j;
      g.cmask |= 1ull << targets[j];
      g.cval |= 1ull << targets[j];
    }
  std::vector<size_t> actj;
  for (size_t j = 0; j < k; ++j)
    if (!ctrl[j]) actj.push_back(j);
  g.na = (int)actj.size();
  for (int a = 0; a < g.na; ++a) g.act[a] = (int)targets[actj[a]];
  const size_t d = (size_t)1 << g.na;
  std::vector<cd> red(d);
  for (size_t ri = 0; ri < d; ++ri) {
    size_t x = cbits;
    for (int a = 0; a < g.na; ++a)
      if ((ri >> a) & 1) x |= (size_t)1 << actj[a];
    red[ri] = dv[x];
  }
  nonunit = 0;
  for (size_t i = 0; i < d; ++i)
    if (!is_one(red[i])) ++nonunit;
  if (nonunit <= 16 && (nonunit < d || g.na == 0)) {
    g.kind = Gate::kSparse;
    uint64_t amask = 0;
    for (int a = 0; a < g.na; ++a) amask |= 1ull << g.act[a];
    g.smask = amask | g.cmask;
    for (size_t ri = 0; ri < d; ++ri) {
      if (is_one(red[ri])) continue;
      uint64_t pat = g.cval;
      for (int a = 0; a < g.na; ++a)
        if ((ri >> a) & 1) pat |= 1ull << g.act[a];
      g.spat.push_back(pat);
      g.m.push_back(red[ri]);
    }
  } else {
    g.kind = Gate::kDiag;
    g.m = std::move(red);
  }
  return enqueue(r, std::move(g));
}

int qs_apply_all(qs_register *r, const double *matrix, size_t rows, size_t cols, size_t arity) {
  if (!r) return fail(QS_ERR_ARGUMENT, 0, 0, "register is NULL");
  if (arity != 1) return fail(QS_INVALID_ARITY, arity, 0, "apply_all needs a unary operation");
  if (!matrix) return fail(QS_ERR_ARGUMENT, 0, 0, "NULL input");
  int rc = validate_dims(rows, cols, 1);
  if (rc) return rc;
  for (int q = 0; q < r->n; ++q) {
    const size_t t = (size_t)q;
    rc = enqueue(r, classify(matrix, &t, 1, 0, 0));
    if (rc) return rc;
  }
  return QS_OK;
}

int qs_measure(qs_register *r, size_t target, double uniform, int *result) {
  if (!r || !result) return fail(QS_ERR_ARGUMENT, 0, 0, "NULL argument");
  if (target >= (size_t)r->n) return fail(QS_INVALID_TARGET, target, 0, "target out of range");
  // a collapse left by the previous measure stays pending (the split fuses
  // it) unless gates were applied since: then it goes first, then the gates
  if (!r->pending.empty()) {
    int rc = flush(r);
    if (rc) return rc;
  }
  return measure_dispatch(r, target, uniform, result);
}

int qs_measure_deferred(qs_register *r, size_t target, double uniform, size_t slot) {
  if (!r) return fail(QS_ERR_ARGUMENT, 0, 0, "NULL argument");
  if (target >= (size_t)r->n) return fail(QS_INVALID_TARGET, target, 0, "target out of range");
  if (slot >= QS_MEASURE_SLOTS) return fail(QS_ERR_ARGUMENT, slot, 0, "slot out of range");
  int rc = flush(r);
  if (rc) return rc;
  double *res_dev = r->d_red + 4 * qs::kReduceBlocks;
  auto go = [&](auto z) {
    using Rt = decltype(z);
    return qs::launch_measure_deferred<Rt>(r->psi, r->n, (int)target, uniform, r->d_red, res_dev,
                                           res_dev + 2, slot_base(r) + slot, r->stream);
  };
  QS_HIP(r->dtype == QS_C128 ? go(0.0) : go(0.0f));
  return QS_OK;
}

int qs_read_slots(qs_register *r, size_t first, size_t count, unsigned char *out) {
  if (!r || (count > 0 && !out)) return fail(QS_ERR_ARGUMENT, 0, 0, "NULL argument");
  if (first > QS_MEASURE_SLOTS || count > QS_MEASURE_SLOTS - first)
    return fail(QS_ERR_ARGUMENT, first, count, "slot range out of range");
  int rc = flush(r);
  if (rc) return rc;
  if (count == 0) return QS_OK;
  QS_HIP(hipMemcpyAsync(out, slot_base(r) + first, count, hipMemcpyDeviceToHost, r->stream));
  QS_HIP(hipStreamSynchronize(r->stream));
  return QS_OK;
}

int qs_measure_sequence(qs_register *r, const size_t *order, size_t count,
                        const double *uniforms, uint64_t *index) {
  if (!r || !index || (count > 0 && (!order || !uniforms)))
    return fail(QS_ERR_ARGUMENT, 0, 0, "NULL argument");
  int rc = flush(r);
  if (rc) return rc;
  *index = 0;
  if (count == 0) return QS_OK;
  // A target out of range fails at its position, after the measurements
  // before it, as a loop of measure calls would.
  size_t valid = count;
  bool distinct = true;
  uint64_t seen = 0;
  for (size_t j = 0; j < count; ++j) {
    if (order[j] >= (size_t)r->n) {
      valid = j;
      break;
    }
    if ((seen >> order[j]) & 1ull) distinct = false;
    seen |= 1ull << order[j];
  }
  int fallback = 1;
  uint64_t idx = 0;
  if (valid > 0 && distinct) {
    rc = r->dtype == QS_C128 ? measure_fast<double>(r, order, valid, uniforms, &idx, &fallback)
                             : measure_fast<float>(r, order, valid, uniforms, &idx, &fallback);
    if (rc) return rc;
  }
  if (fallback) {
    idx = 0;
    for (size_t j = 0; j < valid; ++j) {
      int res = 0;
      rc = measure_dispatch(r, order[j], uniforms[j], &res);
      if (rc) return rc;
      if (res) idx |= 1ull << order[j];
    }
  }
  *index = idx;
  if (valid < count) return fail(QS_INVALID_TARGET, order[valid], 0, "target out of range");
  return QS_OK;
}

int qs_sample_cdf(qs_register *r, const double *uniforms, size_t count, uint64_t *indices) {
  if (!r || (count > 0 && (!uniforms || !indices)))
    return fail(QS_ERR_ARGUMENT, 0, 0, "NULL argument");
  int rc = sync(r);
  if (rc) return rc;
  if (count == 0) return QS_OK;
  const uint64_t chunk = std::min<uint64_t>(r->len, 1ull << 16);
  const uint64_t nchunks = (r->len + chunk - 1) / chunk;
  std::vector<double> sums(nchunks);
  double *d_sums = nullptr;
  QS_HIP(ensure_big(r, std::max<uint64_t>(nchunks, chunk) * sizeof(double), &d_sums));
  hipError_t e = by_dtype(r, [&](auto z) {
    return qs::launch_chunk_sums<decltype(z)>(r->psi, r->n, chunk, d_sums, r->stream);
  });
  if (e == hipSuccess)
    e = hipMemcpyAsync(sums.data(), d_sums, nchunks * sizeof(double), hipMemcpyDeviceToHost,
                       r->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess) return hip_fail(e, "qs_sample_cdf");
  std::vector<double> prefix(nchunks + 1, 0.0);
  for (uint64_t c = 0; c < nchunks; ++c) prefix[c + 1] = prefix[c] + sums[c];
  const double total = prefix[nchunks];
  std::map<uint64_t, std::vector<double>> chunk_probs;
  for (size_t j = 0; j < count && e == hipSuccess; ++j) {
    const double tau = uniforms[j] * total;
    // first chunk whose running sum exceeds tau
    uint64_t c = (uint64_t)(std::upper_bound(prefix.begin() + 1, prefix.end(), tau) -
                            (prefix.begin() + 1));
    if (c >= nchunks) c = nchunks - 1;
    auto it = chunk_probs.find(c);
    if (it == chunk_probs.end()) {
      std::vector<double> pr(chunk);
      e = by_dtype(r, [&](auto z) {
        using R = decltype(z);
        const size_t ab = r->amp_bytes();
        return qs::launch_probabilities<R>((const char *)r->psi + c * chunk * ab,
                                           (int)(63 - __builtin_clzll(chunk)), d_sums, r->stream);
      });
      if (e == hipSuccess)
        e = hipMemcpyAsync(pr.data(), d_sums, chunk * sizeof(double), hipMemcpyDeviceToHost,
                           r->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
      it = chunk_probs.emplace(c, std::move(pr)).first;
    }
    const std::vector<double> &pr = it->second;
    double run = prefix[c];
    uint64_t x = c * chunk + chunk - 1, last_nz = x;
    bool found = false;
    for (uint64_t i = 0; i < chunk; ++i) {
      if (pr[i] > 0.0) last_nz = c * chunk + i;
      run += pr[i];
      if (run > tau) {
        x = c * chunk + i;
        found = true;
        break;
      }
    }
    indices[j] = found ? x : last_nz;
  }
  if (e != hipSuccess) return hip_fail(e, "qs_sample_cdf");
  return QS_OK;
}

int qs_get_state(qs_register *r, double *out) {
  if (!r || !out) return fail(QS_ERR_ARGUMENT, 0, 0, "NULL argument");
  int rc = flush(r);
  if (rc) return rc;
  if (r->dtype == QS_C128) {
    QS_HIP(hipMemcpyAsync(out, r->psi, r->state_bytes(), hipMemcpyDeviceToHost, r->stream));
    QS_HIP(hipStreamSynchronize(r->stream));
    return QS_OK;
  }
  // widened to complex128 through the work buffer, kReadoutAmps at a time
  const uint64_t step = std::min<uint64_t>(r->len, kReadoutAmps);
  double *d = nullptr;
  QS_HIP(ensure_big(r, step * 2 * sizeof(double), &d));
  for (uint64_t off = 0; off < r->len; off += step) {
    const uint64_t cnt = std::min<uint64_t>(step, r->len - off);
    hipError_t e = qs::launch_convert_to_c128<float>(
        d, (const char *)r->psi + off * r->amp_bytes(), cnt, r->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(out + 2 * off, d, cnt * 2 * sizeof(double), hipMemcpyDeviceToHost,
                         r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return hip_fail(e, "qs_get_state");
  }
  return QS_OK;
}

int qs_set_state(qs_register *r, const double *in) {
  if (!r || !in) return fail(QS_ERR_ARGUMENT, 0, 0, "NULL argument");
  int rc = sync(r);
  if (rc) return rc;
  if (r->dtype == QS_C128) {
    QS_HIP(hipMemcpyAsync(r->psi, in, r->state_bytes(), hipMemcpyHostToDevice, r->stream));
    QS_HIP(hipStreamSynchronize(r->stream));
    return QS_OK;
  }
  const uint64_t step = std::min<uint64_t>(r->len, kReadoutAmps);
  double *d = nullptr;
  QS_HIP(ensure_big(r, step * 2 * sizeof(double), &d));
  for (uint64_t off = 0; off < r->len; off += step) {
    const uint64_t cnt = std::min<uint64_t>(step, r->len - off);
    hipError_t e = hipMemcpyAsync(d, in + 2 * off, cnt * 2 * sizeof(double),
                                  hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess)
      e = qs::launch_convert_from_c128<float>((char *)r->psi + off * r->amp_bytes(), d, cnt,
                                              r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return hip_fail(e, "qs_set_state");
  }
  return QS_OK;
}

int qs_get_amplitudes(qs_register *r, const uint64_t *indices, size_t count, double *out) {
  if (!r || (count && (!indices || !out))) return fail(QS_ERR_ARGUMENT, 0, 0, "NULL argument");
  for (size_t i = 0; i < count; ++i)
    if (indices[i] >= r->len)
      return fail(QS_ERR_ARGUMENT, i, (size_t)indices[i], "qs_get_amplitudes: index >= 2^n");
  int rc = flush(r);
  if (rc || count == 0) return rc;
  // indices and amplitudes through the work buffer, `step` at a time
  const size_t step = std::min<size_t>(count, kReadoutAmps);
  double *d = nullptr;
  QS_HIP(ensure_big(r, step * (sizeof(uint64_t) + 2 * sizeof(double)), &d));
  uint64_t *d_idx = (uint64_t *)d;
  double *d_out = d + step;
  for (size_t off = 0; off < count; off += step) {
    const size_t cnt = std::min(step, count - off);
    hipError_t e = hipMemcpyAsync(d_idx, indices + off, cnt * sizeof(uint64_t),
                                  hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess)
      e = by_dtype(r, [&](auto z) {
        return qs::launch_gather_amps<decltype(z)>(d_out, r->psi, d_idx, cnt, r->stream);
      });
    if (e == hipSuccess)
      e = hipMemcpyAsync(out + 2 * off, d_out, cnt * 2 * sizeof(double), hipMemcpyDeviceToHost,
                         r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return hip_fail(e, "qs_get_amplitudes");
  }
  return QS_OK;
}

int qs_get_amplitudes_noalloc(qs_register *r, const uint64_t *indices, size_t count,
                              double *out) {
  if (!r || (count && (!indices || !out))) return fail(QS_ERR_ARGUMENT, 0, 0, "NULL argument");
  if (count > kNoAllocMaxIndices)
    return fail(QS_ERR_ARGUMENT, count, kNoAllocMaxIndices,
                "qs_get_amplitudes_noalloc: more than 65536 indices in one call");
  for (size_t i = 0; i < count; ++i)
    if (indices[i] >= r->len)
      return fail(QS_ERR_ARGUMENT, i, (size_t)indices[i],
                  "qs_get_amplitudes_noalloc: index >= 2^n");
  int rc = flush(r);
  if (rc || count == 0) return rc;
  char *base = nullptr;
  rc = noalloc_scratch(r->device, &base);
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(g_noalloc_mu[r->device]);
  const size_t ib = count * sizeof(uint64_t), ob = count * 2 * sizeof(double);
  uint64_t *d_idx = reinterpret_cast<uint64_t *>(base);
  double *d_out = reinterpret_cast<double *>(base + kNoAllocMaxIndices * sizeof(uint64_t));
  hipError_t e = hipMemcpyAsync(d_idx, indices, ib, hipMemcpyHostToDevice, r->stream);
  if (e == hipSuccess)
    e = by_dtype(r, [&](auto z) {
      return qs::launch_gather_amps<decltype(z)>(d_out, r->psi, d_idx, count, r->stream);
    });
  if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, ob, hipMemcpyDeviceToHost, r->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess) return hip_fail(e, "qs_get_amplitudes_noalloc");
  return QS_OK;
}

int qs_norm_sqr(qs_register *r, double *out) {
  if (!r || !out) return fail(QS_ERR_ARGUMENT, 0, 0, "NULL argument");
  int rc = flush(r);
  if (rc) return rc;
  hipError_t e = by_dtype(r, [&](auto z) { return reduce_norm<decltype(z)>(r, out); });
  if (e != hipSuccess) return hip_fail(e, "qs_norm_sqr");
  return QS_OK;
}

int qs_probabilities(qs_register *r, double *out) {
  if (!r || !out) return fail(QS_ERR_ARGUMENT, 0, 0, "NULL argument");
  int rc = flush(r);
  if (rc) return rc;
  // 2^m amplitudes at a time through the work buffer
  const int m = std::min(r->n, kReadoutBits);
  const uint64_t step = 1ull << m;
  double *d = nullptr;
  QS_HIP(ensure_big(r, step * sizeof(double), &d));
  for (uint64_t off = 0; off < r->len; off += step) {
    hipError_t e = by_dtype(r, [&](auto z) {
      return qs::launch_probabilities<decltype(z)>((const char *)r->psi + off * r->amp_bytes(),
                                                   m, d, r->stream);
    });
    if (e == hipSuccess)
      e = hipMemcpyAsync(out + off, d, step * sizeof(double), hipMemcpyDeviceToHost, r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return hip_fail(e, "qs_probabilities");
  }
  return QS_OK;
}

int qs_inner(qs_register *a, qs_register *b, double *re, double *im) {
  if (!a || !b || !re || !im) return fail(QS_ERR_ARGUMENT, 0, 0, "NULL argument");
  if (a->n != b->n)
    return fail(QS_ERR_ARGUMENT, (size_t)a->n, (size_t)b->n, "registers differ in size");
  int rc = sync(a);
  if (rc) return rc;
  rc = flush(b);
  if (rc) return rc;
  QS_HIP(hipStreamSynchronize(b->stream));
  double *res = a->d_red + 4 * qs::kReduceBlocks;
  hipError_t e = qs::launch_inner(a->psi, a->dtype == QS_C64, b->psi, b->dtype == QS_C64, a->n,
                                  a->d_red, res, a->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(a->h_red, res, 2 * sizeof(double), hipMemcpyDeviceToHost, a->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(a->stream);
  if (e != hipSuccess) return hip_fail(e, "qs_inner");
  *re = a->h_red[0];
  *im = a->h_red[1];
  return QS_OK;
}

int qs_equal(qs_register *a, qs_register *b, double tol, int *equal) {
  if (!a || !b || !equal) return fail(QS_ERR_ARGUMENT, 0, 0, "NULL argument");
  if (a->n != b->n || a->dtype != b->dtype)
    return fail(QS_ERR_ARGUMENT, (size_t)a->n, (size_t)b->n, "registers differ in size or dtype");
  int rc = sync(a);
  if (rc) return rc;
  rc = flush(b);
  if (rc) return rc;
  QS_HIP(hipStreamSynchronize(b->stream));
  double *res = a->d_red + 4 * qs::kReduceBlocks;
  hipError_t e = by_dtype(a, [&](auto z) {
    return qs::launch_neq_count<decltype(z)>(a->psi, b->psi, a->n, tol, a->d_red, res,
                                             a->stream);
  });
  if (e == hipSuccess)
    e = hipMemcpyAsync(a->h_red, res, sizeof(double), hipMemcpyDeviceToHost, a->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(a->stream);
  if (e != hipSuccess) return hip_fail(e, "qs_equal");
  *equal = a->h_red[0] == 0.0;
  return QS_OK;
}

int qs_sync(qs_register *r) {
  if (!r) return fail(QS_ERR_ARGUMENT, 0, 0, "register is NULL");
  return sync(r);
}

int qs_set_fusion(qs_register *r, int enabled) {
  if (!r) return fail(QS_ERR_ARGUMENT, 0, 0, "register is NULL");
  int rc = flush(r);
  if (rc) return rc;
  r->fusion = enabled != 0;
  return QS_OK;
}

int qs_stats(qs_register *r, uint64_t *sweeps, uint64_t *gate_kernels, uint64_t *gates,
             int reset) {
  if (!r) return fail(QS_ERR_ARGUMENT, 0, 0, "register is NULL");
  int rc = flush(r);
  if (rc) return rc;
  if (sweeps) *sweeps = r->st_sweeps;
  if (gate_kernels) *gate_kernels = r->st_kernels;
  if (gates) *gates = r->st_gates;
  if (reset) r->st_sweeps = r->st_kernels = r->st_gates = 0;
  return QS_OK;
}

int qs_stats_sweep_ops(qs_register *r, uint64_t *ops, uint64_t *merged, int reset) {
  if (!r) return fail(QS_ERR_ARGUMENT, 0, 0, "register is NULL");
  int rc = flush(r);
  if (rc) return rc;
  if (ops) *ops = r->st_sweep_ops;
  if (merged) *merged = r->st_merged;
  if (reset) r->st_sweep_ops = r->st_merged = 0;
  return QS_OK;
}

int qs_timing_enable(qs_register *r, int enable) {
  if (!r) return fail(QS_ERR_ARGUMENT, 0, 0, "register is NULL");
  int rc = flush(r);
  if (rc) return rc;
  r->timing = enable != 0;
  return QS_OK;
}

int qs_timing_read(qs_register *r, double *ms, uint64_t *launches, int reset) {
  if (!r) return fail(QS_ERR_ARGUMENT, 0, 0, "register is NULL");
  int rc = flush(r);
  if (rc) return rc;
  QS_HIP(collect_marks(r));
  for (int k = 0; k < QS_K_COUNT; ++k) {
    if (ms) ms[k] = r->t_ms[k];
    if (launches) launches[k] = r->t_n[k];
    if (reset) {
      r->t_ms[k] = 0.0;
      r->t_n[k] = 0;
    }
  }
  return QS_OK;
}

}  // extern "C"
