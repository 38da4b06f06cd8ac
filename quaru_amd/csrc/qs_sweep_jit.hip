// qs_sweep_jit.hip -- specialised sweep kernels compiled at run time
// (interface and modes in qs_sweep_jit.h).
//
// The kernel source is the sweep kernel header itself (embedded at build
// time) plus a generated body: the op functions of the sweep's code
// sequence, called in order with their codes as template arguments.  The
// compile runs on two host threads with hipRTC for gfx950, with the same
// code-generation flags as the library (no packed FP32).
#include "qs_sweep_jit.h"

#include <hip/hiprtc.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <map>
#include <mutex>
#include <sstream>
#include <thread>

#include <unistd.h>

#include "../../include/quaru_hip.h"
#include "qs_sweep_kernel.h"

namespace qs {
namespace sweep {
namespace {

const char kKernelHeader[] =
#include "qs_sweep_kernel_src.inc"
    ;

// Specialised kernels kept per process (a long run over ever-new circuits
// stops compiling there and stays on the generic kernel).
constexpr size_t kMaxKernels = 4096;

struct Entry {
  int state = 0;  // 0 queued or compiling, 1 compiled, 2 failed
  std::vector<char> code;
  std::map<int, hipFunction_t> fn;  // per device
};

// Process-wide state, never destroyed: compile threads may still wait on it
// while the process exits.
struct Jit {
  std::mutex mu;
  std::condition_variable work, done;
  std::map<std::string, Entry> table;
  std::deque<std::pair<std::string, std::string>> queue;  // signature, source
  int busy = 0;  // queued + compiling
  bool started = false, stopping = false;
  int mode = 1;
  std::atomic<uint64_t> compiled{0}, failed{0}, launches{0};
};

Jit &jit() {
  static Jit *j = [] {
    Jit *p = new Jit();
    const char *e = std::getenv("QS_SWEEP_JIT");
    if (e) p->mode = std::max(0, std::min(2, std::atoi(e)));
    return p;
  }();
  return *j;
}

bool compile(const std::string &src, std::vector<char> &code, std::string &log) {
  hiprtcProgram prog;
  const char *headers[] = {kKernelHeader};
  const char *names[] = {"qs_sweep_kernel.h"};
  if (hiprtcCreateProgram(&prog, src.c_str(), "qs_sweep_jit.hip", 1, headers, names) !=
      HIPRTC_SUCCESS)
    return false;
  const char *opts[] = {"--offload-arch=gfx950", "-O3",           "-std=c++17", "-Xclang",
                        "-target-feature",       "-Xclang",       "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  size_t n = 0;
  if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
    log.resize(n);
    hiprtcGetProgramLog(prog, &log[0]);
  }
  bool ok = rc == HIPRTC_SUCCESS;
  if (ok) {
    size_t sz = 0;
    ok = hiprtcGetCodeSize(prog, &sz) == HIPRTC_SUCCESS && sz > 0;
    if (ok) {
      code.resize(sz);
      ok = hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
    }
  }
  hiprtcDestroyProgram(&prog);
  return ok;
}

void worker() {
  Jit &j = jit();
  for (;;) {
    std::pair<std::string, std::string> item;
    {
      std::unique_lock<std::mutex> lk(j.mu);
      j.work.wait(lk, [&] { return !j.queue.empty() || j.stopping; });
      if (j.stopping && j.queue.empty()) return;
      item = std::move(j.queue.front());
      j.queue.pop_front();
    }
    std::vector<char> code;
    std::string log;
    const bool ok = compile(item.second, code, log);
    {
      std::lock_guard<std::mutex> lk(j.mu);
      Entry &e = j.table[item.first];
      e.state = ok ? 1 : 2;
      e.code.swap(code);
      --j.busy;
      if (ok)
        ++j.compiled;
      else {
        ++j.failed;
        std::fprintf(stderr, "quaru-hip: sweep kernel compile failed, generic kernel kept\n%s\n",
                     log.substr(0, 4000).c_str());
      }
    }
    j.done.notify_all();
  }
}

// At exit (an on_exit handler, so the exit status is known): no new
// compiles, queued ones are dropped, and the running ones (at most one per
// compile thread, about 0.5 s each) get 5 s (QS_SWEEP_JIT_EXIT_WAIT) to
// finish before hipRTC's own
// exit-time teardown, which start_workers orders after this handler.  A
// compile still inside hipRTC after that would be torn down under it (a
// crash or a hang), so the rest of the exit is skipped instead: stdio is
// flushed and the process ends at once with its exit status.
void shutdown(int status, void *) {
  Jit &j = jit();
  std::unique_lock<std::mutex> lk(j.mu);
  j.stopping = true;
  j.busy -= (int)j.queue.size();
  j.queue.clear();
  j.work.notify_all();
  static const char *w = std::getenv("QS_SWEEP_JIT_EXIT_WAIT");  // seconds
  const int wait_ms = w ? std::max(0, (int)(std::atof(w) * 1000.0)) : 5000;
  if (!j.done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return j.busy <= 0; })) {
    std::fflush(nullptr);
    _exit(status);
  }
}

std::string source(const std::vector<OpRec> &ops, bool c64, int A, int W);

// hipRTC and the compiler library behind it register parts of their
// exit-time teardown on first use, and exit handlers run in the reverse
// order of registration.  A handler registered before that first use runs
// after the teardown, while compile threads may still be inside
// hiprtcCompileProgram: the process then crashes or hangs at exit.  One
// compile of a sweep kernel on this thread before `shutdown` is registered
// puts that teardown first, so `shutdown` runs (and waits for the compile
// threads) before it.  The kernel is discarded.
void warm_up() {
  std::vector<OpRec> ops(3);
  ops[0].code = CODE_U1;
  ops[1].code = CODE_C1 + kMaxA;  // a lane-bit form
  ops[2].code = CODE_NEG;
  std::vector<char> code;
  std::string log;
  (void)compile(source(ops, false, 4, kWaveBits), code, log);
}

// hipRTC runs one compile at a time (8 specialised kernels on 8 threads:
// 5.1 s wall for 5.3 s of CPU), so two threads keep it busy; at exit at
// most two compiles are still running.
constexpr int kCompileThreads = 2;

void start_workers(Jit &j) {
  if (j.started) return;
  j.started = true;
  warm_up();
  on_exit(shutdown, nullptr);
  for (int i = 0; i < kCompileThreads; ++i) std::thread(worker).detach();
}

std::string source(const std::vector<OpRec> &ops, bool c64, int A, int W) {
  std::ostringstream o;
  const char *R = c64 ? "float" : "double";
  const int waves = (A <= 4 || c64) ? 3 : 1;
  o << "#include \"qs_sweep_kernel.h\"\n"
       "using namespace qs::sweep;\n"
       "using R = "
    << R << ";\nusing I = uint32_t;\nconstexpr int A = " << A << ", W = " << W
    << ";\n"
       "struct Ops {\n"
       "  __device__ __forceinline__ void operator()(Amp<R, 1 << A> &a, const Blob<R> &b,\n"
       "                                             const Ctx<I> &c, const Regs<A, I> &g) const {\n"
       "    const QS_CONST OpRec *ops = b.ops;\n"
       "    const QS_CONST R *m = b.pool;\n"
       "    uint32_t code = 0;\n";
  for (size_t i = 0; i < ops.size(); ++i) {
    const OpRec &op = ops[i];
    const bool uni = op.umask != 0;
    if (op.code == CODE_NONE && (!uni || op.code_fail == CODE_NONE)) continue;
    if (uni || op.bmask) {
      o << "    code = op_setup<R, I>(ops[" << i << "], b.pool, c, m);\n";
      if (uni && op.code_fail != op.code) {
        if (op.code != CODE_NONE)
          o << "    if (code == " << op.code << "u) apply_code<" << op.code << "u, A, W>(a, ops[" << i
            << "], m, c, g);\n";
        if (op.code_fail != CODE_NONE)
          o << "    if (code == " << op.code_fail << "u) apply_code<" << op.code_fail
            << "u, A, W>(a, ops[" << i << "], m, c, g);\n";
      } else {
        o << "    apply_code<" << op.code << "u, A, W>(a, ops[" << i << "], m, c, g);\n";
      }
    } else {
      o << "    m = b.pool + ops[" << i << "].moff;\n";
      o << "    apply_code<" << op.code << "u, A, W>(a, ops[" << i << "], m, c, g);\n";
    }
  }
  o << "    (void)code;\n    (void)m;\n  }\n};\n"
       "extern \"C\" __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu("
    << waves << "))) void qs_sweep_jit(SV2<R>::T *__restrict__ psi, const char *__restrict__ blob_g,\n"
       "                                  uint32_t xcd_chunk) {\n"
       "  sweep_tile<R, A, I, W>(psi, blob_g, xcd_chunk, Ops());\n}\n";
  return o.str();
}

}  // namespace

std::string jit_signature(const std::vector<OpRec> &ops, bool c64, int A, int W) {
  std::string s = c64 ? "f" : "d";
  s += std::to_string(A) + "w" + std::to_string(W) + ":";
  for (const OpRec &op : ops) {
    s += std::to_string(op.code);
    if (op.umask) s += "u" + std::to_string(op.code_fail);
    if (op.bmask) s += "b";
    s += ",";
  }
  return s;
}

hipFunction_t jit_kernel(const std::string &sig, const std::vector<OpRec> &ops, bool c64, int A,
                         int W) {
  Jit &j = jit();
  std::unique_lock<std::mutex> lk(j.mu);
  if (j.mode == 0 || j.stopping) return nullptr;
  auto it = j.table.find(sig);
  if (it == j.table.end()) {
    if (j.table.size() >= kMaxKernels) return nullptr;  // the generic kernel from here on
    start_workers(j);
    it = j.table.emplace(sig, Entry()).first;
    ++j.busy;
    j.queue.emplace_back(sig, source(ops, c64, A, W));
    j.work.notify_one();
  }
  if (j.mode == 2) j.done.wait(lk, [&] { return it->second.state != 0; });
  Entry &e = it->second;
  if (e.state != 1) return nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  auto f = e.fn.find(dev);
  if (f != e.fn.end()) return f->second;
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  if (hipModuleLoadData(&mod, e.code.data()) != hipSuccess ||
      hipModuleGetFunction(&fn, mod, "qs_sweep_jit") != hipSuccess) {
    (void)hipGetLastError();
    e.state = 2;
    ++j.failed;
    return nullptr;
  }
  e.fn[dev] = fn;  // the module stays loaded for the process
  return fn;
}

int jit_mode() {
  Jit &j = jit();
  std::lock_guard<std::mutex> lk(j.mu);
  return j.mode;
}

int jit_set_mode(int mode) {
  Jit &j = jit();
  std::lock_guard<std::mutex> lk(j.mu);
  const int prev = j.mode;
  j.mode = std::max(0, std::min(2, mode));
  return prev;
}

void jit_wait() {
  Jit &j = jit();
  std::unique_lock<std::mutex> lk(j.mu);
  j.done.wait(lk, [&] { return j.busy <= 0; });
}

void jit_stats(uint64_t *compiled, uint64_t *failed, uint64_t *launches) {
  Jit &j = jit();
  if (compiled) *compiled = j.compiled.load();
  if (failed) *failed = j.failed.load();
  if (launches) *launches = j.launches.load();
}

void jit_count_launch() { ++jit().launches; }

}  // namespace sweep
}  // namespace qs

extern "C" {

int qs_sweep_jit_mode(int mode) {
  if (mode < 0) return qs::sweep::jit_mode();
  return qs::sweep::jit_set_mode(mode);
}

int qs_sweep_jit_wait(void) {
  qs::sweep::jit_wait();
  return QS_OK;
}

int qs_sweep_jit_stats(uint64_t *compiled, uint64_t *failed, uint64_t *launches) {
  qs::sweep::jit_stats(compiled, failed, launches);
  return QS_OK;
}

}  // extern "C"
