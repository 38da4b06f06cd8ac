// qs_sweep_jit.hip -- specialised sweep kernels compiled at run time
// (interface and modes in qs_sweep_jit.h).
//
// The kernel source is the sweep kernel header itself (embedded at build
// time) plus a generated body: the op functions of the sweep's code
// sequence, called in order with their codes as template arguments.  Each
// kernel is compiled with hipRTC for gfx950, with the library's
// code-generation flags (no packed FP32), by a helper process
// (bin/qs_sweep_compile beside the library) that host threads start.
#include "qs_sweep_jit.h"

#include <dlfcn.h>
#include <fcntl.h>
#include <poll.h>
#include <signal.h>
#include <spawn.h>
#include <sys/wait.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <vector>
#include <mutex>
#include <sstream>
#include <thread>

#include "../../include/quaru_hip.h"
#include "qs_sweep_kernel.h"

extern char **environ;

namespace qs {
namespace sweep {
namespace {

const char kKernelHeader[] =
#include "qs_sweep_kernel_src.inc"
    ;

// hipRTC options: the library's code-generation flags (no packed FP32).
const char *const kCompileOptions[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang",
                                       "-target-feature",       "-Xclang", "-packed-fp32-ops"};

// Specialised kernels kept per process (a long run over ever-new circuits
// stops compiling there and stays on the generic kernel); a code object is
// 20-30 KB, so the table holds at most about 120 MB of them.
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
  std::vector<pid_t> children;  // helpers started and not yet reaped
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

// The compiler helper: bin/qs_sweep_compile beside the library (or
// QS_SWEEP_COMPILER).  Empty if it cannot be found.
const std::string &helper_path() {
  static const std::string path = [] {
    const char *e = std::getenv("QS_SWEEP_COMPILER");
    if (e && *e) return access(e, X_OK) == 0 ? std::string(e) : std::string();
    Dl_info info;
    if (!dladdr((void *)&helper_path, &info) || !info.dli_fname) return std::string();
    std::string lib = info.dli_fname;
    const size_t slash = lib.rfind('/');
    const std::string dir = slash == std::string::npos ? std::string(".") : lib.substr(0, slash);
    const std::string p = dir + "/bin/qs_sweep_compile";
    return access(p.c_str(), X_OK) == 0 ? p : std::string();
  }();
  return path;
}

void forget_child(pid_t pid);
void track_child(pid_t pid);

// Compile `src` in a child process running the helper: the header and the
// source go in through a pipe, the code object comes back through another,
// and the compiler log through a third.  hipRTC itself never loads into
// this process.
bool compile(const std::string &src, std::vector<char> &code, std::string &log) {
  const std::string &helper = helper_path();
  if (helper.empty()) {
    log = "compiler helper bin/qs_sweep_compile not found beside the library";
    return false;
  }
  int in[2], out[2], err[2];
  if (pipe2(in, O_CLOEXEC)) return false;
  if (pipe2(out, O_CLOEXEC)) {
    close(in[0]);
    close(in[1]);
    return false;
  }
  if (pipe2(err, O_CLOEXEC)) {
    close(in[0]);
    close(in[1]);
    close(out[0]);
    close(out[1]);
    return false;
  }
  posix_spawn_file_actions_t fa;
  posix_spawn_file_actions_init(&fa);
  posix_spawn_file_actions_adddup2(&fa, in[0], 0);
  posix_spawn_file_actions_adddup2(&fa, out[1], 1);
  posix_spawn_file_actions_adddup2(&fa, err[1], 2);
  // the helper inherits no other descriptor of this process (a device file
  // among them)
  posix_spawn_file_actions_addclosefrom_np(&fa, 3);
  posix_spawnattr_t at;
  posix_spawnattr_init(&at);
  sigset_t none;
  sigemptyset(&none);
  posix_spawnattr_setsigmask(&at, &none);
  posix_spawnattr_setflags(&at, POSIX_SPAWN_SETSIGMASK);
  std::vector<char *> argv;
  argv.push_back(const_cast<char *>(helper.c_str()));
  for (const char *o : kCompileOptions) argv.push_back(const_cast<char *>(o));
  argv.push_back(nullptr);
  pid_t pid = 0;
  int rc;
  {
    // registered under the lock, so the exit handler either sees the child
    // or runs before it exists
    std::lock_guard<std::mutex> lk(jit().mu);
    rc = posix_spawn(&pid, helper.c_str(), &fa, &at, argv.data(), environ);
    if (rc == 0) track_child(pid);
  }
  posix_spawn_file_actions_destroy(&fa);
  posix_spawnattr_destroy(&at);
  close(in[0]);
  close(out[1]);
  close(err[1]);
  if (rc != 0) {
    close(in[1]);
    close(out[0]);
    close(err[0]);
    log = "cannot start " + helper;
    return false;
  }
  // input: header length, header, source (the helper reads all of it before
  // it writes anything)
  std::string msg(8, '\0');
  const uint64_t hl = sizeof(kKernelHeader) - 1;
  std::memcpy(&msg[0], &hl, 8);
  msg.append(kKernelHeader, hl);
  msg += src;
  bool wrote = true;
  for (size_t off = 0; off < msg.size();) {
    const ssize_t w = write(in[1], msg.data() + off, msg.size() - off);
    if (w < 0 && errno == EINTR) continue;
    if (w <= 0) {
      wrote = false;
      break;
    }
    off += (size_t)w;
  }
  close(in[1]);
  // output and log, read together so that neither pipe fills up
  pollfd fds[2] = {{out[0], POLLIN, 0}, {err[0], POLLIN, 0}};
  int open_fds = 2;
  char buf[1 << 16];
  while (open_fds > 0) {
    if (poll(fds, 2, -1) < 0) {
      if (errno == EINTR) continue;
      break;
    }
    for (int k = 0; k < 2; ++k) {
      if (fds[k].fd < 0 || !(fds[k].revents & (POLLIN | POLLHUP | POLLERR))) continue;
      const ssize_t r = read(fds[k].fd, buf, sizeof buf);
      if (r < 0 && errno == EINTR) continue;
      if (r <= 0) {
        close(fds[k].fd);
        fds[k].fd = -1;
        --open_fds;
      } else if (k == 0) {
        code.insert(code.end(), buf, buf + r);
      } else if (log.size() < 16384) {
        log.append(buf, (size_t)r);
      }
    }
  }
  for (int k = 0; k < 2; ++k)
    if (fds[k].fd >= 0) close(fds[k].fd);
  forget_child(pid);  // before the wait: its pid stays ours until reaped
  int status = 0;
  while (waitpid(pid, &status, 0) < 0 && errno == EINTR) {
  }
  return wrote && WIFEXITED(status) && WEXITSTATUS(status) == 0 && !code.empty();
}

void worker() {
  // a helper that dies early must not stop the process with SIGPIPE
  sigset_t pipe_set;
  sigemptyset(&pipe_set);
  sigaddset(&pipe_set, SIGPIPE);
  pthread_sigmask(SIG_BLOCK, &pipe_set, nullptr);
  Jit &j = jit();
  for (;;) {
    std::pair<std::string, std::string> item;
    {
      std::unique_lock<std::mutex> lk(j.mu);
      j.work.wait(lk, [&] { return !j.queue.empty() || j.stopping; });
      if (j.stopping) return;
      item = std::move(j.queue.front());
      j.queue.pop_front();
    }
    std::vector<char> code;
    std::string log;
    const bool ok = compile(item.second, code, log);
    {
      std::lock_guard<std::mutex> lk(j.mu);
      if (j.stopping) return;
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

void track_child(pid_t pid) { jit().children.push_back(pid); }

void forget_child(pid_t pid) {
  Jit &j = jit();
  std::lock_guard<std::mutex> lk(j.mu);
  j.children.erase(std::remove(j.children.begin(), j.children.end(), pid), j.children.end());
}

// At exit: no new compiles, and the running ones are stopped (SIGKILL to
// the helpers this library started and has not yet reaped, so none of
// those pids can belong to another process).  Nothing waits: the compile
// threads find their helper gone and return.
void shutdown() {
  Jit &j = jit();
  std::lock_guard<std::mutex> lk(j.mu);
  j.stopping = true;
  j.queue.clear();
  for (pid_t pid : j.children) kill(pid, SIGKILL);
  j.work.notify_all();
}

// Compiles run in child processes, so they run in parallel.
int compile_threads() {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  return (int)std::min(8u, hw);
}

void start_workers(Jit &j) {
  if (j.started) return;
  j.started = true;
  std::atexit(shutdown);
  for (int i = 0, n = compile_threads(); i < n; ++i) std::thread(worker).detach();
}

std::string source(const std::vector<OpRec> &ops, bool c64, int A, int W,
                   const std::vector<int> &rp) {
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
    const int rpi = rp.empty() ? -1 : rp[i];
    const bool uni = op.umask != 0;
    if (op.code == CODE_NONE && (!uni || op.code_fail == CODE_NONE)) continue;
    if (uni || op.bmask) {
      o << "    code = op_setup<R, I>(ops[" << i << "], b.pool, c, m);\n";
      if (uni && op.code_fail != op.code) {
        if (op.code != CODE_NONE)
          o << "    if (code == " << op.code << "u) apply_code<" << op.code << "u, A, W, " << rpi << ">(a, ops[" << i
            << "], m, c, g);\n";
        if (op.code_fail != CODE_NONE)
          o << "    if (code == " << op.code_fail << "u) apply_code<" << op.code_fail
            << "u, A, W, " << rpi << ">(a, ops[" << i << "], m, c, g);\n";
      } else {
        o << "    apply_code<" << op.code << "u, A, W, " << rpi << ">(a, ops[" << i << "], m, c, g);\n";
      }
    } else {
      o << "    m = b.pool + ops[" << i << "].moff;\n";
      o << "    apply_code<" << op.code << "u, A, W, " << rpi << ">(a, ops[" << i << "], m, c, g);\n";
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

std::vector<int> jit_register_patterns(const std::vector<OpRec> &ops, const uint64_t *gb, int A) {
  std::vector<int> rp(ops.size(), 0);
  for (size_t i = 0; i < ops.size(); ++i) {
    int rm = 0, rv = 0;
    for (int j = 0; j < A; ++j) {
      if (ops[i].cmask & gb[j]) rm |= 1 << j;
      if (ops[i].cval & gb[j]) rv |= 1 << j;
    }
    rp[i] = rm | (rv << 8);
  }
  return rp;
}

std::string jit_signature(const std::vector<OpRec> &ops, bool c64, int A, int W,
                          const std::vector<int> &rp) {
  std::string s = c64 ? "f" : "d";
  s += std::to_string(A) + "w" + std::to_string(W) + ":";
  for (size_t i = 0; i < ops.size(); ++i) {
    const OpRec &op = ops[i];
    s += std::to_string(op.code);
    if (op.umask) s += "u" + std::to_string(op.code_fail);
    if (op.bmask) s += "b";
    if (!rp.empty()) s += "r" + std::to_string(rp[i]);
    s += ",";
  }
  return s;
}

hipFunction_t jit_kernel(const std::string &sig, const std::vector<OpRec> &ops, bool c64, int A,
                         int W, const std::vector<int> &rp) {
  Jit &j = jit();
  std::unique_lock<std::mutex> lk(j.mu);
  if (j.mode == 0 || j.stopping) return nullptr;
  auto it = j.table.find(sig);
  if (it == j.table.end()) {
    if (j.table.size() >= kMaxKernels) return nullptr;  // the generic kernel from here on
    if (helper_path().empty()) {
      // no compiler helper: the generic kernel from here on
      std::fprintf(stderr,
                   "quaru-hip: bin/qs_sweep_compile not found beside the library; "
                   "sweeps use the generic kernel\n");
      j.mode = 0;
      return nullptr;
    }
    start_workers(j);
    it = j.table.emplace(sig, Entry()).first;
    ++j.busy;
    j.queue.emplace_back(sig, source(ops, c64, A, W, rp));
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
