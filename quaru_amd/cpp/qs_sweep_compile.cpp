// qs_sweep_compile -- compiles one specialised sweep kernel with hipRTC, in a
// process of its own.  The library (qs_sweep_jit.hip) starts it per kernel,
// so hipRTC never loads into the caller's process: nothing of it is left to
// tear down when that process exits, and compiles run in parallel.
//
// stdin: an 8-byte little-endian length L, then L bytes of the header
// "qs_sweep_kernel.h", then the kernel source up to end of input.  The
// arguments are the compile options.  stdout: the code object; exit status
// 0 on success, 1 with the compiler log on stderr otherwise.
#include <hip/hiprtc.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

int main(int argc, char **argv) {
  std::string in;
  char buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof buf, stdin)) > 0) in.append(buf, got);
  uint64_t hl = 0;
  if (in.size() < 8) {
    std::fprintf(stderr, "qs_sweep_compile: short input\n");
    return 1;
  }
  std::memcpy(&hl, in.data(), 8);
  if (hl > in.size() - 8) {
    std::fprintf(stderr, "qs_sweep_compile: bad header length\n");
    return 1;
  }
  const std::string header = in.substr(8, (size_t)hl);
  const std::string source = in.substr(8 + (size_t)hl);
  hiprtcProgram prog;
  const char *headers[] = {header.c_str()};
  const char *names[] = {"qs_sweep_kernel.h"};
  if (hiprtcCreateProgram(&prog, source.c_str(), "qs_sweep_jit.hip", 1, headers, names) !=
      HIPRTC_SUCCESS) {
    std::fprintf(stderr, "qs_sweep_compile: hiprtcCreateProgram failed\n");
    return 1;
  }
  const hiprtcResult rc = hiprtcCompileProgram(prog, argc - 1, (const char **)(argv + 1));
  size_t n = 0;
  if (rc != HIPRTC_SUCCESS) {
    if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
      std::string log(n, '\0');
      hiprtcGetProgramLog(prog, &log[0]);
      std::fprintf(stderr, "%s\n", log.c_str());
    }
    hiprtcDestroyProgram(&prog);
    return 1;
  }
  std::vector<char> code;
  if (hiprtcGetCodeSize(prog, &n) != HIPRTC_SUCCESS || n == 0) {
    hiprtcDestroyProgram(&prog);
    return 1;
  }
  code.resize(n);
  const bool ok = hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
  hiprtcDestroyProgram(&prog);
  if (!ok || std::fwrite(code.data(), 1, code.size(), stdout) != code.size()) return 1;
  return std::fflush(stdout) == 0 ? 0 : 1;
}
