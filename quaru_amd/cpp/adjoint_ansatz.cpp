// adjoint_ansatz.cpp -- the adjoint gradients through the C++ host API
// (quaru::adjoint in include/quaru/quaru.hpp): the hardware-efficient ansatz
// (per layer RY(theta) and RZ(theta) on every qubit, then a CZ ring) with
// H = sum_i Z_i + 0.5 sum_i Z_i Z_{i+1}.  The CZ gates are the reference's
// Operations (operation::cz); the rotations follow src/openqasm.rs:145-163.
//
// Usage: adjoint_ansatz n layers [c128|c64] [checkpoint_every]
// Reads batch x nparams thetas (whitespace separated, row major) from stdin.
// Prints "energies" and the batch energies, then "grads" and the batch x
// nparams gradients (forward() + backward()), then "vjp" and the gradients
// of sum_b w_b E_b with w_b = b + 1 (backward(grad_energies)), each value
// with 17 significant digits.
#include <quaru/quaru.hpp>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace quaru;

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s n layers [c128|c64] [checkpoint_every]\n", argv[0]);
    return 2;
  }
  const size_t n = std::strtoul(argv[1], nullptr, 10);
  const size_t layers = std::strtoul(argv[2], nullptr, 10);
  const Dtype dtype = (argc > 3 && std::strcmp(argv[3], "c64") == 0) ? Dtype::C64 : Dtype::C128;
  const size_t ckpt = argc > 4 ? std::strtoul(argv[4], nullptr, 10) : 0;

  adjoint::ParamCircuit pc(n);
  size_t p = 0;
  for (size_t l = 0; l < layers; ++l) {
    for (size_t q = 0; q < n; ++q) pc.ry(q, p++);
    for (size_t q = 0; q < n; ++q) pc.rz(q, p++);
    for (size_t q = 0; q < n; ++q)
      if (q != (q + 1) % n) pc.gate(operation::cz(q, (q + 1) % n));
    pc.layer();
  }
  adjoint::Observable h;
  for (size_t i = 0; i < n; ++i) {
    h.masks.push_back(1ull << i);
    h.coeffs.push_back(1.0);
  }
  for (size_t i = 0; i + 1 < n; ++i) {
    h.masks.push_back((1ull << i) | (1ull << (i + 1)));
    h.coeffs.push_back(0.5);
  }

  std::vector<double> thetas;
  double v;
  while (std::scanf("%lf", &v) == 1) thetas.push_back(v);

  adjoint::Expectation ex(pc, h, dtype, ckpt);
  const std::vector<double> e = ex.forward(thetas);
  const std::vector<double> g = ex.backward();
  std::vector<double> w(e.size());
  for (size_t b = 0; b < w.size(); ++b) w[b] = (double)(b + 1);
  ex.forward(thetas);
  const std::vector<double> vjp = ex.backward(w);

  std::printf("energies\n");
  for (double x : e) std::printf("%.17g\n", x);
  std::printf("grads\n");
  for (double x : g) std::printf("%.17g\n", x);
  std::printf("vjp\n");
  for (double x : vjp) std::printf("%.17g\n", x);
  return 0;
}
