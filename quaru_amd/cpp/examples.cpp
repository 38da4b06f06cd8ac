// examples.cpp -- the reference's example programs (examples/grovers.rs,
// examples/quantum_teleportation.rs, examples/superdense_coding.rs, the QFT
// and modular-multiplication gates of examples/shors.rs) run unchanged in
// shape on the GPU register through the C++ host API, each checked against
// what the mathematics fixes.
//
// Usage: examples [c128|c64]
#include <quaru/quaru.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

using namespace quaru;
namespace op = quaru::operation;

namespace {

Dtype g_dtype = Dtype::C128;
int g_failed = 0;

struct Failure {
  std::string what;
};

#define CHECK(cond)                                                                       \
  do {                                                                                    \
    if (!(cond)) throw Failure{std::string(__FILE__ ":") + std::to_string(__LINE__) + \
                               ": " #cond};                                           \
  } while (0)

double amp_tol() { return g_dtype == Dtype::C128 ? 1e-12 : 2e-6; }

void run(const char *name, const std::function<void()> &f) {
  try {
    f();
    std::printf("PASS %s\n", name);
  } catch (const Failure &e) {
    ++g_failed;
    std::printf("FAIL %s: %s\n", name, e.what.c_str());
  } catch (const std::exception &e) {
    ++g_failed;
    std::printf("FAIL %s: exception %s\n", name, e.what());
  }
  std::fflush(stdout);
}

// ---- Grover (examples/grovers.rs:110-230) ---------------------------------

// The oracle: -1 on the diagonal where list[i] == winner (examples/grovers.rs:184-205).
Operation oracle_operation(size_t regsize, size_t winner, const std::vector<size_t> &list) {
  const size_t n = (size_t)1 << regsize;
  Matrix m(n, n);
  for (size_t i = 0; i < n; ++i) m(i, i) = (i < list.size() && list[i] == winner) ? -1.0 : 1.0;
  std::vector<size_t> t(regsize);
  for (size_t i = 0; i < regsize; ++i) t[i] = i;
  return *Operation::create(std::move(m), t);
}

// diag(1, -1, ..., -1) (examples/grovers.rs:207-219).
Operation diffusion_operation(size_t regsize) {
  const size_t n = (size_t)1 << regsize;
  Matrix m(n, n);
  m(0, 0) = 1.0;
  for (size_t i = 1; i < n; ++i) m(i, i) = -1.0;
  std::vector<size_t> t(regsize);
  for (size_t i = 0; i < regsize; ++i) t[i] = i;
  return *Operation::create(std::move(m), t);
}

void diffuser(Register &reg) {  // examples/grovers.rs:175-180
  reg.apply_all(op::hadamard(0));
  reg.apply(diffusion_operation(reg.size()));
  reg.apply_all(op::hadamard(0));
}

size_t iterations_floor(size_t regsize, size_t m) {  // examples/grovers.rs:221-226
  return (size_t)std::floor(M_PI / 4.0 * std::sqrt(std::pow(2.0, (double)regsize) / (double)m));
}

// ---- teleportation (examples/quantum_teleportation.rs:46-125) -------------

void quantum_teleportation(Register &reg) {
  reg.apply(op::hadamard(2));
  reg.apply(op::cnot(2, 1));
  reg.apply(op::cnot(0, 1));
  reg.apply(op::hadamard(0));
  if (reg.measure(1)) reg.apply(op::not_(2));
  if (reg.measure(0)) reg.apply(op::pauli_z(2));
}

QuantumCircuit quantum_teleportation_circuit() {
  QuantumCircuit c;
  c.add_operation(op::hadamard(2));
  c.add_operation(op::cnot(2, 1));
  c.add_operation(op::cnot(0, 1));
  c.add_operation(op::hadamard(0));
  c.add_measurement(1);
  c.add_measurement(0);
  c.add_conditional_operation(op::not_(2), {{1, true}});
  c.add_conditional_operation(op::pauli_z(2), {{0, true}});
  return c;
}

Matrix get_state_of_qubit(const std::vector<c64> &state, size_t n) {
  c64 alpha = 0.0, beta = 0.0;
  for (size_t i = 0; i < state.size(); ++i) ((i >> n) & 1 ? beta : alpha) += state[i];
  const double prob = std::norm(alpha) + std::norm(beta);
  return Matrix::column({alpha / std::sqrt(prob), beta / std::sqrt(prob)});
}

// |0> (x) |0> (x) q0 as the register's state (q0 is qubit 0).
std::vector<c64> teleport_input(c64 a, c64 b) {
  std::vector<c64> s(8, c64(0.0, 0.0));
  s[0] = a;
  s[1] = b;
  return s;
}

// ---- superdense coding (examples/superdense_coding.rs:39-80) --------------

Register superdense_preparation() {
  Register r(std::vector<bool>{false, false}, g_dtype);
  r.apply(op::hadamard(0));
  r.apply(op::cnot(0, 1));
  return r;
}

void superdense_encode(Register &r, bool m0, bool m1) {
  if (!m0 && !m1) r.apply(op::identity(0));
  if (!m0 && m1) r.apply(op::not_(0));
  if (m0 && !m1) r.apply(op::pauli_z(0));
  if (m0 && m1) {
    r.apply(op::pauli_z(0));
    r.apply(op::not_(0));
  }
}

std::pair<bool, bool> superdense_decode(Register &r) {
  r.apply(op::cnot(0, 1));
  r.apply(op::hadamard(0));
  const bool a = r.measure(0);
  const bool b = r.measure(1);
  return {a, b};
}

// ---- Shor's gates (examples/shors.rs:362-373, 434-503) --------------------

Operation qft(size_t n) {
  const size_t m = (size_t)1 << n;
  Matrix mat(m, m);
  const double s = 1.0 / std::sqrt((double)m);
  for (size_t i = 0; i < m; ++i)
    for (size_t j = 0; j < m; ++j)
      mat(i, j) = std::polar(s, 2.0 * M_PI * (double)((i * j) % m) / (double)m);
  std::vector<size_t> t(n);
  for (size_t i = 0; i < n; ++i) t[i] = i;
  return *Operation::create(std::move(mat), t);
}

}  // namespace

int main(int argc, char **argv) {
  if (argc > 1 && std::strcmp(argv[1], "c64") == 0) g_dtype = Dtype::C64;
  std::printf("dtype %s\n", g_dtype == Dtype::C128 ? "c128" : "c64");

  // Grover with dense oracle and diffusion ops as the example builds them:
  // after t iterations the winner's amplitude is sin((2t+1) theta) and every
  // other one cos((2t+1) theta) / sqrt(N - 1), sin(theta) = N^{-1/2}.
  run("grovers_algorithm", [] {
    for (size_t regsize = 2; regsize <= 9; ++regsize) {
      const size_t n_states = (size_t)1 << regsize;
      std::vector<size_t> list(n_states);
      for (size_t i = 0; i < n_states; ++i) list[i] = i;
      for (size_t winner : {(size_t)0, n_states / 3, n_states - 1}) {
        Register reg(std::vector<bool>(regsize, false), g_dtype);
        reg.seed(regsize * 131 + winner);
        reg.apply_all(op::hadamard(0));
        const Operation oracle = oracle_operation(regsize, winner, list);
        const size_t iters = iterations_floor(regsize, 1);
        for (size_t t = 0; t < iters; ++t) {
          reg.apply(oracle);
          diffuser(reg);
        }
        const double theta = std::asin(1.0 / std::sqrt((double)n_states));
        const double aw = std::sin((2.0 * iters + 1.0) * theta);
        const double ao = std::cos((2.0 * iters + 1.0) * theta) / std::sqrt(n_states - 1.0);
        const auto s = reg.state();
        if (g_dtype == Dtype::C128) {
          for (size_t i = 0; i < n_states; ++i)
            CHECK(std::abs(s[i] - c64(i == winner ? aw : ao, 0.0)) < amp_tol());
        } else {
          // complex64: infidelity against the closed form (the dense 2^n
          // matrices sum 2^n products in FP32 per amplitude)
          c64 ov = 0.0;
          double nn = 0.0;
          for (size_t i = 0; i < n_states; ++i) {
            ov += std::conj(c64(i == winner ? aw : ao, 0.0)) * s[i];
            nn += std::norm(s[i]);
          }
          CHECK(1.0 - std::norm(ov) / nn <= 1e-6);
        }
        // The example reads the result most significant qubit first.
        size_t got = 0;
        for (size_t i = reg.size(); i-- > 0;) got |= (size_t)reg.measure(i) << i;
        if (aw * aw > 0.99) CHECK(got == winner);
      }
    }
  });

  run("quantum_teleportation", [] {
    const c64 inputs[][2] = {{c64(0.8, 0.0), c64(0.0, -0.6)},
                             {c64(0.6, 0.0), c64(0.8, 0.0)},
                             {c64(0.5, 0.5), c64(0.5, -0.5)},
                             {c64(1.0, 0.0), c64(0.0, 0.0)}};
    for (const auto &in : inputs)
      for (uint64_t seed = 0; seed < 8; ++seed)
        for (int circuit = 0; circuit < 2; ++circuit) {
          Register reg(std::vector<bool>{false, false, false}, g_dtype);
          reg.seed(seed);
          reg.set_state(teleport_input(in[0], in[1]));
          if (circuit) {
            QuantumCircuit c = quantum_teleportation_circuit();
            reg.apply_circuit(c);
          } else {
            quantum_teleportation(reg);
          }
          const Matrix got = get_state_of_qubit(reg.state(), 2);
          std::vector<c64> want{in[0], in[1]};
          if (g_dtype == Dtype::C128)
            CHECK(math::equal_qubits(got.data, want));
          else
            CHECK(std::abs(got.data[0] - want[0]) + std::abs(got.data[1] - want[1]) < 1e-6);
        }
  });

  run("superdense_coding", [] {
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        Register r = superdense_preparation();
        superdense_encode(r, a, b);
        const auto got = superdense_decode(r);
        CHECK(got.first == (bool)a && got.second == (bool)b);
      }
  });

  // QFT|b> = N^{-1/2} sum_j w^{jb} |j> (the dense k-qubit kernels, k = 3..7).
  run("qft_closed_form", [] {
    for (size_t n = 3; n <= 7; ++n) {
      const size_t m = (size_t)1 << n;
      for (size_t b : {(size_t)1, m / 2 + 1, m - 1}) {
        Register r = Register::from_int(n + 2, b, g_dtype);
        r.apply(qft(n));
        const auto s = r.state();
        for (size_t j = 0; j < s.size(); ++j) {
          const c64 want = j < m ? std::polar(1.0 / std::sqrt((double)m),
                                              2.0 * M_PI * (double)((j * b) % m) / (double)m)
                                 : c64(0.0, 0.0);
          CHECK(std::abs(s[j] - want) < amp_tol());
        }
      }
    }
  });

  // Controlled modular multiplication x -> a x mod N (x < N), built with
  // to_quantum_gate and to_controlled as the Shor example builds it.
  run("controlled_modmul", [] {
    const size_t N = 15, a = 7;
    const Operation mul = op::to_quantum_gate(
        [&](size_t x) { return x < N ? (a * x) % N : x; }, {0, 1, 2, 3});
    const Operation cmul = op::to_controlled(mul, 4);
    for (size_t c = 0; c < 2; ++c)
      for (size_t x = 0; x < 16; ++x) {
        Register r = Register::from_int(6, x | (c << 4), g_dtype);
        r.apply(cmul);
        const size_t want = (c && x < N ? (a * x) % N : x) | (c << 4);
        size_t got = 0;
        for (size_t i = 0; i < 6; ++i) got |= (size_t)r.measure(i) << i;
        CHECK(got == want);
      }
  });

  std::printf("%s: %d failed\n", g_failed ? "FAILED" : "OK", g_failed);
  return g_failed ? 1 : 0;
}
