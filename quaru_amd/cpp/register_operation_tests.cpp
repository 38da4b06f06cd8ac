// register_operation_tests.cpp -- the reference's integration tests
// (tests/register_operation_tests.rs) written against the C++ host API and
// run on the GPU register.  The reference samples its inputs with proptest;
// here every case of each input domain is enumerated instead.
//
// Usage: register_operation_tests [c128|c64]
// Prints one line per test and exits non-zero if any fails.
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
int g_checks = 0;

struct Failure {
  std::string what;
};

#define CHECK(cond)                                                                       \
  do {                                                                                    \
    ++g_checks;                                                                           \
    if (!(cond)) throw Failure{std::string(__FILE__ ":") + std::to_string(__LINE__) + \
                               ": " #cond};                                           \
  } while (0)

// Register equality of the reference (every amplitude within 1e-8); a
// complex64 register is compared at the precision it stores.
bool same(const Register &a, const Register &b) {
  return a.equals(b, g_dtype == Dtype::C128 ? 1e-8 : 1e-6);
}

Register reg_of(const std::vector<bool> &bits) { return Register(bits, g_dtype); }

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

// tests/register_operation_tests.rs:228-246: the (normalised) sums of the
// amplitudes with bit n clear and set.
Matrix get_state_of_qubit(const std::vector<c64> &state, size_t n) {
  c64 alpha = 0.0, beta = 0.0;
  for (size_t i = 0; i < state.size(); ++i) {
    if ((i >> n) & 1)
      beta += state[i];
    else
      alpha += state[i];
  }
  const double prob = std::norm(alpha) + std::norm(beta);
  alpha /= std::sqrt(prob);
  beta /= std::sqrt(prob);
  return Matrix::column({alpha, beta});
}

std::vector<bool> bits_of(unsigned v, size_t n) {
  std::vector<bool> b(n);
  for (size_t i = 0; i < n; ++i) b[i] = (v >> i) & 1;
  return b;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc > 1 && std::strcmp(argv[1], "c64") == 0) g_dtype = Dtype::C64;
  std::printf("dtype %s\n", g_dtype == Dtype::C128 ? "c128" : "c64");

  // tests/register_operation_tests.rs:12-19
  run("measure_on_zero_state_gives_false", [] {
    Register r = reg_of({false});
    CHECK(r.measure(0) == false);
  });

  // tests/register_operation_tests.rs:24-41
  run("from_qubits_test", [] {
    const double f = 1.0 / std::sqrt(2.0);
    Register q = Register::try_from_qubits({Matrix::column({1.0, 0.0}), Matrix::column({0.0, 1.0}),
                                            Matrix::column({f, f})},
                                           g_dtype)
                     .expect("Incorrect input qubits");
    Register r = reg_of({false, true, false});
    r.apply(op::hadamard(2));
    CHECK(same(r, q));
  });

  // tests/register_operation_tests.rs:44-51 (i in 0..3)
  run("test_not_on_arbitrary_qubit", [] {
    for (size_t i = 0; i < 3; ++i) {
      Register r = reg_of({false, false, false});
      r.apply(op::not_(i));
      CHECK(r.measure(i) == true);
    }
  });

  // tests/register_operation_tests.rs:52-95: every 4- and 8-bit basis state.
  run("measure_four_qubits_gives_same_result", [] {
    for (unsigned v = 0; v < 16; ++v) {
      const auto b = bits_of(v, 4);
      Register r = reg_of(b);
      for (size_t i = 0; i < 4; ++i) CHECK(r.measure(i) == b[i]);
    }
  });
  run("measure_eight_qubits_gives_same_result", [] {
    for (unsigned v = 0; v < 256; ++v) {
      const auto b = bits_of(v, 8);
      Register r = reg_of(b);
      for (size_t i = 0; i < 8; ++i) CHECK(r.measure(i) == b[i]);
    }
  });

  // tests/register_operation_tests.rs:97-106
  run("hadamard_hadamard_retains_original_state", [] {
    for (size_t i = 0; i < 3; ++i) {
      Register r = reg_of({false, false, false});
      const Register expected = r;
      const Operation h = op::hadamard(i);
      Register &input = r.apply(h).apply(h);
      CHECK(same(input, expected));
    }
  });

  // tests/register_operation_tests.rs:108-124: every (i, j), several seeds.
  run("any_entangled_bit_measure_same", [] {
    for (size_t i = 0; i < 6; ++i)
      for (size_t j0 = 0; j0 < 6; ++j0)
        for (uint64_t s = 0; s < 4; ++s) {
          const size_t j = i == j0 ? (j0 + 1) % 6 : j0;
          Register r = reg_of(std::vector<bool>(6, false));
          r.seed(1000 * i + 10 * j0 + s);
          r.apply(op::hadamard(i));
          r.apply(op::cnot(i, j));
          CHECK(r.measure(i) == r.measure(j));
        }
  });

  // tests/register_operation_tests.rs:126-138
  run("first_bell_state_measure_equal", [] {
    for (size_t i = 0; i < 5; ++i)
      for (uint64_t s = 0; s < 4; ++s) {
        Register r = reg_of(std::vector<bool>(6, false));
        r.seed(77 * i + s);
        r.apply(op::hadamard(i));
        r.apply(op::cnot(i, i + 1));
        CHECK(r.measure(i) == r.measure(i + 1));
      }
  });

  // tests/register_operation_tests.rs:140-160 with the operation sets of
  // :262-300 (H, X, I on one qubit; CNOT, SWAP on two distinct qubits).
  run("arbitrary_unary_applied_twice_gives_equal", [] {
    for (size_t i = 0; i < 6; ++i)
      for (const Operation &o : {op::hadamard(i), op::not_(i), op::identity(i)}) {
        Register r = reg_of(std::vector<bool>(6, false));
        const Register expected = r;
        r.apply(o);
        r.apply(o);
        CHECK(same(r, expected));
      }
  });
  run("arbitrary_binary_applied_twice_gives_equal", [] {
    for (size_t i = 0; i < 6; ++i)
      for (size_t j = 0; j < 6; ++j) {
        if (i == j) continue;
        for (const Operation &o : {op::cnot(i, j), op::swap(i, j)}) {
          Register r = reg_of(std::vector<bool>(6, false));
          const Register expected = r;
          r.apply(o);
          r.apply(o);
          CHECK(same(r, expected));
        }
      }
  });

  // tests/register_operation_tests.rs:162-178
  run("swap_single_true_qubit", [] {
    for (size_t i = 0; i < 5; ++i)
      for (size_t j = 0; j < 5; ++j) {
        if (i == j) continue;
        std::vector<bool> q(6, false);
        q[i] = true;
        Register r = reg_of(q);
        r.apply(op::swap(i, j));
        for (size_t k = 0; k < 5; ++k) CHECK(r.measure(k) == (k == j));
      }
  });

  // tests/register_operation_tests.rs:180-203: n in 2..=6, every 6-bit input.
  run("toffoli_test", [] {
    for (size_t n = 2; n <= 6; ++n)
      for (unsigned v = 0; v < 64; ++v) {
        const auto init = bits_of(v, 6);
        Register r = reg_of(init);
        std::vector<size_t> controls;
        for (size_t i = 0; i + 1 < n; ++i) controls.push_back(i);
        r.apply(op::cnx(controls, n - 1));
        bool all = true;
        for (size_t i = 0; i + 1 < n; ++i) all = r.measure(i) && all;
        const bool t = r.measure(n - 1);
        CHECK(all ? t != init[n - 1] : t == init[n - 1]);
      }
  });

  // tests/register_operation_tests.rs:205-214
  run("apply_all_test", [] {
    for (size_t n = 2; n <= 6; ++n) {
      Register r1 = reg_of(std::vector<bool>(n, false));
      Register r2 = reg_of(std::vector<bool>(n, false));
      for (size_t i = 0; i < r1.size(); ++i) r1.apply(op::hadamard(i));
      r2.apply_all(op::hadamard(0));
      CHECK(same(r1, r2));
    }
  });

  // tests/register_operation_tests.rs:216-221: apply_all of a binary op panics.
  run("apply_all_panics_if_not_unary", [] {
    for (size_t i = 0; i < 6; ++i)
      for (size_t j = 0; j < 6; ++j) {
        if (i == j) continue;
        for (const Operation &o : {op::cnot(i, j), op::swap(i, j)}) {
          Register r = reg_of(std::vector<bool>(6, false));
          bool panicked = false;
          try {
            r.apply_all(o);
          } catch (const Panic &) {
            panicked = true;
          }
          CHECK(panicked);
          CHECK(r.try_apply_all(o).unwrap_err() == (OperationError{OperationError::InvalidArity, 2}));
        }
      }
  });

  // Error values and their precedence (src/register.rs:140-153, 297-301;
  // src/register.rs:81-102).
  run("operation_errors", [] {
    Register r = reg_of(std::vector<bool>(4, false));
    Matrix m4 = op::cnot(0, 1).matrix();
    auto dup = Operation::create(m4, {2, 2});
    CHECK(dup.has_value());
    CHECK(r.try_apply_fast(*dup).unwrap_err() == (OperationError{OperationError::InvalidTarget, 2}));
    auto range = Operation::create(m4, {7, 1});
    CHECK(r.try_apply(*range).unwrap_err() == (OperationError{OperationError::InvalidTarget, 7}));
    auto both = Operation::create(m4, {9, 9});
    CHECK(r.try_apply(*both).unwrap_err() == (OperationError{OperationError::InvalidTarget, 9}));
    CHECK(!Operation::create(Matrix(2, 2), {0, 1}).has_value());
    CHECK(r.try_measure(4).unwrap_err() == (OperationError{OperationError::InvalidTarget, 4}));
    bool panicked = false;
    try {
      r.apply(*range);
    } catch (const Panic &) {
      panicked = true;
    }
    CHECK(panicked);
    auto bad = Register::try_from_qubits({Matrix::column({1.0, 0.0}), Matrix::column({1.0, 1.0})},
                                         g_dtype);
    CHECK(bad.unwrap_err() == (OperationError{OperationError::InvalidQubit, 1}));
    auto three = Register::try_from_qubits({Matrix::column({1.0, 0.0, 0.0})}, g_dtype);
    CHECK(three.unwrap_err() == (OperationError{OperationError::InvalidQubit, 0}));
  });

  // Register::from_int (src/register.rs:58-63) and a structured cnz on more
  // qubits than a dense matrix could hold.
  run("from_int_and_structured_controls", [] {
    for (size_t k = 0; k < 32; ++k) {
      Register r = Register::from_int(5, k, g_dtype);
      size_t v = 0;
      for (size_t i = 0; i < 5; ++i) v |= (size_t)r.measure(i) << i;
      CHECK(v == k);
    }
    const size_t n = 16;
    Register a = reg_of(std::vector<bool>(n, false));
    a.apply_all(op::hadamard(0));
    std::vector<size_t> controls;
    for (size_t i = 1; i < n; ++i) controls.push_back(i);
    const Operation z = op::cnz(controls, 0);
    CHECK(z.is_structured());
    a.apply(z);
    const auto s = a.state();
    const double amp = std::pow(2.0, -0.5 * n);
    for (size_t i = 0; i < s.size(); ++i) {
      const double want = i == s.size() - 1 ? -amp : amp;
      CHECK(std::abs(s[i] - c64(want, 0.0)) < (g_dtype == Dtype::C128 ? 1e-14 : 1e-7));
    }
  });

  // Deep copy semantics of derive(Clone).
  run("clone_is_deep", [] {
    Register a = reg_of({false, false});
    Register b = a;
    b.apply(op::not_(1));
    CHECK(a.measure(1) == false);
    CHECK(b.measure(1) == true);
  });

  // The state field written and read back (examples/quantum_teleportation.rs:51-58).
  run("state_field_round_trip", [] {
    Register r = reg_of({false, false, false});
    std::vector<c64> v(8);
    for (size_t i = 0; i < 8; ++i) v[i] = c64(0.1 * (double)i, -0.05 * (double)i);
    r.set_state(v);
    const auto w = r.state();
    for (size_t i = 0; i < 8; ++i)
      CHECK(std::abs(w[i] - v[i]) < (g_dtype == Dtype::C128 ? 0.0 : 1e-7) + 1e-300);
    Matrix q = get_state_of_qubit(w, 2);
    CHECK(std::abs(std::norm(q(0, 0)) + std::norm(q(1, 0)) - 1.0) < 1e-12);
  });

  std::printf("%s: %d checks, %d failed tests\n", g_failed ? "FAILED" : "OK", g_checks, g_failed);
  return g_failed ? 1 : 0;
}
