// quaru.hpp -- C++ host API of the MI355X state-vector engine, shaped like
// quaru's Rust crate so that code written against `quaru::register::Register`
// and `quaru::operation` reads the same:
//
//   Rust (quaru 0.2.0)                        C++ (this header)
//   Register::new(&[bool])                    Register(std::vector<bool>)
//   Register::from_int(n, k)                  Register::from_int(n, k)
//   Register::try_from_qubits(&[Array2])      Register::try_from_qubits(...)
//   reg.apply(&op) -> &mut Self               reg.apply(op) -> Register&
//   reg.try_apply_fast(&op) -> Result         reg.try_apply_fast(op) -> Result
//   reg.measure(t) -> bool                    reg.measure(t) -> bool
//   Operation::new(m, targets) -> Option      Operation::create(m, targets)
//   operation::cnot(c, t), cnx, cnz, ...      operation::cnot(c, t), ...
//   OperationError::InvalidTarget(t), ...     OperationError{Kind::InvalidTarget, t}
//   panic via expect()                        throws quaru::Panic
//
// The state vector stays on the GPU; every call goes through the C ABI of
// quaru_hip.h, so this header carries no device code and builds with any
// C++17 compiler.  Header only.
#pragma once

#include <quaru_hip.h>

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <optional>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace quaru {

// ---------------------------------------------------------------- math ----
namespace math {

// Complex<f64> (src/math.rs:84-86).
using c64 = std::complex<double>;

// A dense row-major complex matrix: the stand-in for ndarray's Array2<c64>.
struct Matrix {
  size_t rows = 0, cols = 0;
  std::vector<c64> data;

  Matrix() = default;
  Matrix(size_t r, size_t c) : rows(r), cols(c), data(r * c, c64(0.0, 0.0)) {}
  Matrix(size_t r, size_t c, std::vector<c64> d) : rows(r), cols(c), data(std::move(d)) {
    if (data.size() != r * c) throw std::invalid_argument("Matrix: data size != rows * cols");
  }
  static Matrix eye(size_t d) {
    Matrix m(d, d);
    for (size_t i = 0; i < d; ++i) m(i, i) = 1.0;
    return m;
  }
  // A column vector (2^n x 1), the shape of one qubit or a whole state.
  static Matrix column(std::vector<c64> v) {
    const size_t n = v.size();
    return Matrix(n, 1, std::move(v));
  }
  c64 &operator()(size_t i, size_t j) { return data[i * cols + j]; }
  const c64 &operator()(size_t i, size_t j) const { return data[i * cols + j]; }
  bool operator==(const Matrix &o) const {
    return rows == o.rows && cols == o.cols && data == o.data;
  }
};

// Kronecker product a (x) b: b's index is the low part (ndarray linalg::kron).
inline Matrix kron(const Matrix &a, const Matrix &b) {
  Matrix out(a.rows * b.rows, a.cols * b.cols);
  for (size_t i = 0; i < a.rows; ++i)
    for (size_t j = 0; j < a.cols; ++j)
      for (size_t k = 0; k < b.rows; ++k)
        for (size_t l = 0; l < b.cols; ++l)
          out(i * b.rows + k, j * b.cols + l) = a(i, j) * b(k, l);
  return out;
}

// math::equal_qubits (src/math.rs:21-29): every entry within 1e-8.
inline bool equal_qubits(const std::vector<c64> &a, const std::vector<c64> &b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (std::abs(a[i] - b[i]) >= 1e-8) return false;
  return true;
}

}  // namespace math

using math::c64;
using math::Matrix;

// --------------------------------------------------------------- errors ----

// OperationError (src/register.rs:10-20).  The numeric kinds are the C
// ABI's status codes.
struct OperationError {
  enum Kind {
    InvalidTarget = QS_INVALID_TARGET,
    InvalidDimensions = QS_INVALID_DIMENSIONS,
    InvalidArity = QS_INVALID_ARITY,
    InvalidQubit = QS_INVALID_QUBIT
  };
  Kind kind = InvalidTarget;
  size_t a = 0, b = 0;

  bool operator==(const OperationError &o) const {
    return kind == o.kind && a == o.a && (kind != InvalidDimensions || b == o.b);
  }
  bool operator!=(const OperationError &o) const { return !(*this == o); }
  std::string to_string() const {
    switch (kind) {
      case InvalidTarget: return "InvalidTarget(" + std::to_string(a) + ")";
      case InvalidDimensions:
        return "InvalidDimensions(" + std::to_string(a) + ", " + std::to_string(b) + ")";
      case InvalidArity: return "InvalidArity(" + std::to_string(a) + ")";
      case InvalidQubit: return "InvalidQubit(" + std::to_string(a) + ")";
    }
    return "OperationError";
  }
};

// What Rust's `expect` does on an Err: the plain (non-try) forms throw it.
class Panic : public std::runtime_error {
 public:
  explicit Panic(const std::string &m) : std::runtime_error(m) {}
};

// A failure of the engine itself (HIP, out of memory, no device, bad
// argument).  Never used for an OperationError.
class EngineError : public std::runtime_error {
 public:
  EngineError(int code, const std::string &m) : std::runtime_error(m), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

// Result<T, OperationError>.
template <class T>
class Result {
 public:
  static Result ok(T v) {
    Result r;
    r.v_.emplace(std::move(v));
    return r;
  }
  static Result err(OperationError e) {
    Result r;
    r.e_ = e;
    return r;
  }
  bool is_ok() const { return v_.has_value(); }
  bool is_err() const { return !v_.has_value(); }
  T &unwrap() { return expect("called `Result::unwrap()` on an `Err` value"); }
  T &expect(const char *msg) {
    if (!v_) throw Panic(std::string(msg) + ": " + e_.to_string());
    return *v_;
  }
  const OperationError &unwrap_err() const {
    if (v_) throw Panic("called `Result::unwrap_err()` on an `Ok` value");
    return e_;
  }

 private:
  std::optional<T> v_;
  OperationError e_{};
};

namespace detail {

// Map a C ABI status to an OperationError (codes 1..4) or throw EngineError.
inline std::optional<OperationError> status(int rc) {
  if (rc == QS_OK) return std::nullopt;
  size_t a = 0, b = 0;
  const char *msg = nullptr;
  qs_last_error(&a, &b, &msg);
  if (rc >= QS_INVALID_TARGET && rc <= QS_INVALID_QUBIT)
    return OperationError{static_cast<OperationError::Kind>(rc), a, b};
  throw EngineError(rc, msg ? msg : "quaru_hip error");
}

inline void check(int rc) {
  if (auto e = status(rc)) throw Panic(e->to_string());
}

// SplitMix64 with 53-bit uniforms in [0, 1): the measurement generator.  The
// Python binding uses the same generator, so one seed gives the same
// outcomes from either host language.
class SplitMix64 {
 public:
  explicit SplitMix64(uint64_t seed) : s_(seed) {}
  uint64_t next_u64() {
    s_ += 0x9E3779B97F4A7C15ull;
    uint64_t z = s_;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }

 private:
  uint64_t s_;
};

inline std::string bits_of(size_t i, size_t n) {
  std::string s(n, '0');
  for (size_t b = 0; b < n; ++b)
    if ((i >> b) & 1) s[n - 1 - b] = '1';
  return s;
}

}  // namespace detail

// ------------------------------------------------------------ operation ----

// Operation (src/operation.rs:49-70): a 2^k x 2^k matrix and k targets;
// matrix-index bit j addresses qubit targets[j].
//
// Besides the dense form there is a structured one, Operation::controlled:
// a small matrix on `targets` applied where every control qubit is 1.  It is
// the same map as the dense matrix to_controlled / cnx / cnz build (controls
// as the highest matrix bits) and is what makes a 32-qubit cnz possible.
class Operation {
 public:
  // Operation::new (src/operation.rs:60-69): nullopt unless the matrix is
  // square with side 2^targets.size().
  static std::optional<Operation> create(Matrix matrix, std::vector<size_t> targets) {
    const size_t d = targets.size() < 64 ? (size_t)1 << targets.size() : 0;
    if (matrix.rows != d || matrix.cols != d) return std::nullopt;
    return Operation(std::move(matrix), std::move(targets), {});
  }

  // `matrix` on `targets` where qubit controls[i] equals control_values[i]
  // (every control 1 when control_values is empty).
  static Operation controlled(std::vector<size_t> controls, Matrix matrix,
                              std::vector<size_t> targets,
                              std::vector<bool> control_values = {}) {
    const size_t d = (size_t)1 << targets.size();
    if (matrix.rows != d || matrix.cols != d)
      throw std::invalid_argument("Operation::controlled: matrix side must be 2^targets");
    if (!control_values.empty() && control_values.size() != controls.size())
      throw std::invalid_argument("Operation::controlled: one control value per control");
    Operation o(std::move(matrix), std::move(targets), std::move(controls));
    o.cvals_ = std::move(control_values);
    return o;
  }

  // QuantumOperation::matrix (src/operation.rs:325-327): the dense matrix (a
  // copy).  For a structured operation this materialises 2^arity squared
  // entries, so it is refused beyond 14 qubits.
  Matrix matrix() const {
    if (controls_.empty()) return m_;
    const size_t k = targets_.size(), c = controls_.size();
    if (k + c > 14) throw Panic("dense form of a " + std::to_string(k + c) + "-qubit operation");
    const size_t d_all = (size_t)1 << (k + c), d = (size_t)1 << k;
    Matrix full = Matrix::eye(d_all);
    size_t pattern = 0;  // the control bits are the high matrix bits
    for (size_t i = 0; i < c; ++i)
      if (cvals_.empty() || cvals_[i]) pattern |= (size_t)1 << i;
    const size_t base = pattern << k;
    for (size_t i = 0; i < d; ++i)
      for (size_t j = 0; j < d; ++j) full(base + i, base + j) = m_(i, j);
    return full;
  }
  // QuantumOperation::targets / arity (src/operation.rs:329-337).
  std::vector<size_t> targets() const {
    std::vector<size_t> t = targets_;
    t.insert(t.end(), controls_.begin(), controls_.end());
    return t;
  }
  size_t arity() const { return targets_.size() + controls_.size(); }

  const Matrix &active_matrix() const { return m_; }
  const std::vector<size_t> &active_targets() const { return targets_; }
  const std::vector<size_t> &controls() const { return controls_; }
  // Empty: every control must be 1.
  const std::vector<bool> &control_values() const { return cvals_; }
  bool is_structured() const { return !controls_.empty(); }

  bool operator==(const Operation &o) const {
    return targets() == o.targets() && matrix() == o.matrix();
  }

 private:
  Operation(Matrix m, std::vector<size_t> t, std::vector<size_t> c)
      : m_(std::move(m)), targets_(std::move(t)), controls_(std::move(c)) {}
  Matrix m_;
  std::vector<size_t> targets_, controls_;
  std::vector<bool> cvals_;
};

// Gate constructors (src/operation.rs:340-536; src/openqasm.rs:145-163).
namespace operation {

constexpr double FRAC_1_SQRT_2 = 0.70710678118654752440;  // f64::consts::FRAC_1_SQRT_2

namespace detail {
inline Matrix real2(double a, double b, double c, double d) {
  return Matrix(2, 2, {c64(a, 0), c64(b, 0), c64(c, 0), c64(d, 0)});
}
inline Operation op(Matrix m, std::vector<size_t> t) {
  return *Operation::create(std::move(m), std::move(t));
}
// Dense multi-controlled matrices are built up to 2^12 x 2^12; larger ones
// use the structured form, which acts identically.
constexpr size_t kDenseLimit = 12;
}  // namespace detail

inline Operation identity(size_t target) {  // src/operation.rs:340-345
  return detail::op(detail::real2(1, 0, 0, 1), {target});
}

inline Operation hadamard(size_t target) {  // src/operation.rs:350-355
  const double h = FRAC_1_SQRT_2;
  return detail::op(detail::real2(h, h, h, -h), {target});
}

// src/operation.rs:358-366: kron(H, acc), so a later target is a higher bit.
inline Operation hadamard_transform(const std::vector<size_t> &targets) {
  Matrix m = hadamard(0).matrix();
  for (size_t i = 1; i < targets.size(); ++i) m = math::kron(hadamard(0).matrix(), m);
  return detail::op(std::move(m), targets);
}

// src/operation.rs:372-382: targets == [target, control].
inline Operation cnot(size_t control, size_t target) {
  Matrix m(4, 4);
  m(0, 0) = m(1, 1) = m(2, 3) = m(3, 2) = 1.0;
  return detail::op(std::move(m), {target, control});
}

// src/operation.rs:386-401: column c is the basis vector e_{f(c)}.
inline Operation to_quantum_gate(const std::function<size_t(size_t)> &f,
                                 const std::vector<size_t> &targets) {
  const size_t d = (size_t)1 << targets.size();
  Matrix m(d, d);
  for (size_t c = 0; c < d; ++c) m(f(c), c) = 1.0;
  return detail::op(std::move(m), targets);
}

// src/operation.rs:406-422: identity block first, `op` bottom right; the
// control becomes the new highest matrix bit.
inline Operation to_controlled(const Operation &o, size_t control) {
  const Matrix inner = o.matrix();
  const size_t d = inner.rows;
  Matrix m(2 * d, 2 * d);
  for (size_t i = 0; i < d; ++i) m(i, i) = 1.0;
  for (size_t i = 0; i < d; ++i)
    for (size_t j = 0; j < d; ++j) m(d + i, d + j) = inner(i, j);
  std::vector<size_t> t = o.targets();
  t.push_back(control);
  return detail::op(std::move(m), std::move(t));
}

inline Operation swap(size_t target_1, size_t target_2) {  // src/operation.rs:427-437
  Matrix m(4, 4);
  m(0, 0) = m(1, 2) = m(2, 1) = m(3, 3) = 1.0;
  return detail::op(std::move(m), {target_1, target_2});
}

inline Operation phase(size_t target) {  // src/operation.rs:443-451
  return detail::op(Matrix(2, 2, {1.0, 0.0, 0.0, c64(0.0, 1.0)}), {target});
}

// `not` (src/operation.rs:458-463); `not` is an operator keyword in C++.
inline Operation not_(size_t target) { return detail::op(detail::real2(0, 1, 1, 0), {target}); }

inline Operation pauli_y(size_t target) {  // src/operation.rs:468-476
  return detail::op(Matrix(2, 2, {0.0, c64(0, -1), c64(0, 1), 0.0}), {target});
}

inline Operation pauli_z(size_t target) {  // src/operation.rs:481-486
  return detail::op(detail::real2(1, 0, 0, -1), {target});
}

// src/operation.rs:491-513: targets == [target, controls...]; identity with
// the last two rows exchanged.
inline Operation cnx(const std::vector<size_t> &controls, size_t target) {
  const size_t k = controls.size() + 1;
  if (k > detail::kDenseLimit)
    return Operation::controlled(controls, detail::real2(0, 1, 1, 0), {target});
  const size_t d = (size_t)1 << k;
  Matrix m = Matrix::eye(d);
  m(d - 2, d - 2) = m(d - 1, d - 1) = 0.0;
  m(d - 2, d - 1) = m(d - 1, d - 2) = 1.0;
  std::vector<size_t> t{target};
  t.insert(t.end(), controls.begin(), controls.end());
  return detail::op(std::move(m), std::move(t));
}

// src/operation.rs:520-536: identity with -1 in the last diagonal entry.
inline Operation cnz(const std::vector<size_t> &controls, size_t target) {
  const size_t k = controls.size() + 1;
  if (k > detail::kDenseLimit)
    return Operation::controlled(controls, detail::real2(1, 0, 0, -1), {target});
  const size_t d = (size_t)1 << k;
  Matrix m = Matrix::eye(d);
  m(d - 1, d - 1) = -1.0;
  std::vector<size_t> t{target};
  t.insert(t.end(), controls.begin(), controls.end());
  return detail::op(std::move(m), std::move(t));
}

// OpenQASM U(theta, phi, lambda) (src/openqasm.rs:145-163).
inline std::optional<Operation> u(double theta, double phi, double lambda, size_t target) {
  const c64 i(0.0, 1.0);
  const double c = std::cos(theta / 2.0), s = std::sin(theta / 2.0);
  Matrix m(2, 2,
           {std::exp(-i * (phi + lambda) / 2.0) * c, -std::exp(-i * (phi - lambda) / 2.0) * s,
            std::exp(i * (phi - lambda) / 2.0) * s, std::exp(i * (phi + lambda) / 2.0) * c});
  return Operation::create(std::move(m), {target});
}

// RY(theta) = u(theta, 0, 0); RZ(theta) = u(0, 0, theta); RX = exp(-i theta X/2).
inline Operation ry(double theta, size_t target) {
  const double c = std::cos(theta / 2.0), s = std::sin(theta / 2.0);
  return detail::op(detail::real2(c, -s, s, c), {target});
}
inline Operation rz(double theta, size_t target) {
  return detail::op(Matrix(2, 2, {std::polar(1.0, -theta / 2.0), 0.0, 0.0,
                                  std::polar(1.0, theta / 2.0)}),
                    {target});
}
inline Operation rx(double theta, size_t target) {
  const double c = std::cos(theta / 2.0), s = std::sin(theta / 2.0);
  return detail::op(Matrix(2, 2, {c, c64(0, -s), c64(0, -s), c}), {target});
}
inline Operation cz(size_t a, size_t b) { return cnz({a}, b); }

}  // namespace operation

// QuantumCircuit (src/operation.rs:79-109, 271-322): operations, then
// measurements, then operations conditioned on measured values.  The reduce
// passes of the reference (src/operation.rs:111-269) are not reproduced: the
// register's sweep planner fuses gates on the device.
class QuantumCircuit {
 public:
  using Condition = std::vector<std::pair<size_t, bool>>;

  QuantumCircuit &reset_register() {
    operations_.clear();
    measurement_targets_.clear();
    conditional_operations_.clear();
    return *this;
  }
  const std::vector<Operation> &get_operations() const { return operations_; }
  const std::vector<size_t> &get_measurement_targets() const { return measurement_targets_; }
  const std::vector<std::pair<Operation, Condition>> &get_conditional_operations() const {
    return conditional_operations_;
  }
  void clear_operations() { operations_.clear(); }

  void add_operation(Operation op) {  // src/operation.rs:276-288
    if (touches_measured(op)) throw Panic("Cannot add operation after measurement");
    if (!conditional_operations_.empty())
      throw Panic("Cannot add operation after conditional operation (yet...)");
    operations_.push_back(std::move(op));
  }
  void add_measurement(size_t target) {  // src/operation.rs:292-297
    for (size_t t : measurement_targets_)
      if (t == target) throw Panic("Cannot add measurement after measurement");
    measurement_targets_.push_back(target);
  }
  void add_conditional_operation(Operation op, Condition cond) {  // src/operation.rs:302-321
    if (touches_measured(op)) throw Panic("Cannot add operation after measurement");
    for (const auto &c : cond) {
      bool found = false;
      for (size_t t : measurement_targets_) found = found || t == c.first;
      if (!found) throw Panic("Condition must be a measurement");
    }
    conditional_operations_.emplace_back(std::move(op), std::move(cond));
  }

 private:
  bool touches_measured(const Operation &op) const {
    for (size_t t : op.targets())
      for (size_t m : measurement_targets_)
        if (t == m) return true;
    return false;
  }
  std::vector<Operation> operations_;
  std::vector<size_t> measurement_targets_;
  std::vector<std::pair<Operation, Condition>> conditional_operations_;
};

// ------------------------------------------------------------- register ----

enum class Dtype { C128 = QS_C128, C64 = QS_C64 };

// Register (src/register.rs:22-407): n qubits, 2^n amplitudes resident in GPU
// memory.  Copying a Register deep-copies the state on the device (the
// reference's derive(Clone)).  Measurement uniforms come from a per-register
// SplitMix64 that is seeded from std::random_device (the reference draws
// from the OS-seeded rand::thread_rng, src/register.rs:317); seed() makes a
// run reproducible.
class Register {
 public:
  // Register::new (src/register.rs:41-56): input_bits[i] is qubit i.
  explicit Register(const std::vector<bool> &input_bits, Dtype dtype = Dtype::C128)
      : rng_(std::random_device{}()) {
    std::vector<unsigned char> b(input_bits.begin(), input_bits.end());
    detail::check(qs_register_new(b.data(), b.size(), (int)dtype, &h_));
  }

  // Register::from_int (src/register.rs:58-63).
  static Register from_int(size_t n, size_t k, Dtype dtype = Dtype::C128) {
    qs_register *h = nullptr;
    detail::check(qs_register_from_int(n, k, (int)dtype, &h));
    return Register(h);
  }

  // Register::try_from_qubits (src/register.rs:81-102): each qubit a column
  // of two amplitudes whose squared norms sum to 1 within 1e-6.
  static Result<Register> try_from_qubits(const std::vector<Matrix> &qubits,
                                          Dtype dtype = Dtype::C128) {
    std::vector<double> flat;
    std::vector<size_t> lens;
    for (const Matrix &q : qubits) {
      lens.push_back(q.data.size());
      for (const c64 &z : q.data) {
        flat.push_back(z.real());
        flat.push_back(z.imag());
      }
    }
    qs_register *h = nullptr;
    if (auto e = detail::status(qs_register_from_qubits(flat.data(), lens.data(), lens.size(),
                                                        (int)dtype, &h)))
      return Result<Register>::err(*e);
    return Result<Register>::ok(Register(h));
  }
  static Register from_qubits(const std::vector<Matrix> &qubits, Dtype dtype = Dtype::C128) {
    return std::move(try_from_qubits(qubits, dtype).expect("Invalid input qubits"));
  }

  Register(const Register &o) : rng_(o.rng_) { detail::check(qs_register_clone(o.h_, &h_)); }
  Register(Register &&o) noexcept : h_(o.h_), rng_(o.rng_) { o.h_ = nullptr; }
  Register &operator=(const Register &o) {
    if (this != &o) {
      qs_register *h = nullptr;
      detail::check(qs_register_clone(o.h_, &h));
      qs_register_free(h_);
      h_ = h;
      rng_ = o.rng_;
    }
    return *this;
  }
  Register &operator=(Register &&o) noexcept {
    if (this != &o) {
      qs_register_free(h_);
      h_ = o.h_;
      rng_ = o.rng_;
      o.h_ = nullptr;
    }
    return *this;
  }
  ~Register() { qs_register_free(h_); }

  // Reseed the measurement generator.
  Register &seed(uint64_t s) {
    rng_ = detail::SplitMix64(s);
    return *this;
  }

  // ---- gates (src/register.rs:134-281) ----------------------------------
  using Applied = Result<std::reference_wrapper<Register>>;

  // Register::apply (src/register.rs:134-137): panics on an invalid op.
  Register &apply(const Operation &op) {
    try_apply_fast(op).expect("Could not apply operation");
    return *this;
  }

  // Register::try_apply_fast (src/register.rs:140-191).
  Applied try_apply_fast(const Operation &op) {
    int rc;
    if (op.is_structured()) {
      const Matrix &m = op.active_matrix();
      std::vector<unsigned char> cv(op.control_values().begin(), op.control_values().end());
      rc = qs_apply_controlled(h_, op.controls().data(), cv.empty() ? nullptr : cv.data(),
                               op.controls().size(),
                               reinterpret_cast<const double *>(m.data.data()),
                               op.active_targets().data(), op.active_targets().size());
    } else {
      const Matrix &m = op.active_matrix();
      rc = qs_apply(h_, reinterpret_cast<const double *>(m.data.data()), m.rows, m.cols,
                    op.active_targets().data(), op.active_targets().size());
    }
    if (auto e = detail::status(rc)) return Applied::err(*e);
    return Applied::ok(std::ref(*this));
  }

  // Register::try_apply (src/register.rs:193-250): the same map.
  Applied try_apply(const Operation &op) { return try_apply_fast(op); }

  // `matrix` on `targets` where every qubit in `controls` is 1.
  Register &apply_controlled(const std::vector<size_t> &controls, const Matrix &matrix,
                             const std::vector<size_t> &targets) {
    detail::check(qs_apply_controlled(h_, controls.data(), nullptr, controls.size(),
                                      reinterpret_cast<const double *>(matrix.data.data()),
                                      targets.data(), targets.size()));
    return *this;
  }

  // A diagonal operation given by its 2^k diagonal entries.
  Register &apply_diagonal(const std::vector<c64> &diag, const std::vector<size_t> &targets) {
    if (diag.size() != ((size_t)1 << targets.size()))
      throw std::invalid_argument("apply_diagonal: need 2^k entries");
    detail::check(qs_apply_diagonal(h_, reinterpret_cast<const double *>(diag.data()),
                                    targets.data(), targets.size()));
    return *this;
  }

  // Register::apply_all / try_apply_all (src/register.rs:259-281).
  Register &apply_all(const Operation &op) {
    try_apply_all(op).expect("Could not apply operation");
    return *this;
  }
  Applied try_apply_all(const Operation &op) {
    const Matrix m = op.arity() <= 14 ? op.matrix() : Matrix();
    if (auto e = detail::status(qs_apply_all(h_, reinterpret_cast<const double *>(m.data.data()),
                                             m.rows, m.cols, op.arity())))
      return Applied::err(*e);
    return Applied::ok(std::ref(*this));
  }

  // ---- measurement (src/register.rs:291-342) ----------------------------
  bool measure(size_t target) { return try_measure(target).expect("Could not measure bit"); }
  Result<bool> try_measure(size_t target) { return try_measure_with(target, rng_.uniform()); }

  // try_measure with the uniform supplied by the caller.
  Result<bool> try_measure_with(size_t target, double uniform) {
    int res = 0;
    if (auto e = detail::status(qs_measure(h_, target, uniform, &res)))
      return Result<bool>::err(*e);
    return Result<bool>::ok(res != 0);
  }

  // measure(order[j]) for j = 0.. in one pass; bit order[j] of the result is
  // outcome j.  Draws one uniform per measured qubit.
  uint64_t measure_sequence(const std::vector<size_t> &order) {
    std::vector<double> u(order.size());
    for (double &x : u) x = rng_.uniform();
    return measure_sequence(order, u);
  }
  uint64_t measure_sequence(const std::vector<size_t> &order, const std::vector<double> &uniforms) {
    if (uniforms.size() != order.size())
      throw std::invalid_argument("measure_sequence: one uniform per qubit");
    uint64_t idx = 0;
    detail::check(qs_measure_sequence(h_, order.data(), order.size(), uniforms.data(), &idx));
    return idx;
  }
  uint64_t measure_all() {
    std::vector<size_t> order(size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    return measure_sequence(order);
  }

  // Inverse-CDF samples (FP64 prefix sums) without collapsing the state.
  std::vector<uint64_t> sample(const std::vector<double> &uniforms) {
    std::vector<uint64_t> out(uniforms.size());
    detail::check(qs_sample_cdf(h_, uniforms.data(), uniforms.size(), out.data()));
    return out;
  }

  // ---- circuits (src/register.rs:345-377) -------------------------------
  // Register::apply_circuit kept on the device: measurements decided on the
  // GPU (same uniforms, same rule as measure) and every conditional
  // operation applied as an operation controlled on its condition's measured
  // qubits.  After the collapse, "target measured v" holds exactly where
  // qubit `target` is v, as long as no conditional operation acts on a
  // measured qubit.  add_conditional_operation only refuses targets measured
  // before it (src/operation.rs:302-321); a later add_measurement may name a
  // conditional operation's target, and then a condition on that qubit must
  // read the recorded outcome: such circuits (and those with more
  // measurements than device slots) are decided on the host.  One
  // synchronisation at the end reads the outcomes.
  std::vector<std::pair<size_t, bool>> apply_circuit(QuantumCircuit &circuit) {
    for (const Operation &op : circuit.get_operations()) apply(op);
    const std::vector<size_t> targets = circuit.get_measurement_targets();
    bool host = targets.size() > QS_MEASURE_SLOTS;
    for (const auto &cop : circuit.get_conditional_operations())
      for (size_t t : cop.first.targets())
        for (size_t m : targets) host = host || t == m;
    if (host) return apply_circuit_host(circuit);
    for (size_t j = 0; j < targets.size(); ++j)
      detail::check(qs_measure_deferred(h_, targets[j], rng_.uniform(), j));
    for (const auto &cop : circuit.get_conditional_operations()) {
      std::vector<size_t> ctl;
      std::vector<unsigned char> cv;
      bool never = false;  // (t, true) and (t, false) cannot both be measured
      for (const auto &c : cop.second) {
        size_t k = 0;
        while (k < ctl.size() && ctl[k] != c.first) ++k;
        if (k == ctl.size()) {
          ctl.push_back(c.first);
          cv.push_back(c.second ? 1 : 0);
        } else if (cv[k] != (c.second ? 1 : 0)) {
          never = true;
        }
      }
      if (never) continue;
      const Operation &op = cop.first;
      if (op.is_structured()) {
        for (size_t k = 0; k < op.controls().size(); ++k) {
          ctl.push_back(op.controls()[k]);
          cv.push_back(op.control_values().empty() || op.control_values()[k] ? 1 : 0);
        }
      }
      const Matrix &m = op.active_matrix();
      const int rc = qs_apply_controlled(h_, ctl.data(), cv.data(), ctl.size(),
                                         reinterpret_cast<const double *>(m.data.data()),
                                         op.active_targets().data(), op.active_targets().size());
      if (auto e = detail::status(rc))
        Applied::err(*e).expect("Could not apply operation");
    }
    std::vector<unsigned char> out(targets.size());
    detail::check(qs_read_slots(h_, 0, targets.size(), out.data()));
    std::vector<std::pair<size_t, bool>> measures;
    for (size_t j = 0; j < targets.size(); ++j) measures.emplace_back(targets[j], out[j] != 0);
    return measures;
  }

  // The same with the outcomes decided on the host, one measure at a time
  // (src/register.rs:345-366 as written).
  std::vector<std::pair<size_t, bool>> apply_circuit_host(QuantumCircuit &circuit) {
    std::vector<std::pair<size_t, bool>> measures;
    if (!circuit.get_measurement_targets().empty()) measures = measure_circuit(circuit);
    for (const auto &cop : circuit.get_conditional_operations()) {
      bool ok = true;
      for (const auto &c : cop.second) {
        bool found = false;
        for (const auto &m : measures) found = found || m == c;
        ok = ok && found;
      }
      if (ok) apply(cop.first);
    }
    return measures;
  }
  std::vector<std::pair<size_t, bool>> measure_circuit(QuantumCircuit &circuit) {
    std::vector<std::pair<size_t, bool>> res;
    for (size_t t : circuit.get_measurement_targets()) res.emplace_back(t, measure(t));
    return res;
  }

  // ---- state access -----------------------------------------------------
  // The reference's `pub state` field: a host copy, and a replacement.
  std::vector<c64> state() const {
    std::vector<c64> out((size_t)1 << size());
    detail::check(qs_get_state(h_, reinterpret_cast<double *>(out.data())));
    return out;
  }
  Register &set_state(const std::vector<c64> &v) {
    if (v.size() != ((size_t)1 << size())) throw std::invalid_argument("state length != 2^n");
    detail::check(qs_set_state(h_, reinterpret_cast<const double *>(v.data())));
    return *this;
  }
  // state[i] for the given indices, without copying the whole state.
  std::vector<c64> amplitudes(const std::vector<uint64_t> &indices) const {
    std::vector<c64> out(indices.size());
    detail::check(qs_get_amplitudes(h_, indices.data(), indices.size(),
                                    reinterpret_cast<double *>(out.data())));
    return out;
  }
  double norm_sqr() const {
    double out = 0.0;
    detail::check(qs_norm_sqr(h_, &out));
    return out;
  }
  std::vector<double> probabilities() const {
    std::vector<double> out((size_t)1 << size());
    detail::check(qs_probabilities(h_, out.data()));
    return out;
  }
  // <this|other> accumulated in FP64 (either storage precision).
  c64 inner(const Register &o) const {
    double re = 0.0, im = 0.0;
    detail::check(qs_inner(h_, o.h_, &re, &im));
    return c64(re, im);
  }
  // 1 - |<a|b>|^2 / (<a|a><b|b>).
  double infidelity(const Register &o) const {
    return 1.0 - std::norm(inner(o)) / (norm_sqr() * o.norm_sqr());
  }
  bool equals(const Register &o, double tol) const {
    int eq = 0;
    detail::check(qs_equal(h_, o.h_, tol, &eq));
    return eq != 0;
  }
  // PartialEq (src/register.rs:409-413): every amplitude within 1e-8.
  bool operator==(const Register &o) const { return equals(o, 1e-8); }
  bool operator!=(const Register &o) const { return !equals(o, 1e-8); }

  // Register::size (src/register.rs:404-406).
  size_t size() const { return qs_register_size(h_); }
  Dtype dtype() const { return (Dtype)qs_register_dtype(h_); }

  // ---- printing (src/register.rs:379-401) -------------------------------
  void print_probabilities() const {
    const auto p = probabilities();
    for (size_t i = 0; i < p.size(); ++i)
      std::printf("%s: %.17g%%\n", detail::bits_of(i, size()).c_str(), p[i] * 100.0);
  }
  void print_nonzero_probabilities() const {
    const auto p = probabilities();
    for (size_t i = 0; i < p.size(); ++i)
      if (p[i] > 0.0) std::printf("%s: %.17g%%\n", detail::bits_of(i, size()).c_str(), p[i] * 100.0);
  }
  void print_state() const {
    const auto s = state();
    for (size_t i = 0; i < s.size(); ++i)
      std::printf("%s: %.17g%+.17gi\n", detail::bits_of(i, size()).c_str(), s[i].real(),
                  s[i].imag());
  }

  // ---- engine controls --------------------------------------------------
  Register &sync() {
    detail::check(qs_sync(h_));
    return *this;
  }
  Register &set_fusion(bool on) {
    detail::check(qs_set_fusion(h_, on ? 1 : 0));
    return *this;
  }
  qs_register *handle() const { return h_; }

 private:
  explicit Register(qs_register *h) : h_(h), rng_(std::random_device{}()) {}
  qs_register *h_ = nullptr;
  detail::SplitMix64 rng_;
};


// ------------------------------------------------------------- adjoint ----
// Gradients of E(theta) = <psi(theta)|H|psi(theta)> by the adjoint method,
// behind a Function-style forward/backward pair.  No reference counterpart:
// quaru has no parameterised gates; the rotations follow its OpenQASM U gate
// (src/openqasm.rs:145-163): RY(t) = u(t, 0, 0), RZ(t) = u(0, 0, t), and a
// circuit's fixed gates are the reference's Operations (src/operation.rs).
namespace adjoint {

enum class Axis { X = QS_AXIS_X, Y = QS_AXIS_Y, Z = QS_AXIS_Z };

// Fixed gates (k <= 2) and rotations exp(-i theta_p sigma / 2), with layer
// marks (the checkpoint granularity).
class ParamCircuit {
 public:
  explicit ParamCircuit(size_t n) : n_(n) { detail::check(qs_pcircuit_new(n, &h_)); }
  ParamCircuit(const ParamCircuit &) = delete;
  ParamCircuit &operator=(const ParamCircuit &) = delete;
  ParamCircuit(ParamCircuit &&o) noexcept : h_(o.h_), n_(o.n_) { o.h_ = nullptr; }
  ~ParamCircuit() { qs_pcircuit_free(h_); }

  // A fixed gate: the Operation's dense matrix on its targets.
  ParamCircuit &gate(const Operation &op) {
    const Matrix m = op.matrix();
    const std::vector<size_t> t = op.targets();
    detail::check(qs_pcircuit_gate(h_, reinterpret_cast<const double *>(m.data.data()), t.data(),
                                   t.size()));
    return *this;
  }
  // theta_param about `axis` on `target`.
  ParamCircuit &rotation(Axis axis, size_t target, size_t param) {
    detail::check(qs_pcircuit_rotation(h_, static_cast<int>(axis), target, param));
    return *this;
  }
  ParamCircuit &ry(size_t target, size_t param) { return rotation(Axis::Y, target, param); }
  ParamCircuit &rz(size_t target, size_t param) { return rotation(Axis::Z, target, param); }
  ParamCircuit &rx(size_t target, size_t param) { return rotation(Axis::X, target, param); }
  ParamCircuit &layer() {
    detail::check(qs_pcircuit_layer(h_));
    return *this;
  }
  size_t nparams() const { return qs_pcircuit_nparams(h_); }
  size_t size() const { return n_; }
  const qs_pcircuit *handle() const { return h_; }

 private:
  qs_pcircuit *h_ = nullptr;
  size_t n_ = 0;
};

// H = sum_t coeffs[t] Z^{masks[t]}: term t contributes
// coeffs[t] * (-1)^popcount(x & masks[t]) on basis state x.
struct Observable {
  std::vector<uint64_t> masks;
  std::vector<double> coeffs;
};

// forward(thetas) -> energies; backward(grad_energies) -> grad_thetas.
// thetas and gradients are row-major batch x nparams.  A forward keeps the
// batch's final states (and a checkpoint every `checkpoint_every` layers) on
// the device until the next backward or forward.
class Expectation {
 public:
  Expectation(const ParamCircuit &circuit, Observable h, Dtype dtype = Dtype::C128,
              size_t checkpoint_every = 0)
      : c_(&circuit), h_(std::move(h)), dtype_(dtype), ckpt_(checkpoint_every) {
    if (h_.masks.size() != h_.coeffs.size())
      throw std::invalid_argument("Observable: masks and coeffs differ in length");
  }
  Expectation(const Expectation &) = delete;
  Expectation &operator=(const Expectation &) = delete;
  ~Expectation() { release(); }

  // Energies of every parameter set; nothing stays on the device.
  std::vector<double> energies(const std::vector<double> &thetas) const {
    const size_t b = batch_of(thetas);
    std::vector<double> e(b);
    detail::check(qs_adjoint_forward(c_->handle(), thetas.data(), b, h_.masks.data(),
                                     h_.coeffs.data(), h_.masks.size(), (int)dtype_, 0, nullptr,
                                     e.data()));
    return e;
  }

  std::vector<double> forward(const std::vector<double> &thetas) {
    release();
    batch_ = batch_of(thetas);
    std::vector<double> e(batch_);
    detail::check(qs_adjoint_forward(c_->handle(), thetas.data(), batch_, h_.masks.data(),
                                     h_.coeffs.data(), h_.masks.size(), (int)dtype_, ckpt_, &ctx_,
                                     e.data()));
    return e;
  }

  // dE_b / dtheta_{b,p} (batch x nparams), one reverse sweep.
  std::vector<double> backward() {
    if (!ctx_) throw Panic("backward() needs a forward() first");
    std::vector<double> g(batch_ * c_->nparams());
    const int rc = qs_adjoint_backward(ctx_, g.data());
    release();
    detail::check(rc);
    return g;
  }
  // The vector-Jacobian product: grad_energies[b] * dE_b / dtheta_{b,p}.
  std::vector<double> backward(const std::vector<double> &grad_energies) {
    if (grad_energies.size() != batch_)
      throw std::invalid_argument("backward: one incoming gradient per parameter set");
    std::vector<double> g = backward();
    const size_t p = c_->nparams();
    for (size_t b = 0; b < grad_energies.size(); ++b)
      for (size_t j = 0; j < p; ++j) g[b * p + j] *= grad_energies[b];
    return g;
  }

 private:
  size_t batch_of(const std::vector<double> &thetas) const {
    const size_t p = c_->nparams();
    if (p == 0 || thetas.size() % p != 0 || thetas.empty())
      throw std::invalid_argument("thetas: need batch x nparams values");
    return thetas.size() / p;
  }
  void release() {
    if (ctx_) qs_adjoint_free(ctx_);
    ctx_ = nullptr;
  }

  const ParamCircuit *c_;
  Observable h_;
  Dtype dtype_;
  size_t ckpt_;
  qs_adjoint *ctx_ = nullptr;
  size_t batch_ = 0;
};

}  // namespace adjoint

}  // namespace quaru
