// asan_check.cpp — drives the CPU oracle's C API under AddressSanitizer and
// UndefinedBehaviorSanitizer.
//
// TEST INFRASTRUCTURE ONLY (built by `make -C oracle asan`, run by
// tests/test_oracle_asan.py).  It exercises every entry point of
// ppvm_oracle.h on random inputs: programs with every op code (Cliffords,
// rotations, noise channels, amplitude damping, truncation) at several word
// counts, thresholds, weight caps and preserve sets; clones, overlap, trace
// and pattern membership (valid and invalid patterns), export, fingerprint,
// marked runs and the sharded multi-thread variant; lossy sums with the loss
// channels.  The sanitizers abort on the first invalid access or undefined
// operation; the program prints "asan ok" when every call returned.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ppvm_oracle.h"

namespace {

struct Rng {
  uint64_t s;
  uint64_t next() {  // splitmix64
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  uint32_t below(uint32_t n) { return (uint32_t)(next() % n); }
  double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

const uint32_t kGates[] = {PPVM_OP_X,      PPVM_OP_Y,        PPVM_OP_Z,      PPVM_OP_H,          PPVM_OP_S,
                           PPVM_OP_S_DAG,  PPVM_OP_SQRT_X,   PPVM_OP_SQRT_X_DAG, PPVM_OP_SQRT_Y,
                           PPVM_OP_SQRT_Y_DAG, PPVM_OP_CNOT, PPVM_OP_CZ,     PPVM_OP_CY,         PPVM_OP_RX,
                           PPVM_OP_RY,     PPVM_OP_RZ,       PPVM_OP_RXX,    PPVM_OP_RYY,        PPVM_OP_RZZ,
                           PPVM_OP_PAULI_ERROR, PPVM_OP_DEPOLARIZE1, PPVM_OP_DEPOLARIZE2,
                           PPVM_OP_PAULI_ERROR2, PPVM_OP_AMPLITUDE_DAMPING, PPVM_OP_TRUNCATE};
const uint32_t kLossOps[] = {PPVM_OP_LOSS, PPVM_OP_CORRELATED_LOSS, PPVM_OP_RESET_LOSS};

bool two_qubit(uint32_t c) {
  return c == PPVM_OP_CNOT || c == PPVM_OP_CZ || c == PPVM_OP_CY || c == PPVM_OP_RXX || c == PPVM_OP_RYY ||
         c == PPVM_OP_RZZ || c == PPVM_OP_DEPOLARIZE2 || c == PPVM_OP_PAULI_ERROR2 ||
         c == PPVM_OP_CORRELATED_LOSS;
}

ppvm_op make_op(Rng& r, uint32_t code, uint32_t n) {
  ppvm_op op;
  std::memset(&op, 0, sizeof(op));
  op.code = code;
  op.q0 = r.below(n);
  if (two_qubit(code)) {
    do op.q1 = r.below(n);
    while (op.q1 == op.q0);
  }
  op.theta = (r.unit() - 0.5) * 6.0;
  for (int k = 0; k < 15; ++k) op.p[k] = r.unit() * 0.02;
  return op;
}

std::vector<ppvm_op> program(Rng& r, uint32_t n, size_t len, bool lossy) {
  std::vector<ppvm_op> p;
  const size_t ng = sizeof(kGates) / sizeof(kGates[0]);
  for (size_t i = 0; i < len; ++i) {
    uint32_t code = kGates[r.below((uint32_t)ng)];
    if (lossy && r.below(6) == 0) code = kLossOps[r.below(3)];
    p.push_back(make_op(r, code, n));
    if (r.below(3) == 0) p.push_back(make_op(r, PPVM_OP_TRUNCATE, n));
  }
  return p;
}

void random_words(Rng& r, uint32_t n, uint32_t W, size_t count, std::vector<uint64_t>& x,
                  std::vector<uint64_t>& z) {
  x.assign(count * W, 0);
  z.assign(count * W, 0);
  for (size_t t = 0; t < count; ++t)
    for (uint32_t q = 0; q < n; ++q) {
      if (r.below(4) != 0) continue;
      const uint32_t v = r.below(4);
      if (v & 1) x[t * W + q / 64] |= 1ull << (q % 64);
      if (v & 2) z[t * W + q / 64] |= 1ull << (q % 64);
    }
}

int fails = 0;
void check(bool ok, const char* what) {
  if (!ok) {
    std::printf("check failed: %s\n", what);
    ++fails;
  }
}

void plain_case(Rng& r, uint32_t n, double tau, uint64_t maxw, bool preserve) {
  ppvo_handle h = ppvo_create(n, tau, maxw);
  const uint32_t W = ppvo_words(h);
  std::vector<uint64_t> x, z;
  const size_t nt = 12;
  random_words(r, n, W, nt, x, z);
  std::vector<double> c(nt);
  for (auto& v : c) v = r.unit() * 2.0 - 1.0;
  if (preserve) ppvo_set_preserve(h, x.data(), z.data(), 3);
  ppvo_add_terms(h, x.data(), z.data(), c.data(), nt);
  // without a threshold the sum can double at every rotation: shorter programs
  const std::vector<ppvm_op> prog = program(r, n, tau > 0.0 ? 120 : 40, false);
  check(ppvo_apply(h, prog.data(), prog.size()) == 0, "apply");
  ppvo_handle k = ppvo_clone(h);
  const std::vector<ppvm_op> more = program(r, n, 40, false);
  check(ppvo_apply(k, more.data(), more.size()) == 0, "apply on clone");
  const double ov = ppvo_overlap(h, k) + ppvo_overlap_with_zero(h);
  check(std::isfinite(ov), "overlap finite");
  double tr = 0.0;
  check(ppvo_trace(h, "Z?*", &tr) == 0, "trace Z?*");
  (void)ppvo_trace(h, "[XY]{2}_*Z?", &tr);
  check(ppvo_trace(h, "[[", &tr) != 0, "bad pattern is rejected");
  std::vector<uint8_t> in(nt);
  (void)ppvo_pattern_contains(h, "X*Z?_", x.data(), z.data(), in.data(), nt);
  const uint64_t len = ppvo_len(h);
  std::vector<uint64_t> ex(len * W + 1), ez(len * W + 1);
  std::vector<double> ec(len + 1);
  check(ppvo_export(h, ex.data(), ez.data(), ec.data(), len) == len, "export");
  check(ppvo_export(h, ex.data(), ez.data(), ec.data(), len / 2) == len, "partial export");
  uint64_t fp[3];
  ppvo_fingerprint(h, fp);
  check(fp[0] == len, "fingerprint count");
  (void)ppvo_term_updates(h);
  // marked run and the sharded variant on copies
  ppvo_handle m = ppvo_clone(h);
  std::vector<uint64_t> marks;
  for (size_t i = 0; i < prog.size(); i += 7) marks.push_back(i);
  std::vector<uint64_t> ml(marks.size());
  std::vector<double> mo(marks.size());
  check(ppvo_run_marked(m, prog.data(), prog.size(), marks.data(), marks.size(), ml.data(), mo.data()) ==
            marks.size(),
        "marked run");
  ppvo_handle sh = ppvo_clone(h);
  check(ppvo_apply_sharded(sh, prog.data(), prog.size(), 3) == 0, "sharded apply");
  ppvo_destroy(sh);
  ppvo_destroy(m);
  ppvo_destroy(k);
  ppvo_destroy(h);
}

void lossy_case(Rng& r, uint32_t n, double tau, uint64_t maxw, uint64_t maxl) {
  ppvo_handle h = ppvo_lossy_create(n, tau, maxw, maxl);
  const uint32_t W = ppvo_words(h);
  std::vector<uint64_t> x, z;
  const size_t nt = 8;
  random_words(r, n, W, nt, x, z);
  std::vector<uint64_t> l(nt * W, 0);
  for (size_t t = 0; t < nt; ++t) {
    const uint32_t q = r.below(n);
    const uint64_t b = 1ull << (q % 64);
    if (r.below(3) == 0) {  // lost qubit: x = z = 0 there
      l[t * W + q / 64] |= b;
      x[t * W + q / 64] &= ~b;
      z[t * W + q / 64] &= ~b;
    }
  }
  std::vector<double> c(nt);
  for (auto& v : c) v = r.unit() + 0.1;
  ppvo_lossy_set_preserve(h, x.data(), z.data(), l.data(), 2);
  ppvo_lossy_add_terms(h, x.data(), z.data(), l.data(), c.data(), nt);
  const std::vector<ppvm_op> prog = program(r, n, tau > 0.0 ? 120 : 40, true);
  check(ppvo_apply(h, prog.data(), prog.size()) == 0, "lossy apply");
  double tr = 0.0;
  (void)ppvo_trace(h, "Z?*", &tr);
  const double ov = ppvo_overlap_with_zero(h);
  check(std::isfinite(ov), "lossy overlap finite");
  const uint64_t len = ppvo_len(h);
  std::vector<uint64_t> ex(len * W + 1), ez(len * W + 1), el(len * W + 1);
  std::vector<double> ec(len + 1);
  check(ppvo_lossy_export(h, ex.data(), ez.data(), el.data(), ec.data(), len) == len, "lossy export");
  ppvo_handle k = ppvo_clone(h);
  check(std::isfinite(ppvo_overlap(h, k)), "lossy overlap with clone");
  ppvo_destroy(k);
  ppvo_destroy(h);
}

}  // namespace

int main() {
  Rng r{20261204};
  const uint32_t sizes[] = {5, 64, 70, 130, 300};
  for (uint32_t n : sizes) {
    for (int rep = 0; rep < 3; ++rep) {
      plain_case(r, n, rep == 0 ? 0.0 : 1e-3, rep == 2 ? (uint64_t)(n / 2 + 2) : ~0ull, rep == 1);
      lossy_case(r, n, rep == 0 ? 0.0 : 1e-3, rep == 2 ? (uint64_t)(n / 2 + 2) : ~0ull,
                 rep == 1 ? 2u : ~0ull);
    }
  }
  if (fails) {
    std::printf("%d checks failed\n", fails);
    return 1;
  }
  std::printf("asan ok\n");
  return 0;
}
