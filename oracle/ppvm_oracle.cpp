// ppvm_oracle.cpp — CPU restatement of the reference PauliSum propagation path.
//
// TEST INFRASTRUCTURE ONLY (parity oracle + CPU baseline).  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library,
// and only to check results against or as the timed CPU baseline.
//
// Reference files restated here (paths relative to the reference repository):
//   crates/ppvm-pauli-sum/src/sum/data.rs:192-311   scale / map_insert / map_add / truncate
//   crates/ppvm-pauli-sum/src/sum/clifford.rs:13-217 Clifford conjugation
//   crates/ppvm-pauli-sum/src/sum/rot1.rs:55-143     rx / ry / rz / r
//   crates/ppvm-pauli-sum/src/sum/rot2.rs:63-184     rzz / rxx / ryy
//   crates/ppvm-pauli-sum/src/sum/noise.rs:17-131    diagonal noise channels
//   crates/ppvm-pauli-sum/src/sum/trace.rs:9-37      trace / overlap
//   crates/ppvm-pauli-sum/src/strategy.rs:12-134     CombinedStrategy truncation
//   crates/ppvm-traits/src/traits/coefficient.rs:46-62 mul_sign / sin_cos / cutoff
//   crates/ppvm-traits/src/map/hashmap.rs:40-54,108-127,170-179,242-259 IndexMap ops
//   crates/ppvm-traits/src/traits/clifford.rs:117-165,270-341 Clifford bit maps
//   crates/ppvm-pauli-word/src/phase/clifford.rs:16-180 Clifford sign rules
//   crates/ppvm-pauli-word/src/loss/data.rs:33-316    lossy words (get / set / weights)
//   crates/ppvm-pauli-sum/src/sum/noise.rs:159-274    loss, correlated loss, reset loss
//   crates/ppvm-pauli-sum/src/strategy.rs:137-170     MaxLossWeight
//
// The map is insertion ordered (IndexMap): `+=` on a new key appends, `retain`
// keeps order, Clifford re-insertion keeps order.  Floating point follows the
// reference operation by operation (no FMA contraction: built -ffp-contract=off).

#include "ppvm_oracle.h"

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

inline uint64_t rotl64(uint64_t v, int r) { return (v << r) | (v >> (64 - r)); }

// Word-folding hash for keys (any good 64-bit hash gives the same map
// semantics; only probe lengths depend on it).
inline uint64_t key_hash(const uint64_t* k, uint32_t nw) {
  uint64_t h = 0x243f6a8885a308d3ull;
  for (uint32_t i = 0; i < nw; ++i) h = (rotl64(h, 5) ^ k[i]) * 0x517cc1b727220a95ull;
  h ^= h >> 29;
  h *= 0xbf58476d1ce4e5b9ull;
  h ^= h >> 32;
  return h;
}

struct Sum {
  uint32_t n = 0, W = 1, NW = 2;  // qubits, words per plane, words per key (2W)
  double tau = 0.0;                // CoefficientThreshold
  uint64_t maxw = UINT64_MAX;      // MaxPauliWeight (UINT64_MAX disables)
  bool lossy = false;              // LossyPauliWord keys: NW = 3W (x, z, loss planes)
  uint64_t maxlw = UINT64_MAX;     // MaxLossWeight (lossy sums only)
  std::vector<uint64_t> keys;      // entries * NW (x words then z words)
  std::vector<double> coef;
  std::vector<uint64_t> hashes;
  std::vector<uint32_t> slots;     // open-addressing index over entries
  uint64_t mask = 0;
  std::vector<uint64_t> preserve;  // P * NW
  uint64_t term_updates = 0;
  // scratch for rotation children (sum/data.rs:217-225)
  std::vector<uint64_t> sk;
  std::vector<double> sc;

  size_t len() const { return coef.size(); }
  const uint64_t* key(size_t i) const { return &keys[i * NW]; }
  uint64_t* key(size_t i) { return &keys[i * NW]; }

  void rebuild_index() {
    size_t want = 16;
    while (want < 2 * len() + 2) want <<= 1;
    slots.assign(want, UINT32_MAX);
    mask = want - 1;
    for (size_t i = 0; i < len(); ++i) {
      uint64_t p = hashes[i] & mask;
      while (slots[p] != UINT32_MAX) p = (p + 1) & mask;
      slots[p] = (uint32_t)i;
    }
  }
  void rehash_all() {
    hashes.resize(len());
    for (size_t i = 0; i < len(); ++i) hashes[i] = key_hash(key(i), NW);
    rebuild_index();
  }
  long find(const uint64_t* k, uint64_t h) const {
    if (slots.empty()) return -1;
    uint64_t p = h & mask;
    while (true) {
      uint32_t s = slots[p];
      if (s == UINT32_MAX) return -1;
      if (hashes[s] == h && std::memcmp(key(s), k, NW * 8) == 0) return s;
      p = (p + 1) & mask;
    }
  }
  // entry(k).and_modify(|v| *v += c).or_insert(c)   (map/hashmap.rs:40-44)
  void add_assign(const uint64_t* k, double c) {
    uint64_t h = key_hash(k, NW);
    long i = find(k, h);
    if (i >= 0) {
      coef[i] += c;
      return;
    }
    keys.insert(keys.end(), k, k + NW);
    coef.push_back(c);
    hashes.push_back(h);
    if (2 * len() + 2 > slots.size()) {
      rebuild_index();
    } else {
      uint64_t p = h & mask;
      while (slots[p] != UINT32_MAX) p = (p + 1) & mask;
      slots[p] = (uint32_t)(len() - 1);
    }
  }
  uint64_t weight(const uint64_t* k) const {
    uint64_t w = 0;
    for (uint32_t i = 0; i < W; ++i) w += __builtin_popcountll(k[i] | k[W + i]);
    return w;
  }
  bool xb(const uint64_t* k, uint32_t q) const { return (k[q >> 6] >> (q & 63)) & 1; }
  bool zb(const uint64_t* k, uint32_t q) const { return (k[W + (q >> 6)] >> (q & 63)) & 1; }
  void setx(uint64_t* k, uint32_t q, bool v) const {
    uint64_t m = 1ull << (q & 63);
    k[q >> 6] = v ? (k[q >> 6] | m) : (k[q >> 6] & ~m);
  }
  void setz(uint64_t* k, uint32_t q, bool v) const {
    uint64_t m = 1ull << (q & 63);
    k[W + (q >> 6)] = v ? (k[W + (q >> 6)] | m) : (k[W + (q >> 6)] & ~m);
  }
  bool lb(const uint64_t* k, uint32_t q) const { return lossy && ((k[2 * W + (q >> 6)] >> (q & 63)) & 1); }
};

// ---- Clifford conjugation (sum/clifford.rs; bit maps traits/clifford.rs;
//      signs phase/clifford.rs).  Returns true if the coefficient is negated.
bool clifford_word(const Sum& s, uint32_t code, uint32_t a, uint32_t b, uint64_t* k) {
  bool x = s.xb(k, a), z = s.zb(k, a);
  switch (code) {
    case PPVM_OP_X: return z;              // clifford.rs:39-46
    case PPVM_OP_Y: return x ^ z;          // :49-56
    case PPVM_OP_Z: return x;              // :59-66
    case PPVM_OP_H:                        // :69-87
      s.setx(k, a, z); s.setz(k, a, x);
      return x && z;
    case PPVM_OP_S:                        // :90-107
      s.setz(k, a, x ^ z);
      return x && !z;
    case PPVM_OP_S_DAG:                    // :115-132
      s.setz(k, a, x ^ z);
      return x && z;
    case PPVM_OP_SQRT_X:                   // :135-152
      s.setx(k, a, x ^ z);
      return x && z;
    case PPVM_OP_SQRT_X_DAG:               // :176-193
      s.setx(k, a, x ^ z);
      return !x && z;
    case PPVM_OP_SQRT_Y:                   // :155-173
      s.setx(k, a, z); s.setz(k, a, x);
      return !x && z;
    case PPVM_OP_SQRT_Y_DAG:               // :196-214
      s.setx(k, a, z); s.setz(k, a, x);
      return x && !z;
    case PPVM_OP_CNOT: {                   // bits traits/clifford.rs:142-151, sign phase/clifford.rs:72-85
      bool xc = x, zc = z, xt = s.xb(k, b), zt = s.zb(k, b);
      bool neg = xc && zt && (xt == zc);
      s.setz(k, a, zt ^ zc);
      s.setx(k, b, xc ^ xt);
      return neg;
    }
    case PPVM_OP_CZ: {                     // bits :154-165, sign :88-99
      bool xa = x, za = z, xbb = s.xb(k, b), zbb = s.zb(k, b);
      bool neg = xa && xbb && (za ^ zbb);
      s.setz(k, a, za ^ xbb);
      s.setz(k, b, zbb ^ xa);
      return neg;
    }
    case PPVM_OP_CY: {                     // bits traits/clifford.rs:329-341, sign phase/clifford.rs:168-180
      bool xc = x, zc = z, xt = s.xb(k, b), zt = s.zb(k, b);
      bool neg = xc && (xt ^ zt) && !(zc ^ zt);
      s.setz(k, a, zc ^ xt ^ zt);
      s.setx(k, b, xt ^ xc);
      s.setz(k, b, zt ^ xc);
      return neg;
    }
  }
  return false;
}

// ---- two_qubit_pauli_error anticommuting-index lists (sum/noise.rs:56-106).
// Rows are (P_a, P_b) in the reference's match order; each lists, in the
// reference's summation order, the indices of the 15 channel Paulis
// (IX,IY,IZ,XI,XX,XY,XZ,YI,YX,YY,YZ,ZI,ZX,ZY,ZZ) that anticommute with P_a P_b.
// Kept identical because the order of the FP subtractions fixes the bits.
const int kPe2Rows[16][8] = {
    {-1, -1, -1, -1, -1, -1, -1, -1},  // I I
    {1, 10, 13, 14, 2, 5, 6, 9},       // I X
    {0, 10, 12, 14, 2, 4, 6, 8},       // I Y
    {0, 1, 12, 13, 4, 5, 8, 9},        // I Z
    {10, 11, 12, 13, 14, 7, 8, 9},     // X I
    {1, 11, 12, 2, 5, 6, 7, 8},        // X X
    {0, 11, 13, 2, 4, 6, 7, 9},        // X Y
    {0, 1, 10, 11, 14, 4, 5, 7},       // X Z
    {11, 12, 13, 14, 3, 4, 5, 6},      // Y I
    {1, 10, 11, 12, 2, 3, 4, 9},       // Y X
    {0, 10, 11, 13, 2, 3, 5, 8},       // Y Y
    {0, 1, 11, 14, 3, 6, 8, 9},        // Y Z
    {10, 3, 4, 5, 6, 7, 8, 9},         // Z I
    {1, 13, 14, 2, 3, 4, 7, 8},        // Z X
    {0, 12, 14, 2, 3, 5, 7, 9},        // Z Y
    {0, 1, 10, 12, 13, 3, 6, 7},       // Z Z
};
// Pauli letter index in the rows above: I=0 X=1 Y=2 Z=3.
inline int letter(bool x, bool z) { return x ? (z ? 2 : 1) : (z ? 3 : 0); }

// Factor table indexed by 4*letter(a)+letter(b); NaN-free, 1.0 means "no-op".
void pe2_factors(const double* p, double out[16], bool applies[16]) {
  for (int r = 0; r < 16; ++r) {
    applies[r] = r != 0;
    double acc = 1.0;
    if (r != 0)
      for (int j = 0; j < 8; ++j) acc = acc - p[kPe2Rows[r][j]] * 2.0;
    out[r] = acc;
  }
}

// ---- rotation specs (sum/rot1.rs:55-128, sum/rot2.rs:63-184).
// Returns eps (+1/-1) for an anticommuting key and writes the child key;
// returns 0 for a commuting key.
inline int rot_child(const Sum& s, uint32_t code, uint32_t a, uint32_t b, const uint64_t* k,
                     uint64_t* child) {
  bool xa = s.xb(k, a), za = s.zb(k, a);
  switch (code) {
    case PPVM_OP_RX:
      if (!za) return 0;
      std::memcpy(child, k, s.NW * 8);
      s.setx(child, a, !xa);
      return xa ? -1 : 1;
    case PPVM_OP_RY:
      if (xa == za) return 0;
      std::memcpy(child, k, s.NW * 8);
      s.setx(child, a, !xa);
      s.setz(child, a, !za);
      return za ? -1 : 1;
    case PPVM_OP_RZ:
      if (!xa) return 0;
      std::memcpy(child, k, s.NW * 8);
      s.setz(child, a, !za);
      return za ? 1 : -1;
    case PPVM_OP_RZZ: {
      bool xbb = s.xb(k, b), zbb = s.zb(k, b);
      if (xa == xbb) return 0;
      bool z_anti = xa ? za : zbb;
      std::memcpy(child, k, s.NW * 8);
      s.setz(child, a, !za);
      s.setz(child, b, !zbb);
      return z_anti ? 1 : -1;
    }
    case PPVM_OP_RXX: {
      bool xbb = s.xb(k, b), zbb = s.zb(k, b);
      if (za == zbb) return 0;
      bool x_anti = za ? xa : xbb;
      std::memcpy(child, k, s.NW * 8);
      s.setx(child, a, !xa);
      s.setx(child, b, !xbb);
      return x_anti ? -1 : 1;
    }
    case PPVM_OP_RYY: {
      bool xbb = s.xb(k, b), zbb = s.zb(k, b);
      bool pa = xa ^ za, pb = xbb ^ zbb;
      if (pa == pb) return 0;
      bool x_anti = pa ? xa : xbb;
      std::memcpy(child, k, s.NW * 8);
      s.setx(child, a, !xa);
      s.setz(child, a, !za);
      s.setx(child, b, !xbb);
      s.setz(child, b, !zbb);
      return x_anti ? 1 : -1;
    }
  }
  return 0;
}

// PauliSum::map_insert (sum/data.rs:204-226): parents scaled in place while
// children go to the scratch vector, then each child is merged with `+=`.
void rotate(Sum& s, uint32_t code, uint32_t a, uint32_t b, double theta) {
  // Float::sin_cos (coefficient.rs:55-57): one libm sincos call, the combined
  // form LLVM emits for f64::sin_cos on Linux (glibc's separate sin() can
  // differ from it in the last bit)
  double sn, cs;
  sincos(theta, &sn, &cs);
  s.sk.clear();
  s.sc.clear();
  std::vector<uint64_t> child(s.NW);
  const size_t n = s.len();
  s.term_updates += n;
  for (size_t i = 0; i < n; ++i) {
    int eps = rot_child(s, code, a, b, s.key(i), child.data());
    if (eps == 0) continue;
    double coeff = s.coef[i];
    s.coef[i] *= cs;
    coeff *= (double)eps * sn;  // sin.mul_sign(eps)  (coefficient.rs:47-49)
    s.sk.insert(s.sk.end(), child.begin(), child.end());
    s.sc.push_back(coeff);
  }
  for (size_t j = 0; j < s.sc.size(); ++j) s.add_assign(&s.sk[j * s.NW], s.sc[j]);
}

// AmplitudeDamping for PauliSum (sum/noise.rs:132-157), also a map_insert:
// X and Y on the qubit scale by sqrt(1 - gamma); Z scales by (1 - gamma) and
// branches gamma * c onto the word with I on the qubit.
void amplitude_damping(Sum& s, uint32_t a, double gamma) {
  s.sk.clear();
  s.sc.clear();
  const size_t n = s.len();
  s.term_updates += n;
  const double decay = std::sqrt(1.0 - gamma), keep = 1.0 - gamma;
  for (size_t i = 0; i < n; ++i) {
    const uint64_t* k = s.key(i);
    const bool x = s.xb(k, a), z = s.zb(k, a);
    if (x) {
      s.coef[i] *= decay;
    } else if (z) {
      const double child = s.coef[i] * gamma;
      s.sk.insert(s.sk.end(), k, k + s.NW);
      s.sk[s.sk.size() - s.NW + s.W + (a >> 6)] &= ~(1ull << (a & 63));  // Z -> I
      s.sc.push_back(child);
      s.coef[i] *= keep;
    }
  }
  for (size_t j = 0; j < s.sc.size(); ++j) s.add_assign(&s.sk[j * s.NW], s.sc[j]);
}

void truncate(Sum& s) {
  // snapshot preserved entries (sum/data.rs:283-294)
  std::vector<uint64_t> saved_k;
  std::vector<double> saved_c;
  if (!s.preserve.empty()) {
    size_t P = s.preserve.size() / s.NW;
    for (size_t i = 0; i < s.len(); ++i)
      for (size_t p = 0; p < P; ++p)
        if (std::memcmp(s.key(i), &s.preserve[p * s.NW], s.NW * 8) == 0) {
          saved_k.insert(saved_k.end(), s.key(i), s.key(i) + s.NW);
          saved_c.push_back(s.coef[i]);
        }
  }
  // CoefficientThreshold then MaxPauliWeight retain (strategy.rs:19-30,86-89,133)
  size_t o = 0;
  for (size_t i = 0; i < s.len(); ++i) {
    bool keep = !(std::fabs(s.coef[i]) < s.tau);
    if (keep && s.maxw != UINT64_MAX) keep = s.weight(s.key(i)) <= s.maxw;
    if (!keep) continue;
    if (o != i) {
      std::memmove(s.key(o), s.key(i), s.NW * 8);
      s.coef[o] = s.coef[i];
      s.hashes[o] = s.hashes[i];
    }
    ++o;
  }
  s.keys.resize(o * s.NW);
  s.coef.resize(o);
  s.hashes.resize(o);
  s.rebuild_index();
  // restore dropped preserved entries at the end (sum/data.rs:305-310)
  for (size_t j = 0; j < saved_c.size(); ++j) {
    const uint64_t* k = &saved_k[j * s.NW];
    if (s.find(k, key_hash(k, s.NW)) < 0) s.add_assign(k, saved_c[j]);
  }
}

int apply_one(Sum& s, const ppvm_op& op) {
  const uint32_t code = op.code;
  auto qok = [&](uint32_t q) { return q < s.n; };
  switch (code) {
    case PPVM_OP_NOP: return 0;
    case PPVM_OP_TRUNCATE: truncate(s); return 0;
    case PPVM_OP_X: case PPVM_OP_Y: case PPVM_OP_Z: case PPVM_OP_H: case PPVM_OP_S:
    case PPVM_OP_S_DAG: case PPVM_OP_SQRT_X: case PPVM_OP_SQRT_X_DAG: case PPVM_OP_SQRT_Y:
    case PPVM_OP_SQRT_Y_DAG: {
      if (!qok(op.q0)) return 1;
      s.term_updates += s.len();
      bool moved = false;
      for (size_t i = 0; i < s.len(); ++i) {
        if (clifford_word(s, code, op.q0, 0, s.key(i))) s.coef[i] = -s.coef[i];
      }
      moved = !(code == PPVM_OP_X || code == PPVM_OP_Y || code == PPVM_OP_Z);
      if (moved) s.rehash_all();  // map_add re-inserts in order (sum/data.rs:255-265)
      return 0;
    }
    case PPVM_OP_CNOT: case PPVM_OP_CZ: case PPVM_OP_CY: {
      if (!qok(op.q0) || !qok(op.q1) || op.q0 == op.q1) return 1;
      s.term_updates += s.len();
      for (size_t i = 0; i < s.len(); ++i)
        if (clifford_word(s, code, op.q0, op.q1, s.key(i))) s.coef[i] = -s.coef[i];
      s.rehash_all();
      return 0;
    }
    case PPVM_OP_RX: case PPVM_OP_RY: case PPVM_OP_RZ:
      if (!qok(op.q0)) return 1;
      rotate(s, code, op.q0, 0, op.theta);
      return 0;
    case PPVM_OP_RXX: case PPVM_OP_RYY: case PPVM_OP_RZZ:
      if (!qok(op.q0) || !qok(op.q1) || op.q0 == op.q1) return 1;
      rotate(s, code, op.q0, op.q1, op.theta);
      return 0;
    case PPVM_OP_AMPLITUDE_DAMPING:  // noise.rs:132-157, gamma = p[0]
      if (!qok(op.q0)) return 1;
      amplitude_damping(s, op.q0, op.p[0]);
      return 0;
    case PPVM_OP_PAULI_ERROR: {  // noise.rs:17-39
      if (!qok(op.q0)) return 1;
      s.term_updates += s.len();
      const double one = 1.0;
      const double xf = one - op.p[1] * 2.0 - op.p[2] * 2.0;
      const double zf = one - op.p[0] * 2.0 - op.p[1] * 2.0;
      const double yf = one - op.p[0] * 2.0 - op.p[2] * 2.0;
      for (size_t i = 0; i < s.len(); ++i) {
        bool x = s.xb(s.key(i), op.q0), z = s.zb(s.key(i), op.q0);
        if (x && !z) s.coef[i] *= xf;
        else if (!x && z) s.coef[i] *= zf;
        else if (x && z) s.coef[i] *= yf;
      }
      return 0;
    }
    case PPVM_OP_DEPOLARIZE1: {  // noise.rs:109-118
      if (!qok(op.q0)) return 1;
      s.term_updates += s.len();
      const double f = 1.0 - op.p[0] * (4.0 / 3.0);
      for (size_t i = 0; i < s.len(); ++i)
        if (s.xb(s.key(i), op.q0) || s.zb(s.key(i), op.q0)) s.coef[i] *= f;
      return 0;
    }
    case PPVM_OP_DEPOLARIZE2: {  // noise.rs:121-131
      if (!qok(op.q0) || !qok(op.q1)) return 1;
      s.term_updates += s.len();
      const double f = 1.0 - op.p[0] * (16.0 / 15.0);
      for (size_t i = 0; i < s.len(); ++i) {
        const uint64_t* k = s.key(i);
        if (s.xb(k, op.q0) || s.zb(k, op.q0) || s.xb(k, op.q1) || s.zb(k, op.q1))
          s.coef[i] *= f;
      }
      return 0;
    }
    case PPVM_OP_PAULI_ERROR2: {  // noise.rs:52-107
      if (!qok(op.q0) || !qok(op.q1)) return 1;
      s.term_updates += s.len();
      double fac[16];
      bool app[16];
      pe2_factors(op.p, fac, app);
      for (size_t i = 0; i < s.len(); ++i) {
        const uint64_t* k = s.key(i);
        int r = 4 * letter(s.xb(k, op.q0), s.zb(k, op.q0)) + letter(s.xb(k, op.q1), s.zb(k, op.q1));
        if (app[r]) s.coef[i] *= fac[r];
      }
      return 0;
    }
  }
  return 1;
}

double overlap_zero(const Sum& s) {
  // fold over entries in insertion order (map/hashmap.rs:118-127) accepting
  // terms whose every slot is I or Z (pattern "Z?*", pattern/contains.rs:72-88)
  double acc = 0.0;
  for (size_t i = 0; i < s.len(); ++i) {
    const uint64_t* k = s.key(i);
    bool zero = true;
    for (uint32_t w = 0; w < s.W; ++w) zero &= (k[w] == 0);
    if (s.lossy)  // an L is neither I nor Z
      for (uint32_t w = 0; w < s.W; ++w) zero &= (k[2 * s.W + w] == 0);
    if (zero) acc += s.coef[i];
  }
  return acc;
}

// ---------------------------------------------------------------------------
// Pauli patterns for trace(pattern): grammar of
// crates/ppvm-pauli-word/src/pattern/parse.rs:30-143 and greedy matcher of
// pattern/contains.rs:20-151.  A single-site pattern is kept as the set of
// accepted Pauli codes (bit c for code c = x | z << 1: I 0, X 1, Z 2, Y 3).
// As in contains.rs:39 the "any Pauli or identity" form ('_', '[XYZ]?')
// accepts the identity only.
enum PatKind { PK_POSITION = 0, PK_STAR = 1, PK_REPEAT = 2 };
struct PatItem {
  int kind;
  unsigned accept;
  uint64_t arg;
};
constexpr unsigned kAcceptAnyOrI = 1u;  // pattern/contains.rs:39

// usize::parse on a digit string: non-empty, no overflow
bool parse_usize(const std::string& num, uint64_t& v) {
  if (num.empty()) return false;
  v = 0;
  for (char ch : num) {
    const uint64_t d = (uint64_t)(ch - '0');
    if (v > (UINT64_MAX - d) / 10) return false;
    v = v * 10 + d;
  }
  return true;
}

bool parse_pattern(const char* text, std::vector<PatItem>& out) {
  std::string str(text);
  size_t a = 0, b = str.size();
  while (a < b && std::isspace((unsigned char)str[a])) ++a;  // str::trim
  while (b > a && std::isspace((unsigned char)str[b - 1])) --b;
  const std::string in = str.substr(a, b - a);
  size_t i = 0;
  auto code_of = [](char c) -> int { return c == 'X' ? 1 : c == 'Z' ? 2 : c == 'Y' ? 3 : -1; };
  out.clear();
  while (i < in.size()) {
    const char c = in[i++];
    unsigned accept = 0;
    int nset = 0;  // 1 single, 2 double, 3 any non-identity, 0 any-or-identity
    if (c == '[') {
      unsigned set = 0;
      while (i < in.size()) {
        const char ch = in[i];
        if (ch == ']') {
          ++i;
          break;
        }
        const int cd = code_of(ch);
        if (cd < 0) return false;  // "Expected X, Y, or Z in []"
        set |= 1u << cd;
        ++i;
      }
      nset = __builtin_popcount(set);
      if (nset < 1 || nset > 3) return false;
      accept = set;
    } else if (c == '_') {
      nset = 0;
      accept = kAcceptAnyOrI;
    } else {
      const int cd = code_of(c);
      if (cd < 0) return false;
      nset = 1;
      accept = 1u << cd;
    }
    if (i < in.size() && in[i] == '?') {  // OpPattern::add_identity (data.rs:37-45)
      ++i;
      if (nset == 1 || nset == 2) accept |= 1u;
      else if (nset == 3) accept = kAcceptAnyOrI;
    }
    PatItem it{0, accept, 0};
    if (i < in.size() && in[i] == '*') {
      ++i;
      it.kind = PK_STAR;
    } else if (i < in.size() && in[i] == '{') {
      ++i;
      std::string num;
      while (i < in.size()) {
        const char ch = in[i];
        if (std::isdigit((unsigned char)ch)) {
          num.push_back(ch);
          ++i;
        } else if (ch == '}') {
          ++i;
          break;
        } else {
          return false;
        }
      }
      if (!parse_usize(num, it.arg)) return false;
      it.kind = PK_REPEAT;
    } else if (i < in.size() && std::isdigit((unsigned char)in[i])) {
      std::string num;
      while (i < in.size() && std::isdigit((unsigned char)in[i])) num.push_back(in[i++]);
      if (!parse_usize(num, it.arg)) return false;
      it.kind = PK_POSITION;
    } else {
      return false;  // "Expected '*' or digit after Pauli pattern"
    }
    out.push_back(it);
  }
  return true;
}

bool pattern_contains(const Sum& s, const uint64_t* k, const std::vector<PatItem>& pat) {
  const uint64_t n = s.n;
  uint64_t cur = 0;  // next qubit the word iterator yields (word/data.rs:421-428)
  // Pauli of a slot: I 0, X 1, Z 2, Y 3, and L 4 on a lossy word, which no
  // op pattern accepts and which is not the identity
  auto code = [&](uint64_t q) -> unsigned {
    if (s.lb(k, (uint32_t)q)) return 4u;
    return (unsigned)s.xb(k, (uint32_t)q) | ((unsigned)s.zb(k, (uint32_t)q) << 1);
  };
  for (const PatItem& it : pat) {
    bool next = false;
    if (it.kind == PK_POSITION) {  // match_position, contains.rs:48-70
      while (cur < n) {
        const uint64_t p = cur++;
        const unsigned ch = code(p);
        if (p == it.arg) {
          next = (it.accept >> ch) & 1u;
          break;
        }
        if (ch != 0) break;  // non-identity before the anchored position
      }
    } else if (it.kind == PK_STAR) {  // match_star, contains.rs:72-88
      while (cur < n && ((it.accept >> code(cur)) & 1u)) ++cur;
      next = true;
    } else {  // match_repeat, contains.rs:90-117
      uint64_t matched = 0;
      bool decided = false;
      while (cur < n) {
        const unsigned ch = code(cur);
        if ((it.accept >> ch) & 1u) {
          ++cur;
          ++matched;
          if (matched == it.arg) {
            next = true;
            decided = true;
            break;
          }
        } else {
          next = !(matched < it.arg);
          decided = true;
          break;
        }
      }
      if (!decided) next = matched == it.arg;
    }
    if (!next) return false;
  }
  for (; cur < n; ++cur)  // remaining symbols must be identities (contains.rs:138-148)
    if (code(cur) != 0) return false;
  return true;
}

// trace(pattern): fold over the entries in insertion order (map/hashmap.rs:108-127)
double trace_pattern(const Sum& s, const std::vector<PatItem>& pat) {
  double acc = 0.0;
  for (size_t i = 0; i < s.len(); ++i)
    if (pattern_contains(s, s.key(i), pat)) acc += s.coef[i];
  return acc;
}

// ---------------------------------------------------------------------------
// Sharded multi-thread variant (CPU baseline modelled on the DashMap + rayon
// configuration, crates/ppvm-traits/src/map/dashmap.rs:42-52,119-133,156-165).
// Keys are partitioned over shards by hash; a rotation scales parents and
// routes children to their owner shard in parallel, then every shard merges
// its incoming children in parallel.  Per-key arithmetic is the same as the
// single-thread path (a key receives at most its own scaled value and one
// child), so coefficients are bit-identical.
struct Shard {
  Sum m;
};

// Persistent workers: run(fn) executes fn(t) for t = 0..S-1 (t = 0 on the
// calling thread) and returns when all are done.
class Pool {
 public:
  explicit Pool(int S) : S_(S) {
    for (int t = 1; t < S_; ++t) th_.emplace_back([this, t] { loop(t); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& x : th_) x.join();
  }
  void run(const std::function<void(int)>& fn) {
    {
      std::lock_guard<std::mutex> g(mu_);
      job_ = &fn;
      left_ = S_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    fn(0);
    std::unique_lock<std::mutex> g(mu_);
    done_.wait(g, [this] { return left_ == 0; });
  }

 private:
  void loop(int t) {
    uint64_t seen = 0;
    while (true) {
      const std::function<void(int)>* job;
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
        job = job_;
      }
      (*job)(t);
      {
        std::lock_guard<std::mutex> g(mu_);
        if (--left_ == 0) done_.notify_one();
      }
    }
  }
  int S_;
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int)>* job_ = nullptr;
  int left_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

int apply_sharded(Sum& src, const ppvm_op* ops, size_t nops, int threads) {
  if (threads < 1) threads = 1;
  const int S = threads;
  std::vector<Shard> sh(S);
  for (int t = 0; t < S; ++t) {
    Sum& m = sh[t].m;
    m.n = src.n; m.W = src.W; m.NW = src.NW; m.tau = src.tau; m.maxw = src.maxw;
    m.preserve = src.preserve;
  }
  for (size_t i = 0; i < src.len(); ++i) {
    uint64_t h = key_hash(src.key(i), src.NW);
    sh[h % S].m.add_assign(src.key(i), src.coef[i]);
  }
  uint64_t tu = 0;
  std::vector<std::vector<std::vector<uint64_t>>> outk(S, std::vector<std::vector<uint64_t>>(S));
  std::vector<std::vector<std::vector<double>>> outc(S, std::vector<std::vector<double>>(S));
  Pool pool(S);
  auto par = [&](auto fn) {
    const std::function<void(int)> f = fn;
    pool.run(f);
  };
  for (size_t oi = 0; oi < nops; ++oi) {
    const ppvm_op& op = ops[oi];
    uint64_t live = 0;
    for (int t = 0; t < S; ++t) live += sh[t].m.len();
    bool is_rot = op.code >= PPVM_OP_RX && op.code <= PPVM_OP_RZZ;
    bool is_moving = (op.code >= PPVM_OP_H && op.code <= PPVM_OP_CY);
    if (op.code != PPVM_OP_TRUNCATE && op.code != PPVM_OP_NOP) tu += live;
    if (is_rot) {
      double sn, cs;
      sincos(op.theta, &sn, &cs);
      par([&](int t) {
        Sum& m = sh[t].m;
        std::vector<uint64_t> child(m.NW);
        for (int d = 0; d < S; ++d) { outk[t][d].clear(); outc[t][d].clear(); }
        for (size_t i = 0; i < m.len(); ++i) {
          int eps = rot_child(m, op.code, op.q0, op.q1, m.key(i), child.data());
          if (eps == 0) continue;
          double coeff = m.coef[i];
          m.coef[i] *= cs;
          coeff *= (double)eps * sn;
          int d = (int)(key_hash(child.data(), m.NW) % S);
          outk[t][d].insert(outk[t][d].end(), child.begin(), child.end());
          outc[t][d].push_back(coeff);
        }
      });
      par([&](int d) {
        Sum& m = sh[d].m;
        for (int t = 0; t < S; ++t)
          for (size_t j = 0; j < outc[t][d].size(); ++j) m.add_assign(&outk[t][d][j * m.NW], outc[t][d][j]);
      });
    } else if (is_moving) {
      // keys change shard: apply per shard then re-route
      par([&](int t) {
        Sum& m = sh[t].m;
        for (int d = 0; d < S; ++d) { outk[t][d].clear(); outc[t][d].clear(); }
        for (size_t i = 0; i < m.len(); ++i) {
          uint64_t* k = m.key(i);
          double c = m.coef[i];
          if (clifford_word(m, op.code, op.q0, op.q1, k)) c = -c;
          int d = (int)(key_hash(k, m.NW) % S);
          outk[t][d].insert(outk[t][d].end(), k, k + m.NW);
          outc[t][d].push_back(c);
        }
      });
      par([&](int d) {
        Sum& m = sh[d].m;
        m.keys.clear(); m.coef.clear(); m.hashes.clear(); m.rebuild_index();
        for (int t = 0; t < S; ++t)
          for (size_t j = 0; j < outc[t][d].size(); ++j) m.add_assign(&outk[t][d][j * m.NW], outc[t][d][j]);
      });
    } else {
      std::atomic<int> bad{0};
      par([&](int t) {
        Sum& m = sh[t].m;
        uint64_t before = m.term_updates;
        if (apply_one(m, op)) bad = 1;
        m.term_updates = before;
      });
      if (bad) return 1;
    }
  }
  // gather back (shard order, then insertion order within each shard)
  src.keys.clear(); src.coef.clear(); src.hashes.clear();
  for (int t = 0; t < S; ++t) {
    src.keys.insert(src.keys.end(), sh[t].m.keys.begin(), sh[t].m.keys.end());
    src.coef.insert(src.coef.end(), sh[t].m.coef.begin(), sh[t].m.coef.end());
  }
  src.rehash_all();
  src.term_updates += tu;
  return 0;
}

// ---------------------------------------------------------------------------
// Lossy sums: PauliSum<LossyPauliWord> (crates/ppvm-pauli-word/src/loss/data.rs).
// Keys are 3W words (x plane, z plane, loss plane).  Each op is the reference
// closure written out per term:
//   sum/clifford.rs:37-217 and phase/clifford.rs:16-180 (a lost qubit leaves
//     the word and sign unchanged),
//   sum/rot1.rs:9-128, sum/rot2.rs:17-184 (a lost qubit: no rotation, or the
//     one-qubit rotation on the surviving qubit),
//   sum/noise.rs:17-274 (channels; loss, correlated loss, reset loss),
// with the scratch-vector merge of map_insert (sum/data.rs:204-226,
// map/hashmap.rs:170-179) and the aux-map merge of map_insert_multiple
// (sum/data.rs:181-189,234-247, map/hashmap.rs:181-194,302-308).
enum LPauli { LP_I = 0, LP_X = 1, LP_Z = 2, LP_Y = 3, LP_L = 4 };

// LossyPauliWord::get / set (loss/data.rs:232-316)
int lget(const Sum& s, const uint64_t* k, uint32_t q) {
  if (s.lb(k, q)) return LP_L;
  return (int)s.xb(k, q) | ((int)s.zb(k, q) << 1);
}
void lset(const Sum& s, uint64_t* k, uint32_t q, int p) {
  s.setx(k, q, p == LP_X || p == LP_Y);
  s.setz(k, q, p == LP_Z || p == LP_Y);
  const uint64_t m = 1ull << (q & 63);
  uint64_t& w = k[2 * s.W + (q >> 6)];
  w = p == LP_L ? (w | m) : (w & ~m);
}

// map_insert_multiple (sum/data.rs:234-247, map/hashmap.rs:181-194, consume
// sum/data.rs:181-189 and map/hashmap.rs:302-308): f(key, &coef, kids,
// kid_coefs) returns the number of children (0..3).  The children are merged
// among themselves into an empty insertion-ordered aux map (`+=` in parent,
// then slot order); then the smaller map is consumed into the larger one with
// `+=`: if aux holds more keys than the sum, the sum's entries are added into
// aux in order and aux becomes the sum (aux order first, then the sum's other
// keys), otherwise aux's entries are added into the sum.  A key in both ends
// with p + S (S the children's sum), in either case.
template <class F>
void lmap_insert_multiple(Sum& s, F f) {
  Sum aux;
  aux.n = s.n;
  aux.W = s.W;
  aux.NW = s.NW;
  aux.lossy = s.lossy;
  aux.rebuild_index();
  std::vector<uint64_t> kids(3 * s.NW);
  double kv[3];
  const size_t n = s.len();
  s.term_updates += n;
  for (size_t i = 0; i < n; ++i) {
    const int m = f(s.key(i), s.coef[i], kids.data(), kv);
    for (int j = 0; j < m; ++j) aux.add_assign(&kids[j * s.NW], kv[j]);
  }
  if (aux.len() > s.len()) {
    for (size_t i = 0; i < s.len(); ++i) aux.add_assign(s.key(i), s.coef[i]);
    s.keys.swap(aux.keys);
    s.coef.swap(aux.coef);
    s.hashes.swap(aux.hashes);
    s.rebuild_index();
  } else {
    for (size_t j = 0; j < aux.len(); ++j) s.add_assign(aux.key(j), aux.coef[j]);
  }
}

// map_insert (sum/data.rs:204-226): f(key, &coef, kids, kid_coefs) returns the
// number of children (0 or 1 here); children are collected in a scratch
// vector and merged into the sum with `+=` after the pass, in order.
template <class F>
void lmap_insert(Sum& s, F f) {
  std::vector<uint64_t> ck;
  std::vector<double> cc;
  std::vector<uint64_t> kids(3 * s.NW);
  double kv[3];
  const size_t n = s.len();
  s.term_updates += n;
  for (size_t i = 0; i < n; ++i) {
    const int m = f(s.key(i), s.coef[i], kids.data(), kv);
    for (int j = 0; j < m; ++j) {
      ck.insert(ck.end(), kids.begin() + j * s.NW, kids.begin() + (j + 1) * s.NW);
      cc.push_back(kv[j]);
    }
  }
  for (size_t j = 0; j < cc.size(); ++j) s.add_assign(&ck[j * s.NW], cc[j]);
}

// rotate_1_map_insert_closure (rot1.rs:9-41) on a possibly lost qubit
int lrot1(const Sum& s, uint32_t code1, uint32_t q, const uint64_t* k, double& v, double sn, double cs,
          uint64_t* kid, double* kv) {
  if (s.lb(k, q)) return 0;
  const int eps = rot_child(s, code1, q, 0, k, kid);
  if (eps == 0) return 0;
  kv[0] = v * ((double)eps * sn);
  v *= cs;
  return 1;
}

void ltruncate(Sum& s) {
  std::vector<uint64_t> saved_k;
  std::vector<double> saved_c;
  const size_t P = s.preserve.size() / s.NW;
  for (size_t i = 0; i < s.len() && P; ++i)
    for (size_t p = 0; p < P; ++p)
      if (std::memcmp(s.key(i), &s.preserve[p * s.NW], s.NW * 8) == 0) {
        saved_k.insert(saved_k.end(), s.key(i), s.key(i) + s.NW);
        saved_c.push_back(s.coef[i]);
      }
  // CombinedStrategy(CombinedStrategy(CoefficientThreshold, MaxPauliWeight), MaxLossWeight)
  // (ppvm-python-native/src/interface.rs:58-65; strategy.rs:19-30,86-89,133,160-169)
  auto retain = [&](auto keep) {
    size_t o = 0;
    for (size_t i = 0; i < s.len(); ++i) {
      if (!keep(i)) continue;
      if (o != i) {
        std::memmove(s.key(o), s.key(i), s.NW * 8);
        s.coef[o] = s.coef[i];
        s.hashes[o] = s.hashes[i];
      }
      ++o;
    }
    s.keys.resize(o * s.NW);
    s.coef.resize(o);
    s.hashes.resize(o);
  };
  retain([&](size_t i) { return !(std::fabs(s.coef[i]) < s.tau); });
  if (s.maxw != UINT64_MAX)
    retain([&](size_t i) {  // weight counts X, Y, Z and L (loss/data.rs:134-176)
      uint64_t w = 0;
      const uint64_t* k = s.key(i);
      for (uint32_t p = 0; p < s.W; ++p) w += __builtin_popcountll(k[p] | k[s.W + p] | k[2 * s.W + p]);
      return w <= s.maxw;
    });
  if (s.maxlw != UINT64_MAX)
    retain([&](size_t i) {  // loss_weight (loss/data.rs:178-203)
      uint64_t w = 0;
      const uint64_t* k = s.key(i);
      for (uint32_t p = 0; p < s.W; ++p) w += __builtin_popcountll(k[2 * s.W + p]);
      return w <= s.maxlw;
    });
  s.rebuild_index();
  for (size_t j = 0; j < saved_c.size(); ++j) {  // sum/data.rs:305-310
    const uint64_t* k = &saved_k[j * s.NW];
    if (s.find(k, key_hash(k, s.NW)) < 0) s.add_assign(k, saved_c[j]);
  }
}

int lossy_apply_one(Sum& s, const ppvm_op& op) {
  const uint32_t code = op.code, a = op.q0, b = op.q1;
  auto qok = [&](uint32_t q) { return q < s.n; };
  const bool two = code == PPVM_OP_CNOT || code == PPVM_OP_CZ || code == PPVM_OP_CY || code == PPVM_OP_RXX ||
                   code == PPVM_OP_RYY || code == PPVM_OP_RZZ || code == PPVM_OP_DEPOLARIZE2 ||
                   code == PPVM_OP_PAULI_ERROR2 || code == PPVM_OP_CORRELATED_LOSS;
  if (code == PPVM_OP_NOP) return 0;
  if (code == PPVM_OP_TRUNCATE) {
    ltruncate(s);
    return 0;
  }
  if (!qok(a) || (two && (!qok(b) || a == b))) return 1;
  const uint32_t NW = s.NW;
  switch (code) {
    case PPVM_OP_X: case PPVM_OP_Y: case PPVM_OP_Z:  // scale (clifford.rs:37-66)
      s.term_updates += s.len();
      for (size_t i = 0; i < s.len(); ++i)
        if (!s.lb(s.key(i), a) && clifford_word(s, code, a, 0, s.key(i))) s.coef[i] *= -1.0;
      return 0;
    case PPVM_OP_H: case PPVM_OP_S: case PPVM_OP_S_DAG: case PPVM_OP_SQRT_X: case PPVM_OP_SQRT_X_DAG:
    case PPVM_OP_SQRT_Y: case PPVM_OP_SQRT_Y_DAG:  // map_add (clifford.rs:69-214)
      s.term_updates += s.len();
      for (size_t i = 0; i < s.len(); ++i) {
        if (s.lb(s.key(i), a)) continue;
        if (clifford_word(s, code, a, 0, s.key(i))) s.coef[i] = -s.coef[i];
      }
      s.rehash_all();
      return 0;
    case PPVM_OP_CNOT: case PPVM_OP_CZ: case PPVM_OP_CY:  // map_word! (clifford.rs:13-31)
      s.term_updates += s.len();
      for (size_t i = 0; i < s.len(); ++i) {
        if (s.lb(s.key(i), a) || s.lb(s.key(i), b)) continue;  // phase/clifford.rs:77,92,169
        if (clifford_word(s, code, a, b, s.key(i))) s.coef[i] = -s.coef[i];
      }
      s.rehash_all();
      return 0;
    case PPVM_OP_RX: case PPVM_OP_RY: case PPVM_OP_RZ: {
      double sn, cs;
      sincos(op.theta, &sn, &cs);
      lmap_insert(s, [&](const uint64_t* k, double& v, uint64_t* kid, double* kv) {
        return lrot1(s, code, a, k, v, sn, cs, kid, kv);
      });
      return 0;
    }
    case PPVM_OP_RXX: case PPVM_OP_RYY: case PPVM_OP_RZZ: {
      double sn, cs;
      sincos(op.theta, &sn, &cs);
      const uint32_t one = code == PPVM_OP_RXX ? PPVM_OP_RX : code == PPVM_OP_RYY ? PPVM_OP_RY : PPVM_OP_RZ;
      lmap_insert(s, [&](const uint64_t* k, double& v, uint64_t* kid, double* kv) {
        if (s.lb(k, a)) return lrot1(s, one, b, k, v, sn, cs, kid, kv);  // rot2.rs:28-35
        if (s.lb(k, b)) return lrot1(s, one, a, k, v, sn, cs, kid, kv);
        const int eps = rot_child(s, code, a, b, k, kid);
        if (eps == 0) return 0;
        kv[0] = v * ((double)eps * sn);
        v *= cs;
        return 1;
      });
      return 0;
    }
    case PPVM_OP_PAULI_ERROR: {  // noise.rs:17-39
      s.term_updates += s.len();
      const double xf = 1.0 - op.p[1] * 2.0 - op.p[2] * 2.0;
      const double zf = 1.0 - op.p[0] * 2.0 - op.p[1] * 2.0;
      const double yf = 1.0 - op.p[0] * 2.0 - op.p[2] * 2.0;
      for (size_t i = 0; i < s.len(); ++i) {
        const int p = lget(s, s.key(i), a);
        if (p == LP_X) s.coef[i] *= xf;
        else if (p == LP_Z) s.coef[i] *= zf;
        else if (p == LP_Y) s.coef[i] *= yf;
      }
      return 0;
    }
    case PPVM_OP_DEPOLARIZE1: {  // noise.rs:109-118
      s.term_updates += s.len();
      const double f = 1.0 - op.p[0] * (4.0 / 3.0);
      for (size_t i = 0; i < s.len(); ++i) {
        const int p = lget(s, s.key(i), a);
        if (p != LP_L && p != LP_I) s.coef[i] *= f;
      }
      return 0;
    }
    case PPVM_OP_DEPOLARIZE2: {  // noise.rs:120-131
      s.term_updates += s.len();
      const double f = 1.0 - op.p[0] * (16.0 / 15.0);
      for (size_t i = 0; i < s.len(); ++i) {
        const int pa = lget(s, s.key(i), a), pb = lget(s, s.key(i), b);
        if (pa == LP_L || pb == LP_L) continue;
        if (pa != LP_I || pb != LP_I) s.coef[i] *= f;
      }
      return 0;
    }
    case PPVM_OP_PAULI_ERROR2: {  // noise.rs:52-107: any L -> no noise
      s.term_updates += s.len();
      double fac[16];
      bool app[16];
      pe2_factors(op.p, fac, app);
      for (size_t i = 0; i < s.len(); ++i) {
        const uint64_t* k = s.key(i);
        if (s.lb(k, a) || s.lb(k, b)) continue;
        const int r = 4 * letter(s.xb(k, a), s.zb(k, a)) + letter(s.xb(k, b), s.zb(k, b));
        if (app[r]) s.coef[i] *= fac[r];
      }
      return 0;
    }
    case PPVM_OP_AMPLITUDE_DAMPING: {  // noise.rs:132-157
      const double g = op.p[0];
      lmap_insert(s, [&](const uint64_t* k, double& v, uint64_t* kid, double* kv) {
        const int p = lget(s, k, a);
        if (p == LP_X || p == LP_Y) {
          v *= std::sqrt(1.0 - g);
          return 0;
        }
        if (p != LP_Z) return 0;
        kv[0] = v * g;
        std::memcpy(kid, k, NW * 8);
        lset(s, kid, a, LP_I);
        v *= 1.0 - g;
        return 1;
      });
      return 0;
    }
    case PPVM_OP_LOSS: {  // noise.rs:165-180
      const double p = op.p[0];
      lmap_insert(s, [&](const uint64_t* k, double& v, uint64_t* kid, double* kv) {
        if (lget(s, k, a) == LP_L) {
          kv[0] = v * p;
          std::memcpy(kid, k, NW * 8);
          lset(s, kid, a, LP_I);
          return 1;
        }
        v *= 1.0 - p;
        return 0;
      });
      return 0;
    }
    case PPVM_OP_RESET_LOSS: {  // noise.rs:254-274
      lmap_insert(s, [&](const uint64_t* k, double& v, uint64_t* kid, double* kv) {
        const int p = lget(s, k, a);
        if (p == LP_L) {
          v *= 0.0;
          return 0;
        }
        if (p == LP_I || p == LP_Z) {
          std::memcpy(kid, k, NW * 8);
          lset(s, kid, a, LP_L);
          kv[0] = v;
          return 1;
        }
        return 0;
      });
      return 0;
    }
    case PPVM_OP_CORRELATED_LOSS: {  // noise.rs:182-248 (map_insert_multiple)
      const double* p = op.p;
      lmap_insert_multiple(s, [&](const uint64_t* k, double& v, uint64_t* kid, double* kv) {
        const int pa = lget(s, k, a), pb = lget(s, k, b);
        if (pa == LP_L && pb == LP_L) {
          const double v_il = v * p[2];
          std::memcpy(kid, k, NW * 8);  // (I, L)
          lset(s, kid, a, LP_I);
          lset(s, kid, b, LP_L);
          std::memcpy(kid + NW, k, NW * 8);  // (L, I)
          lset(s, kid + NW, a, LP_L);
          lset(s, kid + NW, b, LP_I);
          const double v_ii = v * p[0];
          std::memcpy(kid + 2 * NW, k, NW * 8);  // (I, I)
          lset(s, kid + 2 * NW, a, LP_I);
          lset(s, kid + 2 * NW, b, LP_I);
          kv[0] = v_il;
          kv[1] = v_il;
          kv[2] = v_ii;
          return 3;
        }
        if (pb == LP_L) {
          std::memcpy(kid, k, NW * 8);
          lset(s, kid, b, LP_I);
          kv[0] = v * p[1];
          v *= 1.0 - p[2];
          return 1;
        }
        if (pa == LP_L) {
          std::memcpy(kid, k, NW * 8);
          lset(s, kid, a, LP_I);
          kv[0] = v * p[1];
          v *= 1.0 - p[2];
          return 1;
        }
        v *= 1.0 - p[1] * 2.0 - p[0];
        return 0;
      });
      return 0;
    }
  }
  return 1;
}

}  // namespace

struct ppvo_sum : Sum {};

extern "C" {

ppvo_handle ppvo_create(uint32_t n_qubits, double min_abs_coeff, uint64_t max_pauli_weight) {
  ppvo_sum* s = new ppvo_sum();
  s->n = n_qubits;
  s->W = (n_qubits + 63) / 64;
  if (s->W == 0) s->W = 1;
  s->NW = 2 * s->W;
  s->tau = min_abs_coeff;
  s->maxw = max_pauli_weight;
  s->rebuild_index();
  return s;
}
void ppvo_destroy(ppvo_handle h) { delete h; }
ppvo_handle ppvo_clone(ppvo_handle h) { return new ppvo_sum(*h); }
uint32_t ppvo_words(ppvo_handle h) { return h->W; }
void ppvo_set_preserve(ppvo_handle h, const uint64_t* x, const uint64_t* z, size_t count) {
  h->preserve.clear();
  for (size_t i = 0; i < count; ++i) {
    h->preserve.insert(h->preserve.end(), x + i * h->W, x + (i + 1) * h->W);
    h->preserve.insert(h->preserve.end(), z + i * h->W, z + (i + 1) * h->W);
  }
}
void ppvo_add_terms(ppvo_handle h, const uint64_t* x, const uint64_t* z, const double* c,
                    size_t count) {
  std::vector<uint64_t> k(h->NW);
  for (size_t i = 0; i < count; ++i) {
    std::memcpy(k.data(), x + i * h->W, h->W * 8);
    std::memcpy(k.data() + h->W, z + i * h->W, h->W * 8);
    h->add_assign(k.data(), c[i]);
  }
}
int ppvo_apply(ppvo_handle h, const ppvm_op* ops, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (h->lossy ? lossy_apply_one(*h, ops[i]) : apply_one(*h, ops[i])) return 1;
  return 0;
}
double ppvo_overlap_with_zero(ppvo_handle h) { return overlap_zero(*h); }
int ppvo_trace(ppvo_handle h, const char* pattern, double* out) {
  std::vector<PatItem> pat;
  if (!pattern || !parse_pattern(pattern, pat)) return 1;
  *out = trace_pattern(*h, pat);
  return 0;
}
int ppvo_pattern_contains(ppvo_handle h, const char* pattern, const uint64_t* x, const uint64_t* z,
                          uint8_t* out, size_t count) {
  std::vector<PatItem> pat;
  if (!pattern || !parse_pattern(pattern, pat)) return 1;
  std::vector<uint64_t> k(h->NW);
  for (size_t i = 0; i < count; ++i) {
    std::memcpy(k.data(), x + i * h->W, h->W * 8);
    std::memcpy(k.data() + h->W, z + i * h->W, h->W * 8);
    out[i] = pattern_contains(*h, k.data(), pat) ? 1 : 0;
  }
  return 0;
}
double ppvo_overlap(ppvo_handle a, ppvo_handle b) {
  // sum over b's entries of b[k] * a.trace(k)  (sum/trace.rs:30-37)
  double acc = 0.0;
  for (size_t i = 0; i < b->len(); ++i) {
    const uint64_t* k = b->key(i);
    long j = a->find(k, key_hash(k, a->NW));
    double t = j >= 0 ? a->coef[j] : 0.0;
    acc += b->coef[i] * t;
  }
  return acc;
}
uint64_t ppvo_len(ppvo_handle h) { return h->len(); }
uint64_t ppvo_export(ppvo_handle h, uint64_t* x, uint64_t* z, double* c, size_t cap) {
  size_t n = std::min(cap, h->len());
  for (size_t i = 0; i < n; ++i) {
    std::memcpy(x + i * h->W, h->key(i), h->W * 8);
    std::memcpy(z + i * h->W, h->key(i) + h->W, h->W * 8);
    c[i] = h->coef[i];
  }
  return h->len();
}
uint64_t ppvo_term_updates(ppvo_handle h) { return h->term_updates; }

// Order-independent digest of the term set (IndexMap equality ignores order):
// per term, k = mix(...mix(mix(seed ^ x_0) ^ x_1)... ^ z_{W-1}) over the x then z
// words, and kc = mix(k ^ coefficient bits); out = {count, sum k, sum kc}, sums
// modulo 2^64.  mix is the SplitMix64 finaliser.
static inline uint64_t fp_mix(uint64_t v) {
  v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
  v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
  return v ^ (v >> 31);
}
void ppvo_fingerprint(ppvo_handle h, uint64_t out[3]) {
  uint64_t ks = 0, kcs = 0;
  const size_t n = h->len();
  for (size_t i = 0; i < n; ++i) {
    const uint64_t* k = h->key(i);
    uint64_t v = 0x9E3779B97F4A7C15ull;
    for (uint32_t p = 0; p < h->NW; ++p) v = fp_mix(v ^ k[p]);  // x, z (, loss) words
    uint64_t cb;
    std::memcpy(&cb, &h->coef[i], 8);
    ks += v;
    kcs += fp_mix(v ^ cb);
  }
  out[0] = n;
  out[1] = ks;
  out[2] = kcs;
}

ppvo_handle ppvo_lossy_create(uint32_t n_qubits, double min_abs_coeff, uint64_t max_pauli_weight,
                              uint64_t max_loss_weight) {
  ppvo_sum* s = ppvo_create(n_qubits, min_abs_coeff, max_pauli_weight);
  s->lossy = true;
  s->maxlw = max_loss_weight;
  s->NW = 3 * s->W;
  s->rebuild_index();
  return s;
}
static void lossy_key(ppvo_handle h, const uint64_t* x, const uint64_t* z, const uint64_t* l, size_t i,
                      uint64_t* k) {
  std::memcpy(k, x + i * h->W, h->W * 8);
  std::memcpy(k + h->W, z + i * h->W, h->W * 8);
  std::memcpy(k + 2 * h->W, l + i * h->W, h->W * 8);
}
void ppvo_lossy_set_preserve(ppvo_handle h, const uint64_t* x, const uint64_t* z, const uint64_t* l,
                             size_t count) {
  h->preserve.assign(count * h->NW, 0);
  for (size_t i = 0; i < count; ++i) lossy_key(h, x, z, l, i, &h->preserve[i * h->NW]);
}
void ppvo_lossy_add_terms(ppvo_handle h, const uint64_t* x, const uint64_t* z, const uint64_t* l,
                          const double* c, size_t count) {
  std::vector<uint64_t> k(h->NW);
  for (size_t i = 0; i < count; ++i) {
    lossy_key(h, x, z, l, i, k.data());
    h->add_assign(k.data(), c[i]);
  }
}
uint64_t ppvo_lossy_export(ppvo_handle h, uint64_t* x, uint64_t* z, uint64_t* l, double* c, size_t cap) {
  size_t n = std::min(cap, h->len());
  for (size_t i = 0; i < n; ++i) {
    std::memcpy(x + i * h->W, h->key(i), h->W * 8);
    std::memcpy(z + i * h->W, h->key(i) + h->W, h->W * 8);
    std::memcpy(l + i * h->W, h->key(i) + 2 * h->W, h->W * 8);
    c[i] = h->coef[i];
  }
  return h->len();
}
size_t ppvo_run_marked(ppvo_handle h, const ppvm_op* ops, size_t n, const uint64_t* marks,
                       size_t n_marks, uint64_t* out_len, double* out_ovl) {
  size_t m = 0;
  for (size_t i = 0; i < n; ++i) {
    if (apply_one(*h, ops[i])) return m;
    while (m < n_marks && marks[m] == i) {
      out_len[m] = h->len();
      out_ovl[m] = overlap_zero(*h);
      ++m;
    }
  }
  return m;
}
int ppvo_apply_sharded(ppvo_handle h, const ppvm_op* ops, size_t n, int threads) {
  return apply_sharded(*h, ops, n, threads);
}

}  // extern "C"
