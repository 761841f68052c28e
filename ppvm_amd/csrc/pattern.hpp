// pattern.hpp — Pauli pattern matcher shared by the engine and the lossy sums.
#pragma once

#include "algebra.hpp"

namespace ppvm {

// ---- trace(pattern): sum of coefficients of terms a Pauli pattern accepts ------
// The pattern is compiled on the host into items (grammar of
// crates/ppvm-pauli-word/src/pattern/parse.rs:30-143); the matcher follows the
// greedy left-to-right walk of pattern/contains.rs:48-151 over qubits 0..n-1.
constexpr int kMaxPatItems = 32;
enum PatKind : u32 { PK_POSITION = 0, PK_STAR = 1, PK_REPEAT = 2 };
struct PatternProg {
  u32 n_items;
  u32 nq;
  u32 kind[kMaxPatItems];
  u32 accept[kMaxPatItems];  // bit c set: Pauli code c = x | z << 1 accepted (I 0, X 1, Z 2, Y 3)
  u64 arg[kMaxPatItems];     // anchored qubit (position) or repeat count
};

// Host side, defined in engine.hip: pattern text -> PatternProg (PPVM_EINVAL
// with the message in ppvm_last_error() if it does not parse), and the
// two_qubit_pauli_error tables (crates/ppvm-pauli-sum/src/sum/noise.rs:56-106).
int compile_pattern(const char* text, u32 nq, PatternProg& prog);
extern const int kPe2Rows[16][8];
extern const int kCodeToLetter[4];

template <int W>
__device__ __forceinline__ u32 pauli_code(const u64 (&k)[2 * W], u64 q) {
  return xbit<W>(k, (u32)q) | (zbit<W>(k, (u32)q) << 1);
}
// true if any qubit in [lo, hi) carries a non-identity Pauli
template <int W>
__device__ __forceinline__ bool any_support(const u64 (&k)[2 * W], u64 lo, u64 hi) {
  u64 acc = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const u64 b0 = (u64)w * 64, b1 = b0 + 64;
    const u64 a = lo > b0 ? lo : b0, b = hi < b1 ? hi : b1;
    if (a >= b) continue;
    const u64 width = b - a;
    const u64 m = (width >= 64 ? ~0ull : ((1ull << width) - 1ull)) << (a - b0);
    acc |= (k[w] | k[W + w]) & m;
  }
  return acc != 0;
}

template <int W>
__device__ bool pattern_match(const u64 (&k)[2 * W], const PatternProg& p) {
  const u64 n = p.nq;
  u64 cur = 0;
  for (u32 t = 0; t < p.n_items; ++t) {
    const u32 acc = p.accept[t];
    const u64 arg = p.arg[t];
    bool next;
    if (p.kind[t] == PK_POSITION) {
      // identities up to the anchor, then the anchor's symbol; an anchor
      // already passed or beyond the word never matches
      if (arg < cur || arg >= n || any_support<W>(k, cur, arg)) return false;
      next = (acc >> pauli_code<W>(k, arg)) & 1u;
      cur = arg + 1;
    } else if (p.kind[t] == PK_STAR) {
      while (cur < n && ((acc >> pauli_code<W>(k, cur)) & 1u)) ++cur;
      next = true;
    } else {
      u64 matched = 0;
      bool decided = false;
      next = false;
      while (cur < n) {
        if ((acc >> pauli_code<W>(k, cur)) & 1u) {
          ++cur;
          if (++matched == arg) {
            next = true;
            decided = true;
            break;
          }
        } else {
          next = matched >= arg;
          decided = true;
          break;
        }
      }
      if (!decided) next = matched == arg;
    }
    if (!next) return false;
  }
  return !any_support<W>(k, cur, n);
}

}  // namespace ppvm
