// algebra.hpp — per-term Pauli algebra shared by the engine and batched kernels:
// worklist records, key loads, bit access on register-resident keys, rotation
// branching (anticommutation, child key, sign) and the map ops (Clifford / noise).
#pragma once

#include "device.hpp"

namespace ppvm {

// Worklist record of one anticommuting arena entry (16 bytes): its coefficient,
// arena index and the facts about its key that the emit and merge steps need,
// taken by the gather while the key is in registers (so no later kernel reads
// the key again except to confirm a candidate pair or to write a child).
struct alignas(16) WMRec {
  double c;
  u32 idx;
  u32 info;  // WM_* bits below
};
constexpr u32 WM_FLAGS = 3u;       // 0 unpaired, 1 merged with its partner, 2 dead entry
constexpr u32 WM_EPS_NEG = 4u;     // the child's sign eps is -1
constexpr u32 WM_ZA = 8u;          // z bit of the first generator qubit (damping: the Z member)
constexpr u32 WM_P_OVER = 16u;     // the key's weight exceeds the cap
constexpr u32 WM_P_PRES = 32u;     // the key is in the preserve set
constexpr u32 WM_C_OVER = 64u;     // the child key's weight exceeds the cap
constexpr u32 WM_C_PRES = 128u;    // the child key is in the preserve set

__device__ __forceinline__ void load_wm(const WMRec* __restrict__ wm, u64 r, double& c, u32& idx, u32& info) {
  const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(wm + r);
  c = __longlong_as_double((long long)v.x);
  idx = (u32)v.y;
  info = (u32)(v.y >> 32);
}
__device__ __forceinline__ void store_wm(WMRec* __restrict__ wm, u64 r, double c, u32 idx, u32 info) {
  *reinterpret_cast<ulonglong2*>(wm + r) = make_ulonglong2((u64)__double_as_longlong(c), (u64)idx | ((u64)info << 32));
}

// The truncation predicate drop_term (device.hpp) with the key's weight test
// and preserve-set membership taken from the record: drop = (|c| < tau or
// weight > cap) and not preserved.
__device__ __forceinline__ bool drop_from_info(double c, bool over, bool pres, const TruncSpec& t) {
  return (fabs(c) < t.tau || over) && !pres;
}

// Word access with a runtime index, without dynamic register indexing.
template <int N>
__device__ __forceinline__ u64 getw(const u64 (&k)[N], u32 i) {
  u64 r = 0;
#pragma unroll
  for (int j = 0; j < N; ++j) r = ((u32)j == i) ? k[j] : r;
  return r;
}
template <int N>
__device__ __forceinline__ void setw(u64 (&k)[N], u32 i, u64 v) {
#pragma unroll
  for (int j = 0; j < N; ++j) k[j] = ((u32)j == i) ? v : k[j];
}
template <int W>
__device__ __forceinline__ u32 xbit(const u64 (&k)[2 * W], u32 q) { return bit_of(getw<2 * W>(k, q >> 6), q); }
template <int W>
__device__ __forceinline__ u32 zbit(const u64 (&k)[2 * W], u32 q) { return bit_of(getw<2 * W>(k, W + (q >> 6)), q); }
template <int W>
__device__ __forceinline__ void put_x(u64 (&k)[2 * W], u32 q, u32 v) {
  u32 wi = q >> 6;
  u64 m = 1ull << (q & 63);
  u64 w = getw<2 * W>(k, wi);
  setw<2 * W>(k, wi, v ? (w | m) : (w & ~m));
}
template <int W>
__device__ __forceinline__ void put_z(u64 (&k)[2 * W], u32 q, u32 v) {
  u32 wi = W + (q >> 6);
  u64 m = 1ull << (q & 63);
  u64 w = getw<2 * W>(k, wi);
  setw<2 * W>(k, wi, v ? (w | m) : (w & ~m));
}

// ---- rotation algebra -------------------------------------------------------
// Anticommutation of P with G (symplectic product parity on the gate qubits).
template <int W>
__device__ __forceinline__ u32 rot_anti(const u64 (&k)[2 * W], const RotSpec& g) {
  if (g.mode == kBranchDamping) return xbit<W>(k, g.a) ^ 1u;  // I or Z on a
  u32 r = (xbit<W>(k, g.a) & g.gaz) ^ (zbit<W>(k, g.a) & g.gax);
  if (g.two) r ^= (xbit<W>(k, g.b) & g.gbz) ^ (zbit<W>(k, g.b) & g.gbx);
  return r;
}
// Child key P xor G (for damping: Z -> I on a).
template <int W>
__device__ __forceinline__ void rot_flip(u64 (&k)[2 * W], const RotSpec& g) {
  if (g.gax) put_x<W>(k, g.a, xbit<W>(k, g.a) ^ 1u);
  if (g.gaz) put_z<W>(k, g.a, zbit<W>(k, g.a) ^ 1u);
  if (g.two) {
    if (g.gbx) put_x<W>(k, g.b, xbit<W>(k, g.b) ^ 1u);
    if (g.gbz) put_z<W>(k, g.b, zbit<W>(k, g.b) ^ 1u);
  }
}
// Sign eps of the child of an anticommuting P:  U^dag P U = cos P + eps sin P',
// eps = -i * phase(P G).  On the single anticommuting qubit q the phase is +i
// when (P_q, G_q) is cyclic in (X, Y, Z) order, -i otherwise.  This reproduces
// the per-gate tables of crates/ppvm-pauli-sum/src/sum/rot1.rs:55-128 and
// sum/rot2.rs:63-184 (checked on all one- and two-qubit inputs by the tests).
__device__ __forceinline__ u32 pauli_rank(u32 x, u32 z) { return x ? (z ? 1u : 0u) : 2u; }  // X0 Y1 Z2
template <int W>
__device__ __forceinline__ double rot_eps(const u64 (&k)[2 * W], const RotSpec& g) {
  if (g.mode == kBranchDamping) return 1.0;
  u32 xa = xbit<W>(k, g.a), za = zbit<W>(k, g.a);
  u32 anti_a = (xa & g.gaz) ^ (za & g.gax);
  u32 px, pz, gx, gz;
  if (anti_a || !g.two) {
    px = xa; pz = za; gx = g.gax; gz = g.gaz;
  } else {
    px = xbit<W>(k, g.b); pz = zbit<W>(k, g.b); gx = g.gbx; gz = g.gbz;
  }
  u32 rp = pauli_rank(px, pz), rg = pauli_rank(gx, gz);
  return (rg == (rp + 1u) % 3u) ? 1.0 : -1.0;
}
// Pair key: both members of a pair {P, P xor G} map to the member whose pivot
// bit (first generator bit) is clear.
template <int W>
__device__ __forceinline__ void pair_key(const u64 (&k)[2 * W], const RotSpec& g, u64 (&r)[2 * W]) {
#pragma unroll
  for (int i = 0; i < 2 * W; ++i) r[i] = k[i];
  u32 piv = g.gax ? xbit<W>(k, g.a) : zbit<W>(k, g.a);
  if (piv) rot_flip<W>(r, g);
}

// Key of entry i from structure-of-arrays planes (plane p at p * cap); the
// batched kernel's arenas.
template <int W>
__device__ __forceinline__ void load_key(const u64* __restrict__ planes, u64 cap, u64 i, u64 (&k)[2 * W]) {
#pragma unroll
  for (int p = 0; p < 2 * W; ++p) k[p] = planes[(u64)p * cap + i];
}

// Key of entry i from an array-of-structures key arena (the single-sum
// engine): the 2W words of an entry are contiguous, x words then z words, so
// one entry is W 16-byte loads from one aligned 16W-byte block.
template <int W>
__device__ __forceinline__ void akey_load(const u64* __restrict__ keys, u64 i, u64 (&k)[2 * W]) {
  const ulonglong2* p = reinterpret_cast<const ulonglong2*>(keys + i * (2 * W));
#pragma unroll
  for (int q = 0; q < W; ++q) {
    const ulonglong2 v = p[q];
    k[2 * q] = v.x;
    k[2 * q + 1] = v.y;
  }
}
template <int W>
__device__ __forceinline__ void akey_store(u64* __restrict__ keys, u64 i, const u64 (&k)[2 * W]) {
  ulonglong2* p = reinterpret_cast<ulonglong2*>(keys + i * (2 * W));
#pragma unroll
  for (int q = 0; q < W; ++q) p[q] = make_ulonglong2(k[2 * q], k[2 * q + 1]);
}


// ---- map ops (Clifford conjugation, diagonal noise) on one key ------------------
// Bit maps and sign rules of crates/ppvm-pauli-sum/src/sum/clifford.rs:37-217 and
// crates/ppvm-pauli-word/src/phase/clifford.rs:16-180.
// Clifford conjugation of one key; returns 1 if the coefficient is negated.
template <int W>
__device__ __forceinline__ u32 clifford_apply(u64 (&k)[2 * W], u32 code, u32 a, u32 b) {
  const u32 x = xbit<W>(k, a), z = zbit<W>(k, a);
  u32 neg = 0;
  switch (code) {
    case M_X: neg = z; break;
    case M_Y: neg = x ^ z; break;
    case M_Z: neg = x; break;
    case M_H: put_x<W>(k, a, z); put_z<W>(k, a, x); neg = x & z; break;
    case M_S: put_z<W>(k, a, x ^ z); neg = x & (z ^ 1u); break;
    case M_S_DAG: put_z<W>(k, a, x ^ z); neg = x & z; break;
    case M_SQRT_X: put_x<W>(k, a, x ^ z); neg = x & z; break;
    case M_SQRT_X_DAG: put_x<W>(k, a, x ^ z); neg = (x ^ 1u) & z; break;
    case M_SQRT_Y: put_x<W>(k, a, z); put_z<W>(k, a, x); neg = (x ^ 1u) & z; break;
    case M_SQRT_Y_DAG: put_x<W>(k, a, z); put_z<W>(k, a, x); neg = x & (z ^ 1u); break;
    case M_CNOT: {
      const u32 xt = xbit<W>(k, b), zt = zbit<W>(k, b);
      neg = x & zt & ((xt == z) ? 1u : 0u);
      put_z<W>(k, a, z ^ zt);
      put_x<W>(k, b, xt ^ x);
      break;
    }
    case M_CZ: {
      const u32 xb = xbit<W>(k, b), zb = zbit<W>(k, b);
      neg = x & xb & (z ^ zb);
      put_z<W>(k, a, z ^ xb);
      put_z<W>(k, b, zb ^ x);
      break;
    }
    case M_CY: {
      const u32 xt = xbit<W>(k, b), zt = zbit<W>(k, b);
      neg = x & (xt ^ zt) & ((z ^ zt) ^ 1u);
      put_z<W>(k, a, z ^ xt ^ zt);
      put_x<W>(k, b, xt ^ x);
      put_z<W>(k, b, zt ^ x);
      break;
    }
    default: break;
  }
  return neg;
}

// Diagonal noise: coefficient *= factors[4*code_a + code_b] where the apply
// mask has that bit (code = x | z << 1; b = 0xffffffff for one-qubit channels).
template <int W>
__device__ __forceinline__ void scale_apply(const u64 (&k)[2 * W], double& c, u32 a, u32 b, u32 mask,
                                            const double* __restrict__ factors) {
  u32 idx = (xbit<W>(k, a) | (zbit<W>(k, a) << 1)) * 4u;
  if (b != 0xffffffffu) idx += xbit<W>(k, b) | (zbit<W>(k, b) << 1);
  if ((mask >> idx) & 1u) c = __dmul_rn(c, factors[idx]);
}

template <int W>
__device__ __forceinline__ void map_apply(u64 (&k)[2 * W], double& c, const MapOp& op,
                                          const MapProgram& prog) {
  const u32 code = op.code_flags & 0xffu;
  if (code == M_SCALE) {
    const u32 tb = (op.code_flags >> 16) & 0xffu;
    scale_apply<W>(k, c, op.a, op.b, prog.apply_mask[tb], prog.factors[tb]);
    return;
  }
  if (clifford_apply<W>(k, code, op.a, op.b)) c = -c;
}

}  // namespace ppvm
