// lossy.hip — PauliSum over lossy Pauli words on gfx950: the C ABI section
// "lossy sums" of include/ppvm_hip.h.
//
// A lossy word (crates/ppvm-pauli-word/src/loss/data.rs:33-45) has, besides
// the x and z planes, a loss plane l; a lost qubit reads as L (l = 1, x = z = 0,
// data.rs:242-256).  Every gate and channel of the reference acts on a term
// through its one or two target qubits only, so each op becomes a rule table
// over the local state of those qubits (I X Z Y L = 0..4 per qubit, 25 states
// for two qubits), built on the host:
//   * the parent's coefficient is multiplied by a factor (1.0, -1.0, cos, a
//     channel factor) and, for Cliffords, its local state replaced;
//   * up to three children are emitted (rotation partner, damping / loss /
//     reset-loss / correlated-loss branches), each the parent word with a new
//     local state and coefficient c_parent * factor.
// That is PauliSum::scale / map_add / map_insert / map_insert_multiple
// (crates/ppvm-pauli-sum/src/sum/data.rs:181-265) with the closures of
// sum/clifford.rs:37-217, sum/rot1.rs:9-143, sum/rot2.rs:17-184 and
// sum/noise.rs:17-274 evaluated on the local state.
//
// Children merge into the insertion-ordered map exactly as the reference's
// scratch vector does (children in parent order, then child slot order; an
// existing key takes `+=`, a new key is appended at its first child).  Rather
// than scattering children into a table, each target gathers: a target word T
// with local state t can only receive from the parents T|s' for the (s', j')
// pairs whose child j' has state t (at most three per op), so it looks those
// parents up in a hash index of the map, orders their contributions by
// (parent index, slot) and adds them in that order.  A child whose word is not
// in the map creates the new entry only if it is the first contributor; new
// entries are placed by a prefix scan over the children in order.  Results are
// bit-identical to the reference's sequential merge.
//
// Layout in HBM per handle: planes u64[3W][cap] x2 (x words, z words, loss
// words; the second copy is the compaction target of truncation), coef
// f64[cap] x2 (a branch op reads the old coefficients while writing the new),
// per-parent child counts and offsets, per-child flags and positions, and an
// open-addressed u32 index (word hash -> entry + 1) of 2^k >= 2 cap slots.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/ppvm_hip.h"
#include "pattern.hpp"

using namespace ppvm;

typedef uint8_t u8;

namespace {

thread_local std::string l_err;

int lfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  l_err = buf;
  return code;
}

#define LHIPCHK(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return lfail(PPVM_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

constexpr int kLB = 256;           // threads per block
constexpr u32 kLGrid = 2048;       // grid of the grid-stride kernels
constexpr u32 kLRedGrid = 512;     // blocks of the coefficient reductions
constexpr u64 kLScanChunk = 1024;  // items per block of the chunked passes and prefix scans
constexpr u32 kLScanMaxBlocks = 65536;
constexpr u64 kLMinCap = 1024;
constexpr u64 kLMaxCap = 1ull << 30;
constexpr u32 kStateL = 4;  // local states: I 0, X 1, Z 2, Y 3 (x | z << 1), L 4

// ---- rule tables ---------------------------------------------------------------------
struct LEntry {
  double pf;      // parent coefficient factor
  double cf[3];   // child coefficient factors (applied to the parent's old coefficient)
  u8 pst;         // parent's new local state
  u8 nch;         // children
  u8 cst[3];      // children's local states
};
struct LPre {     // parents whose child lands on a given local state
  u8 n;
  u8 s[3];        // parent local state
  u8 j[3];        // child slot
};
struct LTable {
  u32 a, b;       // target qubits (b = a for one-qubit ops)
  u32 two;        // 1: local state = 5 * state(a) + state(b)
  u32 branch;     // 1: the op emits children (map_insert); 0: scale / map_add
  u32 moves;      // 1: some parent changes its local state (Cliffords)
  u32 multi;      // 1: map_insert_multiple merge (children summed first, then added once)
  LEntry e[25];
  LPre pre[25];
};

// ---- device helpers ----------------------------------------------------------------------
template <int N>
__device__ __forceinline__ u64 lget(const u64 (&k)[N], u32 i) {
  u64 r = 0;
#pragma unroll
  for (int j = 0; j < N; ++j) r = ((u32)j == i) ? k[j] : r;
  return r;
}
template <int N>
__device__ __forceinline__ void lset(u64 (&k)[N], u32 i, u64 v) {
#pragma unroll
  for (int j = 0; j < N; ++j) k[j] = ((u32)j == i) ? v : k[j];
}

template <int W>
__device__ __forceinline__ void lload(const u64* __restrict__ planes, u64 cap, u64 i, u64 (&k)[3 * W]) {
#pragma unroll
  for (int p = 0; p < 3 * W; ++p) k[p] = planes[(u64)p * cap + i];
}
template <int W>
__device__ __forceinline__ void lstore(u64* __restrict__ planes, u64 cap, u64 i, const u64 (&k)[3 * W]) {
#pragma unroll
  for (int p = 0; p < 3 * W; ++p) planes[(u64)p * cap + i] = k[p];
}

template <int W>
__device__ __forceinline__ u32 lstate(const u64 (&k)[3 * W], u32 q) {
  const u32 w = q >> 6, s = q & 63;
  if ((lget<3 * W>(k, 2 * W + w) >> s) & 1ull) return kStateL;
  return (u32)((lget<3 * W>(k, w) >> s) & 1ull) | ((u32)((lget<3 * W>(k, W + w) >> s) & 1ull) << 1);
}
template <int W>
__device__ __forceinline__ void lput(u64 (&k)[3 * W], u32 q, u32 st) {
  const u32 w = q >> 6;
  const u64 m = 1ull << (q & 63);
  const bool x = st == 1u || st == 3u, z = st == 2u || st == 3u, l = st == kStateL;
  u64 v = lget<3 * W>(k, w);
  lset<3 * W>(k, w, x ? (v | m) : (v & ~m));
  v = lget<3 * W>(k, W + w);
  lset<3 * W>(k, W + w, z ? (v | m) : (v & ~m));
  v = lget<3 * W>(k, 2 * W + w);
  lset<3 * W>(k, 2 * W + w, l ? (v | m) : (v & ~m));
}
template <int W>
__device__ __forceinline__ u32 local_of(const u64 (&k)[3 * W], const LTable& tb) {
  const u32 sa = lstate<W>(k, tb.a);
  return tb.two ? sa * 5u + lstate<W>(k, tb.b) : sa;
}
template <int W>
__device__ __forceinline__ void put_local(u64 (&k)[3 * W], const LTable& tb, u32 t) {
  if (tb.two) {
    lput<W>(k, tb.a, t / 5u);
    lput<W>(k, tb.b, t % 5u);
  } else {
    lput<W>(k, tb.a, t);
  }
}

template <int W>
__device__ __forceinline__ u64 lhash(const u64 (&k)[3 * W]) {
  u64 h = 0x243f6a8885a308d3ull;
#pragma unroll
  for (int p = 0; p < 3 * W; ++p) {
    h = (h ^ k[p]) * 0x9e3779b97f4a7c15ull;
    h ^= h >> 31;
  }
  h *= 0xbf58476d1ce4e5b9ull;
  return h ^ (h >> 29);
}
template <int W>
__device__ __forceinline__ bool lkey_eq(const u64* __restrict__ planes, u64 cap, u64 i, const u64 (&k)[3 * W]) {
  bool eq = true;
#pragma unroll
  for (int p = 0; p < 3 * W; ++p) eq = eq && planes[(u64)p * cap + i] == k[p];
  return eq;
}
// Index lookup: entry index of word k, or ~0.
template <int W>
__device__ __forceinline__ u64 lfind(const u64* __restrict__ planes, u64 cap, const u32* __restrict__ table,
                                     u64 tmask, const u64 (&k)[3 * W]) {
  u64 pos = lhash<W>(k) & tmask;
  while (true) {
    const u32 s = table[pos];
    if (s == 0u) return ~0ull;
    if (lkey_eq<W>(planes, cap, s - 1u, k)) return s - 1u;
    pos = (pos + 1) & tmask;
  }
}

// ---- kernels --------------------------------------------------------------------------------
__global__ void k_lz_fill(u32* __restrict__ p, u64 n) {
  for (u64 i = (u64)blockIdx.x * kLB + threadIdx.x; i < n; i += (u64)gridDim.x * kLB) p[i] = 0u;
}

// Sum of v over the block, written by thread 0 to *dst.
__device__ __forceinline__ void block_total(u64 v, u64* dst) {
  __shared__ u64 red[kLB / 64];
#pragma unroll
  for (u32 d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
  if ((threadIdx.x & 63u) == 0u) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 t = 0;
#pragma unroll
    for (u32 w = 0; w < kLB / 64; ++w) t += red[w];
    *dst = t;
  }
}

// Words of the map are distinct, so an insert never compares keys.  The index
// has at least twice as many slots as words inserted, so every probe ends.
// With `only` set, just the entries with only[i] != 0 are inserted (the parents
// of a branch op: the entries with children).
template <int W>
__global__ __launch_bounds__(kLB) void k_lz_index(const u64* __restrict__ planes, u64 cap, u64 n,
                                                  u32* __restrict__ table, u64 tmask, const u32* __restrict__ only) {
  for (u64 i = (u64)blockIdx.x * kLB + threadIdx.x; i < n; i += (u64)gridDim.x * kLB) {
    if (only && only[i] == 0u) continue;
    u64 k[3 * W];
    lload<W>(planes, cap, i, k);
    u64 pos = lhash<W>(k) & tmask;
    while (atomicCAS(&table[pos], 0u, (u32)(i + 1)) != 0u) pos = (pos + 1) & tmask;
  }
}

// Parent pass: coefficient factor, local state change (map ops), child count
// (branch ops).  cin may equal cout.
// Block b covers [b * chunk, (b + 1) * chunk); for branch ops it also writes
// the block's child total to bsum[b] (the first pass of the offset scan).
template <int W>
__global__ __launch_bounds__(kLB) void k_lz_apply(u64* planes, u64 cap, u64 n, u64 chunk, const double* cin,
                                                  double* cout, LTable tb, u32* __restrict__ cnt,
                                                  u64* __restrict__ bsum) {
  const u64 lo = (u64)blockIdx.x * chunk, hi = min(n, lo + chunk);
  u64 acc = 0;
  for (u64 i = lo + threadIdx.x; i < hi; i += kLB) {
    u64 k[3 * W];
    lload<W>(planes, cap, i, k);
    const u32 t = local_of<W>(k, tb);
    const LEntry& e = tb.e[t];
    cout[i] = __dmul_rn(cin[i], e.pf);
    if (tb.branch) {
      cnt[i] = e.nch;
      acc += e.nch;
    } else if (tb.moves && e.pst != t) {
      put_local<W>(k, tb, e.pst);
      lstore<W>(planes, cap, i, k);
    }
  }
  if (tb.branch) block_total(acc, &bsum[blockIdx.x]);
}

struct Contrib {
  u64 p;      // parent entry
  u32 j;      // child slot
  u32 s;      // parent local state
};

// Contributors to target word T (local state t): parents T|s' present in the map.
template <int W>
__device__ __forceinline__ int gather(const u64* __restrict__ planes, u64 cap, const u32* __restrict__ table,
                                      u64 tmask, const LTable& tb, const u64 (&T)[3 * W], u32 t,
                                      Contrib (&out)[3]) {
  const LPre& pr = tb.pre[t];
  int m = 0;
  for (u32 r = 0; r < pr.n; ++r) {
    u64 P[3 * W];
#pragma unroll
    for (int q = 0; q < 3 * W; ++q) P[q] = T[q];
    put_local<W>(P, tb, pr.s[r]);
    const u64 p = lfind<W>(planes, cap, table, tmask, P);
    if (p == ~0ull) continue;
    Contrib c{p, pr.j[r], pr.s[r]};
    int q = m++;
    while (q > 0 && (out[q - 1].p > c.p || (out[q - 1].p == c.p && out[q - 1].j > c.j))) {
      out[q] = out[q - 1];
      --q;
    }
    out[q] = c;
  }
  return m;
}

// Children pass 1: flag the children that create a new entry (word absent from
// the map, and this child is its first contributor in (parent, slot) order).
// Whether the word is in the map: present[child] as k_lz_merge marked it, or,
// without `present`, a lookup in an index of every word.
template <int W>
__global__ __launch_bounds__(kLB) void k_lz_child_flags(const u64* __restrict__ planes, u64 cap, u64 n, LTable tb,
                                                        const u32* __restrict__ cnt, const u64* __restrict__ off,
                                                        const u32* __restrict__ table, u64 tmask,
                                                        const u32* __restrict__ present, u32* __restrict__ flag) {
  for (u64 i = (u64)blockIdx.x * kLB + threadIdx.x; i < n; i += (u64)gridDim.x * kLB) {
    const u32 nch = cnt[i];
    if (nch == 0u) continue;
    u64 k[3 * W];
    lload<W>(planes, cap, i, k);
    const LEntry& e = tb.e[local_of<W>(k, tb)];
    const u64 base = off[i];
    for (u32 j = 0; j < nch; ++j) {
      u64 T[3 * W];
#pragma unroll
      for (int q = 0; q < 3 * W; ++q) T[q] = k[q];
      put_local<W>(T, tb, e.cst[j]);
      u32 f = 0u;
      if (present ? present[base + j] == 0u : lfind<W>(planes, cap, table, tmask, T) == ~0ull) {
        Contrib c[3];
        const int m = gather<W>(planes, cap, table, tmask, tb, T, e.cst[j], c);
        f = (m > 0 && c[0].p == i && c[0].j == j) ? 1u : 0u;
      }
      flag[base + j] = f;
    }
  }
}

// Children pass 2: write each new entry (word + ordered sum of its contributions)
// at n + its rank among the new entries.
template <int W>
__global__ __launch_bounds__(kLB) void k_lz_child_write(u64* planes, u64 cap, u64 n, LTable tb,
                                                        const u32* __restrict__ cnt, const u64* __restrict__ off,
                                                        const u32* __restrict__ table, u64 tmask,
                                                        const u32* __restrict__ flag, const u64* __restrict__ pos,
                                                        const double* __restrict__ cin, double* __restrict__ cout) {
  for (u64 i = (u64)blockIdx.x * kLB + threadIdx.x; i < n; i += (u64)gridDim.x * kLB) {
    const u32 nch = cnt[i];
    if (nch == 0u) continue;
    u64 k[3 * W];
    lload<W>(planes, cap, i, k);
    const LEntry& e = tb.e[local_of<W>(k, tb)];
    const u64 base = off[i];
    for (u32 j = 0; j < nch; ++j) {
      if (!flag[base + j]) continue;
      u64 T[3 * W];
#pragma unroll
      for (int q = 0; q < 3 * W; ++q) T[q] = k[q];
      put_local<W>(T, tb, e.cst[j]);
      Contrib c[3];
      const int m = gather<W>(planes, cap, table, tmask, tb, T, e.cst[j], c);
      double v = __dmul_rn(cin[c[0].p], tb.e[c[0].s].cf[c[0].j]);
      for (int r = 1; r < m; ++r) v = __dadd_rn(v, __dmul_rn(cin[c[r].p], tb.e[c[r].s].cf[c[r].j]));
      const u64 dst = n + pos[base + j];
      lstore<W>(planes, cap, dst, T);
      cout[dst] = v;
    }
  }
}

// Existing entries that children land on: `+=` of the contributions in order.
// With `present`, each contribution's child is marked there (its word is in
// the map), at off[parent] + slot.
template <int W>
__global__ __launch_bounds__(kLB) void k_lz_merge(const u64* __restrict__ planes, u64 cap, u64 n, LTable tb,
                                                  const u32* __restrict__ table, u64 tmask,
                                                  const double* __restrict__ cin, double* __restrict__ cout,
                                                  const u64* __restrict__ off, u32* __restrict__ present) {
  for (u64 i = (u64)blockIdx.x * kLB + threadIdx.x; i < n; i += (u64)gridDim.x * kLB) {
    u64 k[3 * W];
    lload<W>(planes, cap, i, k);
    const u32 t = local_of<W>(k, tb);
    if (tb.pre[t].n == 0) continue;
    Contrib c[3];
    const int m = gather<W>(planes, cap, table, tmask, tb, k, t, c);
    if (m == 0) continue;
    if (present)
      for (int r = 0; r < m; ++r) present[off[c[r].p] + c[r].j] = 1u;
    if (tb.multi) {
      // map_insert_multiple: the children meet in the aux map first, and their
      // sum is added to the entry once (sum/data.rs:234-247, consume :181-189)
      double sum = __dmul_rn(cin[c[0].p], tb.e[c[0].s].cf[c[0].j]);
      for (int r = 1; r < m; ++r) sum = __dadd_rn(sum, __dmul_rn(cin[c[r].p], tb.e[c[r].s].cf[c[r].j]));
      cout[i] = __dadd_rn(cout[i], sum);
    } else {
      // map_insert: the scratch vector's children are added one by one
      // (sum/data.rs:204-226)
      double v = cout[i];
      for (int r = 0; r < m; ++r) v = __dadd_rn(v, __dmul_rn(cin[c[r].p], tb.e[c[r].s].cf[c[r].j]));
      cout[i] = v;
    }
  }
}

// map_insert_multiple, order (sum/data.rs:181-189): when the aux map of
// distinct children holds more keys than the sum, the result lists aux's keys
// in first-occurrence order, then the sum's keys that received no child, in
// their order.  The three kernels below build that order into the other arena.
//
// Entries that received no child: flag[e] = 1.
template <int W>
__global__ __launch_bounds__(kLB) void k_lz_unhit(const u64* __restrict__ planes, u64 cap, u64 n, LTable tb,
                                                  const u32* __restrict__ table, u64 tmask, u32* __restrict__ flag) {
  for (u64 i = (u64)blockIdx.x * kLB + threadIdx.x; i < n; i += (u64)gridDim.x * kLB) {
    u64 k[3 * W];
    lload<W>(planes, cap, i, k);
    const u32 t = local_of<W>(k, tb);
    u32 f = 1u;
    if (tb.pre[t].n) {
      Contrib c[3];
      f = gather<W>(planes, cap, table, tmask, tb, k, t, c) == 0 ? 1u : 0u;
    }
    flag[i] = f;
  }
}

// Children that are the first contributor of their word: flag[child] = 1.
template <int W>
__global__ __launch_bounds__(kLB) void k_lz_first(const u64* __restrict__ planes, u64 cap, u64 n, LTable tb,
                                                  const u32* __restrict__ cnt, const u64* __restrict__ off,
                                                  const u32* __restrict__ table, u64 tmask, u32* __restrict__ flag) {
  for (u64 i = (u64)blockIdx.x * kLB + threadIdx.x; i < n; i += (u64)gridDim.x * kLB) {
    const u32 nch = cnt[i];
    if (nch == 0u) continue;
    u64 k[3 * W];
    lload<W>(planes, cap, i, k);
    const LEntry& e = tb.e[local_of<W>(k, tb)];
    const u64 base = off[i];
    for (u32 j = 0; j < nch; ++j) {
      u64 T[3 * W];
#pragma unroll
      for (int q = 0; q < 3 * W; ++q) T[q] = k[q];
      put_local<W>(T, tb, e.cst[j]);
      Contrib c[3];
      const int m = gather<W>(planes, cap, table, tmask, tb, T, e.cst[j], c);
      flag[base + j] = (m > 0 && c[0].p == i && c[0].j == j) ? 1u : 0u;
    }
  }
}

// Writes the reordered sum into (qp, qc): the child word of first-occurrence
// rank r at r (its merged value: the entry's, or the new entry's at n + its
// rank among the new entries), then the entries without children at A + their
// rank among them (urank: exclusive ranks, R of them in all).
template <int W>
__global__ __launch_bounds__(kLB) void k_lz_reorder(const u64* __restrict__ planes, u64 cap, u64 n, LTable tb,
                                                    const u32* __restrict__ cnt, const u64* __restrict__ off,
                                                    const u32* __restrict__ table, u64 tmask,
                                                    const u32* __restrict__ first, const u64* __restrict__ arank,
                                                    const u32* __restrict__ isnew, const u64* __restrict__ nrank,
                                                    const u64* __restrict__ urank, u64 R, u64 A,
                                                    const double* __restrict__ cval, u64* __restrict__ qp,
                                                    double* __restrict__ qc) {
  for (u64 i = (u64)blockIdx.x * kLB + threadIdx.x; i < n; i += (u64)gridDim.x * kLB) {
    u64 k[3 * W];
    lload<W>(planes, cap, i, k);
    // exclusive ranks of the entries without a child: entry i has none iff the rank steps after it
    const bool unhit = (i + 1 < n ? urank[i + 1] : R) != urank[i];
    if (unhit) {
      const u64 dst = A + urank[i];
      lstore<W>(qp, cap, dst, k);
      qc[dst] = cval[i];
    }
    const u32 nch = cnt[i];
    if (nch == 0u) continue;
    const LEntry& e = tb.e[local_of<W>(k, tb)];
    const u64 base = off[i];
    for (u32 j = 0; j < nch; ++j) {
      if (!first[base + j]) continue;
      u64 T[3 * W];
#pragma unroll
      for (int q = 0; q < 3 * W; ++q) T[q] = k[q];
      put_local<W>(T, tb, e.cst[j]);
      const u64 src = isnew[base + j] ? n + nrank[base + j] : lfind<W>(planes, cap, table, tmask, T);
      const u64 dst = arank[base + j];
      lstore<W>(qp, cap, dst, T);
      qc[dst] = cval[src];
    }
  }
}

// Truncation flags: 1 = kept, 1 << 32 = dropped but preserved (restored at the
// end in map order, sum/data.rs:296-310), 0 = dropped.
// Chunked like k_lz_apply; bsum[b] = the block's packed flag total, also
// added to *total (zeroed by the host before the launch).
template <int W>
__global__ __launch_bounds__(kLB) void k_lz_keep(const u64* __restrict__ planes, u64 cap, u64 n, u64 chunk,
                                                 const double* __restrict__ coef, double tau, u64 maxw, u64 maxlw,
                                                 const u64* __restrict__ preserve, u32 n_preserve,
                                                 u64* __restrict__ flags, u64* __restrict__ bsum,
                                                 u64* __restrict__ total) {
  const u64 lo = (u64)blockIdx.x * chunk, hi = min(n, lo + chunk);
  u64 acc = 0;
  for (u64 i = lo + threadIdx.x; i < hi; i += kLB) {
    u64 k[3 * W];
    lload<W>(planes, cap, i, k);
    const double c = coef[i];
    bool keep = !(fabs(c) < tau);
    if (keep && maxw != ~0ull) {
      u64 w = 0;
#pragma unroll
      for (int p = 0; p < W; ++p) w += __popcll(k[p] | k[W + p] | k[2 * W + p]);
      keep = w <= maxw;
    }
    if (keep && maxlw != ~0ull) {
      u64 w = 0;
#pragma unroll
      for (int p = 0; p < W; ++p) w += __popcll(k[2 * W + p]);
      keep = w <= maxlw;
    }
    u64 f = keep ? 1ull : 0ull;
    if (!keep) {
      for (u32 r = 0; r < n_preserve; ++r) {
        bool eq = true;
#pragma unroll
        for (int p = 0; p < 3 * W; ++p) eq = eq && preserve[(u64)r * 3 * W + p] == k[p];
        if (eq) {
          f = 1ull << 32;
          break;
        }
      }
    }
    flags[i] = f;
    acc += f;
  }
  block_total(acc, &bsum[blockIdx.x]);
  if (threadIdx.x == 0) atomicAdd((unsigned long long*)total, (unsigned long long)bsum[blockIdx.x]);
}

template <int W>
__global__ __launch_bounds__(kLB) void k_lz_compact(const u64* __restrict__ pin, const double* __restrict__ cin,
                                                    u64* __restrict__ pout, double* __restrict__ cout, u64 cap, u64 n,
                                                    const u64* __restrict__ flags, const u64* __restrict__ off,
                                                    u64 kept) {
  for (u64 i = (u64)blockIdx.x * kLB + threadIdx.x; i < n; i += (u64)gridDim.x * kLB) {
    const u64 f = flags[i];
    if (!f) continue;
    const u64 dst = (f == 1ull) ? (off[i] & 0xffffffffull) : kept + (off[i] >> 32);
#pragma unroll
    for (int p = 0; p < 3 * W; ++p) pout[(u64)p * cap + dst] = pin[(u64)p * cap + i];
    cout[dst] = cin[i];
  }
}

// Block-partial sums of the coefficients of matching terms: mode 0 = every
// qubit I or Z; mode 1 = accepted by the pattern (and no qubit lost).
template <int W>
__global__ __launch_bounds__(kLB) void k_lz_reduce(const u64* __restrict__ planes, u64 cap, u64 n,
                                                   const double* __restrict__ coef, u32 mode, PatternProg prog,
                                                   double* __restrict__ partials) {
  __shared__ double sd[kLB];
  double acc = 0.0;
  for (u64 i = (u64)blockIdx.x * kLB + threadIdx.x; i < n; i += (u64)gridDim.x * kLB) {
    u64 k[3 * W];
    lload<W>(planes, cap, i, k);
    u64 lost = 0;
#pragma unroll
    for (int p = 0; p < W; ++p) lost |= k[2 * W + p];
    if (lost) continue;
    bool hit;
    if (mode == 0) {
      u64 x = 0;
#pragma unroll
      for (int p = 0; p < W; ++p) x |= k[p];
      hit = x == 0;
    } else {
      u64 kz[2 * W];
#pragma unroll
      for (int p = 0; p < 2 * W; ++p) kz[p] = k[p];
      hit = pattern_match<W>(kz, prog);
    }
    if (hit) acc = __dadd_rn(acc, coef[i]);
  }
  sd[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kLB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sd[threadIdx.x] = __dadd_rn(sd[threadIdx.x], sd[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sd[0];
}

// Σ over b's terms of c_b(k) * c_a(k), a indexed in `table`.
template <int W>
__global__ __launch_bounds__(kLB) void k_lz_overlap(const u64* __restrict__ pa, const double* __restrict__ ca,
                                                    u64 cap_a, const u32* __restrict__ table, u64 tmask,
                                                    const u64* __restrict__ pb, const double* __restrict__ cb,
                                                    u64 cap_b, u64 nb, double* __restrict__ partials) {
  __shared__ double sd[kLB];
  double acc = 0.0;
  for (u64 i = (u64)blockIdx.x * kLB + threadIdx.x; i < nb; i += (u64)gridDim.x * kLB) {
    u64 k[3 * W];
    lload<W>(pb, cap_b, i, k);
    const u64 j = lfind<W>(pa, cap_a, table, tmask, k);
    if (j != ~0ull) acc = __dadd_rn(acc, __dmul_rn(cb[i], ca[j]));
  }
  sd[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kLB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sd[threadIdx.x] = __dadd_rn(sd[threadIdx.x], sd[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sd[0];
}

// ---- exclusive prefix scan of n values (u32 or u64) into u64 offsets -----------------------
template <typename T>
__global__ __launch_bounds__(kLB) void k_lz_bsum(const T* __restrict__ in, u64 n, u64 chunk, u64* __restrict__ bsum) {
  __shared__ u64 s[kLB];
  const u64 lo = (u64)blockIdx.x * chunk, hi = min(n, lo + chunk);
  u64 acc = 0;
  for (u64 i = lo + threadIdx.x; i < hi; i += kLB) acc += (u64)in[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int d = kLB / 2; d > 0; d >>= 1) {
    if ((int)threadIdx.x < d) s[threadIdx.x] += s[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = s[0];
}

// One block: exclusive scan of G block sums in place; total to *total.
__global__ __launch_bounds__(1024) void k_lz_bscan(u64* __restrict__ bsum, u32 G, u64* __restrict__ total) {
  __shared__ u64 s[1024];
  u64 carry = 0;
  for (u32 base = 0; base < G; base += 1024) {
    const u32 i = base + threadIdx.x;
    const u64 v = i < G ? bsum[i] : 0ull;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
      const u64 add = (int)threadIdx.x >= d ? s[threadIdx.x - d] : 0ull;
      __syncthreads();
      s[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < G) bsum[i] = carry + s[threadIdx.x] - v;
    const u64 tile = s[1023];
    __syncthreads();
    carry += tile;
  }
  if (threadIdx.x == 0) *total = carry;
}

// Block b scans its chunk [b * chunk, ...) starting from the exclusive block
// offset bsum[b]: tiles of kPer items per thread are staged through LDS (one
// pad slot per kPer so the per-thread runs hit distinct banks), summed per
// thread, combined across the wave with shuffles and across the four waves
// through LDS, and written back coalesced.
constexpr u32 kPer = 4;
constexpr u32 kTile = kLB * kPer;
static_assert(kLScanChunk % kTile == 0, "chunks are whole scan tiles");
__device__ __forceinline__ u32 tpos(u32 e) { return e + e / kPer; }

template <typename T>
__global__ __launch_bounds__(kLB) void k_lz_scan(const T* __restrict__ in, u64 n, u64 chunk,
                                                 const u64* __restrict__ bsum, u64* __restrict__ out) {
  __shared__ u64 tile[kTile + kTile / kPer];
  __shared__ u64 wsum[kLB / 64];
  const u32 t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const u64 lo = (u64)blockIdx.x * chunk, hi = min(n, lo + chunk);
  u64 carry = bsum[blockIdx.x];
  for (u64 base = lo; base < hi; base += kTile) {
#pragma unroll
    for (u32 j = 0; j < kPer; ++j) {
      const u32 e = j * kLB + t;
      const u64 i = base + e;
      tile[tpos(e)] = i < hi ? (u64)in[i] : 0ull;
    }
    __syncthreads();
    u64 loc[kPer];
    u64 sum = 0;
#pragma unroll
    for (u32 j = 0; j < kPer; ++j) {
      const u64 v = tile[tpos(t * kPer + j)];
      loc[j] = sum;
      sum += v;
    }
    u64 incl = sum;
#pragma unroll
    for (u32 d = 1; d < 64; d <<= 1) {
      const u64 y = __shfl_up(incl, d, 64);
      if (lane >= d) incl += y;
    }
    if (lane == 63u) wsum[wave] = incl;
    __syncthreads();
    u64 woff = 0, ttot = 0;
#pragma unroll
    for (u32 w = 0; w < kLB / 64; ++w) {
      const u64 v = wsum[w];
      woff += w < wave ? v : 0ull;
      ttot += v;
    }
    const u64 off = carry + woff + (incl - sum);
#pragma unroll
    for (u32 j = 0; j < kPer; ++j) tile[tpos(t * kPer + j)] = off + loc[j];
    __syncthreads();
#pragma unroll
    for (u32 j = 0; j < kPer; ++j) {
      const u32 e = j * kLB + t;
      const u64 i = base + e;
      if (i < hi) out[i] = tile[tpos(e)];
    }
    carry += ttot;
    __syncthreads();
  }
}

// ---- host ------------------------------------------------------------------------------------
u64 lpow2(u64 v) {
  u64 p = 1;
  while (p < v) p <<= 1;
  return p;
}
u32 lgrid(u64 n) { return (u32)std::max<u64>(1, std::min<u64>((n + kLB - 1) / kLB, kLGrid)); }

}  // namespace

struct ppvm_lossy_sum {
  u32 n = 0, W = 1;  // W: words per plane on the device (1, 2, 4 or 8)
  u32 Wa = 1;        // words per plane at the ABI: ceil(n / 64)
  double tau = 0.0;
  u64 maxw = ~0ull, maxlw = ~0ull;
  int device = 0;
  hipStream_t stream = nullptr;
  bool broken = false;
  u64 cap = 0, len = 0, tcap = 0;
  int cur = 0;
  u64* planes[2] = {nullptr, nullptr};
  double* coef[2] = {nullptr, nullptr};
  u32* cnt = nullptr;     // per entry: child count
  u64* off = nullptr;     // per entry: child offset / truncation flags
  u64* off2 = nullptr;    // per entry: truncation offsets
  u32* flag = nullptr;    // per child: creates a new entry
  u64* pos = nullptr;     // per child: rank among new entries
  u32* flag2 = nullptr;   // per child: first contributor of its word (map_insert_multiple order)
  u64* pos2 = nullptr;    // per child: rank among those
  u32* table = nullptr;   // word index
  u64* bsum = nullptr;    // scan block sums
  u64* dtot = nullptr;    // scan total (device)
  u64* htot = nullptr;    // scan total (pinned host)
  double* partials = nullptr;
  u64* preserve = nullptr;
  u32 n_preserve = 0;
  // Local states some term may hold on each qubit (a superset, widened by
  // every op that runs): bit q of smask[s] for state s (I X Z Y L = 0..4).
  // An op whose rule table leaves every such state unchanged is skipped.
  std::vector<u64> smask[5];
  bool clean = true;  // a truncation ran after the last change (another one drops nothing)
};

namespace {

void lfree(ppvm_lossy_sum* s) {
  for (int b = 0; b < 2; ++b) {
    if (s->planes[b]) (void)hipFree(s->planes[b]);
    if (s->coef[b]) (void)hipFree(s->coef[b]);
    s->planes[b] = nullptr;
    s->coef[b] = nullptr;
  }
  void* bufs[] = {s->cnt, s->off, s->off2, s->flag, s->pos, s->flag2, s->pos2, s->table};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  s->cnt = nullptr;
  s->off = s->off2 = s->pos = s->pos2 = nullptr;
  s->flag = s->flag2 = s->table = nullptr;
}

// Allocates arenas of `cap` entries, keeping the current `len` entries.
int lgrow(ppvm_lossy_sum* s, u64 want) {
  if (want <= s->cap) return PPVM_OK;
  if (want > kLMaxCap) return lfail(PPVM_ENOMEM, "lossy sum needs %llu entries (limit %llu)",
                                    (unsigned long long)want, (unsigned long long)kLMaxCap);
  u64 cap = std::max<u64>(kLMinCap, s->cap);
  while (cap < want) cap *= 2;
  cap = std::min(cap, kLMaxCap);
  const u64 np = 3ull * s->W;
  ppvm_lossy_sum t{};
  bool ok = true;
  for (int b = 0; b < 2 && ok; ++b) {
    ok = hipMalloc(&t.planes[b], np * cap * 8) == hipSuccess && hipMalloc(&t.coef[b], cap * 8) == hipSuccess;
  }
  const u64 tcap = lpow2(2 * cap);
  ok = ok && hipMalloc(&t.cnt, cap * 4) == hipSuccess && hipMalloc(&t.off, cap * 8) == hipSuccess &&
       hipMalloc(&t.off2, cap * 8) == hipSuccess && hipMalloc(&t.flag, cap * 4) == hipSuccess &&
       hipMalloc(&t.pos, cap * 8) == hipSuccess && hipMalloc(&t.flag2, cap * 4) == hipSuccess &&
       hipMalloc(&t.pos2, cap * 8) == hipSuccess && hipMalloc(&t.table, tcap * 4) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    lfree(&t);
    return lfail(PPVM_ENOMEM, "device allocation of a lossy arena of %llu entries failed",
                 (unsigned long long)cap);
  }
  hipError_t e = hipSuccess;
  for (u64 p = 0; p < np && s->len && e == hipSuccess; ++p)
    e = hipMemcpyAsync(t.planes[0] + p * cap, s->planes[s->cur] + p * s->cap, s->len * 8, hipMemcpyDeviceToDevice,
                       s->stream);
  if (e == hipSuccess && s->len)
    e = hipMemcpyAsync(t.coef[0], s->coef[s->cur], s->len * 8, hipMemcpyDeviceToDevice, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) {
    lfree(&t);
    return lfail(PPVM_EHIP, "copying the lossy arena failed: %s", hipGetErrorString(e));
  }
  lfree(s);
  for (int b = 0; b < 2; ++b) {
    s->planes[b] = t.planes[b];
    s->coef[b] = t.coef[b];
  }
  s->cnt = t.cnt;
  s->off = t.off;
  s->off2 = t.off2;
  s->flag = t.flag;
  s->pos = t.pos;
  s->flag2 = t.flag2;
  s->pos2 = t.pos2;
  s->table = t.table;
  s->cap = cap;
  s->tcap = tcap;
  s->cur = 0;
  return PPVM_OK;
}

// Items per block of the chunked passes: kLScanChunk, doubled until the blocks fit bsum.
u64 scan_chunk(u64 n) {
  u64 chunk = kLScanChunk;
  while ((n + chunk - 1) / chunk > kLScanMaxBlocks) chunk *= 2;
  return chunk;
}
u32 scan_blocks(u64 n, u64 chunk) { return (u32)std::max<u64>(1, (n + chunk - 1) / chunk); }

// Block sums (already in s->bsum) -> exclusive block offsets; returns the
// total (read back to the host, one synchronisation) unless total is null.
int lscan_offsets(ppvm_lossy_sum* s, u32 G, u64* total) {
  hipLaunchKernelGGL(k_lz_bscan, dim3(1), dim3(1024), 0, s->stream, s->bsum, G, s->dtot);
  LHIPCHK(hipGetLastError());
  if (!total) return PPVM_OK;
  LHIPCHK(hipMemcpyAsync(s->htot, s->dtot, 8, hipMemcpyDeviceToHost, s->stream));
  LHIPCHK(hipStreamSynchronize(s->stream));
  *total = *s->htot;
  return PPVM_OK;
}

template <typename T>
int lscan(ppvm_lossy_sum* s, const T* in, u64 n, u64* out, u64* total) {
  if (n == 0) {
    *total = 0;
    return PPVM_OK;
  }
  const u64 chunk = scan_chunk(n);
  const u32 G = scan_blocks(n, chunk);
  hipLaunchKernelGGL(k_lz_bsum<T>, dim3(G), dim3(kLB), 0, s->stream, in, n, chunk, s->bsum);
  int rc;
  if ((rc = lscan_offsets(s, G, total))) return rc;
  hipLaunchKernelGGL(k_lz_scan<T>, dim3(G), dim3(kLB), 0, s->stream, in, n, chunk, s->bsum, out);
  LHIPCHK(hipGetLastError());
  return PPVM_OK;
}

// Index of the first n words (or, with `only`, of those with only[i] != 0, at
// most m of them) in the first pow2(2m) slots of table (tcap slots, m = n
// without `only`); returns the slot mask through *tmask.
template <int W>
int lindex_w(ppvm_lossy_sum* s, const u64* planes, u64 cap, u64 n, u32* table, u64 tcap, u64* tmask,
             const u32* only = nullptr, u64 m = 0) {
  const u64 slots = std::min(tcap, lpow2(std::max<u64>(2 * (only ? m : n), 64)));
  *tmask = slots - 1;
  hipLaunchKernelGGL(k_lz_fill, dim3(lgrid(slots)), dim3(kLB), 0, s->stream, table, slots);
  if (n)
    hipLaunchKernelGGL(k_lz_index<W>, dim3(lgrid(n)), dim3(kLB), 0, s->stream, planes, cap, n, table, slots - 1,
                       only);
  LHIPCHK(hipGetLastError());
  return PPVM_OK;
}

#define LDISPATCH(W, fn, ...)                    \
  switch (W) {                                   \
    case 1: return fn<1>(__VA_ARGS__);           \
    case 2: return fn<2>(__VA_ARGS__);           \
    case 4: return fn<4>(__VA_ARGS__);           \
    case 8: return fn<8>(__VA_ARGS__);           \
    default: return lfail(PPVM_EINVAL, "unsupported word count %u", W); \
  }

// One op given its rule table.
template <int W>
int run_op(ppvm_lossy_sum* s, const LTable& tb) {
  const u64 n = s->len;
  if (n == 0) return PPVM_OK;
  u64* P = s->planes[s->cur];
  const double* cin = s->coef[s->cur];
  const u64 chunk = scan_chunk(n);
  const u32 G = scan_blocks(n, chunk);
  if (!tb.branch) {
    hipLaunchKernelGGL(k_lz_apply<W>, dim3(G), dim3(kLB), 0, s->stream, P, s->cap, n, chunk, cin,
                       s->coef[s->cur], tb, s->cnt, s->bsum);
    LHIPCHK(hipGetLastError());
    return PPVM_OK;
  }
  double* cout = s->coef[s->cur ^ 1];
  hipLaunchKernelGGL(k_lz_apply<W>, dim3(G), dim3(kLB), 0, s->stream, P, s->cap, n, chunk, cin, cout, tb, s->cnt,
                     s->bsum);
  LHIPCHK(hipGetLastError());
  u64 M = 0;
  int rc;
  if ((rc = lscan_offsets(s, G, &M))) return rc;
  if (M == 0) {  // the planes stay in buffer cur: only the coefficient buffers trade places
    std::swap(s->coef[s->cur], s->coef[s->cur ^ 1]);
    return PPVM_OK;
  }
  if (n + M > s->cap) {
    // grow (keeps planes and the old coefficients), then redo the op on the new arena
    if ((rc = lgrow(s, n + M))) return rc;
    return run_op<W>(s, tb);
  }
  hipLaunchKernelGGL(k_lz_scan<u32>, dim3(G), dim3(kLB), 0, s->stream, s->cnt, n, chunk, s->bsum, s->off);
  u64 tmask = 0;
  u64 NEW = 0;
  if (!tb.multi) {
    // index of the parents only (at most M words): every lookup below is of a
    // parent; whether a child's word is in the map is marked by the merge,
    // which finds each existing word's contributions
    u32* present = s->flag2;
    hipLaunchKernelGGL(k_lz_fill, dim3(lgrid(M)), dim3(kLB), 0, s->stream, present, M);
    if ((rc = lindex_w<W>(s, P, s->cap, n, s->table, s->tcap, &tmask, s->cnt, M))) return rc;
    hipLaunchKernelGGL(k_lz_merge<W>, dim3(lgrid(n)), dim3(kLB), 0, s->stream, P, s->cap, n, tb, s->table, tmask,
                       cin, cout, s->off, present);
    hipLaunchKernelGGL(k_lz_child_flags<W>, dim3(lgrid(n)), dim3(kLB), 0, s->stream, P, s->cap, n, tb, s->cnt,
                       s->off, s->table, tmask, present, s->flag);
    LHIPCHK(hipGetLastError());
    if ((rc = lscan<u32>(s, s->flag, M, s->pos, &NEW))) return rc;
  } else {
    if ((rc = lindex_w<W>(s, P, s->cap, n, s->table, s->tcap, &tmask))) return rc;
    hipLaunchKernelGGL(k_lz_child_flags<W>, dim3(lgrid(n)), dim3(kLB), 0, s->stream, P, s->cap, n, tb, s->cnt,
                       s->off, s->table, tmask, nullptr, s->flag);
    LHIPCHK(hipGetLastError());
    if ((rc = lscan<u32>(s, s->flag, M, s->pos, &NEW))) return rc;
    hipLaunchKernelGGL(k_lz_merge<W>, dim3(lgrid(n)), dim3(kLB), 0, s->stream, P, s->cap, n, tb, s->table, tmask,
                       cin, cout, nullptr, nullptr);
  }
  if (NEW)
    hipLaunchKernelGGL(k_lz_child_write<W>, dim3(lgrid(n)), dim3(kLB), 0, s->stream, P, s->cap, n, tb, s->cnt,
                       s->off, s->table, tmask, s->flag, s->pos, cin, cout);
  LHIPCHK(hipGetLastError());
  if (tb.multi && NEW) {
    // map_insert_multiple: aux holds NEW + (entries that received a child)
    // keys; if that exceeds n, aux's order comes first (sum/data.rs:181-189).
    // Buffers: flag2 = entries without a child, pos2 = their ranks; then
    // flag2 = first contributors, off2 = their ranks (aux order).
    hipLaunchKernelGGL(k_lz_unhit<W>, dim3(lgrid(n)), dim3(kLB), 0, s->stream, P, s->cap, n, tb, s->table, tmask,
                       s->flag2);
    LHIPCHK(hipGetLastError());
    u64 R = 0;
    if ((rc = lscan<u32>(s, s->flag2, n, s->pos2, &R))) return rc;
    const u64 A = NEW + (n - R);
    if (A > n) {
      hipLaunchKernelGGL(k_lz_first<W>, dim3(lgrid(n)), dim3(kLB), 0, s->stream, P, s->cap, n, tb, s->cnt, s->off,
                         s->table, tmask, s->flag2);
      LHIPCHK(hipGetLastError());
      u64 A2 = 0;
      if ((rc = lscan<u32>(s, s->flag2, M, s->off2, &A2))) return rc;
      if (A2 != A)
        return lfail(PPVM_EHIP, "map_insert_multiple order: %llu first children, %llu expected",
                     (unsigned long long)A2, (unsigned long long)A);
      // cout holds every merged value; the old coefficients (cin's buffer) are
      // no longer read, so they take the reordered values
      u64* qp = s->planes[s->cur ^ 1];
      double* qc = s->coef[s->cur];
      hipLaunchKernelGGL(k_lz_reorder<W>, dim3(lgrid(n)), dim3(kLB), 0, s->stream, P, s->cap, n, tb, s->cnt, s->off,
                         s->table, tmask, s->flag2, s->off2, s->flag, s->pos, s->pos2, R, A, cout, qp, qc);
      LHIPCHK(hipGetLastError());
      s->cur ^= 1;                           // planes[cur] = qp
      std::swap(s->coef[0], s->coef[1]);     // coef[cur] = qc
      s->len = n + NEW;
      return PPVM_OK;
    }
  }
  std::swap(s->coef[s->cur], s->coef[s->cur ^ 1]);
  s->len = n + NEW;
  return PPVM_OK;
}

int run_op_w(ppvm_lossy_sum* s, const LTable& tb) { LDISPATCH(s->W, run_op, s, tb); }

template <int W>
int run_truncate(ppvm_lossy_sum* s) {
  const u64 n = s->len;
  if (n == 0) return PPVM_OK;
  const bool coef_off = !(s->tau > 0.0);
  if (coef_off && s->maxw == ~0ull && s->maxlw == ~0ull) return PPVM_OK;
  const u64 chunk = scan_chunk(n);
  const u32 G = scan_blocks(n, chunk);
  LHIPCHK(hipMemsetAsync(s->dtot, 0, 8, s->stream));
  hipLaunchKernelGGL(k_lz_keep<W>, dim3(G), dim3(kLB), 0, s->stream, s->planes[s->cur], s->cap, n, chunk,
                     s->coef[s->cur], s->tau, s->maxw, s->maxlw, s->preserve, s->n_preserve, s->off, s->bsum,
                     s->dtot);
  LHIPCHK(hipGetLastError());
  LHIPCHK(hipMemcpyAsync(s->htot, s->dtot, 8, hipMemcpyDeviceToHost, s->stream));
  LHIPCHK(hipStreamSynchronize(s->stream));
  const u64 tot = *s->htot;
  const u64 kept = tot & 0xffffffffull, restored = tot >> 32;
  if (kept == n) return PPVM_OK;  // nothing dropped
  int rc;
  if ((rc = lscan_offsets(s, G, nullptr))) return rc;  // the kept total is already known
  hipLaunchKernelGGL(k_lz_scan<u64>, dim3(G), dim3(kLB), 0, s->stream, s->off, n, chunk, s->bsum, s->off2);
  const int o = s->cur ^ 1;
  hipLaunchKernelGGL(k_lz_compact<W>, dim3(lgrid(n)), dim3(kLB), 0, s->stream, s->planes[s->cur], s->coef[s->cur],
                     s->planes[o], s->coef[o], s->cap, n, s->off, s->off2, kept);
  LHIPCHK(hipGetLastError());
  s->cur = o;
  s->len = kept + restored;
  return PPVM_OK;
}

int run_truncate_w(ppvm_lossy_sum* s) { LDISPATCH(s->W, run_truncate, s); }

template <int W>
int run_reduce(ppvm_lossy_sum* s, u32 mode, const PatternProg& prog, double* out) {
  const u64 n = s->len;
  *out = 0.0;
  if (n == 0) return PPVM_OK;
  const u32 G = std::min<u32>(kLRedGrid, lgrid(n));
  hipLaunchKernelGGL(k_lz_reduce<W>, dim3(G), dim3(kLB), 0, s->stream, s->planes[s->cur], s->cap, n,
                     s->coef[s->cur], mode, prog, s->partials);
  LHIPCHK(hipGetLastError());
  std::vector<double> part(G);
  LHIPCHK(hipMemcpyAsync(part.data(), s->partials, G * 8, hipMemcpyDeviceToHost, s->stream));
  LHIPCHK(hipStreamSynchronize(s->stream));
  double acc = 0.0;
  for (double v : part) acc += v;
  *out = acc;
  return PPVM_OK;
}

template <int W>
int run_overlap(ppvm_lossy_sum* a, ppvm_lossy_sum* b, double* out) {
  *out = 0.0;
  if (a->len == 0 || b->len == 0) return PPVM_OK;
  int rc;
  u64 tmask = 0;
  if ((rc = lindex_w<W>(a, a->planes[a->cur], a->cap, a->len, a->table, a->tcap, &tmask))) return rc;
  const u32 G = std::min<u32>(kLRedGrid, lgrid(b->len));
  hipLaunchKernelGGL(k_lz_overlap<W>, dim3(G), dim3(kLB), 0, a->stream, a->planes[a->cur], a->coef[a->cur], a->cap,
                     a->table, tmask, b->planes[b->cur], b->coef[b->cur], b->cap, b->len, a->partials);
  LHIPCHK(hipGetLastError());
  std::vector<double> part(G);
  LHIPCHK(hipMemcpyAsync(part.data(), a->partials, G * 8, hipMemcpyDeviceToHost, a->stream));
  LHIPCHK(hipStreamSynchronize(a->stream));
  double acc = 0.0;
  for (double v : part) acc += v;
  *out = acc;
  return PPVM_OK;
}

// ---- rule tables from the reference closures ----------------------------------------------
// Local one-qubit codes: I 0, X 1, Z 2, Y 3 (x | z << 1), L 4.
inline u32 cx(u32 s) { return s & 1u; }
inline u32 cz(u32 s) { return (s >> 1) & 1u; }
inline u32 code(u32 x, u32 z) { return x | (z << 1); }

void add_child(LEntry& e, u32 st, double f) {
  e.cst[e.nch] = (u8)st;
  e.cf[e.nch] = f;
  ++e.nch;
}

// Clifford conjugation of the local Paulis (sum/clifford.rs:37-217; bit maps of
// crates/ppvm-traits/src/traits/clifford.rs:117-165,270-341, signs of
// crates/ppvm-pauli-word/src/phase/clifford.rs:16-180).  Returns 1 if the
// coefficient is negated.
u32 clifford_local(u32 op, u32& sa, u32& sb) {
  const u32 x = cx(sa), z = cz(sa);
  switch (op) {
    case PPVM_OP_X: return z;
    case PPVM_OP_Y: return x ^ z;
    case PPVM_OP_Z: return x;
    case PPVM_OP_H: sa = code(z, x); return x & z;
    case PPVM_OP_S: sa = code(x, x ^ z); return x & (z ^ 1u);
    case PPVM_OP_S_DAG: sa = code(x, x ^ z); return x & z;
    case PPVM_OP_SQRT_X: sa = code(x ^ z, z); return x & z;
    case PPVM_OP_SQRT_X_DAG: sa = code(x ^ z, z); return (x ^ 1u) & z;
    case PPVM_OP_SQRT_Y: sa = code(z, x); return (x ^ 1u) & z;
    case PPVM_OP_SQRT_Y_DAG: sa = code(z, x); return x & (z ^ 1u);
    case PPVM_OP_CNOT: {
      const u32 xt = cx(sb), zt = cz(sb);
      const u32 neg = x & zt & (xt == z ? 1u : 0u);
      sa = code(x, z ^ zt);
      sb = code(xt ^ x, zt);
      return neg;
    }
    case PPVM_OP_CZ: {
      const u32 xb = cx(sb), zb = cz(sb);
      const u32 neg = x & xb & (z ^ zb);
      sa = code(x, z ^ xb);
      sb = code(xb, zb ^ x);
      return neg;
    }
    case PPVM_OP_CY: {
      const u32 xt = cx(sb), zt = cz(sb);
      const u32 neg = x & (xt ^ zt) & ((z ^ zt) ^ 1u);
      sa = code(x, z ^ xt ^ zt);
      sb = code(xt ^ x, zt ^ x);
      return neg;
    }
  }
  return 0;
}

// One-qubit rotation about axis (X 1, Z 2, Y 3) of a non-lost local Pauli
// (sum/rot1.rs:9-128): returns eps (0 = commutes) and the child's code.
int rot1_local(u32 axis, u32 s, u32& child) {
  const u32 x = cx(s), z = cz(s);
  switch (axis) {
    case 1:  // X
      if (!z) return 0;
      child = code(x ^ 1u, z);
      return x ? -1 : 1;
    case 3:  // Y
      if (x == z) return 0;
      child = code(x ^ 1u, z ^ 1u);
      return z ? -1 : 1;
    default:  // Z
      if (!x) return 0;
      child = code(x, z ^ 1u);
      return z ? 1 : -1;
  }
}
// Two-qubit rotation XX / YY / ZZ of non-lost locals (sum/rot2.rs:63-184).
int rot2_local(u32 op, u32 sa, u32 sb, u32& ca, u32& cb) {
  const u32 xa = cx(sa), za = cz(sa), xb = cx(sb), zb = cz(sb);
  switch (op) {
    case PPVM_OP_RZZ: {
      if (xa == xb) return 0;
      const u32 z_anti = xa ? za : zb;
      ca = code(xa, za ^ 1u);
      cb = code(xb, zb ^ 1u);
      return z_anti ? 1 : -1;
    }
    case PPVM_OP_RXX: {
      if (za == zb) return 0;
      const u32 x_anti = za ? xa : xb;
      ca = code(xa ^ 1u, za);
      cb = code(xb ^ 1u, zb);
      return x_anti ? -1 : 1;
    }
    default: {  // RYY
      const u32 pa = xa ^ za, pb = xb ^ zb;
      if (pa == pb) return 0;
      const u32 x_anti = pa ? xa : xb;
      ca = code(xa ^ 1u, za ^ 1u);
      cb = code(xb ^ 1u, zb ^ 1u);
      return x_anti ? 1 : -1;
    }
  }
}

bool is_two(u32 c) {
  return c == PPVM_OP_CNOT || c == PPVM_OP_CZ || c == PPVM_OP_CY || c == PPVM_OP_RXX || c == PPVM_OP_RYY ||
         c == PPVM_OP_RZZ || c == PPVM_OP_DEPOLARIZE2 || c == PPVM_OP_PAULI_ERROR2 || c == PPVM_OP_CORRELATED_LOSS;
}

// Builds the rule table of one op; returns PPVM_EINVAL for codes that are not gates.
int build_table(const ppvm_op& op, LTable& tb) {
  std::memset(&tb, 0, sizeof(tb));
  const u32 c = op.code;
  tb.a = op.q0;
  tb.two = is_two(c) ? 1u : 0u;
  tb.b = tb.two ? op.q1 : op.q0;
  const u32 ns = tb.two ? 25u : 5u;
  for (u32 t = 0; t < ns; ++t) {
    tb.e[t].pf = 1.0;
    tb.e[t].pst = (u8)t;
  }
  double sn = 0.0, cs = 1.0;
  if (c >= PPVM_OP_RX && c <= PPVM_OP_RZZ) sincos(op.theta, &sn, &cs);  // Float::sin_cos, coefficient.rs:55-57
  for (u32 t = 0; t < ns; ++t) {
    LEntry& e = tb.e[t];
    const u32 sa = tb.two ? t / 5u : t, sb = tb.two ? t % 5u : 0u;
    const bool la = sa == kStateL, lb = tb.two && sb == kStateL;
    switch (c) {
      case PPVM_OP_X: case PPVM_OP_Y: case PPVM_OP_Z: case PPVM_OP_H: case PPVM_OP_S: case PPVM_OP_S_DAG:
      case PPVM_OP_SQRT_X: case PPVM_OP_SQRT_X_DAG: case PPVM_OP_SQRT_Y: case PPVM_OP_SQRT_Y_DAG:
      case PPVM_OP_CNOT: case PPVM_OP_CZ: case PPVM_OP_CY: {
        if (la || lb) break;  // a lost qubit leaves the word as it is (phase/clifford.rs:18-180)
        u32 na = sa, nb = sb;
        if (clifford_local(c, na, nb)) e.pf = -1.0;
        e.pst = (u8)(tb.two ? na * 5u + nb : na);
        if (e.pst != t) tb.moves = 1;
        break;
      }
      case PPVM_OP_RX: case PPVM_OP_RY: case PPVM_OP_RZ: {
        if (la) break;  // rot1.rs:20-23
        const u32 axis = c == PPVM_OP_RX ? 1u : c == PPVM_OP_RY ? 3u : 2u;
        u32 ch = 0;
        const int eps = rot1_local(axis, sa, ch);
        if (!eps) break;
        e.pf = cs;
        add_child(e, ch, (double)eps * sn);  // sin.mul_sign(eps)
        break;
      }
      case PPVM_OP_RXX: case PPVM_OP_RYY: case PPVM_OP_RZZ: {
        const u32 axis = c == PPVM_OP_RXX ? 1u : c == PPVM_OP_RYY ? 3u : 2u;
        u32 ca = sa, cb = sb;
        int eps;
        if (la || lb) {
          // a lost qubit leaves the one-qubit rotation on the other (rot2.rs:66-75);
          // with both lost nothing acts
          if (la && lb) break;
          u32 ch = 0;
          eps = rot1_local(axis, la ? sb : sa, ch);
          if (la) cb = ch; else ca = ch;
        } else {
          eps = rot2_local(c, sa, sb, ca, cb);
        }
        if (!eps) break;
        e.pf = cs;
        add_child(e, ca * 5u + cb, (double)eps * sn);
        break;
      }
      case PPVM_OP_PAULI_ERROR: {  // noise.rs:17-39
        if (la) break;
        const double xf = 1.0 - op.p[1] * 2.0 - op.p[2] * 2.0;
        const double zf = 1.0 - op.p[0] * 2.0 - op.p[1] * 2.0;
        const double yf = 1.0 - op.p[0] * 2.0 - op.p[2] * 2.0;
        if (sa == 1u) e.pf = xf;
        else if (sa == 2u) e.pf = zf;
        else if (sa == 3u) e.pf = yf;
        break;
      }
      case PPVM_OP_DEPOLARIZE1:  // noise.rs:110-118
        if (!la && sa != 0u) e.pf = 1.0 - op.p[0] * (4.0 / 3.0);
        break;
      case PPVM_OP_DEPOLARIZE2:  // noise.rs:120-131
        if (!la && !lb && (sa != 0u || sb != 0u)) e.pf = 1.0 - op.p[0] * (16.0 / 15.0);
        break;
      case PPVM_OP_PAULI_ERROR2: {  // noise.rs:52-107; either qubit lost: no noise
        if (la || lb) break;
        const int row = 4 * kCodeToLetter[sa] + kCodeToLetter[sb];
        if (row == 0) break;
        double acc = 1.0;
        for (int j = 0; j < 8; ++j) acc = acc - op.p[kPe2Rows[row][j]] * 2.0;
        e.pf = acc;
        break;
      }
      case PPVM_OP_AMPLITUDE_DAMPING: {  // noise.rs:132-157
        const double g = op.p[0];
        if (sa == 1u || sa == 3u) {
          e.pf = std::sqrt(1.0 - g);
        } else if (sa == 2u) {
          add_child(e, 0u, g);
          e.pf = 1.0 - g;
        }
        break;
      }
      case PPVM_OP_LOSS: {  // noise.rs:165-180
        const double p = op.p[0];
        if (la) add_child(e, 0u, p);
        else e.pf = 1.0 - p;
        break;
      }
      case PPVM_OP_RESET_LOSS:  // noise.rs:254-274
        if (la) e.pf = 0.0;
        else if (sa == 0u || sa == 2u) add_child(e, kStateL, 1.0);
        break;
      case PPVM_OP_CORRELATED_LOSS: {  // noise.rs:182-248
        const double* p = op.p;
        if (la && lb) {
          add_child(e, 0u * 5u + kStateL, p[2]);
          add_child(e, kStateL * 5u + 0u, p[2]);
          add_child(e, 0u, p[0]);
        } else if (lb) {
          add_child(e, sa * 5u + 0u, p[1]);
          e.pf = 1.0 - p[2];
        } else if (la) {
          add_child(e, 0u * 5u + sb, p[1]);
          e.pf = 1.0 - p[2];
        } else {
          e.pf = 1.0 - p[1] * 2.0 - p[0];
        }
        break;
      }
      default:
        return lfail(PPVM_EINVAL, "unknown op code %u", c);
    }
    if (e.nch) tb.branch = 1;
  }
  // the correlated loss channel goes through map_insert_multiple (noise.rs:193),
  // every other branching op through map_insert
  tb.multi = c == PPVM_OP_CORRELATED_LOSS ? 1u : 0u;
  if (tb.branch) {
    for (u32 t = 0; t < ns; ++t) {
      const LEntry& e = tb.e[t];
      for (u32 j = 0; j < e.nch; ++j) {
        LPre& pr = tb.pre[e.cst[j]];
        if (pr.n >= 3) return lfail(PPVM_EINVAL, "op %u: a word receives from more than three parents", c);
        pr.s[pr.n] = (u8)t;
        pr.j[pr.n] = (u8)j;
        ++pr.n;
      }
    }
  }
  return PPVM_OK;
}

int lvalidate(const ppvm_lossy_sum* s, const ppvm_op& op) {
  const u32 c = op.code;
  if (c == PPVM_OP_NOP || c == PPVM_OP_TRUNCATE) return PPVM_OK;
  const bool known = (c >= PPVM_OP_X && c <= PPVM_OP_CY) || (c >= PPVM_OP_RX && c <= PPVM_OP_RZZ) ||
                     (c >= PPVM_OP_PAULI_ERROR && c <= PPVM_OP_RESET_LOSS);
  if (!known) return lfail(PPVM_EINVAL, "unknown op code %u", c);
  if (op.q0 >= s->n) return lfail(PPVM_EINVAL, "qubit index %u out of range for %u qubits", op.q0, s->n);
  if (is_two(c)) {
    if (op.q1 >= s->n) return lfail(PPVM_EINVAL, "qubit index %u out of range for %u qubits", op.q1, s->n);
    if (op.q1 == op.q0) return lfail(PPVM_EINVAL, "two-qubit op on identical qubits (%u, %u)", op.q0, op.q1);
  }
  if (c >= PPVM_OP_RX && c <= PPVM_OP_RZZ && !std::isfinite(op.theta))
    return lfail(PPVM_EINVAL, "rotation angle is not finite");
  return PPVM_OK;
}

int lready(ppvm_lossy_sum* s) {
  if (!s) return lfail(PPVM_EINVAL, "null handle");
  if (s->broken) return lfail(PPVM_ESTATE, "handle is unusable after an earlier device error");
  return hipSetDevice(s->device) == hipSuccess ? PPVM_OK : lfail(PPVM_EHIP, "hipSetDevice failed");
}

int broken(ppvm_lossy_sum* s, int rc) {
  if (rc == PPVM_EHIP || rc == PPVM_ENOMEM) s->broken = true;
  return rc;
}

// Host copy of the terms (term-major: 3W words, then the coefficient).
int lexport(ppvm_lossy_sum* s, std::vector<u64>& keys, std::vector<double>& c) {
  const u64 n = s->len, np = 3ull * s->W;
  std::vector<u64> pl(np * n);
  c.resize(n);
  if (n) {
    for (u64 p = 0; p < np; ++p)
      LHIPCHK(hipMemcpyAsync(pl.data() + p * n, s->planes[s->cur] + p * s->cap, n * 8, hipMemcpyDeviceToHost,
                             s->stream));
    LHIPCHK(hipMemcpyAsync(c.data(), s->coef[s->cur], n * 8, hipMemcpyDeviceToHost, s->stream));
  }
  LHIPCHK(hipStreamSynchronize(s->stream));
  keys.resize(np * n);
  for (u64 i = 0; i < n; ++i)
    for (u64 p = 0; p < np; ++p) keys[i * np + p] = pl[p * n + i];
  return PPVM_OK;
}

int lupload(ppvm_lossy_sum* s, const std::vector<u64>& keys, const std::vector<double>& c) {
  const u64 n = c.size(), np = 3ull * s->W;
  int rc;
  s->len = 0;
  if ((rc = lgrow(s, n))) return rc;
  std::vector<u64> pl(np * n);
  for (u64 i = 0; i < n; ++i)
    for (u64 p = 0; p < np; ++p) pl[p * n + i] = keys[i * np + p];
  if (n) {
    for (u64 p = 0; p < np; ++p)
      LHIPCHK(hipMemcpyAsync(s->planes[s->cur] + p * s->cap, pl.data() + p * n, n * 8, hipMemcpyHostToDevice,
                             s->stream));
    LHIPCHK(hipMemcpyAsync(s->coef[s->cur], c.data(), n * 8, hipMemcpyHostToDevice, s->stream));
  }
  LHIPCHK(hipStreamSynchronize(s->stream));
  s->len = n;
  return PPVM_OK;
}

struct LKeyHash {
  size_t operator()(const std::vector<u64>& k) const {
    u64 h = 0x9e3779b97f4a7c15ull;
    for (u64 w : k) h = (h ^ w) * 0xbf58476d1ce4e5b9ull + (h >> 31);
    return (size_t)h;
  }
};

// Words (Wa words per plane) must satisfy x = z = 0 wherever l = 1 and have no
// bits at or above n.
int check_words(const ppvm_lossy_sum* s, const uint64_t* x, const uint64_t* z, const uint64_t* l, size_t count) {
  const u32 Wa = s->Wa;
  for (size_t i = 0; i < count; ++i)
    for (u32 p = 0; p < Wa; ++p) {
      const u64 xv = x[i * Wa + p], zv = z[i * Wa + p], lv = l[i * Wa + p];
      u64 valid = ~0ull;
      if (p == Wa - 1 && (s->n & 63)) valid = (1ull << (s->n & 63)) - 1;
      if (((xv | zv | lv) & ~valid) != 0)
        return lfail(PPVM_EINVAL, "term %zu has bits beyond qubit %u", i, s->n - 1);
      if (lv & (xv | zv)) return lfail(PPVM_EINVAL, "term %zu: a lost qubit cannot also carry X, Y or Z", i);
    }
  return PPVM_OK;
}

}  // namespace

extern "C" {

namespace {

// Local-state masks of a list of words (np = 3W words each: x, z, l planes).
void lstates_of(ppvm_lossy_sum* s, const std::vector<u64>& keys, size_t count) {
  const u32 W = s->W, Wa = s->Wa;
  for (auto& m : s->smask) m.assign(Wa, 0ull);
  for (size_t i = 0; i < count; ++i) {
    const u64* k = keys.data() + i * 3 * W;
    for (u32 p = 0; p < Wa; ++p) {
      const u64 x = k[p], z = k[W + p], l = k[2 * W + p];
      const u64 valid = p + 1 < Wa || s->n % 64 == 0 ? ~0ull : (1ull << (s->n % 64)) - 1ull;
      s->smask[0][p] |= ~(x | z | l) & valid;
      s->smask[1][p] |= x & ~z;
      s->smask[2][p] |= z & ~x;
      s->smask[3][p] |= x & z;
      s->smask[4][p] |= l;
    }
  }
}

u32 lstates_at(const ppvm_lossy_sum* s, u32 q) {
  u32 m = 0;
  for (u32 t = 0; t < 5; ++t) m |= (u32)((s->smask[t][q / 64] >> (q % 64)) & 1ull) << t;
  return m;
}

// Whether the op's rule table leaves every local state the sum may hold on
// its qubits unchanged (factor 1, same state, no child): then it changes nothing.
bool lop_is_identity(const ppvm_lossy_sum* s, const LTable& tb) {
  const u32 ma = lstates_at(s, tb.a), mb = tb.two ? lstates_at(s, tb.b) : 1u;
  for (u32 sa = 0; sa < 5; ++sa) {
    if (!((ma >> sa) & 1u)) continue;
    for (u32 sb = 0; sb < (tb.two ? 5u : 1u); ++sb) {
      if (!((mb >> sb) & 1u)) continue;
      const u32 t = tb.two ? sa * 5u + sb : sa;
      const LEntry& e = tb.e[t];
      if (e.pf != 1.0 || e.pst != t || e.nch != 0) return false;
    }
  }
  return true;
}

// Widens the masks by the states an op that ran can leave on its qubits.
void lstates_after(ppvm_lossy_sum* s, const LTable& tb) {
  const u32 ma = lstates_at(s, tb.a), mb = tb.two ? lstates_at(s, tb.b) : 1u;
  u32 na = ma, nb = mb;
  for (u32 sa = 0; sa < 5; ++sa) {
    if (!((ma >> sa) & 1u)) continue;
    for (u32 sb = 0; sb < (tb.two ? 5u : 1u); ++sb) {
      if (!((mb >> sb) & 1u)) continue;
      const LEntry& e = tb.e[tb.two ? sa * 5u + sb : sa];
      u32 outs[4] = {e.pst, 0, 0, 0};
      for (u32 j = 0; j < e.nch; ++j) outs[1 + j] = e.cst[j];
      for (u32 j = 0; j <= e.nch; ++j) {
        na |= 1u << (tb.two ? outs[j] / 5u : outs[j]);
        if (tb.two) nb |= 1u << (outs[j] % 5u);
      }
    }
  }
  for (u32 t = 0; t < 5; ++t) {
    if ((na >> t) & 1u) s->smask[t][tb.a / 64] |= 1ull << (tb.a % 64);
    if (tb.two && ((nb >> t) & 1u)) s->smask[t][tb.b / 64] |= 1ull << (tb.b % 64);
  }
}

}  // namespace

const char* ppvm_lossy_last_error(void) { return l_err.c_str(); }

int ppvm_lossy_create(uint32_t n_qubits, double min_abs_coeff, uint64_t max_pauli_weight, uint64_t max_loss_weight,
                      uint64_t capacity_hint, ppvm_lossy_handle* out) {
  if (!out) return lfail(PPVM_EINVAL, "null output pointer");
  *out = nullptr;
  if (n_qubits == 0) return lfail(PPVM_EINVAL, "n_qubits must be positive");
  const u32 Wa = (n_qubits + 63) / 64;
  const u32 W = Wa <= 1 ? 1 : Wa <= 2 ? 2 : Wa <= 4 ? 4 : Wa <= 8 ? 8 : 0;
  if (!W) return lfail(PPVM_EINVAL, "n_qubits=%u exceeds the supported maximum of 512", n_qubits);
  if (!(min_abs_coeff >= 0.0)) return lfail(PPVM_EINVAL, "min_abs_coeff must be a non-negative number");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return lfail(PPVM_ENODEV, "no HIP device available");
  }
  ppvm_lossy_sum* s = new ppvm_lossy_sum();
  s->n = n_qubits;
  s->W = W;
  s->Wa = Wa;
  s->tau = min_abs_coeff;
  s->maxw = max_pauli_weight;
  s->maxlw = max_loss_weight;
  int rc = PPVM_OK;
  if (hipGetDevice(&s->device) != hipSuccess || hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&s->bsum, (size_t)kLScanMaxBlocks * 8) != hipSuccess || hipMalloc(&s->dtot, 8) != hipSuccess ||
      hipHostMalloc(&s->htot, 8, hipHostMallocDefault) != hipSuccess ||
      hipMalloc(&s->partials, (size_t)kLRedGrid * 8) != hipSuccess) {
    (void)hipGetLastError();
    rc = lfail(PPVM_ENOMEM, "device allocation failed");
  }
  if (!rc) rc = lgrow(s, std::max<u64>(capacity_hint, kLMinCap));
  if (rc) {
    ppvm_lossy_destroy(s);
    return rc;
  }
  lstates_of(s, {}, 0);
  *out = s;
  return PPVM_OK;
}

int ppvm_lossy_destroy(ppvm_lossy_handle s) {
  if (!s) return PPVM_OK;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  lfree(s);
  if (s->bsum) (void)hipFree(s->bsum);
  if (s->dtot) (void)hipFree(s->dtot);
  if (s->htot) (void)hipHostFree(s->htot);
  if (s->partials) (void)hipFree(s->partials);
  if (s->preserve) (void)hipFree(s->preserve);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return PPVM_OK;
}

int ppvm_lossy_words(ppvm_lossy_handle s, uint32_t* w) {
  if (!s || !w) return lfail(PPVM_EINVAL, "null argument");
  *w = s->Wa;
  return PPVM_OK;
}

int ppvm_lossy_clone(ppvm_lossy_handle s, ppvm_lossy_handle* out) {
  if (!out) return lfail(PPVM_EINVAL, "null output pointer");
  *out = nullptr;
  int rc;
  if ((rc = lready(s))) return rc;
  ppvm_lossy_handle c = nullptr;
  if ((rc = ppvm_lossy_create(s->n, s->tau, s->maxw, s->maxlw, s->len, &c))) return rc;
  const u64 np = 3ull * s->W;
  LHIPCHK(hipStreamSynchronize(s->stream));
  for (u64 p = 0; p < np && s->len; ++p)
    LHIPCHK(hipMemcpyAsync(c->planes[c->cur] + p * c->cap, s->planes[s->cur] + p * s->cap, s->len * 8,
                           hipMemcpyDeviceToDevice, c->stream));
  if (s->len)
    LHIPCHK(hipMemcpyAsync(c->coef[c->cur], s->coef[s->cur], s->len * 8, hipMemcpyDeviceToDevice, c->stream));
  if (s->n_preserve) {
    if (hipMalloc(&c->preserve, (size_t)s->n_preserve * np * 8) != hipSuccess) {
      (void)hipGetLastError();
      ppvm_lossy_destroy(c);
      return lfail(PPVM_ENOMEM, "device allocation failed");
    }
    LHIPCHK(hipMemcpyAsync(c->preserve, s->preserve, (size_t)s->n_preserve * np * 8, hipMemcpyDeviceToDevice,
                           c->stream));
    c->n_preserve = s->n_preserve;
  }
  LHIPCHK(hipStreamSynchronize(c->stream));
  c->len = s->len;
  for (int t = 0; t < 5; ++t) c->smask[t] = s->smask[t];
  c->clean = s->clean;
  *out = c;
  return PPVM_OK;
}

int ppvm_lossy_set_preserve(ppvm_lossy_handle s, const uint64_t* x, const uint64_t* z, const uint64_t* l,
                            size_t count) {
  int rc;
  if ((rc = lready(s))) return rc;
  if (count && (!x || !z || !l)) return lfail(PPVM_EINVAL, "null argument");
  if (count > (1u << 20)) return lfail(PPVM_EINVAL, "preserve set of %zu words is too large", count);
  if ((rc = check_words(s, x, z, l, count))) return rc;
  const u32 W = s->W, Wa = s->Wa;
  std::vector<u64> keys((size_t)count * 3 * W, 0ull);
  for (size_t i = 0; i < count; ++i)
    for (u32 p = 0; p < Wa; ++p) {
      keys[i * 3 * W + p] = x[i * Wa + p];
      keys[i * 3 * W + W + p] = z[i * Wa + p];
      keys[i * 3 * W + 2 * W + p] = l[i * Wa + p];
    }
  LHIPCHK(hipStreamSynchronize(s->stream));
  if (s->preserve) (void)hipFree(s->preserve);
  s->preserve = nullptr;
  s->n_preserve = 0;
  s->clean = false;
  if (!count) return PPVM_OK;
  if (hipMalloc(&s->preserve, keys.size() * 8) != hipSuccess) {
    (void)hipGetLastError();
    return lfail(PPVM_ENOMEM, "device allocation failed");
  }
  LHIPCHK(hipMemcpy(s->preserve, keys.data(), keys.size() * 8, hipMemcpyHostToDevice));
  s->n_preserve = (u32)count;
  return PPVM_OK;
}

int ppvm_lossy_add_terms(ppvm_lossy_handle s, const uint64_t* x, const uint64_t* z, const uint64_t* l,
                         const double* c, size_t count) {
  int rc;
  if ((rc = lready(s))) return rc;
  if (count == 0) return PPVM_OK;
  if (!x || !z || !l || !c) return lfail(PPVM_EINVAL, "null argument");
  if ((rc = check_words(s, x, z, l, count))) return rc;
  std::vector<u64> keys;
  std::vector<double> cs;
  if ((rc = lexport(s, keys, cs))) return broken(s, rc);
  const u32 W = s->W, Wa = s->Wa, np = 3 * W;
  std::unordered_map<std::vector<u64>, size_t, LKeyHash> idx;
  for (size_t i = 0; i < cs.size(); ++i)
    idx.emplace(std::vector<u64>(keys.begin() + i * np, keys.begin() + (i + 1) * np), i);
  for (size_t i = 0; i < count; ++i) {
    std::vector<u64> k(np, 0ull);
    for (u32 p = 0; p < Wa; ++p) {
      k[p] = x[i * Wa + p];
      k[W + p] = z[i * Wa + p];
      k[2 * W + p] = l[i * Wa + p];
    }
    auto it = idx.find(k);
    if (it != idx.end()) {
      cs[it->second] += c[i];  // entry(k).and_modify(+=)   (map/hashmap.rs:40-44)
    } else {
      idx.emplace(k, cs.size());
      keys.insert(keys.end(), k.begin(), k.end());
      cs.push_back(c[i]);
    }
  }
  lstates_of(s, keys, cs.size());
  s->clean = false;
  return broken(s, lupload(s, keys, cs));
}

int ppvm_lossy_apply(ppvm_lossy_handle s, const ppvm_op* ops, size_t n) {
  int rc;
  if ((rc = lready(s))) return rc;
  if (n && !ops) return lfail(PPVM_EINVAL, "null program");
  for (size_t i = 0; i < n; ++i)
    if ((rc = lvalidate(s, ops[i]))) return rc;
  for (size_t i = 0; i < n; ++i) {
    const ppvm_op& op = ops[i];
    if (op.code == PPVM_OP_NOP) continue;
    if (op.code == PPVM_OP_TRUNCATE) {
      // on a state a truncation already left (nothing changed since) it drops
      // nothing and moves nothing
      if (s->clean) continue;
      if ((rc = run_truncate_w(s))) return broken(s, rc);
      s->clean = true;
      continue;
    }
    LTable tb;
    if ((rc = build_table(op, tb))) return rc;
    if (lop_is_identity(s, tb)) continue;  // acts on no local state a term holds
    if ((rc = run_op_w(s, tb))) return broken(s, rc);
    lstates_after(s, tb);
    s->clean = false;
  }
  LHIPCHK(hipStreamSynchronize(s->stream));
  return PPVM_OK;
}

int ppvm_lossy_len(ppvm_lossy_handle s, uint64_t* out) {
  if (!out) return lfail(PPVM_EINVAL, "null argument");
  int rc;
  if ((rc = lready(s))) return rc;
  *out = s->len;
  return PPVM_OK;
}

int ppvm_lossy_export_terms(ppvm_lossy_handle s, uint64_t* x, uint64_t* z, uint64_t* l, double* c, size_t cap,
                            size_t* n) {
  if (!n) return lfail(PPVM_EINVAL, "null argument");
  int rc;
  if ((rc = lready(s))) return rc;
  *n = s->len;
  if (cap < s->len) return lfail(PPVM_EINVAL, "export buffer holds %zu terms, the sum has %llu", cap,
                                 (unsigned long long)s->len);
  if (s->len && (!x || !z || !l || !c)) return lfail(PPVM_EINVAL, "null argument");
  std::vector<u64> keys;
  std::vector<double> cs;
  if ((rc = lexport(s, keys, cs))) return broken(s, rc);
  const u32 W = s->W, Wa = s->Wa, np = 3 * W;
  for (size_t i = 0; i < cs.size(); ++i) {
    for (u32 p = 0; p < Wa; ++p) {
      x[i * Wa + p] = keys[i * np + p];
      z[i * Wa + p] = keys[i * np + W + p];
      l[i * Wa + p] = keys[i * np + 2 * W + p];
    }
    c[i] = cs[i];
  }
  return PPVM_OK;
}

int ppvm_lossy_overlap_with_zero(ppvm_lossy_handle s, double* out) {
  if (!out) return lfail(PPVM_EINVAL, "null argument");
  int rc;
  if ((rc = lready(s))) return rc;
  PatternProg prog{};
  switch (s->W) {
    case 1: rc = run_reduce<1>(s, 0, prog, out); break;
    case 2: rc = run_reduce<2>(s, 0, prog, out); break;
    case 4: rc = run_reduce<4>(s, 0, prog, out); break;
    default: rc = run_reduce<8>(s, 0, prog, out); break;
  }
  return broken(s, rc);
}

int ppvm_lossy_trace(ppvm_lossy_handle s, const char* pattern, double* out) {
  if (!out || !pattern) return lfail(PPVM_EINVAL, "null argument");
  int rc;
  if ((rc = lready(s))) return rc;
  PatternProg prog;
  if (compile_pattern(pattern, s->n, prog) != PPVM_OK) return lfail(PPVM_EINVAL, "%s", ppvm_last_error());
  switch (s->W) {
    case 1: rc = run_reduce<1>(s, 1, prog, out); break;
    case 2: rc = run_reduce<2>(s, 1, prog, out); break;
    case 4: rc = run_reduce<4>(s, 1, prog, out); break;
    default: rc = run_reduce<8>(s, 1, prog, out); break;
  }
  return broken(s, rc);
}

int ppvm_lossy_overlap(ppvm_lossy_handle a, ppvm_lossy_handle b, double* out) {
  if (!out) return lfail(PPVM_EINVAL, "null argument");
  int rc;
  if ((rc = lready(a)) || (rc = lready(b))) return rc;
  if (a->n != b->n) return lfail(PPVM_EINVAL, "overlap of sums over %u and %u qubits", a->n, b->n);
  LHIPCHK(hipStreamSynchronize(b->stream));
  switch (a->W) {
    case 1: rc = run_overlap<1>(a, b, out); break;
    case 2: rc = run_overlap<2>(a, b, out); break;
    case 4: rc = run_overlap<4>(a, b, out); break;
    default: rc = run_overlap<8>(a, b, out); break;
  }
  return broken(a, rc);
}

}  // extern "C"
