// device.hpp — device-side building blocks of the gfx950 Pauli-propagation engine.
//
// Data layout in HBM (per PauliSum handle):
//   keys   : u64[cap][2W]   array of structures: the 2W words of an entry are
//                           contiguous (x words 0..W-1, then z words), so a
//                           sparse gather reads one aligned 16W-byte block.
//   coef   : f64[cap]       coefficient; a truncated ("dead") entry holds the
//                           NaN pattern DEAD_BITS and is skipped by every kernel.
//   wl     : WMRec[cap]     per-gate worklist of anticommuting terms, one
//                           16-byte record each (coefficient, arena index,
//                           facts about the key that later steps need).
//   table  : u64[tcap]      open-addressed pair table (generation-tagged slots).
//   masks  : u64[cap/64]    per-pass ballot bitmasks (bit j of word w = item 64w+j)
//   parts  : u64[6][32768]  per-chunk counts and per-block partial sums.
//
// Stream compaction (worklist, children, dead-entry removal) is done without
// atomics or inter-block waiting: a pass over n items is split into G
// contiguous chunks; a count kernel writes each chunk's ballot masks and
// count, and a write kernel re-reads only the masks, takes its chunk offset
// from the counts of the earlier chunks and ranks items inside the chunk with
// an LDS scan.  The output order is the input order (deterministic).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef uint64_t u64;
typedef uint32_t u32;

namespace ppvm {

constexpr int kBlock = 256;        // threads per block (4 wave64) for grid-stride kernels
constexpr int kWaves = kBlock / 64;
constexpr int kMaxParts = 32768;   // max chunks / blocks whose partials are reduced
constexpr int kScanThreads = 1024; // single-block scan / reduction kernels

// Dead-entry marker: a quiet NaN with a fixed payload (sign ignored).
constexpr u64 DEAD_BITS = 0x7ff8dead0000dea1ull;
constexpr u64 ABS_MASK = 0x7fffffffffffffffull;

__device__ __forceinline__ bool is_dead(double c) {
  return (((u64)__double_as_longlong(c)) & ABS_MASK) == DEAD_BITS;
}
__device__ __forceinline__ double dead_value() { return __longlong_as_double((long long)DEAD_BITS); }

// Per-op counters, double-buffered by op parity: op k reads slot k&1 and
// writes slot (k+1)&1, so no kernel ever rewrites a count it iterates over.
struct alignas(64) CountSlot {
  u64 n_entries;  // arena entries (live + dead)
  u64 n_dead;
  u64 pad[6];
};

struct alignas(64) DevState {
  CountSlot slot[2];
  u64 n_wl;           // worklist length of the current rotation
  u64 overflow;       // 0, or (op sequence number + 1) of the first overflow
  u64 term_updates;   // sum over ops of live input terms
  u64 gates;          // ops applied (rotations + map ops)
  u64 rotations;
  u64 scanned;        // entries scanned by rotation scans
  u64 wl_total;       // anticommuting terms over all rotations
  u64 pairs_total;    // pairs merged (two terms each)
  u64 children_total; // new terms appended (after truncation)
  u64 kills_total;    // terms truncated
  u64 map_entries;    // entries visited by map programs
  u64 map_launches;
  u64 scan_planes;    // bit-sliced planes read by rotation scans (entries x planes)
  double overlap;     // last overlap_with_zero result
  u64 tvw;            // words of the bit-sliced planes that are current (entries < 64 tvw)
  u64 transposed;     // entries written into the bit-sliced planes
  u64 pair_fallback;  // records the partitioned pairing sent to the global table
  u64 small_done;     // steps completed by the last LDS-resident run (k_small)
  u64 small_peak;     // largest live count inside that run
  u64 pad[2];
  u64 sup[16];        // OR of the live entries' key words (x words, then z words), written at a sync
};

// Branching op: a rotation exp(-i theta/2 G) with G acting on qubits a (and
// b), or the Z -> (1 - gamma) Z + gamma I branch of amplitude damping on qubit a
// (then G = Z_a, cs = 1 - gamma, sn = gamma, and the worklist holds the terms
// with I or Z on a, paired as {I, Z} versions of one word).
constexpr u32 kBranchRotation = 0, kBranchDamping = 1;
struct RotSpec {
  u32 a, b;
  u32 two;                 // 1 for a two-qubit generator
  u32 gax, gaz, gbx, gbz;  // generator Pauli bits on a and b
  u32 trunc;               // fused truncation after the gate
  u32 test_planes;         // bit-sliced planes the scan reads (x/z of a, b)
  u32 mode;                // kBranchRotation, or kBranchDamping (amplitude damping on a)
  double sn, cs;           // sin(theta), cos(theta) from the host libm
};

struct TruncSpec {
  double tau;              // drop |c| < tau
  u64 maxw;                // drop weight > maxw (UINT64_MAX: off)
  const u64* preserve;     // n_preserve keys of 2W words (never dropped)
  u32 n_preserve;
  u32 W;
};

// ---- map programs (Clifford / noise / truncation), passed by value ---------
constexpr int kMaxMapOps = 48;
constexpr int kMaxFactorTables = 8;
enum MapCode : u32 {
  M_X = 1, M_Y, M_Z, M_H, M_S, M_S_DAG, M_SQRT_X, M_SQRT_X_DAG, M_SQRT_Y, M_SQRT_Y_DAG,
  M_CNOT = 11, M_CZ, M_CY,
  M_SCALE = 30,            // coefficient *= table[4*code_a + code_b] where mask bit set
};
struct MapOp {
  u32 code_flags;          // code | trunc << 8 | table << 16
  u32 a, b;
};
struct MapProgram {
  u32 n_ops;
  u32 trunc_before;        // full truncation check before the first op
  MapOp ops[kMaxMapOps];
  u32 apply_mask[kMaxFactorTables];   // 16-bit mask of table entries that apply
  // dropped ops (identity on every term) that the program books as term
  // updates: idle[j] of them ran before op j (idle[n_ops]: after the last),
  // n_idle in all
  u32 n_idle;
  unsigned short idle[kMaxMapOps + 2];
  double factors[kMaxFactorTables][16];
};

// ---- small helpers --------------------------------------------------------
__device__ __forceinline__ u64 mix64(u64 h) {
  h ^= h >> 31;
  h *= 0x9e3779b97f4a7c15ull;
  h ^= h >> 29;
  h *= 0xbf58476d1ce4e5b9ull;
  h ^= h >> 32;
  return h;
}

template <int W>
__device__ __forceinline__ u64 hash_key(const u64 (&k)[2 * W]) {
  u64 h = 0x6a09e667f3bcc909ull;
#pragma unroll
  for (int i = 0; i < 2 * W; ++i) h = mix64(h ^ k[i]) + (u64)i;
  return mix64(h);
}

template <int W>
__device__ __forceinline__ u32 weight_of(const u64 (&k)[2 * W]) {
  u32 w = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) w += __popcll(k[i] | k[W + i]);
  return w;
}

template <int W>
__device__ __forceinline__ bool is_preserved(const u64 (&k)[2 * W], const TruncSpec& t) {
  for (u32 p = 0; p < t.n_preserve; ++p) {
    bool eq = true;
#pragma unroll
    for (int i = 0; i < 2 * W; ++i) eq &= (t.preserve[p * 2 * W + i] == k[i]);
    if (eq) return true;
  }
  return false;
}

// Truncation predicate of CombinedStrategy(CoefficientThreshold, MaxPauliWeight)
// with the preserve set applied on top.
template <int W>
__device__ __forceinline__ bool drop_term(const u64 (&k)[2 * W], double c, const TruncSpec& t) {
  bool drop = fabs(c) < t.tau;
  if (!drop && t.maxw != ~0ull) drop = weight_of<W>(k) > t.maxw;
  if (drop && t.n_preserve) drop = !is_preserved<W>(k, t);
  return drop;
}

__device__ __forceinline__ u32 bit_of(u64 w, u32 q) { return (u32)(w >> (q & 63)) & 1u; }

// ---- scoped atomics (agent scope: visible across all XCDs) -----------------
__device__ __forceinline__ u64 ld_acquire(const u64* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release(u64* p, u64 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 ld_relaxed(const u64* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 cas_u64(u64* p, u64 expected, u64 desired) {
  __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expected;  // value observed in memory
}
__device__ __forceinline__ u64 add_u64(u64* p, u64 v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 add_u32(u32* p, u32 v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_u32(u32* p, u32 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Deterministic block sum of a u64 (for counters).
__device__ __forceinline__ u64 block_sum_u64(u64 v, u64* lds) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const u32 lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  u64 t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kWaves; ++w) t += lds[w];
  return t;  // valid in thread 0
}

}  // namespace ppvm
