// engine.hip — host implementation of the C ABI in include/ppvm_hip.h.
//
// A handle owns one HIP stream and its device arenas.  Gate calls are queued
// on the host and issued lazily so that (i) a truncation that directly follows
// a gate is fused into that gate's kernels, and (ii) runs of Clifford / noise /
// truncation ops become one k_map launch.  Kernels keep their counts in device
// memory (DevState), so no gate needs a host round trip; the host synchronises
// every kSyncEvery issued ops (or when asked for a result), checks the device
// overflow flag, and if a gate ran out of arena capacity it grows the arena and
// re-issues the gates from the one that overflowed (gates that see the flag do
// nothing, so the arena still holds that gate's input).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/ppvm_hip.h"
#include "kernels.hpp"
#include "small.hpp"

using namespace ppvm;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) return fail(PPVM_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

constexpr u64 kMinCap = 4096;
constexpr u64 kMaxCap = 1ull << 31;  // worklist indices are 32-bit
constexpr int kSyncEvery = 24;
constexpr int kMaxGrid = 1024;
constexpr int kOverlapGrid = 512;
constexpr u32 kGenLimit = (1u << 24) - 2;

enum IssuedKind { IK_ROT = 1, IK_MAP = 2, IK_SKIP = 3, IK_PRES = 4 };

// LDS-resident runs (k_small): steps collected per run, and the map programs among them.
constexpr u32 kSmallMaxSteps = 4096;
constexpr u32 kSmallMaxMaps = 512;
constexpr u32 kSmallThreads = 512;
constexpr size_t kSmallBufferBytes =
    (size_t)kSmallMaxSteps * sizeof(SmallStep) + (size_t)kSmallMaxMaps * sizeof(MapProgram);
struct Issued {
  int kind;
  u64 seq;
  u32 nskip;     // IK_SKIP: rotations in the run
  u32 nskipm;    // IK_SKIP: map ops in the run
  u32 mlo, mhi;  // IK_MAP collected for an LDS run: its source ops in small_mapops
  RotSpec rot;
  MapProgram map;
};

u64 pow2_at_least(u64 v) {
  u64 p = 1;
  while (p < v) p <<= 1;
  return p;
}

bool is_rotation(u32 c) { return c >= PPVM_OP_RX && c <= PPVM_OP_RZZ; }
// Amplitude damping is queued as two internal ops: this branch on the I/Z
// terms of the qubit, then the sqrt(1 - gamma) scaling of its X/Y terms (a map op
// under the public code, which also carries the truncation that follows).
constexpr u32 kOpDampBranch = 1044;
bool is_branch(u32 c) { return is_rotation(c) || c == kOpDampBranch; }
bool is_two_qubit(u32 c) {
  return c == PPVM_OP_CNOT || c == PPVM_OP_CZ || c == PPVM_OP_CY || c == PPVM_OP_RXX ||
         c == PPVM_OP_RYY || c == PPVM_OP_RZZ || c == PPVM_OP_DEPOLARIZE2 ||
         c == PPVM_OP_PAULI_ERROR2;
}
bool is_map_op(u32 c) {
  return (c >= PPVM_OP_X && c <= PPVM_OP_CY) || (c >= PPVM_OP_PAULI_ERROR && c <= PPVM_OP_AMPLITUDE_DAMPING);
}
// ops after which a truncated state stays truncated (sign flips only)
bool keeps_clean(u32 c) { return c == PPVM_OP_X || c == PPVM_OP_Y || c == PPVM_OP_Z; }

}  // namespace

namespace ppvm {

// two_qubit_pauli_error: indices of the channel Paulis
// (IX,IY,IZ,XI,XX,XY,XZ,YI,YX,YY,YZ,ZI,ZX,ZY,ZZ) anticommuting with (P_a,P_b),
// in the summation order of crates/ppvm-pauli-sum/src/sum/noise.rs:56-106
// (rows in the order I, X, Y, Z of P_a then P_b).  The order of the FP
// subtractions fixes the bits, so it is kept as in the reference.
const int kPe2Rows[16][8] = {
    {-1, -1, -1, -1, -1, -1, -1, -1}, {1, 10, 13, 14, 2, 5, 6, 9},   {0, 10, 12, 14, 2, 4, 6, 8},
    {0, 1, 12, 13, 4, 5, 8, 9},       {10, 11, 12, 13, 14, 7, 8, 9}, {1, 11, 12, 2, 5, 6, 7, 8},
    {0, 11, 13, 2, 4, 6, 7, 9},       {0, 1, 10, 11, 14, 4, 5, 7},   {11, 12, 13, 14, 3, 4, 5, 6},
    {1, 10, 11, 12, 2, 3, 4, 9},      {0, 10, 11, 13, 2, 3, 5, 8},   {0, 1, 11, 14, 3, 6, 8, 9},
    {10, 3, 4, 5, 6, 7, 8, 9},        {1, 13, 14, 2, 3, 4, 7, 8},    {0, 12, 14, 2, 3, 5, 7, 9},
    {0, 1, 10, 12, 13, 3, 6, 7},
};
// Pauli code x|z<<1 (I0 X1 Z2 Y3) -> letter index I0 X1 Y2 Z3 of the rows above.
const int kCodeToLetter[4] = {0, 1, 3, 2};

}  // namespace ppvm

struct ppvm_sum {
  u32 n = 0, W = 1;  // W: words per plane inside the engine (1, 2, 4 or 8)
  u32 Wa = 1;        // words per plane at the ABI: ceil(n/64)
  double tau = 0.0;
  u64 maxw = ~0ull;
  bool trunc_active = false;
  int device = 0;
  hipStream_t stream = nullptr;
  u64* keys[2] = {nullptr, nullptr};  // arena keys: 2W words per entry, contiguous (x words, then z words)
  double* coef[2] = {nullptr, nullptr};
  int cur = 0;
  u64 cap = 0;
  WMRec* wl = nullptr;  // per-gate worklist: one 16-byte record per anticommuting entry
  u64* table = nullptr;
  u64 tcap = 0;
  u64* masks = nullptr;     // ballot masks, cap/64 + kMaxParts words
  u64* tp = nullptr;        // bit-sliced key planes, 2 n planes of tstride words (k_rot_count)
  u64* hsh = nullptr;       // pair-key hash per worklist record (k_rot_gather)
  u64* bkt = nullptr;       // worklist records in bucket order (k_pair_scatter_tiled)
  u32* pmeta = nullptr;     // per-(bucket, slice) counts, then bucket bases and sizes
  u64 tstride = 0;
  u64* parts = nullptr;     // 6 x kMaxParts per-chunk counts / partial sums
  double* partials = nullptr;
  DevState* st = nullptr;
  DevState* st_host = nullptr;  // pinned
  u64* preserve = nullptr;
  u64* pres_found = nullptr;  // k_pres_find's list: count, then up to n_preserve arena indices
  u32 n_preserve = 0;
  std::vector<u64> preserve_host;
  u64 op_seq = 0;
  // Union of the x (and z) words of every term the sum may hold: exact after
  // an upload, then widened by every op that can set a bit, never narrowed.
  std::vector<u64> xsup, zsup;
  u32 skip_run = 0;   // rotations dropped since the last issued op (see k_skip_gates)
  u32 skip_maps = 0;  // map ops dropped since then (identity on every term)
  u32 tab_gen = 0;
  bool clean = true;
  bool broken = false;
  bool fresh = false;  // st_host holds the state after the last issued op, and sync_check's work is done
  std::vector<ppvm_op> pending;
  std::vector<Issued> log;
  int issued_since_sync = 0;
  u64 n_bound = 0;  // host-side upper bound of arena entries (grid sizing)
  u64 n_synced = 0; // arena entries at the last host synchronisation (path choice)
  u64 partition_min = 0;  // pair through the radix partition from this many entries on
  u32 bucket_lg_max = 0;  // at most 2^bucket_lg_max buckets in the partition
  bool pair_sort = false; // sort-based pairing inside the partition's buckets
  u32 bucket_target = 0;  // records per bucket aimed at
  ppvm_stats hs{};
  // LDS-resident runs of small sums: while small_on, the planner collects its
  // steps in small_list instead of issuing them, and small_flush runs them in
  // one k_small launch.
  bool small_on = false;
  bool small_off = false;   // off for this handle (PPVM_SMALL_MAX_ENTRIES=0, or a failed launch)
  u64 small_max = 0;        // collect while the synchronised arena holds at most this many entries
  u32 small_threads = kSmallThreads;
  std::vector<Issued> small_list;
  u32 small_maps = 0;       // map programs in small_list
  std::vector<ppvm_op> small_mapops;  // source ops of the collected map programs
  bool sup_stale = false;   // support is the planner's widened estimate after an LDS run
  // step records, then map programs, in one device buffer (filled from pinned staging by one copy)
  unsigned char* d_small = nullptr;
  unsigned char* h_small = nullptr;
  u64 small_runs = 0, small_steps = 0;
  // profiling
  bool prof = false;
  std::vector<hipEvent_t> ev;
  std::vector<int> ev_phase;
  size_t ev_used = 0;
  double prof_ms[8] = {0};
  u64 prof_n[8] = {0};
  u64* part(int i) const { return parts + (size_t)i * kMaxParts; }

  TruncSpec trunc_spec() const {
    TruncSpec t;
    t.tau = tau;
    t.maxw = maxw;
    t.preserve = preserve;
    t.n_preserve = n_preserve;
    t.W = W;
    return t;
  }
};

namespace {

// ---- profiling helpers ---------------------------------------------------------
int prof_begin(ppvm_sum* s, int phase) {
  if (!s->prof) return PPVM_OK;
  while (s->ev.size() < s->ev_used + 2) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    s->ev.push_back(e);
  }
  s->ev_phase.resize(s->ev.size() / 2 + 1);
  s->ev_phase[s->ev_used / 2] = phase;
  HIPCHK(hipEventRecord(s->ev[s->ev_used], s->stream));
  return PPVM_OK;
}
int prof_end(ppvm_sum* s) {
  if (!s->prof) return PPVM_OK;
  HIPCHK(hipEventRecord(s->ev[s->ev_used + 1], s->stream));
  s->ev_used += 2;
  return PPVM_OK;
}
int prof_collect(ppvm_sum* s) {
  for (size_t i = 0; i + 1 < s->ev_used; i += 2) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev[i], s->ev[i + 1]));
    int ph = s->ev_phase[i / 2];
    s->prof_ms[ph] += ms;
    s->prof_n[ph] += 1;
  }
  s->ev_used = 0;
  return PPVM_OK;
}

int grid_for(u64 n, u64 per_block, int maxg) {
  u64 g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > (u64)maxg) g = maxg;
  return (int)g;
}

u32 next_tab_gen(ppvm_sum* s) {
  if (s->tab_gen >= kGenLimit) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1024), dim3(256), 0, s->stream, s->table, s->tcap, 0ull);
    s->tab_gen = 0;
  }
  return ++s->tab_gen;
}

// ---- allocation -------------------------------------------------------------------
void free_buffers(ppvm_sum* s) {
  for (int i = 0; i < 2; ++i) {
    if (s->keys[i]) (void)hipFree(s->keys[i]);
    if (s->coef[i]) (void)hipFree(s->coef[i]);
    s->keys[i] = nullptr;
    s->coef[i] = nullptr;
  }
  if (s->wl) (void)hipFree(s->wl);
  if (s->table) (void)hipFree(s->table);
  if (s->masks) (void)hipFree(s->masks);
  if (s->tp) (void)hipFree(s->tp);
  if (s->hsh) (void)hipFree(s->hsh);
  if (s->bkt) (void)hipFree(s->bkt);
  s->wl = nullptr;
  s->table = nullptr;
  s->masks = nullptr;
  s->tp = nullptr;
  s->hsh = nullptr;
  s->bkt = nullptr;
}

// The bit-sliced planes no longer match the arena (new buffers, rewritten or
// moved keys): the next scan rebuilds them.
int invalidate_tp(ppvm_sum* s) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s->stream, &s->st->tvw, 1ull, 0ull);
  HIPCHK(hipGetLastError());
  return PPVM_OK;
}

// (Re)allocate all capacity-dependent buffers for capacity new_cap, keeping the
// first n_keep entries of the current arena.
int resize(ppvm_sum* s, u64 new_cap, u64 n_keep) {
  if (new_cap > kMaxCap) return fail(PPVM_ENOMEM, "arena capacity %llu exceeds the limit of %llu terms",
                                     (unsigned long long)new_cap, (unsigned long long)kMaxCap);
  const u32 np = 2 * s->W;
  u64* np0 = nullptr;
  double* nc0 = nullptr;
  u64* np1 = nullptr;
  double* nc1 = nullptr;
  if (hipMalloc(&np0, new_cap * np * 8) != hipSuccess || hipMalloc(&nc0, new_cap * 8) != hipSuccess ||
      hipMalloc(&np1, new_cap * np * 8) != hipSuccess || hipMalloc(&nc1, new_cap * 8) != hipSuccess) {
    if (np0) (void)hipFree(np0);
    if (nc0) (void)hipFree(nc0);
    if (np1) (void)hipFree(np1);
    if (nc1) (void)hipFree(nc1);
    (void)hipGetLastError();
    return fail(PPVM_ENOMEM, "device allocation of %llu-term arenas failed", (unsigned long long)new_cap);
  }
  if (n_keep && s->keys[s->cur]) {
    hipLaunchKernelGGL(k_copy_arena, dim3(grid_for(n_keep, 256, 4096)), dim3(256), 0, s->stream,
                       s->keys[s->cur], s->coef[s->cur], np0, nc0, np, n_keep);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  for (int i = 0; i < 2; ++i) {
    if (s->keys[i]) (void)hipFree(s->keys[i]);
    if (s->coef[i]) (void)hipFree(s->coef[i]);
  }
  s->keys[0] = np0;
  s->coef[0] = nc0;
  s->keys[1] = np1;
  s->coef[1] = nc1;
  s->cur = 0;
  if (s->wl) (void)hipFree(s->wl);
  if (s->table) (void)hipFree(s->table);
  if (s->masks) (void)hipFree(s->masks);
  if (s->tp) (void)hipFree(s->tp);
  if (s->hsh) (void)hipFree(s->hsh);
  if (s->bkt) (void)hipFree(s->bkt);
  s->wl = nullptr;
  s->table = nullptr;
  s->masks = nullptr;
  s->tp = nullptr;
  s->hsh = nullptr;
  s->bkt = nullptr;
  s->tcap = pow2_at_least(2 * new_cap);
  s->tstride = new_cap / 64 + 8;  // multiple of 8 words; room for the last 8-word group
  if (hipMalloc(&s->wl, new_cap * sizeof(WMRec)) != hipSuccess ||
      hipMalloc(&s->table, s->tcap * 8) != hipSuccess ||
      hipMalloc(&s->masks, (new_cap / 64 + kMaxParts + 64) * 8) != hipSuccess ||
      hipMalloc(&s->tp, 2 * (size_t)s->n * s->tstride * 8) != hipSuccess ||
      hipMalloc(&s->hsh, new_cap * 8) != hipSuccess || hipMalloc(&s->bkt, new_cap * 8) != hipSuccess) {
    (void)hipGetLastError();
    return fail(PPVM_ENOMEM, "device allocation of work buffers for %llu terms failed",
                (unsigned long long)new_cap);
  }
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(s->tcap, 256, 4096)), dim3(256), 0, s->stream, s->table,
                     s->tcap, 0ull);
  HIPCHK(hipGetLastError());
  s->tab_gen = 0;
  s->cap = new_cap;
  int rc;
  if ((rc = invalidate_tp(s))) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  return PPVM_OK;
}

// ---- kernel launches, dispatched on W ------------------------------------------
// Profiling phases (HIP events per launch).
enum Phase { PH_COUNT = 0, PH_GATHER = 1, PH_PAIR = 2, PH_ECOUNT = 3, PH_EWRITE = 4, PH_MAP = 5,
             PH_FINISH = 6, PH_OTHER = 7 };

// Blocks to launch for a chunked pass over at most `bound` items (the kernels
// derive the exact chunk count from the actual n; extra blocks exit at once).
u32 chunk_blocks(u64 bound, u32 target) {
  u64 g = (bound + target - 1) / target;
  if (g < 1) g = 1;
  if (g > (u64)kMaxParts) g = kMaxParts;
  return (u32)g;
}
constexpr int kFixedGrid = 2048;  // grid-stride kernels with per-block partials
constexpr u32 kArenaGrid = 4096;  // blocks of the chunked arena passes (count, gather)
constexpr u32 kListGrid = 2048;   // blocks of the chunked worklist passes (emit)
constexpr u32 kPairGrid = 256;    // blocks of k_pair_lds (one resident per CU; each loops over buckets)
// Arena size (at the last synchronisation) from which rotations pair through
// the radix partition instead of the single global-table kernel; the
// environment variable PPVM_PARTITION_MIN_ENTRIES overrides it per handle
// (both paths give identical results; 0 forces the partition everywhere).
constexpr u64 kPartitionEntries = 4u << 20;
u64 partition_min_entries() {
  const char* e = std::getenv("PPVM_PARTITION_MIN_ENTRIES");
  if (!e || !*e) return kPartitionEntries;
  char* end = nullptr;
  const unsigned long long v = std::strtoull(e, &end, 10);
  return (end && *end == '\0') ? (u64)v : kPartitionEntries;
}
// Largest bucket count (log2) of the partition; PPVM_PARTITION_MAX_LG lowers it
// (fewer, larger buckets take the multi-pass and global-table branches).
u32 bucket_lg_limit() {
  const char* e = std::getenv("PPVM_PARTITION_MAX_LG");
  if (!e || !*e) return kMaxBucketsLg;
  char* end = nullptr;
  const unsigned long v = std::strtoul(e, &end, 10);
  return (end && *end == '\0' && v < kMaxBucketsLg) ? (u32)v : kMaxBucketsLg;
}

// Records per bucket the partition aims at; PPVM_PARTITION_BUCKET_RECORDS
// lowers it (more, smaller buckets: the per-record scatter and small LDS
// passes at test sizes).
u32 bucket_target_env() {
  const char* e = std::getenv("PPVM_PARTITION_BUCKET_RECORDS");
  if (!e || !*e) return kBucketTarget;
  char* end = nullptr;
  const unsigned long v = std::strtoul(e, &end, 10);
  return (end && *end == '\0' && v >= 1 && v <= kBucketTarget) ? (u32)v : kBucketTarget;
}

// Pairing inside a bucket of the partition: the LDS hash table (default) or
// the sort-based variant (PPVM_PAIR_METHOD=sort: counting sort on the low
// hash bits, then a merge over runs of equal bits).  Both give identical
// results; the default was chosen by timing both on the lattice workload.
bool pair_sort_env() {
  const char* e = std::getenv("PPVM_PAIR_METHOD");
  return e && std::strcmp(e, "sort") == 0;
}

// Sums whose synchronised arena holds at most this many entries run their
// pending steps in LDS (k_small): three quarters of the LDS capacity by
// default, so a run has room to grow (a run that outgrows LDS hands over to
// the arena pipeline); PPVM_SMALL_MAX_ENTRIES overrides it (0 turns the path off;
// both paths give identical results).  PPVM_SMALL_THREADS sets the
// workgroup size (a multiple of 64 up to 1024).
u64 small_cap(u32 W);
u64 small_max_entries(u32 W) {
  const u64 dflt = small_cap(W) / 4 * 3;
  const char* e = std::getenv("PPVM_SMALL_MAX_ENTRIES");
  if (!e || !*e) return dflt;
  char* end = nullptr;
  const unsigned long long v = std::strtoull(e, &end, 10);
  return (end && *end == '\0') ? std::min<u64>((u64)v, small_cap(W)) : dflt;
}
u32 small_threads_env() {
  const char* e = std::getenv("PPVM_SMALL_THREADS");
  if (!e || !*e) return kSmallThreads;
  char* end = nullptr;
  const unsigned long v = std::strtoul(e, &end, 10);
  return (end && *end == '\0' && v >= 64 && v <= kSmallMaxThreads && v % 64 == 0) ? (u32)v : kSmallThreads;
}

// Arena size the launch grids of rotations and map programs are sized for.
// Those kernels loop over chunks (or grid-stride), so a grid smaller than the
// arena is only slower, never wrong; the host's hard bound doubles with every
// rotation since the last synchronisation, so grids follow the synchronised
// size with room for 4x growth instead.
u64 grid_entries(const ppvm_sum* s) { return std::min<u64>(s->n_bound, 4 * s->n_synced + (1u << 16)); }

int launch_finish(ppvm_sum* s, const FinishArgs& a) {
  int rc;
  if ((rc = prof_begin(s, PH_FINISH))) return rc;
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(kScanThreads), 0, s->stream, s->st, a);
  return prof_end(s);
}

template <int W>
int launch_rot(ppvm_sum* s, const Issued& it) {
  const u32 in = (u32)(it.seq & 1), out = in ^ 1u;
  int rc;
  // chunked kernels loop over their chunks, so the grids stay near one
  // resident wave of blocks however large the arena is
  const u64 ne = grid_entries(s);
  const u32 G = std::min<u32>(chunk_blocks(ne, chunk_target(ne, kArenaChunk)), kArenaGrid);
  const u32 Gl = std::min<u32>(chunk_blocks(ne, chunk_target(ne, kListChunk)), kListGrid);
  WMRec* wl = s->wl;
  u64* keys = s->keys[s->cur];
  double* coef = s->coef[s->cur];
  if ((rc = prof_begin(s, PH_COUNT))) return rc;
  hipLaunchKernelGGL(k_rot_count<W>, dim3(G), dim3(kBlock), 0, s->stream, keys, s->tp, s->tstride,
                     s->n, s->st, in, it.rot, s->masks, s->part(0));
  if ((rc = prof_end(s))) return rc;
  if ((rc = prof_begin(s, PH_GATHER))) return rc;
  hipLaunchKernelGGL(k_rot_gather<W>, dim3(G), dim3(kBlock), 0, s->stream, keys, coef, s->st, in,
                     s->masks, s->part(0), wl, it.rot, s->trunc_spec(), s->hsh);
  if ((rc = prof_end(s))) return rc;
  const u32 tg = next_tab_gen(s);
  u32 Gp;
  if ((rc = prof_begin(s, PH_PAIR))) return rc;
  if (s->n_synced >= s->partition_min) {
    // large gate: radix partition + LDS pairing
    Gp = kPairGrid;
    u32* hist = s->pmeta;
    u32* bmeta = s->pmeta + (size_t)kMaxBuckets * kPairSlices;
    const u32 lgm = s->bucket_lg_max;
    const PartSpec ps{lgm, s->bucket_target};
    hipLaunchKernelGGL(k_pair_hist, dim3(kPairSlices), dim3(kPartThreads), 0, s->stream, s->hsh, s->st, in,
                       s->cap, ps, it.seq + 1, hist);
    hipLaunchKernelGGL(k_pair_colscan, dim3((1u << lgm) / 64 + 1), dim3(kPartThreads), 0, s->stream, s->st, in,
                       s->cap, ps, hist, bmeta);
    hipLaunchKernelGGL(k_pair_scatter_tiled, dim3(kPairSlices), dim3(kPartThreads), 0, s->stream, s->hsh, s->st,
                       in, s->cap, ps, hist, s->bkt, bmeta);
    if (s->pair_sort)
      hipLaunchKernelGGL((k_pair_lds<W, true>), dim3(Gp), dim3(kPairThreads), 0, s->stream, s->cap, keys, wl,
                         s->st, in, it.rot, ps, s->hsh, s->bkt, bmeta, s->table, s->tcap, tg, s->trunc_spec(),
                         s->part(1), s->part(2));
    else
      hipLaunchKernelGGL((k_pair_lds<W, false>), dim3(Gp), dim3(kPairThreads), 0, s->stream, s->cap, keys, wl,
                         s->st, in, it.rot, ps, s->hsh, s->bkt, bmeta, s->table, s->tcap, tg, s->trunc_spec(),
                         s->part(1), s->part(2));
  } else {
    Gp = (u32)grid_for(ne, kBlock * 2, kFixedGrid);
    hipLaunchKernelGGL(k_rot_pair<W>, dim3(Gp), dim3(kBlock), 0, s->stream, s->cap, keys, wl, s->st, in,
                       it.rot, s->hsh, s->table, s->tcap, tg, s->trunc_spec(), it.seq + 1, s->part(1),
                       s->part(2));
  }
  if ((rc = prof_end(s))) return rc;
  if ((rc = prof_begin(s, PH_ECOUNT))) return rc;
  hipLaunchKernelGGL(k_emit_count<W>, dim3(Gl), dim3(kBlock), 0, s->stream, coef, s->cap, wl, s->st, in,
                     it.rot, s->trunc_spec(), s->masks, s->part(3), s->part(4));
  if ((rc = prof_end(s))) return rc;
  if ((rc = prof_begin(s, PH_EWRITE))) return rc;
  hipLaunchKernelGGL(k_emit_write<W>, dim3(Gl), dim3(kBlock), 0, s->stream, keys, coef, s->cap, wl, s->st,
                     in, out, it.rot, s->masks, s->part(3), s->part(4), s->part(1), s->part(2), Gp);
  if ((rc = prof_end(s))) return rc;
  HIPCHK(hipGetLastError());
  return PPVM_OK;
}

template <int W>
int launch_map(ppvm_sum* s, const Issued& it) {
  const u32 in = (u32)(it.seq & 1), out = in ^ 1u;
  int rc;
  const u32 Gm = (u32)grid_for(grid_entries(s), kBlock * 4, kFixedGrid);
  if ((rc = prof_begin(s, PH_MAP))) return rc;
  hipLaunchKernelGGL(k_map<W>, dim3(Gm), dim3(kBlock), 0, s->stream, s->keys[s->cur], s->coef[s->cur],
                     s->st, in, it.map, s->trunc_spec(), s->part(0), s->part(1));
  if ((rc = prof_end(s))) return rc;
  FinishArgs fa{};
  fa.mode = F_MAP;
  fa.in_slot = in;
  fa.out_slot = out;
  fa.Ga = Gm;
  fa.aux0 = s->part(0);
  fa.aux1 = s->part(1);
  fa.n_ops = it.map.n_ops + it.map.n_idle;
  if ((rc = launch_finish(s, fa))) return rc;
  HIPCHK(hipGetLastError());
  for (u32 i = 0; i < it.map.n_ops; ++i) {
    const u32 code = it.map.ops[i].code_flags & 0xffu;
    if (code >= PPVM_OP_H && code <= PPVM_OP_CY) return invalidate_tp(s);  // keys changed
  }
  return PPVM_OK;
}

// Preserved entries that the truncation just applied would have dropped move
// to the end of the term order (k_pres_find, k_pres_move).
template <int W>
int launch_pres(ppvm_sum* s, const Issued& it) {
  const u32 in = (u32)(it.seq & 1), out = in ^ 1u;
  int rc;
  const u32 G = (u32)grid_for(grid_entries(s), kBlock * 4, kFixedGrid);
  if ((rc = prof_begin(s, PH_OTHER))) return rc;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s->stream, s->pres_found, 1ull, 0ull);
  hipLaunchKernelGGL(k_pres_find<W>, dim3(G), dim3(kBlock), 0, s->stream, s->keys[s->cur], s->coef[s->cur],
                     s->st, in, s->trunc_spec(), s->pres_found);
  hipLaunchKernelGGL(k_pres_move<W>, dim3(1), dim3(kPresMoveThreads), 0, s->stream, s->keys[s->cur],
                     s->coef[s->cur], s->st, in, out, s->cap, s->n_preserve, s->pres_found, it.seq + 1);
  if ((rc = prof_end(s))) return rc;
  HIPCHK(hipGetLastError());
  return PPVM_OK;
}

// One LDS-resident run of the collected steps (k_small).
template <int W>
int launch_small(ppvm_sum* s, const Issued& it, u32 n_steps) {
  const u32 in = (u32)(it.seq & 1), out = in ^ 1u;
  SmallArgs a{};
  a.keys = s->keys[s->cur];
  a.coef = s->coef[s->cur];
  a.arena_cap = s->cap;
  a.st = s->st;
  a.in_slot = in;
  a.out_slot = out;
  a.steps = reinterpret_cast<const SmallStep*>(s->d_small);
  a.n_steps = n_steps;
  a.maps = reinterpret_cast<const MapProgram*>(s->d_small + (size_t)n_steps * sizeof(SmallStep));
  a.t = s->trunc_spec();
  int rc;
  if ((rc = prof_begin(s, PH_OTHER))) return rc;
  hipLaunchKernelGGL(k_small<W>, dim3(1), dim3(s->small_threads), 0, s->stream, a);
  if ((rc = prof_end(s))) return rc;
  HIPCHK(hipGetLastError());
  return PPVM_OK;
}

template <int W>
int launch_overlap(ppvm_sum* s) {
  int rc;
  const u32 in = (u32)(s->op_seq & 1);
  if ((rc = prof_begin(s, PH_OTHER))) return rc;
  hipLaunchKernelGGL(k_overlap<W>, dim3(kOverlapGrid), dim3(kBlock), 0, s->stream, s->keys[s->cur],
                     s->coef[s->cur], s->st, in, s->partials);
  if ((rc = prof_end(s))) return rc;
  FinishArgs fa{};
  fa.mode = F_OVERLAP;
  fa.Ga = kOverlapGrid;
  fa.dpart = s->partials;
  if ((rc = launch_finish(s, fa))) return rc;
  HIPCHK(hipGetLastError());
  return PPVM_OK;
}

template <int W>
int launch_trace(ppvm_sum* s, const PatternProg& prog) {
  int rc;
  const u32 in = (u32)(s->op_seq & 1);
  if ((rc = prof_begin(s, PH_OTHER))) return rc;
  hipLaunchKernelGGL(k_trace<W>, dim3(kOverlapGrid), dim3(kBlock), 0, s->stream, s->keys[s->cur],
                     s->coef[s->cur], s->st, in, prog, s->partials);
  if ((rc = prof_end(s))) return rc;
  FinishArgs fa{};
  fa.mode = F_OVERLAP;
  fa.Ga = kOverlapGrid;
  fa.dpart = s->partials;
  if ((rc = launch_finish(s, fa))) return rc;
  HIPCHK(hipGetLastError());
  return PPVM_OK;
}

// Σ_{k in b} c_b(k) c_a(k): a's live keys into a's table, then one probe per
// live term of b, all on a's stream (b is synchronised by the caller).
template <int W>
int launch_overlap2(ppvm_sum* a, ppvm_sum* b) {
  int rc;
  const u32 ia = (u32)(a->op_seq & 1), ib = (u32)(b->op_seq & 1);
  const u32 tg = next_tab_gen(a);
  const u64 tmask = a->tcap - 1;
  if ((rc = prof_begin(a, PH_OTHER))) return rc;
  hipLaunchKernelGGL(k_ovl_build<W>, dim3(grid_for(a->n_bound, kBlock * 4, kFixedGrid)), dim3(kBlock), 0,
                     a->stream, a->keys[a->cur], a->coef[a->cur], a->st, ia, a->table, tmask, tg);
  if ((rc = prof_end(a))) return rc;
  if ((rc = prof_begin(a, PH_OTHER))) return rc;
  hipLaunchKernelGGL(k_ovl_probe<W>, dim3(kOverlapGrid), dim3(kBlock), 0, a->stream, a->keys[a->cur],
                     a->coef[a->cur], a->table, tmask, tg, b->keys[b->cur], b->coef[b->cur],
                     b->st, ib, a->partials);
  if ((rc = prof_end(a))) return rc;
  FinishArgs fa{};
  fa.mode = F_OVERLAP;
  fa.Ga = kOverlapGrid;
  fa.dpart = a->partials;
  if ((rc = launch_finish(a, fa))) return rc;
  HIPCHK(hipGetLastError());
  return PPVM_OK;
}

template <int W>
int launch_compact(ppvm_sum* s) {
  int rc;
  const u32 in = (u32)(s->op_seq & 1), out = in ^ 1u;
  const u32 G = chunk_blocks(s->n_bound, kArenaChunk);
  if ((rc = prof_begin(s, PH_OTHER))) return rc;
  hipLaunchKernelGGL(k_compact_count, dim3(G), dim3(kBlock), 0, s->stream, s->coef[s->cur], s->st, in,
                     s->masks, s->part(0));
  if ((rc = prof_end(s))) return rc;
  if ((rc = prof_begin(s, PH_OTHER))) return rc;
  hipLaunchKernelGGL(k_compact_write<W>, dim3(G), dim3(kBlock), 0, s->stream, s->keys[s->cur],
                     s->coef[s->cur], s->keys[s->cur ^ 1], s->coef[s->cur ^ 1], s->st, in, out,
                     s->masks, s->part(0));
  if ((rc = prof_end(s))) return rc;
  HIPCHK(hipGetLastError());
  s->op_seq += 1;
  s->cur ^= 1;
  return invalidate_tp(s);
}

#define PPVM_DISPATCH_W(s, fn, ...)                  \
  switch ((s)->W) {                                  \
    case 1: return fn<1>(s, ##__VA_ARGS__);          \
    case 2: return fn<2>(s, ##__VA_ARGS__);          \
    case 4: return fn<4>(s, ##__VA_ARGS__);          \
    case 8: return fn<8>(s, ##__VA_ARGS__);          \
    default: return fail(PPVM_EINVAL, "unsupported word count %u", (s)->W); \
  }

int issue_rot_w(ppvm_sum* s, const Issued& it) { PPVM_DISPATCH_W(s, launch_rot, it); }
int issue_map_w(ppvm_sum* s, const Issued& it) { PPVM_DISPATCH_W(s, launch_map, it); }
int issue_pres_w(ppvm_sum* s, const Issued& it) { PPVM_DISPATCH_W(s, launch_pres, it); }
int small_w(ppvm_sum* s, const Issued& it, u32 n) { PPVM_DISPATCH_W(s, launch_small, it, n); }
int overlap_w(ppvm_sum* s) { PPVM_DISPATCH_W(s, launch_overlap); }
int compact_w(ppvm_sum* s) { PPVM_DISPATCH_W(s, launch_compact); }
int overlap2_w(ppvm_sum* s, ppvm_sum* b) { PPVM_DISPATCH_W(s, launch_overlap2, b); }
int trace_w(ppvm_sum* s, const PatternProg& prog) { PPVM_DISPATCH_W(s, launch_trace, prog); }

}  // namespace

namespace ppvm {

// Pauli pattern text -> matcher items (crates/ppvm-pauli-word/src/pattern/parse.rs:30-143).
// Symbol sets are kept as accepted Pauli codes; '?' adds the identity to one or
// two letters, and turns the three-letter set into the "any Pauli or identity"
// form, which (like '_') accepts the identity only (pattern/contains.rs:39).
int compile_pattern(const char* text, u32 nq, PatternProg& prog) {
  std::memset(&prog, 0, sizeof(prog));
  prog.nq = nq;
  std::string in(text);
  while (!in.empty() && std::isspace((unsigned char)in.back())) in.pop_back();
  size_t i = 0;
  while (i < in.size() && std::isspace((unsigned char)in[i])) ++i;
  auto letter = [](char c) -> int { return c == 'X' ? 1 : c == 'Z' ? 2 : c == 'Y' ? 3 : -1; };
  auto bad = [&](const char* why) { return fail(PPVM_EINVAL, "invalid Pauli pattern '%s': %s", text, why); };
  auto number = [&](u64& v) -> bool {  // usize::parse: digits only, no overflow
    size_t d = 0;
    v = 0;
    while (i < in.size() && std::isdigit((unsigned char)in[i])) {
      const u64 digit = (u64)(in[i++] - '0');
      if (v > (~0ull - digit) / 10) return false;
      v = v * 10 + digit;
      ++d;
    }
    return d > 0;
  };
  while (i < in.size()) {
    const char c = in[i++];
    u32 set = 0;
    if (c == '[') {
      while (i < in.size() && in[i] != ']') {
        const int l = letter(in[i++]);
        if (l < 0) return bad("expected X, Y or Z inside []");
        set |= 1u << l;
      }
      if (i < in.size()) ++i;  // ']'
      if (set == 0) return bad("empty []");
    } else if (c == '_') {
      set = 0;
    } else {
      const int l = letter(c);
      if (l < 0) return bad("expected X, Y, Z, _ or [");
      set = 1u << l;
    }
    const int letters = __builtin_popcount(set);
    u32 accept = letters ? set : 1u;
    if (i < in.size() && in[i] == '?') {
      ++i;
      accept = (letters == 1 || letters == 2) ? (set | 1u) : 1u;
    }
    if (prog.n_items >= (u32)kMaxPatItems) return bad("too many items");
    const u32 t = prog.n_items;
    prog.accept[t] = accept;
    if (i < in.size() && in[i] == '*') {
      ++i;
      prog.kind[t] = PK_STAR;
    } else if (i < in.size() && in[i] == '{') {
      ++i;
      u64 v;
      if (!number(v)) return bad("expected a count after '{'");
      if (i < in.size()) {
        if (in[i] != '}') return bad("expected '}'");
        ++i;
      }
      prog.kind[t] = PK_REPEAT;
      prog.arg[t] = v;
    } else {
      u64 v;
      if (!number(v)) return bad("expected '*', '{' or a qubit index after a symbol");
      prog.kind[t] = PK_POSITION;
      prog.arg[t] = v;
    }
    prog.n_items = t + 1;
  }
  return PPVM_OK;
}

}  // namespace ppvm

namespace {

int launch_skip(ppvm_sum* s, const Issued& it) {
  const u32 in = (u32)(it.seq & 1), out = in ^ 1u;
  int rc;
  if ((rc = prof_begin(s, PH_OTHER))) return rc;
  hipLaunchKernelGGL(k_skip_gates, dim3(1), dim3(64), 0, s->stream, s->st, in, out, it.nskip + it.nskipm,
                     it.nskip);
  if ((rc = prof_end(s))) return rc;
  HIPCHK(hipGetLastError());
  return PPVM_OK;
}

int issue(ppvm_sum* s, Issued& it, bool log_it) {
  s->fresh = false;
  int rc = it.kind == IK_ROT    ? issue_rot_w(s, it)
           : it.kind == IK_MAP  ? issue_map_w(s, it)
           : it.kind == IK_PRES ? issue_pres_w(s, it)
                                : launch_skip(s, it);
  if (rc) {
    s->broken = true;
    return rc;
  }
  if (it.kind == IK_ROT) {
    s->n_bound = std::min<u64>(s->cap, s->n_bound * 2 + 1);
  } else if (it.kind == IK_PRES) {
    s->n_bound = std::min<u64>(s->cap, s->n_bound + s->n_preserve);
  }
  if (log_it) s->log.push_back(it);
  return PPVM_OK;
}

int read_state(ppvm_sum* s) {
  HIPCHK(hipMemcpyAsync(s->st_host, s->st, sizeof(DevState), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  s->hs.syncs += 1;
  return prof_collect(s);
}

// Synchronise; on overflow grow and replay.  Afterwards the log is empty and
// st_host holds the state after the last issued op.
// Exact support of the live terms, read back at a synchronisation; rotations
// issued later widen it again (sup_after_rot, sup_after_map).  Above this many
// entries the pass costs about what the dropped gates save, and the support
// stays the widened estimate.
constexpr u64 kSupportRefreshMax = 16u << 20;

template <int W>
int launch_support(ppvm_sum* s, u32 G) {
  hipLaunchKernelGGL(k_support<W>, dim3(G), dim3(kBlock), 0, s->stream, s->keys[s->cur], s->coef[s->cur],
                     s->st, (u32)(s->op_seq & 1), s->part(5));
  HIPCHK(hipGetLastError());
  return PPVM_OK;
}

int refresh_support(ppvm_sum* s, u64 n) {
  const u32 W = s->W;
  s->xsup.assign(W, 0ull);
  s->zsup.assign(W, 0ull);
  if (n == 0) return PPVM_OK;
  const u32 G = (u32)grid_for(n, (u64)kBlock * 8, 1024);  // G * 2W <= kMaxParts
  int rc;
  if ((rc = prof_begin(s, PH_OTHER))) return rc;
  switch (W) {
    case 1: rc = launch_support<1>(s, G); break;
    case 2: rc = launch_support<2>(s, G); break;
    case 4: rc = launch_support<4>(s, G); break;
    default: rc = launch_support<8>(s, G); break;
  }
  if (rc) return rc;
  if ((rc = prof_end(s))) return rc;
  std::vector<u64> part((size_t)G * 2 * W);
  HIPCHK(hipMemcpyAsync(part.data(), s->part(5), part.size() * 8, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  for (u32 b = 0; b < G; ++b)
    for (u32 p = 0; p < W; ++p) {
      s->xsup[p] |= part[(size_t)b * 2 * W + p];
      s->zsup[p] |= part[(size_t)b * 2 * W + W + p];
    }
  return PPVM_OK;
}

template <int W>
int launch_fingerprint(ppvm_sum* s, u32 G) {
  hipLaunchKernelGGL(k_fingerprint<W>, dim3(G), dim3(kBlock), 0, s->stream, s->keys[s->cur], s->coef[s->cur],
                     s->st, (u32)(s->op_seq & 1), s->Wa, s->part(5));
  HIPCHK(hipGetLastError());
  return PPVM_OK;
}

// Support of the state after the last issued op into st->sup, so that the
// read-back of the state carries it (no second round trip).
int launch_support_dev(ppvm_sum* s) {
  const u32 W = s->W;
  const u32 G = (u32)grid_for(std::max<u64>(s->n_bound, 1), (u64)kBlock * 8, 1024);
  int rc;
  if ((rc = prof_begin(s, PH_OTHER))) return rc;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s->stream, s->st->sup, (u64)(2 * W), 0ull);
  HIPCHK(hipGetLastError());
  u64* none = nullptr;
  switch (W) {
    case 1: hipLaunchKernelGGL(k_support<1>, dim3(G), dim3(kBlock), 0, s->stream, s->keys[s->cur], s->coef[s->cur],
                               s->st, (u32)(s->op_seq & 1), none); break;
    case 2: hipLaunchKernelGGL(k_support<2>, dim3(G), dim3(kBlock), 0, s->stream, s->keys[s->cur], s->coef[s->cur],
                               s->st, (u32)(s->op_seq & 1), none); break;
    case 4: hipLaunchKernelGGL(k_support<4>, dim3(G), dim3(kBlock), 0, s->stream, s->keys[s->cur], s->coef[s->cur],
                               s->st, (u32)(s->op_seq & 1), none); break;
    default: hipLaunchKernelGGL(k_support<8>, dim3(G), dim3(kBlock), 0, s->stream, s->keys[s->cur],
                                s->coef[s->cur], s->st, (u32)(s->op_seq & 1), none); break;
  }
  HIPCHK(hipGetLastError());
  return prof_end(s);
}

int sync_check(ppvm_sum* s) {
  if (s->broken) return fail(PPVM_ESTATE, "handle is unusable after an earlier device error");
  if (s->fresh) return PPVM_OK;
  int rc;
  // the support rides along with the state read-back when the arena is small
  // enough for the pass to pay (an upper bound decides; it stays valid across
  // compaction and growth, which move live entries without changing them)
  bool sup_ok = s->n_bound <= kSupportRefreshMax;
  if (sup_ok && (rc = launch_support_dev(s))) {
    s->broken = true;
    return rc;
  }
  for (int round = 0; round < 64; ++round) {
    if ((rc = read_state(s))) {
      s->broken = true;
      return rc;
    }
    const u64 of = s->st_host->overflow;
    if (!of) break;
    sup_ok = false;  // the state is rebuilt by the replay below
    const u64 seq = of - 1;
    const CountSlot& c = s->st_host->slot[seq & 1];
    const u64 live = c.n_entries - c.n_dead;
    u64 want = pow2_at_least(std::max<u64>(2 * s->cap, c.n_entries + 2 * live + 1));
    if ((rc = resize(s, want, c.n_entries))) {
      s->broken = true;
      return rc;
    }
    s->hs.regrows += 1;
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s->stream, &s->st->overflow, 1ull, 0ull);
    HIPCHK(hipGetLastError());
    s->n_bound = c.n_entries;
    s->n_synced = s->n_bound;
    std::vector<Issued> redo;
    for (auto& it : s->log)
      if (it.seq >= seq) redo.push_back(it);
    for (auto& it : redo)
      if ((rc = issue(s, it, false))) return rc;
  }
  if (s->st_host->overflow) {
    s->broken = true;
    return fail(PPVM_ENOMEM, "arena overflow persisted after repeated growth");
  }
  s->log.clear();
  s->issued_since_sync = 0;
  const CountSlot& c = s->st_host->slot[s->op_seq & 1];
  s->n_bound = c.n_entries;
  s->n_synced = s->n_bound;
  const u64 live = c.n_entries - c.n_dead;
  s->hs.peak_live = std::max(s->hs.peak_live, live);
  // compaction when dead entries dominate the scans
  if (c.n_dead > 4096 && c.n_dead * 8 > c.n_entries) {
    if ((rc = compact_w(s))) {
      s->broken = true;
      return rc;
    }
    s->hs.compactions += 1;
    if ((rc = read_state(s))) {
      s->broken = true;
      return rc;
    }
    s->n_bound = s->st_host->slot[s->op_seq & 1].n_entries;
    s->n_synced = s->n_bound;
  }
  // proactive growth: keep room for a full gate of children
  const CountSlot& c2 = s->st_host->slot[s->op_seq & 1];
  const u64 live2 = c2.n_entries - c2.n_dead;
  if (c2.n_entries + 2 * live2 > s->cap) {
    u64 want = pow2_at_least(c2.n_entries + 2 * live2 + 1);
    if ((rc = resize(s, want, c2.n_entries))) {
      s->broken = true;
      return rc;
    }
    s->hs.regrows += 1;
  }
  if (c2.n_entries <= kSupportRefreshMax) {
    if (sup_ok) {
      s->xsup.assign(s->st_host->sup, s->st_host->sup + s->W);
      s->zsup.assign(s->st_host->sup + s->W, s->st_host->sup + 2 * s->W);
    } else if ((rc = refresh_support(s, c2.n_entries))) {
      s->broken = true;
      return rc;
    }
  }
  s->fresh = true;
  return PPVM_OK;
}

// ---- op translation -----------------------------------------------------------
RotSpec make_rot(const ppvm_op& op, bool trunc) {
  RotSpec g{};
  if (op.code == kOpDampBranch) {  // Z -> (1 - gamma) Z + gamma I on q0 (noise.rs:146-155)
    g.a = g.b = op.q0;
    g.gaz = 1;
    g.mode = kBranchDamping;
    g.trunc = trunc ? 1u : 0u;
    g.test_planes = 1;
    g.cs = 1.0 - op.p[0];
    g.sn = op.p[0];
    return g;
  }
  g.a = op.q0;
  g.b = op.q1;
  g.two = 0;
  switch (op.code) {
    case PPVM_OP_RX: g.gax = 1; break;
    case PPVM_OP_RY: g.gax = 1; g.gaz = 1; break;
    case PPVM_OP_RZ: g.gaz = 1; break;
    case PPVM_OP_RXX: g.two = 1; g.gax = 1; g.gbx = 1; break;
    case PPVM_OP_RYY: g.two = 1; g.gax = g.gaz = g.gbx = g.gbz = 1; break;
    case PPVM_OP_RZZ: g.two = 1; g.gaz = 1; g.gbz = 1; break;
  }
  if (!g.two) g.b = g.a;
  g.trunc = trunc ? 1u : 0u;
  g.test_planes = g.gaz + g.gax + (g.two ? g.gbz + g.gbx : 0u);
  // Float::sin_cos of the reference (crates/ppvm-traits/src/traits/coefficient.rs:55-57):
  // one libm sincos per gate, the combined call LLVM makes for f64::sin_cos on
  // Linux.  glibc's separate sin() can differ from it in the last bit, so the
  // oracle and the batched path call sincos too.
  sincos(op.theta, &g.sn, &g.cs);
  return g;
}

// Adds one map op (and its factor table) to prog; returns false if full.
bool map_add(MapProgram& p, u32& ntables, const ppvm_op& op, bool trunc) {
  if (p.n_ops >= (u32)kMaxMapOps) return false;
  const bool noise = op.code >= PPVM_OP_PAULI_ERROR && op.code <= PPVM_OP_AMPLITUDE_DAMPING;
  if (noise && ntables >= (u32)kMaxFactorTables) return false;
  MapOp m{};
  m.a = op.q0;
  m.b = op.q1;
  u32 code = 0, table = 0;
  if (!noise) {
    code = op.code;  // Clifford op codes coincide with MapCode values
  } else {
    code = M_SCALE;
    table = ntables++;
    double* f = p.factors[table];
    u32 mask = 0;
    for (int i = 0; i < 16; ++i) f[i] = 1.0;
    if (op.code == PPVM_OP_AMPLITUDE_DAMPING) {  // X, Y: *= sqrt(1 - gamma), noise.rs:141-144
      const double decay = std::sqrt(1.0 - op.p[0]);
      f[1 * 4] = decay;  // code X = 1
      f[3 * 4] = decay;  // code Y = 3
      mask = (1u << 4) | (1u << 12);
      m.b = 0xffffffffu;
    } else if (op.code == PPVM_OP_PAULI_ERROR) {  // noise.rs:17-39
      const double one = 1.0;
      const double xf = one - op.p[1] * 2.0 - op.p[2] * 2.0;
      const double zf = one - op.p[0] * 2.0 - op.p[1] * 2.0;
      const double yf = one - op.p[0] * 2.0 - op.p[2] * 2.0;
      f[1 * 4] = xf;  // code X = 1
      f[2 * 4] = zf;  // code Z = 2
      f[3 * 4] = yf;  // code Y = 3
      mask = (1u << 4) | (1u << 8) | (1u << 12);
      m.b = 0xffffffffu;
    } else if (op.code == PPVM_OP_DEPOLARIZE1) {  // noise.rs:109-118
      const double fac = 1.0 - op.p[0] * (4.0 / 3.0);
      for (int ca = 1; ca < 4; ++ca) {
        f[ca * 4] = fac;
        mask |= 1u << (ca * 4);
      }
      m.b = 0xffffffffu;
    } else if (op.code == PPVM_OP_DEPOLARIZE2) {  // noise.rs:121-131
      const double fac = 1.0 - op.p[0] * (16.0 / 15.0);
      for (int i = 1; i < 16; ++i) {
        f[i] = fac;
        mask |= 1u << i;
      }
    } else {  // PAULI_ERROR2, noise.rs:52-107
      for (int ca = 0; ca < 4; ++ca)
        for (int cb = 0; cb < 4; ++cb) {
          const int row = 4 * kCodeToLetter[ca] + kCodeToLetter[cb];
          if (row == 0) continue;
          double acc = 1.0;
          for (int j = 0; j < 8; ++j) acc = acc - op.p[kPe2Rows[row][j]] * 2.0;
          f[ca * 4 + cb] = acc;
          mask |= 1u << (ca * 4 + cb);
        }
    }
    p.apply_mask[table] = mask;
  }
  m.code_flags = code | ((trunc ? 1u : 0u) << 8) | (table << 16);
  p.ops[p.n_ops++] = m;
  return true;
}

int small_flush(ppvm_sum* s);
bool small_eligible(const ppvm_sum* s);

int issue_and_maybe_sync(ppvm_sum* s, Issued& it) {
  if (s->small_on) {
    if (s->small_list.size() >= kSmallMaxSteps || (it.kind == IK_MAP && s->small_maps >= kSmallMaxMaps)) {
      int rc = small_flush(s);  // full: run what is collected, then go on collecting if still small
      if (rc) return rc;
      s->small_on = small_eligible(s);
    }
    if (s->small_on && it.kind != IK_PRES) {
      s->small_maps += it.kind == IK_MAP ? 1u : 0u;
      s->small_list.push_back(it);
      return PPVM_OK;
    }
  }
  it.seq = s->op_seq++;
  int rc = issue(s, it, true);
  if (rc) return rc;
  if (++s->issued_since_sync >= kSyncEvery) return sync_check(s);
  return PPVM_OK;
}

// ---- support of the sum (host side) ------------------------------------------------
bool sup_bit(const std::vector<u64>& m, u32 q) { return (m[q >> 6] >> (q & 63)) & 1ull; }
void sup_put(std::vector<u64>& m, u32 q, bool v) {
  if (v) m[q >> 6] |= 1ull << (q & 63);
  else m[q >> 6] &= ~(1ull << (q & 63));
}
void sup_from_keys(ppvm_sum* s, const std::vector<u64>& keys, size_t n) {
  const u32 W = s->W;
  s->xsup.assign(W, 0ull);
  s->zsup.assign(W, 0ull);
  for (size_t i = 0; i < n; ++i)
    for (u32 p = 0; p < W; ++p) {
      s->xsup[p] |= keys[i * 2 * W + p];
      s->zsup[p] |= keys[i * 2 * W + W + p];
    }
}
// Can some term anticommute with the rotation's generator?  A term
// anticommutes only if it has x (or z) set on a generator qubit where the
// generator has z (or x), so a clear support bit rules the gate out.
bool rot_may_act(const ppvm_sum* s, const RotSpec& g) {
  bool r = (g.gaz && sup_bit(s->xsup, g.a)) || (g.gax && sup_bit(s->zsup, g.a));
  if (g.two) r = r || (g.gbz && sup_bit(s->xsup, g.b)) || (g.gbx && sup_bit(s->zsup, g.b));
  return r;
}
// Children of a rotation differ from their parents in the generator's bits.
void sup_after_rot(ppvm_sum* s, const RotSpec& g) {
  if (g.mode == kBranchDamping) return;  // Z -> I only clears bits
  if (g.gax) sup_put(s->xsup, g.a, true);
  if (g.gaz) sup_put(s->zsup, g.a, true);
  if (g.two) {
    if (g.gbx) sup_put(s->xsup, g.b, true);
    if (g.gbz) sup_put(s->zsup, g.b, true);
  }
}
// Clifford bit maps applied to the support (a swap is exact, an XOR becomes an
// OR); sign-only gates, noise and truncation leave it as it is.
// A map op (Clifford or diagonal noise) on qubits where every term holds the
// identity changes nothing: Cliffords map I to I with sign +, and every noise
// factor of the identity is 1.
bool map_is_identity(const ppvm_sum* s, const ppvm_op& op) {
  if (sup_bit(s->xsup, op.q0) || sup_bit(s->zsup, op.q0)) return false;
  return !is_two_qubit(op.code) || (!sup_bit(s->xsup, op.q1) && !sup_bit(s->zsup, op.q1));
}

void sup_after_map(ppvm_sum* s, const ppvm_op& op) {
  const u32 a = op.q0, b = op.q1;
  const bool xa = sup_bit(s->xsup, a), za = sup_bit(s->zsup, a);
  switch (op.code) {
    case PPVM_OP_H:
    case PPVM_OP_SQRT_Y:
    case PPVM_OP_SQRT_Y_DAG:
      sup_put(s->xsup, a, za);
      sup_put(s->zsup, a, xa);
      break;
    case PPVM_OP_S:
    case PPVM_OP_S_DAG:
      if (xa) sup_put(s->zsup, a, true);
      break;
    case PPVM_OP_SQRT_X:
    case PPVM_OP_SQRT_X_DAG:
      if (za) sup_put(s->xsup, a, true);
      break;
    case PPVM_OP_CNOT: {  // z_a ^= z_b, x_b ^= x_a
      const bool zb = sup_bit(s->zsup, b);
      if (zb) sup_put(s->zsup, a, true);
      if (xa) sup_put(s->xsup, b, true);
      break;
    }
    case PPVM_OP_CZ: {  // z_a ^= x_b, z_b ^= x_a
      const bool xb = sup_bit(s->xsup, b);
      if (xb) sup_put(s->zsup, a, true);
      if (xa) sup_put(s->zsup, b, true);
      break;
    }
    case PPVM_OP_CY: {  // z_a ^= x_b ^ z_b, x_b ^= x_a, z_b ^= x_a
      const bool xb = sup_bit(s->xsup, b), zb = sup_bit(s->zsup, b);
      if (xb || zb) sup_put(s->zsup, a, true);
      if (xa) {
        sup_put(s->xsup, b, true);
        sup_put(s->zsup, b, true);
      }
      break;
    }
    default:
      break;
  }
}

// A truncation with a preserve set is followed by the op that moves the
// preserved entries it would drop to the end.
bool pres_active(const ppvm_sum* s) { return s->n_preserve > 0 && s->trunc_active; }
int issue_pres(ppvm_sum* s) {
  Issued it{};
  it.kind = IK_PRES;
  return issue_and_maybe_sync(s, it);
}

// Issues the pending run of dropped rotations (their bookkeeping kernel).
int flush_skips(ppvm_sum* s) {
  if (!s->skip_run && !s->skip_maps) return PPVM_OK;
  Issued it{};
  it.kind = IK_SKIP;
  it.nskip = s->skip_run;
  it.nskipm = s->skip_maps;
  s->skip_run = 0;
  s->skip_maps = 0;
  return issue_and_maybe_sync(s, it);
}

// ---- LDS-resident runs of small sums ------------------------------------------------
u64 small_cap(u32 W) {
  return W == 1 ? SmallCap<1>::cap : W == 2 ? SmallCap<2>::cap : W == 4 ? SmallCap<4>::cap : SmallCap<8>::cap;
}

// A pump collects steps for k_small when the state is synchronised (nothing
// issued since, so the run starts from known counts), has no preserve set, and
// the arena is small.
bool small_eligible(const ppvm_sum* s) {
  return !s->small_off && !s->broken && s->n_preserve == 0 && s->log.empty() && s->issued_since_sync == 0 &&
         s->n_synced <= s->small_max;
}

int small_buffers(ppvm_sum* s) {
  if (s->d_small && s->h_small) return PPVM_OK;
  if ((!s->d_small && hipMalloc(&s->d_small, kSmallBufferBytes) != hipSuccess) ||
      (!s->h_small && hipHostMalloc(&s->h_small, kSmallBufferBytes, hipHostMallocDefault) != hipSuccess)) {
    (void)hipGetLastError();
    if (s->d_small) (void)hipFree(s->d_small);
    if (s->h_small) (void)hipHostFree(s->h_small);
    s->d_small = nullptr;
    s->h_small = nullptr;
    return fail(PPVM_ENOMEM, "allocation of the small-run step buffers failed");
  }
  return PPVM_OK;
}

// Runs the collected steps in one k_small launch, synchronises, and issues the
// steps it could not complete (the sum outgrew LDS) through the arena pipeline.
int small_flush(ppvm_sum* s) {
  if (s->small_list.empty()) return PPVM_OK;
  std::vector<Issued> list;
  list.swap(s->small_list);
  s->small_maps = 0;
  s->small_on = false;
  int rc;
  if ((rc = small_buffers(s))) return rc;
  // step records, then the map programs right after them
  SmallStep* hs = reinterpret_cast<SmallStep*>(s->h_small);
  MapProgram* hm = reinterpret_cast<MapProgram*>(s->h_small + list.size() * sizeof(SmallStep));
  u32 nm = 0;
  for (size_t k = 0; k < list.size(); ++k) {
    const Issued& it = list[k];
    SmallStep& st = hs[k];
    std::memset(&st, 0, sizeof(st));
    if (it.kind == IK_ROT) {
      st.kind = SK_ROT;
      st.rot = it.rot;
    } else if (it.kind == IK_MAP) {
      st.kind = SK_MAP;
      st.arg = nm;
      hm[nm++] = it.map;
    } else {
      st.kind = SK_SKIP;
      st.arg = it.nskip;
      st.arg2 = it.nskipm;
    }
  }
  const u32 n = (u32)list.size();
  const size_t bytes = (size_t)n * sizeof(SmallStep) + (size_t)nm * sizeof(MapProgram);
  HIPCHK(hipMemcpyAsync(s->d_small, s->h_small, bytes, hipMemcpyHostToDevice, s->stream));
  Issued run{};
  run.seq = s->op_seq;
  if ((rc = small_w(s, run, n))) {
    // not launched: the slots are as they were; run everything through the arena pipeline
    s->small_off = true;
    for (auto& it : list)
      if ((rc = issue_and_maybe_sync(s, it))) return rc;
    return PPVM_OK;
  }
  s->op_seq += 1;
  if ((rc = read_state(s))) {
    s->broken = true;
    return rc;
  }
  const u64 done = s->st_host->small_done;
  const CountSlot& c = s->st_host->slot[s->op_seq & 1];
  s->n_bound = c.n_entries;
  s->n_synced = c.n_entries;
  s->hs.peak_live = std::max<u64>(s->hs.peak_live, s->st_host->small_peak);
  s->small_runs += 1;
  s->small_steps += done;
  if (done == list.size()) {
    // synchronised with nothing pending: sync_check has nothing to add but the
    // arena's room for a gate of children; the support is the planner's
    // estimate, widened over every collected step, and is read back exactly
    // before the next ops are planned
    s->log.clear();
    s->issued_since_sync = 0;
    if (c.n_entries * 3 + 1 > s->cap && (rc = resize(s, pow2_at_least(c.n_entries * 3 + 1), c.n_entries))) {
      s->broken = true;
      return rc;
    }
    s->fresh = true;
    s->sup_stale = true;
    return PPVM_OK;
  }
  // The sum outgrew LDS at step `done`.  The remaining steps were planned with
  // the support widened over the whole collection, so rotations that act on no
  // term of the actual state would be launched: decide again from the exact
  // support of the state the run left, widened by the remaining steps as they
  // are issued.  Dropping a rotation that meets no anticommuting term changes
  // nothing (a truncation fused with it was planned on a clean state).
  if ((rc = refresh_support(s, c.n_entries))) {
    s->broken = true;
    return rc;
  }
  s->sup_stale = false;
  for (size_t k = done; k < list.size(); ++k) {
    Issued& it = list[k];
    if (it.kind == IK_SKIP) {
      s->skip_run += it.nskip;
      s->skip_maps += it.nskipm;
      continue;
    }
    if (it.kind == IK_ROT && it.rot.mode != kBranchDamping && !rot_may_act(s, it.rot)) {
      ++s->skip_run;
      continue;
    }
    if ((rc = flush_skips(s))) return rc;
    if ((rc = issue_and_maybe_sync(s, it))) return rc;
    if (it.kind == IK_ROT) {
      sup_after_rot(s, it.rot);
    } else if (it.kind == IK_MAP) {
      for (u32 j = it.mlo; j < it.mhi; ++j) sup_after_map(s, s->small_mapops[j]);
    }
  }
  return flush_skips(s);
}

// Turns queued ops into kernel launches.// Turns queued ops into kernel launches.  With final=false, a trailing op whose
// fusion partner may still arrive (a rotation that a truncate could follow, a
// run of map ops) is kept queued.
int pump_ops(ppvm_sum* s, bool final);

// Collection for an LDS-resident run starts at a pump on a small synchronised
// state and goes on over later pumps; the run is launched by the final pump
// that every call reading or changing the state makes first (or earlier, when
// the step list is full).
int pump(ppvm_sum* s, bool final) {
  if (s->broken) return fail(PPVM_ESTATE, "handle is unusable after an earlier device error");
  if (!s->small_on) {
    s->small_list.clear();
    s->small_maps = 0;
    s->small_mapops.clear();
    s->small_on = small_eligible(s);
  }
  if (s->sup_stale && !s->pending.empty() && s->log.empty() && s->issued_since_sync == 0) {
    // an LDS run left the planner's estimate: plan the next ops from the exact
    // support (the state is synchronised and small)
    const u64 n = s->n_synced;
    if (n <= kSupportRefreshMax) {
      int rc = refresh_support(s, n);
      if (rc) {
        s->broken = true;
        return rc;
      }
    }
    s->sup_stale = false;
  }
  int rc = pump_ops(s, final);
  if (!rc && final) rc = small_flush(s);
  if (rc || final) s->small_on = false;
  return rc;
}

int pump_ops(ppvm_sum* s, bool final) {
  size_t i = 0;
  auto& q = s->pending;
  int rc = PPVM_OK;
  while (i < q.size()) {
    const ppvm_op& op = q[i];
    if (op.code == PPVM_OP_NOP) {
      ++i;
      continue;
    }
    if (is_branch(op.code)) {
      if (i + 1 >= q.size() && !final) break;
      const bool trunc_next = i + 1 < q.size() && q[i + 1].code == PPVM_OP_TRUNCATE;
      const bool fused = trunc_next && s->clean && s->trunc_active;
      Issued it{};
      it.kind = IK_ROT;
      it.rot = make_rot(op, fused);
      if (op.code != kOpDampBranch && !rot_may_act(s, it.rot)) {
        // no term anticommutes with the generator: the gate changes nothing,
        // and neither does a truncate fused with it on a clean state
        ++s->skip_run;
        i += fused ? 2 : 1;
        if (fused) {
          while (i < q.size() && q[i].code == PPVM_OP_TRUNCATE) ++i;
        }
        continue;
      }
      if ((rc = flush_skips(s))) return rc;
      if ((rc = issue_and_maybe_sync(s, it))) return rc;
      if (fused && pres_active(s) && (rc = issue_pres(s))) return rc;
      sup_after_rot(s, it.rot);
      i += fused ? 2 : 1;
      if (!fused && s->trunc_active) s->clean = false;
      if (fused) {  // drop repeated truncates (idempotent)
        while (i < q.size() && q[i].code == PPVM_OP_TRUNCATE) ++i;
      }
      continue;
    }
    // a run of map ops and truncates
    size_t j = i;
    bool ends_queue = true;
    while (j < q.size()) {
      if (is_branch(q[j].code)) {
        ends_queue = false;
        break;
      }
      ++j;
    }
    if (ends_queue && !final) {
      // wait for more ops unless the run already fills a program
      size_t nmap = 0;
      for (size_t k = i; k < j; ++k) nmap += is_map_op(q[k].code);
      if (nmap < (size_t)kMaxMapOps) break;
    }
    Issued it{};
    it.kind = IK_MAP;
    it.mlo = it.mhi = (u32)s->small_mapops.size();
    MapProgram& p = it.map;
    std::memset(&p, 0, sizeof(p));
    u32 ntables = 0;
    size_t k = i;
    bool clean = s->clean || !s->trunc_active;
    bool any = false;
    // with a preserve set, each truncation ends its program: the preserved
    // entries it would drop move to the end before the next op runs
    const bool pres = pres_active(s);
    bool event = false;
    while (k < j && q[k].code == PPVM_OP_TRUNCATE) {
      if (s->trunc_active && !clean) p.trunc_before = 1;
      event = event || p.trunc_before;
      clean = true;
      ++k;
    }
    while (k < j && !(pres && event)) {
      const ppvm_op& m = q[k];
      if (m.code == PPVM_OP_NOP) {
        ++k;
        continue;
      }
      const bool tnext = k + 1 < q.size() && q[k + 1].code == PPVM_OP_TRUNCATE && s->trunc_active;
      if (map_is_identity(s, m) && (!tnext || clean) && p.idle[p.n_ops] < 0xffffu) {
        // dropped (a truncation after it would find the state it left, clean);
        // the program books its term updates at its position
        ++p.idle[p.n_ops];
        ++p.n_idle;
        ++k;
        if (tnext) {
          while (k < j && q[k].code == PPVM_OP_TRUNCATE) ++k;
        }
        continue;
      }
      if (!map_add(p, ntables, m, tnext)) break;  // program full: issue and continue
      sup_after_map(s, m);
      if (s->small_on) s->small_mapops.push_back(m);
      any = true;
      ++k;
      if (tnext) {
        clean = true;
        event = true;
        while (k < j && q[k].code == PPVM_OP_TRUNCATE) ++k;
      } else if (s->trunc_active && !keeps_clean(m.code)) {
        clean = false;
      }
    }
    it.mhi = (u32)s->small_mapops.size();
    if (!any && !p.trunc_before && p.n_idle) {
      // only dropped ops: booked with the dropped gates (no truncation point
      // among them, so they see the current live count)
      s->skip_maps += p.n_idle;
    }
    if (any || p.trunc_before) {
      if ((rc = flush_skips(s))) return rc;
      if ((rc = issue_and_maybe_sync(s, it))) return rc;
      if (pres && event && (rc = issue_pres(s))) return rc;
    }
    s->clean = clean;
    i = k;
  }
  q.erase(q.begin(), q.begin() + (long)i);
  return flush_skips(s);
}

int enqueue(ppvm_sum* s, const ppvm_op& op) {
  if (op.code == PPVM_OP_AMPLITUDE_DAMPING) {
    ppvm_op br = op;
    br.code = kOpDampBranch;
    s->pending.push_back(br);
  }
  s->pending.push_back(op);
  if (s->pending.size() >= 4 * (size_t)kMaxMapOps) return pump(s, false);
  return PPVM_OK;
}

int validate(const ppvm_sum* s, const ppvm_op& op) {
  const u32 c = op.code;
  const bool known = c == PPVM_OP_NOP || c == PPVM_OP_TRUNCATE || is_rotation(c) || is_map_op(c);
  if (!known) return fail(PPVM_EINVAL, "unknown op code %u", c);
  if (c == PPVM_OP_NOP || c == PPVM_OP_TRUNCATE) return PPVM_OK;
  if (op.q0 >= s->n) return fail(PPVM_EINVAL, "qubit index %u out of range for %u qubits", op.q0, s->n);
  if (is_two_qubit(c)) {
    if (op.q1 >= s->n) return fail(PPVM_EINVAL, "qubit index %u out of range for %u qubits", op.q1, s->n);
    if (op.q1 == op.q0) return fail(PPVM_EINVAL, "two-qubit op on identical qubits (%u, %u)", op.q0, op.q1);
  }
  if (is_rotation(c) && !std::isfinite(op.theta)) return fail(PPVM_EINVAL, "rotation angle is not finite");
  if (c == PPVM_OP_AMPLITUDE_DAMPING && !std::isfinite(op.p[0]))
    return fail(PPVM_EINVAL, "damping rate is not finite");
  return PPVM_OK;
}

int push_gate(ppvm_sum* s, u32 code, u32 q0, u32 q1, double theta, const double* p, int np) {
  ppvm_op op{};
  op.code = code;
  op.q0 = q0;
  op.q1 = q1;
  op.theta = theta;
  for (int i = 0; i < np; ++i) op.p[i] = p[i];
  int rc = validate(s, op);
  if (rc) return rc;
  return enqueue(s, op);
}

int gate_1q(ppvm_handle h, u32 code, const uint32_t* t, size_t n, double theta) {
  if (!h) return fail(PPVM_EINVAL, "null handle");
  if (n && !t) return fail(PPVM_EINVAL, "null target array");
  for (size_t i = 0; i < n; ++i)
    if (t[i] >= h->n) return fail(PPVM_EINVAL, "qubit index %u out of range for %u qubits", t[i], h->n);
  for (size_t i = 0; i < n; ++i) {
    int rc = push_gate(h, code, t[i], 0, theta, nullptr, 0);
    if (rc) return rc;
  }
  return PPVM_OK;
}
int gate_2q(ppvm_handle h, u32 code, const uint32_t* pr, size_t np, double theta) {
  if (!h) return fail(PPVM_EINVAL, "null handle");
  if (np && !pr) return fail(PPVM_EINVAL, "null pair array");
  for (size_t i = 0; i < np; ++i) {
    if (pr[2 * i] >= h->n || pr[2 * i + 1] >= h->n)
      return fail(PPVM_EINVAL, "qubit pair (%u, %u) out of range for %u qubits", pr[2 * i], pr[2 * i + 1], h->n);
    if (pr[2 * i] == pr[2 * i + 1]) return fail(PPVM_EINVAL, "two-qubit gate on identical qubits (%u)", pr[2 * i]);
  }
  for (size_t i = 0; i < np; ++i) {
    int rc = push_gate(h, code, pr[2 * i], pr[2 * i + 1], theta, nullptr, 0);
    if (rc) return rc;
  }
  return PPVM_OK;
}

// Host copies of the live terms (arena order).
int download_live(ppvm_sum* s, std::vector<u64>& keys, std::vector<double>& cs) {
  int rc;
  if ((rc = pump(s, true))) return rc;
  if ((rc = sync_check(s))) return rc;
  const CountSlot c = s->st_host->slot[s->op_seq & 1];
  const u64 n = c.n_entries;
  const u32 np = 2 * s->W;
  std::vector<u64> pl(n * np);  // term-major, as the arena holds them
  std::vector<double> cf(n);
  if (n) {
    HIPCHK(hipMemcpyAsync(pl.data(), s->keys[s->cur], n * np * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(cf.data(), s->coef[s->cur], n * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  }
  keys.clear();
  cs.clear();
  keys.reserve((n - c.n_dead) * np);
  cs.reserve(n - c.n_dead);
  for (u64 i = 0; i < n; ++i) {
    u64 bits;
    std::memcpy(&bits, &cf[i], 8);
    if ((bits & ABS_MASK) == DEAD_BITS) continue;
    keys.insert(keys.end(), pl.begin() + (size_t)i * np, pl.begin() + (size_t)(i + 1) * np);
    cs.push_back(cf[i]);
  }
  return PPVM_OK;
}

// Replace the arena contents with the given terms (term-major keys).
int upload(ppvm_sum* s, const std::vector<u64>& keys, const std::vector<double>& cs) {
  int rc;
  const u64 n = cs.size();
  const u32 np = 2 * s->W;
  if (n + 2 * n > s->cap) {
    if ((rc = resize(s, pow2_at_least(std::max<u64>(kMinCap, 4 * n)), 0))) return rc;
  }
  if (n) {  // term-major keys are the arena layout
    HIPCHK(hipMemcpyAsync(s->keys[s->cur], keys.data(), (size_t)n * np * 8, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->coef[s->cur], cs.data(), n * 8, hipMemcpyHostToDevice, s->stream));
  }
  CountSlot c{};
  c.n_entries = n;
  c.n_dead = 0;
  s->fresh = false;
  HIPCHK(hipMemcpyAsync(&s->st->slot[s->op_seq & 1], &c, sizeof(CountSlot), hipMemcpyHostToDevice, s->stream));
  if ((rc = invalidate_tp(s))) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  s->n_bound = n;
  s->n_synced = s->n_bound;
  sup_from_keys(s, keys, n);
  s->sup_stale = false;
  return PPVM_OK;
}

struct KeyHash {
  size_t W2;
  size_t operator()(const std::vector<u64>& k) const {
    u64 h = 0x9e3779b97f4a7c15ull;
    for (u64 w : k) h = (h ^ w) * 0xbf58476d1ce4e5b9ull + (h >> 31);
    return (size_t)h;
  }
};

int create_impl(uint32_t n_qubits, double tau, uint64_t maxw, uint64_t cap_hint, ppvm_sum** out) {
  if (!out) return fail(PPVM_EINVAL, "null output pointer");
  *out = nullptr;
  if (n_qubits == 0) return fail(PPVM_EINVAL, "n_qubits must be positive");
  u32 W = (n_qubits + 63) / 64;
  u32 Wp = W <= 1 ? 1 : W <= 2 ? 2 : W <= 4 ? 4 : W <= 8 ? 8 : 0;
  if (!Wp) return fail(PPVM_EINVAL, "n_qubits=%u exceeds the supported maximum of 512", n_qubits);
  if (!(tau >= 0.0)) return fail(PPVM_EINVAL, "min_abs_coeff must be a non-negative number");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return fail(PPVM_ENODEV, "no HIP device available");
  }
  ppvm_sum* s = new ppvm_sum();
  s->n = n_qubits;
  s->W = Wp;
  s->Wa = W;
  s->tau = tau;
  s->maxw = maxw;
  s->trunc_active = tau > 0.0 || maxw != ~0ull;
  s->partition_min = partition_min_entries();
  s->bucket_lg_max = bucket_lg_limit();
  s->pair_sort = pair_sort_env();
  s->bucket_target = bucket_target_env();
  s->small_max = small_max_entries(Wp);
  s->small_off = s->small_max == 0;
  s->small_threads = small_threads_env();
  s->xsup.assign(Wp, 0ull);
  s->zsup.assign(Wp, 0ull);
  s->clean = true;
  HIPCHK(hipGetDevice(&s->device));
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return fail(PPVM_EHIP, "hipStreamCreate failed");
  }
  if (hipMalloc(&s->st, sizeof(DevState)) != hipSuccess ||
      hipHostMalloc(&s->st_host, sizeof(DevState), hipHostMallocDefault) != hipSuccess ||
      hipMalloc(&s->partials, kOverlapGrid * sizeof(double)) != hipSuccess ||
      hipMalloc(&s->parts, 6 * (size_t)kMaxParts * sizeof(u64)) != hipSuccess ||
      hipMalloc(&s->pmeta, ((size_t)kMaxBuckets * kPairSlices + 2 * kMaxBuckets) * sizeof(u32)) != hipSuccess) {
    (void)hipGetLastError();
    delete s;
    return fail(PPVM_ENOMEM, "device allocation of handle state failed");
  }
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s->stream, reinterpret_cast<u64*>(s->st),
                     (u64)(sizeof(DevState) / 8), 0ull);
  std::memset(s->st_host, 0, sizeof(DevState));
  int rc = resize(s, pow2_at_least(std::max<u64>(kMinCap, cap_hint)), 0);
  if (rc) {
    free_buffers(s);
    delete s;
    return rc;
  }
  // the LDS path's step buffers come with the handle, so that its first run
  // does not pay for the allocation; without them the path stays off
  if (!s->small_off && small_buffers(s) != PPVM_OK) s->small_off = true;
  *out = s;
  return PPVM_OK;
}

}  // namespace

// ================================ C ABI ===========================================
extern "C" {

const char* ppvm_last_error(void) { return g_err.c_str(); }

int ppvm_device_info(char* buf, size_t cap) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return fail(PPVM_ENODEV, "no HIP device available");
  }
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipDeviceProp_t p;
  HIPCHK(hipGetDeviceProperties(&p, dev));
  if (buf && cap) snprintf(buf, cap, "%s | %s | CUs=%d | mem=%.1f GiB", p.name, p.gcnArchName,
                           p.multiProcessorCount, (double)p.totalGlobalMem / (1ull << 30));
  return PPVM_OK;
}

int ppvm_create(uint32_t n_qubits, double min_abs_coeff, uint64_t max_pauli_weight, uint64_t capacity_hint,
                ppvm_handle* out) {
  return create_impl(n_qubits, min_abs_coeff, max_pauli_weight, capacity_hint, out);
}

int ppvm_destroy(ppvm_handle h) {
  if (!h) return PPVM_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_buffers(h);
  if (h->st) (void)hipFree(h->st);
  if (h->st_host) (void)hipHostFree(h->st_host);
  if (h->partials) (void)hipFree(h->partials);
  if (h->parts) (void)hipFree(h->parts);
  if (h->pmeta) (void)hipFree(h->pmeta);
  if (h->preserve) (void)hipFree(h->preserve);
  if (h->pres_found) (void)hipFree(h->pres_found);
  if (h->d_small) (void)hipFree(h->d_small);
  if (h->h_small) (void)hipHostFree(h->h_small);
  for (auto e : h->ev) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return PPVM_OK;
}

int ppvm_words(ppvm_handle h, uint32_t* w) {
  if (!h || !w) return fail(PPVM_EINVAL, "null argument");
  *w = h->Wa;
  return PPVM_OK;
}

static int upload_preserve(ppvm_handle h) {
  if (h->preserve) (void)hipFree(h->preserve);
  if (h->pres_found) (void)hipFree(h->pres_found);
  h->preserve = nullptr;
  h->pres_found = nullptr;
  h->n_preserve = 0;
  const size_t cnt = h->preserve_host.size() / (2 * h->W);
  if (cnt) {
    if (cnt >= (1ull << 32)) return fail(PPVM_EINVAL, "preserve set too large");
    HIPCHK(hipMalloc(&h->preserve, h->preserve_host.size() * 8));
    HIPCHK(hipMemcpy(h->preserve, h->preserve_host.data(), h->preserve_host.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&h->pres_found, (cnt + 1) * 8));
    h->n_preserve = (u32)cnt;
  }
  return PPVM_OK;
}

int ppvm_set_preserve(ppvm_handle h, const uint64_t* x, const uint64_t* z, size_t count) {
  if (!h) return fail(PPVM_EINVAL, "null handle");
  if (count && (!x || !z)) return fail(PPVM_EINVAL, "null key array");
  int rc;
  // ops issued so far (and any replay of them) keep the old set
  if ((rc = pump(h, true))) return rc;
  if ((rc = sync_check(h))) return rc;
  const u32 W = h->W, Wa = h->Wa;
  h->preserve_host.assign((size_t)count * 2 * W, 0);
  for (size_t i = 0; i < count; ++i)
    for (u32 p = 0; p < Wa; ++p) {
      h->preserve_host[i * 2 * W + p] = x[i * Wa + p];
      h->preserve_host[i * 2 * W + W + p] = z[i * Wa + p];
    }
  return upload_preserve(h);
}

int ppvm_clone(ppvm_handle h, ppvm_handle* out) {
  if (!h || !out) return fail(PPVM_EINVAL, "null argument");
  int rc;
  if ((rc = pump(h, true))) return rc;
  if ((rc = sync_check(h))) return rc;
  ppvm_sum* c = nullptr;
  if ((rc = create_impl(h->n, h->tau, h->maxw, h->cap, &c))) return rc;
  const CountSlot cs = h->st_host->slot[h->op_seq & 1];
  const u32 np = 2 * h->W;
  if (cs.n_entries) {
    hipLaunchKernelGGL(k_copy_arena, dim3(grid_for(cs.n_entries, 256, 4096)), dim3(256), 0, h->stream,
                       h->keys[h->cur], h->coef[h->cur], c->keys[c->cur], c->coef[c->cur], np, cs.n_entries);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(&c->st->slot[c->op_seq & 1], &cs, sizeof(CountSlot), hipMemcpyHostToDevice));
  c->n_bound = cs.n_entries;
  c->n_synced = c->n_bound;
  c->partition_min = h->partition_min;
  c->xsup = h->xsup;
  c->zsup = h->zsup;
  c->sup_stale = h->sup_stale;
  c->bucket_lg_max = h->bucket_lg_max;
  c->pair_sort = h->pair_sort;
  c->bucket_target = h->bucket_target;
  c->clean = h->clean;
  c->preserve_host = h->preserve_host;
  if ((rc = upload_preserve(c))) {
    ppvm_destroy(c);
    return rc;
  }
  *out = c;
  return PPVM_OK;
}

int ppvm_add_terms(ppvm_handle h, const uint64_t* x, const uint64_t* z, const double* c, size_t count) {
  if (!h) return fail(PPVM_EINVAL, "null handle");
  if (count && (!x || !z || !c)) return fail(PPVM_EINVAL, "null term array");
  if (count == 0) return PPVM_OK;
  const u32 W = h->W, Wa = h->Wa, np = 2 * W;
  // bits beyond n_qubits must be clear
  const u32 nq = h->n;
  for (size_t i = 0; i < count; ++i)
    for (u32 p = 0; p < Wa; ++p) {
      u64 valid = (p * 64 + 64 <= nq) ? ~0ull : (p * 64 >= nq ? 0ull : ((1ull << (nq - p * 64)) - 1));
      if ((x[i * Wa + p] & ~valid) || (z[i * Wa + p] & ~valid))
        return fail(PPVM_EINVAL, "term %zu has bits set beyond n_qubits=%u", i, nq);
    }
  int rc;
  std::vector<u64> keys;
  std::vector<double> cs;
  if ((rc = download_live(h, keys, cs))) return rc;
  // entry(k).and_modify(+=).or_insert(c), in order (map/hashmap.rs:40-44)
  std::unordered_map<std::vector<u64>, size_t, KeyHash> idx(16, KeyHash{np});
  std::vector<u64> k(np);
  for (size_t i = 0; i < cs.size(); ++i) {
    std::copy(keys.begin() + i * np, keys.begin() + (i + 1) * np, k.begin());
    idx.emplace(k, i);
  }
  for (size_t i = 0; i < count; ++i) {
    std::fill(k.begin(), k.end(), 0ull);
    for (u32 p = 0; p < Wa; ++p) {
      k[p] = x[i * Wa + p];
      k[W + p] = z[i * Wa + p];
    }
    auto it = idx.find(k);
    if (it != idx.end()) {
      cs[it->second] += c[i];
    } else {
      idx.emplace(k, cs.size());
      keys.insert(keys.end(), k.begin(), k.end());
      cs.push_back(c[i]);
    }
  }
  if ((rc = upload(h, keys, cs))) return rc;
  h->clean = !h->trunc_active;
  return PPVM_OK;
}

int ppvm_clear(ppvm_handle h) {
  if (!h) return fail(PPVM_EINVAL, "null handle");
  if (h->broken) return fail(PPVM_ESTATE, "handle is unusable after an earlier device error");
  // queued and issued-but-unchecked ops act on terms that are being dropped:
  // discard them (an overflow they raised needs no replay either)
  h->pending.clear();
  h->skip_run = 0;
  h->skip_maps = 0;
  h->log.clear();
  h->issued_since_sync = 0;
  h->small_list.clear();
  h->small_maps = 0;
  h->small_on = false;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, h->stream, &h->st->overflow, 1ull, 0ull);
  HIPCHK(hipGetLastError());
  const std::vector<u64> keys;
  const std::vector<double> cs;
  int rc = upload(h, keys, cs);  // zero entries; synchronises the stream
  if (rc) {
    h->broken = true;
    return rc;
  }
  h->clean = true;
  return PPVM_OK;
}

int ppvm_x(ppvm_handle h, const uint32_t* t, size_t n) { return gate_1q(h, PPVM_OP_X, t, n, 0.0); }
int ppvm_y(ppvm_handle h, const uint32_t* t, size_t n) { return gate_1q(h, PPVM_OP_Y, t, n, 0.0); }
int ppvm_z(ppvm_handle h, const uint32_t* t, size_t n) { return gate_1q(h, PPVM_OP_Z, t, n, 0.0); }
int ppvm_h(ppvm_handle h, const uint32_t* t, size_t n) { return gate_1q(h, PPVM_OP_H, t, n, 0.0); }
int ppvm_s(ppvm_handle h, const uint32_t* t, size_t n) { return gate_1q(h, PPVM_OP_S, t, n, 0.0); }
int ppvm_s_dag(ppvm_handle h, const uint32_t* t, size_t n) { return gate_1q(h, PPVM_OP_S_DAG, t, n, 0.0); }
int ppvm_sqrt_x(ppvm_handle h, const uint32_t* t, size_t n) { return gate_1q(h, PPVM_OP_SQRT_X, t, n, 0.0); }
int ppvm_sqrt_x_dag(ppvm_handle h, const uint32_t* t, size_t n) { return gate_1q(h, PPVM_OP_SQRT_X_DAG, t, n, 0.0); }
int ppvm_sqrt_y(ppvm_handle h, const uint32_t* t, size_t n) { return gate_1q(h, PPVM_OP_SQRT_Y, t, n, 0.0); }
int ppvm_sqrt_y_dag(ppvm_handle h, const uint32_t* t, size_t n) { return gate_1q(h, PPVM_OP_SQRT_Y_DAG, t, n, 0.0); }
int ppvm_cnot(ppvm_handle h, const uint32_t* p, size_t n) { return gate_2q(h, PPVM_OP_CNOT, p, n, 0.0); }
int ppvm_cz(ppvm_handle h, const uint32_t* p, size_t n) { return gate_2q(h, PPVM_OP_CZ, p, n, 0.0); }
int ppvm_cy(ppvm_handle h, const uint32_t* p, size_t n) { return gate_2q(h, PPVM_OP_CY, p, n, 0.0); }
int ppvm_rx(ppvm_handle h, const uint32_t* t, size_t n, double th) { return gate_1q(h, PPVM_OP_RX, t, n, th); }
int ppvm_ry(ppvm_handle h, const uint32_t* t, size_t n, double th) { return gate_1q(h, PPVM_OP_RY, t, n, th); }
int ppvm_rz(ppvm_handle h, const uint32_t* t, size_t n, double th) { return gate_1q(h, PPVM_OP_RZ, t, n, th); }
int ppvm_rxx(ppvm_handle h, const uint32_t* p, size_t n, double th) { return gate_2q(h, PPVM_OP_RXX, p, n, th); }
int ppvm_ryy(ppvm_handle h, const uint32_t* p, size_t n, double th) { return gate_2q(h, PPVM_OP_RYY, p, n, th); }
int ppvm_rzz(ppvm_handle h, const uint32_t* p, size_t n, double th) { return gate_2q(h, PPVM_OP_RZZ, p, n, th); }

int ppvm_r(ppvm_handle h, uint32_t q, double axis_angle, double theta) {
  if (!h) return fail(PPVM_EINVAL, "null handle");
  int rc;
  if ((rc = push_gate(h, PPVM_OP_RZ, q, 0, axis_angle, nullptr, 0))) return rc;
  if ((rc = push_gate(h, PPVM_OP_RX, q, 0, theta, nullptr, 0))) return rc;
  return push_gate(h, PPVM_OP_RZ, q, 0, -axis_angle, nullptr, 0);
}

int ppvm_pauli_error(ppvm_handle h, uint32_t q, const double p[3]) {
  if (!h || !p) return fail(PPVM_EINVAL, "null argument");
  return push_gate(h, PPVM_OP_PAULI_ERROR, q, 0, 0.0, p, 3);
}
int ppvm_depolarize1(ppvm_handle h, uint32_t q, double p) {
  if (!h) return fail(PPVM_EINVAL, "null handle");
  return push_gate(h, PPVM_OP_DEPOLARIZE1, q, 0, 0.0, &p, 1);
}
int ppvm_depolarize2(ppvm_handle h, uint32_t a, uint32_t b, double p) {
  if (!h) return fail(PPVM_EINVAL, "null handle");
  return push_gate(h, PPVM_OP_DEPOLARIZE2, a, b, 0.0, &p, 1);
}
int ppvm_two_qubit_pauli_error(ppvm_handle h, uint32_t a, uint32_t b, const double p[15]) {
  if (!h || !p) return fail(PPVM_EINVAL, "null argument");
  return push_gate(h, PPVM_OP_PAULI_ERROR2, a, b, 0.0, p, 15);
}

int ppvm_amplitude_damping(ppvm_handle h, uint32_t q, double gamma) {
  if (!h) return fail(PPVM_EINVAL, "null handle");
  return push_gate(h, PPVM_OP_AMPLITUDE_DAMPING, q, 0, 0.0, &gamma, 1);
}

int ppvm_truncate(ppvm_handle h) {
  if (!h) return fail(PPVM_EINVAL, "null handle");
  ppvm_op op{};
  op.code = PPVM_OP_TRUNCATE;
  return enqueue(h, op);
}

int ppvm_apply(ppvm_handle h, const ppvm_op* ops, size_t n) {
  if (!h) return fail(PPVM_EINVAL, "null handle");
  if (n && !ops) return fail(PPVM_EINVAL, "null op array");
  for (size_t i = 0; i < n; ++i) {
    int rc = validate(h, ops[i]);
    if (rc) return rc;
  }
  for (size_t i = 0; i < n; ++i) {
    int rc = enqueue(h, ops[i]);
    if (rc) return rc;
  }
  return PPVM_OK;
}

int ppvm_synchronize(ppvm_handle h) {
  if (!h) return fail(PPVM_EINVAL, "null handle");
  int rc;
  if ((rc = pump(h, true))) return rc;
  return sync_check(h);
}

int ppvm_len(ppvm_handle h, uint64_t* out) {
  if (!h || !out) return fail(PPVM_EINVAL, "null argument");
  int rc;
  if ((rc = ppvm_synchronize(h))) return rc;
  const CountSlot& c = h->st_host->slot[h->op_seq & 1];
  *out = c.n_entries - c.n_dead;
  return PPVM_OK;
}

int ppvm_overlap_with_zero(ppvm_handle h, double* out) {
  if (!h || !out) return fail(PPVM_EINVAL, "null argument");
  int rc;
  if ((rc = ppvm_synchronize(h))) return rc;
  if ((rc = overlap_w(h))) return rc;
  if ((rc = read_state(h))) return rc;
  *out = h->st_host->overlap;
  return PPVM_OK;
}

int ppvm_trace(ppvm_handle h, const char* pattern, double* out) {
  if (!h || !pattern || !out) return fail(PPVM_EINVAL, "null argument");
  PatternProg prog;
  int rc;
  if ((rc = compile_pattern(pattern, h->n, prog))) return rc;
  if ((rc = ppvm_synchronize(h))) return rc;
  if ((rc = trace_w(h, prog))) return rc;
  if ((rc = read_state(h))) return rc;
  *out = h->st_host->overlap;
  return PPVM_OK;
}

int ppvm_overlap(ppvm_handle a, ppvm_handle b, double* out) {
  if (!a || !b || !out) return fail(PPVM_EINVAL, "null argument");
  if (a->n != b->n) return fail(PPVM_EINVAL, "overlap of sums over %u and %u qubits", a->n, b->n);
  if (a->device != b->device)
    return fail(PPVM_EINVAL, "overlap of sums on devices %d and %d", a->device, b->device);
  int rc;
  if ((rc = ppvm_synchronize(a))) return rc;
  if (b != a && (rc = ppvm_synchronize(b))) return rc;
  if ((rc = overlap2_w(a, b))) {
    a->broken = true;
    return rc;
  }
  if ((rc = read_state(a))) return rc;
  *out = a->st_host->overlap;
  return PPVM_OK;
}

int ppvm_export_terms(ppvm_handle h, uint64_t* x, uint64_t* z, double* c, size_t cap, size_t* n) {
  if (!h || !n) return fail(PPVM_EINVAL, "null argument");
  std::vector<u64> keys;
  std::vector<double> cs;
  int rc;
  if ((rc = download_live(h, keys, cs))) return rc;
  *n = cs.size();
  if (cap < cs.size()) return fail(PPVM_EINVAL, "export buffer holds %zu terms, %zu needed", cap, cs.size());
  if (cs.size() && (!x || !z || !c)) return fail(PPVM_EINVAL, "null export array");
  const u32 W = h->W, Wa = h->Wa, np = 2 * W;
  for (size_t i = 0; i < cs.size(); ++i) {
    for (u32 p = 0; p < Wa; ++p) {
      x[i * Wa + p] = keys[i * np + p];
      z[i * Wa + p] = keys[i * np + W + p];
    }
    c[i] = cs[i];
  }
  return PPVM_OK;
}

int ppvm_fingerprint(ppvm_handle h, uint64_t out[3]) {
  if (!h || !out) return fail(PPVM_EINVAL, "null argument");
  int rc;
  if ((rc = ppvm_synchronize(h))) return rc;
  const CountSlot c = h->st_host->slot[h->op_seq & 1];
  const u32 G = (u32)grid_for(c.n_entries, (u64)kBlock * 8, 1024);  // 2 G <= kMaxParts
  switch (h->W) {
    case 1: rc = launch_fingerprint<1>(h, G); break;
    case 2: rc = launch_fingerprint<2>(h, G); break;
    case 4: rc = launch_fingerprint<4>(h, G); break;
    default: rc = launch_fingerprint<8>(h, G); break;
  }
  if (rc) return rc;
  std::vector<u64> part((size_t)2 * G);
  HIPCHK(hipMemcpyAsync(part.data(), h->part(5), part.size() * 8, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  u64 ks = 0, kcs = 0;
  for (u32 b = 0; b < G; ++b) {
    ks += part[2 * (size_t)b];
    kcs += part[2 * (size_t)b + 1];
  }
  out[0] = c.n_entries - c.n_dead;
  out[1] = ks;
  out[2] = kcs;
  return PPVM_OK;
}

int ppvm_get_stats(ppvm_handle h, ppvm_stats* out) {
  if (!h || !out) return fail(PPVM_EINVAL, "null argument");
  int rc;
  if ((rc = ppvm_synchronize(h))) return rc;
  const DevState& d = *h->st_host;
  const CountSlot& c = d.slot[h->op_seq & 1];
  *out = h->hs;
  out->term_updates = d.term_updates;
  out->gates = d.gates;
  out->rotations = d.rotations;
  out->n_entries = c.n_entries;
  out->n_live = c.n_entries - c.n_dead;
  out->capacity = h->cap;
  return PPVM_OK;
}

int ppvm_get_counters(ppvm_handle h, uint64_t out[20]) {
  if (!h || !out) return fail(PPVM_EINVAL, "null argument");
  int rc;
  if ((rc = ppvm_synchronize(h))) return rc;
  const DevState& d = *h->st_host;
  const u64 v[20] = {d.term_updates, d.gates,          d.rotations,   d.scanned,
                     d.wl_total,     d.pairs_total,    d.children_total, d.kills_total,
                     d.map_entries,  d.map_launches,   d.scan_planes, h->cap,
                     h->hs.syncs,    h->hs.regrows,    h->hs.compactions, h->hs.peak_live,
                     d.transposed,   d.pair_fallback,  h->small_runs, h->small_steps};
  for (int i = 0; i < 20; ++i) out[i] = v[i];
  return PPVM_OK;
}

int ppvm_profile_enable(ppvm_handle h, int enable) {
  if (!h) return fail(PPVM_EINVAL, "null handle");
  int rc;
  if ((rc = ppvm_synchronize(h))) return rc;
  h->prof = enable != 0;
  for (int i = 0; i < 8; ++i) {
    h->prof_ms[i] = 0.0;
    h->prof_n[i] = 0;
  }
  return PPVM_OK;
}

int ppvm_profile_read(ppvm_handle h, double ms[8], uint64_t launches[8]) {
  if (!h || !ms || !launches) return fail(PPVM_EINVAL, "null argument");
  int rc;
  if ((rc = ppvm_synchronize(h))) return rc;
  for (int i = 0; i < 8; ++i) {
    ms[i] = h->prof_ms[i];
    launches[i] = h->prof_n[i];
  }
  return PPVM_OK;
}

}  // extern "C"
