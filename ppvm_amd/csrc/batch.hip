// batch.hip — batched propagation of many independent instances.
//
// One persistent kernel runs a whole set of instances (circuits, observables
// or parameter points): each workgroup takes the next instance from a queue
// (largest predicted cost first) and applies that instance's program gate by
// gate with workgroup barriers between phases, so no gate costs a kernel
// launch.  An instance's terms live in its own arena pair (ping-pong, SoA
// planes + coefficients); small arenas stay resident in L2 / the Infinity
// Cache while the workgroup works on them.
//
// Per rotation gate (k anticommuting with G, p = k xor G its partner):
//   phase 1  every anticommuting term inserts its pair key into the
//            instance's generation-tagged hash table; the second member of a
//            pair records the partnership for both;
//   phase 2  parents are rewritten in order into the other arena with
//            c_k' = c_k cos + eps_p c_p sin (or c_k cos without a partner),
//            dropping what the fused truncation rejects;
//   phase 3  children of unpaired anticommuting terms (coefficient
//            eps_k c_k sin) are appended in order.
// The output order is the insertion order of the reference's IndexMap
// (parents in place, new keys appended, truncation keeps order), so results
// are bit-identical to the single-instance engine and the CPU oracle.
//
// A launch processes ops [op_lo, op_hi) of every selected instance; instance
// state (current arena, size, progress) lives in device memory, so a program
// can be split into segments whose launches are captured in a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ppvm_hip.h"
#include "algebra.hpp"

using namespace ppvm;

namespace {

thread_local std::string b_err;
int bfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  b_err = buf;
  return code;
}
#define BCHK(expr)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) return bfail(PPVM_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

enum InstStatus : u32 { IS_READY = 0, IS_RUNNING = 1, IS_DONE = 2, IS_OVERFLOW = 3 };

// Device op: rotations carry sin/cos; map ops carry a factor table index.
struct BOp {
  u32 code;      // PPVM op code
  u32 flags;     // bit0 truncate after, bit1 record (len, overlap) after, bit2 skip (fused truncate)
  u32 a, b;
  double sn, cs;
  u32 table;     // noise factor table index
  u32 mask;      // noise factor apply mask
};

struct BInst {
  // program
  const BOp* ops;
  const double* factors;  // 16 doubles per noise table
  u32 n_ops;
  u32 n_qubits;
  // truncation
  double tau;
  u64 maxw;
  // storage
  u64 cap;
  u64* planes[2];
  double* coef[2];
  u32* partner;
  u64* table;
  u64 tcap;
  // marks output
  u64* mark_len;
  double* mark_ovl;
  u32 n_marks;
  // initial state
  u64 n_init;
  // state (device-updated)
  u32 status;
  u32 cur;
  u64 n;
  u32 op_next;
  u32 gen;
  u32 marks_done;
  u32 pad0;
  u64 term_updates;
  u64 peak;
  double ovl_last;
  u64 t_done;     // device wall clock when the instance finished (or stopped)
  // sup_valid: the instance's row of the support array (union of the live
  // terms' x words [0, W) and z words [W, 2W)) is current; clean: every term
  // already passed the truncation criteria
  u32 sup_valid;
  u32 clean;
};
constexpr int kSupWords = 16;  // support words per instance (2W, W <= 8)

constexpr int kBMaxThreads = 1024;

// Block-wide rank of a 0/1 flag over the block (thread order); returns the
// number of flagged threads.  lds needs kBMaxThreads/64 + 1 entries.
__device__ __forceinline__ u32 block_rank(bool f, u32& pos, u32* lds) {
  const u32 lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const u64 b = __ballot(f);
  __syncthreads();
  if (lane == 0) lds[wave] = (u32)__popcll(b);
  __syncthreads();
  u32 base = 0, tot = 0;
  for (u32 w = 0; w < nwave; ++w) {
    const u32 v = lds[w];
    if (w < wave) base += v;
    tot += v;
  }
  pos = base + (u32)__popcll(b & ((1ull << lane) - 1ull));
  return tot;
}

__device__ __forceinline__ double block_sum_d(double v, double* lds) {
  lds[threadIdx.x] = v;
  __syncthreads();
  for (u32 s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) lds[threadIdx.x] = __dadd_rn(lds[threadIdx.x], lds[threadIdx.x + s]);
    __syncthreads();
  }
  const double r = lds[0];
  __syncthreads();
  return r;
}

template <int W>
__device__ __forceinline__ bool bdrop(const u64 (&k)[2 * W], double c, const BInst& I) {
  bool drop = fabs(c) < I.tau;
  if (!drop && I.maxw != ~0ull) drop = weight_of<W>(k) > I.maxw;
  return drop;
}

template <int W>
__device__ __forceinline__ void bstore(u64* planes, double* coef, u64 cap, u64 i, const u64 (&k)[2 * W],
                                       double c) {
#pragma unroll
  for (int p = 0; p < 2 * W; ++p) planes[(u64)p * cap + i] = k[p];
  coef[i] = c;
}

// ---- support of an instance (in LDS while the block works on it) --------------
// s_sup[0, W): OR of the live terms' x words, s_sup[W, 2W): z words.
template <int W>
__device__ __forceinline__ void sup_or(u64 (&acc)[2 * W], const u64 (&k)[2 * W]) {
#pragma unroll
  for (int p = 0; p < 2 * W; ++p) acc[p] |= k[p];
}
// Block OR of every thread's acc into s_sup (lsup: kBMaxThreads/64 x 16 words).
template <int W>
__device__ void sup_publish(u64 (&acc)[2 * W], u64* s_sup, u64* lsup) {
  const u32 lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
#pragma unroll
  for (int p = 0; p < 2 * W; ++p)
    for (int o = 32; o > 0; o >>= 1) acc[p] |= (u64)__shfl_xor((long long)acc[p], o, 64);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int p = 0; p < 2 * W; ++p) lsup[wave * 16 + p] = acc[p];
  }
  __syncthreads();
  if (threadIdx.x < 2 * W) {
    u64 v = 0;
    for (u32 w = 0; w < nwave; ++w) v |= lsup[w * 16 + threadIdx.x];
    s_sup[threadIdx.x] = v;
  }
  __syncthreads();
}
template <int W>
__device__ void b_support(const BInst& I, u64* s_sup, u64* lsup) {
  u64 acc[2 * W];
#pragma unroll
  for (int p = 0; p < 2 * W; ++p) acc[p] = 0;
  const u64* P = I.planes[I.cur];
  for (u64 i = threadIdx.x; i < I.n; i += blockDim.x) {
    u64 k[2 * W];
    load_key<W>(P, I.cap, i, k);
    sup_or<W>(acc, k);
  }
  sup_publish<W>(acc, s_sup, lsup);
}
// Can some live term anticommute with the rotation's generator?  (It needs x
// where the generator has z, or z where it has x, on a generator qubit.)
__device__ __forceinline__ bool b_may_act(const u64* s_sup, u32 W, const BOp& op) {
  u32 ax = 0, az = 0, bx = 0, bz = 0;
  switch (op.code) {
    case PPVM_OP_RX: ax = 1; break;
    case PPVM_OP_RY: ax = az = 1; break;
    case PPVM_OP_RZ: az = 1; break;
    case PPVM_OP_RXX: ax = bx = 1; break;
    case PPVM_OP_RYY: ax = az = bx = bz = 1; break;
    case PPVM_OP_RZZ: az = bz = 1; break;
  }
  auto xb = [&](u32 q) { return (s_sup[q >> 6] >> (q & 63)) & 1ull; };
  auto zb = [&](u32 q) { return (s_sup[W + (q >> 6)] >> (q & 63)) & 1ull; };
  return (az && xb(op.a)) || (ax && zb(op.a)) || (bz && xb(op.b)) || (bx && zb(op.b));
}

// Applies one rotation: arena cur -> cur ^ 1.  Returns false on overflow.
template <int W>
__device__ bool b_rotate(BInst& I, const BOp& op, bool trunc, u32* lds, u64* s_sup, u64* lsup) {
  RotSpec g{};
  g.a = op.a;
  g.b = op.b;
  switch (op.code) {
    case PPVM_OP_RX: g.gax = 1; break;
    case PPVM_OP_RY: g.gax = 1; g.gaz = 1; break;
    case PPVM_OP_RZ: g.gaz = 1; break;
    case PPVM_OP_RXX: g.two = 1; g.gax = 1; g.gbx = 1; break;
    case PPVM_OP_RYY: g.two = 1; g.gax = g.gaz = g.gbx = g.gbz = 1; break;
    case PPVM_OP_RZZ: g.two = 1; g.gaz = 1; g.gbz = 1; break;
  }
  if (!g.two) g.b = g.a;
  g.sn = op.sn;
  g.cs = op.cs;
  const u32 T = blockDim.x;
  const u64 n = I.n, cap = I.cap;
  const u64* P = I.planes[I.cur];
  const double* C = I.coef[I.cur];
  u64* Q = I.planes[I.cur ^ 1];
  double* D = I.coef[I.cur ^ 1];
  u32* partner = I.partner;
  const u32 tgen = I.gen;
  u64 tc = 64;
  while (tc < 2 * n && tc < I.tcap) tc <<= 1;
  const u64 tmask = tc - 1;
  // phase 0: no partner yet
  for (u64 i = threadIdx.x; i < n; i += T) partner[i] = 0xffffffffu;
  __syncthreads();
  // phase 1: pair up anticommuting terms through the table
  for (u64 i = threadIdx.x; i < n; i += T) {
    u64 k[2 * W];
    load_key<W>(P, cap, i, k);
    if (!rot_anti<W>(k, g)) continue;
    u64 rep[2 * W];
    pair_key<W>(k, g, rep);
    const u64 h = hash_key<W>(rep);
    const u64 fp = h >> 56;
    const u64 mine = ((u64)tgen << 40) | (fp << 32) | i;
    u64 pos = h & tmask;
    u64 s = I.table[pos];
    while (true) {
      if ((u32)(s >> 40) != tgen) {
        const u64 old = cas_u64(&I.table[pos], s, mine);
        if (old == s) break;
        s = old;
        continue;
      }
      if (((s >> 32) & 0xffull) == fp) {
        const u64 j = s & 0xffffffffull;
        u64 kj[2 * W], repj[2 * W];
        load_key<W>(P, cap, j, kj);
        pair_key<W>(kj, g, repj);
        bool same = true;
#pragma unroll
        for (int p = 0; p < 2 * W; ++p) same &= (repj[p] == rep[p]);
        if (same) {
          partner[i] = (u32)j;
          partner[j] = (u32)i;
          break;
        }
      }
      pos = (pos + 1) & tmask;
      s = I.table[pos];
    }
  }
  __syncthreads();
  // phase 2: parents, in order (the kept keys also rebuild the support)
  u64 acc[2 * W];
#pragma unroll
  for (int p = 0; p < 2 * W; ++p) acc[p] = 0;
  u64 out = 0;
  for (u64 base = 0; base < n; base += T) {
    const u64 i = base + threadIdx.x;
    bool keep = false;
    u64 k[2 * W];
    double c = 0.0;
    if (i < n) {
      load_key<W>(P, cap, i, k);
      c = C[i];
      if (rot_anti<W>(k, g)) {
        const u32 p = partner[i];
        double nc = __dmul_rn(c, g.cs);
        if (p != 0xffffffffu) {
          u64 kp[2 * W];
          load_key<W>(P, cap, p, kp);
          const double ep = rot_eps<W>(kp, g);
          nc = __dadd_rn(nc, __dmul_rn(C[p], ep * g.sn));
        }
        c = nc;
      }
      keep = !trunc || !bdrop<W>(k, c, I);
    }
    u32 pos;
    const u32 tot = block_rank(keep, pos, lds);
    if (keep) {
      bstore<W>(Q, D, cap, out + pos, k, c);
      sup_or<W>(acc, k);
    }
    out += tot;
  }
  // phase 3: children of unpaired anticommuting terms, in order
  bool overflow = false;
  for (u64 base = 0; base < n; base += T) {
    const u64 i = base + threadIdx.x;
    bool keep = false;
    u64 k[2 * W];
    double cc = 0.0;
    if (i < n && partner[i] == 0xffffffffu) {
      load_key<W>(P, cap, i, k);
      if (rot_anti<W>(k, g)) {
        const double eps = rot_eps<W>(k, g);
        cc = __dmul_rn(C[i], eps * g.sn);
        rot_flip<W>(k, g);
        keep = !trunc || !bdrop<W>(k, cc, I);
      }
    }
    u32 pos;
    const u32 tot = block_rank(keep, pos, lds);
    if (out + tot > cap) {
      overflow = true;
      break;
    }
    if (keep) {
      bstore<W>(Q, D, cap, out + pos, k, cc);
      sup_or<W>(acc, k);
    }
    out += tot;
  }
  __syncthreads();
  if (overflow) return false;
  sup_publish<W>(acc, s_sup, lsup);
  I.n = out;
  I.cur ^= 1u;
  return true;
}

// Map op (Clifford / noise) and/or truncation: arena cur -> cur ^ 1.
template <int W>
__device__ void b_map(BInst& I, const BOp* op, bool trunc, u32* lds, u64* s_sup, u64* lsup) {
  const u32 T = blockDim.x;
  const u64 n = I.n, cap = I.cap;
  const u64* P = I.planes[I.cur];
  const double* C = I.coef[I.cur];
  u64* Q = I.planes[I.cur ^ 1];
  double* D = I.coef[I.cur ^ 1];
  u64 acc[2 * W];
#pragma unroll
  for (int p = 0; p < 2 * W; ++p) acc[p] = 0;
  u64 out = 0;
  for (u64 base = 0; base < n; base += T) {
    const u64 i = base + threadIdx.x;
    bool keep = false;
    u64 k[2 * W];
    double c = 0.0;
    if (i < n) {
      load_key<W>(P, cap, i, k);
      c = C[i];
      if (op) {
        const u32 code = op->code;
        if (code >= PPVM_OP_PAULI_ERROR && code <= PPVM_OP_PAULI_ERROR2) {
          scale_apply<W>(k, c, op->a, op->b, op->mask, I.factors + (u64)op->table * 16);
        } else if (clifford_apply<W>(k, code, op->a, op->b)) {
          c = -c;
        }
      }
      keep = !trunc || !bdrop<W>(k, c, I);
    }
    u32 pos;
    const u32 tot = block_rank(keep, pos, lds);
    if (keep) {
      bstore<W>(Q, D, cap, out + pos, k, c);
      sup_or<W>(acc, k);
    }
    out += tot;
  }
  __syncthreads();
  sup_publish<W>(acc, s_sup, lsup);
  I.n = out;
  I.cur ^= 1u;
}

template <int W>
__device__ double b_overlap(const BInst& I, double* ldsd) {
  double acc = 0.0;
  const u64* P = I.planes[I.cur];
  const double* C = I.coef[I.cur];
  for (u64 i = threadIdx.x; i < I.n; i += blockDim.x) {
    u64 xo = 0;
#pragma unroll
    for (int p = 0; p < W; ++p) xo |= P[(u64)p * I.cap + i];
    if (xo == 0) acc = __dadd_rn(acc, C[i]);
  }
  return block_sum_d(acc, ldsd);
}

// Table slots carry a 24-bit generation tag ((u64)gen << 40); one generation
// per rotation.  When an instance's generation reaches gen_limit (at most
// kBGenLimit) its table is cleared and the count restarts, so a stale slot can
// never carry the current tag.
constexpr u32 kBGenLimit = (1u << 24) - 2;

template <int W>
__global__ __launch_bounds__(kBMaxThreads) void k_batch(BInst* insts, u64* sups, const u32* order, u32 n_order,
                                                        u32* queue, u32 op_lo, u32 op_hi, u32 gen_limit) {
  __shared__ u32 lds[kBMaxThreads / 64 + 1];
  __shared__ double ldsd[kBMaxThreads];
  __shared__ u64 lsup[(kBMaxThreads / 64) * 16];
  __shared__ u64 s_sup[kSupWords];
  __shared__ u32 lds_i;
  while (true) {
    if (threadIdx.x == 0) lds_i = add_u32(queue, 1u);
    __syncthreads();
    const u32 qi = lds_i;
    __syncthreads();
    if (qi >= n_order) break;
    // Every thread keeps a private copy of the instance state; all updates are
    // block-uniform, and thread 0 writes the state back at the end.
    BInst* Ig = &insts[order[qi]];
    u64* sup_g = sups + (u64)order[qi] * kSupWords;
    BInst I = *Ig;
    if (I.status == IS_OVERFLOW || I.status == IS_DONE) continue;
    if (threadIdx.x < kSupWords) s_sup[threadIdx.x] = sup_g[threadIdx.x];
    __syncthreads();
    if (!I.sup_valid) {
      b_support<W>(I, s_sup, lsup);
      I.sup_valid = 1;
    }
    const bool trunc_active = I.tau > 0.0 || I.maxw != ~0ull;
    const u32 end = I.n_ops < op_hi ? I.n_ops : op_hi;
    u32 j = I.op_next > op_lo ? I.op_next : op_lo;
    bool ok = true;
    for (; j < end; ++j) {
      const BOp& op = I.ops[j];
      const bool trunc = (op.flags & 1u) != 0u;
      if (op.flags & 4u) {
        // truncation fused into the previous gate
      } else if (op.code >= PPVM_OP_RX && op.code <= PPVM_OP_RZZ) {
        I.term_updates += I.n;
        // a rotation that anticommutes with no live term changes nothing, and
        // the truncation fused with it has nothing to drop on a clean state
        if (b_may_act(s_sup, W, op) || (trunc && !I.clean)) {
          if (I.gen >= gen_limit) {
            for (u64 t = threadIdx.x; t < I.tcap; t += blockDim.x) I.table[t] = 0ull;
            __syncthreads();
            I.gen = 0u;
          }
          I.gen = I.gen + 1u;
          ok = b_rotate<W>(I, op, trunc, lds, s_sup, lsup);
          if (!ok) break;
          I.clean = (trunc || !trunc_active) ? 1u : 0u;
        }
      } else if (op.code == PPVM_OP_TRUNCATE) {
        b_map<W>(I, nullptr, true, lds, s_sup, lsup);
        I.clean = 1u;
      } else if (op.code != PPVM_OP_NOP) {
        I.term_updates += I.n;
        b_map<W>(I, &op, trunc, lds, s_sup, lsup);
        I.clean = (trunc || !trunc_active) ? 1u : 0u;
      }
      if (I.n > I.peak) I.peak = I.n;
      if (op.flags & 2u) {
        const double ov = b_overlap<W>(I, ldsd);
        if (threadIdx.x == 0 && I.marks_done < I.n_marks) {
          I.mark_len[I.marks_done] = I.n;
          I.mark_ovl[I.marks_done] = ov;
        }
        I.marks_done += 1;
        I.ovl_last = ov;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      Ig->op_next = j;
      Ig->status = !ok ? IS_OVERFLOW : (j >= I.n_ops ? IS_DONE : IS_RUNNING);
      Ig->n = I.n;
      Ig->cur = I.cur;
      Ig->gen = I.gen;
      Ig->marks_done = I.marks_done;
      Ig->term_updates = I.term_updates;
      Ig->peak = I.peak;
      Ig->ovl_last = I.ovl_last;
      Ig->t_done = wall_clock64();
      Ig->sup_valid = I.sup_valid;
      Ig->clean = I.clean;
    }
    if (threadIdx.x < kSupWords) sup_g[threadIdx.x] = s_sup[threadIdx.x];
    __syncthreads();
  }
}

// Resets instance state (the initial terms are already in arena 0) and stamps
// the start of the run.
__global__ void k_batch_reset(BInst* insts, u32 n, u64* t0) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *t0 = wall_clock64();
  if (i >= n) return;
  BInst& I = insts[i];
  I.status = IS_READY;
  I.cur = 0;
  I.n = I.n_init;
  I.op_next = 0;
  I.marks_done = 0;
  I.term_updates = 0;
  I.peak = I.n_init;
  I.ovl_last = 0.0;
  I.t_done = 0;
  I.sup_valid = 0;
  I.clean = 0;
}

__global__ void k_bfill_u32(u32* p, u32 n, u32 v) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

u64 pow2_ge(u64 v) {
  u64 p = 1;
  while (p < v) p <<= 1;
  return p;
}

}  // namespace

struct ppvm_batch {
  u32 W = 1;  // engine words per plane (max over instances, padded to 1/2/4/8)
  u32 n = 0;
  hipStream_t stream = nullptr;
  BInst* d_inst = nullptr;
  u64* d_sup = nullptr;           // kSupWords per instance (see BInst::sup_valid)
  std::vector<BInst> h_inst;      // host mirror (pointers)
  std::vector<void*> allocs;      // per-instance allocations
  BOp* d_ops = nullptr;
  double* d_factors = nullptr;
  u32* d_order = nullptr;
  u32* d_queue = nullptr;         // one counter per segment launch
  u64* d_t0 = nullptr;            // device wall clock at the start of a run
  u32 n_queue = 0;
  std::vector<u32> n_qubits;
  std::vector<u32> Wa;
  std::vector<u64> init_x, init_z;  // engine layout copies of the initial terms (2W words per term)
  std::vector<double> init_c;
  std::vector<u64> init_off;
  hipGraphExec_t graph = nullptr;
  u32 graph_key[4] = {0, 0, 0, 0};
};

namespace {

int alloc_instance(ppvm_batch* b, u32 i, u64 cap) {
  BInst& I = b->h_inst[i];
  const u32 np = 2 * b->W;
  I.cap = cap;
  I.tcap = pow2_ge(2 * cap);
  const size_t bytes = 2 * (size_t)np * cap * 8 + 2 * cap * 8 + ((cap * 4 + 255) & ~(size_t)255) + I.tcap * 8;
  char* base = nullptr;
  if (hipMalloc(&base, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return bfail(PPVM_ENOMEM, "batch instance %u: allocation of %zu bytes failed", i, bytes);
  }
  b->allocs[i] = base;
  char* q = base;
  I.planes[0] = (u64*)q; q += (size_t)np * cap * 8;
  I.planes[1] = (u64*)q; q += (size_t)np * cap * 8;
  I.coef[0] = (double*)q; q += cap * 8;
  I.coef[1] = (double*)q; q += cap * 8;
  I.partner = (u32*)q; q += (cap * 4 + 255) & ~(size_t)255;
  I.table = (u64*)q;
  BCHK(hipMemsetAsync(I.table, 0, I.tcap * 8, b->stream));
  return PPVM_OK;
}

int upload_initial(ppvm_batch* b, u32 i) {
  BInst& I = b->h_inst[i];
  const u32 np = 2 * b->W;
  const u64 n0 = I.n_init;
  std::vector<u64> pl((size_t)np * n0);
  for (u64 t = 0; t < n0; ++t)
    for (u32 p = 0; p < np; ++p) pl[(size_t)p * n0 + t] = (p < b->W ? b->init_x : b->init_z)[(b->init_off[i] + t) * b->W + (p % b->W)];
  for (u32 p = 0; p < np && n0; ++p)
    BCHK(hipMemcpyAsync(I.planes[0] + (size_t)p * I.cap, pl.data() + (size_t)p * n0, n0 * 8, hipMemcpyHostToDevice,
                        b->stream));
  if (n0) BCHK(hipMemcpyAsync(I.coef[0], b->init_c.data() + b->init_off[i], n0 * 8, hipMemcpyHostToDevice, b->stream));
  BCHK(hipStreamSynchronize(b->stream));
  return PPVM_OK;
}

int push_instances(ppvm_batch* b) {
  BCHK(hipMemcpyAsync(b->d_inst, b->h_inst.data(), b->n * sizeof(BInst), hipMemcpyHostToDevice, b->stream));
  BCHK(hipStreamSynchronize(b->stream));
  return PPVM_OK;
}
int pull_instances(ppvm_batch* b) {
  BCHK(hipMemcpyAsync(b->h_inst.data(), b->d_inst, b->n * sizeof(BInst), hipMemcpyDeviceToHost, b->stream));
  BCHK(hipStreamSynchronize(b->stream));
  return PPVM_OK;
}

// Generation limit of the instance tables: kBGenLimit, or a smaller value from
// PPVM_BATCH_GEN_LIMIT (at least 1; lets a test reach the table clear).
u32 batch_gen_limit() {
  const char* e = std::getenv("PPVM_BATCH_GEN_LIMIT");
  if (!e || !*e) return kBGenLimit;
  char* end = nullptr;
  const unsigned long v = std::strtoul(e, &end, 10);
  return (end && *end == '\0' && v >= 1 && v < kBGenLimit) ? (u32)v : kBGenLimit;
}

template <int W>
int launch_batch(ppvm_batch* b, u32 threads, u32 blocks, u32 n_order, u32 seg, u32 op_lo, u32 op_hi) {
  hipLaunchKernelGGL(k_batch<W>, dim3(blocks), dim3(threads), 0, b->stream, b->d_inst, b->d_sup, b->d_order, n_order,
                     b->d_queue + seg, op_lo, op_hi, batch_gen_limit());
  BCHK(hipGetLastError());
  return PPVM_OK;
}

int launch_w(ppvm_batch* b, u32 threads, u32 blocks, u32 n_order, u32 seg, u32 op_lo, u32 op_hi) {
  switch (b->W) {
    case 1: return launch_batch<1>(b, threads, blocks, n_order, seg, op_lo, op_hi);
    case 2: return launch_batch<2>(b, threads, blocks, n_order, seg, op_lo, op_hi);
    case 4: return launch_batch<4>(b, threads, blocks, n_order, seg, op_lo, op_hi);
    case 8: return launch_batch<8>(b, threads, blocks, n_order, seg, op_lo, op_hi);
  }
  return bfail(PPVM_EINVAL, "unsupported word count %u", b->W);
}

bool is_rot(u32 c) { return c >= PPVM_OP_RX && c <= PPVM_OP_RZZ; }
bool is_map(u32 c) {
  return (c >= PPVM_OP_X && c <= PPVM_OP_CY) || (c >= PPVM_OP_PAULI_ERROR && c <= PPVM_OP_PAULI_ERROR2);
}

// Noise factor table, in the reference's FP operation order
// (crates/ppvm-pauli-sum/src/sum/noise.rs:17-131); index 4*code_a + code_b with
// code = x | z << 1.
const int kRows[16][8] = {
    {-1, -1, -1, -1, -1, -1, -1, -1}, {1, 10, 13, 14, 2, 5, 6, 9},   {0, 10, 12, 14, 2, 4, 6, 8},
    {0, 1, 12, 13, 4, 5, 8, 9},       {10, 11, 12, 13, 14, 7, 8, 9}, {1, 11, 12, 2, 5, 6, 7, 8},
    {0, 11, 13, 2, 4, 6, 7, 9},       {0, 1, 10, 11, 14, 4, 5, 7},   {11, 12, 13, 14, 3, 4, 5, 6},
    {1, 10, 11, 12, 2, 3, 4, 9},      {0, 10, 11, 13, 2, 3, 5, 8},   {0, 1, 11, 14, 3, 6, 8, 9},
    {10, 3, 4, 5, 6, 7, 8, 9},        {1, 13, 14, 2, 3, 4, 7, 8},    {0, 12, 14, 2, 3, 5, 7, 9},
    {0, 1, 10, 12, 13, 3, 6, 7},
};
const int kLetter[4] = {0, 1, 3, 2};

void noise_table(const ppvm_op& op, double f[16], u32& mask, u32& b_out) {
  for (int i = 0; i < 16; ++i) f[i] = 1.0;
  mask = 0;
  b_out = op.q1;
  if (op.code == PPVM_OP_PAULI_ERROR) {
    const double one = 1.0;
    f[4] = one - op.p[1] * 2.0 - op.p[2] * 2.0;
    f[8] = one - op.p[0] * 2.0 - op.p[1] * 2.0;
    f[12] = one - op.p[0] * 2.0 - op.p[2] * 2.0;
    mask = (1u << 4) | (1u << 8) | (1u << 12);
    b_out = 0xffffffffu;
  } else if (op.code == PPVM_OP_DEPOLARIZE1) {
    const double fac = 1.0 - op.p[0] * (4.0 / 3.0);
    for (int ca = 1; ca < 4; ++ca) {
      f[ca * 4] = fac;
      mask |= 1u << (ca * 4);
    }
    b_out = 0xffffffffu;
  } else if (op.code == PPVM_OP_DEPOLARIZE2) {
    const double fac = 1.0 - op.p[0] * (16.0 / 15.0);
    for (int i = 1; i < 16; ++i) {
      f[i] = fac;
      mask |= 1u << i;
    }
  } else {
    for (int ca = 0; ca < 4; ++ca)
      for (int cb = 0; cb < 4; ++cb) {
        const int row = 4 * kLetter[ca] + kLetter[cb];
        if (row == 0) continue;
        double acc = 1.0;
        for (int j = 0; j < 8; ++j) acc = acc - op.p[kRows[row][j]] * 2.0;
        f[ca * 4 + cb] = acc;
        mask |= 1u << (ca * 4 + cb);
      }
  }
}

u32 engine_words(u32 w) { return w <= 1 ? 1 : w <= 2 ? 2 : w <= 4 ? 4 : 8; }

int validate_instance(const ppvm_batch_instance& s, size_t i) {
  if (s.n_qubits == 0 || s.n_qubits > 512) return bfail(PPVM_EINVAL, "instance %zu: n_qubits %u out of range", i, s.n_qubits);
  if (s.n_terms && (!s.x || !s.z || !s.c)) return bfail(PPVM_EINVAL, "instance %zu: null term arrays", i);
  if (s.n_ops && !s.ops) return bfail(PPVM_EINVAL, "instance %zu: null program", i);
  if (!(s.min_abs_coeff >= 0.0)) return bfail(PPVM_EINVAL, "instance %zu: bad min_abs_coeff", i);
  for (u32 j = 0; j < s.n_ops; ++j) {
    const ppvm_op& op = s.ops[j];
    const bool two = op.code == PPVM_OP_CNOT || op.code == PPVM_OP_CZ || op.code == PPVM_OP_CY ||
                     op.code == PPVM_OP_RXX || op.code == PPVM_OP_RYY || op.code == PPVM_OP_RZZ ||
                     op.code == PPVM_OP_DEPOLARIZE2 || op.code == PPVM_OP_PAULI_ERROR2;
    if (op.code == PPVM_OP_NOP || op.code == PPVM_OP_TRUNCATE) continue;
    if (op.code == PPVM_OP_AMPLITUDE_DAMPING)
      return bfail(PPVM_EINVAL, "instance %zu op %u: amplitude damping is not supported in batched mode", i, j);
    if (!is_rot(op.code) && !is_map(op.code)) return bfail(PPVM_EINVAL, "instance %zu op %u: unknown code %u", i, j, op.code);
    if (op.q0 >= s.n_qubits || (two && (op.q1 >= s.n_qubits || op.q1 == op.q0)))
      return bfail(PPVM_EINVAL, "instance %zu op %u: qubit out of range", i, j);
  }
  return PPVM_OK;
}

// Appends the instance's device ops (and noise factor tables, 16 doubles
// each) to ops / facs; returns the number of marked ops.
u32 translate_program(const ppvm_batch_instance& s, std::vector<BOp>& ops, std::vector<double>& facs) {
  u32 ntab = 0, marks = 0;
  bool prev_gate = false;
  for (u32 j = 0; j < s.n_ops; ++j) {
    const ppvm_op& op = s.ops[j];
    BOp d{};
    d.code = op.code;
    d.a = op.q0;
    d.b = op.q1;
    if (is_rot(op.code)) {
      sincos(op.theta, &d.sn, &d.cs);  // as the single-sum engine (one libm sincos)
    } else if (op.code >= PPVM_OP_PAULI_ERROR && op.code <= PPVM_OP_PAULI_ERROR2) {
      double f[16];
      u32 mask, bq;
      noise_table(op, f, mask, bq);
      d.table = ntab++;
      d.mask = mask;
      d.b = bq;
      facs.insert(facs.end(), f, f + 16);
    }
    const bool next_trunc = j + 1 < s.n_ops && s.ops[j + 1].code == PPVM_OP_TRUNCATE;
    const bool is_gate = is_rot(op.code) || is_map(op.code);
    const bool marked = (op.reserved & 1u) != 0u;
    // a truncate right after a gate runs inside that gate, unless the state
    // between the two is recorded
    const bool fuse = is_gate && next_trunc && !marked;
    if (fuse) d.flags |= 1u;
    if (op.code == PPVM_OP_TRUNCATE && prev_gate) d.flags |= 4u;  // fused into the gate before
    if (marked) {
      d.flags |= 2u;
      ++marks;
    }
    prev_gate = fuse || (op.code == PPVM_OP_TRUNCATE && prev_gate && !marked);
    ops.push_back(d);
  }
  return marks;
}

// Appends the instance's initial terms to kx / kz / kc in the engine layout
// (W words per plane), merging repeated words with `+=`; returns the count.
u64 dedupe_initial(const ppvm_batch_instance& s, u32 W, std::vector<u64>& kx, std::vector<u64>& kz,
                   std::vector<double>& kc) {
  const u32 wa = (s.n_qubits + 63) / 64;
  const size_t first = kc.size();
  for (u32 t = 0; t < s.n_terms; ++t) {
    std::vector<u64> x(W, 0), z(W, 0);
    for (u32 p = 0; p < wa; ++p) {
      x[p] = s.x[(size_t)t * wa + p];
      z[p] = s.z[(size_t)t * wa + p];
    }
    bool found = false;
    for (size_t e = first; e < kc.size(); ++e) {
      if (std::equal(x.begin(), x.end(), kx.begin() + e * W) && std::equal(z.begin(), z.end(), kz.begin() + e * W)) {
        kc[e] += s.c[t];
        found = true;
        break;
      }
    }
    if (!found) {
      kx.insert(kx.end(), x.begin(), x.end());
      kz.insert(kz.end(), z.begin(), z.end());
      kc.push_back(s.c[t]);
    }
  }
  return kc.size() - first;
}

// ---- allocation-free single-instance runs ----------------------------------------
// Every device buffer of ppvm_batch_run1_noalloc (instance descriptor, program,
// noise tables, arenas, partner list, hash table, queue) lives in this array:
// kScratchBytes per GPU, held from the library's load until the process ends.
// In host code `scratch` names a host-side stand-in; its device address comes
// from hipGetSymbolAddress, once per device (scratch_on_device).
constexpr size_t kScratchBytes = 64ull << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kScratchDevices] = {};

// Device address of `scratch` on the current device (resolved on the first
// call per device, with that device current).
int scratch_on_device(unsigned char** out) {
  int dev = 0;
  BCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices) return bfail(PPVM_ENODEV, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return bfail(PPVM_EHIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return PPVM_OK;
}

__global__ void k_bfill_u64(u64* p, u64 n, u64 v) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) p[i] = v;
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

const char* ppvm_batch_last_error(void) { return b_err.c_str(); }

int ppvm_batch_create(const ppvm_batch_instance* inst, size_t n, ppvm_batch_handle* out) {
  if (!out || (n && !inst)) return bfail(PPVM_EINVAL, "null argument");
  *out = nullptr;
  if (n == 0 || n > (1u << 20)) return bfail(PPVM_EINVAL, "batch size %zu out of range", n);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return bfail(PPVM_ENODEV, "no HIP device available");
  }
  u32 wmax = 1;
  for (size_t i = 0; i < n; ++i) {
    int rc = validate_instance(inst[i], i);
    if (rc) return rc;
    wmax = std::max(wmax, (inst[i].n_qubits + 63) / 64);
  }
  ppvm_batch* b = new ppvm_batch();
  b->W = engine_words(wmax);
  b->n = (u32)n;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return bfail(PPVM_EHIP, "hipStreamCreate failed");
  }
  b->h_inst.assign(n, BInst{});
  b->allocs.assign(n, nullptr);
  // programs: translate ops into device ops (sin/cos from the host libm, noise tables,
  // fused truncation flags)
  std::vector<BOp> ops;
  std::vector<double> facs;
  std::vector<u64> op_off(n);
  std::vector<u64> fac_off(n);
  for (size_t i = 0; i < n; ++i) {
    const ppvm_batch_instance& s = inst[i];
    op_off[i] = ops.size();
    fac_off[i] = facs.size() / 16;
    BInst& I = b->h_inst[i];
    I.n_marks = translate_program(s, ops, facs);
    I.n_ops = s.n_ops;
    I.n_qubits = s.n_qubits;
    I.tau = s.min_abs_coeff;
    I.maxw = s.max_pauli_weight;
  }
  // initial terms: deduplicate with `+=` semantics, engine word layout
  b->init_off.resize(n);
  b->n_qubits.resize(n);
  for (size_t i = 0; i < n; ++i) {
    b->init_off[i] = b->init_c.size();
    b->h_inst[i].n_init = dedupe_initial(inst[i], b->W, b->init_x, b->init_z, b->init_c);
  }
  int rc = PPVM_OK;
  if (hipMalloc(&b->d_inst, n * sizeof(BInst)) != hipSuccess ||
      hipMalloc(&b->d_sup, (size_t)n * kSupWords * sizeof(u64)) != hipSuccess ||
      hipMalloc(&b->d_ops, std::max<size_t>(1, ops.size()) * sizeof(BOp)) != hipSuccess ||
      hipMalloc(&b->d_factors, std::max<size_t>(1, facs.size()) * sizeof(double)) != hipSuccess ||
      hipMalloc(&b->d_order, n * sizeof(u32)) != hipSuccess ||
      hipMalloc(&b->d_t0, sizeof(u64)) != hipSuccess) {
    (void)hipGetLastError();
    ppvm_batch_destroy(b);
    return bfail(PPVM_ENOMEM, "batch descriptor allocation failed");
  }
  if (!ops.empty()) BCHK(hipMemcpy(b->d_ops, ops.data(), ops.size() * sizeof(BOp), hipMemcpyHostToDevice));
  if (!facs.empty()) BCHK(hipMemcpy(b->d_factors, facs.data(), facs.size() * sizeof(double), hipMemcpyHostToDevice));
  for (size_t i = 0; i < n; ++i) {
    BInst& I = b->h_inst[i];
    I.ops = b->d_ops + op_off[i];
    I.factors = b->d_factors + fac_off[i] * 16;
    const ppvm_batch_instance& s = inst[i];
    u64 cap = std::max<u64>(s.capacity, std::max<u64>(1024, 2 * I.n_init));
    cap = pow2_ge(cap);
    if ((rc = alloc_instance(b, (u32)i, cap))) {
      ppvm_batch_destroy(b);
      return rc;
    }
    if (I.n_marks) {
      if (hipMalloc(&I.mark_len, I.n_marks * 8) != hipSuccess || hipMalloc(&I.mark_ovl, I.n_marks * 8) != hipSuccess) {
        (void)hipGetLastError();
        ppvm_batch_destroy(b);
        return bfail(PPVM_ENOMEM, "mark buffer allocation failed");
      }
    }
    if ((rc = upload_initial(b, (u32)i))) {
      ppvm_batch_destroy(b);
      return rc;
    }
  }
  if ((rc = push_instances(b))) {
    ppvm_batch_destroy(b);
    return rc;
  }
  *out = b;
  return PPVM_OK;
}

int ppvm_batch_destroy(ppvm_batch_handle b) {
  if (!b) return PPVM_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  for (size_t i = 0; i < b->allocs.size(); ++i)
    if (b->allocs[i]) (void)hipFree(b->allocs[i]);
  for (auto& I : b->h_inst) {
    if (I.mark_len) (void)hipFree(I.mark_len);
    if (I.mark_ovl) (void)hipFree(I.mark_ovl);
  }
  if (b->d_inst) (void)hipFree(b->d_inst);
  if (b->d_sup) (void)hipFree(b->d_sup);
  if (b->d_ops) (void)hipFree(b->d_ops);
  if (b->d_factors) (void)hipFree(b->d_factors);
  if (b->d_order) (void)hipFree(b->d_order);
  if (b->d_queue) (void)hipFree(b->d_queue);
  if (b->d_t0) (void)hipFree(b->d_t0);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return PPVM_OK;
}

int ppvm_batch_set_capacity(ppvm_batch_handle b, size_t i, uint64_t capacity) {
  if (!b || i >= b->n) return bfail(PPVM_EINVAL, "bad batch handle or index");
  if (capacity > (1ull << 31)) return bfail(PPVM_EINVAL, "capacity too large");
  BCHK(hipStreamSynchronize(b->stream));
  if (b->allocs[i]) (void)hipFree(b->allocs[i]);
  b->allocs[i] = nullptr;
  int rc;
  if ((rc = alloc_instance(b, (u32)i, pow2_ge(std::max<u64>(capacity, 1024))))) return rc;
  if ((rc = upload_initial(b, (u32)i))) return rc;
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  return push_instances(b);
}

int ppvm_batch_run(ppvm_batch_handle b, const ppvm_batch_config* cfg, const uint32_t* order, size_t n_order,
                   double* ms) {
  if (!b || !cfg) return bfail(PPVM_EINVAL, "null argument");
  const u32 threads = cfg->threads ? cfg->threads : 256;
  if (threads % 64 || threads > (u32)kBMaxThreads) return bfail(PPVM_EINVAL, "threads must be a multiple of 64 up to 1024");
  std::vector<u32> ord;
  if (order) {
    for (size_t i = 0; i < n_order; ++i) {
      if (order[i] >= b->n) return bfail(PPVM_EINVAL, "order index %u out of range", order[i]);
      ord.push_back(order[i]);
    }
  } else {
    for (u32 i = 0; i < b->n; ++i) ord.push_back(i);
  }
  if (ord.empty()) return PPVM_OK;
  int dev = 0, cus = 256;
  BCHK(hipGetDevice(&dev));
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const u32 blocks = cfg->blocks ? cfg->blocks : (u32)std::min<size_t>(ord.size(), (size_t)cus * 2);
  u32 max_ops = 0;
  for (u32 i : ord) max_ops = std::max(max_ops, b->h_inst[i].n_ops);
  const u32 seg = cfg->segment_ops ? cfg->segment_ops : std::max<u32>(1, max_ops);
  const u32 n_seg = std::max<u32>(1, (max_ops + seg - 1) / seg);
  if (b->n_queue < n_seg) {
    if (b->d_queue) (void)hipFree(b->d_queue);
    b->d_queue = nullptr;
    BCHK(hipMalloc(&b->d_queue, n_seg * sizeof(u32)));
    b->n_queue = n_seg;
    if (b->graph) {
      (void)hipGraphExecDestroy(b->graph);
      b->graph = nullptr;
    }
  }
  BCHK(hipMemcpyAsync(b->d_order, ord.data(), ord.size() * sizeof(u32), hipMemcpyHostToDevice, b->stream));
  hipEvent_t e0, e1;
  BCHK(hipEventCreate(&e0));
  BCHK(hipEventCreate(&e1));
  auto record = [&](hipStream_t st) -> int {
    hipLaunchKernelGGL(k_batch_reset, dim3((b->n + 255) / 256), dim3(256), 0, st, b->d_inst, b->n, b->d_t0);
    hipLaunchKernelGGL(k_bfill_u32, dim3((n_seg + 255) / 256), dim3(256), 0, st, b->d_queue, n_seg, 0u);
    for (u32 s = 0; s < n_seg; ++s) {
      int rc = launch_w(b, threads, blocks, (u32)ord.size(), s, s * seg, (s + 1) * seg);
      if (rc) return rc;
    }
    return PPVM_OK;
  };
  // initial arenas must hold the initial terms again (a previous run rewrote them)
  for (u32 i : ord) {
    int rc = upload_initial(b, i);
    if (rc) return rc;
  }
  int rc = PPVM_OK;
  BCHK(hipEventRecord(e0, b->stream));
  const u32 key[4] = {threads, blocks, seg, (u32)ord.size()};
  if (cfg->use_graph) {
    if (b->graph && std::memcmp(key, b->graph_key, sizeof(key)) != 0) {
      (void)hipGraphExecDestroy(b->graph);
      b->graph = nullptr;
    }
    if (!b->graph) {
      hipGraph_t g;
      BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      rc = record(b->stream);
      hipError_t ce = hipStreamEndCapture(b->stream, &g);
      if (rc) return rc;
      if (ce != hipSuccess) return bfail(PPVM_EHIP, "graph capture failed: %s", hipGetErrorString(ce));
      BCHK(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
      (void)hipGraphDestroy(g);
      std::memcpy(b->graph_key, key, sizeof(key));
    }
    BCHK(hipGraphLaunch(b->graph, b->stream));
  } else {
    if ((rc = record(b->stream))) return rc;
  }
  BCHK(hipEventRecord(e1, b->stream));
  BCHK(hipEventSynchronize(e1));
  float t = 0.f;
  BCHK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = t;
  return pull_instances(b);
}

int ppvm_batch_results(ppvm_batch_handle b, uint32_t* status, uint64_t* n_final, double* overlap_last,
                       uint64_t* term_updates, uint64_t* peak) {
  if (!b) return bfail(PPVM_EINVAL, "null handle");
  for (u32 i = 0; i < b->n; ++i) {
    const BInst& I = b->h_inst[i];
    if (status) status[i] = I.status == IS_DONE ? 0u : (I.status == IS_OVERFLOW ? 1u : 2u);
    if (n_final) n_final[i] = I.n;
    if (overlap_last) overlap_last[i] = I.ovl_last;
    if (term_updates) term_updates[i] = I.term_updates;
    if (peak) peak[i] = I.peak;
  }
  return PPVM_OK;
}

int ppvm_batch_latency(ppvm_batch_handle b, double* seconds) {
  if (!b || !seconds) return bfail(PPVM_EINVAL, "null argument");
  u64 t0 = 0;
  BCHK(hipMemcpy(&t0, b->d_t0, sizeof(u64), hipMemcpyDeviceToHost));
  int dev = 0, khz = 100000;
  BCHK(hipGetDevice(&dev));
  (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
  for (u32 i = 0; i < b->n; ++i) {
    const BInst& I = b->h_inst[i];
    seconds[i] = I.t_done > t0 ? (double)(I.t_done - t0) / ((double)khz * 1e3) : 0.0;
  }
  return PPVM_OK;
}

int ppvm_batch_marks(ppvm_batch_handle b, size_t i, uint64_t* lens, double* overlaps, size_t cap, size_t* n) {
  if (!b || i >= b->n || !n) return bfail(PPVM_EINVAL, "bad argument");
  const BInst& I = b->h_inst[i];
  *n = std::min<size_t>(I.marks_done, I.n_marks);
  if (cap < *n) return bfail(PPVM_EINVAL, "mark buffer holds %zu, %zu needed", cap, *n);
  if (*n) {
    BCHK(hipMemcpy(lens, I.mark_len, *n * 8, hipMemcpyDeviceToHost));
    BCHK(hipMemcpy(overlaps, I.mark_ovl, *n * 8, hipMemcpyDeviceToHost));
  }
  return PPVM_OK;
}

int ppvm_batch_export(ppvm_batch_handle b, size_t i, uint64_t* x, uint64_t* z, double* c, size_t cap, size_t* n) {
  if (!b || i >= b->n || !n) return bfail(PPVM_EINVAL, "bad argument");
  const BInst& I = b->h_inst[i];
  *n = I.n;
  if (cap < I.n) return bfail(PPVM_EINVAL, "export buffer holds %zu, %llu needed", cap, (unsigned long long)I.n);
  const u32 W = b->W, wa = (I.n_qubits + 63) / 64, np = 2 * W;
  std::vector<u64> pl((size_t)np * I.n);
  for (u32 p = 0; p < np && I.n; ++p)
    BCHK(hipMemcpy(pl.data() + (size_t)p * I.n, I.planes[I.cur] + (size_t)p * I.cap, I.n * 8, hipMemcpyDeviceToHost));
  if (I.n) BCHK(hipMemcpy(c, I.coef[I.cur], I.n * 8, hipMemcpyDeviceToHost));
  for (u64 t = 0; t < I.n; ++t)
    for (u32 p = 0; p < wa; ++p) {
      x[t * wa + p] = pl[(size_t)p * I.n + t];
      z[t * wa + p] = pl[(size_t)(W + p) * I.n + t];
    }
  return PPVM_OK;
}

int ppvm_batch_run1_noalloc(const ppvm_batch_instance* inst, uint32_t threads, uint64_t* x, uint64_t* z,
                            double* c, size_t cap, size_t* n, double* overlap_last, uint64_t* term_updates) {
  if (!inst || !n) return bfail(PPVM_EINVAL, "null argument");
  int rc;
  if ((rc = validate_instance(*inst, 0))) return rc;
  if (threads == 0) threads = 256;
  if (threads % 64 || threads > (u32)kBMaxThreads) return bfail(PPVM_EINVAL, "threads must be a multiple of 64 up to 1024");
  unsigned char* base = nullptr;
  if ((rc = scratch_on_device(&base))) return rc;
  // host-side preparation (no device memory involved)
  const u32 W = engine_words((inst->n_qubits + 63) / 64), np = 2 * W;
  std::vector<BOp> ops;
  std::vector<double> facs;
  (void)translate_program(*inst, ops, facs);
  std::vector<u64> kx, kz;
  std::vector<double> kc;
  const u64 n0 = dedupe_initial(*inst, W, kx, kz, kc);
  const u64 acap = pow2_ge(std::max<u64>(inst->capacity, std::max<u64>(1024, 2 * n0)));
  const u64 tcap = pow2_ge(2 * acap);
  // scratch layout
  size_t off = 0;
  const size_t o_inst = off; off += align256(sizeof(BInst));
  const size_t o_sup = off; off += align256(kSupWords * sizeof(u64));
  const size_t o_ctl = off; off += 256;  // order (u32), queue (u32), t0 (u64)
  const size_t o_ops = off; off += align256(std::max<size_t>(1, ops.size()) * sizeof(BOp));
  const size_t o_fac = off; off += align256(std::max<size_t>(1, facs.size()) * sizeof(double));
  const size_t o_pl = off; off += 2 * (size_t)np * acap * 8;
  const size_t o_cf = off; off += 2 * (size_t)acap * 8;
  const size_t o_pa = off; off += align256(acap * 4);
  const size_t o_tb = off; off += tcap * 8;
  if (off > kScratchBytes)
    return bfail(PPVM_EINVAL, "run needs %zu bytes of scratch (arena of %llu terms, %zu ops); the maximum is %zu",
                 off, (unsigned long long)acap, ops.size(), kScratchBytes);
  BInst I{};
  I.ops = reinterpret_cast<const BOp*>(base + o_ops);
  I.factors = reinterpret_cast<const double*>(base + o_fac);
  I.n_ops = inst->n_ops;
  I.n_qubits = inst->n_qubits;
  I.tau = inst->min_abs_coeff;
  I.maxw = inst->max_pauli_weight;
  I.cap = acap;
  I.planes[0] = reinterpret_cast<u64*>(base + o_pl);
  I.planes[1] = reinterpret_cast<u64*>(base + o_pl + (size_t)np * acap * 8);
  I.coef[0] = reinterpret_cast<double*>(base + o_cf);
  I.coef[1] = reinterpret_cast<double*>(base + o_cf + acap * 8);
  I.partner = reinterpret_cast<u32*>(base + o_pa);
  I.table = reinterpret_cast<u64*>(base + o_tb);
  I.tcap = tcap;
  I.n_marks = 0;  // marked ops still update the last overlap
  I.n_init = n0;
  u32* order = reinterpret_cast<u32*>(base + o_ctl);
  u32* queue = order + 1;
  u64* t0 = reinterpret_cast<u64*>(base + o_ctl + 8);
  // uploads (the scratch array keeps the previous call's contents: everything
  // read before it is written is set here)
  std::vector<u64> pl((size_t)np * n0);
  for (u64 t = 0; t < n0; ++t)
    for (u32 p = 0; p < np; ++p) pl[(size_t)p * n0 + t] = (p < W ? kx : kz)[t * W + (p % W)];
  const hipStream_t st = nullptr;  // the device's null stream
  const u32 zero = 0;
  BCHK(hipMemcpyAsync(base + o_inst, &I, sizeof(BInst), hipMemcpyHostToDevice, st));
  BCHK(hipMemcpyAsync(order, &zero, sizeof(u32), hipMemcpyHostToDevice, st));
  if (!ops.empty()) BCHK(hipMemcpyAsync(base + o_ops, ops.data(), ops.size() * sizeof(BOp), hipMemcpyHostToDevice, st));
  if (!facs.empty()) BCHK(hipMemcpyAsync(base + o_fac, facs.data(), facs.size() * sizeof(double), hipMemcpyHostToDevice, st));
  for (u32 p = 0; p < np && n0; ++p)
    BCHK(hipMemcpyAsync(I.planes[0] + (size_t)p * acap, pl.data() + (size_t)p * n0, n0 * 8, hipMemcpyHostToDevice, st));
  if (n0) BCHK(hipMemcpyAsync(I.coef[0], kc.data(), n0 * 8, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_bfill_u64, dim3((u32)std::min<u64>(1024, (tcap + 255) / 256)), dim3(256), 0, st, I.table, tcap, 0ull);
  hipLaunchKernelGGL(k_bfill_u32, dim3(1), dim3(64), 0, st, queue, 1u, 0u);
  BInst* d_inst = reinterpret_cast<BInst*>(base + o_inst);
  u64* d_sup = reinterpret_cast<u64*>(base + o_sup);
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, st, d_inst, 1u, t0);
  const u32 lim = batch_gen_limit();
  switch (W) {
    case 1: hipLaunchKernelGGL(k_batch<1>, dim3(1), dim3(threads), 0, st, d_inst, d_sup, order, 1u, queue, 0u, inst->n_ops, lim); break;
    case 2: hipLaunchKernelGGL(k_batch<2>, dim3(1), dim3(threads), 0, st, d_inst, d_sup, order, 1u, queue, 0u, inst->n_ops, lim); break;
    case 4: hipLaunchKernelGGL(k_batch<4>, dim3(1), dim3(threads), 0, st, d_inst, d_sup, order, 1u, queue, 0u, inst->n_ops, lim); break;
    default: hipLaunchKernelGGL(k_batch<8>, dim3(1), dim3(threads), 0, st, d_inst, d_sup, order, 1u, queue, 0u, inst->n_ops, lim); break;
  }
  BCHK(hipGetLastError());
  BInst R{};
  BCHK(hipMemcpyAsync(&R, d_inst, sizeof(BInst), hipMemcpyDeviceToHost, st));
  BCHK(hipStreamSynchronize(st));
  if (R.status == IS_OVERFLOW)
    return bfail(PPVM_ENOMEM, "arena of %llu terms overflowed; an allocation-free run cannot grow it (pass a "
                 "larger capacity, up to what %zu bytes of scratch hold)", (unsigned long long)acap, kScratchBytes);
  if (R.status != IS_DONE) return bfail(PPVM_EHIP, "run did not finish (status %u)", R.status);
  *n = R.n;
  if (overlap_last) *overlap_last = R.ovl_last;
  if (term_updates) *term_updates = R.term_updates;
  if (cap < R.n) return bfail(PPVM_EINVAL, "export buffer holds %zu, %llu needed", cap, (unsigned long long)R.n);
  if (R.n) {
    if (!x || !z || !c) return bfail(PPVM_EINVAL, "null export array");
    const u32 wa = (inst->n_qubits + 63) / 64;
    std::vector<u64> out((size_t)np * R.n);
    for (u32 p = 0; p < np; ++p)
      BCHK(hipMemcpyAsync(out.data() + (size_t)p * R.n, R.planes[R.cur] + (size_t)p * acap, R.n * 8,
                          hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(c, R.coef[R.cur], R.n * 8, hipMemcpyDeviceToHost, st));
    BCHK(hipStreamSynchronize(st));
    for (u64 t = 0; t < R.n; ++t)
      for (u32 p = 0; p < wa; ++p) {
        x[t * wa + p] = out[(size_t)p * R.n + t];
        z[t * wa + p] = out[(size_t)(W + p) * R.n + t];
      }
  }
  return PPVM_OK;
}

}  // extern "C"
