// small.hpp — small sums held whole in LDS: one workgroup runs a planned list
// of steps (rotations, fused map programs, runs of rotations that act on no
// term) in a single launch.
//
// A sum of a few hundred to a few thousand terms gives each kernel of the
// arena pipeline only a handful of waves, so a gate costs several launches and
// the host synchronisations between them.  Here the live entries are loaded
// once into LDS (keys as structure of arrays: plane p of entry i at
// p * CAP + i, so a wave reads consecutive words), every step is applied by
// one workgroup with block-wide barriers only, and the result is written back
// to the arena, dense and in order.  The steps are the ones the host planner
// would have issued one by one (the same RotSpec and MapProgram records), and
// each is applied with the arena pipeline's arithmetic and order:
//   rotation: every entry stays in place with its new coefficient (c cos,
//     plus eps sin times the partner's coefficient when the flipped word is
//     present), then the children of unpaired anticommuting entries follow
//     in entry order; truncation, when fused, is applied to both;
//   map program: the program's ops per entry, dropping at each truncation
//     point, entries kept in order.
// The run stops before a step whose result would not fit in LDS (the state
// is then the one after the previous step) and reports how many steps it
// completed; the host issues the rest through the arena pipeline.
#pragma once

#include "kernels.hpp"

namespace ppvm {

constexpr u32 kSmallMaxThreads = 1024;

// Entries held in LDS per word count, and the pair table size (>= 2x).
template <int W> struct SmallCap;
template <> struct SmallCap<1> { static constexpr u32 cap = 2048, tsize = 4096; };
template <> struct SmallCap<2> { static constexpr u32 cap = 1536, tsize = 4096; };
template <> struct SmallCap<4> { static constexpr u32 cap = 768, tsize = 2048; };
template <> struct SmallCap<8> { static constexpr u32 cap = 384, tsize = 1024; };

enum SmallKind : u32 { SK_ROT = 0, SK_MAP = 1, SK_SKIP = 2 };
struct SmallStep {
  u32 kind;
  u32 arg;      // SK_MAP: index into maps, SK_SKIP: rotations in the run
  u32 arg2;     // SK_SKIP: map ops in the run
  u32 pad;
  RotSpec rot;  // SK_ROT
};
constexpr u32 kSmallStage = 32;  // step records staged in LDS at a time

struct SmallArgs {
  u64* keys;                // arena (array of structures, 2W words per entry)
  double* coef;
  u64 arena_cap;            // entries the arena holds: the run stops before a step that would exceed it
  DevState* st;
  u32 in_slot, out_slot;
  const SmallStep* steps;
  u32 n_steps;
  const MapProgram* maps;
  TruncSpec t;
};

// Position of this thread's flagged item among the block's flagged items, and
// the block total.  Wave counts are double-buffered by `par`, so one barrier
// per call suffices: the next call writing the same buffer comes after the
// barrier of the call in between.
__device__ __forceinline__ u32 small_rank(bool f, u32& pos, u32* wcnt, u32& par) {
  const u32 lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const u64 b = __ballot(f);
  u32* buf = wcnt + par * (kSmallMaxThreads / 32);
  if (lane == 0) buf[wave] = (u32)__popcll(b);
  __syncthreads();
  u32 base = 0, tot = 0;
  for (u32 w = 0; w < nwave; ++w) {
    const u32 v = buf[w];
    base += w < wave ? v : 0u;
    tot += v;
  }
  par ^= 1u;
  pos = base + (u32)__popcll(b & ((1ull << lane) - 1ull));
  return tot;
}

template <int W, u32 CAP>
__device__ __forceinline__ void sload(const u64* P, u32 i, u64 (&k)[2 * W]) {
#pragma unroll
  for (int p = 0; p < 2 * W; ++p) k[p] = P[p * CAP + i];
}
template <int W, u32 CAP>
__device__ __forceinline__ void sstore(u64* P, u32 i, const u64 (&k)[2 * W]) {
#pragma unroll
  for (int p = 0; p < 2 * W; ++p) P[p * CAP + i] = k[p];
}

// Block total of a per-thread count.
__device__ __forceinline__ u64 small_total(u64 x, u64* red) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  u64 t = 0;
  for (u32 w = 0; w < blockDim.x / 64; ++w) t += red[w];
  return t;
}

// Tallies of one run (per thread; summed at the end).
struct SmallTally {
  u64 updates, wl, pairs, children, kills;
};

// Values every lane holds alike (read from one LDS address), moved to scalar
// registers so that branches on them are scalar.
__device__ __forceinline__ u32 small_uni(u32 v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double small_uni(double v) {
  const u64 b = (u64)__double_as_longlong(v);
  return __longlong_as_double((long long)(((u64)small_uni((u32)(b >> 32)) << 32) | small_uni((u32)b)));
}
__device__ __forceinline__ RotSpec small_uni(const RotSpec& r) {
  RotSpec g;
  g.a = small_uni(r.a);
  g.b = small_uni(r.b);
  g.two = small_uni(r.two);
  g.gax = small_uni(r.gax);
  g.gaz = small_uni(r.gaz);
  g.gbx = small_uni(r.gbx);
  g.gbz = small_uni(r.gbz);
  g.trunc = small_uni(r.trunc);
  g.test_planes = small_uni(r.test_planes);
  g.mode = small_uni(r.mode);
  g.sn = small_uni(r.sn);
  g.cs = small_uni(r.cs);
  return g;
}

// Two flags ranked at once (one barrier): positions and block totals.
__device__ __forceinline__ void small_rank2(bool f0, bool f1, u32& p0, u32& p1, u32& t0, u32& t1, u32* wcnt,
                                            u32& par) {
  const u32 lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const u64 b0 = __ballot(f0), b1 = __ballot(f1);
  u32* buf = wcnt + par * (kSmallMaxThreads / 32);
  if (lane == 0) {
    buf[2 * wave] = (u32)__popcll(b0);
    buf[2 * wave + 1] = (u32)__popcll(b1);
  }
  __syncthreads();
  u32 s0 = 0, s1 = 0;
  t0 = t1 = 0;
  for (u32 w = 0; w < nwave; ++w) {
    const u32 v0 = buf[2 * w], v1 = buf[2 * w + 1];
    s0 += w < wave ? v0 : 0u;
    s1 += w < wave ? v1 : 0u;
    t0 += v0;
    t1 += v1;
  }
  par ^= 1u;
  const u64 below = (1ull << lane) - 1ull;
  p0 = s0 + (u32)__popcll(b0 & below);
  p1 = s1 + (u32)__popcll(b1 & below);
}

// Entry i of P/C after the rotation (pass 0: the entry itself, kept in place)
// or its child (pass 1: only for unpaired branching entries).  Returns whether
// the item exists (before truncation) and sets keep.
template <int W, u32 CAP>
__device__ __forceinline__ bool small_item(const u64* P, const double* C, const double* D, const u32* partner,
                                           u32 i, u32 pass, const RotSpec& g, const TruncSpec& t, u32 tag,
                                           u64 (&k)[2 * W], double& c, bool& keep) {
  sload<W, CAP>(P, i, k);
  c = C[i];
  const bool damp = g.mode == kBranchDamping;
  const bool anti = rot_anti<W>(k, g) != 0u;
  const u32 pl = anti ? partner[i] : 0u;
  const bool paired = anti && (pl & 0xffff0000u) == tag;
  const bool branches = anti && (!damp || zbit<W>(k, g.a));
  if (pass == 0) {
    if (paired) {
      c = D[i];  // computed by the pairing phase
    } else if (branches) {
      c = __dmul_rn(c, g.cs);
    }
    keep = !g.trunc || !drop_term<W>(k, c, t);
    return true;
  }
  if (!branches || paired) {
    keep = false;
    return false;
  }
  c = __dmul_rn(c, rot_eps<W>(k, g) * g.sn);
  rot_flip<W>(k, g);
  keep = !g.trunc || !drop_term<W>(k, c, t);
  return true;
}

// Both items of entry i at once (one set of loads): the entry with its new
// coefficient, and (candc) the child of an unpaired branching entry.
template <int W, u32 CAP>
__device__ __forceinline__ void small_items(const u64* P, const double* C, const double* D, const u32* partner,
                                            u32 i, const RotSpec& g, const TruncSpec& t, u32 tag,
                                            u64 (&k)[2 * W], double& c, bool& keep, u64 (&kc)[2 * W],
                                            double& cc, bool& keepc, bool& candc) {
  sload<W, CAP>(P, i, k);
  const double c0 = C[i], dp = D[i];
  const u32 pl = partner[i];
  const bool anti = rot_anti<W>(k, g) != 0u;
  const bool paired = anti && (pl & 0xffff0000u) == tag;
  const bool branches = anti && (g.mode != kBranchDamping || zbit<W>(k, g.a));
  c = paired ? dp : (branches ? __dmul_rn(c0, g.cs) : c0);
  keep = !g.trunc || !drop_term<W>(k, c, t);
  candc = branches && !paired;
  keepc = false;
  if (candc) {
    cc = __dmul_rn(c0, rot_eps<W>(k, g) * g.sn);
#pragma unroll
    for (int p = 0; p < 2 * W; ++p) kc[p] = k[p];
    rot_flip<W>(kc, g);
    keepc = !g.trunc || !drop_term<W>(kc, cc, t);
  }
}

// Slot of a pair key in the LDS table (TS a power of two): one 64-bit
// multiply per word and a 32-bit finish, top bits.
template <int W, u32 TS>
__device__ __forceinline__ u32 small_slot(const u64 (&r)[2 * W]) {
  u64 h = r[0];
#pragma unroll
  for (int p = 1; p < 2 * W; ++p) h = (h ^ (h >> 31)) * 0x9E3779B97F4A7C15ull + r[p];
  h = (h ^ (h >> 29)) * 0xBF58476D1CE4E5B9ull;
  const u32 x = (u32)(h >> 32) ^ (u32)h;
  return (x * 0x85EBCA6Bu) >> (32 - __builtin_ctz(TS));
}

// One rotation or damping branch: P/C (n entries) -> Q/D.  Table slots and
// partner links carry the generation gen in their upper 16 bits, so neither
// needs clearing between rotations.  The member of a pair that finds the
// other computes both new coefficients and parks them in D at the members'
// own indices (D is free until the write phase, which only writes at or
// below the index being read).  Returns 1 when the result is in Q/D,
// 0 when no entry anticommutes with the generator (nothing changes: on a
// clean state a fused truncation drops nothing either, so P/C stays
// current), and -1 when the result would exceed lim <= CAP entries (P/C untouched).
template <int W, u32 CAP, u32 TS>
__device__ __forceinline__ int small_rotate(const u64* P, const double* C, u64* Q, double* D, u32* table,
                                            u32* partner, u32* wcnt, u32& par, u32& n, u32 lim, const RotSpec& g,
                                            const TruncSpec& t, u32 gen, SmallTally& ty) {
  const u32 T = blockDim.x, tid = threadIdx.x;
  const u32 tag = gen << 16;
  // pair up: the first member of {P, P xor G} claims a slot, the second links both
  bool any = false;
  for (u32 i = tid; i < n; i += T) {
    u64 k[2 * W];
    sload<W, CAP>(P, i, k);
    if (!rot_anti<W>(k, g)) continue;
    any = true;
    ++ty.wl;
    u64 rep[2 * W];
    pair_key<W>(k, g, rep);
    u32 pos = small_slot<W, TS>(rep);
    u32 v = table[pos];
    while (true) {
      if ((v & 0xffff0000u) != tag) {
        const u32 old = atomicCAS(&table[pos], v, tag | i);
        if (old == v) break;
        v = old;
        continue;
      }
      const u32 j = v & 0xffffu;
      u64 kj[2 * W], rj[2 * W];
      sload<W, CAP>(P, j, kj);
      pair_key<W>(kj, g, rj);
      bool same = true;
#pragma unroll
      for (int p = 0; p < 2 * W; ++p) same &= rj[p] == rep[p];
      if (same) {
        const double ci = C[i], cj = C[j];
        double ni, nj;
        if (g.mode == kBranchDamping) {
          const bool zi = zbit<W>(k, g.a) != 0u;
          const double cz = zi ? ci : cj, cI = zi ? cj : ci;
          const double nz = __dmul_rn(cz, g.cs), nI = __dadd_rn(cI, __dmul_rn(cz, g.sn));
          ni = zi ? nz : nI;
          nj = zi ? nI : nz;
        } else {
          ni = __dadd_rn(__dmul_rn(ci, g.cs), __dmul_rn(cj, rot_eps<W>(kj, g) * g.sn));
          nj = __dadd_rn(__dmul_rn(cj, g.cs), __dmul_rn(ci, rot_eps<W>(k, g) * g.sn));
        }
        D[i] = ni;
        D[j] = nj;
        partner[i] = tag | j;
        partner[j] = tag | i;
        ++ty.pairs;
        break;
      }
      pos = (pos + 1u) & (TS - 1u);
      v = table[pos];
    }
  }
  u32 unused;
  if (small_rank(any, unused, wcnt, par) == 0) return 0;  // one barrier: does any entry anticommute?
  if (n <= T) {
    // one item per thread: the entry and its child ranked together
    const u32 i = tid;
    u64 k[2 * W], kc[2 * W];
    double c = 0.0, cc = 0.0;
    bool keep = false, keepc = false, cand = false, candc = false;
    if (i < n) {
      cand = true;
      small_items<W, CAP>(P, C, D, partner, i, g, t, tag, k, c, keep, kc, cc, keepc, candc);
    }
    ty.kills += ((cand && !keep) ? 1u : 0u) + ((candc && !keepc) ? 1u : 0u);
    ty.children += keepc ? 1u : 0u;
    u32 p0, p1, t0, t1;
    small_rank2(keep, keepc, p0, p1, t0, t1, wcnt, par);
    if (t0 + t1 > lim) return -1;  // uniform
    if (keep) {
      sstore<W, CAP>(Q, p0, k);
      D[p0] = c;
    }
    if (keepc) {
      sstore<W, CAP>(Q, t0 + p1, kc);
      D[t0 + p1] = cc;
    }
    n = t0 + t1;
    return 1;
  }
  // entries in place (pass 0), then the children (pass 1)
  u32 out0 = 0, out1 = 0;
  for (u32 pass = 0; pass < 2; ++pass) {
    for (u32 base = 0; base < n; base += T) {
      const u32 i = base + tid;
      bool keep = false, cand = false;
      u64 k[2 * W];
      double c = 0.0;
      if (i < n) cand = small_item<W, CAP>(P, C, D, partner, i, pass, g, t, tag, k, c, keep);
      ty.kills += (cand && !keep) ? 1u : 0u;
      ty.children += (pass == 1 && keep) ? 1u : 0u;
      u32 pos;
      const u32 tot = small_rank(keep, pos, wcnt, par);
      if (out0 + out1 + tot > lim) return -1;  // uniform
      if (keep) {
        const u32 d = out0 + (pass ? out1 : 0u) + pos;
        sstore<W, CAP>(Q, d, k);
        D[d] = c;
      }
      if (pass) out1 += tot;
      else out0 += tot;
    }
  }
  n = out0 + out1;
  return 1;
}

// One map program (already in LDS): P/C (n entries) -> Q/D (never grows).
// Every lane walks the whole op list, so each op is one uniform LDS read.
template <int W, u32 CAP>
__device__ __forceinline__ void small_map(const u64* P, const double* C, u64* Q, double* D, u32* wcnt, u32& par,
                                          u32& n, const MapProgram& prog, const TruncSpec& t, SmallTally& ty) {
  const u32 T = blockDim.x, tid = threadIdx.x;
  const u32 nops = small_uni(prog.n_ops);
  const bool tb = small_uni(prog.trunc_before) != 0u;
  const bool idle = small_uni(prog.n_idle) != 0u;
  u32 out = 0;
  for (u32 base = 0; base < n; base += T) {
    const u32 i = base + tid;
    bool keep = false;
    u64 k[2 * W];
    double c = 0.0;
    if (i < n) {
      sload<W, CAP>(P, i, k);
      c = C[i];
      bool killed = tb && drop_term<W>(k, c, t);
      if (idle && !killed) ty.updates += small_uni((u32)prog.idle[0]);
      for (u32 j = 0; j < nops; ++j) {
        MapOp op;
        op.code_flags = small_uni(prog.ops[j].code_flags);
        op.a = small_uni(prog.ops[j].a);
        op.b = small_uni(prog.ops[j].b);
        if (!killed) {
          ++ty.updates;
          map_apply<W>(k, c, op, prog);
          if ((op.code_flags >> 8) & 1u) killed = drop_term<W>(k, c, t);
          if (idle && !killed) ty.updates += small_uni((u32)prog.idle[j + 1]);
        }
      }
      keep = !killed;
      ty.kills += killed ? 1u : 0u;
    }
    u32 pos;
    const u32 tot = small_rank(keep, pos, wcnt, par);
    if (keep) {
      sstore<W, CAP>(Q, out + pos, k);
      D[out + pos] = c;
    }
    out += tot;
  }
  n = out;
}

template <int W>
__global__ __launch_bounds__(kSmallMaxThreads) void k_small(SmallArgs a) {
  constexpr u32 CAP = SmallCap<W>::cap, TS = SmallCap<W>::tsize;
  __shared__ u64 lk[2][2 * W * CAP];
  __shared__ double lc[2][CAP];
  __shared__ u32 table[TS];
  __shared__ u32 partner[CAP];
  __shared__ u32 wcnt[2 * (kSmallMaxThreads / 32)];
  __shared__ MapProgram lprog;
  __shared__ SmallStep lsteps[kSmallStage];
  __shared__ u64 red[kSmallMaxThreads / 64];
  DevState* st = a.st;
  if (st->overflow) return;
  const u32 T = blockDim.x, tid = threadIdx.x;
  const CountSlot in = st->slot[a.in_slot];
  const u32 lim = a.arena_cap < CAP ? (u32)a.arena_cap : CAP;  // entries a step may leave
  for (u32 j = tid; j < TS; j += T) table[j] = 0u;
  for (u32 j = tid; j < CAP; j += T) partner[j] = 0u;
  // live entries, in order
  u32 par = 0, n = 0;
  bool fits = true;
  for (u64 base = 0; base < in.n_entries; base += T) {
    const u64 i = base + tid;
    double c = 0.0;
    const bool live = i < in.n_entries && !is_dead(c = a.coef[i]);
    u32 pos;
    const u32 tot = small_rank(live, pos, wcnt, par);
    if (n + tot > lim) {
      fits = false;
      break;
    }
    if (live) {
      u64 k[2 * W];
      akey_load<W>(a.keys, i, k);
      sstore<W, CAP>(lk[0], n + pos, k);
      lc[0][n + pos] = c;
    }
    n += tot;
  }
  if (!fits) {  // nothing done: the counts carry over
    if (tid == 0) {
      st->slot[a.out_slot] = in;
      st->small_done = 0;
      st->small_peak = 0;
    }
    return;
  }
  SmallTally ty{0, 0, 0, 0, 0};
  u64 gates = 0, rots = 0, uni_updates = 0, scanned = 0;  // block-uniform
  u32 cur = 0, gen = 0, s = 0, peak = n;
  for (; s < a.n_steps; ++s) {
    __syncthreads();  // the previous step's writes are visible, and its record read
    if (s % kSmallStage == 0) {
      const u32 m = (a.n_steps - s < kSmallStage ? a.n_steps - s : kSmallStage) * (sizeof(SmallStep) / 4);
      const u32* src = reinterpret_cast<const u32*>(a.steps + s);
      u32* dst = reinterpret_cast<u32*>(lsteps);
      for (u32 j = tid; j < m; j += T) dst[j] = src[j];
      __syncthreads();
    }
    const SmallStep& step = lsteps[s % kSmallStage];
    const u32 kind = small_uni(step.kind);
    if (kind == SK_SKIP) {
      const u32 nskip = small_uni(step.arg), nskipm = small_uni(step.arg2);
      uni_updates += (u64)(nskip + nskipm) * n;
      gates += nskip + nskipm;
      rots += nskip;
      continue;
    }
    const u32 n_in = n;
    if (kind == SK_ROT) {
      const RotSpec g = small_uni(step.rot);
      if (++gen == 0xffffu) {  // generation wrap: clear the tags
        for (u32 j = tid; j < TS; j += T) table[j] = 0u;
        for (u32 j = tid; j < CAP; j += T) partner[j] = 0u;
        __syncthreads();
        gen = 1;
      }
      const int r = small_rotate<W, CAP, TS>(lk[cur], lc[cur], lk[cur ^ 1], lc[cur ^ 1], table, partner, wcnt,
                                             par, n, lim, g, a.t, gen, ty);
      if (r < 0) break;
      cur ^= (u32)r;
      uni_updates += n_in;
      scanned += n_in;
      gates += 1;
      rots += 1;
    } else {
      // the program into LDS: every lane reads each op at the same address
      const u32* src = reinterpret_cast<const u32*>(a.maps + small_uni(step.arg));
      u32* dst = reinterpret_cast<u32*>(&lprog);
      for (u32 j = tid; j < sizeof(MapProgram) / 4; j += T) dst[j] = src[j];
      __syncthreads();
      small_map<W, CAP>(lk[cur], lc[cur], lk[cur ^ 1], lc[cur ^ 1], wcnt, par, n, lprog, a.t, ty);
      gates += small_uni(lprog.n_ops) + small_uni(lprog.n_idle);
      cur ^= 1u;
    }
    peak = n > peak ? n : peak;
  }
  __syncthreads();
  for (u32 i = tid; i < n; i += T) {
    u64 k[2 * W];
    sload<W, CAP>(lk[cur], i, k);
    akey_store<W>(a.keys, i, k);
    a.coef[i] = lc[cur][i];
  }
  const u64 s_upd = small_total(ty.updates, red), s_wl = small_total(ty.wl, red);
  const u64 s_pairs = small_total(ty.pairs, red), s_child = small_total(ty.children, red);
  const u64 s_kills = small_total(ty.kills, red);
  if (tid == 0) {
    CountSlot o = in;
    o.n_entries = n;
    o.n_dead = 0;
    st->slot[a.out_slot] = o;
    st->term_updates += uni_updates + s_upd;
    st->gates += gates;
    st->rotations += rots;
    st->scanned += scanned;
    st->wl_total += s_wl;
    st->pairs_total += s_pairs;
    st->children_total += s_child;
    st->kills_total += s_kills;
    st->tvw = 0;  // the bit-sliced planes no longer match the arena
    st->small_done = s;
    st->small_peak = peak;
  }
}

}  // namespace ppvm
