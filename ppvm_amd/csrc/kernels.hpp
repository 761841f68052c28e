// kernels.hpp — HIP kernels of the gfx950 Pauli-propagation engine.
//
// One rotation gate exp(-i theta/2 G) is three launches over the arena:
//   k_rot_scan : stream the test word(s) of every entry; anticommuting live
//                entries are compacted (ballot ranks + decoupled look-back, so
//                the order is the arena order) into the worklist with their
//                full key and coefficient.
//   k_rot_pair : each worklist record inserts its pair key (key with the
//                generator's pivot bit cleared) into a generation-tagged
//                open-addressed table; the second member of a pair to arrive
//                computes both merged coefficients
//                   c_k' = c_k cos + eps_p c_p sin,  c_p' = c_p cos + eps_k c_k sin
//                and writes them in place (with fused truncation).
//   k_rot_emit : unpaired records scale their parent in place and append the
//                child k ^ M (coefficient eps_k c_k sin) at the arena end, in
//                worklist order, dropping children the truncation rejects.
// Commuting entries are never touched: their only traffic is the test word.
//
// Floating point: products and sums use __dmul_rn / __dadd_rn (no contraction),
// the same operation sequence as the reference's `*v *= cos; coeff *= ±sin;
// data[k'] += coeff` (crates/ppvm-pauli-sum/src/sum/rot1.rs:68-75,
// sum/data.rs:222-224), so merged coefficients are bit-identical.
#pragma once

#include "device.hpp"

namespace ppvm {

template <int W>
struct alignas(16) WLRec {
  u64 k[2 * W];
  double c;
  u32 idx;
  u32 flags;  // 1 = merged with its partner by k_rot_pair
};

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
  u32 r = (xbit<W>(k, g.a) & g.gaz) ^ (zbit<W>(k, g.a) & g.gax);
  if (g.two) r ^= (xbit<W>(k, g.b) & g.gbz) ^ (zbit<W>(k, g.b) & g.gbx);
  return r;
}
// Child key P xor G.
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

template <int W>
__device__ __forceinline__ void load_key(const u64* __restrict__ planes, u64 cap, u64 i, u64 (&k)[2 * W]) {
#pragma unroll
  for (int p = 0; p < 2 * W; ++p) k[p] = planes[(u64)p * cap + i];
}

// ---- k_rot_scan --------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(kBlock) void k_rot_scan(const u64* __restrict__ planes,
                                                     const double* __restrict__ coef, u64 cap,
                                                     WLRec<W>* __restrict__ wl, DevState* st,
                                                     u32 in_slot, u32 out_slot, RotSpec g,
                                                     u64* status, u32 gen) {
  __shared__ u32 lds_cnt[kItems * kWaves];
  __shared__ u32 lds_tile;
  __shared__ u64 lds_pref;
  __shared__ u32 lds_last;
  if (ld_relaxed(&st->overflow)) return;
  const u64 n = st->slot[in_slot].n_entries;
  const u64 n_tiles = (n + kTile - 1) / kTile;
  const u32 wa = g.a >> 6, wb = g.b >> 6;
  const bool needx = (g.gaz | g.gbz) != 0u;
  const bool needz = (g.gax | g.gbx) != 0u;
  while (true) {
    long long tile = next_tile(&st->ticket[TK_SCAN], n_tiles, &lds_tile);
    if (tile < 0) break;
    bool anti[kItems];
    double cv[kItems];
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
      u64 i = (u64)tile * kTile + (u64)it * kBlock + threadIdx.x;
      anti[it] = false;
      cv[it] = 0.0;
      if (i < n) {
        u32 r = 0;
        if (needx) {
          u64 xa = planes[(u64)wa * cap + i];
          r ^= bit_of(xa, g.a) & g.gaz;
          if (g.two) {
            u64 xb = (wb == wa) ? xa : planes[(u64)wb * cap + i];
            r ^= bit_of(xb, g.b) & g.gbz;
          }
        }
        if (needz) {
          u64 za = planes[(u64)(W + wa) * cap + i];
          r ^= bit_of(za, g.a) & g.gax;
          if (g.two) {
            u64 zb = (wb == wa) ? za : planes[(u64)(W + wb) * cap + i];
            r ^= bit_of(zb, g.b) & g.gbx;
          }
        }
        if (r) {
          cv[it] = coef[i];
          anti[it] = !is_dead(cv[it]);
        }
      }
    }
    u32 pos[kItems];
    u32 total = tile_ranks(anti, pos, lds_cnt);
    u64 prefix = lookback(status, gen, (u64)tile, total, &lds_pref);
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
      if (!anti[it]) continue;
      u64 i = (u64)tile * kTile + (u64)it * kBlock + threadIdx.x;
      WLRec<W>& r = wl[prefix + pos[it]];
      u64 k[2 * W];
      load_key<W>(planes, cap, i, k);
#pragma unroll
      for (int p = 0; p < 2 * W; ++p) r.k[p] = k[p];
      r.c = cv[it];
      r.idx = (u32)i;
      r.flags = 0u;
    }
  }
  if (last_block(&st->done_count[DK_SCAN], &lds_last) && threadIdx.x == 0) {
    u64 total = 0;
    if (n_tiles) total = ld_acquire(&status[n_tiles - 1]) & ((1ull << 38) - 1);
    st->n_wl = total;
    st->op_kills = 0;
  }
}

// ---- k_rot_pair ---------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(kBlock) void k_rot_pair(double* __restrict__ coef, u64 cap,
                                                     WLRec<W>* __restrict__ wl, DevState* st,
                                                     u32 in_slot, RotSpec g, u64* table,
                                                     u64 tcap_max, u32 tgen, TruncSpec t,
                                                     u64 op_tag) {
  __shared__ u64 lds_sum[kWaves];
  if (ld_relaxed(&st->overflow)) return;
  const u64 n_entries = st->slot[in_slot].n_entries;
  const u64 n_wl = st->n_wl;
  if (n_entries + n_wl > cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) cas_u64(&st->overflow, 0ull, op_tag);
    return;
  }
  u64 tcap = 64;
  while (tcap < 2 * n_wl && tcap < tcap_max) tcap <<= 1;
  const u64 tmask = tcap - 1;
  u64 kills = 0, pairs = 0;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n_wl; i += stride) {
    u64 k[2 * W];
#pragma unroll
    for (int p = 0; p < 2 * W; ++p) k[p] = wl[i].k[p];
    u64 rep[2 * W];
    pair_key<W>(k, g, rep);
    const u64 h = hash_key<W>(rep);
    const u64 fp = h >> 56;
    const u64 mine = ((u64)tgen << 40) | (fp << 32) | i;
    u64 pos = h & tmask;
    u64 s = ld_relaxed(&table[pos]);
    while (true) {
      if ((u32)(s >> 40) != tgen) {
        u64 old = cas_u64(&table[pos], s, mine);
        if (old == s) break;  // claimed: first member of its pair (so far)
        s = old;
        continue;
      }
      if (((s >> 32) & 0xffull) == fp) {
        const u64 j = s & 0xffffffffull;
        u64 kj[2 * W];
#pragma unroll
        for (int p = 0; p < 2 * W; ++p) kj[p] = wl[j].k[p];
        u64 repj[2 * W];
        pair_key<W>(kj, g, repj);
        bool same = true;
#pragma unroll
        for (int p = 0; p < 2 * W; ++p) same &= (repj[p] == rep[p]);
        if (same) {
          // second arrival: merge both members
          const double ci = wl[i].c, cj = wl[j].c;
          const double ei = rot_eps<W>(k, g), ej = rot_eps<W>(kj, g);
          const double ni = __dadd_rn(__dmul_rn(ci, g.cs), __dmul_rn(cj, ej * g.sn));
          const double nj = __dadd_rn(__dmul_rn(cj, g.cs), __dmul_rn(ci, ei * g.sn));
          bool di = false, dj = false;
          if (g.trunc) {
            di = drop_term<W>(k, ni, t);
            dj = drop_term<W>(kj, nj, t);
          }
          coef[wl[i].idx] = di ? dead_value() : ni;
          coef[wl[j].idx] = dj ? dead_value() : nj;
          wl[i].flags = 1u;
          wl[j].flags = 1u;
          kills += (u64)di + (u64)dj;
          pairs += 1;
          break;
        }
      }
      pos = (pos + 1) & tmask;
      s = ld_relaxed(&table[pos]);
    }
  }
  u64 bk = block_sum_u64(kills, lds_sum);
  u64 bp = block_sum_u64(pairs, lds_sum);
  if (threadIdx.x == 0) {
    if (bk) add_u64(&st->op_kills, bk);
    if (bp) add_u64(&st->pairs_total, bp);
  }
}

// ---- k_rot_emit ---------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(kBlock) void k_rot_emit(u64* __restrict__ planes,
                                                     double* __restrict__ coef, u64 cap,
                                                     const WLRec<W>* __restrict__ wl,
                                                     DevState* st, u32 in_slot, u32 out_slot,
                                                     RotSpec g, TruncSpec t, u64* status,
                                                     u32 gen) {
  __shared__ u32 lds_cnt[kItems * kWaves];
  __shared__ u32 lds_tile;
  __shared__ u64 lds_pref;
  __shared__ u64 lds_sum[kWaves];
  __shared__ u32 lds_last;
  if (ld_relaxed(&st->overflow)) return;
  const u64 n_entries = st->slot[in_slot].n_entries;
  const u64 n_wl = st->n_wl;
  if (n_entries + n_wl > cap) return;  // k_rot_pair flagged the overflow
  const u64 n_tiles = (n_wl + kTile - 1) / kTile;
  while (true) {
    long long tile = next_tile(&st->ticket[TK_EMIT], n_tiles, &lds_tile);
    if (tile < 0) break;
    bool child[kItems];
    u64 kills = 0;
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
      u64 w = (u64)tile * kTile + (u64)it * kBlock + threadIdx.x;
      child[it] = false;
      if (w < n_wl && wl[w].flags == 0u) {
        u64 k[2 * W];
#pragma unroll
        for (int p = 0; p < 2 * W; ++p) k[p] = wl[w].k[p];
        const double c = wl[w].c;
        const double eps = rot_eps<W>(k, g);
        const double cp = __dmul_rn(c, g.cs);
        const double cc = __dmul_rn(c, eps * g.sn);
        bool dp = false, keep = true;
        if (g.trunc) {
          dp = drop_term<W>(k, cp, t);
          rot_flip<W>(k, g);
          keep = !drop_term<W>(k, cc, t);
        }
        coef[wl[w].idx] = dp ? dead_value() : cp;
        kills += (u64)dp;
        child[it] = keep;
      }
    }
    u64 bk = block_sum_u64(kills, lds_sum);
    if (threadIdx.x == 0 && bk) add_u64(&st->op_kills, bk);
    u32 pos[kItems];
    u32 total = tile_ranks(child, pos, lds_cnt);
    u64 prefix = lookback(status, gen, (u64)tile, total, &lds_pref);
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
      if (!child[it]) continue;
      u64 w = (u64)tile * kTile + (u64)it * kBlock + threadIdx.x;
      u64 k[2 * W];
#pragma unroll
      for (int p = 0; p < 2 * W; ++p) k[p] = wl[w].k[p];
      const double eps = rot_eps<W>(k, g);
      const double cc = __dmul_rn(wl[w].c, eps * g.sn);
      rot_flip<W>(k, g);
      const u64 dst = n_entries + prefix + pos[it];
#pragma unroll
      for (int p = 0; p < 2 * W; ++p) planes[(u64)p * cap + dst] = k[p];
      coef[dst] = cc;
    }
  }
  if (last_block(&st->done_count[DK_EMIT], &lds_last) && threadIdx.x == 0) {
    u64 appended = 0;
    if (n_tiles) appended = ld_acquire(&status[n_tiles - 1]) & ((1ull << 38) - 1);
    const u64 kills = ld_relaxed(&st->op_kills);
    const u64 dead_in = st->slot[in_slot].n_dead;
    st->slot[out_slot].n_entries = n_entries + appended;
    st->slot[out_slot].n_dead = dead_in + kills;
    st->op_kills = 0;
    st->term_updates += n_entries - dead_in;
    st->gates += 1;
    st->rotations += 1;
    st->scanned += n_entries;
    st->scan_words += n_entries * g.test_words;
    st->wl_total += n_wl;
    st->children_total += appended;
    st->kills_total += kills;
  }
}

// ---- k_map: Clifford / noise / truncation program over every entry ------------
template <int W>
__device__ __forceinline__ void map_apply(u64 (&k)[2 * W], double& c, const MapOp& op,
                                          const MapProgram& prog) {
  const u32 code = op.code_flags & 0xffu;
  const u32 a = op.a, b = op.b;
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
    case M_SCALE: {
      const u32 tb = (op.code_flags >> 16) & 0xffu;
      u32 idx = (x | (z << 1)) * 4u;
      if (b != 0xffffffffu) idx += xbit<W>(k, b) | (zbit<W>(k, b) << 1);
      if ((prog.apply_mask[tb] >> idx) & 1u) c = __dmul_rn(c, prog.factors[tb][idx]);
      break;
    }
    default: break;
  }
  if (neg) c = -c;
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_map(u64* __restrict__ planes,
                                                double* __restrict__ coef, u64 cap,
                                                DevState* st, u32 in_slot, u32 out_slot,
                                                MapProgram prog, TruncSpec t) {
  __shared__ u64 lds_sum[kWaves];
  __shared__ u32 lds_last;
  if (ld_relaxed(&st->overflow)) return;
  const u64 n = st->slot[in_slot].n_entries;
  u64 kills = 0, updates = 0;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    double c = coef[i];
    if (is_dead(c)) continue;
    u64 k[2 * W], k0[2 * W];
    load_key<W>(planes, cap, i, k);
#pragma unroll
    for (int p = 0; p < 2 * W; ++p) k0[p] = k[p];
    const double c0 = c;
    bool killed = prog.trunc_before && drop_term<W>(k, c, t);
    for (u32 j = 0; j < prog.n_ops && !killed; ++j) {
      const MapOp op = prog.ops[j];
      ++updates;
      map_apply<W>(k, c, op, prog);
      if ((op.code_flags >> 8) & 1u) killed = drop_term<W>(k, c, t);
    }
    if (killed) {
      coef[i] = dead_value();
      ++kills;
      continue;
    }
#pragma unroll
    for (int p = 0; p < 2 * W; ++p)
      if (k[p] != k0[p]) planes[(u64)p * cap + i] = k[p];
    if (__double_as_longlong(c) != __double_as_longlong(c0)) coef[i] = c;
  }
  u64 bk = block_sum_u64(kills, lds_sum);
  u64 bu = block_sum_u64(updates, lds_sum);
  if (threadIdx.x == 0) {
    if (bk) add_u64(&st->op_kills, bk);
    if (bu) add_u64(&st->op_updates, bu);
  }
  if (last_block(&st->done_count[DK_MAP], &lds_last) && threadIdx.x == 0) {
    const u64 kills_all = ld_relaxed(&st->op_kills);
    const u64 upd = ld_relaxed(&st->op_updates);
    st->slot[out_slot].n_entries = n;
    st->slot[out_slot].n_dead = st->slot[in_slot].n_dead + kills_all;
    st->op_kills = 0;
    st->op_updates = 0;
    st->term_updates += upd;
    st->gates += prog.n_ops;
    st->kills_total += kills_all;
    st->map_entries += n;
    st->map_launches += 1;
  }
}

// ---- k_overlap: sum of coefficients of terms with an all-zero x plane ---------
template <int W>
__global__ __launch_bounds__(kBlock) void k_overlap(const u64* __restrict__ planes,
                                                    const double* __restrict__ coef, u64 cap,
                                                    DevState* st, u32 in_slot,
                                                    double* partials) {
  __shared__ double lds_d[kBlock];
  __shared__ u32 lds_last;
  const u64 n = st->slot[in_slot].n_entries;
  double acc = 0.0;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    u64 xo = 0;
#pragma unroll
    for (int p = 0; p < W; ++p) xo |= planes[(u64)p * cap + i];
    if (xo == 0) {
      const double c = coef[i];
      if (!is_dead(c)) acc = __dadd_rn(acc, c);
    }
  }
  lds_d[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) lds_d[threadIdx.x] = __dadd_rn(lds_d[threadIdx.x], lds_d[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = lds_d[0];
  if (last_block(&st->done_count[DK_OVERLAP], &lds_last) && threadIdx.x == 0) {
    double tot = 0.0;
    for (u32 b = 0; b < gridDim.x; ++b) tot = __dadd_rn(tot, partials[b]);
    st->overlap = tot;
  }
}

// ---- k_compact: stable removal of dead entries into the other arena ------------
template <int W>
__global__ __launch_bounds__(kBlock) void k_compact(const u64* __restrict__ src_planes,
                                                    const double* __restrict__ src_coef,
                                                    u64* __restrict__ dst_planes,
                                                    double* __restrict__ dst_coef, u64 cap,
                                                    DevState* st, u32 in_slot, u32 out_slot,
                                                    u64* status, u32 gen) {
  __shared__ u32 lds_cnt[kItems * kWaves];
  __shared__ u32 lds_tile;
  __shared__ u64 lds_pref;
  __shared__ u32 lds_last;
  const u64 n = st->slot[in_slot].n_entries;
  const u64 n_tiles = (n + kTile - 1) / kTile;
  while (true) {
    long long tile = next_tile(&st->ticket[TK_COMPACT], n_tiles, &lds_tile);
    if (tile < 0) break;
    bool live[kItems];
    double cv[kItems];
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
      u64 i = (u64)tile * kTile + (u64)it * kBlock + threadIdx.x;
      live[it] = false;
      cv[it] = 0.0;
      if (i < n) {
        cv[it] = src_coef[i];
        live[it] = !is_dead(cv[it]);
      }
    }
    u32 pos[kItems];
    u32 total = tile_ranks(live, pos, lds_cnt);
    u64 prefix = lookback(status, gen, (u64)tile, total, &lds_pref);
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
      if (!live[it]) continue;
      u64 i = (u64)tile * kTile + (u64)it * kBlock + threadIdx.x;
      u64 dst = prefix + pos[it];
#pragma unroll
      for (int p = 0; p < 2 * W; ++p) dst_planes[(u64)p * cap + dst] = src_planes[(u64)p * cap + i];
      dst_coef[dst] = cv[it];
    }
  }
  if (last_block(&st->done_count[DK_COMPACT], &lds_last) && threadIdx.x == 0) {
    u64 total = 0;
    if (n_tiles) total = ld_acquire(&status[n_tiles - 1]) & ((1ull << 38) - 1);
    st->slot[out_slot].n_entries = total;
    st->slot[out_slot].n_dead = 0;
  }
}

// ---- small utility kernels ------------------------------------------------------
__global__ void k_fill_u64(u64* p, u64 n, u64 v) {
  const u64 stride = (u64)gridDim.x * blockDim.x;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Grow: copy the first n entries of every plane and the coefficients from an
// arena of capacity cap_old into one of capacity cap_new.
__global__ void k_copy_arena(const u64* __restrict__ sp, const double* __restrict__ sc, u64 cap_old,
                             u64* __restrict__ dp, double* __restrict__ dc, u64 cap_new, u32 nplanes,
                             u64 n) {
  const u64 stride = (u64)gridDim.x * blockDim.x;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    for (u32 p = 0; p < nplanes; ++p) dp[(u64)p * cap_new + i] = sp[(u64)p * cap_old + i];
    dc[i] = sc[i];
  }
}

}  // namespace ppvm
