// kernels.hpp — HIP kernels of the gfx950 Pauli-propagation engine.
//
// One rotation gate exp(-i theta/2 G) runs as:
//   k_rot_count  : stream the test word(s) of every entry; ballot masks of the
//                  anticommuting entries + one count per block chunk.
//   k_rot_gather : one 16-byte record per anticommuting entry (coefficient,
//                  arena index, the child's sign and the truncation tests of
//                  the key and its child) plus its pair-key hash, in arena
//                  order (dead entries are flagged).
//   k_pair_hist, k_pair_colscan, k_pair_scatter_tiled, k_pair_lds : radix-partition the worklist
//                  on the hash of the pair key (key with the generator's pivot
//                  bit cleared) and pair each bucket in an LDS hash table; the
//                  second member of a pair to arrive computes both merged
//                  coefficients
//                    c_k' = c_k cos + eps_p c_p sin,  c_p' = c_p cos + eps_k c_k sin
//                  and writes them in place (with fused truncation).
//   k_emit_count : unpaired records scale their parent in place and decide
//                  whether their child k ^ M (coefficient eps_k c_k sin)
//                  survives the truncation; merged records store their new
//                  coefficient in the arena; masks + counts.
//   k_emit_write : append the surviving children at the arena end, in
//                  worklist order (parent key read from the arena, generator
//                  bits flipped); publish the new counts and statistics.
// Commuting entries are never touched: their only traffic is the test word.
//
// Floating point: products and sums use __dmul_rn / __dadd_rn (no contraction),
// the same operation sequence as the reference's `*v *= cos; coeff *= ±sin;
// data[k'] += coeff` (crates/ppvm-pauli-sum/src/sum/rot1.rs:68-75,
// sum/data.rs:222-224), so merged coefficients are bit-identical.
#pragma once

#include "algebra.hpp"
#include "pattern.hpp"

namespace ppvm {

// ---- chunked passes ------------------------------------------------------------
// A pass over n items is split into G contiguous block chunks whose sizes are
// multiples of 256 (4 mask words) and at most kChunkMaxWords mask words; G is
// derived on the device from n and a target chunk size.  A count kernel writes each
// chunk's ballot masks and count; the matching write kernel derives its
// chunk's offset from the counts of the chunks before it and ranks the items
// inside the chunk with a scan of the mask popcounts in LDS.
constexpr int kChunkMaxWords = 1024;
constexpr u32 kArenaChunk = 8192;   // largest target items per chunk, arena passes
constexpr u32 kListChunk = 2048;    // largest target items per chunk, worklist passes
constexpr u32 kChunksAimed = 1024;  // chunks aimed at for a pass over n items

// Target chunk size of the rotation passes for n items: about kChunksAimed
// chunks of at least 256 items, at most `max_target` items each, so that a
// gate on a small arena still spreads over hundreds of workgroups.  The
// passes loop over their chunks, so a launch sized from an upper bound of n
// covers any actual n (blocks beyond the chunk count exit at once).
__host__ __device__ __forceinline__ u32 chunk_target(u64 n, u32 max_target) {
  u64 t = (n + kChunksAimed - 1) / kChunksAimed;
  t = (t + 255) & ~255ull;
  if (t < 256) t = 256;
  if (t > max_target) t = max_target;
  return (u32)t;
}

// Effective chunk count for n items with a target chunk size `target`
// (computed on the device from the actual n, so the launch may carry extra
// blocks; those exit at once).  Chunk sizes are multiples of 256 and at most
// kChunkMaxWords * 64 items, which holds for n <= kMaxParts * kChunkMaxWords * 64.
__device__ __forceinline__ u32 chunk_count(u64 n, u32 target) {
  u64 g = (n + target - 1) / target;
  if (g < 1) g = 1;
  if (g > (u64)kMaxParts) g = kMaxParts;
  return (u32)g;
}
__device__ __forceinline__ void block_chunk(u64 n, u32 G, u32 b, u64& lo, u64& hi) {
  u64 per = (n + G - 1) / G;
  per = (per + 255) & ~255ull;
  if (per > (u64)kChunkMaxWords * 64) per = (u64)kChunkMaxWords * 64;
  lo = (u64)b * per;
  if (lo > n) lo = n;
  hi = lo + per;
  if (hi > n) hi = n;
}
__device__ __forceinline__ u64 lanes_below(u32 lane) { return (1ull << lane) - 1ull; }

// Sum of v over the block, returned to every thread.
__device__ __forceinline__ u64 block_sum_all(u64 v, u64* lds) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const u32 lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  u64 t = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) t += lds[w];
  __syncthreads();
  return t;
}

// Sum of v[0..b), returned to every thread.  Four independent loads per
// thread are in flight per step (b reaches tens of thousands for large arenas).
__device__ __forceinline__ u64 block_sum_prefix(const u64* __restrict__ v, u32 b, u64* lds) {
  u64 s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  u32 j = threadIdx.x;
  for (; j + 3u * kBlock < b; j += 4u * kBlock) {
    s0 += v[j];
    s1 += v[j + kBlock];
    s2 += v[j + 2u * kBlock];
    s3 += v[j + 3u * kBlock];
  }
  for (; j < b; j += kBlock) s0 += v[j];
  return block_sum_all((s0 + s1) + (s2 + s3), lds);
}

// Sums of four arrays of lengths na, nb, nc, nd in one pass (loads of all four
// in flight together); valid in thread 0.
__device__ __forceinline__ void block_sums4(const u64* __restrict__ a, u32 na, const u64* __restrict__ b, u32 nb,
                                            const u64* __restrict__ c, u32 nc, const u64* __restrict__ d, u32 nd,
                                            u64* lds4, u64 (&out)[4]) {
  u64 v[4] = {0, 0, 0, 0};
  const u32 n = max(max(na, nb), max(nc, nd));
  for (u32 j = threadIdx.x; j < n; j += kBlock) {
    if (j < na) v[0] += a[j];
    if (j < nb) v[1] += b[j];
    if (j < nc) v[2] += c[j];
    if (j < nd) v[3] += d[j];
  }
  const u32 lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_down(v[q], o, 64);
  }
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) lds4[q * kWaves + wave] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    u64 t = 0;
    for (int w = 0; w < kWaves; ++w) t += lds4[q * kWaves + w];
    out[q] = t;
  }
}

// Sum of cnt[0..b) (the chunk's offset), returned to every thread.
__device__ __forceinline__ u64 chunk_offset(const u64* __restrict__ cnt, u32 b, u64* lds) {
  return block_sum_prefix(cnt, b, lds);
}

// Loads the chunk's nw mask words into lm and writes their exclusive popcount
// prefix into loff.
__device__ __forceinline__ void chunk_mask_scan(const u64* __restrict__ masks, u64 w0, u32 nw,
                                                u64* lm, u32* loff, u32* lt) {
  const u32 tid = threadIdx.x;
  const u32 q = (nw + kBlock - 1) / kBlock;
  const u32 j0 = tid * q;
  u32 local = 0;
  for (u32 j = j0; j < j0 + q && j < nw; ++j) {
    const u64 m = masks[w0 + j];
    lm[j] = m;
    loff[j] = local;
    local += (u32)__popcll(m);
  }
  // exclusive prefix of `local` over the block: wave scans, then wave totals
  const u32 lane = tid & 63u, wave = tid >> 6;
  u32 incl = local;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const u32 y = __shfl_up(incl, o, 64);
    if (lane >= (u32)o) incl += y;
  }
  if (lane == 63u) lt[wave] = incl;
  __syncthreads();
  u32 excl = incl - local;
  for (u32 w = 0; w < wave; ++w) excl += lt[w];
  for (u32 j = j0; j < j0 + q && j < nw; ++j) loff[j] += excl;
  __syncthreads();
}

// Position (0..63) of the k-th set bit of m (k < popcount(m)).
__device__ __forceinline__ u32 select_bit(u64 m, u32 k) {
  u32 pos = 0;
  u32 c = (u32)__popcll(m & 0xffffffffull);
  if (k >= c) { k -= c; m >>= 32; pos += 32; }
  c = (u32)__popcll(m & 0xffffull);
  if (k >= c) { k -= c; m >>= 16; pos += 16; }
  c = (u32)__popcll(m & 0xffull);
  if (k >= c) { k -= c; m >>= 8; pos += 8; }
  c = (u32)__popcll(m & 0xfull);
  if (k >= c) { k -= c; m >>= 4; pos += 4; }
  c = (u32)__popcll(m & 0x3ull);
  if (k >= c) { k -= c; m >>= 2; pos += 2; }
  if (k >= (u32)(m & 1ull)) pos += 1;
  return pos;
}

// Chunk-local item of rank r among the set bits of the chunk's masks (after
// chunk_mask_scan): returns the item's offset from the chunk start.  The word
// is the last one whose exclusive prefix is <= r (words without set bits share
// their successor's prefix, so the last such word holds the bit).
__device__ __forceinline__ u32 chunk_rank_pos(const u64* lm, const u32* loff, u32 nw, u32 r) {
  u32 a = 0, b = nw;
  while (b - a > 1) {
    const u32 mid = (a + b) >> 1;
    if (loff[mid] <= r) a = mid; else b = mid;
  }
  return a * 64u + select_bit(lm[a], r - loff[a]);
}

// 64 x 64 bit transpose across a wave: lane r holds row r on entry, lane c
// holds column c on return (bit r of the result = bit c of row r).  Six
// butterfly stages swap the off-diagonal blocks of size 32, 16, ..., 1.
__device__ __forceinline__ u64 transpose64(u64 x, u32 lane) {
  const u64 lowmask[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                          0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const u32 j = 32u >> s;
    const u64 y = __shfl_xor(x, (int)j, 64);
    const u64 lo = lowmask[s];
    x = (lane & j) ? ((x & ~lo) | ((y & ~lo) >> j)) : ((x & lo) | ((y & lo) << j));
  }
  return x;
}

// ---- k_rot_count ----------------------------------------------------------------
// Bit-sliced planes: tp[p * tstride + w] holds, for the 64 entries of word w
// (entry i = 64 w + bit), bit q of the x plane (p = q) or of the z plane
// (p = nq + q).  Words w < st->tvw are current; the scan reads only the 1-4
// planes the generator needs there (one thread per 64 entries).  Later words
// (children appended since the last scan, or everything after a compaction,
// upload or key-moving Clifford) are tested from the key planes, one lane per
// entry, and transposed into the bit-sliced planes on the way; the gather
// kernel then advances st->tvw.  Dead entries are filtered later (their
// records are flagged by k_rot_gather, which reads every candidate's
// coefficient anyway).
template <int W>
__global__ __launch_bounds__(kBlock) void k_rot_count(const u64* __restrict__ keys,
                                                      u64* __restrict__ tp, u64 tstride, u32 nq,
                                                      const DevState* __restrict__ st, u32 in_slot,
                                                      RotSpec g, u64* __restrict__ masks,
                                                      u64* __restrict__ cnt) {
  __shared__ u64 lds[kWaves];
  if (st->overflow) return;
  const u32 lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const u64 n = st->slot[in_slot].n_entries;
  const u64 tvw = st->tvw;
  const u32 G = chunk_count(n, chunk_target(n, kArenaChunk));
  const u64* pax = tp + (u64)g.a * tstride;
  const u64* paz = tp + (u64)(nq + g.a) * tstride;
  const u64* pbx = tp + (u64)g.b * tstride;
  const u64* pbz = tp + (u64)(nq + g.b) * tstride;
  for (u32 ch = blockIdx.x; ch < G; ch += gridDim.x) {
    u64 lo, hi;
    block_chunk(n, G, ch, lo, hi);
    const u64 w0 = lo >> 6, w1 = (hi + 63) >> 6;  // lo is a multiple of 256
    const u64 wv = tvw < w0 ? w0 : (tvw > w1 ? w1 : tvw);  // [w0, wv): current words
    u64 count = 0;
    // current words: entries below 64 tvw <= n are whole words, and chunk
    // boundaries other than n are multiples of 256, so no word is partial here
    for (u64 w = w0 + threadIdx.x; w < wv; w += kBlock) {
      u64 m = 0;
      if (g.mode == kBranchDamping) {
        m = ~pax[w];  // I or Z on a
      } else {
        if (g.gaz) m ^= pax[w];
        if (g.gax) m ^= paz[w];
        if (g.two) {
          if (g.gbz) m ^= pbx[w];
          if (g.gbx) m ^= pbz[w];
        }
      }
      masks[w] = m;
      count += (u64)__popcll(m);
    }
    // remaining words: test from the key planes ...
    for (u64 w = wv + wave; w < w1; w += kWaves) {
      const u64 i = w * 64 + lane;
      const bool in = i < hi;
      u64 k[2 * W];
#pragma unroll
      for (int p = 0; p < 2 * W; ++p) k[p] = 0;
      if (in) akey_load<W>(keys, i, k);
      const u64 m = __ballot(in && rot_anti<W>(k, g) != 0u);
      if (lane == 0) {
        masks[w] = m;
        count += (u64)__popcll(m);
      }
    }
    const u64 tot = block_sum_all(count, lds);
    if (threadIdx.x == 0) cnt[ch] = tot;
  }
  // ... and refresh the bit-sliced planes.  The stale words [tvw, ceil(n/64))
  // are shared out over every wave of the grid, one (group of 8 aligned words
  // = 512 entries, key plane word) item per wave step, so that every plane
  // receives whole 64-byte lines; a group that starts below tvw recomputes
  // current words from the same keys (equal values, so no conflict with
  // concurrent readers).
  {
    const u64 nwords = (n + 63) >> 6;
    const u64 waves = (u64)gridDim.x * kWaves;
    const u64 items_lo = (tvw >> 3) * (2 * W), items_hi = ((nwords + 7) >> 3) * (2 * W);
    for (u64 it = items_lo + (u64)blockIdx.x * kWaves + wave; it < items_hi; it += waves) {
      const u64 wb = (it / (2 * W)) << 3;  // group of 8 words
      const int p = (int)(it % (2 * W));   // key plane word
      u64 r[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const u64 i = (wb + (u64)j) * 64 + lane;
        r[j] = i < n ? keys[i * (2 * W) + (u64)p] : 0ull;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) r[j] = transpose64(r[j], lane);
      const u32 q = (u32)(p % W) * 64u + lane;
      if (q < nq) {
        u64* dst = tp + ((p < W ? 0ull : (u64)nq) + q) * tstride + wb;
#pragma unroll
        for (int j = 0; j < 8; ++j) dst[j] = r[j];
      }
    }
  }
}

// Hash of a record's pair key as the pairing kernels use it; kDeadHash marks a
// dead entry (a live key that hashes to it is moved one below).
constexpr u64 kDeadHash = ~0ull;
__device__ __forceinline__ u64 pair_hash(u64 h, bool dead) {
  return dead ? kDeadHash : (h == kDeadHash ? kDeadHash - 1ull : h);
}

// Facts about an anticommuting key that the emit and merge steps need (the
// WM_* bits of its worklist record): the child's sign, the damping member, and
// the weight-cap and preserve-set tests of the key and of its child.
template <int W>
__device__ __forceinline__ u32 rec_info(const u64 (&k)[2 * W], const RotSpec& g, const TruncSpec& t) {
  u32 info = 0u;
  if (rot_eps<W>(k, g) < 0.0) info |= WM_EPS_NEG;
  if (zbit<W>(k, g.a)) info |= WM_ZA;
  const bool capped = t.maxw != ~0ull;
  if (capped && weight_of<W>(k) > t.maxw) info |= WM_P_OVER;
  if (t.n_preserve && is_preserved<W>(k, t)) info |= WM_P_PRES;
  if (capped || t.n_preserve) {
    u64 ch[2 * W];
#pragma unroll
    for (int p = 0; p < 2 * W; ++p) ch[p] = k[p];
    rot_flip<W>(ch, g);
    if (capped && weight_of<W>(ch) > t.maxw) info |= WM_C_OVER;
    if (t.n_preserve && is_preserved<W>(ch, t)) info |= WM_C_PRES;
  }
  return info;
}

// ---- k_rot_gather ---------------------------------------------------------------
// Writes one 16-byte record (coefficient, arena index, facts about the key)
// per anticommuting entry into the worklist in arena order, and the hash of
// its pair key into hsh; records of dead entries are kept (flags 2, hash
// kDeadHash) so the ranks match the masks.  Block 0 also publishes the
// worklist length.
template <int W>
__global__ __launch_bounds__(kBlock) void k_rot_gather(const u64* __restrict__ keys,
                                                       const double* __restrict__ coef,
                                                       DevState* st, u32 in_slot,
                                                       const u64* __restrict__ masks,
                                                       const u64* __restrict__ cnt,
                                                       WMRec* __restrict__ wm, RotSpec g, TruncSpec t,
                                                       u64* __restrict__ hsh) {
  __shared__ u64 lds[kWaves];
  __shared__ u64 lm[kChunkMaxWords];
  __shared__ u32 loff[kChunkMaxWords];
  __shared__ u32 lt[kBlock];
  if (st->overflow) return;
  const u64 n = st->slot[in_slot].n_entries;
  const u32 G = chunk_count(n, chunk_target(n, kArenaChunk));
  if (blockIdx.x == 0) {
    const u64 tot = chunk_offset(cnt, G, lds);
    if (threadIdx.x == 0) {
      st->n_wl = tot;
      // k_rot_count refreshed the bit-sliced planes of every word from the old
      // tvw on; the last word stays stale while it is partial
      const u64 nw_full = n >> 6;
      if (nw_full > st->tvw) st->transposed += n - (st->tvw << 6);
      st->tvw = nw_full;
    }
  }
  for (u32 ch = blockIdx.x; ch < G; ch += gridDim.x) {
    if (cnt[ch] == 0) continue;  // no anticommuting entry in this chunk (block-uniform)
    u64 lo, hi;
    block_chunk(n, G, ch, lo, hi);
    const u32 nw = (u32)((hi - lo + 63) / 64);
    const u64 base_off = chunk_offset(cnt, ch, lds);
    chunk_mask_scan(masks, lo >> 6, nw, lm, loff, lt);
    // one thread per selected entry (rank order), so every lane has a load in
    // flight and consecutive threads write consecutive records
    const u32 total = nw ? loff[nw - 1] + (u32)__popcll(lm[nw - 1]) : 0u;
    for (u32 r = threadIdx.x; r < total; r += kBlock) {
      const u64 i = lo + chunk_rank_pos(lm, loff, nw, r);
      u64 k[2 * W];
      akey_load<W>(keys, i, k);
      const double c = coef[i];
      const bool dead = is_dead(c);
      store_wm(wm, base_off + r, c, (u32)i, dead ? 2u : rec_info<W>(k, g, t));
      u64 rep[2 * W];
      pair_key<W>(k, g, rep);
      hsh[base_off + r] = pair_hash(hash_key<W>(rep), dead);
    }
    __syncthreads();  // lm / loff are reused by the next chunk
  }
}

// ---- pairing ----------------------------------------------------------------------
// Both members of a pair {P, P xor G} share the pair key and hence its 64-bit
// hash h, which k_rot_gather writes per record (kDeadHash for dead entries).
// Two implementations, chosen per gate by the host from the arena size:
//  * k_rot_pair: every record inserts into one global open-addressed table
//    (slot = gen(24) | fingerprint(8) | record(32), generation tags make
//    clearing unnecessary) with device-scope CAS; cheapest for small gates.
//  * the partitioned path: the worklist is radix-partitioned on the top lg
//    bits of h into 2^lg buckets (k_pair_hist counts per (slice, bucket),
//    k_pair_colscan turns the counts into per-slice offsets, k_pair_scatter_tiled
//    places (low 32 bits of h, record) bucket-contiguously) and k_pair_lds
//    pairs each bucket in an LDS table, so global memory sees a few 8-byte
//    words per record plus the merged pairs' two records.
// Bits of h on the partitioned path: bucket = top lg bits (lg <= 14); LDS
// slot = bits 0..13; fingerprint = bits 14..25; sub-bin = bits 26..31.
constexpr u32 kPairSlices = 256;     // worklist slices of the partition kernels
constexpr u32 kPartThreads = 1024;   // threads per block of the partition kernels
constexpr u32 kMaxBucketsLg = 14;
constexpr u32 kMaxBuckets = 1u << kMaxBucketsLg;
constexpr u32 kBucketTarget = 6144;  // records per bucket aimed at (buckets average 3072-6144)
constexpr u32 kPairThreads = 1024;   // threads per block of k_pair_lds
constexpr u32 kLdsSlotsLg = 14;
constexpr u32 kLdsSlots = 1u << kLdsSlotsLg;  // 64 KB of u32 slots
constexpr u32 kPassMax = kLdsSlots / 2;       // records per LDS pass (load factor <= 1/2)
constexpr u32 kPerThread = kPassMax / kPairThreads;
constexpr u32 kSubBins = 64;
constexpr u32 kSortBins = 8192;               // digit counters of the sort-based pairing (hash bits 0..12)
constexpr u32 kSortPer = kSortBins / kPairThreads;
static_assert(kSortBins <= kPassMax && kPassMax * 8 <= kLdsSlots * 4,
              "the sort-based pairing keeps its counters in cand and its sorted entries in tab");
constexpr u32 kPosLimit = (1u << 20) - 1;     // bucket positions held by a slot (20 bits)

// Bucket count of a gate's partition: the smallest 2^lg (lg <= lg_max) whose
// buckets hold at most `target` records on average (target = kBucketTarget
// unless a test lowers it to reach the many-bucket branches).
struct PartSpec {
  u32 lg_max;
  u32 target;
};
__device__ __forceinline__ u32 pair_bucket_lg(u64 n_wl, PartSpec ps) {
  u32 lg = 0;
  while (lg < ps.lg_max && ((u64)ps.target << lg) < n_wl) ++lg;
  return lg;
}
__device__ __forceinline__ u32 bucket_of(u64 h, u32 lg) { return lg ? (u32)(h >> (64u - lg)) : 0u; }

// Exclusive prefix of v over the threads of a kPartThreads block.
__device__ __forceinline__ u32 block_excl_scan_u32(u32 v, u32* wsum) {
  const u32 lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  u32 incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const u32 y = __shfl_up(incl, o, 64);
    if (lane >= (u32)o) incl += y;
  }
  if (lane == 63u) wsum[wave] = incl;
  __syncthreads();
  u32 pre = 0;
  for (u32 w = 0; w < wave; ++w) pre += wsum[w];
  __syncthreads();
  return pre + incl - v;
}

// Records i, j whose fingerprints agree: if their pair keys (read from the
// arena through the records' indices) agree too, both merged coefficients
//   c_i' = c_i cos + eps_j c_j sin,  c_j' = c_j cos + eps_i c_i sin
// (damping: the Z member keeps c_Z cos, the I member receives c_I + c_Z sin),
// after fused truncation, go into the records themselves with the merged flag
// (one 16-byte store each); k_emit_count, which streams the worklist in arena
// order anyway, moves them into the arena.
template <int W>
__device__ __forceinline__ bool pair_try_merge(const u64* __restrict__ keys, WMRec* __restrict__ wm, u64 i,
                                               u64 j, const RotSpec& g, const TruncSpec& t, u64& kills) {
  double ci, cj;
  u32 idxi, idxj, fi, fj;
  load_wm(wm, i, ci, idxi, fi);
  load_wm(wm, j, cj, idxj, fj);
  u64 k[2 * W], kj[2 * W];
  akey_load<W>(keys, idxi, k);
  akey_load<W>(keys, idxj, kj);
  u64 rep[2 * W], repj[2 * W];
  pair_key<W>(k, g, rep);
  pair_key<W>(kj, g, repj);
  bool same = true;
#pragma unroll
  for (int p = 0; p < 2 * W; ++p) same &= (repj[p] == rep[p]);
  if (!same) return false;
  double ni, nj;
  if (g.mode == kBranchDamping) {
    const bool zi = (fi & WM_ZA) != 0u;
    const double cz = zi ? ci : cj, cI = zi ? cj : ci;
    const double nz = __dmul_rn(cz, g.cs), nI = __dadd_rn(cI, __dmul_rn(cz, g.sn));
    ni = zi ? nz : nI;
    nj = zi ? nI : nz;
  } else {
    const double ei = (fi & WM_EPS_NEG) ? -1.0 : 1.0, ej = (fj & WM_EPS_NEG) ? -1.0 : 1.0;
    ni = __dadd_rn(__dmul_rn(ci, g.cs), __dmul_rn(cj, ej * g.sn));
    nj = __dadd_rn(__dmul_rn(cj, g.cs), __dmul_rn(ci, ei * g.sn));
  }
  bool di = false, dj = false;
  if (g.trunc) {
    di = drop_from_info(ni, (fi & WM_P_OVER) != 0u, (fi & WM_P_PRES) != 0u, t);
    dj = drop_from_info(nj, (fj & WM_P_OVER) != 0u, (fj & WM_P_PRES) != 0u, t);
  }
  store_wm(wm, i, di ? dead_value() : ni, idxi, (fi & ~WM_FLAGS) | 1u);
  store_wm(wm, j, dj ? dead_value() : nj, idxj, (fj & ~WM_FLAGS) | 1u);
  kills += (u64)di + (u64)dj;
  return true;
}

// Record i (pair-key hash h) into the global generation-tagged table: claims
// an empty slot (first member of its pair so far) or merges with the member
// already there.  Returns 1 if it merged.
template <int W>
__device__ __forceinline__ u64 pair_global(const u64* __restrict__ keys, WMRec* __restrict__ wm, u64 i, u64 h,
                                           u64* table, u64 tmask, u32 tgen, const RotSpec& g,
                                           const TruncSpec& t, u64& kills) {
  const u64 hg = mix64(h ^ 0x5bd1e9955bd1e995ull);
  const u64 fp = hg >> 56;
  const u64 mine = ((u64)tgen << 40) | (fp << 32) | i;
  u64 pos = hg & tmask;
  // a plain load may return a stale (older) value; that can only make a slot
  // look empty, and the CAS then returns the value actually in memory
  u64 sv = table[pos];
  while (true) {
    if ((u32)(sv >> 40) != tgen) {
      const u64 old = cas_u64(&table[pos], sv, mine);
      if (old == sv) return 0;
      sv = old;
      continue;
    }
    if (((sv >> 32) & 0xffull) == fp && pair_try_merge<W>(keys, wm, i, sv & 0xffffffffull, g, t, kills))
      return 1;
    pos = (pos + 1) & tmask;
    sv = table[pos];
  }
}

__device__ __forceinline__ u64 pair_table_mask(u64 n_wl, u64 tcap_max) {
  u64 tcap = 64;
  while (tcap < 2 * n_wl && tcap < tcap_max) tcap <<= 1;
  return tcap - 1;
}

// ---- k_rot_pair -----------------------------------------------------------------
// Global-table pairing, one record per thread (grid-stride).
template <int W>
__global__ __launch_bounds__(kBlock) void k_rot_pair(u64 cap, const u64* __restrict__ keys,
                                                     WMRec* __restrict__ wm, DevState* st,
                                                     u32 in_slot, RotSpec g, const u64* __restrict__ hsh,
                                                     u64* table, u64 tcap_max, u32 tgen, TruncSpec t,
                                                     u64 op_tag, u64* __restrict__ part_kills,
                                                     u64* __restrict__ part_pairs) {
  __shared__ u64 lds_sum[kWaves];
  if (st->overflow) return;
  const u64 n_entries = st->slot[in_slot].n_entries;
  const u64 n_wl = st->n_wl;
  if (n_entries + n_wl > cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) cas_u64(&st->overflow, 0ull, op_tag);
    return;
  }
  const u64 tmask = pair_table_mask(n_wl, tcap_max);
  u64 kills = 0, pairs = 0;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n_wl; i += stride) {
    const u64 h = hsh[i];
    if (h != kDeadHash) pairs += pair_global<W>(keys, wm, i, h, table, tmask, tgen, g, t, kills);
  }
  const u64 bk = block_sum_u64(kills, lds_sum);
  const u64 bp = block_sum_u64(pairs, lds_sum);
  if (threadIdx.x == 0) {
    part_kills[blockIdx.x] = bk;
    part_pairs[blockIdx.x] = bp;
  }
}

// ---- k_pair_hist ----------------------------------------------------------------
// Slice s of the worklist counts its records per bucket: hist[s * B + b].
// Also raises the overflow flag when the gate's children might not fit.
__global__ __launch_bounds__(kPartThreads) void k_pair_hist(const u64* __restrict__ hsh, DevState* st,
                                                           u32 in_slot, u64 cap, PartSpec ps, u64 op_tag,
                                                           u32* __restrict__ hist) {
  __shared__ u32 h[kMaxBuckets];
  if (st->overflow) return;
  const u64 n_entries = st->slot[in_slot].n_entries;
  const u64 n_wl = st->n_wl;
  if (n_entries + n_wl > cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) cas_u64(&st->overflow, 0ull, op_tag);
    return;
  }
  const u32 lg = pair_bucket_lg(n_wl, ps), nb = 1u << lg;
  for (u32 b = threadIdx.x; b < nb; b += kPartThreads) h[b] = 0u;
  __syncthreads();
  const u64 lo = n_wl * blockIdx.x / kPairSlices, hi = n_wl * (blockIdx.x + 1) / kPairSlices;
#pragma unroll 4
  for (u64 r = lo + threadIdx.x; r < hi; r += kPartThreads) {
    const u64 v = hsh[r];
    if (v != kDeadHash) atomicAdd(&h[bucket_of(v, lg)], 1u);
  }
  __syncthreads();
  u32* row = hist + (u64)blockIdx.x * nb;
  for (u32 b = threadIdx.x; b < nb; b += kPartThreads) row[b] = h[b];
}

// ---- k_pair_colscan ---------------------------------------------------------------
// Per bucket, the exclusive prefix of the slice counts (in place) and the
// bucket size (bmeta[kMaxBuckets + b]).  A block owns 64 buckets; its 16 waves
// each take 16 consecutive slices.
constexpr u32 kScanGroups = kPartThreads / 64;
constexpr u32 kSlicesPerGroup = kPairSlices / kScanGroups;
__global__ __launch_bounds__(kPartThreads) void k_pair_colscan(const DevState* __restrict__ st, u32 in_slot,
                                                              u64 cap, PartSpec ps, u32* __restrict__ hist,
                                                              u32* __restrict__ bmeta) {
  __shared__ u32 gsum[kScanGroups][64];
  if (st->overflow) return;
  const u64 n_entries = st->slot[in_slot].n_entries;
  const u64 n_wl = st->n_wl;
  if (n_entries + n_wl > cap) return;
  const u32 lg = pair_bucket_lg(n_wl, ps), nb = 1u << lg;
  const u32 b0 = blockIdx.x * 64u;
  if (b0 >= nb) return;
  const u32 lane = threadIdx.x & 63u, grp = threadIdx.x >> 6;
  const u32 b = b0 + lane;
  const bool on = b < nb;
  const u32 s0 = grp * kSlicesPerGroup;
  u32 v[kSlicesPerGroup], sum = 0;
#pragma unroll
  for (u32 q = 0; q < kSlicesPerGroup; ++q) {
    v[q] = on ? hist[(u64)(s0 + q) * nb + b] : 0u;
    sum += v[q];
  }
  gsum[grp][lane] = sum;
  __syncthreads();
  u32 run = 0;
  for (u32 q = 0; q < grp; ++q) run += gsum[q][lane];
  if (on) {
#pragma unroll
    for (u32 q = 0; q < kSlicesPerGroup; ++q) {
      hist[(u64)(s0 + q) * nb + b] = run;
      run += v[q];
    }
    if (grp == kScanGroups - 1u) bmeta[kMaxBuckets + b] = run;
  }
}

// ---- per-record scatter (more than 2^kTiledLgMax buckets) --------------------------
// Slice s writes bkt[pos] = record | (low 32 bits of h) << 32, at its offset
// in the record's bucket: bucket base (prefix of the bucket sizes) plus the
// counts of the slices before s.  Block 0 publishes the bucket bases in
// bmeta[0..B).  `off` is LDS for kMaxBuckets counters.
__device__ __forceinline__ void pair_scatter_records(const u64* __restrict__ hsh, u64 n_wl, u32 lg,
                                                     const u32* __restrict__ hist, u64* __restrict__ bkt,
                                                     u32* __restrict__ bmeta, u32* off, u32* wsum) {
  const u32 nb = 1u << lg;
  const u32 s = blockIdx.x;
  // bucket bases: exclusive scan of the sizes, kPer consecutive buckets per thread
  constexpr u32 kPer = kMaxBuckets / kPartThreads;
  const u32 b0 = threadIdx.x * kPer;
  const u32* row = hist + (u64)s * nb;
  u32 loc[kPer], pre[kPer], sum = 0;
#pragma unroll
  for (u32 q = 0; q < kPer; ++q) {
    const bool on = b0 + q < nb;
    loc[q] = on ? bmeta[kMaxBuckets + b0 + q] : 0u;
    pre[q] = on ? row[b0 + q] : 0u;
    sum += loc[q];
  }
  u32 base = block_excl_scan_u32(sum, wsum);
#pragma unroll
  for (u32 q = 0; q < kPer; ++q) {
    if (b0 + q < nb) {
      if (s == 0) bmeta[b0 + q] = base;
      off[b0 + q] = base + pre[q];
    }
    base += loc[q];
  }
  __syncthreads();
  const u64 lo = n_wl * s / kPairSlices, hi = n_wl * (s + 1) / kPairSlices;
#pragma unroll 4
  for (u64 r = lo + threadIdx.x; r < hi; r += kPartThreads) {
    const u64 v = hsh[r];
    if (v == kDeadHash) continue;
    const u32 pos = atomicAdd(&off[bucket_of(v, lg)], 1u);
    bkt[pos] = r | ((u64)(u32)v << 32);
  }
}

// ---- k_pair_scatter_tiled ---------------------------------------------------------
// The scatter for up to 2^kTiledLgMax buckets: slice s takes its records in
// tiles of kScatterTile, sorts each tile by bucket in LDS (local counts, scan,
// placement) and then writes the tile in that order, so consecutive threads
// store consecutive addresses of one bucket's run instead of one 8-byte store
// per bucket line (the per-record scatter, used for more buckets, does the
// latter).
// Results are the same multiset per bucket; the order inside a bucket does
// not matter to k_pair_lds.
constexpr u32 kTiledLgMax = 12;
constexpr u32 kScatterTile = 8192;
static_assert(kScatterTile * 2 >= kMaxBuckets, "the tile buffer holds the per-record scatter's counters");
constexpr u32 kTilePer = kScatterTile / kPartThreads;
constexpr u32 kTiledBuckets = 1u << kTiledLgMax;
constexpr u32 kTiledPer = kTiledBuckets / kPartThreads;
__global__ __launch_bounds__(kPartThreads) void k_pair_scatter_tiled(const u64* __restrict__ hsh,
                                                                    const DevState* __restrict__ st, u32 in_slot,
                                                                    u64 cap, PartSpec ps,
                                                                    const u32* __restrict__ hist,
                                                                    u64* __restrict__ bkt, u32* __restrict__ bmeta) {
  __shared__ u32 off[kTiledBuckets];     // next global position per bucket (this slice)
  __shared__ u32 tcnt[kTiledBuckets];    // records per bucket in the tile
  __shared__ u32 tstart[kTiledBuckets];  // tile position of each bucket's first record
  __shared__ u64 buf[kScatterTile];      // the tile in bucket order
  __shared__ unsigned short bid[kScatterTile];
  __shared__ u32 wsum[kPartThreads / 64];
  __shared__ u32 ttot;
  if (st->overflow) return;
  const u64 n_entries = st->slot[in_slot].n_entries;
  const u64 n_wl = st->n_wl;
  if (n_entries + n_wl > cap) return;
  const u32 lg = pair_bucket_lg(n_wl, ps), nb = 1u << lg;
  if (lg > kTiledLgMax) {
    // too many buckets for the LDS tile: one 4-byte counter per bucket in
    // the tile buffer (kScatterTile u64 = kMaxBuckets u32) and a scattered
    // store per record
    pair_scatter_records(hsh, n_wl, lg, hist, bkt, bmeta, reinterpret_cast<u32*>(buf), wsum);
    return;
  }
  const u32 s = blockIdx.x, tid = threadIdx.x;
  // bucket bases (exclusive scan of the sizes; block 0 publishes them) and this
  // slice's offset inside each bucket
  {
    const u32 b0 = tid * kTiledPer;
    const u32* row = hist + (u64)s * nb;
    u32 loc[kTiledPer], pre[kTiledPer], sum = 0;
#pragma unroll
    for (u32 q = 0; q < kTiledPer; ++q) {
      const bool on = b0 + q < nb;
      loc[q] = on ? bmeta[kMaxBuckets + b0 + q] : 0u;
      pre[q] = on ? row[b0 + q] : 0u;
      sum += loc[q];
    }
    u32 base = block_excl_scan_u32(sum, wsum);
#pragma unroll
    for (u32 q = 0; q < kTiledPer; ++q) {
      if (b0 + q < nb) {
        if (s == 0) bmeta[b0 + q] = base;
        off[b0 + q] = base + pre[q];
      }
      base += loc[q];
    }
  }
  const u64 lo = n_wl * s / kPairSlices, hi = n_wl * (s + 1) / kPairSlices;
  for (u64 t0 = lo; t0 < hi; t0 += kScatterTile) {
    const u32 tn = (u32)min((u64)kScatterTile, hi - t0);
    for (u32 b = tid; b < nb; b += kPartThreads) tcnt[b] = 0u;
    __syncthreads();
    u64 myv[kTilePer];
    u32 myb[kTilePer], myr[kTilePer];
#pragma unroll
    for (u32 q = 0; q < kTilePer; ++q) {
      const u32 e = tid + q * kPartThreads;
      myb[q] = 0xffffffffu;
      if (e < tn) {
        const u64 v = hsh[t0 + e];
        if (v != kDeadHash) {
          const u32 b = bucket_of(v, lg);
          myb[q] = b;
          myr[q] = atomicAdd(&tcnt[b], 1u);
          myv[q] = (t0 + e) | ((u64)(u32)v << 32);
        }
      }
    }
    __syncthreads();
    {  // tile positions: exclusive scan of the tile counts
      const u32 b0 = tid * kTiledPer;
      u32 c[kTiledPer], sum = 0;
#pragma unroll
      for (u32 q = 0; q < kTiledPer; ++q) {
        c[q] = b0 + q < nb ? tcnt[b0 + q] : 0u;
        sum += c[q];
      }
      u32 base = block_excl_scan_u32(sum, wsum);
#pragma unroll
      for (u32 q = 0; q < kTiledPer; ++q) {
        if (b0 + q < nb) tstart[b0 + q] = base;
        base += c[q];
      }
      if (tid == kPartThreads - 1u) ttot = base;
    }
    __syncthreads();
#pragma unroll
    for (u32 q = 0; q < kTilePer; ++q) {
      if (myb[q] != 0xffffffffu) {
        const u32 p = tstart[myb[q]] + myr[q];
        buf[p] = myv[q];
        bid[p] = (unsigned short)myb[q];
      }
    }
    __syncthreads();
    const u32 tot = ttot;
    for (u32 p = tid; p < tot; p += kPartThreads) {
      const u32 b = bid[p];
      bkt[(u64)off[b] + (p - tstart[b])] = buf[p];
    }
    __syncthreads();
    for (u32 b = tid; b < nb; b += kPartThreads) off[b] += tcnt[b];
    __syncthreads();
  }
}

// ---- k_pair_lds -----------------------------------------------------------------
// One block per bucket (blocks loop over buckets).  A bucket of at most
// kPassMax records is paired in one pass: (A) every thread loads its records'
// bucket entries; (B) each record walks its linear probe sequence in the LDS
// table and claims the first empty slot (slot = bucket position + 1 |
// fingerprint << 20) or stops at the first slot with its fingerprint, which
// makes it a candidate; (C) candidates are spread over the threads, load both
// records and merge if the pair keys agree; (D) a candidate whose keys differ
// (fingerprint collision) walks again from its home slot, testing every
// fingerprint match, until it merges or claims a slot.  Claimed slots are
// never released within a pass, so the probe sequence of a pair's second
// member always passes the first member's slot.  Larger buckets are paired
// in several passes over ranges of sub-bins; a single sub-bin that is still
// too large, or a bucket whose positions exceed 20 bits, uses the global
// table.  Pair members always share bucket and sub-bin.
template <int W>
__device__ __forceinline__ u64 lds_probe_sync(const u64* __restrict__ keys, WMRec* __restrict__ wm,
                                              u32* tab, const u64* __restrict__ bkt, u32 base, u32 e, u64 v,
                                              const RotSpec& g, const TruncSpec& t, u64& kills) {
  const u32 lo = (u32)(v >> 32);
  const u32 i = (u32)v;
  const u32 fp = (lo >> kLdsSlotsLg) & 0xfffu;
  const u32 mine = (e + 1u) | (fp << 20);
  u32 pos = lo & (kLdsSlots - 1u);
  while (true) {
    u32 sv = __hip_atomic_load(&tab[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (sv == 0u) {
      u32 exp = 0u;
      if (__hip_atomic_compare_exchange_strong(&tab[pos], &exp, mine, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP))
        return 0;
      sv = exp;
    }
    if ((sv >> 20) == fp && (sv & kPosLimit) != e + 1u) {
      const u32 j = (u32)bkt[base + (sv & kPosLimit) - 1u];
      if (pair_try_merge<W>(keys, wm, i, j, g, t, kills)) return 1;
    }
    pos = (pos + 1u) & (kLdsSlots - 1u);
  }
}

template <int W, bool kSort>
__global__ __launch_bounds__(kPairThreads) void k_pair_lds(u64 cap, const u64* __restrict__ keys,
                                                           WMRec* __restrict__ wm, DevState* st,
                                                           u32 in_slot, RotSpec g, PartSpec ps,
                                                           const u64* __restrict__ hsh,
                                                           const u64* __restrict__ bkt,
                                                           const u32* __restrict__ bmeta, u64* table,
                                                           u64 tcap_max, u32 tgen, TruncSpec t,
                                                           u64* __restrict__ part_kills,
                                                           u64* __restrict__ part_pairs) {
  __shared__ __align__(16) u32 tab[kLdsSlots];
  __shared__ u32 cand[kPassMax];
  __shared__ u32 lrec[kPassMax];  // record index of each bucket position (single pass)
  __shared__ u32 sub[kSubBins];
  __shared__ u32 plo[kSubBins], phi[kSubBins], pfb[kSubBins];
  __shared__ u32 npass, ncand;
  __shared__ u64 lds_sum[kPairThreads / 64];
  if (st->overflow) return;
  const u64 n_entries = st->slot[in_slot].n_entries;
  const u64 n_wl = st->n_wl;
  if (n_entries + n_wl > cap) return;
  const u32 lg = pair_bucket_lg(n_wl, ps), nb = 1u << lg;
  const u64 tmask = pair_table_mask(n_wl, tcap_max);
  const u32 tid = threadIdx.x;
  uint4* t4 = reinterpret_cast<uint4*>(tab);
  u64 kills = 0, pairs = 0, fallback = 0;
  for (u32 b = blockIdx.x; b < nb; b += gridDim.x) {
    const u32 base = bmeta[b], T = bmeta[kMaxBuckets + b];
    if (T == 0u) continue;
    if (kSort && T <= kPassMax) {
      // ---- single pass, sort-based: counting sort of the bucket on hash bits
      // 0..12 in LDS (cand holds the 8192 digit counters, tab the sorted
      // entries), then each run of equal digits is searched for equal 32-bit
      // hashes, confirmed on the keys and merged (a key has at most one
      // partner, so a record merges at most once).
      static_assert(kPairThreads == kPartThreads, "block_excl_scan_u32 scans kPartThreads threads");
      u32* cnt = cand;
      u64* srt = reinterpret_cast<u64*>(tab);
      for (u32 e = tid; e < kSortBins; e += kPairThreads) cnt[e] = 0u;
      __syncthreads();
      u64 v[kPerThread];
      u32 rk[kPerThread];
#pragma unroll
      for (u32 q = 0; q < kPerThread; ++q) {
        const u32 e = tid + q * kPairThreads;
        v[q] = e < T ? bkt[base + e] : 0ull;
        rk[q] = e < T ? atomicAdd(&cnt[(u32)(v[q] >> 32) & (kSortBins - 1u)], 1u) : 0u;
      }
      __syncthreads();
      {  // digit starts: exclusive scan of the counters, kSortPer per thread
        const u32 b0 = tid * kSortPer;
        u32 c[kSortPer], sum = 0;
#pragma unroll
        for (u32 q = 0; q < kSortPer; ++q) {
          c[q] = cnt[b0 + q];
          sum += c[q];
        }
        u32 run = block_excl_scan_u32(sum, reinterpret_cast<u32*>(lds_sum));
#pragma unroll
        for (u32 q = 0; q < kSortPer; ++q) {
          cnt[b0 + q] = run;
          run += c[q];
        }
      }
      __syncthreads();
#pragma unroll
      for (u32 q = 0; q < kPerThread; ++q) {
        const u32 e = tid + q * kPairThreads;
        if (e < T) srt[cnt[(u32)(v[q] >> 32) & (kSortBins - 1u)] + rk[q]] = v[q];
      }
      __syncthreads();
      for (u32 p = tid; p < T; p += kPairThreads) {
        const u32 d = (u32)(srt[p] >> 32) & (kSortBins - 1u);
        if (p > 0 && ((u32)(srt[p - 1] >> 32) & (kSortBins - 1u)) == d) continue;  // not a run start
        u32 end = p + 1;
        while (end < T && ((u32)(srt[end] >> 32) & (kSortBins - 1u)) == d) ++end;
        if (end - p > 32u) {  // a run beyond the merged-member mask: pair it through the global table
          for (u32 a = p; a < end; ++a) {
            const u32 i = (u32)srt[a];
            pairs += pair_global<W>(keys, wm, i, hsh[i], table, tmask, tgen, g, t, kills);
            ++fallback;
          }
          continue;
        }
        u32 done = 0;  // run members already merged (runs are short: the bucket has ~T/8192 per digit)
        for (u32 a = p; a + 1 < end; ++a) {
          if ((done >> (a - p)) & 1u) continue;
          for (u32 b = a + 1; b < end; ++b) {
            if ((done >> (b - p)) & 1u) continue;
            if ((u32)(srt[a] >> 32) != (u32)(srt[b] >> 32)) continue;
            if (pair_try_merge<W>(keys, wm, (u32)srt[a], (u32)srt[b], g, t, kills)) {
              ++pairs;
              done |= 1u << (b - p);
              break;
            }
          }
        }
      }
      __syncthreads();  // counters and sorted entries are rewritten for the next bucket
      continue;
    }
    if (T <= kPassMax) {
      // ---- single pass
      for (u32 e = tid; e < kLdsSlots / 4; e += kPairThreads) t4[e] = make_uint4(0u, 0u, 0u, 0u);
      if (tid == 0) ncand = 0u;
      u64 v[kPerThread];
#pragma unroll
      for (u32 q = 0; q < kPerThread; ++q) {
        const u32 e = tid + q * kPairThreads;
        v[q] = e < T ? bkt[base + e] : 0ull;
        if (e < T) lrec[e] = (u32)v[q];
      }
      __syncthreads();
#pragma unroll
      for (u32 q = 0; q < kPerThread; ++q) {
        const u32 e = tid + q * kPairThreads;
        if (e < T) {
          const u32 lo = (u32)(v[q] >> 32);
          const u32 fp = (lo >> kLdsSlotsLg) & 0xfffu;
          const u32 mine = (e + 1u) | (fp << 20);
          u32 pos = lo & (kLdsSlots - 1u);
          while (true) {
            u32 sv = __hip_atomic_load(&tab[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (sv == 0u) {
              u32 exp = 0u;
              if (__hip_atomic_compare_exchange_strong(&tab[pos], &exp, mine, __ATOMIC_RELAXED,
                                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
                break;  // claimed
              sv = exp;
            }
            if ((sv >> 20) == fp) {  // candidate partner
              const u32 c = atomicAdd(&ncand, 1u);
              cand[c] = e | (((sv & kPosLimit) - 1u) << 16);
              break;
            }
            pos = (pos + 1u) & (kLdsSlots - 1u);
          }
        }
      }
      __syncthreads();
      const u32 nc = ncand;
      bool retry_any = false;
      for (u32 c = tid; c < nc; c += kPairThreads) {
        const u32 ce = cand[c] & 0xffffu, cj = cand[c] >> 16;
        const u32 i = lrec[ce], j = lrec[cj];
        if (pair_try_merge<W>(keys, wm, i, j, g, t, kills)) {
          ++pairs;
        } else {
          cand[c] |= 0x80000000u;  // fingerprint collision: walk again below
          retry_any = true;
        }
      }
      if (__syncthreads_or(retry_any)) {
        for (u32 c = tid; c < nc; c += kPairThreads) {
          if (!(cand[c] & 0x80000000u)) continue;
          const u32 ce = cand[c] & 0xffffu;
          pairs += lds_probe_sync<W>(keys, wm, tab, bkt, base, ce, bkt[base + ce], g, t, kills);
        }
      }
      __syncthreads();  // table and candidate list are reset for the next bucket
      continue;
    }
    // ---- several passes (large bucket)
    if (T > kPosLimit) {
      if (tid == 0) {
        plo[0] = 0u;
        phi[0] = kSubBins;
        pfb[0] = 1u;
        npass = 1u;
      }
    } else {
      if (tid < kSubBins) sub[tid] = 0u;
      __syncthreads();
      for (u32 e = tid; e < T; e += kPairThreads) atomicAdd(&sub[(u32)(bkt[base + e] >> 58)], 1u);
      __syncthreads();
      if (tid == 0) {
        u32 np = 0, start = 0, acc = 0;
        for (u32 q = 0; q < kSubBins; ++q) {
          const u32 c = sub[q];
          if (acc > 0u && acc + c > kPassMax) {
            plo[np] = start;
            phi[np] = q;
            pfb[np] = acc > kPassMax ? 1u : 0u;
            ++np;
            start = q;
            acc = 0u;
          }
          acc += c;
        }
        plo[np] = start;
        phi[np] = kSubBins;
        pfb[np] = acc > kPassMax ? 1u : 0u;
        npass = np + 1u;
      }
    }
    __syncthreads();
    const u32 np = npass;
    for (u32 q = 0; q < np; ++q) {
      const u32 s0 = plo[q], s1 = phi[q];
      if (!pfb[q]) {
        for (u32 e = tid; e < kLdsSlots / 4; e += kPairThreads) t4[e] = make_uint4(0u, 0u, 0u, 0u);
        __syncthreads();
        for (u32 e = tid; e < T; e += kPairThreads) {
          const u64 v = bkt[base + e];
          const u32 sb = (u32)(v >> 58);
          if (sb >= s0 && sb < s1) pairs += lds_probe_sync<W>(keys, wm, tab, bkt, base, e, v, g, t, kills);
        }
        __syncthreads();  // the table is cleared again by the next pass
      } else {
        for (u32 e = tid; e < T; e += kPairThreads) {
          const u64 v = bkt[base + e];
          const u32 sb = (u32)(v >> 58);
          if (sb >= s0 && sb < s1) {
            const u32 i = (u32)v;
            pairs += pair_global<W>(keys, wm, i, hsh[i], table, tmask, tgen, g, t, kills);
            ++fallback;
          }
        }
      }
    }
    __syncthreads();  // plan arrays are rewritten for the next bucket
  }
  if (fallback) add_u64(&st->pair_fallback, fallback);
  // deterministic block sums
  for (int o = 32; o > 0; o >>= 1) {
    kills += __shfl_down(kills, o, 64);
    pairs += __shfl_down(pairs, o, 64);
  }
  __syncthreads();
  if ((tid & 63u) == 0u) lds_sum[tid >> 6] = kills;
  __syncthreads();
  if (tid == 0) {
    u64 s = 0;
    for (u32 w = 0; w < kPairThreads / 64; ++w) s += lds_sum[w];
    part_kills[blockIdx.x] = s;
  }
  __syncthreads();
  if ((tid & 63u) == 0u) lds_sum[tid >> 6] = pairs;
  __syncthreads();
  if (tid == 0) {
    u64 s = 0;
    for (u32 w = 0; w < kPairThreads / 64; ++w) s += lds_sum[w];
    part_pairs[blockIdx.x] = s;
  }
}

// ---- k_emit_count -----------------------------------------------------------------
// Streams the 16-byte records: unpaired records scale their parent in place
// and decide whether their child survives the fused truncation (from the
// record's key facts, without the key); merged records move their final value
// into the arena; masks + counts of the surviving children.
template <int W>
__global__ __launch_bounds__(kBlock) void k_emit_count(double* __restrict__ coef, u64 cap,
                                                       const WMRec* __restrict__ wm,
                                                       const DevState* __restrict__ st, u32 in_slot,
                                                       RotSpec g, TruncSpec t, u64* __restrict__ masks,
                                                       u64* __restrict__ cnt, u64* __restrict__ kills_out) {
  __shared__ u64 lds[kWaves];
  if (st->overflow) return;
  const u64 n_entries = st->slot[in_slot].n_entries;
  const u64 n_wl = st->n_wl;
  if (n_entries + n_wl > cap) return;
  const u32 G = chunk_count(n_wl, chunk_target(n_wl, kListChunk));
  const u32 lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  for (u32 ch = blockIdx.x; ch < G; ch += gridDim.x) {
    u64 lo, hi;
    block_chunk(n_wl, G, ch, lo, hi);
    const u32 nw = (u32)((hi - lo + 63) / 64);
    u64 count = 0, kills = 0;
    for (u32 j = wave; j < nw; j += kWaves) {
      const u64 w = lo + (u64)j * 64 + lane;
      bool keep = false;
      double c = 0.0;
      u32 idx = 0, info = 2u;  // lanes past the chunk end act as dead records
      if (w < hi) load_wm(wm, w, c, idx, info);
      const u32 fl = info & WM_FLAGS;
      if (fl == 0u) {
        // unpaired: the parent is scaled in place and its child survives
        // unless truncated (damping: only Z members branch, I members stay)
        const bool branches = g.mode != kBranchDamping || (info & WM_ZA) != 0u;
        const double eps = (info & WM_EPS_NEG) ? -1.0 : 1.0;
        const double cp = branches ? __dmul_rn(c, g.cs) : c;
        const double cc = __dmul_rn(c, eps * g.sn);
        bool dp = false;
        keep = branches;
        if (g.trunc) {
          dp = drop_from_info(cp, (info & WM_P_OVER) != 0u, (info & WM_P_PRES) != 0u, t);
          keep = branches && !drop_from_info(cc, (info & WM_C_OVER) != 0u, (info & WM_C_PRES) != 0u, t);
        }
        if (dp || branches) coef[idx] = dp ? dead_value() : cp;
        kills += (u64)dp;
      } else if (fl == 1u) {
        coef[idx] = c;  // merged by the pairing kernel (value final, truncation applied)
      }
      const u64 m = __ballot(keep);
      if (lane == 0) masks[(lo >> 6) + j] = m;
      count += (u64)__popcll(m);
    }
    const u64 tc = block_sum_all(lane == 0 ? count : 0ull, lds);  // count is wave-uniform
    const u64 tk = block_sum_all(kills, lds);
    if (threadIdx.x == 0) {
      cnt[ch] = tc;
      kills_out[ch] = tk;
    }
  }
}

// ---- k_emit_write -----------------------------------------------------------------
// Appends the surviving children (the parent's key read from the arena, the
// generator's bits flipped); block 0 also publishes the op's new counts and
// statistics.
template <int W>
__global__ __launch_bounds__(kBlock) void k_emit_write(u64* __restrict__ keys, double* __restrict__ coef,
                                                       u64 cap, const WMRec* __restrict__ wm,
                                                       DevState* st, u32 in_slot, u32 out_slot,
                                                       RotSpec g, const u64* __restrict__ masks,
                                                       const u64* __restrict__ cnt,
                                                       const u64* __restrict__ ekills,
                                                       const u64* __restrict__ pkills,
                                                       const u64* __restrict__ ppairs, u32 Gp) {
  __shared__ u64 lds[kWaves];
  __shared__ u64 lds4[4 * kWaves];
  __shared__ u64 lm[kChunkMaxWords];
  __shared__ u32 loff[kChunkMaxWords];
  __shared__ u32 lt[kBlock];
  if (st->overflow) return;
  const CountSlot in = st->slot[in_slot];
  const u64 n_entries = in.n_entries;
  const u64 n_wl = st->n_wl;
  if (n_entries + n_wl > cap) return;
  const u32 G = chunk_count(n_wl, chunk_target(n_wl, kListChunk));
  if (blockIdx.x == 0) {
    u64 sums[4];
    block_sums4(cnt, G, ekills, G, pkills, Gp, ppairs, Gp, lds4, sums);
    const u64 appended = sums[0], kills = sums[1] + sums[2], pairs = sums[3];
    if (threadIdx.x == 0) {
      st->slot[out_slot].n_entries = n_entries + appended;
      st->slot[out_slot].n_dead = in.n_dead + kills;
      st->term_updates += n_entries - in.n_dead;
      st->gates += 1;
      st->rotations += 1;
      st->scanned += n_entries;
      st->scan_planes += n_entries * g.test_planes;
      st->wl_total += n_wl;
      st->pairs_total += pairs;
      st->children_total += appended;
      st->kills_total += kills;
    }
  }
  for (u32 ch = blockIdx.x; ch < G; ch += gridDim.x) {
    u64 lo, hi;
    block_chunk(n_wl, G, ch, lo, hi);
    const u32 nw = (u32)((hi - lo + 63) / 64);
    const u64 base_off = n_entries + chunk_offset(cnt, ch, lds);
    chunk_mask_scan(masks, lo >> 6, nw, lm, loff, lt);
    // one thread per surviving child (rank order): dense loads, coalesced stores
    const u32 total = nw ? loff[nw - 1] + (u32)__popcll(lm[nw - 1]) : 0u;
    for (u32 r = threadIdx.x; r < total; r += kBlock) {
      const u64 w = lo + chunk_rank_pos(lm, loff, nw, r);
      double c;
      u32 idx, info;
      load_wm(wm, w, c, idx, info);
      u64 k[2 * W];
      akey_load<W>(keys, idx, k);  // the parent (below n_entries; children go above)
      const double eps = (info & WM_EPS_NEG) ? -1.0 : 1.0;
      const double cc = __dmul_rn(c, eps * g.sn);
      rot_flip<W>(k, g);
      const u64 dst = base_off + r;
      akey_store<W>(keys, dst, k);
      coef[dst] = cc;
    }
    __syncthreads();  // lm / loff are reused by the next chunk
  }
}

// ---- k_support: union of the live terms' x and z words --------------------------
// Per block, the OR of every live entry's 2W key words, written to
// out[block * 2W + p] for the host to OR, or (out null) OR-ed into st->sup,
// which a k_fill_u64 cleared before.  Lets the host judge exactly which
// rotations can act on some term (see k_skip_gates).
template <int W>
__global__ __launch_bounds__(kBlock) void k_support(const u64* __restrict__ keys,
                                                    const double* __restrict__ coef,
                                                    DevState* __restrict__ st, u32 in_slot,
                                                    u64* __restrict__ out) {
  __shared__ u64 lds[2 * W][kWaves];
  const u64 n = st->slot[in_slot].n_entries;
  u64 acc[2 * W];
#pragma unroll
  for (int p = 0; p < 2 * W; ++p) acc[p] = 0;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    if (is_dead(coef[i])) continue;
#pragma unroll
    for (int p = 0; p < 2 * W; ++p) acc[p] |= keys[i * (2 * W) + (u64)p];
  }
  const u32 lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
  for (int p = 0; p < 2 * W; ++p) {
    for (int o = 32; o > 0; o >>= 1) acc[p] |= (u64)__shfl_xor((long long)acc[p], o, 64);
    if (lane == 0) lds[p][wave] = acc[p];
  }
  __syncthreads();
  if (threadIdx.x < 2 * W) {
    u64 v = 0;
    for (int w = 0; w < kWaves; ++w) v |= lds[threadIdx.x][w];
    if (out) {
      out[(u64)blockIdx.x * (2 * W) + threadIdx.x] = v;
    } else if (v) {
      atomicOr(reinterpret_cast<unsigned long long*>(&st->sup[threadIdx.x]), (unsigned long long)v);
    }
  }
}

// ---- k_fingerprint: order-independent digest of the live terms ------------------
// Per live entry, k = fmix(...fmix(fmix(seed ^ x_0) ^ x_1)... ^ z_{wa-1}) over the
// wa ABI words of each plane, kc = fmix(k ^ coefficient bits); per block the
// wrap-around sums of k and kc go to out[2 block], out[2 block + 1] (the host
// adds the blocks).  Integer sums, so the result does not depend on the order.
__device__ __forceinline__ u64 fp_fmix(u64 v) {
  v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
  v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
  return v ^ (v >> 31);
}
template <int W>
__global__ __launch_bounds__(kBlock) void k_fingerprint(const u64* __restrict__ keys,
                                                        const double* __restrict__ coef,
                                                        const DevState* __restrict__ st, u32 in_slot, u32 wa,
                                                        u64* __restrict__ out) {
  __shared__ u64 lds[kWaves];
  const u64 n = st->slot[in_slot].n_entries;
  u64 ks = 0, kcs = 0;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double c = coef[i];
    if (is_dead(c)) continue;
    u64 v = 0x9E3779B97F4A7C15ull;
    for (u32 p = 0; p < wa; ++p) v = fp_fmix(v ^ keys[i * (2 * W) + p]);
    for (u32 p = 0; p < wa; ++p) v = fp_fmix(v ^ keys[i * (2 * W) + W + p]);
    ks += v;
    kcs += fp_fmix(v ^ (u64)__double_as_longlong(c));
  }
  const u64 a = block_sum_all(ks, lds);
  const u64 b = block_sum_all(kcs, lds);
  if (threadIdx.x == 0) {
    out[2 * (u64)blockIdx.x] = a;
    out[2 * (u64)blockIdx.x + 1] = b;
  }
}

// ---- k_skip_gates: gates that act on no term -------------------------------------
// The host drops a run of rotations whose generators anticommute with no Pauli
// present anywhere in the sum, and of map ops on qubits where every term holds
// the identity (judged from the union of the terms' x and z supports); such
// gates leave every term and coefficient unchanged.  This kernel carries the
// counts over to the next op slot and books the term updates the gates (n_rot
// of them rotations) represent.
__global__ void k_skip_gates(DevState* st, u32 in_slot, u32 out_slot, u32 n_gates, u32 n_rot) {
  if (threadIdx.x != 0 || st->overflow) return;
  const CountSlot in = st->slot[in_slot];
  st->slot[out_slot] = in;
  st->term_updates += (u64)n_gates * (in.n_entries - in.n_dead);
  st->gates += n_gates;
  st->rotations += n_rot;
}

// ---- preserved strings that a truncation would drop ---------------------------------
// The reference's truncation removes every entry its strategy drops and then
// re-inserts the preserved ones among them, so those leave their place and
// follow every other entry, in their previous order (sum/data.rs:283-310).
// Truncation here keeps such entries where they are; after each truncation
// k_pres_find lists them (found[0] = count, found[1..] = arena indices, the
// host clears the count first) and k_pres_move appends a copy of each at the
// end of the arena, in index order, and marks the original dead.
constexpr u32 kPresMoveThreads = 256;

template <int W>
__global__ __launch_bounds__(kBlock) void k_pres_find(const u64* __restrict__ keys,
                                                      const double* __restrict__ coef,
                                                      const DevState* __restrict__ st, u32 in_slot,
                                                      TruncSpec t, u64* __restrict__ found) {
  if (st->overflow) return;
  const u64 n = st->slot[in_slot].n_entries;
  const bool wcap = t.maxw != ~0ull;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double c = coef[i];
    if (is_dead(c)) continue;
    const bool small = fabs(c) < t.tau;
    if (!small && !wcap) continue;
    u64 k[2 * W];
    akey_load<W>(keys, i, k);
    if (!small && weight_of<W>(k) <= t.maxw) continue;
    if (!is_preserved<W>(k, t)) continue;
    const u64 slot = atomicAdd((unsigned long long*)found, 1ull);
    if (slot < t.n_preserve) found[1 + slot] = i;  // live keys are distinct: slot < n_preserve
  }
}

template <int W>
__global__ __launch_bounds__(kPresMoveThreads) void k_pres_move(u64* __restrict__ keys, double* __restrict__ coef,
                                                                DevState* st, u32 in_slot, u32 out_slot, u64 cap,
                                                                u32 n_preserve, const u64* __restrict__ found,
                                                                u64 op_tag) {
  if (st->overflow) return;
  const CountSlot in = st->slot[in_slot];
  const u64 m = found[0] < (u64)n_preserve ? found[0] : (u64)n_preserve;
  if (in.n_entries + m > cap) {  // nothing written: the host grows the arena and replays this op
    if (threadIdx.x == 0) cas_u64(&st->overflow, 0ull, op_tag);
    return;
  }
  const u64* idx = found + 1;
  for (u64 j = threadIdx.x; j < m; j += kPresMoveThreads) {
    const u64 src = idx[j];
    u64 r = 0;  // rank of src among the listed indices
    for (u64 q = 0; q < m; ++q) r += idx[q] < src ? 1u : 0u;
    u64 k[2 * W];
    akey_load<W>(keys, src, k);
    akey_store<W>(keys, in.n_entries + r, k);
    coef[in.n_entries + r] = coef[src];
    coef[src] = dead_value();
  }
  if (threadIdx.x == 0) {
    CountSlot o = in;
    o.n_entries = in.n_entries + m;
    o.n_dead = in.n_dead + m;
    st->slot[out_slot] = o;
  }
}

// ---- k_finish: single-block reduction / bookkeeping ---------------------------
enum FinishMode : u32 { F_MAP = 0, F_OVERLAP = 1 };
struct FinishArgs {
  u32 mode;
  u32 in_slot, out_slot;
  u32 Ga;              // partial-sum entries to reduce
  u32 n_ops;           // map program length (F_MAP)
  const u64* aux0;     // kills (F_MAP)
  const u64* aux1;     // term updates (F_MAP)
  const double* dpart; // overlap partials (F_OVERLAP)
};

__global__ __launch_bounds__(kScanThreads) void k_finish(DevState* st, FinishArgs a) {
  __shared__ u64 lds[kScanThreads / 64];
  __shared__ double lds_d[kScanThreads];
  const u32 tid = threadIdx.x;
  if (a.mode == F_OVERLAP) {
    // ordered sum of per-block partials (deterministic)
    double v = 0.0;
    const u32 per = (a.Ga + kScanThreads - 1) / kScanThreads;
    for (u32 j = tid * per; j < (tid + 1) * per && j < a.Ga; ++j) v = __dadd_rn(v, a.dpart[j]);
    lds_d[tid] = v;
    __syncthreads();
    if (tid == 0) {
      double tot = 0.0;
      for (u32 j = 0; j < kScanThreads; ++j) tot = __dadd_rn(tot, lds_d[j]);
      st->overlap = tot;
    }
    return;
  }
  if (st->overflow) return;
  u64 s0 = 0, s1 = 0;
  for (u32 j = tid; j < a.Ga; j += kScanThreads) {
    s0 += a.aux0[j];
    s1 += a.aux1[j];
  }
  for (int o = 32; o > 0; o >>= 1) {
    s0 += __shfl_down(s0, o, 64);
    s1 += __shfl_down(s1, o, 64);
  }
  const u32 lane = tid & 63u, wave = tid >> 6;
  __shared__ u64 lds1[kScanThreads / 64];
  if (lane == 0) {
    lds[wave] = s0;
    lds1[wave] = s1;
  }
  __syncthreads();
  if (tid != 0) return;
  u64 kills = 0, upd = 0;
  for (u32 w = 0; w < kScanThreads / 64; ++w) {
    kills += lds[w];
    upd += lds1[w];
  }
  const CountSlot in = st->slot[a.in_slot];
  st->slot[a.out_slot].n_entries = in.n_entries;
  st->slot[a.out_slot].n_dead = in.n_dead + kills;
  st->term_updates += upd;
  st->gates += a.n_ops;
  st->kills_total += kills;
  st->map_entries += in.n_entries;
  st->map_launches += 1;
}

// ---- k_map: Clifford / noise / truncation program over every entry ------------
template <int W>
__global__ __launch_bounds__(kBlock) void k_map(u64* __restrict__ keys,
                                                double* __restrict__ coef,
                                                const DevState* __restrict__ st, u32 in_slot,
                                                MapProgram prog, TruncSpec t,
                                                u64* __restrict__ part_kills,
                                                u64* __restrict__ part_updates) {
  __shared__ u64 lds_sum[kWaves];
  if (st->overflow) return;
  const u64 n = st->slot[in_slot].n_entries;
  u64 kills = 0, updates = 0;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    double c = coef[i];
    if (is_dead(c)) continue;
    u64 k[2 * W], k0[2 * W];
    akey_load<W>(keys, i, k);
#pragma unroll
    for (int p = 0; p < 2 * W; ++p) k0[p] = k[p];
    const double c0 = c;
    bool killed = prog.trunc_before && drop_term<W>(k, c, t);
    if (!killed) updates += prog.idle[0];
    for (u32 j = 0; j < prog.n_ops && !killed; ++j) {
      const MapOp op = prog.ops[j];
      ++updates;
      map_apply<W>(k, c, op, prog);
      if ((op.code_flags >> 8) & 1u) killed = drop_term<W>(k, c, t);
      if (!killed) updates += prog.idle[j + 1];
    }
    if (killed) {
      coef[i] = dead_value();
      ++kills;
      continue;
    }
#pragma unroll
    for (int p = 0; p < 2 * W; ++p)
      if (k[p] != k0[p]) keys[i * (2 * W) + (u64)p] = k[p];
    if (__double_as_longlong(c) != __double_as_longlong(c0)) coef[i] = c;
  }
  const u64 bk = block_sum_u64(kills, lds_sum);
  const u64 bu = block_sum_u64(updates, lds_sum);
  if (threadIdx.x == 0) {
    part_kills[blockIdx.x] = bk;
    part_updates[blockIdx.x] = bu;
  }
}

// ---- k_overlap: sum of coefficients of terms with an all-zero x plane ---------
template <int W>
__global__ __launch_bounds__(kBlock) void k_overlap(const u64* __restrict__ keys,
                                                    const double* __restrict__ coef,
                                                    const DevState* __restrict__ st, u32 in_slot,
                                                    double* __restrict__ partials) {
  __shared__ double lds_d[kBlock];
  const u64 n = st->slot[in_slot].n_entries;
  double acc = 0.0;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    u64 xo = 0;
#pragma unroll
    for (int p = 0; p < W; ++p) xo |= keys[i * (2 * W) + (u64)p];
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
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_trace(const u64* __restrict__ keys,
                                                  const double* __restrict__ coef,
                                                  const DevState* __restrict__ st, u32 in_slot,
                                                  PatternProg prog, double* __restrict__ partials) {
  __shared__ double lds_d[kBlock];
  const u64 n = st->slot[in_slot].n_entries;
  double acc = 0.0;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double c = coef[i];
    if (is_dead(c)) continue;
    u64 k[2 * W];
    akey_load<W>(keys, i, k);
    if (pattern_match<W>(k, prog)) acc = __dadd_rn(acc, c);
  }
  lds_d[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) lds_d[threadIdx.x] = __dadd_rn(lds_d[threadIdx.x], lds_d[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = lds_d[0];
}

// ---- overlap of two sums: sum over b's live terms of c_b(k) * c_a(k) ------------
// Build pass: every live entry of a is inserted (key hash -> arena index) into
// a's pair table under a fresh generation.  Live keys of an arena are unique,
// and the table holds at least twice the arena capacity, so each probe ends.
template <int W>
__global__ __launch_bounds__(kBlock) void k_ovl_build(const u64* __restrict__ keys,
                                                      const double* __restrict__ coef,
                                                      const DevState* __restrict__ st, u32 in_slot,
                                                      u64* table, u64 tmask, u32 tgen) {
  const u64 n = st->slot[in_slot].n_entries;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    if (is_dead(coef[i])) continue;
    u64 k[2 * W];
    akey_load<W>(keys, i, k);
    const u64 h = hash_key<W>(k);
    const u64 mine = ((u64)tgen << 40) | ((h >> 56) << 32) | i;
    u64 pos = h & tmask;
    u64 s = table[pos];
    while (true) {
      if ((u32)(s >> 40) != tgen) {
        const u64 old = cas_u64(&table[pos], s, mine);
        if (old == s) break;
        s = old;
        continue;
      }
      pos = (pos + 1) & tmask;
      s = table[pos];
    }
  }
}

// Probe pass over b: per-block partial sums of c_b * c_a for keys present in a.
template <int W>
__global__ __launch_bounds__(kBlock) void k_ovl_probe(const u64* __restrict__ ka,
                                                      const double* __restrict__ ca,
                                                      const u64* __restrict__ table, u64 tmask, u32 tgen,
                                                      const u64* __restrict__ kb,
                                                      const double* __restrict__ cb,
                                                      const DevState* __restrict__ st_b, u32 slot_b,
                                                      double* __restrict__ partials) {
  __shared__ double lds_d[kBlock];
  const u64 n = st_b->slot[slot_b].n_entries;
  double acc = 0.0;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double c = cb[i];
    if (is_dead(c)) continue;
    u64 k[2 * W];
    akey_load<W>(kb, i, k);
    const u64 h = hash_key<W>(k);
    const u64 fp = h >> 56;
    u64 pos = h & tmask;
    while (true) {
      const u64 s = table[pos];
      if ((u32)(s >> 40) != tgen) break;  // not in a
      if (((s >> 32) & 0xffull) == fp) {
        const u64 j = s & 0xffffffffull;
        u64 kj[2 * W];
        akey_load<W>(ka, j, kj);
        bool same = true;
#pragma unroll
        for (int p = 0; p < 2 * W; ++p) same &= (kj[p] == k[p]);
        if (same) {
          acc = __dadd_rn(acc, __dmul_rn(c, ca[j]));
          break;
        }
      }
      pos = (pos + 1) & tmask;
    }
  }
  lds_d[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) lds_d[threadIdx.x] = __dadd_rn(lds_d[threadIdx.x], lds_d[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = lds_d[0];
}

// ---- dead-entry compaction into the other arena --------------------------------
__global__ __launch_bounds__(kBlock) void k_compact_count(const double* __restrict__ coef,
                                                          const DevState* __restrict__ st, u32 in_slot,
                                                          u64* __restrict__ masks, u64* __restrict__ cnt) {
  __shared__ u64 lds[kWaves];
  const u32 lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const u64 n = st->slot[in_slot].n_entries;
  const u32 G = chunk_count(n, kArenaChunk);
  if (blockIdx.x >= G) return;
  u64 lo, hi;
  block_chunk(n, G, blockIdx.x, lo, hi);
  const u32 nw = (u32)((hi - lo + 63) / 64);
  u64 count = 0;
  for (u32 j = wave; j < nw; j += kWaves) {
    const u64 i = lo + (u64)j * 64 + lane;
    const bool live = i < hi && !is_dead(coef[i]);
    const u64 m = __ballot(live);
    if (lane == 0) masks[(lo >> 6) + j] = m;
    count += (u64)__popcll(m);
  }
  const u64 tot = block_sum_all(lane == 0 ? count : 0ull, lds);  // count is wave-uniform
  if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_compact_write(const u64* __restrict__ sk,
                                                          const double* __restrict__ sc,
                                                          u64* __restrict__ dk, double* __restrict__ dc,
                                                          DevState* st, u32 in_slot, u32 out_slot,
                                                          const u64* __restrict__ masks,
                                                          const u64* __restrict__ cnt) {
  __shared__ u64 lds[kWaves];
  __shared__ u64 lm[kChunkMaxWords];
  __shared__ u32 loff[kChunkMaxWords];
  __shared__ u32 lt[kBlock];
  const u32 lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const u64 n = st->slot[in_slot].n_entries;
  const u32 G = chunk_count(n, kArenaChunk);
  if (blockIdx.x >= G) return;
  u64 lo, hi;
  block_chunk(n, G, blockIdx.x, lo, hi);
  const u32 nw = (u32)((hi - lo + 63) / 64);
  const u64 base_off = chunk_offset(cnt, blockIdx.x, lds);
  if (blockIdx.x == 0) {
    const u64 tot = chunk_offset(cnt, G, lds);
    if (threadIdx.x == 0) {
      st->slot[out_slot].n_entries = tot;
      st->slot[out_slot].n_dead = 0;
    }
  }
  chunk_mask_scan(masks, lo >> 6, nw, lm, loff, lt);
  for (u32 j = wave; j < nw; j += kWaves) {
    const u64 m = lm[j];
    if ((m >> lane) & 1ull) {
      const u64 i = lo + (u64)j * 64 + lane;
      const u64 dst = base_off + loff[j] + (u64)__popcll(m & lanes_below(lane));
      u64 k[2 * W];
      akey_load<W>(sk, i, k);
      akey_store<W>(dk, dst, k);
      dc[dst] = sc[i];
    }
  }
}

// ---- small utility kernels ------------------------------------------------------
__global__ void k_fill_u64(u64* p, u64 n, u64 v) {
  const u64 stride = (u64)gridDim.x * blockDim.x;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Grow / clone: copy the first n entries (nw key words each, then the
// coefficients) into another arena; key arenas do not depend on the capacity.
__global__ void k_copy_arena(const u64* __restrict__ sk, const double* __restrict__ sc,
                             u64* __restrict__ dk, double* __restrict__ dc, u32 nw, u64 n) {
  const u64 stride = (u64)gridDim.x * blockDim.x;
  const u64 nk = n * nw;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < nk; i += stride) dk[i] = sk[i];
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dc[i] = sc[i];
}

}  // namespace ppvm
