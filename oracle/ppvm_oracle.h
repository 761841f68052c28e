/*
 * ppvm_oracle.h — CPU restatement of the reference PauliSum propagation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle and the CPU
 * baseline of the measurement harness.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only to check results
 * against (or as the timed CPU baseline) — never as part of the product path.
 *
 * It restates, in plain C++, the algorithm of the files listed in
 * ppvm_oracle.cpp: an insertion-ordered hash map (IndexMap semantics), rotation
 * branching through a scratch vector merged with `+=`, Clifford re-insertion,
 * order-preserving `retain` truncation, and the insertion-order fold of
 * overlap_with_zero.  Ops are the `ppvm_op` records of include/ppvm_hip.h.
 */
#ifndef PPVM_ORACLE_H
#define PPVM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/ppvm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ppvo_sum* ppvo_handle;

ppvo_handle ppvo_create(uint32_t n_qubits, double min_abs_coeff, uint64_t max_pauli_weight);
void ppvo_destroy(ppvo_handle h);
ppvo_handle ppvo_clone(ppvo_handle h);
uint32_t ppvo_words(ppvo_handle h);
void ppvo_set_preserve(ppvo_handle h, const uint64_t* x, const uint64_t* z, size_t count);
void ppvo_add_terms(ppvo_handle h, const uint64_t* x, const uint64_t* z, const double* c,
                    size_t count);
/* Returns 0 on success, nonzero (and does nothing more) on a bad op. */
int ppvo_apply(ppvo_handle h, const ppvm_op* ops, size_t n);
double ppvo_overlap_with_zero(ppvo_handle h);
double ppvo_overlap(ppvo_handle a, ppvo_handle b);
/* trace(pattern) (pattern grammar of crates/ppvm-pauli-word/src/pattern/parse.rs);
 * returns nonzero if the pattern does not parse. */
int ppvo_trace(ppvo_handle h, const char* pattern, double* out);
/* Pattern membership of `count` words (x, z: count*W words, term-major). */
int ppvo_pattern_contains(ppvo_handle h, const char* pattern, const uint64_t* x, const uint64_t* z,
                          uint8_t* out, size_t count);
uint64_t ppvo_len(ppvo_handle h);
/* Writes min(cap, len) terms in insertion order; returns len. */
uint64_t ppvo_export(ppvo_handle h, uint64_t* x, uint64_t* z, double* c, size_t cap);
uint64_t ppvo_term_updates(ppvo_handle h);
/* Order-independent digest of the live terms: {count, sum of key digests,
 * sum of key-and-coefficient-bit digests} (modulo 2^64). */
void ppvo_fingerprint(ppvo_handle h, uint64_t out[3]);
/* Replays `ops` and, after every op whose index is listed in `marks`
 * (ascending), records len and overlap_with_zero into out_len/out_ovl.
 * Returns the number of marks recorded. */
size_t ppvo_run_marked(ppvo_handle h, const ppvm_op* ops, size_t n, const uint64_t* marks,
                       size_t n_marks, uint64_t* out_len, double* out_ovl);
/* Multi-threaded sharded variant (hash-partitioned shards, one worker thread
 * per shard, modelled on the DashMap + rayon configuration): replays `ops`
 * on a copy of h's terms with `threads` workers and writes the result back
 * into h.  Term sets and coefficients are identical to the single-thread
 * path; only the order of the entries differs. */
int ppvo_apply_sharded(ppvo_handle h, const ppvm_op* ops, size_t n, int threads);

/* Lossy sums (PauliSum<LossyPauliWord>): words carry a loss plane l besides
 * x and z (W words each).  ppvo_apply, ppvo_len, ppvo_overlap_with_zero,
 * ppvo_trace and ppvo_overlap accept lossy handles; ppvo_apply takes the loss
 * channel op codes on them. */
ppvo_handle ppvo_lossy_create(uint32_t n_qubits, double min_abs_coeff, uint64_t max_pauli_weight,
                              uint64_t max_loss_weight);
void ppvo_lossy_set_preserve(ppvo_handle h, const uint64_t* x, const uint64_t* z, const uint64_t* l,
                             size_t count);
void ppvo_lossy_add_terms(ppvo_handle h, const uint64_t* x, const uint64_t* z, const uint64_t* l,
                          const double* c, size_t count);
uint64_t ppvo_lossy_export(ppvo_handle h, uint64_t* x, uint64_t* z, uint64_t* l, double* c, size_t cap);

#ifdef __cplusplus
}
#endif

#endif
