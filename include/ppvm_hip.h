/*
 * ppvm_hip.h — C ABI of the MI355X (gfx950) Pauli-propagation engine.
 *
 * This is the drop-in boundary for the `PauliSum` Heisenberg-propagation path of
 * QuEraComputing/ppvm.  Every entry point below replaces one method of the
 * reference's `PauliSum` surface (Rust traits, PyO3 binding); the citation in
 * each comment names the reference interface it stands in for.  All pointers are
 * host memory: terms are copied in and out; device arenas are owned by the handle.
 *
 * Term layout across the boundary: a term is W = ceil(n_qubits/64) u64 words of
 * x bits followed (in a separate array) by W u64 words of z bits.  Qubit i is bit
 * (i mod 64) of word i/64, the little-endian Lsb0 layout of the reference's
 * `PauliWord<[u8;N]>` planes (crates/ppvm-pauli-word/src/word/data.rs:43-68).
 * The Pauli on a qubit is (x,z): I=(0,0) X=(1,0) Z=(0,1) Y=(1,1)
 * (crates/ppvm-pauli-word/src/word/data.rs:203-213).
 *
 * Semantics follow the reference exactly:
 *  - gates never truncate by themselves; `*_many` style calls (a target array)
 *    are a sequential loop of single gates with no truncation in between
 *    (crates/ppvm-traits/src/traits/clifford.rs:175-218,
 *     crates/ppvm-traits/src/traits/branch/rot1.rs:25-45);
 *  - `ppvm_truncate` drops |c| < min_abs_coeff, then weight > max_pauli_weight
 *    (crates/ppvm-pauli-sum/src/strategy.rs:19-30,86-89,133), never dropping a
 *    preserved string (crates/ppvm-pauli-sum/src/sum/data.rs:275-311).
 *
 * Errors: every call returns PPVM_OK (0) or an error code; `ppvm_last_error()`
 * gives a thread-local message.  Where the reference panics (qubit index out of
 * range, word/data.rs:203-213) or raises ValueError (odd pair list,
 * crates/ppvm-python-native/src/lib.rs:12-22) this ABI returns PPVM_EINVAL.
 *
 * Threading: a handle is used by one host thread at a time (the reference takes
 * `&mut self` on every gate); different handles may be driven concurrently.
 * Work is queued on the handle's own HIP stream and may run asynchronously;
 * calls that return data (len, overlap, export) synchronise.
 */
#ifndef PPVM_HIP_H
#define PPVM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
enum {
  PPVM_OK = 0,
  PPVM_EINVAL = 1,  /* bad argument (index out of range, odd pairs, size)     */
  PPVM_ENOMEM = 2,  /* device or host allocation failed                      */
  PPVM_EHIP = 3,    /* HIP runtime error                                     */
  PPVM_ESTATE = 4,  /* handle in an unusable state (e.g. after a HIP error)  */
  PPVM_ENODEV = 5   /* no usable gfx950 device                               */
};

/* ---- op codes (for ppvm_apply and the batched programs) ----------------- */
enum ppvm_opcode {
  PPVM_OP_NOP = 0,
  /* Clifford conjugation, crates/ppvm-pauli-sum/src/sum/clifford.rs:37-217 */
  PPVM_OP_X = 1, PPVM_OP_Y = 2, PPVM_OP_Z = 3, PPVM_OP_H = 4, PPVM_OP_S = 5,
  PPVM_OP_S_DAG = 6, PPVM_OP_SQRT_X = 7, PPVM_OP_SQRT_X_DAG = 8,
  PPVM_OP_SQRT_Y = 9, PPVM_OP_SQRT_Y_DAG = 10,
  PPVM_OP_CNOT = 11, PPVM_OP_CZ = 12, PPVM_OP_CY = 13,
  /* Pauli rotations exp(-i theta/2 G), crates/ppvm-pauli-sum/src/sum/rot1.rs:55-128,
     crates/ppvm-pauli-sum/src/sum/rot2.rs:63-184 */
  PPVM_OP_RX = 20, PPVM_OP_RY = 21, PPVM_OP_RZ = 22,
  PPVM_OP_RXX = 23, PPVM_OP_RYY = 24, PPVM_OP_RZZ = 25,
  /* diagonal noise channels, crates/ppvm-pauli-sum/src/sum/noise.rs:17-131 */
  PPVM_OP_PAULI_ERROR = 40, PPVM_OP_DEPOLARIZE1 = 41, PPVM_OP_DEPOLARIZE2 = 42,
  PPVM_OP_PAULI_ERROR2 = 43,
  /* amplitude damping (gamma in p[0]), crates/ppvm-pauli-sum/src/sum/noise.rs:132-157 */
  PPVM_OP_AMPLITUDE_DAMPING = 44,
  /* loss channels, lossy sums only (ppvm_lossy_apply), crates/ppvm-pauli-sum/src/sum/noise.rs:159-274:
     LOSS (q0, p[0]), CORRELATED_LOSS (q0, q1, p[0..2]), RESET_LOSS (q0) */
  PPVM_OP_LOSS = 45, PPVM_OP_CORRELATED_LOSS = 46, PPVM_OP_RESET_LOSS = 47,
  /* explicit truncation, crates/ppvm-pauli-sum/src/sum/data.rs:275 */
  PPVM_OP_TRUNCATE = 60
};

/* One operation of a program.  q1 is the second qubit of two-qubit ops;
 * theta the rotation angle; p[] the channel probabilities (3 for
 * PAULI_ERROR, 1 for DEPOLARIZE1/2, 15 for PAULI_ERROR2, gamma for
 * AMPLITUDE_DAMPING). */
typedef struct ppvm_op {
  uint32_t code;
  uint32_t q0;
  uint32_t q1;
  uint32_t reserved;
  double theta;
  double p[15];
} ppvm_op;

/* Counters kept by a handle (device side, read back on request). */
typedef struct ppvm_stats {
  uint64_t term_updates;   /* sum over gates of live input terms             */
  uint64_t gates;          /* gates applied (rotations + map ops)            */
  uint64_t rotations;      /* rotation gates applied                         */
  uint64_t n_entries;      /* arena entries (live + dead)                    */
  uint64_t n_live;         /* live terms                                     */
  uint64_t capacity;       /* arena capacity in entries                      */
  uint64_t peak_live;      /* largest live count seen at a sync point        */
  uint64_t syncs;          /* host synchronisations performed                */
  uint64_t regrows;        /* arena growths (overflow replays included)      */
  uint64_t compactions;    /* dead-entry compactions                         */
} ppvm_stats;

typedef struct ppvm_sum* ppvm_handle;

/* Message for the last failing call on this thread ("" if none). */
const char* ppvm_last_error(void);

/* Library / device information: writes a short description (device name,
 * CU count, gfx arch) into buf. */
int ppvm_device_info(char* buf, size_t cap);

/* Replaces PauliSum::builder().n_qubits().strategy(CombinedStrategy(
 *   CoefficientThreshold(min_abs_coeff), MaxPauliWeight(max_pauli_weight)))
 *   .capacity().build()  (crates/ppvm-pauli-sum/src/sum/data.rs:66-92,
 *   crates/ppvm-python-native/src/interface.rs:80-111).
 * max_pauli_weight = UINT64_MAX disables the weight cap (strategy.rs:86-88).
 * capacity_hint: initial arena size in terms (0 = default). */
int ppvm_create(uint32_t n_qubits, double min_abs_coeff, uint64_t max_pauli_weight,
                uint64_t capacity_hint, ppvm_handle* out);
int ppvm_destroy(ppvm_handle h);
/* Deep copy, replaces Clone / __copy__ (interface.rs:373-387). */
int ppvm_clone(ppvm_handle h, ppvm_handle* out);
/* Words per plane W for this handle. */
int ppvm_words(ppvm_handle h, uint32_t* w);
/* Preserve set, replaces the `preserve_strings` builder argument
 * (crates/ppvm-pauli-sum/src/sum/data.rs:75-77). x,z: count*W words. */
int ppvm_set_preserve(ppvm_handle h, const uint64_t* x, const uint64_t* z, size_t count);

/* Replaces `ps += (word, c)` = entry(k).and_modify(+=).or_insert(c)
 * (crates/ppvm-pauli-sum/src/sum/ops.rs:140-151, map/hashmap.rs:40-44),
 * one call per term in order.  x,z: count*W words (term-major). */
int ppvm_add_terms(ppvm_handle h, const uint64_t* x, const uint64_t* z, const double* c,
                   size_t count);

/* Empties the sum and keeps its arenas (capacity, preserve set, truncation
 * strategy and counters are unchanged); replaces `ps.data_mut().clear()`
 * (ACMap::clear, crates/ppvm-traits/src/traits/map.rs:23).  Queued gates are
 * discarded.  Lets a caller rerun a circuit on a pre-sized handle without
 * reallocating. */
int ppvm_clear(ppvm_handle h);

/* Clifford gates on a target list; replace Clifford / CliffordExtensions and
 * their *_many forms (crates/ppvm-traits/src/traits/clifford.rs:30-79,175-218). */
int ppvm_x(ppvm_handle h, const uint32_t* targets, size_t n);
int ppvm_y(ppvm_handle h, const uint32_t* targets, size_t n);
int ppvm_z(ppvm_handle h, const uint32_t* targets, size_t n);
int ppvm_h(ppvm_handle h, const uint32_t* targets, size_t n);
int ppvm_s(ppvm_handle h, const uint32_t* targets, size_t n);
int ppvm_s_dag(ppvm_handle h, const uint32_t* targets, size_t n);
int ppvm_sqrt_x(ppvm_handle h, const uint32_t* targets, size_t n);
int ppvm_sqrt_x_dag(ppvm_handle h, const uint32_t* targets, size_t n);
int ppvm_sqrt_y(ppvm_handle h, const uint32_t* targets, size_t n);
int ppvm_sqrt_y_dag(ppvm_handle h, const uint32_t* targets, size_t n);
/* pairs: 2*npairs qubit indices (control, target)... */
int ppvm_cnot(ppvm_handle h, const uint32_t* pairs, size_t npairs);
int ppvm_cz(ppvm_handle h, const uint32_t* pairs, size_t npairs);
int ppvm_cy(ppvm_handle h, const uint32_t* pairs, size_t npairs);

/* Rotations; replace RotationOne::{rx,ry,rz}_many (traits/branch/rot1.rs:8-45)
 * and RotationTwo::{rxx,ryy,rzz}_many (traits/branch/rot2.rs:6-47).
 * sin/cos of theta are taken once per call on the host, as the reference does
 * (sum/rot1.rs:58). */
int ppvm_rx(ppvm_handle h, const uint32_t* targets, size_t n, double theta);
int ppvm_ry(ppvm_handle h, const uint32_t* targets, size_t n, double theta);
int ppvm_rz(ppvm_handle h, const uint32_t* targets, size_t n, double theta);
int ppvm_rxx(ppvm_handle h, const uint32_t* pairs, size_t npairs, double theta);
int ppvm_ryy(ppvm_handle h, const uint32_t* pairs, size_t npairs, double theta);
int ppvm_rzz(ppvm_handle h, const uint32_t* pairs, size_t npairs, double theta);
/* RotXY::r = rz(phi); rx(theta); rz(-phi)  (sum/rot1.rs:135-143). */
int ppvm_r(ppvm_handle h, uint32_t q, double axis_angle, double theta);

/* Diagonal noise channels, crates/ppvm-pauli-sum/src/sum/noise.rs:17-131. */
int ppvm_pauli_error(ppvm_handle h, uint32_t q, const double p[3]);
int ppvm_depolarize1(ppvm_handle h, uint32_t q, double p);
int ppvm_depolarize2(ppvm_handle h, uint32_t a, uint32_t b, double p);
int ppvm_two_qubit_pauli_error(ppvm_handle h, uint32_t a, uint32_t b, const double p[15]);
/* Replaces AmplitudeDamping::amplitude_damping (crates/ppvm-pauli-sum/src/sum/noise.rs:132-157,
 * crates/ppvm-python-native/src/interface.rs:365-368): X, Y on q scale by
 * sqrt(1 - gamma); Z scales by (1 - gamma) and branches gamma * c onto the word
 * with I on q (merged with += like a rotation child). */
int ppvm_amplitude_damping(ppvm_handle h, uint32_t q, double gamma);

/* Replaces PauliSum::truncate (crates/ppvm-pauli-sum/src/sum/data.rs:275-311). */
int ppvm_truncate(ppvm_handle h);

/* Replaces overlap_with_zero = trace("Z?*")
 * (crates/ppvm-python-native/src/interface.rs:122-130, sum/trace.rs:19-21):
 * the sum of coefficients of terms whose x plane is all zero. */
int ppvm_overlap_with_zero(ppvm_handle h, double* out);
/* Replaces trace(pattern) (crates/ppvm-python-native/src/interface.rs:122-126,
 * sum/trace.rs:9-22): the sum of coefficients of terms the Pauli pattern accepts,
 * with the grammar of crates/ppvm-pauli-word/src/pattern/parse.rs:30-143 and
 * the matcher of pattern/contains.rs:20-151 ("Z?*" = ppvm_overlap_with_zero).
 * A pattern that does not parse (the reference panics) returns PPVM_EINVAL. */
int ppvm_trace(ppvm_handle h, const char* pattern, double* out);
/* Replaces PauliSum::overlap (crates/ppvm-pauli-sum/src/sum/trace.rs:30-37):
 * sum over terms k of b of b[k] * a[k].  Both handles must have equal n_qubits. */
int ppvm_overlap(ppvm_handle a, ppvm_handle b, double* out);

/* Replaces len() (sum/data.rs:165). */
int ppvm_len(ppvm_handle h, uint64_t* out);
/* Replaces terms() / iter() (interface.rs:393-395): live terms in insertion
 * order.  If cap < len, nothing is written, *n = len and PPVM_EINVAL is returned. */
int ppvm_export_terms(ppvm_handle h, uint64_t* x, uint64_t* z, double* c, size_t cap,
                      size_t* n);

/* Apply a whole program (a sequence of ops, PPVM_OP_TRUNCATE included) in one
 * call.  Equivalent to the corresponding sequence of single calls above. */
int ppvm_apply(ppvm_handle h, const ppvm_op* ops, size_t n);

/* Wait for all queued work; check device-side overflow and regrow if needed. */
int ppvm_synchronize(ppvm_handle h);
/* Order-independent digest of the live terms, for comparing a sum with one
 * held by another backend the way PartialEq compares two sums (IndexMap
 * equality ignores order; crates/ppvm-pauli-sum/src/sum/data.rs:349-356)
 * without copying the terms out: out = {count, sum_k d(k), sum_k e(k, c_k)}
 * modulo 2^64, where d chains the SplitMix64 finaliser m over the key words
 * (d = m(...m(m(0x9E3779B97F4A7C15 ^ x_0) ^ x_1)... ^ z_{W-1}), x words then
 * z words, W = ceil(n_qubits/64)) and e(k, c) = m(d(k) ^ bits of c). */
int ppvm_fingerprint(ppvm_handle h, uint64_t out[3]);
int ppvm_get_stats(ppvm_handle h, ppvm_stats* out);
/* Device counters for the measurement harness, in this order: term_updates,
 * gates, rotations, scanned entries, anticommuting (worklist) terms, merged
 * pairs, appended children, truncated terms, map-program entries, map
 * launches, bit-sliced planes read by scans (entries x planes), capacity,
 * syncs, regrows, compactions, peak live terms, entries transposed into the
 * bit-sliced planes, worklist records that the partitioned pairing sent to
 * the global table (buckets too large for one workgroup's LDS), LDS-resident
 * runs of a small sum, and the steps those runs completed. */
int ppvm_get_counters(ppvm_handle h, uint64_t out[20]);

/* Per-phase kernel timing (HIP events on the handle's stream), for the
 * measurement harness.  Phases: 0 rotation scan, 1 rotation pair/merge,
 * 2 rotation emit, 3 map program (Clifford/noise/truncate), 4 overlap,
 * 5 compaction.  When enabled, every kernel launch is bracketed by events
 * and its duration accumulated; ms[i] and launches[i] receive the totals. */
int ppvm_profile_enable(ppvm_handle h, int enable);
int ppvm_profile_read(ppvm_handle h, double ms[8], uint64_t launches[8]);

/* ---- batched mode -------------------------------------------------------------
 * No counterpart in the reference (it propagates one PauliSum per call); this
 * runs many independent instances, each with the semantics above, in one
 * persistent kernel.  An instance is an initial term list, a program of
 * ppvm_op records (PPVM_OP_TRUNCATE included) and its truncation strategy.
 * Bit 0 of ppvm_op.reserved marks ops after which (len, overlap_with_zero)
 * is recorded. */
typedef struct ppvm_batch_instance {
  uint32_t n_qubits;
  uint32_t n_terms;            /* initial terms, `+=` semantics                 */
  const uint64_t* x;           /* n_terms * ceil(n_qubits/64) words             */
  const uint64_t* z;
  const double* c;
  const ppvm_op* ops;
  uint32_t n_ops;
  uint32_t reserved;
  double min_abs_coeff;
  uint64_t max_pauli_weight;   /* UINT64_MAX: no weight cap                     */
  uint64_t capacity;           /* arena capacity in terms (0: default)          */
} ppvm_batch_instance;

typedef struct ppvm_batch_config {
  uint32_t threads;            /* workgroup size: multiple of 64, <= 1024        */
  uint32_t blocks;             /* persistent grid (0: min(instances, 2 x CUs))   */
  uint32_t segment_ops;        /* ops per launch segment (0: whole program)      */
  uint32_t use_graph;          /* capture the launch sequence in a hipGraph      */
} ppvm_batch_config;

typedef struct ppvm_batch* ppvm_batch_handle;

const char* ppvm_batch_last_error(void);
int ppvm_batch_create(const ppvm_batch_instance* inst, size_t n, ppvm_batch_handle* out);
int ppvm_batch_destroy(ppvm_batch_handle b);
/* Re-sizes one instance's arenas (e.g. after it overflowed). */
int ppvm_batch_set_capacity(ppvm_batch_handle b, size_t i, uint64_t capacity);
/* Runs the instances listed in order[0..n_order) (NULL: all, in index order;
 * the queue hands them out in this order) from their initial terms; *ms gets
 * the device time of the run. */
int ppvm_batch_run(ppvm_batch_handle b, const ppvm_batch_config* cfg, const uint32_t* order,
                   size_t n_order, double* ms);
/* Per instance: status (0 done, 1 arena overflow, 2 not finished), final term
 * count, last recorded overlap, term updates, peak term count.  Arrays of
 * length n (any may be NULL). */
int ppvm_batch_results(ppvm_batch_handle b, uint32_t* status, uint64_t* n_final,
                       double* overlap_last, uint64_t* term_updates, uint64_t* peak);
/* Per instance: seconds from the start of the last run until the instance
 * finished (device wall clock). */
int ppvm_batch_latency(ppvm_batch_handle b, double* seconds);
int ppvm_batch_marks(ppvm_batch_handle b, size_t i, uint64_t* lens, double* overlaps, size_t cap,
                     size_t* n);
int ppvm_batch_export(ppvm_batch_handle b, size_t i, uint64_t* x, uint64_t* z, double* c, size_t cap,
                      size_t* n);

/* One instance through the batched kernel without allocating device memory.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code that must not call hipMalloc, such as serving loops or harnesses that
 * forbid it).  Same results as ppvm_batch_create + ppvm_batch_run +
 * ppvm_batch_results + ppvm_batch_export on a batch holding only `inst`
 * (terms, order and coefficient bits; `capacity` sizes the arena as there).
 * Every device buffer lives in a fixed 64 MiB scratch area per GPU that the
 * library holds until the process ends; a run needing more (arena, hash
 * table, program) fails with PPVM_EINVAL, and an arena overflow, which this
 * entry point cannot grow, fails with PPVM_ENOMEM.  Inputs and outputs are
 * host memory; work runs on the device's null stream and the call returns
 * when it is done.  Calls on one GPU share the scratch area: calls that
 * overlap (other host threads) race and must be serialised by the caller.
 * threads: workgroup size (0: 256).  x, z, c receive the final terms
 * (cap terms at most; *n = count); overlap_last and term_updates may be NULL. */
int ppvm_batch_run1_noalloc(const ppvm_batch_instance* inst, uint32_t threads, uint64_t* x, uint64_t* z,
                            double* c, size_t cap, size_t* n, double* overlap_last, uint64_t* term_updates);

/* ---- lossy sums -----------------------------------------------------------------
 * Replaces PauliSum<LossyPauliWord> (crates/ppvm-pauli-word/src/loss/data.rs:33-45,
 * crates/ppvm-python-native/src/interface.rs:13-66, ppvm-python/src/ppvm/paulisum.py:420-520):
 * Pauli words over {I, X, Y, Z, L}, where L marks a lost qubit.  A word crosses
 * the boundary as three planes x, z, l of W words each (W as for ppvm_words;
 * qubit i at bit i % 64 of word i / 64); a lost qubit has l = 1 and x = z = 0.
 * Gates treat a lost qubit as the reference does (Cliffords and channels skip
 * the term, a two-qubit rotation acts on the surviving qubit only).  The
 * truncation strategy is CombinedStrategy(CoefficientThreshold, MaxPauliWeight,
 * MaxLossWeight) (strategy.rs:12-30,137-170); the Pauli weight counts L.
 * Error messages are read with ppvm_lossy_last_error. */
typedef struct ppvm_lossy_sum* ppvm_lossy_handle;
const char* ppvm_lossy_last_error(void);
/* max_pauli_weight / max_loss_weight = UINT64_MAX disable those caps. */
int ppvm_lossy_create(uint32_t n_qubits, double min_abs_coeff, uint64_t max_pauli_weight,
                      uint64_t max_loss_weight, uint64_t capacity_hint, ppvm_lossy_handle* out);
int ppvm_lossy_destroy(ppvm_lossy_handle h);
int ppvm_lossy_clone(ppvm_lossy_handle h, ppvm_lossy_handle* out);
int ppvm_lossy_words(ppvm_lossy_handle h, uint32_t* w);
/* Preserve set (sum/data.rs:75-77, restore order of sum/data.rs:275-311). */
int ppvm_lossy_set_preserve(ppvm_lossy_handle h, const uint64_t* x, const uint64_t* z, const uint64_t* l,
                            size_t count);
/* `ps += (word, c)` per term in order (sum/ops.rs:140-151). */
int ppvm_lossy_add_terms(ppvm_lossy_handle h, const uint64_t* x, const uint64_t* z, const uint64_t* l,
                         const double* c, size_t count);
/* A program of ppvm_op records: every op code of ppvm_apply plus PPVM_OP_LOSS,
 * PPVM_OP_CORRELATED_LOSS and PPVM_OP_RESET_LOSS.  Gates never truncate by
 * themselves; PPVM_OP_TRUNCATE replaces PauliSum::truncate. */
int ppvm_lossy_apply(ppvm_lossy_handle h, const ppvm_op* ops, size_t n);
int ppvm_lossy_len(ppvm_lossy_handle h, uint64_t* out);
/* Terms in insertion order; cap < len writes nothing, sets *n = len, PPVM_EINVAL. */
int ppvm_lossy_export_terms(ppvm_lossy_handle h, uint64_t* x, uint64_t* z, uint64_t* l, double* c,
                            size_t cap, size_t* n);
/* trace("Z?*"): terms with every qubit I or Z (an L never matches a pattern,
 * pattern/contains.rs:120-151). */
int ppvm_lossy_overlap_with_zero(ppvm_lossy_handle h, double* out);
int ppvm_lossy_trace(ppvm_lossy_handle h, const char* pattern, double* out);
/* sum over terms k of b of b[k] * a[k] (sum/trace.rs:30-37). */
int ppvm_lossy_overlap(ppvm_lossy_handle a, ppvm_lossy_handle b, double* out);

/* ---- learned surrogate ---------------------------------------------------------
 * No counterpart in the reference.  An ensemble of MLPs (in -> H -> H -> H -> 1,
 * GELU) with BF16 weights/activations and FP32 accumulation on MFMA, trained
 * with Adam (FP32 master weights) on expectation values produced by the
 * propagator.  Parameter vector layout per member, for l = 0..3:
 * W_l [N_l][K_l] then b_l [N_l], with K = (32, H, H, H), N = (H, H, H, 32);
 * inputs beyond in_dim and output rows beyond 0 are padding (kept zero). */
typedef struct ppvm_mlp* ppvm_mlp_handle;
const char* ppvm_mlp_last_error(void);
int ppvm_mlp_create(uint32_t in_dim, uint32_t hidden, uint32_t members, uint32_t max_batch, uint64_t seed,
                    ppvm_mlp_handle* out);
int ppvm_mlp_destroy(ppvm_mlp_handle h);
uint64_t ppvm_mlp_param_count(ppvm_mlp_handle h);
int ppvm_mlp_get_params(ppvm_mlp_handle h, uint32_t member, float* out, size_t cap);
int ppvm_mlp_set_params(ppvm_mlp_handle h, uint32_t member, const float* in, size_t n);
/* Full-batch training (MSE), `steps` Adam steps with cosine decay from lr;
 * final_loss receives the last step's loss per member. */
int ppvm_mlp_train(ppvm_mlp_handle h, const float* X, const float* Y, uint32_t n, uint32_t steps, float lr,
                   uint32_t use_graph, float* final_loss);
/* Ensemble mean and standard deviation per row (per_member: members x n, may be NULL). */
int ppvm_mlp_predict(ppvm_mlp_handle h, const float* X, uint32_t n, float* mean, float* sd, float* per_member);
/* One forward + backward pass: gradients of the MSE loss for one member. */
int ppvm_mlp_grads(ppvm_mlp_handle h, const float* X, const float* Y, uint32_t n, uint32_t member, float* grads,
                   size_t cap, float* loss);
/* C[M][N] = bf16(A)[M][K] * bf16(B)[N][K]^T through the MFMA GEMM kernel (tests). */
int ppvm_mlp_gemm_check(const float* A, const float* B, int M, int N, int K, float* C);

#ifdef __cplusplus
}
#endif

#endif /* PPVM_HIP_H */
