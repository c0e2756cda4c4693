/*
 * damavand_gpu.h -- C ABI of libdamavand-gpu.so, the MI355X (gfx950) statevector
 * library behind damavand's apply_method="gpu".
 *
 * Two layers are exported.
 *
 * 1. Compatibility layer: the exact symbols, argument types and semantics that
 *    the reference's Rust crate binds (src/qubit_backend/circuit.rs:27-51,
 *    src/qubit_backend/circuit_gpu.rs:3-24, circuit_distributed_gpu.rs:6-33),
 *    defined by damavand-gpu/rust_communication.cu in the reference.  The Rust
 *    crate links `damavand-gpu` by name (build.rs:7-13), so this library can
 *    replace the CUDA one without a change on the Rust side.  These calls keep
 *    the reference's contract: one process-global state, `int` sizes (so at
 *    most 2^30 amplitudes), synchronous, void returns.  Errors are printed to
 *    stderr and kept for dmv_last_error(); the process is never exit()ed.
 *
 * 2. The dmv_* layer: opaque handles, 64-bit sizes (up to 33 qubits of
 *    complex128 on one GPU), complex128 or complex64, int status codes
 *    (DMV_OK == 0) with a thread-local message from dmv_last_error().  Host
 *    buffers are owned by the caller and are only used during the call; device
 *    memory is owned by the handle.  One handle must not be used from two
 *    threads at once; different handles may (a call on two states, such as
 *    dmv_inner_product or dmv_circuit_adjoint, uses both of them).  Each state
 *    and circuit owns its device buffers, library initialisation and the
 *    tuning table are locked, and all work is ordered on one library stream.
 *    Profiling totals taken while several threads launch include the
 *    other threads' kernels queued between a call's begin and end events.
 */
#ifndef DAMAVAND_GPU_H
#define DAMAVAND_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* 1. Compatibility layer (names and C types of rust_communication.cu)      */
/* ------------------------------------------------------------------------ */

/* rust_communication.cu:24-29 (bound at circuit.rs:30) */
int get_number_of_available_gpus(void);

/* rust_communication.cu:31-37 (bound at circuit.rs:33): total device memory in MiB. */
double get_memory_for_gpu(int gpu_id);

/* rust_communication.cu:39-47: 1 if gpu `source` can access `target` (0 with one GPU). */
int peer_access_allowed(int source_gpu_id, int target_gpu_id);

/* rust_communication.cu:49-104 (bound at circuit.rs:31): prints per-kernel
 * launch counts and mean durations recorded with HIP events. */
void print_timers(void);

/* rust_communication.cu:143-198 (bound at circuit.rs:35-39).  Allocates the
 * process-global state of num_amplitudes_per_gpu amplitudes (a power of two)
 * and sets it to |0...0> when is_first_node is nonzero, else to all zeros.
 * The reference's C side declares `bool`; the Rust side passes c_int, which
 * is what is declared here.  num_gpus_per_node must be 1: the state lives on
 * one GPU. */
void init_quantum_state(int num_amplitudes_per_gpu, int num_gpus_per_node, int is_first_node);

/* rust_communication.cu:339-359 (bound at circuit_gpu.rs:6-13).  Applies the
 * row-major 2x2 matrix (re[4], im[4]) to `target_qubit`, controlled by
 * `control_qubit` (-1 for none), on the global state. */
void apply_one_qubit_gate_gpu_local(double *gate_matrix_real, double *gate_matrix_imaginary,
                                    int num_qubits, int num_amplitudes_per_gpu,
                                    int control_qubit, int target_qubit);

/* rust_communication.cu:361-382 (bound at circuit_gpu.rs:15-22).  Multi-GPU
 * partner update: not supported on one GPU; records an error. */
void apply_one_qubit_gate_gpu_distributed(double *gate_matrix_real, double *gate_matrix_imaginary,
                                          int num_qubits, int num_amplitudes_per_gpu,
                                          int control_qubit, int target_qubit);

/* rust_communication.cu:330-337 (bound at circuit.rs:41-44, which declares a
 * pointer return that the caller ignores).  Writes |a_i|^2 into the caller's
 * buffer of num_amplitudes_per_gpu doubles.  The state is kept (the
 * reference's cudaDeviceReset() at rust_communication.cu:246 is not repeated). */
void measure_on_gpu(int num_amplitudes_per_gpu, double *probabilities);

/* rust_communication.cu:200-252 and :254-328: both variants compute the same
 * probabilities as measure_on_gpu. */
void sequential_measure_on_gpu(int num_amplitudes_per_gpu, double *probabilities);
void concurrent_measure_on_gpu(int num_amplitudes_per_gpu, double *probabilities);

/* rust_communication.cu:450-482 (bound at circuit.rs:46-50): copies the state
 * to the caller's real and imaginary arrays. */
void retrieve_amplitudes_on_host(int num_amplitudes_per_gpu, double *local_amplitudes_real,
                                 double *local_amplitudes_imaginary);

/* rust_communication.cu:385-398: loads caller amplitudes into the state (the
 * reference copies the real part into the imaginary buffer at
 * quantum_amplitudes.cu:93-96; here both parts are loaded as given). */
void load_amplitudes_local_on_device(int num_amplitudes_per_gpu, double *local_amplitudes_real,
                                     double *local_amplitudes_imaginary);

/* rust_communication.cu:400-448 and :106-141 (bound at
 * circuit_distributed_gpu.rs:6-33): multi-GPU exchange; not supported on one
 * GPU; record an error. */
void split_amplitudes_between_gpus(int num_amplitudes_per_gpu, double *local_amplitudes_real,
                                   double *local_amplitudes_imaginary,
                                   double *partner_amplitudes_real,
                                   double *partner_amplitudes_imaginary);
void exchange_amplitudes_between_gpus(int current_gpu_rank, int partner_gpu_rank,
                                      int num_amplitudes_per_gpu);

/* ------------------------------------------------------------------------ */
/* 2. dmv_* layer                                                           */
/* ------------------------------------------------------------------------ */

enum {
    DMV_OK = 0,
    DMV_ERR_ARG = 1,        /* invalid argument (qubit index, size, null pointer) */
    DMV_ERR_HIP = 2,        /* HIP runtime error */
    DMV_ERR_NO_DEVICE = 3,  /* no GPU visible */
    DMV_ERR_MEMORY = 4,     /* state does not fit in device memory */
    DMV_ERR_STATE = 5,      /* call out of order (e.g. no state) */
    DMV_ERR_UNSUPPORTED = 6 /* feature outside this library (multi-GPU) */
};

enum { DMV_C128 = 0, DMV_C64 = 1 };

/* Gate kinds.  0-9 are the reference gates (src/qubit_backend/gates.rs);
 * U3 and CZ are additions. */
enum {
    DMV_GATE_H = 0, DMV_GATE_X = 1, DMV_GATE_Y = 2, DMV_GATE_Z = 3,
    DMV_GATE_RX = 4, DMV_GATE_RY = 5, DMV_GATE_RZ = 6, DMV_GATE_CNOT = 7,
    DMV_GATE_S = 8, DMV_GATE_T = 9, DMV_GATE_ID = 10, DMV_GATE_U3 = 11, DMV_GATE_CZ = 12
};

typedef struct dmv_state dmv_state;
typedef struct dmv_circuit dmv_circuit;

/* Message of the last failed dmv_* or compatibility call on this thread. */
const char *dmv_last_error(void);
/* Library version string. */
const char *dmv_version(void);

/* Number of visible GPUs (0 without a device; no error). */
int dmv_device_count(int *count);
/* Free and total device memory in bytes of the current device. */
int dmv_device_memory(uint64_t *free_bytes, uint64_t *total_bytes);
/* Largest qubit count whose state (plus the library's workspace) fits in free
 * device memory for `dtype`.  Replaces the host-RAM check of circuit.rs:115-128. */
int dmv_max_qubits(int dtype, int *max_qubits);
/* Device bytes a state of num_qubits needs (amplitudes + workspace). */
int dmv_state_bytes(int num_qubits, int dtype, uint64_t *bytes);

/* ---- states (device-resident amplitudes, interleaved re/im) ---- */
int dmv_state_create(int num_qubits, int dtype, dmv_state **out);
int dmv_state_destroy(dmv_state *s);
int dmv_state_num_qubits(const dmv_state *s, int *num_qubits);
/* Sets the state to the basis state |basis_index>. */
int dmv_state_reset(dmv_state *s, uint64_t basis_index);
/* Copies the state to caller arrays of 2^n doubles each (split re/im). */
int dmv_state_get(dmv_state *s, double *re, double *im);
/* Loads caller amplitudes (2^n each); no normalisation. */
int dmv_state_set(dmv_state *s, const double *re, const double *im);
/* dst <- src (same size and dtype). */
int dmv_state_copy(dmv_state *dst, const dmv_state *src);
/* Raw device pointer of the amplitude buffer (2^n amplitudes) and its byte
 * size.  The order is swizzled (dmv_state_layout) so that strided passes
 * spread over the HBM channels; dmv_state_get/dmv_state_set always use plain
 * arrays. */
int dmv_state_device_ptr(dmv_state *s, void **ptr, uint64_t *bytes);
/* Physical pieces behind the amplitude buffer: 1 for one allocation, 8 when
 * the buffer is eight physical pieces 16 GiB apart mapped back to back (done
 * for states of 2 GiB up to 64 GiB when the device has 128 GiB free at
 * creation; passes over such a state stream about 15% faster on MI355X).
 * Either way dmv_state_device_ptr is one contiguous buffer.
 * No counterpart in the reference. */
int dmv_state_pieces(const dmv_state *s, int *pieces);
/* The element order of dmv_state_device_ptr.  *shift = 0: plain order.
 * Otherwise, with s = i >> shift, amplitude i is element
 *     i ^ (P(s) << block_bits) ^ (F(s >> 6) << shift),
 * where P maps bit j of its argument to bit (7 j mod 11) and F maps bit j to
 * bit (6 - fold_bits + j mod fold_bits) (F = 0 when fold_bits = 0); both are
 * XOR-linear.  shift = 17, block_bits = 6 for complex128 (2 MiB spans of 1 KiB
 * blocks), 18 and 7 for complex64, on states of more than one span.  By
 * default fold_bits = 6 on states of 2^15 spans and more (32 qubits in
 * complex128, 33 in complex64) and 0 below.  Each state keeps the order it
 * was created with; operations on two states require equal orders.
 * Any pointer argument may be NULL.
 * No counterpart in the reference (its states are plain split re/im arrays). */
int dmv_state_layout(const dmv_state *s, int *shift, int *block_bits, int *fold_bits);

/* ---- single operations (one kernel launch each) ---- */
/* Row-major 2x2 (re[4], im[4]) on target, controlled on every qubit of
 * control_mask (0 for none).  Mirrors apply_one_qubit_gate_kernel_local
 * (damavand-gpu/kernels.cu:120-172) without its in-place race. */
int dmv_apply_matrix(dmv_state *s, const double *m_re, const double *m_im, uint64_t control_mask,
                     int target);
/* A gate by kind (params: theta for RX/RY/RZ; theta, phi, lambda for U3). */
int dmv_apply_gate(dmv_state *s, int kind, int control, int target, const double *params);

/* ---- measurement ---- */
/* |a_i|^2 for all i into a host buffer of 2^n doubles (circuit.rs:565-589).
 * This and dmv_state_get copy more than 64 KiB through a 16 MiB pinned host
 * buffer of the state (allocated on first use, freed with the state). */
int dmv_probabilities(dmv_state *s, double *host_probabilities);
/* sum |a_i|^2. */
int dmv_norm_sqr(dmv_state *s, double *out);
/* Exact <Z_q> for q = 0..n-1 into out[n]. */
int dmv_expect_z(dmv_state *s, double *out);
/* Exact <H> for H = sum_q a[q] Z_q + sum_t b[t] Z_{pi[t]} Z_{pj[t]}; a has n entries (may be null). */
int dmv_expect_zz(dmv_state *s, const double *a, int n_pairs, const int *pi, const int *pj,
                  const double *b, double *out);
/* Exact <H> for a diagonal H over up to 12 qubits: H|x> = table[j(x)] |x>
 * with j(x) = sum_k bit_{wires[k]}(x) 2^k; table has 2^n_wires entries and
 * the wires are distinct.  Covers any product of Z on those qubits (table
 * entries +-1), sums of them, and projectors (marginal probabilities).
 * No reference counterpart (extract_expectation_values, circuit.rs:494-513,
 * reads single qubits from samples). */
int dmv_expect_diag(dmv_state *s, int n_wires, const int *wires, const double *table, double *out);
/* <a|b> (complex, into out[2]). */
int dmv_inner_product(dmv_state *a, dmv_state *b, double *out_re_im);
/* k samples with the device splitmix64 stream of `seed`: the uniform of
 * sample s is ((mix64(seed + (s+1)*0x9E3779B97F4A7C15) >> 11) + 1) * 2^-53.
 * Semantics of src/utils.rs:258-277: the smallest j with u*total <= cum[j+1]. */
int dmv_sample(dmv_state *s, int64_t k, uint64_t seed, int64_t *out);
/* The same with caller-given uniforms in (0, 1]. */
int dmv_sample_uniforms(dmv_state *s, int64_t k, const double *u, int64_t *out);
/* dmv_sample without any device allocation (no hipMalloc or hipFree, no
 * allocating library call): the same samples for the same state, k and seed.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under stream capture, serving loops, harnesses that forbid it).
 * The k results are staged in a fixed device array of 2^20 int64 (8 MiB,
 * held on each GPU used until the process ends); k > 2^20 returns
 * DMV_ERR_ARG.  Block masses and their prefix use the state's own workspace,
 * allocated by dmv_state_create.  All calls on one GPU share the staging
 * array, which keeps the last call's results; the library serialises these
 * calls with a lock, and calls from other processes have their own copy. */
int dmv_sample_static(dmv_state *s, int64_t k, uint64_t seed, int64_t *out);

/* ---- circuits (gate list + fused execution plan) ---- */
int dmv_circuit_create(int num_qubits, dmv_circuit **out);
int dmv_circuit_destroy(dmv_circuit *c);
/* Appends a gate (circuit.rs:608-691).  control = -1 for none; params as in
 * dmv_apply_gate; is_observable marks a Pauli gate as an observable that
 * forward() skips (circuit.rs:347-349).  Returns the gate index in *index. */
int dmv_circuit_add_gate(dmv_circuit *c, int kind, int control, int target, const double *params,
                         int is_observable, int *index);
int dmv_circuit_num_gates(const dmv_circuit *c, int *count);
/* Number of parametrised gates (RX, RY, RZ), in insertion order (circuit.rs:308-322). */
int dmv_circuit_num_parameters(const dmv_circuit *c, int *count);
/* set_parameters semantics of circuit.rs:308-322: zip(parametrised gates, p),
 * truncated to the shorter. */
int dmv_circuit_set_parameters(dmv_circuit *c, const double *p, int count);
int dmv_circuit_get_parameters(const dmv_circuit *c, double *p, int count);
/* Target qubits of the observable gates, in insertion order (circuit.rs:494-513). */
int dmv_circuit_observables(const dmv_circuit *c, int *targets, int capacity, int *count);
/* Removes every gate (circuit.rs:303-306 clears the gate list). */
int dmv_circuit_clear(dmv_circuit *c);

/* forward() (circuit.rs:341-375): applies the non-observable gates in order
 * to the state, without resetting it.  mode 0: fused passes (the default
 * path); mode 1: one kernel per gate. */
enum { DMV_FORWARD_FUSED = 0, DMV_FORWARD_PER_GATE = 1 };
int dmv_circuit_forward(dmv_circuit *c, dmv_state *s, int mode);

/* Adjoint gradient of E = <psi|H|psi>, psi = U(theta)|0>, with H the Z/ZZ
 * observable of dmv_expect_zz.  psi and lambda are two states of the
 * circuit's size and dtype (their contents are overwritten).  Writes E to
 * *energy and dE/dtheta_k for the k-th parametrised gate into grad[k].
 * No reference counterpart (circuit.rs:529-546 has only a commented-out
 * parameter-shift). */
int dmv_circuit_adjoint(dmv_circuit *c, dmv_state *psi, dmv_state *lambda, const double *a,
                        int n_pairs, const int *pi, const int *pj, const double *b,
                        double *energy, double *grad);

/* dmv_circuit_adjoint with the diagonal observable of dmv_expect_diag.  With
 * table = the upstream gradient of a vector of marginal probabilities, grad
 * is that vector's vector-Jacobian product (one reverse sweep). */
int dmv_circuit_adjoint_diag(dmv_circuit *c, dmv_state *psi, dmv_state *lambda, int n_wires, const int *wires,
                             const double *table, double *energy, double *grad);

/* Fused-plan statistics for a circuit on a state size/dtype: number of HBM
 * passes, total phases and ops, and algorithmic HBM bytes of one forward. */
int dmv_circuit_plan_stats(dmv_circuit *c, int dtype, int *passes, int *phases, int *ops,
                           double *hbm_bytes);

/* The fused plan's operation sequence, for checking the scheduler on the CPU
 * (no device needed): for op i, kind[i] (0 general, 1 diagonal, 2 swap),
 * target[i], control_mask[i], matrix m[8*i .. 8*i+7] (re/im of m00, m01, m10,
 * m11) and pass[i]; *count receives the number of ops (up to capacity).  The
 * ops applied in this order reproduce the circuit. */
int dmv_circuit_plan_ops(dmv_circuit *c, int dtype, int tile_qubits, int *kind, int *target,
                         uint64_t *control_mask, double *m, int *pass, int capacity, int *count);

/* The complete device plan (passes, phases with their register/thread
 * layouts, ops with their masks) as JSON text, for checking the tile
 * executor's index arithmetic on the CPU.  adjoint selects the reverse-sweep
 * plan; tile_qubits / threads_log2 / max_gen_per_pass of 0 take the tuned
 * defaults.  Writes at most `capacity` bytes (NUL-terminated) and the full
 * length into *length.  No device needed. */
int dmv_circuit_plan_json(dmv_circuit *c, int dtype, int adjoint, int tile_qubits, int threads_log2,
                          int max_gen_per_pass, char *buffer, uint64_t capacity, uint64_t *length);

/* ---- batched forward + adjoint: many small circuits in one persistent launch ---- */
/* Circuit i has n[i] qubits and the gates gate_kind/control/target/param in
 * [gate_offset[i], gate_offset[i+1]); its observable is the Z/ZZ chain
 * H = sum_q a_q Z_q + sum_{q<n-1} b_q Z_q Z_{q+1} with a_q at
 * coef_offset[i] + q and b_q at coef_offset[i] + n[i] + q.  Outputs:
 * energy[i] and the gradient of its parametrised gates at
 * grad[param_offset[i] + k].  gate_offset has count + 1 non-decreasing
 * entries starting at >= 0, coef_offset entries are >= 0, and param_offset
 * must be the exclusive prefix sum of the circuits' parameter counts
 * (param_offset[0] = 0), so grad holds sum_i (parameters of circuit i)
 * values; anything else returns DMV_ERR_ARG before the device is touched.
 * Circuits have 2 to 24 qubits, and a circuit's reverse sweep may hold at
 * most 128 parametrised gates in one fused pass (the kernel keeps their
 * per-wave sums in LDS); a batch with more returns DMV_ERR_ARG. */
int dmv_batch_forward_adjoint(int count, int dtype, const int *n, const int *gate_offset,
                              const int *gate_kind, const int *gate_control,
                              const int *gate_target, const double *gate_param,
                              const int *coef_offset, const double *coef,
                              const int *param_offset, double *energy, double *grad);

/* ---- profiling ---- */
/* Kernel classes timed with HIP events when profiling is on. */
enum {
    DMV_KC_TILE = 0,     /* fused tile pass */
    DMV_KC_GATE = 1,     /* single-gate kernel */
    DMV_KC_SMALL = 2,    /* whole-state-in-LDS kernel */
    DMV_KC_REDUCE = 3,   /* norm / <Z> / <H> / probabilities */
    DMV_KC_SAMPLE = 4,   /* sampler */
    DMV_KC_ADJOINT = 5,  /* adjoint tile pass */
    DMV_KC_BATCH = 6,    /* batched persistent kernel */
    DMV_KC_COPY = 7,     /* device copy (bandwidth reference) */
    DMV_KC_COUNT = 8
};
int dmv_profile_enable(int on);
int dmv_profile_reset(void);
/* Launch count, summed milliseconds and summed algorithmic HBM bytes of a class. */
int dmv_profile_get(int kernel_class, int64_t *launches, double *total_ms, double *bytes);
/* Reads and writes `bytes` of device memory with a streaming copy kernel and
 * returns its duration; the measured HBM copy ceiling. */
int dmv_copy_benchmark(uint64_t bytes, int repeats, double *ms_per_copy);
/* Blocks until all library work on the device is done. */
int dmv_synchronize(void);

/* ---- tuning ---- */
/* Sets a tunable, e.g. "tile_qubits" / "adjoint_tile_qubits" (9..13),
 * "tile_threads_log2" (6..9), "max_gen_per_pass" (default: 8 or 12, chosen per
 * circuit for forward plans by their pass and phase counts; 8 for reverse
 * sweeps), "nontemporal" (0/1: streaming loads and stores; default 1 for
 * forward passes over states of 2 GiB and more, for reverse sweeps whose
 * two states together take 512 MiB and more, and for batches whose scratch
 * states take 512 MiB and more),
 * "gate_pairs_per_thread" (1, 2, 4), "batch_tile_qubits",
 * "batch_threads_log2", "batch_wgs_per_cu", "graph" (0/1: replay forward
 * passes from a captured hipGraph), "half_lds" (0/1, default 1: forward
 * passes of the default tile shapes split their LDS transposes in two
 * rounds through half a tile of LDS, so three workgroups share a CU).
 * A negative value unsets the tunable, so its default applies again;
 * dmv_tune_get returns -1 when unset. */
int dmv_tune_set(const char *name, int64_t value);
int dmv_tune_get(const char *name, int64_t *value);
/* Autotuning of launch shapes.  The reference chooses grid and block sizes
 * by an occupancy query before every launch (damavand-gpu/utils.cuh:62-133);
 * here each shape is measured once, on the caller's workload, and reused:
 *  - the choice is kept by the circuit (forward, adjoint) until
 *    dmv_circuit_clear, and in a process-wide table keyed by
 *    "<arch>/n<qubits>/<c128|c64>/<forward|adjoint|batch>";
 *  - the table is read on first use from the JSON-lines file named by the
 *    environment variable DMV_TUNE_CACHE (later lines win), and every new
 *    choice is appended to that file, so later processes start from it;
 *  - a plan takes each parameter from the first of: the circuit's own
 *    choice, the tunables (dmv_tune_set / DMV_<NAME>), the table, the
 *    defaults.
 * The candidates run on scratch states the call allocates (the caller's
 * states are left as they were) and no tunable is changed. */
/* Times the fused forward of `c` for tile shapes x GEN budgets per pass
 * (then streaming stores on the best); writes the chosen tile qubits. */
int dmv_autotune(dmv_circuit *c, dmv_state *s, int *chosen_tile_qubits);
/* The same for the reverse sweep of dmv_circuit_adjoint in `dtype`. */
int dmv_autotune_adjoint(dmv_circuit *c, int dtype, int *chosen_tile_qubits);
/* Times the batched kernel (tile shape x workgroups per CU) on this batch
 * (arguments as dmv_batch_forward_adjoint) and keeps the fastest in the
 * table under the qubits of its largest circuit. */
int dmv_batch_autotune(int count, int dtype, const int *n, const int *gate_offset, const int *gate_kind,
                       const int *gate_control, const int *gate_target, const double *gate_param,
                       const int *coef_offset, const double *coef, const int *param_offset,
                       int *chosen_tile_qubits);
/* Empties the tuning table; the next lookup reads DMV_TUNE_CACHE again.
 * Circuits keep their own choices. */
int dmv_tune_forget(void);
/* The tile shape the next forward (adjoint = 0) or reverse-sweep plan of `c`
 * in `dtype` will use, after the precedence above.  No device needed. */
int dmv_circuit_tile_shape(dmv_circuit *c, int dtype, int adjoint, int *tile_qubits, int *threads_log2,
                           int *max_gen_per_pass, int *nontemporal);

#ifdef __cplusplus
}
#endif

#endif /* DAMAVAND_GPU_H */
