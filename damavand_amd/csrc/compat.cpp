// Compatibility layer: the extern "C" symbols the reference's Rust crate
// binds (src/qubit_backend/circuit.rs:27-51, circuit_gpu.rs:3-24,
// circuit_distributed_gpu.rs:6-33), with the reference's argument types.
// They drive one process-global complex128 state on one GPU, as
// damavand-gpu/rust_communication.cu:19-22 does with its global vectors.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "kernels.hpp"
#include "runtime.hpp"

using namespace dmv;

namespace dmv {
int profile_get(int kc, int64_t *launches, double *ms, double *bytes);
}

static dmv_state *g_state = nullptr;

static void report(const char *where)
{
    fprintf(stderr, "damavand-gpu: %s: %s\n", where, last_error());
}

static int log2_exact(int n_amp)
{
    if (n_amp <= 0 || (n_amp & (n_amp - 1)) != 0)
        return -1;
    int q = 0;
    while ((1 << q) < n_amp)
        q++;
    return q;
}

static bool check_global(int num_amplitudes_per_gpu, const char *where)
{
    if (!g_state) {
        set_error("no state: init_quantum_state was not called");
        report(where);
        return false;
    }
    if (num_amplitudes_per_gpu != (1 << g_state->n)) {
        set_error("num_amplitudes_per_gpu %d does not match the state (%d)", num_amplitudes_per_gpu,
                  1 << g_state->n);
        report(where);
        return false;
    }
    return true;
}

extern "C" {

int get_number_of_available_gpus(void)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return c;
}

double get_memory_for_gpu(int gpu_id)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, gpu_id) != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipGetDeviceProperties(%d) failed", gpu_id);
        report("get_memory_for_gpu");
        return 0.0;
    }
    return (double)prop.totalGlobalMem / 1048576.0;
}

int peer_access_allowed(int source_gpu_id, int target_gpu_id)
{
    int ok = 0;
    if (source_gpu_id == target_gpu_id)
        return 0;
    if (hipDeviceCanAccessPeer(&ok, source_gpu_id, target_gpu_id) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return ok;
}

void print_timers(void)
{
    static const char *names[DMV_KC_COUNT] = {"tile", "gate", "small", "reduce", "sample", "adjoint", "batch", "copy"};
    for (int kc = 0; kc < DMV_KC_COUNT; kc++) {
        int64_t launches = 0;
        double ms = 0, bytes = 0;
        if (profile_get(kc, &launches, &ms, &bytes) != DMV_OK || launches == 0)
            continue;
        printf("%s launches %lld mean %.6f ms algorithmic %.3f GB/s\n", names[kc], (long long)launches,
               ms / launches, ms > 0 ? bytes / (ms * 1e6) : 0.0);
    }
}

void init_quantum_state(int num_amplitudes_per_gpu, int num_gpus_per_node, int is_first_node)
{
    const int n = log2_exact(num_amplitudes_per_gpu);
    if (n < 1 || n > 30) {
        set_error("num_amplitudes_per_gpu must be a power of two in [2, 2^30], got %d", num_amplitudes_per_gpu);
        report("init_quantum_state");
        return;
    }
    if (num_gpus_per_node != 1) {
        set_error("the state lives on one GPU; num_gpus_per_node must be 1, got %d", num_gpus_per_node);
        report("init_quantum_state");
        return;
    }
    if (g_state && g_state->n != n) {
        state_destroy(g_state);
        g_state = nullptr;
    }
    if (!g_state && state_create(n, DMV_C128, &g_state) != DMV_OK) {
        g_state = nullptr;
        report("init_quantum_state");
        return;
    }
    if (state_reset(g_state, 0, is_first_node == 0) != DMV_OK)
        report("init_quantum_state");
}

void apply_one_qubit_gate_gpu_local(double *gate_matrix_real, double *gate_matrix_imaginary, int num_qubits,
                                    int num_amplitudes_per_gpu, int control_qubit, int target_qubit)
{
    if (!check_global(num_amplitudes_per_gpu, "apply_one_qubit_gate_gpu_local"))
        return;
    if (num_qubits != g_state->n) {
        set_error("num_qubits %d does not match the state (%d)", num_qubits, g_state->n);
        report("apply_one_qubit_gate_gpu_local");
        return;
    }
    const uint64_t mask = control_qubit >= 0 ? (uint64_t(1) << control_qubit) : 0;
    if (control_qubit >= num_qubits) {
        set_error("control qubit %d out of range", control_qubit);
        report("apply_one_qubit_gate_gpu_local");
        return;
    }
    if (dmv_apply_matrix(g_state, gate_matrix_real, gate_matrix_imaginary, mask, target_qubit) != DMV_OK)
        report("apply_one_qubit_gate_gpu_local");
}

void apply_one_qubit_gate_gpu_distributed(double *, double *, int, int, int, int)
{
    set_error("apply_one_qubit_gate_gpu_distributed: multi-GPU partner updates are not part of this library");
    report("apply_one_qubit_gate_gpu_distributed");
}

void measure_on_gpu(int num_amplitudes_per_gpu, double *probabilities)
{
    if (!check_global(num_amplitudes_per_gpu, "measure_on_gpu"))
        return;
    if (!probabilities || state_probs(g_state, probabilities) != DMV_OK)
        report("measure_on_gpu");
}

void sequential_measure_on_gpu(int num_amplitudes_per_gpu, double *probabilities)
{
    measure_on_gpu(num_amplitudes_per_gpu, probabilities);
}

void concurrent_measure_on_gpu(int num_amplitudes_per_gpu, double *probabilities)
{
    measure_on_gpu(num_amplitudes_per_gpu, probabilities);
}

void retrieve_amplitudes_on_host(int num_amplitudes_per_gpu, double *local_amplitudes_real,
                                 double *local_amplitudes_imaginary)
{
    if (!check_global(num_amplitudes_per_gpu, "retrieve_amplitudes_on_host"))
        return;
    if (state_get(g_state, local_amplitudes_real, local_amplitudes_imaginary) != DMV_OK)
        report("retrieve_amplitudes_on_host");
}

void load_amplitudes_local_on_device(int num_amplitudes_per_gpu, double *local_amplitudes_real,
                                     double *local_amplitudes_imaginary)
{
    if (!check_global(num_amplitudes_per_gpu, "load_amplitudes_local_on_device"))
        return;
    if (!local_amplitudes_real || !local_amplitudes_imaginary ||
        state_set(g_state, local_amplitudes_real, local_amplitudes_imaginary) != DMV_OK)
        report("load_amplitudes_local_on_device");
}

void split_amplitudes_between_gpus(int, double *, double *, double *, double *)
{
    set_error("split_amplitudes_between_gpus: the state lives on one GPU");
    report("split_amplitudes_between_gpus");
}

void exchange_amplitudes_between_gpus(int, int, int)
{
    set_error("exchange_amplitudes_between_gpus: the state lives on one GPU");
    report("exchange_amplitudes_between_gpus");
}

}  // extern "C"
