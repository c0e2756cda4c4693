/* Replays, in C, the call sequence the reference's Rust crate makes through
 * its extern "C" block when it runs a circuit with apply_method = "gpu":
 *
 *   Circuit::new        get_number_of_available_gpus, get_memory_for_gpu,
 *                       init_quantum_state            (circuit.rs:131-195)
 *   Circuit::forward    apply_one_qubit_gate_gpu_local per gate, matrices as
 *                       re[4] / im[4] row-major        (circuit_gpu.rs:31-60)
 *   Circuit::measure    measure_on_gpu, print_timers   (circuit.rs:565-589)
 *   state getters       retrieve_amplitudes_on_host    (circuit.rs:380-405)
 *
 * Case 1 is the reference's own known answer (circuit.rs:811-836): H(0),
 * H(1), CNOT(0 -> 1) on 2 qubits gives 1/2 on every amplitude.  Case 2 is the
 * RX-layer workload of the reference's examples: L layers of RX(pi/(i+1)) on
 * qubit i leave a product state with <Z_i> = cos(L pi/(i+1)).  Case 3 checks
 * that re-initialisation gives |0...0> again.  Exit status 0 on success. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "damavand_gpu.h"

static int failures = 0;

static void check(int ok, const char *what, double got, double want)
{
    if (!ok) {
        fprintf(stderr, "FAIL %s: got %.17g want %.17g\n", what, got, want);
        failures++;
    }
}

/* gates.rs formulas, row-major [m00, m01, m10, m11] */
static void hadamard(double re[4], double im[4])
{
    const double s = 1.0 / sqrt(2.0);
    re[0] = s; re[1] = s; re[2] = s; re[3] = -s;
    im[0] = im[1] = im[2] = im[3] = 0.0;
}

static void pauli_x(double re[4], double im[4])
{
    re[0] = 0.0; re[1] = 1.0; re[2] = 1.0; re[3] = 0.0;
    im[0] = im[1] = im[2] = im[3] = 0.0;
}

static void rotation_x(double theta, double re[4], double im[4])
{
    const double c = cos(theta / 2.0), s = sin(theta / 2.0);
    re[0] = c; re[1] = 0.0; re[2] = 0.0; re[3] = c;
    im[0] = 0.0; im[1] = -s; im[2] = -s; im[3] = 0.0;
}

int main(void)
{
    const int gpus = get_number_of_available_gpus();
    check(gpus >= 1, "get_number_of_available_gpus", gpus, 1);
    const double mib = get_memory_for_gpu(0);
    check(mib > 1024.0, "get_memory_for_gpu (MiB)", mib, 1024.0);

    /* case 1: circuit.rs:811-836 */
    {
        const int n = 2, n_amp = 1 << n;
        double re[4], im[4], probs[4], ar[4], ai[4];
        init_quantum_state(n_amp, 1, 1);
        hadamard(re, im);
        apply_one_qubit_gate_gpu_local(re, im, n, n_amp, -1, 0);
        apply_one_qubit_gate_gpu_local(re, im, n, n_amp, -1, 1);
        pauli_x(re, im); /* CNOT = X on the target under the control */
        apply_one_qubit_gate_gpu_local(re, im, n, n_amp, 0, 1);
        measure_on_gpu(n_amp, probs);
        retrieve_amplitudes_on_host(n_amp, ar, ai);
        for (int i = 0; i < n_amp; i++) {
            check(fabs(ar[i] - 0.5) <= 1e-15 && fabs(ai[i]) <= 1e-15, "KAT amplitude", ar[i], 0.5);
            check(fabs(probs[i] - 0.25) <= 1e-15, "KAT probability", probs[i], 0.25);
        }
        print_timers();
    }

    /* case 2: RX layers, closed form */
    {
        const int n = 16, n_amp = 1 << n, layers = 3;
        double re[4], im[4];
        double *probs = (double *)malloc(sizeof(double) * n_amp);
        init_quantum_state(n_amp, 1, 1);
        for (int l = 0; l < layers; l++)
            for (int q = 0; q < n; q++) {
                rotation_x(M_PI / (q + 1), re, im);
                apply_one_qubit_gate_gpu_local(re, im, n, n_amp, -1, q);
            }
        measure_on_gpu(n_amp, probs);
        double total = 0.0;
        for (int j = 0; j < n_amp; j++)
            total += probs[j];
        check(fabs(total - 1.0) <= 1e-12, "sum of probabilities", total, 1.0);
        for (int q = 0; q < n; q++) {
            double z = 0.0;
            for (int j = 0; j < n_amp; j++)
                z += ((j >> q) & 1) ? -probs[j] : probs[j];
            const double want = cos(layers * M_PI / (q + 1));
            check(fabs(z - want) <= 1e-12, "<Z_q> of the RX layers", z, want);
        }
        free(probs);
    }

    /* case 3: re-initialisation */
    {
        const int n = 3, n_amp = 1 << n;
        double re[4], im[4], ar[8], ai[8];
        init_quantum_state(n_amp, 1, 1);
        hadamard(re, im);
        apply_one_qubit_gate_gpu_local(re, im, n, n_amp, -1, 2);
        init_quantum_state(n_amp, 1, 1);
        retrieve_amplitudes_on_host(n_amp, ar, ai);
        for (int i = 0; i < n_amp; i++)
            check(ar[i] == (i == 0 ? 1.0 : 0.0) && ai[i] == 0.0, "reset amplitude", ar[i], i == 0);
    }

    if (failures) {
        fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    printf("compat harness: all checks passed\n");
    return 0;
}
