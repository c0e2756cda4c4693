/*
 * CPU oracle for the damavand statevector path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C with OpenMP, the reference's `multithreading`
 * backend and the `Circuit` semantics around it. It is the checker that the
 * HIP library is compared against, and the CPU baseline that bench.py times.
 * Nothing in the product library links, loads or calls it; only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
 *
 * Build: `make -C oracle` (gcc, -O2 -fopenmp -ffp-contract=off so that no FMA
 * contraction changes the rounding of the num-complex arithmetic restated here).
 *
 * Every function cites the reference lines it follows (paths relative to the
 * reference repository root).  Parity status of each function is listed in the
 * oracle/__init__.py header together with the checks that pin it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    double re, im;
} orc_cplx;

/* num-complex `Mul` for Complex<f64>: (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re).
 * Used at src/qubit_backend/circuit_multithreading.rs:47,49. */
static inline orc_cplx orc_mul(orc_cplx a, orc_cplx b)
{
    orc_cplx r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}

static inline orc_cplx orc_add(orc_cplx a, orc_cplx b)
{
    orc_cplx r;
    r.re = a.re + b.re;
    r.im = a.im + b.im;
    return r;
}

/* Gate kinds, numbered as in include/damavand_gpu.h (DMV_GATE_*). */
enum {
    ORC_H = 0, ORC_X = 1, ORC_Y = 2, ORC_Z = 3, ORC_RX = 4, ORC_RY = 5, ORC_RZ = 6,
    ORC_CNOT = 7, ORC_S = 8, ORC_T = 9, ORC_ID = 10, ORC_U3 = 11, ORC_CZ = 12
};

/*
 * Row-major 2x2 matrix [m00, m01, m10, m11] of a gate, as re[4], im[4].
 * Follows src/qubit_backend/gates.rs: Identity :40-49, Hadamard :98-119,
 * CNOT :170-174 (the X block), PauliX :224-233, PauliY :289-298,
 * PauliZ :353-362, RotationX :418-439, RotationY :490-511, RotationZ :562-577,
 * S :627-636, T :685-697.  U3 and CZ have no reference counterpart: U3 is the
 * textbook U3(theta, phi, lambda) = [[cos t/2, -e^{i l} sin t/2],
 * [e^{i p} sin t/2, e^{i(p+l)} cos t/2]] and CZ is a controlled Z.
 * Returns 0, or -1 for an unknown kind.
 */
int orc_gate_matrix(int kind, const double *p, double *re, double *im)
{
    double t = p ? p[0] : 0.0;
    memset(re, 0, 4 * sizeof(double));
    memset(im, 0, 4 * sizeof(double));
    switch (kind) {
    case ORC_ID:
        re[0] = 1.; re[3] = 1.;
        return 0;
    case ORC_H: {
        double s = 1. / sqrt(2.0);
        re[0] = s; re[1] = s; re[2] = s; re[3] = -1. / sqrt(2.0);
        return 0;
    }
    case ORC_X:
    case ORC_CNOT:
        re[1] = 1.; re[2] = 1.;
        return 0;
    case ORC_Y:
        im[1] = -1.; im[2] = 1.;
        return 0;
    case ORC_Z:
    case ORC_CZ:
        re[0] = 1.; re[3] = -1.;
        return 0;
    case ORC_RX:
        re[0] = cos(t / 2.); im[1] = -sin(t / 2.);
        im[2] = -sin(t / 2.); re[3] = cos(t / 2.);
        return 0;
    case ORC_RY:
        re[0] = cos(t / 2.); re[1] = -sin(t / 2.);
        re[2] = sin(t / 2.); re[3] = cos(t / 2.);
        return 0;
    case ORC_RZ:
        re[0] = cos(-t / 2.); im[0] = sin(-t / 2.);
        re[3] = cos(t / 2.); im[3] = sin(t / 2.);
        return 0;
    case ORC_S:
        re[0] = 1.; im[3] = 1.;
        return 0;
    case ORC_T:
        re[0] = 1.; re[3] = sqrt(2.) / 2.; im[3] = sqrt(2.) / 2.;
        return 0;
    case ORC_U3: {
        double th = p[0], ph = p[1], la = p[2];
        double c = cos(th / 2.), s = sin(th / 2.);
        re[0] = c;
        re[1] = -cos(la) * s; im[1] = -sin(la) * s;
        re[2] = cos(ph) * s; im[2] = sin(ph) * s;
        re[3] = cos(ph + la) * c; im[3] = sin(ph + la) * c;
        return 0;
    }
    default:
        return -1;
    }
}

/* Initial state |0...0>: src/qubit_backend/circuit.rs:172-175 (reset :263-266). */
void orc_init_zero(orc_cplx *s, int64_t n_amp)
{
    memset(s, 0, (size_t)n_amp * sizeof(orc_cplx));
    s[0].re = 1.;
}

/* Partner index of amplitude i for a gap of 2^target with one amplitude per
 * "node": src/qubit_backend/circuit.rs:781-795, called with
 * num_amplitudes_per_node = 1 at circuit_multithreading.rs:40-41. */
static inline int64_t orc_partner(int64_t i, int64_t gap)
{
    if (i % (2 * gap) < gap)
        return i + gap;
    return i - gap;
}

/*
 * One gate on the state, out of place via a full clone, exactly as
 * src/qubit_backend/circuit_multithreading.rs:9-54:
 *   partner = clone(state)                                   (:31)
 *   for every index i in parallel                            (:33)
 *     apply only if the control bit of i is 1 (or no control) (:35-39, :43)
 *     j = partner index (:40-41)
 *     j > i : a[i] = a[i]*m00 + p[j]*m01                     (:46-47)
 *     else  : a[i] = p[j]*m10 + a[i]*m11                     (:48-49)
 * `partner` is caller-provided scratch of n_amp amplitudes (the reference
 * allocates a fresh clone per gate).  control < 0 means no control.
 */
void orc_apply_gate(orc_cplx *state, orc_cplx *partner, int64_t n_amp,
                    const double *m_re, const double *m_im, int control, int target)
{
    orc_cplx m00 = {m_re[0], m_im[0]}, m01 = {m_re[1], m_im[1]};
    orc_cplx m10 = {m_re[2], m_im[2]}, m11 = {m_re[3], m_im[3]};
    int64_t gap = (int64_t)1 << target;
    memcpy(partner, state, (size_t)n_amp * sizeof(orc_cplx));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_amp; i++) {
        int64_t apply = control >= 0 ? ((i >> control) & 1) : 1;
        if (apply != 1)
            continue;
        int64_t j = orc_partner(i, gap);
        orc_cplx upper = partner[j];
        if (j > i)
            state[i] = orc_add(orc_mul(state[i], m00), orc_mul(upper, m01));
        else
            state[i] = orc_add(orc_mul(upper, m10), orc_mul(state[i], m11));
    }
}

/*
 * Generalisation of orc_apply_gate to a set of control qubits (all must be 1).
 * No reference counterpart: the reference gates carry at most one control
 * (gates.rs:13-21 get_control_qubit).  The update rule is the one above.
 */
void orc_apply_gate_mask(orc_cplx *state, orc_cplx *partner, int64_t n_amp,
                         const double *m_re, const double *m_im, uint64_t control_mask,
                         int target)
{
    orc_cplx m00 = {m_re[0], m_im[0]}, m01 = {m_re[1], m_im[1]};
    orc_cplx m10 = {m_re[2], m_im[2]}, m11 = {m_re[3], m_im[3]};
    int64_t gap = (int64_t)1 << target;
    memcpy(partner, state, (size_t)n_amp * sizeof(orc_cplx));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_amp; i++) {
        if (((uint64_t)i & control_mask) != control_mask)
            continue;
        int64_t j = orc_partner(i, gap);
        orc_cplx upper = partner[j];
        if (j > i)
            state[i] = orc_add(orc_mul(state[i], m00), orc_mul(upper, m01));
        else
            state[i] = orc_add(orc_mul(upper, m10), orc_mul(state[i], m11));
    }
}

/* Probabilities: src/qubit_backend/circuit.rs:579-581, norm_sqr = re*re + im*im. */
void orc_measure(const orc_cplx *s, int64_t n_amp, double *prob)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_amp; i++)
        prob[i] = s[i].re * s[i].re + s[i].im * s[i].im;
}

/*
 * Inverse-CDF sampling with injected uniforms u[s] in (0, 1].
 * src/utils.rs:269-277 builds cum = [0, p0, p0+p1, ...] by sequential
 * addition, then src/utils.rs:258-268 draws xsi = u * cum.last() and returns
 * index-1 for the first index with xsi <= cum[index] (None if there is none).
 * The reference rebuilds the cumulative for every sample (sample_local,
 * circuit.rs:475-476); the result is the same for a shared cumulative, so
 * this builds it once.  out[s] = -1 stands for the reference's None.
 */
void orc_sample_uniforms(const double *prob, int64_t n_amp, const double *u, int64_t k,
                         int64_t *out)
{
    double *cum = (double *)malloc((size_t)(n_amp + 1) * sizeof(double));
    cum[0] = 0.;
    for (int64_t i = 0; i < n_amp; i++)
        cum[i + 1] = cum[i] + prob[i];
    double total = cum[n_amp];
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < k; s++) {
        double xsi = u[s] * total;
        int64_t r = -1;
        for (int64_t idx = 0; idx <= n_amp; idx++) {
            if (xsi <= cum[idx]) {
                r = idx - 1;
                break;
            }
        }
        out[s] = r;
    }
    free(cum);
}

/* The reference's sampler cost model, for timing only: one cumulative of
 * n_amp+1 entries rebuilt per sample, single thread (src/utils.rs:269-277,
 * circuit.rs:475-476).  Same result as orc_sample_uniforms. */
void orc_sample_uniforms_rebuild(const double *prob, int64_t n_amp, const double *u,
                                 int64_t k, int64_t *out)
{
    double *cum = (double *)malloc((size_t)(n_amp + 1) * sizeof(double));
    for (int64_t s = 0; s < k; s++) {
        cum[0] = 0.;
        for (int64_t i = 0; i < n_amp; i++)
            cum[i + 1] = cum[i] + prob[i];
        double xsi = u[s] * cum[n_amp];
        int64_t r = -1;
        for (int64_t idx = 0; idx <= n_amp; idx++) {
            if (xsi <= cum[idx]) {
                r = idx - 1;
                break;
            }
        }
        out[s] = r;
    }
    free(cum);
}

/*
 * Exact <Z_q> for every qubit q: sum_j p_j (1 - 2 bit_q(j)).  No reference
 * counterpart (the reference only reads +-1 from samples, circuit.rs:494-513);
 * this is the exact value those sample averages estimate.
 */
void orc_expect_z(const orc_cplx *s, int n_qubits, double *out)
{
    int64_t n_amp = (int64_t)1 << n_qubits;
    for (int q = 0; q < n_qubits; q++) {
        double acc = 0.;
#pragma omp parallel for reduction(+ : acc) schedule(static)
        for (int64_t i = 0; i < n_amp; i++) {
            double p = s[i].re * s[i].re + s[i].im * s[i].im;
            acc += ((i >> q) & 1) ? -p : p;
        }
        out[q] = acc;
    }
}

/*
 * Exact <H> for a diagonal Pauli-Z observable
 *   H = sum_i a[i] Z_i + sum_t b[t] Z_{pi[t]} Z_{pj[t]}.
 * No reference counterpart; h(j) evaluated term by term.
 */
double orc_expect_zz(const orc_cplx *s, int n_qubits, const double *a, int n_pairs,
                     const int *pi, const int *pj, const double *b)
{
    int64_t n_amp = (int64_t)1 << n_qubits;
    double acc = 0.;
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (int64_t i = 0; i < n_amp; i++) {
        double h = 0.;
        for (int q = 0; q < n_qubits; q++)
            h += ((i >> q) & 1) ? -a[q] : a[q];
        for (int t = 0; t < n_pairs; t++) {
            int par = (int)(((i >> pi[t]) ^ (i >> pj[t])) & 1);
            h += par ? -b[t] : b[t];
        }
        double p = s[i].re * s[i].re + s[i].im * s[i].im;
        acc += h * p;
    }
    return acc;
}

/* Number of OpenMP threads the oracle runs with (1 without OpenMP). */
int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0)
        omp_set_num_threads(n);
#else
    (void)n;
#endif
}
