/*
 * cpu_ref.h -- TEST INFRASTRUCTURE: the CPU oracle of the parity tests.
 *
 * A plain C++ restatement of MITHRA's single-domain FDTD/PIC time march, used only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to check and time the
 * HIP path.  Nothing in the product links it.  Its entry points mirror the device C-ABI
 * of include/mithra_gpu.h one to one (same parameter structs, same host layouts) so a test
 * drives both with the same call sequence and compares.
 *
 * Every routine keeps the reference's arithmetic in source order (the library is built
 * with -O3 -ffp-contract=off on x86-64, no FMA), so on identical inputs it reproduces the
 * reference bit for bit; tests/test_oracle_golden.py pins that against the fixtures the
 * real reference wrote (tests/golden).  Work that is independent per node or per
 * particle runs under OpenMP, which does not change any result; the current deposition
 * and the power reduction stay sequential in the reference's order.
 */
#ifndef MITHRA_CPU_REF_H_
#define MITHRA_CPU_REF_H_

#include "../include/mithra_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_dev oracle_dev;

int oracle_create(const mithra_grid* grid, const mithra_coeffs* coeffs, oracle_dev** out);
int oracle_destroy(oracle_dev* dev);
int oracle_set_threads(int n);
int oracle_set_potentials(oracle_dev* dev, int level, const double* a, const double* phi);
int oracle_get_potentials(oracle_dev* dev, int level, double* a, double* phi);
int oracle_get_fields(oracle_dev* dev, float* e, float* b);
int oracle_set_particles(oracle_dev* dev, const double* charge11, long n);
int oracle_get_particles(oracle_dev* dev, double* charge11, long n);
long oracle_particle_count(const oracle_dev* dev);
int oracle_field_update(oracle_dev* dev, double time, const mithra_seed* seed);
int oracle_field_evaluate(oracle_dev* dev);
int oracle_set_beams(oracle_dev* dev, const mithra_beam* optical, int n_optical, const mithra_beam* external,
                     int n_external);
void oracle_beam_field(const mithra_bunch_params* p, const mithra_beam* beam, const double r[3], double time_bunch,
                       double et[3], double bt[3]);
int oracle_bunch_begin(oracle_dev* dev);
int oracle_bunch_update(oracle_dev* dev, const mithra_bunch_params* p, const mithra_undulator* und,
                        int n_und, double time_bunch);
int oracle_bunch_counters(oracle_dev* dev, long* transverse_out, long* z_crossings);
int oracle_power_setup(oracle_dev* dev, const mithra_power_params* p, unsigned rows);
int oracle_power_sample(oracle_dev* dev, unsigned n_time, unsigned row);
int oracle_power_fetch(oracle_dev* dev, unsigned first, unsigned count, double* out);
int oracle_field_shift(oracle_dev* dev);
int oracle_current_reset(oracle_dev* dev);
int oracle_current_update(oracle_dev* dev);

int oracle_bunching(oracle_dev* dev, double k, double* b_out);

/* Known-answer helpers for single functions. */
double oracle_halton(unsigned i, unsigned j);
double oracle_signal(const mithra_seed* seed, double t, double phase);
void oracle_seed_fields(const mithra_seed* seed, const double r[3], double time, double a[3]);
void oracle_undulator_field(const mithra_bunch_params* p, const mithra_undulator* und, int n_und,
                            const double r[3], double time_bunch, double et[3], double bt[3]);

#ifdef __cplusplus
}
#endif

#endif
