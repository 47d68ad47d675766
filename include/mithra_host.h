/*
 * mithra_host.h -- C-ABI of the MITHRA drop-in host (libmithra.so).
 *
 * Wraps the C++ host that keeps the reference's job-file parser (reference src/readdata.cpp,
 * src/datainput.cpp), its set-up (Solver::initialize, src/solver.cpp:547-595) and its main
 * loop (Solver::solve, src/solver.cpp:1212-1414) over the device C-ABI of mithra_gpu.h.  The
 * reference's only front end is `main` (src/mithra.cpp:32-102); these calls expose the same
 * steps to other programs (the parity tests and the benchmark bind them with ctypes).
 * Functions return 0 or a negative status; mithra_job_error() gives the last message.
 */
#ifndef MITHRA_HOST_H_
#define MITHRA_HOST_H_

#include "mithra_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mithra_job mithra_job;

/* read_file + cleanJobFile + ParseDarius::setJobParameters (src/mithra.cpp:48-75). */
int mithra_job_open(const char* path, mithra_job** out);
int mithra_job_close(mithra_job* job);
const char* mithra_job_error(const mithra_job* job);

/* Particles for a `type = other` bunch, which the reference leaves to an external program
 * (src/solver.cpp:1164-1165): n lab-frame records q x y z gbx gby gbz. */
int mithra_job_set_external_bunch(mithra_job* job, const double* charge7, long n);

/* Device generation (before initialize): on = 1 generates every `ellipsoid` bunch-initialization
 * on the GPU (mithra_bunch_ellipsoid) and runs the particle loops of Solver::lorentzBoostBunch
 * and Solver::shiftBackInTime there (src/solver.cpp:1130-1176, :263-423, :1185-1206); the records
 * then stay on the device and the domain takes them over without a host copy.  Other bunch types
 * than ellipsoid and `other` are refused in this mode.  No counterpart in the reference. */
int mithra_job_device_bunch(mithra_job* job, int on);
/* A `type = other` bunch handed over on the device (lab-frame records, e.g. from
 * mithra_bunch_flattop): switches device generation on; the job boosts the records in place and
 * does not own the bunch, which must outlive the job's set-up and upload. */
int mithra_job_set_external_device_bunch(mithra_job* job, mithra_bunch* bunch);

/* Solver::initialize (src/solver.cpp:547-595). */
int mithra_job_initialize(mithra_job* job);

/* Views of the initialised state in the device C-ABI's types. */
int mithra_job_grid(mithra_job* job, mithra_grid* out);
int mithra_job_coeffs(mithra_job* job, mithra_coeffs* out);
int mithra_job_seed(mithra_job* job, mithra_seed* out);
int mithra_job_bunch_params(mithra_job* job, mithra_bunch_params* out);
int mithra_job_undulators(mithra_job* job, mithra_undulator* out, int max);   /* returns the count */
/* times[0..5] = time_, timeBunch_, bunch time step, field time step, totalTime_, nUpdateBunch_;
 * returns the number of bunch sub-steps per field step. */
int mithra_job_times(mithra_job* job, double* times);
/* power[0..3] = N, Nl, Nf, pc; then z[N], w[Nl] follow when the buffer holds 4 + N + Nl
 * values.  Returns 0 when no radiation power is sampled, 1 otherwise. */
int mithra_job_power(mithra_job* job, double* power, int max);
long mithra_job_particle_count(mithra_job* job);
int mithra_job_particles(mithra_job* job, double* charge11, long n);
/* level 0 = A^n, 1 = A^{n-1}; phi may be NULL. */
int mithra_job_potentials(mithra_job* job, int level, double* a, double* phi);

/* The initialised pre-step state in a device domain made from mithra_job_grid/coeffs: A^n and
 * A^{n-1} (phi) when the seed set them (a new domain holds zeros, src/solver.cpp:828-839), and the
 * particles (handed over on the device when they were generated there). */
int mithra_job_fill_domain(mithra_job* job, mithra_dev* dev);

/* Solver::solve on the GPU for at most max_steps steps (all when negative), writing the
 * reference's output files.  Returns the number of steps run, or a negative status. */
long mithra_job_run(mithra_job* job, long max_steps, int precision, int quiet);

#ifdef __cplusplus
}
#endif

#endif
