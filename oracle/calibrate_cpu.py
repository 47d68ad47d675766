"""CPU-baseline calibration (test infrastructure; run where the reference is built, never timed
by the product): seconds per step of the reference itself (oracle/_ref/ref_harness, the
unmodified reference with a step clock, under MPICH's mpiexec with 1, 2, 4 and 8 ranks, z-slab
domain decomposition as the reference does it, src/solver.cpp:619-641) beside the CPU oracle
(oracle/_build/libcpu_ref.so) on 1 and 8 threads, on the same job files:

  c1        64 x 64 x 256 nodes, 2e5 particles (the reference's CPU-runnable configuration)
  c2_slice  256 x 256 x 256 nodes, 2,097,152 particles: the production mesh's cross-section and
            cells over an eighth of its 2048 planes, with the particle density kept

Both bunches are the reference's own Gaussian ellipsoid (halton quiet start), which the drop-in
host reproduces bit for bit, so both sides march identical states.  Writes the table to
profiles/cpu_baseline_calibration.md.

Usage:  python oracle/calibrate_cpu.py [--steps N]
"""
import argparse
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

MPIEXEC = "/opt/conda/bin/mpiexec"
HARNESS = os.path.join(HERE, "_ref", "ref_harness")

JOB = """MESH
{{
  length-scale            = MICROMETER
  time-scale              = PICOSECOND
  mesh-lengths            = ( {lx:.10g}, {lx:.10g}, {lz:.10g} )
  mesh-resolution         = ( {dx:.10g}, {dx:.10g}, {dz:.10g} )
  mesh-center             = ( 0.0, 0.0, 0.0 )
  total-time              = 1.0e6
  mesh-truncation-order   = 2
  space-charge            = false
  solver                  = NSFD
}}
BUNCH
{{
  bunch-initialization
  {{
    type                  = ellipsoid
    distribution          = gaussian
    charge                = {charge:.10g}
    number-of-particles   = {n}
    gamma                 = {gamma:.10g}
    direction             = ( 0.0, 0.0, 1.0 )
    position              = ( 0.0, 0.0, 0.0 )
    sigma-position        = ( {sr:.10g}, {sr:.10g}, {sz:.10g} )
    sigma-momentum        = ( {sgb:.10g}, {sgb:.10g}, {sgbz:.10g} )
    transverse-truncation = {tt:.10g}
    longitudinal-truncation = {lt:.10g}
    bunching-factor       = 0.0
  }}
}}
UNDULATOR
{{
  static-undulator
  {{
    undulator-parameter   = {K:.10g}
    period                = {lu:.10g}
    length                = 300
    polarization-angle    = 0.0
  }}
}}
FEL-OUTPUT
{{
  radiation-power
  {{
    sample                = true
    type                  = at-point
    directory             = ./
    base-name             = power-sampling/power
    plane-position        = {zp:.10g}
    normalized-frequency  = 1.00
  }}
}}
"""


def job_text(nodes, d, n, gamma, K, lu, sr, sz, sgb, zp, charge):
    nx, nz = nodes
    dx, dz = d
    return JOB.format(lx=(nx - 1.5) * dx, lz=(nz - 1.5) * dz, dx=dx, dz=dz, charge=charge, n=n, gamma=gamma,
                      sr=sr, sz=sz, sgb=sgb, sgbz=1e-4 * gamma, tt=3 * sr, lt=3 * sz, K=K, lu=lu, zp=zp)


WORKLOADS = {
    # the reference's CPU-runnable configuration: 100 MeV, sigma 100 um / 50 um, 1 mm mrad
    "c1": dict(nodes=(64, 256), d=(380.0, 1.6), n=200_000, gamma=196.7, K=1.0, lu=1.5e4, sr=100.0, sz=50.0,
               sgb=0.01, zp=150.0, charge=6.2415e8),
    # the production mesh (256 x 256 x 2048 nodes, 16.8M particles, 1 GeV, K 1.4, lambda_u 1.8 cm) with
    # the reference's own Gaussian ellipsoid in place of the flat-top slice
    "c2": dict(nodes=(256, 2048), d=(1.2, 2.0e-4), n=16_777_216, gamma=1957.9, K=1.4, lu=1.8e4, sr=50.0,
               sz=0.05, sgb=0.02, zp=0.1977, charge=1.75e7),
    # the production mesh's cells and cross-section, 256 of its 2048 planes, 1/8 of its particles
    "c2_slice": dict(nodes=(256, 256), d=(1.2, 2.0e-4), n=2_097_152, gamma=1957.9, K=1.4, lu=1.8e4, sr=50.0,
                     sz=0.005, sgb=0.02, zp=0.02, charge=2.2e6),
}


def time_reference(workdir, text, ranks, steps, timeout=3600):
    """Seconds per step of the reference on `ranks` MPI ranks, steps 1..steps of the job `text`."""
    path = os.path.join(workdir, "job.job")
    with open(path, "w") as f:
        f.write(text)
    clock = os.path.join(workdir, "clock.txt")
    env = dict(os.environ, MITHRA_STEP_CLOCK=clock, MITHRA_MAX_STEPS=str(steps + 1))
    with open(os.path.join(workdir, "log.txt"), "w") as log:
        subprocess.run([MPIEXEC, "-launcher", "fork", "-n", str(ranks), HARNESS, "job.job"], cwd=workdir, env=env,
                       stdout=log, stderr=subprocess.STDOUT, check=True, timeout=timeout)
    t = [float(line.split()[1]) for line in open(clock)]
    return (t[-1] - t[1]) / (len(t) - 2)   # steps 1..steps (step 0 warms the caches)


def time_oracle(workdir, text, threads, steps):
    from golden_state import load_oracle
    from mithra_amd import abi
    from mithra_amd.job import Job
    import bench
    import ctypes as C
    path = os.path.join(workdir, "job.job")
    with open(path, "w") as f:
        f.write(text)
    cwd = os.getcwd()
    os.chdir(workdir)
    try:
        job = Job("job.job")
        lib = load_oracle()
        lib.oracle_set_threads.restype, lib.oracle_set_threads.argtypes = C.c_int, [C.c_int]
        lib.oracle_set_threads(threads)
        dom = bench.open_domain(job, abi.PRECISION_FP64, lib=lib, prefix="oracle_", rows=steps + 1)
        m = bench.March(job)
        stamps = []
        for _ in range(steps + 1):
            stamps.append(time.perf_counter())
            dom.field_update(m.time, m.seed)
            dom.bunch_begin()
            tb = m.tb
            for _ in range(m.t["n_sub"]):
                dom.bunch_update(m.bp, m.und, tb)
                tb += m.t["bunch_time_step"]
            dom.power_sample(m.n, m.n)
            dom.field_shift()
            dom.current_reset()
            dom.current_update()
            m.advance()
        stamps.append(time.perf_counter())
        dom.close()
        job.close()
    finally:
        os.chdir(cwd)
    return (stamps[-1] - stamps[1]) / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--workloads", default="c1,c2_slice")
    args = ap.parse_args()
    subprocess.run(["make", "-s", "-C", HERE, "ref", "cpu_ref"], check=True)
    rows = []
    for name in args.workloads.split(","):
        w = WORKLOADS[name]
        text = job_text(**w)
        nodes = w["nodes"][0] ** 2 * w["nodes"][1]
        for ranks in (1, 2, 4, 8):
            with tempfile.TemporaryDirectory(prefix="calib_") as work:
                s = time_reference(work, text, ranks, args.steps)
            rows.append((name, nodes, w["n"], f"reference, {ranks} MPI rank{'s' if ranks > 1 else ''}", s))
            print(rows[-1], flush=True)
        for threads in (1, 8):
            with tempfile.TemporaryDirectory(prefix="calib_") as work:
                s = time_oracle(work, text, threads, args.steps)
            rows.append((name, nodes, w["n"], f"CPU oracle, {threads} OpenMP thread{'s' if threads > 1 else ''}", s))
            print(rows[-1], flush=True)
    out = os.path.join(REPO, "profiles", "cpu_baseline_calibration.md")
    with open(out, "w") as f:
        f.write("# CPU baseline calibration: the reference against the CPU oracle\n\n")
        f.write("Command: `python oracle/calibrate_cpu.py --steps %d` in the 8-core build container (the reference is "
                "built from source there; g++ -O3, MPICH 3.3.2).  Seconds per field step, averaged over steps 1..%d "
                "(step 0 left out); cell-updates/s = nodes / (s/step).  Both sides march the same job from the "
                "same state (the reference's own Gaussian-ellipsoid bunch, which the drop-in host reproduces).\n\n"
                % (args.steps, args.steps))
        f.write("| workload | nodes | particles | code | s/step | cell-updates/s | vs reference, 1 rank |\n")
        f.write("|---|---:|---:|---|---:|---:|---:|\n")
        base = {}
        for name, nodes, n, code, s in rows:
            if code.startswith("reference, 1 "):
                base[name] = s
            f.write(f"| {name} | {nodes} | {n} | {code} | {s:.4f} | {nodes / s:.3e} | {base[name] / s:.2f}x |\n")
    print("wrote", out)


if __name__ == "__main__":
    main()
