"""Benchmark: MITHRA FDTD/PIC time march, production single instance, FP64 drop-in, one MI355X.

Workload (production single-instance configuration): 256 x 256 x 2048 nodes in the boosted frame,
16,777,216 macro-particles of a 1 GeV beam (gamma 1957.9, 0.01 % energy spread, sigma_r 50 um)
in a planar undulator (lambda_u 1.8 cm, K 1.4, 300 periods), NSFD solver with second-order Mur
boundaries, no space charge.  As in the reference's own LCLS-scale example
(prj/FEL-LCLS/job-files/FEL-LCLS.job) the mesh resolves the radiation (dz = 0.2 nm lab,
lambda_r / 23) and holds a slice of the flat-top bunch: 0.28 um of it, loaded by a seeded
quiet start (four beamlets lambda_r / 4 apart, shot-noise modulation per lambda_r bucket),
3 kA.  Particles are synthetic; the set-up runs through the drop-in host (job-file parser,
Lorentz boost, coefficients) exactly as `mithra job.job` does, with the bunch generated and boosted
on the device.

One step = one iteration of Solver::solve()'s main loop on the GPU: field update (interior,
Mur, E/B), bunch push, P(z) sample, shift, current reset, deposition.  By default the steps run
pipelined (mithra_dev_step_pipelined): E/B of a step is evaluated with the next interior update's
J-free part in one kernel, and the deposit adds a4 J to it; the timed steps form one complete run
(its first step makes the interior update, its last the separate E/B and J deposit), with the
same FP64 results as the plain step.  --plain times mithra_dev_step instead.  By default the timed
run is the production configuration's full 2000 steps: the push gathers grid fields only for
particles whose emission flag is set (once their lab-frame z passes -dist, src/solver.cpp:1459),
a fraction that grows from 0 to 1 by step ~1500, so a short run would time a cheaper push.

Prints ONE JSON line.  value = cell-updates/s (nodes x steps / s); particle-pushes/s and the
roofline of the dominant kernel (the fused stencil + E/B kernel; the interior stencil with
--plain) are reported beside it, with the CPU
oracle timed on a bounded sample of the same workload on this host.  scan_points_per_hour is
the batched parameter scan's rate (64 seeded instances of uneven size, each a job taken to its
P(z) and bunching at 128 positions, host set-up included), measured in the same run beside the
same scan through the single-instance path.  surrogate_points_per_hour is the surrogate with
FDTD/PIC fallback on the space-charge, seeded generator's held-out points, counting only points
whose returned P(z) is within 2 % of their FP64 solve (a reduced workflow: 1024 training points,
600 epochs, 128 held-out points; the data generation and training are reported, not counted).
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mithra_amd import abi  # noqa: E402
from mithra_amd.job import Job  # noqa: E402

METRIC = "cell-updates/s + particle-pushes/s; FEL scan points/hour at ≤2% P(z) error"
HBM_PEAK_GBS = 8000.0
N_PARTICLES = 16_777_216
GAMMA = 1957.9
SLICE_UM = 0.28          # lab-frame length of the simulated slice of the flat-top
CURRENT_A = 3000.0

JOB = """MESH
{{
  length-scale            = MICROMETER
  time-scale              = PICOSECOND
  mesh-lengths            = ( 305.4, 305.4, 0.4093 )
  mesh-resolution         = ( 1.2, 1.2, 2.0e-4 )
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
    type                  = other
    charge                = {charge}
    number-of-particles   = {n}
    gamma                 = {gamma}
    direction             = ( 0.0, 0.0, 1.0 )
  }}
}}
UNDULATOR
{{
  static-undulator
  {{
    undulator-parameter   = 1.4
    period                = 1.8e4
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
    directory             = {dir}/
    base-name             = power-sampling/power
    plane-position        = 0.1977
    normalized-frequency  = 1.00
  }}
}}
"""


# Space-charge and seeded workload (C3): 128 x 128 x 1024 nodes, 4,194,304 particles of a 30 MeV
# beam (gamma 59.7) with a 500 A peak current, Gaussian sigma_r 200 um, sigma_z 100 um, 0.2 % energy
# spread, 1 mm mrad normalised emittance, in a planar undulator (lambda_u 2 cm, K 0.8), NSFD with
# second-order Mur boundaries, space charge, and a 1 MW Gaussian-beam seed of 2 mm waist at the
# resonant wavelength injected by TF/SF.  The mesh resolves the radiation (dz = lambda_r / 16 lab)
# over 1022.5 cells, so it holds the middle 237 um of the bunch: the particles are loaded within
# +-0.35 of the window (+-83 um) with the charge the 500 A Gaussian carries there.
C3_NODES = (128, 128, 1024)
C3_PARTICLES = 4_194_304
C3_GAMMA = 59.7
C3_K, C3_LU = 0.8, 2.0e4
C3_LAMBDA_R = C3_LU / (2.0 * C3_GAMMA ** 2) * (1.0 + C3_K ** 2 / 2.0)    # 3.70 um
C3_DZ = C3_LAMBDA_R / 16.0
C3_DX = 20.0             # above NSFD's 1.02 sqrt(2) gamma_frame dz = 17.4 um: the mesh stays as written
C3_SIGMA_R, C3_SIGMA_Z = 200.0, 100.0
C3_PEAK_A = 500.0
C3_SEED_W0 = 2000.0
C3_SEED_A0 = 1.263e-5    # e E0 lambda_r / (2 pi m c^2), E0 = sqrt(4 Z0 P / (pi w0^2)) for P = 1 MW
C3_JOB = """MESH
{{
  length-scale            = MICROMETER
  time-scale              = PICOSECOND
  mesh-lengths            = ( {lx:.10g}, {lx:.10g}, {lz:.10g} )
  mesh-resolution         = ( {dx:.10g}, {dx:.10g}, {dz:.10g} )
  mesh-center             = ( 0.0, 0.0, 0.0 )
  total-time              = 1.0e6
  mesh-truncation-order   = 2
  space-charge            = true
  solver                  = NSFD
}}
BUNCH
{{
  bunch-initialization
  {{
    type                  = other
    charge                = {charge:.10g}
    number-of-particles   = {n}
    gamma                 = {gamma:.10g}
    direction             = ( 0.0, 0.0, 1.0 )
  }}
}}
FIELD
{{
  field-initialization
  {{
    type                  = gaussian-beam
    position              = ( 0.0, 0.0, {zseed:.10g} )
    direction             = ( 0.0, 0.0, 1.0 )
    polarization          = ( 0.0, 1.0, 0.0 )
    radius-parallel       = {rseed:.10g}
    radius-perpendicular  = {rseed:.10g}
    strength-parameter    = {a0:.10g}
    signal-type           = gaussian
    offset                = 0.0
    pulse-length          = 1.0e12
    wavelength            = {wseed:.10g}
    CEP                   = 0.0
  }}
}}
UNDULATOR
{{
  static-undulator
  {{
    undulator-parameter   = {K:.10g}
    period                = {lu:.10g}
    length                = 100
    polarization-angle    = 0.0
  }}
}}
FEL-OUTPUT
{{
  radiation-power
  {{
    sample                = true
    type                  = at-point
    directory             = {dir}/
    base-name             = power-sampling/power
    plane-position        = {zpower:.10g}
    normalized-frequency  = 1.00
  }}
}}
"""


def c3_bunch(n, seed=3):
    """Lab-frame records q x y z gbx gby gbz (um) of the C3 bunch: the part of the 500 A Gaussian
    inside +-0.35 of the mesh's z window, with its charge."""
    import math
    rng = np.random.default_rng(seed)
    lz = (C3_NODES[2] - 1.5) * C3_DZ
    half = 0.35 * lz
    z = rng.normal(0.0, C3_SIGMA_Z, 2 * n)
    z = z[np.abs(z) <= half][:n]
    while z.size < n:  # the truncated Gaussian keeps about half of the draws
        extra = rng.normal(0.0, C3_SIGMA_Z, 2 * n)
        z = np.concatenate([z, extra[np.abs(extra) <= half]])[:n]
    inside = math.erf(half / (math.sqrt(2.0) * C3_SIGMA_Z))
    coulomb = C3_PEAK_A / 2.99792458e8 * math.sqrt(2.0 * math.pi) * C3_SIGMA_Z * 1e-6 * inside
    electrons = coulomb / 1.602e-19
    sgb = 1.0e-6 / (C3_SIGMA_R * 1e-6)                     # emittance / sigma_r
    rec = np.empty((n, 7))
    rec[:, 0] = electrons / n
    rec[:, 1] = rng.normal(0.0, C3_SIGMA_R, n)
    rec[:, 2] = rng.normal(0.0, C3_SIGMA_R, n)
    rec[:, 3] = z
    rec[:, 4] = rng.normal(0.0, sgb, n)
    rec[:, 5] = rng.normal(0.0, sgb, n)
    rec[:, 6] = np.sqrt(C3_GAMMA ** 2 - 1.0) * (1.0 + 2e-3 * rng.normal(0.0, 1.0, n))
    return rec, electrons


def c3_job_text(workdir, electrons, n):
    lx, lz = (C3_NODES[0] - 1.5) * C3_DX, (C3_NODES[2] - 1.5) * C3_DZ   # (int)(L / d) + 2 nodes
    return C3_JOB.format(lx=lx, lz=lz, dx=C3_DX, dz=C3_DZ, charge=electrons, n=n, gamma=C3_GAMMA,
                         zseed=lz / 4.0, rseed=C3_SEED_W0, a0=C3_SEED_A0, wseed=C3_LAMBDA_R, K=C3_K, lu=C3_LU,
                         zpower=lz / 8.0, dir=workdir)


def set_up_c3(workdir, n=C3_PARTICLES):
    log("generating the C3 bunch (space charge, Gaussian seed) and running the host set-up")
    bunch, electrons = c3_bunch(n)
    path = os.path.join(workdir, "c3.job")
    with open(path, "w") as f:
        f.write(c3_job_text(workdir, electrons, n))
    job = Job(path, external_bunch=bunch)
    del bunch
    return job


def production_bunch(seed=2):
    """The flat-top slice of the production workload, generated on the device (lab frame, lengths in
    um): 16.8M records of a 3 kA beam over SLICE_UM, four beamlets lambda_r / 4 apart per base
    particle, one shot-noise modulation per lambda_r bucket, sigma_r 50 um, gb_perp spread 0.02,
    0.01 % energy spread, centred where the frame shift puts it mid-mesh."""
    k = 1.4
    lam = 1.8e4 / (2.0 * GAMMA ** 2) * (1.0 + k * k / 2.0)     # resonant wavelength, um
    electrons = CURRENT_A * SLICE_UM * 1e-6 / 2.99792458e8 / 1.602e-19
    f = abi.Flattop()
    f.n, f.charge, f.gamma, f.lambda_, f.length = N_PARTICLES, electrons, GAMMA, lam, SLICE_UM
    f.z_center = 0.33 * SLICE_UM / 2.0        # bunch speed in the frame before the undulator x half slice
    f.sigma_r, f.sigma_gb_perp, f.energy_spread = 50.0, 0.02, 1.0e-4
    f.electrons_per_bucket = electrons * lam / SLICE_UM
    bunch = abi.DeviceBunch()
    bunch.flattop(f, seed)
    return bunch, electrons


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def set_up(workdir):
    log("generating the quiet-start bunch on the device and running the host set-up")
    bunch, electrons = production_bunch()
    path = os.path.join(workdir, "production.job")
    with open(path, "w") as f:
        f.write(JOB.format(charge=electrons, n=N_PARTICLES, gamma=GAMMA, dir=workdir))
    return Job(path, external_bunch=bunch)   # the job keeps the bunch, boosted in place


def open_domain(job, precision, lib=None, prefix="mithra_dev_", rows=64):
    """A domain holding the job's pre-step state: the GPU's through the host (potentials only when a
    seed set them, particles device to device when generated there), the oracle's from host copies."""
    g, c = job.grid(), job.coeffs()
    if lib is None:
        dom = abi.Domain(g, c, precision=precision)
        job.fill(dom)
    else:
        dom = abi.Domain(g, c, lib=lib, prefix=prefix)
        an, fn = job.potentials(0)
        am, fm = job.potentials(1)
        dom.set_potentials(0, an, fn)
        dom.set_potentials(1, am, fm)
        dom.set_particles(job.particles())
    N, Nl, Nf, pc, z, w = job.power()
    t = job.times()
    bp = job.bunch_params()
    dom.power_setup(abi.make_power_params(z, w, Nf, pc, bp.gamma, bp.beta, bp.c0), rows)
    return dom


class March:
    """Host time bookkeeping of Solver::solve() (time_ and timeBunch_ accumulate)."""

    def __init__(self, job):
        self.t = job.times()
        self.seed, self.bp, self.und = job.seed(), job.bunch_params(), job.undulators()
        self.time, self.tb, self.n = self.t["time"], self.t["time_bunch"], 0

    def advance(self):
        for _ in range(self.t["n_sub"]):
            self.tb += self.t["bunch_time_step"]
        self.time += self.t["dt"]
        self.n += 1


def run_gpu(job, steps, warmup, precision, reorder=16, pipelined=True, deterministic=False):
    """Warm-up, then `steps` timed steps, then a few steps with per-phase HIP events.  Pipelined:
    each group of steps is one complete pipelined run (mithra_dev_step_pipelined, more = 1 on all
    but its last step), so the timed region starts and ends in the reference's between-step state
    and includes the run's first interior update and last separate E/B pass."""
    dom = open_domain(job, precision)
    dom.set_lazy_fields(1)   # E/B where the bunch and the sampler read them (the reference's rule)
    dom.set_reorder_interval(reorder)
    if deterministic:
        dom.set_deterministic(True)
    (ci, ce), tune_ms = dom.autotune(reps=2)
    log("autotune (ms per launch, tile shapes rows x planes): " + "; ".join(
        f"{abi.tile_shape(v)[0]}x{abi.tile_shape(v)[1]} interior {tune_ms[0, v]:.3f} eval {tune_ms[1, v]:.3f}"
        for v in range(abi.TILE_VARIANTS)))
    tiles = {"interior": list(abi.tile_shape(ci)), "eval": list(abi.tile_shape(ce))}
    if pipelined:
        cf, fused_ms = dom.autotune_fused(reps=2)
        log("autotune of the fused stencil + E/B kernel (ms per launch): " + "; ".join(
            f"{abi.fused_tile_shape(v)[0]}x{abi.fused_tile_shape(v)[1]} {fused_ms[v]:.3f}"
            for v in range(abi.FUSED_TILE_VARIANTS)))
        tiles["fused"] = list(abi.fused_tile_shape(cf))
    m = March(job)
    n_sub, dtb = m.t["n_sub"], m.t["bunch_time_step"]
    rows = 64

    def step(more, timed=False):
        r = None
        if pipelined:
            r = dom.step_pipelined(m.time, m.seed, m.bp, m.und, m.tb, dtb, n_sub, m.n, m.n % rows,
                                   more=more, timed=timed)
        elif timed:
            r = dom.step_timed(m.time, m.seed, m.bp, m.und, m.tb, dtb, n_sub, m.n, m.n % rows)
        else:
            dom.step(m.time, m.seed, m.bp, m.und, m.tb, dtb, n_sub, m.n, m.n % rows)
        m.advance()
        return r

    for i in range(warmup):
        step(i < warmup - 1)
    dom.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        step(i < steps - 1)
    dom.synchronize()
    elapsed = time.perf_counter() - t0
    # per-phase device times, HIP events on the domain's stream, over a further run of real
    # steps; pipelined, its first and last steps (interior update, separate E/B) are left out
    n_timed = max(4, min(steps, 10))
    rs = np.array([step(i < n_timed - 1, timed=True) for i in range(n_timed)])
    phases = (rs[1:-1] if pipelined else rs).mean(axis=0)
    counters = dom.bunch_counters()
    P = dom.power_fetch(0, min(rows, warmup + steps), 1, 1)[:, 0, 0]
    emitting = float(np.mean(dom.get_particles()[:, 10]))
    dom.close()
    return elapsed, phases, counters, P, tiles, emitting


def run_surrogate(n_train, work):
    """The surrogate workflow at a reduced size (see mithra_amd/surrogate.py:run_workflow): its
    points/hour within 2 % of FP64 beside the single-instance and batched solvers'."""
    from mithra_amd import surrogate
    n_hold = max(32, n_train // 8)
    log(f"surrogate: {n_train} training points, {n_hold} held out, 600 epochs")
    out = surrogate.run_workflow(work, n_train=n_train, n_hold=n_hold, epochs=600, n_single=8, n_outside=0)
    r = out["points_per_hour"]
    log(f"surrogate: {r['surrogate_with_fallback_within_2pct']:.3g} points/h within 2 %, "
        f"{r['gain_vs_single']:.0f}x the single-instance path, P(z) error median {out['surrogate_p_err']['median']:.4f}")
    return {"n_train": n_train, "n_hold": n_hold, "epochs": 600, "train_s": out["train_s"],
            "data_solve_s": out["solve_s"]["train_setup"] + out["solve_s"]["train_march"],
            "p_err_median": out["surrogate_p_err"]["median"], "p_err_within_2pct": out["surrogate_p_err"]["within_2pct"],
            "points_per_hour": r}


def run_scan(n, work):
    """Batched parameter scan (seed 4: meshes 32-128 transverse x 128-1024 longitudinal cells,
    1e4-1e6 macro-particles, FP32-bulk fields, FP64 particles): the scan once as one batch and
    once instance by instance through the single-domain path.  A scan point is a job taken to
    its P(z) and b(z) at 128 positions; the times include each instance's host set-up."""
    from mithra_amd import scan
    insts = scan.generate(n, seed=4)
    log(f"scan: {n} instances, {sum(i.n_perp ** 2 * i.n_z for i in insts):.3g} nodes, "
        f"{sum(i.n_particles for i in insts):.3g} particles")
    t0 = time.perf_counter()
    prep = scan.prepare(insts, work, abi.PRECISION_FP32_BULK)
    t_setup_b = time.perf_counter() - t0
    _, t_batch, tuned = scan.run_batch(prep)
    for p in prep:
        p.dom.close()
        p.job.close()
    log(f"scan batched: set-up {t_setup_b:.1f} s, march {t_batch:.2f} s (tile shape {tuned[0]})")
    t0 = time.perf_counter()
    prep = scan.prepare(insts, work, abi.PRECISION_FP32_BULK)
    t_setup_s = time.perf_counter() - t0
    _, t_single = scan.run_single(prep)
    for p in prep:
        p.dom.close()
        p.job.close()
    log(f"scan single-instance: set-up {t_setup_s:.1f} s, march {t_single:.2f} s")
    return {"instances": n, "precision": "fp32-bulk fields, fp64 particles",
            "batched_setup_s": t_setup_b, "batched_march_s": t_batch, "single_setup_s": t_setup_s,
            "single_march_s": t_single, "tile_shape": list(abi.tile_shape(tuned[0])),
            "points_per_hour_batched": 3600.0 * n / (t_setup_b + t_batch),
            "points_per_hour_single": 3600.0 * n / (t_setup_s + t_single),
            "points_per_hour_batched_march_only": 3600.0 * n / t_batch}


def run_cpu_baseline(job, steps, threads):
    """The CPU oracle (C++ -O3, OpenMP over nodes/particles, sequential deposition) on a bounded
    sample of the same workload: `steps` full steps of the same mesh and bunch."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from golden_state import load_oracle

    lib = load_oracle()
    import ctypes as C
    lib.oracle_set_threads.restype, lib.oracle_set_threads.argtypes = C.c_int, [C.c_int]
    used = lib.oracle_set_threads(threads)
    dom = open_domain(job, abi.PRECISION_FP64, lib=lib, prefix="oracle_", rows=max(steps, 1))
    m = March(job)
    t0 = time.perf_counter()
    for _ in range(steps):
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
    elapsed = time.perf_counter() - t0
    dom.close()
    return elapsed, used


def run_reference_baseline(workload, ranks, steps, work):
    """Seconds per step of the reference itself: oracle/_ref/ref_harness, the MITHRA reference
    built from its sources with a step clock around its own loop, on `ranks` MPI ranks (its z-slab
    decomposition, src/solver.cpp:619-641) over a job of oracle/calibrate_cpu.py's workloads
    (its own Gaussian-ellipsoid bunch generator).  None when the build or MPICH is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import calibrate_cpu
    if not (os.path.exists(calibrate_cpu.HARNESS) and os.path.exists(calibrate_cpu.MPIEXEC)):
        return None
    w = calibrate_cpu.WORKLOADS[workload]
    d = tempfile.mkdtemp(prefix="ref_", dir=work)
    t0 = time.perf_counter()
    s = calibrate_cpu.time_reference(d, calibrate_cpu.job_text(**w), ranks, steps, timeout=600)
    log(f"reference {workload} on {ranks} rank(s): {s:.3f} s/step ({time.perf_counter() - t0:.1f} s with set-up)")
    return s, w["nodes"][0] ** 2 * w["nodes"][1], w["n"]


def cpu_baseline_leg(job, args, nodes, n_part, n_sub, work):
    """The CPU baseline beside the GPU figure, on this host's cores (rank 0, N = 1): the reference
    itself on the production mesh over args.cpu_threads MPI ranks, its one-core rate on an eighth
    of that mesh, and the CPU oracle (this build's C++ restatement) on the same threads."""
    g = job.grid()
    out = None
    if args.workload == "production":
        log(f"CPU baseline: the reference on {args.cpu_threads} MPI ranks, {args.cpu_steps} steps")
        ref = run_reference_baseline("c2", args.cpu_threads, args.cpu_steps, work)
        if ref:
            s, rn, rp = ref
            out = {"value": rn / s, "unit": "cell-updates/s", "cores": args.cpu_threads, "ranks": args.cpu_threads,
                   "kind": "reference",
                   "sample": f"the MITHRA reference built from source (g++ -O3, MPICH), {args.cpu_steps} steps after "
                             f"one untimed, {args.cpu_threads} MPI ranks (z slabs) on the production mesh "
                             f"{g.N0}x{g.N1}x{g.N2} with a {rp}-particle bunch from its own Gaussian-ellipsoid "
                             f"generator (in place of the flat-top slice), {s:.3f} s/step",
                   "particle_pushes_per_s": rp * n_sub / s}
            one = run_reference_baseline("c2_slice", 1, 2, work)
            if one:
                s1, n1, p1 = one
                out["one_core"] = {"value": n1 / s1, "unit": "cell-updates/s", "cores": 1, "ranks": 1,
                                   "sample": f"the reference on one core, 2 steps of the production cells over "
                                             f"256 of the 2048 planes ({n1} nodes, {p1} particles), "
                                             f"{s1:.3f} s/step"}
    log(f"CPU oracle sample: {args.cpu_steps} steps on {args.cpu_threads} threads")
    cpu_t, used = run_cpu_baseline(job, args.cpu_steps, args.cpu_threads)
    port = {"value": nodes * args.cpu_steps / cpu_t, "unit": "cell-updates/s", "cores": used, "kind": "port",
            "sample": f"{args.cpu_steps} full steps of the same {g.N0}x{g.N1}x{g.N2} mesh and {n_part}-particle "
                      f"bunch with the CPU oracle (C++ -O3, OpenMP over nodes and particles, sequential "
                      f"deposition), {cpu_t:.1f} s",
            "particle_pushes_per_s": n_part * n_sub * args.cpu_steps / cpu_t}
    if out is None:
        return port
    out["port"] = port
    return out


def traffic_from_profiles(kernel, precision):
    """HBM bytes per launch of `kernel` (interior, stencil_eb) from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE and WRITE_SIZE, calibrated on the stream probe),
    when present."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_production_{precision}_traffic.md")
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                if line.startswith("bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000,
                    help="timed steps (default: the production run's full 2000 steps, about 15 s)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp64")
    ap.add_argument("--cpu-steps", type=int, default=3, help="steps of the CPU-oracle sample (0 skips it)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--reorder", type=int, default=16,
                    help="deposits between reorders of the particle arrays into cell-row order (0: never)")
    ap.add_argument("--workload", choices=("production", "c3"), default="production",
                    help="production (256x256x2048, no space charge) or c3 (128x128x1024, space charge, "
                         "Gaussian seed)")
    ap.add_argument("--plain", action="store_true",
                    help="time the plain step (mithra_dev_step) instead of the pipelined one")
    ap.add_argument("--surrogate", type=int, default=1024,
                    help="training points of the surrogate workflow measured for its points/hour (0 skips it)")
    ap.add_argument("--deterministic", action="store_true",
                    help="fixed-order current deposition (bit-for-bit repeatable runs; plain steps)")
    ap.add_argument("--scan", type=int, default=64,
                    help="instances of the batched parameter scan measured for scan points/hour (0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the z-slab multi-GPU "
              "decomposition is a design, not part of this build)", file=sys.stderr)
        sys.exit(2)
    precision = abi.PRECISION_FP64 if args.precision == "fp64" else abi.PRECISION_FP32_BULK
    with tempfile.TemporaryDirectory(prefix="mithra_bench_") as work:
        t_setup = time.perf_counter()
        job = set_up(work) if args.workload == "production" else set_up_c3(work)
        t_setup = time.perf_counter() - t_setup
        g = job.grid()
        nodes = g.N0 * g.N1 * g.N2
        interior = (g.N0 - 2) * (g.N1 - 2) * (g.N2 - 2)
        n_part = job.particles().shape[0]
        n_sub = job.times()["n_sub"]
        log(f"set-up done in {t_setup:.1f} s; timing {args.steps} GPU steps after {args.warmup} warm-up")
        # the library pipelines FP64 runs (with or without space charge); FP32-bulk takes the plain step
        pipelined = not args.plain and not args.deterministic and precision == abi.PRECISION_FP64
        if args.workload == "production" and args.warmup + args.steps < 1500:
            log(f"note: steps {args.warmup}..{args.warmup + args.steps - 1} end before most particles emit "
                "(about step 1500); the push then skips its grid gather and the step is cheaper than the "
                "run's average")
        elapsed, phases, counters, P, tiles, emitting = run_gpu(job, args.steps, args.warmup, precision,
                                                                args.reorder, pipelined, args.deterministic)
        log(f"GPU: {1e3 * elapsed / args.steps:.3f} ms/step; phases {np.round(phases, 3).tolist()}")
        ms_step = 1e3 * elapsed / args.steps
        cell_rate = nodes * args.steps / elapsed
        push_rate = n_part * n_sub * args.steps / elapsed
        # the dominant field kernel and its algorithmic bytes per interior node
        real = 8 if precision == abi.PRECISION_FP64 else 4
        ncomp = 4 if args.workload == "c3" else 3
        if pipelined:   # reads A^{n+1}, A^n (and phi's two levels); writes X (and phi's) and E/B (6 floats)
            kernel, phase, bytes_per_node = "stencil_eb", "eval_eb", 3 * ncomp * real + 24
            kernel_name = ("k_stencil_eb (E/B fused with the next NSFD update's stencil"
                           + (", A and phi)" if ncomp == 4 else ")"))
        else:           # reads A^n, A^{n-1}, J (and phi's three levels with space charge); writes A^{n+1}
            kernel, phase, bytes_per_node = "interior", "interior", 4 * ncomp * real
            kernel_name = "k_interior (NSFD stencil" + (", A and phi)" if ncomp == 4 else ")")
        t_k = phases[abi.PHASES.index(phase)] * 1e-3
        achieved = bytes_per_node * interior / t_k / 1e9
        scan_out = run_scan(args.scan, work) if args.scan > 0 and args.workload == "production" else None
        sur_out = (run_surrogate(args.surrogate, work) if args.surrogate > 0 and args.workload == "production"
                   else None)
        copy_gbs = abi.copy_bandwidth()
        stream_gbs = abi.stream_bandwidth()
        cpu = None
        if args.cpu_steps > 0:
            cpu = cpu_baseline_leg(job, args, nodes, n_part, n_sub, work)
    out = {
        "metric": METRIC,
        "value": cell_rate,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == abi.PRECISION_FP64 else "f32 fields, f64 particles",
        "data": "synthetic: seeded quiet-start bunch generated on the device, job set-up through the drop-in host",
        "config": {"workload": "production single instance (256x256x2048 nodes, 16.8M macro-particles, "
                               "1 GeV, K 1.4, lambda_u 1.8 cm, NSFD, Mur order 2, no space charge)"
                               if args.workload == "production" else
                               "C3: 128x128x1024 nodes, 4.2M macro-particles, 30 MeV, 500 A peak, K 0.8, lambda_u "
                               "2 cm, NSFD, Mur order 2, space charge, 1 MW Gaussian-beam seed at lambda_r (TF/SF)",
                   "nodes": nodes, "particles": n_part, "sub_steps": n_sub,
                   "precision": args.precision, "parallelism": "single GPU, whole domain resident",
                   "reorder_interval": args.reorder,
                   "deposit": "fixed order" if args.deterministic else "row tiles, LDS adds",
                   "timed_steps": [args.warmup, args.warmup + args.steps - 1],
                   "emitting_fraction_after": emitting,
                   "step": "pipelined (E/B fused with the next update's stencil)" if pipelined else "plain"},
        "particle_pushes_per_s": push_rate,
        "scan_points_per_hour": scan_out["points_per_hour_batched"] if scan_out else None,
        "scan": scan_out,
        "surrogate_points_per_hour": (sur_out["points_per_hour"]["surrogate_with_fallback_within_2pct"]
                                      if sur_out else None),
        "surrogate": sur_out,
        "phase_ms": {k: round(float(v), 4) for k, v in zip(abi.PHASES, phases)},
        "roofline": {"bound": "hbm", "kernel": kernel_name, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "frac_of_measured_stream": achieved / stream_gbs, "measured_stream_3r1w_GBps": stream_gbs,
                     "measured_copy_GBps": copy_gbs,
                     "traffic": traffic_from_profiles(kernel, args.precision) if args.workload == "production" else None,
                     "algorithmic_bytes_per_launch": bytes_per_node * interior},
        "cpu_baseline": cpu,
        "setup_s": t_setup,
        "tiles_rows_x_planes": tiles,
        "bunch_counters": {"transverse_escapes": counters[0], "z_crossings": counters[1]},
        "power_last": float(P[-1]) if len(P) else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
