"""FEL parameter scans on one GPU: the seeded instance generator, each instance's set-up through
the drop-in host, and the two ways to march a scan -- instance after instance through the
single-domain path, or all instances together as one batch.

An instance is a boosted-frame slice simulation of a planar undulator FEL, written as an
ordinary job file (so any instance can be run, or checked, on its own):

* beam energy gamma, undulator parameter K, period lambda_u and relative energy spread are
  drawn from the scan ranges; the resonant wavelength lambda_r = lambda_u (1 + K^2/2) / (2 gamma^2)
  sets the lab-frame cell length dz = lambda_r / 16;
* the mesh has n_perp x n_perp x n_z cells, the transverse cell dx = 2 gamma dz (above the
  NSFD solver's dx >= 1.44 gamma dz, so the reference keeps the mesh as written);
* the bunch is Gaussian with sigma_r and sigma_z a fraction of the mesh (one eighth by default),
  loaded by the reference's own quiet start (halton numbers, beamlets lambda_r / 4 apart) with
  shot noise, or with an imposed initial bunching at lambda_r instead, generated on the device,
  n_particles macro-particles of a beam of the given peak current (1 kA by default), normalised
  transverse momentum spread emittance / sigma_r (2e-3 by default);
* the march runs n_z field steps (with dt = dz/c in the frame, the light-crossing time of the
  window) times `passes`, and P(z) and the bunching factor are sampled at SAMPLES evenly spaced steps, the
  power on one plane ahead of the bunch at the resonant frequency.
"""
import dataclasses
import math
import os
import time

import numpy as np

from . import abi
from .job import Job

SAMPLES = 128
C0_UM_PS = 299.792458          # speed of light in um / ps

JOB = """MESH
{{
  length-scale            = MICROMETER
  time-scale              = PICOSECOND
  mesh-lengths            = ( {lx:.10g}, {lx:.10g}, {lz:.10g} )
  mesh-resolution         = ( {dx:.10g}, {dx:.10g}, {dz:.10g} )
  mesh-center             = ( 0.0, 0.0, 0.0 )
  total-time              = 1.0e6
  mesh-truncation-order   = 2
  space-charge            = {sc}
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
    shot-noise            = {shot}
    bunching-factor       = {bf:.10g}
    bunching-factor-phase = {bfp:.10g}
  }}
}}
{field}UNDULATOR
{{
  static-undulator
  {{
    undulator-parameter   = {K:.10g}
    period                = {lu:.10g}
    length                = 100000
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
    base-name             = power-sampling/power-{idx}
    plane-position        = {zp:.10g}
    normalized-frequency  = 1.00
  }}
}}
"""

SEED = """FIELD
{{
  field-initialization
  {{
    type                  = gaussian-beam
    position              = ( 0.0, 0.0, {zseed:.10g} )
    direction             = ( 0.0, 0.0, 1.0 )
    polarization          = ( 0.0, 1.0, 0.0 )
    radius-parallel       = {w0:.10g}
    radius-perpendicular  = {w0:.10g}
    strength-parameter    = {a0:.10g}
    signal-type           = gaussian
    offset                = 0.0
    pulse-length          = 1.0e12
    wavelength            = {lam:.10g}
    CEP                   = 0.0
  }}
}}
"""


@dataclasses.dataclass
class Instance:
    """One scan point (lab frame; lengths in um)."""
    n_perp: int
    n_z: int
    n_particles: int
    gamma: float
    K: float
    lambda_u: float
    spread: float
    emittance: float = 0.0    # normalised emittance (um rad); 0: transverse gamma-beta spread 2e-3
    current: float = 1000.0   # peak current (A)
    sr_frac: float = 0.125    # sigma_r / transverse mesh length
    sz_frac: float = 0.125    # sigma_z / longitudinal mesh length
    bunching: float = 0.0     # imposed initial bunching at lambda_r (amplitude, with `phase`); 0: shot noise
    phase: float = 0.0
    space_charge: bool = False
    seed_power: float = 0.0   # peak power (W) of a Gaussian-beam seed at lambda_r; 0: no seed
    seed_w0: float = 10.0     # the seed's waist radius in units of sigma_r
    passes: int = 1           # field steps in units of n_z (window light-crossing times)
    quiet: bool = False       # without an imposed bunching: no shot noise either (pure quiet start)

    @property
    def lambda_r(self):
        return self.lambda_u * (1.0 + self.K ** 2 / 2.0) / (2.0 * self.gamma ** 2)

    @property
    def dz(self):
        return self.lambda_r / 16.0

    @property
    def dx(self):
        return 2.0 * self.gamma * self.dz

    @property
    def n_steps(self):
        return self.n_z * self.passes

    @property
    def every(self):
        return max(1, self.n_steps // SAMPLES)


def generate(n=64, seed=4, n_perp=(32, 128), n_z=(128, 1024), n_particles=(1e4, 1e6)):
    """The scan: sizes log-uniform, gamma U[100, 2000], K U[0.5, 2], lambda_u U[1, 3] cm,
    energy spread log-uniform in [1e-4, 1e-2]."""
    rng = np.random.default_rng(seed)

    def logu(lo, hi):
        return math.exp(rng.uniform(math.log(lo), math.log(hi)))

    out = []
    for i in range(n):
        out.append(Instance(n_perp=int(round(logu(*n_perp))), n_z=int(round(logu(*n_z))),
                            n_particles=int(round(logu(*n_particles))), gamma=rng.uniform(100.0, 2000.0),
                            K=rng.uniform(0.5, 2.0), lambda_u=rng.uniform(1.0e4, 3.0e4),
                            spread=logu(1e-4, 1e-2)))
    return out


def bunch_sizes(inst):
    """sigma_r, sigma_z (lab frame, um), the transverse momentum spread and the electrons of the
    instance's Gaussian bunch of peak current inst.current."""
    lx, lz = (inst.n_perp - 1) * inst.dx, (inst.n_z - 1) * inst.dz
    sr, sz = lx * inst.sr_frac, lz * inst.sz_frac
    sgb = inst.emittance / sr if inst.emittance > 0.0 else 2e-3
    electrons = inst.current * (math.sqrt(2.0 * math.pi) * sz * 1e-6) / 2.99792458e8 / 1.602e-19
    return sr, sz, sgb, electrons


def seed_strength(power, w0, lam):
    """Strength parameter a0 = e E0 lambda / (2 pi m c^2) of a Gaussian beam of peak power `power`
    (W) and waist w0 (um) at wavelength lam (um): E0 = sqrt(4 Z0 P / (pi w0^2))."""
    z0 = 376.730313668
    e0 = math.sqrt(4.0 * z0 * power / (math.pi * (w0 * 1e-6) ** 2))
    return 1.602e-19 * e0 * lam * 1e-6 / (2.0 * math.pi * 9.109e-31 * 2.99792458e8 ** 2)


def job_text(inst, idx, workdir):
    """The instance as a job file: a Gaussian ellipsoid loaded by the reference's own quiet start
    (halton numbers, four beamlets lambda_r / 4 apart), with shot noise, or, for an instance with
    an imposed bunching, with that bunching factor and phase instead; space charge and a
    Gaussian-beam seed at lambda_r (waist seed_w0 sigma_r, a quarter of the window ahead of the
    mesh centre) when the instance asks for them."""
    # the reference makes (int)(L / d) + 2 nodes of a mesh length L (src/solver.cpp:610-612)
    lx, lz = (inst.n_perp - 1.5) * inst.dx, (inst.n_z - 1.5) * inst.dz
    sr, sz, sgb, electrons = bunch_sizes(inst)
    bunched = inst.bunching > 0.0
    field = ""
    if inst.seed_power > 0.0:
        w0 = inst.seed_w0 * sr
        field = SEED.format(zseed=lz / 4.0, w0=w0, a0=seed_strength(inst.seed_power, w0, inst.lambda_r),
                            lam=inst.lambda_r)
    return JOB.format(sc="true" if inst.space_charge else "false", field=field,
                      lx=lx, lz=lz, dx=inst.dx, dz=inst.dz, charge=electrons, n=inst.n_particles // 4 * 4,
                      gamma=inst.gamma, sr=sr, sz=sz, sgb=sgb, sgbz=inst.spread * math.sqrt(inst.gamma ** 2 - 1.0),
                      tt=4.0 * sr, lt=3.5 * sz, shot="false" if bunched or inst.quiet else "true",
                      bf=inst.bunching if bunched else 0.0, bfp=math.degrees(inst.phase) if bunched else 0.0,
                      K=inst.K, lu=inst.lambda_u, dir=workdir, idx=idx, zp=0.375 * lz)


def set_up(inst, idx, workdir):
    """Writes the instance's job file and runs the drop-in host's set-up on it, the bunch generated
    and boosted on the device."""
    path = os.path.join(workdir, f"scan-{idx}.job")
    with open(path, "w") as f:
        f.write(job_text(inst, idx, workdir))
    return Job(path, device_bunch=True)


def open_domain(job, precision, rows):
    """A device domain holding the job's state before its first step (the particles handed over on
    the device)."""
    g, c = job.grid(), job.coeffs()
    dom = abi.Domain(g, c, precision=precision)
    job.fill(dom)
    N, Nl, Nf, pc, z, w = job.power()
    bp = job.bunch_params()
    dom.power_setup(abi.make_power_params(z, w, Nf, pc, bp.gamma, bp.beta, bp.c0), rows)
    return dom


def bunching_wavenumber(job):
    """Frame wavenumber of the sampled (resonant) frequency: omega / c."""
    N, Nl, Nf, pc, z, w = job.power()
    return float(w[0]) / job.bunch_params().c0


class Prepared:
    """One instance set up on the device: its job, domain and march constants.  The host set-up
    (job) and the device set-up (open) are separate steps, so that many jobs can be set up on
    host threads while the device calls stay on one thread."""

    def __init__(self, inst, idx, workdir, precision, open_now=True):
        self.inst, self.idx, self.precision = inst, idx, precision
        self.job = set_up(inst, idx, workdir)
        self.t = self.job.times()
        self.bp, self.und, self.seed = self.job.bunch_params(), self.job.undulators(), self.job.seed()
        self.k = bunching_wavenumber(self.job)
        self.dom = None
        if open_now:
            self.open()

    def open(self):
        self.dom = open_domain(self.job, self.precision, self.inst.n_steps)

    def profile(self, bins):
        """Initial current profile: the charge per bin of the particles' z over the mesh's z range,
        in the frame, before the first step."""
        q = self.job.particles()
        g = self.job.grid()
        h, _ = np.histogram(q[:, 3], bins=bins, range=(g.zmin, g.zmax), weights=q[:, 0])
        return h / max(h.max(), 1e-300)

    def result(self, bunching):
        """P(z) and b(z) at the SAMPLES sample steps."""
        P = self.dom.power_fetch(0, self.inst.n_steps, 1, 1)[:, 0, 0]
        idx = np.arange(SAMPLES) * self.inst.every
        idx = idx[idx < self.inst.n_steps]
        return P[idx], np.asarray(bunching)[:len(idx)]


def prepare(instances, workdir, precision=abi.PRECISION_FP32_BULK, log=None, workers=8):
    """Sets the instances up: the host set-up of the jobs on `workers` host threads (the host
    library's set-up of one job shares no mutable state with another's, and ctypes releases
    the interpreter lock during its calls), then each domain's creation and upload on this
    thread, in order."""
    from concurrent.futures import ThreadPoolExecutor

    def host(item):
        i, inst = item
        return Prepared(inst, i, workdir, precision, open_now=False)

    if workers <= 1:
        out = [host(item) for item in enumerate(instances)]
    else:
        with ThreadPoolExecutor(max_workers=min(workers, max(1, len(instances)))) as ex:
            out = list(ex.map(host, enumerate(instances)))
    for n, p in enumerate(out, 1):
        p.open()
        if log and n % 8 == 0:
            log(f"set up {n}/{len(instances)} instances")
    return out


def march_single(p):
    """The plain single-instance GPU path: mithra_dev_step per field step, the bunching factor
    sampled after the steps the batch samples."""
    t = dict(p.t)
    time_, tb = t["time"], t["time_bunch"]
    bunch = []
    for s in range(p.inst.n_steps):
        p.dom.step(time_, p.seed, p.bp, p.und, tb, t["bunch_time_step"], t["n_sub"], s, s)
        if s % p.inst.every == 0 and len(bunch) < SAMPLES:
            bunch.append(p.dom.bunching(p.k))
        for _ in range(t["n_sub"]):
            tb += t["bunch_time_step"]
        time_ += t["dt"]
    p.dom.synchronize()
    return p.result(bunch)


def run_single(prepared):
    """All instances one after the other; returns (results, seconds)."""
    t0 = time.perf_counter()
    res = [march_single(p) for p in prepared]
    return res, time.perf_counter() - t0


def make_batch(prepared, variant=None, use_graph=True, autotune=True):
    b = abi.Batch([p.dom for p in prepared])
    for i, p in enumerate(prepared):
        b.instance(i, p.t["time"], p.t["time_bunch"], p.t["bunch_time_step"], p.t["n_sub"], p.inst.n_steps, p.bp,
                   p.und, p.k, p.inst.every, SAMPLES)
        b.seed(i, p.seed)
    tuned = None
    if autotune:
        tuned = b.autotune()
    if variant is not None or not use_graph:
        b.set_tiles(variant if variant is not None else (tuned[0] if tuned else 8), use_graph)
    return b, tuned


def run_batch(prepared, variant=None, use_graph=True, autotune=True):
    """All instances as one batch; returns (results, seconds of the march, autotune result)."""
    b, tuned = make_batch(prepared, variant, use_graph, autotune)
    t0 = time.perf_counter()
    b.run(max(p.inst.n_steps for p in prepared))
    seconds = time.perf_counter() - t0
    res = [p.result(b.bunching(i)) for i, p in enumerate(prepared)]
    b.close()
    return res, seconds, tuned


def main():
    """`python -m mithra_amd.scan [--instances N] [--mode batch|single] [--seed S] [--eager]`: sets up
    the seeded scan through the drop-in host and marches it, printing one JSON line of timings.
    --eager launches the batch's kernels one by one instead of replaying the captured HIP graph
    (same kernels, same results; for kernel traces)."""
    import argparse
    import json
    import sys
    import tempfile
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--mode", choices=("batch", "single"), default="batch")
    ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp32")
    ap.add_argument("--eager", action="store_true", help="no captured graph (batch mode)")
    args = ap.parse_args()
    precision = abi.PRECISION_FP64 if args.precision == "fp64" else abi.PRECISION_FP32_BULK
    insts = generate(args.instances, seed=args.seed)
    with tempfile.TemporaryDirectory(prefix="mithra_scan_") as work:
        t0 = time.perf_counter()
        prep = prepare(insts, work, precision, log=lambda s: print(f"[scan] {s}", file=sys.stderr, flush=True))
        t_setup = time.perf_counter() - t0
        if args.mode == "batch":
            _, t_march, tuned = run_batch(prep, use_graph=not args.eager)
        else:
            _, t_march = run_single(prep)
            tuned = None
        for p in prep:
            p.dom.close()
            p.job.close()
    print(json.dumps({"instances": args.instances, "seed": args.seed, "mode": args.mode, "graph": not args.eager,
                      "precision": args.precision, "setup_s": t_setup, "march_s": t_march,
                      "points_per_hour": 3600.0 * args.instances / (t_setup + t_march),
                      "tile_variant": tuned[0] if tuned else None}), flush=True)


if __name__ == "__main__":
    main()
