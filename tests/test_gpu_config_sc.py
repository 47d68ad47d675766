"""The space-charge and seeded path (FdTdSC with a Gaussian-beam TF/SF seed) at a size between the
golden fixtures (12 x 12 x 77) and the C3 workload (128 x 128 x 1024): the GPU's fused step against
the oracle's call-by-call march, from the same drop-in host set-up.

Configuration: 32 x 32 x 128 nodes with the cells of the reference's CPU-runnable case (380 um x
380 um x 1.6 um lab at gamma 196.7, NSFD, Mur order 2), space charge, a Gaussian-beam seed of the
C3 workload's shape (radius a quarter of the transverse mesh, wavelength 20 cells), and a 100 pC
Gaussian bunch of 20000 particles spanning about an eighth of the mesh per axis.  Bar: after 8
steps, A, phi, J, rho and the particles within 1e-12 relative L2 of the oracle, P(z) within 1e-10
of its peak.  The seed's analytic field uses the device's exp/cos/atan, a few ulp from the host's,
which the bar absorbs."""
import math
import os
import tempfile

import numpy as np
import pytest

from golden_state import load_oracle, rel_l2
from mithra_amd.job import Job
from test_gpu_config_small import March, open_domain

pytestmark = pytest.mark.gpu

NODES = (32, 32, 128)
N_PART = 20_000
GAMMA = 196.7
DX, DZ = 380.0, 1.6
STEPS = 8

JOB = """MESH
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
    strength-parameter    = 2.0e-5
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
    undulator-parameter   = 1.0
    period                = 1.5e4
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


def bunch(lx, lz):
    rng = np.random.default_rng(11)
    electrons = 100e-12 / 1.602e-19
    rec = np.empty((N_PART, 7))
    rec[:, 0] = electrons / N_PART
    rec[:, 1] = rng.normal(0.0, lx / 16.0, N_PART)
    rec[:, 2] = rng.normal(0.0, lx / 16.0, N_PART)
    rec[:, 3] = rng.normal(0.0, lz / 16.0, N_PART)
    rec[:, 4] = rng.normal(0.0, 1e-3, N_PART)
    rec[:, 5] = rng.normal(0.0, 1e-3, N_PART)
    rec[:, 6] = math.sqrt(GAMMA ** 2 - 1.0) * (1.0 + 1e-3 * rng.normal(0.0, 1.0, N_PART))
    return rec, electrons


def test_space_charge_seeded_configuration():
    lib = load_oracle()
    lib.oracle_set_threads(min(16, os.cpu_count() or 1))
    lx, lz = (NODES[0] - 1.5) * DX, (NODES[2] - 1.5) * DZ   # (int)(L / d) + 2 nodes, src/solver.cpp:610-612
    with tempfile.TemporaryDirectory() as work:
        rec, electrons = bunch(lx, lz)
        path = os.path.join(work, "sc.job")
        with open(path, "w") as f:
            f.write(JOB.format(lx=lx, lz=lz, dx=DX, dz=DZ, charge=electrons, n=N_PART, gamma=GAMMA, zseed=lz / 4.0,
                               rseed=lx / 4.0, wseed=20.0 * DZ, zpower=lz / 8.0, dir=work))
        job = Job(path, external_bunch=rec)
        g = job.grid()
        assert (g.N0, g.N1, g.N2) == NODES and g.space_charge == 1
        assert abs(job.seed().amplitude) > 1e-50
        gpu = open_domain(job)
        ora = open_domain(job, lib=lib, prefix="oracle_")
        mg, mo = March(job), March(job)
        for _ in range(STEPS):
            gpu.step(mg.time, mg.seed, mg.bp, mg.und, mg.tb, mg.t["bunch_time_step"], mg.t["n_sub"], mg.step, mg.step)
            mg.advance()
            ora.field_update(mo.time, mo.seed)
            ora.bunch_begin()
            tb = mo.tb
            for _ in range(mo.t["n_sub"]):
                ora.bunch_update(mo.bp, mo.und, tb)
                tb += mo.t["bunch_time_step"]
            ora.power_sample(mo.step, mo.step)
            ora.field_shift()
            ora.current_reset()
            ora.current_update()
            mo.advance()
        out = {}
        for lv in (0, 2):  # A^n and phi^n; J and rho
            a_g, f_g = gpu.get_potentials(lv)
            a_o, f_o = ora.get_potentials(lv)
            out["A" if lv == 0 else "J"] = rel_l2(a_g, a_o)
            out["phi" if lv == 0 else "rho"] = rel_l2(f_g, f_o)
        out["particles"] = rel_l2(gpu.get_particles()[:, 1:10], ora.get_particles()[:, 1:10])
        P_g = gpu.power_fetch(0, STEPS, 1, 1)[:, 0, 0]
        P_o = ora.power_fetch(0, STEPS, 1, 1)[:, 0, 0]
        out["P"] = float(np.max(np.abs(P_g - P_o)) / max(np.max(np.abs(P_o)), 1e-300))
        print(f"GPU vs oracle after {STEPS} steps: {out}", flush=True)
        for key in ("A", "phi", "J", "rho", "particles"):
            assert out[key] <= 1e-12, key
        assert out["P"] <= 1e-10
        job.close()
