"""Parity at the production size: 256 x 256 x 2048 nodes and 16,777,216 particles, the bench's own
workload and set-up (seeded quiet-start bunch through the drop-in host).  The pipelined GPU step
is marched for 3 steps (first step with the interior update, last with the separate E/B and J
deposit) against the oracle's call-by-call march of the reference's loop body on the host's
cores.  Bar, as for the golden cases after few steps: every potential level (A^n, A^{n-1}, J)
and the particles within 1e-12 relative L2 of the oracle, the particles' emission flags equal,
P(z) within 1e-10 of its peak.  The deposit's summation order is the only arithmetic
difference of the FP64 path."""
import os
import sys
import tempfile

import numpy as np
import pytest

from golden_state import load_oracle, rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

pytestmark = pytest.mark.gpu

STEPS = 3


def test_production_size_pipelined_against_oracle():
    import bench
    from mithra_amd import abi

    lib = load_oracle()
    lib.oracle_set_threads(min(16, os.cpu_count() or 1))
    with tempfile.TemporaryDirectory() as work:
        job = bench.set_up(work)
        g = job.grid()
        assert (g.N0, g.N1, g.N2) == (256, 256, 2048)
        gpu = bench.open_domain(job, abi.PRECISION_FP64, rows=STEPS)
        gpu.set_lazy_fields(1)
        ora = bench.open_domain(job, abi.PRECISION_FP64, lib=lib, prefix="oracle_", rows=STEPS)
        mg, mo = bench.March(job), bench.March(job)
        n_sub, dtb = mg.t["n_sub"], mg.t["bunch_time_step"]
        for s in range(STEPS):
            gpu.step_pipelined(mg.time, mg.seed, mg.bp, mg.und, mg.tb, dtb, n_sub, mg.n, mg.n,
                               more=s < STEPS - 1)
            assert gpu.pipeline_pending() == (s < STEPS - 1)
            mg.advance()
            ora.field_update(mo.time, mo.seed)
            ora.bunch_begin()
            tb = mo.tb
            for _ in range(n_sub):
                ora.bunch_update(mo.bp, mo.und, tb)
                tb += dtb
            ora.power_sample(mo.n, mo.n)
            ora.field_shift()
            ora.current_reset()
            ora.current_update()
            mo.advance()
            print(f"step {s} done", flush=True)
        out = {}
        for lv, name in ((0, "A^n"), (1, "A^n-1"), (2, "J")):
            out[name] = rel_l2(gpu.get_potentials(lv)[0], ora.get_potentials(lv)[0])
        pg, po = gpu.get_particles(), ora.get_particles()
        out["particles"] = rel_l2(pg[:, 1:10], po[:, 1:10])
        flags_equal = bool(np.array_equal(pg[:, 10], po[:, 10]))
        P_g = gpu.power_fetch(0, STEPS, 1, 1)[:, 0, 0]
        P_o = ora.power_fetch(0, STEPS, 1, 1)[:, 0, 0]
        out["P"] = float(np.max(np.abs(P_g - P_o)) / max(np.max(np.abs(P_o)), 1e-300))
        print(f"production size, GPU pipelined vs oracle after {STEPS} steps: {out}, flags equal {flags_equal}",
              flush=True)
        for key in ("A^n", "A^n-1", "J", "particles"):
            assert out[key] <= 1e-12, key
        assert flags_equal
        assert out["P"] <= 1e-10
        gpu.close()
        ora.close()
        job.close()
