"""The drop-in host's parser and set-up against the reference's own state before step 0.

Every quantity the reference computes in Solver::initialize (src/solver.cpp:547-595) -- mesh,
boost, update coefficients, bunch generation and boost, seeded initial potentials, sampler set-up
-- must come out bit-identical: the host follows the reference's arithmetic in source order.
"""
import os
import shutil

import numpy as np
import pytest

from golden_state import CASES, GOLDEN, Case
from mithra_amd.job import Job

JOBS = os.path.join(GOLDEN, "jobs")


@pytest.fixture(scope="module", params=CASES)
def setup(request, tmp_path_factory):
    name = request.param
    work = tmp_path_factory.mktemp(name)
    shutil.copy(os.path.join(JOBS, name + ".job"), work)
    if "file-name" in open(os.path.join(JOBS, name + ".job")).read():
        import sys
        sys.path.insert(0, GOLDEN)
        from bunchfile import write_bunch_file
        write_bunch_file(os.path.join(work, "bunch_e.txt"), 1000, seed=5)
    cwd = os.getcwd()
    os.chdir(work)
    try:
        job = Job(name + ".job")
    finally:
        os.chdir(cwd)
    return Case(name), job


def test_mesh_and_coefficients(setup):
    case, job = setup
    g, ref = job.grid(), case.grid
    for f, _ in ref._fields_:
        assert getattr(g, f) == getattr(ref, f), f
    c, rc = job.coeffs(), case.coeffs
    for f, _ in rc._fields_:
        a, b = getattr(c, f), getattr(rc, f)
        assert (list(a) == list(b)) if hasattr(a, "__len__") else a == b, f
    bp = job.bunch_params()
    for f, _ in bp._fields_:
        assert getattr(bp, f) == getattr(case.bunch, f), f


def test_seed_undulators_and_sampler(setup):
    case, job = setup
    s = job.seed()
    for f in ("seed_type", "signal_type", "nR", "amplitude", "t0", "s", "f0", "cep", "gamma", "beta",
              "dt_shift", "c0"):
        assert getattr(s, f) == getattr(case.seed, f), f
    for f in ("position", "direction", "polarization", "radius"):
        assert list(getattr(s, f)) == list(getattr(case.seed, f)), f
    und = job.undulators()
    assert len(und) == len(case.undulators)
    for u, v in zip(und, case.undulators):
        for f, _ in u._fields_:
            assert getattr(u, f) == getattr(v, f), f
    t = job.times()
    assert t["n_sub"] == case.n_sub and t["dt"] == case.dt and t["bunch_time_step"] == case.dtb
    assert t["time"] == case.time0 and t["time_bunch"] == case.time_bunch0
    N, Nl, Nf, pc, z, w = job.power()
    assert (N, Nl, Nf, pc) == (case.power_N, case.power_Nl, case.power_Nf, case.power.pc)
    assert np.array_equal(z, case.power._arrays[0]) and np.array_equal(w, case.power._arrays[1])


def test_particles_and_initial_fields(setup):
    case, job = setup
    assert np.array_equal(job.particles(), case.arr("pre0_particles").reshape(-1, 11))
    a0, f0 = job.potentials(0)
    a1, f1 = job.potentials(1)
    assert np.array_equal(a0, case.arr("pre0_an"))
    assert np.array_equal(a1, case.arr("pre0_anm1"))
    if case.grid.space_charge:
        assert np.array_equal(f0, case.arr("pre0_fn")) and np.array_equal(f1, case.arr("pre0_fnm1"))
