"""Diagnostics writers of the drop-in against the reference's own files, and the device gathers
they read.

1. Jobs with every writer the reference drives from its main loop (field sampling at points and
   over a line, field visualization over the whole domain and in xy / yz / xz planes, field
   profiles by rhythm and by time, bunch sampling, bunch visualization, bunch profiles by time
   and rhythm, lab-frame screens; with and without space charge) run through the drop-in for the
   fixture's 30 steps write the same set of files as the reference (tests/golden/diag/*.tar.xz,
   written by tests/golden/make_diag.py), with the same text and numbers to the bar of
   tests/diag_files.py.  The field profiles of the space-charge job print A and phi only, and the
   other job's profile fires on the steps of its whole-domain visualization: the reference prints
   E/B as last lazily evaluated, which for nodes no particle or sampler touched in that step is
   an older step's value (src/fdtd.cpp:1546-1595 reads en_ without evaluating it), while the
   drop-in prints that step's E/B everywhere.
2. mithra_dev_sample_nodes returns exactly the whole-array values at the listed nodes,
   mithra_dev_bunch_sums the charge-weighted sums of the records to rounding, and
   mithra_dev_screen_crossings the crossings a brute-force loop over the records finds."""
import os
import shutil

import numpy as np
import pytest

from diag_files import compare, reference_files
from golden_state import GOLDEN
from mithra_amd import abi
from mithra_amd.job import Job

pytestmark = pytest.mark.gpu

DIAG = os.path.join(GOLDEN, "diag")
JOBS = sorted(f[:-4] for f in os.listdir(DIAG) if f.endswith(".job"))
STEPS = 30


def run_job(name, work):
    shutil.copy(os.path.join(DIAG, name + ".job"), work)
    cwd = os.getcwd()
    os.chdir(work)
    try:
        job = Job(name + ".job")
        job.run(STEPS)
        job.close()
    finally:
        os.chdir(cwd)
    ours = {}
    for root, _, names in os.walk(work):
        for f in names:
            rel = os.path.relpath(os.path.join(root, f), work)
            if rel != name + ".job":
                ours[rel] = open(os.path.join(root, f)).read()
    return ours


@pytest.mark.parametrize("name", JOBS)
def test_diagnostics_files_match_reference(name, tmp_path):
    ref = reference_files(os.path.join(DIAG, name + ".tar.xz"))
    ours = run_job(name, str(tmp_path))
    print(f"{name}: {len(ref)} reference files, {len(ours)} written", flush=True)
    assert sorted(ours) == sorted(ref)
    for rel in sorted(ref):
        ok, msg, exact = compare(ref[rel], ours[rel])
        print(f"  {rel}: {'ok' if ok else 'DIFFERS'}, {exact:.4f} of the numbers print identically", flush=True)
        assert ok, f"{rel}: {msg}"


def open_case(name, tmp_path):
    shutil.copy(os.path.join(DIAG, name + ".job"), tmp_path)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        job = Job(name + ".job")
    finally:
        os.chdir(cwd)
    dom = abi.Domain(job.grid(), job.coeffs())
    job.fill(dom)
    return job, dom


def one_step(job, dom, t):
    dom.field_update(t["time"], job.seed())
    dom.bunch_advance(job.bunch_params(), job.undulators(), t["time_bunch"], t["bunch_time_step"], t["n_sub"])


def test_sample_nodes_and_bunch_sums(tmp_path):
    job, dom = open_case("diag_nsfd_o2_sc", tmp_path)
    t = job.times()
    for s in range(3):
        one_step(job, dom, dict(t, time=t["time"] + s * t["dt"], time_bunch=t["time_bunch"] + s * t["n_sub"] *
                                t["bunch_time_step"]))
        if s < 2:
            dom.field_shift()
            dom.current_reset()
            dom.current_update()
    g = job.grid()
    rng = np.random.default_rng(3)
    nodes = rng.integers(0, g.nodes, 500)
    e, b, a, phi = dom.sample_nodes(nodes)
    E, B = dom.get_fields()
    A, PHI = dom.get_potentials(0)
    assert np.array_equal(e, E.reshape(-1, 3)[nodes]) and np.array_equal(b, B.reshape(-1, 3)[nodes])
    assert np.array_equal(a, A.reshape(-1, 3)[nodes]) and np.array_equal(phi, PHI[nodes])
    with pytest.raises(abi.MithraError):
        dom.sample_nodes(np.array([g.nodes]))
    q = dom.get_particles()
    w, r, gb = q[:, 0], q[:, 1:4], q[:, 7:10]
    want = np.concatenate([[w.sum()], (w[:, None] * r).sum(0), (r * r * w[:, None]).sum(0),
                           (w[:, None] * gb).sum(0), (gb * gb * w[:, None]).sum(0)])
    scale = np.concatenate([[np.abs(w).sum()], (np.abs(w[:, None] * r)).sum(0), (r * r * np.abs(w[:, None])).sum(0),
                            (np.abs(w[:, None] * gb)).sum(0), (gb * gb * np.abs(w[:, None])).sum(0)])
    got = dom.bunch_sums()
    err = np.max(np.abs(got - want) / scale)
    print(f"bunch sums: max difference from numpy {err:.2e} of the sums of magnitudes", flush=True)
    assert err < 1e-13
    assert np.array_equal(dom.bunch_sums(), got)  # the same state gives the same sums
    dom.close()
    job.close()


def test_screen_crossings_brute_force(tmp_path):
    job, dom = open_case("diag_nsfd_o2", tmp_path)
    t = job.times()
    bp = job.bunch_params()
    steps, total = 0, 0
    while steps < 22:
        tt = dict(t, time=t["time"] + steps * t["dt"],
                  time_bunch=t["time_bunch"] + steps * t["n_sub"] * t["bunch_time_step"])
        one_step(job, dom, tt)
        q = dom.get_particles()
        tb = tt["time_bunch"] + t["n_sub"] * t["bunch_time_step"]
        p = abi.ScreenParams(bp.gamma, bp.beta, bp.c0, tb, t["dt"], bp.dt_shift)
        lz = np.array([-30000.0, 0.0, 25000.0])
        scr, idx, rec, count = dom.screen_crossings(lz, p)
        # brute force in the reference's expressions (src/solver.cpp:2220-2250)
        lzm = bp.gamma * (q[:, 6] + bp.beta * bp.c0 * (tb - t["dt"] + bp.dt_shift))
        lzp = bp.gamma * (q[:, 3] + bp.beta * bp.c0 * (tb + bp.dt_shift))
        want = sorted((s, i) for s in range(3) for i in np.nonzero((lzm < lz[s]) & (lzp >= lz[s]))[0])
        assert count == len(want) == len(scr)
        assert sorted(zip(scr.tolist(), idx.tolist())) == want
        for k in range(len(scr)):
            s, i = scr[k], idx[k]
            x = q[i, 4] + (lz[s] - lzm[i]) / (lzp[i] - lzm[i]) * (q[i, 1] - q[i, 4])
            g = q[i, 7:10]
            gz = bp.gamma * (g[2] + bp.beta * np.sqrt(1.0 + (g[0] * g[0] + g[1] * g[1] + g[2] * g[2])))
            assert rec[k, 0] == x and rec[k, 3] == g[0] and rec[k, 5] == gz
        if count:
            # a pass with too little room reports the full count and writes only what fits
            _, _, _, c2 = dom.screen_crossings(lz, p, cap=1)
            assert c2 == count
        total += count
        dom.field_shift()
        dom.current_reset()
        dom.current_update()
        steps += 1
    print(f"screens: {total} crossings over {steps} steps", flush=True)
    assert total > 1000
    dom.close()
    job.close()
