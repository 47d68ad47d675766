"""The drop-in end to end on the GPU: job file -> host set-up -> device time march -> the
reference's power-sampling file, compared with the file the reference itself wrote."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from golden_state import GPU_CASES, GOLDEN, Case
from mithra_amd import abi
from mithra_amd.job import Job

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _stage(name, tmp_path):
    shutil.copy(os.path.join(GOLDEN, "jobs", name + ".job"), tmp_path)
    if "file-name" in open(os.path.join(GOLDEN, "jobs", name + ".job")).read():
        sys.path.insert(0, GOLDEN)
        from bunchfile import write_bunch_file
        write_bunch_file(os.path.join(tmp_path, "bunch_e.txt"), 1000, seed=5)


def _compare(case, path, rows):
    ours = np.loadtxt(path)[:rows]
    ref = case.power_file()[:rows]
    assert ours.shape == ref.shape
    assert np.allclose(ours[:, 0], ref[:, 0], rtol=1e-13, atol=0)
    scale = np.max(np.abs(ref[:, 1]))
    assert np.max(np.abs(ours[:, 1] - ref[:, 1])) <= 1e-6 * scale


@pytest.mark.parametrize("name", GPU_CASES)
def test_job_run_matches_reference_power(name, tmp_path):
    print(f"[gpu job] {name}", flush=True)
    _stage(name, tmp_path)
    case = Case(name)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        job = Job(name + ".job")
        assert job.run(max_steps=case.n_steps, precision=abi.PRECISION_FP64) == case.n_steps
        job.close()
    finally:
        os.chdir(cwd)
    _compare(case, os.path.join(tmp_path, "power-sampling", "power-0.txt"), case.n_steps)


def test_cli_full_run(tmp_path):
    """`mithra job.job` runs the whole total-time and writes the reference's file layout."""
    name = "nsfd_o2_sc_seeded"
    _stage(name, tmp_path)
    exe = os.path.join(ROOT, "mithra_amd", "bin", "mithra")
    r = subprocess.run([exe, name + ".job"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    path = os.path.join(tmp_path, "power-sampling", "power-0.txt")
    with open(path) as f:
        first = f.readline()
    assert first.startswith(" " * 10) and first.rstrip("\n").endswith("\t")   # width(40), tabs
    case = Case(name)
    _compare(case, path, case.n_steps)
    assert len(np.loadtxt(path)) > case.n_steps       # continued past the fixture's 100 steps


def test_cli_rejects_bad_job(tmp_path):
    bad = os.path.join(tmp_path, "bad.job")
    with open(bad, "w") as f:
        f.write("MESH\n{\n  not-a-key = 1\n}\n")
    r = subprocess.run([os.path.join(ROOT, "mithra_amd", "bin", "mithra"), bad], capture_output=True, text=True)
    assert r.returncode == 1 and "not-a-key is not defined in solver group." in r.stdout
