"""Generate the golden fixtures of the parity tests from the real MITHRA reference.

Test infrastructure, run in the development container only (the reference sources are not
on the GPU machines).  It builds oracle/_ref/ref_harness (oracle/Makefile), runs it on every
job in tests/golden/jobs/ and packs what it recorded into small compressed fixtures:

  <case>.npz        state before step 0, 1 and N (N = 100, or the job's `# golden-steps: N`)
                    (A^n, A^{n-1}, J, phi, rho, particles, all
                    update coefficients) and after the field update of steps 0 and 1
                    (A^{n+1}, E, B) -- see oracle/ref_harness.cpp for the exact hook points
  <case>_power.txt  the reference's own P(z) output file (radiation.cpp:222-229, precision 15)

Usage:  python tests/golden/make_golden.py [case ...]
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
from bunchfile import write_bunch_file  # noqa: E402
from dumpio import read_dump  # noqa: E402

HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")
N_STEPS = 100
PRE_KEYS = ("an", "anm1", "jn", "fn", "fnm1", "rn", "particles")
POST_KEYS = ("anp1", "fnp1", "en", "bn")


def job_steps(job):
    """Steps a fixture records: N_STEPS, or a `# golden-steps: N` line of the job file (the
    reference's reader drops everything after '#', src/readdata.cpp:44)."""
    for line in open(job):
        if line.startswith("# golden-steps:"):
            return int(line.split(":")[1])
    return N_STEPS


def run_case(job, steps=None):
    name = os.path.splitext(os.path.basename(job))[0]
    steps = job_steps(job) if steps is None else steps
    work = tempfile.mkdtemp(prefix=f"golden_{name}_")
    try:
        shutil.copy(job, work)
        if "file-name" in open(job).read():
            write_bunch_file(os.path.join(work, "bunch_e.txt"), 1000, seed=5)
        env = dict(os.environ, MITHRA_DUMP_DIR=work, MITHRA_DUMP_STEPS=f"0,1,{steps}",
                   MITHRA_MAX_STEPS=str(steps))
        with open(os.path.join(work, "log.txt"), "w") as log:
            subprocess.run([HARNESS, os.path.basename(job)], cwd=work, env=env, check=True,
                           stdout=log, stderr=subprocess.STDOUT, timeout=600)
        out = {}
        pre0 = read_dump(os.path.join(work, "pre_000000.bin"))
        for k, v in pre0.items():
            out[("pre0_" + k) if k in PRE_KEYS else k] = np.asarray(v)
        for tag, step in (("pre1_", 1), ("preN_", steps)):
            d = read_dump(os.path.join(work, f"pre_{step:06d}.bin"))
            for k in PRE_KEYS + ("nTime", "time", "timeBunch"):
                if k in d:
                    out[tag + k] = np.asarray(d[k])
        for tag, step in (("post0_", 0), ("post1_", 1)):
            d = read_dump(os.path.join(work, f"post_{step:06d}.bin"))
            for k in POST_KEYS:
                if k in d:
                    out[tag + k] = np.asarray(d[k])
        counts = {t: out[t + "particles"].size // 11 for t in ("pre0_", "pre1_", "preN_")}
        if len(set(counts.values())) != 1:
            raise RuntimeError(f"{name}: particle count changed {counts}; the bunch crossed a z "
                               "boundary, which duplicates particles on one rank")
        out["n_steps"] = np.asarray(steps)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        shutil.copy(os.path.join(work, "power-sampling", "power-0.txt"),
                    os.path.join(HERE, name + "_power.txt"))
        print(f"{name}: nodes {out['N0']}x{out['N1']}x{out['N2']}, particles {counts['pre0_']}, "
              f"steps {steps}")
    finally:
        shutil.rmtree(work, ignore_errors=True)


def main(argv):
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    jobs = sorted(os.path.join(HERE, "jobs", j) for j in os.listdir(os.path.join(HERE, "jobs"))
                  if j.endswith(".job"))
    if argv:
        jobs = [j for j in jobs if os.path.splitext(os.path.basename(j))[0] in argv]
    for job in jobs:
        run_case(job)


if __name__ == "__main__":
    main(sys.argv[1:])
