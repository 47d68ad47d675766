"""Fixtures of the diagnostics writers: the reference's own output files for the jobs in
tests/golden/diag/ (field sampling at points and over a line, field visualization over the whole
domain and in the three plane orientations, field profiles, bunch sampling, bunch visualization,
bunch profiles and lab-frame screens, with and without space charge).

Test infrastructure, run in the development container only: it builds the reference with
oracle/Makefile, runs oracle/_ref/ref_harness (the reference driven unmodified; the harness only
stops it after the job's `# golden-steps: N` field steps) on each job and packs every file the
run wrote, except the log, into tests/golden/diag/<job>.tar.xz.

Usage:  python tests/golden/make_diag.py [job ...]
"""
import os
import shutil
import subprocess
import sys
import tarfile
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
DIAG = os.path.join(HERE, "diag")
HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")
sys.path.insert(0, HERE)
from make_golden import job_steps  # noqa: E402


def run(name):
    job = os.path.join(DIAG, name + ".job")
    steps = job_steps(job)
    work = tempfile.mkdtemp(prefix=f"diag_{name}_")
    try:
        shutil.copy(job, work)
        env = dict(os.environ, MITHRA_MAX_STEPS=str(steps))
        env.pop("MITHRA_DUMP_STEPS", None)
        with open(os.path.join(work, "log.txt"), "w") as log:
            subprocess.run([HARNESS, name + ".job"], cwd=work, env=env, check=True, stdout=log,
                           stderr=subprocess.STDOUT, timeout=600)
        files = []
        for root, _, names in os.walk(work):
            for f in names:
                rel = os.path.relpath(os.path.join(root, f), work)
                if rel not in ("log.txt", name + ".job"):
                    files.append(rel)
        out = os.path.join(DIAG, name + ".tar.xz")
        with tarfile.open(out, "w:xz") as tar:
            for rel in sorted(files):
                info = tar.gettarinfo(os.path.join(work, rel), arcname=rel)
                info.mtime = 0
                info.uid = info.gid = 0
                info.uname = info.gname = ""
                with open(os.path.join(work, rel), "rb") as fh:
                    tar.addfile(info, fh)
        print(f"{name}: {len(files)} files after {steps} steps -> {os.path.getsize(out)} bytes")
    finally:
        shutil.rmtree(work, ignore_errors=True)


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    names = sys.argv[1:] or sorted(os.path.splitext(f)[0] for f in os.listdir(DIAG) if f.endswith(".job"))
    for name in names:
        run(name)


if __name__ == "__main__":
    main()
