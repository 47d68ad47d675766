"""Known-answer tables of the reference's analytic seed (test infrastructure, run in the
development container only; the reference sources are not on the GPU machines).

Builds oracle/_ref/kat_harness (oracle/Makefile), which evaluates the reference's own
Signal::self for every signal type and Seed::fields for every seed type and signal (see
oracle/kat_harness.cpp), and packs its output into tests/golden/kat_seed_signal.npz.

Usage:  python tests/golden/make_kat.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
from dumpio import read_dump  # noqa: E402


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    with tempfile.TemporaryDirectory() as work:
        path = os.path.join(work, "kat.bin")
        subprocess.run([os.path.join(REPO, "oracle", "_ref", "kat_harness"), path], check=True, timeout=120)
        d = read_dump(path)
    out = {k: np.atleast_1d(np.asarray(v)) for k, v in d.items()}
    np.savez_compressed(os.path.join(HERE, "kat_seed_signal.npz"), **out)
    print(f"kat_seed_signal.npz: {len(out)} arrays")


if __name__ == "__main__":
    main()
