"""Comparison of diagnostics files written by the drop-in with the reference's own files.

Test infrastructure.  Two files match when their text outside the numbers is identical
(tags, names, separators, padding) and every number agrees:

  * integers exactly (extents, counts, file-name digits);
  * numbers printed with 4 significant decimals (the samplers, VTK files and field profiles)
    within one unit of the last printed digit, or within 1e-9 of the largest magnitude in the
    same column of the file (values that are zero up to rounding of the march);
  * numbers printed with 15 decimals (bunch profiles, screens, power) within 1e-10 of the
    largest magnitude in the same column: the bar of the coupled march itself.
"""
import re
import tarfile

import numpy as np

NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")


def reference_files(archive):
    out = {}
    with tarfile.open(archive, "r:xz") as tar:
        for m in tar.getmembers():
            if m.isfile():
                out[m.name] = tar.extractfile(m).read().decode()
    return out


def _tokens(text):
    lines = text.split("\n")
    skel, nums = [], []
    for li, line in enumerate(lines):
        skel.append(NUM.sub("#", line))
        for ci, tok in enumerate(NUM.findall(line)):
            nums.append((li, ci, tok))
    return skel, nums


def _decimals(tok):
    m = re.match(r"[-+]?\d*\.(\d*)", tok)
    return len(m.group(1)) if m else None


def compare(ref, ours):
    """(ok, message, exact_fraction)."""
    sr, nr = _tokens(ref)
    so, no = _tokens(ours)
    if sr != so:
        for i, (a, b) in enumerate(zip(sr, so)):
            if a != b:
                return False, f"text differs at line {i}: {a!r} vs {b!r}", 0.0
        return False, f"line counts differ: {len(sr)} vs {len(so)}", 0.0
    if len(nr) != len(no):
        return False, "number counts differ", 0.0
    colmax = {}
    for (li, ci, tok) in nr:
        colmax[ci] = max(colmax.get(ci, 0.0), abs(float(tok)))
    exact = 0
    for (li, ci, a), (_, _, b) in zip(nr, no):
        if a == b:
            exact += 1
            continue
        dec = _decimals(a)
        if dec is None:
            return False, f"line {li} column {ci}: integer {a} vs {b}", 0.0
        x, y = float(a), float(b)
        if dec <= 6:
            expo = int(np.floor(np.log10(max(abs(x), abs(y))))) if max(abs(x), abs(y)) > 0 else 0
            tol = max(1.01 * 10.0 ** (expo - dec), 1e-9 * colmax[ci])
        else:
            tol = 1e-10 * colmax[ci]
        if abs(x - y) > tol:
            return False, f"line {li} column {ci}: {a} vs {b} (tolerance {tol:.3e})", 0.0
    return True, "", exact / max(len(nr), 1)
