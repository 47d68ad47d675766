"""Set-up of the diagnostics writers on the host (no GPU): the files and directories the
reference creates while it initializes (src/solver.cpp:848-1126, :2145-2199), the reference's
errors for a writer without a rhythm, and the file comparison the GPU tests use.
"""
import os
import shutil

import pytest

from diag_files import compare, reference_files
from golden_state import GOLDEN
from mithra_amd import abi
from mithra_amd.job import Job

DIAG = os.path.join(GOLDEN, "diag")


def open_job(name, text, tmp_path):
    path = os.path.join(tmp_path, name + ".job")
    with open(path, "w") as f:
        f.write(text)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        return Job(name + ".job")
    finally:
        os.chdir(cwd)


def test_set_up_creates_the_reference_files(tmp_path):
    job = open_job("diag_nsfd_o2", open(os.path.join(DIAG, "diag_nsfd_o2.job")).read(), tmp_path)
    job.close()
    made = sorted(os.path.relpath(os.path.join(r, f), tmp_path) for r, _, fs in os.walk(tmp_path) for f in fs)
    # files opened at set-up (and still empty): sampling streams, screens, power
    assert made == sorted(["diag_nsfd_o2.job", "field-sampling/field-0.txt", "bunch-sampling/bunch.txt",
                           "screen/bunch-p0-screen0.txt", "screen/bunch-p0-screen1.txt",
                           "screen/bunch-p0-screen2.txt", "power-sampling/power-0.txt"])
    for d in ("vtk", "profile", "bunch-vtk", "bunch-profile"):
        assert os.path.isdir(os.path.join(tmp_path, d))
    ref = reference_files(os.path.join(DIAG, "diag_nsfd_o2.tar.xz"))
    for f in made[1:]:
        assert f in ref or f == "diag_nsfd_o2.job"


@pytest.mark.parametrize("group,key,message", [
    ("field-sampling", "rhythm                = 47.0",
     "The sampling rhythm of the field is zero although sampling is activated !!!"),
    ("field-visualization", "rhythm                = 275.2",
     "The visualization rhythm of the field is zero although visualization is activated !!!"),
    ("field-profile", "rhythm                = 275.2",
     "The profiling rhythm of the field is zero and no time is set although profiling of the field is activated !!!"),
    ("bunch-sampling", "rhythm                = 20.14",
     "The sampling rhythm of the bunch is zero although sampling is activated !!!"),
])
def test_writer_without_rhythm_is_the_reference_error(group, key, message, tmp_path):
    text = open(os.path.join(DIAG, "diag_nsfd_o2.job")).read()
    at = text.index(group)
    assert key in text[at:]
    text = text[:at] + text[at:].replace(key, "rhythm                = 0.0", 1)
    with pytest.raises(abi.MithraError, match=message.replace("!", r"\!").replace("(", r"\(")):
        open_job("norhythm", text, tmp_path)


def test_unknown_field_name_is_rejected(tmp_path):
    text = open(os.path.join(DIAG, "diag_nsfd_o2.job")).read().replace("field                 = Az", "field = Q", 1)
    with pytest.raises(abi.MithraError, match="Q is an unknown sampling field"):
        open_job("badfield", text, tmp_path)


def test_points_outside_the_mesh_are_dropped(tmp_path):
    text = open(os.path.join(DIAG, "diag_nsfd_o2.job")).read()
    # every point outside: no sampling file is opened (src/solver.cpp:913-921)
    for p in ("( 37.0, -121.5, 6.1 )", "( -200.3, 150.7, -12.9 )", "( 5.0, 5.0, 2.2 )"):
        text = text.replace(p, "( 480.0, 0.0, 0.0 )")
    job = open_job("outside", text, tmp_path)
    job.close()
    assert not os.path.exists(os.path.join(tmp_path, "field-sampling", "field-0.txt"))


def test_file_comparison_catches_planted_mistakes():
    ref = reference_files(os.path.join(DIAG, "diag_nsfd_o2.tar.xz"))
    text = ref["field-sampling/field-0.txt"]
    assert compare(text, text)[0]
    # one unit in the last printed digit passes, two fail
    first = text.split("\t")[1]
    mant, expo = first.split("e")
    bump = lambda d: f"{float(mant) + d:.4f}e{expo}"  # noqa: E731
    assert compare(text, text.replace(first, bump(1e-4), 1))[0]
    assert not compare(text, text.replace(first, bump(3e-4), 1))[0]
    # a changed tag, a missing line, a 15-digit value off by 1e-8 of its column
    vts = ref["vtk/field-all-21.pvts"]
    assert not compare(vts, vts.replace("GhostLevel", "Ghostlevel"))[0]
    prof = ref["bunch-profile/bunch-p0-8.txt"]
    assert not compare(prof, "\n".join(prof.split("\n")[:-3]) + "\n")[0]
    line = prof.split("\n")[5]
    x = line.split("\t")[1]
    y = f"{float(x) * (1 + 1e-8):.15e}"
    assert not compare(prof, prof.replace(x, y, 1))[0]
