"""Both C-ABI libraries load and export every entry point include/*.h declares (no GPU work)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBS = {"mithra_gpu.h": "libmithra_gpu.so", "mithra_host.h": "libmithra.so"}


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mithra_\w+)\s*\(", text)))


@pytest.mark.parametrize("header", sorted(LIBS))
def test_exports(header):
    path = os.path.join(ROOT, "mithra_amd", "lib", LIBS[header])
    assert os.path.exists(path), f"{path} not built"
    lib = C.CDLL(path)
    names = declared(header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings():
    lib = C.CDLL(os.path.join(ROOT, "mithra_amd", "lib", "libmithra_gpu.so"))
    lib.mithra_strerror.restype = C.c_char_p
    assert lib.mithra_strerror(0) == b"success"
    assert b"not supported" in lib.mithra_strerror(-2)


def test_host_rejects_bad_jobs(tmp_path):
    """Parser errors come back as status codes with the reference's message texts."""
    from mithra_amd.job import Job
    from mithra_amd.abi import MithraError
    cases = {
        "MESH\n{\n  bogus = 3\n}\n": "bogus is not defined in solver group.",
        "NOPE\n{\n}\n": "NOPE is not a defined group.",
        "MESH\n{\n  mesh-truncation-order = 3\n}\n": "Mesh truncation order can not be different from one or two.",
        "MESH\n{\n  space-charge = maybe\n}\n": "got unexpected input",
    }
    for text, msg in cases.items():
        p = tmp_path / "j.job"
        p.write_text(text)
        with pytest.raises(MithraError) as e:
            Job(str(p), initialize=False)
        assert msg in str(e.value)
