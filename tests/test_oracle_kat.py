"""The oracle's analytic seed against the reference's own known-answer tables.

tests/golden/kat_seed_signal.npz holds values computed by the reference itself
(oracle/kat_harness.cpp, built from the reference sources): Signal::self for all five signal
types (src/classes.cpp:534-575) over times that cross every branch, and Seed::fields for all four
seed types with every signal (src/classes.cpp:740-855), called in sequence on one output vector as
the TF/SF loop calls it (src/fdtd.cpp:239, :315-371).  Both sides are plain g++ -O3 builds over
the same libm, so the oracle must reproduce every value bit for bit."""
import ctypes as C
import os

import numpy as np
import pytest

from golden_state import load_oracle
from mithra_amd import abi

KAT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kat_seed_signal.npz")
SIGNALS = ("neumann", "gaussian", "secant-hyperbolic", "flat-top", "inverse-gaussian")
SEEDS = ("plane-wave", "truncated-plane-wave", "gaussian-beam", "super-gaussian-beam")


@pytest.fixture(scope="module")
def kat():
    return np.load(KAT)


@pytest.fixture(scope="module")
def lib():
    L = load_oracle()
    L.oracle_signal.restype = C.c_double
    L.oracle_signal.argtypes = [C.POINTER(abi.Seed), C.c_double, C.c_double]
    L.oracle_seed_fields.restype = None
    L.oracle_seed_fields.argtypes = [C.POINTER(abi.Seed), C.POINTER(C.c_double), C.c_double,
                                     C.POINTER(C.c_double)]
    return L


def signal_struct(params, sd=None):
    sd = sd or abi.Seed()
    sd.signal_type = int(params[0])
    sd.t0, sd.s, sd.f0, sd.cep = params[1], params[2], params[3], params[4]
    sd.nR = int(params[5])
    sd.sigma_invg[:] = [params[6], params[7]]
    return sd


@pytest.mark.parametrize("idx", range(10))
def test_signal_bit_exact(kat, lib, idx):
    p = kat[f"signal{idx}_params"]
    sd = signal_struct(p)
    t, ph, ref = kat[f"signal{idx}_t"], kat[f"signal{idx}_phase"], kat[f"signal{idx}_value"]
    got = np.array([lib.oracle_signal(C.byref(sd), a, b) for a, b in zip(t, ph)])
    assert np.array_equal(got, ref), f"{SIGNALS[int(p[0])]}: max |diff| {np.max(np.abs(got - ref))}"
    # the table reaches every branch: beyond the 10 s cut-off, and for the flat-top and
    # inverse-gaussian types both rising edges and the flat part
    u = (t - p[1]) / p[2]
    assert np.any(np.abs(u) > 10.0) and np.all(ref[np.abs(u) > 10.0] == 0.0)
    for region in (u <= -0.5, np.abs(u) < 0.5, u > 0.5):
        assert np.any(region & (ref != 0.0))


def test_signal_closed_forms(kat):
    """What the formulas fix without the reference: inside the flat part the flat-top signal is
    the bare carrier, and the Gaussian signal's intensity falls to one half at t0 +- s/2 (the
    half-maximum definition of s, src/classes.h:177-178) up to the rounding of its 1.3863."""
    for idx in range(10):
        p = kat[f"signal{idx}_params"]
        t, ph, v = kat[f"signal{idx}_t"], kat[f"signal{idx}_phase"], kat[f"signal{idx}_value"]
        carrier = np.cos(2 * 3.1415926535 * p[3] * (t - p[1]) + p[4] + ph)
        if int(p[0]) == abi_signal("flat-top"):
            flat = np.abs(t - p[1]) <= p[2] / 2
            assert np.allclose(v[flat], carrier[flat], rtol=0, atol=1e-15)
        if int(p[0]) == abi_signal("gaussian"):
            env = np.exp(-1.3863 * ((t - p[1]) / p[2]) ** 2)
            assert np.allclose(v, np.where(np.abs(t - p[1]) > 10 * p[2], 0.0, carrier * env), rtol=1e-13,
                               atol=1e-300)
            assert abs(np.exp(-1.3863 * 0.25) ** 2 - 0.5) < 1e-4


def abi_signal(name):
    return SIGNALS.index(name)


@pytest.mark.parametrize("idx", range(40))
def test_seed_fields_bit_exact(kat, lib, idx):
    desc, sig = kat[f"seed{idx}_desc"], kat[f"seed{idx}_signal"]
    sd = signal_struct(sig)
    sd.seed_type = int(desc[0])
    sd.amplitude = desc[1]
    sd.radius[:] = [desc[2], desc[3]]
    sd.order[:] = [int(desc[4]), int(desc[5])]
    sd.gamma, sd.beta, sd.dt_shift, sd.c0 = desc[6], desc[7], desc[8], desc[9]
    sd.position[:] = list(desc[10:13])
    sd.direction[:] = list(desc[13:16])
    sd.polarization[:] = list(desc[16:19])
    r = kat["seed_r"].reshape(-1, 3)
    times = kat["seed_time"]
    ref = kat[f"seed{idx}_a"].reshape(-1, 3)
    a = (C.c_double * 3)(0.0, 0.0, 0.0)  # one vector for every call, like the TF/SF loop's atemp
    got = np.empty_like(ref)
    for q in range(len(times)):
        rr = (C.c_double * 3)(*r[q])
        lib.oracle_seed_fields(C.byref(sd), rr, times[q], a)
        got[q] = a[:]
    name = f"{SEEDS[sd.seed_type]}/{SIGNALS[sd.signal_type]}"
    assert np.array_equal(got, ref), f"{name}: max |diff| {np.max(np.abs(got - ref))}"
    nonzero = np.any(ref != 0.0, axis=1)
    assert nonzero.any() and (~nonzero).any(), name  # both the field and its zero branch
