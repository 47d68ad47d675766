"""The device's analytic seed against the reference's own known-answer tables.

tests/golden/kat_seed_signal.npz holds Signal::self for all five signal types and Seed::fields
for all four seed types with every signal, computed by the reference itself
(oracle/kat_harness.cpp).  The device evaluates the same functions the TF/SF injection uses
(mithra_signal_evaluate, mithra_seed_evaluate).  The device's cos, exp, pow, atan and cosh differ
from the host libm by an ulp or so, so values agree to 1e-12 of each table's largest magnitude,
and every branch decision (the 10 s window, the 1e-6 signal threshold, the truncation radius)
must agree exactly.  The super-gaussian type adds into the caller's previous output: the table
evaluates its points in turn into one vector, and the device's point evaluator does the same."""
import numpy as np
import pytest

from mithra_amd import abi
from test_oracle_kat import KAT, SEEDS, SIGNALS, signal_struct

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def kat():
    return np.load(KAT)


@pytest.mark.parametrize("idx", range(10))
def test_device_signal(kat, idx):
    p = kat[f"signal{idx}_params"]
    sd = signal_struct(p)
    t, ph, ref = kat[f"signal{idx}_t"], kat[f"signal{idx}_phase"], kat[f"signal{idx}_value"]
    got = abi.signal_evaluate(sd, t, ph)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"signal {SIGNALS[int(p[0])]} (set {idx // 5}): {t.size} values, max error {err:.2e}", flush=True)
    assert np.array_equal(got == 0.0, ref == 0.0)
    assert err <= TOL


def seed_struct(kat, idx):
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
    return sd


@pytest.mark.parametrize("idx", range(40))
def test_device_seed_fields(kat, idx):
    sd = seed_struct(kat, idx)
    name = f"{SEEDS[sd.seed_type]}/{SIGNALS[sd.signal_type]} (geometry {idx // 20})"
    r, times = kat["seed_r"].reshape(-1, 3), kat["seed_time"]
    ref = kat[f"seed{idx}_a"].reshape(-1, 3)
    got = abi.seed_evaluate(sd, r, times)
    scale = np.max(np.abs(ref))
    err = np.max(np.abs(got - ref)) / scale
    print(f"{name}: {len(times)} points, {np.count_nonzero(np.any(ref != 0, axis=1))} non-zero, "
          f"max error {err:.2e}", flush=True)
    assert np.array_equal(got == 0.0, ref == 0.0), name
    assert err <= TOL, name
