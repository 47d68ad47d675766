"""Bunches generated and boosted on the device (mithra_bunch_*, the drop-in host's device mode).

1. The reference's ellipsoid generator (Bunch::initializeEllipsoid, src/classes.cpp:104-298) and
   the particle loops of Solver::lorentzBoostBunch / shiftBackInTime on the device, against the
   reference's own particles before step 0 in every golden fixture with an ellipsoid bunch: the
   same records in the same order, within 1e-13 of the position / momentum scale (the halton
   numbers are exact; the device's log, sin and cos are an ulp or so from the host's), and the
   same frame time shift dt_ to 1e-14.
2. A job set up from a device-generated bunch marches to the reference's own P(z) file: within
   1e-6 of its peak after the fixture's steps (the bar of the drop-in).
3. The seeded flat-top generator against what its definition fixes: the count and the charge, the
   slice bounds, the four beamlets a quarter wavelength apart inside one bucket's modulation, the
   Gaussian moments, and one record set per seed."""
import os
import shutil

import numpy as np
import pytest

from golden_state import CASES, GOLDEN, Case, march, rel_l2
from mithra_amd import abi
from mithra_amd.job import Job

pytestmark = pytest.mark.gpu

ELLIPSOID_CASES = [c for c in CASES if "type                  = ellipsoid" in
                   open(os.path.join(GOLDEN, "jobs", c + ".job")).read() and "3D-crystal" not in
                   open(os.path.join(GOLDEN, "jobs", c + ".job")).read()]


def staged_job(name, tmp_path, **kw):
    shutil.copy(os.path.join(GOLDEN, "jobs", name + ".job"), tmp_path)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        return Job(name + ".job", **kw)
    finally:
        os.chdir(cwd)


@pytest.mark.parametrize("name", ELLIPSOID_CASES)
def test_device_ellipsoid_matches_reference(name, tmp_path):
    case = Case(name)
    ref = case.arr("pre0_particles").reshape(-1, 11)
    dev = staged_job(name, tmp_path, device_bunch=True)
    host = staged_job(name, tmp_path)
    q = dev.particles()
    print(f"{name}: {q.shape[0]} device records, {ref.shape[0]} reference records", flush=True)
    assert q.shape == ref.shape
    assert np.array_equal(q[:, 0], ref[:, 0]) and np.array_equal(q[:, 10], ref[:, 10])
    scale_r = np.max(np.abs(ref[:, 1:4]))
    scale_g = np.max(np.abs(ref[:, 7:10]))
    err_r = np.max(np.abs(q[:, 1:4] - ref[:, 1:4])) / scale_r
    err_g = np.max(np.abs(q[:, 7:10] - ref[:, 7:10])) / scale_g
    print(f"  positions {err_r:.2e}, momenta {err_g:.2e} of their scales", flush=True)
    assert err_r <= 1e-13 and err_g <= 1e-13
    bd, bh = dev.bunch_params(), host.bunch_params()
    assert abs(bd.dt_shift - bh.dt_shift) <= 1e-14 * abs(bh.dt_shift)
    assert bh.dt_shift == case.dt_shift
    dev.close()
    host.close()


@pytest.mark.parametrize("name", [c for c in ELLIPSOID_CASES if c in ("nsfd_o2", "fd_o1_array")])
def test_march_from_device_bunch_matches_reference_power(name, tmp_path):
    case = Case(name)
    job = staged_job(name, tmp_path, device_bunch=True)
    dom = abi.Domain(job.grid(), job.coeffs())
    job.fill(dom)
    dom.power_setup(case.power, case.n_steps)
    march(case, dom, case.n_steps)
    P = dom.power_fetch(0, case.n_steps, case.power_N, case.power_Nl)[:, 0, 0]
    P_ref = case.power_file()[: case.n_steps, 1]
    err = np.max(np.abs(P - P_ref)) / np.max(np.abs(P_ref))
    print(f"{name}: P(z) from the device bunch vs the reference's file: {err:.2e} of its peak", flush=True)
    assert err <= 1e-6
    dom.close()
    job.close()


def flattop_spec(n=400_000, length=0.28, lam=4.7e-3, ne=2.0e5):
    f = abi.Flattop()
    f.n, f.charge, f.gamma, f.lambda_, f.length, f.z_center = n, 1.0e7, 1957.9, lam, length, 0.05
    f.sigma_r, f.sigma_gb_perp, f.energy_spread, f.electrons_per_bucket = 50.0, 0.02, 1e-4, ne
    return f


def test_flattop_generator_definition():
    f = flattop_spec()
    b = abi.DeviceBunch()
    b.flattop(f, seed=7)
    q = b.records()
    assert q.shape == (f.n, 11)
    assert np.allclose(q[:, 0], f.charge / f.n, rtol=0, atol=0) and abs(q[:, 0].sum() - f.charge) < 1e-6 * f.charge
    z = q[:, 3].reshape(-1, 4)
    lo, hi = f.z_center - f.length / 2, f.z_center + f.length / 2
    amp_max = f.lambda_ / np.pi * np.sqrt(-2.0 * np.log(2.0 ** -53)) / np.sqrt(f.electrons_per_bucket)
    assert z.min() >= lo - f.lambda_ - amp_max and z.max() <= hi + amp_max
    # beamlets: z_ii + lambda/4 ii differ only by their modulation (within 2 amp of each other)
    base = z + f.lambda_ / 4.0 * np.arange(4)
    assert np.max(np.ptp(base, axis=1)) <= 2.0 * amp_max
    # one x, y, gb per base particle; Gaussian moments
    for col, s in ((1, f.sigma_r), (2, f.sigma_r), (7, f.sigma_gb_perp), (8, f.sigma_gb_perp)):
        v = q[:, col].reshape(-1, 4)
        assert np.all(v == v[:, :1])
        m, sd = v[:, 0].mean(), v[:, 0].std()
        n = v.shape[0]
        assert abs(m) < 5 * s / np.sqrt(n) and abs(sd - s) < 5 * s / np.sqrt(2 * n)
    gz = q[:, 9].reshape(-1, 4)[:, 0]
    g0 = np.sqrt(f.gamma ** 2 - 1.0)
    assert abs(gz.mean() / g0 - 1.0) < 5e-4 / np.sqrt(len(gz)) * 10 and abs(gz.std() / g0 - f.energy_spread) < 0.05 * f.energy_spread
    # the quiet start removes the bunching at lambda down to the imposed shot noise
    kz = 2.0 * np.pi / f.lambda_ * q[:, 3]
    bunching = abs(np.exp(1j * kz).mean())
    print(f"flat-top: bunching at lambda {bunching:.3e}", flush=True)
    assert bunching < 20.0 / np.sqrt(f.electrons_per_bucket * f.length / f.lambda_)
    # one record set per seed
    b2 = abi.DeviceBunch()
    b2.flattop(f, seed=7)
    assert np.array_equal(b2.records(), q)
    b3 = abi.DeviceBunch()
    b3.flattop(f, seed=8)
    assert not np.array_equal(b3.records()[:, 1], q[:, 1])
    for x in (b, b2, b3):
        x.close()
