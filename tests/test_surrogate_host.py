"""Host side of the surrogate workflow (no GPU): the space-charge, seeded generator's points and
job files, the analytic current-profile input, the target transform and the per-point error and
decision arithmetic that the workflow's points-per-hour figures rest on."""
import math

import numpy as np

from mithra_amd import scan, surrogate


def test_generator_draws_seeded_space_charge_points_in_range():
    pts = surrogate.generate(200, seed=5)
    again = surrogate.generate(200, seed=5)
    assert pts == again  # the same seed, the same scan
    g = np.array([p.gamma for p in pts])
    K = np.array([p.K for p in pts])
    cur = np.array([p.current for p in pts])
    assert 40.0 <= g.min() and g.max() <= 80.0 and 0.6 <= K.min() and K.max() <= 1.0
    assert 250.0 <= cur.min() and cur.max() <= 1000.0
    assert all(p.space_charge and p.quiet and p.seed_power == surrogate.SEED_POWER for p in pts)
    assert all(p.n_steps == surrogate.PASSES * p.n_z and p.every == surrogate.PASSES for p in pts)
    far = surrogate.generate(50, seed=7, outside=True)
    assert min(p.gamma for p in far) >= 90.0 and min(p.K for p in far) >= 1.1
    assert min(p.current for p in far) >= 1500.0


def test_job_text_carries_space_charge_seed_and_quiet_start(tmp_path):
    p = surrogate.generate(1, seed=3)[0]
    text = scan.job_text(p, 0, str(tmp_path))
    assert "space-charge            = true" in text
    assert "shot-noise            = false" in text and "bunching-factor       = 0" in text
    assert "gaussian-beam" in text
    # the seed's strength parameter is that of a 1 MW Gaussian beam of waist 10 sigma_r at lambda_r
    sr = scan.bunch_sizes(p)[0]
    a0 = scan.seed_strength(surrogate.SEED_POWER, 10.0 * sr, p.lambda_r)
    assert f"strength-parameter    = {a0:.10g}" in text
    e0 = math.sqrt(4.0 * 376.730313668 * 1e6 / (math.pi * (10.0 * sr * 1e-6) ** 2))
    assert abs(a0 - 1.602e-19 * e0 * p.lambda_r * 1e-6 / (2 * math.pi * 9.109e-31 * 2.99792458e8 ** 2)) < 1e-12 * a0


def test_profile_is_the_truncated_gaussian_over_the_window():
    p = surrogate.generate(1, seed=9)[0]
    h = surrogate.profile(p)
    assert h.shape == (surrogate.PROFILE_BINS,) and h.max() == 1.0 and h.min() >= 0.0
    assert np.allclose(h, h[::-1], atol=1e-12)  # centred bunch
    lz = (p.n_z - 1.5) * p.dz
    sz = scan.bunch_sizes(p)[1]
    centres = -lz / 2 + (np.arange(surrogate.PROFILE_BINS) + 0.5) * lz / surrogate.PROFILE_BINS
    inside = np.abs(centres) < 3.4 * sz
    want = np.exp(-0.5 * (centres / sz) ** 2)
    # bin integrals of a Gaussian against its midpoint values, normalised at the peak
    assert np.allclose(h[inside], want[inside] / want.max(), rtol=2e-3, atol=2e-3)
    assert np.all(h[np.abs(centres) > 3.6 * sz] == 0.0)
    x = surrogate.features([p])
    assert x.shape == (1, surrogate.DIMS[0]) and np.array_equal(x[0, 8:], h)


def test_targets_round_trip_and_point_error():
    rng = np.random.default_rng(0)
    P = rng.uniform(0.0, 3e5, (4, scan.SAMPLES))
    b = rng.uniform(0.0, 0.3, (4, scan.SAMPLES))
    y = surrogate.targets(list(zip(P, b)))
    assert np.allclose(surrogate.power_from_targets(y), P, rtol=1e-12)
    assert np.array_equal(y[:, scan.SAMPLES:], b)
    assert surrogate.p_error(P[0], P[0]) == 0.0
    assert abs(surrogate.p_error(1.02 * P[0], P[0]) - 0.02) < 1e-12
