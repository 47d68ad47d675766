"""Closed-form checks of the CPU oracle, independent of the reference's dumps.

Each test states a property the mathematics fixes, chosen so that a plausible slip (wrong sign,
swapped index, dropped term) fails it:
  - halton(i, j) is one minus the radical inverse of j+1 in base prime(i);
  - the zigzag deposit of one particle sums to q (rnp - rnm) in every component (charge
    conservation: the node weights of each half-segment add up to one);
  - with a pure magnetic field the Boris rotation conserves |gamma beta|;
  - at dt = dz/c the NSFD update advances a transversely uniform A(z - ct) by exactly one cell
    per step away from the boundaries (the reference's dispersion design, its manual eq. for
    the NSFD coefficients), which a wrong a[0] or a[3] breaks;
  - zero sources and zero initial fields stay zero.
"""
import ctypes as C

import numpy as np
import pytest

from golden_state import load_oracle
from mithra_amd import abi

PRIMES = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29]


@pytest.fixture(scope="module")
def lib():
    L = load_oracle()
    L.oracle_halton.restype, L.oracle_halton.argtypes = C.c_double, [C.c_uint, C.c_uint]
    return L


def radical_inverse(n, b):
    x, f = 0.0, 1.0 / b
    while n > 0:
        n, d = divmod(n, b)
        x += d * f
        f /= b
    return x


def test_halton_is_radical_inverse(lib):
    for i, p in enumerate(PRIMES):
        for j in (0, 1, 2, 7, 100, 4095, 123456):
            assert abs(lib.oracle_halton(i, j) - (1.0 - radical_inverse(j + 1, p))) < 1e-15


def small_grid(nx=10, ny=10, nz=24, solver=abi.SOLVER_FD, sc=0):
    g = abi.Grid()
    g.N0, g.N1, g.N2, g.solver, g.truncation_order, g.space_charge = nx, ny, nz, solver, 2, sc
    g.dx, g.dy, g.dz = 2.0, 2.0, 1.0
    g.xmin, g.ymin, g.zmin = -(nx - 1) / 2 * g.dx, -(ny - 1) / 2 * g.dy, -(nz - 1) / 2 * g.dz
    g.xmax, g.ymax, g.zmax = -g.xmin, -g.ymin, -g.zmin
    g.zp0, g.zp1, g.Lz = g.zmin, g.zmin + (nz - 1) * g.dz, (nz - 1) * g.dz
    g.dt = 1.0
    return g


def test_deposit_conserves_charge(lib):
    g = small_grid()
    c = abi.Coeffs()
    dom = abi.Domain(g, c, lib=lib, prefix="oracle_")
    rng = np.random.default_rng(3)
    for trial in range(50):
        q = rng.uniform(0.5, 2.0)
        rm = rng.uniform([-5, -5, -6], [5, 5, 6])
        rp = rm + rng.uniform(-0.9, 0.9, 3) * np.array([g.dx, g.dy, g.dz])
        rec = np.zeros(11)
        rec[0], rec[1:4], rec[4:7] = q, rp, rm
        dom.set_particles(rec[None])
        dom.current_reset()
        dom.current_update()
        j, _ = dom.get_potentials(2)
        tot = j.reshape(-1, 3).sum(axis=0)
        assert np.allclose(tot, q * (rp - rm), rtol=1e-12, atol=1e-14), trial


def test_boris_conserves_momentum_in_pure_b(lib):
    """Undulator field only (no grid E): the rotation is a pure magnetic one in the frame
    where E = 0, so choose beta = 0 (no boost) and check |gb| after many sub-steps."""
    g = small_grid()
    dom = abi.Domain(g, abi.Coeffs(), lib=lib, prefix="oracle_")
    und = [abi.Undulator(0, 1000, 2.0, 30.0, 0.0, 0.0, 0.0)]
    bp = abi.BunchParams(1.0, 0.0, 299.792458, 0.0, -0.01, -3.0, 0.05)
    rec = np.zeros((4, 11))
    rec[:, 0] = 1.0
    rec[:, 1:4] = [[0.1, 0.2, 1.0], [0.3, -0.2, 2.0], [-0.4, 0.1, 3.0], [0.0, 0.0, 4.0]]
    rec[:, 7:10] = [[0.5, 0.1, 3.0], [0.0, 0.2, 5.0], [0.1, 0.0, 1.0], [0.3, 0.3, 0.3]]
    rec[:, 10] = 0.0      # emission flag 0, beta 0: the grid field is never gathered
    dom.set_particles(rec)
    n0 = np.linalg.norm(rec[:, 7:10], axis=1)
    for s in range(200):
        dom.bunch_update(bp, und, 0.01 * s)
    p = dom.get_particles()
    assert np.allclose(np.linalg.norm(p[:, 7:10], axis=1), n0, rtol=1e-12)
    assert not np.allclose(p[:, 7:10], rec[:, 7:10])        # and it did rotate


def nsfd_coeffs(g, c0):
    dt, dx, dy, dz = g.dt, g.dx, g.dy, g.dz
    beta = (1.0 + 0.02 / ((dz / dx) ** 2 + (dz / dy) ** 2)) / 4.0
    alpha = 1.0 - 2.0 * beta
    c = abi.Coeffs()
    c.a[0] = 2.0 * (1.0 - alpha * (c0 * dt / dx) ** 2 - alpha * (c0 * dt / dy) ** 2 - (c0 * dt / dz) ** 2)
    c.a[1] = (c0 * dt / dx) ** 2
    c.a[2] = (c0 * dt / dy) ** 2
    c.a[3] = (c0 * dt / dz) ** 2 - 2.0 * (beta * (c0 * dt / dx) ** 2 + beta * (c0 * dt / dy) ** 2)
    c.alpha, c.beta = alpha, beta / alpha
    return c


def test_nsfd_plane_wave_moves_one_cell_per_step(lib):
    """A_x = f(z - c t) with f a smooth pulse: after n steps the interior equals the initial
    profile shifted by n cells (up to rounding), far from the transverse boundaries."""
    g = small_grid(nx=40, ny=40, nz=80, solver=abi.SOLVER_NSFD)
    g.truncation_order = 1      # faces only; zero face coefficients clamp the boundary nodes
    c0 = 1.0
    g.dt = g.dz / c0
    c = nsfd_coeffs(g, c0)
    dom = abi.Domain(g, c, lib=lib, prefix="oracle_")
    z = g.zmin + np.arange(g.N2) * g.dz
    f = lambda s: np.exp(-((s + 20.0) / 2.0) ** 2)  # noqa: E731
    shape = (g.N2, g.N0, g.N1, 3)
    an = np.zeros(shape)
    anm1 = np.zeros(shape)
    an[..., 0] = f(z)[:, None, None]
    anm1[..., 0] = f(z + c0 * g.dt)[:, None, None]
    dom.set_potentials(0, an.reshape(-1))
    dom.set_potentials(1, anm1.reshape(-1))
    steps = 15
    for _ in range(steps):
        dom.field_update(0.0, None)
        dom.field_shift()
        dom.current_reset()
    a, _ = dom.get_potentials(0)
    a = a.reshape(shape)[..., 0]
    expect = f(z - steps * c0 * g.dt)
    core = a[10:60, 18:22, 18:22]      # out of reach (c t = 15 cells) of every boundary
    assert np.allclose(core, expect[10:60, None, None], atol=1e-12)


def test_zero_stays_zero(lib):
    g = small_grid(solver=abi.SOLVER_NSFD, sc=1)
    g.truncation_order = 1
    dom = abi.Domain(g, nsfd_coeffs(g, 1.0), lib=lib, prefix="oracle_")
    for _ in range(5):
        dom.field_update(0.0, None)
        dom.field_shift()
        dom.current_reset()
    a, phi = dom.get_potentials(0)
    e, b = dom.get_fields()
    assert not a.any() and not phi.any() and not e.any() and not b.any()


def test_bunching_closed_form(lib):
    """Particles at z_n = n lambda / M (M equal charges, one period) have bunching 0; all at one
    phase have bunching 1; a density 1 + a cos(kz) sampled finely has bunching a / 2."""
    g = small_grid()
    dom = abi.Domain(g, abi.Coeffs(), lib=lib, prefix="oracle_")
    lam = 3.0
    k = 2 * np.pi / lam
    M = 64
    rec = np.zeros((M, 11))
    rec[:, 0] = 1.0
    rec[:, 3] = np.arange(M) * lam / M
    dom.set_particles(rec)
    assert dom.bunching(k) < 1e-14
    rec[:, 3] = 1.234 + lam * np.arange(M)
    dom.set_particles(rec)
    assert abs(dom.bunching(k) - 1.0) < 1e-13
    z = (np.arange(4096) + 0.5) * lam / 4096
    rec = np.zeros((4096, 11))
    rec[:, 0] = 1.0 + 0.3 * np.cos(k * z)
    rec[:, 3] = z
    dom.set_particles(rec)
    assert abs(dom.bunching(k) - 0.15) < 1e-12
