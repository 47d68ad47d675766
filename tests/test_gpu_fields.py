"""E/B evaluation on the GPU against the oracle on random potentials (FdTd::fieldEvaluate,
src/fdtd.cpp:818-845; FdTdSC::fieldEvaluate, src/fdtdSC.cpp:1110-1145).

The device divides by the fixed steps dt, 2dx, 2dy, 2dz with a reciprocal and one FMA
residual step, which yields the correctly rounded quotient; the oracle divides.  Several
million quotients on random data must therefore give bit-identical float E and B, for every
tile shape of the kernel."""
import numpy as np
import pytest

from golden_state import load_oracle
from mithra_amd import abi

pytestmark = pytest.mark.gpu


def grid(nx, ny, nz, sc):
    g = abi.Grid()
    g.N0, g.N1, g.N2, g.solver, g.truncation_order, g.space_charge = nx, ny, nz, abi.SOLVER_NSFD, 2, sc
    g.dx, g.dy, g.dz = 1.2e-6, 1.3e-6, 2.0e-10
    g.xmin, g.ymin, g.zmin = -(nx - 1) / 2 * g.dx, -(ny - 1) / 2 * g.dy, -(nz - 1) / 2 * g.dz
    g.xmax, g.ymax, g.zmax = -g.xmin, -g.ymin, -g.zmin
    g.zp0, g.zp1, g.Lz = g.zmin, g.zmin + (nz - 1) * g.dz, (nz - 1) * g.dz
    g.dt = g.dz / 299792458.0
    return g


@pytest.mark.parametrize("sc", [0, 1], ids=["fdtd", "fdtdSC"])
def test_field_evaluate_bit_exact(sc):
    g = grid(40, 72, 48, sc)
    rng = np.random.default_rng(5 + sc)
    n = g.nodes
    # magnitudes spanning many binades, with exact zeros, so quotients of every kind occur
    a_n = rng.standard_normal(3 * n) * 10.0 ** rng.uniform(-14, -4, 3 * n)
    a_p = a_n + rng.standard_normal(3 * n) * 10.0 ** rng.uniform(-18, -6, 3 * n)
    a_n[rng.random(3 * n) < 0.05] = 0.0
    phi = rng.standard_normal(n) * 1e3 if sc else None
    doms = [abi.Domain(g, abi.Coeffs(), lib=load_oracle(), prefix="oracle_")]
    for v in range(abi.TILE_VARIANTS):
        d = abi.Domain(g, abi.Coeffs())
        d.set_tiles(0, v)
        doms.append(d)
    out = []
    for d in doms:
        d.set_potentials(0, a_n, phi)
        d.set_potentials(2, a_p, phi)
        d.field_evaluate()
        out.append(d.get_fields())
    e_o, b_o = out[0]
    assert np.count_nonzero(e_o) > 0.9 * e_o.size
    for v, (e, b) in enumerate(out[1:]):
        assert np.array_equal(e, e_o), f"E, tile variant {v}"
        assert np.array_equal(b, b_o), f"B, tile variant {v}"


@pytest.mark.parametrize("ny", [65, 70, 130])
def test_field_update_bit_exact_every_shape(ny):
    """Interior update, Mur faces/edges/corners and E/B of one unseeded field update against the
    oracle, bit for bit, for every tile shape (FdTd::fieldUpdate, src/fdtd.cpp:270-303 and
    377-725).  Without a TF/SF seed the LDS-staged shapes update the Mur y faces inside the
    interior kernel; N1 = 65 puts node N1-1 in a 64-node tile column of its own, so that face
    stays with the boundary pass, and 130 spreads the mesh over three tile columns."""
    g = grid(20, ny, 40, 0)
    rng = np.random.default_rng(ny)
    c = abi.Coeffs()
    for name, n in (("a", 6), ("bB", 5), ("cB", 5), ("dB", 5), ("eE", 5), ("fE", 5), ("gE", 5), ("hC", 17)):
        getattr(c, name)[:] = rng.uniform(-1.0, 1.0, n)
    c.hC[0] = 2.5  # the corner update divides by hC[0]
    c.alpha, c.beta = 1.1, 0.07
    n = g.nodes
    levels = [rng.standard_normal(3 * n) * 1e-6 for _ in range(3)]
    doms = [abi.Domain(g, c, lib=load_oracle(), prefix="oracle_")]
    for v in range(abi.TILE_VARIANTS):
        d = abi.Domain(g, c)
        d.set_tiles(v, v)
        doms.append(d)
    out = []
    for d in doms:
        for lvl, a in enumerate(levels):
            d.set_potentials(lvl, a)
        d.field_update(0.0, None)
        out.append((d.get_potentials(2)[0], *d.get_fields()))
    ref = out[0]
    for v, got in enumerate(out[1:]):
        for name, a, b in zip(("A", "E", "B"), got, ref):
            assert np.array_equal(a, b), f"{name}, tile variant {v} {abi.tile_shape(v)}"
