"""Fixed-order current deposition (mithra_dev_set_deterministic).

The default deposit sums each node in LDS with atomic adds, so two FP64 runs from one state agree
to rounding only (FdTd::currentUpdate, src/fdtd.cpp:38-185, sums in particle-list order).  In
fixed order every node sums its eight cells' listed halves in an order set by the particles:

1. on random particles (dense cells, boundary cells, halves outside the deposit region, moves of
   up to a cell; with and without space charge, ragged and wide rows) the fixed-order J and rho
   are the oracle's to the rounding of the summation order, in FP64 and FP32-bulk, written once
   and added a second time;
2. the same deposit from two domains holding the same particles is identical bit for bit, and
   so is a 40-step march of golden jobs run twice, in every potential level, E/B, particle
   record and P(z) sample; that march stays within the golden march's bars of the oracle and
   the reference's P(z) file;
3. a pipelined request runs plain steps while the mode is on, the mode cannot change mid-way
   through a pipelined run, and a batch refuses a fixed-order domain;
4. at the production size two fixed-order marches from one set-up agree in every bit."""
import numpy as np
import pytest

from golden_state import Case, march, rel_l2
from mithra_amd import abi
from test_gpu_deposit import CASES, deposit, grid, particles

pytestmark = pytest.mark.gpu


def deposit_fixed(g, rec, precision=abi.PRECISION_FP64, twice=False):
    dom = abi.Domain(g, abi.Coeffs(), precision=precision)
    dom.set_deterministic(True)
    dom.set_particles(rec)
    dom.current_reset()
    dom.current_update()
    if twice:
        dom.current_update()
    return dom.get_potentials(2)


@pytest.mark.parametrize("name,shape,sc,n", CASES, ids=[c[0] for c in CASES])
def test_fixed_order_deposit_matches_oracle(name, shape, sc, n):
    g = grid(*shape, sc=sc)
    rec = particles(g, n, seed=len(name))
    print(f"fixed-order deposit case {name} {shape} sc={sc} n={n}", flush=True)
    jg, rg = deposit_fixed(g, rec)
    jo, ro = deposit(g, rec, "oracle")
    print(f"  J vs oracle {rel_l2(jg, jo):.2e}" + (f", rho {rel_l2(rg, ro):.2e}" if sc else ""), flush=True)
    assert np.count_nonzero(jo) > 0
    assert rel_l2(jg, jo) < 1e-13
    assert np.max(np.abs(jg - jo)) <= 1e-12 * np.max(np.abs(jo))
    if sc:
        assert rel_l2(rg, ro) < 1e-13
    # another domain, the same particles: the same bits
    jg2, rg2 = deposit_fixed(g, rec)
    assert np.array_equal(jg, jg2)
    if sc:
        assert np.array_equal(rg, rg2)


def test_fixed_order_fp32_bulk_and_accumulate():
    g = grid(11, 300, 8, sc=1)
    rec = particles(g, 30000, seed=7)
    jg, rg = deposit_fixed(g, rec, precision=abi.PRECISION_FP32_BULK)
    jo, ro = deposit(g, rec, "oracle")
    assert rel_l2(jg, jo) < 1e-6 and rel_l2(rg, ro) < 1e-6
    once, r1 = deposit_fixed(g, rec)
    twice, r2 = deposit_fixed(g, rec, twice=True)
    assert rel_l2(twice, 2.0 * once) < 1e-14 and rel_l2(r2, 2.0 * r1) < 1e-14


def snapshot(case, dom, steps):
    P = dom.power_fetch(0, steps, case.power_N, case.power_Nl)
    pot = [dom.get_potentials(level) for level in (0, 1, 2)]
    return pot, dom.get_fields(), dom.get_particles(), P


@pytest.mark.parametrize("name", ["nsfd_o2", "nsfd_o2_sc"])
def test_fixed_order_march_repeats_every_bit(name):
    case = Case(name)
    steps = 40
    runs = []
    for _ in range(2):
        dom = case.open("gpu")
        dom.set_deterministic(True)
        march(case, dom, steps)
        runs.append(snapshot(case, dom, steps))
        dom.close()
    (pa, fa, qa, Pa), (pb, fb, qb, Pb) = runs
    for (a, f), (b, h) in zip(pa, pb):
        assert np.array_equal(a, b)
        if case.grid.space_charge:
            assert np.array_equal(f, h)
    assert all(np.array_equal(x, y) for x, y in zip(fa, fb))
    assert np.array_equal(qa, qb) and np.array_equal(Pa, Pb)
    # the same march against the oracle and the reference's P(z) file
    ora = case.open("oracle")
    march(case, ora, steps)
    for level in (0, 1, 2):
        a_o, _ = ora.get_potentials(level)
        assert rel_l2(pa[level][0], a_o) < 1e-10
    q_o = ora.get_particles()
    assert rel_l2(qa[:, 1:10], q_o[:, 1:10]) < 1e-10
    P_ref = case.power_file()[:steps, 1]
    P_g = Pa[:, 0, 0]
    scale = np.max(np.abs(P_ref))
    print(f"{name}: 2 fixed-order runs identical; A vs oracle {rel_l2(pa[0][0], ora.get_potentials(0)[0]):.2e}, "
          f"P max {np.max(np.abs(P_g - P_ref)) / scale:.2e} of peak", flush=True)
    assert np.max(np.abs(P_g - P_ref)) <= 1e-6 * scale


def test_fixed_order_runs_plain_steps_and_guards_its_switch():
    case = Case("nsfd_o2")
    dom = case.open("gpu")
    t = case.time0
    tb = case.time_bunch0
    dom.step_pipelined(t, case.seed, case.bunch, case.undulators, tb, case.dtb, case.n_sub, 0, 0, more=1)
    assert dom.pipeline_pending() == 1
    with pytest.raises(abi.MithraError):
        dom.set_deterministic(True)  # mid-way through a pipelined run
    dom.step_pipelined(t + case.dt, case.seed, case.bunch, case.undulators, tb + case.n_sub * case.dtb, case.dtb,
                       case.n_sub, 1, 1, more=0)
    assert dom.pipeline_pending() == 0
    dom.set_deterministic(True)
    dom.step_pipelined(t + 2 * case.dt, case.seed, case.bunch, case.undulators, tb + 2 * case.n_sub * case.dtb,
                       case.dtb, case.n_sub, 2, 2, more=1)
    assert dom.pipeline_pending() == 0  # a plain step
    b = abi.Batch([dom])
    with pytest.raises(abi.MithraError, match="not supported"):
        b.instance(0, t, tb, case.dtb, case.n_sub, 4, case.bunch, case.undulators, 1.0, 1, 4)
    b.close()
    dom.close()


def test_fixed_order_repeats_at_production_size():
    """256 x 256 x 2048 nodes, 16,777,216 particles (the bench's set-up): two domains from one
    set-up march 3 fixed-order steps each and agree in every bit of J, A^n and the particles."""
    import os
    import sys
    import tempfile
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    with tempfile.TemporaryDirectory() as work:
        job = bench.set_up(work)
        out = []
        for _ in range(2):
            dom = bench.open_domain(job, abi.PRECISION_FP64, rows=3)
            dom.set_deterministic(True)
            m = bench.March(job)
            for _ in range(3):
                dom.step(m.time, m.seed, m.bp, m.und, m.tb, m.t["bunch_time_step"], m.t["n_sub"], m.n, m.n)
                m.advance()
            out.append((dom.get_potentials(2)[0], dom.get_potentials(0)[0], dom.get_particles()))
            dom.close()
        job.close()
    (ja, aa, qa), (jb, ab, qb) = out
    print(f"production size, fixed order: J non-zero {np.count_nonzero(ja)}, identical "
          f"{np.array_equal(ja, jb) and np.array_equal(aa, ab) and np.array_equal(qa, qb)}", flush=True)
    assert np.count_nonzero(ja) > 0
    assert np.array_equal(ja, jb) and np.array_equal(aa, ab) and np.array_equal(qa, qb)
