"""The pipelined step (mithra_dev_step_pipelined) against the plain step (mithra_dev_step).

The pipelined step reorders Solver::solve()'s loop body (src/solver.cpp:1300-1399): the E/B
evaluation of step n is fused with the next interior update without its source term, X, and the
deposit writes X + a4 J instead of J (and X + a5 rho for phi under space charge, whose E takes the
gradient of phi^n in the same kernel).  The interior update sums its terms left to right with
the source last (src/database.cpp:44-134), so in FP64 the split changes no bit.

Bit-exact checks need a deterministic deposit.  The row-tiled deposit adds into LDS with
atomics, so a node that receives three or more contributions may round differently from run to
run.  The exact tests therefore march a sparse subset of the reference's particles, at least
four z-cells apart, so that no node receives more than two contributions (and a + b = b + a).
The full bunch is compared to the rounding of the summation order, as the other multi-step tests
do.  FP32-bulk domains, and domains without particles, run the plain step whatever `more`
asks, which is checked too."""
import numpy as np
import pytest

from golden_state import GPU_CASES, Case, rel_l2
from mithra_amd import abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=GPU_CASES)
def case(request):
    print(f"[gpu pipelined] case {request.param}", flush=True)
    return Case(request.param)


def sparse_particles(case, cells=4):
    """Particles of the reference's bunch at least `cells` z-cells apart."""
    q = case.arr("pre0_particles").reshape(-1, 11)
    q = q[np.argsort(q[:, 3], kind="stable")]
    keep, last = [], -np.inf
    for r in range(q.shape[0]):
        if q[r, 3] >= last + cells * case.grid.dz:
            keep.append(r)
            last = q[r, 3]
    return q[keep]


def run(case, dom, n, pipelined, timed=False):
    """n steps from the reference's state before step 0; pipelined: more = 1 on all but the last."""
    time, tb = case.time0, case.time_bunch0
    ms = []
    for step in range(n):
        if pipelined:
            r = dom.step_pipelined(time, case.seed, case.bunch, case.undulators, tb, case.dtb, case.n_sub,
                                   step, step, more=step < n - 1, timed=timed)
            ms.append(r)
        else:
            dom.step(time, case.seed, case.bunch, case.undulators, tb, case.dtb, case.n_sub, step, step)
        for _ in range(case.n_sub):
            tb += case.dtb
        time += case.dt
    return ms


def state(case, dom, n):
    out = {f"level{lv}": dom.get_potentials(lv) for lv in (0, 1, 2)}
    out["eb"] = dom.get_fields()
    out["particles"] = dom.get_particles()
    out["power"] = dom.power_fetch(0, n, case.power_N, case.power_Nl)
    return out


def flatten(s):
    for k, v in s.items():
        if isinstance(v, tuple):
            for i, a in enumerate(v):
                if a is not None:
                    yield f"{k}[{i}]", np.asarray(a)
        else:
            yield k, np.asarray(v)


def open_sparse(case, precision=abi.PRECISION_FP64):
    dom = case.open("gpu", precision=precision)
    dom.set_particles(sparse_particles(case))
    return dom


@pytest.mark.parametrize("steps", [2, 7])
def test_pipelined_bit_exact_sparse(case, steps):
    """FP64: a pipelined run of `steps` steps leaves every level, E/B, particle and P(z) sample
    bit-identical to the plain steps (deterministic deposit, see the module docstring)."""
    plain, pipe = open_sparse(case), open_sparse(case)
    assert plain.f["particle_count"](plain.h) >= 3
    run(case, plain, steps, False)
    run(case, pipe, steps, True)
    assert not pipe.pipeline_pending()
    a, b = state(case, plain, steps), state(case, pipe, steps)
    for (k, x), (_, y) in zip(flatten(a), flatten(b)):
        assert np.array_equal(x, y), f"{case.name}: {k} differs after {steps} steps ({rel_l2(y, x):.3e})"


def test_pipelined_full_bunch(case):
    """The reference's whole bunch over its fixture's step count: the pipelined run agrees with the
    plain steps to the rounding of the deposit's summation order, held to the 100-step bar of the
    parity tests (1e-10 relative L2; two plain runs differ by the same order)."""
    plain, pipe = case.open("gpu"), case.open("gpu")
    n = case.n_steps
    run(case, plain, n, False)
    run(case, pipe, n, True)
    a, b = state(case, plain, n), state(case, pipe, n)
    for (k, x), (_, y) in zip(flatten(a), flatten(b)):
        d = rel_l2(y, x)
        print(f"{case.name} {k}: {d:.3e}", flush=True)
        assert d < 1e-10, k
    for key in ("level0", "level1", "level2"):   # and the oracle-checked plain path's own bar
        assert rel_l2(b[key][0], case.arr(("preN_an", "preN_anm1", "preN_jn")[int(key[-1])])) < 1e-10


def test_pending_state_guards_the_potentials(case):
    """Between a more = 1 step and the run's last step the potentials are refused, the particles
    and E/B stay readable, and a plain step ends the run like a more = 0 step."""
    dom = open_sparse(case)
    run(case, dom, 1, True)  # one step, more = 0: nothing pending
    assert not dom.pipeline_pending()
    ref = open_sparse(case)
    time, tb = case.time0, case.time_bunch0
    dom2 = open_sparse(case)
    dom2.step_pipelined(time, case.seed, case.bunch, case.undulators, tb, case.dtb, case.n_sub, 0, 0, more=True)
    assert dom2.pipeline_pending()
    with pytest.raises(RuntimeError):
        dom2.get_potentials(0)
    with pytest.raises(RuntimeError):
        dom2.field_update(time, case.seed)
    with pytest.raises(RuntimeError):
        dom2.current_update()
    dom2.get_particles()
    dom2.get_fields()
    for _ in range(case.n_sub):
        tb += case.dtb
    time += case.dt
    dom2.step(time, case.seed, case.bunch, case.undulators, tb, case.dtb, case.n_sub, 1, 1)
    assert not dom2.pipeline_pending()
    run(case, ref, 2, False)
    for (k, x), (_, y) in zip(flatten(state(case, ref, 2)), flatten(state(case, dom2, 2))):
        assert np.array_equal(x, y), k


def test_fused_tile_shapes_identical(case):
    """Every tile shape of the fused kernel gives the same bits, with and without space charge;
    autotune_fused keeps the state."""
    ref = None
    for v in range(abi.FUSED_TILE_VARIANTS):
        dom = open_sparse(case)
        dom.set_fused_tile(v)
        run(case, dom, 3, True)
        s = list(flatten(state(case, dom, 3)))
        if ref is None:
            ref = s
        else:
            for (k, x), (_, y) in zip(ref, s):
                assert np.array_equal(x, y), f"fused shape {v}: {k}"
    dom = open_sparse(case)
    run(case, dom, 2, False)
    before = list(flatten(state(case, dom, 2)))
    chosen, ms = dom.autotune_fused(reps=1)
    assert 0 <= chosen < abi.FUSED_TILE_VARIANTS and np.all(ms > 0)
    for (k, x), (_, y) in zip(before, flatten(state(case, dom, 2))):
        assert np.array_equal(x, y), k


def test_pipelined_request_on_fp32_bulk_runs_plain(case):
    """FP32-bulk domains take the plain step when more = 1 is asked: nothing is left pending and
    the march is the plain FP32 march."""
    plain = open_sparse(case, abi.PRECISION_FP32_BULK)
    pipe = open_sparse(case, abi.PRECISION_FP32_BULK)
    time, tb = case.time0, case.time_bunch0
    pipe.step_pipelined(time, case.seed, case.bunch, case.undulators, tb, case.dtb, case.n_sub, 0, 0, more=True)
    assert not pipe.pipeline_pending()
    plain.step(time, case.seed, case.bunch, case.undulators, tb, case.dtb, case.n_sub, 0, 0)
    for (k, x), (_, y) in zip(flatten(state(case, plain, 1)), flatten(state(case, pipe, 1))):
        assert np.array_equal(x, y), k


def test_timed_phases(case):
    """step_pipelined(timed) reports the phases; after a pipelined step the interior phase is
    empty (the previous step's kernels marched A^{n+1})."""
    dom = case.open("gpu")
    ms = run(case, dom, 3, True, timed=True)
    assert all(len(m) == len(abi.PHASES) for m in ms)
    assert all(x >= 0 for m in ms for x in m)
    assert ms[1][0] < ms[0][0]


def test_no_particles_runs_plain(case):
    """A domain without particles cannot hand its J to X (there is no deposit to add it), so a
    more = 1 request runs the plain step: nothing is left pending and the march equals the plain
    one bit for bit."""
    plain, pipe = case.open("gpu"), case.open("gpu")
    for dom in (plain, pipe):
        dom.set_particles(np.zeros((0, 11)))
    time, tb = case.time0, case.time_bunch0
    for step in range(3):
        pipe.step_pipelined(time, case.seed, case.bunch, case.undulators, tb, case.dtb, case.n_sub, step, step,
                            more=True)
        assert not pipe.pipeline_pending()
        plain.step(time, case.seed, case.bunch, case.undulators, tb, case.dtb, case.n_sub, step, step)
        for _ in range(case.n_sub):
            tb += case.dtb
        time += case.dt
    for (k, x), (_, y) in zip(flatten(state(case, plain, 3)), flatten(state(case, pipe, 3))):
        assert np.array_equal(x, y), k
