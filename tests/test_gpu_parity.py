"""Parity of the HIP path (libmithra_gpu.so, through its C-ABI) against the CPU oracle and the
reference's own fixtures, FP64 drop-in mode.

Bar: fields, currents and particles after 100 steps within 1e-10
relative L2 of the oracle, P(z) within 1e-6 relative of the reference's power file.  The field
update itself is checked bit-exact after one step (same operation order, no contraction); the
push uses the device's libm and the deposition FP64 atomics, whose summation order differs from
the reference's, so those are held to the tolerance rather than to bit equality.
"""
import numpy as np
import pytest

from golden_state import GPU_CASES, Case, march, rel_l2
from mithra_amd import abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=GPU_CASES)
def case(request):
    print(f"[gpu parity] case {request.param}", flush=True)
    return Case(request.param)


def test_one_step(case):
    gpu = case.open("gpu")
    ora = case.open("oracle")
    march(case, gpu, 1)
    march(case, ora, 1)
    a_g, f_g = gpu.get_potentials(0)
    a_o, f_o = ora.get_potentials(0)
    e_g, b_g = gpu.get_fields()
    seeded = abs(case.seed.amplitude) > 1e-50
    print(f"A {rel_l2(a_g, a_o):.3e}  E {rel_l2(e_g, case.arr('post0_en')):.3e}  "
          f"B {rel_l2(b_g, case.arr('post0_bn')):.3e}", flush=True)
    if not seeded:
        # interior stencil, Mur faces/edges/corners and E/B: no transcendental on the path,
        # same operation order, no contraction -> bit-exact with the reference
        assert np.array_equal(a_g, a_o)
        assert np.array_equal(a_g, case.arr("post0_anp1"))
        assert np.array_equal(e_g, case.arr("post0_en"))
        assert np.array_equal(b_g, case.arr("post0_bn"))
        if case.grid.space_charge:
            assert np.array_equal(f_g, f_o)
    else:
        # the TF/SF source evaluates the analytic seed with the device's cos/exp/atan
        assert rel_l2(a_g, a_o) < 1e-13
        assert rel_l2(e_g, case.arr("post0_en")) < 1e-6      # float E, 1-ulp flips
        assert rel_l2(b_g, case.arr("post0_bn")) < 1e-6
        if case.grid.space_charge:
            assert rel_l2(f_g, f_o) < 1e-13
    j_g, r_g = gpu.get_potentials(2)
    j_o, r_o = ora.get_potentials(2)
    p_g, p_o = gpu.get_particles(), ora.get_particles()
    print(f"J {rel_l2(j_g, j_o):.3e}  particles {rel_l2(p_g, p_o):.3e}", flush=True)
    assert rel_l2(j_g, j_o) < 1e-13
    assert rel_l2(p_g, p_o) < 1e-13
    assert np.array_equal(p_g[:, 10], p_o[:, 10])
    if case.grid.space_charge:
        assert rel_l2(r_g, r_o) < 1e-14


def test_hundred_steps(case):
    gpu = case.open("gpu")
    ora = case.open("oracle")
    march(case, gpu, case.n_steps)
    march(case, ora, case.n_steps)
    for level in (0, 1, 2):
        a_g, f_g = gpu.get_potentials(level)
        a_o, f_o = ora.get_potentials(level)
        print(f"level {level}: A vs oracle {rel_l2(a_g, a_o):.3e}", flush=True)
        assert rel_l2(a_g, a_o) < 1e-10, f"A level {level}"
        assert rel_l2(a_g, case.arr(("preN_an", "preN_anm1", "preN_jn")[level])) < 1e-10
        if case.grid.space_charge:
            assert rel_l2(f_g, f_o) < 1e-10, f"phi level {level}"
    p_g, p_o = gpu.get_particles(), ora.get_particles()
    assert rel_l2(p_g[:, 1:10], p_o[:, 1:10]) < 1e-10
    assert np.array_equal(p_g[:, 10], p_o[:, 10])
    P_g = gpu.power_fetch(0, case.n_steps, case.power_N, case.power_Nl)[:, 0, 0]
    P_ref = case.power_file()[: case.n_steps, 1]
    scale = np.max(np.abs(P_ref))
    print(f"particles {rel_l2(p_g[:, 1:10], p_o[:, 1:10]):.3e}  "
          f"P max rel {np.max(np.abs(P_g - P_ref)) / scale:.3e}", flush=True)
    assert np.max(np.abs(P_g - P_ref)) <= 1e-6 * scale
    nz = np.abs(P_ref) > 1e-3 * scale
    assert np.all(np.abs(P_g[nz] - P_ref[nz]) <= 1e-6 * np.abs(P_ref[nz]))


def test_fused_sub_steps_match_single_steps(case):
    """mithra_dev_bunch_advance (n sub-steps in one launch) == n mithra_dev_bunch_update calls."""
    a = case.open("gpu")
    b = case.open("gpu")
    for dom in (a, b):
        dom.field_update(case.time0, case.seed)
    a.bunch_begin()
    tb = case.time_bunch0
    for _ in range(case.n_sub):
        a.bunch_update(case.bunch, case.undulators, tb)
        tb += case.dtb
    b.bunch_advance(case.bunch, case.undulators, case.time_bunch0, case.dtb, case.n_sub)
    assert np.array_equal(a.get_particles(), b.get_particles())


def test_bunching_factor(case):
    """Bunching factor of the beam after the march, GPU vs oracle, at a wavelength of ten
    cells."""
    gpu = case.open("gpu")
    ora = case.open("oracle")
    march(case, gpu, 20)
    march(case, ora, 20)
    k = 2.0 * np.pi / (10.0 * case.grid.dz)
    bg, bo = gpu.bunching(k), ora.bunching(k)
    print(f"bunching {bg:.6e} vs {bo:.6e}", flush=True)
    assert abs(bg - bo) <= 1e-9 * max(bo, 1e-300)


def test_cell_sort_changes_no_particle_result(case):
    """The particle arrays reordered by cell row after every deposit versus never.  After one step
    (the push precedes the first deposit) every particle record, returned in load order, is
    identical and J agrees to the rounding of the summation order; four steps later the whole
    state still agrees to that rounding."""
    sorted_dom = case.open("gpu")
    sorted_dom.set_reorder_interval(1)
    plain = case.open("gpu")
    plain.set_reorder_interval(0)
    ends = [march(case, dom, 1) for dom in (sorted_dom, plain)]
    assert np.array_equal(sorted_dom.get_particles(), plain.get_particles())
    a_s, _ = sorted_dom.get_potentials(2)
    a_p, _ = plain.get_potentials(2)
    assert rel_l2(a_s, a_p) < 1e-13
    for dom, (t, tb) in zip((sorted_dom, plain), ends):
        march(case, dom, 4, start_step=1, time=t, time_bunch=tb)
    assert rel_l2(sorted_dom.get_particles(), plain.get_particles()) < 1e-13
    assert rel_l2(sorted_dom.get_potentials(0)[0], plain.get_potentials(0)[0]) < 1e-12


def test_tile_shapes_give_identical_fields(case):
    """Every tile shape of the interior update and the E/B evaluation gives bit-identical
    potentials and fields for the field update of the reference's first step (the field
    kernels are deterministic; the autotuner may pick any shape)."""
    ref = None
    for v in range(abi.TILE_VARIANTS):
        dom = case.open("gpu")
        dom.set_tiles(v, v)
        dom.field_update(case.time0, case.seed)
        out = (dom.get_potentials(2)[0], *dom.get_fields())
        if ref is None:
            ref = out
            assert rel_l2(out[0], case.arr("post0_anp1")) < 1e-13
        else:
            for a, b in zip(out, ref):
                assert np.array_equal(a, b), f"tile variant {v} {abi.tile_shape(v)}"


def test_autotune_keeps_state(case):
    dom = case.open("gpu")
    march(case, dom, 2)
    before = dom.get_potentials(2)[0], dom.get_potentials(0)[0]
    (ci, ce), ms = dom.autotune(reps=1)
    assert 0 <= ci < abi.TILE_VARIANTS and 0 <= ce < abi.TILE_VARIANTS
    assert np.all(ms > 0)
    after = dom.get_potentials(2)[0], dom.get_potentials(0)[0]
    assert all(np.array_equal(a, b) for a, b in zip(before, after))


def test_lazy_fields_change_no_result(case):
    """E/B evaluated only on the planes the bunch and the power sampler read gives the same
    march as evaluating every node (to the rounding of the deposit's summation order)."""
    lazy = case.open("gpu")
    lazy.set_lazy_fields(1)
    full = case.open("gpu")
    for dom in (lazy, full):
        march(case, dom, 8)
    assert rel_l2(lazy.get_particles(), full.get_particles()) < 1e-13
    assert rel_l2(lazy.get_potentials(0)[0], full.get_potentials(0)[0]) < 1e-12
    p_l = lazy.power_fetch(0, 8, case.power_N, case.power_Nl)
    p_f = full.power_fetch(0, 8, case.power_N, case.power_Nl)
    assert rel_l2(p_l, p_f) < 1e-12


def test_super_gaussian_injection_runs_in_call_order():
    """The super-gaussian seed's TF/SF injection: each call injects the running vector of the
    calls before it in the reference's loop order (Seed::fields adds into the output vector that
    FdTd::fieldUpdate zeroes once per step, src/classes.cpp:814-851, src/fdtd.cpp:239).  Against
    the oracle, which runs the loop sequentially, after one and after ten steps."""
    case = Case("nsfd_o2_supergauss")
    assert case.seed.seed_type == 3 and abs(case.seed.amplitude) > 1e-50
    gpu, ora = case.open("gpu"), case.open("oracle")
    for n in (1, 9):
        march(case, gpu, n)
        march(case, ora, n)
        a_g, _ = gpu.get_potentials(0)
        a_o, _ = ora.get_potentials(0)
        print(f"after {n} more step(s): A vs oracle {rel_l2(a_g, a_o):.3e}", flush=True)
        assert rel_l2(a_g, a_o) < 1e-12
