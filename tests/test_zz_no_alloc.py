"""The bunching factor through the entry point that allocates no device memory
(mithra_dev_bunching_noalloc) against the allocating one (mithra_dev_bunching), through the C-ABI.

Both run the same kernel on the same fixed grid and add the partial sums in the same order, so
they must agree bit for bit.  Two calls in a row share the static scratch; the second uses fewer
particles, so stale partials of the first call would show as a difference.  The oracle's value is
checked too (its own closed-form pins are in test_oracle_invariants)."""
import numpy as np
import pytest

from golden_state import Case, march

pytestmark = pytest.mark.gpu


def test_noalloc_bunching_matches_allocating_entry():
    case = Case("nsfd_o2")
    gpu = case.open("gpu")
    ora = case.open("oracle")
    march(case, gpu, 3)
    march(case, ora, 3)
    k = 2.0 * np.pi / (10.0 * case.grid.dz)
    full = gpu.get_particles()
    print(f"no-alloc bunching: {full.shape[0]} particles, k = {k:.6e}", flush=True)
    b_alloc = gpu.bunching(k)
    b_static = gpu.bunching_noalloc(k)
    b_oracle = ora.bunching(k)
    print(f"  alloc {b_alloc!r} static {b_static!r} oracle {b_oracle!r}", flush=True)
    assert b_static == b_alloc
    assert abs(b_static - b_oracle) <= 1e-9 * max(b_oracle, 1e-300)

    # second call, smaller input: a third of the particles
    part = full[: full.shape[0] // 3].copy()
    gpu.set_particles(part)
    ora.set_particles(part)
    print(f"no-alloc bunching: {part.shape[0]} particles", flush=True)
    b_alloc2 = gpu.bunching(k)
    b_static2 = gpu.bunching_noalloc(k)
    b_oracle2 = ora.bunching(k)
    print(f"  alloc {b_alloc2!r} static {b_static2!r} oracle {b_oracle2!r}", flush=True)
    assert b_static2 == b_alloc2
    assert abs(b_static2 - b_oracle2) <= 1e-9 * max(b_oracle2, 1e-300)
    assert b_static2 != b_static


def test_noalloc_bunching_repeats_and_second_domain():
    """Repeated calls give the same value, and a second domain on the same GPU shares the array."""
    case = Case("fd_o1_array")
    a = case.open("gpu")
    b = Case("nsfd_o2").open("gpu")
    k = 2.0 * np.pi / (7.0 * case.grid.dz)
    va = [a.bunching_noalloc(k) for _ in range(3)]
    vb = b.bunching_noalloc(k)
    print(f"no-alloc bunching repeats {va}, second domain {vb}", flush=True)
    assert va[0] == va[1] == va[2] == a.bunching(k)
    assert vb == b.bunching(k)
