"""FP32-bulk mode: potentials, J and E/B in FP32, particle positions, momenta and the push in
FP64.  Its error against the FP64 oracle is measured with the reference's own P(z) metric,
integral |P - P0| dz / integral P0 dz (doc/MITHRA_EXAMPLES/MITHRA_EXAMPLES.tex:182-185), and
must stay far inside the 2 % budget of the scan path."""
import numpy as np
import pytest

from golden_state import GPU_CASES, Case, march, rel_l2
from mithra_amd import abi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", GPU_CASES)
def test_fp32_bulk_error(name):
    print(f"[fp32] {name}", flush=True)
    case = Case(name)
    gpu = case.open("gpu", precision=abi.PRECISION_FP32_BULK)
    ora = case.open("oracle")
    march(case, gpu, case.n_steps)
    march(case, ora, case.n_steps)
    P32 = gpu.power_fetch(0, case.n_steps, case.power_N, case.power_Nl)[:, 0, 0]
    P64 = ora.power_fetch(0, case.n_steps, case.power_N, case.power_Nl)[:, 0, 0]
    err = np.sum(np.abs(P32 - P64)) / np.sum(np.abs(P64))
    a32, _ = gpu.get_potentials(0)
    a64, _ = ora.get_potentials(0)
    p32, p64 = gpu.get_particles(), ora.get_particles()
    print(f"P(z) L1 error {err:.3e}, A rel L2 {rel_l2(a32, a64):.3e}, "
          f"particles rel L2 {rel_l2(p32[:, 1:10], p64[:, 1:10]):.3e}", flush=True)
    assert err < 1e-3
    assert rel_l2(a32, a64) < 1e-4
