"""The surrogate's BF16-MFMA kernels against the CPU oracle (oracle/mlp_ref.py).

* GEMM: every operand orientation the network uses, ragged sizes; the products of
  bf16-rounded inputs are exact, so only the FP32 accumulation order differs from float64.
* forward and gradients of a 4-layer, 512-wide member: against the oracle that rounds where the
  device rounds, and against the plain network (relative error <= 1e-2, the cost of BF16
  weights and activations).  The first comparison is not exact: the device's FP32 sums over 512
  terms differ from float64 by ~1e-6, which moves a few activations across a BF16 rounding
  boundary (a 2^-8 step each); over three hidden layers that leaves ~3e-4 in the outputs.
* training: Adam steps of the whole ensemble reduce the loss on a fixed regression task."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp_ref  # noqa: E402
from golden_state import rel_l2  # noqa: E402
from mithra_amd import abi  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,N,K,ta,tb", [
    (64, 64, 32, False, False), (70, 130, 45, False, False), (264, 512, 97, True, False),
    (97, 264, 512, False, True), (512, 256, 300, True, False), (1, 3, 1, False, False)])
def test_gemm_matches_float64(M, N, K, ta, tb):
    rng = np.random.default_rng(M * 7 + N)
    A = rng.normal(0.0, 1.0, (K, M) if ta else (M, K)).astype(np.float32)
    B = rng.normal(0.0, 1.0, (N, K) if tb else (K, N)).astype(np.float32)
    C, ms = abi.gemm_bf16(A, B, ta, tb)
    Ar, Br = mlp_ref.bf16(A), mlp_ref.bf16(B)
    ref = (Ar.T if ta else Ar) @ (Br.T if tb else Br)
    print(f"gemm {M}x{N}x{K} ta={ta} tb={tb}: rel {rel_l2(C, ref):.2e}, {ms:.3f} ms", flush=True)
    assert np.max(np.abs(C - ref)) <= 1e-5 * (np.max(np.abs(ref)) + 1e-30) * np.sqrt(K)


DIMS = [264, 512, 512, 512, 256]


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0.0, 1.0, (n, DIMS[0])).astype(np.float32)
    t = rng.normal(0.0, 1.0, (n, DIMS[-1])).astype(np.float32)
    return x, t


def test_forward_matches_oracle():
    mlp = abi.MLP(DIMS, members=2, seed=3)
    x, _ = _batch(96, 1)
    y = mlp.forward(x)
    p = mlp.params()
    for e in range(2):
        pe = p[e * mlp.n_params:(e + 1) * mlp.n_params]
        ref_bf, _ = mlp_ref.forward(pe, DIMS, x, bf16_products=True)
        ref_32, _ = mlp_ref.forward(pe, DIMS, x, bf16_products=False)
        print(f"member {e}: vs bf16 oracle {rel_l2(y[e], ref_bf):.2e}, vs fp oracle {rel_l2(y[e], ref_32):.2e}",
              flush=True)
        assert rel_l2(y[e], ref_bf) <= 1e-3
        assert rel_l2(y[e], ref_32) <= 1e-2
    assert rel_l2(y[0], y[1]) > 0.1  # the members differ


def test_gradients_match_oracle():
    mlp = abi.MLP(DIMS, members=1, seed=4)
    x, t = _batch(128, 2)
    g, loss = mlp.gradients(0, x, t)
    p = mlp.params(0)
    g_bf, l_bf = mlp_ref.gradients(p, DIMS, x, t, bf16_products=True)
    g_32, l_32 = mlp_ref.gradients(p, DIMS, x, t, bf16_products=False)
    assert abs(loss - l_bf) <= 1e-3 * l_bf
    o = 0
    for a, b in zip(DIMS[:-1], DIMS[1:]):
        for name, n in (("W", a * b), ("b", b)):
            sl = slice(o, o + n)
            print(f"{name} {a}x{b}: vs bf16 oracle {rel_l2(g[sl], g_bf[sl]):.2e}, "
                  f"vs fp oracle {rel_l2(g[sl], g_32[sl]):.2e}", flush=True)
            assert rel_l2(g[sl], g_bf[sl]) <= 5e-3
            assert rel_l2(g[sl], g_32[sl]) <= 1e-2
            o += n


def test_training_reduces_loss():
    rng = np.random.default_rng(9)
    mlp = abi.MLP(DIMS, members=5, seed=5)
    x = rng.normal(0.0, 1.0, (256, DIMS[0])).astype(np.float32)
    w = rng.normal(0.0, 0.1, (DIMS[0], DIMS[-1]))
    t = np.tanh(x @ w).astype(np.float32)
    first = mlp.train_step(x, t, 1e-3)
    for _ in range(60):
        last = mlp.train_step(x, t, 1e-3)
    print("loss", first, "->", last, flush=True)
    assert np.all(last < 0.2 * first)


def test_surrogate_workflow_small():
    """The whole exercise at a small size: batched solves produce the data, the ensemble trains
    on the GPU, predicts, the spread gate decides, and every error is measured against FP64
    re-solves of the held-out points."""
    import tempfile
    from mithra_amd import surrogate
    with tempfile.TemporaryDirectory() as w:
        out = surrogate.run_workflow(w, n_train=256, n_hold=32, epochs=20, n_single=4)
    print(out, flush=True)
    assert out["fp32_bulk_p_err_vs_fp64"]["max"] < 1e-4
    assert out["final_loss"] < 1.0
    assert 0.0 <= out["points_per_hour"]["accepted_at_limit"] <= 1.0
    assert out["points_per_hour"]["batched"] > out["points_per_hour"]["single_instance_gpu"]
