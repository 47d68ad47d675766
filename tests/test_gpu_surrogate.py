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


@pytest.mark.parametrize("tile", [1, 2])
@pytest.mark.parametrize("M,N,K,ta,tb", [
    (64, 64, 32, False, False), (70, 130, 45, False, False), (264, 512, 97, True, False),
    (97, 264, 512, False, True), (512, 256, 300, True, False), (1, 3, 1, False, False),
    (300, 200, 136, True, True), (136, 264, 256, False, False), (129, 257, 33, True, True)])
def test_gemm_matches_float64(M, N, K, ta, tb, tile):
    """Both workgroup tiles, every operand orientation (16-byte staging where the strides allow
    it, element loads at ragged edges and odd strides)."""
    rng = np.random.default_rng(M * 7 + N)
    A = rng.normal(0.0, 1.0, (K, M) if ta else (M, K)).astype(np.float32)
    B = rng.normal(0.0, 1.0, (N, K) if tb else (K, N)).astype(np.float32)
    C, ms = abi.gemm_bf16(A, B, ta, tb, tile=tile)
    Ar, Br = mlp_ref.bf16(A), mlp_ref.bf16(B)
    ref = (Ar.T if ta else Ar) @ (Br.T if tb else Br)
    print(f"gemm {M}x{N}x{K} ta={ta} tb={tb} tile={tile}: rel {rel_l2(C, ref):.2e}, {ms:.3f} ms", flush=True)
    assert np.max(np.abs(C - ref)) <= 1e-5 * (np.max(np.abs(ref)) + 1e-30) * np.sqrt(K)


def test_gemm_rejects_operands_past_their_arrays():
    import ctypes as C
    lib = abi.load_gpu_library()
    abi.gemm_bf16(np.zeros((2, 2), np.float32), np.zeros((2, 2), np.float32))  # declares the argument types
    a = np.zeros(32, np.uint16)
    c = np.zeros(16, np.float32)
    u16p, fp = C.POINTER(C.c_ushort), C.POINTER(C.c_float)
    ms = C.c_float()
    # 4 x 8 times 8 x 4: B needs 32 elements, 31 given -> MITHRA_E_ARG before any device work
    rc = lib.mithra_mlp_gemm(a.ctypes.data_as(u16p), 8, 1, 32, a.ctypes.data_as(u16p), 4, 1, 31,
                             c.ctypes.data_as(fp), 4, 4, 8, 0, 1, C.byref(ms))
    assert rc == -1


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


def test_epoch_graph_matches_host_steps():
    """train_epochs with one mini-batch per epoch (the whole set, so the permutation only orders
    the rows) against the same epochs as host-driven train_step calls at the cosine schedule's
    rates: the captured graph gathers, steps the members and schedules Adam as the host does.
    The two differ only by the row order of the sums over the batch."""
    rng = np.random.default_rng(11)
    x = rng.normal(0.0, 1.0, (128, DIMS[0])).astype(np.float32)
    t = np.tanh(x @ rng.normal(0.0, 0.1, (DIMS[0], DIMS[-1]))).astype(np.float32)
    a, b = abi.MLP(DIMS, members=3, seed=6), abi.MLP(DIMS, members=3, seed=6)
    epochs, lr = 12, 2e-3
    a.set_data(x, t)
    loss = a.train_epochs(epochs, 128, lr, 1e-6, seed=1)
    host = []
    for ep in range(epochs):
        rate = max(lr * 0.5 * (1.0 + np.cos(np.pi * ep / epochs)), 1e-6)
        host.append(float(np.mean(b.train_step(x, t, np.float32(rate)))))
    pa, pb = a.params(), b.params()
    print(f"epoch losses {loss[0]:.4f} -> {loss[-1]:.4f}, host {host[0]:.4f} -> {host[-1]:.4f}; "
          f"params rel {rel_l2(pa, pb):.2e}", flush=True)
    assert np.allclose(loss, host, rtol=1e-3)
    assert rel_l2(pa, pb) <= 1e-4
    # the FP32 masters agree to ~5e-5; their BF16 copies then differ by a rounding step (2^-8
    # relative) at a few weights, which moves the outputs by a few 1e-3
    ya, yb = a.forward(x), b.forward(x)
    assert rel_l2(ya, yb) <= 1e-2


def test_epoch_training_reduces_loss():
    rng = np.random.default_rng(12)
    x = rng.normal(0.0, 1.0, (1024, DIMS[0])).astype(np.float32)
    t = np.tanh(x @ rng.normal(0.0, 0.1, (DIMS[0], DIMS[-1]))).astype(np.float32)
    mlp = abi.MLP(DIMS, members=5, seed=7)
    with pytest.raises(abi.MithraError):
        mlp.train_epochs(1, 64, 1e-3)  # no data set yet
    mlp.set_data(x, t)
    loss = mlp.train_epochs(40, 128, 1e-3, seed=2)
    print("epoch loss", loss[0], "->", loss[-1], flush=True)
    assert np.all(np.isfinite(loss)) and loss[-1] < 0.1 * loss[0]


def test_surrogate_workflow_small():
    """The whole exercise with a barely trained ensemble (256 points, 20 epochs): batched solves
    produce the data, the ensemble trains on the GPU, predicts, and the spread gate sends the
    points it is unsure of to the solver; every returned P(z) the gate refused comes from the
    solver and is within 2 % of the FP64 re-solve."""
    import tempfile
    from mithra_amd import surrogate
    with tempfile.TemporaryDirectory() as w:
        out = surrogate.run_workflow(w, n_train=256, n_hold=32, epochs=20, n_single=4, n_outside=0)
    print(out, flush=True)
    r = out["points_per_hour"]
    assert out["fp32_bulk_p_err_vs_fp64"]["max"] < 1e-4
    assert out["final_loss"] < 1.0
    assert 0.0 <= r["accepted_at_limit"] <= 1.0
    assert r["returned_within_2pct"] >= 1.0 - r["accepted_at_limit"]
    assert r["batched"] > r["single_instance_gpu"]


def test_surrogate_gain_over_the_solver():
    """What the trained network gains: 1024 space-charge, seeded points train the ensemble (600
    epochs); on 128 held-out points every returned P(z) -- the surrogate's where the spread gate
    accepts, the solver's elsewhere -- is judged against its FP64 re-solve, and only points within
    2 % count.  The surrogate with fallback must deliver at least 50 times the points per hour of
    the plain single-instance GPU path and more than the batched solver alone; points beyond
    the training ranges must mostly go to the solver."""
    import tempfile
    from mithra_amd import surrogate
    with tempfile.TemporaryDirectory() as w:
        out = surrogate.run_workflow(w, n_train=1024, n_hold=128, epochs=600, n_single=8, n_outside=32)
    r = out["points_per_hour"]
    print({k: out[k] for k in ("surrogate_p_err", "gate", "outside_ranges", "train_s")}, flush=True)
    print(r, flush=True)
    assert r["returned_within_2pct"] >= 0.95
    assert r["gain_vs_single"] >= 50.0
    assert r["gain_vs_batched"] > 1.0
    assert out["outside_ranges"]["accepted"] <= 0.5
