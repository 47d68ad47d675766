"""Learned surrogate of FEL scan points, with FDTD/PIC fallback.

A scan point is described by 8 beam/undulator scalars (gamma, K, lambda_u, energy spread,
normalised emittance, peak current, sigma_r, sigma_z) and its initial current profile (256 bins
over the mesh's z range); the surrogate predicts P(z) and the bunching factor b(z) at the
SAMPLES sample steps of the march.  It is an ensemble of MEMBERS multilayer perceptrons
(4 layers, 512 wide, SiLU) trained on the GPU by the library's BF16-MFMA kernels on runs that
the batched solver itself produces.  At use, the ensemble's spread decides: a point whose spread
in P exceeds SPREAD_LIMIT of the predicted peak power is re-solved by FDTD/PIC.

The points come from the space-charge, seeded generator: every point has space charge and a
1 MW Gaussian-beam seed at its resonant wavelength (waist ten times sigma_r), scalars drawn
around a 30 MeV, 500 A, 2 cm, K 0.8 undulator beam.

Targets are log10(1 + P / 1 W) and b, each column standardised over the training set; inputs are
the scalars (logs of the scale-like ones) standardised, and the profile normalised to peak 1.
"""
import math
import os
import time

import numpy as np

from . import abi, scan

MEMBERS = 5
HIDDEN = 512
PROFILE_BINS = 256
SPREAD_LIMIT = 0.02
DIMS = [8 + PROFILE_BINS, HIDDEN, HIDDEN, HIDDEN, 2 * scan.SAMPLES]
SEED_POWER = 1.0e6     # W
PASSES = 4             # window light-crossing times per march (4 n_z steps)


def generate(n, seed, n_perp=32, n_z=128, n_particles=16384, outside=False):
    """Scan points of the space-charge, seeded generator on one mesh shape, n_perp x n_perp x n_z
    cells; gamma U[40, 80], K U[0.6, 1.0], lambda_u U[1.5, 2.5] cm, energy spread log-U[1e-3,
    4e-3], normalised emittance log-U[0.5, 2] um, peak current U[250, 1000] A, sigma_r and sigma_z
    fractions of the window.  (The mesh is not among the surrogate's inputs, so it must not vary:
    P sampled over a transverse window depends on the window's size.)  Every point has space
    charge and the seed (SEED_POWER at lambda_r, waist 10 sigma_r); the bunch is loaded by the
    reference's quiet start without shot noise, so P(z) is a function of the inputs.  The march
    is PASSES window crossings, P(z) and b(z) sampled at SAMPLES evenly spaced steps.
    outside: points beyond the training ranges (gamma U[90, 120], K U[1.1, 1.4], peak current
    U[1500, 2500] A), for checking that the spread gate sends them to the solver."""
    rng = np.random.default_rng(seed)

    def logu(lo, hi):
        return math.exp(rng.uniform(math.log(lo), math.log(hi)))

    out = []
    for _ in range(n):
        gamma = rng.uniform(90.0, 120.0) if outside else rng.uniform(40.0, 80.0)
        K = rng.uniform(1.1, 1.4) if outside else rng.uniform(0.6, 1.0)
        lu, spread, emit = rng.uniform(1.5e4, 2.5e4), logu(1e-3, 4e-3), logu(0.5, 2.0)
        current = rng.uniform(1500.0, 2500.0) if outside else rng.uniform(250.0, 1000.0)
        out.append(scan.Instance(
            n_perp=n_perp, n_z=n_z, n_particles=n_particles, gamma=gamma, K=K, lambda_u=lu, spread=spread,
            emittance=emit, current=current, sr_frac=rng.uniform(1.0 / 10.0, 1.0 / 6.0),
            sz_frac=rng.uniform(1.0 / 24.0, 1.0 / 16.0), space_charge=True, seed_power=SEED_POWER,
            passes=PASSES, quiet=True))
    return out


def profile(inst, bins=PROFILE_BINS):
    """Initial current profile of a point as its description gives it: the charge per bin of its
    Gaussian bunch (truncated at 3.5 sigma_z, as the job loads it) over the mesh's z range,
    normalised to peak 1.  A point is judged without generating its particles."""
    lz = (inst.n_z - 1.5) * inst.dz
    sz = scan.bunch_sizes(inst)[1]
    edges = np.linspace(-lz / 2.0, lz / 2.0, bins + 1)
    cut = np.clip(edges, -3.5 * sz, 3.5 * sz)
    erf = np.vectorize(math.erf)
    cdf = 0.5 * (1.0 + erf(cut / (math.sqrt(2.0) * sz)))
    h = np.diff(cdf)
    return h / max(h.max(), 1e-300)


def scalars(inst):
    return np.array([math.log(inst.gamma), inst.K, math.log(inst.lambda_u), math.log(inst.spread),
                     math.log(inst.emittance), inst.current / 1000.0, inst.sr_frac * 8.0, inst.sz_frac * 20.0])


def features(insts):
    """Inputs of the points: their scalars and initial current profiles."""
    return np.array([np.concatenate([scalars(i), profile(i)]) for i in insts])


def targets(results):
    """results: [(P samples, b samples)] -> n x (2 SAMPLES)."""
    return np.array([np.concatenate([np.log10(1.0 + np.maximum(P, 0.0)), b]) for P, b in results])


def power_from_targets(y):
    return 10.0 ** y[..., :scan.SAMPLES] - 1.0


def solve(insts, workdir, precision=abi.PRECISION_FP32_BULK, chunk=256, log=None):
    """FDTD/PIC solutions of the points through the batched solver, `chunk` at a time; returns
    (results, seconds of set-up, seconds of marching)."""
    res, t_setup, t_march = [], 0.0, 0.0
    for c0 in range(0, len(insts), chunk):
        part = insts[c0:c0 + chunk]
        t0 = time.perf_counter()
        prep = scan.prepare(part, workdir, precision)
        t_setup += time.perf_counter() - t0
        r, secs, _ = scan.run_batch(prep, autotune=False)
        t_march += secs
        res.extend(r)
        for p in prep:
            p.dom.close()
            p.job.close()
        if log:
            log(f"solved {c0 + len(part)}/{len(insts)} points")
    return res, t_setup, t_march


class Surrogate:
    """The trained ensemble with its normalisations."""

    def __init__(self, seed=1):
        self.mlp = abi.MLP(DIMS, members=MEMBERS, seed=seed)
        self.x_mean = self.x_std = self.y_mean = self.y_std = None

    def fit(self, X, Y, epochs=2000, batch=256, lr=1e-3, seed=0, log=None):
        """Adam with a cosine-decayed rate over `epochs` epochs of the standardised set, kept on
        the device and run as one captured graph per epoch; returns each epoch's mean loss."""
        self.x_mean, self.x_std = X.mean(0), X.std(0) + 1e-6
        self.x_mean[8:], self.x_std[8:] = 0.0, 1.0          # the profile stays as it is
        self.y_mean, self.y_std = Y.mean(0), Y.std(0) + 1e-6
        xn = ((X - self.x_mean) / self.x_std).astype(np.float32)
        yn = ((Y - self.y_mean) / self.y_std).astype(np.float32)
        self.mlp.set_data(xn, yn)
        history = list(self.mlp.train_epochs(epochs, batch, lr, 1e-6, seed))
        if log:
            for ep in range(max(1, epochs // 10) - 1, epochs, max(1, epochs // 10)):
                log(f"epoch {ep + 1}/{epochs}: loss {history[ep]:.4f}")
        return history

    def predict(self, X):
        """(mean prediction, per-member predictions) in target units."""
        xn = ((X - self.x_mean) / self.x_std).astype(np.float32)
        y = self.mlp.forward(xn) * self.y_std + self.y_mean
        return y.mean(0), y

    def decide(self, X):
        """(predicted P, predicted b, accept): accept where the ensemble's spread in P stays within
        SPREAD_LIMIT of the predicted peak power."""
        mean, members = self.predict(X)
        P_members = power_from_targets(members)
        P = power_from_targets(mean)
        spread = P_members.std(0).max(axis=1)
        peak = np.maximum(P.max(axis=1), 1e-30)
        return P, mean[:, scan.SAMPLES:], spread <= SPREAD_LIMIT * peak


def p_error(P, P_ref):
    """P(z) error of one point: relative L1 over the sample steps."""
    return float(np.sum(np.abs(P - P_ref)) / max(np.sum(np.abs(P_ref)), 1e-300))


def run_workflow(workdir, n_train=4096, n_hold=512, epochs=2000, train_seed=5, hold_seed=6, n_single=32,
                 log=None, save=None, batch=256, lr=1e-3, n_outside=64):
    """The whole surrogate exercise on one GPU: batched FP32-bulk solves of n_train points, the
    held-out points solved in FP32-bulk and re-solved in FP64 (the reference for every error),
    training, the spread-gated decision, and scan points/hour of three ways to evaluate the
    held-out scan: the plain single-instance GPU path (measured on n_single points), the batched
    solver, and surrogate + batched fallback.  A point counts toward a rate only when its result
    (the surrogate's P(z) where the gate accepts it, the fallback's elsewhere) is within 2 % of
    the FP64 solve.  n_outside points beyond the training ranges (FP64-solved) show what the gate
    does with them.  Returns a summary dict."""
    out = {"members": MEMBERS, "dims": DIMS, "n_train": n_train, "n_hold": n_hold, "epochs": epochs,
           "batch": batch, "generator": {"space_charge": True, "seed_power_W": SEED_POWER, "passes": PASSES}}
    train, hold = generate(n_train, train_seed), generate(n_hold, hold_seed)
    res_tr, s1, m1 = solve(train, workdir, abi.PRECISION_FP32_BULK, log=log)
    res_h32, s2, m2 = solve(hold, workdir, abi.PRECISION_FP32_BULK)
    res_h64, s3, m3 = solve(hold, workdir, abi.PRECISION_FP64)
    out["solve_s"] = {"train_setup": s1, "train_march": m1, "hold_fp32_setup": s2, "hold_fp32_march": m2,
                      "hold_fp64_setup": s3, "hold_fp64_march": m3}
    P64 = np.array([P for P, _ in res_h64])
    b64 = np.array([b for _, b in res_h64])
    err32 = np.array([p_error(P, r) for (P, _), r in zip(res_h32, P64)])
    out["fp32_bulk_p_err_vs_fp64"] = {"max": float(err32.max()), "median": float(np.median(err32))}
    Xtr, Ytr, Xh = features(train), targets(res_tr), features(hold)
    if save:  # the generated data set: features, targets, FP64 held-out P(z) and b(z)
        np.savez(save, Xtr=Xtr, Ytr=Ytr, Xh=Xh, Ph64=P64, bh64=b64)
    # what a constant prediction (the training set's mean curve) would score
    Pmean = power_from_targets(Ytr.mean(0))
    out["mean_curve_p_err_median"] = float(np.median([p_error(Pmean, r) for r in P64]))
    model = Surrogate(seed=1)
    t0 = time.perf_counter()
    hist = model.fit(Xtr, Ytr, epochs=epochs, batch=batch, lr=lr, log=log)
    out["train_s"] = time.perf_counter() - t0
    out["final_loss"] = float(hist[-1])
    t0 = time.perf_counter()
    mean, members = model.predict(Xh)
    out["infer_s"] = time.perf_counter() - t0
    P = power_from_targets(mean)
    spread = power_from_targets(members).std(0).max(axis=1) / np.maximum(P.max(axis=1), 1e-30)
    err = np.array([p_error(a, r) for a, r in zip(P, P64)])
    out["surrogate_p_err"] = {"median": float(np.median(err)), "within_2pct": float(np.mean(err <= 0.02)),
                              "p90": float(np.quantile(err, 0.9))}
    out["surrogate_b_err_median"] = float(np.median(np.abs(mean[:, scan.SAMPLES:] - b64).max(axis=1)))
    gate = {}
    for lim in (0.01, 0.02, 0.05, 0.1):
        acc = spread <= lim
        gate[str(lim)] = {"accepted": float(acc.mean()),
                          "accepted_err_median": float(np.median(err[acc])) if acc.any() else None,
                          "accepted_within_2pct": float(np.mean(err[acc] <= 0.02)) if acc.any() else None}
    out["gate"] = gate
    if n_outside:
        far = generate(n_outside, hold_seed + 1, outside=True)
        Pf = np.array([P for P, _ in solve(far, workdir, abi.PRECISION_FP64)[0]])
        Pp, _, acc_f = model.decide(features(far))
        err_f = np.array([p_error(a, r) for a, r in zip(Pp, Pf)])
        out["outside_ranges"] = {"points": n_outside, "accepted": float(acc_f.mean()),
                                 "err_median": float(np.median(err_f)),
                                 "accepted_within_2pct": float(np.mean(err_f[acc_f] <= 0.02)) if acc_f.any() else None}
    # scan points/hour on the held-out scan
    sub = hold[:n_single]
    t0 = time.perf_counter()
    prep = scan.prepare(sub, workdir, abi.PRECISION_FP32_BULK)
    t_set = time.perf_counter() - t0
    _, t_march = scan.run_single(prep)
    for p in prep:
        p.dom.close()
        p.job.close()
    single = (t_set + t_march) / len(sub)
    t0 = time.perf_counter()
    _, _, accept = model.decide(features(hold))
    rejected = [h for h, a in zip(hold, accept) if not a]
    res_fb = solve(rejected, workdir, abi.PRECISION_FP32_BULK)[0] if rejected else []
    t_sur = time.perf_counter() - t0
    # the returned P(z): the surrogate's where accepted, the fallback's elsewhere
    final = err.copy()
    fb = iter(res_fb)
    for i, a in enumerate(accept):
        if not a:
            final[i] = p_error(next(fb)[0], P64[i])
    good = final <= 0.02
    rates = {"single_instance_gpu": 3600.0 / single,
             "batched": 3600.0 * n_hold / (s2 + m2),
             "surrogate_with_fallback": 3600.0 * n_hold / t_sur,
             "surrogate_with_fallback_within_2pct": 3600.0 * float(good.sum()) / t_sur,
             "accepted_at_limit": float(accept.mean()), "spread_limit": SPREAD_LIMIT,
             "returned_within_2pct": float(good.mean()),
             "accepted_err_median": float(np.median(err[accept])) if accept.any() else None}
    rates["gain_vs_single"] = rates["surrogate_with_fallback_within_2pct"] / rates["single_instance_gpu"]
    rates["gain_vs_batched"] = rates["surrogate_with_fallback_within_2pct"] / rates["batched"]
    out["points_per_hour"] = rates
    return out


def bench_gemm(log=print):
    """Times the surrogate's GEMM alone: the training products (256-row mini-batches through the
    512-wide layers, one member) and a large square product, in TFLOP/s and as a fraction of the
    2.5 PFLOP/s dense BF16 peak."""
    out = []
    rng = np.random.default_rng(0)
    for M, N, K, ta, tb, reps in ((256, 512, 512, False, False, 200), (512, 512, 256, True, False, 200),
                                  (256, 512, 512, False, True, 200), (8192, 8192, 8192, False, False, 10)):
        A = rng.normal(0.0, 1.0, (K, M) if ta else (M, K)).astype(np.float32)
        B = rng.normal(0.0, 1.0, (N, K) if tb else (K, N)).astype(np.float32)
        for tile in (1, 2):
            _, ms = abi.gemm_bf16(A, B, ta, tb, reps=reps, tile=tile)
            tf = 2.0 * M * N * K / (ms * 1e-3) / 1e12
            row = {"M": M, "N": N, "K": K, "ta": ta, "tb": tb, "tile": tile, "ms": ms, "tflops": tf,
                   "frac_of_2500": tf / 2500.0}
            log(row)
            out.append(row)
    return out


def main():
    """`python -m mithra_amd.surrogate [--train N] [--hold M] [--epochs E]`: the whole surrogate
    exercise on one GPU (see run_workflow), printing its summary as one JSON line."""
    import argparse
    import json
    import sys
    import tempfile
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--train", type=int, default=4096)
    ap.add_argument("--hold", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=2000)
    ap.add_argument("--single", type=int, default=32, help="held-out points timed on the single-instance path")
    ap.add_argument("--save", default=None, help="write the generated data set to this .npz file")
    ap.add_argument("--bench-gemm", action="store_true", help="time the GEMM alone and exit")
    ap.add_argument("--outside", type=int, default=64, help="points beyond the training ranges")
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    args = ap.parse_args()
    if args.bench_gemm:
        bench_gemm(lambda r: print(json.dumps(r), flush=True))
        return
    with tempfile.TemporaryDirectory(prefix="mithra_surrogate_") as work:
        out = run_workflow(work, n_train=args.train, n_hold=args.hold, epochs=args.epochs, n_single=args.single,
                           log=lambda s: print(f"[surrogate] {s}", file=sys.stderr, flush=True), save=args.save,
                           batch=args.batch, lr=args.lr, n_outside=args.outside)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
