"""Learned surrogate of FEL scan points, with FDTD/PIC fallback.

A scan point is described by 8 beam/undulator scalars (gamma, K, lambda_u, energy spread,
normalised emittance, peak current, sigma_r, sigma_z) and its initial current profile (256 bins
over the mesh's z range, from the particles of the set-up); the surrogate predicts P(z) and the bunching
factor b(z) at the SAMPLES sample steps of the march.  It is an ensemble of MEMBERS multilayer
perceptrons (4 layers, 512 wide, SiLU) trained on the GPU by the library's BF16-MFMA kernels on
runs that the batched solver itself produces.  At use, the ensemble's spread decides: a point
whose spread in P exceeds SPREAD_LIMIT of the predicted peak power is re-solved by FDTD/PIC.

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


def generate(n, seed, n_perp=32, n_z=128, n_particles=16384):
    """Scan points for the surrogate: every beam/undulator scalar varies on one mesh shape,
    n_perp x n_perp x n_z cells with n_z = SAMPLES, so P(z) and b(z) are sampled at every step.
    (The mesh is not among the surrogate's inputs, so it must not vary: P sampled over a
    transverse window depends on the window's size.)  The bunch is loaded by the reference's
    quiet start with an imposed initial bunching (amplitude 0.5-5 %, random phase), which the
    current profile shows: P(z) is then a function of the inputs rather than of shot noise.  The
    bunch (sigma_z at most 1/16 of the window) stays clear of the power plane 3/8 of the window
    ahead of its centre, so the plane sees radiation, not the bunch's own field.  The halton
    quiet start makes the loading deterministic: every point scales the same n_particles
    halton-sequence draws by its own sizes and spreads."""
    rng = np.random.default_rng(seed)

    def logu(lo, hi):
        return math.exp(rng.uniform(math.log(lo), math.log(hi)))

    out = []
    for _ in range(n):
        out.append(scan.Instance(
            n_perp=n_perp, n_z=n_z,
            n_particles=n_particles, gamma=rng.uniform(100.0, 2000.0),
            K=rng.uniform(0.5, 2.0), lambda_u=rng.uniform(1.0e4, 3.0e4), spread=logu(1e-4, 1e-2),
            emittance=logu(0.5, 5.0), current=rng.uniform(200.0, 3000.0),
            sr_frac=rng.uniform(1.0 / 12.0, 1.0 / 6.0), sz_frac=rng.uniform(1.0 / 24.0, 1.0 / 16.0),
            bunching=rng.uniform(0.005, 0.05), phase=rng.uniform(0.0, 2.0 * math.pi)))
    return out


def scalars(inst):
    return np.array([math.log(inst.gamma), inst.K, math.log(inst.lambda_u), math.log(inst.spread),
                     math.log(inst.emittance), inst.current / 1000.0, inst.sr_frac * 8.0, inst.sz_frac * 20.0])


def features(insts, profiles):
    """Inputs of the points: their scalars and initial current profiles (from the set-up, see
    solve)."""
    return np.array([np.concatenate([scalars(i), p]) for i, p in zip(insts, profiles)])


def targets(results):
    """results: [(P samples, b samples)] -> n x (2 SAMPLES)."""
    return np.array([np.concatenate([np.log10(1.0 + np.maximum(P, 0.0)), b]) for P, b in results])


def power_from_targets(y):
    return 10.0 ** y[..., :scan.SAMPLES] - 1.0


def solve(insts, workdir, precision=abi.PRECISION_FP32_BULK, chunk=256, log=None, profiles=None):
    """FDTD/PIC solutions of the points through the batched solver, `chunk` at a time; returns
    (results, seconds of set-up, seconds of marching).  profiles (a list, optional) receives each
    point's initial current profile, read from its set-up outside the timed parts."""
    res, t_setup, t_march = [], 0.0, 0.0
    for c0 in range(0, len(insts), chunk):
        part = insts[c0:c0 + chunk]
        t0 = time.perf_counter()
        prep = scan.prepare(part, workdir, precision)
        t_setup += time.perf_counter() - t0
        if profiles is not None:
            profiles.extend(p.profile(PROFILE_BINS) for p in prep)
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
        self.x_mean, self.x_std = X.mean(0), X.std(0) + 1e-6
        self.x_mean[8:], self.x_std[8:] = 0.0, 1.0          # the profile stays as it is
        self.y_mean, self.y_std = Y.mean(0), Y.std(0) + 1e-6
        xn = ((X - self.x_mean) / self.x_std).astype(np.float32)
        yn = ((Y - self.y_mean) / self.y_std).astype(np.float32)
        rng = np.random.default_rng(seed)
        n = len(X)
        history = []
        for ep in range(epochs):
            order = rng.permutation(n)
            rate = lr * 0.5 * (1.0 + math.cos(math.pi * ep / epochs))   # cosine decay
            losses = []
            for b0 in range(0, n - batch + 1, batch):
                idx = order[b0:b0 + batch]
                losses.append(self.mlp.train_step(xn[idx], yn[idx], max(rate, 1e-6)))
            history.append(float(np.mean(losses)))
            if log and (ep + 1) % max(1, epochs // 10) == 0:
                log(f"epoch {ep + 1}/{epochs}: loss {history[-1]:.4f}")
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
                 log=None, save=None):
    """The whole surrogate exercise on one GPU: batched FP32-bulk solves of n_train points, the
    held-out points solved in FP32-bulk and re-solved in FP64 (the reference for every error),
    training, the spread-gated decision, and scan points/hour of three ways to evaluate the
    held-out scan: the plain single-instance GPU path (measured on n_single points), the batched
    solver, and surrogate + batched fallback.  Returns a summary dict."""
    out = {"members": MEMBERS, "dims": DIMS, "n_train": n_train, "n_hold": n_hold, "epochs": epochs}
    train, hold = generate(n_train, train_seed), generate(n_hold, hold_seed)
    prof_tr, prof_h = [], []
    res_tr, s1, m1 = solve(train, workdir, abi.PRECISION_FP32_BULK, log=log, profiles=prof_tr)
    res_h32, s2, m2 = solve(hold, workdir, abi.PRECISION_FP32_BULK, profiles=prof_h)
    res_h64, s3, m3 = solve(hold, workdir, abi.PRECISION_FP64)
    out["solve_s"] = {"train_setup": s1, "train_march": m1, "hold_fp32_setup": s2, "hold_fp32_march": m2,
                      "hold_fp64_setup": s3, "hold_fp64_march": m3}
    P64 = np.array([P for P, _ in res_h64])
    b64 = np.array([b for _, b in res_h64])
    out["fp32_bulk_p_err_vs_fp64"] = {"max": max(p_error(P, r) for (P, _), r in zip(res_h32, P64)),
                                      "median": float(np.median([p_error(P, r) for (P, _), r in zip(res_h32, P64)]))}
    Xtr, Ytr, Xh = features(train, prof_tr), targets(res_tr), features(hold, prof_h)
    if save:  # the generated data set: features, targets, FP64 held-out P(z) and b(z)
        np.savez(save, Xtr=Xtr, Ytr=Ytr, Xh=Xh, Ph64=P64, bh64=b64)
    model = Surrogate(seed=1)
    t0 = time.perf_counter()
    hist = model.fit(Xtr, Ytr, epochs=epochs, log=log)
    out["train_s"] = time.perf_counter() - t0
    out["final_loss"] = hist[-1]
    t0 = time.perf_counter()
    mean, members = model.predict(Xh)
    out["infer_s"] = time.perf_counter() - t0
    P = power_from_targets(mean)
    spread = power_from_targets(members).std(0).max(axis=1) / np.maximum(P.max(axis=1), 1e-30)
    err = np.array([p_error(a, r) for a, r in zip(P, P64)])
    out["surrogate_p_err"] = {"median": float(np.median(err)), "within_2pct": float(np.mean(err <= 0.02))}
    out["surrogate_b_err_median"] = float(np.median(np.abs(mean[:, scan.SAMPLES:] - b64).max(axis=1)))
    gate = {}
    for lim in (0.02, 0.05, 0.1, 0.2):
        acc = spread <= lim
        gate[str(lim)] = {"accepted": float(acc.mean()),
                          "accepted_err_median": float(np.median(err[acc])) if acc.any() else None,
                          "accepted_within_2pct": float(np.mean(err[acc] <= 0.02)) if acc.any() else None}
    out["gate"] = gate
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
    accept = spread <= SPREAD_LIMIT
    t0 = time.perf_counter()
    model.predict(Xh)
    rejected = [h for h, a in zip(hold, accept) if not a]
    if rejected:
        solve(rejected, workdir, abi.PRECISION_FP32_BULK)
    t_sur = time.perf_counter() - t0
    out["points_per_hour"] = {"single_instance_gpu": 3600.0 / single,
                              "batched": 3600.0 * n_hold / (s2 + m2),
                              "surrogate_with_fallback": 3600.0 * n_hold / t_sur,
                              "accepted_at_limit": float(accept.mean()), "spread_limit": SPREAD_LIMIT}
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
    args = ap.parse_args()
    with tempfile.TemporaryDirectory(prefix="mithra_surrogate_") as work:
        out = run_workflow(work, n_train=args.train, n_hold=args.hold, epochs=args.epochs, n_single=args.single,
                           log=lambda s: print(f"[surrogate] {s}", file=sys.stderr, flush=True), save=args.save)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
