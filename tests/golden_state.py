"""Golden fixtures -> C-ABI parameter structs, and the one-domain step sequence.

Test infrastructure.  A fixture (tests/golden/<case>.npz, written by make_golden.py from the
real reference) holds the reference's state before step 0 and every coefficient it used;
`Case` rebuilds the C-ABI structs from it and `march` replays Solver::solve()'s loop body
(src/solver.cpp:1300-1399) on any backend with the mithra_dev_* API.
"""
import ctypes as C
import os

import numpy as np

from mithra_amd import abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# the march fixtures: one per job file (other .npz files hold known-answer tables)
CASES = sorted(f[:-4] for f in os.listdir(GOLDEN)
               if f.endswith(".npz") and os.path.exists(os.path.join(GOLDEN, "jobs", f[:-4] + ".job")))
GPU_CASES = list(CASES)
# Cases with optical undulators or external-field waves (laser beams in the push)
BEAM_CASES = [c for c in CASES if "optical-undulator" in open(os.path.join(GOLDEN, "jobs", c + ".job")).read() or
              "EXTERNAL-FIELD" in open(os.path.join(GOLDEN, "jobs", c + ".job")).read()]
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(REPO, "oracle", "_build", "libcpu_ref.so")


def load_oracle():
    if not os.path.exists(ORACLE_LIB):
        raise RuntimeError(f"{ORACLE_LIB} missing: run make -C oracle")
    return C.CDLL(ORACLE_LIB)


def n_substeps(n_update):
    """Iterations of `for (Double t = 0.0; t < nUpdateBunch_; t += 1.0)` (src/solver.cpp:1316)."""
    n, t = 0, 0.0
    while t < n_update:
        n += 1
        t += 1.0
    return n


class Case:
    def __init__(self, name):
        self.name = name
        d = np.load(os.path.join(GOLDEN, name + ".npz"))
        self.d = d
        s = lambda k: float(d[k])  # noqa: E731
        g = abi.Grid()
        g.N0, g.N1, g.N2 = int(d["N0"]), int(d["N1"]), int(d["N2"])
        g.solver, g.truncation_order = int(d["solver"]), int(d["truncationOrder"])
        g.space_charge = int(d["spaceCharge"])
        for k in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax", "zp0", "zp1"):
            setattr(g, k, s(k))
        g.dx, g.dy, g.dz, g.dt = s("uf_dx"), s("uf_dy"), s("uf_dz"), s("uf_dt")
        g.Lz = s("meshLength2")
        self.grid = g
        c = abi.Coeffs()
        for k, n in (("a", 6), ("bB", 5), ("cB", 5), ("dB", 5), ("eE", 5), ("fE", 5), ("gE", 5),
                     ("hC", 17)):
            getattr(c, k)[:] = [float(v) for v in d["uf_" + k][:n]]
        c.alpha, c.beta = s("af_alpha"), s("af_beta")
        self.coeffs = c
        sd = abi.Seed()
        sd.seed_type, sd.signal_type = int(d["seed_type"]), int(d["seed_signal_type"])
        sd.nR = int(d["seed_nR"])
        sd.amplitude = s("seed_amplitude")
        sd.position[:] = list(d["seed_position"])
        sd.direction[:] = list(d["seed_direction"])
        sd.polarization[:] = list(d["seed_polarization"])
        sd.radius[:] = list(d["seed_radius"][:2])
        if "seed_sigma_invg" in d:  # fixtures of jobs whose seed needs them
            sd.sigma_invg[:] = [float(v) for v in d["seed_sigma_invg"][:2]]
            sd.order[:] = [int(v) for v in d["seed_order"][:2]]
        sd.t0, sd.s, sd.f0, sd.cep = s("seed_t0"), s("seed_s"), s("seed_f0"), s("seed_cep")
        sd.gamma, sd.beta, sd.dt_shift, sd.c0 = (s("seed_gamma"), s("seed_beta"), s("seed_dt"),
                                                 s("seed_c0"))
        self.seed = sd
        und = d["undulators"].reshape(-1, 7)
        self.undulators = [abi.Undulator(int(u[0]), int(u[4]), u[1], u[2], u[3], u[5], u[6])
                           for u in und]
        self.bunch = abi.BunchParams(s("gamma"), s("beta"), s("c0"), s("dt_shift"), s("ub_r1"),
                                     s("ub_r2"), s("ub_dtb"))
        # laser beams of optical undulators and external fields (27 values each, ref_harness.cpp)
        self.optical, self.external = [], []
        for v in (d["beams"].reshape(-1, 27) if "beams" in d else []):
            b = abi.Beam()
            b.beam_type, b.signal_type = int(v[1]), int(v[2])
            b.order[:] = [int(v[3]), int(v[4])]
            b.nR, b.amplitude = int(v[5]), v[6]
            b.position[:], b.direction[:], b.polarization[:] = list(v[7:10]), list(v[10:13]), list(v[13:16])
            b.radius[:], b.zR[:], b.l = list(v[16:18]), list(v[18:20]), v[20]
            b.t0, b.s, b.f0, b.cep = v[21], v[22], v[23], v[24]
            b.sigma_invg[:] = list(v[25:27])
            (self.optical if v[0] == 0 else self.external).append(b)
        pw = np.asarray(d["power"], dtype=np.float64)
        N, Nz, Nl, Nf = (int(v) for v in pw[:4])
        self.power_N, self.power_Nl, self.power_Nf = N, Nl, Nf
        self.power = abi.make_power_params(pw[5 + Nl:5 + Nl + N], pw[5:5 + Nl], Nf, pw[4],
                                           s("gamma"), s("beta"), s("c0"))
        self.n_sub = n_substeps(s("nUpdateBunch"))
        self.dt, self.dtb = s("uf_dt"), s("bunchTimeStep")
        self.time0, self.time_bunch0 = s("time"), s("timeBunch")
        self.n_steps = int(d["n_steps"])
        self.gamma, self.beta, self.c0, self.dt_shift = s("gamma"), s("beta"), s("c0"), s("dt_shift")

    def arr(self, key):
        return np.asarray(self.d[key])

    def power_file(self):
        return np.loadtxt(os.path.join(GOLDEN, self.name + "_power.txt"))

    def open(self, backend="gpu", precision=abi.PRECISION_FP64):
        """A Domain holding the reference's state before step 0."""
        if backend == "oracle":
            dom = abi.Domain(self.grid, self.coeffs, lib=load_oracle(), prefix="oracle_")
        else:
            dom = abi.Domain(self.grid, self.coeffs, precision=precision)
        sc = self.grid.space_charge
        dom.set_potentials(0, self.arr("pre0_an"), self.arr("pre0_fn") if sc else None)
        dom.set_potentials(1, self.arr("pre0_anm1"), self.arr("pre0_fnm1") if sc else None)
        dom.set_potentials(2, self.arr("pre0_jn"), self.arr("pre0_rn") if sc else None)
        dom.set_particles(self.arr("pre0_particles"))
        dom.power_setup(self.power, max(self.n_steps, 64))  # room for the tests' own marches
        if self.optical or self.external:
            dom.set_beams(self.optical, self.external)
        return dom


def march(case, dom, n_steps, start_step=0, time=None, time_bunch=None):
    """Solver::solve() main-loop body (src/solver.cpp:1300-1399) for n_steps on one domain.
    Returns (time, time_bunch) after the last step; P rows are stored per step."""
    time = case.time0 if time is None else time
    tb = case.time_bunch0 if time_bunch is None else time_bunch
    for step in range(start_step, start_step + n_steps):
        dom.field_update(time, case.seed)
        dom.bunch_begin()
        for _ in range(case.n_sub):
            dom.bunch_update(case.bunch, case.undulators, tb)
            tb += case.dtb
        dom.power_sample(step, step)
        dom.field_shift()
        dom.current_reset()
        dom.current_update()
        time += case.dt
    return time, tb


def z_column(case, time_bunch_at_sample):
    """The first column the reference writes per P(z) row (src/radiation.cpp:226)."""
    z = np.asarray(case.power._arrays[0])
    return case.gamma * (z + case.beta * case.c0 * (time_bunch_at_sample + case.dt_shift))


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))
