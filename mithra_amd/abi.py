"""ctypes mirror of include/mithra_gpu.h and a thin wrapper over the device C-ABI.

The structs here are field-for-field the C structs of the header; the wrapper class
`Domain` binds either the HIP library (libmithra_gpu.so, prefix ``mithra_dev_``) or any
library with the same entry points under another prefix (the parity tests bind the CPU
oracle with prefix ``oracle_``), so both can be driven by one call sequence.
"""
import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
GPU_LIB = os.path.join(LIB_DIR, "libmithra_gpu.so")
HOST_LIB = os.path.join(LIB_DIR, "libmithra.so")

SOLVER_FD, SOLVER_NSFD = 0, 1
PRECISION_FP64, PRECISION_FP32_BULK = 0, 1
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


class Grid(C.Structure):
    _fields_ = [("N0", C.c_int), ("N1", C.c_int), ("N2", C.c_int), ("solver", C.c_int),
                ("truncation_order", C.c_int), ("space_charge", C.c_int),
                ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double),
                ("ymax", C.c_double), ("zmin", C.c_double), ("zmax", C.c_double),
                ("zp0", C.c_double), ("zp1", C.c_double), ("dx", C.c_double),
                ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double), ("Lz", C.c_double)]

    @property
    def nodes(self):
        return self.N0 * self.N1 * self.N2


class Coeffs(C.Structure):
    _fields_ = [("a", C.c_double * 6), ("bB", C.c_double * 5), ("cB", C.c_double * 5),
                ("dB", C.c_double * 5), ("eE", C.c_double * 5), ("fE", C.c_double * 5),
                ("gE", C.c_double * 5), ("hC", C.c_double * 17), ("alpha", C.c_double),
                ("beta", C.c_double)]


class Seed(C.Structure):
    _fields_ = [("seed_type", C.c_int), ("signal_type", C.c_int), ("order", C.c_int * 2),
                ("nR", C.c_uint), ("amplitude", C.c_double), ("position", C.c_double * 3),
                ("direction", C.c_double * 3), ("polarization", C.c_double * 3),
                ("radius", C.c_double * 2), ("t0", C.c_double), ("s", C.c_double),
                ("f0", C.c_double), ("cep", C.c_double), ("sigma_invg", C.c_double * 2),
                ("gamma", C.c_double), ("beta", C.c_double), ("dt_shift", C.c_double),
                ("c0", C.c_double)]


class Undulator(C.Structure):
    _fields_ = [("type", C.c_int), ("length", C.c_uint), ("k", C.c_double), ("lu", C.c_double),
                ("rb", C.c_double), ("dist", C.c_double), ("theta", C.c_double)]


class BunchParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("beta", C.c_double), ("c0", C.c_double),
                ("dt_shift", C.c_double), ("r1", C.c_double), ("r2", C.c_double),
                ("dtb", C.c_double)]


class PowerParams(C.Structure):
    _fields_ = [("N", C.c_uint), ("Nl", C.c_uint), ("Nf", C.c_uint), ("pc", C.c_double),
                ("gamma", C.c_double), ("beta", C.c_double), ("c0", C.c_double),
                ("z", _dp), ("w", _dp)]


class Beam(C.Structure):
    """mithra_beam: the laser beam of an optical undulator or an external field."""
    _fields_ = [("beam_type", C.c_int), ("signal_type", C.c_int), ("order", C.c_int * 2), ("nR", C.c_uint),
                ("amplitude", C.c_double), ("position", C.c_double * 3), ("direction", C.c_double * 3),
                ("polarization", C.c_double * 3), ("radius", C.c_double * 2), ("zR", C.c_double * 2),
                ("l", C.c_double), ("t0", C.c_double), ("s", C.c_double), ("f0", C.c_double), ("cep", C.c_double),
                ("sigma_invg", C.c_double * 2)]


class ScreenParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("gamma", "beta", "c0", "time_bunch", "time_step", "dt_shift")]


class MithraError(RuntimeError):
    pass


PHASES = ("interior", "tfsf", "boundaries", "eval_eb", "push", "power", "current_reset", "deposit")


TILE_VARIANTS = 10
FUSED_TILE_VARIANTS = 4   # shapes of the pipelined step's fused kernel (MITHRA_FUSED_TILE_VARIANTS)


def tile_shape(variant):
    """(rows, planes) of a field-kernel tile variant."""
    lib = load_gpu_library()
    f = lib.mithra_tile_shape
    f.restype, f.argtypes = C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    r, k = C.c_int(), C.c_int()
    if f(variant, C.byref(r), C.byref(k)) != 0:
        raise MithraError(f"no tile variant {variant}")
    return r.value, k.value


def copy_bandwidth(nbytes=2 << 30, reps=10, elem_bytes=32):
    """Measured device copy rate (GB/s, read + write) of a plain streaming kernel moving
    elem_bytes (8, 16, 32) per lane."""
    lib = load_gpu_library()
    f = lib.mithra_copy_bandwidth_elem
    f.restype, f.argtypes = C.c_int, [C.c_ulong, C.c_int, C.c_int, C.POINTER(C.c_double)]
    out = C.c_double()
    rc = f(nbytes, reps, elem_bytes, C.byref(out))
    if rc != 0:
        raise MithraError(f"mithra_copy_bandwidth: status {rc}")
    return out.value


def fused_tile_shape(variant):
    """(rows, planes) of a tile variant of the pipelined step's fused kernel."""
    lib = load_gpu_library()
    f = lib.mithra_fused_tile_shape
    f.restype, f.argtypes = C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    r, k = C.c_int(), C.c_int()
    if f(variant, C.byref(r), C.byref(k)) != 0:
        raise MithraError(f"no fused tile variant {variant}")
    return r.value, k.value


def stream_bandwidth(nbytes=2 << 30, reps=10):
    """Measured rate (GB/s, all bytes moved) of three streams read and one written, the interior
    stencil's access pattern without its arithmetic."""
    lib = load_gpu_library()
    f = lib.mithra_stream_bandwidth
    f.restype, f.argtypes = C.c_int, [C.c_ulong, C.c_int, C.POINTER(C.c_double)]
    out = C.c_double()
    rc = f(nbytes, reps, C.byref(out))
    if rc != 0:
        raise MithraError(f"mithra_stream_bandwidth: status {rc}")
    return out.value


def signal_evaluate(seed, t, phase):
    """Signal::self of the seed's signal at times t with phases `phase`, on the device."""
    lib = load_gpu_library()
    f = lib.mithra_signal_evaluate
    f.restype, f.argtypes = C.c_int, [C.POINTER(Seed), _dp, _dp, C.c_long, _dp]
    t = np.ascontiguousarray(t, dtype=np.float64)
    phase = np.ascontiguousarray(phase, dtype=np.float64)
    out = np.empty(t.size)
    rc = f(C.byref(seed), _ptr(t), _ptr(phase), t.size, _ptr(out))
    if rc != 0:
        raise MithraError(f"mithra_signal_evaluate: status {rc}")
    return out


def seed_evaluate(seed, r, time):
    """Seed::fields at boosted-frame points r (n x 3) and times (n), on the device; (n x 3)."""
    lib = load_gpu_library()
    f = lib.mithra_seed_evaluate
    f.restype, f.argtypes = C.c_int, [C.POINTER(Seed), _dp, _dp, C.c_long, _dp]
    r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 3)
    time = np.ascontiguousarray(time, dtype=np.float64)
    out = np.empty((time.size, 3))
    rc = f(C.byref(seed), _ptr(r), _ptr(time), time.size, _ptr(out))
    if rc != 0:
        raise MithraError(f"mithra_seed_evaluate: status {rc}")
    return out


def _declare(lib, prefix):
    vp = C.c_void_p
    sig = {
        "create": (C.c_int, [C.POINTER(Grid), C.POINTER(Coeffs), C.POINTER(vp)]),
        "destroy": (C.c_int, [vp]),
        "set_potentials": (C.c_int, [vp, C.c_int, _dp, _dp]),
        "get_potentials": (C.c_int, [vp, C.c_int, _dp, _dp]),
        "get_fields": (C.c_int, [vp, _fp, _fp]),
        "set_particles": (C.c_int, [vp, _dp, C.c_long]),
        "get_particles": (C.c_int, [vp, _dp, C.c_long]),
        "particle_count": (C.c_long, [vp]),
        "field_update": (C.c_int, [vp, C.c_double, C.POINTER(Seed)]),
        "field_evaluate": (C.c_int, [vp]),
        "bunch_begin": (C.c_int, [vp]),
        "bunch_update": (C.c_int, [vp, C.POINTER(BunchParams), C.POINTER(Undulator), C.c_int,
                                   C.c_double]),
        "bunch_counters": (C.c_int, [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
        "power_setup": (C.c_int, [vp, C.POINTER(PowerParams), C.c_uint]),
        "power_sample": (C.c_int, [vp, C.c_uint, C.c_uint]),
        "power_fetch": (C.c_int, [vp, C.c_uint, C.c_uint, _dp]),
        "field_shift": (C.c_int, [vp]),
        "current_reset": (C.c_int, [vp]),
        "current_update": (C.c_int, [vp]),
        "bunching": (C.c_int, [vp, C.c_double, C.POINTER(C.c_double)]),
        "set_beams": (C.c_int, [vp, C.POINTER(Beam), C.c_int, C.POINTER(Beam), C.c_int]),
    }
    if prefix == "mithra_dev_":
        sig["set_reorder_interval"] = (C.c_int, [vp, C.c_int])
        sig["set_deterministic"] = (C.c_int, [vp, C.c_int])
        sig["set_tiles"] = (C.c_int, [vp, C.c_int, C.c_int])
        sig["set_lazy_fields"] = (C.c_int, [vp, C.c_int])
        sig["autotune"] = (C.c_int, [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_float)])
        sig["bunch_advance"] = (C.c_int, [vp, C.POINTER(BunchParams), C.POINTER(Undulator), C.c_int,
                                          C.c_double, C.c_double, C.c_int])
        sig["synchronize"] = (C.c_int, [vp])
        sig["step_timed"] = (C.c_int, [vp, C.c_double, C.POINTER(Seed), C.POINTER(BunchParams),
                                       C.POINTER(Undulator), C.c_int, C.c_double, C.c_double,
                                       C.c_int, C.c_uint, C.c_int, C.POINTER(C.c_float)])
        sig["step"] = (C.c_int, [vp, C.c_double, C.POINTER(Seed), C.POINTER(BunchParams),
                                 C.POINTER(Undulator), C.c_int, C.c_double, C.c_double, C.c_int,
                                 C.c_uint, C.c_int])
        sig["step_pipelined"] = (C.c_int, [vp, C.c_double, C.POINTER(Seed), C.POINTER(BunchParams),
                                           C.POINTER(Undulator), C.c_int, C.c_double, C.c_double,
                                           C.c_int, C.c_uint, C.c_int, C.c_int,
                                           C.POINTER(C.c_float)])
        sig["pipeline_pending"] = (C.c_int, [vp])
        sig["set_fused_tile"] = (C.c_int, [vp, C.c_int])
        sig["autotune_fused"] = (C.c_int, [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_float)])
        sig["bunching_noalloc"] = (C.c_int, [vp, C.c_double, C.POINTER(C.c_double)])
        lp = C.POINTER(C.c_long)
        sig["sample_nodes"] = (C.c_int, [vp, lp, C.c_long, _fp, _fp, _dp, _dp])
        sig["bunch_sums"] = (C.c_int, [vp, _dp])
        sig["screen_crossings"] = (C.c_int, [vp, _dp, C.c_int, C.POINTER(ScreenParams), C.c_long, lp,
                                             C.POINTER(C.c_int), lp, _dp])
    fns = {}
    for name, (res, args) in sig.items():
        f = getattr(lib, prefix + name)
        if name == "create" and prefix == "mithra_dev_":
            args = [C.POINTER(Grid), C.POINTER(Coeffs), C.c_int, C.POINTER(vp)]
        f.restype, f.argtypes = res, args
        fns[name] = f
    return fns


_DECLARED = {}
_DECLARE_LOCK = threading.Lock()


def _functions(lib, prefix):
    """The library's entry points under `prefix`, declared once: ctypes function objects are
    shared, and re-assigning their argument types while another thread calls them is a race."""
    key = (id(lib), prefix)
    with _DECLARE_LOCK:
        if key not in _DECLARED:
            _DECLARED[key] = (lib, _declare(lib, prefix))
        return _DECLARED[key][1]


_GPU = None


def load_gpu_library(path=GPU_LIB):
    """Load libmithra_gpu.so; raises MithraError when it is not built (there is no fallback)."""
    global _GPU
    with _DECLARE_LOCK:
        if _GPU is None:
            _GPU = _open_gpu_library(path)
    return _GPU


def _open_gpu_library(path):
    if not os.path.exists(path):
        raise MithraError(f"{path} is missing: run __graft_entry__.build() first")
    lib = C.CDLL(path)
    lib.mithra_strerror.restype = C.c_char_p
    lib.mithra_strerror.argtypes = [C.c_int]
    return lib


def _ptr(a, ctype=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(ctype))


class Domain:
    """One simulation domain behind the device C-ABI (or a library with the same API)."""

    def __init__(self, grid, coeffs, lib=None, prefix="mithra_dev_", precision=PRECISION_FP64):
        self.lib = lib if lib is not None else load_gpu_library()
        self.prefix = prefix
        self.f = _functions(self.lib, prefix)
        self.grid = grid
        self.h = C.c_void_p()
        if prefix == "mithra_dev_":
            self._check(self.f["create"](C.byref(grid), C.byref(coeffs), precision, C.byref(self.h)))
        else:
            self._check(self.f["create"](C.byref(grid), C.byref(coeffs), C.byref(self.h)))
        self._keep = []

    def _check(self, rc):
        if rc != 0:
            msg = ""
            if hasattr(self.lib, "mithra_strerror"):
                msg = self.lib.mithra_strerror(rc).decode()
            raise MithraError(f"{self.prefix}: status {rc} {msg}")
        return rc

    def close(self):
        if self.h:
            self._check(self.f["destroy"](self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state transfer ------------------------------------------------------------
    def set_potentials(self, level, a, phi=None):
        a = np.ascontiguousarray(a, dtype=np.float64)
        phi = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64)
        self._check(self.f["set_potentials"](self.h, level, _ptr(a), _ptr(phi)))

    def get_potentials(self, level):
        n = self.grid.nodes
        a = np.empty(3 * n)
        phi = np.empty(n) if self.grid.space_charge else None
        self._check(self.f["get_potentials"](self.h, level, _ptr(a), _ptr(phi)))
        return a, phi

    def get_fields(self):
        n = self.grid.nodes
        e = np.empty(3 * n, dtype=np.float32)
        b = np.empty(3 * n, dtype=np.float32)
        self._check(self.f["get_fields"](self.h, _ptr(e, C.c_float), _ptr(b, C.c_float)))
        return e, b

    def set_particles(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        self._check(self.f["set_particles"](self.h, _ptr(q), q.size // 11))

    def get_particles(self):
        n = self.f["particle_count"](self.h)
        q = np.empty(11 * n)
        self._check(self.f["get_particles"](self.h, _ptr(q), n))
        return q.reshape(n, 11)

    # -- the step --------------------------------------------------------------------
    def field_evaluate(self):
        """E/B from A^n and A^{n+1} on every node (FdTd::fieldEvaluate)."""
        self._check(self.f["field_evaluate"](self.h))

    def field_update(self, time, seed=None):
        self._check(self.f["field_update"](self.h, time, None if seed is None else C.byref(seed)))

    def bunch_begin(self):
        self._check(self.f["bunch_begin"](self.h))

    def bunch_update(self, params, undulators, time_bunch):
        arr = (Undulator * max(1, len(undulators)))(*undulators)
        self._check(self.f["bunch_update"](self.h, C.byref(params), arr, len(undulators),
                                           time_bunch))

    def bunch_advance(self, params, undulators, time_bunch, bunch_time_step, n_sub):
        """rnm = rnp, then n_sub sub-steps in one launch (device library only)."""
        arr = (Undulator * max(1, len(undulators)))(*undulators)
        self._check(self.f["bunch_advance"](self.h, C.byref(params), arr, len(undulators),
                                            time_bunch, bunch_time_step, n_sub))

    def step(self, time, seed, params, undulators, time_bunch, bunch_time_step, n_sub, n_time,
             power_row=-1):
        """One whole main-loop iteration on the device (device library only)."""
        arr = (Undulator * max(1, len(undulators)))(*undulators)
        self._check(self.f["step"](self.h, time, None if seed is None else C.byref(seed),
                                   C.byref(params), arr, len(undulators), time_bunch,
                                   bunch_time_step, n_sub, n_time, power_row))

    def step_timed(self, time, seed, params, undulators, time_bunch, bunch_time_step, n_sub, n_time,
                   power_row=-1):
        """step() with per-phase device times in ms (PHASES order)."""
        arr = (Undulator * max(1, len(undulators)))(*undulators)
        ms = (C.c_float * 8)()
        self._check(self.f["step_timed"](self.h, time, None if seed is None else C.byref(seed),
                                         C.byref(params), arr, len(undulators), time_bunch,
                                         bunch_time_step, n_sub, n_time, power_row, ms))
        return list(ms)

    def step_pipelined(self, time, seed, params, undulators, time_bunch, bunch_time_step, n_sub,
                       n_time, power_row=-1, more=True, timed=False):
        """step() pipelined across calls (device library only): with more=True the E/B pass is
        fused with the next interior update and the deposit adds its J to that update, so the
        next call skips the interior update; the potentials are unavailable until a call with
        more=False ends the run.  FP64 only, with or without space charge (FP32-bulk, no
        particles or a fixed-order deposit: more=True runs the plain step).  Returns the phase
        times in ms when timed, else None."""
        arr = (Undulator * max(1, len(undulators)))(*undulators)
        ms = (C.c_float * 8)() if timed else None
        self._check(self.f["step_pipelined"](self.h, time, None if seed is None else C.byref(seed),
                                             C.byref(params), arr, len(undulators), time_bunch,
                                             bunch_time_step, n_sub, n_time, power_row, int(more),
                                             ms))
        return list(ms) if timed else None

    def pipeline_pending(self):
        """True while a pipelined run awaits its last (more=False) step."""
        r = self.f["pipeline_pending"](self.h)
        if r < 0:
            self._check(r)
        return bool(r)

    def set_fused_tile(self, variant):
        """Tile shape variant of the pipelined step's fused kernel."""
        self._check(self.f["set_fused_tile"](self.h, variant))

    def autotune_fused(self, reps=3):
        """Times every tile shape of the pipelined step's fused kernel and keeps the fastest.
        Returns (chosen, ms[variants])."""
        chosen = C.c_int()
        ms = (C.c_float * FUSED_TILE_VARIANTS)()
        self._check(self.f["autotune_fused"](self.h, reps, C.byref(chosen), ms))
        return chosen.value, np.array(ms[:])

    def synchronize(self):
        self._check(self.f["synchronize"](self.h))

    def bunch_counters(self):
        a, b = C.c_long(), C.c_long()
        self._check(self.f["bunch_counters"](self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def power_setup(self, params, rows):
        self._check(self.f["power_setup"](self.h, C.byref(params), rows))

    def power_sample(self, n_time, row):
        self._check(self.f["power_sample"](self.h, n_time, row))

    def power_fetch(self, first, count, n_planes, n_freq):
        out = np.empty(count * n_planes * n_freq)
        self._check(self.f["power_fetch"](self.h, first, count, _ptr(out)))
        return out.reshape(count, n_planes, n_freq)

    def field_shift(self):
        self._check(self.f["field_shift"](self.h))

    def current_reset(self):
        self._check(self.f["current_reset"](self.h))

    def current_update(self):
        self._check(self.f["current_update"](self.h))

    def set_reorder_interval(self, n):
        """Deposits between reorders of the particle arrays (0: never); device library only."""
        self._check(self.f["set_reorder_interval"](self.h, n))

    def set_deterministic(self, on=True):
        """Fixed-order current deposition (bit-for-bit repeatable FP64 runs); device library only."""
        self._check(self.f["set_deterministic"](self.h, 1 if on else 0))

    def set_lazy_fields(self, lazy):
        """E/B only where the bunch and the power sampler read them (1), or everywhere (0)."""
        self._check(self.f["set_lazy_fields"](self.h, int(lazy)))

    def set_tiles(self, interior, evaluation):
        """Tile shape variants of the interior update and the E/B evaluation."""
        self._check(self.f["set_tiles"](self.h, interior, evaluation))

    def autotune(self, reps=3):
        """Times every tile shape of both field kernels; keeps the fastest.  Returns
        (chosen (interior, eval), ms[2][variants])."""
        chosen = (C.c_int * 2)()
        ms = (C.c_float * (2 * TILE_VARIANTS))()
        self._check(self.f["autotune"](self.h, reps, chosen, ms))
        return (chosen[0], chosen[1]), np.array(ms[:]).reshape(2, TILE_VARIANTS)

    def set_beams(self, optical, external):
        """Laser beams of the optical undulators (table order) and of the external fields."""
        o = (Beam * max(len(optical), 1))(*optical)
        e = (Beam * max(len(external), 1))(*external)
        self._check(self.f["set_beams"](self.h, o, len(optical), e, len(external)))

    def bunching(self, k):
        """|sum q exp(i k z)| / |sum q| over the particles."""
        b = C.c_double()
        self._check(self.f["bunching"](self.h, k, C.byref(b)))
        return b.value

    def bunching_noalloc(self, k):
        """bunching() through the entry point that allocates no device memory (device library
        only)."""
        b = C.c_double()
        self._check(self.f["bunching_noalloc"](self.h, k, C.byref(b)))
        return b.value


    def sample_nodes(self, nodes):
        """(E, B, A^n, phi^n) at the listed node indices (device library only)."""
        nodes = np.ascontiguousarray(nodes, dtype=np.int64)
        n = nodes.size
        e, b = np.empty((n, 3), np.float32), np.empty((n, 3), np.float32)
        a, phi = np.empty((n, 3)), np.empty(n)
        self._check(self.f["sample_nodes"](self.h, _ptr(nodes, C.c_long), n, _ptr(e, C.c_float),
                                           _ptr(b, C.c_float), _ptr(a), _ptr(phi)))
        return e, b, a, phi

    def bunch_sums(self):
        """[sum q, sum q r (3), sum r^2 q (3), sum q gb (3), sum gb^2 q (3)] (device library only)."""
        out = np.empty(13)
        self._check(self.f["bunch_sums"](self.h, _ptr(out)))
        return out

    def screen_crossings(self, lz, params, cap=None):
        """(screen, particle index, records[n, 6]) of the particles that crossed the lab-frame
        screens at lz during the last field step (device library only)."""
        lz = np.ascontiguousarray(lz, dtype=np.float64)
        cap = self.f["particle_count"](self.h) * len(lz) if cap is None else cap
        scr, idx = np.empty(max(cap, 1), np.int32), np.empty(max(cap, 1), np.int64)
        rec = np.empty((max(cap, 1), 6))
        count = C.c_long()
        self._check(self.f["screen_crossings"](self.h, _ptr(lz), len(lz), C.byref(params), cap, C.byref(count),
                                               _ptr(scr, C.c_int), _ptr(idx, C.c_long), _ptr(rec)))
        n = min(count.value, cap)
        return scr[:n], idx[:n], rec[:n], count.value


class Ellipsoid(C.Structure):
    _fields_ = [("n", C.c_long), ("first_index", C.c_uint), ("distribution", C.c_int), ("shot_noise", C.c_int),
                ("charge", C.c_double), ("gamma", C.c_double), ("beta_vector", C.c_double * 3),
                ("position", C.c_double * 3), ("sigma_position", C.c_double * 3), ("sigma_gb", C.c_double * 3),
                ("tran_trunc", C.c_double), ("long_trunc", C.c_double), ("lambda_", C.c_double),
                ("bunching", C.c_double), ("bunching_phase", C.c_double)]


class Flattop(C.Structure):
    _fields_ = [("n", C.c_long), ("charge", C.c_double), ("gamma", C.c_double), ("lambda_", C.c_double),
                ("length", C.c_double), ("z_center", C.c_double), ("sigma_r", C.c_double),
                ("sigma_gb_perp", C.c_double), ("energy_spread", C.c_double), ("electrons_per_bucket", C.c_double)]


class DeviceBunch:
    """Charge records generated on the device (mithra_bunch_*, include/mithra_gpu.h)."""

    def __init__(self):
        self.lib = load_gpu_library()
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        for name, res, args in (
            ("mithra_bunch_create", C.c_int, [C.POINTER(vp)]),
            ("mithra_bunch_destroy", C.c_int, [vp]),
            ("mithra_bunch_count", C.c_long, [vp]),
            ("mithra_bunch_get", C.c_int, [vp, dp, C.c_long]),
            ("mithra_bunch_ellipsoid", C.c_int, [vp, C.POINTER(Ellipsoid), C.POINTER(C.c_long)]),
            ("mithra_bunch_flattop", C.c_int, [vp, C.POINTER(Flattop), C.c_ulonglong]),
        ):
            f = getattr(self.lib, name)
            f.restype, f.argtypes = res, args
        self.h = vp()
        self._check(self.lib.mithra_bunch_create(C.byref(self.h)))

    def _check(self, rc):
        if rc != 0:
            raise MithraError(f"status {rc}: {self.lib.mithra_strerror(rc).decode()}")

    def ellipsoid(self, spec):
        added = C.c_long()
        self._check(self.lib.mithra_bunch_ellipsoid(self.h, C.byref(spec), C.byref(added)))
        return added.value

    def flattop(self, spec, seed):
        self._check(self.lib.mithra_bunch_flattop(self.h, C.byref(spec), seed))

    def count(self):
        return self.lib.mithra_bunch_count(self.h)

    def records(self):
        n = self.count()
        q = np.empty((n, 11))
        self._check(self.lib.mithra_bunch_get(self.h, _ptr(q), n))
        return q

    def close(self):
        if self.h:
            self.lib.mithra_bunch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_power_params(z, w, Nf, pc, gamma, beta, c0):
    """PowerParams with its two arrays kept alive on the returned object."""
    z = np.ascontiguousarray(z, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    p = PowerParams(len(z), len(w), int(Nf), pc, gamma, beta, c0, _ptr(z), _ptr(w))
    p._arrays = (z, w)
    return p


class Batch:
    """A batch of device domains marched together (mithra_batch_*, include/mithra_gpu.h).

    `domains` are `Domain`s of the device library, already filled (potentials, particles,
    power set-up); they must stay alive and unused while the batch runs."""

    def __init__(self, domains):
        self.lib = load_gpu_library()
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        for name, res, args in (
            ("mithra_batch_create", C.c_int, [C.POINTER(vp), C.c_int, C.POINTER(vp)]),
            ("mithra_batch_destroy", C.c_int, [vp]),
            ("mithra_batch_instance", C.c_int, [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                                C.c_long, C.POINTER(BunchParams), C.POINTER(Undulator), C.c_int,
                                                C.c_double, C.c_long, C.c_int]),
            ("mithra_batch_seed", C.c_int, [vp, C.c_int, C.POINTER(Seed)]),
            ("mithra_batch_set_tiles", C.c_int, [vp, C.c_int, C.c_int]),
            ("mithra_batch_autotune", C.c_int, [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_float)]),
            ("mithra_batch_run", C.c_int, [vp, C.c_long, C.POINTER(C.c_long)]),
            ("mithra_batch_steps", C.c_int, [vp, C.c_int, C.POINTER(C.c_long)]),
            ("mithra_batch_bunching", C.c_int, [vp, C.c_int, dp]),
        ):
            f = getattr(self.lib, name)
            f.restype, f.argtypes = res, args
        self.domains = list(domains)
        handles = (vp * len(self.domains))(*[d.h for d in self.domains])
        self.h = vp()
        self._check(self.lib.mithra_batch_create(handles, len(self.domains), C.byref(self.h)))
        self.slots = [0] * len(self.domains)

    def _check(self, rc):
        if rc != 0:
            raise MithraError(f"status {rc}: {self.lib.mithra_strerror(rc).decode()}")

    def instance(self, i, time, time_bunch, bunch_time_step, n_sub, n_steps, bunch_params, undulators,
                 bunching_k=0.0, bunching_every=0, bunching_slots=0):
        und = (Undulator * max(1, len(undulators)))(*undulators)
        self._check(self.lib.mithra_batch_instance(self.h, i, time, time_bunch, bunch_time_step, n_sub, n_steps,
                                                   C.byref(bunch_params), und, len(undulators), bunching_k,
                                                   bunching_every, bunching_slots))
        self.slots[i] = bunching_slots

    def seed(self, i, seed):
        """Instance i's TF/SF seed (after instance(), before the first run)."""
        self._check(self.lib.mithra_batch_seed(self.h, i, C.byref(seed)))

    def set_tiles(self, variant, use_graph=True):
        self._check(self.lib.mithra_batch_set_tiles(self.h, variant, 1 if use_graph else 0))

    def autotune(self, reps=2):
        chosen = C.c_int()
        ms = (C.c_float * TILE_VARIANTS)()
        self._check(self.lib.mithra_batch_autotune(self.h, reps, C.byref(chosen), ms))
        return chosen.value, list(ms)

    def run(self, steps):
        done = C.c_long()
        self._check(self.lib.mithra_batch_run(self.h, steps, C.byref(done)))
        return done.value

    def steps(self, i):
        s = C.c_long()
        self._check(self.lib.mithra_batch_steps(self.h, i, C.byref(s)))
        return s.value

    def bunching(self, i):
        out = np.zeros(max(1, self.slots[i]))
        self._check(self.lib.mithra_batch_bunching(self.h, i, _ptr(out)))
        return out[:self.slots[i]]

    def close(self):
        if self.h:
            self.lib.mithra_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bf16_bits(x):
    """float32 -> bfloat16 bit patterns (round to nearest even), as the device converts."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    r = ((u >> 16) & 1) + 0x7FFF
    return ((u + r) >> 16).astype(np.uint16)


def bf16_round(x):
    """float32 values rounded to bfloat16 and back."""
    return (bf16_bits(x).astype(np.uint32) << 16).view(np.float32)


class MLP:
    """Ensemble of multilayer perceptrons on the GPU (mithra_mlp_*, include/mithra_gpu.h)."""

    def __init__(self, dims, members=5, seed=1):
        self.lib = load_gpu_library()
        vp, fp = C.c_void_p, C.POINTER(C.c_float)
        for name, res, args in (
            ("mithra_mlp_create", C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int), C.c_ulonglong, C.POINTER(vp)]),
            ("mithra_mlp_destroy", C.c_int, [vp]),
            ("mithra_mlp_param_count", C.c_long, [vp]),
            ("mithra_mlp_set_params", C.c_int, [vp, C.c_int, fp]),
            ("mithra_mlp_get_params", C.c_int, [vp, C.c_int, fp]),
            ("mithra_mlp_forward", C.c_int, [vp, fp, C.c_long, fp]),
            ("mithra_mlp_gradients", C.c_int, [vp, C.c_int, fp, fp, C.c_long, fp, fp]),
            ("mithra_mlp_train_step", C.c_int, [vp, fp, fp, C.c_long, C.c_float, fp]),
            ("mithra_mlp_set_data", C.c_int, [vp, fp, fp, C.c_long]),
            ("mithra_mlp_train_epochs", C.c_int, [vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_ulonglong, fp]),
        ):
            f = getattr(self.lib, name)
            f.restype, f.argtypes = res, args
        self.dims, self.members = list(dims), members
        d = (C.c_int * len(dims))(*dims)
        self.h = vp()
        self._check(self.lib.mithra_mlp_create(members, len(dims) - 1, d, seed, C.byref(self.h)))
        self.n_params = self.lib.mithra_mlp_param_count(self.h)

    def _check(self, rc):
        if rc != 0:
            raise MithraError(f"status {rc}: {self.lib.mithra_strerror(rc).decode()}")

    @staticmethod
    def _f(a):
        return np.ascontiguousarray(a, dtype=np.float32)

    def params(self, member=-1):
        n = self.n_params * (self.members if member < 0 else 1)
        out = np.empty(n, np.float32)
        self._check(self.lib.mithra_mlp_get_params(self.h, member, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, p, member=-1):
        p = self._f(p)
        self._check(self.lib.mithra_mlp_set_params(self.h, member, p.ctypes.data_as(C.POINTER(C.c_float))))

    def forward(self, x):
        x = self._f(x)
        n = x.shape[0]
        y = np.empty((self.members, n, self.dims[-1]), np.float32)
        fp = C.POINTER(C.c_float)
        self._check(self.lib.mithra_mlp_forward(self.h, x.ctypes.data_as(fp), n, y.ctypes.data_as(fp)))
        return y

    def gradients(self, member, x, t):
        x, t = self._f(x), self._f(t)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        fp = C.POINTER(C.c_float)
        self._check(self.lib.mithra_mlp_gradients(self.h, member, x.ctypes.data_as(fp), t.ctypes.data_as(fp),
                                                  x.shape[0], g.ctypes.data_as(fp), C.byref(loss)))
        return g, loss.value

    def train_step(self, x, t, lr):
        x, t = self._f(x), self._f(t)
        loss = np.empty(self.members, np.float32)
        fp = C.POINTER(C.c_float)
        self._check(self.lib.mithra_mlp_train_step(self.h, x.ctypes.data_as(fp), t.ctypes.data_as(fp), x.shape[0],
                                                   lr, loss.ctypes.data_as(fp)))
        return loss

    def set_data(self, x, t):
        """The training set, kept on the device for train_epochs."""
        x, t = self._f(x), self._f(t)
        fp = C.POINTER(C.c_float)
        self._check(self.lib.mithra_mlp_set_data(self.h, x.ctypes.data_as(fp), t.ctypes.data_as(fp), x.shape[0]))

    def train_epochs(self, epochs, batch, lr_max, lr_min=1e-6, seed=0):
        """Adam over the resident set, one captured graph per epoch; returns each epoch's mean loss."""
        loss = np.empty(epochs, np.float32)
        self._check(self.lib.mithra_mlp_train_epochs(self.h, epochs, batch, lr_max, lr_min, seed,
                                                     loss.ctypes.data_as(C.POINTER(C.c_float))))
        return loss

    def close(self):
        if self.h:
            self.lib.mithra_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gemm_bf16(A, B, transpose_a=False, transpose_b=False, reps=1, tile=0):
    """C = op(A) op(B) through the surrogate's MFMA GEMM (inputs rounded to BF16; tile 0 by size,
    1 for 64 x 64 workgroups, 2 for 128 x 128); returns (C, mean ms per launch)."""
    lib = load_gpu_library()
    f = lib.mithra_mlp_gemm
    u16p, fp = C.POINTER(C.c_ushort), C.POINTER(C.c_float)
    f.restype = C.c_int
    f.argtypes = [u16p, C.c_long, C.c_long, C.c_long, u16p, C.c_long, C.c_long, C.c_long, fp, C.c_int, C.c_int,
                  C.c_int, C.c_int, C.c_int, fp]
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    M, K = (A.shape[1], A.shape[0]) if transpose_a else A.shape
    K2, N = (B.shape[1], B.shape[0]) if transpose_b else B.shape
    assert K == K2
    sam, sak = (1, A.shape[1]) if transpose_a else (A.shape[1], 1)
    sbk, sbn = (1, B.shape[1]) if transpose_b else (B.shape[1], 1)
    a16, b16 = bf16_bits(A), bf16_bits(B)
    out = np.empty((M, N), np.float32)
    ms = C.c_float()
    rc = f(a16.ctypes.data_as(u16p), sam, sak, a16.size, b16.ctypes.data_as(u16p), sbk, sbn, b16.size,
           out.ctypes.data_as(fp), M, N, K, tile, reps, C.byref(ms))
    if rc != 0:
        raise MithraError(f"mithra_mlp_gemm: status {rc}")
    return out, ms.value
