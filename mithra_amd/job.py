"""Python front end of the drop-in host (libmithra.so, include/mithra_host.h).

`Job(path)` parses a reference job file and runs the reference's set-up; the views return the
state in the device C-ABI's types, and `run()` marches on the GPU writing the reference's output
files.  Relative paths inside the job file resolve against the current directory, as for the
reference's `mithra job.job`.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HOST = None


def load_host_library(path=abi.HOST_LIB):
    global _HOST
    if _HOST is None:
        if not os.path.exists(path):
            raise abi.MithraError(f"{path} is missing: run __graft_entry__.build() first")
        lib = C.CDLL(path)
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        for name, res, args in (
            ("mithra_job_open", C.c_int, [C.c_char_p, C.POINTER(vp)]),
            ("mithra_job_close", C.c_int, [vp]),
            ("mithra_job_error", C.c_char_p, [vp]),
            ("mithra_job_set_external_bunch", C.c_int, [vp, dp, C.c_long]),
            ("mithra_job_initialize", C.c_int, [vp]),
            ("mithra_job_grid", C.c_int, [vp, C.POINTER(abi.Grid)]),
            ("mithra_job_coeffs", C.c_int, [vp, C.POINTER(abi.Coeffs)]),
            ("mithra_job_seed", C.c_int, [vp, C.POINTER(abi.Seed)]),
            ("mithra_job_bunch_params", C.c_int, [vp, C.POINTER(abi.BunchParams)]),
            ("mithra_job_undulators", C.c_int, [vp, C.POINTER(abi.Undulator), C.c_int]),
            ("mithra_job_times", C.c_int, [vp, dp]),
            ("mithra_job_power", C.c_int, [vp, dp, C.c_int]),
            ("mithra_job_particle_count", C.c_long, [vp]),
            ("mithra_job_particles", C.c_int, [vp, dp, C.c_long]),
            ("mithra_job_potentials", C.c_int, [vp, C.c_int, dp, dp]),
            ("mithra_job_run", C.c_long, [vp, C.c_long, C.c_int, C.c_int]),
            ("mithra_job_fill_domain", C.c_int, [vp, vp]),
            ("mithra_job_device_bunch", C.c_int, [vp, C.c_int]),
            ("mithra_job_set_external_device_bunch", C.c_int, [vp, vp]),
        ):
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args
        _HOST = lib
    return _HOST


class Job:
    """device_bunch: generate ellipsoid bunches and run the set-up's particle loops on the GPU.
    external_bunch: the records of a `type = other` bunch, an (n, 7) array of lab-frame q x y z
    gbx gby gbz on the host, or an abi.DeviceBunch already on the device (which the set-up boosts in
    place; it must outlive the set-up and fill())."""

    def __init__(self, path, external_bunch=None, initialize=True, device_bunch=False):
        self.lib = load_host_library()
        self.h = C.c_void_p()
        rc = self.lib.mithra_job_open(str(path).encode(), C.byref(self.h))
        self._check(rc)
        self._device_bunch = None
        if isinstance(external_bunch, abi.DeviceBunch):
            self._device_bunch = external_bunch
            self._check(self.lib.mithra_job_set_external_device_bunch(self.h, external_bunch.h))
        elif device_bunch:
            self._check(self.lib.mithra_job_device_bunch(self.h, 1))
        if external_bunch is not None and not isinstance(external_bunch, abi.DeviceBunch):
            b = np.ascontiguousarray(external_bunch, dtype=np.float64).reshape(-1, 7)
            self._check(self.lib.mithra_job_set_external_bunch(
                self.h, b.ctypes.data_as(C.POINTER(C.c_double)), b.shape[0]))
        if initialize:
            self._check(self.lib.mithra_job_initialize(self.h))

    def _check(self, rc):
        if rc < 0:
            raise abi.MithraError(self.lib.mithra_job_error(self.h).decode())
        return rc

    def close(self):
        if self.h:
            self.lib.mithra_job_close(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _view(self, fn, cls):
        out = cls()
        self._check(getattr(self.lib, fn)(self.h, C.byref(out)))
        return out

    def grid(self):
        return self._view("mithra_job_grid", abi.Grid)

    def coeffs(self):
        return self._view("mithra_job_coeffs", abi.Coeffs)

    def seed(self):
        return self._view("mithra_job_seed", abi.Seed)

    def bunch_params(self):
        return self._view("mithra_job_bunch_params", abi.BunchParams)

    def undulators(self):
        arr = (abi.Undulator * 256)()
        n = self._check(self.lib.mithra_job_undulators(self.h, arr, 256))
        return list(arr[:n])

    def times(self):
        """dict of time, time_bunch, bunch_time_step, dt, total_time, n_update, n_sub."""
        t = np.zeros(6)
        n = self._check(self.lib.mithra_job_times(self.h, t.ctypes.data_as(C.POINTER(C.c_double))))
        keys = ("time", "time_bunch", "bunch_time_step", "dt", "total_time", "n_update")
        d = dict(zip(keys, t.tolist()))
        d["n_sub"] = n
        return d

    def power(self):
        """(N, Nl, Nf, pc, z, w) of the radiation-power sampler, or None."""
        buf = np.zeros(4 + 4096)
        has = self._check(self.lib.mithra_job_power(self.h, buf.ctypes.data_as(C.POINTER(C.c_double)), buf.size))
        if not has:
            return None
        N, Nl, Nf = (int(v) for v in buf[:3])
        return N, Nl, Nf, float(buf[3]), buf[4:4 + N].copy(), buf[4 + N:4 + N + Nl].copy()

    def particles(self):
        n = self.lib.mithra_job_particle_count(self.h)
        q = np.empty((n, 11))
        self._check(self.lib.mithra_job_particles(self.h, q.ctypes.data_as(C.POINTER(C.c_double)), n))
        return q

    def potentials(self, level):
        g = self.grid()
        a = np.empty(3 * g.nodes)
        phi = np.empty(g.nodes) if g.space_charge else None
        dp = C.POINTER(C.c_double)
        self._check(self.lib.mithra_job_potentials(self.h, level, a.ctypes.data_as(dp),
                                                   None if phi is None else phi.ctypes.data_as(dp)))
        return a, phi

    def fill(self, dom):
        """Writes the pre-step state into a device domain (abi.Domain of the GPU library): the
        potentials when a seed set them, the particles (device to device when generated there)."""
        self._check(self.lib.mithra_job_fill_domain(self.h, dom.h))

    def run(self, max_steps=-1, precision=abi.PRECISION_FP64, quiet=True):
        return self._check(self.lib.mithra_job_run(self.h, max_steps, precision, 1 if quiet else 0))
