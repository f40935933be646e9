"""1dsweptcuda_amd — MI355X-native 1D swept-rule solver (heat, KS, Euler).

Python front end over the C ABI of ``lib/libswept1d.so`` (declared in
``include/swept1d.h``).  The compute path is the HIP library; this module only
moves numpy arrays across the boundary and raises on every error code.  There
is no CPU fallback: if the library is not built, or a GPU call fails, the call
raises.

The package directory name starts with a digit, so import it with
``importlib.import_module("1dsweptcuda_amd")``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libswept1d.so")
BIN_DIR = os.path.join(_HERE, "bin")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "swept1d.h")

HEAT, KS, EULER = 0, 1, 2
F64, F32 = 0, 1
CLASSIC, SWEPT, ALT, TRAPEZOID = 0, 1, 2, 3
BC_DEFAULT, BC_MIRROR, BC_PERIODIC, BC_DIRICHLET = 0, 1, 2, 3

# sw1d_schedule kinds (Heat / KS / Euler classic and swept bookkeeping).
SCHED_HEAT_CLASSIC, SCHED_HEAT_SWEPT = 0, 1
SCHED_KS_CLASSIC, SCHED_KS_SWEPT, SCHED_KS_REGISTER = 2, 3, 4
SCHED_EULER_CLASSIC, SCHED_EULER_SWEPT = 5, 6

# sw1d_kernel_stats kinds
K_CLASSIC, K_UP, K_DOWN, K_DIAMOND, K_SPLIT = 0, 1, 2, 3, 4


class SweptError(RuntimeError):
    """A libswept1d call returned a negative status."""

    def __init__(self, code: int, where: str, message: str):
        super().__init__(f"{where} failed ({code}): {message}")
        self.code = code


class Problem(C.Structure):
    _fields_ = [
        ("equation", C.c_int32),
        ("precision", C.c_int32),
        ("algorithm", C.c_int32),
        ("bc", C.c_int32),
        ("nx", C.c_int64),
        ("tpb", C.c_int32),
        ("node_width", C.c_int32),
        ("dt", C.c_double),
        ("dx", C.c_double),
        ("alpha", C.c_double),
        ("tile_depth", C.c_int32),
        ("reserved", C.c_int32),
    ]


_lib = None


def build() -> None:
    """Compile libswept1d.so and the six programs in-tree (gfx950)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-j8"], check=True)


def lib():
    """Load libswept1d.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SweptError(-5, "load", f"{LIB_PATH} is not built (run build())")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.sw1d_abi_version.restype = C.c_int
        L.sw1d_last_error.restype = C.c_char_p
        L.sw1d_device_check.restype = C.c_int
        L.sw1d_plan_create.argtypes = [C.POINTER(P), C.POINTER(Problem)]
        L.sw1d_plan_destroy.argtypes = [P]
        L.sw1d_set_state.argtypes = [P, P]
        L.sw1d_get_state.argtypes = [P, P]
        L.sw1d_set_state_device.argtypes = [P, P]
        L.sw1d_get_state_device.argtypes = [P, P]
        L.sw1d_advance.argtypes = [P, C.c_int64]
        L.sw1d_synchronize.argtypes = [P]
        L.sw1d_advance_timed.argtypes = [P, C.c_int64, C.POINTER(C.c_double)]
        L.sw1d_plan_info.argtypes = [P, C.POINTER(C.c_int64)]
        L.sw1d_plan_constants.argtypes = [P, C.POINTER(C.c_double), C.c_int32]
        L.sw1d_plan_stream.argtypes = [P, C.POINTER(P)]
        L.sw1d_set_profiling.argtypes = [P, C.c_int32]
        L.sw1d_kernel_stats.argtypes = [P, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.sw1d_schedule.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32,
                                    C.c_double, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                    C.c_int64, C.POINTER(C.c_int64)]
        L.sw1d_initial_condition.argtypes = [C.c_int32, C.c_int32, C.c_int64, C.c_double, C.c_double,
                                             C.c_uint64, P]
        _lib = L
    return _lib


def _check(rc: int, where: str) -> int:
    if rc < 0:
        raise SweptError(rc, where, lib().sw1d_last_error().decode())
    return rc


def last_error() -> str:
    return lib().sw1d_last_error().decode()


def device_check() -> bool:
    return _check(lib().sw1d_device_check(), "sw1d_device_check") == 1


def schedule(kind: int, dt: float, tf_arg: float, freq: float, tpb: int,
             dt_sched: float | None = None):
    """Rows the reference program writes: (steps computed, time printed)."""
    if dt_sched is None:
        dt_sched = dt
    count = C.c_int64(0)
    L = lib()
    L.sw1d_schedule(kind, dt, tf_arg, freq, tpb, dt_sched, None, None, 0, C.byref(count))
    steps = (C.c_int64 * count.value)()
    times = (C.c_double * count.value)()
    _check(L.sw1d_schedule(kind, dt, tf_arg, freq, tpb, dt_sched, steps, times, count.value,
                           C.byref(count)), "sw1d_schedule")
    return list(steps), list(times)


IC_KS_NOISY, IC_HEAT_FOURIER = 0, 1


def initial_condition(kind: int, n: int, length: float, param: float, seed: int,
                      precision: int = F64) -> np.ndarray:
    """Seeded synthetic initial condition drawn by the library (splitmix64 on
    the host; sw1d_initial_condition)."""
    out = np.empty(n, _dtype(precision))
    _check(lib().sw1d_initial_condition(kind, precision, n, length, param, seed,
                                        out.ctypes.data_as(C.c_void_p)), "sw1d_initial_condition")
    return out


def ks_noisy_ic(n: int, seed: int, dx: float = 0.5, sigma: float = 1e-3, precision: int = F64):
    """2cos(19 pi x/128) + sigma N(0,1) at x = k dx (the KS benchmark workload)."""
    return initial_condition(IC_KS_NOISY, n, dx, sigma, seed, precision)


def heat_fourier_ic(n: int, seed: int, L: float = 1.0, precision: int = F64):
    """1 + sum_{k=1..4} a_k sin(2 pi k x/L), a_k ~ U[0,1) (the periodic heat workload)."""
    return initial_condition(IC_HEAT_FOURIER, n, L, 0.0, seed, precision)


def _dtype(precision: int):
    return np.float64 if precision == F64 else np.float32


class Plan:
    """One solver instance on the current GPU (owns device state + a stream).

    node_width: 0 takes the reference's tpb; "auto" takes the autotuner's
    width for this equation, precision and size (tuned/node_widths.json),
    or tpb when that table has no entry."""

    def __init__(self, equation: int, nx: int, dt: float, *, precision: int = F64,
                 algorithm: int = SWEPT, tpb: int = 64, node_width=0,
                 bc: int = BC_DEFAULT, dx: float = 0.0, alpha: float = 0.0, tile_depth: int = 0):
        if isinstance(node_width, str):
            if node_width != "auto":
                raise ValueError(f"node_width must be an int or 'auto', got {node_width!r}")
            from . import autotune
            node_width = autotune.node_width_for(equation, nx, precision)
        self.prob = Problem(equation, precision, algorithm, bc, nx, tpb, node_width, dt, dx, alpha,
                            tile_depth, 0)
        self.dtype = _dtype(precision)
        self.equation = equation
        self.nx = nx
        self._h = C.c_void_p()
        _check(lib().sw1d_plan_create(C.byref(self._h), C.byref(self.prob)), "sw1d_plan_create")

    @property
    def state_len(self) -> int:
        return self.nx * (3 if self.equation == EULER else 1)

    def close(self):
        if self._h:
            lib().sw1d_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        if a.size != self.state_len:
            raise ValueError(f"state must hold {self.state_len} values, got {a.size}")
        _check(lib().sw1d_set_state(self._h, a.ctypes.data_as(C.c_void_p)), "sw1d_set_state")

    def get_state(self) -> np.ndarray:
        out = np.empty(self.state_len, self.dtype)
        _check(lib().sw1d_get_state(self._h, out.ctypes.data_as(C.c_void_p)), "sw1d_get_state")
        return out

    def set_state_device(self, ptr: int) -> None:
        _check(lib().sw1d_set_state_device(self._h, C.c_void_p(ptr)), "sw1d_set_state_device")

    def get_state_device(self, ptr: int) -> None:
        _check(lib().sw1d_get_state_device(self._h, C.c_void_p(ptr)), "sw1d_get_state_device")

    def advance(self, nsteps: int) -> None:
        _check(lib().sw1d_advance(self._h, nsteps), "sw1d_advance")

    def synchronize(self) -> None:
        _check(lib().sw1d_synchronize(self._h), "sw1d_synchronize")

    def advance_timed(self, nsteps: int) -> float:
        ms = C.c_double(0)
        _check(lib().sw1d_advance_timed(self._h, nsteps, C.byref(ms)), "sw1d_advance_timed")
        return ms.value

    def info(self) -> dict:
        v = (C.c_int64 * 5)()
        _check(lib().sw1d_plan_info(self._h, v), "sw1d_plan_info")
        return {"node_width": v[0], "tile_steps": v[1], "values_per_point": v[2],
                "device_bytes": v[3], "cycle_steps": v[4]}

    def constants(self) -> list:
        v = (C.c_double * 16)()
        k = _check(lib().sw1d_plan_constants(self._h, v, 16), "sw1d_plan_constants")
        return list(v[:k])

    def stream(self) -> int:
        s = C.c_void_p()
        _check(lib().sw1d_plan_stream(self._h, C.byref(s)), "sw1d_plan_stream")
        return s.value or 0

    def set_profiling(self, on: bool) -> None:
        _check(lib().sw1d_set_profiling(self._h, 1 if on else 0), "sw1d_set_profiling")

    def kernel_stats(self, kind: int):
        n = C.c_int64(0)
        ms = C.c_double(0)
        _check(lib().sw1d_kernel_stats(self._h, kind, C.byref(n), C.byref(ms)), "sw1d_kernel_stats")
        return n.value, ms.value


def program(equation: int, precision: int) -> str:
    """Path of the drop-in executable, e.g. bin/KSDoubleOut."""
    name = ["Heat", "KS", "Euler"][equation] + ("Double" if precision == F64 else "Single") + "Out"
    return os.path.join(BIN_DIR, name)


class Batch:
    """Many independent periodic heat / KS problems advanced by one persistent
    launch (sw1d_batch_*).  State is the concatenation of the fields."""

    def __init__(self, equations, nx, dt, dx, *, precision: int = F64, tile_points=None, waves=None):
        L = lib()
        P = C.c_void_p
        L.sw1d_batch_create.argtypes = [C.POINTER(P), C.c_int32, C.c_int64, P, P, P, P, P]
        L.sw1d_batch_set_workgroup.argtypes = [P, C.c_int32]
        L.sw1d_batch_destroy.argtypes = [P]
        L.sw1d_batch_set_state.argtypes = [P, P]
        L.sw1d_batch_get_state.argtypes = [P, P]
        L.sw1d_batch_advance.argtypes = [P, C.c_int64]
        L.sw1d_batch_advance_timed.argtypes = [P, C.c_int64, C.POINTER(C.c_double)]
        L.sw1d_batch_info.argtypes = [P, C.POINTER(C.c_int64)]
        self.eq = np.ascontiguousarray(equations, np.int32)
        self.nx = np.ascontiguousarray(nx, np.int64)
        self.dt = np.ascontiguousarray(dt, np.float64)
        self.dx = np.ascontiguousarray(dx, np.float64)
        self.tp = None if tile_points is None else np.ascontiguousarray(tile_points, np.int32)
        self.dtype = _dtype(precision)
        self.offsets = np.concatenate([[0], np.cumsum(self.nx)])
        self._h = C.c_void_p()
        ptr = lambda a: a.ctypes.data_as(P) if a is not None else None
        _check(L.sw1d_batch_create(C.byref(self._h), precision, len(self.nx), ptr(self.eq),
                                   ptr(self.nx), ptr(self.dt), ptr(self.dx), ptr(self.tp)),
               "sw1d_batch_create")
        if waves is not None:
            self.set_workgroup(waves)

    def set_workgroup(self, waves: int) -> None:
        """Waves per workgroup of the persistent launch (2, 4 or 8): the tiles a
        workgroup works on at once.  Results do not depend on it."""
        _check(lib().sw1d_batch_set_workgroup(self._h, int(waves)), "sw1d_batch_set_workgroup")

    def close(self):
        if self._h:
            lib().sw1d_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, packed) -> None:
        a = np.ascontiguousarray(packed, dtype=self.dtype)
        if a.size != self.offsets[-1]:
            raise ValueError("packed state has the wrong length")
        _check(lib().sw1d_batch_set_state(self._h, a.ctypes.data_as(C.c_void_p)), "sw1d_batch_set_state")

    def get_state(self) -> np.ndarray:
        out = np.empty(int(self.offsets[-1]), self.dtype)
        _check(lib().sw1d_batch_get_state(self._h, out.ctypes.data_as(C.c_void_p)), "sw1d_batch_get_state")
        return out

    def advance(self, nsteps: int) -> None:
        _check(lib().sw1d_batch_advance(self._h, nsteps), "sw1d_batch_advance")

    def advance_timed(self, nsteps: int) -> float:
        ms = C.c_double(0)
        _check(lib().sw1d_batch_advance_timed(self._h, nsteps, C.byref(ms)), "sw1d_batch_advance_timed")
        return ms.value

    def info(self) -> dict:
        v = (C.c_int64 * 4)()
        _check(lib().sw1d_batch_info(self._h, v), "sw1d_batch_info")
        return {"total_points": v[0], "workgroups": v[1], "problems": v[2], "device_bytes": v[3]}

    def split(self, packed):
        return [packed[self.offsets[i]:self.offsets[i + 1]] for i in range(len(self.nx))]


class Surrogate:
    """The learned KS surrogate (sw1d_nn_*): periodic 1D CNN, BF16 MFMA on the GPU."""

    def __init__(self, nx: int = 1024, max_batch: int = 64, seed: int = 0):
        L = lib()
        P = C.c_void_p
        L.sw1d_nn_last_error.restype = C.c_char_p
        L.sw1d_nn_create.argtypes = [C.POINTER(P), C.c_int32, C.c_int32, C.c_uint64]
        L.sw1d_nn_destroy.argtypes = [P]
        L.sw1d_nn_set_scales.argtypes = [P, C.c_double, C.c_double]
        L.sw1d_nn_set_inference.argtypes = [P, C.c_int32]
        L.sw1d_nn_get_params.argtypes = [P, P]
        L.sw1d_nn_set_params.argtypes = [P, P]
        L.sw1d_nn_forward.argtypes = [P, P, C.c_int32, P]
        L.sw1d_nn_gradient.argtypes = [P, P, P, C.c_int32, C.POINTER(C.c_double), P]
        L.sw1d_nn_train.argtypes = [P, P, P, C.c_int64, C.c_int32, C.c_int64, C.c_float, C.c_uint64,
                                    C.POINTER(C.c_double)]
        L.sw1d_surrogate_advance.argtypes = [P, P, C.c_int64, C.c_int32, C.c_int32, C.c_double,
                                             C.POINTER(C.c_double)]
        L.sw1d_nn_dataset.argtypes = [P, P, C.c_int32, C.c_int32, C.POINTER(C.c_int64),
                                      C.POINTER(C.c_double)]
        L.sw1d_nn_fit.argtypes = [P, C.c_int64, C.c_int32, C.c_int64, C.c_float, C.c_uint64,
                                  C.POINTER(C.c_double)]
        L.sw1d_nn_eval.argtypes = [P, C.c_int64, C.c_int64, C.POINTER(C.c_double)]
        L.sw1d_nn_dataset_samples.argtypes = [P, C.c_int64, C.c_int64, P, P]
        self.nx, self.max_batch = nx, max_batch
        self._h = C.c_void_p()
        self._chk(L.sw1d_nn_create(C.byref(self._h), nx, max_batch, seed), "sw1d_nn_create")

    @staticmethod
    def param_count() -> int:
        return lib().sw1d_nn_param_count()

    @staticmethod
    def device_check() -> bool:
        return lib().sw1d_nn_device_check() == 1

    def _chk(self, rc, where):
        if rc < 0:
            raise SweptError(rc, where, lib().sw1d_nn_last_error().decode())
        return rc

    def close(self):
        if self._h:
            lib().sw1d_nn_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scales(self, in_scale: float, out_scale: float) -> None:
        self._chk(lib().sw1d_nn_set_scales(self._h, in_scale, out_scale), "sw1d_nn_set_scales")

    # inference kernels (sw1d_nn_set_inference): all give the same bits
    INFER_AUTO, INFER_PER_LAYER, INFER_WHOLE = 0, 1, 2

    def set_inference(self, mode: int) -> None:
        self._chk(lib().sw1d_nn_set_inference(self._h, int(mode)), "sw1d_nn_set_inference")

    def get_params(self) -> np.ndarray:
        p = np.empty(self.param_count(), np.float32)
        self._chk(lib().sw1d_nn_get_params(self._h, p.ctypes.data_as(C.c_void_p)), "sw1d_nn_get_params")
        return p

    def set_params(self, p) -> None:
        a = np.ascontiguousarray(p, np.float32)
        self._chk(lib().sw1d_nn_set_params(self._h, a.ctypes.data_as(C.c_void_p)), "sw1d_nn_set_params")

    def forward(self, x) -> np.ndarray:
        a = np.ascontiguousarray(x, np.float32).reshape(-1, self.nx)
        y = np.empty_like(a)
        self._chk(lib().sw1d_nn_forward(self._h, a.ctypes.data_as(C.c_void_p), a.shape[0],
                                        y.ctypes.data_as(C.c_void_p)), "sw1d_nn_forward")
        return y

    def gradient(self, x, t):
        a = np.ascontiguousarray(x, np.float32).reshape(-1, self.nx)
        b = np.ascontiguousarray(t, np.float32).reshape(-1, self.nx)
        g = np.empty(self.param_count(), np.float32)
        loss = C.c_double(0)
        self._chk(lib().sw1d_nn_gradient(self._h, a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p),
                                         a.shape[0], C.byref(loss), g.ctypes.data_as(C.c_void_p)),
                  "sw1d_nn_gradient")
        return loss.value, g

    def train(self, d_inputs: int, d_targets: int, nsamples: int, batch: int, iters: int, lr: float,
              seed: int = 0) -> float:
        loss = C.c_double(0)
        self._chk(lib().sw1d_nn_train(self._h, C.c_void_p(d_inputs), C.c_void_p(d_targets), nsamples, batch,
                                      iters, lr, seed, C.byref(loss)), "sw1d_nn_train")
        return loss.value

    def dataset(self, batch: "Batch", nsnap: int, k_steps: int = 20):
        """Snapshots of a KS batch every k_steps as the training set; returns
        (number of samples, (in_scale, out_scale))."""
        n = C.c_int64(0)
        sc = (C.c_double * 2)()
        self._chk(lib().sw1d_nn_dataset(self._h, batch._h, nsnap, k_steps, C.byref(n), sc), "sw1d_nn_dataset")
        return n.value, (sc[0], sc[1])

    def dataset_samples(self, first: int, count: int):
        """Stored training pairs [first, first+count): (x, t), each [count, nx]
        float32; sample s*B + b is snapshot s of problem b."""
        x = np.empty((count, self.nx), np.float32)
        t = np.empty((count, self.nx), np.float32)
        self._chk(lib().sw1d_nn_dataset_samples(self._h, first, count, x.ctypes.data_as(C.c_void_p),
                                                t.ctypes.data_as(C.c_void_p)), "sw1d_nn_dataset_samples")
        return x, t

    def fit(self, iters: int, batch: int = 32, lr: float = 1e-3, nfit: int = 0, seed: int = 0) -> float:
        loss = C.c_double(0)
        self._chk(lib().sw1d_nn_fit(self._h, nfit, batch, iters, lr, seed, C.byref(loss)), "sw1d_nn_fit")
        return loss.value

    def eval(self, first: int, count: int) -> float:
        e = C.c_double(0)
        self._chk(lib().sw1d_nn_eval(self._h, first, count, C.byref(e)), "sw1d_nn_eval")
        return e.value

    def advance(self, batch: "Batch", surrogate_steps: int, k_steps: int = 20, judge_every: int = 10,
                tol: float = 1e-2) -> dict:
        st = (C.c_double * 4)()
        self._chk(lib().sw1d_surrogate_advance(self._h, batch._h, surrogate_steps, k_steps, judge_every, tol,
                                               st), "sw1d_surrogate_advance")
        return {"network_steps": int(st[0]), "judged": int(st[1]), "fallbacks": int(st[2]),
                "max_judged_error": st[3]}
