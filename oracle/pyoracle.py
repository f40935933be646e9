"""ctypes front end of the CPU oracle (oracle/swept1d_oracle.c).

TEST INFRASTRUCTURE.  Imported only by tests/, by the smoke check in
__graft_entry__.py and by the CPU-baseline leg of bench.py, always as the
checker.  Nothing in the HIP library imports or links it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SW1D_ORACLE_LIB selects another build of the same library (the sanitizer
# build of the Makefile's `sanitize` target)
LIB_PATH = os.environ.get("SW1D_ORACLE_LIB") or os.path.join(_HERE, "_build", "libswept1d_oracle.so")

_lib = None


def build() -> None:
    """Compile the oracle with its Makefile (output under oracle/_build/)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    """Load the oracle library, building it only when nothing is built yet."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_P = C.c_void_p
_I64 = C.c_int64
_INT = C.c_int
_D = C.c_double
_U64 = C.c_uint64


def _declare(L):
    for s, real in (("_d", C.c_double), ("_f", C.c_float)):
        getattr(L, "orc_heat_step" + s).argtypes = [_P, _P, _I64, real, real, _INT, _INT]
        getattr(L, "orc_heat_run" + s).argtypes = [_P, _I64, _I64, real, real, _INT, _INT]
        getattr(L, "orc_ks_stage" + s).argtypes = [_P, _P, _I64, _P, _INT, _INT]
        getattr(L, "orc_ks_run" + s).argtypes = [_P, _I64, _I64, _P, _INT]
        getattr(L, "orc_euler_stage" + s).argtypes = [_P, _P, _I64, _P, _P, _INT, _INT]
        getattr(L, "orc_euler_run" + s).argtypes = [_P, _I64, _I64, _P, _P, _INT]
        getattr(L, "orc_euler_outputs" + s).argtypes = [_P, _I64, real, _P, _P, _P, _P]
        getattr(L, "orc_heat_consts" + s).argtypes = [_D, _D, _D, _P]
        getattr(L, "orc_ks_consts" + s).argtypes = [_D, _D, _P]
        getattr(L, "orc_euler_consts" + s).argtypes = [_D, _I64, _P, _P]
        getattr(L, "orc_heat_ic_ref" + s).argtypes = [_P, _I64, _D]
        getattr(L, "orc_ks_ic_ref" + s).argtypes = [_P, _I64, _D]
        getattr(L, "orc_euler_ic_ref" + s).argtypes = [_P, _I64, _P]
    L.orc_ic_heat_fourier.argtypes = [_P, _I64, _D, _U64, _P]
    L.orc_ic_ks_noisy.argtypes = [_P, _I64, _D, _D, _U64]
    L.orc_splitmix64_next.argtypes = [_P]
    L.orc_splitmix64_next.restype = _U64


def _sfx(dtype):
    return "_d" if np.dtype(dtype) == np.float64 else "_f"


def _ptr(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(_P)


# ------------------------------------------------------------- constants ---

def heat_consts(dt, dtype=np.float64, alpha=8.418e-5, ds=0.001):
    out = np.zeros(2, dtype)
    getattr(lib(), "orc_heat_consts" + _sfx(dtype))(dt, alpha, ds, _ptr(out))
    return out


def ks_consts(dt, dtype=np.float64, dx=0.5):
    out = np.zeros(5, dtype)
    getattr(lib(), "orc_ks_consts" + _sfx(dtype))(dt, dx, _ptr(out))
    return out


def euler_consts(dt, n, dtype=np.float64):
    k = np.zeros(3, dtype)
    bd = np.zeros(6, dtype)
    getattr(lib(), "orc_euler_consts" + _sfx(dtype))(dt, n, _ptr(k), _ptr(bd))
    return k, bd


# ------------------------------------------------------------------- ICs ---

def heat_ic_ref(n, dtype=np.float64, ds=0.001):
    u = np.zeros(n, dtype)
    getattr(lib(), "orc_heat_ic_ref" + _sfx(dtype))(_ptr(u), n, ds)
    return u


def ks_ic_ref(n, dtype=np.float64, dx=0.5):
    u = np.zeros(n, dtype)
    getattr(lib(), "orc_ks_ic_ref" + _sfx(dtype))(_ptr(u), n, dx)
    return u


def euler_ic_ref(n, dtype=np.float64, dt=1e-5):
    _, bd = euler_consts(dt, n, dtype)
    q = np.zeros(3 * n, dtype)
    getattr(lib(), "orc_euler_ic_ref" + _sfx(dtype))(_ptr(q), n, _ptr(bd))
    return q


def ic_heat_fourier(n, seed, L=1.0):
    u = np.zeros(n, np.float64)
    amp = np.zeros(4, np.float64)
    lib().orc_ic_heat_fourier(_ptr(u), n, L, seed, _ptr(amp))
    return u, amp


def ic_ks_noisy(n, seed, dx=0.5, sigma=1e-3):
    u = np.zeros(n, np.float64)
    lib().orc_ic_ks_noisy(_ptr(u), n, dx, sigma, seed)
    return u


# ---------------------------------------------------------------- stages ---

def heat_step(u, fo, fo2, bc=0, nthreads=1):
    out = np.empty_like(u)
    getattr(lib(), "orc_heat_step" + _sfx(u.dtype))(_ptr(u), _ptr(out), u.size, fo, fo2, bc, nthreads)
    return out


def heat_run(u, nsteps, fo, fo2, bc=0, nthreads=1):
    u = np.ascontiguousarray(u).copy()
    getattr(lib(), "orc_heat_run" + _sfx(u.dtype))(_ptr(u), u.size, nsteps, fo, fo2, bc, nthreads)
    return u


def ks_stage(u_in, u_out, coef, final, nthreads=1):
    coef = np.ascontiguousarray(coef, u_in.dtype)
    getattr(lib(), "orc_ks_stage" + _sfx(u_in.dtype))(_ptr(u_in), _ptr(u_out), u_in.size,
                                                       _ptr(coef), int(final), nthreads)


def ks_run(u, nsteps, coef, nthreads=1):
    u = np.ascontiguousarray(u).copy()
    coef = np.ascontiguousarray(coef, u.dtype)
    getattr(lib(), "orc_ks_run" + _sfx(u.dtype))(_ptr(u), u.size, nsteps, _ptr(coef), nthreads)
    return u


def euler_run(q, nsteps, k, bd, nthreads=1):
    """q: interleaved (rho, rho*u, E) per cell, length 3n."""
    q = np.ascontiguousarray(q).copy()
    k = np.ascontiguousarray(k, q.dtype)
    bd = np.ascontiguousarray(bd, q.dtype)
    getattr(lib(), "orc_euler_run" + _sfx(q.dtype))(_ptr(q), q.size // 3, nsteps, _ptr(k), _ptr(bd),
                                                     nthreads)
    return q


def euler_stage(q_in, q_out, k, bd, final, nthreads=1):
    k = np.ascontiguousarray(k, q_in.dtype)
    bd = np.ascontiguousarray(bd, q_in.dtype)
    getattr(lib(), "orc_euler_stage" + _sfx(q_in.dtype))(_ptr(q_in), _ptr(q_out), q_in.size // 3,
                                                          _ptr(k), _ptr(bd), int(final), nthreads)


def euler_outputs(q, mgam):
    n = q.size // 3
    outs = [np.zeros(n, q.dtype) for _ in range(4)]
    getattr(lib(), "orc_euler_outputs" + _sfx(q.dtype))(_ptr(q), n, mgam, *[_ptr(o) for o in outs])
    return outs  # rho, velocity, energy, pressure


def splitmix64_stream(seed, count):
    s = C.c_uint64(seed)
    return [lib().orc_splitmix64_next(C.byref(s)) for _ in range(count)]
