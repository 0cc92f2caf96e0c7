"""ctypes front end of the CPU oracle (libppfv_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path. Only the tests, the
smoke entry point and the benchmark's CPU-baseline leg may import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libppfv_oracle.so")

_dp = C.POINTER(C.c_double)


class _Problem(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("xe", _dp), ("ye", _dp),
                ("G", _dp), ("Dxx", _dp), ("Dxy", _dp), ("Dyy", _dp), ("inv_tau", _dp),
                ("bc_type", C.c_int * 4), ("bc_val", _dp * 4), ("dt", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        for name, args in [
            ("orc_vertex", [C.c_void_p, _dp, _dp]),
            ("orc_assemble", [C.c_void_p, _dp, _dp, _dp, _dp]),
            ("orc_a_sigma", [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]),
            ("orc_band_solve", [C.c_int, C.c_int, _dp, _dp, _dp]),
            ("orc_matvec", [C.c_int, C.c_int, _dp, _dp, _dp]),
            ("orc_step", [C.c_void_p, _dp, _dp]),
            ("orc_run", [C.c_void_p, _dp, C.c_int]),
        ]:
            fn = getattr(_lib, name)
            fn.argtypes = args
        for name in ("orc_band_solve", "orc_step", "orc_run"):
            getattr(_lib, name).restype = C.c_int
        for name in ("orc_vertex", "orc_assemble", "orc_a_sigma", "orc_matvec"):
            getattr(_lib, name).restype = None
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


class OracleProblem:
    """Holds contiguous copies of a Problem's arrays and the C struct over them."""

    def __init__(self, prob):
        self.nx, self.ny = prob.nx, prob.ny
        self._keep = []

        def keep(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self._keep.append(a)
            return a
        s = _Problem()
        s.nx, s.ny = prob.nx, prob.ny
        s.xe, s.ye = _p(keep(prob.xe)), _p(keep(prob.ye))
        s.G, s.Dxx, s.Dxy = _p(keep(prob.G)), _p(keep(prob.Dxx)), _p(keep(prob.Dxy))
        s.Dyy, s.inv_tau = _p(keep(prob.Dyy)), _p(keep(prob.inv_tau))
        for k in range(4):
            s.bc_type[k] = int(prob.bc_type[k])
            s.bc_val[k] = _p(keep(prob.bc_val[k]))
        s.dt = float(prob.dt)
        self.s = s
        self.ref = C.byref(s)


def vertex(op: OracleProblem, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    V = np.empty((op.nx + 1, op.ny + 1))
    lib().orc_vertex(op.ref, _p(f), _p(V))
    return V


def assemble(op: OracleProblem, f, V=None):
    f = np.ascontiguousarray(f, dtype=np.float64)
    if V is None:
        V = vertex(op, f)
    N = op.nx * op.ny
    bands = np.empty((5, N))
    R = np.empty(N)
    lib().orc_assemble(op.ref, _p(f), _p(np.ascontiguousarray(V)), _p(bands), _p(R))
    return bands, R


def a_sigma(op: OracleProblem, V, i, j, inbr):
    a, b, s = C.c_double(), C.c_double(), C.c_double()
    lib().orc_a_sigma(op.ref, _p(np.ascontiguousarray(V)), i, j, inbr,
                      C.byref(a), C.byref(b), C.byref(s))
    return a.value, b.value, s.value


def band_solve(nx, ny, bands, R):
    x = np.empty(nx * ny)
    rc = lib().orc_band_solve(nx, ny, _p(np.ascontiguousarray(bands)), _p(np.ascontiguousarray(R)), _p(x))
    if rc:
        raise RuntimeError(f"orc_band_solve failed: {rc}")
    return x


def matvec(nx, ny, bands, x):
    y = np.empty(nx * ny)
    lib().orc_matvec(nx, ny, _p(np.ascontiguousarray(bands)), _p(np.ascontiguousarray(x)), _p(y))
    return y


def step(op: OracleProblem, f):
    out = np.array(f, dtype=np.float64, copy=True)
    rc = lib().orc_step(op.ref, _p(out), _p(out))
    if rc:
        raise RuntimeError(f"orc_step failed: {rc}")
    return out


def run(op: OracleProblem, f, nsteps):
    out = np.array(f, dtype=np.float64, copy=True)
    rc = lib().orc_run(op.ref, _p(out), int(nsteps))
    if rc:
        raise RuntimeError(f"orc_run failed: {rc}")
    return out


def run_schedule(problems, f):
    """Steps of a time-dependent Equation: step n uses problems[n], the inputs
    the Equation holds after update(t_n) (G, D, 1/tau and the Dirichlet vertex
    values at t_n; source/Solver.cc:286-289 refreshes Lambda and the vertex
    table after each solve, for the next step). Returns the state after each
    step."""
    out = []
    f = np.array(f, dtype=np.float64, copy=True)
    for p in problems:
        f = step(OracleProblem(p), f)
        out.append(f)
    return out
