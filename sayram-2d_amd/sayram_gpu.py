"""ctypes front end of libsayram_gpu.so (C ABI in include/sayram_gpu.h).

This mirrors the reference's ``Solver`` surface (source/Solver.h:18-82):
``Solver(problem)`` uploads the grid, G, D, 1/tau, boundary data and f0,
``update()`` takes one implicit step (Solver::update, source/Solver.cc:270-290),
``t()`` and ``f()`` read the state. ``Batch`` runs many instances per step.

There is no CPU fallback: loading fails loudly when the HIP library is
missing, and every call raises ``SayramError`` on a non-zero status.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsayram_gpu.so")

SG_OK, SG_ERR_ARG, SG_ERR_HIP, SG_ERR_NOMEM, SG_ERR_NOCONV, SG_ERR_BC, SG_ERR_STATE = 0, -1, -2, -3, -4, -5, -6
SG_FP64, SG_MIXED, SG_PC32 = 0, 1, 2

_dp = C.POINTER(C.c_double)


class SgGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("x_edges", _dp), ("y_edges", _dp)]


class SgCoeffs(C.Structure):
    _fields_ = [("G", _dp), ("Dxx", _dp), ("Dxy", _dp), ("Dyy", _dp), ("inv_tau", _dp)]


class SgBC(C.Structure):
    _fields_ = [("type", C.c_int * 4), ("values", _dp * 4)]


class SgOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("rtol", C.c_double), ("max_iters", C.c_int),
                ("inner_rtol", C.c_double), ("max_inner", C.c_int), ("check_every", C.c_int),
                ("smoother", C.c_int), ("coarse_smoother", C.c_int), ("use_graph", C.c_int),
                ("line_tiles", C.c_int * 4),
                ("coarse_tiles", C.c_int * 4),
                ("fuse_cells", C.c_int), ("streams", C.c_int),
                ("group_cells", C.c_int), ("head_iters", C.c_int), ("no_clamp", C.c_int),
                ("warm_start", C.c_int)]


class SgStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("last_iters", C.c_int), ("last_refinements", C.c_int),
                ("last_rel_residual", C.c_double), ("min_f", C.c_double),
                ("clamp_count", C.c_int64), ("total_iters", C.c_int64), ("status", C.c_int),
                ("last_nonfinite", C.c_int), ("last_clamps", C.c_int64), ("last_min_x", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# every symbol include/sayram_gpu.h declares: (name, restype, argtypes)
_V, _S = C.c_void_p, C.c_size_t
SYMBOLS = [
    ("sg_abi_version", C.c_int, []),
    ("sg_last_error", C.c_char_p, []),
    ("sg_device_count", C.c_int, []),
    ("sg_create", C.c_int, [C.POINTER(SgGrid), C.POINTER(SgCoeffs), C.POINTER(SgBC), _dp,
                            C.c_double, C.POINTER(SgOpts), C.POINTER(_V)]),
    ("sg_set_bc_values", C.c_int, [_V, C.c_int, _dp]),
    ("sg_set_coeffs", C.c_int, [_V, C.POINTER(SgCoeffs)]),
    ("sg_step", C.c_int, [_V, C.c_int]),
    ("sg_get_f", C.c_int, [_V, _dp]),
    ("sg_set_f", C.c_int, [_V, _dp, C.c_int64]),
    ("sg_time", C.c_double, [_V]),
    ("sg_stats", C.c_int, [_V, C.POINTER(SgStats)]),
    ("sg_destroy", None, [_V]),
    ("sg_batch_create", C.c_int, [C.POINTER(SgOpts), _S, C.POINTER(_V)]),
    ("sg_instance_bytes", _S, [C.c_int, C.c_int, C.c_int]),
    ("sg_batch_add", C.c_int, [_V, C.POINTER(SgGrid), C.POINTER(SgCoeffs), C.POINTER(SgBC), _dp,
                               C.c_double, C.POINTER(C.c_int)]),
    ("sg_batch_remove", C.c_int, [_V, C.c_int]),
    ("sg_batch_step", C.c_int, [_V, C.c_int]),
    ("sg_batch_step_until", C.c_int, [_V, C.POINTER(C.c_int64), C.c_int, C.c_int]),
    ("sg_batch_get_f", C.c_int, [_V, C.c_int, _dp]),
    ("sg_batch_get_f_static", C.c_int, [_V, C.c_int, _dp]),
    ("sg_batch_set_f", C.c_int, [_V, C.c_int, _dp, C.c_int64]),
    ("sg_batch_set_bc_values", C.c_int, [_V, C.c_int, C.c_int, _dp]),
    ("sg_batch_set_coeffs", C.c_int, [_V, C.c_int, C.POINTER(SgCoeffs)]),
    ("sg_batch_stats", C.c_int, [_V, C.c_int, C.POINTER(SgStats)]),
    ("sg_batch_size", C.c_int, [_V]),
    ("sg_batch_bytes_in_use", _S, [_V]),
    ("sg_batch_destroy", None, [_V]),
    ("sg_debug_assemble", C.c_int, [_V, _dp, _dp]),
    ("sg_batch_time_kernel", C.c_int, [_V, C.c_int]),
    ("sg_batch_kernel_time", C.c_int, [_V, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                       C.POINTER(C.c_double)]),
    ("sg_batch_kernel_time_id", C.c_int, [_V, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                          C.POINTER(C.c_double)]),
    ("sg_batch_event_overhead", C.c_int, [_V, C.POINTER(C.c_double)]),
    ("sg_kernel_name", C.c_char_p, [C.c_int]),
    ("sg_kernel_count", C.c_int, []),
    ("sg_solver_batch", C.c_int, [_V, C.POINTER(_V)]),
]

KERNEL_IDS = {"assemble": 0, "rap": 1, "line_l0": 2, "line_coarse": 3, "restrict": 4,
              "prolong": 5, "spmv_v": 6, "spmv_t": 7, "vec_update": 8, "ir_residual": 9,
              "finalize": 10, "coarse_cycle": 11, "finish": 12}
TIME_ALL = len(KERNEL_IDS)   # sg_batch_time_kernel id that times every family


class SayramError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sayram_gpu error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Loads libsayram_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it first (python -c "
                              "'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SYMBOLS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != SG_OK:
        raise SayramError(rc, lib().sg_last_error().decode(errors="replace"))
    return rc


def _p(a):
    return a.ctypes.data_as(_dp)


def make_opts(precision=SG_FP64, rtol=0.0, max_iters=0, inner_rtol=0.0, max_inner=0,
              check_every=0, smoother=0, use_graph=1, line_tiles=(0, 0, 0, 0), fuse_cells=0,
              streams=0, group_cells=0, head_iters=0, no_clamp=0, warm_start=0, coarse_tiles=(0, 0, 0, 0),
              coarse_smoother=0):
    o = SgOpts()
    o.precision, o.rtol, o.max_iters = precision, rtol, max_iters
    o.inner_rtol, o.max_inner, o.check_every = inner_rtol, max_inner, check_every
    o.smoother, o.use_graph = smoother, use_graph
    o.coarse_smoother = int(coarse_smoother)
    for k in range(4):
        o.line_tiles[k] = int(line_tiles[k])
        o.coarse_tiles[k] = int(coarse_tiles[k])
    o.fuse_cells = int(fuse_cells)
    o.streams = int(streams)
    o.group_cells = int(group_cells)
    o.head_iters = int(head_iters)
    o.no_clamp = int(no_clamp)
    o.warm_start = int(warm_start)
    return o


def _coeff_struct(keep, G=None, Dxx=None, Dxy=None, Dyy=None, inv_tau=None):
    """SgCoeffs over contiguous copies (None -> NULL: that array is kept)."""
    def k(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return _p(a)
    return SgCoeffs(k(G), k(Dxx), k(Dxy), k(Dyy), k(inv_tau))


class _Inputs:
    """Contiguous FP64 copies of a problem's arrays plus the ABI structs."""

    def __init__(self, prob):
        self.keep = []

        def k(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self.keep.append(a)
            return _p(a)
        self.grid = SgGrid(prob.nx, prob.ny, k(prob.xe), k(prob.ye))
        self.coeffs = SgCoeffs(k(prob.G), k(prob.Dxx), k(prob.Dxy), k(prob.Dyy), k(prob.inv_tau))
        self.bc = SgBC()
        for s in range(4):
            self.bc.type[s] = int(prob.bc_type[s])
            self.bc.values[s] = k(prob.bc_val[s]) if prob.bc_val[s] is not None else None
        self.f0 = k(prob.f0)
        self.dt = float(prob.dt)


class Solver:
    """The reference's Solver surface over the C ABI."""

    def __init__(self, prob, opts=None):
        self.nx, self.ny = prob.nx, prob.ny
        inp = _Inputs(prob)
        h = C.c_void_p()
        _check(lib().sg_create(C.byref(inp.grid), C.byref(inp.coeffs), C.byref(inp.bc), inp.f0,
                               inp.dt, C.byref(opts or make_opts()), C.byref(h)))
        self.h = h

    def update(self, nsteps=1):
        return _check(lib().sg_step(self.h, int(nsteps)))

    def step_status(self, nsteps=1):
        return lib().sg_step(self.h, int(nsteps))

    def t(self):
        return lib().sg_time(self.h)

    def f(self):
        out = np.empty((self.nx, self.ny))
        _check(lib().sg_get_f(self.h, _p(out)))
        return out

    def set_f(self, f, istep):
        f = np.ascontiguousarray(f, dtype=np.float64)
        if f.shape != (self.nx, self.ny):
            raise ValueError(f"f has shape {f.shape}, the instance {(self.nx, self.ny)}")
        _check(lib().sg_set_f(self.h, _p(f), int(istep)))

    def set_bc_values(self, side, vals):
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        _check(lib().sg_set_bc_values(self.h, int(side), _p(vals)))

    def set_coeffs(self, G=None, Dxx=None, Dxy=None, Dyy=None, inv_tau=None):
        """New per-cell arrays, (nx, ny); None keeps an array's current values."""
        keep = []
        c = _coeff_struct(keep, G, Dxx, Dxy, Dyy, inv_tau)
        _check(lib().sg_set_coeffs(self.h, C.byref(c)))

    def stats(self):
        s = SgStats()
        _check(lib().sg_stats(self.h, C.byref(s)))
        return s.as_dict()

    def debug_assemble(self):
        N = self.nx * self.ny
        bands = np.empty((5, N))
        R = np.empty(N)
        _check(lib().sg_debug_assemble(self.h, _p(bands), _p(R)))
        return bands, R

    def close(self):
        if self.h:
            lib().sg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many independent instances advanced by the same launches."""

    def __init__(self, opts=None, hbm_budget=0):
        h = C.c_void_p()
        _check(lib().sg_batch_create(C.byref(opts or make_opts()), int(hbm_budget), C.byref(h)))
        self.h = h
        self.shapes = {}

    @staticmethod
    def instance_bytes(nx, ny, precision=SG_FP64):
        return int(lib().sg_instance_bytes(int(nx), int(ny), int(precision)))

    def add(self, prob):
        inp = _Inputs(prob)
        slot = C.c_int(-1)
        _check(lib().sg_batch_add(self.h, C.byref(inp.grid), C.byref(inp.coeffs), C.byref(inp.bc),
                                  inp.f0, inp.dt, C.byref(slot)))
        self.shapes[slot.value] = (prob.nx, prob.ny)
        return slot.value

    def try_add(self, prob):
        """Like add, but returns None when the HBM budget is exhausted."""
        try:
            return self.add(prob)
        except SayramError as e:
            if e.code == SG_ERR_NOMEM:
                return None
            raise

    def remove(self, slot):
        _check(lib().sg_batch_remove(self.h, int(slot)))
        self.shapes.pop(slot, None)

    def step(self, nsteps=1):
        return _check(lib().sg_batch_step(self.h, int(nsteps)))

    def step_status(self, nsteps=1):
        return lib().sg_batch_step(self.h, int(nsteps))

    def step_until(self, targets, nsteps=1):
        t = np.ascontiguousarray(targets, dtype=np.int64)
        return lib().sg_batch_step_until(self.h, t.ctypes.data_as(C.POINTER(C.c_int64)), len(t),
                                         int(nsteps))

    def f(self, slot):
        nx, ny = self.shapes[slot]
        out = np.empty((nx, ny))
        _check(lib().sg_batch_get_f(self.h, int(slot), _p(out)))
        return out

    def f_static(self, slot):
        """f of one instance through the no-allocation read-out
        (sg_batch_get_f_static)."""
        nx, ny = self.shapes[slot]
        out = np.empty((nx, ny))
        _check(lib().sg_batch_get_f_static(self.h, int(slot), _p(out)))
        return out

    def set_f(self, slot, f, istep):
        f = np.ascontiguousarray(f, dtype=np.float64)
        if slot not in self.shapes:
            raise SayramError(SG_ERR_ARG, f"no instance in slot {slot}")
        if f.shape != self.shapes[slot]:
            raise ValueError(f"f has shape {f.shape}, the instance {self.shapes[slot]}")
        _check(lib().sg_batch_set_f(self.h, int(slot), _p(f), int(istep)))

    def set_bc_values(self, slot, side, vals):
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        _check(lib().sg_batch_set_bc_values(self.h, int(slot), int(side), _p(vals)))

    def set_coeffs(self, slot, G=None, Dxx=None, Dxy=None, Dyy=None, inv_tau=None):
        keep = []
        c = _coeff_struct(keep, G, Dxx, Dxy, Dyy, inv_tau)
        _check(lib().sg_batch_set_coeffs(self.h, int(slot), C.byref(c)))

    def stats(self, slot):
        s = SgStats()
        _check(lib().sg_batch_stats(self.h, int(slot), C.byref(s)))
        return s.as_dict()

    def size(self):
        return lib().sg_batch_size(self.h)

    def bytes_in_use(self):
        return int(lib().sg_batch_bytes_in_use(self.h))

    def time_kernel(self, name):
        """Times every dispatch of one kernel family ("all": every family; None: off)."""
        kid = -1 if not name else TIME_ALL if name == "all" else KERNEL_IDS[name]
        _check(lib().sg_batch_time_kernel(self.h, kid))

    def kernel_time(self, name=None):
        """(total ms, launches, algorithmic bytes) of the timed families, or of one."""
        ms, n, b = C.c_double(), C.c_int64(), C.c_double()
        if name is None:
            _check(lib().sg_batch_kernel_time(self.h, C.byref(ms), C.byref(n), C.byref(b)))
        else:
            _check(lib().sg_batch_kernel_time_id(self.h, KERNEL_IDS[name], C.byref(ms), C.byref(n),
                                                 C.byref(b)))
        return ms.value, n.value, b.value

    def event_overhead_ms(self):
        """Mean HIP-event interval around a one-workgroup kernel on the batch's
        stream (the bracket overhead of every timed dispatch)."""
        ms = C.c_double()
        _check(lib().sg_batch_event_overhead(self.h, C.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            lib().sg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
