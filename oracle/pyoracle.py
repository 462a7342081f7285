"""ctypes binding of the FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SMP, MMP1, MMP2, MMP3 = 0, 1, 2, 3
VARIANTS = {"smp": SMP, "mmp1": MMP1, "mmp2": MMP2, "mmp3": MMP3}


class _Scale(C.Structure):
    _fields_ = [("s", C.c_double), ("Ngauss", C.c_int), ("Nfft", C.c_int),
                ("T", C.c_int), ("F", C.c_int)]


class _Dict(C.Structure):
    _fields_ = [("fs", C.c_double), ("tMax", C.c_double), ("root", C.c_double),
                ("N", C.c_int), ("L", C.c_int), ("entries", C.c_longlong),
                ("scales", C.POINTER(_Scale))]


class _Atom(C.Structure):
    _fields_ = [("params", C.c_double * 6), ("k", C.c_int), ("f", C.c_int),
                ("t", C.c_int), ("pad", C.c_int)]


class _Settings(C.Structure):
    _fields_ = [("iterationMax", C.c_int), ("residualEnergy", C.c_double),
                ("variant", C.c_int)]


class _MapView(C.Structure):
    _fields_ = [("map", C.POINTER(C.c_double)), ("C", C.c_int), ("F", C.c_int), ("T", C.c_int),
                ("pad", C.c_int)]


_TRACE_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_int, C.c_int,
                        C.POINTER(_MapView), C.c_int)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        L.eo_dict_build.argtypes = [C.c_double, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.POINTER(_Dict)]
        L.eo_dict_free.argtypes = [C.POINTER(_Dict)]
        L.eo_rfft.argtypes = [dp, C.c_int, dp]
        L.eo_column.argtypes = [C.POINTER(_Dict), C.c_int, dp, C.c_int, dp]
        L.eo_product.argtypes = [C.POINTER(_Dict), C.c_int, C.c_int, C.c_int, dp, C.c_int, C.c_int]
        L.eo_product.restype = C.c_double
        L.eo_mmp1_constraint.argtypes = [dp, C.c_int]
        L.eo_atom_params.argtypes = [C.POINTER(_Dict), C.c_int, C.c_int, C.c_int, dp, C.c_int,
                                     C.POINTER(_Atom)]
        L.eo_subtract.argtypes = [C.c_double, C.POINTER(_Atom), dp, C.c_int, C.c_int]
        L.eo_heap_new.argtypes = [dp, C.c_size_t]
        L.eo_heap_new.restype = C.c_void_p
        L.eo_heap_update.argtypes = [C.c_void_p, C.c_size_t, C.c_double]
        L.eo_heap_peek.argtypes = [C.c_void_p, C.POINTER(C.c_size_t), dp]
        L.eo_heap_second.argtypes = [C.c_void_p]
        L.eo_heap_second.restype = C.c_double
        L.eo_heap_free.argtypes = [C.c_void_p]
        L.eo_decompose_limited.argtypes = [C.POINTER(_Dict), dp, C.c_int, C.POINTER(_Settings),
                                           C.POINTER(_Atom), C.POINTER(C.c_int), dp, dp, C.c_int,
                                           C.c_int]
        L.eo_decompose_traced.argtypes = [C.POINTER(_Dict), dp, C.c_int, C.POINTER(_Settings),
                                          C.POINTER(_Atom), C.POINTER(C.c_int), dp, dp, C.c_int,
                                          C.c_int, _TRACE_FN, C.c_void_p]
        L.eo_decompose_follow.argtypes = [C.POINTER(_Dict), dp, C.c_int, C.POINTER(_Settings),
                                          C.POINTER(_Atom), C.POINTER(C.c_int), dp, dp, C.c_int,
                                          C.c_int, _TRACE_FN, C.c_void_p, C.POINTER(C.c_int), C.c_double,
                                          C.POINTER(C.c_int)]
        L.eo_waveform_norm2.argtypes = [C.c_double, C.c_int, dp, C.c_int, dp]
        L.eo_fresh_products.argtypes = [C.POINTER(_Dict), C.c_int, dp, C.c_int, C.c_int, dp]
        L.eo_plant_narrow_dirty_rule.argtypes = [C.c_int]
        L.eo_strerror.restype = C.c_char_p
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class Scale:
    s: float
    Ngauss: int
    Nfft: int
    T: int
    F: int


class Dictionary:
    """Optimal Gabor dictionary (gabor.cpp:213-252)."""

    def __init__(self, fs, N, energy_error, scale_min=0.0, scale_max=float("inf"),
                 freq_max=float("inf")):
        self._d = _Dict()
        st = lib().eo_dict_build(float(fs), int(N), float(energy_error), float(scale_min),
                                 float(scale_max), float(freq_max), C.byref(self._d))
        if st != 0:
            raise OracleError(st)
        self.fs = self._d.fs
        self.N = self._d.N
        self.tMax = self._d.tMax
        self.root = self._d.root
        self.entries = self._d.entries
        self.scales = [Scale(self._d.scales[i].s, self._d.scales[i].Ngauss, self._d.scales[i].Nfft,
                             self._d.scales[i].T, self._d.scales[i].F) for i in range(self._d.L)]

    def __del__(self):
        try:
            lib().eo_dict_free(C.byref(self._d))
        except Exception:
            pass

    def column(self, k, x, t):
        x = np.ascontiguousarray(x, dtype=np.float64)
        v = np.zeros(2 * self.scales[k].F)
        lib().eo_column(C.byref(self._d), k, _dp(x), t, _dp(v))
        return v[0::2] + 1j * v[1::2]

    def product(self, k, f, t, values, constraint=False):
        v = np.empty(2 * len(values))
        v[0::2] = np.real(values)
        v[1::2] = np.imag(values)
        return lib().eo_product(C.byref(self._d), k, f, t, _dp(v), len(values), int(constraint))

    def fresh_products(self, k, x, constraint=False):
        """Products [F][T] of scale k from fresh columns of x [C][N]: column +
        product only, no stored map or lazy index."""
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64)
        sc = self.scales[k]
        out = np.zeros(sc.F * sc.T)
        lib().eo_fresh_products(C.byref(self._d), k, _dp(x), x.shape[0], int(constraint), _dp(out))
        return out.reshape(sc.F, sc.T)

    def atom_params(self, k, f, t, values):
        v = np.empty(2 * len(values))
        v[0::2] = np.real(values)
        v[1::2] = np.imag(values)
        out = (_Atom * len(values))()
        lib().eo_atom_params(C.byref(self._d), k, f, t, _dp(v), len(values), out)
        return np.array([list(a.params) for a in out])


class OracleError(RuntimeError):
    def __init__(self, code):
        self.code = code
        super().__init__(lib().eo_strerror(code).decode())


def rfft(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = len(x)
    X = np.zeros(2 * (n // 2 + 1))
    lib().eo_rfft(_dp(x), n, _dp(X))
    return X[0::2] + 1j * X[1::2]


def mmp1_constraint(values):
    v = np.empty(2 * len(values))
    v[0::2] = np.real(values)
    v[1::2] = np.imag(values)
    lib().eo_mmp1_constraint(_dp(v), len(values))
    return v[0::2] + 1j * v[1::2]


def subtract(fs, params, x, fit=False):
    """Returns (new residual, params) after gabor.cpp:143-187."""
    a = _Atom()
    for i in range(6):
        a.params[i] = params[i]
    x = np.array(x, dtype=np.float64)
    lib().eo_subtract(float(fs), C.byref(a), _dp(x), len(x), int(fit))
    return x, np.array(list(a.params))


def waveform_norm2(fs, N, params):
    """||w||^2 of the waveform subtract() removes, per atom of params [n][6]."""
    p = np.ascontiguousarray(np.atleast_2d(params), dtype=np.float64)
    out = np.zeros(p.shape[0])
    lib().eo_waveform_norm2(float(fs), int(N), _dp(p), p.shape[0], _dp(out))
    return out


class Heap:
    """Indexed max-heap restatement (heap.hpp:20-94)."""

    def __init__(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        self._h = lib().eo_heap_new(_dp(v), len(v))

    def update(self, key, value):
        lib().eo_heap_update(self._h, key, float(value))

    def peek(self):
        k = C.c_size_t()
        v = C.c_double()
        lib().eo_heap_peek(self._h, C.byref(k), C.byref(v))
        return k.value, v.value

    def second(self):
        return lib().eo_heap_second(self._h)

    def __del__(self):
        try:
            lib().eo_heap_free(self._h)
        except Exception:
            pass


@dataclass
class Result:
    params: np.ndarray      # [C][I][6]
    index: np.ndarray       # [C][I][3] (k, f, t)
    counts: np.ndarray      # [C]
    energies: np.ndarray    # [R][I+1]
    gaps: np.ndarray        # [R][I]
    status: int
    followed: np.ndarray = None   # [R][I]: 1 where the loop took the followed entry


@dataclass
class TraceStep:
    iteration: int          # atoms subtracted so far
    row: int                # SMP channel, else 0
    residual: np.ndarray    # [C][N] after the subtractions
    maps: list              # per scale: stored map [C][F][T] complex, before the lazy update


def plant_narrow_dirty_rule(on):
    """Planted mistake (tests only): re-evaluate |t - tA| <= 3 s instead of 3 (s + sA)."""
    lib().eo_plant_narrow_dirty_rule(int(on))


def decompose(d: Dictionary, signal, variant, iteration_max, energy_percent,
              threads=0, iter_limit=None, check=True, trace=None, follow=None,
              follow_tol=1e-4) -> Result:
    """trace: a list that receives one TraceStep per subtracting iteration.
    follow: [R][I][3] (k, f, t) per row (channels for SMP, else one row), -1
    for none; where that entry differs from the oracle's own best and its
    product is within follow_tol of the best, the loop takes it (a near-tie
    branch) and Result.followed marks it."""
    sig = np.ascontiguousarray(np.atleast_2d(signal), dtype=np.float64)
    Cn, N = sig.shape
    assert N == d.N
    I = int(iteration_max)
    st = _Settings(I, 1.0 - 0.01 * float(energy_percent), int(variant))
    atoms = (_Atom * (Cn * I))()
    counts = np.zeros(Cn, dtype=np.int32)
    rows = Cn if variant == SMP else 1
    energies = np.zeros(rows * (I + 1))
    gaps = np.zeros(rows * I)
    def on_step(_user, iteration, row, residual, nc, n, views, L):
        res = np.ctypeslib.as_array(residual, shape=(nc * n,)).reshape(nc, n).copy()
        maps = []
        for k in range(L):
            v = views[k]
            a = np.ctypeslib.as_array(v.map, shape=(2 * v.C * v.F * v.T,))
            maps.append((a[0::2] + 1j * a[1::2]).reshape(v.C, v.F, v.T).copy())
        trace.append(TraceStep(iteration, row, res, maps))

    cb = _TRACE_FN(on_step) if trace is not None else _TRACE_FN()
    followed = np.zeros(rows * I, dtype=np.int32)
    fp = None
    if follow is not None:
        fa = np.ascontiguousarray(follow, dtype=np.int32)
        assert fa.shape == (rows, I, 3), (fa.shape, rows, I)
        fp = fa.ctypes.data_as(C.POINTER(C.c_int))
    status = lib().eo_decompose_follow(
        C.byref(d._d), _dp(sig), Cn, C.byref(st), atoms,
        counts.ctypes.data_as(C.POINTER(C.c_int)), _dp(energies), _dp(gaps), int(threads),
        int(iter_limit if iter_limit is not None else 2**31 - 1), cb, None, fp, float(follow_tol),
        followed.ctypes.data_as(C.POINTER(C.c_int)))
    if check and status != 0:
        raise OracleError(status)
    params = np.zeros((Cn, I, 6))
    index = np.zeros((Cn, I, 3), dtype=np.int64)
    for c in range(Cn):
        for i in range(I):
            a = atoms[c * I + i]
            params[c, i] = list(a.params)
            index[c, i] = (a.k, a.f, a.t)
    return Result(params, index, counts, energies.reshape(rows, I + 1), gaps.reshape(rows, I),
                  status, followed.reshape(rows, I))
