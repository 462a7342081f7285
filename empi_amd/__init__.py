"""empi_amd: Matching Pursuit decomposition on AMD MI355X (gfx950).

Python view of the C ABI in include/empi_hip.h (libempi_hip.so, built in
tree by `python -m empi_amd.build`).  The library is required: there is no
CPU fallback, and every entry point raises if the HIP engine is missing.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libempi_hip.so")

SMP, MMP1, MMP2, MMP3 = 0, 1, 2, 3
VARIANTS = {"smp": SMP, "mmp1": MMP1, "mmp2": MMP2, "mmp3": MMP3}

OK = 0
STATUS_NAMES = {
    0: "ok", 1: "signalIsEmpty", 2: "energyMapIsEmpty",
    3: "signalFileIsTooLongForThisDecomposition", 4: "invalidArgument",
    5: "insufficientMemory", 6: "internalError", 7: "deviceError", 8: "unsupportedSize",
}

EXPORTS = [
    "empi_dict_info", "empi_ws_create", "empi_ws_compute", "empi_ws_find_best", "empi_ws_atom_count",
    "empi_ws_subtract", "empi_ws_destroy", "empi_subtract_atom_from_signal",
    "empi_plan_create", "empi_plan_run", "empi_plan_item_result", "empi_plan_item_atom_count",
    "empi_plan_entries_evaluated", "empi_plan_set_profiling", "empi_plan_kernel_classes",
    "empi_plan_kernel_stats", "empi_plan_destroy", "empi_strerror", "empi_version",
    "empi_mlp_create", "empi_mlp_destroy", "empi_mlp_param_count", "empi_mlp_get_params",
    "empi_mlp_set_params", "empi_mlp_forward", "empi_mlp_forward16", "empi_mlp_gradients", "empi_mlp_train",
    "empi_plan_tuning_size", "empi_plan_autotune", "empi_plan_get_tuning", "empi_plan_set_tuning",
    "empi_proposal_feature_dim", "empi_proposal_bound_factor", "empi_plan_set_proposal", "empi_plan_proposal_stats",
    "empi_plan_set_training_capacity", "empi_plan_training_rows", "empi_decompose_noalloc", "empi_plan_stage",
    "empi_debug_argmax",
]

PROPOSAL_OFF, PROPOSAL_ON, PROPOSAL_COLLECT = 0, 1, 2


class DictParams(C.Structure):
    _fields_ = [("energyError", C.c_double), ("scaleMin", C.c_double),
                ("scaleMax", C.c_double), ("freqMax", C.c_double)]


class Atom(C.Structure):
    _fields_ = [("params", C.c_double * 6), ("scale", C.c_int), ("freq", C.c_int),
                ("time", C.c_int), ("reserved", C.c_int)]


class ScaleInfo(C.Structure):
    _fields_ = [("scale", C.c_double), ("Nfft", C.c_int), ("tCount", C.c_int),
                ("fCount", C.c_int), ("reserved", C.c_int)]


class Item(C.Structure):
    _fields_ = [("channels", C.c_int), ("samples", C.c_int), ("fs", C.c_double),
                ("variant", C.c_int), ("iterationMax", C.c_int), ("energyPercent", C.c_double),
                ("dict", DictParams)]


class EmpiError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        name = STATUS_NAMES.get(status, "internalError")
        super().__init__(f"{name}" + (f" ({where})" if where else ""))


_lib = None


def lib():
    """Loads libempi_hip.so; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP engine not built: {LIB_PATH} (python -m empi_amd.build)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.empi_dict_info.argtypes = [C.POINTER(DictParams), C.c_double, C.c_int, C.POINTER(ScaleInfo),
                                     C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]
        L.empi_ws_create.argtypes = [C.POINTER(DictParams), C.c_double, C.c_int, C.c_int, C.c_int,
                                     C.POINTER(vp)]
        L.empi_ws_compute.argtypes = [vp, dp]
        L.empi_ws_find_best.argtypes = [vp, C.POINTER(Atom)]
        L.empi_ws_atom_count.argtypes = [vp]
        L.empi_ws_atom_count.restype = C.c_size_t
        L.empi_ws_subtract.argtypes = [vp, C.POINTER(Atom), C.c_int, dp]
        L.empi_ws_destroy.argtypes = [vp]
        L.empi_ws_destroy.restype = None
        L.empi_subtract_atom_from_signal.argtypes = [C.c_double, C.POINTER(Atom), dp, C.c_int, C.c_int]
        L.empi_plan_create.argtypes = [C.POINTER(Item), C.c_int, C.POINTER(vp)]
        L.empi_plan_run.argtypes = [vp, vp, C.c_int, vp]
        L.empi_plan_stage.argtypes = [vp, vp]
        L.empi_plan_item_result.argtypes = [vp, C.c_int, C.POINTER(Atom), C.POINTER(C.c_int), dp]
        L.empi_plan_item_atom_count.argtypes = [vp, C.c_int]
        L.empi_plan_item_atom_count.restype = C.c_size_t
        L.empi_plan_entries_evaluated.argtypes = [vp]
        L.empi_plan_entries_evaluated.restype = C.c_double
        L.empi_plan_set_profiling.argtypes = [vp, C.c_int]
        L.empi_plan_kernel_classes.restype = C.c_int
        L.empi_plan_kernel_stats.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_longlong),
                                             dp, dp, dp]
        L.empi_plan_destroy.argtypes = [vp]
        L.empi_plan_destroy.restype = None
        fp = C.POINTER(C.c_float)
        L.empi_mlp_create.argtypes = [C.c_int, C.c_uint, C.POINTER(vp)]
        L.empi_mlp_destroy.argtypes = [vp]
        L.empi_mlp_destroy.restype = None
        L.empi_mlp_param_count.argtypes = [vp]
        L.empi_mlp_param_count.restype = C.c_size_t
        L.empi_mlp_get_params.argtypes = [vp, fp]
        L.empi_mlp_set_params.argtypes = [vp, fp]
        L.empi_mlp_forward.argtypes = [vp, vp, C.c_int, vp, C.c_int]
        L.empi_mlp_forward16.argtypes = [vp, vp, C.c_int, vp, C.c_int]
        L.empi_mlp_gradients.argtypes = [vp, fp, fp, C.c_int, fp, fp]
        L.empi_mlp_train.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint, fp]
        L.empi_plan_tuning_size.restype = C.c_int
        L.empi_plan_autotune.argtypes = [vp, vp, C.c_int, C.c_int, vp]
        L.empi_plan_get_tuning.argtypes = [vp, C.POINTER(C.c_int)]
        L.empi_plan_set_tuning.argtypes = [vp, C.POINTER(C.c_int)]
        L.empi_proposal_feature_dim.restype = C.c_int
        L.empi_proposal_bound_factor.argtypes = [C.c_double, C.c_double, C.c_int]
        L.empi_proposal_bound_factor.restype = C.c_double
        L.empi_plan_set_proposal.argtypes = [vp, C.c_int, vp, C.c_int, C.c_double]
        L.empi_plan_proposal_stats.argtypes = [vp, C.POINTER(C.c_ulonglong)]
        L.empi_plan_set_training_capacity.argtypes = [vp, C.c_size_t, C.c_uint]
        L.empi_plan_training_rows.argtypes = [vp, fp, fp, C.c_size_t, C.POINTER(C.c_size_t)]
        L.empi_decompose_noalloc.argtypes = [C.POINTER(Item), vp, C.c_int, C.POINTER(Atom), C.POINTER(C.c_int),
                                             dp, vp]
        ip = C.POINTER(C.c_int)
        L.empi_debug_argmax.argtypes = [C.c_int, ip, ip, ip, ip, ip, fp, ip, fp]
        L.empi_strerror.argtypes = [C.c_int]
        L.empi_strerror.restype = C.c_char_p
        L.empi_version.restype = C.c_char_p
        _lib = L
    return _lib


def _check(status, where=""):
    if status != OK:
        raise EmpiError(status, where)


def dict_params(energy_error, scale_min=0.0, scale_max=math.inf, freq_max=math.inf):
    return DictParams(float(energy_error), float(scale_min), float(scale_max), float(freq_max))


def _atoms_to_arrays(atoms, shape):
    params = np.array([list(a.params) for a in atoms], dtype=np.float64).reshape(shape + (6,))
    index = np.array([(a.scale, a.freq, a.time) for a in atoms], dtype=np.int64).reshape(shape + (3,))
    return params, index


@dataclass
class ItemResult:
    status: int
    params: np.ndarray      # [C][I][6]
    index: np.ndarray       # [C][I][3] (scale, bin, time)
    counts: np.ndarray      # [C]
    energies: np.ndarray    # [rows][I+1]


@dataclass
class ItemSpec:
    channels: int
    samples: int
    fs: float
    variant: int
    iteration_max: int
    energy_percent: float
    energy_error: float
    scale_min: float = 0.0
    scale_max: float = math.inf
    freq_max: float = math.inf

    def to_c(self):
        return Item(self.channels, self.samples, float(self.fs), int(self.variant),
                    int(self.iteration_max), float(self.energy_percent),
                    dict_params(self.energy_error, self.scale_min, self.scale_max, self.freq_max))


def debug_argmax(maps):
    """Test hook: the engine's argmax keys over planted maps on the device.
    maps: list of (products [T][G][B] float32, bins [G][B] int, cells per
    tile).  Returns ((scale, bin, time), best product)."""
    L = len(maps)
    T = np.array([m[0].shape[0] for m in maps], np.int32)
    G = np.array([m[0].shape[1] for m in maps], np.int32)
    B = np.array([m[0].shape[2] for m in maps], np.int32)
    nc = np.array([m[2] for m in maps], np.int32)
    bins = np.ascontiguousarray(np.concatenate([np.asarray(m[1], np.int32).ravel() for m in maps]))
    prods = np.ascontiguousarray(np.concatenate([np.asarray(m[0], np.float32).ravel() for m in maps]))
    out = np.zeros(3, np.int32)
    best = np.zeros(1, np.float32)
    ip = C.POINTER(C.c_int)
    _check(lib().empi_debug_argmax(L, T.ctypes.data_as(ip), G.ctypes.data_as(ip), nc.ctypes.data_as(ip),
                                   B.ctypes.data_as(ip), bins.ctypes.data_as(ip),
                                   prods.ctypes.data_as(C.POINTER(C.c_float)), out.ctypes.data_as(ip),
                                   best.ctypes.data_as(C.POINTER(C.c_float))))
    return tuple(int(v) for v in out), float(best[0])


def decompose_noalloc(item, signal=None, device_ptr=None, stream=None) -> ItemResult:
    """One item through empi_decompose_noalloc (no device allocation: the
    library's fixed scratch array).  signal: float32 host array [C][N], or
    device_ptr: integer device address of the same layout.  The returned
    status is the item's; a call-level failure raises."""
    Cn, I = item.channels, item.iteration_max
    rows = Cn if item.variant == SMP else 1
    atoms = (Atom * (Cn * I))()
    counts = np.zeros(Cn, dtype=np.int32)
    energies = np.zeros(rows * (I + 1))
    spec = item.to_c()
    st_ptr = C.c_void_p(stream) if stream else None
    if device_ptr is not None:
        ptr, on_device = C.c_void_p(int(device_ptr)), 1
    else:
        buf = np.ascontiguousarray(signal, dtype=np.float32).ravel()
        if buf.size != Cn * item.samples:
            raise ValueError(f"expected {Cn * item.samples} samples, got {buf.size}")
        ptr, on_device = buf.ctypes.data_as(C.c_void_p), 0
    status = lib().empi_decompose_noalloc(C.byref(spec), ptr, on_device, atoms,
                                          counts.ctypes.data_as(C.POINTER(C.c_int)),
                                          energies.ctypes.data_as(C.POINTER(C.c_double)), st_ptr)
    if status not in (OK, 1, 2):
        raise EmpiError(status, "empi_decompose_noalloc")
    params, index = _atoms_to_arrays(atoms, (Cn, I))
    return ItemResult(status, params, index, counts, energies.reshape(rows, I + 1))


class Plan:
    """Batched device-resident decomposition (empi_plan_*)."""

    def __init__(self, items):
        self.items = list(items)
        arr = (Item * len(self.items))(*[it.to_c() for it in self.items])
        h = C.c_void_p()
        _check(lib().empi_plan_create(arr, len(self.items), C.byref(h)), "empi_plan_create")
        self._h = h
        self.total_samples = sum(it.channels * it.samples for it in self.items)

    def run(self, signal=None, device_ptr=None, stream=None):
        """signal: float32 host array with every item's [C][N] samples in
        order; or device_ptr: integer device address of the same layout."""
        st_ptr = C.c_void_p(stream) if stream else None
        if device_ptr is not None:
            status = lib().empi_plan_run(self._h, C.c_void_p(int(device_ptr)), 1, st_ptr)
        else:
            buf = np.ascontiguousarray(signal, dtype=np.float32).ravel()
            if buf.size != self.total_samples:
                raise ValueError(f"expected {self.total_samples} samples, got {buf.size}")
            status = lib().empi_plan_run(self._h, buf.ctypes.data_as(C.c_void_p), 0, st_ptr)
        _check(status, "empi_plan_run")

    def result(self, i) -> ItemResult:
        it = self.items[i]
        Cn, I = it.channels, it.iteration_max
        rows = Cn if it.variant == SMP else 1
        atoms = (Atom * (Cn * I))()
        counts = np.zeros(Cn, dtype=np.int32)
        energies = np.zeros(rows * (I + 1))
        status = lib().empi_plan_item_result(self._h, i, atoms,
                                             counts.ctypes.data_as(C.POINTER(C.c_int)),
                                             energies.ctypes.data_as(C.POINTER(C.c_double)))
        params, index = _atoms_to_arrays(atoms, (Cn, I))
        return ItemResult(status, params, index, counts, energies.reshape(rows, I + 1))

    def atom_count(self, i):
        return lib().empi_plan_item_atom_count(self._h, i)

    def entries_evaluated(self):
        return lib().empi_plan_entries_evaluated(self._h)

    def set_profiling(self, on=True):
        _check(lib().empi_plan_set_profiling(self._h, int(on)), "empi_plan_set_profiling")

    def kernel_stats(self):
        """Per sub-transform class of the last run: dicts with sub_length,
        launches, total_ms, entries, entry_channels (classes never launched
        are omitted)."""
        out = []
        for c in range(lib().empi_plan_kernel_classes()):
            q = C.c_int()
            n = C.c_longlong()
            ms, e, ec = C.c_double(), C.c_double(), C.c_double()
            _check(lib().empi_plan_kernel_stats(self._h, c, C.byref(q), C.byref(n), C.byref(ms),
                                                C.byref(e), C.byref(ec)), "empi_plan_kernel_stats")
            if n.value or e.value:
                out.append({"class": c, "sub_length": q.value, "launches": n.value,
                            "total_ms": ms.value, "entries": e.value, "entry_channels": ec.value})
        return out

    def autotune(self, signal=None, device_ptr=None, repeats=2):
        """Times every column-kernel variant on this batch's full map build
        and keeps the fastest per class; returns the tuning table."""
        if device_ptr is not None:
            status = lib().empi_plan_autotune(self._h, C.c_void_p(int(device_ptr)), 1, int(repeats), None)
        else:
            buf = np.ascontiguousarray(signal, dtype=np.float32).ravel()
            status = lib().empi_plan_autotune(self._h, buf.ctypes.data_as(C.c_void_p), 0, int(repeats), None)
        _check(status, "empi_plan_autotune")
        return self.tuning()

    def tuning(self):
        out = np.zeros(lib().empi_plan_tuning_size(), dtype=np.int32)
        _check(lib().empi_plan_get_tuning(self._h, out.ctypes.data_as(C.POINTER(C.c_int))))
        return out

    def set_tuning(self, table):
        t = np.ascontiguousarray(table, dtype=np.int32)
        assert t.size == lib().empi_plan_tuning_size()
        _check(lib().empi_plan_set_tuning(self._h, t.ctypes.data_as(C.POINTER(C.c_int))), "empi_plan_set_tuning")

    def set_proposal(self, mode=PROPOSAL_ON, net=None, top_k=8, delta=1e-3):
        """Learned proposal (see empi_plan_set_proposal); net: an MLP or None.
        The plan keeps a reference to the network."""
        self._net = net
        _check(lib().empi_plan_set_proposal(self._h, int(mode), net._h if net is not None else None,
                                            int(top_k), float(delta)), "empi_plan_set_proposal")

    def proposal_stats(self):
        out = (C.c_ulonglong * 5)()
        _check(lib().empi_plan_proposal_stats(self._h, out), "empi_plan_proposal_stats")
        keys = ["fallback_iterations", "fallback_regions", "certified_iterations", "dirty_checked",
                "proposed_regions"]
        return dict(zip(keys, [int(v) for v in out]))

    def set_training_capacity(self, rows, stride=1):
        _check(lib().empi_plan_set_training_capacity(self._h, int(rows), int(stride)),
               "empi_plan_set_training_capacity")
        self._train_cap = int(rows)

    def training_rows(self):
        """(X [rows][feature_dim], target [rows]) collected since the last read."""
        cap = getattr(self, "_train_cap", 0)
        d = lib().empi_proposal_feature_dim()
        X = np.zeros((cap, d), dtype=np.float32)
        t = np.zeros(cap, dtype=np.float32)
        n = C.c_size_t()
        _check(lib().empi_plan_training_rows(self._h, X.ctypes.data_as(C.POINTER(C.c_float)),
                                             t.ctypes.data_as(C.POINTER(C.c_float)), cap, C.byref(n)),
               "empi_plan_training_rows")
        return X[:n.value], t[:n.value]

    def close(self):
        if getattr(self, "_h", None):
            lib().empi_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Workspace:
    """Per-call Workspace operator API (empi_ws_*), reference granularity."""

    def __init__(self, fs, channels, samples, energy_error, scale_min=0.0, scale_max=math.inf,
                 freq_max=math.inf, constraint=0):
        h = C.c_void_p()
        dp = dict_params(energy_error, scale_min, scale_max, freq_max)
        _check(lib().empi_ws_create(C.byref(dp), float(fs), int(channels), int(samples),
                                    int(constraint), C.byref(h)), "empi_ws_create")
        self._h = h
        self.fs, self.channels, self.samples = float(fs), int(channels), int(samples)

    def compute(self, signal):
        x = np.ascontiguousarray(np.atleast_2d(signal), dtype=np.float64)
        assert x.shape == (self.channels, self.samples)
        _check(lib().empi_ws_compute(self._h, x.ctypes.data_as(C.POINTER(C.c_double))),
               "empi_ws_compute")

    def find_best(self):
        out = (Atom * self.channels)()
        _check(lib().empi_ws_find_best(self._h, out), "empi_ws_find_best")
        return list(out)

    def atom_count(self):
        return lib().empi_ws_atom_count(self._h)

    def subtract(self, atom, channel, residual):
        """Subtracts atom from residual (float64 numpy array, updated in place)."""
        assert residual.dtype == np.float64 and residual.flags.c_contiguous
        _check(lib().empi_ws_subtract(self._h, C.byref(atom), int(channel),
                                      residual.ctypes.data_as(C.POINTER(C.c_double))),
               "empi_ws_subtract")

    def close(self):
        if getattr(self, "_h", None):
            lib().empi_ws_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dictionary(fs, samples, energy_error, scale_min=0.0, scale_max=math.inf, freq_max=math.inf):
    """Host-only dictionary description: (list of (s, Nfft, T, F), entries)."""
    dp = dict_params(energy_error, scale_min, scale_max, freq_max)
    n = C.c_int()
    e = C.c_longlong()
    _check(lib().empi_dict_info(C.byref(dp), float(fs), int(samples), None, 0, C.byref(n), C.byref(e)),
           "empi_dict_info")
    arr = (ScaleInfo * max(1, n.value))()
    _check(lib().empi_dict_info(C.byref(dp), float(fs), int(samples), arr, n.value, C.byref(n), C.byref(e)),
           "empi_dict_info")
    return [(a.scale, a.Nfft, a.tCount, a.fCount) for a in arr[:n.value]], e.value


class MLP:
    """Region-scoring network (empi_mlp_*): input_dim -> 256 -> 256 -> 1."""

    def __init__(self, input_dim, seed=0):
        h = C.c_void_p()
        _check(lib().empi_mlp_create(int(input_dim), int(seed), C.byref(h)), "empi_mlp_create")
        self._h = h
        self.input_dim = int(input_dim)

    @property
    def n_params(self):
        return lib().empi_mlp_param_count(self._h)

    def params(self):
        out = np.zeros(self.n_params, dtype=np.float32)
        _check(lib().empi_mlp_get_params(self._h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        _check(lib().empi_mlp_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.zeros(X.shape[0], dtype=np.float32)
        _check(lib().empi_mlp_forward(self._h, X.ctypes.data_as(C.c_void_p), X.shape[0],
                                      y.ctypes.data_as(C.c_void_p), 0), "empi_mlp_forward")
        return y

    def forward16(self, X):
        """The proposal's inference pass (FP16 inputs and hidden layer, matrix cores)."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.zeros(X.shape[0], dtype=np.float32)
        _check(lib().empi_mlp_forward16(self._h, X.ctypes.data_as(C.c_void_p), X.shape[0],
                                        y.ctypes.data_as(C.c_void_p), 0), "empi_mlp_forward16")
        return y

    def gradients(self, X, t):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_float()
        _check(lib().empi_mlp_gradients(self._h, X.ctypes.data_as(C.POINTER(C.c_float)),
                                        t.ctypes.data_as(C.POINTER(C.c_float)), X.shape[0],
                                        g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)))
        return g, loss.value

    def train(self, X, t, epochs=1, batch=4096, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        _check(lib().empi_mlp_train(self._h, X.ctypes.data_as(C.c_void_p), t.ctypes.data_as(C.c_void_p),
                                    X.shape[0], 0, int(epochs), int(batch), float(lr), int(seed),
                                    C.byref(loss)), "empi_mlp_train")
        return loss.value

    def close(self):
        if getattr(self, "_h", None):
            lib().empi_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def proposal_bound_factor(scale, fs, nfft):
    """c in the certified bound P <= c * window energy (host-only)."""
    return lib().empi_proposal_bound_factor(float(scale), float(fs), int(nfft))


def feature_dim():
    """Input width of the proposal's region-scoring network."""
    return lib().empi_proposal_feature_dim()


def version():
    return lib().empi_version().decode()
