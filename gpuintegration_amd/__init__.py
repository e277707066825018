"""MI355X-native PAGANI adaptive cubature -- Python binding of libpagani.so.

The product is the HIP library behind the C ABI in include/pagani.h and the
C++ drop-in headers in include/pagani/.  This module binds the C ABI with
ctypes so tests and benchmarks can drive it; the phase-level calls take torch
device tensors (torch only provides device memory here).  Loading fails
loudly when the library has not been built: there is no CPU fallback.
"""
import ctypes
import os
from dataclasses import dataclass

__all__ = [
    "load", "Workspace", "IntegrationResult", "FAMILY", "FP64", "FP32",
    "PaganiError", "LIB_PATH",
]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpagani.so")

MIN_DIM, MAX_DIM = 2, 16
GENZ_MAX_DIM = 10  # the precompiled Genz descriptor path
NSETS, NRULES, MAX_SLICES = 9, 5, 8
FP64, FP32 = 64, 32
OK, NOT_CONVERGED, EBADARG, EHIP, ENOMEM, EUNSUPPORTED = 0, 1, -1, -2, -3, -4
SUSPENDED = 2

# Genz family codes (include/pagani.h)
FAMILY = {
    "oscillatory": 1,
    "product_peak": 2,
    "corner_peak": 3,
    "gaussian": 4,
    "c0": 5,
    "discontinuous": 6,
}


class PaganiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pagani error {code}: {msg}")
        self.code = code


class _Result(ctypes.Structure):
    _fields_ = [
        ("estimate", ctypes.c_double),
        ("errorest", ctypes.c_double),
        ("neval", ctypes.c_size_t),
        ("nregions", ctypes.c_size_t),
        ("nFinishedRegions", ctypes.c_size_t),
        ("status", ctypes.c_int),
        ("lastPhase", ctypes.c_int),
        ("chi_sq", ctypes.c_double),
        ("iters", ctypes.c_size_t),
    ]


class _Genz(ctypes.Structure):
    _fields_ = [
        ("family", ctypes.c_int),
        ("ndim", ctypes.c_int),
        ("a", ctypes.c_double * MAX_DIM),
        ("u", ctypes.c_double * MAX_DIM),
    ]


class _BatchItem(ctypes.Structure):
    _fields_ = [
        ("f", _Genz),
        ("epsrel", ctypes.c_double),
        ("epsabs", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("relerr_classification", ctypes.c_int),
        ("init_k", ctypes.c_size_t),
    ]


class EvalConsts(ctypes.Structure):
    _fields_ = [
        ("w", (ctypes.c_double * NRULES) * NSETS),
        ("scale", (ctypes.c_double * NRULES) * NSETS),
        ("norm", (ctypes.c_double * NRULES) * NSETS),
        ("lows", ctypes.c_double * MAX_DIM),
        ("range", ctypes.c_double * MAX_DIM),
        ("jac", ctypes.c_double),
        ("lambda1", ctypes.c_double),
        ("lambda2", ctypes.c_double),
        ("ratio", ctypes.c_double),
        ("gtab", ctypes.c_void_p),
        ("ndim", ctypes.c_int),
        ("feval", ctypes.c_int),
        ("set_begin", ctypes.c_int * (NSETS + 1)),
        ("slices", ctypes.c_int),
        ("slice_begin", ctypes.c_int * (MAX_SLICES + 1)),
        ("lambda3", ctypes.c_double),
        ("lambda4", ctypes.c_double),
        ("lambda5", ctypes.c_double),
        ("h0", ctypes.c_double * MAX_DIM),
        ("abi_version", ctypes.c_uint32),
        ("levels", ctypes.c_uint32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("eval_ms", ctypes.c_double),
        ("refine_ms", ctypes.c_double),
        ("finalize_ms", ctypes.c_double),
        ("compact_ms", ctypes.c_double),
        ("classify_ms", ctypes.c_double),
        ("wall_ms", ctypes.c_double),
        ("regions_evaluated", ctypes.c_uint64),
        ("eval_launches", ctypes.c_uint64),
        ("iterations", ctypes.c_uint64),
        ("classify_passes", ctypes.c_uint64),
        ("max_regions", ctypes.c_uint64),
        ("integrals", ctypes.c_uint64),
        ("alloc_ms", ctypes.c_double),
        ("allocations", ctypes.c_uint64),
        ("device_ms", ctypes.c_double),
        ("device_iterations", ctypes.c_uint64),
        ("device_regions", ctypes.c_uint64),
        ("device_rounds", ctypes.c_uint64),
        ("budget_fixes", ctypes.c_uint64),
        ("graph_launches", ctypes.c_uint64),
        ("segment_integrals", ctypes.c_uint64),
        ("segment_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ClassRes(ctypes.Structure):
    _fields_ = [
        ("pass_mem", ctypes.c_int),
        ("pass_errorest_budget", ctypes.c_int),
        ("threshold", ctypes.c_double),
        ("percent_mem_active", ctypes.c_double),
        ("range_low", ctypes.c_double),
        ("range_high", ctypes.c_double),
        ("num_active", ctypes.c_size_t),
        ("finished_errorest", ctypes.c_double),
        ("max_budget_perc_to_cover", ctypes.c_double),
        ("max_active_perc", ctypes.c_double),
        ("threshold_passes", ctypes.c_int),
    ]


_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_sz = ctypes.c_size_t

MC_BINS = 500
MC_MAX_DIM = 16
MC_RNG = {"default": 0, "lcg": 1}


class McParams(ctypes.Structure):
    _fields_ = [("ncall", ctypes.c_double), ("total_iters", ctypes.c_int),
                ("adjust_iters", ctypes.c_int), ("skip_iters", ctypes.c_int),
                ("generator", ctypes.c_int), ("one_axis", ctypes.c_int),
                ("adjust_params", ctypes.c_int), ("seed", ctypes.c_uint64)]


class McResult(ctypes.Structure):
    _fields_ = [("estimate", ctypes.c_double), ("errorest", ctypes.c_double),
                ("chi_sq", ctypes.c_double), ("status", ctypes.c_int), ("pad", ctypes.c_int),
                ("iters", ctypes.c_uint64), ("runs", ctypes.c_uint64),
                ("neval", ctypes.c_uint64), ("ncall", ctypes.c_double)]


class McStats(ctypes.Structure):
    _fields_ = [("sample_ms", ctypes.c_double), ("wall_ms", ctypes.c_double),
                ("samples", ctypes.c_uint64), ("launches", ctypes.c_uint64),
                ("integrals", ctypes.c_uint64)]

# name -> (restype, argtypes)
_SIGS = {
    "pagani_rule_tables": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp,
                                          ctypes.POINTER(ctypes.c_int), _dp]),
    "pagani_ws_create": (_vp, [ctypes.c_int, ctypes.c_int]),
    "pagani_create_error": (ctypes.c_char_p, []),
    "pagani_ws_destroy": (None, [_vp]),
    "pagani_ws_last_error": (ctypes.c_char_p, [_vp]),
    "pagani_ws_ndim": (ctypes.c_int, [_vp]),
    "pagani_ws_set_device_mem": (ctypes.c_int, [_vp, _sz]),
    "pagani_ws_device_mem": (_sz, [_vp]),
    "pagani_ws_set_mem_model": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pagani_ws_set_slices": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pagani_ws_set_device_loop": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pagani_ws_set_graphs": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pagani_ws_set_tuned_slices": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pagani_ws_set_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pagani_ws_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(Stats)]),
    "pagani_ws_reset_stats": (ctypes.c_int, [_vp]),
    "pagani_ws_eval_consts": (ctypes.c_int, [_vp, _dp, _dp, _sz, ctypes.POINTER(EvalConsts)]),
    "pagani_ws_stream": (_vp, [_vp]),
    "pagani_ws_trace": (_sz, [_vp, _sz, _vp, _vp, _vp, _vp]),
    "pagani_ws_live_regions": (ctypes.c_int, [_vp, ctypes.POINTER(_sz), ctypes.POINTER(_vp),
                                              ctypes.POINTER(_vp)]),
    "pagani_integrate_genz": (ctypes.c_int, [_vp, ctypes.POINTER(_Genz), ctypes.c_int,
                                             ctypes.c_double, ctypes.c_double, _dp, _dp,
                                             ctypes.c_int, _sz, ctypes.POINTER(_Result)]),
    "pagani_integrate_cb": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, ctypes.c_double,
                                           _dp, _dp, ctypes.c_int, _vp, _sz,
                                           ctypes.POINTER(_Result)]),
    "pagani_uniform_split_dev": (ctypes.c_int, [_vp, _sz, _vp, _vp]),
    "pagani_eval_genz_dev": (ctypes.c_int, [_vp, ctypes.POINTER(_Genz), ctypes.c_int, _sz,
                                            _vp, _vp, _dp, _dp, _vp, _vp, _vp]),
    "pagani_refine_dev": (ctypes.c_int, [_vp, _sz, _vp, _vp, _vp, _vp, ctypes.c_double,
                                         ctypes.c_int, _vp, _dp, ctypes.POINTER(_sz)]),
    "pagani_refine_dev_noalloc": (ctypes.c_int, [_vp, _sz, _vp, _vp, _vp, _vp, ctypes.c_double,
                                                 ctypes.c_int, _vp, _vp, _vp]),
    "pagani_filter_dev": (ctypes.c_int, [_vp, _sz, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                         _vp, _vp, _vp, ctypes.POINTER(_sz)]),
    "pagani_split_dev": (ctypes.c_int, [_vp, _sz, _vp, _vp, _vp, _vp, _vp]),
    "pagani_filter_split_dev": (ctypes.c_int, [_vp, _sz, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                               _vp, _vp, _vp, ctypes.POINTER(_sz)]),
    "pagani_sum_dev": (ctypes.c_int, [_vp, _vp, _sz, _dp]),
    "pagani_dot_dev": (ctypes.c_int, [_vp, _vp, _vp, _sz, _dp]),
    "pagani_minmax_dev": (ctypes.c_int, [_vp, _vp, _sz, _dp, _dp]),
    "pagani_scan_flags_dev": (ctypes.c_int, [_vp, _vp, _sz, _vp, ctypes.POINTER(_sz)]),
    "pagani_classify_dev": (ctypes.c_int, [_vp, _vp, _sz, ctypes.c_double, _dp,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           _vp, ctypes.POINTER(ClassRes)]),
    "do_integration_from_c": (ctypes.c_int, [_dp]),
    "pagani_ws_footprint": (_sz, [_vp]),
    "pagani_ws_trim": (ctypes.c_int, [_vp]),
    "pagani_ws_set_iteration_limit": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "pagani_ws_suspended": (ctypes.c_int, [_vp]),
    "pagani_integrate_cb_resume": (ctypes.c_int, [_vp, _vp, _vp, ctypes.POINTER(_Result)]),
    "pagani_integrate_genz_resume": (ctypes.c_int, [_vp, ctypes.POINTER(_Genz), ctypes.c_int,
                                                    ctypes.POINTER(_Result)]),
    "pagani_ws_save_state": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "pagani_ws_load_state": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "pagani_ws_write_regions_csv": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "pagani_ws_set_snapshot_iter": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "pagani_ws_set_predict_split": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pagani_ws_set_device_search": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pagani_ws_set_level_form": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pagani_ws_set_debug": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p,
                                           ctypes.c_int]),
    "pagani_rule_point": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    "pagani_ws_snapshot_regions": (ctypes.c_int, [_vp, ctypes.POINTER(_sz), ctypes.POINTER(_vp),
                                                  ctypes.POINTER(_vp)]),
    "pagani_batch_create": (_vp, [ctypes.c_int, ctypes.c_int]),
    "pagani_batch_destroy": (None, [_vp]),
    "pagani_batch_last_error": (ctypes.c_char_p, [_vp]),
    "pagani_batch_set_device_mem": (ctypes.c_int, [_vp, _sz]),
    "pagani_batch_set_mem_model": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pagani_batch_set_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pagani_batch_set_tuned_slices": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pagani_batch_set_segments": (ctypes.c_int, [_vp, ctypes.c_size_t]),
    "pagani_batch_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(Stats)]),
    "pagani_batch_reset_stats": (ctypes.c_int, [_vp]),
    "pagani_integrate_batch": (ctypes.c_int, [_vp, ctypes.POINTER(_BatchItem), _sz,
                                              ctypes.POINTER(_Result),
                                              ctypes.POINTER(ctypes.c_int)]),
    "pagani_mc_create": (_vp, [ctypes.c_int, ctypes.c_int]),
    "pagani_mc_destroy": (None, [_vp]),
    "pagani_mc_last_error": (ctypes.c_char_p, [_vp]),
    "pagani_mc_set_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pagani_mc_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(McStats)]),
    "pagani_mc_reset_stats": (ctypes.c_int, [_vp]),
    "pagani_mc_integrate_cb": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, ctypes.c_double,
                                              _dp, _dp, ctypes.POINTER(McParams),
                                              ctypes.POINTER(McResult)]),
    "pagani_mc_integrate_genz": (ctypes.c_int, [_vp, ctypes.POINTER(_Genz), ctypes.c_double,
                                                ctypes.c_double, _dp, _dp,
                                                ctypes.POINTER(McParams),
                                                ctypes.POINTER(McResult)]),
    "pagani_mc_trace": (ctypes.c_int, [_vp, _dp, _dp, _sz, ctypes.POINTER(_sz)]),
    "pagani_mc_grid": (ctypes.c_int, [_vp, _dp]),
    "pagani_mc_sample_genz": (ctypes.c_int, [_vp, ctypes.POINTER(_Genz), ctypes.c_double, _dp,
                                             _dp, _dp, ctypes.c_uint64, ctypes.c_int,
                                             ctypes.c_int, _dp, _dp, _dp]),
}

_LIB = None


def load():
    """Load libpagani.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libpagani.so not built at {LIB_PATH}; run __graft_entry__.build() "
            "(the HIP library is required, there is no CPU fallback)")
    # One HIP runtime per process: when torch is present it brings its own
    # libamdhip64; load it first so libpagani.so binds to the same runtime
    # (two runtimes in one process cannot both initialise the device).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def exported_symbols():
    return list(_SIGS)


def rule_tables(ndim):
    """Host-side rule tables of the library: (w, scale, norm, lambda, set_begin, gtab)."""
    import numpy as np
    lib = load()
    w = np.zeros(45)
    sc = np.zeros(45)
    nm = np.zeros(45)
    lam = np.zeros(5)
    sb = (ctypes.c_int * (NSETS + 1))()
    F = lib.pagani_rule_tables(int(ndim), None, None, None, None, None, None)
    if F < 0:
        raise PaganiError(F, f"unsupported ndim {ndim}")
    g = np.zeros((F - 1 - 4 * ndim) * ndim)
    lib.pagani_rule_tables(int(ndim), w.ctypes.data_as(_dp), sc.ctypes.data_as(_dp),
                           nm.ctypes.data_as(_dp), lam.ctypes.data_as(_dp), sb,
                           g.ctypes.data_as(_dp))
    return (w.reshape(9, 5), sc.reshape(9, 5), nm.reshape(9, 5), lam, list(sb),
            g.reshape(-1, ndim))


def rule_point(ndim, p, axis):
    """Generator coordinate `axis` of rule point p in the library's point
    order (the point of a region with centre c and width w is c - g w)."""
    g = ctypes.c_double(0.0)
    rc = load().pagani_rule_point(int(ndim), int(p), int(axis), ctypes.byref(g))
    if rc != 0:
        raise PaganiError(rc, "pagani_rule_point")
    return g.value


@dataclass
class IntegrationResult:
    """numint::integration_result (common/integration_result.hh:11-21)."""
    estimate: float
    errorest: float
    neval: int
    nregions: int
    nFinishedRegions: int
    status: int
    lastPhase: int
    chi_sq: float
    iters: int

    @classmethod
    def _from(cls, r):
        return cls(r.estimate, r.errorest, r.neval, r.nregions, r.nFinishedRegions, r.status,
                   r.lastPhase, r.chi_sq, r.iters)


def _arr(vals, n=None):
    if vals is None:
        return None
    vals = [float(v) for v in vals]
    return (ctypes.c_double * len(vals))(*vals)


def genz(family, a, u=None):
    fam = FAMILY[family] if isinstance(family, str) else int(family)
    a = [float(v) for v in a]
    d = len(a)
    if not (MIN_DIM <= d <= MAX_DIM):
        raise ValueError(f"ndim {d} outside [{MIN_DIM}, {MAX_DIM}]")
    u = [0.0] * d if u is None else [float(v) for v in u]
    g = _Genz()
    g.family = fam
    g.ndim = d
    for i in range(d):
        g.a[i] = a[i]
        g.u[i] = u[i]
    return g


def _ptr(t):
    """Device pointer of a torch tensor (None passes NULL)."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


class Workspace:
    """Owns a pagani_ws: rule tables, region pool and stream for one ndim."""

    def __init__(self, ndim, device=-1):
        self.lib = load()
        self.ndim = int(ndim)
        self._ws = self.lib.pagani_ws_create(self.ndim, int(device))
        if not self._ws:
            why = self.lib.pagani_create_error()
            raise PaganiError(EHIP, f"pagani_ws_create({ndim}) failed: "
                                    f"{why.decode() if why else 'unknown'}")

    def close(self):
        if getattr(self, "_ws", None):
            self.lib.pagani_ws_destroy(self._ws)
            self._ws = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- error handling --------------------------------------------------------
    def _check(self, rc):
        if rc < 0:
            msg = self.lib.pagani_ws_last_error(self._ws)
            raise PaganiError(rc, msg.decode() if msg else "")
        return rc

    # -- configuration -------------------------------------------------------
    def set_device_mem(self, nbytes):
        self._check(self.lib.pagani_ws_set_device_mem(self._ws, int(nbytes)))

    def device_mem(self):
        return self.lib.pagani_ws_device_mem(self._ws)

    def set_mem_model(self, model):
        """0 / "reference": the reference's 8(10d+8)+16 B per region;
        1 / "actual": this library's own footprint."""
        m = {"reference": 0, "actual": 1}.get(model, model)
        self._check(self.lib.pagani_ws_set_mem_model(self._ws, int(m)))

    def set_device_loop(self, on=True):
        """Run small iterations on the device in rounds (default on)."""
        self._check(self.lib.pagani_ws_set_device_loop(self._ws, 1 if on else 0))

    def set_slices(self, s):
        self._check(self.lib.pagani_ws_set_slices(self._ws, int(s)))

    def set_graphs(self, on=True):
        """Replay full-length device-loop rounds from HIP graphs (default on)."""
        self._check(self.lib.pagani_ws_set_graphs(self._ws, 1 if on else 0))

    def set_tuned_slices(self, on=True):
        """Launch shapes from the autotuner's table where it has an entry
        (default on); off: the fixed rule."""
        self._check(self.lib.pagani_ws_set_tuned_slices(self._ws, 1 if on else 0))

    def set_timing(self, on=True):
        self._check(self.lib.pagani_ws_set_timing(self._ws, 1 if on else 0))

    def stats(self):
        s = Stats()
        self._check(self.lib.pagani_ws_get_stats(self._ws, ctypes.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        self._check(self.lib.pagani_ws_reset_stats(self._ws))

    def eval_consts(self, n=1 << 20, lows=None, highs=None):
        c = EvalConsts()
        self._check(self.lib.pagani_ws_eval_consts(self._ws, _arr(lows), _arr(highs), int(n),
                                                   ctypes.byref(c)))
        return c

    def trace(self, cap=1000):
        import numpy as np
        n = np.zeros(cap, dtype=np.uint64)
        m = np.zeros(cap, dtype=np.uint64)
        e = np.zeros(cap)
        r = np.zeros(cap)
        k = self.lib.pagani_ws_trace(self._ws, cap, n.ctypes.data, m.ctypes.data, e.ctypes.data,
                                     r.ctypes.data)
        k = min(k, cap)
        return dict(n=n[:k].astype(np.int64), m=m[:k].astype(np.int64), estimate=e[:k],
                    errorest=r[:k])

    # -- integration ---------------------------------------------------------
    def integrate_genz(self, family, a, u=None, epsrel=1e-3, epsabs=1e-12, lows=None,
                       highs=None, precision=FP64, relerr=True, init_k=0):
        g = genz(family, a, u)
        if g.ndim != self.ndim:
            raise ValueError("integrand ndim != workspace ndim")
        res = _Result()
        rc = self.lib.pagani_integrate_genz(self._ws, ctypes.byref(g), int(precision),
                                            float(epsrel), float(epsabs), _arr(lows),
                                            _arr(highs), 1 if relerr else 0, int(init_k),
                                            ctypes.byref(res))
        if rc < 0 and rc != ENOMEM:
            self._check(rc)
        return IntegrationResult._from(res)

    def footprint(self):
        """Device bytes held by the region pool."""
        return self.lib.pagani_ws_footprint(self._ws)

    def trim(self):
        """Release the region pool (it regrows on the next call)."""
        self._check(self.lib.pagani_ws_trim(self._ws))

    # -- suspend / checkpoint / resume ----------------------------------------
    def set_iteration_limit(self, k):
        """Suspend integrate/resume calls after k iterations (0: no limit)."""
        self._check(self.lib.pagani_ws_set_iteration_limit(self._ws, int(k)))

    def suspended(self):
        return bool(self.lib.pagani_ws_suspended(self._ws))

    def resume_genz(self, family, a, u=None, precision=FP64):
        """Continue the held run; returns (result, suspended_again)."""
        g = genz(family, a, u)
        res = _Result()
        rc = self.lib.pagani_integrate_genz_resume(self._ws, ctypes.byref(g), int(precision),
                                                   ctypes.byref(res))
        if rc < 0 and rc != ENOMEM:
            self._check(rc)
        return IntegrationResult._from(res), rc == SUSPENDED

    def save_state(self, path):
        self._check(self.lib.pagani_ws_save_state(self._ws, os.fsencode(path)))

    def load_state(self, path):
        self._check(self.lib.pagani_ws_load_state(self._ws, os.fsencode(path)))

    def write_regions_csv(self, path):
        self._check(self.lib.pagani_ws_write_regions_csv(self._ws, os.fsencode(path)))

    def set_snapshot_iter(self, it):
        """Keep a copy of the list evaluated in iteration `it` when no region
        has finished by then (None: never)."""
        v = (1 << 64) - 1 if it is None else int(it)
        self._check(self.lib.pagani_ws_set_snapshot_iter(self._ws, v))

    def set_predict_split(self, on=True):
        """Relerr classification by the predict_split rule in every iteration
        (off before iteration 15 while nothing finished and the list is at
        most 15M regions, on after), instead of the relerr argument."""
        self._check(self.lib.pagani_ws_set_predict_split(self._ws, 1 if on else 0))

    def set_device_search(self, on=True):
        """Run the classifier's threshold search on the device (default) or,
        with on=False, on the host (one probe pass per round trip)."""
        self._check(self.lib.pagani_ws_set_device_search(self._ws, 1 if on else 0))

    def set_level_form(self, on=True):
        """Keep the uniform start's widths as levels (two bytes per axis, the
        default) or, with on=False, as doubles; the results are the same."""
        self._check(self.lib.pagani_ws_set_level_form(self._ws, 1 if on else 0))

    def set_debug(self, level, directory=".", id="default", append=False):
        """The reference's debug recorders for the following integrate calls:
        level 1 writes the per-iteration CSV (and, with an id, the per-phase
        time breakdown), 2 also every region's estimate and error, 3 also the
        generator table and, for lists under 1024 regions, every rule point.
        id=None gives the plain integrate overload's files (cuda_iters.csv);
        append=True appends without header lines (collect_mult_runs)."""
        self._check(self.lib.pagani_ws_set_debug(
            self._ws, int(level), os.fsencode(directory),
            None if id is None else id.encode(), 1 if append else 0))

    def snapshot_size(self):
        n = _sz(0)
        L = _vp()
        H = _vp()
        self._check(self.lib.pagani_ws_snapshot_regions(self._ws, ctypes.byref(n),
                                                        ctypes.byref(L), ctypes.byref(H)))
        return n.value

    # -- phases on torch device tensors ----------------------------------------
    def uniform_split(self, k, L, H):
        self._check(self.lib.pagani_uniform_split_dev(self._ws, int(k), _ptr(L), _ptr(H)))

    def eval_genz(self, family, a, u, n, L, H, est, err, dim, lows=None, highs=None,
                  precision=FP64):
        g = genz(family, a, u)
        self._check(self.lib.pagani_eval_genz_dev(self._ws, ctypes.byref(g), int(precision),
                                                  int(n), _ptr(L), _ptr(H), _arr(lows),
                                                  _arr(highs), _ptr(est), _ptr(err), _ptr(dim)))

    def refine(self, n, est, err_in, err_out, parent_est, epsrel, relerr, active):
        sums = (ctypes.c_double * 4)()
        na = ctypes.c_size_t(0)
        self._check(self.lib.pagani_refine_dev(self._ws, int(n), _ptr(est), _ptr(err_in),
                                               _ptr(err_out), _ptr(parent_est), float(epsrel),
                                               1 if relerr else 0, _ptr(active), sums,
                                               ctypes.byref(na)))
        return list(sums), na.value

    def filter(self, n, L, H, active, est, err, dim, L2, H2, pest, perr, dim2):
        m = ctypes.c_size_t(0)
        self._check(self.lib.pagani_filter_dev(self._ws, int(n), _ptr(L), _ptr(H), _ptr(active),
                                               _ptr(est), _ptr(err), _ptr(dim), _ptr(L2),
                                               _ptr(H2), _ptr(pest), _ptr(perr), _ptr(dim2),
                                               ctypes.byref(m)))
        return m.value

    def split(self, m, L, H, dim, L2, H2):
        self._check(self.lib.pagani_split_dev(self._ws, int(m), _ptr(L), _ptr(H), _ptr(dim),
                                              _ptr(L2), _ptr(H2)))

    def filter_split(self, n, L, H, active, est, err, dim, L2, H2, pest, perr):
        m = ctypes.c_size_t(0)
        self._check(self.lib.pagani_filter_split_dev(self._ws, int(n), _ptr(L), _ptr(H),
                                                     _ptr(active), _ptr(est), _ptr(err),
                                                     _ptr(dim), _ptr(L2), _ptr(H2), _ptr(pest),
                                                     _ptr(perr), ctypes.byref(m)))
        return m.value

    def sum(self, x, n=None):
        out = ctypes.c_double(0)
        n = x.numel() if n is None else n
        self._check(self.lib.pagani_sum_dev(self._ws, _ptr(x), int(n), ctypes.byref(out)))
        return out.value

    def dot(self, flags, x, n=None):
        out = ctypes.c_double(0)
        n = x.numel() if n is None else n
        self._check(self.lib.pagani_dot_dev(self._ws, _ptr(flags), _ptr(x), int(n),
                                            ctypes.byref(out)))
        return out.value

    def minmax(self, x, n=None):
        mn, mx = ctypes.c_double(0), ctypes.c_double(0)
        n = x.numel() if n is None else n
        self._check(self.lib.pagani_minmax_dev(self._ws, _ptr(x), int(n), ctypes.byref(mn),
                                               ctypes.byref(mx)))
        return mn.value, mx.value

    def scan_flags(self, flags, offsets, n=None):
        tot = ctypes.c_size_t(0)
        n = flags.numel() if n is None else n
        self._check(self.lib.pagani_scan_flags_dev(self._ws, _ptr(flags), int(n), _ptr(offsets),
                                                   ctypes.byref(tot)))
        return tot.value

    def classify(self, err, n, epsrel, est_hist, iter_errorest, iter_finished_errorest,
                 total_finished_errorest, flags):
        res = ClassRes()
        self._check(self.lib.pagani_classify_dev(self._ws, _ptr(err), int(n), float(epsrel),
                                                 _arr(est_hist), float(iter_errorest),
                                                 float(iter_finished_errorest),
                                                 float(total_finished_errorest), _ptr(flags),
                                                 ctypes.byref(res)))
        return {k: getattr(res, k) for k, _ in ClassRes._fields_}


class Batch:
    """Owns a pagani_batch: independent Genz integrals integrated from one call
    by `workers` host threads with their own streams and workspaces."""

    def __init__(self, device=-1, workers=1):
        self.lib = load()
        self._b = self.lib.pagani_batch_create(int(device), int(workers))
        if not self._b:
            why = self.lib.pagani_batch_last_error(None)
            raise PaganiError(EHIP, f"pagani_batch_create failed: "
                                    f"{why.decode() if why else 'unknown'}")

    def close(self):
        if getattr(self, "_b", None):
            self.lib.pagani_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc < 0:
            msg = self.lib.pagani_batch_last_error(self._b)
            raise PaganiError(rc, msg.decode() if msg else "")
        return rc

    def set_device_mem(self, nbytes):
        self._check(self.lib.pagani_batch_set_device_mem(self._b, int(nbytes)))

    def set_mem_model(self, model):
        m = {"reference": 0, "actual": 1}.get(model, model)
        self._check(self.lib.pagani_batch_set_mem_model(self._b, int(m)))

    def set_timing(self, on=True):
        self._check(self.lib.pagani_batch_set_timing(self._b, 1 if on else 0))

    def set_tuned_slices(self, on=True):
        self._check(self.lib.pagani_batch_set_tuned_slices(self._b, 1 if on else 0))

    def set_segments(self, regions):
        """The persistent kernel (pagani_batch_set_segments): FP64 items whose
        initial split fits `regions` regions run whole, one workgroup each, in
        one launch per (dimension, family); 0 turns it off."""
        self._check(self.lib.pagani_batch_set_segments(self._b, int(regions)))

    def stats(self):
        s = Stats()
        self._check(self.lib.pagani_batch_get_stats(self._b, ctypes.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        self._check(self.lib.pagani_batch_reset_stats(self._b))

    def integrate(self, items, precision=FP64, relerr=True, init_k=0, epsabs=1e-20):
        """items: sequence of (family, a, u, epsrel) or (family, a, u, epsrel,
        epsabs).  Returns (results, codes); raises on a hard error (a negative
        code other than ENOMEM, which still fills the item's result)."""
        n = len(items)
        arr = (_BatchItem * max(n, 1))()
        for i, it in enumerate(items):
            fam, a, u, eps = it[:4]
            arr[i].f = genz(fam, a, u)
            arr[i].epsrel = float(eps)
            arr[i].epsabs = float(it[4]) if len(it) > 4 else float(epsabs)
            arr[i].precision = int(precision)
            arr[i].relerr_classification = 1 if relerr else 0
            arr[i].init_k = int(init_k)
        res = (_Result * max(n, 1))()
        codes = (ctypes.c_int * max(n, 1))()
        rc = self.lib.pagani_integrate_batch(self._b, arr, n, res, codes)
        if rc < 0 and rc != ENOMEM:
            self._check(rc)
        return [IntegrationResult._from(res[i]) for i in range(n)], list(codes)[:n]


class MCubes:
    """Owns a pagani_mc handle: the m-Cubes VEGAS integrator for one
    dimension (cuda/mcubes/vegasT.cuh of the reference)."""

    def __init__(self, ndim, device=-1):
        self.lib = load()
        self.ndim = int(ndim)
        self._h = self.lib.pagani_mc_create(self.ndim, int(device))
        if not self._h:
            why = self.lib.pagani_mc_last_error(None)
            raise PaganiError(EHIP, f"pagani_mc_create failed: "
                                    f"{why.decode() if why else 'unknown'}")

    def close(self):
        if getattr(self, "_h", None):
            self.lib.pagani_mc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc < 0:
            msg = self.lib.pagani_mc_last_error(self._h)
            raise PaganiError(rc, msg.decode() if msg else "")
        return rc

    def set_timing(self, on=True):
        self._check(self.lib.pagani_mc_set_timing(self._h, 1 if on else 0))

    def stats(self):
        st = McStats()
        self._check(self.lib.pagani_mc_get_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in McStats._fields_}

    def reset_stats(self):
        self._check(self.lib.pagani_mc_reset_stats(self._h))

    @staticmethod
    def params(ncall, total_iters=15, adjust_iters=15, skip_iters=5, generator="default",
               one_axis=False, adjust_params=False, seed=0):
        return McParams(float(ncall), int(total_iters), int(adjust_iters), int(skip_iters),
                        MC_RNG.get(generator, generator), int(one_axis), int(adjust_params),
                        int(seed))

    def integrate_genz(self, family, a, u=None, epsrel=1e-3, epsabs=1e-20, ncall=1e6,
                       total_iters=15, adjust_iters=15, skip_iters=5, generator="default",
                       one_axis=False, adjust_params=False, seed=0, lows=None, highs=None):
        """Returns a dict with estimate, errorest, chi_sq, status, iters, runs,
        neval, ncall."""
        g = genz(family, a, u)
        p = self.params(ncall, total_iters, adjust_iters, skip_iters, generator, one_axis,
                        adjust_params, seed)
        r = McResult()
        self._check(self.lib.pagani_mc_integrate_genz(self._h, ctypes.byref(g), float(epsrel),
                                                      float(epsabs), _arr(lows, self.ndim),
                                                      _arr(highs, self.ndim), ctypes.byref(p),
                                                      ctypes.byref(r)))
        return {k: getattr(r, k) for k, _ in McResult._fields_ if k != "pad"}

    def trace(self, cap=1000):
        import numpy as np
        ti = np.zeros(cap)
        tsi = np.zeros(cap)
        n = _sz(0)
        self._check(self.lib.pagani_mc_trace(self._h, ti.ctypes.data_as(_dp),
                                             tsi.ctypes.data_as(_dp), cap, ctypes.byref(n)))
        k = min(n.value, cap)
        return ti[:k], tsi[:k]

    def grid(self):
        import numpy as np
        xi = np.zeros(self.ndim * (MC_BINS + 1))
        self._check(self.lib.pagani_mc_grid(self._h, xi.ctypes.data_as(_dp)))
        return xi

    def sample_genz(self, family, a, u, ncall, xi, key, generator="lcg", hist_dims=0,
                    lows=None, highs=None):
        """One sampling pass on grid xi: (ti, tsi unscaled, hist)."""
        import numpy as np
        g = genz(family, a, u)
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        ti = ctypes.c_double()
        tsi = ctypes.c_double()
        hist = np.zeros(max(hist_dims, 1) * MC_BINS)
        self._check(self.lib.pagani_mc_sample_genz(
            self._h, ctypes.byref(g), float(ncall), _arr(lows, self.ndim),
            _arr(highs, self.ndim), xi.ctypes.data_as(_dp), int(key),
            MC_RNG.get(generator, generator), int(hist_dims), ctypes.byref(ti),
            ctypes.byref(tsi), hist.ctypes.data_as(_dp)))
        return ti.value, tsi.value, hist[: hist_dims * MC_BINS]


def do_integration_from_c():
    """The reference's C interop shim: (status, estimate)."""
    lib = load()
    out = ctypes.c_double(0)
    rc = lib.do_integration_from_c(ctypes.byref(out))
    return rc, out.value
