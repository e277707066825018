"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.  The oracle is the checker, never the
thing measured or shipped.
"""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_GENZ_SO = os.path.join(ROOT, "oracle", "_ref", "libref_genz.so")

FAM = {
    "oscillatory": 1,
    "product_peak": 2,
    "corner_peak": 3,
    "gaussian": 4,
    "c0": 5,
    "discontinuous": 6,
    "const": 10,
    "product": 11,
    "fun6": 12,
    "box": 13,
    "nan": 14,
    "monomial": 15,
}

# per-region extra record of orc_eval_regions: s_0..s_4, A_0..A_4 (sum |w_r f|),
# volume, e_1, e_2, e_3, fourth-difference magnitude, and (kernel-point mode)
# the sensitivity of f to one rounding of its inner sum, sum_p |w_0 J| c_p
RSUM = 16

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint8)
_sp = ctypes.POINTER(ctypes.c_size_t)


class OrcResult(ctypes.Structure):
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


class OrcTrace(ctypes.Structure):
    _fields_ = [
        ("cap", ctypes.c_size_t),
        ("len", ctypes.c_size_t),
        ("n", _sp),
        ("m", _sp),
        ("estimate", _dp),
        ("errorest", _dp),
    ]


class OrcOpts(ctypes.Structure):
    _fields_ = [
        ("predict_split", ctypes.c_int),
        ("max_iters", ctypes.c_size_t),
        ("keep_list", ctypes.c_int),
        ("list_n", ctypes.c_size_t),
        ("list_L", _dp),
        ("list_H", _dp),
        ("list_est", _dp),
        ("list_err", _dp),
        ("list_dim", _ip),
        ("budget_fixes", ctypes.c_size_t),
    ]


class OrcClassifier(ctypes.Structure):
    _fields_ = [
        ("epsrel", ctypes.c_double),
        ("epsabs", ctypes.c_double),
        ("required_digits", ctypes.c_int),
        ("hist", ctypes.c_double * 3),
        ("iters_collected", ctypes.c_size_t),
        ("max_budget", ctypes.c_double),
        ("max_active", ctypes.c_double),
        ("total_mem", ctypes.c_size_t),
        ("ndim", ctypes.c_int),
    ]


class OrcClassRes(ctypes.Structure):
    _fields_ = [
        ("pass_mem", ctypes.c_int),
        ("pass_errorest_budget", ctypes.c_int),
        ("threshold", ctypes.c_double),
        ("errorest_budget_covered", ctypes.c_double),
        ("percent_mem_active", ctypes.c_double),
        ("range_low", ctypes.c_double),
        ("range_high", ctypes.c_double),
        ("num_active", ctypes.c_size_t),
        ("finished_errorest", ctypes.c_double),
        ("max_budget_perc_to_cover", ctypes.c_double),
        ("max_active_perc", ctypes.c_double),
        ("threshold_passes", ctypes.c_int),
    ]


MC_BINS = 500
MC_MAXDIM = 16
MC_RNG = {"default": 0, "lcg": 1}


class McRng(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("lcg", ctypes.c_uint32), ("s0", ctypes.c_uint64),
                ("s1", ctypes.c_uint64)]


class McParams(ctypes.Structure):
    _fields_ = [("ncall", ctypes.c_double), ("total_iters", ctypes.c_int),
                ("adjust_iters", ctypes.c_int), ("skip_iters", ctypes.c_int),
                ("generator", ctypes.c_int), ("one_axis", ctypes.c_int),
                ("adjust_params", ctypes.c_int), ("seed", ctypes.c_uint64)]


class McSetup(ctypes.Structure):
    _fields_ = [("ng", ctypes.c_uint32), ("npg", ctypes.c_uint32), ("ncubes", ctypes.c_double),
                ("calls", ctypes.c_double), ("dv2g", ctypes.c_double), ("dxg", ctypes.c_double),
                ("xnd", ctypes.c_double), ("xjac", ctypes.c_double),
                ("lo", ctypes.c_double * 16), ("dx", ctypes.c_double * 16)]


class McResult(ctypes.Structure):
    _fields_ = [("estimate", ctypes.c_double), ("errorest", ctypes.c_double),
                ("chi_sq", ctypes.c_double), ("status", ctypes.c_int),
                ("iters", ctypes.c_size_t), ("runs", ctypes.c_size_t), ("ncall", ctypes.c_double)]


class McTrace(ctypes.Structure):
    _fields_ = [("cap", ctypes.c_size_t), ("len", ctypes.c_size_t), ("ti", _dp), ("tsi", _dp)]


def _load():
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError(f"oracle not built: {ORACLE_SO} (run __graft_entry__.build())")
    lib = ctypes.CDLL(ORACLE_SO)
    lib.orc_feval.argtypes = [ctypes.c_int]
    lib.orc_feval.restype = ctypes.c_int
    lib.orc_rule_tables.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _ip]
    lib.orc_generators.argtypes = [ctypes.c_int, _dp, _ip]
    lib.orc_integrand.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
    lib.orc_integrand.restype = ctypes.c_double
    lib.orc_exact.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
    lib.orc_exact.restype = ctypes.c_double
    lib.orc_eval_regions.argtypes = [
        ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_size_t, _dp, _dp,
        _dp, _dp, _ip, _dp, _dp]
    lib.orc_uniform_split.argtypes = [ctypes.c_int, ctypes.c_size_t, _dp, _dp]
    lib.orc_uniform_split.restype = None
    lib.orc_refine.argtypes = [ctypes.c_size_t, _dp, _dp, _dp, ctypes.c_double,
                               ctypes.c_int, _up]
    lib.orc_refine.restype = None
    lib.orc_filter.argtypes = [ctypes.c_int, ctypes.c_size_t, _dp, _dp, _up, _dp,
                               _dp, _ip, _dp, _dp, _dp, _dp, _ip]
    lib.orc_filter.restype = ctypes.c_size_t
    lib.orc_split.argtypes = [ctypes.c_int, ctypes.c_size_t, _dp, _dp, _ip, _dp, _dp]
    lib.orc_split.restype = None
    lib.orc_clf_init.argtypes = [ctypes.POINTER(OrcClassifier), ctypes.c_int,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_size_t]
    lib.orc_clf_init.restype = None
    lib.orc_clf_store.argtypes = [ctypes.POINTER(OrcClassifier), ctypes.c_double]
    lib.orc_clf_store.restype = None
    lib.orc_clf_sig_digits_same.argtypes = [ctypes.POINTER(OrcClassifier)]
    lib.orc_clf_estimate_converged.argtypes = [ctypes.POINTER(OrcClassifier)]
    lib.orc_clf_criteria_met.argtypes = [ctypes.POINTER(OrcClassifier), ctypes.c_size_t]
    lib.orc_mem_required.argtypes = [ctypes.c_size_t, ctypes.c_int]
    lib.orc_mem_required.restype = ctypes.c_size_t
    lib.orc_free_mem.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t]
    lib.orc_free_mem.restype = ctypes.c_size_t
    lib.orc_clf_classify.argtypes = [
        ctypes.POINTER(OrcClassifier), _dp, ctypes.c_size_t, ctypes.c_double,
        ctypes.c_double, ctypes.c_double, _up, ctypes.POINTER(OrcClassRes)]
    lib.orc_clf_classify.restype = None
    lib.orc_integrate.argtypes = [
        ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double, _dp,
        _dp, ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int,
        ctypes.POINTER(OrcResult), ctypes.POINTER(OrcTrace)]
    lib.orc_integrate_ex.argtypes = [
        ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double, _dp,
        _dp, ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int,
        ctypes.POINTER(OrcResult), ctypes.POINTER(OrcTrace), ctypes.POINTER(OrcOpts)]
    lib.orc_free.argtypes = [ctypes.c_void_p]
    lib.orc_free.restype = None
    lib.orc_eval_regions_kform.argtypes = [
        ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, ctypes.c_size_t, _dp, _dp,
        _dp, _dp, _ip, _dp, _dp]
    lib.orc_cuhre_seq.argtypes = [
        ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double,
        ctypes.c_size_t, ctypes.POINTER(OrcResult)]
    # m-Cubes (mcubes_oracle.h)
    lib.orc_mc_lcg_step.argtypes = [ctypes.c_uint32]
    lib.orc_mc_lcg_step.restype = ctypes.c_uint32
    lib.orc_mc_splitmix64.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
    lib.orc_mc_splitmix64.restype = ctypes.c_uint64
    lib.orc_mc_iter_key.argtypes = [ctypes.c_uint64, ctypes.c_int]
    lib.orc_mc_iter_key.restype = ctypes.c_uint64
    lib.orc_mc_rng_init.argtypes = [ctypes.POINTER(McRng), ctypes.c_int, ctypes.c_uint64,
                                    ctypes.c_uint64]
    lib.orc_mc_rng_init.restype = None
    lib.orc_mc_rng_fill.argtypes = [ctypes.POINTER(McRng), ctypes.c_int, _dp]
    lib.orc_mc_rng_fill.restype = None
    lib.orc_mc_get_indx.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint32),
                                    ctypes.c_int, ctypes.c_uint32]
    lib.orc_mc_get_indx.restype = None
    lib.orc_mc_rebin.argtypes = [ctypes.c_double, ctypes.c_int, _dp, _dp, _dp]
    lib.orc_mc_rebin.restype = None
    lib.orc_mc_grid_init.argtypes = [ctypes.c_int, _dp]
    lib.orc_mc_grid_init.restype = None
    lib.orc_mc_refine.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
    lib.orc_mc_refine.restype = None
    lib.orc_mc_setup_init.argtypes = [ctypes.POINTER(McSetup), ctypes.c_int, ctypes.c_double,
                                      _dp, _dp]
    lib.orc_mc_setup_init.restype = None
    lib.orc_mc_sample.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.POINTER(McSetup), _dp,
                                  ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
    lib.orc_mc_sample.restype = None
    lib.orc_mc_integrate.argtypes = [
        ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp,
        ctypes.POINTER(McParams), ctypes.POINTER(McResult), ctypes.POINTER(McTrace), _dp]
    lib.orc_num_threads.restype = ctypes.c_int
    lib.orc_set_num_threads.argtypes = [ctypes.c_int]
    lib.orc_set_num_threads.restype = None
    return lib


LIB = _load()


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def uptr(a):
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_up)


def feval(d):
    return LIB.orc_feval(d)


def rule_tables(d):
    w = np.zeros(45)
    lam = np.zeros(5)
    sc = np.zeros(45)
    nm = np.zeros(45)
    cnt = np.zeros(9, dtype=np.int32)
    assert LIB.orc_rule_tables(d, dptr(w), dptr(lam), dptr(sc), dptr(nm), iptr(cnt)) == 0
    return w.reshape(9, 5), lam, sc.reshape(9, 5), nm.reshape(9, 5), cnt


def generators(d):
    F = feval(d)
    g = np.zeros(F * d)
    s = np.zeros(F, dtype=np.int32)
    assert LIB.orc_generators(d, dptr(g), iptr(s)) == 0
    return g.reshape(F, d), s


def params(a, u=None):
    a = np.asarray(a, dtype=np.float64)
    if u is None:
        u = np.zeros_like(a)
    return np.ascontiguousarray(np.concatenate([a, np.asarray(u, dtype=np.float64)]))


def integrand(fam, d, prm, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return LIB.orc_integrand(FAM.get(fam, fam), d, dptr(prm), dptr(x))


def exact(fam, d, prm):
    return LIB.orc_exact(FAM.get(fam, fam), d, dptr(prm))


def uniform_split(d, k):
    n = k ** d
    L = np.zeros(n * d)
    H = np.zeros(n * d)
    LIB.orc_uniform_split(d, k, dptr(L), dptr(H))
    return L, H, n


def eval_regions(d, fam, prm, L, H, n, lows=None, highs=None, want_extra=False):
    est = np.zeros(n)
    err = np.zeros(n)
    dim = np.zeros(n, dtype=np.int32)
    fd = np.zeros(n * d) if want_extra else None
    rs = np.zeros(n * RSUM) if want_extra else None
    lw = None if lows is None else np.ascontiguousarray(lows, dtype=np.float64)
    hg = None if highs is None else np.ascontiguousarray(highs, dtype=np.float64)
    rc = LIB.orc_eval_regions(d, FAM.get(fam, fam), dptr(prm), dptr(lw), dptr(hg), n,
                              dptr(L), dptr(H), dptr(est), dptr(err), iptr(dim),
                              dptr(fd), dptr(rs))
    assert rc == 0
    if want_extra:
        return est, err, dim, fd.reshape(n, d), rs.reshape(n, RSUM)
    return est, err, dim


def eval_regions_kform(d, fam, prm, L, H, n, lows=None, highs=None, plain=False):
    """Per-region evaluation at the library kernel's own points (centre,
    origin, axis scale, fma(-g, h, c)) with f and the sums in long double.
    Returns est, err, dim, fourth differences, the per-region record."""
    est = np.zeros(n)
    err = np.zeros(n)
    dim = np.zeros(n, dtype=np.int32)
    fd = np.zeros(n * d)
    rs = np.zeros(n * RSUM)
    lw = None if lows is None else np.ascontiguousarray(lows, dtype=np.float64)
    hg = None if highs is None else np.ascontiguousarray(highs, dtype=np.float64)
    rc = LIB.orc_eval_regions_kform(d, FAM.get(fam, fam), dptr(prm), int(plain), dptr(lw),
                                    dptr(hg), n, dptr(L), dptr(H), dptr(est), dptr(err),
                                    iptr(dim), dptr(fd), dptr(rs))
    assert rc == 0
    return est, err, dim, fd.reshape(n, d), rs.reshape(n, RSUM)


def eval_floor(rs):
    """vol * sum_p |w_0 J| c_p: the region's estimate moves by this much
    (times 2^-53 per rounding) for one rounding of the inner sum of f at
    every point -- the least any FP64 evaluation of the integrand differs
    from the long-double one (kernel-point mode only)."""
    return rs[:, 10] * rs[:, 15]


def est_scale(rs):
    """vol * sum_p |w_0 f_p|: the magnitude that bounds the rounding of a
    region's estimate."""
    return rs[:, 10] * rs[:, 5]


def err_bounds(d, rs, rel=1e-12):
    """Rounding bounds of a region's error estimate from the oracle's record.

    Returns (value_a, value_b, tol_a, tol_b, cond, ambiguous) where value_a =
    vol*e1 is used when 5e1 <= e2 and 5e2 <= e3 (cond), value_b = 5 vol
    max(e1, e2, e3) otherwise (Sample.cuh:282-288); tol_* bound their rounding
    given |ds_r| <= rel * A_r on every rule sum, and `ambiguous` marks regions
    whose branch condition lies within that rounding."""
    _, _, sc, nm, _ = rule_tables(d)
    A = rs[:, 5:10]
    vol = rs[:, 10]
    e = rs[:, 11:14]
    dA = rel * A
    de = np.zeros_like(e)
    for r in (1, 2, 3):
        de[:, r - 1] = np.max(nm[:, r][None, :] * (dA[:, r + 1][:, None] +
                                                   np.abs(sc[:, r])[None, :] * dA[:, r][:, None]),
                              axis=1)
    e1, e2, e3 = e[:, 0], e[:, 1], e[:, 2]
    cond = (5 * e1 <= e2) & (5 * e2 <= e3)
    amb = (np.abs(5 * e1 - e2) <= 5 * de[:, 0] + de[:, 1]) | \
          (np.abs(5 * e2 - e3) <= 5 * de[:, 1] + de[:, 2])
    va = vol * e1
    vb = vol * 5 * np.max(e, axis=1)
    ta = vol * de[:, 0]
    tb = vol * 5 * np.max(de, axis=1)
    return va, vb, ta, tb, cond, amb


def refine(est, err, pest, epsrel, relerr=True):
    n = est.size
    err2 = err.copy()
    act = np.zeros(n, dtype=np.uint8)
    LIB.orc_refine(n, dptr(est), dptr(err2), dptr(pest), epsrel, int(relerr), uptr(act))
    return err2, act


def filter_regions(d, n, L, H, act, est, err, dim):
    m_max = int(act.sum())
    m_alloc = max(m_max, 1)
    L2 = np.zeros(m_alloc * d)
    H2 = np.zeros(m_alloc * d)
    pe = np.zeros(m_alloc)
    pr = np.zeros(m_alloc)
    d2 = np.zeros(m_alloc, dtype=np.int32)
    m = LIB.orc_filter(d, n, dptr(L), dptr(H), uptr(act), dptr(est), dptr(err), iptr(dim),
                       dptr(L2), dptr(H2), dptr(pe), dptr(pr), iptr(d2))
    return m, L2[: m * d], H2[: m * d], pe[:m], pr[:m], d2[:m]


def split_regions(d, m, L, H, dim):
    L2 = np.zeros(2 * m * d)
    H2 = np.zeros(2 * m * d)
    LIB.orc_split(d, m, dptr(L), dptr(H), iptr(dim), dptr(L2), dptr(H2))
    return L2, H2


@dataclass
class Result:
    estimate: float
    errorest: float
    neval: int
    nregions: int
    status: int
    iters: int


def integrate(d, fam, prm, epsrel, epsabs, lows=None, highs=None, relerr=True,
              total_mem=int(16e9), init_k=0, user_regions=False, trace_cap=0,
              predict_split=False, max_iters=0, keep_list=False):
    """The oracle's Workspace::integrate loop.  With keep_list the final list
    comes back as a dict (n, L, H and, when it was evaluated, est/err/dim);
    max_iters > 0 stops before that iteration (status "suspended")."""
    res = OrcResult()
    opts = OrcOpts(int(predict_split), int(max_iters), int(keep_list), 0)
    tr = None
    arrs = None
    if trace_cap:
        arrs = (np.zeros(trace_cap, dtype=np.uint64), np.zeros(trace_cap, dtype=np.uint64),
                np.zeros(trace_cap), np.zeros(trace_cap))
        tr = OrcTrace(trace_cap, 0, arrs[0].ctypes.data_as(_sp), arrs[1].ctypes.data_as(_sp),
                      dptr(arrs[2]), dptr(arrs[3]))
    lw = None if lows is None else np.ascontiguousarray(lows, dtype=np.float64)
    hg = None if highs is None else np.ascontiguousarray(highs, dtype=np.float64)
    rc = LIB.orc_integrate_ex(d, FAM.get(fam, fam), dptr(prm), epsrel, epsabs, dptr(lw),
                              dptr(hg), int(relerr), int(total_mem), init_k, int(user_regions),
                              ctypes.byref(res), ctypes.byref(tr) if tr is not None else None,
                              ctypes.byref(opts))
    assert rc in (0, 2)
    out = Result(res.estimate, res.errorest, res.neval, res.nregions, res.status, res.iters)
    out.suspended = rc == 2
    out.budget_fixes = opts.budget_fixes
    if keep_list:
        n = opts.list_n
        lst = dict(n=n)
        for k, ln in (("L", n * d), ("H", n * d), ("est", n), ("err", n), ("dim", n)):
            p = getattr(opts, "list_" + k)
            lst[k] = np.ctypeslib.as_array(p, shape=(ln,)).copy() if p and ln else (
                None if not p else np.zeros(0))
            LIB.orc_free(ctypes.cast(p, ctypes.c_void_p))
        out.final_list = lst
    if trace_cap:
        k = tr.len
        return out, dict(n=arrs[0][:k].astype(np.int64), m=arrs[1][:k].astype(np.int64),
                         estimate=arrs[2][:k].copy(), errorest=arrs[3][:k].copy())
    return out


def cuhre_seq(d, fam, prm, epsrel, epsabs, maxeval):
    res = OrcResult()
    assert LIB.orc_cuhre_seq(d, FAM.get(fam, fam), dptr(prm), epsrel, epsabs, maxeval,
                             ctypes.byref(res)) == 0
    return Result(res.estimate, res.errorest, res.neval, res.nregions, res.status, res.iters)


def ref_closed_form_lib():
    """The reference's own closed forms, compiled into oracle/_ref (None if absent)."""
    if not os.path.exists(REF_GENZ_SO):
        return None
    lib = ctypes.CDLL(REF_GENZ_SO)
    lib.ref_genz_closed_form.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
    lib.ref_genz_closed_form.restype = ctypes.c_double
    lib.ref_genz_phi.argtypes = [ctypes.c_double]
    lib.ref_genz_phi.restype = ctypes.c_double
    return lib


# ---------------------------------------------------------------------------
# m-Cubes oracle (oracle/mcubes_oracle.c)
# ---------------------------------------------------------------------------

def mc_lcg_step(s):
    return LIB.orc_mc_lcg_step(s)


def mc_splitmix64(seed, count):
    z = ctypes.c_uint64(seed)
    return [LIB.orc_mc_splitmix64(ctypes.byref(z)) for _ in range(count)]


def mc_iter_key(seed, it):
    return LIB.orc_mc_iter_key(seed, it)


def mc_randoms(kind, key, cube, count, per_sample=None):
    """count numbers of one cube's stream, drawn per_sample at a time (one
    sample's axes; default: all in one draw)."""
    g = McRng()
    LIB.orc_mc_rng_init(ctypes.byref(g), MC_RNG.get(kind, kind), key, cube)
    per = per_sample or count
    out = np.zeros(count)
    buf = np.zeros(per)
    for i in range(0, count, per):
        LIB.orc_mc_rng_fill(ctypes.byref(g), per, dptr(buf))
        k = min(per, count - i)
        out[i:i + k] = buf[:k]
    return out


def mc_get_indx(m, d, ng):
    kg = (ctypes.c_uint32 * d)()
    LIB.orc_mc_get_indx(m, kg, d, ng)
    return list(kg)


def mc_rebin(rc, r, xi):
    """r[1..nd] (r[0] ignored), xi[0..nd]; returns the new xi."""
    nd = len(xi) - 1
    r = np.ascontiguousarray(r, dtype=np.float64)
    xi = np.array(xi, dtype=np.float64)
    xin = np.zeros(nd + 1)
    LIB.orc_mc_rebin(rc, nd, dptr(r), dptr(xin), dptr(xi))
    return xi


def mc_grid_init(d):
    xi = np.zeros(d * (MC_BINS + 1))
    LIB.orc_mc_grid_init(d, dptr(xi))
    return xi


def mc_refine(d, hd, hist, xi):
    hist = np.array(hist, dtype=np.float64)
    xi = np.array(xi, dtype=np.float64)
    LIB.orc_mc_refine(d, hd, dptr(hist), dptr(xi))
    return hist, xi


def mc_setup(d, ncall, lows=None, highs=None):
    s = McSetup()
    lw = None if lows is None else np.ascontiguousarray(lows, dtype=np.float64)
    hg = None if highs is None else np.ascontiguousarray(highs, dtype=np.float64)
    LIB.orc_mc_setup_init(ctypes.byref(s), d, ncall, dptr(lw), dptr(hg))
    return s


def mc_sample(fam, d, prm, setup, xi, key, generator, hd):
    ti = ctypes.c_double()
    tsi = ctypes.c_double()
    hist = np.zeros(max(hd, 1) * MC_BINS)
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    LIB.orc_mc_sample(FAM.get(fam, fam), d, dptr(prm), ctypes.byref(setup), dptr(xi), key,
                      MC_RNG.get(generator, generator), hd, ctypes.byref(ti), ctypes.byref(tsi),
                      dptr(hist))
    return ti.value, tsi.value, hist[: hd * MC_BINS]


def mc_integrate(fam, d, prm, epsrel, epsabs, ncall, total_iters=15, adjust_iters=15,
                 skip_iters=5, generator="lcg", one_axis=False, adjust_params=False, seed=0,
                 lows=None, highs=None, trace_cap=0):
    p = McParams(ncall, total_iters, adjust_iters, skip_iters, MC_RNG.get(generator, generator),
                 int(one_axis), int(adjust_params), seed)
    res = McResult()
    tr = None
    arrs = None
    if trace_cap:
        arrs = (np.zeros(trace_cap), np.zeros(trace_cap))
        tr = McTrace(trace_cap, 0, dptr(arrs[0]), dptr(arrs[1]))
    lw = None if lows is None else np.ascontiguousarray(lows, dtype=np.float64)
    hg = None if highs is None else np.ascontiguousarray(highs, dtype=np.float64)
    xi = np.zeros(d * (MC_BINS + 1))
    assert LIB.orc_mc_integrate(FAM.get(fam, fam), d, dptr(prm), epsrel, epsabs, dptr(lw),
                                dptr(hg), ctypes.byref(p), ctypes.byref(res),
                                ctypes.byref(tr) if tr is not None else None, dptr(xi)) == 0
    out = dict(estimate=res.estimate, errorest=res.errorest, chi_sq=res.chi_sq,
               status=res.status, iters=res.iters, runs=res.runs, ncall=res.ncall, xi=xi)
    if trace_cap:
        out["ti"] = arrs[0][: tr.len].copy()
        out["tsi"] = arrs[1][: tr.len].copy()
    return out
