"""CPU checks of the HIP library (no kernel launches).

The library must load, export every entry point include/pagani.h declares,
build rule tables bit-identical to the oracle's, and fail cleanly (NULL plus
a reason, never a crash) where no device is visible.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_api as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pagani.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text)
    skip = {"sizeof", "if", "for", "while", "return"}
    return sorted({n for n in names if n not in skip and (n.startswith("pagani_") or
                                                          n == "do_integration_from_c")
                   and not n.endswith("_fn")})


def test_library_exports_every_declared_symbol(pg):
    lib = pg.load()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert set(pg.exported_symbols()) >= set(names) - {"pagani_genz_eval_launch"}


def test_test_helper_library_loads():
    lib = ctypes.CDLL(os.path.join(ROOT, "tests", "cxx", "libreftests.so"))
    for n in ("reftest_integrate", "reftest_integrate_split", "reftest_sample"):
        assert hasattr(lib, n)


@pytest.mark.parametrize("d", range(2, 17))
def test_rule_tables_bit_identical_to_oracle(pg, d):
    w, sc, nm, lam, sb, g = pg.rule_tables(d)
    ow, olam, osc, onm, cnt = O.rule_tables(d)
    og, oset = O.generators(d)
    assert np.array_equal(w, ow) and np.array_equal(lam, olam)
    assert np.array_equal(sc[:, 1:4], osc[:, 1:4]) and np.array_equal(nm[:, 1:4], onm[:, 1:4])
    assert np.array_equal(g, og[1 + 4 * d:])
    assert sb == [int(v) for v in np.concatenate([[0], np.cumsum(cnt)])]


@pytest.mark.parametrize("d", [2, 3, 5, 8, 10, 16])
def test_rule_point_order_is_the_oracles(pg, d):
    """pagani_rule_point (the debug recorder's point order) lists the
    oracle's generators: the centre, +-lambda1 and +-lambda2 per axis in the
    fourth difference's order (Sample.cuh:194-218), then the table."""
    og, _ = O.generators(d)
    got = np.array([[pg.rule_point(d, p, j) for j in range(d)] for p in range(len(og))])
    assert np.array_equal(got, og)


def test_rule_point_rejects_bad_arguments(pg):
    lib = pg.load()
    g = ctypes.c_double()
    F = lib.pagani_rule_tables(4, None, None, None, None, None, None)
    for args in ((1, 0, 0), (17, 0, 0), (4, -1, 0), (4, F, 0), (4, 0, -1), (4, 0, 4)):
        assert lib.pagani_rule_point(*args, ctypes.byref(g)) == pg.EBADARG, args
    assert lib.pagani_rule_point(4, 0, 0, None) == pg.EBADARG
    assert lib.pagani_rule_point(4, F - 1, 3, ctypes.byref(g)) == 0


def test_rule_tables_reject_bad_dims(pg):
    lib = pg.load()
    for d in (0, 1, 17, 18):
        assert lib.pagani_rule_tables(d, None, None, None, None, None, None) < 0


def test_create_without_device_fails_cleanly(pg):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a device is visible")
    except ImportError:
        pass
    lib = pg.load()
    assert not lib.pagani_ws_create(8, 0)
    assert lib.pagani_create_error()
    with pytest.raises(pg.PaganiError):
        pg.Workspace(8)
    assert not lib.pagani_ws_create(1, 0)  # bad ndim
    assert b"ndim" in lib.pagani_create_error()
    assert not lib.pagani_batch_create(0, 1)
    assert lib.pagani_batch_last_error(None)
    assert not lib.pagani_batch_create(0, 0)  # bad worker count
    assert b"workers" in lib.pagani_batch_last_error(None)
    with pytest.raises(pg.PaganiError):
        pg.Batch(0, 2)
    # null handles are rejected, not dereferenced
    assert lib.pagani_integrate_batch(None, None, 0, None, None) == pg.EBADARG
    assert lib.pagani_batch_get_stats(None, None) == pg.EBADARG


def test_null_workspace_is_rejected(pg):
    lib = pg.load()
    assert lib.pagani_ws_set_device_mem(None, 1) == pg.EBADARG
    assert lib.pagani_ws_get_stats(None, None) == pg.EBADARG
    assert lib.pagani_integrate_genz(None, None, 64, 1e-3, 1e-12, None, None, 1, 0,
                                     None) == pg.EBADARG
    assert lib.do_integration_from_c(None) == 1  # NULL result pointer


def test_workload_generators():
    from gpuintegration_amd import workloads as W
    fam, a, u = W.gaussian_8d(2024)
    assert fam == "gaussian" and len(a) == 8
    assert all(9.5 <= v <= 10.5 for v in a) and all(0.2 <= v <= 0.8 for v in u)
    assert abs(W.gaussian_exact(a, u) - O.exact("gaussian", 8, O.params(a, u))) <= \
        1e-14 * W.gaussian_exact(a, u)
    b = W.batch(512, 7)
    assert len(b) == 512
    assert {x[0] for x in b} == set(W.FAMILIES)
    assert {len(x[1]) for x in b} == set(range(3, 9))
    assert all(1e-6 <= x[3] <= 1e-3 for x in b)
    assert W.eval_flops_per_region("gaussian", 8) == 90770


def test_workload_closed_forms_match_oracle():
    """The package's analytic values (used by the benchmark to report error)
    agree with the oracle's closed forms, which are pinned to the reference's
    constants and to brute-force quadrature."""
    from gpuintegration_amd import workloads as W
    for fam, a, u, _ in W.batch(96, seed=7):
        want = O.exact(fam, len(a), O.params(a, u))
        got = W.exact(fam, a, u)
        assert got == pytest.approx(want, rel=1e-12, abs=1e-300), (fam, len(a))


def test_binding_structs_match_the_header(pg, tmp_path):
    """Every struct the Python binding passes across the C ABI has the
    header's size and field offsets (a C program built with gcc from
    include/pagani.h prints them)."""
    import subprocess
    structs = {
        "pagani_result": pg._Result, "pagani_genz": pg._Genz,
        "pagani_batch_item": pg._BatchItem, "pagani_eval_consts": pg.EvalConsts,
        "pagani_stats": pg.Stats, "pagani_class_res": pg.ClassRes,
        "pagani_mc_params": pg.McParams, "pagani_mc_result": pg.McResult,
        "pagani_mc_stats": pg.McStats,
    }
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "pagani.h"',
             "int main(void) {"]
    for cname, py in structs.items():
        lines.append('  printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in py._fields_:
            lines.append('  printf("%s.%s %%zu\\n", offsetof(%s, %s));'
                         % (cname, fname, cname, fname))
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = dict(line.split() for line in
               subprocess.run([str(exe)], capture_output=True, text=True,
                              check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, (cname, fname)
