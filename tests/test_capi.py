"""C-ABI boundary checks that run without a GPU: the library builds, loads and exports every
function declared in include/*.h; the ctypes structs match the C structs' layout."""
import ctypes
import os
import re

import pytest

import rls_amd
from rls_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            src = open(os.path.join(inc, f)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(rls_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("librls_hip.so not built")
    lib = ctypes.CDLL(_capi.LIB_PATH)
    decl = declared_functions()
    assert decl, "no declarations found"
    for name in decl:
        assert hasattr(lib, name), f"{name} declared in include/ but not exported"
    assert set(_capi.EXPORTED) <= set(decl)


def test_header_constants_match_python_mirror():
    """Every RLS_* integer constant that rls_amd._capi mirrors has the header's value."""
    src = open(os.path.join(ROOT, "include", "rls_hip.h")).read()
    consts = {m.group(1): int(m.group(2))
              for m in re.finditer(r"#define\s+(RLS_[A-Z0-9_]+)\s+(-?\d+)\b", src)}
    mirrored = [k for k in dir(_capi) if k.startswith("RLS_") and isinstance(getattr(_capi, k), int)]
    assert len(mirrored) >= 14
    for name in mirrored:
        assert name in consts, f"{name} is not defined in include/rls_hip.h"
        assert getattr(_capi, name) == consts[name], name


def test_control_defaults_match_reference():
    """rls_tri_control_default mirrors PsTriRecyControl's defaults (psTriRecy.jl:350-370)."""
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("librls_hip.so not built")
    lib = _capi.load()
    c = _capi.TriControl()
    lib.rls_tri_control_default(ctypes.byref(c), 3.0)
    py = rls_amd.PsTriRecyControl(normC=3.0)
    assert c.tol == py.tol == 1e-3
    assert c.shift == pytest.approx(py.shift) and py.shift == pytest.approx(9.0 * 2.220446049250313e-16 * 100)
    assert (c.gap_tau, c.gap_tau_scale, c.recy_tau) == (py.gap_tau, py.gap_tau_scale, py.recy_tau)
    assert (c.r, c.p, c.p_min) == (py.r, py.p, py.p_min) == (6, 50, 1)
    assert c.lobpcg_tol_scale == py.lobpcg_tol_scale and c.lobpcg_side == b"L"
    assert c.maxit == 2000 and c.implict_recycle == 1 and c.implict_update == 0


def _header_fields(struct):
    """Field names of `typedef struct <struct> {...}` in include/rls_hip.h, in order."""
    src = open(os.path.join(ROOT, "include", "rls_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    body = re.search(r"typedef struct %s \{(.*?)\}\s*%s;" % (struct, struct), src, flags=re.S).group(1)
    names = []
    for decl in body.split(";"):
        decl = re.sub(r"\[\d+\]", "", decl).strip()
        if not decl:
            continue
        parts = [p.strip() for p in decl.split(",")]
        names.append(re.findall(r"\*?\s*(\w+)$", parts[0])[0])
        names += [p.lstrip("*").strip() for p in parts[1:]]
    return names


def _c_layout(struct, fields):
    """(sizeof, {field: offsetof}) of a header struct, computed by the C compiler."""
    import shutil
    import subprocess
    import tempfile
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = [f'printf("%zu\\n", sizeof({struct}));']
    lines += [f'printf("%zu\\n", offsetof({struct}, {f}));' for f in fields]
    prog = ("#include <stddef.h>\n#include <stdio.h>\n#include \"rls_hip.h\"\nint main(void){"
            + "".join(lines) + "return 0;}\n")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        open(c, "w").write(prog)
        exe = os.path.join(d, "l")
        subprocess.run([cc, "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    return out[0], dict(zip(fields, out[1:]))


@pytest.mark.parametrize("struct,mirror", [("rls_tri_control", "TriControl"), ("rls_tri_stats", "TriStats"),
                                           ("rls_tri_state", "TriState"), ("rls_prec_control", "PrecControl"),
                                           ("rls_prec_stats", "PrecStats")])
def test_ctypes_mirror_matches_header_layout(struct, mirror):
    fields = _header_fields(struct)
    size, offs = _c_layout(struct, fields)
    cls = getattr(_capi, mirror)
    assert [f[0] for f in cls._fields_] == fields
    assert ctypes.sizeof(cls) == size
    for f in fields:
        assert getattr(cls, f).offset == offs[f], f


_JULIA_SIZES = {"Float64": 8, "Int64": 8, "UInt64": 8, "Int32": 4, "UInt8": 1, "Cchar": 1}


def _julia_struct(name):
    """(field names, offsets, size) of a Julia `struct` block in INTEGRATION.md, laid out with
    the C rules Julia applies to isbits structs (each field at its natural alignment)."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    body = re.search(r"(?:mutable\s+)?struct %s\n(.*?)\nend" % name, text, flags=re.S).group(1)
    body = re.sub(r"#.*", "", body)
    fields = re.findall(r"(\w+)::([\w{}]+)", body)
    offs, off, align = [], 0, 1
    for _, ty in fields:
        sz = 8 if ty.startswith("Ptr") else _JULIA_SIZES[ty]
        off = (off + sz - 1) // sz * sz
        offs.append(off)
        off += sz
        align = max(align, sz)
    size = (off + align - 1) // align * align
    return [f for f, _ in fields], offs, size


@pytest.mark.parametrize("struct,jl", [("rls_tri_control", "CTriControl"), ("rls_tri_stats", "CTriStats"),
                                       ("rls_prec_control", "CPrecControl"), ("rls_prec_stats", "CPrecStats")])
def test_julia_shim_structs_match_header(struct, jl):
    """The Julia binding in INTEGRATION.md passes every header field at the C offset, so no
    field (e.g. resume) is read from undefined padding."""
    fields = _header_fields(struct)
    size, offs = _c_layout(struct, fields)
    names, joffs, jsize = _julia_struct(jl)
    assert names == fields
    assert jsize == size
    assert joffs == [offs[f] for f in fields]


def test_grid_mirror_matches_oracle():
    from oracle import rls_oracle as O
    import numpy as np
    for args in [(-4.0, 6.5, -5.0, 2.0, 200, 200, 20), (0.45, 0.55, 0.0, 0.1, 20, 20, 20), (0, 1, 0, 1, 5, 6, 3)]:
        a, b = rls_amd.recursion_grid(*args), O.recursion_grid(*args)
        assert np.array_equal(a.Z, b.Z)
        assert np.array_equal(rls_amd.recursion_natural_permutation(a), O.recursion_natural_permutation(b))
    with pytest.raises(ValueError):
        rls_amd.recursion_grid(0, 1, 0, 1, 10, 10, 3)


def test_context_without_gpu_fails_loudly():
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("librls_hip.so not built")
    import subprocess, sys
    code = "import rls_amd._capi as c; c.Context(0)"
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       env={**os.environ, "HIP_VISIBLE_DEVICES": "-1"})
    assert r.returncode != 0 and "rls_ctx_create" in r.stderr


def test_prec_control_defaults_match_reference():
    """rls_prec_control_default mirrors PsPrecRecyControl's defaults (psPrecRecy.jl:220-261)
    and the Python mirror's."""
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("librls_hip.so not built")
    from rls_amd import prec as RP
    lib = _capi.load()
    c = _capi.PrecControl()
    lib.rls_prec_control_default(ctypes.byref(c), 7.5)
    py = RP.PsPrecRecyControl(main_normC=7.5)
    for f in ("tol", "lobpcg_tol_scale", "main_normC", "main_tau", "gap_tau", "tol_pred", "lobpcg_pred_tol_scale",
              "pred_tau", "idx_cutoff_pred", "lambda_lower_min", "lambda_upper_max", "hsl_tol", "hsl_cntl_1",
              "hsl_cntl_2", "gamma_tol", "lambda_n_tol", "r_main", "p_main", "p_min_main", "r_pred", "p_pred",
              "p_min_pred", "hsl_icntl_1", "hsl_icntl_2", "pred_conv_thresh_type", "maxit"):
        assert getattr(c, f) == getattr(py, f), f
    assert c.shift == pytest.approx(py.shift)
    assert (c.tol, c.main_tau, c.r_pred, c.p_pred, c.hsl_cntl_2, c.lambda_upper_max) == (1e-3, 1.666, 20, 20, 1e-5, 30.0)
    assert c.main_implict_recycle == 0 and c.pred_implict_recycle == 1 and c.pred_implict_update == 1
    assert c.pred_orth_type == _capi.RLS_ORTH_LAPACK and c.pred_conv_first == 0
    import math
    assert math.isnan(c.recycle_fail_dist_thresh)
