"""C ABI checks that need no GPU: the library loads, exports every function
declared in include/ccfv.h, and validates its arguments before touching a
device."""
import ctypes
import os
import re

import pytest

import ccfv
from ccfv import types as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "ccfv.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(ccfv_\w+)\s*\(", src, re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = header_functions()
    for n in ("ccfv_create", "ccfv_step", "ccfv_coarse", "ccfv_fine",
              "ccfv_plain_solve", "ccfv_parareal_solve",
              "ccfv_parareal_correction", "ccfv_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(ccfv.LIB_PATH)
    for n in header_functions():
        assert hasattr(lib, n), n
    assert sorted(ccfv.exported_symbols()) == header_functions()


def test_struct_layouts():
    assert ctypes.sizeof(T.Problem) == 96 + 8 + 5 * 8
    assert ctypes.sizeof(T.IC) == 8 + 16 * 8
    assert ctypes.sizeof(T.ParaReport) == 16 + 64 * 8 + 24 + 8
    assert ctypes.sizeof(T.KernelStat) == 24


def test_reference_presets_match_the_c_side():
    lib = ccfv.load()
    p = T.Problem()
    assert lib.ccfv_problem_reference(ctypes.byref(p), 128, 0.01) == 0
    q = T.reference_problem(128, 0.01)
    for f, _ in T.Problem._fields_:
        a, b = getattr(p, f), getattr(q, f)
        if f in ("vel", "src"):
            assert list(a) == list(b)
        else:
            assert a == b, f
    ic = T.IC()
    assert lib.ccfv_ic_reference(ctypes.byref(ic), 0.01) == 0
    assert list(ic.p) == list(T.reference_ic().p)


@pytest.mark.parametrize("kw,msg", [
    ({"nx": 33}, "unsupported grid"),
    ({"ny": 6}, "unsupported grid"),
    ({"nu": -1.0}, "diffusion"),
    ({"bc": 7}, "boundary"),
    ({"vel_kind": 9}, "velocity"),
    ({"x_hi": -2.0}, "box"),
    ({"flags": T.FLAG_STEP3_POSTREMAP}, "reference"),
    ({"src_kind": 3}, "source kind"),
    ({"src_kind": T.SRC_CONSTANT, "src": (float("nan"),)}, "source parameters"),
])
def test_create_rejects_bad_problems(kw, msg):
    p = T.rotation_problem(64).copy(**kw)
    with pytest.raises(ccfv.CcfvError) as ei:
        ccfv.Context(p)
    assert ei.value.code == T.CCFV_EINVAL
    assert msg in str(ei.value)


def test_create_without_device_fails_loudly():
    if ccfv.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(ccfv.CcfvError) as ei:
        ccfv.Context(T.rotation_problem(64))
    assert ei.value.code == T.CCFV_EDEVICE


def test_null_arguments_are_errors():
    lib = ccfv.load()
    assert lib.ccfv_step(None, 0.0, 0.1) == T.CCFV_EINVAL
    assert lib.ccfv_plain_solve(None, 0.0, 1.0, 4) == T.CCFV_EINVAL
    assert b"NULL" in lib.ccfv_last_error()


def test_batch_rejects_bad_problems():
    with pytest.raises(ccfv.CcfvError) as ei:
        ccfv.Batch([T.rotation_problem(64), T.rotation_problem(64).copy(nx=36)])
    assert ei.value.code == T.CCFV_EINVAL
    assert "multiples of 8" in str(ei.value)
    with pytest.raises(ccfv.CcfvError) as ei:
        ccfv.Batch([T.rotation_problem(64).copy(nu=-1.0)])
    assert "problem 0" in str(ei.value)
    lib = ccfv.load()
    assert lib.ccfv_batch_create(None, 1, None) == T.CCFV_EINVAL
    assert lib.ccfv_batch_plain(None, None, None, None) == T.CCFV_EINVAL
    assert lib.ccfv_batch_parareal(None, None, None, None, None, 1, 0.0,
                                   None) == T.CCFV_EINVAL
    lib.ccfv_batch_destroy(None)


def test_batch_without_device_fails_loudly():
    if ccfv.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(ccfv.CcfvError) as ei:
        ccfv.Batch([T.rotation_problem(64)])
    assert ei.value.code == T.CCFV_EDEVICE


def test_ctypes_mirrors_match_the_c_header(tmp_path):
    """sizeof / offsetof of every ABI struct, from the C compiler, against
    the ctypes mirrors."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    structs = {"ccfv_problem": T.Problem, "ccfv_ic": T.IC,
               "ccfv_parareal_report": T.ParaReport,
               "ccfv_kernel_stat": T.KernelStat,
               "ccfv_server_config": T.ServerConfig, "ccfv_request": T.Request,
               "ccfv_server_status": T.ServerStatus}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ccfv.h"',
             "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "sizes.c"
    src.write_text("\n".join(lines))
    exe = str(tmp_path / "sizes")
    subprocess.run([gcc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src),
                    "-o", exe], check=True)
    got = dict(l.split() for l in subprocess.run([exe], capture_output=True,
                                                 text=True).stdout.splitlines())
    for cname, cls in structs.items():
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)


def test_server_arguments_checked_before_the_device():
    lib = ccfv.load()
    assert lib.ccfv_server_create(None, None) == T.CCFV_EINVAL
    cfg = T.ServerConfig(1 << 20, 5000, 0, 0, 0)
    h = ctypes.c_void_p()
    assert lib.ccfv_server_create(ctypes.byref(cfg), ctypes.byref(h)) == T.CCFV_EINVAL
    assert b"max_active" in lib.ccfv_last_error()
    assert lib.ccfv_server_step(None, None) == T.CCFV_EINVAL
    lib.ccfv_server_destroy(None)


def test_request_bytes_is_the_resident_footprint():
    from ccfv import workloads
    s = workloads.rotation_solve(256)
    r = s.request(kmax=4, tol=0.0)
    b = ccfv.request_bytes(r)
    f = 8 * 256 * 256
    # U (ns+1) + G ns + F ns + work field, plus the increment partials
    assert b == (3 * 16 + 2) * f + 2 * 8 * 128 * 17
    bad = s.request(kmax=4, tol=0.0)
    bad.problem.nx = 36
    with pytest.raises(ccfv.CcfvError) as ei:
        ccfv.request_bytes(bad)
    assert ei.value.code == T.CCFV_EINVAL
