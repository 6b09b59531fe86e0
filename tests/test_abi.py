"""The C-ABI boundary without a GPU: libcollop.so loads, exports every entry
point include/collop.h declares, the ctypes mirrors of the structs match the
C layout, and the error paths that need no device behave as documented.
"""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

import collop
from collop import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "collop.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(collop_\w+)\s*\(", text, re.M)))


def test_header_declares_expected_entry_points():
    names = declared()
    for n in ("collop_create", "collop_collision_f64", "collop_collision_f32c", "collop_l2_project",
              "collop_grad_gather", "collop_loggrad_gather", "collop_run_steps"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (collop_\w+)", out))
    assert set(declared()) <= exported
    # the ctypes signature table covers the whole header
    assert {n for n, _, _ in _lib.SIGNATURES} == set(declared())


def test_struct_layout_matches_c():
    src = r"""
#include <stddef.h>
#include <stdio.h>
#include "collop.h"
int main(void) {
  printf("%zu %zu\n", sizeof(collop_step_params), sizeof(collop_step_stats));
  printf("%zu %zu %zu %zu\n", offsetof(collop_step_params, use_anderson), offsetof(collop_step_params, precision),
         offsetof(collop_step_params, f32_floor_factor), offsetof(collop_step_params, f32_keep_window));
  printf("%zu %zu %zu\n", offsetof(collop_step_stats, residual), offsetof(collop_step_stats, momentum),
         offsetof(collop_step_stats, seconds));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        vals = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    P, S = _lib.StepParams, _lib.StepStats
    assert vals[0] == C.sizeof(P) and vals[1] == C.sizeof(S)
    assert vals[2:6] == [P.use_anderson.offset, P.precision.offset, P.f32_floor_factor.offset, P.f32_keep_window.offset]
    assert vals[6:9] == [S.residual.offset, S.momentum.offset, S.seconds.offset]


def test_errors_without_device():
    lib = _lib.load()
    assert lib.collop_version() >= 100
    h = C.c_void_p()
    st = lib.collop_create(0, C.byref(h))
    if st == 0:
        lib.collop_destroy(h)
        pytest.skip("a HIP device is visible here")
    assert st == _lib.ERR_HIP
    assert not h.value
    # null context: invalid argument, no crash
    assert lib.collop_set_mesh(None, None, 0, None, 0, None, None, None, None, 0, 0, 2, 6) == _lib.ERR_INVALID
    assert lib.collop_destroy(None) == 0
    assert lib.collop_run_steps(None, None, 1, None) == _lib.ERR_INVALID
    assert lib.collop_last_error(None) == b"null context"
    with pytest.raises(collop.CollopError):
        collop.Context(0)


def test_product_path_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_LIB", None)
    with pytest.raises(collop.CollopError):
        _lib.load(str(tmp_path / "missing.so"))
