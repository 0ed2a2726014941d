"""The C-ABI boundary, checked without a GPU: the library loads, exports every
function the public headers declare, and the struct layouts generated code
depends on are the reference's (include/physis/physis_cuda.h:18-106,
include/physis/physis_common.h:86-99)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include", "physis")
LIB = os.path.join(ROOT, "physis_amd", "lib", "libphysis_rt_hip.so")
HEADERS = ["physis_common.h", "physis_hip.h", "physis_hip_ext.h"]

# Declared in physis_common.h for the DSL but implemented by generated code or
# by targets other than this one.
NOT_EXPORTED = set()


def declared_functions():
    names = []
    for h in HEADERS:
        text = open(os.path.join(INC, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\bextern\s+[^;()]*?\b([A-Za-z_]\w*)\s*\(", text):
            names.append(m.group(1))
    return sorted(set(names) - NOT_EXPORTED)


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], check=True, capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-C", os.path.join(ROOT, "physis_amd", "csrc"), "-j8", "-s"], check=True)
    return LIB


def test_library_loads(built):
    ctypes.CDLL(built)


def test_every_declared_function_is_exported(built):
    decl = declared_functions()
    assert "__PSGridNew" in decl and "__PSReduceGridFloat" in decl and "PSInit" in decl
    missing = [n for n in decl if n not in exported_symbols()]
    assert not missing, "declared but not exported: %s" % missing


def test_struct_layouts_match_reference(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stddef.h>
#define PHYSIS_HIP
#include "physis/physis.h"
_Static_assert(sizeof(PSIndex) == 4, "PSIndex is int32");
_Static_assert(sizeof(__PSDomain) == 48, "domain");
_Static_assert(offsetof(__PSDomain, local_min) == 24, "domain local_min");
_Static_assert(offsetof(__PSGrid, p0) == 0, "p0");
_Static_assert(offsetof(__PSGrid, dim) == 8, "dim");
_Static_assert(offsetof(__PSGrid, elm_size) == 20, "elm_size");
_Static_assert(offsetof(__PSGrid, num_dims) == 24, "num_dims");
_Static_assert(offsetof(__PSGrid, num_elms) == 32, "num_elms");
_Static_assert(offsetof(__PSGrid, dev) == 40, "dev");
_Static_assert(sizeof(__PSGrid3DFloat_dev) == 24, "dev view");
_Static_assert(offsetof(__PSGrid3DFloat_dev, p0) == 16, "dev view p0");
_Static_assert(offsetof(__PSGrid2DFloat_dev, p0) == 8, "2d view p0");
_Static_assert(PS_MAX == 0 && PS_MIN == 1 && PS_SUM == 2 && PS_PROD == 3, "reduce ops");
int main(void) {
  PSGrid3DFloat g = 0;
  (void)g;
  return __PSGridGetOffset3DDev((const int[3]){4, 5, 6}, 1, 2, 3) == 1 + 2 * 4 + 3 * 20 ? 0 : 1;
}
''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def test_offset_helpers_follow_reference_periodic_rule(tmp_path):
    # (i + n) % n per axis (include/physis/physis_ref.h:69-81).
    src = tmp_path / "periodic.c"
    src.write_text(r'''
#define PHYSIS_HIP
#include "physis/physis.h"
int main(void) {
  int d[3] = {5, 4, 3};
  if (__PSGridGetOffsetPeriodic3DDev(d, -1, 0, 0) != 4) return 1;
  if (__PSGridGetOffsetPeriodic3DDev(d, 0, -1, 0) != 15) return 2;
  if (__PSGridGetOffsetPeriodic3DDev(d, 0, 0, 3) != 0) return 3;
  if (__PSGridGetOffsetPeriodic3DDev(d, 5, 4, -3) != 0) return 4;
  return 0;
}
''')
    exe = tmp_path / "periodic"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def test_python_binding_fails_loudly_without_library(monkeypatch):
    from physis_amd import runtime
    monkeypatch.setattr(runtime, "_lib", None)
    monkeypatch.setattr(runtime, "LIB_PATH", "/nonexistent/libphysis_rt_hip.so")
    with pytest.raises(RuntimeError, match="not built"):
        runtime.lib()
