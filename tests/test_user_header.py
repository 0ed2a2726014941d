"""The DSL user header (include/physis/physis_user.h): DSL sources type-check
as plain C and C++ with PHYSIS_USER defined, and the element-typed accessors
catch type errors.  CPU only (needs gcc/g++)."""
import glob
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
CASES = "/root/reference/tests/system_tests/test_cases"


def programs():
    own = sorted(glob.glob(os.path.join(HERE, "golden", "programs", "*.c")))
    ref = sorted(p for p in glob.glob(os.path.join(CASES, "test_*.c"))
                 if ".manual." not in p and not p.endswith(".module_base.c"))
    return own + ref


def syntax(compiler, path):
    if shutil.which(compiler[0]) is None:
        pytest.skip("%s is not installed" % compiler[0])
    return subprocess.run(compiler + ["-fsyntax-only", "-DPHYSIS_USER", "-I", INCLUDE, "-Werror=implicit-function-declaration",
                                      path], capture_output=True, text=True)


@pytest.mark.parametrize("compiler", [["gcc", "-std=gnu99"], ["gcc", "-std=c11", "-D_POSIX_C_SOURCE=200809L"],
                                      ["g++", "-std=c++17", "-x", "c++"]], ids=["c99", "c11", "c++17"])
def test_dsl_programs_type_check(compiler):
    """This repository's DSL programs and, where the reference checkout is
    present, its whole system-test corpus (grids of every element type, user
    point types, red-black maps, reductions, module units)."""
    failed = []
    for path in programs():
        r = syntax(compiler, path)
        if r.returncode != 0:
            failed.append((os.path.basename(path), r.stderr[-400:]))
    assert not failed, failed


BAD = """
#include "physis/physis.h"
struct Point { float a; };
DeclareGrid3D(Point, struct Point);
static void kernel(const int x, const int y, const int z, PSGrid3DFloat g, PSGrid3DPoint h) {
  %s
}
int main(int argc, char *argv[]) { PSInit(&argc, &argv, 3, 4, 4, 4); return 0; }
"""


@pytest.mark.parametrize("body,ok", [
    ("float v = PSGridGet(g, x, y, z); PSGridEmit(g, v);", True),
    ("float v = PSGridGet(h, x, y, z).a; PSGridEmitUtype(h.a, v);", True),
    ("float v = PSGridGet(g, x, y, z).a;", False),          # member access on a float grid
    ("struct Point p = PSGridGet(g, x, y, z);", False),     # float where a point is expected
    ("PSGridEmit(g, \"text\");", False),                    # emit of a pointer into a float grid
    ("float v = PSGridGet(g, x, y, z, undefined_index);", False),
])
def test_accessors_carry_the_element_type(tmp_path, body, ok):
    src = tmp_path / "k.c"
    src.write_text(BAD % body)
    r = syntax(["gcc", "-std=gnu99"], str(src))
    assert (r.returncode == 0) == ok, r.stderr[-400:]
