"""physisc-hip (CPU only): the clamped-face read recognition behind a star
plan's compute_exact.

A clamped-face read `if (x == 0) w = PSGridGet(g, x, y, z); else w =
PSGridGet(g, x-1, y, z);` equals the staged grid at the clamped coordinate,
so where the fused skeleton supplies clamped-coordinate neighbours the
translator may fold the condition away (`!__kExactX && ...`).  Anything that
is not exactly that form must keep its condition: the tests below translate
small programs and check which conditionals were rewritten.  The numerics of
the rewritten plans are checked on the GPU by the corpus tests (the
reference's Neumann-condition test and diffusion benchmark)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PHYSISC = os.path.join(ROOT, "physis_amd", "bin", "physisc-hip")

PROGRAM = """
#include "physis/physis.h"
static void kernel(const int x, const int y, const int z, PSGrid3DFloat g1, PSGrid3DFloat g2, float k) {
  int nx, ny, nz;
  nx = PSGridDim(g1, 0);
  ny = PSGridDim(g1, 1);
  nz = PSGridDim(g1, 2);
  %(extra)s
  float c, w, e, n, s, b, t;
  c = PSGridGet(g1, x, y, z);
  %(body)s
  PSGridEmit(g2, k * (c + w + e + n + s + b + t));
  return;
}
int main(int argc, char *argv[]) {
  PSInit(&argc, &argv, 3, 16, 16, 16);
  PSGrid3DFloat g1 = PSGrid3DFloatNew(16, 16, 16);
  PSGrid3DFloat g2 = PSGrid3DFloatNew(16, 16, 16);
  PSDomain3D d = PSDomain3DNew(0, 16, 0, 16, 0, 16);
  PSStencilRun(PSStencilMap(kernel, d, g1, g2, 0.1f), PSStencilMap(kernel, d, g2, g1, 0.1f), 2);
  PSGridFree(g1);
  PSGridFree(g2);
  PSFinalize();
  return 0;
}
"""

CLAMPED = """
  if (x == 0) w = PSGridGet(g1, x, y, z); else w = PSGridGet(g1, x-1, y, z);
  if (x == nx-1) e = PSGridGet(g1, x, y, z); else e = PSGridGet(g1, x+1, y, z);
  if (y == 0) n = PSGridGet(g1, x, y, z); else n = PSGridGet(g1, x, y-1, z);
  if (y == PSGridDim(g1, 1) - 1) s = PSGridGet(g1, x, y, z); else s = PSGridGet(g1, x, y+1, z);
  if (z == 0) b = PSGridGet(g1, x, y, z); else b = PSGridGet(g1, x, y, z-1);
  if (z == nz-1) t = PSGridGet(g1, x, y, z); else t = PSGridGet(g1, x, y, z+1);
"""


def translate(tmp_path, body, extra=""):
    if not os.path.exists(PHYSISC):
        pytest.skip("physisc-hip is not built (make -C physis_amd/translator)")
    src = tmp_path / "k.c"
    src.write_text(PROGRAM % {"body": body, "extra": extra})
    out = tmp_path / "k.hip"
    subprocess.run([PHYSISC, "-o", str(out), str(src)], check=True, capture_output=True)
    return out.read_text()


def exact_body(hip):
    """The text of compute_exact, or None when the plan has none."""
    at = hip.find("compute_exact(")
    if at < 0:
        return None
    return hip[at:hip.find("\n  }\n", at)]


def test_all_six_clamped_faces_fold(tmp_path):
    body = exact_body(translate(tmp_path, CLAMPED))
    assert body is not None
    for axis, n in (("X", 2), ("Y", 2), ("Z", 2)):
        assert body.count("!__kExact" + axis) == n, axis
    # compute() itself keeps every condition unchanged.
    hip = translate(tmp_path, CLAMPED)
    plain = hip[hip.find("compute(const C"):hip.find("kExactNeighbours")]
    assert "__kExact" not in plain and plain.count("if") == 6


@pytest.mark.parametrize("body", [
    # offset on the wrong side of the face
    "if (x == 0) w = PSGridGet(g1, x, y, z); else w = PSGridGet(g1, x+1, y, z);",
    # the face value is not the point itself
    "if (x == 0) w = PSGridGet(g1, x+1, y, z); else w = PSGridGet(g1, x-1, y, z);",
    # another axis in the condition than in the offset
    "if (y == 0) w = PSGridGet(g1, x, y, z); else w = PSGridGet(g1, x-1, y, z);",
    # upper face tested against another axis' extent
    "if (x == ny-1) w = PSGridGet(g1, x, y, z); else w = PSGridGet(g1, x+1, y, z);",
    # different targets in the two branches
    "if (x == 0) w = PSGridGet(g1, x, y, z); else e = PSGridGet(g1, x-1, y, z);",
    # a face other than 0 or n-1
    "if (x == 1) w = PSGridGet(g1, x, y, z); else w = PSGridGet(g1, x-1, y, z);",
])
def test_other_conditionals_are_kept(tmp_path, body):
    hip = translate(tmp_path, body + "\n  e = n = s = b = t = c;")
    assert exact_body(hip) is None


def test_extent_variable_assigned_twice_is_not_trusted(tmp_path):
    body = "if (x == nx-1) e = PSGridGet(g1, x, y, z); else e = PSGridGet(g1, x+1, y, z);\n  w = n = s = b = t = c;"
    assert exact_body(translate(tmp_path, body)) is not None
    for extra in ("nx = nx - 1;", "nx += 1;", "nx++;", "--nx;", "int *p = &nx; (void)p;"):
        assert exact_body(translate(tmp_path, body, extra=extra)) is None, extra


def test_periodic_kernels_have_no_exact_form(tmp_path):
    body = ("if (x == 0) w = PSGridGet(g1, x, y, z); else w = PSGridGet(g1, x-1, y, z);\n"
            "  e = PSGridGetPeriodic(g1, x+1, y, z);\n  n = s = b = t = c;")
    assert exact_body(translate(tmp_path, body)) is None


WAVE = """
#include "physis/physis.h"
static void wave(const int x, const int y, const int z, PSGrid3DFloat a, PSGrid3DFloat b, float k) {
  float c, w;
  c = PSGridGet(a, x, y, z);
  w = PSGridGetPeriodic(a, x - 1, y, z);
  PSGridEmit(b, 2.0f * c - %(read)s + k * w);
  return;
}
int main(int argc, char *argv[]) {
  PSInit(&argc, &argv, 3, 16, 16, 16);
  PSGrid3DFloat u = PSGrid3DFloatNew(16, 16, 16);
  PSGrid3DFloat v = PSGrid3DFloatNew(16, 16, 16);
  PSDomain3D d = PSDomain3DNew(0, 16, 0, 16, 0, 16);
  PSStencilRun(PSStencilMap(wave, d, u, v, 0.1f), PSStencilMap(wave, d, v, u, 0.1f), 4);
  PSFinalize();
  return 0;
}
"""


def translate_text(tmp_path, text):
    if not os.path.exists(PHYSISC):
        pytest.skip("physisc-hip is not built (make -C physis_amd/translator)")
    src = tmp_path / "w.c"
    src.write_text(text)
    out = tmp_path / "w.hip"
    subprocess.run([PHYSISC, "-o", str(out), str(src)], check=True, capture_output=True)
    return out.read_text()


@pytest.mark.parametrize("read", ["PSGridGet(b, x, y, z)", "PSGridGetPeriodic(b, x + 1, y, z)"])
def test_kernel_reading_its_output_grid_is_not_fused(tmp_path, read):
    """A ping-pong pair of a kernel that reads the grid it emits (a leapfrog
    wave: b = 2a - b + ...) must run as separate steps: in a fused pass the
    second map would read its output grid while it holds another time level
    than two separate steps see."""
    hip = translate_text(tmp_path, WAVE % {"read": read})
    assert "__PSRunPingPong_wave" not in hip and "__PSFusable_wave" not in hip
    assert "fusable pairs" not in hip


def test_star_kernel_not_reading_its_output_is_fused(tmp_path):
    hip = translate_text(tmp_path, WAVE % {"read": "c"})
    assert "__PSFusable_wave" in hip and "fusable pairs" in hip
    # Rows of any supported width qualify (the stride tile serves ragged
    # rows), not just rows of whole 16-byte vectors.
    fus = hip[hip.find("static int __PSFusable_wave"):]
    fus = fus[:fus.find("\n}\n")]
    assert "psk::FusedPairSupported" in fus and "% 4" not in fus


def test_fusable_check_rejects_grids_aliasing_the_swapped_pair(tmp_path):
    """Another grid the body reads may not be one of the two swapped grids:
    __PSFusable compares it with both at run time."""
    text = (WAVE % {"read": "PSGridGet(g, x, y, z)"}).replace(
        "PSGrid3DFloat b, float k", "PSGrid3DFloat b, PSGrid3DFloat g, float k").replace(
        "u, v, 0.1f), PSStencilMap(wave, d, v, u, 0.1f)", "u, v, u, 0.1f), PSStencilMap(wave, d, v, u, u, 0.1f)")
    hip = translate_text(tmp_path, text)
    fus = hip[hip.find("static int __PSFusable_wave"):]
    fus = fus[:fus.find("\n}\n")]
    assert "a->g == a->a || a->g == a->b" in fus


COEF_KERNEL = """
#include "physis/physis.h"
static void relax(const int x, const int y, const int z, PSGrid3DFloat p, PSGrid3DFloat q, PSGrid3DFloat a,
                  PSGrid3DFloat b, PSGrid3DFloat c, PSGrid3DFloat d, PSGrid2DFloat m) {
  float s = PSGridGet(a, x, y, z) * PSGridGet(p, x + 1, y, z) + PSGridGet(b, x, y, z) * PSGridGet(p, x, y - 1, z)
            + PSGridGet(c, x, y, z + 1) * PSGridGet(p, x, y, z) + PSGridGet(d, x, y, z) * PSGridGet(m, x, y)
            + PSGridGetPeriodic(a, x - 1, y, z);
  PSGridEmit(q, s);
  return;
}
int main(int argc, char *argv[]) {
  PSInit(&argc, &argv, 3, 16, 16, 16);
  PSGrid3DFloat p = PSGrid3DFloatNew(16, 16, 16);
  PSGrid3DFloat q = PSGrid3DFloatNew(16, 16, 16);
  PSGrid3DFloat a = PSGrid3DFloatNew(16, 16, 16);
  PSGrid3DFloat b = PSGrid3DFloatNew(16, 16, 16);
  PSGrid3DFloat c = PSGrid3DFloatNew(16, 16, 17);
  PSGrid3DFloat d = PSGrid3DFloatNew(16, 16, 16);
  PSGrid2DFloat m = PSGrid2DFloatNew(16, 16);
  PSDomain3D dom = PSDomain3DNew(1, 15, 1, 15, 1, 15);
  PSStencilRun(PSStencilMap(relax, dom, p, q, a, b, c, d, m), 1);
  PSFinalize();
  return 0;
}
"""


def test_plan_reading_many_grids_has_a_uniform_extents_variant(tmp_path):
    """A 3-D kernel reading grids at the point itself gets a second plan
    instantiation (__kU) in which every rank-3 access indexes by the staged
    grid's extents in 32 bits (psk::Get3/Put3), launched only when
    psk::UniformExtents holds for every rank-3 grid of the map (here c has
    other extents, so that map keeps the per-grid path at run time).  The
    2-D grid and the periodic read keep their own offsets; the staged grid
    is read from the LDS ring."""
    hip = translate_text(tmp_path, COEF_KERNEL)
    start = hip.find("struct __PSPlan_relax")
    plan = hip[start:hip.find("\n};\n", start)]
    assert "template <class SF, bool __kU = false>\nstruct __PSPlan_relax" in hip
    for g, at in (("a", "x, y, z"), ("b", "x, y, z"), ("c", "x, y, z + 1"), ("d", "x, y, z")):
        assert "psk::Get3<__kU>(%s, __ref, %s)" % (g, at) in plan, g
    assert "psk::Put3<__kU>(q, __ref, x, y, z," in plan
    assert "m.load(psk::Off2(m, x, y))" in plan
    assert "a.load(psk::OffP3(a, x - 1, y, z))" in plan
    assert "const auto &__ref = __a.p;" in plan
    assert "__ctx.at(0, 1, 0, 0)" in plan
    launch = hip[hip.find("static void __PSLaunchT_relax"):]
    assert "psk::UniformExtents(a.p, a.q, a.a, a.b, a.c, a.d))" in launch
    assert "LaunchRing<__PSPlan_relax<SF, true>, 3>(__PSArgs_relax<SF, true>(s)" in launch
    assert "LaunchRing<__PSPlan_relax<SF>, 3>(a, box, st, parity)" in launch


POINT_KERNEL = """
#include "physis/physis.h"
static void relax(const int x, const int y, const int z, PSGrid3DFloat p, PSGrid3DFloat q, PSGrid3DFloat a,
                  PSGrid3DFloat b, float w) {
  float s = PSGridGet(a, x, y, z) * (PSGridGet(p, x + 1, y - 1, z) + PSGridGet(p, x, y, z - 1))
            + PSGridGet(b, %(bidx)s) * PSGridGet(p, x, y, z);
  PSGridEmit(q, %(emit)s);
  return;
}
int main(int argc, char *argv[]) {
  PSInit(&argc, &argv, 3, 16, 16, 16);
  PSGrid3DFloat p = PSGrid3DFloatNew(16, 16, 16);
  PSGrid3DFloat q = PSGrid3DFloatNew(16, 16, 16);
  PSGrid3DFloat a = PSGrid3DFloatNew(16, 16, 16);
  PSGrid3DFloat b = PSGrid3DFloatNew(16, 16, 16);
  PSDomain3D d = PSDomain3DNew(1, 15, 1, 15, 1, 15);
  PSStencilRun(PSStencilMap(relax, d, p, q, a, b, 0.5f), PSStencilMap(relax, d, q, p, a, b, 0.5f), 4);
  PSFinalize();
  return 0;
}
"""


def test_point_form_kernel_gets_update_and_fused_run(tmp_path):
    """A kernel reading its staged grid within radius 1 (diagonals included)
    and other grids only at the point gets update() for psk::PointTB: the
    staged reads become __PSat(dx, dy, dz), the point grids __PSpt(n) in
    parameter order, the emit a return; the run tries the fused sweeps
    first."""
    hip = translate_text(tmp_path, POINT_KERNEL % {"bidx": "x, y, z", "emit": "s"})
    upd = hip[hip.find("static float update("):hip.find("\n  }\n", hip.find("static float update("))]
    assert "((float)__PSpt(0)) * (((float)__PSat(1, -1, 0)) +((float)__PSat(0, 0, -1)))" in upd
    assert "((float)__PSpt(1)) *((float)__PSat(0, 0, 0))" in upd
    assert "return (float)(s);" in upd and "float w = __s.w;" in upd
    assert "static constexpr int NG = 2;" in hip
    assert "__PSPointFusable_relax(&s0, &s1)" in hip and "__PSRunPointTB_relax(&s0, iter)" in hip
    assert "a->dom.local_min[d] != 1 || a->dom.local_max[d] != g->dim[d] - 1" in hip


@pytest.mark.parametrize("bidx,emit", [("x, y, z + 1", "s"), ("x, y, z", "s + (float)x")])
def test_kernels_outside_the_point_form_run_unfused(tmp_path, bidx, emit):
    """A second grid read off the point, or an index name used outside a
    grid read (update() has no coordinates), keeps the per-sweep path."""
    hip = translate_text(tmp_path, POINT_KERNEL % {"bidx": bidx, "emit": emit})
    assert "update(" not in hip and "__PSPointFusable_" not in hip


REDUCE_KERNEL = """
#include "physis/physis.h"
%(ret)s energy(const int x, const int y, const int z, PSGrid3DFloat g, PSGrid3DDouble h, float w) {
  float d = PSGridGet(g, x + 1, y, z) - PSGridGet(g, x, y, z);
  %(body)s
}
int main(int argc, char *argv[]) {
  PSInit(&argc, &argv, 3, 16, 16, 16);
  PSGrid3DFloat g = PSGrid3DFloatNew(16, 16, 16);
  PSGrid3DDouble h = PSGrid3DDoubleNew(16, 16, 16);
  %(ret)s v;
  PSReduce(&v, PS_SUM, energy, PSDomain3DNew(0, 15, 0, 16, 0, 16), g, h, 0.5f);
  PSFinalize();
  return 0;
}
"""


def test_reduction_kernel_lowered_to_a_value_plan(tmp_path):
    """PSReduce(&v, op, kernel, dom, args...) becomes __PSReduceMap_<k>(&v,
    op, dom, args...); the kernel becomes a plan whose value() is the body
    with its grid reads lowered and its return kept (no staged tile), run
    through psk::RunReduceMap with float grids dispatched by storage."""
    hip = translate_text(tmp_path, REDUCE_KERNEL % {"ret": "float",
                                                     "body": "return w * d * d + (float)PSGridGet(h, x, y, z);"})
    assert "__PSReduceMap_energy(& v, PS_SUM, PSDomain3DNew(0, 15, 0, 16, 0, 16), g, h, 0.5f);" in hip
    val = hip[hip.find("static float value("):hip.find("\n  }\n", hip.find("static float value("))]
    assert "return w * d * d + (float)psk::Get3<__kU>(h, __ref, x, y, z);" in val and "__ctx" not in val
    assert "psk::GridRef<psk::Native<double>> h;" in hip and "psk::GridRef<SF> g;" in hip
    assert "psk::RunReduceMap<__PSRPlan_energy<SF>>(__a, __box, __op, &__r" in hip
    # One element index per point when every grid has the same extents.
    assert "if (psk::UniformExtents(__a.g, __a.h)) {" in hip
    assert "psk::RunReduceMap<__PSRPlan_energy<SF, true>>(__u, __box, __op, &__r" in hip
    assert "case PS_HIP_STORAGE_BF16: __PSReduceMapT_energy<psk::BF16>" in hip
    # The kernel's definition does not reach the host code.
    assert "energy(const int x" not in hip


@pytest.mark.parametrize("ret,body,msg", [
    ("int", "return (int)d;", "must return float or double"),
    ("float", "PSGridEmit(g, d); return d;", "must return its value"),
])
def test_reduction_kernels_outside_the_interface_are_rejected(tmp_path, ret, body, msg):
    if not os.path.exists(PHYSISC):
        pytest.skip("physisc-hip is not built (make -C physis_amd/translator)")
    src = tmp_path / "r.c"
    src.write_text(REDUCE_KERNEL % {"ret": ret, "body": body})
    out = subprocess.run([PHYSISC, "-o", str(tmp_path / "r.hip"), str(src)], capture_output=True, text=True)
    assert out.returncode != 0 and msg in out.stderr


def test_psgridreduce_is_an_alias_of_psreduce(tmp_path):
    """PSGridReduce(&v, op, g) and PSGridReduce(&v, op, kernel, dom, ...)
    lower exactly as PSReduce does."""
    text = (REDUCE_KERNEL % {"ret": "float", "body": "return d;"}).replace("PSReduce(", "PSGridReduce(")
    text = text.replace("  PSFinalize();", "  float t;\n  PSGridReduce(&t, PS_MAX, g);\n  PSFinalize();")
    hip = translate_text(tmp_path, text)
    assert "__PSReduceMap_energy(& v, PS_SUM, PSDomain3DNew(0, 15, 0, 16, 0, 16), g, h, 0.5f);" in hip
    assert "__PSReduceGridFloat(& t, PS_MAX, g)" in hip
