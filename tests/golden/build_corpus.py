"""Builds the reference's DSL system tests for the MI355X target.

Each tests/system_tests/test_cases/test_*.c of the reference checkout is
translated with physisc-hip and compiled with hipcc into tests/_corpus_bin/
(ignored by git; the executables travel with the tree to the GPU machine,
where tests/test_gpu_corpus.py runs them and compares their stdout with the
SHA-256 of the reference's own manual.ref outputs in manual_ref.json).
Programs this target does not support would be listed in SKIPPED with the
reason; there are none.  A module test (test_*.module.c, a DSL unit without
main) is linked with its plain C driver test_*.module_base.c.  The
reference's diffusion benchmark (examples/diffusion-benchmark) is built the
same way: its Physis program translated, its C++ harness compiled for the
Physis variant and linked in (tests/_corpus_bin/diffusion_benchmark_physis).
The reference's Himeno program is built the same way (himeno_physis, and
himeno_dropin_timing for the benchmark, see build_himeno).  The DSL programs of this repository under
tests/golden/programs are built into the same directory.
"""
import json
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CASES = "/root/reference/tests/system_tests/test_cases"
OUT = os.path.join(ROOT, "tests", "_corpus_bin")

SKIPPED = {}

# The diffusion benchmark of the reference (examples/diffusion-benchmark):
# its Physis program, translated by physisc-hip, linked with the benchmark's
# own C++ harness built for the Physis variant (main.cc with -DPHYSIS).
# A second build with TRACE_KERNEL (the reference's tracing configuration
# key): its PSStencilRun waits for its kernels and reports their time with
# --physis-trace, which times the drop-in path without the host copies.
BENCH_DIR = "/root/reference/examples/diffusion-benchmark"
BENCH_EXE = os.path.join(OUT, "diffusion_benchmark_physis")
BENCH_TRACED_EXE = os.path.join(OUT, "diffusion_benchmark_physis_traced")
# The same traced program driven by a timing harness of this repository
# (tests/golden/dropin_timing.cc): a warm run, then the timed run.
DROPIN_EXE = os.path.join(OUT, "diffusion_dropin_timing")


def build_benchmark(phc):
    if not os.path.isdir(BENCH_DIR):
        return True
    cfg = os.path.join(OUT, "trace_kernel.cfg")
    with open(cfg, "w") as f:
        f.write("TRACE_KERNEL = true\n")
    try:
        for exe, config in ((BENCH_EXE, None), (BENCH_TRACED_EXE, cfg)):
            phc.build(os.path.join(BENCH_DIR, "diffusion3d_physis.c"), exe, config=config,
                      extra_cxx=[os.path.join(BENCH_DIR, f) for f in ("main.cc", "diffusion3d.cc", "baseline.cc")],
                      cxx_flags=["-DPHYSIS", "-I" + BENCH_DIR, "-I" + os.path.join(ROOT, "include"), "-w"])
            os.remove(exe + ".hip")
        phc.build(os.path.join(BENCH_DIR, "diffusion3d_physis.c"), DROPIN_EXE, config=cfg,
                  extra_cxx=[os.path.join(HERE, "dropin_timing.cc")], cxx_flags=["-w"])
        os.remove(DROPIN_EXE + ".hip")
        return True
    except subprocess.CalledProcessError:
        return False


# The reference's Himeno program (examples/himeno/himenobmtxpa_physis.c, the
# 19-point Jacobi of the Himeno workload), translated unchanged with its
# ENABLE_DUMP block on (it writes p after the four rehearsal sweeps to
# himeno.physis.dat) and its wall clock replaced by himeno_clock.c, so that
# its one-minute measured loop runs 18 sweeps.
HIMENO_SRC = "/root/reference/examples/himeno/himenobmtxpa_physis.c"
HIMENO_EXE = os.path.join(OUT, "himeno_physis")
# The same program translated with TRACE_KERNEL and its main renamed, driven
# by a timing harness of this repository (tests/golden/himeno_timing.cc).
HIMENO_TIMING_EXE = os.path.join(OUT, "himeno_dropin_timing")


def build_himeno(phc):
    if not os.path.exists(HIMENO_SRC):
        return True
    defines = ["ENABLE_DUMP", "gettimeofday=ps_himeno_clock"]
    hip = HIMENO_EXE + ".hip"
    try:
        phc.translate(HIMENO_SRC, hip, None, defines)
        phc.compile_hip(hip, HIMENO_EXE, extra_c=[os.path.join(HERE, "himeno_clock.c")], defines=defines)
        os.remove(hip)
        cfg = os.path.join(OUT, "trace_kernel.cfg")
        with open(cfg, "w") as f:
            f.write("TRACE_KERNEL = true\n")
        hip = HIMENO_TIMING_EXE + ".hip"
        phc.translate(HIMENO_SRC, hip, cfg, ["main=himeno_program_main"])
        phc.compile_hip(hip, HIMENO_TIMING_EXE, extra_cxx=[os.path.join(HERE, "himeno_timing.cc")],
                        cxx_flags=["-I" + os.path.join(ROOT, "include"), "-w"])
        os.remove(hip)
        return True
    except subprocess.CalledProcessError:
        return False


# DSL programs of this repository (tests/golden/programs), built the same way.
OWN = os.path.join(HERE, "programs")


def build_own(phc):
    failed = []
    for name in sorted(os.listdir(OWN)):
        if not name.endswith(".c"):
            continue
        exe = os.path.join(OUT, name[:-2])
        try:
            phc.build(os.path.join(OWN, name), exe)
            os.remove(exe + ".hip")
        except subprocess.CalledProcessError:
            failed.append(name)
    return failed


def main(jobs=8):
    sys.path.insert(0, ROOT)
    from physis_amd import phc
    os.makedirs(OUT, exist_ok=True)
    own_failed = build_own(phc)
    if own_failed:
        print("own programs failed: %s" % own_failed, file=sys.stderr)
    if not os.path.isdir(CASES):
        print("reference corpus not present; built only this repository's programs", file=sys.stderr)
        return 1 if own_failed else 0
    stems = []
    for name in sorted(os.listdir(CASES)):
        if not name.startswith("test_") or not name.endswith(".c") or ".manual." in name or \
                name.endswith(".module_base.c"):
            continue
        stem = name[:-2]
        if not any(k in stem for k in SKIPPED):
            stems.append(stem)

    def one(stem):
        exe = os.path.join(OUT, stem)
        # A module test is a DSL unit without main plus a plain C driver.
        extra = [os.path.join(CASES, stem[:-len(".module")] + ".module_base.c")] if stem.endswith(".module") else []
        try:
            phc.build(os.path.join(CASES, stem + ".c"), exe, extra_c=extra)
            os.remove(exe + ".hip")
            return stem, True
        except subprocess.CalledProcessError:
            return stem, False

    with ThreadPoolExecutor(max_workers=jobs) as pool:
        results = list(pool.map(one, stems))
    built = [s for s, ok in results if ok]
    failed = [s for s, ok in results if not ok]
    (built if build_benchmark(phc) else failed).append("diffusion_benchmark_physis")
    (built if build_himeno(phc) else failed).append("himeno_physis")
    with open(os.path.join(OUT, "manifest.json"), "w") as f:
        json.dump({"built": built, "failed": failed}, f, indent=1)
    print("built %d, failed %d %s" % (len(built), len(failed), failed), file=sys.stderr)
    return 1 if failed or own_failed else 0


if __name__ == "__main__":
    sys.exit(main())
