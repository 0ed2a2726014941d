"""Translate-and-build driver for Physis DSL programs on the MI355X target.

The reference builds a DSL program as `physisc-cuda foo.c` followed by nvcc
and a link against libphysis_rt_cuda.a (docs/compilation.md:11-53).  Here:

    python -m physis_amd.phc foo.c -o foo            # translate + hipcc + link
    python -m physis_amd.phc foo.c --emit-only       # just write foo.hip

The executable links libphysis_rt_hip.so from physis_amd/lib through an
rpath relative to the executable when it is placed under physis_amd/, else
an absolute rpath.
"""
import argparse
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PHYSISC = os.path.join(HERE, "bin", "physisc-hip")
LIB_DIR = os.path.join(HERE, "lib")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-w"]


def translate(src, out_hip, config=None, defines=()):
    if not os.path.exists(PHYSISC):
        raise RuntimeError("physisc-hip is not built (make -C physis_amd/translator)")
    cmd = [PHYSISC, src, "-o", out_hip]
    if config:
        cmd += ["--config", config]
    for d in defines:
        cmd.append("-D" + d)
    subprocess.run(cmd, check=True)
    return out_hip


def compile_hip(hip, exe, extra_c=(), extra_cxx=(), cxx_flags=(), defines=()):
    """hipcc the generated file and link it with libphysis_rt_hip; extra_c
    are plain C files of the program (a module's host side), compiled with
    the system C compiler and linked in, as the reference's test driver does
    for its module tests (run_system_tests.sh.cmake:605-631); extra_cxx are
    plain C++ files (a benchmark harness), compiled with the system C++
    compiler and cxx_flags; defines (NAME or NAME=V) also apply to the
    headers the generated file includes."""
    out_dir = os.path.dirname(os.path.abspath(exe))
    rel = os.path.relpath(LIB_DIR, out_dir)
    rpath = "$ORIGIN/" + rel if os.path.abspath(out_dir).startswith(HERE) else LIB_DIR
    objs = []
    for c in extra_c:
        obj = exe + "." + os.path.splitext(os.path.basename(c))[0] + ".o"
        subprocess.run(["gcc", "-O2", "-c", c, "-o", obj], check=True)
        objs.append(obj)
    for c in extra_cxx:
        obj = exe + "." + os.path.splitext(os.path.basename(c))[0] + ".o"
        subprocess.run(["g++", "-O3", "-ffp-contract=off"] + list(cxx_flags) + ["-c", c, "-o", obj], check=True)
        objs.append(obj)
    cmd = [HIPCC] + HIP_FLAGS + ["-D" + d for d in defines] + ["-I", INCLUDE, "-x", "hip", hip, "-x", "none"] + objs + \
          ["-L", LIB_DIR, "-lphysis_rt_hip", "-Wl,-rpath," + rpath, "-o", exe]
    subprocess.run(cmd, check=True)
    for obj in objs:
        os.remove(obj)
    return exe


def build(src, exe, config=None, defines=(), keep_hip=None, extra_c=(), extra_cxx=(), cxx_flags=()):
    hip = keep_hip or exe + ".hip"
    translate(src, hip, config, defines)
    return compile_hip(hip, exe, extra_c, extra_cxx, cxx_flags)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("source")
    ap.add_argument("-o", "--output")
    ap.add_argument("--config")
    ap.add_argument("-D", dest="defines", action="append", default=[])
    ap.add_argument("--emit-only", action="store_true")
    ap.add_argument("--with-c", dest="extra_c", action="append", default=[],
                    help="plain C source of the program to compile and link in")
    args = ap.parse_args(argv)
    stem = os.path.splitext(os.path.basename(args.source))[0]
    if args.emit_only:
        translate(args.source, args.output or stem + ".hip", args.config, args.defines)
        return 0
    build(args.source, args.output or stem, args.config, args.defines, extra_c=args.extra_c)
    return 0


if __name__ == "__main__":
    sys.exit(main())
