"""Regenerates tests/golden/manual_ref.json from the reference's own ref oracles.

The reference holds hand-written sequential C oracles for its system tests
(tests/system_tests/test_cases/*.manual.ref.c).  `make -C oracle ref` compiles
them from their sources into oracle/_ref/; this script runs each one and keeps
the SHA-256 of its complete stdout (one "%f" line per grid point, the text the
reference's run_system_tests.sh diffs, run_system_tests.sh.cmake:808-856) plus
every 128th line for diagnostics.  Run it in a container that has the
reference checkout; the JSON it writes is the committed fixture.
"""
import hashlib
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
SAMPLE_EVERY = 128


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref", "-j8"], check=True,
                   stdout=subprocess.DEVNULL)
    out = {}
    for name in sorted(os.listdir(REF)):
        if not name.endswith(".manual.ref.exe"):
            continue
        test = name[: -len(".manual.ref.exe")]
        text = subprocess.run([os.path.join(REF, name)], check=True, capture_output=True).stdout
        lines = text.decode().splitlines()
        out[test] = {
            "sha256": hashlib.sha256(text).hexdigest(),
            "lines": len(lines),
            "sample_every": SAMPLE_EVERY,
            "samples": lines[::SAMPLE_EVERY],
        }
    with open(os.path.join(HERE, "manual_ref.json"), "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
    print("wrote %d fixtures" % len(out), file=sys.stderr)


if __name__ == "__main__":
    main()
