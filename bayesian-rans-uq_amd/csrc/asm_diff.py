"""Compare one kernel's gfx950 assembly between two builds.

`make -C bayesian-rans-uq_amd/csrc asm` writes the device assembly of both
kernel parts to build/kernels_f64.s and build/kernels_f32.s. Keep a copy
from a build that passed the GPU tests. After an edit, build the new
assembly and compare a kernel's body with

    python3 asm_diff.py OLD.s NEW.s grid_solve_kernel

The pattern selects the first kernel whose mangled name contains it. Basic
block labels and branch-helper labels are renumbered before the comparison,
and comments are dropped. The exit status is 0 when the bodies are the same
instruction for instruction, 1 when they differ (the first lines of the
difference are printed) and 2 when no kernel matches. The multi-workgroup
kernels are checked this way before an edit that reaches them is run on a
GPU.
"""
from __future__ import annotations

import difflib
import re
import sys


def kernel_body(path: str, pattern: str):
    name, body, inside = None, [], False
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\n")
            if not inside:
                m = re.match(r"^(_Z\S*):", line)
                if m and pattern in m.group(1):
                    name, inside = m.group(1), True
                continue
            if line.startswith(".Lfunc_end"):
                break
            body.append(line)
    return name, body


def normalised(lines):
    out = []
    for line in lines:
        line = line.split(";")[0].rstrip()
        if not line.strip():
            continue
        line = re.sub(r"\.LBB\d+_\d+", ".LBB", line)
        line = re.sub(r"\.Lpost_getpc\d+", ".Lpost_getpc", line)
        out.append(line)
    return out


def main(argv) -> int:
    if len(argv) != 4:
        print(__doc__)
        return 2
    old_path, new_path, pattern = argv[1:]
    old_name, old = kernel_body(old_path, pattern)
    new_name, new = kernel_body(new_path, pattern)
    if old_name is None or new_name is None:
        print(f"no kernel matching {pattern!r} in {old_path if old_name is None else new_path}")
        return 2
    a, b = normalised(old), normalised(new)
    print(f"{old_name}: {len(a)} / {len(b)} instructions and directives")
    if a == b:
        print("identical")
        return 0
    diff = list(difflib.unified_diff(a, b, lineterm="", n=0))
    print(f"different ({len(diff)} diff lines)")
    print("\n".join(diff[:40]))
    return 1


if __name__ == "__main__":
    sys.exit(main(sys.argv))
