"""Print the reference's Fortran sources in module-dependency order.

Build helper for oracle/ref_build.mk (this container only).  The reference's
CMakeLists.txt:62-304 lists its sources by layer and lets CMake's Fortran
dependency scanner order the compilation; flang driven by a plain Makefile
needs that order spelled out, so this script scans every source for
``MODULE x`` definitions and ``USE x`` statements (including those in files
pulled in by ``INCLUDE "*.func.h"``) and prints a topological order.

Usage: python3 oracle/ref_order.py REFDIR file1.F90 file2.F90 ...
"""
import os
import re
import sys

_MOD = re.compile(r"^\s*MODULE\s+(\w+)\s*$", re.I | re.M)
_USE = re.compile(r"^\s*USE\s+(\w+)", re.I | re.M)
_INC = re.compile(r'^\s*INCLUDE\s+["\']([^"\']+)["\']', re.I | re.M)


def _text(path, incdirs):
    with open(path, errors="replace") as f:
        txt = f.read()
    # strip comments so that commented-out USE lines do not create edges
    txt = "\n".join(line.split("!", 1)[0] for line in txt.splitlines())
    for inc in _INC.findall(txt):
        for d in incdirs:
            p = os.path.join(d, inc)
            if os.path.exists(p):
                txt += "\n" + _text(p, incdirs)
                break
    return txt


def main():
    ref = sys.argv[1]
    srcs = sys.argv[2:]
    incdirs = [os.path.join(ref, "micro"), os.path.join(ref, "stubs", "include")]
    defines, uses = {}, {}
    for s in srcs:
        t = _text(os.path.join(ref, s), incdirs)
        for m in _MOD.findall(t):
            if m.upper() != "PROCEDURE":
                defines[m.upper()] = s
        uses[s] = {u.upper() for u in _USE.findall(t)}
    order, state = [], {}

    def visit(s):
        if state.get(s) == 2:
            return
        if state.get(s) == 1:
            return  # cycle through a self-use; the CMake order resolves it
        state[s] = 1
        for u in sorted(uses[s]):
            d = defines.get(u)
            if d is not None and d != s:
                visit(d)
        state[s] = 2
        order.append(s)

    for s in srcs:
        visit(s)
    print(" ".join(order))


if __name__ == "__main__":
    main()
