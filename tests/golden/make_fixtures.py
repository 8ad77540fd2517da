#!/usr/bin/env python3
"""Write the reference-derived fixtures of tests/golden/ (run in the container that
mounts the reference; the fixtures are committed, the tests never read the
reference at run time).

* GRD_gz.dat, GRD_gzh.dat  -- the reference's 96-level vertical grid
                               (run/data/vgrid/, raw little-endian float64 data)
* reference_constants.json -- numeric constants of the path parsed from the
                               reference sources: problemsize.h constants, the
                               GlobalParams defaults, the Roh & Satoh coefficient
                               arrays and the Bergeron tables.
* snapshot_files.json      -- names and reader ranks of the snapshot files the
                               stand-alone run reads (data/, ref_verify/).
"""
import json
import os
import re
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))

NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?"


def read(rel):
    with open(os.path.join(REF, rel)) as fh:
        return fh.read()


def strip_comments(s):
    s = re.sub(r"/\*.*?\*/", "", s, flags=re.S)
    return re.sub(r"//[^\n]*", "", s)


def main():
    for name in ("GRD_gz.dat", "GRD_gzh.dat"):
        shutil.copyfile(os.path.join(REF, "run", "data", "vgrid", name), os.path.join(OUT, name))

    ps = strip_comments(read("include/problemsize.h"))
    consts = {m.group(1): float(m.group(2)) for m in re.finditer(r"constexpr double (CONST_\w+)\s*=\s*(%s)" % NUM, ps)}

    hdr = strip_comments(read("include/mod_mp_nsw6.h"))
    body = hdr[hdr.index("struct GlobalParams"):hdr.index("extern GlobalParams para")]
    defaults = {m.group(1): float(m.group(2)) for m in re.finditer(r"double\s+(\w+)\s*=\s*(%s)\s*;" % NUM, body)}
    arrays = {}
    for m in re.finditer(r"double\s+(coef_[ab])\[10\]\s*=\s*\{([^}]*)\}", body):
        arrays[m.group(1)] = [float(x) for x in m.group(2).split(",")]

    src = read("src/mod_mp_nsw6.cpp")
    k = src.index("void Bergeron_param( const View2D")
    seg = src[k:]
    tabs = {}
    for name in ("arr_a1_tab", "arr_a2_tab"):
        m = re.search(r"double %s\[32\]\s*=\s*\{([^}]*)\}" % name, seg)
        tabs[name.replace("arr_", "")] = [float(x) for x in m.group(1).split(",")]

    out = {"constants": consts, "globalparams_defaults": defaults, "roh_coefficients": arrays,
           "bergeron": tabs,
           "source": "parsed from include/problemsize.h, include/mod_mp_nsw6.h, src/mod_mp_nsw6.cpp"}
    with open(os.path.join(OUT, "reference_constants.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(f"wrote {len(consts)} constants, {len(defaults)} defaults, tables {list(tabs)}")

    # the snapshot files of the stand-alone run: reader rank and name, in read order
    main_src = strip_comments(read("src/main.cpp"))
    files = [{"rank": int(m.group(1)), "dir": m.group(2), "name": m.group(3)}
             for m in re.finditer(r'read_data_(\d)d\("(data|ref_verify)/(\w+)\.dat"', main_src)]
    with open(os.path.join(OUT, "snapshot_files.json"), "w") as fh:
        json.dump({"files": files, "source": "parsed from src/main.cpp"}, fh, indent=1)
    print(f"wrote {len(files)} snapshot file names")


if __name__ == "__main__":
    main()
