"""Writes tests/golden/examples_1d.json and examples_hs_1d.json from the reference's Examples.

Run in a tree that has the reference sources (default /root/reference):

    python tests/golden/make_fixtures.py [REFERENCE_ROOT]

For each example it records the input parameters the nonlocal-density fill
and its Newton solve consume (read from the positional `@` lines of
dft_input.dat and from dft_surfaces.dat) and the "Key Output Parameters" of
the example's README, which the reference's test harness checks with an
absolute tolerance of 1e-3 and niter +-4 (Examples/CMakeLists.txt:14-17,
Examples/check_results/main.cpp:120-139).  The fixture holds data only.
examples_1d.json has the LJ 9-3 wall examples; examples_hs_1d.json the hard
wall ones (Lhard_surf = 1, no attraction; HS8_FMT4_1D with the FMT4 functional) that use no mesh coarsening
(HS6_FMT1_1D and HS7_FMT1_1D set Nzone = 4 with Mesh_coarsening = 1).
"""
import json
import os
import re
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from tramonto_amd.examples import read_tramonto_input  # noqa: E402

EXAMPLES = ["LJ1_FMT1_WCA_1D", "LJ2_FMT1_WCA_1D", "LJ3_FMT1_WCA_1D", "LJ5_FMT1_WCA2_1D", "MIX2_FMT1_WCA_1D",
            "MIX11_FMT1_WCA_1D"]
HARD_WALL_EXAMPLES = ["HS1_FMT1_1D", "HS2_FMT1_1D", "HS3_FMT1_1D", "HS4_FMT1_1D", "HS5_FMT1_1D", "MIX1_FMT1_1D",
                      "HS8_FMT4_1D"]


def parse_example(root, name):
    d = os.path.join(root, "Examples", name)
    g = read_tramonto_input(d)
    with open(os.path.join(d, "README")) as f:
        txt = f.read()
    key = txt[txt.index("Key Output Parameters"):]
    out = {}
    m = re.search(r"niter\s*=\s*(\d+)", key)
    out["niter"] = int(m.group(1))
    ads = re.findall(r"adsorption(?:\[(\d+)\])?\s*=\s*([-\d.eE+]+)", key)
    out["adsorption"] = [float(a[1]) for a in ads]
    for name_, pat in (("energy", r"energy\s*=\s*([-\d.eE+]+)"), ("force", r"force\s*=\s*([-\d.eE+]+)")):
        m = re.search(pat, key)
        if m:
            out[name_] = float(m.group(1))
            out[name_ + "_text"] = m.group(1)
    g["key_outputs"] = out
    g["source"] = f"Examples/{name}/dft_input.dat, dft_surfaces.dat, README"
    return g


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    for names, fname in ((EXAMPLES, "examples_1d.json"), (HARD_WALL_EXAMPLES, "examples_hs_1d.json")):
        data = {name: parse_example(root, name) for name in names}
        out = os.path.join(os.path.dirname(os.path.abspath(__file__)), fname)
        with open(out, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        print("wrote", out)


if __name__ == "__main__":
    main()
