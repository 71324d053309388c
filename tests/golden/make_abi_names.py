"""Writes reference_abi_names.json: every LinProbMgr / ParameterList C-ABI name the
reference's C and C++ sources call or define (src/*.c, src/*.cpp, src/*.h), i.e. the
symbols the reference's C driver and NOX group link against.

Polymer managers (dft_poly_*) are outside the hard-sphere path and are listed
separately.  Run in a container that holds the reference at /root/reference:

    python tests/golden/make_abi_names.py
"""
import glob
import json
import os
import re

SRC = "/root/reference/src"
PAT = re.compile(r"\b(dft_(?:linprobmgr|hardsphere_lin_prob_mgr|basic_lin_prob_mgr|parameterlist|poly_lin_prob_mgr)"
                 r"_[A-Za-z0-9_]+)\s*\(")


def main():
    names = set()
    for fn in sorted(glob.glob(os.path.join(SRC, "*.c")) + glob.glob(os.path.join(SRC, "*.cpp")) +
                     glob.glob(os.path.join(SRC, "*.h"))):
        txt = open(fn, errors="replace").read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)   # commented-out prototypes are not linked
        txt = re.sub(r"//[^\n]*", "", txt)
        names.update(m.group(1) for m in PAT.finditer(txt))
    # the header's misspelt declaration (dft_hardsphere_lin_prob_mgr_wrapper.h:58) is never defined
    # or called; the definition and the caller use ...seta22blockisdiagonal
    names.discard("dft_hardsphere_lin_prob_mgr_seta22blockisdiagona")
    poly = sorted(n for n in names if n.startswith("dft_poly_"))
    hs = sorted(n for n in names if not n.startswith("dft_poly_"))
    out = {"hard_sphere_path": hs, "polymer_out_of_scope": poly}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_abi_names.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(len(hs), "names on the hard-sphere path,", len(poly), "polymer names")


if __name__ == "__main__":
    main()
