"""Writes tests/golden/examples_1d.json from the reference's Examples.

Run in a tree that has the reference sources (default /root/reference):

    python tests/golden/make_fixtures.py [REFERENCE_ROOT]

For each example it records the input parameters the nonlocal-density fill
and its Newton solve consume (read from the positional `@` lines of
dft_input.dat and from dft_surfaces.dat) and the "Key Output Parameters" of
the example's README, which the reference's test harness checks with an
absolute tolerance of 1e-3 and niter +-4 (Examples/CMakeLists.txt:14-17,
Examples/check_results/main.cpp:120-139).  The fixture holds data only.
"""
import json
import os
import re
import sys

EXAMPLES = ["LJ1_FMT1_WCA_1D", "LJ2_FMT1_WCA_1D", "LJ3_FMT1_WCA_1D", "LJ5_FMT1_WCA2_1D", "MIX2_FMT1_WCA_1D",
            "MIX11_FMT1_WCA_1D"]

NUM = re.compile(r"^[-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?$")


def at_lines(path):
    out = []
    with open(path) as f:
        for line in f:
            if not line.startswith("@"):
                continue
            toks = line[1:].split()
            vals = []
            for t in toks:
                if NUM.match(t):
                    vals.append(float(t))
                else:
                    break
            label = " ".join(toks[len(vals):])
            out.append((vals, label))
    return out


def find(lines, prefix, start=0):
    for k in range(start, len(lines)):
        if lines[k][1].startswith(prefix):
            return lines[k][0], k
    raise KeyError(prefix)


def parse_example(root, name):
    d = os.path.join(root, "Examples", name)
    L = at_lines(os.path.join(d, "dft_input.dat"))
    g = {}
    v, _ = find(L, "Length_ref")
    g["vext_max"] = v[4]
    g["ndim"] = int(find(L, "Ndim")[0][0])
    g["size_x"] = find(L, "Size_x")[0][: g["ndim"]]
    g["esize_x"] = find(L, "Esize_x")[0][: g["ndim"]]
    g["type_bc_x"] = [int(t) for t in find(L, "Type_bc(x0")[0][:2]]
    v, _ = find(L, "Type_func")
    g["type_func"] = int(v[0])
    v, _ = find(L, "Type_attr")
    g["type_attr"], g["type_pairpot"] = int(v[0]), int(v[1])
    v, _ = find(L, "Nwall_types")
    g["nwall_type"], g["nwall"] = int(v[0]), int(v[1])
    g["surf_type"] = int(find(L, "Surf_type")[0][0])
    g["orientation"] = int(find(L, "Orientation[iwall_type]")[0][0])
    g["wall_param"] = find(L, "WallParam[iwall_type]")[0][0]
    g["ipot_wf_n"] = int(find(L, "Type of neutral surface-fluid")[0][0])
    g["lhard_surf"] = int(find(L, "Lhard_surf")[0][0])
    g["type_vext"] = int(find(L, "Type_vext")[0][0])
    g["vext_potential_id"] = int(find(L, "Vext_PotentialID")[0][0])
    v, _ = find(L, "Ncomp Mix_type")
    g["ncomp"], g["mix_type"] = int(v[0]), int(v[1])
    C = g["ncomp"]
    for key, lab in (("sigma_ff", "Sigma_ff"), ("eps_ff", "Eps_ff"), ("cut_ff", "Cut_ff")):
        g[key] = find(L, lab)[0][: C * C]
    for key, lab in (("rho_w", "Rho_w"), ("sigma_ww", "Sigma_ww"), ("eps_ww", "Eps_ww"), ("cut_ww", "Cut_ww")):
        g[key] = find(L, lab)[0][0]
    for key, lab in (("sigma_wf", "Sigma_wf"), ("eps_wf", "Eps_wf"), ("cut_wf", "Cut_wf")):
        g[key] = find(L, lab)[0][:C]
    g["rho_b"] = find(L, "Rho_b_0")[0][:C]
    v, _ = find(L, "Iguess")
    g["iguess"] = int(v[0])
    v, _ = find(L, "NL_Solver")
    g["nl_solver"], g["max_nl_iter"], g["att_in_a22"] = int(v[0]), int(v[1]), int(v[3])
    v, _ = find(L, "NL_rel_tol NL_abs_tol")
    g["nl_rel_tol"], g["nl_abs_tol"], g["nl_update_scaling"] = v[0], v[1], v[2]
    with open(os.path.join(d, "dft_surfaces.dat")) as f:
        rows = [r.split() for r in f if r.strip()]
    g["walls"] = [{"type": int(r[0]), "pos": [float(t) for t in r[2:2 + g["ndim"]]]} for r in rows]
    with open(os.path.join(d, "README")) as f:
        txt = f.read()
    key = txt[txt.index("Key Output Parameters"):]
    out = {}
    m = re.search(r"niter\s*=\s*(\d+)", key)
    out["niter"] = int(m.group(1))
    ads = re.findall(r"adsorption(?:\[(\d+)\])?\s*=\s*([-\d.eE+]+)", key)
    out["adsorption"] = [float(a[1]) for a in ads]
    m = re.search(r"energy\s*=\s*([-\d.eE+]+)", key)
    if m:
        out["energy"] = float(m.group(1))
    g["key_outputs"] = out
    g["source"] = f"Examples/{name}/dft_input.dat, dft_surfaces.dat, README"
    return g


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = {name: parse_example(root, name) for name in EXAMPLES}
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "examples_1d.json")
    with open(out, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    print("wrote", out)


if __name__ == "__main__":
    main()
