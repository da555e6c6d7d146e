"""Writes tests/golden/reference_values.json from output files recorded in the reference
repository (read as text; nothing from the reference is imported or executed).

Sources:
  examples/1D_4/GS.tsv                        E0 and averages of the 4-site chain
  tests/test_diagonalization/S_format.ref     E0 and Lehmann poles of a 2x2 plaquette + 8 baths
  tests/test_diagonalization/test_pole_distribution.py   its full-precision parameters
  tests/test_diagonalization/S_format_complex.ref        E0 and poles of the 8-site chain
  examples/1D_4/sef.tsv                       VCA self-energy functional of the 4-site chain
  examples/1D_2_4b/cdmft.tsv                  converged CDMFT bath of the 2-site + 4-bath chain

Run once in a container where the reference is mounted:
  python tests/golden/extract_reference_values.py /root/reference
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_values.json")


def read(rel):
    with open(os.path.join(REF, rel)) as f:
        return f.read()


def gs_tsv():
    lines = read("examples/1D_4/GS.tsv").splitlines()
    head = lines[0].split("\t")
    out = []
    for ln in lines[1:3]:
        row = dict(zip(head, ln.split("\t")))
        out.append({k: row[k] for k in head if k and k != "githash"})
    return out


def tsv_rows(rel, keys, n=None):
    lines = [ln for ln in read(rel).splitlines() if ln.strip()]
    head = lines[0].split("\t")
    out = []
    for ln in lines[1:] if n is None else lines[1:1 + n]:
        if ln.startswith(head[0] + "\t"):
            continue  # repeated header
        row = dict(zip(head, ln.split("\t")))
        out.append({k: row[k].strip() for k in keys if k in row})
    return out


def parse_poles(text):
    """Blocks 'w\tL\tM' followed by M lines 'e  v_1 ... v_L' (real or '(re,im)')."""
    lines = text.splitlines()
    blocks = []
    i = 0
    while i < len(lines):
        ln = lines[i]
        if ln.startswith("w\t"):
            _, L, M = ln.split("\t")[:3]
            L, M = int(L), int(M)
            e, q = [], []
            for k in range(M):
                parts = lines[i + 1 + k].split("\t")
                e.append(float(parts[0]))
                row = []
                for p in parts[1:1 + L]:
                    p = p.strip()
                    if p.startswith("("):
                        re_, im_ = p.strip("()").split(",")
                        row.append([float(re_), float(im_)])
                    else:
                        row.append([float(p), 0.0])
                q.append(row)
            blocks.append({"L": L, "M": M, "e": e, "Q": q})
            i += M + 1
        else:
            i += 1
    return blocks


def gs_energy(text):
    m = re.search(r"GS_energy:\s*(\S+)\s+GS_sector:\s*(\S+)", text)
    return float(m.group(1)), m.group(2)


def plaquette_params():
    src = read("tests/test_diagonalization/test_pole_distribution.py")
    block = src.split('set_parameters(')[1].split('"""')[1]
    params = {}
    for ln in block.splitlines():
        ln = ln.strip()
        if "=" in ln:
            k, v = ln.split("=")
            params[k.strip()] = float(v)
    return params


def main():
    sref = read("tests/test_diagonalization/S_format.ref")
    cref = read("tests/test_diagonalization/S_format_complex.ref")
    E_p, sec_p = gs_energy(sref)
    E_c, sec_c = gs_energy(cref)
    data = {
        "chain4": {
            "source": "examples/1D_4/GS.tsv:1-3 (model examples/1D_4/model_1D_4.py)",
            "rows": gs_tsv(),
        },
        "plaquette_8bath": {
            "source": "tests/test_diagonalization/S_format.ref (GS_energy line, 'w' pole blocks); "
                      "parameters from tests/test_diagonalization/test_pole_distribution.py; "
                      "model tests/test_diagonalization/cluster_2x2_2C_8b_C2v_L.py",
            "E0": E_p, "sector": sec_p, "params": plaquette_params(),
            "blocks": parse_poles(sref),
        },
        "chain8_complex": {
            "source": "tests/test_diagonalization/S_format_complex.ref (model_1D_8.py, ti=1, U=1e-9)",
            "E0": E_c, "sector": sec_c, "blocks": parse_poles(cref),
        },
    }
    data["chain4_sef"] = {
        "source": "examples/1D_4/sef.tsv (model examples/1D_4/model_1D_4.py; V, V0m, V1m = 1e-9)",
        "rows": tsv_rows("examples/1D_4/sef.tsv", ["omega", "U", "mu", "t", "cdw", "cdw_1", "E0_1",
                                                    "sector_1"]),
    }
    data["chain2_4b_cdmft"] = {
        "source": "examples/1D_2_4b/cdmft.tsv (model examples/1D_2_4b/model_1D_2_4b.py, "
                  "start values examples/1D_2_4b/QCM_params.def)",
        "rows": tsv_rows("examples/1D_2_4b/cdmft.tsv",
                         ["iterations", "dist_function", "distance", "U", "mu", "t", "eb1_1", "eb2_1",
                          "tb1_1", "tb2_1", "E0_1", "sector_1", "ave_mu", "version"]),
    }
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
