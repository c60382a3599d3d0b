"""Writes sigma_tables.json: the half-level b values of the reference's two
fixed sigma coordinates, read from the text of its source file.

    python tests/golden/make_sigma_tables.py [path/to/Vert_Coordinate.jl]

The default path is the reference checkout at /root/reference.  The file is
read as text (nothing is run): the literal `b = [...]` array inside
Compute_V197_Sigma (Vert_Coordinate.jl:143-154) and Compute_Old_Model_Sigma
(:157-165), together with each function's `nd`.  The fixture pins the copies
of those tables in the library and in the oracle against typing errors.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT = "/root/reference/src/Atmos_Spectral/Vert_Coordinate.jl"
FUNCTIONS = {"v197": "Compute_V197_Sigma", "mcm": "Compute_Old_Model_Sigma"}


def parse(text, function):
    start = text.index(f"function {function}()")
    end = text.index("\nend", start)
    body = text[start:end]
    nd = int(re.search(r"\bnd\s*=\s*(\d+)", body).group(1))
    literal = re.search(r"\bb\s*=\s*\[(.*?)\]", body, re.S).group(1)
    values = [float(x) for x in re.split(r"[;\s]+", literal.strip()) if x]
    if len(values) != nd + 1:
        raise ValueError(f"{function}: {len(values)} values for nd = {nd}")
    return {"nd": nd, "b": values, "source": f"Vert_Coordinate.jl {function}"}


def main(path=DEFAULT):
    with open(path) as f:
        text = f.read()
    tables = {name: parse(text, fn) for name, fn in FUNCTIONS.items()}
    out = os.path.join(HERE, "sigma_tables.json")
    with open(out, "w") as f:
        json.dump(tables, f, indent=1)
    print("wrote", out, {k: v["nd"] for k, v in tables.items()})


if __name__ == "__main__":
    main(*sys.argv[1:])
