"""Regenerate the basis-set data files sto3g.json and p631ss.json.

The numbers are the published STO-3G (Hehre, Stewart, Pople 1969) and
6-31G** (Hariharan, Pople 1973) exponents and contraction coefficients as
PyQuante distributes them in PyQuante/Basis/sto3g.py and
PyQuante/Basis/p631ss.py.  Those modules are dict literals; they are read
here as text with ast.literal_eval (nothing is imported or executed) and
written out as JSON for elements H..Ar.  Run once, in a checkout that has
the reference sources at /root/reference; the JSON files are committed.
"""
import ast
import json
import os
import sys

REF = "/root/reference/PyQuante/Basis"
HERE = os.path.dirname(os.path.abspath(__file__))


def read_basis(path):
    src = open(path).read()
    tree = ast.parse(src)
    for node in tree.body:
        if isinstance(node, ast.Assign) and node.targets[0].id == "basis_data":
            return ast.literal_eval(node.value)
    raise RuntimeError("no basis_data in %s" % path)


def main():
    for mod, out in (("sto3g.py", "sto3g.json"), ("p631ss.py", "p631ss.json")):
        data = read_basis(os.path.join(REF, mod))
        keep = {}
        for z, shells in sorted(data.items()):
            if z > 18:
                continue
            if any(sym not in ("S", "P", "D") for sym, _ in shells):
                continue
            keep[str(z)] = [[sym, [[float(e), float(c)] for e, c in prims]]
                            for sym, prims in shells]
        with open(os.path.join(HERE, out), "w") as f:
            rows = ['"%s": %s' % (z, json.dumps(keep[z]))
                    for z in sorted(keep, key=int)]
            f.write("{\n" + ",\n".join(rows) + "\n}\n")
        print(out, len(keep), "elements")


if __name__ == "__main__":
    sys.exit(main())
