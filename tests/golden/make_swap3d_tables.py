"""Write the edge-swap triangulation tables of the reference as a fixture.

Reads src/Omega_h_swap3d_tables.hpp under the reference checkout given as the
first argument (default /root/reference) and writes swap3d_tables.json next to
this script: for each loop size n = 3..7 the unique triangles (loop vertex
triples), the triangulations ("meshes", as unique-triangle indices), the
unique interior edges (loop vertex pairs) and each triangulation's interior
edges (as unique-edge indices), in the reference's order.  The tests compare
the oracle's and the library's tables with this file.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def ints(text):
    return [int(x) for x in re.findall(r"-?\d+", text)]


def array_body(src, name):
    m = re.search(r"\b" + re.escape(name) + r"\s*\[[^=]*=\s*\{(.*?)\};", src, re.S)
    if not m:
        raise KeyError(name)
    return re.sub(r"//[^\n]*", "", m.group(1))


def main(ref):
    with open(os.path.join(ref, "src", "Omega_h_swap3d_tables.hpp")) as f:
        src = f.read()
    out = {}
    for n in range(3, 8):
        tris = ints(array_body(src, f"triangles_{n}"))
        meshes = ints(array_body(src, f"meshes_{n}"))
        rec = {"triangles": [tris[i:i + 3] for i in range(0, len(tris), 3)],
               "meshes": [meshes[i:i + n - 2] for i in range(0, len(meshes), n - 2)]}
        if n >= 4:
            ue = ints(array_body(src, f"unique_edges_{n}"))
            rec["unique_edges"] = [ue[i:i + 2] for i in range(0, len(ue), 2)]
            rec["mesh_edges"] = [ints(array_body(src, f"edges2unique_{n}_{k}"))
                                 for k in range(len(rec["meshes"]))]
        else:
            rec["unique_edges"] = []
            rec["mesh_edges"] = [[]]
        out[str(n)] = rec
    with open(os.path.join(HERE, "swap3d_tables.json"), "w") as f:
        json.dump(out, f)
        f.write("\n")
    print({n: (len(r["triangles"]), len(r["meshes"]), len(r["unique_edges"]))
           for n, r in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
