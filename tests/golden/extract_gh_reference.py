"""Write tests/golden/gh_reference.json: the Gauss-Hermite nodes and weights
that the reference hard-codes in set_weight_abs (src/ff_dmd_I.cpp:1841-2134),
read from the reference source as text (numbers only; nothing is executed).

Run in a container that has the reference mounted:
    python tests/golden/extract_gh_reference.py /root/reference
"""
import json
import os
import re
import sys


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = open(os.path.join(root, "src", "ff_dmd_I.cpp")).read()
    start = src.index("ForceField_DMD_I::set_weight_abs")
    end = src.index("create_2nd_neigh_lst", start)
    body = src[start:end]
    parts = re.split(r"no_i==(\d+)\)", body)
    out = {}
    for k in range(1, len(parts), 2):
        n = int(parts[k])
        nodes, weights = {}, {}
        for m in re.finditer(r"(wi_0|xi)\[(\d+)\]=([-0-9.eE+]+);", parts[k + 1]):
            (weights if m.group(1) == "wi_0" else nodes)[int(m.group(2))] = m.group(3)
        out[str(n)] = {"nodes": [nodes[i] for i in range(n)],
                       "weights": [weights[i] for i in range(n)]}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gh_reference.json")
    with open(dst, "w") as fh:
        json.dump({"source": "src/ff_dmd_I.cpp:1841-2134 (set_weight_abs)",
                   "rules": out}, fh, indent=1)
    print("wrote", dst, "rules:", sorted(int(k) for k in out))


if __name__ == "__main__":
    main()
