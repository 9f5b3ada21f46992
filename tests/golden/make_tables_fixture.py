"""Write tests/golden/strassen_tables.json: the Strassen tables of the reference, as data.

Reads (as text) the reference's table file and the product lists of its ABC drivers:
  src/stra_internal/3t/strassen                       U, V, W (4 x 7 each)
  src/stra_internal/3t/stra_mult.cxx:426-591          7 products of stra_contract_blis (L=1)
  src/stra_internal/3t/stra_mult_2level.cxx:341-1458  49 products of stra_contract_2level_blis (L=2)
and stores, per product, the sub-block ids and coefficients of the A, B and C sums.  Run it in a
container that has the reference mounted; the JSON it writes is committed and read by the tests.
"""
import json
import os
import re
import sys

REF = os.environ.get("TBLIS_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "strassen_tables.json")


def parse_table():
    rows = []
    with open(os.path.join(REF, "src/stra_internal/3t/strassen")) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            rows.append([int(x) for x in line.split()])
    assert len(rows) == 12, len(rows)
    return rows[0:4], rows[4:8], rows[8:12]


def first_defs(path, start_line, count):
    """Parse product definitions from start_line, stopping after `count` products."""
    with open(os.path.join(REF, path)) as f:
        lines = f.readlines()
    sub = re.compile(r"std::array<unsigned,\s*\d+>\s*([ABC])(\d+)_subid\s*=\s*\{([^}]*)\};")
    coe = re.compile(r"std::array<T,\s*\d+>\s*([ABC])(\d+)_coeff_list\s*=\s*\{([^}]*)\};")
    prods = {}
    for line in lines[start_line - 1:]:
        m = sub.search(line)
        if m:
            op, p = m.group(1), int(m.group(2))
            if op in prods.get(p, {}) and "sub" in prods[p][op]:
                break
            prods.setdefault(p, {}).setdefault(op, {})["sub"] = [int(x) for x in m.group(3).split(",")]
        m = coe.search(line)
        if m:
            op, p = m.group(1), int(m.group(2))
            prods.setdefault(p, {}).setdefault(op, {})["coeff"] = [float(x) for x in m.group(3).split(",")]
        if len(prods) == count and all(len(prods[p]) == 3 and all("coeff" in prods[p][o] for o in "ABC")
                                       for p in prods):
            break
    out = []
    for p in range(count):
        out.append({op: [[s, c] for s, c in zip(prods[p][op]["sub"], prods[p][op]["coeff"])] for op in "ABC"})
    return out


def main():
    U, V, W = parse_table()
    l1 = first_defs("src/stra_internal/3t/stra_mult.cxx", 308, 7)
    l2 = first_defs("src/stra_internal/3t/stra_mult_2level.cxx", 241, 49)
    data = {
        "source": "flame/tblis-strassen: src/stra_internal/3t/strassen, stra_mult.cxx:426-591, "
                  "stra_mult_2level.cxx:341-1458",
        "U": U, "V": V, "W": W, "level1": l1, "level2": l2,
    }
    with open(OUT, "w") as f:
        json.dump(data, f, indent=0)
    print(f"wrote {OUT}: {len(l1)} + {len(l2)} products")


if __name__ == "__main__":
    sys.exit(main())
