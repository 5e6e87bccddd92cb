"""Writes dltc_drivers.json: the equations of the reference's hard-coded DLTC
CC drivers (DLTC/ccd-t2.cc, DLTC/ccsd-t1.cc, DLTC/ccsd-t2.cc) as data.

Each driver declares tensors (`T[X] = Tensor4(PAR, PAR, HOL, HOL, X)`),
relabels their indices (`setIndexName4(T[X], p3, p4, h1, h2)`), records
summations `s[i] = Summation(T[C], T[A], coef)` (C += coef * A) and
multiplications `m[i] = Multiplication(T[C], T[A], T[B], coef)`
(C += coef * A * B) with the labels current at that point, and runs them in
layers (`exec(SUM|MUL, i, tpN)` between barriers).  This script reads those
lines and records, per driver: tensor kinds, every statement with its index
labels and coefficient, and the layers.  It runs where the reference tree is
present; the JSON it writes is what dtcf_amd.dltc_driver executes.

Run: python dtcf_amd/data/extract_dltc_drivers.py /root/reference/DLTC [out.json]
"""
import json
import os
import re
import sys

DRIVERS = {"ccd_t2": "ccd-t2.cc", "ccsd_t1": "ccsd-t1.cc", "ccsd_t2": "ccsd-t2.cc"}

RE_TENSOR = re.compile(r"T\[(\w+)\]\s*=\s*Tensor[24]\(([^)]*)\)")
RE_LABEL = re.compile(r"setIndexName[24]\(T\[(\w+)\],([^)]*)\)")
RE_SUM = re.compile(r"s\[(\d+)\]\s*=\s*Summation\(T\[(\w+)\],\s*T\[(\w+)\],\s*([-+0-9.eE]+)\)")
RE_MUL = re.compile(r"m\[(\d+)\]\s*=\s*Multiplication\(T\[(\w+)\],\s*T\[(\w+)\],\s*T\[(\w+)\],"
                    r"\s*([-+0-9.eE]+)\)")
RE_POOL = re.compile(r"TaskPool\s+(\w+)\s*=\s*gtc_create")
RE_EXEC = re.compile(r"exec\((SUM|MUL),\s*(\d+),\s*(\w+)\)")


def extract(path):
    tensors, label, sums, muls, pools = {}, {}, {}, {}, []
    layer_of = {}
    for ln, line in enumerate(open(path), 1):
        m = RE_TENSOR.search(line)
        if m:
            args = [a.strip() for a in m.group(2).split(",")]
            assert args[-1] == m.group(1), (path, ln)
            tensors[m.group(1)] = args[:-1]
            continue
        m = RE_LABEL.search(line)
        if m:
            label[m.group(1)] = [a.strip() for a in m.group(2).split(",")]
            continue
        m = RE_SUM.search(line)
        if m:
            i, c, a, coef = int(m.group(1)), m.group(2), m.group(3), float(m.group(4))
            sums[i] = {"c": c, "c_idx": label[c], "a": a, "a_idx": label[a], "coef": coef, "line": ln}
            continue
        m = RE_MUL.search(line)
        if m:
            i, c, a, b, coef = int(m.group(1)), m.group(2), m.group(3), m.group(4), float(m.group(5))
            muls[i] = {"c": c, "c_idx": label[c], "a": a, "a_idx": label[a], "b": b,
                       "b_idx": label[b], "coef": coef, "line": ln}
            continue
        m = RE_POOL.search(line)
        if m:
            layer_of[m.group(1)] = len(pools)
            pools.append([])
            continue
        m = RE_EXEC.search(line)
        if m:
            pools[layer_of[m.group(3)]].append([m.group(1), int(m.group(2))])
    for st in list(sums.values()) + list(muls.values()):
        for key in ("c", "a", "b"):
            if key in st:
                assert len(st[key + "_idx"]) == len(tensors[st[key]]), (path, st)
    return {"tensors": tensors, "sums": [sums[i] for i in sorted(sums)],
            "muls": [muls[i] for i in sorted(muls)], "layers": pools}


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/DLTC"
    out = {}
    for name, fn in DRIVERS.items():
        d = extract(os.path.join(src, fn))
        d["source"] = f"DLTC/{fn}"
        assert [i for i in range(len(d["sums"]))] == sorted(x[1] for l in d["layers"] for x in l if x[0] == "SUM")
        assert [i for i in range(len(d["muls"]))] == sorted(x[1] for l in d["layers"] for x in l if x[0] == "MUL")
        out[name] = d
        print(f"{name}: {len(d['tensors'])} tensors, {len(d['muls'])} multiplications, "
              f"{len(d['sums'])} summations, {len(d['layers'])} layers")
    here = os.path.dirname(os.path.abspath(__file__))
    dst = sys.argv[2] if len(sys.argv) > 2 else os.path.join(here, "dltc_drivers.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
