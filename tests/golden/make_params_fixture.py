#!/usr/bin/env python3
"""Extract the reference's constant tables (data) into tests/golden/params_ref.json
and tests/golden/params32_ref.json.

Reads include/yell/params_64.hpp and include/yell/params_32.hpp under
/root/reference as text and records P / Pn / primitive_roots (and, for the
32-bit build, invkMaxPolyDegree): the first 64 entries verbatim and a SHA-256
digest of all entries (decimal, comma-joined), so the generated tables can be
checked without keeping the reference's files here.
"""
import hashlib
import json
import os
import re

REF = "/root/reference/include/yell/params_64.hpp"
REF32 = "/root/reference/include/yell/params_32.hpp"


def arrays(text, names=("P", "Pn", "primitive_roots")):
    out = {}
    for name in names:
        m = re.search(r"\b" + name + r"\[kMaxNbModuli\]\s*=\s*\{([^}]*)\}", text)
        out[name] = [int(x.strip().rstrip("ULL"), 0) for x in m.group(1).split(",") if x.strip()]
    return out


def digest(vals):
    return hashlib.sha256(",".join(str(v) for v in vals).encode()).hexdigest()


def main():
    tabs = arrays(open(REF).read(), ("P", "Pn", "primitive_roots", "invkMaxPolyDegree"))
    fix = {"source": "include/yell/params_64.hpp:27,31,34,38", "count": len(tabs["P"])}
    for k, v in tabs.items():
        fix[k + "_first64"] = [str(x) for x in v[:64]]
        fix[k + "_sha256"] = digest(v)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "params_ref.json")
    with open(path, "w") as f:
        json.dump(fix, f, indent=1)
    tabs = arrays(open(REF32).read(), ("P", "Pn", "primitive_roots", "invkMaxPolyDegree"))
    fix = {"source": "include/yell/params_32.hpp:21,25,28,33", "count": len(tabs["P"])}
    for k, v in tabs.items():
        fix[k + "_first64"] = [str(x) for x in v[:64]]
        fix[k + "_sha256"] = digest(v)
    with open(os.path.join(os.path.dirname(path), "params32_ref.json"), "w") as f:
        json.dump(fix, f, indent=1)


if __name__ == "__main__":
    main()
