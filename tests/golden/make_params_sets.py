"""Writes params_sets.json: the moduli sets that GRNS 1.1.4 ships as params.h files
(src/params.h and src/params/*/params.*.h), read as text from the reference tree.  Each entry keeps
the set's file name, RNS_MODULI_SIZE, RNS_MODULI_PRODUCT_LOG2 and the moduli in file order.

Run in a container that holds the reference at /root/reference:
    python tests/golden/make_params_sets.py
"""
import glob
import json
import os
import re

REF = "/root/reference/src"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "params_sets.json")


def parse(path):
    text = open(path).read()
    size = int(re.search(r"#define\s+RNS_MODULI_SIZE\s+\((\d+)\)", text).group(1))
    log2 = int(re.search(r"#define\s+RNS_MODULI_PRODUCT_LOG2\s+\((\d+)\)", text).group(1))
    vals = re.search(r"#define\s+RNS_MODULI_VALUES\s+\{([^}]*)\}", text).group(1)
    moduli = [int(v) for v in re.findall(r"\d+", vals)]
    assert len(moduli) == size, path
    return {"file": os.path.relpath(path, os.path.dirname(REF)), "n": size, "log2": log2, "moduli": moduli}


def main():
    files = [os.path.join(REF, "params.h")] + sorted(glob.glob(os.path.join(REF, "params", "*", "params.*.h")))
    sets = [parse(p) for p in files]
    with open(OUT, "w") as f:
        json.dump(sets, f, indent=0)
    print(OUT, len(sets), "sets")


if __name__ == "__main__":
    main()
