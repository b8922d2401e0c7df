"""Writes faddeeva_kat.json: the known-answer table of w(z) that the reference's
vendored Faddeeva package holds in its self-test (qgf2_src/Faddeeva.cc:1918-2075,
values from WolframAlpha/Maple according to that file).  Run in a checkout that
has the reference under /root/reference; the JSON is committed."""
import json
import math
import os
import re

SRC = "/root/reference/qgf2_src/Faddeeva.cc"
HERE = os.path.dirname(os.path.abspath(__file__))


def num(tok):
    tok = tok.strip()
    if tok in ("Inf", "+Inf"):
        return math.inf
    if tok == "-Inf":
        return -math.inf
    if tok == "NaN":
        return math.nan
    return float(tok)


def table(text, name):
    start = text.index("cmplx %s[NTST] = {" % name)
    body = text[start:text.index("};", start)]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    pairs = re.findall(r"C\(\s*([^,]+?)\s*,\s*([^)]+?)\s*\)", body)
    return [(num(a), num(b)) for a, b in pairs]


def main():
    text = open(SRC).read()
    w_test = text[text.index("w(z) tests"):]
    z = table(w_test, "z")
    w = table(w_test, "w")
    assert len(z) == len(w) == 57, (len(z), len(w))
    rows = [{"z": list(a), "w": list(b)} for a, b in zip(z, w)
            if all(math.isfinite(v) for v in a + b)]
    out = {"source": "qgf2_src/Faddeeva.cc:1918-2075 (w(z) self-test table, finite entries)", "cases": rows}
    with open(os.path.join(HERE, "faddeeva_kat.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(len(rows), "finite cases")


if __name__ == "__main__":
    main()
