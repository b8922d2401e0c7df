"""Transcribes the field and force values that the reference's example1 printed
for its shipped samples (example1.log next to each ex.dat) into JSON fixtures,
and copies those ex.dat files (the reference's solved models) next to them.
Run where the reference tree is present; the outputs are committed."""
import json
import os
import re
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLES = {
    # fixture name: (directory, ex.dat name, object radius used for the torque scale)
    "sample": ("", "ex.dat", 5e-3),
    "analysis_sample1": ("analysis_sample1", "ex1.dat", 5e-3),
    "analysis_sample2": ("analysis_sample2", "ex2.dat", 5e-3),
    "verification": ("verification", "veri.dat", 5e-3),
}
NUM = r"[-+]?\d\.\d+e[-+]\d+"


def parse(path):
    lines = open(path).read().split("\n")
    out = {"field": {}, "force": {}}
    for i, l in enumerate(lines):
        if l.startswith("sound pressure and particle velocity at r="):
            r = [float(x) for x in re.findall(r"[-+]?\d*\.?\d+(?:e[-+]?\d+)?", l.split("=")[1])]
            t = re.search(r"type=(\d)", lines[i + 1]).group(1)
            vals = []
            for j in range(4):
                a = re.findall(NUM, lines[i + 2 + j])
                vals.append([float(a[0]), float(a[1])])
            out["r"] = r
            out["field"][t] = {"p": vals[0], "v": vals[1:]}
        if l.startswith("type=") and "Gauss" in l and i + 2 < len(lines) and lines[i + 1].startswith("F = ("):
            t = re.search(r"type=(\d)", l).group(1)
            out["force"][t] = {"F": [float(x) for x in re.findall(NUM, lines[i + 1])],
                               "N": [float(x) for x in re.findall(NUM, lines[i + 2])][:3]}
    out["rc"] = [0.0, 0.0, 0.0]
    return out


def main():
    for name, (d, dat, a) in SAMPLES.items():
        src = os.path.join(REF, d)
        j = parse(os.path.join(src, "example1.log"))
        j["source"] = os.path.join(d, "example1.log") + " of the reference (example1.c on %s)" % dat
        j["radius"] = a
        assert set(j["field"]) == {"0", "1"} and set(j["force"]) == {"0", "1"}, name
        json.dump(j, open(os.path.join(HERE, name + "_example1.json"), "w"), indent=1)
        if name != "sample":
            shutil.copyfile(os.path.join(src, dat), os.path.join(HERE, name + "_" + dat))
        print(name, "ok")


if __name__ == "__main__":
    main()
