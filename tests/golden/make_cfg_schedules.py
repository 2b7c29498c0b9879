"""Writes reference_schedules.json: the `algorithm` schedule strings of the
reference's three configuration files (html/cdi.cfg, html/default/cdi.cfg,
html/default/pulse.cfg), read as text from the reference tree, with the token
count each expands to under the reference's AlgoParser
(src/independant/readConfig.cc:82-163).  The counts are the known answers
stated for these files: 100 x 25 + 21 = 2521, 200 x 22 + 100 = 4500 and
3000 + 20 = 3020 (iterations and shrink-wrap tokens).

Run from the repository root:  python tests/golden/make_cfg_schedules.py
"""
import json
import os
import re

REF = "/root/reference"
FILES = {"html/cdi.cfg": 2521, "html/default/cdi.cfg": 4500, "html/default/pulse.cfg": 3020}


def main():
    out = []
    for rel, count in FILES.items():
        text = open(os.path.join(REF, rel)).read()
        m = re.search(r'^\s*algorithm\s*=\s*"([^"]*)"', text, re.M)
        out.append({"file": rel, "algorithm": m.group(1), "tokens": count})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_schedules.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
