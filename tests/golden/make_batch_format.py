"""Writes batch_output_format.json: for every file of the reference's batch-mode
example output ("Example Output for Batch Mode/", a 257x257-node run), its line
count, fields per line and first two lines -- the shape and number formats the
diagnostic writers must reproduce.  Data only; run where the reference tree is
present:  python tests/golden/make_batch_format.py /root/reference
"""
import json
import os
import sys

src = os.path.join(sys.argv[1] if len(sys.argv) > 1 else "/root/reference",
                   "Example Output for Batch Mode")
out = {}
for name in sorted(os.listdir(src)):
    if not name.endswith(".txt"):
        continue
    with open(os.path.join(src, name)) as f:
        lines = f.read().split("\n")
    if lines and lines[-1] == "":
        lines = lines[:-1]
    # first lines, cut to 120 characters: enough to show the number format
    out[name] = {"lines": len(lines), "fields": len(lines[0].split()),
                 "trailing_space": lines[0].endswith(" "),
                 "head": [ln[:120] for ln in lines[:2]]}
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "batch_output_format.json")
with open(dst, "w") as f:
    json.dump(out, f, indent=1)
print("wrote", dst, len(out), "files")
