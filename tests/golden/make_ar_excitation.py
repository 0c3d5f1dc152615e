"""Write tests/golden/ar_excitation_tables.json: the six tabulated e-Ar
electronic-excitation cross sections of the reference's
`References/PhelpsAr_e_cross_section.pdf` (its section 1, tables (2)-(7):
threshold, then (energy eV, cross section 1e-20 m^2) rows, data of Chutjian &
Cartwright, Phys. Rev. A 23 (1981) 2178), plus the elastic and ionisation
tables (1) and (8) for reference.

The PDF's text operators are read as data (zlib-inflated content streams, the
strings of Tj/TJ operators); nothing in the reference is executed.  Run from
the repository root in a tree where /root/reference exists:

    python tests/golden/make_ar_excitation.py
"""
import json
import os
import re
import zlib

SRC = "/root/reference/References/PhelpsAr_e_cross_section.pdf"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ar_excitation_tables.json")


def page_text(data):
    chunks = []
    for m in re.finditer(rb"stream\r?\n", data):
        start = m.end()
        end = data.find(b"endstream", start)
        try:
            d = zlib.decompressobj().decompress(data[start:end])
        except zlib.error:
            continue
        if b"BT" not in d:
            continue
        for t in re.finditer(rb"\[(.*?)\]\s*TJ|\((.*?)\)\s*Tj|(T\*|Td|TD|Tm)", d, re.S):
            if t.group(1) is not None:
                chunks.append(b"".join(re.findall(rb"\((.*?)(?<!\\)\)", t.group(1))))
            elif t.group(2) is not None:
                chunks.append(t.group(2))
            else:
                chunks.append(b"\n")
    return b"".join(chunks).decode("latin1")


def tables(text):
    # a table starts at "Threshold = X eV" (or at the elastic header) and runs
    # over the following "energy value" rows
    heads = [(m.start(), float(m.group(1)))
             for m in re.finditer(r"Threshold = ([0-9.]+) eV", text)]
    out = []
    for k, (pos, thr) in enumerate(heads):
        end = heads[k + 1][0] if k + 1 < len(heads) else len(text)
        body = text[pos:end]
        # stop at the reference list / next section header
        cut = re.search(r"\\\(\d\\\) ", body[20:])
        if cut:
            body = body[:20 + cut.start()]
        rows = []
        for a, b in re.findall(r"(?m)^\s*([0-9]+\.[0-9]+)\s+([0-9]+\.[0-9]+)\s*$", body):
            if rows and float(a) <= rows[-1][0]:
                break  # past the table: the swarm-data section that follows
            rows.append((float(a), float(b)))
        out.append({"threshold_eV": thr, "rows": rows})
    return out


def main():
    text = page_text(open(SRC, "rb").read())
    t = tables(text)
    # tables in document order: 6 excitation levels, then ionisation
    exc = [x for x in t[:6]]
    for x in exc:
        assert x["rows"][0][0] == x["threshold_eV"] and x["rows"][0][1] == 0.0, x
        es = [r[0] for r in x["rows"]]
        assert es == sorted(es), x
    iz = t[6]
    assert iz["threshold_eV"] == 15.755
    doc = {
        "source": "References/PhelpsAr_e_cross_section.pdf, section 1 (recommended electron "
                  "collision cross sections for Ar), tables (2)-(7) electronic excitation "
                  "[Chutjian & Cartwright, Phys. Rev. A 23 (1981) 2178] and (8) ionisation "
                  "[Rapp & Englander-Golden, J. Chem. Phys. 43 (1965) 1464]",
        "unit": "energy eV, cross section 1e-20 m^2 (angstrom^2)",
        "excitation": exc,
        "ionization": iz,
    }
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
    print("wrote", OUT, [len(x["rows"]) for x in exc], len(iz["rows"]))


if __name__ == "__main__":
    main()
