"""Extracts the DIA interaction tables that the reference printed for its
regression case (12 directions x 25 frequencies, FR1 = 0.04177248 Hz) from
dayfiles/wam_prot_coarse_ST6:878-937 into dia_12x25.json, together with the
run's grid size and maximum CFL number (:872-874, preproc_prot_coarse_ST6:102).
Run in a container that holds the reference: python tests/golden/make_dia_golden.py
"""
import json
import os
import re

REF = "/root/reference/dayfiles/wam_prot_coarse_ST6"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dia_12x25.json")


def main():
    lines = open(REF, encoding="latin-1").read().splitlines()
    i = next(n for n, l in enumerate(lines) if "ACL1       ACL2" in l)
    acl = [float(x) for x in lines[i + 1].split()]
    rows = []
    j = i + 4
    while lines[j].strip():
        p = lines[j].split()
        rows.append([int(p[0]), int(p[1]), int(p[2]), int(p[3]), int(p[4])] + [float(x) for x in p[5:10]])
        j += 1
    k = next(n for n, l in enumerate(lines) if "K   K1W   K2W  K11W  K21W" in l)
    ang = []
    j = k + 1
    while lines[j].strip():
        ang.append([int(x) for x in lines[j].split()])
        j += 1
    f = next(n for n, l in enumerate(lines) if "TAIL ARRAY FRH" in l)
    frh = [float(x) for x in lines[f + 1].split()]
    cfl = float(re.search(r"total_cfl =\s*([0-9.Ee+-]+)", "\n".join(lines)).group(1))
    nadv = int(re.search(r"NADV =\s*([0-9]+)", "\n".join(lines)).group(1))
    doc = dict(source="dayfiles/wam_prot_coarse_ST6:872-937", kl=12, ml=25, fr1=0.04177248,
               acl1=acl[0], acl2=acl[1], cl11=acl[2], cl21=acl[3], dal1=acl[4], dal2=acl[5],
               freq_rows=rows, angular_rows=ang, frh=frh, max_cfl=cfl, nadv=nadv, nsea=1599)
    with open(OUT, "w") as fh:
        json.dump(doc, fh, indent=1)
    print("wrote", OUT, len(rows), "frequency rows,", len(ang), "angular rows")


if __name__ == "__main__":
    main()
