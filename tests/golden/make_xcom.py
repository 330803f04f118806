"""Extract the numeric rows of the NIST XCOM tables the reference repository holds
(data/xcom/photon-Z7.dat, photon-Z8.dat; columns: photon energy (MeV), coherent,
incoherent, photoelectric, pair (nuclear field), pair (electron field), totals; barn/atom)
into tests/golden/xcom-Z{7,8}.txt.  The reference's code never reads them; the tests use
them as independent data for the photon cross sections."""
import os
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/data/xcom"
OUT = os.path.dirname(os.path.abspath(__file__))
for Z in (7, 8):
    rows = []
    with open(os.path.join(SRC, "photon-Z%d.dat" % Z)) as f:
        for line in f:
            parts = line.split()
            try:
                vals = [float(v) for v in parts]
            except ValueError:
                continue
            if len(vals) == 8:
                rows.append(vals)
    with open(os.path.join(OUT, "xcom-Z%d.txt" % Z), "w") as f:
        f.write("# NIST XCOM, Z=%d: E(MeV) coherent incoherent photoelectric pair_nuc pair_el "
                "total_w total_wo (barn/atom)\n" % Z)
        for r in rows:
            f.write(" ".join("%.4e" % v for v in r) + "\n")
    print(Z, len(rows))
