"""Writes tests/golden/coulomb_atoms_reference_cdf.dat: the reference's reversible-Monte-Carlo
data for two +1 point charges (output/2018_JCP_149_064113/coulomb_atoms/
ReferenceDataCoulombAtoms.dat), its numbers unchanged, with the file's notice kept.  Run once
in a container that holds the reference tree; nothing imports the reference."""
import os

SRC = "/root/reference/jellyfysh/output/2018_JCP_149_064113/coulomb_atoms/ReferenceDataCoulombAtoms.dat"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "coulomb_atoms_reference_cdf.dat")


def main():
    lines = open(SRC).read().splitlines()
    notice = [l for l in lines if l.startswith("#")]
    data = [l for l in lines if l.strip() and not l.startswith("#")]
    head = ["# Cumulative distribution of the separation of two +1 point charges (merged-image Coulomb, unit",
            "# periodic cube, beta = 2) from reversible Monte Carlo: column 1 separation, column 2 the",
            "# probability of a separation at or below it.  Numbers of the reference's",
            "# output/2018_JCP_149_064113/coulomb_atoms/ReferenceDataCoulombAtoms.dat, unchanged",
            "# (tests/golden/copy_coulomb_atoms_cdf.py); that file's notice follows."]
    with open(DST, "w") as f:
        f.write("\n".join(head + notice + data) + "\n")
    print("%d rows -> %s" % (len(data), DST))


if __name__ == "__main__":
    main()
