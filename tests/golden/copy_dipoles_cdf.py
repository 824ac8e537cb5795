"""Writes tests/golden/dipoles_reference_cdf_13.dat and _14.dat: the reference's
reversible-Monte-Carlo data for two dipoles (output/2018_JCP_149_064113/dipoles/
ReferenceDataDipoles_13.dat and _14.dat), their numbers unchanged, with the files' notice kept.
Run once in a container that holds the reference tree; nothing imports the reference."""
import os

SRC = "/root/reference/jellyfysh/output/2018_JCP_149_064113/dipoles/ReferenceDataDipoles_%s.dat"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dipoles_reference_cdf_%s.dat")
WHAT = {"13": "equal leaf numbers (atoms 1 and 3: the +1 charges of the two dipoles)",
        "14": "different leaf numbers (atoms 1 and 4: a +1 and a -1 charge of different dipoles)"}


def main():
    for part in ("13", "14"):
        lines = open(SRC % part).read().splitlines()
        notice = [l for l in lines if l.startswith("#")]
        data = [l for l in lines if l.strip() and not l.startswith("#")]
        head = ["# Cumulative distribution of the separation of two leaf units of different dipoles with",
                "# %s, two dipoles in the unit periodic cube at" % WHAT[part],
                "# beta = 1, from reversible Monte Carlo: column 1 separation, column 2 the probability of a",
                "# separation at or below it.  Numbers of the reference's",
                "# output/2018_JCP_149_064113/dipoles/ReferenceDataDipoles_%s.dat, unchanged" % part,
                "# (tests/golden/copy_dipoles_cdf.py); that file's notice follows."]
        with open(DST % part, "w") as f:
            f.write("\n".join(head + notice + data) + "\n")
        print("%d rows -> %s" % (len(data), DST % part))


if __name__ == "__main__":
    main()
