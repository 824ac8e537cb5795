"""Writes tests/golden/hard_disk_dipoles_reference_cdf_px.dat and _py.dat: the reference's
Newtonian-ECMC data for the polarization of 81 hard-disk dipoles (output/hard_disk_dipoles/
ReferenceDataPx_81Dipoles_NewtonianECMC.dat and ...Py...), their numbers unchanged, with the
files' notice kept.  Run once in a container that holds the reference tree; nothing imports the
reference."""
import os

SRC = "/root/reference/jellyfysh/output/hard_disk_dipoles/ReferenceData%s_81Dipoles_NewtonianECMC.dat"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hard_disk_dipoles_reference_cdf_%s.dat")


def main():
    for part in ("Px", "Py"):
        lines = open(SRC % part).read().splitlines()
        notice = [l for l in lines if l.startswith("#")]
        data = [l for l in lines if l.strip() and not l.startswith("#")]
        head = ["# Cumulative distribution of the %s component of the polarization of 81 hard-disk dipoles" % part[1],
                "# (config_files/hard_disk_dipoles/hard_disk_dipoles_cells.ini), from Newtonian event-chain Monte",
                "# Carlo: column 1 the polarization component, column 2 the probability of a value at or below it.",
                "# Numbers of the reference's output/hard_disk_dipoles/ReferenceData%s_81Dipoles_NewtonianECMC.dat," % part,
                "# unchanged (tests/golden/copy_hard_disk_dipoles_cdf.py); that file's notice follows."]
        dst = DST % part.lower()
        with open(dst, "w") as f:
            f.write("\n".join(head + notice + data) + "\n")
        print("%d rows -> %s" % (len(data), dst))


if __name__ == "__main__":
    main()
