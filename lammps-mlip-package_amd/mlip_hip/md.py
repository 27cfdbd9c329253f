"""LAMMPS-style output of device-resident NVE runs (the request pool's
per-step rows: energy, virial[6], kinetic-energy tensor[6]).

thermo() forms the quantities `thermo_style custom step temp epair emol
etotal press vol pxx pyy pzz pxy pxz pyz` prints in metal units (compute
temp with 3N - 3 degrees of freedom, compute pressure from the kinetic
tensor plus the virial), as the reference's regression runs print them
(test/regression/gtinv-197-stress.in, log.lammps.*); write_dump() writes the
`write_dump all custom dump.atom id type x y z` file of those runs
(test/regression/gtinv-197.in)."""
import numpy as np

MVV2E = 1.0364269e-4        # g/mol A^2/ps^2 -> eV (LAMMPS metal units)
FTM2V = 1.0 / 1.0364269e-4  # eV/A / (g/mol) -> A/ps^2
BOLTZ = 8.617343e-5         # eV/K
NKTV2P = 1.6021765e6        # eV/A^3 -> bar

COLUMNS = ("Step", "Temp", "E_pair", "E_mol", "TotEng", "Press", "Volume", "Pxx", "Pyy", "Pzz", "Pxy", "Pxz", "Pyz")


def thermo(natoms, volume, row):
    """Thermo quantities of one pool row (energy, virial xx yy zz xy xz yz,
    kinetic tensor m v_a v_b in the same order, all eV)."""
    row = np.asarray(row, dtype=np.float64)
    energy, virial, ket = row[0], row[1:7], row[7:13]
    ke = 0.5 * (ket[0] + ket[1] + ket[2])
    dof = 3 * natoms - 3
    p = (ket + virial) / volume * NKTV2P
    return {"Temp": 2.0 * ke / (dof * BOLTZ) if dof > 0 else 0.0, "E_pair": float(energy), "E_mol": 0.0,
            "TotEng": float(energy + ke), "Press": float((p[0] + p[1] + p[2]) / 3.0), "Volume": float(volume),
            "Pxx": float(p[0]), "Pyy": float(p[1]), "Pzz": float(p[2]), "Pxy": float(p[3]), "Pxz": float(p[4]),
            "Pyz": float(p[5])}


def thermo_log(rows, structure, every=1, columns=COLUMNS):
    """The thermo block of a LAMMPS log for a pool result's rows (one per
    step from step 0), every `every` steps and the last."""
    lines = [" ".join("%12s" % c if c != "Step" else "%8s" % c for c in columns)]
    n = len(rows)
    for step in range(n):
        if step % every and step != n - 1:
            continue
        t = thermo(structure.natoms, structure.volume, rows[step])
        t["Step"] = step
        lines.append(" ".join("%8d" % step if c == "Step" else "%12.8g" % t[c] for c in columns))
    return "\n".join(lines) + "\n"


def write_dump(path, step, structure, positions, line_format="%d %d %10.15g %10.15g %10.15g"):
    """LAMMPS text dump (write_dump ... custom id type x y z) of unwrapped
    positions, atom ids 1..N in structure order."""
    x = np.asarray(positions, dtype=np.float64)
    c = structure.cell
    o = structure.origin
    tri = abs(c[1, 0]) + abs(c[2, 0]) + abs(c[2, 1]) > 0.0
    with open(path, "w") as fh:
        fh.write("ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\n" % (step, x.shape[0]))
        if tri:
            # bounding box of the triclinic cell, then the tilt factors
            xy, xz, yz = c[1, 0], c[2, 0], c[2, 1]
            xlo, xhi, ylo, yhi = o[0], o[0] + c[0, 0], o[1], o[1] + c[1, 1]
            xlob = xlo + min(0.0, xy, xz, xy + xz)
            xhib = xhi + max(0.0, xy, xz, xy + xz)
            ylob, yhib = ylo + min(0.0, yz), yhi + max(0.0, yz)
            fh.write("ITEM: BOX BOUNDS xy xz yz pp pp pp\n")
            fh.write("%-1.16e %-1.16e %-1.16e\n" % (xlob, xhib, xy))
            fh.write("%-1.16e %-1.16e %-1.16e\n" % (ylob, yhib, xz))
            fh.write("%-1.16e %-1.16e %-1.16e\n" % (o[2], o[2] + c[2, 2], yz))
        else:
            fh.write("ITEM: BOX BOUNDS pp pp pp\n")
            for k in range(3):
                fh.write("%-1.16e %-1.16e\n" % (o[k], o[k] + c[k, k]))
        fh.write("ITEM: ATOMS id type x y z \n")
        for i in range(x.shape[0]):
            fh.write(line_format % (i + 1, structure.types[i], x[i, 0], x[i, 1], x[i, 2]) + " \n")
