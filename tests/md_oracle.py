"""Velocity-Verlet NVE driven by the CPU oracle -- test infrastructure only.

LAMMPS fix nve in metal units (FixNVE::initial_integrate / final_integrate:
half kick and drift before the force evaluation, half kick after), thermo
quantities as LAMMPS computes them (compute temp with 3N - 3 degrees of
freedom, compute pressure with the kinetic tensor).  It pins the device
integrator of the request pool and its units against the regression logs
and dumps of the reference (test/regression/*.in, log.lammps.*)."""
import numpy as np

from mlip_hip import workloads as W

MVV2E = 1.0364269e-4       # g/mol A^2/ps^2 -> eV
FTM2V = 1.0 / 1.0364269e-4  # eV/A / (g/mol) -> A/ps^2
BOLTZ = 8.617343e-5        # eV/K
NKTV2P = 1.6021765e6       # eV/A^3 -> bar


def thermo(natoms, volume, energy, virial, ke_tensor):
    """Temp, E_pair, TotEng, Press and the six pressure components from one
    evaluation (virial and kinetic tensor in xx, yy, zz, xy, xz, yz order)."""
    ke = 0.5 * (ke_tensor[0] + ke_tensor[1] + ke_tensor[2])
    dof = 3 * natoms - 3
    p = (np.asarray(ke_tensor) + np.asarray(virial)) / volume * NKTV2P
    return {"Temp": 2.0 * ke / (dof * BOLTZ) if dof > 0 else 0.0, "E_pair": energy, "TotEng": energy + ke,
            "Press": (p[0] + p[1] + p[2]) / 3.0, "Pxx": p[0], "Pyy": p[1], "Pzz": p[2], "Pxy": p[3],
            "Pxz": p[4], "Pyz": p[5]}


def ke_tensor(mass, v):
    return MVV2E * np.array([np.sum(mass * v[:, 0] * v[:, 0]), np.sum(mass * v[:, 1] * v[:, 1]),
                             np.sum(mass * v[:, 2] * v[:, 2]), np.sum(mass * v[:, 0] * v[:, 1]),
                             np.sum(mass * v[:, 0] * v[:, 2]), np.sum(mass * v[:, 1] * v[:, 2])])


def run_nve(oracle, s, mass_of_type, dt, nsteps, record=(), v0=None):
    """Returns final positions, velocities and {step: thermo} for the steps
    in `record` (step 0 included when listed)."""
    x = s.positions.copy()
    v = np.zeros_like(x) if v0 is None else np.array(v0, dtype=np.float64)
    m = np.asarray(mass_of_type, dtype=np.float64)[s.types - 1]

    def evaluate(pos):
        cur = W.Structure(s.cell, pos, s.types, origin=s.origin)
        return oracle.compute(W.build_neighbor_list(cur, oracle.cutoff), s.types - 1)

    r = evaluate(x)
    out = {}
    if 0 in record:
        out[0] = thermo(s.natoms, s.volume, r["energy"], r["virial"], ke_tensor(m, v))
    dtf = 0.5 * dt * FTM2V
    for step in range(1, nsteps + 1):
        v += dtf * r["forces"] / m[:, None]
        x += dt * v
        r = evaluate(x)
        v += dtf * r["forces"] / m[:, None]
        if step in record:
            out[step] = thermo(s.natoms, s.volume, r["energy"], r["virial"], ke_tensor(m, v))
    return x, v, out


def read_dump(path):
    """id-sorted positions of a LAMMPS `write_dump ... id type x y z` file."""
    lines = open(path).read().splitlines()
    k = next(i for i, ln in enumerate(lines) if ln.startswith("ITEM: ATOMS"))
    rows = [ln.split() for ln in lines[k + 1:] if ln.strip()]
    rows.sort(key=lambda t: int(t[0]))
    return np.array([[float(t[2]), float(t[3]), float(t[4])] for t in rows])


def periodic_diff(a, b, cell):
    """a - b reduced to the nearest periodic image."""
    d = np.asarray(a) - np.asarray(b)
    frac = d @ np.linalg.inv(cell)
    frac -= np.round(frac)
    return frac @ cell
