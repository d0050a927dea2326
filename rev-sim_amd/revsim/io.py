"""File formats of the reference's water workflow.

- PDB water boxes (the reference builds its system from water/tip3p.pdb through Molly,
  water/water.jl:980-991): CRYST1 box and ATOM records, residues HOH with atoms O, H1, H2, in Å.
- OpenMM force-field XML in the layout of water/trained_force_fields/*.xml, which the reference's
  validation script loads (water/validation_sims.py:16): charges, harmonic bond and angle, O-only
  Lennard-Jones.
- The experimental RDF table (water/soper_2013_water_rdf.txt) read as read_rdf_data does
  (water/water.jl:616-629), and its O-O curve resampled onto the library's RDF bins.

Units here are the library's: nm, ps, kJ/mol, e.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np

from .system import KB, MASS_H, MASS_O, THETA0, WaterBox

ANGSTROM = 0.1  # nm


# ------------------------------------------------------------------------------------------ PDB

def read_pdb(path: str, temperature: float | None = None, seed: int = 12) -> WaterBox:
    """Water box from a PDB file: box edges from CRYST1, atoms from ATOM/HETATM records in file
    order, which must be O, H1, H2 per HOH residue (the order the library and the reference use,
    water/water.jl:550-556). Velocities are Maxwell-Boltzmann at `temperature` from a seeded
    generator (the role of random_velocity, water/water.jl:1593), or zero if temperature is None.
    Coordinates are wrapped into the box."""
    box = None
    names, xyz = [], []
    with open(path) as f:
        for line in f:
            rec = line[:6].strip()
            if rec == "CRYST1":
                a, b, c = (float(line[6 + 9 * k: 15 + 9 * k]) for k in range(3))
                angles = [float(line[33 + 7 * k: 40 + 7 * k]) for k in range(3)]
                if any(abs(t - 90.0) > 1e-6 for t in angles):
                    raise ValueError(f"{path}: only orthorhombic boxes are supported, CRYST1 angles {angles}")
                box = np.array([a, b, c]) * ANGSTROM
            elif rec in ("ATOM", "HETATM"):
                names.append(line[12:16].strip())
                xyz.append([float(line[30:38]), float(line[38:46]), float(line[46:54])])
    if box is None:
        raise ValueError(f"{path}: no CRYST1 record")
    if len(names) % 3:
        raise ValueError(f"{path}: {len(names)} atoms is not a whole number of 3-site waters")
    for m in range(len(names) // 3):
        got = names[3 * m: 3 * m + 3]
        if got[0] not in ("O", "OW") or got[1] not in ("H1", "HW1") or got[2] not in ("H2", "HW2"):
            raise ValueError(f"{path}: residue {m + 1} has atoms {got}, expected O, H1, H2")
    x = np.asarray(xyz, dtype=np.float64) * ANGSTROM
    x = np.mod(x, box)
    coords = np.ascontiguousarray(x.reshape(-1))
    n = len(names) // 3
    if temperature is None:
        v = np.zeros_like(coords)
    else:
        m = np.tile([MASS_O, MASS_H, MASS_H], n)
        rng = np.random.Generator(np.random.Philox(seed))
        v = (rng.standard_normal((3 * n, 3)) * np.sqrt(KB * temperature / m)[:, None]).reshape(-1)
    return WaterBox(coords, np.ascontiguousarray(v), box.astype(np.float64))


def write_pdb(path: str, wb: WaterBox) -> None:
    """PDB with CRYST1 and one HOH residue per molecule (O, H1, H2), coordinates in Å with the
    format's three decimals, wrapped into the box."""
    box = np.asarray(wb.box, dtype=np.float64)
    x = np.mod(np.asarray(wb.coords).reshape(-1, 3), box) / ANGSTROM
    lines = [f"CRYST1{box[0] / ANGSTROM:9.3f}{box[1] / ANGSTROM:9.3f}{box[2] / ANGSTROM:9.3f}"
             f"{90.0:7.2f}{90.0:7.2f}{90.0:7.2f} P 1           1 "]
    for a in range(x.shape[0]):
        res = a // 3 + 1
        name = ("O", "H1", "H2")[a % 3]
        lines.append(f"ATOM  {(a + 1) % 100000:5d}  {name:<3s} HOH A{res % 10000:4d}    "
                     f"{x[a, 0]:8.3f}{x[a, 1]:8.3f}{x[a, 2]:8.3f}  1.00  0.00")
    lines.append(f"TER   {(x.shape[0] + 1) % 100000:5d}      HOH A{(x.shape[0] // 3) % 10000:4d}")
    lines.append("END")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


# ----------------------------------------------------------------------------------- force field

def write_forcefield_xml(path: str, theta, theta0: float = THETA0, mass_O: float = MASS_O,
                         mass_H: float = MASS_H) -> None:
    """OpenMM force-field XML for θ = (σ_O, ε_O, q_H, k_bond, r0_bond, k_angle) and the angle rest
    value θ0, laid out as the reference's trained force fields (q_O = -2 q_H; H atoms carry no
    Lennard-Jones: σ = 1, ε = 0)."""
    sig, eps, qH, kb, r0, ka = (float(v) for v in theta)
    qO = -2.0 * qH
    ff = ET.Element("ForceField")
    at = ET.SubElement(ff, "AtomTypes")
    ET.SubElement(at, "Type", name="tip3p-O", **{"class": "tip3p-O"}, element="O", mass=repr(mass_O))
    ET.SubElement(at, "Type", name="tip3p-H", **{"class": "tip3p-H"}, element="H", mass=repr(mass_H))
    rs = ET.SubElement(ET.SubElement(ff, "Residues"), "Residue", name="HOH")
    ET.SubElement(rs, "Atom", name="O", type="tip3p-O", charge=repr(qO))
    ET.SubElement(rs, "Atom", name="H1", type="tip3p-H", charge=repr(qH))
    ET.SubElement(rs, "Atom", name="H2", type="tip3p-H", charge=repr(qH))
    ET.SubElement(rs, "Bond", atomName1="O", atomName2="H1")
    ET.SubElement(rs, "Bond", atomName1="O", atomName2="H2")
    hb = ET.SubElement(ff, "HarmonicBondForce")
    ET.SubElement(hb, "Bond", type1="tip3p-O", type2="tip3p-H", length=repr(r0), k=repr(kb))
    ha = ET.SubElement(ff, "HarmonicAngleForce")
    ET.SubElement(ha, "Angle", type1="tip3p-H", type2="tip3p-O", type3="tip3p-H", angle=repr(float(theta0)),
                  k=repr(ka))
    nb = ET.SubElement(ff, "NonbondedForce", coulomb14scale=repr(1.0 / 1.2), lj14scale="0.5")
    ET.SubElement(nb, "UseAttributeFromResidue", name="charge")
    ET.SubElement(nb, "Atom", type="tip3p-O", sigma=repr(sig), epsilon=repr(eps))
    ET.SubElement(nb, "Atom", type="tip3p-H", sigma="1", epsilon="0")
    ET.indent(ff, space="  ")
    ET.ElementTree(ff).write(path)


def read_forcefield_xml(path: str) -> tuple[np.ndarray, float]:
    """(θ, θ0) from a water force-field XML (the reference's trained_force_fields/*.xml layout):
    σ_O, ε_O from the O NonbondedForce atom, q_H = -q_O/2 from the residue's O charge (the
    reference's convention, water/water.jl:548-549), the O-H bond and the H-O-H angle."""
    root = ET.parse(path).getroot()
    q = {a.get("name"): float(a.get("charge")) for a in root.iter("Atom") if a.get("charge") is not None}
    lj = {a.get("type"): a for a in root.find("NonbondedForce").iter("Atom")}
    o = lj.get("tip3p-O")
    if o is None or "O" not in q:
        raise ValueError(f"{path}: no tip3p-O Lennard-Jones atom or O charge")
    bond = root.find("HarmonicBondForce").find("Bond")
    ang = root.find("HarmonicAngleForce").find("Angle")
    theta = np.array([float(o.get("sigma")), float(o.get("epsilon")), -q["O"] / 2.0,
                      float(bond.get("k")), float(bond.get("length")), float(ang.get("k"))])
    return theta, float(ang.get("angle"))


# ---------------------------------------------------------------------------------------- RDF

def read_rdf_data(path: str):
    """Columns of the experimental RDF table as read_rdf_data reads them (water/water.jl:616-629):
    the first two rows (r = 0.00, 0.01 Å) skipped; r in nm and g_OO, g_OH, g_HH (columns 1, 3, 5).
    The reference then keeps g_OO[86:334] (248 values, 0.261-1.002 nm)."""
    rows = []
    with open(path) as f:
        for line in f.readlines()[2:]:
            cols = line.split()
            if cols:
                rows.append([float(c) for c in cols])
    a = np.asarray(rows)
    return a[:, 0] * ANGSTROM, a[:, 1], a[:, 3], a[:, 5]


def rdf_on_bins(r: np.ndarray, g: np.ndarray, n_bins: int, rmax: float) -> np.ndarray:
    """A tabulated g(r) linearly interpolated at the library's bin centres (b + 1/2) rmax/n_bins,
    zero below the table's first point, for use as the RDF loss target."""
    xb = (np.arange(n_bins) + 0.5) * (rmax / n_bins)
    out = np.interp(xb, r, g, left=0.0, right=float(g[-1]))
    out[xb < r[0]] = 0.0
    return out


def maxwell_boltzmann(n_molecules: int, temperature: float, seed: int) -> np.ndarray:
    """Velocities for n waters at `temperature` (random_velocity, water/water.jl:1593)."""
    m = np.tile([MASS_O, MASS_H, MASS_H], n_molecules)
    rng = np.random.Generator(np.random.Philox(seed))
    return np.ascontiguousarray((rng.standard_normal((3 * n_molecules, 3)) *
                                 np.sqrt(KB * temperature / m)[:, None]).reshape(-1))


def oh_geometry(wb: WaterBox) -> tuple[np.ndarray, np.ndarray]:
    """O-H distances (2 per molecule) and H-O-H angles of a water box, minimum image."""
    x = np.asarray(wb.coords).reshape(-1, 3, 3)
    box = np.asarray(wb.box)
    d = x[:, 1:, :] - x[:, :1, :]
    d -= box * np.round(d / box)
    r = np.linalg.norm(d, axis=2)
    cos = np.sum(d[:, 0] * d[:, 1], axis=1) / (r[:, 0] * r[:, 1])
    return r.reshape(-1), np.arccos(np.clip(cos, -1.0, 1.0))


__all__ = ["read_pdb", "write_pdb", "write_forcefield_xml", "read_forcefield_xml", "read_rdf_data",
           "rdf_on_bins", "maxwell_boltzmann", "oh_geometry"]
