"""Synthetic protein generator for the docking workloads.

Each protein is a C-alpha self-avoiding random walk (3.8 A steps) confined to a
sphere; residue types are uniform over the 20 standard amino acids; heavy atoms are
built from ideal internal-coordinate templates with uniform random chi angles.
Output is 80-column PDB text whose atom-type keys (columns 12-21 with spaces
collapsed, e.g. ``"CA ALA"``) all hit the CHARMM radius/charge and ACE tables, and
whose element field (columns 77-78) is filled so that the reference's hydrogen
filter (``src/pdb_entry.cpp:41``) reads it.

Seed rule for the named workloads: ``seed = 1000*config + 10*index + role``
(role 0 = receptor, 1 = ligand).
"""
import math

import numpy as np

RESIDUES = ["ALA", "ARG", "ASN", "ASP", "CYS", "GLN", "GLU", "GLY", "HIS", "ILE",
            "LEU", "LYS", "MET", "PHE", "PRO", "SER", "THR", "TRP", "TYR", "VAL"]

# Side-chain templates as a Z-matrix: (name, ref1, ref2, ref3, bond A, angle deg, dihedral)
# A dihedral given as a string "chiK" is drawn uniformly per residue; "chiK+d" adds d deg.
_SC = {
    "ALA": [],
    "GLY": [],
    "SER": [("OG", "CB", "CA", "N", 1.42, 111.0, "chi1")],
    "CYS": [("SG", "CB", "CA", "N", 1.81, 114.0, "chi1")],
    "VAL": [("CG1", "CB", "CA", "N", 1.53, 110.5, "chi1"),
            ("CG2", "CB", "CA", "N", 1.53, 110.5, "chi1+120")],
    "THR": [("OG1", "CB", "CA", "N", 1.43, 109.5, "chi1"),
            ("CG2", "CB", "CA", "N", 1.53, 110.5, "chi1-120")],
    "ILE": [("CG1", "CB", "CA", "N", 1.53, 110.4, "chi1"),
            ("CG2", "CB", "CA", "N", 1.53, 110.5, "chi1-120"),
            ("CD1", "CG1", "CB", "CA", 1.52, 113.8, "chi2")],
    "LEU": [("CG", "CB", "CA", "N", 1.53, 116.3, "chi1"),
            ("CD1", "CG", "CB", "CA", 1.52, 110.7, "chi2"),
            ("CD2", "CG", "CB", "CA", 1.52, 110.7, "chi2+120")],
    "MET": [("CG", "CB", "CA", "N", 1.52, 114.1, "chi1"),
            ("SD", "CG", "CB", "CA", 1.81, 112.7, "chi2"),
            ("CE", "SD", "CG", "CB", 1.79, 100.8, "chi3")],
    "PRO": [("CG", "CB", "CA", "N", 1.50, 104.5, 30.0),
            ("CD", "CG", "CB", "CA", 1.51, 105.5, -35.0)],
    "PHE": [("CG", "CB", "CA", "N", 1.50, 114.0, "chi1"),
            ("CD1", "CG", "CB", "CA", 1.39, 120.7, "chi2"),
            ("CD2", "CG", "CB", "CA", 1.39, 120.7, "chi2+180"),
            ("CE1", "CD1", "CG", "CB", 1.39, 120.7, 180.0),
            ("CE2", "CD2", "CG", "CB", 1.39, 120.7, 180.0),
            ("CZ", "CE1", "CD1", "CG", 1.39, 120.0, 0.0)],
    "TYR": [("CG", "CB", "CA", "N", 1.51, 113.8, "chi1"),
            ("CD1", "CG", "CB", "CA", 1.39, 121.0, "chi2"),
            ("CD2", "CG", "CB", "CA", 1.39, 121.0, "chi2+180"),
            ("CE1", "CD1", "CG", "CB", 1.39, 121.2, 180.0),
            ("CE2", "CD2", "CG", "CB", 1.39, 121.2, 180.0),
            ("CZ", "CE1", "CD1", "CG", 1.38, 119.6, 0.0),
            ("OH", "CZ", "CE1", "CD1", 1.38, 119.9, 180.0)],
    "TRP": [("CG", "CB", "CA", "N", 1.50, 113.6, "chi1"),
            ("CD1", "CG", "CB", "CA", 1.37, 127.1, "chi2"),
            ("CD2", "CG", "CB", "CA", 1.43, 126.6, "chi2+180"),
            ("NE1", "CD1", "CG", "CB", 1.38, 110.2, 180.0),
            ("CE2", "CD2", "CG", "CB", 1.41, 107.2, 180.0),
            ("CE3", "CD2", "CG", "CB", 1.40, 133.9, 0.0),
            ("CZ2", "CE2", "CD2", "CG", 1.40, 122.4, 180.0),
            ("CZ3", "CE3", "CD2", "CG", 1.39, 118.7, 180.0),
            ("CH2", "CZ2", "CE2", "CD2", 1.37, 117.5, 0.0)],
    "HIS": [("CG", "CB", "CA", "N", 1.50, 113.7, "chi1"),
            ("ND1", "CG", "CB", "CA", 1.38, 122.7, "chi2"),
            ("CD2", "CG", "CB", "CA", 1.36, 131.0, "chi2+180"),
            ("CE1", "ND1", "CG", "CB", 1.32, 109.0, 180.0),
            ("NE2", "CD2", "CG", "CB", 1.37, 107.0, 180.0)],
    "ASP": [("CG", "CB", "CA", "N", 1.52, 113.0, "chi1"),
            ("OD1", "CG", "CB", "CA", 1.25, 119.2, "chi2"),
            ("OD2", "CG", "CB", "CA", 1.25, 118.2, "chi2+180")],
    "ASN": [("CG", "CB", "CA", "N", 1.52, 112.6, "chi1"),
            ("OD1", "CG", "CB", "CA", 1.23, 120.8, "chi2"),
            ("ND2", "CG", "CB", "CA", 1.33, 116.4, "chi2+180")],
    "GLU": [("CG", "CB", "CA", "N", 1.52, 114.1, "chi1"),
            ("CD", "CG", "CB", "CA", 1.52, 113.0, "chi2"),
            ("OE1", "CD", "CG", "CB", 1.25, 119.0, "chi3"),
            ("OE2", "CD", "CG", "CB", 1.25, 118.0, "chi3+180")],
    "GLN": [("CG", "CB", "CA", "N", 1.52, 114.0, "chi1"),
            ("CD", "CG", "CB", "CA", 1.52, 112.6, "chi2"),
            ("OE1", "CD", "CG", "CB", 1.23, 120.9, "chi3"),
            ("NE2", "CD", "CG", "CB", 1.33, 116.5, "chi3+180")],
    "LYS": [("CG", "CB", "CA", "N", 1.52, 114.1, "chi1"),
            ("CD", "CG", "CB", "CA", 1.52, 111.5, "chi2"),
            ("CE", "CD", "CG", "CB", 1.52, 111.7, "chi3"),
            ("NZ", "CE", "CD", "CG", 1.49, 111.9, "chi4")],
    "ARG": [("CG", "CB", "CA", "N", 1.52, 114.1, "chi1"),
            ("CD", "CG", "CB", "CA", 1.52, 111.5, "chi2"),
            ("NE", "CD", "CG", "CB", 1.46, 112.0, "chi3"),
            ("CZ", "NE", "CD", "CG", 1.33, 124.2, "chi4"),
            ("NH1", "CZ", "NE", "CD", 1.33, 120.0, 0.0),
            ("NH2", "CZ", "NE", "CD", 1.33, 120.0, 180.0)],
}


def _place(a, b, c, bond, angle_deg, dih_deg):
    """Atom d with |cd| = bond, angle(b,c,d) = angle, dihedral(a,b,c,d) = dih (NeRF)."""
    ang = math.radians(angle_deg)
    dih = math.radians(dih_deg)
    bc = c - b
    bc /= np.linalg.norm(bc)
    n = np.cross(b - a, bc)
    nn = np.linalg.norm(n)
    if nn < 1e-8:  # degenerate: pick any perpendicular
        n = np.cross(bc, np.array([1.0, 0.0, 0.0]))
        if np.linalg.norm(n) < 1e-8:
            n = np.cross(bc, np.array([0.0, 1.0, 0.0]))
        nn = np.linalg.norm(n)
    n /= nn
    m = np.cross(n, bc)
    d2 = np.array([-bond * math.cos(ang), bond * math.sin(ang) * math.cos(dih),
                   bond * math.sin(ang) * math.sin(dih)])
    return c + d2[0] * bc + d2[1] * m + d2[2] * n


def ca_walk(n_res, radius, rng, step=3.8, min_dist=4.0, max_tries=200):
    """Self-avoiding C-alpha walk of ``n_res`` points inside a sphere of ``radius``."""
    pts = [np.zeros(3)]
    cell = {}

    def key(p):
        return tuple(np.floor(p / min_dist).astype(int))

    def ok(p, upto):
        k = key(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in cell.get((k[0] + dx, k[1] + dy, k[2] + dz), ()):
                        if j < upto and np.sum((pts[j] - p) ** 2) < min_dist * min_dist:
                            return False
        return True

    cell.setdefault(key(pts[0]), []).append(0)
    while len(pts) < n_res:
        cur = pts[-1]
        placed = False
        for _ in range(max_tries):
            v = rng.normal(size=3)
            v /= np.linalg.norm(v)
            if len(pts) >= 2:  # keep a protein-like virtual bond angle (> 80 deg)
                prev = pts[-2] - cur
                if np.dot(v, prev / np.linalg.norm(prev)) > 0.17:
                    continue
            p = cur + step * v
            if np.dot(p, p) > radius * radius:
                continue
            if not ok(p, len(pts) - 1):
                continue
            pts.append(p)
            cell.setdefault(key(p), []).append(len(pts) - 1)
            placed = True
            break
        if not placed:  # backtrack a few residues
            nb = min(len(pts) - 1, 5)
            for _ in range(nb):
                q = pts.pop()
                cell[key(q)].remove(len(pts))
    return np.array(pts)


def build_protein(n_res, seed, radius=None):
    """Return a list of (atom_name, res_name, res_seq, element, xyz) tuples."""
    rng = np.random.default_rng(seed)
    if radius is None:
        # ~160 A^3 per residue for the C-alpha cloud
        radius = (3.0 * n_res * 160.0 / (4.0 * math.pi)) ** (1.0 / 3.0)
    ca = ca_walk(n_res, radius, rng)
    ca = ca + rng.uniform(-20.0, 20.0, size=3)  # arbitrary absolute frame
    types = rng.integers(0, 20, size=n_res)
    atoms = []
    grid = {}  # 3 A cells -> [(res_index, xyz)] for side-chain clash checks

    def gkey(p):
        return (int(math.floor(p[0] / 3.0)), int(math.floor(p[1] / 3.0)),
                int(math.floor(p[2] / 3.0)))

    def clashes(p, i):
        k = gkey(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for (j, q) in grid.get((k[0] + dx, k[1] + dy, k[2] + dz), ()):
                        if j != i and np.sum((p - q) ** 2) < 2.6 * 2.6:
                            return True
        return False

    # per-peptide frame: u along CA(i)->CA(i+1), w perpendicular (the side C and O sit on)
    pep_u = []
    pep_w = []
    for i in range(n_res - 1):
        u = ca[i + 1] - ca[i]
        u /= np.linalg.norm(u)
        ref = pep_u[-1] if pep_u else np.array([0.3, 0.5, 0.8])
        w = np.cross(u, ref)
        if np.linalg.norm(w) < 1e-6:
            w = np.cross(u, np.array([0.8, -0.5, 0.3]))
        w /= np.linalg.norm(w)
        if i % 2:
            w = -w
        pep_u.append(u)
        pep_w.append(w)
    # backbone of every residue first, so side chains are checked against all of it
    bb = []
    for i in range(n_res):
        pos = {"CA": ca[i]}
        # peptide i-1 -> i supplies N(i); peptide i -> i+1 supplies C(i) and O(i)
        u, w = (pep_u[i - 1], pep_w[i - 1]) if i > 0 else (-pep_u[0], pep_w[0])
        pos["N"] = ca[i] + 1.46 * (-0.819 * u + 0.574 * w)
        u, w = (pep_u[i], pep_w[i]) if i + 1 < n_res else (-pep_u[i - 1], pep_w[i - 1])
        pos["C"] = ca[i] + 1.52 * (0.819 * u + 0.574 * w)
        o_dir = w - 0.35 * u
        pos["O"] = pos["C"] + 1.23 * o_dir / np.linalg.norm(o_dir)
        for nm in ("N", "CA", "C", "O"):
            grid.setdefault(gkey(pos[nm]), []).append((i, pos[nm]))
        bb.append(pos)
    for i in range(n_res):
        res = RESIDUES[types[i]]
        pos = bb[i]
        order = ["N", "CA", "C", "O"]
        if res != "GLY":
            pos["CB"] = _place(pos["C"], pos["N"], ca[i], 1.53, 110.5, 122.5)
            order.append("CB")
            for _attempt in range(30):  # resample chi angles while the side chain clashes
                chi = rng.uniform(-180.0, 180.0, size=4)
                sc = {}
                for (nm, r1, r2, r3, b, a, d) in _SC[res]:
                    if isinstance(d, str):
                        k = int(d[3])
                        off = float(d[4:]) if len(d) > 4 else 0.0
                        d = chi[k - 1] + off
                    get = lambda x: sc[x] if x in sc else pos[x]
                    sc[nm] = _place(get(r3), get(r2), get(r1), b, a, d)
                if not any(clashes(p, i) for p in sc.values()):
                    break
            pos.update(sc)
            order.extend(nm for (nm, *_r) in _SC[res])
            for nm in order[4:]:
                grid.setdefault(gkey(pos[nm]), []).append((i, pos[nm]))
        for nm in order:
            atoms.append((nm, res, i + 1, nm[0], pos[nm]))
    return atoms


def pdb_text(atoms, chain="A"):
    lines = []
    for serial, (nm, res, seq, el, xyz) in enumerate(atoms, start=1):
        name = (" " + nm).ljust(4) if len(nm) < 4 else nm
        lines.append("ATOM  %5d %-4s %3s %c%4d    %8.3f%8.3f%8.3f%6.2f%6.2f          %2s  "
                     % (serial, name, res, chain, seq, xyz[0], xyz[1], xyz[2], 1.0, 0.0, el))
    lines.append("END")
    return "\n".join(lines) + "\n"


def write_pdb(path, n_res, seed, radius=None):
    atoms = build_protein(n_res, seed, radius)
    with open(path, "w") as f:
        f.write(pdb_text(atoms))
    return len(atoms)


def workload_seed(config, index, role):
    return 1000 * config + 10 * index + role
