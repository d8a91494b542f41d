"""Synthetic systems and frames for the five workloads of the energy pipeline.

Every coordinate is rounded to an FP32-representable value, so the FP64 and
FP32 paths start from identical inputs.  Water is the 3-site rigid model
(O-H 0.9572 A, HOH 104.52 deg, q_O -0.834, q_H +0.417, O sigma 3.1507 A,
epsilon 0.1521 kcal/mol, H without LJ); waters sit on a jittered cubic lattice
(+-0.3 A) with uniform random orientations.  Solution systems keep every water
oxygen at least ``cavity`` A from the solute so that all solute-solvent
energies fall inside the soft region of the energy grid.

Workloads (names used by bench.py and the tests):

* ``soln12``  -- 12-site rigid solute + 1000 waters, 31.0 A, PME 32^3, p 4
* ``refs20``  -- 4000 waters, 49.3 A, 20-site test particle, PME 48^3, 600 bins
* ``soln300`` -- 300-site flexible chain + 20000 waters, 84.5 A, PME 96^3, p 6
* ``soln30``  -- 30-site solute + 4000 waters, 49.3 A, PME 48^3
* ``mixed``   -- 7000 waters + 800 10-site cosolvents + 100 Na+ + 100 Cl-,
  60 A, PME 64^3, test particles of 5 / 20 / 60 / 150 sites
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .system import (ATOM_MASS, CHARGE_SCALE, ErmSystem, Hist, PT_SOLUTE, PT_SOLVENT,
                     PT_TEST_RIGID, SLT_REFS_RIGID, SLT_SOLN, lj_tables, mol_charges,
                     uvspec_of)

OH = 0.9572
HOH = np.deg2rad(104.52)
WATER_Q = (-0.834, 0.417, 0.417)
WATER_SIG = (3.1507, 0.0, 0.0)
WATER_EPS = (0.1521, 0.0, 0.0)
WATER_MASS = (ATOM_MASS["O"], ATOM_MASS["H"], ATOM_MASS["H"])
SEED0 = 20261202


def f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def water_body() -> np.ndarray:
    h = HOH / 2.0
    return np.array([[0.0, 0.0, 0.0],
                     [OH * np.sin(h), 0.0, OH * np.cos(h)],
                     [-OH * np.sin(h), 0.0, OH * np.cos(h)]])


def random_rotations(rng, n) -> np.ndarray:
    """Uniform rotations (Shoemake's subgroup algorithm) as 3x3 matrices."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    q = np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], axis=1)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def rigid_blob(rng, n, radius, min_sep=1.0) -> np.ndarray:
    """n sites uniform in a ball of `radius` with pairwise separation >= min_sep."""
    pts: list = []
    while len(pts) < n:
        p = (rng.random(3) * 2 - 1) * radius
        if np.dot(p, p) > radius * radius:
            continue
        if all(np.linalg.norm(p - q) >= min_sep for q in pts):
            pts.append(p)
    x = np.array(pts)
    return x - x.mean(axis=0)


def chain(rng, n, bond=1.5, min_sep=2.0) -> np.ndarray:
    """Self-avoiding random chain with fixed bond length."""
    while True:
        x = [np.zeros(3)]
        ok = True
        for i in range(1, n):
            for _ in range(200):
                d = rng.normal(size=3)
                p = x[-1] + bond * d / np.linalg.norm(d)
                if i < 3 or np.min(np.linalg.norm(np.array(x[:-2]) - p, axis=1)) >= min_sep:
                    x.append(p)
                    break
            else:
                ok = False
                break
        if ok:
            a = np.array(x)
            return a - a.mean(axis=0)


def neutral_charges(rng, n, qmax=0.5) -> np.ndarray:
    q = rng.uniform(-0.5, 0.5, n)
    q -= q.mean()
    m = np.abs(q).max()
    if m > qmax:
        q *= qmax / m
    q -= q.mean()
    return q


@dataclasses.dataclass
class MolKind:
    name: str
    q: np.ndarray        # raw charges (e)
    sig: np.ndarray
    eps: np.ndarray
    mass: np.ndarray
    body: np.ndarray     # body-frame coordinates (n x 3)
    flexible: bool = False


def water_kind() -> MolKind:
    return MolKind("water", np.array(WATER_Q), np.array(WATER_SIG), np.array(WATER_EPS),
                   np.array(WATER_MASS), water_body())


def solute_kind(rng, n, radius=3.5, name="solute") -> MolKind:
    body = rigid_blob(rng, n, radius) if n > 1 else np.zeros((1, 3))
    q = neutral_charges(rng, n) if n > 1 else np.zeros(1)
    sig = rng.uniform(3.0, 4.0, n)
    eps = rng.uniform(0.05, 0.2, n)
    elems = ["C", "N", "O", "S"]
    mass = np.array([ATOM_MASS[elems[i % 4]] for i in range(n)])
    return MolKind(name, q, sig, eps, mass, body)


@dataclasses.dataclass
class Synthetic:
    """A system plus a deterministic frame generator."""
    system: ErmSystem
    L: float
    kinds: list            # MolKind per molecule type (1-based moltype order)
    moltype: np.ndarray    # per molecule, 1-based
    seed: int
    cavity: float = 3.3
    solute_mol: int = -1   # index of the solute / test particle

    @property
    def cell(self) -> np.ndarray:
        return np.diag([self.L, self.L, self.L]).astype(np.float64)

    def frame(self, i: int) -> np.ndarray:
        """sitepos (natom x 3, f64 holding FP32 values) of frame i."""
        rng = np.random.default_rng([self.seed, 7919, i])
        s = self.system
        pos = np.zeros((s.natom, 3))
        L = self.L
        solute_sites = None
        if self.solute_mol >= 0 and s.sluvid[self.solute_mol] == PT_SOLUTE:
            kind = self.kinds[self.moltype[self.solute_mol] - 1]
            body = chain(rng, len(kind.q)) if kind.flexible else kind.body
            R = random_rotations(rng, 1)[0]
            c = rng.random(3) * L
            solute_sites = f32(body @ R.T + c)
            b = s.mol_begin[self.solute_mol]
            pos[b:b + len(kind.q)] = solute_sites
        mols = [m for m in range(s.nmol) if s.sluvid[m] == PT_SOLVENT]
        nsolv = len(mols)
        n = int(np.ceil(nsolv ** (1.0 / 3.0)))
        while True:
            grid = (np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1)
                    .reshape(-1, 3) + 0.5) * (L / n)
            centres = grid + rng.uniform(-0.3, 0.3, grid.shape)
            if solute_sites is not None:
                d = centres[:, None, :] - solute_sites[None, :, :]
                d -= L * np.round(d / L)
                keep = np.min(np.linalg.norm(d, axis=2), axis=1) >= self.cavity
                centres = centres[keep]
            if len(centres) >= nsolv:
                break
            n += 1
        sel = rng.permutation(len(centres))[:nsolv]
        centres = centres[sel]
        rot = random_rotations(rng, nsolv)
        for j, m in enumerate(mols):
            kind = self.kinds[self.moltype[m] - 1]
            b = s.mol_begin[m]
            pos[b:b + len(kind.q)] = kind.body @ rot[j].T + centres[j]
        return f32(pos)


def build(kinds, counts, sluvid_of_kind, ms, splodr, L, seed, slttype, hists,
          name, corrcal=0, self_hist=None, **params) -> Synthetic:
    """Assemble an ErmSystem from molecule kinds (moltype = kind index + 1)."""
    moltype, sluvid, mol_begin = [], [], [0]
    q, sig, eps, mass = [], [], [], []
    for t, (kind, cnt) in enumerate(zip(kinds, counts)):
        for _ in range(cnt):
            moltype.append(t + 1)
            sluvid.append(sluvid_of_kind[t])
            mol_begin.append(mol_begin[-1] + len(kind.q))
            q.extend(kind.q)
            sig.extend(kind.sig)
            eps.extend(kind.eps)
            mass.extend(kind.mass)
    moltype = np.array(moltype, dtype=np.int32)
    sluvid = np.array(sluvid, dtype=np.int32)
    mol_begin = np.array(mol_begin, dtype=np.int32)
    charge = CHARGE_SCALE * np.array(q)
    ljtype, epsm, sig2 = lj_tables(np.array(eps), np.array(sig))
    solute = int(np.nonzero(sluvid > 0)[0][0])
    bf = None
    if sluvid[solute] == PT_TEST_RIGID:
        bf = kinds[moltype[solute] - 1].body.copy()
    sysm = ErmSystem(slttype=slttype, ms=tuple(ms), splodr=splodr, corrcal=corrcal,
                     mol_begin=mol_begin, charge=charge, ljtype=ljtype, lj_eps=epsm,
                     lj_sig2=sig2, sluvid=sluvid, uvspec=uvspec_of(moltype, sluvid),
                     mol_charge=mol_charges(charge, mol_begin), sitemass=np.array(mass),
                     bfcoord=bf, hists=hists, self_hist=self_hist, name=name, **params)
    return Synthetic(sysm, L, kinds, moltype, seed, solute_mol=solute)


def soln12(nwater=1000, L=31.0, K=32, seed=SEED0 + 0) -> Synthetic:
    rng = np.random.default_rng(seed)
    slt = solute_kind(rng, 12)
    return build([slt, water_kind()], [1, nwater], [PT_SOLUTE, PT_SOLVENT], (K, K, K), 4, L, seed,
                 SLT_SOLN, [Hist()], "soln12", self_hist=Hist())


def refs20(nwater=4000, L=49.3, K=48, nsite=20, pecore=200, seed=SEED0 + 1) -> Synthetic:
    rng = np.random.default_rng(seed)
    slt = solute_kind(rng, nsite)
    return build([water_kind(), slt], [nwater, 1], [PT_SOLVENT, PT_TEST_RIGID], (K, K, K), 4, L,
                 seed, SLT_REFS_RIGID, [Hist(pecore=pecore)], "refs20", corrcal=1,
                 self_hist=Hist(pecore=pecore))


def soln300(nwater=20000, L=84.5, K=96, nsite=300, seed=SEED0 + 2) -> Synthetic:
    rng = np.random.default_rng(seed)
    body = chain(rng, nsite)
    q = rng.normal(0.0, 0.3, nsite)
    q -= q.mean()
    elems = ["C", "N", "O", "S"]
    slt = MolKind("chain", q, rng.uniform(3.0, 4.0, nsite), rng.uniform(0.05, 0.2, nsite),
                  np.array([ATOM_MASS[elems[i % 4]] for i in range(nsite)]), body, flexible=True)
    return build([slt, water_kind()], [1, nwater], [PT_SOLUTE, PT_SOLVENT], (K, K, K), 6, L, seed,
                 SLT_SOLN, [Hist(ecdmin=-40.0, ecfmns=-5.0, eccore=30.0)], "soln300",
                 self_hist=Hist())


def soln30(nwater=4000, L=49.3, K=48, nsite=30, seed=SEED0 + 3) -> Synthetic:
    rng = np.random.default_rng(seed)
    slt = solute_kind(rng, nsite, radius=4.5)
    return build([slt, water_kind()], [1, nwater], [PT_SOLUTE, PT_SOLVENT], (K, K, K), 4, L, seed,
                 SLT_SOLN, [Hist()], "soln30", self_hist=Hist())


def tiny_refs(nwater=60, L=14.0, K=16, nsite=5, seed=SEED0 + 9, pecore=50, cut=6.5) -> Synthetic:
    """A small test-particle system for fast parity tests."""
    rng = np.random.default_rng(seed)
    slt = solute_kind(rng, nsite, radius=2.0)
    return build([water_kind(), slt], [nwater, 1], [PT_SOLVENT, PT_TEST_RIGID], (K, K, K), 4, L,
                 seed, SLT_REFS_RIGID, [Hist(pecore=pecore)], "tiny_refs", corrcal=1,
                 self_hist=Hist(pecore=pecore), elecut=cut, upljcut=cut, lwljcut=cut - 1.5,
                 screen=0.5, block_threshold=3.0)


def tiny_soln(nwater=60, L=14.0, K=16, nsite=5, seed=SEED0 + 8, cut=6.5) -> Synthetic:
    rng = np.random.default_rng(seed)
    slt = solute_kind(rng, nsite, radius=2.0)
    return build([slt, water_kind()], [1, nwater], [PT_SOLUTE, PT_SOLVENT], (K, K, K), 4, L, seed,
                 SLT_SOLN, [Hist(ecdmin=-40.0, eccore=30.0)], "tiny_soln", self_hist=Hist(),
                 elecut=cut, upljcut=cut, lwljcut=cut - 1.5, screen=0.5, block_threshold=3.0)


# ions (CHARMM-like Lennard-Jones parameters)
ION_NA = dict(q=1.0, sig=2.43, eps=0.047, mass=ATOM_MASS["Na"])
ION_CL = dict(q=-1.0, sig=4.04, eps=0.150, mass=ATOM_MASS["Cl"])


def ion_kind(name, d) -> MolKind:
    return MolKind(name, np.array([d["q"]]), np.array([d["sig"]]), np.array([d["eps"]]),
                   np.array([d["mass"]]), np.zeros((1, 3)))


def mixed_hists():
    """150 bins per species (600 per solute over the four species): water and
    cosolvent linear over [-102.5, 52.5] (0.1 near 0, 0.25, then 5) + 32 log
    bins to 1e4; the ions' linear range reaches -302.5 (the reference stops
    with eng_min below the first edge; a 150-site solute binds a cosolvent
    molecule by more than 30 kcal/mol)."""
    neutral = Hist(ecdmin=-100.0, ecfmns=-5.0, ecdcen=0.0, eccore=45.0, eclbin=5.0, ecfbin=0.25,
                   ec0bin=0.1, pecore=32)
    ion = Hist(ecdmin=-300.0, ecfmns=-5.0, ecdcen=0.0, eccore=300.0, eclbin=20.0, ecfbin=0.25,
               ec0bin=0.1, pecore=30)
    return [neutral, neutral, ion, ion]


MIXED_SOLUTE_SITES = (5, 20, 60, 150)


def mixed(nsite=20, nwater=7000, ncosolv=800, nion=100, L=60.0, K=64, seed=SEED0 + 4) -> Synthetic:
    """Heterogeneous insertion system: 7000 waters, 800 rigid 10-site neutral
    cosolvent molecules, 100 Na+ and 100 Cl- in a 60 A box; one rigid test
    particle of ``nsite`` sites (5, 20, 60 or 150).  The solvent of every
    solute size is identical (same kinds, counts and seed), so frames agree
    on all solvent sites."""
    rng = np.random.default_rng(seed)
    cos = solute_kind(rng, 10, radius=2.0, name="cosolvent")
    radius = {5: 2.0, 20: 3.5, 60: 5.0, 150: 6.5}.get(nsite, 1.2 * nsite ** (1.0 / 3.0) + 1.0)
    slt = solute_kind(np.random.default_rng([seed, nsite]), nsite, radius=radius)
    kinds = [water_kind(), cos, ion_kind("Na", ION_NA), ion_kind("Cl", ION_CL), slt]
    return build(kinds, [nwater, ncosolv, nion, nion, 1],
                 [PT_SOLVENT, PT_SOLVENT, PT_SOLVENT, PT_SOLVENT, PT_TEST_RIGID], (K, K, K), 4, L, seed,
                 SLT_REFS_RIGID, mixed_hists(), f"mixed{nsite}", corrcal=1)


WORKLOADS = {"soln12": soln12, "refs20": refs20, "soln300": soln300, "soln30": soln30, "mixed": mixed}


def write_dcd(path, frames, cells=None, big_endian=False, title="ermod_amd synthetic"):
    """A DCD trajectory as CHARMM writes it (Fortran unformatted records):
    the 'CORD' header (icntrl(11) = 1 when cells are given), one title line,
    the atom count, then per frame the unit-cell record (A, 0, B, 0, 0, C for
    the diagonal cell, the layout trajectory-formats/dcd.f90 reads into
    cell(1,1), (1,2), (2,2), (1,3), (2,3), (3,3)) and FP32 x, y, z records."""
    import struct
    e = ">" if big_endian else "<"
    frames = np.asarray(frames, dtype=np.float32)
    nf, n = frames.shape[0], frames.shape[1]

    def rec(payload):
        return struct.pack(e + "i", len(payload)) + payload + struct.pack(e + "i", len(payload))

    icntrl = [0] * 20
    icntrl[0] = nf
    icntrl[2] = 1
    icntrl[10] = 1 if cells is not None else 0
    icntrl[19] = 24
    hdr = b"CORD" + struct.pack(e + "20i", *icntrl)
    out = [rec(hdr), rec(struct.pack(e + "i", 1) + title.encode().ljust(80)[:80]), rec(struct.pack(e + "i", n))]
    for k in range(nf):
        if cells is not None:
            c = np.asarray(cells[k], dtype=np.float64)
            out.append(rec(struct.pack(e + "6d", c[0, 0], c[0, 1], c[1, 1], c[0, 2], c[1, 2], c[2, 2])))
        for d in range(3):
            out.append(rec(frames[k, :, d].astype(e + "f4").tobytes()))
    with open(path, "wb") as f:
        f.write(b"".join(out))
