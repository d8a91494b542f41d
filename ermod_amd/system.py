"""System description: the engmain data contract of the energy pipeline.

The arrays here are what ERmod's setup (setconf.F90:538-837) and enginit
(engproc.F90:37-274) leave in the engmain module globals
(engmain.F90:252-301) for the hot path to read: scaled charges, LJ types and
combined tables, molecule ranges, solute/solvent ids, species, masses and the
body-frame coordinates of a test particle.  Both the device library wrapper
(ermod_amd.api) and the CPU oracle take an ``ErmSystem``.

Indices are 0-based; ``mol_begin`` has nmol+1 entries.
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

# setconf.F90:829 -- (e^2 N_A / (4 pi eps0) in kcal/mol A)^(1/2)
CHARGE_SCALE = 18.22261721
# setconf.F90:313 -- kT in kcal/mol from K
GAS_CONST = 8.314510e-3
CAL_PER_J = 4.184
# setconf.F90:561 -- Rmin/2 -> sigma
SGMCNV = 1.7817974362806784

SLT_SOLN, SLT_REFS_RIGID, SLT_REFS_FLEX = 1, 2, 3
ES_NVT, ES_NPT = 1, 2
# cltype and boxshp (engmain.F90:306,314); Ewald (1) is refused by the reference
EL_COULOMB, EL_PME, EL_PPPM = 0, 2, 3
SYS_NONPERIODIC, SYS_PERIODIC = 0, 1
# insertion modes (engmain.F90:320-327)
INSORG_ORIGIN, INSORG_NOCHANGE, INSORG_AGGCEN, INSORG_REFSTR = 0, 1, 2, 3
(INSPOS_RANDOM, INSPOS_NOCHANGE, INSPOS_SPHERE, INSPOS_SLAB_GENERIC, INSPOS_SLAB_SYMMETRIC, INSPOS_RMSD,
 INSPOS_GAUSS) = range(7)
INSROT_RANDOM, INSROT_NOCHANGE = 0, 1
PT_SOLVENT, PT_SOLUTE, PT_TEST_RIGID, PT_TEST_FLEX = 0, 1, 2, 3

# getmass (setconf.F90:999-1053)
ATOM_MASS = {
    "H": 1.00794, "C": 12.0107, "O": 15.9994, "N": 14.00674, "S": 32.066,
    "P": 30.973761, "He": 4.0026, "B": 10.811, "Li": 6.941, "Na": 22.989770,
    "K": 39.0983, "F": 18.9984032, "Cl": 35.4527, "Br": 79.904, "I": 126.90447,
    "Ca": 40.078, "Zn": 65.409, "Fe": 55.845, "Cu": 63.546,
}


@dataclasses.dataclass
class Hist:
    """&hist namelist for one species (engproc.F90:66-67)."""
    ecdmin: float = -19.0
    ecfmns: float = -5.0
    ecdcen: float = 0.0
    eccore: float = 14.0
    ecdmax: float = 1.0e4
    eclbin: float = 0.1
    ecfbin: float = 0.1
    ec0bin: float = 0.1
    finfac: float = 10.0
    pecore: int = 50


@dataclasses.dataclass
class ErmSystem:
    # parameters (&ene_param, engmain.F90:331-343)
    slttype: int
    ms: tuple
    splodr: int = 4
    estype: int = ES_NVT
    ljswitch: int = 0
    corrcal: int = 0
    selfcal: int = 0
    wgtslf: int = 0
    force_calculation: int = 0
    elecut: float = 12.0
    lwljcut: float = 10.0
    upljcut: float = 12.0
    screen: float = 0.35
    temp: float = 300.0 * GAS_CONST / CAL_PER_J
    block_threshold: float = 4.0
    cltype: int = EL_PME           # EL_COULOMB (bare), EL_PME or EL_PPPM
    boxshp: int = SYS_PERIODIC     # SYS_NONPERIODIC with bare Coulomb only
    # topology
    mol_begin: np.ndarray = None      # int32 nmol+1
    charge: np.ndarray = None         # f64 natom, scaled
    ljtype: np.ndarray = None         # int32 natom
    lj_eps: np.ndarray = None         # f64 ntype x ntype
    lj_sig2: np.ndarray = None        # f64 ntype x ntype
    sluvid: np.ndarray = None         # int32 nmol
    uvspec: np.ndarray = None         # int32 nmol
    mol_charge: np.ndarray = None     # f64 nmol
    sitemass: np.ndarray = None       # f64 natom
    bfcoord: Optional[np.ndarray] = None  # f64 nsite x 3 of the test particle
    # energy coordinate, one Hist per solvent species (and one for self)
    hists: list = None
    self_hist: Optional[Hist] = None
    name: str = ""

    @property
    def natom(self) -> int:
        return int(self.mol_begin[-1])

    @property
    def nmol(self) -> int:
        return len(self.mol_begin) - 1

    @property
    def ntype(self) -> int:
        return self.lj_eps.shape[0]

    @property
    def numslv(self) -> int:
        return len(self.hists)

    def physical(self) -> np.ndarray:
        """tagpt: molecules present in the trajectory (engproc.F90:336-347)."""
        return np.nonzero((self.sluvid == PT_SOLVENT) | (self.sluvid == PT_SOLUTE))[0].astype(np.int32)

    def solutes(self) -> np.ndarray:
        return np.nonzero(self.sluvid > 0)[0].astype(np.int32)

    def nsite(self, mol: int) -> int:
        return int(self.mol_begin[mol + 1] - self.mol_begin[mol])

    def solute_mass(self, mol: int) -> np.ndarray:
        return self.sitemass[self.mol_begin[mol]:self.mol_begin[mol + 1]]


def lj_tables(eps: np.ndarray, sig: np.ndarray, cmbrule: int = 0):
    """LJ type assignment and combination (setconf.F90:755-823).

    Types are the distinct (sigma, epsilon) pairs in order of first
    appearance; sigma^2 is combined by the arithmetic (cmbrule 0) or
    geometric (1) rule, epsilon always geometrically.
    """
    types: list = []
    ljtype = np.empty(len(eps), dtype=np.int32)
    for i, (e, s) in enumerate(zip(eps, sig)):
        for t, (tl, te) in enumerate(types):
            if tl == s and te == e:
                ljtype[i] = t
                break
        else:
            types.append((s, e))
            ljtype[i] = len(types) - 1
    ln = np.array([t[0] for t in types], dtype=np.float64)
    en = np.array([t[1] for t in types], dtype=np.float64)
    n = len(types)
    sig2 = np.empty((n, n))
    epsm = np.empty((n, n))
    for i in range(n):
        if cmbrule == 0:
            sig2[:, i] = ((ln + ln[i]) / 2.0) ** 2
        else:
            sig2[:, i] = ln * ln[i]
        epsm[:, i] = np.sqrt(en * en[i])
    return ljtype, epsm, sig2


def uvspec_of(moltype: np.ndarray, sluvid: np.ndarray) -> np.ndarray:
    """Solvent species numbering (engproc.F90:96-103) for one solute type."""
    uv = moltype.astype(np.int32).copy()
    slt = sluvid != PT_SOLVENT
    solute_type = int(moltype[slt][0])
    if int(slt.sum()) == 1:
        uv[slt] = 0
        uv[(~slt) & (moltype > solute_type)] -= 1
    return uv


def mol_charges(charge: np.ndarray, mol_begin: np.ndarray) -> np.ndarray:
    """setconf.F90:832-834, sum in site order."""
    out = np.empty(len(mol_begin) - 1)
    for i in range(len(out)):
        s = 0.0
        for a in range(mol_begin[i], mol_begin[i + 1]):
            s = s + float(charge[a])
        out[i] = s
    return out


def site_lj(system: ErmSystem):
    """Per-site LJ epsilon and sigma: the diagonals of the combined tables
    (both combination rules give sigma_ii^2 = sigma_i^2)."""
    t = system.ljtype
    return system.lj_eps[t, t].copy(), np.sqrt(system.lj_sig2[t, t])


def with_solute_params(system: ErmSystem, mol: int, q=None, eps=None, sigma=None,
                       cmbrule: int = 0) -> ErmSystem:
    """A copy of ``system`` whose molecule ``mol`` has per-site charges
    ``q`` (scaled units), LJ ``eps`` and ``sigma`` (None keeps a field);
    types and tables are rebuilt with combination rule ``cmbrule``
    (setconf.F90:755-823), charges re-summed per molecule."""
    b, e = int(system.mol_begin[mol]), int(system.mol_begin[mol + 1])
    charge = system.charge.astype(np.float64).copy()
    ep, sg = site_lj(system)
    if q is not None:
        charge[b:e] = np.asarray(q, dtype=np.float64)
    if eps is not None:
        ep[b:e] = np.asarray(eps, dtype=np.float64)
    if sigma is not None:
        sg[b:e] = np.asarray(sigma, dtype=np.float64)
    ljtype, epsm, sig2 = lj_tables(ep, sg, cmbrule)
    return dataclasses.replace(system, charge=charge, ljtype=ljtype, lj_eps=epsm, lj_sig2=sig2,
                               mol_charge=mol_charges(charge, system.mol_begin))


def cubic_cell(L: float) -> np.ndarray:
    """3x3 cell, columns = box vectors, stored column-major as 9 values."""
    return np.diag([L, L, L]).astype(np.float64)
