"""ctypes front end of the oracle (oracle/ercpu.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, where it checks or times the product; the
product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liberc.so")
REFLIB = os.path.join(HERE, "_ref", "libref.so")

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class OrcHist(C.Structure):
    _fields_ = [("ecdmin", C.c_double), ("ecfmns", C.c_double), ("ecdcen", C.c_double),
                ("eccore", C.c_double), ("ecdmax", C.c_double), ("eclbin", C.c_double),
                ("ecfbin", C.c_double), ("ec0bin", C.c_double), ("finfac", C.c_double),
                ("pecore", C.c_int32)]


class OrcSystem(C.Structure):
    _fields_ = [("slttype", C.c_int32), ("estype", C.c_int32), ("ljswitch", C.c_int32),
                ("splodr", C.c_int32), ("ms", C.c_int32 * 3), ("corrcal", C.c_int32),
                ("selfcal", C.c_int32), ("wgtslf", C.c_int32), ("force_calculation", C.c_int32),
                ("elecut", C.c_double), ("lwljcut", C.c_double), ("upljcut", C.c_double),
                ("screen", C.c_double), ("temp", C.c_double), ("block_threshold", C.c_double),
                ("natom", C.c_int32), ("nmol", C.c_int32), ("ntype", C.c_int32),
                ("mol_begin", _i32p), ("charge", _f64p), ("ljtype", _i32p), ("lj_eps", _f64p),
                ("lj_sig2", _f64p), ("sluvid", _i32p), ("uvspec", _i32p), ("mol_charge", _f64p),
                ("numslv", C.c_int32), ("uvmax", _i32p), ("uvsoft", _i32p), ("uvcrd", _f64p),
                ("esmax", C.c_int32), ("escrd", _f64p), ("cltype", C.c_int32), ("boxshp", C.c_int32)]


class OrcInsertion(C.Structure):
    _fields_ = [("insorigin", C.c_int32), ("insposition", C.c_int32), ("insorient", C.c_int32),
                ("lwreg", C.c_double), ("upreg", C.c_double)]


class OrcAcc(C.Structure):
    _fields_ = [("edens", _f64p), ("ecorr", _f64p), ("eself", _f64p), ("minuv", _f64p),
                ("maxuv", _f64p), ("slnuv", _f64p), ("engnorm", C.c_double),
                ("engsmpl", C.c_double), ("avslf", C.c_double), ("voffset", C.c_double),
                ("ermax", C.c_int32), ("esmax", C.c_int32), ("numslv", C.c_int32),
                ("voffset_initialized", C.c_int32)]


def build_lib() -> None:
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build_lib()
        L = C.CDLL(LIB)
        vp = C.c_void_p
        L.orc_spline_coeff.argtypes = [C.c_int32, _f64p]
        L.orc_spline_value.argtypes = [_f64p, C.c_int32, C.c_double]
        L.orc_spline_value.restype = C.c_double
        L.orc_build_edges.argtypes = [C.POINTER(OrcHist), C.c_int32, _f64p, C.c_int32, _i32p, _i32p]
        L.orc_binsearch.argtypes = [_f64p, C.c_int32, C.c_double]
        L.orc_hash_double.argtypes = [_f64p, C.c_int32]
        L.orc_hash_double.restype = C.c_uint64
        L.orc_inverted_cell.argtypes = [_f64p, _f64p, _f64p]
        L.orc_fft_r2c_3d.argtypes = [_i32p, _f64p, _f64p]
        L.orc_fft_c2r_3d.argtypes = [_i32p, _f64p, _f64p]
        L.orc_spline_axis.argtypes = [C.c_int32, C.c_int32, _f64p]
        L.orc_greenfunc.argtypes = [_i32p, C.c_int32, _f64p, C.c_double, C.c_double, _f64p, _f64p,
                                    _f64p, _f64p]
        L.orc_aggregate_center.argtypes = [C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_insertion_pose.argtypes = [C.c_int32, _f64p, _f64p, _f64p, C.POINTER(OrcInsertion), _f64p,
                                         C.c_uint64, C.c_uint64, _f64p]
        L.orc_set_intermediate.argtypes = [vp, _i32p]
        L.orc_get_intermediate.argtypes = [vp, _f64p, _f64p, _f64p]
        L.orc_fit_a_rotate_b.argtypes = [C.c_int32, _f64p, _f64p, _f64p, C.c_int32, _f64p, _f64p]
        L.orc_rmsd_nofit.argtypes = [C.c_int32, _f64p, _f64p, _f64p]
        L.orc_rmsd_nofit.restype = C.c_double
        L.orc_refstr_pose.argtypes = [C.c_int32, _f64p, _f64p, C.POINTER(OrcInsertion), _f64p, _f64p, _f64p,
                                      _f64p, C.c_uint64, C.c_uint64, _f64p, _f64p]
        L.orc_gfb_pppm.argtypes = [C.c_int32, _f64p]
        L.orc_pppm_greenfunc.argtypes = [_i32p, C.c_int32, _f64p, _f64p, C.c_double, C.c_double, _f64p]
        L.orc_realcal_bare.argtypes = [vp, C.c_int32, _f64p, _f64p]
        L.orc_spline_site.argtypes = [_f64p, _f64p, _i32p, C.c_int32, _f64p, _f64p, _i32p]
        L.orc_place_solute.argtypes = [C.c_int32, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]
        L.orc_rng_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32]
        L.orc_rng_uniform.restype = C.c_double
        L.orc_random_pose.argtypes = [C.c_uint64, C.c_uint64, _f64p, _f64p]
        L.orc_create.argtypes = [C.POINTER(OrcSystem), _i32p]
        L.orc_create.restype = vp
        L.orc_destroy.argtypes = [vp]
        L.orc_slvmax.argtypes = [vp]
        L.orc_set_frame.argtypes = [vp, _f64p, _f64p, C.c_double]
        L.orc_uv_energy.argtypes = [vp, C.c_int32, _f64p, _f64p]
        L.orc_realcal_proc.argtypes = [vp, C.c_int32, _f64p, _f64p]
        L.orc_realcal_brute.argtypes = [vp, C.c_int32, _f64p, _f64p]
        L.orc_realcal_self.argtypes = [vp, C.c_int32, _f64p, _f64p]
        L.orc_recpcal.argtypes = [vp, C.c_int32, _f64p, _f64p, _f64p, _f64p]
        L.orc_update_histogram.argtypes = [vp, C.c_int32, C.c_double, _f64p, _i32p]
        L.orc_engclear.argtypes = [vp]
        L.orc_engstore.argtypes = [vp]
        L.orc_get_acc.argtypes = [vp, C.POINTER(OrcAcc)]
        L.orc_flceng.argtypes = [vp]
        L.orc_flceng.restype = _f64p
        L.orc_solute_gradient.argtypes = [vp, C.c_int32, _f64p, C.c_int32, _f64p, _f64p, _f64p, _f64p]
        L.orc_bench_insertions.argtypes = [C.POINTER(vp), C.c_int32, C.c_int32, _i32p, _i32p, C.POINTER(_f64p),
                                           C.POINTER(_f64p), _f64p, C.c_int64, C.c_uint64, C.c_uint64, C.c_double,
                                           C.POINTER(C.c_int64)]
        _lib = L
    return _lib


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def P(a, t=_f64p):
    return None if a is None else a.ctypes.data_as(t)


class OracleError(RuntimeError):
    def __init__(self, code, where):
        super().__init__(f"oracle status {code} in {where}")
        self.code = code


def check(code, where):
    if code:
        raise OracleError(code, where)


def spline_coeff(order):
    c = np.zeros(order + 1)
    lib().orc_spline_coeff(order, P(c))
    return c


def spline_value(order, x):
    c = spline_coeff(order)
    return np.array([lib().orc_spline_value(P(c), order, float(v)) for v in np.atleast_1d(x)])


def host_molecules(system, hostspec):
    """com_aggregate's molecule selection: every physical molecule whose
    species (uvspec) is listed in hostspec."""
    s = system
    return i32([(int(s.sluvid[m]) <= 1 and int(s.uvspec[m]) in set(hostspec)) for m in range(s.nmol)])


def aggregate_center(system, sitepos, hostspec):
    """com_aggregate (insertion.F90:539-566) of a frame."""
    s = system
    out = np.zeros(3)
    lib().orc_aggregate_center(s.nmol, P(i32(s.mol_begin), _i32p), P(host_molecules(s, hostspec), _i32p),
                               P(f64(s.sitemass)), P(f64(sitepos)), P(out))
    return out


def insertion_pose(bfcoord, mass, cell, ins, seed, index, aggcen=None):
    """instslt's placement of insertion `index` of stream `seed`; ins =
    (insorigin, insposition, insorient, lwreg, upreg)."""
    bf = f64(bfcoord)
    n = bf.size // 3
    st = OrcInsertion(int(ins[0]), int(ins[1]), int(ins[2]), float(ins[3]), float(ins[4]))
    out = np.zeros((n, 3))
    cm = f64(np.asarray(cell).T.copy())
    ag = None if aggcen is None else f64(aggcen)
    lib().orc_insertion_pose(n, P(bf), P(f64(mass)), P(cm), C.byref(st), P(ag), C.c_uint64(seed),
                             C.c_uint64(index), P(out))
    return out


def fit_a_rotate_b(refa, a, w, b):
    """bestfit's fit_a_rotate_b: b moved by the superposition of a onto refa."""
    refa, a, b, w = f64(refa), f64(a), f64(b), f64(w)
    out = np.zeros_like(b)
    lib().orc_fit_a_rotate_b(len(a), P(refa), P(a), P(w), len(b), P(b), P(out))
    return out


def rmsd_nofit(a, b, w):
    return lib().orc_rmsd_nofit(len(f64(a)), P(f64(a)), P(f64(b)), P(f64(w)))


def reffit(frame_pos, host_atoms, host_ref, host_w, slt_ref, slt_w, bfcoord):
    """reffit (insertion.F90:598-622): (fitted solute, best-fit reference
    solute) for a frame."""
    hc = f64(np.asarray(frame_pos)[np.asarray(host_atoms)])
    best = fit_a_rotate_b(hc, host_ref, host_w, slt_ref)
    fitted = fit_a_rotate_b(best, bfcoord, slt_w, bfcoord)
    return fitted, best


def gauss_constants(cell, upreg):
    """uniform_gauss_mixture's first-call constants: (L/2)/upreg per axis,
    z0, z1 (insertion.F90:240-246)."""
    from math import erf, sqrt, pi
    cell = np.asarray(cell, dtype=np.float64)
    gl = [(float(np.sqrt(np.sum(cell[:, d] ** 2))) / 2.0) / upreg for d in range(3)]
    sp = sqrt(pi)
    return np.array(gl + [8.0 * gl[0] * gl[1] * gl[2], (sp * sp * sp) * erf(gl[0]) * erf(gl[1]) * erf(gl[2])])


def refstr_pose(mass, cell, ins, fitted, best, w, gconst, seed, index):
    """A reference-structure insertion: (coordinates, weight)."""
    fitted = f64(fitted)
    n = len(fitted)
    st = OrcInsertion(int(ins[0]), int(ins[1]), int(ins[2]), float(ins[3]), float(ins[4]))
    out = np.zeros((n, 3))
    wt = C.c_double()
    cm = f64(np.asarray(cell).T.copy())
    check(lib().orc_refstr_pose(n, P(f64(mass)), P(cm), C.byref(st), P(fitted), P(f64(best)), P(f64(w)),
                                P(f64(gconst)), C.c_uint64(seed), C.c_uint64(index), P(out), C.byref(wt)),
          "refstr_pose")
    return out, wt.value


def gfb_pppm(order):
    """calc_gfb_pppm (recpcal.F90:168-186): gf_b(1..order)."""
    out = np.zeros(order)
    lib().orc_gfb_pppm(order, P(out))
    return out


def pppm_greenfunc(K, order, cell, screen):
    """recpcal_pppm_greenfunc on the (K3, K2, K1/2+1) half spectrum."""
    cell = np.asarray(cell, dtype=np.float64)
    inv = np.linalg.inv(cell)
    out = np.zeros((K[2], K[1], K[0] // 2 + 1))
    invcm = f64(inv.T.copy())
    cm = f64(cell.T.copy())
    check(lib().orc_pppm_greenfunc(P(i32(K), _i32p), order, P(cm), P(invcm), float(np.linalg.det(cell)),
                                   float(screen), P(out)), "pppm_greenfunc")
    return out


def hist_struct(h) -> OrcHist:
    return OrcHist(h.ecdmin, h.ecfmns, h.ecdcen, h.eccore, h.ecdmax, h.eclbin, h.ecfbin,
                   h.ec0bin, h.finfac, int(h.pecore))


def build_edges(h, self_energy=False):
    out = np.zeros(10002)
    pemax, pesoft = C.c_int32(), C.c_int32()
    hs = hist_struct(h)
    check(lib().orc_build_edges(C.byref(hs), int(self_energy), P(out), len(out), C.byref(pemax),
                                C.byref(pesoft)), "build_edges")
    return out[:pemax.value].copy(), pemax.value, pesoft.value


def binsearch(edges, v):
    e = f64(edges)
    return lib().orc_binsearch(P(e), len(e), float(v))


def hash_double(v):
    a = f64(v).ravel()
    return lib().orc_hash_double(P(a), len(a))


def fft_r2c(grid):
    g = f64(grid)  # shape (K3, K2, K1), x fastest
    K3, K2, K1 = g.shape
    K = i32([K1, K2, K3])
    out = np.zeros((K3, K2, K1 // 2 + 1), dtype=np.complex128)
    check(lib().orc_fft_r2c_3d(P(K, _i32p), P(g), out.ctypes.data_as(_f64p)), "fft_r2c")
    return out


def fft_c2r(spec, K1):
    s = np.ascontiguousarray(spec, dtype=np.complex128)
    K3, K2, _ = s.shape
    K = i32([K1, K2, K3])
    out = np.zeros((K3, K2, K1))
    check(lib().orc_fft_c2r_3d(P(K, _i32p), s.ctypes.data_as(_f64p), P(out)), "fft_c2r")
    return out


def random_pose(seed, index):
    r, q = np.zeros(3), np.zeros(4)
    lib().orc_random_pose(C.c_uint64(seed), C.c_uint64(index), P(r), P(q))
    return r, q


def place_solute(bfcoord, mass, cell, r, q):
    bf = f64(bfcoord)
    out = np.zeros_like(bf)
    lib().orc_place_solute(len(bf), P(bf), P(f64(mass)), P(f64(cell).T.copy()), P(f64(r)),
                           P(f64(q)), P(out))
    return out


class Oracle:
    """One oracle context over an ErmSystem (the per-frame pipeline)."""

    def __init__(self, system):
        s = system
        self.system = s
        self._keep = []
        edges, uvmax, uvsoft = [], [], []
        for h in s.hists:
            e, pemax, pesoft = build_edges(h)
            edges.append(e)
            uvmax.append(pemax)
            uvsoft.append(pesoft)
        self.uvcrd = f64(np.concatenate(edges))
        self.uvmax = i32(uvmax)
        self.uvsoft = i32(uvsoft)
        self.escrd, self.esmax = None, 0
        if s.selfcal:
            self.escrd, self.esmax, _ = build_edges(s.self_hist, self_energy=True)
            self.escrd = f64(self.escrd)
        arrs = dict(mol_begin=i32(s.mol_begin), charge=f64(s.charge), ljtype=i32(s.ljtype),
                    lj_eps=f64(s.lj_eps), lj_sig2=f64(s.lj_sig2), sluvid=i32(s.sluvid),
                    uvspec=i32(s.uvspec), mol_charge=f64(s.mol_charge))
        self._keep.append(arrs)
        st = OrcSystem()
        st.slttype, st.estype, st.ljswitch, st.splodr = s.slttype, s.estype, s.ljswitch, s.splodr
        st.ms[:] = list(s.ms)
        st.corrcal, st.selfcal, st.wgtslf = s.corrcal, s.selfcal, s.wgtslf
        st.force_calculation = s.force_calculation
        st.elecut, st.lwljcut, st.upljcut = s.elecut, s.lwljcut, s.upljcut
        st.screen, st.temp, st.block_threshold = s.screen, s.temp, s.block_threshold
        st.natom, st.nmol, st.ntype = s.natom, s.nmol, s.ntype
        st.mol_begin = P(arrs["mol_begin"], _i32p)
        st.charge = P(arrs["charge"])
        st.ljtype = P(arrs["ljtype"], _i32p)
        st.lj_eps = P(arrs["lj_eps"])
        st.lj_sig2 = P(arrs["lj_sig2"])
        st.sluvid = P(arrs["sluvid"], _i32p)
        st.uvspec = P(arrs["uvspec"], _i32p)
        st.mol_charge = P(arrs["mol_charge"])
        st.numslv = s.numslv
        st.uvmax = P(self.uvmax, _i32p)
        st.uvsoft = P(self.uvsoft, _i32p)
        st.uvcrd = P(self.uvcrd)
        st.esmax = self.esmax
        st.escrd = P(self.escrd)
        st.cltype, st.boxshp = s.cltype, s.boxshp
        self._st = st
        status = C.c_int32()
        self.h = lib().orc_create(C.byref(st), C.byref(status))
        check(status.value, "create")
        self.slvmax = lib().orc_slvmax(self.h)
        self.ermax = int(self.uvmax.sum())

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_frame(self, sitepos, cell, weight=1.0):
        pos = f64(sitepos)
        cm = f64(np.asarray(cell).T.copy())  # column-major 3x3
        check(lib().orc_set_frame(self.h, P(pos), P(cm), float(weight)), "set_frame")

    def uv_energy(self, solute, solute_pos=None):
        out = np.zeros(self.slvmax + 1)
        sp = None if solute_pos is None else f64(solute_pos)
        check(lib().orc_uv_energy(self.h, int(solute), P(sp), P(out)), "uv_energy")
        return out

    def realcal(self, solute, solute_pos=None, brute=False):
        out = np.zeros(self.slvmax)
        sp = None if solute_pos is None else f64(solute_pos)
        fn = lib().orc_realcal_brute if brute else lib().orc_realcal_proc
        check(fn(self.h, int(solute), P(sp), P(out)), "realcal")
        return out

    def realcal_bare(self, solute, solute_pos=None):
        out = np.zeros(self.slvmax)
        sp = None if solute_pos is None else f64(solute_pos)
        check(lib().orc_realcal_bare(self.h, int(solute), P(sp), P(out)), "realcal_bare")
        return out

    def realcal_self(self, solute, solute_pos=None):
        v = C.c_double()
        sp = None if solute_pos is None else f64(solute_pos)
        check(lib().orc_realcal_self(self.h, int(solute), P(sp), C.byref(v)), "realcal_self")
        return v.value

    def recpcal(self, solute, solute_pos=None, want_phi=False):
        out = np.zeros(self.slvmax)
        self_e = C.c_double()
        K = self.system.ms
        phi = np.zeros((K[2], K[1], K[0])) if want_phi else None
        sp = None if solute_pos is None else f64(solute_pos)
        check(lib().orc_recpcal(self.h, int(solute), P(sp), P(out), C.byref(self_e), P(phi)),
              "recpcal")
        return (out, self_e.value, phi) if want_phi else (out, self_e.value)

    def solute_gradient(self, solute, solute_pos=None, cmbrule=0, q=None, eps=None, sigma=None):
        """(u, du/dq, du/deps, du/dsigma) of one configuration; q in scaled
        units, None parameters taken from the topology."""
        ns = int(self.system.mol_begin[solute + 1] - self.system.mol_begin[solute])
        out = np.zeros(1 + 3 * ns)
        sp = None if solute_pos is None else f64(solute_pos)
        qq = None if q is None else f64(q)
        ee = None if eps is None else f64(eps)
        ss = None if sigma is None else f64(sigma)
        check(lib().orc_solute_gradient(self.h, int(solute), P(sp), int(cmbrule), P(qq), P(ee), P(ss),
                                        P(out)), "solute_gradient")
        return out[0], out[1:1 + ns], out[1 + ns:1 + 2 * ns], out[1 + 2 * ns:]

    def update_histogram(self, solute, uvengy, weight=1.0):
        ins = np.zeros(self.ermax, dtype=np.int32)
        u = f64(uvengy)
        check(lib().orc_update_histogram(self.h, int(solute), float(weight), P(u), P(ins, _i32p)),
              "update_histogram")
        return ins

    def set_intermediate(self, zerodns):
        self._zerodns = None if zerodns is None else i32(zerodns)
        check(lib().orc_set_intermediate(self.h, P(self._zerodns, _i32p)), "set_intermediate")

    def intermediate(self):
        """(edintm, uvintm, norm_intm, soln_intm, soln_total)."""
        ed = np.zeros(self.ermax)
        uv = np.zeros(self.system.numslv + 1)
        sc = np.zeros(3)
        check(lib().orc_get_intermediate(self.h, P(ed), P(uv), P(sc)), "get_intermediate")
        return ed, uv, sc[0], sc[1], sc[2]

    def engclear(self):
        lib().orc_engclear(self.h)

    def engstore(self):
        check(lib().orc_engstore(self.h), "engstore")

    def acc(self):
        a = OrcAcc()
        lib().orc_get_acc(self.h, C.byref(a))
        n, ns = a.ermax, a.numslv
        out = dict(edens=np.ctypeslib.as_array(a.edens, (n,)).copy(),
                   minuv=np.ctypeslib.as_array(a.minuv, (ns + 1,)).copy(),
                   maxuv=np.ctypeslib.as_array(a.maxuv, (ns + 1,)).copy(),
                   slnuv=np.ctypeslib.as_array(a.slnuv, (ns + 1,)).copy(),
                   engnorm=a.engnorm, engsmpl=a.engsmpl, avslf=a.avslf, voffset=a.voffset)
        if a.ecorr:
            out["ecorr"] = np.ctypeslib.as_array(a.ecorr, (n, n)).copy()  # [iduv][iduvp]
        if a.eself:
            out["eself"] = np.ctypeslib.as_array(a.eself, (a.esmax,)).copy()
        out["flceng"] = np.ctypeslib.as_array(lib().orc_flceng(self.h), (ns + 1,)).copy()
        return out


def host_threads():
    """Cores this process may run on: the scheduler's affinity set, capped by
    OMP_NUM_THREADS when that is set."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def bench_insertions(systems, solute_mols, frame, cell, n, seconds, seed=0, first_index=0, threads=None):
    """Times engconst's test-particle loop on the CPU: up to n random
    insertions (insertion i of systems[i % len(systems)]), one oracle context
    per (thread, system) holding the frame, OpenMP over insertions, stopped
    after `seconds`.  Returns (insertions done per system, seconds, threads)."""
    import time
    nt = threads or host_threads()
    ntypes = len(systems)
    ctxs = []
    for _ in range(nt):
        for s in systems:
            o = Oracle(s)
            o.set_frame(frame, cell)
            ctxs.append(o)
    handles = (C.c_void_p * len(ctxs))(*[o.h for o in ctxs])
    bfs = [f64(s.bfcoord) for s in systems]
    masses = [f64(s.solute_mass(m)) for s, m in zip(systems, solute_mols)]
    bfp = (_f64p * ntypes)(*[P(b) for b in bfs])
    mp = (_f64p * ntypes)(*[P(m) for m in masses])
    mols = i32(solute_mols)
    nsite = i32([len(b) for b in bfs])
    cm = f64(np.asarray(cell).T.copy())
    done = (C.c_int64 * ntypes)()
    t0 = time.perf_counter()
    st = lib().orc_bench_insertions(handles, nt, ntypes, P(mols, _i32p), P(nsite, _i32p), bfp, mp, P(cm),
                                    int(n), C.c_uint64(seed), C.c_uint64(first_index), float(seconds), done)
    dt = time.perf_counter() - t0
    check(st, "bench_insertions")
    for o in ctxs:
        o.close()
    return [int(d) for d in done], dt, nt


def ref_lib():
    """The reference's spline.F90 + hash_real.c compiled from its sources
    (oracle/_ref/libref.so), or None where it was not built."""
    if not os.path.exists(REFLIB):
        return None
    L = C.CDLL(REFLIB)
    L.ref_spline_init.argtypes = [C.c_int32]
    L.ref_spline_value.argtypes = [C.c_double]
    L.ref_spline_value.restype = C.c_double
    L.hash_double_.argtypes = [_f64p, _i32p, C.POINTER(C.c_uint64)]
    return L
