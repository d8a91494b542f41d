"""ctypes front end of libermgpu.so (include/ermgpu.h).

``ErmGPU`` mirrors the reference's engproc interface (engproc.F90): engclear,
set_frame (the per-frame part of engconst), get_uv_energy, engconst's hot
loop, update_histogram and engstore, plus realcal_proc / realcal_self /
recpcal for stage-wise checks.  It raises ``ErmError`` carrying the library's
status code; nothing here computes energies on the CPU -- a missing or
unloadable library is an error.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .system import ErmSystem

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libermgpu.so")

ERM_FP64, ERM_FP32 = 0, 1
STATUS = {0: "ok", 1: "eng_min", 2: "eng_sft", 3: "mbin", 4: "eng_ecd", 5: "rcp_cns", 6: "noncuboid",
          7: "eng_bug", 8: "eng_cns", 9: "eng_siz", 10: "arg", 11: "unsupported", 12: "device", 13: "eof",
          14: "ins_limit"}
ERM_E_EOF = 13
STAGES = ("prep", "pme_z", "pme_y", "pme_x", "mol_sum", "pair", "self", "bin", "corr", "total")

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class ErmParams(C.Structure):
    _fields_ = [("slttype", C.c_int32), ("estype", C.c_int32), ("ljswitch", C.c_int32),
                ("splodr", C.c_int32), ("ms", C.c_int32 * 3), ("corrcal", C.c_int32),
                ("selfcal", C.c_int32), ("wgtslf", C.c_int32), ("force_calculation", C.c_int32),
                ("elecut", C.c_double), ("lwljcut", C.c_double), ("upljcut", C.c_double),
                ("screen", C.c_double), ("temp", C.c_double), ("block_threshold", C.c_double),
                ("precision", C.c_int32), ("max_batch", C.c_int32), ("device", C.c_int32),
                ("pme_path", C.c_int32), ("cltype", C.c_int32), ("boxshp", C.c_int32)]


class ErmHist(C.Structure):
    _fields_ = [("ecdmin", C.c_double), ("ecfmns", C.c_double), ("ecdcen", C.c_double),
                ("eccore", C.c_double), ("ecdmax", C.c_double), ("eclbin", C.c_double),
                ("ecfbin", C.c_double), ("ec0bin", C.c_double), ("finfac", C.c_double),
                ("pecore", C.c_int32)]


class ErmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


EXPORTS = ("erm_build_edges", "erm_hash_double", "erm_create", "erm_destroy", "erm_last_error",
           "erm_set_topology", "erm_set_bins", "erm_engclear", "erm_set_frame", "erm_set_frame_device",
           "erm_get_uv_energy", "erm_realcal_proc", "erm_realcal_self", "erm_recpcal", "erm_engconst",
           "erm_insert_random", "erm_update_histogram", "erm_engstore", "erm_fetch", "erm_sizes",
           "erm_stage_times", "erm_solute_gradient", "erm_set_tuning", "erm_get_tuning", "erm_autotune",
           "erm_insert_random_multi", "erm_set_insertion", "erm_dcd_open", "erm_dcd_read", "erm_dcd_close",
           "erm_read_permutation", "erm_traj_open", "erm_traj_next", "erm_traj_close", "erm_set_intermediate",
           "erm_fetch_intermediate", "erm_set_refstructure", "erm_insert_flexible", "erm_set_pair_path",
           "erm_get_pair_path", "erm_update_histogram_static",
           "erm_group_create", "erm_group_destroy", "erm_group_last_error", "erm_group_set_frame",
           "erm_group_set_frame_device", "erm_group_insert_random", "erm_group_set_tuning", "erm_group_set_graph",
           "erm_group_stage_times")

_lib = None


def lib():
    """Loads libermgpu.so; raises if it is not built (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is not built: run `python -m ermod_amd.build`")
        L = C.CDLL(LIBPATH)
        vp = C.c_void_p
        L.erm_build_edges.argtypes = [C.POINTER(ErmHist), C.c_int32, _f64p, C.c_int32, _i32p, _i32p]
        L.erm_hash_double.argtypes = [_f64p, C.c_int32]
        L.erm_hash_double.restype = C.c_uint64
        L.erm_create.argtypes = [C.POINTER(ErmParams), C.POINTER(vp)]
        L.erm_destroy.argtypes = [vp]
        L.erm_destroy.restype = None
        L.erm_last_error.argtypes = [vp]
        L.erm_last_error.restype = C.c_char_p
        L.erm_set_topology.argtypes = [vp, C.c_int32, C.c_int32, _i32p, _f64p, _i32p, C.c_int32, _f64p,
                                       _f64p, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.erm_set_bins.argtypes = [vp, C.c_int32, _i32p, _i32p, _f64p, C.c_int32, _f64p]
        L.erm_engclear.argtypes = [vp]
        L.erm_set_frame.argtypes = [vp, _f64p, _f64p, C.c_double]
        L.erm_set_frame_device.argtypes = [vp, C.c_void_p, _f64p, C.c_double]
        L.erm_get_uv_energy.argtypes = [vp, C.c_int32, C.c_int32, _f64p, _f64p]
        L.erm_realcal_proc.argtypes = [vp, C.c_int32, _f64p, _f64p]
        L.erm_realcal_self.argtypes = [vp, C.c_int32, _f64p, _f64p]
        L.erm_recpcal.argtypes = [vp, C.c_int32, _f64p, _f64p, _f64p]
        L.erm_engconst.argtypes = [vp, C.c_int32, _i32p, _f64p, _f64p, _f64p]
        L.erm_insert_random.argtypes = [vp, C.c_int32, C.c_uint64, C.c_uint64, _f64p]
        L.erm_update_histogram.argtypes = [vp, C.c_int32, _i32p, _f64p, _f64p]
        L.erm_update_histogram_static.argtypes = [vp, C.c_int32, _i32p, _f64p, _f64p]
        L.erm_engstore.argtypes = [vp]
        L.erm_fetch.argtypes = [vp, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]
        L.erm_sizes.argtypes = [vp, _i32p, _i32p]
        L.erm_stage_times.argtypes = [vp, _f64p]
        L.erm_set_tuning.argtypes = [vp] + [C.c_int32] * 8
        L.erm_get_tuning.argtypes = [vp, _i32p]
        L.erm_autotune.argtypes = [vp, C.c_int32, C.c_uint64, _f64p]
        L.erm_set_pair_path.argtypes = [vp, C.c_int32]
        L.erm_group_create.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(vp)]
        L.erm_group_destroy.argtypes = [vp]
        L.erm_group_destroy.restype = None
        L.erm_group_last_error.argtypes = [vp]
        L.erm_group_last_error.restype = C.c_char_p
        L.erm_group_set_frame.argtypes = [vp, _f64p, _f64p, C.c_double]
        L.erm_group_set_frame_device.argtypes = [vp, C.c_void_p, _f64p, C.c_double]
        L.erm_group_insert_random.argtypes = [vp, _i32p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.erm_group_set_tuning.argtypes = [vp, C.c_int32, C.c_int32]
        L.erm_group_set_graph.argtypes = [vp, C.c_int32]
        L.erm_group_stage_times.argtypes = [vp, _f64p]
        L.erm_get_pair_path.argtypes = [vp]
        L.erm_insert_random_multi.argtypes = [C.POINTER(vp), C.c_int32, _i32p, C.POINTER(C.c_uint64),
                                              C.POINTER(C.c_uint64)]
        L.erm_solute_gradient.argtypes = [vp, C.c_int32, _i32p, _f64p, C.c_uint64, C.c_uint64, _f64p,
                                          C.c_int32, _f64p, _f64p, _f64p, _f64p]
        L.erm_set_insertion.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_int32,
                                        _i32p]
        L.erm_set_refstructure.argtypes = [vp, C.c_int32, _i32p, _f64p, _f64p, _f64p, _f64p]
        L.erm_insert_flexible.argtypes = [vp, C.c_int32, _f64p, _f64p, C.c_uint64, C.c_uint64, _f64p]
        L.erm_set_intermediate.argtypes = [vp, _i32p]
        L.erm_fetch_intermediate.argtypes = [vp, _f64p, _f64p, _f64p]
        L.erm_dcd_open.argtypes = [C.c_char_p, C.POINTER(vp), _i32p, _i32p]
        L.erm_dcd_read.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(C.c_float), _f64p]
        L.erm_dcd_close.argtypes = [vp]
        L.erm_dcd_close.restype = None
        L.erm_read_permutation.argtypes = [C.c_char_p, C.c_int32, _i32p, _i32p]
        L.erm_traj_open.argtypes = [vp, C.c_char_p, C.c_int32, _i32p, C.c_char_p, C.c_int32, C.POINTER(vp)]
        L.erm_traj_next.argtypes = [vp, _f64p, _f64p]
        L.erm_traj_close.argtypes = [vp]
        L.erm_traj_close.restype = None
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a, t=_f64p):
    return None if a is None else a.ctypes.data_as(t)


def read_dcd(path, natom=None, periodic=True):
    """All frames of a DCD file through the library's host reader (as
    trajectory-formats/dcd.f90 reads them): (coords [F, natom, 3] float32,
    cells [F, 3, 3])."""
    h = C.c_void_p()
    nf, hc = C.c_int32(), C.c_int32()
    st = lib().erm_dcd_open(os.fsencode(path), C.byref(h), C.byref(nf), C.byref(hc))
    if st:
        raise ErmError(st, "erm_dcd_open")
    n = nf.value if natom is None else natom
    frames, cells = [], []
    try:
        while True:
            xyz = np.zeros(3 * n, dtype=np.float32)
            cm = np.zeros(9)
            st = lib().erm_dcd_read(h, n, int(bool(periodic)), xyz.ctypes.data_as(C.POINTER(C.c_float)), _p(cm))
            if st == ERM_E_EOF:
                break
            if st:
                raise ErmError(st, "erm_dcd_read")
            frames.append(xyz.reshape(3, n).T.copy())
            cells.append(cm.reshape(3, 3).T.copy())
    finally:
        lib().erm_dcd_close(h)
    return np.array(frames).reshape(-1, n, 3), np.array(cells).reshape(-1, 3, 3)


def read_permutation(path, natom):
    """PermIndex (setconf.F90:866-905) as a 0-based array, or None when the
    file does not exist."""
    perm = np.zeros(natom, dtype=np.int32)
    found = C.c_int32()
    st = lib().erm_read_permutation(os.fsencode(path), natom, _p(perm, _i32p), C.byref(found))
    if st:
        raise ErmError(st, "erm_read_permutation")
    return perm if found.value else None


class Trajectory:
    """Frames of a DCD file streamed into an ErmGPU context (erm_traj_*)."""

    def __init__(self, engine, path, natom, perm=None, weights=None, skpcnf=1):
        self.engine = engine
        self._perm = None if perm is None else _i32(perm)
        h = C.c_void_p()
        engine._check(lib().erm_traj_open(engine.h, os.fsencode(path), int(natom), _p(self._perm, _i32p),
                                          None if weights is None else os.fsencode(weights), int(skpcnf),
                                          C.byref(h)), "traj_open")
        self.h = h

    def next(self):
        """Loads the next frame; returns (cell, weight), or None at the end."""
        cm = np.zeros(9)
        w = C.c_double()
        st = lib().erm_traj_next(self.h, _p(cm), C.byref(w))
        if st == ERM_E_EOF:
            return None
        self.engine._check(st, "traj_next")
        return cm.reshape(3, 3).T.copy(), w.value

    def close(self):
        if getattr(self, "h", None):
            lib().erm_traj_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def hist_struct(h) -> ErmHist:
    return ErmHist(h.ecdmin, h.ecfmns, h.ecdcen, h.eccore, h.ecdmax, h.eclbin, h.ecfbin, h.ec0bin,
                   h.finfac, int(h.pecore))


def build_edges(h, self_energy=False):
    """enginit's bin edges of one species (engproc.F90:124-237)."""
    out = np.zeros(10002)
    pemax, pesoft = C.c_int32(), C.c_int32()
    hs = hist_struct(h)
    st = lib().erm_build_edges(C.byref(hs), int(self_energy), _p(out), len(out), C.byref(pemax),
                               C.byref(pesoft))
    if st:
        raise ErmError(st, "erm_build_edges")
    return out[:pemax.value].copy(), pemax.value, pesoft.value


def insert_random_multi(engines, n, seeds, first_indices):
    """erm_insert_random on several contexts at once (solutes of different
    sizes on one frame, each on its own stream)."""
    k = len(engines)
    hs = (C.c_void_p * k)(*[e.h for e in engines])
    nn = np.ascontiguousarray(n, dtype=np.int32)
    ss = (C.c_uint64 * k)(*[int(x) for x in seeds])
    ff = (C.c_uint64 * k)(*[int(x) for x in first_indices])
    st = lib().erm_insert_random_multi(hs, k, _p(nn, _i32p), ss, ff)
    if st:
        msgs = [lib().erm_last_error(e.h) for e in engines]
        raise ErmError(st, "insert_random_multi: " + "; ".join(m.decode() for m in msgs if m))


class ErmGroup:
    """Several rigid test particles (one ErmGPU context each, same solvent)
    inserted in one launch sequence (erm_group_*)."""

    def __init__(self, engines):
        self.engines = list(engines)
        k = len(self.engines)
        hs = (C.c_void_p * k)(*[e.h for e in self.engines])
        out = C.c_void_p()
        st = lib().erm_group_create(hs, k, C.byref(out))
        if st:
            raise ErmError(st, "group_create: " + lib().erm_last_error(self.engines[0].h).decode())
        self.h = out.value

    def _check(self, st, where):
        if st:
            raise ErmError(st, f"{where}: {lib().erm_group_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            lib().erm_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_frame(self, sitepos, cell, weight=1.0):
        pos = _f64(sitepos)
        cm = _f64(np.asarray(cell).T.copy())
        self._check(lib().erm_group_set_frame(self.h, _p(pos), _p(cm), float(weight)), "group_set_frame")

    def set_frame_device(self, dev_ptr: int, cell, weight=1.0):
        cm = _f64(np.asarray(cell).T.copy())
        self._check(lib().erm_group_set_frame_device(self.h, C.c_void_p(dev_ptr), _p(cm), float(weight)),
                    "group_set_frame_device")

    def insert_random(self, n, seeds, first_indices):
        k = len(self.engines)
        nn = np.ascontiguousarray(n, dtype=np.int32)
        ss = (C.c_uint64 * k)(*[int(x) for x in seeds])
        ff = (C.c_uint64 * k)(*[int(x) for x in first_indices])
        self._check(lib().erm_group_insert_random(self.h, _p(nn, _i32p), ss, ff), "group_insert_random")

    def set_tuning(self, max_batch=0, persistent=None):
        """persistent: None keeps it; True / False every / no heavy stage
        as one persistent grid; or the bit mask (1 z columns, 2 pairs)."""
        if persistent is None:
            ps = -1
        elif isinstance(persistent, bool):
            ps = 3 if persistent else 0
        else:
            ps = int(persistent)
        self._check(lib().erm_group_set_tuning(self.h, int(max_batch), ps), "group_set_tuning")

    def set_graph(self, on: bool):
        """Replay the batch sequence from a captured HIP graph when the
        shape repeats (the first call of a shape runs eagerly and captures)."""
        self._check(lib().erm_group_set_graph(self.h, int(bool(on))), "group_set_graph")

    def stage_times(self):
        out = np.zeros(len(STAGES))
        lib().erm_group_stage_times(self.h, _p(out))
        return dict(zip(STAGES, out))


def hash_double(v):
    a = _f64(v).ravel()
    return lib().erm_hash_double(_p(a), len(a))


class ErmGPU:
    """One device context over an ErmSystem."""

    def __init__(self, system: ErmSystem, precision: str = "fp64", max_batch: int = 0,
                 device: int = 0, pme_path: str = "auto"):
        s = system
        self.system = s
        self.precision = precision
        P = ErmParams()
        P.slttype, P.estype, P.ljswitch, P.splodr = s.slttype, s.estype, s.ljswitch, s.splodr
        P.ms[:] = list(s.ms)
        P.corrcal, P.selfcal, P.wgtslf = s.corrcal, s.selfcal, s.wgtslf
        P.force_calculation = s.force_calculation
        P.elecut, P.lwljcut, P.upljcut = s.elecut, s.lwljcut, s.upljcut
        P.screen, P.temp, P.block_threshold = s.screen, s.temp, s.block_threshold
        P.precision = ERM_FP32 if precision == "fp32" else ERM_FP64
        P.max_batch = max_batch
        P.device = device
        P.pme_path = {"auto": 0, "fused": 1, "unfused": 2, "register": 3}[pme_path]
        P.cltype, P.boxshp = s.cltype, s.boxshp
        h = C.c_void_p()
        st = lib().erm_create(C.byref(P), C.byref(h))
        if st:
            raise ErmError(st, "erm_create")
        self.h = h
        self._keep = dict(mol_begin=_i32(s.mol_begin), charge=_f64(s.charge), ljtype=_i32(s.ljtype),
                          lj_eps=_f64(s.lj_eps), lj_sig2=_f64(s.lj_sig2), sluvid=_i32(s.sluvid),
                          uvspec=_i32(s.uvspec), mol_charge=_f64(s.mol_charge),
                          sitemass=_f64(s.sitemass),
                          bf=None if s.bfcoord is None else _f64(s.bfcoord))
        k = self._keep
        self._check(lib().erm_set_topology(h, s.natom, s.nmol, _p(k["mol_begin"], _i32p), _p(k["charge"]),
                                           _p(k["ljtype"], _i32p), s.ntype, _p(k["lj_eps"]),
                                           _p(k["lj_sig2"]), _p(k["sluvid"], _i32p),
                                           _p(k["uvspec"], _i32p), _p(k["mol_charge"]),
                                           _p(k["sitemass"]), _p(k["bf"])), "set_topology")
        edges, uvmax, uvsoft = [], [], []
        for hh in s.hists:
            e, pemax, pesoft = build_edges(hh)
            edges.append(e)
            uvmax.append(pemax)
            uvsoft.append(pesoft)
        self.uvcrd = _f64(np.concatenate(edges))
        self.uvmax = _i32(uvmax)
        self.uvsoft = _i32(uvsoft)
        self.escrd, self.esmax = None, 0
        if s.selfcal:
            e, self.esmax, _ = build_edges(s.self_hist, self_energy=True)
            self.escrd = _f64(e)
        self._check(lib().erm_set_bins(h, s.numslv, _p(self.uvmax, _i32p), _p(self.uvsoft, _i32p),
                                       _p(self.uvcrd), self.esmax, _p(self.escrd)), "set_bins")
        sv, em = C.c_int32(), C.c_int32()
        lib().erm_sizes(h, C.byref(sv), C.byref(em))
        self.slvmax, self.ermax = sv.value, em.value
        self.test_mol = s.nmol - 1

    # ------------------------------------------------------------------
    def _check(self, st, where):
        if st:
            msg = lib().erm_last_error(self.h)
            raise ErmError(st, f"{where}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            lib().erm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------------
    def engclear(self):
        self._check(lib().erm_engclear(self.h), "engclear")

    def set_frame(self, sitepos, cell, weight=1.0):
        pos = _f64(sitepos)
        cm = _f64(np.asarray(cell).T.copy())
        self._check(lib().erm_set_frame(self.h, _p(pos), _p(cm), float(weight)), "set_frame")

    def set_frame_device(self, dev_ptr: int, cell, weight=1.0):
        cm = _f64(np.asarray(cell).T.copy())
        self._check(lib().erm_set_frame_device(self.h, C.c_void_p(dev_ptr), _p(cm), float(weight)),
                    "set_frame_device")

    def get_uv_energy(self, solute, solute_pos=None, n=None):
        if solute_pos is not None:
            sp = _f64(solute_pos)
            n = sp.size // (3 * self.system.nsite(solute)) if n is None else n
        else:
            sp, n = None, (1 if n is None else n)
        out = np.zeros((n, self.slvmax + 1))
        self._check(lib().erm_get_uv_energy(self.h, int(solute), n, _p(sp), _p(out)), "get_uv_energy")
        return out

    def realcal_proc(self, solute, solute_pos=None):
        out = np.zeros(self.slvmax)
        sp = None if solute_pos is None else _f64(solute_pos)
        self._check(lib().erm_realcal_proc(self.h, int(solute), _p(sp), _p(out)), "realcal_proc")
        return out

    def realcal_self(self, solute, solute_pos=None):
        v = C.c_double()
        sp = None if solute_pos is None else _f64(solute_pos)
        self._check(lib().erm_realcal_self(self.h, int(solute), _p(sp), C.byref(v)), "realcal_self")
        return v.value

    def recpcal(self, solute, solute_pos=None):
        out = np.zeros(self.slvmax)
        v = C.c_double()
        sp = None if solute_pos is None else _f64(solute_pos)
        self._check(lib().erm_recpcal(self.h, int(solute), _p(sp), _p(out), C.byref(v)), "recpcal")
        return out, v.value

    def engconst(self, solutes=None, solute_pos=None, weights=None, want_uv=False):
        """Hot loop of engconst: energies + histogram for n configurations."""
        if solute_pos is not None:
            sp = _f64(solute_pos)
            n = sp.size // (3 * self.system.nsite(self.test_mol))
        else:
            sp = None
        sl = None if solutes is None else _i32(solutes)
        if sl is not None:
            n = len(sl)
        w = None if weights is None else _f64(weights)
        out = np.zeros((n, self.slvmax + 1)) if want_uv else None
        self._check(lib().erm_engconst(self.h, n, _p(sl, _i32p), _p(sp), _p(w), _p(out)), "engconst")
        return out

    def insert_random(self, n, seed, first_index=0, want_uv=False):
        out = np.zeros((n, self.slvmax + 1)) if want_uv else None
        self._check(lib().erm_insert_random(self.h, int(n), C.c_uint64(seed), C.c_uint64(first_index),
                                            _p(out)), "insert_random")
        return out

    def insert_flexible(self, confs, seed, first_index=0, weights=None, want_uv=False):
        """Random placement of the given body configurations of a flexible
        test particle (slttype 3) with their weights (SltWght)."""
        cf = _f64(confs)
        n = cf.size // (3 * self.system.nsite(self.test_mol))
        w = None if weights is None else _f64(weights)
        out = np.zeros((n, self.slvmax + 1)) if want_uv else None
        self._check(lib().erm_insert_flexible(self.h, n, _p(cf), _p(w), C.c_uint64(seed), C.c_uint64(first_index),
                                              _p(out)), "insert_flexible")
        return out

    def update_histogram(self, uvengy, solutes=None, weights=None):
        u = _f64(uvengy)
        n = u.shape[0]
        sl = None if solutes is None else _i32(solutes)
        w = None if weights is None else _f64(weights)
        self._check(lib().erm_update_histogram(self.h, n, _p(sl, _i32p), _p(w), _p(u)),
                    "update_histogram")

    def update_histogram_static(self, uvengy, solutes=None, weights=None):
        """update_histogram with the library's static scratch (allocates no
        device memory; n limited by the scratch size)."""
        u = _f64(uvengy)
        n = u.shape[0]
        sl = None if solutes is None else _i32(solutes)
        w = None if weights is None else _f64(weights)
        self._check(lib().erm_update_histogram_static(self.h, n, _p(sl, _i32p), _p(w), _p(u)),
                    "update_histogram_static")

    def engstore(self):
        self._check(lib().erm_engstore(self.h), "engstore")

    def fetch(self):
        E, ns = self.ermax, self.system.numslv
        edens = np.zeros(E)
        ecorr = np.zeros((E, E)) if self.system.corrcal else None
        eself = np.zeros(max(self.esmax, 1))
        mn, mx, sl = np.zeros(ns + 1), np.zeros(ns + 1), np.zeros(ns + 1)
        sc = np.zeros(4)
        self._check(lib().erm_fetch(self.h, _p(edens), _p(ecorr), _p(eself), _p(mn), _p(mx), _p(sl),
                                    _p(sc)), "fetch")
        out = dict(edens=edens, minuv=mn, maxuv=mx, slnuv=sl, engnorm=sc[0], engsmpl=sc[1],
                   avslf=sc[2], voffset=sc[3])
        if ecorr is not None:
            out["ecorr"] = ecorr  # [iduv][iduvp]: the column-major ecorr(iduvp, iduv)
        if self.esmax:
            out["eself"] = eself[:self.esmax]
        return out

    def solute_gradient(self, solutes=None, solute_pos=None, n=None, seed=0, first_index=0,
                        weights=None, cmbrule=0, q=None, eps=None, sigma=None):
        """Weighted sums over configurations of the solute-solvent energy u and
        its gradient with respect to the solute's per-site scaled charge, LJ
        epsilon and sigma: (sum w, sum w u, sum w du/dq, sum w du/deps,
        sum w du/dsigma).  soln: ``solutes``; refs: ``solute_pos`` or ``n``
        random poses from (seed, first_index)."""
        ns = self.system.nsite(int(self.system.solutes()[0]))
        sl = None if solutes is None else _i32(solutes)
        sp = None if solute_pos is None else _f64(solute_pos)
        if sl is not None:
            n = len(sl)
        elif sp is not None:
            n = sp.size // (3 * ns)
        if n is None:
            raise ValueError("give solutes, solute_pos or n")
        w = None if weights is None else _f64(weights)
        qq = None if q is None else _f64(q)
        ee = None if eps is None else _f64(eps)
        ss = None if sigma is None else _f64(sigma)
        acc = np.zeros(2 + 3 * ns)
        self._check(lib().erm_solute_gradient(self.h, int(n), _p(sl, _i32p), _p(sp), C.c_uint64(seed),
                                              C.c_uint64(first_index), _p(w), int(cmbrule), _p(qq),
                                              _p(ee), _p(ss), _p(acc)), "solute_gradient")
        return acc[0], acc[1], acc[2:2 + ns], acc[2 + ns:2 + 2 * ns], acc[2 + 2 * ns:]

    def set_refstructure(self, host_atoms, host_crd, host_weight, slt_crd, slt_weight):
        """The reference structure of the REFSTR insertion modes
        (load_refstructure's arrays, insertion.F90:623-723): 0-based host
        atoms with reference coordinates and weights, the test particle's
        reference coordinates and weights."""
        ha = _i32(host_atoms)
        self._ref = [ha, _f64(host_crd), _f64(host_weight), _f64(slt_crd), _f64(slt_weight)]
        self._check(lib().erm_set_refstructure(self.h, len(ha), _p(ha, _i32p), *[_p(x) for x in self._ref[1:]]),
                    "set_refstructure")

    def set_intermediate(self, zerodns):
        """The intermediate-state method (extensions/Intermediate/engproc.F90)
        with the bins empty in the reference solvent marked; None: off."""
        self._zerodns = None if zerodns is None else _i32(zerodns)
        self._check(lib().erm_set_intermediate(self.h, _p(self._zerodns, _i32p)), "set_intermediate")

    def fetch_intermediate(self):
        """(edintm, uvintm, norm_intm, soln_intm, soln_total)."""
        ed = np.zeros(self.ermax)
        uv = np.zeros(self.system.numslv + 1)
        sc = np.zeros(3)
        self._check(lib().erm_fetch_intermediate(self.h, _p(ed), _p(uv), _p(sc)), "fetch_intermediate")
        return ed, uv, sc[0], sc[1], sc[2]

    def set_insertion(self, insorigin, insposition, insorient=0, lwreg=0.0, upreg=0.0, hostspec=()):
        """Placement of device-generated poses (instslt, insertion.F90:103-366):
        the reference's insorigin / insposition / insorient numbering, region
        bounds and the host species (uvspec numbering) of the aggregate
        centre."""
        hs = _i32(list(hostspec)) if len(hostspec) else None
        self._check(lib().erm_set_insertion(self.h, int(insorigin), int(insposition), int(insorient),
                                            float(lwreg), float(upreg), 0 if hs is None else len(hs),
                                            _p(hs, _i32p)), "set_insertion")

    def set_tuning(self, max_batch=0, pme_path="auto", use_graph=False, slab_planes=0, group_planes=0,
                   slab_threads=0, column_threads=0, fold_y=None):
        """fold_y: True / False moves the inverse y transform into the x slab
        or not; None keeps the current setting."""
        pp = {"auto": 0, "fused": 1, "unfused": 2, "register": 3}[pme_path]
        fy = -1 if fold_y is None else int(bool(fold_y))
        self._check(lib().erm_set_tuning(self.h, int(max_batch), pp, int(bool(use_graph)), int(slab_planes),
                                         int(group_planes), int(slab_threads), int(column_threads), fy),
                    "set_tuning")

    def tuning(self):
        out = np.zeros(10, dtype=np.int32)
        self._check(lib().erm_get_tuning(self.h, _p(out, _i32p)), "get_tuning")
        names = {0: "auto", 1: "fused", 2: "unfused", 3: "register"}
        return dict(max_batch=int(out[0]), pme_path=names[int(out[1])], use_graph=bool(out[2]),
                    slab_threads=int(out[3]), slab_planes=int(out[4]), group_planes=int(out[5]),
                    slab_cap=int(out[6]), group_cap=int(out[7]), column_threads=int(out[8]),
                    fold_y=bool(out[9]))

    def set_pair_path(self, path):
        """'tiles' (compacted per-wave pair lists) or 'molecule' (one thread
        per candidate molecule)."""
        self._check(lib().erm_set_pair_path(self.h, {"tiles": 0, "molecule": 1}[path]), "set_pair_path")

    def pair_path(self):
        return {0: "tiles", 1: "molecule"}[lib().erm_get_pair_path(self.h)]

    def autotune(self, n_trial=512, seed=12345):
        """Times batch sizes, PME paths and graph replay on the current frame;
        keeps the fastest (accumulators untouched).  Returns (tuning, ms)."""
        ms = np.zeros(1)
        self._check(lib().erm_autotune(self.h, int(n_trial), C.c_uint64(seed), _p(ms)), "autotune")
        return self.tuning(), float(ms[0])

    def stage_times(self):
        out = np.zeros(len(STAGES))
        lib().erm_stage_times(self.h, _p(out))
        return dict(zip(STAGES, out))
