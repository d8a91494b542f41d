"""The Fortran host (ermod_amd/fortran/ermgpu_host, built by amdflang over
the ISO_C_BINDING module ermgpu) drives the device pipeline end to end;
its raw accumulators match the oracle bit for bit and its output files
(engref, corref, slfeng per division; weight_refs and uvrange.tt) carry the
reference's formats (engproc.F90:560-656) and the oracle's values."""
import dataclasses
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle as O
from ermod_amd import synth
from ermod_amd.api import build_edges

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ermod_amd", "fortran", "ermgpu_host")


def write_input(path, sy, frames, nins, seed, ndiv=1, s=None, nframes=None):
    s = s or sy.system
    h = s.hists[0]
    nframes = len(frames) if nframes is None else nframes
    with open(path, "wb") as f:
        f.write(struct.pack("<6i", s.natom, s.nmol, s.ntype, nframes, nins, ndiv))
        f.write(struct.pack("<4i3i4i6d5i", s.slttype, s.estype, s.ljswitch, s.splodr, *s.ms, s.corrcal,
                            s.selfcal, s.wgtslf, s.force_calculation, s.elecut, s.lwljcut, s.upljcut,
                            s.screen, s.temp, s.block_threshold, 0, 16, 0, s.cltype, s.boxshp))
        f.write(struct.pack("<9di", h.ecdmin, h.ecfmns, h.ecdcen, h.eccore, h.ecdmax, h.eclbin, h.ecfbin,
                            h.ec0bin, h.finfac, h.pecore))
        # the host reads 3 x (sites of the last molecule) body-frame values
        bf = s.bfcoord if s.bfcoord is not None else np.zeros((s.natom - s.mol_begin[-2], 3))
        for a, t in ((s.mol_begin, np.int32), (s.charge, np.float64), (s.ljtype, np.int32),
                     (s.lj_eps, np.float64), (s.lj_sig2, np.float64), (s.sluvid, np.int32),
                     (s.uvspec, np.int32), (s.mol_charge, np.float64), (s.sitemass, np.float64),
                     (bf, np.float64), (sy.cell.T, np.float64)):
            f.write(np.ascontiguousarray(a, dtype=t).tobytes())
        f.write(struct.pack("<q", seed))
        for fr in frames:
            f.write(np.ascontiguousarray(fr, dtype=np.float64).tobytes())


def test_fortran_host_end_to_end(tmp_path):
    print("\n[gpu case] fortran host tiny_refs", flush=True)
    assert os.path.exists(EXE), "build the Fortran host first (python -m ermod_amd.build)"
    sy = synth.tiny_refs()
    s = sy.system
    frames = [sy.frame(i) for i in range(3)]
    nins, seed = 12, 31
    inp = tmp_path / "in.bin"
    write_input(inp, sy, frames, nins, seed)
    r = subprocess.run([EXE, str(inp), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(tmp_path / "raw.bin", "rb").read()
    slvmax, ermax = struct.unpack("<2i", raw[:8])
    off = 8
    scal = np.frombuffer(raw, np.float64, 4, off); off += 32
    edens = np.frombuffer(raw, np.float64, ermax, off); off += 8 * ermax
    ecorr = np.frombuffer(raw, np.float64, ermax * ermax, off).reshape(ermax, ermax); off += 8 * ermax * ermax
    uv = np.frombuffer(raw, np.float64, (slvmax + 1) * nins, off).reshape(nins, slvmax + 1)
    o = O.Oracle(s)
    m = sy.solute_mol
    for fi, fr in enumerate(frames):
        o.set_frame(fr, sy.cell)
        for i in range(nins):
            rr, q = O.random_pose(seed, fi * nins + i)
            sp = O.place_solute(s.bfcoord, s.solute_mass(m), sy.cell, rr, q)
            u = o.uv_energy(m, sp)
            if fi == 0:
                np.testing.assert_allclose(uv[i], u, rtol=1e-11, atol=1e-9)
            o.update_histogram(m, u)
    a = o.acc()
    assert np.array_equal(edens, a["edens"]) and np.array_equal(ecorr, a["ecorr"])
    assert scal[0] == a["engnorm"] and scal[1] == a["engsmpl"]
    # engref.tt: (g15.7, i5, g25.15) per bin; corref.tt: one record of ermax^2 doubles
    e, pemax, pesoft = build_edges(s.hists[0])
    lines = open(tmp_path / "engref.tt").read().splitlines()
    assert len(lines) == ermax
    for iduv, line in enumerate(lines):
        assert len(line) == 45
        rep, spec, dens = float(line[:15]), int(line[15:20]), float(line[20:])
        assert spec == 1
        assert abs(dens - a["edens"][iduv] / a["engnorm"]) <= 1e-14 * max(1.0, abs(dens))
        if iduv + 1 <= pesoft:
            wr = (e[iduv] + e[iduv + 1]) / 2.0
        elif iduv + 1 < pemax:
            wr = np.sqrt(e[iduv] * e[iduv + 1])
        else:
            wr = e[iduv]
        assert abs(rep - wr) <= 1e-6 * max(1.0, abs(wr))
    prog = open(tmp_path / "progress.tt").read().split()
    assert [int(v) for v in prog] == [1, 2, 3]
    rec = open(tmp_path / "corref.tt", "rb").read()
    n1 = struct.unpack("<i", rec[:4])[0]
    assert n1 == 8 * ermax * ermax and struct.unpack("<i", rec[-4:])[0] == n1
    cor = np.frombuffer(rec[4:-4], np.float64).reshape(ermax, ermax)
    assert np.array_equal(cor, a["ecorr"] / a["engnorm"])


def test_fortran_host_divisions_and_files(tmp_path):
    """Two divisions with the self-energy distribution: per-division files,
    the weights of both divisions and the overall energy range."""
    print("\n[gpu case] fortran host tiny_refs, 2 divisions, selfcal", flush=True)
    sy = synth.tiny_refs()
    s = dataclasses.replace(sy.system, selfcal=1, self_hist=sy.system.hists[0])
    frames = [sy.frame(i) for i in range(4)]
    nins, seed = 10, 77
    inp = tmp_path / "in.bin"
    write_input(inp, sy, frames, nins, seed, ndiv=2, s=s)
    r = subprocess.run([EXE, str(inp), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    o = O.Oracle(s)
    m = sy.solute_mol
    per_div = []
    for div in range(2):
        o.engclear()
        for fi in range(2 * div, 2 * div + 2):
            o.set_frame(frames[fi], sy.cell)
            for i in range(nins):
                rr, q = O.random_pose(seed, fi * nins + i)
                sp = O.place_solute(s.bfcoord, s.solute_mass(m), sy.cell, rr, q)
                o.update_histogram(m, o.uv_energy(m, sp))
        a = o.acc()
        per_div.append({k: (v.copy() if hasattr(v, "copy") else v) for k, v in a.items()})
    for div, a in enumerate(per_div, start=1):
        lines = open(tmp_path / f"engref.{div:02d}").read().splitlines()
        dens = np.array([float(l[20:]) for l in lines])
        np.testing.assert_allclose(dens, a["edens"] / a["engnorm"], rtol=1e-14, atol=1e-300)
        slf = open(tmp_path / f"slfeng.{div:02d}").read().splitlines()
        assert len(slf) == len(a["eself"]) and all(len(l) == 40 for l in slf)
        np.testing.assert_allclose([float(l[15:]) for l in slf], a["eself"] / a["engnorm"], rtol=1e-14,
                                   atol=1e-300)
        rec = open(tmp_path / f"corref.{div:02d}", "rb").read()
        cor = np.frombuffer(rec[4:-4], np.float64).reshape(a["ecorr"].shape)
        assert np.array_equal(cor, a["ecorr"] / a["engnorm"])
    w = open(tmp_path / "weight_refs").read().splitlines()
    assert len(w) == 2 and all(len(l) == 25 for l in w)
    for div, (line, a) in enumerate(zip(w, per_div), start=1):
        assert int(line[:5]) == div
        assert float(line[5:]) == pytest.approx(a["engnorm"] / a["engsmpl"], rel=1e-7)
    uvr = open(tmp_path / "uvrange.tt").read().splitlines()
    assert uvr[0] == " species     minimum        maximum"
    a = per_div[-1]
    for pti, line in enumerate(uvr[1:]):
        assert int(line[:5]) == pti
        mn = float(line[5:20])
        assert mn == pytest.approx(a["minuv"][pti], abs=1e-5)


def test_fortran_host_frames_from_dcd(tmp_path):
    """The same run with the frames read from a DCD file in permuted atom
    order (PermIndex beside it) through erm_traj_*: identical raw
    accumulators and engref.tt."""
    print("\n[gpu case] fortran host tiny_refs from DCD + PermIndex", flush=True)
    sy = synth.tiny_refs()
    s = sy.system
    frames = [sy.frame(i) for i in range(3)]
    nins, seed = 12, 31
    a, b = tmp_path / "stream", tmp_path / "dcd"
    a.mkdir()
    b.mkdir()
    write_input(a / "in.bin", sy, frames, nins, seed)
    r = subprocess.run([EXE, str(a / "in.bin"), str(a)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ntraj = s.natom - s.nsite(sy.solute_mol)
    perm = np.random.default_rng(2).permutation(ntraj)
    synth.write_dcd(b / "t.dcd", [f[:ntraj][perm] for f in frames], [sy.cell] * 3)
    (b / "PermIndex").write_text("".join(f"{i + 1} {p + 1}\n" for i, p in enumerate(perm)))
    write_input(b / "in.bin", sy, [], nins, seed, nframes=3)
    r = subprocess.run([EXE, str(b / "in.bin"), str(b), str(b / "t.dcd")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert (a / "raw.bin").read_bytes() == (b / "raw.bin").read_bytes()
    assert (a / "engref.tt").read_text() == (b / "engref.tt").read_text()


def test_fortran_host_solution(tmp_path):
    """A solution system: every frame's solute through erm_engconst, one
    flcuv.tt line per frame ('(i9, 9999f15.5)', the per-species sum of the
    solute-solvent energy), engsln.tt, weight_soln and aveuv.tt."""
    print("\n[gpu case] fortran host tiny_soln", flush=True)
    sy = synth.tiny_soln()
    s = sy.system
    frames = [sy.frame(i) for i in range(3)]
    inp = tmp_path / "in.bin"
    write_input(inp, sy, frames, 0, 1)
    r = subprocess.run([EXE, str(inp), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    o = O.Oracle(s)
    m = sy.solute_mol
    flc = []
    for fr in frames:
        o.set_frame(fr, sy.cell)
        u = o.uv_energy(m)
        o.update_histogram(m, u)
        flc.append(u[1:].sum())
    lines = open(tmp_path / "flcuv.tt").read().splitlines()
    assert len(lines) == 3 and all(len(l) == 24 for l in lines)
    for f, (line, want) in enumerate(zip(lines, flc), start=1):
        assert int(line[:9]) == f
        assert abs(float(line[9:]) - want) <= 5e-6 + 1e-12 * abs(want)
    a = o.acc()
    raw = open(tmp_path / "raw.bin", "rb").read()
    slvmax, ermax = struct.unpack("<2i", raw[:8])
    edens = np.frombuffer(raw, np.float64, ermax, 8 + 32)
    assert np.array_equal(edens, a["edens"])
    dens = np.array([float(l[20:]) for l in open(tmp_path / "engsln.tt").read().splitlines()])
    np.testing.assert_allclose(dens, a["edens"] / a["engnorm"], rtol=1e-14, atol=1e-300)
    w = open(tmp_path / "weight_soln").read().splitlines()
    assert len(w) == 1 and int(w[0][:5]) == 1
    aveuv = open(tmp_path / "aveuv.tt").read().splitlines()
    assert len(aveuv) == 1 and int(aveuv[0][:5]) == 1
    assert abs(float(aveuv[0][5:]) - a["slnuv"][1] / a["engnorm"]) <= 5e-6


def test_fortran_host_intermediate(tmp_path):
    """The intermediate-state method through the Fortran host: zerodns from
    a reference engref.tt, engint.NN per division, Prob_Result and avint.tt
    against the oracle."""
    print("\n[gpu case] fortran host tiny_soln intermediate state", flush=True)
    sy = synth.tiny_soln()
    s = sy.system
    m = sy.solute_mol
    frames = [sy.frame(i) for i in range(4)]
    o0 = O.Oracle(s)
    occ = 0
    for fr in frames:
        o0.set_frame(fr, sy.cell)
        occ = occ + (o0.update_histogram(m, o0.uv_energy(m)) > 0)
    mask = np.zeros(o0.ermax, np.int32)
    mask[np.nonzero((occ >= 1) & (occ <= 3))[0][0]] = 1
    refs = tmp_path / "refs"
    refs.mkdir()
    with open(refs / "engref.tt", "w") as f:
        for iduv in range(o0.ermax):
            f.write(f"{0.0:15.7g}{1:5d}{0.0 if mask[iduv] else 0.25:25.15g}\n")
    (tmp_path / "intermediate").write_text(str(refs / "engref") + "\n")
    inp = tmp_path / "in.bin"
    write_input(inp, sy, frames, 0, 1, ndiv=2)
    r = subprocess.run([EXE, str(inp), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    o = O.Oracle(s)
    o.set_intermediate(mask)
    avint = []
    for div in range(2):
        o.engclear()
        for fr in frames[2 * div:2 * div + 2]:
            o.set_frame(fr, sy.cell)
            o.update_histogram(m, o.uv_energy(m))
        ed, uv, norm, sint, stot = o.intermediate()
        dens = np.array([float(l[20:]) for l in open(tmp_path / f"engint.{div + 1:02d}").read().splitlines()])
        np.testing.assert_allclose(dens, ed / norm, rtol=1e-14, atol=1e-300)
        avint.append(uv[1] / norm)
    prob = open(tmp_path / "Prob_Result").read().splitlines()
    assert abs(float(prob[0].split("=")[1]) - sint / stot) <= 5e-7
    av = open(tmp_path / "avint.tt").read().splitlines()
    assert len(av) == 2
    for line, want in zip(av, avint):
        assert abs(float(line[5:]) - want) <= 5e-6
