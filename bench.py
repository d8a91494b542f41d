"""Benchmark: solute-solvent mol-pair energies/s and insertions/s of
test-particle insertion on one MI355X.

Headline workload (``mixed5000``, FP32 pair/PME with FP64 sums): a mixed
solvent of 7000 rigid 3-site waters, 800 rigid 10-site cosolvent molecules,
100 Na+ and 100 Cl- (8000 molecules, 29200 sites) in a 60 A periodic box;
per frame 5000 random insertions split evenly over four rigid test particles
of 5, 20, 60 and 150 sites, each with its own 600-bin energy coordinate
(four solvent species x 150 bins) and 600 x 600 correlation matrix on FP64
MFMA; cutoff 12 A, Ewald alpha 0.35, PME 64^3 order 4.  A step is one frame
(5000 insertions x 8000 molecules = 4e7 mol-pair energies).  Frames are
synthetic and resident in HBM before the timed region.

Also measured (secondary keys): the FP32 error of the headline workload
against its FP64 run on the same poses (largest per-molecule |dE| and the
bin-flip fraction), and the ``refs20`` workload (4000 waters in 49.3 A,
20-site test particle, 1000 insertions per frame, PME 48^3, 600 bins) in
FP64 (eager and graph replay) and FP32.

Prints one JSON line.  ``roofline`` is the dominant kernel of the headline
run, timed with HIP events on the library's stream over the timed region;
``stages`` gives every stage against its roofline; ``cpu_baseline`` is the
oracle (the FP64 C restatement) on the same frame and test particles over
the host's cores (OpenMP over insertions), bounded in time.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8000.0          # GB/s, MI355X HBM3E
FP64_MFMA_PEAK = 78.6      # TFLOP/s dense FP64 MFMA
FP64_PEAK = 78.6           # TFLOP/s dense FP64 vector
FP32_PEAK = 157.3          # TFLOP/s dense FP32 vector

STAGE_KERNELS = {"prep": "k_pose + k_solute_prep", "pme_z": "k_pme_zcol_fft", "pme_y": "k_pme_ycol_reg",
                 "pme_x": "k_pme_xslab", "mol_sum": "k_mol_sum", "pair": "k_pair_tile", "self": "k_selfreal + k_u0",
                 "bin": "k_bin + k_scalars + k_edens", "corr": "k_corr"}


def pair_count(sy, pos, P, rc=12.0):
    """Mean number of solute-site x solvent-site pairs within the cutoff over
    the configurations P (numpy, minimum image)."""
    s = sy.system
    L = sy.L
    idx = np.concatenate([np.arange(s.mol_begin[m], s.mol_begin[m + 1]) for m in s.physical()])
    v = pos[idx]
    tot = 0
    for sp in P:
        for a in range(0, len(sp), 16):
            d = sp[a:a + 16, None, :] - v[None, :, :]
            d -= L * np.round(d / L)
            tot += int(np.count_nonzero((d * d).sum(-1) <= rc * rc))
    return tot / len(P)


def sample_poses(sy, n, seed):
    """Test-particle coordinates of insertions 0..n-1 of stream `seed`, placed
    as the device places them (rotation about the centre of mass at a
    uniform position); only used to count pairs for the roofline."""
    s = sy.system
    m = sy.solute_mol
    mass = s.solute_mass(m)
    bf = s.bfcoord - (mass[:, None] * s.bfcoord).sum(0) / mass.sum()
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        out.append(bf @ R.T + rng.random(3) * sy.L)
    return out


def stage_work(sy, T, slvmax, ermax, npair):
    """Algorithmic work per configuration of each stage and per launch (the
    solvent's site data a launch reads once), FP32 (T = 4) or FP64 (T = 8):
    bytes for HBM-bound stages, flops for the pair kernel (40 per pair
    within the cutoff) and the correlation matrix (2 ermax^2, FP64 MFMA)."""
    s = sy.system
    K1, K2, K3 = s.ms
    H = K1 // 2 + 1
    p = s.splodr
    ns = s.nsite(sy.solute_mol)
    nv = int(sum(s.nsite(m) for m in s.physical()))
    hk = 2 * T * H * K2 * K3                          # half spectrum of one configuration
    return {
        # spread along z + z FFT + Green's function: the spectrum it writes,
        # the per-site x / y spectra and z weights it reads
        "pme_z": ("hbm", hk + ns * (2 * T * (H + K2) + (T + 4) * p + T), 0),
        "pme_y": ("hbm", 2 * hk, 0),                   # inverse y transform in place
        # x transform + gather: the potential grid (T K^3) and the per-site
        # sums it writes; solvent site data (base, charge, 3p weights, index)
        # once per launch
        "pme_x": ("hbm", T * K1 * K2 * K3 + 8 * nv, nv * (8 + T * (3 * p + 1))),
        "mol_sum": ("hbm", 8 * nv + 8 * (slvmax + 1), 16 * slvmax),
        "pair": ("valu", 40.0 * npair, 0),
        "bin": ("hbm", 8 * (slvmax + 1) + 8 * ermax, 4 * slvmax),
        "corr": ("mfma", 2.0 * ermax * ermax, 0),
    }


def measure_peaks(torch, dev):
    """Peak rates measured on this device: a 1 GiB device-to-device copy
    (bytes read + written per second), FP32 FMA (v_fma_f32; the packed FP32
    forms are disabled at compile time in this library, so the published
    157.3 TFLOP/s, which counts them, is not reachable by the pair kernel) and FP64 MFMA
    (v_mfma_f64_16x16x4f64), the last two from libermpeak.so's
    microbenchmarks; each the best of five timed launches."""
    import ctypes as C
    lib = C.CDLL(os.path.join(ROOT, "ermod_amd", "libermpeak.so"))
    lib.erm_peak_fma_f32.restype = C.c_double
    lib.erm_peak_mfma_f64.restype = C.c_double
    for f in (lib.erm_peak_fma_f32, lib.erm_peak_mfma_f64):
        f.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]
    st = torch.cuda.current_stream(dev)

    def best(launch):
        ms = []
        work = launch()
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            work = launch()
            e1.record(st)
            e1.synchronize()
            ms.append(e0.elapsed_time(e1))
        if work <= 0:
            raise RuntimeError("peak microbenchmark launch failed")
        return work / (min(ms) * 1e-3)

    n = (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    copy = best(lambda: (b.copy_(a), 2.0 * n * 8)[1]) / 1e9
    del a, b
    blocks = 2048
    o32 = torch.empty(blocks * 256, dtype=torch.float32, device=dev)
    o64 = torch.empty(blocks * 256, dtype=torch.float64, device=dev)
    fma = best(lambda: lib.erm_peak_fma_f32(C.c_void_p(st.cuda_stream), blocks, 50, C.c_void_p(o32.data_ptr())))
    mfma = best(lambda: lib.erm_peak_mfma_f64(C.c_void_p(st.cuda_stream), blocks, 20, C.c_void_p(o64.data_ptr())))
    torch.cuda.synchronize()
    return {"hbm_copy_GBps": copy, "fp32_fma_TFLOPs": fma / 1e12, "fp64_mfma_TFLOPs": mfma / 1e12,
            "method": "best of 5 launches timed with HIP events: 1 GiB float64 device copy (2 x bytes); "
                      "2048 x 256 threads of 8 independent v_fma_f32 chains; 2048 x 4 waves of 4 independent "
                      "v_mfma_f64_16x16x4f64 chains"}


MEASURED = {}   # filled from measure_peaks() when it runs


def roofline_of(stage, bound, work_cfg, work_launch, ms_total, n_cfg, n_launch, T):
    peak, unit = {"hbm": (HBM_PEAK, "GB/s"), "valu": (FP32_PEAK if T == 4 else FP64_PEAK, "TFLOP/s"),
                  "mfma": (FP64_MFMA_PEAK, "TFLOP/s")}[bound]
    total = work_cfg * n_cfg + work_launch * n_launch
    per_launch = total / max(n_launch, 1)
    avg_ms = ms_total / max(n_launch, 1)
    ach = per_launch / (avg_ms * 1e-3) / (1e9 if unit == "GB/s" else 1e12)
    out = {"kernel": STAGE_KERNELS.get(stage, stage), "stage": stage, "bound": bound, "achieved": ach,
           "peak": peak, "unit": unit, "frac": ach / peak, "avg_launch_ms": avg_ms,
           ("bytes_per_launch" if unit == "GB/s" else "flops_per_launch"): per_launch,
           "configs_per_launch": n_cfg / max(n_launch, 1), "launches": n_launch}
    key = {"hbm": "hbm_copy_GBps", "mfma": "fp64_mfma_TFLOPs", "valu": "fp32_fma_TFLOPs" if T == 4 else None}[bound]
    if key and key in MEASURED:
        out["peak_measured"] = MEASURED[key]
        out["frac_of_measured"] = ach / MEASURED[key]
    return out


def measured_traffic(workload, precision, kernel, configs):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes, each scaled by the
    same counter on a device copy of known size), at `configs`
    configurations per launch; None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", f"{workload}_{precision}_hbm_traffic_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.strip().split()
        if len(f) >= 2 and f[0] == "traffic" and f[1] == kernel:
            kv = dict(x.split("=", 1) for x in f[2:] if "=" in x)
            try:
                per_cfg = (float(kv["read_bytes_per_launch"]) + float(kv["write_bytes_per_launch"])) / float(
                    kv["configs_per_launch"])
            except (KeyError, ValueError, ZeroDivisionError):
                return None
            return per_cfg * configs
    return None


# ------------------------------------------------------------ workloads
def mixed_setup(torch, dev, nframes):
    from ermod_amd import synth
    syss = [synth.mixed(ns) for ns in synth.MIXED_SOLUTE_SITES]
    big = syss[-1]
    host = [big.frame(i) for i in range(nframes)]
    dpos = [torch.from_numpy(f.reshape(-1)).to(dev) for f in host]
    torch.cuda.synchronize()
    return syss, host, dpos


# the shape the autotuner settles on for the mixed-solvent workload, used
# when it is skipped (profiling runs)
FIXED_TUNING = dict(max_batch=1024, pme_path="register", slab_planes=6, group_planes=9, slab_threads=512,
                    column_threads=256, fold_y=True)


def run_mixed(torch, syss, dpos, prec, steps, warmup, per_frame, tune=True, schedule="group", graph_steps=0):
    """Timed frames of the mixed-solvent workload.  schedule "group": the four
    solutes' contexts bound in one group, each frame one launch sequence on
    one stream (z-column and pair launches persistent over all solutes or
    one per solute; y/x transforms and gather one launch over all);
    "multi": the four contexts concurrently, one stream each; "sequential":
    one after another (clean per-kernel traces).  Each context is autotuned
    on the first frame (untimed); the group's batch size and launch mode
    are those the warm-up frames time fastest.  Returns the timing, the
    summed stage times and the tunings."""
    from ermod_amd.api import ErmGPU, ErmGroup, insert_random_multi
    nfu = len(dpos)
    share = [per_frame // len(syss)] * len(syss)
    engines = [ErmGPU(sy.system, prec) for sy in syss]
    seeds = [11, 12, 13, 14]
    group = None

    def frame(i):
        if group is not None:
            group.set_frame_device(dpos[i % nfu].data_ptr(), syss[0].cell)
            group.insert_random(share, seeds, [i * s for s in share])
            return
        for e, sy in zip(engines, syss):
            e.set_frame_device(dpos[i % nfu].data_ptr(), sy.cell)
        if schedule == "sequential":
            for k, e in enumerate(engines):
                e.insert_random(share[k], seeds[k], first_index=i * share[k])
        else:
            insert_random_multi(engines, share, seeds, [i * s for s in share])

    group_tuning = None
    try:
        for e, sy in zip(engines, syss):
            e.set_frame_device(dpos[0].data_ptr(), sy.cell)
            if tune:
                e.autotune(n_trial=share[0], seed=99)
            else:
                e.set_tuning(**FIXED_TUNING)
        for e in engines:
            e.set_tuning(max_batch=0, pme_path=e.tuning()["pme_path"], use_graph=False)
        if schedule == "group":
            group = ErmGroup(engines)
            total = sum(share)
            if tune:
                # the group's y/x transforms and gather run in the slab shape
                # of its first member, autotuned alone on a quarter of the
                # insertions: the shapes of the fixed tuning compete with it
                t0m = engines[0].tuning()
                own = dict(slab_planes=t0m["slab_cap"], group_planes=t0m["group_cap"],
                           slab_threads=t0m["slab_threads"], fold_y=t0m["fold_y"])
                # (slab_planes caps the slab depth: 16 gives 8-plane slabs of
                # 16 sites per thread on this workload, 8 gives 4-plane slabs)
                shapes = [own] + [dict(slab_planes=sp, group_planes=gp, slab_threads=512, fold_y=fy)
                                  for sp, gp, fy in ((16, 8, True), (16, 8, False), (8, 8, True), (6, 9, True))]

                def set_shape(shape):
                    engines[0].set_tuning(max_batch=0, pme_path=t0m["pme_path"], use_graph=False, **shape)

                def frames_min(fs, measure):
                    # the smallest of a few frames' measures
                    t = None
                    for f in fs:
                        torch.cuda.synchronize()
                        t0 = time.perf_counter()
                        frame(f)
                        torch.cuda.synchronize()
                        tf = measure(time.perf_counter() - t0)
                        t = tf if t is None else min(t, tf)
                    return t

                # batch size and persistence by wall time per frame, in the
                # first member's own shape
                best = None
                set_shape(own)
                for bmax in ((total + 1) // 2, total):
                    for pers in (0, 1, 2, 3):
                        group.set_tuning(max_batch=bmax, persistent=pers)
                        frame(0)
                        t = frames_min((1, 2), lambda w: w)
                        if best is None or t < best[0]:
                            best = (t, bmax, pers)
                group.set_tuning(max_batch=best[1], persistent=best[2])
                # the slab shape by the device time of the stages it changes
                # (y and x transforms and the gather, HIP events on the
                # stream): frame wall times vary by more than the shapes
                # differ
                xy = lambda w: (lambda st: st["pme_y"] + st["pme_x"])(group.stage_times())
                sbest = None
                shape_ms = []
                for si, shape in enumerate(shapes):
                    set_shape(shape)
                    frame(0)
                    t = frames_min((1, 2, 3), xy)
                    shape_ms.append(round(t, 3))
                    if sbest is None or t < sbest[0]:
                        sbest = (t, si)
                set_shape(shapes[sbest[1]])
                group_tuning = {"max_batch": best[1], "persistent": best[2],
                                "slab_shape": "first member's own" if sbest[1] == 0 else shapes[sbest[1]],
                                "slab_shape_xy_ms": round(sbest[0], 3),
                                "shapes_xy_ms": [dict(shape, xy_ms=t) for shape, t in zip(shapes, shape_ms)]}
            else:
                # the group tuning the warm-up timing settles on for this
                # workload (profiling runs): one batch, both grids persistent
                group_tuning = {"max_batch": total, "persistent": 3}
            group.set_tuning(max_batch=group_tuning["max_batch"], persistent=group_tuning["persistent"])
        for i in range(warmup):
            frame(i)
        for e in engines:
            e.engclear()
        torch.cuda.synchronize()
        st = [dict() for _ in engines]
        # Python's garbage collector paused over the timed frames (as timeit
        # does): a collection of the process's heap stalls the host for tens
        # of milliseconds
        gc.collect()
        gc.disable()
        t0 = time.perf_counter()
        tf = [t0]
        for i in range(warmup, warmup + steps):
            frame(i)
            tf.append(time.perf_counter())
            if group is not None:
                for key, v in group.stage_times().items():
                    st[0][key] = st[0].get(key, 0.0) + v
            else:
                for k, e in enumerate(engines):
                    for key, v in e.stage_times().items():
                        st[k][key] = st[k].get(key, 0.0) + v
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        gc.enable()
        print("[bench] frame wall ms:", " ".join("%.2f" % ((b - a) * 1e3) for a, b in zip(tf, tf[1:])),
              file=sys.stderr, flush=True)
        tun = [e.tuning() for e in engines]
        acc = [e.fetch() for e in engines]
        slvmax, ermax = engines[0].slvmax, engines[0].ermax
        graph_dt = None
        if group is not None and graph_steps > 0:
            # the same frames replayed from the group's captured graph (one
            # eager call captures it, untimed)
            group.set_graph(True)
            frame(warmup + steps)
            torch.cuda.synchronize()
            gc.collect()
            gc.disable()
            t0 = time.perf_counter()
            for i in range(warmup + steps + 1, warmup + steps + 1 + graph_steps):
                frame(i)
            torch.cuda.synchronize()
            graph_dt = time.perf_counter() - t0
            gc.enable()
    finally:
        if group is not None:
            group.close()
        for e in engines:
            e.close()
    return dict(dt=dt, stage=st, tuning=tun, acc=acc, share=share, slvmax=slvmax, ermax=ermax,
                group_tuning=group_tuning, graph_dt=graph_dt, graph_steps=graph_steps)


FP32_EMAX = 1e4   # kcal/mol: the last finite bin edge of every energy coordinate here


def mixed_fp32_error(syss, dpos, per_frame):
    """FP32 against FP64 on one frame's insertions (every solute, the same
    poses): the largest per-molecule |dE| where |E| <= 1e4 kcal/mol (the
    binned range), the largest relative error above it (overflow bin,
    overlapping sites), the largest |dE| per magnitude band, and the
    bin-flip fraction over the four species' edges."""
    from ermod_amd.api import ErmGPU, build_edges
    share = per_frame // len(syss)
    worst, rel, flips, n = 0.0, 0.0, 0, 0
    bands = {}
    per = {}
    edges_lo = [-1.0, 10.0, 100.0, 1e3, 1e4]
    for k, sy in enumerate(syss):
        s = sy.system
        out = {}
        for prec in ("fp64", "fp32"):
            with ErmGPU(s, prec) as g:
                g.set_frame_device(dpos[-1].data_ptr(), sy.cell)
                out[prec] = g.insert_random(share, 31 + k, first_index=0, want_uv=True)[:, 1:]
        spec = s.uvspec[s.physical()]
        err = np.abs(out["fp32"] - out["fp64"])
        mag = np.abs(out["fp64"])
        f = 0
        for t, h in enumerate(s.hists, start=1):
            edges = build_edges(h)[0]
            sel = spec == t
            f += int(np.count_nonzero(np.searchsorted(edges, out["fp32"][:, sel], side="left")
                                      != np.searchsorted(edges, out["fp64"][:, sel], side="left")))
        inr = mag <= FP32_EMAX
        w = float(err[inr].max()) if inr.any() else 0.0
        r = float((err[~inr] / mag[~inr]).max()) if (~inr).any() else 0.0
        per[str(s.nsite(sy.solute_mol))] = {"max_abs_err_binned": w, "max_rel_err_overflow": r, "bin_flips": f,
                                            "overflow_entries": int((~inr).sum())}
        for i, lo in enumerate(edges_lo):
            hi = edges_lo[i + 1] if i + 1 < len(edges_lo) else np.inf
            sel = (mag > lo) & (mag <= hi)
            key = f"|E|<={hi:g}" if np.isfinite(hi) else f"|E|>{lo:g}"
            m = float(err[sel].max()) if sel.any() else 0.0
            if np.isinf(hi) and sel.any():
                m = float((err[sel] / mag[sel]).max())
                key += " (relative)"
            bands[key] = max(bands.get(key, 0.0), m)
        worst, rel = max(worst, w), max(rel, r)
        flips += f
        n += err.size
    return {"max_abs_err_kcal_mol": worst, "bar_kcal_mol": 1e-4, "binned_range_kcal_mol": FP32_EMAX,
            "max_rel_err_above_range": rel, "max_err_by_band": bands, "bin_flips": flips, "entries": n,
            "bin_flip_fraction": flips / n, "insertions": share * len(syss), "per_solute_sites": per}


def cpu_baseline_mixed(syss, host_frame, seconds):
    """The oracle on the same frame and the four test particles (insertion i
    of solute i mod 4), OpenMP over insertions on the host's cores."""
    import oracle as O
    mols = [sy.solute_mol for sy in syss]
    done, dt, nt = O.bench_insertions([sy.system for sy in syss], mols, host_frame, syss[0].cell,
                                      10 ** 7, seconds, seed=7)
    nins = sum(done)
    slv = syss[0].system.physical().size
    return {"value": nins * slv / dt, "unit": "mol-pair energies/s", "cores": nt, "kind": "port",
            "sample": f"oracle (FP64 C restatement), {nt} OpenMP threads over insertions, one context per "
                      f"thread and solute; frame 0 of the mixed-solvent workload, {nins} insertions "
                      f"({'/'.join(str(d) for d in done)} of the 5/20/60/150-site solutes) in {dt:.1f} s",
            "insertions_per_s": nins / dt, "nproc": os.cpu_count()}


def refs20_pass(torch, dev, prec, steps, warmup, insertions, graph=False, tune=True):
    from ermod_amd import synth
    from ermod_amd.api import ErmGPU
    sy = synth.refs20()
    nfu = min(steps + warmup, 8)
    frames = [torch.from_numpy(sy.frame(i).reshape(-1)).to(dev) for i in range(nfu)]
    g = ErmGPU(sy.system, prec)
    try:
        g.set_frame_device(frames[0].data_ptr(), sy.cell)
        if tune:
            g.autotune(n_trial=insertions, seed=20261204)
        g.set_tuning(max_batch=0, pme_path=g.tuning()["pme_path"], use_graph=graph)
        for i in range(warmup):
            g.set_frame_device(frames[i % nfu].data_ptr(), sy.cell)
            g.insert_random(insertions, 20261203, first_index=i * insertions)
        g.engclear()
        # the faster of two windows of `steps` frames: a window that met a
        # host stall (seen in the full bench sequence: an eager window at a
        # third of its speed in isolation) does not stand for the kernels
        dt = None
        for w in range(2):
            torch.cuda.synchronize()
            gc.collect()
            gc.disable()
            t0 = time.perf_counter()
            for i in range(warmup + w * steps, warmup + (w + 1) * steps):
                g.set_frame_device(frames[i % nfu].data_ptr(), sy.cell)
                g.insert_random(insertions, 20261203, first_index=i * insertions)
            torch.cuda.synchronize()
            dw = time.perf_counter() - t0
            gc.enable()
            dt = dw if dt is None else min(dt, dw)
        tun = g.tuning()
        slv = g.slvmax
    finally:
        g.close()
    n = steps * insertions
    return {"value": n * slv / dt, "insertions_per_s": n / dt, "ms_per_step": dt / steps * 1e3,
            "batch": tun["max_batch"], "pme_path": tun["pme_path"], "use_graph": tun["use_graph"],
            "timing": f"faster of two windows of {steps} frames"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed frames")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="fp32", choices=["fp64", "fp32"])
    ap.add_argument("--insertions", type=int, default=5000, help="insertions per frame (split over 4 solutes)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default batch and PME path)")
    ap.add_argument("--no-error", action="store_true", help="skip the FP32-vs-FP64 error pass")
    ap.add_argument("--no-refs20", action="store_true", help="skip the refs20 secondary passes")
    ap.add_argument("--schedule", default="group", choices=["group", "multi", "sequential"],
                    help="group: one launch sequence over the four solutes (default); multi: four contexts on "
                         "four streams; sequential: one context after another")
    ap.add_argument("--no-peaks", action="store_true", help="skip the peak-rate microbenchmarks")
    ap.add_argument("--no-graph", action="store_true", help="skip the group's graph-replay pass")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="after the run, one device copy of MB megabytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; runs on several GPUs (frames sharded over "
                                   "ranks, one all-reduce of the accumulators per division) are not implemented",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    dev = torch.device("cuda:0")
    nfu = min(args.steps + args.warmup, 6) + 1   # distinct frames (cycled) + one for the error pass
    syss, host, dpos = mixed_setup(torch, dev, nfu)
    r = run_mixed(torch, syss, dpos[:-1], args.precision, args.steps, args.warmup, args.insertions,
                  tune=not args.no_tune, schedule=args.schedule,
                  graph_steps=0 if args.no_graph else max(2, args.steps // 2))
    slvmax, ermax = r["slvmax"], r["ermax"]
    n_ins = args.steps * sum(r["share"])
    value = n_ins * slvmax / r["dt"]
    T = 4 if args.precision == "fp32" else 8

    # stage rooflines over the timed region: the work of every solute's
    # configurations against the stage's summed time; a launch is one batch
    # of the group (all solutes), or one batch of a context
    stage_rows = {}
    tot = {}
    grouped = args.schedule == "group"
    nb_group = int(np.ceil(sum(r["share"]) / r["group_tuning"]["max_batch"])) if grouped else 0
    for k, sy in enumerate(syss):
        pos0 = host[0]
        npair = pair_count(sy, pos0, sample_poses(sy, 3, 100 + k))
        work = stage_work(sy, T, slvmax, ermax, npair)
        B = r["tuning"][k]["max_batch"]
        n_cfg = args.steps * r["share"][k]
        n_launch = args.steps * int(np.ceil(r["share"][k] / B))
        folded = r["tuning"][0 if grouped else k]["fold_y"]
        if folded:
            work.pop("pme_y")   # the inverse y transform runs inside the x slab
        for stg, (bound, wc, wl) in work.items():
            a = tot.setdefault(stg, dict(bound=bound, work=0.0, ms=0.0, cfg=0, launch=0, wl=wl))
            a["work"] += wc * n_cfg
            a["cfg"] += n_cfg
            if not grouped:
                a["work"] += wl * n_launch
                a["ms"] += r["stage"][k].get(stg, 0.0)
                a["launch"] += n_launch
    pmask = r["group_tuning"]["persistent"] if grouped else 0
    if grouped:
        # a group batch launches the y/x transforms once, the z columns and
        # pairs once (persistent) or once per solute, the rest once per solute
        nrun = sum(1 for x in r["share"] if x > 0)
        for stg, a in tot.items():
            once = stg in ("pme_y", "pme_x") or (stg == "pme_z" and pmask & 1) or (stg == "pair" and pmask & 2)
            a["launch"] = args.steps * nb_group * (1 if once else nrun)
            a["work"] += a["wl"] * args.steps * nb_group
            a["ms"] = r["stage"][0].get(stg, 0.0)
    if not args.no_peaks:
        MEASURED.update(measure_peaks(torch, dev))
    for stg, a in tot.items():
        if a["ms"] <= 0:
            continue
        stage_rows[stg] = roofline_of(stg, a["bound"], a["work"] / a["cfg"], 0.0, a["ms"], a["cfg"], a["launch"], T)
        stage_rows[stg]["ms_total"] = round(a["ms"], 3)
        if (stg == "pme_z" and pmask & 1) or (stg == "pair" and pmask & 2):
            stage_rows[stg]["kernel"] += "_runs"
    dom = max(stage_rows, key=lambda k: stage_rows[k]["ms_total"])
    roof = dict(stage_rows[dom])
    roof["traffic"] = measured_traffic("mixed5000", args.precision, roof["kernel"], roof["configs_per_launch"])
    for row in stage_rows.values():
        row["traffic"] = measured_traffic("mixed5000", args.precision, row["kernel"], row["configs_per_launch"])
    stage_ms_total = {k: round(sum(st.get(k, 0.0) for st in r["stage"]), 3) for k in r["stage"][0]}

    out = {"metric": "solute-solvent mol-pair energies/s", "value": value, "unit": "mol-pair energies/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["dt"] / args.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 pair/PME, f64 sums" if args.precision == "fp32" else "f64",
           "data": "synthetic",
           "config": {"workload": "mixed5000: 7000 waters + 800 10-site cosolvent + 100 Na+ + 100 Cl- "
                                  "(8000 molecules) in 60 A; per frame 5000 random insertions split over "
                                  "rigid 5/20/60/150-site test particles, 600 bins and a 600x600 correlation "
                                  "matrix (FP64 MFMA) per solute; cutoff 12 A, alpha 0.35, PME 64^3 p4",
                      "frames": args.steps, "insertions_per_frame": sum(r["share"]),
                      "precision": args.precision,
                      "tuning": [{k: t[k] for k in ("max_batch", "pme_path", "slab_planes", "group_planes",
                                                    "slab_threads", "column_threads", "fold_y")}
                                 for t in r["tuning"]],
                      "schedule": {"group": "one group on one stream: y/x transforms and gather one launch over "
                                            "the four solutes; z columns "
                                            + ("one persistent launch over the four" if pmask & 1 else
                                               "one launch per solute") + "; pairs "
                                            + ("one persistent launch over the four" if pmask & 2 else
                                               "one launch per solute"),
                                   "multi": "four contexts, one stream each",
                                   "sequential": "four contexts, one after another"}[args.schedule],
                      "group_tuning": r["group_tuning"]},
           "insertions_per_s": n_ins / r["dt"],
           "roofline": roof, "stages": stage_rows, "stage_ms": stage_ms_total,
           "peaks_measured": dict(MEASURED) or None,
           "group_graph": ({"value": r["graph_steps"] * sum(r["share"]) * slvmax / r["graph_dt"],
                            "ms_per_step": r["graph_dt"] / r["graph_steps"] * 1e3, "steps": r["graph_steps"],
                            "note": "the same workload replayed from the group's captured HIP graph"}
                           if r.get("graph_dt") else None),
           "engsmpl": [a["engsmpl"] for a in r["acc"]]}
    # refs20 before the FP32-error pass (which allocates and frees some 11 GB
    # of FP64 spectra)
    if not args.no_refs20:
        st = max(2, args.steps // 2)
        out["refs20"] = {"workload": "refs20: 4000 waters in 49.3 A, rigid 20-site test particle, 1000 random "
                                     "insertions per frame, PME 48^3 p4, 600 bins",
                         "fp64": refs20_pass(torch, dev, "fp64", st, 2, 1000, tune=not args.no_tune),
                         "fp64_graph": refs20_pass(torch, dev, "fp64", st, 2, 1000, graph=True,
                                                   tune=not args.no_tune),
                         "fp32": refs20_pass(torch, dev, "fp32", st, 2, 1000, tune=not args.no_tune)}
    if not args.no_error:
        out["fp32_error"] = mixed_fp32_error(syss, dpos, args.insertions)
    if not args.no_cpu:
        cb = cpu_baseline_mixed(syss, host[0], args.cpu_seconds)
        cb["gpu_over_cpu"] = value / cb["value"]
        out["cpu_baseline"] = cb
    if args.copy_calibration > 0:
        n = args.copy_calibration * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = 2 * n * 8
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
