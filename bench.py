"""Benchmark: solute-solvent mol-pair energies/s and insertions/s of the
reference-solvent insertion workload on one MI355X.

Workload (``refs20``): 4000 rigid 3-site waters in a 49.3 A periodic box;
per frame, 1000 insertions of a rigid 20-site test particle at uniform random
positions and orientations (device-generated, counter-based); real-space
cutoff 12 A, Ewald alpha 0.35, PME 48^3 with order-4 B-splines; 600-bin
energy coordinate; per-insertion histograms accumulated into the 600 x 600
correlation matrix on FP64 MFMA.  A step is one frame (1000 insertions x
4000 molecules = 4e6 mol-pair energies).  Frames are synthetic and uploaded
to HBM before the timed region.

Prints one JSON line (rank 0).  ``roofline`` refers to the dominant kernel,
timed with HIP events on the library's stream over the timed region;
``cpu_baseline`` is the oracle (the FP64 CPU restatement) on a bounded sample
of the same workload on this host.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8000.0          # GB/s, MI355X
FP64_PEAK = 78.6           # TFLOP/s dense FP64 (vector and MFMA)
FP32_PEAK = 157.3          # TFLOP/s dense FP32


def stage_bytes(sy, prec_bytes, n_ins, slvmax, ermax):
    """Algorithmic HBM bytes per launch (n_ins configurations) of the
    register-path stages: Hk = the half-spectrum grid H*K2*K3 of complex T
    per configuration (written by the z columns, read and rewritten by the y
    transform, read by the x transform + gather, which also writes and
    re-reads the per-site sums and writes the per-molecule energies; the
    solvent's site data is read once per launch)."""
    K1, K2, K3 = sy.system.ms
    H = K1 // 2 + 1
    T = prec_bytes
    nsites = 3 * slvmax
    p = sy.system.splodr
    hk = 2 * T * H * K2 * K3
    return {
        "pme_z": n_ins * hk,
        "pme_y": n_ins * 2 * hk,
        "pme_x_gather": n_ins * (hk + 16 * nsites + 8 * slvmax) + nsites * (3 * p * T + 8 + T),
        "bin": n_ins * (8 * slvmax + 8 * ermax),
    }


def measured_traffic(stage, precision, configs):
    """HBM bytes per launch of the x-slab stage from the committed counter
    summary (two rocprofv3 --pmc passes, FETCH_SIZE and WRITE_SIZE, each
    calibrated on a device copy of known size), scaled to `configs`
    configurations per launch; None when no summary covers the case."""
    path = os.path.join(ROOT, "profiles", "refs20_fp64_hbm_traffic_pmc.md")
    if stage != "pme_x_gather" or precision != "fp64" or not os.path.exists(path):
        return None
    kv = {}
    for line in open(path):
        if "=" in line and " " not in line.strip():
            k, v = line.strip().split("=", 1)
            kv[k] = float(v)
    keys = ("xslab_read_bytes_per_launch", "xslab_write_bytes_per_launch",
            "mol_sum_read_bytes_per_launch", "mol_sum_write_bytes_per_launch")
    if not all(k in kv for k in keys) or not kv.get("configs_per_launch"):
        return None
    return sum(kv[k] for k in keys) * configs / kv["configs_per_launch"]


def pair_count(sy, pos, P, rc=12.0):
    """Exact solute-site x solvent-site pairs within the cutoff for the
    configurations P (numpy, minimum image)."""
    s = sy.system
    L = sy.L
    phys = s.physical()
    idx = np.concatenate([np.arange(s.mol_begin[m], s.mol_begin[m + 1]) for m in phys])
    v = pos[idx]
    tot = 0
    for sp in P:
        d = sp[:, None, :] - v[None, :, :]
        d -= L * np.round(d / L)
        tot += int(np.count_nonzero((d * d).sum(-1) <= rc * rc))
    return tot / len(P)


def cpu_baseline(sy, frame0, n_ins_max, seconds, seed):
    import oracle as O
    s = sy.system
    o = O.Oracle(s)
    o.set_frame(frame0, sy.cell)
    m = sy.solute_mol
    t0 = time.perf_counter()
    done = 0
    while done < n_ins_max:
        r, q = O.random_pose(seed, done)
        sp = O.place_solute(s.bfcoord, s.solute_mass(m), sy.cell, r, q)
        u = o.uv_energy(m, sp)
        o.update_histogram(m, u)
        done += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    slv = o.slvmax
    return {"value": done * slv / dt, "unit": "mol-pair energies/s", "cores": 1, "kind": "port",
            "sample": f"oracle (FP64 C restatement, 1 thread) on frame 0 of the same workload: "
                      f"{done} insertions with update_histogram in {dt:.1f} s",
            "insertions_per_s": done / dt}


def mixed_pass(torch, dev, frames=4, per_frame=5000, prec="fp32"):
    """The heterogeneous insertion workload: 7000 waters + 800 10-site
    cosolvent molecules + 100 Na+ + 100 Cl- in a 60 A box (K = 64), per frame
    5000 insertions split over four rigid solutes of 5, 20, 60 and 150 sites,
    one context each (600 bins apiece), run concurrently on four streams
    from one device copy of the frame; FP32.  Each context is autotuned on
    the first frame (untimed)."""
    from ermod_amd import synth
    from ermod_amd.api import ErmGPU, insert_random_multi
    sizes = synth.MIXED_SOLUTE_SITES
    syss = [synth.mixed(ns) for ns in sizes]
    big = max(syss, key=lambda sy: sy.system.natom)
    host = [big.frame(i) for i in range(frames + 1)]
    dpos = [torch.from_numpy(f.reshape(-1)).to(dev) for f in host]
    share = [per_frame // len(sizes)] * len(sizes)
    engines = [ErmGPU(sy.system, prec) for sy in syss]
    try:
        for e, sy in zip(engines, syss):
            e.set_frame_device(dpos[0].data_ptr(), sy.cell)
            e.autotune(n_trial=share[0], seed=99)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for f in range(1, frames + 1):
            for e, sy in zip(engines, syss):
                e.set_frame_device(dpos[f].data_ptr(), sy.cell)
            insert_random_multi(engines, share, [11, 12, 13, 14], [f * s for s in share])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        nmol = syss[0].system.nmol - 1
        nins = frames * sum(share)
        return {"workload": "7000 waters + 800 cosolvent + 100 Na+ + 100 Cl- in 60 A, K 64; solutes of "
                            "5/20/60/150 sites, 1250 insertions each per frame, concurrent contexts",
                "precision": prec, "frames": frames, "mol_pairs_per_s": nins * nmol / dt,
                "insertions_per_s": nins / dt, "ms_per_frame": dt / frames * 1e3}
    finally:
        for e in engines:
            e.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--insertions", type=int, default=1000)
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default batch and PME path)")
    ap.add_argument("--no-graph", action="store_true", help="skip the graph-replay pass")
    ap.add_argument("--slab-planes", type=int, default=0, help="fixed slab depth (with --batch)")
    ap.add_argument("--group-planes", type=int, default=0, help="fixed planes per x group (with --batch)")
    ap.add_argument("--slab-threads", type=int, default=0, help="fixed x-slab workgroup size (with --batch)")
    ap.add_argument("--column-threads", type=int, default=0, help="fixed z-column workgroup size (with --batch)")
    ap.add_argument("--no-variants", action="store_true", help="skip the PPPM and bare-Coulomb passes")
    ap.add_argument("--fold-y", type=int, default=-1, choices=[-1, 0, 1],
                    help="inverse y transform inside the x slab (with --batch; -1 keeps the default)")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the heterogeneous pass (four solute sizes in a mixed solvent, concurrently)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="after the run, one device copy of MB megabytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs (frame sharding with "
                                   "an all-reduce of the accumulators) are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from ermod_amd import synth
    from ermod_amd.api import ErmGPU

    sy = synth.refs20()
    s = sy.system
    nf = args.warmup + args.steps
    nfu = min(nf, 8)   # distinct synthetic frames, cycled
    t_gen = time.perf_counter()
    frames = [sy.frame(i) for i in range(nfu)]
    gen_s = time.perf_counter() - t_gen
    dev = torch.device("cuda:0")
    dframes = [torch.from_numpy(f.reshape(-1)).to(dev) for f in frames]
    torch.cuda.synchronize()

    def run(prec, steps, warmup, graph=False, system=None):
        g = ErmGPU(system or s, prec, max_batch=args.batch)
        tuning = None
        g.set_frame_device(dframes[0].data_ptr(), sy.cell)
        if not args.no_tune and args.batch == 0:
            tuning, _ = g.autotune(n_trial=args.insertions, seed=args.seed + 1)
        else:
            g.set_tuning(max_batch=args.batch, slab_planes=args.slab_planes, group_planes=args.group_planes,
                         fold_y=None if args.fold_y < 0 else bool(args.fold_y),
                         slab_threads=args.slab_threads, column_threads=args.column_threads)
        g.set_tuning(max_batch=0, pme_path=g.tuning()["pme_path"], use_graph=graph)  # keeps batch and slab shape
        for i in range(warmup):
            g.set_frame_device(dframes[i % nfu].data_ptr(), sy.cell)
            g.insert_random(args.insertions, args.seed, first_index=i * args.insertions)
        g.engclear()
        torch.cuda.synchronize()
        st = {k: 0.0 for k in ("prep", "pme_z", "pme_y", "pme_x_gather", "pair", "bin", "corr", "total")}
        t0 = time.perf_counter()
        for i in range(warmup, warmup + steps):
            g.set_frame_device(dframes[i % nfu].data_ptr(), sy.cell)
            g.insert_random(args.insertions, args.seed, first_index=i * args.insertions)
            for k, v in g.stage_times().items():
                st[k] += v
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        acc = g.fetch()
        tun = g.tuning()
        slv, erm = g.slvmax, g.ermax
        g.close()
        return dt, st, acc, tun, slv, erm

    dt, st, acc, tun, slvmax, ermax = run(args.precision, args.steps, args.warmup)
    bmax = tun["max_batch"]
    n_ins = args.steps * args.insertions
    value = n_ins * slvmax / dt
    launches = args.steps * int(np.ceil(args.insertions / bmax))
    T = 8 if args.precision == "fp64" else 4
    sb = stage_bytes(sy, T, bmax, slvmax, ermax)
    # dominant stage over the timed region
    stages = {k: v for k, v in st.items() if k not in ("total",)}
    dom = max(stages, key=stages.get)
    avg_ms = stages[dom] / launches
    kname = {"pme_z": "k_pme_zcol_reg", "pme_y": "k_pme_ycol_reg", "pme_x_gather": "k_pme_xslab + k_mol_sum",
             "bin": "k_bin", "prep": "k_pose + k_solute_prep", "corr": "k_corr"}.get(dom, dom)
    if tun["pme_path"] != "register" and dom.startswith("pme"):
        kname = dom + " (" + tun["pme_path"] + " path)"
    if dom in sb:
        ach = sb[dom] / (avg_ms * 1e-3) / 1e9
        roof = {"kernel": kname, "stage": dom, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK, "unit": "GB/s",
                "frac": ach / HBM_PEAK, "traffic": measured_traffic(dom, args.precision, bmax),
                "avg_launch_ms": avg_ms,
                "bytes_per_launch": sb[dom], "configs_per_launch": bmax}
    else:
        roof = {"kernel": dom, "avg_launch_ms": avg_ms}
    if dom == "pair":
        import oracle as O
        Ps = []
        for i in range(4):
            r, q = O.random_pose(args.seed, i)
            Ps.append(O.place_solute(s.bfcoord, s.solute_mass(sy.solute_mol), sy.cell, r, q))
        npair = pair_count(sy, frames[0], Ps, 12.0)
        flops = 40.0 * npair * bmax
        peak = FP64_PEAK if args.precision == "fp64" else FP32_PEAK
        ach = flops / (avg_ms * 1e-3) / 1e12
        roof = {"kernel": "pair", "bound": "valu", "achieved": ach, "peak": peak, "unit": "TFLOP/s",
                "frac": ach / peak, "traffic": None, "avg_launch_ms": avg_ms,
                "flops_per_launch": flops, "pairs_per_insertion": npair}
    stage_share = {k: round(v / max(st["total"], 1e-9), 4) for k, v in stages.items()}

    out = {"metric": "solute-solvent mol-pair energies/s", "value": value, "unit": "mol-pair energies/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if args.precision == "fp64" else "f32 pair/PME, f64 sums",
           "data": "synthetic",
           "config": {"workload": "refs20: 4000 waters in 49.3 A, rigid 20-site test particle, "
                                  f"{args.insertions} random insertions per frame, cutoff 12 A, "
                                  "alpha 0.35, PME 48^3 p4, 600 bins, correlation on FP64 MFMA",
                      "frames": args.steps, "insertions_per_frame": args.insertions,
                      "batch": bmax, "pme_path": tun["pme_path"], "slab_planes": tun["slab_planes"],
                      "group_planes": tun["group_planes"], "slab_threads": tun["slab_threads"],
                      "column_threads": tun["column_threads"], "fold_y": tun["fold_y"],
                      "precision": args.precision,
                      "tuning": "autotuned" if not args.no_tune and args.batch == 0 else "fixed"},
           "insertions_per_s": n_ins / dt,
           "roofline": roof, "stage_ms": {k: round(v, 3) for k, v in st.items()},
           "stage_share": stage_share, "engsmpl": acc["engsmpl"]}
    if not args.no_graph:
        dtg, _, accg, tung, _, _ = run(args.precision, args.steps, args.warmup, graph=True)
        out["graph_replay"] = {"value": n_ins * slvmax / dtg, "insertions_per_s": n_ins / dtg,
                               "ms_per_step": dtg / args.steps * 1e3, "use_graph": tung["use_graph"],
                               "edens_equal": bool(np.array_equal(accg["edens"], acc["edens"]))}
    if not args.no_fp32 and args.precision == "fp64":
        dt32, st32, _, tun32, _, _ = run("fp32", max(1, args.steps // 2), 1)
        n32 = max(1, args.steps // 2) * args.insertions
        out["fp32"] = {"value": n32 * slvmax / dt32, "insertions_per_s": n32 / dt32,
                       "batch": tun32["max_batch"], "pme_path": tun32["pme_path"],
                       "stage_ms": {k: round(v, 3) for k, v in st32.items()}}
    if not args.no_mixed and args.precision == "fp64":
        out["heterogeneous"] = mixed_pass(torch, dev)
    if not args.no_variants and args.precision == "fp64":
        # the same workload with PPPM and with bare Coulomb (site cutoff at
        # elecut; its cut molecules reach lower energies, hence the wider bins)
        from ermod_amd.system import EL_COULOMB, EL_PPPM
        var = {}
        nst = max(1, args.steps // 4)
        for name, cl in (("pppm", EL_PPPM), ("bare_coulomb", EL_COULOMB)):
            sv = dataclasses.replace(s, cltype=cl)
            if cl == EL_COULOMB:
                sv = dataclasses.replace(sv, hists=[dataclasses.replace(h, ecdmin=-60.0) for h in s.hists])
            dtv, stv, _, tunv, _, _ = run(args.precision, nst, 1, system=sv)
            nv = nst * args.insertions
            var[name] = {"value": nv * slvmax / dtv, "insertions_per_s": nv / dtv, "batch": tunv["max_batch"],
                         "pme_path": tunv["pme_path"] if cl != EL_COULOMB else None,
                         "stage_ms": {k: round(v, 3) for k, v in stv.items()}}
        out["electrostatics"] = var
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sy, frames[0], 100000, args.cpu_seconds, args.seed)
        out["cpu_baseline"]["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
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
