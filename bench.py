#!/usr/bin/env python3
"""Benchmark of the plain XPBD step on one MI355X.

Workload (one "step" = one frame = solver.step(1/60, 8, 1), the reference's
timed "physics ms" call): 1,048,576 particles (128 x 128 x 64 lattice,
r = 0.005 m) falling onto a watertight 81,920-triangle icosphere plus the
ground plane.  Inputs are resident on the GPU before timing starts; warm-up
frames run first and the particle state is then reset to the initial lattice,
so the K timed frames are frames 1..K of the configuration.

Prints ONE JSON line: particle-substeps/s (whole job), the dominant kernel's
achieved HBM bandwidth against the 8 TB/s peak (per-kernel durations from HIP
events recorded on the solver's stream over a second, profiled pass of the
same frames), and the oracle's single-thread CPU rate on a bounded sample.
Two more objects cover the other measured rows: "batched" (256 scenes of
1,024-65,536 particles in one batched solver vs 256 plain solvers stepped in
sequence, same frames, final states compared) and "adjoint" (differentiable
rollout of 32,768 particles over 64 substeps with checkpoints every 8: forward
and reverse times and their ratio).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "newton-xpbd-vulkan_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
SUBSTEPS = 8
DT = 1.0 / 60.0


def stage_bytes(stage, N, C, Cm, U, tiles, mesh_bytes, passes, bins=128):
    """Algorithmic HBM bytes of ONE launch of a stage: every array the kernel
    must touch, counted once (neighbour re-reads are cache hits).  The state
    lives in sorted slot order; the sort runs `passes` LSD passes of `bins`
    digit values over `tiles` tiles."""
    hist = tiles * bins * 4
    if stage == "integrate":      # P,V,flags,ids in; (key, slot) pair out; plane contacts;
        # reset of the last substep's cell rows (their keys, zeroed rows)
        return N * (16 + 16 + 4 + 4 + 8) + 56 * (C - Cm) + 4 * N + 8 * U
    if stage == "sort_upsweep":   # pass 0 reads the pairs, the others the keys
        return (8 * N + 4 * N * (passes - 1)) / passes + hist
    if stage == "sort_binscan":
        return 2 * hist
    if stage == "sort_scatter":   # mean over the passes; the last also integrates and stores P,V, moves flags,id, writes rank
        return (16 * N * passes + (16 + 16 + 4 + 4) * 2 * N + 4 * N) / passes + hist
    if stage == "cell_ranges":    # sorted keys, new rows, key record
        return 4 * N + 8 * U + 4 * N
    if stage == "shape_solve":    # contact record, slot remap, particle P,V gather, delta RMW, owner
        return C * (56 + 8 + 32 + 24 + 4)
    if stage == "pp_solve_apply":  # vals, flags, Xs/Vs, x0, delta, P/V out; cell ranges
        return N * (4 + 4 + 32 + 16 + 12 + 32) + 8 * U
    if stage == "shape_apply":     # per contact: particle slot, owner; per owner (taken as every
        # contact): owner clear, shape delta rw, delta rw, Xs/Vs, slot + x0, P/V out
        return C * (4 + 4 + 4 + 24 + 24 + 32 + 4 + 16 + 32)
    if stage == "mesh_contacts":
        return N * (16 + 4) + 56 * Cm + mesh_bytes
    return 0


# profiler stage -> kernels it brackets (for the measured-traffic lookup)
STAGE_KERNELS = {
    "mesh_contacts": ["k_mesh_enqueue", "k_mesh_walk", "k_mesh_parity"],
    "integrate": ["k_integrate"],
    "sort_upsweep": ["k_sort_upsweep"],
    "sort_binscan": ["k_sort_binscan"],
    "sort_scatter": ["k_sort_scatter"],
    "cell_ranges": ["k_fill_zero", "k_clear_ranges", "k_cell_ranges", "k_remap_contacts"],
    "shape_solve": ["k_shape_solve"],
    "pp_solve_apply": ["k_pp_apply"],
    "shape_apply": ["k_shape_apply"],
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_step_pile1M_icosphere.txt")


def summarize_pmc(fetch_dir, write_dir, cfetch_dir, cwrite_dir):
    """Per-dispatch HBM traffic from two rocprofv3 --pmc runs of the workload
    (FETCH_SIZE, WRITE_SIZE; KiB per dispatch), each scaled by what the same
    counter reads for a plain 256 MiB device copy.  Prints the summary that
    profiles/ keeps and bench.py reads."""
    import csv
    from collections import defaultdict

    def per_kernel(d, counter):
        acc = defaultdict(list)
        with open(os.path.join(d, "pmc_counter_collection.csv")) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] == counter:
                    name = r["Kernel_Name"].split("(")[0].replace("void ", "").strip()
                    acc[name].append(float(r["Counter_Value"]))
        return acc

    copy_kib = 256 * 1024
    cf = per_kernel(cfetch_dir, "FETCH_SIZE")
    cw = per_kernel(cwrite_dir, "WRITE_SIZE")
    ck = max(cf, key=lambda k: sum(cf[k]) / len(cf[k]))
    f_factor = copy_kib / (sum(cf[ck]) / len(cf[ck]))
    w_factor = copy_kib / (sum(cw[ck]) / len(cw[ck]))
    fe = per_kernel(fetch_dir, "FETCH_SIZE")
    wr = per_kernel(write_dir, "WRITE_SIZE")
    print(f"fetch_factor {f_factor:.4f}")
    print(f"write_factor {w_factor:.4f}")
    for k in sorted(fe):
        if not k.startswith("xpbd::"):
            continue
        fk = sum(fe[k]) / len(fe[k])
        wk = sum(wr.get(k, [0.0])) / max(len(wr.get(k, [0.0])), 1)
        print(f"kernel {k.replace(' ', '')} fetch_kib {fk:.1f} write_kib {wk:.1f} dispatches {len(fe[k])}")


def measured_traffic(stage):
    """HBM bytes per launch of a profiler stage from the committed --pmc
    summary, or None when it does not cover the stage."""
    if not os.path.exists(TRAFFIC_FILE) or stage not in STAGE_KERNELS:
        return None
    ff = wf = None
    rows = []
    with open(TRAFFIC_FILE) as f:
        for line in f:
            t = line.split()
            if len(t) == 2 and t[0] == "fetch_factor":
                ff = float(t[1])
            elif len(t) == 2 and t[0] == "write_factor":
                wf = float(t[1])
            elif len(t) >= 8 and t[0] == "kernel":
                rows.append((t[1], float(t[3]), float(t[5]), float(t[7])))
    if ff is None or wf is None:
        return None
    total, wsum, launches, hits = 0.0, 0.0, 0.0, 0
    for name, fk, wk, nd in rows:
        if any(("::" + k + "<") in name or name.endswith("::" + k) for k in STAGE_KERNELS[stage]):
            b = (ff * fk + wf * wk) * 1024.0
            total += b
            wsum += b * nd
            launches += nd
            hits += 1
    if not hits:
        return None
    # a sort stage launches one of several pass variants: their dispatch-weighted mean
    return wsum / launches if stage.startswith("sort_") else total


def reset(s, sc):
    p = sc["particles"]
    s.set_particles(p["q"], p["qd"], p["inv_mass"], p["radii"], p["flags"])


def bench_batched(vkxpbd, scenes, frames):
    """256 scenes (log-uniform 1,024..65,536 particles) advanced by one batched
    solver (frame graphs, autotuned block size) against 256 plain solvers
    stepped one after another, same frames; final states compared."""
    sc = scenes.batch_of_piles()
    counts = sc["counts"]
    total = int(counts.sum())
    b = scenes.build_batch_solver(vkxpbd, sc)
    block = b.autotune(2, DT, SUBSTEPS, 1)
    b.step(DT, SUBSTEPS, 1)                       # capture
    p = sc["particles"]
    b.set_particles(p["q"], p["qd"], p["inv_mass"], p["radii"], p["flags"])
    t0 = time.perf_counter()
    for _ in range(frames):
        b.step(DT, SUBSTEPS, 1)
    tb = time.perf_counter() - t0
    xb = b.get_positions()
    del b
    singles = [scenes.build_solver(vkxpbd, scenes.scene_slice(sc, k)) for k in range(len(counts))]
    for s in singles:                             # capture each solver's frame graph
        s.step(DT, SUBSTEPS, 1)
    for k, s in enumerate(singles):
        q = scenes.scene_slice(sc, k)["particles"]
        s.set_particles(q["q"], q["qd"], q["inv_mass"], q["radii"], q["flags"])
    t0 = time.perf_counter()
    for _ in range(frames):
        for s in singles:
            s.step(DT, SUBSTEPS, 1)
    ts = time.perf_counter() - t0
    xs = np.concatenate([s.get_positions() for s in singles])
    del singles
    return {"workload": sc["name"], "scenes": int(len(counts)), "particles": total, "frames_timed": frames,
            "value": total * SUBSTEPS * frames / tb, "unit": "particle-substeps/s",
            "ms_per_frame": 1e3 * tb / frames, "sequential_ms_per_frame": 1e3 * ts / frames,
            "speedup_vs_sequential": ts / tb, "pp_block": int(block),
            "max_abs_diff_vs_sequential_m": float(np.max(np.abs(xb.astype(np.float64) - xs)))}


def bench_adjoint(vkxpbd, scenes, reps=3):
    """Differentiable rollout of 32,768 particles, 64 substeps of 1/480 s,
    checkpoints every 8: forward and reverse pass times (GPU-resident
    tensors), and their ratio."""
    import torch
    from vkxpbd.autograd import rollout
    sc = scenes.rollout_pile_32768()
    p = sc["particles"]
    s = scenes.build_solver(vkxpbd, sc)
    dev = torch.device("cuda:0")
    x0 = torch.tensor(p["q"], device=dev)
    tgt = torch.tensor(sc["target"], device=dev)
    fw, bw = [], []
    for k in range(reps + 1):
        v0 = torch.tensor(p["qd"], device=dev, requires_grad=True)
        relax = torch.tensor(0.9, device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xs, _ = rollout(s, x0, v0, relax, 1.0 / 480.0, 64, 8)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss = ((xs - tgt) ** 2).sum(dim=1).mean()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if k > 0:
            fw.append(t1 - t0)
            bw.append(t2 - t1)
    f, b = float(np.median(fw)), float(np.median(bw))
    return {"workload": "rollout_32768_64_substeps_checkpoint8", "forward_ms": 1e3 * f, "backward_ms": 1e3 * b,
            "bwd_fwd_ratio": b / f, "loss": float(loss.detach().cpu()),
            "grad_relax": float(relax.grad.cpu()), "grad_v0_max": float(v0.grad.abs().max().cpu())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--batched-frames", type=int, default=10)
    ap.add_argument("--summarize-pmc", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR"),
                    help="print the per-kernel traffic summary from rocprofv3 --pmc output directories")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy (3 times): the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counts to bytes")
    args = ap.parse_args()
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc)
        return
    if args.copy_calibration:
        import torch
        n = 64 * 1024 * 1024
        src = torch.arange(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design, not a run)", file=sys.stderr)
        sys.exit(2)

    import vkxpbd
    from vkxpbd import scenes

    sc = scenes.pile_1m_icosphere()
    N = sc["particles"]["q"].shape[0]
    verts, tris = sc["meshes"][0]
    s = scenes.build_solver(vkxpbd, sc)
    for _ in range(args.warmup):
        s.step(DT, SUBSTEPS, 1)
    reset(s, sc)
    contacts = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.step(DT, SUBSTEPS, 1)          # synchronous: returns after the GPU finished
        contacts.append(s.last_contact_count)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = N * SUBSTEPS * args.steps / elapsed
    keys = s.get_sorted_cells()
    U = int(np.count_nonzero(keys[1:] != keys[:-1]) + 1)
    del s

    out = {
        "metric": "particle-substeps/s @1M (plain HIP step)",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 128 x 128 x 64 lattice over an 81,920-triangle icosphere)",
        "config": {"workload": sc["name"], "particles": N, "triangles": int(tris.shape[0]), "substeps": SUBSTEPS,
                   "iterations": 1, "frames_timed": args.steps, "dt": DT, "parallelism": "single-gpu"},
    }

    if not args.no_profile:
        sp = scenes.build_solver(vkxpbd, sc, enable_profile=True)
        for _ in range(args.warmup):
            sp.step(DT, SUBSTEPS, 1)
        reset(sp, sc)
        sp.profile_reset()
        cm = []
        for _ in range(args.steps):
            sp.step(DT, SUBSTEPS, 1)
        shapes = sp.get_contact_shapes()
        cm.append(int(np.count_nonzero(shapes > 0)))
        prof = sp.profile_stages()
        # radix passes per substep (three for 21-bit cell keys or one plane bit
        # per axis, four for two plane bits), as launched
        passes = max(1, round(prof["sort_scatter"][1] / (args.steps * SUBSTEPS)))
        bins = 256 if (sp.wrap_planes and passes == 3) else 128   # 24-bit keys: 8-bit digits; 21 or 27: 7-bit
        del sp
        C = float(np.mean(contacts))
        Cm = float(cm[-1]) * (C / max(contacts[-1], 1))
        tiles = (N + 2047) // 2048   # 2,048-key radix tiles
        mesh_bytes = 2 * tris.shape[0] * 32 + tris.shape[0] * 16 + verts.shape[0] * 24
        kern = {}
        for name, (ms, launches) in prof.items():
            if launches == 0 or name == "readback":
                continue
            avg_ms = ms / launches
            b = stage_bytes(name, N, C, Cm, U, tiles, mesh_bytes, passes, bins)
            kern[name] = {"avg_us": 1e3 * avg_ms, "launches": int(launches), "alg_bytes": b,
                          "gbs": b / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0}
        # the dominant kernel: the single-kernel stage with the most GPU time
        # (the mesh stage is three kernels, reported per stage below)
        single = [k for k in kern if len(STAGE_KERNELS.get(k, [])) == 1]
        dom = max(single, key=lambda k: kern[k]["avg_us"] * kern[k]["launches"])
        d = kern[dom]
        out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": d["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": d["gbs"] / HBM_PEAK_GBS, "traffic": measured_traffic(dom),
                           "alg_bytes_per_launch": d["alg_bytes"], "avg_us": d["avg_us"]}
        out["kernels"] = {}
        for k, v in kern.items():
            e = {"avg_us": round(v["avg_us"], 3), "launches": v["launches"], "alg_GBs": round(v["gbs"], 1)}
            tr = measured_traffic(k)
            if tr is not None and v["avg_us"] > 0:   # measured HBM bytes (rocprofv3 --pmc summary) per launch
                e["traffic_MB"] = round(tr / 1e6, 2)
                e["traffic_GBs"] = round(tr / (v["avg_us"] * 1e-6) / 1e9, 1)
                e["traffic_frac_of_peak"] = round(tr / (v["avg_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
            out["kernels"][k] = e
        # whole-step roofline with the reference-layout byte model per substep,
        # N(376 + 24m) + 156C + 16U + 8U' (mean contacts, final occupied cells)
        b_sub = N * (376 + 24) + 156 * C + 16 * U + 8 * U
        out["step_roofline"] = {"bytes_per_substep": b_sub, "achieved_GBs": b_sub * value / N / 1e9,
                                "frac": b_sub * value / N / 1e9 / HBM_PEAK_GBS}
        out["gpu_ms_per_frame_profiled"] = sum(v[0] for v in prof.values()) / args.steps
        out["contacts_mean"] = C
        out["occupied_cells"] = U

    if not args.no_batched:
        out["batched"] = bench_batched(vkxpbd, scenes, args.batched_frames)
    if not args.no_adjoint:
        out["adjoint"] = bench_adjoint(vkxpbd, scenes)

    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        cfg = scenes.solver_config(vkxpbd, sc["r"])
        o = oracle_lib.OracleSolver()
        o.init(cfg, N)
        o.add_mesh(verts, tris)
        o.finalize_meshes()
        p = sc["particles"]
        o.set_particles(p["q"], p["qd"], p["inv_mass"], p["radii"], p["flags"])
        cpu_frames = 3
        c0 = time.perf_counter()
        for _ in range(cpu_frames):
            o.step(DT, SUBSTEPS, 1)
        c1 = time.perf_counter()
        out["cpu_baseline"] = {"value": N * SUBSTEPS * cpu_frames / (c1 - c0), "unit": "particle-substeps/s",
                               "cores": 1, "kind": "port",
                               "sample": f"frames 1-{cpu_frames} of the same workload ({8 * cpu_frames} substeps from "
                                         "the initial state), fp32 CPU restatement, 1 thread",
                               "seconds": c1 - c0}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
