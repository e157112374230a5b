#!/usr/bin/env python3
"""Benchmark of the plain XPBD step on one MI355X.

Workload (one "step" = one frame = solver.step(1/60, 8, 1), the reference's
timed "physics ms" call): 1,048,576 particles (128 x 128 x 64 lattice,
r = 0.005 m) falling onto a watertight 81,920-triangle icosphere plus the
ground plane, 60 frames.  Inputs are resident on the GPU before timing starts.
The K timed frames are the last K frames of those 60 (frames 61-K..60, the
densest, most expensive part of the fall; untimed frames advance the pile to
the start of the window, the last W of them as warm-up); for K >= 60 the
state is reset after the warm-up and frames 1..K are timed.  Per-frame times
of frames 1-60 are reported beside it.

Prints ONE JSON line: particle-substeps/s (whole job), the dominant kernel's
achieved HBM bandwidth against the 8 TB/s peak (per-kernel durations from HIP
events recorded on the solver's stream over a second, eager pass of the same
frames; algorithmic bytes as the reference-layout model counts them), and the
oracle's CPU rate on a bounded sample, on all host threads and on one.  More
objects cover the other measured rows: "batched" (256 scenes of 1,024-65,536
particles in one batched solver vs 256 plain solvers stepped in sequence,
same frames, final states compared), "adjoint" (differentiable rollout of
32,768 particles over 64 substeps with checkpoints every 8: forward and
reverse times and their ratio), "host_io" (get_positions at 1M) and
"animated_collider" (update_mesh of the 81,920-triangle icosphere, then a
frame).
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


def survey_bytes(stage, N, C, Cp, Cm, U, U_prev, mesh_bytes):
    """Algorithmic HBM bytes of ONE launch of a stage in the reference-layout
    model (vec4 state, f32 scalars, u32 keys and ids, 44-byte contact
    records; every array touched once per stage, neighbour re-reads counted
    as cache hits), one iteration per substep.  Stages of this build that fuse
    several reference stages count the sum of theirs; the radix sort is the
    model's fixed 36 bytes per particle for the whole sort of a substep."""
    if stage == "integrate":        # integrate + cell index + clear of the last rows
        return (124 + 24) * N + 8 * U_prev
    if stage == "sort":             # every pass of a substep together
        return 36 * N
    if stage == "cell_ranges":      # grid offsets
        return 4 * N + 8 * U
    if stage == "shape_solve":      # the mesh contacts' shape solve
        return 112 * Cm
    if stage == "pp_solve_apply":   # particle-particle solve + apply, and the plane contacts' shape solve
        return (72 + 80) * N + 8 * U + 112 * Cp
    if stage == "mesh_contacts":    # mesh contacts (the mesh tables once per frame, i.e. an eighth per
        # substep) + plane contacts
        return 24 * N + 44 * Cm + mesh_bytes / SUBSTEPS + 24 * N + 44 * Cp
    return None


def stage_bytes(stage, N, C, Cm, U, tiles, mesh_bytes, passes, bins=128, keyed=(SUBSTEPS - 1) / SUBSTEPS):
    Cp = C - Cm
    """This build's own count of the HBM bytes of ONE launch of a stage:
    every array the kernel must touch in its actual layout, counted once
    (neighbour re-reads are cache hits).  The state lives in sorted slot
    order; the sort runs `passes` LSD passes of `bins` digit values over
    `tiles` tiles.  `keyed`: the share of particle-pair solves and shape-delta
    applies that also key the next substep (k_integrate runs once per frame):
    the ids read and the (key, slot) pairs written by the first, the cell
    rows reset from their keys by the second."""
    hist = tiles * bins * 4
    if stage == "integrate":      # P,V,flags,ids in; (key, slot) pair out;
        # reset of the last substep's cell rows (their keys, zeroed rows)
        return N * (16 + 16 + 4 + 4 + 8) + 4 * N + 8 * U
    if stage == "sort_upsweep":   # pass 0 reads the pairs, the others the keys
        return (8 * N + 4 * N * (passes - 1)) / passes + hist
    if stage == "sort_binscan":
        return 2 * hist
    if stage == "sort_scatter":   # keys and slots in and out, per pass
        return 16 * N + hist
    if stage == "cell_ranges":    # sorted keys, new rows, key record; the state's move into the
        # sorted order: slot, P,V,flags,id gathered, integrated P,V, flags, id and rank stored
        return 4 * N + 8 * U + 4 * N + N * (4 + 32 + 4 + 4 + 32 + 4 + 4 + 4)
    if stage == "shape_solve":    # mesh contacts: record, substep-start P,V and flags, delta RMW, owner;
        # plane contacts: shape and slot read, skipped
        return Cm * (56 + 4 + 32 + 4 + 24 + 4) + Cp * 8
    if stage == "pp_solve_apply":  # vals, flags, Xs/Vs, x0, delta, P/V out; cell ranges; pair sums
        # kept for particles with a plane contact; the next substep's keys
        return N * (4 + 4 + 32 + 16 + 12 + 32) + 8 * U + 12 * Cp + keyed * N * (4 + 8)
    if stage == "shape_apply":     # per contact: slot, new slot, slot store, owner; per owner (the
        # mesh contacts): owner clear, shape delta rw, delta rw, Xs/Vs, slot + x0, flags, pair sum, P/V out;
        # the next substep's cell-row reset
        return C * 16 + Cm * (4 + 24 + 24 + 32 + 4 + 16 + 4 + 12 + 32) + keyed * (4 * N + 8 * U)
    if stage == "mesh_contacts":   # particles in, mesh and plane contacts out, mesh tables
        return N * (16 + 4) + 56 * C + mesh_bytes
    return 0


# profiler stage -> kernels it brackets (for the measured-traffic lookup)
STAGE_KERNELS = {
    "mesh_contacts": ["k_mesh_enqueue", "k_mesh_walk"],
    "integrate": ["k_integrate"],
    "sort_upsweep": ["k_sort_upsweep"],
    "sort_binscan": ["k_sort_binscan"],
    "sort_scatter": ["k_sort_scatter"],
    "cell_ranges": ["k_fill_zero", "k_clear_ranges", "k_cell_ranges", "k_cell_ranges_gather", "k_remap_contacts"],
    "shape_solve": ["k_shape_solve"],
    "pp_solve_apply": ["k_pp_apply"],
    "shape_apply": ["k_shape_apply"],
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_step_pile1M_icosphere.txt")


READ_COUNTERS = ["TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_32B_sum", "TCC_BUBBLE_sum", "TCC_EA0_RDREQ_DRAM_sum"]
WRITE_COUNTERS = ["TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum", "TCC_EA0_WRREQ_DRAM_sum"]


def summarize_pmc(read_dir, write_dir, copy_read_dir, copy_write_dir):
    """Per-dispatch HBM traffic from two rocprofv3 --pmc runs of the workload:
    the L2-to-memory read requests by size (32-byte, 128-byte and the rest)
    and the write requests by size (64-byte and the rest), plus the requests
    destined for DRAM.  The bytes of a read request that is neither 32- nor
    128-byte, and of a write request that is not 64-byte, come from the same
    counters over a plain 256 MiB device copy (known bytes).  Prints the
    summary profiles/ keeps and bench.py reads."""
    import csv
    from collections import defaultdict

    def per_kernel(d):
        acc = defaultdict(lambda: defaultdict(float))
        n = defaultdict(set)
        for fn in os.listdir(d):
            if not fn.endswith("counter_collection.csv"):
                continue
            with open(os.path.join(d, fn)) as f:
                for r in csv.DictReader(f):
                    name = r["Kernel_Name"].split("(")[0].replace("void ", "").replace(" ", "").strip()
                    acc[name][r["Counter_Name"]] += float(r["Counter_Value"])
                    n[name].add(r["Dispatch_Id"])
        return {k: ({c: v / len(n[k]) for c, v in cs.items()}, len(n[k])) for k, cs in acc.items()}

    copy_bytes = 256 * 1024 * 1024
    cr, cw = per_kernel(copy_read_dir), per_kernel(copy_write_dir)
    ck = max(cr, key=lambda k: cr[k][0].get("TCC_EA0_RDREQ_sum", 0.0))
    r = cr[ck][0]
    r32, r128 = r["TCC_EA0_RDREQ_32B_sum"], r["TCC_BUBBLE_sum"]
    plain = r["TCC_EA0_RDREQ_sum"] - r32 - r128
    read_plain = (copy_bytes - 32 * r32 - 128 * r128) / plain
    w = cw[ck][0]
    w64 = w["TCC_EA0_WRREQ_64B_sum"]
    w32 = w["TCC_EA0_WRREQ_sum"] - w64
    write64 = (copy_bytes - 32 * w32) / w64
    print(f"copy_kernel {ck}")
    print(f"read_plain_bytes {read_plain:.2f}")
    print(f"write_64b_bytes {write64:.2f}")
    rd, wr = per_kernel(read_dir), per_kernel(write_dir)
    for k in sorted(rd):
        if not k.startswith("xpbd::"):
            continue
        rc, nd = rd[k]
        wc = wr.get(k, ({}, 0))[0]
        r32, r128 = rc.get("TCC_EA0_RDREQ_32B_sum", 0.0), rc.get("TCC_BUBBLE_sum", 0.0)
        rpl = rc.get("TCC_EA0_RDREQ_sum", 0.0) - r32 - r128
        w64 = wc.get("TCC_EA0_WRREQ_64B_sum", 0.0)
        w32 = wc.get("TCC_EA0_WRREQ_sum", 0.0) - w64
        print(f"kernel {k} rd32 {r32:.0f} rdplain {rpl:.0f} rd128 {r128:.0f} "
              f"rddram {rc.get('TCC_EA0_RDREQ_DRAM_sum', 0.0):.0f} wr32 {w32:.0f} wr64 {w64:.0f} "
              f"wrdram {wc.get('TCC_EA0_WRREQ_DRAM_sum', 0.0):.0f} dispatches {nd}")


def measured_traffic(stage):
    """HBM bytes per launch of a profiler stage (reads + writes) from the
    committed request-count summary, or None when it does not cover the
    stage."""
    if not os.path.exists(TRAFFIC_FILE) or stage not in STAGE_KERNELS:
        return None
    rp = w64b = None
    rows = []
    with open(TRAFFIC_FILE) as f:
        for line in f:
            t = line.split()
            if len(t) == 2 and t[0] == "read_plain_bytes":
                rp = float(t[1])
            elif len(t) == 2 and t[0] == "write_64b_bytes":
                w64b = float(t[1])
            elif len(t) >= 18 and t[0] == "kernel":
                v = {t[i]: float(t[i + 1]) for i in range(2, len(t) - 1, 2)}
                rows.append((t[1], v))
    if rp is None or w64b is None:
        return None
    total, wsum, launches, hits = 0.0, 0.0, 0.0, 0
    for name, v in rows:
        if any(("::" + k + "<") in name or name.endswith("::" + k) for k in STAGE_KERNELS[stage]):
            b = 32 * v["rd32"] + rp * v["rdplain"] + 128 * v["rd128"] + 32 * v["wr32"] + w64b * v["wr64"]
            total += b
            wsum += b * v["dispatches"]
            launches += v["dispatches"]
            hits += 1
    if not hits:
        return None
    # a sort stage launches one of several pass variants: their dispatch-weighted mean
    return wsum / launches if stage.startswith("sort_") else total


def reset(s, sc):
    p = sc["particles"]
    s.set_particles(p["q"], p["qd"], p["inv_mass"], p["radii"], p["flags"])


def bench_batched(vkxpbd, scenes, frames, sequential=True):
    """256 scenes (log-uniform 1,024..65,536 particles) advanced by one batched
    solver (frame graphs, autotuned block size) against 256 plain solvers
    stepped one after another, same frames; final states compared
    (sequential=False: the batched solver alone)."""
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
    if not sequential:
        return {"workload": sc["name"], "scenes": int(len(counts)), "particles": total, "frames_timed": frames,
                "value": total * SUBSTEPS * frames / tb, "unit": "particle-substeps/s",
                "ms_per_frame": 1e3 * tb / frames, "pp_block": int(block)}
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
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched solver's frames (no plain step, no sequential comparison): "
                         "the command profiles/ uses for the batched kernels")
    ap.add_argument("--summarize-pmc", nargs=4, metavar=("READ_DIR", "WRITE_DIR", "COPY_READ_DIR", "COPY_WRITE_DIR"),
                    help="print the per-kernel traffic summary from rocprofv3 --pmc output directories")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy (3 times): the known-size kernel that gives "
                         "the bytes of a rocprofv3 read or write request of each size")
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

    if args.batched_only:
        b = bench_batched(vkxpbd, scenes, args.batched_frames, sequential=False)
        print(json.dumps({"metric": "particle-substeps/s (batched HIP)", "value": b["value"], "unit": b["unit"],
                          "n_gpus": 1, "batched": b}), flush=True)
        return

    sc = scenes.pile_1m_icosphere()
    N = sc["particles"]["q"].shape[0]
    verts, tris = sc["meshes"][0]
    frames_total = sc["frames"]   # 60
    K, W = args.steps, args.warmup
    # the timed window: the last K of the 60 frames (W warm-up frames just
    # before it), or frames 1..K after a reset when K >= 60
    first = frames_total - K + 1 if K < frames_total else 1
    lead = max(first - 1 - W, 0)           # untimed frames before the warm-up

    def advance(s, frames):
        for _ in range(frames):
            s.step(DT, SUBSTEPS, 1)

    s = scenes.build_solver(vkxpbd, sc)
    advance(s, lead)
    advance(s, W)
    if first == 1:
        reset(s, sc)
    contacts = []
    t0 = time.perf_counter()
    for _ in range(K):
        s.step(DT, SUBSTEPS, 1)          # synchronous: returns after the GPU finished
        contacts.append(s.last_contact_count)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = N * SUBSTEPS * K / elapsed
    keys = s.get_sorted_cells()
    U = int(np.count_nonzero(keys[1:] != keys[:-1]) + 1)
    # host I/O: one get_positions of all 1,048,576 particles (device scatter
    # into particle order, one DMA to pinned memory), median of five
    io = []
    for _ in range(5):
        a0 = time.perf_counter()
        s.get_positions()
        io.append(time.perf_counter() - a0)
    # animated collider: update_mesh of the icosphere, moved a little each
    # call (vertex upload, device refit, grid rebuilds), median of ten; then
    # one frame, which replays the captured frame graph
    um = []
    for k in range(10):
        v = (verts + np.float32([0.0, 0.0, 1e-4 * (k + 1)])).astype(np.float32)
        a0 = time.perf_counter()
        s.update_mesh(0, v.reshape(-1))
        um.append(time.perf_counter() - a0)
    a0 = time.perf_counter()
    s.step(DT, SUBSTEPS, 1)
    frame_after_update = time.perf_counter() - a0
    del s

    # per-frame times over the whole 60-frame fall (graph replays)
    sf = scenes.build_solver(vkxpbd, sc)
    advance(sf, 2)
    reset(sf, sc)
    ft = []
    for _ in range(frames_total):
        a0 = time.perf_counter()
        sf.step(DT, SUBSTEPS, 1)
        ft.append(time.perf_counter() - a0)
    del sf
    ft = 1e3 * np.asarray(ft)

    out = {
        "metric": "particle-substeps/s @1M (plain HIP step)",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 128 x 128 x 64 lattice over an 81,920-triangle icosphere)",
        "config": {"workload": sc["name"], "particles": N, "triangles": int(tris.shape[0]), "substeps": SUBSTEPS,
                   "iterations": 1, "frames_timed": f"{first}-{first + K - 1}", "dt": DT,
                   "parallelism": "single-gpu"},
        "frames_ms": {"1-20": round(float(ft[:20].mean()), 4), "21-40": round(float(ft[20:40].mean()), 4),
                      "41-60": round(float(ft[40:60].mean()), 4), "1-60": round(float(ft.mean()), 4),
                      "particle_substeps_per_s_1_60": N * SUBSTEPS / (float(ft.mean()) * 1e-3)},
        "host_io": {"get_positions_ms": round(1e3 * float(np.median(io)), 3), "particles": N},
        "animated_collider": {"update_mesh_ms": round(1e3 * float(np.median(um)), 3),
                              "triangles": int(tris.shape[0]),
                              "frame_after_update_ms": round(1e3 * frame_after_update, 3)},
    }

    if not args.no_profile:
        # the same frames again, eagerly on one stream with HIP events around
        # every stage (the mesh stage does not overlap there)
        sp = scenes.build_solver(vkxpbd, sc, enable_profile=True)
        advance(sp, lead + W)
        if first == 1:
            reset(sp, sc)
        sp.profile_reset()
        cm = []
        advance(sp, K)
        shapes = sp.get_contact_shapes()
        cm.append(int(np.count_nonzero(shapes > 0)))
        prof = sp.profile_stages()
        # radix passes per substep (three for 21-bit cell keys or one plane bit
        # per axis, four for two plane bits), as launched
        passes = max(1, round(prof["sort_scatter"][1] / (K * SUBSTEPS)))
        bins = 256 if (sp.wrap_planes and passes == 3) else 128   # 24-bit keys: 8-bit digits; 21 or 27: 7-bit
        del sp
        C = float(np.mean(contacts))
        Cm = float(cm[-1]) * (C / max(contacts[-1], 1))
        Cp = C - Cm
        tiles = (N + 2047) // 2048   # 2,048-key radix tiles
        mesh_bytes = 2 * tris.shape[0] * 32 + tris.shape[0] * 16 + verts.shape[0] * 24
        kern = {}
        for name, (ms, launches) in prof.items():
            if launches == 0 or name == "readback":
                continue
            avg_ms = ms / launches
            own = stage_bytes(name, N, C, Cm, U, tiles, mesh_bytes, passes, bins)
            sv = survey_bytes(name, N, C, Cp, Cm, U, U, mesh_bytes)
            kern[name] = {"avg_us": 1e3 * avg_ms, "launches": int(launches), "own_bytes": own, "survey_bytes": sv}
        # the sort stage per substep (every pass: upsweep, bin scan, scatter)
        sort_ms = sum(prof[k][0] for k in ("sort_upsweep", "sort_binscan", "sort_scatter")) / (K * SUBSTEPS)
        sort_b = survey_bytes("sort", N, C, Cp, Cm, U, U, mesh_bytes)
        # the dominant kernel: the single-kernel stage with the most GPU time
        # (the mesh stage is three kernels, reported per stage below)
        single = [k for k in kern if len(STAGE_KERNELS.get(k, [])) == 1]
        dom = max(single, key=lambda k: kern[k]["avg_us"] * kern[k]["launches"])
        d = kern[dom]
        ach = d["survey_bytes"] / (d["avg_us"] * 1e-6) / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": ach / HBM_PEAK_GBS, "traffic": measured_traffic(dom),
                           "alg_bytes_per_launch": d["survey_bytes"],
                           "bytes_model": "reference layout, 152N + 8U + 112Cp (the plane contacts' shape solve runs inside)",
                           "avg_us": d["avg_us"]}
        out["kernels"] = {}
        for k, v in kern.items():
            e = {"avg_us": round(v["avg_us"], 3), "launches": v["launches"],
                 "own_GBs": round(v["own_bytes"] / (v["avg_us"] * 1e-6) / 1e9, 1)}
            if v["survey_bytes"] is not None:
                e["survey_GBs"] = round(v["survey_bytes"] / (v["avg_us"] * 1e-6) / 1e9, 1)
                e["survey_frac_of_peak"] = round(e["survey_GBs"] / HBM_PEAK_GBS, 4)
            tr = measured_traffic(k)
            if tr is not None and v["avg_us"] > 0:   # measured HBM bytes (rocprofv3 --pmc summary) per launch
                e["traffic_MB"] = round(tr / 1e6, 2)
                e["traffic_GBs"] = round(tr / (v["avg_us"] * 1e-6) / 1e9, 1)
                e["traffic_frac_of_peak"] = round(tr / (v["avg_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
            out["kernels"][k] = e
        out["sort_stage"] = {"us_per_substep": round(1e3 * sort_ms, 3), "launches_per_substep": 3 * passes,
                             "survey_bytes": sort_b, "survey_GBs": round(sort_b / (sort_ms * 1e-3) / 1e9, 1),
                             "survey_frac_of_peak": round(sort_b / (sort_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        # whole-step roofline with the reference-layout byte model per substep,
        # N(376 + 24m) + 156C + 16U + 8U' (mean contacts, final occupied cells)
        b_sub = N * (376 + 24) + 156 * C + 16 * U + 8 * U
        out["step_roofline"] = {"bytes_per_substep": b_sub, "achieved_GBs": b_sub * value / N / 1e9,
                                "frac": b_sub * value / N / 1e9 / HBM_PEAK_GBS}
        out["gpu_ms_per_frame_profiled"] = sum(v[0] for v in prof.values()) / K
        out["contacts_mean"] = C
        out["occupied_cells"] = U

    if not args.no_batched:
        out["batched"] = bench_batched(vkxpbd, scenes, args.batched_frames)
    if not args.no_adjoint:
        out["adjoint"] = bench_adjoint(vkxpbd, scenes)

    if not args.no_cpu_baseline:
        # the oracle (fp32 restatement) on the host's cores: frames 1..F of
        # the same workload from the initial state, on every host thread it
        # uses (at most 16) and on one
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        cfg = scenes.solver_config(vkxpbd, sc["r"])
        p = sc["particles"]
        res = {}
        for threads, frames in (("all", 8), ("1", 3)):
            if threads == "1":
                os.environ["ORC_THREADS"] = "1"
            else:
                os.environ.pop("ORC_THREADS", None)
            o = oracle_lib.OracleSolver()
            o.init(cfg, N)
            o.add_mesh(verts, tris)
            o.finalize_meshes()
            o.set_particles(p["q"], p["qd"], p["inv_mass"], p["radii"], p["flags"])
            c0 = time.perf_counter()
            for _ in range(frames):
                o.step(DT, SUBSTEPS, 1)
            c1 = time.perf_counter()
            res[threads] = (N * SUBSTEPS * frames / (c1 - c0), frames, c1 - c0)
            del o
        os.environ.pop("ORC_THREADS", None)
        cores = min(os.cpu_count() or 1, 16)
        v_all, f_all, s_all = res["all"]
        v_one, f_one, s_one = res["1"]
        out["cpu_baseline"] = {"value": v_all, "unit": "particle-substeps/s", "cores": cores, "kind": "port",
                               "sample": f"frames 1-{f_all} of the same workload ({SUBSTEPS * f_all} substeps from "
                                         f"the initial state), fp32 CPU restatement on {cores} threads",
                               "seconds": s_all,
                               "single_thread": {"value": v_one, "cores": 1, "seconds": s_one,
                                                 "sample": f"frames 1-{f_one}, one thread"}}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
