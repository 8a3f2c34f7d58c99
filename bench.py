#!/usr/bin/env python3
"""Benchmark: bodies/sec of AABB + broadphase + narrowphase on one MI355X.

Workload: 1,048,576 mixed convex bodies, FP64 input (40 % sphere, 20 % box,
15 % ellipsoid, 15 % capsule, 10 % cylinder; size log-normal(ln 0.01, 0.3);
uniform random rotations; centres uniform in a cube sized for ~4 AABB
neighbours per body).  A step is one full pass (AABB generation, uniform-grid
broadphase, MPR narrowphase, contact compaction) with the scene resident in
HBM; the step ends when the contact count is known on the host.  `value` is
the step recorded once as a HIP graph and replayed (the scene as a one-scene
batch); `scene_path` is the same step through the stage-by-stage scene API,
which reads totals back between stages as the reference does, and provides
the per-stage HIP-event times.

Prints ONE JSON line.  The CPU baseline is the oracle (our C++/OpenMP
restatement of the reference's shipped OpenMP path) timed on a bounded
sample on this host's cores.
"""
import argparse
import gc
import importlib.util
import json
import os
import sys
import time

import numpy as np

# The CPU baseline binds its OpenMP threads to neighbouring cores (the plan
# of record for the reference's OpenMP path); set before any OpenMP runtime
# loads in this process.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "collision-detection_amd")

HBM_PEAK_GBS = 8000.0
FP64_VECTOR_PEAK_TFLOPS = 78.6


def _package():
    name = "collision_detection_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def stage_bytes(s, c):
    """Algorithmic HBM bytes per launch of the single-kernel stages (compulsory reads + writes, packed sizes).

    r = bytes per real of the AABBs (the input real type).  AABB: per shape
    type 1 + body id 4 + local pos 3r + dims 3r + local rot 4r + body pos 3r
    + body rot 4r in, AABB 6r out.  Bin count: AABB 6r in, count 4 out.  Key
    emit: AABB 6r + offset 4 in, 8 per bin entry out.  Pair count pass: the
    (bin, body) entry 8 and the AABB 6r per entry in, a count 4 per entry
    out.  Pair emit pass: the entry's count 4 in, per pair the partner's
    entry 8 in and the pair 8 out.
    """
    n, m, p = s.n, c.bin_entries, c.pairs
    r = s.dtype.itemsize
    return {
        "aabb": n * (1 + 4 + 17 * r) + n * 6 * r,
        "bp_count": n * 6 * r + n * 4,
        "bp_emit": n * (6 * r + 4) + m * 8,
        "bp_pairs": m * (8 + 6 * r) + m * 4,
        "bp_pairemit": m * 4 + p * (8 + 8),
    }


# Kernel behind each single-kernel stage (names as rocprofv3 lists them).
STAGE_KERNEL = {
    "aabb": "k_aabb<double>", "bp_count": "k_count<double>", "bp_emit": "k_emit_keys<double>",
    "bp_pairs": "k_pairs<double>", "bp_pairemit": "k_pair_emit<double>"
}


def read_traffic(stage):
    """HBM bytes per launch of a stage's kernel from the committed rocprofv3 PMC summary, or None.

    The summary (profiles/pmc_traffic_mixed1M.txt) holds key = value lines:
    <stage>.fetch_kib / <stage>.write_kib (FETCH_SIZE / WRITE_SIZE per
    dispatch) and copy.fetch_kib / copy.write_kib / copy.bytes for a plain
    device copy of known size, which scales the counters to bytes.
    """
    path = os.path.join(ROOT, "profiles", "pmc_traffic_mixed1M.txt")
    if not os.path.exists(path):
        return None
    kv = {}
    for line in open(path):
        line = line.split("#", 1)[0].strip()
        if "=" in line:
            k, v = line.split("=", 1)
            try:
                kv[k.strip()] = float(v)
            except ValueError:
                pass
    need = [f"{stage}.fetch_kib", f"{stage}.write_kib", "copy.fetch_kib", "copy.write_kib", "copy.bytes"]
    if not all(k in kv for k in need):
        return None
    f_scale = kv["copy.bytes"] / (kv["copy.fetch_kib"] * 1024.0)
    w_scale = kv["copy.bytes"] / (kv["copy.write_kib"] * 1024.0)
    return int(kv[f"{stage}.fetch_kib"] * 1024.0 * f_scale + kv[f"{stage}.write_kib"] * 1024.0 * w_scale)


def read_resolve_fp64():
    """FP64 rate of the MPR re-solve kernel from the committed rocprofv3 summary, or None.

    profiles/pmc_fp64_resolve_mixed1M.txt holds the kernel's FP64 instruction
    counts per dispatch (SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64, per wave) and
    its average duration from the kernel-trace summary of the same tree.
    Issued FLOP = 64 x (ADD + MUL + TRANS + 2 FMA), inactive lanes included.
    """
    path = os.path.join(ROOT, "profiles", "pmc_fp64_resolve_mixed1M.txt")
    if not os.path.exists(path):
        return None
    kv = {}
    for line in open(path):
        line = line.split("#", 1)[0].strip()
        if "=" in line:
            k, v = line.split("=", 1)
            try:
                kv[k.strip()] = float(v)
            except ValueError:
                pass
    need = ["resolve.add_f64", "resolve.mul_f64", "resolve.fma_f64", "resolve.trans_f64", "resolve.kernel_us"]
    if not all(k in kv for k in need):
        return None
    flop = 64.0 * (kv["resolve.add_f64"] + kv["resolve.mul_f64"] + kv["resolve.trans_f64"] + 2.0 * kv["resolve.fma_f64"])
    tflops = flop / (kv["resolve.kernel_us"] * 1e-6) / 1e12
    return {"bound": "valu_fp64", "kernel": "k_np_resolve<double, double>", "achieved": round(tflops, 2),
            "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tflops / FP64_VECTOR_PEAK_TFLOPS, 4),
            "flop_issued_per_launch": int(flop), "avg_launch_ms": round(kv["resolve.kernel_us"] / 1000.0, 4),
            "source": "profiles/pmc_fp64_resolve_mixed1M.txt (rocprofv3 counters and kernel time)"}


def contact_error(got, ref, scale):
    """Contacts of the GPU against the FP64 oracle's, matched by body-id pair.

    normal_err: max-norm of N_gpu - N_oracle (unit vectors); depth_rel_err:
    |d_gpu - d_oracle| / |d_oracle| over contacts deeper than 1e-3 x scale
    (relative error is meaningless at grazing depth), depth_abs_err over all
    matched contacts in units of scale (the typical shape size)."""
    gn, ga, gb, gd, gi = got
    key = lambda ids: (ids[:, 0].astype(np.int64) << 32) | ids[:, 1].astype(np.int64)
    kg, kr = key(gi), key(ref.bids)
    both, ig, ir = np.intersect1d(kg, kr, assume_unique=True, return_indices=True)
    out = {"contacts": int(len(both)), "only_gpu": int(len(kg) - len(both)), "only_oracle": int(len(kr) - len(both))}
    if len(both) == 0:
        return out
    ne = np.abs(gn[ig].astype(np.float64) - ref.norm[ir]).max(axis=1)
    dg, dr = gd[ig].astype(np.float64), ref.dpth[ir].astype(np.float64)
    deep = np.abs(dr) > 1e-3 * scale
    de = np.abs(dg - dr)
    rel = de[deep] / np.abs(dr[deep]) if deep.any() else np.zeros(1)
    out.update({"normal_err_max": float(ne.max()), "normal_err_p99": float(np.percentile(ne, 99)),
                "depth_rel_err_max": float(rel.max()), "depth_rel_err_p99": float(np.percentile(rel, 99)),
                "depth_abs_err_max_over_scale": float(de.max() / scale)})
    return out


def _host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "omp_proc_bind": os.environ.get("OMP_PROC_BIND", "")}


def _oracle_passes(oracle, s, flags, min_seconds):
    reps, t_total, parts = 0, 0.0, [0.0, 0.0, 0.0]
    pairs = contacts = 0
    while t_total < min_seconds:
        t0 = time.perf_counter()
        ab = oracle.aabb(s)
        t1 = time.perf_counter()
        bp = oracle.broadphase(ab, s.grid, flags)
        t2 = time.perf_counter()
        npr = oracle.narrowphase(s, bp.pairs)
        t3 = time.perf_counter()
        parts[0] += t1 - t0
        parts[1] += t2 - t1
        parts[2] += t3 - t2
        t_total += t3 - t0
        reps += 1
        pairs, contacts = len(bp.pairs), npr.count
    return reps, t_total, parts, pairs, contacts, bp.pairs


def _oracle_scene_rate(oracle, ss, grids, min_seconds):
    """Full oracle passes over a fixed list of scenes, one at a time: scenes/s."""
    done, t0 = 0, time.perf_counter()
    while True:
        for s, g in zip(ss, grids):
            ab = oracle.aabb(s)
            bp = oracle.broadphase(ab, g, oracle.DEFAULT_FLAGS)
            oracle.narrowphase(s, bp.pairs)
            done += 1
        if time.perf_counter() - t0 >= min_seconds:
            break
    dt = time.perf_counter() - t0
    return done / dt, done, dt


def cpu_leg(oracle, scenes, s, checks, min_seconds=10.0, all_parallel_seconds=5.0):
    """The CPU side of the run, after every GPU leg.

    Baselines: the oracle (C++/OpenMP restatement of the reference's OpenMP
    path), OpenMP on all host threads bound close, on bounded samples: the
    mixed-1M workload (~10 s with the reference's default parallel flags,
    ~5 s all-parallel), the 8,192-sphere workload, the broadphase of a
    2,097,152-sphere block of the 16M lattice, and per-scene loops over
    fixed samples of the batched and served scene sets.
    Checks: the GPU's contacts of each workload against the FP64 oracle
    (normal / depth error), and the FLOP census of the pairs the GPU's
    reference-order MPR launch ran on (the roofline's algorithmic FLOPs).
    """
    reps, t_total, parts, pairs, contacts, pair_list = _oracle_passes(oracle, s, oracle.DEFAULT_FLAGS, min_seconds)
    out = {
        "value": round(s.n * reps / t_total, 1),
        "unit": "bodies/s",
        "cores": oracle.max_threads(),
        "kind": "port",
        "sample": (f"{reps} full passes over the same {s.n}-body workload ({t_total:.1f} s), FP64, OpenMP on all "
                   f"host threads, reference default parallel flags (1,0,1,0,0,0,0,0,1,0); per pass aabb "
                   f"{parts[0] / reps:.3f}s broadphase {parts[1] / reps:.3f}s narrowphase {parts[2] / reps:.3f}s; "
                   f"{pairs} pairs, {contacts} contacts"),
        "host": _host_info(),
    }
    if all_parallel_seconds > 0:
        r2, t2, p2, _, _, _ = _oracle_passes(oracle, s, oracle.ALL_PARALLEL, all_parallel_seconds)
        out["all_parallel_flags"] = {
            "value": round(s.n * r2 / t2, 1), "unit": "bodies/s",
            "sample": (f"{r2} passes ({t2:.1f} s), every broadphase step parallel; per pass aabb {p2[0] / r2:.3f}s "
                       f"broadphase {p2[1] / r2:.3f}s narrowphase {p2[2] / r2:.3f}s")}
    out["mpr_steps"] = mpr_step_histogram(oracle, s, pair_list)
    ref1 = oracle.narrowphase(s, pair_list)
    configs = {}
    precision = {}
    c1 = checks.get("C1")
    if c1 is not None:
        precision["C1_mixed_1M"] = dict(mode=c1["mode"], **contact_error(c1["contacts"], ref1, 0.01))
        census = oracle.mpr_flops(s, pair_list)
        q = c1["mpr_queue"]
        out["mpr_flop_census"] = {
            "pairs": int(len(pair_list)), "flops_all_pairs": int(census.sum()),
            "mpr_launch_pairs": int(len(q)), "flops_mpr_launch": int(census[q].sum()),
            "note": "FP64 +, -, *, /, sqrt of function_MPR_Store per pair after world placement (oracle count)"}
    c0 = checks.get("C0")
    if c0 is not None:
        s0 = c0["scene"]
        t0 = time.perf_counter()
        r0 = 0
        while time.perf_counter() - t0 < 3.0:
            ab = oracle.aabb(s0)
            bp = oracle.broadphase(ab, s0.grid, oracle.DEFAULT_FLAGS)
            ref0 = oracle.narrowphase(s0, bp.pairs)
            r0 += 1
        dt = time.perf_counter() - t0
        configs["C0_spheres_8k"] = {"value": round(s0.n * r0 / dt, 1), "unit": "bodies/s",
                                    "sample": f"{r0} full passes over the 8,192-sphere FP64 workload ({dt:.1f} s)",
                                    "gpu_bodies_per_s": c0.get("gpu_bodies_per_s")}
        precision["C0_spheres_8k"] = dict(mode=c0["mode"], **contact_error(c0["contacts"], ref0, 0.01))
    c2 = checks.get("C2")
    if c2 is not None:
        blk = scenes.lattice_spheres(side=128)
        t0 = time.perf_counter()
        bp = oracle.broadphase(oracle.aabb(blk), blk.grid, oracle.DEFAULT_FLAGS)
        dt = time.perf_counter() - t0
        configs["C2_lattice_broadphase"] = {
            "value": round(blk.n / dt, 1), "unit": "bodies/s",
            "sample": (f"AABB + broadphase of a 128^3 block (2,097,152 FP32 spheres, same spacing and radii) of the "
                       f"16M lattice, one pass ({dt:.1f} s), {len(bp.pairs)} pairs"),
            "gpu_bodies_per_s": c2.get("gpu_broadphase_bodies_per_s")}
        sub = c2["scene"]
        s64 = sub.astype(np.float64)
        ref2 = oracle.narrowphase(s64, oracle.broadphase(oracle.aabb(s64), sub.grid).pairs)
        precision["C2_lattice_subset"] = dict(mode=c2["mode"], workload=c2["workload"],
                                              **contact_error(c2["contacts"], ref2, 0.005))
    c3 = checks.get("C3")
    if c3 is not None:
        ss, grids = c3["scenes"], c3["grids"]
        rate, done, dt = _oracle_scene_rate(oracle, ss, grids, 3.0)
        configs["C3_batched"] = {"value": round(rate, 2), "unit": "scenes/s",
                                 "sample": f"{len(ss)} fixed scenes of the batched set, full pass each, one at a "
                                           f"time ({done} scene passes, {dt:.1f} s)"}
        errs = [contact_error(got, oracle.narrowphase(sc, oracle.broadphase(oracle.aabb(sc), g).pairs), 0.01)
                for sc, g, got in zip(ss, grids, c3["contacts"])]
        precision["C3_batched_sample"] = {
            "mode": c3["mode"], "scenes": len(errs), "contacts": sum(e["contacts"] for e in errs),
            "only_gpu": sum(e["only_gpu"] for e in errs), "only_oracle": sum(e["only_oracle"] for e in errs),
            "normal_err_max": max(e.get("normal_err_max", 0.0) for e in errs),
            "depth_rel_err_max": max(e.get("depth_rel_err_max", 0.0) for e in errs)}
    c4 = checks.get("C4")
    if c4 is not None:
        ss = c4["scenes"]
        rate, done, dt = _oracle_scene_rate(oracle, ss, c4["grids"], 3.0)
        configs["C4_served"] = {"value": round(rate, 2), "unit": "scenes/s",
                                "sample": f"{len(ss)} scenes drawn from the served size distribution, full pass "
                                          f"each, one at a time ({done} scene passes, {dt:.1f} s)"}
    out["configs"] = configs
    out["precision_vs_fp64_oracle"] = precision
    return out


def mpr_step_histogram(oracle, s, pairs):
    """Work per candidate pair of the reference narrowphase, in support steps
    (pairs of support-map calls, contact points included), from the oracle."""
    st = oracle.mpr_steps(s, pairs)
    m = st[st > 0]
    edges = [(1, 1), (2, 3), (4, 8), (9, 16), (17, 32), (33, 38), (39, 73)]
    return {"pairs": int(len(st)), "mpr_pairs": int(len(m)), "mean_steps": round(float(m.mean()), 2) if len(m) else 0.0,
            "hist": {f"{a}-{b}" if a != b else str(a): int(((m >= a) & (m <= b)).sum()) for a, b in edges},
            "note": "39+ steps: the refinement ran to its 34-iteration cap"}


BP_STAGES = ["bp_bounds", "bp_count", "bp_scan", "bp_emit", "bp_sort", "bp_runs", "bp_pairs", "bp_pairscan",
             "bp_pairemit", "bp_pairsort"]


def dropin_leg(pkg, ctx, s, mode, steps):
    """The reference's call sequence through the drop-in calls: host arrays in,
    host arrays out, every call re-uploading what it reads (PCIe-inclusive)."""
    def one():
        aabb = ctx.generate_aabb(s)
        pairs = ctx.detect_possible_collisions(aabb, s.grid)
        out = ctx.do_narrowphase(s, pairs, mode=mode)
        return len(pairs), len(out[3])
    counts = one()
    t0 = time.perf_counter()
    for _ in range(steps):
        one()
    dt = (time.perf_counter() - t0) / steps
    return {"ms_per_pass": round(dt * 1e3, 3), "bodies_per_s": round(s.n / dt, 1), "passes": steps,
            "pairs": counts[0], "contacts": counts[1],
            "calls": "cdx_generate_aabb -> cdx_detect_possible_collisions + read -> cdx_do_narrowphase + read, "
                     "numpy host buffers"}


def lattice_leg(pkg, scenes, ctx, side, steps):
    """16,777,216-sphere FP32 granular pack (side 256): AABB + broadphase, then the full step.

    The scene's broadphase takes the cell search (near-uniform sizes).  Two
    byte counts are set against its time:
    * the reference decomposition's bytes at the reference grid,
      N*(24+4+8+28) + M*(8+16+4+24) + 8A + 8P (count, scan, key emit; key
      sort, bin starts, pair emit): the work the bin path moves, so the
      fraction states how fast the broadphase delivers the reference's result;
    * the cell search's own algorithmic bytes (_cell_algo_bytes), each array
      counted once per kernel that reads or writes it.
    The bin search at the reference grid and at the autotuned grid are timed
    beside it.
    """
    log(f"lattice leg: generating {side}^3 spheres")
    s = scenes.lattice_spheres(side=side)
    sc = pkg.Scene(ctx, np.float32).load(s)
    avg, wall, c = _bp_measure(sc, steps)
    cells = _cell_grid_cells(sc.read_aabb(shifted=True), s.n)
    bp_ms = sum(avg[k] for k in BP_STAGES)
    n, m, a, p = s.n, c.bin_entries, c.active_bins, c.pairs
    algo = _bp_algo_bytes(n, m, a, p)
    own = _cell_algo_bytes(n, p, cells, 4)
    gbs = algo / (bp_ms * 1e-3) / 1e9
    sc.set_mode(pkg.NP_NATIVE)
    t0 = time.perf_counter()
    sc.step()
    sc.sync()
    step_ms = (time.perf_counter() - t0) * 1e3
    np_ms = sum(v for k, v in sc.stage_ms().items() if k.startswith("np_"))
    cc = sc.counts()
    out = {
        "workload": f"lattice-{side}^3: {n} FP32 spheres, jittered cubic lattice, r in {{0.005, 0.0075}}",
        "pair_search": "cells" if c.flags & pkg.FLAG_CELL_SEARCH else "bins",
        "bins": list(s.grid), "bin_entries": int(m), "active_bins": int(a), "pairs": int(p),
        "cell_grid_cells": int(cells), "contacts": int(cc.contacts),
        "aabb_ms": round(avg["aabb"], 4), "broadphase_ms": round(bp_ms, 4),
        "broadphase_wall_ms_p50": round(float(np.median(wall)) * 1e3, 3),
        "broadphase_stage_ms": {k: round(avg[k], 4) for k in BP_STAGES},
        "broadphase_algorithmic_bytes": int(algo), "broadphase_GBs": round(gbs, 1),
        "broadphase_hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
        "broadphase_bytes_basis": "the reference bin decomposition at the reference grid "
                                  "(N*64 + M*52 + 8A + 8P bytes)",
        "cell_search_algorithmic_bytes": int(own),
        "cell_search_GBs": round(own / (bp_ms * 1e-3) / 1e9, 1),
        "cell_search_hbm_frac": round(own / (bp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "full_step_ms": round(step_ms, 3), "narrowphase_ms": round(np_ms, 3),
        "bodies_per_s_full_step": round(n / (step_ms * 1e-3), 1),
    }
    sc.close()
    # The bin search, at the reference grid and at the grid the per-size-class
    # autotuner picks (the pair list does not depend on the grid).
    log("lattice leg: bin search")
    bt = pkg.Batch(ctx, np.float32)
    tuned = bt.scene_bins(bt.add_scene(s, bins=(0, 0, 0)))
    bt.close()
    for key, grid in (("bin_search_reference_grid", s.grid), ("bin_search_autotuned_grid", tuned)):
        sc = pkg.Scene(ctx, np.float32).load(s, bins=grid)
        sc.set_pair_search(pkg.PAIRS_BINS)
        tavg, twall, tc = _bp_measure(sc, steps)
        sc.close()
        if tc.pairs != p:
            raise RuntimeError(f"the bin search found {tc.pairs} pairs, the cell search {p}")
        t_ms = sum(tavg[k] for k in BP_STAGES)
        t_algo = _bp_algo_bytes(n, tc.bin_entries, tc.active_bins, tc.pairs)
        out[key] = {
            "bins": list(grid), "bin_entries": int(tc.bin_entries), "active_bins": int(tc.active_bins),
            "pairs": int(tc.pairs), "broadphase_ms": round(t_ms, 4),
            "broadphase_wall_ms_p50": round(float(np.median(twall)) * 1e3, 3),
            "broadphase_stage_ms": {k: round(tavg[k], 4) for k in BP_STAGES},
            "broadphase_algorithmic_bytes": int(t_algo),
            "broadphase_GBs": round(t_algo / (t_ms * 1e-3) / 1e9, 1),
            "broadphase_hbm_frac": round(t_algo / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        }
    del s
    # a 64^3 block of the same lattice, full step in NATIVE FP32, for the
    # FP32 contact error against the FP64 oracle
    sub = scenes.lattice_spheres(side=64)
    sc = pkg.Scene(ctx, np.float32).load(sub)
    sc.set_mode(pkg.NP_NATIVE)
    sc.step()
    sc.sync()
    check = {"mode": "native-fp32", "scene": sub, "contacts": sc.read_contacts(),
             "workload": "lattice-64^3 (262,144 FP32 spheres, same spacing and radii)",
             "gpu_broadphase_bodies_per_s": round(n / (bp_ms * 1e-3), 1)}
    sc.close()
    return out, check


CELL_KX = 2  # fine cells along x per coarse cell of the library's cell search


def _cell_grid_cells(shifted_aabb, n):
    """Cells of the cell search's grid, as the library's scene broadphase plans
    it: coarse edge (1 + 1e-3) x the widest extent (at most 2N + 64 coarse
    cells), CELL_KX times finer along x."""
    lo, hi = shifted_aabb[:n].astype(np.float64), shifted_aabb[n:].astype(np.float64)
    w = float(np.float32((hi - lo).max()))
    ext = np.maximum(hi.max(axis=0), 0.0)
    target = 2.0 * n + 64.0
    edge = w * (1.0 + 1e-3)
    for _ in range(64):
        g = np.floor(ext / edge) + 1.0
        if float(np.prod(g)) <= target:
            break
        edge *= max(float(np.cbrt(np.prod(g) / target)), 1.0 + 1e-3)
    return int((np.floor(ext[0] * CELL_KX / edge) + 1.0) * g[1] * g[2])


def _cell_algo_bytes(n, p, cells, rs):
    """Bytes the cell search moves, each array once per kernel touching it:
    bin count pass (AABB 6rs, count 4) and its scan (4 + 4); cell keys (AABB,
    entry 8, histogram read and write 8), table clear and scan (4 + 8 per
    cell); radix sort (8 read twice, 8 written, per pass); gather (entry,
    AABB, the two entry halves 2h, last index 4 per cell after its clear);
    window maxima (4 + 4 per cell); search (own halves, count 4, 4 per
    partner into the slab); count scan (4 + 4); placement (count, offset,
    4 per partner read back, 8 per pair written)."""
    h = 16 if rs == 4 else 32
    bits = max(1, int(np.ceil(np.log2(max(cells, 2)))))
    passes = -(-bits // 11)
    per_n = (6 * rs + 4) + 8 + (6 * rs + 16) + 24 * passes + (8 + 6 * rs + 2 * h) + (2 * h + 4) + 8 + 8
    return n * per_n + 16 * p + 24 * cells


def _bp_algo_bytes(n, m, a, p):
    return n * (24 + 4 + 8 + 28) + m * (8 + 16 + 4 + 24) + 8 * a + 8 * p


def _bp_measure(sc, steps):
    """AABB + broadphase `steps` times after one untimed pass: mean stage times (HIP events), wall times, counts."""
    sc.set_profiling(True)
    sc.aabb()
    sc.broadphase()
    sc.sync()
    c = sc.counts()
    tot = {}
    wall = []
    for _ in range(steps):
        t0 = time.perf_counter()
        sc.aabb()
        sc.broadphase()
        sc.sync()
        wall.append(time.perf_counter() - t0)
        for k, v in sc.stage_ms().items():
            tot[k] = tot.get(k, 0.0) + v
    return {k: v / steps for k, v in tot.items()}, wall, c


def _gen_threaded(fn, args, workers=16):
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(lambda a: fn(*a), args))


def batched_leg(pkg, scenes, ctx, count, steps):
    """512 independent scenes (log-uniform 1,024-65,536 mixed bodies, U[2,8] neighbours), FP64 input, MIXED.

    One batched step = AABB + broadphase + narrowphase over every scene as one
    recorded HIP graph, grids from the per-size-class autotuner; it ends when
    the per-scene contact counts are on the host.  The plain per-scene path
    steps the same scenes one at a time (one resident cdx_scene each, same
    grid), each step ending when its contact count is on the host.
    """
    log(f"batched leg: generating {count} scenes")
    n, k = scenes.batched_sizes(count)
    ss = _gen_threaded(scenes.fleet_scene, [(int(n[i]), float(k[i]), 0x5EED0003 + 1 + i) for i in range(count)])
    b = pkg.Batch(ctx, np.float64)
    t0 = time.perf_counter()
    for s in ss:
        b.add_scene(s, bins=(0, 0, 0))
    t_add = time.perf_counter() - t0
    b.set_mode(pkg.NP_MIXED, 1e-2)
    for _ in range(2):
        b.step(True)
        tot = b.finish()
    lat = []
    dev = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        b.step(True)
        b.finish()
        lat.append(time.perf_counter() - t1)
        dev.append(b.step_ms())
    total = time.perf_counter() - t0
    per_scene = [b.scene_counts(j) for j in range(count)]
    grids = {}
    for j, s in enumerate(ss):
        cls = int(np.log2(max(s.n, 1)) // 2)
        grids.setdefault(cls, []).append(b.scene_bins(j))
    # plain per-scene path on a sample
    log("batched leg: plain per-scene path")
    plain = []
    mismatch = 0
    for j, s in enumerate(ss):
        sc = pkg.Scene(ctx, np.float64).load(s, bins=b.scene_bins(j))
        sc.set_mode(pkg.NP_MIXED, 1e-2)
        sc.step()
        sc.sync()
        t1 = time.perf_counter()
        sc.step()
        sc.sync()
        plain.append(time.perf_counter() - t1)
        cc = sc.counts()
        if cc.pairs != per_scene[j].pairs or cc.contacts != per_scene[j].contacts:
            mismatch += 1
        sc.close()
    sample = list(range(0, count, max(1, count // 16)))[:16]
    check = {"mode": "mixed", "scenes": [ss[j] for j in sample], "grids": [b.scene_bins(j) for j in sample],
             "contacts": [b.read_contacts(j) for j in sample]}
    lat_ms = np.array(lat) * 1e3
    plain_sps = count / float(np.sum(plain))
    sps = count * steps / total
    out = {
        "workload": f"{count} scenes, log-uniform 1024-65536 mixed bodies, U[2,8] neighbours, FP64 input, "
                    "FP32 MPR + FP64 re-solve, autotuned grids",
        "bodies": int(sum(s.n for s in ss)), "pairs": int(tot.pairs), "contacts": int(tot.contacts),
        "scenes_per_s": round(sps, 1), "bodies_per_s": round(sps * sum(s.n for s in ss) / count, 1),
        "p50_step_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_step_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "device_step_ms_p50": round(float(np.median(dev)), 3),
        "plain_scenes_per_s": round(plain_sps, 1),
        "plain_step_ms_p50": round(float(np.median(plain)) * 1e3, 3),
        "speedup_vs_plain": round(sps / plain_sps, 2),
        "per_scene_counts_match_plain": mismatch == 0,
        "add_scenes_s": round(t_add, 2),
        "autotuned_grids_by_size_class": {str(c): {"scenes": len(g), "distinct_grids": len({tuple(x) for x in g}),
                                                   "first": g[0]} for c, g in sorted(grids.items())},
        "device_bytes": int(b.device_bytes()),
    }
    b.close()
    return out, check


def served_leg(pkg, scenes, ctx, args):
    """Poisson step requests against 4,096 resident scenes (log-uniform 256-262,144 mixed bodies, FP64, MIXED).

    Each request replaces every body pose of one uniformly chosen scene with
    the next of two precomputed pose sets (displacement N(0, (0.05 size)^2),
    rotation by N(0, 0.05) rad about a random axis) and returns its contacts.
    Producer threads write each request's poses into a pinned buffer of the
    server ahead of its arrival (as a caller producing poses in place would);
    latency runs from the submit call at the arrival time to the result being
    ready on the host.
    The plain per-scene path runs the same requests one at a time on
    resident cdx_scene objects (pose upload + step + contact readback).
    """
    import threading
    count = args.served_scenes
    sizes = scenes.served_sizes(count)
    log(f"served leg: generating {count} scenes ({int(sizes.sum())} bodies)")
    rng0 = np.random.default_rng(0x5EED0005)
    seeds = rng0.integers(0, 2**62, count)

    def make(i):
        s = scenes.fleet_scene(int(sizes[i]), 4.0, int(seeds[i]))
        p1, r1 = scenes.pose_step(s, np.random.default_rng(int(seeds[i]) + 1))
        return s, p1, r1

    srv = pkg.Server(ctx, np.float64, mode=pkg.NP_MIXED, max_batch_shapes=args.served_batch_shapes,
                     max_pending=args.served_max_pending, use_graph=not args.served_eager)
    poses = [None] * count
    t0 = time.perf_counter()
    from concurrent.futures import ThreadPoolExecutor
    refused = 0
    with ThreadPoolExecutor(16) as ex:
        for c0 in range(0, count, 256):  # bounded look-ahead keeps host memory in check
            for i, (s, p1, r1) in zip(range(c0, min(count, c0 + 256)), ex.map(make, range(c0, min(count, c0 + 256)))):
                try:
                    sid = srv.add_scene(s, bins=(0, 0, 0))
                    assert sid == i - refused
                    poses[i] = ((s.pos, s.rot), (p1, r1))
                except pkg.CdxError as e:
                    if e.status != 4:
                        raise
                    refused += 1
            log(f"served leg: {c0 + 256} scenes added")
    t_setup = time.perf_counter() - t0
    ids = [i for i in range(count) if poses[i] is not None]
    sid_of = {i: j for j, i in enumerate(ids)}
    mem = srv.memory()
    log(f"served leg: {len(ids)} scenes resident ({refused} refused) in {t_setup:.1f}s, {mem}")

    # plain per-scene path
    rng = np.random.default_rng(7)
    sample = rng.choice(ids, min(args.plain_sample, len(ids)), replace=False)
    plain = []
    for i in sample:
        (p0, r0), (p1, r1) = poses[i]
        s = scenes.fleet_scene(int(sizes[i]), 4.0, int(seeds[i]))
        sc = pkg.Scene(ctx, np.float64).load(s, bins=srv.scene_info(sid_of[i])["bins"])
        sc.set_mode(pkg.NP_MIXED, 1e-2)
        idx = np.arange(s.nb, dtype=np.uint32)
        sc.step()
        sc.sync()
        t1 = time.perf_counter()
        sc.update_bodies(idx, p1, r1)
        sc.step()
        sc.sync()
        sc.read_contacts()  # a request returns its contacts, on both paths
        plain.append(time.perf_counter() - t1)
        sc.close()
    plain = np.array(plain)
    # the sample's mean cost, weighted to the whole fleet's mean size
    plain_sps = len(plain) / float(plain.sum())
    cpu_ids = [int(i) for i in sample[:16]]
    check = {"scenes": [scenes.fleet_scene(int(sizes[i]), 4.0, int(seeds[i])) for i in cpu_ids],
             "grids": [srv.scene_info(sid_of[i])["bins"] for i in cpu_ids]}

    # warm up every class (graph capture, pinned pose and result blocks of
    # every size bucket) with one request per scene
    state = {i: 0 for i in ids}
    warm = []
    for i in ids:
        warm.append(srv.submit(sid_of[i], *poses[i][1]))
        if len(warm) >= 64:
            for r in warm:
                srv.wait(r, 60_000)
                srv.release(r)
            warm = []
    for r in warm:
        srv.wait(r, 60_000)
        srv.release(r)
    srv.stats(reset=True)

    points = []
    for rate in [float(x) for x in args.served_rates.split(",") if x]:
        rr = np.random.default_rng(int(rate))
        n_req = int(rate * args.served_seconds)
        arrivals = np.cumsum(rr.exponential(1.0 / rate, n_req))
        targets = rr.choice(ids, n_req)
        which = np.zeros(n_req, np.int8)
        for r, i in enumerate(targets):  # each request of a scene takes the other pose set
            state[i] ^= 1
            which[r] = state[i]
        # Producers fill pinned pose buffers ahead of the arrivals (the caller
        # writes its new poses in place); a request is submitted at its
        # arrival time, and its latency runs from that submit call.
        ready = {}
        cond = threading.Condition()
        ahead = 256

        def producer(start, step):
            for r in range(start, n_req, step):
                with cond:
                    cond.wait_for(lambda: r < progress[0] + ahead)
                i = targets[r]
                pos, rot = poses[i][which[r]]
                h, bp, br = srv.pose_buffer(sid_of[i], len(pos))
                np.copyto(bp, pos)
                np.copyto(br, rot)
                with cond:
                    ready[r] = h
                    cond.notify_all()

        outstanding = []
        lock = threading.Lock()
        done_flag = [False]
        progress = [0]

        def reaper():
            while True:
                with lock:
                    item = outstanding.pop(0) if outstanding else None
                if item is None:
                    if done_flag[0]:
                        return
                    time.sleep(0.0002)
                    continue
                srv.wait(item, 120_000)
                srv.release(item)

        prods = [threading.Thread(target=producer, args=(k, 4)) for k in range(4)]
        for t in prods:
            t.start()
        th = threading.Thread(target=reaper)
        th.start()
        with cond:  # let the producers fill the first buffers before the clock starts
            cond.wait_for(lambda: len(ready) >= min(ahead // 2, n_req), timeout=30)
        srv.stats(reset=True)
        refused_q = 0
        lag = []
        gc.disable()  # no collector pauses in the load generator while arrivals run
        t_start = time.perf_counter()
        for r in range(n_req):
            now = time.perf_counter() - t_start
            if arrivals[r] > now:
                time.sleep(arrivals[r] - now)
            with cond:
                cond.wait_for(lambda: r in ready)
                h = ready.pop(r)
                progress[0] = r + 1
                cond.notify_all()
            lag.append(time.perf_counter() - t_start - arrivals[r])
            try:
                rid = srv.submit_buffer(h)
            except pkg.CdxError:
                refused_q += 1
                srv.drop_buffer(h)
                continue
            with lock:
                outstanding.append(rid)
        t_submit_end = time.perf_counter() - t_start
        done_flag[0] = True
        th.join()
        for t in prods:
            t.join()
        gc.enable()
        gc.collect()
        st = srv.stats(reset=True)
        wall = time.perf_counter() - t_start
        pt = {"rate_rps": rate, "requests": int(st["completed"]), "queue_refused": refused_q,
              "scenes_per_s": round(st["completed"] / wall, 1),
              "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3), "max_ms": round(st["max_ms"], 3),
              "steps": int(st["steps"]), "requests_per_step": round(st["stepped_requests"] / max(st["steps"], 1), 2),
              "device_busy_frac": round(st["busy_ms"] / 1e3 / wall, 3),
              "worker_host_ms_per_step": {k: round(v / max(st["steps"], 1), 3) for k, v in st["host_ms"].items()},
              "classes": [{"completed": int(k["completed"]), "steps": int(k["steps"]),
                           "requests_per_step": round(k["completed"] / max(k["steps"], 1), 2),
                           "p50_ms": round(k["p50_ms"], 3), "p99_ms": round(k["p99_ms"], 3),
                           "busy_ms_per_step": round(k["busy_ms"] / max(k["steps"], 1), 3)} for k in st["classes"]],
              "submit_lag_p99_ms": round(float(np.percentile(lag, 99)) * 1e3, 3),
              "submit_end_lag_s": round(t_submit_end - float(arrivals[-1]), 3)}
        log(f"served rate {rate}: {pt}")
        points.append(pt)
    ok = [p for p in points if p["p99_ms"] <= 2.0]
    knee = max(ok, key=lambda p: p["scenes_per_s"]) if ok else None
    best = max(points, key=lambda p: p["scenes_per_s"]) if points else None
    out = {
        "workload": f"{count} resident scenes, log-uniform 256-262144 mixed bodies (mean {sizes.mean():.0f}), "
                    "FP64 input, FP32 MPR + FP64 re-solve, Poisson arrivals, full pose update per request",
        "resident_scenes": len(ids), "refused_scenes": refused, "resident_bodies": int(sizes[ids].sum()),
        "hbm_committed_GB": round(mem["hbm_committed"] / 1e9, 2), "device_GB": round(mem["device_bytes"] / 1e9, 2),
        "setup_s": round(t_setup, 1),
        "plain_scenes_per_s": round(plain_sps, 1),
        "plain_p50_ms": round(float(np.median(plain)) * 1e3, 3),
        "plain_p99_ms": round(float(np.percentile(plain, 99)) * 1e3, 3),
        "plain_sample": int(len(plain)),
        "sweep": points,
        "knee_p99_le_2ms": knee,
        "speedup_vs_plain_at_knee": round(knee["scenes_per_s"] / plain_sps, 2) if knee else None,
        "best_scenes_per_s": best["scenes_per_s"] if best else None,
        "speedup_vs_plain_best": round(best["scenes_per_s"] / plain_sps, 2) if best else None,
    }
    srv.close()
    return out, check


def copy_calibration(nbytes, reps=5):
    """A plain device copy of nbytes (read nbytes, write nbytes) for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = nbytes // 4
    src = torch.rand(n, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ms = []
    for _ in range(reps):
        ev[0].record()
        torch.mul(src, 1.0, out=dst)
        ev[1].record()
        torch.cuda.synchronize()
        ms.append(ev[0].elapsed_time(ev[1]))
    best = min(ms)
    print(json.dumps({"copy_bytes": n * 4, "kernel": "torch elementwise mul by 1 (read n, write n)",
                      "best_ms": round(best, 4), "GBs": round(2 * n * 4 / (best * 1e-3) / 1e9, 1)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", choices=["mixed", "native", "fast"], default="mixed")
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--dropin-passes", type=int, default=3, help="passes through the drop-in calls (0 skips)")
    ap.add_argument("--lattice-side", type=int, default=256, help="0 skips the 16M-sphere broadphase leg")
    ap.add_argument("--lattice-steps", type=int, default=5)
    ap.add_argument("--batched-scenes", type=int, default=512, help="0 skips the batched leg")
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--served-scenes", type=int, default=4096, help="0 skips the served leg")
    ap.add_argument("--served-rates", type=str, default="500,1000,2000,3000,3500,4000,4500,5000,6000",
                    help="Poisson arrival rates (requests/s) swept by the served leg")
    ap.add_argument("--served-seconds", type=float, default=3.0, help="seconds of arrivals per rate")
    ap.add_argument("--served-batch-shapes", type=int, default=1 << 19)
    ap.add_argument("--served-max-pending", type=int, default=4096,
                    help="queued requests before the server refuses new ones")
    ap.add_argument("--served-eager", action="store_true", help="launch served steps without HIP graphs")
    ap.add_argument("--plain-sample", type=int, default=256, help="scenes timed on the plain per-scene path")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a plain device copy of this many bytes (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_bytes:
        copy_calibration(args.copy_bytes)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this pipeline runs on one GPU (one MI355X per process); "
                                   "multi-GPU sharding of scenes is not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    pkg = _package()
    import collision_detection_amd.scenes as scenes

    log("generating workload")
    s = scenes.mixed_bodies(n=args.n, seed=0x5EED0001)
    mode = {"mixed": pkg.NP_MIXED, "native": pkg.NP_NATIVE, "fast": pkg.NP_FAST}[args.mode]
    ctx = pkg.Context(0)
    sc = pkg.Scene(ctx, s.dtype).load(s)
    sc.set_mode(mode, 1e-2)
    sc.set_profiling(True)
    log(f"n={s.n} grid={s.grid} mode={args.mode}")
    for _ in range(args.warmup):
        sc.step()
        sc.sync()
    c = sc.counts() if args.warmup > 0 else None
    if c is not None:
        log(f"pairs={c.pairs} contacts={c.contacts} bin_entries={c.bin_entries} resolved={c.resolved}")
    lat = []
    stage_tot = {k: 0.0 for k in pkg.STAGES}
    t_start = time.perf_counter()
    for k in range(args.steps):
        t0 = time.perf_counter()
        sc.step()
        sc.sync()
        lat.append(time.perf_counter() - t0)
        for name, v in sc.stage_ms().items():
            stage_tot[name] += v
        if c is None:
            c = sc.counts()
    total = time.perf_counter() - t_start
    c2 = sc.counts()
    assert c2.pairs == c.pairs and c2.contacts == c.contacts, "non-deterministic step"
    ms_step = 1000.0 * total / args.steps
    stage_avg = {k: v / args.steps for k, v in stage_tot.items()}
    log("stage ms: " + ", ".join(f"{k}={v:.3f}" for k, v in stage_avg.items()))

    # The same step as a one-scene batch: buffers sized by capacity, the
    # launch sequence recorded once as a HIP graph and replayed, totals read
    # back once per step (no host round trips between stages).
    b1 = pkg.Batch(ctx, s.dtype)
    b1.add_scene(s)
    b1.set_mode(mode, 1e-2)
    for _ in range(max(args.warmup, 2)):
        b1.step(True)
        gc = b1.finish()
    glat = []
    g0 = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        b1.step(True)
        gc = b1.finish()
        glat.append(time.perf_counter() - t0)
    gtotal = time.perf_counter() - g0
    assert gc.pairs == c.pairs and gc.contacts == c.contacts, "graph step disagrees with the scene step"
    graph_path = {"ms_per_step": round(1000.0 * gtotal / args.steps, 4),
                  "bodies_per_s": round(s.n * args.steps / gtotal, 1),
                  "p50_step_ms": round(float(np.percentile(np.array(glat) * 1e3, 50)), 4),
                  "p99_step_ms": round(float(np.percentile(np.array(glat) * 1e3, 99)), 4)}
    log(f"graph path: {graph_path}")
    b1.close()

    checks = {"C1": {"mode": args.mode, "contacts": sc.read_contacts(), "mpr_queue": sc.read_mpr_queue()}}

    # The HBM-bound single-kernel stages: algorithmic bytes over their
    # HIP-event times on the scene stream (the slowest one leads).
    bytes_ = stage_bytes(s, c)
    hbm = max(bytes_, key=lambda k: stage_avg.get(k, 0.0))
    hbm_ms = stage_avg[hbm]
    hbm_gbs = bytes_[hbm] / (hbm_ms * 1e-3) / 1e9 if hbm_ms > 0 else 0.0
    hbm_stage = {
        "bound": "hbm", "kernel": STAGE_KERNEL[hbm], "stage": hbm, "achieved": round(hbm_gbs, 1),
        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(hbm_gbs / HBM_PEAK_GBS, 4),
        "traffic": read_traffic(hbm), "algorithmic_bytes": int(bytes_[hbm]), "avg_launch_ms": round(hbm_ms, 4),
        "all_single_kernel_stages": {
            k: {"ms": round(stage_avg[k], 4), "GBs": round(b / (stage_avg[k] * 1e-3) / 1e9, 1) if stage_avg[k] else None}
            for k, b in bytes_.items()},
    }

    # C0: the 8,192-sphere FP64 workload, full step (precision check and
    # GPU rate beside its CPU baseline).
    s0 = scenes.small_spheres()
    sc0 = pkg.Scene(ctx, s0.dtype).load(s0)
    sc0.set_mode(mode, 1e-2)
    sc0.step()
    sc0.sync()
    t0 = time.perf_counter()
    for _ in range(20):
        sc0.step()
        sc0.sync()
    c0_rate = s0.n * 20 / (time.perf_counter() - t0)
    checks["C0"] = {"mode": args.mode, "scene": s0, "contacts": sc0.read_contacts(),
                    "gpu_bodies_per_s": round(c0_rate, 1)}
    sc0.close()

    dropin = None
    if args.dropin_passes > 0:
        dropin = dropin_leg(pkg, ctx, s, mode, args.dropin_passes)
        assert dropin["pairs"] == c.pairs and dropin["contacts"] == c.contacts, "drop-in calls disagree with the step"
        log(f"drop-in calls: {dropin}")

    batched = None
    if args.batched_scenes > 0:
        batched, checks["C3"] = batched_leg(pkg, scenes, ctx, args.batched_scenes, args.batched_steps)
        log(f"batched: {json.dumps(batched)}")

    served = None
    if args.served_scenes > 0:
        served, checks["C4"] = served_leg(pkg, scenes, ctx, args)
        log(f"served: {json.dumps(served)}")

    lattice = None
    if args.lattice_side > 0:
        lattice, checks["C2"] = lattice_leg(pkg, scenes, ctx, args.lattice_side, args.lattice_steps)
        log(f"lattice: {lattice}")

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        log("timing the CPU baseline and checking contacts against the FP64 oracle")
        cpu = cpu_leg(oracle, scenes, s, checks)

    # Roofline of the step's dominant kernel: the reference-order FP64 MPR
    # (k_np_resolve), timed alone with HIP events on the scene stream; its
    # algorithmic FLOPs are the oracle's census of the reference algorithm's
    # FP64 operations on exactly the pairs that launch ran on.
    mpr_ms = stage_avg.get("np_mpr", 0.0)
    census = cpu.get("mpr_flop_census") if cpu else None
    flops = census["flops_mpr_launch"] if census else None
    tflops = flops / (mpr_ms * 1e-3) / 1e12 if flops and mpr_ms > 0 else None
    roofline = {
        "bound": "mfma",
        "compute_unit": ("FP64 vector ALU; the kernel issues no MFMA instruction: MI355X's FP64 vector peak equals "
                         "its FP64 matrix peak, 78.6 TFLOP/s"),
        "achieved": round(tflops, 3) if tflops is not None else None,
        "peak": FP64_VECTOR_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(tflops / FP64_VECTOR_PEAK_TFLOPS, 4) if tflops is not None else None,
        "traffic": read_traffic("np_mpr"),
        "kernel": "k_np_resolve<double, double>",
        "stage": "np_mpr",
        "avg_launch_ms": round(mpr_ms, 4),
        "share_of_step": round(mpr_ms / sum(stage_avg[k] for k in pkg.STAGES if k != "np_mpr"), 3),
        "pairs_per_launch": int(len(checks["C1"]["mpr_queue"])),
        "algorithmic_flops_per_launch": int(flops) if flops else None,
        "flops_unit": "FP64 +, -, *, /, sqrt of the reference MPR per pair (oracle census), summed over the launch",
        "issued_fp64_counters": read_resolve_fp64(),
    }

    lat_ms = np.array(lat) * 1000.0
    scene_path = {"ms_per_step": round(ms_step, 4), "bodies_per_s": round(s.n * args.steps / total, 1),
                  "p50_step_ms": round(float(np.percentile(lat_ms, 50)), 4),
                  "p99_step_ms": round(float(np.percentile(lat_ms, 99)), 4)}
    out = {
        "metric": "bodies/sec AABB+broad+MPR at 1M mixed shapes; scenes/sec + p50/p99 step lat",
        "value": graph_path["bodies_per_s"],
        "unit": "bodies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": graph_path["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.mode == "native" else "f64+f32",
        "data": "synthetic",
        "config": {"workload": "mixed-1M: 1,048,576 mixed convex bodies, FP64 input, ~4 AABB neighbours/body",
                   "n_bodies": s.n, "bins": list(s.grid), "narrowphase": args.mode,
                   "pairs": int(c.pairs), "contacts": int(c.contacts), "fp64_resolved": int(c.resolved)},
        "p50_step_ms": graph_path["p50_step_ms"],
        "p99_step_ms": graph_path["p99_step_ms"],
        "stage_ms": {k: round(v, 4) for k, v in stage_avg.items()},
        "narrowphase_pairs_per_s": round(c.pairs / (1e-3 * (stage_avg["np_main"] + stage_avg["np_resolve"] +
                                                             stage_avg["np_compact"])), 1),
        "scene_path": scene_path,
        "dropin_path": dropin,
        "roofline": roofline,
        "hbm_stage_roofline": hbm_stage,
        "cpu_baseline": cpu,
        "lattice_16M_broadphase": lattice,
        "batched_512": batched,
        "served_4096": served,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
