#!/usr/bin/env python3
"""Throughput of the SPH step on one MI355X: particle-updates/s.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--config dam1m]

Workload (default): plain WCSPH dam break with 1,048,576 particles
(128x64x128 lattice at 0.005 m, h = 0.01 m, box 2.56x1.28x2.56 m with a
sphere + cone obstacle on a 256x128x256 distance grid).  A "step" is one
full time step (hash/sort/cells/reorder/density/force+integrate) of all
particles, inputs resident in HBM.  K steps are timed with HIP events on the
simulation's stream after W warm-up steps; per-stage device times come from
events recorded around each stage inside the same timed region.

Besides the headline (plain 1M), the line carries the other forms of the
same step measured in the same run (`forms`): the 1M scene with the learned
boundary attached, and 64 independent scenes of uneven size (log-uniform
4,096..262,144 particles, seed 1) run as 64 separate handles, as one batch
(one launch per stage, hipGraph-captured step) and as one batch with a
learned boundary per scene.  `--no-forms` skips them.

Prints one JSON line (stdout).  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-updates/s at 1M (plain vs batched+neural-SDF); achieved HBM GB/s"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "dam1m": "plain WCSPH dam break, 1,048,576 FP32 particles (128x64x128 lattice, 0.005 m, h=0.01 m), "
             "2.56x1.28x2.56 m box with sphere (r=0.3 m) + cone (r=0.25 m, 0.5 m) obstacle on a "
             "256x128x256 FP32 distance grid, dt=1e-4 s",
    "dam16k": "plain WCSPH dam break, 16,384 FP32 particles (32x16x32 lattice, 0.01 m, h=0.02 m), "
              "0.64^3 m box on a 64^3 FP32 distance grid, dt=1e-4 s",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def touched_cells(params, sorted_keys: np.ndarray) -> int:
    """Distinct cell-table entries read by the neighbour loops: the 3x3x3
    neighbourhoods of the occupied cells (clipped to the grid)."""
    nx, ny, nz = (int(params.grid_dim[a]) for a in range(3))
    occ = np.unique(sorted_keys.astype(np.int64))
    x, y, z = occ % nx, (occ // nx) % ny, occ // (nx * ny)
    seen = []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                xx, yy, zz = x + dx, y + dy, z + dz
                ok = (xx >= 0) & (xx < nx) & (yy >= 0) & (yy < ny) & (zz >= 0) & (zz < nz)
                seen.append((xx + nx * (yy + ny * zz))[ok])
    return int(np.unique(np.concatenate(seen)).size)


# Algorithmic (compulsory) bytes per particle of each single-kernel stage:
# each minimal unpadded FP32 field read or written once per launch, neighbour
# re-reads counted as on-chip hits -- position 12, velocity 12, density 4,
# key 4, id 4.
#   density: read x, write rho                                   16 B
#   force + integration + collision + next key: read x, v, rho;
#     write x, v, key, id                                         60 B
#   reorder: read key, slot, x, v, id; write x, v, id            64 B
# plus, for the neighbour passes, 8 B per touched cell-table entry (T_c) and
# 4 B per distinct boundary texel of each grid they sample.
STAGE_BYTES_PER_PARTICLE = {
    "density": 12 + 4,
    "force": 12 + 12 + 4 + 12 + 12 + 4 + 4,
    "reorder": 4 + 4 + 12 + 12 + 4 + 12 + 12 + 4,
    "cell_ranges": 4 + 0,
}


def touched_texels(sc, sim) -> dict:
    """Distinct boundary-grid nodes the trilinear lookups of one step touch
    (the 8 corners of each lookup): the distance and wall-density lookup at
    every particle whose cell has a wall term, and, where the particle is in
    the wall-penalty zone (distance > -r0), the six central-difference
    lookups of the distance (+-dims/(dims-1) texels) and the viscosity at
    the centre."""
    from cudasphnew_amd import Field
    x = sim.read(Field.POSITION).astype(np.float32)
    wall = sim.read(Field.WALL_SOURCE) != 0
    dims = np.array(sc.boundary_dims, np.int64)              # (nx, ny, nz)
    o = np.array(sc.desc.boundary_origin[:], np.float32)
    sp = np.float32(sc.desc.boundary_spacing)
    L = (dims - 1).astype(np.float32) * sp
    t = (x[wall] - o) / L * dims.astype(np.float32) - np.float32(0.5)

    def corners(tt):
        tt = np.clip(tt, -1.0, dims.astype(np.float32))
        f = np.floor(tt).astype(np.int64)
        out = []
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    c = np.clip(f + np.array([dx, dy, dz]), 0, dims - 1)
                    out.append(c[:, 0] + dims[0] * (c[:, 1] + dims[1] * c[:, 2]))
        return np.concatenate(out)

    dist = sc.distance.reshape(-1)
    cen = corners(t)
    # distance at the particle (trilinear) decides the penalty zone
    tt = np.clip(t, -1.0, dims.astype(np.float32))
    f = np.floor(tt)
    a = tt - f
    fi = f.astype(np.int64)
    d = np.zeros(t.shape[0], np.float64)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                c = np.clip(fi + np.array([dx, dy, dz]), 0, dims - 1)
                w = np.where(dx, a[:, 0], 1 - a[:, 0]) * np.where(dy, a[:, 1], 1 - a[:, 1]) * \
                    np.where(dz, a[:, 2], 1 - a[:, 2])
                d += w * dist[c[:, 0] + dims[0] * (c[:, 1] + dims[1] * c[:, 2])]
    # (a constant -r0 neighbourhood interpolates to -r0 up to rounding here;
    # the device's a + t (b - a) form keeps it exact)
    pz = d > -float(sc.desc.rest_distance) * (1.0 - 1e-6)
    step = dims / (dims - 1)
    grad = [corners(t[pz] + s * step * e) for e in np.eye(3) for s in (1.0, -1.0)]
    t_d = np.unique(np.concatenate([cen] + grad)).size
    t_v = np.unique(corners(t[pz])).size if pz.any() else 0
    t_rho = np.unique(cen).size
    return {"T_D": int(t_d), "T_nu": int(t_v), "T_rho": int(t_rho), "wall_particles": int(wall.sum()),
            "penalty_zone": int(pz.sum())}


def copy_probe(nbytes: int) -> None:
    """One device-to-device copy of known size through the HIP runtime the
    library itself uses (libamdhip64), to calibrate rocprof byte counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("copy probe: hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)       # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy probe: {nbytes} bytes read, {nbytes} bytes written")


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_topology() -> dict:
    """Logical CPUs this process may run on and the physical cores they
    belong to (distinct (physical id, core id) pairs in /proc/cpuinfo)."""
    allowed = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count()))
    cores, cur = set(), {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if cur.get("processor") in allowed:
                    cores.add((cur.get("physical id", 0), cur.get("core id", cur["processor"])))
                cur = {}
                continue
            k, v = (t.strip() for t in line.split(":", 1))
            if k in ("processor", "physical id", "core id"):
                cur[k] = int(v)
        if cur.get("processor") in allowed:
            cores.add((cur.get("physical id", 0), cur.get("core id", cur["processor"])))
    except (OSError, ValueError):
        pass
    return {"logical_cpus": len(allowed), "physical_cores": len(cores) or None}


def _cpu_rate(scene, seconds: float, threads: int):
    from oracle import oracle as O
    sim = O.OracleSim(scene, real=1, threads=threads)
    sim.step(1)                                  # warm-up
    steps, total = 0, 0.0
    while total < seconds and steps < 200:
        total += sim.step(1)
        steps += 1
    return scene.n * steps / total, steps, total


def cpu_baseline(scene, seconds: float, threads: int):
    """FP32 C++ transcription of the reference step on the host's cores:
    all the threads this process may use (the reported value) and one."""
    topo = cpu_topology()
    v, steps, total = _cpu_rate(scene, seconds, threads)
    v1, steps1, total1 = _cpu_rate(scene, seconds / 2, 1)
    return {
        "value": v,
        "unit": "particle-updates/s",
        "cores": threads,
        "cpu": f"{cpu_model()}: {topo['logical_cpus']} logical CPUs on {topo['physical_cores']} physical cores "
               f"available to this process",
        "physical_cores": topo["physical_cores"],
        "kind": "port",
        "sample": f"{steps} steps of the same {scene.n}-particle workload after 1 warm-up step, "
                  f"FP32 C++ transcription of the reference step (oracle/, OpenMP, {threads} threads), "
                  f"{total:.1f} s",
        "single_thread": {"value": v1, "unit": "particle-updates/s", "cores": 1,
                          "sample": f"{steps1} steps, 1 thread, {total1:.1f} s"},
    }


def kernel_name(stage: str, tuning: dict, one_scene: bool = True) -> str:
    """The kernel (as rocprofv3 names it, template arguments included) that a
    stage launched under the given variants."""
    b, one = tuning["block"], "true" if one_scene else "false"
    if stage == "force":
        if tuning["hit_masks"]:
            return f"k_force_masks<{tuning['hit_list']},{b},{one}>"
        return f"k_force<{b},{one}>" if tuning["force_lanes"] == 1 else f"k_force_coop<{tuning['force_lanes']}>"
    if stage == "density":
        if tuning["hit_masks"]:
            return f"k_density_masks<{b},{one}>"
        return f"k_density<{b},{one}>" if tuning["density_lanes"] == 1 else \
            f"k_density_coop<{tuning['density_lanes']}>"
    if stage == "reorder":
        return f"k_reorder<{'false' if one_scene else 'true'}>"
    return f"k_{stage}"


def profiled_traffic(config: str, kernel: str):
    """HBM bytes per launch of a kernel from the committed rocprofv3 summary
    (calibrated FETCH_SIZE / WRITE_SIZE), or None."""
    path = os.path.join(ROOT, "profiles", f"plain_{config}_step_kernels.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            for item in line.split(":", 1)[1].split():
                k, v = item.split("=")
                if k == kernel:
                    return int(v)
    return None


def profiled_busy(config: str, kernel: str):
    """VALU / texture-address / texture-data busy fractions of a kernel from
    the committed rocprofv3 summary table, or None."""
    path = os.path.join(ROOT, "profiles", f"plain_{config}_step_kernels.md")
    if not os.path.exists(path):
        return None
    head = None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if line.startswith("| kernel |"):
            head = cells
        elif head and line.startswith("|") and cells and cells[0] == kernel:
            row = dict(zip(head, cells))
            try:
                return {"valu": float(row["VALU busy"]), "ta": float(row["TA busy"]),
                        "td": float(row["TD busy"])}
            except (KeyError, ValueError):
                return None
    return None


def time_handles(sims, steps: int) -> float:
    """Wall time per step of separate handles stepped round-robin (each on
    its own stream), synchronised at both ends."""
    for s in sims:
        s.sync()
    t = time.perf_counter()
    for _ in range(steps):
        for s in sims:
            s.Advance(1)
    for s in sims:
        s.sync()
    return (time.perf_counter() - t) / steps * 1e3


def flow_states(S, sim, sc, done, points=(1000, 2000), steps=50):
    """Step time of the 1M dam break later in the run: at each point the
    variants are autotuned on the live state and `steps` steps replay the
    captured step graph.  The column compresses as it collapses (mean
    density ~1155 kg/m^3 at step 1000, ~2100 at step 2000), so each particle
    has more neighbours than near rest density, where the headline is
    measured."""
    out = {}
    for p in points:
        if p <= done:
            continue
        sim.Advance(p - done)
        sim.sync()
        tun = sim.autotune(reps=2)
        sim.Advance(2)
        sim.sync()
        ms = sim.time_steps(steps) / steps
        done = p + 2 + steps
        rho = sim.read(S.Field.DENSITY)
        out[f"step_{p}"] = {"value": round(sc.n / (ms / 1e3), 1), "ms_per_step": round(ms, 5),
                            "mean_density": round(float(rho.mean()), 1),
                            "p99_density": round(float(np.percentile(rho, 99)), 1), "kernel_variants": tun}
        log(f"flow at step {p}: {ms:.4f} ms/step, mean density {rho.mean():.1f}, variants {tun}")
    out["note"] = ("the 1M simulation continued into the collapse; variants autotuned at each point, "
                   f"{steps} graph-replayed steps")
    return out


def bench_forms(S, a, plain_ms: float, sc) -> dict:
    from cudasphnew_amd import neural as N
    forms = {}
    # 1M scene with the learned boundary
    t0 = time.time()
    net = N.train_for_scene(sc.desc, steps=a.train_steps)
    train_ms = net.time_train(100, 65536) / 100        # 100 more steps at the final learning rate
    sim = S.ParticleSimulation(sc).Init()
    rep = sim.attach_neural(net)
    sim.Advance(max(a.warmup, 1))
    tun = sim.autotune(reps=3)
    sim.use_graph(not a.no_graph)
    sim.Advance(2)
    ms = sim.time_steps(a.steps) / a.steps
    sim.sync()
    forms["plain_1m_neural"] = {
        "value": round(sc.n / (ms / 1e3), 1), "ms_per_step": round(ms, 5), "particles": sc.n,
        "vs_plain": round(plain_ms / ms, 4), "train_steps": a.train_steps,
        "cells": {"none": rep["cells_none"], "network": rep["cells_neural"], "grid": rep["cells_grid"]},
        "band_mean_error_h": round(rep["mean_error"] / float(sc.desc.h), 5), "kernel_variants": tun,
        "train_ms_per_step": round(train_ms, 5),
        "train_roofline": {"bound": "mfma", "kernel": "k_mlp_train + k_mlp_adam",
                           "achieved": round(MLP_TRAIN_FLOP * 65536 / (train_ms / 1e3) / 1e12, 2),
                           "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                           "frac": round(MLP_TRAIN_FLOP * 65536 / (train_ms / 1e3) / 1e12 / MFMA_BF16_PEAK_TFLOPS, 5),
                           "dtype": "bf16 (FP32 accumulate)"},
    }
    log(f"1M + learned boundary: {ms:.4f} ms/step ({time.time() - t0:.0f}s incl. training)")
    del sim
    # mesh -> signed distance field, the reference's example configuration
    # (main.cpp:91-103): icosphere fitted to [-0.75, 0.75]^3, compact support
    # 0.017, dx = support / 3, rest distance = 2 support; 266^3 nodes
    import numpy as np
    from cudasphnew_amd import meshsdf as MS
    meshes = np.load(os.path.join(ROOT, "tests", "golden", "meshes.npz"))
    for name in ("icosphere", "bunny"):
        v = MS.fit(meshes[f"{name}_vertices"], (-0.75,) * 3, (0.75,) * 3)
        fcs = meshes[f"{name}_faces"]
        cs = np.float32(0.017)
        cfgm = MS.config(float(cs * np.float32(2)), float(cs), float(cs / np.float32(3)))
        MS.generate(v, fcs, cfgm)                      # warm-up
        ms_runs = [MS.generate(v, fcs, cfgm)["ms"] for _ in range(3)]
        bm = MS.generate(v, fcs, cfgm)
        nodes = int(np.prod(bm["dims"]))
        entry = {"mesh": name, "triangles": int(fcs.shape[0]), "dims": list(bm["dims"]), "nodes": nodes,
                 "nodes_within_maxdist": int(bm["present"].sum()), "ms": round(min(ms_runs), 4),
                 "nodes_per_s": round(nodes / (min(ms_runs) / 1e3), 1)}
        if name == "icosphere" and not a.no_cpu_baseline:
            from oracle import oracle as O
            t1 = time.perf_counter()
            O.mesh_sdf(v, fcs, cfgm.rest_distance, cfgm.compact_support, cfgm.dx, bm["dims"], bm["lo"], bm["hi"])
            entry["cpu_oracle_ms"] = round((time.perf_counter() - t1) * 1e3, 1)
            entry["cpu_oracle"] = "FP64 restatement (oracle/mesh_oracle.cpp), OpenMP"
        forms[f"boundary_map_{name}"] = entry
        log(f"boundary map {name}: {entry['dims']} in {entry['ms']} ms")
    # two-scale step (AdvanceTwoScale): 262,144 coarse particles with the
    # split region around the sphere
    sc2 = S.make_scene("two_scale")
    sim2 = S.ParticleSimulation(sc2).Init()
    sim2.Advance(1)
    tun2 = sim2.autotune(reps=3)            # coarse density/force variants
    sim2.AdvanceTwoScale(max(a.warmup, 2))
    sim2.sync()
    # eager launches (the library runs two-scale steps eagerly: replayed from
    # a graph they measured slower)
    ms2 = sim2.time_steps(a.steps, two_scale=True) / a.steps
    sim2.enable_timing(True)
    ms2_staged = sim2.time_steps(a.steps, two_scale=True) / a.steps
    st2 = sim2.stage_times()
    sim2.enable_timing(False)
    c2 = sim2.counts()
    forms["two_scale_262k"] = {
        "value": round(sc2.n / (ms2 / 1e3), 1), "ms_per_step": round(ms2, 5),
        "ms_per_step_stage_events": round(ms2_staged, 5),
        "particles": sc2.n,
        "sub_particles": int(c2.num_sub), "split": int(c2.num_split), "boundary": int(c2.num_boundary),
        "kernel_variants": tun2,
        "stage_ms": {k: round(v, 5) for k, v in st2.items() if v > 0},
        "note": "coarse particle-updates/s of AdvanceTwoScale (sub-particle pass included), eager launches; "
                "stage times from a run with events",
    }
    log(f"two-scale 262K: {ms2:.4f} ms/step (with stage events {ms2_staged:.4f}), {c2.num_sub} sub-particles")
    del sim2
    # 64 scenes
    t0 = time.time()
    scenes = [S.make_scene("batch", index=i, seed=1) for i in range(64)]
    n = sum(x.n for x in scenes)
    log(f"64 scenes, {n} particles ({time.time() - t0:.0f}s)")
    sims = [S.ParticleSimulation(x).Init() for x in scenes]
    for x in sims:
        x.Advance(max(a.warmup, 1))
    ms_h = time_handles(sims, a.batch_steps)
    forms["handles_64"] = {"value": round(n / (ms_h / 1e3), 1), "ms_per_step": round(ms_h, 5), "particles": n,
                           "note": "64 separate handles stepped round-robin, wall clock"}
    log(f"64 handles: {ms_h:.4f} ms/step")
    for x in sims:
        x.close()
    del sims
    for neural in (False, True):
        b = S.Batch(scenes)
        if neural:
            t1 = time.time()
            bad = 0
            for i, x in enumerate(scenes):
                nt = N.train_for_scene(x.desc, steps=a.batch_train_steps, n_samples=1 << 20)
                rep = b.attach_neural(i, nt, val=N.validation_samples(x.desc, 1 << 18))
                bad += rep["cells_grid"]
            log(f"64 networks trained and attached in {time.time() - t1:.0f}s ({bad} grid-fallback cells)")
        b.use_graph(True)
        b.step(max(a.warmup, 1))
        tun = b.autotune(reps=2)
        b.step(2)
        ms_b = b.time_steps(a.batch_steps) / a.batch_steps
        b.sync()
        key = "batched_64_neural" if neural else "batched_64"
        forms[key] = {"value": round(n / (ms_b / 1e3), 1), "ms_per_step": round(ms_b, 5), "particles": n,
                      "vs_handles": round(ms_h / ms_b, 4), "vs_plain_1m": round((n / ms_b) / (1048576 / plain_ms), 4),
                      "kernel_variants": tun, "graph": True}
        if neural:
            forms[key]["train_steps_per_scene"] = a.batch_train_steps
        log(f"{key}: {ms_b:.4f} ms/step")
        b.close()
    return forms


# Algorithmic FLOPs of one training sample: forward 2 (99*64 + 64*64 + 64*64
# + 64) = 29,184, training about three times the forward.
MLP_FWD_FLOP = 2 * (99 * 64 + 64 * 64 + 64 * 64 + 64)
MLP_TRAIN_FLOP = 3 * MLP_FWD_FLOP
MFMA_BF16_PEAK_TFLOPS = 2500.0


def train_profile(S, a) -> int:
    from cudasphnew_amd import neural as N
    sc = S.make_scene(a.config)
    h = float(sc.desc.h)
    net = N.NeuralSDF.for_scene(sc.desc, seed=1, out_scale=float(sc.desc.rest_distance))
    net.set_samples(N.grid_samples(sc.desc, 1 << 22, near_frac=0.3, near_width=3 * h, seed=2))
    net.train(5, 65536)
    ms = net.time_train(a.train_profile, 65536) / a.train_profile
    tflops = MLP_TRAIN_FLOP * 65536 / (ms / 1e3) / 1e12
    print(json.dumps({"mlp_train_ms_per_step": round(ms, 5), "batch": 65536,
                      "algorithmic_tflops": round(tflops, 2), "mfma_bf16_peak_tflops": MFMA_BF16_PEAK_TFLOPS,
                      "frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 5)}), flush=True)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="dam1m", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="run this many steps with no timing events and exit (for counter passes)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 256 MiB device copy (calibrates rocprof byte counters)")
    ap.add_argument("--no-forms", action="store_true", help="skip the learned-boundary and batched forms")
    ap.add_argument("--no-graph", action="store_true", help="time the 1M steps eagerly instead of as graph replays")
    ap.add_argument("--train-profile", type=int, default=0,
                    help="train the 1M scene's network for this many steps (batch 65536), report and exit")
    ap.add_argument("--two-scale-profile", type=int, default=0,
                    help="run this many two-scale steps of the 262,144-particle two-scale workload and exit "
                         "(for kernel traces of the sub-particle pass)")
    ap.add_argument("--batch-profile", type=int, default=0,
                    help="run this many batched steps of the 64-scene workload (graph replay, autotuned) and exit "
                         "(for kernel traces of the batched step)")
    ap.add_argument("--train-steps", type=int, default=20000, help="training steps of the 1M scene's network")
    ap.add_argument("--batch-train-steps", type=int, default=2000, help="training steps per batch scene network")
    ap.add_argument("--batch-steps", type=int, default=50, help="timed steps of the batched forms")
    ap.add_argument("--tuning", default="auto",
                    help="density/force kernel variants: 'auto' (autotune after the warm-up) or "
                         "'D,F,M,C[,B]' = density lanes, force lanes, hit masks (0/1), hit-list capacity, "
                         "workgroup size (128 or 256; default 256)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the step shards by spatial slabs only in design)", file=sys.stderr)
        return 2

    import cudasphnew_amd as S

    if a.train_profile:
        return train_profile(S, a)
    if a.copy_probe:
        copy_probe(256 << 20)
    if a.batch_profile:
        scenes = [S.make_scene("batch", index=i, seed=1) for i in range(64)]
        b = S.Batch(scenes)
        b.use_graph(True)
        b.step(3)
        b.sync()
        if a.tuning == "auto":
            tun = b.autotune(reps=2)
        else:
            b.set_tuning(*(int(v) for v in a.tuning.split(",")))
            tun = b.get_tuning()
        b.step(a.batch_profile)
        b.sync()
        log(f"batch profile: {a.batch_profile} steps, 64 scenes, {sum(x.n for x in scenes)} particles, {tun}")
        return 0
    if a.two_scale_profile:
        sc2 = S.make_scene("two_scale")
        sim2 = S.ParticleSimulation(sc2).Init()
        sim2.Advance(1)
        sim2.autotune(reps=3)
        sim2.AdvanceTwoScale(a.two_scale_profile)
        sim2.sync()
        c2 = sim2.counts()
        log(f"two-scale profile: {a.two_scale_profile} steps, {sc2.n} coarse, {c2.num_split} split, "
            f"{c2.num_boundary} boundary, {c2.num_sub} sub-particles")
        return 0
    t0 = time.time()
    sc = S.make_scene(a.config)
    log(f"scene {a.config}: {sc.n} particles, grid {tuple(sc.params.grid_dim)}, "
        f"boundary {sc.boundary_dims} ({time.time() - t0:.1f}s)")
    sim = S.ParticleSimulation(sc).Init()
    sim.Advance(max(a.warmup, 1))
    sim.sync()
    log(f"warm-up {a.warmup} steps done")
    if a.tuning == "auto":
        tuning = sim.autotune(reps=3)
    else:
        sim.set_tuning(*(int(v) for v in a.tuning.split(",")))
        tuning = sim.get_tuning()
    sim.Advance(2)
    sim.sync()
    log(f"kernel variants: {tuning}")
    if a.profile_steps:
        sim.Advance(a.profile_steps)
        sim.sync()
        log(f"profile: {a.profile_steps} steps done")
        return 0
    # the timed steps replay the captured step graph; the per-stage
    # breakdown comes from a second run of the same number of steps with an
    # event around every stage (launched eagerly: events split the graph)
    sim.use_graph(not a.no_graph)
    sim.Advance(2)
    sim.sync()
    ms = sim.time_steps(a.steps)
    sim.enable_timing(True)
    ms_staged = sim.time_steps(a.steps)
    stages = sim.stage_times()
    sim.enable_timing(False)
    log(f"timed {a.steps} steps: {ms:.2f} ms ({'graph' if not a.no_graph else 'eager'}); "
        f"with stage events {ms_staged:.2f} ms; stages (ms/step): "
        + ", ".join(f"{k}={v:.4f}" for k, v in stages.items() if v > 0))

    n = sc.n
    value = n * a.steps / (ms / 1e3)
    keys = sim.read(S.Field.SORTED_KEYS)
    tc = touched_cells(sc.params, keys)
    single = {k: v for k, v in stages.items() if k in STAGE_BYTES_PER_PARTICLE and v > 0}
    dom = max(single, key=single.get)
    tex = touched_texels(sc, sim)
    algo = {k: STAGE_BYTES_PER_PARTICLE[k] * n for k in single}
    if "density" in algo:
        algo["density"] += 8 * tc + 4 * tex["T_rho"]
    if "force" in algo:
        algo["force"] += 8 * tc + 4 * (tex["T_D"] + tex["T_nu"])
    algo_bytes = algo[dom]
    t_dom = single[dom] / 1e3
    achieved = algo_bytes / t_dom / 1e9
    kname = kernel_name(dom, tuning)
    roofline = {
        "bound": "hbm",
        "kernel": kname,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": profiled_traffic(a.config, kname),
        # the step's neighbour kernels are bound by arithmetic and by the
        # texture path of their gathers, not by HBM bytes: busy fractions
        # of the VALU and texture units from the same profile
        "profiled_busy": profiled_busy(a.config, kname),
        "algorithmic_bytes_per_launch": int(algo_bytes),
        "avg_launch_ms": round(single[dom], 5),
        "timing": "an event pair on the handle's stream around each stage of eager steps (each interval "
                  "includes the stage's dispatch gap; rocprofv3's in-step kernel durations are in profiles/)",
        "touched_cells": tc,
        "touched_texels": tex,
        # every neighbour pass and streaming pass of the step on the same footing
        "stages": {k: {"kernel": kernel_name(k, tuning), "avg_launch_ms": round(single[k], 5),
                       "algorithmic_bytes": int(algo[k]),
                       "frac": round(algo[k] / (single[k] / 1e3) / 1e9 / HBM_PEAK_GBS, 5),
                       "traffic": profiled_traffic(a.config, kernel_name(k, tuning)),
                       "profiled_busy": profiled_busy(a.config, kernel_name(k, tuning))}
                   for k in single},
    }
    forms = None
    if not a.no_forms:
        # the same simulation continued into the collapse (steps so far:
        # warm-up, 2 + 2 settling steps, the graph-timed and the event-timed
        # runs)
        flow = flow_states(S, sim, sc, max(a.warmup, 1) + 4 + 2 * a.steps)
        del sim
        forms = bench_forms(S, a, ms / a.steps, sc)
        forms["plain_1m_flow"] = flow
    cpu = None
    if not a.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", min(16, os.cpu_count() or 1)))
        log(f"cpu baseline: FP32 transcription, {threads} threads, ~{a.cpu_seconds}s")
        cpu = cpu_baseline(sc, a.cpu_seconds, threads)
        log(f"cpu baseline {cpu['value']:.4g} particle-updates/s")
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms / a.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[a.config], "form": "plain", "particles": n,
                   "kernel_variants": tuning},
        "stage_ms": {k: round(v, 5) for k, v in stages.items() if v > 0},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
