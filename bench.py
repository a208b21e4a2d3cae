"""Benchmark: batched multi-pose collision vs per-pose collideWith at 512^3, PBA EDT Gvox/s.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one batched collision check of 1024 poses of the 7-link DH arm
(157,556 link points) against a 512^3 BitVectorVoxelMap built from a 640x480
depth image (cfg4).  Inputs (joint values) and outputs (counts, masks) stay in
HBM during the timed region.  The line also carries the per-pose collideWith
rates (robot as a BitVectorVoxelMap and as a BitVectorVoxelList) on the same GPU,
the PBA EDT on the 512^3 and 1024^3 fields, the clearance cost forward and
backward, the sensor scan and the other kernels of the path, each with its
rate against HBM peak, and the CPU oracle timed on this host.

Roofline fractions: `frac` of the batched call is measured HBM traffic per call
(rocprofv3 FETCH_SIZE / WRITE_SIZE, committed under profiles/) over its kernel
time: the call is latency- and cache-bound, and the exhaustive test's bytes it
skips are reported as an effective figure only.  The EDT reports both the
24 B/voxel algorithmic fraction and the measured-traffic fraction per field.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "gpu-voxels_amd", "python"))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profiled_traffic(key):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by the copy-kernel
    calibration), read from the committed summary; None when the workload has not been profiled."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 2 and f[0] == key:
            return float(f[-1])
    return None


def edt_field(torch, scenes, edge, which):
    """Packed DistanceVoxel obstacle field (int32 view of uint32) on the GPU: the cfg1 occupancy at the
    origin, or i.i.d. voxels of the given density (seeds 31/32/33 for 0.1%/1%/10%)."""
    d3 = (edge, edge, edge)
    uninit = 0xFFFFFFFF if edge == 1024 else 0x3FFFFFFF
    if which == "cfg3a":
        ids = scenes.points_to_voxel_ids(scenes.cfg1()["env_points"], d3, 0.01)
    else:
        ids = scenes.cfg3_random(d3, which, {0.001: 31, 0.01: 32, 0.1: 33}[which])
    ids = torch.from_numpy(np.asarray(ids, np.int64)).cuda()
    packed = (ids % edge) | (((ids // edge) % edge) << 10) | ((ids // (edge * edge)) << 20)
    src = torch.full((edge ** 3,), uninit - (1 << 32) if uninit >= (1 << 31) else uninit, dtype=torch.int32,
                     device="cuda")
    src[ids] = (packed - ((packed >> 31) << 32)).to(torch.int32)
    torch.cuda.synchronize()
    return src


def scope_ms(gvx, torch, scopes, fn, reps=5, setup=None):
    """Average GPU time (HIP events on the library's stream) of the named scopes over `reps` calls
    of fn, after one untimed call; setup runs before every call, outside the scopes."""
    for r in range(reps + 1):
        if setup is not None:
            setup()
        torch.cuda.synchronize()
        if r == 1:
            gvx.profile_reset()
        fn()
    torch.cuda.synchronize()
    return sum(gvx.profile_get(s)[0] for s in scopes)


def bench_paths(torch, gvx, scenes, c2, env, dev, cpu=True):
    """Voxelization, clear, map x map and list x map collision, list insert with sort + dedup,
    each with its algorithmic bytes per unit (12 B point read + 8 B word RMW per voxelized point,
    sizeof(voxel) per cleared voxel, sizeof(self) + sizeof(other) per collided voxel, 4 B id + 32 B
    entry + the gathered map voxel per list entry, 8-bit radix passes x 20 B + 96 B gather per
    sorted element).  With cpu, the oracle's serial voxelization and map x map collision run on the
    same inputs (1 core) and are checked against the GPU maps and counts."""
    out = []
    dims, side = c2["dims"], c2["side"]
    nvox = dims[0] * dims[1] * dims[2]

    def rec(name, workload, ms, units, unit, bpu, extra=None, needed=None):
        """needed: bytes the kernel cannot avoid moving when it skips the second map's voxel where the
        first one already rules a collision out; the fraction of peak is then taken on those bytes"""
        moved = needed if needed is not None else bpu * units
        gbs = moved / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        r = {"path": name, "workload": workload, "ms": round(ms, 4), "units": units, "unit": unit,
             "bytes_per_unit": bpu, "rate_per_s": round(units / (ms * 1e-3), 1) if ms > 0 else None,
             "gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
        if needed is not None:
            r["bytes_needed"] = int(needed)
            r["alg_gbs"] = round(bpu * units / (ms * 1e-3) / 1e9, 1)
        if extra:
            r.update(extra)
        out.append(r)
        log(f"bench: {name}: {ms:.4f} ms, {units} {unit}s, {gbs:.0f} GB/s ({100 * gbs / HBM_PEAK_GBS:.1f}% of peak)")

    pts = torch.from_numpy(c2["env_points"]).to(dev)
    npts = int(c2["env_points"].shape[0])
    prob = gvx.Map(gvx.PROBAB_VOXELMAP, dims, side)
    ms = scope_ms(gvx, torch, ["insert_points"], lambda: prob.insert_points(pts, gvx.eBVM_OCCUPIED), setup=prob.clear)
    extra = None
    threads = max(1, min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "64"))))
    if cpu:
        from oracle import oracle as O
        om = O.new_map(O.PROB_MAP, dims)
        t0 = time.perf_counter()
        O.map_insert_points(O.PROB_MAP, om, dims, side, c2["env_points"], gvx.eBVM_OCCUPIED)
        cs = time.perf_counter() - t0
        gm = prob.read()
        extra = {"cpu_baseline": {"ms": round(cs * 1e3, 2), "cores": 1, "kind": "port",
                                  "bit_exact": bool(np.array_equal(om, gm))}}
        # the same oracle with the points bucketed by voxel range over every core
        om = O.new_map(O.PROB_MAP, dims)
        t0 = time.perf_counter()
        O.map_insert_points(O.PROB_MAP, om, dims, side, c2["env_points"], gvx.eBVM_OCCUPIED, threads=threads)
        cs = time.perf_counter() - t0
        extra["cpu_baseline"]["all_cores"] = {"ms": round(cs * 1e3, 2), "cores": threads,
                                              "bit_exact": bool(np.array_equal(om, gm))}
        del om, gm
    rec("insertPointCloud -> ProbVoxelMap", f"cfg2 depth points ({npts}) into 512^3", ms, npts, "point", 20, extra)
    bit = gvx.Map(gvx.BITVECTOR_VOXELMAP, dims, side)
    ms = scope_ms(gvx, torch, ["insert_points"], lambda: bit.insert_points(pts, gvx.eBVM_OCCUPIED), setup=bit.clear)
    rec("insertPointCloud -> BitVectorVoxelMap", f"cfg2 depth points ({npts}) into 512^3", ms, npts, "point", 20)
    del pts

    arm = c2["arm"]
    nlink = int(arm["link_sizes"].sum())
    meanings = np.arange(4, 4 + len(arm["link_sizes"]), dtype=np.uint32)
    robot = gvx.Map(gvx.BITVECTOR_VOXELMAP, dims, side)
    ms = scope_ms(gvx, torch, ["insert_robot"],
                  lambda: gvx.map_insert_robot(robot, arm["link_points"], arm["link_sizes"], arm["dh"], c2["q"],
                                               c2["base"], meanings), setup=robot.clear)
    rec("FK + insertMetaPointCloud -> BitVectorVoxelMap", f"cfg2 7-link arm ({nlink} points), meanings 4..10",
        ms, nlink, "point", 24)

    ms = scope_ms(gvx, torch, ["clear_map"], bit.clear)
    rec("clearMap BitVectorVoxelMap", "512^3 (4 GiB)", ms, nvox, "voxel", 32)
    ms = scope_ms(gvx, torch, ["clear_map"], prob.clear)
    rec("clearMap ProbVoxelMap", "512^3 (128 MiB)", ms, nvox, "voxel", 1)

    prob.insert_points(torch.from_numpy(c2["env_points"]).to(dev), gvx.eBVM_OCCUPIED)
    n_pass = int((prob.read() >= 127).sum())           # threshold 1.0 -> probability 127
    hits = []
    ms = scope_ms(gvx, torch, ["collide"], lambda: hits.append(robot.collide(prob, 1.0)))
    extra = {"collisions": int(hits[-1]), "prob_voxels_passing": n_pass}
    if cpu:
        from oracle import oracle as O
        hr, hp = robot.read(), prob.read()
        t0 = time.perf_counter()
        oc = O.map_collide(O.BITVECTOR_MAP, hr, O.PROB_MAP, hp, dims, 1.0)
        cs = time.perf_counter() - t0
        extra["cpu_baseline"] = {"ms": round(cs * 1e3, 1), "cores": 1, "kind": "port",
                                 "gvox_per_s": round(nvox / cs / 1e9, 3), "count_bit_exact": bool(oc == hits[-1])}
        hr = robot.read()  # the serial call marked its copy; the voxels over every core
        t0 = time.perf_counter()
        oc = O.map_collide(O.BITVECTOR_MAP, hr, O.PROB_MAP, hp, dims, 1.0, threads=threads)
        cs = time.perf_counter() - t0
        extra["cpu_baseline"]["all_cores"] = {"ms": round(cs * 1e3, 1), "cores": threads,
                                              "gvox_per_s": round(nvox / cs / 1e9, 3),
                                              "count_bit_exact": bool(oc == hits[-1])}
        del hr, hp
    rec("collideWith BitVector x Prob", "cfg2 robot vs cfg2 environment, 512^3, threshold 1.0", ms, nvox, "voxel",
        33, extra, needed=nvox + 32 * n_pass)
    hits.clear()
    ms = scope_ms(gvx, torch, ["collide"], lambda: hits.append(robot.collide(env, 1.0)))
    rec("collideWith BitVector x BitVector", "cfg2 robot vs cfg4 environment, 512^3", ms, nvox, "voxel", 64,
        {"collisions": int(hits[-1])}, needed=32 * nvox)
    hits.clear()
    ms = scope_ms(gvx, torch, ["collide_types"], lambda: hits.append(robot.collide_types(prob, 1.0)))
    rec("collideWithTypes BitVector x Prob", "cfg2 robot vs cfg2 environment, 512^3", ms, nvox, "voxel", 33,
        {"collisions": int(hits[-1][0])}, needed=nvox + 32 * n_pass)
    hits.clear()
    ms = scope_ms(gvx, torch, ["collide_types"], lambda: hits.append(robot.collide_types(env, 1.0)))
    rec("collideWithTypes BitVector x BitVector", "cfg2 robot vs cfg4 environment, 512^3", ms, nvox, "voxel", 64,
        {"collisions": int(hits[-1][0])}, needed=32 * nvox)
    del robot, bit, prob
    torch.cuda.empty_cache()

    c1 = scenes.cfg1()
    d1, s1 = c1["dims"], c1["side"]
    lp = torch.from_numpy(c1["list_points"]).to(dev)
    nl = int(c1["list_points"].shape[0])
    holder = {}

    def fresh_list():
        holder["l"] = gvx.VoxelList(d1, s1)

    def list_insert():
        holder["l"].insert_points(lp, sizes=c1["list_sizes"], meanings=c1["list_meanings"])

    ms = scope_ms(gvx, torch, ["list_insert_points", "list_make_unique"], list_insert, setup=fresh_list)
    passes = 3                                   # 8-bit digits of ids below 2^21 (128^3)
    rec("VoxelList insert + sort + dedup", f"cfg1 list: {nl} points on 5 spheres, meanings 1-5, 128^3", ms, nl,
        "element", passes * 20 + 96, {"entries": len(holder["l"]),
                                     "note": "the unique count is read once, after the call's one synchronisation"})
    lst = holder["l"]
    m1 = gvx.Map(gvx.BITVECTOR_VOXELMAP, d1, s1)
    m1.insert_points(c1["env_points"], 0)
    hits.clear()
    ms = scope_ms(gvx, torch, ["list_collide_map"], lambda: hits.append(lst.collide_map(m1, 1.0)))
    rec("VoxelList collideWith BitVectorVoxelMap", f"cfg1 list ({len(lst)} entries) vs cfg1 128^3 map", ms,
        len(lst), "entry", 68, {"collisions": int(hits[-1][0])})
    big = torch.from_numpy(c2["env_points"]).to(dev)

    def big_insert():
        holder["l"].insert_points(big, meaning=gvx.eBVM_OCCUPIED)

    holder["l"] = None
    ms = scope_ms(gvx, torch, ["list_insert_points", "list_make_unique"], big_insert,
                  setup=lambda: holder.__setitem__("l", gvx.VoxelList(dims, side)))
    passes = 4                                   # ids below 2^27 (512^3)
    rec("VoxelList insert + sort + dedup", f"cfg2 depth points ({npts}) into a 512^3 list", ms, npts, "element",
        passes * 20 + 96, {"entries": len(holder["l"]), "note": "the unique count is read once, after the call's one synchronisation"})
    del lst, m1, holder, lp, big
    torch.cuda.empty_cache()
    return out


EDT_SEEDS = {0.001: 31, 0.01: 32, 0.1: 33}


def edt_key(edge, which):
    return f"pba_{'cfg3a' if which == 'cfg3a' else f'{which * 100:g}pct'}_{edge}"


def bench_edt(torch, gvx, scenes, edges):
    """PBA EDT on the cfg3 fields: HIP-event time of the whole call and of each pass, Gvox/s, the
    24 B/voxel algorithmic fraction of HBM peak and, where profiled, the measured-traffic fraction."""
    sweep = []
    for edge in edges:
        for which in ("cfg3a", 0.001, 0.01, 0.1):
            src = edt_field(torch, scenes, edge, which)
            d3 = (edge, edge, edge)
            dist = gvx.Map(gvx.DISTANCE_VOXELMAP, d3, 0.01)
            reps = 10 if edge == 512 else 5
            for r in range(reps + 1):
                dist.write(src)                      # device-to-device: PBA works in place
                if r == 1:
                    gvx.profile_reset()
                dist.pba()
            pba_ms, _ = gvx.profile_get("pba")
            parts = {k2: round(gvx.profile_get(k2)[0], 4) for k2 in ("pba_z", "pba_y", "pba_x")}
            nvox = edge ** 3
            gbs = 24.0 * nvox / (pba_ms * 1e-3) / 1e9
            key = edt_key(edge, which)
            traffic = profiled_traffic(key)
            rec = {"grid": f"{edge}^3", "obstacles": which if which == "cfg3a" else f"{which:.1%} i.i.d.",
                   "ms": round(pba_ms, 4), "gvox_per_s": round(nvox / (pba_ms * 1e-3) / 1e9, 2),
                   "alg_bytes": 24 * nvox, "alg_frac": round(gbs / HBM_PEAK_GBS, 4),
                   "traffic": traffic, "traffic_key": key,
                   "measured_frac": (round(traffic / (pba_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                     if traffic else None),
                   "pass_ms": parts}
            sweep.append(rec)
            log(f"bench: EDT {edge}^3 {rec['obstacles']}: {pba_ms:.3f} ms {parts}, {rec['gvox_per_s']} Gvox/s, "
                f"{100 * rec['alg_frac']:.1f}% of HBM peak at 24 B/voxel")
            del dist, src
            torch.cuda.empty_cache()
    return sweep


def bench_cost(torch, gvx, scenes, dev):
    """cfg5: hinge-squared clearance cost on the 512^3 cfg3(a) EDT (float metres), 256 trajectories x
    32 waypoints x 7 links x 64 spheres.  Algorithmic bytes per sphere: 8 x 4 B trilinear gather + 16 B
    sphere + 4 B (forward) or 12 B dc/dx (backward)."""
    d = (512, 512, 512)
    side = 0.01
    ids = scenes.points_to_voxel_ids(scenes.cfg1()["env_points"], d, side)
    coords = np.stack([ids % 512, (ids // 512) % 512, ids // (512 * 512)], axis=1).astype(np.uint32)
    m = gvx.Map(gvx.DISTANCE_VOXELMAP, d, side)
    m.insert_coords(coords, 1)
    m.pba()
    grid = torch.empty(512 ** 3, device=dev, dtype=torch.float32)
    m.metric(grid)
    del m
    c5 = scenes.cfg5()
    cost = gvx.Cost(grid, d, side, c5["dh"], c5["spheres"], eps=c5["eps"], weight=c5["w"], base=c5["base"])
    q = torch.from_numpy(c5["q"]).to(dev)
    nt, nw, nl = c5["q"].shape
    spl = c5["spheres"].shape[1]
    n_sph = nt * nw * nl * spl
    out = torch.empty(nt, device=dev, dtype=torch.float32)
    dq = torch.empty_like(q)
    dxs = torch.empty((nt, nw, nl, spl, 3), device=dev, dtype=torch.float32)
    reps = 10
    for r in range(reps + 1):
        if r == 1:
            gvx.profile_reset()
        cost.forward(q, out)
    fwd_ms = gvx.profile_get("cost_forward")[0]
    for r in range(reps + 1):
        if r == 1:
            gvx.profile_reset()
        cost.backward(q, dq, dxs)
    bwd_ms = gvx.profile_get("cost_backward")[0]
    torch.cuda.synchronize()
    recs = {}
    for name, ms, bps in (("forward", fwd_ms, 52), ("backward", bwd_ms, 60)):
        alg = bps * n_sph
        traffic = profiled_traffic(f"cost_{name}_cfg5")
        recs[name] = {"ms": round(ms, 4), "spheres_per_s": round(n_sph / (ms * 1e-3), 1),
                      "roofline": {"bound": "hbm", "achieved": round(alg / (ms * 1e-3) / 1e9, 1),
                                   "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                   "alg_bytes_per_launch": alg, "traffic": traffic,
                                   "measured_frac": (round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                                     if traffic else None)}}
        log(f"bench: cost {name}: {ms:.4f} ms, {n_sph / (ms * 1e-3) / 1e9:.2f} G spheres/s, "
            f"{alg / (ms * 1e-3) / 1e9:.0f} GB/s algorithmic")
    res = {"workload": "cfg5: 256 trajectories x 32 waypoints x 7 links x 64 spheres on the 512^3 cfg3(a) EDT",
           "spheres": n_sph, "cost_sum": float(out.sum().item()), **recs}
    del cost, grid, q, out, dq, dxs
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--per-pose-sample", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--schedule", type=int, default=-1, help="batched launch schedule (default: autotune)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-edt", action="store_true")
    ap.add_argument("--no-cost", action="store_true")
    ap.add_argument("--no-sensor", action="store_true")
    ap.add_argument("--no-paths", action="store_true")
    ap.add_argument("--edt-1024", type=int, default=1, help="also run the 1024^3 EDT fields (1/0)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy this many MiB device-to-device 5 times (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU (the batched check shards by pose, the EDT by z slab, "
              "neither needs a second GPU here); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import gvx
    from gvx import scenes

    if args.copy_probe:
        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)                      # blit kernel: reads and writes the size once
            torch.add(src, 1.0, out=dst)        # elementwise kernel: 16-byte vector reads and writes
        torch.cuda.synchronize()
        log(f"copy probe: 5 x {args.copy_probe} MiB copy and 5 x elementwise add")
        return

    gvx.init(0)
    gvx.profile_enable(True)
    dev = torch.device("cuda:0")
    result = {
        "metric": "poses/s batched vs per-pose collideWith @512^3; EDT Gvox/s; % HBM peak",
        "value": None, "unit": "poses/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded depth-image scene and DH arm)",
    }

    # ------------------------------------------------------------ cfg4 workload
    log("bench: building cfg2 environment and cfg4 arm")
    c2 = scenes.cfg2()
    c4 = scenes.cfg4(n_poses=1024)
    dims, side = c2["dims"], c2["side"]
    env = gvx.Map(gvx.BITVECTOR_VOXELMAP, dims, side)
    env.insert_points(c2["env_points"], gvx.eBVM_OCCUPIED)
    arm = c4["arm"]
    n_pts = int(arm["link_sizes"].sum())
    n_poses = c4["q"].shape[0]
    counts = None
    if not args.no_batch:
        batch = gvx.Batch(env, arm["link_points"], arm["link_sizes"], arm["dh"], c4["base"], max_poses=n_poses)
        q_dev = torch.from_numpy(c4["q"]).to(dev)
        counts = torch.empty(n_poses, dtype=torch.int32, device=dev)
        masks = torch.empty((n_poses, 32), dtype=torch.uint8, device=dev)
        torch.cuda.synchronize()
        batch.collide(c4["q"])                # host call: the poses autotune measures on
        sched = batch.autotune(n_poses) if args.schedule < 0 else args.schedule
        log(f"bench: autotuned schedule {sched}")

        def step():
            # one batch of 1024 poses; calls queue on the library stream, the timed region
            # ends with one synchronisation
            batch.collide_async(q_dev, counts, masks)

        def timed(k):
            gvx.sync()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(k):
                step()
            gvx.sync()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) * 1e3 / k

        gvx.profile_enable(False)            # no event records between the timed calls
        sweep = []
        for sch in range(16):                # 0-7 direct launches, 8-15 graph replays
            batch.set_schedule(sch)
            timed(3)
            sweep.append(round(timed(20), 4))
        log(f"bench: ms per batch by schedule {sweep}")
        batch.set_schedule(sched)
        for _ in range(args.warmup):
            step()
        ms_step = timed(args.steps)
        gvx.profile_enable(True)             # GPU time of one call (prologue + collide) from HIP events
        gvx.profile_reset()
        timed(args.steps)
        kern_ms, launches = gvx.profile_get("batch_collide")
        poses_s = n_poses / (ms_step * 1e-3)
        hits = int((counts.cpu().numpy() > 0).sum())
        n_tests, n_live, n_live_pp = batch.last_stats()
        # latency of one synchronous call (device buffers, returns after the GPU work)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            batch.collide(q_dev, counts, masks)
        t1 = time.perf_counter()
        ms_sync = (t1 - t0) * 1e3 / args.steps
        log(f"bench: batched {ms_step:.4f} ms/step (GPU {kern_ms:.4f} ms, {launches} launches; synchronous call "
            f"{ms_sync:.4f} ms), {poses_s:.0f} poses/s, {hits} colliding poses; broad phase kept {n_live} of "
            f"{n_tests} (chunk, pose) pairs = {n_live_pp} of {n_pts * n_poses} point-poses")
        gpu_counts = counts.cpu().numpy().astype(np.uint32)

        # per-pose reference pattern on the same GPU: clearMap + FK + insertMetaPointCloud + collideWith
        robot = gvx.Map(gvx.BITVECTOR_VOXELMAP, dims, side)
        k = args.per_pose_sample
        qs = c4["q"][:k]
        batch.collide_per_pose(robot, qs[:2])  # warm-up
        gvx.profile_reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pp_counts = batch.collide_per_pose(robot, qs)
        t1 = time.perf_counter()
        pp_ms = (t1 - t0) * 1e3 / k
        pp_gpu_ms, _ = gvx.profile_get("per_pose_collide")
        pp_poses_s = 1e3 / pp_ms
        bit_exact = bool(np.array_equal(pp_counts, gpu_counts[:k]))
        log(f"bench: per-pose {pp_ms:.3f} ms/pose ({pp_gpu_ms:.3f} ms GPU), counts bit-exact: {bit_exact}")
        del robot

        # the same pattern with the robot as a BitVectorVoxelList: clear, FK + insertMetaPointCloud into
        # the list (voxelize, sort, dedup), VoxelList::collideWith(environment)
        meanings = np.arange(4, 4 + len(arm["link_sizes"]), dtype=np.uint32)
        lst = gvx.VoxelList(dims, side)
        lc = []

        def list_pose(qv):
            lst.clear()
            gvx.list_insert_robot(lst, arm["link_points"], arm["link_sizes"], arm["dh"], qv, c4["base"],
                                  meanings=meanings)
            return lst.collide_map(env, 1.0)[0]

        for qv in qs[:2]:
            list_pose(qv)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for qv in qs:
            lc.append(list_pose(qv))
        t1 = time.perf_counter()
        pl_ms = (t1 - t0) * 1e3 / k
        list_exact = bool(np.array_equal(np.array(lc, np.uint64), gpu_counts[:k].astype(np.uint64)))
        log(f"bench: per-pose list {pl_ms:.3f} ms/pose, counts bit-exact: {list_exact}")
        del lst

        # roofline of the batched call (prologue + k_batch_collide): measured HBM bytes per call
        # (rocprofv3) over its kernel time.  The exhaustive test's bytes (12 B link point + 4 B
        # bitmap word per point-pose) are an effective figure: the broad phase skips most of them.
        alg_bytes = 16.0 * n_pts * n_poses
        traffic = profiled_traffic("batch_cfg4")
        achieved = traffic / (kern_ms * 1e-3) / 1e9 if (traffic and kern_ms > 0) else None
        alg_gbs = alg_bytes / (kern_ms * 1e-3) / 1e9 if kern_ms > 0 else 0.0
        result.update({
            "value": round(poses_s, 1), "ms_per_step": round(ms_step, 4),
            "config": {
                "workload": "cfg4: 1024 poses x 7-link DH arm (%d link points) vs 512^3 BitVectorVoxelMap "
                            "from a 640x480 depth image (%d points), 1 cm voxels" % (n_pts, c2["env_points"].shape[0]),
                "poses": n_poses, "link_points": n_pts, "map": "512^3", "schedule": sched,
                "launch": ("graph replay" if sched >= 8 else "direct") +
                          " (autotuned over 8 kernel shapes x {direct, hipGraph replay})",
                "schedule_sweep_ms": sweep,
            },
            "per_pose": {"poses_per_s": round(pp_poses_s, 2), "ms_per_pose": round(pp_ms, 4),
                         "sample_poses": k, "counts_bit_exact": bit_exact,
                         "pattern": "clearMap + FK + insertMetaPointCloud + collideWith (BitVectorVoxelMap robot)"},
            "per_pose_list": {"poses_per_s": round(1e3 / pl_ms, 2), "ms_per_pose": round(pl_ms, 4),
                              "sample_poses": k, "counts_bit_exact": list_exact,
                              "pattern": "clear + FK + insertMetaPointCloud + collideWith (BitVectorVoxelList robot)"},
            "speedup_vs_per_pose": round(poses_s / pp_poses_s, 1),
            "speedup_vs_per_pose_list": round(poses_s * pl_ms * 1e-3, 1),
            "ms_per_call_sync": round(ms_sync, 4),
            "roofline": {"bound": "latency", "achieved": round(achieved, 1) if achieved else None,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                         "traffic": traffic, "traffic_key": "batch_cfg4",
                         "kernel": "batched collision call (k_batch_prologue + k_batch_collide)",
                         "kernel_ms": round(kern_ms, 4),
                         "effective": {"alg_bytes_per_launch": alg_bytes, "alg_gbs": round(alg_gbs, 1),
                                       "note": "bytes of the exhaustive test (16 B per point-pose); the broad "
                                               "phase skips most of them, so this is not a fraction of peak"},
                         "broad_phase": {"pairs_tested": n_tests, "pairs_kept": n_live,
                                         "point_poses_kept": n_live_pp, "point_poses": n_pts * n_poses}},
        })

    # ------------------------------------------------------------ cfg4 arm vs a dense environment
    if not args.no_batch:
        # the cfg3 10% i.i.d. field (seed 33, 12.8 M occupied voxels) as a 512^3 BitVector environment:
        # the call's cost must not grow with the environment (its hit matrix is cleared word by word)
        ids = scenes.cfg3_random(dims, 0.1, EDT_SEEDS[0.1])
        coords = np.stack([ids % 512, (ids // 512) % 512, ids // (512 * 512)], axis=1).astype(np.uint32)
        denv = gvx.Map(gvx.BITVECTOR_VOXELMAP, dims, side)
        denv.insert_coords(coords, gvx.eBVM_OCCUPIED)
        dbatch = gvx.Batch(denv, arm["link_points"], arm["link_sizes"], arm["dh"], c4["base"], max_poses=n_poses)
        dbatch.set_schedule(sched)
        dcounts = torch.empty(n_poses, dtype=torch.int32, device=dev)
        dmasks = torch.empty((n_poses, 32), dtype=torch.uint8, device=dev)
        gvx.profile_enable(False)
        for _ in range(args.warmup + 1):
            dbatch.collide_async(q_dev, dcounts, dmasks)
        gvx.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            dbatch.collide_async(q_dev, dcounts, dmasks)
        gvx.sync()
        torch.cuda.synchronize()
        d_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        gvx.profile_enable(True)
        d_tests, d_live, d_live_pp = dbatch.last_stats()
        # the call's time follows the point-poses the broad phase keeps (the collision work), not the
        # number of occupied environment voxels: compare the time per kept point-pose with cfg4's
        drec = {"workload": "cfg4 arm, 1024 poses, vs the 512^3 10% i.i.d. field (seed 33) as a BitVectorVoxelMap",
                "occupied_voxels": int(len(ids)), "schedule": sched, "ms_per_step": round(d_ms, 4),
                "poses_per_s": round(n_poses / (d_ms * 1e-3), 1), "ms_vs_cfg4": round(d_ms / ms_step, 3),
                "colliding_poses": int((dcounts.cpu().numpy() > 0).sum()),
                "broad_phase": {"pairs_tested": d_tests, "pairs_kept": d_live, "point_poses_kept": d_live_pp},
                "ns_per_kept_point_pose": round(d_ms * 1e6 / max(d_live_pp, 1), 5),
                "cfg4_ns_per_kept_point_pose": round(ms_step * 1e6 / max(n_live_pp, 1), 5),
                "traffic": profiled_traffic("batch_dense_env"), "traffic_key": "batch_dense_env"}
        if not args.no_cpu:
            from oracle import oracle as O
            bits = np.zeros(512 ** 3 // 8, np.uint8)
            np.bitwise_or.at(bits, ids // 8, (1 << (ids % 8)).astype(np.uint8))
            ns = 32
            thr = max(1, min(len(os.sched_getaffinity(0)), 16))
            oc, _ = O.batch_collide(bits, dims, side, arm["link_points"], arm["link_sizes"], arm["dh"], c4["base"],
                                    c4["q"][:ns], threads=thr)
            drec["counts_bit_exact"] = bool(np.array_equal(oc, dcounts.cpu().numpy()[:ns].astype(np.uint32)))
            drec["checked_poses"] = ns
            del bits
        result["batch_dense_env"] = drec
        log(f"bench: batched vs dense environment {d_ms:.4f} ms/step ({d_ms / ms_step:.2f}x cfg4), "
            f"{drec['colliding_poses']} colliding poses, {d_live_pp} kept point-poses "
            f"({drec['ns_per_kept_point_pose']} ns each; cfg4 {drec['cfg4_ns_per_kept_point_pose']} ns)" +
            (f", counts bit-exact on {drec['checked_poses']} poses: {drec['counts_bit_exact']}"
             if "counts_bit_exact" in drec else ""))
        del dbatch, denv, dcounts, dmasks, coords, ids
        torch.cuda.empty_cache()

    # ------------------------------------------------------------ EDT (cfg3 at 512^3 and 1024^3)
    if not args.no_edt:
        sweep = bench_edt(torch, gvx, scenes, (512, 1024) if args.edt_1024 else (512,))
        head = next(r for r in sweep if r["grid"] == "512^3" and r["obstacles"] == "1.0% i.i.d.")
        nvox = 512 ** 3
        result["edt"] = {"workload": "cfg3(b): PBA exact EDT, 512^3, 1% i.i.d. obstacles (seed 32): every pass "
                                     "does its full work (no empty region to skip)",
                         "ms": head["ms"], "gvox_per_s": head["gvox_per_s"],
                         "roofline": {"bound": "hbm", "achieved": round(24.0 * nvox / (head["ms"] * 1e-3) / 1e9, 1),
                                      "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": head["alg_frac"],
                                      "traffic": head["traffic"], "measured_frac": head["measured_frac"],
                                      "bytes_per_launch": 24.0 * nvox},
                         "pass_ms": head["pass_ms"]}
        result["edt_sweep"] = sweep

    # ------------------------------------------------------------ clearance cost (cfg5)
    if not args.no_cost:
        result["cost"] = bench_cost(torch, gvx, scenes, dev)

    # ------------------------------------------------------------ sensor scan (insertSensorData)
    if not args.no_sensor:
        cam = scenes.CAM_POS.astype(np.float32)
        pts = c2["env_points"]
        pm = gvx.Map(gvx.PROBAB_VOXELMAP, dims, side)
        pts_dev = torch.from_numpy(pts).to(dev)
        reps = 5
        for r in range(reps + 1):
            pm.clear()                                   # every scan starts from an unknown map
            if r == 1:
                gvx.profile_reset()
            pm.insert_sensor_data(pts_dev, cam)
        tgt_ms = gvx.profile_get("sensor_targets")[0]
        rays_ms = gvx.profile_get("sensor_rays")[0]
        apply_ms = gvx.profile_get("sensor_apply")[0]
        hits_ms = gvx.profile_get("sensor_hits")[0]
        # rays and cells walked (RayCaster visits 1 + |dx| + |dy| + |dz| cells per ray)
        vox = np.floor(pts / np.float32(side)).astype(np.int64)
        sv = np.floor(cam / np.float32(side)).astype(np.int64)
        ok = np.all(np.isfinite(pts), axis=1) & np.all((vox >= 0) & (vox < np.array(dims)), axis=1)
        n_rays = int(ok.sum())
        ray_len = 1 + np.abs(vox[ok] - sv).sum(axis=1)
        cells = int(ray_len.sum())
        vid = (vox[ok, 2] * dims[1] + vox[ok, 1]) * dims[0] + vox[ok, 0]
        _, first = np.unique(vid, return_index=True)   # one walk per distinct target voxel
        cells_walked = int(ray_len[first].sum())
        scan_ms = tgt_ms + rays_ms + apply_ms + hits_ms
        rec = {"workload": "cfg2 depth image (%d points) ray-cast from the camera into a 512^3 ProbVoxelMap" % len(pts),
               "ms": round(scan_ms, 4), "rays": n_rays, "cells": cells, "target_voxels": int(len(first)),
               "cells_walked": cells_walked,
               "rays_per_s": round(n_rays / (scan_ms * 1e-3), 1), "cells_per_s": round(cells / (rays_ms * 1e-3), 1),
               "pass_ms": {"sensor_targets": round(tgt_ms, 4), "sensor_rays": round(rays_ms, 4),
                           "sensor_apply": round(apply_ms, 4),
                           "sensor_hits": round(hits_ms, 4)}}
        if not args.no_cpu:
            from oracle import oracle as O
            om = O.new_map(O.PROB_MAP, dims)
            t0 = time.perf_counter()
            O.prob_insert_sensor(om, dims, side, cam, pts)
            cpu_s = time.perf_counter() - t0
            rec["cpu_oracle"] = {"ms": round(cpu_s * 1e3, 1), "cores": 1,
                                 "bit_exact": bool(np.array_equal(om, pm.read()))}
        result["sensor"] = rec
        log(f"bench: sensor scan {scan_ms:.3f} ms ({n_rays} rays, {cells} cells: "
            f"{cells / (rays_ms * 1e-3) / 1e9:.1f} G cells/s)" +
            (f", CPU oracle {rec['cpu_oracle']['ms']} ms, bit-exact {rec['cpu_oracle']['bit_exact']}"
             if "cpu_oracle" in rec else ""))
        del pm, pts_dev

    # ------------------------------------------------------------ the other kernels of the path
    if not args.no_paths:
        result["paths"] = bench_paths(torch, gvx, scenes, c2, env, dev, cpu=not args.no_cpu)

    # ------------------------------------------------------------ HBM ceiling (device copy)
    n = (1 << 30) // 4
    a_ = torch.ones(n, dtype=torch.float32, device=dev)
    b_ = torch.empty_like(a_)
    b_.copy_(a_)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        b_.copy_(a_)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 10 * 2 * (1 << 30) / (e0.elapsed_time(e1) * 1e-3) / 1e9
    result["hbm_copy_ceiling_gbs"] = round(copy_gbs, 1)
    log(f"bench: device copy 1 GiB: {copy_gbs:.0f} GB/s (read + write)")
    del a_, b_

    # ------------------------------------------------------------ CPU oracle baselines
    if not args.no_cpu:
        from oracle import oracle as O
        threads = max(1, min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "64"))))
        if counts is not None:
            env_ids = scenes.points_to_voxel_ids(c2["env_points"], dims, side)
            env_bits = np.zeros(512 ** 3 // 8, np.uint8)
            np.bitwise_or.at(env_bits, env_ids // 8, (1 << (env_ids % 8)).astype(np.uint8))
            done, t0 = 0, time.perf_counter()
            while done < n_poses and (time.perf_counter() - t0) < args.cpu_seconds:
                sl = c4["q"][done:done + 8]
                oc, _ = O.batch_collide(env_bits, dims, side, arm["link_points"], arm["link_sizes"], arm["dh"],
                                        c4["base"], sl)
                assert np.array_equal(oc, gpu_counts[done:done + 8]), "oracle mismatch"
                done += sl.shape[0]
            cpu_s = time.perf_counter() - t0
            result["cpu_baseline"] = {"value": round(done / cpu_s, 2), "unit": "poses/s", "cores": 1,
                                      "kind": "port",
                                      "sample": f"first {done} of the 1024 cfg4 poses, serial C oracle "
                                                f"({cpu_s:.1f} s), counts checked against the GPU",
                                      "host_cpus": os.cpu_count()}
            log(f"bench: CPU oracle {done / cpu_s:.2f} poses/s on 1 core")
            # the same oracle over all of this process's host cores (OpenMP over poses)
            t0 = time.perf_counter()
            oc, _ = O.batch_collide(env_bits, dims, side, arm["link_points"], arm["link_sizes"], arm["dh"],
                                    c4["base"], c4["q"], threads=threads)
            mt_s = time.perf_counter() - t0
            mt_exact = bool(np.array_equal(oc, gpu_counts))
            result["cpu_baseline"]["all_cores"] = {"value": round(n_poses / mt_s, 2), "unit": "poses/s",
                                                   "cores": threads, "sample": f"all {n_poses} cfg4 poses "
                                                   f"({mt_s:.1f} s)", "counts_bit_exact": mt_exact}
            log(f"bench: CPU oracle {n_poses / mt_s:.2f} poses/s on {threads} threads, counts bit-exact: {mt_exact}")
        if not args.no_edt:
            # EDT: the oracle's exact PBA on the headline field (512^3, 1% i.i.d.), 1 core and all cores
            d3 = (512, 512, 512)
            idx = scenes.cfg3_random(d3, 0.01, EDT_SEEDS[0.01])
            u = O.dist_uninit(d3)
            src = np.full(512 ** 3, u, np.uint32)
            src[idx] = (idx % 512) | (((idx // 512) % 512) << 10) | ((idx // (512 * 512)) << 20)
            m1 = src.copy()
            t0 = time.perf_counter()
            O.pba(m1, d3)
            s1 = time.perf_counter() - t0
            m2 = src.copy()
            t0 = time.perf_counter()
            O.pba(m2, d3, threads=threads)
            s2 = time.perf_counter() - t0
            same = bool(np.array_equal(m1, m2))
            result["edt"]["cpu_baseline"] = {
                "value": round(512 ** 3 / s1 / 1e9, 4), "unit": "Gvox/s", "cores": 1, "kind": "port",
                "sample": f"the whole 512^3 1% field, serial C oracle PBA ({s1:.1f} s)",
                "all_cores": {"value": round(512 ** 3 / s2 / 1e9, 4), "unit": "Gvox/s", "cores": threads,
                              "seconds": round(s2, 2), "equal_to_serial": same}}
            log(f"bench: CPU oracle EDT 512^3 1%: {s1:.2f} s on 1 core, {s2:.2f} s on {threads} threads")
            del src, m1, m2

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
