#!/usr/bin/env python
"""Benchmark of the static priority search tree hot path on one MI355X.

Workload (N = 1): a 2^26-point tree of float32 intervals with uint32 IDs
(dim1 ~ U[0,1), dim2 = min(dim1 + Exp(0.02), 1), id = index); one step =
one batch of 2048 queries: 1024 two-sided stabbing queries (the
reference's twoSidedLeftSearch(q, q): dim1 <= q <= dim2, q ~ U[0,1)) and
1024 three-sided queries (threeSidedSearch(a, a + w, c), a, c ~ U[0,1),
w = exp(U[ln 1e-4, ln 0.5])), run through the batched path (count pass,
device prefix sum, write pass; IDs returned). Reported beside it, outside
the timed steps: the plain one-launch-per-query path (preallocated
output), interval parallel search on the same data, two-sided right
queries, tree construction, and the CPU baseline (the oracle's stack
walks, single thread, on the same tree).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "static-priority-search-tree_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBPS = 8000.0
BYTES_PER_ID = 21  # 3 x f32 (dim1, dim2, med) + u32 id + 1 B bitcode read, u32 written
# kernel launches of one write pass of the walk: four walk rounds and the subtree copy kernel
WRITE_LAUNCHES = 5
# level index write pass: each reported u32 ID is read once from a sorted run and written once
COPY_BYTES_PER_ID = 8


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def iso_bench(args, P, D, torch, dev, stream):
    """Isosurface form (StaticPSTGPUArr): a side^3 float32 volume (sum of 64
    Gaussians + N(0, 0.01) noise, seed 1) or a raw float32 dataset file,
    4^3-voxel metacells, one tree per 64-thread block, 200 isovalues evenly
    spaced over [vmin, vmax] (both ends included) as twoSidedLeftSearch(q, q)."""
    import numpy as np
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
    if args.volume_file:
        dims = tuple(args.volume_dims) if args.volume_dims else D.DATASET_DIMS[args.dataset]
        nv = dims[0] * dims[1] * dims[2]
        vol = torch.empty(nv, dtype=torch.float32, device=dev)
        ev[0].record(stream)
        P.read_vertices_to_device(args.volume_file, nv, vol, "f32", stream=stream)
        ev[1].record(stream)
        source = f"{os.path.basename(args.volume_file)} ({dims[0]}x{dims[1]}x{dims[2]} f32 raw file)"
    else:
        side = args.iso_side
        dims = (side, side, side)
        nv = side ** 3
        prm = torch.from_numpy(P.gaussian_volume_params(64, 1).ravel()).to(dev)
        vol = torch.empty(nv, dtype=torch.float32, device=dev)
        ev[0].record(stream)
        P.synth_gaussian_volume(vol, dims, prm, 64, 0.01, 1, stream=stream)
        ev[1].record(stream)
        source = f"{side}^3 f32 volume (64 Gaussians + N(0, 0.01) noise, seed 1)"
    nmc, g = P.calcNumMetacells(dims, (4, 4, 4))
    rec = P.point_dtype("f32", "u32")
    mc = torch.empty(nmc * rec.itemsize, dtype=torch.uint8, device=dev)
    P.form_metacells_into(vol, dims, (4, 4, 4), mc, stream=stream)  # warm
    ev[2].record(stream)
    P.form_metacells_into(vol, dims, (4, 4, 4), mc, stream=stream)
    ev[3].record(stream)
    torch.cuda.synchronize()
    vmin, vmax = float(vol.min().item()), float(vol.max().item())
    vol_keep = vol  # the isosurface extraction below reads the volume again
    del vol
    t0 = time.perf_counter()
    forest = P.StaticPSTGPUArr(mc.data_ptr(), nmc, 64, "f32", "u32")
    torch.cuda.synchronize()
    forest_ms = (time.perf_counter() - t0) * 1e3
    iso = np.linspace(vmin, vmax, args.iso_queries).astype(np.float32)
    qs = np.zeros(len(iso), P.query_dtype("f32"))
    qs["min_dim1"] = qs["max_dim1"] = qs["min_dim2"] = iso
    qs["kind"] = 1
    nq = len(qs)
    d_q = P.to_device(qs, dev)
    counts = torch.zeros(nq, dtype=torch.int64, device=dev)
    offsets = torch.zeros(nq + 1, dtype=torch.int64, device=dev)
    forest.batch_count(d_q, nq, counts, stream=stream)
    torch.cuda.synchronize()
    total = int(counts.sum().item())
    out = torch.empty(total + 1, dtype=torch.int32, device=dev)
    forest.batch_search(d_q, nq, counts, offsets, out, total + 1, stream=stream)
    torch.cuda.synchronize()
    # the batch on a second handle with autotuned walk settings (the first
    # keeps the defaults for the plain path below)
    tf = P.StaticPSTGPUArr(mc.data_ptr(), nmc, 64, "f32", "u32")
    tuned = tf.autotune(d_q, nq, stream=stream)
    tf.batch_search(d_q, nq, counts, offsets, out, total + 1, stream=stream)
    reps = 3
    ev[4].record(stream)
    for _ in range(reps):
        tf.batch_search(d_q, nq, counts, offsets, out, total + 1, stream=stream)
    ev[5].record(stream)
    torch.cuda.synchronize()
    batch_ms = ev[4].elapsed_time(ev[5]) / reps
    tf.close()
    ok = int(offsets[nq].item()) == total
    # spot check against the metacells themselves
    cells = mc.cpu().numpy().view(rec)
    host_counts = counts.cpu().numpy()
    for k in (0, nq // 3, nq // 2, nq - 1):
        ok &= int(np.count_nonzero((cells["dim1"] <= iso[k]) & (iso[k] <= cells["dim2"]))) == int(host_counts[k])
    # plain path and IPS on a sample of the isovalues
    sel = np.linspace(0, nq - 1, min(nq, args.iso_plain)).astype(int)
    cap = int(host_counts.max()) + 1
    pout = torch.empty(cap, dtype=torch.int32, device=dev)
    forest.search_into(1, 0, iso[sel[0]], iso[sel[0]], pout, cap)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in sel:
        ok &= forest.search_into(1, 0, iso[k], iso[k], pout, cap) == int(host_counts[k])
    plain_qps = len(sel) / (time.perf_counter() - t0)
    ctr = torch.zeros(1, dtype=torch.int64, device=dev)
    P.ips_into(mc, nmc, iso[sel[0]], pout, cap, ctr, stream=stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for k in sel:
        P.ips_into(mc, nmc, iso[k], pout, cap, ctr, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ips_ms = e0.elapsed_time(e1) / len(sel)
    # isosurface of the active metacells of the middle isovalue (marching
    # tetrahedra): count, scan, write
    k_mid = sel[len(sel) // 2]
    q_mid = iso[k_mid]
    nact = forest.search_into(1, 0, q_mid, q_mid, pout, cap)
    ids_d = pout[:nact]
    mcd = (4, 4, 4)
    icounts = torch.zeros(max(nact, 1), dtype=torch.int64, device=dev)
    ioffs = torch.zeros(max(nact, 1) + 1, dtype=torch.int64, device=dev)
    P.isosurface_count(vol_keep, dims, mcd, ids_d, nact, q_mid, icounts, stream=stream)
    torch.cuda.synchronize()
    ntri = int(icounts[:nact].sum().item())
    tri = torch.empty(max(ntri, 1) * 9, dtype=torch.float32, device=dev)
    iso_runs = []
    for _ in range(4):  # the first pass is untimed warm-up
        ie = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ie[0].record(stream)
        P.isosurface_count(vol_keep, dims, mcd, ids_d, nact, q_mid, icounts, stream=stream)
        ie[1].record(stream)
        ioffs[1:nact + 1] = torch.cumsum(icounts[:nact], 0)
        ie[2].record(stream)
        P.isosurface_write(vol_keep, dims, mcd, ids_d, nact, q_mid, ioffs, tri, ntri, stream=stream)
        ie[3].record(stream)
        torch.cuda.synchronize()
        iso_runs.append([ie[k].elapsed_time(ie[k + 1]) for k in range(3)])
    iso_split = sorted(iso_runs[1:], key=sum)[1]  # median of the timed passes
    iso_ms = sum(iso_split)
    ok &= int(ioffs[nact].item()) == ntri
    del vol_keep
    forest.close()
    mc_ms = ev[2].elapsed_time(ev[3])
    mc_bytes = nv * 4 + nmc * rec.itemsize
    log(f"[bench] iso: volume {ev[0].elapsed_time(ev[1]):.1f} ms ({source}), metacells {mc_ms:.2f} ms "
        f"({mc_bytes / (mc_ms * 1e-3) / 1e9:.0f} GB/s), forest {forest_ms:.1f} ms, "
        f"{nq} isovalues batched {batch_ms:.2f} ms ({total} IDs), plain {plain_qps:.0f} q/s, "
        f"IPS {1e3 / ips_ms:.0f} q/s, isosurface of {nact} active metacells {iso_ms:.2f} ms ({ntri} triangles; "
        f"passes {[[round(x, 2) for x in r] for r in iso_runs]})")
    return {
        "workload": f"{source}, 4^3-voxel metacells "
                    f"({nmc} intervals), StaticPSTGPUArr with 64 threads per block, {nq} isovalues over "
                    "[vmin, vmax] as twoSidedLeftSearch(q, q)",
        "volume_ms": round(ev[0].elapsed_time(ev[1]), 3),
        "metacell_ms": round(mc_ms, 3),
        "metacell_GBps": round(mc_bytes / (mc_ms * 1e-3) / 1e9, 1),
        "forest_build_ms": round(forest_ms, 2),
        "batch_ms": round(batch_ms, 3),
        "batch_qps": round(nq / (batch_ms * 1e-3), 1),
        "batch_autotuned": tuned,
        "reported_ids": total,
        "plain_qps": round(plain_qps, 1),
        "ips_qps": round(1e3 / ips_ms, 1),
        "isosurface": {"isovalue": float(q_mid), "active_metacells": nact, "triangles": ntri,
                       "ms": round(iso_ms, 3), "count_ms": round(iso_split[0], 3),
                       "scan_ms": round(iso_split[1], 3), "write_ms": round(iso_split[2], 3),
                       "triangle_GBps": round(ntri * 36 / (iso_split[2] * 1e-3) / 1e9, 1),
                       "method": "marching tetrahedra, count + scan + write (median of 3 after one warm pass)"},
        "results_checked": bool(ok),
    }


def serve_bench(args, P, D, torch, dev):
    """Serving form: 8 resident trees of 2^20 ... 2^27 points, alternating
    float32 intervals and int32 geometric-width points; queries arrive as a
    seeded Poisson stream at args.serve_qps, pick a tree with probability
    proportional to its size, and draw a query from that tree's mix (float
    trees: stabbing (q, q) and three-sided as in the main workload; int32
    trees: the two-/three-sided mix of the int32 batch workload). The
    library's server batches them on the fly and admits them by a 32 GB
    output budget. Latency = arrival to results ready in device memory."""
    import numpy as np
    trees, keep, mixes, sizes = [], [], [], []
    for k, lg in enumerate(range(20, 28)):
        n = 1 << lg
        if k % 2 == 0:
            pts = D.interval_points(n, seed=100 + lg)
            t = "f32"
            qs = D.float_query_mix(2048, 2048, seed=200 + lg)
            qs["kind"][:2048] = 1
        else:
            pts = D.int_points(n, seed=100 + lg)
            t = "i32"
            qs = D.int_query_mix(4096, seed=200 + lg)
        d = P.to_device(pts, dev)
        trees.append(P.StaticPSTGPU(d, n, t, "u32"))
        torch.cuda.synchronize()
        del d, pts
        mixes.append(qs)
        sizes.append(n)
    torch.cuda.empty_cache()
    index_bytes = 0
    if args.level_index:
        for tr, n in zip(trees, sizes):
            tr.set_level_index(True)
            index_bytes += P.StaticPSTGPU.level_index_bytes(n, tr.t, "u32")
        torch.cuda.synchronize()
    srv = P.Server(trees, args.serve_budget_gb << 30, max_batch=args.serve_max_batch)
    rng = np.random.default_rng(args.serve_seed)
    nq = int(args.serve_qps * args.serve_seconds)
    gaps = rng.exponential(1.0 / args.serve_qps, nq)
    arrive = np.cumsum(gaps)
    prob = np.array(sizes, np.float64) / sum(sizes)
    tree_of = rng.choice(len(trees), nq, p=prob)
    pick = rng.integers(0, 4096, nq)
    arr_at, lat, ids = {}, [], 0
    i = 0
    t0 = time.perf_counter()
    while len(lat) < nq:
        now = time.perf_counter() - t0
        while i < nq and arrive[i] <= now:
            tk = srv.submit(int(tree_of[i]), mixes[tree_of[i]][pick[i]])
            arr_at[tk] = arrive[i]
            i += 1
        srv.step()
        for tk in srv.poll():
            tnow = time.perf_counter() - t0
            _, cnt = srv.result(tk)
            ids += cnt
            lat.append(tnow - arr_at.pop(tk))
            srv.release(tk)
        if time.perf_counter() - t0 > args.serve_seconds * 4 + 20:
            break
    elapsed = time.perf_counter() - t0
    srv.close()
    lat_ms = np.array(lat) * 1e3
    # plain path on a sample of the same mix
    samp = np.arange(0, nq, max(1, nq // 200))[:200]
    cap = 1 << 27
    pout = torch.empty(cap, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for j in samp:
        q = mixes[tree_of[j]][pick[j]]
        trees[tree_of[j]].search_into(int(q["kind"]), q["min_dim1"], q["max_dim1"], q["min_dim2"], pout, cap)
    plain_qps = len(samp) / (time.perf_counter() - tp)
    for tr in trees:
        tr.close()
    done = len(lat)
    log(f"[bench] serve: {done}/{nq} queries at {args.serve_qps} q/s offered, "
        f"{done / elapsed:.0f} q/s served, p50 {np.percentile(lat_ms, 50):.2f} ms, "
        f"p99 {np.percentile(lat_ms, 99):.2f} ms, plain {plain_qps:.0f} q/s")
    return {
        "workload": "8 resident trees 2^20..2^27 (alternating f32 intervals / int32 geometric widths), "
                    f"Poisson arrivals at {args.serve_qps} q/s for {args.serve_seconds} s, tree chosen with "
                    "probability proportional to size, 32 GB output budget, IDs returned"
                    + (", level index on every tree" if args.level_index else ""),
        "level_index_bytes": index_bytes,
        "offered_qps": args.serve_qps, "completed": done, "submitted": nq,
        "served_qps": round(done / elapsed, 1),
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "reported_ids": int(ids),
        "plain_qps_same_mix": round(plain_qps, 1),
    }


def int_batch_bench(args, P, D, torch, dev):
    """Batched form on int32: a 2^27-point tree (dim1 ~ U{0..2^30-1}, dim2 =
    dim1 + Geom(2^-12), uint32 IDs = index), 8192 mixed queries (half
    two-sided, half three-sided, log-uniform widths), count + scan + write
    captured as one HIP graph, launch settings from the autotuner."""
    import numpy as np
    n = 1 << args.int_log2n
    pts = D.int_points(n, seed=3)
    d = P.to_device(pts, dev)
    t0 = time.perf_counter()
    tree = P.StaticPSTGPU(d, n, "i32", "u32")
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    qs = D.int_query_mix(args.int_queries, seed=3)
    nq = len(qs)
    dq = P.to_device(qs, dev)
    counts = torch.zeros(nq, dtype=torch.int64, device=dev)
    offs = torch.zeros(nq + 1, dtype=torch.int64, device=dev)
    tree.batch_count(dq, nq, counts)
    torch.cuda.synchronize()
    total = int(counts.sum().item())
    out = torch.empty(total + 1, dtype=torch.int32, device=dev)
    host_counts = counts.cpu().numpy()
    ok = True
    for k in (0, 1, nq // 2, nq - 1):
        a, b, c, kind = qs["min_dim1"][k], qs["max_dim1"][k], qs["min_dim2"][k], int(qs["kind"][k])
        m = (pts["dim2"] >= c) & ((pts["dim1"] <= b) if kind == 1 else (pts["dim1"] >= a))
        if kind == 0:
            m &= pts["dim1"] <= b
        ok &= int(np.count_nonzero(m)) == int(host_counts[k])
    del d
    tuned = tree.autotune(dq, nq)
    s = torch.cuda.Stream()
    g = tree.batch_graph(dq, nq, counts, offs, out, total + 1, stream=s)
    g.launch(s)
    s.synchronize()
    reps = args.int_steps
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        g.launch(s)
    e1.record(s)
    s.synchronize()
    graph_ms = e0.elapsed_time(e1) / reps
    ok &= int(offs[nq].item()) == total
    e0.record(s)
    for _ in range(reps):
        tree.batch_search(dq, nq, counts, offs, out, total + 1, stream=s)
    e1.record(s)
    s.synchronize()
    eager_ms = e0.elapsed_time(e1) / reps
    g.close()
    index = None
    if args.level_index:
        ks = (0, 1, nq // 2, nq - 1)
        oh, ch = offs.cpu().numpy(), counts.cpu().numpy()
        ref = {k: np.sort(out[oh[k]:oh[k] + ch[k]].cpu().numpy()) for k in ks}
        t0 = time.perf_counter()
        tree.set_level_index(True)
        torch.cuda.synchronize()
        ib_ms = (time.perf_counter() - t0) * 1e3
        g = tree.batch_graph(dq, nq, counts, offs, out, total + 1, stream=s)
        g.launch(s)
        s.synchronize()
        e0.record(s)
        for _ in range(reps):
            g.launch(s)
        e1.record(s)
        s.synchronize()
        ig_ms = e0.elapsed_time(e1) / reps
        ok &= int(offs[nq].item()) == total
        oh = offs.cpu().numpy()
        ok &= all(np.array_equal(np.sort(out[oh[k]:oh[k] + ch[k]].cpu().numpy()), ref[k]) for k in ks)
        e0.record(s)
        for _ in range(reps):
            tree.batch_search(dq, nq, counts, offs, out, total + 1, stream=s)
        e1.record(s)
        s.synchronize()
        ie_ms = e0.elapsed_time(e1) / reps
        g.close()
        index = {"build_ms": round(ib_ms, 2), "graph_ms_per_batch": round(ig_ms, 3),
                 "graph_qps": round(nq / (ig_ms * 1e-3), 1), "eager_ms_per_batch": round(ie_ms, 3),
                 "reported_ids_per_s": round(total / (ig_ms * 1e-3), 1),
                 "copy_GBps": round(total * COPY_BYTES_PER_ID / (ig_ms * 1e-3) / 1e9, 1)}
        log(f"[bench] int32 level index: build {ib_ms:.1f} ms, graph {ig_ms:.2f} ms "
            f"({nq / (ig_ms * 1e-3):.0f} q/s), eager {ie_ms:.2f} ms")
    tree.close()
    log(f"[bench] int32 2^{args.int_log2n}: build {build_ms:.1f} ms, {nq} queries, {total} IDs, graph "
        f"{graph_ms:.2f} ms ({nq / (graph_ms * 1e-3):.0f} q/s), eager {eager_ms:.2f} ms, tuned {tuned}")
    return {
        "workload": f"2^{args.int_log2n}-point int32 StaticPSTGPU (geometric widths, p = 2^-12), {nq} queries "
                    "(50% two-sided, 50% three-sided, log-uniform widths), count + scan + write as one HIP "
                    "graph, autotuned launch settings, IDs returned",
        "build_ms": round(build_ms, 2),
        "reported_ids_per_batch": total,
        "graph_ms_per_batch": round(graph_ms, 3),
        "graph_qps": round(nq / (graph_ms * 1e-3), 1),
        "eager_ms_per_batch": round(eager_ms, 3),
        "reported_ids_per_s": round(total / (graph_ms * 1e-3), 1),
        "alg_GBps": round(total * BYTES_PER_ID / (graph_ms * 1e-3) / 1e9, 1),
        "autotuned": tuned,
        "level_index": index,
        "results_checked": bool(ok),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log2n", type=int, default=26)
    ap.add_argument("--nq2", type=int, default=1024)
    ap.add_argument("--nq3", type=int, default=1024)
    ap.add_argument("--plain", type=int, default=256, help="queries timed on the plain path")
    ap.add_argument("--ips", type=int, default=64, help="queries timed for interval parallel search")
    ap.add_argument("--cpu-sample", type=int, default=512, help="queries in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--iso-side", type=int, default=1024, help="isosurface volume side (0: skip)")
    ap.add_argument("--volume-file", default="", help="raw float32 volume for the isosurface leg "
                    "(default: the synthetic volume)")
    ap.add_argument("--dataset", default="Vortex", choices=sorted(["Isabel", "Radiation", "TeraShake", "Vortex"]),
                    help="dataset whose grid dimensions the volume file has")
    ap.add_argument("--volume-dims", type=int, nargs=3, default=None, help="grid dimensions of the volume file")
    ap.add_argument("--iso-queries", type=int, default=200)
    ap.add_argument("--iso-plain", type=int, default=50, help="isovalues timed on the plain path and IPS")
    ap.add_argument("--int-log2n", type=int, default=27, help="int32 batch workload size (0: skip)")
    ap.add_argument("--int-queries", type=int, default=8192)
    ap.add_argument("--int-steps", type=int, default=5)
    ap.add_argument("--serve-qps", type=float, default=20000.0, help="offered load of the serving run (0: skip)")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-budget-gb", type=int, default=32)
    ap.add_argument("--serve-max-batch", type=int, default=1024)
    ap.add_argument("--serve-seed", type=int, default=4)
    ap.add_argument("--sort-queries", type=int, default=0,
                    help="hand the batch to the library pre-ordered by min_dim1/max_dim1 "
                         "(the library groups queries itself)")
    ap.add_argument("--ungrouped", type=int, default=1,
                    help="also time the batch with one walk per query (grouping off)")
    ap.add_argument("--level-index", type=int, default=1,
                    help="1: the headline batch runs through the level index (the walk is timed beside it)")
    ap.add_argument("--profile-mode", action="store_true",
                    help="build + one batched step + a 1 GiB copy (for counter scaling), no JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by replicating the tree or splitting points; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import pst_amd as P
    import pst_data as D

    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    n = 1 << args.log2n
    log(f"[bench] generating {n} points")
    pts = D.interval_points(n, seed=0)
    d_pts = P.to_device(pts, dev)
    torch.cuda.synchronize()

    # ---- construction
    # the first build of a process also loads the construction kernels and
    # maps fresh memory; later builds show the steady construction time
    log("[bench] building tree")
    t0 = time.perf_counter()
    tree = P.StaticPSTGPU(d_pts, n, "f32", "u32")
    torch.cuda.synchronize()
    cold_build_ms = (time.perf_counter() - t0) * 1e3
    warm = []
    for _ in range(3):
        t0 = time.perf_counter()
        again = P.StaticPSTGPU(d_pts, n, "f32", "u32")
        torch.cuda.synchronize()
        warm.append((time.perf_counter() - t0) * 1e3)
        again.close()
    build_ms = float(np.median(warm))
    S = (1 << (args.log2n + 1)) - 1
    build_bytes = n * 12 + S * 17
    log(f"[bench] build {build_ms:.1f} ms (median of 3; first build of the process {cold_build_ms:.1f} ms)")

    # ---- queries
    qs = D.float_query_mix(args.nq3, args.nq2, seed=0)
    qs["kind"][:args.nq2] = 1  # two-sided left stabbing (q, q)
    if args.sort_queries:
        # order each kind by the bound that fixes where in dim1 its search runs
        for lo, hi in ((0, args.nq2), (args.nq2, args.nq2 + args.nq3)):
            key = qs["max_dim1"][lo:hi] if lo == 0 else qs["min_dim1"][lo:hi]
            qs[lo:hi] = qs[lo:hi][np.argsort(key, kind="stable")]
    nq = len(qs)
    d_q = P.to_device(qs, dev)
    counts = torch.zeros(nq, dtype=torch.int64, device=dev)
    offsets = torch.zeros(nq + 1, dtype=torch.int64, device=dev)
    # size the output from one count pass
    tree.batch_count(d_q, nq, counts, stream=stream)
    torch.cuda.synchronize()
    total = int(counts.sum().item())
    log(f"[bench] batch of {nq}: {total} reported IDs")
    out = torch.empty(total + 1, dtype=torch.int32, device=dev)

    # sanity (outside timing): batch counts equal brute-force counts for a few queries
    host_counts = counts.cpu().numpy()
    ok = True
    d1, d2 = pts["dim1"], pts["dim2"]
    for k in (0, 1, args.nq2, args.nq2 + 1):
        a, b, c, kind = qs["min_dim1"][k], qs["max_dim1"][k], qs["min_dim2"][k], int(qs["kind"][k])
        if kind == 1:
            want = int(np.count_nonzero((d1 <= b) & (d2 >= c)))
        else:
            want = int(np.count_nonzero((d1 >= a) & (d1 <= b) & (d2 >= c)))
        ok &= want == int(host_counts[k])

    if args.profile_mode:
        if args.level_index:
            tree.set_level_index(True)
        tree.batch_search(d_q, nq, counts, offsets, out, total + 1, stream=stream)
        torch.cuda.synchronize()
        # plain copy of a known size: the scale for HBM byte counters
        x = torch.ones(1 << 28, dtype=torch.int32, device=dev)
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        log(f"[bench] profile mode done: {total} IDs, copy of {x.numel() * 4} bytes")
        tree.close()
        return

    def timed_steps():
        for _ in range(args.warmup):
            tree.batch_search(d_q, nq, counts, offsets, out, total + 1, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            tree.batch_search(d_q, nq, counts, offsets, out, total + 1, stream=stream)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def pass_times():
        # per-pass timing with events on the launch stream (roofline figure)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        tc = tw = 0.0
        reps = max(1, args.steps)
        for _ in range(reps):
            ev[0].record(stream)
            tree.batch_count(d_q, nq, counts, stream=stream)
            ev[1].record(stream)
            tree.batch_write(d_q, nq, offsets, out, caps_d=counts, stream=stream)
            ev[2].record(stream)
            torch.cuda.synchronize()
            tc += ev[0].elapsed_time(ev[1])
            tw += ev[1].elapsed_time(ev[2])
        return tc / reps, tw / reps

    def sample_results(ks):
        c, o = counts.cpu().numpy(), offsets.cpu().numpy()
        return {k: np.sort(out[o[k]:o[k] + c[k]].cpu().numpy()) for k in ks}

    # ---- timed steps on the tree walk: batched count + scan + write
    elapsed = timed_steps()
    ok &= int(offsets[nq].item()) == total
    walk_ms = elapsed * 1e3 / args.steps
    w_count, w_write = pass_times()
    sample_ks = sorted({0, 1, args.nq2 // 2, args.nq2 - 1, args.nq2, args.nq2 + args.nq3 // 2, nq - 1})
    walk_ref = sample_results(sample_ks)
    log(f"[bench] walk: {walk_ms:.2f} ms per batch ({nq / (walk_ms * 1e-3):.0f} q/s), count pass {w_count:.2f} ms, "
        f"write pass {w_write:.2f} ms")
    # HBM bytes per launch of the walk's write pass on this workload, from
    # the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (copy-calibrated)
    walk_traffic = None
    tpath = os.path.join(ROOT, "profiles", "write_pass_k_search_round_k_copy_tasks_f32u32_2e26_batch2048.traffic.json")
    if args.log2n == 26 and args.nq2 == 1024 and args.nq3 == 1024 and os.path.exists(tpath):
        with open(tpath) as f:
            walk_traffic = round(json.load(f)["traffic_bytes_per_launch"])
    walk = {"ms_per_step": round(walk_ms, 3), "qps": round(nq / (walk_ms * 1e-3), 2),
            "count_pass_ms": round(w_count, 3), "write_pass_ms": round(w_write, 3),
            "write_pass_alg_GBps": round(total * BYTES_PER_ID / (w_write * 1e-3) / 1e9, 1),
            "write_pass_traffic_bytes_per_launch": walk_traffic,
            "write_launches": WRITE_LAUNCHES}

    # ---- the same batch with one walk per query (grouping off), for comparison
    ungrouped_qps = None
    if args.ungrouped:
        tree.set_grouping(1)
        tree.batch_search(d_q, nq, counts, offsets, out, total + 1, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tree.batch_search(d_q, nq, counts, offsets, out, total + 1, stream=stream)
        torch.cuda.synchronize()
        ungrouped_qps = nq / (time.perf_counter() - t0)
        ok &= int(offsets[nq].item()) == total
        tree.set_grouping(8)
        log(f"[bench] ungrouped batch {ungrouped_qps:.0f} q/s")

    # ---- level index: build it, then the same batch through it
    index = None
    if args.level_index:
        t0 = time.perf_counter()
        tree.set_level_index(True)
        torch.cuda.synchronize()
        first_index_ms = (time.perf_counter() - t0) * 1e3
        iw = []
        for _ in range(2):
            tree.set_level_index(False)
            t0 = time.perf_counter()
            tree.set_level_index(True)
            torch.cuda.synchronize()
            iw.append((time.perf_counter() - t0) * 1e3)
        out.fill_(-1)
        elapsed = timed_steps()
        ok &= int(offsets[nq].item()) == total
        got = sample_results(sample_ks)
        same = all(np.array_equal(got[k], walk_ref[k]) for k in sample_ks)
        ok &= same
        t_count, t_write = pass_times()
        index = {"build_ms": round(min(iw), 2), "first_build_ms": round(first_index_ms, 2),
                 "bytes": P.StaticPSTGPU.level_index_bytes(n, "f32", "u32"),
                 "same_results_as_walk": bool(same)}
        log(f"[bench] level index: build {min(iw):.1f} ms (first {first_index_ms:.1f} ms), "
            f"{index['bytes'] / 2**30:.1f} GiB, results as the walk's: {same}")
    else:
        t_count, t_write = w_count, w_write
    ms_per_step = elapsed * 1e3 / args.steps
    qps = nq * args.steps / elapsed
    ids_per_s = total * args.steps / elapsed
    if args.level_index:
        alg_bytes = total * COPY_BYTES_PER_ID
        launches = 1
        kernel = ("write pass: k_lv_plan<float, unsigned int, true> + offset scan + k_lv_mark + "
                  "k_lv_copy<unsigned int> (the copy streams every ID; the pass time is its upper bound)")
        tpath = os.path.join(ROOT, "profiles", "write_pass_k_lv_copy_f32u32_2e26_batch2048.traffic.json")
    else:
        alg_bytes = total * BYTES_PER_ID
        launches = WRITE_LAUNCHES
        kernel = ("write pass: k_search_round<float, unsigned int, false, true, 8> x 4 "
                  "+ k_copy_tasks<float, unsigned int, false, 8> x 1")
    achieved = alg_bytes / (t_write * 1e-3) / 1e9
    traffic, traffic_src = None, None
    if args.log2n == 26 and args.nq2 == 1024 and args.nq3 == 1024 and os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic, traffic_src = round(tj["traffic_bytes_per_launch"]), tj["source"]
    log(f"[bench] count pass {t_count:.2f} ms, write pass {t_write:.2f} ms, "
        f"write-pass alg {achieved:.0f} GB/s")

    # ---- plain path: one blocking search per query, preallocated output
    np_plain = min(args.plain, nq)
    sel = np.concatenate([np.arange(np_plain // 2), args.nq2 + np.arange(np_plain - np_plain // 2)])
    cap = int(host_counts[sel].max()) + 1
    pout = torch.empty(cap, dtype=torch.int32, device=dev)
    for k in sel[:4]:
        tree.search_into(int(qs["kind"][k]), qs["min_dim1"][k], qs["max_dim1"][k], qs["min_dim2"][k], pout, cap)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_ids = 0
    for k in sel:
        got = tree.search_into(int(qs["kind"][k]), qs["min_dim1"][k], qs["max_dim1"][k],
                               qs["min_dim2"][k], pout, cap)
        plain_ids += got
        ok &= got == int(host_counts[k])
    plain_s = time.perf_counter() - t0
    plain_qps = len(sel) / plain_s
    log(f"[bench] plain path {plain_qps:.0f} q/s")

    # ---- two-sided right (q, q): outputs average n/2
    right_q = D.stabbing_queries(8, seed=5)
    rq = np.zeros(len(right_q), P.query_dtype("f32"))
    rq["min_dim1"] = right_q
    rq["min_dim2"] = right_q
    rq["kind"] = 2
    d_rq = P.to_device(rq, dev)
    rc = torch.zeros(len(rq), dtype=torch.int64, device=dev)
    ro = torch.zeros(len(rq) + 1, dtype=torch.int64, device=dev)
    tree.batch_count(d_rq, len(rq), rc, stream=stream)
    torch.cuda.synchronize()
    rtotal = int(rc.sum().item())
    rout = torch.empty(rtotal + 1, dtype=torch.int32, device=dev)
    tree.batch_search(d_rq, len(rq), rc, ro, rout, rtotal + 1, stream=stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tree.batch_search(d_rq, len(rq), rc, ro, rout, rtotal + 1, stream=stream)
    torch.cuda.synchronize()
    right_s = time.perf_counter() - t0
    del rout

    # ---- interval parallel search on the same data (stabbing queries)
    nips = min(args.ips, args.nq2)
    ips_out = torch.empty(int(host_counts[:nips].max()) + 1, dtype=torch.int32, device=dev)
    ips_ctr = torch.zeros(1, dtype=torch.int64, device=dev)
    P.ips_into(d_pts, n, qs["max_dim1"][0], ips_out, ips_out.numel(), ips_ctr, stream=stream)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for k in range(nips):
        P.ips_into(d_pts, n, qs["max_dim1"][k], ips_out, ips_out.numel(), ips_ctr, stream=stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    ips_ms = ev0.elapsed_time(ev1) / nips
    ok &= int(ips_ctr.item()) == int(host_counts[nips - 1])
    ips_gbps = (n * 12 + host_counts[:nips].mean() * 4) / (ips_ms * 1e-3) / 1e9
    log(f"[bench] IPS {1e3 / ips_ms:.0f} q/s ({ips_gbps:.0f} GB/s)")

    # ---- CPU baseline: the oracle's walks on the same tree (exported)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O  # CPU checker, used here only as the timed CPU baseline
        ex_bytes = P.lib().pst_gpu_export_bytes(tree.h)
        tr = O.empty_tree("f32", "u32", n)
        assert tr.blob.nbytes >= ex_bytes
        import ctypes
        off = (ctypes.c_size_t * 6)()
        P._check(P.lib().pst_gpu_export_tree(tree.h, tr.blob.ctypes.data, ex_bytes, off))
        ns = min(args.cpu_sample, nq)
        csel = np.concatenate([np.arange(ns // 2), args.nq2 + np.arange(ns - ns // 2)])
        kinds = np.where(qs["kind"][csel] == 1, 1, 0)
        tot, secs = O.time_queries(tr, kinds, qs["min_dim1"][csel], qs["max_dim1"][csel],
                                   qs["min_dim2"][csel], threads=1)
        ok &= tot == int(host_counts[csel].sum())
        cpu = {"value": round(ns / secs, 3), "unit": "queries/s", "cores": 1, "kind": "port",
               "sample": f"{ns} of the batch's queries ({ns // 2} two-sided stabbing, "
                         f"{ns - ns // 2} three-sided) on the same 2^{args.log2n}-point tree, "
                         f"oracle stack walk (g++ -O2), one thread, {secs:.1f} s, "
                         f"{tot} IDs"}
        log(f"[bench] CPU {ns / secs:.2f} q/s over {secs:.1f} s")
        del tr

    tree.close()
    del d_pts, out
    torch.cuda.empty_cache()
    iso = iso_bench(args, P, D, torch, dev, stream) if (args.iso_side > 0 or args.volume_file) else None
    torch.cuda.empty_cache()
    intb = int_batch_bench(args, P, D, torch, dev) if args.int_log2n > 0 else None
    torch.cuda.empty_cache()
    serve = serve_bench(args, P, D, torch, dev) if args.serve_qps > 0 else None
    res = {
        "metric": "queries/s + reported IDs/s (GB/s vs HBM peak), 2^26-pt PST; p50/p99 latency",
        "value": round(qps, 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 generator, pst_data.py)",
        "config": {
            "workload": f"batched search on a 2^{args.log2n}-point f32/u32 StaticPSTGPU: "
                        f"{args.nq2} twoSidedLeftSearch(q, q) + {args.nq3} threeSidedSearch per step, "
                        "count + scan + write, IDs returned",
            "points": n, "queries_per_step": nq, "reported_ids_per_step": total,
        },
        "reported_ids_per_s": round(ids_per_s, 1),
        "alg_GBps_step": round(ids_per_s * BYTES_PER_ID / 1e9, 1),
        "path": "level index" if args.level_index else "tree walk",
        "count_pass_ms": round(t_count, 3),
        "write_pass_ms": round(t_write, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                     "kernel": kernel, "launches_per_pass": launches,
                     "avg_launch_ms": round(t_write / launches, 3),
                     "alg_bytes_per_launch": round(alg_bytes / launches),
                     "alg_bytes_per_id": COPY_BYTES_PER_ID if args.level_index else BYTES_PER_ID,
                     "traffic_source": traffic_src},
        "level_index": index,
        "walk": walk,
        "cpu_baseline": cpu,
        "plain_per_query_qps": round(plain_qps, 2),
        "batched_over_plain": round(qps / plain_qps, 2),
        "ungrouped_batch_qps": None if ungrouped_qps is None else round(ungrouped_qps, 2),
        "right_qq_qps": round(len(rq) / right_s, 2),
        "right_qq_ids_per_query": rtotal // len(rq),
        "ips_qps": round(1e3 / ips_ms, 2),
        "ips_GBps": round(float(ips_gbps), 1),
        "build_ms": round(build_ms, 2),
        "first_build_ms": round(cold_build_ms, 2),
        "build_GBps_io_floor": round(build_bytes / (build_ms * 1e-3) / 1e9, 1),
        "results_checked": bool(ok and (iso is None or iso["results_checked"])
                                and (intb is None or intb["results_checked"])),
        "isosurface": iso,
        "batched_int32": intb,
        "serving": serve,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
