"""Benchmark: inside/outside labels (GWN > 0.5) of a 512^3 query grid against the 1M-triangle open mesh, one MI355X --
learned surrogate + exact fallback, against the plain exact method on the same grid.

Workload: seeded noisy torus R=1, r=0.35, 1353 x 473 quads (2 triangles each), sigma = 0.01 r Gaussian vertex noise,
400 disk holes of radius 0.05 -> 1,004,810 triangles, 31,376 unit-weight boundary segments; queries = the 512^3
lattice over the bounding box expanded by 10 % (134,217,728 points, generated in-kernel); FP32.

  plain      one eval_grid (HipGridGwn): full GWN (fractional + integer term) of every lattice point, labels = w > 0.5
  surrogate  a per-mesh MLP (192 Fourier features up to 64 periods, 4 x 256 hard-SiLU, BF16 MFMA) trained once on 2^24
             exact GWN samples (70 % drawn within 0.05 of the surface, 30 % uniform in the box, minus those in the
             near-surface mask); each step labels the lattice with the fused inference kernel and routes to the exact
             kernels every point in the mask (voxels of edge eps/4 within eps = 0.10 % of the box diagonal of a
             triangle), every point with |pred - 0.5| <= delta, and a 1/1024 audit sample, which take their exact
             labels: integer term, then a far-field decision over a tree of the boundary's loops (second-order
             expansion with a rigorous remainder bound per accepted node; loops too close are summed exactly), and
             the full boundary sum only for the points the decision leaves open. delta comes from 2^23 held-out points (2 x the worst wrong-side margin +
             0.02). Training happens before the timed region and is reported (train_s).
The headline value is surrogate steps/s x queries; mislabels = labels that differ from the plain run's (must be 0).

Reported beside it:
  roofline      the dominant kernel of a surrogate step, re-launched alone on the same inputs and timed with HIP
                events on its stream: the fused lattice inference (394,240 FLOP per point: the network's 492,032
                less layer 0, which on an axis-aligned lattice is the sum of three per-axis table rows; against the
                2.5 PFLOP/s BF16 MFMA peak) unless the exact stage takes longer, else the full boundary sum over the routed
                points (routed x segments x 56 FLOP: the 53 adds and multiplies of the reference's pair expression,
                FMA = 2, plus 2 sqrt and 1 atan2, against the 157.3 TFLOP/s FP32 vector peak)
  cpu_baseline  the reference's own eval_gwnr_mesh_batch_fractional<float> (oracle/_ref, compiled from its headers)
                with its OpenMPDispatcher on this host's cores, on a bounded seeded sample of the grid's queries,
                scaled to queries/s; the integer term is counted as free on the CPU (no Embree here), which favours
                the CPU side
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
FLOP_PER_PAIR = 56
MLP_FLOP_PER_POINT = 2 * (192 * 256 + 3 * 256 * 256 + 256)
# the lattice kernel's own work per point: layer 0 as two adds of table rows per unit, the three 256 x 256 layers and
# the output dot; plus, per table row (one lattice index of one axis), a coordinate's 64 features times 256 weights
MLP_FLOP_PER_POINT_LATTICE = 2 * 256 + 2 * 3 * 256 * 256 + 2 * 256
LATTICE_TABLE_FLOP_PER_ROW = 2 * 64 * 256
METRIC = "GWN queries/s, 1M-tri open mesh; surrogate+fallback vs plain, 0 mislabels"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu() -> dict:
    """The host's CPU as Linux reports it: logical CPUs, this process's affinity, physical cores (distinct (package,
    core) pairs of /proc/cpuinfo) and whether SMT is on (more than one logical CPU per core)."""
    info = {"logical_cpus": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        cores, model, phys, core = set(), None, None, None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None and core is not None:
                cores.add((phys, core))
                phys = core = None
        if phys is not None and core is not None:
            cores.add((phys, core))
        if cores:
            info["physical_cores"] = len(cores)
            info["smt"] = (os.cpu_count() or 0) > len(cores)
        info["model"] = model
    except OSError:  # pragma: no cover
        pass
    return info


def cpu_baseline(pos32_sample, seg32, x0, budget_s: float, grid_desc: str) -> dict | None:
    """The reference's FP32 fractional term on the host cores, over a bounded sample of the benchmark's queries."""
    try:
        from oracle import Reference
    except Exception as e:  # pragma: no cover
        log(f"cpu baseline unavailable: {e}")
        return None
    ref = Reference()
    if not ref.available():
        log("cpu baseline unavailable: oracle/_ref/libref.so not built")
        return None
    threads = ref.omp_threads()
    # calibrate on a small slice, then size the sample to ~budget_s
    n0 = 256
    t = time.perf_counter()
    ref.frac(pos32_sample[:n0], seg32, x0, np.float32, mode=1)
    dt = time.perf_counter() - t
    n = int(min(len(pos32_sample), max(n0, n0 * budget_s / max(dt, 1e-6))))
    t = time.perf_counter()
    ref.frac(pos32_sample[:n], seg32, x0, np.float32, mode=1)
    dt = time.perf_counter() - t
    return {
        "value": n / dt,
        "unit": "queries/s",
        "cores": threads,
        "host": host_cpu(),
        "kind": "reference",
        "sample": f"{n} seeded queries of the {grid_desc} x {len(seg32)} segments, "
                  f"eval_gwnr_mesh_batch_fractional<float> with OpenMPDispatcher ({threads} threads), {dt:.1f} s; "
                  f"integer term not timed (taken as 0)",
    }


def lattice_points(start, dx, dy, dz, nx, ny, nz, idx: np.ndarray) -> np.ndarray:
    """Float32 coordinates of lattice cells idx, evaluated as the kernels do (start + dx*ix + dy*iy + dz*iz, left to
    right per component)."""
    f = np.float32
    ix = (idx % nx).astype(f)
    iy = ((idx // nx) % ny).astype(f)
    iz = (idx // (nx * ny)).astype(f)
    s, a, b, c = (np.asarray(v, dtype=f) for v in (start, dx, dy, dz))
    return (((s[None, :] + a[None, :] * ix[:, None]) + b[None, :] * iy[:, None]) + c[None, :] * iz[:, None]).astype(f)


def lattice_torch(start, dx, dy, dz, n: int, dev):
    """All n^3 lattice points on the device, in the kernels' order and float32 arithmetic (no contraction: torch's
    elementwise kernels round each operation)."""
    import torch

    f = torch.float32
    s, a, b, c = (torch.tensor(np.asarray(v, dtype=np.float32), device=dev) for v in (start, dx, dy, dz))
    r = torch.arange(n, device=dev, dtype=f)
    P = torch.empty((n, n, n, 3), dtype=f, device=dev)
    P[...] = s
    P += a * r[None, None, :, None]
    P += b * r[None, :, None, None]
    P += c * r[:, None, None, None]
    return P.reshape(-1, 3)


def audit_hash(idx, audit_every: int, seed: int):
    """The routing kernel's audit choice with audit seed `seed`: hash32(i * 0x9E3779B97F4A7C15 + seed) % m == 0, on
    int64 device tensors (wrapping multiplication; logical shifts by masking)."""

    def s64(c: int) -> int:
        return c - (1 << 64) if c >= 1 << 63 else c

    def lsr33(x):
        return (x >> 33) & ((1 << 31) - 1)

    x = idx * s64(0x9E3779B97F4A7C15) + seed
    x = x ^ lsr33(x)
    x = x * s64(0xFF51AFD7ED558CCD)
    x = x ^ lsr33(x)
    x = x * s64(0xC4CEB9FE1A85EC53)
    x = x ^ lsr33(x)
    return (x & 0xFFFFFFFF) % audit_every == 0


def copy_calibration() -> int:
    import torch

    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")  # 1 GiB
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": a.numel() * 4, "copies": 3}), flush=True)
    return 0


def bench_batched(args, agwn, mg, torch) -> int:
    """The uneven many-mesh configuration: `--batch-meshes` seeded open meshes (noisy spheres and tori alternating,
    triangle counts log-uniform in [100, 2e5], 1..200 holes), queries per mesh log-uniform in [1e3, 1e6] uniform in
    each mesh's 10 %-expanded box, FP32, x0 = normalize(0.31, 0.57, 0.76); a step is one HipMeshBatch.eval of every
    query (two persistent launches replayed from a HIP graph)."""
    import math

    k = args.batch_meshes
    t = time.perf_counter()
    geo, qn, boxes = mg.batch_workload(k, seed=7, workers=min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:  # host extraction releases the GIL (ctypes)
        segs = list(ex.map(lambda g: agwn.build_boundary_segments(*g), geo))
    meshes = [(v, idx, b) for (v, idx), b in zip(geo, segs)]
    gen_s = time.perf_counter() - t
    q_off = np.concatenate([[0], np.cumsum(qn)]).astype(np.int64)
    nq = int(q_off[-1])
    # queries generated on the device: mesh id per query, then uniform in that mesh's box
    mid = torch.repeat_interleave(torch.arange(k, device="cuda"), torch.tensor(qn, device="cuda"))
    lo_k = torch.from_numpy(np.array([b[0] for b in boxes], np.float32)).cuda()
    hi_k = torch.from_numpy(np.array([b[1] for b in boxes], np.float32)).cuda()
    gen = torch.Generator(device="cuda").manual_seed(11)
    pos = torch.rand((nq, 3), device="cuda", generator=gen)
    pos.mul_(hi_k[mid] - lo_k[mid]).add_(lo_k[mid])
    del mid
    log(f"batched inputs: {k} meshes, {nq} queries ({gen_s:.1f} s to generate)")
    t = time.perf_counter()
    batch = agwn.HipMeshBatch(meshes)
    build_s = time.perf_counter() - t
    x0 = mg.example_x0(np.float32)
    out = torch.empty(nq, dtype=torch.float32, device="cuda")
    t = time.perf_counter()
    batch.eval(pos, q_off, x0, out=out)  # tunes, prepares, captures
    torch.cuda.synchronize()
    first_s = time.perf_counter() - t
    info = batch.info()
    log(f"built in {build_s:.1f} s, first eval (tune + prepare + capture) {first_s:.1f} s: {info}")
    for _ in range(args.warmup):
        batch.eval(pos, q_off, x0, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.eval(pos, q_off, x0, out=out)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    pairs = float(sum(int(qn[m]) * int(len(meshes[m][2])) for m in range(k)))
    line = {
        "metric": "batched GWN queries/s (uneven meshes, persistent launches)",
        "value": nq / (ms * 1e-3), "unit": "queries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded noisy open spheres/tori)",
        "config": {"workload": f"{k} open meshes, {info['triangles']} triangles, {info['unit_segments']} boundary "
                               f"segments, {nq} queries", "parallelism": "single GPU"},
        "pairs_per_s": pairs / (ms * 1e-3), "tuned": info, "build_s": build_s, "first_eval_s": first_s,
    }
    batch.close()
    print(json.dumps(line), flush=True)
    return 0


def bench_plain256(args, agwn, mg, torch) -> int:
    """The plain method at scale: the 1M-triangle holed torus (its 31,376-segment boundary) and the 256^3 lattice over
    its 10 %-expanded box, 16,777,216 queries; the full GWN in FP32 (the grid evaluator, and the array entry on the
    same float32 lattice points), in FP64 (array entry) and eval_gwnr_mesh_batch_robust with robust::int64_grid on the
    FP32 points. Each method: a warm-up call on the first 65,536 queries (per-direction preparation), then one timed
    call on all of them. The labels (GWN > 0.5) of every method are compared with the FP64 ones on every query."""
    t = time.perf_counter()
    verts, idx = mg.holed_torus_mesh()
    v32 = verts.astype(np.float32)
    seg32 = agwn.build_boundary_segments(v32, idx)
    seg64 = agwn.build_boundary_segments(verts, idx)
    lo, hi = mg.expanded_bbox(verts)
    n = 256
    nq, nseg = n ** 3, len(seg32)
    start, dx, dy, dz = mg.grid_axes(lo, hi, n)
    p32 = torch.from_numpy(mg.grid_points(lo, hi, n, np.float32)).cuda()
    p64 = torch.from_numpy(mg.grid_points(lo, hi, n)).cuda()
    s32 = torch.from_numpy(np.ascontiguousarray(seg32, dtype=np.float32)).cuda()
    s64 = torch.from_numpy(np.ascontiguousarray(seg64, dtype=np.float64)).cuda()
    x0 = np.array([1.0, 0.0, 0.0])
    log(f"plain256 inputs: {len(idx) // 3} triangles, {nseg} segments, {nq} queries ({time.perf_counter() - t:.1f} s)")
    grid = agwn.HipGridGwn().set_scene(v32, idx, seg32)
    m32 = agwn.HipMesh(v32, idx, dtype=np.float32)
    m64 = agwn.HipMesh(verts, idx, dtype=np.float64)
    rob = agwn.RobustMeshGwn(v32, idx, seg32, precision="int64_grid", dtype=np.float32)
    w = 65536
    outs = {"fp32_grid": torch.empty(nq, dtype=torch.float32, device="cuda"),
            "fp32_array": torch.empty(nq, dtype=torch.float32, device="cuda"),
            "fp64_array": torch.empty(nq, dtype=torch.float64, device="cuda"),
            "robust_int64_grid": torch.empty(nq, dtype=torch.float32, device="cuda")}
    # the array methods run in 8 slices of 2^21 queries, each timed and logged (the slowest takes seconds per slice)
    sl = [(k * nq // 8, (k + 1) * nq // 8) for k in range(8)]
    runs = {
        "fp32_grid": (lambda: grid.eval_grid(start, dx, dy, dz, n, n, 4, out=outs["fp32_grid"][: 4 * n * n]),
                      [lambda: grid.eval_grid(start, dx, dy, dz, n, n, n, out=outs["fp32_grid"])]),
        "fp32_array": (lambda: m32.full(p32[:w], x0, s32, out=outs["fp32_array"][:w]),
                       [lambda a=a, b=b: m32.full(p32[a:b], x0, s32, out=outs["fp32_array"][a:b]) for a, b in sl]),
        "fp64_array": (lambda: m64.full(p64[:w], x0, s64, out=outs["fp64_array"][:w]),
                       [lambda a=a, b=b: m64.full(p64[a:b], x0, s64, out=outs["fp64_array"][a:b]) for a, b in sl]),
        "robust_int64_grid": (lambda: rob.eval(p32[:w], out=outs["robust_int64_grid"][:w]),
                              [lambda a=a, b=b: rob.eval(p32[a:b], out=outs["robust_int64_grid"][a:b])
                               for a, b in sl]),
    }
    res = {}
    for name, (warm, parts) in runs.items():
        warm()
        torch.cuda.synchronize()
        ms = 0.0
        for k, part in enumerate(parts):
            t0 = time.perf_counter()
            part()
            torch.cuda.synchronize()
            ms += (time.perf_counter() - t0) * 1e3
            if len(parts) > 1:
                log(f"{name}: slice {k + 1}/{len(parts)} done, {ms:.1f} ms so far")
        res[name] = {"ms": ms, "queries_per_s": nq / (ms * 1e-3), "pairs_per_s": nq * float(nseg) / (ms * 1e-3)}
        log(f"{name}: {ms:.1f} ms = {nq / ms * 1e3 / 1e6:.2f} M queries/s")
    ref = outs["fp64_array"] > 0.5
    for name in ("fp32_grid", "fp32_array", "robust_int64_grid"):
        diff = ((outs[name] > 0.5) != ref).nonzero().flatten()
        res[name]["labels_differing_from_fp64"] = int(diff.numel())
        first = diff[:16].tolist()  # which queries, with both values (for a look with the CPU reference)
        res[name]["first_differing"] = [[i, float(outs[name][i]), float(outs["fp64_array"][i])] for i in first]
    d = (outs["fp32_array"].double() - outs["fp64_array"]).abs()
    res["fp32_array"]["max_abs_diff_from_fp64"] = float(d.max())
    log(f"labels differing from FP64: " + ", ".join(f"{k} {v['labels_differing_from_fp64']}" for k, v in res.items()
                                                    if "labels_differing_from_fp64" in v))
    line = {
        "metric": "GWN queries/s, 256^3 grid over the 1M-tri holed torus (plain FP32, FP64, robust int64_grid)",
        "value": res["fp32_grid"]["queries_per_s"], "unit": "queries/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": res["fp32_grid"]["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded holed torus + 256^3 lattice)",
        "config": {"workload": f"256^3 lattice over the {len(idx) // 3}-triangle holed torus ({nseg} boundary "
                               f"segments), full GWN", "parallelism": "single GPU"},
        "methods": res,
    }
    for h in (grid, m32, m64, rob):
        h.close()
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--train-samples", type=int, default=1 << 24)
    ap.add_argument("--epochs", type=int, default=16)
    ap.add_argument("--val-samples", type=int, default=1 << 23,
                    help="held-out points the routing band is calibrated on (the band must hold on every grid point)")
    ap.add_argument("--max-freq", type=float, default=64.0)
    ap.add_argument("--eps-frac", type=float, default=0.0010,
                    help="near-surface mask radius as a fraction of the box diagonal")
    ap.add_argument("--voxel-frac", type=float, default=0.25, help="mask voxel edge as a fraction of the mask radius")
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--audit-every", type=int, default=1024)
    ap.add_argument("--params-cache", type=str, default=None,
                    help="npz file: load the trained surrogate (parameters, delta, fit record) from it if it exists, "
                         "else train and save there (for profiler runs that should not retrain)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["surrogate", "batched", "plain256"], default="surrogate",
                    help="surrogate: the metric's 512^3 surrogate+fallback vs plain run (default); batched: the "
                         "uneven many-mesh configuration in persistent launches; plain256: the plain FP32, FP64 and "
                         "robust int64_grid evaluations of the 256^3 lattice over the same torus")
    ap.add_argument("--batch-meshes", type=int, default=4096)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 1 GiB device buffer 3 times (the known-size reference for rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()

    if args.copy_calibration:
        return copy_calibration()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (queries are independent, so the grid "
              "would shard by z-slab with the mesh state replicated per GPU). Run with --gpus 1.", file=sys.stderr)
        return 2

    import torch

    import antipodal_amd as agwn
    from antipodal_amd import meshgen as mg

    if not agwn.device_available():
        raise SystemExit(f"no HIP device: {agwn.last_error()}")
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    if args.workload == "batched":
        return bench_batched(args, agwn, mg, torch)
    if args.workload == "plain256":
        return bench_plain256(args, agwn, mg, torch)

    t = time.perf_counter()
    verts, idx = mg.holed_torus_mesh()
    v32 = verts.astype(np.float32)
    tb = time.perf_counter()
    seg32 = agwn.build_boundary_segments(v32, idx)
    boundary_s = time.perf_counter() - tb
    lo, hi = mg.expanded_bbox(verts)
    nres = args.grid
    start, dx, dy, dz = mg.grid_axes(lo, hi, nres)
    nq, nseg, ntri = nres ** 3, len(seg32), len(idx) // 3
    # the evaluator's integration / pole axis: the smallest axis (all equal -> x), x0 = its normalised step
    x0 = np.array([1.0, 0.0, 0.0], dtype=np.float32)
    log(f"inputs: {ntri} triangles, {nseg} segments, {nq} queries ({time.perf_counter() - t:.1f} s)")

    t = time.perf_counter()
    grid = agwn.HipGridGwn().set_scene(v32, idx, seg32)
    mesh = agwn.HipMesh(v32, idx, dtype=np.float32)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t
    log(f"scene + LBVH builds {build_s * 1e3:.1f} ms")

    # ---- plain exact method on the lattice -------------------------------------------------------------------------
    w = torch.empty(nq, dtype=torch.float32, device=dev)
    grid.eval_grid(start, dx, dy, dz, nres, nres, nres, out=w)  # warm-up (prepares the per-direction BVH state)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
        grid.eval_grid(start, dx, dy, dz, nres, nres, nres, out=w)
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3 / args.plain_steps
    exact = (w > 0.5).to(torch.int8)
    log(f"plain: {plain_ms:.1f} ms/step = {nq / plain_ms * 1e3 / 1e6:.2f} Mq/s")

    # ---- surrogate: fit (outside the timed region) -------------------------------------------------------------------
    cache = Path(args.params_cache) if args.params_cache else None
    if cache is not None and cache.exists():
        with np.load(cache, allow_pickle=False) as z:
            fit = json.loads(str(z["fit"]))
            net = agwn.GwnSurrogate(lo.astype(np.float32), hi.astype(np.float32), max_freq=fit["max_freq"])
            net.set_params(z["params"])
            delta = float(z["delta"])
        net.set_surface(v32, idx, fit["eps"], fit.get("voxel", 0.0))
        log(f"surrogate loaded from {cache}")
    else:
        net, delta, fit = agwn.fit_surrogate(v32, idx, seg32, x0, lo.astype(np.float32), hi.astype(np.float32),
                                             n_train=args.train_samples, n_val=args.val_samples, epochs=args.epochs,
                                             batch=args.batch, lr=args.lr, seed=0, mesh=mesh,
                                             max_freq=args.max_freq,
                                             eps=args.eps_frac * float(np.linalg.norm(hi - lo)),
                                             voxel=args.voxel_frac * args.eps_frac * float(np.linalg.norm(hi - lo)))
        fit = {k: v for k, v in fit.items() if k not in ("val_pred", "val_exact")}
        if cache is not None:
            np.savez(cache, params=net.get_params(), delta=delta, fit=json.dumps(fit))
    log(f"fit: {fit['fit_s']:.1f} s (train {fit['train_s']:.1f} s), losses {fit['losses'][0]:.4g} -> "
        f"{fit['losses'][-1]:.4g}, validation wrong {fit['val_wrong']}/{fit['val_points']}, delta {delta:.4f}")

    # ---- surrogate + fallback steps ----------------------------------------------------------------------------------
    lab = torch.empty(nq, dtype=torch.int8, device=dev)
    st = None
    for k in range(args.warmup):
        _, st = net.label_grid(grid, start, dx, dy, dz, nres, nres, nres, delta, audit_every=args.audit_every,
                               audit_seed=1000 + k, labels=lab)
        log(f"warmup {k + 1}/{args.warmup}: {st}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for k in range(args.steps):
        _, st = net.label_grid(grid, start, dx, dy, dz, nres, nres, nres, delta, audit_every=args.audit_every,
                               audit_seed=k, labels=lab)
        stats.append(st)
    torch.cuda.synchronize()
    total_s = time.perf_counter() - t0
    ms_per_step = total_s * 1e3 / args.steps
    mislabels = int((lab != exact).sum().item())
    st = stats[-1]
    log(f"surrogate+fallback: {ms_per_step:.1f} ms/step = {nq / ms_per_step * 1e3 / 1e6:.2f} Mq/s, "
        f"speedup {plain_ms / ms_per_step:.2f}x, routed {st['fallback_fraction']:.4f}, mislabels {mislabels}")

    # ---- dominant kernel: the step's launches, re-run alone on the same inputs with HIP events on their stream ------
    # The last step's routed set is rebuilt exactly: the lattice points (same float32 formula as the kernels), the
    # mask flags and network outputs of the same kernels, the band, and the audit hash; its size must match the
    # step's own count.
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    P = lattice_torch(start, dx, dy, dz, nres, dev)
    near = net.near_surface(P).bool()
    pred = torch.empty(nq, dtype=torch.float32, device=dev)
    net.predict_grid(start, dx, dy, dz, nres, nres, nres, out=pred, stream=stream)
    ev0.record(stream)
    net.predict_grid(start, dx, dy, dz, nres, nres, nres, out=pred, stream=stream)
    ev1.record(stream)
    ev1.synchronize()
    infer_ms = ev0.elapsed_time(ev1)  # the route variant also writes labels and the list: a lower bound
    band = ~near & ((pred - 0.5).abs() <= delta)
    routed_mask = near | band
    if args.audit_every:
        ids = torch.arange(nq, device=dev, dtype=torch.int64)
        routed_mask |= ~routed_mask & audit_hash(ids, args.audit_every, args.steps - 1)
        del ids
    rp_t = torch.nonzero(routed_mask).flatten()
    if len(rp_t) != st["routed"]:
        log(f"warning: rebuilt routed set has {len(rp_t)} points, the step routed {st['routed']}")
    pts_r = P[rp_t].contiguous()
    del P
    seg_t = torch.from_numpy(seg32).to(dev)
    frac_r = torch.empty(len(rp_t), dtype=torch.float32, device=dev)
    agwn.eval_gwnr_mesh_batch_fractional(pts_r, x0, seg_t, out=frac_r, stream=stream)
    ev0.record(stream)
    agwn.eval_gwnr_mesh_batch_fractional(pts_r, x0, seg_t, out=frac_r, stream=stream)
    ev1.record(stream)
    ev1.synchronize()
    frac_ms = ev0.elapsed_time(ev1)
    cnt_r = torch.empty(len(rp_t), dtype=torch.int32, device=dev)
    ev0.record(stream)
    agwn.eval_gwnr_mesh_batch_integer(mesh, pts_r, x0, out=cnt_r, stream=stream)
    ev1.record(stream)
    ev1.synchronize()
    int_ms = ev0.elapsed_time(ev1)
    rp = rp_t
    # sanity on the same points: full = fractional + integer
    full_r = torch.empty_like(frac_r)
    agwn.eval_gwnr_mesh_batch(mesh, pts_r, x0, seg_t, out=full_r, stream=stream)
    chk = (full_r - frac_r - cnt_r.to(torch.float32)).abs().max().item()
    if not chk < 1e-4:
        raise SystemExit(f"full != fractional + integer (max diff {chk})")
    # the step's exact stage on exactly the routed points: integer term, far-field decision, full boundary sum for the
    # points it leaves open (the point-labelling entry with delta = +inf routes every point; it also runs inference
    # on them, 5 % of the grid's, checks the boundary against the one its plan was built from, and its host
    # synchronisations are inside the interval: an upper bound of the step's exact stage)
    lab_r = torch.empty(len(rp_t), dtype=torch.int8, device=dev)
    net.label(mesh, pts_r, x0, seg_t, float("inf"), labels=lab_r, stream=stream)
    ev0.record(stream)
    net.label(mesh, pts_r, x0, seg_t, float("inf"), labels=lab_r, stream=stream)
    ev1.record(stream)
    ev1.synchronize()
    stage_ms = ev0.elapsed_time(ev1)
    if not torch.equal(lab_r, exact[rp_t]):
        raise SystemExit("exact-stage labels of the routed points differ from the plain run's")
    log(f"kernels per step: inference {infer_ms:.2f} ms; on the {len(rp)} routed points: exact stage at most {stage_ms:.2f} ms "
        f"(integer terms by packet traversal in index order, decision, full sums; timed through the point-labelling "
        f"entry, which also runs inference on them); on the same points the public "
        f"integer entry (rays sorted by their projected origin, then packets) takes {int_ms:.2f} ms and the full "
        f"boundary sum {frac_ms:.2f} ms")
    if stage_ms > infer_ms:
        achieved = len(rp) * nseg * FLOP_PER_PAIR / (frac_ms * 1e-3) / 1e12
        roof = {"bound": "valu", "kernel": "k_frac_f32_chain (full boundary sum on the routed points)",
                "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "kernel_ms": frac_ms,
                "flop_per_launch": len(rp) * nseg * FLOP_PER_PAIR, "launch_points": len(rp)}
        traffic_key = "k_frac_f32_chain"
    else:
        flop = nq * MLP_FLOP_PER_POINT_LATTICE + 3 * nres * LATTICE_TABLE_FLOP_PER_ROW
        achieved = flop / (infer_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "kernel": "k_mlp_infer (lattice tables)", "achieved": achieved,
                "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_BF16_TFLOPS, "kernel_ms": infer_ms,
                "flop_per_launch": flop, "launch_points": nq,
                "network_equivalent_tflops": nq * MLP_FLOP_PER_POINT / (infer_ms * 1e-3) / 1e12}
        traffic_key = "k_mlp_infer"
    # HBM bytes per launch of that kernel, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes of this
    # workload (profiles/traffic.json, scaled by the copy calibration); null when none is committed
    roof["traffic"] = None
    tj = ROOT / "profiles" / "traffic.json"
    if tj.exists():
        try:
            tr = json.loads(tj.read_text())
            if tr.get("workload_grid") == nres:
                roof["traffic"] = tr[traffic_key]["hbm_bytes_per_launch"]
        except (KeyError, ValueError):
            pass
    roof["traffic_unit"] = "bytes per launch (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE, see profiles/)"

    cpu = None
    if not args.no_cpu:
        rng = np.random.default_rng(0)
        sample = lattice_points(start, dx, dy, dz, nres, nres, nres,
                                np.sort(rng.choice(nq, size=min(nq, 1_000_000), replace=False)))
        log("cpu baseline ...")
        cpu = cpu_baseline(sample, seg32, x0, args.cpu_seconds, f"{nres}^3 grid")
        if cpu:
            log(f"cpu baseline {cpu['value']:.1f} queries/s on {cpu['cores']} threads")

    value = nq / (ms_per_step * 1e-3)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic (seeded holed torus + {nres}^3 lattice; surrogate trained on seeded samples)",
        "config": {
            "workload": f"inside/outside labels of a {nres}^3 grid over the 1,004,810-triangle holed noisy torus: "
                        f"surrogate (BF16 MLP, 4 x 256 hard-SiLU units) + exact FP32 fallback",
            "deviations": ["hard-SiLU z*clamp(z/4+1/2,0,1) in place of SiLU (cheaper, routes fewer points)",
                           "FP32 pairs accumulated in FP64 in place of FP32 two-sum (cheaper and more accurate)",
                           "FP64 re-evaluation of pairs chosen by a conditioning test, not by a 1e-4 distance"],
            "triangles": ntri,
            "segments": nseg,
            "queries": nq,
            "parallelism": "single GPU",
        },
        "plain": {"value": nq / (plain_ms * 1e-3), "unit": "queries/s", "ms_per_step": plain_ms,
                  "steps": args.plain_steps, "method": "exact full GWN (fractional + integer), FP32"},
        "speedup_vs_plain": plain_ms / ms_per_step,
        # one-off preprocessing: boundary extraction (host), scene + LBVH builds, and the surrogate fit (samples,
        # their exact labels, training, calibration); a single evaluation that pays all of it, and the number of
        # evaluations after which the surrogate's fit has paid for itself against plain
        "with_preprocessing": {
            "value": nq / (boundary_s + build_s + fit["fit_s"] + ms_per_step * 1e-3),
            "unit": "queries/s",
            "preprocessing_s": boundary_s + build_s + fit["fit_s"],
            "boundary_s": boundary_s, "build_s": build_s, "fit_s": fit["fit_s"],
            "break_even_evaluations": fit["fit_s"] / max((plain_ms - ms_per_step) * 1e-3, 1e-9),
        },
        "mislabels": mislabels,
        "surrogate": {
            "delta": delta,
            "eps": fit["eps"],
            "max_freq": args.max_freq,
            "fallback_fraction": st["fallback_fraction"],
            "band": st["band"],
            "near_surface": st["near"],
            "audit": st["audit"],
            "audit_disagree": sum(s["audit_disagree"] for s in stats),
            "train_samples": args.train_samples,
            "epochs": args.epochs,
            "batch": args.batch,
            "train_s": fit["train_s"],
            "fit_s": fit["fit_s"],
            "loss_first_last": [fit["losses"][0], fit["losses"][-1]],
            "val_wrong": fit["val_wrong"],
            "val_exact_disagree": fit.get("val_exact_disagree", 0),
            "val_points": fit["val_points"],
        },
        "kernels_ms": {"k_mlp_infer": infer_ms, "exact_stage_on_routed": stage_ms,
                       "integer_entry_on_routed_for_comparison": int_ms,
                       "full_boundary_sum_on_routed_for_comparison": frac_ms},
        "roofline": roof,
        "build_ms": build_s * 1e3,
        "cpu_baseline": cpu,
    }
    net.close()
    mesh.close()
    grid.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
