"""Benchmark: full GWN (fractional + integer term) of a 256^3 query grid against the 1M-triangle open mesh, one MI355X.

Workload (the plain-method GPU configuration): seeded noisy torus R=1, r=0.35, 1353 x 473 quads (2 triangles each),
sigma = 0.01 r Gaussian vertex noise, 400 disk holes of radius 0.05 -> 1,004,810 triangles, 31,376 unit-weight
boundary segments; queries = the 256^3 lattice over the bounding box expanded by 10 %; x0 = normalize(0.31,0.57,0.76);
FP32. A step is one eval_gwnr_mesh_batch over all 16,777,216 queries, inputs resident in HBM (device pointers through
the C ABI, on torch's current stream).

Reported beside it:
  roofline      the dominant kernel (the fractional boundary sum, k_frac_f32): algorithmic FLOP per launch =
                n_queries * n_segments * 56 (the per-pair count fixed in the design notes), over its average launch
                time measured with HIP events on the launch stream, against MI355X FP32 vector peak 157.3 TFLOP/s
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
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP32_TFLOPS = 157.3
FLOP_PER_PAIR = 56


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(verts32, idx, pos32_sample, seg32, x0, budget_s: float) -> dict | None:
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
        "kind": "reference",
        "sample": f"{n} seeded queries of the 256^3 grid x {len(seg32)} segments, eval_gwnr_mesh_batch_fractional<float> "
                  f"with OpenMPDispatcher ({threads} threads), {dt:.1f} s; integer term not timed (taken as 0)",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 1 GiB device buffer 3 times (the known-size reference for rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch

        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")  # 1 GiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": a.numel() * 4, "copies": 3}), flush=True)
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (queries shard trivially by "
              "z-slab; see the design notes). Run with --gpus 1.", file=sys.stderr)
        return 2

    import torch

    import antipodal_amd as agwn
    from antipodal_amd import meshgen as mg

    if not agwn.device_available():
        raise SystemExit(f"no HIP device: {agwn.last_error()}")
    dev = torch.device("cuda:0")

    t = time.perf_counter()
    verts, idx = mg.holed_torus_mesh()
    v32 = verts.astype(np.float32)
    seg32 = agwn.build_boundary_segments(v32, idx)
    lo, hi = mg.expanded_bbox(verts)
    n = args.grid
    pos32 = mg.grid_points(lo, hi, n, dtype=np.float32)
    x0 = mg.example_x0(np.float32)
    nq, nseg, ntri = len(pos32), len(seg32), len(idx) // 3
    log(f"inputs: {ntri} triangles, {nseg} segments, {nq} queries ({time.perf_counter() - t:.1f} s)")

    pos_t = torch.from_numpy(pos32).to(dev)
    seg_t = torch.from_numpy(seg32).to(dev)
    out_t = torch.empty(nq, dtype=torch.float32, device=dev)
    frac_t = torch.empty(nq, dtype=torch.float32, device=dev)

    t = time.perf_counter()
    mesh = agwn.HipMesh(v32, idx)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t
    log(f"LBVH build {build_s * 1e3:.1f} ms")

    for w in range(args.warmup):
        agwn.eval_gwnr_mesh_batch(mesh, pos_t, x0, seg_t, out=out_t)
        torch.cuda.synchronize()
        log(f"warmup {w + 1}/{args.warmup} done")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        agwn.eval_gwnr_mesh_batch(mesh, pos_t, x0, seg_t, out=out_t)
    torch.cuda.synchronize()
    total_s = time.perf_counter() - t0
    ms_per_step = total_s * 1e3 / args.steps
    log(f"timed {args.steps} steps: {ms_per_step:.2f} ms/step")

    # dominant kernel: the fractional boundary sum, timed with HIP events on its launch stream
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(1, min(3, args.steps))
    ev0.record(stream)
    for _ in range(reps):
        agwn.eval_gwnr_mesh_batch_fractional(pos_t, x0, seg_t, out=frac_t, stream=stream)
    ev1.record(stream)
    ev1.synchronize()
    frac_ms = ev0.elapsed_time(ev1) / reps
    int_t = torch.empty(nq, dtype=torch.int32, device=dev)
    ev0.record(stream)
    agwn.eval_gwnr_mesh_batch_integer(mesh, pos_t, x0, out=int_t, stream=stream)
    ev1.record(stream)
    ev1.synchronize()
    int_ms = ev0.elapsed_time(ev1)
    pairs = nq * nseg
    achieved = pairs * FLOP_PER_PAIR / (frac_ms * 1e-3) / 1e12
    log(f"fractional {frac_ms:.2f} ms ({pairs / (frac_ms * 1e-3) / 1e12:.3f} Tpair/s), integer {int_ms:.2f} ms")

    # sanity: the full GWN must be frac + count
    chk = (out_t - frac_t - int_t.to(torch.float32)).abs().max().item()
    if not chk < 1e-4:
        raise SystemExit(f"full != fractional + integer (max diff {chk})")

    cpu = None
    if not args.no_cpu:
        rng = np.random.default_rng(0)
        sample = pos32[np.sort(rng.choice(nq, size=min(nq, 200_000), replace=False))]
        log("cpu baseline ...")
        cpu = cpu_baseline(v32, idx, sample, seg32, x0, args.cpu_seconds)
        if cpu:
            log(f"cpu baseline {cpu['value']:.1f} queries/s on {cpu['cores']} threads")

    # HBM bytes per launch of the dominant kernel, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    # (profiles/traffic.json, scaled by the copy calibration); null when no profile of this workload is committed
    traffic = None
    tj = ROOT / "profiles" / "traffic.json"
    if tj.exists() and n == 256:
        try:
            traffic = json.loads(tj.read_text())["k_frac_f32"]["hbm_bytes_per_launch"]
        except (KeyError, ValueError):
            traffic = None

    value = nq / (ms_per_step * 1e-3)
    line = {
        "metric": "GWN queries/s, 1M-tri open mesh; surrogate+fallback vs plain, 0 mislabels",
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
        "data": "synthetic (seeded holed torus + 256^3 lattice)",
        "config": {
            "workload": f"plain full GWN (fractional + integer) of a {n}^3 grid, holed noisy torus",
            "triangles": ntri,
            "segments": nseg,
            "queries": nq,
            "method": "plain (exact, no surrogate)",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "valu",
            "kernel": "k_frac_f32",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE, see profiles/)",
            "kernel_ms": frac_ms,
            "flop_per_launch": pairs * FLOP_PER_PAIR,
        },
        "integer_term_ms": int_ms,
        "lbvh_build_ms": build_s * 1e3,
        "cpu_baseline": cpu,
    }
    mesh.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
