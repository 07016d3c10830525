"""Benchmark: FaSTCC sparse contraction on one MI355X.

One step = one full contraction (index build, planning, count pass for the
exact planner, accumulate, compaction into the SoA output) of the synthetic
workload with its inputs already resident in HBM.

Default workload: the order-4 self-contraction T(6186x24x77x32, 5.3M nnz,
mode-0 Zipf(1.1), lognormal values) on modes {0,1}, exact-symbolic planning.
Prints one JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "fastcc-benchmark_amd"), ROOT]

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4"])
    ap.add_argument("--planner", default="exact", choices=["exact", "learned"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=0, help="left-tile sample stride for the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--tile-hint", type=int, default=0, help="rows per left tile (0 = planner)")
    ap.add_argument("--col-window", type=int, default=0, help="columns per dense unit (0 = planner)")
    ap.add_argument("--workgroups", type=int, default=0, help="persistent workgroups (0 = planner)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is described in DESIGN.md, not implemented)"}))
        sys.exit(2)

    import numpy as np
    import torch
    if args.copy_calibration:
        n = args.copy_calibration << 20
        src = torch.ones(n // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n}))
        return
    from fastcc_amd import Context, DeviceTensor
    from fastcc_amd import workloads as W

    wl = W.FULL[args.workload]()
    ta, tb = wl.tensors()
    dev = torch.device("cuda:0")
    ca = torch.from_numpy(ta.coords.view(np.int32)).to(dev).contiguous()
    va = torch.from_numpy(ta.vals).to(dev).contiguous()
    da = DeviceTensor.from_torch(ca, va)
    if wl.self_contraction:
        db = da
    else:
        cb = torch.from_numpy(tb.coords.view(np.int32)).to(dev).contiguous()
        vb = torch.from_numpy(tb.vals).to(dev).contiguous()
        db = DeviceTensor.from_torch(cb, vb)
    torch.cuda.synchronize()

    ctx = Context(0)
    opts = dict(planner=args.planner, tile_hint=args.tile_hint, col_window=args.col_window,
                workgroups=args.workgroups)
    if args.planner == "learned":
        params = os.path.join(ROOT, "fastcc-benchmark_amd", "fastcc_amd", "mlp_params.npy")
        ctx.set_mlp_params(np.load(params))

    def step():
        r = ctx.contract(da, list(wl.a_modes), db, list(wl.b_modes), **opts)
        n, st = r.nnz, r.stats
        r.free()
        return n, st

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    nnz = 0
    for _ in range(args.steps):
        nnz, st = step()
        stats.append(st)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(args.steps, 1)

    products = stats[-1]["products"]
    bytes_alg = stats[-1]["bytes_alg"]
    k_ms = float(np.mean([s["ms_numeric_kernel"] for s in stats]))
    dev_ms = float(np.mean([s["ms_total"] for s in stats]))
    # dominant kernel: numeric accumulate; FP64-bound when 2*products/peak exceeds bytes/HBM peak
    t_fp64 = 2.0 * products / (FP64_PEAK_TFLOPS * 1e12)
    t_hbm = bytes_alg / (HBM_PEAK_GBS * 1e9)
    if t_fp64 >= t_hbm:
        achieved = 2.0 * products / (k_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": "k_numeric",
                "note": "fp64 multiply-adds of the numeric accumulate kernel; fp64 peak (no MFMA used)"}
    else:
        achieved = bytes_alg / (k_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_numeric"}

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        stride = args.cpu_stride or {"c1": 1, "c2": 1, "c3": 4096, "c4": 32}[args.workload]
        r = oracle.contract_tiled(ta, wl.a_modes, tb, wl.b_modes, tile_size=128, threads=threads,
                                  high_precision=False, sample_stride=stride, sink=args.workload in ("c3", "c4"),
                                  sort=False)
        cpu = {"value": r.nnz / r.seconds, "unit": "output nnz/s", "cores": r.threads, "kind": "port",
               "sample": f"left tiles t % {stride} == 0 at T=128 ({r.nnz} outputs, {r.products} products, "
                         f"{r.seconds:.2f} s); CPU {cpu_model()}",
               "seconds": r.seconds}

    line = {
        "metric": "output nnz/s (exact-symbolic planning)" if args.planner == "exact" else
                  "output nnz/s (learned planning)",
        "value": nnz / dt,
        "unit": "output nnz/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, csrc/fcc_gen.c)",
        "config": {"workload": wl.name, "description": wl.description, "planner": args.planner,
                   "output_nnz": int(nnz), "products": int(products)},
        "hbm_gbs_alg": bytes_alg / dt / 1e9,
        "device_ms": dev_ms,
        "phases_ms": {k: float(np.mean([s[k] for s in stats])) for k in
                      ("ms_index", "ms_plan", "ms_symbolic", "ms_numeric", "ms_overflow")},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
