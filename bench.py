"""Benchmark of the sparse 3D FFT path on one MI355X.

Workload (the headline configuration): 128^3 grid, C2C, FP64, sparse set =
all centred triplets with x^2+y^2+z^2 <= 40^2 (267,761 values), values
U[-1,1) from SplitMix64. One step = one backward (values -> internal GPU
space buffer) plus one forward with full scaling (space buffer -> values),
device-resident input and output, SPFFT_EXEC_ASYNCHRONOUS, rotating over
`--rotate` independent transforms so the working set (about 77 MB of buffers
per transform) exceeds the 256 MB Infinity Cache.

Prints one JSON line on stdout (everything else goes to stderr):
  value    = pairs/s (whole job), timed with torch.cuda.synchronize() around
             exactly --steps pairs after --warmup untimed pairs;
  roofline = dominant kernel: algorithmic bytes per launch / mean launch
             time from HIP events recorded on the library's own stream during
             the timed region (spfft_amd_profiling_*);
  cpu_baseline = the C oracle (restatement of the reference's host path)
             timed on the host's cores on a bounded sample of the same
             workload.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(name, ns, ng, r2c, real_bytes):
    """Algorithmic (compulsory) HBM bytes one launch of a kernel moves.

    Per direction the compulsory traffic is values (2 reals per value), the
    int32 value->slot map and the dense space grid. The z kernels own the
    value and map terms, the x / xy kernels own the dense grid; the generic
    path's y kernels only move intermediates and own none."""
    cplx = 2 * real_bytes
    grid = ng * (real_bytes if r2c else cplx)
    if name.startswith("k_z_"):
        return ns * (cplx + 4)
    if name.startswith("k_x_") or name.startswith("k_xy_"):
        return grid
    return 0


def read_profile(lib):
    names = ctypes.create_string_buffer(64 * 32)
    counts = (ctypes.c_longlong * 32)()
    total = (ctypes.c_double * 32)()
    n = lib.spfft_amd_profiling_report(32, names, counts, total)
    out = {}
    for k in range(n):
        nm = names.raw[64 * k:64 * (k + 1)].split(b"\0")[0].decode()
        out[nm] = (counts[k], total[k])
    return out


def cpu_baseline(t, dims, values, seconds):
    """The C oracle on the host cores: warm once, then pairs until `seconds`."""
    import numpy as np
    from oracle import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    plan = orc.OraclePlan(False, dims, t, nthreads=threads)
    space = plan.backward(values)
    plan.forward(space, True)
    pairs = 0
    t0 = time.perf_counter()
    while True:
        space = plan.backward(values)
        plan.forward(space, True)
        pairs += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return {"value": pairs / el, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"{pairs} backward+forward pairs of the same 128^3 r=40 C2C FP64 workload "
                      f"({el:.1f} s), C/OpenMP restatement of the reference host path"}


def batched_bench(args, sp, torch, np, t, dims, ns, ng, plain_value):
    """Batched form: --batch bands of the same 128^3 r=40 set through one
    multi_transform_backward/forward call each, FP64 and FP32; band-pairs/s
    and the FP32 / plain-FP64 ratio. Device-resident, values as above."""
    from tests.workloads import complex_values
    out = {"bands": args.batch}
    for prec in ("double", "float"):
        real = np.float64 if prec == "double" else np.float32
        cls = sp.Transform if prec == "double" else sp.TransformFloat
        trs, ins, outs = [], [], []
        for k in range(args.batch):
            tr = cls(-1, sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, *dims, ns, sp.SPFFT_INDEX_TRIPLETS, t)
            tr.set_execution_mode(sp.SPFFT_EXEC_ASYNCHRONOUS)
            v = complex_values(0x5EED0100 + k, ns)
            ins.append(torch.from_numpy(v.view(np.float64).astype(real)).cuda())
            outs.append(torch.empty(2 * ns, dtype=ins[-1].dtype, device="cuda"))
            trs.append(tr)
        locs = [sp.SPFFT_PU_GPU] * args.batch
        scal = [sp.SPFFT_FULL_SCALING] * args.batch

        def step():
            sp.multi_transform_backward(trs, ins, locs)
            sp.multi_transform_forward(trs, locs, outs, scal)

        for _ in range(3):
            step()
        torch.cuda.synchronize()
        steps = max(5, args.steps // args.batch)
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        rate = steps * args.batch / el
        rb = np.dtype(real).itemsize
        pair_bytes = 2 * (ns * (2 * rb + 4) + ng * 2 * rb)
        out[prec] = {"band_pairs_per_s": rate, "ms_per_call_pair": el / steps * 1e3,
                     "frac_hbm": rate * pair_bytes / 1e9 / HBM_PEAK_GBS}
        log(f"batched {prec}: {rate:.1f} band-pairs/s ({args.batch} bands per call)")
        del trs, ins, outs
        torch.cuda.empty_cache()
    out["fp32_vs_plain_fp64"] = out["float"]["band_pairs_per_s"] / plain_value
    out["fp64_vs_plain_fp64"] = out["double"]["band_pairs_per_s"] / plain_value
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--rotate", type=int, default=8)
    ap.add_argument("--precision", choices=["double", "float"], default="double")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch", type=int, default=64,
                    help="transforms per multi_transform call in the batched section (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the batched "
              "path shards by bands/k-points with no communication and is not run on "
              "several GPUs here.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import spfft_amd as sp
    from tests.workloads import complex_values, sphere_triplets

    dims = (128, 128, 128)
    radius = 40
    t = sphere_triplets(radius)
    ns = len(t)
    ng = dims[0] * dims[1] * dims[2]
    real = np.float64 if args.precision == "double" else np.float32
    real_bytes = np.dtype(real).itemsize
    cls = sp.Transform if args.precision == "double" else sp.TransformFloat

    trs, ins, outs = [], [], []
    for k in range(args.rotate):
        tr = cls(-1, sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, *dims, ns, sp.SPFFT_INDEX_TRIPLETS, t)
        tr.set_execution_mode(sp.SPFFT_EXEC_ASYNCHRONOUS)
        v = complex_values(0x5EED0002 + k, ns)
        ins.append(torch.from_numpy(v.view(np.float64).astype(real)).cuda())
        outs.append(torch.empty(2 * ns, dtype=ins[-1].dtype, device="cuda"))
        trs.append(tr)

    def pair(i):
        k = i % args.rotate
        trs[k].backward(ins[k].data_ptr(), sp.SPFFT_PU_GPU)
        trs[k].forward(sp.SPFFT_PU_GPU, outs[k].data_ptr(), sp.SPFFT_FULL_SCALING)

    for i in range(args.warmup):
        pair(i)
    torch.cuda.synchronize()
    log(f"warmup done ({args.warmup} pairs)")

    # Timed region: exactly --steps pairs, synchronised on both sides.
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        pair(i)
    t_enq = time.perf_counter() - t0
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = args.steps / elapsed
    log(f"timed {args.steps} pairs in {elapsed * 1e3:.1f} ms -> {value:.1f} pairs/s "
        f"(host enqueue {t_enq / args.steps * 1e6:.1f} us/pair)")

    # Same steps again with per-kernel HIP events on the library's stream.
    sp.lib.spfft_amd_profiling_enable(1)
    for i in range(args.steps):
        pair(i)
    torch.cuda.synchronize()
    sp.lib.spfft_amd_profiling_enable(0)
    prof = read_profile(sp.lib)

    kernels = {}
    for nm, (cnt, tot) in prof.items():
        mean_s = tot / cnt * 1e-3
        b = kernel_bytes(nm, ns, ng, False, real_bytes)
        kernels[nm] = {"launches": cnt, "mean_us": mean_s * 1e6,
                       "alg_bytes": b, "gbs": b / mean_s / 1e9 if b else 0.0}
        log(f"  {nm:14s} {cnt:6d} launches  mean {mean_s * 1e6:8.2f} us  "
            f"alg {b / 1e6:7.2f} MB  {kernels[nm]['gbs']:8.1f} GB/s")
    dominant = max(kernels, key=lambda k: kernels[k]["mean_us"] * kernels[k]["launches"])
    d = kernels[dominant]
    pair_kernel_s = sum(k["mean_us"] * k["launches"] for k in kernels.values()) / args.steps * 1e-6
    pair_bytes = 2 * (ns * (2 * real_bytes + 4) + ng * 2 * real_bytes)

    # Measured HBM bytes per launch of the dominant kernel, from rocprofv3
    # FETCH_SIZE / WRITE_SIZE runs, when a summary for this kernel exists.
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dominant}_{args.precision}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")

    result = {
        "metric": "bwd+fwd pairs/s @128^3 FP64 (sparse sphere r=40, C2C)",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "double" else "f32",
        "data": "synthetic",
        "config": {
            "workload": f"128^3 C2C {'FP64' if args.precision == 'double' else 'FP32'} sparse "
                        f"sphere r=40 ({ns} values), backward+forward(full scaling), "
                        f"device-resident, async, rotating over {args.rotate} transforms",
            "grid": list(dims), "num_values": ns, "rotate": args.rotate,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": d["gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": d["gbs"] / HBM_PEAK_GBS,
            "traffic": traffic,
        },
        "pair_roofline": {
            "alg_bytes_per_pair": pair_bytes,
            "kernel_time_per_pair_us": pair_kernel_s * 1e6,
            "achieved": pair_bytes / pair_kernel_s / 1e9,
            "frac": pair_bytes / pair_kernel_s / 1e9 / HBM_PEAK_GBS,
            "wall_achieved": pair_bytes * value / 1e9,
            "wall_frac": pair_bytes * value / 1e9 / HBM_PEAK_GBS,
        },
        "kernels": kernels,
    }
    if args.batch > 0:
        result["batched"] = batched_bench(args, sp, torch, np, t, dims, ns, ng, value)
    if not args.no_cpu_baseline:
        v0 = complex_values(0x5EED0002, ns)
        result["cpu_baseline"] = cpu_baseline(t, dims, v0, args.cpu_seconds)
        log(f"cpu baseline: {result['cpu_baseline']['value']:.2f} pairs/s on "
            f"{result['cpu_baseline']['cores']} threads")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
