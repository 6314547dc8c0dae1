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


def host_cpu():
    """CPU model and physical core count of this host (from /proc/cpuinfo)."""
    model, cores = "unknown", set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model == "unknown":
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
    except OSError:
        pass
    return model, len(cores) or (os.cpu_count() or 1)


def cpu_baseline(t, dims, values, seconds):
    """The C oracle (restatement of the reference's SPFFT_PU_HOST path) on the
    host: warm once, then backward+forward pairs until `seconds`. Threads =
    OMP_NUM_THREADS (the CPU share this job may use), else the physical cores."""
    from oracle import oracle as orc
    model, physical = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or physical
    orc.use_fft_library()  # the FFTW3 API as the reference calls it, when the host has one
    plan = orc.OraclePlan(False, dims, t, nthreads=threads)
    provider = orc.fft_provider()
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
            "cpu_model": model, "physical_cores": physical, "fft_provider": provider,
            "sample": f"{pairs} backward+forward pairs of the same 128^3 r=40 C2C FP64 workload "
                      f"({el:.1f} s) through the C/OpenMP restatement of the reference host "
                      f"path on {threads} threads, 1D FFTs by {provider}"}


def _bench_transforms(sp, torch, np, trs, ins, outs, steps, r2c=False):
    """Seconds per multi_transform backward+forward call over all transforms."""
    locs = [sp.SPFFT_PU_GPU] * len(trs)
    scal = [sp.SPFFT_FULL_SCALING] * len(trs)

    def step():
        sp.multi_transform_backward(trs, ins, locs)
        sp.multi_transform_forward(trs, locs, outs, scal)

    for _ in range(5):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def batched_bench(args, sp, torch, np):
    """Batched forms, each one multi_transform_backward + _forward call over
    all bands (one fused launch chain per direction), device-resident:
      - 128^3, sphere r=40, 64 bands, FP64 and FP32 (the FP32 / plain-FP64
        ratio is taken on this one);
      - 96^3, sphere r=30, 64 bands, FP64 and FP32, values N(0,1) exp(-|G|^2/900);
      - 96^3, 16 k-points x 32 bands, each k-point its own sphere |G+k|^2 <= r^2
        with k ~ U[-1/2,1/2)^3 and r ~ U[22,30] (uneven sets), FP32;
      - 192^3, sphere r=60, 8 bands, c ~ N(0,1), FP64 and FP32 (the shape of the
        differentiated workload, whose energy gradient is one such pair).
    Band-pairs/s and the fraction of the HBM roofline (compulsory bytes:
    values + map + dense grid per direction)."""
    from tests.workloads import complex_values, normal, sphere_triplets, uniform, uniform01
    out = {}
    cases = [("128^3 r=40 x64", 128, [40.0], [(0.0, 0.0, 0.0)], 64, ("double", "float"), "uniform"),
             ("96^3 r=30 x64", 96, [30.0], [(0.0, 0.0, 0.0)], 64, ("double", "float"), "gauss")]
    ks = uniform(0x5EED0004, 48).reshape(16, 3) * 0.5
    rs = 22.0 + 8.0 * uniform01(0x5EED0005, 16)
    cases.append(("96^3 16 k-points x32", 96, list(rs), [tuple(k) for k in ks], 32, ("float",),
                  "gauss"))
    # the differentiated workload's shape: 8 bands of 192^3 r=60, c ~ N(0,1)
    cases.append(("192^3 r=60 x8", 192, [60.0], [(0.0, 0.0, 0.0)], 8, ("double", "float"),
                  "normal"))
    for name, n, radii, kpts, bands, precs, dist in cases:
        if args.batch_cases and not any(c in name for c in args.batch_cases):
            continue
        dims = (n, n, n)
        sets = [sphere_triplets(r, kshift=k) for r, k in zip(radii, kpts)]
        rec = {"grid": n, "bands_per_set": bands, "sets": len(sets),
               "values_per_set": [len(t) for t in sets]}
        for prec in precs:
            real = np.float64 if prec == "double" else np.float32
            cls = sp.Transform if prec == "double" else sp.TransformFloat
            trs, ins, outs = [], [], []
            for si, t in enumerate(sets):
                g2 = np.sum(t.astype(np.float64) ** 2, axis=1)
                for b in range(bands):
                    tr = cls(-1, sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, *dims, len(t),
                             sp.SPFFT_INDEX_TRIPLETS, t)
                    tr.set_execution_mode(sp.SPFFT_EXEC_ASYNCHRONOUS)
                    seed = 0x5EED0100 + 1000 * si + b
                    if dist == "uniform":
                        v = complex_values(seed, len(t))
                    elif dist == "normal":
                        v = normal(seed, len(t)) + 1j * normal(seed + 7919, len(t))
                    else:
                        w = np.exp(-g2 / 900.0)
                        v = (normal(seed, len(t)) + 1j * normal(seed + 7919, len(t))) * w
                    ins.append(torch.from_numpy(v.view(np.float64).astype(real)).cuda())
                    outs.append(torch.empty(2 * len(t), dtype=ins[-1].dtype, device="cuda"))
                    trs.append(tr)
            # at least 20 calls (tens of ms of GPU work), so that clock ramps and
            # first-call costs (tuning, graph capture) stay out of the average
            steps = max(20, args.steps // len(trs))
            sec = _bench_transforms(sp, torch, np, trs, ins, outs, steps)
            rb = np.dtype(real).itemsize
            nv = sum(len(t) for t in sets) * bands
            # compulsory bytes: values and dense grid per band, the value map
            # once per distinct index set (bands of a set share it)
            maps = sum(len(t) for t in sets) * 4
            pair_bytes = 2 * (nv * 2 * rb + maps + len(trs) * n ** 3 * 2 * rb)
            rate = len(trs) / sec
            err = batch_error(sp, torch, np, trs, ins, outs, sets, bands, dims, prec)
            rec[prec] = {"band_pairs_per_s": rate, "ms_per_call_pair": sec * 1e3,
                         "frac_hbm": pair_bytes / sec / 1e9 / HBM_PEAK_GBS, **err}
            log(f"batched {name} {prec}: {rate:.1f} band-pairs/s, "
                f"{rec[prec]['frac_hbm'] * 100:.1f}% of HBM roofline, worst sampled rel L2 "
                f"vs FP64 {err['max_rel_l2_vs_fp64']:.2e}")
            del trs, ins, outs
            torch.cuda.empty_cache()
        out[name] = rec
    return out


def batch_error(sp, torch, np, trs, ins, outs, sets, bands, dims, prec):
    """Worst relative L2 error of sampled bands of the batch just timed
    (backward field and scaled forward values) against the library's FP64
    single-transform path on the same inputs; that path matches the FP64
    oracle to 1e-12 in the GPU parity tests. The batch runs once more with the
    sampled bands' fields in separate buffers (others in their space domain)."""
    n = dims[0]
    ntr = len(trs)
    sample = sorted({0, ntr // 3, (2 * ntr) // 3, ntr - 1})
    real = torch.float64 if prec == "double" else torch.float32
    fields = {k: torch.empty(2 * n ** 3, dtype=real, device="cuda") for k in sample}
    spaces = [fields[k].data_ptr() if k in fields else trs[k].space_domain_data(sp.SPFFT_PU_GPU)
              for k in range(ntr)]
    sp.multi_transform_backward(trs, ins, spaces)
    sp.multi_transform_forward(trs, spaces, outs, [sp.SPFFT_FULL_SCALING] * ntr)
    torch.cuda.synchronize()
    worst = 0.0
    for k in sample:
        t = sets[k // bands]
        ref = sp.Transform(-1, sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, *dims, len(t),
                           sp.SPFFT_INDEX_TRIPLETS, t)
        vin = ins[k].to(torch.float64)
        field = torch.empty(2 * n ** 3, dtype=torch.float64, device="cuda")
        ref.backward(vin.data_ptr(), field.data_ptr())
        vout = torch.empty(2 * len(t), dtype=torch.float64, device="cuda")
        ref.forward(field.data_ptr(), vout.data_ptr(), sp.SPFFT_FULL_SCALING)
        torch.cuda.synchronize()
        e_b = float(torch.linalg.vector_norm(fields[k].to(torch.float64) - field) /
                    torch.linalg.vector_norm(field))
        e_f = float(torch.linalg.vector_norm(outs[k].to(torch.float64) - vout) /
                    torch.linalg.vector_norm(vout))
        worst = max(worst, e_b, e_f)
        del ref, field, vout
    return {"max_rel_l2_vs_fp64": worst, "error_bands_sampled": sample}


def launch_floor(torch, kernel_us):
    """Single-launch floor at the size of the dominant kernel's grid: a fill
    (write only) and a copy of the 128^3 complex FP64 grid (33.5 MB) by
    torch's own kernels, rotating over 8 buffers, timed with HIP events.
    `kernel_frac_of_fill` = fill time / dominant kernel time."""
    n = 2 * 128 ** 3
    bufs = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(8)]
    out = {"bytes": n * 8}
    for name, fn in (("fill", lambda i: bufs[i % 8].fill_(1.0)),
                     ("copy", lambda i: bufs[i % 8].copy_(bufs[(i + 3) % 8]))):
        for i in range(20):
            fn(i)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for i in range(200):
            fn(i)
        e1.record()
        torch.cuda.synchronize()
        out[f"{name}_us"] = e0.elapsed_time(e1) / 200 * 1e3
    out["fill_gbs"] = out["bytes"] / out["fill_us"] / 1e3
    out["kernel_frac_of_fill"] = out["fill_us"] / kernel_us
    del bufs
    torch.cuda.empty_cache()
    log(f"launch floor: fill {out['fill_us']:.2f} us ({out['fill_gbs']:.0f} GB/s), copy "
        f"{out['copy_us']:.2f} us for 33.5 MB; dominant kernel at "
        f"{out['kernel_frac_of_fill'] * 100:.0f}% of the fill rate")
    return out


def plain_lengths_bench(args, sp, torch, np):
    """Single C2C FP64 transforms at other plane-wave lengths, sphere radius
    0.3125 N (40 at 128): 120, 100, 144 and 132 (z kernels and the split plane
    stage; 120, 100 and 132 on mixed radix schedules, 132 with a radix-11
    pass) and 143 (11 x 13: generic stage kernels). Backward+forward pairs/s
    and the fraction of the HBM roofline of the pair's compulsory bytes,
    rotating over 8 transforms; `kernels` names the kernels one pair
    launched."""
    from tests.workloads import complex_values, sphere_triplets
    out = {}
    for n in (120, 100, 144, 132, 143):
        dims = (n, n, n)
        t = sphere_triplets(0.3125 * n)
        trs, ins, outs = [], [], []
        for k in range(args.rotate):
            tr = sp.Transform(-1, sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, *dims, len(t),
                              sp.SPFFT_INDEX_TRIPLETS, t)
            tr.set_execution_mode(sp.SPFFT_EXEC_ASYNCHRONOUS)
            ins.append(torch.from_numpy(complex_values(0x5EED0006 + k, len(t)).view(np.float64))
                       .cuda())
            outs.append(torch.empty(2 * len(t), dtype=torch.float64, device="cuda"))
            trs.append(tr)

        def pair(i):
            k = i % args.rotate
            trs[k].backward(ins[k].data_ptr(), sp.SPFFT_PU_GPU)
            trs[k].forward(sp.SPFFT_PU_GPU, outs[k].data_ptr(), sp.SPFFT_FULL_SCALING)

        for i in range(args.warmup):
            pair(i)
        torch.cuda.synchronize()
        steps = max(50, args.steps // 2)
        t0 = time.perf_counter()
        for i in range(steps):
            pair(i)
        torch.cuda.synchronize()
        rate = steps / (time.perf_counter() - t0)
        sp.lib.spfft_amd_profiling_enable(1)
        pair(0)
        torch.cuda.synchronize()
        kernels = sorted(read_profile(sp.lib))
        sp.lib.spfft_amd_profiling_enable(0)
        path = "fixed-length" if any(k.startswith("k_z_backward_") and k != "k_z_backward"
                                     for k in kernels) else "generic"
        pair_bytes = 2 * (20 * len(t) + 16 * n ** 3)
        out[f"{n}^3"] = {"values": len(t), "pairs_per_s": rate,
                         "frac_hbm_wall": pair_bytes * rate / 1e9 / HBM_PEAK_GBS,
                         "path": path, "kernels": kernels}
        log(f"plain C2C {n}^3 r={0.3125 * n:.1f} ({len(t)} values, {path} kernels): "
            f"{rate:.1f} pairs/s, {out[f'{n}^3']['frac_hbm_wall'] * 100:.1f}% of HBM roofline")
        del trs, ins, outs
        torch.cuda.empty_cache()
    return out


def plain_r2c_bench(args, sp, torch, np):
    """128^3 R2C FP64, half sphere r=40 (x > 0, or x = 0 and y > 0, or
    x = y = 0 and z >= 0): single transforms, backward+forward pairs/s."""
    from tests.workloads import complex_values, sphere_triplets
    dims = (128, 128, 128)
    t = sphere_triplets(40, half=True)
    trs, ins, outs = [], [], []
    for k in range(args.rotate):
        tr = sp.Transform(-1, sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_R2C, *dims, len(t),
                          sp.SPFFT_INDEX_TRIPLETS, t)
        tr.set_execution_mode(sp.SPFFT_EXEC_ASYNCHRONOUS)
        ins.append(torch.from_numpy(complex_values(0x5EED0003 + k, len(t)).view(np.float64)).cuda())
        outs.append(torch.empty(2 * len(t), dtype=torch.float64, device="cuda"))
        trs.append(tr)

    def pair(i):
        k = i % args.rotate
        trs[k].backward(ins[k].data_ptr(), sp.SPFFT_PU_GPU)
        trs[k].forward(sp.SPFFT_PU_GPU, outs[k].data_ptr(), sp.SPFFT_FULL_SCALING)

    for i in range(args.warmup):
        pair(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        pair(i)
    torch.cuda.synchronize()
    rate = args.steps / (time.perf_counter() - t0)
    pair_bytes = 2 * (len(t) * 20 + 128 ** 3 * 8)
    log(f"plain R2C 128^3 r=40: {rate:.1f} pairs/s")
    return {"values": len(t), "pairs_per_s": rate,
            "frac_hbm_wall": pair_bytes * rate / 1e9 / HBM_PEAK_GBS}


def _counter_rows(folder):
    """(kernel name, counter value) rows of a rocprofv3 --pmc csv output."""
    import csv
    import glob
    rows = []
    for f in sorted(glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True)):
        with open(f) as fh:
            for row in csv.DictReader(fh):
                name = row["Kernel_Name"]
                for pre in ("void spfft::(anonymous namespace)::", "spfft::(anonymous namespace)::"):
                    name = name.replace(pre, "")
                rows.append((name, float(row["Counter_Value"])))
    return rows


def traffic_summary(fetch_dir, write_dir, copy_mib, precision, command):
    """Measured HBM bytes per launch from two rocprofv3 --pmc runs of this
    script (FETCH_SIZE, then WRITE_SIZE, both in KiB per dispatch), each
    scaled by what the counter reads for the device-to-device copy of
    --copy-mib MiB that the profiled runs include (the runtime's plain copy
    kernel: it reads and writes exactly that many bytes). Merges one record
    per library kernel into profiles/traffic_kernels.json under the key
    <kernel>_<length>_<precision> (R2C instances as <kernel>_r2c; precision
    and length from the kernel's template arguments); keys of earlier runs
    that this run does not measure are kept."""
    import statistics
    expect_kib = copy_mib * 1024.0
    per = {}
    scale = {}
    for counter, folder in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = _counter_rows(folder)
        copies = [v for n, v in rows if "copyBuffer" in n]
        if not copies:
            raise SystemExit(f"no copy kernel dispatches under {folder}")
        big = [v for v in copies if v >= 0.5 * max(copies)]
        scale[counter] = expect_kib / statistics.median(big)
        for n, v in rows:
            if not n.startswith("k_"):
                continue
            base = n.split("<")[0]
            if ", true>" in n.split("(")[0]:  # R2C variants are profiled as <name>_r2c
                base = base.replace("_fast", "") + "_r2c"
            targs = n.split("(")[0].split("<", 1)[1].split(">")[0].split(",") if "<" in n else []
            prec = targs[0].strip() if targs else precision
            length = targs[1].strip() if len(targs) > 1 else "any"
            key = f"{base}_{length}_{prec}"
            per.setdefault(key, {}).setdefault(counter, []).append(v)
    path = os.path.join(ROOT, "profiles", "traffic_kernels.json")
    doc = {"runs": {}, "kernels": {}}
    if os.path.exists(path):
        with open(path) as fh:
            doc = json.load(fh)
    run = f"{fetch_dir} + {write_dir}"
    doc["runs"][run] = {"command": command, "copy_mib": copy_mib,
                        "fetch_scale": scale["FETCH_SIZE"], "write_scale": scale["WRITE_SIZE"]}
    out = doc["kernels"]
    for k, d in per.items():
        if "FETCH_SIZE" not in d or "WRITE_SIZE" not in d:
            continue
        f = statistics.mean(d["FETCH_SIZE"])
        w = statistics.mean(d["WRITE_SIZE"])
        out[k] = {
            "fetch_kib_per_dispatch": round(f, 1), "write_kib_per_dispatch": round(w, 1),
            "bytes_per_launch": round(1024.0 * (f * scale["FETCH_SIZE"] + w * scale["WRITE_SIZE"])),
            "dispatches": len(d["FETCH_SIZE"]), "run": run,
        }
        log(f"{k:40s} fetch {f:10.1f} KiB x{scale['FETCH_SIZE']:.2f}  write {w:10.1f} KiB "
            f"x{scale['WRITE_SIZE']:.2f}  -> {out[k]['bytes_per_launch'] / 1e6:8.2f} MB/launch")
    doc["kernels"] = dict(sorted(out.items()))
    with open(path, "w") as fh:
        json.dump(doc, fh, indent=1)
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
    ap.add_argument("--batch", type=int, default=1,
                    help="run the batched section (0: skip)")
    ap.add_argument("--batch-cases", nargs="*", default=[],
                    help="substrings selecting batched cases (default: all)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the measurements, 10 device-to-device copies of this many MiB "
                         "(the calibration dispatches for --pmc FETCH_SIZE / WRITE_SIZE runs)")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarise two rocprofv3 --pmc outputs (FETCH_SIZE, WRITE_SIZE) of "
                         "runs with --copy-mib into profiles/traffic_kernels.json and exit")
    ap.add_argument("--traffic-command", default="",
                    help="the profiling command recorded in the traffic summaries")
    args = ap.parse_args()
    if args.traffic:
        traffic_summary(args.traffic[0], args.traffic[1], args.copy_mib or 256, args.precision,
                        args.traffic_command)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the batched "
              "path shards by bands/k-points with no communication and is not run on "
              "several GPUs here.", file=sys.stderr)
        sys.exit(2)

    os.environ.setdefault("SPFFT_AMD_TUNE_LOG", "1")  # batched kernel choices go to the log
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

    # The API's default mode (SPFFT_EXEC_SYNCHRONOUS: every call waits for
    # its work) on the same rotation, and a warm single transform (no
    # rotation: its ~77 MB of buffers stay in the 256 MB Infinity Cache).
    extra = {}
    for k in range(args.rotate):
        trs[k].set_execution_mode(sp.SPFFT_EXEC_SYNCHRONOUS)
    nsync = max(10, args.steps // 4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(nsync):
        pair(i)
    torch.cuda.synchronize()
    extra["plain_synchronous"] = {"pairs_per_s": nsync / (time.perf_counter() - t0),
                                  "pairs": nsync, "rotate": args.rotate}
    for k in range(args.rotate):
        trs[k].set_execution_mode(sp.SPFFT_EXEC_ASYNCHRONOUS)
    for i in range(args.warmup):
        pair(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        pair(0)
    torch.cuda.synchronize()
    extra["plain_warm_cache"] = {"pairs_per_s": args.steps / (time.perf_counter() - t0),
                                 "pairs": args.steps, "rotate": 1}
    log(f"synchronous mode: {extra['plain_synchronous']['pairs_per_s']:.1f} pairs/s; warm "
        f"single transform: {extra['plain_warm_cache']['pairs_per_s']:.1f} pairs/s")

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
    tfile = os.path.join(ROOT, "profiles", "traffic_kernels.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f)["kernels"].get(f"{dominant}_128_{args.precision}", {}).get(
                "bytes_per_launch")

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
        **extra,
    }
    result["launch_floor"] = launch_floor(torch, d["mean_us"])
    result["plain_r2c"] = plain_r2c_bench(args, sp, torch, np)
    result["plain_other_lengths"] = plain_lengths_bench(args, sp, torch, np)
    if args.batch > 0:
        b = batched_bench(args, sp, torch, np)
        main_case = b.get("128^3 r=40 x64")
        if main_case:
            b["fp32_vs_plain_fp64"] = main_case["float"]["band_pairs_per_s"] / value
            b["fp64_vs_plain_fp64"] = main_case["double"]["band_pairs_per_s"] / value
        result["batched"] = b
    if not args.no_cpu_baseline:
        v0 = complex_values(0x5EED0002, ns)
        result["cpu_baseline"] = cpu_baseline(t, dims, v0, args.cpu_seconds)
        log(f"cpu baseline: {result['cpu_baseline']['value']:.2f} pairs/s on "
            f"{result['cpu_baseline']['cores']} threads")
    print(json.dumps(result), flush=True)
    if args.copy_mib > 0:
        n = args.copy_mib << 20
        src = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
