#!/usr/bin/env python3
"""Throughput of 2D block-equivalence labeling on one MI355X.

Workload: 8192x8192 uint8 binary images, the six inputs of the single-image
configuration (i.i.d. Bernoulli p in {0.1, 0.3, 0.5, 0.7, 0.9} and Gaussian-
smoothed noise blobs, sigma 4 px, 50% quantile), 8-connectivity, uint32
labels.  One step = one full labeling of one image (inputs cycle), inputs and
outputs resident in HBM, no host synchronisation inside the timed region.

Prints one JSON line: value = Gpixel/s over the timed steps; `roofline` for the
labeling pipeline (algorithmic 5 B per pixel: 1 B in + 4 B label out, against
8 TB/s), measured with HIP events on the stream the kernels run on; and
`cpu_baseline`: the oracle's OpenMP restatement of the reference CPU labeler
(TLabelEquivalenceX2) timed on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_GBS = 8000.0
ALGO_BYTES_PER_PX = 5.0
METRIC = "Gpixel/s 2D 8192², Gvoxel/s 3D 512³, batched vs per-image; % HBM roofline"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_inputs(side):
    import labeling_amd as L

    out = []
    for p in (0.1, 0.3, 0.5, 0.7, 0.9):
        out.append((f"bern_p{p}", L.bernoulli((side, side), p, seed=0)))
    out.append(("blob_s4", L.blobs2d(side, side, 4.0, 0.5, seed=0)))
    return out


def traffic_from_profiles():
    path = os.path.join(ROOT, "profiles", "bleq2d_8192_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get("hbm_bytes_per_image")
    except (OSError, ValueError):
        return None


def cpu_baseline(inputs, seconds_budget):
    from oracle import pyoracle as O

    threads = min(16, os.cpu_count() or 1)
    total_px, total_s, runs = 0, 0.0, 0
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < seconds_budget or runs < len(inputs):
        name, img = inputs[runs % len(inputs)]
        _, info = O.bleq2d(img, threads=threads, with_info=True)
        total_px += img.size
        total_s += info["seconds"]
        runs += 1
        if runs >= 6 * len(inputs):
            break
    return {
        "value": round(total_px / total_s / 1e9, 4),
        "unit": "Gpixel/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{runs} labelings of the 8192x8192 inputs (cycled), oracle OpenMP restatement of "
                  f"TLabelEquivalenceX2 on the padded image, DoLabel region only",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--side", type=int, default=8192)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-times", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported (a single image or batch is "
              f"labelled on one device; see DESIGN.md for how the path would shard)", file=sys.stderr)
        sys.exit(2)

    import torch

    import labeling_amd as L

    side = a.side
    log(f"[bench] generating inputs {side}x{side}")
    inputs = make_inputs(side)
    dev = torch.device("cuda:0")
    d_inputs = [torch.from_numpy(img).to(dev) for _, img in inputs]
    d_labels = torch.empty((side, side), dtype=torch.int32, device=dev)
    d_count = torch.zeros(1, dtype=torch.int32, device=dev)
    lab = L.Labeler(0)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    def step(i):
        x = d_inputs[i % len(d_inputs)]
        lab.label2d_dev(x.data_ptr(), side, side, side, d_labels.data_ptr(), "bleq", 0, d_count.data_ptr(), sptr)

    log(f"[bench] warmup {a.warmup}")
    for i in range(a.warmup):
        step(i)
    torch.cuda.synchronize()

    log(f"[bench] timed {a.steps} steps")
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for i in range(a.steps):
        step(i)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = ev0.elapsed_time(ev1)
    ms_per_step = ms_total / a.steps
    px = side * side
    value = px * a.steps / (ms_total / 1e3) / 1e9

    # per-kernel device times (separate pass; events between the kernels)
    kernels = {}
    per_input = {}
    if not a.no_kernel_times:
        lab.set_profiling(True)
        reps = max(1, min(a.steps, 30))
        for i in range(reps):
            step(i)
            kt = lab.kernel_times()
            for name, us in kt:
                kernels.setdefault(name, []).append(us)
            per_input.setdefault(inputs[i % len(inputs)][0], []).append(sum(us for _, us in kt))
        lab.set_profiling(False)
    kernel_avg = {k: sum(v) / len(v) for k, v in kernels.items()}
    path_us = sum(kernel_avg.values()) if kernel_avg else ms_per_step * 1e3
    achieved = ALGO_BYTES_PER_PX * px / (path_us * 1e-6) / 1e9
    traffic = traffic_from_profiles()

    cpu = None
    if not a.no_cpu_baseline:
        log("[bench] cpu baseline")
        cpu = cpu_baseline(inputs, a.cpu_seconds)

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpixel/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": f"bleq 2D single image {side}x{side} uint8, 8-connectivity, uint32 labels; inputs cycle "
                        "Bernoulli p=0.1,0.3,0.5,0.7,0.9 and blobs sigma=4 q=0.5",
            "global_batch": 1,
            "pixels_per_step": px,
            "wall_s": round(wall, 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_GBS, 4),
            "traffic": traffic,
            "scope": "labeling pipeline (all kernels of one image), algorithmic 5 B/px",
            "path_us": round(path_us, 2),
            "kernels_us": {k: round(v, 2) for k, v in kernel_avg.items()},
            "per_input_us": {k: round(sum(v) / len(v), 2) for k, v in per_input.items()},
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    lab.close()


if __name__ == "__main__":
    main()
