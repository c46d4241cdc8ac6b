"""Benchmark: 3D radial kooshball SENSE NUFFT forward+adjoint on one MI355X.

Workload (the headline configuration): 128^3 complex64 image, 32 smooth coil
maps, 20,000 Fibonacci-sphere spokes x 256 readout points (5.12 M samples per
coil), Kaiser-Bessel width 4, oversampling 1.25 (grid 160^3). One step = one
SENSE forward (32 coils, image -> k-space) plus one adjoint (k-space -> image),
inputs resident in HBM. Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

``roofline`` is reported for the dominant kernel of the step, timed live with
HIP events around its launches inside the timed loop; its algorithmic bytes
per launch are stated in ``roofline_detail``. ``cpu_baseline`` times the
reference CPU path (the oracle's restatement of ungrid_torch/grid_torch + torch
FFT) on a bounded sample of the same workload and scales it to the full step.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, C, NSPOKES, NREAD, WIDTH, OS = 128, 32, 20000, 256, 4, 1.25
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(dev):
    from mr_recon_amd import synth

    t0 = time.time()
    trj = synth.kooshball_3d(NSPOKES, NREAD, N).to(dev)
    mps = synth.coil_maps(C, (N, N, N), seed=0).to(dev)
    img = synth.random_image((N, N, N), seed=1).to(dev)
    log(f"inputs ready in {time.time() - t0:.1f}s")
    return trj, mps, img


def kernel_bytes(M):
    """Algorithmic HBM bytes per launch of the sample-side kernels (one launch covers all
    C coils): interp reads the C grids and the coordinates and writes the C sample
    vectors; gridding (partial) reads the samples and coordinates and produces the C
    grids' worth of output; the reduce reads nothing algorithmic beyond one grid write."""
    G = math.ceil(OS * N)
    return {"interp": 8 * C * G ** 3 + 4 * 3 * M + 8 * C * M,
            "grid_partial": 8 * C * M + 4 * 3 * M + 8 * C * G ** 3,
            "grid_reduce": 8 * C * G ** 3}


def committed_traffic(kernel):
    """HBM bytes per launch of ``kernel`` measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes (scaled by a 1 GiB copy) and recorded in profiles/ by the profiling recipe;
    None when no record exists for this kernel."""
    import re

    path = os.path.join(ROOT, "profiles", f"{kernel}_3d128_c32_traffic.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return None if m is None else {"bytes": float(m.group(1)), "source": os.path.relpath(path, ROOT)}


def step_bytes(M):
    """Algorithmic HBM bytes of one SENSE forward+adjoint (x, maps, 7 grid passes, samples, coords)."""
    G = math.ceil(OS * N)
    return 16 * N ** 3 + 24 * C * N ** 3 + 56 * C * G ** 3 + 16 * C * M + 8 * 3 * M


def cpu_baseline(frac_spokes=1 / 4):
    """Reference CPU path on 1 coil and a spoke subset; scaled to 32 coils and all spokes."""
    from mr_recon_amd import synth
    from oracle.nufft_ref import TritonNufftRef

    nsp = int(NSPOKES * frac_spokes)
    trj = synth.kooshball_3d(nsp, NREAD, N).reshape(1, -1, 3)
    img = synth.random_image((1, 1, N, N, N), seed=1)
    ref = TritonNufftRef((N, N, N), oversamp=OS, width=WIDTH)
    # FFT-side work (per coil, independent of the sample count)
    t0 = time.perf_counter()
    g = ref.forward_FT_only(img)
    x = ref.adjoint_iFT_only(g)
    t_fft = time.perf_counter() - t0
    # sample-side work (interp + grid on nsp spokes, 2^18-sample chunks to bound memory)
    from oracle import interp_ref

    t_scaled = ref._scale_trj(trj, (1, 1, N, N, N))
    t0 = time.perf_counter()
    y = interp_ref.ungrid(g[0], t_scaled[0], (WIDTH,) * 3, beta=ref.beta, batch_size=2 ** 18)
    gg = interp_ref.grid(y, t_scaled[0], g.shape[-3:], (WIDTH,) * 3, beta=ref.beta, batch_size=2 ** 18)
    t_samp = time.perf_counter() - t0
    del x, gg
    per_step = C * (t_fft + t_samp / frac_spokes)
    return {
        "value": 1.0 / per_step,
        "unit": "NUFFT fwd+adj/s",
        "cores": torch.get_num_threads(),
        "kind": "port",
        "sample": (f"1 of {C} coils; FFT/pad/crop side on the full 160^3 grid ({t_fft:.2f}s) + interp/grid on "
                   f"{nsp} of {NSPOKES} spokes ({t_samp:.2f}s, 2^18-sample chunks); scaled linearly to {C} coils "
                   f"and all spokes = {per_step:.1f}s per fwd+adj"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed steps, run one elementwise kernel that reads 1 GiB and writes 1 GiB "
                         "(calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build runs the NUFFT on a single MI355X "
              "(multi-GPU sharding is a design note, not implemented). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from mr_recon_amd.fourier import hip_nufft
    from mr_recon_amd.linops import sense_linop

    dev = torch.device("cuda:0")
    trj, mps, img = make_inputs(dev)
    t0 = time.time()
    A = sense_linop(trj, mps, nufft=hip_nufft((N, N, N), oversamp=OS, width=WIDTH))
    torch.cuda.synchronize()
    log(f"plan built in {time.time() - t0:.2f}s; M={A.plan.npts} per coil, items={A.plan.max_items}")
    M = A.plan.npts

    for _ in range(args.warmup):
        A.adjoint(A.forward(img))
    torch.cuda.synchronize()

    A.timers = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        A.adjoint(A.forward(img))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    stages = {k: sum(s.elapsed_time(e) for s, e in v) / len(v) for k, v in A.timers.items()}
    A.timers = None
    ms = dt / args.steps * 1e3
    log("stage averages (ms): " + ", ".join(f"{k}={v:.3f}" for k, v in stages.items()))

    kb = kernel_bytes(M)
    dom = max(kb, key=lambda k: stages.get(k, 0.0))
    dom_ms = stages[dom]
    ib = kb[dom]
    achieved = ib / (dom_ms * 1e-3) / 1e9
    rec = {
        "metric": "NUFFT fwd+adj/s (3D 128^3 x 32 coils)",
        "value": args.steps / dt,
        "unit": "NUFFT fwd+adj/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded kooshball trajectory, Gaussian-bump coil maps, random image)",
        "config": {"workload": "3D kooshball SENSE NUFFT fwd+adj", "image": [N, N, N], "coils": C,
                   "spokes": NSPOKES, "readout": NREAD, "samples_per_coil": M, "width": WIDTH, "oversamp": OS,
                   "grid": [math.ceil(OS * N)] * 3},
        "step_GBps_algorithmic": step_bytes(M) / (ms * 1e-3) / 1e9,
        "stage_ms": {k: round(v, 4) for k, v in stages.items()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "roofline_detail": {"kernel": dom, "avg_ms": dom_ms, "algorithmic_bytes_per_launch": ib,
                            "bytes_formula": {"interp": "8*C*G^3 + 4*3*M + 8*C*M",
                                              "grid_partial": "8*C*M + 4*3*M + 8*C*G^3",
                                              "grid_reduce": "8*C*G^3"}[dom]},
    }
    traffic = committed_traffic(dom)
    if traffic is not None:
        rec["roofline"]["traffic"] = traffic["bytes"]
        rec["roofline_detail"]["traffic_source"] = traffic["source"]
    if args.copy_probe:
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 1.0, out=dst)  # a copy through the shader cores (a D2D memcpy may use DMA)
        torch.cuda.synchronize()
        del src, dst
    if not args.no_cpu_baseline:
        t0 = time.time()
        rec["cpu_baseline"] = cpu_baseline()
        log(f"cpu baseline took {time.time() - t0:.1f}s")
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
