"""Writes c5_fd.json: FP64 central differences (step 1e-4) of the trajectory-gradient loss
defined in c5_problem.py, evaluated with the CPU oracle: along one image direction with the
workload's 10 CG iterations, and along named trajectory directions with ITERS_K iterations,
together with the fp32 oracle's losses and its autograd projections on the same trajectory
directions (the rounding scale of the reference's own precision). Run from the repository root:

    python tests/golden/make_c5_fd.py

(about 6 minutes on 8 cores).
"""

import json
import os
import sys
import time

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

import c5_problem as P  # noqa: E402


def main():
    torch.set_num_threads(os.cpu_count() or 1)
    t0 = time.time()
    inp = P.build_inputs()
    rec = {"h": P.H, "y_abs_sum": float(inp["y"].abs().to(torch.float64).sum()),
           "loss": P.loss64(inp), "loss_fp32_oracle": P.oracle32(inp, P.ITERS, grad=False)[0], "k": {}, "x": {}}
    print(f"base loss {rec['loss']:.12e} ({time.time() - t0:.0f}s)", flush=True)
    kin = inp["kin"].to(torch.float64)
    rec["k_iters"] = P.ITERS_K
    rec["loss_k"] = P.loss64(inp, iters=P.ITERS_K)
    rec["loss_k_fp32_oracle"], g32 = P.oracle32(inp, P.ITERS_K)
    print(f"loss with {P.ITERS_K} iterations {rec['loss_k']:.12e}; fp32 oracle gradient ({time.time() - t0:.0f}s)",
          flush=True)
    for name, v in P.k_directions(inp["kin"]):
        lp = P.loss64(inp, k=kin + P.H * v, iters=P.ITERS_K)
        lm = P.loss64(inp, k=kin - P.H * v, iters=P.ITERS_K)
        rec["k"][name] = {"lp": lp, "lm": lm, "fd": (lp - lm) / (2 * P.H), "moved": int((v != 0).sum()),
                          "fp32_oracle": float(torch.sum(g32 * v))}
        print(f"k {name}: fd {rec['k'][name]['fd']:.10e}, fp32 oracle {rec['k'][name]['fp32_oracle']:.10e} "
              f"({time.time() - t0:.0f}s)", flush=True)
    v = P.x_direction()
    xin = inp["xin"].to(torch.complex128)
    lp = P.loss64(inp, x=xin + P.H * v)
    lm = P.loss64(inp, x=xin - P.H * v)
    rec["x"]["random"] = {"lp": lp, "lm": lm, "fd": (lp - lm) / (2 * P.H)}
    print(f"x random: fd {rec['x']['random']['fd']:.10e} ({time.time() - t0:.0f}s)", flush=True)
    rec["torch"] = torch.__version__
    with open(os.path.join(HERE, "c5_fd.json"), "w") as f:
        json.dump(rec, f, indent=1)


if __name__ == "__main__":
    main()
