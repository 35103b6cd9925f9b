"""Time the dense solve alone on a random complex matrix of the benchmark's size.

    python -m multifebe_amd.lu_timing [n] [split16,mixed,fp64] [repeats]

A = randn + i randn + n I (well conditioned), one right-hand side; prints one line per solve with
the wall time, the LU rate (8/3 n^3 real flops) and the solver's report.
"""
from __future__ import annotations

import sys
import time


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    n = int(argv[0]) if len(argv) > 0 else 10374
    modes = argv[1].split(",") if len(argv) > 1 else ["mixed", "fp64"]
    reps = int(argv[2]) if len(argv) > 2 else 2
    import torch
    from . import bem
    g = torch.Generator(device="cuda").manual_seed(0)
    A0 = (torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
          + 1j * torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
    A0 += n * torch.eye(n, dtype=torch.complex128, device="cuda")
    b0 = torch.randn(n, dtype=torch.complex128, device="cuda", generator=g)
    for mode in modes:
        pm = {"mixed": bem.PREC_FP32_REFINE, "split16": bem.PREC_SPLIT16_REFINE,
              "fp64": bem.PREC_FP64}[mode]
        for rep in range(reps):
            A = A0.clone()
            b = b0.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x, r = bem.gesv_device(A, b, pm)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            print(f"{mode} n={n} rep {rep}: {t * 1e3:.1f} ms, {8 / 3 * n ** 3 / t / 1e12:.2f} TFLOP/s "
                  f"(LU-equivalent), {r}", flush=True)


if __name__ == "__main__":
    main()
