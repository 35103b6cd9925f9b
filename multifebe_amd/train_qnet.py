"""Train the quadrature-rule network on the GPU and save its weights.

    python -m multifebe_amd.train_qnet --tau 1e-10 --out multifebe_amd/data/qnet_tau1e-10.npz

Training pairs come from closed quad9 surfaces of the families the solver meets (projected-cube
spheres, cubes, cylinders with h = 2r, and a thin 1 x 1 x 0.01 slab whose opposite faces are
nearly coincident), static and harmonic (a0 up to 4), labelled by the reference's adaptive rule
at tau_ref.  The saved file holds the FP32 parameters and the training report.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np


def slab(n_xy=20, n_z=2, size=1.0, thickness=0.01):
    """closed quad9 box size x size x thickness (n_xy intervals in x and y, n_z through the
    thickness); the elements of the two large faces are 2 x (size / n_xy) wide"""
    from . import mesh as M
    m = M.cube(n_xy)
    # squash the unit cube to the slab (z scaled), keep the top/bottom meshes
    x = m.x.copy()
    x[:, 0] *= size
    x[:, 1] *= size
    x[:, 2] *= thickness
    return M.Mesh(x=x, conn=m.conn.copy())


def default_samples():
    from . import learned as L, mesh as M
    return [
        L.Sample(M.sphere(8), harmonic_a0=(0.5, 2.0, 4.0), far_pairs=3000),
        L.Sample(M.sphere(12), harmonic_a0=(1.0, 3.0), near_fraction=0.5, far_pairs=3000),
        L.Sample(M.cube(8), harmonic_a0=(1.0, 3.5), far_pairs=3000),
        L.Sample(M.cylinder(6), harmonic_a0=(0.5, 2.5), far_pairs=3000),
        L.Sample(slab(12), harmonic_a0=(1.0, 3.0), near_fraction=0.3, far_pairs=3000),
    ]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--tau", type=float, default=1e-10)
    ap.add_argument("--tau-ref", type=float, default=1e-12)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    from . import learned as L
    t0 = time.perf_counter()
    ts = L.build_training_set(default_samples(), args.tau, args.tau_ref, seed=args.seed)
    t1 = time.perf_counter()
    print(f"training set: {ts.X.shape[0]} pairs in {t1 - t0:.1f} s; classes "
          f"{L.class_histogram(ts.y).tolist()}", flush=True)
    net, rep = L.train(ts, epochs=args.epochs, batch=2048, lr=2e-3, seed=args.seed)
    t2 = time.perf_counter()
    rep.update({"tau": args.tau, "tau_ref": args.tau_ref, "epochs": args.epochs,
                "seed": args.seed, "label_s": t1 - t0, "train_s": t2 - t1,
                "sources": [list(map(float, s)) if s[2] is not None else [s[0], s[1], 0.0, s[3]]
                            for s in ts.source]})
    print(json.dumps(rep), flush=True)
    np.savez(args.out, params=net.params, report=json.dumps(rep))


if __name__ == "__main__":
    main()
