"""Training data and training of the quadrature-rule network (learned rule choice).

Labels come from the reference's own adaptive quasi-singular integration
(fbem_bem_staela3d_sbie_ext_adp / fbem_bem_harela3d_sbie_ext_adp, lib/fbem/src/bem_staela3d.f90:
973-1096, bem_harela3d.f90:1050-1172) run by the library on a labelling mesh created with a tight
qsi_relative_error: for every sampled (collocation node, element) pair, each precalculated
Gauss-Legendre rule (ext_pre, bem_staela3d.f90:592-650) is compared with it, and the label is the
cheapest rule whose blocks are within the target tolerance (the adaptive class when none is).
Everything runs on the GPU; pairs are sampled with a fixed seed.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import bem
from . import mesh as M
from .qnet import N_CLASSES, QNet


@dataclass
class Sample:
    mesh: M.Mesh
    harmonic_a0: tuple = ()          # a0 = omega R / c_s values (harmonic); () -> static only
    near_fraction: float = 1.0       # share of the near pairs (node within 2 element sizes)
    far_pairs: int = 2000            # far pairs sampled per mesh
    length: float = 1.0              # R in a0 = omega R / c_s


@dataclass
class TrainingSet:
    X: object                        # torch float32 (n, 5), device
    y: object                        # torch int32 (n,), device
    source: list = field(default_factory=list)


MATERIAL = dict(E=1e9, nu=0.3, xi=0.05, rho=2000.0)


def _pairs(m: M.Mesh, s: Sample, rng) -> np.ndarray:
    """pair ids e * nn + i of non-singular pairs: a share of the near ones and some far ones"""
    conn = m.conn
    # element centre and size from the corner nodes (sampling only)
    cen = m.x[conn[:, 8]]
    size = np.max(np.linalg.norm(m.x[conn[:, [0, 1, 2, 3]]] - m.x[conn[:, [2, 3, 0, 1]]], axis=2), axis=1)
    near, far = [], []
    for e in range(m.ne):
        d = np.linalg.norm(m.x - cen[e], axis=1)
        own = np.zeros(m.nn, bool)
        own[conn[e]] = True
        ids = e * m.nn + np.arange(m.nn)
        isn = (d < 2.0 * size[e]) & ~own
        near.append(ids[isn])
        far.append(ids[~isn & ~own])
    near = np.concatenate(near)
    far = np.concatenate(far)
    near = rng.permutation(near)[: int(round(s.near_fraction * len(near)))]
    far = rng.permutation(far)[: s.far_pairs]
    return np.concatenate([near, far]).astype(np.int32)


def build_training_set(samples, tau: float, tau_ref: float = 1e-12, seed: int = 0) -> TrainingSet:
    import torch
    rng = np.random.default_rng(seed)
    mu_r, mu_c, lam_c = M.elastic_constants(MATERIAL["E"], MATERIAL["nu"], MATERIAL["xi"])
    cs = np.sqrt(mu_r / MATERIAL["rho"])
    Xs, ys, src = [], [], []
    for s in samples:
        m = s.mesh
        if m.ctype is None:
            m = M.static_cube_bcs(m)
        sysm = bem.BemSystem(m, bem.Settings(qsi_relative_error=tau_ref))
        pairs = _pairs(m, s, rng)
        cases = [None] + [a0 * cs / s.length for a0 in s.harmonic_a0]
        for omega in cases:
            if omega is None:
                lab, _ = sysm.pair_labels(pairs, tau, mu=mu_r, nu=MATERIAL["nu"])
                kw = 0.0
            else:
                lab, _ = sysm.pair_labels(pairs, tau, omega=omega, lam=lam_c, mu_c=mu_c,
                                          rho=MATERIAL["rho"])
                kw = omega / cs
            f = sysm.pair_features(pairs, kw)
            keep = lab >= 0
            Xs.append(f[keep])
            ys.append(lab[keep])
            src.append((m.nn, m.ne, omega, int(keep.sum().item())))
        sysm.close()
    return TrainingSet(torch.cat(Xs), torch.cat(ys), src)


def class_histogram(y) -> np.ndarray:
    y = y.cpu().numpy() if hasattr(y, "cpu") else np.asarray(y)
    return np.bincount(y, minlength=N_CLASSES)


def train(ts: TrainingSet, epochs=40, batch=2048, lr=2e-3, seed=0, holdout=0.1):
    """Train on (1 - holdout) of the rows; returns (net, report) with the held-out accuracy and the
    share of held-out rows where the network's rule at the given confidence is cheaper than the
    label (an under-resolved choice)."""
    import torch
    n = ts.X.shape[0]
    g = torch.Generator(device="cpu").manual_seed(seed)
    perm = torch.randperm(n, generator=g).to(ts.X.device)
    nh = int(holdout * n)
    hold, tr = perm[:nh], perm[nh:]
    net = QNet(seed=seed)
    losses = net.train(ts.X[tr], ts.y[tr], epochs=epochs, batch=batch, lr=lr, seed=seed)
    report = {"rows": n, "train_rows": n - nh, "holdout_rows": nh, "loss_first": float(losses[0]),
              "loss_last": float(losses[-1]), "classes": class_histogram(ts.y).tolist()}
    if nh:
        probs, _ = net.infer(ts.X[hold], 0.5)
        y = ts.y[hold]
        report["holdout_accuracy"] = float((probs.argmax(dim=1).to(torch.int32) == y).float().mean())
        for conf in (0.9, 0.99, 0.999):
            _, c = net.infer(ts.X[hold], conf)
            report[f"under_resolved@{conf}"] = float((c < y).float().mean())
            report[f"mean_class@{conf}"] = float(c.float().mean())
        report["mean_label"] = float(y.float().mean())
    return net, report
