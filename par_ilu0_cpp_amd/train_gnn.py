#!/usr/bin/env python3
"""Train the learned-compensation network on solver-generated labels.

Training set: 3D 7-point diffusion matrices whose cell coefficient is
exp(sigma * g), g a unit-variance Gaussian random field (correlation
length in cells), with uniform[-1, 1] right-hand sides.  For every training
matrix the solver runs ILU(lof)-GMRES(m) with constant omega over
learned.OMEGA_GRID; the omega with the fewest steps is the label of all its
rows.  The network is then fitted by Adam on the GPU with the library's
own forward/backward kernels, and the held-out matrices report steps
with and without the learned compensation.

    python -m par_ilu0_cpp_amd.train_gnn --train 24 --side 24 --epochs 30 \
        --eval-side 48 --out par_ilu0_cpp_amd/weights/gnn_compensation.json
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from par_ilu0_cpp_amd import generators as gen  # noqa: E402
from par_ilu0_cpp_amd.learned import WEIGHTS, DeviceSystem, Gnn, best_omega, transpose_map  # noqa: E402


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_system(side, seed, lof, m, sigma, corr):
    a = gen.grf_diffusion3d(side, sigma=sigma, corr_len=corr, seed=seed)
    b = gen.rhs_uniform(a.n, seed=10_000 + seed)
    return DeviceSystem(a, b, lof=lof, m=m)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=24, help="number of training matrices")
    ap.add_argument("--side", type=int, default=24, help="grid side of the training matrices")
    ap.add_argument("--eval", type=int, default=4, help="held-out matrices")
    ap.add_argument("--eval-side", type=int, default=48)
    ap.add_argument("--lof", type=int, default=1)
    ap.add_argument("--m", type=int, default=50)
    ap.add_argument("--sigma", type=float, default=1.5)
    ap.add_argument("--corr", type=float, default=8.0)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--out", type=Path, default=WEIGHTS)
    args = ap.parse_args(argv)
    import torch

    t0 = time.time()
    data = []
    for k in range(args.train):
        s = make_system(args.side, k, args.lof, args.m, args.sigma, args.corr)
        w, table = best_omega(s)
        a = s.csr
        rpt, cit, tmap = transpose_map(a.row_ptr, a.col_idx)
        tt = [torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in (rpt, cit, tmap)]
        target = torch.full((a.n,), float(w), dtype=torch.float32, device="cuda")
        data.append((s, tt, target, w, table))
        log(f"[train] matrix {k}: label omega={w} steps={table}")
    log(f"[train] labels in {time.time() - t0:.1f}s")
    net = Gnn()
    hist = []
    for ep in range(args.epochs):
        losses = []
        for s, tt, target, _, _ in data:
            losses.append(net.train_step(s.n, s.ptr("rp"), s.ptr("ci"), s.ptr("va"), tt[0].data_ptr(),
                                         tt[1].data_ptr(), tt[2].data_ptr(), target.data_ptr(), args.lr))
        hist.append(float(np.mean(losses)))
        log(f"[train] epoch {ep}: loss {hist[-1]:.5f}")
    evals = []
    for k in range(args.eval):
        s = make_system(args.eval_side, 50_000 + k, args.lof, args.m, args.sigma, args.corr)
        ok0, its0, res0 = s.solve_with_omega(None)
        okl, itsl, resl, fb = s.solve_learned(net, budget=its0)
        evals.append({"plain_its": its0, "learned_its": itsl, "fallback": fb, "plain_res": res0,
                      "learned_res": resl})
        log(f"[eval] {k}: plain {its0} learned {itsl} fallback={fb} res={resl:.2e}")
    meta = {
        "train": args.train, "side": args.side, "lof": args.lof, "m": args.m, "sigma": args.sigma,
        "corr": args.corr, "epochs": args.epochs, "lr": args.lr, "loss_history": hist,
        "labels": [d[3] for d in data], "eval_side": args.eval_side, "eval": evals,
    }
    net.save(args.out, meta)
    print(json.dumps({"loss_first": hist[0] if hist else None, "loss_last": hist[-1] if hist else None,
                      "eval": evals}))


if __name__ == "__main__":
    main()
