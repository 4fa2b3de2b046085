"""Train the second-order surrogate on fields the exact kernels generate and
write its weights (python -m cfdwavemaker_amd.train_surrogate --help).

Training states are drawn from the batched-workload distribution (N 128-2048,
Hs 1-8 m, Tp 6-16 s, depth 20-200 m, long- and short-crested); held-out
states measure the error with and without the network (the method without it
is linear theory: second-order correction zero)."""
import argparse
import json
import sys
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-states", type=int, default=64)
    ap.add_argument("--test-states", type=int, default=8)
    ap.add_argument("--ncol", type=int, default=48)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--width", type=int, default=256)
    ap.add_argument("--nhidden", type=int, default=3)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    from cfdwavemaker_amd import surrogate as S

    def log(m):
        print(m, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    Xtr, Ytr = S.build_dataset(range(a.train_states), ncol=a.ncol, log=log)
    Xte, Yte = S.build_dataset(range(10000, 10000 + a.test_states), ncol=a.ncol, log=log)
    log(f"dataset {Xtr.shape} / {Xte.shape} in {time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    m, curve = S.train(Xtr, Ytr, steps=a.steps, batch=a.batch, lr=a.lr, width=a.width, nhidden=a.nhidden)
    log(f"trained {a.steps} steps in {time.perf_counter() - t1:.1f} s; loss {curve[curve > 0]}")
    import torch
    P = m.forward(torch.tensor(Xte, device="cuda")).cpu().numpy()
    err = P - Yte                                   # dimensionless (per-sea scales)
    # error of the method without the network (linear theory) = the exact correction itself
    def rel(e, ref):
        return float(np.sqrt(np.mean(e ** 2)) / np.sqrt(np.mean(ref ** 2)))
    rep = {
        "train_rows": int(Xtr.shape[0]), "test_rows": int(Xte.shape[0]),
        "eta2_rel_rms_err_net": rel(err[:, 0], Yte[:, 0]),
        "u2_rel_rms_err_net": rel(err[:, 1::4], Yte[:, 1::4]),
        "w2_rel_rms_err_net": rel(err[:, 3::4], Yte[:, 3::4]),
        "p2_rel_rms_err_net": rel(err[:, 4::4], Yte[:, 4::4]),
        "eta2_max_err_net_over_max_exact": float(np.abs(err[:, 0]).max() / np.abs(Yte[:, 0]).max()),
        "final_loss": float(curve[curve > 0][-1]),
    }
    print(json.dumps(rep), flush=True)
    if a.out:
        m.save(a.out, report=json.dumps(rep))


if __name__ == "__main__":
    main()
