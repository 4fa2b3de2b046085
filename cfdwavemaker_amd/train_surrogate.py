"""Train the second-order surrogate on fields the exact kernels generate and
write its weights (python -m cfdwavemaker_amd.train_surrogate --help).

Training states are drawn from the batched-workload distribution (N 128-2048,
Hs 1-8 m, Tp 6-16 s, depth 20-200 m, long- and short-crested); held-out
states measure the error with and without the network (the method without it
is linear theory: second-order correction zero).  --budget trains one network
per feature set (local linear features; + the quadrature of eta1; + the exact
eta2) and reports, per set, the held-out errors and the elevation error of a
learned eta2 on the N=1024 short-crested sea's 512 x 512 columns against the
1e-3 Hs gate: the error budget that decides which parts are learned."""
import argparse
import json
import sys
import time

import numpy as np


def rel(e, ref):
    return float(np.sqrt(np.mean(e ** 2)) / np.sqrt(np.mean(ref ** 2)))


def heldout_report(P, Yte):
    err = P - Yte                                   # dimensionless (per-sea scales)
    return {
        "eta2_rel_rms_err_net": rel(err[:, 0], Yte[:, 0]),
        "u2_rel_rms_err_net": rel(err[:, 1::4], Yte[:, 1::4]),
        "v2_rel_rms_err_net": rel(err[:, 2::4], Yte[:, 2::4]),
        "w2_rel_rms_err_net": rel(err[:, 3::4], Yte[:, 3::4]),
        "p2_rel_rms_err_net": rel(err[:, 4::4], Yte[:, 4::4]),
        "eta2_max_err_net_over_max_exact": float(np.abs(err[:, 0]).max() / np.abs(Yte[:, 0]).max()),
    }


def cfg5_eta_error(net_cols, m, ncol=128):
    """Max |eta2_net - eta2_exact| / Hs of a network's own eta2 on ncol x ncol
    columns of the N=1024 short-crested sea (1 km x 1 km, t = 0.25 s)."""
    import torch
    from cfdwavemaker_amd import engine
    from cfdwavemaker_amd import surrogate as S
    from cfdwavemaker_amd.seastate import config_sea
    sea = config_sea(5)
    d = sea.as_dict()
    es = engine.Sea(d, precision=engine.FP32, with_direct=False)
    g = np.linspace(0.0, 1000.0, ncol)
    xx, yy = (q.ravel() for q in np.meshgrid(g, g, indexing="ij"))
    xd = torch.tensor(xx, dtype=torch.float64, device="cuda")
    yd = torch.tensor(yy, dtype=torch.float64, device="cuda")
    X, _, e2 = S.linear_fields(es, xd, yd, 0.25, sea.depth, S.peak_wavenumber(d))
    sx, sy = S.sea_scales(d)
    Xd = (X.cpu().numpy() / sx)[:, net_cols].astype(np.float32)
    P = m.forward(torch.tensor(Xd, device="cuda")).cpu().numpy()
    err = np.abs(P[:, 0] * sy[0] - e2.cpu().numpy())
    es.close()
    return float(err.max() / sea.hs)


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
    ap.add_argument("--feature-set", default="local+quad+eta2")
    ap.add_argument("--mode", default="hybrid", choices=["hybrid", "learned"])
    ap.add_argument("--budget", action="store_true", help="train every feature set and report its errors")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    import torch
    from cfdwavemaker_amd import surrogate as S

    def log(m):
        print(m, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    Xtr, Ytr = S.build_dataset(range(a.train_states), ncol=a.ncol, log=log)
    Xte, Yte = S.build_dataset(range(10000, 10000 + a.test_states), ncol=a.ncol, log=log)
    log(f"dataset {Xtr.shape} / {Xte.shape} in {time.perf_counter() - t0:.1f} s")
    sets = list(S.FEATURE_SETS) if a.budget else [a.feature_set]
    budget = {}
    for name in sets:
        cols = S.FEATURE_SETS[name]
        t1 = time.perf_counter()
        m, curve = S.train(Xtr[:, cols], Ytr, steps=a.steps, batch=a.batch, lr=a.lr, width=a.width,
                           nhidden=a.nhidden)
        log(f"[{name}] trained {a.steps} steps in {time.perf_counter() - t1:.1f} s")
        P = m.forward(torch.tensor(np.ascontiguousarray(Xte[:, cols]), device="cuda")).cpu().numpy()
        rep = {"features": len(cols), "train_rows": int(Xtr.shape[0]), "test_rows": int(Xte.shape[0]),
               "final_loss": float(curve[curve > 0][-1])}
        rep.update(heldout_report(P, Yte))
        rep["cfg5_learned_eta2_max_err_over_hs"] = cfg5_eta_error(cols, m)
        budget[name] = rep
        print(json.dumps({name: rep}), flush=True)
        if a.out and name == a.feature_set:
            sur = S.Surrogate(m, cols, a.mode)
            sur.save(a.out, report=json.dumps(rep))
    if a.budget:
        print(json.dumps({"budget": budget}), flush=True)


if __name__ == "__main__":
    main()
