"""Train the transfer-rate surrogate on the GPU and write the model file.

    python -m cnt_film_monte_carlo_amd.train_surrogate --out model.npz --report report.json

Labels come from the exact first-order rate on the device (exciton.Transfer); training runs
the library's BF16 MFMA kernels (kmc_mlp_train).  The model file holds the FP32 parameters,
the domain, the residual map and the training report.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

from . import surrogate


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--out", default=str(surrogate.DEFAULT_MODEL))
    ap.add_argument("--report", default=None)
    ap.add_argument("--n-train", type=int, default=2_000_000)
    ap.add_argument("--n-val", type=int, default=2_000_000)
    ap.add_argument("--epochs", type=int, default=1200)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--lr-min", type=float, default=1e-6)
    ap.add_argument("--seed", type=int, default=1001)
    ap.add_argument("--tol", type=float, default=0.05)
    ap.add_argument("--remap", default=None, metavar="MODEL",
                    help="keep MODEL's weights; rebuild its residual map from fresh held-out "
                         "pairs and write it to --out")
    a = ap.parse_args(argv)

    def log(msg):
        print(f"[train-surrogate] {msg}", flush=True)

    t0 = time.time()
    if a.remap:
        from . import exciton
        s = surrogate.RateSurrogate.load(a.remap)
        models = {c: exciton.CntModel(*c, 10) for c in surrogate.CHIRALITIES}
        st = surrogate.validate(s, models, a.n_val, a.seed + 104729, a.tol, log=log)
        s.meta.setdefault("train", {}).update(st)
        s.save(a.out)
        log(json.dumps(st))
        if a.report:
            with open(a.report, "w") as f:
                json.dump({"report": s.meta["train"]}, f, indent=1)
        return 0
    s, rep = surrogate.train_surrogate(a.n_train, a.n_val, a.epochs, a.batch, a.lr, a.lr_min,
                                       a.seed, a.tol, log=log)
    s.save(a.out)
    d = rep.__dict__
    d["total_seconds"] = time.time() - t0
    log(json.dumps(d))
    if a.report:
        with open(a.report, "w") as f:
            json.dump({"report": d, "loss_curve": s.meta.get("loss_curve", [])}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
