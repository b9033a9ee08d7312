"""Train the learned pool-size network on the device.

    python -m hip_peel.train_pools [--runs 4000] [--steps 4000] [--seed 0]

1. peels ``--runs`` meshes drawn from the scene generators of the 256
   random triangles, the torus knot, the stack and the icosphere instances
   over their size ranges (``pools.mesh_specs``; meshes built in ``--workers``
   processes) with conservative pools and logs each pool's use
   (``pools.collect_runs``);
2. splits the runs 70/15/15 into training, calibration and test sets;
3. trains the network with the library's own device kernels
   (``pools.PoolTrainer``: ``pool_mlp_gradient`` + ``pool_mlp_adam``) on
   log1p(use) - log(faces), with under-predictions weighted 4x;
4. calibrates one log margin per pool on the calibration set
   (``pools.calibrate_margin``) and a per-face floor for closed meshes from
   the closed-surface runs of the training set (``pools.closed_mesh_floor``);
5. reports, on the test set, the fraction of runs that would overflow and the
   pool bytes against ``EventCapacities.for_mesh``, and writes the model
   (``hip_peel/pool_model.npz``) and the logged runs
   (``hip_peel/pool_runs.npz``).
"""

import argparse
import json
import time
from pathlib import Path

import numpy as np

from . import pools
from .events import EventCapacities

HERE = Path(__file__).resolve().parent


def pool_bytes(caps):
    """Approximate bytes of the five predicted pools (rows of capacities)."""
    w = np.array([pools.ELEMENT_BYTES[p] for p, _ in pools.TARGETS], dtype=np.float64)
    return np.asarray(caps, dtype=np.float64) @ w


def conservative(faces):
    rows = []
    for n in faces:
        c = EventCapacities.for_mesh(int(n))
        rows.append([getattr(c, p) for p, _ in pools.TARGETS])
    return np.array(rows, dtype=np.float64)


def train(features, usage, faces, steps, seed=0, lr=3e-3, under=4.0, device="cuda:0", kinds=None):
    """Returns (model, report)."""
    n = len(features)
    rng = np.random.default_rng(seed)
    order = rng.permutation(n)
    a, b = int(0.7 * n), int(0.85 * n)
    tr, cal, te = order[:a], order[a:b], order[b:]
    y = np.log1p(usage)
    per_face = y - features[:, :1]  # the network's target: log1p(use) - log(faces)
    x_mean, x_std = features[tr].mean(0), np.maximum(features[tr].std(0), 1e-6)
    y_mean, y_std = per_face[tr].mean(0), np.maximum(per_face[tr].std(0), 1e-6)
    trainer = pools.PoolTrainer(features.shape[1], y.shape[1], device=device, seed=seed)
    t0 = time.perf_counter()
    curve = trainer.fit((features[tr] - x_mean) / x_std, (per_face[tr] - y_mean) / y_std, steps, lr=lr,
                        under=under, log_every=max(1, steps // 20))
    seconds = time.perf_counter() - t0
    params = trainer.params.cpu().numpy()
    floor = pools.closed_mesh_floor(usage[tr], faces[tr], kinds[tr]) if kinds is not None else None
    model = pools.PoolModel(params, x_mean, x_std, y_mean, y_std, np.zeros(y.shape[1]), floor=floor)
    margin = pools.calibrate_margin(y[cal], model.predict_log_usage(features[cal], device))
    model.margin = margin
    pred = model.predict_log_usage(features[te], device)
    caps = np.maximum(pools.MIN_CAPACITY, np.ceil(np.expm1(pred + margin)))
    overflow = (caps < usage[te]).any(axis=1)
    ratio = pool_bytes(caps).sum() / pool_bytes(conservative(faces[te])).sum()
    median_ratio = float(np.median(pool_bytes(caps) / pool_bytes(conservative(faces[te]))))
    report = {
        "runs": int(n),
        "train": int(len(tr)),
        "calibration": int(len(cal)),
        "test": int(len(te)),
        "steps": int(steps),
        "train_seconds": round(seconds, 3),
        "loss_curve": [round(v, 5) for v in curve],
        "margin": [round(float(m), 4) for m in margin],
        "test_overflow_fraction": float(overflow.mean()),
        "test_pool_bytes_vs_conservative": float(ratio),
        "test_median_pool_bytes_ratio": median_ratio,
        "test_conservative_overflow_fraction": float((conservative(faces[te]) < usage[te]).any(axis=1).mean()),
    }
    model.meta = {"runs": n, "steps": steps, "final_loss": curve[-1] if curve else float("nan"),
                  "test_overflow_fraction": report["test_overflow_fraction"],
                  "test_pool_bytes_vs_conservative": ratio}
    return model, report


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--runs", type=int, default=4000)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--device", default="cuda:0")
    ap.add_argument("--workers", type=int, default=12, help="processes building the meshes")
    ap.add_argument("--out", type=Path, default=pools.MODEL_PATH)
    ap.add_argument("--runs-out", type=Path, default=HERE / "pool_runs.npz")
    ap.add_argument("--reuse-runs", action="store_true", help="train on the saved runs instead of peeling")
    args = ap.parse_args(argv)
    if args.reuse_runs:
        with np.load(args.runs_out, allow_pickle=False) as z:
            features, usage, kinds, faces = z["features"], z["usage"], z["kinds"], z["faces"]
    else:
        t0 = time.perf_counter()
        features, usage, kinds, faces = pools.collect_runs(args.runs, args.seed, args.device, progress=250,
                                                           workers=args.workers)
        print(f"collected {len(features)} runs in {time.perf_counter() - t0:.1f} s", flush=True)
        np.savez_compressed(args.runs_out, features=features, usage=usage, kinds=kinds, faces=faces)
    model, report = train(features, usage, faces, args.steps, seed=args.seed, device=args.device, kinds=kinds)
    model.save(args.out)
    report["kinds"] = {str(k): int((kinds == k).sum()) for k in np.unique(kinds)}
    print(json.dumps(report))


if __name__ == "__main__":
    main()
