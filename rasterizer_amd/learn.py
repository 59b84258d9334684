"""Training of the batch-scoring MLP that orders occluder batches.

Data: frames of a scene through the batched multi-view path with labels on
(rz_views_*).  Each frame is drawn twice, the second time after a small
camera move, so that the second step's features carry a previous frame's
labels as the reference scene's camera path would; the second step's labels
(blocks whose final depth the batch keeps) are the targets.

Training: pairs of batches of one frame with different labels; the network
should score the batch with the larger label higher (mean softplus ranking
loss); Adam steps run on the GPU through the library's own forward and
backward kernels (rz_mlp_*).
"""

from __future__ import annotations

import math
import time
from pathlib import Path

import numpy as np

from . import Mlp, Rasterizer, Views, scenes

WEIGHTS = Path(__file__).resolve().parent / "order_mlp.npz"


def random_camera(rng: np.random.Generator, half: float = 50.0):
    eye = rng.uniform(-half, half, 3).astype(np.float32)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    up = (0.0, 1.0, 0.0) if abs(d[1]) < 0.95 else (1.0, 0.0, 0.0)
    return eye, d.astype(np.float32), up


def street_camera(scene, rng: np.random.Generator):
    """A street-level camera in the city scene (scenes.config2): eye 2 units
    above the terrain at a random street intersection or along a street, a
    random horizontal direction tilted down as the scene's camera path is."""
    hf, pitch, grid = scene.meta["heightfield"], scene.meta["pitch"], scene.meta["grid"]
    along = rng.uniform(1.0, grid - 1.0)
    cross = float(rng.integers(1, grid))
    x, z = (along, cross) if rng.random() < 0.5 else (cross, along)
    x, z = x * pitch, z * pitch
    eye = np.array([x, float(hf.height(x, z)) + 2.0, z], np.float32)
    ang = rng.uniform(0.0, 2.0 * math.pi)
    d = np.array([math.cos(ang), -0.08, math.sin(ang)])
    d /= np.linalg.norm(d)
    return eye, d.astype(np.float32), (0.0, 1.0, 0.0)


def generate(scene, frames: int, seed: int, width: int = 1280, height: int = 720, views_per_launch: int = 64,
             log=print, camera=None):
    """Features (list per frame of (nBatches, 16)) and labels (per frame).
    camera(rng) -> (eye, direction, up); default: random_camera."""
    rng = np.random.default_rng(seed)
    camera = camera or random_camera
    owner = Rasterizer(64, 64)
    V = Views(owner, [(width, height)] * views_per_launch)
    V.enable_labels(True)
    feats, labels = [], []
    t0 = time.perf_counter()
    done = 0
    while done < frames:
        k = min(views_per_launch, frames - done)
        cams = [camera(rng) for _ in range(views_per_launch)]
        for v in range(views_per_launch):
            eye, d, up = cams[v]
            occ = [scene.occluders[i] for i in scenes.front_to_back(scene.occluders, eye)]
            V.set_occluders(v, occ)
            V.set_mvp(v, scenes.view_projection(eye - 0.5 * d, d, up, math.pi / 3, width / height))
        V.run(graph=True)  # previous frame: the eye half a unit back
        for v in range(views_per_launch):
            eye, d, up = cams[v]
            V.set_mvp(v, scenes.view_projection(eye, d, up, math.pi / 3, width / height))
        V.run(graph=True)
        V.sync()
        for v in range(k):
            lab, f = V.labels(v)
            feats.append(f.copy())
            labels.append(lab.copy())
        done += k
    log(f"labelled {frames} frames in {time.perf_counter() - t0:.1f} s")
    return feats, labels


def frame_pairs(labels: np.ndarray, rng: np.random.Generator, count: int) -> np.ndarray:
    """Pairs (i, j) of one frame with labels[i] > labels[j]."""
    n = labels.size
    i = rng.integers(0, n, count)
    j = rng.integers(0, n, count)
    keep = labels[i] != labels[j]
    i, j = i[keep], j[keep]
    hi = np.where(labels[i] > labels[j], i, j)
    lo = np.where(labels[i] > labels[j], j, i)
    return np.stack([hi, lo], 1)


LABEL_FEATURE = 10  # log2(1 + previous step's label)


def without_labels(feats):
    """Features with the previous-label feature held at 0 (for models that
    run without label tracking)."""
    out = []
    for f in feats:
        g = np.array(f, np.float32, copy=True)
        g[:, LABEL_FEATURE] = 0.0
        out.append(g)
    return out


def normalisation(feats) -> np.ndarray:
    x = np.concatenate(feats)
    mean = x.mean(0)
    std = x.std(0)
    inv = np.where(std > 1e-6, 1.0 / np.maximum(std, 1e-6), 1.0)
    return np.concatenate([mean, inv]).astype(np.float32)


def train(feats, labels, steps: int = 400, frames_per_step: int = 32, pairs_per_frame: int = 256, lr: float = 2e-3,
          seed: int = 1, log=print, use_labels: bool = True):
    """use_labels=False trains without the previous-label feature (held at 0),
    for ordering without label tracking."""
    rng = np.random.default_rng(seed)
    if not use_labels:
        feats = without_labels(feats)
    m = Mlp(seed=seed)
    m.set_uses_labels(use_labels)
    norm = normalisation(feats)
    m.set_params(norm=norm)
    losses = []
    t0 = time.perf_counter()
    nf = len(feats)
    for step in range(steps):
        pick = rng.choice(nf, min(frames_per_step, nf), replace=False)
        xs, ps, base = [], [], 0
        for f in pick:
            xs.append(feats[f])
            ps.append(frame_pairs(labels[f], rng, pairs_per_frame) + base)
            base += feats[f].shape[0]
        x = np.concatenate(xs)
        pr = np.concatenate(ps).astype(np.uint32)
        if pr.size == 0:
            continue
        losses.append(m.train_step(x, pr, lr * (0.5 if step > steps * 3 // 4 else 1.0)))
    log(f"trained {steps} steps in {time.perf_counter() - t0:.1f} s: loss {losses[0]:.4f} -> "
        f"{np.mean(losses[-20:]):.4f}")
    return m, norm, losses


def pair_agreement(m: Mlp, feats, labels, seed: int = 5, pairs_per_frame: int = 512) -> float:
    """Fraction of held-out within-frame pairs with different labels that the
    scores rank the same way."""
    rng = np.random.default_rng(seed)
    if not getattr(m, "uses_labels", True):
        feats = without_labels(feats)
    ok = tot = 0
    for f, lab in zip(feats, labels):
        pr = frame_pairs(lab, rng, pairs_per_frame)
        if pr.size == 0:
            continue
        s = m.forward(f)
        ok += int(np.sum(s[pr[:, 0]] > s[pr[:, 1]]))
        tot += pr.shape[0]
    return ok / max(tot, 1)


def save(m: Mlp, path: Path = WEIGHTS, **meta) -> None:
    w, norm = m.params()
    np.savez(path, params=w, norm=norm, uses_labels=np.asarray(int(getattr(m, "uses_labels", True))),
             **{k: np.asarray(v) for k, v in meta.items()})


def load(path: Path = WEIGHTS) -> Mlp:
    d = np.load(path, allow_pickle=False)
    m = Mlp(seed=0)
    m.set_params(d["params"], d["norm"])
    m.set_uses_labels(bool(int(d["uses_labels"])) if "uses_labels" in d.files else True)
    return m


def main() -> int:
    import argparse

    ap = argparse.ArgumentParser(description="train the batch-ordering MLP")
    ap.add_argument("--scene", choices=["boxes", "boxes3", "city"], default="boxes",
                    help="boxes: the 1,024-box scene, random cameras in its cube (1280x720); "
                         "boxes3: the 40,000-box scene of the batched multi-view workload, cameras inside its "
                         "cube as its views are placed (1024x1024); "
                         "city: the city scene, street-level cameras (1920x1080)")
    ap.add_argument("--frames", type=int, default=10000)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--no-labels", action="store_true",
                    help="train without the previous-label feature, for ordering without label tracking")
    ap.add_argument("--out", default=str(WEIGHTS))
    args = ap.parse_args()
    n_test = max(args.frames // 10, 64)
    if args.scene == "boxes":
        scene = scenes.config1(seed=1)
        feats, labels = generate(scene, args.frames + n_test, seed=10)
    elif args.scene == "boxes3":
        scene = scenes.config3(seed=3, sizes=[])
        feats, labels = generate(scene, args.frames + n_test, seed=13, width=1024, height=1024, views_per_launch=32,
                                 camera=lambda rng: random_camera(rng, half=0.8 * 171.0))
    else:
        scene = scenes.config2(seed=1, frames=8, n_queries=16)
        feats, labels = generate(scene, args.frames + n_test, seed=12, width=1920, height=1080, views_per_launch=32,
                                 camera=lambda rng: street_camera(scene, rng))
    tr_f, tr_l = feats[:args.frames], labels[:args.frames]
    te_f, te_l = feats[args.frames:], labels[args.frames:]
    m, norm, losses = train(tr_f, tr_l, steps=args.steps, use_labels=not args.no_labels)
    agree = pair_agreement(m, te_f, te_l)
    print(f"held-out pair agreement {agree:.3f} over {len(te_f)} frames")
    save(m, Path(args.out), agreement=agree, frames=args.frames, steps=args.steps, final_loss=np.mean(losses[-20:]),
         scene=args.scene)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
