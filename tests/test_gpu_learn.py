"""GPU draw-order learning pieces against the oracle: per-batch labels (bit
exact), draws in an explicit batch order (bit exact against the oracle drawing
the same batches in the same order), and the 16 batch features (against a
numpy restatement of their definitions on the oracle's per-quad setup)."""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest

from rasterizer_amd import Rasterizer, Views, scenes

pytestmark = pytest.mark.gpu

SIZES = [(320, 192), (256, 256), (640, 360), (128, 64)]


@pytest.fixture(scope="module")
def scene():
    return scenes.config1(seed=4, n_boxes=512)


def _setup(scene, seed):
    rng = np.random.default_rng(seed)
    out = []
    for w, h in SIZES:
        eye = rng.uniform(-40, 40, 3).astype(np.float32)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        mvp = scenes.view_projection(eye, d, (0, 1, 0) if abs(d[1]) < 0.95 else (1, 0, 0), np.pi / 3, w / h)
        occ = [scene.occluders[i] for i in scenes.front_to_back(scene.occluders, eye)]
        out.append((mvp, occ))
    return out


def _batches(occ):
    """(occluder index, batch index within it) of each batch, in draw order."""
    return [(i, b) for i, o in enumerate(occ) for b in range(o.packets.size // 32)]


def _oracle_batch_draw(oracle_mod, w, h, mvp, occ, status, order):
    o = oracle_mod.OracleRasterizer(w, h)
    o.clear()
    o.setModelViewProjection(mvp)
    bl = _batches(occ)
    for k in order:
        i, b = bl[k]
        if status[i]:
            src = occ[i]
            part = SimpleNamespace(packets=src.packets[32 * b:32 * b + 32], m_refMin=src.m_refMin,
                                   m_refMax=src.m_refMax)
            o.rasterize(part, status[i] == 2)
    return o


def test_labels_match_oracle(oracle_mod, scene):
    owner = Rasterizer(64, 64)
    V = Views(owner, SIZES)
    setup = _setup(scene, 1)
    for k, (mvp, occ) in enumerate(setup):
        V.set_mvp(k, mvp)
        V.set_occluders(k, occ)
    V.enable_labels(True)
    # Two steps with different cameras: the second step's draw starts from
    # the first step's depth and winners, which its first writes replace.
    second = _setup(scene, 11)
    oracles = [oracle_mod.OracleRasterizer(w, h) for w, h in SIZES]
    for step, cams in enumerate((setup, second)):
        for k, (mvp, _) in enumerate(cams):
            V.set_mvp(k, mvp)
        V.run(graph=True)
        V.sync()
        for k, (w, h) in enumerate(SIZES):
            mvp, occ = cams[k][0], setup[k][1]
            o = oracles[k]
            o.clear()
            o.setModelViewProjection(mvp)
            st, want = o.frame_labels(occ)
            assert np.array_equal(V.status(k), st), f"step {step} view {k}: statuses"
            got, _ = V.labels(k)
            assert got.shape == want.shape
            assert np.array_equal(got, want), f"step {step} view {k}: {int((got != want).sum())} labels differ"
            # Every written block with a non-zero pixel has at least one winner.
            d, hz = o.raw()
            nb = (w // 8) * (h // 8)
            live = (hz[:nb] != 1) & (d.reshape(nb, 64).max(axis=1) > 0)
            assert want.sum() >= live.sum()
            print(f"step {step} view {k}: {want.size} batches, {int((want > 0).sum())} labelled, "
                  f"{int(live.sum())} live blocks", flush=True)


def test_batch_order_matches_oracle(oracle_mod, scene, report):
    owner = Rasterizer(64, 64)
    V = Views(owner, SIZES)
    setup = _setup(scene, 2)
    rng = np.random.default_rng(9)
    orders = []
    for k, (mvp, occ) in enumerate(setup):
        V.set_mvp(k, mvp)
        V.set_occluders(k, occ)
        n = V.n_batches(k)
        order = np.arange(n) if k == 0 else rng.permutation(n)
        orders.append(order)
        V.set_batch_order(k, order)
    V.run(graph=True)
    V.sync()
    for k, (w, h) in enumerate(SIZES):
        mvp, occ = setup[k]
        o0 = oracle_mod.OracleRasterizer(w, h)
        o0.clear()
        o0.setModelViewProjection(mvp)
        st = o0.draw_frame(occ, gated=False)
        o = _oracle_batch_draw(oracle_mod, w, h, mvp, occ, st, orders[k])
        dg, hg = V.raw(k)
        do, ho = o.raw()
        assert np.array_equal(hg, ho), f"view {k}: hiZ {report(hg, ho)}"
        assert np.array_equal(dg, do), f"view {k}: depth {report(dg, do)}"
        if k == 0:
            # Batch items in occluder order draw exactly the occluder-order frame.
            d0, h0 = o0.raw()
            assert np.array_equal(dg, d0) and np.array_equal(hg, h0)
        else:
            d0, _ = o0.raw()
            diff = np.abs(dg.astype(int) - d0.astype(int))
            print(f"view {k}: shuffled batch order vs occluder order: {int((diff > 0).sum())} px differ, "
                  f"max {int(diff.max())} LSB", flush=True)


def _features_ref(oracle_mod, w, h, mvp, occ, st, prev_labels):
    """The feature definitions of rz_features_kernel restated in numpy on the
    oracle's per-quad setup (float64; compared with a tolerance)."""
    o = oracle_mod.OracleRasterizer(w, h)
    o.setModelViewProjection(mvp)
    pre, raw = o.mvp_state()
    bX, bY = w // 8, h // 8
    vd = raw[12:15].astype(np.float64)
    vd /= np.linalg.norm(vd)
    rows = []
    nbt = sum(x.packets.size // 32 for x in occ)
    k = 0
    for i, oc in enumerate(occ):
        sd = o.setup_dump(oc, st[i] == 2) if st[i] else None
        pk = oc.packets.reshape(-1, 4, 8)
        ext = (oc.m_refMax - oc.m_refMin)[:3].astype(np.float64)
        for b in range(pk.shape[0]):
            f = np.zeros(16)
            boxes, zs, nds, convex = [], [], [], 0
            cen = np.zeros(3)
            for lane in range(8):
                wds = pk[b, :, lane]
                X = (wds.view(np.int32) >> 21).astype(np.float64) + 1024
                Y = ((wds >> 10) & 2047).astype(np.float64)
                Z = (wds & 1023).astype(np.float64)
                P = oc.m_refMin[:3].astype(np.float64) + np.stack([X / 2047, Y / 2047, Z / 1023], 1) * ext
                cen += P.sum(0) / 32
                if sd is None or not sd[8 * b + lane, 0]:
                    continue
                q = sd[8 * b + lane]
                fb, rx, ry = int(q[22]), int(q[4]), int(q[5])
                col, row = fb % bX, fb // bX
                if col + rx <= bX:
                    box = (col, col + rx - 1, row, min(row + ry - 1, bY - 1))
                else:
                    box = (0, bX - 1, row, min(row + ry, bY - 1))
                boxes.append(box)
                zs.append(q[6] / 65535)
                convex += int(q[1]) == 6
                e1, e2, e3 = P[1] - P[0], P[2] - P[0], P[3] - P[0]
                n = np.cross(e1, e2) + np.cross(e2, e3)
                nl = np.linalg.norm(n)
                nds.append(n @ vd / nl if nl > 0 else 0.0)
            n = len(boxes)
            f[0] = n / 8
            if n:
                c0 = min(b_[0] for b_ in boxes); c1 = max(b_[1] for b_ in boxes)
                r0 = min(b_[2] for b_ in boxes); r1 = max(b_[3] for b_ in boxes)
                f[1] = np.log2(1 + (c1 - c0 + 1) * (r1 - r0 + 1))
                f[2], f[3], f[4] = min(zs), max(zs), np.mean(zs)
                f[5] = np.log2(1 + sum((b_[1] - b_[0] + 1) * (b_[3] - b_[2] + 1) for b_ in boxes))
                f[7] = (0.5 * (c0 + c1) + 0.5) / bX
                f[8] = (0.5 * (r0 + r1) + 0.5) / bY
                f[11], f[12], f[13] = np.mean(nds), min(nds), max(nds)
                f[15] = convex / n
            else:
                f[7] = f[8] = 0.5
            f[6] = 1.0 if st[i] == 2 else 0.0
            wv = cen @ raw[12:15].astype(np.float64) + raw[15]
            f[9] = np.log2(1 + abs(wv))
            f[10] = np.log2(1 + prev_labels[k])
            f[14] = k / (nbt - 1) if nbt > 1 else 0.0
            rows.append(f)
            k += 1
    return np.array(rows)


def test_features_match_definition(oracle_mod, scene):
    owner = Rasterizer(64, 64)
    V = Views(owner, SIZES[:2])
    setup = _setup(scene, 3)
    for k in range(2):
        V.set_mvp(k, setup[k][0])
        V.set_occluders(k, setup[k][1])
    V.enable_labels(True)
    V.run(graph=True)
    V.sync()
    prev = [V.labels(k)[0] for k in range(2)]
    V.run(graph=True)  # the second step's features see the first step's labels
    V.sync()
    for k, (w, h) in enumerate(SIZES[:2]):
        mvp, occ = setup[k]
        o = oracle_mod.OracleRasterizer(w, h)
        o.clear()
        o.setModelViewProjection(mvp)
        st = o.draw_frame(occ, gated=False)
        want = _features_ref(oracle_mod, w, h, mvp, occ, st, prev[k])
        _, got = V.labels(k)
        assert got.shape == want.shape
        err = np.abs(got - want)
        assert np.all(err <= 1e-4 + 1e-4 * np.abs(want)), (np.unravel_index(np.argmax(err), err.shape), err.max())


def _scores_of_drawn(m, feats, occ, st):
    """Model scores of each batch of a view; batches of occluders the oracle
    did not draw (status 0) score -inf."""
    scores = m.forward(feats)
    drawn = np.repeat(np.asarray(st) != 0, [int(o.packets.size) // 32 for o in occ])
    assert drawn.size == scores.size
    return np.where(drawn, scores, -np.inf).astype(np.float32)


def test_learned_order_is_score_order_and_exact(oracle_mod, scene, report):
    """With a model set, each view draws its batches by descending model score
    (ties in occluder order); the buffer equals the oracle drawing the batches
    in that order, and the difference to occluder order is reported as pixels
    and maximum LSB."""
    from rasterizer_amd import Mlp

    owner = Rasterizer(64, 64)
    V = Views(owner, SIZES)
    setup = _setup(scene, 4)
    for k, (mvp, occ) in enumerate(setup):
        V.set_mvp(k, mvp)
        V.set_occluders(k, occ)
    m = Mlp(seed=11)
    V.set_model(m)
    V.run(graph=True)
    V.sync()
    V.run(graph=True)  # second step: features see the first step's labels
    V.sync()
    for k, (w, h) in enumerate(SIZES):
        mvp, occ = setup[k]
        _, feats = V.labels(k)
        o0 = oracle_mod.OracleRasterizer(w, h)
        o0.clear()
        o0.setModelViewProjection(mvp)
        st = o0.draw_frame(occ, gated=False)
        scores = _scores_of_drawn(m, feats, occ, st)
        want = np.lexsort((np.arange(scores.size), -scores))
        got = V.order(k)
        assert np.array_equal(got, want), f"view {k}: order is not descending score"
        o = _oracle_batch_draw(oracle_mod, w, h, mvp, occ, st, got)
        dg, hg = V.raw(k)
        do, ho = o.raw()
        written = np.repeat(ho[: (w // 8) * (h // 8)] != 1, 64)
        assert np.array_equal(hg, ho), f"view {k}: hiZ {report(hg, ho)}"
        assert np.array_equal(dg[written], do[written]), f"view {k}: depth {report(dg[written], do[written])}"
        d0, _ = o0.raw()
        diff = np.abs(dg.astype(int) - d0.astype(int))[written]
        print(f"view {k}: learned vs occluder order {int((diff > 0).sum())} px differ, max "
              f"{int(diff.max()) if diff.size else 0} LSB", flush=True)


def test_learned_occluder_order_is_score_order_and_exact(oracle_mod, scene, report):
    """Ordering whole occluders: each occluder's score is its best batch's
    score, occluders are drawn by descending score (ties in occluder order),
    and the buffer equals the oracle drawing the occluders in that order; the
    statuses stay in occluder order."""
    from rasterizer_amd import Mlp

    owner = Rasterizer(64, 64)
    V = Views(owner, SIZES)
    setup = _setup(scene, 5)
    for k, (mvp, occ) in enumerate(setup):
        V.set_mvp(k, mvp)
        V.set_occluders(k, occ)
    m = Mlp(seed=12)
    V.set_model(m, by_occluder=True)
    for _ in range(2):
        V.run(graph=True)
        V.sync()
    for k, (w, h) in enumerate(SIZES):
        mvp, occ = setup[k]
        _, feats = V.labels(k)
        o0 = oracle_mod.OracleRasterizer(w, h)
        o0.clear()
        o0.setModelViewProjection(mvp)
        scores = _scores_of_drawn(m, feats, occ, o0.draw_frame(occ, gated=False))
        nb = [int(o.packets.size) // 32 for o in occ]
        bounds = np.concatenate([[0], np.cumsum(nb)]).astype(int)
        fmax = np.float32(3.402823466e38)
        best = np.array([max(scores[bounds[i]:bounds[i + 1]].max(), -fmax) if nb[i] else -fmax
                         for i in range(len(occ))], np.float32)
        want = np.lexsort((np.arange(best.size), -best))
        got = V.order(k)
        assert np.array_equal(got, want), f"view {k}: occluder order is not descending best-batch score"
        o = oracle_mod.OracleRasterizer(w, h)
        o.clear()
        o.setModelViewProjection(mvp)
        st_sorted = o.draw_frame([occ[i] for i in got], gated=False)
        st = np.empty_like(st_sorted)
        st[got] = st_sorted
        assert np.array_equal(V.status(k), st), f"view {k}: statuses"
        dg, hg = V.raw(k)
        do, ho = o.raw()
        written = np.repeat(ho[: (w // 8) * (h // 8)] != 1, 64)
        assert np.array_equal(hg, ho), f"view {k}: hiZ {report(hg, ho)}"
        assert np.array_equal(dg[written], do[written]), f"view {k}: depth {report(dg[written], do[written])}"
    V.set_model(None)
    V.run(graph=True)
    V.sync()
    with pytest.raises(Exception):
        V.order(0)


def test_learned_occluder_order_in_gated_views(oracle_mod, scene, report):
    """The reference's query-gated loop (Main.cpp:192-206) over the occluders
    in the model's order: one query per occluder against the buffer drawn so
    far, as the reference.  Each view equals the oracle's gated loop over the
    occluders in the order the device chose, from the buffer the previous
    step left; statuses come back in submission order."""
    from rasterizer_amd import Mlp

    owner = Rasterizer(64, 64)
    V = Views(owner, SIZES)
    V.set_query_gated(True)
    setup = _setup(scene, 6)
    for k, (mvp, occ) in enumerate(setup):
        V.set_mvp(k, mvp)
        V.set_occluders(k, occ)
    m = Mlp(seed=13)
    V.set_model(m, by_occluder=True)
    V.run(graph=False)
    V.sync()
    before = [V.raw(k) for k in range(len(SIZES))]
    V.run(graph=False)  # second step: features see the first step's labels
    V.sync()
    drawn = 0
    for k, (w, h) in enumerate(SIZES):
        mvp, occ = setup[k]
        _, feats = V.labels(k)
        oc = oracle_mod.OracleRasterizer(w, h)
        oc.clear()
        oc.setModelViewProjection(mvp)
        scores = _scores_of_drawn(m, feats, occ, oc.draw_frame(occ, gated=False))
        nb = [int(o.packets.size) // 32 for o in occ]
        bounds = np.concatenate([[0], np.cumsum(nb)]).astype(int)
        fmax = np.float32(3.402823466e38)
        best = np.array([max(scores[bounds[i]:bounds[i + 1]].max(), -fmax) if nb[i] else -fmax
                         for i in range(len(occ))], np.float32)
        got = V.order(k)
        assert np.array_equal(got, np.lexsort((np.arange(best.size), -best))), f"view {k}: not score order"
        o = oracle_mod.OracleRasterizer(w, h)
        o.set_raw(*before[k])
        o.clear()
        o.setModelViewProjection(mvp)
        st_sorted = o.draw_frame([occ[i] for i in got], gated=True)
        st = np.empty_like(st_sorted)
        st[got] = st_sorted
        assert np.array_equal(V.status(k), st), f"view {k}: statuses"
        dg, hg = V.raw(k)
        do, ho = o.raw()
        assert np.array_equal(hg, ho), f"view {k}: hiZ {report(hg, ho)}"
        assert np.array_equal(dg, do), f"view {k}: depth {report(dg, do)}"
        drawn += int((st > 0).sum())
    assert drawn > 0
    # Batch-granular orders would query per batch: refused in gated mode.
    V.set_model(m, by_occluder=False)
    with pytest.raises(Exception):
        V.run(graph=False)


def _reject_rates(scene, models, size=(1280, 720), nv=32, seed=77):
    """hiZ/edge reject rate of (quad, block) visits over nv held-out cameras
    of the scene (eyes in its cube, random directions, as the models were
    trained), each camera's step preceded by one half a unit back (the
    previous frame whose labels the features read), for each order: 'plain'
    (occluders front to back, Main.cpp:184-190), 'random' (a random batch
    permutation) and each named model (batches by descending score)."""
    from rasterizer_amd import learn

    w, h = size
    owner = Rasterizer(64, 64)
    rng = np.random.default_rng(seed)
    V = Views(owner, [(w, h)] * nv)
    V.enable_stats(True)
    cams = [learn.random_camera(rng) for _ in range(nv)]
    for v, (eye, d, up) in enumerate(cams):
        V.set_occluders(v, [scene.occluders[i] for i in scenes.front_to_back(scene.occluders, eye)])
    out = {}
    for mode in ["plain", "random"] + list(models):
        V.set_model(models[mode] if mode in models else None)
        for v in range(nv):
            V.set_batch_order(v, rng.permutation(V.n_batches(v)) if mode == "random" else None)
        for step in range(2):
            for v, (eye, d, up) in enumerate(cams):
                V.set_mvp(v, scenes.view_projection(eye + (step - 1) * 0.5 * d, d, up, np.pi / 3, w / h))
            V.run(graph=True)
            ms = V.sync()
        st = V.stats()
        ts = V.tile_stats()
        # The summed per-tile counters hold the same block counts.
        assert (int(ts[13]), int(ts[14]), int(ts[15])) == (st["visited"], st["rejected"], st["written"])
        out[mode] = (st["rejected"] / max(st["visited"], 1), st["visited"], ms)
    V.set_model(None)
    return out


def test_trained_order_gains_rejects(scene):
    """Training end to end on the GPU (labels, features, the MLP's forward and
    backward kernels, Adam), then what a trained network gains over the
    order without it: the shipped models (trained the same way on 10,000
    frames of the 1,024-box scene at 1280x720, one with the previous-label
    feature and one without) reject more (quad, block) visits at the hiZ and
    edge tests than front-to-back order (the reference's) and than a random
    order, on held-out cameras."""
    from rasterizer_amd import learn

    feats, labels = learn.generate(scene, 384, seed=21, width=640, height=360, views_per_launch=64)
    m, norm, losses = learn.train(feats[:320], labels[:320], steps=150, frames_per_step=32)
    agree = learn.pair_agreement(m, feats[320:], labels[320:])
    print(f"loss {losses[0]:.4f} -> {np.mean(losses[-10:]):.4f}, held-out pair agreement {agree:.3f}", flush=True)
    assert np.mean(losses[-10:]) < losses[0]
    assert agree > 0.6

    box_scene = scenes.config1(seed=1)
    models = {"learned": learn.load(), "learned, no labels": learn.load(learn.WEIGHTS.with_name("order_mlp_nolabels.npz"))}
    res = _reject_rates(box_scene, models)
    for mode, (rate, visited, ms) in res.items():
        print(f"{mode:20s}: hiZ/edge reject rate {rate:.4f} of {visited:,} visited (quad, block) pairs, "
              f"{ms:.3f} ms", flush=True)
    for mode in models:
        assert res[mode][0] > res["plain"][0], f"{mode} does not beat front-to-back order"
        assert res[mode][0] > res["random"][0]
