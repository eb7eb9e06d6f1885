"""GPU tests of the learned warm start against its CPU oracle
(oracle/learn_oracle.py): features bit-exact, forward / backward / Adam of the bf16
MLP within stated tolerances, training that lowers the loss and a refined guess
that helps NDT.

Tolerances: the GPU accumulates in float32 (split-K partial sums, then bias), the
oracle in float64 with the same bf16 roundings of weights, activations and
backpropagated signals; a float32 difference that moves a value across a bf16
rounding boundary changes it by one bf16 ulp (2^-8 relative).  Forward outputs
must agree within 1e-2 of their largest magnitude and every gradient within 2e-2
relative Frobenius error (observed: about 1e-3)."""
import numpy as np
import pytest

import learn_oracle as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_scene(A):
    w = A.make_workload("cfg0", seed=3, scans=6)
    return w


def test_features_match_oracle(A, ctx, small_scene):
    w = small_scene
    sc = A.Scans(ctx, w.scans)
    bm = A.BevMap(ctx, w.map_xyzw)
    g = A.perturb_poses(w.gt, np.random.default_rng(0))
    idx = np.array([0, 1, 2, 3, 4, 5, 2, 0], np.int32)
    gg = np.concatenate([g, g[:2]])
    F = bm.features(sc, idx, gg)
    lut = L.log1p_lut()
    for b, s in enumerate(idx):
        assert np.array_equal(F[b, :32768], L.scan_grid(w.scans[s], lut)), b
        assert np.array_equal(F[b, 32768:], L.map_grid(w.map_xyzw, gg[b], lut)), b
    assert (F[:, :32768] != 0).sum() > 1000 and (F[:, 32768:] != 0).sum() > 10000


def test_features_of_dense_inputs_match_oracle(A, ctx):
    """Scans and map crops of more than 65535 points take the kernels' two-pass
    32-bit counters instead of the packed 16-bit ones; a cell of more than 65535
    points saturates.  Both paths, and the packed one beside them, bit-exact."""
    r = np.random.default_rng(7)

    def cloud(n, lo, hi):
        p = np.zeros((n, 4), np.float32)
        p[:, :3] = r.uniform(lo, hi, size=(n, 3))
        p[:, 3] = 1.0
        return p

    # one 0.5 m x 0.5 m x 1.5 m cell with 66,000 points plus 5,000 spread ones
    dense = np.concatenate([cloud(66000, (3.05, -2.45, 0.6), (3.45, -2.05, 1.9)),
                            cloud(5000, (-15.0, -15.0, -2.0), (15.0, 15.0, 9.0))])
    small = cloud(20000, (-15.0, -15.0, -2.0), (15.0, 15.0, 9.0))
    scans = [dense, small]
    # map: 90,000 points in a 24 m x 24 m patch, and a guess in its middle
    m = cloud(90000, (-12.0, -12.0, -2.0), (12.0, 12.0, 9.0))
    sc = A.Scans(ctx, scans)
    bm = A.BevMap(ctx, m)
    g = np.stack([A.pose(0.3, -0.2, 0.1, 0.0, 0.0, 0.2), A.pose(40.0, 40.0, 0.0, 0.0, 0.0, 0.0)])
    F = bm.features(sc, np.array([0, 1], np.int32), g)
    lut = L.log1p_lut()
    for b in range(2):
        assert np.array_equal(F[b, :32768], L.scan_grid(scans[b], lut)), b
        assert np.array_equal(F[b, 32768:], L.map_grid(m, g[b], lut)), b
    assert F[0, :32768].max() == lut[65535]          # the saturated cell
    assert (F[0, 32768:] != 0).sum() > 10000          # the dense crop counted
    assert (F[1, 32768:] != 0).sum() == 0             # a footprint beside the map


def test_features_world_origin(A, ctx, small_scene):
    """Guesses in world coordinates: the map grid is taken relative to the origin."""
    w = small_scene
    origin = np.array([4.2e5, 5.6e6, 80.0])
    sc = A.Scans(ctx, w.scans[:2])
    bm = A.BevMap(ctx, w.map_xyzw, origin)
    g = A.perturb_poses(w.gt[:2], np.random.default_rng(1))
    gw = g.copy()
    gw[:, :3, 3] += origin
    F = bm.features(sc, np.array([0, 1], np.int32), gw)
    for b in range(2):
        assert np.array_equal(F[b, 32768:], L.map_grid(w.map_xyzw, g[b]))


def _rp_dict(rp):
    return dict(radius=rp.radius, yaws=rp.yaws, yaw_step=rp.yaw_step, refine_yaw=rp.refine_yaw, z_search=rp.z_search,
                level=rp.level)


@pytest.mark.parametrize("kw,origin", [(dict(), (0.0, 0.0, 0.0)), (dict(refine_yaw=0, radius=8, yaws=11),
                                                                    (4.2e5, 5.6e6, 80.0)),
                                       (dict(yaws=1, radius=3, z_search=0, level=0), (0.0, 0.0, 0.0))])
def test_register_matches_oracle(A, ctx, small_scene, kw, origin):
    """Registered (and leveled) guesses and best scores bit for bit equal to the
    oracle's, for guesses off by up to +-1 m / +-3 degrees and +-0.5 degrees of roll
    and pitch, a guess beside the map (no cell in common: the guess comes back
    unchanged with score 0) and a scan above the grid's height range, with a world
    origin far from the map-local frame."""
    w = small_scene
    high = w.scans[0].copy()
    high[:, 2] += 40.0
    scans = list(w.scans) + [high]
    sc = A.Scans(ctx, scans)
    o = np.array(origin)
    bm = A.BevMap(ctx, w.map_xyzw, o)
    g = A.perturb_poses(w.gt, np.random.default_rng(4), xy=1.0, z=0.3, yaw_deg=3.0, rollpitch_deg=0.5)
    gl = np.concatenate([g, g[:1], w.gt[:1]])
    gl[len(w.scans), 0, 3] += 500.0
    idx = np.array(list(range(len(w.scans))) + [0, len(w.scans)], np.int32)
    gw = gl.copy()
    gw[:, :3, 3] += o
    rp = A.register_params(**kw)
    out, score = bm.register(sc, idx, gw, rp)
    lut = L.log1p_lut()
    for b, s in enumerate(idx):
        gl_b = gw[b].copy()
        gl_b[:3, 3] -= o
        el, ew, v = L.register(scans[s], w.map_xyzw, gl_b, gw[b], _rp_dict(rp), lut)
        assert score[b] == v and np.array_equal(out[b], ew), (b, score[b], v)
    # no cell in common: the guess stays, apart from the height search (the map
    # beside it is empty, the scan above the grid still has heights, and above the
    # leveling's height range too)
    assert score[len(w.scans)] == 0 and np.array_equal(out[len(w.scans)], gw[len(w.scans)])
    assert score[-1] == 0 and np.array_equal(out[-1][:, :3], gw[-1][:, :3])
    if rp.level:
        # leveling turned the registered guesses towards the truth's roll and pitch
        # (on these 2 k-point scans from 0.42 to 0.18 degrees: most 1 m cells hold
        # one or two scan points, so their lowest heights are noisy)
        def tilt(G, T):
            E = np.linalg.inv(T) @ G
            return np.hypot(np.arctan2(E[2, 1], E[2, 2]), np.arcsin(np.clip(-E[2, 0], -1, 1)))
        gt_w = w.gt.copy()
        gt_w[:, :3, 3] += o
        before = np.mean([tilt(gw[b], gt_w[b]) for b in range(len(w.scans))])
        after = np.mean([tilt(out[b], gt_w[b]) for b in range(len(w.scans))])
        assert after < 0.5 * before, (before, after)
    assert (score[:len(w.scans)] > 100).all()
    off = bm.register(sc, idx, gw, A.register_params(radius=0))[0]
    assert np.array_equal(off, gw)
    with pytest.raises(A.NdtError):
        bm.register(sc, idx, gw, A.register_params(yaws=4))
    with pytest.raises(A.NdtError):
        bm.register(sc, idx, gw, A.register_params(radius=9))


def _random_model(A, ctx, f0, f1, f2, seed):
    m = A.Model(ctx, f0, f1, f2, seed=seed)
    r = np.random.default_rng(seed)
    P = {"W1": r.normal(size=(f1, f0)) * np.sqrt(2 / f0), "b1": r.normal(size=f1) * 0.1,
         "W2": r.normal(size=(f2, f1)) * np.sqrt(2 / f1), "b2": r.normal(size=f2) * 0.1,
         "W3": r.normal(size=(6, f2)) * np.sqrt(1 / f2), "b3": r.normal(size=6) * 0.1}
    for k, key in enumerate(["W1", "b1", "W2", "b2", "W3", "b3"]):
        m.set(k, P[key].astype(np.float32))
        P[key] = P[key].astype(np.float32).astype(np.float64)
    return m, P, r


@pytest.mark.parametrize("dims", [(256, 128, 128, 128), (1024, 256, 128, 256), (65536, 1024, 256, 128)])
def test_forward_backward_adam_match_oracle(A, ctx, dims):
    f0, f1, f2, B = dims
    m, P, r = _random_model(A, ctx, f0, f1, f2, f0 + B)
    xb = L.bf16_bits(np.abs(r.normal(size=(B, f0))).astype(np.float32))
    x = L.bf16_value(xb).astype(np.float64)
    T = r.normal(size=(B, 6)).astype(np.float32)
    wts = np.array([1, 1, 0.5, 2, 2, 1], np.float32)
    y = m.forward(xb)
    loss = m.backward(xb, T, wts)
    lo, go, fo = L.mlp_backward(x, P, T.astype(np.float64), wts.astype(np.float64), bf16=True)
    assert np.abs(y - fo["Y"]).max() <= 1e-2 * np.abs(fo["Y"]).max()
    assert abs(loss - lo) <= 1e-2 * lo
    for k, key in enumerate(["W1", "b1", "W2", "b2", "W3", "b3"]):
        gg = m.grad(k).astype(np.float64)
        rel = np.linalg.norm(gg - go[key]) / np.linalg.norm(go[key])
        assert rel < 2e-2, (key, rel)
    # one Adam step on the stored gradients
    w0 = [m.get(k).astype(np.float64) for k in range(6)]
    g0 = [m.grad(k).astype(np.float64) for k in range(6)]
    m.adam_step(1e-3)
    for k in range(6):
        we, _, _ = L.adam_step(w0[k], np.zeros_like(w0[k]), np.zeros_like(w0[k]), g0[k], 1, 1e-3)
        assert np.abs(m.get(k) - we).max() < 1e-6, k


def test_model_save_load(A, ctx, tmp_path):
    m, P, r = _random_model(A, ctx, 256, 128, 128, 7)
    p = tmp_path / "m.bin"
    m.save(str(p))
    m2 = A.Model(ctx, path=str(p))
    assert m2.dims == (256, 128, 128) and m2.get_register().radius == 0
    m.set_register(A.register_params(radius=5, yaws=7))
    m.save(str(p))
    r3 = A.Model(ctx, path=str(p)).get_register()
    assert (r3.radius, r3.yaws, r3.yaw_step, r3.refine_yaw, r3.z_search, r3.level) == \
        (5, 7, A.register_params().yaw_step, 1, 1, 1)
    m.set_register(A.register_params(z_search=0, level=0))
    m.save(str(p))
    r4 = A.Model(ctx, path=str(p)).get_register()
    assert (r4.z_search, r4.level) == (0, 0)
    # a file with the earlier trailer (no level) loads with level 0
    raw = bytearray(p.read_bytes())
    k = raw.rindex(b"NDTREG03")
    old = raw[:k] + b"NDTREG02" + raw[k + 8:k + 8 + 16] + raw[k + 8 + 20:]
    p.write_bytes(bytes(old))
    r5 = A.Model(ctx, path=str(p)).get_register()
    assert (r5.radius, r5.z_search, r5.level, r5.yaw_step) == (6, 0, 0, A.register_params().yaw_step)
    with pytest.raises(A.NdtError):
        m.set_register(A.register_params(level=2))
    xb = L.bf16_bits(np.abs(r.normal(size=(128, 256))).astype(np.float32))
    assert np.array_equal(m.forward(xb), m2.forward(xb))


def test_invalid_model_arguments(A, ctx):
    with pytest.raises(A.NdtError):
        A.Model(ctx, 100, 128, 128)
    m = A.Model(ctx, 256, 128, 128)
    with pytest.raises(A.NdtError):
        m.forward(np.zeros((100, 256), np.uint16))


@pytest.fixture(scope="module")
def trained(A, ctx):
    """A proposer trained on 24,000 scans of the cfg2 map (8 guesses each, half within
    +-3 m / +-10 deg yaw, half within +-1 m / +-3 deg), labelled by NDT from the true
    pose, built in three batches of scans; 256 held-out scans (other sensor poses)
    for evaluation."""
    w = A.make_workload("cfg2", seed=1, scans=4)
    prm = A.params(A.NDT_MODE_NDT, origin=w.origin)
    grid = A.Grid(ctx, w.map_xyzw, prm)
    bm = A.BevMap(ctx, w.map_xyzw, w.origin)
    rng = np.random.default_rng(5)
    ds = ds_reg = None
    rp = A.register_params()
    for c0 in range(0, 24000, 8000):
        tr, tr_gt, _ = A.make_scans(w.scene, w.map_xyzw, A.scan_sizes(1000, 60000, 8000, 99 + c0),
                                    100000 + c0 + np.arange(8000))
        trS = A.Scans(ctx, tr)
        lab = trS.align(grid, tr_gt, prm)
        labels = np.stack([r["final_f64"] for r in lab])
        keep = np.nonzero([r["converged"] for r in lab])[0]
        idx = np.repeat(keep, 8).astype(np.int32)
        wide = rng.uniform(size=len(idx)) < 0.5
        guesses = np.where(wide[:, None, None], A.perturb_poses(tr_gt[idx], rng),
                           A.perturb_poses(tr_gt[idx], rng, xy=1.0, z=1.0, yaw_deg=3.0))
        # registered as bench.py does: a coarse search (+-4 m, +-10 deg), then rp
        coarse = A.register_params(radius=8, yaws=11, yaw_step=float(np.radians(2.0)), level=0)
        reg, _ = bm.register(trS, idx, bm.register(trS, idx, guesses, coarse)[0], rp)
        if ds is None:
            ds = A.Dataset(ctx, bm, trS, idx, guesses, labels[idx])
            ds_reg = A.Dataset(ctx, bm, trS, idx, reg, labels[idx])
        else:
            ds.add(trS, idx, guesses, labels[idx])
            ds_reg.add(trS, idx, reg, labels[idx])
        del tr, trS
    model = A.Model(ctx, seed=1)
    losses = model.train(ds, A.train_params(epochs=8, batch=512))
    model_reg = A.Model(ctx, seed=1)
    model_reg.set_register(rp)
    losses_reg = model_reg.train(ds_reg, A.train_params(epochs=8, batch=512, weights=[0.1, 0.1, 1, 1, 1, 1]))
    ds.close()
    ds_reg.close()
    ev, ev_gt, _ = A.make_scans(w.scene, w.map_xyzw, A.scan_sizes(1000, 60000, 256, 77), 900000 + np.arange(256))
    evS = A.Scans(ctx, ev)
    eg = A.perturb_poses(ev_gt, np.random.default_rng(9))
    return dict(w=w, prm=prm, grid=grid, bm=bm, model=model, losses=losses, evS=evS, ev_gt=ev_gt, eg=eg, ev=ev,
                model_reg=model_reg, losses_reg=losses_reg, rp=rp)


def test_predictions_depend_only_on_weights_and_inputs(A, ctx, trained):
    """Predictions for a batch, with a prediction for another batch in between,
    equal those of a fresh model loaded with the same weights."""
    t = trained
    other = A.Scans(ctx, t["ev"][:40])
    g2 = A.perturb_poses(t["ev_gt"], np.random.default_rng(21))
    m = t["model"]
    p1 = m.predict(t["bm"], t["evS"], t["eg"])
    po = m.predict(t["bm"], other, g2[:40])
    p2 = m.predict(t["bm"], t["evS"], g2)
    fresh = A.Model(ctx, seed=3)
    for k in range(6):
        fresh.set(k, m.get(k))
    assert np.array_equal(p1, fresh.predict(t["bm"], t["evS"], t["eg"]))
    assert np.array_equal(p2, fresh.predict(t["bm"], t["evS"], g2))
    assert np.array_equal(po, fresh.predict(t["bm"], other, g2[:40]))


def test_learned_step_replays(A, ctx, trained):
    """The fallback policy runs as one captured graph per batch: replays with other
    guesses in between give the same results, equal to those of a fresh model with
    the same weights (its own capture) and to the uncaptured race's two starts."""
    t = trained
    ev, eg, m = t["evS"], t["eg"], t["model"]
    pol = A.LEARNED_FALLBACK | A.LEARNED_KEEP_ROTATION
    g2 = A.perturb_poses(t["ev_gt"], np.random.default_rng(33), xy=1.0, z=1.0, yaw_deg=3.0)
    r1, u1 = m.align(t["grid"], t["bm"], ev, eg, t["prm"], pol)
    r2, u2 = m.align(t["grid"], t["bm"], ev, g2, t["prm"], pol)
    r3, u3 = m.align(t["grid"], t["bm"], ev, eg, t["prm"], pol)
    fresh = A.Model(ctx, seed=3)
    for k in range(6):
        fresh.set(k, m.get(k))
    r4, u4 = fresh.align(t["grid"], t["bm"], ev, g2, t["prm"], pol)
    assert np.array_equal(u1, u3) and np.array_equal(u2, u4)
    for a, b in zip(r1, r3):
        assert np.array_equal(a["final_f64"], b["final_f64"]) and a["passes"] == b["passes"]
    for a, b in zip(r2, r4):
        assert np.array_equal(a["final_f64"], b["final_f64"]) and a["passes"] == b["passes"]
    assert any(not np.array_equal(a["final_f64"], b["final_f64"]) for a, b in zip(r1, r2))


def _success(results, gt):
    return np.array([np.linalg.norm(r["final_f64"][:3, 3] - g[:3, 3]) < 0.05 for r, g in zip(results, gt)])


def test_learned_warm_start_gains(A, trained):
    """What the trained network adds on held-out scans with +-3 m / +-10 deg guesses:
    refined guesses closer to the truth, and NDT that reaches the truth more often
    and converges at least as often as plain NDT."""
    t = trained
    assert t["losses"][-1] < 0.5 * t["losses"][0]
    pred = t["model"].predict(t["bm"], t["evS"], t["eg"])
    raw_err = np.linalg.norm(t["eg"][:, :2, 3] - t["ev_gt"][:, :2, 3], axis=1)
    pred_err = np.linalg.norm(pred[:, :2, 3] - t["ev_gt"][:, :2, 3], axis=1)
    assert pred_err.mean() < 0.85 * raw_err.mean(), (raw_err.mean(), pred_err.mean())
    plain = t["evS"].align(t["grid"], t["eg"], t["prm"])
    race, used = t["model"].align(t["grid"], t["bm"], t["evS"], t["eg"], t["prm"], A.LEARNED_RACE)
    sp, sr = _success(plain, t["ev_gt"]).mean(), _success(race, t["ev_gt"]).mean()
    assert sr >= sp + 0.08, (sp, sr)
    assert np.mean([r["converged"] for r in race]) >= np.mean([r["converged"] for r in plain])


@pytest.mark.parametrize("which,fine_pass", [("model", 0), ("model_reg", 0), ("model_reg", 12)])
def test_policies_return_plain_or_learned_runs(A, trained, which, fine_pass):
    """Every result of either policy, with the model's full pose or only its
    translation, is bit for bit the NDT run from the given guess or the one from the
    refined guess, as used_learned says; the race keeps the better final score; an
    unknown policy is rejected.  With a registration the refined guess corrects the
    registered guess, whose rotation the translation-only variant keeps."""
    t = dict(trained)
    if fine_pass:
        t["prm"] = A.NdtParams.from_buffer_copy(t["prm"])
        t["prm"].fine_pass = fine_pass
    ev, eg, m = t["evS"], t["eg"], t[which]
    pred = m.predict(t["bm"], ev, eg)
    base = eg if which == "model" else t["bm"].register(ev, np.arange(ev.S, dtype=np.int32), eg, t["rp"])[0]
    plain = ev.align(t["grid"], eg, t["prm"])
    from_pred = ev.align(t["grid"], pred, t["prm"])
    pred_t = pred.copy()
    pred_t[:, :3, :3] = base[:, :3, :3]
    from_pred_t = ev.align(t["grid"], pred_t, t["prm"])
    for policy in (A.LEARNED_RACE, A.LEARNED_FALLBACK, A.LEARNED_RACE | A.LEARNED_KEEP_ROTATION,
                   A.LEARNED_FALLBACK | A.LEARNED_KEEP_ROTATION):
        learned = from_pred_t if policy & A.LEARNED_KEEP_ROTATION else from_pred
        res, used = m.align(t["grid"], t["bm"], ev, eg, t["prm"], policy)
        for s, r in enumerate(res):
            ref = learned[s] if used[s] else plain[s]
            assert np.array_equal(r["final_f64"], ref["final_f64"]) and r["iterations"] == ref["iterations"], (policy, s)
            if (policy & 1) == A.LEARNED_RACE and plain[s]["converged"] == learned[s]["converged"]:
                assert r["score"] == max(plain[s]["score"], learned[s]["score"])
    with pytest.raises(A.NdtError):
        m.align(t["grid"], t["bm"], ev, eg, t["prm"], 7)


def test_registered_prediction_corrects_the_registered_guess(A, trained):
    """With a registration set, the model sees the map around the registered guess
    and its correction applies to that guess: predictions equal registered guess *
    C(model output on the features at the registered guess), to rounding."""
    t = trained
    m, ev, eg = t["model_reg"], t["evS"], t["eg"]
    idx = np.arange(ev.S, dtype=np.int32)
    reg, _ = t["bm"].register(ev, idx, eg, t["rp"])
    y = m.forward(t["bm"].features(ev, idx, reg)).astype(np.float64)
    pred = m.predict(t["bm"], ev, eg)
    for s in range(ev.S):
        ref = reg[s] @ L.pose_matrix(y[s] * L.CORR_SCALE)
        assert np.abs(pred[s] - ref).max() < 1e-9, s


def test_registered_warm_start_saves_passes(A, trained):
    """Held-out scans with +-1 m / +-3 degree guesses: the registered model's refined
    guesses lie much closer to the truth, and the fallback policy (translation from
    the model) spends at least a fifth fewer derivative passes than plain NDT, probe
    included, converging and reaching the truth at least as often."""
    t = trained
    g = A.perturb_poses(t["ev_gt"], np.random.default_rng(12), xy=1.0, z=1.0, yaw_deg=3.0)
    pred = t["model_reg"].predict(t["bm"], t["evS"], g)
    raw = np.linalg.norm(g[:, :3, 3] - t["ev_gt"][:, :3, 3], axis=1)
    ref = np.linalg.norm(pred[:, :3, 3] - t["ev_gt"][:, :3, 3], axis=1)
    assert ref.mean() < 0.5 * raw.mean(), (raw.mean(), ref.mean())
    plain = t["evS"].align(t["grid"], g, t["prm"])
    res, used = t["model_reg"].align(t["grid"], t["bm"], t["evS"], g, t["prm"],
                                     A.LEARNED_FALLBACK | A.LEARNED_KEEP_ROTATION)
    pp, pl = np.mean([r["passes"] for r in plain]), np.mean([r["passes"] for r in res])
    assert pl < 0.8 * pp, (pp, pl)
    assert np.mean([r["converged"] for r in res]) >= np.mean([r["converged"] for r in plain])
    assert _success(res, t["ev_gt"]).mean() >= _success(plain, t["ev_gt"]).mean()
