"""Learned safety on the GPU: the fused BF16 matrix-core forward and the
training kernels' gradient against the numpy oracle (oracle/mlp.py), training
of a tessellated solid's network with zero overestimates of the served bound
on fresh points, and the learned transport: trajectories bit-identical to the
FP64 path, no audited violation, and the work the network saves (ablation
against the same kernel with a network that never serves a bound)."""
import time

import numpy as np
import pytest

import vecgeom_amd as vg
from oracle import mlp
from vecgeom_amd import sampling, workloads

pytestmark = pytest.mark.gpu

UNIT_GEOM = np.array([0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 0.0])  # centre 0, half-widths 1, scale 1, margin 0


def _points(seed, n, half=1.0):
    return sampling.points_in_box(seed, 1, n, (half, half, half), 1.0)


@pytest.mark.parametrize("seed,scale", [(1, 1.0), (2, 3.0)])
def test_forward_matches_oracle(ctx, seed, scale):
    """vgm_mlp_predict (fused forward, weights in LDS, layer-to-layer in
    registers) against the oracle with the same BF16 rounding points."""
    print(f"[learned] forward seed={seed} scale={scale}", flush=True)
    s = ctx.box(1.0, 1.0, 1.0)
    p = mlp.random_params(seed, scale=scale)
    ctx.mlp_set(s, p.astype(np.float32), UNIT_GEOM)
    x, y, z = _points(seed, 4099)  # not a multiple of 64: partial last wave
    got = ctx.mlp_predict(s, x, y, z)
    ref, _ = mlp.forward(p.astype(np.float32).astype(np.float64), np.stack([x, y, z], axis=1), emulate=True)
    err = np.abs(got - ref)
    tol = 2e-2 * (1.0 + np.abs(ref))
    print(f"[learned] forward max err {err.max():.3e} (|y| max {np.abs(ref).max():.3f})", flush=True)
    assert np.all(err <= tol), (int(np.argmax(err - tol)), err.max())
    assert np.mean(err) < 2e-3 * (1.0 + np.mean(np.abs(ref)))
    # served bound = max(0, prediction - margin)
    g = UNIT_GEOM.copy()
    g[7] = 0.25
    ctx.mlp_set(s, p.astype(np.float32), g)
    b = ctx.mlp_bound(s, x, y, z)
    assert np.allclose(b, np.maximum(0.0, got - 0.25), rtol=0, atol=1e-12)


@pytest.mark.parametrize("seed", [5, 6])
def test_gradient_matches_oracle(ctx, seed):
    """Training kernels' gradient (BF16 products, FP32 accumulation) against
    the oracle's float64 gradient with the same rounding points."""
    print(f"[learned] gradient seed={seed}", flush=True)
    rng = np.random.default_rng(seed)
    p = mlp.random_params(seed).astype(np.float32)
    n = 512
    u = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    t = rng.uniform(0.0, 0.5, n).astype(np.float32)
    g, loss = ctx.mlp_gradient(p, u, t, over_weight=4.0)
    lref, gref = mlp.loss_and_grad(p.astype(np.float64), u.astype(np.float64), t.astype(np.float64), 4.0)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    cos = float(np.dot(g, gref) / (np.linalg.norm(g) * np.linalg.norm(gref)))
    print(f"[learned] gradient rel err {rel:.3e} cos {cos:.6f} loss {loss:.6f} / {lref:.6f}", flush=True)
    assert abs(loss - lref) <= 1e-3 * lref
    assert rel < 2e-2 and cos > 0.9998
    # every parameter block agrees on its own
    offs = np.cumsum([0, 128 * 51, 128, 16384, 128, 16384, 128, 16384, 128, 128, 1])
    for a, b in zip(offs[:-1], offs[1:]):
        na = np.linalg.norm(gref[a:b])
        assert np.linalg.norm(g[a:b] - gref[a:b]) <= 3e-2 * na + 1e-7, (a, b)


def _hull(ctx):
    return ctx.tessellated(workloads.ellipsoid_hull(2000, (12.0, 50.0, 30.0), 4))


def test_training_bound_is_conservative(ctx):
    """Train on the GPU; on 400k fresh points the served bound never exceeds
    the exact safety, and the prediction error is small against the size."""
    print("[learned] training", flush=True)
    s = _hull(ctx)
    rep = ctx.mlp_train(s, samples=2_000_000, epochs=3, validation=500_000)
    print(f"[learned] report {rep}", flush=True)
    assert rep["margin"] > 0 and rep["rms_err"] < 0.02 * rep["scale"]
    pv = ctx.place(ctx.logical_volume(s))
    mn, mx = ctx.extent(s)
    half, cen = 0.5 * (mx - mn), 0.5 * (mx + mn)
    x, y, z = sampling.points_in_box(99, 7, 400_000, half, 1.3, cen)
    P = vg.SOA3D(x, y, z)
    inside = ctx.contains(pv, P).astype(bool)
    exact = np.where(inside, ctx.safety_to_out(pv, P), ctx.safety_to_in(pv, P))
    bound = ctx.mlp_bound(s, x, y, z)
    viol = int(np.sum(bound > exact))
    useful = float(np.mean(bound > 0.5 * exact))
    print(f"[learned] violations {viol}, bound above half the exact safety on {useful:.1%}", flush=True)
    assert viol == 0
    assert useful > 0.5


def _batch(ctx, total, seed=13):
    import torch

    fams = ("box", "tube", "polycone", "tessellated")
    specs, counts = workloads.batched_solids(total, families=fams, per_family=2, seed=seed)
    shapes, offs, steps, arr = workloads.build(ctx, specs, counts, seed=seed)
    n = int(offs[-1])
    b = vg.Batch(ctx, shapes, offs, steps)
    init = [torch.from_numpy(a.copy()).cuda() for a in arr]
    side0 = torch.zeros(n, dtype=torch.int8, device="cuda")
    torch.cuda.synchronize()
    b.classify(init[0], init[1], init[2], side0)
    tess = [int(sh) for (kind, _), sh in zip(specs, shapes) if kind == "tessellated"]
    return b, init, side0, tess, n


def _run(b, init, side0, n, nsteps, **kw):
    import torch

    st = [t.clone() for t in init]
    side = side0.clone()
    cr = torch.zeros(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.transport(*st, side, cr, nsteps, step0=0, seed=5, **kw)
    dt = time.perf_counter() - t0
    return [x.cpu().numpy() for x in st] + [side.cpu().numpy(), cr.cpu().numpy()], dt


def test_learned_transport_bit_exact_and_ablation(ctx):
    """FP64 transport vs learned transport (trained networks on the
    tessellated solids): identical trajectories, zero audited violations.
    Ablation: the same learned kernel with networks that never serve a bound
    (every step needing a query goes to the exact fallback)."""
    print("[learned] transport", flush=True)
    b, init, side0, tess, n = _batch(ctx, 300_000)
    for s in tess:
        ctx.mlp_train(s, samples=1_000_000, epochs=2, validation=300_000)
    ref, t_fp64 = _run(b, init, side0, n, 40)
    got, t_learned = _run(b, init, side0, n, 40, learned=True, audit=True)
    stats = b.learned_stats()
    print(f"[learned] stats {stats}", flush=True)
    for k, (a, c) in enumerate(zip(ref, got)):
        bad = np.flatnonzero(a != c)
        assert bad.size == 0, (k, bad.size, bad[:5])
    assert stats["violations"] == 0
    assert stats["served"] > 0 and stats["queries"] >= stats["served"]
    # ablation: a network whose bound is never positive
    saved = {s: ctx.mlp_get(s) for s in tess}
    for s in tess:
        p, g = saved[s]
        g2 = g.copy()
        g2[7] = 1e30
        ctx.mlp_set(s, np.zeros_like(p), g2)
    got0, t_null = _run(b, init, side0, n, 40, learned=True)
    stats0 = b.learned_stats()
    print(f"[learned] ablation stats {stats0}; wall fp64 {t_fp64:.3f}s learned {t_learned:.3f}s "
          f"null network {t_null:.3f}s", flush=True)
    for a, c in zip(ref, got0):
        assert np.array_equal(a, c)
    assert stats0["served"] == 0
    assert stats0["fallbacks"] > stats["fallbacks"]
    for s in tess:
        ctx.mlp_set(s, *saved[s])
    b.close()


@pytest.mark.parametrize("mode", ["learned", "fp64"])
def test_graph_captured_transport(ctx, mode):
    """VGM_GRAPH: the transport call (learned families: two kernels per
    track-step) captured once and replayed.  Four launches of 10 track-steps,
    step0 advancing; eager library and torch work with synchronisations
    between the replays; after every replay the state equals the eager
    (uncaptured) run of the same launches bit for bit."""
    import torch

    print(f"[learned] graph-captured transport mode={mode}", flush=True)
    b, init, side0, tess, n = _batch(ctx, 200_000, seed=17)
    kw = {}
    if mode == "learned":
        for s in tess:
            ctx.mlp_train(s, samples=500_000, epochs=1, validation=200_000)
        kw = {"learned": True}
    runs = []
    for graph in (False, True):
        st = [t.clone() for t in init]
        side = side0.clone()
        cr = torch.zeros(n, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        snaps = []
        for k in range(4):
            b.transport(*st, side, cr, 10, step0=10 * k, seed=5, graph=graph, **kw)
            # eager work between replays
            tmp = torch.zeros(n, dtype=torch.int8, device="cuda")
            b.classify(st[0], st[1], st[2], tmp)
            assert int(tmp.sum().item()) >= 0
            torch.cuda.synchronize()
            snaps.append([x.cpu().numpy().copy() for x in st + [side, cr]])
        runs.append(snaps)
    for k, (e, g) in enumerate(zip(*runs)):
        for j, (a, c) in enumerate(zip(e, g)):
            bad = np.flatnonzero(a != c)
            assert bad.size == 0, (k, j, bad.size)
    b.close()


def test_save_load_roundtrip(ctx):
    s = ctx.box(2.0, 3.0, 4.0)
    p = mlp.random_params(9).astype(np.float32)
    g = np.array([0.1, -0.2, 0.3, 2.6, 3.9, 5.2, 10.77, 0.05])
    ctx.mlp_set(s, p, g)
    p2, g2 = ctx.mlp_get(s)
    assert np.array_equal(p, p2) and np.array_equal(g, g2)
    with pytest.raises(vg.VgmError):
        ctx.mlp_get(ctx.box(1.0, 1.0, 1.0))
