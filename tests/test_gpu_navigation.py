"""GPU navigation vs the CPU oracle on the barrel geometry, bit-exact:
located states, one step, 200-step transport and safeties."""
import numpy as np
import pytest

import oracle
import vecgeom_amd as vg
from vecgeom_amd import workloads

pytestmark = pytest.mark.gpu

N = 50000


@pytest.fixture(scope="module")
def setup():
    import torch

    ctx = vg.Context(0)
    w = workloads.build_barrel(ctx)
    nav = vg.Navigation(ctx, w)
    geo = oracle.Geometry()
    wo = workloads.build_barrel(geo)
    assert geo.nav_close(wo, use_bvh=True) == nav.nnodes
    pts, dirs = workloads.barrel_tracks(N, seed=9)
    yield ctx, nav, geo, pts, dirs
    ctx.close()


def test_locate(setup):
    print("[nav] locate", flush=True)
    ctx, nav, geo, pts, dirs = setup
    st = np.zeros(N, np.uint32)
    nav.locate(*pts, st)
    assert np.array_equal(st, geo.nav_locate(pts))


def test_step_and_safety(setup):
    print("[nav] step + safety", flush=True)
    ctx, nav, geo, pts, dirs = setup
    st = geo.nav_locate(pts)
    saf = np.zeros(N)
    nav.safety(*pts, st, saf)
    assert np.array_equal(saf, geo.nav_safety(pts, st))
    s_g, l_g, b_g = st.copy(), np.zeros(N, np.uint32), np.zeros(N, np.uint8)
    step = np.zeros(N)
    nav.step(*pts, *dirs, s_g, l_g, b_g, step)
    s_o, l_o, b_o = st.copy(), np.zeros(N, np.uint32), np.zeros(N, np.uint8)
    ref = geo.nav_step(pts, dirs, s_o, l_o, b_o)
    assert np.array_equal(step, ref)
    assert np.array_equal(s_g, s_o) and np.array_equal(l_g, l_o) and np.array_equal(b_g, b_o)


@pytest.mark.parametrize("limited", [False, True], ids=["unlimited", "limit5mm"])
def test_compute_step_without_relocation(setup, limited):
    """ComputeStep (VNavigator.h:494-540) bit-exact against the oracle's,
    and ComputeStepAndNextVolume identical to ComputeStepAndPropagatedState.
    Without relocation the state after a mother exit stays the current one,
    so the states differ from the relocated step exactly on the boundary
    steps that leave a volume."""
    print(f"[nav] compute step limited={limited}", flush=True)
    ctx, nav, geo, pts, dirs = setup
    st = geo.nav_locate(pts)
    lim = np.full(N, 5.0) if limited else None
    s_g, l_g, b_g = st.copy(), np.zeros(N, np.uint32), np.zeros(N, np.uint8)
    step = np.zeros(N)
    nav.compute_step(*pts, *dirs, s_g, l_g, b_g, step, step_limit=lim)
    s_o, l_o, b_o = st.copy(), np.zeros(N, np.uint32), np.zeros(N, np.uint8)
    ref = geo.nav_compute_step(pts, dirs, s_o, l_o, b_o, step_limit=lim)
    assert np.array_equal(step, ref)
    assert np.array_equal(s_g, s_o) and np.array_equal(l_g, l_o) and np.array_equal(b_g, b_o)
    # the relocating step: same step lengths, and the same state wherever no
    # volume is left (daughter entries keep the pushed daughter on both)
    s_r, l_r, b_r = st.copy(), np.zeros(N, np.uint32), np.zeros(N, np.uint8)
    step_r = np.zeros(N)
    nav.compute_step_and_next_volume(*pts, *dirs, s_r, l_r, b_r, step_r, step_limit=lim)
    s_p, l_p, b_p = st.copy(), np.zeros(N, np.uint32), np.zeros(N, np.uint8)
    step_p = np.zeros(N)
    nav.step(*pts, *dirs, s_p, l_p, b_p, step_p, step_limit=lim)
    assert np.array_equal(step_r, step_p) and np.array_equal(s_r, s_p) and np.array_equal(l_r, l_p)
    assert np.array_equal(step_r, step)
    # physics-limited steps relocate nothing: both keep the state
    off = b_g == 0
    assert np.array_equal(s_r[off], s_g[off]) and np.array_equal(s_g[off], st[off])
    # boundary steps: ComputeStep either keeps the state (the mother is left)
    # or pushes a daughter of it, with the current state as the last exited
    on = ~off & (st != 0)
    assert np.array_equal(l_g[on], st[on])
    entered = on & (s_g != st)
    mothers = {int(s): nav.node(int(s))[2] for s in np.unique(s_g[entered])}
    assert all(mothers[int(s)] == int(m) for s, m in zip(s_g[entered], st[entered]))
    if limited:
        assert off.sum() > 0 and np.all(step[off] == 5.0)


@pytest.mark.parametrize("mask", ["all", "half"])
def test_step_with_safeties(setup, mask):
    """ComputeStepsAndSafetiesAndPropagatedStates: the safety of the start
    point (where asked, else 0) and the step, both equal to the oracle's
    separate safety and step."""
    print(f"[nav] step with safeties mask={mask}", flush=True)
    ctx, nav, geo, pts, dirs = setup
    st = geo.nav_locate(pts)
    calc = None if mask == "all" else (np.arange(N) % 2 == 0).astype(np.uint8)
    s_g, l_g, b_g = st.copy(), np.zeros(N, np.uint32), np.zeros(N, np.uint8)
    step, saf = np.zeros(N), np.full(N, -7.0)
    nav.step_safety(*pts, *dirs, s_g, l_g, b_g, step, saf, calcsafety=calc)
    ref_saf = geo.nav_safety(pts, st)
    if calc is not None:
        ref_saf = np.where(calc != 0, ref_saf, 0.0)
    assert np.array_equal(saf, ref_saf)
    s_o, l_o, b_o = st.copy(), np.zeros(N, np.uint32), np.zeros(N, np.uint8)
    ref = geo.nav_step(pts, dirs, s_o, l_o, b_o)
    assert np.array_equal(step, ref)
    assert np.array_equal(s_g, s_o) and np.array_equal(l_g, l_o) and np.array_equal(b_g, b_o)


def test_state_path_round_trip(setup):
    """NavStatePath <-> state index for every node of the placement tree:
    the path runs from the world down the mother chain, and converts back to
    the same index; a path that is not in the tree is an error."""
    print("[nav] state <-> path", flush=True)
    ctx, nav, geo, pts, dirs = setup
    world = nav.state_to_path(1)
    assert len(world) == 1
    for s in range(1, nav.nnodes + 1):
        path = nav.state_to_path(s)
        pv, level, mother = nav.node(s)
        assert len(path) == level + 1 and path[-1] == pv and path[0] == world[0]
        if level > 0:
            assert nav.state_to_path(mother) == path[:-1]
        assert nav.path_to_state(path) == s
    # a module directly under the world skips its layer: not in the tree
    layer_path = nav.state_to_path(nav.path_to_state(nav.state_to_path(2)))
    deep = next(s for s in range(1, nav.nnodes + 1) if nav.node(s)[1] == 2)
    with pytest.raises(vg.VgmError):
        nav.path_to_state([world[0], nav.node(deep)[0]])
    with pytest.raises(vg.VgmError):
        nav.path_to_state([layer_path[-1]])
    # located states: the path's top volume is the state's placed volume
    st = geo.nav_locate([a[:200] for a in pts])
    for s in st[st > 0]:
        assert nav.path_to_state(nav.state_to_path(int(s))) == int(s)


@pytest.mark.parametrize("device_ptrs", [False, True], ids=["host", "device"])
def test_transport_200_steps(setup, device_ptrs):
    print(f"[nav] transport device_ptrs={device_ptrs}", flush=True)
    import torch

    ctx, nav, geo, pts, dirs = setup
    st = geo.nav_locate(pts)
    ref = [a.copy() for a in pts]
    s_o, l_o, b_o = st.copy(), np.zeros(N, np.uint32), np.zeros(N, np.uint8)
    n_o = np.zeros(N, np.int32)
    geo.nav_transport(*ref, *dirs, s_o, l_o, b_o, 200, n_o)
    arrs = [a.copy() for a in pts] + [a.copy() for a in dirs] + [st.copy(), np.zeros(N, np.uint32),
                                                                    np.zeros(N, np.uint8), np.zeros(N, np.int32)]
    if device_ptrs:
        arrs = [torch.from_numpy(a).cuda() for a in arrs]
        torch.cuda.synchronize()
    nav.transport(*arrs[:6], arrs[6], arrs[7], arrs[8], 200, arrs[9])
    got = [a.cpu().numpy() if device_ptrs else a for a in arrs]
    for k, r in enumerate(ref):
        assert np.array_equal(got[k], r), k
    assert np.array_equal(got[6], s_o) and np.array_equal(got[7], l_o) and np.array_equal(got[8], b_o)
    assert np.array_equal(got[9], n_o)
    assert n_o.sum() > 20 * N


def test_graph_captured_step_loop(setup):
    """The step loop as one-step kernels launched one by one and as a
    captured HIP graph: both equal the persistent transport and the oracle bit
    for bit.  The graph is replayed three times on the same arrays (each
    replay continues the transport by 60 steps), with eager library and torch
    work and synchronisations between the replays, and every replay is
    checked against the eager one-kernel transport of the same steps."""
    print("[nav] graph-captured step loop", flush=True)
    import torch

    ctx, nav, geo, pts, dirs = setup
    st = geo.nav_locate(pts)

    def fresh():
        a = [a.copy() for a in pts] + [a.copy() for a in dirs] + [st.copy(), np.zeros(N, np.uint32),
                                                                  np.zeros(N, np.uint8), np.zeros(N, np.int32)]
        out = [torch.from_numpy(x).cuda() for x in a]
        torch.cuda.synchronize()
        return out

    def state(arrs):
        return [x.cpu().numpy().copy() for x in arrs]

    eager, graph, kern = fresh(), fresh(), fresh()
    ref = [a.copy() for a in pts]
    s_o, l_o, b_o, n_o = st.copy(), np.zeros(N, np.uint32), np.zeros(N, np.uint8), np.zeros(N, np.int32)
    for rep in range(3):
        nav.transport(*eager[:6], eager[6], eager[7], eager[8], 60, eager[9])
        nav.transport(*graph[:6], graph[6], graph[7], graph[8], 60, graph[9], graph=True)
        nav.transport(*kern[:6], kern[6], kern[7], kern[8], 60, kern[9], step_kernels=True)
        geo.nav_transport(*ref, *dirs, s_o, l_o, b_o, 60, n_o)
        # eager work between replays: a locate and a torch reduction on other arrays
        tmp = torch.zeros(N, dtype=torch.int32, device="cuda")
        nav.locate(*[torch.from_numpy(a).cuda() for a in pts], tmp)
        assert int((tmp > 0).sum().item()) > 0
        torch.cuda.synchronize()
        e, g, k = state(eager), state(graph), state(kern)
        for j in range(10):
            assert np.array_equal(g[j], e[j]), (rep, j)
            assert np.array_equal(k[j], e[j]), (rep, j)
        for j in range(3):
            assert np.array_equal(e[j], ref[j]), (rep, j)
        assert np.array_equal(e[6], s_o) and np.array_equal(e[9], n_o), rep


def test_approach_next_daughter(setup):
    print("[nav] approach", flush=True)
    ctx, nav, geo, pts, dirs = setup
    st = geo.nav_locate(pts)
    big = np.finfo(np.float64).max
    # from the located points, then from the boundary after one step (last set)
    s_o, l_o, b_o = st.copy(), np.zeros(N, np.uint32), np.zeros(N, np.uint8)
    stp = geo.nav_step(pts, dirs, s_o, l_o, b_o)
    moved = tuple(np.ascontiguousarray(pts[k] + stp * dirs[k]) for k in range(3))
    for p, s, last in ((pts, st, np.zeros(N, np.uint32)), (moved, s_o, l_o)):
        ref = np.full(N, big)
        geo.nav_approach(p, dirs, s, last, ref)
        got = np.full(N, big)
        nav.approach(*p, *dirs, s.copy(), last.copy(), got)
        assert np.array_equal(got, ref)
        assert np.mean(ref < big) > 0.2


@pytest.mark.parametrize("reuse", [False, True], ids=["fresh", "carried"])
@pytest.mark.parametrize("L", [2.0, 40.0])
def test_transport_limited(setup, L, reuse):
    print(f"[nav] limited transport L={L} reuse={reuse}", flush=True)
    ctx, nav, geo, pts, dirs = setup
    st = geo.nav_locate(pts)
    lim = np.full(N, L)
    outs = []
    for run in (geo.nav_transport_limited, nav.transport_limited):
        x, y, z = (a.copy() for a in pts)
        s, last, b = st.copy(), np.zeros(N, np.uint32), np.zeros(N, np.uint8)
        steps, geo_steps = np.zeros(N, np.int32), np.zeros(N, np.int32)
        run(x, y, z, *dirs, lim, s, last, b, 100, steps, geo_steps, reuse_safety=reuse)
        outs.append((x, y, z, s, last, b, steps, geo_steps))
    for k, (r, g) in enumerate(zip(*outs)):
        assert np.array_equal(g, r), k
    assert outs[0][7].sum() < outs[0][6].sum()


@pytest.mark.parametrize("directional", [False, True])
def test_level_locate(setup, directional):
    """LevelLocate in the world and in the layers (placed at the identity, so
    global points are their local ones), on random points and on the surface
    points where one step stops, with and without the direction and with an
    excluded daughter: the daughter's state and its local point bit-exact
    against the oracle."""
    print(f"[nav] level locate directional={directional}", flush=True)
    ctx, nav, geo, pts, dirs = setup
    st = geo.nav_locate(pts)
    s, last, bnd = st.copy(), np.zeros(N, np.uint32), np.zeros(N, np.uint8)
    step = geo.nav_step(pts, dirs, s, last, bnd)
    on = tuple(np.ascontiguousarray(pts[k] + step * dirs[k]) for k in range(3))
    layer = np.array([n if geo.nav_node(int(n))[1] <= 1 else geo.nav_node(int(n))[2] for n in st], np.uint32)
    world = np.where(st != 0, 1, 0).astype(np.uint32)
    d = dirs if directional else None
    found = 0
    for p, state, excl in ((pts, world, None), (pts, layer, None), (on, layer, None), (on, layer, s)):
        ex = None if excl is None else np.ascontiguousarray(excl, dtype=np.uint32)
        c_o, q_o = geo.nav_level_locate(p, state, d, ex)
        c_g = np.zeros(N, np.uint32)
        q_g = [np.zeros(N) for _ in range(3)]
        nav.level_locate(*p, state, c_g, *q_g, dirs=d, excl=ex)
        assert np.array_equal(c_g, c_o)
        for k in range(3):
            assert np.array_equal(q_g[k], q_o[k])
        found += int(np.count_nonzero(c_g))
    assert found > N
