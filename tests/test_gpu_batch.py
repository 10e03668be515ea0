"""GPU parity of the batched path (include/nanomap_amd_batch.h) against
separate plain launches of the same frames.

32 agents with uneven image sizes, ranges and voxel sizes in one
100x100x10 m world, each agent with its own map.  A batch casts all agents'
frames in one graph-captured launch sequence; the plain path (itself
bit-exact against the oracle, test_gpu_cast.py) processes the same frames
one by one.  Maps must agree bit for bit (leaf set, FP32 log-odds bit
patterns, active masks), and so must the per-frame statistics.
"""
import numpy as np
import pytest
import torch

from nanomap_amd import workloads as W

pytestmark = pytest.mark.gpu

RD, RN = 1, 0
STAT_KEYS = ("points", "rays", "visits", "voxels", "leaves", "hit_voxels",
             "mixed_voxels", "map_leaves", "errors")


def _leaf_capacity(res):
    return {0.05: 1 << 19, 0.1: 1 << 17}.get(res, 1 << 16)


def _maps(nm, agents):
    return [nm.Map(a["res"], *W.CLAMP, leaf_capacity=_leaf_capacity(a["res"]))
            for a in agents]


def _assert_same_map(a, b, k):
    ao, av, am = a.download()
    bo, bv, bm = b.download()
    assert ao.shape == bo.shape, (k, ao.shape, bo.shape)
    assert np.array_equal(ao, bo), k
    assert np.array_equal(av.view(np.uint32), bv.view(np.uint32)), k
    assert np.array_equal(am, bm), k


def _plain(nm, ctx, agents, clouds, maps, rounding):
    stats = []
    for a, c, m in zip(agents, clouds, maps):
        ctx.submit(m, W.agent_frame(nm, a, rounding), c)
        stats.append(ctx.stats())
    return stats


@pytest.fixture(scope="module")
def agents(gpu):
    nm = gpu
    scene, a0 = W.agents_32(nm, frame_index=0)
    _, a1 = W.agents_32(nm, scene=scene, frame_index=1)
    return nm, a0, a1


def test_agents_32_batch_bit_exact_vs_plain(agents):
    nm, a0, a1 = agents
    P = max(a["cloud"].shape[0] for a in a0)
    T = max(a["range"] / a["res"] for a in a0) + 1
    ctx = nm.Context(P, T)
    batch = nm.Batch(32, sum(a["cloud"].shape[0] for a in a0), T, 1 << 21)
    bmaps, pmaps = _maps(nm, a0), _maps(nm, a0)
    for step, agents in enumerate((a0, a1)):
        clouds = [torch.from_numpy(a["cloud"]).cuda() for a in agents]
        print("batch step %d: %d agents, %d points" % (
            step, len(agents), sum(c.shape[0] for c in clouds)), flush=True)
        batch.submit([(W.agent_frame(nm, a, RD), c, m)
                      for a, c, m in zip(agents, clouds, bmaps)])
        bstats = batch.stats()
        pstats = _plain(nm, ctx, agents, clouds, pmaps, RD)
        # an agent whose pixels are all capped at its range (open space)
        # casts no ray (Handler.cpp:71); the batch as a whole must not
        assert sum(bs["rays"] for bs in bstats) > 10 ** 6
        for k, (bs, ps) in enumerate(zip(bstats, pstats)):
            assert bs["overflow"] == 0 and ps["overflow"] == 0
            assert bs["errors"] == 0
            assert bs["rays"] == 0 or bs["map_leaves"] > 0, (step, k, bs)
            for key in STAT_KEYS:
                assert bs[key] == ps[key], (step, k, key, bs[key], ps[key])
    for k, (bm, pm) in enumerate(zip(bmaps, pmaps)):
        _assert_same_map(bm, pm, k)


def _subset(agents, n=8):
    """The n agents with the fewest points (fast repeated runs)."""
    order = sorted(range(len(agents)), key=lambda k: agents[k]["cloud"].shape[0])
    return [agents[k] for k in order[:n]]


@pytest.mark.parametrize("rounding", [RD, RN])
def test_every_tuning_and_graph_mode_agree(agents, rounding):
    nm, a0, _ = agents
    agents = _subset(a0, 10)
    clouds = [torch.from_numpy(a["cloud"]).cuda() for a in agents]
    P = max(c.shape[0] for c in clouds)
    T = max(a["range"] / a["res"] for a in agents) + 1
    ctx = nm.Context(P, T)
    pmaps = _maps(nm, agents)
    _plain(nm, ctx, agents, clouds, pmaps, rounding)
    batch = nm.Batch(len(agents), sum(c.shape[0] for c in clouds), T, 1 << 19)
    for layout in range(4):
        for queue, gm, lds in ((0, 4, 4096), (1, 2, 2048), (2, 8, 4096),
                               (0, 4, 2048), (3, 6, 4096), (3, 4, 2048)):
            for graph in (True, False):
                print("tuning layout=%d queue=%d grid_mult=%d lds=%d graph=%d"
                      % (layout, queue, gm, lds, graph), flush=True)
                batch.set_tuning(layout, queue, gm, lds)
                batch.set_graph(graph)
                bmaps = _maps(nm, agents)
                batch.submit([(W.agent_frame(nm, a, rounding), c, m)
                              for a, c, m in zip(agents, clouds, bmaps)])
                torch.cuda.synchronize()
                for k, (bm, pm) in enumerate(zip(bmaps, pmaps)):
                    _assert_same_map(bm, pm, k)
                for m in bmaps:
                    m.close()


def test_graph_replays_with_eager_work_between(agents):
    """The captured graph is replayed for new frames of the same shape, with
    unrelated eager work and host synchronizations between replays; every
    replay must match the plain path."""
    nm, a0, a1 = agents
    agents0, agents1 = _subset(a0, 6), None
    ids = [a0.index(a) for a in agents0]
    agents1 = [a1[k] for k in ids]
    T = max(a["range"] / a["res"] for a in agents0) + 1
    P = max(a["cloud"].shape[0] for a in agents0 + agents1)
    ctx = nm.Context(P, T)
    batch = nm.Batch(6, 2 * sum(a["cloud"].shape[0] for a in agents0), T,
                     1 << 19)
    bmaps, pmaps = _maps(nm, agents0), _maps(nm, agents0)
    for rep, agents in enumerate((agents0, agents1, agents0, agents1)):
        print("replay %d" % rep, flush=True)
        clouds = [torch.from_numpy(a["cloud"]).cuda() for a in agents]
        batch.submit([(W.agent_frame(nm, a, RD), c, m)
                      for a, c, m in zip(agents, clouds, bmaps)])
        x = torch.randn(1 << 20, device="cuda")
        y = (x * 2).sum().item()  # eager work + a host synchronization
        assert np.isfinite(y)
        _plain(nm, ctx, agents, clouds, pmaps, RD)
        torch.cuda.synchronize()
        for k, (bm, pm) in enumerate(zip(bmaps, pmaps)):
            _assert_same_map(bm, pm, k)


def test_autotuner_picks_a_valid_tuning(agents):
    nm, a0, _ = agents
    agents = _subset(a0, 8)
    clouds = [torch.from_numpy(a["cloud"]).cuda() for a in agents]
    T = max(a["range"] / a["res"] for a in agents) + 1
    batch = nm.Batch(8, sum(c.shape[0] for c in clouds), T, 1 << 19)
    maps = _maps(nm, agents)
    frames = [(W.agent_frame(nm, a, RD), c, m)
              for a, c, m in zip(agents, clouds, maps)]
    best, ms = batch.tune(frames, reps=1)
    print("tuned", best, ms, flush=True)
    assert 0 <= best["layout"] < 4 and 0 <= best["queue"] <= 3 and ms > 0
    assert best["lds_slots"] in (2048, 4096)
    assert batch.get_tuning() == best
    # tuning does not touch the maps
    assert all(m.leaf_count() == 0 for m in maps)
    batch.submit(frames)
    ctx = nm.Context(max(c.shape[0] for c in clouds), T)
    pmaps = _maps(nm, agents)
    _plain(nm, ctx, agents, clouds, pmaps, RD)
    for k, (bm, pm) in enumerate(zip(maps, pmaps)):
        _assert_same_map(bm, pm, k)


def test_batch_rejects_bad_input(agents):
    nm, a0, _ = agents
    a = _subset(a0, 2)
    clouds = [torch.from_numpy(x["cloud"]).cuda() for x in a]
    batch = nm.Batch(2, 10 ** 6, 400.0, 1 << 16)
    m = nm.Map(a[0]["res"], *W.CLAMP, leaf_capacity=1 << 12)
    f0 = W.agent_frame(nm, a[0], RD)
    f1 = W.agent_frame(nm, a[1], RD)
    with pytest.raises(nm.NanoMapError):  # one map twice
        batch.submit([(f0, clouds[0], m), (f1, clouds[1], m)])
    with pytest.raises(nm.NanoMapError):  # host cloud
        batch.submit([(f0, a[0]["cloud"], m)])
    fv = W.agent_frame(nm, a[0], RD)
    fv.filter = nm.FILTER_VOXEL
    with pytest.raises(nm.NanoMapError):  # voxel filter not batched
        batch.submit([(fv, clouds[0], m)])
    f2 = W.agent_frame(nm, a[1], RN)
    m2 = nm.Map(a[1]["res"], *W.CLAMP, leaf_capacity=1 << 12)
    with pytest.raises(nm.NanoMapError):  # mixed rounding modes
        batch.submit([(f0, clouds[0], m), (f2, clouds[1], m2)])
    with pytest.raises(nm.NanoMapError):  # stats before any batch
        batch.stats()


def test_batch_with_per_agent_log_odds(agents):
    """Agents with different hit/miss log-odds in one batch: each job's
    table of miss-only sums and its heavy-voxel sums use its own values."""
    nm, a0, _ = agents
    agents = _subset(a0, 10)
    clouds = [torch.from_numpy(a["cloud"]).cuda() for a in agents]
    P = max(c.shape[0] for c in clouds)
    T = max(a["range"] / a["res"] for a in agents) + 1
    odds = [(0.85, -0.4), (0.7, -0.3), (1.2, -0.55), (0.4, -0.9)]

    def frame(k, a):
        f = W.agent_frame(nm, a, RD)
        f.logodds_hit, f.logodds_miss = odds[k % len(odds)]
        return f

    ctx = nm.Context(P, T)
    pmaps = _maps(nm, agents)
    for k, (a, c, m) in enumerate(zip(agents, clouds, pmaps)):
        ctx.submit(m, frame(k, a), c)
    batch = nm.Batch(len(agents), sum(c.shape[0] for c in clouds), T, 1 << 19)
    bmaps = _maps(nm, agents)
    batch.submit([(frame(k, a), c, m)
                  for k, (a, c, m) in enumerate(zip(agents, clouds, bmaps))])
    torch.cuda.synchronize()
    for k, (bm, pm) in enumerate(zip(bmaps, pmaps)):
        _assert_same_map(bm, pm, k)
