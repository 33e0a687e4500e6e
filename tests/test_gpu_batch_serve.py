"""GPU tests of the batched and served forms: every simulation's result is
bit-identical to the same simulation run alone (oracle trajectory = solo drop-in
trajectory, which test_gpu_dropin pins), whatever else shares the launches."""
import ctypes as C

import numpy as np
import pytest

from galaxy_amd import Config, NbodyGpu, ics
from galaxy_amd.batch import Batch
from galaxy_amd.serve import Server

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    g = NbodyGpu()
    yield g
    g.close()


def ocfg(oracle, c: Config):
    return oracle.make_config(theta=c.theta, softening=c.softening_kpc, g=c.g, bucket=c.bucket,
                              max_depth=c.max_depth)


def mixed_sims():
    disc = Config(theta=0.5, softening_kpc=0.01, g=1.0, bucket=4, max_depth=16)
    coll = Config(theta=0.5, softening_kpc=0.005, g=1.0, bucket=4, max_depth=16)
    return [
        (ics.rotating_disc(700, 1001, scale=0.7), disc, 1e-3),
        (ics.uniform_collapse(5000, 1002), coll, 5e-4),
        (ics.rotating_disc(33, 1003), disc, 1e-3),
        (ics.rotating_disc(2049, 1004, scale=1.8), Config(theta=0.6, softening_kpc=0.01, g=1.0,
                                                          bucket=1, max_depth=16), 1e-3),
        (ics.uniform_collapse(64, 1005), coll, 5e-4),
    ]


@pytest.mark.parametrize("use_graph,sort_tile,block,waves,bucket", [
    (True, 2048, 256, 32, None), (False, 1024, 256, 32, None), (True, 1024, 64, 16, 8),
    (True, 2048, 128, 24, 16)])
def test_batch_equals_solo(oracle, gpu, use_graph, sort_tile, block, waves, bucket):
    """Every launch shape the autotuner can pick, and other leaf sizes, give the
    oracle's trajectory for every simulation of the batch."""
    sims = mixed_sims()
    if bucket is not None:
        sims = [(s, Config(c.theta, c.softening_kpc, c.g, bucket, c.max_depth), dt) for s, c, dt in sims]
    print(f"batch graph={use_graph} sort_tile={sort_tile} block={block} waves={waves} bucket={bucket}",
          flush=True)
    b = Batch(gpu, [s for s, _, _ in sims], [c for _, c, _ in sims], [dt for _, _, dt in sims],
              use_graph=use_graph, sort_tile=sort_tile, traverse_block=block, traverse_waves_per_cu=waves)
    b.step(2)
    b.step(1)
    for i, (s0, cfg, dt) in enumerate(sims):
        s, a = b.read(i)
        rs, ra = oracle.f32_evolve(s0, 3, dt, ocfg(oracle, cfg))
        assert np.array_equal(s.view(np.uint32), rs.view(np.uint32)), i
        assert np.array_equal(a.view(np.uint32), ra.view(np.uint32)), i
    st, hsh = b.force_evidence(1)
    t = oracle.tree_from_state(b.read(1)[0], ocfg(oracle, sims[1][1]))
    _, ostats, ohash = oracle.traverse_f32(*t.pack(), evidence=True)
    assert np.array_equal(st, ostats) and np.array_equal(hsh, ohash)
    b.close()


@pytest.mark.parametrize("topo_tile,streams", [(512, 1), (1024, 1), (1024, 2)])
def test_batch_mixed_leaf_sizes(oracle, gpu, topo_tile, streams):
    """Leaf sizes on both sides of 32 in one batch: simulations with small leaves
    have their columns counted in the boundary pass and their tiles' places in the
    preorder from the scan, those with larger leaves take a ticket and a decoupled
    look-back in the column pass, all in the same launches, at both topology tile
    sizes.  Every simulation follows the oracle's trajectory and evidence."""
    sims = [(ics.rotating_disc(5000, 3001, scale=0.8), 1), (ics.uniform_collapse(3000, 3002), 48),
            (ics.rotating_disc(2500, 3003, scale=1.5), 4), (ics.uniform_collapse(4100, 3004), 200),
            (ics.rotating_disc(1500, 3005), 32), (ics.rotating_disc(2600, 3006, scale=0.6), 33)]
    cfgs = [Config(theta=0.5, softening_kpc=0.01, g=1.0, bucket=bk, max_depth=16) for _, bk in sims]
    print(f"mixed leaf sizes topo_tile={topo_tile} streams={streams}", flush=True)
    b = Batch(gpu, [s for s, _ in sims], cfgs, 1e-3, use_graph=True, topo_tile=topo_tile, streams=streams)
    b.step(2)
    for i, ((s0, _), cfg) in enumerate(zip(sims, cfgs)):
        s, a = b.read(i)
        rs, ra = oracle.f32_evolve(s0, 2, 1e-3, ocfg(oracle, cfg))
        assert np.array_equal(s.view(np.uint32), rs.view(np.uint32)), i
        assert np.array_equal(a.view(np.uint32), ra.view(np.uint32)), i
        st, hsh = b.force_evidence(i)
        t = oracle.tree_from_state(s, ocfg(oracle, cfg))
        _, ostats, ohash = oracle.traverse_f32(*t.pack(), evidence=True)
        assert np.array_equal(st, ostats) and np.array_equal(hsh, ohash), i
    b.close()


def test_batch_two_target_walks_with_and_without_containment_test(oracle, gpu):
    """One persistent wavefront per CU, so that the set has items enough for two
    targets per lane.  Simulations whose thresholds reject every cell holding its
    target (theta <= 2/3) and whose self terms vanish walk without a containment
    test (with G = 1 and G = 2, and at theta 0.66, where the margin between a
    containing cell's d2 <= 2 w^2 and its threshold is smallest); theta 0.75, and a
    softening of 1e-15 whose self term m / eps^3 overflows, keep the test.  Every
    trajectory equals the oracle's."""
    sims = [(ics.rotating_disc(30000, 4001), Config(theta=0.5, softening_kpc=0.01, g=1.0, bucket=4, max_depth=16)),
            (ics.rotating_disc(30000, 4002, scale=1.4),
             Config(theta=0.5, softening_kpc=0.01, g=2.0, bucket=4, max_depth=16)),
            (ics.uniform_collapse(20000, 4003), Config(theta=0.75, softening_kpc=0.005, g=1.0, bucket=4, max_depth=16)),
            (ics.uniform_collapse(20000, 4005), Config(theta=0.66, softening_kpc=0.005, g=1.0, bucket=4, max_depth=16)),
            (ics.rotating_disc(20000, 4004, scale=0.8),
             Config(theta=0.5, softening_kpc=1e-15, g=1.0, bucket=4, max_depth=16))]
    print("two-target walks: theta 0.5 (G 1, G 2), theta 0.75, theta 0.66, softening 1e-15", flush=True)
    b = Batch(gpu, [s for s, _ in sims], [c for _, c in sims], 1e-3, use_graph=True, traverse_block=64,
              traverse_waves_per_cu=1)
    b.step(2)
    for i, (s0, cfg) in enumerate(sims):
        s, a = b.read(i)
        rs, ra = oracle.f32_evolve(s0, 2, 1e-3, ocfg(oracle, cfg))
        assert np.array_equal(s.view(np.uint32), rs.view(np.uint32)), i
        assert np.array_equal(a.view(np.uint32), ra.view(np.uint32)), i
    b.close()


def paired_disc(n, seed):
    """Bodies in tight pairs (1e-5 apart): with one body per leaf every pair hangs
    from a chain of single-child cells, about 5.4 cells per position."""
    rng = np.random.default_rng(seed)
    base = ics.rotating_disc(n // 2, seed)
    s = np.zeros((2 * (n // 2), 8), np.float32)
    s[0::2] = base
    s[1::2] = base
    s[1::2, 0] += np.float32(1e-5) * rng.standard_normal(n // 2).astype(np.float32)
    s[1::2, 1] += np.float32(1e-5) * rng.standard_normal(n // 2).astype(np.float32)
    return s


@pytest.mark.parametrize("topo_tile", [512, 1024])
def test_batch_dense_trees_take_the_global_link_path(oracle, gpu, topo_tile):
    """Tiles with more than 2 cells per position exceed the link kernel's LDS stage
    and run its column-per-lane global-memory rounds, beside simulations whose
    tiles fit the stage; every simulation follows the oracle's trajectory and
    evidence."""
    c1 = Config(theta=0.5, softening_kpc=0.01, g=1.0, bucket=1, max_depth=16)
    c4 = Config(theta=0.5, softening_kpc=0.01, g=1.0, bucket=4, max_depth=16)
    sims = [(paired_disc(3000, 4001), c1), (ics.rotating_disc(2500, 4002), c1), (paired_disc(1800, 4003), c4),
            (ics.rotating_disc(4000, 4004, scale=0.8), c4)]
    print(f"dense trees topo_tile={topo_tile}", flush=True)
    b = Batch(gpu, [s for s, _ in sims], [c for _, c in sims], 1e-3, use_graph=True, topo_tile=topo_tile)
    b.step(2)
    for i, (s0, cfg) in enumerate(sims):
        s, a = b.read(i)
        rs, ra = oracle.f32_evolve(s0, 2, 1e-3, ocfg(oracle, cfg))
        assert np.array_equal(s.view(np.uint32), rs.view(np.uint32)), i
        assert np.array_equal(a.view(np.uint32), ra.view(np.uint32)), i
        st, hsh = b.force_evidence(i)
        t = oracle.tree_from_state(s, ocfg(oracle, cfg))
        _, ostats, ohash = oracle.traverse_f32(*t.pack(), evidence=True)
        assert np.array_equal(st, ostats) and np.array_equal(hsh, ohash), i
    b.close()


@pytest.mark.parametrize("n,use_graph", [(2, True), (700, True), (1024, False), (1025, True)])
def test_batch_of_one_small_simulation(oracle, gpu, n, use_graph):
    """A batch of one simulation of at most 1,024 bodies builds in one launch per
    step, kick/drift and final kick included; 1,025 bodies take the tiled
    kernels.  Both give the oracle's trajectory, and so does the per-phase step
    that times the phases (separate launches)."""
    cfg = Config(theta=0.5, softening_kpc=0.01, g=1.0, bucket=4, max_depth=16)
    s0 = ics.rotating_disc(n, 2000 + n, scale=0.7)
    print(f"single-simulation batch n={n} graph={use_graph}", flush=True)
    b = Batch(gpu, [s0], [cfg], [1e-3], use_graph=use_graph)
    b.step(3)
    b.phase_ms()  # one more step, phase by phase
    b.step(1)
    s, a = b.read(0)
    rs, ra = oracle.f32_evolve(s0, 5, 1e-3, ocfg(oracle, cfg))
    assert np.array_equal(s.view(np.uint32), rs.view(np.uint32))
    assert np.array_equal(a.view(np.uint32), ra.view(np.uint32))
    st, hsh = b.force_evidence(0)
    t = oracle.tree_from_state(s, ocfg(oracle, cfg))
    _, ostats, ohash = oracle.traverse_f32(*t.pack(), evidence=True)
    assert np.array_equal(st, ostats) and np.array_equal(hsh, ohash)
    b.close()


@pytest.mark.parametrize("use_graph,block,waves,streams", [(True, 256, 20, 1), (False, 64, 16, 1),
                                                          (True, 256, 16, 2), (False, 128, 20, 2)])
def test_batch_many_sims_per_xcd_work(oracle, gpu, use_graph, block, waves, streams):
    """Twenty simulations of uneven size: no simulation exceeds an eighth of the
    traversal items, so the items are distributed by XCD (one range and queue per
    XCD, stealing at the end).  With streams=2, two groups of ten step concurrently
    on two streams (each group holds a simulation larger than an eighth of its own
    items, so it uses the grid-wide range).  Every simulation follows the oracle's
    trajectory, and the evidence traversal matches too."""
    disc = Config(theta=0.5, softening_kpc=0.01, g=1.0, bucket=4, max_depth=16)
    sizes = [3000, 40, 2100, 700, 1500, 129, 900, 2600, 64, 1800, 333, 1200, 2900, 517, 80, 2400, 1000,
             260, 1700, 600]
    assert max((n + 63) // 64 for n in sizes) * 8 <= sum((n + 63) // 64 for n in sizes)
    states = [ics.rotating_disc(n, 5000 + n, scale=0.6 + (n % 7) * 0.2) for n in sizes]
    print(f"batch of {len(sizes)} sims graph={use_graph} block={block} waves={waves} streams={streams}",
          flush=True)
    b = Batch(gpu, states, disc, 1e-3, use_graph=use_graph, traverse_block=block, traverse_waves_per_cu=waves,
              streams=streams)
    b.step(2)
    b.step(1)
    for i, s0 in enumerate(states):
        s, a = b.read(i)
        rs, ra = oracle.f32_evolve(s0, 3, 1e-3, ocfg(oracle, disc))
        assert np.array_equal(s.view(np.uint32), rs.view(np.uint32)), i
        assert np.array_equal(a.view(np.uint32), ra.view(np.uint32)), i
    k = sizes.index(2600)
    st, hsh = b.force_evidence(k)
    t = oracle.tree_from_state(b.read(k)[0], ocfg(oracle, disc))
    _, ostats, ohash = oracle.traverse_f32(*t.pack(), evidence=True)
    assert np.array_equal(st, ostats) and np.array_equal(hsh, ohash)
    b.close()


@pytest.mark.parametrize("streams", [1, 2])
def test_server_results_same_alone_and_in_mix(oracle, gpu, streams):
    sims = mixed_sims()
    steps = [10, 30, 20, 10, 40]
    print(f"server mix streams={streams}", flush=True)
    srv = Server(gpu, hbm_budget_bytes=8 << 30, steps_per_round=10, streams=streams)
    ids = [srv.submit(s, c, dt, k) for (s, c, dt), k in zip(sims, steps)]
    done = srv.drain()
    assert sorted(d.id for d in done) == sorted(ids)
    mixed = {i: srv.take(i) for i in ids}
    for i, ((s0, cfg, dt), k) in zip(ids, zip(sims, steps)):
        rs, ra = oracle.f32_evolve(s0, k, dt, ocfg(oracle, cfg))
        s, a = mixed[i]
        assert np.array_equal(s.view(np.uint32), rs.view(np.uint32)), i
        assert np.array_equal(a.view(np.uint32), ra.view(np.uint32)), i
    # the same request alone in a fresh server
    alone = Server(gpu, hbm_budget_bytes=8 << 30, steps_per_round=10)
    j = alone.submit(sims[1][0], sims[1][1], sims[1][2], steps[1])
    alone.drain()
    s, a = alone.take(j)
    assert np.array_equal(s.view(np.uint32), mixed[ids[1]][0].view(np.uint32))


def test_server_worker_thread_continuous_serving(oracle, gpu):
    """The library's worker thread runs the rounds while requests arrive from the
    caller's thread: every result still equals the oracle's trajectory, rounds with
    an unchanged in-flight set replay HIP graphs (two long requests run many rounds
    together), and manual rounds are refused while the worker runs."""
    import time

    sims = mixed_sims()
    steps = [60, 60, 20, 10, 40]
    print("server worker thread: 5 requests, staggered", flush=True)
    srv = Server(gpu, hbm_budget_bytes=8 << 30, steps_per_round=10, streams=2)
    srv.start()
    with pytest.raises(Exception, match="worker thread"):
        srv.round()
    ids, done = [], {}
    for (s, c, dt), k in zip(sims, steps):
        ids.append(srv.submit(s, c, dt, k))
        for d in srv.poll(wait=0.002):
            done[d.id] = d
        time.sleep(0.003)
    t_end = time.time() + 60
    while len(done) < len(ids) and time.time() < t_end:
        for d in srv.poll(wait=0.05):
            done[d.id] = d
    srv.stop()
    assert sorted(done) == sorted(ids) and all(d.status == 0 for d in done.values())
    assert all(d.admit_time >= d.submit_time and d.finish_time >= d.admit_time for d in done.values())
    for rid, (s0, cfg, dt), k in zip(ids, sims, steps):
        rs, ra = oracle.f32_evolve(s0, k, dt, ocfg(oracle, cfg))
        s, a = srv.take(rid)
        assert np.array_equal(s.view(np.uint32), rs.view(np.uint32)), rid
        assert np.array_equal(a.view(np.uint32), ra.view(np.uint32)), rid
    # the worker can be restarted and drives later submissions
    rid = srv.submit(*sims[2], 10)
    srv.start()
    got = []
    t_end = time.time() + 30
    while not got and time.time() < t_end:
        got = srv.poll(wait=0.05)
    srv.stop()
    assert [d.id for d in got] == [rid]
    rs, _ = oracle.f32_evolve(sims[2][0], 10, sims[2][2], ocfg(oracle, sims[2][1]))
    assert np.array_equal(srv.take(rid)[0].view(np.uint32), rs.view(np.uint32))
    srv.close()


def test_server_budget_admission_and_staggered_arrivals(oracle, gpu):
    sims = mixed_sims()
    one = Server.request_bytes(5000)
    # budget holds the 5000-body request plus a little: requests queue behind it
    srv = Server(gpu, hbm_budget_bytes=one + Server.request_bytes(2049), steps_per_round=10)
    a = srv.submit(*sims[1], 30)
    srv.round()
    b = srv.submit(*sims[0], 10)      # arrives while `a` is in flight
    c = srv.submit(*sims[3], 10)
    q, infl, res = srv.stats()
    assert infl == 1 and q == 2
    done = srv.drain()
    assert {d.id for d in done} == {a, b, c}
    for rid, (s0, cfg, dt), k in [(a, sims[1], 30), (b, sims[0], 10), (c, sims[3], 10)]:
        rs, _ = oracle.f32_evolve(s0, k, dt, ocfg(oracle, cfg))
        assert np.array_equal(srv.take(rid)[0].view(np.uint32), rs.view(np.uint32))
    with pytest.raises(Exception, match="HBM budget"):
        Server(gpu, hbm_budget_bytes=1 << 20, steps_per_round=10).submit(*sims[1], 10)


def test_server_reuses_retired_arenas(oracle, gpu):
    """The budget admits one request at a time, so each later request runs in the
    arena a retired one left behind (same size, then a smaller size within 2x), with
    that request's data still in it; results still equal the oracle's."""
    cfg = Config(theta=0.5, softening_kpc=0.01, g=1.0, bucket=4, max_depth=16)
    srv = Server(gpu, hbm_budget_bytes=Server.request_bytes(6000) + 1, steps_per_round=10)
    for n, seed, k in [(6000, 11, 10), (6000, 12, 20), (4000, 13, 10)]:
        s0 = ics.rotating_disc(n, seed)
        print(f"server arena reuse: n={n} seed={seed} steps={k}", flush=True)
        rid = srv.submit(s0, cfg, 1e-3, k)
        srv.drain()
        rs, ra = oracle.f32_evolve(s0, k, 1e-3, ocfg(oracle, cfg))
        s, a = srv.take(rid)
        assert np.array_equal(s.view(np.uint32), rs.view(np.uint32)), n
        assert np.array_equal(a.view(np.uint32), ra.view(np.uint32)), n


def test_tuning_rejects_unsupported_shapes(gpu):
    sims = mixed_sims()[:1]
    for kw, msg in [(dict(sort_tile=512), "sort_tile"), (dict(traverse_block=96), "traverse_block"),
                    (dict(traverse_waves_per_cu=33), "traverse_waves_per_cu"), (dict(streams=3), "streams")]:
        with pytest.raises(Exception, match=msg):
            Batch(gpu, [sims[0][0]], sims[0][1], sims[0][2], **kw)


def test_autotune_picks_a_measured_choice_and_caches_it(gpu, tmp_path):
    from galaxy_amd import autotune

    states = [ics.rotating_disc(n, 2000 + n) for n in (900, 3000, 12000)]
    cfg = Config(theta=0.5, softening_kpc=0.01, g=1.0, bucket=4, max_depth=16)
    cache = str(tmp_path / "tune.json")
    print("autotune small ensemble", flush=True)
    r = autotune.tune(gpu, states, cfg, 1e-3, cache=cache, steps=1)
    assert not r["cached"]
    measured = {tuple(sorted(c.items())): t for c, t in r["measured"]}
    best = tuple(sorted(autotune.asdict(r["choice"]).items()))
    assert measured[best] == min(measured.values())
    again = autotune.tune(gpu, states, cfg, 1e-3, cache=cache, steps=1)
    assert again["cached"] and again["choice"] == r["choice"]