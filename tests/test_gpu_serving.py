"""Batched / served MD (reaxamd.serving): the result of a request is the same alone and in the mix (bit for bit: every engine
is deterministic and owns its buffers and stream), requests of uneven size run concurrently, and a Poisson-arrival trace
reports latencies."""
import numpy as np
import pytest


def small_batch(count=10, seed=3):
    from reaxamd import serving
    return serving.batch_systems(count, nmin=100, nmax=700, seed=seed)


def requests(systems_, steps=20):
    from reaxamd.serving import Request
    return [Request(rid=k, system=s, steps=steps, t_target=2500.0, dt=0.1, seed=100 + k) for k, s in enumerate(systems_)]


@pytest.mark.gpu
def test_same_result_alone_and_in_the_mix(ffield_text):
    from reaxamd.serving import Server
    syst = small_batch()
    print("systems", [s.n for s in syst], flush=True)
    alone = Server(ffield_text, slots=1)
    ra = alone.run(requests(syst))
    alone.close()
    mix = Server(ffield_text, slots=4)
    rb = mix.run(requests(syst))
    mix.close()
    overlap = sum(1 for a in rb for b in rb if a.rid < b.rid and a.start < b.end and b.start < a.end)
    print("overlapping request pairs in the mix:", overlap, flush=True)
    assert overlap > 0
    for a, b in zip(sorted(ra, key=lambda r: r.rid), sorted(rb, key=lambda r: r.rid)):
        for k in a.result["energy"]:
            assert a.result["energy"][k] == b.result["energy"][k], (a.rid, k)
        assert np.array_equal(a.result["grad"], b.result["grad"]), a.rid
        assert np.array_equal(a.result["charges"], b.result["charges"]), a.rid
        assert np.array_equal(a.result["positions"], b.result["positions"]), a.rid


@pytest.mark.gpu
def test_poisson_trace_reports_latency(ffield_text):
    from reaxamd.serving import Server, latency_stats
    syst = small_batch(8, seed=5)
    reqs = requests(syst, steps=10)
    rng = np.random.default_rng(1)
    t = 0.0
    for r in reqs:
        t += rng.exponential(0.02)
        r.arrival = t
    srv = Server(ffield_text, slots=3)
    done = srv.run(reqs, realtime=True)
    srv.close()
    st = latency_stats(done)
    print(st, flush=True)
    assert all(r.end > r.start >= r.arrival - 1e-6 for r in done)
    assert st["p99_s"] >= st["p50_s"] > 0


@pytest.mark.gpu
def test_engine_reuse_matches_fresh_engine(ffield_text):
    """An engine that ran one system and then receives another with the same atom count and ids (1..n) must give the fresh
    engine's answer once its charge history is reset: the stock extrapolation of the new system starts from a zero history."""
    from reaxamd.serving import Request, new_engine, run_one
    from reaxamd import systems
    a = systems.ch4_o2_gas(81, density=0.3, seed=11)
    b = systems.ch4_o2_gas(81, density=0.3, seed=12)
    assert np.array_equal(a.tags, b.tags)
    ra_req, rb_req = Request(0, a, 15, t_target=2500.0, dt=0.1, seed=1), Request(1, b, 15, t_target=2500.0, dt=0.1, seed=2)
    e = new_engine(ffield_text, "mixed")
    run_one(e, ffield_text, ra_req, "mixed")
    reused = run_one(e, ffield_text, rb_req, "mixed")
    e.close()
    f = new_engine(ffield_text, "mixed")
    fresh = run_one(f, ffield_text, rb_req, "mixed")
    f.close()
    assert np.array_equal(reused["charges"], fresh["charges"])
    assert np.array_equal(reused["positions"], fresh["positions"])
    assert np.array_equal(reused["grad"], fresh["grad"])


@pytest.mark.gpu
def test_packed_server_same_result_alone_and_in_the_mix(ffield_text):
    """Dynamic batching: requests of uneven sizes and step counts packed in one context give, bit for bit, what each gives
    alone (a pack of one); the pack's chunks end at every request's own step count."""
    from reaxamd.serving import Request, Server, new_engine, run_pack_chunks
    syst = small_batch(8, seed=4)

    def reqs():
        return [Request(rid=k, system=s, steps=6 + 3 * k, t_target=2500.0, dt=0.1, seed=200 + k) for k, s in enumerate(syst)]
    srv = Server(ffield_text, slots=1, packed=True)
    mixed = srv.run(reqs())
    srv.close()
    assert all(r.end >= r.start for r in mixed)
    e = new_engine(ffield_text, "mixed")
    for k in (0, 3, 7):
        alone = run_pack_chunks(e, ffield_text, [reqs()[k]])[0].result
        m = mixed[k].result
        for name in alone["energy"]:
            assert alone["energy"][name] == m["energy"][name], (k, name)
        for key in ("grad", "charges", "positions"):
            assert np.array_equal(alone[key], m[key]), (k, key)
    e.close()


@pytest.mark.gpu
def test_packed_server_poisson_trace(ffield_text):
    from reaxamd.serving import Server, latency_stats
    syst = small_batch(12, seed=6)
    reqs = requests(syst, steps=12)
    rng = np.random.default_rng(2)
    t = 0.0
    for r in reqs:
        t += rng.exponential(0.01)
        r.arrival = t
    srv = Server(ffield_text, slots=2, packed=True)
    done = srv.run(reqs, realtime=True)
    srv.close()
    st = latency_stats(done)
    print(st, flush=True)
    assert all(r.end > r.start >= r.arrival - 1e-6 and r.result for r in done)
    assert st["p99_s"] >= st["p50_s"] > 0
