"""Serving on one device (fz_server_*): a case's result is the same alone and
in a mix of cases that arrive while others run, admission respects the
footprint budget in arrival order, and the footprint model matches what a
case actually allocates."""
import numpy as np
import pytest

import fizzy_amd as F
from fizzy_amd import cases, serve

pytestmark = pytest.mark.gpu

FIELDS = ["u", "v", "w", "p", "psi", "phi", "kappa"]


def mix():
    return [(serve.bubble_case(32, two_d=True, density_ratio=500.0), 23),
            (cases.rising_bubble_3d(ng=(24, 24, 48)), 17),
            (serve.bubble_case(16, 32, density_ratio=3.0), 30),
            (cases.hysing_2d(), 12),
            (dict(serve.bubble_case(32, 32, density_ratio=20.0), precision="mixed", refine=1), 15)]


def alone(spec, nsteps):
    dc = F.DeviceCase(spec)
    try:
        dc.init()
        dc.step(nsteps)
        return {k: dc.get(k) for k in FIELDS}, dc.scalars()
    finally:
        dc.close()


@pytest.mark.parametrize("batch", ["1", "8"])
def test_results_are_the_same_alone_and_in_the_mix(batch, monkeypatch):
    """(batch: FZ_SERVE_BATCH, the most steps a short case gets per round)"""
    monkeypatch.setenv("FZ_SERVE_BATCH", batch)
    reqs = mix()
    ref = [alone(s, n) for s, n in reqs]
    free, _ = serve_memory()
    srv = serve.Server(int(0.5 * free))
    try:
        ids = [srv.submit(*reqs[0]), srv.submit(*reqs[1])]
        for _ in range(5):
            srv.round()
        ids.append(srv.submit(*reqs[2]))  # arrives while the first two run
        for _ in range(3):
            srv.round()
        ids.append(srv.submit(*reqs[3]))
        ids.append(srv.submit(*reqs[4]))
        while any(srv.status(i)[0] != 2 for i in ids):
            srv.round()
        for i, (fields, sc) in zip(ids, ref):
            dc = srv.case(i)
            for k in FIELDS:
                assert np.array_equal(dc.get(k), fields[k]), (i, k)
            assert np.array_equal(dc.scalars(), sc)
            assert len(srv.step_times(i)) == reqs[ids.index(i)][1]
            srv.release(i)
        assert srv.memory()[0] == 0
    finally:
        srv.close()


def serve_memory():
    import ctypes as C
    f, t = C.c_size_t(), C.c_size_t()
    F.call("fz_device_memory", C.byref(f), C.byref(t))
    return f.value, t.value


def test_admission_in_arrival_order_within_the_budget():
    specs = [(cases.rising_bubble_3d(ng=(32, 32, 64)), 400), (serve.bubble_case(24, 48), 3),
             (serve.bubble_case(24, 48, density_ratio=50.0), 2)]
    b = [serve.footprint(s) for s, _ in specs]
    assert b[1] == b[2]
    srv = serve.Server(b[0] + b[1] + b[2] // 2)  # the first two fit together, the third once the second leaves
    try:
        ids = [srv.submit(s, n) for s, n in specs]
        for _ in range(50):  # both admitted; running once their set-ups are done
            srv.round()
            if srv.status(ids[1])[0] != 0:
                break
        assert [srv.status(i)[0] for i in ids] == [1, 1, 0]
        while srv.status(ids[1])[0] != 2:
            srv.round()
        assert srv.status(ids[2])[0] == 0  # finished cases hold their memory until released
        srv.release(ids[1])
        for _ in range(50):  # admitted now; running once its set-up (on the server's thread) is done
            srv.round()
            if srv.status(ids[2])[0] != 0:
                break
        assert srv.status(ids[2])[0] == 1
        used, budget = srv.memory()
        assert used == b[0] + b[2] <= budget
    finally:
        srv.close()


@pytest.mark.parametrize("spec", [cases.rising_bubble_3d(ng=(64, 64, 128)), serve.bubble_case(128, two_d=True),
                                  cases.vof_droplets(ng=(48, 48, 48)),
                                  dict(cases.rising_bubble_3d(ng=(64, 64, 128)), precision="mixed", refine=1)],
                         ids=["bubble64x64x128", "bubble2d128", "vof48", "mixed64x64x128"])
def test_footprint_model_matches_the_allocation(spec):
    before, _ = serve_memory()
    dc = F.DeviceCase(spec)
    try:
        dc.init()
        dc.step(11)  # every boundary map and reduction buffer exists after a check step
        dc.sync()
        after, _ = serve_memory()
        used = before - after
        est = serve.footprint(spec)
        print(spec["ng"], "allocated", used, "model", est)
        assert 0.9 * used <= est <= 1.25 * used + (8 << 20)
    finally:
        dc.close()


def test_split_step_matches_the_whole_step():
    """fz_case_step_begin/_end (the server's way of stepping) give the fields
    and scalars of fz_case_step bit for bit across check steps (icheck = 4:
    the next dt comes from the read-back that end() takes), and begin twice
    without end() is refused."""
    import ctypes as C
    spec = dict(serve.bubble_case(24, 48, density_ratio=100.0), icheck=4)
    ref = alone(spec, 11)
    dc = F.DeviceCase(spec)
    try:
        dc.init()
        pending, killed = C.c_int(), C.c_int()
        seen = []
        for _ in range(11):
            F.call("fz_case_step_begin", dc._h, C.byref(pending))
            seen.append(pending.value)
            if pending.value:
                with pytest.raises(F.FizzyError):
                    F.call("fz_case_step_begin", dc._h, C.byref(pending))
            F.call("fz_case_step_end", dc._h, C.byref(killed))
            assert killed.value == 0
        assert seen == [0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]
        for k in FIELDS:
            assert np.array_equal(dc.get(k), ref[0][k]), k
        assert np.array_equal(dc.scalars(), ref[1])
    finally:
        dc.close()


def test_a_request_of_zero_steps_finishes():
    srv = serve.Server(2 * serve.footprint(serve.bubble_case(16, 32)))
    try:
        a = srv.submit(serve.bubble_case(16, 32), 0)
        b = srv.submit(serve.bubble_case(16, 32), 2)
        for _ in range(50):
            srv.round()
            if srv.status(a)[0] == 2 and srv.status(b)[0] == 2:
                break
        assert srv.status(a)[0] == 2 and len(srv.step_times(a)) == 0
        assert srv.status(b)[0] == 2 and len(srv.step_times(b)) == 2
    finally:
        srv.close()


def test_a_failed_request_stops_only_itself():
    """A request whose set-up fails (mixed precision with non-homogeneous
    pressure boundary values is refused by fz_case_create_ex) ends in state 3
    with its error message, its memory returns at once, and the requests
    beside it finish with the fields they have alone."""
    good = serve.bubble_case(16, 32, density_ratio=30.0)
    bad = dict(serve.bubble_case(16, 32), precision="mixed", refine=1, cbcpre=list("PPPPND"),
               bcpre=[0.0, 0.0, 0.0, 0.0, 0.0, 0.5])
    ref = alone(good, 6)
    srv = serve.Server(serve.footprint(good) + serve.footprint(bad))
    try:
        a = srv.submit(bad, 4)
        b = srv.submit(good, 6)
        for _ in range(100):
            srv.round()
            if srv.status(a)[0] in (2, 3) and srv.status(b)[0] in (2, 3):
                break
        assert srv.status(a)[0] == serve.Server.FAILED
        assert "homogeneous" in srv.error(a)
        assert srv.error(b) == ""
        with pytest.raises(F.FizzyError):
            srv.case(a)
        assert srv.status(b)[0] == serve.Server.FINISHED
        dc = srv.case(b)
        for k in FIELDS:
            assert np.array_equal(dc.get(k), ref[0][k]), k
        srv.release(a)
        srv.release(b)
        assert srv.memory()[0] == 0
    finally:
        srv.close()


def test_close_while_cases_are_being_set_up():
    """Closing the server with requests still queued for (or in) set-up on
    its thread returns, and a new server works after it."""
    specs = [serve.bubble_case(64, 128), serve.bubble_case(32, 64), serve.bubble_case(48, 96)]
    srv = serve.Server(sum(serve.footprint(s) for s in specs))
    for s in specs:
        srv.submit(s, 5)
    srv.round()  # waits for the first set-up only
    srv.close()
    srv = serve.Server(serve.footprint(specs[1]))
    try:
        rid = srv.submit(specs[1], 2)
        for _ in range(50):
            srv.round()
            if srv.status(rid)[0] == 2:
                break
        assert srv.status(rid)[0] == 2
    finally:
        srv.close()


def test_serving_mix_at_its_workload():
    """The serving configuration itself: a seeded Poisson-arrival stream (mean
    gap 50 ms) of 8 requests of the mix -- three 128^2x256 and one 64^2x128
    3-D bubbles, two 256^2 and two 64^2 2-D ones, density ratios from 7.8 to
    998 -- of 200 mixed-precision steps each, run through the server at their
    arrival times. Every result equals the case run alone, bit for bit."""
    reqs = serve.serving_requests(8, mean_interarrival=0.05, nsteps=200, precision="mixed")
    grids = [tuple(s["ng"]) for _, s, _ in reqs]
    assert (128, 128, 256) in grids and (256, 2, 256) in grids
    assert max(s["rho12"][0] for _, s, _ in reqs) > 990.0
    ref = [alone(s, n) for _, s, n in reqs]
    free, _ = serve_memory()
    srv = serve.Server(int(0.5 * free))
    got = {}

    def keep(rid):
        assert srv.status(rid)[0] == serve.Server.FINISHED, srv.error(rid)
        dc = srv.case(rid)
        got[rid] = ({k: dc.get(k) for k in FIELDS}, dc.scalars(), len(srv.step_times(rid)))

    try:
        wall, ids = serve.run_stream(srv, reqs, release=True, on_finish=keep)
        print("stream wall time", wall)
        for rid, (fields, sc), (_, _, n) in zip(ids, ref, reqs):
            gf, gs, nsteps = got[rid]
            assert nsteps == n
            for k in FIELDS:
                assert np.array_equal(gf[k], fields[k]), (rid, k)
            assert np.array_equal(gs, sc), rid
        assert srv.memory()[0] == 0
    finally:
        srv.close()


def test_same_shaped_requests_are_stepped_in_batches():
    """Requests of one shape (four FP64 2-D bubbles and three mixed-precision
    3-D ones with the FP64 projection, each set with its own density ratio),
    arriving while others run: the server steps them together once they have
    taken their first step (fz_case_batch), and every result still equals
    the case run alone, bit for bit. With batching off (FZ_SERVE_GROUP=0)
    the results are the same."""
    two_d = [dict(serve.bubble_case(32, two_d=True, density_ratio=r), icheck=5) for r in (900.0, 40.0, 3.0, 250.0)]
    three_d = [dict(serve.bubble_case(16, 24, density_ratio=r), precision="mixed", refine=-1, icheck=5)
               for r in (500.0, 8.0, 60.0)]
    reqs = [(two_d[0], 40), (three_d[0], 30), (two_d[1], 36), (three_d[1], 25), (two_d[2], 31), (two_d[3], 45),
            (three_d[2], 28)]
    ref = [alone(sp, n) for sp, n in reqs]
    free, _ = serve_memory()
    for group in ("1", "0"):
        import os
        os.environ["FZ_SERVE_GROUP"] = group
        try:
            srv = serve.Server(int(0.5 * free))
        finally:
            del os.environ["FZ_SERVE_GROUP"]
        try:
            ids = [srv.submit(*r) for r in reqs[:3]]
            for k, req in enumerate(reqs[3:]):
                for _ in range(1 + k // 2):
                    srv.round()
                ids.append(srv.submit(*req))
            while any(srv.status(i)[0] != 2 for i in ids):
                srv.round()
            times = {}
            for i, (fields, sc), (_, n) in zip(ids, ref, reqs):
                dc = srv.case(i)
                for k in FIELDS:
                    assert np.array_equal(dc.get(k), fields[k]), (group, i, k)
                assert np.array_equal(dc.scalars(), sc), (group, i)
                times[i] = list(srv.step_times(i))
                assert len(times[i]) == n
                srv.release(i)
            if group == "1":
                # a batched step is timed once for all its cases: the first two
                # 2-D requests share the durations of the steps they took together
                shared = set(times[ids[0]]) & set(times[ids[2]])
                assert len(shared) >= 5, shared
        finally:
            srv.close()
