"""The serving loop (fem_server_*): each solve's result is the same alone and in a mix, admission
respects the device-memory budget, and cancellation and refusal behave as documented.

Tolerance: a solve in the mix must equal the same solve alone to 1e-12 relative L2. In a batch
each solve's blocks run exactly the launches of its own iteration (on its own buffers), and
unbatched each solve runs them on its own stream, so the tests also expect bitwise equality, and
check each against the oracle's direct solution to the 1e-8 parity bar.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def specs():
    """Refinement pairs (h, h/2) of heat tet4 and elasticity tet10 boxes, seeded materials."""
    rng = np.random.default_rng(17)
    out = []
    for n in ([4, 2, 2], [8, 4, 4]):
        nt = 6 * int(np.prod(n))
        out.append(("heat_tet4", n, dict(coef0=rng.uniform(0.5, 2.0, nt), fixed={"x0": 1.0, "x1": 0.0})))
        out.append(("elastic_tet10", n, dict(coef0=rng.uniform(1.0, 2.0, nt), coef1=rng.uniform(0.2, 0.35, nt),
                                             fixed={"x0": True}, body=(0, 0, -1))))
    for n in ([8, 4, 4], [16, 8, 8]):
        nt = 6 * int(np.prod(n))
        out.append(("heat_tet4", n, dict(coef0=rng.uniform(0.5, 2.0, nt), fixed={"y0": 0.0, "z1": 2.0}, source=1.0)))
    return out


def make_box(fem, kind, n, kw):
    length = [float(v) / max(n) for v in n]
    return fem.Box(kind, n, length, kw["coef0"], kw.get("coef1"), fixed=kw["fixed"], source=kw.get("source", 0.0),
                   body=kw.get("body", (0.0, 0.0, 0.0)))


# levels=3: V-cycles even on the smallest boxes, whose automatic hierarchy would be one dense level
OPTS = dict(method="multigrid", precision="fp32", tol=1e-10, levels=3)


def test_each_result_is_the_same_alone_and_in_the_mix(fem, oracle):
    cases = specs()
    alone = []
    for kind, n, kw in cases:
        s = fem.Solver(make_box(fem, kind, n, kw), **OPTS)
        x, st = s.solve(s.rhs())
        alone.append(x)
    srv = fem.Server(hbm_budget=8 << 30, max_active=3)
    try:
        ids = [srv.submit(make_box(fem, kind, n, kw), **OPTS) for kind, n, kw in cases]
        for (kind, n, kw), rid, xa in zip(cases, ids, alone):
            x, st, info = srv.wait(rid, len(xa), timeout=120)
            print(kind, n, st.iterations, f"{info.admit_time:.3f} {info.start_time:.3f} {info.finish_time:.3f}",
                  info.footprint, info.device_bytes)
            assert np.linalg.norm(x - xa) <= 1e-12 * np.linalg.norm(xa)
            assert np.array_equal(x, xa)
            assert info.device_bytes <= info.footprint
        # every solve took its memory from the arena reserved at the start: no device allocation
        # ran beside the running solves
        arena, peak, misses = srv.memory()
        assert arena >= 8 << 30 and 0 < peak <= arena and misses == 0
        # the iterations went out in batches (more solve iterations than batches); whether kernels
        # merge here depends on which solves happen to overlap (these small solves are shorter
        # than a setup), which the test below makes certain
        bs = srv.batching()
        print(bs)
        assert bs["batches"] > 0 and bs["iterations"] > bs["batches"]
    finally:
        srv.close()
    # and each agrees with the oracle's direct solve
    for (kind, n, kw), xa in zip(cases, alone):
        length = [float(v) / max(n) for v in n]
        if kind == "heat_tet4":
            pr = oracle.Problem.heat(n, length, kw["coef0"], fixed=kw["fixed"], source=kw.get("source", 0.0), tet4=True)
        else:
            pr = oracle.Problem.elastic(10, n, length, kw["coef0"], kw["coef1"], clamped=list(kw["fixed"]),
                                        body=kw["body"])
        xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
        assert np.linalg.norm(xa - xd) <= 1e-8 * np.linalg.norm(xd)


def test_identical_solves_in_flight_share_their_launches(fem):
    """Five copies of one box in flight together: their iterations go out in batches whose kernels
    merge (fewer launches than the same iterations one solve at a time), replayed as captured graphs
    once the member set repeats. The solves run to their iteration limit (a negative tolerance,
    which no residual meets), so each runs longer than a setup and they overlap whatever the order
    of their setups; each ends with the reference's non-convergence text (multigrid.rs:730-733)."""
    rng = np.random.default_rng(29)
    n = [16, 8, 8]
    k = rng.uniform(0.5, 2.0, 6 * int(np.prod(n)))
    box = fem.Box("heat_tet4", n, [1.0, 0.5, 0.5], k, fixed={"x0": 1.0, "x1": 0.0})
    nfree = len(fem.Model(box).free_list())
    srv = fem.Server(hbm_budget=8 << 30, max_active=5)
    try:
        ids = [srv.submit(box, method="multigrid", precision="fp32", tol=-1.0, levels=3, max_iter=80,
                          refine=False) for _ in range(5)]
        for rid in ids:
            with pytest.raises(fem.FemError, match="^conjugate gradients did not converge: the residual is .* after 80 iterations$"):
                srv.wait(rid, nfree, timeout=120)
        bs = srv.batching()
        print(bs)
        assert bs["batches"] < bs["iterations"]
        assert bs["launches"] < bs["solo_launches"] and bs["graphs"] >= 1
    finally:
        srv.close()


def test_admission_keeps_the_footprint_within_the_budget(fem):
    cases = [c for c in specs() if c[0] == "heat_tet4"]
    boxes = [make_box(fem, kind, n, kw) for kind, n, kw in cases]
    # a budget that holds the largest solve but not two of them
    probe = fem.Server(hbm_budget=8 << 30, max_active=1)
    try:
        fps = []
        for b in boxes:
            rid = probe.submit(b, **OPTS)
            _, _, info = probe.wait(rid, fem.Model(b).free_list().size, timeout=120)
            fps.append(info.footprint)
    finally:
        probe.close()
    budget = max(fps) + min(fps) - 1
    srv = fem.Server(hbm_budget=budget, max_active=8)
    try:
        ids = [srv.submit(b, **OPTS) for b in boxes]
        spans = []
        for b, rid, fp in zip(boxes, ids, fps):
            _, _, info = srv.wait(rid, fem.Model(b).free_list().size, timeout=120)
            spans.append((info.admit_time, info.finish_time, fp))
        # at any admission instant, the footprints of the solves then admitted fit the budget
        for t, _, _ in spans:
            held = sum(fp for a, f, fp in spans if a <= t < f)
            assert held <= budget
    finally:
        srv.close()


def test_a_cancelled_solve_ends_with_the_reference_text(fem):
    kind, n, kw = specs()[-1]
    box = make_box(fem, kind, n, kw)
    srv = fem.Server(hbm_budget=8 << 30, max_active=2)
    try:
        rid = srv.submit(box, method="multigrid", precision="fp32", tol=1e-300, max_iter=100000)
        assert srv.cancel(rid)
        with pytest.raises(fem.FemError, match="^cancelled$"):
            srv.wait(rid, fem.Model(box).free_list().size, timeout=120)
    finally:
        srv.close()


def test_a_solve_beyond_the_budget_is_refused(fem):
    kind, n, kw = specs()[-1]
    srv = fem.Server(hbm_budget=1 << 16, max_active=2)
    try:
        with pytest.raises(fem.FemError, match="more than the server's budget"):
            srv.submit(make_box(fem, kind, n, kw), **OPTS)
    finally:
        srv.close()


def test_freeing_the_server_ends_a_waiting_request_and_its_waiter(fem):
    """A waiter blocked on an unfinished request returns with the server's stop message when another
    thread frees the server. One solve at a time: request A's setup (a forced dense direct factor
    of 27,869 unknowns) outlasts the 0.3 s before the free, so request B is still queued then."""
    import threading
    import time

    n = [30, 30, 30]
    k = np.exp(np.random.default_rng(3).standard_normal(int(np.prod(n))))
    big = fem.Box("heat_hex8", n, [1.0, 1.0, 1.0], k, fixed={"x0": 0.0, "x1": 1.0})
    kind, m, kw = specs()[0]
    small = make_box(fem, kind, m, kw)
    srv = fem.Server(hbm_budget=48 << 30, max_active=1)
    srv.submit(big, method="direct", direct_solver="dense")
    rid = srv.submit(small, **OPTS)
    got = {}

    def waiter():
        try:
            srv.wait(rid, fem.Model(small).free_list().size, timeout=-1)
            got["result"] = "finished"
        except fem.FemError as e:
            got["result"] = str(e)

    t = threading.Thread(target=waiter)
    t.start()
    time.sleep(0.3)
    srv.close()
    t.join(timeout=120)
    assert not t.is_alive()
    assert got["result"] == "the server stopped before the solve finished"


def test_unbatched_and_marching_solves_match_their_solves_alone(fem):
    """The loop without batching (each solve launches its own iterations), and a batching loop whose
    mix holds hex8 boxes (their marching kernels are not batched: those solves run their own
    iterations beside the batches): every result equals the same solve alone, bitwise."""
    rng = np.random.default_rng(23)
    cases = [(k, n, kw) for k, n, kw in specs() if max(n) <= 8]
    hexes = []
    for n in ([8, 8, 8], [12, 6, 6]):
        k = np.exp(rng.standard_normal(int(np.prod(n))))
        hexes.append(fem.Box("heat_hex8", n, [1.0, 1.0, 1.0], k, fixed={"x0": 0.0, "x1": 1.0}))
    boxes = [make_box(fem, kind, n, kw) for kind, n, kw in cases] + hexes
    alone = []
    for b in boxes:
        s = fem.Solver(b, **OPTS)
        alone.append(s.solve(s.rhs())[0])
    for batched in (False, True):
        srv = fem.Server(hbm_budget=8 << 30, max_active=4, batched=batched)
        try:
            ids = [srv.submit(b, **OPTS) for b in boxes]
            for b, rid, xa in zip(boxes, ids, alone):
                x, st, _ = srv.wait(rid, len(xa), timeout=120)
                assert np.array_equal(x, xa), (batched, b.kind)
            bs = srv.batching()
            print("batched" if batched else "unbatched", bs)
            if batched:
                assert bs["batches"] > 0
            else:
                assert bs["batches"] == 0
        finally:
            srv.close()
