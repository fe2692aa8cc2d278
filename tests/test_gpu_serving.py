"""GPU tests of the continuous-batching solve server (bnmf_server_*).

Each served result must equal the same request solved alone, bitwise, and
must equal the batch/drop-in solver on the same inputs; admission must hold
the HBM budget; a request above the whole budget is rejected."""
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inst(hip, x, y, z, seed, dtype=np.float32, with_S0=True):
    p = hip.Params(40.0, 1.0, [1.0] * x, [1.0] * z)
    X = hip.sample_X(seed, (x, y, z), p).astype(dtype)
    S0 = hip.init_S(seed + 1, X, z, dtype=dtype) if with_S0 else None
    return hip.Instance(X=X, z=z, params=p, S0=S0, seed=seed + 1)


SHAPES = [(64, 80, 8), (130, 70, 17), (96, 200, 32), (40, 300, 12), (257, 65, 64), (33, 47, 5)]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_served_equals_alone_and_batch(hip, dtype):
    insts = [_inst(hip, x, y, z, 10 + n, dtype) for n, (x, y, z) in enumerate(SHAPES)]
    iters = 21
    appr = dtype == np.float32
    # alone: one server, one request at a time
    alone = []
    with hip.Server(1 << 30, dtype=dtype, use_psi_appr=appr) as srv:
        for inst in insts:
            t = srv.submit(inst, iters, log_marginal=True)
            assert srv.wait(t, 60.0) == hip.OK
            alone.append(srv.collect(t)[:2])
    # in the mix: all submitted together, staggered, different iteration counts
    extra = [_inst(hip, 150, 150, 24, 99, dtype), _inst(hip, 60, 500, 9, 98, dtype)]
    with hip.Server(1 << 30, dtype=dtype, use_psi_appr=appr) as srv:
        tickets = []
        for n, inst in enumerate(insts):
            tickets.append(srv.submit(inst, iters, log_marginal=True))
            if n == 2:
                for e in extra:
                    srv.submit(e, 37)
                time.sleep(0.01)
        mixed = []
        for t in tickets:
            assert srv.wait(t, 60.0) == hip.OK
            mixed.append(srv.collect(t)[:2])
        st = srv.stats()
        assert st["completed"] >= len(insts)
    for n in range(len(insts)):
        assert np.array_equal(alone[n][0], mixed[n][0]), SHAPES[n]
        assert alone[n][1] == mixed[n][1]
    # the batch solver (and hence the drop-in bld_mult) on the same inputs
    with hip.Batch(insts, dtype=dtype, use_psi_appr=appr) as b:
        b.solve(iters)
        lms = b.log_marginal()
        for n in range(len(insts)):
            assert np.array_equal(b.S(n), alone[n][0]), SHAPES[n]
            assert lms[n] == pytest.approx(alone[n][1], rel=1e-14)


def test_device_init_matches_host_init(hip):
    # a request without S0 draws it on the GPU from the same Philox stream
    inst = _inst(hip, 70, 90, 11, 5, np.float64, with_S0=False)
    with hip.Server(1 << 28, dtype=np.float64, use_psi_appr=False) as srv:
        t = srv.submit(inst, 0)
        assert srv.wait(t, 30.0) == hip.OK
        S0_dev = srv.collect(t)[0]
    S0_host = hip.init_S(inst.seed, inst.X, inst.z)
    assert np.allclose(S0_dev, S0_host, rtol=1e-13, atol=0)
    assert np.allclose(S0_dev.sum(2), inst.X, rtol=1e-12, atol=1e-12)


def test_budget_admission_and_rejection(hip):
    a = _inst(hip, 200, 200, 32, 1)
    need = hip.Server.request_bytes((200, 200, 32))
    assert need > 200 * 200 * 32 * 4
    # budget for one request at a time: all finish, never two active
    with hip.Server(int(need * 1.5), dtype=np.float32) as srv:
        ts = [srv.submit(a, 10) for _ in range(4)]
        for t in ts:
            assert srv.wait(t, 60.0) == hip.OK
            srv.collect(t)
        st = srv.stats()
        assert st["peak_active"] == 1 and st["completed"] == 4
        assert st["peak_bytes"] <= int(need * 1.5) + 256
        # larger than the whole budget: rejected at once
        big = _inst(hip, 400, 400, 32, 2)
        t = srv.submit(big, 5)
        assert srv.wait(t, 5.0) == hip.ERR_REJECTED
        with pytest.raises(hip.RequestRejected):
            srv.collect(t)
        assert srv.stats()["rejected"] == 1


def test_uncollected_results_hold_memory(hip):
    # S stays in HBM until collected: the second request waits for the first
    a = _inst(hip, 128, 128, 16, 3)
    need = hip.Server.request_bytes((128, 128, 16))
    with hip.Server(int(need * 1.2), dtype=np.float32) as srv:
        t1 = srv.submit(a, 5)
        t2 = srv.submit(a, 5)
        assert srv.wait(t1, 30.0) == hip.OK
        assert srv.wait(t2, 0.3) == hip.ERR_TIMEOUT
        S1 = srv.collect(t1)[0]
        assert srv.wait(t2, 30.0) == hip.OK
        S2 = srv.collect(t2)[0]
        assert np.array_equal(S1, S2)


def test_concurrent_clients(hip):
    insts = [_inst(hip, 90 + 7 * n, 70 + 5 * n, 6 + 3 * n, 40 + n) for n in range(8)]
    out = {}
    with hip.Server(1 << 30, dtype=np.float32) as srv:
        def client(n):
            t = srv.submit(insts[n], 15)
            srv.wait(t, 60.0)
            out[n] = srv.collect(t)[0]
        th = [threading.Thread(target=client, args=(n,)) for n in range(8)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    for n in range(8):
        ref = hip.bld_mult(insts[n].X, insts[n].z, insts[n].params, 15, True, S0=insts[n].S0,
                           dtype=np.float32)
        assert np.array_equal(out[n], ref)
