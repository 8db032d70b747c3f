"""The request server (tbk_server_*): continuous batching of DoS and sigma_xy
requests on one GPU.  A DoS result must equal the single-problem path (same
vectors; moment sums in another order: 1e-12 relative) and be the same bits
alone and in any mix; a sigma_xy result must be the direct path's Gamma bit
for bit; admission must respect the memory budget."""
import math
import time

import numpy as np
import pytest

from tests.test_gpu_batch import model_problem
from tight_binding_amd import _native as Nn
from tight_binding_amd import kpm
from tight_binding_amd import lattice as L
from tight_binding_amd import spectral as S

pytestmark = pytest.mark.gpu


def dos_requests(k=8, seed=4):
    rng = np.random.default_rng(seed)
    models = ["square", "honeycomb", "lieb", "kagome"]
    out = []
    for i in range(k):
        H, a = model_problem(models[i % 4], int(rng.integers(6, 40)), float(rng.uniform(0, 2)), seed=50 + i)
        out.append(dict(H=H, n=int(rng.choice([2, 5, 40, 64, 101, 200])), a=a, b=0.05 * (i % 2),
                        R=int(rng.choice([8, 16])), seed=900 + i,
                        precision=kpm.FP64 if i % 3 else kpm.FP32))
    return out


def run_all(srv, reqs, stagger=0.0):
    tickets = []
    for r in reqs:
        tickets.append(srv.submit_dos(r["H"], r["n"], r["a"], r["b"], r["R"], r["seed"], r["precision"]))
        if stagger:
            time.sleep(stagger)
    out = []
    for t in tickets:
        info = srv.wait(t, 60.0)
        assert info["state"] == Nn.TBK_REQ_DONE, info
        mu, info = srv.take(t)
        out.append(mu)
    return out


def test_dos_requests_match_single_path(gpu_ctx):
    reqs = dos_requests()
    with kpm.Server(0, lazy_arena=True, chunk_steps=8) as srv:
        res = run_all(srv, reqs)
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0 and st["used_bytes"] == 0
    for r, mu in zip(reqs, res):
        M = gpu_ctx.matrix(r["H"])
        K = gpu_ctx.random_block(r["H"].n, r["R"], seed=r["seed"], precision=r["precision"])
        ref, _, _ = gpu_ctx.moments(M, K, r["n"], a=r["a"], b=r["b"])
        assert mu.size == r["n"]
        assert np.max(np.abs(mu - ref)) <= 1e-12 * np.max(np.abs(ref)), (r["n"], np.max(np.abs(mu - ref)))


def test_alone_and_in_mix_are_bit_identical():
    reqs = dos_requests(10, seed=8)
    target = 6
    with kpm.Server(0, lazy_arena=True, chunk_steps=4) as srv:
        alone = run_all(srv, [reqs[target]])[0]
    # the same request amid others, submitted while they are running
    with kpm.Server(0, lazy_arena=True, chunk_steps=4) as srv:
        mixed = run_all(srv, reqs, stagger=0.002)
    assert np.array_equal(alone, mixed[target])
    # and with another chunk length and slot width
    with kpm.Server(0, lazy_arena=True, chunk_steps=16, cpl=1) as srv:
        again = run_all(srv, reqs[::-1])
    for i, mu in enumerate(again[::-1]):
        with kpm.Server(0, lazy_arena=True, chunk_steps=16, cpl=1) as s2:
            solo = run_all(s2, [reqs[i]])[0]
        assert np.array_equal(solo, mu)
        if i >= 2:
            break


def landau_kubo(L0=8, N=24, R=4):
    c = L.UnitCell.model("square")
    c.add_hopping_nn("A", 1, 1.0)
    lat = L.PartialLattice(c, (L0, L0))
    lat.set_gauge(L.GAUGE_LANDAU, 2 * math.pi / 8)
    lat.set_disorder(1.0, 3)
    H = lat.get_Hamiltonian_sparse(True)
    Hs = L.rescale(H, 5.2, 0.0)
    X, Y = lat.op_position(0), lat.op_position(1)
    vx, vy = L.commutator_diag(X, Hs), L.commutator_diag(Y, Hs)
    return Hs, vx, vy, S.lorentz(N, 6.0)


@pytest.mark.parametrize("prec", [kpm.GAMMA_FP64, kpm.GAMMA_BF16X3])
def test_kubo_request_matches_direct_path(gpu_ctx, prec):
    Hs, vx, vy, g = landau_kubo()
    N = g.size
    ref, _ = kpm.kubo_gamma(gpu_ctx, gpu_ctx.matrix(Hs), gpu_ctx.matrix(vx), gpu_ctx.matrix(vy),
                            gpu_ctx.random_block(Hs.n, 4, seed=77), N, precision=prec, g=g)
    with kpm.Server(0, lazy_arena=True) as srv:
        t = srv.submit_kubo(Hs, vx, vy, N, R=4, seed=77, precision=prec, g=g)
        assert srv.wait(t, 60.0)["state"] == Nn.TBK_REQ_DONE
        (z, G), info = srv.take(t)
    assert info["result_doubles"] == 2 * N + 2 * N * N
    assert np.array_equal(G, ref)
    zref, _ = kpm.kubo_zeta(gpu_ctx, ref, S.kubo_nodes(N))
    assert np.max(np.abs(z - zref)) <= 1e-12 * np.max(np.abs(zref))


def test_mixed_dos_and_kubo_overlap_and_results_unchanged():
    reqs = dos_requests(6, seed=13)
    Hs, vx, vy, g = landau_kubo(10, 32, 4)
    with kpm.Server(0, lazy_arena=True, chunk_steps=4) as srv:
        alone = run_all(srv, reqs)
        tk = srv.submit_kubo(Hs, vx, vy, g.size, R=4, seed=5, g=g)
        assert srv.wait(tk, 60.0)["state"] == Nn.TBK_REQ_DONE
        (z_alone, G_alone), _ = srv.take(tk)
    with kpm.Server(0, lazy_arena=True, chunk_steps=4, max_kubo=2) as srv:
        tk = srv.submit_kubo(Hs, vx, vy, g.size, R=4, seed=5, g=g)
        mixed = run_all(srv, reqs)
        assert srv.wait(tk, 60.0)["state"] == Nn.TBK_REQ_DONE
        (z_mix, G_mix), _ = srv.take(tk)
        st = srv.stats()
    assert st["kubo_done"] == 1 and st["completed"] == len(reqs) + 1
    assert np.array_equal(G_alone, G_mix) and np.array_equal(z_alone, z_mix)
    for a, b in zip(alone, mixed):
        assert np.array_equal(a, b)


def test_admission_respects_the_memory_budget():
    reqs = dos_requests(8, seed=21)
    # footprint of the largest request, from a probe server
    with kpm.Server(0, lazy_arena=True) as srv:
        fp = []
        for r in reqs:
            t = srv.submit_dos(r["H"], r["n"], r["a"], r["b"], r["R"], r["seed"], r["precision"])
            fp.append(srv.wait(t, 60.0)["footprint_bytes"])
            srv.take(t)
    budget = int(2.2 * max(fp))
    with kpm.Server(0, hbm_budget_bytes=budget, chunk_steps=2) as srv:
        tickets = [srv.submit_dos(r["H"], r["n"], r["a"], r["b"], r["R"], r["seed"], r["precision"]) for r in reqs]
        peak = 0
        while True:
            st = srv.stats()
            peak = max(peak, st["used_bytes"])
            if st["completed"] + st["failed"] == len(reqs):
                break
            time.sleep(0.0005)
        for t in tickets:
            srv.take(t)
        assert st["failed"] == 0
    assert peak <= budget
    # a request larger than the whole budget is refused at submission
    with kpm.Server(0, hbm_budget_bytes=max(fp) // 2) as srv:
        big = reqs[int(np.argmax(fp))]
        with pytest.raises(Nn.NativeError) as e:
            srv.submit_dos(big["H"], big["n"], big["a"], big["b"], big["R"], big["seed"], big["precision"])
        assert e.value.code == Nn.TBK_E_OOM


def test_server_errors():
    H, a = model_problem("square", 6, 0.0, 1)
    with kpm.Server(0, lazy_arena=True, max_moments=64) as srv:
        with pytest.raises(Nn.NativeError) as e:
            srv.submit_dos(H, 1, a)
        assert e.value.code == Nn.TBK_E_INVALID
        with pytest.raises(Nn.NativeError) as e:
            srv.submit_dos(H, 100, a)
        assert e.value.code == Nn.TBK_E_UNSUPPORTED
        with pytest.raises(Nn.NativeError) as e:
            srv.submit_dos(H, 10, float("nan"))
        assert e.value.code == Nn.TBK_E_INVALID
        with pytest.raises(Nn.NativeError) as e:  # sigma_xy operators come rescaled: a, b required
            srv.submit_kubo(H, H, H, 8, a=0.0)
        assert e.value.code == Nn.TBK_E_INVALID
        t = srv.submit_dos(H, 10, a)
        with pytest.raises(Nn.NativeError):
            srv.info(t + 100)
        srv.wait(t)
        mu, info = srv.take(t)
        assert info["state"] == Nn.TBK_REQ_DONE and mu.size == 10


@pytest.mark.parametrize("prec", [kpm.GAMMA_BF16X3, kpm.GAMMA_BF16X3_FP32])
def test_presplit_kubo_request_matches_direct_path(gpu_ctx, prec):
    """A split-BF16x3 request with N >= 256 runs on pre-split stacks in the
    server as in the direct path, from the fp64 or the fp32 recursion: the
    same Gamma bits."""
    Hs, vx, vy, g = landau_kubo(L0=10, N=256, R=2)
    N = g.size
    ref, st = kpm.kubo_gamma(gpu_ctx, gpu_ctx.matrix(Hs), gpu_ctx.matrix(vx), gpu_ctx.matrix(vy),
                             gpu_ctx.random_block(Hs.n, 2, seed=31), N, precision=prec, g=g)
    assert st["recursion_fp32"] == (1 if prec == kpm.GAMMA_BF16X3_FP32 else 0)
    with kpm.Server(0, lazy_arena=True) as srv:
        t = srv.submit_kubo(Hs, vx, vy, N, R=2, seed=31, precision=prec, g=g)
        assert srv.wait(t, 60.0)["state"] == Nn.TBK_REQ_DONE
        (z, G), info = srv.take(t)
    assert np.array_equal(G, ref)


@pytest.mark.parametrize("prec,tol", [(kpm.GAMMA_FP64, 1e-12), (kpm.GAMMA_BF16X3, 1e-5),
                                      (kpm.GAMMA_BF16X3_FP32, 1e-5)])
def test_kubo_request_over_budget_runs_in_vector_groups(gpu_ctx, prec, tol):
    """A sigma_xy request whose stacks exceed the server's budget runs its
    random vectors in groups (Gamma is a sum over vectors) instead of being
    refused: its footprint drops below the budget, and Gamma and zeta equal
    the whole-block result to rounding (the groups' Gammas are added in
    group order).  R = 12 is not a multiple of the group width: the last
    group is partial."""
    Hs, vx, vy, g = landau_kubo(40, 64, 12)
    N = g.size
    with kpm.Server(0, lazy_arena=True) as srv:
        t = srv.submit_kubo(Hs, vx, vy, N, R=12, seed=31, precision=prec, g=g)
        assert srv.wait(t, 60.0)["state"] == Nn.TBK_REQ_DONE
        (z_full, G_full), info_full = srv.take(t)
    budget = int(0.6 * info_full["footprint_bytes"])
    with kpm.Server(0, hbm_budget_bytes=budget) as srv:
        t = srv.submit_kubo(Hs, vx, vy, N, R=12, seed=31, precision=prec, g=g)
        info = srv.wait(t, 60.0)
        assert info["state"] == Nn.TBK_REQ_DONE, info
        assert info["footprint_bytes"] <= budget < info_full["footprint_bytes"]
        (z, G), _ = srv.take(t)
    assert np.max(np.abs(G - G_full)) <= tol * np.max(np.abs(G_full))
    # zeta weights Gamma by up to (1 - e^2)^-2 at the outermost nodes; sigma
    # uses the nodes inside the N/24 cut at each end (snippets.h:420-424)
    cut = N // 24
    zi, zf = z[cut:N - cut], z_full[cut:N - cut]
    ztol = 1e-10 if prec == kpm.GAMMA_FP64 else 1e-4
    assert np.max(np.abs(zi - zf)) <= ztol * np.max(np.abs(zf))
    # the whole-block server result is the direct path's
    if prec == kpm.GAMMA_FP64:
        ref, _ = kpm.kubo_gamma(gpu_ctx, gpu_ctx.matrix(Hs), gpu_ctx.matrix(vx), gpu_ctx.matrix(vy),
                                gpu_ctx.random_block(Hs.n, 12, seed=31), N, precision=prec, g=g)
        assert np.array_equal(G_full, ref)


def test_kubo_fp64_request_over_budget_runs_in_row_ranges(gpu_ctx):
    """One random vector, fp64 stacks larger than the budget: the stacks hold
    a range of rows of every entry at a time, the chains run once per range,
    and the ranges' Gammas add up to the whole-stack Gamma (1e-12), which is
    the direct path's."""
    Hs, vx, vy, g = landau_kubo(40, 48, 1)
    N = g.size
    with kpm.Server(0, lazy_arena=True) as srv:
        t = srv.submit_kubo(Hs, vx, vy, N, R=1, seed=9, g=g)
        assert srv.wait(t, 60.0)["state"] == Nn.TBK_REQ_DONE
        (z_full, G_full), info_full = srv.take(t)
    budget = int(0.6 * info_full["footprint_bytes"])
    with kpm.Server(0, hbm_budget_bytes=budget) as srv:
        t = srv.submit_kubo(Hs, vx, vy, N, R=1, seed=9, g=g)
        info = srv.wait(t, 60.0)
        assert info["state"] == Nn.TBK_REQ_DONE, info
        assert info["footprint_bytes"] <= budget
        (z, G), _ = srv.take(t)
    assert np.max(np.abs(G - G_full)) <= 1e-12 * np.max(np.abs(G_full))
    cut = N // 24
    assert np.max(np.abs(z[cut:N - cut] - z_full[cut:N - cut])) <= 1e-10 * np.max(np.abs(z_full[cut:N - cut]))
    ref, _ = kpm.kubo_gamma(gpu_ctx, gpu_ctx.matrix(Hs), gpu_ctx.matrix(vx), gpu_ctx.matrix(vy),
                            gpu_ctx.random_block(Hs.n, 1, seed=9), N, g=g)
    assert np.array_equal(G_full, ref)


def _square_lattice(Lc, W, seed):
    c = L.UnitCell.model("square")
    c.add_hopping_nn("A", 1, 1.0)
    lat = L.PartialLattice(c, (Lc, Lc))
    lat.set_disorder(W, seed)
    return lat


def _served_vs_single(gpu_ctx, srv, lat, n, R, seed, a):
    t = srv.submit_dos(lat, n, a, 0.0, R, seed, kpm.FP64, builder=L.BUILD_PERIODIC, extend=(1, 1))
    info = srv.wait(t, 120.0)
    assert info["state"] == Nn.TBK_REQ_DONE, info
    mu, info = srv.take(t)
    M, _ = gpu_ctx.assemble(lat, L.BUILD_PERIODIC, (1, 1), False)
    K = gpu_ctx.random_block(M.info()["rows"], R, seed=seed)
    ref, _, _ = gpu_ctx.moments(M, K, n, a=a)
    K.close()
    M.close()
    assert np.max(np.abs(mu - ref)) <= 1e-12 * np.max(np.abs(ref))
    return info


def test_lazy_arena_grows_and_retires_idle_slabs(gpu_ctx):
    """lazy_arena with a 6 GiB budget: a small request makes the first 4 GiB
    slab; a 4.6 GB request that fits neither that slab nor the room left
    beside it retires the idle slab and gets one of its own; both results
    equal the single path."""
    gib = 1 << 30
    with kpm.Server(0, hbm_budget_bytes=6 * gib, lazy_arena=True) as srv:
        st = srv.stats()
        assert st["arena_bytes"] == 0 and st["arena_slabs"] == 0
        _served_vs_single(gpu_ctx, srv, _square_lattice(128, 1.0, 5), 16, 8, 31, 4.6)
        st = srv.stats()
        assert st["arena_slabs"] == 1 and st["arena_bytes"] == 4 * gib
        info = _served_vs_single(gpu_ctx, srv, _square_lattice(1750, 1.0, 6), 16, 32, 32, 4.6)
        assert info["footprint_bytes"] > 4 * gib
        st = srv.stats()
        assert st["arena_slabs"] == 1 and info["footprint_bytes"] <= st["arena_bytes"] <= 6 * gib
        assert st["used_bytes"] == 0


def test_preallocated_arena_holds_the_budget(gpu_ctx):
    """The default arena is the whole budget, allocated at creation; requests
    are carved from it and their results equal the single path."""
    budget = 3 << 30
    with kpm.Server(0, hbm_budget_bytes=budget) as srv:
        st = srv.stats()
        assert st["arena_slabs"] == 1 and budget - (1 << 16) < st["arena_bytes"] <= budget
        for i in range(3):
            _served_vs_single(gpu_ctx, srv, _square_lattice(200 + 50 * i, 0.5, 9 + i), 24, 16, 40 + i, 4.3)
        assert srv.stats()["arena_bytes"] == st["arena_bytes"]


def _with_values(H, values):
    from tight_binding_amd.sparse import CSC
    return CSC(H.n, H.colptr.copy(), H.rowidx.copy(), np.asarray(values).copy())


def test_server_cancel():
    """tbk_server_cancel: a queued request leaves the queue, a running DoS
    request at the next chunk, a sigma_xy request when its enqueued work
    ends; each ends failed with TBK_E_CANCELLED (take raises it); unknown and
    finished tickets are refused; the server keeps serving."""
    H, a = model_problem("square", 256, 0.5, 5)
    with kpm.Server(0, lazy_arena=True, max_slots=1, chunk_steps=1) as srv:
        t1 = srv.submit_dos(H, 4096, a, 0.0, 16, 1)   # takes the only slot: 2048 one-step chunks
        t2 = srv.submit_dos(H, 64, a, 0.0, 8, 2)      # waits for the slot
        assert srv.info(t2)["state"] == Nn.TBK_REQ_QUEUED
        srv.cancel(t2)
        info2 = srv.info(t2)
        assert info2["state"] == Nn.TBK_REQ_FAILED and info2["error_code"] == Nn.TBK_E_CANCELLED
        t0 = time.time()
        while srv.info(t1)["state"] != Nn.TBK_REQ_RUNNING and time.time() - t0 < 10:
            time.sleep(0.001)
        assert srv.info(t1)["state"] == Nn.TBK_REQ_RUNNING
        srv.cancel(t1)
        info1 = srv.wait(t1, 60.0)
        assert info1["state"] == Nn.TBK_REQ_FAILED and info1["error_code"] == Nn.TBK_E_CANCELLED
        for t in (t1, t2):
            with pytest.raises(Nn.NativeError) as e:
                srv.take(t)
            assert e.value.code == Nn.TBK_E_CANCELLED
        with pytest.raises(Nn.NativeError) as e:
            srv.cancel(123456)
        assert e.value.code == Nn.TBK_E_INVALID
        # still serving: a new request equals the single path
        t3 = srv.submit_dos(H, 64, a, 0.0, 8, 3)
        assert srv.wait(t3, 60.0)["state"] == Nn.TBK_REQ_DONE
        with pytest.raises(Nn.NativeError) as e:
            srv.cancel(t3)
        assert e.value.code == Nn.TBK_E_INVALID
        mu, _ = srv.take(t3)
        st = srv.stats()
    ctx = kpm.Context(0)
    ref, _, _ = ctx.moments(ctx.matrix(H), ctx.random_block(H.n, 8, seed=3), 64, a=a)
    ctx.close()
    assert np.max(np.abs(mu - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert st["cancelled"] == 2 and st["completed"] == 1

    # sigma_xy: one running, one queued behind it (max_kubo = 1); the first
    # may already have ended when its cancel arrives, then cancel is refused
    Hs, vx, vy, g = landau_kubo(L0=64, N=512, R=16)
    with kpm.Server(0, lazy_arena=True, max_kubo=1, max_kubo_moments=512) as srv:
        k1 = srv.submit_kubo(Hs, vx, vy, g.size, R=16, seed=4, g=g)
        k2 = srv.submit_kubo(Hs, vx, vy, g.size, R=16, seed=5, g=g)
        srv.cancel(k2)
        info = srv.wait(k2, 60.0)
        assert info["state"] == Nn.TBK_REQ_FAILED and info["error_code"] == Nn.TBK_E_CANCELLED
        try:
            srv.cancel(k1)
            info = srv.wait(k1, 60.0)
            assert info["state"] == Nn.TBK_REQ_FAILED and info["error_code"] == Nn.TBK_E_CANCELLED
        except Nn.NativeError as e:
            assert e.code == Nn.TBK_E_INVALID and srv.info(k1)["state"] == Nn.TBK_REQ_DONE
        assert srv.stats()["cancelled"] >= 1


def test_server_rejects_non_finite_and_non_hermitian_matrices():
    """Submission checks: a NaN or infinite value in H, v_x or v_y, and an H
    that is not Hermitian, are refused with TBK_E_INVALID; an H that is
    Hermitian to rounding (1e-14 relative asymmetry) is accepted."""
    H, a = model_problem("honeycomb", 8, 1.0, 6)
    Hs, vx, vy, g = landau_kubo()
    with kpm.Server(0, lazy_arena=True) as srv:
        for bad in (np.nan, np.inf):
            v = H.values.copy()
            v[3] = bad
            with pytest.raises(Nn.NativeError) as e:
                srv.submit_dos(_with_values(H, v), 32, a, 0.0, 8, 1)
            assert e.value.code == Nn.TBK_E_INVALID and "non-finite" in str(e.value)
        # one off-diagonal entry changed: no longer Hermitian
        off = next(p for j in range(H.n) for p in range(H.colptr[j], H.colptr[j + 1]) if H.rowidx[p] != j)
        v = H.values.copy()
        v[off] += 0.25
        with pytest.raises(Nn.NativeError) as e:
            srv.submit_dos(_with_values(H, v), 32, a, 0.0, 8, 1)
        assert e.value.code == Nn.TBK_E_INVALID and "Hermitian" in str(e.value)
        v = H.values.copy()
        v[off] *= 1 + 1e-14
        t = srv.submit_dos(_with_values(H, v), 32, a, 0.0, 8, 1)
        assert srv.wait(t, 60.0)["state"] == Nn.TBK_REQ_DONE
        srv.take(t)
        # sigma_xy: a non-finite velocity value
        w = vx.values.copy()
        w[5] = np.nan
        with pytest.raises(Nn.NativeError) as e:
            srv.submit_kubo(Hs, _with_values(vx, w), vy, g.size, R=4, seed=1, g=g)
        assert e.value.code == Nn.TBK_E_INVALID
