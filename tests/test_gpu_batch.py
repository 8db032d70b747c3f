"""Batched ragged moment problems (tbk_batch_*) on the GPU.

Each problem of a batch must give the moments of the single-problem path
(same vectors T_n, moment sums in another order: within 1e-12 relative in
fp64), the oracle's moments within 1e-10, the same bits alone and in any
mix, the same bits with and without hipGraph capture and across graph
replays, and agree across launch shapes."""
import numpy as np
import pytest

import oracle
from oracle import lattice_oracle as O
from tight_binding_amd import kpm
from tight_binding_amd import lattice as L

pytestmark = pytest.mark.gpu

COORD = {"square": 4, "honeycomb": 3, "lieb": 4, "kagome": 4}


def model_problem(model, Lc, W, seed, t=1.0):
    c = L.UnitCell.model(model)
    c.add_hopping_nn("A", 1, t)
    if model == "kagome":
        c.add_hopping_nn("B", 1, t)  # the B-C bonds (A's shell holds A-B and A-C)
    lat = L.PartialLattice(c, (Lc, Lc))
    if W:
        lat.set_disorder(W, seed)
    H = lat.get_Hamiltonian_periodic_extend_bare_sparse((1, 1))
    a = 1.02 * (COORD[model] * abs(t) + W / 2)
    return H, a


def mixed_batch(n=7, base_seed=11):
    rng = np.random.default_rng(base_seed)
    models = ["square", "honeycomb", "lieb", "kagome"]
    probs = []
    for i in range(n):
        m = models[i % 4]
        H, a = model_problem(m, int(rng.integers(5, 23)), float(rng.uniform(0, 3)), seed=100 + i)
        probs.append((H, a, 0.1 * (i % 3) - 0.1, int(rng.choice([2, 3, 30, 64, 130, 257])), 1000 + i))
    return probs


def make_batch(ctx, probs, R, precision=kpm.FP64):
    return kpm.Batch(ctx, [p[0] for p in probs], [p[1] for p in probs], [p[2] for p in probs],
                     [p[3] for p in probs], R, precision, seeds=[p[4] for p in probs])


def single(ctx, H, a, b, n, seed, R, precision=kpm.FP64):
    M = ctx.matrix(H)
    K = ctx.random_block(H.n, R, seed=seed, precision=precision)
    mu, _, _ = ctx.moments(M, K, n, a=a, b=b)
    return mu, K


def test_batch_matches_single_problem_path(gpu_ctx):
    probs = mixed_batch()
    B = make_batch(gpu_ctx, probs, R=4)
    mus, st = B.run()
    assert st["launches"] == max((p[3] + 1) // 2 for p in probs)
    for (H, a, b, n, seed), mu in zip(probs, mus):
        ref, _ = single(gpu_ctx, H, a, b, n, seed, 4)
        assert mu.size == n
        assert np.max(np.abs(mu - ref)) <= 1e-12 * np.max(np.abs(ref)), (n, np.max(np.abs(mu - ref)))
    B.close()


def test_batch_matches_oracle(gpu_ctx):
    probs = mixed_batch(5, base_seed=3)
    B = make_batch(gpu_ctx, probs, R=3)
    mus, _ = B.run()
    for (H, a, b, n, seed), mu in zip(probs, mus):
        K = gpu_ctx.random_block(H.n, 3, seed=seed)
        kets = K.download()
        Hs = O.rescale(O.CSC(H.n, H.colptr, H.rowidx, H.values), a, b)
        ref = oracle.cheb_braket(Hs, n, kets, kets)
        assert np.max(np.abs(mu - ref)) <= 1e-10 * np.max(np.abs(ref))
    B.close()


def test_alone_and_in_mix_are_bit_identical(gpu_ctx):
    probs = mixed_batch(6, base_seed=5)
    Bm = make_batch(gpu_ctx, probs, R=8)
    mix, _ = Bm.run()
    Bm.close()
    for i in (0, 3, 5):
        Ba = make_batch(gpu_ctx, [probs[i]], R=8)
        alone, _ = Ba.run()
        Ba.close()
        assert np.array_equal(alone[0], mix[i])
    # another mix with the problems reordered and one repeated
    order = [4, 1, 1, 0, 3]
    Br = make_batch(gpu_ctx, [probs[j] for j in order], R=8)
    mix2, _ = Br.run()
    Br.close()
    for k, j in enumerate(order):
        assert np.array_equal(mix2[k], mix[j])


def test_graph_replays_match_eager(gpu_ctx):
    """Graph capture per 64-step chunk: the first (capturing) run, two
    replays with eager work and a synchronisation between them, and the eager
    run all give the same bits."""
    probs = mixed_batch(5, base_seed=9)
    B = make_batch(gpu_ctx, probs, R=4)
    eager, st0 = B.run(graphs=False)
    assert st0["graphs"] == 0
    first, st1 = B.run(graphs=True)
    assert st1["graphs"] == (st1["launches"] + 63) // 64
    for rep in range(2):
        # unrelated eager work on the same stream between replays
        H, a = model_problem("square", 9, 1.0, 2)
        single(gpu_ctx, H, a, 0.0, 20, 3, 4)
        gpu_ctx.synchronize()
        again, _ = B.run(graphs=True)
        for x, y, z in zip(eager, first, again):
            assert np.array_equal(x, y) and np.array_equal(x, z)
    B.close()


@pytest.mark.parametrize("R", [1, 8, 16, 32])
def test_launch_shapes_agree(gpu_ctx, R):
    probs = mixed_batch(4, base_seed=R)
    B = make_batch(gpu_ctx, probs, R=R)
    ref, _ = B.run()
    Rp = 1 << (R - 1).bit_length()
    for cpl in (1, 2, 4):
        if Rp % cpl or Rp // cpl > 64 or (cpl == 4 and Rp < 8) or (cpl == 2 and Rp < 2):
            continue
        for bpc in (2, 8):
            B.set_shape(cpl, bpc)
            mus, st = B.run(graphs=True)
            assert st["cpl"] == cpl and st["blocks_per_cu"] == bpc
            for m, r in zip(mus, ref):
                assert np.max(np.abs(m - r)) <= 1e-12 * np.max(np.abs(r))
    B.close()


def test_autotune_and_fp32_batch(gpu_ctx):
    probs = mixed_batch(6, base_seed=21)
    B64 = make_batch(gpu_ctx, probs, R=8)
    ref, _ = B64.run()
    B64.close()
    B32 = make_batch(gpu_ctx, probs, R=8, precision=kpm.FP32)
    ms = B32.autotune(steps=4)
    assert ms > 0
    mus, st = B32.run(graphs=True)
    assert st["cpl"] in (1, 2, 4)
    for m, r in zip(mus, ref):
        assert np.max(np.abs(m - r)) <= 1e-5 * np.max(np.abs(r))
    # fp32 batch = fp32 single-problem path to moment-sum rounding
    H, a, b, n, seed = probs[2]
    s32, _ = single(gpu_ctx, H, a, b, n, seed, 8, kpm.FP32)
    assert np.max(np.abs(mus[2] - s32)) <= 1e-12 * np.max(np.abs(s32))
    B32.close()


def test_batch_errors(gpu_ctx):
    from tight_binding_amd import _native as N

    H, a = model_problem("square", 6, 0.0, 1)
    Hc = L.PartialLattice(L.UnitCell.model("square").add_hopping_nn("A", 1, 1.0), (6, 6)).get_Hamiltonian_sparse(True)
    with pytest.raises(N.NativeError) as e:
        kpm.Batch(gpu_ctx, [H, Hc], [a, a], [0, 0], [8, 8], 4)
    assert e.value.code == N.TBK_E_INVALID
    with pytest.raises(N.NativeError) as e:
        kpm.Batch(gpu_ctx, [H], [a], [0], [1], 4)
    assert e.value.code == N.TBK_E_INVALID
    B = kpm.Batch(gpu_ctx, [H], [a], [0], [8], 8)
    with pytest.raises(N.NativeError) as e:
        B.set_shape(3, 4)
    assert e.value.code == N.TBK_E_UNSUPPORTED
    B.close()


def test_persistent_batch_matches_oracle_and_per_step_launches(gpu_ctx):
    """TBK_BATCH_PERSISTENT: each chunk of 64 steps as one persistent launch
    (problems on XCD queues, items in step order).  The vectors are the
    per-step launches' (the same step_row sums), the moments summed in
    another grouping: within 1e-12 of the per-step batch and 1e-10 of the
    oracle, over step counts that span several chunks (257 moments = 129
    steps) and problems larger than one 4096-row share."""
    probs = mixed_batch(7, base_seed=5)
    H, a = model_problem("kagome", 40, 1.3, seed=77)  # 4800 rows: two shares
    probs.append((H, a, 0.05, 257, 4242))
    B = make_batch(gpu_ctx, probs, R=3)
    mus_p, st = B.run(persistent=True)
    assert st["launches"] == (max((p[3] + 1) // 2 for p in probs) + 63) // 64
    mus_s, _ = B.run()
    for (H, a, b, n, seed), mp, ms in zip(probs, mus_p, mus_s):
        assert np.max(np.abs(mp - ms)) <= 1e-12 * np.max(np.abs(ms)), n
        K = gpu_ctx.random_block(H.n, 3, seed=seed)
        kets = K.download()
        Hs = O.rescale(O.CSC(H.n, H.colptr, H.rowidx, H.values), a, b)
        ref = oracle.cheb_braket(Hs, n, kets, kets)
        assert np.max(np.abs(mp - ref)) <= 1e-10 * np.max(np.abs(ref)), n
        K.close()
    B.close()


@pytest.mark.parametrize("precision", [kpm.FP64, kpm.FP32])
def test_persistent_batch_alone_mixed_and_graphs_are_the_same_bits(gpu_ctx, precision):
    """A problem's persistent-launch moments do not depend on the rest of
    the batch (its own XCD queue placement changes, its partial grouping
    does not), nor on hipGraph capture: the same bits alone, in a mix, and
    over several graph replays with eager runs between them."""
    probs = mixed_batch(9, base_seed=8)
    B = make_batch(gpu_ctx, probs, R=5, precision=precision)
    mixed, _ = B.run(persistent=True)
    reps = []
    for _ in range(3):
        g, st = B.run(graphs=True, persistent=True)
        assert st["graphs"] == st["launches"]
        reps.append(g)
        e, _ = B.run(persistent=True)
        reps.append(e)
    for r in reps:
        for x, y in zip(mixed, r):
            assert np.array_equal(x, y)
    for i in (0, 4, 8):
        A = make_batch(gpu_ctx, [probs[i]], R=5, precision=precision)
        alone, _ = A.run(persistent=True)
        assert np.array_equal(alone[0], mixed[i]), i
        A.close()
    B.close()


@pytest.mark.parametrize("precision", [kpm.FP64, kpm.FP32])
def test_synced_batch_matches_per_step_launches_alone_mixed_and_graphs(gpu_ctx, precision):
    """TBK_BATCH_SYNCED: each chunk of 64 steps as one persistent launch over
    the whole GPU, grid barriers between steps, doing the per-step launches'
    rows (the same vectors) with one moment partial per wave: moments within
    1e-12 (fp64) / 1e-6 (fp32) of the per-step launches, (fp64) within 1e-10
    of the oracle, over step counts that span several chunks (257 moments =
    129 steps); the same bits eager and over several graph replays with eager
    runs between them, and alone or in the mix."""
    from tight_binding_amd import _native as N

    probs = mixed_batch(9, base_seed=8)
    H, a = model_problem("kagome", 40, 1.3, seed=77)
    probs.append((H, a, 0.05, 257, 4242))
    B = make_batch(gpu_ctx, probs, R=5, precision=precision)
    ref, _ = B.run()
    y, st = B.run(synced=True)
    assert st["launches"] == (max((p[3] + 1) // 2 for p in probs) + 63) // 64
    assert st["workgroups"] % 8 == 0 and st["workgroups"] > 0
    tol = 1e-12 if precision == kpm.FP64 else 1e-6
    for x, z in zip(ref, y):
        assert np.max(np.abs(x - z)) <= tol * np.max(np.abs(x))
    for _ in range(3):
        g, stg = B.run(graphs=True, synced=True)
        assert stg["graphs"] == stg["launches"]
        e, _ = B.run(synced=True)
        for r in (g, e):
            for x, z in zip(y, r):
                assert np.array_equal(x, z)
    for i in (0, 4, 9):
        A = make_batch(gpu_ctx, [probs[i]], R=5, precision=precision)
        alone, _ = A.run(synced=True)
        assert np.array_equal(alone[0], y[i]), i
        A.close()
    if precision == kpm.FP64:
        for i in (0, 5, 9):
            Hh, ah, bh, n, seed = probs[i]
            K = gpu_ctx.random_block(Hh.n, 5, seed=seed)
            kets = K.download()
            Hs = O.rescale(O.CSC(Hh.n, Hh.colptr, Hh.rowidx, Hh.values), ah, bh)
            want = oracle.cheb_braket(Hs, n, kets, kets)
            assert np.max(np.abs(y[i] - want)) <= 1e-10 * np.max(np.abs(want)), i
            K.close()
    with pytest.raises(N.NativeError) as err:
        B.run(persistent=True, synced=True)
    assert err.value.code == N.TBK_E_INVALID
    B.close()


def test_synced_batch_follows_a_new_launch_shape(gpu_ctx):
    """A synchronised run after the launch shape changed (set_shape: more
    workgroups per problem) uses a plan rebuilt for the new shape: its
    moments again match the per-step launches of that shape."""
    probs = mixed_batch(6, base_seed=21)
    B = make_batch(gpu_ctx, probs, R=4)
    B.set_shape(1, 4)
    B.run(synced=True)
    for cpl, bpc in ((1, 8), (2, 4)):
        B.set_shape(cpl, bpc)
        ref, _ = B.run()
        y, _ = B.run(synced=True)
        g, _ = B.run(graphs=True, synced=True)
        for x, z, w in zip(ref, y, g):
            assert np.max(np.abs(x - z)) <= 1e-12 * np.max(np.abs(x))
            assert np.array_equal(z, w)
    B.close()


@pytest.mark.parametrize("precision", [kpm.FP64, kpm.FP32])
def test_persistent_forms_with_complex_hamiltonians(gpu_ctx, precision):
    """Complex Hamiltonians (open square and honeycomb lattices in a Landau
    gauge) through both persistent forms: moments within 1e-12 (fp64) /
    1e-6 (fp32) of the per-step launches, and (fp64) within 1e-10 of the
    oracle."""
    rng = np.random.default_rng(31)
    probs = []
    for i, (model, Lc) in enumerate((("square", 14), ("honeycomb", 9), ("square", 21), ("honeycomb", 17))):
        c = L.UnitCell.model(model)
        c.add_hopping_nn("A", 1, 1.0)
        lat = L.PartialLattice(c, (Lc, Lc + 2))
        lat.set_gauge(L.GAUGE_LANDAU, 2 * np.pi / 7)
        lat.set_disorder(float(rng.uniform(0, 2)), 40 + i)
        H = lat.get_Hamiltonian_sparse(True)
        probs.append((H, 1.02 * (COORD[model] + 1.0), 0.0, [64, 130, 257, 30][i], 600 + i))
    B = make_batch(gpu_ctx, probs, R=6, precision=precision)
    ref, _ = B.run()
    tol = 1e-12 if precision == kpm.FP64 else 1e-6
    for run in (dict(persistent=True), dict(synced=True), dict(graphs=True, synced=True)):
        got, _ = B.run(**run)
        for x, z in zip(ref, got):
            assert np.max(np.abs(x - z)) <= tol * np.max(np.abs(x)), run
    if precision == kpm.FP64:
        got, _ = B.run(synced=True)
        for (Hh, ah, bh, n, seed), y in zip(probs, got):
            K = gpu_ctx.random_block(Hh.n, 6, seed=seed)
            kets = K.download()
            Hs = O.rescale(O.CSC(Hh.n, Hh.colptr, Hh.rowidx, Hh.values), ah, bh)
            want = oracle.cheb_braket(Hs, n, kets, kets)
            assert np.max(np.abs(y - want)) <= 1e-10 * np.max(np.abs(want))
            K.close()
    B.close()
