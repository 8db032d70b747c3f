"""Benchmark of the Chebyshev moment recursion on one MI355X.

Workload (default): graphene honeycomb, 1024 x 1024 cells periodic
(2,097,152 sites), t = -2.7 eV, on-site disorder uniform in [-0.25, 0.25] eV,
R = 64 random-phase vectors as one block, N = 2048 moments, fp64 recursion
(complex128 vectors, real fp64 H).  H and the vector block are resident in
HBM before the timed region.

A step is one application of H~ = (H - b)/a to the whole 64-vector block,
with the three-term update and the moment partials fused (one kernel
launch).  With bras = kets each step yields two moments
(mu_{2k} = 2<T_k|T_k> - mu_0, mu_{2k+1} = 2<T_{k+1}|T_k> - mu_1), so N = 2048
moments take 1024 steps.  --steps K times exactly K steps (2K moments).

Prints one JSON line.  roofline.achieved uses the algorithmic bytes of a
step: nnz*(8+4) + (n+1)*4 (CSR matrix) + 3*n*R*16 (read T_k, read T_{k-1},
write T_{k+1}).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# DRAM bytes per step of the default workload (1024 x 1024 cells, R = 64)
# from rocprofv3 --pmc: TCC_EA0_RDREQ_DRAM_32B and TCC_EA0_WRREQ_WRITE_DRAM_32B
# count 32-byte sectors and read 1.000x the traffic of stream kernels of known
# size.  A figure holds for the launch shape it was profiled at only (other
# shapes read 4.4-7.3 GB per step), so it is reported only when the tuner
# installs one of the profiled shapes.  The tuner's two usual choices differ
# in xcd_split only (row ranges dealt per XCD, or not); both were profiled
# in one run: [(bytes, shape, profile summary)].
_C2_SHAPE = {"blocks_per_cu": (0, 4), "chunk_rows": (0, 256), "chunk_period": (0,), "stream_nt": (0, 1),
             "use_tiled": (0,), "row_lane_bytes": (0, 16), "row_slab_cols": (0,), "slab_cols": (0,)}
STEP_TRAFFIC = {
    (1024, 64): [
        (7.794e9, dict(_C2_SHAPE, xcd_split=(0, 1)), "profiles/cheb_step_fp64_honeycomb2M_R64_pmc_counters.md"),
        (8.016e9, dict(_C2_SHAPE, xcd_split=(-1,)), "profiles/cheb_step_fp64_honeycomb2M_R64_pmc_counters.md"),
    ],
}


def step_traffic(cells, R, tuning):
    """(bytes, source) of the profiled DRAM traffic when `tuning` is one of
    the profiled launch shapes, else (None, reason)."""
    entries = STEP_TRAFFIC.get((cells, R))
    if not entries:
        return None, None
    t = tuning or {}
    for nbytes, shape, source in entries:
        if all(t.get(k, 0) in allowed for k, allowed in shape.items()):
            return nbytes, ("rocprofv3 --pmc DRAM sector counters of this workload at this launch shape (%s)"
                            % source)
    return None, "not profiled at this launch shape (%s holds the profiled shapes)" % entries[0][2]


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def build_c2(L=1024, t=-2.7, W=0.5, seed=2026):
    from tight_binding_amd import lattice as Lm

    cell = Lm.UnitCell.model("honeycomb")
    cell.add_hopping_nn("A", 1, t)
    lat = Lm.PartialLattice(cell, (L, L))
    lat.set_disorder(W, seed)
    H = lat.get_Hamiltonian_periodic_extend_bare_sparse((1, 1))
    # Gershgorin bounds, contracted like the reference's (E_max-E_min)/(2-eps)
    emax = 3 * abs(t) + W / 2
    a = (2 * emax) / (2 - 0.01)
    return H, a


def build_c3(L=512, W=2.0, flux=1.0 / 64, seed=2027):
    """Square L x L open, |t| = 1, Landau-gauge field of `flux` quanta per
    plaquette, on-site disorder uniform in [-W/2, W/2]; complex H and the
    position operators X, Y."""
    from tight_binding_amd import lattice as Lm

    cell = Lm.UnitCell.model("square")
    cell.add_hopping_nn("A", 1, 1.0)
    lat = Lm.PartialLattice(cell, (L, L))
    lat.set_gauge(Lm.GAUGE_LANDAU, 2 * math.pi * flux)
    lat.set_disorder(W, seed)
    H = lat.get_Hamiltonian_sparse(is_complex=True)
    return H, lat.op_position(0), lat.op_position(1), 4.0 + W / 2


MFMA_PEAK_TFLOPS = {"fp64": 78.6, "bf16": 2500.0}


def kubo_leg(ctx, cells, N, R, repeats):
    """sigma_xy solves of the Kubo-Bastin workload with H, v_x, v_y resident:
    each solve = random vectors, both stacks, the Gamma contraction, zeta at
    the N Chebyshev nodes (device) and the integral to 201 Fermi energies."""
    from tight_binding_amd import kpm
    from tight_binding_amd import kubo

    t0 = time.perf_counter()
    H, X, Y, emax = build_c3(cells)
    solver = kubo.KuboSolver(ctx, H, X, Y, -1.02 * emax, 1.02 * emax, volume=float(H.n), lam=6.0)
    log("kubo: assembled and uploaded n=%d nnz=%d a=%.4f in %.1f s" % (H.n, H.nnz, solver.a, time.perf_counter() - t0))
    ef = np.linspace(-4.0, 4.0, 201)
    solver.solve(16, R=R, seed=1, fermi_energies=ef)  # loads the kernels
    saved = ctx.get_tuning()
    tuning, tms = solver.tune(R)  # launch shape of the stack recursion (restored after the leg)
    log("kubo: recursion tuning %s, %.3f ms per step" % (json.dumps(tuning), tms))
    out = {
        "workload": "Kubo-Bastin sigma_xy, square %dx%d open (%d sites), Landau flux 1/64 per plaquette, W=2, "
                    "R=%d, N=%d, Lorentz lambda=6, 201 Fermi energies in [-4, 4]" % (cells, cells, H.n, R, N),
        "sites": H.n, "R": R, "N": N, "solves_per_precision": repeats,
    }
    sigmas = {}
    for prec, name, unit in ((kpm.GAMMA_FP64, "fp64", "fp64"), (kpm.GAMMA_BF16X3, "bf16x3", "bf16"),
                             (kpm.GAMMA_BF16X3_FP32, "bf16x3_fp32rec", "bf16")):
        lat, rec_ms, con_ms, sig_s = [], [], [], []
        for r in range(repeats):
            res = solver.solve(N, R=R, seed=100 + r, precision=prec, fermi_energies=ef)
            st = res["stats"]
            lat.append(res["timings"]["total_s"])
            rec_ms.append(st["recursion_ms"])
            con_ms.append(st["contraction_ms"])
            sig_s.append(res["timings"]["sigma_s"])
        sigmas[name] = res["sigma"]
        # st["flops"] = 8 N^2 K, the complex GEMM.  Issued MFMA flops: split-bf16
        # makes three products of every real product; fp64 forms each complex
        # product from three real ones (Gauss), 6 N^2 K
        flops = st["flops"] * (3.0 if unit == "bf16" else 0.75)
        c = float(np.median(con_ms))
        r_ms = float(np.median(rec_ms))
        out[name] = {
            "latency_s_p50": float(np.percentile(lat, 50)),
            "latency_s_p99": float(np.percentile(lat, 99)),
            "latency_s_all": lat,
            "recursion_ms": r_ms,
            "recursion_GBps": st["recursion_bytes"] / (r_ms * 1e-3) / 1e9,
            "recursion_frac_hbm": st["recursion_bytes"] / (r_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "contraction_ms": c,
            "contraction_TFLOPs_issued": flops / (c * 1e-3) / 1e12,
            "contraction_TFLOPs_complex_gemm": st["flops"] / (c * 1e-3) / 1e12,
            "mfma_frac": flops / (c * 1e-3) / 1e12 / MFMA_PEAK_TFLOPS[unit],
            "zeta_and_integral_ms": 1e3 * float(np.median(sig_s)),
            "stack_GB": st["stack_bytes"] / 1e9,
            "recursion": "fp32" if st["recursion_fp32"] else "fp64",
        }
        log("kubo %s: %s" % (name, json.dumps(out[name])))
    s64 = sigmas["fp64"]
    for name in ("bf16x3", "bf16x3_fp32rec"):
        out[name]["sigma_rel_err_vs_fp64"] = float(np.max(np.abs(sigmas[name] - s64)) / np.max(np.abs(s64)))
    out["sigma_fp64_sample"] = [float(x) for x in s64[::25]]
    out["launch_tuning"] = tuning
    ctx.set_tuning(**saved)
    solver.close()
    return out


BATCH_COORD = {"square": 4, "honeycomb": 3, "lieb": 4, "kagome": 4}


def build_batch_problems(n, seed=2028, lmin=32, lmax=512):
    """n problems: model uniform over square, honeycomb, Lieb, Kagome
    (nearest-neighbour t = 1, periodic), L ~ U{lmin..lmax} cells per side,
    on-site disorder uniform in [-W/2, W/2] with W ~ U[0, 3], N moments from
    {256, 512, 1024}; a from the Gershgorin bound.  Returns
    [(H, a, N, seed, model, L, W)]."""
    from tight_binding_amd import lattice as Lm

    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        model = ["square", "honeycomb", "lieb", "kagome"][int(rng.integers(0, 4))]
        Lc = int(rng.integers(lmin, lmax + 1))
        W = float(rng.uniform(0.0, 3.0))
        N = int(rng.choice([256, 512, 1024]))
        c = Lm.UnitCell.model(model)
        c.add_hopping_nn("A", 1, 1.0)
        if model == "kagome":
            c.add_hopping_nn("B", 1, 1.0)
        lat = Lm.PartialLattice(c, (Lc, Lc))
        lat.set_disorder(W, seed + i)
        H = lat.get_Hamiltonian_periodic_extend_bare_sparse((1, 1))
        out.append((H, 1.02 * (BATCH_COORD[model] + W / 2), N, 5000 + i, model, Lc, W))
    return out


def batch_leg(ctx, nprob, R=8):
    """Batched ragged DoS: nprob problems in one fp32 batch (fp64 moments),
    autotuned launch shape, each 64-step chunk a hipGraph; against the plain
    per-solve GPU path (one tbk_moments call per problem) on the same inputs."""
    from tight_binding_amd import kpm

    t0 = time.perf_counter()
    probs = build_batch_problems(nprob)
    sites = sum(p[0].n for p in probs)
    moments = sum(p[2] for p in probs)
    log("batch: %d problems, %d sites, %d moments, assembled in %.1f s" % (nprob, sites, moments,
                                                                          time.perf_counter() - t0))
    t0 = time.perf_counter()
    B = kpm.Batch(ctx, [p[0] for p in probs], [p[1] for p in probs], 0.0, [p[2] for p in probs], R, kpm.FP32,
                  seeds=[p[3] for p in probs])
    create_s = time.perf_counter() - t0
    tune_ms = B.autotune(steps=6)
    B.run(graphs=True)  # captures the chunk graphs
    runs = []
    for _ in range(2):
        mus, st = B.run(graphs=True)
        runs.append(st)
    st = min(runs, key=lambda d: d["kernel_ms"])
    gbps = st["bytes_total"] / (st["kernel_ms"] * 1e-3) / 1e9
    log("batch: kernel %.1f ms, wall %.1f ms, %.1f GB/s, shape cpl=%d bpc=%d, %d workgroups (create %.1f s, "
        "tune %.3f ms/step)" % (st["kernel_ms"], st["wall_ms"], gbps, st["cpl"], st["blocks_per_cu"], st["workgroups"],
                                create_s, tune_ms))
    # the same batch as persistent launches (one per 64-step chunk, problems
    # on XCD queues), each chunk a hipGraph
    mus_p, _ = B.run(graphs=True, persistent=True)
    pruns = [B.run(graphs=True, persistent=True)[1] for _ in range(2)]
    stp = min(pruns, key=lambda d: d["kernel_ms"])
    gbps_p = stp["bytes_total"] / (stp["kernel_ms"] * 1e-3) / 1e9
    dmu_p = max(float(np.max(np.abs(a - b)) / np.max(np.abs(b))) for a, b in zip(mus_p, mus))
    log("batch persistent: kernel %.1f ms, wall %.1f ms, %.1f GB/s, %d launches of %d workgroups, "
        "max |dmu|/max|mu| vs per-step launches %.2e" % (stp["kernel_ms"], stp["wall_ms"], gbps_p, stp["launches"],
                                                         stp["workgroups"], dmu_p))
    # one grid-synchronised persistent launch per chunk (TBK_BATCH_SYNCED), at
    # the tuned shape and at 4 values per lane, 4 workgroups per CU (its best
    # on this batch); the tuned shape is restored afterwards
    tuned_shape = (st["cpl"], st["blocks_per_cu"])
    sty, mus_y = None, None
    for shape in dict.fromkeys([tuned_shape, (4, 4)]):
        B.set_shape(*shape)
        m_y, _ = B.run(synced=True)
        for _ in range(2):
            r = B.run(graphs=True, synced=True)[1]
            if sty is None or r["kernel_ms"] < sty["kernel_ms"]:
                sty, mus_y = dict(r, shape=shape), m_y
    B.set_shape(*tuned_shape)
    gbps_y = sty["bytes_total"] / (sty["kernel_ms"] * 1e-3) / 1e9
    dmu_y = max(float(np.max(np.abs(a - b)) / np.max(np.abs(b))) for a, b in zip(mus_y, mus))
    log("batch synced: kernel %.1f ms, wall %.1f ms, %.1f GB/s, %d launches of %d workgroups, "
        "max |dmu|/max|mu| vs per-step launches %.2e" % (sty["kernel_ms"], sty["wall_ms"], gbps_y, sty["launches"],
                                                         sty["workgroups"], dmu_y))
    B.close()

    # plain per-solve path: one moments call per problem, matrices resident
    mats = [ctx.matrix(p[0]) for p in probs]
    ctx.synchronize()
    k_ms, w_s = 0.0, 0.0
    errs = []
    for i, (p, M) in enumerate(zip(probs, mats)):
        K = ctx.random_block(p[0].n, R, seed=p[3], precision=kpm.FP32)
        ctx.synchronize()
        t1 = time.perf_counter()
        mu1, s1, _ = ctx.moments(M, K, p[2], a=p[1])
        w_s += time.perf_counter() - t1
        k_ms += s1["kernel_ms"]
        K.close()
        if i < 4:  # fp32 error against the fp64 recursion on the same vectors
            K64 = ctx.random_block(p[0].n, R, seed=p[3])
            mu64, _, _ = ctx.moments(M, K64, p[2], a=p[1])
            errs.append(float(np.max(np.abs(mus[i] - mu64)) / np.max(np.abs(mu64))))
            assert np.max(np.abs(mus[i] - mu1)) <= 1e-9 * np.max(np.abs(mu1))
            K64.close()
    for M in mats:
        M.close()
    out = {
        "workload": "%d independent DoS problems (square/honeycomb/Lieb/Kagome, L~U{32..512}, W~U[0,3], "
                    "N in {256,512,1024}), R=%d, fp32 recursion with fp64 moments" % (nprob, R),
        "sites_total": sites, "moments_total": moments,
        "batched_kernel_ms": st["kernel_ms"], "batched_wall_ms": st["wall_ms"],
        "batched_moments_per_s": moments / (st["wall_ms"] * 1e-3),
        "batched_GBps": gbps, "batched_frac_hbm": gbps / HBM_PEAK_GBS,
        "launch_shape": {"cpl": st["cpl"], "blocks_per_cu": st["blocks_per_cu"], "workgroups": st["workgroups"]},
        "graphs_per_run": st["graphs"], "step_launches": st["launches"],
        "persistent": {"kernel_ms": stp["kernel_ms"], "wall_ms": stp["wall_ms"], "GBps": gbps_p,
                       "frac_hbm": gbps_p / HBM_PEAK_GBS, "launches": stp["launches"],
                       "workgroups": stp["workgroups"], "moments_per_s": moments / (stp["wall_ms"] * 1e-3),
                       "max_rel_dmu_vs_per_step_launches": dmu_p},
        "synced": {"kernel_ms": sty["kernel_ms"], "wall_ms": sty["wall_ms"], "GBps": gbps_y,
                   "shape": {"cpl": sty["shape"][0], "blocks_per_cu": sty["shape"][1]},
                   "frac_hbm": gbps_y / HBM_PEAK_GBS, "launches": sty["launches"], "workgroups": sty["workgroups"],
                   "moments_per_s": moments / (sty["wall_ms"] * 1e-3), "max_rel_dmu_vs_per_step_launches": dmu_y},
        "per_solve_kernel_ms": k_ms, "per_solve_wall_ms": 1e3 * w_s,
        "per_solve_moments_per_s": moments / w_s,
        "speedup_wall": (1e3 * w_s) / st["wall_ms"],
        "fp32_max_rel_err_vs_fp64_sample": max(errs) if errs else None,
    }
    log("batch: %s" % json.dumps(out))
    return out


SERVE_MODELS = {"square": (2, 4), "honeycomb": (2, 3), "lieb": (2, 4), "kagome": (2, 4), "cubic": (3, 6),
                "fcc": (3, 12)}


SERVE_ATOMS = {"honeycomb": 2, "lieb": 3, "kagome": 3}
SERVE_BANDS = ((1e3, 1e4), (1e4, 1e5), (1e5, 1e6), (1e6, 4.2e6))
SERVE_SIGMA_FLOP_CAP = 4e14


def make_requests(n, qps, seed=2030, sigma_flop_cap=SERVE_SIGMA_FLOP_CAP):
    """n served requests, each carrying its own lattice: Poisson arrivals at
    `qps`; 70% DoS, 30% sigma_xy; model uniform over square, honeycomb,
    Lieb, Kagome, simple cubic and FCC (nearest neighbours, t = 1); sites
    log-uniform in [1e3, 4e6] (L = round((sites/atoms)^(1/dim)) cells a side);
    W ~ U[0, 2]; N ~ U{256..2048}; R in {8, 16, 32}; half fp64, half mixed
    precision (fp32 recursion / split-bf16 Gamma).  DoS lattices are
    periodic; sigma_xy lattices are open with a Landau gauge A = (0, B x, 0),
    B = 2 pi / 16.  A sigma_xy draw whose contraction 8 N^2 n R exceeds
    sigma_flop_cap is drawn again (the returned count says how often): the
    largest sigma_xy of the distribution, 4e6 sites at N = 2048 and R = 32,
    is 4.3e18 flop, over an hour at the bf16 MFMA peak."""
    rng = np.random.default_rng(seed)
    models = list(SERVE_MODELS)
    t = 0.0
    out, redrawn = [], 0
    for i in range(n):
        t += float(rng.exponential(1.0 / qps))
        dos = bool(rng.uniform() < 0.7)
        mixed = bool(rng.uniform() < 0.5)
        while True:
            model = models[int(rng.integers(0, len(models)))]
            dim, z = SERVE_MODELS[model]
            atoms = SERVE_ATOMS.get(model, 1)
            target = math.exp(rng.uniform(math.log(1e3), math.log(4e6)))
            Lc = max(2, int(round((target / atoms) ** (1.0 / dim))))
            sites = atoms * Lc ** dim
            W = float(rng.uniform(0.0, 2.0))
            N = int(rng.integers(256, 2049))
            R = int(rng.choice([8, 16, 32]))
            if dos or 8.0 * N * N * sites * R <= sigma_flop_cap:
                break
            redrawn += 1
        out.append(dict(kind="dos" if dos else "kubo", t=t, model=model, dim=dim, L=Lc, sites=sites, W=W, N=N, R=R,
                        mixed=mixed, seed=7000 + i, a=1.02 * (z + W / 2)))
    return out, redrawn


def serve_lattice(r):
    """The lattice request r carries (disorder keyed by the request's seed)."""
    from tight_binding_amd import lattice as Lm

    c = Lm.UnitCell.model(r["model"])
    c.add_hopping_nn("A", 1, 1.0)
    if r["model"] == "kagome":
        c.add_hopping_nn("B", 1, 1.0)
    lat = Lm.PartialLattice(c, (r["L"],) * r["dim"])
    if r["kind"] == "kubo":
        lat.set_gauge(Lm.GAUGE_LANDAU, 2 * math.pi / 16)
    lat.set_disorder(r["W"], r["seed"] + 100000)
    return lat


def percentiles_by(lat, reqs, key):
    """p50 / p99 latency (ms) and count of the requests in each group."""
    out = {}
    for name in sorted({key(r) for r in reqs}):
        sel = np.array([key(r) == name for r in reqs])
        if sel.any():
            x = lat[sel]
            out[name] = {"n": int(sel.sum()), "p50_ms": float(np.percentile(x, 50)),
                         "p99_ms": float(np.percentile(x, 99))}
    return out


def size_band(sites):
    for lo, hi in SERVE_BANDS:
        if lo <= sites < hi:
            return "%.0e-%.0e" % (lo, hi)
    return "other"


def serve_leg(qps, n_requests, saturation_only=False):
    """Served mixed workload: continuous batching in the request server, the
    requests carrying lattices that the server assembles on the device at
    admission (sigma_xy: H, H~ = H/a and v_x, v_y formed there too).  Latency
    percentiles at a fixed Poisson arrival rate (overall, per kind, per size
    band), the makespan of the same requests all submitted at once, and the
    plain per-solve GPU path on the same requests (device assembly, then one
    moments or Gamma call per request) for comparison."""
    import threading

    from tight_binding_amd import kpm
    from tight_binding_amd import lattice as Lm
    from tight_binding_amd import spectral as S

    reqs, redrawn = make_requests(n_requests, qps)
    lats = [serve_lattice(r) for r in reqs]
    nd = sum(r["kind"] == "dos" for r in reqs)
    log("serve: %d requests at %.1f QPS (%d DoS, %d sigma_xy, %.1fM sites in all, largest %d DoS / %d sigma_xy; "
        "%d sigma_xy draws above %.0e flop redrawn)"
        % (len(reqs), qps, nd, len(reqs) - nd, sum(r["sites"] for r in reqs) / 1e6,
           max([r["sites"] for r in reqs if r["kind"] == "dos"] or [0]),
           max([r["sites"] for r in reqs if r["kind"] == "kubo"] or [0]), redrawn, SERVE_SIGMA_FLOP_CAP))
    g_of = {}

    def gweights(N):
        return g_of.setdefault(N, S.lorentz(N, 6.0))

    def submit(srv, r, lat):
        if r["kind"] == "dos":
            return srv.submit_dos(lat, r["N"], r["a"], 0.0, r["R"], r["seed"], kpm.FP32 if r["mixed"] else kpm.FP64,
                                  builder=Lm.BUILD_PERIODIC, extend=(1,) * r["dim"])
        return srv.submit_kubo_lattice(lat, r["N"], r["a"], 0.0, R=r["R"], seed=r["seed"],
                                       precision=kpm.GAMMA_BF16X3_FP32 if r["mixed"] else kpm.GAMMA_FP64,
                                       g=gweights(r["N"]))

    created = {}

    def serve(arrivals, label):
        t_c = time.perf_counter()
        srv = kpm.Server(0, max_kubo=2, max_kubo_moments=2048)
        created[label] = time.perf_counter() - t_c
        tickets = [None] * len(reqs)
        start = time.perf_counter()

        def feeder():
            for i, r in enumerate(reqs):
                if arrivals:
                    dt = r["t"] - (time.perf_counter() - start)
                    if dt > 0:
                        time.sleep(dt)
                tickets[i] = submit(srv, r, lats[i])

        th = threading.Thread(target=feeder)
        th.start()
        th.join()
        results, lat, infos = [], [], []
        for i, t in enumerate(tickets):
            info = srv.wait(t, 600.0)
            if info["state"] != 2:
                raise SystemExit("bench: served request %d did not finish: %s" % (i, info))
            res, info = srv.take(t)
            results.append(res)
            infos.append(info)
            lat.append(info["latency_ms"])
            if (i + 1) % 16 == 0:
                log("serve: %s run, %d of %d requests done, %.1f s" % (label, i + 1, len(reqs),
                                                                        time.perf_counter() - start))
        span = time.perf_counter() - start
        st = srv.stats()
        srv.close()
        return results, np.asarray(lat), span, st, infos

    if saturation_only:
        res_s, lat_s, span_s, st_s, _ = serve(False, "saturation")
        log("serve: saturation run (all submitted at once) %.2f s" % span_s)
        return {"served_makespan_s": span_s, "requests": len(reqs), "chunks": st_s["chunks"]}
    res_p, lat_p, span_p, st_p, info_p = serve(True, "Poisson")
    log("serve: Poisson run %.1f s, latency p50 %.1f ms p99 %.1f ms" % (span_p, np.percentile(lat_p, 50),
                                                                       np.percentile(lat_p, 99)))
    res_s, lat_s, span_s, st_s, _ = serve(False, "saturation")
    log("serve: saturation run (all submitted at once) %.2f s" % span_s)

    # plain per-solve GPU path: device assembly and one call per request, in
    # the precision asked for
    ctx = kpm.Context(0)

    def plain_dos(r, lat, prec):
        M, _ = ctx.assemble(lat, Lm.BUILD_PERIODIC, (1,) * r["dim"], False)
        K = ctx.random_block(r["sites"], r["R"], seed=r["seed"], precision=prec)
        mu, _, _ = ctx.moments(M, K, r["N"], a=r["a"])
        K.close()
        M.close()
        return mu

    def plain_kubo(r, lat, prec):
        M, _ = ctx.assemble(lat, Lm.BUILD_OPEN, None, True)
        Hs = ctx.rescale(M, r["a"], 0.0)
        M.close()
        vx, vy = ctx.velocity(Hs, lat, 0), ctx.velocity(Hs, lat, 1)
        K = ctx.random_block(r["sites"], r["R"], seed=r["seed"])
        G, st = kpm.kubo_gamma(ctx, Hs, vx, vy, K, r["N"], precision=prec, g=gweights(r["N"]))
        z, _ = kpm.kubo_zeta(ctx, G, S.kubo_nodes(r["N"]))
        for m in (K, Hs, vx, vy):
            m.close()
        return z, G, st

    t1 = time.perf_counter()
    plain = []
    for i, (r, lat) in enumerate(zip(reqs, lats)):
        if r["kind"] == "dos":
            plain.append(plain_dos(r, lat, kpm.FP32 if r["mixed"] else kpm.FP64))
        else:
            plain.append(plain_kubo(r, lat, kpm.GAMMA_BF16X3_FP32 if r["mixed"] else kpm.GAMMA_FP64))
        if (i + 1) % 16 == 0:
            log("serve: plain path, %d of %d requests, %.1f s" % (i + 1, len(reqs), time.perf_counter() - t1))
    span_plain = time.perf_counter() - t1
    ctx.release_work()
    log("serve: plain per-solve path %.2f s" % span_plain)

    # checks: served == plain (same requests, same precision; Gamma to the
    # rounding of its summation order, which vector groups change)
    for r, a, b in zip(reqs, res_s, plain):
        if r["kind"] == "dos":
            assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(b)), r
        else:
            tol = 1e-5 if r["mixed"] else 1e-12
            assert np.max(np.abs(a[1] - b[1])) <= tol * np.max(np.abs(b[1])), r
    # mixed precision against fp64 on the same requests
    err_dos, err_sig = [], []
    for r, lat, a in zip(reqs, lats, res_s):
        if not r["mixed"]:
            continue
        if r["kind"] == "dos":
            mu64 = plain_dos(r, lat, kpm.FP64)
            err_dos.append(float(np.max(np.abs(a - mu64)) / np.max(np.abs(mu64))))
        elif len(err_sig) < 4:
            G64 = plain_kubo(r, lat, kpm.GAMMA_FP64)[1]
            ef = np.linspace(-0.7, 0.7, 201)
            s64 = S.kubo_sigma(G64, ef)[0]
            smx = S.kubo_sigma(a[1], ef)[0]
            err_sig.append(float(np.max(np.abs(smx - s64)) / np.max(np.abs(s64))))
    ctx.close()
    kinds = np.array([r["kind"] for r in reqs])
    moments = sum(r["N"] for r in reqs if r["kind"] == "dos")
    split = [dict(sites=r["sites"], N=r["N"], R=r["R"], vector_groups=inf["vector_groups"],
                  row_ranges=inf["row_ranges"])
             for r, inf in zip(reqs, info_p) if r["kind"] == "kubo" and inf["vector_groups"] * inf["row_ranges"] > 1]
    out = {
        "workload": "%d requests, each carrying its lattice (assembled on the device at admission), Poisson arrivals "
                    "at %.1f QPS, 70%% DoS / 30%% sigma_xy, models square/honeycomb/Lieb/Kagome/cubic/FCC, sites "
                    "log-uniform in [1e3, 4e6], W ~ U[0,2], N ~ U{256..2048}, R in {8,16,32}, half fp64 and half "
                    "mixed precision; sigma_xy draws with 8 N^2 n R > %.0e flop redrawn (%d times)"
                    % (len(reqs), qps, SERVE_SIGMA_FLOP_CAP, redrawn),
        "latency_ms_p50": float(np.percentile(lat_p, 50)),
        "latency_ms_p99": float(np.percentile(lat_p, 99)),
        "latency_ms_p50_dos": float(np.percentile(lat_p[kinds == "dos"], 50)),
        "latency_ms_p99_dos": float(np.percentile(lat_p[kinds == "dos"], 99)),
        "latency_ms_p50_sigma": float(np.percentile(lat_p[kinds == "kubo"], 50)) if (kinds == "kubo").any() else None,
        "latency_ms_p99_sigma": float(np.percentile(lat_p[kinds == "kubo"], 99)) if (kinds == "kubo").any() else None,
        "latency_by_size_band": percentiles_by(lat_p, reqs, lambda r: size_band(r["sites"])),
        "latency_by_kind_and_band": percentiles_by(lat_p, reqs,
                                                   lambda r: "%s %s" % (r["kind"], size_band(r["sites"]))),
        "largest_sigma_sites": max([r["sites"] for r in reqs if r["kind"] == "kubo"] or [0]),
        "sigma_requests_split_to_fit": split,
        "poisson_run_s": span_p,
        "served_makespan_s": span_s,
        "served_requests_per_s": len(reqs) / span_s,
        "served_dos_moments_per_s": moments / span_s,
        "plain_per_solve_s": span_plain,
        "plain_requests_per_s": len(reqs) / span_plain,
        "throughput_ratio_served_over_plain": span_plain / span_s,
        "chunks": st_s["chunks"],
        "server_create_s": created,
        "mixed_dos_max_rel_err_vs_fp64": max(err_dos) if err_dos else None,
        "mixed_sigma_max_rel_err_vs_fp64": max(err_sig) if err_sig else None,
    }
    log("serve: %s" % json.dumps(out))
    return out


def c1_leg(ctx, repeats=5):
    """The reference's CPU-runnable case end to end: square 100x100 periodic,
    t = 1, on-site energies uniform in [-0.5, 0.5], a = 4.6, R = 16 random-phase
    vectors, N = 512 moments, Jackson kernel, DoS at 1024 energies.  Device
    path (resident H, device kets, fused recursion, device reconstruction)
    against the oracle's restatement of the reference CPU loop (one moment per
    step, single thread) on the same kets."""
    import oracle
    from oracle import lattice_oracle as O
    from tight_binding_amd import kpm
    from tight_binding_amd import lattice as Lm
    from tight_binding_amd import spectral as S

    cell = Lm.UnitCell.model("square")
    cell.add_hopping_nn("A", 1, 1.0)
    lat = Lm.PartialLattice(cell, (100, 100))
    lat.set_disorder(1.0, 2031)
    H = lat.get_Hamiltonian_periodic_extend_bare_sparse((1, 1))
    a, N, R = 4.6, 512, 16
    e = S.linspace(-1.0, 1.0, 1024)
    g = S.jackson(N)
    M = ctx.matrix(H)
    K = ctx.random_block(H.n, R, seed=2032)
    ctx.moments(M, K, N, a=a)  # warm
    times, kms = [], []
    for _ in range(repeats):
        ctx.synchronize()
        t0 = time.perf_counter()
        mu, st, _ = ctx.moments(M, K, N, a=a)
        dos = kpm.density_of_states(ctx, mu, g, e, R)  # reconstruction on the device
        times.append(time.perf_counter() - t0)
        kms.append(st["kernel_ms"])
    kets = K.download()
    Hs = O.rescale(O.CSC(H.n, H.colptr, H.rowidx, H.values), a, 0.0)
    t0 = time.perf_counter()
    mu_cpu = oracle.cheb_braket(Hs, N, kets, kets)
    with np.errstate(divide="ignore", invalid="ignore"):
        dos_cpu = -S.green_imag(mu_cpu, g, e) / (R * math.pi)
    cpu_s = time.perf_counter() - t0
    err = float(np.max(np.abs(mu - mu_cpu)) / np.max(np.abs(mu_cpu)))
    fin = np.isfinite(dos_cpu) & np.isfinite(dos)
    K.close()
    M.close()
    out = {
        "workload": "DoS, square 100x100 periodic (10^4 sites), W=1, a=4.6, R=16, N=512, Jackson, 1024 energies",
        "gpu_solve_ms_median": 1e3 * float(np.median(times)),
        "gpu_kernel_ms_median": float(np.median(kms)),
        "cpu_oracle_solve_s": cpu_s, "cpu_cores": 1,
        "speedup_vs_cpu": cpu_s / float(np.median(times)),
        "mu_max_rel_diff_vs_cpu": err,
        # both DoS diverge at e = +-1 (the 1/sqrt(1-e^2) factor): compare the
        # finite values and require the same non-finite positions
        "dos_max_abs_diff_vs_cpu": float(np.max(np.abs(dos[fin] - dos_cpu[fin]))),
        "dos_nonfinite_positions_agree": bool(np.array_equal(np.isfinite(dos), np.isfinite(dos_cpu))),
    }
    log("c1: %s" % json.dumps(out))
    return out


def kubo_cpu_sample(ctx, cells=64, N=128, R=32):
    """The sigma_xy pipeline at a reduced size on the CPU (oracle stacks in C,
    one thread; Gamma as one numpy product; zeta by the separable host form)
    beside the device path at the same size: the baseline for the C3 solve,
    whose CPU cost grows as N^2 x sites x R for the contraction."""
    import oracle
    from oracle import spectral_oracle as SO
    from tight_binding_amd import kpm
    from tight_binding_amd import kubo
    from tight_binding_amd import spectral as S

    H, X, Y, emax = build_c3(cells)
    solver = kubo.KuboSolver(ctx, H, X, Y, -1.02 * emax, 1.02 * emax, volume=float(H.n), lam=6.0)
    ef = np.linspace(-4.0, 4.0, 201)
    solver.solve(N, R=R, seed=9, fermi_energies=ef)
    t0 = time.perf_counter()
    res = solver.solve(N, R=R, seed=9, fermi_energies=ef)
    gpu_s = time.perf_counter() - t0
    K = ctx.random_block(H.n, R, seed=9)
    kets = K.download()
    K.close()
    g = S.lorentz(N, 6.0)
    t0 = time.perf_counter()
    TL = oracle.cheb_on_bases(solver.Hs, N, kets)
    w = oracle.cheb_on_bases(solver.vy, 2, kets)[1]
    TR = oracle.cheb_on_bases(solver.Hs, N, w)
    aR = np.stack([oracle.cheb_on_bases(solver.vx, 2, TR[m])[1] for m in range(N)])
    G = SO.kubo_gamma(TL * g[:, None, None], aR * g[:, None, None])
    sig = S.kubo_sigma(G, (ef - solver.b) / solver.a, solver.volume)[0]
    cpu_s = time.perf_counter() - t0
    solver.close()
    out = {
        "workload": "sigma_xy, square %dx%d open, Landau flux 1/64, W=2, R=%d, N=%d" % (cells, cells, R, N),
        "gpu_solve_s": gpu_s, "cpu_solve_s": cpu_s, "cpu_cores": "1 (stacks) + numpy BLAS threads (product)",
        "sigma_max_rel_diff": float(np.max(np.abs(sig - res["sigma"])) / np.max(np.abs(res["sigma"]))),
    }
    log("kubo cpu sample: %s" % json.dumps(out))
    return out


def cpu_baseline(H, R, a, seconds):
    """The oracle (C restatement of the reference's CPU recursion, single
    thread) timed on a bounded number of steps of the same workload."""
    import oracle
    from oracle import lattice_oracle as O

    Hs = O.rescale(O.CSC(H.n, H.colptr, H.rowidx, H.values), a, 0.0)
    kets = np.full((H.n, R), 1.0 / math.sqrt(H.n), dtype=np.complex128)
    # one timed probe step decides how many fit the time budget
    t0 = time.perf_counter()
    oracle.cheb_braket(Hs, 3, kets, kets)  # T1 + one recursion step
    per_step = (time.perf_counter() - t0) / 2
    steps = max(1, min(64, int(seconds / max(per_step, 1e-3)) - 1))
    t0 = time.perf_counter()
    oracle.cheb_braket(Hs, steps + 2, kets, kets)
    dt = time.perf_counter() - t0
    total_steps = steps + 1  # T1 and `steps` recursion steps
    return {
        "value": total_steps / dt,
        "unit": "steps/s",
        "cores": 1,
        "kind": "port",
        "sample": "%d Chebyshev steps (T1 + %d recursion steps) of the same 2,097,152-site x 64-vector "
                  "fp64 workload, oracle/cheb_oracle.c (restatement of numerical.h:363-381, single "
                  "thread), %.1f s" % (total_steps, steps, dt),
    }


def cpu_baseline_all_cores(H, R, a, seconds):
    """The same recursion with each step split over the block columns on all
    host cores (OpenMP; oracle_cheb_steps_threads): a clearly labelled
    all-cores variant beside the single-thread port, on a bounded number of
    steps of the same workload."""
    import oracle
    from oracle import lattice_oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    Hs = O.rescale(O.CSC(H.n, H.colptr, H.rowidx, H.values), a, 0.0)
    kets = np.full((H.n, R), 1.0 / math.sqrt(H.n), dtype=np.complex128)
    t0 = time.perf_counter()
    oracle.cheb_steps_threads(Hs, 1, kets, threads)  # T1 + one recursion step
    per_step = (time.perf_counter() - t0) / 2
    steps = max(1, min(64, int(seconds / max(per_step, 1e-3)) - 1))
    t0 = time.perf_counter()
    oracle.cheb_steps_threads(Hs, steps, kets, threads)
    dt = time.perf_counter() - t0
    return {
        "value": (steps + 1) / dt,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d Chebyshev steps of the same workload, oracle_cheb_steps_threads (the oracle's loop with "
                  "each step split over the 64 block columns, OpenMP, %d threads), %.1f s" % (steps + 1, threads, dt),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1024)
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--cells", type=int, default=1024)
    ap.add_argument("--R", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=3, help="time this many plain block copies")
    ap.add_argument("--no-extra", action="store_true", help="skip the fp32 and one-moment-per-step variants")
    ap.add_argument("--no-autotune", action="store_true", help="use the default launch shape")
    ap.add_argument("--tuning", default=None,
                    help="C2 leg: install this launch shape (JSON of tbk_tuning fields) instead of autotuning")
    ap.add_argument("--kubo-repeats", type=int, default=3, help="sigma_xy solves per precision (0: skip)")
    ap.add_argument("--kubo-cells", type=int, default=512)
    ap.add_argument("--kubo-N", type=int, default=512)
    ap.add_argument("--kubo-R", type=int, default=32)
    ap.add_argument("--kubo-only", action="store_true", help="run only the sigma_xy leg")
    ap.add_argument("--batch-problems", type=int, default=256, help="problems of the batched leg (0: skip)")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched leg")
    ap.add_argument("--serve-requests", type=int, default=96, help="requests of the served leg (0: skip)")
    ap.add_argument("--serve-qps", type=float, default=2.0)
    ap.add_argument("--serve-only", action="store_true", help="run only the served leg")
    ap.add_argument("--serve-saturation-only", action="store_true",
                    help="served leg: only the all-at-once run (no Poisson run, no per-solve comparison)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path (a multi-GPU run "
              "would give each GPU R/G of the vectors and all-reduce the moments)", file=sys.stderr)
        sys.exit(2)

    from tight_binding_amd import kpm

    if args.serve_only:
        print(json.dumps({"serve": serve_leg(args.serve_qps, max(1, args.serve_requests),
                                             saturation_only=args.serve_saturation_only)}), flush=True)
        return

    if args.batch_only:
        ctx = kpm.Context(0)
        print(json.dumps({"batch": batch_leg(ctx, max(1, args.batch_problems))}), flush=True)
        ctx.close()
        return

    if args.kubo_only:
        ctx = kpm.Context(0)
        print(json.dumps({"kubo": kubo_leg(ctx, args.kubo_cells, args.kubo_N, args.kubo_R,
                                           max(1, args.kubo_repeats))}), flush=True)
        ctx.close()
        return

    t0 = time.perf_counter()
    H, a = build_c2(args.cells)
    log("assembled H: n=%d nnz=%d a=%.6f in %.1f s" % (H.n, H.nnz, a, time.perf_counter() - t0))

    ctx = kpm.Context(0)
    M = ctx.matrix(H)
    K = ctx.random_block(H.n, args.R, seed=7)
    ctx.synchronize()
    log("resident: %.2f GB on device" % (ctx.memory() / 1e9))

    copy_gbps = None
    if args.copy_probe > 0:
        # plain device copy of one vector block: the measured bandwidth ceiling
        # for streaming traffic, and the counter calibration kernel
        K2 = ctx.random_block(H.n, args.R, seed=8)
        nbytes = 2.0 * H.n * args.R * 16  # read + write
        times = [K.copy_to(K2) for _ in range(args.copy_probe)]
        copy_gbps = nbytes / (min(times) * 1e-3) / 1e9
        log("block copy (tbk_block_copy): %.3f ms best of %d, %.1f GB/s (read+write)"
            % (min(times), len(times), copy_gbps))
        K2.close()
    # attainable references on buffers of one vector block: a non-temporal
    # copy, the step's two-reads-one-write mix without gathers, a read
    attain = ctx.bandwidth_probe(H.n * args.R * 16, repeats=5)
    log("attainable: copy %.1f, 2 reads + 1 write %.1f, read %.1f GB/s"
        % (attain["copy_GBps"], attain["stream_2r1w_GBps"], attain["read_GBps"]))

    # launch-shape autotuning for this matrix and block (outside the timed region)
    tuning = None
    if args.tuning:
        tuning = dict(ctx.get_tuning(), **json.loads(args.tuning))
        ctx.set_tuning(**tuning)
        log("tuning installed: %s" % json.dumps(tuning))
    elif not args.no_autotune:
        t0 = time.perf_counter()
        tuning, tms = ctx.autotune(M, K, a=a, steps=16)
        log("autotune: %.3f ms/step with %s (%.1f s)" % (tms, json.dumps(tuning), time.perf_counter() - t0))

    # warmup
    ctx.moments(M, K, 2 * args.warmup, a=a)
    ctx.synchronize()

    # timed: exactly K steps (2K moments, doubling path).  One HIP event pair
    # on the library's stream brackets the K step launches, so kernel_ms / K
    # is the mean launch-to-launch time of a step (per-launch event pairs
    # would add each launch's dispatch delay to its time).
    n_mom = 2 * args.steps
    ctx.synchronize()
    t0 = time.perf_counter()
    mu, st, _ = ctx.moments(M, K, n_mom, a=a)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    assert st["steps"] == args.steps
    per_launch_ms = st["kernel_ms"] / st["steps"]
    achieved = st["bytes_total"] / (st["kernel_ms"] * 1e-3) / 1e9
    log("fp64: %d steps, wall %.3f s, kernel %.3f ms/step, %.1f GB/s (%.1f%% of 8 TB/s), mu0=%.6f"
        % (st["steps"], wall, per_launch_ms, achieved, 100 * achieved / HBM_PEAK_GBS, mu[0]))
    if not (abs(mu[0] - args.R) < 1e-8 and np.all(np.abs(mu) <= mu[0] * (1 + 1e-9))):
        raise SystemExit("bench: moments fail the mu_0 = R, |mu_n| <= mu_0 check")

    extra = {}
    if not args.no_extra:
        # one moment per step (distinct bras in general, k = 4 blocks per
        # step), with a launch shape tuned for that mode
        saved = ctx.get_tuning()
        tuning_plain = None
        if not args.no_autotune:
            tuning_plain, _ = ctx.autotune(M, K, a=a, steps=16, plain=True)
        ctx.moments(M, K, 8, a=a, plain=True)
        mu_p, stp, _ = ctx.moments(M, K, args.steps + 1, a=a, plain=True)
        ctx.set_tuning(**saved)
        gbs_p = stp["bytes_total"] / (stp["kernel_ms"] * 1e-3) / 1e9
        extra["plain_fp64"] = {
            "steps_per_s": stp["steps"] / (stp["kernel_ms"] * 1e-3),
            "GBps": gbs_p,
            "frac_hbm": gbs_p / HBM_PEAK_GBS,
            "bytes_per_step": stp["bytes_per_step"],
            "launch_tuning": tuning_plain,
        }
        K32 = ctx.random_block(H.n, args.R, seed=7, precision=kpm.FP32)
        saved = ctx.get_tuning()
        if not args.no_autotune:
            ctx.autotune(M, K32, a=a, steps=16)
        ctx.moments(M, K32, 8, a=a)
        mu32, st32, _ = ctx.moments(M, K32, n_mom, a=a)
        ctx.set_tuning(**saved)
        extra["fp32_recursion"] = {
            "steps_per_s": st32["steps"] / (st32["kernel_ms"] * 1e-3),
            "GBps": st32["bytes_total"] / (st32["kernel_ms"] * 1e-3) / 1e9,
            "max_abs_dmu_over_max_mu_vs_fp64": float(np.max(np.abs(mu32 - mu)) / np.max(np.abs(mu))),
        }
        K32.close()
        log("extra: %s" % json.dumps(extra))

    cpu = cpu_all = None
    if not args.no_cpu:
        log("cpu baseline (oracle, 1 thread, ~%.0f s)" % args.cpu_seconds)
        cpu = cpu_baseline(H, args.R, a, args.cpu_seconds)
        log("cpu: %.4f steps/s" % cpu["value"])
        cpu_all = cpu_baseline_all_cores(H, args.R, a, args.cpu_seconds)
        log("cpu, all cores (%d threads): %.4f steps/s" % (cpu_all["cores"], cpu_all["value"]))

    M.close()
    K.close()
    c1_out = c1_leg(ctx) if not args.no_cpu else None
    kubo_out = None
    if args.kubo_repeats > 0:
        kubo_out = kubo_leg(ctx, args.kubo_cells, args.kubo_N, args.kubo_R, args.kubo_repeats)
        if not args.no_cpu:
            kubo_out["cpu_sample"] = kubo_cpu_sample(ctx)
    batch_out = None
    if args.batch_problems > 0:
        ctx.release_work()
        batch_out = batch_leg(ctx, args.batch_problems)
    serve_out = None
    if args.serve_requests > 0:
        ctx.release_work()
        serve_out = serve_leg(args.serve_qps, args.serve_requests)

    value = st["steps"] / (st["kernel_ms"] * 1e-3)
    out = {
        "metric": "Chebyshev SpMM steps/s + %HBM roofline; sigma_xy solve p50/p99 latency",
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": per_launch_ms,
        "wall_ms_per_step": 1e3 * wall / st["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded honeycomb lattice, site-keyed disorder, device-generated random-phase vectors",
        "config": {
            "workload": "KPM DoS moments, honeycomb %dx%d cells periodic (%d sites), t=-2.7 eV, W=0.5 eV, "
                        "R=%d, N=%d moments, fp64" % (args.cells, args.cells, H.n, args.R, n_mom),
            "sites": H.n, "nnz": H.nnz, "R": args.R, "moments": n_mom, "moments_per_step": 2,
            "parallelism": "single GPU",
            "launch_tuning": tuning,
            "kernel_kind": st["kernel_kind"], "chunk_rows": st["chunk_rows"], "chunk_period": st["chunk_period"],
        },
        "moments_per_s": n_mom / (st["kernel_ms"] * 1e-3),
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": step_traffic(args.cells, args.R, tuning)[0],
            "traffic_source": step_traffic(args.cells, args.R, tuning)[1],
            "bytes_per_step": st["bytes_per_step"],
        },
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "copy_GBps_same_block": copy_gbps,
        "attainable_bw": {
            "stream_2r1w_GBps": attain["stream_2r1w_GBps"], "copy_nt_GBps": attain["copy_GBps"],
            "read_GBps": attain["read_GBps"], "buffer_bytes": attain["bytes"],
            "step_over_stream_2r1w": achieved / attain["stream_2r1w_GBps"],
        },
    }
    if extra:
        out["variants"] = extra
    if c1_out:
        out["c1_dos"] = c1_out
    if kubo_out:
        out["kubo_sigma_xy"] = kubo_out
    if batch_out:
        out["batched_ragged"] = batch_out
    if serve_out:
        out["served_mixed"] = serve_out
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
