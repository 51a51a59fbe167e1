"""Benchmark: RTE solves/s and p50/p99 latency at 32^3, L=5 (mixed FP32/FP64),
with the exitance error against an all-FP64 solve of the same inputs.

Workload (one "step"): 32 independent heterogeneous media of 32^3 voxels
(0.1 cm), L=5, g=0.9, mu_s=1, mu_a=0.01, lognormal sigma=0.3, 1-3 spherical
mu_a inclusions, 5 pencil beams per medium -> 160 solves, Neumann series to a
1e-6 relative change of total exitance (cap 30 terms).  The media are resident
in HBM before the timed region; every step rebuilds their path tables.

    python bench.py [--gpus 1] [--steps 2] [--warmup 1]
    python bench.py --serve [--requests 2000] [--rate 200]   # serving stream (2000 requests, 200/s)
    python bench.py --gradient                                # gradient + Fisher at 24^3
    python bench.py --fd [--omega 6.28e8]                    # the batch with modulated beams

Prints one JSON line.  One GPU only: --gpus N > 1 exits with a message.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 (vector/matrix dense) published peak
PEAK_HBM_GBS = 8000.0
# HBM bytes per solve and term of k_prop at 32^3 L=5 (mixed, cached), from rocprofv3
# FETCH_SIZE + WRITE_SIZE calibrated against a device copy: a beam in a group of five,
# three and two at the benchmark's full size
# (profiles/k_prop_32cube_L5_bench_hbm_traffic.md), alone on 4 media
# (profiles/k_prop_32cube_L5_4media_hbm_traffic.md)
TRAFFIC_PER_SOLVE_TERM = {5: 0.960e9, 3: 1.508e9, 2: 2.273e9, 1: 4.41e9}
DEFAULT_GROUP = 5          # the library's default number of beams per propagation group (L <= 5)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


REF_INI = """[Runtime]
TransFile = trans.out
ReflectFile = refl.out
PhaseFile = phase.out
[Algorithm]
nL = {L}
nTerms = 30
[Geometry]
nX = {n}
nY = {n}
nZ = {n}
xMin = {a!r}
xMax = {b!r}
yMin = {a!r}
yMax = {b!r}
zMin = 0.0
zMax = {z!r}
subVox = 0
subThresh = 0.0
propThresh = 10.0
selfSubVox = 1
[Phantom]
mua = 0.01
mus = 1.0
g = 0.9
n = 1.0
[Source]
posX = 0.0
posY = 0.0
diam = 0.2
magnitude = 1.0
modulation = 0.0
"""


def reference_cpu_propagation(n=10, L=5):
    """The reference CPU package itself (built from its sources into
    oracle/_ref/rte_cpu; absent -> None) on a uniform n^3 medium of the
    workload's optics: seconds per propagation A, from the times its progress
    lines appear (read through a pty, so they are line-buffered): the loop from
    "Neuman series iteration" to "Neumann series: n" holds n - 1 propagations
    and scatters plus the final propagation.  Run before this process touches
    the GPU."""
    import pty
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_ref", "rte_cpu")
    if not os.access(exe, os.X_OK):
        return None
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "in.ini"), "w") as f:
            f.write(REF_INI.format(L=L, n=n, a=-0.05 * n, b=0.05 * n, z=0.1 * n))
        master, slave = pty.openpty()
        t0 = time.perf_counter()
        proc = subprocess.Popen([exe, "in.ini"], cwd=d, stdout=slave, stderr=subprocess.DEVNULL,
                                env=dict(os.environ, OMP_NUM_THREADS=str(cores)))
        os.close(slave)
        buf, stamps = b"", []
        while True:
            try:
                chunk = os.read(master, 4096)
            except OSError:
                break
            if not chunk:
                break
            now = time.perf_counter() - t0
            buf += chunk
            while b"\n" in buf:
                line, buf = buf.split(b"\n", 1)
                stamps.append((now, line.decode(errors="replace").strip()))
        rc = proc.wait()
        os.close(master)
    try:
        t_loop = [t for t, l in stamps if "Neuman series iteration" in l][0]
        t_done, done = [(t, l) for t, l in stamps if l.startswith("Neumann series:")][0]
        nprop = int(done.split(":")[1])
    except (IndexError, ValueError):
        log("reference CPU run gave no timing (rc %d)" % rc)
        return None
    return dict(per_prop_s=(t_done - t_loop) / nprop, n=n, L=L, propagations=nprop, cores=cores,
                total_s=stamps[-1][0] if stamps else 0.0)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _port_per_propagation(P, W, L, budget_s, rows=256):
    """Seconds per full propagation of the FP64 oracle on problem P, timed on a
    growing sample of destination voxels until budget_s of CPU work."""
    N = P.N
    done, el = 0, 0.0
    while el < budget_s and done < N:
        n = min(rows, N - done)
        _, t = P.propagate_range_timed(W, L, done, done + n)   # the pair loop's own time
        el += t
        done += n
        rows *= 2
    return el * N / done, done, el


def cpu_baseline(shape, bounds, mu_a, mu_s, g, L, terms, budget_s=12.0, budget_1core_s=8.0):
    """FP64 oracle (C, OpenMP) timed on this host's cores, all of them and one,
    on a sample of destination voxels of one propagation of the workload's first
    medium; scaled to solves/s with the GPU's term count."""
    from oracle import oracle as O
    N = int(np.prod(shape))
    P = O.Problem(shape, bounds, mu_a, mu_s, g)
    W = np.random.default_rng(0).standard_normal(((L + 1) ** 2, N))
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.set_threads(cores)
    per_prop, done, el = _port_per_propagation(P, W, L, budget_s)
    O.set_threads(1)
    per_prop1, done1, el1 = _port_per_propagation(P, W, L, budget_1core_s, rows=64)
    O.set_threads(cores)
    return dict(value=1.0 / (per_prop * terms), unit="solves/s", cores=cores, kind="port",
                value_1core=1.0 / (per_prop1 * terms), cpu_model=cpu_model(),
                sample="FP64 oracle propagation of the workload's first %dx%dx%d L=%d medium over %d of %d "
                       "destination voxels on %d cores (%.1f s; %d voxels on 1 core, %.1f s), scaled to a full "
                       "propagation (%.1f s; %.1f s on 1 core) x %.1f Neumann terms per solve"
                       % (shape + (L, done, N, cores, el, done1, el1, per_prop, per_prop1, terms)))


def reference_calibration(L, n=10):
    """The reference CPU package (oracle/_ref/rte_cpu, built from its sources)
    against the FP64 oracle on the same uniform n^3 problem and the same cores:
    seconds per propagation of each and their ratio.  The reference runs
    uniform media only, so this ratio, not an extrapolation, is what relates
    the oracle's 32^3 timing to the reference's code."""
    ref = reference_cpu_propagation(n, L)
    if ref is None:
        return None
    from oracle import oracle as O
    shape, bounds = (n, n, n), (-0.05 * n, 0.05 * n, -0.05 * n, 0.05 * n, 0.0, 0.1 * n)
    P = O.Problem(shape, bounds, 0.01, 1.0, 0.9)
    W = np.random.default_rng(1).standard_normal(((L + 1) ** 2, n ** 3))
    O.set_threads(ref["cores"])
    port, _, _ = _port_per_propagation(P, W, L, 4.0)
    return dict(problem="uniform %d^3, L=%d, the workload's optics" % (n, L), cores=ref["cores"],
                reference_s_per_propagation=ref["per_prop_s"], port_s_per_propagation=port,
                reference_over_port=ref["per_prop_s"] / port,
                note="reference timed between its own progress lines over %d propagations"
                     % ref["propagations"])


def serve(args):
    """Serving stream: seeded Poisson arrivals of mixed-shape solves submitted
    in real time to the continuous-batching server; throughput, p50/p99
    latency, and the plain one-at-a-time FP32 path timed on a stratified
    sample of the same stream."""
    import threading
    from doi_amd import media, rte

    reqs = media.serving_requests(args.requests, seed=4, rate=args.rate)
    classes = sorted(set((r["shape"][0], r["L"]) for r in reqs))
    ctx = rte.Context(0)
    srv = rte.Server(ctx)
    # warm-up: one solve of every shape class (per-shape tables, buckets)
    warm = {}
    for r in reqs:
        warm.setdefault((r["shape"][0], r["L"]), r)
    for r in warm.values():
        srv.wait(srv.submit(r["shape"], r["bounds"], r["mu_a"], r["mu_s"], r["g"], r["beam"], r["params"]))
    log("serving %d requests at %.0f/s, classes %s" % (len(reqs), args.rate, classes))
    ids = [None] * len(reqs)
    results = [None] * len(reqs)
    ready = threading.Semaphore(0)

    def collect():
        for k in range(len(reqs)):
            ready.acquire()
            results[k] = srv.wait(ids[k])
            if (k + 1) % 250 == 0:
                log("%d done" % (k + 1))

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    for k, r in enumerate(reqs):
        dt = t0 + r["t"] - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        while True:
            try:
                ids[k] = srv.submit(r["shape"], r["bounds"], r["mu_a"], r["mu_s"], r["g"], r["beam"], r["params"])
                break
            except rte.RteError as e:
                if e.code != rte.E_BUSY:
                    raise
                time.sleep(0.001)
        ready.release()
    th.join()
    wall = time.perf_counter() - t0
    info = srv.info()
    srv.close()
    lat = np.array([res[2]["latency_ms"] for res in results])
    served = len(reqs) / wall
    # plain path on a sample: 2 solves per class, FP32, marched path integrals, one at a time
    per_class = {}
    for c in classes:
        sample = [r for r in reqs if (r["shape"][0], r["L"]) == c][:args.plain_sample]
        ts = []
        for r in sample:
            p = rte.make_params(r["L"], precision=rte.PREC_FP32, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30,
                                path_mode=rte.PATH_MARCH)
            t1 = time.perf_counter()
            ctx.solve(r["shape"], r["bounds"], r["mu_a"], r["mu_s"], r["g"], r["beam"], p)
            ts.append(time.perf_counter() - t1)
        per_class["%d^3 L%d" % c] = float(np.mean(ts))
        log("plain %d^3 L%d: %.4f s" % (c[0], c[1], per_class["%d^3 L%d" % c]))
    t_plain = sum(per_class["%d^3 L%d" % (r["shape"][0], r["L"])] for r in reqs)
    plain = len(reqs) / t_plain
    # error of served results against all-FP64 solves with the same term count:
    # the first request of every (grid, L) class
    errs = {}
    for c in classes:
        k = next(i for i, r in enumerate(reqs) if (r["shape"][0], r["L"]) == c)
        r, (status, img, stt) = reqs[k], results[k]
        p64 = rte.make_params(r["L"], precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=0.0,
                              max_terms=stt["terms"])
        ref, _ = ctx.solve(r["shape"], r["bounds"], r["mu_a"], r["mu_s"], r["g"], r["beam"], p64)
        errs["%d^3 L%d" % c] = float(np.max(np.abs(img - ref)) / np.max(np.abs(ref)))
    res = {
        "metric": "RTE serving stream: solves/s + p50/p99 latency (mixed grids 8-32^3, L 3/5)",
        "value": served,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": len(warm),
        "ms_per_step": 1000.0 * wall,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 running sums)",
        "data": "synthetic (seeded serving stream)",
        "config": {"workload": "serving stream: %d requests, Poisson %.0f/s, edges {8,12,16,24,32}, L {3,5}, "
                               "g U[0.7,0.9], mu_s U[0.5,1.5], mu_a U[0.005,0.05], sigma 0.2, tol 1e-6 cap 30"
                               % (len(reqs), args.rate)},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max())},
        "arrival_span_s": reqs[-1]["t"],
        "server": {"steps": info["steps"], "buckets": info["buckets"], "failed": info["failed"]},
        "plain_fp32_solves_per_s": plain,
        "served_over_plain": served / plain,
        "plain_s_per_solve_by_class": per_class,
        "rel_err_vs_fp64": max(errs.values()),
        "rel_err_vs_fp64_by_class": errs,
    }
    print(json.dumps(res), flush=True)
    ctx.close()


def served_latency(maps, beams5, n, L, rate, nreq, seed=11):
    """Request latency at 32^3 L=5: solves of the workload's own media and
    beams submitted one at a time to the continuous-batching server with
    seeded Poisson arrivals at `rate` per second; submit-to-result latency of
    each, served throughput, and the arrival span."""
    import threading
    from doi_amd import media, rte
    bounds = media.cube_bounds(n)
    prm = rte.make_params(L, precision=rte.PREC_MIXED, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30)
    st = media.Stream(seed)
    gaps = -np.log1p(-st.uniform(nreq)) / rate
    t_arr = np.cumsum(gaps) - gaps[0]
    pick = [(k % len(maps), (k // len(maps)) % len(beams5)) for k in range(nreq)]
    ctx = rte.Context(0)
    srv = rte.Server(ctx)
    ma, ms = maps[0]
    srv.wait(srv.submit((n, n, n), bounds, ma, ms, 0.9, beams5[0], prm))      # warm-up of the bucket
    ids, res = [None] * nreq, [None] * nreq
    ready = threading.Semaphore(0)

    def collect():
        for k in range(nreq):
            ready.acquire()
            res[k] = srv.wait(ids[k])

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    for k, (mi, bi) in enumerate(pick):
        dt = t0 + t_arr[k] - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        ma, ms = maps[mi]
        ids[k] = srv.submit((n, n, n), bounds, ma, ms, 0.9, beams5[bi], prm)
        ready.release()
    th.join()
    wall = time.perf_counter() - t0
    srv.close()
    ctx.close()
    lat = np.array([r[2]["latency_ms"] for r in res])
    assert all(r[0] == rte.REQ_DONE for r in res)
    return {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)), "max": float(lat.max()),
            "requests": nreq, "rate_per_s": rate, "served_solves_per_s": nreq / wall,
            "kind": "request latency: 32^3 L=5 requests of the workload's media and beams, Poisson arrivals "
                    "at %.0f/s, submit to result through the continuous-batching server" % rate}


def cpu_baseline_fd(shape, bounds, mu_a, mu_s, g, L, omega, terms, budget_s=12.0):
    """cpu_baseline for the frequency-domain workload: the oracle's complex
    propagation on a sample of destination voxels, scaled as cpu_baseline."""
    from oracle import oracle as O
    N = int(np.prod(shape))
    P = O.Problem(shape, bounds, mu_a, mu_s, g)
    rng = np.random.default_rng(0)
    W = rng.standard_normal(((L + 1) ** 2, N)) + 1j * rng.standard_normal(((L + 1) ** 2, N))
    t0 = time.perf_counter()
    P.propagate_c_range(W, L, omega, 0, 0)
    setup = time.perf_counter() - t0
    rows, done, el = 128, 0, 0.0
    while el < budget_s and done < N:
        n = min(rows, N - done)
        t0 = time.perf_counter()
        P.propagate_c_range(W, L, omega, done, done + n)
        el += time.perf_counter() - t0 - setup
        done += n
        rows *= 2
    per_prop = el * N / done
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=1.0 / (per_prop * terms), unit="solves/s", cores=cores, kind="port",
                sample="FP64 oracle complex propagation of one 32^3 L=5 medium over %d of %d destination voxels "
                       "(%.1f s), scaled to a full propagation (%.1f s) x %.1f Neumann terms per solve"
                       % (done, N, el, per_prop, terms))


def frequency_domain(args):
    """The batch of 32 lognormal media x 5 beams at 32^3, L=5, with every beam
    intensity-modulated (omega rad/s):
    complex Neumann series through the complex propagation group."""
    from doi_amd import media, rte
    n, L, g = args.n, args.L, 0.9
    shape, bounds = (n, n, n), media.cube_bounds(n)
    maps = media.lognormal_media_with_inclusions(args.media, seed=5, n=n)
    beams5 = media.top_face_beams(seed=5, n=n, per_medium=args.beams)
    ctx = rte.Context(0)
    mm = [ctx.medium(shape, bounds, ma, ms, g) for ma, ms in maps]
    med_list = [m for m in mm for _ in range(args.beams)]
    beam_list = [b for _ in mm for b in beams5]
    nsolve = len(med_list)
    prm = rte.make_params(L, precision=rte.PREC_MIXED, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30)
    for i in range(args.warmup):
        log("fd warmup step", i)
        ctx.solve_fd(med_list, beam_list, [args.omega] * nsolve, prm)
    ctx.profile(True)
    times, lat = [], []
    for i in range(args.steps):
        for m in mm:
            m.reset_cache()
        t0 = time.perf_counter()
        imgs, _, stats = ctx.solve_fd(med_list, beam_list, [args.omega] * nsolve, prm)
        times.append(time.perf_counter() - t0)
        lat += [s["latency_ms"] for s in stats]
        log("fd step %d: %.3f s, terms %s" % (i, times[-1], sorted(set(s["terms"] for s in stats))))
    prof = ctx.get_profile()
    ctx.profile(False)
    ms_per_step = 1000.0 * float(np.mean(times))
    terms_mean = float(np.mean([s["terms"] for s in stats]))
    launches = max(prof["prop_launches"], 1)
    avg_ms = prof["prop_ms"] / launches
    achieved = prof["prop_flops"] / launches / (avg_ms * 1e-3) / 1e12
    p64 = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30)
    ref, _, _ = ctx.solve_fd([mm[0]] * args.beams, beams5, [args.omega] * args.beams, p64)
    err = max(float(np.max(np.abs(imgs[b] - ref[b])) / np.max(np.abs(ref[b]))) for b in range(args.beams))
    # phase where the exitance is well resolved (>= 1 % of its peak); the corner
    # pixels of an L = 5 expansion at g = 0.9 can go slightly negative
    big = np.abs(ref[0]) >= 1e-2 * np.max(np.abs(ref[0]))
    ph = np.angle(ref[0][big])
    res = {"metric": "frequency-domain RTE solves/s at 32^3 L=5 (modulated beams), <=1e-4 rel err vs FP64",
           "value": nsolve / (ms_per_step / 1000.0), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32 complex (fp64 running sums)",
           "data": "synthetic (seeded lognormal media with inclusions)",
           "config": {"workload": "32 lognormal media with inclusions x 5 beams, %d^3, L=%d, every beam modulated at omega = %.4g rad/s"
                                  % (n, L, args.omega), "solves_per_step": nsolve},
           "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
           "terms_mean": terms_mean, "rel_err_vs_fp64": err,
           "phase_range_deg_medium0_beam0_resolved_pixels": [float(np.degrees(ph.min())), float(np.degrees(ph.max()))],
           "roofline": {"bound": "valu", "pipe": "valu-fp32", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                        "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                        "kernel": "k_prop complex group (cached path)", "avg_launch_ms": avg_ms}}
    if not args.no_cpu:
        log("cpu baseline (complex)")
        res["cpu_baseline"] = cpu_baseline_fd(shape, bounds, maps[0][0], maps[0][1], g, L, args.omega, terms_mean)
    print(json.dumps(res), flush=True)
    for m in mm:
        m.close()
    ctx.close()


def gradient_and_fisher(args):
    """Differentiation workload: 24^3 heterogeneous medium, L=5, mixed
    precision; the loss gradient with respect to every voxel's mu_a and mu_s
    (one forward + one adjoint series), its error against the same gradient in
    FP64, and the Fisher matrix of 64 block parameters on 576 detectors (J by
    the reference's per-parameter derivative series, and exact by one adjoint
    series per detector; Gram on MFMA, FP64 Cholesky)."""
    from doi_amd import media, rte
    n, L, g = 24, 5, 0.8
    shape, bounds = (n, n, n), media.cube_bounds(n)
    maps = lambda seed: media.heterogeneous(n, seed, sigma=0.2, n_incl=1, radius=(3, 3), factor=(2.0, 2.0))
    ma, ms = maps(3)
    ma2, ms2 = maps(4)
    beam = (0.05, 0.05, 0.2, 1.0)
    ctx = rte.Context(0)
    pm = rte.make_params(L, precision=rte.PREC_MIXED, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30)
    target, _ = ctx.solve(shape, bounds, ma2, ms2, g, beam, pm)
    med = ctx.medium(shape, bounds, ma, ms, g)
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    part = ((i // 6) + 4 * (j // 6) + 16 * (k // 6)).transpose(2, 1, 0).ravel()
    ctx.gradient(med, beam, pm, target)          # warm-up (tables, moments)
    tg, tf = [], []
    for _ in range(max(args.steps, 1)):
        med.reset_cache()
        t0 = time.perf_counter()
        loss, img, ga, gs, st = ctx.gradient(med, beam, pm, target)
        tg.append(time.perf_counter() - t0)
        med.reset_cache()
        t0 = time.perf_counter()
        F, Cf, J, _, stf = ctx.fisher(med, beam, pm, part, 64, wrt=0, jacobian=1)
        tf.append(time.perf_counter() - t0)
        log("gradient %.3f s, Fisher (reference Jacobian) %.3f s" % (tg[-1], tf[-1]))
    # the same gradient with every step in FP64: the mixed path's error
    p64 = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30)
    loss64, img64, ga64, gs64, st64 = ctx.gradient(med, beam, p64, target)
    gscale = max(np.max(np.abs(ga64)), np.max(np.abs(gs64)))
    grad_err = max(np.max(np.abs(ga - ga64)), np.max(np.abs(gs - gs64))) / gscale
    # the exact Fisher information: one adjoint series per detector
    med.reset_cache()
    t0 = time.perf_counter()
    Fx, Cx, Jx, _, stx = ctx.fisher(med, beam, pm, part, 64, wrt=0, jacobian=0)
    tfx = time.perf_counter() - t0
    log("exact Fisher %.3f s" % tfx)
    ev = np.linalg.eigvalsh(F)
    evx = np.linalg.eigvalsh(Fx)
    res = {"metric": "gradient and Fisher time at 24^3 L=5", "value": 1.0 / float(np.mean(tg)),
           "unit": "gradients/s", "n_gpus": 1, "steps": len(tg), "warmup": 1,
           "ms_per_step": 1000.0 * float(np.mean(tg)), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32 (fp64 running sums)", "data": "synthetic (seeded lognormal map with an inclusion)",
           "config": {"workload": "24^3, L=5, g=0.8, loss vs second seeded map; Fisher 64 mu_a blocks "
                                  "(6^3 voxels), 576 detectors"},
           "terms": st["terms"], "gradient_s": float(np.mean(tg)),
           "gradient_rel_err_vs_fp64": float(grad_err), "terms_fp64": st64["terms"],
           "loss_rel_err_vs_fp64": float(abs(loss - loss64) / abs(loss64)),
           "fisher_reference_jacobian_s": float(np.mean(tf)),
           "fisher_reference_jacobian_eig_min_max": [float(ev[0]), float(ev[-1])],
           "fisher_exact_s": tfx, "fisher_exact_eig_min_max": [float(evx[0]), float(evx[-1])],
           "jacobian_reference_vs_exact_rel": float(np.linalg.norm(J - Jx) / np.linalg.norm(Jx))}
    print(json.dumps(res), flush=True)
    med.close()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--media", type=int, default=32)
    ap.add_argument("--beams", type=int, default=5)
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--L", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--group", type=int, default=0, help="max beams per propagation group (0 = library default)")
    ap.add_argument("--serve", action="store_true", help="run the serving stream instead of the batch workload")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--plain-sample", type=int, default=2)
    ap.add_argument("--serve-rate", type=float, default=8.0,
                    help="arrival rate of the 32^3 L=5 request-latency measurement (0: skip it)")
    ap.add_argument("--serve-requests", type=int, default=80)
    ap.add_argument("--gradient", action="store_true", help="time the gradient and Fisher workload instead")
    ap.add_argument("--fd", action="store_true", help="time the frequency-domain (modulated beam) workload")
    ap.add_argument("--omega", type=float, default=2.0 * np.pi * 1.0e8, help="modulation, rad/s (--fd)")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="only run 3 plain device copies of this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.calib_copy:
        import torch
        a = torch.ones(args.calib_copy // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": a.numel() * 4, "copies": 3}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (independent solves "
              "shard across GPUs with no communication). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if args.serve:
        return serve(args)
    if args.gradient:
        return gradient_and_fisher(args)
    if args.fd:
        return frequency_domain(args)
    ref_cal = None
    if not args.no_cpu:
        # before this process touches the GPU
        log("reference CPU package against the oracle (uniform 10^3 L=5)")
        ref_cal = reference_calibration(args.L)
    from doi_amd import media, rte

    n, L = args.n, args.L
    shape, bounds = (n, n, n), media.cube_bounds(n)
    g = 0.9
    log("generating %d media of %d^3" % (args.media, n))
    maps = media.lognormal_media_with_inclusions(args.media, seed=5, n=n)
    beams5 = media.top_face_beams(seed=5, n=n, per_medium=args.beams)
    ctx = rte.Context(0)
    if args.group:
        ctx.set_group_size(args.group)
    mm = [ctx.medium(shape, bounds, ma, ms, g) for ma, ms in maps]
    med_list = [m for m in mm for _ in range(args.beams)]
    beam_list = [b for _ in mm for b in beams5]
    nsolve = len(med_list)
    prm = rte.make_params(L, precision=rte.PREC_MIXED, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30,
                          path_mode=rte.PATH_CACHED)

    def step():
        for m in mm:
            m.reset_cache()           # every step rebuilds the per-medium path tables
        return ctx.solve_media(med_list, beam_list, prm)

    for i in range(args.warmup):
        log("warmup step", i)
        step()
    ctx.profile(True)
    lat, times, imgs, stats = [], [], None, None
    for i in range(args.steps):
        for m in mm:
            m.reset_cache()
        t0 = time.perf_counter()
        imgs, stats = ctx.solve_media(med_list, beam_list, prm)
        dt = time.perf_counter() - t0
        times.append(dt)
        lat += [s["latency_ms"] for s in stats]
        log("step %d: %.3f s, terms %s" % (i, dt, sorted(set(s["terms"] for s in stats))))
    prof = ctx.get_profile()
    ctx.profile(False)
    ms_per_step = 1000.0 * float(np.mean(times))
    value = nsolve / (ms_per_step / 1000.0)
    terms_mean = float(np.mean([s["terms"] for s in stats]))

    # roofline of the dominant kernel, the propagation group of the default size
    # (k_prop<float, double, L, 0, 5>; the library allows at most 3 at L = 6 and 2 from
    # L = 7): its algorithmic flops / its launch time
    gsz = min(args.group or DEFAULT_GROUP, args.beams, 5 if L <= 5 else (3 if L == 6 else 2))
    launches = max(prof["group_launches"][gsz], 1)
    avg_ms = prof["group_ms"][gsz] / launches
    flops_per_launch = prof["group_flops"][gsz] / launches
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    pairs_per_s = prof["prop_pairs"] / (prof["prop_ms"] * 1e-3) if prof["prop_ms"] > 0 else 0.0
    all_launches = max(prof["prop_launches"], 1)
    achieved_all = prof["prop_flops"] / (prof["prop_ms"] * 1e-3) / 1e12 if prof["prop_ms"] > 0 else 0.0
    K = (L + 1) ** 2
    npairs = float(n ** 3) * (n ** 3 - 1)
    # beam-terms a launch processes: the library counts (16 + k 4K) flops per pair of a group of k
    beam_terms = flops_per_launch / ((16 + gsz * 4 * K) * npairs) * gsz
    traffic = None
    if n == 32 and L == 5 and gsz in TRAFFIC_PER_SOLVE_TERM:
        traffic = TRAFFIC_PER_SOLVE_TERM[gsz] * beam_terms

    out = {}
    # error against an all-FP64 solve of medium 0's beams (same term cap)
    if not args.no_check:
        p64 = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30,
                              path_mode=rte.PATH_CACHED)
        ref, _ = ctx.solve_media([mm[0]] * args.beams, beams5, p64)
        errs = [float(np.max(np.abs(imgs[b] - ref[b])) / np.max(np.abs(ref[b]))) for b in range(args.beams)]
        out["rel_err_vs_fp64"] = max(errs)
        mm[0].reset_cache()
        # the plain path: one solve at a time, FP32 throughout, path integral
        # marched per pair in every term (the reference's structure)
        pplain = rte.make_params(L, precision=rte.PREC_FP32, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30,
                                 path_mode=rte.PATH_MARCH)
        t0 = time.perf_counter()
        plain_img, _ = ctx.solve_media([mm[0]], [beams5[0]], pplain)
        t_plain = time.perf_counter() - t0
        out["plain_fp32_solves_per_s"] = 1.0 / t_plain
        out["plain_rel_err_vs_fp64"] = float(np.max(np.abs(plain_img[0] - ref[0])) / np.max(np.abs(ref[0])))
        out["batched_over_plain"] = value / out["plain_fp32_solves_per_s"]
        # the same one-at-a-time FP32 solve with cached path integrals (each solve
        # builds its medium's table, as an independent solve must)
        pcached = rte.make_params(L, precision=rte.PREC_FP32, stop_rule=rte.STOP_DELTA, tol=1e-6, max_terms=30,
                                  path_mode=rte.PATH_CACHED)
        tc = []
        for b in range(2):
            mm[0].reset_cache()
            t0 = time.perf_counter()
            ctx.solve_media([mm[0]], [beams5[b]], pcached)
            tc.append(time.perf_counter() - t0)
        out["plain_cached_fp32_solves_per_s"] = 1.0 / float(np.mean(tc))
        out["batched_over_plain_cached"] = value / out["plain_cached_fp32_solves_per_s"]

    res = {
        "metric": "RTE solves/s + p50/p99 latency at 32^3 L=5, <=1e-4 rel exitance err vs FP64",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 running sums)",
        "data": "synthetic (seeded lognormal media with inclusions)",
        "config": {"workload": "%d lognormal media with inclusions x %d beams, %d^3 voxels 0.1 cm, L=%d, g=0.9, tol 1e-6 cap 30"
                               % (args.media, args.beams, n, L),
                   "solves_per_step": nsolve, "grid": n, "L": L, "precision": "mixed",
                   "path": "cached path integrals"},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "kind": "in-batch: from the batch call's start to each solve's final reduction, "
                               "all %d solves submitted together (request latency under arrivals: --serve)"
                               % nsolve},
        "terms_mean": terms_mean,
        "voxel_pair_propagations_per_s": pairs_per_s,
        "roofline": {"bound": "valu", "pipe": "valu-fp32", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "kernel": "k_prop<float, double, %d, 0, %d> (groups of %d beams, cached path)" % (L, gsz, gsz),
                     "launches": launches, "avg_launch_ms": avg_ms, "flops_per_launch": flops_per_launch,
                     "beam_terms_per_launch": beam_terms,
                     "all_propagation_launches": {"launches": all_launches,
                                                  "avg_launch_ms": prof["prop_ms"] / all_launches,
                                                  "achieved": achieved_all,
                                                  "frac": achieved_all / PEAK_FP32_TFLOPS}},
    }
    res.update(out)
    res["ktab_ms_total"] = prof["ktab_ms"]
    if args.serve_rate > 0:
        for m in mm:
            m.close()
        mm = []
        log("served latency: %d requests at %.0f/s" % (args.serve_requests, args.serve_rate))
        res["served_latency_ms"] = served_latency(maps, beams5, n, L, args.serve_rate, args.serve_requests)
    if not args.no_cpu:
        log("cpu baseline")
        res["cpu_baseline"] = cpu_baseline(shape, bounds, maps[0][0], maps[0][1], g, L, terms_mean)
        if ref_cal is not None:
            res["cpu_baseline"]["reference_calibration"] = ref_cal
    print(json.dumps(res), flush=True)
    for m in mm:
        m.close()
    ctx.close()


if __name__ == "__main__":
    main()
