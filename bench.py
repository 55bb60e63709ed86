"""Benchmark: LJ31 SAD moves/s on one MI355X.

Workload (the LJ31 precision configuration): 16384 independent SAD walkers,
each an LJ cluster of N = 31 atoms in a hard sphere of radius 3.0 sigma,
started from a random configuration with all pairs >= 0.8 sigma; single-atom
moves uniform in a cube of half-width 0.1 sigma; SAD with min_T = 0.05
epsilon, energy bins of 0.01 epsilon, --max-allowed-energy 0; FP32 pair terms
accumulated in FP64, FP64 full-energy refresh every 1024 moves.

A step is one launch that advances every walker by --moves-per-step moves;
the default 20 steps of 500000 moves time each walker's 1e7-move run (the
configuration's timing run, from SAD's start-up phase, when the energy range
is still being discovered, to its steady state).  --warmup untimed launches
of 2000 moves come first.  `value` is total moves / wall seconds of the timed
steps, bracketed by stream synchronisation; the kernel time per launch comes
from HIP events on the library's stream.

The same line carries the FP64 variant's throughput, the relative energy
error of a sample of walkers against an FP64 recompute, the roofline figures
of the run kernel, and CPU baselines: the C oracle (the same algorithm and
configuration, one walker per thread on every host core, each walker running
the same 1e7-move trajectory) timed on this host, for both precisions.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sad-monte-carlo_amd"))

import sadmc_amd as S  # noqa: E402

N_ATOMS = 31
RADIUS = 3.0
FLOP_PER_MOVE = 30 * (N_ATOMS - 1)  # 2(N-1) pair terms x 15 flop (division counted as one)
FP32_VECTOR_PEAK = 157.3            # TFLOP/s, MI355X
FP64_VECTOR_PEAK = 78.6
HBM_PEAK_GBS = 8000.0
BYTES_PER_MOVE = 80                 # lnw at two bins, read-modify-write of hist, lnw, e_tot, e_sq
PROFILE = os.path.join(ROOT, "profiles", "k_run_lj31_fp32_sad_16384w.md")
LEARNED_CV_PROFILE = os.path.join(ROOT, "profiles", "learned_cv_lj31_16384w.md")
WARMUP_MOVES = 2000                 # moves per warmup launch


def log(msg):
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def lj_params(precision, lanes):
    return S.make_params(system=S.LJ, n=N_ATOMS, lanes=lanes, method=S.SAD, method_param=0.05,
                         move_param=0.1, energy_bin=0.01, has_max_e=1, max_e=0.0, precision=precision,
                         refresh=1024 if precision == S.FP32_PAIRS else 0)


def lj_energy_fp64(c):
    x = c.reshape(3, -1)
    d = x[:, :, None] - x[:, None, :]
    r2 = (d * d).sum(0)
    iu = np.triu_indices(x.shape[1], 1)
    s = 1.0 / r2[iu]
    s3 = s * s * s
    return float(np.sum(4.0 * (s3 * s3 - s3)))


def measured_traffic(a):
    """HBM bytes per launch of the run kernel for this workload, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary kept in profiles/ (scaled by
    the copy-kernel calibration there); None if it is not there."""
    try:
        for line in open(PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def recorded_learned_cv():
    """The second half of the metric, wall time to Cv(T) within 2 % for plain
    and learned walkers: a run of 15+ minutes of GPU time (`--lj31-cv`, then
    `--learned-cv ... --learned-cv-ref`), so the default line carries the
    result recorded in profiles/ with its source instead of re-measuring it;
    None if the summary is not there."""
    try:
        for line in open(LEARNED_CV_PROFILE):
            if line.startswith("learned_cv:"):
                res = json.loads(line.split(":", 1)[1])
                res["source"] = os.path.relpath(LEARNED_CV_PROFILE, os.path.dirname(os.path.abspath(__file__)))
                res["measured_in_this_run"] = False
                return res
    except OSError:
        pass
    return None


def run_gpu(precision, walkers, lanes, warmup, steps, mps, label):
    """Time `steps` launches of `mps` moves of every walker, after `warmup`
    short launches (WARMUP_MOVES each) of the same walkers."""
    ctx = S.Context(0)
    sid = ctx.add_set(lj_params(precision, lanes), np.arange(walkers, dtype=np.uint64), geom=RADIUS)
    ctx.init()
    for _ in range(warmup):
        ctx.run(WARMUP_MOVES)
    log(f"[{label}] timing {steps} steps of {mps} moves x {walkers} walkers")
    ctx.sync()
    kernel_ms = []
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.run(mps, sync=False)
        ctx.sync()
        kernel_ms.append(ctx.last_run_ms())
        log(f"[{label}] step {k + 1}/{steps}: {kernel_ms[-1]:.1f} ms")
    wall = time.perf_counter() - t0
    moves = steps * mps * walkers
    # relative energy error of a sample of walkers against an FP64 recompute
    st, _ = ctx.download_raw(sid, with_bins=False)
    rel = []
    for w in range(0, walkers, max(1, walkers // 256)):
        e_ref = lj_energy_fp64(ctx.coords(sid, w))
        rel.append(abs(st[w]["E"] - e_ref) / abs(e_ref))
    ctx.close()
    log(f"[{label}] {moves / wall / 1e9:.3f} G moves/s; kernel {np.mean(kernel_ms):.2f} ms/launch")
    return dict(value=moves / wall, wall=wall, kernel_ms=float(np.mean(kernel_ms)),
                max_rel_energy_error=float(np.max(rel)), mean_rel_energy_error=float(np.mean(rel)))


def cpu_baseline(precision, moves, rounds):
    """The C oracle (same algorithm and configuration, the reference's
    summation order), one walker per thread over every core, each walker
    running the same `moves`-move trajectory as a GPU walker; built here with
    -O3 -march=native."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes as C
    import oracle as O
    src = os.path.join(ROOT, "oracle", "sadmc_oracle.c")
    so = os.path.join("/tmp", f"liboracle_native_{os.getpid()}.so")
    subprocess.run(["gcc", "-O3", "-march=native", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-shared", "-o", so,
                    src, "-lm", "-lpthread"], check=True)
    lib = C.CDLL(so)
    lib.or_bench.argtypes = [C.POINTER(O.Params), C.c_int32, C.c_uint64, C.c_uint64]
    lib.or_bench.restype = C.c_double
    p = O.make_params(system=O.LJ, n=N_ATOMS, radius=RADIUS, method=O.SAD, method_param=0.05, move_param=0.1,
                      energy_bin=0.01, has_max_e=1, max_e=0.0, precision=precision,
                      refresh=1024 if precision == 1 else 0, order_g=0)
    threads = len(os.sched_getaffinity(0))
    wall = 0.0
    for r in range(rounds):
        wall += lib.or_bench(C.byref(p), threads, moves, r * threads)
        log(f"[cpu {'fp32' if precision else 'fp64'}] round {r + 1}/{rounds}: {wall:.1f} s")
    os.remove(so)
    cpu = subprocess.run(["bash", "-c", "grep -m1 'model name' /proc/cpuinfo | cut -d: -f2"], capture_output=True,
                         text=True).stdout.strip()
    n = rounds * threads
    return dict(value=n * moves / wall, unit="moves/s", cores=threads, kind="port",
                sample=f"{n} walkers (seeds 0..{n - 1}) x {moves} moves each, {threads} threads, C oracle in the "
                       f"reference's summation order, -O3 -march=native, {wall:.1f} s on {cpu}")


def learned_cv(budget, walkers, lanes, chunk, out_path, system="lj38", ref_path=None):
    """Wall time to Cv(T) within 2% for LJ38 (sphere 3.2 sigma, SAD min_T =
    0.03, 16384 walkers; or LJ13, sphere 2.5, min_T = 0.05): plain cube-move
    walkers against learned-proposal walkers.  The reference curve is the plain walkers' Cv after `budget`
    seconds of GPU time.  The learned walkers are the same walkers: they run
    the cube move while plain-walker records are collected (64 per walker,
    1e6 in all), the network is trained, and then they continue with the
    learned move; their clock includes the collection and the training.  Cv
    comes from the device entropy reduction at checkpoints every `chunk`
    moves; the time to 2% is the first checkpoint after which every later
    checkpoint stays within 2% of the reference on the temperature window."""
    import analysis as A
    T = np.linspace(0.05, 0.40, 71)
    windows = {"T0.05-0.40": T >= 0.05, "T0.10-0.40": T >= 0.10, "T0.15-0.40": T >= 0.15,
               "T0.30-0.40": T >= 0.30 - 1e-9}

    n, radius, min_t = {"lj38": (38, 3.2, 0.03), "lj13": (13, 2.5, 0.05), "lj31": (31, 3.0, 0.05)}[system]
    # LJ31: the precision configuration (FP32 pairs, FP64 refresh every 1024 moves), as --lj31-cv runs it
    prec = dict(precision=S.FP32_PAIRS, refresh=1024) if system == "lj31" else {}

    def params(kind):
        return S.make_params(system=S.LJ, n=n, lanes=lanes, method=S.SAD, method_param=min_t, move_param=0.1,
                             energy_bin=0.01, has_max_e=1, max_e=0.0, move_kind=kind, **prec)

    def checkpoint(ctx, sid, t, tag, curves):
        try:
            E, Sg, _ = ctx.mean_entropy(sid)
        except S.SadmcError as e:  # no energy window common to all walkers yet
            log(f"[{tag}] t={t:7.1f} s  no estimate: {e}")
            return
        cv = A.heat_capacity(T, E, Sg)
        curves.append((t, cv))
        log(f"[{tag}] t={t:7.1f} s  Cv(0.1)={cv[10]:8.2f} Cv(0.2)={cv[30]:8.2f} Cv(0.3)={cv[50]:8.2f}")

    seeds = np.arange(walkers, dtype=np.uint64)
    if ref_path:
        # the plain walkers' curves from a --lj31-cv run: its last checkpoint is the reference
        with open(ref_path) as f:
            r = json.load(f)
        assert np.allclose(r["T"], T)
        plain = [(c["gpu_s"], np.array(c["cv"])) for c in r["checkpoints"]]
    else:
        # plain walkers
        ctx = S.Context(0)
        sid = ctx.add_set(params(S.MOVE_CUBE), seeds, geom=radius)
        ctx.init()
        plain, t = [], 0.0
        while t < budget:
            ctx.run(chunk)
            t += ctx.last_run_ms() / 1e3
            checkpoint(ctx, sid, t, "plain", plain)
        ctx.close()
    ref = plain[-1][1]
    # learned walkers: cube moves while records are collected, training, learned moves
    ctx = S.Context(0)
    sid = ctx.add_set(params(S.MOVE_CUBE), seeds, geom=radius)
    ctx.init()
    learned, t = [], 0.0
    t0 = time.perf_counter()
    for k in range(16):
        ctx.run(chunk // 16)
        t += ctx.last_run_ms() / 1e3
        tc = time.perf_counter()
        ctx.prop_collect(sid, 1000 + k, 4)
        t += time.perf_counter() - tc
    tc = time.perf_counter()
    ctx.prop_init(0)
    losses = ctx.prop_train(3000, 4096, 3e-3)
    w, _ = ctx.prop_get_weights()
    st, bins = ctx.download_raw(sid)
    coords = np.concatenate([ctx.coords(sid, k) for k in range(walkers)])
    ctx.close()
    ctx = S.Context(0)
    ctx.prop_set_weights(w)
    sid = ctx.add_set(params(S.MOVE_LEARNED), seeds, geom=radius)
    ctx.init()
    ctx.upload_raw(sid, st, bins, coords=coords)
    t += time.perf_counter() - tc
    log(f"[learned] collected 1e6 records and trained in {t:.1f} s (loss {losses[:20].mean():.3f} -> "
        f"{losses[-100:].mean():.3f})")
    checkpoint(ctx, sid, t, "learned", learned)
    while t < budget:
        ctx.run(chunk // 2)
        t += ctx.last_run_ms() / 1e3
        checkpoint(ctx, sid, t, "learned", learned)
    ctx.close()

    def time_to(curves, mask):
        dev = [float(np.max(np.abs(cv[mask] / ref[mask] - 1.0))) for _, cv in curves]
        for i in range(len(curves)):
            if all(d < 0.02 for d in dev[i:]):
                return curves[i][0]
        return None

    res = {}
    for name, mask in windows.items():
        tp, tl = time_to(plain[:-1], mask), time_to(learned, mask)
        res[name] = {"plain_s": tp, "learned_s": tl,
                     "learned_over_plain": (tl / tp) if (tp and tl) else None}
    with open(out_path, "w") as f:
        json.dump({"T": T.tolist(), "reference": ref.tolist(),
                   "plain": [(t, cv.tolist()) for t, cv in plain],
                   "learned": [(t, cv.tolist()) for t, cv in learned]}, f)
    return res


LIT_DIR = os.path.join(ROOT, "tests", "golden", "lj31_cv")
# the literature files the reference holds; True: total heat capacity per atom
LIT_SOURCES = [("REM", "LJ31_Cv_Reference", True), ("RESTMC", "LJ31_Cv_Reference_alt", False),
               ("MCET", "LJ31_Cv_Reference_3", True), ("EMC", "LJ31_Cv_Reference_4", True),
               ("t-REM", "tRem_Ref", False)]


def lj31_literature():
    """name -> (T, configurational Cv), with plotting/final_heat_capacity.py:33-53's (c - 3/2) * 31."""
    out = {}
    for name, f, per_atom in LIT_SOURCES:
        T, c = np.loadtxt(os.path.join(LIT_DIR, f + ".csv"), delimiter=",", unpack=True)
        o = np.argsort(T)
        out[name] = (T[o], ((c - 1.5) * 31 if per_atom else c)[o])
    return out


def combine_entropy(parts, width):
    """One mean S(E) from several (E, S, count) reductions that share their
    normalising window: the counts-weighted mean on their common bin grid."""
    base = min(float(E[0]) for E, _, _ in parts)
    idx = [np.rint((E - base) / width).astype(np.int64) for E, _, _ in parts]
    n = int(max(i.max() for i in idx)) + 1
    tot, cnt = np.zeros(n), np.zeros(n)
    for i, (_, S_, c) in zip(idx, parts):
        np.add.at(tot, i, S_ * c)
        np.add.at(cnt, i, c)
    keep = cnt > 0
    return base + np.arange(n)[keep] * width, tot[keep] / cnt[keep]


def lj31_cv(total, walkers, lanes, chunk, out_path):
    """Cv(T) of the LJ31 configuration (sphere 3.0, SAD min_T 0.05, bins 0.01,
    FP32 pairs with FP64 sums and an FP64 refresh every 1024 moves) from two
    independent halves of the walkers (seeds 0.. and 1000000..), every `chunk`
    moves up to `total`: the halves against each other (plain vs plain) and
    the combined curve against the reference's literature curves."""
    import analysis as A
    T = np.linspace(0.05, 0.40, 71)
    half = walkers // 2
    sp = lj_params(S.FP32_PAIRS, lanes)
    ctx = S.Context(0)
    sa = ctx.add_set(sp, np.arange(half, dtype=np.uint64), geom=RADIUS)
    sb = ctx.add_set(sp, np.arange(half, dtype=np.uint64) + 1000000, geom=RADIUS)
    ctx.init()
    checkpoints = []
    done, t = 0, 0.0
    while done < total:
        k = min(chunk, total - done)
        ctx.run(k)
        done += k
        t += ctx.last_run_ms() / 1e3
        # one energy window common to both halves, so their S share the constant
        st = np.concatenate([ctx.download_raw(s, with_bins=False)[0] for s in (sa, sb)])
        window = (float(st["too_lo"].max()), float(st["too_hi"].min()))
        curves = []
        grids = []
        for s in (sa, sb):
            E, Sg, cnt = ctx.mean_entropy(s, window)
            grids.append((E, Sg, cnt))
            curves.append(A.heat_capacity(T, E, Sg))
        Ec, Sc = combine_entropy(grids, 0.01)
        cv = A.heat_capacity(T, Ec, Sc)
        checkpoints.append(dict(moves=done, gpu_s=t, cv_a=curves[0].tolist(), cv_b=curves[1].tolist(),
                                cv=cv.tolist()))
        log(f"[lj31 cv] {done} moves, {t:.1f} s: Cv(0.1) {cv[10]:.1f} Cv(0.2) {cv[30]:.1f} Cv(0.3) {cv[50]:.1f}; "
            f"halves differ by {np.max(np.abs(curves[0] / curves[1] - 1)):.3f}")
    ctx.close()
    final = np.array(checkpoints[-1]["cv"])
    windows = {"T0.05-0.10": (0.05, 0.10), "T0.10-0.20": (0.10, 0.20), "T0.20-0.30": (0.20, 0.30),
               "T0.30-0.40": (0.30, 0.40)}
    lit = lj31_literature()
    vs_lit = {}
    for name, (lt, lc) in lit.items():
        row = {}
        for w, (a, b) in windows.items():
            m = (T >= a - 1e-9) & (T <= b + 1e-9) & (T >= lt[0]) & (T <= lt[-1])
            ref = np.interp(T[m], lt, lc)
            row[w] = float(np.max(np.abs(final[m] / ref - 1.0)))
        vs_lit[name] = row
    spread = {}
    for w, (a, b) in windows.items():
        m = (T >= a - 1e-9) & (T <= b + 1e-9)
        c = np.array([np.interp(T[m], lt, lc) for lt, lc in lit.values()])
        spread[w] = float(np.max((c.max(0) - c.min(0)) / c.mean(0)))
    halves = {w: float(np.max(np.abs(np.array(checkpoints[-1]["cv_a"])[(T >= a - 1e-9) & (T <= b + 1e-9)]
                                     / np.array(checkpoints[-1]["cv_b"])[(T >= a - 1e-9) & (T <= b + 1e-9)] - 1)))
              for w, (a, b) in windows.items()}

    def time_to_2pct(key, a, b):
        m = (T >= a - 1e-9) & (T <= b + 1e-9)
        dev = [float(np.max(np.abs(np.array(c[key])[m] / final[m] - 1))) for c in checkpoints[:-1]]
        for i in range(len(dev)):
            if all(d < 0.02 for d in dev[i:]):
                return checkpoints[i]["gpu_s"]
        return None

    with open(out_path, "w") as f:
        json.dump({"T": T.tolist(), "checkpoints": checkpoints,
                   "literature": {k: [v[0].tolist(), v[1].tolist()] for k, v in lit.items()}}, f)
    return {"walkers": walkers, "moves_per_walker": total, "gpu_s": checkpoints[-1]["gpu_s"],
            "max_rel_dev_vs_literature": vs_lit, "literature_spread": spread, "halves_max_rel_dev": halves,
            "plain_time_to_2pct_s": {w: time_to_2pct("cv", a, b) for w, (a, b) in windows.items()},
            "Cv_at": {str(x): float(final[int(round((x - 0.05) / 0.005))]) for x in (0.05, 0.1, 0.2, 0.3, 0.35)}}


# The joint autotuner's choice for the mixed batch below (lanes per walker by
# size class, wavefronts per workgroup), as `--batched` measures it; the
# default run times the batch at this configuration instead of re-tuning.
BATCH_TUNED = ({"LJ13": 8, "LJ19": 4, "LJ31": 4, "LJ38": 8, "WCA32": 4, "WCA64": 8, "WCA128": 16, "WCA256": 32}, 1)


def batched(moves, per_size, launches, tuned=None):
    """The mixed batch: LJ clusters N = 13, 19, 31, 38 in spheres of radius
    2.5, 2.7, 3.0, 3.2 sigma and periodic WCA fluids N = 32, 64, 128, 256
    with the packing fraction drawn per walker from U[0.3, 0.5]; `per_size`
    walkers each, SAD min_T = 0.1, cube moves of half-width 0.05 sigma, FP64
    energy refresh every 1024 moves.  Lanes per walker come from the
    autotuner (timed on as many walkers as the batch holds); the whole batch
    runs as one persistent launch captured in a HIP graph, timed against the
    per-set launches of the same walkers.  tuned = (lanes by class name,
    wavefronts per workgroup) skips the autotuner."""
    rng = np.random.default_rng(7)
    specs = []
    for n, R in ((13, 2.5), (19, 2.7), (31, 3.0), (38, 3.2)):
        specs.append((dict(system=S.LJ, n=n, energy_bin=0.01, has_max_e=1, max_e=0.0), np.full(per_size, R)))
    for n in (32, 64, 128, 256):
        eta = rng.uniform(0.3, 0.5, per_size)
        specs.append((dict(system=S.WCA, n=n, energy_bin=0.05, has_max_e=1, max_e=2.5 * n),
                      np.cbrt(n / (6.0 * eta / np.pi))))
    sets = []
    for k, (kw, geom) in enumerate(specs):
        sp = S.make_params(method=S.SAD, method_param=0.1, move_param=0.05, refresh=1024, **kw)
        seeds = np.arange(per_size, dtype=np.uint64) + 100000 * k
        if tuned:
            sp.lanes = tuned[0][f"{'LJ' if kw['system'] == S.LJ else 'WCA'}{kw['n']}"]
            sets.append((sp, seeds, geom))
            continue
        # start from each class timed alone (on as many walkers as the whole batch holds)
        trial = np.arange(len(specs) * per_size, dtype=np.uint64)
        best, rates = S.autotune_lanes(sp, trial, np.resize(geom, len(trial)), warm_moves=20000, trial_moves=2000)
        sp.lanes = best
        log(f"[batched] {'LJ' if kw['system'] == S.LJ else 'WCA'}{kw['n']}: alone, lanes {best} "
            + " ".join(f"{g}:{r / 1e9:.2f}" for g, r in sorted(rates.items())))
        sets.append((sp, seeds, geom))
    if tuned:
        waves = tuned[1]
    else:
        # then jointly: every class's lanes and the workgroup's wavefronts, the whole batch timed
        lanes, waves, rate = S.autotune_batch(sets, warm_moves=20000, trial_moves=2000)
        for (sp, _, _), g in zip(sets, lanes):
            sp.lanes = g
        log(f"[batched] joint: lanes {lanes}, {waves} wavefronts per workgroup, {rate / 1e9:.2f} G moves/s")
    out = {}
    for mode in ("per_set", "graph"):
        ctx = S.Context(0)
        for sp, seeds, geom in sets:
            ctx.add_set(sp, seeds, geom=geom)
        ctx.init()
        ctx.set_block_waves(-1, waves)
        # one untimed launch (captures the graph)
        if mode == "graph":
            ctx.run_batched(moves, graph=True)
        else:
            ctx.run(moves)
        ms = []
        for _ in range(launches):
            if mode == "graph":
                ctx.run_batched(moves, graph=True)
            else:
                ctx.run(moves)
            ms.append(ctx.last_run_ms())
        ctx.close()
        total = len(sets) * per_size * moves
        out[mode] = {"moves_per_s": total / (np.mean(ms) / 1e3), "ms_per_launch": float(np.mean(ms))}
        log(f"[batched] {mode}: {out[mode]['moves_per_s'] / 1e9:.3f} G moves/s")
    out["lanes"] = {f"{'LJ' if sp.system == S.LJ else 'WCA'}{sp.n}": sp.lanes for sp, _, _ in sets}
    out["waves_per_workgroup"] = waves
    # algorithmic work of one launch: 30 (N-1) flop per LJ move, 48 (N-1) per WCA move (pair terms of the
    # old and new position, minimum image for WCA); 80 B of bin traffic per move
    flop = sum(per_size * moves * (30 if sp.system == S.LJ else 48) * (sp.n - 1) for sp, _, _ in sets)
    nmov = len(sets) * per_size * moves
    t = out["graph"]["ms_per_launch"] / 1e3
    out["roofline"] = {"valu_fp64": {"achieved": flop / t / 1e12, "peak": FP64_VECTOR_PEAK, "unit": "TFLOP/s",
                                     "frac": flop / t / 1e12 / FP64_VECTOR_PEAK},
                       "hbm": {"achieved": BYTES_PER_MOVE * nmov / t / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": BYTES_PER_MOVE * nmov / t / 1e9 / HBM_PEAK_GBS}}
    return out


def tempering(rounds, ladders, replicas, lanes, cpu_seconds):
    """Parallel tempering of LJ31 clusters (sphere R = 3.0): `ladders`
    independent ladders of `replicas` canonical replicas at temperatures
    0.05 * 1.15^r, cube moves of half-width 0.1 sigma, one round = 31 moves
    per replica and one sweep of neighbour swaps.  Value: replica moves/s over
    one launch of `rounds` rounds (inputs resident on the device).  The CPU
    baseline runs the oracle's ladder, one ladder per host thread, for about
    `cpu_seconds`."""
    temps = 0.05 * 1.15 ** np.arange(replicas)
    sp = S.make_params(system=S.LJ, n=N_ATOMS, lanes=lanes, move_param=0.1)
    seeds = np.arange(ladders, dtype=np.uint64) + 1000
    ctx = S.Context(0)
    pt = ctx.pt_add(sp, temps, seeds, geom=RADIUS)
    ctx.pt_run(pt, 20)
    t0 = time.perf_counter()
    ctx.pt_run(pt, rounds)
    host_wall = time.perf_counter() - t0
    wall = ctx.last_run_ms() / 1e3  # the launch on the device (HIP events)
    reps = ctx.pt_replicas(pt)
    ctx.close()
    moves = ladders * replicas * N_ATOMS * rounds
    swaps = reps["accepted_swaps"].sum() / max(1, (reps["accepted_swaps"] + reps["rejected_swaps"]).sum())
    out = {"moves_per_s": moves / wall, "kernel_s": wall, "host_wall_s": host_wall, "ladders": ladders,
           "replicas": replicas, "lanes": lanes,
           "rounds": rounds, "swap_acceptance": float(swaps),
           # FP64 pair terms, no bin traffic: the VALU roofline bounds it
           "roofline": {"bound": "valu-fp64", "achieved": FLOP_PER_MOVE * moves / wall / 1e12,
                        "peak": FP64_VECTOR_PEAK, "unit": "TFLOP/s",
                        "frac": FLOP_PER_MOVE * moves / wall / 1e12 / FP64_VECTOR_PEAK}}
    log(f"[tempering] gpu {out['moves_per_s'] / 1e9:.3f} G replica moves/s, swap acceptance {swaps:.3f}")
    if cpu_seconds > 0:
        import threading
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        threads = min(16, os.cpu_count() or 1)
        op = O.make_params(system=O.LJ, n=N_ATOMS, move_param=0.1, radius=RADIUS, order_g=lanes)
        pts = [O.Tempering(op, temps, seed=int(s), coords=S.init_coords(S.LJ, N_ATOMS, RADIUS, int(s)))
               for s in seeds[:threads]]
        t0 = time.perf_counter()
        pts[0].run(10)
        per = max(1, int(10 * cpu_seconds / max(time.perf_counter() - t0, 1e-6) / 2))
        t0 = time.perf_counter()
        ths = [threading.Thread(target=p.run, args=(per,)) for p in pts]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        cw = time.perf_counter() - t0
        cmoves = threads * replicas * N_ATOMS * per
        out["cpu_baseline"] = {"value": cmoves / cw, "unit": "replica moves/s", "cores": threads, "kind": "port",
                               "sample": f"{threads} ladders x {per} rounds on {threads} threads"}
        out["gpu_over_cpu"] = out["moves_per_s"] / out["cpu_baseline"]["value"]
        log(f"[tempering] cpu {cmoves / cw / 1e6:.2f} M replica moves/s on {threads} threads")
    return out


def binning_path(moves, walkers, lanes, cpu_seconds, linear=False):
    """The binning binary's walker (energy_binning.rs over a histogram) on the
    LJ31 configuration: 16384 walkers, N = 31 in a sphere of radius 3.0, cube
    moves 0.1, SAD min_T = 0.05, bins 0.01, max_allowed_energy 0, FP64.
    One untimed launch of 2000 moves, then one timed launch of `moves` moves
    per walker.  CPU baseline: the oracle's walker, one walker per host
    thread, for about `cpu_seconds`."""
    kind = S.BIN_LINEAR if linear else S.BIN_HISTOGRAM
    sp = S.make_params(system=S.LJ, n=N_ATOMS, lanes=lanes, method=S.SAD, method_param=0.05, move_param=0.1,
                       energy_bin=0.01, has_max_e=1, max_e=0.0, binning=kind)
    seeds = np.arange(walkers, dtype=np.uint64)
    ctx = S.Context(0)
    s = ctx.bin_add_set(sp, seeds, geom=RADIUS)
    ctx.bin_run(s, WARMUP_MOVES)
    t0 = time.perf_counter()
    ctx.bin_run(s, moves)
    host_wall = time.perf_counter() - t0
    wall = ctx.last_run_ms() / 1e3  # the launch on the device (HIP events)
    lens = [int(ctx.bin_state(s, w)["len"]) for w in range(0, walkers, max(1, walkers // 64))]
    ctx.close()
    total = walkers * moves
    gbs = BYTES_PER_MOVE * total / wall / 1e9
    out = {"moves_per_s": total / wall, "kernel_s": wall, "host_wall_s": host_wall, "walkers": walkers,
           "lanes": lanes, "moves": moves,
           "binning": "linear" if linear else "histogram",
           "warmup_moves": WARMUP_MOVES, "mean_bins": float(np.mean(lens)),
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": gbs / HBM_PEAK_GBS, "bytes_per_move": BYTES_PER_MOVE}}
    log(f"[binning] gpu {out['moves_per_s'] / 1e9:.3f} G moves/s ({wall:.2f} s)")
    if cpu_seconds > 0:
        # the same segment of each walker's run (moves WARMUP_MOVES .. WARMUP_MOVES + moves) for as many
        # walkers per thread as fit in about cpu_seconds; the rate of a thread is its timed moves over its
        # timed seconds, the threads run side by side
        import threading
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        threads = min(16, os.cpu_count() or 1)

        def params(seed):
            return O.make_params(system=O.LJ, n=N_ATOMS, method=O.SAD, method_param=0.05, move_param=0.1,
                                 energy_bin=0.01, has_max_e=1, max_e=0.0, radius=RADIUS, order_g=lanes,
                                 seed=int(seed), binning=1 if linear else 0)
        probe = O.BWalker(params(0))
        probe.run(WARMUP_MOVES)
        t0 = time.perf_counter()
        probe.run(moves)
        per_walker = time.perf_counter() - t0
        del probe
        reps = max(1, int(round(cpu_seconds / max(per_walker, 1e-6))))
        timed = [0.0] * threads

        def work(t):
            for r in range(reps):
                w = O.BWalker(params(t + threads * r))
                w.run(WARMUP_MOVES)
                t1 = time.perf_counter()
                w.run(moves)
                timed[t] += time.perf_counter() - t1
                del w
        ths = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        rate = sum(reps * moves / x for x in timed)
        out["cpu_baseline"] = {"value": rate, "unit": "moves/s", "cores": threads, "kind": "port",
                               "sample": f"{threads} threads x {reps} walkers, each timed over moves "
                                         f"{WARMUP_MOVES}..{WARMUP_MOVES + moves} (the GPU's timed segment)"}
        out["gpu_over_cpu"] = out["moves_per_s"] / rate
        log(f"[binning] cpu {rate / 1e6:.2f} M moves/s on {threads} threads")
    return out


def zmc(rounds, runs, lanes, cpu_seconds):
    """Energy-ceiling replicas (energy_replicas.rs) of LJ31 clusters in a
    sphere of radius 3.0: `runs` independent runs of up to 16 replicas, 8
    lanes each, min_T = 0.05, a new replica after 64 unique visitors; one
    untimed launch of 200 rounds (replicas split there), then one timed
    launch of `rounds` rounds.  Value: replica moves/s (N = 31 moves per
    replica per round; the unbounded replica's randomization counts as N, as
    the reference counts it).  CPU baseline: the oracle's runs, one run per
    host thread, over the same rounds."""
    sp = S.make_params(system=S.LJ, n=N_ATOMS, lanes=lanes, move_param=0.1)
    seeds = np.arange(runs, dtype=np.uint64) + 500
    ctx = S.Context(0)
    t0 = time.perf_counter()
    zid = ctx.zmc_add(sp, seeds, geom=RADIUS, min_T=0.05, isbnb=64, max_replicas=16)
    init_s = time.perf_counter() - t0
    ctx.zmc_run(zid, 200)
    m0 = sum(ctx.zmc_state(zid, s)[1] for s in range(runs))
    t0 = time.perf_counter()
    ctx.zmc_run(zid, rounds)
    host_wall = time.perf_counter() - t0
    wall = ctx.last_run_ms() / 1e3  # the launch on the device (HIP events)
    m1 = sum(ctx.zmc_state(zid, s)[1] for s in range(runs))
    reps = [len(ctx.zmc_replicas(zid, s)) for s in range(runs)]
    ctx.close()
    out = {"moves_per_s": (m1 - m0) / wall, "kernel_s": wall, "host_wall_s": host_wall, "runs": runs,
           "lanes": lanes, "rounds": rounds,
           "host_init_s": init_s, "replicas_mean": float(np.mean(reps)), "replicas_max": int(np.max(reps)),
           "roofline": {"bound": "valu-fp64", "achieved": FLOP_PER_MOVE * (m1 - m0) / wall / 1e12,
                        "peak": FP64_VECTOR_PEAK, "unit": "TFLOP/s",
                        "frac": FLOP_PER_MOVE * (m1 - m0) / wall / 1e12 / FP64_VECTOR_PEAK}}
    log(f"[zmc] gpu {out['moves_per_s'] / 1e9:.3f} G replica moves/s, {np.mean(reps):.1f} replicas per run")
    if cpu_seconds > 0:
        # the same rounds of each run (200 untimed, then `rounds` timed) for runs_per_thread runs per host
        # thread; a thread's rate is its timed moves over its timed seconds, the threads run side by side
        import threading
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        threads = min(16, os.cpu_count() or 1)
        op = O.make_params(system=O.LJ, n=N_ATOMS, move_param=0.1, radius=RADIUS, order_g=lanes)
        probe = O.Zmc(op, seed=int(seeds[0]), min_T=0.05, isbnb=64)
        t0 = time.perf_counter()
        probe.run(200 + rounds)
        per_run = time.perf_counter() - t0 + 0.1
        del probe
        reps = max(1, min(64, int(cpu_seconds / per_run)))
        timed = [0.0] * threads
        done = [0] * threads

        def work(t):
            for k in range(reps):
                z = O.Zmc(op, seed=int(seeds[(t + threads * k) % len(seeds)]), min_T=0.05, isbnb=64)
                z.run(200)
                m = z.moves()
                t1 = time.perf_counter()
                z.run(rounds)
                timed[t] += time.perf_counter() - t1
                done[t] += z.moves() - m
        log(f"[zmc] cpu baseline: {threads} threads x {reps} runs")
        ths = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        rate = sum(d / x for d, x in zip(done, timed))
        out["cpu_baseline"] = {"value": rate, "unit": "replica moves/s", "cores": threads, "kind": "port",
                               "sample": f"{threads} threads x {reps} runs, each timed over rounds "
                                         f"200..{200 + rounds} (the GPU's timed rounds)"}
        out["gpu_over_cpu"] = out["moves_per_s"] / rate
        log(f"[zmc] cpu {rate / 1e6:.2f} M replica moves/s on {threads} threads")
    return out


def entropy(moves, walkers, lanes, repeats):
    """The mean S(E) of a walker set reduced on the device (the LJ31 SAD
    configuration after `moves` moves per walker), timed against the same
    reduction on the host from downloaded bins (analysis.mean_entropy, the
    restatement of parse-binning.py:150-171 over walkers).  Algorithmic
    bytes: each walker's 16-byte hot records over its bins (read once)."""
    sys.path.insert(0, os.path.join(ROOT, "sad-monte-carlo_amd"))
    import analysis
    ctx = S.Context(0)
    sid = ctx.add_set(lj_params(S.FP32_PAIRS, lanes), np.arange(walkers, dtype=np.uint64), geom=RADIUS)
    ctx.init()
    ctx.run(moves)
    E, Sd, cnt = ctx.mean_entropy(sid)
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        ctx.mean_entropy(sid)
        times.append(time.perf_counter() - t0)
    dev = float(np.median(times))
    st, rows = ctx.download_raw(sid)
    t0 = time.perf_counter()
    pairs = []
    for w in range(walkers):
        b = S.logical_bins(st[w], rows[w])
        pairs.append((st[w], b))
    Eh, Sh, ch = analysis.mean_entropy(pairs)
    host = time.perf_counter() - t0
    ctx.close()
    nbins = int(np.sum(st["len"]))
    gbs = 16.0 * nbins / dev / 1e9
    same = len(E) == len(Eh) and np.allclose(E, Eh, atol=1e-9) and np.array_equal(cnt, ch)
    out = {"device_s": dev, "host_s": host, "walkers": walkers, "moves": moves, "bins_total": nbins,
           "grid_bins": int(len(E)), "same_grid_as_host": bool(same),
           "max_abs_diff_vs_host": float(np.max(np.abs(Sd - Sh))) if same else None,
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": gbs / HBM_PEAK_GBS, "bytes_per_bin": 16},
           "cpu_baseline": {"value": host, "unit": "s", "cores": 1, "kind": "port",
                            "sample": "analysis.mean_entropy over the same walkers' downloaded bins"},
           "gpu_over_cpu": host / dev}
    log(f"[entropy] device {dev * 1e3:.2f} ms, host {host:.2f} s over {walkers} walkers, {nbins} bins")
    return out


def learned_rate(moves, walkers, lanes):
    """Cost of the learned move: the learned-proposal workload (LJ38 in a
    3.2 sigma sphere, SAD min_T 0.03, bins 0.01, max energy 0, FP64 pair
    terms) with an untrained network (prop_init(7)), learned moves against
    cube moves on the same walkers: one untimed launch of 2e5 moves, then
    `moves` moves timed by HIP events."""
    out = {}
    for name, kind in (("cube", S.MOVE_CUBE), ("learned", S.MOVE_LEARNED)):
        sp = S.make_params(system=S.LJ, n=38, lanes=lanes, method=S.SAD, method_param=0.03, move_param=0.1,
                           energy_bin=0.01, has_max_e=1, max_e=0.0, move_kind=kind)
        ctx = S.Context(0)
        sid = ctx.add_set(sp, np.arange(walkers, dtype=np.uint64), geom=3.2)
        ctx.prop_init(7)
        ctx.init()
        ctx.run(200000)
        ctx.run(moves)
        ms = ctx.last_run_ms()
        ctx.close()
        out[name] = walkers * moves / (ms / 1e3)
        log(f"[learned rate] {name}: {out[name] / 1e9:.3f} G moves/s ({ms:.1f} ms)")
    out["learned_over_cube_cost"] = out["cube"] / out["learned"]
    return out


def other_paths(walkers, lanes):
    """Short runs of the engine's other kernels on the same box as the
    headline line, each on its own workload (the `--batched`, `--binning`,
    `--zmc`, `--tempering` and `--entropy` modes at reduced lengths, GPU
    only): moves/s (or seconds) and the fraction of the roofline that bounds
    each kernel."""
    def summary(res, value_key="moves_per_s", unit="moves/s"):
        r = res["roofline"]
        return {"value": res[value_key], "unit": unit, "roofline_bound": r["bound"], "roofline_frac": r["frac"]}

    out = {}
    log("[paths] mixed LJ/WCA batch, graph-captured, at the recorded autotuned configuration")
    b = batched(20000, 4096, 3, tuned=BATCH_TUNED)
    out["k_batched"] = {"workload": "8 size classes x 4096 walkers (LJ13/19/31/38, WCA32/64/128/256), 20000 "
                                    "moves per launch, 3 launches",
                        "value": b["graph"]["moves_per_s"], "unit": "moves/s",
                        "per_set_launches_moves_per_s": b["per_set"]["moves_per_s"],
                        "roofline_bound": "valu-fp64", "roofline_frac": b["roofline"]["valu_fp64"]["frac"]}
    log("[paths] binning-path walkers")
    r = binning_path(50000, walkers, lanes, 0.0)
    out["k_bin"] = {"workload": f"LJ31 histogram binning, {walkers} walkers, moves 2000-52000", **summary(r)}
    log("[paths] energy-ceiling replicas")
    r = zmc(3000, 2048, lanes, 0.0)
    out["k_zmc"] = {"workload": "LJ31, 2048 runs of up to 16 replicas, rounds 200-3200", **summary(r)}
    log("[paths] parallel tempering")
    r = tempering(4000, 1024, 16, lanes, 0.0)
    out["k_pt"] = {"workload": "LJ31, 1024 ladders x 16 replicas, 4000 rounds", **summary(r)}
    log("[paths] walker-set entropy reduction")
    r = entropy(100000, walkers, lanes, 5)
    out["k_entropy"] = {"workload": f"LJ31, {walkers} walkers after 1e5 moves", **summary(r, "device_s", "s")}
    return out


def copy_probe(nbytes):
    """A plain device-to-device copy of known size, for scaling the HBM byte
    counters (FETCH_SIZE, WRITE_SIZE) of a profiler run."""
    import torch
    a = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copied {nbytes} bytes three times")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--moves-per-step", type=int, default=500000)
    ap.add_argument("--walkers", type=int, default=16384)
    ap.add_argument("--lanes", type=int, default=8)
    ap.add_argument("--cpu-rounds", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-wide", action="store_true", help="skip the FP64 run with 8x the walkers at 4 lanes")
    ap.add_argument("--no-paths", action="store_true", help="skip the short runs of the other kernels")
    ap.add_argument("--copy-probe", type=int, default=0, help="only run a device copy of this many bytes")
    ap.add_argument("--learned-cv", type=float, default=0.0,
                    help="only run the LJ38 wall-time-to-Cv comparison with this GPU-time budget (s) per method")
    ap.add_argument("--learned-cv-out", default="learned_cv.json")
    ap.add_argument("--learned-cv-system", default="lj38", choices=["lj38", "lj13", "lj31"])
    ap.add_argument("--learned-cv-ref", default=None,
                    help="with --learned-cv: the plain curves from a --lj31-cv output instead of a plain run")
    ap.add_argument("--learned-cv-chunk", type=float, default=1e6, help="with --learned-cv: moves per checkpoint")
    ap.add_argument("--learned-rate", type=int, default=0,
                    help="only time this many learned and cube moves of the learned-proposal workload")
    ap.add_argument("--batched", type=int, default=0,
                    help="only run the mixed LJ/WCA batch (this many moves per launch)")
    ap.add_argument("--binning", type=int, default=0,
                    help="only run the binning-path LJ31 walkers for this many moves per walker")
    ap.add_argument("--binning-linear", action="store_true", help="with --binning: the linear interpolation")
    ap.add_argument("--zmc", type=int, default=0,
                    help="only run LJ31 energy-ceiling replica runs for this many rounds")
    ap.add_argument("--pt-ladders", type=int, default=1024,
                    help="with --tempering: ladders of 16 replicas (1024 x 16 = the 16384 walkers of the LJ31 case)")
    ap.add_argument("--zmc-runs", type=int, default=2048,
                    help="with --zmc: independent runs (2048 = 8 runs per CU, all resident at once)")
    ap.add_argument("--entropy", type=int, default=0,
                    help="only time the walker-set S(E) reduction after this many moves per walker")
    ap.add_argument("--lj31-cv", type=float, default=0.0,
                    help="only run the LJ31 Cv(T) comparison: this many moves per walker (two seed halves)")
    ap.add_argument("--lj31-cv-chunk", type=float, default=1e7, help="with --lj31-cv: moves between checkpoints")
    ap.add_argument("--lj31-cv-out", default="lj31_cv.json")
    ap.add_argument("--tempering", type=int, default=0,
                    help="only run LJ31 parallel tempering ladders for this many rounds")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (walkers are independent chains and shard by walker); "
              f"--gpus {a.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.learned_rate > 0:
        res = learned_rate(a.learned_rate, a.walkers, a.lanes)
        print(json.dumps({"metric": "LJ38 learned-proposal moves/s on 1 GPU", "unit": "moves/s",
                          "higher_is_better": True, "value": res["learned"], "walkers": a.walkers,
                          "lanes": a.lanes, "moves": a.learned_rate, **res}))
        return
    if a.batched > 0:
        res = batched(a.batched, 4096, a.steps)
        print(json.dumps({"metric": "mixed LJ/WCA SAD batch moves/s on 1 GPU", "unit": "moves/s",
                          "higher_is_better": True, "walkers_per_size": 4096, "moves_per_launch": a.batched,
                          "launches": a.steps, **res}))
        return
    if a.binning > 0:
        res = binning_path(a.binning, a.walkers, a.lanes, 0.0 if a.no_cpu else 15.0, a.binning_linear)
        print(json.dumps({"metric": "LJ31 SAD moves/s of the binning-path walker on 1 GPU", "unit": "moves/s",
                          "higher_is_better": True, "value": res["moves_per_s"], **res}))
        return
    if a.zmc > 0:
        res = zmc(a.zmc, a.zmc_runs, a.lanes, 0.0 if a.no_cpu else 15.0)
        print(json.dumps({"metric": "LJ31 energy-ceiling replica moves/s on 1 GPU", "unit": "moves/s",
                          "higher_is_better": True, "value": res["moves_per_s"], **res}))
        return
    if a.entropy > 0:
        res = entropy(a.entropy, a.walkers, a.lanes, 5)
        print(json.dumps({"metric": "LJ31 walker-set S(E) reduction time on 1 GPU", "unit": "s",
                          "higher_is_better": False, "value": res["device_s"], **res}))
        return
    if a.tempering > 0:
        res = tempering(a.tempering, a.pt_ladders, 16, a.lanes, 0.0 if a.no_cpu else 15.0)
        print(json.dumps({"metric": "LJ31 parallel-tempering replica moves/s on 1 GPU", "unit": "moves/s",
                          "higher_is_better": True, "value": res["moves_per_s"], **res}))
        return
    if a.lj31_cv > 0:
        res = lj31_cv(int(a.lj31_cv), a.walkers, a.lanes, int(a.lj31_cv_chunk), a.lj31_cv_out)
        print(json.dumps({"metric": "LJ31 Cv(T): plain walkers vs the literature and vs each other", "unit": "",
                          **res}))
        return
    if a.learned_cv > 0:
        lanes = a.lanes if a.learned_cv_system != "lj13" else 4
        res = learned_cv(a.learned_cv, a.walkers, lanes, int(a.learned_cv_chunk), a.learned_cv_out,
                         a.learned_cv_system, a.learned_cv_ref)
        print(json.dumps({"metric": f"{a.learned_cv_system.upper()} wall time to Cv(T) within 2% (plain vs learned)",
                          "unit": "s",
                          "higher_is_better": False, "budget_s": a.learned_cv, "walkers": a.walkers,
                          "windows": res}))
        return
    main_run = run_gpu(S.FP32_PAIRS, a.walkers, a.lanes, a.warmup, a.steps, a.moves_per_step, "lj31 fp32 pairs")
    moves_per_launch = a.walkers * a.moves_per_step
    achieved = FLOP_PER_MOVE * moves_per_launch / (main_run["kernel_ms"] * 1e-3) / 1e12
    achieved_gbs = BYTES_PER_MOVE * moves_per_launch / (main_run["kernel_ms"] * 1e-3) / 1e9
    traffic = measured_traffic(a) if (a.walkers, a.moves_per_step) == (16384, 500000) else None
    run_moves = a.steps * a.moves_per_step
    variants = {}
    if not a.no_fp64:
        r = run_gpu(S.FP64, a.walkers, a.lanes, a.warmup, a.steps, a.moves_per_step, "lj31 fp64")
        variants["fp64"] = dict(value=r["value"], ms_per_step=r["wall"] * 1e3 / a.steps, kernel_ms=r["kernel_ms"],
                                max_rel_energy_error=r["max_rel_energy_error"],
                                roofline_frac=FLOP_PER_MOVE * moves_per_launch / (r["kernel_ms"] * 1e-3) / 1e12
                                / FP64_VECTOR_PEAK)
        if not a.no_wide:
            # the same FP64 walk with 8x the walkers at 4 lanes: enough waves per SIMD to hide the latency
            # of each move's dependent chain (informative; the configuration's figure is `value` above)
            # (its bins take about 113 GB of HBM; where that does not fit, the line says so and goes on)
            ww, wl, wsteps, wmps = 8 * a.walkers, 4, 2, 250000
            try:
                r = run_gpu(S.FP64, ww, wl, a.warmup, wsteps, wmps, "lj31 fp64 wide")
                variants["fp64_wide"] = dict(value=r["value"], walkers=ww, lanes_per_walker=wl,
                                             timed_moves_per_walker=wsteps * wmps, kernel_ms=r["kernel_ms"],
                                             max_rel_energy_error=r["max_rel_energy_error"],
                                             roofline_frac=FLOP_PER_MOVE * ww * wmps / (r["kernel_ms"] * 1e-3)
                                             / 1e12 / FP64_VECTOR_PEAK)
            except S.SadmcError as e:
                log(f"[lj31 fp64 wide] not run: {e}")
                variants["fp64_wide_error"] = str(e)
    out = {
        "metric": "LJ31 SAD moves/s on 1 GPU",
        "value": main_run["value"],
        "unit": "moves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": main_run["wall"] * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair terms, f64 sums",
        "data": "synthetic",
        "config": {
            "workload": "LJ31 SAD precision configuration: 16384 walkers x 1e7 moves, N=31, sphere R=3.0, cube "
                        "moves 0.1, min_T=0.05, bins 0.01, max_allowed_energy 0, FP32 pairs + FP64 sums, FP64 "
                        "refresh every 1024 moves",
            "walkers": a.walkers, "lanes_per_walker": a.lanes, "moves_per_step": a.moves_per_step,
            "timed_moves_per_walker": run_moves, "warmup_moves_per_step": WARMUP_MOVES,
        },
        # 900 flop / 80 B = 11.25 flop/B, under the FP32 ridge of 157.3 / 8 = 19.7 flop/B: the roofline that
        # bounds this kernel is HBM; the VALU fraction is given beside it
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_move": BYTES_PER_MOVE, "flop_per_move": FLOP_PER_MOVE,
                     "kernel_ms_per_launch": main_run["kernel_ms"],
                     "valu": {"achieved": achieved, "peak": FP32_VECTOR_PEAK, "unit": "TFLOP/s",
                              "frac": achieved / FP32_VECTOR_PEAK}},
        "energy_error": {"max_rel": main_run["max_rel_energy_error"], "mean_rel": main_run["mean_rel_energy_error"]},
        "variants": variants,
        "learned_cv": recorded_learned_cv(),
    }
    if not a.no_paths:
        out["paths"] = other_paths(a.walkers, a.lanes)
    if not a.no_cpu:
        log("cpu baseline: C oracle, one walker per thread")
        out["cpu_baseline"] = cpu_baseline(1, run_moves, a.cpu_rounds)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        if "fp64" in variants:
            c64 = cpu_baseline(0, run_moves, a.cpu_rounds)
            variants["fp64"]["cpu_baseline"] = c64
            variants["fp64"]["gpu_over_cpu"] = variants["fp64"]["value"] / c64["value"]
            if "fp64_wide" in variants:
                variants["fp64_wide"]["gpu_over_cpu"] = variants["fp64_wide"]["value"] / c64["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
