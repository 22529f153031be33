"""Benchmark on one MI355X, printed as one JSON line.

1. particle-steps/s of the FP64 DEM step on the dense periodic shear box
   (1,000,000 grains, Hertz-Mindlin, 5% bonded pairs): the line's value.
2. Darcy ms-to-1e-8 on the consolidation box (200,000 grains, 48 x 48 x 96
   Darcy grid, 20 kPa load): the "darcy" object of the line.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 200] [--n 1000000]
                    [--darcy-steps 500] [--no-darcy]

Untimed preparation: dense packings grown on the GPU (radius growth, then
quenched relaxation under gravity), then the run configuration (walls,
driven layer, bonds, fluid).  Timed: exactly --steps steps after --warmup
steps, with every input resident in HBM; per-kernel device time from HIP
events on the library's stream in a separate run.  The Darcy solve time of
each fluid step runs from the first Jacobi sweep to convergence (HIP events
around the sweep loop).
"""
import argparse
import json
import os
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def measured_traffic():
    """HBM bytes per step of the contact stage from the committed rocprofv3
    counter summary (FETCH_SIZE + WRITE_SIZE passes, calibrated against a
    copy kernel of known size, cross-checked by the 128-B read request
    count), or None if it is missing."""
    path = os.path.join(ROOT, "profiles", "shear_box_1m_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    total = 0.0
    found = 0
    for line in open(path):
        cells = [c.strip() for c in line.split("|")]
        if len(cells) > 4 and ("k_forces" in cells[1] or "k_topology" in cells[1]):
            mb = [c for c in cells if c.endswith(" MB")]  # the scaled bytes column
            if not mb:
                continue  # a row of another table (instruction counters)
            total += float(mb[-1].split()[0]) * 1e6
            found += 1
    return total if found == 2 else None


def contact_bytes_per_particle(zbar):
    # state in 80 B + slot mask 4 B + outputs 112 B, 52 B per contact (index 4,
    # tangential spring read 24 + write 24)
    return 196.0 + 52.0 * zbar


def percentiles(v):
    v = numpy.asarray(v, dtype=numpy.float64)
    return {"mean": float(v.mean()), "p50": float(numpy.percentile(v, 50)),
            "p95": float(numpy.percentile(v, 95)), "min": float(v.min()), "max": float(v.max())}


def run_darcy(engine, s, warmup, steps):
    """Fluid steps one at a time; per step the Jacobi solve time (first sweep
    to convergence, device clock) and its sweep count."""
    import sphere_amd
    engine.load(s.to_state())
    for _ in range(warmup):
        engine.step(1)
    ms = []
    sweeps = []
    prev = engine.darcy_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        engine.step(1)
        cur = engine.darcy_stats()
        ms.append(cur.solve_ms - prev.solve_ms)
        sweeps.append(cur.sweeps - prev.sweeps)
        prev = cur
    elapsed = time.perf_counter() - t0
    return numpy.array(ms), numpy.array(sweeps), elapsed, prev


def darcy_packing(args, engine):
    import sphere_amd
    from sphere_amd import workloads as W
    if args.darcy_load:
        s = sphere_amd.sim(fluid=True)
        s.readbin(args.darcy_load, verbose=False)
        s.fluid = True
        log("loaded the consolidation box from %s" % args.darcy_load)
    else:
        log("preparing the consolidation box (%d grains)" % args.darcy_n)
        s = W.prepare_consolidation(engine, args.darcy_n, args.darcy_seed, 300, 3, 400, log=log)
    if args.darcy_save:
        os.makedirs(args.darcy_save, exist_ok=True)
        s.writebin(folder=args.darcy_save, verbose=False)
        log("wrote %s/%s.bin" % (args.darcy_save, s.sid))
    return s


def darcy_section(args):
    import sphere_amd
    engine = sphere_amd.Engine(0, nc=32)
    s = darcy_packing(args, engine)
    n = int(s.np[0])
    log("consolidation box: %d warm-up + %d timed fluid steps" % (args.darcy_warmup,
                                                                   args.darcy_steps))
    ms, sweeps, elapsed, st = run_darcy(engine, s, args.darcy_warmup, args.darcy_steps)
    cells = int(s.num[0]) * int(s.num[1]) * int(s.num[2])
    sweep_bytes = 40.0 * cells
    per_sweep_s = ms.sum() / 1000.0 / max(1, sweeps.sum())
    hist_edges = [0, 1, 11, 21, 51, 101, 201, 501, 1001, 2001, 5001, 10**9]
    hist = numpy.histogram(sweeps, bins=hist_edges)[0]
    log("Darcy: %.3f ms per solve (p50 %.3f, p95 %.3f), %.1f sweeps per solve, %.2f us per "
        "sweep, %.1f ms per step overall" % (ms.mean(), numpy.percentile(ms, 50),
                                             numpy.percentile(ms, 95), sweeps.mean(),
                                             1e6 * per_sweep_s, 1000.0 * elapsed / len(ms)))
    out = {
        "metric": "Darcy ms-to-1e-8 (Jacobi solve per fluid step, first sweep to convergence)",
        "plain": {"ms": percentiles(ms), "sweeps": percentiles(sweeps),
                  "sweeps_histogram": {"edges": hist_edges[:-1], "counts": [int(c) for c in hist]},
                  "nonconverged": int(st.nonconverged)},
        "particle_steps_per_s": n * len(ms) / elapsed,
        "ms_per_step": 1000.0 * elapsed / len(ms),
        "steps": len(ms), "warmup": args.darcy_warmup,
        "config": {"workload": "Darcy-coupled consolidation: %d grains, radii U[0.009, 0.011] m, "
                               "20 kPa top wall, linear contacts k_n=k_t=1.16e9 mu=0.5, Darcy grid "
                               "%dx%dx%d, k_c=3.5e-13, mu_f=1.04e-7, beta_f=1.43e-8, p=0 top, "
                               "no flow at the base, periodic x/y, tolerance 1e-8"
                               % (n, s.num[0], s.num[1], s.num[2]),
                   "grains": n, "grid": [int(v) for v in s.num], "dt": float(s.time_dt[0])},
        "sweep_roofline": {"bound": "hbm", "achieved": sweep_bytes / per_sweep_s / 1e9,
                           "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": sweep_bytes / per_sweep_s / 1e9 / HBM_PEAK_GBS,
                           "bytes_per_sweep": sweep_bytes, "us_per_sweep": 1e6 * per_sweep_s,
                           "note": "the sweep's fields (%.1f MB) stay in the Infinity Cache"
                                   % (sweep_bytes / 1e6)},
    }
    engine.close()
    out["cpu_baseline"] = None if args.no_cpu or args.darcy_cpu_steps <= 0 else \
        darcy_cpu_leg(args, s)
    params = None
    if not args.no_net:
        params, out["learned"] = learned_section(args, s, ms, sweeps)
    out["ensemble"] = None if args.no_ensemble else darcy_ensemble_section(args, params)
    return out


def darcy_ensemble_section(args, params):
    """Independent Darcy-coupled shear boxes (config 5: grains log-uniform in
    [5,000, 60,000], per-box friction, wall stress and shear velocity, Darcy
    grids scaled to the box) stepped together as one Ensemble (one launch per
    stage of the fluid step, the Jacobi sweeps of all solves batched, the
    learned correction's forward passes as multi-box launches) against the
    same boxes stepped one at a time with sph_step, from the same states, with
    the network trained on the consolidation boxes."""
    import sphere_amd
    from sphere_amd import workloads as W
    t0 = time.time()
    sims, info = W.darcy_ensemble_boxes(count=args.ens_boxes, seed=args.ens_seed + 2)
    total = sum(p["grains"] for p in info)
    log("Darcy ensemble: %d boxes, %d grains in all, prepared in %.1f s" % (len(sims), total,
                                                                          time.time() - t0))
    states = [s.to_state() for s in sims]
    # Per-box autotune of the learned correction: from its state, each box
    # runs --dens-tune-steps fluid steps with the network and without (after
    # one untimed step each), and keeps the network where its solves took
    # less device time.  Both the ensemble and the one-at-a-time runs use
    # the choice, so every box's results can be compared bit for bit.
    use_net = [params is not None] * len(states)
    tune = None
    if params is not None and args.dens_tune_steps > 0:
        t0 = time.time()
        trial = []
        for k, st in enumerate(states):
            ms = []
            for on in (True, False):
                e = sphere_amd.Engine(0, nc=32)
                if on:
                    e.set_darcy_net(params, args.net_corrections)
                e.load(st)
                e.step(1)
                a = e.darcy_stats().solve_ms
                e.step(args.dens_tune_steps)
                ms.append((e.darcy_stats().solve_ms - a) / args.dens_tune_steps)
                e.close()
            use_net[k] = ms[0] < ms[1]
            trial.append(ms)
        off = [k for k in range(len(states)) if not use_net[k]]
        tune = {"trial_steps": args.dens_tune_steps, "boxes_with_network": len(states) - len(off),
                "boxes_without": off, "seconds": time.time() - t0,
                "solve_ms_with_without": [[round(a, 4), round(b, 4)] for a, b in trial]}
        log("Darcy ensemble: the network kept on %d of %d boxes (autotune, %.1f s; off on %s)" %
            (len(states) - len(off), len(states), tune["seconds"], off))

    def engines():
        out = []
        for k, st in enumerate(states):
            e = sphere_amd.Engine(0, nc=32)
            if use_net[k]:
                e.set_darcy_net(params, args.net_corrections)
            e.load(st)
            out.append(e)
        return out

    steps, warm = args.dens_steps, args.dens_warmup
    eng = engines()
    ens = sphere_amd.Ensemble(eng)
    ens.step(warm)
    t0 = time.perf_counter()
    ens.step(steps)
    t_ens = time.perf_counter() - t0
    einfo = ens.info()
    ens.close()
    one = engines()
    for e in one:
        e.step(warm)
    t0 = time.perf_counter()
    for e in one:
        e.step(steps)
    t_one = time.perf_counter() - t0
    dx = dp = 0.0
    same_sweeps = True
    sweeps = corr = fsteps = 0
    per_box = []
    for k, (a, b) in enumerate(zip(eng, one)):
        n = info[k]["grains"]
        V = int(numpy.prod(info[k]["grid"]))
        ra, rb = a.read(), b.read()
        dx = max(dx, float(numpy.max(numpy.abs(ra.get("x", n, 4) - rb.get("x", n, 4)))))
        dp = max(dp, float(numpy.max(numpy.abs(ra.get("p_f", V, 1) - rb.get("p_f", V, 1)))))
        sa, sb = a.darcy_stats(), b.darcy_stats()
        same_sweeps = same_sweeps and sa.sweeps == sb.sweeps
        sweeps += sa.sweeps
        corr += sa.corrections
        fsteps += sa.fluid_steps
        per_box.append(sa.sweeps / max(1, sa.fluid_steps))
    for e in eng + one:
        e.close()
    v_ens = total * steps / t_ens
    v_one = total * steps / t_one
    log("Darcy ensemble: %.4g particle-steps/s together, %.4g one at a time (%.2fx); max |x| "
        "difference %g, max |p_f| difference %g, sweep counts equal: %s; %.1f sweeps and %.1f "
        "corrections per solve (per box: min %.0f, median %.0f, max %.0f sweeps)"
        % (v_ens, v_one, v_ens / v_one, dx, dp, same_sweeps, sweeps / max(1, fsteps),
           corr / max(1, fsteps), numpy.min(per_box), numpy.median(per_box), numpy.max(per_box)))
    sizes = sorted(p["grains"] for p in info)
    grids = sorted(set(tuple(p["grid"]) for p in info))
    # the CPU oracle runs the plain solve (no network)
    cpu = None if args.no_cpu else ensemble_cpu_leg(args, sims, info, 32, 3)
    return {
        "metric": "aggregate particle-steps/s (independent Darcy-coupled boxes, FP64)",
        "cpu_oracle_two_boxes": cpu,
        "together": v_ens, "one_at_a_time": v_one, "speedup": v_ens / v_one,
        "steps": steps, "warmup": warm,
        "max_abs_position_difference_vs_one_at_a_time": dx,
        "max_abs_pressure_difference_vs_one_at_a_time": dp,
        "sweep_counts_equal": bool(same_sweeps),
        "sweeps_per_solve": sweeps / max(1, fsteps), "corrections_per_solve": corr / max(1, fsteps),
        "sweeps_per_solve_per_box_min_median_max": [float(numpy.min(per_box)),
                                                     float(numpy.median(per_box)),
                                                     float(numpy.max(per_box))],
        "learned_correction": params is not None,
        "correction_autotune": tune,
        "ensemble_info": einfo,
        "config": {"workload": "%d independent Darcy-coupled shear boxes, grains log-uniform in "
                               "[5000, 60000] (radii U[0.009, 0.011] m, linear contacts), mu ~ "
                               "U[0.3, 0.7], sigma0 ~ U[5, 80] kPa, top layer driven at U[1e-3, "
                               "1e-1] m/s; Darcy grid c x c x 2c (c = 16, 24 or 32) per box, "
                               "tolerance 1e-8" % len(info),
                   "boxes": len(info), "grains_total": total,
                   "grains_min_median_max": [sizes[0], sizes[len(sizes) // 2], sizes[-1]],
                   "grids": [list(g) for g in grids]},
    }


def darcy_cpu_leg(args, s):
    """The consolidation box on the CPU oracle: 2 untimed fluid steps, then
    --darcy-cpu-steps timed ones; the Jacobi solve time per fluid step (the
    same span the GPU figure covers) and the step time."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _oracle import Oracle
    threads = host_threads(args)
    ora = Oracle(s.to_state(), threads=threads, nc=32)
    ora.step(2)
    a = ora.darcy_stats()
    t0 = time.perf_counter()
    ora.step(args.darcy_cpu_steps)
    dt = time.perf_counter() - t0
    b = ora.darcy_stats()
    k = args.darcy_cpu_steps
    solve_ms = (b.solve_ms - a.solve_ms) / k
    log("consolidation box on the CPU oracle (%d threads): %.1f ms per solve, %.1f sweeps, %.1f ms "
        "per fluid step" % (threads, solve_ms, (b.sweeps - a.sweeps) / k, 1e3 * dt / k))
    return {"ms_per_solve": solve_ms, "sweeps_per_solve": (b.sweeps - a.sweeps) / float(k),
            "ms_per_step": 1e3 * dt / k, "cores": threads, "kind": "port",
            "sample": "2 untimed + %d timed fluid steps of the FP64 C oracle from the loaded box"
                      % k, "host": host_info(threads)}


def net_flops(nx, ny, nz):
    """Multiply-adds x 2 of one forward pass of the correction network."""
    from sphere_amd import learn
    levels = [0, 1, 2, 3, 3, 2, 1, 0, 0]
    V0 = nx * ny * nz
    return sum(2.0 * taps * ci * co * V0 / 8 ** lv
               for (ci, co, taps), lv in zip(learn.LAYERS, levels))


def size_check(args, params, steps=3):
    """The trained network on a 96 x 96 x 192 grid: a pressure pulse relaxing
    in a grain-free Darcy box (0.96 x 0.96 x 1.92 m, p = 0 at the top, no
    flow at the base, periodic sides), solved to the 1e-8 test with and
    without the learned correction.  Without grains the porosity is 0.9
    in every cell (the porosity kernel recomputes it each fluid step), so
    the consolidation box's fields cannot be carried over; the check is
    that the network, trained at 48 x 48 x 96, runs at eight times the
    voxels and the solve ends at the plain solve's pressure."""
    import sphere_amd
    s = sphere_amd.sim(np=0, nd=3, nw=0, sid="size96", fluid=True)
    s.defaultParams()
    s.defineWorldBoundaries([0.96, 0.96, 1.92], dx=0.01)
    s.initFluid(mu=1e-3, cfd_solver=1)
    s.setFluidCompressibility(1e-8)
    s.setPermeabilityPrefactor(1e-12, verbose=False)
    s.setTolerance(1e-8)
    s.setMaxIterations(200000)
    z = (numpy.arange(192) + 0.5) * 0.01
    s.p_f[:] = 1e3 * numpy.exp(-((z - 0.9) / 0.1) ** 2)[None, None, :]
    s.p_f[:, :, -1] = 0.0
    s.initTemporal(total=1.0, file_dt=1.0, dt=1e-4)
    V = 96 * 96 * 192
    out = {}
    pf = []
    for use in (False, True):
        e = sphere_amd.Engine(0, nc=32)
        if use:
            e.set_darcy_net(params, args.net_corrections)
        e.load(s.to_state())
        st0 = e.darcy_stats()
        e.step(steps)
        st = e.darcy_stats()
        pf.append(e.read().get("p_f", V, 1).copy())
        e.close()
        out["learned" if use else "plain"] = {
            "ms_per_solve": (st.solve_ms - st0.solve_ms) / steps,
            "sweeps_per_solve": (st.sweeps - st0.sweeps) / float(steps),
            "corrections_per_solve": st.corrections / float(steps), "nonconverged": int(st.nonconverged)}
    out["final_pressure_rel_diff_vs_plain"] = float(numpy.max(numpy.abs(pf[1] - pf[0])) /
                                                   max(1e-300, numpy.max(numpy.abs(pf[0]))))
    out["speedup_ms"] = out["plain"]["ms_per_solve"] / max(1e-9, out["learned"]["ms_per_solve"])
    out["workload"] = ("pressure pulse in a grain-free 96 x 96 x 192 Darcy box, %d fluid steps, "
                       "network trained on the 48 x 48 x 96 consolidation boxes" % steps)
    log("size check 96x96x192: plain %.2f ms, %.0f sweeps; learned %.2f ms, %.0f sweeps; pressure "
        "difference %.2e" % (out["plain"]["ms_per_solve"], out["plain"]["sweeps_per_solve"],
                             out["learned"]["ms_per_solve"], out["learned"]["sweeps_per_solve"],
                             out["final_pressure_rel_diff_vs_plain"]))
    return out


def learned_section(args, s, plain_ms, plain_sweeps):
    """Train the correction network on --net-pairs (input, error) pairs from
    the solves of --net-boxes consolidation boxes grown from other seeds
    (held out from the timed run): about 70% from plain sweeps of each
    solve, the rest, in --net-rounds more rounds, from the iterates the
    corrected solve meets.  Then time the timed box's run with it (same
    packing, warm-up and step count as the plain run).  The same network is
    also timed on its first training box (in sample) for comparison."""
    import torch
    import sphere_amd
    from sphere_amd import learn
    from sphere_amd import workloads as W
    t0 = time.time()
    seeds = [args.darcy_seed + 1 + k for k in range(args.net_boxes)]
    boxes = [W.prepare_consolidation(sphere_amd.Engine(0, nc=32), args.darcy_n, sd, 300, 3, 400)
             for sd in seeds]
    log("Darcy network: %d training boxes grown from seeds %s in %.1f s" % (len(seeds), seeds,
                                                                           time.time() - t0))
    t0 = time.time()
    trainers = []
    for sb in boxes:
        tr = sphere_amd.Engine(0, nc=32)
        tr.load(sb.to_state())
        tr.step(args.darcy_warmup)
        trainers.append(tr)
    net = learn.make_net(trainers[0])
    nx, ny, nz = learn.grid_of(trainers[0])
    store = learn.PairStore(args.net_pairs, nx * ny * nz)
    plain_target = int(0.7 * args.net_pairs)
    per_box = (plain_target + len(trainers) - 1) // len(trainers)
    for tr in trainers:  # plain iterates: 7 pairs per fluid step
        target = min(store.n + per_box, plain_target)
        while store.n < target:
            before = store.n
            learn.collect(tr, 1, store=store)
            if store.n == before:
                break
            if store.n > target:
                store.n = target
    X, Y = store.view()
    params, _ = learn.train(net, X, Y, iters=args.net_iters, lr=1e-3)
    rounds = 1
    for r in range(args.net_rounds):
        # the iterates the corrected solve meets (8 pairs per fluid step)
        # and 3 plain ones, round by round over the boxes
        target = plain_target + (args.net_pairs - plain_target) * (r + 1) // args.net_rounds
        k = 0
        while store.n < target:
            before = store.n
            learn.collect(trainers[k % len(trainers)], 1, sweeps=(2, 42, 162), net=net, params=params,
                          corrections=args.net_corrections, store=store)
            k += 1
            if store.n == before and k > 4 * len(trainers):
                break
            if store.n > target:
                store.n = target
        X, Y = store.view()
        params, _ = learn.train(net, X, Y, iters=args.net_iters, lr=5e-4, params=params, seed=r + 1)
        rounds += 1
    pairs = int(store.n)
    del X, Y, store
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    train_s = time.time() - t0
    for tr in trainers:
        tr.close()
    s_train = boxes[0]
    log("Darcy network: %d rounds on %d (input, error) pairs in %.1f s" % (rounds, pairs, train_s))
    # inference time of one forward pass (HIP events on torch's stream)
    nx, ny, nz = int(s.num[0]), int(s.num[1]), int(s.num[2])
    V = nx * ny * nz
    dev = torch.device("cuda:0")
    pt = torch.tensor(params, device=dev)
    inp = torch.randn((V, 3), device=dev)
    o = torch.empty(V, device=dev)
    for _ in range(3):
        net.forward(pt, inp, o)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        net.forward(pt, inp, o)
    e1.record()
    torch.cuda.synchronize()
    fwd_s = e0.elapsed_time(e1) / 20 / 1000.0
    net.close()
    # the timed run with the network (held out), and the same network on its
    # training box (in sample) beside it, with that box's plain solve
    engine = sphere_amd.Engine(0, nc=32)
    engine.set_darcy_net(params, args.net_corrections)
    ms, sweeps, elapsed, st = run_darcy(engine, s, args.darcy_warmup, args.darcy_steps)
    engine.close()
    insample = {}
    for use in (False, True):
        e = sphere_amd.Engine(0, nc=32)
        if use:
            e.set_darcy_net(params, args.net_corrections)
        ims, isw, _, _ = run_darcy(e, s_train, args.darcy_warmup, args.darcy_steps)
        e.close()
        insample["learned" if use else "plain"] = (ims, isw)
    # the converged pressure of one solve from the same state, with and without
    pf = []
    for use in (False, True):
        e = sphere_amd.Engine(0, nc=32)
        if use:
            e.set_darcy_net(params, args.net_corrections)
        e.load(s.to_state())
        e.step(args.darcy_warmup + 1)
        pf.append(e.read().get("p_f", V, 1).copy())
        e.close()
    p_err = float(numpy.max(numpy.abs(pf[1] - pf[0])) / max(1e-300, numpy.max(numpy.abs(pf[0]))))
    size = size_check(args, params)
    flops = net_flops(nx, ny, nz)
    log("Darcy with the network: %.3f ms per solve (p50 %.3f), %.1f sweeps, %.1f corrections per "
        "solve; plain %.3f ms, %.1f sweeps; forward %.1f us; pressure difference %.2e" %
        (ms.mean(), numpy.percentile(ms, 50), sweeps.mean(), st.corrections / float(args.darcy_warmup +
                                                                                     args.darcy_steps),
         plain_ms.mean(), plain_sweeps.mean(), 1e6 * fwd_s, p_err))
    (ip_ms, ip_sw), (il_ms, il_sw) = insample["plain"], insample["learned"]
    log("in sample (the training box): %.3f ms, %.1f sweeps with the network; plain %.3f ms, "
        "%.1f sweeps (%.2fx)" % (il_ms.mean(), il_sw.mean(), ip_ms.mean(), ip_sw.mean(),
                                 ip_ms.mean() / il_ms.mean()))
    return params, {
        "held_out": "trained on boxes grown from seeds %s, timed on the seed-%d box"
                    % (seeds, args.darcy_seed),
        "ms": percentiles(ms), "sweeps": percentiles(sweeps),
        "in_sample": {"ms": percentiles(il_ms), "sweeps": percentiles(il_sw),
                      "plain_ms": percentiles(ip_ms), "plain_sweeps": percentiles(ip_sw),
                      "speedup_ms": float(ip_ms.mean() / il_ms.mean()),
                      "sweep_ratio_plain_over_learned": float(ip_sw.mean() / il_sw.mean())},
        "corrections_per_solve": st.corrections / float(args.darcy_warmup + args.darcy_steps),
        "nonconverged": int(st.nonconverged),
        "sweep_ratio_plain_over_learned": float(plain_sweeps.mean() / sweeps.mean()),
        "speedup_ms": float(plain_ms.mean() / ms.mean()),
        "final_pressure_rel_diff_vs_plain": p_err,
        "max_corrections": args.net_corrections,
        "size_check_96x96x192": size,
        "training": {"pairs": pairs, "boxes": len(seeds), "rounds": rounds,
                     "iters_per_round": args.net_iters,
                     "seconds": train_s, "params": int(sphere_amd.Net.param_count())},
        "net_roofline": {"bound": "mfma", "achieved": flops / fwd_s / 1e12, "peak": 2500.0,
                         "unit": "TFLOP/s", "frac": flops / fwd_s / 1e12 / 2500.0,
                         "flops_per_forward": flops, "us_per_forward": 1e6 * fwd_s,
                         "dtype": "bf16 operands, fp32 accumulation"},
    }


def ensemble_section(args):
    """Independent shear boxes stepped together (Ensemble) against the same
    boxes stepped one at a time (sph_step per box), from the same states."""
    import sphere_amd
    from sphere_amd import workloads as W
    t0 = time.time()
    sims, params = W.ensemble_boxes(count=args.ens_boxes, seed=args.ens_seed)
    total = sum(p["grains"] for p in params)
    log("ensemble: %d boxes, %d grains in all, prepared in %.1f s" % (len(sims), total, time.time() - t0))
    states = [s.to_state() for s in sims]

    def engines():
        out = []
        for st in states:
            e = sphere_amd.Engine(0, nc=64)
            e.load(st)
            out.append(e)
        return out

    # together
    eng = engines()
    ens = sphere_amd.Ensemble(eng)
    ens.step(args.ens_warmup)
    t0 = time.perf_counter()
    ens.step(args.ens_steps)
    t_ens = time.perf_counter() - t0
    ens.close()
    # one at a time
    one = engines()
    for e in one:
        e.step(args.ens_warmup)
    t0 = time.perf_counter()
    for e in one:
        e.step(args.ens_steps)
    t_one = time.perf_counter() - t0
    # per-box results
    diff = 0.0
    for k, (a, b) in enumerate(zip(eng, one)):
        n = params[k]["grains"]
        xa = a.read().get("x", n, 4)
        xb = b.read().get("x", n, 4)
        diff = max(diff, float(numpy.max(numpy.abs(xa - xb))))
    for e in eng + one:
        e.close()
    v_ens = total * args.ens_steps / t_ens
    v_one = total * args.ens_steps / t_one
    log("ensemble: %.4g particle-steps/s together, %.4g one at a time (%.2fx); max |x| difference %g" %
        (v_ens, v_one, v_ens / v_one, diff))
    sizes = sorted(p["grains"] for p in params)
    cpu = None if args.no_cpu else ensemble_cpu_leg(args, sims, params, 64, 10)
    return {
        "cpu_oracle_two_boxes": cpu,
        "metric": "aggregate particle-steps/s (independent boxes, FP64)",
        "together": v_ens, "one_at_a_time": v_one, "speedup": v_ens / v_one,
        "steps": args.ens_steps, "warmup": args.ens_warmup,
        "max_abs_position_difference_vs_one_at_a_time": diff,
        "config": {"workload": "%d independent periodic shear boxes, grains log-uniform in [5000, 60000] "
                               "(lognormal radii, Hertz-Mindlin E=7e9 nu=0.3), mu ~ U[0.3, 0.7], "
                               "sigma0 ~ U[5, 80] kPa, top-layer velocity ~ U[1e-3, 1e-1] m/s; "
                               "DEM only" % len(params),
                   "boxes": len(params), "grains_total": total,
                   "grains_min_median_max": [sizes[0], sizes[len(sizes) // 2], sizes[-1]]},
    }


def ensemble_cpu_leg(args, sims, params, nc, steps):
    """Two representative boxes of an ensemble (the median-sized and the
    largest) on the CPU oracle: 2 untimed + `steps` timed steps each, with
    all the job's host cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _oracle import Oracle
    threads = host_threads(args)
    order = sorted(range(len(params)), key=lambda k: params[k]["grains"])
    out = []
    for k in (order[len(order) // 2], order[-1]):
        ora = Oracle(sims[k].to_state(), threads=threads, nc=nc)
        ora.step(2)
        t0 = time.perf_counter()
        ora.step(steps)
        dt = time.perf_counter() - t0
        rec = {"box": k, "grains": params[k]["grains"],
               "particle_steps_per_s": params[k]["grains"] * steps / dt, "ms_per_step": 1e3 * dt / steps}
        if "grid" in params[k]:
            rec["grid"] = params[k]["grid"]
        out.append(rec)
        log("ensemble box %d (%d grains) on the CPU oracle (%d threads): %.4g particle-steps/s" %
            (k, params[k]["grains"], threads, rec["particle_steps_per_s"]))
    return {"boxes": out, "cores": threads, "kind": "port",
            "sample": "2 untimed + %d timed steps of the FP64 C oracle per box, from the prepared state"
                      % steps, "host": host_info(threads)}


def host_threads(args):
    """Threads of the CPU baseline: all the host cores this job may use (the
    OpenMP thread count the job is given, else every core)."""
    if args.cpu_threads:
        return args.cpu_threads
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def host_info(threads):
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "threads_used": threads,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS", "")}


def settling_leg(args, threads):
    """Config 1 on the GPU and on the CPU oracle: dry settling of 10,000
    grains, all 2,000 steps on each (and the first 200 on one CPU thread)."""
    import sphere_amd
    from sphere_amd import workloads as W
    s = W.settling(n=10000, seed=1, steps=2000)
    st = s.to_state()
    e = sphere_amd.Engine(0)
    e.load(st)
    e.step(1)
    e.load(st)
    t0 = time.perf_counter()
    e.step(2000)
    gpu_s = time.perf_counter() - t0
    e.close()
    cpu_v, cpu_s = cpu_baseline(st, 10000, 0, 2000, threads, nc=32)
    one_v, one_s = cpu_baseline(st, 10000, 0, 200, 1, nc=32)
    log("settling 10k x 2000 steps: GPU %.3f s (%.3g particle-steps/s), CPU oracle %.2f s on %d "
        "threads (%.3g), one thread %.3g" % (gpu_s, 2e7 / gpu_s, cpu_s, threads, cpu_v, one_v))
    return {"workload": "dry settling: 10000 grains, radii U[0.009, 0.011] m, jittered lattice, "
                        "linear contacts k_n=k_t=1.16e7, mu=0.5, frictional walls, 2000 steps",
            "gpu_particle_steps_per_s": 2e7 / gpu_s, "gpu_seconds": gpu_s,
            "cpu_particle_steps_per_s": cpu_v, "cpu_seconds": cpu_s, "cpu_threads": threads,
            "cpu_single_core_particle_steps_per_s": one_v,
            "cpu_single_core_sample": "the first 200 steps (%.1f s)" % one_s}


def cpu_baseline(state_buf, n, warm, steps, threads, nc=64):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _oracle import Oracle
    ora = Oracle(state_buf, threads=threads, nc=nc)
    ora.step(warm)
    t0 = time.perf_counter()
    ora.step(steps)
    dt = time.perf_counter() - t0
    return n * steps / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--prof-steps", type=int, default=200,
                    help="steps after the timed ones with per-kernel HIP-event timing")
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--stage-steps", type=int, default=400)
    ap.add_argument("--z-target", type=float, default=4.2)
    ap.add_argument("--relax-rounds", type=int, default=4)
    ap.add_argument("--relax-steps", type=int, default=400)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps-1", type=int, default=2,
                    help="timed steps of the one-thread CPU baseline")
    ap.add_argument("--no-settling", action="store_true",
                    help="skip the 10,000-grain settling leg of the CPU baseline")
    ap.add_argument("--darcy-cpu-steps", type=int, default=5,
                    help="fluid steps of the consolidation box on the CPU oracle")
    ap.add_argument("--skin", type=float, default=None,
                    help="neighbour-list skin as a fraction of the largest radius (0: cell search)")
    ap.add_argument("--separate-rebuild", action="store_true",
                    help="list rebuilds as separate sort, reorder and list launches")
    ap.add_argument("--save-packing", default="",
                    help="write the prepared shear box to this folder (<sid>.bin)")
    ap.add_argument("--load-packing", default="",
                    help="start from a shear box written by --save-packing")
    ap.add_argument("--prep-only", action="store_true", help="stop after preparing")
    ap.add_argument("--no-darcy", action="store_true", help="skip the consolidation box")
    ap.add_argument("--no-shear", action="store_true",
                    help="skip the shear box (the line then carries only the Darcy object)")
    ap.add_argument("--darcy-n", type=int, default=200_000)
    ap.add_argument("--darcy-seed", type=int, default=3)
    ap.add_argument("--darcy-steps", type=int, default=500)
    ap.add_argument("--darcy-warmup", type=int, default=20)
    ap.add_argument("--no-net", action="store_true", help="skip the learned Darcy correction")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip both ensemble sections (DEM boxes; Darcy-coupled boxes)")
    ap.add_argument("--dens-steps", type=int, default=40, help="timed steps of the Darcy ensemble")
    ap.add_argument("--dens-warmup", type=int, default=10)
    ap.add_argument("--dens-tune-steps", type=int, default=3,
                    help="trial fluid steps per box of the learned-correction autotune (0: the "
                         "network on every box)")
    ap.add_argument("--ensemble", action="store_true",
                    help="with --no-shear --no-darcy: only the ensemble of DEM boxes")
    ap.add_argument("--ens-boxes", type=int, default=64)
    ap.add_argument("--ens-seed", type=int, default=5)
    ap.add_argument("--ens-steps", type=int, default=200)
    ap.add_argument("--ens-warmup", type=int, default=20)
    ap.add_argument("--net-iters", type=int, default=10000, help="Adam steps per training round")
    ap.add_argument("--net-pairs", type=int, default=20000,
                    help="(input, error) training pairs in all")
    ap.add_argument("--net-boxes", type=int, default=3,
                    help="consolidation boxes (other seeds) the pairs come from")
    ap.add_argument("--net-rounds", type=int, default=2,
                    help="training rounds on iterates of the corrected solve")
    ap.add_argument("--net-corrections", type=int, default=8, help="corrections per solve at most")
    ap.add_argument("--darcy-save", default="", help="write the prepared consolidation box here")
    ap.add_argument("--darcy-load", default="", help="start from a box written by --darcy-save")
    ap.add_argument("--no-ns", action="store_true", help="skip the Navier-Stokes fluid section")
    ap.add_argument("--ns-steps", type=int, default=20, help="timed Navier-Stokes steps")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="finish with three device-to-device copies of MIB MiB, the "
                         "known-size kernel that scales HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the DEM step and the Darcy solve run on one "
              "device (multi-GPU runs are not part of this build); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    if args.no_shear and not args.no_darcy and not args.no_net:
        # torch carries its own HIP runtime: initialise it before the library
        # loads, so that both use that one runtime in this process.  (With
        # the shear box, the Darcy section runs in a child process, so that
        # the DEM step is timed on the system HIP runtime the library is
        # built against, without torch.)
        import torch
        torch.cuda.init()
    import sphere_amd
    from sphere_amd import workloads as W
    if args.no_shear and args.ensemble and args.no_darcy:
        d = ensemble_section(args)
        print(json.dumps({"metric": d["metric"], "value": d["together"], "unit": "particle-steps/s",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
                          "data": "synthetic", "ensemble": d}), flush=True)
        return
    if args.no_shear and args.no_darcy:
        d = navier_stokes_section(args)
        print(json.dumps({"metric": "ms per Navier-Stokes fluid step (48 x 48 x 96, no grains)",
                          "value": d["ms_per_step"], "unit": "ms", "n_gpus": 1,
                          "higher_is_better": False, "dtype": "f64", "data": "synthetic",
                          "navier_stokes": d}), flush=True)
        return
    if args.no_shear:
        if args.prep_only:
            darcy_packing(args, sphere_amd.Engine(0, nc=32))
            return
        d = darcy_section(args)
        print(json.dumps({"metric": d["metric"], "value": d["plain"]["ms"]["mean"], "unit": "ms",
                          "n_gpus": 1, "higher_is_better": False, "dtype": "f64",
                          "data": "synthetic", "darcy": d}), flush=True)
        if args.copy_calibration:
            copy_calibration(args.copy_calibration)
        return
    # the largest lognormal grains touch more than 32 others: 64 contact slots
    engine = sphere_amd.Engine(0, nc=64)
    run_engine = engine
    if args.skin is not None or args.separate_rebuild:
        run_engine = sphere_amd.Engine(0, nc=64, skin=args.skin,
                                       fused_rebuild=not args.separate_rebuild)
    if args.load_packing:
        s = sphere_amd.sim()
        s.readbin(args.load_packing, verbose=False)
        if int(s.np[0]) != args.n:
            raise SystemExit("%s holds %d grains, not --n %d" % (args.load_packing, s.np[0], args.n))
        log("loaded the shear box from %s" % args.load_packing)
    else:
        log("preparing the shear box (%d grains)" % args.n)
        s = W.prepare_shear_box(engine, args.n, args.seed, args.stage_steps, args.z_target,
                                args.relax_rounds, args.relax_steps, log=log)
    if args.save_packing:
        os.makedirs(args.save_packing, exist_ok=True)
        s.writebin(folder=args.save_packing, verbose=False)
        log("wrote %s/%s.bin" % (args.save_packing, s.sid))
    if args.prep_only:
        if not args.no_darcy:
            darcy_packing(args, sphere_amd.Engine(0, nc=32))
        return
    state = s.to_state()
    run_engine.load(state)
    log("warm-up %d steps" % args.warmup)
    run_engine.step(args.warmup)
    warm_buf = run_engine.read()
    n0 = run_engine.count_contacts()
    log("timed %d steps" % args.steps)
    t0 = time.perf_counter()
    run_engine.step(args.steps)
    elapsed = time.perf_counter() - t0
    n1 = run_engine.count_contacts()
    # per-kernel device times (HIP events between the launches) in a run of
    # their own after the timed one, so that the events do not slow it
    run_engine.set_profiling(1)
    log("profiled %d steps" % args.prof_steps)
    run_engine.step(args.prof_steps)
    ktimes = {}
    for k, name in enumerate(sphere_amd._abi.K_NAMES):
        ms, cnt = run_engine.kernel_time(k)
        if cnt:
            ktimes[name] = {"ms_total": round(ms, 3), "launches": int(cnt),
                            "us_avg": round(1000.0 * ms / cnt, 2)}
    run_engine.set_profiling(0)
    nls = run_engine.neighbor_stats()
    lists = None
    if nls.enabled:
        lists = {"skin_m": nls.skin, "rebuilds_total": int(nls.rebuilds),
                 "full_search_steps_total": int(nls.full_steps),
                 "entries_per_grain": round(nls.entries / float(args.n), 2),
                 "longest": int(nls.longest)}
        log("neighbour lists: skin %.3g m, %d rebuilds, %d full-search steps, %.2f entries per "
            "grain, longest %d" % (nls.skin, nls.rebuilds, nls.full_steps,
                                   nls.entries / float(args.n), nls.longest))
    n = args.n
    value = n * args.steps / elapsed
    zbar = 0.5 * (n0 + n1) / n
    # the contact stage is two launches per step: topology (search and slot
    # bookkeeping) and contact (laws, walls, sums); its duration is the sum of
    # their average launch durations
    c_avg_s = sum(ktimes[k]["ms_total"] / ktimes[k]["launches"] / 1000.0
                  for k in ("topology", "contact") if k in ktimes)
    c_bytes = contact_bytes_per_particle(zbar) * n
    achieved = c_bytes / c_avg_s / 1e9
    log("elapsed %.3f s, %.1f us/step, Zbar %.2f, contact kernel %.1f us" %
        (elapsed, 1e6 * elapsed / args.steps, zbar, 1e6 * c_avg_s))

    cpu = None
    if not args.no_cpu:
        threads = host_threads(args)
        log("CPU oracle baseline: %d threads, 2 + %d steps; 1 thread, 1 + %d steps" %
            (threads, args.cpu_steps, args.cpu_steps_1))
        cv, cdt = cpu_baseline(warm_buf, n, 2, args.cpu_steps, threads)
        cv1, cdt1 = cpu_baseline(warm_buf, n, 1, args.cpu_steps_1, 1)
        cpu = {"value": cv, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": "the same 1M-grain shear box after the GPU warm-up: 2 untimed + %d timed "
                         "steps of the FP64 C oracle (OpenMP over %d threads, %.1f s)"
                         % (args.cpu_steps, threads, cdt),
               "single_core": {"value": cv1, "cores": 1,
                               "sample": "1 untimed + %d timed steps on one thread (%.1f s)"
                                         % (args.cpu_steps_1, cdt1)},
               "host": host_info(threads)}
        if not args.no_settling:
            cpu["settling_10k"] = settling_leg(args, threads)

    line = {
        "metric": "particle-steps/s (1M grains, FP64); Darcy ms-to-1e-8, CNN-precond vs plain",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "dense periodic shear box: %d grains, lognormal radii (median 0.01 m, "
                               "sigma_ln 0.2), Hertz-Mindlin E=7e9 nu=0.3 mu=0.5, 10 kPa top wall, "
                               "top layer at 0.01 m/s, 5%% bonded pairs" % n,
                   "grains": n, "bonds": int(s.nb0[0]), "contacts_per_grain": round(zbar, 3),
                   "grid": [int(v) for v in s.num], "dt": float(s.time_dt[0]),
                   "neighbour_lists": lists},
        "roofline": {"bound": "hbm", "kernel": "contact stage (k_topology_lists + k_forces)",
                     "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(),
                     "traffic_source": "profiles/shear_box_1m_hbm_traffic.md",
                     "bytes_per_step": c_bytes, "us_per_step": 1e6 * c_avg_s,
                     "measured_copy_gbs": copy_ceiling(),
                     "measured_copy_note": "device-to-device hipMemcpy of 1 GiB, read + write bytes "
                                           "per second, median of 10: the stream-copy ceiling of "
                                           "this GPU"},
        "kernels": ktimes,
        "cpu_baseline": cpu,
    }
    if run_engine is not engine:
        run_engine.close()
    engine.close()
    line["darcy"] = None if args.no_darcy else darcy_child()
    line["ensemble"] = None if args.no_ensemble else ensemble_section(args)
    line["navier_stokes"] = None if args.no_ns else navier_stokes_section(args)
    print(json.dumps(line), flush=True)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)


def navier_stokes_section(args):
    """The Navier-Stokes fluid step (cfd_solver = 0) without grains on the
    consolidation box's 48 x 48 x 96 grid: a water column over a closed
    bottom under a fixed-pressure top, from a hydrostatic pressure and a
    random velocity perturbation; --ns-steps timed steps after 2 untimed
    ones, and the CPU oracle (one thread) on the first two steps."""
    import sphere_amd
    from sphere_amd import _abi
    shape = (48, 48, 96)
    L = (1.1088, 1.1088, 2.2176)
    dt = 1.0e-4
    P = _abi.ns_params(*shape, L, dt, mu=8.9e-4, rho_f=1000.0, g=(0.0, 0.0, -9.81), bc_bot=1,
                       bc_top=0, tolerance=1.0e-8, maxiter=100000)
    rng = numpy.random.default_rng(11)
    v = 0.01 * rng.standard_normal(shape + (3,))
    dz = L[2] / shape[2]
    p = numpy.broadcast_to(1000.0 * 9.81 * dz * (shape[2] - 1 - numpy.arange(shape[2])),
                           shape).copy()
    ns = sphere_amd.NavierStokes(P, v, p)
    ns.step(2)
    a = ns.stats()
    t0 = time.perf_counter()
    ns.step(args.ns_steps)
    el = time.perf_counter() - t0
    b = ns.stats()
    ns.close()
    iters = (b.iterations - a.iterations) / float(args.ns_steps)
    cells = shape[0] * shape[1] * shape[2]
    out = {"workload": "Navier-Stokes water column without grains, %d x %d x %d cells (the "
                       "consolidation box's grid), closed bottom, fixed-pressure top, hydrostatic "
                       "start with a 0.01 m/s random velocity perturbation, dt 1e-4 s, Jacobi "
                       "tolerance 1e-8" % shape,
           "steps": args.ns_steps, "ms_per_step": 1000.0 * el / args.ns_steps,
           "jacobi_iterations_per_solve": iters,
           "us_per_jacobi_iteration": 1e6 * el / max(1.0, b.iterations - a.iterations),
           "nonconverged_timed": int(b.nonconverged - a.nonconverged)}
    log("Navier-Stokes 48x48x96 without grains: %.3f ms per step, %.0f Jacobi iterations per "
        "solve, %.2f us per iteration" % (out["ms_per_step"], iters, out["us_per_jacobi_iteration"]))
    if not args.no_cpu:
        # the oracle's cost per Jacobi iteration: the first step of the same
        # column with at most 300 iterations (one thread)
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from _ns_oracle import NsOracle
        Pc = _abi.ns_params(*shape, L, dt, mu=8.9e-4, rho_f=1000.0, g=(0.0, 0.0, -9.81), bc_bot=1,
                            bc_top=0, tolerance=1.0e-8, maxiter=300)
        o = NsOracle(Pc, v, p)
        t0 = time.perf_counter()
        o.step(1)
        cel = time.perf_counter() - t0
        st = o.stats()
        o.close()
        us_it = 1e6 * cel / st.iterations
        out["cpu_oracle"] = {"us_per_jacobi_iteration": us_it, "cores": 1, "kind": "port",
                             "ms_per_step_at_gpu_iterations": us_it * iters / 1000.0,
                             "sample": "the first step of the same column on the FP64 C oracle "
                                       "(one thread), stopped at 300 Jacobi iterations (%.1f s)"
                                       % cel}
        log("Navier-Stokes on the CPU oracle: %.0f us per Jacobi iteration (one thread), %.0f ms "
            "per step at the GPU's iteration count" % (us_it, us_it * iters / 1000.0))
    out["cells"] = cells
    out["coupled"] = navier_stokes_coupled(args)
    return out


def navier_stokes_coupled(args):
    """The fluid attached to a DEM context (sph_ns_attach): a settled packing
    of 20,000 grains (the consolidation box's recipe with 16 growth cells
    across, periodic in x and y, a wall on top) in water on the packing's
    grid, each DEM step running the fluid iteration and adding its force on
    the grains; --ns-steps timed steps after 2 untimed ones."""
    import sphere_amd
    from sphere_amd import _abi
    from sphere_amd import workloads as W
    n = 20_000
    s, f0, _ = W.consolidation_base(n=n, seed=4, cells_xy=16)
    e = sphere_amd.Engine(0, nc=32)
    st = W.EngineStepper(e)
    W.grow_to_jamming(s, st, f0, z_target=4.2, steps=300, fine_df=0.005)
    s.w_x[0] = s.L[2]
    s.g[2] = -9.81
    W.relax_packing(s, st, 2, 300)
    e.close()
    nx, ny, nz = (int(v) for v in s.num)
    P = _abi.ns_params(nx, ny, nz, s.L, float(s.time_dt[0]), mu=8.9e-4, rho_f=1000.0, g=s.g, bc_bot=1,
                       bc_top=0, tolerance=1.0e-8, maxiter=100000, t0=float(s.time_current[0]))
    dz = s.L[2] / nz
    p = numpy.broadcast_to(1000.0 * 9.81 * dz * (nz - 1 - numpy.arange(nz)), (nx, ny, nz)).copy()
    e = sphere_amd.Engine(0, nc=32)
    e.load(s.to_state())
    ns = sphere_amd.NavierStokes(P, numpy.zeros((nx, ny, nz, 3)), p)
    e.attach_fluid(ns)
    e.step(2)
    a = ns.stats()
    t0 = time.perf_counter()
    e.step(args.ns_steps)
    el = time.perf_counter() - t0
    b = ns.stats()
    phi = ns.read()[2]
    e.close()
    ns.close()
    out = {"workload": "%d settled grains (radii U[0.009, 0.011] m, periodic in x and y, a wall "
                       "on top) in water on their %d x %d x %d grid, the Navier-Stokes fluid "
                       "attached (its force moves the grains), dt %.3g s, Jacobi tolerance 1e-8"
                       % (n, nx, ny, nz, float(s.time_dt[0])),
           "grains": n, "grid": [nx, ny, nz], "steps": args.ns_steps,
           "ms_per_step": 1000.0 * el / args.ns_steps,
           "particle_steps_per_s": n * args.ns_steps / el,
           "jacobi_iterations_per_step": (b.iterations - a.iterations) / float(args.ns_steps),
           "porosity_min_mean": [float(phi.min()), float(phi.mean())]}
    log("Navier-Stokes with %d grains moving: %.3f ms per step (%.3g particle-steps/s), %.0f Jacobi "
        "iterations per step" % (n, out["ms_per_step"], out["particle_steps_per_s"],
                                 out["jacobi_iterations_per_step"]))
    return out


def darcy_child():
    """The Darcy section (plain and learned solves) in a child process with
    torch, which trains the correction network; returns its record."""
    import subprocess
    argv = []
    skip = False
    for a in sys.argv[1:]:
        if skip:
            skip = False
            continue
        if a == "--copy-calibration":
            skip = True
            continue
        if a.startswith("--copy-calibration="):
            continue
        argv.append(a)
    cmd = [sys.executable, "-u", os.path.abspath(__file__)] + argv + ["--no-shear"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, universal_newlines=True)
    if r.returncode != 0:
        raise SystemExit("the Darcy section failed (exit status %d)" % r.returncode)
    return json.loads(r.stdout.strip().splitlines()[-1])["darcy"]


def copy_ceiling(mib=1024, reps=10):
    """Read + write bytes per second of a device-to-device hipMemcpy of
    `mib` MiB (the median of `reps` copies timed with HIP events)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib * (1 << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or \
            hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    hip.hipEventCreate(ctypes.byref(e0))
    hip.hipEventCreate(ctypes.byref(e1))
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    rates = []
    for _ in range(reps):
        hip.hipEventRecord(e0, None)
        hip.hipMemcpyAsync(b, a, ctypes.c_size_t(n), 3, None)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float(0.0)
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        if ms.value > 0:
            rates.append(2.0 * n / (ms.value * 1e-3) / 1e9)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    gbs = float(numpy.median(rates)) if rates else None
    if gbs:
        log("stream-copy ceiling: %.0f GB/s (device-to-device copy of %d MiB)" % (gbs, mib))
    return gbs


def copy_calibration(mib):
    """hipMemcpy device-to-device of a known size (the runtime's plain copy
    kernel), three times: its FETCH_SIZE / WRITE_SIZE counters give the
    scale of the HBM traffic counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib * (1 << 20)
    a = ctypes.c_void_p()
    b = ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or \
            hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log("copy calibration: 3 device-to-device copies of %d B (read %d B + write %d B each)" %
        (n, n, n))

if __name__ == "__main__":
    main()
