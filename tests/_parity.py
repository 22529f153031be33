"""Shared checks of the GPU parity tests: the library's state, sort arrays
and contact rows against the oracle's.

Bar: cell IDs, sort permutation, cellStart/cellEnd and contact index lists
bit-exact; positions, velocities, forces and the rest of the FP64 state
within a normwise relative difference of 1e-10 per step (the kernels keep the
reference's operation order without FMA contraction, so the observed
difference is zero)."""
import numpy

from _oracle import Oracle

TOL_PER_STEP = 1e-10

PARTICLE_FIELDS = [("x", 4), ("vel", 4), ("angvel", 3), ("force", 3), ("torque", 3),
                   ("angpos", 3), ("xyzsum", 3), ("es_dot", 1), ("es", 1), ("ev_dot", 1),
                   ("ev", 1), ("p", 1)]


def rel_diff(a, b):
    a = numpy.asarray(a, dtype=numpy.float64)
    b = numpy.asarray(b, dtype=numpy.float64)
    scale = numpy.max(numpy.abs(b)) if b.size else 0.0
    if scale == 0.0:
        return float(numpy.max(numpy.abs(a))) if a.size else 0.0
    return float(numpy.max(numpy.abs(a - b)) / scale)


def compare_state(g, o, n, steps, fluid=False, label=""):
    report = {}
    for name, w in PARTICLE_FIELDS:
        report[name] = rel_diff(g.get(name, n, w), o.get(name, n, w))
    if fluid:
        h = o.state
        cells = h.num[0] * h.num[1] * h.num[2]
        for name, w in (("p_f", 1), ("phi", 1), ("dphi", 1), ("v_f", 3)):
            report[name] = rel_diff(g.get(name, cells, w), o.get(name, cells, w))
        report["f_p"] = rel_diff(g.get("f_p", n, 3), o.get("f_p", n, 3))
    worst = max(report.values()) if report else 0.0
    assert worst <= TOL_PER_STEP * steps, "%s: state differs: %r" % (label, report)
    for k in range(g.state.nw):
        for c in range(4):
            assert abs(g.state.w_nx[4 * k + c] - o.state.w_nx[4 * k + c]) <= \
                TOL_PER_STEP * steps * max(1.0, abs(o.state.w_nx[4 * k + c]))
            assert abs(g.state.w_mvfd[4 * k + c] - o.state.w_mvfd[4 * k + c]) <= \
                TOL_PER_STEP * steps * max(1.0, abs(o.state.w_mvfd[4 * k + c]))
    return report


def compare_sort(engine, ora):
    cg, ig, sg, eg = engine.read_sort()
    co, io, so, eo = ora.read_sort()
    assert numpy.array_equal(cg, co), "cell IDs differ"
    assert numpy.array_equal(ig, io), "sort permutation differs"
    assert numpy.array_equal(sg, so), "cellStart differs"
    nonempty = so != 0xffffffff
    assert numpy.array_equal(eg[nonempty], eo[nonempty]), "cellEnd differs"


def compare_contacts(engine, ora, steps):
    cg, dg, tg = engine.read_contacts()
    co, do, to = ora.read_contacts()
    assert numpy.array_equal(cg, co), "contact index lists differ"
    occ = co != ora.header().np
    assert numpy.array_equal(dg[occ], do[occ]), "distmod differs"
    assert rel_diff(tg, to) <= TOL_PER_STEP * steps, "delta_t differs"
    return int(numpy.count_nonzero(occ))


def run_pair(engine, s, chunks, per_chunk, fluid=False, label="", check_contacts=True):
    engine.load(s.to_state())
    ora = Oracle(s.to_state(), nc=engine.nc)
    # in list mode the cell sort is only current on rebuild steps; without
    # grains no sort runs (the reference leaves the arrays unset too)
    check_sort = (fluid or not engine.neighbor_stats().enabled) and int(s.np[0]) > 0
    done = 0
    ncont = 0
    for _ in range(chunks):
        engine.step(per_chunk)
        ora.step(per_chunk)
        done += per_chunk
        if check_sort:
            compare_sort(engine, ora)
        if check_contacts:
            ncont = compare_contacts(engine, ora, done)
        compare_state(engine.read(), ora.read(), int(s.np[0]), done, fluid=fluid, label=label)
    return ora, ncont


