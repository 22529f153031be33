"""The oracle's cell sort, contact search and slot bookkeeping against what
the mathematics fixes: a stable sort by (cell, original index), the set of
overlapping pairs found by brute force over all periodic images, and slot
sequences computed by hand from the rule of src/contactsearch.cuh:263-299.
CPU only."""
import numpy
import pytest

import sphere_amd
from _cases import dense_case, lattice_positions
from _oracle import Oracle


def free_case(x, r, L, num, periodic=0, model=2):
    """Grains at rest, no gravity, no walls: one step moves nothing that
    matters for the sort (forces are computed at the step's start)."""
    n = len(r)
    s = sphere_amd.sim(np=n, sid="free", nw=0)
    s.x[:] = x
    s.radius[:] = r
    s.defaultParams(mu_s=0.5, mu_d=0.5, k_n=1e6, k_t=1e6)
    s.contactmodel[0] = model
    s.g[:] = 0.0
    s.periodic[0] = periodic
    s.origo[:] = 0.0
    s.L[:] = L
    s.num[:] = num
    s.initTemporal(total=1.0, file_dt=1.0, dt=1e-6)
    return s


def expected_sort(x, L, num):
    cell_len = numpy.asarray(L) / numpy.asarray(num)
    ijk = numpy.floor(x / cell_len).astype(numpy.int64)
    cid = (ijk[:, 2] * num[1]) * num[0] + ijk[:, 1] * num[0] + ijk[:, 0]
    order = numpy.argsort(cid, kind="stable")
    ncell = int(numpy.prod(num))
    start = numpy.full(ncell, 0xffffffff, dtype=numpy.uint32)
    end = numpy.full(ncell, 0xffffffff, dtype=numpy.uint32)
    sc = cid[order]
    for i, c in enumerate(sc):
        if i == 0 or c != sc[i - 1]:
            start[c] = i
        if i == len(sc) - 1 or sc[i + 1] != c:
            end[c] = i + 1
    return sc.astype(numpy.uint32), order.astype(numpy.uint32), start, end


@pytest.mark.parametrize("seed", [0, 1])
def test_sort_random_and_faces(seed):
    """Cell IDs, the stable permutation and cellStart/cellEnd; a third of the
    grains sit exactly on cell faces (x = k * L/num)."""
    rng = numpy.random.default_rng(seed)
    L = numpy.array([1.0, 0.75, 0.5])
    num = numpy.array([8, 6, 4])
    n = 600
    x = rng.uniform(0.0, 1.0, (n, 3)) * L * 0.999
    face = rng.random(n) < 0.33
    k = rng.integers(0, num, size=(n, 3))
    x[face] = (k[face] * (L / num))
    r = numpy.full(n, 1e-4)
    s = free_case(x, r, L, num)
    o = Oracle(s.to_state())
    o.step(1)
    cid, idx, st, en = o.read_sort()
    ecid, eidx, est, een = expected_sort(x, L, num)
    assert numpy.array_equal(cid, ecid)
    assert numpy.array_equal(idx, eidx)
    assert numpy.array_equal(st, est)
    ne = est != 0xffffffff
    assert numpy.array_equal(en[ne], een[ne])


def brute_force_pairs(x, r, L, periodic):
    """Overlapping pairs over all periodic images (x and y for periodic 1, x
    for periodic 2): {(a, b)} with b != a."""
    n = len(r)
    shifts = [numpy.zeros(3)]
    if periodic in (1, 2):
        shifts += [numpy.array([L[0], 0, 0]), numpy.array([-L[0], 0, 0])]
    if periodic == 1:
        base = list(shifts)
        for sy in (L[1], -L[1]):
            shifts += [b + numpy.array([0, sy, 0]) for b in base]
    pairs = set()
    for a in range(n):
        for sh in shifts:
            d = x[a] - (x + sh)
            dist = numpy.sqrt(numpy.sum(d * d, axis=1))
            hit = numpy.nonzero(dist - (r[a] + r) < 0.0)[0]
            for b in hit:
                if b != a:
                    pairs.add((a, int(b)))
    return pairs


@pytest.mark.parametrize("periodic", [0, 1, 2])
def test_contact_set_matches_brute_force(periodic):
    """After one step every row holds exactly the grains overlapping it at the
    step's start (cells are wider than a diameter, so the 27-cell search
    misses nothing)."""
    s = dense_case(n=343, seed=5 + periodic, model=2, periodic=periodic, vmag=0.0, wmag=0.0,
                   gravity=False, spacing_factor=0.99)
    x0 = s.x.copy()
    n = int(s.np[0])
    o = Oracle(s.to_state())
    o.step(1)
    rows = o.read_contacts()[0].reshape(n, 32)
    got = {(a, int(b)) for a in range(n) for b in rows[a] if b != n}
    exp = brute_force_pairs(x0, s.radius, s.L, periodic)
    assert len(exp) > 300
    assert got == exp


def slot_case(extra):
    """Grain 0 in the middle cell of a 3x1x1 row of cells, its partners given
    as (cell offset along x, x position) pairs."""
    r = 0.01
    L = numpy.array([0.09, 0.03, 0.03])
    num = numpy.array([3, 1, 1])
    x = [[0.045, 0.015, 0.015]]
    for px in extra:
        x.append([px, 0.015, 0.015])
    return free_case(numpy.array(x), numpy.full(len(x), r), L, num)


def row_of(o, n, nc=32):
    return o.read_contacts()[0].reshape(n, nc)[0]


def test_slots_discovery_order():
    """Partners of grain 0 in cells x-1 (grain 1), x (grain 3) and x+1
    (grain 2) are met in that order and take slots 31, 30, 29."""
    s = slot_case([0.029, 0.061, 0.0451])  # grain 1 in cell 0, grain 2 in cell 2, grain 3 in cell 1
    o = Oracle(s.to_state())
    o.step(1)
    row = row_of(o, 4)
    assert row[31] == 1 and row[30] == 3 and row[29] == 2
    assert numpy.all(row[:29] == 4)


def test_slots_insert_remove_reinsert():
    """Topology runs before interact removes separated contacts, so a grain
    arriving in the step its predecessor leaves takes the next lower slot;
    the freed slot goes to the next newcomer."""
    base = [0.029, 0.061, 0.0451, 0.085]  # grain 4 far away at first
    s = slot_case(base)
    n = 5
    o = Oracle(s.to_state())
    o.step(1)
    row = row_of(o, n)
    assert (row[31], row[30], row[29]) == (1, 3, 2)
    contacts, distmod, delta_t = o.read_contacts()
    # grain 3 leaves (cell 2, 0.03 m away), grain 4 arrives in cell 0: the
    # search meets 1, 4 (cell 0) and 2 (cell 2); slot 30 still holds the
    # stale grain 3, so 4 takes 28; interact then empties slot 30
    s2 = slot_case(base)
    s2.x[3] = [0.075, 0.015, 0.015]
    s2.x[4] = [0.028, 0.015, 0.020]
    o2 = Oracle(s2.to_state())
    o2.write_contacts(contacts, distmod, delta_t)
    o2.step(1)
    row = row_of(o2, n)
    assert (row[31], row[30], row[29], row[28]) == (1, n, 2, 4)
    assert numpy.all(row[:28] == n)
    contacts, distmod, delta_t = o2.read_contacts()
    # grain 3 returns and takes the highest empty slot, 30
    s3 = slot_case(base)
    s3.x[4] = [0.028, 0.015, 0.020]
    o3 = Oracle(s3.to_state())
    o3.write_contacts(contacts, distmod, delta_t)
    o3.step(1)
    row = row_of(o3, n)
    assert (row[31], row[30], row[29], row[28]) == (1, 3, 2, 4)


def test_slots_full_row_written_at_nc():
    """A row full of stale partners has no empty slot: the k-th contact found
    this step is written to slot k (the counter of contacts found so far),
    and the stale partners are removed by interact afterwards."""
    s = slot_case([0.029, 0.061])
    # 32 far grains to fill the row with stale entries
    far = numpy.column_stack([numpy.full(32, 0.085), numpy.full(32, 0.015),
                              numpy.linspace(0.001, 0.029, 32)])
    x = numpy.vstack([s.x, far])
    n = x.shape[0]
    s = free_case(x, numpy.concatenate([[0.01, 0.01, 0.01], numpy.full(32, 0.0001)]),
                  numpy.array([0.09, 0.03, 0.03]), numpy.array([3, 1, 1]))
    o = Oracle(s.to_state())
    contacts = numpy.full(n * 32, n, dtype=numpy.uint32)
    contacts[:32] = numpy.arange(3, 35, dtype=numpy.uint32)
    o.write_contacts(contacts, numpy.zeros(n * 32 * 3), numpy.zeros(n * 32 * 3))
    o.step(1)
    row = row_of(o, n)
    assert row[0] == 1 and row[1] == 2
    assert numpy.all(row[2:] == n)
