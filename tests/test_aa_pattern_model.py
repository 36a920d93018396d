"""The one-array (AA-pattern) update rule of the fused step, checked on the
CPU before any kernel runs it: a model of the even step (own slots read,
collided, written back swapped), the odd step (gather from the neighbours'
swapped slots or, where the source is not fluid, from the cell's own slot;
scatter to exactly the slots read) and the out-of-place gather that reads a
state after an odd number of steps, against a plain two-array push step
(the corrected step of the oracle: bounce-back into the cell's own reverse
slot, the moving-wall term where a wall face is hit).  The collision is an
arbitrary nonlinear per-cell map, so the check covers every index and
bounce-back decision independently of the BGK arithmetic; the odd step's read
and write sets are shown identical per cell and disjoint between cells, so
the in-place update has no race."""
import numpy as np
import pytest

LATTICES = {
    9: ([(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 0, 1), (0, 0, -1), (1, 0, 1), (-1, 0, -1), (1, 0, -1),
         (-1, 0, 1)]),
    19: ([(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, 0),
          (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1),
          (0, 1, -1), (0, -1, 1)]),
}


def setup(q, X, Y, Z, seed, walls, solids):
    c = np.array(LATTICES[q])
    rev = np.array([next(j for j in range(q) if (c[j] == -c[k]).all()) for k in range(q)])
    rng = np.random.default_rng(seed)
    flags = np.zeros((Z, Y, X), np.int64)  # 0 fluid, 1 wall face, 7 solid
    if walls:
        flags[0], flags[-1] = 1, 1
    if solids:
        s = rng.random((Z, Y, X)) < 0.15
        flags[(flags == 0) & s] = 7
    a = rng.uniform(0.5, 0.9, (Z, Y, X))
    b = rng.uniform(-0.1, 0.1, (q, Z, Y, X))
    term = rng.uniform(-0.01, 0.01, (q, Z, Y, X))
    f0 = rng.uniform(0.0, 1.0, (q, Z, Y, X))
    return c, rev, flags, a, b, term, f0


def nbr(x, y, z, d, X, Y, Z):
    return ((z + d[2]) % Z, (y + d[1]) % Y, (x + d[0]) % X)


def collide(f, cell, a, b):
    """An arbitrary nonlinear map of the cell's populations."""
    s = f.sum()
    return a[cell] * f + b[(slice(None),) + cell] + 0.05 * np.sin(s) * f[::-1]


def push_step(F, c, rev, flags, a, b, term):
    q, Z, Y, X = F.shape
    N = np.zeros_like(F)
    for z in range(Z):
        for y in range(Y):
            for x in range(X):
                cell = (z, y, x)
                if flags[cell]:
                    continue
                fs = collide(F[(slice(None),) + cell], cell, a, b)
                for j in range(q):
                    d = nbr(x, y, z, c[j], X, Y, Z)
                    if flags[d] == 0:
                        N[(j,) + d] = fs[j]
                    else:
                        N[(rev[j],) + cell] = fs[j] - (term[(j,) + cell] if flags[d] == 1 else 0.0)
    return N


def aa_slots(cell, c, rev, flags, X, Y, Z):
    """(population, cell) read for population k: S[rev k][x - c_k], or S[k][x]
    where x - c_k is not fluid."""
    z, y, x = cell
    out = []
    for k in range(len(c)):
        src = nbr(x, y, z, -c[k], X, Y, Z)
        out.append((k,) + cell if flags[src] else (rev[k],) + src)
    return out


def aa_even(A, c, rev, flags, a, b, term):
    q, Z, Y, X = A.shape
    for z in range(Z):
        for y in range(Y):
            for x in range(X):
                cell = (z, y, x)
                if flags[cell]:
                    A[(slice(None),) + cell] = 0.0
                    continue
                fs = collide(A[(slice(None),) + cell].copy(), cell, a, b)
                for j in range(q):
                    d = nbr(x, y, z, c[j], X, Y, Z)
                    if flags[d] == 1:
                        fs[j] -= term[(j,) + cell]
                for k in range(q):
                    A[(rev[k],) + cell] = fs[k]


def aa_odd(A, c, rev, flags, a, b, term):
    """In place, cell by cell; returns every cell's read and write sets."""
    q, Z, Y, X = A.shape
    snapshot = A.copy()
    reads, writes = {}, {}
    for z in range(Z):
        for y in range(Y):
            for x in range(X):
                cell = (z, y, x)
                if flags[cell]:
                    continue
                slot = aa_slots(cell, c, rev, flags, X, Y, Z)
                f = np.array([snapshot[s] for s in slot])
                fs = collide(f, cell, a, b)
                for j in range(q):
                    d = nbr(x, y, z, c[j], X, Y, Z)
                    if flags[d] == 1:
                        fs[j] -= term[(j,) + cell]
                reads[cell] = set(slot)
                writes[cell] = set(slot[rev[j]] for j in range(q))
                for j in range(q):
                    A[slot[rev[j]]] = fs[j]
    return reads, writes


def aa_gather(A, c, rev, flags):
    q, Z, Y, X = A.shape
    B = np.zeros_like(A)
    for z in range(Z):
        for y in range(Y):
            for x in range(X):
                cell = (z, y, x)
                if flags[cell]:
                    continue
                for k, s in enumerate(aa_slots(cell, c, rev, flags, X, Y, Z)):
                    B[(k,) + cell] = A[s]
    return B


@pytest.mark.parametrize("q,X,Y,Z,walls,solids", [(9, 7, 1, 6, True, False), (9, 6, 1, 7, True, True),
                                                  (19, 5, 4, 6, True, True), (19, 4, 5, 4, False, True),
                                                  (19, 4, 3, 5, False, False)])
def test_aa_steps_equal_two_array_steps(q, X, Y, Z, walls, solids):
    c, rev, flags, a, b, term, f0 = setup(q, X, Y, Z, 3, walls, solids)
    fluid = flags == 0
    F = f0.copy()
    A = f0.copy()
    for step in range(6):
        F = push_step(F, c, rev, flags, a, b, term)
        if step % 2 == 0:
            aa_even(A, c, rev, flags, a, b, term)
            got = aa_gather(A, c, rev, flags)  # a read after an odd number of steps
        else:
            reads, writes = aa_odd(A, c, rev, flags, a, b, term)
            got = A
            all_reads = [s for r in reads.values() for s in r]
            assert len(all_reads) == len(set(all_reads))  # no slot read by two cells
            assert all(reads[cell] == writes[cell] for cell in reads)  # each cell writes what it read
        assert np.array_equal(got[:, fluid], F[:, fluid]), step
