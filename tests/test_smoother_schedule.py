"""The phase smoother's LDS schedule has no cross-thread hazard (CPU model).

smooth3_kernel (rust-simulation-tools_amd/csrc/mg.hip) streams a 32 x 32 or
44 x 44 region through an 8-plane LDS ring and runs six half-sweeps per step with a
single workgroup barrier, at the end of the step.  That is exact only if, within
one barrier interval, no LDS word written by one thread is read or written by
another.  This model replays the kernel's accesses for one workgroup -- the
plane stores (each thread stores the four nodes of its own column pairs), the
write-back reads of the finished plane t - 6 and, for every active update, the
seven stencil reads and the store -- and checks that property on interior,
face and corner tiles.  A planted schedule (row-mapped plane stores, as the
kernel had while it kept a barrier after them) must be caught.
"""
from collections import defaultdict

import pytest

K, H, RING = 6, 6, 8
PP, PF = 2, 4               # pairs and plane values per thread


def conflicts(n, tile, z0, z1, pair_stores=True, region=32, shuffle=None):
    """Words of one workgroup touched by two threads with a write in one
    barrier interval, for a tile at core origin `tile` and planes z0 .. z1-1 of
    an n^3 level, with a region x region workgroup.  Thread t owns pair
    shuffle[t] of the row-major order (the kernel's 44 x 44 region assigns
    pairs to lanes by LDS bank, not row by row); the schedule's correctness must
    not depend on that assignment."""
    RW = RH = region
    RN = RW * RH
    TX = TY = RW - 2 * H    # tile core
    PPR, PR = RW // 2, RH // 2  # pairs per row, rows between a thread's pairs
    NP = RN // 4            # threads owning pairs
    NT = (NP + 63) // 64 * 64
    WX = TX + 2             # write-back rectangle width
    WB = (WX * (TY + 2) + NT - 1) // NT
    owner = list(range(NP)) if shuffle is None else shuffle(NP)

    def pair_of(tid):
        p = owner[tid]
        return p // PPR, 2 * (p % PPR)

    x0, y0 = tile
    xl, yl = x0 - H, y0 - H
    x1, y1 = min(n - 1, x0 + TX), min(n - 1, y0 + TY)
    wx0, wx1 = (0 if x0 == 1 else x0), (n if x1 == n - 1 else x1)
    wy0, wy1 = (0 if y0 == 1 else y0), (n if y1 == n - 1 else y1)
    wz0, wz1 = (0 if z0 == 1 else z0), (n if z1 == n - 1 else z1)
    t0, t1 = z0 - K + 1, z1 + K - 1
    seen = defaultdict(lambda: (set(), set()))  # (interval, word) -> (readers, writers)
    interval = [0]

    def slot(z):
        return ((z + 2 * RING) % RING) * RN

    def touch(write, tid, word):
        readers, writers = seen[(interval[0], word)]
        (writers if write else readers).add(tid)

    def store_plane(z):
        for tid in range(NT):
            if pair_stores:
                if tid >= NP:
                    continue
                ry0, rx0 = pair_of(tid)
                for m in range(PP):
                    for c in range(2):
                        touch(True, tid, slot(z) + (ry0 + PR * m) * RW + rx0 + c)
            else:
                for j in range(PF):
                    if tid + NT * j < RN:
                        touch(True, tid, slot(z) + tid + NT * j)

    def write_back(z):
        if z < wz0 or z >= wz1:
            return
        for tid in range(NT):
            for j in range(WB):
                v = tid + NT * j
                r, c = divmod(v, WX)
                if r < wy1 - wy0 and c < wx1 - wx0:
                    touch(False, tid, slot(z) + (wy0 + r - yl) * RW + (wx0 + c - xl))

    def half_sweep(k, t):
        z = t - k + 1
        if not (1 <= z <= n - 2 and z0 - (K - k) <= z <= z1 - 1 + (K - k)):
            return
        for tid in range(NP):
            ry0, rx0 = pair_of(tid)
            c = (xl + rx0 + yl + ry0 + t) & 1  # the pair's active column
            for m in range(PP):
                ry, rx = ry0 + PR * m, rx0 + c
                x, y = xl + rx, yl + ry
                if min(rx, RW - 1 - rx, ry, RH - 1 - ry) < k or not (1 <= x <= n - 2 and
                                                                      1 <= y <= n - 2):
                    continue
                li = ry * RW + rx
                for w in (slot(z) + li + d for d in (0, -1, 1, -RW, RW)):
                    touch(False, tid, w)
                touch(False, tid, slot(z - 1) + li)
                touch(False, tid, slot(z + 1) + li)
                touch(True, tid, slot(z) + li)

    store_plane(t0 - 1)
    store_plane(t0)
    interval[0] += 1  # the barrier after the prologue
    for t in range(t0, t1 + 1):
        store_plane(t + 1)
        write_back(t - K)
        for k in range(1, K + 1):
            half_sweep(k, t)
        interval[0] += 1  # the step's one barrier
    write_back(z1)
    return sum(1 for readers, writers in seen.values()
               if writers and (len(writers) > 1 or readers - writers))


@pytest.mark.parametrize("n,tile,zr,region", [
    (161, (1, 1), (1, 14), 32),        # corner tile, first chunk
    (161, (61, 81), (97, 110), 32),    # interior tile, a chunk in the middle
    (161, (141, 141), (148, 160), 32), # far corner, last chunk
    (41, (21, 1), (1, 40), 32),        # face tile of a small level, whole column
    (161, (1, 1), (1, 12), 44),        # the 44 x 44 region: corner, first chunk
    (161, (65, 97), (81, 92), 44),     # interior tile
    (161, (129, 129), (150, 160), 44), # last tile (30-node core), last chunk
])
def test_one_barrier_per_step_is_race_free(n, tile, zr, region):
    assert conflicts(n, tile, *zr, region=region) == 0


def test_any_pair_assignment_is_race_free():
    import random

    def shuffled(n):
        owner = list(range(n))
        random.Random(7).shuffle(owner)
        return owner
    assert conflicts(161, (65, 97), 81, 90, region=44, shuffle=shuffled) == 0
    assert conflicts(161, (1, 1), 1, 10, region=44, shuffle=shuffled) == 0


@pytest.mark.parametrize("region", [32, 44])
def test_model_catches_a_planted_hazard(region):
    # plane t + 1 stored by row instead of by the reading thread's own pairs:
    # other threads read it in the same interval
    assert conflicts(161, (65, 81), 97, 104, pair_stores=False, region=region) > 1000
