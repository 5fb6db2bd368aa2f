"""The tile schedule of the LU-SGS sweeps (lusgs_tile_kernel in
combustioncodegpu_amd/csrc/kernels_ns.hip), emulated on the CPU: tiles of
4 x 8 x 8 cells, one launch per tile hyperplane, the tile and its upstream
face layers staged at the start of a launch (values final before it), lane
(jl, kl) updating cell il = s - jl - kl at local step s, every cell's slot
written once, the tile written back at the end; staged cell (i, j, k) sits
in LDS slot i + 5 j + 46 k (one pad slot per k plane).  With the recurrence
y_c = (r_c + sum over upstream neighbours n of (y_n / 2 + 1/4)) / 3 (forward)
and y_c = r_c + (...) / 3 (backward) the emulation must reproduce a plain
sequential sweep exactly, on blocks whose extents are not multiples of the
tile, and every neighbour it reads must be staged or already updated."""
import itertools

import numpy as np

TI, TJ, TK = 4, 8, 8
EI, EJ, EK = TI + 1, TJ + 1, TK + 1
STEPS = TI + TJ + TK - 2
SK = EI * EJ + 1                    # k-plane stride of the LDS slots
SLOTS = (EK - 1) * SK + EI * EJ     # slots per staged variable


def slot(i, j, k):
    return k * SK + j * EI + i


def test_slot_map_is_one_to_one_and_in_range():
    slots = [slot(i, j, k) for k in range(EK) for j in range(EJ) for i in range(EI)]
    assert len(set(slots)) == EI * EJ * EK and min(slots) == 0 and max(slots) == SLOTS - 1


def sequential(r, m, fwd, dim):
    y = np.zeros_like(r)
    order = list(itertools.product(range(m[2]), range(m[1]), range(m[0])))
    for k, j, i in (order if fwd else order[::-1]):
        c = (i, j, k)
        acc = r[c] if fwd else 0.0
        for d in range(dim):
            n = list(c)
            n[d] += -1 if fwd else 1
            if 0 <= n[d] < m[d]:
                acc += 0.5 * y[tuple(n)] + 0.25
        y[c] = acc / 3.0 if fwd else r[c] + acc / 3.0
    return y


def tiled(r, m, fwd, dim):
    dq = r.copy()
    nt = [-(-m[0] // TI), -(-m[1] // TJ), -(-m[2] // TK)]
    ntp = sum(nt) - 2
    o = 1 if fwd else 0
    for n in range(ntp):
        tplane = n if fwd else ntp - 1 - n
        before = dq.copy()
        writes = []
        for t in range(nt[1] * nt[2]):
            t1, t2 = t % nt[1], t // nt[1]
            t0 = tplane - t1 - t2
            if t2 >= nt[2] or t0 < 0 or t0 >= nt[0]:
                continue
            org = (t0 * TI - o, t1 * TJ - o, t2 * TK - o)
            ext = np.full(SLOTS, np.nan)
            for n in range(EI * EJ * EK):
                li, lj, lk = n % EI, (n // EI) % EJ, n // (EI * EJ)
                c = (org[0] + li, org[1] + lj, org[2] + lk)
                if all(0 <= c[q] < m[q] for q in range(3)):
                    ext[slot(li, lj, lk)] = before[c]
            done = set()
            for step in range(STEPS):
                upd = {}
                for lane in range(64):
                    jl, kl = lane % TJ, lane // TJ
                    il = (step if fwd else STEPS - 1 - step) - jl - kl
                    c = (t0 * TI + il, t1 * TJ + jl, t2 * TK + kl)
                    if not (0 <= il < TI and all(c[q] < m[q] for q in range(3))):
                        continue
                    e = slot(il + o, jl + o, kl + o)
                    acc = ext[e] if fwd else 0.0
                    for d, es in enumerate((1, EI, SK)[:dim]):
                        if (c[d] == 0) if fwd else (c[d] == m[d] - 1):
                            continue
                        en = e - es if fwd else e + es
                        nb = list(c)
                        nb[d] += -1 if fwd else 1
                        inside = (t0 * TI <= nb[0] < t0 * TI + TI and t1 * TJ <= nb[1] < t1 * TJ + TJ
                                  and t2 * TK <= nb[2] < t2 * TK + TK)
                        assert 0 <= en < SLOTS and not np.isnan(ext[en])
                        assert en in done or not inside, "in-tile neighbour read before its update"
                        acc += 0.5 * ext[en] + 0.25
                    upd[e] = acc / 3.0 if fwd else ext[e] + acc / 3.0
                for e, x in upd.items():
                    assert e not in done
                    ext[e] = x
                    done.add(e)
            for cc in range(TI * TJ * TK):
                il, jc, kc = cc % TI, (cc // TI) % TJ, cc // (TI * TJ)
                c = (t0 * TI + il, t1 * TJ + jc, t2 * TK + kc)
                if all(c[q] < m[q] for q in range(3)):
                    writes.append((c, ext[slot(il + o, jc + o, kc + o)]))
        for c, x in writes:
            dq[c] = x
    return dq


def test_tile_schedule_equals_sequential_sweep():
    rng = np.random.default_rng(3)
    for m in [(9, 17, 3), (6, 9, 10), (7, 11, 1)]:
        dim = 2 if m[2] == 1 else 3
        r = rng.random(m)
        fw = tiled(r, m, True, dim)
        assert np.array_equal(fw, sequential(r, m, True, dim)), m
        assert np.array_equal(tiled(fw, m, False, dim), sequential(fw, m, False, dim)), m
