"""Seeded synthetic workloads (grids, species, initial fields, BCs).

The reference initialises from input decks (Initialize.cu:9-111,
Field.cu:26-77) that are not shipped; these generators produce the same
kind of data deterministically so that the HIP path and the CPU oracle see
identical bytes.  Random numbers come from a counter-based SplitMix64 hash
(Steele, Lea & Flood 2014), which is easy to reproduce bit-exactly in any
language.  Every field is filled over the ghost-padded range, as the
reference's initialisation does (Field.cu:48-75).

Workloads:
  c0   2-D 128x64 air, Mach-2 inflow with a Gaussian density bump,
       inflow / outflow / slip walls, MUSCL-minmod, AUSM+, viscous, RK3.
  c1   3-D 256x128x128, 9 synthetic species (NASA-7, Lennard-Jones),
       smooth seeded T/p/u fields, Dirichlet mass fractions smoothed,
       outflow on all faces, Van Albada, AUSM+, viscous + Wilke, RK3.
  c3   64 blocks on a 4x4x4 face-matched lattice, extents drawn from
       [16, 160] per lattice axis (about 40M points), c1 species and fields.
All accept reduced sizes for tests.
"""
from __future__ import annotations

import concurrent.futures as cf
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import abi

R_U = 8314.462618  # Constants.h:19
NGG = 2

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n doubles in [0,1) from SplitMix64 of counters offset+1..offset+n of
    (seed, stream)."""
    with np.errstate(over="ignore"):
        base = _mix(np.array([(seed * 0x100000001B3 + stream * 0x9E3779B1 + 1) & 0xFFFFFFFFFFFFFFFF],
                             dtype=np.uint64))[0]
        ctr = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = _mix(base + ctr * _GOLDEN)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclass
class BlockSpec:
    mx: int
    my: int
    mz: int
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    bc: list = field(default_factory=list)     # (range_start, range_end, type, cond)
    inner: list = field(default_factory=list)  # (src[6], tar[6], target_block 1-based)

    @property
    def padded_shape(self):
        return (self.mz + 2 * NGG, self.my + 2 * NGG, self.mx + 2 * NGG)

    @property
    def n_padded(self):
        a, b, c = self.padded_shape
        return a * b * c

    @property
    def n_interior(self):
        return self.mx * self.my * self.mz


@dataclass
class Case:
    name: str
    params: dict
    species: dict | None
    blocks: list
    inflow: list
    wall: list
    bv0: list   # per block (6, pz, py, px)
    yk0: list   # per block (ns, pz, py, px) or None

    @property
    def n_spec(self):
        return 0 if self.species is None else len(self.species["mw"])

    @property
    def n_var(self):
        return 5 + self.n_spec

    @property
    def n_cells(self):
        return sum(b.n_interior for b in self.blocks)

    def marshal(self):
        return abi.Marshalled(self)


def outflow_all_faces(mx, my, mz):
    """Outflow on the 6 faces of a block (0-based vertex ranges)."""
    X, Y, Z = mx - 1, my - 1, mz - 1
    return [
        ((0, 0, 0), (0, Y, Z), abi.BC_OUTFLOW, 0),
        ((X, 0, 0), (X, Y, Z), abi.BC_OUTFLOW, 0),
        ((0, 0, 0), (X, 0, Z), abi.BC_OUTFLOW, 0),
        ((0, Y, 0), (X, Y, Z), abi.BC_OUTFLOW, 0),
        ((0, 0, 0), (X, Y, 0), abi.BC_OUTFLOW, 0),
        ((0, 0, Z), (X, Y, Z), abi.BC_OUTFLOW, 0),
    ]


def cartesian_grid(mx, my, mz, dx, perturb=0.0, seed=0):
    """Vertex coordinates (k,j,i order flattened i-fastest); optional smooth
    curvilinear perturbation of amplitude perturb*dx (positive Jacobian)."""
    k, j, i = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    x = i * dx
    y = j * dx
    z = k * dx
    if perturb:
        ph = uniform(seed, 77, 6) * 2 * math.pi
        x = x + perturb * dx * np.sin(2 * math.pi * j / max(my - 1, 1) + ph[0]) * np.cos(
            2 * math.pi * k / max(mz - 1, 1) + ph[1])
        y = y + perturb * dx * np.sin(2 * math.pi * i / max(mx - 1, 1) + ph[2]) * np.cos(
            2 * math.pi * k / max(mz - 1, 1) + ph[3])
        if mz > 1:
            z = z + perturb * dx * np.sin(2 * math.pi * i / max(mx - 1, 1) + ph[4]) * np.cos(
                2 * math.pi * j / max(my - 1, 1) + ph[5])
    return [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (x, y, z)]


def padded_index_coords(mx, my, mz):
    k, j, i = np.meshgrid(np.arange(-NGG, mz + NGG), np.arange(-NGG, my + NGG), np.arange(-NGG, mx + NGG),
                          indexing="ij")
    return i.astype(np.float64), j.astype(np.float64), k.astype(np.float64)


# ------------------------------------------------------------------ species ---
def synthetic_species(n_spec=9, seed=2024, two_range=True, t_low=200.0):
    """n_spec synthetic species: W ~ U[2,44] kg/kmol; NASA-7 with t_low
    (200 K by default), t_mid 1000 K, t_high 3000 K and cp/R in [2.5, 6] on
    [200, 3000] K (checked on 2801 samples of each set); a5 ~ U[-3e4, 3e4] K;
    Lennard-Jones sigma ~ U[2.5,4] A, eps/k ~ U[30,600] K.

    two_range=True gives the high set its own coefficients, as real CHEMKIN
    data has: cp_hi(T) = cp_lo(T) + (T - t_mid)(b1 + b2 (T - t_mid)) with
    seeded b1, b2, and a5 of the high set chosen so that h is continuous at
    t_mid (cp is continuous there too, its slope is not).  two_range=False
    uses the low set for both ranges."""
    u = uniform(seed, 1, 8 * n_spec)
    mw = 2.0 + 42.0 * u[:n_spec]
    sigma = 2.5 + 1.5 * u[n_spec:2 * n_spec]
    eps_k = 30.0 + 570.0 * u[2 * n_spec:3 * n_spec]
    a5 = -3e4 + 6e4 * u[3 * n_spec:4 * n_spec]
    T0, D, TM = 200.0, 2800.0, 1000.0
    tt = np.linspace(200.0, 3000.0, 2801)

    def cp_of(a, t):
        return a[0] + t * (a[1] + t * (a[2] + t * (a[3] + t * a[4])))

    def h_of(a, t):  # h/R
        return t * (a[0] + t * (a[1] / 2 + t * (a[2] / 3 + t * (a[3] / 4 + t * a[4] / 5)))) + a[5]

    coeff = np.zeros((n_spec, 7))
    coeff_hi = np.zeros((n_spec, 7))
    for l in range(n_spec):
        attempt = 0
        while True:
            r = uniform(seed, 100 + 31 * l + attempt, 5)
            c = np.array([2.5 + 1.0 * r[0], 2.0 * r[1], -0.8 + 1.6 * r[2], -0.3 + 0.6 * r[3], -0.2 + 0.4 * r[4]])
            tau = (tt - T0) / D
            cpr = c[0] + tau * (c[1] + tau * (c[2] + tau * (c[3] + tau * c[4])))
            if cpr.min() >= 2.5 and cpr.max() <= 6.0:
                break
            attempt += 1
        # expand sum_m c_m ((T-T0)/D)^m into powers of T
        a = np.zeros(5)
        for m in range(5):
            for n in range(m + 1):
                a[n] += c[m] * math.comb(m, n) * (-T0) ** (m - n) / D ** m
        coeff[l, :5] = a
        coeff[l, 5] = a5[l]
        coeff[l, 6] = 0.0
        coeff_hi[l] = coeff[l]
        if two_range:
            attempt = 0
            while True:
                r = uniform(seed, 500 + 31 * l + attempt, 2)
                b1 = -6e-4 + 1.2e-3 * r[0]     # per K
                b2 = -2e-7 + 4e-7 * r[1]       # per K^2
                hi = coeff[l].copy()
                # (T - TM)(b1 + b2 (T - TM)) in powers of T
                hi[0] += -b1 * TM + b2 * TM * TM
                hi[1] += b1 - 2 * b2 * TM
                hi[2] += b2
                hi[5] += h_of(coeff[l], TM) - h_of(hi, TM)
                if 2.5 <= cp_of(hi, tt).min() and cp_of(hi, tt).max() <= 6.0:
                    break
                attempt += 1
            coeff_hi[l] = hi
    return {
        "mw": mw, "t_low": np.full(n_spec, float(t_low)), "t_mid": np.full(n_spec, TM),
        "t_high": np.full(n_spec, 3000.0), "coeff_lo": coeff, "coeff_hi": coeff_hi,
        "eps_k": eps_k, "sigma": sigma,
    }


# ---------------------------------------------------------------------- C0 ---
def c0(mx=128, my=64, seed=7, dx=1e-3, r_air=287.0, gamma=1.4, cfl=0.5, rk_stages=3, limiter=0):
    """2-D air, uniform Mach-2 inflow (T=300 K, p=101325 Pa) with a seeded
    Gaussian density bump (10 %, sigma = 8 cells) at uniform pressure;
    inflow at i-min, outflow at i-max, slip walls at j-min / j-max."""
    mz = 1
    x, y, z = cartesian_grid(mx, my, mz, dx)
    T0, p0 = 300.0, 101325.0
    rho0 = p0 / (r_air * T0)
    u0 = 2.0 * math.sqrt(gamma * r_air * T0)
    ii, jj, _ = padded_index_coords(mx, my, mz)
    c = uniform(seed, 3, 2)
    ic = 0.25 * (mx - 1) + 0.5 * (mx - 1) * c[0]
    jc = 0.25 * (my - 1) + 0.5 * (my - 1) * c[1]
    sig = 8.0
    r2 = (ii - ic) ** 2 + (jj - jc) ** 2
    rho = rho0 * (1.0 + 0.1 * np.exp(-r2 / (2 * sig * sig)))
    bv = np.zeros((6,) + rho.shape)
    bv[0] = rho
    bv[1] = u0
    bv[4] = p0
    bv[5] = p0 / (rho * r_air)
    X, Y = mx - 1, my - 1
    bc = [
        ((0, 0, 0), (0, Y, 0), abi.BC_INFLOW, 0),
        ((X, 0, 0), (X, Y, 0), abi.BC_OUTFLOW, 0),
        ((0, 0, 0), (X, 0, 0), abi.BC_SYMMETRY, 0),
        ((0, Y, 0), (X, Y, 0), abi.BC_SYMMETRY, 0),
    ]
    blk = BlockSpec(mx, my, mz, x, y, z, bc=bc)
    params = dict(inviscid_scheme=3, reconstruction=2, limiter=limiter, viscous_order=2, rk_stages=rk_stages,
                  cfl=cfl, prandtl=0.72, schmidt=0.9, gamma_air=gamma, r_air=r_air)
    inflow = [dict(mach=2.0, pressure=p0, temperature=T0, u=1.0, v=0.0, w=0.0)]
    return Case("c0", params, None, [blk], inflow, [], [np.ascontiguousarray(bv)], [None])


# ---------------------------------------------------------------------- C1 ---
def smooth_field(seed, stream, shape_idx, lengths):
    """S = 1/4 sum_{m=1..4} sin(2 pi k_m . x / L + phi_m), k_m in {1..4}^3."""
    ii, jj, kk = shape_idx
    r = uniform(seed, stream, 16)
    s = np.zeros_like(ii)
    for m in range(4):
        kv = [1 + int(4 * r[4 * m + d]) for d in range(3)]
        phi = 2 * math.pi * r[4 * m + 3]
        s += np.sin(2 * math.pi * (kv[0] * ii / lengths[0] + kv[1] * jj / lengths[1] + kv[2] * kk / lengths[2]) + phi)
    return 0.25 * s


def _box_filter(a, passes):
    for _ in range(passes):
        for ax in (0, 1, 2):
            ap = np.concatenate([np.take(a, [0], axis=ax), a, np.take(a, [-1], axis=ax)], axis=ax)
            sl = [slice(None)] * 3
            parts = []
            for o in range(3):
                sl[ax] = slice(o, o + a.shape[ax])
                parts.append(ap[tuple(sl)])
            a = (parts[0] + parts[1] + parts[2]) / 3.0
    return a


def _threads(n):
    return max(1, min(n, os.cpu_count() or 1, 8))


def smoothed_dirichlet(seed, stream, n_spec, shape, passes=4):
    """Per point Dirichlet(1) mass fractions (normalised exponentials), then
    `passes` separable 3-point box filters (edge-replicated), renormalised.
    Species are drawn and filtered independently (the filter is linear per
    species), in parallel threads on large grids."""
    n = int(np.prod(shape))
    y = np.empty((n_spec,) + tuple(shape))

    def draw(l):
        y[l] = -np.log1p(-uniform(seed, stream, n, offset=l * n)).reshape(shape)

    def smooth(l):
        y[l] = _box_filter(y[l], passes)

    with cf.ThreadPoolExecutor(_threads(n_spec)) as ex:
        list(ex.map(draw, range(n_spec)))
        y /= y.sum(axis=0, keepdims=True)
        list(ex.map(smooth, range(n_spec)))
    return y / y.sum(axis=0, keepdims=True)


def _smooth_fields(seed, n_spec, sp, idx, L):
    """The c1 field generator on index coordinates idx = (i, j, k) arrays."""
    ii, jj, kk = idx
    with cf.ThreadPoolExecutor(_threads(5)) as ex:
        s = list(ex.map(lambda st: smooth_field(seed, st, (ii, jj, kk), L), (10, 11, 12, 13, 14)))
    T = 1400.0 + 1100.0 * s[0]
    p = 101325.0 * (1.0 + 0.1 * s[1])
    amp = 200.0 / math.sqrt(3.0)
    u, v, w = amp * s[2], amp * s[3], amp * s[4]
    if sp is not None:
        Y = smoothed_dirichlet(seed, 20, n_spec, ii.shape)
        r_mix = R_U * np.einsum("l...,l->...", Y, 1.0 / sp["mw"])
    else:
        Y = None
        r_mix = np.full(ii.shape, R_U / 28.9647)
    rho = p / (r_mix * T)
    return np.stack([rho, u, v, w, p, T]), Y


def c1(mx=256, my=128, mz=128, n_spec=9, seed=11, dx=1e-3, perturb=0.0, cfl=0.5, rk_stages=3, limiter=1,
       bc="outflow", species_seed=2024, viscous=2, quirks=abi.QUIRKS_REFERENCE, t_low=200.0, reconstruction=2):
    """3-D block with n_spec synthetic species and smooth seeded fields:
    T = 1400 + 1100 S_T, p = 101325 (1 + 0.1 S_p), u_d = 200/sqrt(3) S_ud,
    Y smoothed Dirichlet(1); outflow on all faces."""
    sp = synthetic_species(n_spec, species_seed, t_low=t_low) if n_spec > 0 else None
    x, y, z = cartesian_grid(mx, my, mz, dx, perturb=perturb, seed=seed)
    L = (max(mx - 1, 1), max(my - 1, 1), max(mz - 1, 1))
    bv, Y = _smooth_fields(seed, n_spec, sp, padded_index_coords(mx, my, mz), L)
    blk = BlockSpec(mx, my, mz, x, y, z, bc=outflow_all_faces(mx, my, mz))
    params = dict(inviscid_scheme=3, reconstruction=reconstruction, limiter=limiter, viscous_order=viscous,
                  rk_stages=rk_stages, cfl=cfl, prandtl=0.72, schmidt=0.9, quirks=quirks)
    return Case("c1", params, sp, [blk], [], [], [np.ascontiguousarray(bv)],
                [None if Y is None else np.ascontiguousarray(Y)])


# ---------------------------------------------------------------------- C3 ---
def lattice_sizes(nb, lo, hi, seed):
    """Per lattice axis a, nb[a] block extents drawn uniformly from the
    integers [lo, hi]; block (a, b, c) has extents (X[a], Y[b], Z[c]), so
    neighbours share whole faces."""
    out = []
    for ax in range(3):
        u = uniform(seed, 40 + ax, nb[ax])
        out.append([lo + int((hi - lo + 1) * x) for x in u])
    return out


def lattice(sizes, n_spec=9, seed=11, dx=1e-3, cfl=0.5, rk_stages=3, limiter=1, species_seed=2024, viscous=2,
            quirks=abi.QUIRKS_REFERENCE, name="lattice"):
    """Blocks on a face-matched lattice: sizes[a] lists the vertex extents
    along axis a; neighbouring blocks share their interface vertex layer and
    are joined by inner faces (both directions); the outer faces of the
    lattice are outflow.  Grid, fields and species come from the c1
    generator evaluated on the global vertex grid, so the blocks together
    hold exactly the fields of one block of the global extent."""
    nbx, nby, nbz = (len(s) for s in sizes)
    G = [sum(s) - (len(s) - 1) for s in sizes]   # global vertex counts
    org = [np.concatenate([[0], np.cumsum(np.array(s) - 1)[:-1]]).astype(int) for s in sizes]
    sp = synthetic_species(n_spec, species_seed) if n_spec > 0 else None
    gx, gy, gz = cartesian_grid(G[0], G[1], G[2], dx)
    gshape = (G[2], G[1], G[0])
    gxyz = [a.reshape(gshape) for a in (gx, gy, gz)]
    idx = padded_index_coords(G[0], G[1], G[2])
    L = (max(G[0] - 1, 1), max(G[1] - 1, 1), max(G[2] - 1, 1))
    bv_g, y_g = _smooth_fields(seed, n_spec, sp, idx, L)
    del idx

    def bid(a, b, c):
        return a + nbx * (b + nby * c)

    blocks, bvs, yks = [], [], []
    for c in range(nbz):
        for b in range(nby):
            for a in range(nbx):
                mx, my, mz = sizes[0][a], sizes[1][b], sizes[2][c]
                ox, oy, oz = org[0][a], org[1][b], org[2][c]
                xs = [np.ascontiguousarray(g[oz:oz + mz, oy:oy + my, ox:ox + mx]).ravel() for g in gxyz]
                X, Y, Z = mx - 1, my - 1, mz - 1
                bc, inner = [], []
                # (axis, side, neighbour lattice index or None)
                nbr = [(0, 0, (a - 1, b, c) if a > 0 else None), (0, 1, (a + 1, b, c) if a < nbx - 1 else None),
                       (1, 0, (a, b - 1, c) if b > 0 else None), (1, 1, (a, b + 1, c) if b < nby - 1 else None),
                       (2, 0, (a, b, c - 1) if c > 0 else None), (2, 1, (a, b, c + 1) if c < nbz - 1 else None)]
                ext = (mx, my, mz)
                for ax, side, nb_ in nbr:
                    if nb_ is None:
                        rs = [0, 0, 0]
                        re = [X, Y, Z]
                        if side == 0:
                            re[ax] = 0
                        else:
                            rs[ax] = ext[ax] - 1
                        bc.append((tuple(rs), tuple(re), abi.BC_OUTFLOW, 0))
                        continue
                    text = (sizes[0][nb_[0]], sizes[1][nb_[1]], sizes[2][nb_[2]])
                    src_r, tar_r = [], []
                    first = True
                    for q in range(3):
                        if q == ax:
                            s_v = 1 if side == 0 else ext[q]
                            t_v = text[q] if side == 0 else 1
                            src_r += [s_v, s_v]
                            tar_r += [t_v, t_v]
                        elif first:  # first in-plane axis flagged negative
                            src_r += [-1, -ext[q]]
                            tar_r += [-1, -text[q]]
                            first = False
                        else:
                            src_r += [1, ext[q]]
                            tar_r += [1, text[q]]
                    inner.append((src_r, tar_r, bid(*nb_) + 1))
                blocks.append(BlockSpec(mx, my, mz, *xs, bc=bc, inner=inner))
                sl = np.s_[oz:oz + mz + 2 * NGG, oy:oy + my + 2 * NGG, ox:ox + mx + 2 * NGG]
                bvs.append(np.ascontiguousarray(bv_g[(slice(None),) + sl]))
                yks.append(None if y_g is None else np.ascontiguousarray(y_g[(slice(None),) + sl]))
    params = dict(inviscid_scheme=3, reconstruction=2, limiter=limiter, viscous_order=viscous,
                  rk_stages=rk_stages, cfl=cfl, prandtl=0.72, schmidt=0.9, quirks=quirks)
    return Case(name, params, sp, blocks, [], [], bvs, yks)


def c1_walls(mx=24, my=20, mz=16, n_spec=9, seed=11, dx=1e-3, rk_stages=3, limiter=1, t_low=500.0):
    """The C1 fields and species on one block with every boundary kind:
    inflow at i-min (T = 1400 K, p = 101325 Pa, 150 m/s along x, the mean
    composition of the field), outflow at i-max, an isothermal no-slip wall
    (600 K) at j-min, an adiabatic no-slip wall at j-max, slip walls at k-min
    and k-max.  t_low defaults to 500 K so that the field (300-2500 K) also
    reaches the NASA-7 extension below t_low."""
    c = c1(mx, my, mz, n_spec=n_spec, seed=seed, dx=dx, rk_stages=rk_stages, limiter=limiter, t_low=t_low)
    blk = c.blocks[0]
    X, Y, Z = mx - 1, my - 1, mz - 1
    blk.bc = [
        ((0, 0, 0), (0, Y, Z), abi.BC_INFLOW, 0),
        ((X, 0, 0), (X, Y, Z), abi.BC_OUTFLOW, 0),
        ((0, 0, 0), (X, 0, Z), abi.BC_WALL, 0),
        ((0, Y, 0), (X, Y, Z), abi.BC_WALL, 1),
        ((0, 0, 0), (X, Y, 0), abi.BC_SYMMETRY, 0),
        ((0, 0, Z), (X, Y, Z), abi.BC_SYMMETRY, 0),
    ]
    yk = c.yk0[0] if n_spec else None
    ymean = None if yk is None else np.ascontiguousarray(yk.reshape(n_spec, -1).mean(axis=1))
    inflow = [dict(mach=-1.0, velocity=150.0, pressure=101325.0, temperature=1400.0, u=1.0, v=0.0, w=0.0,
                   yk=ymean)]
    wall = [dict(adiabatic=0, temperature=600.0), dict(adiabatic=1)]
    return Case("c1_walls", dict(c.params), c.species, [blk], inflow, wall, c.bv0, c.yk0)


def c3(nb=(4, 4, 4), lo=16, hi=160, seed=2, n_spec=9, **kw):
    """64 blocks (4 x 4 x 4 lattice) with extents drawn from [16, 160]: about
    40M points, the c1 species and field generator."""
    return lattice(lattice_sizes(nb, lo, hi, seed), n_spec=n_spec, name="c3", **kw)


def cell_stages(case: Case, steps: int) -> int:
    """Interior cells x stages x steps (the metric's work unit)."""
    stages = 3 if case.params.get("rk_stages", 3) == 3 else 1
    return case.n_cells * stages * steps


def split_i(case: Case, ma: int) -> Case:
    """Split a single-block case along i into two blocks that share the
    vertex layer i = ma-1, joined by an inner face in the reference's inner
    file convention (1-based, the first in-plane axis flagged negative).
    Only blocks whose six faces are all outflow are supported."""
    assert len(case.blocks) == 1
    b = case.blocks[0]
    assert all(t == abi.BC_OUTFLOW for _, _, t, _ in b.bc)
    mx, my, mz = b.mx, b.my, b.mz
    mb = mx - ma + 1
    shape = (mz, my, mx)
    xs = [a.reshape(shape) for a in (b.x, b.y, b.z)]
    xa = [np.ascontiguousarray(a[:, :, :ma]).ravel() for a in xs]
    xb = [np.ascontiguousarray(a[:, :, ma - 1:]).ravel() for a in xs]
    dim3 = mz > 1
    zr = (1, mz) if dim3 else (1, 1)

    def faces(m_x, keep_min, keep_max):
        out = []
        for f in outflow_all_faces(m_x, my, mz):
            rs, re = f[0], f[1]
            if rs[0] == re[0] == 0 and not keep_min:
                continue
            if rs[0] == re[0] == m_x - 1 and not keep_max:
                continue
            if not dim3 and rs[2] == re[2] and (rs[0] != re[0] and rs[1] != re[1]):
                continue
            out.append(f)
        return out

    inner_a = ([ma, ma, -1, -my, zr[0], zr[1]], [1, 1, -1, -my, zr[0], zr[1]], 2)
    inner_b = ([1, 1, -1, -my, zr[0], zr[1]], [ma, ma, -1, -my, zr[0], zr[1]], 1)
    blk_a = BlockSpec(ma, my, mz, *xa, bc=faces(ma, True, False), inner=[inner_a])
    blk_b = BlockSpec(mb, my, mz, *xb, bc=faces(mb, False, True), inner=[inner_b])
    bv, yk = case.bv0[0], case.yk0[0]
    sl_a = np.s_[..., 0:ma + 2 * NGG]
    sl_b = np.s_[..., ma - 1:ma - 1 + mb + 2 * NGG]
    bv_a, bv_b = np.ascontiguousarray(bv[sl_a]), np.ascontiguousarray(bv[sl_b])
    yk_a = None if yk is None else np.ascontiguousarray(yk[sl_a])
    yk_b = None if yk is None else np.ascontiguousarray(yk[sl_b])
    return Case(case.name + "_split", dict(case.params), case.species, [blk_a, blk_b], case.inflow, case.wall,
                [bv_a, bv_b], [yk_a, yk_b])
