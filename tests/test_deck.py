"""The reference input-deck reader (combustioncodegpu_amd/deck.py) on the CPU.

Pinned by: hand-laid CHEMKIN records in the standard fixed columns with
known values and closed-form molecular weights (2 x 14.007 = 28.014,
2 x 1.008 + 15.999 = 18.015); the key-file rules of Parameter.cpp:41-71
(comments, first definition wins, typed values); round trips of the seeded
workloads through the writer; the initial field against a brute-force box
test over the oracle's ghost coordinates; the library's host-side ghost grid
and inflow states against the oracle bit for bit.
"""
import os

import numpy as np
import pytest

from combustioncodegpu_amd import abi, cases, deck, lib
from tests import deck_cases


# tests of the pure text readers; every other test calls the library's
# host-side setup (no device) and needs libcfx.so built
NO_LIBRARY = {"test_parameter_files", "test_file_names_pad_the_rank_with_spaces", "test_mechanism_therm_tran"}


@pytest.fixture(autouse=True)
def _library_built(request):
    if request.node.originalname not in NO_LIBRARY and not os.path.exists(lib.LIB_PATH):
        pytest.skip("libcfx.so is not built (python -m combustioncodegpu_amd.build)")


def _oracle(case):
    from oracle.pyoracle import Oracle
    return Oracle(case)


# --------------------------------------------------------- host-side setup ---
@pytest.mark.parametrize("make", [lambda: cases.c0(mx=20, my=12), lambda: cases.c1(14, 10, 8, perturb=0.3),
                                  lambda: cases.lattice([[6, 7], [5, 6], [5, 5]])])
def test_ghost_grid_matches_oracle(oracle_lib, make):
    case = make()
    g = lib.ghost_grid(case)
    o = _oracle(case)
    for b, blk in enumerate(case.blocks):
        oc = o.download(b, abi.COORD)
        assert np.array_equal(g[b], oc)
        # interior points are the input vertices
        n = np.s_[:, 2:2 + blk.mz, 2:2 + blk.my, 2:2 + blk.mx]
        want = np.stack([np.asarray(a).reshape(blk.mz, blk.my, blk.mx) for a in (blk.x, blk.y, blk.z)])
        assert np.array_equal(g[b][n], want)


def test_oracle_metric_fallback_on_folded_ghosts(oracle_lib):
    """The grid of the GPU fallback test does fold its ghosts at i = -2: the
    oracle's metrics there are copies of i = -1 (Mesh.cpp:226-488)."""
    from tests.test_gpu_parity import folded_ghost_case
    case = folded_ghost_case()
    X, Y, Z = lib.ghost_grid(case)[0]
    dy = (Y[2:-2, 4:-2, :] - Y[2:-2, 2:-4, :]) * 0.5    # d y / d j at j = 1..my-2, interior k
    assert (dy[:, :, 0] < 0).all() and (dy[:, :, 2:] > 0).all()
    m = _oracle(case).download(0, abi.METRIC)
    assert np.array_equal(m[:, 2:-2, 3:-3, 0], m[:, 2:-2, 3:-3, 1]) and (m[9] > 0).all()


def test_ghost_grid_cartesian_continues_linearly(oracle_lib):
    """On a uniform Cartesian block the ghost vertices continue the grid."""
    case = cases.c1(9, 8, 7, dx=1e-3)
    X, Y, Z = lib.ghost_grid(case)[0]
    ii, jj, kk = cases.padded_index_coords(9, 8, 7)
    for a, idx in ((X, ii), (Y, jj), (Z, kk)):
        assert np.abs(a - idx * 1e-3).max() < 1e-15


@pytest.mark.parametrize("make", [lambda: cases.c0(mx=16, my=10), lambda: cases.c1_walls(14, 12, 10),
                                  lambda: cases.c1_walls(14, 12, 10, n_spec=2)])
def test_inflow_state_matches_oracle_ghosts(oracle_lib, make):
    """The derived inflow state equals what the oracle's inflow condition
    writes into the ghost layers of the inflow face."""
    case = make()
    st, yk = lib.inflow_state(case, 0)
    o = _oracle(case)
    o.upload_initial()
    o.init_conservative()
    bv = o.download(0, abi.BV)
    blk = case.blocks[0]
    k = 0 if blk.mz == 1 else blk.mz // 2
    pt = (slice(None), k + 2, blk.my // 2 + 2, 1)   # i = -1, away from the other faces
    assert np.array_equal(bv[pt], st)
    if case.n_spec:
        assert np.array_equal(o.download(0, abi.YK)[pt], yk)


# ------------------------------------------------------------- key files ---
def test_parameter_files(tmp_path):
    s = tmp_path / "input_files" / "setup"
    s.mkdir(parents=True)
    (s / "0_global_control.txt").write_text(
        "// comment\n! comment\n\nint total_step = 100\ndouble cfl = 0.25\nbool steady = 1\n"
        "string mechanism_file = chem/mech.inp\nint total_step = 7\nint inviscid_scheme = 3\n")
    (s / "2_scheme.txt").write_text("double cfl = 0.9\nint limiter = 1   trailing words\n")
    p = deck.read_parameters(str(tmp_path))
    assert p["total_step"] == 100 and p["cfl"] == 0.25   # first definition wins
    assert p["steady"] is True and p["mechanism_file"] == "chem/mech.inp" and p["limiter"] == 1
    assert p["ngg"] == 2 and p["n_var"] == 5
    (s / "2_scheme.txt").write_text("int inviscid_scheme = 51\n")
    (s / "0_global_control.txt").write_text("int total_step = 1\n")
    assert deck.read_parameters(str(tmp_path))["ngg"] == 3   # WENO tags (5x) widen the halo


def test_file_names_pad_the_rank_with_spaces():
    assert deck.grid_path("/r").endswith(os.path.join("grid", "grid   0.grd"))
    assert deck.grid_path("/r", 12).endswith("grid  12.grd")


# ------------------------------------------------------------- chemistry ---
def _therm_head(name, comp, tlow, thigh, tmid):
    line = name.ljust(18) + "121286" + comp.ljust(20) + "G" + tlow.rjust(10) + thigh.rjust(10) + tmid.rjust(10)
    return line.ljust(79) + "1"


N2_ROWS = [" 0.02926640E+02 0.14879768E-02-0.05684760E-05 0.10097038E-09-0.06753351E-13    2",
           "-0.09227977E+04 0.05980528E+02 0.03298677E+02 0.14082404E-02-0.03963222E-04    3",
           " 0.05641515E-07-0.02444854E-10-0.10208999E+04 0.03950372E+02                   4"]
N2_HI = [2.926640, 1.4879768e-3, -5.684760e-7, 1.0097038e-10, -6.753351e-15, -922.7977, 5.980528]
N2_LO = [3.298677, 1.4082404e-3, -3.963222e-6, 5.641515e-9, -2.444854e-12, -1020.8999, 3.950372]


def _write_chem(tmp_path):
    mech = tmp_path / "mech.inp"
    mech.write_text("Elements\n--------\n3 elements\n--------\nO H\nn\nSpecies\n--------\n3 species\n--------\n"
                    "n2 h2o\nAR extra\nReactions\nH2+O2=2OH 1 0 0\n")
    head_h2o = _therm_head("H2O", "H   2O   1", "300.000", "5000.000", "")
    head_co = _therm_head("CO", "C   1O   1", "300.000", "5000.000", "1000.000")
    head_ar = _therm_head("AR", "AR  1", "200.000", "6000.000", "1500.000")
    rows = [" 1.00000000E+00 2.00000000E-03 3.00000000E-06 4.00000000E-09 5.00000000E-12    2",
            " 6.00000000E+03 7.00000000E+00 8.00000000E+00 9.00000000E-03 1.00000000E-05    3",
            " 1.10000000E-08 1.20000000E-11 1.30000000E+04 1.40000000E+01                   4"]
    therm = ["THERMO ALL", "   200.000  1000.000  6000.000", _therm_head("N2", "N   2", "300.000", "5000.000",
                                                                            "1000.000")] + N2_ROWS
    therm += ["! a comment between records", head_co] + rows + [head_h2o] + rows + [head_ar] + rows + ["END"]
    (tmp_path / "therm.dat").write_text("\n".join(therm) + "\n")
    (tmp_path / "tran.dat").write_text("! comment\nCO  1  98.100  3.650  0.000  1.950  1.800\n"
                                       "N2  1  97.530  3.621  0.000  1.760  4.000\n"
                                       "H2O 2 572.400  2.605  1.844  0.000  4.000\n"
                                       "AR  0 136.500  3.330  0.000  0.000  0.000\nEND\n")
    return mech


def test_mechanism_therm_tran(tmp_path):
    mech = _write_chem(tmp_path)
    elements, species = deck.read_mechanism(str(mech))
    assert elements == ["O", "H", "N"]
    assert species == ["N2", "H2O", "AR"]   # upper-cased, capped at the declared count
    with pytest.raises(deck.DeckError, match="AR not declared"):
        deck.read_therm(str(tmp_path / "therm.dat"), species, elements)
    species = species[:2]
    th = deck.read_therm(str(tmp_path / "therm.dat"), species, elements)
    assert th["mw"][0] == 2 * 14.007 and abs(th["mw"][1] - 18.015) < 1e-12
    assert th["mw"][1] == 1.008 * 2 + 15.999   # summed in sorted element order: H, N, O
    assert np.array_equal(th["coeff_hi"][0], N2_HI) and np.array_equal(th["coeff_lo"][0], N2_LO)
    assert list(th["coeff_hi"][1]) == [1.0, 2e-3, 3e-6, 4e-9, 5e-12, 6e3, 7.0]
    assert list(th["coeff_lo"][1]) == [8.0, 9e-3, 1e-5, 1.1e-8, 1.2e-11, 1.3e4, 14.0]
    assert list(th["t_low"]) == [300.0, 300.0] and list(th["t_high"]) == [5000.0, 5000.0]
    # a blank T_mid keeps the default; the global line after THERMO changes nothing
    assert list(th["t_mid"]) == [1000.0, 1000.0]
    assert th["elem_comp"].tolist() == [[0, 0, 2], [1, 2, 0]]
    eps, sig = deck.read_tran(str(tmp_path / "tran.dat"), species)
    assert list(eps) == [97.53, 572.4] and list(sig) == [3.621, 2.605]
    with pytest.raises(deck.DeckError, match="no transport data"):
        deck.read_tran(str(tmp_path / "tran.dat"), species + ["OH"])


def test_heavier_elements_and_missing_weights(tmp_path):
    """Molecular weights of species with heavier elements (Br, Zn) from the
    standard atomic weights: HBr = 1.008 + 79.904, ZnO = 65.38 + 15.999;
    elements without a standard weight (Tc) and the two the reference never
    matches (Ag, Hf) raise."""
    def chem(elems, sp_lines):
        mech = tmp_path / "m.inp"
        mech.write_text("Elements\n--------\n%d elements\n--------\n%s\nSpecies\n--------\n%d species\n"
                        "--------\n%s\nReactions\n" % (len(elems), " ".join(elems), len(sp_lines),
                                                         " ".join(n for n, _ in sp_lines)))
        therm = ["THERMO ALL", "   200.000  1000.000  6000.000"]
        for name, comp in sp_lines:
            therm += [_therm_head(name, comp, "300.000", "5000.000", "1000.000")] + N2_ROWS
        (tmp_path / "t.dat").write_text("\n".join(therm + ["END"]) + "\n")
        e, s_ = deck.read_mechanism(str(mech))
        return deck.read_therm(str(tmp_path / "t.dat"), s_, e)

    th = chem(["H", "BR", "ZN", "O"], [("HBR", "H   1BR  1"), ("ZNO", "ZN  1O   1")])
    assert th["mw"][0] == 1.008 + 79.904
    assert th["mw"][1] == 15.999 + 65.38   # sorted element order: O before ZN
    for el in ("TC", "AG", "HF"):
        with pytest.raises(deck.DeckError, match="no standard atomic weight"):
            chem(["H", el], [("X", "H   1%-3s1" % el)])


# ------------------------------------------------------------ round trips ---
def _same_blocks(a, b):
    assert len(a.blocks) == len(b.blocks)
    for p, q in zip(a.blocks, b.blocks):
        assert (p.mx, p.my, p.mz) == (q.mx, q.my, q.mz)
        for u, v in ((p.x, q.x), (p.y, q.y), (p.z, q.z)):
            assert np.array_equal(np.asarray(u).ravel(), v)
        assert [tuple(map(tuple, f[:2])) + f[2:] for f in p.bc] == [tuple(map(tuple, f[:2])) + f[2:] for f in q.bc]
        assert [(list(s), list(t), n) for s, t, n in p.inner] == [(list(s), list(t), n) for s, t, n in q.inner]


def _rounded_species(sp):
    out = {k: np.array([float(f"{v:15.8E}") for v in np.ravel(sp[k])]).reshape(np.shape(sp[k]))
           for k in ("coeff_lo", "coeff_hi")}
    out["eps_k"] = np.array([float(f"{v:.3f}") for v in sp["eps_k"]])
    out["sigma"] = np.array([float(f"{v:.3f}") for v in sp["sigma"]])
    return out


def test_walls_deck_round_trip(oracle_lib, tmp_path):
    case, init = deck_cases.walls_deck(str(tmp_path))
    got = deck.load_case(str(tmp_path))
    _same_blocks(case, got)
    for k in ("inviscid_scheme", "reconstruction", "limiter", "viscous_order", "rk_stages", "cfl", "prandtl",
              "schmidt"):
        assert got.params[k] == case.params[k], k
    assert got.wall == [dict(adiabatic=0, temperature=600.0), dict(adiabatic=1, temperature=300.0)]
    assert len(got.inflow) == 1 and np.array_equal(got.inflow[0]["yk"], case.inflow[0]["yk"])
    assert all(got.inflow[0][k] == case.inflow[0][k] for k in ("mach", "pressure", "temperature", "velocity"))
    want = _rounded_species(case.species)
    for k, v in want.items():
        assert np.array_equal(got.species[k], v), k
    comp_mw = [sum(w * n for w, n in zip((15.999, 1.008, 14.007), c)) for c in deck_cases.COMP]
    assert np.allclose(got.species["mw"], comp_mw, rtol=1e-15, atol=0)
    assert got.species["names"] == deck_cases.NAMES
    # writing the loaded case again gives the same deck
    again = tmp_path / "again"
    deck.write_deck(str(again), got, species_names=deck_cases.NAMES, elements=deck_cases.ELEMENTS,
                    elem_comp=deck_cases.COMP, init=init)
    for rel in ("chemistry/therm.dat", "chemistry/tran.dat", "setup/5_boundary_condition.txt",
                "boundary_condition/boundary   0.txt", "grid/grid   0.grd"):
        assert (tmp_path / "input_files" / rel).read_bytes() == (again / "input_files" / rel).read_bytes(), rel


def test_initial_groups_against_brute_force(oracle_lib, tmp_path):
    """Every padded point takes the first group (after group 0) whose box
    holds its ghost-grid coordinates, else group 0 (Field.cu:48-72)."""
    _, init = deck_cases.walls_deck(str(tmp_path))
    got = deck.load_case(str(tmp_path))
    o = _oracle(got)
    X, Y, Z = o.download(0, abi.COORD)
    states = []
    for g in range(len(init)):
        probe = cases.Case("probe", got.params, got.species, got.blocks, [got.inflow[0]] if g == 0 else
                           [dict(init[g][0])], [], [], [])
        states.append(lib.inflow_state(probe, 0))
    shp = X.shape
    n_hot = 0
    for idx in np.ndindex(*shp):
        g = 0
        for l in range(1, len(init)):
            (xs, xe), (ys, ye), (zs, ze) = init[l][1]
            if xs <= X[idx] <= xe and ys <= Y[idx] <= ye and zs <= Z[idx] <= ze:
                g = l
                break
        n_hot += g == 1
        assert np.array_equal(got.bv0[0][(slice(None),) + idx], states[g][0])
        assert np.array_equal(got.yk0[0][(slice(None),) + idx], states[g][1])
    assert 0 < n_hot < np.prod(shp)


def test_lattice_deck_round_trip(oracle_lib, tmp_path):
    case, _ = deck_cases.lattice_deck(str(tmp_path))
    got = deck.load_case(str(tmp_path))
    _same_blocks(case, got)
    assert got.inflow == [] and got.wall == []


def test_deck_case_steps_on_the_oracle(oracle_lib, tmp_path):
    deck_cases.walls_deck(str(tmp_path))
    got = deck.load_case(str(tmp_path))
    o = _oracle(got)
    o.upload_initial()
    o.init_conservative()
    o.step(2)
    assert np.isfinite(o.download(0, abi.CV)).all()
    assert o.newton_failures() == 0


def test_boundary_labels(oracle_lib, tmp_path):
    """Named labels of 5_boundary_condition.txt take their kind; default
    labels 2/5/6 without a line keep theirs; lines for labels no face uses
    are ignored; the condition index is the label's rank within its kind."""
    case = cases.c1_walls(12, 10, 8)
    labels = {(abi.BC_WALL, 0): 7, (abi.BC_WALL, 1): 2, (abi.BC_INFLOW, 0): 5, (abi.BC_OUTFLOW, 0): 6}
    deck.write_deck(str(tmp_path), case, species_names=deck_cases.NAMES, elements=deck_cases.ELEMENTS,
                    elem_comp=deck_cases.COMP, labels=labels, init=[(case.inflow[0], None)])
    bc_file = tmp_path / "input_files" / "setup" / "5_boundary_condition.txt"
    bc_file.write_text("// kinds\nwall hot_wall 7\noutflow unused 44\nend\n")
    got = deck.load_case(str(tmp_path))
    kinds = {f[2:]: None for f in got.blocks[0].bc}
    assert (abi.BC_WALL, 0) in kinds and (abi.BC_WALL, 1) in kinds
    faces = got.blocks[0].bc
    assert faces[2][2:] == (abi.BC_WALL, 0)   # label 7: named first
    assert faces[3][2:] == (abi.BC_WALL, 1)   # label 2: default after the named ones
    w7 = got.wall[0]
    assert w7 == dict(adiabatic=0, temperature=600.0)


def test_missing_records_raise(oracle_lib, tmp_path):
    case = cases.c1_walls(12, 10, 8)
    deck.write_deck(str(tmp_path), case, species_names=deck_cases.NAMES, elements=deck_cases.ELEMENTS,
                    elem_comp=deck_cases.COMP, init=[(case.inflow[0], None)])
    f = tmp_path / "input_files" / "setup" / "7_wall.txt"
    f.write_text("label 2\noption thermal_type = isothermal\nend\n")
    with pytest.raises(deck.DeckError, match="wall label 21"):
        deck.load_case(str(tmp_path))


def test_number_parser():
    """The library's grid-coordinate parser against Python's float() (both
    correctly rounded): signs, exponents, mixed whitespace, a cap on the
    count, a malformed token, and a text long enough to be cut into chunks
    for several threads."""
    got = lib.parse_numbers(b" 1 +2.5\n-3e-3\t4.0E+02 \r\n.5 ", 10)
    assert got.tolist() == [1.0, 2.5, -3e-3, 400.0, 0.5]
    assert lib.parse_numbers(b"1 2 3 4 5", 3).tolist() == [1.0, 2.0, 3.0]
    with pytest.raises(ValueError, match="token 2"):
        lib.parse_numbers(b"1 2 1.0D+00 4", 10)
    rng = np.random.default_rng(5)
    v = rng.standard_normal(2_000_000) * 10.0 ** rng.integers(-12, 12, 2_000_000)
    text = " \n".join(f"{x:.17g}" for x in v).encode()
    assert len(text) > 4 * (1 << 22)
    assert np.array_equal(lib.parse_numbers(text, v.size), v)


def test_air_2d_deck(oracle_lib, tmp_path):
    """The 2-D air workload (inflow, outflow, slip walls; no species files)
    through the deck: same blocks and conditions, the initial field is the
    inflow state, and the oracle steps it."""
    case = cases.c0(mx=24, my=12)
    deck.write_deck(str(tmp_path), case, init=[(case.inflow[0], None)])
    assert not (tmp_path / "input_files" / "chemistry").exists()
    got = deck.load_case(str(tmp_path))
    _same_blocks(case, got)
    assert got.species is None and got.n_spec == 0 and got.yk0 == [None]
    st, _ = lib.inflow_state(got, 0)
    assert np.array_equal(got.bv0[0], np.broadcast_to(st[:, None, None, None], got.bv0[0].shape))
    assert abs(st[1] - 2.0 * np.sqrt(1.4 * deck.R_U / deck.MW_AIR * 300.0)) < 1e-9   # Mach 2 at 300 K
    o = _oracle(got)
    o.upload_initial()
    o.init_conservative()
    o.step(3)
    assert np.isfinite(o.download(0, abi.CV)).all()


def test_grid_file_errors(tmp_path):
    with pytest.raises(deck.DeckError, match="cannot read"):
        deck.read_grid(str(tmp_path))
    g = tmp_path / "input_files" / "grid"
    g.mkdir(parents=True)
    path = g / "grid   0.grd"
    path.write_text("1\n3 3 1\n" + " ".join(["0.5"] * 26) + "\n")
    with pytest.raises(deck.DeckError, match="holds 26 coordinates, 27 needed"):
        deck.read_grid(str(tmp_path))
    path.write_text("1\n3 3 1\n" + " ".join(["0.5"] * 5 + ["1.0D+00"] + ["0.5"] * 21) + "\n")
    with pytest.raises(deck.DeckError, match="token 5"):
        deck.read_grid(str(tmp_path))
    path.write_text("1\n3 3 1\n" + " ".join(str(float(v)) for v in range(27)) + "\n")
    (mx, my, mz, x, y, z), = deck.read_grid(str(tmp_path))
    assert (mx, my, mz) == (3, 3, 1) and x.tolist() == list(range(9)) and z.tolist() == list(range(18, 27))


def test_grid_scale(oracle_lib, tmp_path):
    """`double gridScale` multiplies every coordinate (Mesh.cpp:537-543,
    704-716) unless it is 1 within 1e-15."""
    case, init = deck_cases.lattice_deck(str(tmp_path), sizes=((6, 5), (5,), (5,)))
    ctl = tmp_path / "input_files" / "setup" / "0_global_control.txt"
    ctl.write_text(ctl.read_text().replace("double gridScale = 1.0", "double gridScale = 0.001"))
    got = deck.load_case(str(tmp_path))
    for p, q in zip(case.blocks, got.blocks):
        assert np.array_equal(q.x, np.asarray(p.x) * 0.001) and np.array_equal(q.z, np.asarray(p.z) * 0.001)


def test_reference_time_scheme_without_rk_key(oracle_lib, tmp_path):
    """A deck without the (library-only) rk_stages key runs the reference's
    explicit Euler with local time steps."""
    deck_cases.walls_deck(str(tmp_path), box=False)
    ctl = tmp_path / "input_files" / "setup" / "0_global_control.txt"
    ctl.write_text("\n".join(l for l in ctl.read_text().splitlines() if "rk_stages" not in l) + "\n")
    got = deck.load_case(str(tmp_path))
    assert got.params["rk_stages"] == 1
    o = _oracle(got)
    o.upload_initial()
    o.init_conservative()
    o.step(2)
    assert np.isfinite(o.download(0, abi.CV)).all()
