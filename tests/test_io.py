"""Input formats (CPU: host-only parsing in libsil_hip.so): Gmsh MSH2
structured grids and .silpro problem files against the oracle's literal
restatement of the reference readers (oracle/io_oracle.py), on synthetic
meshes (shuffled node and element order, sparse node tags, boundary
elements, stretched axes) and every error the reference raises; the
magnetostatics cell inputs against their rule."""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
from oracle import io_oracle as O  # noqa: E402


@pytest.fixture(scope="module")
def IO():
    import similie_amd.io as IO

    return IO


def axis(n, rng, origin=0.0):
    return origin + np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 2.0, n - 1))]) * 0.1


def write_msh2(coords, cell_tags, rng, dim, sparse=True, extras=True, node_perm=True, fmt="2.2 0 8",
               tweak=None):
    """a MSH2 text of the rectilinear grid `coords` with one physical tag per
    cell (x fastest); node tags sparse, orders shuffled, boundary elements
    interleaved"""
    shape = [len(c) for c in coords]
    z = coords[2] if dim == 3 else [0.25]
    nz = len(z)
    nodes = []
    for k in range(nz):
        for j in range(shape[1]):
            for i in range(shape[0]):
                lin = i + shape[0] * (j + shape[1] * k)
                tag = 3 * lin + 7 if sparse else lin + 1
                nodes.append((tag, coords[0][i], coords[1][j], z[k]))
    tag_of = {}
    for t, x, y, zz in nodes:
        tag_of[(x, y, zz)] = t
    if node_perm:
        nodes = [nodes[q] for q in rng.permutation(len(nodes))]
    cells = []
    cshape = [s - 1 for s in shape]
    ck = cshape[2] if dim == 3 else 1
    for k in range(ck):
        for j in range(cshape[1]):
            for i in range(cshape[0]):
                lin = i + cshape[0] * (j + cshape[1] * k)
                if dim == 3:
                    corners = [(i + a, j + b, k + c) for c in (0, 1) for b in (0, 1) for a in (0, 1)]
                    ids = [tag_of[(coords[0][a], coords[1][b], z[c])] for a, b, c in corners]
                    cells.append((5, [int(cell_tags[lin]), 1], ids))
                else:
                    corners = [(i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)]
                    ids = [tag_of[(coords[0][a], coords[1][b], z[0])] for a, b in corners]
                    cells.append((3, [int(cell_tags[lin]), 1], ids))
    if extras:
        t0 = nodes[0][0]
        t1 = nodes[1][0]
        t2 = nodes[2][0]
        cells += [(15, [99, 1], [t0]), (1, [98, 2], [t0, t1]), (2, [97, 3], [t0, t1, t2])]
    cells = [cells[q] for q in rng.permutation(len(cells))]
    if tweak:
        nodes, cells = tweak(nodes, cells)
    lines = ["$MeshFormat", fmt, "$EndMeshFormat", "$PhysicalNames", "1", '3 1 "core"', "$EndPhysicalNames",
             "$Nodes", str(len(nodes))]
    lines += [f"{t} {float(x)!r} {float(y)!r} {float(zz)!r}" for t, x, y, zz in nodes]
    lines += ["$EndNodes", "$Elements", str(len(cells))]
    lines += [" ".join(str(v) for v in [e + 1, typ, len(tags), *tags, *ids]) for e, (typ, tags, ids) in
              enumerate(cells)]
    lines += ["$EndElements", ""]
    return "\n".join(lines)


def check_grid(IO, text):
    g = IO.parse_msh2(text)
    axes, tags, plane = O.read_structured_grid(text)
    assert g.ndim == len(axes)
    for a in range(g.ndim):
        np.testing.assert_array_equal(g.coords[a], np.array(axes[a]))
    np.testing.assert_array_equal(g.cell_tags, np.array(tags, dtype=np.int32))
    assert g.plane == plane
    return g


@pytest.mark.parametrize("sparse", [True, False])
def test_hexahedral_mesh(IO, sparse):
    rng = np.random.default_rng(1)
    coords = [axis(5, rng), axis(4, rng, -1.0), axis(6, rng, 2.0)]
    tags = rng.integers(1, 6, 4 * 3 * 5)
    g = check_grid(IO, write_msh2(coords, tags, rng, 3, sparse=sparse))
    np.testing.assert_array_equal(g.cell_tags, tags)
    for a in range(3):
        np.testing.assert_allclose(g.coords[a], coords[a], rtol=0, atol=0)


def test_quadrilateral_mesh(IO):
    rng = np.random.default_rng(2)
    coords = [axis(7, rng), axis(5, rng)]
    tags = rng.integers(1, 4, 6 * 4)
    g = check_grid(IO, write_msh2(coords, tags, rng, 2))
    np.testing.assert_array_equal(g.cell_tags, tags)
    assert g.plane == 0.25 and g.shape == (7, 5)


def test_reading_a_file(IO, tmp_path):
    rng = np.random.default_rng(3)
    coords = [axis(4, rng)] * 3
    tags = rng.integers(1, 3, 27)
    p = tmp_path / "grid.msh"
    p.write_text(write_msh2(coords, tags, rng, 3))
    g = IO.read_msh2(p)
    np.testing.assert_array_equal(g.cell_tags, tags)


def test_coordinates_merged_within_tolerance(IO):
    rng = np.random.default_rng(4)
    coords = [axis(4, rng)] * 3
    tags = np.arange(27)

    def jitter(nodes, cells):
        return [(t, x + 4e-13, y, z) if q % 2 else (t, x, y, z) for q, (t, x, y, z) in enumerate(nodes)], cells

    check_grid(IO, write_msh2(coords, tags, rng, 3, tweak=jitter))


def reject(IO, text, match):
    with pytest.raises(ValueError, match=match):
        O.read_structured_grid(text)
    L = __import__("similie_amd._lib", fromlist=["SilError"])
    with pytest.raises(L.SilError, match=match):
        IO.parse_msh2(text)


def test_rejections(IO):
    rng = np.random.default_rng(5)
    coords = [axis(4, rng)] * 3
    tags = np.ones(27, dtype=int)
    good = write_msh2(coords, tags, rng, 3)
    reject(IO, good.replace("2.2 0 8", "4.1 0 8"), "only Gmsh msh2 ASCII")
    reject(IO, good.replace("2.2 0 8", "2.2 1 8"), "only Gmsh msh2 ASCII")
    # a missing cell, a missing node, a cell spanning two grid cells, a duplicated cell
    reject(IO, write_msh2(coords, tags, rng, 3, tweak=lambda n, c: (n, [e for e in c if e[0] != 5][:0] +
                                                                     [e for e in c if e[0] == 5][1:] +
                                                                     [e for e in c if e[0] != 5])),
           "do not cover a full rectilinear grid")
    reject(IO, write_msh2(coords, tags, rng, 3, tweak=lambda n, c: (n[1:], c)),
           "do not form a full rectilinear grid|unknown node")

    def stretch(n, c):
        # the cell at the origin (the only one holding node 0, tag 7) gets the
        # far corner of the grid (node 63, tag 196) instead of node 0
        hexes = [e for e in c if e[0] == 5]
        other = [e for e in c if e[0] != 5]
        first = next(e for e in hexes if 7 in e[2])
        ids = [196 if t == 7 else t for t in first[2]]
        return n, other + [(5, first[1], ids)] + [e for e in hexes if e is not first]

    reject(IO, write_msh2(coords, tags, rng, 3, tweak=stretch), "does not match a single rectilinear grid cell")

    def dup(n, c):
        hexes = [e for e in c if e[0] == 5]
        return n, hexes[:-1] + [hexes[0]]

    reject(IO, write_msh2(coords, tags, rng, 3, tweak=dup), "duplicated hexahedron")
    # mixed topology and unsupported elements
    reject(IO, write_msh2(coords, tags, rng, 3, tweak=lambda n, c: (n, c + [(3, [1, 1], [n[0][0], n[1][0],
                                                                                           n[2][0], n[3][0]])])),
           "quadrilaterals or hexahedra")
    reject(IO, write_msh2(coords, tags, rng, 3, tweak=lambda n, c: (n, c + [(4, [1, 1], [n[0][0], n[1][0],
                                                                                           n[2][0], n[3][0]])])),
           "quadrilaterals or hexahedra")
    reject(IO, write_msh2(coords, tags, rng, 3, tweak=lambda n, c: (n, c + [(9, [1, 1], [n[0][0]] * 6)])),
           "unsupported Gmsh element type 9")
    # a node off the grid
    reject(IO, write_msh2(coords, tags, rng, 3, tweak=lambda n, c: ([(n[0][0], n[0][1] + 1e-3, n[0][2],
                                                                        n[0][3])] + n[1:], c)),
           "rectilinear|single rectilinear grid cell")


def test_quadrilaterals_off_one_plane_are_rejected(IO):
    rng = np.random.default_rng(6)
    text = write_msh2([axis(3, rng), axis(3, rng)], np.ones(4, dtype=int), rng, 2,
                      tweak=lambda n, c: ([(n[0][0], n[0][1], n[0][2], 1.0)] + n[1:], c))
    reject(IO, text, "single plane")


# ------------------------------------------------------------------ silpro ---

SILPRO = r'''
# a problem file
Problem {
  Name "An inductor \"test\"";   // trailing comment
  Physics NonLinearMagnetostatics;
  Solver MinimizeStrongFormulationResidual;
}

Solver {
  MaxIterations 50000;
  RelativeTolerance 1e-9;
  UseMatrixFree yes;
  Criterion PotentialAndMomentsTemporalDerivative;
  Preconditioner spd_isai;
  VectorPotentialGaugePenalty 1e6;
  RelativeTolerance 1e-3;
}

Preprocess {
  Name SingleElectricalConductorMaterialWithSingleLinearMagneticMaterial;
  NonlinearBHCurve EIcore;
  MagneticMaterialTags {
    Tag1 "Input/Core B";
    Tag0 "Input/Core A";
  }
}
'''


def test_silpro_tree_matches_the_oracle(IO):
    s = IO.Silpro.parse(SILPRO)
    tree = O.parse_silpro(SILPRO)
    assert s.value("Problem.Name") == tree["sections"]["Problem"]["values"]["Name"] == 'An inductor "test"'
    # the first of a repeated key wins
    assert s.value("Solver.RelativeTolerance") == "1e-9"
    assert s.values("Preprocess.MagneticMaterialTags") == list(
        tree["sections"]["Preprocess"]["sections"]["MagneticMaterialTags"]["values"][k]
        for k in sorted(tree["sections"]["Preprocess"]["sections"]["MagneticMaterialTags"]["values"]))
    assert s.values("Preprocess.MagneticMaterialTags") == ["Input/Core A", "Input/Core B"]
    assert s.values("PostProcess.DiagnosticRegionTags") == []
    assert s.value("Preprocess.Missing", default="x") == "x"


def test_silpro_settings_and_defaults(IO):
    st = IO.Silpro.parse(SILPRO).problem_settings()
    assert st["physics"] == "NonLinearMagnetostatics"
    assert st["max_iterations"] == 50000 and st["relative_tolerance"] == 1e-9
    assert st["use_matrix_free"] is True and st["criterion"] == "PotentialAndMomentsTemporalDerivative"
    assert st["preconditioner"] == "SpdIsai" and st["vector_potential_gauge_penalty"] == 1e6
    # MinimizeStrongFormulationResidualProblem defaults (onelab_interface.hpp:96-111)
    d = IO.Silpro.parse("Problem { } Solver { }").problem_settings()
    assert d == {"physics": "LinearMagnetostatics", "max_iterations": 10000, "relative_tolerance": 1e-10,
                 "jacobi_max_block_size": 1, "use_matrix_free": True, "criterion": "MomentsTemporalDerivative",
                 "preconditioner": "Jacobi", "vector_potential_gauge_penalty": 20.0, "sor_relaxation_factor": 1.2,
                 "chebyshev_iterations": 8, "chebyshev_lower_bound": 0.05, "chebyshev_upper_bound": 1.5,
                 "ir_iterations": 8, "ir_relaxation_factor": 1.0}
    ss = IO.Silpro.parse(SILPRO).solver_settings()
    assert ss.preconditioner == "spd_isai" and ss.max_iterations == 50000 and ss.relative_tolerance == 1e-9


@pytest.mark.parametrize("text, oracle_match, match", [
    ("Problem { Name x", "unexpected end", "line 1, column 17: the file ends before the ';' that ends 'Name'"),
    ("Problem { Name x }", "expected ';'", "line 1, column 18: the value of 'Name' must be followed by ';'"),
    ("Problem Name x; }", "expected '\\{'", "line 1, column 9: 'Problem' must be followed by '\\{'"),
    ('Problem { Name "open; }', "unterminated string", "line 1, column 16: a quoted string is never closed"),
    ("Problem {\n  Name x;\n", "unexpected end", "line 3, column 1: the file ends inside block 'Problem'"),
    ("{ }", "expected identifier", "found '\\{' where a block name should be"),
    ("Problem { ; }", "expected identifier", "found ';' where a key or a block name should be"),
    ("Problem { Name ; }", "expected identifier", "found ';' where the value of 'Name' should be"),
    ("Problem", "unexpected end", "the file ends after 'Problem', where '\\{' should open its block"),
])
def test_silpro_syntax_errors(IO, text, oracle_match, match):
    """The inputs the format rejects are rejected, with the position of the
    first offending symbol."""
    with pytest.raises(ValueError, match=oracle_match):
        O.parse_silpro(text)
    L = __import__("similie_amd._lib", fromlist=["SilError"])
    with pytest.raises(L.SilError, match=match):
        IO.Silpro.parse(text)


def test_silpro_lexical_corners_match_the_oracle(IO):
    """Corners of the format read the same way as the oracle's restatement:
    '#' and '"' inside a word, '//' ending a word, a backslash keeping any
    character inside a string, comments between every two symbols, repeated
    keys and repeated blocks (the first wins), nested blocks."""
    text = ('A{k1 v#1;k2 a"b; k3 word//comment\n;k4 "x\\\\y\\"z";#c\n}'
            ' B // c\n { k 1 ; k 2 ; C { t "u" ; } C { t v; d w; } }\n'
            'A { k1 other; k9 new; } B{k3 3;}')
    s = IO.Silpro.parse(text)
    tree = O.parse_silpro(text)
    a = tree["sections"]["A"]["values"]
    assert s.value("A.k1") == a["k1"] == "v#1"
    assert s.value("A.k2") == a["k2"] == 'a"b'
    assert s.value("A.k3") == a["k3"] == "word"
    assert s.value("A.k4") == a["k4"] == 'x\\y"z'
    assert s.value("A.k9", default="-") == "-" and "k9" not in a
    b = tree["sections"]["B"]
    assert s.value("B.k") == b["values"]["k"] == "1"
    assert s.value("B.C.t") == b["sections"]["C"]["values"]["t"] == "u"
    assert s.value("B.C.d", default="-") == "-" and "d" not in b["sections"]["C"]["values"]
    assert s.value("B.k3", default="-") == "-"
    assert s.values("A") == [a[k] for k in sorted(a)]


def test_reference_inductor_problem_file(IO):
    """examples/onelab_inductor/inductor.silpro, the reference's own problem
    file (a copy under tests/golden): physics, solver settings, tag lists."""
    path = Path(__file__).resolve().parent / "golden" / "inductor.silpro"
    s = IO.Silpro.read(path)
    st = s.problem_settings()
    assert st["physics"] == "NonLinearMagnetostatics"
    assert st["max_iterations"] == 100000 and st["relative_tolerance"] == 1e-8
    assert st["jacobi_max_block_size"] == 1 and st["use_matrix_free"] is True
    assert st["criterion"] == "MomentsTemporalDerivative" and st["preconditioner"] == "SpdIsai"
    assert st["vector_potential_gauge_penalty"] == 1e7
    ss = s.solver_settings()
    assert ss.preconditioner == "spd_isai" and ss.max_iterations == 100000 and ss.relative_tolerance == 1e-8
    assert s.value("Problem.Name") == "Structured inductor magnetostatics"
    assert s.value("Preprocess.Name") == "SingleElectricalConductorMaterialWithSingleLinearMagneticMaterial"
    assert s.value("Preprocess.NonlinearBHCurve") == "EIcore"
    assert s.value("Preprocess.CurrentDensityZ") == "Input/90SimiLie/0Coil current density magnitude z [A/m^2]"
    assert s.values("Preprocess.MagneticMaterialTags") == ["Input/90SimiLie/2E-core physical tag",
                                                           "Input/90SimiLie/3I-core physical tag"]
    assert s.values("Preprocess.PositiveElectricalConductorTags") == ["Input/90SimiLie/4Left coil physical tag"]
    assert s.values("Preprocess.NegativeElectricalConductorTags") == ["Input/90SimiLie/5Right coil physical tag"]
    assert s.values("Preprocess.PositiveElectricalConductorXTags") == [
        "Input/90SimiLie/8Positive x end-turn physical tag"]
    assert s.values("Preprocess.NegativeElectricalConductorXTags") == [
        "Input/90SimiLie/9Negative x end-turn physical tag"]
    assert s.value("PostProcess.Name") == "ForceDensityDiagnostics"
    assert s.values("PostProcess.DiagnosticRegionTags") == ["Input/90SimiLie/6Air-gap physical tag"]
    tree = O.parse_silpro(path.read_text())
    assert tree["sections"]["Solver"]["values"]["Preconditioner"] == "SpdIsai"


@pytest.mark.parametrize("text, match", [
    ("Solver { }", "the file has no 'Problem' block"),
    ("Problem { } Blocks { }", "the file has no 'Solver' block"),
    ("Problem { Physics Acoustics; } Solver { }", "Physics: 'Acoustics' is not supported"),
    ("Problem { Solver Newton; } Solver { }", "Problem.Solver: 'Newton' is not supported"),
    ("Problem { } Solver { UseMatrixFree maybe; }", "UseMatrixFree: 'maybe' is not a boolean"),
    ("Problem { } Solver { MaxIterations 12x; }", "MaxIterations: '12x' is not a non-negative integer"),
    ("Problem { } Solver { MaxIterations -3; }", "MaxIterations: '-3' is not a non-negative integer"),
    ("Problem { } Solver { RelativeTolerance 1e-8x; }", "RelativeTolerance: '1e-8x' is not a number"),
    ("Problem { } Solver { RelativeTolerance inf; }", "RelativeTolerance: 'inf' is not a number"),
    ("Problem { } Solver { Criterion Energy; }", "Criterion: 'Energy' is not supported"),
    ("Problem { } Solver { Preconditioner Multigrid; }", "'Multigrid' names no strong-formulation preconditioner"),
])
def test_silpro_setting_errors(IO, text, match):
    L = __import__("similie_amd._lib", fromlist=["SilError"])
    with pytest.raises(L.SilError, match=match):
        IO.Silpro.parse(text).problem_settings()


def test_unimplemented_preconditioners_are_reported(IO):
    L = __import__("similie_amd._lib", fromlist=["SilError"])
    with pytest.raises(L.SilError, match="not implemented"):
        IO.Silpro.parse("Problem { } Solver { Preconditioner Ssor; }").solver_settings()


def test_cell_inputs_rule(IO):
    rng = np.random.default_rng(7)
    tags = rng.integers(0, 9, 500)
    mu, cur, nl = IO.magnetostatics_cell_inputs(tags, 1.0, core_mu=1000.0, current_density_x=3.0,
                                                current_density_z=5.0, magnetic_material_tags=[1, 2],
                                                positive_x_tags=[3], negative_x_tags=[4], positive_z_tags=[5, 2],
                                                negative_z_tags=[6], nonlinear=True)
    core = np.isin(tags, [1, 2])
    np.testing.assert_array_equal(mu, np.where(core, 1000.0, 1.0))
    np.testing.assert_array_equal(nl, core.astype(float))
    np.testing.assert_array_equal(cur[:, 0], np.where(tags == 3, 3.0, np.where(tags == 4, -3.0, 0.0)))
    np.testing.assert_array_equal(cur[:, 2], np.where(tags == 5, 5.0, np.where(tags == 6, -5.0, 0.0)))
    np.testing.assert_array_equal(cur[:, 1], 0.0)
