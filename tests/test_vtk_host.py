"""VTU output (src/mesh/vtk_writer.rs:104-265) and the field containers
(src/mesh/fields.rs).  Numbers follow Rust's f64 Display: shortest
round-trip digits in positional notation; the cases below are the ones
that differ from Python's repr (integers without ".0", no exponents, signed
zero, specials).  No compute entry point of the library is called."""
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import geo_simulator_amd as G
from geo_simulator_amd.vtk import rust_f64


@pytest.mark.parametrize("v,s", [
    (1.0, "1"), (0.1, "0.1"), (1e-7, "0.0000001"), (1e21, "1000000000000000000000"),
    (1.2345678901234568e17, "123456789012345680"), (-0.0, "-0"), (0.0, "0"),
    (2.5, "2.5"), (-1280.0, "-1280"), (1e16, "10000000000000000"), (1.2345e-5, "0.000012345"),
    (float("nan"), "NaN"), (float("inf"), "inf"), (-float("inf"), "-inf"),
    (5e-324, "0." + "0" * 323 + "5"),
    (1.7976931348623157e308, "17976931348623157" + "0" * 292),
])
def test_rust_display(v, s):
    assert rust_f64(v) == s


def test_rust_display_round_trips():
    rng = np.random.default_rng(1)
    for v in np.concatenate([rng.standard_normal(200) * 10.0 ** rng.integers(-30, 30, 200),
                             rng.uniform(-1e6, 1e6, 200)]):
        s = rust_f64(v)
        assert "e" not in s and float(s) == v


def test_write_vtu_layout(tmp_path):
    mesh = G.ImprovedMeshGenerator.generate_cube(2, 1, 1, 1000.0, 500.0, 600.0)
    n, ne = mesh.num_nodes(), mesh.num_elements()
    T = np.linspace(0.0, 1280.0, n)
    mesh.field_data.add_field(G.ScalarField("temperature", T))
    u = np.arange(3 * n, dtype=float) * 1e-6
    mesh.field_data.add_vector_field(G.VectorField.from_dof_vector("displacement", u))
    mesh.cell_data.add_field(G.ScalarField("log_k", np.full(ne, 0.5)))
    fn = tmp_path / "mesh.vtu"
    G.VtkWriter.write_vtu(mesh, str(fn))
    text = fn.read_text().split("\n")
    assert text[0] == '<?xml version="1.0"?>'
    assert text[3] == f'    <Piece NumberOfPoints="{n}" NumberOfCells="{ne}">'
    assert text[6] == "          " + " ".join(rust_f64(x) for x in mesh.nodes[0])
    root = ET.parse(str(fn)).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    pts = np.array(piece.find("Points/DataArray").text.split(), float).reshape(-1, 3)
    assert np.array_equal(pts, mesh.nodes)
    arrays = {a.get("Name"): a for a in piece.find("Cells")}
    conn = np.array(arrays["connectivity"].text.split(), int).reshape(-1, 10)
    assert np.array_equal(conn, mesh.elements)
    assert np.array_equal(np.array(arrays["offsets"].text.split(), int), 10 * np.arange(1, ne + 1))
    assert set(arrays["types"].text.split()) == {"24"}
    pd = {a.get("Name"): a for a in piece.find("PointData")}
    assert np.array_equal(np.array(pd["temperature"].text.split(), float), T)
    assert pd["displacement"].get("NumberOfComponents") == "3"
    assert np.array_equal(np.array(pd["displacement"].text.split(), float), u)
    cd = {a.get("Name"): a for a in piece.find("CellData")}
    assert np.array_equal(np.array(cd["log_k"].text.split(), float), np.full(ne, 0.5))
    # no fields: no PointData / CellData sections
    bare = G.ImprovedMeshGenerator.generate_cube(1, 1, 1, 1.0, 1.0, 1.0)
    G.VtkWriter.write_vtu(bare, str(tmp_path / "bare.vtu"))
    t = (tmp_path / "bare.vtu").read_text()
    assert "<PointData>" not in t and "<CellData>" not in t


def test_vector_field_from_dof_vector_checks_length():
    with pytest.raises(ValueError):
        G.VectorField.from_dof_vector("u", np.zeros(4))
    f = G.VectorField.from_dof_vector("u", np.arange(6.0))
    assert f.len() == 2 and np.array_equal(f.data[1], [3.0, 4.0, 5.0])
