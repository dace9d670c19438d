"""ASCII VTU output (src/mesh/vtk_writer.rs:104-265): the reference's text
layout line for line, with its number formatting.

Numbers are written as Rust's `Display` writes an f64: the shortest decimal
digits that read back to the same double, in positional notation (never an
exponent), "NaN", "inf", "-inf"; a float with no fractional part has no
decimal point ("3", "-0").  Python's repr yields the same shortest digits."""
from __future__ import annotations

import math

import numpy as np


def rust_f64(v: float) -> str:
    """Rust's `format!("{}", v)` for an f64."""
    v = float(v)
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    sign = "-" if math.copysign(1.0, v) < 0 else ""
    r = repr(abs(v))
    mant, _, exp = r.partition("e")
    e = int(exp) if exp else 0
    if "." in mant:
        ip, fp = mant.split(".")
    else:
        ip, fp = mant, ""
    if fp == "0":
        fp = ""
    digits = (ip + fp).lstrip("0")
    point = len(ip) + e  # position of the decimal point within ip + fp
    if not digits:
        return sign + "0"
    # shift so the digit string has no leading zeros
    lead = len(ip + fp) - len((ip + fp).lstrip("0"))
    point -= lead
    digits = digits.rstrip("0") or "0"
    if point <= 0:
        s = "0." + "0" * (-point) + digits
    elif point >= len(digits):
        s = digits + "0" * (point - len(digits))
    else:
        s = digits[:point] + "." + digits[point:]
    return sign + s


def _vec_line(values) -> str:
    return "".join(rust_f64(x) + " " for x in values)


class VtkWriter:
    """VtkWriter (src/mesh/vtk_writer.rs)."""

    @staticmethod
    def write_vtu(mesh, filename: str):
        """VtkWriter::write_vtu (src/mesh/vtk_writer.rs:104-265): points,
        Tet10 cells (VTK type 24), then point and cell data when present."""
        nodes = np.asarray(mesh.nodes, dtype=np.float64)
        elems = np.asarray(mesh.elements, dtype=np.int64)
        ne = elems.shape[0]
        out = ['<?xml version="1.0"?>',
               '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian">',
               "  <UnstructuredGrid>",
               f'    <Piece NumberOfPoints="{nodes.shape[0]}" NumberOfCells="{ne}">',
               "      <Points>",
               '        <DataArray type="Float64" NumberOfComponents="3" format="ascii">']
        for p in nodes:
            out.append(f"          {rust_f64(p[0])} {rust_f64(p[1])} {rust_f64(p[2])}")
        out += ["        </DataArray>", "      </Points>", "      <Cells>",
                '        <DataArray type="Int32" Name="connectivity" format="ascii">',
                "          " + "".join(f"{int(i)} " for i in elems.reshape(-1)),
                "        </DataArray>",
                '        <DataArray type="Int32" Name="offsets" format="ascii">',
                "          " + "".join(f"{10 * i} " for i in range(1, ne + 1)),
                "        </DataArray>",
                '        <DataArray type="UInt8" Name="types" format="ascii">',
                "          " + "24 " * ne,
                "        </DataArray>",
                "      </Cells>"]
        for tag, fd in (("PointData", getattr(mesh, "field_data", None)),
                        ("CellData", getattr(mesh, "cell_data", None))):
            if fd is None or fd.is_empty():
                continue
            out.append(f"      <{tag}>")
            for name in fd.field_names():
                f = fd.get_field(name)
                out += [f'        <DataArray type="Float64" Name="{f.name}" format="ascii">',
                        "          " + _vec_line(f.data), "        </DataArray>"]
            for name in fd.vector_field_names():
                f = fd.get_vector_field(name)
                out += [f'        <DataArray type="Float64" Name="{f.name}" '
                        f'NumberOfComponents="3" format="ascii">',
                        "          " + _vec_line(f.data.reshape(-1)), "        </DataArray>"]
            out.append(f"      </{tag}>")
        out += ["    </Piece>", "  </UnstructuredGrid>", "</VTKFile>"]
        with open(filename, "w") as fh:
            fh.write("\n".join(out) + "\n")
