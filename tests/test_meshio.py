"""Mesh ingest (dsv.meshio) against mesh files the HDF5 C library wrote in the reference's
layout (`domain/mesh_utils.py:155-267`; tests/golden/make_h5_fixtures.py): the arrays read back
bit for bit from every storage variant, an inline-XML XDMF, and the error paths (CPU)."""
import os

import numpy as np
import pytest

from dsv import meshio

H5 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "h5")
EXP = np.load(os.path.join(H5, "mesh_expected.npz"), allow_pickle=False)


@pytest.mark.parametrize("variant", ["default", "chunked", "latest", "bigendian"])
def test_reads_every_storage_variant_bit_exact(variant):
    f = meshio.H5File(os.path.join(H5, f"mesh_{variant}.h5"))
    pts, cells, data = f["/mesh/points"], f["mesh/cells"], f["/mesh/cell_data"]
    assert pts.dtype == np.float64 and cells.dtype == np.int64 and data.dtype == np.int32
    assert np.array_equal(pts, EXP["points"]) and np.array_equal(cells, EXP["cells"])
    assert np.array_equal(data, EXP["cell_data"])
    conn, xyz, tags = meshio.read_xdmf(os.path.join(H5, f"mesh_{variant}.xdmf"))
    assert conn.dtype == np.int32 and np.array_equal(conn, EXP["cells"])
    assert np.array_equal(xyz, EXP["points"]) and np.array_equal(tags, EXP["cell_data"])


def test_inline_xml_xdmf(tmp_path):
    xyz = EXP["points"][:5]
    conn = np.array([[0, 1, 2, 3], [1, 2, 3, 4]])
    txt = f"""<?xml version="1.0"?>
<Xdmf Version="3.0"><Domain><Grid Name="Grid" GridType="Uniform">
<Topology TopologyType="Tetrahedron" NumberOfElements="2">
<DataItem DataType="Int" Dimensions="2 4" Format="XML">{' '.join(map(str, conn.ravel()))}</DataItem>
</Topology><Geometry GeometryType="XYZ">
<DataItem DataType="Float" Dimensions="5 3" Format="XML" Precision="8">{' '.join(repr(float(x)) for x in xyz.ravel())}</DataItem>
</Geometry></Grid></Domain></Xdmf>"""
    p = tmp_path / "inline.xdmf"
    p.write_text(txt)
    c, x, tags = meshio.read_xdmf(str(p))
    assert np.array_equal(c, conn) and np.array_equal(x, xyz) and tags is None


def test_errors(tmp_path):
    bad = tmp_path / "bad.h5"
    bad.write_bytes(b"not an hdf5 file" * 64)
    with pytest.raises(ValueError):
        meshio.H5File(str(bad))
    f = meshio.H5File(os.path.join(H5, "mesh_default.h5"))
    with pytest.raises(KeyError):
        f["/mesh/nothing"]
    with pytest.raises(ValueError):
        f["/mesh"]   # a group, not a dataset
