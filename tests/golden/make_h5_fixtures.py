"""Writes the mesh-file fixtures of tests/golden/h5/ (run from the repository root:
`python tests/golden/make_h5_fixtures.py`).

The reference reads and writes meshes as XDMF + HDF5 (`domain/mesh_utils.py:155-267`, written
with h5py): /mesh/points float64 (N x 3), /mesh/cells int64 (C x 4), /mesh/cell_data int32 (C),
and an XDMF file naming them.  h5py is not installed here; the HDF5 C library is (under
/opt/conda), so tests/golden/h5_writer.c writes the same layout with it, in four storage
variants (default = h5py's defaults, chunked, latest file format, big-endian types).  The
XDMF text follows the reference's _write_xdmf template.  mesh_expected.npz holds the arrays
written, for the reader's bit-exact check."""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "dissolve-2.0_amd"))

from dsv import meshes  # noqa: E402

CONDA = "/opt/conda"
VARIANTS = ("default", "chunked", "latest", "bigendian")

XDMF = """<?xml version="1.0"?>
<Xdmf Version="3.0">
  <Domain>
    <Grid Name="Grid" GridType="Uniform">
      <Topology TopologyType="Tetrahedron" NumberOfElements="{nc}">
        <DataItem DataType="Int" Dimensions="{nc} 4" Format="HDF">
          {h5}:/mesh/cells
        </DataItem>
      </Topology>
      <Geometry GeometryType="XYZ">
        <DataItem DataType="Float" Dimensions="{np} 3" Format="HDF" Precision="8">
          {h5}:/mesh/points
        </DataItem>
      </Geometry>
      <Attribute Name="region" AttributeType="Scalar" Center="Cell">
        <DataItem DataType="Int" Dimensions="{nc}" Format="HDF">
          {h5}:/mesh/cell_data
        </DataItem>
      </Attribute>
    </Grid>
  </Domain>
</Xdmf>
"""


def main():
    conn, xyz = meshes.kuhn_cube(3)
    rng = np.random.default_rng(4)
    xyz = meshes.jitter(xyz, 1 / 3, 0.2, rng, ~meshes.boundary_vertex_mask(xyz))
    region = np.where(xyz[conn].mean(1)[:, 2] > 0.5, 1, 2).astype(np.int32)
    cells = conn.astype(np.int64)
    out = os.path.join(HERE, "h5")
    os.makedirs(out, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "h5_writer")
        subprocess.run(["gcc", "-O2", f"-I{CONDA}/include", "-o", exe,
                        os.path.join(HERE, "h5_writer.c"), f"-L{CONDA}/lib", "-lhdf5",
                        f"-Wl,-rpath,{CONDA}/lib"], check=True)
        files = []
        for name, arr in (("points", xyz), ("cells", cells), ("data", region)):
            p = os.path.join(tmp, name + ".bin")
            np.ascontiguousarray(arr).tofile(p)
            files.append(p)
        for k, v in enumerate(VARIANTS):
            h5 = f"mesh_{v}.h5"
            subprocess.run([exe, str(k), os.path.join(out, h5), *files, str(len(xyz)),
                            str(len(cells))], check=True)
            with open(os.path.join(out, f"mesh_{v}.xdmf"), "w") as f:
                f.write(XDMF.format(nc=len(cells), np=len(xyz), h5=h5))
    np.savez_compressed(os.path.join(out, "mesh_expected.npz"), points=xyz, cells=cells,
                        cell_data=region)
    print("wrote", [f"mesh_{v}" for v in VARIANTS], len(cells), "cells")


if __name__ == "__main__":
    main()
