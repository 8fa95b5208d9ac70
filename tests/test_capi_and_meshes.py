"""CPU checks of the C-ABI library (loads, exports every declared symbol), the host shim's
loud failure without a GPU, and the mesh generators."""
import ctypes
import os
import re

import numpy as np
import pytest

from dsv import _lib, meshes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "dsv.h")).read()
    return sorted(set(re.findall(r"\b(dsv_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTS)


def test_loader_declares_signatures_and_last_error_is_empty():
    lib = _lib.load()
    assert lib.dsv_last_error() == b""


def test_setters_reject_a_null_mesh():
    """The option setters validate their mesh before touching any device state."""
    lib = _lib.load()
    for name, args in (("dsv_set_species_narrowing", (None, 1)),
                       ("dsv_set_operator_reuse", (None, 1)),
                       ("dsv_set_operator_form", (None, 1)),
                       ("dsv_set_solver_options", (None, 1e-3, 12, 8, 0))):
        assert getattr(lib, name)(*args) != 0, name
        assert b"NULL mesh" in lib.dsv_last_error(), name


def test_stats_struct_matches_the_header():
    """dsv_stats as the loader declares it: every field of include/dsv.h, in order."""
    txt = open(os.path.join(ROOT, "include", "dsv.h")).read()
    body = txt[txt.index("typedef struct dsv_stats {"):txt.index("} dsv_stats;")]
    fields = re.findall(r"\b(?:int32_t|double)\s+([a-z_]+)\s*(?:\[\d+\])?;", body)
    assert fields == [f[0] for f in _lib.Stats._fields_]


def test_mesh_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from dsv.core import Mesh
    conn, xyz = meshes.kuhn_cube(2)
    with pytest.raises(_lib.DsvLibraryError):
        Mesh(conn, xyz)


def test_kuhn_counts_and_volume():
    conn, xyz = meshes.kuhn_cube(20)
    assert conn.shape == (48000, 4) and xyz.shape == (9261, 3)
    vol = meshes.tet_signed_volume(conn, xyz)
    assert np.all(np.abs(vol) > 0)
    assert abs(np.abs(vol).sum() - 1.0) < 1e-12


def test_gyroid_small_and_compact():
    conn, xyz = meshes.gyroid_lattice(n=16)
    assert conn.min() == 0 and conn.max() == len(xyz) - 1
    used = np.zeros(len(xyz), bool)
    used[conn.ravel()] = True
    assert used.all()
    c = xyz[conn].mean(1)
    g = np.sin(c[:, 0]) * np.cos(c[:, 1]) + np.sin(c[:, 1]) * np.cos(c[:, 2]) + \
        np.sin(c[:, 2]) * np.cos(c[:, 0])
    assert np.all(g < 0.3)


def test_graded_disc_small():
    conn, xyz = meshes.graded_disc(seed=0, scale=6.0)
    q = meshes.tet_quality(conn, xyz)
    assert q.min() >= 0.05
    r = np.hypot(xyz[:, 0], xyz[:, 1])
    assert r.max() <= 1.0 + 1e-12
    vol = np.abs(meshes.tet_signed_volume(conn, xyz)).sum()
    assert 0.9 * np.pi * 0.2 < vol < 1.1 * np.pi * 0.2


def test_locality_order_is_a_renumbering():
    conn, xyz = meshes.kuhn_cube(6)
    c2, x2 = meshes.locality_order(conn, xyz)
    assert sorted(map(tuple, np.round(x2, 12))) == sorted(map(tuple, np.round(xyz, 12)))
    v1 = np.sort(np.abs(meshes.tet_signed_volume(conn, xyz)))
    v2 = np.sort(np.abs(meshes.tet_signed_volume(c2, x2)))
    np.testing.assert_allclose(v1, v2)


def test_solve_transport_drops_cache_entry_and_reraises(monkeypatch):
    """A failure inside the GPU branch is reported, the species' cached assembler is dropped
    (`numerics/solvers.py:264-271`) and the error propagates: there is no CPU fallback."""
    import types
    from dsv import solvers

    def boom(*a, **k):
        raise RuntimeError("device lost")

    monkeypatch.setattr(solvers, "_solve_gpu_asm", boom)
    solvers._GPU_ASM_CACHE["O2"] = {"asm": None}
    u = types.SimpleNamespace(x=types.SimpleNamespace(array=np.zeros(4),
                                                      scatter_forward=lambda: None))
    with pytest.raises(RuntimeError, match="device lost"):
        solvers.solve_transport("O2", None, None, u, use_gpu=True, de_field=object())
    assert "O2" not in solvers._GPU_ASM_CACHE
