"""Mesh and point-sample files through the library's host readers and writers (no device needed).

read_mesh: Thea's OFF/OBJ readers and face triangulation (Graphics/MeshCodecOff.hpp,
Graphics/MeshCodecObj.hpp, Algorithms/MeshTriangles.hpp:173-228). write_pts/read_pts: MeshSample's
point files (Tools/MeshSample/MeshSample.cpp:314-348, :560-601).
"""
import ctypes as C
import os

import numpy as np

from . import _capi as capi


def read_mesh(path):
    """(vertices float32 (nv, 3), triangles int32 (nt, 3), triangle_faces int64 (nt,), num_faces):
    the file's vertices, Thea's triangulation of its faces in face order, and each triangle's face
    index (the face's position in the file)."""
    m = capi.MeshFile()
    capi.check(capi.lib().thg_mesh_file_read(os.fsencode(path), C.byref(m)))
    try:
        nv, nt = m.num_vertices, m.num_triangles
        v = np.ctypeslib.as_array(m.vertices, shape=(max(nv, 1) * 3,))[:nv * 3].reshape(nv, 3).copy()
        t = np.ctypeslib.as_array(m.triangles, shape=(max(nt, 1) * 3,))[:nt * 3].reshape(nt, 3).copy()
        f = np.ctypeslib.as_array(m.triangle_faces, shape=(max(nt, 1),))[:nt].copy()
        return v, t, f, int(m.num_faces)
    finally:
        capi.lib().thg_mesh_file_free(C.byref(m))


def write_pts(path, positions, normals=None, ids=None):
    """MeshSample's point file: "x y z [nx ny nz] [id]" per line, C++ stream float format."""
    p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    n = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    i = None if ids is None else np.ascontiguousarray(ids, np.int64)
    for a in (n, i):
        if a is not None and len(a) != len(p):
            raise ValueError("one normal / id per point")
    capi.check(capi.lib().thg_pts_write(os.fsencode(path), len(p), p.ctypes.data_as(C.c_void_p),
                                        None if n is None else n.ctypes.data_as(C.c_void_p),
                                        None if i is None else i.ctypes.data_as(C.c_void_p)))


def read_pts(path):
    """The first three numbers of each non-empty, non-comment line, float32 (n, 3)."""
    n = C.c_int64()
    ptr = C.POINTER(C.c_float)()
    capi.check(capi.lib().thg_pts_read(os.fsencode(path), C.byref(n), C.byref(ptr)))
    try:
        return np.ctypeslib.as_array(ptr, shape=(max(n.value, 1) * 3,))[:n.value * 3].reshape(-1, 3).copy()
    finally:
        capi.lib().thg_free(ptr)
