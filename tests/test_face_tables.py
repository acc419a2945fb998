"""The FP32-bulk HLLC-Batten face list (csrc/hydra_faces.h, host code):
each face of a valid mesh is listed once, by the lower-numbered cell (a
boundary face by its cell); both cells of an interior face point to that
one entry, the higher-numbered one with the negation bit; a pair of cells
that does not share exactly one face keeps an entry per side.  The header
is compiled here with g++ into a small test library."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from hydraphase import mesh as M

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "multiphase-code_amd", "csrc")

DRIVER = r"""
#include "hydra_faces.h"
#include <cstring>
extern "C" int faces(int ne, const int *nc, int *fslot, int *fbase, int *flist)
{
    std::vector<int> a, b, c;
    hydra::build_face_tables(ne, nc, a, b, c);
    std::memcpy(fslot, a.data(), sizeof(int) * a.size());
    std::memcpy(fbase, b.data(), sizeof(int) * b.size());
    std::memcpy(flist, c.data(), sizeof(int) * c.size());
    return (int)c.size();
}
"""


@pytest.fixture(scope="module")
def faces(tmp_path_factory):
    d = tmp_path_factory.mktemp("faces")
    src, lib = d / "faces.cpp", d / "libfaces.so"
    src.write_text(DRIVER)
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", f"-I{CSRC}", str(src), "-o", str(lib)],
                   check=True)
    L = ctypes.CDLL(str(lib))
    ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    L.faces.argtypes = [ctypes.c_int, ip, ip, ip, ip]

    def run(nc):
        nc = np.ascontiguousarray(nc, dtype=np.int32)
        ne = nc.shape[1]
        fslot = np.zeros(4 * ne, np.int32)
        fbase = np.zeros(ne + 1, np.int32)
        flist = np.zeros(4 * ne, np.int32)
        nf = L.faces(ne, nc, fslot, fbase, flist)
        return fslot.reshape(4, ne), fbase, flist[:nf]
    return run


def test_valid_mesh_lists_each_face_once(faces):
    m = M.box_mesh(5, 4, 3, jitter=0.1, seed=2)
    nc = np.asarray(m.nc)
    ne = m.neles
    fslot, fbase, flist = faces(nc)
    nbound = int((nc >= ne).sum())
    assert flist.size == (4 * ne + nbound) // 2
    # cell by cell, in order, each cell's own faces
    assert fbase[0] == 0 and fbase[-1] == flist.size and (np.diff(fbase) >= 0).all()
    owner = np.repeat(np.arange(ne), np.diff(fbase))
    assert np.array_equal(flist >> 2, owner)
    for i in range(ne):
        for f in range(4):
            j = nc[f, i]
            k, neg = fslot[f, i] >> 1, fslot[f, i] & 1
            if j >= ne or i < j:
                assert neg == 0 and flist[k] == 4 * i + f
            else:
                fb = int(np.flatnonzero(nc[:, j] == i)[0])
                assert neg == 1 and flist[k] == 4 * j + fb and fslot[fb, j] == k << 1


def test_degenerate_pairs_keep_an_entry_per_side(faces):
    # cells 0 and 1 joined by two faces, cell 2 naming itself; ids >= 3 are ghosts
    nc = np.array([[1, 0, 2], [1, 0, 3], [3, 3, 4], [4, 4, 5]], dtype=np.int32)
    fslot, fbase, flist = faces(nc)
    assert flist.size == 12
    assert ((fslot & 1) == 0).all()
    assert np.array_equal(np.sort(flist), np.arange(12))
