"""Host logic of the transport schedule on meshes whose upwind graphs have
cycles (Kuhn cubes twisted about the vertical axis, cells positively
oriented; jittered Kuhn meshes stay acyclic while their cells are positively
oriented): gdn_sweep_schedule lags the edges that close cycles (class 2 -> 3)
and levels the rest.

Pinned by graph facts computed independently with scipy: every cyclic
strongly connected component of a velocity's upwind graph loses at least one
edge, every lagged edge lies inside such a component, and every remaining
inflow edge goes from a lower level to a higher one (so the remaining graph
is acyclic and the level order is a topological order)."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

import gdon3d as G
from helpers import kuhn_msh, twisted_msh


def schedule(path, kinetic):
    lib = G.load_library()
    m = G.Mesh(path=path)
    G.set_cavity_mode(1, 1, 1, (1.0, -1.0, 0.0))
    mdl = G.Model("cavity_mode", kinetic=kinetic)
    ne, nv = m.nelem, mdl.s.nb_v
    cp = lib.gdn_upwind_classes(C.byref(m.s), C.byref(mdl.s))
    cls0 = np.ctypeslib.as_array(cp, shape=(ne * 4 * nv,)).copy()
    cls = cls0.copy()
    nbr = m.arrays()["elem2elem"].astype(np.int32)
    items = np.zeros(nv * ne, np.int32)
    lptr = np.zeros(nv * ne + 1, np.int32)
    nl = C.c_int(0)
    f = lib.gdn_sweep_schedule
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.POINTER(C.c_int)]
    nlag = f(ne, nv, nbr.ctypes.data, cls.ctypes.data, items.ctypes.data,
             lptr.ctypes.data, C.byref(nl))
    return dict(ne=ne, nv=nv, nbr=nbr.reshape(ne, 4), cls0=cls0.reshape(ne, 4, nv),
                cls=cls.reshape(ne, 4, nv), items=items, lptr=lptr[:nl.value + 1],
                nlag=nlag)


def scc_labels(e2e, cls_v):
    """Strong components of the upwind graph e -> nb (outflow faces of e)."""
    ne = e2e.shape[0]
    src, f = np.nonzero((cls_v == 1) & (e2e >= 0))
    g = sp.csr_matrix((np.ones(len(src)), (src, e2e[src, f])), shape=(ne, ne))
    return connected_components(g, directed=True, connection="strong")[1]


@pytest.mark.parametrize("n,alpha,kinetic", [(3, 3.0, "d3q4"), (4, 4.0, "d3q4"),
                                             (3, 3.0, "d3q6"), (5, 6.0, "d3q4")])
def test_lagged_edges_break_every_cycle(n, alpha, kinetic):
    s = schedule(twisted_msh(n, alpha), kinetic)
    ne, nv, e2e = s["ne"], s["nv"], s["nbr"]
    assert s["nlag"] > 0
    assert np.count_nonzero(s["cls"] == 3) == s["nlag"]
    changed = s["cls"] != s["cls0"]
    assert np.all(s["cls0"][changed] == 2) and np.all(s["cls"][changed] == 3)
    level = np.empty(nv * ne, np.int64)
    for L in range(len(s["lptr"]) - 1):
        level[s["items"][s["lptr"][L]:s["lptr"][L + 1]]] = L
    assert sorted(s["items"].tolist()) == list(range(nv * ne))
    for v in range(nv):
        lab = scc_labels(e2e, s["cls0"][:, :, v])
        sizes = np.bincount(lab)
        lev = level[v * ne:(v + 1) * ne]
        # every remaining interior inflow edge goes upwards in level
        e, f = np.nonzero((s["cls"][:, :, v] == 2) & (e2e >= 0))
        assert np.all(lev[e2e[e, f]] < lev[e])
        # lagged edges lie inside cyclic components, and each such component lost one
        e, f = np.nonzero(s["cls"][:, :, v] == 3)
        assert np.all(lab[e] == lab[e2e[e, f]])
        assert np.all(sizes[lab[e]] > 1)
        assert set(np.flatnonzero(sizes > 1)) <= set(lab[e].tolist())


def test_acyclic_mesh_has_no_lagged_edges():
    s = schedule(kuhn_msh(6, 0.2, 23), "d3q4")
    assert s["nlag"] == 0
    assert np.array_equal(s["cls"], s["cls0"])
