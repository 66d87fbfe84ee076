"""Cluster plan of the transport sweep (host logic, no device).

The cluster sweep solves the items of one cluster in order inside one
wavefront and takes the traces of earlier in-cluster producers from LDS; a
persistent launch over the clusters in plan order cannot deadlock when every
producer outside a cluster lies in an earlier cluster. Checked on Kuhn
meshes (plain and jittered), on twisted meshes whose upwind graphs have
cycles (lagged edges are not dependencies), and on a two-instance union:
every item appears exactly once; a cluster holds at most 6 items of one
velocity and one instance in upwind-level order; every face code names the
true producer (in-cluster: an earlier member; outside: the header lists that
face of that item); every producer lies earlier in the same cluster or in an
earlier cluster. On Kuhn meshes the clusters are the n^3 cubes.
"""
import ctypes as C

import numpy as np
import pytest

import gdon3d as G
from helpers import kuhn_msh, twisted_msh


def mesh_graph(path, kinetic):
    lib = G.load_library()
    m = G.Mesh(path=path)
    G.set_cavity_mode(1, 1, 1, (1.0, -1.0, 0.0))
    mdl = G.Model("cavity_mode", kinetic=kinetic)
    ne, nv = m.nelem, mdl.s.nb_v
    cp = lib.gdn_upwind_classes(C.byref(m.s), C.byref(mdl.s))
    cls = np.ctypeslib.as_array(cp, shape=(ne * 4 * nv,)).copy()
    a = m.arrays()
    nbr = a["elem2elem"].astype(np.int32)
    x = a["node"].reshape(-1, 3)
    vtx = x[a["elem2qnode"][: 10 * ne].reshape(ne, 10)[:, :4]].astype(np.float32)
    m.free()
    return dict(ne=ne, nv=nv, nbr=nbr, cls=cls, vtx=np.ascontiguousarray(vtx.reshape(-1)))


def schedule(g):
    lib = G.load_library()
    ne, nv = g["ne"], g["nv"]
    cls = g["cls"].copy()
    items = np.zeros(nv * ne, np.int32)
    lptr = np.zeros(nv * ne + 1, np.int32)
    nl = C.c_int(0)
    f = lib.gdn_sweep_schedule
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.POINTER(C.c_int)]
    f(ne, nv, g["nbr"].ctypes.data, cls.ctypes.data, items.ctypes.data, lptr.ctypes.data,
      C.byref(nl))
    return cls, items, lptr[: nl.value + 1], nl.value


def plan(g, ranges=None):
    lib = G.load_library()
    ne, nv = g["ne"], g["nv"]
    cls, items, lptr, nl = schedule(g)
    head = np.zeros(4 * nv * ne, np.int32)
    perm = np.zeros(nv * ne, np.int32)
    code = np.zeros(nv * ne, np.int32)
    stats = np.zeros(2, np.int32)
    rg = np.zeros(1, np.int32) if ranges is None else np.asarray(ranges, np.int32)
    f = lib.gdh_cluster_plan
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 2 + [C.c_void_p, C.c_int] + \
        [C.c_void_p] * 3 + [C.c_int] + [C.c_void_p] * 4
    ncl = f(ne, nv, g["vtx"].ctypes.data, g["nbr"].ctypes.data, rg.ctypes.data,
            0 if ranges is None else len(ranges) - 1, cls.ctypes.data, items.ctypes.data,
            lptr.ctypes.data, nl, head.ctypes.data, perm.ctypes.data, code.ctypes.data,
            stats.ctypes.data)
    assert ncl > 0
    return dict(cls=cls.reshape(ne, 4, nv), items=items, lptr=lptr,
                head=head[: 4 * ncl].reshape(ncl, 4), perm=perm, code=code, stats=stats)


def check_plan(g, p, ranges=None):
    ne, nv = g["ne"], g["nv"]
    nbr = g["nbr"].reshape(ne, 4)
    cls, items, head, perm, code = p["cls"], p["items"], p["head"], p["perm"], p["code"]
    lev = np.zeros(nv * ne, np.int64)
    for L in range(len(p["lptr"]) - 1):
        lev[items[p["lptr"][L]:p["lptr"][L + 1]]] = L
    assert np.array_equal(np.sort(perm), np.arange(nv * ne))
    n = head[:, 1] & 255
    nx = (head[:, 1] >> 8) & 255
    assert n.min() >= 1 and n.max() <= 6 and nx.max() <= 8
    assert np.array_equal(head[:, 0], np.concatenate(([0], np.cumsum(n)[:-1])))
    ve = items[perm]  # v * ne + e of each cluster-order item
    clus = np.repeat(np.arange(len(head)), n)
    where = np.empty(nv * ne, np.int64)  # cluster of item (v, e)
    where[ve] = clus
    pos = np.empty(nv * ne, np.int64)  # cluster-order position of item (v, e)
    pos[ve] = np.arange(nv * ne)
    inst = None if ranges is None else np.searchsorted(ranges, np.arange(ne), side="right") - 1
    for k in range(len(head)):
        i0 = head[k, 0]
        mem = ve[i0:i0 + n[k]]
        v = mem // ne
        e = mem % ne
        assert np.all(v == v[0])
        assert np.all(np.diff(lev[mem]) >= 0)
        if inst is not None:
            assert np.all(inst[e] == inst[e[0]])
        xb = int(head[k, 2]) & 0xFFFFFFFF | (int(head[k, 3]) & 0xFFFFFFFF) << 32
        seen = []
        for j in range(n[k]):
            for f in range(4):
                nib = int(code[i0 + j] >> (4 * f)) & 15
                u = nbr[e[j], f]
                if u < 0 or cls[e[j], f, v[0]] != 2:
                    assert nib == 0
                    continue
                if nib >= 8:
                    assert nib - 8 < j and e[nib - 8] == u
                else:
                    assert (xb >> (8 * nib)) & 255 == 4 * j + f
                    assert where[v[0] * ne + u] < k
                    # a producer read from outside its cluster publishes
                    pu = pos[v[0] * ne + u]
                    assert code[pu] >> 16 & 1
                    seen.append(nib)
            # publish bit: some downstream neighbour lies in another cluster
            out = [nbr[e[j], f] for f in range(4)
                   if nbr[e[j], f] >= 0 and cls[e[j], f, v[0]] == 1]
            ext = any(where[v[0] * ne + u] != k for u in out)
            assert (code[i0 + j] >> 16 & 1) == int(ext)
        assert sorted(seen) == list(range(nx[k]))


@pytest.mark.parametrize("n,jit,kinetic", [(4, 0.0, "d3q4"), (5, 0.1, "d3q4"),
                                           (4, 0.2, "d3q6"), (6, 0.33, "d3q4")])
def test_cluster_plan_kuhn(n, jit, kinetic):
    g = mesh_graph(kuhn_msh(n, jit, 17), kinetic)
    p = plan(g)
    check_plan(g, p)
    if jit <= 0.1 and kinetic == "d3q4":
        # clusters are the n^3 cubes of every velocity
        assert len(p["head"]) == 4 * n ** 3
        assert p["stats"][1] == g["nv"] * g["ne"]


@pytest.mark.parametrize("n,alpha,kinetic", [(3, 3.0, "d3q4"), (4, 4.0, "d3q4"),
                                             (3, 3.0, "d3q6")])
def test_cluster_plan_with_cycles(n, alpha, kinetic):
    g = mesh_graph(twisted_msh(n, alpha), kinetic)
    p = plan(g)
    assert (p["cls"] == 3).any()
    check_plan(g, p)


def test_cluster_plan_union_of_instances():
    a = mesh_graph(kuhn_msh(3, 0.0, 0), "d3q4")
    b = mesh_graph(kuhn_msh(4, 0.1, 5), "d3q4")
    ne = a["ne"] + b["ne"]
    nb = np.where(b["nbr"] >= 0, b["nbr"] + a["ne"], b["nbr"])
    g = dict(ne=ne, nv=4, nbr=np.concatenate([a["nbr"], nb]).astype(np.int32),
             cls=np.concatenate([a["cls"], b["cls"]]),
             vtx=np.concatenate([a["vtx"], b["vtx"]]))
    ranges = [0, a["ne"], ne]
    p = plan(g, ranges)
    check_plan(g, p, np.asarray(ranges))
    assert len(p["head"]) == 4 * (27 + 64)


@pytest.mark.parametrize("mesh,kinetic", [("t4_cube_8", "d3q4"), ("t4_cube_8", "d3q6"),
                                          ("t4_cube_8_torus_128", "d3q4")])
def test_cluster_plan_unstructured(mesh, kinetic):
    """The reference's unstructured meshes (fixtures in tests/golden): most
    clusters are split (grid cells cut across the flow), the invariants hold."""
    import os
    g = mesh_graph(os.path.join(os.path.dirname(__file__), "golden", mesh + ".msh"), kinetic)
    p = plan(g)
    check_plan(g, p)


_PLAN_CHECKSUM = r"""
import sys, hashlib
for p in sys.argv[1:4]:
    sys.path.insert(0, p)
import test_cpu_cluster_plan as T
g = T.mesh_graph(sys.argv[4], "d3q4")
p = T.plan(g)
h = hashlib.sha256()
h.update(g["cls"].tobytes())
for k in ("cls", "items", "lptr", "head", "perm", "code", "stats"):
    h.update(p[k].tobytes())
print(h.hexdigest())
"""


def test_setup_identical_for_any_thread_count():
    """The upwind classes (per face), the sweep schedule with its lagged
    cycle edges (per velocity) and the cluster plan (per velocity, face codes
    over ranges of clusters) run on the host's threads: each is the same bit
    for bit with 1, 3 and 8 threads, on a twisted mesh whose upwind graphs
    have cycles."""
    import os
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    root = os.path.dirname(here)
    paths = [os.path.join(root, "dg-cfl-less-maxwell-solver_amd"), os.path.join(root, "oracle"),
             here]
    msh = twisted_msh(8, 2.5)
    out = set()
    for t in ("1", "3", "8"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        r = subprocess.run([sys.executable, "-c", _PLAN_CHECKSUM] + paths + [msh], env=env,
                           check=True, capture_output=True, text=True)
        out.add(r.stdout.strip().splitlines()[-1])
    assert len(out) == 1
