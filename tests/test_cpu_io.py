"""Field output (include/io/io_hdf5.h): io_save_all writes the reference's
datasets (src/io/io_hdf5.c:60-368: connectivity_000000 = elem2qnode,
nodes_000000 = T10 node coordinates, f00000.. = each component at each
global node, from the first element of its node2elem list) as raw binary,
and an XDMF file pointing into them. Host only: the state is set on the host."""
import ctypes as C
import os
import re
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import gdon3d as G
from helpers import kuhn_msh


@pytest.mark.parametrize("kinetic", [None, "d3q4"])
def test_io_save_all_round_trip(tmp_path, kinetic):
    lib = G.load_library()
    lib.io_save_all.argtypes = [C.POINTER(G.GdnSimulation), C.c_char_p, C.c_int]
    mesh = G.Mesh(path=kuhn_msh(2, 0.1, 3))
    sim = G.Simulation(mesh, G.Model("plane_wave", kinetic=kinetic))
    wn = np.arange(sim.s.wlen, dtype=np.float64) * 0.25 + 1.0
    sim.set_wn(wn)
    base = str(tmp_path / "run")
    lib.io_save_all(C.byref(sim.s), base.encode(), 7)
    a = mesh.arrays()
    ne, m = mesh.nelem, sim.s.m
    conn = a["elem2qnode"][:10 * ne]
    npts = int(conn.max()) + 1
    raw = open(base + "_mesh.bin", "rb").read()
    assert np.frombuffer(raw[:4], np.int32)[0] == ne
    assert np.array_equal(np.frombuffer(raw[4:4 + 40 * ne], np.int32), conn)
    nodes = np.frombuffer(raw[4 + 40 * ne:], np.float64)
    assert nodes.size == 3 * npts
    data = np.fromfile(base + "_data_0007.bin", np.float64).reshape(m, npts)
    # the first element holding each node, its local index, the stored value
    first = {}
    for e in range(ne):
        for i in range(10):
            first.setdefault(int(conn[10 * e + i]), 10 * e + i)
    neq = sim.s.neq
    for node, ipg in list(first.items())[::7]:
        for c in range(m):
            ref = wn[c * neq + ipg] if kinetic else wn[ipg * m + c]
            assert data[c, node] == ref
    # node coordinates agree with the element nodes' coordinates
    xml = open(base + "_simulation_0007.xmf").read()
    root = ET.fromstring(re.sub(r"<!DOCTYPE[^>]*>", "", xml))
    topo = root.find(".//Topology")
    assert topo.get("TopologyType") == "Tet_10" and int(topo.get("NumberOfElements")) == ne
    atts = root.findall(".//Attribute")
    assert len(atts) == 6
    items = root.findall(".//DataItem[@Format='Binary']")
    assert all(os.path.exists(tmp_path / it.text.strip()) for it in items)
    sim.free()
