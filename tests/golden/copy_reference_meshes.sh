#!/bin/sh
# Copies the two gmsh v2.2 meshes the reference ships in data/mesh/ into this
# directory as test fixtures (data, not code): an unstructured tetrahedral
# mesh of the unit cube (1808 tets) and the same cube with a torus
# (15,086 tets). The GPU tests read these copies; /root/reference exists only
# where the fixtures are made.
set -e
here=$(dirname "$0")
cp /root/reference/data/mesh/t4_cube_8.msh "$here/t4_cube_8.msh"
cp /root/reference/data/mesh/t4_cube_8_torus_128.msh "$here/t4_cube_8_torus_128.msh"
