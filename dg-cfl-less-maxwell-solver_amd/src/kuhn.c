/*
 * Synthetic Kuhn cube meshes and a gmsh v2.2 ASCII writer. No counterpart in
 * the reference, which only reads meshes; the writer emits exactly what the
 * reader of src/mesh/mesh.c:623-726 accepts ($MeshFormat 2.2 0 8, nodes and
 * elements numbered from 1, type-4 tetrahedra with 2 tags).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mesh/kuhn.h>
#include <mesh/t10.h>

uint64_t gdn_splitmix64(uint64_t *state)
{
	uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

double gdn_uniform01(uint64_t *state)
{
	return (double)(gdn_splitmix64(state) >> 11) * 0x1.0p-53;
}

/* axis permutations in lexicographic order and their parity */
static const int perms[6][3] = { { 0, 1, 2 }, { 0, 2, 1 }, { 1, 0, 2 },
				 { 1, 2, 0 }, { 2, 0, 1 }, { 2, 1, 0 } };
static const int odd[6] = { 0, 1, 1, 0, 0, 1 };

void mesh_kuhn_cube(gdn_mesh *m, int n, double jitter, uint64_t seed)
{
	gdn_gdn_mesh_free(m);
	const int nv1 = n + 1;
	const double h = 1.0 / n;
	m->nbnodes = nv1 * nv1 * nv1;
	m->node = calloc(3 * (size_t)m->nbnodes, sizeof(gdn_real));
	m->nbelems = 6 * n * n * n;
	m->elem2node = calloc(NNO * (size_t)m->nbelems, sizeof(int));
	if (!m->node || !m->elem2node) {
		fprintf(stderr, "gdon3d kuhn: out of memory\n");
		abort();
	}
	uint64_t st = seed;
	for (int k = 0; k <= n; k++)
		for (int j = 0; j <= n; j++)
			for (int i = 0; i <= n; i++) {
				const int id = i + nv1 * (j + nv1 * k);
				const int ijk[3] = { i, j, k };
				const int interior = i > 0 && i < n && j > 0 &&
						     j < n && k > 0 && k < n;
				for (int d = 0; d < 3; d++) {
					double x = ijk[d] * h;
					if (interior && jitter != 0.0)
						x += jitter * h *
						     (2.0 * gdn_uniform01(&st) - 1.0);
					m->node[3 * id + d] = x;
				}
			}
	int e = 0;
	for (int k = 0; k < n; k++)
		for (int j = 0; j < n; j++)
			for (int i = 0; i < n; i++)
				for (int p = 0; p < 6; p++) {
					int c[3] = { i, j, k };
					int vid[4];
					vid[0] = c[0] + nv1 * (c[1] + nv1 * c[2]);
					for (int s = 0; s < 3; s++) {
						c[perms[p][s]]++;
						vid[s + 1] =
							c[0] + nv1 * (c[1] + nv1 * c[2]);
					}
					if (odd[p]) {
						const int t = vid[2];
						vid[2] = vid[3];
						vid[3] = t;
					}
					memcpy(m->elem2node + NNO * e, vid, sizeof(vid));
					e++;
				}
	snprintf(m->file_name, sizeof(m->file_name), "kuhn_%d", n);
	m->is_read = true;
}

int mesh_write_msh22_file(const gdn_mesh *m, const char *filename)
{
	if (!m->elem2node)
		return -1;
	FILE *f = fopen(filename, "w");
	if (!f)
		return -1;
	fprintf(f, "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n",
		m->nbnodes);
	for (int i = 0; i < m->nbnodes; i++)
		fprintf(f, "%d %.17g %.17g %.17g\n", i + 1, m->node[3 * i],
			m->node[3 * i + 1], m->node[3 * i + 2]);
	fprintf(f, "$EndNodes\n$Elements\n%d\n", m->nbelems);
	for (int e = 0; e < m->nbelems; e++) {
		const int *v = m->elem2node + NNO * e;
		fprintf(f, "%d 4 2 1 1 %d %d %d %d\n", e + 1, v[0] + 1, v[1] + 1,
			v[2] + 1, v[3] + 1);
	}
	fprintf(f, "$EndElements\n");
	return fclose(f) == 0 ? 0 : -1;
}

int mesh_kuhn_cube_write_msh22(const char *filename, int n, double jitter,
			       uint64_t seed)
{
	gdn_mesh m;
	memset(&m, 0, sizeof(m));
	mesh_kuhn_cube(&m, n, jitter, seed);
	const int rc = mesh_write_msh22_file(&m, filename);
	gdn_gdn_mesh_free(&m);
	return rc;
}
