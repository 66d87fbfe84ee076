/*
 * Tetrahedral mesh: gmsh v2.2 ASCII reader, connectivity, element geometry.
 *
 * Behaviour follows the reference's src/mesh/mesh.c (reader :372-726,
 * connectivity :811-1268, geometry accessors :29-210) array for array; the
 * construction differs: faces and T10 nodes are matched with hash tables in
 * O(N) instead of the reference's qsort passes, producing the same face
 * order, the same neighbour pairing and the same global T10 node numbering
 * (first appearance when scanning elements and local nodes in order, which is
 * what the reference's stable coordinate sort yields). The matching runs on
 * the host's threads (OpenMP): the keys are dealt into one partition per
 * thread by their hash, each thread scans every element in order and
 * handles the keys of its partition, so every key sees its occurrences in
 * the sequential order and the numbering is the sequential one (prefix
 * sums where it counts); the arrays are bit-identical for any thread count.
 */
#define _GNU_SOURCE
#include <assert.h>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mesh/mesh.h>
#include <mesh/t10.h>
#include <mesh/tetrageometry.h>

#ifdef _OPENMP
#include <omp.h>
static int nparts(void)
{
	return omp_get_max_threads();
}
static int part_id(void)
{
	return omp_get_thread_num();
}
#else
static int nparts(void)
{
	return 1;
}
static int part_id(void)
{
	return 0;
}
#endif

static void die(const char *what)
{
	fprintf(stderr, "gdon3d mesh: %s\n", what);
	abort();
}

static void *xcalloc(size_t n, size_t sz)
{
	void *p = calloc(n ? n : 1, sz);
	if (!p)
		die("out of memory");
	return p;
}

/* ------------------------------------------------------------ accessors */

gdn_real mesh_get_mass_ref_ij(const gdn_mesh *msh, const int i, const int j)
{
	(void)msh;
	return t10_get_mass_ref_entry(i, j);
}

gdn_real mesh_get_inv_mass_ref_ij(const gdn_mesh *msh, const int i, const int j)
{
	(void)msh;
	return t10_get_inv_mass_ref_entry(i, j);
}

gdn_real mesh_get_internal_ref_ijk(const gdn_mesh *msh, const int i,
				   const int j, const int k)
{
	(void)msh;
	return t10_get_internal_ref_entry(i, j, k);
}

int mesh_get_nb_elem(const gdn_mesh *msh)
{
	return msh->nbelems;
}

int mesh_get_nb_face(const gdn_mesh *msh)
{
	return msh->nbfaces;
}

int mesh_get_nb_loc_qnode(const gdn_mesh *msh)
{
	(void)msh;
	return NQN;
}

int mesh_get_gauss_point_id(const gdn_mesh *msh, const int id_elem,
			    const int id_loc_node)
{
	(void)msh;
	return id_elem * NQN + id_loc_node;
}

void mesh_get_elem_node(const gdn_mesh *msh, const int ie,
			gdn_real (*t4_node)[3])
{
	for (int v = 0; v < NNO; v++) {
		const gdn_real *x = msh->node + 3 * msh->elem2qnode[ie * NQN + v];
		t4_node[v][0] = x[0];
		t4_node[v][1] = x[1];
		t4_node[v][2] = x[2];
	}
}

void mesh_get_elem_det(const gdn_mesh *msh, const int id_elem, gdn_real *det)
{
	gdn_real dtau[3][3], codtau[3][3], xref[3] = { 0, 0, 0 }, nodes[NNO][3];
	mesh_get_elem_node(msh, id_elem, nodes);
	gdn_tetra_ref2phy(nodes, xref, NULL, -1, NULL, dtau, codtau, NULL, NULL);
	*det = geometry_dot_product(dtau[0], codtau[0]);
}

void mesh_get_elem_codtau(const gdn_mesh *msh, const int id_elem,
			  gdn_real codtau[3][3])
{
	gdn_real dtau[3][3], xref[3] = { 0, 0, 0 }, nodes[NNO][3];
	mesh_get_elem_node(msh, id_elem, nodes);
	gdn_tetra_ref2phy(nodes, xref, NULL, -1, NULL, dtau, codtau, NULL, NULL);
}

void mesh_get_face_vn(const gdn_mesh *msh, const int id_face, gdn_real *vn)
{
	gdn_real dtau[3][3], codtau[3][3], xref[3] = { 0, 0, 0 }, nodes[NNO][3];
	const int ieL = msh->face2elem[E2F * id_face + 0];
	const int ifL = msh->face2elem[E2F * id_face + 1];
	mesh_get_elem_node(msh, ieL, nodes);
	gdn_tetra_ref2phy(nodes, xref, NULL, ifL, NULL, dtau, codtau, NULL, vn);
	/* the reference tests the global face id here (src/mesh/mesh.c:120) */
	const gdn_real cr = (id_face == 3) ? sqrt(3) : 1;
	for (int d = 0; d < 3; d++)
		vn[d] *= cr;
}

int mesh_get_ipg_elem_right_from_face(const gdn_mesh *msh, const int id_face,
				      const int a)
{
	const int ieR = msh->face2elem[E2F * id_face + 2];
	return ieR * NQN + msh->fnode2enode[id_face * 2 * NQV + NQV + a];
}

int mesh_get_ipg_elem_left_from_face(const gdn_mesh *msh, const int id_face,
				     const int a)
{
	const int ieL = msh->face2elem[E2F * id_face + 0];
	return ieL * NQN + msh->fnode2enode[id_face * 2 * NQV + a];
}

bool mesh_get_bool_is_boundary_face(const gdn_mesh *msh, const int id_face)
{
	return msh->face2elem[E2F * id_face + 2] < 0;
}

/* src/mesh/mesh.c:149-210: hmin = min_e vol_e / surf_e, with the reference's
 * consistency check of the Jacobian and Heron measures (it aborts on
 * negatively oriented elements). */
gdn_real mesh_get_t4_hmin(const gdn_mesh *msh)
{
	gdn_real hmin = FLT_MAX;
	/* surface weight of each face: area of the reference face */
	const gdn_real face_w[4] = { 0.5, 0.5, 0.5, 0.86602540378443865 };

	for (int ie = 0; ie < msh->nbelems; ie++) {
		gdn_real nodes[NNO][3], xref[3] = { 0, 0, 0 }, dtau[3][3],
			codtau[3][3], vnds[3];
		mesh_get_elem_node(msh, ie, nodes);
		gdn_tetra_ref2phy(nodes, xref, NULL, -1, NULL, dtau, codtau, NULL,
				  NULL);
		const gdn_real det = geometry_dot_product(dtau[0], codtau[0]);
		gdn_real vol = 0, surf = 0;
		for (int v = 0; v < NNO; v++)
			vol += t4_get_wpg_ref_entry(v) * det;
		for (int f = 0; f < NFA; f++) {
			gdn_tetra_ref2phy(nodes, xref, NULL, f, NULL, dtau, codtau,
					  NULL, vnds);
			surf += sqrt(vnds[0] * vnds[0] + vnds[1] * vnds[1] +
				     vnds[2] * vnds[2]) *
				face_w[f];
		}
		gdn_real surf2, vol2;
		geometry_get_tetra_area_and_vol(nodes, &surf2, &vol2);
		if (!(fabs(surf - surf2) < 1e-12) || !(fabs(vol - vol2) < 1e-12)) {
			fprintf(stderr,
				"gdon3d mesh: element %d is degenerate or negatively "
				"oriented (det=%g)\n",
				ie, det);
			abort();
		}
		const gdn_real h = vol / surf;
		if (h < hmin)
			hmin = h;
	}
	return hmin;
}

/* ----------------------------------------------------------------- free */

void gdn_gdn_mesh_free(gdn_mesh *m)
{
	free(m->boundaryface);
	free(m->macrointerface);
	free(m->phys_group_id);
	free(m->physical_group);
	free(m->elem2node);
	free(m->elem2qnode);
	free(m->elem2elem);
	free(m->face2elem);
	free(m->fnode2enode);
	free(m->node2elem);
	free(m->node);
	free(m->qnode);
	char name[1024];
	memcpy(name, m->file_name, sizeof(name));
	memset(m, 0, sizeof(*m));
	memcpy(m->file_name, name, sizeof(name));
}

int gdn_gdn_mesh_is_empty(gdn_mesh *m)
{
	return m->nbelems == 0;
}

/* --------------------------------------------------------------- reader */

typedef struct {
	FILE *f;
	char *buf;
	size_t cap;
} reader;

/* next line including its '\n'; NULL at end of file */
static const char *rd_line(reader *r)
{
	ssize_t n = getline(&r->buf, &r->cap, r->f);
	return n < 0 ? NULL : r->buf;
}

static void rd_expect(reader *r, const char *pattern)
{
	const char *l = rd_line(r);
	if (!l || strcmp(l, pattern) != 0) {
		fprintf(stderr, "gdon3d mesh: expected \"%.*s\"\n",
			(int)strcspn(pattern, "\n"), pattern);
		abort();
	}
}

void mesh_read_from_msh22_file(gdn_mesh *m, const char *filename)
{
	gdn_gdn_mesh_free(m);
	reader r = { fopen(filename, "r"), NULL, 0 };
	if (!r.f) {
		fprintf(stderr, "gdon3d mesh: cannot open %s\n", filename);
		abort();
	}
	snprintf(m->file_name, sizeof(m->file_name), "%s", filename);

	/* $MeshFormat may be preceded by anything (src/mesh/mesh.c:330-360) */
	const char *l;
	while ((l = rd_line(&r)) && strcmp(l, "$MeshFormat\n") != 0)
		;
	if (!l)
		die("no $MeshFormat block");
	l = rd_line(&r);
	if (!l)
		die("truncated $MeshFormat");
	char *end;
	const double version = strtod(l, &end);
	const long file_type = strtol(end, &end, 10);
	if (!(version > 2 && version < 3) || file_type != 0)
		die("only gmsh 2.x ASCII files are supported");
	rd_expect(&r, "$EndMeshFormat\n");

	l = rd_line(&r);
	if (l && strcmp(l, "$PhysicalNames\n") == 0) {
		m->has_physical_groups = true;
		l = rd_line(&r);
		const long nnames = l ? strtol(l, NULL, 10) : 0;
		for (long i = 0; i < nnames; i++)
			if (!rd_line(&r))
				die("truncated $PhysicalNames");
		rd_expect(&r, "$EndPhysicalNames\n");
		rd_expect(&r, "$Nodes\n");
	} else if (!l || strcmp(l, "$Nodes\n") != 0) {
		die("$PhysicalNames or $Nodes not located");
	}

	l = rd_line(&r);
	m->nbnodes = l ? (int)strtol(l, NULL, 10) : 0;
	m->node = xcalloc(3 * (size_t)m->nbnodes, sizeof(gdn_real));
	for (int i = 0; i < m->nbnodes; i++) {
		l = rd_line(&r);
		if (!l)
			die("truncated $Nodes");
		const long id = strtol(l, &end, 10);
		if (id != i + 1)
			die("nodes must be numbered 1..N consecutively");
		for (int d = 0; d < 3; d++)
			m->node[3 * i + d] = strtod(end, &end);
	}
	rd_expect(&r, "$EndNodes\n");
	rd_expect(&r, "$Elements\n");

	l = rd_line(&r);
	const int nall = l ? (int)strtol(l, NULL, 10) : 0;
	m->elem2node = xcalloc(NNO * (size_t)nall, sizeof(int));
	if (m->has_physical_groups) {
		m->physical_group = xcalloc((size_t)nall, sizeof(int));
		m->physical_group_is_alloc = true;
	}
	int ntet = 0;
	for (int i = 0; i < nall; i++) {
		l = rd_line(&r);
		if (!l)
			die("truncated $Elements");
		const long id = strtol(l, &end, 10);
		if (id != i + 1)
			die("elements must be numbered consecutively");
		const long type = strtol(end, &end, 10);
		if (type != ECODE)
			continue; /* only 4-node tetrahedra are kept */
		const long ntags = strtol(end, &end, 10);
		long tag0 = 0;
		for (long t = 0; t < ntags; t++) {
			const long tag = strtol(end, &end, 10);
			if (t == 0)
				tag0 = tag;
		}
		if (m->has_physical_groups)
			m->physical_group[ntet] = (int)tag0;
		for (int v = 0; v < NNO; v++) {
			const long nid = strtol(end, &end, 10);
			if (nid < 1 || nid > m->nbnodes)
				die("element references a missing node");
			m->elem2node[NNO * ntet + v] = (int)nid - 1;
		}
		ntet++;
	}
	rd_expect(&r, "$EndElements\n");
	m->nbelems = ntet;
	m->is_read = true;
	free(r.buf);
	fclose(r.f);
}

/* ---------------------------------------------------------- hash tables */

/* Open-addressing map from a key of up to 3 ints to an int. */
typedef struct {
	int *keys; /* 3 per slot, keys[3 s] = -1 when empty */
	int *vals;
	size_t mask;
} imap;

static void imap_init(imap *h, size_t n)
{
	size_t cap = 16;
	while (cap < 2 * n + 16)
		cap <<= 1;
	h->keys = malloc(3 * cap * sizeof(int));
	h->vals = malloc(cap * sizeof(int));
	if (!h->keys || !h->vals)
		die("out of memory");
	for (size_t s = 0; s < cap; s++)
		h->keys[3 * s] = -1;
	h->mask = cap - 1;
}

static void imap_free(imap *h)
{
	free(h->keys);
	free(h->vals);
}

static size_t imap_hash(int a, int b, int c)
{
	uint64_t x = (uint64_t)(uint32_t)a * 0x9E3779B97F4A7C15ull;
	x ^= (uint64_t)(uint32_t)b * 0xC2B2AE3D27D4EB4Full + (x >> 29);
	x ^= (uint64_t)(uint32_t)c * 0x165667B19E3779F9ull + (x >> 31);
	x ^= x >> 33;
	return (size_t)x;
}

/* partition (of np) of key (a,b,c): bits the slot index does not use */
static int imap_part(int a, int b, int c, int np)
{
	return (int)((imap_hash(a, b, c) >> 40) % (uint64_t)np);
}

/* returns the slot of (a,b,c); *found tells whether it was present */
static size_t imap_slot(imap *h, int a, int b, int c, int *found)
{
	size_t s = imap_hash(a, b, c) & h->mask;
	for (;;) {
		int *k = h->keys + 3 * s;
		if (k[0] == -1) {
			*found = 0;
			return s;
		}
		if (k[0] == a && k[1] == b && k[2] == c) {
			*found = 1;
			return s;
		}
		s = (s + 1) & h->mask;
	}
}

static void sort3(int v[3])
{
	int t;
	if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
	if (v[1] > v[2]) { t = v[1]; v[1] = v[2]; v[2] = t; }
	if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
}

/* --------------------------------------------------------- connectivity */

static void build_bounds(gdn_mesh *m)
{
	for (int d = 0; d < 3; d++)
		m->xmin[d] = m->xmax[d] = m->node[d];
	for (int i = 0; i < m->nbnodes; i++)
		for (int d = 0; d < 3; d++) {
			const gdn_real x = m->node[3 * i + d];
			if (x < m->xmin[d])
				m->xmin[d] = x;
			if (x > m->xmax[d])
				m->xmax[d] = x;
		}
}

/* elem2elem by face matching (src/mesh/mesh.c:811-873) */
static void build_elem2elem(gdn_mesh *m)
{
	const int ne = m->nbelems;
	m->elem2elem = xcalloc(NFA * (size_t)ne, sizeof(int));
	for (int i = 0; i < NFA * ne; i++)
		m->elem2elem[i] = -1;
	const int np = nparts();
	int nfaces = 0, bad = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(np) reduction(+ : nfaces, bad)
#endif
	{
		const int p = part_id();
		imap h;
		imap_init(&h, (size_t)NFA * ne / np + 16);
		for (int e = 0; e < ne; e++) {
			for (int f = 0; f < NFA; f++) {
				int v[3];
				for (int k = 0; k < NVE; k++)
					v[k] = m->elem2node[NNO * e +
							    t4_get_ref_face2node_entry(f, k)];
				sort3(v);
				if (imap_part(v[0], v[1], v[2], np) != p)
					continue;
				int found;
				const size_t s = imap_slot(&h, v[0], v[1], v[2], &found);
				if (!found) {
					memcpy(h.keys + 3 * s, v, sizeof(v));
					h.vals[s] = NFA * e + f;
					nfaces++;
				} else {
					const int other = h.vals[s];
					if (other < 0) {
						bad++;
						continue;
					}
					m->elem2elem[NFA * e + f] = other / NFA;
					m->elem2elem[other] = e;
					h.vals[s] = -1;
				}
			}
		}
		imap_free(&h);
	}
	if (bad)
		die("a face is shared by more than two elements");
	m->nbfaces = nfaces;
}

/* src/mesh/mesh.c:971-1016: face f owned by the higher element id */
static void build_face2elem(gdn_mesh *m)
{
	m->face2elem = xcalloc(E2F * (size_t)m->nbfaces, sizeof(int));
	int nf = 0;
	for (int e = 0; e < m->nbelems; e++) {
		for (int f = 0; f < NFA; f++) {
			const int e2 = m->elem2elem[NFA * e + f];
			if (e2 >= e)
				continue;
			int *rec = m->face2elem + E2F * nf;
			rec[0] = e;
			rec[1] = f;
			rec[2] = e2;
			rec[3] = -1;
			if (e2 >= 0) {
				for (int f2 = 0; f2 < NFA; f2++)
					if (m->elem2elem[NFA * e2 + f2] == e) {
						rec[3] = f2;
						break;
					}
				assert(rec[3] != -1);
			}
			nf++;
		}
	}
	assert(nf == m->nbfaces);
}

/* src/mesh/mesh.c:875-929 */
static void build_face_lists(gdn_mesh *m)
{
	int nb = 0;
	for (int f = 0; f < m->nbfaces; f++)
		nb += m->face2elem[E2F * f + 2] == -1;
	m->nboundaryfaces = nb;
	m->nmacrointerfaces = m->nbfaces - nb;
	m->boundaryface = nb ? xcalloc((size_t)nb, sizeof(int)) : NULL;
	m->macrointerface =
		m->nmacrointerfaces ?
			xcalloc((size_t)m->nmacrointerfaces, sizeof(int)) :
			NULL;
	int ib = 0, ii = 0;
	for (int f = 0; f < m->nbfaces; f++) {
		if (m->face2elem[E2F * f + 2] == -1)
			m->boundaryface[ib++] = f;
		else
			m->macrointerface[ii++] = f;
	}
}

/* T10 nodes, numbered by first appearance (src/mesh/mesh.c:1018-1115);
 * coordinates are the affine images of the reference nodes. Position
 * pos = NQN e + i: each partition records, per position, the first position
 * of its key; the first positions, counted in order, give the numbers. */
static void build_qnodes(gdn_mesh *m)
{
	static const int edge_v[6][2] = { { 0, 1 }, { 1, 2 }, { 0, 2 },
					   { 0, 3 }, { 1, 3 }, { 2, 3 } };
	const int ne = m->nbelems;
	const size_t npos = (size_t)NQN * ne;
	int *first = xcalloc(npos, sizeof(int)); /* the key's first position */
	int *id = xcalloc(npos, sizeof(int)); /* 1 at first positions, then numbers */
	const int np = nparts();
#ifdef _OPENMP
#pragma omp parallel num_threads(np)
#endif
	{
		const int p = part_id();
		imap h;
		imap_init(&h, npos / np + 16);
		for (int e = 0; e < ne; e++) {
			const int *vid = m->elem2node + NNO * e;
			for (int i = 0; i < NQN; i++) {
				int key[3];
				if (i < NNO) {
					key[0] = vid[i];
					key[1] = key[2] = -2;
				} else {
					int a = vid[edge_v[i - NNO][0]], b = vid[edge_v[i - NNO][1]];
					key[0] = a < b ? a : b;
					key[1] = a < b ? b : a;
					key[2] = -3;
				}
				if (imap_part(key[0], key[1], key[2], np) != p)
					continue;
				const int pos = NQN * e + i;
				int found;
				const size_t s = imap_slot(&h, key[0], key[1], key[2], &found);
				if (!found) {
					memcpy(h.keys + 3 * s, key, sizeof(key));
					h.vals[s] = pos;
					id[pos] = 1;
				}
				first[pos] = h.vals[s];
			}
		}
		imap_free(&h);
	}
	int nq = 0; /* numbers of the first positions, in order */
	for (size_t pos = 0; pos < npos; pos++) {
		const int is_first = id[pos];
		id[pos] = nq;
		nq += is_first;
	}
	m->elem2qnode = xcalloc(npos, sizeof(int));
	gdn_real *xq = xcalloc(3 * (size_t)(nq > 0 ? nq : 1), sizeof(gdn_real));
#ifdef _OPENMP
#pragma omp parallel for num_threads(np) schedule(static)
#endif
	for (int e = 0; e < ne; e++) {
		const int *vid = m->elem2node + NNO * e;
		for (int i = 0; i < NQN; i++) {
			const int pos = NQN * e + i;
			m->elem2qnode[pos] = id[first[pos]];
			if (first[pos] != pos)
				continue;
			gdn_real t4[NNO][3];
			for (int v = 0; v < NNO; v++)
				for (int d = 0; d < 3; d++)
					t4[v][d] = m->node[3 * vid[v] + d];
			gdn_real xr[3];
			t10_get_ref_node_coord(i, xr);
			gdn_tetra_ref2phy(t4, xr, NULL, 0, xq + 3 * (size_t)id[pos], NULL, NULL, NULL,
					  NULL);
		}
	}
	free(first);
	free(id);
	free(m->node);
	m->node = xq;
	m->nbqnodes_in = nq;
}

/* src/mesh/mesh.c:1122-1175 */
static void build_fnode2enode(gdn_mesh *m)
{
	const int ne = m->nbelems, nf = m->nbfaces;
	m->fnode2enode = xcalloc(2 * NQV * (size_t)nf, sizeof(int));
	m->elem2qnode = realloc(m->elem2qnode,
				((size_t)NQN * ne + (size_t)NQV * m->nboundaryfaces) *
					sizeof(int));
	if (!m->elem2qnode)
		die("out of memory");
	/* boundary face number of each face (its place in boundaryface) */
	int *bfn = xcalloc((size_t)(nf > 0 ? nf : 1), sizeof(int));
	int nbf = 0;
	for (int f = 0; f < nf; f++) {
		bfn[f] = nbf;
		nbf += m->face2elem[E2F * f + 2] < 0;
	}
	assert(nbf == m->nboundaryfaces);
	int bad = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nparts()) schedule(static) reduction(+ : bad)
#endif
	for (int f = 0; f < nf; f++) {
		const int ieL = m->face2elem[E2F * f + 0];
		const int ifL = m->face2elem[E2F * f + 1];
		const int ieR = m->face2elem[E2F * f + 2];
		for (int a = 0; a < NQV; a++) {
			const int iL = t10_get_ref_face2node_entry(ifL, a);
			m->fnode2enode[2 * NQV * f + a] = iL;
			const int q = m->elem2qnode[NQN * ieL + iL];
			if (ieR < 0) {
				m->fnode2enode[2 * NQV * f + NQV + a] = a;
				m->elem2qnode[NQN * ne + NQV * bfn[f] + a] = q;
			} else {
				int iR = 0;
				while (iR < NQN && m->elem2qnode[NQN * ieR + iR] != q)
					iR++;
				bad += iR == NQN;
				m->fnode2enode[2 * NQV * f + NQV + a] = iR;
			}
		}
	}
	free(bfn);
	if (bad)
		die("face nodes of neighbouring elements do not match");
}

/* src/mesh/mesh.c:1178-1231: node2elem rows of width max_node2elem,
 * -1 terminated; each thread fills the rows of its nodes (q mod np),
 * scanning the elements in order */
static void build_node2elem(gdn_mesh *m)
{
	const int nq = m->nbqnodes_in;
	const size_t npos = (size_t)NQN * m->nbelems;
	const int np = nparts();
	int *count = xcalloc((size_t)nq, sizeof(int));
	int mx = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(np) reduction(max : mx)
#endif
	{
		const int p = part_id();
		for (size_t i = 0; i < npos; i++) {
			const int q = m->elem2qnode[i];
			if (q % np == p && ++count[q] > mx)
				mx = count[q];
		}
	}
	m->max_node2elem = mx + 1;
	m->node2elem = malloc(((size_t)m->max_node2elem + 1) * nq * sizeof(int));
	if (!m->node2elem)
		die("out of memory");
	for (size_t i = 0; i < (size_t)m->max_node2elem * nq; i++)
		m->node2elem[i] = -1;
	memset(count, 0, (size_t)nq * sizeof(int));
#ifdef _OPENMP
#pragma omp parallel num_threads(np)
#endif
	{
		const int p = part_id();
		for (int e = 0; e < m->nbelems; e++)
			for (int i = 0; i < NQN; i++) {
				const int q = m->elem2qnode[NQN * e + i];
				if (q % np == p)
					m->node2elem[(size_t)m->max_node2elem * q + count[q]++] = e;
			}
	}
	free(count);
}

void mesh_build_connectivity(gdn_mesh *m)
{
	if (!m->is_read || m->is_built)
		die("mesh_build_connectivity needs a read, not yet built mesh");
	build_bounds(m);
	build_elem2elem(m);
	build_face2elem(m);
	build_face_lists(m);
	build_qnodes(m);
	build_fnode2enode(m);
	build_node2elem(m);
	free(m->elem2node);
	m->elem2node = NULL;
	m->is_built = true;
}

void gdn_gdn_mesh_print(gdn_mesh *m)
{
	printf("[mesh] %s: %d elements, %d T10 nodes, %d faces (%d boundary)\n",
	       m->file_name, m->nbelems, m->nbqnodes_in, m->nbfaces,
	       m->nboundaryfaces);
	for (int f = 0; f < m->nbfaces; f++)
		printf("\tface %-6d L %-6d loc%d  R %-6d loc%d\n", f,
		       m->face2elem[E2F * f], m->face2elem[E2F * f + 1],
		       m->face2elem[E2F * f + 2], m->face2elem[E2F * f + 3]);
}
