/*
 * ORACLE -- test infrastructure only.
 *
 * A plain C99 + OpenMP restatement of the reference gdon3d CPU algorithm
 * (p-gerhard/dg-cfl-less-maxwell-solver), used as the checker of the GPU
 * library and as the CPU baseline. Nothing in the product links or calls it;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * It follows the reference's assembled-matrix formulation step by step:
 *  - gmsh v2.2 reader and connectivity with the reference's qsort-based face
 *    and T10-node matching (src/mesh/mesh.c:372-1268);
 *  - COO assembly of inverse mass, volume and face terms in the reference's
 *    insertion order (src/simulation/simulation.c:218-604), compressed to CSR
 *    with duplicates summed at their first occurrence (src/maths/sparse.c:
 *    176-213, CSparse cs_compress + cs_dupl semantics);
 *  - CSR SpMV with per-row private accumulators (src/maths/sparse.c:387-402);
 *  - the Williamson LSRK3 loop (src/dg/explicit.c:39-117);
 *  - the theta = 1/2 scheme: A = M^-1 F per velocity (sparse product),
 *    M1 = I + theta dt A, M2 = I - (1-theta) dt A, sol = M1^-1 (M2 f)
 *    (src/dg/relaxation.c:24-127). The reference factors M1 with SuiteSparse
 *    KLU (BTF + per-block LU with partial pivoting, src/maths/sparse.c:
 *    411-466), which is absent here; this file restates that published
 *    method: strongly connected components of the matrix graph (Tarjan) in
 *    dependency order, dense LU with partial pivoting per component,
 *    block forward substitution;
 *  - node-local relaxation with the D3Q4/D3Q6 equilibria (src/dg/dg.c:
 *    109-147, src/models/d3q4.c, src/models/d3q6.c);
 *  - the nodal-quadrature L2 error (src/simulation/simulation.c:606-666).
 *
 * The P2 tables are recomputed here by exact polynomial integration in
 * double precision (independently of the library's generated tables); the
 * face mass is rounded to 10 significant digits as the reference stores it
 * (src/mesh/t10.c:1317-1330).
 */
#define _GNU_SOURCE
#include <assert.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NQN 10
#define NQV 6
#define NFA 4

static void die(const char *m)
{
	fprintf(stderr, "oracle: %s\n", m);
	abort();
}

static void *xc(size_t n, size_t s)
{
	void *p = calloc(n ? n : 1, s);
	if (!p)
		die("out of memory");
	return p;
}

static double now(void)
{
	struct timeval tv;
	gettimeofday(&tv, NULL);
	return tv.tv_sec + 1e-6 * tv.tv_usec;
}

/* =================================================================== */
/*                          P2 reference tables                        */
/* =================================================================== */

/* polynomial in x,y,z of degree <= 4: coefficient of x^a y^b z^c */
typedef struct {
	double c[5][5][5];
} poly;

static double fact(int n)
{
	double r = 1;
	for (int i = 2; i <= n; i++)
		r *= i;
	return r;
}

static void p_zero(poly *p)
{
	memset(p, 0, sizeof(*p));
}

static void p_mul(const poly *a, const poly *b, poly *r)
{
	p_zero(r);
	for (int i = 0; i < 5; i++)
		for (int j = 0; j < 5; j++)
			for (int k = 0; k < 5; k++) {
				if (a->c[i][j][k] == 0)
					continue;
				for (int l = 0; i + l < 5; l++)
					for (int m = 0; j + m < 5; m++)
						for (int n = 0; k + n < 5; n++)
							r->c[i + l][j + m][k + n] +=
								a->c[i][j][k] * b->c[l][m][n];
			}
}

static void p_der(const poly *a, int d, poly *r)
{
	p_zero(r);
	for (int i = 0; i < 5; i++)
		for (int j = 0; j < 5; j++)
			for (int k = 0; k < 5; k++) {
				int e[3] = { i, j, k };
				if (!e[d])
					continue;
				double v = a->c[i][j][k] * e[d];
				e[d]--;
				r->c[e[0]][e[1]][e[2]] += v;
			}
}

/* int over the unit tetrahedron: a! b! c! / (a+b+c+3)! */
static double p_int3(const poly *a)
{
	double s = 0;
	for (int i = 0; i < 5; i++)
		for (int j = 0; j < 5; j++)
			for (int k = 0; k < 5; k++)
				if (a->c[i][j][k] != 0)
					s += a->c[i][j][k] * fact(i) * fact(j) *
					     fact(k) / fact(i + j + k + 3);
	return s;
}

/* int over the unit triangle (z exponent 0): a! b! / (a+b+2)! */
static double p_int2(const poly *a)
{
	double s = 0;
	for (int i = 0; i < 5; i++)
		for (int j = 0; j < 5; j++)
			if (a->c[i][j][0] != 0)
				s += a->c[i][j][0] * fact(i) * fact(j) / fact(i + j + 2);
	return s;
}

/* linear polynomial c0 + cx x + cy y + cz z */
static void p_lin(poly *p, double c0, double cx, double cy, double cz)
{
	p_zero(p);
	p->c[0][0][0] = c0;
	p->c[1][0][0] = cx;
	p->c[0][1][0] = cy;
	p->c[0][0][1] = cz;
}

static double T_wpg[NQN], T_mass[NQN][NQN], T_imass[NQN][NQN];
static double T_int[NQN][NQN][3], T_face[NQV][NQV];
static int T_ready;

static const double ref_node[NQN][3] = { { 0, 0, 0 },	   { 1, 0, 0 },
					 { 0, 1, 0 },	   { 0, 0, 1 },
					 { 0.5, 0, 0 },	   { 0.5, 0.5, 0 },
					 { 0, 0.5, 0 },	   { 0, 0, 0.5 },
					 { 0.5, 0, 0.5 }, { 0, 0.5, 0.5 } };
static const int face2node[NFA][NQV] = { { 0, 3, 2, 7, 9, 6 },
					 { 0, 1, 3, 4, 8, 7 },
					 { 0, 2, 1, 6, 5, 4 },
					 { 1, 2, 3, 5, 9, 8 } };
static const int t4face[NFA][3] = { { 0, 3, 2 }, { 0, 1, 3 }, { 0, 2, 1 }, { 1, 2, 3 } };
#define NREF3 0.5773502691896258420811705
static const double ref_normal[NFA][3] = {
	{ -1, 0, 0 }, { 0, -1, 0 }, { 0, 0, -1 }, { NREF3, NREF3, NREF3 }
};

static void tables_init(void)
{
	if (T_ready)
		return;
	poly lam[4], phi[NQN], tri[6], tl[3], t1, t2;
	p_lin(&lam[0], 1, -1, -1, -1);
	p_lin(&lam[1], 0, 1, 0, 0);
	p_lin(&lam[2], 0, 0, 1, 0);
	p_lin(&lam[3], 0, 0, 0, 1);
	const int ev[6][2] = { { 0, 1 }, { 1, 2 }, { 0, 2 }, { 0, 3 }, { 1, 3 }, { 2, 3 } };
	for (int v = 0; v < 4; v++) { /* l (2 l - 1) */
		poly q = lam[v];
		for (int i = 0; i < 5; i++)
			for (int j = 0; j < 5; j++)
				for (int k = 0; k < 5; k++)
					q.c[i][j][k] *= 2;
		q.c[0][0][0] -= 1;
		p_mul(&lam[v], &q, &phi[v]);
	}
	for (int e = 0; e < 6; e++) {
		p_mul(&lam[ev[e][0]], &lam[ev[e][1]], &phi[4 + e]);
		for (int i = 0; i < 5; i++)
			for (int j = 0; j < 5; j++)
				for (int k = 0; k < 5; k++)
					phi[4 + e].c[i][j][k] *= 4;
	}
	for (int i = 0; i < NQN; i++) {
		T_wpg[i] = p_int3(&phi[i]);
		for (int j = 0; j < NQN; j++) {
			p_mul(&phi[i], &phi[j], &t1);
			T_mass[i][j] = p_int3(&t1);
			for (int d = 0; d < 3; d++) {
				p_der(&phi[i], d, &t2);
				p_mul(&t2, &phi[j], &t1);
				T_int[i][j][d] = p_int3(&t1);
			}
		}
	}
	/* inverse mass: Gauss-Jordan with partial pivoting */
	double a[NQN][2 * NQN];
	for (int i = 0; i < NQN; i++)
		for (int j = 0; j < NQN; j++) {
			a[i][j] = T_mass[i][j];
			a[i][NQN + j] = i == j;
		}
	for (int c = 0; c < NQN; c++) {
		int p = c;
		for (int r = c + 1; r < NQN; r++)
			if (fabs(a[r][c]) > fabs(a[p][c]))
				p = r;
		for (int k = 0; k < 2 * NQN; k++) {
			double t = a[c][k];
			a[c][k] = a[p][k];
			a[p][k] = t;
		}
		double d = a[c][c];
		for (int k = 0; k < 2 * NQN; k++)
			a[c][k] /= d;
		for (int r = 0; r < NQN; r++)
			if (r != c) {
				double l = a[r][c];
				for (int k = 0; k < 2 * NQN; k++)
					a[r][k] -= l * a[c][k];
			}
	}
	for (int i = 0; i < NQN; i++)
		for (int j = 0; j < NQN; j++)
			T_imass[i][j] = a[i][NQN + j];
	/* face mass of the P2 triangle, (v0,v1,v2,e01,e12,e20) */
	p_lin(&tl[0], 1, -1, -1, 0);
	p_lin(&tl[1], 0, 1, 0, 0);
	p_lin(&tl[2], 0, 0, 1, 0);
	for (int v = 0; v < 3; v++) {
		poly q = tl[v];
		for (int i = 0; i < 5; i++)
			for (int j = 0; j < 5; j++)
				q.c[i][j][0] *= 2;
		q.c[0][0][0] -= 1;
		p_mul(&tl[v], &q, &tri[v]);
	}
	for (int e = 0; e < 3; e++) {
		p_mul(&tl[e], &tl[(e + 1) % 3], &tri[3 + e]);
		for (int i = 0; i < 5; i++)
			for (int j = 0; j < 5; j++)
				tri[3 + e].c[i][j][0] *= 4;
	}
	for (int i = 0; i < NQV; i++)
		for (int j = 0; j < NQV; j++) {
			p_mul(&tri[i], &tri[j], &t1);
			double v = p_int2(&t1);
			char buf[64];
			snprintf(buf, sizeof(buf), "%.9e", v); /* 10 significant digits */
			T_face[i][j] = fabs(v) < 1e-12 ? 0.0 : strtod(buf, NULL);
		}
	T_ready = 1;
}

void orc_tables(double *wpg, double *mass, double *imass, double *internal,
		double *face)
{
	tables_init();
	memcpy(wpg, T_wpg, sizeof(T_wpg));
	memcpy(mass, T_mass, sizeof(T_mass));
	memcpy(imass, T_imass, sizeof(T_imass));
	memcpy(internal, T_int, sizeof(T_int));
	memcpy(face, T_face, sizeof(T_face));
}

/* =================================================================== */
/*                        mesh and connectivity                        */
/* =================================================================== */

typedef struct {
	int nelem, nnodes, nq, nfaces, nbf;
	int *e2n; /* 4 per element (T4), kept */
	double *xv; /* T4 vertex coordinates */
	int *e2e, *f2e, *bface, *e2q, *fn2en;
	double *xq; /* T10 node coordinates */
} omesh;

static char *getln(FILE *f, char **buf, size_t *cap)
{
	return getline(buf, cap, f) < 0 ? NULL : *buf;
}

/* gmsh 2.2 ASCII (src/mesh/mesh.c:372-726) */
static int read_msh(omesh *m, const char *path)
{
	FILE *f = fopen(path, "r");
	if (!f)
		return -1;
	char *b = NULL, *l, *end;
	size_t cap = 0;
	while ((l = getln(f, &b, &cap)) && strcmp(l, "$MeshFormat\n"))
		;
	if (!l)
		die("no $MeshFormat");
	l = getln(f, &b, &cap);
	double ver = strtod(l, &end);
	if (!(ver > 2 && ver < 3) || strtol(end, NULL, 10) != 0)
		die("unsupported format");
	getln(f, &b, &cap); /* $EndMeshFormat */
	l = getln(f, &b, &cap);
	if (!strcmp(l, "$PhysicalNames\n")) {
		int n = atoi(getln(f, &b, &cap));
		for (int i = 0; i < n + 2; i++)
			getln(f, &b, &cap); /* names, $End, $Nodes */
	}
	m->nnodes = atoi(getln(f, &b, &cap));
	m->xv = xc(3 * (size_t)m->nnodes, sizeof(double));
	for (int i = 0; i < m->nnodes; i++) {
		l = getln(f, &b, &cap);
		strtol(l, &end, 10);
		for (int d = 0; d < 3; d++)
			m->xv[3 * i + d] = strtod(end, &end);
	}
	getln(f, &b, &cap); /* $EndNodes */
	getln(f, &b, &cap); /* $Elements */
	int nall = atoi(getln(f, &b, &cap));
	m->e2n = xc(4 * (size_t)nall, sizeof(int));
	int ne = 0;
	for (int i = 0; i < nall; i++) {
		l = getln(f, &b, &cap);
		strtol(l, &end, 10);
		long type = strtol(end, &end, 10);
		if (type != 4)
			continue;
		long nt = strtol(end, &end, 10);
		for (long t = 0; t < nt; t++)
			strtol(end, &end, 10);
		for (int v = 0; v < 4; v++)
			m->e2n[4 * ne + v] = (int)strtol(end, &end, 10) - 1;
		ne++;
	}
	m->nelem = ne;
	free(b);
	fclose(f);
	return 0;
}

/* affine image of a reference point; Jacobian, cofactors, face normal
 * (src/mesh/tetrageometry.c:166-260, same order of operations) */
static void ref2phy(const double pn[4][3], const double xr[3], double xp[3],
		    double dtau[3][3], double cod[3][3], int ifa, double vn[3])
{
	const double g[4][4] = { { -1, -1, -1, 1 - xr[0] - xr[1] - xr[2] },
				 { 1, 0, 0, xr[0] },
				 { 0, 1, 0, xr[1] },
				 { 0, 0, 1, xr[2] } };
	if (xp)
		for (int d = 0; d < 3; d++) {
			xp[d] = 0;
			for (int i = 0; i < 4; i++)
				xp[d] += g[i][3] * pn[i][d];
		}
	if (!dtau)
		return;
	for (int r = 0; r < 3; r++) {
		for (int c = 0; c < 3; c++)
			dtau[r][c] = 0;
		for (int i = 0; i < 4; i++)
			for (int c = 0; c < 3; c++)
				dtau[r][c] += pn[i][r] * g[i][c];
	}
	cod[0][0] = dtau[1][1] * dtau[2][2] - dtau[1][2] * dtau[2][1];
	cod[0][1] = -dtau[1][0] * dtau[2][2] + dtau[1][2] * dtau[2][0];
	cod[0][2] = dtau[1][0] * dtau[2][1] - dtau[1][1] * dtau[2][0];
	cod[1][0] = -dtau[0][1] * dtau[2][2] + dtau[0][2] * dtau[2][1];
	cod[1][1] = dtau[0][0] * dtau[2][2] - dtau[0][2] * dtau[2][0];
	cod[1][2] = -dtau[0][0] * dtau[2][1] + dtau[0][1] * dtau[2][0];
	cod[2][0] = dtau[0][1] * dtau[1][2] - dtau[0][2] * dtau[1][1];
	cod[2][1] = -dtau[0][0] * dtau[1][2] + dtau[0][2] * dtau[1][0];
	cod[2][2] = dtau[0][0] * dtau[1][1] - dtau[0][1] * dtau[1][0];
	if (vn && ifa >= 0)
		for (int r = 0; r < 3; r++) {
			vn[r] = 0;
			for (int c = 0; c < 3; c++)
				vn[r] += cod[r][c] * ref_normal[ifa][c];
		}
}

typedef struct {
	int node[3], left, locfaceleft;
} face3;

static int cmp_face(const void *a, const void *b)
{
	const face3 *f1 = a, *f2 = b;
	int r = f1->node[0] - f2->node[0];
	if (!r)
		r = f1->node[1] - f2->node[1];
	if (!r)
		r = f1->node[2] - f2->node[2];
	return r;
}

static int cmp_int(const void *a, const void *b)
{
	return *(const int *)a - *(const int *)b;
}

typedef struct {
	double x[3];
	int id_elem, id_node, rank_e, rank_c;
} qnode;

static int cmp_real(double a, double b)
{
	return a < b ? -1 : a > b ? 1 : 0;
}

static int cmp_coord(const void *a, const void *b)
{
	const qnode *n1 = a, *n2 = b;
	int r = cmp_real(n1->x[0], n2->x[0]);
	if (!r)
		r = cmp_real(n1->x[1], n2->x[1]);
	if (!r)
		r = cmp_real(n1->x[2], n2->x[2]);
	return r;
}

static int cmp_elem(const void *a, const void *b)
{
	const qnode *n1 = a, *n2 = b;
	int r = n1->id_elem - n2->id_elem;
	return r ? r : n1->id_node - n2->id_node;
}

/* src/mesh/mesh.c:811-1175 */
static void build(omesh *m)
{
	const int ne = m->nelem;
	face3 *fc = xc(4 * (size_t)ne, sizeof(face3));
	for (int e = 0; e < ne; e++)
		for (int f = 0; f < 4; f++) {
			face3 *p = fc + 4 * e + f;
			for (int k = 0; k < 3; k++)
				p->node[k] = m->e2n[4 * e + t4face[f][k]];
			p->left = e;
			p->locfaceleft = f;
			qsort(p->node, 3, sizeof(int), cmp_int);
		}
	qsort(fc, 4 * (size_t)ne, sizeof(face3), cmp_face);
	m->e2e = xc(4 * (size_t)ne, sizeof(int));
	for (int i = 0; i < 4 * ne; i++)
		m->e2e[i] = -1;
	m->nfaces = 0;
	for (int i = 0; i < 4 * ne; i++) {
		if (i != 4 * ne - 1 && !cmp_face(fc + i, fc + i + 1)) {
			m->e2e[4 * fc[i].left + fc[i].locfaceleft] = fc[i + 1].left;
			m->e2e[4 * fc[i + 1].left + fc[i + 1].locfaceleft] = fc[i].left;
			i++;
		}
		m->nfaces++;
	}
	free(fc);
	m->f2e = xc(4 * (size_t)m->nfaces, sizeof(int));
	int nf = 0;
	for (int e = 0; e < ne; e++)
		for (int f = 0; f < 4; f++) {
			int e2 = m->e2e[4 * e + f];
			if (e2 >= e)
				continue;
			int *r = m->f2e + 4 * nf++;
			r[0] = e;
			r[1] = f;
			r[2] = e2;
			r[3] = -1;
			if (e2 >= 0)
				for (int f2 = 0; f2 < 4; f2++)
					if (m->e2e[4 * e2 + f2] == e) {
						r[3] = f2;
						break;
					}
		}
	assert(nf == m->nfaces);
	m->nbf = 0;
	for (int f = 0; f < m->nfaces; f++)
		m->nbf += m->f2e[4 * f + 2] == -1;
	m->bface = xc((size_t)m->nbf, sizeof(int));
	for (int f = 0, k = 0; f < m->nfaces; f++)
		if (m->f2e[4 * f + 2] == -1)
			m->bface[k++] = f;

	/* T10 nodes (src/mesh/mesh.c:1018-1115) */
	const size_t nl = (size_t)NQN * ne;
	qnode *nd = xc(nl, sizeof(qnode));
	for (int e = 0; e < ne; e++) {
		double pn[4][3];
		for (int v = 0; v < 4; v++)
			for (int d = 0; d < 3; d++)
				pn[v][d] = m->xv[3 * m->e2n[4 * e + v] + d];
		for (int i = 0; i < NQN; i++) {
			qnode *q = nd + NQN * e + i;
			q->id_elem = e;
			q->id_node = i;
			q->rank_e = NQN * e + i;
			ref2phy(pn, ref_node[i], q->x, NULL, NULL, -1, NULL);
		}
	}
	qsort(nd, nl, sizeof(qnode), cmp_coord);
	qnode *set = xc(nl, sizeof(qnode));
	int iu = 0;
	set[0] = nd[0];
	for (size_t k = 1; k < nl; k++) {
		nd[k].rank_c = (int)k;
		if (cmp_coord(&set[iu], &nd[k]))
			set[++iu] = nd[k];
		nd[k].rank_e = -1;
	}
	m->nq = iu + 1;
	qsort(set, m->nq, sizeof(qnode), cmp_elem);
	for (int k = 0; k < m->nq; k++)
		nd[set[k].rank_c].rank_e = k;
	for (size_t k = 1, id = 0; k < nl; k++) {
		if (!cmp_coord(&nd[id], &nd[k]))
			nd[k].rank_e = nd[id].rank_e;
		else
			id = k;
	}
	qsort(nd, nl, sizeof(qnode), cmp_elem);
	m->xq = xc(3 * (size_t)m->nq, sizeof(double));
	for (int k = 0; k < m->nq; k++)
		memcpy(m->xq + 3 * k, set[k].x, sizeof(set[k].x));
	m->e2q = xc(nl + (size_t)NQV * m->nbf, sizeof(int));
	for (size_t k = 0; k < nl; k++)
		m->e2q[k] = nd[k].rank_e;
	free(nd);
	free(set);

	/* face node maps (src/mesh/mesh.c:1122-1175) */
	m->fn2en = xc(2 * NQV * (size_t)m->nfaces, sizeof(int));
	int bf = -1;
	for (int f = 0; f < m->nfaces; f++) {
		int eL = m->f2e[4 * f], fL = m->f2e[4 * f + 1], eR = m->f2e[4 * f + 2];
		if (eR < 0)
			bf++;
		for (int a = 0; a < NQV; a++) {
			int iL = face2node[fL][a];
			m->fn2en[2 * NQV * f + a] = iL;
			int q = m->e2q[NQN * eL + iL];
			if (eR < 0) {
				m->fn2en[2 * NQV * f + NQV + a] = a;
				m->e2q[nl + NQV * bf + a] = q;
			} else {
				int j = 0;
				while (j < NQN && m->e2q[NQN * eR + j] != q)
					j++;
				if (j == NQN)
					die("face nodes do not match");
				m->fn2en[2 * NQV * f + NQV + a] = j;
			}
		}
	}
}

static void elem_nodes(const omesh *m, int e, double pn[4][3])
{
	for (int v = 0; v < 4; v++)
		for (int d = 0; d < 3; d++)
			pn[v][d] = m->xq[3 * m->e2q[NQN * e + v] + d];
}

static void elem_geom(const omesh *m, int e, double cod[3][3], double *det)
{
	double pn[4][3], dt[3][3], xr[3] = { 0, 0, 0 };
	elem_nodes(m, e, pn);
	ref2phy(pn, xr, NULL, dt, cod, -1, NULL);
	*det = dt[0][0] * cod[0][0] + dt[0][1] * cod[0][1] + dt[0][2] * cod[0][2];
}

/* src/mesh/mesh.c:149-210 */
static double hmin_of(const omesh *m)
{
	double h = 3.4028234663852886e38; /* FLT_MAX */
	const double fw[4] = { 0.5, 0.5, 0.5, 0.86602540378443865 };
	for (int e = 0; e < m->nelem; e++) {
		double pn[4][3], dt[3][3], cod[3][3], vn[3], xr[3] = { 0, 0, 0 };
		elem_nodes(m, e, pn);
		ref2phy(pn, xr, NULL, dt, cod, -1, NULL);
		double det = dt[0][0] * cod[0][0] + dt[0][1] * cod[0][1] +
			     dt[0][2] * cod[0][2];
		double vol = 0, surf = 0;
		for (int v = 0; v < 4; v++)
			vol += (1.0 / 24.0) * det;
		for (int f = 0; f < 4; f++) {
			ref2phy(pn, xr, NULL, dt, cod, f, vn);
			surf += sqrt(vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2]) * fw[f];
		}
		if (vol / surf < h)
			h = vol / surf;
		if (!(det > 0))
			die("negatively oriented element");
	}
	return h;
}

void *orc_mesh_load(const char *path)
{
	tables_init();
	omesh *m = xc(1, sizeof(omesh));
	if (read_msh(m, path)) {
		free(m);
		return NULL;
	}
	build(m);
	return m;
}

/* counts: nelem, nnodes(T4), nq, nfaces, nbf */
void orc_mesh_info(const void *h, int *out)
{
	const omesh *m = h;
	out[0] = m->nelem;
	out[1] = m->nnodes;
	out[2] = m->nq;
	out[3] = m->nfaces;
	out[4] = m->nbf;
}

void orc_mesh_arrays(const void *h, int *f2e, int *fn2en, int *e2q, double *xq,
		     int *bface, int *e2e)
{
	const omesh *m = h;
	memcpy(f2e, m->f2e, sizeof(int) * 4 * (size_t)m->nfaces);
	memcpy(fn2en, m->fn2en, sizeof(int) * 12 * (size_t)m->nfaces);
	memcpy(e2q, m->e2q, sizeof(int) * ((size_t)NQN * m->nelem + NQV * (size_t)m->nbf));
	memcpy(xq, m->xq, sizeof(double) * 3 * (size_t)m->nq);
	memcpy(bface, m->bface, sizeof(int) * (size_t)m->nbf);
	memcpy(e2e, m->e2e, sizeof(int) * 4 * (size_t)m->nelem);
}

double orc_mesh_hmin(const void *h)
{
	return hmin_of(h);
}

void orc_mesh_free(void *h)
{
	omesh *m = h;
	if (!m)
		return;
	free(m->e2n);
	free(m->xv);
	free(m->e2e);
	free(m->f2e);
	free(m->bface);
	free(m->e2q);
	free(m->fn2en);
	free(m->xq);
	free(m);
}

/* =================================================================== */
/*                               physics                               */
/* =================================================================== */

/* upwind flux (src/models/maxwell.c:152-192) */
static void flux_upwind(const double *wL, const double *wR, const double *vn,
			double *fl)
{
	const double nx = vn[0], ny = vn[1], nz = vn[2];
	const double overr = 1.0 / (sqrt(nx * nx + ny * ny + nz * nz) + 1e-16);
	const double nxy = overr * nx * ny, nxz = overr * nx * nz,
		     nyz = overr * ny * nz, nxx = overr * nx * nx,
		     nyy = overr * ny * ny, nzz = overr * nz * nz;
	double Es[3], Hs[3], Ed[3], Hd[3];
	for (int d = 0; d < 3; d++) {
		Es[d] = 0.5 * (wR[d] + wL[d]);
		Hs[d] = 0.5 * (wR[3 + d] + wL[3 + d]);
		Ed[d] = 0.5 * (wR[d] - wL[d]);
		Hd[d] = 0.5 * (wR[3 + d] - wL[3 + d]);
	}
	fl[0] = nz * Hs[1] - ny * Hs[2] - (nyy + nzz) * Ed[0] + nxy * Ed[1] + nxz * Ed[2];
	fl[1] = -nz * Hs[0] + nx * Hs[2] + nxy * Ed[0] - (nxx + nzz) * Ed[1] + nyz * Ed[2];
	fl[2] = -nx * Hs[1] + ny * Hs[0] + nxz * Ed[0] + nyz * Ed[1] - (nxx + nyy) * Ed[2];
	fl[3] = -nz * Es[1] + ny * Es[2] - (nyy + nzz) * Hd[0] + nxy * Hd[1] + nxz * Hd[2];
	fl[4] = nz * Es[0] - nx * Es[2] + nxy * Hd[0] - (nxx + nzz) * Hd[1] + nyz * Hd[2];
	fl[5] = -ny * Es[0] + nx * Es[1] + nxz * Hd[0] + nyz * Hd[1] - (nxx + nyy) * Hd[2];
}

/* (-n x H_L, 0) (src/models/maxwell.c:205-223) */
static void flux_metal(const double *wL, const double *wR, const double *vn,
		       double *fl)
{
	(void)wR;
	fl[0] = vn[2] * wL[4] - vn[1] * wL[5];
	fl[1] = -vn[2] * wL[3] + vn[0] * wL[5];
	fl[2] = -vn[0] * wL[4] + vn[1] * wL[3];
	fl[3] = fl[4] = fl[5] = 0;
}

/* exact solutions: kind numbering and parameters as include/gdn_hip.h */
typedef struct {
	int kind;
	double p[16];
} oimposed;

#define PI10 0.3141592654e1
static double sg(double x) { return sin(2. * M_PI * (x - 0.5)); }
static double sf(double y, double z)
{
	return exp(-2 * M_PI * (y - 0.5)) * sin(2 * M_PI * (z - 0.5)) +
	       exp(-2 * M_PI * (z - 0.5)) * sin(2 * M_PI * (y - 0.5));
}
static double sd1(double y, double z)
{
	return -0.2e1 * PI10 * exp(-0.2e1 * PI10 * (y - 0.5e0)) * sin(0.2e1 * PI10 * (z - 0.5e0)) +
	       0.2e1 * exp(-0.2e1 * PI10 * (z - 0.5e0)) * PI10 * cos(0.2e1 * PI10 * (y - 0.5e0));
}
static double sd2(double y, double z)
{
	return 0.2e1 * exp(-0.2e1 * PI10 * (y - 0.5e0)) * PI10 * cos(0.2e1 * PI10 * (z - 0.5e0)) -
	       0.2e1 * exp(-0.2e1 * PI10 * (z - 0.5e0)) * PI10 * sin(0.2e1 * PI10 * (y - 0.5e0));
}

static void exact(const oimposed *ip, const double *x, double t, double *w)
{
	switch (ip->kind) {
	case 1: { /* src/models/maxwell.c:95-107 */
		double s = cos(5 * M_PI * (x[0] - t));
		w[0] = 0; w[1] = 0; w[2] = s; w[3] = 0; w[4] = -s; w[5] = 0;
		break;
	}
	case 2: /* src/models/maxwell.c:118-133 */
		w[0] = 0;
		w[1] = 0;
		w[2] = sin(PI10 * x[0]) * sin(PI10 * x[1]) * cos(sqrt(2.) * PI10 * t);
		w[3] = -sin(sqrt(2.) * PI10 * t) * sqrt(2.) * sin(PI10 * x[0]) * cos(PI10 * x[1]) / 2.;
		w[4] = sin(sqrt(2.) * PI10 * t) * sqrt(2.) * cos(PI10 * x[0]) * sin(PI10 * x[1]) / 2.;
		w[5] = 0;
		break;
	case 3: /* src/models/maxwell.c:72-84 */
		w[0] = sg(x[0]) + sf(x[1], x[2]);
		w[1] = sg(x[1]) + sf(x[2], x[0]);
		w[2] = sg(x[2]) + sf(x[0], x[1]);
		w[3] = (sd1(x[2], x[0]) - sd2(x[0], x[1])) * t;
		w[4] = (sd1(x[0], x[1]) - sd2(x[1], x[2])) * t;
		w[5] = (sd1(x[1], x[2]) - sd2(x[2], x[0])) * t;
		break;
	case 4: { /* plane wave: E = e s, H = h s, s = sin(k.x - om t + ph) */
		const double *k = ip->p, *e = ip->p + 3, *hh = ip->p + 6;
		double s = sin(k[0] * x[0] + k[1] * x[1] + k[2] * x[2] - ip->p[9] * t + ip->p[10]);
		for (int d = 0; d < 3; d++) {
			w[d] = e[d] * s;
			w[3 + d] = hh[d] * s;
		}
		break;
	}
	case 5: { /* PEC cube mode: E = a (c s s, s c s, s s c) cos(om t) */
		const double *k = ip->p, *a = ip->p + 3, om = ip->p[6];
		double s[3], c[3];
		for (int d = 0; d < 3; d++) {
			s[d] = sin(k[d] * x[d]);
			c[d] = cos(k[d] * x[d]);
		}
		double ct = cos(om * t), st = sin(om * t) / om;
		w[0] = a[0] * c[0] * s[1] * s[2] * ct;
		w[1] = a[1] * s[0] * c[1] * s[2] * ct;
		w[2] = a[2] * s[0] * s[1] * c[2] * ct;
		w[3] = -st * (k[1] * a[2] - k[2] * a[1]) * s[0] * c[1] * c[2];
		w[4] = -st * (k[2] * a[0] - k[0] * a[2]) * c[0] * s[1] * c[2];
		w[5] = -st * (k[0] * a[1] - k[1] * a[0]) * c[0] * c[1] * s[2];
		break;
	}
	case 6: { /* Gaussian pulse E_z = amp exp(-|x - c|^2 / (2 sigma^2)) */
		const double *c = ip->p;
		double r2 = 0;
		for (int d = 0; d < 3; d++)
			r2 += (x[d] - c[d]) * (x[d] - c[d]);
		for (int q = 0; q < 6; q++)
			w[q] = 0;
		w[2] = ip->p[4] * exp(-r2 / (2 * ip->p[3] * ip->p[3]));
		break;
	}
	default:
		die("unknown exact solution");
	}
}

/* kinetic sets (src/models/d3q4.c, src/models/d3q6.c) */
typedef struct {
	int kind, nv;
	double vi[24], lambda, omega;
} okin;

static void kin_set(okin *k, int kind)
{
	memset(k, 0, sizeof(*k));
	k->kind = kind;
	k->omega = 1.999999999999;
	if (kind == 1) {
		const double L = 1.732050807568877293;
		const double v[12] = { L, L, L, L, -L, -L, -L, L, -L, -L, -L, L };
		k->nv = 4;
		k->lambda = L;
		memcpy(k->vi, v, sizeof(v));
	} else if (kind == 2) {
		const double L = 3.;
		const double v[18] = { L, 0, 0, 0, L, 0, 0, 0, L, -L, 0, 0, 0, -L, 0, 0, 0, -L };
		k->nv = 6;
		k->lambda = L;
		memcpy(k->vi, v, sizeof(v));
	}
}

/* f_eq via the physical flux qM = flux(wf, wf, e_k) with wf = (D / eps, B)
 * the fields of the conserved state w = (D, B) (wf = w without permittivity) */
static void kin_feq_eps(const okin *k, const double *w, double ieps, double *feq)
{
	const double e[3][3] = { { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } };
	const double wf[6] = { w[0] * ieps, w[1] * ieps, w[2] * ieps, w[3], w[4], w[5] };
	double fl[3][6];
	for (int d = 0; d < 3; d++)
		flux_upwind(wf, wf, e[d], fl[d]);
	if (k->kind == 1) {
		const double l2 = k->lambda * k->lambda;
		for (int iw = 0; iw < 6; iw++)
			for (int iv = 0; iv < 4; iv++)
				feq[iw * 4 + iv] = (w[iw] + (fl[0][iw] * k->vi[3 * iv + 0] +
							    fl[1][iw] * k->vi[3 * iv + 1] +
							    fl[2][iw] * k->vi[3 * iv + 2]) / l2) / 4;
	} else {
		const double l = k->lambda, l2 = l * l;
		for (int iw = 0; iw < 6; iw++) {
			const double q0 = fl[0][iw], q1 = fl[1][iw], q2 = fl[2][iw];
			feq[iw * 6 + 0] = +w[iw] / 6. + q0 / (2. * l) + 0 / (6. * l2) + 0 / (6. * l2);
			feq[iw * 6 + 1] = +w[iw] / 6. + q1 / (2. * l) - 0 / (3. * l2) + 0 / (6. * l2);
			feq[iw * 6 + 2] = +w[iw] / 6. + q2 / (2. * l) + 0 / (6. * l2) - 0 / (3. * l2);
			feq[iw * 6 + 3] = +w[iw] / 6. - q0 / (2. * l) + 0 / (6. * l2) + 0 / (6. * l2);
			feq[iw * 6 + 4] = +w[iw] / 6. - q1 / (2. * l) - 0 / (3. * l2) + 0 / (6. * l2);
			feq[iw * 6 + 5] = +w[iw] / 6. - q2 / (2. * l) + 0 / (6. * l2) - 0 / (3. * l2);
		}
	}
}

/* Maxwell eigen-projectors (src/models/maxwell.c:234-575) from the 6x6
 * matrix A(m), m = n / |n|, A w = (-m x H, m x E): P^- = (A^2 - A) / 2 onto
 * the incoming characteristics, P^{+,0} = I - P^- */
static void proj_matrices(const double *n, double pneg[6][6], double ppos[6][6])
{
	const double r = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
	const double m[3] = { n[0] / r, n[1] / r, n[2] / r };
	const double N[3][3] = { { 0, -m[2], m[1] }, { m[2], 0, -m[0] }, { -m[1], m[0], 0 } };
	double A[6][6] = { { 0 } }, A2[6][6];
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++) {
			A[i][3 + j] = -N[i][j];
			A[3 + i][j] = N[i][j];
		}
	for (int i = 0; i < 6; i++)
		for (int j = 0; j < 6; j++) {
			A2[i][j] = 0;
			for (int k = 0; k < 6; k++)
				A2[i][j] += A[i][k] * A[k][j];
		}
	for (int i = 0; i < 6; i++)
		for (int j = 0; j < 6; j++) {
			pneg[i][j] = 0.5 * (A2[i][j] - A[i][j]);
			ppos[i][j] = (i == j ? 1.0 : 0.0) - pneg[i][j];
		}
}

void orc_proj(const double *w, const double *n, int neg, double *res)
{
	double pn[6][6], pp[6][6];
	proj_matrices(n, pn, pp);
	for (int i = 0; i < 6; i++) {
		res[i] = 0;
		for (int j = 0; j < 6; j++)
			res[i] += (neg ? pn[i][j] : pp[i][j]) * w[j];
	}
}

/* the velocity moments of f (src/models/d3q4.c:105-135,
 * src/models/d3q6.c:100-141): w, then per field nm = nv - 1 moments
 * X, Y, Z (D3Q6 also X^2 - Y^2, X^2 - Z^2) in q[nm iw + k] */
static void kin_moments(const okin *k, const double *f, double *w, double *q)
{
	const int nm = k->nv - 1;
	for (int iw = 0; iw < 6; iw++) {
		w[iw] = 0;
		for (int c = 0; c < nm; c++)
			q[nm * iw + c] = 0;
		for (int iv = 0; iv < k->nv; iv++) {
			const double fi = f[iw * k->nv + iv], *v = k->vi + 3 * iv;
			w[iw] += fi;
			for (int d = 0; d < 3; d++)
				q[nm * iw + d] += fi * v[d];
			if (nm == 5) {
				q[nm * iw + 3] += fi * (v[0] * v[0] - v[1] * v[1]);
				q[nm * iw + 4] += fi * (v[0] * v[0] - v[2] * v[2]);
			}
		}
	}
}

/* the physical flux in moments: A(e_d) w for X, Y, Z, 0 beyond */
static void kin_qM(const okin *k, const double *w, double *qM)
{
	const int nm = k->nv - 1;
	const double e[3][3] = { { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } };
	double fl[3][6];
	for (int d = 0; d < 3; d++)
		flux_upwind(w, w, e[d], fl[d]);
	for (int iw = 0; iw < 6; iw++)
		for (int c = 0; c < nm; c++)
			qM[nm * iw + c] = c < 3 ? fl[c][iw] : 0;
}

/* f from its moments (src/models/d3q4.c:137-157, src/models/d3q6.c:143-208) */
static void kin_from_moments(const okin *k, const double *w, const double *q, double *f)
{
	const double l = k->lambda, l2 = l * l;
	for (int iw = 0; iw < 6; iw++) {
		if (k->nv == 4) {
			for (int iv = 0; iv < 4; iv++)
				f[iw * 4 + iv] = (w[iw] + (q[3 * iw] * k->vi[3 * iv] +
							   q[3 * iw + 1] * k->vi[3 * iv + 1] +
							   q[3 * iw + 2] * k->vi[3 * iv + 2]) / l2) / 4;
		} else {
			const double *r = q + 5 * iw;
			f[iw * 6 + 0] = +w[iw] / 6. + r[0] / (2. * l) + r[3] / (6. * l2) + r[4] / (6. * l2);
			f[iw * 6 + 1] = +w[iw] / 6. + r[1] / (2. * l) - r[3] / (3. * l2) + r[4] / (6. * l2);
			f[iw * 6 + 2] = +w[iw] / 6. + r[2] / (2. * l) + r[3] / (6. * l2) - r[4] / (3. * l2);
			f[iw * 6 + 3] = +w[iw] / 6. - r[0] / (2. * l) + r[3] / (6. * l2) + r[4] / (6. * l2);
			f[iw * 6 + 4] = +w[iw] / 6. - r[1] / (2. * l) - r[3] / (3. * l2) + r[4] / (6. * l2);
			f[iw * 6 + 5] = +w[iw] / 6. - r[2] / (2. * l) + r[3] / (6. * l2) - r[4] / (3. * l2);
		}
	}
}

/* The boundary state of get_fR's use_feq = false branch
 * (src/models/d3q4.c:303-332, src/models/d3q6.c:358-379) with the model
 * projectors of the reference examples (get_proj_pos =
 * maxwell_get_proj_pos_null_prod_w, get_proj_neg =
 * maxwell_get_proj_neg_prod_w): (wL, yL) = Q(fL); wR = P^- wI (D3Q4; the
 * reference overwrites the sum with that term) or P^{+,0} wL + P^- wI (D3Q6);
 * fR = Q^-1(wR, yL). */
void orc_fr_projector(int kind, const double *fL, const double *vn, const double *wI,
		      double *fR)
{
	okin k;
	kin_set(&k, kind);
	const int nm = k.nv - 1;
	double wL[6], q[30], qM[30], wR[6], pn[6][6], pp[6][6];
	kin_moments(&k, fL, wL, q);
	kin_qM(&k, wL, qM);
	for (int c = 0; c < 6 * nm; c++)
		q[c] -= qM[c]; /* yL */
	proj_matrices(vn, pn, pp);
	for (int i = 0; i < 6; i++) {
		wR[i] = 0;
		for (int j = 0; j < 6; j++)
			wR[i] += pn[i][j] * wI[j] + (kind == 2 ? pp[i][j] * wL[j] : 0.0);
	}
	kin_qM(&k, wR, qM);
	for (int c = 0; c < 6 * nm; c++)
		q[c] += qM[c];
	kin_from_moments(&k, wR, q, fR);
}

static void kin_feq(const okin *k, const double *w, double *feq)
{
	kin_feq_eps(k, w, 1.0, feq);
}

/* =================================================================== */
/*                          sparse matrices                            */
/* =================================================================== */

typedef struct {
	size_t nz, cap;
	int *i, *j;
	double *x;
} coo;

typedef struct {
	int n;
	int64_t *p;
	int *j;
	double *x;
} csr;

static void coo_add(coo *t, int i, int j, double x)
{
	if (t->nz == t->cap) {
		t->cap = t->cap ? 2 * t->cap : 1024;
		t->i = realloc(t->i, t->cap * sizeof(int));
		t->j = realloc(t->j, t->cap * sizeof(int));
		t->x = realloc(t->x, t->cap * sizeof(double));
		if (!t->i || !t->j || !t->x)
			die("out of memory");
	}
	t->i[t->nz] = i;
	t->j[t->nz] = j;
	t->x[t->nz] = x;
	t->nz++;
}

static void coo_free(coo *t)
{
	free(t->i);
	free(t->j);
	free(t->x);
	memset(t, 0, sizeof(*t));
}

static void csr_free(csr *a)
{
	free(a->p);
	free(a->j);
	free(a->x);
	memset(a, 0, sizeof(*a));
}

/* rows in insertion order, duplicates summed into their first occurrence
 * in insertion order (cs_compress + cs_dupl on the transposed triplets) */
static void coo_to_csr(coo *t, int n, csr *a)
{
	a->n = n;
	a->p = xc((size_t)n + 1, sizeof(int64_t));
	for (size_t k = 0; k < t->nz; k++)
		a->p[t->i[k] + 1]++;
	for (int r = 0; r < n; r++)
		a->p[r + 1] += a->p[r];
	int *tj = xc(t->nz, sizeof(int));
	double *tx = xc(t->nz, sizeof(double));
	int64_t *pos = xc((size_t)n, sizeof(int64_t));
	memcpy(pos, a->p, sizeof(int64_t) * n);
	for (size_t k = 0; k < t->nz; k++) {
		int64_t q = pos[t->i[k]]++;
		tj[q] = t->j[k];
		tx[q] = t->x[k];
	}
	free(pos);
	coo_free(t);
	int64_t *first = xc((size_t)n, sizeof(int64_t));
	for (int c = 0; c < n; c++)
		first[c] = -1;
	int64_t nz = 0;
	for (int r = 0; r < n; r++) {
		int64_t q0 = nz;
		for (int64_t k = a->p[r]; k < a->p[r + 1]; k++) {
			int c = tj[k];
			if (first[c] >= q0) {
				tx[first[c]] += tx[k];
			} else {
				first[c] = nz;
				tj[nz] = c;
				tx[nz] = tx[k];
				nz++;
			}
		}
		a->p[r] = q0;
	}
	a->p[n] = nz;
	free(first);
	a->j = realloc(tj, sizeof(int) * (size_t)(nz ? nz : 1));
	a->x = realloc(tx, sizeof(double) * (size_t)(nz ? nz : 1));
}

/* res = A x (src/maths/sparse.c:387-402, per-row accumulator) */
static void spmv(const csr *a, const double *x, double *res)
{
#pragma omp parallel for schedule(static)
	for (int r = 0; r < a->n; r++) {
		double s = 0.;
		for (int64_t k = a->p[r]; k < a->p[r + 1]; k++)
			s += a->x[k] * x[a->j[k]];
		res[r] = s;
	}
}

/* C = alpha I + beta A B (rows of A times B, columns in first-appearance
 * order; A here is the block-diagonal inverse mass) */
static void spgemm_shift(const csr *A, const csr *B, double alpha, double beta,
			 csr *C)
{
	const int n = A->n;
	C->n = n;
	C->p = xc((size_t)n + 1, sizeof(int64_t));
	int64_t cap = B->p[n] + n + 16;
	C->j = xc((size_t)cap, sizeof(int));
	C->x = xc((size_t)cap, sizeof(double));
	int64_t *mark = xc((size_t)n, sizeof(int64_t));
	for (int c = 0; c < n; c++)
		mark[c] = -1;
	int64_t nz = 0;
	for (int r = 0; r < n; r++) {
		int64_t q0 = nz;
		C->p[r] = q0;
		for (int64_t k = A->p[r]; k < A->p[r + 1]; k++) {
			int m = A->j[k];
			double a = A->x[k];
			for (int64_t l = B->p[m]; l < B->p[m + 1]; l++) {
				int c = B->j[l];
				if (mark[c] < q0) {
					if (nz == cap) {
						cap *= 2;
						C->j = realloc(C->j, sizeof(int) * cap);
						C->x = realloc(C->x, sizeof(double) * cap);
						if (!C->j || !C->x)
							die("out of memory");
					}
					mark[c] = nz;
					C->j[nz] = c;
					C->x[nz] = 0;
					nz++;
				}
				C->x[mark[c]] += a * B->x[l];
			}
		}
		for (int64_t k = q0; k < nz; k++)
			C->x[k] *= beta;
		/* + alpha I */
		if (mark[r] >= q0) {
			C->x[mark[r]] = alpha + C->x[mark[r]];
		} else {
			if (nz == cap) {
				cap *= 2;
				C->j = realloc(C->j, sizeof(int) * cap);
				C->x = realloc(C->x, sizeof(double) * cap);
			}
			C->j[nz] = r;
			C->x[nz] = alpha;
			nz++;
		}
	}
	C->p[n] = nz;
	free(mark);
}

/* ------------------ sparse LU: strongly connected blocks -------------- */

typedef struct {
	int n, nblk;
	int *blk_ptr; /* nblk + 1 offsets into order */
	int *order; /* unknowns grouped by block, blocks in dependency order */
	int *inblk; /* position of unknown within its block */
	int *blkof;
	int64_t *lu_ptr; /* dense LU of each block (row-major) */
	double *lu;
	int *piv; /* pivot rows per block position */
	const csr *A;
} blocklu;

/* Tarjan, iterative; blocks are emitted after everything they reference */
static void scc(const csr *A, blocklu *F)
{
	const int n = A->n;
	int *index = xc((size_t)n, sizeof(int)), *low = xc((size_t)n, sizeof(int));
	char *onst = xc((size_t)n, 1);
	int *st = xc((size_t)n, sizeof(int)), sp = 0;
	int *cs = xc((size_t)n, sizeof(int));
	int64_t *ce = xc((size_t)n, sizeof(int64_t));
	int csp = 0, idx = 1;
	F->order = xc((size_t)n, sizeof(int));
	F->blk_ptr = xc((size_t)n + 1, sizeof(int));
	F->blkof = xc((size_t)n, sizeof(int));
	F->inblk = xc((size_t)n, sizeof(int));
	int nord = 0, nb = 0;
	for (int s = 0; s < n; s++) {
		if (index[s])
			continue;
		cs[csp] = s;
		ce[csp] = A->p[s];
		csp++;
		index[s] = low[s] = idx++;
		st[sp++] = s;
		onst[s] = 1;
		while (csp) {
			int v = cs[csp - 1];
			if (ce[csp - 1] < A->p[v + 1]) {
				int w = A->j[ce[csp - 1]++];
				if (w == v || A->x[ce[csp - 1] - 1] == 0.0)
					continue;
				if (!index[w]) {
					index[w] = low[w] = idx++;
					st[sp++] = w;
					onst[w] = 1;
					cs[csp] = w;
					ce[csp] = A->p[w];
					csp++;
				} else if (onst[w] && index[w] < low[v]) {
					low[v] = index[w];
				}
			} else {
				csp--;
				if (csp && low[v] < low[cs[csp - 1]])
					low[cs[csp - 1]] = low[v];
				if (low[v] == index[v]) {
					F->blk_ptr[nb] = nord;
					int w, k = 0;
					do {
						w = st[--sp];
						onst[w] = 0;
						F->order[nord++] = w;
						F->blkof[w] = nb;
						F->inblk[w] = k++;
					} while (w != v);
					nb++;
				}
			}
		}
	}
	F->blk_ptr[nb] = nord;
	F->nblk = nb;
	free(index);
	free(low);
	free(onst);
	free(st);
	free(cs);
	free(ce);
}

static void blocklu_factor(const csr *A, blocklu *F)
{
	memset(F, 0, sizeof(*F));
	F->n = A->n;
	F->A = A;
	scc(A, F);
	F->lu_ptr = xc((size_t)F->nblk + 1, sizeof(int64_t));
	for (int b = 0; b < F->nblk; b++) {
		int64_t s = F->blk_ptr[b + 1] - F->blk_ptr[b];
		F->lu_ptr[b + 1] = F->lu_ptr[b] + s * s;
	}
	F->lu = xc((size_t)F->lu_ptr[F->nblk], sizeof(double));
	F->piv = xc((size_t)F->n, sizeof(int));
#pragma omp parallel for schedule(dynamic, 256)
	for (int b = 0; b < F->nblk; b++) {
		const int r0 = F->blk_ptr[b], s = F->blk_ptr[b + 1] - r0;
		double *M = F->lu + F->lu_ptr[b];
		for (int r = 0; r < s; r++) {
			const int row = F->order[r0 + r];
			for (int64_t k = A->p[row]; k < A->p[row + 1]; k++)
				if (F->blkof[A->j[k]] == b)
					M[(int64_t)r * s + F->inblk[A->j[k]]] += A->x[k];
		}
		int *pv = F->piv + r0;
		for (int c = 0; c < s; c++) {
			int p = c;
			for (int r = c + 1; r < s; r++)
				if (fabs(M[(int64_t)r * s + c]) > fabs(M[(int64_t)p * s + c]))
					p = r;
			pv[c] = p;
			if (p != c)
				for (int k = 0; k < s; k++) {
					double t = M[(int64_t)c * s + k];
					M[(int64_t)c * s + k] = M[(int64_t)p * s + k];
					M[(int64_t)p * s + k] = t;
				}
			const double d = M[(int64_t)c * s + c];
			if (d == 0.0)
				die("singular diagonal block");
			for (int r = c + 1; r < s; r++) {
				double l = M[(int64_t)r * s + c] / d;
				M[(int64_t)r * s + c] = l;
				for (int k = c + 1; k < s; k++)
					M[(int64_t)r * s + k] -= l * M[(int64_t)c * s + k];
			}
		}
	}
}

/* solve A x = b in place (x holds b on entry) */
static void blocklu_solve(const blocklu *F, double *x)
{
	const csr *A = F->A;
	double tmp[4096];
	for (int b = 0; b < F->nblk; b++) {
		const int r0 = F->blk_ptr[b], s = F->blk_ptr[b + 1] - r0;
		double *y = s <= 4096 ? tmp : xc((size_t)s, sizeof(double));
		for (int r = 0; r < s; r++) {
			const int row = F->order[r0 + r];
			double v = x[row];
			for (int64_t k = A->p[row]; k < A->p[row + 1]; k++)
				if (F->blkof[A->j[k]] != b)
					v -= A->x[k] * x[A->j[k]];
			y[r] = v;
		}
		const double *M = F->lu + F->lu_ptr[b];
		const int *pv = F->piv + r0;
		for (int c = 0; c < s; c++)
			if (pv[c] != c) {
				double t = y[c];
				y[c] = y[pv[c]];
				y[pv[c]] = t;
			}
		for (int r = 1; r < s; r++)
			for (int c = 0; c < r; c++)
				y[r] -= M[(int64_t)r * s + c] * y[c];
		for (int r = s - 1; r >= 0; r--) {
			for (int c = r + 1; c < s; c++)
				y[r] -= M[(int64_t)r * s + c] * y[c];
			y[r] /= M[(int64_t)r * s + r];
		}
		/* unknown inblk k of the block is order[r0 + k] */
		for (int r = 0; r < s; r++)
			x[F->order[r0 + r]] = y[r];
		if (y != tmp)
			free(y);
	}
}

static void blocklu_free(blocklu *F)
{
	free(F->blk_ptr);
	free(F->order);
	free(F->inblk);
	free(F->blkof);
	free(F->lu_ptr);
	free(F->lu);
	free(F->piv);
}

/* =================================================================== */
/*                     simulation: assembly and solve                  */
/* =================================================================== */

typedef struct {
	omesh *m;
	int scheme; /* 0 RK3, 1 kinetic */
	int bflux; /* 1 upwind, 2 metal */
	oimposed ip;
	okin kin;
	int nw, nv, mm, neq, itermax;
	double t, tmax, dt, cfl, hmin, lambda;
	double *wn;
	double t_asm, t_run;
	int nthreads;
	int fr_projector; /* boundary state: 0 f_eq(wI), 1 the projector form */
	double *eps; /* permittivity per element (NULL: 1) */
	/* assembled operators */
	csr Minv, F; /* RK3 */
	csr M1[8], M2[8];
	blocklu LU[8];
	int assembled;
	int nofactor; /* kinetic: assemble M1, M2 without factoring M1 */
} osim;

void *orc_sim_create(void *mesh, int scheme, int imposed_kind,
		     const double *imposed_par, int kin_kind, int bflux)
{
	osim *s = xc(1, sizeof(osim));
	s->m = mesh;
	s->scheme = scheme;
	s->bflux = bflux ? bflux : 1;
	s->ip.kind = imposed_kind;
	if (imposed_par)
		memcpy(s->ip.p, imposed_par, sizeof(s->ip.p));
	s->nw = 6;
	if (scheme) {
		kin_set(&s->kin, kin_kind);
		s->nv = s->kin.nv;
		s->mm = s->nv * s->nw;
		s->lambda = s->kin.lambda;
	} else {
		s->nv = 1;
		s->mm = s->nw;
		s->lambda = 1.0;
	}
	s->neq = NQN * s->m->nelem + NQV * s->m->nbf;
	s->wn = xc((size_t)s->mm * s->neq, sizeof(double));
	s->hmin = hmin_of(s->m);
	s->nthreads = 1;
	return s;
}

void orc_sim_set_lambda(void *h, double lambda)
{
	((osim *)h)->lambda = lambda;
}

/* permittivity per element (copied; NULL: 1); before solve / operators */
void orc_sim_set_eps(void *h, const double *eps)
{
	osim *s = h;
	if (s->assembled)
		die("orc_sim_set_eps after the operators were assembled");
	free(s->eps);
	s->eps = NULL;
	if (eps) {
		s->eps = xc(s->m->nelem, sizeof(double));
		memcpy(s->eps, eps, sizeof(double) * s->m->nelem);
	}
}

void orc_sim_set_fr_projector(void *h, int on)
{
	((osim *)h)->fr_projector = on;
}

/* 1: the kinetic assembly skips the factorization of M1 (residual checks
 * on meshes whose strong components are too large for dense blocks); the
 * solves then refuse to run */
void orc_sim_set_nofactor(void *h, int on)
{
	osim *s = h;
	if (s->assembled)
		die("orc_sim_set_nofactor after assembly");
	s->nofactor = on;
}

void orc_sim_set_threads(void *h, int n)
{
	((osim *)h)->nthreads = n > 0 ? n : 1;
}

/* src/simulation/simulation.c:144-164 */
void orc_set_time(void *h, double tmax, double dt, double cfl)
{
	osim *s = h;
	s->tmax = tmax;
	if (cfl != -1) {
		s->cfl = cfl;
		s->dt = (s->hmin * cfl) / s->lambda;
	} else {
		s->dt = dt;
		s->cfl = (s->lambda * s->dt) / s->hmin;
	}
	s->itermax = (int)floor(tmax / s->dt);
	s->tmax = s->itermax * s->dt;
}

/* neq, m, itermax, tmax, dt, hmin, cfl, t, t_asm, t_run */
void orc_sim_info(void *h, double *out)
{
	osim *s = h;
	out[0] = s->neq;
	out[1] = s->mm;
	out[2] = s->itermax;
	out[3] = s->tmax;
	out[4] = s->dt;
	out[5] = s->hmin;
	out[6] = s->cfl;
	out[7] = s->t;
	out[8] = s->t_asm;
	out[9] = s->t_run;
}

/* kinetic layout index (src/simulation/simulation.c:22-40) */
static inline size_t kidx(const osim *s, int im, int ipg)
{
	return (size_t)im * s->neq + ipg;
}

/* permittivity of element e and of DoF ipg (ghosts: their L element) */
static double eps_elem(const osim *s, int e)
{
	return s->eps ? s->eps[e] : 1.0;
}

static double eps_dof(const osim *s, int ipg)
{
	const omesh *m = s->m;
	if (!s->eps)
		return 1.0;
	if (ipg < NQN * m->nelem)
		return s->eps[ipg / NQN];
	return s->eps[m->f2e[4 * m->bface[(ipg - NQN * m->nelem) / NQV]]];
}

/* the kinetic state of imposed fields w = (E, H) at DoF ipg: f_eq of the
 * conserved fields (eps E, H) with the flux of (E, H) */
static void kin_state(const osim *s, int ipg, const double *w, double *f)
{
	const double ep = eps_dof(s, ipg);
	const double wc[6] = { ep * w[0], ep * w[1], ep * w[2], w[3], w[4], w[5] };
	kin_feq_eps(&s->kin, wc, 1.0 / ep, f);
}

/* src/dg/dg.c:32-59 */
static void init_sol(osim *s)
{
	const omesh *m = s->m;
	double w[6], f[48];
	for (int e = 0; e < m->nelem; e++)
		for (int i = 0; i < NQN; i++) {
			int ipg = NQN * e + i;
			const double *x = m->xq + 3 * m->e2q[ipg];
			exact(&s->ip, x, 0., w);
			if (s->scheme) {
				kin_state(s, ipg, w, f);
				for (int im = 0; im < s->mm; im++)
					s->wn[kidx(s, im, ipg)] = f[im];
			} else {
				memcpy(s->wn + (size_t)6 * ipg, w, sizeof(w));
			}
		}
}

/* src/dg/dg.c:61-107 */
static void face_vn(const omesh *m, int F, double vn[3]);

static void update_boundary(osim *s, double t)
{
	const omesh *m = s->m;
	double w[6], f[48];
	for (int bf = 0; bf < m->nbf; bf++) {
		int F = m->bface[bf];
		for (int a = 0; a < NQV; a++) {
			int ipgL = m->f2e[4 * F] * NQN + m->fn2en[2 * NQV * F + a];
			int ipgR = NQN * m->nelem + NQV * bf + a;
			const double *x = m->xq + 3 * m->e2q[ipgL];
			exact(&s->ip, x, t, w);
			if (s->scheme && s->fr_projector) {
				/* src/dg/dg.c:95-100: fL at the L node, the face's normal */
				double fl[48], vn[3];
				for (int im = 0; im < s->mm; im++)
					fl[im] = s->wn[kidx(s, im, ipgL)];
				face_vn(m, F, vn);
				orc_fr_projector(s->kin.kind, fl, vn, w, f);
				for (int im = 0; im < s->mm; im++)
					s->wn[kidx(s, im, ipgR)] = f[im];
			} else if (s->scheme) {
				kin_state(s, ipgR, w, f);
				for (int im = 0; im < s->mm; im++)
					s->wn[kidx(s, im, ipgR)] = f[im];
			} else {
				memcpy(s->wn + (size_t)6 * ipgR, w, sizeof(w));
			}
		}
	}
}

/* src/simulation/simulation.c:218-261 (inverse mass; per component for
 * the macro scheme, scalar for the kinetic one) */
static void asm_inv_mass(osim *s, coo *t)
{
	const omesh *m = s->m;
	const int nw = s->scheme ? 1 : s->nw;
	for (int e = 0; e < m->nelem; e++) {
		double cod[3][3], det;
		elem_geom(m, e, cod, &det);
		for (int i = 0; i < NQN; i++)
			for (int j = 0; j < NQN; j++) {
				double v = T_imass[i][j] / det;
				/* RK3, eps dE/dt = curl H: the E rows of M^-1 carry
				 * 1/eps (the kinetic transport has no permittivity) */
				for (int iw = 0; iw < nw; iw++)
					coo_add(t, (NQN * e + i) * nw + iw,
						(NQN * e + j) * nw + iw,
						!s->scheme && iw < 3 ? v / eps_elem(s, e) : v);
			}
	}
}

static void kin_flux(double wL, double wR, const double *vn, const double *v,
		     double *fl)
{
	double x = v[0] * vn[0] + v[1] * vn[1] + v[2] * vn[2];
	double p = x > 0 ? x : 0, q = x - p;
	*fl = p * wL + q * wR;
}

/* src/simulation/simulation.c:263-324. The reference leaves the state
 * vector uninitialised before the first column (it zeroes w[nb_w]); the
 * intended w = 0 is used. */
static void asm_internal(osim *s, coo *t)
{
	const omesh *m = s->m;
	for (int e = 0; e < m->nelem; e++) {
		double cod[3][3], det;
		elem_geom(m, e, cod, &det);
		for (int i = 0; i < NQN; i++)
			for (int j = 0; j < NQN; j++) {
				double c[3] = { 0, 0, 0 };
				for (int l = 0; l < 3; l++)
					for (int k = 0; k < 3; k++)
						c[l] += cod[l][k] * T_int[i][j][k];
				int ipg = NQN * e + i, jpg = NQN * e + j;
				if (s->scheme) {
					for (int iv = 0; iv < s->nv; iv++) {
						double ft;
						kin_flux(1., 1., c, s->kin.vi + 3 * iv, &ft);
						coo_add(&t[iv], ipg, jpg, -ft);
					}
				} else {
					double w[6] = { 0, 0, 0, 0, 0, 0 }, fl[6];
					for (int k = 0; k < 6; k++) {
						w[k] = 1;
						for (int l = 0; l < 6; l++) {
							flux_upwind(w, w, c, fl);
							coo_add(&t[0], 6 * ipg + k, 6 * jpg + l, -fl[l]);
						}
						w[k] = 0;
					}
				}
			}
	}
}

/* scaled normal of the L element's face (src/simulation/simulation.c:377-395) */
static void face_vn(const omesh *m, int F, double vn[3])
{
	int eL = m->f2e[4 * F], fL = m->f2e[4 * F + 1];
	double pn[4][3], dt[3][3], cod[3][3], xr[3] = { 0, 0, 0 };
	elem_nodes(m, eL, pn);
	ref2phy(pn, xr, NULL, dt, cod, fL, vn);
	double cr = fL == 3 ? sqrt(3) : 1;
	for (int d = 0; d < 3; d++)
		vn[d] *= cr;
}

/* src/simulation/simulation.c:356-491 */
static void asm_flux(osim *s, coo *t)
{
	const omesh *m = s->m;
	const int start = NQN * m->nelem;
	int fc = -1;
	double wL[6] = { 0 }, wR[6] = { 0 }, fl[6];
	for (int F = 0; F < m->nfaces; F++) {
		int eL = m->f2e[4 * F], eR = m->f2e[4 * F + 2];
		if (eR < 0)
			fc++;
		double vn[3];
		face_vn(m, F, vn);
		const int *fnL = m->fn2en + 2 * NQV * F, *fnR = fnL + NQV;
		for (int side = 0; side < 2; side++) {
			double *wx = side ? wR : wL;
			for (int k = 0; k < 6; k++)
				for (int l = 0; l < 6; l++) {
					wx[l] = 1;
					if (eR >= 0 || s->bflux == 1)
						flux_upwind(wL, wR, vn, fl);
					else
						flux_metal(wL, wR, vn, fl);
					wx[l] = 0;
					for (int a = 0; a < NQV; a++) {
						int ipg = fnL[a] + eL * NQN;
						for (int b = 0; b < NQV; b++) {
							int jpg = side == 0 ? fnL[b] + eL * NQN :
								  eR >= 0 ? fnR[b] + eR * NQN :
									    fnR[b] + start + fc * NQV;
							double v = fl[k] * T_face[a][b];
							if (v != 0.)
								coo_add(t, 6 * ipg + k, 6 * jpg + l, v);
						}
					}
					for (int a = 0; a < NQV && eR >= 0; a++) {
						int ipg = fnR[a] + eR * NQN;
						for (int b = 0; b < NQV; b++) {
							int jpg = side == 0 ? fnL[b] + eL * NQN :
									      fnR[b] + eR * NQN;
							double v = fl[k] * T_face[a][b];
							if (v != 0.)
								coo_add(t, 6 * ipg + k, 6 * jpg + l, -v);
						}
					}
				}
		}
	}
}

/* src/simulation/simulation.c:493-604 */
static void asm_flux_kinetic(osim *s, coo *t)
{
	const omesh *m = s->m;
	const int start = NQN * m->nelem;
	int fc = -1;
	for (int F = 0; F < m->nfaces; F++) {
		int eL = m->f2e[4 * F], eR = m->f2e[4 * F + 2];
		if (eR < 0)
			fc++;
		double vn[3];
		face_vn(m, F, vn);
		const int *fnL = m->fn2en + 2 * NQV * F, *fnR = fnL + NQV;
		for (int iv = 0; iv < s->nv; iv++) {
			const double *v = s->kin.vi + 3 * iv;
			double fl;
			kin_flux(1., 0., vn, v, &fl);
			for (int a = 0; a < NQV; a++)
				for (int b = 0; b < NQV; b++) {
					double x = fl * T_face[a][b];
					if (x != 0.)
						coo_add(&t[iv], fnL[a] + eL * NQN, fnL[b] + eL * NQN, x);
				}
			for (int a = 0; a < NQV && eR >= 0; a++)
				for (int b = 0; b < NQV; b++) {
					double x = fl * T_face[a][b];
					if (x != 0.)
						coo_add(&t[iv], fnR[a] + eR * NQN, fnL[b] + eL * NQN, -x);
				}
			kin_flux(0., 1., vn, v, &fl);
			for (int a = 0; a < NQV; a++)
				for (int b = 0; b < NQV; b++) {
					int jpg = eR >= 0 ? fnR[b] + eR * NQN : fnR[b] + start + fc * NQV;
					double x = fl * T_face[a][b];
					if (x != 0.)
						coo_add(&t[iv], fnL[a] + eL * NQN, jpg, x);
				}
			for (int a = 0; a < NQV && eR >= 0; a++)
				for (int b = 0; b < NQV; b++) {
					double x = fl * T_face[a][b];
					if (x != 0.)
						coo_add(&t[iv], fnR[a] + eR * NQN, fnR[b] + eR * NQN, -x);
				}
		}
	}
}

/* src/simulation/simulation.c:326-354: ghost rows M = 1, F = 0 */
static void asm_extend(osim *s, coo *mass, coo *flux)
{
	const omesh *m = s->m;
	int start = NQN * m->nelem;
	for (int bf = 0; bf < m->nbf; bf++)
		for (int a = 0; a < NQV; a++) {
			int ipg = start + bf * NQV + a;
			if (s->scheme) {
				coo_add(mass, ipg, ipg, 1);
				for (int iv = 0; iv < s->nv; iv++)
					coo_add(&flux[iv], ipg, ipg, 0);
			} else {
				for (int iw = 0; iw < 6; iw++) {
					coo_add(mass, 6 * ipg + iw, 6 * ipg + iw, 1);
					coo_add(&flux[0], 6 * ipg + iw, 6 * ipg + iw, 0);
				}
			}
		}
}

static void assemble(osim *s)
{
	if (s->assembled)
		return;
	coo mass = { 0 }, flux[8];
	memset(flux, 0, sizeof(flux));
	asm_inv_mass(s, &mass);
	asm_internal(s, flux);
	if (s->scheme)
		asm_flux_kinetic(s, flux);
	else
		asm_flux(s, flux);
	asm_extend(s, &mass, flux);
	if (!s->scheme) {
		coo_to_csr(&mass, s->mm * s->neq, &s->Minv);
		coo_to_csr(&flux[0], s->mm * s->neq, &s->F);
	} else {
		csr Mi, Fv;
		coo_to_csr(&mass, s->neq, &Mi);
		const double theta = 0.5;
		for (int iv = 0; iv < s->nv; iv++) {
			coo_to_csr(&flux[iv], s->neq, &Fv);
			spgemm_shift(&Mi, &Fv, 1.0, theta * s->dt, &s->M1[iv]);
			spgemm_shift(&Mi, &Fv, 1.0, -(1 - theta) * s->dt, &s->M2[iv]);
			csr_free(&Fv);
			if (!s->nofactor)
				blocklu_factor(&s->M1[iv], &s->LU[iv]);
		}
		csr_free(&Mi);
	}
	s->assembled = 1;
}

/* src/dg/dg.c:109-147 */
static void relax_all(osim *s, int with_ghosts)
{
	const int n = with_ghosts ? s->neq : NQN * s->m->nelem;
	const double om = s->kin.omega;
	/* serial, as the reference (its omp pragma is commented out) */
	for (int ipg = 0; ipg < n; ipg++) {
		double f[48], w[6], feq[48];
		for (int im = 0; im < s->mm; im++)
			f[im] = s->wn[kidx(s, im, ipg)];
		for (int iw = 0; iw < 6; iw++) {
			w[iw] = 0;
			for (int iv = 0; iv < s->nv; iv++)
				w[iw] += f[iw * s->nv + iv];
		}
		kin_feq_eps(&s->kin, w, 1.0 / eps_dof(s, ipg), feq);
		for (int im = 0; im < s->mm; im++)
			s->wn[kidx(s, im, ipg)] = om * feq[im] + (1. - om) * f[im];
	}
}

static void set_threads(const osim *s)
{
#ifdef _OPENMP
	omp_set_num_threads(s->nthreads);
#else
	(void)s;
#endif
}

/* src/dg/explicit.c:39-117 */
static void solve_rk3(osim *s, int nsteps)
{
	const double A[3] = { 0., -5. / 9., -153. / 128. };
	const double B[3] = { 1. / 3., 15. / 16., 8. / 15. };
	const double C[3] = { 0., 1. / 3., 3. / 4. };
	const size_t len = (size_t)s->mm * s->neq;
	double *k2 = xc(len, sizeof(double)), *dw = xc(len, sizeof(double)),
	       *tmp = xc(len, sizeof(double));
	double t0 = now();
	assemble(s);
	s->t_asm = now() - t0;
	init_sol(s);
	double t = 0;
	int iter = 0;
	t0 = now();
	while (t < s->tmax - 1e-12 && (nsteps < 0 || iter < nsteps)) {
		s->t = t;
		for (int i = 0; i < 3; i++) {
			update_boundary(s, s->t + C[i] * s->dt);
			spmv(&s->F, s->wn, tmp);
			spmv(&s->Minv, tmp, dw);
			const double Ai = A[i], Bi = B[i], dt = s->dt;
#pragma omp parallel for schedule(static)
			for (size_t j = 0; j < len; j++) {
				k2[j] = i == 0 ? -dt * dw[j] : Ai * k2[j] - dt * dw[j];
				s->wn[j] = s->wn[j] + Bi * k2[j];
			}
		}
		iter++;
		t += s->dt;
	}
	s->t = t;
	s->t_run = now() - t0;
	free(k2);
	free(dw);
	free(tmp);
}

/* src/dg/relaxation.c:93-185 */
static void solve_relax(osim *s, int nsteps)
{
	double t0 = now();
	if (s->nofactor)
		die("the kinetic solve needs the factorization of M1");
	assemble(s);
	s->t_asm = now() - t0;
	init_sol(s);
	double *sol = xc((size_t)s->mm * s->neq, sizeof(double));
	double t = 0;
	int iter = 0;
	t0 = now();
	while (t < s->tmax - 1e-12 && (nsteps < 0 || iter < nsteps)) {
		update_boundary(s, t + s->dt / 2);
		for (int iw = 0; iw < 6; iw++) {
#pragma omp parallel for schedule(static)
			for (int iv = 0; iv < s->nv; iv++) {
				size_t off = (size_t)(iw * s->nv + iv) * s->neq;
				spmv(&s->M2[iv], s->wn + off, sol + off);
				blocklu_solve(&s->LU[iv], sol + off);
			}
		}
		double *p = s->wn;
		s->wn = sol;
		sol = p;
		relax_all(s, 1);
		t += s->dt;
		s->t = t;
		iter++;
	}
	update_boundary(s, t);
	s->t_run = now() - t0;
	free(sol);
}

/* run the whole solve (nsteps < 0) or only the first nsteps steps */
void orc_solve(void *h, int nsteps)
{
	osim *s = h;
	tables_init();
	set_threads(s);
	if (s->scheme)
		solve_relax(s, nsteps);
	else
		solve_rk3(s, nsteps);
}

void orc_get_wn(void *h, double *out)
{
	osim *s = h;
	memcpy(out, s->wn, sizeof(double) * (size_t)s->mm * s->neq);
}

void orc_set_wn(void *h, const double *in)
{
	osim *s = h;
	memcpy(s->wn, in, sizeof(double) * (size_t)s->mm * s->neq);
}

/* one application dw = M^-1 F w of the assembled RK3 operator */
void orc_rk3_operator(void *h, const double *w, double *dw)
{
	osim *s = h;
	tables_init();
	set_threads(s);
	assemble(s);
	double *tmp = xc((size_t)s->mm * s->neq, sizeof(double));
	spmv(&s->F, w, tmp);
	spmv(&s->Minv, tmp, dw);
	free(tmp);
}

/* sol = M1^-1 (M2 f) for velocity iv on one micro-field of length neq */
void orc_transport(void *h, int iv, const double *f, double *sol)
{
	osim *s = h;
	if (s->nofactor)
		die("orc_transport needs the factorization of M1");
	tables_init();
	set_threads(s);
	assemble(s);
	spmv(&s->M2[iv], f, sol);
	blocklu_solve(&s->LU[iv], sol);
}

/* residual || M1 x - M2 f ||_2 / || M2 f ||_2 of a transport solution */
double orc_transport_residual(void *h, int iv, const double *f, const double *x)
{
	osim *s = h;
	assemble(s);
	double *r1 = xc((size_t)s->neq, sizeof(double)), *r2 = xc((size_t)s->neq, sizeof(double));
	spmv(&s->M1[iv], x, r1);
	spmv(&s->M2[iv], f, r2);
	double num = 0, den = 0;
	for (int i = 0; i < s->neq; i++) {
		num += (r1[i] - r2[i]) * (r1[i] - r2[i]);
		den += r2[i] * r2[i];
	}
	free(r1);
	free(r2);
	return sqrt(num / den);
}

/* number of diagonal blocks of M1 for velocity iv and the largest size */
void orc_block_stats(void *h, int iv, int *out)
{
	osim *s = h;
	if (s->nofactor)
		die("orc_block_stats needs the factorization of M1");
	assemble(s);
	const blocklu *F = &s->LU[iv];
	int mx = 0;
	for (int b = 0; b < F->nblk; b++)
		if (F->blk_ptr[b + 1] - F->blk_ptr[b] > mx)
			mx = F->blk_ptr[b + 1] - F->blk_ptr[b];
	out[0] = F->nblk;
	out[1] = mx;
}

/* upwind dependencies read off the assembled M1 of velocity iv: for each
 * element, the other elements its rows reference (-2 for ghost DoFs),
 * 4 slots per element, -1 padded */
void orc_m1_dependencies(void *h, int iv, int *out)
{
	osim *s = h;
	assemble(s);
	const csr *A = &s->M1[iv];
	const int ne = s->m->nelem;
	for (int e = 0; e < ne; e++) {
		int k = 0, *o = out + 4 * e;
		o[0] = o[1] = o[2] = o[3] = -1;
		for (int r = NQN * e; r < NQN * e + NQN; r++)
			for (int64_t q = A->p[r]; q < A->p[r + 1]; q++) {
				if (A->x[q] == 0.0)
					continue;
				const int c = A->j[q];
				const int ce = c < NQN * ne ? c / NQN : -2;
				if (ce == e)
					continue;
				int seen = 0;
				for (int u = 0; u < k; u++)
					seen |= o[u] == ce && ce != -2;
				if (seen || (ce == -2 && k && o[k - 1] == -2))
					continue;
				if (k < 4)
					o[k++] = ce;
			}
	}
}

void orc_relax(void *h, int with_ghosts)
{
	osim *s = h;
	set_threads(s);
	relax_all(s, with_ghosts);
}

void orc_update_boundary(void *h, double t)
{
	update_boundary(h, t);
}

void orc_init_solution(void *h)
{
	tables_init();
	init_sol(h);
}

/* src/simulation/simulation.c:606-666 */
double orc_error_l2(void *h)
{
	osim *s = h;
	const omesh *m = s->m;
	double err[6] = { 0 }, wex[6], wl[6];
	for (int e = 0; e < m->nelem; e++) {
		double cod[3][3], det;
		elem_geom(m, e, cod, &det);
		for (int i = 0; i < NQN; i++) {
			int ipg = NQN * e + i;
			exact(&s->ip, m->xq + 3 * m->e2q[ipg], s->t, wex);
			for (int iw = 0; iw < 6; iw++) {
				if (s->scheme) {
					wl[iw] = 0;
					for (int iv = 0; iv < s->nv; iv++)
						wl[iw] += s->wn[kidx(s, iw * s->nv + iv, ipg)];
					if (iw < 3)
						wl[iw] /= eps_elem(s, e); /* E = D / eps */
				} else {
					wl[iw] = s->wn[(size_t)6 * ipg + iw];
				}
				double d = wl[iw] - wex[iw];
				err[iw] += d * d * T_wpg[i] * det;
			}
		}
	}
	double e2 = 0;
	for (int iw = 0; iw < 6; iw++)
		e2 += sqrt(fabs(err[iw])) / 6;
	return e2;
}

/* Discrete energy 1/2 w^T M w with M the assembled P2 mass matrix
 * (src/simulation/simulation.c:218-261: det_e times the reference mass per
 * component) over the element DoFs; w the macro fields, or sum_v f_v
 * (src/models/d3q4.c:82-95) for the kinetic scheme. */
double orc_energy(void *h)
{
	tables_init();
	osim *s = h;
	const omesh *m = s->m;
	double acc = 0;
	for (int e = 0; e < m->nelem; e++) {
		double cod[3][3], det;
		elem_geom(m, e, cod, &det);
		for (int iw = 0; iw < 6; iw++) {
			double w[NQN];
			for (int i = 0; i < NQN; i++) {
				int ipg = NQN * e + i;
				if (s->scheme) {
					w[i] = 0;
					for (int iv = 0; iv < s->nv; iv++)
						w[i] += s->wn[kidx(s, iw * s->nv + iv, ipg)];
				} else {
					w[i] = s->wn[(size_t)6 * ipg + iw];
				}
			}
			/* 1/2 (eps |E|^2 + |H|^2) (RK3), 1/2 (|D|^2 / eps + |B|^2) */
			const double wt = iw >= 3 ? 1.0
					  : s->scheme ? 1.0 / eps_elem(s, e) : eps_elem(s, e);
			for (int i = 0; i < NQN; i++)
				for (int j = 0; j < NQN; j++)
					acc += 0.5 * det * wt * w[i] * T_mass[i][j] * w[j];
		}
	}
	return acc;
}

void orc_sim_free(void *h)
{
	osim *s = h;
	if (!s)
		return;
	csr_free(&s->Minv);
	csr_free(&s->F);
	for (int iv = 0; iv < 8; iv++) {
		csr_free(&s->M1[iv]);
		csr_free(&s->M2[iv]);
		if (s->assembled && s->scheme && iv < s->nv)
			blocklu_free(&s->LU[iv]);
	}
	free(s->wn);
	free(s->eps);
	free(s);
}
