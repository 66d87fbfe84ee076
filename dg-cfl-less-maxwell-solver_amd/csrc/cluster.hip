/*
 * Cluster plan of the transport sweep (host).
 *
 * One velocity's transport solves M1 sol = M2 f block row by block row in a
 * topological order of its upwind graph (src/dg/relaxation.c:24-110; KLU's
 * block triangular form of M1 is that order). The cluster sweep gives the
 * elements of one small compact group to one wave, which solves them in
 * upwind-level order and passes their values to each other through LDS;
 * only values that cross from one cluster to another go through the
 * hand-off words in memory, so most links of the dependency chain cost an
 * LDS access instead of a round trip between workgroups.
 *
 * Clusters are cells of a grid over each instance's bounding box whose
 * cell volume is that of six average elements (for a Kuhn n^3 x 6 mesh of
 * the unit cube, exactly the n^3 cubes), cut into pieces of at most CL_MAX
 * elements. A persistent launch over clusters taken in a topological order
 * of the cluster graph of each velocity cannot deadlock when that graph is
 * acyclic. Per velocity, the clusters that lie in a strongly connected
 * component of more than one cluster are split into single elements (the
 * element graph is acyclic once cycles are lagged), and a cluster with more
 * than CL_XMAX inflow faces from outside is cut, in level order, into
 * pieces that have at most CL_XMAX. Neither step can close a cycle: a
 * cycle through the result maps onto a closed walk through the clusters it
 * came from. A single element left by these steps then joins the cluster
 * of its producers when they all lie in one cluster with room (it brings no
 * in-edge from outside, so no cycle can form): where grid cells cut across
 * the flow (unstructured meshes) clusters then follow it.
 */
#include "ctx.h"

#include <numeric>

static uint64_t spread3(uint64_t x) /* 21 bits -> every third bit */
{
	x &= 0x1fffff;
	x = (x | x << 32) & 0x1f00000000ffffull;
	x = (x | x << 16) & 0x1f0000ff0000ffull;
	x = (x | x << 8) & 0x100f00f00f00f00full;
	x = (x | x << 4) & 0x10c30c30c30c30c3ull;
	x = (x | x << 2) & 0x1249249249249249ull;
	return x;
}

/* Grid cells of each instance (the whole mesh when not a batch), numbered
 * in Morton order of the cells, instance after instance, and cut into
 * pieces of at most CL_MAX elements: the base cluster of each element. */
static std::vector<int> base_clusters(int ne, const float *vtx, const std::vector<int> &rng,
				       int &nbase)
{
	std::vector<int> bc(ne, -1);
	std::vector<std::pair<uint64_t, int>> key;
	nbase = 0;
	for (size_t r = 0; r + 1 < rng.size(); r++) {
		const int e0 = rng[r], e1 = rng[r + 1];
		if (e1 <= e0)
			continue;
		double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
		for (int e = e0; e < e1; e++)
			for (int k = 0; k < 12; k++) {
				const double x = vtx[12L * e + k];
				lo[k % 3] = std::min(lo[k % 3], x);
				hi[k % 3] = std::max(hi[k % 3], x);
			}
		double ext[3], vol = 1.0;
		for (int d = 0; d < 3; d++) {
			ext[d] = std::max(hi[d] - lo[d], 1e-300);
			vol *= ext[d];
		}
		const double h = std::cbrt(6.0 * vol / (e1 - e0));
		int m[3];
		for (int d = 0; d < 3; d++)
			m[d] = (int)std::max(1.0, std::min(2097151.0, std::round(ext[d] / h)));
		key.resize(e1 - e0);
		for (int e = e0; e < e1; e++) {
			uint64_t code = 0;
			for (int d = 0; d < 3; d++) {
				double s = 0.0;
				for (int i = 0; i < 4; i++)
					s += vtx[12L * e + 3 * i + d];
				const double u = (0.25 * s - lo[d]) / ext[d] * m[d];
				const long cell = std::max(0L, std::min((long)m[d] - 1, (long)std::floor(u)));
				code |= spread3((uint64_t)cell) << d;
			}
			key[e - e0] = { code, e };
		}
		std::sort(key.begin(), key.end());
		int cnt = 0;
		for (size_t k = 0; k < key.size(); k++) {
			if (k == 0 || key[k].first != key[k - 1].first || cnt == CL_MAX) {
				nbase++;
				cnt = 0;
			}
			bc[key[k].second] = nbase - 1;
			cnt++;
		}
	}
	return bc;
}

/* Strongly connected components (Tarjan, iterative) of the graph with
 * adjacency ap/aj over n nodes: component of each node; returns the
 * component sizes. */
static std::vector<int> scc(int n, const std::vector<int> &ap, const std::vector<int> &aj,
			    std::vector<int> &comp)
{
	std::vector<int> idx(n, -1), low(n, 0), stk, size;
	std::vector<char> on(n, 0);
	std::vector<std::pair<int, int>> call;
	comp.assign(n, -1);
	int counter = 0;
	for (int s = 0; s < n; s++) {
		if (idx[s] >= 0)
			continue;
		idx[s] = low[s] = counter++;
		stk.push_back(s);
		on[s] = 1;
		call.push_back({ s, ap[s] });
		while (!call.empty()) {
			const int u = call.back().first, k = call.back().second;
			if (k < ap[u + 1]) {
				call.back().second = k + 1;
				const int w = aj[k];
				if (idx[w] < 0) {
					idx[w] = low[w] = counter++;
					stk.push_back(w);
					on[w] = 1;
					call.push_back({ w, ap[w] });
				} else if (on[w]) {
					low[u] = std::min(low[u], idx[w]);
				}
				continue;
			}
			if (low[u] == idx[u]) {
				const int id = (int)size.size();
				int cnt = 0, w;
				do {
					w = stk.back();
					stk.pop_back();
					on[w] = 0;
					comp[w] = id;
					cnt++;
				} while (w != u);
				size.push_back(cnt);
			}
			call.pop_back();
			if (!call.empty()) {
				const int p = call.back().first;
				low[p] = std::min(low[p], low[u]);
			}
		}
	}
	return size;
}

/* adjacency (CSR) of the graph between groups g[] induced by the element
 * edges producer -> consumer of velocity v */
template <typename Dep>
static void group_graph(int ng, int ne, const std::vector<int> &g, Dep dep,
			std::vector<int> &ap, std::vector<int> &aj)
{
	ap.assign(ng + 1, 0);
	for (int e = 0; e < ne; e++)
		for (int f = 0; f < 4; f++) {
			const int u = dep(e, f);
			if (u >= 0 && g[u] != g[e])
				ap[g[u] + 1]++;
		}
	for (int k = 0; k < ng; k++)
		ap[k + 1] += ap[k];
	aj.resize(ap[ng]);
	std::vector<int> fill(ap.begin(), ap.end() - 1);
	for (int e = 0; e < ne; e++)
		for (int f = 0; f < 4; f++) {
			const int u = dep(e, f);
			if (u >= 0 && g[u] != g[e])
				aj[fill[g[u]]++] = g[e];
		}
}

bool cluster_plan(int ne, int nv, const float *vtx, const int *nbr, const std::vector<int> &rng,
		  const uint8_t *cls, const int *items_ve, const int *level_ptr, int nlevels,
		  ClusterPlan &P)
{
	int nbase = 0;
	const std::vector<int> bc = base_clusters(ne, vtx, rng, nbase);
	/* upwind level and level-order position of each item (v, e) */
	std::vector<int> lev((size_t)nv * ne), ipos((size_t)nv * ne);
	for (int L = 0; L < nlevels; L++)
		for (int k = level_ptr[L]; k < level_ptr[L + 1]; k++) {
			lev[items_ve[k]] = L;
			ipos[items_ve[k]] = k;
		}
	struct CItem {
		int lev, v, key, first, n;
	};
	std::vector<CItem> cis;
	std::vector<int> mem; /* member elements of the cluster items */
	std::vector<int> ap, aj, comp, grp(ne), cid(ne), order(ne), ckey;
	std::vector<int> clev, indeg, queue;
	P.levels = 0;
	P.multi = 0;
	for (int v = 0; v < nv; v++) {
		/* the producer across face f of element e, or -1 */
		auto dep = [&](int e, int f) -> int {
			const int nb = nbr[4L * e + f];
			return nb >= 0 && cls[(4L * e + f) * nv + v] == 2 ? nb : -1;
		};
		group_graph(nbase, ne, bc, dep, ap, aj);
		const std::vector<int> csize = scc(nbase, ap, aj, comp);
		/* groups: base clusters, or single elements inside a cycle */
		for (int e = 0; e < ne; e++)
			grp[e] = csize[comp[bc[e]]] > 1 ? nbase + e : bc[e];
		std::iota(order.begin(), order.end(), 0);
		std::sort(order.begin(), order.end(), [&](int a, int b) {
			if (grp[a] != grp[b])
				return grp[a] < grp[b];
			const int la = lev[(size_t)v * ne + a], lb = lev[(size_t)v * ne + b];
			return la != lb ? la < lb : a < b;
		});
		/* cut each group in level order: at most CL_MAX elements and
		 * CL_XMAX inflow faces from outside the piece */
		int ncid = 0, cnt = 0, nxc = 0; /* current piece: elements, faces from outside */
		ckey.clear();
		for (int k = 0; k < ne; k++) {
			const int e = order[k];
			const bool same = k > 0 && grp[e] == grp[order[k - 1]];
			/* inflow faces of e from outside the current piece, were e
			 * added to it (a producer in the same group comes earlier in
			 * `order`, so its piece is already set) */
			int ex = 0, all = 0;
			for (int f = 0; f < 4; f++) {
				const int u = dep(e, f);
				if (u < 0)
					continue;
				all++;
				ex += !(same && grp[u] == grp[e] && cid[u] == ncid - 1);
			}
			if (!same || cnt == CL_MAX || nxc + ex > CL_XMAX) {
				ncid++;
				ckey.push_back(bc[e]);
				cnt = 0;
				nxc = 0;
				ex = all;
			}
			cid[e] = ncid - 1;
			cnt++;
			nxc += ex;
		}
		/* flow-following merge, in upwind-level order: a single-element
		 * piece whose producers all lie in one piece of fewer than CL_MAX
		 * elements joins it. Its in-edges then all come from inside that
		 * piece, so a cycle through the merged piece would be one through
		 * the piece and the single element before: none can form. This
		 * recovers clusters where grid cells cut across the flow
		 * (unstructured meshes). */
		{
			std::vector<int> psize(ncid, 0), lorder(ne);
			for (int e = 0; e < ne; e++)
				psize[cid[e]]++;
			std::iota(lorder.begin(), lorder.end(), 0);
			std::sort(lorder.begin(), lorder.end(), [&](int a, int b) {
				const int la = lev[(size_t)v * ne + a], lb = lev[(size_t)v * ne + b];
				return la != lb ? la < lb : a < b;
			});
			for (const int e : lorder) {
				if (psize[cid[e]] != 1)
					continue;
				int c = -1, np = 0;
				bool one = true;
				for (int f = 0; f < 4; f++) {
					const int u = dep(e, f);
					if (u < 0)
						continue;
					np++;
					if (c < 0)
						c = cid[u];
					else if (cid[u] != c)
						one = false;
				}
				if (!one || np == 0 || psize[c] >= CL_MAX)
					continue;
				psize[cid[e]]--;
				cid[e] = c;
				psize[c]++;
			}
			/* members of each piece in upwind-level order */
			std::sort(order.begin(), order.end(), [&](int a, int b) {
				if (cid[a] != cid[b])
					return cid[a] < cid[b];
				const int la = lev[(size_t)v * ne + a], lb = lev[(size_t)v * ne + b];
				return la != lb ? la < lb : a < b;
			});
		}
		/* longest path over the pieces (Kahn); all placed = acyclic */
		group_graph(ncid, ne, cid, dep, ap, aj);
		clev.assign(ncid, 0);
		indeg.assign(ncid, 0);
		for (int x : aj)
			indeg[x]++;
		queue.clear();
		for (int g = 0; g < ncid; g++)
			if (!indeg[g])
				queue.push_back(g);
		for (size_t h = 0; h < queue.size(); h++) {
			const int g = queue[h];
			for (int k = ap[g]; k < ap[g + 1]; k++) {
				const int x = aj[k];
				clev[x] = std::max(clev[x], clev[g] + 1);
				if (--indeg[x] == 0)
					queue.push_back(x);
			}
		}
		if ((int)queue.size() != ncid)
			return false;
		/* pieces are runs of `order` */
		for (int k = 0; k < ne;) {
			int k1 = k;
			while (k1 < ne && cid[order[k1]] == cid[order[k]])
				k1++;
			const int g = cid[order[k]];
			cis.push_back({ clev[g], v, ckey[g], (int)mem.size(), k1 - k });
			for (int j = k; j < k1; j++)
				mem.push_back(order[j]);
			P.levels = std::max(P.levels, clev[g] + 1);
			k = k1;
		}
	}
	std::sort(cis.begin(), cis.end(), [](const CItem &a, const CItem &b) {
		if (a.lev != b.lev)
			return a.lev < b.lev;
		if (a.v != b.v)
			return a.v < b.v;
		return a.key != b.key ? a.key < b.key : a.first < b.first;
	});
	P.head.resize(cis.size());
	P.perm.resize((size_t)nv * ne);
	P.code.assign((size_t)nv * ne, 0);
	std::vector<int> jloc(ne, -1);
	long pos = 0;
	for (size_t ci = 0; ci < cis.size(); ci++) {
		const CItem &q = cis[ci];
		for (int j = 0; j < q.n; j++)
			jloc[mem[q.first + j]] = j;
		int nx = 0;
		unsigned xb[2] = { 0, 0 };
		for (int j = 0; j < q.n; j++) {
			const int e = mem[q.first + j];
			int code = 0;
			for (int f = 0; f < 4; f++) {
				const int nb = nbr[4L * e + f];
				if (nb < 0 || cls[(4L * e + f) * nv + q.v] != 2)
					continue;
				int nib;
				if (jloc[nb] >= 0 && jloc[nb] < j &&
				    mem[q.first + jloc[nb]] == nb) {
					nib = 8 + jloc[nb];
				} else {
					if (nx == CL_XMAX)
						return false;
					nib = nx;
					xb[nx / 4] |= (unsigned)(4 * j + f) << (8 * (nx % 4));
					nx++;
				}
				code |= nib << (4 * f);
			}
			/* bit 16: some downstream neighbour lies outside the cluster,
			 * so the item's values go to the hand-off words */
			for (int f = 0; f < 4; f++) {
				const int nb = nbr[4L * e + f];
				if (nb >= 0 && cls[(4L * e + f) * nv + q.v] == 1 && jloc[nb] < 0)
					code |= 1 << 16;
			}
			P.perm[pos + j] = ipos[(size_t)q.v * ne + e];
			P.code[pos + j] = code;
		}
		for (int j = 0; j < q.n; j++)
			jloc[mem[q.first + j]] = -1;
		P.head[ci] = make_int4((int)pos, q.n | nx << 8, (int)xb[0], (int)xb[1]);
		P.multi += q.n > 1 ? q.n : 0;
		pos += q.n;
	}
	return pos == (long)nv * ne;
}

bool cluster_plan(const gdh_ctx *c, const uint8_t *cls, const int *items_ve,
		  const int *level_ptr, int nlevels, ClusterPlan &P)
{
	if ((long)c->vtx.size() != 12L * c->ne || (long)c->nbr_h.size() != 4L * c->ne)
		return false;
	const std::vector<int> rng = c->bt_e0.size() >= 2 ? c->bt_e0 : std::vector<int>{ 0, c->ne };
	return cluster_plan(c->ne, c->kp.nv, c->vtx.data(), c->nbr_h.data(), rng, cls, items_ve,
			    level_ptr, nlevels, P);
}

extern "C" int gdh_cluster_plan(int ne, int nv, const float *vtx, const int *nbr,
				const int *ranges, int nranges, const uint8_t *cls,
				const int *items_ve, const int *level_ptr, int nlevels,
				int *head, int *perm, int *code, int stats[2])
{
	const std::vector<int> rng = nranges > 0 ? std::vector<int>(ranges, ranges + nranges + 1)
						 : std::vector<int>{ 0, ne };
	ClusterPlan P;
	if (!cluster_plan(ne, nv, vtx, nbr, rng, cls, items_ve, level_ptr, nlevels, P))
		return -1;
	for (size_t k = 0; k < P.head.size(); k++) {
		head[4 * k] = P.head[k].x;
		head[4 * k + 1] = P.head[k].y;
		head[4 * k + 2] = P.head[k].z;
		head[4 * k + 3] = P.head[k].w;
	}
	std::copy(P.perm.begin(), P.perm.end(), perm);
	std::copy(P.code.begin(), P.code.end(), code);
	stats[0] = P.levels;
	stats[1] = (int)P.multi;
	return (int)P.head.size();
}
