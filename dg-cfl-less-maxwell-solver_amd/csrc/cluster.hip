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
#include <thread>

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

/* a stable counting sort of the elements in `in` by key[e] (keys below
 * nkey): `out` lists them by key, in their order in `in` for equal keys */
static void sort_by_key(const std::vector<int> &in, const std::vector<int> &key, int nkey,
			std::vector<int> &cnt, std::vector<int> &out)
{
	cnt.assign((size_t)nkey + 1, 0);
	for (const int e : in)
		cnt[key[e] + 1]++;
	for (int k = 0; k < nkey; k++)
		cnt[k + 1] += cnt[k];
	out.resize(in.size());
	for (const int e : in)
		out[cnt[key[e]]++] = e;
}

struct CItem {
	int lev, v, key, first, n;
};

/* The clusters of one velocity: pieces of the base clusters (single
 * elements where a base cluster lies on a cycle of the cluster graph),
 * their members in upwind-level order (mem, from cis[].first) and their
 * levels in the cluster graph. lev: upwind level of every item (v, e).
 * False when the cluster graph is cyclic. */
static bool velocity_clusters(int v, int ne, int nv, int nbase, const std::vector<int> &bc,
			      const int *nbr, const uint8_t *cls, const std::vector<int> &lev,
			      int nlevels, std::vector<CItem> &cis, std::vector<int> &mem,
			      int &levels)
{
	/* the producer across face f of element e, or -1 */
	auto dep = [&](int e, int f) -> int {
		const int nb = nbr[4L * e + f];
		return nb >= 0 && cls[(4L * e + f) * nv + v] == 2 ? nb : -1;
	};
	std::vector<int> ap, aj, comp, grp(ne), cid(ne), order, lorder(ne), ckey, cnt;
	std::vector<int> clev, indeg, queue;
	/* the elements in upwind-level order (level, then element index) */
	const std::vector<int> lv(lev.begin() + (size_t)v * ne, lev.begin() + (size_t)(v + 1) * ne);
	{
		std::vector<int> all(ne);
		std::iota(all.begin(), all.end(), 0);
		sort_by_key(all, lv, nlevels, cnt, lorder);
	}
	group_graph(nbase, ne, bc, dep, ap, aj);
	const std::vector<int> csize = scc(nbase, ap, aj, comp);
	/* groups: base clusters, or single elements inside a cycle */
	for (int e = 0; e < ne; e++)
		grp[e] = csize[comp[bc[e]]] > 1 ? nbase + e : bc[e];
	/* by group, each group in upwind-level order */
	sort_by_key(lorder, grp, nbase + ne, cnt, order);
	/* cut each group in level order: at most CL_MAX elements and
	 * CL_XMAX inflow faces from outside the piece */
	int ncid = 0, nin = 0, nxc = 0; /* current piece: elements, faces from outside */
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
		if (!same || nin == CL_MAX || nxc + ex > CL_XMAX) {
			ncid++;
			ckey.push_back(bc[e]);
			nin = 0;
			nxc = 0;
			ex = all;
		}
		cid[e] = ncid - 1;
		nin++;
		nxc += ex;
	}
	/* flow-following merge, in upwind-level order: a single-element
	 * piece whose producers all lie in one piece of fewer than CL_MAX
	 * elements joins it. Its in-edges then all come from inside that
	 * piece, so a cycle through the merged piece would be one through
	 * the piece and the single element before: none can form. This
	 * recovers clusters where grid cells cut across the flow
	 * (unstructured meshes). */
	std::vector<int> psize(ncid, 0);
	for (int e = 0; e < ne; e++)
		psize[cid[e]]++;
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
	sort_by_key(lorder, cid, ncid, cnt, order);
	/* longest path over the pieces (Kahn); all placed = acyclic */
	group_graph(ncid, ne, cid, dep, ap, aj);
	clev.assign(ncid, 0);
	indeg.assign(ncid, 0);
	for (int x : aj)
		indeg[x]++;
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
	levels = 0;
	for (int k = 0; k < ne;) {
		int k1 = k;
		while (k1 < ne && cid[order[k1]] == cid[order[k]])
			k1++;
		const int g = cid[order[k]];
		cis.push_back({ clev[g], v, ckey[g], (int)mem.size(), k1 - k });
		for (int j = k; j < k1; j++)
			mem.push_back(order[j]);
		levels = std::max(levels, clev[g] + 1);
		k = k1;
	}
	return true;
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
	/* the velocities' clusters, one host thread each (the work of one
	 * velocity is independent of the others'), joined in velocity order */
	std::vector<std::vector<CItem>> cv(nv);
	std::vector<std::vector<int>> mv(nv);
	std::vector<int> lv(nv, 0);
	std::vector<char> ok(nv, 0);
	{
		std::vector<std::thread> th;
		for (int v = 0; v < nv; v++)
			th.emplace_back([&, v] {
				ok[v] = velocity_clusters(v, ne, nv, nbase, bc, nbr, cls, lev, nlevels,
							  cv[v], mv[v], lv[v]);
			});
		for (auto &t : th)
			t.join();
	}
	std::vector<CItem> cis;
	std::vector<int> mem;
	P.levels = 0;
	P.multi = 0;
	for (int v = 0; v < nv; v++) {
		if (!ok[v])
			return false;
		const int off = (int)mem.size();
		for (CItem q : cv[v]) {
			q.first += off;
			cis.push_back(q);
		}
		mem.insert(mem.end(), mv[v].begin(), mv[v].end());
		P.levels = std::max(P.levels, lv[v]);
	}
	/* order (level, velocity, base cluster, first member): cis is in
	 * first-member order, so stable counting sorts by the other three keys
	 * from the last to the first give it */
	{
		std::vector<int> idx(cis.size()), k1(cis.size()), cnt, tmp;
		std::iota(idx.begin(), idx.end(), 0);
		for (size_t i = 0; i < cis.size(); i++)
			k1[i] = cis[i].key;
		sort_by_key(idx, k1, nbase, cnt, tmp);
		for (size_t i = 0; i < cis.size(); i++)
			k1[i] = cis[i].v;
		sort_by_key(tmp, k1, nv, cnt, idx);
		for (size_t i = 0; i < cis.size(); i++)
			k1[i] = cis[i].lev;
		sort_by_key(idx, k1, P.levels, cnt, tmp);
		std::vector<CItem> sorted(cis.size());
		for (size_t i = 0; i < cis.size(); i++)
			sorted[i] = cis[tmp[i]];
		cis.swap(sorted);
	}
	const size_t ncl = cis.size();
	P.head.resize(ncl);
	P.perm.resize((size_t)nv * ne);
	P.code.assign((size_t)nv * ne, 0);
	std::vector<long> start(ncl + 1, 0);
	for (size_t ci = 0; ci < ncl; ci++)
		start[ci + 1] = start[ci] + cis[ci].n;
	/* each cluster's face codes and header, on the host's threads (ranges
	 * of clusters; each thread marks its cluster's members in its own
	 * element array) */
	const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
	std::vector<char> fine(nth, 1);
	std::vector<long> multi(nth, 0);
	auto codes = [&](int w) {
		std::vector<int> jloc(ne, -1);
		const size_t c0 = ncl * w / nth, c1 = ncl * (w + 1) / nth;
		for (size_t ci = c0; ci < c1; ci++) {
			const CItem &q = cis[ci];
			const long pos = start[ci];
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
						if (nx == CL_XMAX) {
							fine[w] = 0;
							return;
						}
						nib = nx;
						xb[nx / 4] |= (unsigned)(4 * j + f) << (8 * (nx % 4));
						nx++;
					}
					code |= nib << (4 * f);
				}
				/* bit 16: some downstream neighbour lies outside the
				 * cluster, so the item's values go to the hand-off
				 * words */
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
			multi[w] += q.n > 1 ? q.n : 0;
		}
	};
	{
		std::vector<std::thread> th;
		for (int w = 0; w < nth; w++)
			th.emplace_back(codes, w);
		for (auto &t : th)
			t.join();
	}
	for (int w = 0; w < nth; w++) {
		if (!fine[w])
			return false;
		P.multi += multi[w];
	}
	return start[ncl] == (long)nv * ne;
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
