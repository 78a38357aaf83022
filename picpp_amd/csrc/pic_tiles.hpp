// Deposit tiles shared by the FP32 and FP64 deposit kernels: z-row tiles of cells (node moments)
// and (y, z) plane tiles of dual cells (cell moments, density), with the segment lookup of a
// dual cell in the cell-sorted particle order.
#pragma once
#include "pic_device.hpp"

namespace picgpu {

struct DualTile {
	int i, j, k0;
};

__device__ inline DualTile dual_tile(const Grid& g, int T) {
	int nt = (g.N[2] + T - 1) / T;
	int b = blockIdx.x;
	DualTile t;
	t.k0 = (b % nt) * T;
	b /= nt;
	t.j = b % g.N[1];
	t.i = b / g.N[1];
	return t;
}

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Dual cells: base b collects the segments (species, cell b + a, octant 1 - a), a in {0,1}^3.
// A dual-cell workgroup serves Tj x Tk bases (i, j0 .. j0+Tj-1, k0 .. k0+Tk-1) of one x plane.
// The octants of a source cell with one x bit are contiguous (octant = x<<2 | y<<1 | z), and
// each x half feeds exactly one base plane, so a plane tile reads half cells once; only the
// tile's j and k edges are shared with its neighbours. Workgroups are renumbered so that each
// XCD (blocks are dealt to the 8 XCDs in turn) walks a contiguous run of tiles and finds the
// shared edges in its own L2.
struct PlaneTile {
	int i, j0, k0;
	bool ok;
};

__device__ inline PlaneTile plane_tile(const Grid& g, int Tj, int Tk) {
	const int nj = (g.N[1] + Tj - 1) / Tj, nk = (g.N[2] + Tk - 1) / Tk;
	const long long n = (long long)g.N[0] * nj * nk;
	const long long per = (n + 7) / 8;
	long long L = (long long)(blockIdx.x % 8) * per + blockIdx.x / 8;
	PlaneTile t;
	t.ok = L < n;
	if (!t.ok) L = 0;
	t.k0 = (int)(L % nk) * Tk;
	L /= nk;
	t.j0 = (int)(L % nj) * Tj;
	t.i = (int)(L / nj);
	return t;
}

// The 8 segments of one (dual cell, species): combined index v in [0, p[8]) lies in segment
// a = max{q : p[q] <= v}; its particle index is v + off[a].
struct DualSegments {
	int p[9];
	int off[8];
};

// The 16 loads are issued together: a source cell outside the grid reads key 0 and takes an
// empty segment (a load under a per-segment branch would wait for each in turn).
__device__ inline void dual_segments(const Grid& g, const int* __restrict__ start, int bi, int bj, int bk, int s, DualSegments& d) {
	int lo[8], hi[8];
	bool in[8];
#pragma unroll
	for (int a = 0; a < 8; ++a) {
		const int ax = (a >> 2) & 1, ay = (a >> 1) & 1, az = a & 1;
		const int ci = bi + ax, cj = bj + ay, ck = bk + az;
		in[a] = ci < g.N[0] && cj < g.N[1] && ck < g.N[2];
		const long long key = in[a] ? ((((long long)s * g.ncells + cell_id(g, ci, cj, ck)) << 3) | (7 - a)) : 0;
		lo[a] = start[key];
		hi[a] = start[key + 1];
	}
	d.p[0] = 0;
#pragma unroll
	for (int a = 0; a < 8; ++a) {
		const int l = in[a] ? lo[a] : 0, len = in[a] ? hi[a] - lo[a] : 0;
		d.p[a + 1] = d.p[a] + len;
		d.off[a] = l - d.p[a];
	}
}

// A box of a field's values staged in LDS: extent e[3] from the origin o[3] (ghosted indices),
// 3 components per point.
struct Box {
	int o[3], e[3];
	__device__ int at(int i, int j, int k) const { return ((i - o[0]) * e[1] + (j - o[1])) * e[2] + (k - o[2]); }
	__device__ bool holds(int i, int j, int k) const {
		return i >= o[0] && i + 1 < o[0] + e[0] && j >= o[1] && j + 1 < o[1] + e[1] && k >= o[2] && k + 1 < o[2] + e[2];
	}
};

// Stages the box (indices clamped into [0, lim]) of a component-plane field into LDS.
__device__ inline void stage_box(const Box& b, const double* __restrict__ f, long long plane, int lim0, int lim1, int lim2,
                                 bool nodes, const Grid& g, double* lds) {
	const int n = b.e[0] * b.e[1] * b.e[2];
	for (int q = threadIdx.x; q < n; q += blockDim.x) {
		const int dk = q % b.e[2], dj = (q / b.e[2]) % b.e[1], di = q / (b.e[2] * b.e[1]);
		const int i = clampi(b.o[0] + di, 0, lim0), j = clampi(b.o[1] + dj, 0, lim1), k = clampi(b.o[2] + dk, 0, lim2);
		const long long id = nodes ? node_id(g, i, j, k) : cell_id(g, i, j, k);
		lds[3 * q + 0] = f[id];
		lds[3 * q + 1] = f[plane + id];
		lds[3 * q + 2] = f[2 * plane + id];
	}
}

}  // namespace picgpu
