// Device-side geometry and per-particle arithmetic shared by the HIP kernels.
//
// Layout (ghosted, as the reference lays out its grid, PIC++/initialization.cpp:6618-7222):
//   N = n + 4 cells per axis; nodes 0..N (N+1 per axis), cells 0..N-1;
//   node i sits at x0 + i*dx with x0 = L - 2 dx (PIC++/initialization.cpp:194);
//   owned nodes and cells are 2..n+1 per axis.
// Grid arrays are structure-of-arrays: component c of a node field lives at
// field[c * nnodes + node_id], node_id = (i*(N1+1) + j)*(N2+1) + k.
// GMRES vectors are compact over owned nodes: vec[c * nown + (i*n1 + j)*n2 + k].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace picgpu {

constexpr int kMaxSpecies = 8;

struct Grid {
	int n[3];
	int N[3];
	double d[3];
	double x0[3];
	double L[3];
	double volume;
	// nown: unknowns of the Maxwell solve per component, nu0 * n1 * n2 owned-node columns
	long long nnodes, ncells, nown;
	// x boundary: 0 periodic (nu0 = n0), 1 reflecting wall left + constant-field free right end
	// (nu0 = n0 + 2: nodes 2 .. n0+3, the wall node 2 and the boundary nodes n0+2, n0+3)
	int wall, nu0;
	double e_right[3], b_right[3];
	// 1: one cell along z and no ghost layers there (the reference's 2-D XY grids, znumberAdded =
	// 1, PIC++/initialization.cpp:170-174): N[2] = 1, every z stencil collapses onto the plane
	// k = 0 and particles keep their z unwrapped. Explicit scheme only.
	int cz;
};

// First owned index along axis a: 2 behind the two ghost layers, 0 along a collapsed z.
__host__ __device__ inline int owned_lo(const Grid& g, int a) { return (a == 2 && g.cz) ? 0 : 2; }

struct SpeciesTable {
	int count;
	double mass[kMaxSpecies];
	double charge[kMaxSpecies];
	double weight[kMaxSpecies];
};

__host__ __device__ inline long long node_id(const Grid& g, int i, int j, int k) {
	return ((long long)i * (g.N[1] + 1) + j) * (g.N[2] + 1) + k;
}
__host__ __device__ inline long long cell_id(const Grid& g, int i, int j, int k) {
	return ((long long)i * g.N[1] + j) * g.N[2] + k;
}
__host__ __device__ inline long long own_id(const Grid& g, int i, int j, int k) {
	return ((long long)i * g.n[1] + j) * g.n[2] + k;
}
// The operators' halo layout of node data: the Maxwell unknowns (nu0 x n1 x n2 owned nodes)
// with one node on each side, halo index h <-> ghosted node h + 1 per axis.
__host__ __device__ inline long long halo_count(const Grid& g) {
	return (long long)(g.nu0 + 2) * (g.n[1] + 2) * (g.n[2] + 2);
}
__host__ __device__ inline long long halo_id(const Grid& g, int hi, int hj, int hk) {
	return ((long long)hi * (g.n[1] + 2) + hj) * (g.n[2] + 2) + hk;
}
// periodic owned index (0..n-1) of a ghosted index
__host__ __device__ inline int wrap_owned(int i, int n) {
	int r = (i - 2) % n;
	return r < 0 ? r + n : r;
}

// correlationBspline, splineOrder 0 (PIC++/correlation.cpp:798-835).
__device__ inline double bspline(double x, double hw, double left, double right) {
	if (x < left - hw) return 0.0;
	if (x > right + hw) return 0.0;
	double c;
	if (x < left + hw) {
		c = 0.5 * (x + hw - left) / hw;
	} else if (x > right - hw) {
		c = 0.5 * (right - (x - hw)) / hw;
	} else {
		c = 1.0;
	}
	if (c < 0.0 && c > -1e-14) c = 0.0;
	return c;
}

// One axis of the node and cell correlation maps
// (updateCorrelationMapNodeX / updateCorrelationMapCellX, PIC++/correlation.cpp:1112-1145, :1283-1312).
struct AxisMap {
	int nbase;     // first node index
	int cbase;     // first cell index
	double nw[2];  // node weights
	double cw[2];  // cell weights
};

__device__ inline AxisMap axis_map(double x, double x0, double d) {
	AxisMap m;
	double hw = d / 2;
	int cc = (int)floor((x - x0) / d);
	double mid = ((x0 + cc * d) + (x0 + (cc + 1) * d)) / 2;
	m.cbase = cc + ((x < mid) ? -1 : 0);
	m.cw[0] = bspline(x, hw, x0 + m.cbase * d, x0 + (m.cbase + 1) * d);
	m.cw[1] = bspline(x, hw, x0 + (m.cbase + 1) * d, x0 + (m.cbase + 2) * d);
	int nc = (int)floor(((x - x0) / d) + 0.5);
	m.nbase = nc + ((x < x0 + nc * d) ? -1 : 0);
	m.nw[0] = bspline(x, hw, x0 + (m.nbase - 0.5) * d, x0 + (m.nbase + 0.5) * d);
	m.nw[1] = bspline(x, hw, x0 + (m.nbase + 1 - 0.5) * d, x0 + (m.nbase + 1 + 0.5) * d);
	return m;
}

// The node and cell bases of axis_map alone, with the same arithmetic (so the same key).
__device__ inline void axis_bases(double x, double x0, double d, int& nbase, int& cbase) {
	const int cc = (int)floor((x - x0) / d);
	const double mid = ((x0 + cc * d) + (x0 + (cc + 1) * d)) / 2;
	cbase = cc + ((x < mid) ? -1 : 0);
	const int nc = (int)floor(((x - x0) / d) + 0.5);
	nbase = nc + ((x < x0 + nc * d) ? -1 : 0);
}

// Sort key of a particle: (species, node-map cell, octant). The octant bit per axis is
// 1 when the cell-map base equals the node-map base (upper half of the cell), so the
// particles of one key share both stencils.
__device__ inline uint32_t particle_key(const Grid& g, int sp, const AxisMap* m) {
	int bx = m[0].cbase - m[0].nbase + 1;
	int by = m[1].cbase - m[1].nbase + 1;
	int bz = m[2].cbase - m[2].nbase + 1;
	long long cid = cell_id(g, m[0].nbase, m[1].nbase, m[2].nbase);
	return (uint32_t)((((long long)sp * g.ncells + cid) << 3) | (bx << 2) | (by << 1) | bz);
}

// particle_key of the position x (axis_bases per axis)
__device__ inline uint32_t position_key(const Grid& g, int sp, const double x[3]) {
	AxisMap m[3];
	for (int a = 0; a < 3; ++a) axis_bases(x[a], g.x0[a], g.d[a], m[a].nbase, m[a].cbase);
	if (g.cz) {  // one cell along z: k = 0, octant bit 0, whatever the unwrapped z
		m[2].nbase = 0;
		m[2].cbase = -1;
	}
	return particle_key(g, sp, m);
}

// correlationGeneralEfield / correlationBfield (PIC++/correlation.cpp:23-58, :103-164):
// 8-point gathers, x outermost, in the reference's summation order.
__device__ inline void gather_node(const Grid& g, const double* __restrict__ f, const AxisMap* m, double out[3]) {
	out[0] = out[1] = out[2] = 0.0;
	for (int i = 0; i < 2; ++i)
		for (int j = 0; j < 2; ++j)
			for (int k = 0; k < 2; ++k) {
				long long id = node_id(g, m[0].nbase + i, m[1].nbase + j, m[2].nbase + k);
				double c = m[0].nw[i] * m[1].nw[j] * m[2].nw[k];
				out[0] = out[0] + f[id] * c;
				out[1] = out[1] + f[g.nnodes + id] * c;
				out[2] = out[2] + f[2 * g.nnodes + id] * c;
			}
}

__device__ inline void gather_cell(const Grid& g, const double* __restrict__ f, const AxisMap* m, double out[3]) {
	out[0] = out[1] = out[2] = 0.0;
	for (int i = 0; i < 2; ++i)
		for (int j = 0; j < 2; ++j)
			for (int k = 0; k < 2; ++k) {
				long long id = cell_id(g, m[0].cbase + i, m[1].cbase + j, m[2].cbase + k);
				double c = m[0].cw[i] * m[1].cw[j] * m[2].cw[k];
				out[0] = out[0] + f[id] * c;
				out[1] = out[1] + f[g.ncells + id] * c;
				out[2] = out[2] + f[2 * g.ncells + id] * c;
			}
}

}  // namespace picgpu
