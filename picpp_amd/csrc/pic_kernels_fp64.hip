// FP64 kernels of the implicit step (the path held to the reference's arithmetic).
//
// Particles are stored structure-of-arrays and sorted by key = (species, node-map
// cell, octant); `start[key]` is the first particle of each key. Deposition is
// deterministic without atomics: one lane owns one cell (node-map moments) or one
// dual cell (cell-map moments), walks its particles in sorted order and keeps the
// 8 corner sums in registers; a second pass gathers the 8 corner sums of the cells
// around each grid point in a fixed order.
//
// FP contraction is off in this file so the per-particle arithmetic follows the
// reference operation by operation.
#pragma clang fp contract(off)

#include "pic_device.hpp"
#include "pic_kernels.hpp"
#include "pic_recip.hpp"

namespace picgpu {

constexpr double kPi = 3.141592653589793115997963468544185161590576171875;  // 4*atan2(1,1)
constexpr double kCprime = 1.025;                                              // PIC++/constants.h:45
constexpr double kCprime2 = 1.025 * 1.025;

__device__ inline void load_particle(const ParticleArrays& P, long long i, double x[3], double p[3]) {
	x[0] = P.x[0][i];
	x[1] = P.x[1][i];
	x[2] = P.x[2][i];
	p[0] = P.p[0][i];
	p[1] = P.p[1][i];
	p[2] = P.p[2][i];
}

// Maps of the three axes. Stencil bases are clamped into the ghosted arrays so that no
// access can leave them; returns false when a clamp was needed (the particle was
// further than the ghost layers allow, which the callers report as an error).
__device__ inline bool maps3(const Grid& g, const double x[3], AxisMap m[3]) {
	bool ok = true;
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		m[a] = axis_map(x[a], g.x0[a], g.d[a]);
		if (!(m[a].nbase >= 0 && m[a].nbase <= g.N[a] - 1)) {
			m[a].nbase = m[a].nbase < 0 ? 0 : g.N[a] - 1;
			ok = false;
		}
		if (!(m[a].cbase >= 0 && m[a].cbase <= g.N[a] - 2)) {
			m[a].cbase = m[a].cbase < 0 ? 0 : g.N[a] - 2;
			ok = false;
		}
	}
	return ok;
}

// ------------------------------------------------------------------ moments at x^n
// Node values from the corner sums of the 8 cells around each node, written to the
// ghosted node arrays J (3 planes) and mu (9 planes).
__global__ void k_node_assemble(Grid g, const double* __restrict__ cs, double* __restrict__ J, double* __restrict__ mu) {
	long long nid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (nid >= g.nnodes) return;
	int k = (int)(nid % (g.N[2] + 1));
	int j = (int)((nid / (g.N[2] + 1)) % (g.N[1] + 1));
	int i = (int)(nid / ((long long)(g.N[2] + 1) * (g.N[1] + 1)));
	double s[12];
#pragma unroll
	for (int q = 0; q < 12; ++q) s[q] = 0.0;
	for (int a = 0; a < 2; ++a)
		for (int b = 0; b < 2; ++b)
			for (int cc = 0; cc < 2; ++cc) {
				int ci = i - a, cj = j - b, ck = k - cc;
				if (ci < 0 || cj < 0 || ck < 0 || ci >= g.N[0] || cj >= g.N[1] || ck >= g.N[2]) continue;
				long long cid = cell_id(g, ci, cj, ck);
				int corner = a * 4 + b * 2 + cc;
#pragma unroll
				for (int q = 0; q < 12; ++q) s[q] += cs[((long long)q * 8 + corner) * g.ncells + cid];
			}
#pragma unroll
	for (int q = 0; q < 3; ++q) J[q * g.nnodes + nid] = s[q];
#pragma unroll
	for (int q = 0; q < 9; ++q) mu[q * g.nnodes + nid] = s[3 + q];
}

// rhohat per cell from the corner sums of the 8 dual cells around it.
__global__ void k_cell_assemble(Grid g, const double* __restrict__ rs, double* __restrict__ rho) {
	long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.ncells) return;
	int k = (int)(c % g.N[2]);
	int j = (int)((c / g.N[2]) % g.N[1]);
	int i = (int)(c / ((long long)g.N[2] * g.N[1]));
	double s = 0.0;
	for (int a = 0; a < 2; ++a)
		for (int b = 0; b < 2; ++b)
			for (int cc = 0; cc < 2; ++cc) {
				int bi = i - a, bj = j - b, bk = k - cc;
				if (bi < 0 || bj < 0 || bk < 0) continue;
				s += rs[(long long)(a * 4 + b * 2 + cc) * g.ncells + cell_id(g, bi, bj, bk)];
			}
	rho[c] = s;
}

// Pressure tensor at [i][j][k], j, k in {0, 1}: sum over every dual cell b with
// b.x + ti == i of its (ti, j, k) partial. One block per (i, j, k); fixed-order tree.
__global__ void __launch_bounds__(256) k_pressure_reduce(Grid g, const double* __restrict__ ps, double* __restrict__ Pi) {
	int i = blockIdx.x / 4;
	int tj = (blockIdx.x / 2) % 2;
	int tk = blockIdx.x % 2;
	__shared__ double red[6][256];
	double acc[6] = {0, 0, 0, 0, 0, 0};
	long long plane = (long long)g.N[1] * g.N[2];
	for (int ti = 0; ti < 2; ++ti) {
		int bi = i - ti;
		if (bi < 0 || bi >= g.N[0]) continue;
		int t = ti * 4 + tj * 2 + tk;
		for (long long r = threadIdx.x; r < plane; r += blockDim.x) {
			long long b = (long long)bi * plane + r;
#pragma unroll
			for (int q = 0; q < 6; ++q) acc[q] += ps[((long long)t * 6 + q) * g.ncells + b];
		}
	}
#pragma unroll
	for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = acc[q];
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < w)
#pragma unroll
			for (int q = 0; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		long long c = cell_id(g, i, tj, tk);
		const int sym[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
		for (int q = 0; q < 9; ++q) Pi[q * g.ncells + c] = red[sym[q]][0];
	}
}

// Cell-local pressure tensor (PIC_PRESSURE_CELL): Pi at cell c from the partials of the 8
// dual cells around it, the deposit the reference's pressure loop intends
// (PIC++/momentum_collecting.cpp:565, :577 with [curI][curJ][curK]), assembled as rhohat is.
template <typename T>
__global__ void k_pressure_cell(Grid g, const T* __restrict__ ps, double* __restrict__ Pi) {
	long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.ncells) return;
	int k = (int)(c % g.N[2]);
	int j = (int)((c / g.N[2]) % g.N[1]);
	int i = (int)(c / ((long long)g.N[2] * g.N[1]));
	double acc[6] = {0, 0, 0, 0, 0, 0};
	for (int a = 0; a < 2; ++a)
		for (int b = 0; b < 2; ++b)
			for (int cc = 0; cc < 2; ++cc) {
				int bi = i - a, bj = j - b, bk = k - cc;
				if (bi < 0 || bj < 0 || bk < 0) continue;
				const long long src = cell_id(g, bi, bj, bk);
				const int t = a * 4 + b * 2 + cc;
#pragma unroll
				for (int q = 0; q < 6; ++q) acc[q] += (double)ps[((long long)t * 6 + q) * g.ncells + src];
			}
	const int sym[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
#pragma unroll
	for (int q = 0; q < 9; ++q) Pi[q * g.ncells + c] = acc[sym[q]];
}

// ------------------------------------------------------------------ folds and ghosts
// One-rank periodic folds, one axis per launch (momentum_exchange.cpp:172-195, :446-460,
// :807-811, :1657-1660). Node arrays: pairs (N - i, 4 - i), i = 0..4. Cell arrays: pairs
// (N-1-i, lo-i), i < pairs. Every thread owns one (line, pair, component).
// Thread order: along x and y folds the z index varies fastest (the planes' rows are
// contiguous); along z the pair index, then y.
__global__ void k_fold(Grid g, double* __restrict__ f, int ncomp, int axis, int nodes, int pairs, int hi0, int lo0) {
	int ext[3];
	for (int a = 0; a < 3; ++a) ext[a] = g.N[a] + (nodes ? 1 : 0);
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	const int p1 = axis == 0 ? 1 : 0;  // the slower of the two plane axes (x or y)
	long long total = (long long)ext[p1] * ext[axis == 2 ? 1 : 2] * pairs * ncomp;
	if (t >= total) return;
	int hi[3], lo[3], pr;
	if (axis == 2) {
		pr = (int)(t % pairs);
		long long r = t / pairs;
		hi[1] = lo[1] = (int)(r % ext[1]);
		r /= ext[1];
		hi[0] = lo[0] = (int)(r % ext[0]);
		t = r / ext[0];
	} else {
		hi[2] = lo[2] = (int)(t % ext[2]);
		long long r = t / ext[2];
		hi[p1] = lo[p1] = (int)(r % ext[p1]);
		r /= ext[p1];
		pr = (int)(r % pairs);
		t = r / pairs;
	}
	const int comp = (int)t;
	hi[axis] = hi0 - pr;
	lo[axis] = lo0 - pr;
	long long len = nodes ? g.nnodes : g.ncells;
	long long ih = nodes ? node_id(g, hi[0], hi[1], hi[2]) : cell_id(g, hi[0], hi[1], hi[2]);
	long long il = nodes ? node_id(g, lo[0], lo[1], lo[2]) : cell_id(g, lo[0], lo[1], lo[2]);
	double v = f[comp * len + ih] + f[comp * len + il];
	f[comp * len + ih] = v;
	f[comp * len + il] = v;
}

// Periodic ghost fill of node (exchangeGeneralEfield, fields_exchange.cpp:158-170) or cell
// (exchangeGeneralBfield, :629-637) arrays: every ghost takes its owned periodic image.
// With the wall boundary the x ghosts behind the wall copy index 2 (fields_exchange.cpp:171-183,
// :639-651) and the ghosts beyond the free right end keep their values (y and z stay periodic).
// With wall_zero the ghosts behind the wall are zeroed instead (exchangeGeneralScalarCellFieldX,
// one-rank SUPER_CONDUCTOR_LEFT branch, fields_exchange.cpp:1468-1478).
__global__ void k_ghost_fill(Grid g, double* __restrict__ f, int ncomp, int nodes, int wall_zero) {
	long long len = nodes ? g.nnodes : g.ncells;
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= len) return;
	int e1 = g.N[1] + (nodes ? 1 : 0), e2 = g.N[2] + (nodes ? 1 : 0);
	int k = (int)(t % e2);
	int j = (int)((t / e2) % e1);
	int i = (int)(t / ((long long)e1 * e2));
	bool ghost = i < 2 || j < 2 || k < 2 || i > g.n[0] + 1 || j > g.n[1] + 1 || k > g.n[2] + 1;
	if (!ghost) return;
	if (g.wall && wall_zero && i < 2) {
		for (int c = 0; c < ncomp; ++c) f[c * len + t] = 0.0;
		return;
	}
	int si = g.wall ? (i < 2 ? 2 : i) : wrap_owned(i, g.n[0]) + 2;
	int sj = wrap_owned(j, g.n[1]) + 2, sk = wrap_owned(k, g.n[2]) + 2;
	if (si == i && sj == j && sk == k) return;
	long long src = nodes ? node_id(g, si, sj, sk) : cell_id(g, si, sj, sk);
	for (int c = 0; c < ncomp; ++c) f[c * len + t] = f[c * len + src];
}

// ------------------------------------------------------------------ smoothing (PIC++/smoothing.cpp)
// One pass of the 27-point kernel m = f(a) f(b) f(c), f = (s/2, 1-s, s/2), into `out` over the
// reference's loop ranges; the boundary planes it leaves unsmoothed are copied. A second launch
// with `back` set copies the pass back into f over the same range.
// Cells (smoothCellParameter, smoothing.cpp:20-93): i in [2, N0-2], j, k in [2, N-3]; with the
// wall the planes i = 2 and i = N0-2 are copied.
__global__ void k_smooth_cells(Grid g, double* __restrict__ in, double* __restrict__ out, Smooth27 sm, int back) {
	const int ni = g.N[0] - 3, nj = g.n[1], nk = g.n[2];
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= (long long)ni * nj * nk) return;
	int k = 2 + (int)(t % nk);
	int j = 2 + (int)((t / nk) % nj);
	int i = 2 + (int)(t / ((long long)nk * nj));
	long long c = cell_id(g, i, j, k);
	if (back) {
		in[c] = out[c];
		return;
	}
	double acc;
	if (!g.wall || (i > 2 && i < g.N[0] - 2)) {
		acc = 0.0;
		for (int a = -1; a <= 1; ++a)
			for (int b = -1; b <= 1; ++b)
				for (int d = -1; d <= 1; ++d) acc += in[cell_id(g, i + a, j + b, k + d)] * sm.m[(a + 1) * 9 + (b + 1) * 3 + (d + 1)];
	} else {
		acc = in[c];
	}
	out[c] = acc;
}

// Nodes (smoothVectorNodeParameter, smoothMatrixNodeParameter, smoothing.cpp:139-364): i in
// [2, N0-2] (from 3 behind a wall), j, k in [2, N-2]; with the wall the plane i = N0-2 is copied.
// ncomp component planes.
__global__ void k_smooth_nodes(Grid g, double* __restrict__ in, double* __restrict__ out, int ncomp, Smooth27 sm, int back) {
	const int i0 = g.wall ? 3 : 2;
	const int ni = g.N[0] - 1 - i0, nj = g.n[1] + 1, nk = g.n[2] + 1;
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= (long long)ni * nj * nk * ncomp) return;
	const long long pts = (long long)ni * nj * nk;
	int comp = (int)(t / pts);
	long long r = t % pts;
	int k = 2 + (int)(r % nk);
	int j = 2 + (int)((r / nk) % nj);
	int i = i0 + (int)(r / ((long long)nk * nj));
	const double* f = in + (long long)comp * g.nnodes;
	long long o = (long long)comp * g.nnodes + node_id(g, i, j, k);
	if (back) {
		in[o] = out[o];
		return;
	}
	double acc;
	if (!(g.wall && i == g.N[0] - 2)) {
		acc = 0.0;
		for (int a = -1; a <= 1; ++a)
			for (int b = -1; b <= 1; ++b)
				for (int d = -1; d <= 1; ++d) acc += f[node_id(g, i + a, j + b, k + d)] * sm.m[(a + 1) * 9 + (b + 1) * 3 + (d + 1)];
	} else {
		acc = f[node_id(g, i, j, k)];
	}
	out[o] = acc;
}

// ------------------------------------------------------------------ wall boundary (PIC_BOUNDARY_WALL)
// Cell-map deposits that land behind the wall go to their mirror cells: cell c <= 1 -> 3 - c
// (momentum_collecting.cpp:557-567 for rhohat, :570-574 for the pressure tensor, :885-898
// for the post-push density). One thread per (mirror pair, y-z column, component); `flip`
// marks components that change sign under the reflection (the x momentum sums).
__global__ void k_wall_mirror_cells(Grid g, double* __restrict__ f, int ncomp, const int* __restrict__ flip) {
	long long cols = (long long)g.N[1] * g.N[2];
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= cols * ncomp) return;
	int comp = (int)(t / cols);
	long long col = t % cols;
	long long plane = cols;
	double sgn = (flip && flip[comp]) ? -1.0 : 1.0;
	double* fc = f + (long long)comp * g.ncells;
	for (int c = 1; c >= 0; --c) {
		fc[(3 - c) * plane + col] += sgn * fc[c * plane + col];
		fc[c * plane + col] = 0.0;
	}
}

// Right end of the post-push density (sumCellParametersX / sumCellVectorParametersX, non-periodic
// branch, momentum_exchange.cpp:960-985, :1340-1350): cells N-2, N-1 add into N-3; the zeroed
// planes are those with zero_mask set (charge, concentration, energy; not the momentum sums).
__global__ void k_wall_right_fold(Grid g, double* __restrict__ f, int ncomp, const int* __restrict__ zero_mask) {
	long long cols = (long long)g.N[1] * g.N[2];
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= cols * ncomp) return;
	int comp = (int)(t / cols);
	long long col = t % cols;
	double* fc = f + (long long)comp * g.ncells;
	const int N0 = g.N[0];
	for (int i = 0; i <= 1; ++i) {
		fc[(long long)(N0 - 3) * cols + col] += fc[(long long)(N0 - 2 + i) * cols + col];
		if (zero_mask[comp]) fc[(long long)(N0 - 2 + i) * cols + col] = 0.0;
	}
}

// Zeroes components [c0, c1) of a node array on the x planes i0 .. i1 (all y, z).
__global__ void k_zero_node_planes(Grid g, double* __restrict__ f, int c0, int c1, int i0, int i1) {
	long long plane = (long long)(g.N[1] + 1) * (g.N[2] + 1);
	long long n = plane * (i1 - i0 + 1) * (c1 - c0);
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	int comp = c0 + (int)(t / (plane * (i1 - i0 + 1)));
	long long r = t % (plane * (i1 - i0 + 1));
	f[(long long)comp * g.nnodes + (long long)i0 * plane + r] = 0.0;
}

// ------------------------------------------------------------------ moment post-processing
// evaluateDivPressureTensor (fields.cpp:3158-3223) and J -= eta dt divPi (momentum_collecting.cpp:623-636).
__global__ void k_div_pressure_update_j(Grid g, const double* __restrict__ Pi, double* __restrict__ J, double* __restrict__ divPi,
                                        double eta, double dt) {
	long long nid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (nid >= g.nnodes) return;
	int k = (int)(nid % (g.N[2] + 1));
	int j = (int)((nid / (g.N[2] + 1)) % (g.N[1] + 1));
	int i = (int)(nid / ((long long)(g.N[2] + 1) * (g.N[1] + 1)));
	double out[3] = {0.0, 0.0, 0.0};
	if (!(i == 0 || i == g.N[0] || j == 0 || j == g.N[1] || k == 0 || k == g.N[2])) {
		auto Pv = [&](int a, int b, int c, int l, int m) { return Pi[(l * 3 + m) * g.ncells + cell_id(g, a, b, c)]; };
		for (int col = 0; col < 3; ++col) {
			double rx = (Pv(i, j, k, 0, col) + Pv(i, j - 1, k, 0, col) + Pv(i, j, k - 1, 0, col) + Pv(i, j - 1, k - 1, 0, col)) * 0.25;
			double lx = (Pv(i - 1, j, k, 0, col) + Pv(i - 1, j - 1, k, 0, col) + Pv(i - 1, j, k - 1, 0, col) +
			             Pv(i - 1, j - 1, k - 1, 0, col)) * 0.25;
			double by = (Pv(i, j, k, 1, col) + Pv(i - 1, j, k, 1, col) + Pv(i, j, k - 1, 1, col) + Pv(i - 1, j, k - 1, 1, col)) * 0.25;
			double fy = (Pv(i, j - 1, k, 1, col) + Pv(i - 1, j - 1, k, 1, col) + Pv(i, j - 1, k - 1, 1, col) +
			             Pv(i - 1, j - 1, k - 1, 1, col)) * 0.25;
			double tz = (Pv(i, j, k, 2, col) + Pv(i, j - 1, k, 2, col) + Pv(i - 1, j, k, 2, col) + Pv(i - 1, j - 1, k, 2, col)) * 0.25;
			double bz = (Pv(i, j, k - 1, 2, col) + Pv(i, j - 1, k - 1, 2, col) + Pv(i - 1, j, k - 1, 2, col) +
			             Pv(i - 1, j - 1, k - 1, 2, col)) * 0.25;
			out[col] = ((rx - lx) / g.d[0]) + ((by - fy) / g.d[1]) + ((tz - bz) / g.d[2]);
		}
	}
	for (int c = 0; c < 3; ++c) {
		divPi[c * g.nnodes + nid] = out[c];
		J[c * g.nnodes + nid] = J[c * g.nnodes + nid] - out[c] * eta * dt;
	}
}

// evaluateDivFlux (fields.cpp:3060-3156) and rhohat -= dt theta divJ (momentum_collecting.cpp:651-665).
__global__ void k_div_flux_update_rho(Grid g, const double* __restrict__ J, double* __restrict__ rho, double dt, double theta) {
	long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.ncells) return;
	int k = (int)(c % g.N[2]);
	int j = (int)((c / g.N[2]) % g.N[1]);
	int i = (int)(c / ((long long)g.N[2] * g.N[1]));
	auto F = [&](int a, int b, int cc, int l) { return J[l * g.nnodes + node_id(g, a, b, cc)]; };
	double rX = (F(i + 1, j, k, 0) + F(i + 1, j + 1, k, 0) + F(i + 1, j, k + 1, 0) + F(i + 1, j + 1, k + 1, 0)) * 0.25;
	double lX = (F(i, j, k, 0) + F(i, j + 1, k, 0) + F(i, j, k + 1, 0) + F(i, j + 1, k + 1, 0)) * 0.25;
	double rY = (F(i, j + 1, k, 1) + F(i + 1, j + 1, k, 1) + F(i, j + 1, k + 1, 1) + F(i + 1, j + 1, k + 1, 1)) * 0.25;
	double lY = (F(i, j, k, 1) + F(i + 1, j, k, 1) + F(i, j, k + 1, 1) + F(i + 1, j, k + 1, 1)) * 0.25;
	double rZ = (F(i + 1, j, k + 1, 2) + F(i, j, k + 1, 2) + F(i, j + 1, k + 1, 2) + F(i + 1, j + 1, k + 1, 2)) * 0.25;
	double lZ = (F(i + 1, j, k, 2) + F(i, j, k, 2) + F(i, j + 1, k, 2) + F(i + 1, j + 1, k, 2)) * 0.25;
	double divJ = ((rX - lX) / g.d[0]) + ((rY - lY) / g.d[1]) + ((rZ - lZ) / g.d[2]);
	rho[c] -= dt * theta * divJ;
}

// ------------------------------------------------------------------ Maxwell system
// Right part of createInternalEquation (fields.cpp:2635-2715) with evaluateRotB (:2799-2860)
// and evaluateGradDensity (:3229-3259), for owned nodes, into the compact GMRES layout.
// With the wall boundary the unknowns run over x nodes 2 .. n0+3: node 2 carries the
// super-conducting wall rows (createSuperConductorLeftEquation, fields.cpp:690-767: the x row
// E_x(2) - E_x(3) = -pi (sum of the 4 adjacent rhohat), tangential rows E = 0), nodes n0+2 and
// n0+3 the free right rows (createFreeRightEquation, :769-781: E = e_right).
__global__ void k_rhs(Grid g, const double* __restrict__ E, const double* __restrict__ B, const double* __restrict__ J,
                      const double* __restrict__ rho, double theta, double dt, double* __restrict__ rhs) {
	long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= g.nown) return;
	int k = (int)(o % g.n[2]) + 2;
	int j = (int)((o / g.n[2]) % g.n[1]) + 2;
	int i = (int)(o / ((long long)g.n[2] * g.n[1])) + 2;
	if (g.wall && i == 2) {
		auto r = [&](int a, int bb, int c) { return rho[cell_id(g, a, bb, c)]; };
		rhs[o] = -kPi * (r(i, j, k) + r(i, j - 1, k) + r(i, j, k - 1) + r(i, j - 1, k - 1));
		rhs[g.nown + o] = 0.0;
		rhs[2 * g.nown + o] = 0.0;
		return;
	}
	if (g.wall && i >= g.n[0] + 2) {
		for (int a = 0; a < 3; ++a) rhs[a * g.nown + o] = g.e_right[a];
		return;
	}
	auto b = [&](int a, int bb, int c, int l) { return B[l * g.ncells + cell_id(g, a, bb, c)]; };
	double rX[3], lX[3], rY[3], lY[3], rZ[3], lZ[3];
	for (int l = 0; l < 3; ++l) {
		rX[l] = (b(i, j, k, l) + b(i, j - 1, k, l) + b(i, j, k - 1, l) + b(i, j - 1, k - 1, l)) / 4.0;
		lX[l] = (b(i - 1, j, k, l) + b(i - 1, j - 1, k, l) + b(i - 1, j, k - 1, l) + b(i - 1, j - 1, k - 1, l)) / 4.0;
		rY[l] = (b(i, j, k, l) + b(i - 1, j, k, l) + b(i, j, k - 1, l) + b(i - 1, j, k - 1, l)) / 4.0;
		lY[l] = (b(i, j - 1, k, l) + b(i - 1, j - 1, k, l) + b(i, j - 1, k - 1, l) + b(i - 1, j - 1, k - 1, l)) / 4.0;
		rZ[l] = (b(i, j, k, l) + b(i - 1, j, k, l) + b(i, j - 1, k, l) + b(i - 1, j - 1, k, l)) / 4.0;
		lZ[l] = (b(i, j, k - 1, l) + b(i - 1, j, k - 1, l) + b(i, j - 1, k - 1, l) + b(i - 1, j - 1, k - 1, l)) / 4.0;
	}
	double rb[3];
	rb[0] = (rY[2] - lY[2]) / g.d[1] - (rZ[1] - lZ[1]) / g.d[2];
	rb[1] = (rZ[0] - lZ[0]) / g.d[2] - (rX[2] - lX[2]) / g.d[0];
	rb[2] = (rX[1] - lX[1]) / g.d[0] - (rY[0] - lY[0]) / g.d[1];
	auto r = [&](int a, int bb, int c) { return rho[cell_id(g, a, bb, c)]; };
	double dRX = (r(i, j, k) + r(i, j - 1, k) + r(i, j, k - 1) + r(i, j - 1, k - 1)) * 0.25;
	double dLX = (r(i - 1, j, k) + r(i - 1, j - 1, k) + r(i - 1, j, k - 1) + r(i - 1, j - 1, k - 1)) * 0.25;
	double dRY = (r(i, j, k) + r(i - 1, j, k) + r(i, j, k - 1) + r(i - 1, j, k - 1)) * 0.25;
	double dLY = (r(i, j - 1, k) + r(i - 1, j - 1, k) + r(i, j - 1, k - 1) + r(i - 1, j - 1, k - 1)) * 0.25;
	double dRZ = (r(i, j, k) + r(i, j - 1, k) + r(i - 1, j, k) + r(i - 1, j - 1, k)) * 0.25;
	double dLZ = (r(i, j, k - 1) + r(i, j - 1, k - 1) + r(i - 1, j, k - 1) + r(i - 1, j - 1, k - 1)) * 0.25;
	double gr[3] = {(dRX - dLX) / g.d[0], (dRY - dLY) / g.d[1], (dRZ - dLZ) / g.d[2]};
	long long nid = node_id(g, i, j, k);
	for (int a = 0; a < 3; ++a) {
		double t1 = (rb[a] * kCprime - (J[a * g.nnodes + nid] * 4 * kPi)) * (theta * dt);
		double t2 = gr[a] * kCprime2 * theta * theta * dt * dt * 4 * kPi;
		rhs[a * g.nown + o] = (E[a * g.nnodes + nid] + t1) - t2;
	}
}

// Copy of the folded dielectric tensor at the unknowns' nodes (9 planes, compact layout).
__global__ void k_pack_mu(Grid g, const double* __restrict__ mu, double* __restrict__ muc) {
	long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= g.nown) return;
	int k = (int)(o % g.n[2]) + 2;
	int j = (int)((o / g.n[2]) % g.n[1]) + 2;
	int i = (int)(o / ((long long)g.n[2] * g.n[1])) + 2;
	long long nid = node_id(g, i, j, k);
	for (int q = 0; q < 9; ++q) muc[q * g.nown + o] = mu[q * g.nnodes + nid];
}

// mu in the halo layout: the ghosted array's values on the unknowns and one node around them.
// The rows read mu at the neighbouring nodes, ghost nodes included (the reference indexes
// dielectricTensor[i +- 1][j][k] directly). Those hold the periodic images after the folds,
// and keep their values through the smoothing passes, which refresh no mu ghosts.
__global__ void k_pack_mu_halo(Grid g, const double* __restrict__ mu, double* __restrict__ muh) {
	const long long H = halo_count(g);
	long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (h >= H) return;
	int hk = (int)(h % (g.n[2] + 2));
	int hj = (int)((h / (g.n[2] + 2)) % (g.n[1] + 2));
	int hi = (int)(h / ((long long)(g.n[2] + 2) * (g.n[1] + 2)));
	long long nid = node_id(g, hi + 1, hj + 1, hk + 1);
	for (int q = 0; q < 9; ++q) muh[q * H + h] = mu[q * g.nnodes + nid];
}

// The implicit Maxwell operator applied matrix-free on the compact periodic layout:
// the 37 entries of createInternalEquation{X,Y,Z} (fields.cpp:1396-1509, :1961-2072,
// :2521-2632), in push_back order, with mu read at the neighbouring nodes (halo layout).
struct Neigh {
	int i[3], j[3], k[3];  // owned index of offset -1, 0, +1 per axis
};

__device__ inline long long oidx(const Grid& g, const Neigh& nb, int di, int dj, int dk) {
	return ((long long)nb.i[di + 1] * g.n[1] + nb.j[dj + 1]) * g.n[2] + nb.k[dk + 1];
}

__global__ void k_apply_operator(Grid g, const double* __restrict__ muh, const double* __restrict__ v, double C,
                                 double* __restrict__ out) {
	long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= g.nown) return;
	int k = (int)(o % g.n[2]);
	int j = (int)((o / g.n[2]) % g.n[1]);
	int i = (int)(o / ((long long)g.n[2] * g.n[1]));
	if (g.wall && (i == 0 || i >= g.n[0])) {
		// wall rows (x: E_x(2) - E_x(3), tangential: identity) and free right rows (identity)
		const long long L = g.nown;
		out[o] = i == 0 ? v[o] + -1.0 * v[o + (long long)g.n[1] * g.n[2]] : v[o];
		out[L + o] = v[L + o];
		out[2 * L + o] = v[2 * L + o];
		return;
	}
	Neigh nb;
	nb.i[0] = i == 0 ? g.nu0 - 1 : i - 1;
	nb.i[1] = i;
	nb.i[2] = i == g.nu0 - 1 ? 0 : i + 1;
	nb.j[0] = j == 0 ? g.n[1] - 1 : j - 1;
	nb.j[1] = j;
	nb.j[2] = j == g.n[1] - 1 ? 0 : j + 1;
	nb.k[0] = k == 0 ? g.n[2] - 1 : k - 1;
	nb.k[1] = k;
	nb.k[2] = k == g.n[2] - 1 ? 0 : k + 1;
	const long long L = g.nown;
	const long long H = halo_count(g);
	auto M = [&](const int* off, int l, int m) { return muh[(l * 3 + m) * H + halo_id(g, i + 1 + off[0], j + 1 + off[1], k + 1 + off[2])]; };
	const int here[3] = {0, 0, 0};
	auto V = [&](long long id, int l) { return v[l * L + id]; };
	double D2[3] = {g.d[0] * g.d[0], g.d[1] * g.d[1], g.d[2] * g.d[2]};
	const int others[3][2] = {{1, 2}, {0, 2}, {0, 1}};
	const int face_order[3][3] = {{0, 1, 2}, {1, 0, 2}, {2, 1, 0}};
	const int trans_order[3][2] = {{1, 2}, {0, 2}, {1, 0}};
	const int plane_order[3][2] = {{1, 2}, {0, 2}, {1, 0}};
	const int sgn[4][2] = {{1, 1}, {-1, -1}, {-1, 1}, {1, -1}};
	for (int c = 0; c < 3; ++c) {
		double r = 0.0;
		// diagonal (same node, same component)
		double e = 1.0 - M(here, c, c);
		for (int ax = 0; ax < 3; ++ax) {
			if (ax == c) {
				if (c == 0)
					e += C * ((2.0 - 2 * M(here, c, c)) / D2[ax]);
				else
					e += C * (2.0 - 2 * M(here, c, c)) / D2[ax];
			} else {
				if (c == 0 || (c == 1 && ax == 2) || (c == 2 && ax == 1))
					e += C * 2.0 / D2[ax];
				else
					e += C * (2.0 / D2[ax]);
			}
		}
		r += e * V(o, c);
		for (int t = 0; t < 2; ++t) {
			int m = others[c][t];
			double val = -M(here, c, m);
			val -= C * 2 * M(here, c, m) / D2[c];
			r += val * V(o, m);
		}
		for (int s = 1; s >= -1; s -= 2) {
			int off[3] = {0, 0, 0};
			off[c] = s;
			long long q = oidx(g, nb, off[0], off[1], off[2]);
			for (int t = 0; t < 3; ++t) {
				int m = face_order[c][t];
				double val = (m == c) ? C * (-1.0 + M(off, c, c)) / D2[c] : C * M(off, c, m) / D2[c];
				r += val * V(q, m);
			}
		}
		for (int t = 0; t < 2; ++t) {
			int ax = trans_order[c][t];
			double val = -C / D2[ax];
			int off[3] = {0, 0, 0};
			off[ax] = 1;
			r += val * V(oidx(g, nb, off[0], off[1], off[2]), c);
			off[ax] = -1;
			r += val * V(oidx(g, nb, off[0], off[1], off[2]), c);
		}
		for (int t = 0; t < 2; ++t) {
			int ax = plane_order[c][t];
			double den;
			if (c == 2 && ax == 1)
				den = 4 * g.d[2] * g.d[1];
			else if (c == 1 && ax == 2)
				den = 4 * g.d[1] * g.d[2];
			else
				den = 4 * g.d[0] * g.d[ax == 0 ? c : ax];
			int axA = c, axB = ax;
			if ((c == 1 && ax == 0) || (c == 2 && ax == 0)) {
				axA = 0;
				axB = c;
			}
			if (c == 2 && ax == 1) {
				axA = 1;
				axB = 2;
			}
			for (int cr = 0; cr < 4; ++cr) {
				int off[3] = {0, 0, 0};
				off[axA] = sgn[cr][0];
				off[axB] = sgn[cr][1];
				long long q = oidx(g, nb, off[0], off[1], off[2]);
				bool pos = sgn[cr][0] * sgn[cr][1] > 0;
				for (int m = 0; m < 3; ++m) {
					double val = pos ? C * M(off, ax, m) / den : -C * M(off, ax, m) / den;
					r += val * V(q, m);
				}
			}
		}
		out[c * L + o] = r;
	}
}

// ------------------------------------------------------------------ Krylov vector kernels
// Deterministic dot product: fixed grid, fixed per-thread order, fixed tree.
__global__ void __launch_bounds__(256) k_dot_partial(const double* __restrict__ a, const double* __restrict__ b, long long len,
                                                     double* __restrict__ partial) {
	__shared__ double red[256];
	double s = 0.0;
	for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < len; t += (long long)gridDim.x * blockDim.x)
		s += a[t] * b[t];
	red[threadIdx.x] = s;
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Final stage: out = sum(partial); if do_sqrt, out = sqrt(out).
__global__ void __launch_bounds__(256) k_dot_final(const double* __restrict__ partial, int n, double* __restrict__ out, int do_sqrt) {
	__shared__ double red[256];
	double s = 0.0;
	for (int t = threadIdx.x; t < n; t += blockDim.x) s += partial[t];
	red[threadIdx.x] = s;
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0) *out = do_sqrt ? sqrt(red[0]) : red[0];
}

// w = w - h * v, h read from device memory
__global__ void k_axpy_dev(double* __restrict__ w, const double* __restrict__ v, const double* __restrict__ h, long long len) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= len) return;
	w[t] -= (*h) * v[t];
}

// w = w / h if h > 0, else 0 (arnoldiIterations, specialmath.cpp:338-362)
__global__ void k_normalize_dev(double* __restrict__ w, const double* __restrict__ h, long long len) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= len) return;
	double hv = *h;
	w[t] = hv > 0 ? w[t] / hv : 0.0;
}

// b = b / norm (norm on device); copy into v0
__global__ void k_div_copy(double* __restrict__ b, double* __restrict__ v0, const double* __restrict__ norm, long long len) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= len) return;
	double x = b[t] / *norm;
	b[t] = x;
	v0[t] = x;
}

// x = sum_m (v_m * y_m) * norm (specialmath.cpp:625-637)
__global__ void k_combine(double* __restrict__ x, const double* const* __restrict__ V, const double* __restrict__ y, int m,
                          const double* __restrict__ norm, long long len) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= len) return;
	double s = 0.0;
	double nm = *norm;
	for (int q = 0; q < m; ++q) s += V[q][t] * y[q] * nm;
	x[t] = s;
}

// E^{n+theta} on every ghosted node from the compact solution, and
// newE = (E^{n+theta} - (1 - theta) E^n) / theta (fields.cpp:256-339).
__global__ void k_unpack_e(Grid g, const double* __restrict__ x, const double* __restrict__ E, double theta,
                           double* __restrict__ Etheta, double* __restrict__ Enew) {
	long long nid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (nid >= g.nnodes) return;
	int k = (int)(nid % (g.N[2] + 1));
	int j = (int)((nid / (g.N[2] + 1)) % (g.N[1] + 1));
	int i = (int)(nid / ((long long)(g.N[2] + 1) * (g.N[1] + 1)));
	if (g.wall) {
		// evaluateElectricField with the wall (fields.cpp:256-339): E^{n+theta} from the solution on
		// nodes 2 .. N-1, the ghosts behind the wall copy node 2 (exchangeEfield), node N holds e_right
		long long o = ((long long)((i < 2 ? 2 : i) - 2) * g.n[1] + wrap_owned(j, g.n[1])) * g.n[2] + wrap_owned(k, g.n[2]);
		long long src = i < 2 ? node_id(g, 2, j, k) : nid;
		for (int c = 0; c < 3; ++c) {
			double et = i == g.N[0] ? g.e_right[c] : x[c * g.nown + o];
			Etheta[c * g.nnodes + nid] = et;
			Enew[c * g.nnodes + nid] = i == g.N[0] ? g.e_right[c] : (et - E[c * g.nnodes + src] * (1 - theta)) / theta;
		}
		return;
	}
	long long o = own_id(g, wrap_owned(i, g.n[0]), wrap_owned(j, g.n[1]), wrap_owned(k, g.n[2]));
	for (int c = 0; c < 3; ++c) {
		double et = x[c * g.nown + o];
		Etheta[c * g.nnodes + nid] = et;
		Enew[c * g.nnodes + nid] = (et - E[c * g.nnodes + nid] * (1 - theta)) / theta;
	}
}

// evaluateMagneticField with evaluateRotEgeneral (fields.cpp:2717-2773, :2862-2953), all cells.
__global__ void k_update_b(Grid g, const double* __restrict__ Et, const double* __restrict__ B, double dt,
                           double* __restrict__ Bnew) {
	long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.ncells) return;
	int k = (int)(c % g.N[2]);
	int j = (int)((c / g.N[2]) % g.N[1]);
	int i = (int)(c / ((long long)g.N[2] * g.N[1]));
	if (g.wall && i >= g.N[0] - 2) {  // the free right end holds b_right (fields.cpp:2739-2743)
		for (int l = 0; l < 3; ++l) Bnew[l * g.ncells + c] = g.b_right[l];
		return;
	}
	if (g.wall && i < 2) i = 2;  // behind the wall: a copy of cell 2 (fields.cpp:2747-2757)
	auto e = [&](int a, int b, int cc, int l) { return Et[l * g.nnodes + node_id(g, a, b, cc)]; };
	double rX[3], lX[3], rY[3], lY[3], rZ[3], lZ[3];
	for (int l = 0; l < 3; ++l) {
		rX[l] = (e(i + 1, j, k, l) + e(i + 1, j + 1, k, l) + e(i + 1, j, k + 1, l) + e(i + 1, j + 1, k + 1, l)) / 4.0;
		lX[l] = (e(i, j, k, l) + e(i, j + 1, k, l) + e(i, j, k + 1, l) + e(i, j + 1, k + 1, l)) / 4.0;
		rY[l] = (e(i, j + 1, k, l) + e(i + 1, j + 1, k, l) + e(i, j + 1, k + 1, l) + e(i + 1, j + 1, k + 1, l)) / 4.0;
		lY[l] = (e(i, j, k, l) + e(i + 1, j, k, l) + e(i, j, k + 1, l) + e(i + 1, j, k + 1, l)) / 4.0;
		rZ[l] = (e(i, j, k + 1, l) + e(i + 1, j, k + 1, l) + e(i, j + 1, k + 1, l) + e(i + 1, j + 1, k + 1, l)) / 4.0;
		lZ[l] = (e(i, j, k, l) + e(i + 1, j, k, l) + e(i, j + 1, k, l) + e(i + 1, j + 1, k, l)) / 4.0;
	}
	double r[3];
	r[0] = (rY[2] - lY[2]) / g.d[1] - (rZ[1] - lZ[1]) / g.d[2];
	r[1] = (rZ[0] - lZ[0]) / g.d[2] - (rX[2] - lX[2]) / g.d[0];
	r[2] = (rX[1] - lX[1]) / g.d[0] - (rY[0] - lY[0]) / g.d[1];
	const long long self = cell_id(g, i, j, k);
	for (int l = 0; l < 3; ++l) Bnew[l * g.ncells + c] = B[l * g.ncells + self] - r[l] * (kCprime * dt);
}

// ------------------------------------------------------------------ mover
__device__ inline void cross(const double a[3], const double b[3], double out[3]) {
	out[0] = a[1] * b[2] - a[2] * b[1];
	out[1] = -a[0] * b[2] + a[2] * b[0];
	out[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ inline double dnorm(const double a[3], const double b[3]) {
	double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
	return sqrt(x * x + y * y + z * z);
}

// moveParticle (PIC++/mover.cpp:194-491) with the one-rank periodic wrap of exchangeParticles
// (mover.cpp:1007-1021), followed by the new sort key. alpha is recomputed from E^n, B^n at x^n
// as evaluateParticlesRotationTensor computed it (to rounding). The maps of the predictor and
// Picard positions, the velocities and alpha take reciprocals in place of divisions
// (pic_recip.hpp); the new key keeps the reference's divisions. One thread per particle.
// 8-point gathers of the mover (correlationGeneralEfield / correlationBfield,
// PIC++/correlation.cpp:23-58, :103-164) with fused multiply-adds
__device__ inline void gather_node_f(const Grid& g, const double* __restrict__ f, const AxisMap* m, double out[3]) {
	out[0] = out[1] = out[2] = 0.0;
#pragma unroll
	for (int t = 0; t < 8; ++t) {
		const int i = (t >> 2) & 1, j = (t >> 1) & 1, k = t & 1;
		const long long id = node_id(g, m[0].nbase + i, m[1].nbase + j, m[2].nbase + k);
		const double c = m[0].nw[i] * m[1].nw[j] * m[2].nw[k];
		out[0] = fma(f[id], c, out[0]);
		out[1] = fma(f[g.nnodes + id], c, out[1]);
		out[2] = fma(f[2 * g.nnodes + id], c, out[2]);
	}
}

__device__ inline void gather_cell_f(const Grid& g, const double* __restrict__ f, const AxisMap* m, double out[3]) {
	out[0] = out[1] = out[2] = 0.0;
#pragma unroll
	for (int t = 0; t < 8; ++t) {
		const int i = (t >> 2) & 1, j = (t >> 1) & 1, k = t & 1;
		const long long id = cell_id(g, m[0].cbase + i, m[1].cbase + j, m[2].cbase + k);
		const double c = m[0].cw[i] * m[1].cw[j] * m[2].cw[k];
		out[0] = fma(f[id], c, out[0]);
		out[1] = fma(f[g.ncells + id], c, out[1]);
		out[2] = fma(f[2 * g.ncells + id], c, out[2]);
	}
}

// Returns whether the particle changed segment (its sort key) or was removed. The arithmetic is
// contracted to fused multiply-adds (within an ulp of the reference's separate roundings).
template <bool WALL>
__device__ inline bool push_particle(const Grid& g, const SpeciesTable& sp, const Recip& R, const ParticleArrays& P,
                                     const int* __restrict__ species_start, const double* __restrict__ E,
                                     const double* __restrict__ Etheta, const double* __restrict__ B, double dt, double eta,
                                     uint32_t* __restrict__ keys, int* __restrict__ status, int* __restrict__ dead,
                                     uint32_t* __restrict__ esc, long long i, bool& bounced) {
#pragma clang fp contract(fast)
	const double xw = g.x0[0] + 2 * g.d[0];  // the wall, xgrid[2]
	const double rest = (1.0 - eta) * dt;
	bool reflected = false;
	int s = 0;
	while (s + 1 < sp.count && i >= species_start[s + 1]) ++s;
	double m = sp.mass[s], q = sp.charge[s];
	double x[3], p[3];
	load_particle(P, i, x, p);
	const uint32_t key0 = position_key(g, s, x);
	AxisMap mp[3];
	if (!maps3_fast(g, R, x, mp)) atomicOr(status, 8);
	double En[3], Bf[3], v[3], gm, a[9];
	gather_node_f(g, E, mp, En);
	gather_cell_f(g, B, mp, Bf);
	velocity_gamma_r(p, m, R.im2[s], v, gm);
	double beta = 0.5 * q * dt / m;
	alpha_tensor_r(beta, v, gm, En, Bf, a);
	double Ef[3];
	gather_node_f(g, Etheta, mp, Ef);
	double oldP[3] = {p[0], p[1], p[2]};
	double newP[3] = {p[0], p[1], p[2]};
	double oldX[3] = {x[0], x[1], x[2]};
	double etaDt = eta * dt;
	double vxb[3];
	cross(v, Bf, vxb);
	for (int c = 0; c < 3; ++c) newP[c] += (Ef[c] + vxb[c]) * q * dt;
	double newV[3];
	evaluate_velocity_r(newP, m, R.im2[s], newV);
	double mid[3];
	for (int c = 0; c < 3; ++c) mid[c] = v[c] * (1 - eta) + newV[c] * eta;
	for (int c = 0; c < 3; ++c) x[c] += mid[c] * etaDt;
	// A predictor or Picard position behind the wall ends the push there: the particle is
	// reflected with |v_x| of the predictor velocity and the rest of the step (mover.cpp:248-258,
	// :315-331; Particle::setMomentumByV, particle.cpp:337-364).
	auto reflect_at_wall = [&]() {
		x[0] = 2 * xw - x[0] + fabs(mid[0] * rest);
		x[1] = x[1] + mid[1] * rest;
		x[2] = x[2] + mid[2] * rest;
		double vr[3] = {fabs(newV[0]), newV[1], newV[2]};
		double v2 = vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2];
		if (v2 < 0.0001) {
			for (int c = 0; c < 3; ++c) p[c] = vr[c] * m;
		} else {
			double gr = 1 / sqrt(1 - v2);
			for (int c = 0; c < 3; ++c) p[c] = vr[c] * (m * gr);
		}
		reflected = true;
	};
	if (WALL && x[0] < xw) reflect_at_wall();
	double vhat[3];
	for (int c = 0; c < 3; ++c) vhat[c] = (a[c * 3 + 0] * gm) * v[0] + (a[c * 3 + 1] * gm) * v[1] + (a[c * 3 + 2] * gm) * v[2];
	double prevP[3] = {p[0], p[1], p[2]};
	double pnorm = sqrt(oldP[0] * oldP[0] + oldP[1] * oldP[1] + oldP[2] * oldP[2]);
	double err = dnorm(prevP, newP);
	double prevX[3] = {x[0], x[1], x[2]};
	double newX[3] = {x[0], x[1], x[2]};
	double cerr = g.d[0];
	int it = 0;
	while (!reflected && err > 1e-14 * pnorm && cerr > 1e-8 * g.d[0] && it < 3) {
		++it;
		for (int c = 0; c < 3; ++c) {
			prevP[c] = newP[c];
			prevX[c] = newX[c];
		}
		double rE[3];
		for (int c = 0; c < 3; ++c) rE[c] = a[c * 3 + 0] * Ef[0] + a[c * 3 + 1] * Ef[1] + a[c * 3 + 2] * Ef[2];
		for (int c = 0; c < 3; ++c) mid[c] = vhat[c] + rE[c] * beta;
		for (int c = 0; c < 3; ++c) x[c] = oldX[c] + (mid[c] * etaDt);
		for (int c = 0; c < 3; ++c) newX[c] = x[c];
		if (WALL && x[0] < xw) {
			reflect_at_wall();
			break;
		}
		if (!maps3_fast(g, R, x, mp)) atomicOr(status, 8);
		gather_node_f(g, Etheta, mp, Ef);
		gather_cell_f(g, B, mp, Bf);
		double mxb[3];
		cross(mid, Bf, mxb);
		for (int c = 0; c < 3; ++c) newP[c] = oldP[c] + (Ef[c] + mxb[c]) * (q * dt);
		err = dnorm(prevP, newP);
		cerr = dnorm(prevX, newX);
	}
	if (!reflected) {
		for (int c = 0; c < 3; ++c) {
			p[c] = newP[c];
			x[c] = oldX[c] + mid[c] * dt;
		}
		if (WALL && x[0] < xw) {  // final position behind the wall (mover.cpp:435-441)
			x[0] = 2 * xw - x[0];
			p[0] = -p[0];
			bounced = true;
		}
	}
	if (reflected) bounced = true;
	if (WALL) {
		// leaving through the free right end removes the particle (mover.cpp:447-458, exchangeParticles
		// keeps no right escapes without periodicity, :1007-1021)
		const double xr = g.x0[0] + (g.N[0] - 2) * g.d[0];
		if (x[0] > xr) {
			for (int c = 0; c < 3; ++c) {
				P.x[c][i] = x[c];
				P.p[c][i] = p[c];
			}
			keys[i] = 0xFFFFFFFFu;
			esc[atomicAdd(dead, 1)] = (uint32_t)i;  // listed for replacement (y, z left unwrapped)
			return true;
		}
	}
	// one-rank periodic wrap (mover.cpp:435-490 classification, :1007-1021 shift)
	for (int c = WALL ? 1 : 0; c < 3; ++c) {
		double lo = g.x0[c] + 2 * g.d[c];
		double hi = g.x0[c] + (g.N[c] - 1 - 1) * g.d[c];
		if (x[c] < lo)
			x[c] += g.L[c];
		else if (x[c] > hi)
			x[c] -= g.L[c];
		if (!(x[c] >= lo && x[c] <= hi)) atomicOr(status, 1);  // moved more than a box length
	}
	for (int c = 0; c < 3; ++c) {
		P.x[c][i] = x[c];
		P.p[c][i] = p[c];
	}
	if (!maps3(g, x, mp)) atomicOr(status, 8);
	int bits[3];
	for (int c = 0; c < 3; ++c) {
		bits[c] = mp[c].cbase - mp[c].nbase + 1;
		if (bits[c] < 0 || bits[c] > 1) atomicOr(status, 2);
	}
	const uint32_t key = particle_key(g, s, mp);
	keys[i] = key;
	return key != key0;
}

template <bool WALL>
__global__ void __launch_bounds__(256) k_push(Grid g, SpeciesTable sp, Recip R, ParticleArrays P, const int* __restrict__ species_start,
                                              const double* __restrict__ E, const double* __restrict__ Etheta,
                                              const double* __restrict__ B, double dt, double eta, uint32_t* __restrict__ keys,
                                              uint64_t* __restrict__ mask, int* __restrict__ status, int* __restrict__ dead,
                                              uint32_t* __restrict__ esc, uint64_t* __restrict__ bounce_mask) {
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	bool mv = false, bounced = false;
	if (i < P.count) mv = push_particle<WALL>(g, sp, R, P, species_start, E, Etheta, B, dt, eta, keys, status, dead, esc, i, bounced);
	// one mover bit per particle: bit i & 63 of word i >> 6, written by the wave's first lane
	// (waves are 64 consecutive particles; lanes past the end contribute zeros)
	const unsigned long long bits = __ballot(mv);
	if ((threadIdx.x & 63) == 0 && i < P.count) mask[i >> 6] = bits;
	if (WALL) {  // and one bit per particle reflected at the wall in this push
		const unsigned long long rb = __ballot(bounced);
		if ((threadIdx.x & 63) == 0 && i < P.count) bounce_mask[i >> 6] = rb;
	}
}

// Sort key of every particle at its current position (updateParticleCorrelationMaps).
__global__ void k_keys(Grid g, SpeciesTable sp, ParticleArrays P, const int* __restrict__ species_start, uint32_t* __restrict__ keys,
                       int* __restrict__ status) {
	long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= P.count) return;
	int s = 0;
	while (s + 1 < sp.count && i >= species_start[s + 1]) ++s;
	double x[3];
	for (int c = 0; c < 3; ++c) {
		x[c] = P.x[c][i];
		if (c == 2 && g.cz) continue;  // z unwrapped along a collapsed axis
		double lo = g.x0[c] + 2 * g.d[c];
		double hi = g.x0[c] + (g.N[c] - 2) * g.d[c];
		if (!(x[c] >= lo && x[c] <= hi)) atomicOr(status, 4);
	}
	if (g.cz) {  // one cell along z: the stencils of x and y, the key of plane 0
		for (int c = 0; c < 2; ++c) {
			const AxisMap m = axis_map(x[c], g.x0[c], g.d[c]);
			if (!(m.nbase >= 0 && m.nbase <= g.N[c] - 1 && m.cbase >= 0 && m.cbase <= g.N[c] - 2)) atomicOr(status, 8);
		}
		keys[i] = position_key(g, s, x);
		return;
	}
	AxisMap mp[3];
	if (!maps3(g, x, mp)) atomicOr(status, 8);
	for (int c = 0; c < 3; ++c)
		if (mp[c].cbase - mp[c].nbase + 1 < 0 || mp[c].cbase - mp[c].nbase + 1 > 1) atomicOr(status, 2);
	keys[i] = particle_key(g, s, mp);
}

// Apply the sort permutation to the particle arrays.
__global__ void k_permute(ParticleArrays src, ParticleArrays dst, const uint32_t* __restrict__ perm) {
	long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= src.count) return;
	long long s = perm[i];
	for (int c = 0; c < 3; ++c) {
		dst.x[c][i] = src.x[c][s];
		dst.p[c][i] = src.p[c][s];
	}
	dst.id[i] = src.id[s];
}

// start[key] = first sorted index with that key; start[nkeys] = count.
__global__ void k_segments(const uint32_t* __restrict__ keys, long long count, long long nkeys, int* __restrict__ start) {
	long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i > count) return;
	long long prev = (i == 0) ? -1 : (long long)keys[i - 1];
	long long cur = (i == count) ? nkeys : (long long)keys[i];
	for (long long k = prev + 1; k <= cur; ++k) start[k] = (int)i;
}


// Cell values of the density quantities: nq planes of 8 corner sums -> nq cell planes.
__global__ void k_density_assemble(Grid g, const double* __restrict__ cs, int nq, double* __restrict__ out) {
	long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.ncells) return;
	int k = (int)(c % g.N[2]);
	int j = (int)((c / g.N[2]) % g.N[1]);
	int i = (int)(c / ((long long)g.N[2] * g.N[1]));
	for (int q = 0; q < nq; ++q) {
		double s = 0.0;
		for (int a = 0; a < 2; ++a)
			for (int b = 0; b < 2; ++b)
				for (int cc = 0; cc < 2; ++cc) {
					int bi = i - a, bj = j - b, bk = k - cc;
					if (bi < 0 || bj < 0 || bk < 0) continue;
					s += cs[((long long)q * 8 + a * 4 + b * 2 + cc) * g.ncells + cell_id(g, bi, bj, bk)];
				}
		out[(long long)q * g.ncells + c] = s;
	}
}

// bulk velocity = bulk momentum / energy, or 0 (momentum_collecting.cpp:991-1008)
__global__ void k_bulk_divide(Grid g, double* __restrict__ dens, int nspecies) {
	long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.ncells) return;
	for (int s = 0; s < nspecies; ++s) {
		double en = dens[((long long)s * 5 + 1) * g.ncells + c];
		for (int a = 0; a < 3; ++a) {
			double* bv = &dens[((long long)s * 5 + 2 + a) * g.ncells + c];
			*bv = en > 0 ? *bv / en : 0.0;
		}
	}
}

// ------------------------------------------------------------------ energy (updateEnergy, simulation.cpp:1506-1640)
// partial sums per block: [0] E energy, [1] B energy, [2..4] field momentum
__global__ void __launch_bounds__(256) k_field_energy(Grid g, const double* __restrict__ E, const double* __restrict__ B,
                                                      double* __restrict__ partial) {
	__shared__ double red[5][256];
	double acc[5] = {0, 0, 0, 0, 0};
	// x range of updateEnergy (simulation.cpp:1522-1530): owned nodes and cells 2 .. n0+1, or
	// 0 .. N-2 with the wall boundary (its ghost layers included)
	const int nx = g.wall ? g.N[0] - 1 : g.n[0], x_lo = g.wall ? 0 : 2;
	const long long cols = (long long)nx * g.n[1] * g.n[2];
	for (long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x; o < cols; o += (long long)gridDim.x * blockDim.x) {
		int k = (int)(o % g.n[2]) + 2;
		int j = (int)((o / g.n[2]) % g.n[1]) + 2;
		int i = (int)(o / ((long long)g.n[2] * g.n[1])) + x_lo;
		long long nid = node_id(g, i, j, k), cid = cell_id(g, i, j, k);
		double e[3], b[3], ea[3];
		for (int c = 0; c < 3; ++c) {
			e[c] = E[c * g.nnodes + nid];
			b[c] = B[c * g.ncells + cid];
			ea[c] = (E[c * g.nnodes + node_id(g, i, j, k)] + E[c * g.nnodes + node_id(g, i, j, k + 1)] +
			         E[c * g.nnodes + node_id(g, i, j + 1, k)] + E[c * g.nnodes + node_id(g, i, j + 1, k + 1)] +
			         E[c * g.nnodes + node_id(g, i + 1, j, k)] + E[c * g.nnodes + node_id(g, i + 1, j, k + 1)] +
			         E[c * g.nnodes + node_id(g, i + 1, j + 1, k)] + E[c * g.nnodes + node_id(g, i + 1, j + 1, k + 1)]) / 8;
		}
		acc[0] += (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) * g.volume / (8 * kPi);
		acc[1] += (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]) * g.volume / (8 * kPi);
		double cr[3];
		cross(ea, b, cr);
		for (int c = 0; c < 3; ++c) acc[2 + c] += (cr[c] / (4 * kPi)) * g.volume;
	}
	for (int q = 0; q < 5; ++q) red[q][threadIdx.x] = acc[q];
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < w)
			for (int q = 0; q < 5; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		for (int q = 0; q < 5; ++q) partial[q * gridDim.x + blockIdx.x] = red[q][0];
}

// partial sums per block: [0] kinetic energy sum w (gamma - 1) m, [1..3] particle momentum
__global__ void __launch_bounds__(256) k_particle_energy(SpeciesTable sp, ParticleArrays P, const int* __restrict__ species_start,
                                                         double* __restrict__ partial) {
	__shared__ double red[4][256];
	double acc[4] = {0, 0, 0, 0};
	for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < P.count; i += (long long)gridDim.x * blockDim.x) {
		int s = 0;
		while (s + 1 < sp.count && i >= species_start[s + 1]) ++s;
		double m = sp.mass[s], w = sp.weight[s];
		double p[3] = {P.p[0][i], P.p[1][i], P.p[2][i]};
		double p2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
		double gm = sqrt(p2 / (m * m) + 1.0);
		acc[0] += ((gm - 1) * m) * w;
		for (int c = 0; c < 3; ++c) acc[1 + c] += p[c] * w;
	}
	for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = acc[q];
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < w)
			for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		for (int q = 0; q < 4; ++q) partial[q * gridDim.x + blockIdx.x] = red[q][0];
}

// generic deterministic final reduce of `rows` rows of `n` partials each
__global__ void __launch_bounds__(256) k_rows_final(const double* __restrict__ partial, int n, int rows, double* __restrict__ out) {
	__shared__ double red[256];
	int r = blockIdx.x;
	if (r >= rows) return;
	double s = 0.0;
	for (int t = threadIdx.x; t < n; t += blockDim.x) s += partial[(long long)r * n + t];
	red[threadIdx.x] = s;
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0) out[r] = red[0];
}

}  // namespace picgpu

// ------------------------------------------------------------------ integer helpers
namespace picgpu {

// Run lengths of the sorted keys: the first index of every run writes its length.
__global__ void k_count_keys(const uint32_t* __restrict__ keys, long long count, int* __restrict__ counts) {
	long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= count) return;
	uint32_t k = keys[i];
	if (i > 0 && keys[i - 1] == k) return;
	long long j = i + 1;
	while (j < count && keys[j] == k) ++j;
	counts[k] = (int)(j - i);
}

__global__ void k_iota(uint32_t* __restrict__ v, long long count) {
	long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= count) return;
	v[i] = (uint32_t)i;
}

__global__ void k_add(double* __restrict__ x, const double* __restrict__ d, long long len) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < len) x[t] += d[t];
}

__global__ void k_sub(double* __restrict__ r, const double* __restrict__ b, const double* __restrict__ y, long long len) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < len) r[t] = b[t] - y[t];
}

// ------------------------------------------------------------------ launch wrappers
namespace {
inline unsigned blocks(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
}  // namespace

void launch_node_assemble(const Grid& g, const double* scratch, double* J, double* mu, hipStream_t s) {
	hipLaunchKernelGGL(k_node_assemble, dim3(blocks(g.nnodes, 256)), dim3(256), 0, s, g, scratch, J, mu);
}
void launch_cell_assemble(const Grid& g, const double* rho_scratch, double* rho, hipStream_t s) {
	hipLaunchKernelGGL(k_cell_assemble, dim3(blocks(g.ncells, 256)), dim3(256), 0, s, g, rho_scratch, rho);
}
void launch_pressure_cell(const Grid& g, const double* pi_scratch, double* Pi, hipStream_t s) {
	hipLaunchKernelGGL(k_pressure_cell<double>, dim3(blocks(g.ncells, 256)), dim3(256), 0, s, g, pi_scratch, Pi);
}
void launch_pressure_cell32(const Grid& g, const float* pi_scratch, double* Pi, hipStream_t s) {
	hipLaunchKernelGGL(k_pressure_cell<float>, dim3(blocks(g.ncells, 256)), dim3(256), 0, s, g, pi_scratch, Pi);
}
void launch_pressure_reduce(const Grid& g, const double* pi_scratch, double* Pi, hipStream_t s) {
	hipLaunchKernelGGL(k_pressure_reduce, dim3(g.N[0] * 4), dim3(256), 0, s, g, pi_scratch, Pi);
}
void launch_fold_nodes(const Grid& g, double* f, int ncomp, hipStream_t s) {
	for (int axis = g.wall ? 1 : 0; axis < 3; ++axis) {  // no x fold behind a wall (momentum_exchange.cpp:173)
		int o1 = (axis + 1) % 3, o2 = (axis + 2) % 3;
		long long total = (long long)(g.N[o1] + 1) * (g.N[o2] + 1) * 5 * ncomp;
		hipLaunchKernelGGL(k_fold, dim3(blocks(total, 256)), dim3(256), 0, s, g, f, ncomp, axis, 1, 5, g.N[axis], 4);
	}
}
void launch_fold_cells(const Grid& g, double* f, int ncomp, int pairs, int lo0, hipStream_t s) {
	for (int axis = g.wall ? 1 : 0; axis < 3; ++axis) {  // no x fold behind a wall (momentum_exchange.cpp:173)
		int o1 = (axis + 1) % 3, o2 = (axis + 2) % 3;
		long long total = (long long)g.N[o1] * g.N[o2] * pairs * ncomp;
		hipLaunchKernelGGL(k_fold, dim3(blocks(total, 256)), dim3(256), 0, s, g, f, ncomp, axis, 0, pairs, g.N[axis] - 1, lo0);
	}
}
void launch_ghost_fill(const Grid& g, double* f, int ncomp, bool nodes, hipStream_t s, bool wall_zero) {
	hipLaunchKernelGGL(k_ghost_fill, dim3(blocks(nodes ? g.nnodes : g.ncells, 256)), dim3(256), 0, s, g, f, ncomp, nodes ? 1 : 0,
	                   wall_zero ? 1 : 0);
}
Smooth27 smoothing_matrix(double sp, bool collapsed_z) {
	Smooth27 m;
	const double f[3] = {sp / 2, 1 - sp, sp / 2};  // createSmoothingMatrix, smoothing.cpp:105-137
	const double fz[3] = {0.0, 1.0, 0.0};          // its factors along a one-cell z (:121-130)
	for (int a = 0; a < 3; ++a)
		for (int b = 0; b < 3; ++b)
			for (int c = 0; c < 3; ++c) m.m[a * 9 + b * 3 + c] = f[a] * f[b] * (collapsed_z ? fz[c] : f[c]);
	return m;
}
void launch_smooth_cells(const Grid& g, double* f, double* tmp, const Smooth27& m, hipStream_t s) {
	long long n = (long long)(g.N[0] - 3) * g.n[1] * g.n[2];
	hipLaunchKernelGGL(k_smooth_cells, dim3(blocks(n, 256)), dim3(256), 0, s, g, f, tmp, m, 0);
	hipLaunchKernelGGL(k_smooth_cells, dim3(blocks(n, 256)), dim3(256), 0, s, g, f, tmp, m, 1);
}
void launch_smooth_nodes(const Grid& g, double* f, int ncomp, double* tmp, const Smooth27& m, hipStream_t s) {
	long long n = (long long)(g.N[0] - 1 - (g.wall ? 3 : 2)) * (g.n[1] + 1) * (g.n[2] + 1) * ncomp;
	hipLaunchKernelGGL(k_smooth_nodes, dim3(blocks(n, 256)), dim3(256), 0, s, g, f, tmp, ncomp, m, 0);
	hipLaunchKernelGGL(k_smooth_nodes, dim3(blocks(n, 256)), dim3(256), 0, s, g, f, tmp, ncomp, m, 1);
}
void launch_div_pressure_update_j(const Grid& g, const double* Pi, double* J, double* divPi, double eta, double dt, hipStream_t s) {
	hipLaunchKernelGGL(k_div_pressure_update_j, dim3(blocks(g.nnodes, 256)), dim3(256), 0, s, g, Pi, J, divPi, eta, dt);
}
void launch_div_flux_update_rho(const Grid& g, const double* J, double* rho, double dt, double theta, hipStream_t s) {
	hipLaunchKernelGGL(k_div_flux_update_rho, dim3(blocks(g.ncells, 256)), dim3(256), 0, s, g, J, rho, dt, theta);
}
void launch_rhs(const Grid& g, const double* E, const double* B, const double* J, const double* rho, double theta, double dt,
                double* rhs, hipStream_t s) {
	hipLaunchKernelGGL(k_rhs, dim3(blocks(g.nown, 256)), dim3(256), 0, s, g, E, B, J, rho, theta, dt, rhs);
}
void launch_pack_mu(const Grid& g, const double* mu, double* muc, hipStream_t s) {
	hipLaunchKernelGGL(k_pack_mu, dim3(blocks(g.nown, 256)), dim3(256), 0, s, g, mu, muc);
}
void launch_apply_operator(const Grid& g, const double* muh, const double* v, double C, double* out, hipStream_t s) {
	hipLaunchKernelGGL(k_apply_operator, dim3(blocks(g.nown, 256)), dim3(256), 0, s, g, muh, v, C, out);
}
void launch_pack_mu_halo(const Grid& g, const double* mu, double* muh, hipStream_t s) {
	hipLaunchKernelGGL(k_pack_mu_halo, dim3(blocks(halo_count(g), 256)), dim3(256), 0, s, g, mu, muh);
}
void launch_dot(const double* a, const double* b, long long len, double* partial, double* out, bool do_sqrt, hipStream_t s) {
	hipLaunchKernelGGL(k_dot_partial, dim3(kDotBlocks), dim3(256), 0, s, a, b, len, partial);
	hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, s, partial, kDotBlocks, out, do_sqrt ? 1 : 0);
}
void launch_axpy_dev(double* w, const double* v, const double* h, long long len, hipStream_t s) {
	hipLaunchKernelGGL(k_axpy_dev, dim3(blocks(len, 256)), dim3(256), 0, s, w, v, h, len);
}
void launch_normalize_dev(double* w, const double* h, long long len, hipStream_t s) {
	hipLaunchKernelGGL(k_normalize_dev, dim3(blocks(len, 256)), dim3(256), 0, s, w, h, len);
}
void launch_div_copy(double* b, double* v0, const double* norm, long long len, hipStream_t s) {
	hipLaunchKernelGGL(k_div_copy, dim3(blocks(len, 256)), dim3(256), 0, s, b, v0, norm, len);
}
void launch_combine(double* x, const double* const* V, const double* y, int m, const double* norm, long long len, hipStream_t s) {
	hipLaunchKernelGGL(k_combine, dim3(blocks(len, 256)), dim3(256), 0, s, x, V, y, m, norm, len);
}
void launch_add(double* x, const double* d, long long len, hipStream_t s) {
	hipLaunchKernelGGL(k_add, dim3(blocks(len, 256)), dim3(256), 0, s, x, d, len);
}
void launch_sub(double* r, const double* b, const double* y, long long len, hipStream_t s) {
	hipLaunchKernelGGL(k_sub, dim3(blocks(len, 256)), dim3(256), 0, s, r, b, y, len);
}
void launch_unpack_e(const Grid& g, const double* x, const double* E, double theta, double* Etheta, double* Enew, hipStream_t s) {
	hipLaunchKernelGGL(k_unpack_e, dim3(blocks(g.nnodes, 256)), dim3(256), 0, s, g, x, E, theta, Etheta, Enew);
}
void launch_update_b(const Grid& g, const double* Etheta, const double* B, double dt, double* Bnew, hipStream_t s) {
	hipLaunchKernelGGL(k_update_b, dim3(blocks(g.ncells, 256)), dim3(256), 0, s, g, Etheta, B, dt, Bnew);
}
void launch_push(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P, const int* species_start, const double* E,
                 const double* Etheta, const double* B, double dt, double eta, uint32_t* keys, uint64_t* mask, int* status,
                 int* dead, uint32_t* esc, uint64_t* bounce_mask, hipStream_t s) {
	if (P.count == 0) return;
	hipLaunchKernelGGL(g.wall ? k_push<true> : k_push<false>, dim3(blocks(P.count, 256)), dim3(256), 0, s, g, sp, make_recip(g, sp), P,
	                   species_start, E, Etheta, B, dt, eta, keys, mask, status, dead, esc, bounce_mask);
}
void launch_wall_mirror_cells(const Grid& g, double* f, int ncomp, const int* flip, hipStream_t s) {
	long long n = (long long)g.N[1] * g.N[2] * ncomp;
	hipLaunchKernelGGL(k_wall_mirror_cells, dim3(blocks(n, 256)), dim3(256), 0, s, g, f, ncomp, flip);
}
void launch_wall_right_fold(const Grid& g, double* f, int ncomp, const int* zero_mask, hipStream_t s) {
	long long n = (long long)g.N[1] * g.N[2] * ncomp;
	hipLaunchKernelGGL(k_wall_right_fold, dim3(blocks(n, 256)), dim3(256), 0, s, g, f, ncomp, zero_mask);
}
void launch_zero_node_planes(const Grid& g, double* f, int c0, int c1, int i0, int i1, hipStream_t s) {
	long long n = (long long)(g.N[1] + 1) * (g.N[2] + 1) * (i1 - i0 + 1) * (c1 - c0);
	hipLaunchKernelGGL(k_zero_node_planes, dim3(blocks(n, 256)), dim3(256), 0, s, g, f, c0, c1, i0, i1);
}
void launch_keys(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P, const int* species_start, uint32_t* keys,
                 int* status, hipStream_t s) {
	if (P.count == 0) return;
	hipLaunchKernelGGL(k_keys, dim3(blocks(P.count, 256)), dim3(256), 0, s, g, sp, P, species_start, keys, status);
}
void launch_permute(const ParticleArrays& src, const ParticleArrays& dst, const uint32_t* perm, hipStream_t s) {
	if (src.count == 0) return;
	hipLaunchKernelGGL(k_permute, dim3(blocks(src.count, 256)), dim3(256), 0, s, src, dst, perm);
}
void launch_count_keys(const uint32_t* keys, long long count, int* counts, hipStream_t s) {
	if (count == 0) return;
	hipLaunchKernelGGL(k_count_keys, dim3(blocks(count, 256)), dim3(256), 0, s, keys, count, counts);
}
void launch_iota(uint32_t* v, long long count, hipStream_t s) {
	if (count == 0) return;
	hipLaunchKernelGGL(k_iota, dim3(blocks(count, 256)), dim3(256), 0, s, v, count);
}
void launch_density_assemble(const Grid& g, const double* scratch, int nq, double* out, hipStream_t s) {
	hipLaunchKernelGGL(k_density_assemble, dim3(blocks(g.ncells, 256)), dim3(256), 0, s, g, scratch, nq, out);
}
void launch_bulk_divide(const Grid& g, double* dens, int nspecies, hipStream_t s) {
	hipLaunchKernelGGL(k_bulk_divide, dim3(blocks(g.ncells, 256)), dim3(256), 0, s, g, dens, nspecies);
}
void launch_field_energy(const Grid& g, const double* E, const double* B, double* partial, hipStream_t s) {
	hipLaunchKernelGGL(k_field_energy, dim3(kEnergyBlocks), dim3(256), 0, s, g, E, B, partial);
}
void launch_particle_energy(const SpeciesTable& sp, const ParticleArrays& P, const int* species_start, double* partial,
                            hipStream_t s) {
	hipLaunchKernelGGL(k_particle_energy, dim3(kEnergyBlocks), dim3(256), 0, s, sp, P, species_start, partial);
}
void launch_rows_final(const double* partial, int n, int rows, double* out, hipStream_t s) {
	hipLaunchKernelGGL(k_rows_final, dim3(rows), dim3(256), 0, s, partial, n, rows, out);
}

}  // namespace picgpu
