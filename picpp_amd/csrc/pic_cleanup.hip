// Divergence cleanup of E and Fourier short-wave filtering of the new fields (FP64).
//
// Divergence cleanup (Simulation::cleanupDivergence, PIC++/divergencecleanup.cpp:20-392, the GMRES
// branch; evaluateDivergenceCleaningField :553-618; updateFieldByCleaning :531-551): a cell-centred
// potential phi solves L phi = div E - 4 pi rho on the owned cells of the periodic box, with L the
// 27-point operator of createDivergenceCleanupInternalEquation (:620-813), which is exactly the
// divergence (evaluateDivEgeneral, PIC++/fields.cpp:2967-2978) of the node gradient below; then
// E -= grad phi at the nodes, so that div E = 4 pi rho to the solver's tolerance.
//
// Fourier filtering (Simulation::filterFields / filterFieldGeneral, PIC++/fields_filtering.cpp:
// 21-295, periodic branch): each component of newE and newB over the owned points goes through
// the separable discrete Fourier transform of fourierTranslation (PIC++/fourier.cpp:13-192: x, then
// y, then z; the inverse divides by the length of each axis), its modes with |k| above
// 2 pi / (cut dx) are multiplied by -1, and the inverse transform's real part replaces the owned
// values. The mode numbers are the unfolded indices 0 .. n-1 of fourierTranslation's output
// (k = m 2 pi / L), and the y component takes its k_y from the z mode index, as the reference does.
#include "pic_cleanup.hpp"
#include "pic_device.hpp"

#include <utility>

namespace picgpu {

namespace {
constexpr double kPiC = 3.141592653589793115997963468544185161590576171875;  // 4*atan2(1,1)
inline unsigned nb(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// owned cell o = (i, j, k), 0-based, of the n0 x n1 x n2 owned block
__device__ inline void owned_ijk(const Grid& g, long long o, int& i, int& j, int& k) {
	k = (int)(o % g.n[2]);
	j = (int)((o / g.n[2]) % g.n[1]);
	i = (int)(o / ((long long)g.n[2] * g.n[1]));
}
__device__ inline int wrapn(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
}  // namespace

// right part of the cleanup equation per owned cell: -(4 pi rho - div E) (cleanUpRightPart,
// PIC++/divergencecleanup.cpp:832-837, divergence of evaluateDivEgeneral)
__global__ void k_cleanup_rhs(Grid g, const double* __restrict__ E, const double* __restrict__ rho, double* __restrict__ rhs) {
	const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= (long long)g.n[0] * g.n[1] * g.n[2]) return;
	int a, b, c;
	owned_ijk(g, o, a, b, c);
	const int i = a + 2, j = b + 2, k = c + 2;
	const double* Ex = E;
	const double* Ey = E + g.nnodes;
	const double* Ez = E + 2 * g.nnodes;
	auto n = [&](int ii, int jj, int kk) { return node_id(g, ii, jj, kk); };
	const double rx = (Ex[n(i + 1, j, k)] + Ex[n(i + 1, j + 1, k)] + Ex[n(i + 1, j, k + 1)] + Ex[n(i + 1, j + 1, k + 1)]) / 4.0;
	const double lx = (Ex[n(i, j, k)] + Ex[n(i, j + 1, k)] + Ex[n(i, j, k + 1)] + Ex[n(i, j + 1, k + 1)]) / 4.0;
	const double ry = (Ey[n(i, j + 1, k)] + Ey[n(i + 1, j + 1, k)] + Ey[n(i, j + 1, k + 1)] + Ey[n(i + 1, j + 1, k + 1)]) / 4.0;
	const double ly = (Ey[n(i, j, k)] + Ey[n(i + 1, j, k)] + Ey[n(i, j, k + 1)] + Ey[n(i + 1, j, k + 1)]) / 4.0;
	const double rz = (Ez[n(i, j, k + 1)] + Ez[n(i + 1, j, k + 1)] + Ez[n(i, j + 1, k + 1)] + Ez[n(i + 1, j + 1, k + 1)]) / 4.0;
	const double lz = (Ez[n(i, j, k)] + Ez[n(i + 1, j, k)] + Ez[n(i, j + 1, k)] + Ez[n(i + 1, j + 1, k)]) / 4.0;
	const double div = ((rx - lx) / g.d[0]) + ((ry - ly) / g.d[1]) + ((rz - lz) / g.d[2]);
	rhs[o] = -(4 * kPiC * rho[cell_id(g, i, j, k)] - div);
}

// The 27-point cleanup operator on the owned cells with periodic images (createDivergenceCleanup
// InternalEquation, 3-D branch with _27points; exchangeLargeVector's one-rank periodic copy)
__global__ void __launch_bounds__(256) k_cleanup_apply(Grid g, CleanupStencil S, const double* __restrict__ v, double* __restrict__ out) {
	const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= (long long)g.n[0] * g.n[1] * g.n[2]) return;
	int i, j, k;
	owned_ijk(g, o, i, j, k);
	double s = 0.0;
#pragma unroll
	for (int a = -1; a <= 1; ++a)
#pragma unroll
		for (int b = -1; b <= 1; ++b)
#pragma unroll
			for (int c = -1; c <= 1; ++c) {
				const int ii = wrapn(i + a, g.n[0]), jj = wrapn(j + b, g.n[1]), kk = wrapn(k + c, g.n[2]);
				s += S.w[(a + 1) * 9 + (b + 1) * 3 + (c + 1)] * v[((long long)ii * g.n[1] + jj) * g.n[2] + kk];
			}
	out[o] = s;
}

// E -= grad phi at the owned nodes: the node gradient of the 8 cells around each node
// (evaluateDivergenceCleaningField and updateFieldByCleaning, PIC++/divergencecleanup.cpp:553-618,
// 531-551); phi is on the owned cells, its ghosts the periodic images
__global__ void k_cleanup_correct(Grid g, const double* __restrict__ phi, double* __restrict__ E) {
	const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= (long long)g.n[0] * g.n[1] * g.n[2]) return;
	int a, b, c;
	owned_ijk(g, o, a, b, c);
	// node (a+2, b+2, c+2) sits between cells a+1 .. a+2 (ghosted), i.e. owned cells a-1 .. a
	auto P = [&](int di, int dj, int dk) {
		const int ii = wrapn(a - 1 + di, g.n[0]), jj = wrapn(b - 1 + dj, g.n[1]), kk = wrapn(c - 1 + dk, g.n[2]);
		return phi[((long long)ii * g.n[1] + jj) * g.n[2] + kk];
	};
	// P(1, ., .) is the reference's [i] (the node's upper cell), P(0, ., .) its [prevI]
	const double cx = -(P(1, 1, 1) + P(1, 0, 1) + P(1, 1, 0) + P(1, 0, 0) - P(0, 1, 1) - P(0, 0, 1) - P(0, 1, 0) - P(0, 0, 0)) /
	                  (4 * g.d[0]);
	const double cy = -(P(1, 1, 1) + P(0, 1, 1) + P(1, 1, 0) + P(0, 1, 0) - P(1, 0, 1) - P(0, 0, 1) - P(1, 0, 0) - P(0, 0, 0)) /
	                  (4 * g.d[1]);
	const double cz = -(P(1, 1, 1) + P(1, 0, 1) + P(0, 1, 1) + P(0, 0, 1) - P(1, 1, 0) - P(1, 0, 0) - P(0, 1, 0) - P(0, 0, 0)) /
	                  (4 * g.d[2]);
	const long long id = node_id(g, a + 2, b + 2, c + 2);
	E[id] += cx;
	E[g.nnodes + id] += cy;
	E[2 * g.nnodes + id] += cz;
}

// right part of the magnetic cleanup per owned node: div newB (evaluateDivBgeneral,
// PIC++/fields.cpp:3019-3030, of the 8 cells around the node; -cleanUpRightPartMagnetic,
// PIC++/divergencecleanup.cpp:1302-1306)
__global__ void k_cleanup_rhs_b(Grid g, const double* __restrict__ B, double* __restrict__ rhs) {
	const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= (long long)g.n[0] * g.n[1] * g.n[2]) return;
	int a, b, c;
	owned_ijk(g, o, a, b, c);
	const int i = a + 2, j = b + 2, k = c + 2;
	const double* Bx = B;
	const double* By = B + g.ncells;
	const double* Bz = B + 2 * g.ncells;
	auto q = [&](int ii, int jj, int kk) { return cell_id(g, ii, jj, kk); };
	const double rx = (Bx[q(i, j - 1, k - 1)] + Bx[q(i, j, k - 1)] + Bx[q(i, j - 1, k)] + Bx[q(i, j, k)]) / 4.0;
	const double lx = (Bx[q(i - 1, j - 1, k - 1)] + Bx[q(i - 1, j, k - 1)] + Bx[q(i - 1, j - 1, k)] + Bx[q(i - 1, j, k)]) / 4.0;
	const double ry = (By[q(i - 1, j, k - 1)] + By[q(i, j, k - 1)] + By[q(i - 1, j, k)] + By[q(i, j, k)]) / 4.0;
	const double ly = (By[q(i - 1, j - 1, k - 1)] + By[q(i, j - 1, k - 1)] + By[q(i - 1, j - 1, k)] + By[q(i, j - 1, k)]) / 4.0;
	const double rz = (Bz[q(i - 1, j - 1, k)] + Bz[q(i, j - 1, k)] + Bz[q(i - 1, j, k)] + Bz[q(i, j, k)]) / 4.0;
	const double lz = (Bz[q(i - 1, j - 1, k - 1)] + Bz[q(i, j - 1, k - 1)] + Bz[q(i - 1, j, k - 1)] + Bz[q(i, j, k - 1)]) / 4.0;
	rhs[o] = ((rx - lx) / g.d[0]) + ((ry - ly) / g.d[1]) + ((rz - lz) / g.d[2]);
}

// B -= grad phi at the owned cells from the node potential at the cell's 8 corners
// (evaluateDivergenceCleaningFieldMagnetic and updateFieldByCleaningMagnetic,
// PIC++/divergencecleanup.cpp:1076-1159); phi on the owned nodes, its ghosts the periodic images
__global__ void k_cleanup_correct_b(Grid g, const double* __restrict__ phi, double* __restrict__ B) {
	const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= (long long)g.n[0] * g.n[1] * g.n[2]) return;
	int a, b, c;
	owned_ijk(g, o, a, b, c);
	// cell a sits between owned nodes a and a + 1
	auto P = [&](int di, int dj, int dk) {
		const int ii = wrapn(a + di, g.n[0]), jj = wrapn(b + dj, g.n[1]), kk = wrapn(c + dk, g.n[2]);
		return phi[((long long)ii * g.n[1] + jj) * g.n[2] + kk];
	};
	const double cx = -(P(1, 0, 0) + P(1, 1, 0) + P(1, 0, 1) + P(1, 1, 1) - P(0, 0, 0) - P(0, 1, 0) - P(0, 0, 1) - P(0, 1, 1)) /
	                  (4 * g.d[0]);
	const double cy = -(P(0, 1, 0) + P(1, 1, 0) + P(0, 1, 1) + P(1, 1, 1) - P(0, 0, 0) - P(1, 0, 0) - P(0, 0, 1) - P(1, 0, 1)) /
	                  (4 * g.d[1]);
	const double cz = -(P(0, 0, 1) + P(0, 1, 1) + P(1, 0, 1) + P(1, 1, 1) - P(0, 0, 0) - P(0, 1, 0) - P(1, 0, 0) - P(1, 1, 0)) /
	                  (4 * g.d[2]);
	const long long id = cell_id(g, a + 2, b + 2, c + 2);
	B[id] += cx;
	B[g.ncells + id] += cy;
	B[2 * g.ncells + id] += cz;
}

// ---------------------------------------------------------------- Fourier filtering
// component comp of a ghosted field (nodes or cells) at the owned points -> re, im = 0
__global__ void k_dft_load(Grid g, const double* __restrict__ f, long long plane, int nodes, double* __restrict__ re,
                           double* __restrict__ im) {
	const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= (long long)g.n[0] * g.n[1] * g.n[2]) return;
	int a, b, c;
	owned_ijk(g, o, a, b, c);
	const long long id = nodes ? node_id(g, a + 2, b + 2, c + 2) : cell_id(g, a + 2, b + 2, c + 2);
	re[o] = f[plane + id];
	im[o] = 0.0;
}

__global__ void k_dft_store(Grid g, const double* __restrict__ re, double* __restrict__ f, long long plane, int nodes) {
	const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= (long long)g.n[0] * g.n[1] * g.n[2]) return;
	int a, b, c;
	owned_ijk(g, o, a, b, c);
	const long long id = nodes ? node_id(g, a + 2, b + 2, c + 2) : cell_id(g, a + 2, b + 2, c + 2);
	f[plane + id] = re[o];
}

// One axis of the transform: out[.. m ..] = sum_t in[.. t ..] exp(sign 2 pi i m t / n) (divided by n
// for the inverse), one thread per output point. tw[q] = (cos, sin)(2 pi q / n) of the axis.
__global__ void __launch_bounds__(256) k_dft_axis(Grid g, int axis, int inverse, const double* __restrict__ tw,
                                                  const double* __restrict__ ire, const double* __restrict__ iim,
                                                  double* __restrict__ ore, double* __restrict__ oim) {
	const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= (long long)g.n[0] * g.n[1] * g.n[2]) return;
	int ijk[3];
	owned_ijk(g, o, ijk[0], ijk[1], ijk[2]);
	const int n = g.n[axis];
	const int m = ijk[axis];
	const long long stride = axis == 0 ? (long long)g.n[1] * g.n[2] : (axis == 1 ? g.n[2] : 1);
	const long long base = o - (long long)m * stride;
	const double sgn = inverse ? 1.0 : -1.0;
	double sr = 0.0, si = 0.0;
	int q = 0;  // m t mod n
	for (int t = 0; t < n; ++t) {
		const double c = tw[2 * q], s = sgn * tw[2 * q + 1];
		const double xr = ire[base + t * stride], xi = iim[base + t * stride];
		sr += xr * c - xi * s;
		si += xr * s + xi * c;
		q += m;
		if (q >= n) q -= n;
	}
	if (inverse) {
		sr /= n;
		si /= n;
	}
	ore[o] = sr;
	oim[o] = si;
}

// modes with |k| > kcut multiplied by -1 (filterFieldGeneral); k from the unfolded mode indices,
// the y component's k_y from the z index
__global__ void k_dft_flip(Grid g, int comp, double kcut, double* __restrict__ re, double* __restrict__ im) {
	const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= (long long)g.n[0] * g.n[1] * g.n[2]) return;
	int a, b, c;
	owned_ijk(g, o, a, b, c);
	const double kx = a * 2 * kPiC / g.L[0];
	const double ky = (comp == 1 ? c : b) * 2 * kPiC / g.L[1];
	const double kz = c * 2 * kPiC / g.L[2];
	const double kw = sqrt(kx * kx + ky * ky + kz * kz);
	if (kw > kcut) {
		re[o] = re[o] * (-1.0);
		im[o] = im[o] * (-1.0);
	}
}

CleanupStencil cleanup_stencil(const Grid& g) {
	const double X = g.d[0] * g.d[0], Y = g.d[1] * g.d[1], Z = g.d[2] * g.d[2];
	CleanupStencil S{};
	for (int a = -1; a <= 1; ++a)
		for (int b = -1; b <= 1; ++b)
			for (int c = -1; c <= 1; ++c) {
				const int nz = (a != 0) + (b != 0) + (c != 0);
				double w;
				if (nz == 0) {
					w = -1 / (2 * X) - 1 / (2 * Y) - 1 / (2 * Z);
				} else if (nz == 3) {
					w = 1 / (16 * X) + 1 / (16 * Y) + 1 / (16 * Z);
				} else if (nz == 2) {  // edges: + along the two offset axes, - along the third
					w = (a ? 1 / (8 * X) : -1 / (8 * X)) + (b ? 1 / (8 * Y) : -1 / (8 * Y)) + (c ? 1 / (8 * Z) : -1 / (8 * Z));
				} else {  // faces: + along the offset axis, - along the others
					w = (a ? 1 / (4 * X) : -1 / (4 * X)) + (b ? 1 / (4 * Y) : -1 / (4 * Y)) + (c ? 1 / (4 * Z) : -1 / (4 * Z));
				}
				S.w[(a + 1) * 9 + (b + 1) * 3 + (c + 1)] = w;
			}
	return S;
}

void launch_cleanup_rhs(const Grid& g, const double* E, const double* rho, double* rhs, hipStream_t s) {
	const long long n = (long long)g.n[0] * g.n[1] * g.n[2];
	hipLaunchKernelGGL(k_cleanup_rhs, dim3(nb(n, 256)), dim3(256), 0, s, g, E, rho, rhs);
}
void launch_cleanup_apply(const Grid& g, const CleanupStencil& S, const double* v, double* out, hipStream_t s) {
	const long long n = (long long)g.n[0] * g.n[1] * g.n[2];
	hipLaunchKernelGGL(k_cleanup_apply, dim3(nb(n, 256)), dim3(256), 0, s, g, S, v, out);
}
void launch_cleanup_correct(const Grid& g, const double* phi, double* E, hipStream_t s) {
	const long long n = (long long)g.n[0] * g.n[1] * g.n[2];
	hipLaunchKernelGGL(k_cleanup_correct, dim3(nb(n, 256)), dim3(256), 0, s, g, phi, E);
}
void launch_cleanup_rhs_b(const Grid& g, const double* B, double* rhs, hipStream_t s) {
	const long long n = (long long)g.n[0] * g.n[1] * g.n[2];
	hipLaunchKernelGGL(k_cleanup_rhs_b, dim3(nb(n, 256)), dim3(256), 0, s, g, B, rhs);
}
void launch_cleanup_correct_b(const Grid& g, const double* phi, double* B, hipStream_t s) {
	const long long n = (long long)g.n[0] * g.n[1] * g.n[2];
	hipLaunchKernelGGL(k_cleanup_correct_b, dim3(nb(n, 256)), dim3(256), 0, s, g, phi, B);
}
void launch_filter_component(const Grid& g, double* f, long long plane, bool nodes, int comp, double kcut, const double* const tw[3],
                             double* re, double* im, double* re2, double* im2, hipStream_t s) {
	const long long n = (long long)g.n[0] * g.n[1] * g.n[2];
	const unsigned blocks = nb(n, 256);
	hipLaunchKernelGGL(k_dft_load, dim3(blocks), dim3(256), 0, s, g, f, plane, nodes ? 1 : 0, re, im);
	// forward x, y, z then inverse x, y, z (fourierTranslation), ping-ponging between the buffers
	double* a_re = re;
	double* a_im = im;
	double* b_re = re2;
	double* b_im = im2;
	for (int inv = 0; inv < 2; ++inv) {
		for (int axis = 0; axis < 3; ++axis) {
			hipLaunchKernelGGL(k_dft_axis, dim3(blocks), dim3(256), 0, s, g, axis, inv, tw[axis], a_re, a_im, b_re, b_im);
			std::swap(a_re, b_re);
			std::swap(a_im, b_im);
		}
		if (inv == 0) hipLaunchKernelGGL(k_dft_flip, dim3(blocks), dim3(256), 0, s, g, comp, kcut, a_re, a_im);
	}
	hipLaunchKernelGGL(k_dft_store, dim3(blocks), dim3(256), 0, s, g, a_re, f, plane, nodes ? 1 : 0);
}

}  // namespace picgpu
