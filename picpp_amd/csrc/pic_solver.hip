// Krylov kernels of the mixed-precision Maxwell solve: FP32 inner GMRES with FP64
// iterative refinement (outer residual, Hessenberg least squares and dot-product
// accumulation stay FP64).
//
// The operator is applied in factored form. With w = mu E at every node, the 37 entries of
// createInternalEquation{X,Y,Z} (PIC++/fields.cpp:1396-1509, :1961-2072, :2521-2632)
// regroup exactly into
//   (A E)_c = E_c - w_c - C lap_h E_c + C d2_c w_c
//             + C / (4 dc dd) sum_{d != c} [w_d(+c+d) + w_d(-c-d) - w_d(-c+d) - w_d(+c-d)],
// C = (c' theta dt)^2, lap_h the 7-point Laplacian, d2_c the second difference along c.
// With the wall boundary the x range of the unknowns is nodes 2 .. n0+3 and the first and the
// last two x planes carry the boundary rows; the interior rows then never wrap along x.
#include "pic_device.hpp"
#include "pic_solver.hpp"

namespace picgpu {

// w = mu v on the halo layout: mu of the node (ghost nodes included), v of its periodic image
// among the unknowns. The stencil then reads w at the neighbouring nodes without a wrap.
template <class T>
__global__ void k_mu_times(Grid g, const T* __restrict__ mu, const T* __restrict__ v, T* __restrict__ w) {
	const long long H = halo_count(g);
	long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (h >= H) return;
	const int n0 = g.nu0, n1 = g.n[1], n2 = g.n[2];
	int hk = (int)(h % (n2 + 2));
	int hj = (int)((h / (n2 + 2)) % (n1 + 2));
	int hi = (int)(h / ((long long)(n2 + 2) * (n1 + 2)));
	int i = hi == 0 ? n0 - 1 : (hi == n0 + 1 ? 0 : hi - 1);
	int j = hj == 0 ? n1 - 1 : (hj == n1 + 1 ? 0 : hj - 1);
	int k = hk == 0 ? n2 - 1 : (hk == n2 + 1 ? 0 : hk - 1);
	const long long L = g.nown;
	const long long o = ((long long)i * n1 + j) * n2 + k;
	T e0 = v[o], e1 = v[L + o], e2 = v[2 * L + o];
#pragma unroll
	for (int c = 0; c < 3; ++c) w[c * H + h] = mu[(c * 3 + 0) * H + h] * e0 + mu[(c * 3 + 1) * H + h] * e1 + mu[(c * 3 + 2) * H + h] * e2;
}

template <class T>
__global__ void __launch_bounds__(256) k_stencil(Grid g, const T* __restrict__ v, const T* __restrict__ w, T C, T* __restrict__ out) {
	long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= g.nown) return;
	const int n0 = g.nu0, n1 = g.n[1], n2 = g.n[2];
	int k = (int)(o % n2);
	int j = (int)((o / n2) % n1);
	int i = (int)(o / ((long long)n2 * n1));
	if (g.wall && (i == 0 || i >= g.n[0])) {
		// wall rows (x: E_x(2) - E_x(3), tangential: identity) and free right rows (identity),
		// PIC++/fields.cpp:690-781
		const long long L = g.nown;
		out[o] = i == 0 ? v[o] + (T)(-1.0) * v[o + (long long)n1 * n2] : v[o];
		out[L + o] = v[L + o];
		out[2 * L + o] = v[2 * L + o];
		return;
	}
	int I[3] = {i == 0 ? n0 - 1 : i - 1, i, i == n0 - 1 ? 0 : i + 1};
	int Jj[3] = {j == 0 ? n1 - 1 : j - 1, j, j == n1 - 1 ? 0 : j + 1};
	int K[3] = {k == 0 ? n2 - 1 : k - 1, k, k == n2 - 1 ? 0 : k + 1};
	auto id = [&](int a, int b, int c) { return ((long long)I[a + 1] * n1 + Jj[b + 1]) * n2 + K[c + 1]; };
	const long long L = g.nown;
	const long long H = halo_count(g);
	auto hid = [&](int a, int b, int c) { return halo_id(g, i + 1 + a, j + 1 + b, k + 1 + c); };
	T id2[3] = {(T)(1.0 / (g.d[0] * g.d[0])), (T)(1.0 / (g.d[1] * g.d[1])), (T)(1.0 / (g.d[2] * g.d[2]))};
	T q4[3][3];
	for (int a = 0; a < 3; ++a)
		for (int b = 0; b < 3; ++b) q4[a][b] = (T)(1.0 / (4.0 * g.d[a] * g.d[b]));
	long long nb[3][2] = {{id(-1, 0, 0), id(1, 0, 0)}, {id(0, -1, 0), id(0, 1, 0)}, {id(0, 0, -1), id(0, 0, 1)}};
	long long hb[3][2] = {{hid(-1, 0, 0), hid(1, 0, 0)}, {hid(0, -1, 0), hid(0, 1, 0)}, {hid(0, 0, -1), hid(0, 0, 1)}};
#pragma unroll
	for (int c = 0; c < 3; ++c) {
		const T* vc = v + c * L;
		const T* wc = w + c * H;
		T vo = vc[o], wo = wc[hid(0, 0, 0)];
		T lap = 0;
#pragma unroll
		for (int a = 0; a < 3; ++a) lap += (vc[nb[a][0]] - 2 * vo + vc[nb[a][1]]) * id2[a];
		T d2w = (wc[hb[c][0]] - 2 * wo + wc[hb[c][1]]) * id2[c];
		T cross = 0;
#pragma unroll
		for (int d = 0; d < 3; ++d) {
			if (d == c) continue;
			const T* wd = w + d * H;
			int pp[3] = {0, 0, 0}, mm[3] = {0, 0, 0}, mp[3] = {0, 0, 0}, pm[3] = {0, 0, 0};
			pp[c] = 1;
			pp[d] = 1;
			mm[c] = -1;
			mm[d] = -1;
			mp[c] = -1;
			mp[d] = 1;
			pm[c] = 1;
			pm[d] = -1;
			cross += (wd[hid(pp[0], pp[1], pp[2])] + wd[hid(mm[0], mm[1], mm[2])] - wd[hid(mp[0], mp[1], mp[2])] -
			          wd[hid(pm[0], pm[1], pm[2])]) * q4[c][d];
		}
		out[c * L + o] = vo - wo - C * lap + C * (d2w + cross);
	}
}

// sum over blocks of a . b in FP64 (inputs T)
template <class T>
__global__ void __launch_bounds__(256) k_dotT(const T* __restrict__ a, const T* __restrict__ b, long long len, double* __restrict__ partial) {
	__shared__ double red[256];
	double s = 0.0;
	for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < len; t += (long long)gridDim.x * blockDim.x)
		s += (double)a[t] * (double)b[t];
	red[threadIdx.x] = s;
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_final(const double* __restrict__ partial, int n, double* __restrict__ out, int do_sqrt) {
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

template <class T>
__global__ void k_axpyT(T* __restrict__ w, const T* __restrict__ v, const double* __restrict__ h, long long len) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < len) w[t] -= (T)(*h) * v[t];
}

template <class T>
__global__ void k_normalizeT(T* __restrict__ w, const double* __restrict__ h, long long len) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= len) return;
	double hv = *h;
	w[t] = hv > 0 ? (T)((double)w[t] / hv) : (T)0;
}

// v0 = r / norm in T
template <class T>
__global__ void k_scaled_copy(const double* __restrict__ r, T* __restrict__ v0, const double* __restrict__ norm, long long len) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < len) v0[t] = (T)(r[t] / *norm);
}

// x += scale * sum_m y_m V_m  (FP64 accumulation)
template <class T>
__global__ void k_update_x(double* __restrict__ x, const T* const* __restrict__ V, const double* __restrict__ y, int m,
                           const double* __restrict__ scale, long long len) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= len) return;
	double s = 0.0;
	for (int q = 0; q < m; ++q) s += (double)V[q][t] * y[q];
	x[t] += s * (*scale);
}

__global__ void k_residual(const double* __restrict__ b, const double* __restrict__ Ax, double* __restrict__ r, long long len) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < len) r[t] = b[t] - Ax[t];
}

__global__ void k_to_f32(const double* __restrict__ a, float* __restrict__ b, long long n) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < n) b[t] = (float)a[t];
}

namespace {
inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
}  // namespace

template <class T>
void launch_apply_factored(const Grid& g, const T* mu, const T* v, T* w, double C, T* out, hipStream_t s) {
	hipLaunchKernelGGL(k_mu_times<T>, dim3(nblk(halo_count(g), 256)), dim3(256), 0, s, g, mu, v, w);
	hipLaunchKernelGGL(k_stencil<T>, dim3(nblk(g.nown, 256)), dim3(256), 0, s, g, v, w, (T)C, out);
}
template <class T>
void launch_dotT(const T* a, const T* b, long long len, double* partial, double* out, bool do_sqrt, hipStream_t s) {
	hipLaunchKernelGGL(k_dotT<T>, dim3(kSolverDotBlocks), dim3(256), 0, s, a, b, len, partial);
	hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, s, partial, kSolverDotBlocks, out, do_sqrt ? 1 : 0);
}
template <class T>
void launch_axpyT(T* w, const T* v, const double* h, long long len, hipStream_t s) {
	hipLaunchKernelGGL(k_axpyT<T>, dim3(nblk(len, 256)), dim3(256), 0, s, w, v, h, len);
}
template <class T>
void launch_normalizeT(T* w, const double* h, long long len, hipStream_t s) {
	hipLaunchKernelGGL(k_normalizeT<T>, dim3(nblk(len, 256)), dim3(256), 0, s, w, h, len);
}
template <class T>
void launch_scaled_copy(const double* r, T* v0, const double* norm, long long len, hipStream_t s) {
	hipLaunchKernelGGL(k_scaled_copy<T>, dim3(nblk(len, 256)), dim3(256), 0, s, r, v0, norm, len);
}
template <class T>
void launch_update_x(double* x, const T* const* V, const double* y, int m, const double* scale, long long len, hipStream_t s) {
	hipLaunchKernelGGL(k_update_x<T>, dim3(nblk(len, 256)), dim3(256), 0, s, x, V, y, m, scale, len);
}
void launch_residual(const double* b, const double* Ax, double* r, long long len, hipStream_t s) {
	hipLaunchKernelGGL(k_residual, dim3(nblk(len, 256)), dim3(256), 0, s, b, Ax, r, len);
}
void launch_to_f32(const double* a, float* b, long long n, hipStream_t s) {
	hipLaunchKernelGGL(k_to_f32, dim3(nblk(n, 256)), dim3(256), 0, s, a, b, n);
}

template void launch_apply_factored<float>(const Grid&, const float*, const float*, float*, double, float*, hipStream_t);
template void launch_apply_factored<double>(const Grid&, const double*, const double*, double*, double, double*, hipStream_t);
template void launch_dotT<float>(const float*, const float*, long long, double*, double*, bool, hipStream_t);
template void launch_dotT<double>(const double*, const double*, long long, double*, double*, bool, hipStream_t);
template void launch_axpyT<float>(float*, const float*, const double*, long long, hipStream_t);
template void launch_normalizeT<float>(float*, const double*, long long, hipStream_t);
template void launch_scaled_copy<float>(const double*, float*, const double*, long long, hipStream_t);
template void launch_update_x<float>(double*, const float* const*, const double*, int, const double*, long long, hipStream_t);

}  // namespace picgpu
