// Diagnostics of the reference's general.dat rows: field sums and maxima over updateEnergy's
// range (PIC++/simulation.cpp:1506-1640, updateParameters :1707-1880) and the theoretical
// energy and momentum bookkeeping (updateTheoreticalEnergy :1399-1504, the mover's reflection
// updates PIC++/mover.cpp:256, 329, 439, injectNewParticles :994, addToPreserveChargeGlobal
// PIC++/initialization.cpp:8319). Every sum runs in a fixed order: the same bits in every run.
#include <math.h>

#include "pic_general.hpp"

namespace picgpu {

namespace {

constexpr double kPiG = 3.141592653589793115997963468544185161590576171875;  // 4 atan2(1, 1)

inline unsigned nbg(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__device__ inline double at(const double* f, const int d[3], int i, int j, int k) {
	return f[((long long)i * d[1] + j) * d[2] + k];
}

__device__ inline void cross3(const double a[3], const double b[3], double o[3]) {
	o[0] = a[1] * b[2] - a[2] * b[1];
	o[1] = a[2] * b[0] - a[0] * b[2];
	o[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ inline void field_e(const FieldView& v, int i, int j, int k, double e[3]) {
	for (int c = 0; c < 3; ++c) e[c] = at(v.e[c], v.de[c], i, j, k);
}

__device__ inline void field_b(const FieldView& v, int i, int j, int k, double b[3]) {
	for (int c = 0; c < 3; ++c) b[c] = at(v.b[c], v.db[c], i, j, k);
}

// Block sums of `rows` per-thread values (256 threads), tree order fixed; max for the rows at and
// above first_max.
template <int ROWS>
__device__ inline void block_reduce(double (&acc)[ROWS], int first_max, double (*red)[256]) {
	for (int q = 0; q < ROWS; ++q) red[q][threadIdx.x] = acc[q];
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if ((int)threadIdx.x < w)
			for (int q = 0; q < ROWS; ++q)
				red[q][threadIdx.x] = q < first_max ? red[q][threadIdx.x] + red[q][threadIdx.x + w]
				                                    : fmax(red[q][threadIdx.x], red[q][threadIdx.x + w]);
		__syncthreads();
	}
}

__global__ void __launch_bounds__(256) k_general_fields(Grid g, FieldView v, double* __restrict__ partial) {
	__shared__ double red[kGeneralRows][256];
	double acc[kGeneralRows] = {};
	const int nx = g.wall ? g.N[0] - 1 : g.n[0], x_lo = g.wall ? 0 : 2;
	const long long cols = (long long)nx * g.n[1] * g.n[2];
	for (long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x; o < cols; o += (long long)gridDim.x * blockDim.x) {
		const int k = (int)(o % g.n[2]) + owned_lo(g, 2);
		const int j = (int)((o / g.n[2]) % g.n[1]) + 2;
		const int i = (int)(o / ((long long)g.n[2] * g.n[1])) + x_lo;
		double e[3], b[3], ea[3];
		field_e(v, i, j, k, e);
		field_b(v, i, j, k, b);
		if (v.average_e) {
			for (int c = 0; c < 3; ++c) {
				const double* f = v.e[c];
				const int* d = v.de[c];
				ea[c] = (at(f, d, i, j, k) + at(f, d, i, j, k + 1) + at(f, d, i, j + 1, k) + at(f, d, i, j + 1, k + 1) +
				         at(f, d, i + 1, j, k) + at(f, d, i + 1, j, k + 1) + at(f, d, i + 1, j + 1, k) +
				         at(f, d, i + 1, j + 1, k + 1)) / 8;
			}
		} else {
			for (int c = 0; c < 3; ++c) ea[c] = e[c];
		}
		double cr[3];
		cross3(ea, b, cr);
		for (int c = 0; c < 3; ++c) {
			acc[c] += e[c] * e[c];
			acc[3 + c] += b[c] * b[c];
			acc[6 + c] += cr[c];
		}
		acc[9] = fmax(acc[9], e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
		acc[10] = fmax(acc[10], b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
	}
	block_reduce<kGeneralRows>(acc, 9, red);
	if (threadIdx.x == 0)
		for (int q = 0; q < kGeneralRows; ++q) partial[q * gridDim.x + blockIdx.x] = red[q][0];
}

// One block per row: the blocks' partials in a fixed order (sums, maxima for rows 9 and 10).
__global__ void __launch_bounds__(256) k_general_final(const double* __restrict__ partial, int n, double* __restrict__ out) {
	__shared__ double red[1][256];
	const int r = blockIdx.x;
	const bool mx = r >= 9;
	double a[1] = {0.0};
	for (int t = threadIdx.x; t < n; t += blockDim.x) {
		const double v = partial[(long long)r * n + t];
		a[0] = mx ? fmax(a[0], v) : a[0] + v;
	}
	block_reduce<1>(a, mx ? 0 : 1, red);
	if (threadIdx.x == 0) out[r] = red[0][0];
}

__device__ inline int species_of(const SpeciesTable& sp, const int* species_start, const uint32_t* keys, long long ncells,
                                 long long i) {
	if (i >= species_start[sp.count]) return (int)((long long)(keys[i] >> 3) / ncells);
	int s = 0;
	while (s + 1 < sp.count && i >= species_start[s + 1]) ++s;
	return s;
}

// w (gamma - 1) m and w p of particle i of species s (the FP32 form keeps u = p / m)
__device__ inline void particle_terms(const SpeciesTable& sp, const ParticleArrays& P, const ParticleArrays32& Q, bool fp32,
                                      int s, long long i, double t[4]) {
	const double m = sp.mass[s], w = sp.weight[s];
	if (fp32) {
		const float u[3] = {Q.u[0][i], Q.u[1][i], Q.u[2][i]};
		const float u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
		const float gm1 = u2 / (sqrtf(1.0f + u2) + 1.0f);
		t[0] = (double)gm1 * m * w;
		for (int c = 0; c < 3; ++c) t[1 + c] = (double)u[c] * m * w;
	} else {
		const double p[3] = {P.p[0][i], P.p[1][i], P.p[2][i]};
		const double gm = sqrt((p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) / (m * m) + 1.0);
		t[0] = ((gm - 1) * m) * w;
		for (int c = 0; c < 3; ++c) t[1 + c] = p[c] * w;
	}
}

__global__ void __launch_bounds__(256) k_theo_particles(SpeciesTable sp, ParticleArrays P, ParticleArrays32 Q, bool fp32,
                                                        const int* __restrict__ species_start, const uint32_t* __restrict__ keys,
                                                        long long ncells, const uint32_t* __restrict__ idx, long long first,
                                                        long long count, double sign, double* __restrict__ theo) {
	__shared__ double red[4][256];
	double acc[4] = {};
	for (long long q = threadIdx.x; q < count; q += blockDim.x) {
		const long long i = idx ? (long long)idx[q] : first + q;
		double t[4];
		particle_terms(sp, P, Q, fp32, species_of(sp, species_start, keys, ncells, i), i, t);
		for (int c = 0; c < 4; ++c) acc[c] += t[c];
	}
	block_reduce<4>(acc, 4, red);
	if (threadIdx.x == 0)
		for (int c = 0; c < 4; ++c) theo[c] += sign * red[c][0];
}

__global__ void __launch_bounds__(256) k_theo_bounce(SpeciesTable sp, ParticleArrays P, ParticleArrays32 Q, bool fp32,
                                                     const int* __restrict__ species_start, const uint64_t* __restrict__ mask,
                                                     long long nwords, long long n, double* __restrict__ partial) {
	__shared__ double red[1][256];
	const long long chunk = (nwords + gridDim.x - 1) / gridDim.x;
	const long long w0 = (long long)blockIdx.x * chunk, w1 = min(w0 + chunk, nwords);
	double a[1] = {0.0};
	for (long long w = w0 + threadIdx.x; w < w1; w += blockDim.x) {
		unsigned long long bits = mask[w];
		while (bits) {
			const int b = __ffsll((long long)bits) - 1;
			bits &= bits - 1;
			const long long i = w * 64 + b;
			if (i >= n) break;
			int s = 0;
			while (s + 1 < sp.count && i >= species_start[s + 1]) ++s;
			const double px = fp32 ? (double)Q.u[0][i] * sp.mass[s] : P.p[0][i];
			a[0] += px * (2 * sp.weight[s]);
		}
	}
	block_reduce<1>(a, 1, red);
	if (threadIdx.x == 0) partial[blockIdx.x] = red[0][0];
}

__global__ void __launch_bounds__(256) k_theo_add_partials(const double* __restrict__ partial, int n, double* __restrict__ target) {
	__shared__ double red[1][256];
	double a[1] = {0.0};
	for (int t = threadIdx.x; t < n; t += blockDim.x) a[0] += partial[t];
	block_reduce<1>(a, 1, red);
	if (threadIdx.x == 0) *target += red[0][0];
}

__global__ void __launch_bounds__(256) k_theo_poynting(Grid g, FieldView v, double scale, double* __restrict__ theo) {
	__shared__ double red[3][256];
	double acc[3] = {};
	const long long cols = (long long)g.n[1] * g.n[2];
	const int il = 2, ir = g.N[0] - 2;
	for (long long o = threadIdx.x; o < cols; o += blockDim.x) {
		const int k = (int)(o % g.n[2]) + owned_lo(g, 2), j = (int)(o / g.n[2]) + 2;
		double el[3], bl[3], er[3], br[3], cl[3], cr[3];
		field_e(v, il, j, k, el);
		field_b(v, il, j, k, bl);
		field_e(v, ir, j, k, er);
		field_b(v, ir, j, k, br);
		cross3(el, bl, cl);
		cross3(er, br, cr);
		for (int c = 0; c < 3; ++c) acc[c] += cl[c] - cr[c];
	}
	block_reduce<3>(acc, 3, red);
	if (threadIdx.x == 0) {
		theo[0] += red[0][0] * scale;
		for (int c = 0; c < 3; ++c) theo[1 + c] += red[c][0] * scale;
	}
}

}  // namespace

FieldView implicit_view(const Grid& g, const double* E, const double* B) {
	FieldView v{};
	for (int c = 0; c < 3; ++c) {
		v.e[c] = E + c * g.nnodes;
		v.b[c] = B + c * g.ncells;
		for (int a = 0; a < 3; ++a) {
			v.de[c][a] = g.N[a] + 1;
			v.db[c][a] = g.N[a];
		}
	}
	v.average_e = true;
	return v;
}

FieldView yee_view(const YeeGrid& y) {
	FieldView v{};
	for (int c = 0; c < 3; ++c) {
		v.e[c] = y.E[c];
		v.b[c] = y.B[c];
		for (int a = 0; a < 3; ++a) {
			v.de[c][a] = y.dE[c][a];
			v.db[c][a] = y.dB[c][a];
		}
	}
	v.average_e = false;
	return v;
}

void launch_general_fields(const Grid& g, const FieldView& v, double* partial, double* out, hipStream_t s) {
	hipLaunchKernelGGL(k_general_fields, dim3(kEnergyBlocks), dim3(256), 0, s, g, v, partial);
	hipLaunchKernelGGL(k_general_final, dim3(kGeneralRows), dim3(256), 0, s, partial, kEnergyBlocks, out);
}

void launch_theo_particles(const SpeciesTable& sp, const ParticleArrays& P, const ParticleArrays32& Q, bool fp32,
                           const int* species_start, const uint32_t* keys, long long ncells, const uint32_t* idx,
                           long long first, long long count, double sign, double* theo, hipStream_t s) {
	if (count <= 0) return;
	hipLaunchKernelGGL(k_theo_particles, dim3(1), dim3(256), 0, s, sp, P, Q, fp32, species_start, keys, ncells, idx, first, count,
	                   sign, theo);
}

void launch_theo_bounce(const SpeciesTable& sp, const ParticleArrays& P, const ParticleArrays32& Q, bool fp32,
                        const int* species_start, const uint64_t* bounce_mask, long long n, double* partial, double* theo,
                        hipStream_t s) {
	if (n <= 0) return;
	const long long nwords = (n + 63) / 64;
	hipLaunchKernelGGL(k_theo_bounce, dim3(kEnergyBlocks), dim3(256), 0, s, sp, P, Q, fp32, species_start, bounce_mask, nwords, n,
	                   partial);
	hipLaunchKernelGGL(k_theo_add_partials, dim3(1), dim3(256), 0, s, partial, kEnergyBlocks, theo + 1);
}

void launch_theo_poynting(const Grid& g, const FieldView& v, double dt, double* theo, hipStream_t s) {
	hipLaunchKernelGGL(k_theo_poynting, dim3(1), dim3(256), 0, s, g, v, dt * g.d[1] * g.d[2] / (4 * kPiG), theo);
}

}  // namespace picgpu
