// The explicit (BUNEMAN / Tristan) step of PIC++ on the Yee grid, FP64:
// - tristanEvaluateBhalfStep / tristanEvaluateE (PIC++/fields.cpp:21-150, curls at
//   fields.cpp:3273-3338) as one thread per point of the owned box;
// - exchangeBuneman{E,B}field and exchangeBunemanFlux, one-rank periodic branches
//   (PIC++/fields_exchange.cpp:2523-4200, momentum_exchange.cpp:1806-2310), one thread per
//   line along the exchanged axis;
// - moveParticleTristan (PIC++/mover.cpp:718-808) with the field interpolation of
//   correlationBunemanEBfieldsWithoutMaps (correlation.cpp:441-626), fused with the zig-zag
//   current of the same move (addParticleFluxZigzag, momentum_collecting.cpp:85-249).
//
// The zig-zag current is summed without floating-point atomics: every contribution is
// rounded to a 64-bit fixed-point value and added with integer atomics. Integer addition is
// associative, so the sums, and everything after them, are the same in every run whatever the
// order of the atomics. The scale, a power of two, is chosen per step from the largest number
// of particles in a cell and a bound on the particles' new speeds (their largest |u| before the
// step, as the previous push recorded it, plus what the largest |E| can add), so that no entry
// can overflow: each contribution is rounded by at most half a unit of 2^-62 times that bound
// on an entry. With m particles in the fullest cell and speeds up to v_max that unit is about
// 2^-53 m v_max |q w / V|, against contributions of |q w / V| |v|: the relative rounding of a
// particle's contribution is about 1e-16 m v_max / |v|, and the sums' errors average out over
// the particles of an entry.
#include <math.h>

#include "pic_explicit.hpp"

namespace picgpu {

namespace {

constexpr double kCprimeX = 1.025;  // speed_of_light_correction (PIC++/constants.h:45)

inline unsigned nb(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__device__ inline long long yi(const int d[3], int i, int j, int k) { return ((long long)i * d[1] + j) * d[2] + k; }

// tristanEvaluateBhalfStep over the owned box i, j, k in [2, N-2]: Bx on i <= N-2, j, k < N-2;
// By on j <= N-2; Bz on k <= N-2 (fields.cpp:26-74). Along a collapsed z only k = 0, without
// the z derivatives (fields.cpp:36-39, :3280-3282).
__global__ void k_b_half(Grid g, YeeGrid y, double cdtd2) {
	const int ni = g.N[0] - 3, nj = g.N[1] - 3, nk = g.cz ? 1 : g.N[2] - 3;
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= (long long)ni * nj * nk) return;
	const int k = owned_lo(g, 2) + (int)(t % nk), j = 2 + (int)((t / nk) % nj), i = 2 + (int)(t / ((long long)nk * nj));
	const int hx = g.N[0] - 2, hy = g.N[1] - 2, hz = g.cz ? 1 : g.N[2] - 2;
	const double* Ex = y.E[0];
	const double* Ey = y.E[1];
	const double* Ez = y.E[2];
	if (j < hy && k < hz) {  // evaluateBunemanRotEx (fields.cpp:3273-3284)
		double r = 0;
		r += (Ez[yi(y.dE[2], i, j + 1, k)] - Ez[yi(y.dE[2], i, j, k)]) / g.d[1];
		if (!g.cz) r -= (Ey[yi(y.dE[1], i, j, k + 1)] - Ey[yi(y.dE[1], i, j, k)]) / g.d[2];
		double& b = y.B[0][yi(y.dB[0], i, j, k)];
		b = b - r * cdtd2;
	}
	if (i < hx && k < hz) {  // evaluateBunemanRotEy (:3286-3294)
		double r = -((Ez[yi(y.dE[2], i + 1, j, k)] - Ez[yi(y.dE[2], i, j, k)]) / g.d[0]);
		if (!g.cz) r += (Ex[yi(y.dE[0], i, j, k + 1)] - Ex[yi(y.dE[0], i, j, k)]) / g.d[2];
		double& b = y.B[1][yi(y.dB[1], i, j, k)];
		b = b - r * cdtd2;
	}
	if (i < hx && j < hy) {  // evaluateBunemanRotEz (:3296-3304)
		double r = (Ey[yi(y.dE[1], i + 1, j, k)] - Ey[yi(y.dE[1], i, j, k)]) / g.d[0];
		r -= (Ex[yi(y.dE[0], i, j + 1, k)] - Ex[yi(y.dE[0], i, j, k)]) / g.d[1];
		double& b = y.B[2][yi(y.dB[2], i, j, k)];
		b = b - r * cdtd2;
	}
}

// tristanEvaluateE over the owned box: Ex on i < N-2, Ey on j < N-2, Ez on k < N-2
// (fields.cpp:81-150, periodic x). Along a collapsed z only k = 0, without the z derivatives
// (fields.cpp:96-99, :3313-3315, :3323-3325).
__global__ void k_e(Grid g, YeeGrid y, double dt, double four_pi, double e0y, double e0z) {
	const int ni = g.N[0] - 3, nj = g.N[1] - 3, nk = g.cz ? 1 : g.N[2] - 3;
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= (long long)ni * nj * nk) return;
	const int k = owned_lo(g, 2) + (int)(t % nk), j = 2 + (int)((t / nk) % nj), i = 2 + (int)(t / ((long long)nk * nj));
	const int hx = g.N[0] - 2, hy = g.N[1] - 2, hz = g.cz ? 1 : g.N[2] - 2;
	const double* Bx = y.B[0];
	const double* By = y.B[1];
	const double* Bz = y.B[2];
	if (i < hx) {  // evaluateBunemanRotBx (fields.cpp:3306-3317)
		double r = 0;
		r += (Bz[yi(y.dB[2], i, j, k)] - Bz[yi(y.dB[2], i, j - 1, k)]) / g.d[1];
		if (!g.cz) r -= (By[yi(y.dB[1], i, j, k)] - By[yi(y.dB[1], i, j, k - 1)]) / g.d[2];
		const long long q = yi(y.dE[0], i, j, k);
		y.E[0][q] = y.E[0][q] + (kCprimeX * r - four_pi * y.J[0][q]) * dt;
	}
	if (j < hy) {  // evaluateBunemanRotBy (:3319-3327)
		double r = -((Bz[yi(y.dB[2], i, j, k)] - Bz[yi(y.dB[2], i - 1, j, k)]) / g.d[0]);
		if (!g.cz) r += (Bx[yi(y.dB[0], i, j, k)] - Bx[yi(y.dB[0], i, j, k - 1)]) / g.d[2];
		const long long q = yi(y.dE[1], i, j, k);
		y.E[1][q] = y.E[1][q] + (kCprimeX * r - four_pi * y.J[1][q]) * dt;
		// the wall: zero at the conductor node 2 and E0 at the free end (fields.cpp:117-122)
		if (g.wall && i <= 2) y.E[1][q] = 0;
		if (g.wall && i >= g.N[0] - 2) y.E[1][q] = e0y;
	}
	if (k < hz) {  // evaluateBunemanRotBz (:3329-3338)
		double r = (By[yi(y.dB[1], i, j, k)] - By[yi(y.dB[1], i - 1, j, k)]) / g.d[0];
		r -= (Bx[yi(y.dB[0], i, j, k)] - Bx[yi(y.dB[0], i, j - 1, k)]) / g.d[1];
		const long long q = yi(y.dE[2], i, j, k);
		y.E[2][q] = y.E[2][q] + (kCprimeX * r - four_pi * y.J[2][q]) * dt;
		if (g.wall && i <= 2) y.E[2][q] = 0;  // (fields.cpp:136-141)
		if (g.wall && i >= g.N[0] - 2) y.E[2][q] = e0z;
	}
}

struct Yee3 {
	double* f[3];
	int d[3][3];
};

// One line along axis ax of component c; lines of the three components one after another.
__device__ inline bool line_of(const Yee3& y, int ax, long long t, int& c, int idx[3]) {
	const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
	for (c = 0; c < 3; ++c) {
		const long long lines = (long long)y.d[c][o1] * y.d[c][o2];
		if (t < lines) {
			idx[o1] = (int)(t / y.d[c][o2]);
			idx[o2] = (int)(t % y.d[c][o2]);
			return true;
		}
		t -= lines;
	}
	return false;
}

// Periodic ghost copy along ax: an axis of N points (cells) takes f[N-2+i] = f[2+i],
// f[i] = f[N-4+i] (i = 0, 1); an axis of N + 1 points (nodes) f[N-1+i] = f[3+i],
// f[i] = f[N-4+i], then f[N-2] = f[2].
__global__ void k_yee_exchange(Yee3 y, int ax, int Nax) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	int c, idx[3];
	if (!line_of(y, ax, t, c, idx)) return;
	const int* d = y.d[c];
	double* f = y.f[c];
	const bool node = d[ax] == Nax + 1;
	auto at = [&](int v) -> double& {
		idx[ax] = v;
		return f[yi(d, idx[0], idx[1], idx[2])];
	};
	for (int i = 0; i <= 1; ++i) {
		if (node)
			at(Nax - 1 + i) = at(3 + i);
		else
			at(Nax - 2 + i) = at(2 + i);
		at(i) = at(Nax - 4 + i);
	}
	if (node) at(Nax - 2) = at(2);
}

// The wall's ghost planes along x (the one-rank non-periodic branches of exchangeBuneman*AlongX,
// PIC++/fields_exchange.cpp:2537-4215, SUPER_CONDUCTOR_LEFT): planes 0 and 1 take zero for the
// components in zero_mask (bit c) and plane 2's value for the others; the right end keeps its
// values.
__global__ void k_yee_wall_x(Yee3 y, int zero_mask) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	int c, idx[3];
	if (!line_of(y, 0, t, c, idx)) return;
	const int* d = y.d[c];
	double* f = y.f[c];
	const double v = (zero_mask >> c) & 1 ? 0.0 : f[yi(d, 2, idx[1], idx[2])];
	for (int i = 0; i <= 1; ++i) f[yi(d, i, idx[1], idx[2])] = v;
}

// Periodic fold of the current along ax: N points f[N-1-i] += f[3-i], f[3-i] = f[N-1-i]
// (i = 0..3); N + 1 points f[N-i] += f[4-i], f[4-i] = f[N-i] (i = 0..4).
__global__ void k_yee_fold(Yee3 y, int ax, int Nax) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	int c, idx[3];
	if (!line_of(y, ax, t, c, idx)) return;
	const int* d = y.d[c];
	double* f = y.f[c];
	const bool node = d[ax] == Nax + 1;
	auto at = [&](int v) -> double& {
		idx[ax] = v;
		return f[yi(d, idx[0], idx[1], idx[2])];
	};
	if (node) {
		for (int i = 0; i <= 4; ++i) {
			at(Nax - i) += at(4 - i);
			at(4 - i) = at(Nax - i);
		}
	} else {
		for (int i = 0; i <= 3; ++i) {
			at(Nax - 1 - i) += at(3 - i);
			at(3 - i) = at(Nax - 1 - i);
		}
	}
}

// Along a collapsed z every exchange and every current sum copies plane 0 onto plane 1 of the
// arrays with two planes (Ex, Ey, Jx, Jy, Bz; the 2-D branches of exchangeBuneman*AlongZ and
// sumBuneman*AlongZ, PIC++/fields_exchange.cpp:2823-2829, momentum_exchange.cpp:1980-1986); the
// others have one plane.
__global__ void k_yee_copy_z(Yee3 y) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	int c, idx[3];
	if (!line_of(y, 2, t, c, idx)) return;
	const int* d = y.d[c];
	if (d[2] != 2) return;
	double* f = y.f[c];
	f[yi(d, idx[0], idx[1], 1)] = f[yi(d, idx[0], idx[1], 0)];
}

__global__ void k_fixed_to_double(const long long* __restrict__ acc, double* __restrict__ J, long long n, double inv_scale) {
	long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	J[t] = (double)acc[t] * inv_scale;
}

// A 64-bit DPP move: lanes whose source is outside the pattern or whose row is masked off
// read 0.
template <int CTRL, int ROWS>
__device__ inline long long dpp64(long long v) {
	const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(unsigned long long)v, CTRL, ROWS, 0xF, false);
	const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)((unsigned long long)v >> 32), CTRL, ROWS, 0xF, false);
	return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ inline double wave_max(double v) {
	for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o));
	return v;
}

// The largest of the wave's 64 non-negative doubles, given and returned as bit patterns (which
// order like the values), in every lane: the DPP steps of wave_run_sums below with max for the
// sum (a lane whose source is outside a pattern reads 0, which changes no maximum), after which
// lane 63 holds the maximum. Needs all 64 lanes of the wave active.
__device__ inline long long wave_max_bits(long long b) {
	b = max(b, dpp64<0x111, 0xF>(b));  // row_shr:1
	b = max(b, dpp64<0x112, 0xF>(b));  // row_shr:2
	b = max(b, dpp64<0x114, 0xF>(b));  // row_shr:4
	b = max(b, dpp64<0x118, 0xF>(b));  // row_shr:8
	b = max(b, dpp64<0x142, 0xA>(b));  // row_bcast:15
	b = max(b, dpp64<0x143, 0xC>(b));  // row_bcast:31
	const int lo = __builtin_amdgcn_readlane((int)(unsigned)(unsigned long long)b, 63);
	const int hi = __builtin_amdgcn_readlane((int)(unsigned)((unsigned long long)b >> 32), 63);
	return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ inline int wave_max(int v) {
	for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o));
	return v;
}

// Largest particle count of a cell: one atomic per wave.
__global__ void k_max_cell_count(Grid g, int nspecies, const int* __restrict__ start, int* __restrict__ out) {
	long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	int n = 0;
	if (c < g.ncells)
		for (int s = 0; s < nspecies; ++s) {
			const long long k0 = ((long long)s * g.ncells + c) << 3;
			n += start[k0 + 8] - start[k0];
		}
	n = wave_max(n);
	if ((threadIdx.x & 63) == 0 && n > 0) atomicMax(out, n);
}

// Largest |v| of n values into *out (the bits of a non-negative double): a grid-stride loop,
// then one atomic per wave.
__global__ void k_abs_max(const double* __restrict__ v, long long n, unsigned long long* __restrict__ out) {
	double m = 0;
	for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
		m = fmax(m, fabs(v[i]));
	m = wave_max(m);
	if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// Per species the largest |u|^2 of its particles: the wave's maximum with one atomic when all
// its lanes hold one species (lanes without a particle pass species -1), else one atomic per
// lane. The atomics go to one of kU2Slots rows picked by the workgroup (a single address taking
// one atomic from every wave of a 268 M particle launch serialises them: 48 ms instead of 17 ms
// for the push), folded by k_fold_u2_slots.
__device__ inline void species_u2_max(int s, double u2, unsigned long long* slots) {
	unsigned long long* row = slots + (blockIdx.x % kU2Slots) * kMaxSpecies;
	const int s0 = __builtin_amdgcn_readfirstlane(s);
	if (__all(s == s0)) {
		const long long m = wave_max_bits(__double_as_longlong(u2));
		if ((threadIdx.x & 63) == 0 && s0 >= 0) atomicMax(row + s0, (unsigned long long)m);
	} else if (s >= 0) {
		atomicMax(row + s, (unsigned long long)__double_as_longlong(u2));
	}
}

// The same for particles added after a push (injected: species from the sort key; replacements
// in escaped slots: species from the ranges below species_start[nspecies]), idx[0 .. count) or
// first .. first + count - 1.
__global__ void k_species_u2_added(SpeciesTable sp, ParticleArrays P, const int* __restrict__ species_start,
                                   const uint32_t* __restrict__ keys, long long ncells, const uint32_t* __restrict__ idx,
                                   long long first, long long count, unsigned long long* __restrict__ out) {
	const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	int s = -1;
	double u2 = 0;
	if (t < count) {
		const long long i = idx ? (long long)idx[t] : first + t;
		if (i >= species_start[sp.count]) {
			s = (int)((long long)(keys[i] >> 3) / ncells);
		} else {
			s = 0;
			while (s + 1 < sp.count && i >= species_start[s + 1]) ++s;
		}
		const double p0 = P.p[0][i], p1 = P.p[1][i], p2 = P.p[2][i];
		u2 = (p0 * p0 + p1 * p1 + p2 * p2) / (sp.mass[s] * sp.mass[s]);
	}
	species_u2_max(s, u2, out);
}

// The slots' maxima per species into out[kMaxSpecies]: one workgroup of kU2Slots threads.
__global__ void k_fold_u2_slots(const unsigned long long* __restrict__ slots, unsigned long long* __restrict__ out) {
	__shared__ unsigned long long r[kU2Slots];
	for (int s = 0; s < kMaxSpecies; ++s) {
		r[threadIdx.x] = slots[threadIdx.x * kMaxSpecies + s];
		__syncthreads();
		for (int h = kU2Slots / 2; h > 0; h >>= 1) {
			if ((int)threadIdx.x < h) r[threadIdx.x] = max(r[threadIdx.x], r[threadIdx.x + h]);
			__syncthreads();
		}
		if (threadIdx.x == 0) out[s] = r[0];
		__syncthreads();
	}
}

__global__ void k_species_u2(SpeciesTable sp, ParticleArrays P, const int* __restrict__ species_start,
                             unsigned long long* __restrict__ out) {
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	int s = -1;
	double u2 = 0;
	if (i < P.count) {
		s = 0;
		while (s + 1 < sp.count && i >= species_start[s + 1]) ++s;
		const double p0 = P.p[0][i], p1 = P.p[1][i], p2 = P.p[2][i];
		u2 = (p0 * p0 + p1 * p1 + p2 * p2) / (sp.mass[s] * sp.mass[s]);
	}
	species_u2_max(s, u2, out);
}

// The workgroup's current box in LDS. A workgroup's 256 particles are consecutive in the
// sorted order; when they start in one (i, j) row of cells k0 .. k0 + K - 1 (K <= kJBoxK), both
// parts of every zig-zag path deposit within one cell of that row: Jx on cells i0-1 .. i0+1,
// nodes j0-1 .. j0+2, k0-1 .. k0+K+1 (Jy, Jz alike with their own staggering). The
// workgroup sums into LDS (64-bit integer atomics) and adds the box to the global sums once;
// any other contribution goes to the global sums directly. All sums are integers, so the
// result does not depend on the path.
constexpr int kJBoxK = 20, kJBoxL = kJBoxK + 3;  // points along k per box row
constexpr int kJBoxX = 3 * 4 * kJBoxL, kJBoxY = 4 * 3 * kJBoxL, kJBoxZ = 4 * 4 * kJBoxL;

struct JBox {
	bool use;
	int o[3];    // ghosted index of local (0, 0, 0): (i0 - 1, j0 - 1, k0 - 1)
	int kext;    // K + 3 (points along k in use)
};

__device__ inline void jbox_extent(int comp, int e[2]) {  // extents along x, y
	e[0] = comp == 0 ? 3 : 4;
	e[1] = comp == 1 ? 3 : 4;
}

// Per-launch constants of the Tristan push, formed once on the host: reciprocals of the cell
// sizes and the species' factors.
struct TristanConsts {
	double inv_d[3], inv_2d[3];
	double hs[kMaxSpecies];      // q dt / 2
	double beta[kMaxSpecies];    // q dt / (2 m)
	double inv_m2[kMaxSpecies];  // 1 / m^2
	double fcd[kMaxSpecies];     // q w / V
	double fcd_dt[kMaxSpecies];  // q w / (V dt)
	double scale;                // fixed-point scale of the current, a power of two
	double u2_limit[kMaxSpecies];  // the new |u|^2 the scale allows
};

// v * scale rounded to the nearest integer, ties to even (llrint). The scale is a power of two,
// so the product is exact, and it keeps every contribution below 2^50 in magnitude: adding
// 1.5 * 2^52 leaves the rounded integer in the low bits of the sum.
__device__ inline long long to_fixed(double v, double scale) {
	return __double_as_longlong(v * scale + 6755399441055744.0) - 0x4338000000000000LL;
}

// The 12 fixed-point values of one part of a zig-zag path whose stencil starts at cell
// (a, b, c): Jx at (a, b..b+1, c..c+1), Jy at (a..a+1, b, c..c+1), Jz at (a..a+1, b..b+1, c)
// with the bilinear weights of the part's midpoint (addParticleFluxZigzag,
// momentum_collecting.cpp:176-245).
__device__ inline void quantize_part(const double F[3], const double W[3], double scale, long long q[12]) {
	const double om[3] = {1.0 - W[0], 1.0 - W[1], 1.0 - W[2]};
	q[0] = to_fixed(F[0] * om[1] * om[2], scale);
	q[1] = to_fixed(F[0] * W[1] * om[2], scale);
	q[2] = to_fixed(F[0] * om[1] * W[2], scale);
	q[3] = to_fixed(F[0] * W[1] * W[2], scale);
	q[4] = to_fixed(F[1] * om[0] * om[2], scale);
	q[5] = to_fixed(F[1] * W[0] * om[2], scale);
	q[6] = to_fixed(F[1] * om[0] * W[2], scale);
	q[7] = to_fixed(F[1] * W[0] * W[2], scale);
	q[8] = to_fixed(F[2] * om[0] * om[1], scale);
	q[9] = to_fixed(F[2] * om[0] * W[1], scale);
	q[10] = to_fixed(F[2] * W[0] * om[1], scale);
	q[11] = to_fixed(F[2] * W[0] * W[1], scale);
}

__device__ inline void add64(long long* p, long long q) { atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)q); }

// Adds one part's 12 values: into the workgroup's box when the whole 2 x 2 stencil of every
// component lies in it, else into the global sums.
__device__ inline void add_part(const YeeGrid& y, long long* box, const JBox& jb, int a, int b, int c, const long long q[12]) {
	const int la = a - jb.o[0], lb = b - jb.o[1], lc = c - jb.o[2];
	if (jb.use && (unsigned)la <= 2u && (unsigned)lb <= 2u && lc >= 0 && lc + 1 < jb.kext) {
		constexpr int L = kJBoxL;
		long long* bx = box + (la * 4 + lb) * L + lc;                  // Jx: 3 x 4 x L
		long long* by = box + kJBoxX + (la * 3 + lb) * L + lc;         // Jy: 4 x 3 x L
		long long* bz = box + kJBoxX + kJBoxY + (la * 4 + lb) * L + lc;  // Jz: 4 x 4 x L
		add64(bx, q[0]);
		add64(bx + L, q[1]);
		add64(bx + 1, q[2]);
		add64(bx + L + 1, q[3]);
		add64(by, q[4]);
		add64(by + 3 * L, q[5]);
		add64(by + 1, q[6]);
		add64(by + 3 * L + 1, q[7]);
		add64(bz, q[8]);
		add64(bz + L, q[9]);
		add64(bz + 4 * L, q[10]);
		add64(bz + 5 * L, q[11]);
		return;
	}
	long long* Jx = y.Jacc[0];
	long long* Jy = y.Jacc[1];
	long long* Jz = y.Jacc[2];
	add64(Jx + yi(y.dE[0], a, b, c), q[0]);
	add64(Jx + yi(y.dE[0], a, b + 1, c), q[1]);
	add64(Jx + yi(y.dE[0], a, b, c + 1), q[2]);
	add64(Jx + yi(y.dE[0], a, b + 1, c + 1), q[3]);
	add64(Jy + yi(y.dE[1], a, b, c), q[4]);
	add64(Jy + yi(y.dE[1], a + 1, b, c), q[5]);
	add64(Jy + yi(y.dE[1], a, b, c + 1), q[6]);
	add64(Jy + yi(y.dE[1], a + 1, b, c + 1), q[7]);
	add64(Jz + yi(y.dE[2], a, b, c), q[8]);
	add64(Jz + yi(y.dE[2], a, b + 1, c), q[9]);
	add64(Jz + yi(y.dE[2], a + 1, b, c), q[10]);
	add64(Jz + yi(y.dE[2], a + 1, b + 1, c), q[11]);
}

// What a particle leaves for the wave's segmented sum: the first part of its path (with the
// second part added when both parts share the cell), or cell = -1 for nothing.
struct PartSum {
	int cell;  // linear ghosted index of the stencil's cell
	int a, b, c;
	long long q[12];
	bool moved;  // the particle's sort key changed
	int species;  // -1 for no particle
	double u2;    // |p|^2 / m^2 after the move
	bool bounced;  // reflected at the wall in this move
};

// The particles' wall of the Tristan mover with SUPER_CONDUCTOR_LEFT: x = xgrid[reflectingWallPoint]
// (PIC++/constants.h:64), cell 20 (the fields' conductor stays at node 2).
constexpr int kWallCell = 20;

template <int CTRL, int ROWS>
__device__ inline void seg_step(long long q[12], bool take) {
#pragma unroll
	for (int v = 0; v < 12; ++v) {
		const long long t = dpp64<CTRL, ROWS>(q[v]);
		q[v] += take ? t : 0;
	}
}

// Inclusive sums over runs of lanes with the same cell (the particles are sorted by cell, so
// a wave holds a few runs): after it the last lane of every run holds the run's total.
// Steps of 1, 2, 4, 8 lanes inside rows of 16 (row_shr), then the last lane of row 0 and 2
// into rows 1 and 3 (row_bcast:15) and the last lane of row 1 into rows 2 and 3
// (row_bcast:31), each taken only when the source lane lies in the same run. Integer sums, so
// the result is the same as any other order. Needs all 64 lanes of the wave active.
__device__ inline void wave_run_sums(long long q[12], int lane, int dist) {
	seg_step<0x111, 0xF>(q, dist >= 1);
	seg_step<0x112, 0xF>(q, dist >= 2);
	seg_step<0x114, 0xF>(q, dist >= 4);
	seg_step<0x118, 0xF>(q, dist >= 8);
	seg_step<0x142, 0xA>(q, dist > (lane & 15));
	seg_step<0x143, 0xC>(q, dist > lane - 32);
}

template <bool WALL, bool FLAT>
__device__ void tristan_particle(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P, const int* __restrict__ species_start,
                                 const YeeGrid& y, double dt, const TristanConsts& K, uint32_t* __restrict__ keys,
                                 int* __restrict__ status, int* __restrict__ dead, uint32_t* __restrict__ esc, long long i,
                                 long long* box, const JBox& jb, PartSum& ps);

// FLAT: one cell along z (2-D XY grids), a template parameter so that the 3-D stencils keep
// constant offsets
template <bool WALL, bool FLAT>
__global__ void __launch_bounds__(256) k_tristan_push(Grid g, SpeciesTable sp, ParticleArrays P, const int* __restrict__ species_start,
                                                      YeeGrid y, double dt, TristanConsts K, uint32_t* __restrict__ keys,
                                                      uint64_t* __restrict__ mask, int* __restrict__ status,
                                                      unsigned long long* __restrict__ u2_slots, int* __restrict__ dead,
                                                      uint32_t* __restrict__ esc, uint64_t* __restrict__ bounce_mask) {
	__shared__ long long box[kJBoxX + kJBoxY + kJBoxZ];
	__shared__ int row_ends[2][3];
	const long long b0 = (long long)blockIdx.x * blockDim.x;
	const long long idx = b0 + threadIdx.x;
	const int last = (int)min((long long)blockDim.x - 1, P.count - 1 - b0);
	if ((int)threadIdx.x == 0 || (int)threadIdx.x == last) {
		const int w = (int)threadIdx.x == 0 ? 0 : 1;
		for (int a = 0; a < 3; ++a) row_ends[w][a] = (int)floor((P.x[a][idx] - g.x0[a]) / g.d[a]);
	}
	for (int q = threadIdx.x; q < kJBoxX + kJBoxY + kJBoxZ; q += blockDim.x) box[q] = 0;
	__syncthreads();
	JBox jb;
	{
		const int K = row_ends[1][2] - row_ends[0][2] + 1;
		jb.use = !FLAT && row_ends[0][0] == row_ends[1][0] && row_ends[0][1] == row_ends[1][1] && K >= 1 && K <= kJBoxK;
		for (int a = 0; a < 3; ++a) jb.o[a] = row_ends[0][a] - 1;
		jb.kext = K + 3;
	}
	PartSum ps;
	ps.cell = -1;
	ps.moved = false;
	ps.species = -1;
	ps.u2 = 0;
	ps.bounced = false;
	tristan_particle<WALL, FLAT>(g, sp, P, species_start, y, dt, K, keys, status, dead, esc, idx, box, jb, ps);
	species_u2_max(ps.species, ps.u2, u2_slots);
	if (WALL) {  // one bit per particle reflected at the wall
		const unsigned long long rb = __ballot(ps.bounced);
		if ((threadIdx.x & 63) == 0 && idx < P.count) bounce_mask[idx >> 6] = rb;
	}
	{
		const int lane = (int)(threadIdx.x & 63);
		// one mover bit per particle: bit i & 63 of word i >> 6 (waves are 64 consecutive
		// particles; lanes past the end contribute zeros)
		const unsigned long long bits = __ballot(ps.moved);
		if (lane == 0 && idx < P.count) mask[idx >> 6] = bits;
		const int prev = __shfl_up(ps.cell, 1);
		const unsigned long long heads = __ballot(lane == 0 || prev != ps.cell);
		const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1;
		const int head = 63 - __clzll(heads & upto);
		wave_run_sums(ps.q, lane, lane - head);
		const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1);
		if (tail && ps.cell >= 0) add_part(y, box, jb, ps.a, ps.b, ps.c, ps.q);
	}
	__syncthreads();
	if (!jb.use) return;
	for (int q = threadIdx.x; q < kJBoxX + kJBoxY + kJBoxZ; q += blockDim.x) {
		const long long v = box[q];
		if (v == 0) continue;
		const int comp = q < kJBoxX ? 0 : (q < kJBoxX + kJBoxY ? 1 : 2);
		const int r = q - (comp == 0 ? 0 : (comp == 1 ? kJBoxX : kJBoxX + kJBoxY));
		int e[2];
		jbox_extent(comp, e);
		const int lc = r % kJBoxL, lab = r / kJBoxL, lb = lab % e[1], la = lab / e[1];
		atomicAdd(reinterpret_cast<unsigned long long*>(y.Jacc[comp] + yi(y.dE[comp], jb.o[0] + la, jb.o[1] + lb, jb.o[2] + lc)),
		          (unsigned long long)v);
	}
}

// One particle of k_tristan_push (none when i >= P.count): the second part of a path that
// changes cell goes to the sums here; the first part, with the second when both share the
// cell, is left in ps for the wave's run sums. Divisions by the cell sizes, by dt and by m^2 are
// products with the host's reciprocals.
template <bool WALL, bool FLAT>
__device__ __forceinline__ void tristan_particle(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P,
                                                 const int* __restrict__ species_start, const YeeGrid& y, double dt,
                                                 const TristanConsts& K, uint32_t* __restrict__ keys, int* __restrict__ status,
                                                 int* __restrict__ dead, uint32_t* __restrict__ esc, long long i, long long* box,
                                                 const JBox& jb, PartSum& ps) {
	for (int v = 0; v < 12; ++v) ps.q[v] = 0;
	if (i >= P.count) return;
	int s = 0;
	while (s + 1 < sp.count && i >= species_start[s + 1]) ++s;
	const double m = sp.mass[s];
	double x[3], p[3];
	for (int a = 0; a < 3; ++a) {
		x[a] = P.x[a][i];
		p[a] = P.p[a][i];
	}
	const uint32_t key0 = position_key(g, s, x);
	// correlationBunemanEBfieldsWithoutMaps: the lower cell and node of the stencil per axis,
	// with the weights of the lower and the upper point
	int lc[3], ln[3];
	double wc[3][2], wn[3][2];
	bool ok = true;
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		const double g0 = g.x0[a], d = g.d[a], id = K.inv_d[a];
		int cnt = (int)floor((x[a] - g0) * id);
		const double midc = ((g0 + cnt * d) + (g0 + (cnt + 1) * d)) / 2;
		if (x[a] < midc) {
			lc[a] = cnt - 1;
			wc[a][1] = 0.5 + (x[a] - (g0 + cnt * d)) * id;
			wc[a][0] = 1.0 - wc[a][1];
		} else {
			lc[a] = cnt;
			wc[a][0] = 0.5 + ((g0 + (cnt + 1) * d) - x[a]) * id;
			wc[a][1] = 1.0 - wc[a][0];
		}
		cnt = (int)floor(((x[a] - g0) * id) + 0.5);
		if (x[a] < g0 + cnt * d) {
			ln[a] = cnt - 1;
			wn[a][1] = 0.5 + (x[a] - ((g0 + (cnt - 1) * d) + (g0 + cnt * d)) / 2) * id;
			wn[a][0] = 1.0 - wn[a][1];
		} else {
			ln[a] = cnt;
			wn[a][0] = 0.5 + (((g0 + cnt * d) + (g0 + (cnt + 1) * d)) / 2 - x[a]) * id;
			wn[a][1] = 1.0 - wn[a][0];
		}
		if (FLAT && a == 2) {  // one plane along z: weight 1 at k = 0 (the TWO_D_XY sums, correlation.cpp:627-661)
			lc[a] = ln[a] = 0;
			wc[a][0] = wn[a][0] = 1.0;
			wc[a][1] = wn[a][1] = 0.0;
			continue;
		}
		// stencils inside the arrays: cells 0 .. N-1, nodes 0 .. N
		if (lc[a] < 0 || lc[a] + 1 > g.N[a] - 1 || ln[a] < 0 || ln[a] + 1 > g.N[a]) ok = false;
	}
	constexpr int zs = FLAT ? 0 : 1;  // the block's second z plane (plane 0 again, at weight 0, when collapsed)
	if (!ok) {  // no access outside the arrays: report and leave the particle
		atomicOr(status, 8);
		return;
	}
	// E component c sits on cells along c and nodes along the others, B the opposite: each
	// takes a 2 x 2 x 2 block from its lower corner (indices below 2^31, checked on the host)
	double Ef[3], Bf[3];
#pragma unroll
	for (int c = 0; c < 3; ++c) {
		int e0[3], b0[3];
		double we[3][2], wb[3][2];
#pragma unroll
		for (int a = 0; a < 3; ++a) {
			e0[a] = a == c ? lc[a] : ln[a];
			b0[a] = a == c ? ln[a] : lc[a];
			for (int t = 0; t < 2; ++t) {
				we[a][t] = a == c ? wc[a][t] : wn[a][t];
				wb[a][t] = a == c ? wn[a][t] : wc[a][t];
			}
		}
		const int* dE = y.dE[c];
		const int* dB = y.dB[c];
		const double* pe = y.E[c] + ((e0[0] * dE[1] + e0[1]) * dE[2] + e0[2]);
		const double* pb = y.B[c] + ((b0[0] * dB[1] + b0[1]) * dB[2] + b0[2]);
		const int se = dE[1] * dE[2], sb = dB[1] * dB[2];
		double e = 0, b = 0;
#pragma unroll
		for (int sx = 0; sx < 2; ++sx)
#pragma unroll
			for (int sy = 0; sy < 2; ++sy)
#pragma unroll
				for (int sz = 0; sz < 2; ++sz) {
					e += pe[sx * se + sy * dE[2] + sz * zs] * we[0][sx] * we[1][sy] * we[2][sz];
					b += pb[sx * sb + sy * dB[2] + sz * zs] * wb[0][sx] * wb[1][sy] * wb[2][sz];
				}
		Ef[c] = e;
		Bf[c] = b;
	}
	// moveParticleTristan: half kick, rotation, half kick
	const double hs = K.hs[s];
	double dp[3];
	for (int a = 0; a < 3; ++a) {
		dp[a] = Ef[a] * hs;
		p[a] = p[a] + dp[a];
	}
	const double gamma = sqrt((p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) * K.inv_m2[s] + 1.0);
	const double bs = K.beta[s] / gamma;
	const double f = 2.0 / sqrt(1.0 + bs * bs * (Bf[0] * Bf[0] + Bf[1] * Bf[1] + Bf[2] * Bf[2]));
	double c1[3] = {p[1] * Bf[2] - p[2] * Bf[1], -p[0] * Bf[2] + p[2] * Bf[0], p[0] * Bf[1] - p[1] * Bf[0]};
	double tm[3];
	for (int a = 0; a < 3; ++a) tm[a] = p[a] + c1[a] * bs;
	double c2[3] = {tm[1] * Bf[2] - tm[2] * Bf[1], -tm[0] * Bf[2] + tm[2] * Bf[0], tm[0] * Bf[1] - tm[1] * Bf[0]};
	const double fb = f * bs;
	for (int a = 0; a < 3; ++a) p[a] += c2[a] * fb + dp[a];
	const double u2 = (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) * K.inv_m2[s];
	ps.species = s;
	ps.u2 = u2;
	if (u2 > K.u2_limit[s]) atomicOr(status, 16);  // beyond what the current's scale was set for
	const double gn = sqrt(u2 + 1.0);
	const double iv = 1.0 / (m * gn);
	double v[3];
	for (int a = 0; a < 3; ++a) {
		v[a] = p[a] * iv;
		x[a] += v[a] * dt;
	}
	const double xw = g.x0[0] + kWallCell * g.d[0];  // xgrid[reflectingWallPoint]
	if (WALL && x[0] < xw) {  // reflected at the wall (mover.cpp:746-750, reflectMomentumX)
		x[0] = 2 * xw - x[0];
		p[0] = -p[0];
		v[0] = -v[0];
		ps.bounced = true;
	}
	// addParticleFluxZigzag at the unwrapped position
	double loc[3], prev[3], r[3];
	int cur[3], pre[3];
	for (int a = 0; a < 3; ++a) {
		loc[a] = x[a] - g.x0[a];
		cur[a] = (int)floor(loc[a] * K.inv_d[a]);
		prev[a] = loc[a] - v[a] * dt;
	}
	// the wall cell's path of a reflected particle starts from the mirror of its previous
	// position (momentum_collecting.cpp:93-99)
	if (WALL && cur[0] == kWallCell && v[0] > 0 && prev[0] < kWallCell * g.d[0]) prev[0] = 2 * kWallCell * g.d[0] - prev[0];
	for (int a = 0; a < 3; ++a) {
		pre[a] = (int)floor(prev[a] * K.inv_d[a]);
		r[a] = (loc[a] + prev[a]) / 2;
		if (FLAT && a == 2) {  // the TWO_D_XY path (momentum_collecting.cpp:235-277): the z midpoint, plane 0
			cur[a] = pre[a] = 0;
			continue;
		}
		if (cur[a] != pre[a]) r[a] = fmax(g.x0[a] + pre[a] * g.d[a], g.x0[a] + cur[a] * g.d[a]) - g.x0[a];
		// J stencils of both parts inside J's arrays (cells 0 .. N-2 along the component's
		// own axis, nodes 0 .. N-1 along the others)
		if (cur[a] < 0 || pre[a] < 0 || cur[a] > g.N[a] - 2 || pre[a] > g.N[a] - 2) ok = false;
	}
	if (!ok) {
		atomicOr(status, 8);
		return;
	}
	const double fcd = K.fcd[s], fdt = K.fcd_dt[s];
	double F1[3], F2[3], W1[3], W2[3];
	for (int a = 0; a < 3; ++a) F1[a] = fdt * (r[a] - prev[a]);
	F2[0] = fdt * (loc[0] - r[0]);
	F2[1] = fcd * v[1] - F1[1];
	F2[2] = fcd * v[2] - F1[2];
	for (int a = 0; a < 3; ++a) {
		W1[a] = (prev[a] + r[a]) * K.inv_2d[a] - pre[a];
		W2[a] = (loc[a] + r[a]) * K.inv_2d[a] - cur[a];
	}
	if (FLAT) W1[2] = W2[2] = 0.0;  // Jx, Jy without a z factor: all on plane 0
	quantize_part(F1, W1, K.scale, ps.q);
	{
		long long q2[12];
		quantize_part(F2, W2, K.scale, q2);
		if (cur[0] == pre[0] && cur[1] == pre[1] && cur[2] == pre[2]) {
			for (int v = 0; v < 12; ++v) ps.q[v] += q2[v];
		} else {
			add_part(y, box, jb, cur[0], cur[1], cur[2], q2);
		}
	}
	ps.a = pre[0];
	ps.b = pre[1];
	ps.c = pre[2];
	ps.cell = (int)(((long long)pre[0] * g.N[1] + pre[1]) * g.N[2] + pre[2]);
	if (WALL && x[0] > g.x0[0] + (g.N[0] - 2) * g.d[0]) {
		// out through the free right end (mover.cpp:760-767): its current is summed above, as the
		// reference's flux loop still holds it; listed for removal or replacement (y, z unwrapped)
		for (int a = 0; a < 3; ++a) {
			P.x[a][i] = x[a];
			P.p[a][i] = p[a];
		}
		keys[i] = 0xFFFFFFFFu;
		esc[atomicAdd(dead, 1)] = (uint32_t)i;
		ps.moved = true;
		return;
	}
	// one-rank periodic wrap (mover.cpp:770-806 classification, exchangeParticles' shift); x too
	// unless walled
	for (int a = WALL ? 1 : 0; a < (FLAT ? 2 : 3); ++a) {  // z unwrapped when collapsed (mover.cpp:794)
		const double lo = g.x0[a] + 2 * g.d[a], hi = g.x0[a] + (g.N[a] - 2) * g.d[a];
		if (x[a] < lo)
			x[a] += g.L[a];
		else if (x[a] > hi)
			x[a] -= g.L[a];
		if (!(x[a] >= lo && x[a] <= hi)) atomicOr(status, 1);  // moved more than a box length
	}
	for (int a = 0; a < 3; ++a) {
		P.x[a][i] = x[a];
		P.p[a][i] = p[a];
	}
	const uint32_t key = position_key(g, s, x);
	keys[i] = key;
	ps.moved = key != key0;
}

}  // namespace

void yee_dims(const Grid& g, bool b_field, int c, int dims[3]) {
	for (int a = 0; a < 3; ++a) dims[a] = ((a == c) != b_field) ? g.N[a] : g.N[a] + 1;
}

void launch_tristan_b_half(const Grid& g, const YeeGrid& y, double dt, hipStream_t s) {
	const long long n = (long long)(g.N[0] - 3) * (g.N[1] - 3) * (g.cz ? 1 : g.N[2] - 3);
	hipLaunchKernelGGL(k_b_half, dim3(nb(n, 256)), dim3(256), 0, s, g, y, kCprimeX * dt / 2.0);
}

void launch_tristan_e(const Grid& g, const YeeGrid& y, double dt, hipStream_t s) {
	const long long n = (long long)(g.N[0] - 3) * (g.N[1] - 3) * (g.cz ? 1 : g.N[2] - 3);
	const double pi = 4 * atan2(1.0, 1.0);  // PIC++/constants.h:28-29
	hipLaunchKernelGGL(k_e, dim3(nb(n, 256)), dim3(256), 0, s, g, y, dt, 4 * pi, g.e_right[1], g.e_right[2]);
}

// The exchanges (or, with fold, the current's folds) along the three axes. With the wall, x
// has no folds and its ghost planes follow the wall's rule (zero_mask: the components zeroed).
static void yee_axes(const Grid& g, Yee3& q, bool fold, hipStream_t s, int zero_mask = 0) {
	for (int ax = 0; ax < 3; ++ax) {
		const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
		long long lines = 0;
		for (int c = 0; c < 3; ++c) lines += (long long)q.d[c][o1] * q.d[c][o2];
		if (g.wall && ax == 0) {
			if (!fold) hipLaunchKernelGGL(k_yee_wall_x, dim3(nb(lines, 256)), dim3(256), 0, s, q, zero_mask);
		} else if (g.cz && ax == 2)
			hipLaunchKernelGGL(k_yee_copy_z, dim3(nb(lines, 256)), dim3(256), 0, s, q);
		else if (fold)
			hipLaunchKernelGGL(k_yee_fold, dim3(nb(lines, 256)), dim3(256), 0, s, q, ax, g.N[ax]);
		else
			hipLaunchKernelGGL(k_yee_exchange, dim3(nb(lines, 256)), dim3(256), 0, s, q, ax, g.N[ax]);
	}
}

void launch_yee_exchange(const Grid& g, const YeeGrid& y, bool b_field, hipStream_t s) {
	Yee3 q;
	for (int c = 0; c < 3; ++c) {
		q.f[c] = b_field ? y.B[c] : y.E[c];
		for (int a = 0; a < 3; ++a) q.d[c][a] = b_field ? y.dB[c][a] : y.dE[c][a];
	}
	yee_axes(g, q, false, s, b_field ? 6 : 1);  // the wall zeroes By, Bz and Ex behind it
}

// smoothBunemanEfieldGeneral (PIC++/smoothing.cpp:535-733) on current component c: the 27-point
// kernel of createSmoothingMatrix over the component's range (along its own axis [2, N - 2),
// along the others [2, N - 2]), read from src and written to dst; with the wall the plane i = 2
// (and i = N0 - 2 for the y and z components) is copied instead (the reference's non-periodic
// branches). The sum runs over the 27 neighbours in the reference's order; along a collapsed z
// over the 9 of the plane k = 0 (kwidth = 0, smoothing.cpp:546-561).
__global__ void k_smooth_yee(Grid g, const double* __restrict__ src, double* __restrict__ dst, int comp, Smooth27 sm, int wall,
                             int d1, int d2, int r1, int r2, long long n) {
	const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	const int k = owned_lo(g, 2) + (int)(t % r2), j = 2 + (int)((t / r2) % r1), i = 2 + (int)(t / ((long long)r2 * r1));
	const int cw = g.cz ? 0 : 1;
	const long long o = ((long long)i * d1 + j) * d2 + k;
	if (wall && (i == 2 || (comp != 0 && i == g.N[0] - 2))) {
		dst[o] = src[o];
		return;
	}
	double acc = 0;
	for (int a = -1; a <= 1; ++a)
		for (int b = -1; b <= 1; ++b)
			for (int c = -cw; c <= cw; ++c)
				acc += src[((long long)(i + a) * d1 + (j + b)) * d2 + (k + c)] * sm.m[(a + 1) * 9 + (b + 1) * 3 + (c + 1)];
	dst[o] = acc;
}

__global__ void k_copy_range(const double* __restrict__ src, double* __restrict__ dst, int d1, int d2, int r1, int r2,
                             long long n, int k0) {
	const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	const int k = k0 + (int)(t % r2), j = 2 + (int)((t / r2) % r1), i = 2 + (int)(t / ((long long)r2 * r1));
	const long long o = ((long long)i * d1 + j) * d2 + k;
	dst[o] = src[o];
}

void launch_yee_current(const Grid& g, const YeeGrid& y, double inv_scale, hipStream_t s) {
	Yee3 q;
	for (int c = 0; c < 3; ++c) {
		const long long n = yee_count(y.dE[c]);
		hipLaunchKernelGGL(k_fixed_to_double, dim3(nb(n, 256)), dim3(256), 0, s, y.Jacc[c], y.J[c], n, inv_scale);
		q.f[c] = y.J[c];
		for (int a = 0; a < 3; ++a) q.d[c][a] = y.dE[c][a];
	}
	yee_axes(g, q, true, s);
}

void launch_smooth_yee_current(const Grid& g, const YeeGrid& y, const Smooth27& sm, double* const tmp[3], hipStream_t s) {
	Yee3 q;
	for (int c = 0; c < 3; ++c) {
		const int* d = y.dE[c];
		int r[3];
		for (int a = 0; a < 3; ++a) r[a] = (a == 2 && g.cz) ? 1 : (a == c ? g.N[a] - 4 : g.N[a] - 3);
		const long long n = (long long)r[0] * r[1] * r[2];
		hipLaunchKernelGGL(k_smooth_yee, dim3(nb(n, 256)), dim3(256), 0, s, g, y.J[c], tmp[c], c, sm, g.wall, d[1], d[2], r[1], r[2], n);
		hipLaunchKernelGGL(k_copy_range, dim3(nb(n, 256)), dim3(256), 0, s, tmp[c], y.J[c], d[1], d[2], r[1], r[2], n, owned_lo(g, 2));
		q.f[c] = y.J[c];
		for (int a = 0; a < 3; ++a) q.d[c][a] = d[a];
	}
	yee_axes(g, q, false, s, 1);  // exchangeBunemanEfield on the current
}

void launch_species_u2_added(const SpeciesTable& sp, const ParticleArrays& P, const int* species_start, const uint32_t* keys,
                             long long ncells, const uint32_t* idx, long long first, long long count, unsigned long long* u2_slots,
                             hipStream_t s) {
	if (count <= 0) return;
	hipLaunchKernelGGL(k_species_u2_added, dim3(nb(count, 256)), dim3(256), 0, s, sp, P, species_start, keys, ncells, idx, first,
	                   count, u2_slots);
}

void launch_explicit_bounds(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P, const int* species_start,
                            const int* start, const YeeGrid& y, bool particles, unsigned long long* u2_slots,
                            ExplicitBounds* out, hipStream_t s) {
	hipLaunchKernelGGL(k_max_cell_count, dim3(nb(g.ncells, 256)), dim3(256), 0, s, g, sp.count, start, &out->maxcount);
	for (int c = 0; c < 3; ++c) {
		const long long n = yee_count(y.dE[c]);
		hipLaunchKernelGGL(k_abs_max, dim3(std::min(nb(n, 256), 1024u)), dim3(256), 0, s, y.E[c], n, out->emax + c);
	}
	if (particles && P.count > 0)
		hipLaunchKernelGGL(k_species_u2, dim3(nb(P.count, 256)), dim3(256), 0, s, sp, P, species_start, u2_slots);
	hipLaunchKernelGGL(k_fold_u2_slots, dim3(1), dim3(kU2Slots), 0, s, u2_slots, out->u2max);
}

void launch_tristan_push(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P, const int* species_start,
                         const YeeGrid& y, double dt, double scale, const double* u2_limit, uint32_t* keys, uint64_t* mask,
                         int* status, unsigned long long* u2_slots, int* dead, uint32_t* esc, uint64_t* bounce_mask,
                         hipStream_t s) {
	if (P.count == 0) return;
	TristanConsts K{};
	for (int a = 0; a < 3; ++a) {
		K.inv_d[a] = 1.0 / g.d[a];
		K.inv_2d[a] = 1.0 / (2 * g.d[a]);
	}
	for (int q = 0; q < sp.count; ++q) {
		const double m = sp.mass[q], c = sp.charge[q];
		K.hs[q] = c * dt * 0.5;
		K.beta[q] = c * dt / (2.0 * m);
		K.inv_m2[q] = 1.0 / (m * m);
		K.fcd[q] = c * sp.weight[q] / g.volume;
		K.fcd_dt[q] = K.fcd[q] / dt;
	}
	K.scale = scale;
	for (int q = 0; q < kMaxSpecies; ++q) K.u2_limit[q] = q < sp.count ? u2_limit[q] : 0.0;
	auto kern = g.wall ? (g.cz ? k_tristan_push<true, true> : k_tristan_push<true, false>)
	                   : (g.cz ? k_tristan_push<false, true> : k_tristan_push<false, false>);
	hipLaunchKernelGGL(kern, dim3(nb(P.count, 256)), dim3(256), 0, s, g, sp, P, species_start, y, dt, K, keys, mask, status, u2_slots,
	                   dead, esc, bounce_mask);
}

}  // namespace picgpu
