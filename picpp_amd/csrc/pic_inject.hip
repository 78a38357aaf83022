// Injection at the free right end and replacement of escaped particles
// (PIC++/mover.cpp:910-1000, PIC++/initialization.cpp:8281-8338, :8479-8570).
//
// The reference draws from the C library's rand(); here every draw comes from a
// counter-based generator, Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC 2011), keyed by the run's seed, so that a particle's
// numbers depend only on (step, slab or escape rank, column) and the result does not depend
// on the launch shape. Counters:
//   (step, 2 << 32 | slab, 0, 0)     the slab's y, z offsets (tempDeltaY, tempDeltaZ)
//   (step, 0 << 32 | slab, col, 0)   the thermal momentum of the slab's particle in column col
//   (step, 1 << 32 | rank, 0, 0)     the thermal momentum of the rank-th replacement
//   (step, 1 << 32 | rank, 1, 0)     its y, z redraws when the escaped particle was outside
// and, for species from kT / mc^2 = 0.1 on, the fourth word numbering the Juttner sampler's
// attempts, 1 << 40 for the direction (rest_u).
// A 64-bit output w becomes the uniform (w >> 11) 2^-53 in [0, 1); normals come in pairs from
// the Box-Muller transform of two uniforms.
//
// Coordinates follow the reference's arithmetic (grid node i at x0 + i d, no fused
// multiply-add), so injected positions are exact in FP64.
#pragma clang fp contract(off)

#include "pic_inject.hpp"

#include <cmath>

namespace picgpu {

namespace {

constexpr double kTwoPi = 6.283185307179586231995926937088370323181152343750;  // 2 * 4*atan2(1,1)
constexpr double kRelativisticPrecision = 0.0001;                                // PIC++/constants.h:39

__device__ inline void philox4x64_10(uint64_t c[4], uint64_t k0, uint64_t k1) {
	const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
	const uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
	for (int r = 0; r < 10; ++r) {
		if (r > 0) {
			k0 += W0;
			k1 += W1;
		}
		const uint64_t lo0 = M0 * c[0], hi0 = __umul64hi(M0, c[0]);
		const uint64_t lo1 = M1 * c[2], hi1 = __umul64hi(M1, c[2]);
		const uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
		c[0] = n0;
		c[1] = lo1;
		c[2] = n2;
		c[3] = lo0;
	}
}

__device__ inline double uniform53(uint64_t w) { return (double)(w >> 11) * 0x1.0p-53; }

__device__ inline void box_muller(uint64_t wa, uint64_t wb, double& n0, double& n1) {
	const double r = sqrt(-2.0 * log(1.0 - uniform53(wa)));
	const double a = kTwoPi * uniform53(wb);
	n0 = r * cos(a);
	n1 = r * sin(a);
}

// |u| = |p| / (m c) from a species' tabulated Juttner CDF for the uniform y
// (dichotomySolver, PIC++/specialmath.cpp:931-947: halve the index interval [0, points - 1]
// until it is one step wide, then interpolate linearly; an exact hit returns the point).
__device__ inline double juttner_inverse(const double* tab, double y) {
	const double* u = tab;
	const double* F = tab + kJuttnerPoints;
	int lo = 0, hi = kJuttnerPoints - 1;
	while (lo + 1 != hi) {
		const int mid = (lo + hi) / 2;
		if (F[mid] > y)
			hi = mid;
		else if (F[mid] < y)
			lo = mid;
		else
			return u[mid];
	}
	return (u[hi] - u[lo]) * (y - F[lo]) / (F[hi] - F[lo]) + u[lo];
}

// The rest-frame u = p / (m c) of createParticle (PIC++/initialization.cpp:8505-8535). Below
// kT / mc^2 = 0.1 the Maxwellian u_i = sqrt(theta) N(0, 1) from w (two Box-Muller pairs).
// From 0.1 on, an isotropic direction (u_z = |u| (2 U - 1), phi = 2 pi U) and |u| from the
// species' table when the run's theta is below 2 (maxwellJuttnerMomentumColdDistribution,
// PIC++/random.cpp:43-48), else from the rejection sampler of maxwellJuttnerMomentumHotDistribution
// (:50-66: |u| = -theta log(U2 U3 U4), kept when exp((|u| - sqrt(1 + u^2)) / theta) >= U1).
// Counters: (c0, c1, c2, a) for attempt a of the rejection (a = 0 also the table's uniform),
// (c0, c1, c2, 1 << 40) for the direction.
__device__ inline void rest_u(const InjectPhysics& ph, int s, const uint64_t c[4], double u[3]) {
	const double theta = ph.theta[s];
	uint64_t w[4] = {c[0], c[1], c[2], c[3]};
	philox4x64_10(w, ph.seed, 0);
	if (theta < 0.1) {
		double n[4];
		box_muller(w[0], w[1], n[0], n[1]);
		box_muller(w[2], w[3], n[2], n[3]);
		const double st = sqrt(theta);
		u[0] = st * n[0];
		u[1] = st * n[1];
		u[2] = st * n[2];
		return;
	}
	double a;
	const double t1 = ph.theta_run[s];
	if (t1 < 2) {
		a = juttner_inverse(ph.juttner + (size_t)s * 2 * kJuttnerPoints, uniform53(w[0]));
	} else {
		// U1 is drawn once and kept for every attempt, as the reference does; the loop ends after
		// 64 attempts at most
		const double x1 = uniform53(w[0]);
		a = -t1 * log(uniform53(w[1]) * uniform53(w[2]) * uniform53(w[3]));
		for (uint64_t att = 1; att < 64 && exp((a - sqrt(1 + a * a)) / t1) < x1; ++att) {
			w[0] = c[0];
			w[1] = c[1];
			w[2] = c[2];
			w[3] = att;
			philox4x64_10(w, ph.seed, 0);
			a = -t1 * log(uniform53(w[1]) * uniform53(w[2]) * uniform53(w[3]));
		}
	}
	uint64_t d[4] = {c[0], c[1], c[2], 1ull << 40};
	philox4x64_10(d, ph.seed, 0);
	const double uz = a * (2 * uniform53(d[0]) - 1);
	const double phi = kTwoPi * uniform53(d[1]);
	const double un = sqrt(a * a - uz * uz);
	u[0] = un * cos(phi);
	u[1] = un * sin(phi);
	u[2] = uz;
}

// Momentum p = m gamma v of a particle of the flowing plasma: the rest-frame u of createParticle
// (rest_u), its velocity v' = u / gamma' (Particle::getVelocity, particle.cpp:216-232), the
// relativistic velocity addition of V0 (Particle::addVelocity, particle.cpp:291-331: along V0,
// (V + v'_par) / (1 + V v'_par); across, v'_perp / (gamma_V (1 + V v'_par))) and p from the new
// velocity (Particle::setMomentumByV, particle.cpp:337-364).
__device__ inline void flow_momentum(const InjectPhysics& ph, double m, int s, const uint64_t c[4], double p[3]) {
	double u[3];
	rest_u(ph, s, c, u);
	const double g1 = sqrt((u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) + 1.0);
	double v[3] = {u[0] / g1, u[1] / g1, u[2] / g1};
	const double V2 = ph.flow[0] * ph.flow[0] + ph.flow[1] * ph.flow[1] + ph.flow[2] * ph.flow[2];
	if (V2 > 0.0) {
		const double V = sqrt(V2);
		const double e[3] = {ph.flow[0] / V, ph.flow[1] / V, ph.flow[2] / V};
		const double gV = 1.0 / sqrt(1.0 - V2);
		const double vpar = v[0] * e[0] + v[1] * e[1] + v[2] * e[2];
		const double den = 1.0 + V * vpar;
		const double npar = (V + vpar) / den;
#pragma unroll
		for (int c = 0; c < 3; ++c) {
			const double vperp = v[c] - vpar * e[c];
			v[c] = npar * e[c] + vperp / (gV * den);
		}
	}
	const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
	const double f = v2 < kRelativisticPrecision ? m : m * (1.0 / sqrt(1.0 - v2));
#pragma unroll
	for (int c = 0; c < 3; ++c) p[c] = v[c] * f;
}

// Stores a particle at index i in either form, with its sort key.
__device__ inline void store_particle(const Grid& g, int s, const SpeciesTable& sp, const double x[3], const double p[3], int id,
                                      const ParticleArrays& P, const ParticleArrays32& Q, bool fp32, long long i,
                                      uint32_t* keys) {
	if (fp32) {
		int cc[3];
		float f[3];
#pragma unroll
		for (int a = 0; a < 3; ++a) {
			const double r = (x[a] - g.x0[a]) / g.d[a];
			const double fl = floor(r);
			int ci = (int)fl;
			float fa = (float)(r - fl);
			if (fa >= 1.0f) {
				fa = 0.0f;
				ci += 1;
			}
			if (ci == g.n[a] + 2) {
				if (a == 0 && g.wall) {  // on the right boundary node: kept in the last owned cell
					ci = g.n[a] + 1;
					fa = 0x1.fffffep-1f;
				} else {  // y, z = 2L is the periodic image of L
					ci = 2;
				}
			}
			cc[a] = ci < 2 ? 2 : (ci > g.n[a] + 1 ? g.n[a] + 1 : ci);
			f[a] = fa;
			Q.off[a][i] = fa;
			Q.u[a][i] = (float)(p[a] / sp.mass[s]);
		}
		const int cell = (int)cell_id(g, cc[0], cc[1], cc[2]);
		Q.cell[i] = cell;
		Q.id[i] = id;
		keys[i] = (uint32_t)((((long long)s * g.ncells + cell) << 3) | ((f[0] >= 0.5f) << 2) | ((f[1] >= 0.5f) << 1) |
		                     (f[2] >= 0.5f));
	} else {
#pragma unroll
		for (int a = 0; a < 3; ++a) {
			P.x[a][i] = x[a];
			P.p[a][i] = p[a];
		}
		P.id[i] = id;
		keys[i] = position_key(g, s, x);
	}
}

__global__ void k_inject(Grid g, SpeciesTable sp, InjectSlabs S, InjectPhysics ph, ParticleArrays P, ParticleArrays32 Q,
                         int fp32, long long first, uint32_t* __restrict__ keys) {
	const int ncols = g.n[1] * g.n[2];
	const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= (long long)S.count * ncols) return;
	const int sl = (int)(t / ncols), col = (int)(t % ncols);
	const int s = S.species[sl];
	const int j = 2 + col / g.n[2], k = owned_lo(g, 2) + col % g.n[2];
	uint64_t w[4] = {ph.step, (2ull << 32) | S.slab[sl], 0, 0};
	philox4x64_10(w, ph.seed, 0);
	const double dy = g.d[1] * uniform53(w[0]), dz = g.d[2] * uniform53(w[1]);
	double x[3];
	x[0] = S.x[sl];
	x[1] = (g.x0[1] + j * g.d[1]) + dy;  // ygrid[j] + tempDeltaY (PIC++/mover.cpp:976-977)
	x[2] = (g.x0[2] + k * g.d[2]) + dz;
	const uint64_t c[4] = {ph.step, (0ull << 32) | S.slab[sl], (uint64_t)col, 0};
	double p[3];
	flow_momentum(ph, sp.mass[s], s, c, p);
	store_particle(g, s, sp, x, p, ph.id_base + (int)t, P, Q, fp32 != 0, first + t, keys);
}

__global__ void k_replace(Grid g, SpeciesTable sp, InjectPhysics ph, const int* __restrict__ species_start,
                          const uint32_t* __restrict__ esc, int count, ParticleArrays P, ParticleArrays32 Q, int fp32,
                          uint32_t* __restrict__ keys) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= count) return;
	const long long i = esc[r];
	int s = 0;
	while (s + 1 < sp.count && i >= species_start[s + 1]) ++s;
	double x[3];
	x[0] = g.x0[0] + (g.N[0] - 2) * g.d[0];  // xgrid[xnumberAdded - 1 - additionalBinNumber]
	if (fp32) {
		x[1] = g.x0[1] + (double)Q.off[1][i] * g.d[1];
		x[2] = g.x0[2] + (double)Q.off[2][i] * g.d[2];
	} else {
		x[1] = P.x[1][i];
		x[2] = P.x[2][i];
	}
	uint64_t wy[4] = {ph.step, (1ull << 32) | (uint64_t)r, 1, 0};
	philox4x64_10(wy, ph.seed, 0);
	// y, z of the escaped particle; outside the box, redrawn inside its first or last cell
	// (PIC++/initialization.cpp:8294-8310); a collapsed z is kept as it is
#pragma unroll
	for (int a = 1; a < (g.cz ? 2 : 3); ++a) {
		const double lo = g.x0[a] + 2 * g.d[a], hi = g.x0[a] + (g.N[a] - 2) * g.d[a];
		if (x[a] < lo) x[a] = lo + g.d[a] * uniform53(wy[2 * (a - 1)]);
		if (x[a] > hi) x[a] = hi - g.d[a] * uniform53(wy[2 * (a - 1) + 1]);
	}
	const uint64_t c[4] = {ph.step, (1ull << 32) | (uint64_t)r, 0, 0};
	double p[3];
	flow_momentum(ph, sp.mass[s], s, c, p);
	store_particle(g, s, sp, x, p, ph.id_base + r, P, Q, fp32 != 0, i, keys);
}

__global__ void k_philox(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3, uint64_t k0, uint64_t k1, long long n,
                         uint64_t* __restrict__ out) {
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	uint64_t c[4] = {c0, c1 + (uint64_t)i, c2, c3};
	philox4x64_10(c, k0, k1);
#pragma unroll
	for (int w = 0; w < 4; ++w) out[4 * i + w] = c[w];
}

inline unsigned nblocks(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

void juttner_table(double theta, double* out) {
	double* u = out;
	double* F = out + kJuttnerPoints;
	const double umax = 20 * theta;
	const double du = umax / (kJuttnerPoints - 1);
	const int sub = 10;
	const double sdu = du / sub;
	u[0] = 0;
	F[0] = 0;
	for (int i = 1; i < kJuttnerPoints; ++i) {
		u[i] = u[i - 1] + du;
		F[i] = F[i - 1];
		double lu = u[i - 1];
		for (int j = 0; j < sub; ++j) {  // juttnerFunction, PIC++/specialmath.cpp:949-951
			F[i] += std::exp(-std::sqrt(1 + lu * lu) / theta) * lu * lu * sdu;
			lu += sdu;
		}
	}
	const double last = F[kJuttnerPoints - 1];
	for (int i = 0; i < kJuttnerPoints; ++i) F[i] /= last;
	F[kJuttnerPoints - 1] = 1.0;
}

std::vector<InjectSlabs> plan_injection(const Grid& g, int nspecies, const int* ppc, double flow_x, double dt, double* len,
                                        unsigned& slab_counter) {
	std::vector<InjectSlabs> out;
	const double dx = g.d[0];
	const double xr = g.x0[0] + (g.N[0] - 2) * dx;  // xgrid[xnumberAdded - 1 - additionalBinNumber]
	InjectSlabs cur{};
	for (int s = 0; s < nspecies; ++s) {
		len[s] += fabs(flow_x * dt);
		if (ppc[s] > 0 && ppc[s] * len[s] >= dx) {
			const int cnt = (int)(len[s] * ppc[s] / dx);
			for (int i = cnt; i > 0; --i) {
				len[s] -= dx / ppc[s];
				if (cur.count == kMaxSlabs) {
					out.push_back(cur);
					cur = InjectSlabs{};
				}
				cur.species[cur.count] = s;
				cur.x[cur.count] = xr - len[s];
				cur.slab[cur.count] = slab_counter++;
				cur.count++;
			}
		}
	}
	if (cur.count > 0) out.push_back(cur);
	return out;
}

void launch_inject(const Grid& g, const SpeciesTable& sp, const InjectSlabs& slabs, const InjectPhysics& ph,
                   const ParticleArrays& P, const ParticleArrays32& Q, bool fp32, long long first, uint32_t* keys,
                   hipStream_t s) {
	const long long total = (long long)slabs.count * g.n[1] * g.n[2];
	if (total == 0) return;
	hipLaunchKernelGGL(k_inject, dim3(nblocks(total, 256)), dim3(256), 0, s, g, sp, slabs, ph, P, Q, fp32 ? 1 : 0, first, keys);
}

void launch_replace(const Grid& g, const SpeciesTable& sp, const InjectPhysics& ph, const int* species_start,
                    const uint32_t* esc, int count, const ParticleArrays& P, const ParticleArrays32& Q, bool fp32,
                    uint32_t* keys, hipStream_t s) {
	if (count == 0) return;
	hipLaunchKernelGGL(k_replace, dim3(nblocks(count, 256)), dim3(256), 0, s, g, sp, ph, species_start, esc, count, P, Q,
	                   fp32 ? 1 : 0, keys);
}

void launch_philox(unsigned long long c0, unsigned long long c1, unsigned long long c2, unsigned long long c3,
                   unsigned long long k0, unsigned long long k1, long long n, unsigned long long* out, hipStream_t s) {
	if (n <= 0) return;
	hipLaunchKernelGGL(k_philox, dim3(nblocks(n, 256)), dim3(256), 0, s, (uint64_t)c0, (uint64_t)c1, (uint64_t)c2,
	                   (uint64_t)c3, (uint64_t)k0, (uint64_t)k1, n, reinterpret_cast<uint64_t*>(out));
}

}  // namespace picgpu
