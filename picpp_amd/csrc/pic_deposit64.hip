// FP64 deposits of the parity path with cooperative lanes: the moments at x^n (J and mu per
// node, rhohat and the pressure partials per dual cell) and the post-push density.
//
// Work distribution (the same tiles as the FP32 deposits): a workgroup serves T cells of one z
// row (node moments) or a (y, z) plane tile of dual cells (cell moments, density) with a group
// of G lanes per cell. The lanes of a group take the cell's particles round-robin (lane l:
// particles l, l+G, ...), so that a wave's particle loads are coalesced; each lane sums its
// particles in that fixed order, and the group then sums the lanes through LDS in lane order.
// The field stencils come from boxes of the tile's nodes and cells staged in LDS.
//
// The per-particle arithmetic is the reference's algorithm (CIC maps, 8-point gathers, gamma,
// alpha, the deposit products) in forms that avoid FP64 divisions: a particle's cell and map
// bases come from its sort key, which the reference's own divisions computed, and its weights
// from its offset in that cell times 1/d, in the closed forms of correlationBspline. Each weight,
// velocity and alpha entry then differs from the reference's by about an ulp, and the sums over
// particles are regrouped; the results agree with the reference to rounding.

// Contraction off: the sums use explicit fma where fused, so the code fixes every rounding.
#pragma clang fp contract(off)

#include "pic_device.hpp"
#include "pic_kernels.hpp"
#include "pic_recip.hpp"
#include "pic_tiles.hpp"

namespace picgpu {

namespace {

constexpr double kPi64 = 3.141592653589793115997963468544185161590576171875;  // 4*atan2(1,1)

// Sums 8 per-lane values over each group of G consecutive lanes (8 <= G, G divides the
// workgroup) through LDS: group lane a < 8 gets the sum of value a over lanes 0, 1, ..., G-1
// in that order. Every thread of the workgroup must call it; red holds 9 doubles per thread.
__device__ inline double group_sum8_64(const double v[8], double* red, int lt, int lane, int G) {
	__syncthreads();  // the previous call's reads are done
	double* row = red + lt * 9;
#pragma unroll
	for (int q = 0; q < 8; ++q) row[q] = v[q];
	__syncthreads();
	double s = 0.0;
	if (lane < 8) {
		const double* col = red + (lt - lane) * 9 + lane;
		for (int l = 0; l < G; ++l) s += col[l * 9];
	}
	return s;
}


// The segment of combined index v among a dual cell's 8 segments and its particle index. The
// segment bounds p[] are nondecreasing, so the sums of the comparisons telescope to segment a
// and its offset off[a] (a form that keeps the tables in registers).
__device__ inline int segment_of(const DualSegments& d, int v, int& particle) {
	int seg = 0, off = d.off[0];
#pragma unroll
	for (int q = 1; q < 8; ++q) {
		const int past = v >= d.p[q] ? 1 : 0;
		seg += past;
		off += past * (d.off[q] - d.off[q - 1]);
	}
	particle = v + off;
	return seg;
}

__device__ inline void load_particle64(const ParticleArrays& P, long long i, double x[3], double p[3]) {
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		x[a] = P.x[a][i];
		p[a] = P.p[a][i];
	}
}

constexpr int kLanes64 = 8;  // lanes per cell or dual cell
constexpr int kTile64 = 16;  // cells per z-row workgroup: 128 lanes
constexpr int kPlaneTj = 4, kPlaneTk = 4;  // dual cells per plane-tile workgroup: 128 lanes
static_assert(kPlaneTj * kPlaneTk * kLanes64 == 128, "plane tiles fill the 128-lane workgroup");
static_assert(kTile64 * kLanes64 == 128, "row tiles fill the 128-lane workgroup");

}  // namespace

// the 8 products w0[i] w1[j] w2[k], t = 4 i + 2 j + k
__device__ inline void corner_products(const double w0[2], const double w1[2], const double w2[2], double W[8]) {
#pragma unroll
	for (int i = 0; i < 2; ++i)
#pragma unroll
		for (int j = 0; j < 2; ++j) {
			const double wij = w0[i] * w1[j];
			W[i * 4 + j * 2 + 0] = wij * w2[0];
			W[i * 4 + j * 2 + 1] = wij * w2[1];
		}
}

// 8-point interpolation (correlationGeneralEfield / correlationBfield, PIC++/correlation.cpp:23-58,
// :103-164) from a staged box, stencil base (i, j, k)
__device__ inline void interp_box(const Box& b, const double* lds, int i, int j, int k, const double W[8], double out[3]) {
	out[0] = out[1] = out[2] = 0.0;
	const int b0 = b.at(i, j, k), sj = b.e[2], si = b.e[1] * b.e[2];
#pragma unroll
	for (int t = 0; t < 8; ++t) {
		const double* v = lds + 3 * (b0 + ((t >> 2) & 1) * si + ((t >> 1) & 1) * sj + (t & 1));
		out[0] = fma(v[0], W[t], out[0]);
		out[1] = fma(v[1], W[t], out[1]);
		out[2] = fma(v[2], W[t], out[2]);
	}
}

// The partner lane's value in lane pairs (0, 1), (2, 3), ...: a DPP quad permutation [1, 0, 3, 2]
// of both halves of the double. Both lanes of every pair must be active.
__device__ inline double swap_pair(double v) {
	const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0xB1, 0xF, 0xF, false);
	const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0xB1, 0xF, 0xF, false);
	return __hiloint2double(hi, lo);
}

// Group sums of 8 per-lane values split by lane parity: group lane a < 8 gets value a summed over
// the even lanes 0, 2, ... (even) and over the odd lanes 1, 3, ... (odd), in lane order.
__device__ inline void group_sum8_pairs(const double v[8], double* red, int lt, int lane, int G, double& even, double& odd) {
	__syncthreads();
	double* row = red + lt * 9;
#pragma unroll
	for (int q = 0; q < 8; ++q) row[q] = v[q];
	__syncthreads();
	even = odd = 0.0;
	if (lane < 8) {
		const double* col = red + (lt - lane) * 9 + lane;
		for (int l = 0; l < G; l += 2) {
			even += col[l * 9];
			odd += col[(l + 1) * 9];
		}
	}
}

// evaluateParticlesRotationTensor + the node loop of updateElectroMagneticParameters
// (PIC++/mover.cpp:810-841, PIC++/momentum_collecting.cpp:397-480): per cell, the 8 corner sums
// of J (3) and of the dielectric tensor mu (9), written as out[(q * 8 + corner) * ncells + cell].
// A group's particles all have the node map of its cell; E comes from the tile's box of nodes,
// B from its box of cells. The 96 sums of a lane would fill the register file, so the lanes of a
// pair split them: each lane forms its own particle's 12 scaled values and 8 node weights, keeps
// one half of the values (J and mu's first row on even lanes, mu's other rows on odd lanes) and
// trades the other half for its partner's, then adds both particles to its 48 sums.
__global__ void __launch_bounds__(128) k_node_moments64(Grid g, SpeciesTable sp, Recip R, ParticleArrays P, const int* __restrict__ start,
                                                        const double* __restrict__ E, const double* __restrict__ B, double dt,
                                                        double theta, int T, double* __restrict__ out) {
	__shared__ double red[128 * 9];
	__shared__ double Eb[3 * 2 * 2 * (kTile64 + 1)];
	__shared__ double Bb[3 * 3 * 3 * (kTile64 + 2)];
	const int G = kLanes64;
	static_assert(kLanes64 % 2 == 0, "lane pairs");
	const DualTile tile = dual_tile(g, T);
	const Box be{{tile.i, tile.j, tile.k0}, {2, 2, T + 1}};
	const Box bb{{tile.i - 1, tile.j - 1, tile.k0 - 1}, {3, 3, T + 2}};
	stage_box(be, E, g.nnodes, g.N[0], g.N[1], g.N[2], true, g, Eb);
	stage_box(bb, B, g.ncells, g.N[0] - 1, g.N[1] - 1, g.N[2] - 1, false, g, Bb);
	__syncthreads();
	const int lt = threadIdx.x;
	const int kb = lt / G, lane = lt % G, half = lane & 1;
	const int bk = tile.k0 + kb;
	const bool cell_ok = kb < T && bk < g.N[2];
	const int ci[3] = {tile.i, tile.j, cell_ok ? bk : tile.k0};
	const long long c = cell_id(g, ci[0], ci[1], ci[2]);
	double xn[3];  // node c (the cell's lower corner)
#pragma unroll
	for (int a = 0; a < 3; ++a) xn[a] = g.x0[a] + ci[a] * g.d[a];
	double acc[8][6];
#pragma unroll
	for (int a = 0; a < 8; ++a)
#pragma unroll
		for (int q = 0; q < 6; ++q) acc[a][q] = 0.0;
	if (cell_ok) {
		for (int s = 0; s < sp.count; ++s) {
			const long long key0 = ((long long)s * g.ncells + c) << 3;
			const int lo = start[key0], hi = start[key0 + 8];
			const double m = sp.mass[s], q = sp.charge[s], w = sp.weight[s];
			const double beta = 0.5 * q * dt / m;
			const double omv = w * theta * dt * dt * 2 * kPi64 * q * q / m * R.ivol;
			const double pcv = q * w * R.ivol;
			// the group's lanes run the same number of rounds, so both lanes of a pair trade values
			for (int base = lo; base < hi; base += G) {
				const int i = base + lane;
				const bool valid = i < hi;
				double x[3], p[3];
				load_particle64(P, valid ? i : hi - 1, x, p);
				AxisW aw[3];
#pragma unroll
				for (int a = 0; a < 3; ++a) aw[a] = axis_weights((x[a] - xn[a]) * R.id[a]);
				double Wn[8], Wc[8], Ef[3], Bf[3], v[3], gm, al[9];
				corner_products(aw[0].nw, aw[1].nw, aw[2].nw, Wn);
				corner_products(aw[0].cw, aw[1].cw, aw[2].cw, Wc);
				interp_box(be, Eb, ci[0], ci[1], ci[2], Wn, Ef);
				interp_box(bb, Bb, ci[0] - 1 + aw[0].hi, ci[1] - 1 + aw[1].hi, ci[2] - 1 + aw[2].hi, Wc, Bf);
				velocity_gamma_r(p, m, R.im2[s], v, gm);
				alpha_tensor_r(beta, v, gm, Ef, Bf, al);
				const double u[3] = {v[0] * gm, v[1] * gm, v[2] * gm};
				double q12[12];  // pcv alpha u (J), then -omv alpha (mu); zero for a lane without a particle
#pragma unroll
				for (int r = 0; r < 3; ++r)
					q12[r] = valid ? pcv * fma(al[r * 3 + 2], u[2], fma(al[r * 3 + 1], u[1], al[r * 3 + 0] * u[0])) : 0.0;
#pragma unroll
				for (int r = 0; r < 9; ++r) q12[3 + r] = valid ? -omv * al[r] : 0.0;
				double mine[6], other[6], Wo[8];
#pragma unroll
				for (int k = 0; k < 6; ++k) {
					mine[k] = half ? q12[6 + k] : q12[k];
					other[k] = swap_pair(half ? q12[k] : q12[6 + k]);
				}
#pragma unroll
				for (int t = 0; t < 8; ++t) Wo[t] = swap_pair(Wn[t]);
#pragma unroll
				for (int t = 0; t < 8; ++t)
#pragma unroll
					for (int k = 0; k < 6; ++k) acc[t][k] = fma(other[k], Wo[t], fma(mine[k], Wn[t], acc[t][k]));
			}
		}
	}
#pragma unroll
	for (int k = 0; k < 6; ++k) {
		double v8[8], even, odd;
#pragma unroll
		for (int a = 0; a < 8; ++a) v8[a] = acc[a][k];
		group_sum8_pairs(v8, red, lt, lane, G, even, odd);
		if (cell_ok && lane < 8) {
			out[((long long)k * 8 + lane) * g.ncells + c] = even;
			out[((long long)(6 + k) * 8 + lane) * g.ncells + c] = odd;
		}
	}
}

// One particle of a dual cell b (cell-map base): its cell weights Wc (3/2 - g, g - 1/2 per axis
// from its offset g from node b, in [1/2, 3/2)) and, with fields, alpha u (rv) for the cell
// moments. Its node map starts at node b + a (a: its segment), node weights (1 - (g - a), g - a).
// The cell-moment and density kernels share this code.
struct DualParticle {
	double Wc[8];
	double p[3];
	double gm;
	double rv[3];
};

template <bool FIELDS>
__device__ inline void dual_particle(const Grid& g, const Recip& R, const ParticleArrays& P, long long pi, int seg, const double xb[3],
                                     const int bi3[3], double m, double im2, double beta, const Box& be, const double* Eb,
                                     const Box& bb, const double* Bb, DualParticle& d) {
	double x[3];
	load_particle64(P, pi, x, d.p);
	const int sa[3] = {(seg >> 2) & 1, (seg >> 1) & 1, seg & 1};
	double nw[3][2], cw[3][2];
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		const double gb = (x[a] - xb[a]) * R.id[a];
		cw[a][0] = 1.5 - gb;
		cw[a][1] = gb - 0.5;
		nw[a][0] = 1.0 - (gb - sa[a]);
		nw[a][1] = gb - sa[a];
	}
	corner_products(cw[0], cw[1], cw[2], d.Wc);
	double vel[3];
	velocity_gamma_r(d.p, m, im2, vel, d.gm);
	if (FIELDS) {
		double Wn[8], Ef[3], Bf[3], al[9];
		corner_products(nw[0], nw[1], nw[2], Wn);
		interp_box(be, Eb, bi3[0] + sa[0], bi3[1] + sa[1], bi3[2] + sa[2], Wn, Ef);
		interp_box(bb, Bb, bi3[0], bi3[1], bi3[2], d.Wc, Bf);
		alpha_tensor_r(beta, vel, d.gm, Ef, Bf, al);
		const double u[3] = {vel[0] * d.gm, vel[1] * d.gm, vel[2] * d.gm};
#pragma unroll
		for (int r = 0; r < 3; ++r) d.rv[r] = fma(al[r * 3 + 2], u[2], fma(al[r * 3 + 1], u[1], al[r * 3 + 0] * u[0]));
	}
}

// Cell loop of updateElectroMagneticParameters (PIC++/momentum_collecting.cpp:521-584) per dual
// cell b (cell-map base): the 8 corner sums of rhohat, rho_out[corner * ncells + b], and the 48
// pressure partials (corner, 6 symmetric components), pi_out[(corner * 6 + r) * ncells + b].
// The dual cell's particles come from the 8 segments (cell b + a, octant 7 - a).
__global__ void __launch_bounds__(128) k_cell_moments64(Grid g, SpeciesTable sp, Recip R,
                                                        ParticleArrays P, const int* __restrict__ start,
                                                        const double* __restrict__ E, const double* __restrict__ B, double dt,
                                                        int Tj, int Tk, double* __restrict__ rho_out, double* __restrict__ pi_out) {
	__shared__ double red[128 * 9];
	__shared__ double Eb[3 * 3 * (kPlaneTj + 2) * (kPlaneTk + 2)];
	__shared__ double Bb[3 * 2 * (kPlaneTj + 1) * (kPlaneTk + 1)];
	const int G = kLanes64;
	const PlaneTile tile = plane_tile(g, Tj, Tk);
	const Box be{{tile.i, tile.j0, tile.k0}, {3, Tj + 2, Tk + 2}};
	const Box bb{{tile.i, tile.j0, tile.k0}, {2, Tj + 1, Tk + 1}};
	if (tile.ok) {
		stage_box(be, E, g.nnodes, g.N[0], g.N[1], g.N[2], true, g, Eb);
		stage_box(bb, B, g.ncells, g.N[0] - 1, g.N[1] - 1, g.N[2] - 1, false, g, Bb);
	}
	__syncthreads();
	const int lt = threadIdx.x;
	const int T = Tj * Tk;
	const int kb = lt / G, lane = lt % G;
	const int bj = tile.j0 + kb / Tk, bk = tile.k0 + kb % Tk;
	const bool base_ok = tile.ok && kb < T && bj < g.N[1] && bk < g.N[2];
	const long long b = base_ok ? cell_id(g, tile.i, bj, bk) : 0;
	const int bi3[3] = {tile.i, bj, bk};
	double xb[3];  // node b
#pragma unroll
	for (int a = 0; a < 3; ++a) xb[a] = g.x0[a] + bi3[a] * g.d[a];
	// lane pairs split the sums (as k_node_moments64): even lanes rhohat and Pi xx, xy, xz, odd
	// lanes Pi yy, yz, zz; the cell weights and the other half's products are traded
	const int half = lane & 1;
	double acc[8][4];
#pragma unroll
	for (int t = 0; t < 8; ++t)
#pragma unroll
		for (int q = 0; q < 4; ++q) acc[t][q] = 0.0;
	if (base_ok) {
		for (int s = 0; s < sp.count; ++s) {
			DualSegments d;
			dual_segments(g, start, tile.i, bj, bk, s, d);
			const double m = sp.mass[s], q = sp.charge[s], w = sp.weight[s];
			const double beta = 0.5 * q * dt / m;
			const double pcv = q * w * R.ivol;
			for (int base = 0; base < d.p[8]; base += G) {  // the same rounds on every lane of the group
				const int v = base + lane;
				const bool valid = v < d.p[8];
				int pi;
				const int seg = segment_of(d, valid ? v : d.p[8] - 1, pi);
				DualParticle dp;
				dual_particle<true>(g, R, P, pi, seg, xb, bi3, m, R.im2[s], beta, be, Eb, bb, Bb, dp);
				const double Tt[6] = {dp.rv[0] * dp.rv[0], dp.rv[0] * dp.rv[1], dp.rv[0] * dp.rv[2],
				                      dp.rv[1] * dp.rv[1], dp.rv[1] * dp.rv[2], dp.rv[2] * dp.rv[2]};
				double Wf[8], Wo[8], mine[4], other[4];  // Wf = pcv Wc: the rhohat contributions
#pragma unroll
				for (int t = 0; t < 8; ++t) Wf[t] = valid ? pcv * dp.Wc[t] : 0.0;
				// per-particle factors of the 4 sums of each half: 1 (rhohat) and Pi xx, xy, xz; Pi yy, yz, zz
				mine[0] = half ? Tt[3] : 1.0;
				mine[1] = half ? Tt[4] : Tt[0];
				mine[2] = half ? Tt[5] : Tt[1];
				mine[3] = half ? 0.0 : Tt[2];
				other[0] = swap_pair(half ? 1.0 : Tt[3]);
				other[1] = swap_pair(half ? Tt[0] : Tt[4]);
				other[2] = swap_pair(half ? Tt[1] : Tt[5]);
				other[3] = swap_pair(half ? Tt[2] : 0.0);
#pragma unroll
				for (int t = 0; t < 8; ++t) Wo[t] = swap_pair(Wf[t]);
#pragma unroll
				for (int t = 0; t < 8; ++t)
#pragma unroll
					for (int k = 0; k < 4; ++k) acc[t][k] = fma(Wo[t], other[k], fma(Wf[t], mine[k], acc[t][k]));
			}
		}
	}
	// even lanes: q 0 rhohat, 1..3 Pi xx, xy, xz; odd lanes: q 0..2 Pi yy, yz, zz
#pragma unroll
	for (int k = 0; k < 4; ++k) {
		double v8[8], even, odd;
#pragma unroll
		for (int t = 0; t < 8; ++t) v8[t] = acc[t][k];
		group_sum8_pairs(v8, red, lt, lane, G, even, odd);
		if (base_ok && lane < 8) {
			if (k == 0)
				rho_out[(long long)lane * g.ncells + b] = even;
			else
				pi_out[((long long)lane * 6 + (k - 1)) * g.ncells + b] = even;
			if (k < 3) pi_out[((long long)lane * 6 + (3 + k)) * g.ncells + b] = odd;
		}
	}
}

// updateDensityParameters (PIC++/momentum_collecting.cpp:760-1020) per dual cell: per species
// [concentration, energy, bulk momentum x3] x 8 corners, then rho x 8 corners, as planes
// out[(plane * 8 + corner) * ncells + b]; the particle weights are k_cell_moments64's.
__global__ void __launch_bounds__(128) k_density64(Grid g, SpeciesTable sp, Recip R, ParticleArrays P, const int* __restrict__ start, int Tj,
                                                   int Tk, double* __restrict__ out) {
	__shared__ double red[128 * 9];
	const int G = kLanes64;
	const PlaneTile tile = plane_tile(g, Tj, Tk);
	const int lt = threadIdx.x;
	const int T = Tj * Tk;
	const int kb = lt / G, lane = lt % G;
	const int bj = tile.j0 + kb / Tk, bk = tile.k0 + kb % Tk;
	const bool base_ok = tile.ok && kb < T && bj < g.N[1] && bk < g.N[2];
	const long long b = base_ok ? cell_id(g, tile.i, bj, bk) : 0;
	const int bi3[3] = {tile.i, bj, bk};
	double xb[3];
#pragma unroll
	for (int a = 0; a < 3; ++a) xb[a] = g.x0[a] + bi3[a] * g.d[a];
	const Box none{{0, 0, 0}, {1, 1, 1}};
	double rho[8];
#pragma unroll
	for (int t = 0; t < 8; ++t) rho[t] = 0.0;
	for (int s = 0; s < sp.count; ++s) {  // every thread runs every species: the sums reduce per species
		double conc[8], en[8], bulk[8][3];
#pragma unroll
		for (int t = 0; t < 8; ++t) {
			conc[t] = en[t] = 0.0;
			bulk[t][0] = bulk[t][1] = bulk[t][2] = 0.0;
		}
		if (base_ok) {
			DualSegments d;
			dual_segments(g, start, tile.i, bj, bk, s, d);
			const double m = sp.mass[s], q = sp.charge[s], w = sp.weight[s];
			const double wv = w * R.ivol, pcv = q * w * R.ivol;
			for (int v = lane; v < d.p[8]; v += G) {
				int pi;
				const int seg = segment_of(d, v, pi);
				DualParticle dp;
				dual_particle<false>(g, R, P, pi, seg, xb, bi3, m, R.im2[s], 0.0, none, nullptr, none, nullptr, dp);
				const double full = dp.gm * m;
#pragma unroll
				for (int t = 0; t < 8; ++t) {
					const double c = wv * dp.Wc[t];
#pragma unroll
					for (int r = 0; r < 3; ++r) bulk[t][r] = fma(dp.p[r], c, bulk[t][r]);
					rho[t] += pcv * dp.Wc[t];
					conc[t] += c;
					en[t] = fma(full, c, en[t]);
				}
			}
		}
		const long long base = (long long)s * 5 * 8;
		{
			const double sum = group_sum8_64(conc, red, lt, lane, G);
			if (base_ok && lane < 8) out[(base + 0 * 8 + lane) * g.ncells + b] = sum;
		}
		{
			const double sum = group_sum8_64(en, red, lt, lane, G);
			if (base_ok && lane < 8) out[(base + 1 * 8 + lane) * g.ncells + b] = sum;
		}
#pragma unroll
		for (int r = 0; r < 3; ++r) {
			double v8[8];
#pragma unroll
			for (int t = 0; t < 8; ++t) v8[t] = bulk[t][r];
			const double sum = group_sum8_64(v8, red, lt, lane, G);
			if (base_ok && lane < 8) out[(base + (2 + r) * 8 + lane) * g.ncells + b] = sum;
		}
	}
	{
		const double sum = group_sum8_64(rho, red, lt, lane, G);
		if (base_ok && lane < 8) out[((long long)sp.count * 5 * 8 + lane) * g.ncells + b] = sum;
	}
}

namespace {
// z-row tiles of kTile64 cells; plane tiles of 4 x 4 dual cells (a multiple of 8 blocks, one
// run of tiles per XCD)
unsigned row_blocks(const Grid& g, int T) { return (unsigned)(g.N[0] * g.N[1] * ((g.N[2] + T - 1) / T)); }
unsigned plane_blocks(const Grid& g, int Tj, int Tk) {
	long long n = (long long)g.N[0] * ((g.N[1] + Tj - 1) / Tj) * ((g.N[2] + Tk - 1) / Tk);
	return (unsigned)(8 * ((n + 7) / 8));
}
}  // namespace


void launch_node_moments(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P, const int* start, const double* E,
                         const double* B, double dt, double theta, double* scratch, hipStream_t s) {
	hipLaunchKernelGGL(k_node_moments64, dim3(row_blocks(g, kTile64)), dim3(128), 0, s, g, sp, make_recip(g, sp), P, start, E, B, dt, theta,
	                   kTile64, scratch);
}
void launch_cell_moments(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P, const int* start, const double* E,
                         const double* B, double dt, double* rho_scratch, double* pi_scratch, hipStream_t s) {
	hipLaunchKernelGGL(k_cell_moments64, dim3(plane_blocks(g, kPlaneTj, kPlaneTk)), dim3(128), 0, s, g, sp, make_recip(g, sp), P, start, E,
	                   B, dt, kPlaneTj, kPlaneTk, rho_scratch, pi_scratch);
}
void launch_density(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P, const int* start, double* scratch,
                    hipStream_t s) {
	hipLaunchKernelGGL(k_density64, dim3(plane_blocks(g, kPlaneTj, kPlaneTk)), dim3(128), 0, s, g, sp, make_recip(g, sp), P, start, kPlaneTj, kPlaneTk,
	                   scratch);
}

}  // namespace picgpu
