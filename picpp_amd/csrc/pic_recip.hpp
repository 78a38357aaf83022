// FP64 per-particle arithmetic without divisions, shared by the FP64 deposits and the FP64 mover:
// reciprocals formed once per launch on the host, map weights from a particle's offset in its
// cell, one reciprocal per velocity and per alpha tensor. Each result is within about an ulp of
// the reference's quotient; the FP64 path's parity bar (1e-12 against the reference) absorbs it.
#pragma once
#include "pic_device.hpp"

namespace picgpu {

// Reciprocals that replace divisions in the per-particle arithmetic: each product is within an
// ulp of the reference's quotient, which the 1e-12 parity bar absorbs.
struct Recip {
	double id[3];     // 1 / d per axis
	double ivol;      // 1 / cell volume
	double im2[kMaxSpecies];  // 1 / m^2 per species
};

inline Recip make_recip(const Grid& g, const SpeciesTable& sp) {
	Recip R{};
	for (int a = 0; a < 3; ++a) R.id[a] = 1.0 / g.d[a];
	R.ivol = 1.0 / g.volume;
	for (int q = 0; q < kMaxSpecies; ++q) R.im2[q] = 1.0 / (sp.mass[q] * sp.mass[q]);
	return R;
}

// Particle::gammaFactor / getVelocity (PIC++/particle.cpp:216-235, :381-390) with 1 / m^2 and
// one reciprocal of m gamma
__device__ inline void velocity_gamma_r(const double p[3], double m, double im2, double v[3], double& g) {
	double p2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
	g = sqrt(p2 * im2 + 1.0);
	double r = 1.0 / (m * g);
	v[0] = p[0] * r;
	v[1] = p[1] * r;
	v[2] = p[2] * r;
}

// evaluateAlphaRotationTensor (PIC++/mover.cpp:851-908) with one reciprocal of the denominator:
// alpha_ij = [delta_ij + bs^2 B_i B_j - bs sum_l eps_jil B_l] / (G (1 + bs^2 B^2)).
__device__ inline void alpha_tensor_r(double beta, const double v[3], double gamma, const double E[3], const double B[3],
                                      double a[9]) {
	double G = (beta * (E[0] * v[0] + E[1] * v[1] + E[2] * v[2])) + gamma;
	double bs = beta / G;
	double b2 = bs * bs;
	double iden = 1.0 / (G * (1 + b2 * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2])));
	// the Levi-Civita terms of the reference's index order (PIC++/initialization.cpp:779-793)
	a[0] = (1.0 + (b2 * B[0] * B[0])) * iden;
	a[1] = ((0.0 + (b2 * B[0] * B[1])) - (bs * -1.0 * B[2])) * iden;
	a[2] = ((0.0 + (b2 * B[0] * B[2])) - (bs * 1.0 * B[1])) * iden;
	a[3] = ((0.0 + (b2 * B[1] * B[0])) - (bs * 1.0 * B[2])) * iden;
	a[4] = (1.0 + (b2 * B[1] * B[1])) * iden;
	a[5] = ((0.0 + (b2 * B[1] * B[2])) - (bs * -1.0 * B[0])) * iden;
	a[6] = ((0.0 + (b2 * B[2] * B[0])) - (bs * -1.0 * B[1])) * iden;
	a[7] = ((0.0 + (b2 * B[2] * B[1])) - (bs * 1.0 * B[0])) * iden;
	a[8] = (1.0 + (b2 * B[2] * B[2])) * iden;
}

// Weights of one axis from the offset g = (x - x_n) / d of a particle from node n, in the closed
// forms of correlationBspline (PIC++/correlation.cpp:824-835) for a particle between nodes n and
// n + 1: node weights (1 - g, g) on nodes n, n + 1; cell weights (1/2 - g, 1/2 + g) on cells
// n - 1, n when g < 1/2, else (3/2 - g, g - 1/2) on cells n, n + 1. The bases come from the
// particle's sort key (the reference's own divisions), so only the weights' rounding differs
// from the reference's quotients, by an ulp.
struct AxisW {
	double nw[2], cw[2];
	int hi;  // 1: the cell map starts at the particle's own cell
};

__device__ inline AxisW axis_weights(double gofs) {
	AxisW w;
	w.nw[0] = 1.0 - gofs;
	w.nw[1] = gofs;
	w.hi = gofs >= 0.5 ? 1 : 0;
	w.cw[0] = (w.hi ? 1.5 : 0.5) - gofs;
	w.cw[1] = gofs - (w.hi ? 0.5 : -0.5);
	return w;
}

// Node and cell maps of one axis at position x (updateCorrelationMapNodeX / CellX,
// PIC++/correlation.cpp:1112-1145, :1283-1312) from t = (x - x0) / d as a product: node base
// floor(t), cell base one below it in the lower half of the cell; the weights from the offset
// in the cell. At a cell face the base may come out one below or above the reference's, whose
// weights then put the whole particle on the same node or cell: the maps are continuous there.
// Bases are clamped into the ghosted arrays; returns false when a clamp was needed.
__device__ inline bool axis_map_fast(double x, double x0, double d, double id, int N, AxisMap& m) {
	const int n = (int)floor((x - x0) * id);
	const AxisW w = axis_weights((x - (x0 + n * d)) * id);
	m.nbase = n;
	m.cbase = n - 1 + w.hi;
	m.nw[0] = w.nw[0];
	m.nw[1] = w.nw[1];
	m.cw[0] = w.cw[0];
	m.cw[1] = w.cw[1];
	bool ok = true;
	if (!(m.nbase >= 0 && m.nbase <= N - 1)) {
		m.nbase = m.nbase < 0 ? 0 : N - 1;
		ok = false;
	}
	if (!(m.cbase >= 0 && m.cbase <= N - 2)) {
		m.cbase = m.cbase < 0 ? 0 : N - 2;
		ok = false;
	}
	return ok;
}

__device__ inline bool maps3_fast(const Grid& g, const Recip& R, const double x[3], AxisMap m[3]) {
	bool ok = true;
#pragma unroll
	for (int a = 0; a < 3; ++a) ok = axis_map_fast(x[a], g.x0[a], g.d[a], R.id[a], g.N[a], m[a]) && ok;
	return ok;
}

// Particle::evaluateVelocity (PIC++/particle.cpp:366-379): p / m below the relativistic
// threshold p^2 < 1e-4 m^2, else p / (m gamma)
__device__ inline void evaluate_velocity_r(const double p[3], double m, double im2, double v[3]) {
	const double p2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
	double r;
	if (p2 < 0.0001 * m * m)
		r = 1.0 / m;
	else
		r = 1.0 / (m * sqrt(p2 * im2 + 1.0));
	v[0] = p[0] * r;
	v[1] = p[1] * r;
	v[2] = p[2] * r;
}

}  // namespace picgpu
