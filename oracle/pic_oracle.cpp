// FP64 CPU restatement of the PIC++ implicit step (oracle). TEST INFRASTRUCTURE ONLY:
// the checker for the HIP product path; never linked into the product library.
//
// Every function cites the reference file:line whose arithmetic it restates. The
// arithmetic is written in the same operation order as the reference so that the
// restatement and the compiled reference agree to a few ulps.
#include "pic_oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>

namespace picoracle {

namespace {
const double kPi = 4 * std::atan2(1.0, 1.0);               // PIC++/constants.h:28
const double kCprime = 1.025;                               // constants.h:45
const double kCprime2 = kCprime * kCprime;                  // constants.h:46
const double kRelPrecision = 0.0001;                        // constants.h:39
const double kVelErr = 1e-14;                               // constants.h:33
const double kCoordErr = 1e-8;                              // constants.h:34
const int kPicardIterations = 3;                            // mover.cpp:222-223
const int kMaxGmres = 50;                                   // constants.h:55
const double kMaxError = 1e-8;                              // constants.h:31
}  // namespace

Oracle::Oracle(int nx, int ny, int nz, double dx, double dy, double dz, double dt_, const std::vector<Species>& sp)
    : dt(dt_), theta(0.5), eta(0.5), cprime(kCprime), species(sp) {
	n[0] = nx;
	n[1] = ny;
	n[2] = nz;
	d[0] = dx;
	d[1] = dy;
	d[2] = dz;
	for (int a = 0; a < 3; ++a) {
		if (n[a] < 5) throw std::runtime_error("oracle: every axis needs at least 5 cells");
		N[a] = n[a] + 2 * kGhost;
		L[a] = n[a] * d[a];
		// leftX = xsizeGeneral + firstAbsoluteXindex * deltaX, firstAbsoluteXindex = -2
		// (PIC++/initialization.cpp:88-95, :194)
		x0[a] = L[a] + (-kGhost) * d[a];
	}
	volume = d[0] * d[1] * d[2];
	E.assign(nnodes() * 3, 0);
	Etemp = E;
	Enew = E;
	J = E;
	Jm = E;
	divPi = E;
	divPim = E;
	mu.assign(nnodes() * 9, 0);
	B.assign(ncells() * 3, 0);
	Bnew = B;
	rhohat.assign(ncells(), 0);
	rhom = rhohat;
	rho = rhohat;
	Pi.assign(ncells() * 9, 0);
	conc.assign(species.size(), std::vector<double>(ncells(), 0));
	energy = conc;
	bulk.assign(species.size(), std::vector<double>(ncells() * 3, 0));
	rhs.assign(ncells() * 3, 0);
}

// ---------------------------------------------------------------- shape functions
// correlationBspline, splineOrder 0 branch (PIC++/correlation.cpp:798-835).
double Oracle::bspline(double x, double hw, double left, double right) {
	if (right < left) throw std::runtime_error("bspline: right < left");
	if (x < left - hw) return 0;
	if (x > right + hw) return 0;
	double c;
	if (x < left + hw) {
		c = 0.5 * (x + hw - left) / hw;
	} else if (x > right - hw) {
		c = 0.5 * (right - (x - hw)) / hw;
	} else {
		c = 1;
	}
	if (c < 0 && c > -1e-14) c = 0;
	return c;
}

// updateCorrelationMapCell{X,Y,Z} (correlation.cpp:1112-1258) and
// updateCorrelationMapNode{X,Y,Z} (correlation.cpp:1283-1429), splineOrder 0.
// The particle half-width is deltaX/2 (initialization.cpp:8491).
void Oracle::update_maps(Particle& p) const {
	for (int a = 0; a < 3; ++a) {
		double x = p.x[a];
		double hw = d[a] / 2;
		int cc = (int)std::floor((x - x0[a]) / d[a]);
		double mid = ((x0[a] + cc * d[a]) + (x0[a] + (cc + 1) * d[a])) / 2;  // middleXgrid
		int cs = (x < mid) ? -1 : 0;
		for (int t = 0; t < 2; ++t) {
			int i = cc + cs + t;
			p.ci[a][t] = i;
			p.cw[a][t] = bspline(x, hw, x0[a] + i * d[a], x0[a] + (i + 1) * d[a]);
		}
		int nc = (int)std::floor(((x - x0[a]) / d[a]) + 0.5);
		int ns = (x < x0[a] + nc * d[a]) ? -1 : 0;
		for (int t = 0; t < 2; ++t) {
			int i = nc + ns + t;
			p.ni[a][t] = i;
			p.nw[a][t] = bspline(x, hw, x0[a] + (i - 0.5) * d[a], x0[a] + (i + 0.5) * d[a]);
		}
	}
}

// correlationGeneralEfield (correlation.cpp:103-164).
void Oracle::gather_node(const Particle& p, const std::vector<double>& f, double out[3]) const {
	out[0] = out[1] = out[2] = 0;
	for (int i = 0; i < 2; ++i)
		for (int j = 0; j < 2; ++j)
			for (int k = 0; k < 2; ++k) {
				const double* e = &f[node(p.ni[0][i], p.ni[1][j], p.ni[2][k]) * 3];
				double c = p.nw[0][i] * p.nw[1][j] * p.nw[2][k];
				out[0] = out[0] + e[0] * c;
				out[1] = out[1] + e[1] * c;
				out[2] = out[2] + e[2] * c;
			}
}

// correlationBfield (correlation.cpp:23-58).
void Oracle::gather_cell(const Particle& p, const std::vector<double>& f, double out[3]) const {
	out[0] = out[1] = out[2] = 0;
	for (int i = 0; i < 2; ++i)
		for (int j = 0; j < 2; ++j)
			for (int k = 0; k < 2; ++k) {
				const double* b = &f[cell(p.ci[0][i], p.ci[1][j], p.ci[2][k]) * 3];
				double c = p.cw[0][i] * p.cw[1][j] * p.cw[2][k];
				out[0] = out[0] + b[0] * c;
				out[1] = out[1] + b[1] * c;
				out[2] = out[2] + b[2] * c;
			}
}

// evaluateAlphaRotationTensor (mover.cpp:851-908); Levi-Civita table initialization.cpp:779-793.
void Oracle::alpha_tensor(double beta, const double v[3], double gamma, const double Ef[3], const double Bf[3],
                          double out[9]) {
	static const double eps[3][3][3] = {{{0, 0, 0}, {0, 0, 1}, {0, -1, 0}},
	                                    {{0, 0, -1}, {0, 0, 0}, {1, 0, 0}},
	                                    {{0, 1, 0}, {-1, 0, 0}, {0, 0, 0}}};
	double G = (beta * (Ef[0] * v[0] + Ef[1] * v[1] + Ef[2] * v[2])) + gamma;
	double bs = beta / G;
	double b2 = bs * bs;
	double den = G * (1 + b2 * (Bf[0] * Bf[0] + Bf[1] * Bf[1] + Bf[2] * Bf[2]));
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j) {
			double r = (i == j ? 1.0 : 0.0) + (b2 * Bf[i] * Bf[j]);
			for (int l = 0; l < 3; ++l)
				if (eps[j][i][l] != 0) r -= (bs * eps[j][i][l] * Bf[l]);
			r /= den;
			out[i * 3 + j] = r;
		}
}

// ---------------------------------------------------------------- ghosts and folds
// exchangeGeneralEfield{X,Y,Z}, one-rank periodic branch (fields_exchange.cpp:158-170, :313-321, :471-482).
void Oracle::exchange_node_field(std::vector<double>& f) const {
	for (int ax = 0; ax < 3; ++ax) {
		int NN = N[ax];
		int ext[3] = {N[0] + 1, N[1] + 1, N[2] + 1};
		int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
		for (int a = 0; a < ext[o1]; ++a)
			for (int b = 0; b < ext[o2]; ++b) {
				auto at = [&](int s) {
					int c[3];
					c[ax] = s;
					c[o1] = a;
					c[o2] = b;
					return node(c[0], c[1], c[2]) * 3;
				};
				for (int i = 0; i <= 1; ++i) {
					for (int l = 0; l < 3; ++l) f[at(NN - 1 + i) + l] = f[at(3 + i) + l];
					for (int l = 0; l < 3; ++l) f[at(i) + l] = f[at(NN - 4 + i) + l];
				}
				for (int l = 0; l < 3; ++l) f[at(NN - 2) + l] = f[at(2) + l];
			}
	}
}

// exchangeGeneralBfield{X,Y,Z}, one-rank periodic branch (fields_exchange.cpp:629-637, :768-782, :912-926).
void Oracle::exchange_cell_field(std::vector<double>& f) const {
	for (int ax = 0; ax < 3; ++ax) {
		int NN = N[ax];
		int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
		for (int a = 0; a < N[o1]; ++a)
			for (int b = 0; b < N[o2]; ++b) {
				auto at = [&](int s) {
					int c[3];
					c[ax] = s;
					c[o1] = a;
					c[o2] = b;
					return cell(c[0], c[1], c[2]) * 3;
				};
				for (int i = 0; i <= 1; ++i) {
					for (int l = 0; l < 3; ++l) f[at(NN - 2 + i) + l] = f[at(2 + i) + l];
					for (int l = 0; l < 3; ++l) f[at(i) + l] = f[at(NN - 4 + i) + l];
				}
			}
	}
}

// sumNodeVectorParameters{X,Y,Z} / sumNodeMatrixParameters{X,Y,Z}, one-rank periodic branch
// (momentum_exchange.cpp:172-195, :446-460): F[N-i] += F[4-i]; F[4-i] = F[N-i], i = 0..4.
void Oracle::fold_nodes(std::vector<double>& f, int nc) const {
	for (int ax = 0; ax < 3; ++ax) {
		int NN = N[ax];
		int ext[3] = {N[0] + 1, N[1] + 1, N[2] + 1};
		int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
		for (int a = 0; a < ext[o1]; ++a)
			for (int b = 0; b < ext[o2]; ++b) {
				auto at = [&](int s) {
					int c[3];
					c[ax] = s;
					c[o1] = a;
					c[o2] = b;
					return node(c[0], c[1], c[2]) * nc;
				};
				for (int i = 0; i <= 2 * 1 + 2; ++i)
					for (int l = 0; l < nc; ++l) {
						f[at(NN - i) + l] += f[at(4 - i) + l];
						f[at(4 - i) + l] = f[at(NN - i) + l];
					}
			}
	}
}

// Cell folds, one-rank periodic branch: pairs (N-1-i, lo_base-i), i < pairs.
// chargeDensityHat / density parameters: pairs 4, lo_base 3 (momentum_exchange.cpp:807-811, :940-952);
// pressureTensor: pairs 3, lo_base 2 (momentum_exchange.cpp:1657-1660, :1690-1692, :1727-1729).
void Oracle::fold_cells(std::vector<double>& f, int nc, int pairs, int /*unused*/, int lo_base) const {
	for (int ax = 0; ax < 3; ++ax) {
		int NN = N[ax];
		int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
		for (int a = 0; a < N[o1]; ++a)
			for (int b = 0; b < N[o2]; ++b) {
				auto at = [&](int s) {
					int c[3];
					c[ax] = s;
					c[o1] = a;
					c[o2] = b;
					return cell(c[0], c[1], c[2]) * nc;
				};
				for (int i = 0; i < pairs; ++i)
					for (int l = 0; l < nc; ++l) {
						f[at(NN - 1 - i) + l] = f[at(NN - 1 - i) + l] + f[at(lo_base - i) + l];
						f[at(lo_base - i) + l] = f[at(NN - 1 - i) + l];
					}
			}
	}
}

// ---------------------------------------------------------------- particles: maps, alpha
void Oracle::update_particle_maps() {
	for (auto& p : particles) update_maps(p);
}

static inline void velocity_gamma(const Particle& p, double m, double v[3], double& g) {
	// Particle::gammaFactor / getVelocity (particle.cpp:216-235, :381-390)
	double p2 = p.p[0] * p.p[0] + p.p[1] * p.p[1] + p.p[2] * p.p[2];
	g = std::sqrt(p2 / (m * m) + 1.0);
	for (int a = 0; a < 3; ++a) v[a] = p.p[a] / (m * g);
}

// evaluateParticlesRotationTensor (mover.cpp:810-841): E^n, B^n at x^n.
void Oracle::evaluate_rotation_tensor() {
	for (auto& p : particles) {
		const Species& s = species[p.sp];
		double beta = 0.5 * s.charge * dt / s.mass;
		double v[3], g;
		velocity_gamma(p, s.mass, v, g);
		double e[3], b[3];
		gather_node(p, E, e);
		gather_cell(p, B, b);
		alpha_tensor(beta, v, g, e, b, p.alpha);
	}
}

// ---------------------------------------------------------------- moments
// evaluateDivPressureTensor (fields.cpp:3158-3223).
void Oracle::div_pressure(int i, int j, int k, double out[3]) const {
	out[0] = out[1] = out[2] = 0;
	if (i == 0 || i == N[0] || j == 0 || j == N[1] || k == 0 || k == N[2]) return;
	auto P = [&](int a, int b, int c, int l, int m) { return Pi[cell(a, b, c) * 9 + l * 3 + m]; };
	for (int col = 0; col < 3; ++col) {
		double rx = (P(i, j, k, 0, col) + P(i, j - 1, k, 0, col) + P(i, j, k - 1, 0, col) + P(i, j - 1, k - 1, 0, col)) * 0.25;
		double lx = (P(i - 1, j, k, 0, col) + P(i - 1, j - 1, k, 0, col) + P(i - 1, j, k - 1, 0, col) +
		             P(i - 1, j - 1, k - 1, 0, col)) * 0.25;
		double by = (P(i, j, k, 1, col) + P(i - 1, j, k, 1, col) + P(i, j, k - 1, 1, col) + P(i - 1, j, k - 1, 1, col)) * 0.25;
		double fy = (P(i, j - 1, k, 1, col) + P(i - 1, j - 1, k, 1, col) + P(i, j - 1, k - 1, 1, col) +
		             P(i - 1, j - 1, k - 1, 1, col)) * 0.25;
		double tz = (P(i, j, k, 2, col) + P(i, j - 1, k, 2, col) + P(i - 1, j, k, 2, col) + P(i - 1, j - 1, k, 2, col)) * 0.25;
		double bz = (P(i, j, k - 1, 2, col) + P(i, j - 1, k - 1, 2, col) + P(i - 1, j, k - 1, 2, col) +
		             P(i - 1, j - 1, k - 1, 2, col)) * 0.25;
		out[col] = ((rx - lx) / d[0]) + ((by - fy) / d[1]) + ((tz - bz) / d[2]);
	}
}

// evaluateDivFlux (fields.cpp:3060-3156).
double Oracle::div_flux(int i, int j, int k) const {
	auto F = [&](int a, int b, int c, int l) { return J[node(a, b, c) * 3 + l]; };
	double rX = (F(i + 1, j, k, 0) + F(i + 1, j + 1, k, 0) + F(i + 1, j, k + 1, 0) + F(i + 1, j + 1, k + 1, 0)) * 0.25;
	double lX = (F(i, j, k, 0) + F(i, j + 1, k, 0) + F(i, j, k + 1, 0) + F(i, j + 1, k + 1, 0)) * 0.25;
	double rY = (F(i, j + 1, k, 1) + F(i + 1, j + 1, k, 1) + F(i, j + 1, k + 1, 1) + F(i + 1, j + 1, k + 1, 1)) * 0.25;
	double lY = (F(i, j, k, 1) + F(i + 1, j, k, 1) + F(i, j, k + 1, 1) + F(i + 1, j, k + 1, 1)) * 0.25;
	double rZ = (F(i + 1, j, k + 1, 2) + F(i, j, k + 1, 2) + F(i, j + 1, k + 1, 2) + F(i + 1, j + 1, k + 1, 2)) * 0.25;
	double lZ = (F(i + 1, j, k, 2) + F(i, j, k, 2) + F(i, j + 1, k, 2) + F(i + 1, j + 1, k, 2)) * 0.25;
	return ((rX - lX) / d[0]) + ((rY - lY) / d[1]) + ((rZ - lZ) / d[2]);
}

// updateElectroMagneticParameters, implicit branch, periodic (momentum_collecting.cpp:365-758).
void Oracle::update_em_parameters() {
	std::fill(J.begin(), J.end(), 0.0);
	std::fill(Jm.begin(), Jm.end(), 0.0);
	std::fill(mu.begin(), mu.end(), 0.0);
	std::fill(divPi.begin(), divPi.end(), 0.0);
	std::fill(divPim.begin(), divPim.end(), 0.0);
	for (auto& p : particles) {  // :397-480
		const Species& s = species[p.sp];
		double v[3], g;
		velocity_gamma(p, s.mass, v, g);
		double u[3] = {v[0] * g, v[1] * g, v[2] * g};
		double rv[3];
		for (int a = 0; a < 3; ++a) rv[a] = p.alpha[a * 3 + 0] * u[0] + p.alpha[a * 3 + 1] * u[1] + p.alpha[a * 3 + 2] * u[2];
		double omega = s.weight * theta * dt * dt * 2 * kPi * s.charge * s.charge / s.mass;
		double pc = s.charge * s.weight;
		for (int i = 0; i < 2; ++i)
			for (int j = 0; j < 2; ++j)
				for (int k = 0; k < 2; ++k) {
					size_t nd = node(p.ni[0][i], p.ni[1][j], p.ni[2][k]);
					double corr = p.nw[0][i] * p.nw[1][j] * p.nw[2][k] / volume;
					double* tgt = (pc > 0) ? &J[nd * 3] : &Jm[nd * 3];
					for (int a = 0; a < 3; ++a) tgt[a] += rv[a] * (pc * corr);
					for (int a = 0; a < 9; ++a) mu[nd * 9 + a] = mu[nd * 9 + a] - p.alpha[a] * (omega * corr);
				}
	}
	std::fill(rhohat.begin(), rhohat.end(), 0.0);
	std::fill(rho.begin(), rho.end(), 0.0);
	std::fill(rhom.begin(), rhom.end(), 0.0);
	std::fill(Pi.begin(), Pi.end(), 0.0);
	for (auto& p : particles) {  // :521-584
		const Species& s = species[p.sp];
		double v[3], g;
		velocity_gamma(p, s.mass, v, g);
		double u[3] = {v[0] * g, v[1] * g, v[2] * g};
		double rv[3];
		for (int a = 0; a < 3; ++a) rv[a] = p.alpha[a * 3 + 0] * u[0] + p.alpha[a * 3 + 1] * u[1] + p.alpha[a * 3 + 2] * u[2];
		double T[9] = {rv[0] * rv[0], rv[0] * rv[1], rv[0] * rv[2], 0, rv[1] * rv[1], rv[1] * rv[2], 0, 0, rv[2] * rv[2]};
		T[3] = T[1];
		T[6] = T[2];
		T[7] = T[5];
		double pc = s.charge * s.weight;
		for (int i = 0; i < 2; ++i)
			for (int j = 0; j < 2; ++j)
				for (int k = 0; k < 2; ++k) {
					int ci = p.ci[0][i], cj = p.ci[1][j], ck = p.ci[2][k];
					double corr = p.cw[0][i] * p.cw[1][j] * p.cw[2][k] / volume;
					if (pc > 0)
						rhohat[cell(ci, cj, ck)] += pc * corr;
					else
						rhom[cell(ci, cj, ck)] += pc * corr;
					// pressureTensor[curI][j][k]: the stencil indices j, k, not curJ, curK
					// (momentum_collecting.cpp:565, :577).
					double* P = &Pi[cell(ci, j, k) * 9];
					for (int a = 0; a < 9; ++a) P[a] += pc * corr * T[a];
				}
	}
	fold_cells(rhohat, 1, 4, 0, 3);  // sumChargeDensityHat{X,Y,Z}
	fold_cells(rhom, 1, 4, 0, 3);
	fold_cells(Pi, 9, 3, 0, 2);      // sumCellMatrixParameters{X,Y,Z}
	fold_nodes(J, 3);                // sumNodeVectorParameters{X,Y,Z}
	fold_nodes(Jm, 3);
	fold_nodes(divPi, 3);
	fold_nodes(divPim, 3);
	for (size_t q = 0; q < J.size(); ++q) {
		J[q] += Jm[q];
		divPi[q] += divPim[q];
	}
	for (size_t q = 0; q < rhohat.size(); ++q) rhohat[q] += rhom[q];
	for (int i = 0; i <= N[0]; ++i)  // :623-636
		for (int j = 0; j <= N[1]; ++j)
			for (int k = 0; k <= N[2]; ++k) {
				double dp[3];
				div_pressure(i, j, k, dp);
				size_t nd = node(i, j, k) * 3;
				for (int a = 0; a < 3; ++a) {
					divPi[nd + a] = dp[a];
					J[nd + a] = J[nd + a] - dp[a] * eta * dt;
				}
			}
	for (int i = 0; i < N[0]; ++i)  // :651-665
		for (int j = 0; j < N[1]; ++j)
			for (int k = 0; k < N[2]; ++k) rhohat[cell(i, j, k)] -= dt * theta * div_flux(i, j, k);
	fold_nodes(mu, 9);  // sumNodeMatrixParameters{X,Y,Z} (:742-744)
}

// ---------------------------------------------------------------- Maxwell system
// evaluateRotB (fields.cpp:2799-2860).
void Oracle::rot_b(int i, int j, int k, double out[3]) const {
	auto b = [&](int a, int bb, int c, int l) { return B[cell(a, bb, c) * 3 + l]; };
	double rX[3], lX[3], rY[3], lY[3], rZ[3], lZ[3];
	for (int l = 0; l < 3; ++l) {
		rX[l] = (b(i, j, k, l) + b(i, j - 1, k, l) + b(i, j, k - 1, l) + b(i, j - 1, k - 1, l)) / 4.0;
		lX[l] = (b(i - 1, j, k, l) + b(i - 1, j - 1, k, l) + b(i - 1, j, k - 1, l) + b(i - 1, j - 1, k - 1, l)) / 4.0;
		rY[l] = (b(i, j, k, l) + b(i - 1, j, k, l) + b(i, j, k - 1, l) + b(i - 1, j, k - 1, l)) / 4.0;
		lY[l] = (b(i, j - 1, k, l) + b(i - 1, j - 1, k, l) + b(i, j - 1, k - 1, l) + b(i - 1, j - 1, k - 1, l)) / 4.0;
		rZ[l] = (b(i, j, k, l) + b(i - 1, j, k, l) + b(i, j - 1, k, l) + b(i - 1, j - 1, k, l)) / 4.0;
		lZ[l] = (b(i, j, k - 1, l) + b(i - 1, j, k - 1, l) + b(i, j - 1, k - 1, l) + b(i - 1, j - 1, k - 1, l)) / 4.0;
	}
	out[0] = (rY[2] - lY[2]) / d[1] - (rZ[1] - lZ[1]) / d[2];
	out[1] = (rZ[0] - lZ[0]) / d[2] - (rX[2] - lX[2]) / d[0];
	out[2] = (rX[1] - lX[1]) / d[0] - (rY[0] - lY[0]) / d[1];
}

// evaluateGradDensity (fields.cpp:3229-3259).
void Oracle::grad_rhohat(int i, int j, int k, double out[3]) const {
	auto r = [&](int a, int b, int c) { return rhohat[cell(a, b, c)]; };
	double rX = (r(i, j, k) + r(i, j - 1, k) + r(i, j, k - 1) + r(i, j - 1, k - 1)) * 0.25;
	double lX = (r(i - 1, j, k) + r(i - 1, j - 1, k) + r(i - 1, j, k - 1) + r(i - 1, j - 1, k - 1)) * 0.25;
	double rY = (r(i, j, k) + r(i - 1, j, k) + r(i, j, k - 1) + r(i - 1, j, k - 1)) * 0.25;
	double lY = (r(i, j - 1, k) + r(i - 1, j - 1, k) + r(i, j - 1, k - 1) + r(i - 1, j - 1, k - 1)) * 0.25;
	double rZ = (r(i, j, k) + r(i, j - 1, k) + r(i - 1, j, k) + r(i - 1, j - 1, k)) * 0.25;
	double lZ = (r(i, j, k - 1) + r(i, j - 1, k - 1) + r(i - 1, j, k - 1) + r(i - 1, j - 1, k - 1)) * 0.25;
	out[0] = (rX - lX) / d[0];
	out[1] = (rY - lY) / d[1];
	out[2] = (rZ - lZ) / d[2];
}

// Right part of createInternalEquation (fields.cpp:2635-2715) for interior nodes;
// ghost rows are createFakeEquation with a zero right part (fields.cpp:3261-3271).
void Oracle::build_rhs() {
	std::fill(rhs.begin(), rhs.end(), 0.0);
	for (int i = kGhost; i <= N[0] - 3; ++i)
		for (int j = kGhost; j <= N[1] - 3; ++j)
			for (int k = kGhost; k <= N[2] - 3; ++k) {
				double rb[3], gr[3];
				rot_b(i, j, k, rb);
				grad_rhohat(i, j, k, gr);
				size_t nd = node(i, j, k) * 3;
				for (int a = 0; a < 3; ++a) {
					double t1 = (rb[a] * kCprime - (J[nd + a] * 4 * kPi)) * (theta * dt);
					double t2 = gr[a] * kCprime2 * theta * theta * dt * dt * 4 * kPi;
					rhs[lvec(i, j, k) * 3 + a] = (E[nd + a] + t1) - t2;
				}
			}
}

// createInternalEquation{X,Y,Z}, live branch _27_points = false
// (fields.cpp:1396-1509, :1961-2072, :2521-2632). Entries in push_back order.
void Oracle::row_entries(int i, int j, int k, int c, std::vector<double>& val, std::vector<int>& idx) const {
	val.clear();
	idx.clear();
	double C = (theta * dt) * (theta * dt) * kCprime2;
	double dx2 = d[0] * d[0], dy2 = d[1] * d[1], dz2 = d[2] * d[2];
	double D2[3] = {dx2, dy2, dz2};
	auto M = [&](int a, int b, int cc, int l, int m) { return mu[node(a, b, cc) * 9 + l * 3 + m]; };
	auto push = [&](double v, int a, int b, int cc, int l) {
		val.push_back(v);
		idx.push_back((int)(lvec(a, b, cc) * 3 + l));
	};
	int p[3] = {i, j, k};
	auto off = [&](int ax, int s) {
		int q[3] = {p[0], p[1], p[2]};
		q[ax] += s;
		return std::vector<int>{q[0], q[1], q[2]};
	};
	auto off2 = [&](int ax1, int s1, int ax2, int s2) {
		int q[3] = {p[0], p[1], p[2]};
		q[ax1] += s1;
		q[ax2] += s2;
		return std::vector<int>{q[0], q[1], q[2]};
	};
	// component orders of the same-node and face entries for each row
	static const int others[3][2] = {{1, 2}, {0, 2}, {0, 1}};
	static const int face_order[3][3] = {{0, 1, 2}, {1, 0, 2}, {2, 1, 0}};
	static const int trans_order[3][2] = {{1, 2}, {0, 2}, {1, 0}};  // -C/D^2 face entries, axis order
	// diagonal
	double e = 1.0 - M(i, j, k, c, c);
	for (int ax = 0; ax < 3; ++ax) {
		if (ax == c) {
			if (c == 0)
				e += C * ((2.0 - 2 * M(i, j, k, c, c)) / D2[ax]);
			else
				e += C * (2.0 - 2 * M(i, j, k, c, c)) / D2[ax];
		} else {
			if (c == 0 || (c == 1 && ax == 2) || (c == 2 && ax == 1))
				e += C * 2.0 / D2[ax];
			else
				e += C * (2.0 / D2[ax]);
		}
	}
	push(e, i, j, k, c);
	for (int t = 0; t < 2; ++t) {
		int m = others[c][t];
		double v = -M(i, j, k, c, m);
		v -= C * 2 * M(i, j, k, c, m) / D2[c];
		push(v, i, j, k, m);
	}
	for (int s : {1, -1}) {
		auto q = off(c, s);
		for (int t = 0; t < 3; ++t) {
			int m = face_order[c][t];
			double v;
			if (m == c)
				v = C * (-1.0 + M(q[0], q[1], q[2], c, c)) / D2[c];
			else
				v = C * M(q[0], q[1], q[2], c, m) / D2[c];
			push(v, q[0], q[1], q[2], m);
		}
	}
	for (int t = 0; t < 2; ++t) {
		int ax = trans_order[c][t];
		double v = -C / D2[ax];
		auto qp = off(ax, 1);
		auto qm = off(ax, -1);
		push(v, qp[0], qp[1], qp[2], c);
		push(v, qm[0], qm[1], qm[2], c);
	}
	// edge corners: plane (c, d) for the two other axes, in the reference's plane order
	static const int plane_order[3][2] = {{1, 2}, {0, 2}, {1, 0}};
	for (int t = 0; t < 2; ++t) {
		int ax = plane_order[c][t];
		double den;
		if (c == 2 && ax == 1)
			den = 4 * d[2] * d[1];
		else if (c == 1 && ax == 2)
			den = 4 * d[1] * d[2];
		else
			den = 4 * d[0] * d[ax == 0 ? c : ax];
		// corners in order (+,+), (-,-), (-,+), (+,-) over (row axis, plane axis);
		// the row axis is the first index of the reference's loops except for the
		// xy-plane of the Y row and the yz-plane of the Z row (same corners, same signs).
		int sgn[4][2] = {{1, 1}, {-1, -1}, {-1, 1}, {1, -1}};
		int axA = c, axB = ax;
		// The reference writes corner offsets as (x-ish first) in each block; the
		// resulting (offset, sign) set is symmetric so only the visiting order matters.
		if (c == 1 && ax == 0) {
			axA = 0;
			axB = 1;
		}
		if (c == 2 && ax == 0) {
			axA = 0;
			axB = 2;
		}
		if (c == 2 && ax == 1) {
			axA = 1;
			axB = 2;
		}
		for (int cr = 0; cr < 4; ++cr) {
			auto q = off2(axA, sgn[cr][0], axB, sgn[cr][1]);
			double sign = sgn[cr][0] * sgn[cr][1];
			for (int m = 0; m < 3; ++m) {
				double v = (sign > 0) ? C * M(q[0], q[1], q[2], ax, m) / den : -C * M(q[0], q[1], q[2], ax, m) / den;
				push(v, q[0], q[1], q[2], m);
			}
		}
	}
}

// multiplySpecialMatrixVector over interior rows (specialmath.cpp:202-252).
void Oracle::apply_operator(const std::vector<double>& v, std::vector<double>& out) const {
	out.assign(ncells() * 3, 0.0);
	std::vector<double> val;
	std::vector<int> idx;
	for (int i = kGhost; i <= N[0] - 3; ++i)
		for (int j = kGhost; j <= N[1] - 3; ++j)
			for (int k = kGhost; k <= N[2] - 3; ++k)
				for (int c = 0; c < 3; ++c) {
					row_entries(i, j, k, c, val, idx);
					double r = 0;
					for (size_t q = 0; q < val.size(); ++q) r += val[q] * v[idx[q]];
					out[lvec(i, j, k) * 3 + c] = r;
				}
}

// exchangeLargeVector one-rank periodic branches (mpi_util.cpp:5835-5907, :261-271, :335-346, ...).
void Oracle::exchange_large_vector(std::vector<double>& v) const {
	for (int ax = 0; ax < 3; ++ax) {
		int NN = N[ax];
		int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
		auto at = [&](int s, int a, int b) {
			int c[3];
			c[ax] = s;
			c[o1] = a;
			c[o2] = b;
			return lvec(c[0], c[1], c[2]) * 3;
		};
		for (int a = 0; a < N[o1]; ++a)  // send to left / front / bottom
			for (int b = 0; b < N[o2]; ++b)
				for (int i = 0; i <= 1; ++i)
					for (int l = 0; l < 3; ++l) v[at(NN - 2 + i, a, b) + l] = v[at(2 + i, a, b) + l];
		for (int a = 0; a < N[o1]; ++a)  // send to right / back / top
			for (int b = 0; b < N[o2]; ++b)
				for (int i = 0; i <= 1; ++i)
					for (int l = 0; l < 3; ++l) v[at(1 - i, a, b) + l] = v[at(NN - 3 - i, a, b) + l];
	}
}

// scalarMultiplyLargeVectors (specialmath.cpp:673-709): sequential i, j, k, l sum over interior.
double Oracle::dot_interior(const std::vector<double>& a, const std::vector<double>& b) const {
	double r = 0;
	for (int i = kGhost; i <= N[0] - 3; ++i)
		for (int j = kGhost; j <= N[1] - 3; ++j)
			for (int k = kGhost; k <= N[2] - 3; ++k)
				for (int l = 0; l < 3; ++l) {
					size_t q = lvec(i, j, k) * 3 + l;
					r += a[q] * b[q];
				}
	return r;
}

// generalizedMinimalResidualMethod + arnoldiIterations (specialmath.cpp:254-671):
// x0 = 0, no restart, modified Gram-Schmidt with ghost zeroing, dense Givens update of Q.
void Oracle::gmres(std::vector<double>& b, std::vector<double>& x, double precision, int max_iterations) {
	size_t len = ncells() * 3;
	x.assign(len, 0.0);
	exchange_large_vector(b);
	double norm = std::sqrt(dot_interior(b, b));
	gmres_iterations = 0;
	if (norm == 0) return;
	for (auto& q : b) q /= norm;
	double normRight = std::sqrt(dot_interior(b, b));
	int dim = 3 * n[0] * n[1] * n[2];
	std::vector<std::vector<double>> V;
	V.push_back(b);
	std::vector<std::vector<double>> H, Q, R, oldQ, oldR;
	std::vector<double> y;
	int nn = 2;
	double beta = 1.0;
	double rel = 1;
	double maxRel = precision / dim;
	auto interior = [&](int i, int j, int k) {
		return i >= kGhost && i <= N[0] - 3 && j >= kGhost && j <= N[1] - 3 && k >= kGhost && k <= N[2] - 3;
	};
	while (rel > std::max(maxRel, 1e-15) && nn < std::min(max_iterations, dim + 3)) {
		// arnoldiIterations (:254-365)
		std::vector<std::vector<double>> Hn(nn, std::vector<double>(nn - 1, 0.0));
		for (int r = 0; r < nn - 1; ++r)
			for (int c = 0; c < nn - 2; ++c) Hn[r][c] = H[r][c];
		std::vector<double> w;
		apply_operator(V[nn - 2], w);
		exchange_large_vector(w);
		for (int m = 0; m < nn - 1; ++m) {
			double h = dot_interior(V[m], w);
			Hn[m][nn - 2] = h;
			for (int i = 0; i < N[0]; ++i)
				for (int j = 0; j < N[1]; ++j)
					for (int k = 0; k < N[2]; ++k)
						for (int l = 0; l < 3; ++l) {
							size_t q = lvec(i, j, k) * 3 + l;
							if (interior(i, j, k))
								w[q] -= h * V[m][q];
							else
								w[q] = 0;
						}
		}
		Hn[nn - 1][nn - 2] = std::sqrt(dot_interior(w, w));
		if (Hn[nn - 1][nn - 2] > 0) {
			for (auto& q : w) q /= Hn[nn - 1][nn - 2];
		} else {
			for (auto& q : w) q = 0;
		}
		exchange_large_vector(w);
		V.push_back(w);
		H = Hn;
		// Givens update (:487-575)
		if (nn == 2) {
			double rho_ = H[0][0], sigma = H[1][0];
			double module = std::sqrt(rho_ * rho_ + sigma * sigma);
			double cs = rho_ / module, sn = sigma / module;
			Q = {{cs, sn}, {-sn, cs}};
			R = {{module}, {0}};
		} else {
			R.assign(nn, std::vector<double>(nn - 1, 0.0));
			for (int r = 0; r < nn; ++r)
				for (int c = 0; c < nn - 2; ++c) R[r][c] = (r < nn - 1) ? oldR[r][c] : 0;
			Q.assign(nn, std::vector<double>(nn, 0.0));
			for (int r = 0; r < nn - 1; ++r)
				for (int c = 0; c < nn - 1; ++c) Q[r][c] = oldQ[r][c];
			Q[nn - 1][nn - 1] = 1;
			for (int r = 0; r < nn; ++r) {
				R[r][nn - 2] = 0;
				for (int c = 0; c < nn; ++c) R[r][nn - 2] += Q[r][c] * H[c][nn - 2];
			}
			double rho_ = R[nn - 2][nn - 2], sigma = R[nn - 1][nn - 2];
			double module = std::sqrt(rho_ * rho_ + sigma * sigma);
			double cs = rho_ / module, sn = sigma / module;
			R[nn - 2][nn - 2] = module;
			R[nn - 1][nn - 2] = 0;
			for (int c = 0; c < nn - 1; ++c) {
				Q[nn - 2][c] = cs * oldQ[nn - 2][c];
				Q[nn - 1][c] = -sn * oldQ[nn - 2][c];
			}
			Q[nn - 2][nn - 1] = sn;
			Q[nn - 1][nn - 1] = cs;
		}
		y.assign(nn - 1, 0.0);
		for (int i = nn - 2; i >= 0; --i) {
			y[i] = beta * Q[i][0];
			for (int j = nn - 2; j > i; --j) y[i] -= R[i][j] * y[j];
			if (R[i][i] > 0)
				y[i] /= R[i][i];
			else
				y[i] = 0;
		}
		double error = std::fabs(beta * Q[nn - 1][0]);
		rel = error / normRight;
		oldQ = Q;
		oldR = R;
		nn++;
	}
	nn = nn - 1;
	gmres_iterations = nn - 1;
	for (size_t q = 0; q < len; ++q) {
		double s = 0;
		for (int m = 0; m < nn - 1; ++m) s += V[m][q] * y[m] * norm;
		x[q] = s;
	}
	exchange_large_vector(x);
	for (auto& q : b) q *= norm;
}

// evaluateElectricField, implicit branch (fields.cpp:152-376).
void Oracle::evaluate_electric_field() {
	build_rhs();
	std::vector<double> x;
	gmres(rhs, x, kMaxError, kMaxGmres);
	exchange_large_vector(x);  // fields.cpp:211
	for (int i = 0; i < N[0]; ++i)
		for (int j = 0; j < N[1]; ++j)
			for (int k = 0; k < N[2]; ++k)
				for (int l = 0; l < 3; ++l) Etemp[node(i, j, k) * 3 + l] = x[lvec(i, j, k) * 3 + l];
	for (size_t q = 0; q < Enew.size(); ++q) Enew[q] = (Etemp[q] - E[q] * (1 - theta)) / theta;
}

// evaluateRotEgeneral (fields.cpp:2862-2953).
void Oracle::rot_e(const std::vector<double>& Ef, int i, int j, int k, double out[3]) const {
	auto e = [&](int a, int b, int c, int l) { return Ef[node(a, b, c) * 3 + l]; };
	double rX[3], lX[3], rY[3], lY[3], rZ[3], lZ[3];
	for (int l = 0; l < 3; ++l) {
		rX[l] = (e(i + 1, j, k, l) + e(i + 1, j + 1, k, l) + e(i + 1, j, k + 1, l) + e(i + 1, j + 1, k + 1, l)) / 4.0;
		lX[l] = (e(i, j, k, l) + e(i, j + 1, k, l) + e(i, j, k + 1, l) + e(i, j + 1, k + 1, l)) / 4.0;
		rY[l] = (e(i, j + 1, k, l) + e(i + 1, j + 1, k, l) + e(i, j + 1, k + 1, l) + e(i + 1, j + 1, k + 1, l)) / 4.0;
		lY[l] = (e(i, j, k, l) + e(i + 1, j, k, l) + e(i, j, k + 1, l) + e(i + 1, j, k + 1, l)) / 4.0;
		rZ[l] = (e(i, j, k + 1, l) + e(i + 1, j, k + 1, l) + e(i, j + 1, k + 1, l) + e(i + 1, j + 1, k + 1, l)) / 4.0;
		lZ[l] = (e(i, j, k, l) + e(i + 1, j, k, l) + e(i, j + 1, k, l) + e(i + 1, j + 1, k, l)) / 4.0;
	}
	out[0] = (rY[2] - lY[2]) / d[1] - (rZ[1] - lZ[1]) / d[2];
	out[1] = (rZ[0] - lZ[0]) / d[2] - (rX[2] - lX[2]) / d[0];
	out[2] = (rX[1] - lX[1]) / d[0] - (rY[0] - lY[0]) / d[1];
}

// evaluateMagneticField (fields.cpp:2717-2773): newB = B - c' dt curl(E^{n+theta}).
void Oracle::evaluate_magnetic_field() {
	for (int i = 0; i < N[0]; ++i)
		for (int j = 0; j < N[1]; ++j)
			for (int k = 0; k < N[2]; ++k) {
				double r[3];
				rot_e(Etemp, i, j, k, r);
				size_t c = cell(i, j, k) * 3;
				for (int l = 0; l < 3; ++l) Bnew[c + l] = B[c + l] - r[l] * (kCprime * dt);
			}
}

// ---------------------------------------------------------------- mover
// Particle::evaluateVelocity (particle.cpp:366-379), with the nonrelativistic branch.
static inline void evaluate_velocity(const double p[3], double m, double v[3]) {
	double p2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
	if (p2 < kRelPrecision * m * m) {
		for (int a = 0; a < 3; ++a) v[a] = p[a] / m;
	} else {
		double g = std::sqrt(p2 + m * m) / m;
		for (int a = 0; a < 3; ++a) v[a] = p[a] / (m * g);
	}
}

static inline void cross(const double a[3], const double b[3], double out[3]) {
	// Vector3d::vectorMult (vector3d.cpp:82-88)
	out[0] = a[1] * b[2] - a[2] * b[1];
	out[1] = -a[0] * b[2] + a[2] * b[0];
	out[2] = a[0] * b[1] - a[1] * b[0];
}

// moveParticle(Particle*), periodic branch (mover.cpp:194-491). Returns true if escaped.
bool Oracle::move_one(Particle& p) {
	const Species& s = species[p.sp];
	update_maps(p);
	double Ef[3], Bf[3];
	gather_node(p, Etemp, Ef);
	gather_cell(p, B, Bf);
	double v[3], g;
	velocity_gamma(p, s.mass, v, g);
	double beta = 0.5 * s.charge * dt / s.mass;
	double oldP[3] = {p.p[0], p.p[1], p.p[2]};
	double newP[3] = {p.p[0], p.p[1], p.p[2]};
	double oldX[3] = {p.x[0], p.x[1], p.x[2]};
	double etaDt = eta * dt;
	double vxb[3];
	cross(v, Bf, vxb);
	for (int a = 0; a < 3; ++a) newP[a] += (Ef[a] + vxb[a]) * s.charge * dt;
	double newV[3];
	evaluate_velocity(newP, s.mass, newV);
	double mid[3];
	for (int a = 0; a < 3; ++a) mid[a] = v[a] * (1 - eta) + newV[a] * eta;
	for (int a = 0; a < 3; ++a) p.x[a] += mid[a] * etaDt;
	// velocityHat = (rotationTensor * gamma) * velocity
	double vhat[3];
	for (int a = 0; a < 3; ++a)
		vhat[a] = (p.alpha[a * 3 + 0] * g) * v[0] + (p.alpha[a * 3 + 1] * g) * v[1] + (p.alpha[a * 3 + 2] * g) * v[2];
	double prevP[3] = {p.p[0], p.p[1], p.p[2]};
	double pnorm = std::sqrt(oldP[0] * oldP[0] + oldP[1] * oldP[1] + oldP[2] * oldP[2]);
	auto dnorm = [](const double a[3], const double b[3]) {
		double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
		return std::sqrt(x * x + y * y + z * z);
	};
	double err = dnorm(prevP, newP);
	double prevX[3] = {p.x[0], p.x[1], p.x[2]};
	double newX[3] = {p.x[0], p.x[1], p.x[2]};
	double cerr = d[0];
	int it = 0;
	while (err > kVelErr * pnorm && cerr > kCoordErr * d[0] && it < kPicardIterations) {
		++it;
		for (int a = 0; a < 3; ++a) {
			prevP[a] = newP[a];
			prevX[a] = newX[a];
		}
		double rE[3];
		for (int a = 0; a < 3; ++a) rE[a] = p.alpha[a * 3 + 0] * Ef[0] + p.alpha[a * 3 + 1] * Ef[1] + p.alpha[a * 3 + 2] * Ef[2];
		for (int a = 0; a < 3; ++a) mid[a] = vhat[a] + rE[a] * beta;
		for (int a = 0; a < 3; ++a) p.x[a] = oldX[a] + (mid[a] * etaDt);
		for (int a = 0; a < 3; ++a) newX[a] = p.x[a];
		update_maps(p);
		gather_node(p, Etemp, Ef);
		gather_cell(p, B, Bf);
		double mxb[3];
		cross(mid, Bf, mxb);
		for (int a = 0; a < 3; ++a) newP[a] = oldP[a] + (Ef[a] + mxb[a]) * (s.charge * dt);
		err = dnorm(prevP, newP);
		cerr = dnorm(prevX, newX);
	}
	for (int a = 0; a < 3; ++a) p.p[a] = newP[a];
	for (int a = 0; a < 3; ++a) p.x[a] = oldX[a] + mid[a] * dt;
	// escape classification (mover.cpp:435-490), periodic single rank
	for (int a = 0; a < 3; ++a) {
		double lo = x0[a] + kGhost * d[a];
		double hi = x0[a] + (N[a] - 1 - 1) * d[a];
		if (p.x[a] < lo) {
			p.escaped = true;
			esc_[2 * a].push_back(p);
			return true;
		}
		if (p.x[a] > hi) {
			p.escaped = true;
			esc_[2 * a + 1].push_back(p);
			return true;
		}
	}
	return false;
}

// moveParticles (mover.cpp:21-55).
void Oracle::move_particles() {
	for (auto& l : esc_) l.clear();
	for (auto& p : particles) move_one(p);
}

// eraseEscapedPaticles (mover.cpp:167-192) + exchangeParticles one-rank periodic
// branch (mover.cpp:1002-1210): the order in which wrapped particles are re-appended
// is kept, since it sets the deposition summation order.
void Oracle::erase_and_exchange_particles() {
	std::vector<Particle> kept;
	kept.reserve(particles.size());
	for (auto& p : particles)
		if (!p.escaped) kept.push_back(p);
	particles.swap(kept);
	auto lo = [&](int a) { return x0[a] + kGhost * d[a]; };
	auto hi = [&](int a) { return x0[a] + (N[a] - 2) * d[a]; };
	for (int ax = 0; ax < 3; ++ax) {
		std::vector<Particle> tmp;
		for (auto& p : esc_[2 * ax]) {
			p.x[ax] += L[ax];
			p.escaped = false;
			tmp.push_back(p);
		}
		for (auto& p : esc_[2 * ax + 1]) {
			p.x[ax] -= L[ax];
			p.escaped = false;
			tmp.push_back(p);
		}
		esc_[2 * ax].clear();
		esc_[2 * ax + 1].clear();
		for (auto& p : tmp) {
			bool moved = false;
			for (int b = ax + 1; b < 3 && !moved; ++b) {
				if (p.x[b] < lo(b)) {
					p.escaped = true;
					esc_[2 * b].push_back(p);
					moved = true;
				} else if (p.x[b] > hi(b)) {
					p.escaped = true;
					esc_[2 * b + 1].push_back(p);
					moved = true;
				}
			}
			if (!moved) {
				p.escaped = false;
				particles.push_back(p);
			}
		}
	}
}

// updateDensityParameters (momentum_collecting.cpp:760-1020).
void Oracle::update_density_parameters() {
	for (size_t s = 0; s < species.size(); ++s) {
		std::fill(conc[s].begin(), conc[s].end(), 0.0);
		std::fill(energy[s].begin(), energy[s].end(), 0.0);
		std::fill(bulk[s].begin(), bulk[s].end(), 0.0);
	}
	std::fill(rho.begin(), rho.end(), 0.0);
	std::fill(rhom.begin(), rhom.end(), 0.0);
	for (auto& p : particles) {
		const Species& s = species[p.sp];
		double v[3], g;
		velocity_gamma(p, s.mass, v, g);
		double full = g * s.mass;
		double pc = s.charge * s.weight;
		for (int i = 0; i < 2; ++i)
			for (int j = 0; j < 2; ++j)
				for (int k = 0; k < 2; ++k) {
					size_t c = cell(p.ci[0][i], p.ci[1][j], p.ci[2][k]);
					double corr = p.cw[0][i] * p.cw[1][j] * p.cw[2][k] / volume;
					for (int a = 0; a < 3; ++a) bulk[p.sp][c * 3 + a] += p.p[a] * s.weight * corr;
					if (pc > 0)
						rho[c] += pc * corr;
					else
						rhom[c] += pc * corr;
					conc[p.sp][c] += s.weight * corr;
					energy[p.sp][c] += s.weight * full * corr;
				}
	}
	for (size_t s = 0; s < species.size(); ++s) {
		fold_cells(conc[s], 1, 4, 0, 3);
		fold_cells(energy[s], 1, 4, 0, 3);
	}
	fold_cells(rho, 1, 4, 0, 3);
	fold_cells(rhom, 1, 4, 0, 3);
	for (size_t s = 0; s < species.size(); ++s) fold_cells(bulk[s], 3, 4, 0, 3);
	for (size_t q = 0; q < rho.size(); ++q) rho[q] += rhom[q];
	for (size_t s = 0; s < species.size(); ++s)
		for (size_t q = 0; q < ncells(); ++q) {
			if (energy[s][q] > 0) {
				for (int a = 0; a < 3; ++a) bulk[s][q * 3 + a] = bulk[s][q * 3 + a] / energy[s][q];
			} else {
				for (int a = 0; a < 3; ++a) bulk[s][q * 3 + a] = 0;
			}
		}
}

// updateFields (fields.cpp:400-439).
void Oracle::update_fields() {
	E = Enew;
	B = Bnew;
}

// updateEnergy (simulation.cpp:1506-1640), scaleFactor = plasma_period = 1.
void Oracle::update_energy() {
	double we = 0, wb = 0, wp = 0;
	double mom[3] = {0, 0, 0};
	for (int i = kGhost; i < N[0] - 2; ++i)
		for (int j = kGhost; j < N[1] - 2; ++j)
			for (int k = kGhost; k < N[2] - 2; ++k) {
				const double* e = &E[node(i, j, k) * 3];
				we += (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) * volume / (8 * kPi);
			}
	for (int i = kGhost; i < N[0] - 2; ++i)
		for (int j = kGhost; j < N[1] - 2; ++j)
			for (int k = kGhost; k < N[2] - 2; ++k) {
				const double* b = &B[cell(i, j, k) * 3];
				wb += (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]) * volume / (8 * kPi);
			}
	for (int i = kGhost; i < N[0] - 2; ++i)
		for (int j = kGhost; j < N[1] - 2; ++j)
			for (int k = kGhost; k < N[2] - 2; ++k) {
				double e[3] = {0, 0, 0};
				for (int l = 0; l < 3; ++l)
					e[l] = (E[node(i, j, k) * 3 + l] + E[node(i, j, k + 1) * 3 + l] + E[node(i, j + 1, k) * 3 + l] +
					        E[node(i, j + 1, k + 1) * 3 + l] + E[node(i + 1, j, k) * 3 + l] +
					        E[node(i + 1, j, k + 1) * 3 + l] + E[node(i + 1, j + 1, k) * 3 + l] +
					        E[node(i + 1, j + 1, k + 1) * 3 + l]) / 8;
				double c[3];
				cross(e, &B[cell(i, j, k) * 3], c);
				for (int l = 0; l < 3; ++l) mom[l] += (c[l] / (4 * kPi)) * volume;
			}
	for (auto& p : particles) {
		const Species& s = species[p.sp];
		double p2 = p.p[0] * p.p[0] + p.p[1] * p.p[1] + p.p[2] * p.p[2];
		double g = std::sqrt(p2 / (s.mass * s.mass) + 1.0);
		wp += ((g - 1) * s.mass) * s.weight;
		for (int l = 0; l < 3; ++l) mom[l] += p.p[l] * s.weight;
	}
	energy_particles = wp;
	energy_e = we;
	energy_b = wb;
	for (int l = 0; l < 3; ++l) momentum[l] = mom[l];
}

void Oracle::initial_sequence() {
	exchange_node_field(E);
	exchange_node_field(Etemp);
	exchange_node_field(Enew);
	exchange_cell_field(B);
	exchange_cell_field(Bnew);
	update_particle_maps();
	evaluate_rotation_tensor();
	update_em_parameters();
	update_density_parameters();
	update_fields();
	update_energy();
}

void Oracle::step() {
	evaluate_rotation_tensor();
	update_em_parameters();
	evaluate_electric_field();
	exchange_node_field(E);
	exchange_node_field(Etemp);
	exchange_node_field(Enew);
	evaluate_magnetic_field();
	exchange_cell_field(Bnew);
	exchange_node_field(Etemp);
	exchange_node_field(Enew);
	move_particles();
	erase_and_exchange_particles();
	update_particle_maps();
	update_density_parameters();
	exchange_node_field(Enew);
	update_fields();
	exchange_node_field(Enew);
	exchange_node_field(E);
	exchange_cell_field(B);
	update_energy();
}

}  // namespace picoracle
