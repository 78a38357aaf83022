// FP64 CPU restatement of the PIC++ explicit (BUNEMAN / Tristan) step: Yee-grid field
// updates, the Boris-type mover and the zig-zag current deposit.
//
// TEST INFRASTRUCTURE ONLY: the checker of the HIP explicit path, loaded by tests/ (through
// oracle_capi.cpp); the product path never calls it. Periodic, single rank, 3-D,
// additionalBinNumber = 1, smoothingCount = 0.
#include <algorithm>
#include <cmath>

#include "pic_oracle.hpp"

namespace picoracle {

namespace {
const double kPiE = 4 * std::atan2(1.0, 1.0);  // PIC++/constants.h:28
const double kFourPi = 4 * kPiE;                // constants.h:29
const double kCprimeE = 1.025;                  // speed_of_light_correction, constants.h:45
}  // namespace

// E, J component c: N cells along axis c, N + 1 nodes along the others; B component c: N + 1
// along axis c, N along the others (PIC++/initialization.cpp:7146-7174).
void Oracle::yee_dims(bool b_field, int c, int dims[3]) const {
	for (int a = 0; a < 3; ++a) dims[a] = ((a == c) != b_field) ? N[a] : N[a] + 1;
}

void Oracle::yee_init() {
	for (int c = 0; c < 3; ++c) {
		int de[3], db[3];
		yee_dims(false, c, de);
		yee_dims(true, c, db);
		yE[c].assign((size_t)de[0] * de[1] * de[2], 0.0);
		yJ[c].assign(yE[c].size(), 0.0);
		yB[c].assign((size_t)db[0] * db[1] * db[2], 0.0);
	}
}

// exchangeBuneman{Ex,Ey,Ez,Bx,By,Bz}Along{X,Y,Z}, one-rank periodic branches
// (PIC++/fields_exchange.cpp:2523-4200), axes in the order X, Y, Z over the whole array.
// Along an axis of N points (cells): f[N-2+i] = f[2+i], f[i] = f[N-4+i], i = 0, 1. Along an
// axis of N + 1 points (nodes): f[N-1+i] = f[3+i], f[i] = f[N-4+i], then f[N-2] = f[2].
void Oracle::exchange_yee(std::vector<double>& f, const int dims[3]) const {
	for (int ax = 0; ax < 3; ++ax) {
		const int L = dims[ax], Nn = N[ax];
		const bool node = L == Nn + 1;
		const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
		int idx[3];
		auto at = [&](int t) -> double& {
			idx[ax] = t;
			return f[yidx(dims, idx[0], idx[1], idx[2])];
		};
		for (int u = 0; u < dims[o1]; ++u)
			for (int v = 0; v < dims[o2]; ++v) {
				idx[o1] = u;
				idx[o2] = v;
				for (int i = 0; i <= 1; ++i) {
					if (node) {
						at(Nn - 1 + i) = at(3 + i);
					} else {
						at(Nn - 2 + i) = at(2 + i);
					}
					at(i) = at(Nn - 4 + i);
				}
				if (node) at(Nn - 2) = at(2);
			}
	}
}

// sumBunemanJ{x,y,z}Along{X,Y,Z}, one-rank periodic branches (PIC++/momentum_exchange.cpp:
// 1806-2310): along an axis of N points, f[N-1-i] += f[3-i], f[3-i] = f[N-1-i], i = 0..3;
// along an axis of N + 1 points, f[N-i] += f[4-i], f[4-i] = f[N-i], i = 0..4.
void Oracle::fold_yee(std::vector<double>& f, const int dims[3]) const {
	for (int ax = 0; ax < 3; ++ax) {
		const int L = dims[ax], Nn = N[ax];
		const bool node = L == Nn + 1;
		const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
		int idx[3];
		auto at = [&](int t) -> double& {
			idx[ax] = t;
			return f[yidx(dims, idx[0], idx[1], idx[2])];
		};
		// the reference's loops run X: j, k; Y: i, k; Z: i, j (outer to inner); the pairs are
		// disjoint, so the loop order does not change the sums
		for (int u = 0; u < dims[o1]; ++u)
			for (int v = 0; v < dims[o2]; ++v) {
				idx[o1] = u;
				idx[o2] = v;
				if (node) {
					for (int i = 0; i <= 4; ++i) {
						at(Nn - i) += at(4 - i);
						at(4 - i) = at(Nn - i);
					}
				} else {
					for (int i = 0; i <= 3; ++i) {
						at(Nn - 1 - i) += at(3 - i);
						at(3 - i) = at(Nn - 1 - i);
					}
				}
			}
	}
}

// tristanEvaluateBhalfStep (PIC++/fields.cpp:21-79) with evaluateBunemanRotE{x,y,z}
// (fields.cpp:3273-3304): B -= c' dt / 2 curl E on the owned ranges.
void Oracle::tristan_b_half() {
	const int lo = 2;
	int dB[3][3], dE[3][3];
	for (int c = 0; c < 3; ++c) {
		yee_dims(true, c, dB[c]);
		yee_dims(false, c, dE[c]);
	}
	const double cdtd2 = kCprimeE * dt / 2.0;
	auto Ex = [&](int i, int j, int k) { return yE[0][yidx(dE[0], i, j, k)]; };
	auto Ey = [&](int i, int j, int k) { return yE[1][yidx(dE[1], i, j, k)]; };
	auto Ez = [&](int i, int j, int k) { return yE[2][yidx(dE[2], i, j, k)]; };
	const int hx = N[0] - 2, hy = N[1] - 2, hz = N[2] - 2;
	for (int i = lo; i <= hx; ++i)
		for (int j = lo; j < hy; ++j)
			for (int k = lo; k < hz; ++k) {
				double r = 0;
				r += (Ez(i, j + 1, k) - Ez(i, j, k)) / d[1];
				r -= (Ey(i, j, k + 1) - Ey(i, j, k)) / d[2];
				double& b = yB[0][yidx(dB[0], i, j, k)];
				b = b - r * cdtd2;
			}
	for (int i = lo; i < hx; ++i)
		for (int j = lo; j <= hy; ++j)
			for (int k = lo; k < hz; ++k) {
				double r = -((Ez(i + 1, j, k) - Ez(i, j, k)) / d[0]);
				r += (Ex(i, j, k + 1) - Ex(i, j, k)) / d[2];
				double& b = yB[1][yidx(dB[1], i, j, k)];
				b = b - r * cdtd2;
			}
	for (int i = lo; i < hx; ++i)
		for (int j = lo; j < hy; ++j)
			for (int k = lo; k <= hz; ++k) {
				double r = (Ey(i + 1, j, k) - Ey(i, j, k)) / d[0];
				r -= (Ex(i, j + 1, k) - Ex(i, j, k)) / d[1];
				double& b = yB[2][yidx(dB[2], i, j, k)];
				b = b - r * cdtd2;
			}
}

// tristanEvaluateE (PIC++/fields.cpp:81-150) with evaluateBunemanRotB{x,y,z}
// (fields.cpp:3306-3338): E += (c' curl B - 4 pi J) dt on the owned ranges, periodic x.
void Oracle::tristan_e() {
	const int lo = 2;
	int dB[3][3], dE[3][3];
	for (int c = 0; c < 3; ++c) {
		yee_dims(true, c, dB[c]);
		yee_dims(false, c, dE[c]);
	}
	auto Bx = [&](int i, int j, int k) { return yB[0][yidx(dB[0], i, j, k)]; };
	auto By = [&](int i, int j, int k) { return yB[1][yidx(dB[1], i, j, k)]; };
	auto Bz = [&](int i, int j, int k) { return yB[2][yidx(dB[2], i, j, k)]; };
	const int hx = N[0] - 2, hy = N[1] - 2, hz = N[2] - 2;
	for (int i = lo; i < hx; ++i)
		for (int j = lo; j <= hy; ++j)
			for (int k = lo; k <= hz; ++k) {
				double r = 0;
				r += (Bz(i, j, k) - Bz(i, j - 1, k)) / d[1];
				r -= (By(i, j, k) - By(i, j, k - 1)) / d[2];
				const size_t q = yidx(dE[0], i, j, k);
				yE[0][q] = yE[0][q] + (kCprimeE * r - kFourPi * yJ[0][q]) * dt;
			}
	for (int i = lo; i <= hx; ++i)
		for (int j = lo; j < hy; ++j)
			for (int k = lo; k <= hz; ++k) {
				double r = -((Bz(i, j, k) - Bz(i - 1, j, k)) / d[0]);
				r += (Bx(i, j, k) - Bx(i, j, k - 1)) / d[2];
				const size_t q = yidx(dE[1], i, j, k);
				yE[1][q] = yE[1][q] + (kCprimeE * r - kFourPi * yJ[1][q]) * dt;
			}
	for (int i = lo; i <= hx; ++i)
		for (int j = lo; j <= hy; ++j)
			for (int k = lo; k < hz; ++k) {
				double r = (By(i, j, k) - By(i - 1, j, k)) / d[0];
				r -= (Bx(i, j, k) - Bx(i, j - 1, k)) / d[1];
				const size_t q = yidx(dE[2], i, j, k);
				yE[2][q] = yE[2][q] + (kCprimeE * r - kFourPi * yJ[2][q]) * dt;
			}
}

// correlationBunemanEBfieldsWithoutMaps, 3-D branch (PIC++/correlation.cpp:441-626): per axis
// a cell map (left and right cell around the point, weights from the cell centres) and a
// node map (left and right node); each component is interpolated with the map of its own
// staggering, summed over (x, y, z) in {left, right}^3 with z fastest.
void Oracle::interp_yee(const double x[3], double Ef[3], double Bf[3]) const {
	int lc[3], rc[3], ln[3], rn[3];
	double wlc[3], wrc[3], wln[3], wrn[3];
	for (int a = 0; a < 3; ++a) {
		auto grid = [&](int i) { return x0[a] + i * d[a]; };
		auto mid = [&](int i) { return (grid(i) + grid(i + 1)) / 2; };
		int cnt = (int)std::floor((x[a] - grid(0)) / d[a]);
		if (x[a] < mid(cnt)) {
			lc[a] = cnt - 1;
			rc[a] = cnt;
			wrc[a] = 0.5 + (x[a] - grid(cnt)) / d[a];
			wlc[a] = 1.0 - wrc[a];
		} else {
			lc[a] = cnt;
			rc[a] = cnt + 1;
			wlc[a] = 0.5 + (grid(cnt + 1) - x[a]) / d[a];
			wrc[a] = 1.0 - wlc[a];
		}
		cnt = (int)std::floor(((x[a] - grid(0)) / d[a]) + 0.5);
		if (x[a] < grid(cnt)) {
			ln[a] = cnt - 1;
			rn[a] = cnt;
			wrn[a] = 0.5 + (x[a] - mid(cnt - 1)) / d[a];
			wln[a] = 1.0 - wrn[a];
		} else {
			ln[a] = cnt;
			rn[a] = cnt + 1;
			wln[a] = 0.5 + (mid(cnt) - x[a]) / d[a];
			wrn[a] = 1.0 - wln[a];
		}
	}
	int dE[3][3], dB[3][3];
	for (int c = 0; c < 3; ++c) {
		yee_dims(false, c, dE[c]);
		yee_dims(true, c, dB[c]);
	}
	for (int c = 0; c < 3; ++c) Ef[c] = Bf[c] = 0.0;
	for (int sx = 0; sx < 2; ++sx)
		for (int sy = 0; sy < 2; ++sy)
			for (int sz = 0; sz < 2; ++sz) {
				const int s[3] = {sx, sy, sz};
				int ci[3], ni[3];
				double cw[3], nw[3];
				for (int a = 0; a < 3; ++a) {
					ci[a] = s[a] ? rc[a] : lc[a];
					ni[a] = s[a] ? rn[a] : ln[a];
					cw[a] = s[a] ? wrc[a] : wlc[a];
					nw[a] = s[a] ? wrn[a] : wln[a];
				}
				// E component c: cell map along c, node maps along the others; B the opposite
				for (int c = 0; c < 3; ++c) {
					int ie[3], ib[3];
					double we[3], wb[3];
					for (int a = 0; a < 3; ++a) {
						ie[a] = a == c ? ci[a] : ni[a];
						we[a] = a == c ? cw[a] : nw[a];
						ib[a] = a == c ? ni[a] : ci[a];
						wb[a] = a == c ? nw[a] : cw[a];
					}
					Ef[c] += yE[c][yidx(dE[c], ie[0], ie[1], ie[2])] * we[0] * we[1] * we[2];
					Bf[c] += yB[c][yidx(dB[c], ib[0], ib[1], ib[2])] * wb[0] * wb[1] * wb[2];
				}
			}
}

// moveParticleTristan (PIC++/mover.cpp:718-808): half electric kick, magnetic rotation,
// half kick, then x += v dt; escape classification as sortParticleToEscaped (periodic).
bool Oracle::move_tristan(Particle& p) {
	const Species& sp = species[p.sp];
	const double m = sp.mass, q = sp.charge;
	double Ef[3], Bf[3];
	interp_yee(p.x, Ef, Bf);
	double dp[3];
	const double s = q * dt * 0.5;
	for (int a = 0; a < 3; ++a) dp[a] = Ef[a] * s;
	for (int a = 0; a < 3; ++a) p.p[a] = p.p[a] + dp[a];
	const double p2 = p.p[0] * p.p[0] + p.p[1] * p.p[1] + p.p[2] * p.p[2];
	const double gamma = std::sqrt(p2 / (m * m) + 1.0);
	const double beta = q * dt / (2.0 * m);  // Particle::beta (particle.cpp:47)
	const double bs = beta / gamma;
	const double f = 2.0 / std::sqrt(1.0 + bs * bs * (Bf[0] * Bf[0] + Bf[1] * Bf[1] + Bf[2] * Bf[2]));
	double mom[3] = {p.p[0], p.p[1], p.p[2]};
	// Vector3d::vectorMult (vector3d.cpp:82-88): a x b
	auto cross = [](const double a[3], const double b[3], double out[3]) {
		out[0] = a[1] * b[2] - a[2] * b[1];
		out[1] = -a[0] * b[2] + a[2] * b[0];
		out[2] = a[0] * b[1] - a[1] * b[0];
	};
	double c1[3], tmp[3], c2[3];
	cross(mom, Bf, c1);
	for (int a = 0; a < 3; ++a) tmp[a] = mom[a] + c1[a] * bs;
	cross(tmp, Bf, c2);
	const double fb = f * bs;
	for (int a = 0; a < 3; ++a) mom[a] += c2[a] * fb + dp[a];
	for (int a = 0; a < 3; ++a) p.p[a] = mom[a];
	const double p2n = mom[0] * mom[0] + mom[1] * mom[1] + mom[2] * mom[2];
	const double gn = std::sqrt(p2n / (m * m) + 1.0);
	for (int a = 0; a < 3; ++a) p.x[a] += (mom[a] / (m * gn)) * dt;
	for (int a = 0; a < 3; ++a) {
		const double lo = x0[a] + kGhost * d[a];
		const double hi = x0[a] + (N[a] - 2) * d[a];
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

void Oracle::move_particles_tristan() {
	for (auto& l : esc_) l.clear();
	for (auto& p : particles) move_tristan(p);
}

// addParticleFluxZigzag, 3-D branch (PIC++/momentum_collecting.cpp:85-249): the straight
// path from x - v dt to x is split at the cell boundary it crosses (Umeda's zig-zag), each
// part deposits its flux with linear weights at its midpoint.
void Oracle::zigzag(const Particle& p) {
	const Species& sp = species[p.sp];
	const double m = sp.mass;
	const double p2 = p.p[0] * p.p[0] + p.p[1] * p.p[1] + p.p[2] * p.p[2];
	const double g = std::sqrt(p2 / (m * m) + 1.0);
	double v[3];
	for (int a = 0; a < 3; ++a) v[a] = p.p[a] / (m * g);
	const double fcd = sp.charge * sp.weight / volume;
	double loc[3], prev[3], r[3];
	int cur[3], pre[3];
	for (int a = 0; a < 3; ++a) {
		loc[a] = p.x[a] - x0[a];
		cur[a] = (int)std::floor(loc[a] / d[a]);
		prev[a] = loc[a] - v[a] * dt;
		pre[a] = (int)std::floor(prev[a] / d[a]);
		r[a] = (loc[a] + prev[a]) / 2;
		if (cur[a] != pre[a]) {
			const double gp = x0[a] + pre[a] * d[a], gc = x0[a] + cur[a] * d[a];
			r[a] = std::max(gp, gc) - x0[a];
		}
	}
	double F1[3], F2[3], W1[3], W2[3];
	for (int a = 0; a < 3; ++a) F1[a] = fcd * (r[a] - prev[a]) / dt;
	F2[0] = fcd * (loc[0] - r[0]) / dt;
	F2[1] = fcd * v[1] - F1[1];
	F2[2] = fcd * v[2] - F1[2];
	for (int a = 0; a < 3; ++a) {
		W1[a] = (prev[a] + r[a]) / (2 * d[a]) - pre[a];
		W2[a] = (loc[a] + r[a]) / (2 * d[a]) - cur[a];
	}
	int dJ[3][3];
	for (int c = 0; c < 3; ++c) yee_dims(false, c, dJ[c]);
	auto Jx = [&](int i, int j, int k) -> double& { return yJ[0][yidx(dJ[0], i, j, k)]; };
	auto Jy = [&](int i, int j, int k) -> double& { return yJ[1][yidx(dJ[1], i, j, k)]; };
	auto Jz = [&](int i, int j, int k) -> double& { return yJ[2][yidx(dJ[2], i, j, k)]; };
	for (int part = 0; part < 2; ++part) {
		const int* c = part ? cur : pre;
		const double* F = part ? F2 : F1;
		const double* W = part ? W2 : W1;
		const double om[3] = {1.0 - W[0], 1.0 - W[1], 1.0 - W[2]};
		const int i = c[0], j = c[1], k = c[2];
		Jx(i, j, k) += F[0] * om[1] * om[2];
		Jx(i, j + 1, k) += F[0] * W[1] * om[2];
		Jx(i, j, k + 1) += F[0] * om[1] * W[2];
		Jx(i, j + 1, k + 1) += F[0] * W[1] * W[2];
		Jy(i, j, k) += F[1] * om[0] * om[2];
		Jy(i + 1, j, k) += F[1] * W[0] * om[2];
		Jy(i, j, k + 1) += F[1] * om[0] * W[2];
		Jy(i + 1, j, k + 1) += F[1] * W[0] * W[2];
		Jz(i, j, k) += F[2] * om[0] * om[1];
		Jz(i, j + 1, k) += F[2] * om[0] * W[1];
		Jz(i + 1, j, k) += F[2] * W[0] * om[1];
		Jz(i + 1, j + 1, k) += F[2] * W[0] * W[1];
	}
}

// tristanUpdateFlux (PIC++/momentum_collecting.cpp:20-83): zero J, deposit every particle
// of the vector (escaped ones too, at their unwrapped positions), fold the ghost planes.
void Oracle::tristan_flux() {
	for (int c = 0; c < 3; ++c) std::fill(yJ[c].begin(), yJ[c].end(), 0.0);
	for (const auto& p : particles) zigzag(p);
	for (int c = 0; c < 3; ++c) {
		int dJ[3];
		yee_dims(false, c, dJ);
		fold_yee(yJ[c], dJ);
	}
}

// The loop body of the BUNEMAN solver (PIC++/simulation.cpp:163-197, smoothingCount = 0),
// then eraseEscapedPaticles and exchangeParticles (:258-263).
void Oracle::explicit_step() {
	int dE[3][3], dB[3][3];
	for (int c = 0; c < 3; ++c) {
		yee_dims(false, c, dE[c]);
		yee_dims(true, c, dB[c]);
	}
	tristan_b_half();
	for (int c = 0; c < 3; ++c) exchange_yee(yB[c], dB[c]);
	move_particles_tristan();
	tristan_b_half();
	for (int c = 0; c < 3; ++c) exchange_yee(yB[c], dB[c]);
	tristan_flux();
	tristan_e();
	for (int c = 0; c < 3; ++c) exchange_yee(yE[c], dE[c]);
	erase_and_exchange_particles();
}

}  // namespace picoracle
