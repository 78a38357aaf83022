// C entry points of the FP64 oracle, for ctypes. TEST INFRASTRUCTURE ONLY.
#include <cstring>
#include <stdexcept>
#include <string>

#include "pic_oracle.hpp"

using picoracle::Oracle;
using picoracle::Particle;
using picoracle::Species;

extern "C" {

void* oracle_create(int nx, int ny, int nz, double dx, double dy, double dz, double dt, int nspecies,
                    const double* mass, const double* charge, const double* weight, const int* type) {
	std::vector<Species> sp(nspecies);
	for (int s = 0; s < nspecies; ++s) sp[s] = Species{mass[s], charge[s], weight[s], type[s]};
	try {
		return new Oracle(nx, ny, nz, dx, dy, dz, dt, sp);
	} catch (const std::exception&) {
		return nullptr;
	}
}

void oracle_destroy(void* h) { delete static_cast<Oracle*>(h); }

// Ghosted fields: E on (N+1)^3 nodes x 3, B on N^3 cells x 3 (component fastest).
void oracle_set_fields(void* h, const double* E, const double* B) {
	Oracle& o = *static_cast<Oracle*>(h);
	std::memcpy(o.E.data(), E, o.E.size() * 8);
	o.Etemp = o.E;
	o.Enew = o.E;
	std::memcpy(o.B.data(), B, o.B.size() * 8);
	o.Bnew = o.B;
}

void oracle_set_particles(void* h, long np, const int* sp, const int* id, const double* x, const double* p) {
	Oracle& o = *static_cast<Oracle*>(h);
	o.particles.resize(np);
	for (long q = 0; q < np; ++q) {
		Particle& a = o.particles[q];
		std::memset(&a, 0, sizeof(a));
		a.sp = sp[q];
		a.id = id[q];
		for (int c = 0; c < 3; ++c) {
			a.x[c] = x[q * 3 + c];
			a.p[c] = p[q * 3 + c];
		}
	}
}

long oracle_num_particles(void* h) { return (long)static_cast<Oracle*>(h)->particles.size(); }

// Particles in the oracle's current order; alpha is 9 doubles per particle.
void oracle_get_particles(void* h, int* sp, int* id, double* x, double* p, double* alpha) {
	Oracle& o = *static_cast<Oracle*>(h);
	for (size_t q = 0; q < o.particles.size(); ++q) {
		const Particle& a = o.particles[q];
		if (sp) sp[q] = a.sp;
		if (id) id[q] = a.id;
		for (int c = 0; c < 3; ++c) {
			if (x) x[q * 3 + c] = a.x[c];
			if (p) p[q * 3 + c] = a.p[c];
		}
		if (alpha)
			for (int c = 0; c < 9; ++c) alpha[q * 9 + c] = a.alpha[c];
	}
}

// Copy a named grid array out; returns its element count, or -1 for an unknown name.
long oracle_get_array(void* h, const char* name, double* out) {
	Oracle& o = *static_cast<Oracle*>(h);
	std::string n(name);
	const std::vector<double>* v = nullptr;
	if (n == "E") v = &o.E;
	else if (n == "Etheta") v = &o.Etemp;
	else if (n == "Enew") v = &o.Enew;
	else if (n == "B") v = &o.B;
	else if (n == "Bnew") v = &o.Bnew;
	else if (n == "J") v = &o.J;
	else if (n == "mu") v = &o.mu;
	else if (n == "divPi") v = &o.divPi;
	else if (n == "rhohat") v = &o.rhohat;
	else if (n == "rho") v = &o.rho;
	else if (n == "Pi") v = &o.Pi;
	else if (n == "rhs") v = &o.rhs;
	else if (n.rfind("conc", 0) == 0) v = &o.conc.at(std::stoi(n.substr(4)));
	else if (n.rfind("energy", 0) == 0) v = &o.energy.at(std::stoi(n.substr(6)));
	else if (n.rfind("bulkv", 0) == 0) v = &o.bulk.at(std::stoi(n.substr(5)));
	if (!v) return -1;
	if (out) std::memcpy(out, v->data(), v->size() * 8);
	return (long)v->size();
}

void oracle_get_energy(void* h, double* out7) {
	Oracle& o = *static_cast<Oracle*>(h);
	out7[0] = o.energy_particles;
	out7[1] = o.energy_e;
	out7[2] = o.energy_b;
	out7[3] = o.energy_particles + o.energy_e + o.energy_b;
	out7[4] = o.momentum[0];
	out7[5] = o.momentum[1];
	out7[6] = o.momentum[2];
}

int oracle_gmres_iterations(void* h) { return static_cast<Oracle*>(h)->gmres_iterations; }

// Phases, named after the reference Simulation methods they restate.
void oracle_initial_sequence(void* h) { static_cast<Oracle*>(h)->initial_sequence(); }
void oracle_step(void* h, int nsteps) {
	for (int s = 0; s < nsteps; ++s) static_cast<Oracle*>(h)->step();
}
void oracle_update_particle_maps(void* h) { static_cast<Oracle*>(h)->update_particle_maps(); }
void oracle_evaluate_rotation_tensor(void* h) { static_cast<Oracle*>(h)->evaluate_rotation_tensor(); }
void oracle_update_em_parameters(void* h) { static_cast<Oracle*>(h)->update_em_parameters(); }
void oracle_build_rhs(void* h) { static_cast<Oracle*>(h)->build_rhs(); }
void oracle_evaluate_electric_field(void* h) { static_cast<Oracle*>(h)->evaluate_electric_field(); }
void oracle_evaluate_magnetic_field(void* h) { static_cast<Oracle*>(h)->evaluate_magnetic_field(); }
void oracle_exchange_fields(void* h) {
	Oracle& o = *static_cast<Oracle*>(h);
	o.exchange_node_field(o.E);
	o.exchange_node_field(o.Etemp);
	o.exchange_node_field(o.Enew);
	o.exchange_cell_field(o.B);
	o.exchange_cell_field(o.Bnew);
}
void oracle_move_particles(void* h) {
	Oracle& o = *static_cast<Oracle*>(h);
	o.move_particles();
	o.erase_and_exchange_particles();
}
void oracle_update_density_parameters(void* h) {
	Oracle& o = *static_cast<Oracle*>(h);
	o.update_particle_maps();
	o.update_density_parameters();
}
void oracle_update_fields(void* h) { static_cast<Oracle*>(h)->update_fields(); }
void oracle_update_energy(void* h) { static_cast<Oracle*>(h)->update_energy(); }

// Operator on a GMRES-layout vector (N^3 x 3, ghosts periodic): out = A v over interior rows.
void oracle_apply_operator(void* h, const double* v, double* out) {
	Oracle& o = *static_cast<Oracle*>(h);
	std::vector<double> vin(v, v + o.ncells() * 3), vout;
	o.apply_operator(vin, vout);
	std::memcpy(out, vout.data(), vout.size() * 8);
}

// One row of the Maxwell matrix as (value, flat GMRES index) pairs; returns the entry count.
int oracle_row_entries(void* h, int i, int j, int k, int c, double* val, int* idx) {
	Oracle& o = *static_cast<Oracle*>(h);
	std::vector<double> v;
	std::vector<int> q;
	o.row_entries(i, j, k, c, v, q);
	for (size_t t = 0; t < v.size(); ++t) {
		val[t] = v[t];
		idx[t] = q[t];
	}
	return (int)v.size();
}

// Solve A x = b (GMRES-layout vectors) with the reference's GMRES; returns iterations.
int oracle_gmres(void* h, double* b, double* x, double precision, int max_iterations) {
	Oracle& o = *static_cast<Oracle*>(h);
	std::vector<double> vb(b, b + o.ncells() * 3), vx;
	o.gmres(vb, vx, precision, max_iterations);
	std::memcpy(x, vx.data(), vx.size() * 8);
	return o.gmres_iterations;
}

void oracle_set_mu(void* h, const double* mu) {
	Oracle& o = *static_cast<Oracle*>(h);
	std::memcpy(o.mu.data(), mu, o.mu.size() * 8);
}

double oracle_bspline(double x, double hw, double left, double right) { return Oracle::bspline(x, hw, left, right); }

// ---- explicit (BUNEMAN / Tristan) path. Yee arrays by name: Ex Ey Ez Bx By Bz Jx Jy Jz, each
// in the reference's own layout (dims from oracle_yee_dims).
static std::vector<double>* yee_array(Oracle& o, const std::string& n) {
	if (n.size() != 2) return nullptr;
	int c = n[1] - 'x';
	if (c < 0 || c > 2) return nullptr;
	if (n[0] == 'E') return &o.yE[c];
	if (n[0] == 'B') return &o.yB[c];
	if (n[0] == 'J') return &o.yJ[c];
	return nullptr;
}
void oracle_yee_init(void* h) { static_cast<Oracle*>(h)->yee_init(); }
void oracle_yee_dims(void* h, const char* name, int* dims) {
	Oracle& o = *static_cast<Oracle*>(h);
	o.yee_dims(name[0] == 'B', name[1] - 'x', dims);
}
long oracle_yee_get(void* h, const char* name, double* out) {
	Oracle& o = *static_cast<Oracle*>(h);
	std::vector<double>* v = yee_array(o, name);
	if (!v) return -1;
	if (out) std::memcpy(out, v->data(), v->size() * 8);
	return (long)v->size();
}
long oracle_yee_set(void* h, const char* name, const double* in) {
	Oracle& o = *static_cast<Oracle*>(h);
	std::vector<double>* v = yee_array(o, name);
	if (!v) return -1;
	std::memcpy(v->data(), in, v->size() * 8);
	return (long)v->size();
}
// exchangeBunemanBfield + exchangeBunemanEfield (the pre-loop exchanges, PIC++/simulation.cpp:83-85)
void oracle_yee_exchange(void* h) {
	Oracle& o = *static_cast<Oracle*>(h);
	for (int c = 0; c < 3; ++c) {
		int d[3];
		o.yee_dims(true, c, d);
		o.exchange_yee(o.yB[c], d);
		o.yee_dims(false, c, d);
		o.exchange_yee(o.yE[c], d);
	}
}
void oracle_explicit_step(void* h, int nsteps) {
	for (int s = 0; s < nsteps; ++s) static_cast<Oracle*>(h)->explicit_step();
}
void oracle_tristan_b_half(void* h) { static_cast<Oracle*>(h)->tristan_b_half(); }
void oracle_tristan_move(void* h) { static_cast<Oracle*>(h)->move_particles_tristan(); }
void oracle_tristan_flux(void* h) { static_cast<Oracle*>(h)->tristan_flux(); }
void oracle_tristan_e(void* h) { static_cast<Oracle*>(h)->tristan_e(); }
void oracle_erase_and_exchange(void* h) { static_cast<Oracle*>(h)->erase_and_exchange_particles(); }
void oracle_interp_yee(void* h, const double* x, double* E, double* B) { static_cast<Oracle*>(h)->interp_yee(x, E, B); }

void oracle_alpha(double beta, const double* v, double gamma, const double* E, const double* B, double* out9) {
	Oracle::alpha_tensor(beta, v, gamma, E, B, out9);
}

}  // extern "C"
