// FP64 CPU restatement of the PIC++ implicit step (oracle).
//
// TEST INFRASTRUCTURE ONLY. This file and everything under oracle/ is the
// checker for the HIP product path: only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it. The product path never calls it.
//
// It restates, in plain C++ on ghosted arrays laid out exactly as the
// reference lays them out, the algorithm of one periodic, single-rank
// implicit step of PIC++ (3-D, splineOrder = 0, additionalBinNumber = 1),
// including the reference's quirks (see each function's citation).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace picoracle {

constexpr int kGhost = 2;  // 1 + additionalBinNumber (PIC++/constants.h:61-62)

struct Species {
	double mass, charge, weight;
	int type;  // PIC++ ParticleTypes index (particleTypes.h:4)
};

struct Particle {
	double x[3];
	double p[3];
	double alpha[9];
	int sp;
	int id;
	bool escaped;
	// CorrelationMap (PIC++/particle.h:45-55): 2 indices + 2 weights per axis.
	int ni[3][2], ci[3][2];
	double nw[3][2], cw[3][2];
};

class Oracle {
public:
	Oracle(int nx, int ny, int nz, double dx, double dy, double dz, double dt, const std::vector<Species>& species);

	// geometry
	int n[3];       // physical cells per axis
	int N[3];       // ghosted cells per axis (n + 4)
	double d[3];    // cell size
	double x0[3];   // coordinate of node 0 (PIC++/initialization.cpp:194, :1988)
	double L[3];    // box size
	double dt, theta, eta, cprime, volume;
	std::vector<Species> species;

	// ghosted grid arrays, component-fastest
	std::vector<double> E, Etemp, Enew;        // nodes (N+1)^3 x 3
	std::vector<double> J, Jm, divPi, divPim;  // nodes x 3
	std::vector<double> mu;                    // nodes x 9
	std::vector<double> B, Bnew;               // cells N^3 x 3
	std::vector<double> rhohat, rhom, rho;     // cells
	std::vector<double> Pi;                    // cells x 9
	std::vector<std::vector<double>> conc, energy, bulk;  // per species, cells (bulk x 3)
	std::vector<double> rhs;                   // N^3 x 3 (Maxwell right part)
	std::vector<Particle> particles;

	int gmres_iterations = 0;
	double energy_particles = 0, energy_e = 0, energy_b = 0, momentum[3] = {0, 0, 0};

	size_t node(int i, int j, int k) const { return ((size_t)i * (N[1] + 1) + j) * (N[2] + 1) + k; }
	size_t cell(int i, int j, int k) const { return ((size_t)i * N[1] + j) * N[2] + k; }
	size_t nnodes() const { return (size_t)(N[0] + 1) * (N[1] + 1) * (N[2] + 1); }
	size_t ncells() const { return (size_t)N[0] * N[1] * N[2]; }

	// shape functions (PIC++/correlation.cpp:798-850, :1112-1452)
	static double bspline(double x, double hw, double left, double right);
	void update_maps(Particle& p) const;
	void gather_node(const Particle& p, const std::vector<double>& f, double out[3]) const;
	void gather_cell(const Particle& p, const std::vector<double>& f, double out[3]) const;
	static void alpha_tensor(double beta, const double v[3], double gamma, const double Ef[3], const double Bf[3],
	                         double out[9]);

	// Simulation-level phases (names follow the reference methods)
	void exchange_node_field(std::vector<double>& f) const;  // exchangeGeneralEfield
	void exchange_cell_field(std::vector<double>& f) const;  // exchangeGeneralBfield
	void update_particle_maps();                              // updateParticleCorrelationMaps
	void evaluate_rotation_tensor();                          // evaluateParticlesRotationTensor
	void update_em_parameters();                              // updateElectroMagneticParameters
	void build_rhs();                                         // createInternalEquation (right part)
	void row_entries(int i, int j, int k, int c, std::vector<double>& val, std::vector<int>& idx) const;
	void apply_operator(const std::vector<double>& v, std::vector<double>& out) const;
	void gmres(std::vector<double>& b, std::vector<double>& x, double precision, int max_iterations);
	void evaluate_electric_field();                           // evaluateElectricField
	void evaluate_magnetic_field();                           // evaluateMagneticField
	void move_particles();                                    // moveParticles
	void erase_and_exchange_particles();                      // eraseEscapedPaticles + exchangeParticles
	void update_density_parameters();                         // updateDensityParameters
	void update_fields();                                     // updateFields
	void update_energy();                                     // updateEnergy

	void initial_sequence();  // PIC++/simulation.cpp:86-123
	void step();              // PIC++/simulation.cpp:205-390

	// ---- the explicit (BUNEMAN / Tristan) path, explicit_oracle.cpp ----
	// Yee ("Buneman") grid, ghosted, laid out as the reference allocates it
	// (PIC++/initialization.cpp:7145-7175): E and J component c have N cells along axis c and
	// N + 1 along the others; B component c has N + 1 along axis c and N along the others.
	std::vector<double> yE[3], yB[3], yJ[3];
	void yee_dims(bool b_field, int c, int dims[3]) const;
	size_t yidx(const int dims[3], int i, int j, int k) const { return ((size_t)i * dims[1] + j) * dims[2] + k; }
	void yee_init();                                          // allocate the Yee arrays (zero)
	void exchange_yee(std::vector<double>& f, const int dims[3]) const;  // exchangeBuneman{E,B}field
	void fold_yee(std::vector<double>& f, const int dims[3]) const;      // exchangeBunemanFlux
	void tristan_b_half();                                    // tristanEvaluateBhalfStep
	void tristan_e();                                         // tristanEvaluateE
	void interp_yee(const double x[3], double Ef[3], double Bf[3]) const;  // correlationBunemanEBfieldsWithoutMaps
	void move_particles_tristan();                            // moveParticles, BUNEMAN branch
	void tristan_flux();                                      // tristanUpdateFlux
	void explicit_step();                                     // PIC++/simulation.cpp:163-197, then erase/exchange

	size_t lvec(int i, int j, int k) const { return ((size_t)i * N[1] + j) * N[2] + k; }  // GMRES vector cell index

private:
	std::vector<Particle> esc_[6];  // left, right, front, back, bottom, top
	void fold_nodes(std::vector<double>& f, int ncomp) const;
	void fold_cells(std::vector<double>& f, int ncomp, int pairs, int hi_off, int lo_base) const;
	void exchange_large_vector(std::vector<double>& v) const;
	double dot_interior(const std::vector<double>& a, const std::vector<double>& b) const;
	void div_pressure(int i, int j, int k, double out[3]) const;
	double div_flux(int i, int j, int k) const;
	void rot_b(int i, int j, int k, double out[3]) const;
	void grad_rhohat(int i, int j, int k, double out[3]) const;
	void rot_e(const std::vector<double>& Ef, int i, int j, int k, double out[3]) const;
	bool move_one(Particle& p);
	bool move_tristan(Particle& p);
	void zigzag(const Particle& p);
};

}  // namespace picoracle
