// Golden-vector harness for the reference PIC++ explicit (BUNEMAN / Tristan) step.
//
// Test infrastructure only: links against the unmodified reference objects (compiled from
// PIC++/*.cpp by oracle/Makefile.ref) and drives the public Simulation methods in the order
// of the reference step loop's BUNEMAN branch (PIC++/simulation.cpp:83-85 pre-loop exchanges,
// :163-197 loop body with smoothingCount = 0, :258-263 erase and exchange, :388 clearing the
// escape lists), dumping the Yee fields, the Yee current and the particles after every step.
//
// Input (binary, little endian): nx ny nz (int32), dx dt (f64), nsteps nspecies (int32), per
// species type (int32) mass charge weight (f64), np (int64), the Yee arrays Ex Ey Ez Bx By Bz
// in the reference's layouts (f64, see oracle/explicit_oracle.cpp), then np particles:
// species (int32), x y z px py pz (f64).
// Output: named float64 arrays, as oracle/ref_harness.cpp writes them.
#include <mpi.h>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "particle.h"
#include "simulation.h"
#include "boundaryFieldEvaluator.h"
#include "vector3d.h"

namespace {

FILE* g_out = nullptr;

void dump(const std::string& name, const std::vector<double>& data, std::vector<uint64_t> dims) {
	uint32_t len = (uint32_t)name.size();
	fwrite(&len, 4, 1, g_out);
	fwrite(name.data(), 1, len, g_out);
	uint32_t nd = (uint32_t)dims.size();
	fwrite(&nd, 4, 1, g_out);
	fwrite(dims.data(), 8, nd, g_out);
	fwrite(data.data(), 8, data.size(), g_out);
}

std::string tag(const char* base, int step) {
	char buf[128];
	snprintf(buf, sizeof(buf), "s%d/%s", step, base);
	return std::string(buf);
}

template <class T>
T rd(FILE* f) {
	T v;
	if (fread(&v, sizeof(T), 1, f) != 1) {
		fprintf(stderr, "short input\n");
		exit(2);
	}
	return v;
}

struct Yee {
	double*** a;
	int d[3];
};

void read_yee(FILE* in, const Yee& y) {
	for (int i = 0; i < y.d[0]; ++i)
		for (int j = 0; j < y.d[1]; ++j)
			for (int k = 0; k < y.d[2]; ++k) y.a[i][j][k] = rd<double>(in);
}

void dump_yee(const Yee& y, const std::string& name) {
	std::vector<double> v;
	v.reserve((size_t)y.d[0] * y.d[1] * y.d[2]);
	for (int i = 0; i < y.d[0]; ++i)
		for (int j = 0; j < y.d[1]; ++j)
			for (int k = 0; k < y.d[2]; ++k) v.push_back(y.a[i][j][k]);
	dump(name, v, {(uint64_t)y.d[0], (uint64_t)y.d[1], (uint64_t)y.d[2]});
}

void dump_particles(Simulation& s, const std::string& name) {
	std::vector<double> d;
	for (size_t p = 0; p < s.particles.size(); ++p) {
		Particle* q = s.particles[p];
		Vector3d m = q->getMomentum();
		d.push_back((double)q->number);
		d.push_back(q->coordinates.x);
		d.push_back(q->coordinates.y);
		d.push_back(q->coordinates.z);
		d.push_back(m.x);
		d.push_back(m.y);
		d.push_back(m.z);
	}
	dump(name, d, {(uint64_t)s.particles.size(), 7});
}

}  // namespace

int main(int argc, char** argv) {
	MPI_Init(&argc, &argv);
	if (argc < 3) {
		fprintf(stderr, "usage: ref_harness_explicit input.bin output.bin\n");
		return 2;
	}
	FILE* in = fopen(argv[1], "rb");
	if (!in) {
		fprintf(stderr, "cannot open %s\n", argv[1]);
		return 2;
	}
	int nx = rd<int32_t>(in), ny = rd<int32_t>(in), nz = rd<int32_t>(in);
	double dx = rd<double>(in), dt = rd<double>(in);
	int nsteps = rd<int32_t>(in);
	int nspecies = rd<int32_t>(in);
	std::vector<int> sp_type(nspecies);
	std::vector<double> sp_mass(nspecies), sp_charge(nspecies), sp_weight(nspecies);
	for (int s = 0; s < nspecies; ++s) {
		sp_type[s] = rd<int32_t>(in);
		sp_mass[s] = rd<double>(in);
		sp_charge[s] = rd<double>(in);
		sp_weight[s] = rd<double>(in);
	}
	int64_t np = rd<int64_t>(in);

	int dims[3] = {1, 1, 1};
	int period[3] = {1, 1, 1};
	MPI_Comm comm;
	MPI_Cart_create(MPI_COMM_WORLD, 3, dims, period, 0, &comm);
	int ppb[8] = {1, 1, 0, 0, 0, 0, 0, 0};
	double conc[8] = {1, 1, 0, 0, 0, 0, 0, 0};
	// the 37-argument constructor (PIC++/initialization.cpp:506-516) with the BUNEMAN solver
	Simulation* sim = new Simulation(nx, ny, nz, dx, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0,
	                                 nsteps + 10, 1e30, 1000000, 1, 1000000, 1000000, 0, 0.1, false, 8, ppb, conc,
	                                 1, 1, 0, dt, 0.01, 1.0, 1.0, BUNEMAN, comm);
	Simulation& s = *sim;
	s.timing = false;
	s.createArrays();
	s.initialize();
	s.deltaT = dt;
	s.smoothingCount = 0;
	const int NX = s.xnumberAdded, NY = s.ynumberAdded, NZ = s.znumberAdded;
	// PIC++/initialization.cpp:7146-7174
	Yee Ex{s.bunemanEx, {NX, NY + 1, NZ + 1}}, Ey{s.bunemanEy, {NX + 1, NY, NZ + 1}}, Ez{s.bunemanEz, {NX + 1, NY + 1, NZ}};
	Yee Bx{s.bunemanBx, {NX + 1, NY, NZ}}, By{s.bunemanBy, {NX, NY + 1, NZ}}, Bz{s.bunemanBz, {NX, NY, NZ + 1}};
	Yee Jx{s.bunemanJx, {NX, NY + 1, NZ + 1}}, Jy{s.bunemanJy, {NX + 1, NY, NZ + 1}}, Jz{s.bunemanJz, {NX + 1, NY + 1, NZ}};
	for (const Yee* y : {&Ex, &Ey, &Ez, &Bx, &By, &Bz}) read_yee(in, *y);
	for (int64_t p = 0; p < np; ++p) {
		int sp = rd<int32_t>(in);
		double x = rd<double>(in), y = rd<double>(in), z = rd<double>(in);
		double px = rd<double>(in), py = rd<double>(in), pz = rd<double>(in);
		Particle* q = new Particle((int)p, sp_mass[sp], sp_charge[sp] > 0 ? 1 : -1, sp_charge[sp], sp_weight[sp],
		                           (ParticleTypes)sp_type[sp], x, y, z, px, py, pz, s.deltaX / 2, s.deltaY / 2,
		                           s.deltaZ / 2, dt);
		s.particles.push_back(q);
	}
	// optional trailer: smoothing passes of the current every step (smoothingCount,
	// smoothingParameter; smoothBunemanEfieldGeneral after tristanUpdateFlux,
	// PIC++/simulation.cpp:185-188, PIC++/smoothing.cpp:535)
	int32_t smooth_count = 0;
	double smooth_param = 0.0;
	if (fread(&smooth_count, sizeof(int32_t), 1, in) == 1) smooth_param = rd<double>(in);
	s.smoothingCount = smooth_count;
	s.smoothingParameter = smooth_param;
	// optional trailer: x boundary (1: SUPER_CONDUCTOR_LEFT with the fields E0, B0 at the free
	// right end, as initializeFluxFromRight sets up, PIC++/initialization.cpp:5201-5210; the
	// Tristan mover reflects at xgrid[reflectingWallPoint]), then injection (V0, per species the
	// particles per bin and kT / (m c^2)), each step injectNewParticles and addToPreserveChargeGlobal
	// after the particle exchange (PIC++/simulation.cpp:274-289)
	int32_t boundary = 0, inject = 0;
	if (fread(&boundary, sizeof(int32_t), 1, in) != 1) boundary = 0;
	if (boundary == 1) {
		Vector3d e_right, b_right;
		e_right.x = rd<double>(in);
		e_right.y = rd<double>(in);
		e_right.z = rd<double>(in);
		b_right.x = rd<double>(in);
		b_right.y = rd<double>(in);
		b_right.z = rd<double>(in);
		s.boundaryConditionTypeX = SUPER_CONDUCTOR_LEFT;
		s.E0 = e_right;  // tristanEvaluateE holds E0 at the free end (PIC++/fields.cpp:117-141)
		s.B0 = b_right;
		s.rightBoundaryFieldEvaluator = new ConstantBoundaryFieldEvaluator(e_right, b_right);
		if (fread(&inject, sizeof(int32_t), 1, in) != 1) inject = 0;
	}
	if (inject == 1) {
		s.V0.x = rd<double>(in);
		s.V0.y = rd<double>(in);
		s.V0.z = rd<double>(in);
		for (int t = 0; t < s.typesNumber; ++t) s.types[t].particlesPerBin = 0;
		for (int q = 0; q < nspecies; ++q) {
			const int ppc = rd<int32_t>(in);
			const double theta = rd<double>(in);
			int t = 0;
			while (t < s.typesNumber && s.types[t].type != (ParticleTypes)sp_type[q]) ++t;
			if (t == s.typesNumber) {
				fprintf(stderr, "species type %d not found\n", sp_type[q]);
				return 2;
			}
			ParticleTypeContainer& tc = s.types[t];
			tc.mass = sp_mass[q];
			tc.charge = sp_charge[q];
			tc.chargeCount = sp_charge[q] > 0 ? 1 : -1;
			tc.particlesPerBin = ppc;
			tc.particesDeltaX = s.deltaX / ppc;
			tc.injectionLength = tc.particesDeltaX - 0.0001 * s.deltaX;  // PIC++/simulation.cpp:125-127
			const double T = theta * sp_mass[q] / s.kBoltzman_normalized;
			tc.temperatureX = tc.temperatureY = tc.temperatureZ = T;
			tc.concentration = sp_weight[q] * ppc / s.volumeB();
		}
		s.preserveChargeGlobal = true;
	}
	fclose(in);
	s.particlesNumber = (int)np;

	g_out = fopen(argv[2], "wb");
	std::vector<double> meta = {(double)NX, (double)NY, (double)NZ, s.deltaX, s.deltaT, s.xgrid[0], s.ygrid[0], s.zgrid[0],
	                            s.cellVolume};
	dump("meta", meta, {meta.size()});

	// pre-loop exchanges (PIC++/simulation.cpp:83-85)
	s.exchangeBunemanBfield(s.bunemanBx, s.bunemanBy, s.bunemanBz);
	s.exchangeBunemanEfield(s.bunemanEx, s.bunemanEy, s.bunemanEz);
	const char* en[3] = {"Ex", "Ey", "Ez"};
	const char* bn[3] = {"Bx", "By", "Bz"};
	const char* jn[3] = {"Jx", "Jy", "Jz"};
	const Yee* ey[3] = {&Ex, &Ey, &Ez};
	const Yee* by[3] = {&Bx, &By, &Bz};
	const Yee* jy[3] = {&Jx, &Jy, &Jz};
	dump_particles(s, tag("particles", 0));
	for (int c = 0; c < 3; ++c) {
		dump_yee(*ey[c], tag(en[c], 0));
		dump_yee(*by[c], tag(bn[c], 0));
	}
	// updateEnergy with the Yee fields (getBunemanElectricField / getBunemanMagneticField,
	// PIC++/simulation.cpp:1506-1640): particle, E, B and total energy, total momentum
	auto dump_energy = [&](int step) {
		s.updateEnergy();
		std::vector<double> e = {s.particleEnergy, s.electricFieldEnergy, s.magneticFieldEnergy, s.energy,
		                         s.globalMomentum.x, s.globalMomentum.y, s.globalMomentum.z};
		dump(tag("energy", step), e, {7});
	};
	s.currentIteration = 0;
	dump_energy(0);
	for (int step = 1; step <= nsteps; ++step) {
		// PIC++/simulation.cpp:163-197, BUNEMAN branch
		s.tristanEvaluateBhalfStep();
		s.exchangeBunemanBfield(s.bunemanBx, s.bunemanBy, s.bunemanBz);
		s.moveParticles();
		s.tristanEvaluateBhalfStep();
		s.exchangeBunemanBfield(s.bunemanBx, s.bunemanBy, s.bunemanBz);
		s.tristanUpdateFlux();
		for (int n = 0; n < s.smoothingCount; ++n) s.smoothBunemanEfieldGeneral(s.bunemanJx, s.bunemanJy, s.bunemanJz);
		if (step <= 2)
			for (int c = 0; c < 3; ++c) dump_yee(*jy[c], tag(jn[c], step));
		s.tristanEvaluateE();
		s.exchangeBunemanEfield(s.bunemanEx, s.bunemanEy, s.bunemanEz);
		// :258-263, :274-289, :388
		s.eraseEscapedPaticles();
		s.exchangeParticles();
		if (inject == 1) {
			for (int t = 0; t < s.typesNumber; ++t) s.types[t].injectionLength += fabs(s.V0.x * s.deltaT);
			s.injectNewParticles();
			s.addToPreserveChargeGlobal();
		}
		s.removeEscapedParticles();
		s.time += s.deltaT;
		s.currentIteration++;
		dump_particles(s, tag("particles", step));
		for (int c = 0; c < 3; ++c) {
			dump_yee(*ey[c], tag(en[c], step));
			dump_yee(*by[c], tag(bn[c], step));
		}
		dump_energy(step);
	}
	fclose(g_out);
	MPI_Finalize();
	return 0;
}
