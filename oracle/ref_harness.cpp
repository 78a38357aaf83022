// Golden-vector harness for the reference PIC++ implicit step.
//
// Test infrastructure only: links against the unmodified reference objects
// (compiled from PIC++/*.cpp by oracle/Makefile.ref) and drives the public
// Simulation methods in the order of the reference step loop
// (PIC++/simulation.cpp:94-123 pre-loop, :205-390 loop body, implicit branch,
// smoothingCount = 0), dumping every per-phase array the hot path produces.
//
// With a third argument (a path prefix) it first writes the reference's backup files of the
// loaded state (outputSimulationBackup): prefix + general.dat, E.dat, B.dat, particles.dat.
//
// Input  (binary, little endian): see read_input() below.
// Output (binary): a sequence of named float64 arrays, read by
//                  tests/golden/make_golden.py.
#include <mpi.h>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cmath>
#include <string>
#include <vector>

#include "output.h"
#include "simulation.h"
#include "boundaryFieldEvaluator.h"
#include "particle.h"
#include "vector3d.h"
#include "matrix3d.h"

namespace {

FILE* g_out = nullptr;

void dump(const char* name, const std::vector<double>& data, std::vector<uint64_t> dims) {
	uint32_t len = (uint32_t)std::string(name).size();
	fwrite(&len, 4, 1, g_out);
	fwrite(name, 1, len, g_out);
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

void dump_node_vec(Simulation& s, Vector3d*** f, const std::string& name) {
	int NX = s.xnumberAdded + 1, NY = s.ynumberAdded + 1, NZ = s.znumberAdded + 1;
	std::vector<double> d;
	d.reserve((size_t)NX * NY * NZ * 3);
	for (int i = 0; i < NX; ++i)
		for (int j = 0; j < NY; ++j)
			for (int k = 0; k < NZ; ++k)
				for (int l = 0; l < 3; ++l) d.push_back(f[i][j][k][l]);
	dump(name.c_str(), d, {(uint64_t)NX, (uint64_t)NY, (uint64_t)NZ, 3});
}

void dump_node_mat(Simulation& s, Matrix3d*** f, const std::string& name) {
	int NX = s.xnumberAdded + 1, NY = s.ynumberAdded + 1, NZ = s.znumberAdded + 1;
	std::vector<double> d;
	for (int i = 0; i < NX; ++i)
		for (int j = 0; j < NY; ++j)
			for (int k = 0; k < NZ; ++k)
				for (int a = 0; a < 3; ++a)
					for (int b = 0; b < 3; ++b) d.push_back(f[i][j][k].matrix[a][b]);
	dump(name.c_str(), d, {(uint64_t)NX, (uint64_t)NY, (uint64_t)NZ, 3, 3});
}

void dump_cell_vec(Simulation& s, Vector3d*** f, const std::string& name) {
	int NX = s.xnumberAdded, NY = s.ynumberAdded, NZ = s.znumberAdded;
	std::vector<double> d;
	for (int i = 0; i < NX; ++i)
		for (int j = 0; j < NY; ++j)
			for (int k = 0; k < NZ; ++k)
				for (int l = 0; l < 3; ++l) d.push_back(f[i][j][k][l]);
	dump(name.c_str(), d, {(uint64_t)NX, (uint64_t)NY, (uint64_t)NZ, 3});
}

void dump_cell_mat(Simulation& s, Matrix3d*** f, const std::string& name) {
	int NX = s.xnumberAdded, NY = s.ynumberAdded, NZ = s.znumberAdded;
	std::vector<double> d;
	for (int i = 0; i < NX; ++i)
		for (int j = 0; j < NY; ++j)
			for (int k = 0; k < NZ; ++k)
				for (int a = 0; a < 3; ++a)
					for (int b = 0; b < 3; ++b) d.push_back(f[i][j][k].matrix[a][b]);
	dump(name.c_str(), d, {(uint64_t)NX, (uint64_t)NY, (uint64_t)NZ, 3, 3});
}

void dump_cell_scalar(Simulation& s, double*** f, const std::string& name) {
	int NX = s.xnumberAdded, NY = s.ynumberAdded, NZ = s.znumberAdded;
	std::vector<double> d;
	for (int i = 0; i < NX; ++i)
		for (int j = 0; j < NY; ++j)
			for (int k = 0; k < NZ; ++k) d.push_back(f[i][j][k]);
	dump(name.c_str(), d, {(uint64_t)NX, (uint64_t)NY, (uint64_t)NZ});
}

void dump_rhs(Simulation& s, const std::string& name) {
	int NX = s.xnumberAdded, NY = s.ynumberAdded, NZ = s.znumberAdded;
	std::vector<double> d;
	for (int i = 0; i < NX; ++i)
		for (int j = 0; j < NY; ++j)
			for (int k = 0; k < NZ; ++k)
				for (int l = 0; l < 3; ++l) d.push_back(s.maxwellEquationRightPart[i][j][k][l]);
	dump(name.c_str(), d, {(uint64_t)NX, (uint64_t)NY, (uint64_t)NZ, 3});
}

// Particles in the reference's current vector order: id, x, y, z, px, py, pz.
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
	dump(name.c_str(), d, {(uint64_t)s.particles.size(), 7});
}

void dump_alpha(Simulation& s, const std::string& name) {
	std::vector<double> d;
	for (size_t p = 0; p < s.particles.size(); ++p) {
		Particle* q = s.particles[p];
		d.push_back((double)q->number);
		for (int a = 0; a < 3; ++a)
			for (int b = 0; b < 3; ++b) d.push_back(q->rotationTensor.matrix[a][b]);
	}
	dump(name.c_str(), d, {(uint64_t)s.particles.size(), 10});
}

void dump_species(Simulation& s, int step, const std::vector<int>& types) {
	for (size_t t = 0; t < types.size(); ++t) {
		int tn = types[t];
		char nb[64];
		snprintf(nb, sizeof(nb), "conc%zu", t);
		dump_cell_scalar(s, s.particleConcentrations[tn], tag(nb, step));
		snprintf(nb, sizeof(nb), "energy%zu", t);
		dump_cell_scalar(s, s.particleEnergies[tn], tag(nb, step));
		snprintf(nb, sizeof(nb), "bulkv%zu", t);
		dump_cell_vec(s, s.particleBulkVelocities[tn], tag(nb, step));
	}
}

void dump_energy(Simulation& s, int step) {
	std::vector<double> d = {s.particleEnergy, s.electricFieldEnergy, s.magneticFieldEnergy, s.energy,
	                         s.globalMomentum.x, s.globalMomentum.y, s.globalMomentum.z};
	dump(tag("energy", step).c_str(), d, {7});
}

// The numbers of one general.dat row (outputGeneral, PIC++/output.cpp:3215-3250), unscaled, in
// the row's order without the iteration and the time columns.
void dump_general(Simulation& s, int step) {
	std::vector<double> d = {s.particleEnergy, s.electricFieldEnergy, s.magneticFieldEnergy, s.energy,
	                         s.globalMomentum.x, s.globalMomentum.y, s.globalMomentum.z, s.generalTheoreticalEnergy,
	                         s.generalTheoreticalMomentum.x, s.generalTheoreticalMomentum.y, s.generalTheoreticalMomentum.z,
	                         s.maxEfield.norm(), s.maxBfield.norm(), (double)s.particlesNumber, s.meanSquaredEfield[0],
	                         s.meanSquaredEfield[1], s.meanSquaredEfield[2], s.electricFieldEnergyX, s.electricFieldEnergyY,
	                         s.electricFieldEnergyZ, s.magneticFieldEnergyX, s.magneticFieldEnergyY, s.magneticFieldEnergyZ,
	                         s.electromagneticMomentum.x, s.electromagneticMomentum.y, s.electromagneticMomentum.z,
	                         s.particleMomentum.x, s.particleMomentum.y, s.particleMomentum.z};
	dump(tag("general", step).c_str(), d, {d.size()});
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

}  // namespace

int main(int argc, char** argv) {
	MPI_Init(&argc, &argv);
	if (argc < 3) {
		fprintf(stderr, "usage: ref_harness input.bin output.bin\n");
		return 2;
	}
	FILE* in = fopen(argv[1], "rb");
	if (!in) {
		fprintf(stderr, "cannot open %s\n", argv[1]);
		return 2;
	}
	// ---- input ----
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
	// 37-argument constructor with the IMPLICIT solver (PIC++/initialization.cpp:506-516).
	Simulation* sim = new Simulation(nx, ny, nz, dx, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0,
	                                 nsteps + 10, 1e30, 1000000, 1, 1000000, 1000000, 0, 0.1, false, 8, ppb, conc,
	                                 1, 1, 0, dt, 0.01, 1.0, 1.0, IMPLICIT, comm);
	Simulation& s = *sim;
	s.timing = false;
	s.createArrays();
	s.initialize();
	s.deltaT = dt;

	// Fields: E on (N+1)^3 nodes, B on N^3 cells, both ghosted.
	int NX = s.xnumberAdded, NY = s.ynumberAdded, NZ = s.znumberAdded;
	for (int i = 0; i <= NX; ++i)
		for (int j = 0; j <= NY; ++j)
			for (int k = 0; k <= NZ; ++k) {
				Vector3d e;
				e.x = rd<double>(in);
				e.y = rd<double>(in);
				e.z = rd<double>(in);
				s.Efield[i][j][k] = e;
				s.newEfield[i][j][k] = e;
				s.tempEfield[i][j][k] = e;
			}
	for (int i = 0; i < NX; ++i)
		for (int j = 0; j < NY; ++j)
			for (int k = 0; k < NZ; ++k) {
				Vector3d b;
				b.x = rd<double>(in);
				b.y = rd<double>(in);
				b.z = rd<double>(in);
				s.Bfield[i][j][k] = b;
				s.newBfield[i][j][k] = b;
			}
	for (int64_t p = 0; p < np; ++p) {
		int sp = rd<int32_t>(in);
		double x = rd<double>(in), y = rd<double>(in), z = rd<double>(in);
		double px = rd<double>(in), py = rd<double>(in), pz = rd<double>(in);
		Particle* q = new Particle((int)p, sp_mass[sp], sp_charge[sp] > 0 ? 1 : -1, sp_charge[sp], sp_weight[sp],
		                           (ParticleTypes)sp_type[sp], x, y, z, px, py, pz, s.deltaX / 2, s.deltaY / 2,
		                           s.deltaZ / 2, dt);
		s.particles.push_back(q);
	}
	// optional trailer: x boundary (0 periodic, 1 super-conducting left wall + constant-field
	// free right boundary, as initializeFluxFromRight sets up, PIC++/initialization.cpp:5201-5210)
	int32_t boundary = 0;
	if (fread(&boundary, sizeof(int32_t), 1, in) != 1) boundary = 0;
	Vector3d e_right, b_right;
	if (boundary == 1) {
		e_right.x = rd<double>(in);
		e_right.y = rd<double>(in);
		e_right.z = rd<double>(in);
		b_right.x = rd<double>(in);
		b_right.y = rd<double>(in);
		b_right.z = rd<double>(in);
		s.boundaryConditionTypeX = SUPER_CONDUCTOR_LEFT;
		s.rightBoundaryFieldEvaluator = new ConstantBoundaryFieldEvaluator(e_right, b_right);
	}
	// optional trailer: injection at the right end (config of initializeFluxFromRight): flow
	// velocity V0, then per species the particles per bin and kT / (m c^2). Each step then adds
	// columns of new particles (Simulation::injectNewParticles, PIC++/mover.cpp:910-1000) and
	// replaces right escapes (addToPreserveChargeGlobal, PIC++/initialization.cpp:8281-8338),
	// as PIC++/simulation.cpp:281-289 does.
	int32_t inject = 0;
	if (boundary == 1 && fread(&inject, sizeof(int32_t), 1, in) != 1) inject = 0;
	if (inject == 1) {
		s.V0.x = rd<double>(in);
		s.V0.y = rd<double>(in);
		s.V0.z = rd<double>(in);
		for (int t = 0; t < s.typesNumber; ++t) s.types[t].particlesPerBin = 0;
		for (int q = 0; q < nspecies; ++q) {
			int ppc = rd<int32_t>(in);
			double theta = rd<double>(in);
			int t = 0;
			while (t < s.typesNumber && s.types[t].type != (ParticleTypes)sp_type[q]) ++t;
			if (t == s.typesNumber) {
				fprintf(stderr, "species type %d not found\n", sp_type[q]);
				return 2;
			}
			ParticleTypeContainer& tc = s.types[t];
			tc.mass = sp_mass[q];  // the species as loaded (createParticle reads the type's mass and charge)
			tc.charge = sp_charge[q];
			tc.chargeCount = sp_charge[q] > 0 ? 1 : -1;
			tc.particlesPerBin = ppc;
			tc.particesDeltaX = s.deltaX / ppc;
			tc.injectionLength = tc.particesDeltaX - 0.0001 * s.deltaX;  // PIC++/simulation.cpp:125-127
			double T = theta * sp_mass[q] / s.kBoltzman_normalized;
			tc.temperatureX = tc.temperatureY = tc.temperatureZ = T;
			tc.concentration = sp_weight[q] * ppc / s.volumeB();  // weight = concentration / ppc * volumeB
		}
		// the tabulated Juttner CDFs of the types, from their temperatures (as the reference's
		// set-ups do before creating particles, PIC++/initialization.cpp:5393)
		s.evaluateJuttnerFunctions();
		s.preserveChargeGlobal = true;
	}
	// optional trailer: smoothing passes and parameter (smoothingCount, smoothingParameter;
	// the step loop smooths chargeDensityHat, electricFlux and dielectricTensor,
	// PIC++/simulation.cpp:213-217)
	int32_t smooth_count = 0;
	double smooth_param = 0.0;
	if (fread(&smooth_count, sizeof(int32_t), 1, in) == 1) {
		smooth_param = rd<double>(in);
		s.smoothingCount = smooth_count;
		s.smoothingParameter = smooth_param;
	} else {
		s.smoothingCount = 0;
	}
	// optional trailer: divergence cleanup (bit 1: cleanupDivergence(newEfield, chargeDensity), bit 2:
	// cleanupDivergenceMagnetic(), every step at their call position in the loop,
	// PIC++/simulation.cpp:348-349, after updateDensityParameters and the newE exchange) and the cut
	// wave number of filterFields (> 0: filterFields(cut) next)
	int32_t cleanup = 0, filter_cut = 0;
	if (fread(&cleanup, sizeof(int32_t), 1, in) == 1) filter_cut = rd<int32_t>(in);
	// optional trailer: general rows (1: the loop's diagnostics for general.dat, as simulate()
	// keeps them: updateParameters in the pre-loop sequence and after every step,
	// updateTheoreticalEnergy between eraseEscapedPaticles and exchangeParticles, outputGeneral
	// after every updateEnergy; PIC++/simulation.cpp:96, 138-148, 266-272, 381-384)
	int32_t general = 0;
	if (fread(&general, sizeof(int32_t), 1, in) != 1) general = 0;
	// optional trailer: kBoltzman_normalized x the run's temperature in m c^2 units
	// (Simulation::temperature, which createParticle's Juttner branch choice reads,
	// PIC++/initialization.cpp:8515-8527)
	double run_kT = 0.0;
	if (fread(&run_kT, sizeof(double), 1, in) == 1) s.temperature = run_kT / s.kBoltzman_normalized;
	fclose(in);
	s.particlesNumber = (int)np;
	if (argc > 3) {
		// the reference's backup files of the loaded state (outputSimulationBackup,
		// PIC++/output.cpp:3432), named by the prefix argv[3]
		const std::string pre = argv[3];
		outputSimulationBackup((pre + "general.dat").c_str(), (pre + "E.dat").c_str(), (pre + "B.dat").c_str(),
		                       (pre + "particles.dat").c_str(), &s);
	}

	g_out = fopen(argv[2], "wb");
	std::vector<double> meta = {(double)NX, (double)NY, (double)NZ, s.deltaX, s.deltaT, s.xgrid[0], s.ygrid[0],
	                            s.zgrid[0], s.theta, s.eta, s.cellVolume};
	dump("meta", meta, {meta.size()});

	// ---- pre-loop (PIC++/simulation.cpp:86-123, implicit branch) ----
	s.exchangeEfield();
	s.exchangeGeneralBfield(s.Bfield);
	s.exchangeGeneralBfield(s.newBfield);
	s.updateParticleCorrelationMaps();
	if (general) s.updateParameters();
	s.evaluateParticlesRotationTensor();
	s.updateElectroMagneticParameters();
	s.updateDensityParameters();
	s.updateFields();
	s.currentIteration = 0;
	s.updateEnergy();
	const std::string general_file = std::string(argv[2]) + ".general";
	if (general) {
		FILE* gf = fopen(general_file.c_str(), "w");  // outputGeneral appends
		if (gf) fclose(gf);
		outputGeneral(general_file.c_str(), &s);
		dump_general(s, 0);
	}
	dump_particles(s, tag("particles", 0));
	dump_node_vec(s, s.Efield, tag("E", 0));
	dump_cell_vec(s, s.Bfield, tag("B", 0));
	dump_cell_scalar(s, s.chargeDensity, tag("rho", 0));
	dump_species(s, 0, sp_type);
	dump_energy(s, 0);

	// ---- loop body (PIC++/simulation.cpp:205-390, implicit branch, smoothingCount = 0) ----
	for (int step = 1; step <= nsteps; ++step) {
		bool full = (step <= 2);
		s.evaluateParticlesRotationTensor();
		if (full) dump_alpha(s, tag("alpha", step));
		s.updateElectroMagneticParameters();
		for (int n = 0; n < s.smoothingCount; ++n) {  // PIC++/simulation.cpp:213-217
			s.smoothChargeDensityHat();
			s.smoothFlux();
			s.smoothMatrixNodeParameter(s.dielectricTensor);
		}
		if (full) {
			dump_node_vec(s, s.electricFlux, tag("J", step));
			dump_node_mat(s, s.dielectricTensor, tag("mu", step));
			dump_cell_scalar(s, s.chargeDensityHat, tag("rhohat", step));
			dump_cell_mat(s, s.pressureTensor, tag("Pi", step));
			dump_node_vec(s, s.divPressureTensor, tag("divPi", step));
			s.evaluateMaxwellEquationMatrix();
			dump_rhs(s, tag("rhs", step));
		}
		s.evaluateElectricField();
		s.exchangeEfield();
		s.evaluateMagneticField();
		s.exchangeGeneralBfield(s.newBfield);
		s.exchangeGeneralEfield(s.tempEfield);
		s.exchangeGeneralEfield(s.newEfield);
		if (full) {
			dump_node_vec(s, s.tempEfield, tag("Etheta", step));
			dump_node_vec(s, s.newEfield, tag("Enew", step));
			dump_cell_vec(s, s.newBfield, tag("Bnew", step));
		}
		s.moveParticles();
		s.eraseEscapedPaticles();
		if (general) s.updateTheoreticalEnergy();
		s.exchangeParticles();
		if (inject == 1) {  // PIC++/simulation.cpp:281-289
			for (int t = 0; t < s.typesNumber; ++t) s.types[t].injectionLength += fabs(s.V0.x * s.deltaT);
			s.injectNewParticles();
			s.addToPreserveChargeGlobal();
		}
		s.updateParticleCorrelationMaps();
		s.updateDensityParameters();
		s.exchangeGeneralEfield(s.newEfield);
		if (cleanup & 1) s.cleanupDivergence(s.newEfield, s.chargeDensity);
		if (cleanup & 2) s.cleanupDivergenceMagnetic();
		if (filter_cut > 0) s.filterFields(filter_cut);
		if ((cleanup != 0 || filter_cut > 0) && full) {
			dump_node_vec(s, s.newEfield, tag("Enew_post", step));
			dump_cell_vec(s, s.newBfield, tag("Bnew_post", step));
		}
		s.updateFields();
		s.exchangeGeneralEfield(s.newEfield);
		s.exchangeGeneralEfield(s.Efield);
		s.exchangeGeneralBfield(s.Bfield);
		if (general) s.updateParameters();
		s.removeEscapedParticles();
		s.time += s.deltaT;
		s.currentIteration++;
		s.updateEnergy();
		if (general) {
			outputGeneral(general_file.c_str(), &s);
			dump_general(s, step);
		}
		dump_particles(s, tag("particles", step));
		if (inject == 1) {  // reference type of every particle, in the same order
			std::vector<double> ty;
			for (size_t p = 0; p < s.particles.size(); ++p) ty.push_back((double)s.particles[p]->type);
			dump(tag("ptype", step).c_str(), ty, {(uint64_t)ty.size()});
		}
		dump_node_vec(s, s.Efield, tag("E", step));
		dump_cell_vec(s, s.Bfield, tag("B", step));
		dump_cell_scalar(s, s.chargeDensity, tag("rho", step));
		if (full) dump_species(s, step, sp_type);
		dump_energy(s, step);
	}
	fclose(g_out);
	MPI_Finalize();
	return 0;
}
