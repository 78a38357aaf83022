// Timing harness for the compiled reference's implicit step on the CPU at any number of MPI ranks.
//
// Test and measurement infrastructure only: links against the unmodified reference objects
// (oracle/Makefile.ref) and times the public Simulation methods of the reference step loop
// (PIC++/simulation.cpp:205-390, implicit branch, smoothingCount = 0, no output) phase by phase.
// The ranks form a slab decomposition along x (cartDim = {ranks, 1, 1}, as PIC++/input/MPI_input.dat
// would give), each rank owning its share of the cells (Simulation::setSpaceForProc,
// PIC++/initialization.cpp:71-140) and of the particles.
//
// Input: the file format of oracle/ref_harness.cpp (header, ghosted E and B of the whole box,
// particles), read by every rank. The fields must be uniform (the timed workload starts from a
// uniform B0 and E = 0): every rank fills its own ghosted arrays with that value. A particle
// belongs to the rank whose owned x range holds it.
// Output (rank 0, stdout): one line per phase, "phase <name> <seconds>", then totals. Every phase
// ends with a barrier, so a phase's time is that of the slowest rank.
#include <mpi.h>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "simulation.h"
#include "particle.h"
#include "vector3d.h"

namespace {

template <class T>
T rd(FILE* f) {
	T v;
	if (fread(&v, sizeof(T), 1, f) != 1) {
		fprintf(stderr, "short input\n");
		exit(2);
	}
	return v;
}

struct PhaseClock {
	std::vector<std::string> names;
	std::vector<double> secs;
	MPI_Comm comm;
	double t0 = 0;
	void start() {
		MPI_Barrier(comm);
		t0 = MPI_Wtime();
	}
	void stop(const char* name) {
		MPI_Barrier(comm);
		double dt = MPI_Wtime() - t0;
		for (size_t i = 0; i < names.size(); ++i)
			if (names[i] == name) {
				secs[i] += dt;
				return;
			}
		names.push_back(name);
		secs.push_back(dt);
	}
};

}  // namespace

int main(int argc, char** argv) {
	MPI_Init(&argc, &argv);
	int world = 1, wrank = 0;
	MPI_Comm_size(MPI_COMM_WORLD, &world);
	MPI_Comm_rank(MPI_COMM_WORLD, &wrank);
	if (argc < 2) {
		if (wrank == 0) fprintf(stderr, "usage: ref_timing input.bin [steps]\n");
		MPI_Finalize();
		return 2;
	}
	FILE* in = fopen(argv[1], "rb");
	if (!in) {
		fprintf(stderr, "cannot open %s\n", argv[1]);
		MPI_Abort(MPI_COMM_WORLD, 2);
	}
	int nx = rd<int32_t>(in), ny = rd<int32_t>(in), nz = rd<int32_t>(in);
	double dx = rd<double>(in), dt = rd<double>(in);
	int nsteps = rd<int32_t>(in);
	if (argc > 2) nsteps = atoi(argv[2]);
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

	int dims[3] = {world, 1, 1};
	int period[3] = {1, 1, 1};
	MPI_Comm comm;
	MPI_Cart_create(MPI_COMM_WORLD, 3, dims, period, 0, &comm);
	int ppb[8] = {1, 1, 0, 0, 0, 0, 0, 0};
	double conc[8] = {1, 1, 0, 0, 0, 0, 0, 0};
	// the 37-argument constructor with the IMPLICIT solver (PIC++/initialization.cpp:506-516),
	// as oracle/ref_harness.cpp builds it
	Simulation* sim = new Simulation(nx, ny, nz, dx, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0,
	                                 nsteps + 10, 1e30, 1000000, 1, 1000000, 1000000, 0, 0.1, false, 8, ppb, conc,
	                                 1, world, 0, dt, 0.01, 1.0, 1.0, IMPLICIT, comm);
	Simulation& s = *sim;
	s.timing = false;
	s.createArrays();
	s.initialize();
	s.deltaT = dt;

	// the whole box's ghosted fields: checked uniform, then spread over this rank's arrays
	const int GX = nx + 4, GY = ny + 4, GZ = nz + 4;
	Vector3d e0, b0;
	double nonuniform = 0;
	for (int i = 0; i <= GX; ++i)
		for (int j = 0; j <= GY; ++j)
			for (int k = 0; k <= GZ; ++k) {
				Vector3d e;
				e.x = rd<double>(in);
				e.y = rd<double>(in);
				e.z = rd<double>(in);
				if (i == 0 && j == 0 && k == 0) e0 = e;
				nonuniform = fmax(nonuniform, (e - e0).norm());
			}
	for (int i = 0; i < GX; ++i)
		for (int j = 0; j < GY; ++j)
			for (int k = 0; k < GZ; ++k) {
				Vector3d b;
				b.x = rd<double>(in);
				b.y = rd<double>(in);
				b.z = rd<double>(in);
				if (i == 0 && j == 0 && k == 0) b0 = b;
				nonuniform = fmax(nonuniform, (b - b0).norm());
			}
	if (nonuniform > 0) {
		if (wrank == 0) fprintf(stderr, "ref_timing needs uniform initial fields (deviation %g)\n", nonuniform);
		MPI_Abort(MPI_COMM_WORLD, 2);
	}
	const int NX = s.xnumberAdded, NY = s.ynumberAdded, NZ = s.znumberAdded;
	for (int i = 0; i <= NX; ++i)
		for (int j = 0; j <= NY; ++j)
			for (int k = 0; k <= NZ; ++k) {
				s.Efield[i][j][k] = e0;
				s.newEfield[i][j][k] = e0;
				s.tempEfield[i][j][k] = e0;
			}
	for (int i = 0; i < NX; ++i)
		for (int j = 0; j < NY; ++j)
			for (int k = 0; k < NZ; ++k) {
				s.Bfield[i][j][k] = b0;
				s.newBfield[i][j][k] = b0;
			}
	// this rank's particles: x in its owned cells [xgrid[1 + a], xgrid[NX - 1 - a])
	const double xlo = s.xgrid[1 + additionalBinNumber], xhi = s.xgrid[NX - 1 - additionalBinNumber];
	long long mine = 0;
	for (int64_t p = 0; p < np; ++p) {
		int sp = rd<int32_t>(in);
		double x = rd<double>(in), y = rd<double>(in), z = rd<double>(in);
		double px = rd<double>(in), py = rd<double>(in), pz = rd<double>(in);
		if (x < xlo || x >= xhi) continue;
		Particle* q = new Particle((int)p, sp_mass[sp], sp_charge[sp] > 0 ? 1 : -1, sp_charge[sp], sp_weight[sp],
		                           (ParticleTypes)sp_type[sp], x, y, z, px, py, pz, s.deltaX / 2, s.deltaY / 2,
		                           s.deltaZ / 2, dt);
		s.particles.push_back(q);
		++mine;
	}
	fclose(in);
	long long total = 0;
	MPI_Allreduce(&mine, &total, 1, MPI_LONG_LONG, MPI_SUM, comm);
	if (total != np) {
		if (wrank == 0) fprintf(stderr, "particles distributed: %lld of %lld\n", total, (long long)np);
		MPI_Abort(MPI_COMM_WORLD, 2);
	}
	s.particlesNumber = (int)np;

	PhaseClock clk;
	clk.comm = comm;
	// pre-loop (PIC++/simulation.cpp:86-123, implicit branch), timed apart from the loop
	clk.start();
	s.exchangeEfield();
	s.exchangeGeneralBfield(s.Bfield);
	s.exchangeGeneralBfield(s.newBfield);
	s.updateParticleCorrelationMaps();
	s.evaluateParticlesRotationTensor();
	s.updateElectroMagneticParameters();
	s.updateDensityParameters();
	s.updateFields();
	s.currentIteration = 0;
	s.updateEnergy();
	clk.stop("pre_loop");
	const double pre = clk.secs[0];
	clk.names.clear();
	clk.secs.clear();

	MPI_Barrier(comm);
	const double loop0 = MPI_Wtime();
	for (int step = 1; step <= nsteps; ++step) {
		clk.start();
		s.evaluateParticlesRotationTensor();
		clk.stop("rotation_tensor");  // PIC++/mover.cpp:810-841
		clk.start();
		s.updateElectroMagneticParameters();
		clk.stop("moments");  // PIC++/momentum_collecting.cpp:365-758 with the folds
		clk.start();
		s.evaluateElectricField();
		s.exchangeEfield();
		clk.stop("maxwell_solve");  // PIC++/fields.cpp:152-376, GMRES
		clk.start();
		s.evaluateMagneticField();
		s.exchangeGeneralBfield(s.newBfield);
		s.exchangeGeneralEfield(s.tempEfield);
		s.exchangeGeneralEfield(s.newEfield);
		clk.stop("b_update");  // PIC++/fields.cpp:2717-2773
		clk.start();
		s.moveParticles();
		s.eraseEscapedPaticles();
		s.exchangeParticles();
		clk.stop("push");  // PIC++/mover.cpp:21-491, particle exchange
		clk.start();
		s.updateParticleCorrelationMaps();
		s.updateDensityParameters();
		s.exchangeGeneralEfield(s.newEfield);
		clk.stop("density");  // PIC++/momentum_collecting.cpp:760-1020
		clk.start();
		s.updateFields();
		s.exchangeGeneralEfield(s.newEfield);
		s.exchangeGeneralEfield(s.Efield);
		s.exchangeGeneralBfield(s.Bfield);
		s.removeEscapedParticles();
		s.time += s.deltaT;
		s.currentIteration++;
		clk.stop("commit");  // PIC++/fields.cpp:422-439
		clk.start();
		s.updateEnergy();
		clk.stop("energy");  // PIC++/simulation.cpp:1506-1640
	}
	MPI_Barrier(comm);
	const double loop = MPI_Wtime() - loop0;
	long long after = (long long)s.particles.size(), after_total = 0;
	MPI_Allreduce(&after, &after_total, 1, MPI_LONG_LONG, MPI_SUM, comm);
	if (wrank == 0) {
		printf("ranks %d cells %d %d %d particles %lld steps %d\n", world, nx, ny, nz, (long long)np, nsteps);
		printf("pre_loop %.6f\n", pre);
		for (size_t i = 0; i < clk.names.size(); ++i) printf("phase %s %.6f\n", clk.names[i].c_str(), clk.secs[i]);
		printf("loop %.6f\n", loop);
		printf("particle_steps_per_s %.6e\n", (double)np * nsteps / loop);
		printf("particles_after %lld energy %.17g\n", after_total, s.energy);
	}
	MPI_Finalize();
	return 0;
}
