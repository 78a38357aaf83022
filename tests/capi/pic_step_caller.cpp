// A C++ caller of the C ABI, written the way INTEGRATION.md's first section shows a PIC++
// maintainer wiring the library into Simulation: a pic_config from the simulation's members,
// the reference's ghosted Efield / Bfield arrays converted to component planes, the particles
// as structure of arrays, then the step loop body (PIC++/simulation.cpp:205-390) as one entry
// point per reference method.
//
// Usage: pic_step_caller DIR
// DIR holds config.txt (see read_config), E.bin (Efield[i][j][k][c], (N0+1)(N1+1)(N2+1) x 3
// doubles), B.bin (Bfield[i][j][k][c], N0 N1 N2 x 3), species.bin, ids.bin (int32 per particle),
// x.bin, p.bin (3 doubles per particle, particle-major as Particle::coordinates / momentum).
// Writes out_E.bin, out_B.bin (same layouts), out_ids.bin, out_x.bin, out_p.bin (particle-major,
// the context's order) and out_energy.bin (7 doubles per step) after `steps` steps.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "picpp_amd.h"

namespace {

template <class T>
std::vector<T> read_all(const std::string& path) {
	std::ifstream f(path, std::ios::binary | std::ios::ate);
	if (!f) {
		std::cerr << "cannot open " << path << "\n";
		std::exit(2);
	}
	std::streamsize n = f.tellg();
	f.seekg(0);
	std::vector<T> v(n / sizeof(T));
	f.read(reinterpret_cast<char*>(v.data()), n);
	return v;
}

template <class T>
void write_all(const std::string& path, const std::vector<T>& v) {
	std::ofstream f(path, std::ios::binary);
	f.write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(T));
}

struct Setup {
	int n[3];
	double dx, dt;
	int nspecies;
	double mass[8], charge[8], weight[8];
	int steps;
	int precision, solver, stop_rule;
};

// config.txt: nx ny nz dx dt nspecies, then mass charge weight per species, then steps,
// precision (0 FP64, 1 FP32), solver (0 FP64, 1 mixed), stop rule (0 reference, 1 true residual)
Setup read_config(const std::string& path) {
	std::ifstream f(path);
	Setup s{};
	f >> s.n[0] >> s.n[1] >> s.n[2] >> s.dx >> s.dt >> s.nspecies;
	for (int q = 0; q < s.nspecies; ++q) f >> s.mass[q] >> s.charge[q] >> s.weight[q];
	f >> s.steps >> s.precision >> s.solver >> s.stop_rule;
	if (!f) {
		std::cerr << "bad " << path << "\n";
		std::exit(2);
	}
	return s;
}

int check(pic_ctx* gpu, int rc, const char* what) {
	if (rc != PIC_OK) {
		std::cerr << what << " failed (" << rc << "): " << pic_last_error(gpu) << "\n";
		std::exit(1);
	}
	return rc;
}

}  // namespace

int main(int argc, char** argv) {
	if (argc != 2) {
		std::cerr << "usage: pic_step_caller DIR\n";
		return 2;
	}
	const std::string dir = std::string(argv[1]) + "/";
	Setup s = read_config(dir + "config.txt");
	const int xnumberAdded = s.n[0] + 4, ynumberAdded = s.n[1] + 4, znumberAdded = s.n[2] + 4;

	// the context, created once after createArrays() / initialize()
	pic_ctx* gpu = nullptr;
	pic_config cfg = {};
	cfg.nx = xnumberAdded - 4;
	cfg.ny = ynumberAdded - 4;
	cfg.nz = znumberAdded - 4;
	cfg.dx = s.dx;
	cfg.dy = s.dx;
	cfg.dz = s.dx;
	cfg.dt = s.dt;
	cfg.nspecies = s.nspecies;
	for (int q = 0; q < s.nspecies; ++q) {
		cfg.mass[q] = s.mass[q];
		cfg.charge[q] = s.charge[q];
		cfg.weight[q] = s.weight[q];
	}
	std::vector<int> sp = read_all<int>(dir + "species.bin"), id = read_all<int>(dir + "ids.bin");
	std::vector<double> xa = read_all<double>(dir + "x.bin"), pa = read_all<double>(dir + "p.bin");
	const long long np = (long long)sp.size();
	cfg.capacity = 2 * np;
	cfg.precision = s.precision;
	cfg.stop_rule = s.stop_rule;
	cfg.solver_precision = s.solver;
	cfg.ortho = PIC_ORTHO_MGS;
	cfg.gmres_tolerance = 1e-8;    // maxErrorLevel (PIC++/constants.h:31)
	cfg.gmres_max_iterations = 50; // maxGMRESIterations (PIC++/constants.h:55)
	check(gpu, pic_create(&cfg, nullptr, &gpu), "pic_create");

	// fields: Efield[i][j][k][c] -> component planes
	const long long nnodes = (long long)(xnumberAdded + 1) * (ynumberAdded + 1) * (znumberAdded + 1);
	const long long ncells = (long long)xnumberAdded * ynumberAdded * znumberAdded;
	std::vector<double> Ein = read_all<double>(dir + "E.bin"), Bin = read_all<double>(dir + "B.bin");
	std::vector<double> E(3 * nnodes), B(3 * ncells);
	for (long long q = 0; q < nnodes; ++q)
		for (int c = 0; c < 3; ++c) E[c * nnodes + q] = Ein[3 * q + c];
	for (long long q = 0; q < ncells; ++q)
		for (int c = 0; c < 3; ++c) B[c * ncells + q] = Bin[3 * q + c];
	check(gpu, pic_upload_fields(gpu, E.data(), B.data()), "pic_upload_fields");

	// particles: x[c*np + q] = particles[q]->coordinates[c], p likewise
	std::vector<double> x(3 * np), p(3 * np);
	for (long long q = 0; q < np; ++q)
		for (int c = 0; c < 3; ++c) {
			x[c * np + q] = xa[3 * q + c];
			p[c * np + q] = pa[3 * q + c];
		}
	check(gpu, pic_upload_particles(gpu, np, sp.data(), id.data(), x.data(), p.data()), "pic_upload_particles");
	check(gpu, pic_initialize(gpu), "pic_initialize");

	std::vector<double> energy;
	for (int step = 0; step < s.steps; ++step) {
		// evaluateParticlesRotationTensor(); updateElectroMagneticParameters();
		check(gpu, pic_rotation_and_moments(gpu), "pic_rotation_and_moments");
		// evaluateElectricField();
		pic_solve_info info;
		check(gpu, pic_maxwell_solve(gpu, &info), "pic_maxwell_solve");
		// evaluateMagneticField();
		check(gpu, pic_update_b(gpu), "pic_update_b");
		// moveParticles(); eraseEscapedPaticles(); updateParticleCorrelationMaps(); updateDensityParameters();
		check(gpu, pic_push(gpu), "pic_push");
		// updateFields();
		check(gpu, pic_commit_fields(gpu), "pic_commit_fields");
		// updateEnergy();
		double e[7];
		check(gpu, pic_energy(gpu, e), "pic_energy");
		energy.insert(energy.end(), e, e + 7);
	}

	// read back for output: fields in the reference's layout, particles in the context's order
	check(gpu, (int)std::min<long long>(0, pic_download_array(gpu, PIC_ARRAY_E, E.data())), "pic_download_array E");
	check(gpu, (int)std::min<long long>(0, pic_download_array(gpu, PIC_ARRAY_B, B.data())), "pic_download_array B");
	for (long long q = 0; q < nnodes; ++q)
		for (int c = 0; c < 3; ++c) Ein[3 * q + c] = E[c * nnodes + q];
	for (long long q = 0; q < ncells; ++q)
		for (int c = 0; c < 3; ++c) Bin[3 * q + c] = B[c * ncells + q];
	const long long nout = pic_num_particles(gpu);
	std::vector<int> osp(nout), oid(nout);
	std::vector<double> ox(3 * nout), op(3 * nout), xo(3 * nout), po(3 * nout);
	check(gpu, pic_download_particles(gpu, osp.data(), oid.data(), ox.data(), op.data()), "pic_download_particles");
	for (long long q = 0; q < nout; ++q)
		for (int c = 0; c < 3; ++c) {
			xo[3 * q + c] = ox[c * nout + q];
			po[3 * q + c] = op[c * nout + q];
		}
	pic_destroy(gpu);
	write_all(dir + "out_E.bin", Ein);
	write_all(dir + "out_B.bin", Bin);
	write_all(dir + "out_ids.bin", oid);
	write_all(dir + "out_x.bin", xo);
	write_all(dir + "out_p.bin", po);
	write_all(dir + "out_energy.bin", energy);
	std::printf("pic_step_caller: %d steps, %lld particles\n", s.steps, nout);
	return 0;
}
