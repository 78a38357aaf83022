// Injection of the flow at the free right end of a wall-bounded subdomain, and replacement
// of the particles that leave there (the reference's injectNewParticles,
// PIC++/mover.cpp:910-1000, and addToPreserveChargeGlobal, PIC++/initialization.cpp:8281-8338).
#pragma once
#include "pic_device.hpp"
#include "pic_kernels.hpp"
#include "pic_kernels_fp32.hpp"

#include <vector>

namespace picgpu {

constexpr int kMaxSlabs = 96;

// The slabs of one step: each adds one particle to every (y, z) column of owned cells at
// x = x[slab], all of the slab's particles at the same offset inside their cells.
struct InjectSlabs {
	int count;
	int species[kMaxSlabs];
	double x[kMaxSlabs];
	unsigned slab[kMaxSlabs];  // running slab number: the slab's random stream
};

// Physics and random streams shared by injection and replacement.
// Points of the tabulated Maxwell-Juttner CDF of a species (randomParameter, PIC++/constants.h:8).
constexpr int kJuttnerPoints = 4096;

struct InjectPhysics {
	double flow[3];               // V0, |V0| < 1
	double theta[kMaxSpecies];    // kT / (m c^2) of each species
	double theta_run[kMaxSpecies];  // the run's kT / (m_s c^2): the Juttner branch choice
	// for species with theta >= 0.1 and theta_run < 2: kJuttnerPoints values u then the
	// normalised CDF per species (evaluateJuttnerFunction, PIC++/initialization.cpp:7372-7394)
	const double* juttner;
	unsigned long long seed;      // Philox key
	unsigned long long step;      // step number: first counter word
	int id_base;                  // id of the first particle this launch creates
};

// The table of one species: u_i = i 20 theta / (points - 1) and the running integral of
// u^2 exp(-sqrt(1 + u^2) / theta) over 10 sub-steps per interval, normalised to end at 1
// (evaluateJuttnerFunction's arithmetic). out: values then CDF, 2 kJuttnerPoints doubles.
void juttner_table(double theta, double* out);

// The slabs of one step, in the reference's order (species, then slab), for species with
// ppc[s] > 0: injectionLength len[s] grows by |V0_x| dt, and while ppc len >= dx a slab is
// added at x(n0+2) - len after len -= dx / ppc (PIC++/simulation.cpp:281-283,
// mover.cpp:910-925). Returned in launches of at most kMaxSlabs slabs; slab_counter numbers
// the slabs of the run.
std::vector<InjectSlabs> plan_injection(const Grid& g, int nspecies, const int* ppc, double flow_x, double dt, double* len,
                                        unsigned& slab_counter);

// Writes count * n1 * n2 new particles at indices first.. (FP64 form into P, or FP32 form
// into Q) with their sort keys.
void launch_inject(const Grid& g, const SpeciesTable& sp, const InjectSlabs& slabs, const InjectPhysics& ph,
                   const ParticleArrays& P, const ParticleArrays32& Q, bool fp32, long long first, uint32_t* keys,
                   hipStream_t s);

// Replaces the particles at indices esc[0..count) (ascending) in place: same species, at the
// right boundary x = x(n0+2), y and z of the escaped particle (drawn again inside the box when
// outside it), fresh flow momentum, ids id_base + r, new sort keys. In the FP32 form the push
// leaves the escaped particle's absolute y, z in cell units in off[1], off[2].
void launch_replace(const Grid& g, const SpeciesTable& sp, const InjectPhysics& ph, const int* species_start,
                    const uint32_t* esc, int count, const ParticleArrays& P, const ParticleArrays32& Q, bool fp32,
                    uint32_t* keys, hipStream_t s);

// out[4 * i + w] = Philox4x64-10 of counter (c0, c1 + i, c2, c3) and key (k0, k1), i < n:
// the generator the injection draws from, exposed for tests.
void launch_philox(unsigned long long c0, unsigned long long c1, unsigned long long c2, unsigned long long c3,
                   unsigned long long k0, unsigned long long k1, long long n, unsigned long long* out, hipStream_t s);

}  // namespace picgpu
