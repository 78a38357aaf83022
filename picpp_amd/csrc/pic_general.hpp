// Diagnostics of the reference's general.dat rows (outputGeneral, PIC++/output.cpp:3215-3250):
// the field sums of updateEnergy and updateParameters, and the theoretical energy and momentum
// that the step loop keeps (updateTheoreticalEnergy, PIC++/simulation.cpp:1399-1504, and the
// mover's and the injection's updates of the same members).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pic_device.hpp"
#include "pic_explicit.hpp"
#include "pic_kernels.hpp"
#include "pic_kernels_fp32.hpp"

namespace picgpu {

// Three E and three B component arrays, each indexed ((i d1) + j) d2 + k: the implicit path's
// planes (E on nodes, B on cells) or the Yee arrays.
struct FieldView {
	const double* e[3];
	const double* b[3];
	int de[3][3], db[3][3];
	bool average_e;  // the momentum's E is the mean of the cell's 8 nodes (implicit E on nodes)
};
FieldView implicit_view(const Grid& g, const double* E, const double* B);
FieldView yee_view(const YeeGrid& y);

// Field sums over updateEnergy's range (x from 2, or 0 with the wall, to N - 2, or N - 1 with
// the wall; y and z 2 .. N - 3): rows 0..2 sum E_c^2, 3..5 sum B_c^2, 6..8 sum (E x B)_c, 9 the
// largest |E|^2, 10 the largest |B|^2. partial holds kGeneralRows x kEnergyBlocks doubles; out
// gets the rows, in a fixed order (the same bits in every run).
constexpr int kGeneralRows = 11;
void launch_general_fields(const Grid& g, const FieldView& v, double* partial, double* out, hipStream_t s);

// theo[0] += sign sum w (gamma - 1) m, theo[1..3] += sign sum w p over the particles idx[0 .. count)
// (idx null: first .. first + count - 1). A particle's species comes from the species ranges
// below species_start[nspecies], from its sort key above (appended particles).
void launch_theo_particles(const SpeciesTable& sp, const ParticleArrays& P, const ParticleArrays32& Q, bool fp32,
                           const int* species_start, const uint32_t* keys, long long ncells, const uint32_t* idx,
                           long long first, long long count, double sign, double* theo, hipStream_t s);
// theo[1] += sum over the particles with a bit in bounce_mask of 2 w p_x (their momentum after the
// reflection; the mover's theoreticalMomentum.x update, PIC++/mover.cpp:256, 329, 439). partial
// holds kEnergyBlocks doubles.
void launch_theo_bounce(const SpeciesTable& sp, const ParticleArrays& P, const ParticleArrays32& Q, bool fp32,
                        const int* species_start, const uint64_t* bounce_mask, long long n, double* partial, double* theo,
                        hipStream_t s);
// theo += the Poynting flux through the x boundaries over one step: per (j, k) in 2 .. N - 3,
// (E x B) at x index 2 minus (E x B) at N - 2, times dt dy dz / (4 pi) (energy: the x component).
void launch_theo_poynting(const Grid& g, const FieldView& v, double dt, double* theo, hipStream_t s);

}  // namespace picgpu
