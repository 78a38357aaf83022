// The explicit (BUNEMAN / Tristan) step on the Yee grid: launch wrappers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pic_device.hpp"
#include "pic_kernels.hpp"

namespace picgpu {

// Ghosted Yee ("Buneman") arrays in the reference's layouts (PIC++/initialization.cpp:
// 7146-7174): E and J component c have N cells along axis c and N + 1 nodes along the others;
// B component c has N + 1 along axis c and N along the others. Index ((i d1) + j) d2 + k.
struct YeeGrid {
	double* E[3];
	double* B[3];
	double* J[3];
	long long* Jacc[3];  // fixed-point current accumulators, J's layout
	int dE[3][3], dB[3][3];
};

void yee_dims(const Grid& g, bool b_field, int c, int dims[3]);
inline long long yee_count(const int d[3]) { return (long long)d[0] * d[1] * d[2]; }

// tristanEvaluateBhalfStep: B -= c' dt / 2 curl E on the owned ranges
void launch_tristan_b_half(const Grid& g, const YeeGrid& y, double dt, hipStream_t s);
// tristanEvaluateE: E += (c' curl B - 4 pi J) dt on the owned ranges
void launch_tristan_e(const Grid& g, const YeeGrid& y, double dt, hipStream_t s);
// exchangeBuneman{E,B}field, one-rank periodic: b_field selects B, else E
void launch_yee_exchange(const Grid& g, const YeeGrid& y, bool b_field, hipStream_t s);
// the fixed-point current to J, then exchangeBunemanFlux's periodic folds
void launch_yee_current(const Grid& g, const YeeGrid& y, double inv_scale, hipStream_t s);
// largest number of particles of all species in one cell (start: segment starts per key)
void launch_max_cell_count(const Grid& g, int nspecies, const int* start, int* out, hipStream_t s);
// moveParticleTristan for every particle, the zig-zag current of the move into Jacc (fixed
// point, scale a power of two), the one-rank periodic wrap, the new sort key and one mover bit
// per particle in mask (bit i & 63 of word i >> 6: the key changed)
void launch_tristan_push(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P, const int* species_start,
                         const YeeGrid& y, double dt, double scale, uint32_t* keys, uint64_t* mask, int* status,
                         hipStream_t s);

}  // namespace picgpu
