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
// one pass of smoothBunemanEfieldGeneral on J (tmp: three scratch arrays of J's shapes), then
// exchangeBunemanEfield on J
void launch_smooth_yee_current(const Grid& g, const YeeGrid& y, const Smooth27& sm, double* const tmp[3], hipStream_t s);
// What the fixed-point scale of the explicit push's current is set from: the largest number of
// particles of all species in one cell, per species the largest |u|^2 = |p|^2 / m^2 of its
// particles, and per component the largest |E| on the Yee grid (ghosts included). The maxima of
// non-negative doubles are kept as their bit patterns, which order like the values.
struct ExplicitBounds {
	int maxcount;
	int pad;
	unsigned long long u2max[kMaxSpecies];
	unsigned long long emax[3];
};
// The per-species maxima are gathered in kU2Slots rows of kMaxSpecies (atomics spread over the
// rows by workgroup) and folded into ExplicitBounds::u2max.
constexpr int kU2Slots = 256;
// Fills out (cleared by the caller): the cell count (start: segment starts per key) and |E|;
// when `particles`, the species' |u|^2 from the particles into u2_slots (cleared by the caller;
// else the slots hold what the last push left); then u2max from the slots.
void launch_explicit_bounds(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P, const int* species_start,
                            const int* start, const YeeGrid& y, bool particles, unsigned long long* u2_slots,
                            ExplicitBounds* out, hipStream_t s);
// The |u|^2 of particles added after a push (injection, replacement) into the same rows, so that
// the next step's bound covers them: idx[0 .. count) or first .. first + count - 1.
void launch_species_u2_added(const SpeciesTable& sp, const ParticleArrays& P, const int* species_start, const uint32_t* keys,
                             long long ncells, const uint32_t* idx, long long first, long long count, unsigned long long* u2_slots,
                             hipStream_t s);
// moveParticleTristan for every particle, the zig-zag current of the move into Jacc (fixed
// point, scale a power of two), the one-rank periodic wrap, the new sort key and one mover bit
// per particle in mask (bit i & 63 of word i >> 6: the key changed). Per species, the largest
// new |u|^2 goes to the rows of u2_slots (cleared by the caller) for the next step's scale; a particle whose
// new |u|^2 exceeds u2_limit[species] (the bound the scale was set for) sets status bit 16.
// With the wall (g.wall): particles reflect at x = xgrid[20] (a bit in bounce_mask), and those
// past node N - 2 leave: key 0xFFFFFFFF, listed in esc (count in dead), y and z unwrapped.
void launch_tristan_push(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P, const int* species_start,
                         const YeeGrid& y, double dt, double scale, const double* u2_limit, uint32_t* keys, uint64_t* mask,
                         int* status, unsigned long long* u2_slots, int* dead, uint32_t* esc, uint64_t* bounce_mask,
                         hipStream_t s);

}  // namespace picgpu
