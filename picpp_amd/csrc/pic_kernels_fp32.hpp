// FP32 particle kernels: integer cell + FP32 offset positions, u = gamma v momenta.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pic_device.hpp"

namespace picgpu {

struct ParticleArrays32 {
	int* cell;      // ghosted linear cell id
	float* off[3];  // offset within the cell, [0, 1)
	float* u[3];    // p / (m c)
	int* id;
	long long count;
};

int group_lanes(int nspecies);
void launch_node_moments32(const Grid& g, const SpeciesTable& sp, const ParticleArrays32& P, const int* start, const float* E,
                           const float* B, float dt, float theta, float* scratch, hipStream_t s);
// With dens_scratch given, the same pass also writes the post-push density planes of
// launch_density32 (the particles at x^{n+1}, E and B the fields of step n+1).
void launch_cell_moments32(const Grid& g, const SpeciesTable& sp, const ParticleArrays32& P, const int* start, const float* E,
                           const float* B, float dt, float* rho_scratch, float* pi_scratch, hipStream_t s,
                           float* dens_scratch = nullptr);
void launch_density32(const Grid& g, const SpeciesTable& sp, const ParticleArrays32& P, const int* start, float* scratch,
                      hipStream_t s);
// Per-launch constants of the FP32 push, formed on the host: q dt / m and q dt / (2 m) per
// species in float arithmetic, the cell sizes and their inverses, and the multipliers and
// shifts that divide a cell index by N[2] and N[1].
struct Push32Consts {
	float qdm[kMaxSpecies];
	float beta[kMaxSpecies];
	float inv_d[3], d[3];
	uint32_t m2, m1;
	int l2, l1;
};

// The push leaves a mover bit per particle in mask (word i >> 6, bit i & 63), written by each
// wave of 64 particles.
void launch_push32(const Grid& g, const SpeciesTable& sp, const ParticleArrays32& P, const int* species_start, const float* E,
                   const float* Et, const float* B, float dt, float eta, uint32_t* keys, uint64_t* mask, int* status, int* dead,
                   uint32_t* esc, uint64_t* bounce_mask, hipStream_t s);
// counts[w] = population of mask word w < nwords, counts[nwords] = 0
void launch_mask_counts(const uint64_t* mask, long long nwords, int* counts, hipStream_t s);
// mover bits of the particles lo .. hi-1 appended after the push
void launch_mark_movers(uint64_t* mask, long long lo, long long hi, hipStream_t s);
// prefix: exclusive scan of the word populations (nwords + 1 entries)
void launch_gather_movers(const uint64_t* mask, const int* prefix, const uint32_t* keys, long long n, long long nkeys, uint32_t* mk,
                          uint32_t* mi, hipStream_t s);
void launch_first_of_run(const uint32_t* mk, long long M, int* first, hipStream_t s);
void launch_rebin_counts(const int* start_old, const uint64_t* mask, const int* prefix, const int* arrive, long long nkeys, int* counts,
                         hipStream_t s);
// perm[new position] = old index: per segment its stayers, then its arrivals mi[first[k] ..]
void launch_rebin_place(const int* start_old, const int* start_new, const uint64_t* mask, const int* arrive, const int* first,
                        const uint32_t* mi, long long nkeys, uint32_t* perm, hipStream_t s);
void launch_keys32(const Grid& g, const SpeciesTable& sp, const ParticleArrays32& P, const int* species_start, uint32_t* keys,
                   hipStream_t s);
void launch_permute32(const ParticleArrays32& src, const ParticleArrays32& dst, const uint32_t* perm, hipStream_t s);
void launch_from_fp64(const Grid& g, const SpeciesTable& sp, const double* x, const double* p, long long np,
                      const int* species_start, const ParticleArrays32& P, int* status, hipStream_t s);
void launch_to_fp64(const Grid& g, const SpeciesTable& sp, const ParticleArrays32& P, const int* species_start, double* x,
                    double* p, hipStream_t s);
void launch_f64_to_f32(const double* a, float* b, long long n, hipStream_t s);
void launch_node_assemble32(const Grid& g, const float* scratch, double* J, double* mu, hipStream_t s);
void launch_cell_assemble32(const Grid& g, const float* scratch, int nq, double* out, hipStream_t s);
void launch_pressure_reduce32(const Grid& g, const float* pi_scratch, double* Pi, hipStream_t s);
void launch_particle_energy32(const SpeciesTable& sp, const ParticleArrays32& P, const int* species_start, double* partial,
                              hipStream_t s);

}  // namespace picgpu
