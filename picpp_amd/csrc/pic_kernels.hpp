// Launch wrappers of the HIP kernels (host-callable), grouped by phase.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pic_device.hpp"

namespace picgpu {

struct ParticleArrays {
	double* x[3];
	double* p[3];
	int* id;
	long long count;
};

// moments at x^n (evaluateParticlesRotationTensor + updateElectroMagneticParameters)
void launch_node_moments(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P, const int* start, const double* E,
                         const double* B, double dt, double theta, double* scratch, hipStream_t s);
void launch_node_assemble(const Grid& g, const double* scratch, double* J, double* mu, hipStream_t s);
void launch_cell_moments(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P, const int* start, const double* E,
                         const double* B, double dt, double* rho_scratch, double* pi_scratch, hipStream_t s);
void launch_cell_assemble(const Grid& g, const double* rho_scratch, double* rho, hipStream_t s);
void launch_pressure_reduce(const Grid& g, const double* pi_scratch, double* Pi, hipStream_t s);
// cell-local pressure tensor from the dual-cell partials (FP64 or FP32 partials)
void launch_pressure_cell(const Grid& g, const double* pi_scratch, double* Pi, hipStream_t s);
void launch_pressure_cell32(const Grid& g, const float* pi_scratch, double* Pi, hipStream_t s);
void launch_fold_nodes(const Grid& g, double* f, int ncomp, hipStream_t s);
void launch_fold_cells(const Grid& g, double* f, int ncomp, int pairs, int lo0, hipStream_t s);
// periodic ghost fill; with the wall, x ghosts behind it copy index 2 (or are zeroed with wall_zero)
void launch_ghost_fill(const Grid& g, double* f, int ncomp, bool nodes, hipStream_t s, bool wall_zero = false);
// smoothing passes (PIC++/smoothing.cpp): tmp holds at least ncomp node planes or one cell plane
struct Smooth27 {
	double m[27];
};
Smooth27 smoothing_matrix(double smoothing_parameter, bool collapsed_z = false);
void launch_smooth_cells(const Grid& g, double* f, double* tmp, const Smooth27& m, hipStream_t s);
void launch_smooth_nodes(const Grid& g, double* f, int ncomp, double* tmp, const Smooth27& m, hipStream_t s);
void launch_div_pressure_update_j(const Grid& g, const double* Pi, double* J, double* divPi, double eta, double dt, hipStream_t s);
void launch_div_flux_update_rho(const Grid& g, const double* J, double* rho, double dt, double theta, hipStream_t s);

// Maxwell system
void launch_rhs(const Grid& g, const double* E, const double* B, const double* J, const double* rho, double theta, double dt,
                double* rhs, hipStream_t s);
void launch_pack_mu(const Grid& g, const double* mu, double* muc, hipStream_t s);
// mu on the unknowns and one node around them (halo_count(g) nodes x 9), for the operators
void launch_pack_mu_halo(const Grid& g, const double* mu, double* muh, hipStream_t s);
void launch_apply_operator(const Grid& g, const double* muh, const double* v, double C, double* out, hipStream_t s);
void launch_dot(const double* a, const double* b, long long len, double* partial, double* out, bool do_sqrt, hipStream_t s);
void launch_axpy_dev(double* w, const double* v, const double* h, long long len, hipStream_t s);
void launch_normalize_dev(double* w, const double* h, long long len, hipStream_t s);
void launch_div_copy(double* b, double* v0, const double* norm, long long len, hipStream_t s);
void launch_combine(double* x, const double* const* V, const double* y, int m, const double* norm, long long len, hipStream_t s);
void launch_add(double* x, const double* d, long long len, hipStream_t s);
void launch_sub(double* r, const double* b, const double* y, long long len, hipStream_t s);
void launch_unpack_e(const Grid& g, const double* x, const double* E, double theta, double* Etheta, double* Enew, hipStream_t s);
void launch_update_b(const Grid& g, const double* Etheta, const double* B, double dt, double* Bnew, hipStream_t s);

// particles
// the push writes the new sort keys and one mover bit per particle (mask, as the FP32 push)
void launch_push(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P, const int* species_start, const double* E,
                 const double* Etheta, const double* B, double dt, double eta, uint32_t* keys, uint64_t* mask, int* status,
                 int* dead, uint32_t* esc, uint64_t* bounce_mask, hipStream_t s);
// wall boundary helpers: cell planes 0, 1 folded into their mirrors 3, 2 (flip[c] negates
// component c), the right-end fold of cells N-2, N-1 into N-3 (zero_mask[c] zeroes them), and
// zeroing of node planes i0..i1 for components [c0, c1)
void launch_wall_mirror_cells(const Grid& g, double* f, int ncomp, const int* flip, hipStream_t s);
void launch_wall_right_fold(const Grid& g, double* f, int ncomp, const int* zero_mask, hipStream_t s);
void launch_zero_node_planes(const Grid& g, double* f, int c0, int c1, int i0, int i1, hipStream_t s);
void launch_keys(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P, const int* species_start, uint32_t* keys,
                 int* status, hipStream_t s);
void launch_permute(const ParticleArrays& src, const ParticleArrays& dst, const uint32_t* perm, hipStream_t s);
void launch_count_keys(const uint32_t* keys, long long count, int* counts, hipStream_t s);
void launch_iota(uint32_t* v, long long count, hipStream_t s);
void launch_density(const Grid& g, const SpeciesTable& sp, const ParticleArrays& P, const int* start, double* scratch,
                    hipStream_t s);
void launch_density_assemble(const Grid& g, const double* scratch, int nq, double* out, hipStream_t s);
void launch_bulk_divide(const Grid& g, double* dens, int nspecies, hipStream_t s);

// energy
constexpr int kEnergyBlocks = 512;
void launch_field_energy(const Grid& g, const double* E, const double* B, double* partial, hipStream_t s);
void launch_particle_energy(const SpeciesTable& sp, const ParticleArrays& P, const int* species_start, double* partial,
                            hipStream_t s);
void launch_rows_final(const double* partial, int n, int rows, double* out, hipStream_t s);

constexpr int kDotBlocks = 1024;

// the failure text of this thread's last context-free entry point (pic_energy_noalloc)
const char* noalloc_last_error();

}  // namespace picgpu
