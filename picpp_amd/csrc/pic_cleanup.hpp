// Divergence cleanup and Fourier filtering kernels (pic_cleanup.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "pic_device.hpp"

namespace picgpu {

// weights of the 27-point cleanup operator, index (a+1)*9 + (b+1)*3 + (c+1) for offsets a, b, c
struct CleanupStencil {
	double w[27];
};
CleanupStencil cleanup_stencil(const Grid& g);

// rhs (owned cells) = -(4 pi rho - div E); E ghosted node field, rho ghosted cell plane
void launch_cleanup_rhs(const Grid& g, const double* E, const double* rho, double* rhs, hipStream_t s);
// out = L v on owned cells (periodic)
void launch_cleanup_apply(const Grid& g, const CleanupStencil& S, const double* v, double* out, hipStream_t s);
// E -= grad phi at the owned nodes (ghosts are left for the caller's exchange)
void launch_cleanup_correct(const Grid& g, const double* phi, double* E, hipStream_t s);
// magnetic cleanup: rhs (owned nodes) = div B at the nodes; B -= grad phi at the owned cells
void launch_cleanup_rhs_b(const Grid& g, const double* B, double* rhs, hipStream_t s);
void launch_cleanup_correct_b(const Grid& g, const double* phi, double* B, hipStream_t s);
// filterFieldGeneral on component comp (plane offset `plane`) of a ghosted node or cell field:
// tw[a] holds (cos, sin)(2 pi q / n_a), q < n_a; re, im, re2, im2 are owned-size scratch
void launch_filter_component(const Grid& g, double* f, long long plane, bool nodes, int comp, double kcut, const double* const tw[3],
                             double* re, double* im, double* re2, double* im2, hipStream_t s);

}  // namespace picgpu
