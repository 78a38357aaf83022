// Mixed-precision Krylov kernels (FP32 inner GMRES, FP64 refinement).
#pragma once
#include <hip/hip_runtime.h>

#include "pic_device.hpp"

namespace picgpu {

constexpr int kSolverDotBlocks = 1024;

// out = A v in factored form: w = mu v on the halo layout (scratch of 3 halo_count(g) values,
// mu in the halo layout of launch_pack_mu_halo), then the stencil on v and w.
template <class T>
void launch_apply_factored(const Grid& g, const T* mu, const T* v, T* w, double C, T* out, hipStream_t s);
template <class T>
void launch_dotT(const T* a, const T* b, long long len, double* partial, double* out, bool do_sqrt, hipStream_t s);
template <class T>
void launch_axpyT(T* w, const T* v, const double* h, long long len, hipStream_t s);
template <class T>
void launch_normalizeT(T* w, const double* h, long long len, hipStream_t s);
template <class T>
void launch_scaled_copy(const double* r, T* v0, const double* norm, long long len, hipStream_t s);
template <class T>
void launch_update_x(double* x, const T* const* V, const double* y, int m, const double* scale, long long len, hipStream_t s);
void launch_residual(const double* b, const double* Ax, double* r, long long len, hipStream_t s);
void launch_to_f32(const double* a, float* b, long long n, hipStream_t s);

}  // namespace picgpu
