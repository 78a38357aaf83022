// updateEnergy of caller-owned device state without a context and without any allocation
// (pic_energy_noalloc). The partial sums live in a file-scope device array; each device holds its
// own copy, whose address is looked up once per device with hipGetSymbolAddress.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/picpp_amd.h"
#include "pic_device.hpp"
#include "pic_kernels.hpp"

using namespace picgpu;

namespace {

// [0, 5 B) field rows and [5 B, 9 B) particle rows of block partials (B = kEnergyBlocks), then the
// 9 row totals. Every launch writes all of it before reading it, whatever the input sizes, so
// nothing needs clearing and no size needs more room.
constexpr int kRows = 9;
constexpr long long kScratchDoubles = (long long)kRows * kEnergyBlocks + 16;

}  // namespace

// In host code `scratch` names a host-side stand-in, not device memory: only the address that
// hipGetSymbolAddress returns for the current device (device_scratch) is passed to kernels.
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex scratch_lock;
double* scratch_ptr[kMaxDevices] = {};

thread_local std::string last_error;

int refuse(int code, const char* what) {
	last_error = what;
	return code;
}

// the device address of `scratch` on the current device (looked up on the first call there)
double* device_scratch(int dev) {
	std::lock_guard<std::mutex> hold(scratch_lock);
	if (scratch_ptr[dev] == nullptr) {
		void* p = nullptr;
		if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return nullptr;
		scratch_ptr[dev] = static_cast<double*>(p);
	}
	return scratch_ptr[dev];
}

struct SpeciesStarts {
	long long start[kMaxSpecies + 1];
};

// Kinetic energy sum w (gamma - 1) m and particle momentum sum w p per block, the particles
// grouped by species (PIC++/simulation.cpp:1506-1640); the same loop and block reduction as the
// context's particle-energy kernel.
__global__ void __launch_bounds__(256) k_particle_energy_ranges(SpeciesTable sp, SpeciesStarts st, const double* __restrict__ p0,
                                                                const double* __restrict__ p1, const double* __restrict__ p2,
                                                                long long count, double* __restrict__ partial) {
	__shared__ double red[4][256];
	double acc[4] = {0, 0, 0, 0};
	for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x) {
		int s = 0;
		while (s + 1 < sp.count && i >= st.start[s + 1]) ++s;
		double m = sp.mass[s], w = sp.weight[s];
		double p[3] = {p0[i], p1[i], p2[i]};
		double pp = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
		double gm = sqrt(pp / (m * m) + 1.0);
		acc[0] += ((gm - 1) * m) * w;
		for (int c = 0; c < 3; ++c) acc[1 + c] += p[c] * w;
	}
	for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = acc[q];
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < w)
			for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		for (int q = 0; q < 4; ++q) partial[q * gridDim.x + blockIdx.x] = red[q][0];
}

// out7 from the 9 row totals: particle, E, B, total energy, momentum x, y, z (pic_energy's order)
__global__ void k_energy_combine(const double* __restrict__ rows, double* __restrict__ out7) {
	if (threadIdx.x != 0) return;
	out7[0] = rows[5];
	out7[1] = rows[0];
	out7[2] = rows[1];
	out7[3] = rows[5] + rows[0] + rows[1];
	out7[4] = rows[2] + rows[6];
	out7[5] = rows[3] + rows[7];
	out7[6] = rows[4] + rows[8];
}

}  // namespace

extern "C" {

int pic_energy_noalloc(const pic_config* cfg, const double* E, const double* B, const double* p, const long long* counts,
                       double* out7, void* stream) {
	if (!cfg || !E || !B || !counts || !out7) return refuse(PIC_ERR_ARGUMENT, "pic_energy_noalloc: null argument");
	if (cfg->nx < 5 || cfg->ny < 5 || cfg->nz < 5 || cfg->nspecies < 1 || cfg->nspecies > kMaxSpecies || cfg->dx <= 0 ||
	    cfg->dy <= 0 || cfg->dz <= 0)
		return refuse(PIC_ERR_ARGUMENT, "pic_energy_noalloc: grid or species outside the supported range");
	if (cfg->boundary_x != PIC_BOUNDARY_PERIODIC && cfg->boundary_x != PIC_BOUNDARY_WALL)
		return refuse(PIC_ERR_ARGUMENT, "pic_energy_noalloc: unknown x boundary");
	SpeciesStarts st{};
	long long np = 0;
	for (int s = 0; s < cfg->nspecies; ++s) {
		if (counts[s] < 0) return refuse(PIC_ERR_ARGUMENT, "pic_energy_noalloc: negative particle count");
		st.start[s] = np;
		np += counts[s];
	}
	for (int s = cfg->nspecies; s <= kMaxSpecies; ++s) st.start[s] = np;
	if (np > 0 && !p) return refuse(PIC_ERR_ARGUMENT, "pic_energy_noalloc: particles without a momentum array");
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
		return refuse(PIC_ERR_DEVICE, "pic_energy_noalloc: no usable current device");
	double* sc = device_scratch(dev);
	if (!sc) return refuse(PIC_ERR_DEVICE, "pic_energy_noalloc: hipGetSymbolAddress of the scratch array failed");

	Grid g{};
	const int n[3] = {cfg->nx, cfg->ny, cfg->nz};
	const double d[3] = {cfg->dx, cfg->dy, cfg->dz};
	for (int a = 0; a < 3; ++a) {
		g.n[a] = n[a];
		g.N[a] = n[a] + 4;
		g.d[a] = d[a];
		g.L[a] = n[a] * d[a];
		g.x0[a] = g.L[a] + (-2) * d[a];
	}
	g.volume = d[0] * d[1] * d[2];
	g.nnodes = (long long)(g.N[0] + 1) * (g.N[1] + 1) * (g.N[2] + 1);
	g.ncells = (long long)g.N[0] * g.N[1] * g.N[2];
	g.wall = cfg->boundary_x == PIC_BOUNDARY_WALL ? 1 : 0;
	g.nu0 = n[0] + 2 * g.wall;
	g.nown = (long long)g.nu0 * n[1] * n[2];
	SpeciesTable sp{};
	sp.count = cfg->nspecies;
	for (int s = 0; s < kMaxSpecies; ++s) {
		sp.mass[s] = s < cfg->nspecies ? cfg->mass[s] : 1.0;
		sp.charge[s] = s < cfg->nspecies ? cfg->charge[s] : 0.0;
		sp.weight[s] = s < cfg->nspecies ? cfg->weight[s] : 0.0;
	}
	hipStream_t s = static_cast<hipStream_t>(stream);
	double* partial = sc;
	double* rows = sc + (long long)kRows * kEnergyBlocks;
	launch_field_energy(g, E, B, partial, s);
	const double* p0 = p;
	const double* p1 = p ? p + np : nullptr;
	const double* p2 = p ? p + 2 * np : nullptr;
	hipLaunchKernelGGL(k_particle_energy_ranges, dim3(kEnergyBlocks), dim3(256), 0, s, sp, st, p0, p1, p2, np,
	                   partial + 5 * kEnergyBlocks);
	launch_rows_final(partial, kEnergyBlocks, kRows, rows, s);
	hipLaunchKernelGGL(k_energy_combine, dim3(1), dim3(64), 0, s, rows, out7);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return refuse(PIC_ERR_DEVICE, hipGetErrorString(e));
	last_error.clear();
	return PIC_OK;
}

}  // extern "C"

// what pic_last_error(NULL) reports: the failure of this thread's last context-free call
const char* picgpu::noalloc_last_error() { return last_error.empty() ? "null context" : last_error.c_str(); }
