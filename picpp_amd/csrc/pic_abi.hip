// C ABI of the implicit PIC hot path: context, transfers and the Simulation-level phases.
#include <hip/hip_runtime.h>
#include <climits>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>

#include "../../include/picpp_amd.h"
#include "pic_device.hpp"
#include "pic_kernels.hpp"
#include "pic_explicit.hpp"
#include "pic_kernels_fp32.hpp"
#include "pic_inject.hpp"
#include "pic_precond.hpp"
#include "pic_solver.hpp"
#include "pic_cleanup.hpp"
#include "pic_general.hpp"

using namespace picgpu;

namespace {
constexpr double kCprime2 = 1.025 * 1.025;  // speed_of_light_correction_sqr (PIC++/constants.h:46)

struct DeviceError {
	int code;
	std::string what;
};

#define HIP_CHECK(expr)                                                                          \
	do {                                                                                         \
		hipError_t e_ = (expr);                                                                  \
		if (e_ != hipSuccess) throw DeviceError{PIC_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
	} while (0)

template <class T>
T* dalloc(size_t n) {
	void* p = nullptr;
	if (n == 0) n = 1;
	hipError_t e = hipMalloc(&p, n * sizeof(T));
	if (e != hipSuccess) throw DeviceError{PIC_ERR_MEMORY, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed"};
	return static_cast<T*>(p);
}

}  // namespace

struct pic_ctx {
	pic_config cfg;
	Grid g;
	SpeciesTable sp;
	hipStream_t stream = nullptr;
	bool own_stream = false;
	std::string error;
	double theta = 0.5, eta = 0.5;  // initialTheta, eta (PIC++/constants.h:41, initialization.cpp:625)

	// ghosted fields (component planes)
	double *E = nullptr, *Etheta = nullptr, *Enew = nullptr, *B = nullptr, *Bnew = nullptr;
	double *J = nullptr, *mu = nullptr, *divPi = nullptr, *rhohat = nullptr, *Pi = nullptr, *dens = nullptr;
	// scratch of the deterministic deposits
	double *node_scratch = nullptr, *rho_scratch = nullptr, *pi_scratch = nullptr, *dens_scratch = nullptr;
	// particles (FP64 form, or the FP32 cell + offset form when fp32)
	bool fp32 = false;
	ParticleArrays P{}, Palt{};
	ParticleArrays32 Q{}, Qalt{};
	float *E32 = nullptr, *Et32 = nullptr, *B32 = nullptr;
	float *nscr32 = nullptr, *rscr32 = nullptr, *pscr32 = nullptr, *dscr32 = nullptr;
	long long capacity = 0;
	uint32_t *keys = nullptr, *keys_sorted = nullptr, *perm = nullptr, *perm_sorted = nullptr;
	int *start = nullptr, *counts = nullptr;
	long long nkeys = 0;
	// rebin of the FP32 particles after a push: mover flags, their prefix count, the movers
	// (key, old index) and per-key arrival counts
	uint64_t* mask = nullptr;  // FP32 push: mover bit per particle
	int *wcount = nullptr, *wprefix = nullptr;  // mask word populations and their exclusive scan
	int *start_new = nullptr, *arrive = nullptr, *first = nullptr;
	uint32_t *mk = nullptr, *mi = nullptr, *mk_s = nullptr, *mi_s = nullptr;
	long long mover_capacity = 0;
	int end_bit = 0;
	int mover_bit = 0;  // radix bits of the mover keys, which include the removal key nkeys
	// FP32: the post-push density pass also formed the next step's cell moments (rscr32,
	// pscr32) and FP32 fields (E32, B32) from Enew, Bnew. 1: after pic_push; 2: pic_commit_fields
	// made E, B equal to them, and the next pic_rotation_and_moments uses them. Every other call
	// that can change particles or fields resets it to 0.
	int fused = 0;
	// PIC_SCHEME_EXPLICIT: the Yee arrays and the bounds of the current's fixed-point scale.
	// bounds->u2max holds each species' largest |u|^2 as the last explicit push left it
	// (u2_valid, in u2_slots); every entry point that can change particles clears u2_valid, and the next step
	// then reads it from the particles.
	bool explicit_scheme = false;
	YeeGrid yee{};
	ExplicitBounds* bounds = nullptr;
	unsigned long long* u2_slots = nullptr;  // kU2Slots x kMaxSpecies
	bool u2_valid = false;
	// general.dat bookkeeping: theo = the theoretical energy and momentum (4 doubles, device),
	// formed from the energies when the run starts (theo_init) and updated by every push;
	// gpartial / gout the field sums of pic_general; bounce one bit per particle reflected at
	// the wall by the last push
	double* theo = nullptr;
	bool theo_init = false;
	double* juttner = nullptr;  // the injected species' Juttner tables (kMaxSpecies x 2 x kJuttnerPoints)
	double* gpartial = nullptr;
	double* gout = nullptr;
	uint64_t* bounce = nullptr;
	void* sort_temp = nullptr;
	size_t sort_temp_bytes = 0;
	std::vector<int> species_start_h;
	int* species_start = nullptr;
	int* status = nullptr;
	// wall boundary: particles removed at the free right end, and the plane masks of the density
	// array (5 per species + charge) for the mirror (negated x momentum) and the right-end fold
	int* dead = nullptr;
	int *dens_flip = nullptr, *dens_zero = nullptr;
	// indices of the particles that left through the right end (push order, then ascending)
	uint32_t *esc = nullptr, *esc_sorted = nullptr;
	int esc_bits = 1;
	// injection at the right end: injectionLength per species, step and slab numbers of the
	// random streams, next particle id
	double inj_len[kMaxSpecies] = {};
	unsigned long long inj_step = 0;
	unsigned inj_slabs = 0;
	long long next_id = 0;
	// smoothing scratch: 9 node planes
	double* smooth_tmp = nullptr;
	// divergence cleanup (allocated by its first call): the GMRES basis and vectors over the owned
	// cells; Fourier filtering: two complex owned-size buffers and the twiddles of each axis
	std::vector<double*> Vcl;
	double** Vcl_dev = nullptr;
	double *cl_rhs = nullptr, *cl_phi = nullptr;
	double *ft_re = nullptr, *ft_im = nullptr, *ft_re2 = nullptr, *ft_im2 = nullptr, *ft_tw[3] = {nullptr, nullptr, nullptr};
	// Krylov solver
	double* muh = nullptr;    // mu on the halo layout of the operators
	float* mu32h = nullptr;
	double *rhs = nullptr, *muc = nullptr, *xsol = nullptr, *work = nullptr, *work2 = nullptr, *partial = nullptr, *dscal = nullptr, *ydev = nullptr;
	std::vector<double*> V;
	double** Vdev = nullptr;
	int max_basis = 0;
	// mixed-precision solve: FP32 basis, operator and scratch; FP64 operator scratch
	std::vector<float*> V32;
	float** V32dev = nullptr;
	float *mu32 = nullptr, *w32 = nullptr;
	double* w64 = nullptr;
	// learned preconditioner: FGMRES basis Z = M(V) inside the mixed solver
	std::unique_ptr<UNet> net;
	double mu_scale = 0.0;
	bool use_net = false;
	std::vector<float*> Z32;
	float** Z32dev = nullptr;
	// training pairs (residual r, correction e with A e = r) logged from solves
	bool logging = false;
	int log_capacity = 0, log_count = 0;
	float *log_r = nullptr, *log_e = nullptr;  // [capacity][3 nown]
	std::vector<float*> log_mu;                 // dielectric tensor of each logged solve [9 nown]
	std::vector<int> log_mu_of;                 // pair -> its solve
	float* tmp32 = nullptr;
	// energies
	double* epartial = nullptr;
	double* eout = nullptr;
	// per-kernel timing with events on the context's stream (the reference's `timing` prints)
	struct Pending {
		std::string name;
		hipEvent_t a, b;
		double bytes;
	};
	struct Stat {
		double ms = 0, bytes = 0;
		long long count = 0;
	};
	bool profiling = false;
	std::vector<hipEvent_t> ev_pool;
	size_t ev_next = 0;
	std::vector<Pending> pending;
	std::map<std::string, Stat> prof;

	long long nown3() const { return 3 * g.nown; }
	long long np() const { return fp32 ? Q.count : P.count; }
};

namespace {

// status words: [0] particle check bits, [1] record claimed, [2..4] the FP32 mover's first failing
// particle: failed checks, index, species
constexpr int kStatusWords = 32;

int fail(pic_ctx* c, const DeviceError& e) {
	if (c) c->error = e.what;
	return e.code;
}

void flush_profile(pic_ctx* c) {
	if (c->pending.empty()) return;
	HIP_CHECK(hipStreamSynchronize(c->stream));
	for (auto& p : c->pending) {
		float ms = 0;
		HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
		auto& st = c->prof[p.name];
		st.ms += ms;
		st.bytes += p.bytes;
		st.count += 1;
	}
	c->pending.clear();
	c->ev_next = 0;
}

// Records events around the enclosed launches when profiling is on; `bytes` is the
// algorithmic HBM traffic of one launch.
struct Timed {
	pic_ctx* c;
	const char* name;
	double bytes;
	hipEvent_t a{}, b{};
	bool on;
	Timed(pic_ctx* c_, const char* n, double by) : c(c_), name(n), bytes(by), on(c_->profiling) {
		if (!on) return;
		if (c->ev_next + 2 > c->ev_pool.size()) flush_profile(c);
		a = c->ev_pool[c->ev_next++];
		b = c->ev_pool[c->ev_next++];
		HIP_CHECK(hipEventRecord(a, c->stream));
	}
	~Timed() {
		if (!on) return;
		if (hipEventRecord(b, c->stream) == hipSuccess) c->pending.push_back({name, a, b, bytes});
	}
};

void check_status(pic_ctx* c) {
	int st[kStatusWords] = {};
	HIP_CHECK(hipMemcpyAsync(st, c->status, sizeof(st), hipMemcpyDeviceToHost, c->stream));
	HIP_CHECK(hipStreamSynchronize(c->stream));
	if (st[0]) {
		HIP_CHECK(hipMemsetAsync(c->status, 0, sizeof(st), c->stream));
		std::string why;
		if (st[0] & 1) why += " particle moved more than one box length;";
		if (st[0] & 2) why += " particle maps inconsistent (moved more than half a cell?);";
		if (st[0] & 4) why += " particle outside the owned region;";
		if (st[0] & 8) why += " particle stencil outside the ghost layers;";
		if (st[0] & 16)
			why += " a particle's new momentum exceeds the bound the explicit current's fixed-point scale was set for "
			       "(momenta changed through device pointers since the last step?), the current is not valid;";
		if (st[1] && c->fp32) {  // the FP32 mover's first failing particle, still in its pre-push state
			const long long i = st[3];
			int cell = 0;
			float off[3] = {0, 0, 0}, u[3] = {0, 0, 0};
			if (i >= 0 && i < c->Q.count) {
				HIP_CHECK(hipMemcpy(&cell, c->Q.cell + i, sizeof(int), hipMemcpyDeviceToHost));
				for (int a = 0; a < 3; ++a) {
					HIP_CHECK(hipMemcpy(&off[a], c->Q.off[a] + i, sizeof(float), hipMemcpyDeviceToHost));
					HIP_CHECK(hipMemcpy(&u[a], c->Q.u[a] + i, sizeof(float), hipMemcpyDeviceToHost));
				}
			}
			const Grid& g = c->g;
			const int ck = cell % g.N[2], cj = (cell / g.N[2]) % g.N[1], ci = cell / (g.N[2] * g.N[1]);
			char buf[512];
			snprintf(buf, sizeof(buf),
			         " first: checks %d (1 start stencil, 2 Picard stencil, 4/8/16 new cell x/y/z), index %lld, species %d, "
			         "cell (%d %d %d), offset (%.9g %.9g %.9g), u (%.9g %.9g %.9g)",
			         st[2], i, st[4], ci, cj, ck, off[0], off[1], off[2], u[0], u[1], u[2]);
			why += buf;
		}
		throw DeviceError{PIC_ERR_PARTICLE, "particle check failed:" + why};
	}
}

// Species ranges after a removal or an injection: species q starts at start[] of its first key.
void species_ranges(pic_ctx* c, long long n) {
	hipStream_t s = c->stream;
	for (int q = 0; q < c->sp.count; ++q)
		HIP_CHECK(hipMemcpyAsync(&c->species_start_h[q], c->start + (((long long)q * c->g.ncells) << 3), sizeof(int),
		                         hipMemcpyDeviceToHost, s));
	HIP_CHECK(hipStreamSynchronize(s));
	for (int q = c->sp.count; q <= kMaxSpecies; ++q) c->species_start_h[q] = (int)n;
	HIP_CHECK(hipMemcpyAsync(c->species_start, c->species_start_h.data(), (kMaxSpecies + 1) * sizeof(int), hipMemcpyHostToDevice, s));
	HIP_CHECK(hipStreamSynchronize(s));
}

// Sort particles by key (stable radix sort: deterministic), rebuild start[]. The `removed`
// particles the push took out (wall boundary: key 0xFFFFFFFF) sort to the end and are
// dropped; after a removal or an injection (`resized`) the species ranges are read back from
// start[].
void sort_particles(pic_ctx* c, int removed = 0, bool resized = false) {
	long long n = c->np();
	hipStream_t s = c->stream;
	if (n > 0) {
		{
			Timed t(c, "sort_radix", (double)n * 16);
			launch_iota(c->perm, n, s);
			size_t bytes = c->sort_temp_bytes;
			HIP_CHECK(rocprim::radix_sort_pairs(c->sort_temp, bytes, c->keys, c->keys_sorted, c->perm, c->perm_sorted,
			                                    (size_t)n, 0, c->end_bit, s));
		}
		if (c->fp32) {
			Timed t(c, "permute", (double)n * (2 * 32 + 4));
			launch_permute32(c->Q, c->Qalt, c->perm_sorted, s);
			std::swap(c->Q, c->Qalt);
		} else {
			Timed t(c, "permute", (double)n * (2 * 52 + 4));
			launch_permute(c->P, c->Palt, c->perm_sorted, s);
			std::swap(c->P, c->Palt);
		}
	}
	if (removed > 0) {
		n -= removed;
		c->P.count = c->Palt.count = c->fp32 ? 0 : n;
		c->Q.count = c->Qalt.count = c->fp32 ? n : 0;
	}
	Timed t(c, "segments", (double)n * 4 + (double)(c->nkeys + 1) * 12);
	HIP_CHECK(hipMemsetAsync(c->counts, 0, (c->nkeys + 1) * sizeof(int), s));
	launch_count_keys(c->keys_sorted, n, c->counts, s);
	size_t bytes = c->sort_temp_bytes;
	HIP_CHECK(rocprim::exclusive_scan(c->sort_temp, bytes, c->counts, c->start, 0, (size_t)(c->nkeys + 1), rocprim::plus<int>(), s));
	if (removed > 0 || resized) species_ranges(c, n);
}

void set_count(pic_ctx* c, long long n) {
	c->P.count = c->Palt.count = c->fp32 ? 0 : n;
	c->Q.count = c->Qalt.count = c->fp32 ? n : 0;
}

InjectPhysics inject_physics(const pic_ctx* c) {
	InjectPhysics ph{};
	for (int a = 0; a < 3; ++a) ph.flow[a] = c->cfg.flow[a];
	for (int s = 0; s < kMaxSpecies; ++s) {
		ph.theta[s] = s < c->sp.count ? c->cfg.temperature[s] : 0.0;
		ph.theta_run[s] = s < c->sp.count ? c->cfg.run_kT / c->sp.mass[s] : 0.0;
	}
	ph.juttner = c->juttner;
	ph.seed = c->cfg.seed;
	ph.step = c->inj_step;
	return ph;
}

void rebin_particles(pic_ctx* c, int removed = 0, bool resized = false);

// After a push with the wall boundary (PIC++/simulation.cpp:258-289 order: erase and exchange,
// injection, preserve-charge replacement): particles that left through the right end are
// dropped, or with injection on replaced in place; the injected slabs are appended; the rebin
// restores the cell order and the species ranges.
void wall_particles(pic_ctx* c) {
	hipStream_t s = c->stream;
	int removed = 0;
	HIP_CHECK(hipMemcpyAsync(&removed, c->dead, sizeof(int), hipMemcpyDeviceToHost, s));
	HIP_CHECK(hipStreamSynchronize(s));
	HIP_CHECK(hipMemsetAsync(c->dead, 0, sizeof(int), s));
	bool resized = removed > 0;
	const int escaped = removed;
	if (escaped > 0) {
		// the escapes in index order (the push lists them in atomic order), and their energy and
		// momentum out of the theoretical totals (updateTheoreticalEnergy's escapedParticlesRight)
		size_t bytes = c->sort_temp_bytes;
		HIP_CHECK(rocprim::radix_sort_keys(c->sort_temp, bytes, c->esc, c->esc_sorted, (size_t)escaped, 0, c->esc_bits, s));
		launch_theo_particles(c->sp, c->P, c->Q, c->fp32, c->species_start, c->keys, c->g.ncells, c->esc_sorted, 0, escaped, -1.0,
		                      c->theo, s);
	}
	if (c->cfg.inject) {
		long long n = c->np();
		InjectPhysics ph = inject_physics(c);
		{
			Timed t(c, "inject", 0);
			std::vector<InjectSlabs> plan =
			    plan_injection(c->g, c->sp.count, c->cfg.ppc, c->cfg.flow[0], c->cfg.dt, c->inj_len, c->inj_slabs);
			const long long ncols = (long long)c->g.n[1] * c->g.n[2];
			long long ninj = 0;
			for (const InjectSlabs& b : plan) ninj += b.count * ncols;
			if (n + ninj > c->capacity)
				throw DeviceError{PIC_ERR_MEMORY, "particle capacity exceeded by injection (" + std::to_string(n + ninj) + " > " +
				                                      std::to_string(c->capacity) + ")"};
			if (c->next_id + ninj + removed > 0x7FFFFFFFLL) throw DeviceError{PIC_ERR_MEMORY, "particle ids exhausted"};
			long long at = n;
			for (const InjectSlabs& b : plan) {
				ph.id_base = (int)c->next_id;
				launch_inject(c->g, c->sp, b, ph, c->P, c->Q, c->fp32, at, c->keys, s);
				// the new particles into the theoretical totals (injectNewParticles, PIC++/mover.cpp:994-996)
				launch_theo_particles(c->sp, c->P, c->Q, c->fp32, c->species_start, c->keys, c->g.ncells, nullptr, at,
				                      b.count * ncols, 1.0, c->theo, s);
				if (c->explicit_scheme)  // the next step's current scale covers them
					launch_species_u2_added(c->sp, c->P, c->species_start, c->keys, c->g.ncells, nullptr, at, b.count * ncols,
					                        c->u2_slots, s);
				at += b.count * ncols;
				c->next_id += b.count * ncols;
			}
			launch_mark_movers(c->mask, n, n + ninj, s);  // the new particles arrive in the rebin
			if (ninj > 0) {
				set_count(c, n + ninj);
				resized = true;
			}
		}
		if (removed > 0) {
			Timed t(c, "replace", 0);
			ph.id_base = (int)c->next_id;
			launch_replace(c->g, c->sp, ph, c->species_start, c->esc_sorted, removed, c->P, c->Q, c->fp32, c->keys, s);
			// addToPreserveChargeGlobal's new particles (PIC++/initialization.cpp:8319-8321)
			launch_theo_particles(c->sp, c->P, c->Q, c->fp32, c->species_start, c->keys, c->g.ncells, c->esc_sorted, 0, removed,
			                      1.0, c->theo, s);
			if (c->explicit_scheme)
				launch_species_u2_added(c->sp, c->P, c->species_start, c->keys, c->g.ncells, c->esc_sorted, 0, removed, c->u2_slots,
				                        s);
			c->next_id += removed;
			removed = 0;
		}
		c->inj_step++;
	}
	rebin_particles(c, removed, resized);
}

// Restores the cell-sorted order after a push that left one mover bit per particle in
// c->mask (the FP64 and FP32 pushes, the explicit push), without a full sort (pic_kernels_fp32.hip, rebin
// kernels); more than a quarter of the particles changing segment takes the full sort
// instead. Same result as a stable sort by key within each new segment's stayers and
// arrivals. With the wall boundary the `removed` particles (key 0xFFFFFFFF, flagged as movers
// by the push) are dropped, and particles appended after the push (injection, flagged as
// movers) arrive like any other mover.
void rebin_particles(pic_ctx* c, int removed, bool resized) {
	const long long n = c->np();
	hipStream_t s = c->stream;
	int M = 0;
	if (n > 0) {
		Timed t(c, "rebin_scan", (double)n / 64 * 16);
		const long long nw = (n + 63) / 64;
		launch_mask_counts(c->mask, nw, c->wcount, s);
		size_t bytes = c->sort_temp_bytes;
		HIP_CHECK(rocprim::exclusive_scan(c->sort_temp, bytes, c->wcount, c->wprefix, 0, (size_t)(nw + 1), rocprim::plus<int>(), s));
		HIP_CHECK(hipMemcpyAsync(&M, c->wprefix + nw, sizeof(int), hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipStreamSynchronize(s));
	}
	if (n == 0 || M > c->mover_capacity) {
		sort_particles(c, removed, resized);
		return;
	}
	{
		Timed t(c, "rebin_movers", (double)n / 64 * 12 + (double)M * 36 + (double)(c->nkeys + 1) * 4);
		launch_gather_movers(c->mask, c->wprefix, c->keys, n, c->nkeys, c->mk, c->mi, s);
		if (M > 0) {
			size_t bytes = c->sort_temp_bytes;
			HIP_CHECK(rocprim::radix_sort_pairs(c->sort_temp, bytes, c->mk, c->mk_s, c->mi, c->mi_s, (size_t)M, 0, c->mover_bit, s));
		}
		HIP_CHECK(hipMemsetAsync(c->arrive, 0, (c->nkeys + 1) * sizeof(int), s));
		launch_count_keys(c->mk_s, M, c->arrive, s);
		launch_first_of_run(c->mk_s, M, c->first, s);
	}
	{
		Timed t(c, "rebin_segments", (double)(c->nkeys + 1) * 4 * 6);
		launch_rebin_counts(c->start, c->mask, c->wprefix, c->arrive, c->nkeys, c->counts, s);
		size_t bytes = c->sort_temp_bytes;
		HIP_CHECK(rocprim::exclusive_scan(c->sort_temp, bytes, c->counts, c->start_new, 0, (size_t)(c->nkeys + 1),
		                                  rocprim::plus<int>(), s));
	}
	{
		Timed t(c, "rebin_perm", (double)n * 4 + (double)c->nkeys * 12);
		launch_rebin_place(c->start, c->start_new, c->mask, c->arrive, c->first, c->mi_s, c->nkeys, c->perm_sorted, s);
	}
	const long long n_new = n - removed;
	set_count(c, n_new);  // the permutation fills the first n_new places
	if (c->fp32) {
		Timed t(c, "permute", (double)n_new * (2 * 32 + 4));
		launch_permute32(c->Q, c->Qalt, c->perm_sorted, s);
		std::swap(c->Q, c->Qalt);
	} else {
		Timed t(c, "permute", (double)n_new * (2 * 52 + 4));
		launch_permute(c->P, c->Palt, c->perm_sorted, s);
		std::swap(c->P, c->Palt);
	}
	std::swap(c->start, c->start_new);
	if (removed > 0 || resized) species_ranges(c, n_new);
}

// after_push (FP32): the density pass also forms the next step's cell moments, from the
// FP32 copies of Enew and Bnew (c->fused = 1).
void density(pic_ctx* c, bool after_push = false) {
	hipStream_t s = c->stream;
	int nq = c->sp.count * 5 + 1;
	double np = (double)c->np();
	if (c->fp32 && after_push) {
		const Grid& g = c->g;
		{
			Timed t(c, "fields_to_fp32", (double)(g.nnodes + g.ncells) * 3 * 12);
			launch_f64_to_f32(c->Enew, c->E32, g.nnodes, s);
			launch_f64_to_f32(c->Bnew, c->B32, g.ncells, s);
		}
		{
			Timed t(c, "density_cell_moments", np * 28 + (double)g.ncells * (10 + nq) * 8 * 4);
			launch_cell_moments32(g, c->sp, c->Q, c->start, c->E32, c->B32, (float)c->cfg.dt, c->rscr32, c->pscr32, s, c->dscr32);
		}
		c->fused = 1;
		Timed t(c, "density_assemble", (double)g.ncells * nq * (8 * 4 + 8));
		launch_cell_assemble32(g, c->dscr32, nq, c->dens, s);
	} else if (c->fp32) {
		{
			Timed t(c, "density", np * 28 + (double)c->g.ncells * nq * 8);
			launch_density32(c->g, c->sp, c->Q, c->start, c->dscr32, s);
		}
		Timed t(c, "density_assemble", (double)c->g.ncells * nq * (8 * 4 + 8));
		launch_cell_assemble32(c->g, c->dscr32, nq, c->dens, s);
	} else {
		{
			Timed t(c, "density", np * 48 + (double)c->g.ncells * nq * 8);
			launch_density(c->g, c->sp, c->P, c->start, c->dens_scratch, s);
		}
		Timed t(c, "density_assemble", (double)c->g.ncells * nq * 8 * 9);
		launch_density_assemble(c->g, c->dens_scratch, nq, c->dens, s);
	}
	{
		Timed t(c, "folds", 0);
		if (c->g.wall) {
			launch_wall_mirror_cells(c->g, c->dens, nq, c->dens_flip, s);
			launch_wall_right_fold(c->g, c->dens, nq, c->dens_zero, s);
		}
		launch_fold_cells(c->g, c->dens, nq, 4, 3, s);  // sumCellParameters{X,Y,Z}, sumCellVectorParameters{X,Y,Z}
		launch_bulk_divide(c->g, c->dens, c->sp.count, s);
	}
}

void moments(pic_ctx* c) {
	hipStream_t s = c->stream;
	const Grid& g = c->g;
	double np = (double)c->np();
	const bool cached = c->fused == 2;  // cell moments and FP32 fields from the last push's density pass
	c->fused = 0;
	if (c->fp32) {
		if (!cached) {
			Timed t(c, "fields_to_fp32", (double)(g.nnodes + g.ncells) * 3 * 12);
			launch_f64_to_f32(c->E, c->E32, g.nnodes, s);
			launch_f64_to_f32(c->B, c->B32, g.ncells, s);
		}
		{
			Timed t(c, "node_moments", np * 28 + (double)g.nnodes * 12 * 4);
			launch_node_moments32(g, c->sp, c->Q, c->start, c->E32, c->B32, (float)c->cfg.dt, (float)c->theta, c->nscr32, s);
		}
		{
			Timed t(c, "node_assemble", (double)g.ncells * 96 * 4 + (double)g.nnodes * 12 * 8);
			launch_node_assemble32(g, c->nscr32, c->J, c->mu, s);
		}
		if (!cached) {
			Timed t(c, "cell_moments", np * 28 + (double)g.ncells * 10 * 4);
			launch_cell_moments32(g, c->sp, c->Q, c->start, c->E32, c->B32, (float)c->cfg.dt, c->rscr32, c->pscr32, s);
		}
		{
			Timed t(c, "cell_assemble", (double)g.ncells * (8 * 4 + 8));
			launch_cell_assemble32(g, c->rscr32, 1, c->rhohat, s);
		}
		if (c->cfg.pressure == PIC_PRESSURE_CELL) {
			Timed t(c, "pressure_cell", (double)g.ncells * (48 * 4 + 72));
			launch_pressure_cell32(g, c->pscr32, c->Pi, s);
		} else {
			Timed t(c, "pressure_reduce", (double)g.ncells * 48 * 4);
			HIP_CHECK(hipMemsetAsync(c->Pi, 0, 9 * g.ncells * sizeof(double), s));
			launch_pressure_reduce32(g, c->pscr32, c->Pi, s);
		}
	} else {
		{
			Timed t(c, "node_moments", np * 48 + (double)g.nnodes * 12 * 8);
			launch_node_moments(g, c->sp, c->P, c->start, c->E, c->B, c->cfg.dt, c->theta, c->node_scratch, s);
		}
		{
			Timed t(c, "node_assemble", (double)g.ncells * 96 * 8 + (double)g.nnodes * 12 * 8);
			launch_node_assemble(g, c->node_scratch, c->J, c->mu, s);
		}
		{
			Timed t(c, "cell_moments", np * 48 + (double)g.ncells * 10 * 8);
			launch_cell_moments(g, c->sp, c->P, c->start, c->E, c->B, c->cfg.dt, c->rho_scratch, c->pi_scratch, s);
		}
		{
			Timed t(c, "cell_assemble", (double)g.ncells * 9 * 8);
			launch_cell_assemble(g, c->rho_scratch, c->rhohat, s);
		}
		if (c->cfg.pressure == PIC_PRESSURE_CELL) {
			Timed t(c, "pressure_cell", (double)g.ncells * (48 * 8 + 72));
			launch_pressure_cell(g, c->pi_scratch, c->Pi, s);
		} else {
			Timed t(c, "pressure_reduce", (double)g.ncells * 48 * 8);
			HIP_CHECK(hipMemsetAsync(c->Pi, 0, 9 * g.ncells * sizeof(double), s));
			launch_pressure_reduce(g, c->pi_scratch, c->Pi, s);
		}
	}
	Timed tf(c, "folds", 0);
	if (g.wall) {  // cell-map deposits behind the wall go to their mirror cells
		launch_wall_mirror_cells(g, c->rhohat, 1, nullptr, s);
		launch_wall_mirror_cells(g, c->Pi, 9, nullptr, s);
	}
	launch_fold_cells(g, c->rhohat, 1, 4, 3, s);  // sumChargeDensityHat{X,Y,Z}
	if (c->cfg.pressure == PIC_PRESSURE_CELL)
		launch_fold_cells(g, c->Pi, 9, 4, 3, s);  // the periodic images of chargeDensityHat's folds
	else
		launch_fold_cells(g, c->Pi, 9, 3, 2, s);  // sumCellMatrixParameters{X,Y,Z}
	launch_fold_nodes(g, c->J, 3, s);             // sumNodeVectorParameters{X,Y,Z}
	launch_div_pressure_update_j(g, c->Pi, c->J, c->divPi, c->eta, c->cfg.dt, s);
	// the wall node and the ghosts behind it carry no current (momentum_collecting.cpp:497-506, :638-646)
	if (g.wall) launch_zero_node_planes(g, c->J, 0, 3, 0, 2, s);
	launch_div_flux_update_rho(g, c->J, c->rhohat, c->cfg.dt, c->theta, s);
	launch_fold_nodes(g, c->mu, 9, s);            // sumNodeMatrixParameters{X,Y,Z}
	if (g.wall) launch_zero_node_planes(g, c->mu, 0, 9, 2, 2, s);
}

// The smoothing loop of the step (PIC++/simulation.cpp:213-217): smoothChargeDensityHat with its
// scalar cell exchange, smoothFlux with exchangeGeneralEfield, smoothMatrixNodeParameter on the
// dielectric tensor, whose ghosts the reference does not refresh between passes.
void smooth(pic_ctx* c) {
	if (c->cfg.smoothing_count <= 0) return;
	const Grid& g = c->g;
	hipStream_t s = c->stream;
	Timed t(c, "smoothing", (double)c->cfg.smoothing_count * ((double)g.ncells * 2 + (double)g.nnodes * 12 * 2) * 8 * 2);
	const Smooth27 m = smoothing_matrix(c->cfg.smoothing_parameter);
	for (int n = 0; n < c->cfg.smoothing_count; ++n) {
		launch_smooth_cells(g, c->rhohat, c->smooth_tmp, m, s);
		launch_ghost_fill(g, c->rhohat, 1, false, s, /*wall_zero=*/true);
		launch_smooth_nodes(g, c->J, 3, c->smooth_tmp, m, s);
		launch_ghost_fill(g, c->J, 3, true, s);
		launch_smooth_nodes(g, c->mu, 9, c->smooth_tmp, m, s);
	}
}

double operator_C(const pic_ctx* c) { return (c->theta * c->cfg.dt) * (c->theta * c->cfg.dt) * kCprime2; }

// Givens-updated least squares of the reference GMRES (PIC++/specialmath.cpp:487-600).
struct Givens {
	std::vector<std::vector<double>> H, Q, R, oldQ, oldR;
	std::vector<double> y;
	double rel = 1.0;
	void update(int nn, const std::vector<double>& hcol) {
		// hcol: column nn-2 of the Hessenberg matrix, entries 0..nn-1
		std::vector<std::vector<double>> Hn(nn, std::vector<double>(nn - 1, 0.0));
		for (int r = 0; r < nn - 1; ++r)
			for (int cc = 0; cc < nn - 2; ++cc) Hn[r][cc] = H[r][cc];
		for (int r = 0; r < nn; ++r) Hn[r][nn - 2] = hcol[r];
		H = Hn;
		if (nn == 2) {
			double rho = H[0][0], sigma = H[1][0];
			double module = std::sqrt(rho * rho + sigma * sigma);
			double cs = rho / module, sn = sigma / module;
			Q = {{cs, sn}, {-sn, cs}};
			R = {{module}, {0}};
		} else {
			R.assign(nn, std::vector<double>(nn - 1, 0.0));
			for (int r = 0; r < nn - 1; ++r)
				for (int cc = 0; cc < nn - 2; ++cc) R[r][cc] = oldR[r][cc];
			Q.assign(nn, std::vector<double>(nn, 0.0));
			for (int r = 0; r < nn - 1; ++r)
				for (int cc = 0; cc < nn - 1; ++cc) Q[r][cc] = oldQ[r][cc];
			Q[nn - 1][nn - 1] = 1;
			for (int r = 0; r < nn; ++r) {
				R[r][nn - 2] = 0;
				for (int cc = 0; cc < nn; ++cc) R[r][nn - 2] += Q[r][cc] * H[cc][nn - 2];
			}
			double rho = R[nn - 2][nn - 2], sigma = R[nn - 1][nn - 2];
			double module = std::sqrt(rho * rho + sigma * sigma);
			double cs = rho / module, sn = sigma / module;
			R[nn - 2][nn - 2] = module;
			R[nn - 1][nn - 2] = 0;
			for (int cc = 0; cc < nn - 1; ++cc) {
				Q[nn - 2][cc] = cs * oldQ[nn - 2][cc];
				Q[nn - 1][cc] = -sn * oldQ[nn - 2][cc];
			}
			Q[nn - 2][nn - 1] = sn;
			Q[nn - 1][nn - 1] = cs;
		}
		double beta = 1.0;
		y.assign(nn - 1, 0.0);
		for (int i = nn - 2; i >= 0; --i) {
			y[i] = beta * Q[i][0];
			for (int j = nn - 2; j > i; --j) y[i] -= R[i][j] * y[j];
			if (R[i][i] > 0)
				y[i] /= R[i][i];
			else
				y[i] = 0;
		}
		oldQ = Q;
		oldR = R;
	}
};

// The operator and Krylov storage of one GMRES: op(v, w) sets w = A v on vectors of length len;
// V holds up to max_basis + 1 basis vectors (Vdev: their device addresses).
struct KrylovSystem {
	long long len;
	std::function<void(const double*, double*)> op;
	const char* op_name;
	double op_bytes;  // algorithmic bytes of one application
	std::vector<double*>* V;
	double** Vdev;
	int max_basis;
	int ortho;
};

// One unrestarted GMRES cycle on A x = b (x0 = 0), as generalizedMinimalResidualMethod:
// b is normalised in place and restored. Returns Arnoldi steps; rel = Givens estimate.
int gmres_cycle_on(pic_ctx* c, const KrylovSystem& K, double* b, double* x, double target, int max_iterations, int dim,
                   double& rel) {
	hipStream_t s = c->stream;
	const long long len = K.len;
	std::vector<double*>& V = *K.V;
	double* d_norm = c->dscal + 0;
	double* d_normr = c->dscal + 1;
	double* d_h = c->dscal + 8;
	launch_dot(b, b, len, c->partial, d_norm, true, s);
	double norm = 0;
	HIP_CHECK(hipMemcpyAsync(&norm, d_norm, sizeof(double), hipMemcpyDeviceToHost, s));
	HIP_CHECK(hipStreamSynchronize(s));
	rel = 0;
	if (norm == 0) {
		HIP_CHECK(hipMemsetAsync(x, 0, len * sizeof(double), s));
		return 0;
	}
	launch_div_copy(b, V[0], d_norm, len, s);
	launch_dot(b, b, len, c->partial, d_normr, true, s);
	double normr = 0;
	HIP_CHECK(hipMemcpyAsync(&normr, d_normr, sizeof(double), hipMemcpyDeviceToHost, s));
	HIP_CHECK(hipStreamSynchronize(s));
	Givens gv;
	int nn = 2;
	rel = 1.0;
	int limit = std::min(std::min(max_iterations, dim + 3), K.max_basis + 1);
	std::vector<double> hcol;
	while (rel > target && nn < limit) {
		int j = nn - 2;
		double* w = V[j + 1];
		{
			Timed t(c, K.op_name, K.op_bytes);
			K.op(V[j], w);
		}
		{
			// algorithmic bytes: MGS reads v_m and w twice and writes w once per basis vector
			Timed tg(c, "orthogonalize", (double)len * 8 * (K.ortho == PIC_ORTHO_MGS ? 5.0 * (j + 1) + 3 : 6.0 * (j + 1) + 3));
			if (K.ortho == PIC_ORTHO_MGS) {
				for (int m = 0; m <= j; ++m) {
					launch_dot(V[m], w, len, c->partial, d_h + m, false, s);
					launch_axpy_dev(w, V[m], d_h + m, len, s);
				}
			} else {
				// CGS2: two passes of h = V^T w, w -= V h (same span; reorthogonalised)
				for (int pass = 0; pass < 2; ++pass) {
					double* hp = d_h + (pass == 0 ? 0 : 64);
					for (int m = 0; m <= j; ++m)
						launch_dot(V[m], w, len, c->partial + (m + 1) * kDotBlocks, hp + m, false, s);
					for (int m = 0; m <= j; ++m) launch_axpy_dev(w, V[m], hp + m, len, s);
				}
			}
			launch_dot(w, w, len, c->partial, d_h + j + 1, true, s);
			launch_normalize_dev(w, d_h + j + 1, len, s);
		}
		hcol.assign(nn, 0.0);
		if (K.ortho == PIC_ORTHO_MGS) {
			HIP_CHECK(hipMemcpyAsync(hcol.data(), d_h, nn * sizeof(double), hipMemcpyDeviceToHost, s));
			HIP_CHECK(hipStreamSynchronize(s));
		} else {
			std::vector<double> h1(nn), h2(nn);
			HIP_CHECK(hipMemcpyAsync(h1.data(), d_h, nn * sizeof(double), hipMemcpyDeviceToHost, s));
			HIP_CHECK(hipMemcpyAsync(h2.data(), d_h + 64, (j + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
			HIP_CHECK(hipStreamSynchronize(s));
			for (int m = 0; m <= j; ++m) hcol[m] = h1[m] + h2[m];
			hcol[j + 1] = h1[j + 1];
		}
		gv.update(nn, hcol);
		rel = std::fabs(1.0 * gv.Q[nn - 1][0]) / normr;
		nn++;
	}
	int steps = nn - 2;
	HIP_CHECK(hipMemcpyAsync(c->ydev, gv.y.data(), gv.y.size() * sizeof(double), hipMemcpyHostToDevice, s));
	Timed t(c, "combine", (double)len * 8 * (gv.y.size() + 1));
	launch_combine(x, K.Vdev, c->ydev, (int)gv.y.size(), d_norm, len, s);
	return steps;
}

// GMRES on the Maxwell operator (the 37-entry rows) with the context's basis
int gmres_cycle(pic_ctx* c, double* b, double* x, double target, int max_iterations, int dim, double& rel) {
	const double C = operator_C(c);
	KrylovSystem K{c->nown3(), [c, C](const double* v, double* w) { launch_apply_operator(c->g, c->muh, v, C, w, c->stream); },
	               "apply_operator", (double)c->g.nown * 15 * 8, &c->V, c->Vdev, c->max_basis, c->cfg.ortho};
	return gmres_cycle_on(c, K, b, x, target, max_iterations, dim, rel);
}

// updateEnergy's sums (PIC++/simulation.cpp:1506-1640) without the unit factors: h[0] E energy,
// h[1] B energy, h[2..4] field momentum, h[5] particle energy, h[6..8] particle momentum. The
// explicit scheme takes the Yee components at one (i, j, k) as the reference does
// (getBunemanElectricField / getBunemanMagneticField, PIC++/fields.cpp:3429-3435).
void energy_sums(pic_ctx* c, double h[9]) {
	hipStream_t s = c->stream;
	const Grid& g = c->g;
	if (!c->explicit_scheme) launch_field_energy(g, c->E, c->B, c->epartial, s);
	if (c->fp32)
		launch_particle_energy32(c->sp, c->Q, c->species_start, c->epartial + 5 * kEnergyBlocks, s);
	else
		launch_particle_energy(c->sp, c->P, c->species_start, c->epartial + 5 * kEnergyBlocks, s);
	launch_rows_final(c->epartial, kEnergyBlocks, 9, c->eout, s);
	HIP_CHECK(hipMemcpyAsync(h, c->eout, 9 * sizeof(double), hipMemcpyDeviceToHost, s));
	double gr[kGeneralRows] = {};
	if (c->explicit_scheme) {
		launch_general_fields(g, yee_view(c->yee), c->gpartial, c->gout, s);
		HIP_CHECK(hipMemcpyAsync(gr, c->gout, sizeof(gr), hipMemcpyDeviceToHost, s));
	}
	HIP_CHECK(hipStreamSynchronize(s));
	if (c->explicit_scheme) {
		const double pi = 4 * std::atan2(1.0, 1.0);
		h[0] = (gr[0] + gr[1] + gr[2]) * g.volume / (8 * pi);
		h[1] = (gr[3] + gr[4] + gr[5]) * g.volume / (8 * pi);
		for (int a = 0; a < 3; ++a) h[2 + a] = gr[6 + a] * g.volume / (4 * pi);
	}
}

// The theoretical energy and momentum at the start of the run: the energy and the total
// momentum (updateEnergy at iteration 0, PIC++/simulation.cpp:1630-1633).
void init_theo(pic_ctx* c) {
	double h[9];
	energy_sums(c, h);
	const double t[4] = {h[5] + h[0] + h[1], h[2] + h[6], h[3] + h[7], h[4] + h[8]};
	HIP_CHECK(hipMemcpyAsync(c->theo, t, sizeof(t), hipMemcpyHostToDevice, c->stream));
	HIP_CHECK(hipStreamSynchronize(c->stream));
	c->theo_init = true;
}

}  // namespace

extern "C" {

const char* pic_last_error(const pic_ctx* c) { return c ? c->error.c_str() : noalloc_last_error(); }

int pic_create(const pic_config* cfg, void* stream, pic_ctx** out) {
	if (!cfg || !out) return PIC_ERR_ARGUMENT;
	*out = nullptr;
	// nz = 1: the reference's 2-D XY grid (no ghost layers along z), built for the explicit scheme
	const bool cz = cfg->nz == 1;
	if (cfg->nx < 5 || cfg->ny < 5 || (cfg->nz < 5 && !cz) || cfg->nspecies < 1 || cfg->nspecies > kMaxSpecies || cfg->dt <= 0 ||
	    cfg->dx <= 0 || cfg->dy <= 0 || cfg->dz <= 0 || cfg->capacity < 0)
		return PIC_ERR_ARGUMENT;
	if (cz && cfg->scheme != PIC_SCHEME_EXPLICIT) return PIC_ERR_ARGUMENT;
	if (cfg->precision != PIC_PRECISION_FP64 && cfg->precision != PIC_PRECISION_FP32) return PIC_ERR_ARGUMENT;
	if (cfg->pressure != PIC_PRESSURE_REFERENCE && cfg->pressure != PIC_PRESSURE_CELL) return PIC_ERR_ARGUMENT;
	if (cfg->smoothing_count < 0 || (cfg->smoothing_count > 0 && !(cfg->smoothing_parameter >= 0.0 && cfg->smoothing_parameter <= 1.0)))
		return PIC_ERR_ARGUMENT;
	if (cfg->boundary_x != PIC_BOUNDARY_PERIODIC && cfg->boundary_x != PIC_BOUNDARY_WALL) return PIC_ERR_ARGUMENT;
	if (cfg->scheme != PIC_SCHEME_IMPLICIT && cfg->scheme != PIC_SCHEME_EXPLICIT) return PIC_ERR_ARGUMENT;
	if ((cfg->divergence_cleanup || cfg->filter_cut_wave_number != 0) &&  // built for the periodic implicit step
	    (cfg->scheme != PIC_SCHEME_IMPLICIT || cfg->boundary_x != PIC_BOUNDARY_PERIODIC || cfg->filter_cut_wave_number < 0 ||
	     cfg->divergence_cleanup < 0 || cfg->divergence_cleanup > 3))
		return PIC_ERR_ARGUMENT;
	if (cfg->scheme == PIC_SCHEME_EXPLICIT &&  // the explicit step is built FP64; its wall needs the particles' wall
	                                          // (cell 20) inside the box
	    (cfg->precision != PIC_PRECISION_FP64 || (cfg->boundary_x == PIC_BOUNDARY_WALL && cfg->nx < 20)))
		return PIC_ERR_ARGUMENT;
	if (cfg->inject) {  // injection needs the wall set-up, a flow below c and finite temperatures
		if (cfg->inject != 1 || cfg->boundary_x != PIC_BOUNDARY_WALL) return PIC_ERR_ARGUMENT;
		if (!(cfg->run_kT >= 0.0 && std::isfinite(cfg->run_kT))) return PIC_ERR_ARGUMENT;
		double v2 = cfg->flow[0] * cfg->flow[0] + cfg->flow[1] * cfg->flow[1] + cfg->flow[2] * cfg->flow[2];
		if (!(v2 < 1.0)) return PIC_ERR_ARGUMENT;
		for (int s = 0; s < cfg->nspecies; ++s)
			if (cfg->ppc[s] < 0 || (cfg->ppc[s] > 0 && !(cfg->temperature[s] >= 0.0 && std::isfinite(cfg->temperature[s]))))
				return PIC_ERR_ARGUMENT;
	}
	pic_ctx* c = new pic_ctx();
	try {
		c->cfg = *cfg;
		if (c->cfg.gmres_max_iterations <= 0) c->cfg.gmres_max_iterations = 50;
		if (c->cfg.gmres_tolerance <= 0) c->cfg.gmres_tolerance = 1e-8;
		Grid& g = c->g;
		int n[3] = {cfg->nx, cfg->ny, cfg->nz};
		double d[3] = {cfg->dx, cfg->dy, cfg->dz};
		g.cz = cz ? 1 : 0;
		for (int a = 0; a < 3; ++a) {
			g.n[a] = n[a];
			g.N[a] = n[a] + 2 * owned_lo(g, a);  // two ghost layers each side, none along a collapsed z
			g.d[a] = d[a];
			g.L[a] = n[a] * d[a];
			g.x0[a] = g.L[a] - owned_lo(g, a) * d[a];  // node 0; the first owned node at L
		}
		g.volume = d[0] * d[1] * d[2];
		g.nnodes = (long long)(g.N[0] + 1) * (g.N[1] + 1) * (g.N[2] + 1);
		g.ncells = (long long)g.N[0] * g.N[1] * g.N[2];
		g.wall = cfg->boundary_x == PIC_BOUNDARY_WALL ? 1 : 0;
		g.nu0 = n[0] + 2 * g.wall;
		g.nown = (long long)g.nu0 * n[1] * n[2];
		for (int a = 0; a < 3; ++a) {
			g.e_right[a] = cfg->e_right[a];
			g.b_right[a] = cfg->b_right[a];
		}
		c->sp.count = cfg->nspecies;
		for (int s = 0; s < kMaxSpecies; ++s) {
			c->sp.mass[s] = s < cfg->nspecies ? cfg->mass[s] : 1.0;
			c->sp.charge[s] = s < cfg->nspecies ? cfg->charge[s] : 0.0;
			c->sp.weight[s] = s < cfg->nspecies ? cfg->weight[s] : 0.0;
		}
		if (stream) {
			c->stream = static_cast<hipStream_t>(stream);
		} else {
			HIP_CHECK(hipStreamCreate(&c->stream));
			c->own_stream = true;
		}
		if (cfg->scheme == PIC_SCHEME_EXPLICIT) {
			c->explicit_scheme = true;
			if ((long long)(g.N[0] + 1) * (g.N[1] + 1) * (g.N[2] + 1) > 0x7FFFFFFFLL)
				throw DeviceError{PIC_ERR_ARGUMENT, "explicit scheme: the Yee arrays need fewer than 2^31 points"};
			for (int comp = 0; comp < 3; ++comp) {
				yee_dims(g, false, comp, c->yee.dE[comp]);
				yee_dims(g, true, comp, c->yee.dB[comp]);
				c->yee.E[comp] = dalloc<double>(yee_count(c->yee.dE[comp]));
				c->yee.J[comp] = dalloc<double>(yee_count(c->yee.dE[comp]));
				c->yee.Jacc[comp] = dalloc<long long>(yee_count(c->yee.dE[comp]));
				c->yee.B[comp] = dalloc<double>(yee_count(c->yee.dB[comp]));
				HIP_CHECK(hipMemset(c->yee.E[comp], 0, yee_count(c->yee.dE[comp]) * sizeof(double)));
				HIP_CHECK(hipMemset(c->yee.J[comp], 0, yee_count(c->yee.dE[comp]) * sizeof(double)));
				HIP_CHECK(hipMemset(c->yee.B[comp], 0, yee_count(c->yee.dB[comp]) * sizeof(double)));
			}
			c->bounds = dalloc<ExplicitBounds>(1);
			c->u2_slots = dalloc<unsigned long long>((size_t)kU2Slots * kMaxSpecies);
		}
		c->nkeys = (long long)cfg->nspecies * g.ncells * 8;
		if (c->nkeys >= (1LL << 32) - 1) throw DeviceError{PIC_ERR_ARGUMENT, "grid too large for 32-bit particle keys"};
		c->end_bit = 1;
		while ((1LL << c->end_bit) < c->nkeys) ++c->end_bit;
		c->mover_bit = c->end_bit;
		while ((1LL << c->mover_bit) <= c->nkeys) ++c->mover_bit;
		c->E = dalloc<double>(3 * g.nnodes);
		c->Etheta = dalloc<double>(3 * g.nnodes);
		c->Enew = dalloc<double>(3 * g.nnodes);
		c->B = dalloc<double>(3 * g.ncells);
		c->Bnew = dalloc<double>(3 * g.ncells);
		c->J = dalloc<double>(3 * g.nnodes);
		c->mu = dalloc<double>(9 * g.nnodes);
		c->divPi = dalloc<double>(3 * g.nnodes);
		c->rhohat = dalloc<double>(g.ncells);
		c->Pi = dalloc<double>(9 * g.ncells);
		c->dens = dalloc<double>((5 * cfg->nspecies + 1) * g.ncells);
		c->fp32 = cfg->precision == PIC_PRECISION_FP32;
		long long pcap = c->fp32 ? 0 : cfg->capacity;  // FP64 particle arrays
		if (c->fp32) {
			c->nscr32 = dalloc<float>(96 * g.ncells);
			c->rscr32 = dalloc<float>(8 * g.ncells);
			c->pscr32 = dalloc<float>(48 * g.ncells);
			c->dscr32 = dalloc<float>((5 * cfg->nspecies + 1) * 8 * g.ncells);
			c->E32 = dalloc<float>(4 * g.nnodes);  // node-major float4
			c->Et32 = dalloc<float>(4 * g.nnodes);
			c->B32 = dalloc<float>(4 * g.ncells);
			for (ParticleArrays32* Q : {&c->Q, &c->Qalt}) {
				Q->cell = dalloc<int>(cfg->capacity);
				for (int a = 0; a < 3; ++a) {
					Q->off[a] = dalloc<float>(cfg->capacity);
					Q->u[a] = dalloc<float>(cfg->capacity);
				}
				Q->id = dalloc<int>(cfg->capacity);
				Q->count = 0;
			}
		} else {
			c->node_scratch = dalloc<double>(96 * g.ncells);
			c->rho_scratch = dalloc<double>(8 * g.ncells);
			c->pi_scratch = dalloc<double>(48 * g.ncells);
			c->dens_scratch = dalloc<double>((5 * cfg->nspecies + 1) * 8 * g.ncells);
		}
		for (double* p : {c->E, c->Etheta, c->Enew, c->J, c->divPi}) HIP_CHECK(hipMemset(p, 0, 3 * g.nnodes * sizeof(double)));
		HIP_CHECK(hipMemset(c->mu, 0, 9 * g.nnodes * sizeof(double)));
		for (double* p : {c->B, c->Bnew}) HIP_CHECK(hipMemset(p, 0, 3 * g.ncells * sizeof(double)));
		HIP_CHECK(hipMemset(c->rhohat, 0, g.ncells * sizeof(double)));
		HIP_CHECK(hipMemset(c->Pi, 0, 9 * g.ncells * sizeof(double)));
		HIP_CHECK(hipMemset(c->dens, 0, (5 * cfg->nspecies + 1) * g.ncells * sizeof(double)));
		c->capacity = cfg->capacity;
		for (ParticleArrays* P : {&c->P, &c->Palt}) {
			for (int a = 0; a < 3; ++a) {
				P->x[a] = dalloc<double>(pcap);
				P->p[a] = dalloc<double>(pcap);
			}
			P->id = dalloc<int>(pcap);
			P->count = 0;
		}
		c->keys = dalloc<uint32_t>(c->capacity);
		c->keys_sorted = dalloc<uint32_t>(c->capacity);
		c->perm = dalloc<uint32_t>(c->capacity);
		c->perm_sorted = dalloc<uint32_t>(c->capacity);
		c->start = dalloc<int>(c->nkeys + 1);
		c->counts = dalloc<int>(c->nkeys + 1);
		HIP_CHECK(hipMemset(c->start, 0, (c->nkeys + 1) * sizeof(int)));
		size_t b1 = 0, b2 = 0, b3 = 0, b4 = 0;
		HIP_CHECK(rocprim::radix_sort_pairs(nullptr, b1, c->keys, c->keys_sorted, c->perm, c->perm_sorted,
		                                    (size_t)std::max<long long>(c->capacity, 1), 0, c->end_bit, c->stream));
		HIP_CHECK(rocprim::exclusive_scan(nullptr, b2, c->counts, c->start, 0, (size_t)(c->nkeys + 1),
		                                  rocprim::plus<int>(), c->stream));
		{  // the rebin after a push with mover bits
			const long long words = c->capacity / 64 + 2;
			c->mask = dalloc<uint64_t>(words);
			if (g.wall) c->bounce = dalloc<uint64_t>(words);
			c->wcount = dalloc<int>(words);
			c->wprefix = dalloc<int>(words);
			c->start_new = dalloc<int>(c->nkeys + 1);
			c->arrive = dalloc<int>(c->nkeys + 1);
			c->first = dalloc<int>(c->nkeys + 1);
			c->mover_capacity = c->capacity / 4 + 1;
			for (uint32_t** p : {&c->mk, &c->mi, &c->mk_s, &c->mi_s}) *p = dalloc<uint32_t>(c->mover_capacity);
			HIP_CHECK(rocprim::exclusive_scan(nullptr, b3, c->wcount, c->wprefix, 0, (size_t)words, rocprim::plus<int>(), c->stream));
			HIP_CHECK(rocprim::radix_sort_pairs(nullptr, b4, c->mk, c->mk_s, c->mi, c->mi_s,
			                                    (size_t)c->mover_capacity, 0, c->mover_bit, c->stream));
		}
		c->sort_temp_bytes = std::max(std::max(b1, b2), std::max(b3, b4));
		c->sort_temp = dalloc<char>(c->sort_temp_bytes);
		c->species_start = dalloc<int>(kMaxSpecies + 1);
		c->species_start_h.assign(kMaxSpecies + 1, 0);
		HIP_CHECK(hipMemset(c->species_start, 0, (kMaxSpecies + 1) * sizeof(int)));
		c->status = dalloc<int>(kStatusWords);
		if (cfg->smoothing_count > 0 && cfg->scheme != PIC_SCHEME_EXPLICIT) c->smooth_tmp = dalloc<double>(9 * g.nnodes);
		c->dead = dalloc<int>(1);
		HIP_CHECK(hipMemset(c->dead, 0, sizeof(int)));
		if (g.wall) {
			c->esc = dalloc<uint32_t>(c->capacity);
			c->esc_sorted = dalloc<uint32_t>(c->capacity);
			while ((1LL << c->esc_bits) < c->capacity) ++c->esc_bits;
			size_t b4 = 0;
			HIP_CHECK(rocprim::radix_sort_keys(nullptr, b4, c->esc, c->esc_sorted,
			                                   (size_t)std::max<long long>(c->capacity, 1), 0, c->esc_bits, c->stream));
			if (b4 > c->sort_temp_bytes) {
				hipFree(c->sort_temp);
				c->sort_temp_bytes = b4;
				c->sort_temp = dalloc<char>(c->sort_temp_bytes);
			}
			for (int s = 0; s < cfg->nspecies; ++s)  // injectionLength at the loop start (PIC++/simulation.cpp:125-127)
				if (cfg->inject && cfg->ppc[s] > 0) c->inj_len[s] = cfg->dx / cfg->ppc[s] - 0.0001 * cfg->dx;
			bool hot = false;  // the tabulated Juttner CDFs of the species injected from kT / mc^2 = 0.1 on
			for (int s = 0; s < cfg->nspecies; ++s) hot = hot || (cfg->inject && cfg->ppc[s] > 0 && cfg->temperature[s] >= 0.1);
			if (hot) {
				std::vector<double> tab((size_t)kMaxSpecies * 2 * kJuttnerPoints, 0.0);
				for (int s = 0; s < cfg->nspecies; ++s)
					if (cfg->ppc[s] > 0 && cfg->temperature[s] >= 0.1) juttner_table(cfg->temperature[s], &tab[(size_t)s * 2 * kJuttnerPoints]);
				c->juttner = dalloc<double>(tab.size());
				HIP_CHECK(hipMemcpy(c->juttner, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
			}
		}
		{
			int nq = 5 * cfg->nspecies + 1;
			std::vector<int> flip(nq, 0), zero(nq, 1);
			for (int q = 0; q < 5 * cfg->nspecies; ++q) {
				flip[q] = q % 5 == 2;  // x momentum sum of each species
				zero[q] = q % 5 <= 1;  // concentration, energy (charge is the last plane)
			}
			c->dens_flip = dalloc<int>(nq);
			c->dens_zero = dalloc<int>(nq);
			HIP_CHECK(hipMemcpy(c->dens_flip, flip.data(), nq * sizeof(int), hipMemcpyHostToDevice));
			HIP_CHECK(hipMemcpy(c->dens_zero, zero.data(), nq * sizeof(int), hipMemcpyHostToDevice));
		}
		HIP_CHECK(hipMemset(c->status, 0, kStatusWords * sizeof(int)));
		// Krylov storage: up to max_iterations basis vectors
		c->max_basis = std::min(120, std::max(2, c->cfg.gmres_max_iterations));
		long long len = 3 * g.nown;
		c->rhs = dalloc<double>(len);
		c->muc = dalloc<double>(9 * g.nown);
		c->muh = dalloc<double>(9 * halo_count(g));
		c->xsol = dalloc<double>(len);
		c->work = dalloc<double>(len);
		c->work2 = dalloc<double>(len);
		HIP_CHECK(hipMemset(c->xsol, 0, len * sizeof(double)));
		c->V.resize(c->max_basis + 1);
		for (auto& v : c->V) v = dalloc<double>(len);
		c->Vdev = dalloc<double*>(c->V.size());
		HIP_CHECK(hipMemcpy(c->Vdev, c->V.data(), c->V.size() * sizeof(double*), hipMemcpyHostToDevice));
		c->w64 = dalloc<double>(3 * halo_count(g));
		if (c->cfg.solver_precision == PIC_SOLVER_MIXED) {
			c->mu32 = dalloc<float>(9 * g.nown);
			c->mu32h = dalloc<float>(9 * halo_count(g));
			c->w32 = dalloc<float>(3 * halo_count(g));
			c->V32.resize(c->max_basis + 1);
			for (auto& v : c->V32) v = dalloc<float>(len);
			c->V32dev = dalloc<float*>(c->V32.size());
			HIP_CHECK(hipMemcpy(c->V32dev, c->V32.data(), c->V32.size() * sizeof(float*), hipMemcpyHostToDevice));
			c->tmp32 = dalloc<float>(2 * len);
		}
		c->partial = dalloc<double>((size_t)kDotBlocks * 128);
		c->dscal = dalloc<double>(256);
		c->ydev = dalloc<double>(128);
		c->epartial = dalloc<double>(9 * kEnergyBlocks);
		c->eout = dalloc<double>(16);
		c->theo = dalloc<double>(4);
		HIP_CHECK(hipMemset(c->theo, 0, 4 * sizeof(double)));
		c->gpartial = dalloc<double>((size_t)kGeneralRows * kEnergyBlocks);
		c->gout = dalloc<double>(kGeneralRows);
		HIP_CHECK(hipDeviceSynchronize());
	} catch (const DeviceError& e) {
		std::string msg = e.what;
		int code = e.code;
		pic_destroy(c);
		(void)msg;
		return code;
	}
	*out = c;
	return PIC_OK;
}

void pic_destroy(pic_ctx* c) {
	if (!c) return;
	hipDeviceSynchronize();
	auto f = [](void* p) {
		if (p) hipFree(p);
	};
	for (void* p : {(void*)c->E, (void*)c->Etheta, (void*)c->Enew, (void*)c->B, (void*)c->Bnew, (void*)c->J, (void*)c->mu,
	                (void*)c->divPi, (void*)c->rhohat, (void*)c->Pi, (void*)c->dens, (void*)c->node_scratch,
	                (void*)c->rho_scratch, (void*)c->pi_scratch, (void*)c->dens_scratch, (void*)c->keys, (void*)c->keys_sorted,
	                (void*)c->perm, (void*)c->perm_sorted, (void*)c->start, (void*)c->counts, c->sort_temp,
	                (void*)c->species_start, (void*)c->status, (void*)c->dead, (void*)c->dens_flip, (void*)c->dens_zero, (void*)c->esc, (void*)c->esc_sorted, (void*)c->smooth_tmp, (void*)c->rhs, (void*)c->muc, (void*)c->xsol, (void*)c->work, (void*)c->work2,
	                (void*)c->Vdev, (void*)c->partial, (void*)c->dscal, (void*)c->ydev, (void*)c->epartial, (void*)c->eout,
	                (void*)c->theo, (void*)c->gpartial, (void*)c->gout, (void*)c->bounce, (void*)c->juttner})
		f(p);
	for (ParticleArrays* P : {&c->P, &c->Palt}) {
		for (int a = 0; a < 3; ++a) {
			f(P->x[a]);
			f(P->p[a]);
		}
		f(P->id);
	}
	for (int comp = 0; comp < 3; ++comp) {
		f(c->yee.E[comp]);
		f(c->yee.B[comp]);
		f(c->yee.J[comp]);
		f(c->yee.Jacc[comp]);
	}
	f(c->bounds);
	f(c->u2_slots);
	for (ParticleArrays32* Q : {&c->Q, &c->Qalt}) {
		f(Q->cell);
		for (int a = 0; a < 3; ++a) {
			f(Q->off[a]);
			f(Q->u[a]);
		}
		f(Q->id);
	}
	for (void* p : {(void*)c->E32, (void*)c->Et32, (void*)c->B32, (void*)c->nscr32, (void*)c->rscr32, (void*)c->pscr32,
	                (void*)c->dscr32})
		f(p);
	for (double* v : c->V) f(v);
	for (float* v : c->V32) f(v);
	for (void* p : {(void*)c->V32dev, (void*)c->mu32, (void*)c->mu32h, (void*)c->muh, (void*)c->w32, (void*)c->w64, (void*)c->tmp32, (void*)c->Z32dev,
	                (void*)c->log_r, (void*)c->log_e, (void*)c->mask, (void*)c->wcount, (void*)c->wprefix, (void*)c->start_new,
	                (void*)c->arrive, (void*)c->first, (void*)c->mk, (void*)c->mi, (void*)c->mk_s, (void*)c->mi_s})
		f(p);
	for (float* v : c->Z32) f(v);
	for (float* v : c->log_mu) f(v);
	for (double* v : c->Vcl) f(v);
	for (void* p : {(void*)c->Vcl_dev, (void*)c->cl_rhs, (void*)c->cl_phi, (void*)c->ft_re, (void*)c->ft_im, (void*)c->ft_re2,
	                (void*)c->ft_im2, (void*)c->ft_tw[0], (void*)c->ft_tw[1], (void*)c->ft_tw[2]})
		f(p);
	c->net.reset();
	for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
	if (c->own_stream && c->stream) hipStreamDestroy(c->stream);
	delete c;
}

int pic_upload_fields(pic_ctx* c, const double* E, const double* B) {
	if (!c || !E || !B) return PIC_ERR_ARGUMENT;
	c->fused = 0;
	c->u2_valid = false;
	c->theo_init = false;  // the run starts again from this state
	try {
		const Grid& g = c->g;
		HIP_CHECK(hipMemcpyAsync(c->E, E, 3 * g.nnodes * sizeof(double), hipMemcpyHostToDevice, c->stream));
		HIP_CHECK(hipMemcpyAsync(c->Etheta, E, 3 * g.nnodes * sizeof(double), hipMemcpyHostToDevice, c->stream));
		HIP_CHECK(hipMemcpyAsync(c->Enew, E, 3 * g.nnodes * sizeof(double), hipMemcpyHostToDevice, c->stream));
		HIP_CHECK(hipMemcpyAsync(c->B, B, 3 * g.ncells * sizeof(double), hipMemcpyHostToDevice, c->stream));
		HIP_CHECK(hipMemcpyAsync(c->Bnew, B, 3 * g.ncells * sizeof(double), hipMemcpyHostToDevice, c->stream));
		HIP_CHECK(hipStreamSynchronize(c->stream));
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

}  // extern "C"

namespace {

// Device-resident FP64 particles (grouped by species, cnt = species prefix counts) into
// the context's particle form, then sorted by key.
void load_particles(pic_ctx* c, const std::vector<int>& cnt, const double* x, const double* p, const int* ids) {
	long long np = cnt[c->sp.count];
	if (np > c->capacity) throw DeviceError{PIC_ERR_ARGUMENT, "particle capacity exceeded"};
	hipStream_t s = c->stream;
	c->species_start_h.assign(kMaxSpecies + 1, (int)np);
	for (int q = 0; q <= c->sp.count; ++q) c->species_start_h[q] = cnt[q];
	HIP_CHECK(hipMemcpyAsync(c->species_start, c->species_start_h.data(), (kMaxSpecies + 1) * sizeof(int),
	                         hipMemcpyHostToDevice, s));
	int* idst = c->fp32 ? c->Q.id : c->P.id;
	if (ids) {
		HIP_CHECK(hipMemcpyAsync(idst, ids, np * sizeof(int), hipMemcpyDeviceToDevice, s));
		// new particles are numbered from one past the largest id (the reference's particlesNumber)
		c->next_id = 0;
		if (np > 0) {
			int* dmax = dalloc<int>(1);
			size_t bytes = 0;
			HIP_CHECK(rocprim::reduce(nullptr, bytes, ids, dmax, INT_MIN, (size_t)np, rocprim::maximum<int>(), s));
			char* tmp = dalloc<char>(bytes);
			HIP_CHECK(rocprim::reduce(tmp, bytes, ids, dmax, INT_MIN, (size_t)np, rocprim::maximum<int>(), s));
			int mx = 0;
			HIP_CHECK(hipMemcpyAsync(&mx, dmax, sizeof(int), hipMemcpyDeviceToHost, s));
			HIP_CHECK(hipStreamSynchronize(s));
			hipFree(tmp);
			hipFree(dmax);
			c->next_id = (long long)mx + 1;
		}
	} else {
		launch_iota(reinterpret_cast<uint32_t*>(idst), np, s);
		c->next_id = np;
	}
	if (c->fp32) {
		c->Q.count = np;
		c->Qalt.count = np;
		launch_from_fp64(c->g, c->sp, x, p, np, c->species_start, c->Q, c->status, s);
		check_status(c);
		launch_keys32(c->g, c->sp, c->Q, c->species_start, c->keys, s);
	} else {
		for (int a = 0; a < 3; ++a) {
			HIP_CHECK(hipMemcpyAsync(c->P.x[a], x + a * np, np * sizeof(double), hipMemcpyDeviceToDevice, s));
			HIP_CHECK(hipMemcpyAsync(c->P.p[a], p + a * np, np * sizeof(double), hipMemcpyDeviceToDevice, s));
		}
		c->P.count = np;
		c->Palt.count = np;
		launch_keys(c->g, c->sp, c->P, c->species_start, c->keys, c->status, s);
		check_status(c);
	}
	sort_particles(c);
	HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace

extern "C" {

int pic_upload_particles(pic_ctx* c, long long np, const int* species, const int* ids, const double* x, const double* p) {
	if (!c || np < 0 || np > c->capacity || (np > 0 && (!species || !x || !p))) {
		if (c) c->error = "pic_upload_particles: bad arguments or capacity exceeded";
		return PIC_ERR_ARGUMENT;
	}
	c->fused = 0;
	c->u2_valid = false;
	c->theo_init = false;  // the run starts again from this state
	double* dx = nullptr;
	double* dp = nullptr;
	int* di = nullptr;
	try {
		// group by species on the host so species ranges are contiguous before the key sort
		std::vector<long long> order(np);
		std::vector<int> cnt(c->sp.count + 1, 0);
		for (long long i = 0; i < np; ++i) {
			if (species[i] < 0 || species[i] >= c->sp.count) throw DeviceError{PIC_ERR_ARGUMENT, "species index out of range"};
			cnt[species[i] + 1]++;
		}
		for (int s = 0; s < c->sp.count; ++s) cnt[s + 1] += cnt[s];
		std::vector<int> pos(cnt.begin(), cnt.end() - 1);
		for (long long i = 0; i < np; ++i) order[pos[species[i]]++] = i;
		std::vector<double> hx(3 * np), hp(3 * np);
		std::vector<int> ib(np);
		for (int a = 0; a < 3; ++a)
			for (long long i = 0; i < np; ++i) {
				hx[a * np + i] = x[a * np + order[i]];
				hp[a * np + i] = p[a * np + order[i]];
			}
		for (long long i = 0; i < np; ++i) ib[i] = ids ? ids[order[i]] : (int)order[i];
		dx = dalloc<double>(3 * np);
		dp = dalloc<double>(3 * np);
		di = dalloc<int>(np);
		HIP_CHECK(hipMemcpy(dx, hx.data(), 3 * np * sizeof(double), hipMemcpyHostToDevice));
		HIP_CHECK(hipMemcpy(dp, hp.data(), 3 * np * sizeof(double), hipMemcpyHostToDevice));
		HIP_CHECK(hipMemcpy(di, ib.data(), np * sizeof(int), hipMemcpyHostToDevice));
		load_particles(c, cnt, dx, dp, di);
		hipFree(dx);
		hipFree(dp);
		hipFree(di);
	} catch (const DeviceError& e) {
		hipStreamSynchronize(c->stream);
		for (void* q : {(void*)dx, (void*)dp, (void*)di})
			if (q) hipFree(q);
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_upload_particles_device(pic_ctx* c, const long long* counts, const double* x, const double* p, const int* ids) {
	if (!c || !counts || !x || !p) return PIC_ERR_ARGUMENT;
	c->fused = 0;
	c->u2_valid = false;
	c->theo_init = false;  // the run starts again from this state
	try {
		std::vector<int> cnt(c->sp.count + 1, 0);
		long long np = 0;
		for (int s = 0; s < c->sp.count; ++s) {
			if (counts[s] < 0) throw DeviceError{PIC_ERR_ARGUMENT, "negative particle count"};
			np += counts[s];
			if (np > c->capacity) throw DeviceError{PIC_ERR_ARGUMENT, "particle capacity exceeded"};
			cnt[s + 1] = (int)np;
		}
		load_particles(c, cnt, x, p, ids);
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

long long pic_num_particles(const pic_ctx* c) { return c ? c->np() : -1; }

int pic_download_particles(pic_ctx* c, int* species, int* ids, double* x, double* p) {
	if (!c) return PIC_ERR_ARGUMENT;
	double* dx = nullptr;
	double* dp = nullptr;
	try {
		long long n = c->np();
		HIP_CHECK(hipStreamSynchronize(c->stream));
		const double* xs[3];
		const double* ps[3];
		if (c->fp32) {
			dx = dalloc<double>(3 * n);
			dp = dalloc<double>(3 * n);
			launch_to_fp64(c->g, c->sp, c->Q, c->species_start, dx, dp, c->stream);
			HIP_CHECK(hipStreamSynchronize(c->stream));
			for (int a = 0; a < 3; ++a) {
				xs[a] = dx + a * n;
				ps[a] = dp + a * n;
			}
		} else {
			for (int a = 0; a < 3; ++a) {
				xs[a] = c->P.x[a];
				ps[a] = c->P.p[a];
			}
		}
		for (int a = 0; a < 3; ++a) {
			if (x) HIP_CHECK(hipMemcpy(x + a * n, xs[a], n * sizeof(double), hipMemcpyDeviceToHost));
			if (p) HIP_CHECK(hipMemcpy(p + a * n, ps[a], n * sizeof(double), hipMemcpyDeviceToHost));
		}
		if (ids) HIP_CHECK(hipMemcpy(ids, c->fp32 ? c->Q.id : c->P.id, n * sizeof(int), hipMemcpyDeviceToHost));
		if (species)
			for (int s = 0; s < c->sp.count; ++s)
				for (long long i = c->species_start_h[s]; i < c->species_start_h[s + 1]; ++i) species[i] = s;
		if (dx) hipFree(dx);
		if (dp) hipFree(dp);
	} catch (const DeviceError& e) {
		if (dx) hipFree(dx);
		if (dp) hipFree(dp);
		return fail(c, e);
	}
	return PIC_OK;
}

static int array_ptr(pic_ctx* c, int which, double** ptr, long long* count) {
	const Grid& g = c->g;
	switch (which) {
	case PIC_ARRAY_E: *ptr = c->E; *count = 3 * g.nnodes; break;
	case PIC_ARRAY_ETHETA: *ptr = c->Etheta; *count = 3 * g.nnodes; break;
	case PIC_ARRAY_ENEW: *ptr = c->Enew; *count = 3 * g.nnodes; break;
	case PIC_ARRAY_B: *ptr = c->B; *count = 3 * g.ncells; break;
	case PIC_ARRAY_BNEW: *ptr = c->Bnew; *count = 3 * g.ncells; break;
	case PIC_ARRAY_J: *ptr = c->J; *count = 3 * g.nnodes; break;
	case PIC_ARRAY_MU: *ptr = c->mu; *count = 9 * g.nnodes; break;
	case PIC_ARRAY_DIVPI: *ptr = c->divPi; *count = 3 * g.nnodes; break;
	case PIC_ARRAY_RHOHAT: *ptr = c->rhohat; *count = g.ncells; break;
	case PIC_ARRAY_PI: *ptr = c->Pi; *count = 9 * g.ncells; break;
	case PIC_ARRAY_DENSITY: *ptr = c->dens; *count = (5 * c->sp.count + 1) * g.ncells; break;
	case PIC_ARRAY_RHS: *ptr = c->rhs; *count = 3 * g.nown; break;
	case PIC_ARRAY_SOLUTION: *ptr = c->xsol; *count = 3 * g.nown; break;
	default:
		if (which >= PIC_ARRAY_YEE_EX && which <= PIC_ARRAY_YEE_JZ && c->explicit_scheme) {
			const int k = which - PIC_ARRAY_YEE_EX, comp = k % 3;
			*ptr = k < 3 ? c->yee.E[comp] : (k < 6 ? c->yee.B[comp] : c->yee.J[comp]);
			*count = yee_count(k >= 3 && k < 6 ? c->yee.dB[comp] : c->yee.dE[comp]);
			break;
		}
		c->error = "unknown array id";
		return PIC_ERR_ARGUMENT;
	}
	return PIC_OK;
}

// The caller may write through the pointer: the cell moments formed after the last push are
// not used.
int pic_device_array(pic_ctx* c, int which, double** ptr, long long* count) {
	if (!c || !ptr || !count) return PIC_ERR_ARGUMENT;
	c->fused = 0;
	c->u2_valid = false;
	return array_ptr(c, which, ptr, count);
}

long long pic_download_array(pic_ctx* c, int which, double* out) {
	if (!c) return PIC_ERR_ARGUMENT;
	double* ptr = nullptr;
	long long n = 0;
	int rc = array_ptr(c, which, &ptr, &n);
	if (rc) return rc;
	if (!out) return n;
	try {
		HIP_CHECK(hipStreamSynchronize(c->stream));
		HIP_CHECK(hipMemcpy(out, ptr, n * sizeof(double), hipMemcpyDeviceToHost));
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return n;
}

int pic_initialize(pic_ctx* c) {
	if (!c) return PIC_ERR_ARGUMENT;
	c->fused = 0;
	c->u2_valid = false;
	try {
		const Grid& g = c->g;
		hipStream_t s = c->stream;
		launch_ghost_fill(g, c->E, 3, true, s);
		launch_ghost_fill(g, c->Etheta, 3, true, s);
		launch_ghost_fill(g, c->Enew, 3, true, s);
		launch_ghost_fill(g, c->B, 3, false, s);
		launch_ghost_fill(g, c->Bnew, 3, false, s);
		density(c);
		HIP_CHECK(hipMemcpyAsync(c->E, c->Enew, 3 * g.nnodes * sizeof(double), hipMemcpyDeviceToDevice, s));
		HIP_CHECK(hipMemcpyAsync(c->B, c->Bnew, 3 * g.ncells * sizeof(double), hipMemcpyDeviceToDevice, s));
		HIP_CHECK(hipStreamSynchronize(s));
		init_theo(c);
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_rotation_and_moments(pic_ctx* c) {
	if (!c) return PIC_ERR_ARGUMENT;
	try {
		moments(c);
		smooth(c);
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_field_rhs(pic_ctx* c) {
	if (!c) return PIC_ERR_ARGUMENT;
	c->fused = 0;
	c->u2_valid = false;
	try {
		Timed t(c, "rhs", (double)c->g.nown * (3 + 3 + 3 + 1 + 3 + 9 + 9) * 8);
		launch_rhs(c->g, c->E, c->B, c->J, c->rhohat, c->theta, c->cfg.dt, c->rhs, c->stream);
		launch_pack_mu(c->g, c->mu, c->muc, c->stream);
		launch_pack_mu_halo(c->g, c->mu, c->muh, c->stream);
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

// Appends a training pair (r, e) with A e = r to the log (float32 device vectors).
static void log_pair(pic_ctx* c, const float* r, const float* e) {
	if (!c->logging || c->log_count >= c->log_capacity || c->log_mu.empty()) return;
	long long len = c->nown3();
	HIP_CHECK(hipMemcpyAsync(c->log_r + (long long)c->log_count * len, r, len * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
	HIP_CHECK(hipMemcpyAsync(c->log_e + (long long)c->log_count * len, e, len * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
	c->log_mu_of.push_back((int)c->log_mu.size() - 1);
	++c->log_count;
}

// GMRES with FP64 iterative refinement: each outer step solves A d = r approximately with an
// FP32 basis and FP32 operator (dots accumulated and the Hessenberg least squares kept in
// FP64), then x += d and r = b - A x in FP64, until ||r|| / ||b|| <= tol.
// With the learned preconditioner M the inner solve is flexible GMRES (right preconditioned,
// z_j = M v_j, x = Z y). The rest of the solve switches to plain GMRES (a fallback) when a
// preconditioned cycle converges slowly (Givens estimate above 1e-2 after 8 steps: worse than
// about 0.56 per step; the cycle stops there and its partial update is kept), leaves the true
// residual above half its previous value, or produces a non-finite least-squares solution.
static void mixed_solve(pic_ctx* c, const double* b, double* x, pic_solve_info& inf) {
	hipStream_t s = c->stream;
	const Grid& g = c->g;
	long long len = c->nown3();
	const double C = operator_C(c);
	double* d_bn = c->dscal + 4;
	double* d_rn = c->dscal + 5;
	double* d_h = c->dscal + 8;
	launch_to_f32(c->muc, c->mu32, 9 * g.nown, s);
	launch_to_f32(c->muh, c->mu32h, 9 * halo_count(g), s);
	if (c->logging && c->log_count < c->log_capacity) {
		if (c->mu_scale <= 0) {  // the network's input scale of mu: its rms over the first logged solve
			launch_dotT<float>(c->mu32, c->mu32, 9 * g.nown, c->partial, d_h, false, s);
			double ss = 0;
			HIP_CHECK(hipMemcpyAsync(&ss, d_h, sizeof(double), hipMemcpyDeviceToHost, s));
			HIP_CHECK(hipStreamSynchronize(s));
			c->mu_scale = ss > 0 ? std::sqrt(ss / (9.0 * g.nown)) : 1.0;
		}
		float* m = dalloc<float>(9 * g.nown);
		HIP_CHECK(hipMemcpyAsync(m, c->mu32, 9 * g.nown * sizeof(float), hipMemcpyDeviceToDevice, s));
		c->log_mu.push_back(m);
	}
	bool pre = c->use_net && c->net != nullptr;
	HIP_CHECK(hipMemsetAsync(x, 0, len * sizeof(double), s));
	HIP_CHECK(hipMemcpyAsync(c->work, b, len * sizeof(double), hipMemcpyDeviceToDevice, s));  // r = b
	launch_dotT<double>(b, b, len, c->partial, d_bn, true, s);
	HIP_CHECK(hipMemcpyAsync(d_rn, d_bn, sizeof(double), hipMemcpyDeviceToDevice, s));
	double bnorm = 0;
	HIP_CHECK(hipMemcpyAsync(&bnorm, d_bn, sizeof(double), hipMemcpyDeviceToHost, s));
	HIP_CHECK(hipStreamSynchronize(s));
	double rnorm = bnorm, rel = 0;
	const double inner_floor = 1e-5;  // what an FP32 basis reliably delivers per cycle
	int limit = std::min(c->cfg.gmres_max_iterations, c->max_basis + 1);
	while (bnorm > 0 && rnorm / bnorm > c->cfg.gmres_tolerance && inf.restarts < 30) {
		double target = std::max(inner_floor, 0.5 * c->cfg.gmres_tolerance * bnorm / rnorm);
		launch_scaled_copy<float>(c->work, c->V32[0], d_rn, len, s);
		Givens gv;
		int nn = 2;
		rel = 1.0;
		std::vector<double> hcol;
		const bool cycle_pre = pre;
		bool slow = false;
		while (rel > target && nn < limit) {
			int j = nn - 2;
			float* w = c->V32[j + 1];
			const float* dir = c->V32[j];
			if (cycle_pre) {
				Timed t(c, "precond_net", 0);
				c->net->forward(c->V32[j], c->mu32, (float)c->mu_scale, c->Z32[j], s);
				dir = c->Z32[j];
			}
			{
				Timed t(c, "apply_operator32", (double)g.nown * 15 * 4);
				launch_apply_factored<float>(g, c->mu32h, dir, c->w32, C, w, s);
			}
			log_pair(c, w, dir);
			{
				Timed t(c, "orthogonalize32", (double)len * 4 * (5.0 * (j + 1) + 3));
				for (int m = 0; m <= j; ++m) {
					launch_dotT<float>(c->V32[m], w, len, c->partial, d_h + m, false, s);
					launch_axpyT<float>(w, c->V32[m], d_h + m, len, s);
				}
				launch_dotT<float>(w, w, len, c->partial, d_h + j + 1, true, s);
				launch_normalizeT<float>(w, d_h + j + 1, len, s);
			}
			hcol.assign(nn, 0.0);
			HIP_CHECK(hipMemcpyAsync(hcol.data(), d_h, nn * sizeof(double), hipMemcpyDeviceToHost, s));
			HIP_CHECK(hipStreamSynchronize(s));
			gv.update(nn, hcol);
			rel = std::fabs(gv.Q[nn - 1][0]);
			nn++;
			if (cycle_pre) inf.preconditioned += 1;
			if (!std::isfinite(rel)) break;
			if (cycle_pre && nn - 2 >= 8 && rel > 1e-2) {
				slow = true;
				break;
			}
		}
		inf.iterations += nn - 2;
		inf.restarts += 1;
		bool finite = std::isfinite(rel);
		for (double yv : gv.y) finite = finite && std::isfinite(yv);
		if (!finite) {  // only a preconditioned cycle can get here: drop it and go on without the network
			pre = false;
			inf.fallbacks += 1;
			continue;
		}
		HIP_CHECK(hipMemcpyAsync(c->ydev, gv.y.data(), gv.y.size() * sizeof(double), hipMemcpyHostToDevice, s));
		double rprev = rnorm;
		{
			Timed t(c, "refinement", (double)len * 8 * (gv.y.size() / 2.0 + 2 + 15.0 / 3 + 3));
			launch_update_x<float>(x, cycle_pre ? c->Z32dev : c->V32dev, c->ydev, (int)gv.y.size(), d_rn, len, s);  // x += ||r|| Z y
			launch_apply_factored<double>(g, c->muh, x, c->w64, C, c->work2, s);  // A x in FP64
			launch_residual(b, c->work2, c->work, len, s);                         // r = b - A x
			launch_dotT<double>(c->work, c->work, len, c->partial, d_rn, true, s);
		}
		HIP_CHECK(hipMemcpyAsync(&rnorm, d_rn, sizeof(double), hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipStreamSynchronize(s));
		if (cycle_pre && (slow || !(rnorm <= 0.5 * rprev))) {
			pre = false;
			inf.fallbacks += 1;
		}
		if (nn == 2) break;
	}
	inf.true_residual = bnorm > 0 ? rnorm / bnorm : 0.0;
	inf.estimated_residual = rel;
	if (c->logging && c->log_count < c->log_capacity && bnorm > 0) {  // the solve itself: A x = b
		launch_to_f32(b, c->tmp32, len, s);
		launch_to_f32(x, c->tmp32 + len, len, s);
		log_pair(c, c->tmp32, c->tmp32 + len);
	}
	if (inf.true_residual > c->cfg.gmres_tolerance)
		throw DeviceError{PIC_ERR_SOLVER, "mixed-precision GMRES did not reach the requested true residual"};
}

static int solve_impl(pic_ctx* c, double* b, double* x, pic_solve_info* info) {
	hipStream_t s = c->stream;
	HIP_CHECK(hipStreamSynchronize(s));  // the solve time below excludes work queued before it
	auto t0 = std::chrono::steady_clock::now();
	if (c->cfg.solver_precision == PIC_SOLVER_MIXED) {
		pic_solve_info inf{0, 0, 0, -1, 0, 0, 0};
		mixed_solve(c, b, x, inf);
		HIP_CHECK(hipStreamSynchronize(s));
		inf.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
		if (info) *info = inf;
		return PIC_OK;
	}
	int dim = 3 * c->g.n[0] * c->g.n[1] * c->g.n[2];
	double rel = 0;
	pic_solve_info inf{0, 0, 0, -1, 0, 0, 0};
	long long len = c->nown3();
	if (c->cfg.stop_rule == PIC_STOP_REFERENCE) {
		// generalizedMinimalResidualMethod: b is normalised in place there; a copy keeps rhs intact here
		HIP_CHECK(hipMemcpyAsync(c->work, b, len * sizeof(double), hipMemcpyDeviceToDevice, s));
		double target = std::max(c->cfg.gmres_tolerance / dim, 1e-15);
		inf.iterations = gmres_cycle(c, c->work, x, target, c->cfg.gmres_max_iterations, dim, rel);
		inf.estimated_residual = rel;
	} else {
		// restarted GMRES on the residual until ||b - A x|| / ||b|| <= tol
		const double C = operator_C(c);
		HIP_CHECK(hipMemsetAsync(x, 0, len * sizeof(double), s));
		launch_dot(b, b, len, c->partial, c->dscal + 4, true, s);
		double bnorm = 0;
		HIP_CHECK(hipMemcpyAsync(&bnorm, c->dscal + 4, sizeof(double), hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipStreamSynchronize(s));
		double tr = 0, rnorm = bnorm;
		HIP_CHECK(hipMemcpyAsync(c->work, b, len * sizeof(double), hipMemcpyDeviceToDevice, s));
		while (bnorm > 0) {
			tr = rnorm / bnorm;
			if (tr <= c->cfg.gmres_tolerance || inf.restarts >= 50) break;
			double target = std::max(0.5 * c->cfg.gmres_tolerance * bnorm / rnorm, 1e-15);
			int it = gmres_cycle(c, c->work, c->work2, target, c->cfg.gmres_max_iterations, dim, rel);
			inf.iterations += it;
			inf.restarts += 1;
			launch_add(x, c->work2, len, s);
			launch_apply_operator(c->g, c->muh, x, C, c->work2, s);
			launch_sub(c->work, b, c->work2, len, s);  // r = b - A x
			launch_dot(c->work, c->work, len, c->partial, c->dscal + 5, true, s);
			HIP_CHECK(hipMemcpyAsync(&rnorm, c->dscal + 5, sizeof(double), hipMemcpyDeviceToHost, s));
			HIP_CHECK(hipStreamSynchronize(s));
			if (it == 0) break;
		}
		inf.true_residual = bnorm > 0 ? rnorm / bnorm : 0.0;
		inf.estimated_residual = rel;
		if (inf.true_residual > c->cfg.gmres_tolerance)
			throw DeviceError{PIC_ERR_SOLVER, "GMRES did not reach the requested true residual"};
	}
	HIP_CHECK(hipStreamSynchronize(s));
	inf.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	if (info) *info = inf;
	return PIC_OK;
}

int pic_maxwell_solve(pic_ctx* c, pic_solve_info* info) {
	if (!c) return PIC_ERR_ARGUMENT;
	c->fused = 0;
	c->u2_valid = false;
	try {
		int rc = pic_field_rhs(c);
		if (rc) return rc;
		rc = solve_impl(c, c->rhs, c->xsol, info);
		if (rc) return rc;
		Timed t(c, "unpack_e", (double)c->g.nnodes * 9 * 8);
		launch_unpack_e(c->g, c->xsol, c->E, c->theta, c->Etheta, c->Enew, c->stream);
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_gmres_solve(pic_ctx* c, const double* rhs, double* x, pic_solve_info* info) {
	if (!c || !rhs || !x) return PIC_ERR_ARGUMENT;
	c->fused = 0;
	c->u2_valid = false;
	try {
		return solve_impl(c, const_cast<double*>(rhs), x, info);
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
}

int pic_apply_operator(pic_ctx* c, const double* v, double* y) {
	if (!c || !v || !y) return PIC_ERR_ARGUMENT;
	c->fused = 0;
	c->u2_valid = false;
	try {
		launch_pack_mu(c->g, c->mu, c->muc, c->stream);
		launch_pack_mu_halo(c->g, c->mu, c->muh, c->stream);
		launch_apply_operator(c->g, c->muh, v, operator_C(c), y, c->stream);
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_apply_operator_factored(pic_ctx* c, int precision, const void* v, void* y) {
	if (!c || !v || !y) return PIC_ERR_ARGUMENT;
	if (precision != PIC_PRECISION_FP64 && precision != PIC_PRECISION_FP32) return PIC_ERR_ARGUMENT;
	if (precision == PIC_PRECISION_FP32 && c->mu32h == nullptr) {
		c->error = "the FP32 factored operator needs a context created with PIC_SOLVER_MIXED";
		return PIC_ERR_ARGUMENT;
	}
	c->fused = 0;
	c->u2_valid = false;
	try {
		launch_pack_mu(c->g, c->mu, c->muc, c->stream);
		launch_pack_mu_halo(c->g, c->mu, c->muh, c->stream);
		if (precision == PIC_PRECISION_FP64) {
			launch_apply_factored<double>(c->g, c->muh, static_cast<const double*>(v), c->w64, operator_C(c),
			                              static_cast<double*>(y), c->stream);
		} else {
			launch_to_f32(c->muh, c->mu32h, 9 * halo_count(c->g), c->stream);
			launch_apply_factored<float>(c->g, c->mu32h, static_cast<const float*>(v), c->w32, operator_C(c),
			                             static_cast<float*>(y), c->stream);
		}
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

// maxDivergenceCleanupIterations (PIC++/constants.h:56)
constexpr int kCleanupMaxIterations = 100;

// One divergence cleanup: the right part over the owned points (rhs_fn), the 27-point potential
// equation solved by GMRES with the reference's stop rule, the correction (correct_fn).
// target: the Givens estimate to stop at (0: the reference's GMRES rule).
static int cleanup_solve(pic_ctx* c, pic_solve_info* info, const std::function<void(double*)>& rhs_fn,
                         const std::function<void(const double*)>& correct_fn, double target = 0.0) {
	if (!c) return PIC_ERR_ARGUMENT;
	if (c->explicit_scheme || c->g.wall) {
		c->error = "divergence cleanup is built for the implicit scheme on the periodic box";
		return PIC_ERR_ARGUMENT;
	}
	c->fused = 0;
	c->u2_valid = false;
	try {
		const Grid& g = c->g;
		hipStream_t s = c->stream;
		const long long len = (long long)g.n[0] * g.n[1] * g.n[2];
		const int max_basis = kCleanupMaxIterations - 2;  // n < maxIteration Arnoldi steps (specialmath.cpp:481)
		if (c->Vcl.empty()) {
			c->Vcl.resize(max_basis + 1);
			for (auto& v : c->Vcl) v = dalloc<double>(len);
			c->Vcl_dev = dalloc<double*>(c->Vcl.size());
			HIP_CHECK(hipMemcpy(c->Vcl_dev, c->Vcl.data(), c->Vcl.size() * sizeof(double*), hipMemcpyHostToDevice));
			c->cl_rhs = dalloc<double>(len);
			c->cl_phi = dalloc<double>(len);
		}
		{
			Timed t(c, "cleanup_rhs", (double)g.nnodes * 24 + (double)g.ncells * 8 + (double)len * 8);
			rhs_fn(c->cl_rhs);
		}
		const CleanupStencil S = cleanup_stencil(g);
		KrylovSystem K{len, [c, S](const double* v, double* w) { launch_cleanup_apply(c->g, S, v, w, c->stream); }, "cleanup_apply",
		               (double)len * 16, &c->Vcl, c->Vcl_dev, max_basis, PIC_ORTHO_MGS};
		const int dim = (int)len;  // one component (generalizedMinimalResidualMethod with lnumber = 1)
		if (target <= 0) target = std::max(c->cfg.gmres_tolerance / dim, 1e-15);
		double rel = 0;
		pic_solve_info inf{0, 0, 0, -1, 0, 0, 0};
		HIP_CHECK(hipStreamSynchronize(s));
		auto t0 = std::chrono::steady_clock::now();
		inf.iterations = gmres_cycle_on(c, K, c->cl_rhs, c->cl_phi, target, kCleanupMaxIterations, dim, rel);
		inf.estimated_residual = rel;
		{
			Timed t(c, "cleanup_correct", (double)len * 8 + (double)g.nnodes * 48);
			correct_fn(c->cl_phi);
		}
		HIP_CHECK(hipStreamSynchronize(s));
		inf.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
		if (info) *info = inf;
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_cleanup_divergence(pic_ctx* c, pic_solve_info* info) {
	if (!c) return PIC_ERR_ARGUMENT;
	return cleanup_solve(
	    c, info,
	    [c](double* rhs) { launch_cleanup_rhs(c->g, c->Enew, c->dens + (long long)5 * c->sp.count * c->g.ncells, rhs, c->stream); },
	    [c](const double* phi) {
		    launch_cleanup_correct(c->g, phi, c->Enew, c->stream);
		    launch_ghost_fill(c->g, c->Enew, 3, true, c->stream);  // exchangeGeneralEfield
	    });
}

int pic_cleanup_divergence_magnetic(pic_ctx* c, pic_solve_info* info) {
	if (!c) return PIC_ERR_ARGUMENT;
	return cleanup_solve(
	    c, info, [c](double* rhs) { launch_cleanup_rhs_b(c->g, c->Bnew, rhs, c->stream); },
	    [c](const double* phi) {
		    launch_cleanup_correct_b(c->g, phi, c->Bnew, c->stream);
		    launch_ghost_fill(c->g, c->Bnew, 3, false, c->stream);  // exchangeGeneralBfield
	    },
	    // the reference's BiCGStab reaches rounding level on these systems: the GMRES floor
	    1e-15);
}

int pic_filter_fields(pic_ctx* c, int cut_wave_number) {
	if (!c || cut_wave_number <= 0) return PIC_ERR_ARGUMENT;
	if (c->explicit_scheme || c->g.wall) {
		c->error = "Fourier filtering is built for the implicit scheme on the periodic box";
		return PIC_ERR_ARGUMENT;
	}
	c->fused = 0;
	c->u2_valid = false;
	try {
		const Grid& g = c->g;
		hipStream_t s = c->stream;
		const long long len = (long long)g.n[0] * g.n[1] * g.n[2];
		if (!c->ft_re) {
			for (double** p : {&c->ft_re, &c->ft_im, &c->ft_re2, &c->ft_im2}) *p = dalloc<double>(len);
			for (int a = 0; a < 3; ++a) {
				// (cos, sin)(2 pi q / n), q < n, from the exact remainder q (the reference's recurrence
				// of complexExp products gives the same factors to rounding)
				std::vector<double> tw(2 * g.n[a]);
				for (int q = 0; q < g.n[a]; ++q) {
					const double ang = 2.0 * 3.141592653589793115997963468544185161590576171875 * q / g.n[a];
					tw[2 * q] = std::cos(ang);
					tw[2 * q + 1] = std::sin(ang);
				}
				c->ft_tw[a] = dalloc<double>(tw.size());
				HIP_CHECK(hipMemcpy(c->ft_tw[a], tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice));
			}
		}
		const double kcut = 2 * 3.141592653589793115997963468544185161590576171875 / (cut_wave_number * g.d[0]);
		const double* tw[3] = {c->ft_tw[0], c->ft_tw[1], c->ft_tw[2]};
		Timed t(c, "filter_fields", (double)len * 8 * 2 * 6 * (g.n[0] + g.n[1] + g.n[2]));
		for (int comp = 0; comp < 3; ++comp)
			launch_filter_component(g, c->Enew, comp * g.nnodes, true, comp, kcut, tw, c->ft_re, c->ft_im, c->ft_re2, c->ft_im2, s);
		for (int comp = 0; comp < 3; ++comp)
			launch_filter_component(g, c->Bnew, comp * g.ncells, false, comp, kcut, tw, c->ft_re, c->ft_im, c->ft_re2, c->ft_im2, s);
		launch_ghost_fill(g, c->Enew, 3, true, s);   // exchangeGeneralEfield(newEfield)
		launch_ghost_fill(g, c->Bnew, 3, false, s);  // exchangeGeneralBfield(newBfield)
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_update_b(pic_ctx* c) {
	if (!c) return PIC_ERR_ARGUMENT;
	c->fused = 0;
	c->u2_valid = false;
	try {
		Timed t(c, "update_b", (double)c->g.ncells * 6 * 8 + (double)c->g.nnodes * 3 * 8);
		launch_update_b(c->g, c->Etheta, c->B, c->cfg.dt, c->Bnew, c->stream);
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_push(pic_ctx* c) {
	if (!c) return PIC_ERR_ARGUMENT;
	c->fused = 0;
	c->u2_valid = false;
	try {
		if (!c->theo_init) init_theo(c);
		if (c->fp32) {
			const Grid& g = c->g;
			{
				Timed t(c, "fields_to_fp32", (double)(2 * g.nnodes + g.ncells) * 3 * 12);
				launch_f64_to_f32(c->E, c->E32, g.nnodes, c->stream);
				launch_f64_to_f32(c->Etheta, c->Et32, g.nnodes, c->stream);
				launch_f64_to_f32(c->B, c->B32, g.ncells, c->stream);
			}
			// algorithmic bytes: the particle state read once and written once, 2 x 28 B (the
			// per-particle unit of the step; the key, the mover bits and the field reads are extra)
			Timed t(c, "push", (double)c->np() * 56);
			launch_push32(g, c->sp, c->Q, c->species_start, c->E32, c->Et32, c->B32, (float)c->cfg.dt, (float)c->eta, c->keys,
			              c->mask, c->status, c->dead, c->esc, c->bounce, c->stream);
		} else {
			Timed t(c, "push", (double)c->P.count * 96);  // 2 x 48 B of particle state
			launch_push(c->g, c->sp, c->P, c->species_start, c->E, c->Etheta, c->B, c->cfg.dt, c->eta, c->keys, c->mask, c->status,
			            c->dead, c->esc, c->bounce, c->stream);
		}
		check_status(c);
		{
			// the theoretical energy and momentum (updateTheoreticalEnergy, PIC++/simulation.cpp:1399-1504,
			// after the mover; E and B are still E^n, B^n): the wall's reflections, then the
			// Poynting flux through the x boundaries (zero with periodic x: the same values on both
			// sides); escapes and injected particles follow in wall_particles
			Timed t(c, "theoretical_energy", 0);
			if (c->g.wall)
				launch_theo_bounce(c->sp, c->P, c->Q, c->fp32, c->species_start, c->bounce, c->np(), c->epartial, c->theo,
				                   c->stream);
			launch_theo_poynting(c->g, implicit_view(c->g, c->E, c->B), c->cfg.dt, c->theo, c->stream);
		}
		if (c->g.wall)  // removals, replacements and injection, then the rebin
			wall_particles(c);
		else
			rebin_particles(c);
		density(c, true);
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_commit_fields(pic_ctx* c) {
	if (!c) return PIC_ERR_ARGUMENT;
	c->fused = c->fused == 1 ? 2 : 0;  // E, B become the Enew, Bnew the push's density pass used
	try {
		const Grid& g = c->g;
		HIP_CHECK(hipMemcpyAsync(c->E, c->Enew, 3 * g.nnodes * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
		HIP_CHECK(hipMemcpyAsync(c->B, c->Bnew, 3 * g.ncells * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_energy(pic_ctx* c, double* out7) {
	if (!c || !out7) return PIC_ERR_ARGUMENT;
	try {
		double h[9];
		energy_sums(c, h);
		out7[0] = h[5];
		out7[1] = h[0];
		out7[2] = h[1];
		out7[3] = h[5] + h[0] + h[1];
		out7[4] = h[2] + h[6];
		out7[5] = h[3] + h[7];
		out7[6] = h[4] + h[8];
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_general(pic_ctx* c, const pic_backup_info* info, double* out) {
	if (!c || !info || !out) return PIC_ERR_ARGUMENT;
	try {
		const Grid& g = c->g;
		hipStream_t s = c->stream;
		if (!c->theo_init) init_theo(c);
		double h[9];
		energy_sums(c, h);
		const FieldView v = c->explicit_scheme ? yee_view(c->yee) : implicit_view(g, c->E, c->B);
		launch_general_fields(g, v, c->gpartial, c->gout, s);
		double gr[kGeneralRows], th[4], e000[3], b000[3];
		HIP_CHECK(hipMemcpyAsync(gr, c->gout, sizeof(gr), hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipMemcpyAsync(th, c->theo, sizeof(th), hipMemcpyDeviceToHost, s));
		for (int a = 0; a < 3; ++a) {  // the values updateParameters starts from: E[0][0][0], B[0][0][0] - B0
			HIP_CHECK(hipMemcpyAsync(&e000[a], v.e[a], sizeof(double), hipMemcpyDeviceToHost, s));
			HIP_CHECK(hipMemcpyAsync(&b000[a], v.b[a], sizeof(double), hipMemcpyDeviceToHost, s));
		}
		HIP_CHECK(hipStreamSynchronize(s));
		const double pi = 4 * std::atan2(1.0, 1.0);
		const double ms = info->scale_factor / info->plasma_period, es = ms * ms;
		const double ff = info->plasma_period * std::sqrt(info->scale_factor);
		double e0 = 0, b0 = 0;
		for (int a = 0; a < 3; ++a) {
			e0 += e000[a] * e000[a];
			b0 += (b000[a] - info->B0[a]) * (b000[a] - info->B0[a]);
		}
		out[0] = h[5] * es;
		out[1] = h[0] * es;
		out[2] = h[1] * es;
		out[3] = (h[5] + h[0] + h[1]) * es;
		for (int a = 0; a < 3; ++a) out[4 + a] = (h[2 + a] + h[6 + a]) * ms;
		out[7] = th[0] * es;
		for (int a = 0; a < 3; ++a) out[8 + a] = th[1 + a] * ms;
		out[11] = std::sqrt(std::max(e0, gr[9])) / ff;  // the largest |E| of updateParameters
		out[12] = std::sqrt(std::max(b0, gr[10])) / ff;
		out[13] = (double)c->next_id;  // particlesNumber: the particles numbered so far
		const double ncell = (double)g.n[0] * g.n[1] * g.n[2];
		for (int a = 0; a < 3; ++a) {
			out[14 + a] = std::sqrt(gr[a] / ncell) / ff;
			out[17 + a] = gr[a] * g.volume / (8 * pi) * es;
			out[20 + a] = gr[3 + a] * g.volume / (8 * pi) * es;
			out[23 + a] = h[2 + a] * ms;
			out[26 + a] = h[6 + a] * ms;
		}
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_write_general(pic_ctx* c, const pic_backup_info* info, const char* path) {
	if (!c || !info || !path) return PIC_ERR_ARGUMENT;
	double r[PIC_GENERAL_VALUES];
	const int rc = pic_general(c, info, r);
	if (rc) return rc;
	try {
		FILE* f = std::fopen(path, "a");
		if (!f) throw DeviceError{PIC_ERR_ARGUMENT, std::string("cannot open ") + path};
		const double t = info->time, dx = c->g.d[0];
		// outputGeneral's two fprintf calls (PIC++/output.cpp:3229-3249); shockWaveX and the
		// derivative points keep their initial values, their updates being off in the loop
		if (std::fprintf(f,
		                 "%d %15.10g %15.10g %15.10g %15.10g %15.10g %15.10g %15.10g %15.10g %15.10g %15.10g %15.10g %15.10g "
		                 "%15.10g %15.10g %15.10g %15.10g %d %15.10g %15.10g %15.10g %15.10g ",
		                 info->current_iteration, t, t * info->plasma_period, r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8],
		                 r[9], r[10], r[11], r[12], c->cfg.dt, (int)r[13], -1.0, r[14], r[15], r[16]) < 0 ||
		    std::fprintf(f,
		                 "%15.10g %15.10g %15.10g %15.10g %15.10g %15.10g %15.10g %15.10g %15.10g %15.10g %15.10g %15.10g %d %15.10g "
		                 "%d %15.10g\n",
		                 r[17], r[18], r[19], r[20], r[21], r[22], r[23], r[24], r[25], r[26], r[27], r[28], 0,
		                 0 * dx * info->scale_factor, 0, 0 * dx * info->scale_factor) < 0) {
			std::fclose(f);
			throw DeviceError{PIC_ERR_ARGUMENT, std::string("cannot write ") + path};
		}
		if (std::fclose(f) != 0) throw DeviceError{PIC_ERR_ARGUMENT, std::string("cannot close ") + path};
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_upload_yee(pic_ctx* c, const double* const* fields6) {
	if (!c || !fields6) return PIC_ERR_ARGUMENT;
	if (!c->explicit_scheme) {
		c->error = "pic_upload_yee: the context was not created with PIC_SCHEME_EXPLICIT";
		return PIC_ERR_ARGUMENT;
	}
	c->theo_init = false;  // the run starts again from this state
	try {
		for (int k = 0; k < 6; ++k) {
			const int comp = k % 3;
			if (!fields6[k]) throw DeviceError{PIC_ERR_ARGUMENT, "pic_upload_yee: null field"};
			double* dst = k < 3 ? c->yee.E[comp] : c->yee.B[comp];
			const long long n = yee_count(k < 3 ? c->yee.dE[comp] : c->yee.dB[comp]);
			HIP_CHECK(hipMemcpyAsync(dst, fields6[k], n * sizeof(double), hipMemcpyHostToDevice, c->stream));
		}
		launch_yee_exchange(c->g, c->yee, true, c->stream);  // PIC++/simulation.cpp:83-85
		launch_yee_exchange(c->g, c->yee, false, c->stream);
		HIP_CHECK(hipStreamSynchronize(c->stream));
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_explicit_step(pic_ctx* c) {
	if (!c) return PIC_ERR_ARGUMENT;
	if (!c->explicit_scheme) {
		c->error = "pic_explicit_step: the context was not created with PIC_SCHEME_EXPLICIT";
		return PIC_ERR_ARGUMENT;
	}
	try {
		const Grid& g = c->g;
		hipStream_t s = c->stream;
		const double dt = c->cfg.dt;
		if (!c->theo_init) init_theo(c);
		{
			Timed t(c, "yee_b_half", (double)g.ncells * 3 * 8 * 3);
			launch_tristan_b_half(g, c->yee, dt, s);
			launch_yee_exchange(g, c->yee, true, s);
		}
		// Fixed-point scale of the current. A particle's part of a zig-zag path holds at most
		// 2 |q w / V| |v| (v its new velocity): an entry takes at most 2 parts from each particle
		// whose cell before or after the move touches it (4 + 4 cells; the particles of the 4
		// cells after the move come from their 27 neighbours), so it stays below
		// 512 m max(|q w / V| |v|) with m particles in the fullest cell. The new momentum obeys
		// |u_new| <= |u| + |q| dt |E| / m (the rotation keeps |u|; each half kick adds at most
		// |q| dt |E| / 2m, |E| bounded by the components' largest grid values) and
		// |v| = |u| / sqrt(1 + |u|^2). The scale is the largest power of two keeping that sum
		// below 2^62 and one contribution below 2^50 (the push's rounding to integers): the
		// rounding of a contribution is then at most 2^-62 of the sum's bound, i.e. relative to
		// the particles' actual speeds, not to c.
		ExplicitBounds hb{};
		const size_t slot_bytes = sizeof(unsigned long long) * kU2Slots * kMaxSpecies;
		HIP_CHECK(hipMemsetAsync(c->bounds, 0, sizeof(ExplicitBounds), s));
		if (!c->u2_valid) HIP_CHECK(hipMemsetAsync(c->u2_slots, 0, slot_bytes, s));  // else the last push's maxima
		launch_explicit_bounds(g, c->sp, c->P, c->species_start, c->start, c->yee, !c->u2_valid, c->u2_slots, c->bounds, s);
		HIP_CHECK(hipMemcpyAsync(&hb, c->bounds, sizeof(hb), hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipStreamSynchronize(s));
		auto bits_to_double = [](unsigned long long b) {
			double v;
			memcpy(&v, &b, sizeof(v));
			return v;
		};
		double e2 = 0;
		for (int a = 0; a < 3; ++a) e2 += bits_to_double(hb.emax[a]) * bits_to_double(hb.emax[a]);
		const double emag = std::sqrt(e2);
		double fv = 0, u2_limit[kMaxSpecies] = {};
		for (int q = 0; q < c->sp.count; ++q) {
			const double U = (std::sqrt(bits_to_double(hb.u2max[q])) + std::fabs(c->sp.charge[q]) * dt * emag / c->sp.mass[q]) *
			                 (1.0 + 1e-6);  // margin for the rounding of the particles' own arithmetic
			u2_limit[q] = U * U;
			fv = std::max(fv, std::fabs(c->sp.charge[q] * c->sp.weight[q] / g.volume) * (U / std::sqrt(1.0 + U * U)));
		}
		if (!std::isfinite(fv) || !std::isfinite(emag))
			throw DeviceError{PIC_ERR_PARTICLE, "pic_explicit_step: a particle momentum or the electric field is not finite"};
		int e = 0;
		if (fv > 0) {
			const double bound = 512.0 * std::max(hb.maxcount, 1) * fv;
			e = std::min((int)std::floor(62.0 - std::log2(bound)), (int)std::floor(49.0 - std::log2(fv)));
			e = std::max(-1000, std::min(1000, e));
		}
		const double scale = std::ldexp(1.0, e);
		HIP_CHECK(hipMemsetAsync(c->u2_slots, 0, slot_bytes, s));  // for the push's maxima
		for (int comp = 0; comp < 3; ++comp)
			HIP_CHECK(hipMemsetAsync(c->yee.Jacc[comp], 0, yee_count(c->yee.dE[comp]) * sizeof(long long), s));
		{
			Timed t(c, "tristan_push", (double)c->P.count * 100 + (double)g.ncells * 6 * 8 * 8);
			launch_tristan_push(g, c->sp, c->P, c->species_start, c->yee, dt, scale, u2_limit, c->keys, c->mask, c->status,
			                    c->u2_slots, c->dead, c->esc, c->bounce, s);
		}
		c->u2_valid = true;
		check_status(c);
		if (g.wall)  // the reflections' momentum for the theoretical totals (mover.cpp:749)
			launch_theo_bounce(c->sp, c->P, c->Q, c->fp32, c->species_start, c->bounce, c->np(), c->epartial, c->theo, s);
		{
			Timed t(c, "yee_b_half", (double)g.ncells * 3 * 8 * 3);
			launch_tristan_b_half(g, c->yee, dt, s);
			launch_yee_exchange(g, c->yee, true, s);
		}
		{
			Timed t(c, "yee_current", (double)g.ncells * 3 * 16);
			launch_yee_current(g, c->yee, 1.0 / scale, s);
		}
		if (c->cfg.smoothing_count > 0) {
			// smoothBunemanEfieldGeneral passes (PIC++/simulation.cpp:185-188); the fixed-point
			// accumulators, done with for this step, hold the scratch copies
			Timed t(c, "smoothing", (double)c->cfg.smoothing_count * g.ncells * 3 * 8 * 3);
			const Smooth27 sm = smoothing_matrix(c->cfg.smoothing_parameter, g.cz != 0);
			double* const tmp[3] = {reinterpret_cast<double*>(c->yee.Jacc[0]), reinterpret_cast<double*>(c->yee.Jacc[1]),
			                        reinterpret_cast<double*>(c->yee.Jacc[2])};
			for (int n = 0; n < c->cfg.smoothing_count; ++n) launch_smooth_yee_current(g, c->yee, sm, tmp, s);
		}
		{
			Timed t(c, "yee_e", (double)g.ncells * 3 * 8 * 4);
			launch_tristan_e(g, c->yee, dt, s);
			launch_yee_exchange(g, c->yee, false, s);
		}
		// the Poynting flux through the x boundaries with the new fields (updateTheoreticalEnergy
		// after the BUNEMAN branch, PIC++/simulation.cpp:272); zero with periodic x
		launch_theo_poynting(g, yee_view(c->yee), dt, c->theo, s);
		if (g.wall)  // removals, replacements and injection, then the rebin
			wall_particles(c);
		else
			rebin_particles(c);
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_step(pic_ctx* c, pic_solve_info* info) {
	int rc;
	if (c && c->explicit_scheme) return pic_explicit_step(c);
	if ((rc = pic_rotation_and_moments(c))) return rc;
	if ((rc = pic_maxwell_solve(c, info))) return rc;
	if ((rc = pic_update_b(c))) return rc;
	if ((rc = pic_push(c))) return rc;
	// the reference's call position of cleanupDivergence (PIC++/simulation.cpp:348): after the
	// push's updateDensityParameters, before updateFields
	if ((c->cfg.divergence_cleanup & 1) && (rc = pic_cleanup_divergence(c, nullptr))) return rc;
	if ((c->cfg.divergence_cleanup & 2) && (rc = pic_cleanup_divergence_magnetic(c, nullptr))) return rc;
	if (c->cfg.filter_cut_wave_number > 0 && (rc = pic_filter_fields(c, c->cfg.filter_cut_wave_number))) return rc;
	return pic_commit_fields(c);
}

// ---------------------------------------------------------------- checkpoints and backups

}  // extern "C"

namespace {

constexpr char kCkptMagic[8] = {'P', 'I', 'C', 'A', 'M', 'D', 'C', 'K'};
constexpr char kCkptEnd[8] = {'C', 'K', 'P', 'T', 'E', 'N', 'D', '\0'};
constexpr uint32_t kCkptVersion = 2;  // 2: the theoretical energy and momentum
constexpr size_t kStage = size_t(64) << 20;  // host staging per copy, bytes

struct File {
	FILE* f = nullptr;
	std::string path;
	File(const char* p, const char* mode) : f(std::fopen(p, mode)), path(p) {
		if (!f) throw DeviceError{PIC_ERR_ARGUMENT, std::string("cannot open ") + p};
	}
	~File() {
		if (f) std::fclose(f);
	}
	void put(const void* data, size_t n) {
		if (n && std::fwrite(data, 1, n, f) != n) throw DeviceError{PIC_ERR_ARGUMENT, "short write to " + path};
	}
	void get(void* data, size_t n) {
		if (n && std::fread(data, 1, n, f) != n) throw DeviceError{PIC_ERR_ARGUMENT, "short read from " + path};
	}
	template <class T>
	void put(T v) {
		put(&v, sizeof(T));
	}
	template <class T>
	T get() {
		T v;
		get(&v, sizeof(T));
		return v;
	}
	void close() {
		if (std::fclose(f) != 0) {
			f = nullptr;
			throw DeviceError{PIC_ERR_ARGUMENT, "cannot close " + path};
		}
		f = nullptr;
	}
};

// device bytes to the file (or back), through a bounded host buffer
void put_device(File& out, const void* d, size_t bytes, std::vector<char>& stage) {
	for (size_t o = 0; o < bytes; o += kStage) {
		const size_t n = std::min(kStage, bytes - o);
		HIP_CHECK(hipMemcpy(stage.data(), static_cast<const char*>(d) + o, n, hipMemcpyDeviceToHost));
		out.put(stage.data(), n);
	}
}
void get_device(File& in, void* d, size_t bytes, std::vector<char>& stage) {
	for (size_t o = 0; o < bytes; o += kStage) {
		const size_t n = std::min(kStage, bytes - o);
		in.get(stage.data(), n);
		HIP_CHECK(hipMemcpy(static_cast<char*>(d) + o, stage.data(), n, hipMemcpyHostToDevice));
	}
}

// The arrays of a checkpoint, in file order, with their sizes in bytes for np particles.
std::vector<std::pair<void*, size_t>> checkpoint_arrays(pic_ctx* c, long long np) {
	const Grid& g = c->g;
	std::vector<std::pair<void*, size_t>> a;
	const size_t nn = 3 * g.nnodes * sizeof(double), nc = 3 * g.ncells * sizeof(double);
	a.push_back({c->E, nn});
	a.push_back({c->Etheta, nn});
	a.push_back({c->Enew, nn});
	a.push_back({c->B, nc});
	a.push_back({c->Bnew, nc});
	a.push_back({c->dens, (size_t)(5 * c->sp.count + 1) * g.ncells * sizeof(double)});
	if (c->explicit_scheme)
		for (int comp = 0; comp < 3; ++comp) {
			a.push_back({c->yee.E[comp], yee_count(c->yee.dE[comp]) * sizeof(double)});
			a.push_back({c->yee.B[comp], yee_count(c->yee.dB[comp]) * sizeof(double)});
			a.push_back({c->yee.J[comp], yee_count(c->yee.dE[comp]) * sizeof(double)});
		}
	if (c->fp32) {
		a.push_back({c->Q.cell, np * sizeof(int)});
		for (int d = 0; d < 3; ++d) a.push_back({c->Q.off[d], np * sizeof(float)});
		for (int d = 0; d < 3; ++d) a.push_back({c->Q.u[d], np * sizeof(float)});
		a.push_back({c->Q.id, np * sizeof(int)});
	} else {
		for (int d = 0; d < 3; ++d) a.push_back({c->P.x[d], np * sizeof(double)});
		for (int d = 0; d < 3; ++d) a.push_back({c->P.p[d], np * sizeof(double)});
		a.push_back({c->P.id, np * sizeof(int)});
	}
	return a;
}

void checkpoint_config(pic_ctx* c, File& f, bool write) {
	// what must match between the writing and the reading context
	auto same = [&](auto mine, const char* what) {
		using T = decltype(mine);
		if (write) {
			f.put<T>(mine);
		} else if (f.get<T>() != mine) {
			throw DeviceError{PIC_ERR_ARGUMENT, std::string("checkpoint written with another ") + what};
		}
	};
	for (int a = 0; a < 3; ++a) same((int32_t)c->g.n[a], "grid");
	for (int a = 0; a < 3; ++a) same(c->g.d[a], "cell size");
	same(c->cfg.dt, "time step");
	same((int32_t)c->sp.count, "species count");
	for (int s = 0; s < c->sp.count; ++s) {
		same(c->sp.mass[s], "species mass");
		same(c->sp.charge[s], "species charge");
		same(c->sp.weight[s], "species weight");
	}
	same((int32_t)c->fp32, "precision");
	same((int32_t)c->explicit_scheme, "scheme");
	same((int32_t)c->g.wall, "boundary");
}

}  // namespace

extern "C" {

int pic_checkpoint_save(pic_ctx* c, const char* path) {
	if (!c || !path) return PIC_ERR_ARGUMENT;
	try {
		HIP_CHECK(hipStreamSynchronize(c->stream));
		File f(path, "wb");
		f.put(kCkptMagic, 8);
		f.put<uint32_t>(kCkptVersion);
		checkpoint_config(c, f, true);
		const long long np = c->np();
		f.put<int64_t>(np);
		f.put<int64_t>(c->next_id);
		f.put<uint64_t>(c->inj_step);
		f.put<uint32_t>(c->inj_slabs);
		for (int s = 0; s < kMaxSpecies; ++s) f.put<double>(c->inj_len[s]);
		for (int s = 0; s <= kMaxSpecies; ++s) f.put<int32_t>(c->species_start_h[s]);
		double th[4] = {0, 0, 0, 0};
		if (c->theo_init) {
			HIP_CHECK(hipMemcpy(th, c->theo, sizeof(th), hipMemcpyDeviceToHost));
		}
		f.put<int32_t>(c->theo_init ? 1 : 0);
		for (double v : th) f.put<double>(v);
		// FP32: whether the next step takes its cell moments from the last push's density pass
		f.put<int32_t>(c->fused == 2 ? 1 : 0);
		std::vector<char> stage(kStage);
		for (const auto& a : checkpoint_arrays(c, np)) put_device(f, a.first, a.second, stage);
		f.put(kCkptEnd, 8);
		f.close();
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_checkpoint_load(pic_ctx* c, const char* path) {
	if (!c || !path) return PIC_ERR_ARGUMENT;
	c->fused = 0;
	c->u2_valid = false;
	try {
		HIP_CHECK(hipStreamSynchronize(c->stream));
		File f(path, "rb");
		char magic[8];
		f.get(magic, 8);
		if (std::memcmp(magic, kCkptMagic, 8) != 0) throw DeviceError{PIC_ERR_ARGUMENT, "not a checkpoint of this library"};
		if (f.get<uint32_t>() != kCkptVersion) throw DeviceError{PIC_ERR_ARGUMENT, "unsupported checkpoint version"};
		checkpoint_config(c, f, false);
		const long long np = f.get<int64_t>();
		if (np < 0 || np > c->capacity) throw DeviceError{PIC_ERR_ARGUMENT, "checkpoint particle count exceeds the capacity"};
		const long long next_id = f.get<int64_t>();
		const uint64_t inj_step = f.get<uint64_t>();
		const uint32_t inj_slabs = f.get<uint32_t>();
		double inj_len[kMaxSpecies];
		for (int s = 0; s < kMaxSpecies; ++s) inj_len[s] = f.get<double>();
		std::vector<int> start(kMaxSpecies + 1);
		for (int s = 0; s <= kMaxSpecies; ++s) start[s] = f.get<int32_t>();
		if (start[0] != 0 || start[c->sp.count] != np) throw DeviceError{PIC_ERR_ARGUMENT, "bad species ranges"};
		for (int s = 0; s < c->sp.count; ++s)
			if (start[s] < 0 || start[s] > start[s + 1] || start[s + 1] > np) throw DeviceError{PIC_ERR_ARGUMENT, "bad species ranges"};
		const int32_t theo_init = f.get<int32_t>();
		double th[4];
		for (double& v : th) v = f.get<double>();
		const int32_t fused_cache = f.get<int32_t>();
		// Everything is checked before the first device write, so a refused file leaves the
		// context as it was: the length the arrays need, and the end marker.
		const auto arrays = checkpoint_arrays(c, np);
		long long need = 8;
		for (const auto& a : arrays) need += (long long)a.second;
		const long long here = std::ftell(f.f);
		if (here < 0 || std::fseek(f.f, 0, SEEK_END) != 0) throw DeviceError{PIC_ERR_ARGUMENT, "cannot size the checkpoint"};
		const long long size = std::ftell(f.f);
		if (size != here + need) throw DeviceError{PIC_ERR_ARGUMENT, "checkpoint truncated or corrupt (length)"};
		char end[8];
		if (std::fseek(f.f, size - 8, SEEK_SET) != 0) throw DeviceError{PIC_ERR_ARGUMENT, "cannot read the checkpoint's end"};
		f.get(end, 8);
		if (std::memcmp(end, kCkptEnd, 8) != 0) throw DeviceError{PIC_ERR_ARGUMENT, "checkpoint truncated or corrupt"};
		if (std::fseek(f.f, here, SEEK_SET) != 0) throw DeviceError{PIC_ERR_ARGUMENT, "cannot read the checkpoint"};
		std::vector<char> stage(kStage);
		for (const auto& a : arrays) get_device(f, a.first, a.second, stage);
		HIP_CHECK(hipMemcpyAsync(c->theo, th, sizeof(th), hipMemcpyHostToDevice, c->stream));
		c->theo_init = theo_init != 0;
		set_count(c, np);
		c->next_id = next_id;
		c->inj_step = inj_step;
		c->inj_slabs = inj_slabs;
		for (int s = 0; s < kMaxSpecies; ++s) c->inj_len[s] = inj_len[s];
		c->species_start_h = start;
		hipStream_t s = c->stream;
		HIP_CHECK(hipMemcpyAsync(c->species_start, c->species_start_h.data(), (kMaxSpecies + 1) * sizeof(int),
		                         hipMemcpyHostToDevice, s));
		// the particles were written in key order: the stable sort keeps that order and
		// rebuilds the segment starts
		if (c->fp32) {
			launch_keys32(c->g, c->sp, c->Q, c->species_start, c->keys, s);
		} else {
			launch_keys(c->g, c->sp, c->P, c->species_start, c->keys, c->status, s);
			check_status(c);
		}
		sort_particles(c);
		if (c->fp32 && fused_cache) {
			// the saved run's next step takes its cell moments from the post-push pass: form them
			// again with the same kernel (from the saved particles and Enew, Bnew, which equal E, B
			// after the commit), so that the resumed run repeats its bits
			density(c, true);
			c->fused = 2;
		}
		HIP_CHECK(hipStreamSynchronize(s));
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_write_backup(pic_ctx* c, const pic_backup_info* info, const char* general_file, const char* e_file, const char* b_file,
                     const char* particles_file) {
	if (!c || !info || !general_file || !e_file || !b_file || !particles_file) return PIC_ERR_ARGUMENT;
	if (c->explicit_scheme) {
		c->error = "pic_write_backup: the implicit scheme's fields only (the explicit scheme keeps the Yee arrays)";
		return PIC_ERR_ARGUMENT;
	}
	try {
		for (int s = 0; s < c->sp.count; ++s)
			if (info->species_type[s] < 0 || info->species_type[s] > 3)
				throw DeviceError{PIC_ERR_ARGUMENT, "species_type must be 0 .. 3 (electron, proton, positron, alpha)"};
		const Grid& g = c->g;
		const long long np = c->np();
		{  // outputSimulationBackup's general file (PIC++/output.cpp:3443-3540)
			File f(general_file, "w");
			FILE* o = f.f;
			std::fprintf(o, "%d\n", info->input_type == 0 ? 0 : (info->input_type == 1 ? 1 : 2));
			for (int a = 0; a < 3; ++a) std::fprintf(o, "%d\n", g.n[a]);
			std::fprintf(o, "%d\n", (int)np);
			for (int t = 0; t < 6; ++t) std::fprintf(o, "%d\n", info->particles_per_bin[t]);
			for (int t = 0; t < 6; ++t) std::fprintf(o, "%lf\n", info->concentration[t]);
			for (double v : {info->temperature, info->plasma_period, info->scale_factor, info->time, info->max_time})
				std::fprintf(o, "%15.10g\n", v);
			std::fprintf(o, "%d\n", info->current_iteration);
			std::fprintf(o, "%d\n", info->max_iteration);
			for (double v : {info->size[0], info->size[1], info->size[2], c->theta, c->eta}) std::fprintf(o, "%15.10g\n", v);
			std::fprintf(o, "%d\n", info->debug_mode ? 1 : 0);
			std::fprintf(o, "%d\n", c->explicit_scheme ? 0 : 1);  // solverType: IMPLICIT 1
			std::fprintf(o, "%d\n", g.wall ? 0 : 1);             // boundaryConditionTypeX: PERIODIC 1
			std::fprintf(o, "%d\n", info->maxwell_matrix_size);
			std::fprintf(o, "%15.10g\n", info->ext_j);
			for (const double* v : {info->V0, info->E0, info->B0})
				for (int a = 0; a < 3; ++a) std::fprintf(o, "%15.10g\n", v[a]);
			f.close();
		}
		HIP_CHECK(hipStreamSynchronize(c->stream));
		{  // outputFields with scale 1 over all ghosted nodes (E) and cells (B)
			std::vector<double> E(3 * g.nnodes), B(3 * g.ncells);
			HIP_CHECK(hipMemcpy(E.data(), c->E, E.size() * sizeof(double), hipMemcpyDeviceToHost));
			HIP_CHECK(hipMemcpy(B.data(), c->B, B.size() * sizeof(double), hipMemcpyDeviceToHost));
			File fe(e_file, "w");
			for (int i = 0; i <= g.N[0]; ++i)
				for (int j = 0; j <= g.N[1]; ++j)
					for (int k = 0; k <= g.N[2]; ++k) {
						const long long q = node_id(g, i, j, k);
						std::fprintf(fe.f, "%30.20g %30.20g %30.20g\n", E[q], E[g.nnodes + q], E[2 * g.nnodes + q]);
					}
			fe.close();
			File fb(b_file, "w");
			for (int i = 0; i < g.N[0]; ++i)
				for (int j = 0; j < g.N[1]; ++j)
					for (int k = 0; k < g.N[2]; ++k) {
						const long long q = cell_id(g, i, j, k);
						std::fprintf(fb.f, "%30.20g %30.20g %30.20g\n", B[q], B[g.ncells + q], B[2 * g.ncells + q]);
					}
			fb.close();
		}
		{  // outputBackupParticle per particle (PIC++/output.cpp:3576-3601), ascending id
			std::vector<int> sp(np), ids(np);
			std::vector<double> x(3 * np), p(3 * np);
			if (np > 0) {
				const int rc = pic_download_particles(c, sp.data(), ids.data(), x.data(), p.data());
				if (rc != PIC_OK) return rc;
			}
			std::vector<long long> order(np);
			for (long long i = 0; i < np; ++i) order[i] = i;
			std::sort(order.begin(), order.end(), [&](long long a, long long b) { return ids[a] < ids[b]; });
			File fp(particles_file, "w");
			for (long long i : order) {
				const int s = sp[i];
				std::fprintf(fp.f, "%d ", ids[i]);
				std::fprintf(fp.f, "%15.10g ", c->sp.mass[s]);
				std::fprintf(fp.f, "%15.10g ", c->sp.charge[s]);
				std::fprintf(fp.f, "%15.10g ", c->sp.weight[s]);
				std::fprintf(fp.f, "%d ", info->species_type[s]);
				for (int a = 0; a < 3; ++a) std::fprintf(fp.f, "%15.10g ", x[a * np + i]);
				for (int a = 0; a < 3; ++a) std::fprintf(fp.f, "%15.10g ", p[a * np + i]);
				std::fprintf(fp.f, "%15.10g\n", info->particle_dx);
			}
			fp.close();
		}
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_profile_enable(pic_ctx* c, int on) {
	if (!c) return PIC_ERR_ARGUMENT;
	try {
		flush_profile(c);
		c->prof.clear();
		if (on && c->ev_pool.empty()) {
			c->ev_pool.resize(4096);
			for (auto& e : c->ev_pool) HIP_CHECK(hipEventCreate(&e));
		}
		c->profiling = on != 0;
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_profile_read(pic_ctx* c, int max_entries, char* names, double* total_ms, long long* counts, double* bytes) {
	if (!c) return PIC_ERR_ARGUMENT;
	try {
		flush_profile(c);
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	int i = 0;
	for (auto& kv : c->prof) {
		if (i >= max_entries) break;
		if (names) {
			std::strncpy(names + 32 * i, kv.first.c_str(), 31);
			names[32 * i + 31] = 0;
		}
		if (total_ms) total_ms[i] = kv.second.ms;
		if (counts) counts[i] = kv.second.count;
		if (bytes) bytes[i] = kv.second.bytes;
		++i;
	}
	return i;
}

int pic_synchronize(pic_ctx* c) {
	if (!c) return PIC_ERR_ARGUMENT;
	hipError_t e = hipStreamSynchronize(c->stream);
	if (e != hipSuccess) {
		c->error = hipGetErrorString(e);
		return PIC_ERR_DEVICE;
	}
	return PIC_OK;
}

int pic_philox(pic_ctx* c, unsigned long long c0, unsigned long long c1, unsigned long long c2, unsigned long long c3,
               unsigned long long k0, unsigned long long k1, long long n, unsigned long long* out) {
	if (!c || n < 0 || (n > 0 && !out)) return PIC_ERR_ARGUMENT;
	try {
		launch_philox(c0, c1, c2, c3, k0, k1, n, out, c->stream);
		HIP_CHECK(hipGetLastError());
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

// ------------------------------------------------------------------ learned preconditioner
static void need_net(pic_ctx* c) {
	if (!c->net) throw DeviceError{PIC_ERR_ARGUMENT, "no preconditioner network: call pic_precond_create first"};
}

int pic_precond_create(pic_ctx* c) {
	if (!c) return PIC_ERR_ARGUMENT;
	try {
		if (c->cfg.solver_precision != PIC_SOLVER_MIXED)
			throw DeviceError{PIC_ERR_ARGUMENT, "the learned preconditioner runs inside the mixed-precision solver"};
		try {
			const int dims[3] = {c->g.nu0, c->g.n[1], c->g.n[2]};  // the solver vectors' grid
			c->net.reset(new UNet(dims));
		} catch (const std::exception& e) {
			throw DeviceError{PIC_ERR_ARGUMENT, e.what()};
		}
		if (c->Z32.empty()) {
			c->Z32.resize(c->max_basis + 1);
			for (auto& v : c->Z32) v = dalloc<float>(c->nown3());
			c->Z32dev = dalloc<float*>(c->Z32.size());
			HIP_CHECK(hipMemcpy(c->Z32dev, c->Z32.data(), c->Z32.size() * sizeof(float*), hipMemcpyHostToDevice));
		}
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

long long pic_precond_param_count(void) { return UNet::param_count(); }

int pic_precond_set_params(pic_ctx* c, const float* params, double mu_scale) {
	if (!c || !params || !(mu_scale > 0)) return PIC_ERR_ARGUMENT;
	try {
		need_net(c);
		c->net->set_params(params, c->stream);
		c->mu_scale = mu_scale;
	} catch (const DeviceError& e) {
		return fail(c, e);
	} catch (const std::exception& e) {
		return fail(c, DeviceError{PIC_ERR_DEVICE, e.what()});
	}
	return PIC_OK;
}

int pic_precond_get_params(pic_ctx* c, float* params, double* mu_scale) {
	if (!c || !params) return PIC_ERR_ARGUMENT;
	try {
		need_net(c);
		c->net->get_params(params, c->stream);
		if (mu_scale) *mu_scale = c->mu_scale;
	} catch (const DeviceError& e) {
		return fail(c, e);
	} catch (const std::exception& e) {
		return fail(c, DeviceError{PIC_ERR_DEVICE, e.what()});
	}
	return PIC_OK;
}

int pic_precond_apply(pic_ctx* c, const float* r, const float* mu, float* z) {
	if (!c || !r || !mu || !z) return PIC_ERR_ARGUMENT;
	try {
		need_net(c);
		if (!(c->mu_scale > 0)) throw DeviceError{PIC_ERR_ARGUMENT, "mu_scale is not set (set_params or a logged solve sets it)"};
		c->net->forward(r, mu, (float)c->mu_scale, z, c->stream);
		HIP_CHECK(hipGetLastError());
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_precond_gradient(pic_ctx* c, const float* r, const float* mu, const float* e, double* loss, float* grad) {
	if (!c || !r || !mu || !e) return PIC_ERR_ARGUMENT;
	try {
		need_net(c);
		if (!(c->mu_scale > 0)) throw DeviceError{PIC_ERR_ARGUMENT, "mu_scale is not set"};
		c->net->zero_grad(c->stream);
		c->net->accumulate_gradient(r, mu, (float)c->mu_scale, e, c->stream);
		HIP_CHECK(hipGetLastError());
		double l = 0;
		HIP_CHECK(hipMemcpyAsync(&l, c->net->loss_ptr(), sizeof(double), hipMemcpyDeviceToHost, c->stream));
		HIP_CHECK(hipStreamSynchronize(c->stream));
		if (loss) *loss = l;
		if (grad) c->net->get_grads(grad, c->stream);
	} catch (const DeviceError& ex) {
		return fail(c, ex);
	} catch (const std::exception& ex) {
		return fail(c, DeviceError{PIC_ERR_DEVICE, ex.what()});
	}
	return PIC_OK;
}

int pic_precond_log(pic_ctx* c, int max_pairs) {
	if (!c) return PIC_ERR_ARGUMENT;
	if (max_pairs < 0) return c->log_count;
	try {
		if (c->cfg.solver_precision != PIC_SOLVER_MIXED)
			throw DeviceError{PIC_ERR_ARGUMENT, "training pairs are logged by the mixed-precision solver"};
		HIP_CHECK(hipStreamSynchronize(c->stream));
		for (void* p : {(void*)c->log_r, (void*)c->log_e}) if (p) HIP_CHECK(hipFree(p));
		for (float* v : c->log_mu) HIP_CHECK(hipFree(v));
		c->log_r = c->log_e = nullptr;
		c->log_mu.clear();
		c->log_mu_of.clear();
		c->log_count = 0;
		c->log_capacity = max_pairs;
		c->logging = max_pairs > 0;
		if (c->logging) {
			c->log_r = dalloc<float>((long long)max_pairs * c->nown3());
			c->log_e = dalloc<float>((long long)max_pairs * c->nown3());
		}
	} catch (const DeviceError& e) {
		return fail(c, e);
	}
	return PIC_OK;
}

int pic_precond_train(pic_ctx* c, int epochs, int batch, double lr, double holdout, double* train_loss, double* holdout_loss) {
	if (!c || epochs < 0 || batch < 1 || !(lr > 0) || holdout < 0 || holdout >= 1) return PIC_ERR_ARGUMENT;
	try {
		need_net(c);
		const int n = c->log_count;
		const int ntrain = n - (int)std::floor(holdout * n);
		if (ntrain < 1) throw DeviceError{PIC_ERR_ARGUMENT, "no logged training pairs"};
		const long long len = c->nown3();
		hipStream_t s = c->stream;
		const float ms = (float)c->mu_scale;
		auto sample = [&](int i) {
			c->net->accumulate_gradient(c->log_r + (long long)i * len, c->log_mu[c->log_mu_of[i]], ms, c->log_e + (long long)i * len, s);
		};
		auto read_loss = [&]() {
			double l = 0;
			HIP_CHECK(hipMemcpyAsync(&l, c->net->loss_ptr(), sizeof(double), hipMemcpyDeviceToHost, s));
			HIP_CHECK(hipStreamSynchronize(s));
			return l;
		};
		std::vector<int> order(ntrain);
		for (int i = 0; i < ntrain; ++i) order[i] = i;
		std::mt19937_64 rng(2024);
		double last = 0;
		for (int ep = 0; ep < epochs; ++ep) {
			std::shuffle(order.begin(), order.end(), rng);
			double sum = 0;
			for (int b0 = 0; b0 < ntrain; b0 += batch) {
				int nb = std::min(batch, ntrain - b0);
				c->net->zero_grad(s);
				for (int q = 0; q < nb; ++q) sample(order[b0 + q]);
				c->net->adam_step((float)lr, nb, s);
				sum += read_loss();
			}
			last = sum / ntrain;
		}
		double hl = 0;
		if (ntrain < n) {
			c->net->zero_grad(s);
			for (int i = ntrain; i < n; ++i) sample(i);
			hl = read_loss() / (n - ntrain);
			c->net->zero_grad(s);
		}
		HIP_CHECK(hipGetLastError());
		if (train_loss) *train_loss = last;
		if (holdout_loss) *holdout_loss = hl;
	} catch (const DeviceError& e) {
		return fail(c, e);
	} catch (const std::exception& e) {
		return fail(c, DeviceError{PIC_ERR_DEVICE, e.what()});
	}
	return PIC_OK;
}

int pic_precond_use(pic_ctx* c, int on) {
	if (!c) return PIC_ERR_ARGUMENT;
	if (on && (!c->net || !(c->mu_scale > 0))) {
		c->error = "the learned preconditioner needs a network and its mu_scale";
		return PIC_ERR_ARGUMENT;
	}
	c->use_net = on != 0;
	return PIC_OK;
}

}  // extern "C"

