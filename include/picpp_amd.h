/*
 * picpp_amd — C ABI of the MI355X-native implicit PIC hot path.
 *
 * One context holds one periodic subdomain (one reference MPI rank at its real
 * shape) resident in device memory. Each entry point replaces one Simulation
 * method of the reference PIC++ code and is called in the reference's order
 * (PIC++/simulation.cpp:205-390). All entry points return 0 on success and a
 * negative PIC_ERR_* code on failure (the reference exits on errors instead:
 * e.g. PIC++/particle.cpp:295-304); pic_last_error() describes the failure.
 *
 * Grid layout (as the reference's ghosted arrays, PIC++/initialization.cpp:6618-7222):
 *   N = n + 4 cells per axis, nodes 0..N, cells 0..N-1, owned indices 2..n+1,
 *   node i at x0 + i*dx with x0 = L - 2*dx, L = n*dx (PIC++/initialization.cpp:194).
 *   Field arrays are component planes: a node field is [ncomp][(N0+1)(N1+1)(N2+1)]
 *   with node id (i*(N1+1) + j)*(N2+1) + k; a cell field is [ncomp][N0*N1*N2].
 * Units: c = 1, Gaussian (the reference's normalised units).
 */
#ifndef PICPP_AMD_H
#define PICPP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pic_ctx pic_ctx;

enum {
	PIC_OK = 0,
	PIC_ERR_ARGUMENT = -1,
	PIC_ERR_DEVICE = -2,
	PIC_ERR_MEMORY = -3,
	PIC_ERR_PARTICLE = -4, /* particle outside the box or moved more than half a cell */
	PIC_ERR_SOLVER = -5
};

enum { PIC_PRECISION_FP64 = 0, PIC_PRECISION_FP32 = 1 };

/* GMRES stopping rules */
enum {
	/* the reference's rule: Givens residual estimate <= max(tol / (3 nx ny nz), 1e-15),
	   at most 48 Arnoldi steps, no restart (PIC++/specialmath.cpp:427-481) */
	PIC_STOP_REFERENCE = 0,
	/* restarted GMRES until the true residual ||b - A x|| / ||b|| <= tol */
	PIC_STOP_TRUE_RESIDUAL = 1
};

/* arithmetic of the Maxwell solve */
enum {
	PIC_SOLVER_FP64 = 0,  /* FP64 GMRES on the reference's 37-entry rows */
	/* FP32 inner GMRES (FP32 basis and operator, FP64 dot accumulation and Hessenberg least
	   squares) inside FP64 iterative refinement on the true residual; implies
	   PIC_STOP_TRUE_RESIDUAL */
	PIC_SOLVER_MIXED = 1
};

/* x boundary of the subdomain (y and z are periodic) */
enum {
	PIC_BOUNDARY_PERIODIC = 0,
	/* a perfectly conducting reflecting wall at the left end (SUPER_CONDUCTOR_LEFT,
	   PIC++/mover.cpp:248-272, fields.cpp:690-760) and a free right end held at constant
	   fields e_right, b_right (ConstantBoundaryFieldEvaluator, boundaryFieldEvaluator.cpp:30-41);
	   particles leaving through the right end are removed. The Maxwell unknowns then span
	   nodes 2 .. nx+3 along x: vectors of pic_gmres_solve hold 3 x (nx+2) ny nz values. */
	PIC_BOUNDARY_WALL = 1
};

/* pressure tensor Pi of the implicit current (J -= eta dt div Pi) */
enum {
	/* the reference's deposit: Pi[curI][j][k] with the stencil slot indices j, k in {0, 1}
	   (PIC++/momentum_collecting.cpp:565, :577), so each x plane's pressure lands in the
	   cells j, k <= 1, then the periodic folds of period n + 1 (momentum_exchange.cpp:1646-1660).
	   Needed for parity with the reference. */
	PIC_PRESSURE_REFERENCE = 0,
	/* Pi[curI][curJ][curK]: each particle's pressure in the cells of its own stencil, folded as
	   chargeDensityHat is. The reference's deposit concentrates the pressure of a whole x plane
	   in the corner columns, whose spurious div Pi grows with ny nz; with a relativistic flow
	   (the shock set-up at 64 x 64 cells across) it drives the fields there without bound. */
	PIC_PRESSURE_CELL = 1
};

/* time-stepping scheme */
enum {
	PIC_SCHEME_IMPLICIT = 0, /* the theta-implicit step (the reference's IMPLICIT solver) */
	/* The reference's BUNEMAN solver: explicit Yee fields, the Tristan mover and the zig-zag
	   current (PIC++/simulation.cpp:163-197). FP64; 3-D, or 2-D XY with nz = 1; periodic, or
	   PIC_BOUNDARY_WALL with nx >= 20 (the mover's wall at xgrid[20]) and injection;
	   smoothing_count passes smooth the current; fields are the Yee arrays (PIC_ARRAY_YEE_*),
	   stepped by pic_explicit_step. */
	PIC_SCHEME_EXPLICIT = 1
};

/* orthogonalisation inside the Arnoldi process */
enum {
	PIC_ORTHO_MGS = 0, /* modified Gram-Schmidt, one dot + axpy per basis vector (PIC++/specialmath.cpp:295-332) */
	PIC_ORTHO_CGS2 = 1 /* classical Gram-Schmidt applied twice, batched over the basis */
};

typedef struct {
	/* physical cells per axis, each >= 5; nz = 1 with PIC_SCHEME_EXPLICIT is the reference's 2-D XY
	   grid (znumberGeneral = 1, PIC++/initialization.cpp:170-174): no ghost layers along z, the Yee
	   arrays have 1 or 2 planes along z as the reference allocates them, particles keep z unwrapped */
	int nx, ny, nz;
	double dx, dy, dz;    /* cell sizes */
	double dt;            /* time step (deltaT) */
	int nspecies;         /* <= 8 (PIC++ supports 8 types, initialization.cpp:596-606) */
	double mass[8];       /* per species */
	double charge[8];
	double weight[8];     /* physical particles per macro-particle */
	long long capacity;   /* maximum number of macro-particles */
	int precision;        /* PIC_PRECISION_* of the particle mover and deposit */
	int stop_rule;        /* PIC_STOP_* */
	int ortho;            /* PIC_ORTHO_* */
	double gmres_tolerance; /* maxErrorLevel (PIC++/constants.h:31), 1e-8 */
	int gmres_max_iterations; /* maxGMRESIterations (PIC++/constants.h:55), 50 */
	int solver_precision;   /* PIC_SOLVER_* */
	int boundary_x;         /* PIC_BOUNDARY_* */
	double e_right[3];      /* E held at the free right boundary (PIC_BOUNDARY_WALL) */
	double b_right[3];      /* B held at the free right boundary */
	/* Injection (PIC_BOUNDARY_WALL only), the set-up of initializeFluxFromRight
	   (PIC++/initialization.cpp:5201-5210): with inject = 1 every pic_push, after the push,
	   adds the flow at the right end (Simulation::injectNewParticles, PIC++/mover.cpp:910-1000,
	   with injectionLength advanced by |V0_x| dt, simulation.cpp:281-283) and replaces every
	   particle that left through the right end by a new one at the boundary
	   (addToPreserveChargeGlobal, PIC++/initialization.cpp:8281-8338). New particles get ids
	   from one past the largest id uploaded. Random numbers come from Philox4x64-10 keyed by
	   seed (the reference uses rand()), so injected momenta and in-cell offsets follow the
	   reference's distributions, not its samples. */
	int inject;
	double flow[3];          /* V0, the flow velocity (|V0| < 1) */
	int ppc[8];              /* particlesPerBin of each species; 0 = not injected */
	double temperature[8];   /* kT / (m c^2) of each injected species in the flow frame
	                            (createParticle, initialization.cpp:8479-8570: the Maxwellian
	                            branch below 0.1, the isotropic Juttner branches from 0.1 on) */
	unsigned long long seed;
	int pressure;            /* PIC_PRESSURE_* */
	/* Smoothing after the moments (PIC++/simulation.cpp:213-217): smoothing_count passes over
	   chargeDensityHat, electricFlux and the dielectric tensor of the 27-point kernel with
	   weights (s/2, 1 - s, s/2) per axis, s = smoothing_parameter (PIC++/smoothing.cpp). The
	   shipped input uses 8 passes of s = 0.5; 0 passes is the default. With PIC_SCHEME_EXPLICIT the
	   passes smooth the Yee current after its summation instead (smoothBunemanEfieldGeneral,
	   PIC++/simulation.cpp:185-188, PIC++/smoothing.cpp:535-733). */
	int smoothing_count;
	double smoothing_parameter;
	int scheme;              /* PIC_SCHEME_* */
	/* bit 1: pic_step runs pic_cleanup_divergence after the push, at the reference's call position
	   (after updateDensityParameters and the newE exchange, before updateFields; the reference's
	   calls, PIC++/simulation.cpp:117-118, :348-349, run every divergenceCleanUpParameter = 1 steps
	   and are commented out in its loop); bit 2: then pic_cleanup_divergence_magnetic. Implicit
	   scheme, periodic boundary. */
	int divergence_cleanup;
	/* > 0: pic_step then runs pic_filter_fields(ctx, filter_cut_wave_number) on newE and newB */
	int filter_cut_wave_number;
	/* kBoltzman_normalized x the run's temperature (Simulation::temperature), in units of mass c^2:
	   createParticle's Juttner branches for species with temperature[s] >= 0.1 choose between
	   the tabulated ("cold") and the rejection ("hot") sampler by run_kT / mass[s] < 2, a
	   quirk of the reference that reads the run's temperature there, not the species'
	   (PIC++/initialization.cpp:8515-8527). */
	double run_kT;
} pic_config;

typedef struct {
	int iterations;           /* Arnoldi steps taken (all restarts) */
	int restarts;
	double estimated_residual; /* Givens estimate relative to ||b|| */
	double true_residual;      /* ||b - A x|| / ||b|| when computed, else -1 */
	double seconds;            /* wall time of the solve on the stream */
	int preconditioned;        /* Arnoldi steps that applied the learned preconditioner */
	int fallbacks;             /* switches from the learned preconditioner to plain GMRES */
} pic_solve_info;

/* which arrays pic_download_array / pic_device_array address */
enum {
	PIC_ARRAY_E = 0,       /* Efield, nodes x 3 */
	PIC_ARRAY_ETHETA = 1,  /* tempEfield E^{n+theta}, nodes x 3 */
	PIC_ARRAY_ENEW = 2,    /* newEfield, nodes x 3 */
	PIC_ARRAY_B = 3,       /* Bfield, cells x 3 */
	PIC_ARRAY_BNEW = 4,    /* newBfield, cells x 3 */
	PIC_ARRAY_J = 5,       /* electricFlux (J hat), nodes x 3 */
	PIC_ARRAY_MU = 6,      /* dielectricTensor, nodes x 9 (row-major 3x3) */
	PIC_ARRAY_DIVPI = 7,   /* divPressureTensor, nodes x 3 */
	PIC_ARRAY_RHOHAT = 8,  /* chargeDensityHat, cells */
	PIC_ARRAY_PI = 9,      /* pressureTensor, cells x 9 */
	PIC_ARRAY_DENSITY = 10,/* updateDensityParameters: per species [concentration, energy, bulk v x3], then rho; cells x (5S+1) */
	PIC_ARRAY_RHS = 11,    /* Maxwell right part, owned nodes x 3 (compact n0*n1*n2) */
	PIC_ARRAY_SOLUTION = 12,/* last GMRES solution, owned nodes x 3 */
	/* PIC_SCHEME_EXPLICIT: the ghosted Yee arrays bunemanE{x,y,z}, bunemanB{x,y,z},
	   bunemanJ{x,y,z} in the reference's layouts (PIC++/initialization.cpp:7146-7174): E and J
	   component c have N = n + 4 points along axis c and N + 1 along the others, B component c
	   N + 1 along axis c and N along the others; index ((i * d1) + j) * d2 + k */
	PIC_ARRAY_YEE_EX = 13,
	PIC_ARRAY_YEE_EY = 14,
	PIC_ARRAY_YEE_EZ = 15,
	PIC_ARRAY_YEE_BX = 16,
	PIC_ARRAY_YEE_BY = 17,
	PIC_ARRAY_YEE_BZ = 18,
	PIC_ARRAY_YEE_JX = 19,
	PIC_ARRAY_YEE_JY = 20,
	PIC_ARRAY_YEE_JZ = 21
};

/* Context lifetime. `stream` may be NULL (the library creates one). */
int pic_create(const pic_config* cfg, void* stream, pic_ctx** out);
void pic_destroy(pic_ctx* ctx);
const char* pic_last_error(const pic_ctx* ctx);

/* Host <-> device transfer (synchronous). Field arrays are ghosted component planes. */
int pic_upload_fields(pic_ctx* ctx, const double* E, const double* B);
/* PIC_SCHEME_EXPLICIT: the six Yee arrays Ex, Ey, Ez, Bx, By, Bz (host, the layouts of
   PIC_ARRAY_YEE_*), then the reference's pre-loop exchanges exchangeBunemanBfield and
   exchangeBunemanEfield (PIC++/simulation.cpp:83-85). Replaces pic_upload_fields and
   pic_initialize for that scheme. */
int pic_upload_yee(pic_ctx* ctx, const double* const* fields6);
/* PIC_SCHEME_EXPLICIT: one step of the BUNEMAN loop body (PIC++/simulation.cpp:163-197, then
   eraseEscapedPaticles / exchangeParticles): tristanEvaluateBhalfStep and exchangeBunemanBfield,
   moveParticleTristan for every particle with its zig-zag current (addParticleFluxZigzag, summed
   in 64-bit fixed point: the same sums in every run), the second half step,
   exchangeBunemanFlux, smoothing_count passes of smoothBunemanEfieldGeneral on the current,
   tristanEvaluateE and exchangeBunemanEfield, the one-rank periodic wrap. Replaces the implicit
   phases for that scheme. */
int pic_explicit_step(pic_ctx* ctx);
/* Particles: species[np], ids[np], x[3][np], p[3][np] (component planes, absolute
   coordinates in [L, 2L] per axis, momentum p = m gamma v). */
int pic_upload_particles(pic_ctx* ctx, long long np, const int* species, const int* ids, const double* x, const double* p);
long long pic_num_particles(const pic_ctx* ctx);
/* Particles in the context's current (cell-sorted) order. */
int pic_download_particles(pic_ctx* ctx, int* species, int* ids, double* x, double* p);
/* Returns the element count of the array (out may be NULL to query it). */
long long pic_download_array(pic_ctx* ctx, int which, double* out);
/* Device pointer and element count of an array (no copy). */
int pic_device_array(pic_ctx* ctx, int which, double** ptr, long long* count);

/* ---- Simulation-level phases, in the reference's step order ---- */

/* Pre-loop sequence (PIC++/simulation.cpp:86-123): ghost fill, particle sort,
   updateDensityParameters, updateFields, updateEnergy. */
int pic_initialize(pic_ctx* ctx);
/* evaluateParticlesRotationTensor (PIC++/mover.cpp:810-841) +
   updateElectroMagneticParameters (PIC++/momentum_collecting.cpp:365-758) */
int pic_rotation_and_moments(pic_ctx* ctx);
/* right part of createInternalEquation (PIC++/fields.cpp:2635-2715) */
int pic_field_rhs(pic_ctx* ctx);
/* evaluateElectricField (PIC++/fields.cpp:152-376): GMRES on the matrix-free operator
   of createInternalEquation{X,Y,Z} (fields.cpp:886-2633), generalizedMinimalResidualMethod
   (PIC++/specialmath.cpp:367-671); writes E^{n+theta} and newE. Calls pic_field_rhs. */
int pic_maxwell_solve(pic_ctx* ctx, pic_solve_info* info);
/* evaluateMagneticField (PIC++/fields.cpp:2717-2773) */
int pic_update_b(pic_ctx* ctx);
/* moveParticles (PIC++/mover.cpp:21-55, :194-491) + eraseEscapedPaticles/exchangeParticles
   periodic wrap (mover.cpp:167-192, :1002-1210) + updateParticleCorrelationMaps +
   updateDensityParameters (PIC++/momentum_collecting.cpp:760-1020) */
int pic_push(pic_ctx* ctx);
/* updateFields (PIC++/fields.cpp:422-439) */
int pic_commit_fields(pic_ctx* ctx);
/* updateEnergy (PIC++/simulation.cpp:1506-1640): out[7] = particle, E, B, total energy,
   momentum x, y, z */
int pic_energy(pic_ctx* ctx, double* out7);
/* one full implicit step: all of the above in the reference's order */
int pic_step(pic_ctx* ctx, pic_solve_info* info);

/* The free-function solver boundary: generalizedMinimalResidualMethod
   (PIC++/specialmath.h:14-17) on the context's current operator (dielectric tensor of
   the last pic_rotation_and_moments). rhs and x are device pointers to owned-node
   vectors [3][n0*n1*n2]; rhs is preserved. */
int pic_gmres_solve(pic_ctx* ctx, const double* rhs, double* x, pic_solve_info* info);
/* y = A v on device vectors [3][n0*n1*n2] (multiplySpecialMatrixVector, specialmath.cpp:202-252) */
int pic_apply_operator(pic_ctx* ctx, const double* v, double* y);
/* The same operator in the factored form that PIC_SOLVER_MIXED applies (multiplySpecialMatrixVector,
   PIC++/specialmath.cpp:202-252, on the rows of createInternalEquation{X,Y,Z},
   PIC++/fields.cpp:1396-2632, and the wall rows of fields.cpp:690-781): w = mu v on the nodes
   around the unknowns, then (A v)_c = v_c - w_c - C lap v_c + C d2_c w_c + the mixed second
   differences of w. precision PIC_PRECISION_FP64: v, y are device double vectors
   [3][n0*n1*n2]; PIC_PRECISION_FP32: device float vectors, mu rounded to float (needs a context
   created with PIC_SOLVER_MIXED). Exposes the mixed solver's operator so that it can be checked
   against the 37-entry rows. */
int pic_apply_operator_factored(pic_ctx* ctx, int precision, const void* v, void* y);

/* Simulation::cleanupDivergence(newEfield, chargeDensity) (PIC++/divergencecleanup.cpp:20-392, its
   GMRES branch): solves L phi = div newE - 4 pi rho on the owned cells, L the 27-point operator of
   createDivergenceCleanupInternalEquation (:620-813), with the reference's GMRES (stop rule of
   PIC_STOP_REFERENCE with one component, at most maxDivergenceCleanupIterations = 100,
   PIC++/constants.h:56), then newE -= grad phi (evaluateDivergenceCleaningField,
   updateFieldByCleaning, :531-618) and the newE ghost exchange. rho is the density array's charge
   plane (the last pic_push's updateDensityParameters). After it div newE = 4 pi rho to the solver's
   tolerance. Implicit scheme, periodic boundary (PIC_ERR_ARGUMENT otherwise). info (may be NULL)
   receives the GMRES iterations and the Givens estimate. */
int pic_cleanup_divergence(pic_ctx* ctx, pic_solve_info* info);
/* Simulation::cleanupDivergenceMagnetic (PIC++/divergencecleanup.cpp:939-1306): the same 27-point
   equation for a node potential with div newB at the nodes (evaluateDivBgeneral, fields.cpp:3019-3030)
   as right part, then newB -= grad phi at the cells and the newB ghost exchange, so that div newB = 0
   to the solver's tolerance. The reference solves it with BiCGStab to maxCleanupErrorLevel = 1e-8
   (PIC++/specialmath.cpp:1714), which reaches rounding level on these systems; this build uses the
   GMRES of pic_cleanup_divergence run to its 1e-15 floor (at most 98 Arnoldi steps). B from pic_update_b is solenoidal to rounding (the
   cell curl of node E), so on the implicit step this corrects rounding only. */
int pic_cleanup_divergence_magnetic(pic_ctx* ctx, pic_solve_info* info);
/* Simulation::filterFields(cutWaveNumber), periodic branch (PIC++/fields_filtering.cpp:21-295):
   every component of newE and newB over the owned points is transformed with the separable
   discrete Fourier transform of fourierTranslation (PIC++/fourier.cpp:13-192), the modes with
   |k| > 2 pi / (cut_wave_number dx) are multiplied by -1 (the reference's filterFieldGeneral; k from
   the unfolded mode indices 0 .. n-1, and the y component's k_y from the z mode index, as there),
   transformed back and written to the owned points; then the newE and newB ghost exchanges.
   Implicit scheme, periodic boundary, cut_wave_number > 0. */
int pic_filter_fields(pic_ctx* ctx, int cut_wave_number);

/* Philox4x64-10, the generator of the injection: out (device, 4 n words) receives the
   outputs for counters (c0, c1 + i, c2, c3), i < n, under key (k0, k1). For tests of the
   generator against an independent implementation. */
int pic_philox(pic_ctx* ctx, unsigned long long c0, unsigned long long c1, unsigned long long c2, unsigned long long c3,
               unsigned long long k0, unsigned long long k1, long long n, unsigned long long* out);

/* updateEnergy (the seven values of pic_energy, PIC++/simulation.cpp:1506-1640) of caller-owned
   device state, with no context and no allocation of any kind. No counterpart in the reference:
   it is for callers that cannot allocate device memory (code under HIP graph capture, tight
   serving loops, harnesses that forbid hipMalloc). cfg supplies the grid (nx, ny, nz, dx, dy, dz,
   boundary_x) and the species (nspecies, mass, weight); its other fields are ignored. E: ghosted
   node field [3][(N0+1)(N1+1)(N2+1)] and B: ghosted cell field [3][N0 N1 N2] (device, the layouts
   of PIC_ARRAY_E / PIC_ARRAY_B); p: device momenta [3][np] (component planes, m gamma v) of the
   particles grouped by species, counts[s] (host array) of species s; out7 (device, 7 doubles)
   receives particle, E, B, total energy, momentum x, y, z. Work is queued on `stream` (NULL: the
   null stream) and the call returns without waiting. The partial sums use a 37 KB device array
   that the library holds on every device it runs on until the process ends (9 x 512 block sums
   and 9 totals, the same whatever the sizes, so no size is refused for lack of room); every call
   overwrites it before reading it. Calls that overlap on one device race on that array: order
   them on one stream. PIC_ERR_ARGUMENT for a bad argument, PIC_ERR_DEVICE when the device array
   or a launch fails; pic_last_error(NULL) describes this thread's last failure. */
int pic_energy_noalloc(const pic_config* cfg, const double* E, const double* B, const double* p, const long long* counts,
                       double* out7, void* stream);

/* Blocks until the context's stream is idle. */
int pic_synchronize(pic_ctx* ctx);

/* Particles already in device memory, grouped by species (counts[s] particles of species
   s, in species order): x[3][np], p[3][np] component planes, ids[np] or NULL. Copied and
   sorted on the context's stream. Replaces the reference's particle construction
   (PIC++/initialization.cpp:8145-8234) for inputs generated on the device. */
int pic_upload_particles_device(pic_ctx* ctx, const long long* counts, const double* x, const double* p, const int* ids);

/* Per-kernel timing with HIP events on the context's stream, the counterpart of the
   reference's `timing` prints (PIC++/simulation.cpp:156-160, fields.cpp:189-205).
   pic_profile_enable resets the statistics. pic_profile_read fills up to max_entries
   rows: names (32 chars each), total milliseconds, launch counts and the algorithmic
   bytes of those launches; it returns the number of rows. */
int pic_profile_enable(pic_ctx* ctx, int on);
int pic_profile_read(pic_ctx* ctx, int max_entries, char* names, double* total_ms, long long* counts, double* bytes);

/* ---- Checkpoints and the reference's backup files ---- */

/* Lossless checkpoint of the state at a step boundary, in this library's own binary format
   (little endian; magic "PICAMDCK", version 1): the configuration it was written with, the
   node and cell fields (E, B, their step copies, the density planes), the Yee arrays of the
   explicit scheme, the particles in the context's own precision with their ids, and the
   injection state. pic_checkpoint_load needs a context created with the same grid, time step,
   species, precision, scheme and boundary (PIC_ERR_ARGUMENT otherwise) and restores the state
   bit for bit; the following steps repeat an uninterrupted run's bits (FP64 and the explicit
   scheme; FP32 forms the first step's cell moments afresh, equal to FP32 rounding). The
   reference's backup (outputSimulationBackup, PIC++/output.cpp:3432-3602) is text with 10
   significant digits for particles, and its reader (readBackup, PIC++/input.cpp:341-622) does
   not read what the writer writes, so it cannot resume a run; this format replaces it. */
int pic_checkpoint_save(pic_ctx* ctx, const char* path);
int pic_checkpoint_load(pic_ctx* ctx, const char* path);

/* What the reference's backup header holds beyond the context's own state (Simulation members
   written by outputSimulationBackup, PIC++/output.cpp:3443-3540). */
typedef struct pic_backup_info {
	int input_type;                 /* 0 CGS, 1 Theoretical, 2 other */
	int particles_per_bin[6];       /* types[0..5].particlesPerBin */
	double concentration[6];        /* types[0..5].concentration */
	double temperature, plasma_period, scale_factor;
	double time, max_time;
	int current_iteration, max_iteration;
	double size[3];                 /* xsizeGeneral, ysizeGeneral, zsizeGeneral */
	int debug_mode;
	int maxwell_matrix_size;
	double ext_j;
	double V0[3], E0[3], B0[3];
	int species_type[8];            /* per species: 0 electron, 1 proton, 2 positron, 3 alpha */
	double particle_dx;             /* the particles' dx member */
} pic_backup_info;

/* The reference's backup files for the current state of an implicit-scheme context
   (outputSimulationBackup with one rank, PIC++/output.cpp:3432-3602): the general file, E on
   all ghosted nodes and B on all ghosted cells (outputFields, PIC++/output.cpp:1375-1520,
   "%30.20g" triples, i then j then k), and one line per particle in ascending id order
   (outputBackupParticle). Text identical to the reference's for the same state. */
int pic_write_backup(pic_ctx* ctx, const pic_backup_info* info, const char* general_file, const char* e_file,
                     const char* b_file, const char* particles_file);

/* ---- general.dat rows ----
   The numbers of one row of the reference's general.dat (outputGeneral, PIC++/output.cpp:3215-3250)
   for the context's current state, as the step loop keeps them (updateEnergy,
   PIC++/simulation.cpp:1506-1640; updateParameters :1707-1880; the theoretical energy and momentum
   of updateTheoreticalEnergy :1399-1504 and of the mover's reflections, the injection and the
   replacement of escapes, PIC++/mover.cpp:256, 994, PIC++/initialization.cpp:8319). The context
   starts the theoretical totals from the energy and momentum of the state that the first step,
   pic_initialize or pic_general after an upload finds, and updates them in every pic_push and
   pic_explicit_step. out[PIC_GENERAL_VALUES], in the row's order without the iteration and time
   columns, with the row's unit factors (energies x (scale_factor / plasma_period)^2, momenta x
   scale_factor / plasma_period, field maxima and rms over plasma_period sqrt(scale_factor)):
     0 particle energy, 1 E energy, 2 B energy, 3 total energy, 4-6 total momentum,
     7 theoretical energy, 8-10 theoretical momentum, 11 largest |E|, 12 largest |B|,
     13 particles numbered so far (particlesNumber), 14-16 rms of E_x, E_y, E_z,
     17-19 E_x, E_y, E_z energies, 20-22 B_x, B_y, B_z energies, 23-25 field momentum,
     26-28 particle momentum.
   From info: plasma_period, scale_factor, B0 (updateParameters starts its |B| maximum from
   B[0][0][0] - B0), time and current_iteration (pic_write_general only). */
#define PIC_GENERAL_VALUES 29
int pic_general(pic_ctx* ctx, const pic_backup_info* info, double* out);
/* Appends the row to general_file in outputGeneral's format (two printf calls, "%15.10g"
   fields); shockWaveX (-1) and the derivative points (0) keep their initial values, their
   updates being switched off in the reference's loop (PIC++/simulation.cpp:108-111). */
int pic_write_general(pic_ctx* ctx, const pic_backup_info* info, const char* general_file);

/* ---- Learned preconditioner ----
   No counterpart in the reference, whose GMRES is unpreconditioned
   (PIC++/specialmath.cpp:367-671). A 3-D convolution U-Net (3 levels, 16/32/64 channels,
   3x3x3 kernels, periodic, bfloat16 weights and activations with float32 accumulation on
   MFMA) maps a solver vector r and the dielectric tensor mu to an approximate correction
   z ~ A^-1 r. With pic_precond_use on, the mixed-precision solve (PIC_SOLVER_MIXED) runs
   flexible GMRES with z_j = M(v_j) and falls back to plain GMRES when the network does not
   reduce the residual; the FP64 true residual <= tolerance judges it either way.
   Vectors are device float32: r, z, e [3][nx ny nz], mu [9][nx ny nz] (owned nodes, the
   layout of pic_gmres_solve). Every physical dimension must be a multiple of 4. */
/* Creates the network for the context's grid (He-initialised; the untrained net is z = r).
   Requires solver_precision = PIC_SOLVER_MIXED. */
int pic_precond_create(pic_ctx* ctx);
/* Number of float32 parameters: per convolution W[27][cout][cin] then b[cout], in the order
   enc1 16->16, enc2 16->32, enc3 32->64, dec3 64->32, dec2 32->16, dec1 16->16; then the
   1x1x1 head W[3][16], b[3]. */
long long pic_precond_param_count(void);
/* Host parameter transfer; mu_scale divides mu at the network input. */
int pic_precond_set_params(pic_ctx* ctx, const float* params, double mu_scale);
int pic_precond_get_params(pic_ctx* ctx, float* params, double* mu_scale);
/* z = M(r, mu) on device vectors. */
int pic_precond_apply(pic_ctx* ctx, const float* r, const float* mu, float* z);
/* Loss ||M(r, mu) - e||^2 / ||e||^2 and its gradient with respect to the parameters
   (host array of pic_precond_param_count floats). */
int pic_precond_gradient(pic_ctx* ctx, const float* r, const float* mu, const float* e, double* loss, float* grad);
/* Starts logging training pairs (r, e) with A e = r from the following mixed-precision
   solves: every Arnoldi step gives (A v_j, v_j) and every solve gives (b, x). max_pairs = 0
   stops logging and frees the log. Returns the number of pairs held with max_pairs < 0. */
int pic_precond_log(pic_ctx* ctx, int max_pairs);
/* Adam on the logged pairs: `epochs` passes over the first (1 - holdout) of them in a seeded
   shuffled order, `batch` pairs per step. Reports the mean training loss of the last epoch
   and the mean loss on the held-out pairs. */
int pic_precond_train(pic_ctx* ctx, int epochs, int batch, double lr, double holdout, double* train_loss, double* holdout_loss);
/* Turns the learned preconditioner on (1) or off (0) for the following solves. */
int pic_precond_use(pic_ctx* ctx, int on);

#ifdef __cplusplus
}
#endif

#endif /* PICPP_AMD_H */
