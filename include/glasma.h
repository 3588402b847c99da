/*
 * glasma.h — C ABI of the MI355X-native 3+1D glasma engine (libglasma.so).
 *
 * The reference (pragyasingh1729/SoerenPragya3DGlasma) has no plugin or FFI layer: its "API" is a set
 * of C++ namespace functions acting on global field objects.  Every entry point below replaces one of
 * those functions; the citation gives the reference file:line (paths relative to
 * Pragya_workingdirectory_worksonm1/Workingdirectory/SourceCode/).
 *
 * Conventions
 *   - Plain C types only: pointers, sizes, doubles.  No torch types.
 *   - Return value: 0 on success, negative glasma_status on failure; glasma_last_error() gives text.
 *     The library never calls exit() (the reference exits with status 0 on every error).
 *   - Host field arrays use the reference array-of-structs layout INCLUDING the two z ghost layers:
 *       U, UOld : complex double (re,im interleaved) at 9*(mu + 3*Index3D) + i + 3*j
 *       E       : double at a + 8*(mu + 3*Index3D)
 *       J       : double at a + 8*Index3D
 *     with Index3D = x + Nx*(y + Ny*(z+1)), z in [-1, Nz]   (LATTICE/GaugeLinks.cpp:27-36).
 *     Inputs: ghost layers are ignored (the device keeps no ghost layers; open z boundaries read the
 *     edge layer, which is what SynchronizeGhostCells guarantees before every read in the reference).
 *     Outputs: ghost layers are filled as SynchronizeGhostCells would (LATTICE/GaugeLinks.cpp:59-73).
 *   - A context owns all device memory; no caller pointer is retained past a call.  Calls block until
 *     the device work is done.  A context is not thread-safe; separate contexts are independent.
 *   - Only SU(3) (nc = 3) runs on the device.
 */
#ifndef GLASMA_H_
#define GLASMA_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct glasma_ctx glasma_ctx;

typedef enum glasma_status {
    GLASMA_OK = 0,
    GLASMA_ERR_ARG = -1,     /* invalid argument / unsupported configuration */
    GLASMA_ERR_HIP = -2,     /* HIP runtime error (text in glasma_last_error) */
    GLASMA_ERR_NOMEM = -3,   /* device or host allocation failed */
    GLASMA_ERR_STATE = -4,   /* call out of order (e.g. observable before fields are set) */
    GLASMA_ERR_NUMERIC = -5, /* non-finite value or non-unitary update detected */
    GLASMA_ERR_IO = -6       /* a file could not be read or written, or does not match the context */
} glasma_status;

/* Lattice and dynamics parameters: Lattice::N/a (LATTICE/3DGrid.cpp:14-34), Dynamics::dTStep
 * (DYNAMICS/SUNcDynamics.cpp:30), BOUNDARY_CONDITIONS_FLAG (LATTICE/3DGrid.cpp:10). */
typedef struct glasma_params {
    int nc;       /* gauge group SU(nc); must be 3 */
    int n[3];     /* Nx, Ny, Nz */
    double a[3];  /* lattice spacings ax, ay, az */
    double dt;    /* time step */
    int bc_z;     /* 0 = open (edge-copy ghost layers, the reference default), 1 = periodic */
} glasma_params;

/* MV-model parameters: LATTICE/Parameters.cpp:4-50, COLLIDINGOBJECTS/CollidingObjects.cpp:30-102 */
typedef struct glasma_mv_params {
    double lambda;      /* UV regulator Lambda (Parameters::Lambda) */
    double m_eff;       /* IR regulator m (Parameters::MEff) */
    double sigma_l;     /* longitudinal width in index units (Parameters::SigmaLOverAzL) */
    double sigma_r;     /* (Parameters::SigmaLOverAzR) */
    double nucleus_l;   /* multiplier of the left charge (Parameters::LeftNucleus) */
    double nucleus_r;   /* (Parameters::RightNucleus) */
    int center_l;       /* z centre of the left nucleus (Simulation.cpp:344: 0.85*Nz/2-1) */
    int center_r;       /* (Simulation.cpp:348: 1.15*Nz/2-1) */
} glasma_mv_params;

/* ---- context ---------------------------------------------------------------------------------- */
/* Replaces Lattice::Init + Simulation::Init (LATTICE/SUNcVariables.cpp:71-90, Simulation.cpp:115-130):
 * allocates every field on the current HIP device; U = UOld = 1, E = J = 0, t = 0. */
int glasma_create(const glasma_params* p, glasma_ctx** out);
int glasma_destroy(glasma_ctx* ctx);
const char* glasma_last_error(const glasma_ctx* ctx);   /* ctx may be NULL (creation errors) */
const char* glasma_version(void);
/* device bytes held by the context */
int glasma_memory_bytes(const glasma_ctx* ctx, size_t* bytes);

/* ---- fields (LATTICE/GaugeLinks.cpp, ElectricFields.cpp, ChargeDensity.cpp) ------------------ */
/* Any pointer may be NULL (that field is left unchanged / not returned). */
int glasma_set_fields(glasma_ctx* ctx, const double* U, const double* E, const double* Jp, const double* Jm,
                      const double* JpStat, const double* JmStat);
int glasma_get_fields(glasma_ctx* ctx, double* U, double* E, double* Jp, double* Jm, double* JpStat,
                      double* JmStat);
/* UOld = links at t - dt/2 of the last step (Dynamics::SUNc::SaveOldLink, DYNAMICS/SUNcDynamics.cpp:108-128) */
int glasma_set_uold(glasma_ctx* ctx, const double* UOld);
int glasma_get_uold(glasma_ctx* ctx, double* UOld);
/* Dynamics::t / tSteps (DYNAMICS/SUNcDynamics.cpp:22-50) */
int glasma_get_time(const glasma_ctx* ctx, double* t, long* tsteps);
int glasma_set_time(glasma_ctx* ctx, double t, long tsteps);

/* ---- leapfrog (DYNAMICS/SUNcDynamics.cpp:353-378) ---------------------------------------------- */
/* Dynamics::SUNc::Update applied nsteps times: E update (with the J_z source), J-stat advection,
 * half-step links, UOld save, J advection, half-step links.  One fused HIP kernel per step.
 * UOld is written on the last step only (earlier values are overwritten by the reference anyway). */
int glasma_update(glasma_ctx* ctx, int nsteps);
/* Same, asynchronous on the context stream (no host synchronisation); for timing loops. */
int glasma_update_async(glasma_ctx* ctx, int nsteps);
int glasma_synchronize(glasma_ctx* ctx);
/* Average duration of the step kernel over the last glasma_update_async/glasma_update call, measured
 * with HIP events on the context stream (ms per launch). */
int glasma_last_step_kernel_ms(const glasma_ctx* ctx, double* ms);

/* ---- observables ------------------------------------------------------------------------------ */
/* Observables::Bulk::Update (OBSERVABLES/BulkObservables.cpp:172-187): out9 = ESqr[3], BSqr[3],
 * EDotJ[3] (raw sums; T00..SOURCE follow from BulkObservables.cpp:22-62).  GLASMA_ERR_NUMERIC (with the
 * sums filled in as computed) when a sum is not finite. */
int glasma_bulk(glasma_ctx* ctx, double* out9);
/* No counterpart in the reference: glasma_bulk for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same sums, bit for bit, from the same
 * kernels; the block partials go to a fixed 8 MiB device array of the library (one per device, held until the
 * process ends) instead of the context's reduction buffer, and no allocation is made on the success path.  The
 * partials of one call take Nz * (ceil(Nx*Ny/256) + 1) * 9 + 9 doubles; a lattice needing more than 2^20 (e.g.
 * 256 x 256 x 512) is refused with GLASMA_ERR_ARG.  Calls that overlap on one device (from several threads or
 * contexts) share the array and race; the caller serialises them. */
int glasma_bulk_noalloc(glasma_ctx* ctx, double* out9);
/* GaussLaw::UpdateViolation (OBSERVABLES/GaussLawViolation.cpp:16-124): per_z[Nz] (may be NULL),
 * global max |G^a| and RMS. */
int glasma_gauss(glasma_ctx* ctx, double* per_z, double* gmax, double* grms);
/* Unitarity::UpdateViolation (OBSERVABLES/UnitarityViolation.cpp:13-51) */
int glasma_unitarity(glasma_ctx* ctx, double* umax, double* urms);
/* EnergyMomentumTensor::CreateMap (OBSERVABLES/EnergyMomentumTensor.cpp:92-170): out[10*z + k] =
 * T00 TXX TYY TZZ T0X T0Y T0Z TXY TYZ TZX transverse averages per z. */
int glasma_emt(glasma_ctx* ctx, double* out);
/* IO::SaveFields (IO/SaveConfiguration.cpp:112-158): out[4*z + k], z = 0..Nz-1, the file's columns
 * 0.5 (B_x^2 + B_y^2), 0.5 B_z^2, 0.5 (E_x^2 + E_y^2), 0.5 E_z^2, transverse averages from the
 * elementary plaquettes of the current links and the local E. */
int glasma_field_profile(glasma_ctx* ctx, double* out);
/* IO::SaveTransverselyIntegratedCurrents (IO/SaveConfiguration.cpp:160-204): out[4*r + k] for the
 * (Nz + 2) layers z = r - 1 = -1..Nz (the ghost layers as SynchronizeGhostCells fills them), the
 * file's columns 0.5 <Jm^2>, 0.5 <JmStat^2>, 0.5 <Jp^2>, 0.5 <JpStat^2> (transverse averages). */
int glasma_current_profile(glasma_ctx* ctx, double* out);
/* EnergyMomentumTensor::CreateMapBfCollision (OBSERVABLES/EnergyMomentumTensor.cpp:10-90): out[2*(x +
 * Nx*y) + side], the energy density summed over z / Nz for z <= (Nz - 2) / 2 (side 0, "T00Left") and
 * above (side 1, "T00Right"). */
int glasma_transverse_energy_map(glasma_ctx* ctx, double* out);

/* ---- MV initial conditions ---------------------------------------------------------------------- */
/* Charges of both nuclei from one MT19937 stream seeded as gsl_rng_set (seed 0 -> 4357), Gaussian
 * by gsl_ran_gaussian's polar method; left first, then right, each in y,x,a order scaled by qs
 * (MISC/RNG/GSLRandomNumberGenerator.cpp:12-33, COLLIDINGOBJECTS/CollidingObjects.cpp:335-342,
 * Simulation.cpp:344-348).  xi arrays hold Nx*Ny*8 doubles, index a + 8*(x + Nx*y).  Host only. */
int glasma_draw_charges(unsigned long seed, int nx, int ny, double qs_l, double qs_r, double* xi_l,
                        double* xi_r);
/* Simulation::Run's initial state (Simulation.cpp:327-378): SetWilsonLines for both nuclei
 * (CollidingObjects.cpp:318-371) then SingleSheet (INITIALCONDITIONS/SUNCCURRENTS/RevSingleNucleus.cpp:7-210)
 * for driver 0 = collision (Simulation.cpp:359-361), 1 = left-moving (LeftNucleusSimulation.cpp:357),
 * 2 = right-moving (RightNucleusSimulation.cpp:357).  Resets fields and time first. */
int glasma_mv_initial_state(glasma_ctx* ctx, int driver, const glasma_mv_params* mv, const double* xi_l,
                            const double* xi_r);
/* CollidingObjects::SetWilsonLines for one nucleus ('L' or 'R'): V (complex, AoS with ghosts,
 * 9*Index3D + i + 3j) and the smoothed charge rho (a + 8*Index3D, with ghosts).  Either may be NULL. */
int glasma_set_wilson_lines(glasma_ctx* ctx, char dir, const glasma_mv_params* mv, const double* xi,
                            double* V_out, double* rho_out);

/* ---- physical-charge (TMD / GBW) nuclei (the TMDNUCLEUS sources, SimulationmitPC.cpp:315-370) ----- */
/* Parameters of LATTICE/Parameters.cpp:4-48 that the physical-charge path reads. */
typedef struct glasma_tmd_params {
    double qs0a;        /* Parameters::Qs0A (lattice units) */
    double m_eff;       /* Parameters::MEff, IR regulator of the transverse filter (lattice units) */
    double com_energy;  /* Parameters::CoMEnergy (fm^-1; 500/0.197 in the reference) */
    double a_fm[3];     /* Parameters::aXInFm, aYInFm, aZInFm */
    int atomic_number;  /* Parameters::AtomicNumber: 16, 54, 197 or 208 (NuclearData.cpp) */
    int center_l;       /* z index of the left nucleus' centre (TMDbased.cpp:43: 785) */
    int center_r;       /* (TMDbased.cpp:48: 1261) */
} glasma_tmd_params;
/* The random draws of SimulationmitPC::Run (SimulationmitPC.cpp:326-344) from one MT19937 stream seeded as
 * glasma_draw_charges: the left nucleus' noise (Nz*Ny*Nx*8 Gauss() in z, y, x, a order, index
 * a + 8*(x + Nx*(y + Ny*z))), its nucleons (Woods-Saxon rejection sampling with gsl_rng_uniform,
 * GBWSampling.cpp:21-87; 3*A doubles in fm, centre of mass at the origin), then the right noise and
 * nucleons.  Host only; GLASMA_ERR_ARG for an atomic number without Woods-Saxon parameters. */
int glasma_tmd_draw(unsigned long seed, int nx, int ny, int nz, int atomic_number, double* noise_l,
                    double* noise_r, double* nucleons_l, double* nucleons_r);
/* SimulationmitPC::Run's initial state: per nucleus ChargeFluctuation (CorrelationFunction.cpp:24-95: the noise
 * / sqrt(a^3) filtered by sqrt(GammaK) in 3D momentum space), the charge a^3 sqrt(aZInFm T / a_z) CF with T the
 * nucleus profile (TMDbased.cpp:27-57, GBWSampling.cpp:110-157), the IR filter p^2/(p^2 + m^2), Poisson solve
 * and Wilson lines (TMDbased.cpp:256-303), then SingleSheet as glasma_mv_initial_state; driver as there.  The
 * reference's own build of this path does not compile; this follows its source.  Resets fields and time. */
int glasma_tmd_initial_state(glasma_ctx* ctx, int driver, const glasma_tmd_params* p, const double* noise_l,
                             const double* noise_r, const double* nucleons_l, const double* nucleons_r);

/* ---- Coulomb gauge fixing (GAUGETRANSFORMATION/CoulombGaugeFixing.cpp:32-88 with
 *      ALGORITHMS/FourierAcceleration.cpp:15-148) ------------------------------------------------- */
/* Fourier-accelerated steepest descent on the gauge transformation G (kept in the context across
 * calls, identity at creation), until max |Delta| <= tol or maxit iterations (reference: 1e-15, 1000,
 * alpha = 2).  The gauge-fixed copies U^G, E^G are kept in the context; the dynamical fields are not
 * modified.  stats7 = iterations, max dev, avg dev, max step, avg step, min functional, mean functional. */
int glasma_coulomb_gauge(glasma_ctx* ctx, double tol, int maxit, double alpha, double* stats7);
/* The same with the update the reference selects by #define (CoulombGaugeFixing.cpp:4,55-67):
 * GLASMA_GAUGE_FOURIER (param = alpha; identical to glasma_coulomb_gauge), GLASMA_GAUGE_JACOBI
 * (ALGORITHMS/Jacobi.cpp:13-74, param = step size, 0.3 in the reference) or GLASMA_GAUGE_LOS_ALAMOS
 * (ALGORITHMS/LosAlamos.cpp:13-87, even then odd checkerboard sites, MaxTraceProjection of w; param = the
 * over-relaxation probability, which must be 0, the reference's value). */
/* GLASMA_GAUGE_FOURIER_CONSISTENT (param = alpha, 2): no counterpart in the reference, whose Fourier iteration cannot
 * converge for Nz > 1 (periodic z momenta despite open z, the gauge-fixed links' z = -1 ghost left at the identity,
 * the zero mode passed through unscaled).  w takes only the links that exist (open z: none below layer 0 and none
 * from layer Nz - 1 to its ghost copy; periodic z: the wrapped links, with the current G at both ends), the Poisson
 * step inverts the Laplacian of that graph (open z: Neumann, a cosine transform in z, which equals the periodic
 * transform of the lattice doubled by its mirror image in z; periodic z: periodic), and the zero mode is dropped; it
 * reaches tolerances of 1e-10 and below.
 * Every algorithm returns 0 at the iteration cap as the reference does; when the cap is reached above a tolerance
 * tol > 0, glasma_last_error says so. */
enum { GLASMA_GAUGE_FOURIER = 0, GLASMA_GAUGE_JACOBI = 1, GLASMA_GAUGE_LOS_ALAMOS = 2, GLASMA_GAUGE_FOURIER_CONSISTENT = 3 };
int glasma_coulomb_gauge_algorithm(glasma_ctx* ctx, int algorithm, double tol, int maxit, double param,
                                   double* stats7);
/* Gauss-law restoration (GAUSSLAWFIXING/GaussRestore.cpp:134-182; transverse_only != 0:
 * GaussRestore2D.cpp, E_x and E_y only): E relaxed in place toward the Gauss law with the static charges,
 * three steps per iteration (Gamma = 7/24, 7/48, 7/72), until max |G| <= tol or maxit iterations
 * (reference: 1e-14, 30000).  The violation's ghost layers are edge copies and the fields' ghosts are
 * synchronised after each step, as in GaussRestore2D.cpp, for both variants (GaussRestore.cpp indexes past its
 * allocation at the top layer).  The step sizes are stable for unit spacings.  stats3 = iterations, max |G|,
 * RMS G after a final measurement. */
int glasma_gauss_restore(glasma_ctx* ctx, int transverse_only, double tol, int maxit, double* stats3);
/* GaugeTransformation::Reset / SetIdentity (CoulombGaugeFixing.cpp:26-30) */
int glasma_gauge_reset(glasma_ctx* ctx);
/* Gauge-fixed copies (GaugeFixedVariables::U/E, GAUGETRANSFORMATION/GaugeTransformation.cpp:8-21) and G
 * (complex, 9*Index3D + i + 3j with ghosts).  Any pointer may be NULL. */
int glasma_get_gauge_fixed(glasma_ctx* ctx, double* UG, double* EG, double* G);
int glasma_set_gauge_transformation(glasma_ctx* ctx, const double* G);

/* ---- device FFT (replaces FFT/FFT1D.cpp, FFT2D.cpp, FFT3D.cpp; FFTW conventions: forward e^{-2 pi i},
 *      backward e^{+2 pi i}, unnormalised) -------------------------------------------------------- */
/* Multi-dimensional complex transform of a host array: dims[0] fastest, `batch` interleaved
 * components fastest of all (XIndex = a + batch*(x + Nx*(y + Ny*z)), FFT/FFT3D.cpp:41-47);
 * sign -1 forward, +1 backward.  Runs on the current device; for tests and tools. */
int glasma_fft_host(double* data, int nd, const int* dims, long batch, int sign);

/* ---- bitwise state checksum (no counterpart in the reference; for run-to-run identity checks) ------- */
/* out4 = checksums of the current links U, UOld, E and the four currents, each the wrapping 64-bit sum over
 * every physical site and component of SplitMix64(bits ^ key(component, site)).  Two contexts on the same
 * lattice hold bit-identical fields exactly when (up to a 2^-64 collision) their checksums agree; the value
 * depends only on the field values, not on launch order or tile order. */
int glasma_checksum(glasma_ctx* ctx, uint64_t* out4);

/* ---- synthetic input (no counterpart in the reference; for benchmarks and tests) ---------------- */
/* Fills the current links with exp(i link_scale alpha^a t^a), alpha_a uniform in [-1,1), and E, J+,
 * J-, J+stat, J-stat with uniform values in [-scale, scale), from a counter-based hash of (seed, site). */
int glasma_fill_random(glasma_ctx* ctx, unsigned long seed, double link_scale, double e_scale, double j_scale);

/* ---- ensembles of events (no counterpart in the reference, which runs one event per process) ------- */
/* An ensemble of independent events, each an ordinary context (its own lattice, spacings and time step)
 * reached through glasma_batch_context for fields, initial conditions and observables.
 * glasma_batch_update advances every event by nsteps leapfrog steps (Dynamics::SUNc::Update,
 * DYNAMICS/SUNcDynamics.cpp:353-378) with one kernel launch per step over all events (two when the
 * ensemble mixes lattices whose Nx is and is not a multiple or a divisor of 64), the launches captured
 * once per step count into a HIP graph.  Each event's fields are bit-identical to advancing it
 * alone.  glasma_batch_update_serial advances the events one after another with per-event launches
 * (the baseline).  glasma_batch_last_ms: device time of the last batched call, wall time of the last
 * serial call.  Each event needs fewer than 2^32 sites. */
typedef struct glasma_batch glasma_batch;
int glasma_batch_create(const glasma_params* p, int nevents, glasma_batch** out);   /* p[nevents] */
int glasma_batch_destroy(glasma_batch* b);
int glasma_batch_size(const glasma_batch* b, int* nevents);
int glasma_batch_context(glasma_batch* b, int e, glasma_ctx** ctx);
const char* glasma_batch_last_error(const glasma_batch* b);
int glasma_batch_update(glasma_batch* b, int nsteps);
int glasma_batch_update_serial(glasma_batch* b, int nsteps);
int glasma_batch_last_ms(const glasma_batch* b, double* ms);

/* ---- configuration files -------------------------------------------------------------------------------- */
/* The reference's text configurations: IO::SaveConfiguration (U and E files) and SaveConfigurationJ (J file),
 * IO/SaveConfiguration.cpp:6-93, and IO::LoadConfiguration, IO/LoadConfiguration.cpp:3-202, with
 * Dynamics::LoadedTime (DYNAMICS/SUNcDynamics.cpp:42-45).  Any file name may be NULL (that file is skipped /
 * that field kept).  Records "x y z mu" + 9 complex entries (U), "x y z mu" + 8 components (E), "x y z a
 * J+stat J+ J-stat J-" (J), precision 15, z over the ghost layers -1..Nz (U, E of periodic lattices:
 * 0..Nz-1).  Loading checks record order and count and returns GLASMA_ERR_IO with a message instead of
 * exiting.  Fixed reference bugs: a J record sets its own colour only (LoadCurrents' `for(aIn==0;aIn<3;...)`
 * wrote colours a..2), and the step count of a loaded time is round(t/dt) (LoadedTime computed it from the
 * old time).  The text keeps 15 digits; for a bit-identical restart use glasma_save_state / glasma_load_state
 * (no counterpart in the reference): an exact binary file of U, UOld, E, the four currents and the clock,
 * checked against the context's lattice, spacings, time step and boundary on loading. */
int glasma_save_configuration(glasma_ctx* ctx, const char* ufile, const char* efile, const char* jfile);
int glasma_load_configuration(glasma_ctx* ctx, const char* ufile, const char* efile, const char* jfile, double t);
int glasma_save_state(glasma_ctx* ctx, const char* file);
int glasma_load_state(glasma_ctx* ctx, const char* file);

/* ---- tile-order autotuner (no counterpart in the reference, which has no GPU kernels) -------------------- */
/* The order in which the step's blocks visit the lattice sets how much neighbour-link traffic the XCD-local
 * L2 caches absorb.  glasma_tune_step times the step on the context's current fields under every order
 * that fits the lattice (XCD column bands of several shapes walked in z, plane bands, z slabs, linear), each with one
 * warm-up and `reps` timed launches into scratch buffers -- the fields, UOld and the clock are unchanged, and
 * no order changes any result -- and keeps the fastest for the context's later steps.  Optional outputs
 * (arrays of at least GLASMA_TUNE_MAX entries): the candidates' codes (mode*10000 + band width in tiles *100
 * + band height in rows; mode 0 linear, 1 plane bands, 2 column bands, 3 z slabs), their ms per step, the index of
 * the chosen one.  Needs 880 B per site of scratch
 * device memory while it runs.  glasma_tile_order reports the code in use and whether it was tuned.
 * glasma_batch_tune tunes each distinct lattice of an ensemble once (events of the same lattice share the
 * order) and rebuilds the ensemble's work list. */
#define GLASMA_TUNE_MAX 16
int glasma_tune_step(glasma_ctx* ctx, int reps, int* ncand, int* codes, double* ms, int* best);
int glasma_tile_order(const glasma_ctx* ctx, int* code, int* tuned);
/* The bandwidth ceiling of the step's compulsory traffic on this device: a copy kernel with the step's tiled
 * layout and grid (one 64-site tile per block) reads the current U, E and currents once and writes them once to
 * scratch (880 + 880 bytes per site), `reps` timed launches after one warm-up; ms = average per launch.  The
 * fields are unchanged.  Needs 880 B per site of scratch device memory while it runs. */
int glasma_copy_ceiling(glasma_ctx* ctx, int reps, double* ms);
int glasma_batch_tune(glasma_batch* b, int reps);

/* ---- reverse-mode derivatives (no counterpart in the reference, which has no gradient code) ------------ */
/* Slab energy of the current fields: 1/2 sum over sites with |z - zc| < half of sum_mu c_mu (E_mu^a E_mu^a
 * + 4 Re tr(1 - P_mu)), c_mu = (a_mu / a^3)^2, P_0 = P_yz, P_1 = P_zx, P_2 = P_xy rooted at the site: the
 * local form of Observables::Bulk's T00 (OBSERVABLES/BulkObservables.cpp:22-97). */
int glasma_slab_energy(glasma_ctx* ctx, int zc, int half, double* energy);
/* One leapfrog step's vector-Jacobian product: the arrays hold cotangents of the fields after one step
 * (reference AoS layout with ghosts, ghosts ignored) and are overwritten with the cotangents of the fields
 * before it.  The context's fields are the linearisation point and are left unchanged (UOld is not). */
int glasma_step_vjp(glasma_ctx* ctx, double* Ub, double* Eb, double* Jpb, double* Jmb, double* JpSb, double* JmSb);
/* The left-moving MV sheet over all z (InitialConditions::SingleSheet with 'L',
 * INITIALCONDITIONS/SUNCCURRENTS/RevSingleNucleus.cpp:7-210) built from a given charge density rho (the
 * density CollidingObjects::SetChargeDensities produces, COLLIDINGOBJECTS/CollidingObjects.cpp:34-102;
 * a + 8*Index3D with ghosts, ghosts ignored; smoothing with mv->lambda, mv->m_eff), then nsteps leapfrog
 * steps; energy = glasma_slab_energy of the result. */
int glasma_slab_energy_from_charge(glasma_ctx* ctx, const glasma_mv_params* mv, const double* rho, int nsteps, int zc,
                                   int half, double* energy);
/* The same energy and its gradient with respect to rho (grad_rho: same layout, ghosts zero), by
 * reverse-mode differentiation of the steps (forward states checkpointed every ckpt steps, each segment but the
 * last recomputed once) and of the initial conditions.  The context holds the initial state afterwards.  Both calls keep
 * their device work arrays in the context (counted by glasma_memory_bytes, freed by glasma_destroy), so that
 * repeated calls of one size allocate nothing.  When memory allows, a segment's recompute runs on a second stream
 * of the context (created on first use, default priority) beside the backward sweep; the call returns after both. */
int glasma_slab_energy_gradient(glasma_ctx* ctx, const glasma_mv_params* mv, const double* rho, int nsteps, int ckpt,
                                int zc, int half, double* energy, double* grad_rho);

/* ---- per-site SU(3) algebra on the device, batched (tests and tools) ------------------------- */
/* op: 0 exp(i c alpha t) [in: n*8 alpha, out: n*18], 1 log [in: n*18 U, out: n*8],
 *     2 V alpha V^dag [in: n*18 V, in2: n*8, out: n*8], 3 V^dag alpha V [same],
 *     4 Re tr(i t^a V) [in: n*18, out: n*8]. Complex matrices interleaved, column-major. */
int glasma_su3_batch(int op, long n, double c, const double* in, const double* in2, double* out);

#ifdef __cplusplus
}
#endif

#endif /* GLASMA_H_ */
