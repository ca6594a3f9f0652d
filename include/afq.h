/* afq.h -- C ABI of the gfx950 AFQMC walker-step library (safire_amd/libafq_hip.so).
 *
 * Drop-in boundary for SAFIRE's Cholesky HamiltonianOperations and Slater-determinant
 * operations.  Every entry point names the reference interface it replaces (paths relative to
 * the SAFIRE source tree).
 *
 * Conventions
 *  - All array arguments are DEVICE pointers (HBM of the context's GPU), row-major, complex
 *    values interleaved (re, im) as double[2] (complex128) or float[2] (complex64).
 *  - Every call is ordered on the context's stream (the one given to afq_ctx_create, or a
 *    private non-blocking stream) and returns before the work completes; use
 *    afq_ctx_synchronize() before reading results on the host.  Pass the caller's own stream
 *    to order the calls with its other work.
 *  - Return value: AFQ_OK (0) or a negative afq_status; afq_last_error() gives the text of the
 *    last failure on the calling thread.  Shape errors play the role of the reference's
 *    utils::check -> APP_ABORT (check.hpp:36-46); nothing is written when a check fails.
 *  - Ownership: the caller owns every array it passes.  Hamiltonian tensors are copied into the
 *    afq_ham object at creation and stay resident; scratch space belongs to the context.
 *  - One context per host thread.  Calls on one context are not thread-safe.
 *
 * Layouts (N orbitals, M Cholesky vectors, W walkers, nup/ndown electrons, nel = nup (CLOSED)
 * or nup + ndown (COLLINEAR)):
 *   Likn  double [N][N][M]        DenseFactorized/L (RealDenseHamiltonian.cpp:121-127)
 *   h1    double [N][N]           Hamiltonian/hcore
 *   PsiT  complex [nel][N]        trial, conjugate-transposed (NOMSD.hpp:292-344); alpha rows first
 *   Phi   complex [W][N][n]       walker Slater matrices of one spin block (WalkerSetBase.h:320-328);
 *                                 each walker's [N][n] block is contiguous and consecutive walkers
 *                                 are phi_stride complex elements apart (0: N * n, contiguous), so
 *                                 SlaterMatrices(s) of SAFIRE's walker buffer, a view with the
 *                                 walker row as its stride (WalkerSetBase.h:256-276), is passed
 *                                 in place
 *   G     complex [W][nel][N]     compact mixed density matrix (density_matrix.hpp:576)
 *   X, v  complex [W][M]          fields / force bias
 *   V     complex [W][N][N]       vHS (complex128 in FP64, complex64 in mixed precision)
 *   E     complex [W][3]          (E1, EXX, EJ)
 */
#ifndef AFQ_H
#define AFQ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  AFQ_OK = 0,
  AFQ_ERR_ARG = -1,
  AFQ_ERR_HIP = -2,
  AFQ_ERR_SHAPE = -3,
  AFQ_ERR_ALLOC = -4,
  AFQ_ERR_STATE = -5
} afq_status;

enum { AFQ_CLOSED = 0, AFQ_COLLINEAR = 1 };          /* WALKER_TYPES (config.h) */
enum { AFQ_FP64 = 0, AFQ_MIXED = 1 };                /* compute precision */

typedef struct afq_ctx afq_ctx;
typedef struct afq_ham afq_ham;
typedef struct afq_prop afq_prop;
typedef struct afq_phmsd afq_phmsd;
typedef struct afq_nomsd afq_nomsd;

/* ---- context ------------------------------------------------------------------------------ */

/* Version of this ABI (major * 100 + minor).  2.00: afq_prop_params gained local_energy and
 * afq_prop_te_anchor its follow_walkers argument.  3.00: walker strides (afq_walkers.phi_stride
 * and prop_stride, the phi_stride argument of the Slater-determinant and PHMSD entry points) and
 * afq_tune_force.  3.01: afq_mixed_density_matrix_noalloc.  3.02: multi-determinant NOMSD
 * (afq_nomsd_*, afq_prop_create_nomsd), afq_vbias_full and afq_ham_create_spin. */
int afq_abi_version(void);
/* Text of the last error on this thread ("" when none). */
const char* afq_last_error(void);
/* Creates a context on `device`; `stream` is a hipStream_t or NULL for a private stream.
 * Replaces arch::init's device selection (arch.cpp:60-69). */
int afq_ctx_create(int device, void* stream, afq_ctx** out);
int afq_ctx_destroy(afq_ctx* ctx);
int afq_ctx_synchronize(afq_ctx* ctx);
/* Tile shapes picked by the autotuner so far, one "site:precision:shape -> candidate" line per
 * tuned product (written into buf, truncated to len - 1 characters). */
int afq_tuning_report(char* buf, int len);
/* Bytes of scratch held by the context. */
int afq_ctx_scratch_bytes(afq_ctx* ctx, uint64_t* bytes);
/* Pins the autotuner's choice for a product site (no counterpart in the reference; for tests
 * that run every tile candidate): site "name" or "name:f32" / "name:f64" (the first field or two
 * of an afq_tuning_report key), candidate >= 0 pins, < 0 unpins; site NULL unpins every site.
 * Process-wide.  A pinned index beyond the site's candidates fails the product's call with
 * AFQ_ERR_ARG.  The environment variable AFQ_TUNE_FORCE="site:i,..." sets pins at load time. */
int afq_tune_force(const char* site, int candidate);

/* ---- Cholesky hamiltonian operations (Real3IndexFactorization) ------------------------------ */

/* RealDenseHamiltonian::getHamiltonianOperations (RealDenseHamiltonian.cpp:52-186) +
 * Real3IndexFactorization ctor (Real3IndexFactorization.hpp:43-97): copies Likn, h1, PsiT and
 * builds Lnak, haj and vexx on the device.  precision = AFQ_MIXED also keeps FP32 copies of the
 * Cholesky tensors for the FP32 contractions. */
int afq_ham_create(afq_ctx* ctx, int walker_type, int N, int nup, int ndown, int M,
                   const double* Likn, const double* h1, const double* PsiT, double E0,
                   int precision, afq_ham** out);
/* As afq_ham_create, with spin-dependent integrals (RealDenseHamiltonian.cpp:86-120, hcore
 * [nspin_in_H1][NMO][NMO] and DenseFactorized/L [nspin_in_H2][NMO][NMO][ncv]): Likn
 * [nspin_L][N][N][M], h1 [nspin_h1][N][N], each 1 or 2 (alpha block, then beta; 2 only for
 * COLLINEAR walkers).  Spin s uses block s % nspin: its half-rotated tensors, vexx, vHS block
 * and one-body propagator.  With nspin_L = 2, afq_vhs writes V [2][W][N][N]; with either 2,
 * afq_one_body_propagator writes P1 [2][N][N]. */
int afq_ham_create_spin(afq_ctx* ctx, int walker_type, int N, int nup, int ndown, int M,
                        const double* Likn, int nspin_L, const double* h1, int nspin_h1,
                        const double* PsiT, double E0, int precision, afq_ham** out);
int afq_ham_destroy(afq_ham* ham);
/* Device views of the half-rotated tensors: Lnak complex [M][nup][N] (then [M][ndown][N] for
 * COLLINEAR), haj complex [nel][N], vexx double [N][N]. number_of_cholesky_vectors() is M. */
int afq_ham_tensors(afq_ham* ham, const double** Lnak, const double** haj, const double** vexx,
                    int* real_trial);
/* Real3IndexFactorization::vbias (Real3IndexFactorization.hpp:328-428):
 * v[w][m] = sqrt(dt) * c * sum_{a,k} G[w][a][k] Lnak[m][a][k], c = 2 CLOSED / 1 COLLINEAR.
 * precision AFQ_MIXED contracts in FP32 (FP32 MFMA, FP32 accumulation) and writes FP64 v. */
int afq_vbias(afq_ham* ham, const double* G, int W, double dt, int precision, double* v);
/* Real3IndexFactorization::vHS (Real3IndexFactorization.hpp:282-326):
 * V[s][w][i][k] = sqrt(dt) * sum_m L_s[i][k][m] X[w][m] over the Cholesky tensor's spin blocks
 * s (one unless created with nspin_L = 2).  FP64: X, V complex128.
 * AFQ_MIXED: X, V complex64 (FP32 MFMA). */
int afq_vhs(afq_ham* ham, const void* X, int W, double dt, int precision, void* V);
/* Real3IndexFactorization::energy (Real3IndexFactorization.hpp:138-174, energy_impl :476-590):
 * E[w] = (E0 + c sum G haj, -c/2 sum_{m,a,b} T_ab T_ba, c^2/2 sum_m (sum_a T_aa)^2).
 * AFQ_MIXED runs the exchange contraction in FP32 with FP64 accumulation. */
int afq_energy(afq_ham* ham, const double* G, int W, int addH1, int addEJ, int addEXX,
               int precision, double* E);
/* The per-spin Real3IndexFactorization::energy(spin, E, G, idet, EJn, addH1, addEJ, addEXX)
 * (Real3IndexFactorization.hpp:176-210): G [W][n_s][N] is the compact density matrix of spin
 * block `spin` (0 alpha, 1 beta: COLLINEAR only).  E [W][3] = (E0 (alpha only) + c sum G haj_s,
 * -c/2 sum_{m,a,b} T_ab T_ba, 1/2 sum_m EJn^2) with c = 2 CLOSED / 1 COLLINEAR, and EJn [W][M] =
 * sum_a T^m_aa, the Coulomb vectors a caller combines across spins (NULL when addEJ = 0).
 * precision as afq_energy. */
int afq_energy_spin(afq_ham* ham, int spin, const double* G, int W, int addH1, int addEJ,
                    int addEXX, int precision, double* E, double* EJn);
/* NOMSD::vMF (NOMSD.cpp:437-464) + zero_imag (AFQMCBasePropagator.cpp:105-108): vMF [M]. */
int afq_vmf(afq_ham* ham, double dt, double* vMF);
/* generateP1 + getOneBodyPropagatorMatrix (AFQMCBasePropagator.cpp:43-214,
 * Real3IndexFactorization.hpp:99-124): P1 = exp(-1/2 hermitize(dt (h1 + vexx) + vHS(vMF))),
 * entries with |x| <= 1e-8 set to zero (to_csr, csr_utils.hpp:43-66).  P1 complex [N][N], or
 * [2][N][N] (alpha, beta) for spin-dependent integrals. */
int afq_one_body_propagator(afq_ham* ham, const double* vMF, double dt, double* P1);

/* ---- Slater-determinant operations (det_ops) --------------------------------------------- */

/* det_ops::MixedDensityMatrix (density_matrix.hpp:535-601), herm = true.  ovlp [W] is
 * accumulated: ovlp += log det(PsiT Phi), imaginary part wrapped to [-pi, pi]
 * (determinants_impl.hpp:33-75).  compact != 0: G [W][n][N]; else full G [W][N][N]. */
int afq_mixed_density_matrix(afq_ctx* ctx, const double* PsiT, int n, int N, const double* Phi,
                             int64_t phi_stride, int W, double* G, double* ovlp, int compact);
/* afq_mixed_density_matrix without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results; the workspace, W n n
 * complex (compact) or W n (n + N) complex (full G), comes from a fixed per-GPU array of
 * AFQ_NOALLOC_SCRATCH_BYTES that the library holds until the process ends; larger calls fail
 * with AFQ_ERR_SHAPE.  No tile is timed (the tuner's cached choice or its first candidate runs:
 * the same bits).  Every call on one GPU shares that array, so calls that overlap on one GPU
 * (two streams or two host threads) race. */
#define AFQ_NOALLOC_SCRATCH_BYTES (32u << 20)
int afq_mixed_density_matrix_noalloc(afq_ctx* ctx, const double* PsiT, int n, int N,
                                     const double* Phi, int64_t phi_stride, int W, double* G,
                                     double* ovlp, int compact);
/* det_ops::Log_Overlap (density_matrix.hpp:431-447): ovlp += log det(PsiT Phi). */
int afq_log_overlap(afq_ctx* ctx, const double* PsiT, int n, int N, const double* Phi,
                    int64_t phi_stride, int W, double* ovlp);
/* det_ops::detail::inverse_logdet (density_matrix.hpp:82-111): Ainv = A^{-1} (zero when
 * singular), ovlp += log det A.  A, Ainv complex [nb][n][n]; A == Ainv is allowed. */
int afq_inverse_logdet(afq_ctx* ctx, const double* A, int nb, int n, double* ovlp, double* Ainv);
/* det_ops::detail::propagate_impl + apply_expM (propagate.hpp:41-70, :241-298):
 * Phi <- P1 * (sum_{k<=order} (i op(V))^k / k!) * P1 * Phi, TA = 'N' or 'H' (op = V or V^H, -i).
 * P1 complex128 [N][N]; V complex128 (AFQ_FP64) or complex64 (AFQ_MIXED) [W][N][N]. */
int afq_propagate(afq_ctx* ctx, const double* P1, const void* V, double* Phi, int64_t phi_stride,
                  int W, int N, int n, int order, char TA, int precision);
/* det_ops::orthogonalize (orthogonalize.hpp:131-169): Phi = Q R with R upper triangular,
 * positive real diagonal; Phi <- Q; ldet += log det R. */
int afq_orthogonalize(afq_ctx* ctx, double* Phi, int64_t phi_stride, int W, int N, int n,
                      double* ldet);

/* ---- fields and weights ------------------------------------------------------------------- */

/* Uniform field draws u[w][m] in (0,1) from Philox4x32-10 keyed by (seed, stream, step, w, m).
 * Replaces RandomGenerator::sampleUniformFields (Random.hpp:37-85). */
int afq_field_uniforms(afq_ctx* ctx, uint64_t seed, uint32_t stream, uint32_t step, int W, int M,
                       double* u);
/* assemble_X + construct_X (AFQMCBasePropagator.cpp:686-733, construct_X_kernel.hpp:34-86) and
 * the mean-field factor (:335-336). X holds vbias on entry and the fields on exit; HW, mf [W]. */
int afq_construct_x(afq_ctx* ctx, const double* u, const double* vMF, int W, int M, double bound,
                    int project_vbias, int free_projection, double* X, double* HW, double* mf);

typedef struct {
  double dt, E0, lower_cut, upper_cut, weight_floor, weight_frac, target_pop;
  int free_projection, use_cp;
} afq_update_params;

/* walker_update + bound_walker_weights (WalkerSetUpdate.hpp:70-218), hybrid branch.  Eshift is
 * a device scalar.  weight, phase, eloc (PSEUDO_ELOC_), ovlp (OVLP), theta [W] updated. */
int afq_walker_update(afq_ctx* ctx, int W, const afq_update_params* p, const double* Eshift,
                      const double* new_ovlp, const double* mf, const double* HW, double* weight,
                      double* phase, double* eloc, double* ovlp, double* theta);
/* averagePseudoEnergy (average_energy.hpp:54-63) -> *out (device); when relax > 0 also
 * *Eshift += relax * (*out - *Eshift) (run_afqmc.cpp:91). */
int afq_average_pseudo_energy(afq_ctx* ctx, int W, const double* weight, const double* eloc,
                              double dt, double relax, double* out, double* Eshift);

/* ---- propagator (AFQMCBasePropagator, single-determinant trial) -------------------------- */

typedef struct {
  double dt;                 /* timestep */
  int order;                 /* taylor_n (default 6) */
  double vbias_bound;        /* default 50 */
  double lower_cut, upper_cut;          /* lower/upper_cutoff_scale (1, 10) */
  double weight_floor, weight_frac;     /* 100, 0.1 */
  double target_pop;         /* global target population (<= 0: W) */
  int free_projection, use_cp, project_force_bias;
  uint64_t seed;             /* field RNG key */
  uint32_t stream;           /* field RNG stream index */
  int local_energy;          /* 0: hybrid step (default); 1: the reference's hybrid = false --
                                the local energy of the propagated walkers (Energy instead of
                                Log_Overlap, AFQMCBasePropagator.cpp:368-373) drives the weights
                                with the midpoint rule (WalkerSetUpdate.hpp:120-146) */
} afq_prop_params;

typedef struct {
  double* Phi;     /* complex [W][N][nup] (alpha / CLOSED) */
  double* Phi_b;   /* complex [W][N][ndown] (COLLINEAR beta) or NULL */
  double* weight;  /* complex [W] WEIGHT */
  double* phase;   /* complex [W] PHASE */
  double* eloc;    /* complex [W] PSEUDO_ELOC_ */
  double* ovlp;    /* complex [W] OVLP (log overlap) */
  double* theta;   /* complex [W] THETA */
  int W;
  /* Complex elements between consecutive walkers.  phi_stride: of Phi and of Phi_b (0: each
   * block contiguous, N * nup and N * ndown); prop_stride: of the five property arrays (0 or 1:
   * contiguous).  SAFIRE's walker buffer holds one row per walker, [Phi_a | Phi_b | WEIGHT,
   * PHASE, PSEUDO_ELOC_, E1_, EXX_, EJ_, OVLP, ..., THETA] (WalkerSetBase.cpp:75-121): there
   * Phi = row 0 + data_displ[SM], Phi_b = Phi + N * nup, each property = row 0 + its
   * data_displ, and both strides are the row length walker_buffer.strides()[0]
   * (SlaterMatrices, getProperty: WalkerSetBase.h:256-276, 463-481).  The step reads and writes
   * the Slater blocks in place and stages the properties through 80 bytes per walker of scratch. */
  int64_t phi_stride;
  int64_t prop_stride;
  /* E1_, EXX_, EJ_ columns (complex [W], prop_stride apart) or NULL.  When given, the
   * local-energy step (afq_prop_params.local_energy = 1, the reference's hybrid = false) stores
   * the components of each propagated walker's local energy there (WalkerSetUpdate.hpp:180-184)
   * and population control copies them with their walkers; the hybrid step leaves them alone. */
  double* E1;
  double* EXX;
  double* EJ;
} afq_walkers;

/* AFQMCBasePropagator constructor (AFQMCBasePropagator.h:45-171): builds vMF and P1. */
int afq_prop_create(afq_ham* ham, int W, const afq_prop_params* p, afq_prop** out);
/* The same propagator over a PHMSD trial (Wavefunction<PHMSD>, COLLINEAR): vMF from the trial
 * density, G for vbias, overlaps and energies from the multi-determinant expansion. */
int afq_prop_create_phmsd(afq_phmsd* trial, int W, const afq_prop_params* p, afq_prop** out);
/* As afq_prop_create, with a multi-determinant NOMSD trial (ndet > 1): force bias and energy
 * through the determinant sum, vMF from the trial's full density matrix, and the back-propagated
 * references over every determinant (NOMSD.cpp:39-65, 159-217, 437-521). */
int afq_prop_create_nomsd(afq_nomsd* trial, int W, const afq_prop_params* p, afq_prop** out);
int afq_prop_destroy(afq_prop* prop);
/* Device views: vMF complex [M], P1 complex [N][N]. */
int afq_prop_views(afq_prop* prop, const double** vMF, const double** P1);
/* AFQMCBasePropagator::Propagate (AFQMCBasePropagator.cpp:289-392) for every walker.
 * u: injected uniforms [W][M] (device) or NULL to draw them from Philox with the step counter
 * held on the device (advanced by one per call).  Eshift: device scalar. */
int afq_prop_step(afq_prop* prop, const afq_walkers* w, const double* u, const double* Eshift);
/* AFQMCBasePropagator::Orthogonalize (AFQMCBasePropagator.cpp:735-741). */
int afq_prop_orthogonalize(afq_prop* prop, const afq_walkers* w);
/* Sets the device step counter used by the in-kernel RNG. */
int afq_prop_set_step(afq_prop* prop, uint32_t step);
/* NOMSD::Energy (NOMSD.cpp:159-217) or, for a PHMSD propagator, PHMSD::Energy (energy_alg1,
 * PHMSD.cpp:376-544, FP64): E [W][3], ovlp [W] (log overlap). */
int afq_prop_energy(afq_prop* prop, const afq_walkers* w, int precision, double* E, double* ovlp);
/* One step + optional orthogonalization + optional comb population control + growth
 * estimator (-> *energy_out, device) + Eshift relaxation (relax > 0), captured in a HIP graph
 * on first reuse and replayed afterwards (the loop body of run_afqmc.cpp:69-100).  Uses the
 * in-kernel RNG for the fields and the comb. */
int afq_prop_step_graph(afq_prop* prop, const afq_walkers* w, double* Eshift, int ortho,
                        int popctl, double relax, double* energy_out);
/* population_control (population_control.hpp:216-257) on this GPU's walkers with the comb's
 * uniform r in (0, 1): offspring by |w|, copies into the slots of dropped walkers in walker
 * order, then every weight set to unit magnitude (phase kept).  W <= 16384. */
int afq_prop_population_control(afq_prop* prop, const afq_walkers* w, double r);
/* Bound-trigger statistics since the propagator's creation or the last reset (reset != 0
 * zeroes them after reading), counts[9] in the order of AFQMCBasePropagator::printBoundStatistics
 * (AFQMCBasePropagator.h:195-235): force-bias clamp |vbias| > vbias_bound -- operations (walker,
 * field pairs), upper hits, lower hits (always 0); local-energy clamp outside
 * Eshift - lower_cut sqrt(2/dt), Eshift + upper_cut sqrt(2/dt) (bound_local_energy,
 * WalkerSetUpdate.hpp:46-58) -- operations (finite, constrained walkers), upper, lower; weight
 * clamp |w| > max(weight_floor, weight_frac target_pop) (bound_walker_weights, :199-218) --
 * operations (walkers), upper, lower (0).  Counted on the device by every step, captured steps
 * included (the reference counts the force-bias clamp in its host path only).  Synchronizes. */
int afq_prop_bound_stats(afq_prop* prop, int64_t* counts, int reset);
/* *collapsed = 1 when the last population control found the total weight <= 1e-6 (the
 * population is then left unchanged; the reference aborts).  Synchronizes the stream. */
int afq_prop_status(afq_prop* prop, int* collapsed);

/* ---- back-propagation -------------------------------------------------------------------------
 * Replaces the walker set's back-propagation buffers (WalkerSetBase::resize_bp, storeFields,
 * getWeightFactors, setSlaterMatrixN) and AFQMCBasePropagator::BackPropagate
 * (AFQMCBasePropagator.cpp:404-560) as BackPropEstimator (BackPropEstimator.hpp:150-260) uses
 * them. */
/* From now on every step stores its fields in a ring of nback slots and its path-restoration
 * factor (w / |w|) / scale (WalkerSetUpdate.hpp:136-146) in a ring of 3 nback slots per walker,
 * both indexed by the device step counter (steps must draw their own fields, u = NULL);
 * population control moves the rings and the anchor with their walker. */
int afq_prop_enable_backprop(afq_prop* prop, int nback);
/* Anchor = the walkers' current Slater matrices (BackPropEstimator::setAnchor). */
int afq_prop_bp_anchor(afq_prop* prop, const afq_walkers* w);
/* Back-propagates the trial through the last nbp stored fields (newest first, the adjoint
 * propagator, QR every `ortho` steps and at the end) and writes G [W][nspin][N][N] (complex,
 * device; nspin = 1 CLOSED, 2 COLLINEAR): the full mixed density matrix between the
 * back-propagated trial and each walker's anchor, and weights [W] (complex, device): WEIGHT
 * times the path-restoration factors of the last restore * nbp steps (restore = 0 none,
 * 1 path_restoration, 2 extra_path_restoration; BackPropEstimator.hpp:212-228).  Single-
 * determinant trials only (AFQ_ERR_ARG for PHMSD: use afq_prop_bp_measure). */
int afq_prop_backpropagate(afq_prop* prop, const afq_walkers* w, int nbp, int ortho, int restore,
                           double* G, double* weights);
/* Device views of the rings: fields [W][nback][M] complex, factors [W][3 nback] complex. */
int afq_prop_bp_views(afq_prop* prop, const double** fields, const double** factors, int* nback);

/* BackPropEstimator::backPropagate (BackPropEstimator.hpp:186-241) with the OneRDM and
 * DiagonalTwoRDM observables: every reference determinant of the trial (one for NOMSD, the
 * expansion for PHMSD, PHMSD::getReferences) is back-propagated through the last nbp stored
 * fields as in afq_prop_backpropagate, weighted per walker by conj(ci) <L_d|anchor>
 * e^{conj(log det R_d)} normalized over the references (constructBPMeasurementInputs), and the
 * observables are measured with the path-restored weights (restore as above).  onerdm
 * [nspin][N][N], diag2rdm [nspin + 1][N][N] as afq_rdm_measure writes them; either may be NULL.
 * Device pointers; the call synchronizes the stream once. */
int afq_prop_bp_measure(afq_prop* prop, const afq_walkers* w, int nbp, int ortho, int restore,
                        double* onerdm, double* diag2rdm);
/* MixedEstimator::measure (MixedEstimator.hpp:122-140) with the OneRDM and DiagonalTwoRDM
 * observables: the walkers' full mixed density matrices with every reference determinant,
 * weighted by conj(ci) <D_d|Phi> normalized over the references (constructMixedMeasurementInputs,
 * :40-106), and WEIGHT.  Outputs as for afq_prop_bp_measure. */
int afq_prop_mixed_measure(afq_prop* prop, const afq_walkers* w, double* onerdm, double* diag2rdm);
/* Number of reference determinants of the trial (Wavefunction::total_number_of_references). */
int afq_prop_number_of_references(afq_prop* prop, int* nref);

/* Kernel timing of the step (no counterpart in the reference; for benchmarks): on != 0 makes
 * every later step record two HIP events around its vHS launch, also inside captured step
 * graphs (switching re-captures them); afq_prop_vhs_ms returns the last step's vHS duration in
 * milliseconds, waiting for the end event. */
int afq_prop_vhs_timing(afq_prop* prop, int on);
int afq_prop_vhs_ms(afq_prop* prop, float* ms);

/* ---- back-propagation by operators evolved forward in time (TimeEvolvedBPEstimator.hpp) ------
 * The operator states X, Y, M [W][nspin][N][N] (complex) of the propagator's W walker slots.
 * All three calls need afq_prop_enable_backprop (the field and factor rings) and the full
 * walker count W.
 * afq_prop_te_anchor: TimeEvolvedBPEstimator::setAnchor (:271-276), X = Y = I, M = 0.
 * follow_walkers = 0 leaves the operators with their slots when population control copies
 * walkers, as the reference's estimator-held arrays do (a copy then carries another path's
 * operators); 1 makes population control copy them with their walkers. */
int afq_prop_te_anchor(afq_prop* prop, int follow_walkers);
/* AFQMCBasePropagator::PropagateOperators (AFQMCBasePropagator.cpp:538-684): the nsteps most
 * recent stored fields, oldest first, each X <- P1 expM(V) P1 X and
 * Y <- P1inv expM(-V)^H P1inv Y (FP64, P1inv = exp(+1/2 H1) of generateP1(true)); then, with
 * the walkers' Slater matrices S (orthonormal when the step was orthogonalized),
 * Y <- S S^H Y, M += Y^H X, X <- X - S S^H X.  1 <= nsteps <= nback.  Synchronizes once. */
int afq_prop_te_propagate(afq_prop* prop, const afq_walkers* w, int nsteps);
/* TimeEvolvedBPEstimator::evolveAndMeasure steps 2-3 (:228-262) with
 * constructTimeEvolvedMeasurementInputs (:49-146): per reference d of the trial,
 * G_d = M + (Ref_d^H X)^T (Gc_d conj(Y)), Gc_d the compact mixed density matrix of Ref_d and
 * the walker, coefficient conj(c_d) <Ref_d|Phi> normalized over the references, and WEIGHT
 * times the path-restoration factors of the last restore_steps steps (bp_step, or 2 bp_step
 * for extra_path_restoration, or 0).  transpose_m != 0 adds M^T in place of M: the reference
 * accumulates M = sum Y^H X, whose transpose is the term that makes G_d equal to the
 * back-propagated estimator of afq_prop_bp_measure; 0 reproduces the reference.  Outputs as
 * afq_prop_bp_measure. */
int afq_prop_te_measure(afq_prop* prop, const afq_walkers* w, int restore_steps, int transpose_m,
                        double* onerdm, double* diag2rdm);
/* Device views of X, Y, M (NULL before the first afq_prop_te_anchor) and of P1inv [N][N]
 * (built by this call when needed). */
int afq_prop_te_views(afq_prop* prop, const double** X, const double** Y, const double** M,
                      const double** P1inv);

/* ---- density-matrix observables --------------------------------------------------------------
 * One measurement of OneRDM (onerdm.hpp:43-70) and DiagonalTwoRDM (diagonal_twordm.hpp:40-81)
 * from per-walker density matrices G [W][nspin][N][N] (complex, device; the full mixed density
 * matrices of the walkers, or afq_prop_backpropagate's output), divided by the sum of the
 * weights as MeasurementOutput does (Measurements.hpp:98-121).  c_w = weights[w] * coeff[w]
 * (coeff may be NULL: 1).  onerdm [nspin][N][N]; diag2rdm [nblk][N][N] with the blocks
 * (aaaa), (aabb) for CLOSED and (aaaa), (aabb), (bbbb) for COLLINEAR, before the reference's
 * fold over i <-> j (DiagonalTwoRDM::compress, a host-side layout step).  Either output may be
 * NULL.  accumulate != 0 adds to the outputs (the sum over references of a multi-reference
 * trial: one call per reference with its coefficients, the referenceLoop of
 * MixedEstimator.hpp:77-103).  Device pointers, asynchronous on the context's stream. */
int afq_rdm_measure(afq_ctx* ctx, int walker_type, int N, int W, const double* G,
                    const double* weights, const double* coeff, int accumulate, double* onerdm,
                    double* diag2rdm);

/* ---- particle-hole multi-determinant trial (PHMSD, COLLINEAR walkers) ----------------------- */

/* The expansion in the ph_excitations numbering (Excitations.hpp:203-530, as build_ph_struct
 * makes it, readWfn.cpp:238-340). HOST pointers, copied at creation. Per spin s:
 *   nconf[s] unique configurations, configuration 0 = the reference, the rest ordered by rank;
 *   rank[s][u]; configuration u >= 1 is 2 rank ints at exc[s][exc_start[s][u]]: the hole
 *   positions in the reference, then the particle active orbitals;
 *   refc[s] [nel[s]] the reference occupation, which must be 0 .. nel-1 (woodbury algorithm).
 * Determinant d couples configurations det_a[d], det_b[d] with coefficient ci[d] (complex,
 * permutation sign included). Ranks up to 6. */
typedef struct {
  int nel[2];
  int nconf[2];
  const int* rank[2];
  const int* exc_start[2];
  const int* exc[2];
  const int* refc[2];
  int ndet;
  const int* det_a;
  const int* det_b;
  const double* ci;
} afq_ph_excitations;

/* PHMSD constructor (PHMSD.hpp:63-115). `ham` must be COLLINEAR with the trial's active
 * orbitals as its PsiT rows (nup = nact_alpha, ndown = nact_beta), as Wavefunction.cpp:459-474
 * builds it from PsiT_MO. */
int afq_phmsd_create(afq_ham* ham, const afq_ph_excitations* ex, afq_phmsd** out);
int afq_phmsd_destroy(afq_phmsd* trial);
/* PHMSD::MixedDensityMatrix, compact (PHMSD.cpp:549-790): Phi_a [W][N][nel_a], Phi_b
 * [W][N][nel_b] -> G [W][nact_a + nact_b][N], ovlp [W] = log overlap (PHMSD::vbias feeds this G
 * to afq_vbias). */
int afq_phmsd_mixed_density_matrix(afq_phmsd* trial, const double* Phi_a, const double* Phi_b,
                                   int64_t phi_stride, int W, double* G, double* ovlp);
/* PHMSD::Log_Overlap (PHMSD.cpp:804-900): ovlp [W]. */
int afq_phmsd_log_overlap(afq_phmsd* trial, const double* Phi_a, const double* Phi_b,
                          int64_t phi_stride, int W, double* ovlp);
/* PHMSD::Energy, energy_alg1 (PHMSD.cpp:376-544): E [W][3] = (E1 + E0, EXX, EJ), ovlp [W].
 * phi_stride: as for afq_walkers (0: each block contiguous). */
int afq_phmsd_energy(afq_phmsd* trial, const double* Phi_a, const double* Phi_b,
                     int64_t phi_stride, int W, double* E, double* ovlp);
/* PHMSD::vMF (PHMSD.cpp:1329-1573) + zero_imag (AFQMCBasePropagator.cpp:105-108): vMF [M]. */
int afq_phmsd_vmf(afq_phmsd* trial, double dt, double* vMF);

/* ---- multi-determinant NOMSD (NOMSD.hpp:54-100 with ndet > 1) ----
 * A non-orthogonal expansion sum_d c_d |D_d>: PsiT [ndet][nel][N] complex (device; determinant 0
 * must equal the PsiT `ham` was created with, whose half-rotated tensors it shares), ci [ndet]
 * complex (HOST).  Each further determinant gets its own half-rotated Lnak and haj.  CLOSED or
 * COLLINEAR (Phi_b NULL for CLOSED).  Full density matrices G are [W][nspin][N][N]. */
int afq_nomsd_create(afq_ham* ham, int ndet, const double* PsiT, const double* ci,
                     afq_nomsd** out);
int afq_nomsd_destroy(afq_nomsd* trial);
/* NOMSD::Log_Overlap (NOMSD.cpp:280-333): ovlp [W] = log sum_d conj(c_d) <D_d|Phi>. */
int afq_nomsd_log_overlap(afq_nomsd* trial, const double* Phi_a, const double* Phi_b,
                          int64_t phi_stride, int W, double* ovlp);
/* NOMSD::MixedDensityMatrix, full (NOMSD.cpp:219-278): G = sum_d w_d G_d / sum_d w_d with
 * w_d = conj(c_d) exp(ov_d - log_m[w]); log_m [W] is the scale (the walkers' stored log
 * overlap); ovlp [W] = sum_d w_d + log_m, the value the reference leaves in Ov. */
int afq_nomsd_mixed_density_matrix(afq_nomsd* trial, const double* Phi_a, const double* Phi_b,
                                   int64_t phi_stride, int W, const double* log_m, double* G,
                                   double* ovlp);
/* NOMSD::vbias with ndet > 1 (NOMSD.cpp:39-65): the full mixed density matrix, then
 * afq_vbias_full; v [W][M]. */
int afq_nomsd_vbias(afq_nomsd* trial, const double* Phi_a, const double* Phi_b,
                    int64_t phi_stride, int W, const double* log_m, double dt, int precision,
                    double* v);
/* NOMSD::Energy with ndet > 1 (NOMSD.cpp:159-217): per-determinant local energies from each
 * determinant's half-rotated tensors, combined with weights conj(c_d) <D_d|Phi> / <Psi_T|Phi>;
 * E [W][3] = (E1 + E0, EXX, EJ), ovlp [W] the log overlap. */
int afq_nomsd_energy(afq_nomsd* trial, const double* Phi_a, const double* Phi_b,
                     int64_t phi_stride, int W, int precision, double* E, double* ovlp);
/* NOMSD::vMF with ndet > 1 (NOMSD.cpp:437-521) + zero_imag (AFQMCBasePropagator.cpp:105-108):
 * vMF [M] from the trial's own full density matrix. */
int afq_nomsd_vmf(afq_nomsd* trial, double dt, double* vMF);
/* Real3IndexFactorization::vbias of full density matrices (the ndet > 1 branch,
 * Real3IndexFactorization.hpp:395-423): G [W][nspin][N][N],
 * v[w][m] = sqrt(dt) c sum_s sum_ik G[w][s][i][k] Likn[i][k][m] (c = 2 CLOSED / 1 COLLINEAR);
 * AFQ_MIXED contracts in FP32. */
int afq_vbias_full(afq_ham* ham, const double* G, int W, double dt, int precision, double* v);

#ifdef __cplusplus
}
#endif

#endif /* AFQ_H */
