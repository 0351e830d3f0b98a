/*
 * blfid.h -- C ABI of the MI355X powder-averaged MAS FID engine.
 *
 * This is the boundary a BlochLib caller binds instead of the solid-2.0 FID path.  It replaces
 *   oneFID<SolidSys, M>::FID(powder&, ro, det)   src/QMspins/driver/oneFID.h:200-208, 339-410
 *   oneFID<SolidSys, M>::FID(ro, det, th, ph, ga) src/QMspins/driver/oneFID.h:210-218, 419-431
 *   compute<SolidSys, M>  (gamma-COMPUTE)         src/QMspins/driver/compute.h:57-461
 *   SequenceRun::doFID 1-D fid() branch           examples/solid-2.0/src/sequencerun_funcs.cc:690-704
 * and the MPI powder reduction it sits in        src/QMspins/driver/oneFID.h:351-407,
 *                                                 examples/solid-2.0/src/sequencerun.cc:435-508.
 *
 * Conventions (all as in BlochLib):
 *   - Hilbert space of S spin-1/2 nuclei, N = 2^S; spin 0 is the leftmost Kronecker factor and
 *     m = +1/2 comes first (src/QMspins/multispinop.cc:51-70, singlespinop.cc:142-155).
 *   - Operator matrices (rho, det) are N x N complex, column-major, re/im interleaved doubles
 *     (BlochLib Full matrices are column-major, src/container/matrix/fmatrix.h:97-121).
 *   - Frequencies in Hz, times in s, powder angles in radians, Euler angles of interaction
 *     tensors and the rotor angle in degrees (csa.cc:276, oneFID.h:421).
 *   - Powder orientation k is (theta[k], phi[k], gamma[k]) with weight w[k]; a powder file row
 *     "a b g w" is phi=a, theta=b, gamma=g (src/QMspins/driver/powder.cc:333-389).
 *
 * Errors: every call returns a blfid_status; blfid_last_error() gives a thread-local message.
 * A plan is not thread safe; distinct plans may be used from distinct threads.
 */
#ifndef BLFID_H
#define BLFID_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    BLFID_OK = 0,
    BLFID_ERR_INVALID = 1,     /* bad argument (BLEXCEPTION cases of the reference) */
    BLFID_ERR_NOMEM = 2,       /* host or device allocation failed / over the HBM budget */
    BLFID_ERR_DEVICE = 3,      /* HIP runtime error or no device */
    BLFID_ERR_UNSUPPORTED = 4, /* valid input outside what this engine implements */
    BLFID_ERR_NOCONV = 5       /* eigen-solver did not converge */
} blfid_status;

typedef enum {
    BLFID_FP64 = 0, /* complex128 everywhere: oneFID<SolidSys, hmatrix> (oneFID.h:52) */
    BLFID_C64 = 1,  /* complex64 propagators, products and correlation sums, as solid-2.0's
                     * oneFID<SolidSys, hmatrixs> (oneFID.h:52-60, sequencerun_funcs.cc:602-609);
                     * the complex products run on FP32 matrix cores with FP32 accumulation; the
                     * period eigen-solver, phases, FID points and the powder reduction stay
                     * FP64.  The static path always runs in FP64. */
    BLFID_C64_MFMA = BLFID_C64,
    BLFID_C64_BF16X3 = 2 /* as BLFID_C64, but each FP32 operand of the complex products (Dyson,
                          * B = A W, Q) is split into bf16 hi + lo parts and every real product
                          * is hi*hi + hi*lo + lo*hi on bf16 matrix cores with FP32 accumulation
                          * (about 16 mantissa bits per product).  No reference counterpart: an
                          * opt-in speed/accuracy trade with a looser error than BLFID_C64. */
} blfid_precision;

/* Tensor-described SolidSys (src/QMspins/sys.h:59-62, read by sys.cc:120-204).
 *   isotope[S]       small integer label per spin; equal labels mean homonuclear
 *                    (T_Dip's A(on1)==A(on2) test, spin_ten.cc:441)
 *   csa[7*ncsa]      iso del eta spin alpha beta gamma      (C line, csa.cc:400-440)
 *   dip[6*ndip]      del spin1 spin2 alpha beta gamma       (D line, dip.cc:285-323)
 *   jc[4*nj]         iso spin1 spin2 strong                 (J line, j.cc:171-202) */
typedef struct {
    int nspins;
    const int *isotope;
    int ncsa;
    const double *csa;
    int ndip;
    const double *dip;
    int nj;
    const double *jc;
} blfid_system;

typedef struct blfid_plan blfid_plan;

/* Plan for one spin system and one acquisition: replaces constructing
 * oneFID(sys, npts, sw, wr, rotor_deg, t0) (oneFID.h:103-108) + compute(sys, wr, sw, t0, t0+1/wr)
 * (compute.h:213-235).  wr == 0 or rotor_deg == 0 selects the static FID (oneFID.h:216). */
blfid_status blfid_plan_create(const blfid_system *sys, double wr_hz, double rotor_deg,
                               double sw_hz, int npts, double t0, blfid_precision prec,
                               blfid_plan **out);
void blfid_plan_destroy(blfid_plan *plan);

/* One rf channel of a pulse program over the steps of one rotor period: on step k every spin
 * whose isotope label equals `isotope` gets amp_hz[k] (cos(phase_rad[k]) Ix + sin(phase_rad[k])
 * Iy) + offset_hz[k] Iz, and a step with amp_hz[k] == 0 adds nothing for this channel, its
 * offset included (PulseData::Pulse, examples/solid-2.0/src/pulsedata.cc:110-125; the script's
 * phase in degrees is converted to radians at pulsedata.cc:213-215). */
typedef struct {
    int isotope;
    const double *amp_hz;
    const double *phase_rad;
    const double *offset_hz;
} blfid_rf;

/* Plan for a rotor-synchronised point-to-point ("ptop") FID with rf pulses: the sub-sequence is
 * one rotor period 1/wr cut into `steps` equal steps; step k uses H(t1, t2, wr) at the midpoint
 * of [t0 + k tau, t0 + (k+1) tau], tau = 1/(wr steps), plus the rf of every channel, and
 * U <- exp(+i 2 pi tau (H + H_rf)) U.  FID point m records Tr(rho_m det), rho_{m+1} = U rho_m U^H.
 * Replaces Propogation::propogator (examples/solid-2.0/src/propogation.cc:240-304),
 * subSequence::propogate (sequencerun.cc:164-196) and the ptop branch of SequenceRun::doFID
 * (sequencerun_funcs.cc:624-655).  The sampling interval (blfid_plan_sw) is wr.  A strong
 * (I.S) J coupling between different isotopes where one of them is pulsed is rejected with
 * BLFID_ERR_UNSUPPORTED. */
blfid_status blfid_plan_create_ptop(const blfid_system *sys, double wr_hz, double rotor_deg,
                                    int steps, int npts, double t0, int nrf, const blfid_rf *rf,
                                    blfid_precision prec, blfid_plan **out);

/* The same for a general sub-sequence: `nsteps` steps of lengths step_dt[k] (s), each with the
 * rf tables' entry k, e.g. a list of pulses and delays each split into ceil(length / maxtstep)
 * equal steps as Propogation::propogator does (propogation.cc:257-300).  Step k spans
 * [step_t1[k], step_t1[k] + step_dt[k]] (the rotor phase is taken at its midpoint); with
 * step_t1 == NULL the steps are consecutive from t0.  wr may be 0 (no rotor phase).  One FID
 * point per application of the whole sub-sequence (blfid_plan_sw = 1 / total length). */
blfid_status blfid_plan_create_pulse(const blfid_system *sys, double wr_hz, double rotor_deg,
                                     int nsteps, const double *step_t1, const double *step_dt,
                                     int npts, double t0, int nrf, const blfid_rf *rf,
                                     blfid_precision prec, blfid_plan **out);

/* Pulses applied to rho before the FID of a spinning gamma-COMPUTE plan, as solid-2.0 runs a
 * plain fid() after pulses (SequenceRun::doFID, sequencerun_funcs.cc:693-703): for every
 * crystallite rho <- U rho U^H with U the propagator of the `nsteps` steps (lengths step_dt[k],
 * start times step_t1[k] or consecutive from 0 when NULL, rf channels as in
 * blfid_plan_create_pulse, rotor angle rotor_deg, the plan's spinning speed), computed in FP64;
 * the FID then runs from the plan's start time t0 (oneFID::setStartTime(currentTime),
 * sequencerun_funcs.cc:611), i.e. the plan should be created with t0 = the pulses' end.  The
 * prepared rho is treated as general (the rho == (det + det^H)/2 shortcut of compute.h:75-83 is
 * not taken; for a prepared rho it is an exact floating-point coincidence).  nsteps == 0 removes
 * the preparation.  Operators must be set (again) after this call.  Replaces the per-crystallite
 * PropList[j].propogate(ro) + hFID.FID(ro, ...) of the reference.  Static and pulse-program
 * plans return BLFID_ERR_UNSUPPORTED. */
blfid_status blfid_plan_set_prep(blfid_plan *plan, double rotor_deg, int nsteps,
                                 const double *step_t1, const double *step_dt, int nrf,
                                 const blfid_rf *rf);

/* Spin state: the density operator of every crystallite of one powder, resident in device
 * memory, carried through a pulse program.  Replaces SequenceRun's per-orientation `ro`
 * (examples/solid-2.0/src/sequencerun.h:268, reset at every powder point,
 * sequencerun.cc:512-527) and its in-place propagation by the used sub-sequences
 * (subSequence::propogate, sequencerun.cc:164-196, called from SequenceRun::doFID,
 * sequencerun_funcs.cc:625-744); ro(op) is blfid_state_set (setRo, sequencerun_funcs.cc:
 * 315-331), a used propagator is blfid_state_apply, fid(i)'s single point is blfid_state_trace
 * (sequencerun_funcs.cc:635, 694), and an FID from the current rho is a plan run after
 * blfid_plan_set_state.  Crystallite k has angles (theta[k], phi[k], gamma[k]) and weight w[k].
 * Every state call returns after its work on `stream` is done; a plan run that reads the state
 * must finish before the state is changed (use one stream).  The state holds K N^2 complex128. */
typedef struct blfid_state blfid_state;
blfid_status blfid_state_create(int nspins, int K, const double *theta, const double *phi,
                                const double *gamma, const double *weight, blfid_state **out);
void blfid_state_destroy(blfid_state *state);
int blfid_state_dim(const blfid_state *state);  /* N */
int blfid_state_size(const blfid_state *state); /* K */
/* Device bytes blfid_state_apply may use for its propagators (0: 40% of the free memory). */
blfid_status blfid_state_set_memory_budget(blfid_state *state, double bytes);
/* Every crystallite's rho = rho (N x N, column-major, interleaved). */
blfid_status blfid_state_set(blfid_state *state, const double *rho, void *stream);
/* Crystallite k's rho = rho_all[k] (K consecutive N x N matrices); fetch is the reverse. */
blfid_status blfid_state_load(blfid_state *state, const double *rho_all, void *stream);
blfid_status blfid_state_fetch(blfid_state *state, double *rho_all, void *stream);
/* rho_k <- U_k^repeat rho_k U_k^-repeat for every crystallite k, U_k the propagator of the
 * `nsteps` steps (lengths step_dt, start times step_t1 or consecutive from 0 when NULL, rf
 * channels as in blfid_plan_create_pulse, rotor angle rotor_deg, spinning speed wr_hz) of
 * `sys` at crystallite k's orientation, computed in FP64 (Propogation::propogator,
 * propogation.cc:240-304; a sub-sequence used `repeat` times applies the same U repeat times,
 * subSequence::propogate, sequencerun.cc:183-195).  repeat == 0 changes nothing. */
blfid_status blfid_state_apply(blfid_state *state, const blfid_system *sys, double wr_hz,
                               double rotor_deg, int nsteps, const double *step_t1,
                               const double *step_dt, int nrf, const blfid_rf *rf, int repeat,
                               void *stream);
/* out2 = sum_k w_k Tr(rho_k det) [/ Tr(det det^H) when normalize], summed in crystallite order
 * in groups as the plan's powder sum (sequencerun_funcs.cc:635, 694: fid1D[idx] +=
 * weight trace(ro, detect) / trace(detect, adjoint(detect))). */
blfid_status blfid_state_trace(blfid_state *state, const double *det, int normalize,
                               double *out2, void *stream);
/* rho_k <- Tr(rho_k op^H) op / Tr(op op^H) for every crystallite: the main-section cycler(op)
 * entry of the propagator list, which projects ro onto op where the list is applied
 * (subSequence::propogate with docycler, sequencerun.cc:164-178; doCycler,
 * sequencerun_funcs.cc:232-256).  op: N x N complex128, column-major, not zero. */
blfid_status blfid_state_project(blfid_state *state, const double *op, void *stream);
/* The plan's FID starts from the state's rho_k for crystallite k of the plan's powder (same K;
 * the same orientations are the caller's to keep) instead of the plan's rho: a gamma-COMPUTE,
 * static or pulse-program FID from each crystallite's current rho (sequencerun_funcs.cc:
 * 699-735).  rho passed to blfid_plan_set_operators is then ignored except for its size; the
 * rho == (det + det^H)/2 shortcut is not taken.  Detaches a preparation (blfid_plan_set_prep);
 * NULL detaches the state.  The state must outlive the plan's runs.  Operators must be set
 * (again) after this call. */
blfid_status blfid_plan_set_state(blfid_plan *plan, blfid_state *state);

/* The plan's whole run (every crystallite batch's launches) recorded once as a HIP graph and
 * launched as one: no reference counterpart (the reference has no device launches), for callers
 * that repeat one plan (a fixed acquisition over many parameter sets, the benchmark's steps).
 * Capture allocates the plan's working storage and records, without running anything; launch
 * runs the recorded work on `stream` and blfid_plan_fetch reads the result, identical to
 * blfid_plan_run's.  A state attached to the plan is read at launch time.  Any later change to
 * the plan (operators, powder, preparation, state, workspace, budget, profiling) makes launch
 * return BLFID_ERR_INVALID until the next capture.  Capture needs the plan's own working storage
 * (no shared workspace) and profiling off.  Graph launches and runs of one plan share its
 * buffers: keep them on one stream. */
blfid_status blfid_plan_graph_capture(blfid_plan *plan, int normalize);
blfid_status blfid_plan_graph_launch(blfid_plan *plan, void *stream);

/* The run in slices of the powder: crystallites [c_begin, c_end) of the plan's powder, c_begin
 * == 0 starting the FID and c_end == K finishing it (then blfid_plan_fetch as after a run).
 * Consecutive slices (each starting where the previous one ended) give blfid_plan_run's FID bit
 * for bit.  No reference counterpart: it lets a server interleave the slices of a large request
 * with other requests on the device.  Needs the plan's own working storage; the plan must not
 * change, or be run otherwise, between its first and last slice. */
blfid_status blfid_plan_run_slice(blfid_plan *plan, int c_begin, int c_end, int normalize,
                                  void *stream);

/* Adjusted sweep width n*wr (compute.h:96, oneFID.h:429), steps per rotor period n, N. */
double blfid_plan_sw(const blfid_plan *plan);
int blfid_plan_steps(const blfid_plan *plan);
int blfid_plan_dim(const blfid_plan *plan);

/* Detection and initial density operators (oneFID::FID(.., ro, det), oneFID.h:200-218). */
blfid_status blfid_plan_set_operators(blfid_plan *plan, const double *rho, const double *det);

/* Powder set, uploaded to device memory once (the powder class, powder.h:52-162). */
blfid_status blfid_plan_set_powder(blfid_plan *plan, const double *theta, const double *phi,
                                   const double *gamma, const double *weight, int K);

/* Run the whole powder on the device, asynchronously on `stream` (a hipStream_t, or NULL for
 * the default stream).  normalize != 0 divides by Tr(det det^H) as solid-2.0's doFID does
 * (sequencerun_funcs.cc:698-702); normalize == 0 is oneFID::FID(powder&,..) (oneFID.h:396-402).
 * The result stays on the device until blfid_plan_fetch. */
blfid_status blfid_plan_run(blfid_plan *plan, int normalize, void *stream);

/* Variant of blfid_plan_run that allocates nothing: no reference counterpart.  It is for callers
 * that cannot allocate device memory while they run (HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The working storage is a static device array of
 * BLFID_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends); the crystallites run in
 * batches that fit it, with fewer concurrent scratch slots than blfid_plan_run, and the FID is
 * bit for bit the one blfid_plan_run gives.  The normalised FID (2*npts doubles, re/im) is
 * written to the caller's device buffer fid_out_dev, asynchronously on `stream`.  Plan tables,
 * operators and powder must have been set (those calls allocate, at set-up).  A plan whose single
 * crystallite needs more than the scratch returns BLFID_ERR_NOMEM before any launch.  Every call
 * on one GPU shares that GPU's array: calls that overlap in time on one GPU (two streams, two
 * threads) race and must be serialised by the caller.  Profiling is not recorded here. */
#define BLFID_NOALLOC_SCRATCH_BYTES (256u << 20)
blfid_status blfid_plan_run_noalloc(blfid_plan *plan, int normalize, double *fid_out_dev,
                                    void *stream);

/* Wait for the last run and copy its FID (2*npts doubles, re/im interleaved) to host. */
blfid_status blfid_plan_fetch(blfid_plan *plan, double *fid_out);

/* Device pointer of the last run's FID (2*npts doubles), valid until the next run/destroy. */
const double *blfid_plan_device_fid(const blfid_plan *plan);

/* One-call form of oneFID::FID(powder&, ro, det): set operators + powder, run, fetch. */
blfid_status blfid_powder_fid(blfid_plan *plan, const double *theta, const double *phi,
                              const double *gamma, const double *weight, int K,
                              const double *rho, const double *det, int normalize,
                              double *fid_out);

/* Work accounting of the plan for the current operators (per crystallite):
 * out[0] executed flops per crystallite-step (the propagator counted at its nominal degree-8
 * Taylor form, 7 real GEMMs per block), out[1] reference dense flops per
 * crystallite-step (48N^3 + 14nN^2 + 14PN^2/n), out[2] number of Hilbert-space blocks,
 * out[3] largest block, out[4] device bytes per crystallite of working storage,
 * out[5] fixed device bytes of working storage independent of the batch size. */
blfid_status blfid_plan_stats(const blfid_plan *plan, double *out6);

/* Limit device working memory of a plan (bytes); crystallites are processed in batches that
 * fit.  0 = default (a fraction of free HBM). */
blfid_status blfid_plan_set_memory_budget(blfid_plan *plan, double bytes);

/* Device workspace shared by several plans (crystallite batches are sized to fit it).  A run
 * on a stream other than the workspace's previous user first waits for that user's work, so
 * plans sharing a workspace serialise safely.  Without a workspace a plan grows its own,
 * sized by blfid_plan_set_memory_budget.  The workspace must outlive the plans using it. */
typedef struct blfid_workspace blfid_workspace;
blfid_status blfid_workspace_create(double bytes, blfid_workspace **out);
void blfid_workspace_destroy(blfid_workspace *ws);
blfid_status blfid_plan_set_workspace(blfid_plan *plan, blfid_workspace *ws);

/* Per-kernel timing with HIP events recorded on the run's stream around every launch
 * (enable before blfid_plan_run; costs a few microseconds per launch).  Phases:
 *   0 exp (H build + propagator)   1 dyson   2 period eig   3 B = A W   4 Q/R similarity
 *   5 correlation + FID            6 powder reduce            7 static FID
 *   8 fused propagator + Dyson chain (complex64 gamma-COMPUTE plans, blocks of <= 80 states)
 * Templated kernels come in tile variants M = 1..5 (16M x 16M output tiles); for the period
 * eigen-solver (phase 2) M - 1 is its block-size class (b <= 16, 32, 48, 64, 80, larger).
 * blfid_plan_phase_times waits for the last run and returns, per slot phase*8 + M (M = 0 for
 * untemplated kernels), the summed kernel time in ms, the number of launches and the executed
 * algorithmic flops (unpadded block sizes; the propagator's as executed, its Taylor degree being
 * chosen per task on the device; the correlation's as FFT or direct sum, as run), accumulated
 * since profiling was enabled.  Arrays hold BLFID_NSLOT entries. */
#define BLFID_NPHASE 9
#define BLFID_NSLOT 72
blfid_status blfid_plan_set_profiling(blfid_plan *plan, int on);
blfid_status blfid_plan_phase_times(blfid_plan *plan, double *ms, int *launches, double *flops);
/* Algorithmic HBM bytes per slot (same slots, accumulated likewise): what each kernel must read
 * and write at the least -- its inputs once and its outputs once (U_k written by the propagator;
 * U_k read and A_k written by the Dyson chain; A_n read and W, eigenphases written by the eigen-
 * solver; A_k, W read and B_k written; both blocks of a pair read and Q (R) written; Q (and R or
 * the reverse Q) read and the partial FIDs written by the correlation; partial FIDs read by the
 * powder sum).  Beside counter-measured traffic it shows re-reads and spills. */
blfid_status blfid_plan_phase_bytes(blfid_plan *plan, double *bytes);

/* Launch autotuner (no counterpart in the reference, whose CPU loops have no launch shape).
 * A tuning is BLFID_TUNE_LEN ints:
 *   [0..4]  propagator method of tile class M = 1..5 (blocks of 16(M-1)+1 .. 16M states, class 5
 *           up to 80) in complex64 gamma-COMPUTE plans: 0 the fused propagator + Dyson chain
 *           kernel, 1 the propagator and the Dyson chain as two kernels;
 *   [5]     correlation kernel variant: 0 four FID points per thread up to 1024 points, eight
 *           above; 1 the same with the complex64 four-point kernel at two waves per SIMD; 2 eight
 *           points per thread throughout;
 *   [6..14] persistent-grid cap (workgroups) of the kernel families in profiling-phase order
 *           (exp, dyson, eig, bmul, qr, corr, reduce, static, expdyson); 0 the default 4096;
 *   [15]    Q-product (k_qr) form of the complex64 tile classes 1-4: bit M-1 set takes the other
 *           form (one task per step for classes 1-2, whose default walks the steps of a
 *           crystallite with the pair in LDS; the step-walking form for classes 3-4).
 * Every choice gives the same FID bit for bit except [0..4] and [15], which switch between two
 * complex64 forms (both within the complex64 tolerance of the FP64 form).  Each device has a process-wide
 * table (all zero at start); a plan takes the table of the current device when it is created.
 * blfid_autotune times the candidates of every entry on the leading crystallites of each plan --
 * `crystallites` of them, or more where that many would leave a launch with one task per
 * crystallite and block (or block pair) of a tile class with fewer than 2048 workgroups (a
 * sample smaller than the runs' batches would favour the forms with a task per step), up to the
 * whole powder; it times them with profiling events per kernel family and tile class, one run
 * per candidate, in seven candidate passes (three for [0..5] and [15], four grid caps), keeps a
 * candidate only
 * when it is at least 3% faster than the current choice, writes the result to every plan and to
 * the process table of the current device, and stores the summed time of each pass (ms) in
 * report[0..6] when report is not NULL.  The plans must be gamma-COMPUTE plans with operators
 * and powder set; their profiling counters are cleared (the profiling switch is kept), their
 * last FID is overwritten (run them again before fetching) and their recorded graphs
 * invalidated.  Work is enqueued on `stream` and waited for.  A failure leaves every plan and the
 * process table with the tuning they had. */
#define BLFID_TUNE_LEN 16
#define BLFID_TUNE_PASSES 7
blfid_status blfid_autotune(blfid_plan **plans, int nplans, int crystallites, void *stream,
                            double *report);
blfid_status blfid_tune_get(int *vals);
blfid_status blfid_tune_set(const int *vals);
blfid_status blfid_plan_tune_get(const blfid_plan *plan, int *vals);
blfid_status blfid_plan_tune_set(blfid_plan *plan, const int *vals);

/* Thread-local description of the last error. */
const char *blfid_last_error(void);

/* Engine version string. */
const char *blfid_version(void);

#ifdef __cplusplus
}
#endif
#endif /* BLFID_H */
