/*
 * lulesh_mi355x.h -- C ABI of the MI355X LULESH Lagrange step.
 *
 * One call of lulesh_step() advances the Sedov blast problem by whole
 * cycles, each cycle being TimeIncrement followed by LagrangeLeapFrog, the
 * loop body of the reference drivers:
 *   reference/C/LULESH-OMP.cc:3179-3188           (LULESH 1.0)
 *   reference/C/lulesh-2.0.3/lulesh.cc:2771-2781  (LULESH 2.0.3)
 *   src/Main.hs:90-109, src/LULESH.hs:37-53       (the Accelerate step)
 * Every phase runs as a HIP kernel on one gfx950 GPU; the domain state stays
 * resident in device memory between calls.
 *
 * Plain C types only: the Haskell driver binds these with `foreign import
 * ccall` (the pattern of src/VisIt.hs:71-86) and the C++ harness links them
 * directly.  All calls are synchronous on return and belong to the calling
 * host thread; one domain must not be used by two threads at once.
 *
 * Return codes mirror the reference's exit codes
 * (reference/C/LULESH-OMP.cc:82, lulesh-2.0.3/lulesh.h:41).
 */
#ifndef LULESH_MI355X_H
#define LULESH_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

#define LULESH_OK 0
#define LULESH_VOLUME_ERROR (-1) /* VolumeError: a volume went <= 0 */
#define LULESH_QSTOP_ERROR (-2)  /* QStopError: q exceeded qstop */
#define LULESH_BAD_ARG (-3)
#define LULESH_HIP_ERROR (-4)

/* Initial conditions.  1.0: e(0) = 3.948746e7, dt0 = 1e-7, one region
 * (reference/C/LULESH-OMP.cc:3052-3112).  2.0: e(0) scaled by (nx/45)^3,
 * dt0 from the CFL estimate, num_reg seeded regions with cost-dependent EOS
 * repetitions (reference/C/lulesh-2.0.3/lulesh-init.cc:16-184, 377-484;
 * lulesh.cc:2357-2430). */
#define LULESH_IC_1_0 1
#define LULESH_IC_2_0 2

typedef struct lulesh_domain lulesh_domain;

/* The run options of lulesh-2.0.3's command line
 * (reference/C/lulesh-2.0.3/lulesh-util.cc:60-168; lulesh.cc:2703-2711). */
typedef struct {
    int nx;          /* -s: elements along each edge of the cube (>= 1) */
    int ic_mode;     /* LULESH_IC_1_0 or LULESH_IC_2_0 */
    int num_reg;     /* -r: regions (2.0 only), default 11 */
    int balance;     /* -b: region size balance (2.0 only), default 1 */
    int cost;        /* -c: extra cost of the costly regions, default 1 */
    int max_cycles;  /* -i: stop after this many cycles, default 9999999 */
    double stoptime; /* stop when time >= stoptime, default 1e-2 */
} lulesh_config;

/* Scalars of the current state. */
typedef struct {
    double time;       /* domain.time() */
    double deltatime;  /* domain.deltatime() */
    double dtcourant;  /* domain.dtcourant() */
    double dthydro;    /* domain.dthydro() */
    double e_origin;   /* e(0), the "Final Origin Energy" */
    int cycle;         /* domain.cycle() */
    int done;          /* time >= stoptime or cycle >= max_cycles */
    int num_elem;
    int num_node;
} lulesh_stats;

/* Field identifiers for lulesh_get_field / lulesh_set_field.  Node fields
 * have num_node values, element fields num_elem; both in the reference's
 * order (x fastest, then y, then z). */
enum {
    LULESH_F_X = 0, LULESH_F_Y, LULESH_F_Z,
    LULESH_F_XD, LULESH_F_YD, LULESH_F_ZD,
    LULESH_F_NODAL_MASS,
    LULESH_F_E = 16, LULESH_F_P, LULESH_F_Q, LULESH_F_QL, LULESH_F_QQ,
    LULESH_F_V, LULESH_F_VOLO, LULESH_F_DELV, LULESH_F_VDOV, LULESH_F_AREALG,
    LULESH_F_SS, LULESH_F_ELEM_MASS, LULESH_F_VNEW,
    LULESH_F_DELV_XI, LULESH_F_DELV_ETA, LULESH_F_DELV_ZETA,
    LULESH_F_DELX_XI, LULESH_F_DELX_ETA, LULESH_F_DELX_ZETA
};

/* Fills the reference defaults: nx 30, 2.0 ICs, 11 regions, balance 1,
 * cost 1, no cycle limit, stoptime 1e-2. */
void lulesh_config_default(lulesh_config *cfg);

/* Builds the mesh, boundary conditions, regions and initial conditions of
 * Domain::Domain (reference/C/lulesh-2.0.3/lulesh-init.cc:16-184;
 * reference/C/LULESH-OMP.cc:2960-3171) and uploads them to the GPU. */
int lulesh_create(const lulesh_config *cfg, lulesh_domain **out);
void lulesh_destroy(lulesh_domain *d);

/* The mixed-precision form of the same step (no counterpart in the
 * reference, which is float64 throughout): the element-local arithmetic of
 * the stress integration, hourglass control, monotonic Q and EOS runs in
 * float32, while node coordinates and velocities, element volumes (the
 * Jacobian determinants), the energy update and the dt reductions stay
 * float64.  Every other entry point works on such a domain unchanged; the
 * float32 fields are widened to double on the way out and rounded on the
 * way in.  Its results differ from lulesh_create's by rounding: the origin
 * energy at t = 1e-2 within 1e-5 relative at 120^3. */
int lulesh_create_mixed(const lulesh_config *cfg, lulesh_domain **out);
/* lulesh_create (mixed = 0) or lulesh_create_mixed (mixed = 1) with the
 * initial origin energy multiplied by efactor (and, in 2.0 mode, the initial
 * dt derived from it): one member of a seeded ensemble of Sedov problems. */
int lulesh_create_scaled(const lulesh_config *cfg, double efactor, int mixed,
                         lulesh_domain **out);
/* An unstructured hexahedral mesh (the .lmesh input of
 * reference/CUDA/lulesh-kepler-singlegpu/lulesh.cu:728-925) in LULESH 1.0
 * semantics (cfg->ic_mode must be LULESH_IC_1_0; only stoptime and
 * max_cycles are read; NULL: the defaults with 1.0 ICs -- pass a config).
 * The element face masks are derived from the node sets as the reference's
 * reader derives them: a face whose four nodes lie on the symmetry sets is
 * a symmetry face, and so is one whose four nodes lie on the free-surface
 * set; the initial energy goes into the first element whose mask is the
 * three low symmetry faces.  Corner forces are gathered per node in
 * ascending element order (reference/C/LULESH-OMP.cc:209-268).  The plane-0
 * symmetry check, sweep shapes and the learned closure do not apply.
 * lulesh_create_lmesh reads the file: node and element counts, the node
 * coordinates, 8 nodes per element, 6 face neighbours per element
 * (xi-, xi+, eta-, eta+, zeta-, zeta+), then the x, y, z symmetry node sets
 * and the free-surface node set, each preceded by its size.
 * lulesh_create_mesh takes the same lists as arrays. */
int lulesh_create_lmesh(const char *path, const lulesh_config *cfg, lulesh_domain **out);
int lulesh_create_mesh(int num_elem, int num_node, const double *x, const double *y,
                       const double *z, const int *nodelist, const int *neighbors,
                       int nsx, const int *symm_x, int nsy, const int *symm_y, int nsz,
                       const int *symm_z, int nfree, const int *free_surf,
                       const lulesh_config *cfg, lulesh_domain **out);
/* The element holding the initial energy (0 on the cube). */
int lulesh_origin_element(const lulesh_domain *d);
/* How an unstructured domain's force kernel cuts the mesh, no counterpart in
 * the reference: out[0] element blocks (at most 512 elements each, by
 * recursive coordinate bisection), out[1] interior nodes (all their
 * elements in one block: summed from on-chip corner forces), out[2]
 * boundary nodes (the rest: summed from corner forces in device memory),
 * out[3] element corners of boundary nodes (the corner forces written to
 * device memory per cycle, stress and hourglass).  LULESH_BAD_ARG on the
 * cube. */
int lulesh_force_blocks(const lulesh_domain *d, long long out[4]);
/* The element corners [num_elem][8] (node indices, bottom face
 * counter-clockwise, then the top face): the nodelist of
 * reference/C/LULESH-OMP.cc:2996-3012 for the cube, the file's lists for a
 * mesh read by lulesh_create_lmesh (lulesh.cu:728-848). */
int lulesh_get_nodelist(lulesh_domain *d, int *nodelist);
/* 1 for a domain made by lulesh_create_mixed, else 0. */
int lulesh_is_mixed(const lulesh_domain *d);

/* Advances up to ncycles cycles, stopping early at stoptime/max_cycles or on
 * an error; *cycles_done (may be NULL) receives the cycles completed.
 * Replaces one or more turns of the reference loop
 * (reference/C/lulesh-2.0.3/lulesh.cc:2771-2781) and of `simulate`
 * (src/Main.hs:90-109). */
int lulesh_step(lulesh_domain *d, int ncycles, int *cycles_done);

/* Runs to completion (time >= stoptime or cycle >= max_cycles). */
int lulesh_run(lulesh_domain *d, int *cycles_done);

/* Float64 arithmetic with fused multiply-adds (no counterpart in the
 * reference, whose C code rounds every product and every sum): enable = 1
 * lets the element phases (stress, hourglass, kinematics, Q gradients,
 * limiter, EOS) fuse a * b + c into one rounding and take their quotients
 * through refined hardware reciprocals, which makes the 150^3 step about 14%
 * faster.  Results then differ from the reference by rounding only:
 * measured over full runs at 30^3 and 45^3, the same cycle counts, origin
 * energy within 1e-14 relative and fields within 1e-11 relative (the
 * accepted tolerance is 1e-10).  enable = 0, the default, rounds every value
 * as the reference does, bit for bit.  Float64 domains, structured or
 * unstructured; not the mixed-precision form. */
int lulesh_set_fused_multiply_add(lulesh_domain *d, int enable);

/* State in the Accelerate step's terms (src/Domain.hs:11-21): the domain
 * tuple (x, xd, e, p, q, v, ss, t, dt).  Host arrays, copied in/out. */
int lulesh_upload_state(lulesh_domain *d, const double *x, const double *y,
                        const double *z, const double *xd, const double *yd,
                        const double *zd, const double *e, const double *p,
                        const double *q, const double *v, const double *ss,
                        double time, double deltatime);
int lulesh_download_state(lulesh_domain *d, double *x, double *y, double *z,
                          double *xd, double *yd, double *zd, double *e,
                          double *p, double *q, double *v, double *ss,
                          double *time, double *deltatime);

/* One field, copied to/from a host buffer of num_node or num_elem values. */
int lulesh_get_field(lulesh_domain *d, int field, double *host_out);
int lulesh_set_field(lulesh_domain *d, int field, const double *host_in);

/* Scalars; lulesh_set_scalars restores a checkpoint's time, dt, cycle and
 * constraints (any NULL pointer leaves that value unchanged). */
int lulesh_get_stats(lulesh_domain *d, lulesh_stats *out);
int lulesh_set_scalars(lulesh_domain *d, const double *time,
                       const double *deltatime, const double *dtcourant,
                       const double *dthydro, const int *cycle);

/* Sets one constant of the method by its reference name: "hgcoef", "qstop",
 * "monoq_max_slope", "monoq_limiter_mult", "qlc_monoq", "qqc_monoq", "qqc",
 * "eosvmax", "eosvmin", "pmin", "emin", "dvovmax", "refdens", "e_cut",
 * "p_cut", "q_cut", "u_cut", "v_cut", "stoptime", "dtmax", "dtfixed"
 * (the Domain parameters of reference/C/LULESH-OMP.cc:351-380 and
 * lulesh-2.0.3/lulesh-init.cc:19-41).  Takes effect at the next cycle. */
int lulesh_set_parameter(lulesh_domain *d, const char *name, double value);

/* Launch shape of the nodal-phase kernel: tx x ty element columns per block,
 * nz node planes per block, and the waves per SIMD its register budget
 * leaves room for.  Only the shapes the library was built with are accepted
 * (LULESH_BAD_ARG otherwise); results do not depend on the shape. */
int lulesh_set_sweep_shape(lulesh_domain *d, int tx, int ty, int nz, int waves);
/* Schedule of the Q/EOS kernel (float64 and mixed domains).  0, the
 * default: blocks of consecutive elements in mesh order, each block sharing
 * its elements' EOS repetitions among its threads, the costliest blocks
 * first.  bands > 0: consecutive blocks take element
 * chunks from `bands` bands of the region-sorted EOS order in turn (1 =
 * costly regions first, then the rest; odd counts spread the costly band
 * over the 8 XCDs).  A tuning knob with no counterpart in the reference;
 * results do not depend on it. */
int lulesh_set_eos_bands(lulesh_domain *d, int bands);
int lulesh_get_sweep_shape(lulesh_domain *d, int *tx, int *ty, int *nz,
                           int *waves);

/* Region of each element (1-based, the reference's regNumList). */
int lulesh_get_regions(lulesh_domain *d, int *reg_num_list);

/* Plane-0 energy symmetry of the reference's final check
 * (reference/C/lulesh-2.0.3/lulesh-util.cc:189-208). */
int lulesh_symmetry(lulesh_domain *d, double *max_abs_diff,
                    double *total_abs_diff, double *max_rel_diff);

/* The state as a file for visual verification: replaces writeDomain
 * (cbits/visit.cc:147-175, bound by src/VisIt.hs:71-86), with the same
 * arguments, writing a legacy VTK unstructured grid (VisIt and ParaView read
 * it; the Silo library is not available): hexahedra of the num_elem = n^3
 * cube, node velocities, cell e, p, v, q, time and cycle.  The arrays are
 * host arrays in the reference's order. */
int lulesh_write_domain_arrays(const char *path, int num_elem, int step, double time,
                               const double *x, const double *y, const double *z,
                               const double *xd, const double *yd, const double *zd,
                               const double *e, const double *p, const double *v,
                               const double *q);
/* The same for a domain's current state (downloaded first), the cube or an
 * unstructured mesh (its own connectivity). */
int lulesh_write_domain(lulesh_domain *d, const char *path);

/* Per-kernel timing with HIP events on the domain's stream.  While enabled,
 * every launch of lulesh_step is bracketed by events; lulesh_kernel_times
 * returns, per kernel, the accumulated milliseconds and launch count. */
int lulesh_kernel_count(void);
const char *lulesh_kernel_name(int k);
int lulesh_set_kernel_timing(lulesh_domain *d, int enable);

/* lulesh_step replays its cycles from a captured HIP graph of 32 cycles
 * (captured again whenever a parameter, mode or launch shape changes), and
 * queues the rest one launch at a time.  enable = 0 queues every cycle one
 * launch at a time; the results are the same bits either way.  Default 1.
 * No counterpart in the reference, whose Accelerate term is compiled and
 * run by the Accelerate runtime. */
int lulesh_set_cycle_graphs(lulesh_domain *d, int enable);
int lulesh_kernel_times(lulesh_domain *d, double *ms, long long *launches);

/* Milliseconds between the HIP events bracketing the last lulesh_step call
 * (0 when timing was never enabled). */
double lulesh_last_step_ms(lulesh_domain *d);

/* Message of the last error on this domain ("" when none). */
const char *lulesh_last_error(const lulesh_domain *d);

/* The library's cube root (lulesh_cbrt.h) evaluated on the GPU (op 0), IEEE
 * sqrt (1) and division (2), the EOS's square root for inputs >= 2^-767
 * (3), and two quotients of the fused multiply-add build: a / b (4: the
 * hardware reciprocal refined by two Newton steps, times a) and the
 * hourglass coefficient's a / cbrt(b) for b > 0 normal (5: a times the
 * reciprocal cube root), for n inputs: a math-parity probe for the tests. */
int lulesh_device_math(int op, int n, const double *a, const double *b,
                       double *out);

/* ------------------------------------------------------------------------
 * Ensembles.  Many independent Sedov domains advanced together on one GPU:
 * each phase of a cycle is one kernel over the tiles / element chunks of
 * all members (one block per item, or a persistent grid: the launch shape),
 * every member keeps its own time, dt and stop test
 * (the loop of reference/C/lulesh-2.0.3/lulesh.cc:2771-2781 run per
 * member), and the cycle is replayed from a captured HIP graph.  The
 * reference runs one domain per process; a caller replaces N runs of
 * `lulesh` (src/Main.hs:90-109) by one batch.  Each member's results are
 * bit-identical to a lulesh_create/lulesh_run of the same configuration.
 * ------------------------------------------------------------------------ */
typedef struct lulesh_batch lulesh_batch;

/* An ensemble drawn from a seed: edge sizes uniform in [nx_min, nx_max] and
 * origin-energy factors uniform in [efactor_min, efactor_max], both from
 * the project's splitmix64 stream (nx first, then the factor, per member). */
typedef struct {
    int num_members;
    int nx_min, nx_max;
    double efactor_min, efactor_max;
    unsigned long long seed;
} lulesh_ensemble;

/* 64 members, nx in [16, 64], factors in [0.8, 1.2], seed 0. */
void lulesh_ensemble_default(lulesh_ensemble *e);
/* Writes the num_members edge sizes and energy factors of the ensemble. */
int lulesh_ensemble_draw(const lulesh_ensemble *e, int *nx, double *efactor);

/* n members with configurations cfg[0..n-1] (all with the same ic_mode)
 * and origin-energy factors efactor[0..n-1] (NULL: all 1.0); the factor
 * multiplies the member's initial e(0), and in 2.0 mode its dt0 follows. */
int lulesh_batch_create(const lulesh_config *cfg, const double *efactor, int n,
                        lulesh_batch **out);
void lulesh_batch_destroy(lulesh_batch *b);

/* Advances every member that has cycles left by up to ncycles cycles;
 * *pending (may be NULL) receives the number of members with cycles left
 * afterwards.  A member that fails stops; the others go on. */
int lulesh_batch_step(lulesh_batch *b, int ncycles, int *pending);
/* Steps until no member has cycles left.  Returns LULESH_OK, or the error
 * code of the first failed member. */
int lulesh_batch_run(lulesh_batch *b);

/* A member's scalars; the return value is its error code (LULESH_OK,
 * LULESH_VOLUME_ERROR, LULESH_QSTOP_ERROR) or LULESH_BAD_ARG. */
int lulesh_batch_get_stats(lulesh_batch *b, int member, lulesh_stats *out);
int lulesh_batch_get_field(lulesh_batch *b, int member, int field,
                           double *host_out);

/* Launch shapes: the sweep shape (as lulesh_set_sweep_shape), the number of
 * persistent blocks of the sweep and of the element kernels (0: one block
 * per tile / chunk), and the cycles replayed per captured graph (0: no
 * graph, plain launches).  Results do not depend on any of them. */
int lulesh_batch_set_shape(lulesh_batch *b, int tx, int ty, int nz, int waves,
                           int sweep_blocks, int elem_blocks,
                           int cycles_per_graph);
int lulesh_batch_get_shape(lulesh_batch *b, int *tx, int *ty, int *nz,
                           int *waves, int *sweep_blocks, int *elem_blocks,
                           int *cycles_per_graph);

/* The phases with fused multiply-adds and refined reciprocals, as
 * lulesh_set_fused_multiply_add for a single domain: every member is then
 * bit for bit the fused single-domain run of its configuration.  Default 0
 * (bit-identical to the reference). */
int lulesh_batch_set_fused_multiply_add(lulesh_batch *b, int enable);

/* The autotuner: times trial_cycles cycles of each candidate launch shape
 * (plain launches after one warm-up cycle) from a saved copy of the state,
 * restores the state and keeps the fastest shape with the batch's cycles
 * per graph.  *best_ms (may be NULL) receives its milliseconds per cycle.
 * The default shape is its pick on the 64-member ensemble of nx 16..64. */
int lulesh_batch_autotune(lulesh_batch *b, int trial_cycles, double *best_ms);

/* Milliseconds between the HIP events around the last step/run call. */
double lulesh_batch_last_ms(lulesh_batch *b);
const char *lulesh_batch_last_error(const lulesh_batch *b);

/* ------------------------------------------------------------------------
 * The learned-Q network (no counterpart in the reference): a 12-64-64-2
 * multilayer perceptron with ReLU hidden layers, bf16 weights and
 * activations with fp32 accumulation on the matrix cores, fp32 master
 * weights and Adam for training.  Parameters, fp32, in the order W1[12][64]
 * (input-major), b1[64], W2[64][64], b2[64], W3[64][2], b3[2]
 * (lulesh_mlp_param_count() values).  Inputs are standardised on the device
 * with the given per-input mean and standard deviation.
 * ------------------------------------------------------------------------ */
#define LULESH_MLP_INPUTS 12
#define LULESH_MLP_OUTPUTS 2
typedef struct lulesh_mlp lulesh_mlp;

int lulesh_mlp_param_count(void);
int lulesh_mlp_create(const float *params, const float *in_mean, const float *in_std,
                      lulesh_mlp **out);
void lulesh_mlp_destroy(lulesh_mlp *n);
int lulesh_mlp_get_params(lulesh_mlp *n, float *params);
/* Outputs y[count][2] of the raw inputs x[count][12]. */
int lulesh_mlp_eval(lulesh_mlp *n, const float *x, int count, float *y);
/* Gradient of the loss 0.5/count * sum_s sum_o w_o (y_so - t_so)^2 with
 * respect to the parameters (w = out_weight[2], NULL: 1), and the loss. */
int lulesh_mlp_gradient(lulesh_mlp *n, const float *x, const float *t, int count,
                        const float *out_weight, float *grad, double *loss);
/* lulesh_mlp_gradient without any allocation.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, training loops that keep their samples on the device,
 * harnesses that forbid hipMalloc).  x[count][12], t[count][2], grad[params]
 * and loss[1] (fp32) are DEVICE pointers; out_weight is a host array of 2
 * (NULL: 1, 1).  The kernels are queued on `stream` (a hipStream_t; NULL: the
 * null stream) and not waited for, and the call makes no allocation: the
 * per-wave partial gradients go to a fixed buffer of the library, 5.0 MiB
 * held on each GPU until the process ends, which bounds count to
 * LULESH_MLP_NOALLOC_MAX (a larger count returns LULESH_BAD_ARG).  Every
 * call on one GPU shares that buffer: calls that overlap on one GPU (other
 * host threads, or other streams before this call's kernels finished) race.
 * The current device must be the one the network was created on.  grad and
 * loss are bit-identical to lulesh_mlp_gradient's for the same samples. */
#define LULESH_MLP_NOALLOC_MAX 16384
int lulesh_mlp_gradient_noalloc(lulesh_mlp *n, const float *x, const float *t, int count,
                                const float *out_weight, float *grad, float *loss,
                                void *stream);
/* Adam (beta 0.9 / 0.999, eps 1e-8) on minibatches of `batch` samples drawn
 * without replacement (splitmix64 shuffle from `seed`), `epochs` passes;
 * *final_loss receives the mean minibatch loss of the last pass. */
int lulesh_mlp_train(lulesh_mlp *n, const float *x, const float *t, int count,
                     const float *out_weight, int epochs, int batch, float lr,
                     unsigned long long seed, double *final_loss);
const char *lulesh_mlp_last_error(const lulesh_mlp *n);

/* The learned closure in a domain's Q phase: with net != NULL the linear
 * and quadratic Q of the monotonic limiter (reference/C/LULESH-OMP.cc:
 * 2182-2320) of each compressing element are multiplied by the network's
 * two outputs m_l, m_q (clamped to [0, 3]; 1 is the plain limiter); an
 * element keeps the plain limiter when an input is more than z_max standard
 * deviations from the training mean or its volume / energy is not
 * positive.  net == NULL restores the plain limiter.  The domain uses the
 * network until it is reset: the network must outlive that use.  Float64
 * domains only. */
int lulesh_set_learned_q(lulesh_domain *d, const lulesh_mlp *net, double z_max);
/* The form of the learned closure (no counterpart in the reference); it
 * sets which inputs lulesh_q_features forms and what the network's outputs
 * mean in the Q phase.  LULESH_LEARNED_FACTORS (the default): inputs include
 * the limiter values phi, and the outputs multiply the limiter's ql and qq.
 * LULESH_LEARNED_STANDIN: the network takes the limiter's place; its inputs
 * are the limiter's inputs (the face velocity gradients relative to the
 * element's own, velocity jumps over c, vdov, relative volume, sound speed)
 * and its outputs are the limiting factors themselves,
 * ql = -qlc rho m_l sum(du), qq = qqc rho m_q sum(du^2); the limiter
 * (reference/C/LULESH-OMP.cc:2183-2320) runs only where an element falls
 * back to it.  A network is trained for one form. */
#define LULESH_LEARNED_FACTORS 0
#define LULESH_LEARNED_STANDIN 1
int lulesh_set_learned_q_form(lulesh_domain *d, int form);
/* The network inputs of every element of the current state (features
 * [num_elem][12] of the form set by lulesh_set_learned_q_form, in the order
 * of lulesh_kernels.hip) after forming its
 * kinematics and Q gradients, and q[num_elem][4] = (-rho sum du,
 * rho sum du^2, plain ql, plain qq), the first two 0 for an element without
 * compression.  Overwrites vnew, delv, arealg, vdov and the Q gradients;
 * returns the error code of a non-positive volume, if any. */
int lulesh_q_features(lulesh_domain *d, float *features, double *q);

#ifdef __cplusplus
}
#endif

#endif /* LULESH_MI355X_H */
