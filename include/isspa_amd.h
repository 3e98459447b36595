/*
 * isspa_amd.h -- C ABI of the MI355X-native IS-SPA MD engine.
 *
 * Plain C linkage, POD arguments, int status codes (ISSPA_OK = 0), explicit
 * streams (a hipStream_t passed as void*, NULL = the null stream).  No HIP or
 * torch types appear in these signatures.  Every call that takes a stream is
 * asynchronous unless it says otherwise; device pointers (suffix _d) must be
 * device-resident.  Positions are float4 {x, y, z, charge} (the reference's
 * pos_d layout), velocities float4 {vx, vy, vz, dynamic mass} (vel_d), forces
 * float4 {fx, fy, fz, unused} (for_d).  Force entry points ADD into the force
 * array, as the reference's wrappers do; each call reduces its own
 * contribution per atom in FP64 in a fixed order and adds it once.
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the reference repository).
 */
#ifndef ISSPA_AMD_H
#define ISSPA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ISSPA_OK 0
#define ISSPA_ERR_ARG 1     /* invalid argument or inconsistent sizes */
#define ISSPA_ERR_HIP 2     /* HIP runtime error (message has the HIP string) */
#define ISSPA_ERR_ALLOC 3   /* device allocation failed */
#define ISSPA_ERR_STATE 4   /* call out of order (e.g. replay before capture) */
#define ISSPA_ERR_NOTSUP 5  /* unsupported configuration */

/* Message of the last failing call on this host thread ("" if none).  The
 * reference has no error reporting at all (every wrapper returns void or ms). */
const char *isspa_last_error(void);
int isspa_abi_version(void);

/* ------------------------------------------------------------------------- */
/* Device memory and stream helpers, so that host classes need no HIP types.  */
/* They replace the cudaMalloc / cudaMemcpy / cudaMemset / cudaEvent calls in  */
/* the classes' initialize_gpu and copy_* methods                              */
/* (src/atom_class.cpp:175-246, src/isspa_class.cpp:180-214).                  */
/* ------------------------------------------------------------------------- */
int isspa_dev_malloc(void **ptr_d, size_t bytes);
int isspa_dev_free(void *ptr_d);
int isspa_memcpy_htod(void *dst_d, const void *src_h, size_t bytes, void *stream);
int isspa_memcpy_dtoh(void *dst_h, const void *src_d, size_t bytes, void *stream);
int isspa_memcpy_dtod(void *dst_d, const void *src_d, size_t bytes, void *stream);
/* zero-fill with a kernel (safe inside a captured graph) */
int isspa_zero(void *dst_d, size_t bytes, void *stream);
/* dst4_d[i].xyz += src4_d[i].xyz for i < n (float4 arrays); used to fold the
 * separately kept IS-SPA force (the reference's isspaf_d) into for_d */
int isspa_add4(float *dst4_d, const float *src4_d, int n, void *stream);
int isspa_stream_create(void **stream);
int isspa_stream_destroy(void *stream);
int isspa_stream_sync(void *stream);
/* milliseconds between two points of a stream: begin() records, end() syncs */
int isspa_timer_begin(void **timer, void *stream);
int isspa_timer_end(void *timer, void *stream, float *ms);
/* a reusable pair of events: mark the begin and end of some work on a stream
 * (no waiting), read the elapsed milliseconds later (waits for the end mark
 * only), mark again for the next piece of work */
int isspa_events_create(void **pair);
int isspa_events_mark_begin(void *pair, void *stream);
int isspa_events_mark_end(void *pair, void *stream);
int isspa_events_elapsed(void *pair, float *ms);
int isspa_events_destroy(void *pair);

/* ------------------------------------------------------------------------- */
/* Batched windows.  W independent systems (e.g. umbrella-sampling windows,  */
/* each started by the reference as its own process from its own deck,       */
/* src/total_force_cuda.cpp:36-100) share one model and one set of launches:  */
/* their atoms are concatenated in window order, window w owning              */
/* window_atoms[w] consecutive atoms.  Atoms interact only within their       */
/* window.  system_ids[w] (NULL: w) is window w's Philox counter word 3 and   */
/* lbox[w] (NULL: the model's lbox) its box.  A model created without windows */
/* is a batch of one.  The reference has no batched form.                     */
/* ------------------------------------------------------------------------- */
typedef struct {
    int n_windows;
    const int *window_atoms;     /* int[n_windows] */
    const uint32_t *system_ids;  /* uint32[n_windows] or NULL */
    const float *lbox;           /* float[n_windows] or NULL */
} isspa_windows;

/* ------------------------------------------------------------------------- */
/* IS-SPA solvent model: class isspa's device state (src/isspa_class.h:11-68) */
/* plus isspa_grid_block (src/isspa_force_cuda.cu:362-408).                    */
/* ------------------------------------------------------------------------- */
typedef struct isspa_model isspa_model;

typedef struct {
    int n_atoms;     /* N */
    int n_types;     /* IS-SPA types T */
    int n_mc;        /* MC points per atom */
    int n_g, n_e, n_f;             /* rows of ISSPA_DENSITIES / EFIELD / FORCES */
    float g_r0, g_dr;              /* grid r_b = r0 + b*dr of each table */
    float e_r0, e_dr;
    float f_r0, f_dr;
    float lbox;                    /* cubic box length */
    float e0, p0, q0, o0, ep;      /* model constants, see isspa_default_constants */
    uint64_t seed;                 /* Philox key */
    uint32_t system_id;            /* Philox counter word 3 (batched windows) */
} isspa_model_desc;

/* src/constants.h:6-13 evaluated in FP32: e0, p0, q0, o0, ep */
int isspa_default_constants(float *e0, float *p0, float *q0, float *o0, float *ep);

/* Host arrays: isspa_type int[N] (0-based), rmax float[T], g float[T*n_g],
 * e float[T*n_e], f float[T*n_f] ([type][row], the layout
 * read_isspa_prmtop builds, src/isspa_class.cpp:193-241).  vtot[t] =
 * 4/3 pi rmax^3 0.0074 / n_mc is computed here in double and stored in float
 * (src/isspa_class.cpp:103-113).  Requires n_e >= n_g (the E table is read at
 * the G-table bin). */
int isspa_model_create(const isspa_model_desc *desc, const int *isspa_type, const float *rmax,
                       const float *g, const float *e, const float *f, isspa_model **out);
/* The same over windows: desc->n_atoms is the total, isspa_type is indexed by
 * the concatenated atoms, the tables are shared (concatenate the windows'
 * types).  Every window's points, fields and forces are those of the window
 * alone with system_id = system_ids[w]. */
int isspa_model_create_batch(const isspa_model_desc *desc, const isspa_windows *windows,
                             const int *isspa_type, const float *rmax, const float *g,
                             const float *e, const float *f, isspa_model **out);
int isspa_model_destroy(isspa_model *m);

/* isspa_force_cuda (src/isspa_force_cuda.cu:331-360): MC points for `step`,
 * field pass, per-point epilogue, force pass; f4_d[i].xyz += F_i.
 * If step_d is non-NULL the step is read from that device word instead (graph
 * replay).  The reference's isspaf debug output (always zero there, Q17) is
 * replaced by isspa_model_force_components. */
int isspa_model_force(isspa_model *m, const float *xyzq4_d, float *f4_d, uint32_t step,
                      const uint32_t *step_d, void *stream);
/* The same as isspa_model_force, allocating nothing: its per-call buffers
 * come from a static device array of ISSPA_NOALLOC_SCRATCH_BYTES per GPU
 * (held until the process ends).  No counterpart in the reference: for
 * callers that cannot allocate device memory (stream capture, serving loops,
 * harnesses that forbid hipMalloc).  ISSPA_ERR_ARG if the model's size and
 * tiling need more scratch than the array holds (see the message).  All calls
 * on one GPU share the array: calls that overlap on one GPU race. */
#define ISSPA_NOALLOC_SCRATCH_BYTES (64u << 20)
int isspa_model_force_noalloc(isspa_model *m, const float *xyzq4_d, float *f4_d, uint32_t step,
                              const uint32_t *step_d, void *stream);
/* the same, and the IS-SPA force alone also stored (overwritten, w = 0) in
 * f_isspa4_d by the same reduction kernel. Fills the isspaf_d argument of
 * src/isspa_force_cuda.cu:331-360, which the reference zeroes
 * (src/isspa_force_cuda.cu:343) and never accumulates. */
int isspa_model_force_keep(isspa_model *m, const float *xyzq4_d, float *f4_d, float *f_isspa4_d,
                           uint32_t step, const uint32_t *step_d, void *stream);
/* Per-component forces in FP64 (double[N*3] each, overwritten): LJ-table term,
 * polarisation (dipole + quadrupole + octupole), constant-density dielectric.
 * These are the components of the reference's statistical fixtures. */
int isspa_model_force_components(isspa_model *m, const float *xyzq4_d, double *f_lj_d,
                                 double *f_pol_d, double *f_cdd_d, uint32_t step, void *stream);
/* Internal per-step buffers for inspection, copied to device buffers:
 * pts4_d float4[P] (.w = owner rmax^2) from the model's last force call of any
 * kind; g_d double[P], cnt_d int[P], E_d double[P*3], E0_d double[P*3] from
 * its last isspa_model_force_components call (the force paths skip writing
 * them).  Any pointer may be NULL. */
int isspa_model_get_field(isspa_model *m, float *pts4_d, double *g_d, int *cnt_d, double *E_d,
                          double *E0_d, void *stream);
/* Stage timing for tuning and roofline accounting (no reference counterpart;
 * the reference times whole wrappers with events, src/timing_class.cpp):
 * runs each stage once untimed and then `reps` times back to back between
 * two events on `stream`, and returns the mean milliseconds per launch in
 * ms[5] = {points, field, epilogue, force, force reduction}.  Uses the
 * positions xyzq4_d; leaves caller buffers untouched.  Synchronous. */
int isspa_model_profile(isspa_model *m, const float *xyzq4_d, int reps, float *ms, void *stream);
/* Tuning knobs (0 = automatic): atom splits of the field pass, point chunks of
 * the force pass, applied to every window.  get_tiling reports the largest
 * window's.  The automatic choice aims at no more than `waves` wavefronts
 * per pass, shared among windows in proportion to their work; without a
 * set_target_waves call (or after one with waves = 0) that is one wave per 5
 * MC points, between 2048 and 131072.  A split count divisible by 4 (or 2)
 * lets each field block sum 4 (or 2) splits itself, so the epilogue reads
 * that many times fewer partials. */
int isspa_model_set_tiling(isspa_model *m, int field_splits, int force_chunks);
int isspa_model_get_tiling(const isspa_model *m, int *field_splits, int *force_chunks);
int isspa_model_set_target_waves(isspa_model *m, int waves);

/* ------------------------------------------------------------------------- */
/* All-pairs nonbond: nonbond_force_cuda + nonbond_force_cuda_grid_block      */
/* (src/nonbond_force_cuda.cu:178-222, kernel :31-174).                        */
/* ------------------------------------------------------------------------- */
typedef struct nonbond_model nonbond_model;
/* ityp int[N] (0-based LJ types), nbparm int[T*T] (0-based), lj_a/lj_b
 * float[T(T+1)/2], nexcl_prefix int[N] (inclusive prefix sum), excl_list
 * int[L] (0-based; negative placeholders allowed), isspa_type int[N],
 * rmax float[T_isspa].  The exclusion list is decoded once on the host with
 * the reference's scan (lower-index atom's slice, early exit on a larger
 * entry) and stored as per-atom bit windows, so no list is staged in LDS per
 * block (Q11).  ep is the solvent dielectric (src/constants.h:7).  At most 90
 * LJ types (the T x T pair table is staged in 64 KB of LDS). */
int nonbond_model_create(int n_atoms, int n_lj_types, const int *ityp, const int *nbparm,
                         const float *lj_a, const float *lj_b, const int *nexcl_prefix,
                         const int *excl_list, int excl_len, const int *isspa_type,
                         int n_isspa_types, const float *rmax, float ep, float lbox,
                         nonbond_model **out);
/* The same over windows (pairs only within a window; exclusions must not
 * cross windows). */
int nonbond_model_create_batch(int n_atoms, int n_lj_types, const int *ityp, const int *nbparm,
                               const float *lj_a, const float *lj_b, const int *nexcl_prefix,
                               const int *excl_list, int excl_len, const int *isspa_type,
                               int n_isspa_types, const float *rmax, float ep, float lbox,
                               const isspa_windows *windows, nonbond_model **out);
int nonbond_model_destroy(nonbond_model *m);
/* f4_d float4[N] += the total all-pairs force.  Windows of 4,096 atoms or more
 * (with a symmetric LJ pair table and every exclusion within 64 indices)
 * evaluate each unordered pair once and add it to both atoms; the result
 * agrees with the sum of the components below to FP32 rounding and is the
 * same bit for bit from run to run. */
int nonbond_force(nonbond_model *m, const float *xyzq4_d, float *f4_d, void *stream);
/* components in FP64 (double[N*3] each, overwritten): LJ, Coulomb, IS-SPA pair */
int nonbond_force_components(nonbond_model *m, const float *xyzq4_d, double *f_lj_d,
                             double *f_coul_d, double *f_pair_d, void *stream);

/* ------------------------------------------------------------------------- */
/* Bonded terms: bond_force_cuda (src/bond_force_cuda.cu:58-79),              */
/* angle_force_cuda (src/angle_force_cuda.cu:158-175),                         */
/* dih_force_cuda (src/dih_force_cuda.cu:140-158) and their *_grid_block.      */
/* ------------------------------------------------------------------------- */
typedef struct bonded_model bonded_model;
/* bond_atoms int[nb*4] {i, j, type, -}, bond_params float[nbt*2] {2k, r0};
 * angle_atoms int[na*4] {i, j, k, type}, angle_params float[nat*2] {2k, theta0};
 * dih_atoms int[nd*4] {i, j, +-k, +-l}, dih_types int[nd],
 * dih_params float[ndt*4] {n, V, phase, -}, sc14 float[ndt*2] {scee, scnb};
 * LJ tables as for nonbond_model_create (used by the scaled 1-4 pairs). */
int bonded_model_create(int n_atoms, int nb, const int *bond_atoms, int nbt,
                        const float *bond_params, int na, const int *angle_atoms, int nat,
                        const float *angle_params, int nd, const int *dih_atoms,
                        const int *dih_types, int ndt, const float *dih_params,
                        const float *sc14, int n_lj_types, const int *ityp, const int *nbparm,
                        const float *lj_a, const float *lj_b, float lbox, bonded_model **out);
int bonded_model_destroy(bonded_model *m);
#define ISSPA_TERM_BOND 1
#define ISSPA_TERM_ANGLE 2
#define ISSPA_TERM_DIH 4
#define ISSPA_TERM_ALL 7
/* f4_d += forces of the selected terms (mask of ISSPA_TERM_*); energies_d, if
 * not NULL, receives double[5]: bond, angle, torsion, 1-4 Coulomb, 1-4 LJ. */
int bonded_force(bonded_model *m, int terms, const float *xyzq4_d, float *f4_d,
                 double *energies_d, void *stream);

/* ------------------------------------------------------------------------- */
/* Differentiable IS-SPA + nonbond forces for force matching.  No reference   */
/* counterpart: the reference's force path (src/isspa_force_cuda.cu:110-327,  */
/* src/nonbond_force_cuda.cu:31-174) with its per-type tables replaced by the */
/* analytic functions they are sampled from, differentiable in              */
/* theta_t = (sigma, eps, x0, g0, alpha) per type and the charge per atom:    */
/*   sigma_ts = (sigma + sig_s)/2, eps_ts = sqrt(eps eps_s)                    */
/*   r_lo = max(x0 - sqrt(g0/alpha), core sigma_ts), r_hi = x0 + sqrt((g0-1)/alpha) */
/*   S(d) = smoothstep((d - r_lo)/delta + 1)                                  */
/*   g(d) = S(d) max(0, g0 - alpha (d - x0)^2) (d < r_hi), 1 beyond            */
/*   E(d) = camp exp(-(d - x0)^2/2) S(d)                                      */
/*   F(d) = 24 eps_ts/d (2 (sigma_ts/d)^12 - (sigma_ts/d)^6) S(d)              */
/*   nonbond A_ij = 4 eps_ij sigma_ij^12, B_ij = 4 eps_ij sigma_ij^6 (Lorentz-Berthelot) */
/* IS-SPA types double as LJ types.  MC points are drawn once per call at the */
/* fixed `step` (common random numbers: they depend on positions and rmax    */
/* only).  Computed in FP32 with FP64 accumulation.                           */
/* ------------------------------------------------------------------------- */
typedef struct isspa_diff isspa_diff;
typedef struct {
    int n_atoms, n_types, n_mc;
    float lbox;
    float e0, p0, q0, o0, ep;        /* as isspa_default_constants */
    double sig_s, eps_s, core, delta; /* solvent sphere, core floor, switch width */
    uint64_t seed;
    uint32_t step;
} isspa_diff_desc;
/* isspa_type int[N] (0-based; also the LJ type), rmax float[T], camp double[T]
 * (E-field amplitudes, fixed), exclusions as for nonbond_model_create.  At
 * most 64 types.  `windows` may be NULL (one system). */
int isspa_diff_create(const isspa_diff_desc *desc, const isspa_windows *windows,
                      const int *isspa_type, const float *rmax, const double *camp,
                      const int *nexcl_prefix, const int *excl_list, int excl_len,
                      isspa_diff **out);
int isspa_diff_destroy(isspa_diff *m);
int isspa_diff_set_tiling(isspa_diff *m, int field_splits, int force_chunks);
/* f_d double[N*3] = F_isspa + F_nonbond (overwritten) for positions xyz4_d
 * (float4, .w ignored), charges q_d double[N], theta_d double[T*5] */
int isspa_diff_forces(isspa_diff *m, const float *xyz4_d, const double *q_d,
                      const double *theta_d, double *f_d, void *stream);
/* adjoint for lam_d = dL/dF (double[N*3]): gq_d double[N] = dL/dq,
 * gtheta_d double[T*5] = dL/dtheta (FP64 per-type sums in a fixed order),
 * gatom_d double[N*5] per-atom contributions (may be NULL) */
int isspa_diff_vjp(isspa_diff *m, const float *xyz4_d, const double *q_d,
                   const double *theta_d, const double *lam_d, double *gq_d, double *gtheta_d,
                   double *gatom_d, void *stream);
/* the same without re-running the forward: uses the MC points, per-point
 * records and S_e of the last isspa_diff_forces on this model, which must have
 * been given the same xyz4_d, q_d and theta_d buffers (the caller guarantees
 * their contents are unchanged) under the current tiling; ISSPA_ERR_STATE
 * otherwise.  A plain isspa_diff_vjp in between invalidates that state. */
int isspa_diff_vjp_after_forces(isspa_diff *m, const float *xyz4_d, const double *q_d,
                                const double *theta_d, const double *lam_d, double *gq_d,
                                double *gtheta_d, double *gatom_d, void *stream);

/* ------------------------------------------------------------------------- */
/* Leap-frog + Andersen: leapfrog_cuda (src/leapfrog_cuda.cu:73-86, kernel    */
/* :33-67).  Philox keyed by (seed, step, atom) replaces the curand states     */
/* (init_rand_states, src/init_rand.cu:10-30).  kT in kcal/mol, dt in AMBER    */
/* time units; pnu is the collision probability per step.  Positions are      */
/* wrapped into [0, lbox).                                                    */
/* ------------------------------------------------------------------------- */
int leapfrog_andersen(int n_atoms, float *xyzq4_d, float *vel4_d, const float *f4_d, float kT,
                      float dt, float pnu, float lbox, uint64_t seed, uint32_t system_id,
                      uint32_t step, const uint32_t *step_d, void *stream);

/* ------------------------------------------------------------------------- */
/* Umbrella restraint: us_force_cuda + us_grid_block                          */
/* (src/us_force_cuda.cu:313-343, kernel :14-47), us::populate_mass           */
/* (src/us_class.cpp:136-157).                                                 */
/* ------------------------------------------------------------------------- */
typedef struct us_model us_model;
/* atom_list int[nbias*2] {atom, group 0/1} (atoms unique), mass float[nbias]
 * normalised by the group mass, kumb float[2] = {-k, +k}, x0 the centre. */
int us_model_create(int n_atoms, int nbias, const int *atom_list, const float *mass,
                    const float *kumb, float x0, float lbox, us_model **out);
/* n_res restraints (one per umbrella window): restraint r owns bias entries
 * [bias_offsets[r], bias_offsets[r+1]) of atom_list/mass (global atom indices),
 * force constant k[r] (group 0 gets -k, group 1 +k), centre x0[r], box lbox[r]. */
int us_model_create_batch(int n_atoms, int n_res, const int *bias_offsets, const int *atom_list,
                          const float *mass, const float *k, const float *x0, const float *lbox,
                          us_model **out);
int us_model_destroy(us_model *m);
/* f4_d += restraint forces; com_d, if not NULL, receives double[6 n_res] (COM
 * of group 0 then group 1 per restraint; the CV is |min_image(com0 - com1)|,
 * us::print_cv). */
int us_force(us_model *m, const float *xyzq4_d, float *f4_d, double *com_d, void *stream);

/* ------------------------------------------------------------------------- */
/* Whole MD step (src/total_force_cuda.cpp:102-154) captured in a hipGraph.   */
/* ------------------------------------------------------------------------- */
typedef struct md_engine md_engine;
/* Components may be NULL (skipped).  The engine does not own them.  Buffers
 * are the caller's device arrays; the step counter lives in the engine.  The
 * integrator uses the IS-SPA (else nonbond) model's windows: system_ids and
 * boxes per window; lbox/system_id apply when neither model is given. */
int md_engine_create(isspa_model *isspa, nonbond_model *nb, bonded_model *bonded, us_model *us,
                     int n_atoms, float *xyzq4_d, float *vel4_d, float *f4_d, float kT,
                     float dt, float pnu, float lbox, uint64_t seed, uint32_t system_id,
                     md_engine **out);
int md_engine_destroy(md_engine *e);
int md_engine_set_step(md_engine *e, uint32_t step, void *stream);
/* one step eagerly (same kernels as the graph): forces then integration */
int md_engine_step(md_engine *e, void *stream);
/* capture `steps_per_graph` consecutive steps into a graph (done once) */
int md_engine_capture(md_engine *e, int steps_per_graph, void *stream);
/* replay the captured graph `replays` times; ISSPA_ERR_STATE if the IS-SPA
 * model's tiling was changed after the capture (capture again) */
int md_engine_replay(md_engine *e, int replays, void *stream);
/* forces only (zero, bonded, IS-SPA, nonbond, US) for the current step */
int md_engine_forces(md_engine *e, void *stream);
/* the restraint centres of mass of the last force evaluation, double[6 n_res] */
int md_engine_get_com(md_engine *e, double *com_d, void *stream);

#ifdef __cplusplus
}
#endif
#endif
