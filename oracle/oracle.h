/*
 * oracle.h -- CPU restatement of the IS-SPA MD hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the parity oracle for the HIP path.  Only tests/, the smoke
 * entry point and bench.py's cpu_baseline leg may load it; the product path
 * never calls it.
 *
 * Arithmetic: FP64 throughout.  Inputs that the device holds in FP32
 * (positions, charges, tables, rmax, vtot) are taken as FP32 and widened, and
 * the three discontinuous decisions (in-sphere test of the field pass, LJ-table
 * versus constant-density-dielectric test of the force pass, and minimum-image
 * shifts) are taken on FP32 values computed with exactly the operation order
 * the device uses, so that both sides take the same branch on every pair.
 *
 * Reference files followed (paths relative to the reference repository):
 *   src/isspa_force_cuda.cu:73-327     MC points, field pass, force pass
 *   src/isspa_class.cpp:103-113        vtot normalisation
 *   src/constants.h:6-13               model constants
 *   src/cuda_vector_routines.h:28-68   min_image / wrap
 *   src/nonbond_force_cuda.cu:31-174   all-pairs LJ + Coulomb + IS-SPA pair term
 *   src/bond_force_cuda.cu:15-53       harmonic bonds
 *   src/angle_force_cuda.cu:12-64      harmonic angles
 *   src/dih_force_cuda.cu:13-136       periodic dihedrals + scaled 1-4
 *   src/leapfrog_cuda.cu:33-67         leap-frog + Andersen
 *   src/us_force_cuda.cu:14-47         umbrella COM-distance restraint
 *   src/isspa_force.omp.c:11-105       OpenMP CPU IS-SPA (baseline restatement)
 */
#ifndef ISSPA_ORACLE_H
#define ISSPA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Philox4x32-10 (Salmon et al., SC'11) ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- IS-SPA model constants as the device uses them (FP32 values) ---- */
typedef struct {
    float e0, p0, q0, o0, ep;
} orc_consts;
void orc_default_consts(orc_consts *c);

/* Tables: flat [type][row] arrays of n rows on the grid r_b = r0 + b*dr. */
typedef struct {
    int n_types;
    int n_g, n_e, n_f;
    float g_r0, g_dr, e_r0, e_dr, f_r0, f_dr;
    const float *g;    /* [n_types*n_g] */
    const float *e;    /* [n_types*n_e] */
    const float *f;    /* [n_types*n_f] */
    const float *rmax; /* [n_types] */
    const float *vtot; /* [n_types] */
} orc_tables;

/* A.1: MC points (FP32, bit-identical with the device). pts is float4[P]. */
void orc_mc_points(int n, int nmc, const float *xyzq, const int *itype, const float *rmax,
                   uint64_t seed, uint32_t step, uint32_t sys, float *pts);

/* A.2: field pass.  Outputs per point: g (product), cnt, E (3), E0 (3). */
void orc_isspa_field(int n, int nmc, const float *xyzq, const int *itype, const orc_tables *t,
                     const float *pts, float lbox, const orc_consts *c,
                     double *g, int *cnt, double *E, double *E0);

/* A.2 on P points with explicit owners (owner == NULL: owner of m is m / nmc). */
void orc_isspa_field_subset(int n, long P, int nmc, const int *owner, const float *xyzq,
                            const int *itype, const orc_tables *t, const float *pts, float lbox,
                            const orc_consts *c, double *g, int *cnt, double *E, double *E0);

/* A.3: force pass on receivers [i0,i1).  Per-component outputs (3 per atom,
 * indexed by the absolute atom index): LJ-table term, polarisation
 * (dipole+quadrupole+octupole) term, constant-density-dielectric term. */
void orc_isspa_force(int n, int nmc, const float *xyzq, const int *itype, const orc_tables *t,
                     const float *pts, const double *g, const int *cnt, const double *E,
                     const double *E0, float lbox, const orc_consts *c, int i0, int i1,
                     double *f_lj, double *f_pol, double *f_cdd);

/* Langevin moments <P1>,<P2>,<P3> of a unit dipole in field x (FP64). */
void orc_langevin_moments(double x, double m[3]);

/* Positions of the nonbond, bonded and umbrella terms are taken in FP64
 * (xyzq as double[4N]) so the golden vectors pin the formulas exactly. */
/* A.4: all-pairs nonbond.  Components: LJ, Coulomb (both on non-excluded
 * pairs) and the IS-SPA pair electrostatics on all pairs.  Energies of the LJ
 * and Coulomb parts are returned in energy[0], energy[1] (each unordered pair
 * once). */
void orc_nonbond(int n, const double *xyzq, int n_lj_types, const int *ityp, const int *nbparm,
                 const double *lj_a, const double *lj_b, const int *nexcl_prefix,
                 const int *excl_list, const int *isspa_type, const float *rmax, float lbox,
                 const orc_consts *c, double *f_lj, double *f_coul, double *f_pair,
                 double *energy);

/* A.5 bonded terms.  bond_atoms: int4 {i, j, type, 0}; bond_params: float2 {2k, r0}. */
void orc_bonds(int nb, const int *bond_atoms, const float *bond_params, const double *xyzq,
               float lbox, double *f, double *energy);
/* angle_atoms: int4 {i, j, k, type}; angle_params: float2 {2k, theta0}. */
void orc_angles(int na, const int *angle_atoms, const float *angle_params, const double *xyzq,
                float lbox, double *f, double *energy);
/* dih_atoms: int4 signed {i, j, +-k, +-l}; dih_types: int; dih_params: float4 {n, V, phase, 0};
 * sc14: float2 {scee, scnb}.  energy[0] torsion, energy[1] 1-4 Coulomb, energy[2] 1-4 LJ. */
void orc_dihedrals(int nd, const int *dih_atoms, const int *dih_types, const float *dih_params,
                   const float *sc14, const double *xyzq, int n_lj_types, const int *ityp,
                   const int *nbparm, const double *lj_a, const double *lj_b, float lbox,
                   double *f, double *energy);

/* A.15: leap-frog + Andersen for one step.  pos/vel are float4[n] (vel.w is
 * the dynamic mass); force is double[3n].  Outputs in double. */
void orc_leapfrog(int n, const float *pos, const float *vel, const double *force, float T,
                  float dt, float pnu, float lbox, uint64_t seed, uint32_t step, uint32_t sys,
                  double *pos_out, double *vel_out, int *collided);

/* A.16: umbrella restraint.  atom_list: int2 {atom, group}; mass: float (normalised by
 * group mass); kumb: {-k, +k}.  com: double[6] output; energy = 0.5 k (d - x0)^2. */
void orc_us_force(int nbias, const int *atom_list, const float *mass, const float *kumb,
                  float x0, const double *xyzq, float lbox, double *f, double *com,
                  double *energy);

/* Restatement of cpu_ispa_force (src/isspa_force.omp.c:11-105): same rand()
 * call order and the same OpenMP loop, so at one thread after srand(s) it
 * reproduces the original bit for bit. */
void orc_cpu_ispa_force(float *xyz, float *f, float *w, float *x0, float *g0, float *gr2,
                        float *alpha, float *lj_A, float *lj_B, int *ityp, int nAtoms, int nMC,
                        int nThreads);
void orc_srand(unsigned int s);

/* ---- differentiable model (oracle_diff.c) ----
 * Per-type trainable parameters (sigma, eps, x0, g0, alpha) and fixed
 * constants of the analytic IS-SPA + nonbond model. */
typedef struct {
    int n_types;
    const double *sigma, *eps, *x0, *g0, *alpha; /* [T] trainable */
    const double *camp;                          /* [T] E-field amplitude */
    const float *rmax;                           /* [T] sampling radius (FP32, as the device) */
    const double *vtot;                          /* [T] MC normalisation */
    double sig_s, eps_s, core, delta;            /* solvent sphere, core floor, switch width */
    double e0, p0, q0, o0, ep;
} orc_dparams;

/* val = {g, E, F}(d) of type t; dval[5k+p] = d val[k] / d theta_p */
void orc_diff_type_fn(const orc_dparams *P, int t, double d, double val[3], double dval[15]);
/* Langevin moments and their derivatives */
void orc_langevin_derivs(double x, double m[3], double dm[3]);
/* forces of the analytic model: f_isspa, f_nb double[N*3]; psi (optional)
 * double[P*8] = E(3), E0(3), g, cnt per point.  q double[N] are the charges
 * (the .w of xyz4 is ignored); pre/list the exclusions. */
void orc_diff_forward(int n, int nmc, const float *xyz4, const double *q, const int *itype,
                      const float *pts, float lbox, const orc_dparams *P, const int *pre,
                      const int *list, double *f_isspa, double *f_nb, double *psi);
/* adjoint: lam = dL/dF (F = f_isspa + f_nb) -> gq[N] = dL/dq, gt[T*5] = dL/dtheta */
void orc_diff_vjp(int n, int nmc, const float *xyz4, const double *q, const int *itype,
                  const float *pts, float lbox, const orc_dparams *P, const int *pre,
                  const int *list, const double *lam, double *gq, double *gt);

#ifdef __cplusplus
}
#endif
#endif
