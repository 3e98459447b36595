// C-ABI of the differentiable IS-SPA + nonbond model (include/isspa_amd.h):
// forces as a function of per-type parameters theta and per-atom charges q,
// and their adjoint.  The model is built over an IS-SPA model and a nonbond
// model (window layout, MC points, launch tables, exclusions); its kernels
// are the plain field/force kernels with analytic per-type functions plus
// the adjoint kernels of diff_kernels.hip.
#include <algorithm>
#include <new>
#include <vector>

#include "models.hpp"

#define fail isspa_fail

struct isspa_diff {
    isspa_model *im = nullptr;
    nonbond_model *nb = nullptr;
    int n = 0, T = 0;
    uint32_t step = 0;
    double sig_s = 0, eps_s = 0, core = 0, delta = 0;
    double *camp = nullptr;
    DiffType *dtypes = nullptr;
    float4 *nb_ld = nullptr;
    float4 *xq = nullptr, *arec = nullptr;
    double *se = nullptr, *ftmp = nullptr, *gatom = nullptr;
    double *ppart = nullptr, *apart = nullptr;
    float4 *prec = nullptr;
    long long capF = 0, capC = 0;
    // inputs and tiling generation of the last isspa_diff_forces, whose points,
    // records and S_e isspa_diff_vjp_after_forces reuses
    const void *fwd_x = nullptr, *fwd_q = nullptr, *fwd_theta = nullptr;
    unsigned fwd_gen = 0;
    bool fwd_valid = false;
};

static void diff_free(isspa_diff *m)
{
    if (m->im) isspa_model_destroy(m->im);
    if (m->nb) nonbond_model_destroy(m->nb);
    m->im = nullptr;
    m->nb = nullptr;
    dfree(m->camp); dfree(m->dtypes); dfree(m->nb_ld); dfree(m->xq); dfree(m->arec);
    dfree(m->se); dfree(m->ftmp); dfree(m->gatom); dfree(m->ppart); dfree(m->apart);
    dfree(m->prec);
}

// partial buffers follow the IS-SPA model's current tiling
static int diff_buffers(isspa_diff *m)
{
    isspa_model *im = m->im;
    if (im->F > m->capF) {
        dfree(m->ppart);
        CHK(dalloc(&m->ppart, (size_t)10 * im->F));
        m->capF = im->F;
    }
    if (im->C > m->capC) {
        dfree(m->apart);
        CHK(dalloc(&m->apart, (size_t)6 * im->C));
        m->capC = im->C;
    }
    return ISSPA_OK;
}

extern "C" int isspa_diff_create(const isspa_diff_desc *d, const isspa_windows *win,
                                 const int *isspa_type, const float *rmax, const double *camp,
                                 const int *pre, const int *list, int L, isspa_diff **out)
{
    if (!d || !out || !isspa_type || !rmax || !camp || !pre || (!list && L > 0))
        return fail(ISSPA_ERR_ARG, "isspa_diff_create: null argument");
    *out = nullptr;
    const int T = d->n_types;
    if (T < 1 || T > 64)
        return fail(ISSPA_ERR_NOTSUP, "isspa_diff_create: %d types (1..64 supported: the adjoint stages a T x T float4 table in LDS)", T);
    if (!(d->delta > 0) || !(d->eps_s > 0) || !(d->sig_s > 0))
        return fail(ISSPA_ERR_ARG, "isspa_diff_create: delta, sig_s, eps_s must be > 0");
    isspa_diff *m = new (std::nothrow) isspa_diff;
    if (!m) return fail(ISSPA_ERR_ALLOC, "host allocation");
    m->n = d->n_atoms;
    m->T = T;
    m->step = d->step;
    m->sig_s = d->sig_s;
    m->eps_s = d->eps_s;
    m->core = d->core;
    m->delta = d->delta;
    // IS-SPA model over the windows; its tables are unused (2 zero rows)
    isspa_model_desc md{};
    md.n_atoms = d->n_atoms;
    md.n_types = T;
    md.n_mc = d->n_mc;
    md.n_g = md.n_e = md.n_f = 2;
    md.g_r0 = md.e_r0 = md.f_r0 = 0.0f;
    md.g_dr = md.e_dr = md.f_dr = 1.0f;
    md.lbox = d->lbox;
    md.e0 = d->e0; md.p0 = d->p0; md.q0 = d->q0; md.o0 = d->o0; md.ep = d->ep;
    md.seed = d->seed;
    md.system_id = 0;
    std::vector<float> zeros((size_t)T * 2, 0.0f);
    int rc = isspa_model_create_batch(&md, win, isspa_type, rmax, zeros.data(), zeros.data(),
                                      zeros.data(), &m->im);
    // nonbond model: IS-SPA types double as LJ types; its pair table is
    // written from theta on every call
    std::vector<int> nbparm((size_t)T * T);
    for (int a = 0; a < T; ++a)
        for (int b = 0; b < T; ++b) {
            const int lo = std::min(a, b), hi = std::max(a, b);
            nbparm[(size_t)a * T + b] = hi * (hi + 1) / 2 + lo;
        }
    std::vector<float> lj0((size_t)T * (T + 1) / 2, 0.0f);
    if (rc == ISSPA_OK)
        rc = nonbond_model_create_batch(d->n_atoms, T, isspa_type, nbparm.data(), lj0.data(),
                                        lj0.data(), pre, list, L, isspa_type, T, rmax, d->ep,
                                        d->lbox, win, &m->nb);
    const size_t n = (size_t)d->n_atoms;
    if (rc == ISSPA_OK) rc = upload(&m->camp, camp, (size_t)T);
    if (rc == ISSPA_OK) rc = dalloc(&m->dtypes, (size_t)T);
    if (rc == ISSPA_OK) rc = dalloc(&m->nb_ld, (size_t)T * T);
    if (rc == ISSPA_OK) rc = dalloc(&m->xq, n);
    if (rc == ISSPA_OK) rc = dalloc(&m->arec, 4 * n);
    if (rc == ISSPA_OK) rc = dalloc(&m->se, 3 * n);
    if (rc == ISSPA_OK) rc = dalloc(&m->ftmp, 3 * n);
    if (rc == ISSPA_OK) rc = dalloc(&m->gatom, 5 * n);
    if (rc == ISSPA_OK) rc = dalloc(&m->prec, 4 * (size_t)m->im->P);
    if (rc == ISSPA_OK) rc = diff_buffers(m);
    if (rc != ISSPA_OK) {
        diff_free(m);
        delete m;
        return rc;
    }
    *out = m;
    return ISSPA_OK;
}

extern "C" int isspa_diff_destroy(isspa_diff *m)
{
    if (!m) return ISSPA_OK;
    diff_free(m);
    delete m;
    return ISSPA_OK;
}

extern "C" int isspa_diff_set_tiling(isspa_diff *m, int field_splits, int force_chunks)
{
    if (!m) return fail(ISSPA_ERR_ARG, "null model");
    CHK(isspa_model_set_tiling(m->im, field_splits, force_chunks));
    return diff_buffers(m);
}

// theta -> per-type constants and LJ tables; positions/charges (and receiver
// records when lam is given)
static int diff_prepare(isspa_diff *m, const float *xyz4_d, const double *q_d,
                        const double *theta_d, const double *lam_d, hipStream_t s)
{
    DiffPrepArgs pa{};
    pa.T = m->T;
    pa.theta = theta_d;
    pa.camp = m->camp;
    pa.sig_s = m->sig_s;
    pa.eps_s = m->eps_s;
    pa.core = m->core;
    pa.delta = m->delta;
    pa.types = m->dtypes;
    pa.nb_lj = m->nb->lj;
    pa.nb_ld = m->nb_ld;
    HIPCHK(launch_diff_prep_types(pa, s));
    DiffAtomArgs aa{};
    aa.n = m->n;
    aa.xyz = reinterpret_cast<const float4 *>(xyz4_d);
    aa.q = q_d;
    aa.lam = lam_d;
    aa.itype = m->im->itype;
    aa.vtot = m->im->vtot;
    aa.rmax2 = m->im->rmax2;
    aa.xq = m->xq;
    aa.arec = m->arec;
    HIPCHK(launch_diff_prep_atoms(aa, s));
    return ISSPA_OK;
}

// points, field, epilogue, force partials, then F and S_e
static int diff_isspa_forward(isspa_diff *m, const double *q_d, double *f_d, hipStream_t s)
{
    isspa_model *im = m->im;
    HIPCHK(launch_mc_points(im->wins, im->atom_win, m->xq, im->itype, im->rmax, im->rmax2,
                            im->d.n_atoms, im->d.n_mc, im->d.seed, m->step, nullptr, im->pts, s,
                            im->lay.W, im->rmax_all, im->pbc_flag));
    FieldArgs fa = field_args(im, m->xq);
    fa.dtypes = m->dtypes;
    HIPCHK(launch_field(fa, im->ngblk, s));
    // the adjoint reads the per-point g, count and E
    HIPCHK(launch_epilogue(epilogue_args(im, true), s));
    ForceArgs ka = force_args(im, m->xq);
    ka.dtypes = m->dtypes;
    HIPCHK(launch_force(ka, im->npblk, false, s));
    DiffReduceArgs ra{};
    ra.n = m->n;
    ra.wins = im->wins;
    ra.atom_win = im->atom_win;
    ra.part = im->part_f;
    ra.C = im->C;
    ra.itype = im->itype;
    ra.vtot = im->vtot;
    ra.q = q_d;
    ra.se = m->se;
    ra.f = f_d;
    HIPCHK(launch_diff_reduce(ra, s));
    return ISSPA_OK;
}

extern "C" int isspa_diff_forces(isspa_diff *m, const float *xyz4_d, const double *q_d,
                                 const double *theta_d, double *f_d, void *stream)
{
    if (!m || !xyz4_d || !q_d || !theta_d || !f_d)
        return fail(ISSPA_ERR_ARG, "isspa_diff_forces: null argument");
    hipStream_t s = S(stream);
    m->fwd_valid = false;
    CHK(diff_buffers(m));
    CHK(diff_prepare(m, xyz4_d, q_d, theta_d, nullptr, s));
    CHK(diff_isspa_forward(m, q_d, f_d, s));
    CHK(nonbond_run(m->nb, reinterpret_cast<const float *>(m->xq), false, nullptr, f_d, nullptr,
                    nullptr, s));
    m->fwd_x = xyz4_d;
    m->fwd_q = q_d;
    m->fwd_theta = theta_d;
    m->fwd_gen = m->im->generation;
    m->fwd_valid = true;
    return ISSPA_OK;
}

// the adjoint; with run_forward false the points, records and S_e of the last
// isspa_diff_forces on the same inputs are used as they are
static int diff_vjp(isspa_diff *m, const float *xyz4_d, const double *q_d, const double *theta_d,
                    const double *lam_d, double *gq_d, double *gtheta_d, double *gatom_d,
                    bool run_forward, hipStream_t s)
{
    isspa_model *im = m->im;
    nonbond_model *nb = m->nb;
    CHK(diff_buffers(m));
    CHK(diff_prepare(m, xyz4_d, q_d, theta_d, lam_d, s));
    if (run_forward) {
        m->fwd_valid = false;  // ftmp, not the caller's forces, holds this forward
        CHK(diff_isspa_forward(m, q_d, m->ftmp, s));
    }
    VjpArgs va{};
    va.wins = im->wins;
    va.atom_win = im->atom_win;
    va.fblk = im->fblk;
    va.pblk = im->pblk;
    va.nmc = im->d.n_mc;
    va.Ptot = im->P;
    va.pts = im->pts;
    va.rec4 = reinterpret_cast<const float4 *>(im->rec);
    va.aux4 = rec_aux(im);
    va.arec = m->arec;
    va.xq = m->xq;
    va.itype = im->itype;
    va.rmax2 = im->rmax2;
    va.dtypes = m->dtypes;
    va.ntypes = m->T;
    va.pbc_flag = im->pbc_flag;
    va.g_out = im->g_out;
    va.E_out = im->E_out;
    va.cnt_out = im->cnt_out;
    va.e0 = im->d.e0;
    va.p0 = im->d.p0;
    va.q0 = im->d.q0;
    va.o0 = im->d.o0;
    va.ppart = m->ppart;
    va.F = im->F;
    va.prec = m->prec;
    va.apart = m->apart;
    va.C = im->C;
    HIPCHK(launch_vjp_points(va, im->nfblk, s));
    HIPCHK(launch_vjp_epilogue(va, s));
    HIPCHK(launch_vjp_atoms(va, im->npblk, s));
    HIPCHK(launch_pbc_flag(nb->wins, nb->lay.W, m->xq, 0.0f, nb->pbc_flag, s));
    VjpNbArgs na{};
    na.wins = nb->wins;
    na.blocks = nb->blk;
    na.xq = m->xq;
    na.aux = nb->aux;
    na.arec = m->arec;
    na.excl = nb->excl;
    na.excl_ptr = nb->excl_ptr;
    na.excl_idx = nb->excl_idx;
    na.cfar = nb->cfar;
    na.pbc_flag = nb->pbc_flag;
    na.ld = m->nb_ld;
    na.T = m->T;
    na.part = nb->part;
    na.B = nb->B;
    HIPCHK(launch_vjp_nonbond(na, nb->nblk, s));
    VjpFinishArgs fa{};
    fa.n = m->n;
    fa.wins = im->wins;
    fa.nb_wins = nb->wins;
    fa.atom_win = im->atom_win;
    fa.apart = m->apart;
    fa.C = im->C;
    fa.nbpart = nb->part;
    fa.B = nb->B;
    fa.itype = im->itype;
    fa.vtot = im->vtot;
    fa.lam = lam_d;
    fa.se = m->se;
    fa.theta = theta_d;
    fa.gq = gq_d;
    fa.gatom = gatom_d ? gatom_d : m->gatom;
    HIPCHK(launch_vjp_finish(fa, s));
    HIPCHK(launch_type_reduce(fa.gatom, im->itype, m->n, m->T, gtheta_d, s));
    return ISSPA_OK;
}

extern "C" int isspa_diff_vjp(isspa_diff *m, const float *xyz4_d, const double *q_d,
                              const double *theta_d, const double *lam_d, double *gq_d,
                              double *gtheta_d, double *gatom_d, void *stream)
{
    if (!m || !xyz4_d || !q_d || !theta_d || !lam_d || !gq_d || !gtheta_d)
        return fail(ISSPA_ERR_ARG, "isspa_diff_vjp: null argument");
    return diff_vjp(m, xyz4_d, q_d, theta_d, lam_d, gq_d, gtheta_d, gatom_d, true, S(stream));
}

extern "C" int isspa_diff_vjp_after_forces(isspa_diff *m, const float *xyz4_d, const double *q_d,
                                           const double *theta_d, const double *lam_d,
                                           double *gq_d, double *gtheta_d, double *gatom_d,
                                           void *stream)
{
    if (!m || !xyz4_d || !q_d || !theta_d || !lam_d || !gq_d || !gtheta_d)
        return fail(ISSPA_ERR_ARG, "isspa_diff_vjp_after_forces: null argument");
    if (!m->fwd_valid || m->fwd_x != xyz4_d || m->fwd_q != q_d || m->fwd_theta != theta_d ||
        m->fwd_gen != m->im->generation)
        return fail(ISSPA_ERR_STATE,
                    "isspa_diff_vjp_after_forces: the last isspa_diff_forces on this model did "
                    "not use these buffers under the current tiling");
    return diff_vjp(m, xyz4_d, q_d, theta_d, lam_d, gq_d, gtheta_d, gatom_d, false, S(stream));
}
