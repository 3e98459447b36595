// isspa_model_force_noalloc: isspa_model_force with every per-call device
// buffer (MC points, field partials, point records, force chunk
// partials, minimum-image flags) carved from one file-scope device array
// instead of the model's own allocations.  It allocates nothing and has no
// counterpart in the reference; it is for callers that cannot allocate device
// memory (code under stream capture, serving loops, harnesses that forbid
// hipMalloc).
//
// The array holds kNoallocScratchBytes on every GPU the library runs on, until
// the process ends.  Its device address is looked up once per device with
// hipGetSymbolAddress (the array's name in host code is only a stand-in); the
// kernels receive that address.  Every call on one GPU shares the array and
// leaves its contents behind; nothing in it needs to start at zero, because
// each kernel writes every partial that a later kernel reads.  Calls that
// overlap on one GPU (concurrent streams, or host threads) race on it.
#include <mutex>

#include "models.hpp"

#define fail isspa_fail

static constexpr size_t kNoallocScratchBytes = (size_t)ISSPA_NOALLOC_SCRATCH_BYTES;

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_lock;
unsigned char *g_scratch[kMaxDevices] = {};

// the current device's copy of the array
int device_scratch(unsigned char **out)
{
    int dev = 0;
    hipError_t he = hipGetDevice(&dev);
    if (he != hipSuccess) return fail(ISSPA_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(he));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(ISSPA_ERR_NOTSUP, "isspa_model_force_noalloc: device %d beyond %d", dev,
                    kMaxDevices);
    std::lock_guard<std::mutex> g(g_lock);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        he = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (he != hipSuccess || !p)
            return fail(ISSPA_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                        hipGetErrorString(he));
        g_scratch[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch[dev];
    return ISSPA_OK;
}

// consecutive 256-byte aligned regions of the array
struct Carver {
    size_t used = 0;
    template <class T>
    size_t take(size_t count)
    {
        const size_t at = used;
        used = (used + count * sizeof(T) + 255) & ~(size_t)255;
        return at;
    }
};

}  // namespace

extern "C" int isspa_model_force_noalloc(isspa_model *m, const float *xyzq4_d, float *f4_d,
                                         uint32_t step, const uint32_t *step_d, void *stream)
{
    if (!m || !xyzq4_d || !f4_d) return fail(ISSPA_ERR_ARG, "isspa_model_force_noalloc: null argument");
    const size_t P = (size_t)m->P, F = (size_t)m->F, C = (size_t)m->C;
    Carver cv;
    const size_t o_pts = cv.take<float4>(P);
    const size_t o_rec = cv.take<float>(P * (kRecFloats + kAuxFloats));
    const size_t o_pg = cv.take<double>(F);
    const size_t o_pc = cv.take<int>(F);
    const size_t o_ps = cv.take<double>(3 * F);
    const size_t o_pe = cv.take<double>(3 * F);
    const size_t o_pf = cv.take<double>(6 * C);  // the six sums of the total force
    const size_t o_fl = cv.take<int>((size_t)m->lay.W);
    if (cv.used > kNoallocScratchBytes)
        return fail(ISSPA_ERR_ARG,
                    "isspa_model_force_noalloc: this model and tiling need %zu bytes of scratch "
                    "(%zu points, %zu field partials, %zu force partials); the static scratch "
                    "holds %zu bytes: use isspa_model_force, or fewer field splits / force chunks",
                    cv.used, P, F, C, kNoallocScratchBytes);
    unsigned char *base = nullptr;
    CHK(device_scratch(&base));
    float4 *pts = reinterpret_cast<float4 *>(base + o_pts);
    float *rec = reinterpret_cast<float *>(base + o_rec);
    float4 *aux = reinterpret_cast<float4 *>(rec + P * kRecFloats);
    double *part_g = reinterpret_cast<double *>(base + o_pg);
    int *part_cnt = reinterpret_cast<int *>(base + o_pc);
    double *part_s0 = reinterpret_cast<double *>(base + o_ps);
    double *part_ein = reinterpret_cast<double *>(base + o_pe);
    double *part_f = reinterpret_cast<double *>(base + o_pf);
    int *flag = reinterpret_cast<int *>(base + o_fl);

    hipStream_t s = S(stream);
    const float4 *x = reinterpret_cast<const float4 *>(xyzq4_d);
    // the launches of isspa_model_force (isspa_field_pass, isspa_force_pass)
    // with the buffers above
    HIPCHK(launch_mc_points(m->wins, m->atom_win, x, m->itype, m->rmax, m->rmax2, m->d.n_atoms,
                            m->d.n_mc, m->d.seed, step, step_d, pts, s, m->lay.W, m->rmax_all,
                            flag));
    FieldArgs fa = field_args(m, x);
    fa.pts = pts;
    fa.pbc_flag = flag;
    fa.part_g = part_g;
    fa.part_cnt = part_cnt;
    fa.part_s0 = part_s0;
    fa.part_ein = part_ein;
    HIPCHK(launch_field(fa, m->ngblk, s));
    EpilogueArgs ea = epilogue_args(m, false);
    ea.pts = pts;
    ea.part_g = part_g;
    ea.part_cnt = part_cnt;
    ea.part_s0 = part_s0;
    ea.part_ein = part_ein;
    ea.rec = rec;
    ea.aux = aux;
    HIPCHK(launch_epilogue(ea, s));
    ForceArgs fo = force_args(m, x);
    fo.rec4 = reinterpret_cast<const float4 *>(rec);
    fo.aux4 = aux;
    fo.pbc_flag = flag;
    fo.part = part_f;
    HIPCHK(launch_force(fo, m->npblk, false, s));
    ReduceArgs ra = reduce_args(m, x, reinterpret_cast<float4 *>(f4_d), nullptr, nullptr, nullptr);
    ra.part = part_f;
    HIPCHK(launch_force_reduce(ra, false, s));
    return ISSPA_OK;
}
