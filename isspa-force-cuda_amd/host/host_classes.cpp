// Host classes: allocation, file readers/writers, device copies through the
// C ABI.  Methods keep the reference's names and meaning
// (src/atom_class.cpp, src/isspa_class.cpp, src/us_class.cpp,
// src/config_class.cpp, src/timing_class.cpp, src/{bond,angle,dih}_class.cpp).
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/isspa_amd.h"
#include "force_calls.h"

static double now_sec()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static char *trim(char *s)
{
    while (*s == ' ' || *s == '\t' || *s == '\n' || *s == '\r') ++s;
    char *e = s + strlen(s);
    while (e > s && (e[-1] == ' ' || e[-1] == '\t' || e[-1] == '\n' || e[-1] == '\r')) *--e = 0;
    return s;
}

template <class T>
static T *dev_alloc(size_t n)
{
    void *p = nullptr;
    isspa_check(isspa_dev_malloc(&p, n * sizeof(T)), "isspa_dev_malloc");
    return static_cast<T *>(p);
}

// ---------------------------------------------------------------- atom -----
void atom::allocate()
{
    pos_h = new float4[nAtoms]();
    vel_h = new float4[nAtoms]();
    mass_h = new float4[nAtoms]();
    for_h = new float4[nAtoms]();
    isspaf_h = new float4[nAtoms]();
    ityp_h = new int[nAtoms]();
    nExcludedAtoms_h = new int[nAtoms]();
    excludedAtomsList_h = new int[excludedAtomsListLength > 0 ? excludedAtomsListLength : 1]();
    nonBondedParmIndex_h = new int[nTypes * nTypes]();
    lj_h = new float2[nTypes * (nTypes + 1) / 2]();
    nPairs = nAtoms * nAtoms;
}

void atom::allocate_molecule_arrays() { molPointer_h = new int[nMols > 0 ? nMols : 1](); }

// Box-Muller pair by Marsaglia's polar method on rand(), src/atom_class.cpp:158-173
float atom::rand_gauss()
{
    float v1, v2, r2;
    do {
        v1 = 1.0f - 2.0f * (float)rand() / (float)RAND_MAX;
        v2 = 1.0f - 2.0f * (float)rand() / (float)RAND_MAX;
        r2 = v1 * v1 + v2 * v2;
    } while (r2 > 1.0f || r2 == 0.0f);
    return v1 * (float)sqrt(-2.0 * log(r2) / r2);
}

void atom::initialize_velocities(float T)
{
    for (int i = 0; i < nAtoms; ++i) {
        const float s = sqrtf(T / vel_h[i].w);
        vel_h[i].x = rand_gauss() * s;
        vel_h[i].y = rand_gauss() * s;
        vel_h[i].z = rand_gauss() * s;
    }
}

// Exact restart.  The AMBER text restart holds 7 decimals per value
// (src/atom_class.cpp:292-327), which does not round-trip a float velocity of
// order 0.1.  Beside each restart file the writer keeps "<file>.exact": a
// 16-byte header {"ISSPARST", n, step} and the n x 3 float values as written.
// A resumed run (velRst = 1) reads them when they agree with the text values,
// and continues the Philox counter at `step`, so it repeats bit for bit what
// the uninterrupted run would have done.  Without the file the text values and
// the step in the title line (if any) are used.
struct ExactHeader {
    char magic[8];
    int32_t n;
    uint32_t step;
};
static const char kExactMagic[8] = {'I', 'S', 'S', 'P', 'A', 'R', 'S', 'T'};

static void exact_path(const char *path, char *out, size_t len)
{
    snprintf(out, len, "%s.exact", path);
}

// text values replaced by the exact ones when the side file belongs to them
static bool read_exact(const char *path, int nAtoms, std::vector<float> &vals, unsigned *step)
{
    char ep[MAXCHAR + 8];
    exact_path(path, ep, sizeof ep);
    FILE *fh = fopen(ep, "rb");
    if (!fh) return false;
    ExactHeader h{};
    std::vector<float> ex(3 * (size_t)nAtoms);
    const bool ok = fread(&h, sizeof h, 1, fh) == 1 && !memcmp(h.magic, kExactMagic, 8) &&
                    h.n == nAtoms && fread(ex.data(), sizeof(float), ex.size(), fh) == ex.size();
    fclose(fh);
    if (!ok) return false;
    for (size_t k = 0; k < ex.size(); ++k)
        if (!(fabsf(ex[k] - vals[k]) <= 1e-6f * std::max(1.0f, fabsf(ex[k])))) {
            fprintf(stderr, "%s does not match %s (value %zu: %g vs %g); using the text values\n",
                    ep, path, k, ex[k], vals[k]);
            return false;
        }
    vals = ex;
    *step = h.step;
    return true;
}

// AMBER restart: two header lines, then 6 fields of width 12 per line
static int read_rst7_into(const char *path, int nAtoms, float4 *dst, const char *what,
                          unsigned *step)
{
    FILE *fh = fopen(path, "r");
    if (!fh) {
        fprintf(stderr, "Could not find %s file %s\n", what, path);
        return 1;
    }
    char line[MAXCHAR];
    std::vector<float> vals;
    int header = 0;
    *step = 0;
    while (fgets(line, MAXCHAR, fh) && (int)vals.size() < 3 * nAtoms) {
        if (header < 2) {
            // this writer's title ends in "step <n>"
            const char *at = header == 0 ? strstr(line, " step ") : nullptr;
            if (at) *step = (unsigned)strtoul(at + 6, nullptr, 10);
            ++header;
            continue;
        }
        const size_t len = strlen(line);
        for (int c = 0; c < 6 && (int)vals.size() < 3 * nAtoms; ++c) {
            if ((size_t)(c * 12) >= len) break;
            char tok[13] = {0};
            memcpy(tok, line + c * 12, std::min<size_t>(12, len - (size_t)(c * 12)));
            vals.push_back((float)atof(tok));
        }
    }
    fclose(fh);
    if ((int)vals.size() < 3 * nAtoms) {
        fprintf(stderr, "%s file %s: expected %d values, found %zu\n", what, path, 3 * nAtoms,
                vals.size());
        return 2;
    }
    read_exact(path, nAtoms, vals, step);
    for (int i = 0; i < nAtoms; ++i) {
        dst[i].x = vals[3 * i];
        dst[i].y = vals[3 * i + 1];
        dst[i].z = vals[3 * i + 2];
    }
    return 0;
}

void atom::read_initial_positions(const char *f)
{
    if (read_rst7_into(f, nAtoms, pos_h, "coordinate", &posRstStep)) exit(1);
}

void atom::read_initial_velocities(const char *f)
{
    if (read_rst7_into(f, nAtoms, vel_h, "velocity", &velRstStep)) exit(2);
}

void atom::open_traj_files(const char *forName, const char *posName, const char *velName)
{
    forFile = fopen(forName, "w");
    posFile = fopen(posName, "w");
    velFile = fopen(velName, "w");
    IFFile = fopen("ISSPA_force.xyz", "w");
}

void atom::initialize_gpu(int s)
{
    seed = s;
    pos_d = dev_alloc<float4>(nAtoms);
    vel_d = dev_alloc<float4>(nAtoms);
    for_d = dev_alloc<float4>(nAtoms);
    isspaf_d = dev_alloc<float4>(nAtoms);
    isspa_check(isspa_zero(for_d, nAtoms * sizeof(float4), nullptr), "isspa_zero");
    isspa_check(isspa_zero(isspaf_d, nAtoms * sizeof(float4), nullptr), "isspa_zero");
}

void atom::copy_params_to_gpu()
{
    // parameters reach the device when the nonbond and bonded models are built
}

void atom::copy_pos_vel_to_gpu()
{
    isspa_check(isspa_memcpy_htod(vel_d, vel_h, nAtoms * sizeof(float4), nullptr), "htod");
    isspa_check(isspa_memcpy_htod(pos_d, pos_h, nAtoms * sizeof(float4), nullptr), "htod");
    isspa_check(isspa_stream_sync(nullptr), "sync");
}

void atom::get_pos_vel_for_from_gpu()
{
    isspa_check(isspa_memcpy_dtoh(for_h, for_d, nAtoms * sizeof(float4), nullptr), "dtoh");
    isspa_check(isspa_memcpy_dtoh(isspaf_h, isspaf_d, nAtoms * sizeof(float4), nullptr), "dtoh");
    get_pos_vel_from_gpu();
}

void atom::get_pos_vel_from_gpu()
{
    isspa_check(isspa_memcpy_dtoh(vel_h, vel_d, nAtoms * sizeof(float4), nullptr), "dtoh");
    isspa_check(isspa_memcpy_dtoh(pos_h, pos_d, nAtoms * sizeof(float4), nullptr), "dtoh");
}

// xyz frames, src/atom_class.cpp:248-290
static void write_xyz(FILE *fh, int n, const float4 *v)
{
    if (!fh) return;
    fprintf(fh, "%d\n%d\n", n, n);
    for (int i = 0; i < n; ++i) fprintf(fh, "C %10.6f %10.6f %10.6f\n", v[i].x, v[i].y, v[i].z);
    fflush(fh);
}
void atom::print_for() { write_xyz(forFile, nAtoms, for_h); }
void atom::print_pos() { write_xyz(posFile, nAtoms, pos_h); }
void atom::print_vel() { write_xyz(velFile, nAtoms, vel_h); }
void atom::print_isspaf()
{
    if (!IFFile) return;
    fprintf(IFFile, "%d\n%d\n", nAtoms, nAtoms);
    for (int i = 0; i < nAtoms; ++i)
        fprintf(IFFile, "C %10.6f %10.6f %10.6f %16.8e %16.8e %16.8e\n", pos_h[i].x, pos_h[i].y,
                pos_h[i].z, isspaf_h[i].x, isspaf_h[i].y, isspaf_h[i].z);
    fflush(IFFile);
}

static void write_rst(const char *path, const char *title, int n, const float4 *v, float lbox,
                      unsigned step)
{
    FILE *fh = fopen(path, "w");
    if (!fh) return;
    fprintf(fh, "%s step %u\n%d\n", title, step, n);
    for (int i = 0; i < n; i += 2) {
        fprintf(fh, "%12.7f%12.7f%12.7f", v[i].x, v[i].y, v[i].z);
        if (i + 1 < n) fprintf(fh, "%12.7f%12.7f%12.7f\n", v[i + 1].x, v[i + 1].y, v[i + 1].z);
        else fprintf(fh, "\n");
    }
    fprintf(fh, "%12.7f%12.7f%12.7f%12.7f%12.7f%12.7f\n", lbox, lbox, lbox, 90.0, 90.0, 90.0);
    fclose(fh);
    char ep[MAXCHAR + 8];
    exact_path(path, ep, sizeof ep);
    FILE *bh = fopen(ep, "wb");
    if (!bh) return;
    ExactHeader h{};
    memcpy(h.magic, kExactMagic, 8);
    h.n = n;
    h.step = step;
    std::vector<float> ex(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        ex[3 * i] = v[i].x;
        ex[3 * i + 1] = v[i].y;
        ex[3 * i + 2] = v[i].z;
    }
    fwrite(&h, sizeof h, 1, bh);
    fwrite(ex.data(), sizeof(float), ex.size(), bh);
    fclose(bh);
}

// src/atom_class.cpp:292-327; `step` = steps completed (the Philox step of
// the next one)
void atom::write_rst_files(const char *posName, const char *velName, float lbox, unsigned step)
{
    get_pos_vel_from_gpu();
    write_rst(posName, "Restart Positions from IS-SPA MI355X engine", nAtoms, pos_h, lbox, step);
    write_rst(velName, "Restart Velocities from IS-SPA MI355X engine", nAtoms, vel_h, lbox, step);
}

void atom::free_arrays()
{
    delete[] pos_h; delete[] vel_h; delete[] mass_h; delete[] for_h; delete[] isspaf_h;
    delete[] ityp_h; delete[] nExcludedAtoms_h; delete[] excludedAtomsList_h;
    delete[] nonBondedParmIndex_h; delete[] lj_h; delete[] molPointer_h;
    pos_h = vel_h = mass_h = for_h = isspaf_h = nullptr;
    ityp_h = nExcludedAtoms_h = excludedAtomsList_h = nonBondedParmIndex_h = molPointer_h = nullptr;
    lj_h = nullptr;
    for (FILE **f : {&forFile, &posFile, &velFile, &IFFile})
        if (*f) {
            fclose(*f);
            *f = nullptr;
        }
}

void atom::free_arrays_gpu()
{
    isspa_dev_free(pos_d); isspa_dev_free(vel_d); isspa_dev_free(for_d); isspa_dev_free(isspaf_d);
    pos_d = vel_d = for_d = isspaf_d = nullptr;
    if (nonbond) nonbond_model_destroy(nonbond);
    nonbond = nullptr;
}

// ------------------------------------------------------- bond/angle/dih ----
void bond::allocate()
{
    bondAtoms_h = new int4[nBonds > 0 ? nBonds : 1]();
    bondParams_h = new float2[nTypes > 0 ? nTypes : 1]();
}
void bond::initialize_gpu() {}
void bond::free_arrays() { delete[] bondAtoms_h; delete[] bondParams_h; bondAtoms_h = nullptr; bondParams_h = nullptr; }
void bond::free_arrays_gpu() { if (model) bonded_model_destroy(model); model = nullptr; }

void angle::allocate()
{
    angleAtoms_h = new int4[nAngles > 0 ? nAngles : 1]();
    angleParams_h = new float2[nTypes > 0 ? nTypes : 1]();
}
void angle::initialize_gpu() {}
void angle::free_arrays() { delete[] angleAtoms_h; delete[] angleParams_h; angleAtoms_h = nullptr; angleParams_h = nullptr; }
void angle::free_arrays_gpu() { if (model) bonded_model_destroy(model); model = nullptr; }

void dih::allocate()
{
    dihAtoms_h = new int4[nDihs > 0 ? nDihs : 1]();
    dihTypes_h = new int[nDihs > 0 ? nDihs : 1]();
    dihParams_h = new float4[nTypes > 0 ? nTypes : 1]();
    scaled14Factors_h = new float2[nTypes > 0 ? nTypes : 1]();
}
void dih::initialize_gpu() {}
void dih::free_arrays()
{
    delete[] dihAtoms_h; delete[] dihTypes_h; delete[] dihParams_h; delete[] scaled14Factors_h;
    dihAtoms_h = nullptr; dihTypes_h = nullptr; dihParams_h = nullptr; scaled14Factors_h = nullptr;
}
void dih::free_arrays_gpu() { if (model) bonded_model_destroy(model); model = nullptr; }

// --------------------------------------------------------------- isspa -----
void isspa::allocate(int n, int configMC)
{
    nAtoms = n;
    nMC = configMC;
    isspaTypes_h = new int[n]();
    rmax_h = new float[nTypes]();
    vtot_h = new float[nTypes]();
    isspaForceTable_h = new float[(size_t)nTypes * nRs]();
    isspaForceR_h = new float[nRs]();
    isspaGTable_h = new float[(size_t)nTypes * nGRs]();
    isspaGR_h = new float[nGRs]();
    isspaETable_h = new float[(size_t)nTypes * nERs]();
    isspaER_h = new float[nERs]();
}

// src/isspa_class.cpp:103-113
void isspa::construct_parameter_arrays()
{
    const float PI = 3.1415926535f;
    for (int i = 0; i < nTypes; ++i)
        vtot_h[i] = (float)(4.0 / 3.0 * PI * rmax_h[i] * rmax_h[i] * rmax_h[i] * 0.0074 /
                            ((float)nMC));
}

// src/isspa_class.cpp:115-248: POINTERS (nAtoms, nTypes, nGRs, nERs, nRs at
// columns 0/8/16/24/32), ISSPA_TYPE_INDEX (10I8, 1-based), ISSPA_RMAX
// (5E16.8), tables (row = r then one column per type, width 16)
int isspa::read_isspa_prmtop(const char *path, int configMC)
{
    FILE *fh = fopen(path, "r");
    if (!fh) {
        fprintf(stderr, "could not open IS-SPA topology %s\n", path);
        return 1;
    }
    char line[LONGCHAR];
    auto field = [](const char *l, int col, int w) {
        char tok[32] = {0};
        const size_t len = strlen(l);
        if ((size_t)col < len) memcpy(tok, l + col, std::min<size_t>({(size_t)w, len - col, 31}));
        return std::string(tok);
    };
    int n = -1, status = 0;
    while (fgets(line, LONGCHAR, fh)) {
        if (strncmp(line, "%FLAG", 5) != 0) continue;
        char flag[64] = {0};
        sscanf(line + 5, "%63s", flag);
        auto next = [&]() { return fgets(line, LONGCHAR, fh) != nullptr; };
        if (!strcmp(flag, "POINTERS")) {
            next();  // format
            if (!next()) { status = 2; break; }
            n = atoi(field(line, 0, 8).c_str());
            nTypes = atoi(field(line, 8, 8).c_str());
            nGRs = atoi(field(line, 16, 8).c_str());
            nERs = atoi(field(line, 24, 8).c_str());
            nRs = atoi(field(line, 32, 8).c_str());
            if (n < 1 || nTypes < 1 || nGRs < 2 || nERs < 2 || nRs < 2) { status = 3; break; }
            allocate(n, configMC);
        } else if (n > 0 && !strcmp(flag, "ISSPA_TYPE_INDEX")) {
            next();
            int c = 0;
            while (c < n && next())
                for (int k = 0; k < 10 && c < n; ++k) isspaTypes_h[c++] = atoi(field(line, 8 * k, 8).c_str()) - 1;
            if (c < n) status = 4;
        } else if (n > 0 && !strcmp(flag, "ISSPA_RMAX")) {
            next();
            int c = 0;
            while (c < nTypes && next())
                for (int k = 0; k < 5 && c < nTypes; ++k) rmax_h[c++] = (float)atof(field(line, 16 * k, 16).c_str());
            if (c < nTypes) status = 5;
        } else if (n > 0 && (!strcmp(flag, "ISSPA_DENSITIES") || !strcmp(flag, "ISSPA_EFIELD") ||
                             !strcmp(flag, "ISSPA_FORCES"))) {
            float *tab, *r;
            float2 *par;
            int rows;
            if (!strcmp(flag, "ISSPA_DENSITIES")) { tab = isspaGTable_h; r = isspaGR_h; rows = nGRs; par = &gRparams; }
            else if (!strcmp(flag, "ISSPA_EFIELD")) { tab = isspaETable_h; r = isspaER_h; rows = nERs; par = &eRparams; }
            else { tab = isspaForceTable_h; r = isspaForceR_h; rows = nRs; par = &forceRparams; }
            next();
            int b = 0;
            for (; b < rows && next(); ++b) {
                r[b] = (float)atof(field(line, 0, 16).c_str());
                for (int t = 0; t < nTypes; ++t) tab[(size_t)t * rows + b] = (float)atof(field(line, 16 * (t + 1), 16).c_str());
            }
            if (b < rows) { status = 6; break; }
            par->x = r[0];
            par->y = r[1] - r[0];
        }
    }
    fclose(fh);
    if (n < 1 && !status) status = 7;
    if (status) {
        fprintf(stderr, "IS-SPA topology %s: malformed (code %d)\n", path, status);
        return status;
    }
    for (int i = 0; i < n; ++i)
        if (isspaTypes_h[i] < 0 || isspaTypes_h[i] >= nTypes) {
            fprintf(stderr, "IS-SPA topology %s: type of atom %d out of range\n", path, i);
            return 8;
        }
    construct_parameter_arrays();
    return 0;
}

void isspa::initialize_gpu(int n, int s)
{
    nAtoms = n;
    seed = s;
}

void isspa::free_arrays()
{
    delete[] isspaTypes_h; delete[] rmax_h; delete[] vtot_h; delete[] isspaForceTable_h;
    delete[] isspaForceR_h; delete[] isspaGTable_h; delete[] isspaGR_h; delete[] isspaETable_h;
    delete[] isspaER_h;
    isspaTypes_h = nullptr;
    rmax_h = vtot_h = isspaForceTable_h = isspaForceR_h = isspaGTable_h = isspaGR_h = nullptr;
    isspaETable_h = isspaER_h = nullptr;
}

void isspa::free_arrays_gpu()
{
    if (model) isspa_model_destroy(model);
    model = nullptr;
}

// ------------------------------------------------------------------ us -----
// `k = `, `x0 = `, `cvFile = `, `atomSelect1/2 = ` (1-based comma lists),
// src/us_class.cpp:13-104
int us::initialize(const char *path)
{
    FILE *fh = fopen(path, "r");
    if (!fh) {
        fprintf(stderr, "could not open umbrella file %s\n", path);
        return 1;
    }
    char line[LONGCHAR];
    std::vector<int> g1, g2;
    std::string cv = "cv.dat";
    while (fgets(line, LONGCHAR, fh)) {
        char *hash = strchr(line, '#');
        if (hash) *hash = 0;
        char *eq = strchr(line, '=');
        if (!eq) continue;
        *eq = 0;
        char *key = trim(line), *val = trim(eq + 1);
        if (!strncmp(key, "k", 1) && strlen(key) == 1) k = (float)atof(val);
        else if (!strncmp(key, "x0", 2)) x0 = (float)atof(val);
        else if (!strncmp(key, "cvFile", 6)) cv = val;
        else if (!strncmp(key, "atomSelect1", 11) || !strncmp(key, "atomSelect2", 11)) {
            std::vector<int> &g = key[10] == '1' ? g1 : g2;
            for (char *tok = strtok(val, ","); tok; tok = strtok(nullptr, ",")) g.push_back(atoi(trim(tok)) - 1);
        }
    }
    fclose(fh);
    totalBiasAtoms = (int)(g1.size() + g2.size());
    if (g1.empty() || g2.empty()) {
        fprintf(stderr, "umbrella file %s: both atom groups are required\n", path);
        return 2;
    }
    atomList_h = new int2[totalBiasAtoms];
    int b = 0;
    for (int a : g1) atomList_h[b++] = int2{a, 0};
    for (int a : g2) atomList_h[b++] = int2{a, 1};
    cvFile = fopen(cv.c_str(), "w");
    return 0;
}

// src/us_class.cpp:136-157: masses normalised by the FP32 group totals
void us::populate_mass(float4 *m, int nAtoms)
{
    mass_h = new float[totalBiasAtoms];
    float tot[2] = {0.0f, 0.0f};
    for (int b = 0; b < totalBiasAtoms; ++b) {
        const int a = atomList_h[b].x;
        mass_h[b] = (a >= 0 && a < nAtoms) ? m[a].w : 0.0f;
        tot[atomList_h[b].y] += mass_h[b];
    }
    for (int b = 0; b < totalBiasAtoms; ++b) mass_h[b] /= tot[atomList_h[b].y];
    kumb_h[0] = -k;
    kumb_h[1] = k;
}

void us::initialize_gpu() { groupComPos_d = dev_alloc<double>(6); }

// src/us_class.cpp:107-133
void us::print_cv(int step, float lbox)
{
    isspa_check(isspa_memcpy_dtoh(groupComPos_h, groupComPos_d, 6 * sizeof(double), nullptr), "dtoh");
    const double h = 0.5 * lbox;
    double s = 0.0;
    for (int c = 0; c < 3; ++c) {
        double r = groupComPos_h[c] - groupComPos_h[3 + c];
        if (r > h) r -= lbox;
        else if (r < -h) r += lbox;
        s += r * r;
    }
    if (cvFile) {
        fprintf(cvFile, "%8d  %8.3f\n", step, sqrt(s));
        fflush(cvFile);
    }
}

void us::free_arrays()
{
    delete[] atomList_h;
    delete[] mass_h;
    atomList_h = nullptr;
    mass_h = nullptr;
    if (cvFile) fclose(cvFile);
    cvFile = nullptr;
}

void us::free_arrays_gpu()
{
    isspa_dev_free(groupComPos_d);
    groupComPos_d = nullptr;
    if (model) us_model_destroy(model);
    model = nullptr;
}

// -------------------------------------------------------------- config -----
// src/config_class.cpp:14-133: `key = value`, `#` comments, prefix keys in
// the reference's order; dt fixed at 2 fs (Q23); T converted to kcal/mol.
int config::initialize(const char *path)
{
    FILE *fh = fopen(path, "r");
    if (!fh) {
        fprintf(stderr, "could not open config file %s\n", path);
        return 1;
    }
    char line[MAXCHAR];
    while (fgets(line, MAXCHAR, fh)) {
        char *hash = strchr(line, '#');
        if (hash) *hash = 0;
        char *eq = strchr(line, '=');
        if (!eq) continue;
        *eq = 0;
        char *key = trim(line), *val = trim(eq + 1);
        auto is = [key](const char *k) { return strncmp(key, k, strlen(k)) == 0; };
        auto cp = [val](char *dst) { strncpy(dst, val, MAXCHAR - 1); };
        if (is("prmtop")) cp(prmtopFileName);
        else if (is("isspaPrmtop")) cp(isspaPrmtopFileName);
        else if (is("posFile")) cp(posOutFileName);
        else if (is("posRstFile")) cp(posRstFileName);
        else if (is("velFile")) cp(velOutFileName);
        else if (is("velRstFile")) cp(velRstFileName);
        else if (is("forFile")) cp(forOutFileName);
        else if (is("inputCoord")) cp(inputCoordFileName);
        else if (is("nMC")) nMC = atoi(val);
        else if (is("temperature")) T = (float)atof(val);
        else if (is("seed")) seed = (int)atof(val);
        else if (is("cutoff")) rCut = (float)atof(val);
        else if (is("dielectric")) eps = (float)atof(val);
        else if (is("nSteps")) nSteps = atoi(val);
        else if (is("deltaWrite")) deltaWrite = atoi(val);
        else if (is("boxLength")) lbox = (float)atof(val);
        else if (is("pnu")) pnu = (float)atof(val);
        else if (is("velRst")) velRst = atoi(val);
        else if (is("inputVel")) cp(inputVelFileName);
        else if (is("US")) us = atoi(val);
        else if (is("usCfgFile")) cp(usCfgFileName);
    }
    fclose(fh);
    dtPs = 0.002f;
    dt = dtPs * 20.455f;
    T *= 0.00198717f;
    rCut2 = rCut * rCut;
    rNN = 15.0f;
    rNN2 = rNN * rNN;
    if (deltaWrite < 1) deltaWrite = 1;
    return 0;
}

// -------------------------------------------------------------- timing -----
void timing::initialize()
{
    totalStartSec = now_sec();
    bondTime = angleTime = dihTime = nonbondTime = neighborListTime = 0.0f;
    writeTime = leapFrogTime = isspaTime = usTime = 0.0f;
}
void timing::startWriteTimer() { writeStartSec = now_sec(); }
void timing::stopWriteTimer() { writeTime += (float)(1e3 * (now_sec() - writeStartSec)); }

// src/timing_class.cpp:68-102: ns/day = simulated ns / elapsed days
void timing::print_final(float elapsedns)
{
    isspa_stream_sync(nullptr);
    milliseconds = (float)(1e3 * (now_sec() - totalStartSec));
    const char *unit = "ms";
    float divider = 1.0f;
    if (milliseconds > 1.0e3f) {
        unit = "s";
        divider = 1.0e3f;
    }
    printf("Elapsed CPU/GPU time = %10.2f %s\n", milliseconds / divider, unit);
    printf("Simulation time = %10.2f ns\n", elapsedns);
    day_per_millisecond = 1e-3f / 60.0f / 60.0f / 24.0f;
    printf("Average ns/day = %10.2f\n", elapsedns / (milliseconds * day_per_millisecond));
    const struct { const char *name; float t; } rows[] = {
        {"Bond force calculation time", bondTime}, {"Angle force calculation time", angleTime},
        {"Dihedral force calculation time", dihTime}, {"Nonbond force calculation time", nonbondTime},
        {"Neighbor list calculation time", neighborListTime}, {"US bias force calculation time", usTime},
        {"IS-SPA force calculation time", isspaTime}, {"Leap-frog propogation time", leapFrogTime},
        {"Write trajectory file time", writeTime}};
    float acc = 0.0f;
    for (const auto &r : rows) {
        printf("%s = %10.2f %s (%5.1f %%)\n", r.name, r.t / divider, unit, r.t / milliseconds * 100.0f);
        acc += r.t;
    }
    printf("Total percent accounted for = %10.2f %%\n", acc / milliseconds * 100.0f);
}
