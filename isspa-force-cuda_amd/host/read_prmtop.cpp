// AMBER prmtop reader with the reference's conversions (src/read_prmtop.cpp:10-593):
// charge / sqrt(dielectric); dynamic mass 12 below mass 2 (true mass kept in
// mass_h.w); exclusion counts -> inclusive prefix sum; coordinate indices / 3
// (C truncation keeps the dihedral sign flags); type indices - 1; bond and
// angle force constants doubled.  Sections are cut by their %FORMAT widths.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "force_calls.h"

namespace {

struct Section {
    int per_line = 0, width = 0;
    std::vector<std::string> lines;
};

bool parse_format(const char *s, int &per_line, int &width)
{
    const char *p = strchr(s, '(');
    if (!p) return false;
    ++p;
    per_line = (int)strtol(p, (char **)&p, 10);
    if (!*p) return false;
    ++p;  // a / I / E / F
    width = (int)strtol(p, nullptr, 10);
    return per_line > 0 && width > 0;
}

std::map<std::string, Section> read_sections(FILE *fh)
{
    std::map<std::string, Section> out;
    char line[LONGCHAR];
    Section *cur = nullptr;
    bool want_fmt = false;
    while (fgets(line, LONGCHAR, fh)) {
        size_t len = strlen(line);
        while (len && (line[len - 1] == '\n' || line[len - 1] == '\r')) line[--len] = 0;
        if (!strncmp(line, "%FLAG", 5)) {
            char flag[128] = {0};
            sscanf(line + 5, "%127s", flag);
            cur = &out[flag];
            want_fmt = true;
        } else if (cur && want_fmt) {
            parse_format(line, cur->per_line, cur->width);
            want_fmt = false;
        } else if (cur) {
            cur->lines.emplace_back(line);
        }
    }
    return out;
}

bool values(const std::map<std::string, Section> &s, const char *flag, size_t count,
            std::vector<double> &out)
{
    out.clear();
    auto it = s.find(flag);
    if (it == s.end()) return count == 0;
    const Section &sec = it->second;
    for (const std::string &l : sec.lines) {
        for (int c = 0; c < sec.per_line && out.size() < count; ++c) {
            const size_t at = (size_t)c * sec.width;
            if (at >= l.size()) break;
            out.push_back(atof(l.substr(at, sec.width).c_str()));
        }
        if (out.size() >= count) break;
    }
    return out.size() >= count;
}

}  // namespace

int read_prmtop(const char *path, atom &atoms, bond &bonds, angle &angles, dih &dihs, float eps)
{
    FILE *fh = fopen(path, "r");
    if (!fh) {
        fprintf(stderr, "could not open prmtop %s\n", path);
        return 1;
    }
    auto s = read_sections(fh);
    fclose(fh);
    std::vector<double> v;
    auto need = [&](const char *flag, size_t count) {
        if (!values(s, flag, count, v)) {
            fprintf(stderr, "prmtop %s: section %s truncated or missing (%zu values)\n", path, flag, count);
            return false;
        }
        return true;
    };
    if (!need("POINTERS", 18)) return 2;
    const std::vector<double> p = v;
    atoms.nAtoms = (int)p[0];
    atoms.nTypes = (int)p[1];
    bonds.nBondHs = (int)p[2];
    bonds.nBondnHs = (int)p[3];
    angles.nAngleHs = (int)p[4];
    angles.nAnglenHs = (int)p[5];
    dihs.nDihHs = (int)p[6];
    dihs.nDihnHs = (int)p[7];
    atoms.excludedAtomsListLength = (int)p[10];
    bonds.nTypes = (int)p[15];
    angles.nTypes = (int)p[16];
    dihs.nTypes = (int)p[17];
    bonds.nBonds = bonds.nBondHs + bonds.nBondnHs;
    angles.nAngles = angles.nAngleHs + angles.nAnglenHs;
    dihs.nDihs = dihs.nDihHs + dihs.nDihnHs;
    const int n = atoms.nAtoms, T = atoms.nTypes;
    if (n < 1 || T < 1) {
        fprintf(stderr, "prmtop %s: bad POINTERS (natom=%d ntypes=%d)\n", path, n, T);
        return 3;
    }
    atoms.allocate();
    bonds.allocate();
    angles.allocate();
    dihs.allocate();
    const float sqrtEps = sqrtf(eps);
    if (!need("CHARGE", n)) return 4;
    for (int i = 0; i < n; ++i) atoms.pos_h[i].w = (float)v[i] / sqrtEps;
    if (!need("MASS", n)) return 4;
    for (int i = 0; i < n; ++i) {
        atoms.mass_h[i].w = (float)v[i];
        atoms.vel_h[i].w = (float)v[i] < 2.0f ? 12.0f : (float)v[i];
    }
    if (!need("ATOM_TYPE_INDEX", n)) return 4;
    for (int i = 0; i < n; ++i) atoms.ityp_h[i] = (int)v[i] - 1;
    if (!need("NUMBER_EXCLUDED_ATOMS", n)) return 4;
    for (int i = 0, acc = 0; i < n; ++i) atoms.nExcludedAtoms_h[i] = (acc += (int)v[i]);
    if (!need("NONBONDED_PARM_INDEX", (size_t)T * T)) return 4;
    for (int k = 0; k < T * T; ++k) atoms.nonBondedParmIndex_h[k] = (int)v[k] - 1;
    const int ntri = T * (T + 1) / 2;
    if (!need("LENNARD_JONES_ACOEF", ntri)) return 4;
    for (int k = 0; k < ntri; ++k) atoms.lj_h[k].x = (float)v[k];
    if (!need("LENNARD_JONES_BCOEF", ntri)) return 4;
    for (int k = 0; k < ntri; ++k) atoms.lj_h[k].y = (float)v[k];
    if (!need("BOND_FORCE_CONSTANT", bonds.nTypes)) return 4;
    for (int k = 0; k < bonds.nTypes; ++k) bonds.bondParams_h[k].x = (float)v[k] * 2.0f;
    if (!need("BOND_EQUIL_VALUE", bonds.nTypes)) return 4;
    for (int k = 0; k < bonds.nTypes; ++k) bonds.bondParams_h[k].y = (float)v[k];
    if (!need("ANGLE_FORCE_CONSTANT", angles.nTypes)) return 4;
    for (int k = 0; k < angles.nTypes; ++k) angles.angleParams_h[k].x = (float)v[k] * 2.0f;
    if (!need("ANGLE_EQUIL_VALUE", angles.nTypes)) return 4;
    for (int k = 0; k < angles.nTypes; ++k) angles.angleParams_h[k].y = (float)v[k];
    const char *dflags[3] = {"DIHEDRAL_PERIODICITY", "DIHEDRAL_FORCE_CONSTANT", "DIHEDRAL_PHASE"};
    for (int c = 0; c < 3; ++c) {
        if (!need(dflags[c], dihs.nTypes)) return 4;
        for (int k = 0; k < dihs.nTypes; ++k) (&dihs.dihParams_h[k].x)[c] = (float)v[k];
    }
    if (!need("SCEE_SCALE_FACTOR", dihs.nTypes)) return 4;
    for (int k = 0; k < dihs.nTypes; ++k) dihs.scaled14Factors_h[k].x = (float)v[k];
    if (!need("SCNB_SCALE_FACTOR", dihs.nTypes)) return 4;
    for (int k = 0; k < dihs.nTypes; ++k) dihs.scaled14Factors_h[k].y = (float)v[k];
    auto cdiv3 = [](double x) { return (int)x / 3; };  // C truncation keeps the sign
    struct { const char *flag; int count, off; } bl[2] = {{"BONDS_INC_HYDROGEN", bonds.nBondHs, 0},
                                                       {"BONDS_WITHOUT_HYDROGEN", bonds.nBondnHs, bonds.nBondHs}};
    for (auto &b : bl) {
        if (!need(b.flag, (size_t)3 * b.count)) return 4;
        for (int q = 0; q < b.count; ++q)
            bonds.bondAtoms_h[b.off + q] = int4{cdiv3(v[3 * q]), cdiv3(v[3 * q + 1]), (int)v[3 * q + 2] - 1, 0};
    }
    struct { const char *flag; int count, off; } al[2] = {{"ANGLES_INC_HYDROGEN", angles.nAngleHs, 0},
                                                       {"ANGLES_WITHOUT_HYDROGEN", angles.nAnglenHs, angles.nAngleHs}};
    for (auto &a : al) {
        if (!need(a.flag, (size_t)4 * a.count)) return 4;
        for (int q = 0; q < a.count; ++q)
            angles.angleAtoms_h[a.off + q] = int4{cdiv3(v[4 * q]), cdiv3(v[4 * q + 1]), cdiv3(v[4 * q + 2]), (int)v[4 * q + 3] - 1};
    }
    struct { const char *flag; int count, off; } dl[2] = {{"DIHEDRALS_INC_HYDROGEN", dihs.nDihHs, 0},
                                                       {"DIHEDRALS_WITHOUT_HYDROGEN", dihs.nDihnHs, dihs.nDihHs}};
    for (auto &d : dl) {
        if (!need(d.flag, (size_t)5 * d.count)) return 4;
        for (int q = 0; q < d.count; ++q) {
            dihs.dihAtoms_h[d.off + q] = int4{cdiv3(v[5 * q]), cdiv3(v[5 * q + 1]), cdiv3(v[5 * q + 2]), cdiv3(v[5 * q + 3])};
            dihs.dihTypes_h[d.off + q] = (int)v[5 * q + 4] - 1;
        }
    }
    if (!need("EXCLUDED_ATOMS_LIST", atoms.excludedAtomsListLength)) return 4;
    for (int k = 0; k < atoms.excludedAtomsListLength; ++k) atoms.excludedAtomsList_h[k] = (int)v[k] - 1;
    if (values(s, "SOLVENT_POINTERS", 3, v)) {
        atoms.nMols = (int)v[1];
        atoms.allocate_molecule_arrays();
        if (values(s, "ATOMS_PER_MOLECULE", atoms.nMols, v))
            for (int k = 0; k < atoms.nMols; ++k) atoms.molPointer_h[k] = (int)v[k];
    }
    return 0;
}
