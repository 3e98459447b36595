// class isspa: IS-SPA types, rmax, tables and the device model (counterpart of
// src/isspa_class.h:11-68; the curand states are replaced by Philox keys).
#pragma once

#include "host_types.h"

struct isspa_model;

class isspa
{
  public:
    int nTypes = 0;
    int nMC = 0;
    int nRs = 0, nGRs = 0, nERs = 0;
    int nAtoms = 0;
    int seed = 0;
    unsigned step = 0;  // Philox counter of the MC points, +1 per isspa_force_cuda
    int *isspaTypes_h = nullptr;
    float *rmax_h = nullptr;
    float *vtot_h = nullptr;
    float *isspaForceTable_h = nullptr, *isspaForceR_h = nullptr;
    float *isspaGTable_h = nullptr, *isspaGR_h = nullptr;
    float *isspaETable_h = nullptr, *isspaER_h = nullptr;
    float2 forceRparams{0, 0}, gRparams{0, 0}, eRparams{0, 0};  // {r0, dr}
    isspa_model *model = nullptr;

    // returns 0 on success, nonzero on a missing or malformed file (the
    // reference silently continues on uninitialised memory, quirk Q16)
    int read_isspa_prmtop(const char *isspaPrmtopFileName, int configMC);
    void allocate(int nAtoms, int configMC);
    void construct_parameter_arrays();
    void initialize_gpu(int nAtoms, int seed);
    void free_arrays();
    void free_arrays_gpu();
};
