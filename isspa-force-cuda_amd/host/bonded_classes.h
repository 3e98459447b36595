// classes bond, angle, dih (counterparts of src/bond_class.h, src/angle_class.h,
// src/dih_class.h): term lists and parameters in the reference's layouts.
// The device copies live in a bonded_model built by the first force call.
#pragma once

#include "host_types.h"

struct bonded_model;

class bond
{
  public:
    int nBonds = 0, nBondHs = 0, nBondnHs = 0, nTypes = 0;
    int4 *bondAtoms_h = nullptr;    // {i, j, type, -}
    float2 *bondParams_h = nullptr; // {2k, r0}
    bonded_model *model = nullptr;
    int gridSize = 0, blockSize = 0, minGridSize = 0;
    void allocate();
    void initialize_gpu();
    void free_arrays();
    void free_arrays_gpu();
};

class angle
{
  public:
    int nAngles = 0, nAngleHs = 0, nAnglenHs = 0, nTypes = 0;
    int4 *angleAtoms_h = nullptr;    // {i, j, k, type}
    float2 *angleParams_h = nullptr; // {2k, theta0}
    bonded_model *model = nullptr;
    int gridSize = 0, blockSize = 0, minGridSize = 0;
    void allocate();
    void initialize_gpu();
    void free_arrays();
    void free_arrays_gpu();
};

class dih
{
  public:
    int nDihs = 0, nDihHs = 0, nDihnHs = 0, nTypes = 0;
    int4 *dihAtoms_h = nullptr;          // {i, j, +-k, +-l}
    int *dihTypes_h = nullptr;
    float4 *dihParams_h = nullptr;       // {n, V, phase, -}
    float2 *scaled14Factors_h = nullptr; // {scee, scnb}
    bonded_model *model = nullptr;
    int gridSize = 0, blockSize = 0, minGridSize = 0;
    void allocate();
    void initialize_gpu();
    void free_arrays();
    void free_arrays_gpu();
};
