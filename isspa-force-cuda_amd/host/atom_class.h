// class atom: positions, velocities, forces and nonbonded parameters, host
// and device copies (counterpart of src/atom_class.h:11-106).  Device memory
// is owned through the C ABI; no GPU runtime type appears here.
#pragma once

#include <cstdio>

#include "host_types.h"

struct nonbond_model;

class atom
{
  private:
    FILE *forFile = nullptr;
    FILE *posFile = nullptr;
    FILE *velFile = nullptr;
    FILE *IFFile = nullptr;
    float rand_gauss();

  public:
    int nAtoms = 0;
    int nPairs = 0;
    int nTypes = 0;
    int excludedAtomsListLength = 0;
    int nMols = 0;
    int *molPointer_h = nullptr;
    float4 *pos_h = nullptr, *pos_d = nullptr;      // {x, y, z, charge}
    float4 *for_h = nullptr, *for_d = nullptr;      // {fx, fy, fz, -}
    float4 *isspaf_h = nullptr, *isspaf_d = nullptr;// IS-SPA force alone (debug output)
    float4 *vel_h = nullptr, *vel_d = nullptr;      // {vx, vy, vz, dynamic mass}
    float4 *mass_h = nullptr;                       // .w = true mass
    int *ityp_h = nullptr;
    int *nExcludedAtoms_h = nullptr;                // inclusive prefix sum
    int *excludedAtomsList_h = nullptr;
    int *nonBondedParmIndex_h = nullptr;
    float2 *lj_h = nullptr;                         // {A, B}
    int seed = 0;                                   // Philox key of the Andersen draws
    unsigned step = 0;                              // Philox counter, +1 per leapfrog_cuda
    unsigned posRstStep = 0, velRstStep = 0;        // steps completed, from restart files
    float lbox = 0.0f;                              // set by nonbond_force_cuda_grid_block
    nonbond_model *nonbond = nullptr;               // built by the first nonbond_force_cuda
    // launch geometry kept for interface parity (the engine chooses its own)
    int gridSize = 0, blockSize = 0, minGridSize = 0, nThreads = 0;

    void allocate();
    void allocate_molecule_arrays();
    void read_initial_positions(const char *);
    void read_initial_velocities(const char *);
    void initialize_velocities(float T);
    void open_traj_files(const char *forOutFileName, const char *posOutFileName,
                         const char *velOutFileName);
    void initialize_gpu(int seed);
    void copy_params_to_gpu();
    void copy_pos_vel_to_gpu();
    void get_pos_vel_for_from_gpu();
    void get_pos_vel_from_gpu();
    void print_pos();
    void print_vel();
    void print_for();
    void print_isspaf();
    void write_rst_files(const char *posRstFileName, const char *velRstFileName, float lbox,
                         unsigned step);
    void free_arrays();
    void free_arrays_gpu();
};
