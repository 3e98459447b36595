// The reference's per-step wrapper functions, same names, arguments and
// return value (elapsed milliseconds of the call, measured with events on the
// null stream, synchronous), implemented over the C ABI of include/isspa_amd.h.
//
//   src/isspa_force_cuda.h:3,5      isspa_grid_block, isspa_force_cuda
//   src/nonbond_force_cuda.h:2-3    nonbond_force_cuda, nonbond_force_cuda_grid_block
//   src/bond_force_cuda.h:9,11      bond_force_cuda, bond_force_cuda_grid_block
//   src/angle_force_cuda.h:3-4      angle_force_cuda, angle_force_cuda_grid_block
//   src/dih_force_cuda.h:2-3        dih_force_cuda, dih_force_cuda_grid_block
//   src/leapfrog_cuda.h:4-5         leapfrog_cuda, leapfrog_cuda_grid_block
//   src/us_force_cuda.h:3,5         us_force_cuda, us_grid_block
//   src/init_rand.cuh:2             init_rand_states (no state needed with Philox)
//
// Differences, all deliberate: every wrapper checks the C-ABI status and
// aborts with the library's message on failure (the reference checks no CUDA
// error, quirk Q16); the random streams are Philox counters that advance by
// one step per call instead of curand states.
#pragma once

#include "atom_class.h"
#include "bonded_classes.h"
#include "host_types.h"
#include "isspa_class.h"
#include "us_config_timing.h"

float isspa_force_cuda(float4 *xyz_d, float4 *f_d, float4 *isspaf_d, isspa &isspas, int nAtoms);
void isspa_grid_block(int nAtoms, int nPairs, float lbox, isspa &isspas);
float nonbond_force_cuda(atom &atoms, isspa &isspas, int nAtoms);
extern "C" void nonbond_force_cuda_grid_block(atom &atoms, float rCut2, float lbox);
float bond_force_cuda(float4 *xyz_d, float4 *f_d, int nAtoms, float lbox, bond &bonds);
extern "C" void bond_force_cuda_grid_block(int nBonds, int *gridSize, int *blockSize,
                                           int *minGridSize);
float angle_force_cuda(float4 *xyz_d, float4 *f_d, int nAtoms, float lbox, angle &angles);
extern "C" void angle_force_cuda_grid_block(int nAngles, int *gridSize, int *blockSize,
                                            int *minGridSize);
float dih_force_cuda(atom &atoms, dih &dihs, float lbox);
extern "C" void dih_force_cuda_grid_block(int nDihs, int *gridSize, int *blockSize,
                                          int *minGridSize);
float leapfrog_cuda(atom &atoms, config &configs);
extern "C" void leapfrog_cuda_grid_block(int nAtoms, int *gridSize, int *blockSize,
                                         int *minGridSize);
float us_force_cuda(float4 *xyz_d, float4 *f_d, us &bias, float lbox, int nAtoms);
void us_grid_block(us &bias);
extern "C" void init_rand_states(void *state, int seed, int n);

// AMBER prmtop reader (src/read_prmtop.cpp:10): returns 0 or a nonzero status
// with a message on stderr (missing file, truncated section, bad POINTERS).
int read_prmtop(const char *prmtopFileName, atom &atoms, bond &bonds, angle &angles, dih &dihs,
                float eps);

// abort with the library's last error message when a C-ABI call fails
void isspa_check(int status, const char *what);
