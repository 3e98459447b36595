// classes us, config and timing (counterparts of src/us_class.h:116-149,
// src/config_class.h:160-196, src/timing_class.h:7-30).
#pragma once

#include <cstdio>

#include "host_types.h"

struct us_model;

class us
{
  private:
    FILE *cvFile = nullptr;

  public:
    int2 *atomList_h = nullptr;   // {atom, group}
    float *mass_h = nullptr;      // mass / group mass
    double *groupComPos_d = nullptr;  // device double[6]: COM of group 0, group 1
    double groupComPos_h[6] = {0, 0, 0, 0, 0, 0};
    int totalBiasAtoms = 0;
    float kumb_h[2] = {0, 0};
    float x0 = 0.0f;
    float k = 0.0f;
    us_model *model = nullptr;
    int usGridSize = 0, usBlockSize = 0;
    int initialize(const char *);
    void populate_mass(float4 *, int);
    void initialize_gpu();
    void print_cv(int, float);
    void free_arrays();
    void free_arrays_gpu();
};

class config
{
  public:
    float T = 0.0f;      // kT in kcal/mol after initialize
    float dt = 0.0f;     // AMBER time units
    float dtPs = 0.0f;
    float pnu = 0.0f;    // Andersen collision probability per step (0 if absent, Q14)
    int nSteps = 0;
    int deltaWrite = 1;
    int deltaNN = 10;
    int nMC = 0;
    int seed = 0;
    int us = 0;
    int velRst = 0;
    float lbox = 0.0f;
    float eps = 1.0f;
    float rCut = 0.0f, rCut2 = 0.0f;
    float rNN = 0.0f, rNN2 = 0.0f;
    char prmtopFileName[MAXCHAR] = "";
    char isspaPrmtopFileName[MAXCHAR] = "";
    char inputCoordFileName[MAXCHAR] = "";
    char inputVelFileName[MAXCHAR] = "";
    char usCfgFileName[MAXCHAR] = "";
    char forOutFileName[MAXCHAR] = "";
    char posOutFileName[MAXCHAR] = "";
    char velOutFileName[MAXCHAR] = "";
    char posRstFileName[MAXCHAR] = "";
    char velRstFileName[MAXCHAR] = "";
    int initialize(const char *);
    void set_cuda_constants() {}
};

class timing
{
  public:
    double totalStartSec = 0.0;
    double writeStartSec = 0.0;
    float bondTime = 0, angleTime = 0, dihTime = 0, nonbondTime = 0, neighborListTime = 0;
    float leapFrogTime = 0, isspaTime = 0, usTime = 0, writeTime = 0;
    float milliseconds = 0;
    float day_per_millisecond = 0;
    void initialize();
    void startWriteTimer();
    void stopWriteTimer();
    void print_final(float elapsedns);
};
