// Step driver: the reference's main() (src/total_force_cuda.cpp:28-179) over
// the restated classes and wrappers.
//
//   isspa_md.x config.txt           wrapper-per-term loop, as the reference
//   isspa_md.x config.txt --graph   same step captured once in a hipGraph and
//                                   replayed (forces re-evaluated eagerly on
//                                   write steps so frames hold that step's forces)
//   isspa_md.x config.txt --check   parse the deck and print the counts only
//
// Step order: zero f -> bond -> angle -> dihedral -> IS-SPA -> nonbond -> US
// -> (write every deltaWrite steps, before integration) -> leap-frog.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/isspa_amd.h"
#include "force_calls.h"

int main(int argc, char *argv[])
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s config.txt [--graph|--check]\n", argv[0]);
        return 1;
    }
    const bool graph = argc > 2 && !strcmp(argv[2], "--graph");
    const bool check_only = argc > 2 && !strcmp(argv[2], "--check");
    timing times;
    atom atoms;
    bond bonds;
    angle angles;
    dih dihs;
    config configs;
    isspa isspas;
    us bias;

    if (configs.initialize(argv[1])) return 1;
    if (read_prmtop(configs.prmtopFileName, atoms, bonds, angles, dihs, configs.eps)) return 1;
    printf("Number of atoms from prmtop file: %d\n", atoms.nAtoms);
    printf("Number of atom types from prmtop file: %d\n", atoms.nTypes);
    printf("Number of bonds: %d (%d with hydrogen)\n", bonds.nBonds, bonds.nBondHs);
    printf("Number of angles: %d (%d with hydrogen)\n", angles.nAngles, angles.nAngleHs);
    printf("Number of dihedrals: %d (%d with hydrogen)\n", dihs.nDihs, dihs.nDihHs);
    printf("Length of excluded atoms list: %d\n", atoms.excludedAtomsListLength);
    if (isspas.read_isspa_prmtop(configs.isspaPrmtopFileName, configs.nMC)) return 1;
    if (isspas.nAtoms != atoms.nAtoms) {
        fprintf(stderr, "atom count mismatch: prmtop %d, IS-SPA topology %d\n", atoms.nAtoms,
                isspas.nAtoms);
        return 1;
    }
    printf("Number of ISSPA types: %d, table rows g/E/F: %d/%d/%d\n", isspas.nTypes, isspas.nGRs,
           isspas.nERs, isspas.nRs);
    if (configs.us == 1) {
        if (bias.initialize(configs.usCfgFileName)) return 1;
        bias.populate_mass(atoms.mass_h, atoms.nAtoms);
        printf("Total number of biased atoms: %d\n", bias.totalBiasAtoms);
    }
    if (check_only) return 0;
    if (configs.nMC < 1) {
        fprintf(stderr, "nMC must be >= 1\n");
        return 1;
    }

    atoms.read_initial_positions(configs.inputCoordFileName);
    // a resumed run (velRst = 1) continues the Philox counter of the run that
    // wrote the restart, so the MC points and Andersen draws of its steps are
    // those the uninterrupted run would have used (with the same seed)
    unsigned start = 0;
    if (configs.velRst == 1) {
        atoms.read_initial_velocities(configs.inputVelFileName);
        start = atoms.velRstStep;
        if (atoms.posRstStep != start)
            fprintf(stderr, "warning: position restart at step %u, velocity restart at step %u; "
                            "continuing from step %u\n", atoms.posRstStep, start, start);
        printf("Resuming at step %u\n", start);
    } else {
        srand((unsigned)configs.seed);
        atoms.initialize_velocities(configs.T);
    }
    atoms.open_traj_files(configs.forOutFileName, configs.posOutFileName, configs.velOutFileName);
    atoms.initialize_gpu(configs.seed);
    nonbond_force_cuda_grid_block(atoms, configs.rCut2, configs.lbox);
    bonds.initialize_gpu();
    bond_force_cuda_grid_block(bonds.nBonds, &bonds.gridSize, &bonds.blockSize, &bonds.minGridSize);
    angles.initialize_gpu();
    angle_force_cuda_grid_block(angles.nAngles, &angles.gridSize, &angles.blockSize, &angles.minGridSize);
    dihs.initialize_gpu();
    dih_force_cuda_grid_block(dihs.nDihs, &dihs.gridSize, &dihs.blockSize, &dihs.minGridSize);
    isspas.initialize_gpu(atoms.nAtoms, configs.seed);
    isspa_grid_block(atoms.nAtoms, atoms.nPairs, configs.lbox, isspas);
    isspas.step = atoms.step = start;
    if (configs.us == 1) {
        bias.initialize_gpu();
        us_grid_block(bias);
    }
    times.initialize();
    times.startWriteTimer();
    atoms.copy_params_to_gpu();
    atoms.copy_pos_vel_to_gpu();
    times.stopWriteTimer();

    const size_t fbytes = (size_t)atoms.nAtoms * sizeof(float4);
    md_engine *eng = nullptr;
    bonded_model *all_bonded = nullptr;
    void *stream = nullptr;
    int steps_per_graph = 1;
    if (graph) {
        // one combined bonded model and the lazily built term models
        std::vector<float> A(atoms.nTypes * (atoms.nTypes + 1) / 2), B(A.size());
        for (size_t k = 0; k < A.size(); ++k) {
            A[k] = atoms.lj_h[k].x;
            B[k] = atoms.lj_h[k].y;
        }
        isspa_check(bonded_model_create(atoms.nAtoms, bonds.nBonds, (const int *)bonds.bondAtoms_h,
                                        bonds.nTypes, (const float *)bonds.bondParams_h,
                                        angles.nAngles, (const int *)angles.angleAtoms_h,
                                        angles.nTypes, (const float *)angles.angleParams_h,
                                        dihs.nDihs, (const int *)dihs.dihAtoms_h, dihs.dihTypes_h,
                                        dihs.nTypes, (const float *)dihs.dihParams_h,
                                        (const float *)dihs.scaled14Factors_h, atoms.nTypes,
                                        atoms.ityp_h, atoms.nonBondedParmIndex_h, A.data(),
                                        B.data(), configs.lbox, &all_bonded),
                    "bonded_model_create");
        nonbond_force_cuda(atoms, isspas, atoms.nAtoms);  // builds the nonbond model
        if (configs.us == 1) us_force_cuda(atoms.pos_d, atoms.for_d, bias, configs.lbox, atoms.nAtoms);
        isspa_check(md_engine_create(isspas.model, atoms.nonbond, all_bonded,
                                     configs.us == 1 ? bias.model : nullptr, atoms.nAtoms,
                                     (float *)atoms.pos_d, (float *)atoms.vel_d,
                                     (float *)atoms.for_d, configs.T, configs.dt, configs.pnu,
                                     configs.lbox, (uint64_t)(uint32_t)configs.seed, 0u, &eng),
                    "md_engine_create");
        isspa_check(isspa_stream_create(&stream), "isspa_stream_create");
        isspa_check(md_engine_set_step(eng, start, stream), "md_engine_set_step");
        // several steps per graph replay when they divide the write interval,
        // so every write still falls on a replay boundary
        for (int k : {50, 20, 10, 5, 2})
            if (configs.deltaWrite % k == 0 && configs.nSteps >= k) {
                steps_per_graph = k;
                break;
            }
        isspa_check(md_engine_capture(eng, steps_per_graph, stream), "md_engine_capture");
    }

    for (int step = 0; step < configs.nSteps;) {
        const bool write = step % configs.deltaWrite == 0;
        if (graph) {
            if (write) {
                isspa_check(md_engine_forces(eng, stream), "md_engine_forces");
                // the engine keeps its own restraint centres of mass
                if (configs.us == 1)
                    isspa_check(md_engine_get_com(eng, bias.groupComPos_d, stream), "md_engine_get_com");
                isspa_check(isspa_stream_sync(stream), "sync");
                times.startWriteTimer();
                atoms.get_pos_vel_for_from_gpu();
                atoms.print_for();
                atoms.print_pos();
                atoms.print_vel();
                if (configs.us == 1) bias.print_cv(step, configs.lbox);
                times.stopWriteTimer();
            }
            if (configs.nSteps - step < steps_per_graph) {
                // the last steps, fewer than one replay: a one-step graph
                steps_per_graph = 1;
                isspa_check(md_engine_capture(eng, 1, stream), "md_engine_capture");
            }
            isspa_check(md_engine_replay(eng, 1, stream), "md_engine_replay");
            step += steps_per_graph;
            continue;
        }
        isspa_check(isspa_zero(atoms.for_d, fbytes, nullptr), "isspa_zero");
        times.bondTime += bond_force_cuda(atoms.pos_d, atoms.for_d, atoms.nAtoms, configs.lbox, bonds);
        times.angleTime += angle_force_cuda(atoms.pos_d, atoms.for_d, atoms.nAtoms, configs.lbox, angles);
        times.dihTime += dih_force_cuda(atoms, dihs, configs.lbox);
        times.isspaTime += isspa_force_cuda(atoms.pos_d, atoms.for_d, atoms.isspaf_d, isspas, atoms.nAtoms);
        times.nonbondTime += nonbond_force_cuda(atoms, isspas, atoms.nAtoms);
        if (configs.us == 1)
            times.usTime += us_force_cuda(atoms.pos_d, atoms.for_d, bias, configs.lbox, atoms.nAtoms);
        if (write) {
            times.startWriteTimer();
            atoms.get_pos_vel_for_from_gpu();
            atoms.print_isspaf();
            atoms.print_for();
            atoms.print_pos();
            atoms.print_vel();
            if (configs.us == 1) bias.print_cv(step, configs.lbox);
            times.stopWriteTimer();
        }
        times.leapFrogTime += leapfrog_cuda(atoms, configs);
        ++step;
    }
    if (graph) isspa_check(isspa_stream_sync(stream), "sync");

    times.startWriteTimer();
    atoms.write_rst_files(configs.posRstFileName, configs.velRstFileName, configs.lbox,
                          start + (unsigned)configs.nSteps);
    times.stopWriteTimer();
    times.print_final(configs.nSteps * configs.dtPs * 1.0e-3f);

    if (eng) md_engine_destroy(eng);
    if (all_bonded) bonded_model_destroy(all_bonded);
    if (stream) isspa_stream_destroy(stream);
    atoms.free_arrays();
    atoms.free_arrays_gpu();
    bonds.free_arrays();
    bonds.free_arrays_gpu();
    angles.free_arrays();
    angles.free_arrays_gpu();
    dihs.free_arrays();
    dihs.free_arrays_gpu();
    isspas.free_arrays();
    isspas.free_arrays_gpu();
    if (configs.us == 1) {
        bias.free_arrays();
        bias.free_arrays_gpu();
    }
    return 0;
}
