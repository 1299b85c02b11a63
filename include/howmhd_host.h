/*
 * howmhd_host.h — C-ABI of the host shell around the GPU time step: the input
 * parser, the problem registry, the output writers and the run loop of the
 * `how_mhd` driver.  These run on the CPU in the reference too; they are
 * restated here so that `how_mhd input.in` is a drop-in for
 * `./bin/how_mhd_mpi input.in` at one rank.
 */
#ifndef HOWMHD_HOST_H
#define HOWMHD_HOST_H

#include <stdint.h>
#include "howmhd_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Parsed input file.  Field names and defaults follow the reference
 * `Parameters` (src/parameters.hpp:10-76, src/parameters.cpp:7-33). */
typedef struct howmhd_run {
    int32_t gnx, gny, gnz;          /* nx|gnx, ny|gny, nz|gnz                  */
    int32_t nx, ny, nz;             /* local = global at one rank             */
    int32_t nt, nstep, ntime;       /* nstep > 0: stop after nstep hydro steps */
    int32_t px, py, pz;
    double xsize, ysize, zsize;
    double dx, dy, dz;
    double gam, cour, tend, rhomin, pgmin;
    char problem[64];
    char x1bc[32], x2bc[32], x3bc[32];
    char output_format[32];         /* dat | vtk | vts | none                  */
    /* keys added by this implementation (unknown to the reference parser) */
    char precision[16];             /* fp64 (default) | fp32                   */
    char indicator[16];             /* off (default) | mlp                     */
    char indicator_weights[256];    /* weights file for indicator mlp          */
    uint64_t seed;                  /* turbulence3d random phases               */
    int32_t quiet;                  /* 1: no per-step log lines                */
} howmhd_run;

/* read_parameters + finalize_parameters_after_read
 * (src/parameters.cpp:7-160).  `text` is the whole input file.  Returns 0, or
 * non-zero with a message in errbuf (the reference aborts instead). */
HOWMHD_API int howmhd_host_parse(const char* text, howmhd_run* run, char* errbuf, int errlen);

/* Time-step parameters for howmhd_gpu_create. */
HOWMHD_API int howmhd_host_gpu_params(const howmhd_run* run, howmhd_params* p, char* errbuf, int errlen);

/* init_problem (src/problem.cpp:39-59): fills the active cells of q0, an
 * 8 x (nx+6) x (ny+6) x (nz+6) array (k fastest) that the caller zeroed.
 * Unknown names fail and list the registry in errbuf. */
HOWMHD_API int howmhd_host_init_problem(const howmhd_run* run, double* q0, char* errbuf, int errlen);

/* Names of the registered problems, separated by '\n'. */
HOWMHD_API const char* howmhd_host_problem_names(void);

/* output (src/output.cpp:555-571): writes output/dumpNNNNNN_rank000000.{dat,vts}
 * under `dir` from stage-0 q in the reference layout.  format "none" writes
 * nothing. */
HOWMHD_API int howmhd_host_output(const howmhd_run* run, const double* q0, int step, double time,
                       const char* dir, char* errbuf, int errlen);

/* Reads a learned-indicator weights file (the `indicator_weights` input key;
 * no counterpart in the reference): text, '#' lines are comments, then the
 * header "howmhd-mlp <nparam> <threshold>" and nparam FP32 values.  nparam
 * must equal HOWMHD_MLP_NPARAM (1697) and the threshold lie in (0, 1). */
HOWMHD_API int howmhd_host_load_indicator(const char* path, float* params, int nparam, float* threshold,
                               char* errbuf, int errlen);

#ifdef __cplusplus
}
#endif

#endif
