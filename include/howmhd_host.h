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
    int32_t nt, nstep, ntime;       /* parsed, unused (as in the reference)    */
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
    int32_t checkpoint_every;       /* > 0: checkpoint after every such output */
    char restart[256];              /* checkpoint file to resume from          */
    int32_t max_steps;              /* > 0: stop after max_steps hydro steps   */
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
 * header "howmhd-mlp <nparam> <threshold>" and nparam FP32 values, one per
 * line.  nparam must equal HOWMHD_MLP_NPARAM (1697) and the threshold lie in
 * (0, 1). */
HOWMHD_API int howmhd_host_load_indicator(const char* path, float* params, int nparam, float* threshold,
                               char* errbuf, int errlen);

/* Checkpoint of a run: q(0) and the stage-0 faces in the reference layout
 * (8 and 3 arrays of (nx+6) x (ny+6) x (nz+6) doubles, k fastest), the output
 * step reached, the hydro steps taken and the time.  No counterpart in the
 * reference, which has no checkpoint/restart; for runs longer than one job
 * (the `checkpoint_every` and `restart` input keys).  Binary, native byte
 * order, a 48-byte header ("HOWMHDCK", u32 version 1, i32 nx, ny, nz,
 * i32 out_step, 4 bytes of padding, i64 total_steps, f64 time) and then q0 and
 * the faces.  Writing creates or replaces `path` (its directory must exist);
 * reading checks the magic, the version and that nx, ny, nz are run's. */
HOWMHD_API int howmhd_host_write_checkpoint(const howmhd_run* run, const char* path, const double* q0,
                                 const double* b, int32_t out_step, int64_t total_steps, double time,
                                 char* errbuf, int errlen);
HOWMHD_API int howmhd_host_read_checkpoint(const howmhd_run* run, const char* path, double* q0, double* b,
                                int32_t* out_step, int64_t* total_steps, double* time, char* errbuf,
                                int errlen);

#ifdef __cplusplus
}
#endif

#endif
