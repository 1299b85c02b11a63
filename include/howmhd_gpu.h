/*
 * howmhd_gpu.h — C-ABI of the MI355X (gfx950) HOW-MHD time-step library.
 *
 * The library owns the device state of one single-GPU run: the conserved
 * variables q = (rho, Mx, My, Mz, Bx, By, Bz, E) of every SSPRK(5,4) stage and
 * the face-centred magnetic field (bxb, byb, bzb) used by the 4th-order
 * constrained transport.  Host arrays cross this boundary in the reference
 * layout: q(m,i,j,k) with k fastest, padded by 3 ghost cells per side, i.e.
 * 8 x (nx+6) x (ny+6) x (nz+6) doubles (src/main.cpp:114-120 without the stage
 * index).  On the device only active cells are stored; ghost values are the
 * periodic or open images of active cells, exactly as src/bound.cpp and the
 * serial paths of src/fluxct.cpp leave them.
 *
 * Conventions: every call returns 0 on success and a non-zero HOWMHD_E* code
 * on failure; howmhd_gpu_last_error() then holds a message.  Calls are ordered
 * on the context's stream and synchronise only where a host value is produced
 * (tstep, downloads, diag).  A context belongs to one host thread.
 */
#ifndef HOWMHD_GPU_H
#define HOWMHD_GPU_H

#include <stdint.h>

/* Only the functions declared in these headers are exported; the libraries
 * are built with hidden visibility and bind their own symbols (-Bsymbolic),
 * so the product build and its FP-contraction-off twin can be loaded into
 * one process side by side. */
#define HOWMHD_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

enum {
    HOWMHD_OK = 0,
    HOWMHD_EINVAL = 1,   /* bad argument or unsupported configuration   */
    HOWMHD_EHIP = 2,     /* a HIP runtime call failed                   */
    HOWMHD_ENOMEM = 3,   /* device allocation failed                    */
    HOWMHD_ESTATE = 4    /* call out of order (e.g. ssprk before faces) */
};

enum { HOWMHD_BC_PERIODIC = 0, HOWMHD_BC_OPEN = 1 };

enum {
    HOWMHD_PREC_FP64 = 0,      /* plain method in FP64 with the reference's rounding:
                                  no contraction to FMA, true divisions, the
                                  reference's evaluation order (bit-for-bit equal
                                  to a reference CPU run built without FMA)       */
    HOWMHD_PREC_FP32 = 1,      /* sweeps and CT EMFs in FP32 (with FMA); stage sums,
                                  E->p, floors and the face-B accumulation in FP64 */
    HOWMHD_PREC_FP64_FAST = 2  /* FP64 with FMA contraction and reciprocal forms:
                                  same method, different rounding               */
};

enum {
    HOWMHD_INDICATOR_OFF = 0, /* characteristic WENO5 at every interface            */
    HOWMHD_INDICATOR_MLP = 1  /* learned smoothness indicator routes interfaces it
                                 calls smooth to linear upwind-5 reconstruction   */
};

/* Run parameters.  Mirrors the fields of the reference `Parameters`
 * (src/parameters.hpp:10-76) that the time step reads. */
typedef struct howmhd_params {
    int32_t nx, ny, nz;      /* active cells (one rank: gnx, gny, gnz)           */
    int32_t bc[3];           /* x1bc, x2bc, x3bc: HOWMHD_BC_*                    */
    double dx, dy, dz;       /* xsize/gnx ... (src/parameters.cpp:149-151)        */
    double gam, cour, rhomin, pgmin;
    int32_t precision;       /* HOWMHD_PREC_*                                     */
    int32_t indicator;       /* HOWMHD_INDICATOR_*                                */
} howmhd_params;

/* Diagnostics of stage-0 state, reduced on the device in FP64. */
typedef struct howmhd_diag {
    double sum_cons[8];      /* sum over active cells of each conserved variable */
    double max_divb_ct;      /* max |sum_d D4_d(b_d)/Delta_d| over active cells   */
    double max_divb_cell;    /* max |2nd-order cell-centred divB| (output.cpp:314-347) */
    double min_rho;          /* min over active cells of rho                      */
    double min_p;            /* min of (gam-1)(E - 0.5 rho v^2 - 0.5 B^2), formed as
                                src/output.cpp:300-312 forms the pressure output  */
    /* indicator: flagged interfaces / interfaces of the most recent step that
     * routed (0 when no step has routed since set_indicator) */
    double flagged_fraction;
    /* cells whose density (rho < rhomin) and pressure (p < pgmin) floors the
     * prot of the last step's final stage applied (src/prot.cpp:31, 41; a
     * value that is not a number counts too, as fmax floors it); 0 before the
     * first step after upload_q0 */
    int64_t n_rho_floor, n_p_floor;
} howmhd_diag;

/* Per-kernel device time of the launches issued while timing was on,
 * measured with HIP events on the context's stream. */
enum {
    HOWMHD_K_SWEEP_X = 0, HOWMHD_K_SWEEP_Y, HOWMHD_K_SWEEP_Z, HOWMHD_K_EMF2, HOWMHD_K_EMF,
    HOWMHD_K_FACES, HOWMHD_K_COMBINE, HOWMHD_K_CFL, HOWMHD_K_COUNT
};
typedef struct howmhd_timing {
    double ms[HOWMHD_K_COUNT];         /* summed kernel time, milliseconds */
    int32_t launches[HOWMHD_K_COUNT];  /* number of launches timed        */
} howmhd_timing;

typedef struct howmhd_gpu_ctx howmhd_gpu_ctx;

/* Allocates q (5 stage buffers), faces (5 stage buffers), the sweep / CT
 * scratch and a staging buffer for host transfers on the current HIP device;
 * no later call allocates device memory.  Replaces the allocation of q
 * (src/main.cpp:114-120) and allocate_bfield_ct (src/fluxct.cpp:27-35).
 * Axes must have n == 1 (degenerate) or n >= 3; ny and nz at most 65535
 * (launch grid limit), at most 2^40 cells.  Bad parameters return
 * HOWMHD_EINVAL before any HIP call. */
HOWMHD_API int howmhd_gpu_create(const howmhd_params* p, howmhd_gpu_ctx** out);

/* Frees everything.  Replaces deallocate_bfield_ct (src/fluxct.cpp:37-42). */
HOWMHD_API int howmhd_gpu_destroy(howmhd_gpu_ctx* ctx);

/* Copies stage-0 q from the host (reference layout, ghosts ignored).  The
 * reference's problem initialisers write q(0,...) in place
 * (src/problem.cpp:39-59); this is the device hand-over of that array.
 * Host transfers (this call, upload_faces and the downloads) read or write
 * the caller's memory directly, one variable at a time, through the
 * context's device buffer of one variable's reference-layout size; the
 * layout change runs on the device.  Synchronises. */
HOWMHD_API int howmhd_gpu_upload_q0(howmhd_gpu_ctx* ctx, const double* q0);

/* bound(q,0) + initialize_bfield_from_q + bound_bfield(0)
 * (src/main.cpp:127-130, src/fluxct.cpp:65-107). */
HOWMHD_API int howmhd_gpu_init_faces(howmhd_gpu_ctx* ctx);

/* Stage-0 faces from the host, 3 x (nx+6) x (ny+6) x (nz+6) (bxb, byb, bzb),
 * for restart or tests; ghosts ignored. */
HOWMHD_API int howmhd_gpu_upload_faces(howmhd_gpu_ctx* ctx, const double* b);

/* CFL step: dt = cour / sum_d max(vmax_d,1e-300)/Delta_d over active cells of
 * q(0) (src/tstep.cpp:11-136).  Synchronises. */
HOWMHD_API int howmhd_gpu_tstep(howmhd_gpu_ctx* ctx, double* dt);

/* The CFL dt of howmhd_gpu_tstep for a q(0) the caller holds in device
 * memory, with no context and no device allocation: q0_dev is 8 x (nx+6) x
 * (ny+6) x (nz+6) doubles in the reference layout (k fastest, ghosts not
 * read); p gives the grid, dx/dy/dz, gam, cour and the floors; stream is a
 * hipStream_t (NULL: the default stream), synchronised before returning.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc); it
 * synchronises its stream to return dt, so it is not itself for use inside
 * a stream capture.  The per-block maxima live in a fixed 24.6 KB
 * file-scope device array (one per GPU, held until the process ends), so
 * the result equals howmhd_gpu_tstep's bit for bit on every grid
 * howmhd_gpu_create accepts; other grids (and more than 2^40 active cells)
 * return HOWMHD_EINVAL before any HIP call, with the reason in
 * howmhd_gpu_last_error(NULL).  All calls on one GPU share that array: calls
 * that overlap on one GPU (other host threads, other streams) race. */
HOWMHD_API int howmhd_gpu_tstep_noalloc(const howmhd_params* p, const double* q0_dev, void* stream, double* dt);

/* One SSPRK(5,4) step of size dt (src/ssprk.cpp:1277-1329).  HOWMHD_EINVAL
 * for a negative or non-finite dt; HOWMHD_ESTATE before init_faces or
 * upload_faces, or with indicator mlp before set_indicator. */
HOWMHD_API int howmhd_gpu_ssprk(howmhd_gpu_ctx* ctx, double dt);

/* q(0) to the host in the reference layout, ghosts filled with the
 * periodic/open images (what src/bound.cpp leaves there).  Synchronises. */
HOWMHD_API int howmhd_gpu_download_q0(howmhd_gpu_ctx* ctx, double* q0);

/* Stage-0 faces to the host, 3 x (nx+6) x (ny+6) x (nz+6): bxb, byb, bzb
 * (src/fluxct.cpp:10-12 at stage 0).  Synchronises. */
HOWMHD_API int howmhd_gpu_download_faces(howmhd_gpu_ctx* ctx, double* b);

/* Diagnostics of the stage-0 state: the cell-centred divB of the reference's
 * output (src/output.cpp:154-188, 314-347), the CT face divergence, conserved
 * sums, minima and floor counts.  Two kernels reduce them in FP64 with a
 * fixed grid and a fixed combining order, so the same state gives the same
 * bits; only the 13 results cross to the host.  Synchronises. */
HOWMHD_API int howmhd_gpu_diag(howmhd_gpu_ctx* ctx, howmhd_diag* d);

/* Starts (on != 0, counters reset) or stops per-kernel timing.  No
 * counterpart in the reference, which has no timers (src/ has no clock). */
HOWMHD_API int howmhd_gpu_set_timing(howmhd_gpu_ctx* ctx, int on);

/* Resolves the timed launches into *t.  Synchronises. */
HOWMHD_API int howmhd_gpu_get_timing(howmhd_gpu_ctx* ctx, howmhd_timing* t);

/* Waits for all work queued on the context's stream. */
HOWMHD_API int howmhd_gpu_synchronize(howmhd_gpu_ctx* ctx);

/* Message of the last failing call on ctx (or, when ctx is NULL, of the
 * calling thread's last failing create or howmhd_gpu_tstep_noalloc).  Never
 * NULL. */
HOWMHD_API const char* howmhd_gpu_last_error(const howmhd_gpu_ctx* ctx);

/* Build identification: the target and the rounding of the precision forms
 * ("gfx950: fp64 reference rounding; fp64fast and fp32 with FMA
 * contraction"). */
HOWMHD_API const char* howmhd_gpu_build_kind(void);

#ifdef __cplusplus
}
#endif

#endif /* HOWMHD_GPU_H */
