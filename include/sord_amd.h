/*
 * C ABI of the MI355X SORD solver library (coseis_amd/libsord_amd.so).
 *
 * The reference crosses its drop-in boundary as a process boundary: the
 * Python stager writes ``sord.in`` and runs ``./sord.x`` as a blocking child
 * process (cst/sord.py:538-561, cst/job.py:192-194); the Fortran program reads
 * the text (sord/input_output.f90:21-98), runs the init sequence and the time
 * loop (sord/sord.f90:24-47) and writes its output files into the current
 * directory.  These entry points replace that process with a library handle:
 * same input text, same output files, explicit error codes instead of
 * ``stop`` (which exits with status 0 under most compilers).
 *
 * All functions return 0 on success or one of the SORD_E* codes; the message
 * of the last error on the calling thread is returned by sord_last_error().
 * A handle is not re-entrant; distinct handles may run concurrently (each owns
 * one HIP stream on its device).
 */
#ifndef SORD_AMD_H
#define SORD_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SORD_OK 0
#define SORD_EINPUT 1       /* malformed sord.in or parameter (sord/input_output.f90:51-92, sord/setup.f90:190) */
#define SORD_ECFL 2         /* Courant condition not satisfied (sord/material_model.f90:168) */
#define SORD_EMODULI 3      /* negative moduli (sord/material_model.f90:167) */
#define SORD_EVOLUME 4      /* negative cell volume (sord/grid_generation.f90:192-198) */
#define SORD_ENAN 5         /* NaN/Inf in the solution (sord/statistics.f90:52-60) */
#define SORD_ENANIN 6       /* NaN/Inf in an input file (sord/input_output.f90:333-337) */
#define SORD_EBADOP 7       /* bad i/o op or time function (sord/input_output.f90:252-254, sord/utilities.f90:279-281) */
#define SORD_EFILE 8        /* file cannot be read or written (sord/fortran_io.f90:51-92) */
#define SORD_EHIP 9         /* HIP runtime error */
#define SORD_EUNSUPPORTED 10 /* input this build does not implement (named in the message) */
#define SORD_ESTATE 11      /* call out of order */

/* Parse a sord.in text (Appendix format of cst/sord.py:538-544) and bind the
 * handle to HIP device `device`.  realsize: 4 (float) or 8 (double), the
 * reference's REALSIZE (sord/Makefile.in:7).  Replaces read_parameters
 * (sord/input_output.f90:21-98) and setup_dimensions (sord/setup.f90:6-119). */
int sord_create(const char *sord_in_text, int realsize, int device, void **handle);

/* Initialisation at it = 0: grid, material, PML, rupture, static field i/o,
 * CFL and moduli checks, device upload.  Replaces sord/sord.f90:30-36. */
int sord_init(void *handle);

/* Advance `nsteps` time steps (or to the last step) with the per-step field
 * i/o and statistics.  One iteration of sord/sord.f90:39-47 per step. */
int sord_step(void *handle, int nsteps);

/* Run all remaining steps and flush every output file (sord/sord.f90:39-49). */
int sord_run(void *handle);

/* sord_step without device allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory while stepping (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * step's scratch (the two sets of statistics slots and the partial maxima of
 * the max-norm reduction, at most 1024 blocks) lives in a file-scope device
 * array of SCRATCH_BYTES (16.6 KB) that each device holds until the process
 * ends; output buffers are allocated by sord_init.  Same results as
 * sord_step, and the two may alternate on one handle.  Calls on one device
 * share that array: calls that overlap in time on one device (two threads,
 * two handles) race.  A step that would allocate (none does after sord_init)
 * fails with SORD_ESTATE; SORD_EHIP if the array's address cannot be taken. */
int sord_step_static(void *handle, int nsteps);

/* Current step (0 after init) and total steps. */
int sord_status(void *handle, int *it, int *nt);

/* Borrow a device view of a solver array (the handle keeps ownership; the
 * state vectors move between two buffers each step, so a view of them holds
 * until the next step).  name: "vx" "vy" "vz" "ux" "uy" "uz" (nodes), "wxx"..
 * "wxy" (cell stress of the last step), "lam" "mu" "yy" (cells), "mr" "gam"
 * (nodes); "rho", "vp", "vs" are not kept on the device.  The element
 * (j,k,l), 1-based local index over the padded mesh, is at
 *   ptr[offset + (j-1) + strides[1]*(k-1) + strides[2]*(l-1)].
 * shape[] receives the padded extents. */
int sord_field(void *handle, const char *name, void **dev_ptr, int64_t shape[3], int64_t strides[3],
               int64_t *offset);

/* Adopt a caller-owned device buffer (on the handle's device, `numel`
 * elements of realsize bytes, at least sord_tensor_numel's count) for a
 * solver array, in sord_field's layout.  Replaces the reference's arrays
 * allocated by sord/arrays.f90:6-34 for the caller's use (autograd, ensembles
 * driven from the caller's tensors).  The current contents (halo and padding
 * included) are copied into the buffer, which the kernels then read and
 * write; the caller keeps ownership and must keep it alive until
 * sord_destroy or until it adopts another.  Names: "vx" "vy" "vz" "ux" "uy"
 * "uz" (the buffer holding the current state), the same with ".next" (the
 * buffer the next step writes; the two exchange roles every step, and
 * sord_field reports the current one), "lam" "mu" "yy" (cells), "mr" "gam"
 * (nodes).  Material the caller edits between steps enters the stress, the
 * hourglass and Newton from the next step; the fault-plane moduli and the
 * Courant check stay those of sord_init.  Refused while the adjoint is set
 * up, and the adjoint is refused on a handle with adopted buffers. */
int sord_set_tensor(void *handle, const char *name, void *dev_ptr, int64_t numel);
/* Elements of one padded solver array (the size sord_set_tensor needs). */
int sord_tensor_numel(void *handle, int64_t *numel);

/* Advance `nsteps` steps (with their field i/o) timed with a pair of HIP
 * events on the library's stream around all of them.  Returns the elapsed
 * time of all steps (ms), 0 for ms_kernel (the launches are not instrumented
 * here: a launch that records its own start and end adds gaps to the step;
 * per-launch durations come from sord_bench_boundary), the number of nodes the
 * dominant solver kernel's launch advances, and the kernel's name.  No
 * counterpart in the reference: it is the measurement hook of the benchmark. */
int sord_bench(void *handle, int nsteps, double *ms_total, double *ms_kernel, int64_t *kernel_nodes,
               char *kernel_name, int name_len);

/* Advance nsteps steps as sord_bench does, with the interior launch and
 * every boundary launch stamping a pair of HIP events with the kernel's start
 * and end on the solver stream (one stream only: SORD_ESTATE with side
 * streams).  Returns the number of launches per step and, in [0, cap), each
 * one's damped-axis specialisation (entry 0: the interior launch, mask -1),
 * mean duration (ms), the nodes it advances and the sum over them of the axes
 * along which the node is damped (the PML memory it reads and writes).  No
 * counterpart in the reference: the measurement hook of the per-launch
 * rooflines. */
int sord_bench_boundary(void *handle, int nsteps, int *nlaunch, int *masks, double *ms, double *nodes,
                        double *damped, int cap);

/* Select the execution path: 1 fused interior kernel where it applies
 * (default), 0 stage kernels everywhere. */
int sord_set_fused(void *handle, int on);

/* Where sord_init computes the material model (sord/material_model.f90:6-171),
 * the resampling to nodes (sord/material_resample.f90:6-55) and the volumes
 * of rectilinear cells (sord/grid_generation.f90:183-198): 1 (the default) on
 * the device from the input arrays, 0 on the host.  Both give the same arrays
 * bit for bit.  Before sord_init only (SORD_ESTATE after).  No counterpart in
 * the reference, which initialises on its CPU. */
int sord_set_device_init(void *handle, int on);
/* side = 1: the boundary kernels run on three side streams beside the
 * interior kernel (forked and joined with events); side = 0 (the default):
 * in order on the handle's stream after the interior kernel.  The kernels
 * write disjoint nodes, but with side streams outputs were seen to differ
 * between processes on one machine, so one stream is the default.  No
 * counterpart in the reference (its loops run in order). */
int sord_set_streams(void *handle, int side);

/* Tile shape of the fused kernels: waves per block of the interior kernel
 * (4, 8 or 16; a block advances waves-2 node rows) and the planes one block
 * marches in the interior and boundary kernels; the face slabs of the PML
 * then get their own boundary launches at any mesh size.  Takes effect at
 * the next step; the results do not depend on it.  No counterpart in the reference
 * (its loops have no tiles): the hook of the autotuner. */
int sord_set_tiles(void *handle, int waves, int planes, int planes_boundary);

/* The current plan of the fused kernels: *nlaunch boundary launches per step
 * (-1: the stage kernels run the steps) and, in masks[0..cap), the damped-axis
 * specialisation of each (bit a: the kernel handles PML slabs of axis a; 7
 * is the general kernel).  No counterpart in the reference: a check hook. */
int sord_plan(void *handle, int *nlaunch, int *masks, int cap);

/* Copy nbytes (a multiple of 4) between two device buffers of `device` with a
 * plain kernel that moves one 4-byte word per lane per access, and wait for
 * it.  No counterpart in the reference: the calibration of the HBM traffic
 * counters (FETCH_SIZE, WRITE_SIZE) for the access width of the solver. */
int sord_copy_probe(const void *src, void *dst, int64_t nbytes, int device);

/* ---- batched ensembles ----
 * No counterpart in the reference, which runs one scenario per process
 * (cst/sord.py:558-561): n scenarios with their own sord.in texts (uneven
 * meshes allowed) advance together on one device.  Each step runs the
 * per-scenario phases of the time loop (i/o, sources, statistics) and one
 * launch of the interior kernel over the interior tiles of all scenarios and
 * one launch per boundary class, each tile naming its scenario.  Scenario q
 * writes its output files into dirs[q] (NULL or "" : the current directory).
 * Outputs are those of n separate sord_run calls. */
int sord_batch_create(int n, const char *const *sord_in_texts, const char *const *dirs, int realsize, int device,
                      void **batch);
/* Planes per work item of scenario `index` (interior, boundary).  Before
 * sord_batch_init it takes effect at init; afterwards, set the scenarios'
 * shapes and then call it with index -1 to re-plan the batch.  Performance
 * only (the autotuner's hook). */
int sord_batch_set_tiles(void *batch, int index, int planes, int planes_boundary);
int sord_batch_init(void *batch);
/* on = 1 (default): each step's merged launches (scenario table copy,
 * boundary and interior launches, stream fork and join) are captured once per
 * slot of the table ring into a HIP graph and replayed; 0: launched directly;
 * -1: query only.  *in_use (if not NULL) receives 1 while graphs are in use (a
 * capture that fails switches them off).  Results are the same either way. */
int sord_batch_set_graph(void *batch, int on, int *in_use);
/* side = 1: the merged boundary launches on three side streams beside the
 * interior launch; 0 (default): in order on the batch's stream, as
 * sord_set_streams for one handle.  Drops the graphs (recaptured at the
 * next step). */
int sord_batch_set_streams(void *batch, int side);
/* Advance every scenario that has steps left by up to nsteps steps. */
int sord_batch_step(void *batch, int nsteps);
int sord_batch_run(void *batch);
/* Advance nsteps steps timed with HIP events; returns the elapsed time (ms) and
 * the node-updates done (sum over the scenarios that stepped). */
int sord_batch_bench(void *batch, int nsteps, double *ms_total, int64_t *node_updates);
int sord_batch_size(void *batch, int *n);
void sord_batch_destroy(void *batch);

/* ---- adjoint gradient: no counterpart in the reference (the inversion
 * workload's misfit gradient with respect to the material).  The handle's
 * mesh must be rectilinear with uniform spacing, free-surface faces (bc 0 on
 * all six), no fault, material from inputs only and moment-tensor sources
 * (mxx..mxy field-I/O '.'/'+' ops); sord_init must have been called. */
/* Receivers: nrec nodes as 0-based (j, k, l) of the global mesh; checkpoint
 * the state every checkpoint_every steps (about sqrt(nt) is the usual
 * choice: memory and recomputation both grow as nt/every + every). */
int sord_adjoint_setup(void *handle, int nrec, const int32_t *rec_nodes, int checkpoint_every);
/* Before sord_adjoint_setup: how many of the forward run's last states to
 * keep resident in device memory, so that the backward sweep reads them
 * instead of recomputing their segments: -1 as many as fit in 98 % of the
 * free memory (less 1 GB), 0 none (the default: checkpoints only), n > 0 at
 * most n.  Rounded to whole segments.  No counterpart in the reference. */
int sord_adjoint_set_resident(void *handle, int64_t states);

/* After sord_adjoint_setup: the first resident state (1-based step; 0 when
 * none is kept) and the bytes they take. */
int sord_adjoint_resident(void *handle, int64_t *first_state, int64_t *bytes);

/* Run the remaining steps, storing checkpoints; traces (host, realsize
 * floats) receives [nt][nrec][3]: the velocity at each receiver at the
 * start of each step. */
int sord_adjoint_forward(void *handle, void *traces);
/* Given residual = dJ/dtraces (host, [nt][nrec][3]), accumulate the gradient
 * of J with respect to the cell density, P and S speeds into grad_rho,
 * grad_vp, grad_vs (host, realsize floats, one per cell, Fortran order over
 * the (nx-1, ny-1, nz-1) cells; overwritten). */
int sord_adjoint_backward(void *handle, const void *residual, void *grad_rho, void *grad_vp, void *grad_vs);
/* Wall time (ms, HIP events) of the last adjoint forward and backward. */
int sord_adjoint_timing(void *handle, double *ms_forward, double *ms_backward);

/* Message of the last error on this thread ("" if none). */
const char *sord_last_error(void);

/* Free the handle and its device memory; flushes nothing. */
void sord_destroy(void *handle);

#ifdef __cplusplus
}
#endif
#endif
