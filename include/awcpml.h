/*
 * awcpml.h -- C ABI of the MI355X (gfx950) CPML acoustic time-stepping
 * library libawcpml.so.
 *
 * The library replaces the per-step part of AcousticWaveCPML.jl's xPU
 * solvers: the kernels update_psi_{x,y,z}!, update_p!, inject_sources!,
 * record_receivers! and the forward! driver with its (pold, pcur, pnew)
 * rotation (src/shared/acoustic_2D_xPU.jl:7-131,327-345 and
 * src/shared/acoustic_3D_xPU.jl:8-167,392-407).  The host-side setup that
 * the reference does in FP64 before its time loop (CFL dt, calc_Kab_CPML,
 * fact = dt^2 vel^2, Ricker wavelets, position snapping;
 * src/shared/acoustic_2D_xPU.jl:151-240) stays with the caller, which hands
 * its results to the plan below.  That is the seam a Julia
 * Acoustic{2,3}D_HIP submodule binds with ccall.
 *
 * Conventions
 *  - Every array is a plain host pointer, column-major with x fastest,
 *    exactly the memory order of the reference's Julia arrays.
 *  - Grid-node positions are 1-based, as in the reference
 *    (possrcs/posrecs, src/shared/acoustic_2D_xPU.jl:228-240).
 *  - All coefficients are passed in FP64; a plan of dtype AWCPML_F32 rounds
 *    them to FP32 once, on upload.
 *  - Every call returns an int status: AWCPML_OK (0) or a negative code;
 *    awcpml_last_error() gives the message of the calling thread's last
 *    failure.  Nothing aborts the process.
 *  - A plan owns its device buffers and one HIP stream; calls are
 *    synchronous for the caller.  A plan must not be used by two threads at
 *    once; distinct plans are independent.
 */
#ifndef AWCPML_H
#define AWCPML_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AWCPML_OK 0
#define AWCPML_ERR_ARG -1      /* invalid argument or shape              */
#define AWCPML_ERR_HIP -2      /* HIP runtime error                      */
#define AWCPML_ERR_NOMEM -3    /* device allocation failed               */
#define AWCPML_ERR_STATE -4    /* call out of order (e.g. no sources set) */
#define AWCPML_ERR_RANGE -5    /* a source/receiver outside the model    */

#define AWCPML_F64 0
#define AWCPML_F32 1

typedef struct awcpml_plan awcpml_plan;

/* Create a plan for `nshots` independent shots on one velocity model.
 * ndim 2 (nz ignored) or 3; n >= 3 and n >= halo+2 per axis; halo >= 0.
 * With n < 2*halo+2 the two CPML slabs of an axis overlap; a node in both
 * takes the low side's terms, as the reference's if/elseif does
 * (src/shared/acoustic_2D_xPU.jl:45-56).
 * dtype AWCPML_F64 reproduces the reference's Float64 arithmetic operation
 * for operation; AWCPML_F32 keeps state and arithmetic in FP32.
 * Replaces the array allocation of src/shared/acoustic_2D_xPU.jl:208-215
 * (3D: src/shared/acoustic_3D_xPU.jl:256-267). */
int awcpml_plan_create(int ndim, int64_t nx, int64_t ny, int64_t nz, int halo,
                       int dtype, int nshots, awcpml_plan **out);

int awcpml_plan_destroy(awcpml_plan *plan);

/* Model: fact = dt^2 .* vel.^2 (prod(n) values) and the inverse spacings
 * _dx,_dy(,_dz) and _dx2,_dy2(,_dz2)
 * (src/shared/acoustic_2D_xPU.jl:190-196; 3D: acoustic_3D_xPU.jl:236-245).
 * inv_d and inv_d2 hold ndim values each. */
int awcpml_set_model(awcpml_plan *plan, const double *fact,
                     const double *inv_d, const double *inv_d2);

/* CPML coefficients of one axis (0 = x, 1 = y, 2 = z), as returned by
 * calc_Kab_CPML "ongrd" (a_l, a_r, b_K_l, b_K_r: halo values) and
 * "halfgrd" (a_hl, a_hr, b_K_hl, b_K_hr: halo+1 values), after the free-top
 * override (src/utils.jl:19-57; src/shared/acoustic_2D_xPU.jl:175-185).
 * A side whose a's are all 0 and b_K's all 1 (the free top) is detected
 * and skipped by the kernels: its psi/xi stay exactly zero, as in the
 * reference. */
int awcpml_set_cpml(awcpml_plan *plan, int axis,
                    const double *a_l, const double *a_r,
                    const double *b_K_l, const double *b_K_r,
                    const double *a_hl, const double *a_hr,
                    const double *b_K_hl, const double *b_K_hr);

/* Sources of one shot: pos is nsrcs x ndim (column-major, 1-based grid
 * indices, possrcs of src/shared/acoustic_2D_xPU.jl:228-232), dt2srctf is
 * nt x nsrcs column-major (acoustic_2D_xPU.jl:221-226).  nt is the number
 * of time steps of the run and must be the same for all shots of a plan.
 * Sources at the same node are added in source order, like the reference's
 * plain-loop twin (src/shared/acoustic_2D.jl:81-88). */
int awcpml_set_sources(awcpml_plan *plan, int shot, int nsrcs,
                       const int64_t *pos, int64_t nt, const double *dt2srctf);

/* Receivers of one shot: pos is nrecs x ndim, 1-based
 * (posrecs, src/shared/acoustic_2D_xPU.jl:233-240). */
int awcpml_set_receivers(awcpml_plan *plan, int shot, int nrecs,
                         const int64_t *pos);

/* Zero the wavefields, psi/xi and traces and set the step counter to 0
 * (the @zeros state of acoustic_2D_xPU.jl:208-215). */
int awcpml_reset(awcpml_plan *plan);

/* Run time steps it+1 .. it+nsteps of every shot (the `for it=1:nt` loop of
 * src/shared/acoustic_2D_xPU.jl:327-345 / acoustic_3D_xPU.jl:392-407,
 * including receiver recording).  Fails with AWCPML_ERR_STATE past nt. */
int awcpml_advance(awcpml_plan *plan, int64_t nsteps);

/* One forward! call for an explicit step index `it` (1-based), including
 * the buffer rotation; the do_bench path of acoustic_2D_xPU.jl:261-271
 * times this with it = 1.  Does not move the step counter. */
int awcpml_forward(awcpml_plan *plan, int64_t it);

/* Number of time steps completed since the last reset. */
int awcpml_steps_done(const awcpml_plan *plan, int64_t *out);

/* Receiver traces of one shot, nt x nrecs column-major in FP64
 * (recs.seismograms = Array(traces), acoustic_2D_xPU.jl:441). */
int awcpml_get_traces(awcpml_plan *plan, int shot, double *out);

/* Copy one wavefield of a shot to the host in FP64, prod(n) values.
 * which = 1 is pcur, the field solve2D/solve3D returns
 * (acoustic_2D_xPU.jl:444) and the 3D do_save snapshot writes
 * (acoustic_3D_xPU.jl:480-498); 0 = pold, 2 = pnew. */
int awcpml_get_field(awcpml_plan *plan, int shot, int which, double *out);

/* Convenience for one-shot plans: reset, run nt steps, copy the traces
 * (nt x nrecs) and the final field; either output pointer may be NULL. */
int awcpml_run(awcpml_plan *plan, double *traces, double *pend);

/* Time `nlaunches` launches of the plan's stencil kernel alone (no source
 * or receiver work) between two HIP events on the plan's stream and return
 * the total in milliseconds.  The wavefield state advances without
 * sources; call awcpml_reset afterwards.  Measurement support, no
 * reference counterpart (the reference times whole forward! calls with
 * BenchmarkTools, acoustic_2D_xPU.jl:263-276). */
int awcpml_time_stencil(awcpml_plan *plan, int nlaunches, double *ms_total);

/* Time `nsteps` full steps (awcpml_advance) between two HIP events on the
 * plan's stream, in milliseconds.  Measurement support (see above). */
int awcpml_time_advance(awcpml_plan *plan, int64_t nsteps, double *ms_total);

/* Launch geometry of the stencil kernel: wavefronts per block along y
 * (1, 2, 4 or 8), the length of the z (3D) or y (2D) march per block, and
 * the software pipeline of the march: 1..3 = whole-node slots issued up to
 * prefetch-1 nodes ahead; 4, 5 = two-distance pipeline (HBM streams 3 or 4
 * nodes ahead, cached neighbour loads 1 or 2 nodes ahead).  0 keeps the
 * library's choice.  A block_y that
 * the current shot group does not divide resets the shot group to 1.  Used
 * by the autotuner; no reference counterpart (ParallelStencil picks its
 * launch shape itself). */
int awcpml_set_tiling(awcpml_plan *plan, int block_y, int march_len, int prefetch);
int awcpml_get_tiling(const awcpml_plan *plan, int *block_y, int *march_len,
                      int *prefetch);

/* Shots per block (1, 2, 4 or 8, dividing block_y): the wavefronts of a
 * block then march the same block_y/shot_group rows of shot_group
 * different shots, so the model (fact) rows they share are fetched once
 * per block.  Results do not depend on it.  Multi-shot extension, no
 * reference counterpart. */
int awcpml_set_shot_group(awcpml_plan *plan, int shot_group);
int awcpml_get_shot_group(const awcpml_plan *plan, int *shot_group);

/* Time blocking: with steps = 2, awcpml_advance and
 * awcpml_time_advance advance two time steps per stencil launch wherever
 * that applies, and one step otherwise (an odd remainder, or a plan with a
 * source on a Dirichlet node since its last reset, whose pressure there
 * must rotate through the three field buffers exactly as the reference's
 * does).  Results are bit-identical to one step per launch.  block_rows is
 * the number of wavefronts per block: in 3D 4, 8 or 16 (the block updates
 * block_rows-2 rows and recomputes the first step on the two outer ones),
 * in 2D 4 or 8 (each wavefront marches its own chunk of rows);
 * march_len the chunk along the march axis (z in 3D, y in 2D), prefetch
 * the march positions of loads each
 * wavefront keeps in flight (1..3); 0 keeps the library's choice.  The
 * first call with steps = 2 allocates one more field buffer and one more
 * psi and xi buffer per CPML side.  awcpml_get_time_block returns the
 * steps a launch advances now (1 or 2) and the effective block_rows,
 * march_len and prefetch.  A plan that ran two-step launches on its own is
 * reset before it joins a batch.  No reference counterpart (the reference launches one kernel per
 * time step, acoustic_2D_xPU.jl:327-345, acoustic_3D_xPU.jl:392-407). */
int awcpml_set_time_block(awcpml_plan *plan, int steps, int block_rows, int march_len,
                          int prefetch);
int awcpml_get_time_block(const awcpml_plan *plan, int *steps, int *block_rows,
                          int *march_len, int *prefetch);

/* Tiling of 3D two-step launches (enable != 0; the default for FP32 plans,
 * off for FP64 plans): blocks whose
 * rows and planes no CPML slab reaches own aligned 64-node x tiles, so
 * that every row they store covers whole cache lines; the other blocks own
 * 62-node tiles.  Applies only when no tile's outside x neighbour (x =
 * 64*t-1 or 64*t+64) is an interior node of an active x slab.  Results
 * are bit-identical either way.  Tuning and test support, no reference
 * counterpart. */
int awcpml_set_two_step_align(awcpml_plan *plan, int enable);
int awcpml_get_two_step_align(const awcpml_plan *plan, int *enable);

/* Time `nlaunches` two-step stencil launches alone (no source or receiver
 * work), in milliseconds; fails unless two-step launches apply.  The state
 * advances without sources; call awcpml_reset afterwards.  Measurement
 * support, no reference counterpart. */
int awcpml_time_stencil2(awcpml_plan *plan, int nlaunches, double *ms_total);

/* Force the kernels' 64-bit address offsets on any grid (enable != 0).
 * The library switches to them by itself for fields of 4 GiB and more per
 * shot; results are bit-identical either way.  Two-step launches do not
 * apply to such plans (they take single steps).  Test support for the
 * large-grid code path; no reference counterpart. */
int awcpml_set_wide_offsets(awcpml_plan *plan, int enable);

/* Capture the time loop in HIP graphs: with enable != 0, awcpml_advance and
 * awcpml_time_advance record each (first step, number of steps) sequence
 * of launches once and replay it afterwards, so a solve after
 * awcpml_reset costs one graph launch instead of two launches per step.
 * No counterpart in the reference (its loop launches every kernel from
 * Julia, acoustic_2D_xPU.jl:327-345). */
int awcpml_set_graph(awcpml_plan *plan, int enable);

/* ------------------------------------------------------------------ */
/* Batches: several plans advanced by one launch per step.             */
/* ------------------------------------------------------------------ */

/* Group 1..32 plans that share ndim, dtype and tiling (grids, halos,
 * models, shot counts and nt may all differ) so that every time step of
 * all of them is one stencil launch and one source/receiver launch (or every
 * two steps, see awcpml_batch_set_time_block).  The
 * plans stay owned by the caller and must be fully configured; a plan whose
 * nt is shorter than the batch's longest stops at its own nt.  Results are
 * read per plan with awcpml_get_traces / awcpml_get_field.  Extension
 * beyond the reference, which solves one model per call
 * (src/shared/acoustic_3D_xPU.jl:169). */
typedef struct awcpml_batch awcpml_batch;
int awcpml_batch_create(awcpml_plan *const *plans, int count, awcpml_batch **out);
int awcpml_batch_destroy(awcpml_batch *batch);
/* awcpml_reset on every plan; the batch starts again at step 0. */
int awcpml_batch_reset(awcpml_batch *batch);
/* Run the next nsteps steps of every plan (each stops at its own nt). */
int awcpml_batch_advance(awcpml_batch *batch, int64_t nsteps);
/* Graph capture of the batch's step sequences (see awcpml_set_graph). */
int awcpml_batch_set_graph(awcpml_batch *batch, int enable);
/* awcpml_batch_advance timed with HIP events on the batch's stream (ms). */
int awcpml_batch_time_advance(awcpml_batch *batch, int64_t nsteps, double *ms_total);
/* Time nlaunches batch stencil launches alone (ms); the plans' wavefields
 * advance without sources; call awcpml_batch_reset afterwards. */
int awcpml_batch_time_stencil(awcpml_batch *batch, int nlaunches, double *ms_total);

/* Time blocking for the whole batch: with steps = 2, each pair of steps of
 * every plan is one stencil launch (and two source/receiver launches) while
 * every plan of the batch has two steps left and two-step launches apply
 * to all of them (see awcpml_set_time_block); single steps otherwise.
 * block_rows (3D 4, 8 or 16; 2D 4 or 8) and prefetch (1..3; 0: the
 * library's choice) apply to the whole launch; each plan keeps its own
 * march_len.  Calls awcpml_set_time_block(plan, steps, block_rows,
 * <its march_len>, prefetch) on every plan, so the first call with steps = 2
 * allocates the plans' spare buffers.  Results are bit-identical to single
 * steps.  A batch starts with steps = 2 when all its plans have steps = 2
 * with the same block_rows and prefetch.  After a batch run the plans'
 * buffers may have traded roles; the plans read their results as usual.  awcpml_batch_get_time_block returns
 * the steps a launch advances now and the effective block_rows and
 * prefetch; awcpml_batch_time_stencil2 times nlaunches two-step launches
 * alone (ms; call awcpml_batch_reset afterwards).  No reference
 * counterpart (acoustic_3D_xPU.jl:392-407 launches one kernel per step of
 * one model). */
int awcpml_batch_set_time_block(awcpml_batch *batch, int steps, int block_rows, int prefetch);
int awcpml_batch_get_time_block(const awcpml_batch *batch, int *steps, int *block_rows,
                                int *prefetch);
int awcpml_batch_time_stencil2(awcpml_batch *batch, int nlaunches, double *ms_total);

/* ------------------------------------------------------------------ */
/* Differentiation: the FWI gradient by a hand-written adjoint.        */
/* ------------------------------------------------------------------ */

/* Reset the plan and run all nt steps, keeping a checkpoint of the forward
 * state (the three field buffers, psi, xi) every `every` steps (<= 0:
 * round(sqrt(nt))).  The traces of the run are then available as after
 * awcpml_advance.  Device memory: ceil(nt/every) x (3 fields + slabs).
 * Needs >= 2*halo+4 points on every absorbing axis.  The checkpoints stay
 * valid until the model, the CPML coefficients, the sources or the
 * receivers change.  This is the forward half of a reverse-mode rule: the
 * reference has no gradient yet (README.md:297 lists the adjoint and
 * checkpointing as work in progress); the Julia side would expose the pair
 * (this call and awcpml_adjoint) as a ChainRulesCore rrule of the solve. */
int awcpml_checkpoint_run(awcpml_plan *plan, int64_t every);

/* Pullback of the traces: given trace_bar = dJ/dtraces (for every shot in
 * order, that shot's receivers x nt values, time fastest -- the layout of
 * awcpml_get_traces; count = total receivers x nt), write fact_bar =
 * dJ/dfact summed over the shots (nx*ny*nz values, x fastest).  dt, the
 * CPML coefficients and the source wavelets are held fixed; for the
 * velocity, dJ/dvel = fact_bar * 2 * dt^2 * vel.  Per checkpoint segment
 * the forward is replayed (storing d(pnew)/d(fact) for each step) and then
 * stepped back in reverse time.  Receivers on Dirichlet nodes carry no
 * gradient (those nodes do not depend on fact).  Leaves the plan's forward
 * state at an intermediate step; the traces are unchanged. */
int awcpml_adjoint(awcpml_plan *plan, const double *trace_bar, int64_t count, double *fact_bar);

/* ------------------------------------------------------------------ */
/* Allocation-free solve.                                               */
/* ------------------------------------------------------------------ */

/* Bytes of the library's static device scratch array, one copy per GPU it
 * runs on, held until the process ends. */
#define AWCPML_NOALLOC_SCRATCH_BYTES (64ull << 20)

/* awcpml_run for one shot without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The wavefields, psi/xi and the source/receiver tables live
 * in the library's static scratch array; a problem needing more than
 * AWCPML_NOALLOC_SCRATCH_BYTES of it (3 fields plus the CPML slabs; with
 * halo 20: 2D up to 1631^2 in FP64 or 2323^2 in FP32, 3D up to 109^3 in
 * FP64 or 144^3 in FP32) fails with AWCPML_ERR_ARG; larger problems need a
 * plan (awcpml_plan_create).
 *
 * Device inputs, in the precision of `dtype` (8 or 4 bytes per value):
 *   fact      nx*ny*nz values, x fastest, no padding;
 *   coef      host array of ndim*8 device pointers: per axis a_l, a_r,
 *             b_K_l, b_K_r (halo values), a_hl, a_hr, b_K_hl, b_K_hr
 *             (halo+1 values), as for awcpml_set_cpml (unused if halo = 0);
 *   dt2srctf  nt x nsrcs values, column-major.
 * Host inputs: inv_d, inv_d2 (ndim FP64 values each); srcpos (nsrcs x ndim)
 * and recpos (nrecs x ndim), 1-based, column-major, as for
 * awcpml_set_sources / awcpml_set_receivers; active: ndim x 2 flags (axis
 * major) marking the CPML sides that absorb, or NULL for all of them
 * (a side whose coefficients are a = 0, b_K = 1 may be marked 0 to skip it,
 * as awcpml_set_cpml does; the results then differ at most in the sign of
 * zeros).
 * Device outputs: traces (nt x nrecs values, column-major) and pend (the
 * final pcur, nx*ny*nz values, or NULL).
 * stream: a hipStream_t; the call enqueues its work on it and returns
 * without waiting.  NULL: the null stream, and the call waits for the
 * result.  The results equal awcpml_run's on a one-shot plan with the same
 * inputs and one step per launch.  Every call on one GPU uses that GPU's
 * one scratch array: calls on one stream run in order, calls that overlap
 * in time on one GPU (other streams, other threads) race and corrupt each
 * other. */
int awcpml_run_noalloc(int ndim, int64_t nx, int64_t ny, int64_t nz, int halo, int dtype,
                       const void *fact, const double *inv_d, const double *inv_d2,
                       const void *const *coef, const int *active, int nsrcs,
                       const int64_t *srcpos, int64_t nt, const void *dt2srctf, int nrecs,
                       const int64_t *recpos, void *traces, void *pend, void *stream);

/* Precision of the gradient sum of awcpml_adjoint.  Each reverse step adds
 * lam * q of every shot of a shot group (awcpml_set_adjoint_group) to the
 * group's partial sum in the plan's precision; every `steps` reverse steps
 * (and at the end of every checkpoint segment) the partial sums of all
 * groups are added, in group order, to an FP64 total and reset.
 * steps = 0 (the default): once per checkpoint segment.  steps = -1: never;
 * the whole sum over time and shots stays in the plan's precision (the
 * cheapest, and for FP32 plans on long runs the least accurate).  Extension
 * beyond the reference, which has no gradient (README.md:297). */
int awcpml_set_gradient_fold(awcpml_plan *plan, int64_t steps);

/* Shots per thread of awcpml_adjoint's reverse step: a thread walks
 * `shots` consecutive shots at its node, so fact and the node's gradient
 * accumulator are read once per group and the group's contributions are
 * summed in registers.  0 (the default): the fewest groups whose launch
 * still fills the GPU, and a quarter of that group size for the blocks
 * holding CPML nodes, which are few.  awcpml_get_adjoint_group reports the
 * value in use (for the blocks without CPML nodes).
 * A tuning knob of this library's reverse step; the reference has no
 * gradient (README.md:297). */
int awcpml_set_adjoint_group(awcpml_plan *plan, int shots);
int awcpml_get_adjoint_group(const awcpml_plan *plan, int *shots);

/* Message of the calling thread's last failed call ("" if none). */
const char *awcpml_last_error(void);

/* Library version string. */
const char *awcpml_version(void);

#ifdef __cplusplus
}
#endif

#endif /* AWCPML_H */
