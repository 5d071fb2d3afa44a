/*
 * solvcon_amd: C ABI of the MI355X CESE Euler march (libsolvcon_amd.so).
 *
 * Plain pointers and sizes only.  Every function that can fail returns a
 * status: 0 ok, 1 invalid argument (the reference's std::invalid_argument,
 * ValueError in Python), 2 out of range (std::out_of_range, IndexError),
 * 3 runtime error (std::runtime_error, RuntimeError), 4 HIP failure.  The
 * message of the last failure on the calling thread is read with
 * sc_last_error().  Host pointers are plain host memory; tables use the
 * reference's row-major layout with ghost rows first.
 *
 * The reference binds these operations through pybind11
 * (cpp/solvcon/onedim/pymod/wrap_onedim.cpp,
 * cpp/solvcon/mesh/pymod/wrap_StaticMesh.cpp,
 * cpp/solvcon/multidim/pymod/wrap_multidim.cpp); each declaration below
 * names the reference method it replaces.
 */

#ifndef SOLVCON_AMD_H
#define SOLVCON_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- library */

/* Message of the last failure on this thread; returns its status and
 * clears it. */
int sc_last_error(char * buf, int n);
/* Number of visible GPUs; 0 when there is none (never fails). */
int sc_device_count(void);
/* Name and architecture of the current device. */
int sc_device_name(char * buf, int n);
/* hipDeviceSynchronize. */
int sc_device_sync(void);
/* Stream-copy probe: copies nbytes between two device buffers reps times
 * and returns the average kernel time in ms (no reference counterpart; it
 * calibrates the HBM counters and the achievable bandwidth). */
int sc_stream_copy(int64_t nbytes, int reps, double * ms_avg);

/* ------------------------------------------------------ 1D Euler (onedim) */
/* Array ids for sc1d_*: 0 coord[n], 1 cfl[n], 2 so0[n*3], 3 so1[n*3],
 * 4 gamma[n]. */

/* Euler1DCore(ncoord, time_increment), Euler1DCore.hpp:71-75; even ncoord
 * is rejected (Euler1DCore.cpp:25-28). */
int sc1d_create(int64_t ncoord, double time_increment, void ** out);
/* Euler1DCore::clone, Euler1DCore.hpp:58-62. */
int sc1d_clone(void * h, void ** out);
void sc1d_destroy(void * h);
/* ncoord / time_increment, Euler1DCore.hpp:88-90. */
int sc1d_info(void * h, int64_t * ncoord, double * time_increment);
/* Device pointer of an array (coord/cfl/so0/so1/gamma accessors,
 * Euler1DCore.hpp:91-104). */
int sc1d_device_ptr(void * h, int which, double ** ptr, int64_t * count);
/* Host <-> device copies of one array (the numpy views of
 * wrap_onedim.cpp:89-104). */
int sc1d_upload(void * h, int which, double const * host, int64_t count);
int sc1d_download(void * h, int which, double * host, int64_t count);
/* Euler1DCore::update_cfl, Euler1DCore.cpp:70-94. */
int sc1d_update_cfl(void * h, int odd_plane);
/* Euler1DCore::march_half_so0, Euler1DCore.cpp:96-130. */
int sc1d_march_half_so0(void * h, int odd_plane);
/* Euler1DCore::march_half_so1_alpha<alpha>, Euler1DCore.hpp:284-320. */
int sc1d_march_half_so1(void * h, int alpha, int odd_plane);
/* Euler1DCore::treat_boundary_so0/so1, Euler1DCore.cpp:132-172. */
int sc1d_treat_boundary_so0(void * h);
int sc1d_treat_boundary_so1(void * h);
/* Euler1DCore::setup_march, Euler1DCore.hpp:126. */
int sc1d_setup_march(void * h);
/* march_half1_alpha<alpha> (second == 0) / march_half2_alpha<alpha>
 * (second != 0), Euler1DCore.hpp:322-337. */
int sc1d_march_half(void * h, int alpha, int second);
/* march_alpha<alpha>(steps), Euler1DCore.hpp:339-349. */
int sc1d_march(void * h, int alpha, int64_t steps);
/* The same march enqueued without waiting; sc1d_sync waits. */
int sc1d_march_async(void * h, int alpha, int64_t steps);
int sc1d_sync(void * h);
/* *done = 1 when the handle's stream is idle, else 0 (no reference
 * counterpart; for request loops that poll many solves). */
int sc1d_query(void * h, int * done);
/* density / velocity / pressure / temperature / internal_energy / entropy
 * (which = 0..5) into host_out[ncoord], Euler1DCore.hpp:106-168. */
int sc1d_derived(void * h, int which, double * host_out);

/* No counterpart in the reference.  march_alpha<alpha>(steps) of a tube
 * whose tables the caller holds in device memory (coord, gamma, cfl [n],
 * so0, so1 [n*3], row-major as above), enqueued on `stream` (a hipStream_t;
 * NULL for the default stream) without allocating anything: for callers
 * that cannot call hipMalloc (code under HIP graph capture, tight serving
 * loops).  The other buffer pair of the tiled march is a file-scope device
 * array of 2 x 3 x 131072 doubles (6 MB, held by every GPU the library runs
 * on until the process ends), so ncoord is at most 131071; larger tubes are
 * rejected with SC_EVALUE.  Every call on one GPU shares that array: calls
 * that overlap on one GPU race.  Results equal sc1d_march's bit for bit. */
int sc1d_march_noalloc(double const * coord, double const * gamma, double * so0, double * so1, double * cfl,
                       int64_t ncoord, double time_increment, int alpha, int64_t steps, void * stream);

/* -------------------------------------------------- mesh (StaticMesh) ---- */
/* Table ids for scmesh_*: 0 ndcrd, 1 fccnd, 2 fcnml, 3 fcara, 4 clcnd,
 * 5 clvol, 6 fctpn, 7 cltpn, 8 clgrp, 9 fcnds, 10 fccls, 11 clnds,
 * 12 clfcs, 13 ednds, 14 bndfcs (StaticMesh.hpp:433-451). */

/* StaticMesh(ndim, nnode, nface, ncell), StaticMesh.hpp:290-311. */
int scmesh_create(int ndim, int32_t nnode, int32_t nface, int32_t ncell, void ** out);
void scmesh_destroy(void * h);
/* out[10]: ndim nnode nface ncell nbound ngstnode ngstface ngstcell nedge
 * nbcs (StaticMesh.hpp:321-332). */
int scmesh_counts(void * h, int64_t * out);
/* info[4]: rows (ghost + body), columns (0 for a 1-D table), ghost rows,
 * 1 if int32 else double. */
int scmesh_table_info(void * h, int id, int64_t * info);
/* Device pointer of a table (its first ghost row). */
int scmesh_table_ptr(void * h, int id, void ** ptr);
/* Host <-> device copy of a whole table (nbytes must match). */
int scmesh_download(void * h, int id, void * host, int64_t nbytes);
/* Device bytes of the mesh's tables, ghost rows included (no reference
 * counterpart; admission by actual footprint in the request loop). */
int scmesh_device_bytes(void * h, int64_t * out);
int scmesh_upload(void * h, int id, void const * host, int64_t nbytes);
/* StaticMesh::build_interior, StaticMesh.hpp:363-374. */
int scmesh_build_interior(void * h, int do_metric, int do_edge);
/* StaticMesh::build_edge, StaticMesh_interior.cpp:566-610. */
int scmesh_build_edge(void * h);
/* StaticMesh::build_boundary, StaticMesh_boundary.cpp:103-182. */
int scmesh_build_boundary(void * h);
/* StaticMesh::build_ghost, StaticMesh_boundary.cpp:185-261. */
int scmesh_build_ghost(void * h);
/* StaticMesh::add_bc(name, faces), StaticMesh_boundary.cpp:14-78. */
int scmesh_add_bc(void * h, char const * name, int32_t const * faces, int64_t n);
/* StaticMesh::bc(ibc): name and facn columns 0/1 (either may be null);
 * *n receives the face count (StaticMesh.hpp:394-396). */
int scmesh_bc(void * h, int ibc, char * name, int nname, int32_t * facn0, int32_t * facn1, int64_t * n);

/* ------------------------------------------------- multidim (EulerCore) -- */
/* Table ids for sceuler_*: 0 cevol, 1 cecnd, 2 sfcnd, 3 sfnml, 4 so0c,
 * 5 so0n, 6 so0t, 7 so1c, 8 so1n, 9 stm, 10 cflo, 11 cflc, 12 gamma
 * (euler.hpp:104-117). */

/* EulerCore(mesh, time_increment), euler_ce.cpp:15-29; use_sfc orders the
 * cells along a Hilbert curve inside each cell type (no reference
 * counterpart; results are identical either way). */
int sceuler_create(void * mesh, double time_increment, int use_sfc, void ** out);
void sceuler_destroy(void * h);
/* out[4]: ndim ncell ngstcell neq (euler.hpp:90-94). */
int sceuler_counts(void * h, int64_t * out);
/* sigma0 (0), taumin (1), tauscale (2), time_increment (3),
 * euler.hpp:95-102; the reference fixes time_increment at construction,
 * setting it here (not on a batched solver) has no counterpart.  No
 * counterpart either: 5, the cell count of a rank chunk of the batched march
 * (0: all ranks at once); 6, 1 when march and march_batch replay each whole
 * step from a captured hipGraph, 0 (default) to launch its kernels one by
 * one; read-only 7, the number of exact passes the march's fast/exact
 * kernel pairs have run (a fast kernel met a value that is not finite or
 * lies near the ends of the exponent range; -1 if the count cannot be
 * read); read-only 4, 1 once the first substep has tuned the kernel forms
 * (else -1); read-only 100 + t, the form of the cells of type t: V for the
 * fused kernel in variant V (+ 100 when it reads so0t, + 1000 when it
 * prefetches its sub-face geometry into LDS), 10000 + 100 A + B for the
 * split pair in variants A and B (a variant is the minimum wavefronts per
 * SIMD plus 10 when LDS-staged), 0 when the mesh has no such cells. */
double sceuler_get_param(void * h, int which);
int sceuler_set_param(void * h, int which, double v);
/* info[4]: rows, elements per row, ghost rows, rank of the table. */
int sceuler_table_info(void * h, int id, int64_t * info);
/* Host <-> device copy of a whole table in the reference layout
 * (expose_SimpleArray, cpp/solvcon/python/common.hpp:347-383). */
int sceuler_download(void * h, int id, double * host, int64_t count);
int sceuler_upload(void * h, int id, double const * host, int64_t count);
/* One stage (euler.hpp:119-162): 0 prepare_ce, 1 update, 2 calc_solt,
 * 3 calc_soln, 4 calc_cfl, 5 calc_dsoln, 6 bc_soln, 7 bc_dsoln,
 * 8 march_substep, 9 clear_bc. */
int sceuler_stage(void * h, int stage);
/* EulerCore::march(steps), euler_march.cpp:350-359. */
int sceuler_march(void * h, int64_t steps);
/* The same march enqueued without waiting; sceuler_sync waits. */
int sceuler_march_async(void * h, int64_t steps);
int sceuler_sync(void * h);
/* *done = 1 when the handle's stream is idle, else 0 (no reference
 * counterpart). */
int sceuler_query(void * h, int * done);
/* Device bytes of the solver's own tables (no reference counterpart). */
int sceuler_device_bytes(void * h, int64_t * out);
/* dt that puts the largest calc_cfl number of the current so0n at cfl,
 * reduced on the device; sceuler_set_param(h, 3, dt) then applies it (no
 * reference counterpart: the reference takes dt from its caller). */
int sceuler_cfl_dt(void * h, double cfl, double * dt);
/* Batched solves (no reference counterpart; the per-instance march is
 * EulerCore::march, euler_march.cpp:339-359): an EulerCore over a
 * disjoint-union mesh of ninst instances.  inst_of_cell[ncell] is the
 * instance of every body cell, steps[ninst] the step count of every
 * instance. */
int sceuler_create_batch(void * mesh, int64_t ninst, int32_t const * inst_of_cell, int64_t const * steps, int use_sfc,
                         void ** out);
/* Time increment of every instance (by instance id). */
int sceuler_batch_set_dt(void * h, double const * dt);
/* Set and return in dt_out[ninst] the dt of every instance that puts its
 * largest initial calc_cfl number at cfl. */
int sceuler_batch_cfl_dt(void * h, double cfl, double * dt_out);
/* March every instance its own step count; *ms_total (if not null)
 * receives the device time of the march. */
int sceuler_march_batch(void * h, double * ms_total);
/* Completion time of every instance (by instance id) in the last timed
 * sceuler_march_batch: milliseconds from the march's start until the
 * instance's last step ran, on the device clock (n == ninst; no reference
 * counterpart). */
int sceuler_batch_done_ms(void * h, double * out, int64_t n);
/* march(steps) with HIP-event timing: the average duration of one
 * substep's fused kernels and the whole march, in milliseconds (no
 * reference counterpart; used by the benchmark). */
int sceuler_march_timed(void * h, int64_t steps, double * ms_substep, double * ms_total);
/* EulerCore::init_solution(gamma, rho, v[nv], p), euler_solution.cpp:54-99
 * and the velocity-size check of wrap_multidim.cpp:276-301. */
int sceuler_init_solution(void * h, double gamma, double rho, double const * v, int nv, double p);
/* EulerCore::add_bc(kind, faces, value): kind 0 non-reflective, 1 slip
 * wall, 2 inlet with value = [rho, v(ndim), p, gamma]
 * (euler_bc.cpp:324-391). */
int sceuler_add_bc(void * h, int kind, int32_t const * faces, int64_t n, double const * value, int64_t nv);
/* EulerCore::get_normal_matrix(ifc) into out[ndim*ndim] row-major
 * (euler_bc.cpp:393-425). */
int sceuler_normal_matrix(void * h, int32_t ifc, double * out);
/* FP64 device reductions over the body cells: out[3] = sum of so0n density
 * times cell volume, min cflo, max cflo (the sums of
 * solvcon/multidim/euler/_field.py:138-148). */
int sceuler_reduce(void * h, double * out);

/* GradientElement(mesh, cecnd, icl, tau), GradientElement.cpp:137-213,
 * over a host cecnd table [nrow][ncol].  iout[34]: ndim, clnfc, nfge,
 * valid, rcl[6], faces[8][3]; dout[37]: nfge_inverse, idis[6][3],
 * jdis[6][3]. */
int scge_build(void * mesh, double const * cecnd, int64_t nrow, int64_t ncol, int64_t nghost, int32_t icl, double tau,
               int32_t * iout, double * dout);
/* GradientElement::solve_gradient on a row-major dst[3][3],
 * GradientElement.cpp:447-458. */
int scge_solve(void * mesh, int ndim, double const * dst9, double const * udf3, double * out3);

/* ------------------------------------------- request loop (serving) ----- */
/* No reference counterpart.  The reference runs one EulerCore::march
 * (euler_march.cpp:350-359) or Euler1DCore::march_alpha<2>
 * (Euler1DCore.hpp:339-349) per problem; a request loop calls these to
 * march many independent problems together, each with the same arithmetic
 * and results as its plain march.
 *
 * scserve_*: a resident arena of one dimension (2 or 3) with fixed
 * capacities: type_cap[t] cells of cell type t (t = 0..8, StaticMesh type
 * ids; 0 for none), ghost_cap ghost rows and max_slots solves.  Each
 * substep is one fused launch per cell type over the arena. */
int scserve_create(int ndim, int64_t const * type_cap, int64_t ghost_cap, int64_t max_slots, void ** out);
void scserve_destroy(void * a);
/* Admit a set-up solver (sceuler_create, then its state, boundary
 * conditions and dt) for `steps` whole steps: its tables are copied into
 * free ranges of the arena at the next step boundary.  *slot is -1 when
 * its ranges or a slot do not fit now.  The solver may be destroyed after
 * the call (its stream waits for the copies). */
int scserve_admit(void * a, void * euler, int64_t steps, int64_t * slot);
/* Enqueue `steps` whole steps of every admitted solve with steps left; a
 * solve stops after its own last step and its so0n / so1n are drained to
 * pinned host memory.  Returns without waiting. */
int scserve_step(void * a, int64_t steps);
/* Chunks of steps enqueued and not yet completed. */
int scserve_pending(void * a, int64_t * chunks);
/* *done = 1 when the slot's solve has finished and been drained. */
int scserve_query(void * a, int64_t slot, int * done);
/* so0n [ncell][ndim + 2] and so1n [ncell][ndim + 2][ndim] of a finished
 * solve in its mesh's cell order (either may be NULL); frees the slot. */
int scserve_result(void * a, int64_t slot, double * so0, double * so1);
int scserve_sync(void * a);
/* out[8]: used slots, marching slots, device bytes, cell capacity, ghost
 * capacity, free cells, free ghost rows, slot capacity. */
int scserve_info(void * a, int64_t * out);

/* scserve1d_*: the 1D engine.  Admitted tubes (sc1d_create handles, set
 * up) march together, one tiled launch per round for all of them, each
 * with its own alpha, dt and step count. */
int scserve1d_create(int64_t max_slots, void ** out);
void scserve1d_destroy(void * a);
/* Take over a set-up tube for `steps` steps of march_alpha<alpha>: on
 * success the engine owns the handle (do not destroy it); *slot is -1
 * when every slot is in use. */
int scserve1d_admit(void * a, void * core, int alpha, int64_t steps, int64_t * slot);
/* One round: each marching tube advances up to its tiled step count. */
int scserve1d_step(void * a);
int scserve1d_pending(void * a, int64_t * chunks);
int scserve1d_query(void * a, int64_t slot, int * done);
/* so0 [n*3], so1 [n*3] and cfl [n] of a finished tube (any may be NULL);
 * frees the slot and the tube. */
int scserve1d_result(void * a, int64_t slot, double * so0, double * so1, double * cfl);

/* --------------------------------------------- spacetime scalar solvers -- */
/* Array ids for scst_*: 0 xcoord[nx], 1 so0[nx*nvar], 2 so1[nx*nvar],
 * 3 cfl[nx], with nx = 2*ncelm + 5 (cpp/solvcon/spacetime/core.cpp:48-98).
 * Kernel kinds: 0 the default hooks of Kernel::reset (core.hpp:520-536),
 * 1 LinearScalar (kernel/linear_scalar.hpp), 2 InviscidBurgers
 * (kernel/inviscid_burgers.hpp).  The pybind11 surface replaced is
 * cpp/solvcon/spacetime/pymod/wrap_spacetime.{hpp,cpp}. */

/* Grid(xmin, xmax, nelm) coordinates into xc[nx], core.cpp:14-45 (host
 * code; ValueError texts as the reference's). */
int scst_grid(double xmin, double xmax, int64_t ncelm, double * xc, int64_t nx);
/* Grid(xloc) coordinates into xc[nx], Grid::init_from_array,
 * core.cpp:48-98. */
int scst_grid_from(double const * xloc, int64_t n, double * xc, int64_t nx);
/* Solver / LinearScalarSolver / InviscidBurgersSolver(grid, time_increment
 * [, nvar]), core.hpp:740-745,1038-1056, linear_scalar.hpp:27-36 and
 * inviscid_burgers.hpp:32-41 construct(). */
int scst_create(int kind, double const * xc, int64_t nx, int64_t nvar, double time_increment, void ** out);
/* SolverBase::clone (core.hpp:717-731). */
int scst_clone(void * h, void ** out);
void scst_destroy(void * h);
/* time_increment setter, Field::set_time_increment (core.cpp:109-114). */
int scst_set_time_increment(void * h, double time_increment);
/* Whole-array copies (the so0/so1/cfl/xcoord SimpleArray views). */
int scst_upload(void * h, int which, double const * host, int64_t count);
int scst_download(void * h, int which, double * host, int64_t count);
/* One pass (core.hpp:946-1024): 0 march_half_so0, 1 march_half_so1_alpha,
 * 2 update_cfl, 3 treat_boundary_so0, 4 treat_boundary_so1,
 * 5 march_half1_alpha, 6 march_half2_alpha; alpha 0, 1 or 2. */
int scst_stage(void * h, int stage, int alpha, int odd_plane);
/* march_alpha<alpha>(steps), core.hpp:1026-1036. */
int scst_march(void * h, int alpha, int64_t steps);
int scst_march_async(void * h, int alpha, int64_t steps);
int scst_sync(void * h);
/* One element value at xindex (Selm::xn/xp/tn/tp/so0p/update_cfl,
 * Celm::calc_so0/calc_so1_alpha, core.hpp:458-672): what 0 xn, 1 xp, 2 tn,
 * 3 tp, 4 so0p, 5 cfl (store != 0 writes it), 6 calc_so0, 7 calc_so1. */
int scst_eval(void * h, int what, int alpha, int64_t xindex, int64_t iv, int store, double * out);
/* get_so0p(iv, odd_plane) into host_out[nselm - odd], core.hpp:812-824. */
int scst_so0p(void * h, int64_t iv, int odd_plane, double * host_out, int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* SOLVCON_AMD_H */
