/* C ABI of the MI355X LGF fast-multipole path (libiblgf_gpu.so).
 *
 * The reference (IBLGF-AMR) has no C ABI: the path sits behind compile-time
 * C++ seams.  Each entry point below names the reference interface it stands
 * in for (file:line in IBLGF-AMR).  Conventions:
 *   - every function returns an int status, IBLGF_OK (0) on success; no C++
 *     exception crosses the boundary; iblgf_gpu_last_error() / iblgf_last_error()
 *     return the message of the last failure;
 *   - one host thread per context; all device work is ordered on the
 *     context's stream and finished when the call that reads results back
 *     returns (field_download, sync);
 *   - device memory is owned by the context; host arrays are borrowed for the
 *     duration of the call.
 * Field layout: one FP64 block of n^3 values per octant with data (n = block
 * extent + 2 buffer cells), x fastest, blocks in slot order (level, then
 * Morton key) — the DataField layout of
 * include/iblgf/domain/dataFields/datafield.hpp:166-197.
 */
#ifndef IBLGF_GPU_H
#define IBLGF_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum iblgf_status
{
    IBLGF_OK = 0,
    IBLGF_ERR_ARG = 1,
    IBLGF_ERR_STATE = 2,
    IBLGF_ERR_DEVICE = 3,
    IBLGF_ERR_UNSUPPORTED = 4,
    IBLGF_ERR_INTERNAL = 5
};

typedef struct iblgf_tree    iblgf_tree;
typedef struct iblgf_gpu_ctx iblgf_gpu_ctx;

/* Refinement condition: return 1 to refine the block where the 2:1 balance
 * allows it (Tree::try_2to1), 2 to refine it unconditionally (what
 * Domain::init_refine does for blocks that overlap an immersed-boundary
 * point, domain.hpp:887-891), 0 to keep it.  block_base is the block's lower
 * corner in cells of its own refinement level.  Stands in for
 * Domain::register_refinement_condition (include/iblgf/domain/domain.hpp:890-900,
 * tests/poisson/vortexrings.hpp:471-520). */
typedef int (*iblgf_refine_fn)(void* user, const int block_base[3], int block_extent, int refinement_level,
    int diff_level);

/* ---- octree (host) ------------------------------------------------------ */

/* Domain::initialize + construct_tree (include/iblgf/domain/domain.hpp:106-211,
 * 406-487): base mesh from nregions boxes {base xyz, extent xyz} in base cells. */
int iblgf_tree_create(const int bd_base[3], const int bd_extent[3], int block_extent, double dx_base,
    int nregions, const int* regions, iblgf_tree** out);

/* Tree handed over by the caller (the reference's own server tree): the keys
 * of the octants with data (Octant::has_data) and their flags (bit0 leaf,
 * bit1 correction), any order.  Ancestors without data are created as
 * Tree::insert does (tree.hpp:254-268); lists are rebuilt from the keys
 * (Tree::construct_lists, tree.hpp:559-603).  base_shift = bounding-box base /
 * block extent, as Domain::construct_tree's base_ (domain.hpp:169-190). */
int iblgf_tree_import(const int base_shift[3], int block_extent, int base_level, double dx_base, int nkeys,
    const uint64_t* keys, const uint32_t* flags, iblgf_tree** out);

/* Domain::init_refine (domain.hpp:878-943): refinement with 2:1 balance,
 * mark_correction (domain.hpp:617-735) and mark_leaf_boundary. */
int iblgf_tree_init_refine(iblgf_tree* t, int nref, int level_up_max, iblgf_refine_fn fn, void* user);

/* FmmMaskBuilder::fmm_lgf_mask_build (include/iblgf/fmm/fmm.hpp:392-400),
 * and the IB2xIB / xIB2IB masks (fmm.hpp:70-144).  With immersed-boundary
 * points set, the is_ib / extended_ib flags are computed first
 * (Server::update_ib_flag, include/iblgf/domain/decomposition/server.hpp:544-575).
 * Masks built with subtract_non_leaf = 1 can be exported, but
 * iblgf_gpu_tree_upload refuses such a tree with IBLGF_ERR_UNSUPPORTED: the
 * device solves implement the reference default subtract_non_leaf = false. */
int iblgf_tree_build_masks(iblgf_tree* t, int subtract_non_leaf);

/* Decomposition::adapt_decoposition (include/iblgf/domain/decomposition/
 * decomposition.hpp:182-470) on one rank: n (key, level change) pairs name
 * leaves to refine (+1; done where the 2:1 balance allows, Tree::try_2to1,
 * tree.hpp:330-377, always for immersed-boundary blocks) or to coarsen (-1:
 * a parent whose children with data are all leaves or corrections aimed for
 * deletion becomes the leaf, Domain::delete_all_children, domain.hpp:
 * 512-568); corrections are re-marked (mark_correction, 617-735), base-level
 * blocks within two blocks of a kept one are added (base_mesh_update 1 when
 * any was).  The tree's masks must be built again before an upload. */
int iblgf_tree_adapt(iblgf_tree* t, int n, const uint64_t* keys, const int* level_change, int* base_mesh_update);

int iblgf_tree_info(const iblgf_tree* t, int* nblocks, int* base_level, int* depth, int* block_extent,
    double* dx_base, int* nleaves);

/* Immersed-boundary points (physical coordinates, npoints x 3) of the tree,
 * at refinement level ib_level (the finest level the tree will have), with
 * the regularised delta kernel delta (IBLGF_DELTA_*) and the safety distance
 * of the extended IB region: IB::init / read_points
 * (include/iblgf/domain/ib.hpp:60-121, 179-200).  Set before init_refine:
 * refinement then refines every block whose grown region holds a point. */
enum iblgf_delta
{
    IBLGF_DELTA_YANG3 = 0, /* reference default */
    IBLGF_DELTA_YANG4 = 1,
    IBLGF_DELTA_ROMA = 2   /* 3-point delta, support 1.5 */
};
int iblgf_tree_set_ib(iblgf_tree* t, const double* points, int npoints, int ib_level, int delta, double safety_dis);

/* The sparse delta weights the device uses (host-side, for checking): the
 * influence nodes of Client::update_ib_rank_and_infl
 * (include/iblgf/domain/decomposition/client.hpp:1264-1320) with nonzero
 * weight delta(dist + face offset) (operators.hpp:844-861, 996-1015).
 * sizes4 receives {interp nnz, spread rows, spread nnz, npoints}; every
 * other pointer may be NULL (call once with NULLs for the sizes).
 * interp: rows 3*i+c, row_ptr[3N+1], node index slot*n^3 + q, weight.
 * spread: row_ptr[rows+1], dst {comp, index} per row, col 3*i+c, weight.
 * owned[N]: 1 if a finest leaf strictly contains the point. */
int iblgf_tree_ib_stencil(const iblgf_tree* t, long* sizes4, int* interp_ptr, int* interp_node, double* interp_w,
    int* spread_ptr, int* spread_dst, int* spread_col, double* spread_w, int* owned);

/* Flat block table in slot order; any pointer may be NULL.  Array sizes:
 * key[nb], level[nb], coord[3nb], bbase[3nb], flags[nb] (bit0 leaf, bit1
 * correction, bit2 leaf_boundary, bit3 physical, bit4 is_ib, bit5
 * extended_ib), parent[nb], children[8nb],
 * neighbor[27nb], influence[189nb], influence_num[nb], mask_src[nb],
 * mask_tgt[nb].  Octant::neighbor/influence (include/iblgf/domain/octree/octant.hpp:461-467),
 * Tree::construct_lists (include/iblgf/domain/octree/tree.hpp:559-603). */
int iblgf_tree_export(const iblgf_tree* t, uint64_t* key, int* level, int* coord, int* bbase, uint32_t* flags,
    int* parent, int* children, int* neighbor, int* influence, int* influence_num, uint64_t* mask_src,
    uint64_t* mask_tgt);

void iblgf_tree_destroy(iblgf_tree* t);

/* ---- output and restart files (host) ------------------------------------ */

/* Tree::write (include/iblgf/domain/octree/tree.hpp:832-853): the keys of the
 * octants with data in depth-first order (uint64 each), then one bool leaf
 * flag per key; Simulation::write_tree (include/iblgf/simulation.hpp:89-97). */
int iblgf_tree_write_restart(const iblgf_tree* t, const char* path);

/* Domain::initialize_with_keys + restart_list_construct (include/iblgf/domain/
 * domain.hpp:212-283, 839-849): the tree of a Tree::write file on the domain
 * {bd_base, bd_extent} (cells of the base level); correction and leaf-boundary
 * flags are recomputed (mark_correction, mark_leaf_boundary). */
int iblgf_tree_read_restart(const char* path, const int bd_base[3], const int bd_extent[3], int block_extent,
    double dx_base, iblgf_tree** out);

/* H5io::write_h5 (include/iblgf/IO/chombo/h5_io.hpp:161-173) with
 * Chombo::write_global_metaData / write_level_info (include/iblgf/IO/chombo/
 * chombo_impl.hpp:1755-2026, 2336-2598): a Chombo-layout HDF5 file of the
 * blocks of blocks_list_build (h5_io.hpp:232-263) — refinement level >= 0,
 * corrections only with include_correction, and then only on the base level
 * with base_correction_only (the reference's flow and restart files pass 1, 1,
 * simulation.hpp:99-121).  Field i is named names[i] (one component) or
 * names[i]_c, has ncomp[i] components and is the host array data[i] of
 * [ncomp[i]][nslots][n^3] in slot order.  time and dt are the level
 * attributes (write_h5 passes 0 and 1). */
int iblgf_chombo_write(const iblgf_tree* t, const char* path, int nfields, const char* const* names,
    const int* ncomp, const double* const* data, int include_correction, int base_correction_only, double time,
    double dt);

/* H5io::read_h5 + Chombo::read_u (h5_io.hpp:96-108, chombo_impl.hpp:217-626):
 * field `name` with ncomp components of mesh type mesh_type (IBLGF_MESH_*, or
 * 3 for vertex fields) into data ([ncomp][nslots][n^3]): zero on the blocks of
 * blocks_list_build(true, true), then every leaf receives the file's cells at
 * its level; a coarser leaf averages, per axis, the one or two middle file
 * cells of the 2^d it covers (one along a face field's own axis, one on the
 * other axes of an edge field).  The file's base dx may be the tree's
 * divided by a power of two.  A name the file does not hold is an error
 * (IBLGF_ERR_ARG; the reference reads component 0 then). */
int iblgf_chombo_read(const iblgf_tree* t, const char* path, const char* name, int ncomp, int mesh_type,
    double* data);

/* Morton key helpers: Key<3>::compute_index / coordinate / parent / child
 * (include/iblgf/domain/octree/key.hpp:64-75, 124-132, 254-275). */
uint64_t iblgf_key_make(int x, int y, int z, int level);
int      iblgf_key_coord(uint64_t key, int* xyz);
uint64_t iblgf_key_parent(uint64_t key);
uint64_t iblgf_key_child(uint64_t key, int i);
/* Key::get_infl_keys (key.hpp:226-250); out holds up to 216 keys; returns count. */
int iblgf_key_influence(uint64_t key, uint64_t* out);

/* ---- kernel tables (host) ---------------------------------------------- */

/* LGF_GL_Lookup table (include/iblgf/lgf/lgf_gl_lookup.hpp:33-52) regenerated by
 * quadrature: g has (N+1)(N+2)(N+3)/6 entries, G(x>=y>=z>=0). */
int    iblgf_lgf_table(int nmax, double* g);
/* LGF_GL<3>::get (include/iblgf/lgf/lgf_gl.hpp:98-105) with N_max = 100. */
double iblgf_lgf_get(int x, int y, int z);
/* LGF_GE<3> kernel value (include/iblgf/lgf/lgf_ge.hpp:158-168). */
double iblgf_if_kernel(int x, int y, int z, double alpha);
/* Nli / cell_center_nli transfer matrices, n x 2n row-major
 * (include/iblgf/fmm/fmm_nli.hpp:319-374; cell_center_nli_intrp.hpp:677-787). */
int iblgf_lagrange_matrix(int n, int pts_cap, int relative_pos, double* out);

/* ---- synthetic vortex-ring workload (host) --------------------------------- */

/* The seeded Gaussian-core vortex rings of the benchmark workload, evaluated
 * natively; they play the roles of the reference's test-setup source and
 * refinement condition (tests/poisson/vortexrings.hpp:165-178, 471-520).
 * params = {nrings, source_max, refinement_factor, dx_base, then per ring:
 * centre xyz, radius, core radius, unit axis xyz, circulation}, physical
 * units; omega = circulation / (pi core^2) exp(-rho^2 / core^2) along the
 * circle's tangent, rho the distance to the core circle.
 * iblgf_vortex_rings_refine is an iblgf_refine_fn (user = params): refine
 * where |omega| over the block grown by two cells exceeds
 * source_max * refinement_factor^diff_level. */
int iblgf_vortex_rings_refine(void* user, const int block_base[3], int block_extent, int refinement_level,
    int diff_level);
/* The components in comp_mask (bit c = component c) of omega on the interior
 * cells of every leaf block of t, [nb][n^3] in slot order, zeros elsewhere. */
int iblgf_vortex_rings_vorticity(const iblgf_tree* t, const double* params, int comp_mask, double* out_x,
    double* out_y, double* out_z);

/* ---- device context ------------------------------------------------------ */

/* Replaces the per-rank cudaSetDevice + Convolution_GPU resources
 * (tests/poisson/poisson_gpu.cu:41-44; src/utilities/convolution_GPU.cu:351-395). */
int iblgf_gpu_create(int device, iblgf_gpu_ctx** out);
int iblgf_gpu_destroy(iblgf_gpu_ctx* ctx);
const char* iblgf_gpu_last_error(const iblgf_gpu_ctx* ctx);
const char* iblgf_last_error(void);

/* Mirrors the tree once and keeps it resident; replaces the per-pair
 * host-to-device copies of the reference GPU build
 * (include/iblgf/utilities/convolution_GPU.hpp:166-197, 408-420). */
int iblgf_gpu_tree_upload(iblgf_gpu_ctx* ctx, const iblgf_tree* t);

/* Fields ([nblocks][n^3] FP64 on the device). */
/* No counterpart in the reference (it runs one tree per process).  Several
 * independent trees of one block extent, base level and dx_base mirrored into
 * one context: the slot table is level-major, then by tree (in the order
 * given), then key-sorted, so every launch of a solve covers the blocks of all
 * trees, and each tree's result is that of its own solve.  Fields are
 * [total blocks][n^3] in that order; iblgf_gpu_forest_slots gives, for tree i,
 * the combined slot of each of its blocks in the tree's own slot order
 * (iblgf_tree_export's).  Immersed-boundary systems need a single tree. */
int iblgf_gpu_forest_upload(iblgf_gpu_ctx* ctx, int ntrees, const iblgf_tree* const* trees);
int iblgf_gpu_forest_slots(iblgf_gpu_ctx* ctx, int tree, int* slots);

int iblgf_gpu_field_create(iblgf_gpu_ctx* ctx, int* field_id);
int iblgf_gpu_field_device_ptr(iblgf_gpu_ctx* ctx, int field_id, void** dptr);
int iblgf_gpu_field_upload(iblgf_gpu_ctx* ctx, int field_id, const double* host);
int iblgf_gpu_field_download(iblgf_gpu_ctx* ctx, int field_id, double* host);
/* No counterpart in the reference (its fields are FP64).  FP32 fields of the
 * same layout, for callers that keep their fields in single precision: the
 * AMR2AMR solves (iblgf_gpu_apply_lgf, _batch, _noalloc and
 * iblgf_gpu_solve_sequence) read and write them, converting to and from their
 * FP64 work arrays; every other operation refuses them with IBLGF_ERR_ARG. */
int iblgf_gpu_field_create_f32(iblgf_gpu_ctx* ctx, int* field_id);
int iblgf_gpu_field_upload_f32(iblgf_gpu_ctx* ctx, int field_id, const float* host);
int iblgf_gpu_field_download_f32(iblgf_gpu_ctx* ctx, int field_id, float* host);

/* Simulation::write / read_h5 (include/iblgf/simulation.hpp:99-121, 196-200)
 * on device fields: iblgf_chombo_write / iblgf_chombo_read with the
 * context's tree, field i taking ncomp[i] consecutive ids of field_ids.
 * The fields go through host memory; a forest context is refused
 * (IBLGF_ERR_UNSUPPORTED). */
int iblgf_gpu_chombo_write(iblgf_gpu_ctx* ctx, const char* path, int nfields, const char* const* names,
    const int* ncomp, const int* field_ids, int include_correction, int base_correction_only, double time,
    double dt);
int iblgf_gpu_chombo_read(iblgf_gpu_ctx* ctx, const char* path, const char* name, int ncomp, int mesh_type,
    const int* field_ids);

enum iblgf_kernel
{
    IBLGF_KERNEL_LAPLACE = 0,
    IBLGF_KERNEL_IF = 1
};

/* PoissonSolver::apply_lgf<Src,Tgt> for one scalar component
 * (include/iblgf/solver/poisson/poisson.hpp:130-135, 587-894) and
 * PoissonSolver::apply_lgf_IF (poisson.hpp:427-494).  AMR2AMR masks,
 * use_correction = true, subtract_non_leaf = false.  Laplace writes dst on
 * leaf blocks only; IF writes every block with data.  Asynchronous. */
int iblgf_gpu_apply_lgf(iblgf_gpu_ctx* ctx, int src_field, int dst_field, int kernel, double alpha_base);

/* The component loop of apply_lgf / apply_lgf_IF over a vector field
 * (poisson.hpp:130-145, 419-434: one scalar solve per entry of
 * Source::nFields()), with the k components run through every launch
 * together: src_fields[c] -> dst_fields[c] for c < k (1 <= k <= 64), each
 * result equal to its single-component solve.  The work arrays grow to k
 * components on first use.  Asynchronous. */
int iblgf_gpu_apply_lgf_batch(iblgf_gpu_ctx* ctx, const int* src_fields, const int* dst_fields, int k, int kernel,
    double alpha_base);

/* Mesh type of a field (MeshObject, include/iblgf/domain/dataFields/datafield.hpp):
 * where its samples sit in a cell, which sets the relative positions of the
 * level transfers on each axis (cell_center_nli_intrp.hpp:599-618): cell 1 on
 * every axis; face component c 0 on axis c, 1 elsewhere; edge component c 1
 * on axis c, 0 elsewhere (1 = half a cell off the child's samples). */
enum iblgf_mesh
{
    IBLGF_MESH_CELL = 0,
    IBLGF_MESH_FACE = 1,
    IBLGF_MESH_EDGE = 2
};
/* iblgf_gpu_apply_lgf_batch for fields of a mesh type, as apply_lgf /
 * apply_lgf_IF coarsen, interpolate and correct with Source::mesh_type()
 * (poisson.hpp:632, 456-460, 1177; the integrating factor of the IF-HERK
 * stages acts on face fields, ifherk.hpp:1155-1156, 1184): component c of the
 * batch is field index c % 3 (so a batch of two face vectors is {x, y, z, x,
 * y, z}).  IBLGF_MESH_CELL is iblgf_gpu_apply_lgf_batch. */
int iblgf_gpu_apply_lgf_mesh(iblgf_gpu_ctx* ctx, const int* src_fields, const int* dst_fields, int k, int kernel,
    double alpha_base, int mesh_type);

/* Replay batched solves (iblgf_gpu_apply_lgf, iblgf_gpu_apply_lgf_batch) from
 * captured HIP graphs: the first call with a given kernel, field set, k,
 * precision and alpha prepares its plans and kernel spectra and records its
 * launches into a graph; later identical calls launch that graph (one launch
 * per solve).  Field contents may change between calls; field ids may not be
 * destroyed.  The cache is dropped when the work arrays grow or on re-upload.
 * Calls run eagerly while kernel timing is on.  No counterpart in the
 * reference (it has no device path); it serves time-step loops that repeat the
 * same solves.  Off by default. */
int iblgf_gpu_set_graphs(iblgf_gpu_ctx* ctx, int on);

/* No counterpart in the reference.  A sequence of solves issued in order, each
 * as iblgf_gpu_apply_lgf_batch (kernel, alpha_base, k components
 * src_fields[c] -> dst_fields[c]); with graphs on, the whole sequence is one
 * captured HIP graph, so a caller that repeats the same sequence every time
 * step (e.g. the integrating-factor and Poisson solves of the three IF-HERK
 * stages, ifherk.hpp:1078-1200) launches one graph per step.  Results equal
 * the items' own calls bit for bit. */
/* No counterpart in the reference.  Autotuner: times
 * iblgf_gpu_apply_lgf_batch(src_fields, dst_fields, k, kernel, alpha_base)
 * under every combination of the launch shapes that leave results unchanged
 * bit for bit — the multiply-accumulate grid (mac_policy 0: resident
 * workgroups in rounds; 1: one workgroup per frequency chunk; 2: resident
 * workgroups as a persistent grid claiming (frequency chunk, run of sibling
 * groups) units from a device work queue), the block-transform grid (resident
 * workgroups looping over blocks, or one per block and component) and one or
 * two streams per FMM pass — with `reps` eager calls each after a warm-up,
 * keeps the fastest for the context's later solves and writes one line per
 * timing plus the choice into report (report_len bytes, NUL-terminated;
 * report may be NULL with report_len 0).  The destination fields are
 * overwritten.  iblgf_gpu_launch_shapes reads the current choice. */
int iblgf_gpu_autotune(iblgf_gpu_ctx* ctx, const int* src_fields, const int* dst_fields, int k, int kernel,
    double alpha_base, int reps, char* report, int report_len);
int iblgf_gpu_launch_shapes(iblgf_gpu_ctx* ctx, int* mac_policy, int* dft_grid, int* streams);
/* Sets the two grid shapes directly (mac_policy 0, 1 or 2, dft_grid 0 or 1,
 * as above; streams through iblgf_gpu_set_concurrency). */
int iblgf_gpu_set_launch_shapes(iblgf_gpu_ctx* ctx, int mac_policy, int dft_grid);

typedef struct iblgf_solve_item
{
    int         kernel;
    double      alpha_base;
    int         k;
    const int*  src_fields;
    const int*  dst_fields;
    int         mesh_type; /* IBLGF_MESH_*, as iblgf_gpu_apply_lgf_mesh */
} iblgf_solve_item;
int iblgf_gpu_solve_sequence(iblgf_gpu_ctx* ctx, int nitems, const iblgf_solve_item* items);

/* on = 1: each FMM pass runs its coarse chain (anterpolation, coarser
 * transforms and multiply-accumulates, interpolation above the finest level)
 * on a second stream beside the finest level's work, joining before the last
 * interpolation; on = 0: one stream.  Results are bit for bit the same.  No
 * counterpart in the reference (it has no device path).  On by default. */
int iblgf_gpu_set_concurrency(iblgf_gpu_ctx* ctx, int on);

enum iblgf_precision
{
    IBLGF_PREC_FP64 = 0,  /* plain path: spectra, DFT legs and MAC in FP64 */
    IBLGF_PREC_MIXED = 1,  /* source, kernel and target spectra stored FP32 (kernel spectra
                              generated in FP64; target spectra accumulated in FP64 and rounded
                              once on store), MAC products in FP32; DFT arithmetic, fields and
                              cross-level transfers FP64 */
    /* Spectra, MAC, fields and transfers as IBLGF_PREC_MIXED, every 1-D DFT
     * leg a GEMM on the matrix cores with FP32 intermediates: */
    IBLGF_PREC_BF16X3 = 2,  /* BF16 hi/lo split, three products, FP32 accumulation */
    IBLGF_PREC_FP16X3 = 3,  /* FP16 hi/lo split of power-of-two scaled lines, three products,
                               FP32 accumulation */
    IBLGF_PREC_FP16 = 4,    /* FP16 of power-of-two scaled lines, one product */
    IBLGF_PREC_INT8X4 = 5,  /* four signed 7-bit digits per value, digit products summed
                               exactly in INT32, combined in FP64 */
    /* The forward transforms' legs as IBLGF_PREC_FP16X3 on the matrix cores,
     * the inverse transforms' legs in FP64 as IBLGF_PREC_MIXED (the inverse
     * output, a sum of many contributions, is where reduced-precision legs
     * cost accuracy); spectra and MAC as IBLGF_PREC_MIXED. */
    IBLGF_PREC_FP16X3_FORWARD = 6
};

/* Arithmetic of the subsequent solves (no reference counterpart: the
 * reference computes in FP64 throughout, include/iblgf/types.hpp). */
int iblgf_gpu_set_precision(iblgf_gpu_ctx* ctx, int precision);

/* Builds the kernel spectra a solve will use (first-use work of
 * LGF_Base::dft, include/iblgf/lgf/lgf.hpp:77-98), so timed solves are warm. */
int iblgf_gpu_prepare(iblgf_gpu_ctx* ctx, int kernel, double alpha_base);

int iblgf_gpu_sync(iblgf_gpu_ctx* ctx);
/* The context's HIP stream (hipStream_t), for timing with events. */
int iblgf_gpu_stream(iblgf_gpu_ctx* ctx, void** stream);

/* Kernel-class timing with HIP events on the context's stream, in place of
 * the reference's POISSON_TIMINGS / Fmm::Timings clocks
 * (include/iblgf/solver/poisson/poisson.hpp:621-893; include/iblgf/fmm/fmm.hpp:1197-1240).
 * Classes: 0 forward DFT, 1 spectral MAC, 2 inverse DFT, 3 transfers,
 * 4 other.  read() synchronises, then returns per class the summed event
 * time (ms), the number of timed launches and the work units (blocks for
 * 0/2, source-target pairs for 1, parent-child transfers for 3). */
int iblgf_gpu_timing_enable(iblgf_gpu_ctx* ctx, int enable);
int iblgf_gpu_timing_read(iblgf_gpu_ctx* ctx, double* ms5, long* launches5, long* work5);
int iblgf_gpu_timing_reset(iblgf_gpu_ctx* ctx);

/* Algorithmic counts of the Laplace solve built so far: FMM passes, source
 * transforms, target transforms, (source, target) pairs, parent-child
 * transfers, leaves. */
int iblgf_gpu_counts(iblgf_gpu_ctx* ctx, long* counts6);

/* Multiply-accumulate plan statistics of the same passes (no counterpart in
 * the reference; for roofline accounting of issued against useful products):
 * entries (one source spectrum acting on the targets of one sibling group),
 * entries acting on every target of their group, pairs carried by those,
 * all pairs, then the other entries by target count 1..8. */
int iblgf_gpu_mac_counts(iblgf_gpu_ctx* ctx, long* counts12);

/* ---- block operators (device) ------------------------------------------- */

enum iblgf_block_op
{
    IBLGF_OP_LAPLACE = 0,        /* Operator::laplace, include/iblgf/operators/operators.hpp:1387-1417 */
    IBLGF_OP_GRADIENT = 1,       /* Operator::gradient, cell -> face, operators.hpp:1425-1444 */
    IBLGF_OP_DIVERGENCE = 2,     /* Operator::divergence, face -> cell, operators.hpp:1453-1474 */
    IBLGF_OP_CURL = 3,           /* Operator::curl, face -> edge, operators.hpp:1483-1511 */
    IBLGF_OP_CURL_TRANSPOSE = 4  /* Operator::curl_transpose, edge -> face, operators.hpp:2420-2453 */
};

/* dst = op(src) on the interior of every block with data at one tree level,
 * dx = dx_base / 2^(level - base_level); 1 field for cell data, 3 (x, y, z)
 * for face and edge data; scale is curl_transpose's factor.  dst must not
 * alias src.  Asynchronous. */
int iblgf_gpu_block_op(iblgf_gpu_ctx* ctx, int op, const int* src_fields, const int* dst_fields, int level,
    double scale);

/* Halo fill of one level: buffer cells from the same-level neighbours'
 * interiors (Client::buffer_exchange, include/iblgf/domain/decomposition/client.hpp). */
int iblgf_gpu_buffer_exchange(iblgf_gpu_ctx* ctx, int field, int level);

/* Operator::levelDivergence (face -> cell, with halo fill and leaf/correction
 * boundary clean) and Operator::levelGradient (cell -> face, then halo fill)
 * (operators.hpp:522-538, 588-604). */
int iblgf_gpu_level_divergence(iblgf_gpu_ctx* ctx, const int src_fields[3], int dst_field, int level);
int iblgf_gpu_level_gradient(iblgf_gpu_ctx* ctx, int src_field, const int dst_fields[3], int level);

/* Zero width planes on block sides: kind 0 Operator::clean_leaf_correction_boundary
 * (operators.hpp:329-513; with leaf_only, correction blocks are zeroed whole and
 * count as missing neighbours; sides are cleaned at the base level only),
 * kind 1 Operator::clean_ib_region_boundary (operators.hpp:73-187; sides whose
 * neighbour is missing or not IB). */
int iblgf_gpu_clean_boundary(iblgf_gpu_ctx* ctx, int field, int level, int kind, int leaf_only, int width);

/* ---- adaptation: device mirror refresh -------------------------------------
 * After iblgf_tree_adapt and iblgf_tree_build_masks on the context's tree:
 * the tree re-uploaded (plans and lists rebuilt; kernel spectra, which do
 * not depend on the tree, kept), every field id kept at the new size; the nf
 * fields listed (FP64; one field group such as the components of a vector,
 * comp[i] its field index, mesh[i] its mesh type) carried over block by
 * block by key on the device, and the new leaves filled as Ifherk::adapt does
 * (ifherk.hpp:1047-1074): per field, halo fills of the group at every level,
 * then every parent of a new leaf (refinement level > 0, in new-leaf key
 * order) interpolated into all its children.  Unlisted fields restart at
 * zero; an immersed-boundary system of the context is dropped (create it
 * again).  nparents (or NULL) receives the parents interpolated. */
int iblgf_gpu_tree_readapt(iblgf_gpu_ctx* ctx, const iblgf_tree* t, int nf, const int* fields, const int* mesh,
    const int* comp, int* nparents);
/* Ifherk::up (ifherk.hpp:999-1004) after clean<F>(true): the leaves' ring and
 * the non-leaf blocks of k fields zeroed, children coarsened into parents
 * (field c is field index c % 3) - what Ifherk::adapt does to u first. */
int iblgf_gpu_coarsen(iblgf_gpu_ctx* ctx, const int* fields, int k, int mesh_type);

/* Operator::maxnorm and Operator::blockRootMeanSquare (include/iblgf/
 * operators/operators.hpp:1351-1367, 1149-1167) of the field with components
 * fields[0..k) (1 <= k <= 3) on every block with data, in slot order: per
 * interior point the Euclidean norm over the components; maxnorm[b] its
 * largest value (the reference's bit for bit), rms[b] its root mean square.
 * Either output may be NULL.  The inputs of the setups' adaptation criteria
 * (tests/ns_amr_lgf/ns_amr_lgf.hpp:448-495, Ifherk::update_source_max,
 * ifherk.hpp:401-420) without a download of the fields. */
int iblgf_gpu_block_norms(iblgf_gpu_ctx* ctx, const int* fields, int k, double* maxnorm, double* rms);

/* ---- IF-HERK time integration (Ifherk, ifherk.hpp) ------------------------
 * Velocity fields are face vectors (three fields, component c the faces
 * normal to axis c), vorticity edge vectors.  All asynchronous. */

/* Ifherk::clean_leaf_correction_boundary (ifherk.hpp:1352-1412), the ramped
 * clean of the integrator (not Operator::clean_leaf_correction_boundary):
 * with leaf_only the level's correction blocks are zeroed; at the base level
 * each block is multiplied, node by node, by Operator::smooth2zero's ramp
 * towards every neighbour that is missing (or a correction block, with
 * leaf_only), operators.hpp:653-760. */
int iblgf_gpu_smooth_boundary(iblgf_gpu_ctx* ctx, int field, int level, int leaf_only);
/* Ifherk::up_and_down (ifherk.hpp:991-1013) of k fields of one mesh type
 * (field c is field index c % 3): the leaves' one-point ring and every
 * non-leaf block zeroed (clean<F>(true)), the children coarsened into their
 * parents (source_coarsify, poisson.hpp:1136-1194), the parents interpolated
 * into their correction children (intrp_to_correction_buffer, 1015-1060);
 * as the reference's halo fills cover every field of the list, each field
 * but the last has its buffers of levels base..depth-2 filled once more from
 * its final values. */
int iblgf_gpu_up_and_down(iblgf_gpu_ctx* ctx, const int* fields, int k, int mesh_type);
/* Operator::nonlinear (operators.hpp:2455-2500) on the blocks of one level:
 * dst = 0.25 (w x u) on the interior nodes from face u and edge w (their
 * one-cell buffers read), then every node of dst times scale (the
 * Ifherk::nonlinear scaling, ifherk.hpp:1676-1686).  dst must not alias u
 * or w. */
int iblgf_gpu_nonlinear(iblgf_gpu_ctx* ctx, const int u_fields[3], const int w_fields[3], const int dst_fields[3],
    int level, double scale);
/* Ifherk::time_step (ifherk.hpp:1078-1200) for a tree without
 * immersed-boundary points, in place on the face velocity u: up_and_down of
 * u, then three stages of nonlinear term, integrating-factor solves and
 * projection (lin_sys_with_ib_solve, 1451-1517, whose force solve and
 * pressure correction vanish without IB points), the stage alphas
 * (c_{i+1} - c_i) dt / (Re dx_base^2) (164-175).  dt is the step of the
 * finest level (dt_base / 2^nref); Re may be infinite (every integrating
 * factor the identity).  p_field (or -1) receives the pressure d_i / (a33 dt).
 * frame_vel (or NULL) holds the frame velocity x, y, z of stages 1, 2, 3,
 * constant over the domain (simulation frame_vel at T + c_i dt; a
 * position-dependent frame is not supported).  refresh selects how the step
 * opens (time_step, 1093-1116): 0 with up_and_down of u; 1 with
 * pad_velocity<u, u>(true) (the base-level velocity rebuilt from the STREAM
 * solve of its curl, 1521-1575) and adapt_corr_time_step (three alpha-0
 * projection stages, 1206-1304), as the reference does when the base mesh
 * changed or the velocity-refresh interval has passed; 2 with pad_velocity
 * only (use_adaptation_correction off).  The solves run in the context's
 * precision.  With graphs on, the whole step is captured as one graph per
 * (fields, dt, Re, frame, refresh) and replayed.  The context holds 29 work
 * fields from the first call. */
int iblgf_gpu_ifherk_step(iblgf_gpu_ctx* ctx, const int u_fields[3], int p_field, double dt, double Re,
    const double* frame_vel, int refresh);
/* The same step with the context's immersed-boundary points
 * (iblgf_gpu_ib_create first): lin_sys_with_ib_solve (ifherk.hpp:1451-1517)
 * in full - per stage the IB2xIB integrating factor of r_i - grad d_i, the
 * force solve (LinSysSolver::ib_solve, linsys.hpp:113-129: the projected
 * velocity less the boundary velocity bc_vel, CG_solve to cg_tol or cg_maxit
 * iterations from the stage's previous force), the pressure correction
 * (linsys.hpp:131-145) and the smeared force removed with the pressure
 * gradient.  The stage forces persist in the context across steps (the
 * reference's force_prev, zero at first; the main stages update them, the
 * adaptation-correction stages do not).  bc_vel (or NULL) holds the boundary
 * velocity x, y, z of stages 1..3, constant over the points.  cg_iterations
 * (or NULL) receives the three main stages' CG iteration counts.  The force
 * solves check convergence on the host, so this step runs eagerly. */
int iblgf_gpu_ifherk_step_ib(iblgf_gpu_ctx* ctx, const int u_fields[3], int p_field, double dt, double Re,
    const double* frame_vel, int refresh, const double* bc_vel, double cg_tol, int cg_maxit, int* cg_iterations);
/* Reads (write 0) or sets (write 1) the stage forces of the IF-HERK step,
 * forces[3 stages][npoints][3]. */
int iblgf_gpu_ib_stage_forces(iblgf_gpu_ctx* ctx, double* forces, int write);
/* Boundary velocity per point and stage for the following
 * iblgf_gpu_ifherk_step_ib calls, velocity[3 stages][npoints][3]: the
 * reference's LinSysSolver::subtract_boundary_vel (linsys.hpp:148-158)
 * subtracts bc_vel(idx, t, coordinate(i)) point by point, a function of the
 * stage's time and the point's position (moving or deforming bodies); the
 * caller evaluates it.  It replaces the step's constant bc_vel until called
 * with NULL. */
int iblgf_gpu_ib_set_boundary_velocity(iblgf_gpu_ctx* ctx, const double* velocity);

/* PoissonSolver::apply_lgf / apply_lgf_IF with an immersed-boundary FMM type
 * (poisson.hpp:587-894, 435-494): mask_type IBLGF_MASK_IB2xIB or
 * IBLGF_MASK_xIB2IB; one FMM pass at the finest level from the interior of
 * the finest blocks.  Laplace writes the finest leaves, IF every finest
 * block.  src may equal dst. */
enum iblgf_mask_type
{
    IBLGF_MASK_IB2xIB = 2,
    IBLGF_MASK_xIB2IB = 3
};
int iblgf_gpu_apply_lgf_masked(iblgf_gpu_ctx* ctx, int src_field, int dst_field, int kernel, double alpha_base,
    int mask_type);

/* The fmm_type argument of PoissonSolver::apply_lgf<Src,Tgt>(int fmm_type) and
 * apply_lgf_IF<Src,Tgt>(double alpha, int fmm_type)
 * (include/iblgf/solver/poisson/poisson.hpp:130-135, 427-433, 587-653): the
 * reference's MASK_TYPE values (include/iblgf/domain/octree/octant.hpp:75-85). */
enum iblgf_fmm_type
{
    IBLGF_FMM_STREAM = 0,
    IBLGF_FMM_AMR2AMR = 1,
    IBLGF_FMM_IB2xIB = 2,
    IBLGF_FMM_xIB2IB = 3,
    IBLGF_FMM_IB2AMR = 4,
    IBLGF_FMM_LAPLACIAN_BC = 5,
    IBLGF_FMM_REF_FOURIER_STREAM = 6,
    IBLGF_FMM_REF_FOURIER_LGF = 7
};
/* apply_lgf / apply_lgf_IF with the reference's fmm_type: AMR2AMR runs
 * iblgf_gpu_apply_lgf, IB2xIB and xIB2IB run iblgf_gpu_apply_lgf_masked;
 * STREAM (Laplace kernel) runs one FMM pass at the base level with the
 * stream-function masks (FmmMaskBuilder::fmm_vortex_streamfun_mask,
 * fmm.hpp:199-233: base-level blocks are targets, the non-correction ones
 * sources), the source's base-level interiors as input, no coarsening and no
 * correction, and writes dst on the base-level leaves (poisson.hpp:609-611,
 * 645-647, 760-768); IB2AMR (Laplace kernel) coarsens the source's finest-level
 * interiors to every level and runs one pass per level with the IB2AMR masks
 * (FmmMaskBuilder::fmm_IB2AMR_mask, fmm.hpp:146-195: blocks holding
 * immersed-boundary points, or with descendants that do, are sources, every
 * block of the level a target), writing each level's leaves
 * (poisson.hpp:613-616, 632-633, 645-650); Laplacian_BC, the Fourier-mode
 * types and STREAM or IB2AMR with the IF kernel return IBLGF_ERR_UNSUPPORTED
 * (checked before anything else, so a caller can route them to its own path),
 * any other value IBLGF_ERR_ARG.  Call sites: ifherk.hpp:1471, 1555;
 * linsys.hpp:142, 277, 332. */
int iblgf_gpu_apply_lgf_typed(iblgf_gpu_ctx* ctx, int src_field, int dst_field, int kernel, double alpha_base,
    int fmm_type);

/* No counterpart in the reference.  iblgf_gpu_apply_lgf for callers that
 * cannot allocate device memory during the call (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc): same solve, same
 * kernels and result, but its work fields and spectra live in a fixed device
 * array of IBLGF_NOALLOC_SCRATCH_BYTES that each GPU holds for the life of the
 * process, and the call makes no hipMalloc/hipFree.
 *  - iblgf_gpu_prepare(ctx, kernel, alpha_base) must have run in the current
 *    precision (plans and kernel spectra are allocated there); otherwise
 *    IBLGF_ERR_STATE.
 *  - A tree whose single-component work exceeds the scratch
 *    (4 n^3 FP64 fields + 2 spectra of H P^2 complex values per block, each
 *    array rounded up to 256 bytes: e.g. about 455 blocks at n = 10 in FP64)
 *    is refused with IBLGF_ERR_UNSUPPORTED.
 *  - All calls on one GPU share its scratch, which keeps the last call's
 *    contents (each call clears the part it uses first): calls that overlap
 *    on one GPU, from any context or stream, race.
 *  - Kernel timing must be off. */
#define IBLGF_NOALLOC_SCRATCH_BYTES (64L << 20)
int iblgf_gpu_apply_lgf_noalloc(iblgf_gpu_ctx* ctx, int src_field, int dst_field, int kernel, double alpha_base);

/* ---- immersed-boundary force system (device) ---------------------------- */

/* Sparse delta weights of the tree's immersed-boundary points on the finest
 * leaf blocks of the uploaded tree (Client::update_ib_rank_and_infl,
 * include/iblgf/domain/decomposition/client.hpp:1264-1320).  The tree must be
 * the one uploaded, with points set. */
int iblgf_gpu_ib_create(iblgf_gpu_ctx* ctx, const iblgf_tree* t);
/* npoints, points owned (strictly inside a finest leaf), nonzero weights. */
int iblgf_gpu_ib_info(iblgf_gpu_ctx* ctx, int* npoints, int* nowned, long* nnz);
/* LinSysSolver::smearing (face data; fields cleaned first) and projection
 * (include/iblgf/solver/linsys/linsys.hpp:350-376, 417-445); forces are host
 * arrays npoints x 3. */
int iblgf_gpu_ib_smear(iblgf_gpu_ctx* ctx, const double* force, const int fields[3]);
int iblgf_gpu_ib_project(iblgf_gpu_ctx* ctx, const int fields[3], double* force);
/* fout = E (I - G L^-1 D) H_alpha E^T fin: LinSysSolver::ET_H_S_E (linsys.hpp:259-319). */
int iblgf_gpu_ib_apply(iblgf_gpu_ctx* ctx, const double* fin, double* fout, double alpha);
/* LinSysSolver::CG_solve (linsys.hpp:170-230): f holds the initial guess and
 * receives the solution; stops when sqrt(r.r / f.f) < tol or after maxit
 * iterations.  history (maxit entries, may be NULL) receives the residual
 * per iteration. */
int iblgf_gpu_ib_cg(iblgf_gpu_ctx* ctx, const double* uc, double* f, double alpha, double tol, int maxit,
    int* iterations, double* residual, double* history);

/* Wall time (s), operator applications, preconditioner fallback flag and
 * stop reason (0 converged, 1 iteration limit, 2 breakdown) of the last CG /
 * flexible-CG solve. */
int iblgf_gpu_ib_stats(iblgf_gpu_ctx* ctx, double* seconds, int* applies, int* fell_back, int* stop);
/* Iteration after which the last learned flexible solve continued with the
 * identity because the network converged slower than the last identity
 * solve (-1: it did not switch).  No counterpart in the reference, whose only
 * method is plain CG (linsys.hpp:170-230). */
int iblgf_gpu_ib_switch_iteration(iblgf_gpu_ctx* ctx, int* iteration);
/* Krylov recycling for iblgf_gpu_ib_fcg: keep up to `capacity` (p, A p) pairs
 * of finished restart cycles (mutually orthogonal, oldest replaced first)
 * while alpha stays the same; every flexible solve projects its initial
 * residual onto them and orthogonalises its new directions against them.
 * 0 turns it off and drops the pairs.  No counterpart in the reference (plain
 * CG from the previous force, linsys.hpp:170-230); it serves callers that
 * repeat force solves with one operator (every IF-HERK stage of a time step
 * with fixed geometry). */
int iblgf_gpu_ib_recycle(iblgf_gpu_ctx* ctx, int capacity);

/* ---- learned preconditioner (no reference counterpart: the reference
 * solves with plain CG, linsys.hpp:170-230; this is its replacement for
 * callers that repeat force solves) ------------------------------------- */

/* A per-point MLP (3 layers, hidden width `hidden`, tanh, BF16 weights, FP32
 * accumulation) over the residuals of each point's K nearest surface
 * neighbours in the point's local frame.  Weights are seeded from `seed`. */
int iblgf_gpu_ib_precond_create(iblgf_gpu_ctx* ctx, const iblgf_tree* t, int K, int hidden, uint64_t seed);
/* Record every stride-th iterate of the following iblgf_gpu_ib_cg /
 * iblgf_gpu_ib_fcg solves as training pairs (r_k, f* - f_k); stride 0 stops. */
int iblgf_gpu_ib_precond_collect(iblgf_gpu_ctx* ctx, int stride);
/* Adam steps on random point batches of the recorded pairs, on the device;
 * loss = mean squared error of the last steps, npairs = pairs held. */
int iblgf_gpu_ib_precond_train(iblgf_gpu_ctx* ctx, int steps, int batch_rows, double lr, double* loss, int* npairs);
/* Sizes: K, hidden, feature dimension (6K), parameter count, output scale. */
int iblgf_gpu_ib_precond_info(iblgf_gpu_ctx* ctx, int* K, int* hidden, int* dim, long* nparam, double* out_scale);
/* Checking hooks: FP32 master parameters (get/set), the features of a host
 * residual (npoints x 6K), and loss + parameter gradients of a batch
 * X [rows][6K], T [rows][3] through the device kernels. */
int iblgf_gpu_ib_precond_params(iblgf_gpu_ctx* ctx, float* params, long nparam);
int iblgf_gpu_ib_precond_set_params(iblgf_gpu_ctx* ctx, const float* params, long nparam);
int iblgf_gpu_ib_precond_features(iblgf_gpu_ctx* ctx, const double* r, float* X);
int iblgf_gpu_ib_precond_grad(iblgf_gpu_ctx* ctx, const float* X, const float* T, int rows, float* grads, double* loss);
/* Flexible CG in conjugate-residual form (flexible GCR with 30 kept
 * directions, restarted; minimises |r| each step, so it converges although
 * the FMM-approximated operator is only symmetric to ~1e-4, on which plain
 * CG converges slowly), preconditioned by the learned network when use_mlp is nonzero
 * (the identity otherwise, and wherever the network's direction is not a
 * finite descent direction).  A learned solve is monitored: every 10
 * iterations its mean residual reduction is compared with 0.8 times that of
 * the last identity solve, and below it the identity takes over for the rest
 * of the solve (iblgf_gpu_ib_switch_iteration).  Same stopping rule and
 * outputs as iblgf_gpu_ib_cg. */
int iblgf_gpu_ib_fcg(iblgf_gpu_ctx* ctx, const double* uc, double* f, double alpha, double tol, int maxit,
    int use_mlp, int* iterations, double* residual, double* history);

/* ---- Screened 2-D LGFs of the spanwise Fourier modes ------------------------
 * The kernels of the reference's 2-D + spanwise-Fourier solvers: per Fourier
 * mode the lattice Green's function of (Lap_h - c^2) on a 2-D grid, c = k dx.
 * Both calls take device pointers from the caller, run on `stream` (NULL: the
 * null stream), allocate nothing and return before the work ends. */

/* Helmholtz::get / Helmholtz_Lookup::get for Dim = 2 (include/iblgf/lgf/
 * helmholtz.hpp:143-153, helmholtz_lookup.hpp:60-68, 146-181): table (nmodes
 * x ny x nx doubles, x fastest) receives, for mode k with screening c[k] (host
 * array, 1e-4 <= c[k]) and 0 <= i < nx, 0 <= j < ny, the value at lattice
 * offset (i, j), which equals -G with (-Lap_h + c^2) G = delta.  The trapezoid
 * rule on the integral's half period uses enough nodes for 1e-19 relative
 * error (about nx + ny + 45 / c).  1 <= nmodes <= 256, nx * ny <= 65536. */
int iblgf_gpu_helm2d_table(const double* c, int nmodes, int nx, int ny, double* table, void* stream);

/* The free-space convolution of each mode's source over a uniform nx x ny
 * patch with that mode's table (Convolution<2> over every source/target
 * block pair of one level, include/iblgf/utilities/convolution.hpp:259-260
 * with the Helmholtz kernel):
 *   dst[k][t] = scale * sum_s table[k][|tj - sj|][|ti - si|] * src[k][s],
 * summed directly over every cell s of the patch (so dst satisfies
 * (Lap_h - c^2) dst = scale * src at interior cells).  src and dst (nmodes x
 * ny x nx doubles) must differ; same limits as iblgf_gpu_helm2d_table. */
int iblgf_gpu_helm2d_convolve(
    const double* table, int nmodes, int nx, int ny, const double* src, double* dst, double scale, void* stream);

#ifdef __cplusplus
}
#endif
#endif
