/*
 * fealpy_amd.h -- C-ABI of libfealpy_amd.so, the MI355X (gfx950) kernels of the
 * Lagrange finite-element assemble-and-solve path.
 *
 * Conventions
 *   - Every array argument is a DEVICE pointer (HBM), except where noted.
 *   - Integer index arrays are int32 (cells, DOFs, CSR indptr/indices); sizes
 *     are int64.  CSR indices inside a row are sorted ascending.
 *   - `dtype` selects the floating type of values/vectors: FA_F64 or FA_F32.
 *     Geometry (node coordinates) and quadrature tables are always FP64.
 *   - `stream` is a hipStream_t passed as an opaque pointer (NULL = default
 *     stream).  All work is enqueued on it; functions that must read a result
 *     back to the host say so and synchronise the stream.
 *   - Scratch memory is caller-owned: `*_work_bytes` reports the size a call
 *     needs and the call takes `work`/`work_bytes`.  No entry point allocates
 *     device memory.
 *   - Return value: FA_OK or an error code; fa_last_error() gives the message
 *     (thread-local).
 *   - Sums of element contributions into one global entry are formed in a
 *     fixed order (ascending cell id, as the reference's stable lexsort +
 *     sequential np.add.at), with no floating-point atomics, so results are
 *     bit-for-bit reproducible run to run.
 *
 * Reference interfaces replaced (weihuayi/fealpy, paths relative to its root)
 * are cited on each declaration.
 */
#ifndef FEALPY_AMD_H
#define FEALPY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* fa_stream_t;

typedef enum {
    FA_OK = 0,
    FA_EINVAL = 1,     /* bad argument (shape, size, dtype)                      */
    FA_ENOMEM = 2,     /* caller workspace too small                             */
    FA_EHIP = 3,       /* HIP runtime / launch error                              */
    FA_ENOCONV = 4,    /* iterative solver stopped at maxit                       */
    FA_ECAPACITY = 5   /* a row exceeds a kernel's per-row capacity              */
} fa_status;

enum { FA_F64 = 0, FA_F32 = 1 };

/* coefficient kinds for element kernels */
enum {
    FA_COEF_NONE = 0,   /* operator absent                                        */
    FA_COEF_CONST = 1,  /* scalar `scale`                                         */
    FA_COEF_CELL = 2,   /* scale * coef[c],      coef shape (NC,)                 */
    FA_COEF_CELLQ = 3,  /* scale * coef[c*NQ+q], coef shape (NC, NQ)              */
    FA_COEF_SEP = 4     /* scale * f(x_cq), f an fa_sep_t that `coef` points to   */
};

/* A coefficient or source given as a sum of separable terms, evaluated by
 * the element kernels at each quadrature point (no point array, no callback):
 *
 *   f(x) = sum_t coef[t] * (field[t] ? field[t][c] : 1)
 *                * prod_{k in [term_end[t-1], term_end[t])} g_k(x[dim[k]])
 *
 *   g = sin(a s + b) | cos(a s + b) | exp(a s + b) | a s + b | s^a (a a whole
 *   number >= 0) | sin(pi (a s + b)) | cos(pi (a s + b)), s the coordinate
 *   dim[k] of the point;
 *   x = sum_v bcs[q*(TD+1) + v] * node[cell[c*(TD+1) + v]] (bc_to_point,
 *   fealpy/mesh/mesh_base.py:454-481).
 *
 * It stands for the reference's @cartesian coefficient callables
 * (fealpy/utils/utils.py:21-53, process_coef_func) when they have this form,
 * e.g. 3 pi^2 k_c sin(pi x) sin(pi y) sin(pi z).  `field` arrays are FP64 per
 * cell (NC,); `bcs` is the (NQ, TD+1) FP64 table of the rule in use.  The
 * struct itself lives in device memory. */
#define FA_SEP_MAX_TERMS 16
#define FA_SEP_MAX_FACTORS 48
enum { FA_SEP_SIN = 0, FA_SEP_COS = 1, FA_SEP_EXP = 2, FA_SEP_AFFINE = 3, FA_SEP_POW = 4, FA_SEP_SINPI = 5,
       FA_SEP_COSPI = 6 };
typedef struct {
    int32_t nterms, nfactors;
    int32_t term_end[FA_SEP_MAX_TERMS];
    int32_t kind[FA_SEP_MAX_FACTORS];
    int32_t dim[FA_SEP_MAX_FACTORS];
    double coef[FA_SEP_MAX_TERMS];
    double a[FA_SEP_MAX_FACTORS];
    double b[FA_SEP_MAX_FACTORS];
    const double* field[FA_SEP_MAX_TERMS];
    const double* bcs;
} fa_sep_t;

const char* fa_last_error(void);
int fa_version(void);

/* ---------------------------------------------------------------- pattern --
 * Sparsity of the global matrix of a cell->DOF map (kernel "b").
 * Replaces COOTensor.coalesce().tocsr() on the index part
 * (fealpy/sparse/coo_tensor.py:137-157,184-213) called from
 * BilinearForm.assembly (fealpy/fem/bilinear_form.py:83-105).
 *
 * Step 1 fa_adjacency_build: DOF -> (cell, local) incidence, a CSR
 *   adj_ptr[gdof+1], adj[nc*ldof] whose entries e = cell*ldof + local are
 *   sorted ascending within each DOF.  Reports the largest incidence count.
 * Step 2 fa_pattern_count: per-row count of distinct columns, exclusive scan
 *   into indptr[gdof+1]; synchronises and returns nnz and the longest row.
 * Step 3 fa_pattern_fill: writes indices[nnz] (sorted per row).
 */
size_t fa_adjacency_work_bytes(int64_t nc, int32_t ldof, int64_t gdof);
int fa_adjacency_build(const int32_t* cell2dof, int64_t nc, int32_t ldof, int64_t gdof,
                       int32_t* adj_ptr, int32_t* adj, int32_t* max_degree_host,
                       void* work, size_t work_bytes, fa_stream_t stream);
size_t fa_pattern_work_bytes(int64_t gdof);
int fa_pattern_count(const int32_t* cell2dof, int64_t nc, int32_t ldof, int64_t gdof,
                     const int32_t* adj_ptr, const int32_t* adj, int32_t max_degree,
                     int32_t* indptr, int64_t* nnz_host, int32_t* max_row_host,
                     void* work, size_t work_bytes, fa_stream_t stream);
int fa_pattern_fill(const int32_t* cell2dof, int64_t nc, int32_t ldof, int64_t gdof,
                    const int32_t* adj_ptr, const int32_t* adj, int32_t max_degree,
                    const int32_t* indptr, int32_t* indices, fa_stream_t stream);

/* The CSR pattern of the P1 space of TetrahedronMesh.from_box(nx, ny, nz)
 * (what fa_pattern_count + fa_pattern_fill build from the incidence lists,
 * coo_tensor.py:137-213) straight from the grid: node (ix, iy, iz) =
 * (ix (ny+1) + iy)(nz+1) + iz shares a Kuhn cell with its neighbours along
 * the 14 Kuhn edge directions, whose lexicographic order is ascending column
 * order.  indptr: NN + 1 entries; indices: nnz, which must be the grid's
 * count (FA_EINVAL otherwise; the caller knows it from the grid). */
size_t fa_p1_box_pattern_work_bytes(int nx, int ny, int nz);
int fa_p1_box_pattern(int nx, int ny, int nz, int32_t* indptr, int32_t* indices, int64_t nnz, void* work,
                      size_t work_bytes, fa_stream_t stream);

/* ------------------------------------------------------- element matrices --
 * Scalar Lagrange element matrices of ScalarDiffusionIntegrator (default
 * variant, fealpy/fem/scalar_diffusion_integrator.py:54-63) plus
 * ScalarMassIntegrator (fealpy/fem/scalar_mass_integrator.py:48-54), summed
 * as a GroupIntegrator (fealpy/fem/integrator.py:361-378):
 *   K_c[i][j] = |T_c| sum_q w_q ( kd_cq grad(phi_i).grad(phi_j) + km_cq phi_i phi_j )
 * Geometry per cell (grad lambda, |T|) follows numpy_backend.py:586-629 and
 * tetrahedron_mesh.py:165-175.  Tables (FP64, host-prepared):
 *   qw[nq], phi[nq*ldof] (mass only), gphi_u[nq*ldof*(td+1)] (diffusion only).
 * Output elem[nc*ldof*ldof] in `dtype`.
 */
int fa_element_matrices(int dtype, int td, int ldof, int nq,
                        const double* node, const int32_t* cell, int64_t nc,
                        const double* qw, const double* phi, const double* gphi_u,
                        int diff_kind, double diff_scale, const double* diff_coef,
                        int mass_kind, double mass_scale, const double* mass_coef,
                        void* elem, fa_stream_t stream);

/* Element load vectors of ScalarSourceIntegrator
 * (fealpy/fem/scalar_source_integrator.py:51-58 -> functional.py:27-65):
 *   b_c[i] = |T_c| sum_q w_q f_cq phi_i(q);  output elem[nc*ldof] in `dtype`. */
int fa_element_vectors(int dtype, int td, int ldof, int nq,
                       const double* node, const int32_t* cell, int64_t nc,
                       const double* qw, const double* phi,
                       int src_kind, double src_scale, const double* src,
                       void* elem, fa_stream_t stream);

/* ----------------------------------------------------------- value gather --
 * Kernel "c": CSR values from element matrices by a fixed-order segmented
 * sum over the DOF->cell incidence (no atomics).  Replaces the value part of
 * coalesce (np.add.at in stable-sorted COO order, coo_tensor.py:197-200).
 * values[nnz] in `dtype`; elem as written by fa_element_matrices (`dtype`),
 * cell c's ldof x ldof block starting at c * elem_stride (0: ldof^2). */
int fa_gather_matrix(int dtype, int64_t gdof, int32_t ldof, const int32_t* cell2dof,
                     const int32_t* adj_ptr, const int32_t* adj,
                     const int32_t* indptr, const int32_t* indices, int32_t max_row,
                     const void* elem, int64_t elem_stride, void* values, fa_stream_t stream);
/* Dense global vector from element vectors: LinearForm.assembly(format='dense')
 * (fealpy/fem/linear_form.py:66-86 -> COOTensor.to_dense, coo_tensor.py:116-130). */
int fa_gather_vector(int dtype, int64_t gdof, int32_t ldof,
                     const int32_t* adj_ptr, const int32_t* adj,
                     const void* elem, void* out, fa_stream_t stream);
/* Element matrices applied to a global vector, cell by cell:
 *   v[c][i] = sum_j elem[c*elem_stride + i*ldof + j] * x[cell2dof[c][j]]
 * (the einsum of BilinearForm.__matmul__ without an assembled matrix,
 * fealpy/fem/bilinear_form.py:126-158); fa_gather_vector then sums v into
 * the global result in ascending cell order (its index_add). */
int fa_element_apply(int dtype, int64_t nc, int32_t ldof, const int32_t* cell2dof, const void* elem,
                     int64_t elem_stride, const void* x, void* v, fa_stream_t stream);

/* Element matrices of LinearElasticityIntegrator on a vector Lagrange space
 * (fealpy/fem/linear_elasticity_integrator.py:60-186, isotropic D of
 * material/elastic_material.py:236-249).  `M` is the precontracted
 * reference tensor M[i][j][m][n] = sum_q w_q R[q][i][m] R[q][j][n]
 * (ldof*ldof*(td+1)^2 doubles, R = d phi / d lambda).  Entry of rows (i, a),
 * columns (j, b): cA A_aa + cB sum_{e != a} A_ee on diagonal blocks,
 * cC A_ab + cD A_ba off them, with A_ab = |T| sum_q w_q d_a phi_i d_b phi_j.
 * Layout i*td + a (dof_priority 0, TensorFunctionSpace shape (-1, td)) or
 * a*ldof + i (dof_priority 1).  elem: (nc, td*ldof, td*ldof).  p <= 4. */
int fa_element_elasticity(int dtype, int td, int ldof, const double* node, const int32_t* cell, int64_t nc,
                          const double* M, double cA, double cB, double cC, double cD, int dof_priority,
                          void* elem, fa_stream_t stream);
/* Element vectors of VectorSourceIntegrator (fealpy/fem/vector_source_integrator.py:45-52):
 * elem[c][I(i, a)] = |T| sum_q w_q phi_i(q) src[c][q][a], src (nc, nq, ncomp) FP64. */
int fa_element_vectors_tensor(int dtype, int td, int ldof, int ncomp, int dof_priority, int nq,
                              const double* node, const int32_t* cell, int64_t nc, const double* qw,
                              const double* phi, const double* src, void* elem, fa_stream_t stream);

/* High-order (tetrahedral P2-P4) diffusion + mass element matrices on the
 * matrix cores: the reduced-precision variant of fa_element_matrices for
 * ScalarDiffusionIntegrator / ScalarMassIntegrator.assembly
 * (scalar_diffusion_integrator.py:54-63, scalar_mass_integrator.py:41-54).
 * A_ij = sum_{s<6} |T| G_s sum_q w_q k(c,q) K_s[q][ij] + |T| sum_q w_q m(c,q) K_6[q][ij]
 * with G_s the metric grad lambda_m . grad lambda_n (m <= n in 1..3) and K_s
 * the reference tensors of the reduced gradients (K_6 = phi_i phi_j), built by
 * the caller in the fragment order [group][s][tile][k-step][lane][t]
 * (fealpy_amd/fem/highorder.py).  precision 0: BF16 inputs
 * (v_mfma_f32_32x32x16_bf16), 1: FP32 (v_mfma_f32_32x32x2f32); FP32
 * accumulation and output.  ks = k-steps (16 or 2 points each, zero padded);
 * tpg = 32-column tiles per LDS slice (fa_mfma_tiles_per_group); cell c's
 * block starts at c * elem_stride (0: ldof^2). */
int fa_mfma_tiles_per_group(int precision, int ks);
size_t fa_mfma_work_bytes(int precision, int64_t nc, int ks);
int fa_element_matrices_mfma(int precision, int ldof, int nq, int ks, int tpg, int ngroups,
                             const double* node, const int32_t* cell, int64_t nc, const double* qw,
                             const void* btab, int diff_kind, double diff_scale, const double* diff_coef,
                             int mass_kind, double mass_scale, const double* mass_coef, float* elem,
                             int64_t elem_stride, void* work, size_t work_bytes, fa_stream_t stream);

/* Fused P1 matrix assembly: geometry, element rows and the fixed-order CSR sum
 * in one kernel, no element buffer.  Diffusion and/or mass terms with the
 * coefficient kinds of fa_element_matrices; `phi` is the P1 basis at the nq
 * points (needed for mass).  Every CSR entry is the sequential sum of its cell
 * contributions in ascending cell order.  Replaces, for p = 1,
 * BilinearForm.assembly (fealpy/fem/bilinear_form.py:46-105) over
 * ScalarDiffusionIntegrator / ScalarMassIntegrator.assembly
 * (scalar_diffusion_integrator.py:54-63, scalar_mass_integrator.py:41-54)
 * and COOTensor.coalesce().tocsr() (sparse/coo_tensor.py:137-213).
 * max_row <= 128. */
int fa_assemble_p1_matrix(int dtype, int td, int64_t gdof, int nq, const double* node,
                          const int32_t* cell, const double* qw, const double* phi,
                          int diff_kind, double diff_scale, const double* diff_coef,
                          int mass_kind, double mass_scale, const double* mass_coef,
                          const int32_t* adj_ptr, const int32_t* adj,
                          const int32_t* indptr, const int32_t* indices, int32_t max_row,
                          void* values, fa_stream_t stream);
/* Fused P1 load vector (ScalarSourceIntegrator + LinearForm.assembly,
 * fealpy/fem/scalar_source_integrator.py:44-58, linear_form.py:36-86). */
int fa_assemble_p1_vector(int dtype, int td, int64_t gdof, int nq, const double* node,
                          const int32_t* cell, const double* qw, const double* phi,
                          int src_kind, double src_scale, const double* src,
                          const int32_t* adj_ptr, const int32_t* adj, void* out, fa_stream_t stream);

/* P1 scalar diffusion/mass matrix of TetrahedronMesh.from_box(nx, ny, nz)
 * (tetrahedron_mesh.py:1017-1086) straight into its CSR values, one pass, no
 * incidence list and no element-row buffer: one thread per row walks the
 * cells around its node in ascending cell order, found from the node's grid
 * position, and sums each entry's contributions in that order (the order of
 * BilinearForm.assembly's COO coalesce, bilinear_form.py:46-105,
 * coo_tensor.py:184-213).  FP64 values equal those of fa_p1_incidence_rows +
 * fa_sum_incidence_rows to a few units in the last place (the same element
 * rows and order; a cube that is an axis-aligned box takes the closed form of
 * the Kuhn tetrahedron's gradients and volume, any other the cross products),
 * bitwise reproducible; FP32 values are the FP64 sums rounded once.
 * node_wide: NN x 4 coordinate rows (x, y, z, pad) in from_box node order;
 * axes: null, or the X, Y, Z arrays of fa_box_axes for a grid it reported on
 * its lines (the nodes then are not read); indptr: the space's CSR row
 * pointer; nnz must be the P1 pattern's (checked: FA_EINVAL otherwise).
 * Coefficient arguments as fa_p1_incidence_rows. */
int fa_p1_box_matrix(int dtype, int nx, int ny, int nz, const double* node_wide, const double* axes, int nq,
                     const double* qw, const double* phi, int diff_kind, double diff_scale, const double* diff_coef,
                     int mass_kind, double mass_scale, const double* mass_coef, const int32_t* indptr, int64_t nnz,
                     void* values, fa_stream_t stream);

/* f at n cartesian points (points[n*gd], row-major, FP64) for a field-free
 * fa_sep_t in device memory (its field pointers and bcs are not read):
 * out[i] = sum_t coef[t] prod_k g_k(points[i*gd + dim[k]]).  Replaces the
 * evaluation of a @cartesian callable at interpolation points, as
 * LagrangeFESpace.boundary_interpolate does with gd
 * (fealpy/functionspace/lagrange_fe_space.py:109-140), in one launch. */
int fa_sep_points(const void* sep, int64_t n, int gd, const double* points, double* out, fa_stream_t stream);

/* Grid lines of a from_box node array: axes[0 .. nx] = x of nodes (i, 0, 0),
 * then y of (0, j, 0), then z of (0, 0, k) (nx + ny + nz + 3 doubles), and
 * *on_grid_host = 1 when every node (i, j, k) is at (X[i], Y[j], Z[k]) (a
 * tensor grid, as from_box makes it), 0 otherwise; syncs.  No counterpart in
 * the reference: it lets fa_p1_box_matrix skip the node rows. */
size_t fa_box_axes_work_bytes(void);
int fa_box_axes(int nx, int ny, int nz, const double* node_wide, double* axes, int32_t* on_grid_host, void* work,
                size_t work_bytes, fa_stream_t stream);

/* P1 load vector of a ScalarSourceIntegrator on TetrahedronMesh.from_box(nx,
 * ny, nz) (scalar_source_integrator.py, LinearForm.assembly,
 * linear_form.py:36-86): pass 1 writes each cell's element vector in cell
 * order into `work` (fa_p1_box_vector_work_bytes for the rule's nq points,
 * 32-byte aligned), pass 2 sums each node's contributions over its cells in
 * ascending cell order, found from its grid position.  A separable source
 * (FA_COEF_SEP) is first tabulated per grid line: each factor once per (cube
 * index along its axis, tetrahedron, point), used for the cells whose
 * vertices lie on the grid lines (a tensor grid: all of them).  The values
 * are those of fa_p1_incidence_vectors + fa_sum_incidence_values to a few
 * units in the last place (bit for bit where the compiler contracts the
 * products alike).  Source arguments as fa_p1_incidence_vectors. */
size_t fa_p1_box_vector_work_bytes(int nx, int ny, int nz, int nq);
int fa_p1_box_vector(int dtype, int nx, int ny, int nz, const double* node_wide, int nq, const double* qw,
                     const double* phi, int src_kind, double src_scale, const double* src, void* out, void* work,
                     size_t work_bytes, fa_stream_t stream);

/* Two-pass P1 assembly in incidence order, the same values as
 * fa_assemble_p1_matrix bit for bit (same element rows, same LDS-table sum in
 * ascending cell order: COOTensor.coalesce's order, sparse/coo_tensor.py:
 * 184-213).  Built once per pattern, fa_incidence_maps: inc[c*ldof + i] = the
 * incidence slot k of (cell c, local DOF i) in the DOF -> incidence list
 * (adj_ptr, adj), and pos[k] = the CSR row positions of the cell's DOFs, one
 * byte each (ldof <= 4, rows of at most 256 entries; syncs).  Pass 1, one
 * thread per cell, vertex coordinates `node_wide` as 16-byte rows (NN x 4
 * doubles in 3D, the node array itself in 2D) (ScalarDiffusionIntegrator / ScalarMassIntegrator /
 * ScalarSourceIntegrator .assembly, fem/scalar_*_integrator.py):
 * fa_p1_incidence_rows writes element row i of cell c at rows[inc*ldof ..]
 * (computed in FP64, stored as dtype_rows: FP64, or FP32 for an FP32 matrix,
 * whose pass 2 then sums in FP32; FP64 rows and values are the fused
 * kernel's bit for bit),
 * fa_p1_incidence_vectors the load-vector entry at vals[inc].  Pass 2 streams
 * each row's slots: fa_sum_incidence_rows (CSR values), fa_sum_incidence_values
 * (the load vector, an ordered sum per DOF). */
int fa_incidence_maps(int64_t gdof, int32_t ldof, const int32_t* cell2dof, const int32_t* adj_ptr,
                      const int32_t* adj, const int32_t* indptr, const int32_t* indices, int32_t* inc, uint32_t* pos,
                      void* work, size_t work_bytes, fa_stream_t stream);
int fa_p1_incidence_rows(int dtype_rows, int td, int64_t nc, int nq, const double* node_wide, const int32_t* cell,
                         const double* qw, const double* phi, int diff_kind, double diff_scale,
                         const double* diff_coef, int mass_kind, double mass_scale, const double* mass_coef,
                         const int32_t* inc, void* rows, fa_stream_t stream);
int fa_sum_incidence_rows(int dtype, int dtype_rows, int td, int64_t gdof, const int32_t* adj_ptr,
                          const int32_t* indptr, const uint32_t* pos, const void* rows, int32_t max_row, void* values,
                          fa_stream_t stream);
int fa_p1_incidence_vectors(int td, int64_t nc, int nq, const double* node_wide, const int32_t* cell, const double* qw,
                            const double* phi, int src_kind, double src_scale, const double* src, const int32_t* inc,
                            double* vals, fa_stream_t stream);
int fa_sum_incidence_values(int dtype, int64_t gdof, const int32_t* adj_ptr, const double* vals, void* out,
                            fa_stream_t stream);

/* Matrix-free P1 operator: y = A x with A the matrix fa_assemble_p1_matrix
 * would assemble for the same integrators (element rows recomputed from the
 * cell geometry and coefficients where they are used; nothing is stored).
 * Replaces BilinearForm.__matmul__ with no assembled matrix
 * (fealpy/fem/bilinear_form.py:126-158: element tensors from
 * assembly_local_iterative applied to the gathered u, then index_add).  Each
 * y_r sums its incidences (adj_ptr, adj) in a fixed order, so the result is
 * the same run to run.  With b != NULL, the residual form used by iterative
 * refinement: y = b - A x, y_r = 0 where is_bd[r] != 0, and *norm2 (device) =
 * sum of y_r^2 in a fixed order; `work` holds fa_p1_apply_work_bytes() bytes.
 * node_wide as for fa_p1_incidence_rows.  x and b are FP64, y is stored as
 * dtype_y (FP64, or FP32 for the right-hand side of an FP32 inner solve);
 * sums and the norm are formed in FP64. */
size_t fa_p1_apply_work_bytes(void);
int fa_p1_apply(int dtype_y, int td, int64_t gdof, int nq, const double* node_wide, const int32_t* cell,
                const double* qw, const double* phi, int diff_kind, double diff_scale, const double* diff_coef,
                int mass_kind, double mass_scale, const double* mass_coef, const int32_t* adj_ptr,
                const int32_t* adj, const double* x, const double* b, const uint8_t* is_bd, void* y, double* norm2,
                void* work, size_t work_bytes, fa_stream_t stream);

/* ------------------------------------------------------------- topology --
 * Unique faces / edges of a simplex mesh: MeshDS.construct + flocc
 * (fealpy/mesh/mesh_data_structure.py:428-464, fealpy/mesh/utils.py:81-110).
 * `local` (nloc x m, m = 2 or 3) lists the vertex positions of each local
 * entity of a cell (localFace / localEdge); cell is (nc x nv) int32 over nn
 * vertices.  Entities are numbered in the lexicographic order of their
 * sorted vertex tuples, and stored with the orientation of their first
 * occurrence (cell-major order); entity2cell[g] = [c0, c1, l0, l1] from the
 * first and last occurrence; cell2entity[c][l] = g.  fa_topology_sort
 * (counting sort by the smallest vertex, then a per-vertex LDS sort; syncs)
 * returns the number of entities; fa_topology_fill, with the same workspace,
 * writes entity (n x m), cell2entity (nc x nloc) and entity2cell (n x 4).
 * FA_ECAPACITY when more than 1024 occurrences share a smallest vertex. */
size_t fa_topology_work_bytes(int64_t nocc, int64_t nn);
int fa_topology_sort(int m, int nloc, const int32_t* local, int nv, const int32_t* cell, int64_t nc, int64_t nn,
                     int64_t* nuniq_host, void* work, size_t work_bytes, fa_stream_t stream);
int fa_topology_fill(int m, int nloc, const int32_t* local, int nv, const int32_t* cell, int64_t nc, int64_t nn,
                     int32_t* entity, int32_t* cell2entity, int32_t* entity2cell, void* work, size_t work_bytes,
                     fa_stream_t stream);

/* FP64 residual of iterative refinement on an assembled FP64 CSR matrix A
 * (the matrix before the Dirichlet condition): y = b - A x on the rows with
 * is_bd == 0, 0 on the others, stored as dtype_y (FP64 or FP32), and
 * *norm2 (device) = sum of y^2 in a fixed order.  With x equal to the
 * boundary values on the Dirichlet rows this is the residual of the system
 * DirichletBC.apply produces (fealpy/fem/dirichlet_bc.py:101-121).  Work:
 * fa_csr_residual_work_bytes().  No counterpart in the reference (its cg
 * runs in FP64 throughout, fealpy/solver/cg.py:14-123). */
size_t fa_csr_residual_work_bytes(void);
int fa_csr_residual(int dtype_y, int64_t n, const int32_t* indptr, const int32_t* indices, const double* values,
                    const double* x, const double* b, const uint8_t* is_bd, void* y, double* norm2, void* work,
                    size_t work_bytes, fa_stream_t stream);

/* Correction step of iterative refinement: x += d on every row with
 * is_bd[i] == 0 (x FP64, d of dtype_d, FP64 or FP32); Dirichlet rows keep
 * their boundary values exactly.  No counterpart in the reference (its cg
 * has no refinement); used by the mixed-precision solve. */
int fa_refine_update(int dtype_d, int64_t n, double* x, const void* d, const uint8_t* is_bd, fa_stream_t stream);

/* Physical points of barycentric quadrature points on every cell:
 * out[c][q][d] = sum_v bcs[q][v] node[cell[c][v]][d]  (FP64; (gd, nv) in
 * {(2, 3), (3, 3), (3, 4)}).
 * Replaces mesh.bc_to_point (fealpy/mesh/mesh_base.py:454-481 ->
 * backend/numpy_backend.py:401-407), the input of every cartesian
 * coefficient / source callable. */
int fa_bc_to_point(int gd, int nv, int64_t nc, int nq, const double* node, const int32_t* cell,
                   const double* bcs, double* out, fa_stream_t stream);

/* ------------------------------------------------------------ Dirichlet BC --
 * DirichletBC.apply (fealpy/fem/dirichlet_bc.py:101-234), reference-compatible
 * form: the new matrix drops every entry in a boundary row or boundary column
 * (_mul, :213-229) and gets a unit diagonal on boundary rows (spdiags + add,
 * :123-147); f <- f - A uh, f[bd] <- uh[bd] (apply_vector :149-198).
 * fa_dirichlet_count writes new_indptr[n+1] and returns the new nnz (syncs);
 * fa_dirichlet_fill writes new_indices (skipped when NULL: pattern reuse) and
 * new_values, stored as dtype_out (dtype, or FP32 from FP64 values: the
 * filtered FP32 operator of a mixed-precision solve, each value rounded once).
 * is_bd: uint8[n]; uh: boundary values (zeros off the boundary). */
size_t fa_dirichlet_work_bytes(int64_t n);
int fa_dirichlet_count(int64_t n, const int32_t* indptr, const int32_t* indices, const uint8_t* is_bd,
                       int32_t* new_indptr, int64_t* new_nnz_host,
                       void* work, size_t work_bytes, fa_stream_t stream);
int fa_dirichlet_fill(int dtype, int dtype_out, int64_t n, const int32_t* indptr, const int32_t* indices,
                      const void* values, const uint8_t* is_bd, const int32_t* new_indptr, int32_t* new_indices,
                      void* new_values, fa_stream_t stream);
int fa_dirichlet_rhs(int dtype, int64_t n, const int32_t* indptr, const int32_t* indices, const void* values,
                     const uint8_t* is_bd, const void* uh, const void* f, void* f_out, fa_stream_t stream);
/* DirichletBC.apply with the pattern kept: the pattern-preserving mode of
 * the Dirichlet condition (the reference's apply, dirichlet_bc.py:101-147,
 * drops boundary rows and columns and so changes the pattern).  new_values
 * (same length as values, may be values itself) get the unit row on every
 * boundary row and 0 on boundary columns of interior rows; f_out = f - A uh
 * on interior rows, uh on boundary rows, bit for bit the right-hand side of
 * fa_dirichlet_apply.  The matrix has the same solution as the filtered one,
 * and a new boundary condition on it keeps the pattern (and everything built
 * on the pattern, such as multigrid symbolic products). */
int fa_dirichlet_apply_keep(int dtype, int64_t n, const int32_t* indptr, const int32_t* indices, const void* values,
                            const uint8_t* is_bd, void* new_values, const void* uh, const void* f, void* f_out,
                            fa_stream_t stream);

/* fa_dirichlet_fill (values only, on the pattern of fa_dirichlet_count) and
 * fa_dirichlet_rhs in one pass over A: DirichletBC.apply (fealpy/fem/
 * dirichlet_bc.py:101-121) with A read once; the same results bit for bit. */
int fa_dirichlet_apply(int dtype, int64_t n, const int32_t* indptr, const int32_t* indices, const void* values,
                       const uint8_t* is_bd, const int32_t* new_indptr, void* new_values, const void* uh,
                       const void* f, void* f_out, fa_stream_t stream);

/* ------------------------------------------------------------------ SpMV --
 * y = A x.  Replaces CSRTensor.matmul -> bm.csr_spmm -> scipy csr_matvec
 * (fealpy/sparse/csr_tensor.py:411-452, backend/numpy_backend.py:181-199). */
int fa_spmv(int dtype, int64_t n, int64_t nnz, const int32_t* indptr, const int32_t* indices,
            const void* values, const void* x, void* y, fa_stream_t stream);
/* Y = A X for nvec right-hand sides, X (ncols, nvec) and Y (n, nvec)
 * row-major: the 2-D branch of bm.csr_spmm, SciPy csr_matvecs
 * (backend/numpy_backend.py:181-199). */
int fa_spmm(int dtype, int64_t n, int64_t nnz, int64_t nvec, const int32_t* indptr, const int32_t* indices,
            const void* values, const void* x, void* y, fa_stream_t stream);

/* ------------------------------------------------------------ vector ops --
 * Building blocks of fealpy/solver/cg.py's vector algebra (bm.sum(r*z),
 * x + alpha p, ...).  Reductions use a fixed grid and a fixed tree order
 * (deterministic) and write their FP64 result to device memory `out`. */
size_t fa_reduce_work_bytes(void);
int fa_dot(int dtype, int64_t n, const void* x, const void* y, double* out, void* work, size_t work_bytes,
           fa_stream_t stream);
int fa_axpby(int dtype, int64_t n, double a, const void* x, double b, void* y, fa_stream_t stream);
int fa_cast(int dtype_in, int dtype_out, int64_t n, const void* x, void* y, fa_stream_t stream);

/* ----------------------------------------------------------------- SELL-64 --
 * Solver-side matrix format: sliced ELLPACK whose slice is one 64-lane
 * wavefront of consecutive rows.  Row r = 64 s + i keeps its k-th entry at
 * slice_ptr[s] + 64 k + i (coalesced across the wave); the slice width is its
 * longest row; padding entries have value 0 and a column index < ncols.
 * Built from CSR once per pattern
 * (fa_sell_count + fa_sell_fill with cols), values refreshed per assembly
 * (fa_sell_fill with cols = NULL), optionally cast FP64 -> FP32 on the way. */
typedef struct {
    int64_t n;                 /* rows                                          */
    int64_t ncols;             /* columns                                       */
    const int32_t* slice_ptr;  /* ceil(n/64) + 1 entry offsets                  */
    const int32_t* cols;
    const void* vals;          /* FP64 or FP32, as the call's dtype says        */
} fa_sell_t;

/* Compressed SELL-64 for the FP32 smoother: the row-scaled operator
 * D^-1 A (entries in [-1, 1]) in FP16 with int16 column offsets from the row
 * (same slices as the full SELL; padding: offset 0, value 0).  4 bytes per
 * entry instead of 8.  offs = NULL: a column lies more than 32767 from its
 * row, and the FP16 values use the level's int32 SELL columns (6 bytes an
 * entry). */
typedef struct {
    int64_t n;
    const int32_t* slice_ptr;
    const int16_t* offs;       /* column - row                                  */
    const void* vals;          /* FP16 (a_ij / a_ii)                            */
} fa_sell16_t;

size_t fa_sell_work_bytes(int64_t n);
int fa_sell_count(int64_t n, const int32_t* indptr, int32_t* slice_ptr, int64_t* entries_host,
                  void* work, size_t work_bytes, fa_stream_t stream);
int fa_sell_fill(int dtype_in, int dtype_out, int64_t n, int64_t ncols, const int32_t* indptr,
                 const int32_t* indices, const void* values, const int32_t* slice_ptr, int32_t* cols, void* vals,
                 fa_stream_t stream);
int fa_sell_spmv(int dtype, const fa_sell_t* A, const void* x, void* y, fa_stream_t stream);

/* ------------------------------------------------------- Galerkin R A P --
 * Coarse operator C = R A P of the multigrid hierarchy, the role of
 * `R @ A @ P` with SciPy SpGEMM in GAMGSolver.setup
 * (fealpy/solver/gamg_solver.py:85-101).  All operands CSR (int32 indices,
 * FP64 values).  Symbolic phase (count + fill) once per hierarchy; numeric
 * phase per assembly, summed in a fixed order (no atomics).  A coarse row
 * with no entries (a Dirichlet node outside R's range) gets a unit diagonal.
 * max_cand bounds sum_{i in R row} sum_{j in A row i} |P row j|. */
size_t fa_rap_work_bytes(int64_t nrows);
int fa_rap_count(int64_t nrows, int64_t ncols, const int32_t* r_ptr, const int32_t* r_idx,
                 const int32_t* a_ptr, const int32_t* a_idx, const int32_t* p_ptr, const int32_t* p_idx,
                 int32_t max_cand, int32_t* c_ptr, int64_t* nnz_host, int32_t* max_row_host,
                 void* work, size_t work_bytes, fa_stream_t stream);
int fa_rap_fill(int64_t nrows, const int32_t* r_ptr, const int32_t* r_idx, const int32_t* a_ptr,
                const int32_t* a_idx, const int32_t* p_ptr, const int32_t* p_idx, int32_t max_cand,
                const int32_t* c_ptr, int32_t* c_idx, fa_stream_t stream);
/* fa_rap_values / fa_rap_values_terms: a_val and c_val are of `dtype` (FP64
 * or FP32, the type of the operator being set up); products and sums are
 * formed in FP64. */
int fa_rap_values(int dtype, int64_t nrows, const int32_t* r_ptr, const int32_t* r_idx, const double* r_val,
                  const int32_t* a_ptr, const int32_t* a_idx, const void* a_val,
                  const int32_t* p_ptr, const int32_t* p_idx, const double* p_val,
                  const int32_t* c_ptr, const int32_t* c_idx, int32_t max_row, void* c_val,
                  fa_stream_t stream);

/* Numeric R A P from a precomputed term list, for transfers with at most 3
 * distinct values (nested P1 grids: 1, 1/2; P2 -> P1: 1, 1/2; P3 -> P1: 1,
 * 2/3, 1/3); same product as fa_rap_values, which it replaces per assembly
 * (gamg_solver.py:85-101 forms it with SciPy SpGEMM).  Each coarse entry is
 * sum_k w_k a_val[t_k] over its terms in the order R entry -> A entry -> P
 * entry, w_k = r p a product of two values of `values` (host array).  The
 * terms of coarse entry e are words tptr[e] .. tptr[e+1] - 1 (CSR over the
 * coarse entries, no padding); one 32-bit word each (A entry index in bits
 * 0-27, weight code in bits 28-31), so nnz(A) < 2^28.  The numeric sum of an
 * entry is formed by 16 lanes, lane h adding terms h, h + 16, ... in order,
 * combined by a fixed butterfly: the same value run to run.
 * fa_rap_terms_count: the entry offsets tptr (nnz_c + 1 entries) and the term
 * count; syncs.  fa_rap_terms_fill: the words; syncs (error flags).
 * c_ptr: the coarse pattern of fa_rap_count / fa_rap_fill. */
size_t fa_rap_terms_work_bytes(int64_t nnz_c);
int fa_rap_terms_count(int64_t nrows, const int32_t* r_ptr, const int32_t* r_idx, const double* r_val,
                       const int32_t* a_ptr, const int32_t* a_idx, const int32_t* p_ptr, const int32_t* p_idx,
                       const double* p_val, int32_t max_cand, const int32_t* c_ptr, int64_t nnz_c,
                       const double* values, int nvalues, int32_t* tptr, int64_t* nterms_host, void* work,
                       size_t work_bytes, fa_stream_t stream);
int fa_rap_terms_fill(int64_t nrows, const int32_t* r_ptr, const int32_t* r_idx, const double* r_val,
                      const int32_t* a_ptr, const int32_t* a_idx, const int32_t* p_ptr, const int32_t* p_idx,
                      const double* p_val, int32_t max_cand, const int32_t* c_ptr, int64_t nnz_c,
                      const double* values, int nvalues, const int32_t* tptr, int64_t nterms, uint32_t* terms,
                      void* work, size_t work_bytes, fa_stream_t stream);
int fa_rap_values_terms(int dtype, int64_t nnz_c, const int32_t* tptr, const uint32_t* terms,
                        const double* values, int nvalues, const void* a_val, void* c_val, fa_stream_t stream);
/* Staged form of the same product, for patterns whose every coarse row I
 * has at most 64 R entries, at most 63 coarse entries, at most 1024 terms,
 * and whose fine rows (the A rows of R row I) hold at most 512 entries
 * together (nested P1 grids: 15 fine rows of 15 entries): the terms are
 * 16-bit words, weight code in bits 9-12 and in bits 0-8 the A entry's slot
 * in the concatenation of R row I's A rows (in R's order), with the same
 * offsets tptr and term order as fa_rap_terms_fill.  fa_rap_values_staged
 * copies those A rows and the row's words into LDS per coarse row and sums
 * each entry from there: the values of fa_rap_values_terms bit for bit,
 * without gathers from A.  fa_rap_terms_fill_staged returns FA_ECAPACITY
 * when a coarse row is past these bounds.  a_ptr and c_ptr must be the
 * patterns the words were made for. */
int fa_rap_terms_fill_staged(int64_t nrows, const int32_t* r_ptr, const int32_t* r_idx, const double* r_val,
                             const int32_t* a_ptr, const int32_t* a_idx, const int32_t* p_ptr, const int32_t* p_idx,
                             const double* p_val, int32_t max_cand, const int32_t* c_ptr, int64_t nnz_c,
                             const double* values, int nvalues, const int32_t* tptr, int64_t nterms, uint16_t* terms,
                             void* work, size_t work_bytes, fa_stream_t stream);
int fa_rap_values_staged(int dtype, int64_t nrows, const int32_t* r_ptr, const int32_t* r_idx, const int32_t* a_ptr,
                         const int32_t* c_ptr, const int32_t* tptr, const uint16_t* terms, const double* values,
                         int nvalues, const void* a_val, void* c_val, fa_stream_t stream);

/* ------------------------------------------------ algebraic multigrid --
 * Smoothed-aggregation transfers (the algebraic branch of GAMGSolver,
 * fealpy/solver/gamg_solver.py:104-121; fealpy/solver/amg.py).
 * fa_amg_aggregate: aggregates of the strength graph of A (CSR, values of
 * dtype; j strong for i when |a_ij| >= theta sqrt(|a_ii a_jj|)) grown from a
 * distance-2 maximal independent set computed in parallel rounds (keys: a
 * fixed hash of the row index, the index as tie-break); agg[i] = aggregate of
 * row i, -1 for rows with no strong off-diagonal entry.  Returns the number
 * of aggregates and of rounds; syncs once per round.  Deterministic.
 * fa_amg_smooth_prolongation: values of P = (I - omega D^-1 A) P0 on the
 * pattern (s_indptr, s_indices) of A P0 with values s_values (P0: one unit
 * entry per aggregated row): p_ij = [agg(i) == j] - omega / a_ii (A P0)_ij. */
size_t fa_amg_aggregate_work_bytes(int64_t n, int64_t nnz);
int fa_amg_aggregate(int dtype, int64_t n, const int32_t* indptr, const int32_t* indices, const void* values,
                     double theta, int32_t* agg, int64_t* naggregates_host, int32_t* rounds_host, void* work,
                     size_t work_bytes, fa_stream_t stream);
int fa_amg_smooth_prolongation(int dtype, int64_t n, const int32_t* a_indptr, const int32_t* a_indices,
                               const void* a_values, const int32_t* agg, const int32_t* s_indptr,
                               const int32_t* s_indices, const double* s_values, double omega, double* p_values,
                               fa_stream_t stream);

/* ------------------------------------------------------- multigrid cycle --
 * Geometric multigrid V-cycle used as a CG preconditioner: Chebyshev
 * polynomial smoothing of D^-1 A (pre and post, same polynomial, so the cycle
 * is symmetric), restriction R, prolongation P, dense inverse on the
 * coarsest level.  Replaces GAMGSolver.vcycle (fealpy/solver/gamg_solver.py:
 * 256-301), whose Gauss-Seidel smoothing is a sequential triangular solve.
 * All level arrays are caller-owned device memory of the cycle's dtype. */
typedef struct {
    fa_sell_t A;               /* operator of this level                        */
    const void* dinv;          /* 1 / diag(A)                                   */
    fa_sell_t R;               /* restriction to the next coarser level         */
    fa_sell_t P;               /* prolongation from the next coarser level      */
    double lmax;               /* largest-eigenvalue estimate of D^-1 A         */
    void* b;                   /* level vectors of length A.n                   */
    void* x;
    void* x2;
    void* d;
    void* r;
    fa_sell16_t Ah;            /* compressed D^-1 A for an FP32 cycle, or vals = NULL */
    const void* diag;          /* diag(A) in the cycle dtype (with Ah)          */
    int32_t width;             /* longest row of A (0: unknown); pre-smoothing
                                  fuses its first two steps only for rows of at
                                  most 32 entries                               */
    int32_t grid[3];           /* cell counts of this level's TetrahedronMesh
                                  from_box grid when R and P are the P1
                                  transfers to the nested grid of half the
                                  counts (all counts even, one component); then
                                  the V-cycle applies R and P from the grid
                                  (fine node 2I + o, o a Kuhn edge direction,
                                  weight 1/2; o = 0, weight 1) with the values
                                  of the SELL transfers.  {0, 0, 0} otherwise  */
    const uint8_t* bd_fine;    /* with grid: Dirichlet flags of this level's    */
    const uint8_t* bd_coarse;  /* nodes and of the next coarser level's         */
} fa_mg_level_t;

typedef struct {
    int dtype;                 /* precision of the cycle                        */
    int nlevels;
    const fa_mg_level_t* levels;   /* host array, finest level first            */
    int degree;                /* Chebyshev degree of each smoothing sweep      */
    double lower, upper;       /* smoothing interval [lower, upper] * lmax      */
    const double* coarse_chol; /* inverse of the coarsest A, row-major n x n
                                  FP64 (fa_dense_spd_inverse), or NULL: then
                                  the coarsest level is smoothed from zero
                                  with a Chebyshev polynomial of degree
                                  coarse_degree (any size; SPD, so the cycle
                                  stays a valid CG preconditioner)              */
    int coarse_degree;
} fa_mg_t;

/* ------------------------------------------- one-workgroup PCG solves --
 * Many small independent systems in one launch, one 1024-thread workgroup
 * each: the whole MG-preconditioned CG of a request (FP64 outer recurrence
 * on its FP64 SELL operator, the V-cycle of its FP32 hierarchy) runs inside
 * one workgroup with block barriers between phases, so a request costs one
 * workgroup's time instead of ~400 dependent kernel launches.  The role of
 * the reference's cg (fealpy/solver/cg.py:14-123) with GAMGSolver as M, per
 * request; the batch axis of Form / cg (fem/form.py:39-52, solver/cg.py:
 * 69-123) is served by one workgroup per request instead of lockstep columns,
 * so every request stops on its own test.  Workgroups never wait on each
 * other.  A request's result depends only on its own data (bitwise the same
 * alone and in any batch); it agrees with fa_pcg to the solve tolerance (the
 * reductions are summed in a different fixed order). */
typedef struct {
    fa_sell_t A, R, P;         /* level operator and transfers (FP32 values)    */
    const void* dinv;          /* FP32                                          */
    void* b;                   /* level vectors (FP32, length A.n); the finest  */
    void* x;                   /* level's b is the cycle's input (the FP32 copy */
    void* x2;                  /* of the residual)                              */
    void* d;
    void* r;
    double lmax;
} fa_wg_level_t;

#define FA_WG_MAX_LEVELS 16
typedef struct {
    fa_sell_t A;               /* FP64 operator (SELL-64, FP64 values)          */
    const double* b;           /* right-hand side                               */
    double* x;                 /* in: initial guess; out: solution              */
    double* r;                 /* FP64 work vectors, length A.n each            */
    double* p;
    double* Ap;
    double atol, rtol;
    int64_t maxit;             /* > 0                                           */
    int32_t nlevels;           /* 1 .. FA_WG_MAX_LEVELS                         */
    int32_t degree;            /* Chebyshev degree                              */
    double lower, upper;       /* smoothing interval [lower, upper] * lmax      */
    const double* coarse_inv;  /* inverse of the coarsest operator, row-major   */
    fa_wg_level_t levels[FA_WG_MAX_LEVELS];
    int64_t niter;             /* out                                           */
    double residual;           /* out: sqrt(r . z) at the last test             */
    int32_t status;            /* out: 1 converged, 2 ||b|| < 1e-15, 3 maxit,
                                  4 breakdown                                   */
    int32_t pad;
} fa_wg_solve_t;

/* reqs_host: the descriptors (checked on the host); reqs_dev: a device copy
 * of them, whose out fields the kernel writes (copy them back after the
 * stream has run).  Asynchronous. */
int fa_pcg_wg_batched(int nreq, const fa_wg_solve_t* reqs_host, fa_wg_solve_t* reqs_dev, fa_stream_t stream);

/* Fill a compressed SELL (offsets and FP16 values of D^-1 A) from the CSR
 * (values of `dtype`, FP64 or FP32), on the slices `slice_ptr` of the level's
 * SELL.  With a workspace
 * (>= 4 bytes) it checks the band and syncs: FA_ECAPACITY when a column lies
 * more than 32767 from its row (the level then keeps the full format).
 * work = NULL refreshes the values of an already checked pattern without a
 * sync. */
int fa_sell16_fill(int dtype, int64_t n, const int32_t* indptr, const int32_t* indices, const void* values,
                   const int32_t* slice_ptr, int16_t* offs, void* vals, void* work, size_t work_bytes,
                   fa_stream_t stream);

/* Dense row-major copy of a small CSR matrix (FP64) and its in-place inverse
 * (Gauss-Jordan, SPD so no pivoting; n <= 2048, single workgroup).  The
 * coarsest-level solve of the cycle is then one dense matrix-vector product. */
int fa_csr_to_dense(int64_t n, const int32_t* indptr, const int32_t* indices, const double* values,
                    double* dense, fa_stream_t stream);
int fa_dense_spd_inverse(int64_t n, double* a, fa_stream_t stream);
/* Smoother data of a level.  fa_jacobi_setup: dinv = 1/diag(A) (1 where the
 * diagonal is absent) in `dtype_out`, and the Gershgorin bound of D^-1 A,
 * max_i sum_j |a_ij|/|a_ii| -- an upper bound of its spectrum, the top of the
 * Chebyshev interval; syncs.  fa_mg_lmax: power-iteration estimate of the
 * largest eigenvalue of D^-1 A on level `level` (uses the level's x, x2;
 * diagnostics); syncs.  Work: fa_reduce_work_bytes(). */
int fa_jacobi_setup(int dtype_out, int64_t n, int64_t nnz, const int32_t* indptr, const int32_t* indices,
                    const double* values, void* dinv, double* bound_host, void* work, size_t work_bytes,
                    fa_stream_t stream);
int fa_mg_lmax(const fa_mg_t* mg, int level, int iters, double* lmax_host, void* work, size_t work_bytes,
               fa_stream_t stream);
/* Smoother data of a level in one pass over its SELL copy A (values of
 * dtype_in: FP64, or FP32 for an operator assembled in FP32), no sync:
 * dinv (dtype) as fa_jacobi_setup, diag = 1/dinv (or NULL), the FP16
 * D^-1 A values of Ah (offsets already filled by fa_sell16_fill on the same
 * slices; Ah NULL or Ah->vals NULL: none), the FP32 copy of A's values
 * (vals_f32, or NULL), and the Gershgorin bound written to bound_dev[0]
 * (device memory; read it back with the other levels' bounds in one copy).
 * Replaces the per-level diagonal extraction of GAMGSolver.setup
 * (fealpy/solver/gamg_solver.py:85-120).  Work: fa_reduce_work_bytes(). */
int fa_smoother_setup(int dtype, int dtype_in, const fa_sell_t* A, void* dinv, void* diag, void* vals_f32,
                      const fa_sell16_t* Ah, double* bound_dev, void* work, size_t work_bytes, fa_stream_t stream);
/* z = V(r) on the finest level (r, z: dtype of the cycle). */
int fa_mg_vcycle(const fa_mg_t* mg, const void* r, void* z, fa_stream_t stream);

/* -------------------------------------------------------------------- PCG --
 * cg(A, b, x0, M, atol, rtol, maxit) of fealpy/solver/cg.py:14-123 for one
 * right-hand side with the reference's stopping rule (sqrt(r.z) < atol or
 * < rtol ||b||, or niter >= maxit, tested after each update).  A and the
 * vectors b, x are in `dtype`; M = NULL (plain CG) or a V-cycle of any dtype
 * (FP64 CG with an FP32 cycle is the mixed-precision solver).  x holds x0 on
 * entry.  The host reads the device convergence flag every `check_every`
 * iterations; iterations past convergence are skipped on the device.
 * Returns FA_OK, or FA_ENOCONV (maxit reached or a NaN/Inf breakdown). */
size_t fa_pcg_work_bytes(int dtype, int64_t n);
int fa_pcg(int dtype, const fa_sell_t* A, const fa_mg_t* M, const void* b, void* x, double atol, double rtol,
           int64_t maxit, int32_t check_every, int64_t* niter_host, double* residual_host,
           void* work, size_t work_bytes, fa_stream_t stream);
/* fa_pcg with x0 = 0: x is output only (set to 0 before the first update),
 * and the initial residual is b itself, with no product A x0; the same
 * iterates and result as fa_pcg on a zero x, one SpMV pass fewer (the inner
 * solves of iterative refinement start from zero). */
int fa_pcg_from_zero(int dtype, const fa_sell_t* A, const fa_mg_t* M, const void* b, void* x, double atol,
                     double rtol, int64_t maxit, int32_t check_every, int64_t* niter_host, double* residual_host,
                     void* work, size_t work_bytes, fa_stream_t stream);

/* fa_pcg without a workspace argument.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code in a tight serving
 * loop, harnesses that forbid hipMalloc).  Same contract, arguments and
 * results as fa_pcg (bitwise the same x, niter and residual); the vectors,
 * partial sums and state live in a file-scope device array of
 * FA_PCG_STATIC_BYTES (the workspace of FA_PCG_STATIC_MAX_ROWS FP64 rows),
 * one per GPU, held until the process ends.  A->n above
 * FA_PCG_STATIC_MAX_ROWS returns FA_EINVAL.  It calls no allocator.  The
 * array's device address is looked up on the first call on each device, so
 * `stream` must belong to the current device.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (other streams, other
 * threads) race and must be serialised by the caller. */
#define FA_PCG_STATIC_MAX_ROWS 262144
#define FA_PCG_STATIC_BYTES (5u * 262144u * 8u + 32768u + 4096u)
int fa_pcg_static(int dtype, const fa_sell_t* A, const fa_mg_t* M, const void* b, void* x, double atol, double rtol,
                  int64_t maxit, int32_t check_every, int64_t* niter_host, double* residual_host,
                  fa_stream_t stream);

/* fa_pcg with its iterations replayed from a HIP graph, for callers whose
 * solves are bound by kernel-launch cost (small systems, many concurrent
 * requests).  Same contract, arguments and results as fa_pcg (cg of
 * fealpy/solver/cg.py:14-123); bitwise the same x, niter and residual.
 * Each call records one PCG iteration (SpMV, the V-cycle, the scalar
 * kernels) on the handle's private capture stream into a graph (nothing runs
 * while recording), updates the handle's executable graph (or instantiates
 * it when the kernel sequence changed), then launches it on `stream` once
 * per iteration; the initial residual and first cycle run as in fa_pcg.
 * A handle serves one caller thread at a time; it holds a HIP stream and an
 * executable graph until fa_pcg_graph_destroy. */
typedef struct fa_pcg_graph_s fa_pcg_graph_t;
int fa_pcg_graph_create(fa_pcg_graph_t** out);
void fa_pcg_graph_destroy(fa_pcg_graph_t* graph);
int fa_pcg_graph(int dtype, const fa_sell_t* A, const fa_mg_t* M, const void* b, void* x, double atol, double rtol,
                 int64_t maxit, int32_t check_every, int64_t* niter_host, double* residual_host,
                 void* work, size_t work_bytes, fa_pcg_graph_t* graph, fa_stream_t stream);

/* ----------------------------------------------------------- autotuner --
 * Tile shapes chosen by the project autotuner (fealpy_amd/tune.py), which
 * times the candidates of a kernel family on the caller's own operator and
 * stores the fastest here; no counterpart in the reference.  Families:
 *   FA_TUNE_SPMV_LANES      lanes per CSR row of fa_spmv (4..64), keyed by the
 *                           average row length class (fa_tune_key(f, n, nnz));
 *   FA_TUNE_SUM_ROWS_LANES  lanes per CSR row of P1 assembly pass 2
 *                           (fa_sum_incidence_rows: 4, 8, 16 or 32), keyed by
 *                           (td, longest row) (fa_tune_key(f, td, max_row)).
 * A value of 0 restores the kernel's built-in choice.  Process-wide; the
 * entries are read at each launch. */
enum { FA_TUNE_SPMV_LANES = 0, FA_TUNE_SUM_ROWS_LANES = 1, FA_TUNE_FAMILIES = 2, FA_TUNE_KEYS = 8 };
int fa_tune_set(int family, int key, int value);
int fa_tune_get(int family, int key);
int fa_tune_key(int family, int64_t a, int64_t b);

/* --------------------------------------------------------- batched solves --
 * Many independent Dirichlet problems packed at offsets into one block-
 * diagonal system, so each kernel of a PCG iteration and of its V-cycle runs
 * once for the whole batch (serving).  The reference's counterpart is the
 * batch axis of Form and cg (fealpy/fem/form.py:39-52, fealpy/solver/cg.py:
 * 69-123: columns in lockstep, each with its own alpha and beta).
 * fa_pack_sell: the SELL-64 matrices of the requests into one (items in host
 * and device memory, the same contents): request q's slices go to
 * slice_base.., its entries to entry_base.. with columns shifted by col_off,
 * followed by extra_slices empty slices; slice_ptr gets nslices_total + 1
 * entries.  fa_copy_segments: dst[0, len) = src (or 0 when src is NULL),
 * then zero_tail zeros, for every segment, in one launch.
 * fa_pcg_batched: PCG on the packed system A (FP64 values) with the FP32
 * V-cycle M over the packed hierarchy; the rows come in segments of 256
 * (block_req[k] = request of rows [256k, 256k + 256), block_off[q] = first
 * segment of request q, nreq + 1 entries, both in device memory); b.b, r.z and
 * p.Ap are summed per request in a fixed order, alpha, beta and the stop test
 * (the cg rule: sqrt(r.z) < atol or < rtol ||b_q||) are per request, and a
 * stopped request's x is frozen at its converging iterate.  Per request:
 * iterations, final sqrt(r.z) and status (1 converged, 2 ||b|| = 0, 3 maxit,
 * 4 breakdown).  Syncs every check_every iterations. */
typedef struct {
    const int32_t* sp;      /* the request's slice pointers (nslices + 1)          */
    const int32_t* cols;
    const void* vals;
    int64_t nslices, extra_slices, slice_base, entry_base, entries;
    int32_t col_off;
} fa_pack_sell_t;
typedef struct {
    const void* src;        /* NULL: zeros                                         */
    void* dst;
    int64_t len, zero_tail;
} fa_copy_seg_t;
int fa_pack_sell(int elem_bytes, int nreq, const fa_pack_sell_t* items_host, const fa_pack_sell_t* items_dev,
                 int64_t nslices_total, int64_t entries_total, int32_t* slice_ptr, int32_t* cols, void* vals,
                 fa_stream_t stream);
int fa_copy_segments(int elem_bytes, int nseg, const fa_copy_seg_t* segs_host, const fa_copy_seg_t* segs_dev,
                     fa_stream_t stream);
size_t fa_pcg_batched_work_bytes(int64_t n, int nreq);
int fa_pcg_batched(const fa_sell_t* A, const fa_mg_t* M, int nreq, const int32_t* block_req,
                   const int32_t* block_off, const double* b, double* x, double atol, double rtol, int64_t maxit,
                   int32_t check_every, int64_t* niter_host, double* residual_host, int32_t* status_host,
                   void* work, size_t work_bytes, fa_stream_t stream);

/* ------------------------------------------------------ in-step timing --
 * Measurement only; no counterpart in the reference.  fa_timer_arm(mask)
 * arms the kernel families whose FA_TIMER_* bits are set (bit 0 must be
 * clear; mask 0 disarms) and clears their counts: from then on every launch
 * of an armed family made by this library is bracketed by two HIP events
 * recorded on the stream it is launched on (at most 128 launches per family
 * and arming; launches under stream capture are not timed).
 * fa_timer_read(kind) waits for the recorded events of one family and
 * returns how many launches were timed and their summed duration (ms).
 * Process-wide state, not thread-safe: for benchmarks, not for serving. */
enum {
    FA_TIMER_PCG_SPMV = 1,       /* k_pcg_spmv: the CG operator A p (+ p.Ap)   */
    FA_TIMER_P1_ROWS = 2,        /* P1 matrix, pass 1 (element rows per cell)  */
    FA_TIMER_P1_SUM_ROWS = 3,    /* P1 matrix, pass 2 (ordered CSR row sums)   */
    FA_TIMER_P1_VECTORS = 4,     /* P1 load vector, pass 1                     */
    FA_TIMER_P1_SUM_VALUES = 5,  /* P1 load vector, pass 2                     */
    FA_TIMER_BC_APPLY = 6,       /* Dirichlet filter + right-hand side         */
    FA_TIMER_SELL_SPMV = 7,      /* fa_sell_spmv                               */
    FA_TIMER_VCYCLE = 8,         /* one whole V-cycle (all its kernels)        */
    FA_TIMER_P1_APPLY = 9,       /* matrix-free P1 operator (fa_p1_apply)      */
    FA_TIMER_SMOOTH0 = 10,       /* finest-level FP16 smoothing steps and
                                    residual of the V-cycle (k_cheb16*,
                                    k_residual16 on level 0)                   */
    FA_TIMER_P1_BOX = 11         /* P1 matrix on a from_box grid, one pass     */
};
int fa_timer_arm(unsigned mask);
int fa_timer_read(int kind, int64_t* launches, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif /* FEALPY_AMD_H */
