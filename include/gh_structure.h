/*
 * gh_structure: host-side producer of multigrid problem structure for the
 * uniformly refined quad/hex meshes of the benchmark workloads.
 *
 * It stands in for Gascoigne's CPU side (MeshAgent + CGDisc), which hands the
 * GPU path its inputs: level meshes, CSR structure, assembled matrices, mesh
 * transfer maps and Dirichlet node lists.  Counterparts in the reference:
 *   - coarse mesh reader: HierarchicalMesh3d::read_inp
 *     (src/Mesh/hierarchicalmesh3d.cc:1403-1525), same .inp format
 *   - uniform refinement: HierarchicalMesh3d::global_refine / basic_refine3d
 *     (hierarchicalmesh3d.cc:373-472); numbering nested as in the reference
 *     (coarse nodes keep their numbers, c2f = identity), new nodes ordered
 *     cell midpoints, then face midpoints, then edge midpoints.  The order
 *     inside each group follows first encounter over the cells, not the
 *     reference's hash-map iteration order.
 *   - mesh transfer: GascoigneMeshTransferConstructor3d
 *     (src/Mesh/gascoignemeshtransferconstructor.cc:115-186)
 *   - CSR structure: CGDisc::Structure + ColumnStencil
 *     (src/DofHandler/cgdisc.h:299-315, src/LinAlg/columnstencil.cc:52-70)
 *   - Q1 assembly: CGDisc::Matrix with Laplace<DIM> (cgdisc.h:599-635,
 *     src/Problems/laplace.h) on axis-aligned boxes, coefficient per cell,
 *     then SparseBlockMatrix::dirichlet (src/LinAlg/sparseblockmatrix.xx:205-218)
 */
#ifndef GH_STRUCTURE_H
#define GH_STRUCTURE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ghs_hier ghs_hier;

const char* ghs_last_error(void);

/* Reads a 2D (quad) or 3D (hex) .inp mesh and refines it `prerefine` times.
 * Level numbering of the result: 0 = coarsest (the .inp mesh ... after
 * refinements), nlevels-1 = finest; nlevels = prerefine + 1 - drop_coarse. */
int ghs_build(const char* inp_path, int dim, int prerefine, int drop_coarse, ghs_hier** out);
int ghs_destroy(ghs_hier* h);
int ghs_nlevels(const ghs_hier* h);
/* sizes of level l: nodes, cells, CSR entries */
int ghs_level_sizes(const ghs_hier* h, int l, int64_t* nnodes, int64_t* ncells, int64_t* nnz);
int ghs_level_coords(const ghs_hier* h, int l, double* xyz /* nnodes*3 */);
int ghs_level_cells(const ghs_hier* h, int l, int32_t* cells /* ncells*(4|8) */);
/* CSR structure: rowptr[nnodes+1], col[nnz], sorted columns */
int ghs_level_structure(const ghs_hier* h, int l, int32_t* rowptr, int32_t* col);
/* Dirichlet nodes of the given colours, sorted ascending */
int ghs_level_dirichlet_count(const ghs_hier* h, int l, int ncolours, const int32_t* colours,
                              int64_t* count);
int ghs_level_dirichlet(const ghs_hier* h, int l, int ncolours, const int32_t* colours,
                        int32_t* nodes);
/* -div(k grad u) stiffness with one coefficient per cell (kcell == NULL: k = 1),
 * strong Dirichlet rows and columns of the given colours (identity diagonal) */
int ghs_level_assemble(const ghs_hier* h, int l, const double* kcell, int ncolours,
                       const int32_t* colours, double* val);
/* transfer maps from level l to level l-1 (l >= 1) */
int ghs_level_transfer_sizes(const ghs_hier* h, int l, int64_t* ncoarse, int64_t* n2,
                             int64_t* n4, int64_t* n8);
int ghs_level_transfer(const ghs_hier* h, int l, int32_t* c2f, int32_t* zweier,
                       int32_t* vierer, int32_t* achter);

/* CSR structure from a cell list (CGDisc::Structure without hanging nodes):
 * rowptr[n+1] is filled; col may be NULL to query nnz = rowptr[n]. */
int ghs_structure_from_cells(int64_t n, int64_t ncells, int npc, const int32_t* cells,
                             int32_t* rowptr, int32_t* col);
/* Q1 element-matrix assembly on a given mesh (boxes), same rules as
 * ghs_level_assemble; rowptr/col from ghs_structure_from_cells. */
int ghs_assemble_on(int dim, int64_t n, const double* xyz, int64_t ncells, const int32_t* cells,
                    const double* kcell, const int32_t* rowptr, const int32_t* col,
                    int64_t ndir, const int32_t* dir, double* val);
/* A node order for the ILU smoother (gh_ilu_create, gh_mg_set_ilu) whose
 * triangular solves run in few levels: greedy colouring of the matrix graph in
 * node order (each node takes the smallest colour none of its coupled nodes
 * holds), perm[] = the nodes of colour 0, then colour 1, ..., ascending inside
 * a colour; *ncolours = the colours used (8 on a uniform Q1 hexahedral
 * level).  No counterpart in the reference, whose default ILU order comes from
 * METIS_NodeND (src/Solver/cuthillmckee.cc:70-121, through
 * StdSolver::PermutateIlu, src/Solver/stdsolver.cc:1598-1625). */
int ghs_colour_order(int64_t n, const int32_t* rowptr, const int32_t* col, int32_t* perm,
                     int32_t* ncolours);

/* Patches of level l (l must come from a refinement): the 2x2x2 (2x2) children
 * of one coarser cell, nodes in patch-local lexicographic order
 * a + 3b + 9c (x fastest), the element order of Gascoigne's patch dof
 * handler (used by VankaSmoother::ConstructStructure,
 * src/LinAlg/vankasmoother.cc:47-60).  nodes[npatches*27] (9 in 2D);
 * nnz = entries of the patch-based CSR structure (CGDisc<DIM,2,...>::Structure
 * of the patch discretisations, src/DofHandler/cgdisc.h:299-315,1314-1319). */
int ghs_level_patch_sizes(const ghs_hier* h, int l, int64_t* npatches, int64_t* nnz);
int ghs_level_patches(const ghs_hier* h, int l, int32_t* nodes);
int ghs_level_patch_structure(const ghs_hier* h, int l, int32_t* rowptr, int32_t* col);
/* Stokes with local projection stabilisation on patches (NavierStokesLps<3>
 * at U = 0, symmetrictensor 0, discretisation CGDiscQ13dLps:
 * src/Problems/navierstokes.h:92-270, src/DofHandler/elementlpsintegrator.h),
 * components (p, u, v, w), 4x4 blocks val[nnz*16] (row-major per block) on
 * the patch structure.  Nodes on the given colours get strong Dirichlet
 * conditions for the components in the bit mask dircomps
 * (SparseBlockMatrix::dirichlet).  With u0 (n*4, the Dirichlet lifting) the
 * residual rhs = -A u0 of the unconstrained operator is returned, zero in the
 * constrained entries (the Newton residual StdMultiLevelSolver hands to the
 * linear solve); u0 and rhs are both NULL or both given. */
int ghs_level_assemble_stokes(const ghs_hier* h, int l, double visc, double alpha, int ncolours,
                              const int32_t* colours, int dircomps, const double* u0, double* rhs,
                              double* val);
/* Laplace on the triquadratic patch element (CGDiscQ23d with Laplace<3>):
 * scalar values val[nnz] on the patch structure, strong Dirichlet rows and
 * columns on the nodes of the given colours (identity diagonal). */
int ghs_level_assemble_q2(const ghs_hier* h, int l, int ncolours, const int32_t* colours,
                          double* val);
/* Adaptive Q2 patch hierarchy: the .inp mesh refined `prerefine` times, then
 * `nadapt` (0..2) steps of MeshAgent::refine_nodes (src/Mesh/meshagent.cc:
 * 540-556) on the nodes inside the ball |x - centre| < radius, with patch
 * depth 1: the nodes' patches marked as HierarchicalMesh::vertex_patch_refine
 * marks them, closed so that the levels at any patch corner differ by one at
 * most (src/Mesh/hierarchicalmesh.cc:272-290, hierarchicalmesh3d.cc:
 * 1938-2254), every cell of a marked patch refined.  The multigrid levels
 * are those GascoigneMeshConstructor derives from the refined mesh
 * (LevelMesh3d::construct_lists, src/Mesh/levelmesh3d.cc:703-831): each
 * coarser level merges every group of 64 cells of one grandfather into its 8
 * children where the level jumps allow.  The unit box only (node coordinates
 * on the dyadic lattice).  Levels otherwise as in ghs_build; levels with
 * hanging nodes carry the Q2 constraints. */
/* The hierarchy of a Gascoigne hierarchical mesh file (.gup, written by
 * MeshAgent::write_gup, src/Mesh/meshagent.cc; read as
 * HierarchicalMesh3d::read_gup, src/Mesh/hierarchicalmesh3d.cc:1615-1700
 * does): its macro mesh (level-0 hexes and boundary quads with their
 * colours) and its finest cells, with the multigrid levels Gascoigne derives
 * from them (as ghs_build_adaptive; Q2 hanging-node constraints where cells
 * of different levels meet).  The unit box only: one macro hex, cells on the
 * dyadic lattice. */
int ghs_build_gup(const char* gup_path, int drop_coarse, ghs_hier** out);
int ghs_build_adaptive(const char* inp_path, int prerefine, int drop_coarse, int nadapt,
                       double radius, const double* centre /* 3 */, ghs_hier** out);
/* Q2 hanging-node constraints of level l (HNStructureQ23d layout):
 * edges[ne*4] = hanging node, near, middle, far parent (weights 3/8, 3/4,
 * -1/8); faces[nf*10] = hanging node, 9 parents g_{3a+b} (weights w_a w_b). */
int ghs_level_hn_sizes(const ghs_hier* h, int l, int64_t* nedges, int64_t* nfaces);
int ghs_level_hn(const ghs_hier* h, int l, int32_t* edges, int32_t* faces);
/* Q2 Laplace on level l condensed over its hanging nodes (T^T A T: hanging
 * rows/columns become identity rows, as in the reference's assembled
 * matrices), then strong Dirichlet rows/columns of the given colours.  The
 * first call forms it and returns nnz; the second copies it out (rowptr[n+1],
 * col[nnz] ascending, val[nnz]) and releases it. */
int ghs_level_q2_condensed_nnz(ghs_hier* h, int l, int ncolours, const int32_t* colours,
                               int64_t* nnz);
int ghs_level_q2_condensed(ghs_hier* h, int l, int32_t* rowptr, int32_t* col, double* val);
/* m[i] = integral of the bilinear hat of node i over the rectangular boundary
 * faces of the given colour (boundary mass row sums), m[nnodes] */
int ghs_level_face_mass(const ghs_hier* h, int l, int colour, double* m);

#ifdef __cplusplus
}
#endif
#endif
