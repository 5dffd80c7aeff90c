/*
 * fvmr.h -- C ABI of the MI355X (gfx950) explicit finite-volume
 * residual-and-update path of FVMR (OliveTIen/CFD_Solver_2D_GPU).
 *
 * The ABI replaces the reference's GPU solver surface:
 *   - the solver container  GPU::GPUSolver2      (FVMR/src/solvers/GPUSolver2.h:32-89)
 *   - the per-step driver    CIteration           (FVMR/src/time/CIteration.h:8-9)
 *   - the mesh preprocessing SU2MeshReader/FVM_2D  (FVMR/src/input/SU2MeshReader.cpp:10-251,
 *                                                   FVMR/src/legacy/FVM_2D.cpp:25-198)
 *   - the AoS->SoA converter OldDataConverter     (FVMR/src/solvers/OldDataConverter.cpp:3-195)
 *   - the initial fields     FieldInitializer      (FVMR/src/input/FieldInitializer.cpp:22-177)
 *   - the residual norms     get_residual_host     (FVMR/src/output/residual/Residual.cpp:4-43)
 *
 * Conventions are the reference's: 0-based SoA indices, -1 sentinels, 1-based
 * edge setID (-1 = interior), element slot 3 = -1 (triangles), BC/scheme codes
 * of FVMR/src/global/Constexpr.h.  Arrays are SoA (one pointer per component).
 * Every function returns an int status (FVMR_OK == 0) and never calls exit();
 * the message of the last failure on the calling thread is fvmr_last_error().
 * Host buffers are caller-owned; device buffers are library-owned.  One solver
 * handle owns one HIP stream; a handle is thread-compatible, not thread-safe.
 */
#ifndef FVMR_H
#define FVMR_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define FVMR_OK               0
#define FVMR_ERR_INVALID_ARG  1
#define FVMR_ERR_HIP          2
#define FVMR_ERR_MESH         3
#define FVMR_ERR_UNSUPPORTED  4
#define FVMR_ERR_IO           5
#define FVMR_ERR_NUMERIC      6

/* ---- codes of FVMR/src/global/Constexpr.h:13-57 ------------------------- */
#define FVMR_EQ_EULER            1
#define FVMR_EQ_NS               2
#define FVMR_SOL_LLF          2001
#define FVMR_SOL_ROE          2002
#define FVMR_REC_CONSTANT     3001
#define FVMR_REC_LINEAR       3002
#define FVMR_REC_MUSCL        3003   /* reference falls back to constant */
#define FVMR_GRA_LEASTSQUARE     1
#define FVMR_GRA_GREENGAUSS      2
#define FVMR_EVO_EXPLICIT     3101   /* forward Euler, the reference GPU path */
#define FVMR_EVO_RK3          3103   /* SSP-RK3 (Shu-Osher) in this library   */
#define FVMR_LIM_NONE            0
#define FVMR_LIM_BARTH           1
#define FVMR_LIM_VENKATAKRISHNAN 4   /* new: not in the reference */
#define FVMR_BC_WALL_NONVISCOUS 6001
#define FVMR_BC_WALL_ADIABAT    6002
#define FVMR_BC_WALL_ISOTHERMAL 6003 /* declared by the reference, produced here */
#define FVMR_BC_INLET           6011
#define FVMR_BC_OUTLET          6012
#define FVMR_BC_INF             6013
#define FVMR_BC_DSR             6014
#define FVMR_BC_SYMMETRY        6021
#define FVMR_BC_PERIODIC_0      6100 /* 6100..6109 */

/* Device cell order (results always cross the ABI in the caller's order). */
#define FVMR_RENUMBER_NONE      0    /* the caller's order */
#define FVMR_RENUMBER_RCM       1    /* reverse Cuthill-McKee */
#define FVMR_RENUMBER_HILBERT   2    /* Hilbert curve over the centroids (default) */

/* Device kernel family.  TILE: one work-group per tile of cells, gradients kept in LDS,
 * every face flux evaluated once; it needs the standard geometry (centroids, midpoints,
 * normals recomputable from the nodes, float-valued lengths, as fvmr_mesh_* produce).
 * CELL: one gradient and one stage kernel per RK stage over packed per-cell records.
 * AUTO: TILE when the geometry allows it, else CELL.  Both give the same results. */
#define FVMR_KERNELS_AUTO       0
#define FVMR_KERNELS_CELL       1
#define FVMR_KERNELS_TILE       2

/* precision of the device path */
#define FVMR_PREC_FP64        0     /* plain: every array and operation FP64 */
#define FVMR_PREC_MIXED       1     /* FP32 flux/reconstruction/limiter, FP64
                                       state, residual accumulation, RK update
                                       and dt reduction */
/* viscosity law */
#define FVMR_MU_SUTHERLAND    0     /* reference (always on there) */
#define FVMR_MU_CONSTANT      1     /* new: mu = mu0 */

/* residual norm types, FVMR/src/output/residual/Residual.h */
#define FVMR_NORM_1           1
#define FVMR_NORM_2           2
#define FVMR_NORM_INF         3

/* field initialiser types, FVMR/src/input/FieldInitializer.h:8-13 */
#define FVMR_INIT_UNIFORM        1
#define FVMR_INIT_ISENTROPIC_VORTEX 2
#define FVMR_INIT_SHOCK_TUBE     3
#define FVMR_INIT_DSR            4

/* synthetic mesh families (fvmr_mesh_generate) */
#define FVMR_GEN_PERIODIC_SQUARE 1  /* [x0,x0+L]^2, periodic_0 bottom/top, periodic_1 left/right */
#define FVMR_GEN_WEDGE_CHANNEL   2  /* [0,Lx]x[0,Ly] channel with a ramp on the bottom wall */
#define FVMR_GEN_CYLINDER_OGRID  3  /* O-grid around a cylinder, wall + farfield */
#define FVMR_GEN_NACA_OGRID      4  /* O-grid around a NACA 4-digit airfoil */
#define FVMR_GEN_RECT_CHANNEL    5  /* jittered channel, inlet/outlet/slip walls */

/*
 * Mesh in the reference SoA layout: GPU::NodeSoA (FVMR/src/gpu/datatype/NodeSoA.h:9-52),
 * GPU::ElementSoA (ElementSoA.h:15-105), GPU::EdgeSoA (EdgeSoA.h:9-100) and
 * GPU::BoundarySetMap (BoundarySetMap.h:10-70).  The IDs of all three are the
 * array indices, so the ID arrays are omitted.
 */
typedef struct fvmr_mesh_soa {
    int num_node;
    int num_element;
    int num_edge;
    int num_boundary_set;
    const double *node_xy[2];
    const int    *element_nodes[4];      /* slot 3 = -1 */
    const int    *element_edges[4];      /* slot 3 = -1 */
    const int    *element_neighbors[4];  /* local-edge order; -1 = none (periodic too) */
    const double *element_xy[2];         /* centroid */
    const double *element_volume;        /* area */
    const int    *edge_nodes[2];
    const int    *edge_setID;            /* 1-based boundary set, -1 interior */
    const int    *edge_periodicPair;     /* pair edge index or -1 */
    const int    *edge_elementL;
    const int    *edge_elementR;         /* -1 boundary; periodic -> pair's L */
    const double *edge_length;
    const double *edge_xy[2];            /* midpoint */
    const double *edge_normal[2];        /* (dy,-dx)/length, L -> R */
    const int    *boundary_type;         /* type of set s is boundary_type[s-1] */
} fvmr_mesh_soa;

/* Inputs the reference reads from GlobalPara (FVMR/src/global/GlobalPara.h:10-98). */
typedef struct fvmr_params {
    double gamma;            /* constant.gamma */
    double R;                /* constant R (287.06), "rcpcv" in the kernels */
    double Pr;               /* constant.Pr: used by the viscous dt term */
    double Pr_viscous;       /* Prandtl number of the viscous flux (reference: 0.72) */
    double T0;               /* constant.T0: Sutherland reference temperature */
    double mu0;              /* constant.mu0 (or rho*U*L/Re) */
    double CFL;              /* time.CFL */
    int    equation;         /* FVMR_EQ_* */
    int    flux_scheme;      /* FVMR_SOL_* */
    int    reconstruct;      /* FVMR_REC_* */
    int    gradient;         /* FVMR_GRA_* */
    int    limiter;          /* FVMR_LIM_* */
    int    time_advance;     /* FVMR_EVO_* */
    int    viscosity_model;  /* FVMR_MU_* */
    int    precision;        /* FVMR_PREC_* */
    int    renumber;         /* device cell order: FVMR_RENUMBER_* */
    int    compat_inv2x2;    /* 1: reproduce the reference inv_2x2 determinant bug (LSQ) */
    int    kernels;          /* FVMR_KERNELS_*: device kernel family */
    int    tile_cells;       /* cells per tile of the tile kernels (0: the work-group shape's default, 150
                                in FP64 and 300 in mixed precision for the default shapes; at most 400,
                                capped by the shape's LDS) */
    int    tile_threads;     /* work-items per work-group of the tile kernels (0: the default, 256 in FP64
                                and 512 in mixed precision; 128 or 512 in FP64, 256 in mixed precision) */
    int    reserved;         /* 0 */
    double inf_ruvp[4];      /* boundaryCondition.2D.inf    (rho,u,v,p) */
    double inlet_ruvp[4];    /* boundaryCondition.2D.inlet  */
    double outlet_ruvp[4];   /* boundaryCondition.2D.outlet */
    double wall_temperature; /* isothermal wall temperature (FVMR_BC_WALL_ISOTHERMAL) */
    double venkat_K;         /* Venkatakrishnan constant K */
    double dsr_shock_x;      /* double-Mach-reflection shock foot */
    double dsr_shock_y;
    double dsr_shock_speed;  /* reference: 20.0 */
} fvmr_params;

/* Per-step report of fvmr_step/fvmr_advance. */
typedef struct fvmr_step_info {
    double t;                /* physical time after the step(s) */
    double dt;               /* dt of the last step (after the T-t clamp) */
    double residual[4];      /* norm of dU/dt of the last step (stage 1), if requested */
    int    residual_norm;    /* in: 0 = do not compute, else FVMR_NORM_* */
    int    n_restricted;     /* cells repaired by the restriction in the last step */
    int    steps_done;
    int    reserved;
} fvmr_step_info;

/* Parameters of the synthetic mesh generators (fvmr_mesh_generate). */
typedef struct fvmr_gen_params {
    int    ni, nj;           /* cells (quads) along the two index directions */
    double a, b, c, d;       /* family-specific lengths, see fvmr_mesh_generate */
    double jitter;           /* interior node jitter, fraction of local spacing */
    unsigned long long seed; /* splitmix64 seed for the jitter */
    int    naca_digits;      /* e.g. 12 for NACA0012 */
    int    reserved;
} fvmr_gen_params;

typedef struct fvmr_mesh   fvmr_mesh;    /* host mesh (SU2 ingest / generator) */
typedef struct fvmr_solver fvmr_handle;  /* device solver */

/* ---------------------------------------------------------------------- */
/* library                                                                 */
/* ---------------------------------------------------------------------- */
const char *fvmr_last_error(void);
/* Defaults of FVMR/src/global/GlobalPara.cpp:4-90 (Roe, GG, Barth, linear, Euler, explicit). */
int fvmr_params_default(fvmr_params *p);
/* freestream from Mach/AoA: AirParameterConverter (FVMR/src/input/AirParameterConverter.cpp:3-24) */
int fvmr_ruvp_from_mach(double Ma, double AoA_deg, double T0, double p0, double gamma, double R,
                        double ruvp[4]);

/* ---------------------------------------------------------------------- */
/* host mesh: SU2MeshReader::read_mesh_and_process + FVM_2D::ini* +        */
/* OldDataConverter::Convert_FVM2D_to_HostData, with a hashed edge builder */
/* that reproduces the reference's first-appearance edge order.           */
/* ---------------------------------------------------------------------- */
/* SU2MeshReader::read_mesh_and_process (FVMR/src/input/SU2MeshReader.cpp:10-19).
 * equation selects the wall type of "wall"-named markers (BoundaryManager.cpp:21-24). */
int fvmr_mesh_read_su2(const char *path, int equation, double scale, fvmr_mesh **out);
/* Same pipeline from in-memory lists (SU2 semantics: node/element IDs are the
 * file's 0-based ids; each marker is a list of node-id pairs).
 * elem_types: 5 triangle, 9 quad (split (0,1,3),(3,1,2) as SU2MeshReader.cpp:103-119);
 * elem_conn holds 4 ids per element (ignored 4th for triangles). */
int fvmr_mesh_from_lists(int num_node, const double *x, const double *y, const int *node_ids,
                         int num_elem, const int *elem_types, const int *elem_conn,
                         int num_marker, const char *const *marker_names,
                         const int *marker_offsets, const int *marker_edge_nodes,
                         int equation, fvmr_mesh **out);
/* Deterministic synthetic meshes (see FVMR_GEN_*). */
int fvmr_mesh_generate(int family, const fvmr_gen_params *g, int equation, fvmr_mesh **out);
/* Write a mesh produced by fvmr_mesh_generate/fvmr_mesh_from_lists as SU2 text. */
int fvmr_mesh_write_su2(const fvmr_mesh *m, const char *path);
/* ContinueFileReader::readContinueFile_1 + process (FVMR/src/input/ContinueFileReader.cpp:19-158,
 * 457-494): the mesh of a continue file (nodes, element node lists, boundary sets given by edge
 * ID) and its header; t, istep, CFL may be NULL.  fvmr_mesh_continue_state copies the element
 * states it holds (caller's element order = file order). */
int fvmr_mesh_read_continue(const char *path, int equation, fvmr_mesh **out, double *t, int *istep, double *CFL);
int fvmr_mesh_continue_state(const fvmr_mesh *m, double *const U[4]);
/* FieldWriter::writeContinueFile_1 (FVMR/src/output/FieldWriter.cpp:264-326): the reference's
 * layout (header, nodes, elements with node IDs and U, boundary sets with edge IDs ten per
 * line); numbers with 17 significant digits, so a restart reproduces the state bit for bit. */
int fvmr_write_continue(const fvmr_mesh *m, const char *path, double t, int istep, double CFL, const double *const U[4]);
int fvmr_mesh_soa_view(const fvmr_mesh *m, fvmr_mesh_soa *out);
/* BoundaryV2 edge sets (FVMR/src/gpu/datatype/BoundaryV2.h:13-23): name, type and edge list of set s (0-based). */
int fvmr_mesh_boundary_set(const fvmr_mesh *m, int s, const char **name, int *type,
                           const int **edges, int *num_edges);
/* Post-CCW-swap element node IDs (1-based file ids +1, as FVM_2D keeps them). */
int fvmr_mesh_element_node_ids(const fvmr_mesh *m, const int **ids3);
void fvmr_mesh_free(fvmr_mesh *m);

/* FieldInitializer::setInitialAndBoundaryCondition (FVMR/src/input/FieldInitializer.cpp:22-69):
 * args: vortex {x, y, strength}; shock tube {x, y, nx, ny}; DSR {shock_x, shock_y, angle_deg}. */
int fvmr_field_init(const fvmr_mesh_soa *m, const fvmr_params *p, int type, const double *args,
                    double *const U[4]);

/* ---------------------------------------------------------------------- */
/* device solver: GPUSolver2 + CIteration::iteration_device               */
/* ---------------------------------------------------------------------- */
/* GPUSolver2::allocateMemory + initializeData + host_to_device (GPUSolver2.cpp:54-165,232-239).
 * Copies the mesh to the device (renumbered when p->renumber), state set to zero. */
int fvmr_create(const fvmr_mesh_soa *m, const fvmr_params *p, fvmr_handle **out);
/* GPUSolver2::freeMemory (GPUSolver2.cpp:241-292). */
int fvmr_destroy(fvmr_handle *h);
/* elementField U host->device / device->host (GPUSolver2.cpp:232-239; CDriver.cpp:107-112);
 * arrays are in the caller's (reference) element order. */
int fvmr_set_state(fvmr_handle *h, const double *const U[4]);
int fvmr_get_state(fvmr_handle *h, double *const U[4]);
/* CIteration::iteration_device (CIteration.cpp:89-161): ruvp/Uold -> global dt ->
 * clamp to T - *t -> evolve (forward Euler or SSP-RK3) -> restriction.  *t advanced. */
int fvmr_step(fvmr_handle *h, double *t, double T, fvmr_step_info *info);
/* nsteps x fvmr_step with dt and t kept on the device (no host round trip per step). */
int fvmr_advance(fvmr_handle *h, int nsteps, double *t, double T, fvmr_step_info *info);
/* parity hook: calculateFunctionF_device (EvolveGPU.cu:40-51): rhs = dU/dt of the current state. */
int fvmr_stage_rhs(fvmr_handle *h, double *const rhs[4]);
/* parity hook: Gradient_2 (GradientGPU.cu:513-551): limited gradients of the current state. */
int fvmr_gradient(fvmr_handle *h, double *const Ux[4], double *const Uy[4]);
/* parity hook: get_global_dt_device (CalculateDtGPU.cu:97-103) of the current state (no clamp). */
int fvmr_global_dt(fvmr_handle *h, double *dt);
/* get_residual_host (Residual.cpp:4-43) of the last step's rhs, computed on the device. */
int fvmr_residual(fvmr_handle *h, int norm, double out[4]);
/* The cell permutation used on the device: perm[new] = reference index. */
int fvmr_permutation(fvmr_handle *h, int *perm);
/* parity hook: the FP64 quotient a[i] / b[i] and reciprocal 1 / b[i] as the device kernels form
 * them (hardware reciprocal refined by Newton-Raphson, the quotient corrected by its residual),
 * on n operand pairs on the current device.  Equal to IEEE division on the solver's operands;
 * b = 0 or b = +-inf give NaN where IEEE gives +-inf or 0 (no solver quotient has such a
 * divisor on a valid state). */
int fvmr_device_division(int n, const double *a, const double *b, double *quotient, double *reciprocal);

/* ---------------------------------------------------------------------- */
/* batches: many independent cases advanced together (no reference         */
/* counterpart: the reference runs one case per process)                    */
/* ---------------------------------------------------------------------- */
/* One handle over ncases meshes, packed one after the other (case c's cells keep their order
 * within the case; the device renumbers inside each case and never lets a tile cross a case).
 * The cases must share every setting of fvmr_params except the boundary states
 * (inf/inlet/outlet_ruvp) and CFL, which are per case; each case has its own dt and time.
 * Needs the tile kernels.  fvmr_destroy releases it. */
int fvmr_create_batch(int ncases, const fvmr_mesh_soa *const *meshes, const fvmr_params *params, fvmr_handle **out);
/* Number of cases and their cell ranges in the packed numbering (ncases + 1 entries). */
int fvmr_batch_cases(fvmr_handle *h, int *ncases, int *case_cell_begin);
/* State of case c in its own element order. */
int fvmr_batch_set_state(fvmr_handle *h, int c, const double *const U[4]);
int fvmr_batch_get_state(fvmr_handle *h, int c, double *const U[4]);
/* nsteps steps of every case: t[c] in/out, each case's dt clamped to T[c] - t[c]. */
int fvmr_batch_advance(fvmr_handle *h, int nsteps, double *t, const double *T);
/* Case c after the last fvmr_batch_advance: t, dt of its last step and the cells of the case
 * that the restriction repaired in that step (the other fields are left as they are). */
int fvmr_batch_step_info(fvmr_handle *h, int c, fvmr_step_info *info);

/* ---------------------------------------------------------------------- */
/* output quantities (FieldWriter, FVMR/src/output/FieldWriter.cpp)         */
/* ---------------------------------------------------------------------- */
/* Node fields of the current state, formed on the device: node ruvp = mean of the primitives
 * of the cells listing the node, summed in the caller's element order (update_nodeField_ruvp,
 * FieldWriter.cpp:371-418); Cp = (p - p_inf) / (0.5 rho_inf |V_inf|^2) with the handle's
 * inf_ruvp and Ma = |V| / sqrt(gamma p / rho) (update_nodeField_other_variables, :420-452).
 * num_node long, caller's node order; ruvp, Cp or Ma may be NULL.  (The reference forms them
 * from the primitives refreshed at the start of the last step, one step behind its U; call
 * this before fvmr_step to reproduce that.) */
int fvmr_node_field(fvmr_handle *h, double *const ruvp[4], double *Cp, double *Ma);
/* Pressure force coefficients per boundary set from the node Cp of the current state
 * (ReferenceData::calculate_edgeSet_force, FieldWriter.cpp:174-205): set s lists the caller's
 * edge indices set_edges[set_offsets[s] .. set_offsets[s+1]-1] in set order (the reference's
 * BoundaryV2::edgeSets); Cx[s] = sum of 0.5 (Cp_0 + Cp_1) (1/S_ref) L nx (the reference's CD),
 * Cy[s] likewise with ny (CL).  Edges must belong to a boundary set. */
int fvmr_surface_force(fvmr_handle *h, int nsets, const int *set_offsets, const int *set_edges, double S_ref,
                       double *Cx, double *Cy);
/* fvmr_surface_force with the same arguments and results, bit for bit, that allocates nothing
 * (no hipMalloc, hipFree or allocating library call): no counterpart in the reference; it is for
 * callers that cannot allocate device memory (a lift evaluated inside a captured or
 * allocation-free loop).  Its staging (offsets, edge ids, results) lives in a fixed device array
 * of the library, one per device, held until the process ends: at most 256 sets and 131072 edges
 * per call (larger calls return FVMR_ERR_INVALID_ARG).  Calls on one device share that array,
 * so calls that overlap in time on one device (other threads or streams) race. */
int fvmr_surface_force_noalloc(fvmr_handle *h, int nsets, const int *set_offsets, const int *set_edges,
                               double S_ref, double *Cx, double *Cy);

/* GUI vertex colours of a cell quantity (GPU::setVerticesColor, CudaGrid_kernels.cu:185-232, with
 * getColor_device :22-65 and the kernels :77-123): per node the largest of q_c * n over its n
 * cells (from 0) times (float)(1.0 / n), q = primitive var of the current state as float
 * (0 rho, 1 u, 2 v, 3 p); normalised by [lo, hi] (the reference's render range_1, range_2) and
 * coloured from npts control points colormap[4 i .. 4 i + 3] = {r, g, b, position} (positions
 * increasing; at most 64) by the reference's cubic blend; written as r, g, b to
 * vertices[6 v + 3 .. 6 v + 5] (the x, y, z, r, g, b vertex buffer of the renderer; x, y, z are
 * left as they are).  on_device = 1: vertices is device memory (a mapped GL buffer), 0: host
 * memory. */
int fvmr_vertex_colors(fvmr_handle *h, int var, float lo, float hi, int npts, const float *colormap, float *vertices,
                       int on_device);

/* ---------------------------------------------------------------------- */
/* steady states and the discrete adjoint (no reference counterpart)       */
/* ---------------------------------------------------------------------- */
typedef struct fvmr_adjoint_result {
    double J;            /* lift functional of the state (below) */
    double dJ_dalpha;    /* total derivative, per radian of the freestream angle */
    double res0, res;    /* L2 of (dr/dU)^T psi + dJ/dU at psi = 0 and at the end */
    int    iters;        /* pseudo-time iterations done */
} fvmr_adjoint_result;

/* Replace the boundary states of a one-case handle (the reference reads them once from
 * input.toml; here a freestream angle can be varied without a new handle). */
int fvmr_set_boundary_states(fvmr_handle *h, const double inf[4], const double inlet[4], const double outlet[4]);
/* Drive the current state towards r(U) = 0 by pseudo-time iteration with per-cell time steps
 * dt_c = CFL V_c / sum over the cell's faces of (|u.n| + a) L (face-mean primitives; NS adds a
 * viscous term), the handle's time scheme and no restriction.  Stops after max_iters or once the L2 norm of the
 * density residual, checked every 50 iterations, is at most tol times its first value;
 * res[0] = first norm, res[1] = last (res may be NULL).  FP64 one-case handles. */
int fvmr_steady_solve(fvmr_handle *h, int max_iters, double tol, int *iters, double res[2]);
/* Products with the transposed residual Jacobian at the current state U (caller's cell and node
 * order): out = (dr/dU)^T psi; *dalpha = psi . dr/dalpha, alpha the freestream angle
 * (d(u_inf, v_inf)/dalpha = (-v_inf, u_inf)); dxy[2v], dxy[2v+1] = psi . dr/d(x_v, y_v).
 * dalpha and dxy may be NULL.  FP64 one-case Euler handles without periodic, double Mach
 * reflection or isothermal boundaries. */
int fvmr_adjoint_vjp(fvmr_handle *h, const double *const psi[4], double *const out[4], double *dalpha, double *dxy);
/* Discrete adjoint of the lift functional J = sum over the listed sets of
 * (-sin a Cx + cos a Cy), a = atan2(v_inf, u_inf), Cx, Cy as fvmr_surface_force, at the current
 * (steady) state: psi solves (dr/dU)^T psi = -dJ/dU through the pseudo-time iteration
 * psi <- psi + D ((dr/dU)^T psi + dJ/dU) (per-cell dt of the state, the handle's time scheme, from
 * psi = 0), stopped after max_iters iterations or once the L2 norm of all four components of
 * (dr/dU)^T psi + dJ/dU, checked every 50 iterations, is at most tol times its value at psi = 0.
 * Then dJ/dalpha and, when dJ_dxy is not NULL, dJ/d(x_v, y_v) (2 nv, interleaved).  Same
 * restrictions as fvmr_adjoint_vjp. */
int fvmr_adjoint_solve(fvmr_handle *h, int nsets, const int *set_offsets, const int *set_edges, double S_ref,
                       int max_iters, double tol, double *dJ_dxy, fvmr_adjoint_result *out);
/* The adjoint field psi of the last fvmr_adjoint_solve (caller's cell order). */
int fvmr_adjoint_state(fvmr_handle *h, double *const psi[4]);

/* ---------------------------------------------------------------------- */
/* measurement                                                             */
/* ---------------------------------------------------------------------- */
/* Replay runs of three steps of fvmr_advance / fvmr_batch_advance from captured HIP graphs
 * (default off; never while profiling).  No reference counterpart: the reference launches
 * every kernel from the host each step. */
int fvmr_use_graphs(fvmr_handle *h, int on);
/* Autotuner of the tile kernels (no reference counterpart): times `steps` steps of the uniform
 * freestream of p on mesh m for every candidate work-group shape and tile size (fvmr_autotune_
 * candidates) and returns the fastest pair (set them as p->tile_threads, p->tile_cells) and its
 * ms per step (best_ms may be NULL).  It holds one handle per candidate at once, keeps the GPU
 * busy for at least 40 ms before timing (clocks rise under load) and takes the faster of two
 * timed runs per candidate, one pass forward and one backward. */
int fvmr_autotune(const fvmr_mesh_soa *m, const fvmr_params *p, int steps, int *best_tile_threads,
                  int *best_tile_cells, double *best_ms);
/* The same for a batch (fvmr_create_batch's arguments): one batch handle per candidate; timings
 * (may be NULL) receives the ms per step of every candidate, in fvmr_autotune_candidates' order. */
int fvmr_autotune_batch(int ncases, const fvmr_mesh_soa *const *meshes, const fvmr_params *params, int steps,
                        int *best_tile_threads, int *best_tile_cells, double *best_ms, double *timings);
/* The tuners' candidates for a precision (FVMR_PREC_*): *n pairs (at most 16) of work-group size
 * and tile size. */
int fvmr_autotune_candidates(int precision, int *n, int *tile_threads, int *tile_cells);
/* 1: record HIP events around every kernel launch on the handle's stream. */
int fvmr_profile_enable(fvmr_handle *h, int on);
/* Average ms per launch and launch count of a kernel class since the last reset:
 * kernel: 0 = stage (k_tile or k_stage), 1 = k_tile_dt or k_gradient (with the dt partial in
 * stage 1), 2 = dt finish, 3 = restriction, 4 = residual. */
int fvmr_profile_query(fvmr_handle *h, int kernel, double *avg_ms, int *count);
int fvmr_profile_reset(fvmr_handle *h);
/* Wait for all work of the handle. */
int fvmr_synchronize(fvmr_handle *h);

#ifdef __cplusplus
}
#endif
#endif /* FVMR_H */
