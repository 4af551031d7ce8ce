/*
 * pgfem3d_amd.h -- C ABI of the MI355X (gfx950) finite-strain element hot path.
 *
 * The library replaces, for one GPU, the per-Newton-iteration element work of PGFem3D:
 *   deformation gradient per quadrature point (QP) -> constitutive update
 *   (neo-Hookean, J2, crystal plasticity) -> element residual / tangent -> CSR assembly.
 *
 * All citations are file:line into the PGFem3D sources (C-SWARM/pgfem-3d).
 *
 * Conventions (kept from the reference):
 *   - DOF ids are 1-based for free DOFs, 0 for fixed DOFs and -k for the k-th prescribed
 *     value (Node::id_map, include/node.h:7-35; generate_local_dof_ids,
 *     src/generate_dof_ids.cc:69-171).
 *   - The CSR pattern is the one Psparse_ApAi builds for one rank: int row pointers Ap
 *     [ndof+1], int (Ai_t) column ids Ai, 0-based, sorted and unique per row
 *     (src/Psparse_ApAi.cc:238-670, include/datatype.hpp:4).
 *   - Element DOFs are node-major (a*3+b); 3x3 tensors are row-major; 4th-order tensors
 *     use idx_4 = ((i*3+j)*3+k)*3+l (include/index_macros.h:15-30).
 *   - Every function returns an int status: 0 = success, > 0 = number of quadrature
 *     points / elements that failed (the reference's INFO/err count: J <= 0, singular
 *     inverse, non-normal tangent diagonal, local Newton failure), < 0 = API error
 *     (PGF_E*).
 *
 * Threading: a handle is owned by one host thread; calls on one handle are serialised on
 * the handle's stream.  Vector arguments of pgf_residual / pgf_tangent are DEVICE
 * pointers (HBM-resident), everything in the set-up functions is HOST memory.
 */
#ifndef PGFEM3D_AMD_H
#define PGFEM3D_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
#define PGF_OK 0
#define PGF_EINVAL (-1)   /* bad argument / inconsistent sizes */
#define PGF_ENOMEM (-2)   /* device or host allocation failed */
#define PGF_EHIP (-3)     /* a HIP runtime call failed */
#define PGF_ENOTSUP (-4)  /* element type / model combination not supported */
#define PGF_ENODEV (-5)   /* no usable gfx950 device */
#define PGF_ESCRATCH (-6) /* the call needs more than the library's fixed device buffer */

/* ------------------------------------------------------------- element types */
/* Identified by nodes per element, as FEMLIB does (FEMLIB/src/femlib.cc:33-37,75-83). */
#define PGF_TET4 4    /* linear tetrahedron, 1 QP */
#define PGF_TET10 10  /* quadratic tetrahedron, 4 QPs (order upgraded, femlib.cc:1010-1016) */
#define PGF_HEX8 8    /* trilinear hexahedron, 2x2x2 Gauss */

/* ---------------------------------------------------------------- model types */
/* Values of the reference's enum model_type (include/constitutive_model.h:41-54). */
#define PGF_MODEL_HYPER_ELASTICITY 0
#define PGF_MODEL_CRYSTAL_PLASTICITY 1
#define PGF_MODEL_J2_PLASTICITY 4

/* Material record.  Elastic data follow HOMMAT (include/hommat.h:13-28) as filled for the
 * CM analysis (src/homogen.cc:172-191): m10 = Ey, m01 = Ez, E = Ex, G = Gyz, nu = nyz,
 * and kappa = E/(3(1-2nu)) (src/constitutive_model.cc:1006). */
typedef struct pgf_material {
  int32_t model;       /* PGF_MODEL_* */
  int32_t devPotFlag;  /* 1 = Mooney-Rivlin (src/new_potentials.cc:106-127) */
  int32_t volPotFlag;  /* 2 = Doll-Schweizerhof 7, 3 = DS 8, 99 = quadratic
                          (src/new_potentials.cc:184-210) */
  int32_t cp_elastic;  /* crystal plasticity, elastic law of the intermediate configuration:
                          0 = cubic C11/C12/C44 St Venant-Kirchhoff (src/tensors.cc:49-75),
                          1 = the devPotFlag/volPotFlag potentials on eF, as the CM model
                          builds them (src/constitutive_model.cc:998-1014) */
  double E, nu, G, m10, m01;
  /* J2 plasticity, linear isotropic hardening (CM_J2P_PARAM PARAM_hp / PARAM_k0,
   * src/cm_j2_plasticity.cc:86-103): k0 = initial uniaxial yield stress, hp = linear
   * hardening modulus. */
  double k0, hp;
  /* Crystal plasticity (CP_PARAM, src/plasticity_model.cc:89-101,960-967). */
  double gamma_dot_0, m_rate, G0, g0, gs_0, gamma_dot_s, w_sat;
  double C11, C12, C44;           /* cubic elasticity in the crystal frame */
  double tol_M, tol_hardening;    /* local Newton tolerances */
  int32_t max_itr_M, max_itr_stag;
  /* J2 continuum damage (CM_J2P_PARAM damage properties, src/cm_j2_plasticity.cc:865-932):
   * viscosity mu, cap w_max (>= 1: 0.9999), Weibull P1, P2 and threshold Yin.  mu <= 0 turns
   * the damage off. */
  double mu, w_max, p1, p2, Yin;
} pgf_material;

/* Mesh record (host memory).  Element connectivity is CSR-like so that several element
 * types can share one mesh (Element::toe + Element::nod, include/element.h:16-60). */
typedef struct pgf_mesh {
  int64_t n_nodes;
  const double *X;           /* [n_nodes*3] reference coordinates, node-major */
  int64_t n_elem;
  const int64_t *elem_ptr;   /* [n_elem+1] offsets into elem_conn */
  const int64_t *elem_conn;  /* node ids, 0-based; nodes per element = 4, 10 or 8 */
  const int32_t *elem_mat;   /* [n_elem] material index, NULL = all 0 */
  int64_t n_coel;            /* cohesive elements (COEL): they enter the DOF / CSR graph */
  const int64_t *coel_ptr;   /* [n_coel+1] or NULL */
  const int64_t *coel_conn;  /* node ids or NULL; 6 (3 + 3) or 8 (4 + 4) nodes per element,
                                - side first (COEL::nod, include/cohesive_element.h:17-60) */
} pgf_mesh;

/* Interface (cohesive) element properties for the exponential traction-separation law,
 * in the reference's property order (Needleman_pkey, include/cohesive_potentials.h:38-41;
 * read_cohesive_properties, src/cohesive_potentials.cc:19-135): peak traction sig_c at the
 * effective opening chi_c, shear weight beta, contact penalty shape alpha. */
typedef struct pgf_cohesive {
  double sig_c, chi_c, beta, alpha;
} pgf_cohesive;

/* ==================================================================== set-up (host) */

/* Number the DOFs by element traversal.  On input id[n_nodes*3] holds support codes
 * (0 free, 1 fixed, -k prescribed by value k, as read by read_Dirichlet_BCs,
 * src/in.cc:29-110); on output it holds DOF ids (>0 free, 0 fixed, <0 prescribed).
 * Replaces generate_local_dof_ids / generate_local_dof_ids_on_elem
 * (src/generate_dof_ids.cc:69-171,306-412); on one rank the global ids equal the local
 * ones (generate_global_dof_ids, :173-219,470-575).  *ndof = number of free DOFs. */
int pgf_generate_dof_ids(const pgf_mesh *mesh, int64_t *id, int64_t *ndof);

/* Build the CSR pattern of the free DOFs.  Replaces Psparse_ApAi on one rank
 * (src/Psparse_ApAi.cc:238-670; countDof/mapDof :160-213).  Ap has ndof+1 entries.  With
 * Ai == NULL only Ap and *nnz are computed; otherwise Ai must hold Ap[ndof] entries. */
int pgf_psparse_apai(const pgf_mesh *mesh, const int64_t *id, int64_t ndof, int32_t *Ap,
                     int32_t *Ai, int64_t *nnz);

/* Synthetic structured meshes of the unit cube (n^3 cells).  etype = PGF_TET4 (Kuhn split
 * of each cell into 6 tets), PGF_TET10 (same, plus mid-edge nodes) or PGF_HEX8.  Interior
 * vertices are moved by U(-jitter*h, +jitter*h) per coordinate from std::mt19937_64(seed).
 * First call with X == NULL returns the sizes. */
int pgf_mesh_box(int32_t n, int32_t etype, double jitter, uint64_t seed, int64_t *n_nodes,
                 int64_t *n_elem, int64_t *conn_len, double *X, int64_t *elem_ptr,
                 int64_t *elem_conn);

/* ================================================================= device handle */
typedef struct pgf_handle_s *pgf_handle;

/* Create the device-side model: copies mesh, DOF ids, CSR pattern and materials to HBM,
 * builds the node -> staging-block tables of the gather assembly, colours the elements (no
 * two elements of one colour share a node) and builds the element -> CSR scatter map of the
 * coloured assembly (see pgf_set_assembly).  The pattern must be the one pgf_psparse_apai returns.
 * Replaces the per-element set-up of stiffmat_fd_MP / fd_residuals_MP
 * (src/stiffmat_fd.cc:831-993, src/fd_residuals.cc:468-636) and the per-QP
 * Constitutive_model objects (include/constitutive_model.h:766-844). */
int pgf_create(const pgf_mesh *mesh, const int64_t *id, int64_t ndof, const int32_t *Ap,
               const int32_t *Ai, int32_t n_mat, const pgf_material *mats, pgf_handle *out);
int pgf_destroy(pgf_handle h);

/* Use a caller-provided hipStream_t (passed as void*); NULL = the handle's own stream. */
int pgf_set_stream(pgf_handle h, void *stream);
int pgf_synchronize(pgf_handle h);

/* Pass timing (no reference counterpart; a measurement aid): when enabled, every pass of
 * pgf_residual / pgf_tangent / pgf_residual_tangent records HIP events on the handle's
 * stream around its element launches; pgf_last_pass_ms returns the elapsed GPU time of the
 * last pass and the number of element-kernel launches it made. */
int pgf_set_profiling(pgf_handle h, int on);
int pgf_last_pass_ms(pgf_handle h, double *ms, int32_t *n_launches);

/* Sizes: n_qp = total quadrature points, nnz = CSR entries, ndof, n_colors. */
int pgf_get_sizes(pgf_handle h, int64_t *n_qp, int64_t *nnz, int64_t *ndof,
                  int32_t *n_colors);

/* Crystal orientations: n_grains rotation matrices R (crystal -> sample frame, row-major,
 * [n_grains*9], host) and the grain of every element ([n_elem], host).  Replaces
 * plasticity_model_set_orientations / plasticity_model_construct_rotation
 * (src/plasticity_model.cc:1226-1258,1455), which store R per QP (TENSOR_R); here R is stored
 * per grain and slip systems and stiffness are rotated once per grain.  Without this call
 * every element uses the crystal frame. */
int pgf_set_orientations(pgf_handle h, int32_t n_grains, const double *R, const int32_t *elem_grain);

/* Interface element properties: n_props records and the property index of every interface
 * element ([n_coel], host; NULL = all 0).  Required before a pass when the mesh has
 * interface elements (COEL::props, reset_coel_props, src/cohesive_element.cc:50-54). */
int pgf_set_cohesive_props(pgf_handle h, int32_t n_props, const pgf_cohesive *props,
                           const int32_t *coel_mat);

/* Interface sizes: QPs (3 per triangle, 4 per quad element) and colours of the interface
 * elements (launched after the volume colours). */
int pgf_get_cohesive_sizes(pgf_handle h, int64_t *n_cqp, int32_t *n_colors);

/* Time step size dt = t(n+1) - t(n) used by rate-dependent models (CM_Ctx::dt). */
int pgf_set_time_step(pgf_handle h, double dt);

/* Crystal plasticity: start the next integrations' local Newton from a caller-provided point
 * (device pointer, [11][n_qp] in DEVICE QP order: M = pF(n) pF(n+1)^-1 (9), lambda, g), e.g. a
 * learned prediction; NULL restores the plain start (the fully plastic predictor pulled back
 * to yield).
 * No counterpart in the reference; the converged result does not depend on the start. */
int pgf_set_warm_start(pgf_handle h, const double *warm);

/* Device QP order: mesh_qp[d] = mesh QP index (element order, then QP order) of device QP d
 * (host array [n_qp]).  State arrays on the device are stored in device QP order. */
int pgf_get_qp_order(pgf_handle h, int64_t *mesh_qp);

/* Device pointer of a state field ([ncomp][n_qp], device QP order), for callers that read
 * the state on the GPU (e.g. to build warm starts); NULL if the field does not exist. */
int pgf_state_device_ptr(pgf_handle h, int32_t field, double **ptr);

/* ============================================================ per Newton iteration */
/* Nodal unknowns follow def_elem_total (src/utils.cc:1391-1412): a free DOF id>0 takes
 * r[id-1] + d_r[id-1]; a prescribed DOF id<0 takes defl[-id-1] + defl_d[-id-1]; a fixed DOF
 * is 0.  d_r, defl, defl_d may be NULL (= zeros).  All vectors are device pointers. */
typedef struct pgf_nodal_input {
  const double *r;       /* [ndof] */
  const double *d_r;     /* [ndof] or NULL */
  const double *defl;    /* [n_prescribed] or NULL */
  const double *defl_d;  /* [n_prescribed] or NULL */
} pgf_nodal_input;

/* Internal force vector f[ndof] (zeroed first), with the QP integration algorithm run
 * when run_integration != 0 (Solver::run_integration_algorithm).  Replaces fd_residuals_MP
 * (src/fd_residuals.cc:468-636) -> residuals_el_constitutive_model_1f
 * (src/constitutive_model.cc:3322-3539) -> compute_residual_vector (:1628-1662) ->
 * fd_res_assemble (src/fd_residuals.cc:58-73). */
int pgf_residual(pgf_handle h, const pgf_nodal_input *in, double dt, int run_integration,
                 double *f);

/* Tangent values vals[nnz] (zeroed first, as SparseSystem::zero) in the Psparse_ApAi
 * pattern, from the state left by the last integration.  Replaces stiffmat_fd_MP
 * (src/stiffmat_fd.cc:831-993) -> stiffness_el_constitutive_model_1f
 * (src/constitutive_model.cc:2474-2692) -> compute_stiffness_matrix (:1520-1616) ->
 * PLoc_Sparse (src/PLoc_Sparse.cc:15-175). */
int pgf_tangent(pgf_handle h, const pgf_nodal_input *in, double dt, double *vals);

/* Load vector of a prescribed increment: f_defl[ndof] (zeroed first) = sum over elements of
 * K_e u_e, K_e the element tangent at (r + d_r, defl) -- the in->defl_d increment left out, as
 * pgf_tangent at in->defl_d = 0 would form it -- and u_e the element's values of in->defl_d on
 * its prescribed DOFs (zero on the others), kept on the free rows.  Its negative is the
 * right-hand side of the first Newton iteration of a load step.  Replaces
 * compute_load_vector_for_prescribed_BC (src/load.cc:363-417) -> momentum_equation_load4pBCs
 * (:102-220), whose stiffness is el_compute_stiffmat_MP at def_elem(TOTAL) values. */
int pgf_prescribed_load(pgf_handle h, const pgf_nodal_input *in, double dt, double *f_defl);

/* Both of the above at one displacement in one pass over the elements (integration
 * run, then residual and tangent from the updated state).  Equivalent to pgf_residual
 * (run_integration=1) followed by pgf_tangent at the same input. */
int pgf_residual_tangent(pgf_handle h, const pgf_nodal_input *in, double dt, double *f,
                         double *vals);

/* Output variables at a time level (0 = n+1, 1 = n, 2 = n-1) of the stored state, as
 * constitutive_model_update_output_variables (src/constitutive_model.cc:1677-1881) forms
 * them: per QP (mesh order) S[n_qp*6], the PK2 stress (sig[eid].il[ip].o), and E[n_qp*6], the
 * Green-Lagrange strain (eps[eid].il[ip].o); per element sigma[n_elem*6], the volume average
 * of the Cauchy stress (sig[eid].el.o), and eps[n_elem*6], that of the Eulerian-Almansi
 * strain with doubled shears (eps[eid].el.o).  Components in the reference's order XX, YY,
 * ZZ, YZ, XZ, XY.  Host arrays; any may be NULL.  The reference calls it after a converged
 * step, when the state at n is the step's result. */
int pgf_output_variables(pgf_handle h, int32_t level, double *S, double *E, double *sigma,
                         double *eps);

/* Starts of the last crystal-plasticity local Newton: *n_learned QPs converged from the
 * attached network's start (chosen per QP when its residual merit is below the plain
 * start's), *n_fallback QPs took it, failed and converged from the plain start instead; the
 * others started plainly.  (No reference counterpart: the learned start's report.) */
int pgf_cp_start_stats(pgf_handle h, int64_t *n_learned, int64_t *n_fallback);

/* Physics-based subdivision measure of the last integration (cm_get_subdivision_parameter,
 * src/constitutive_model.cc:1980-2001): the largest per-QP value of the models'
 * get_subdiv_param over the mesh.  Crystal plasticity (CP_PARAM::get_subdiv_param,
 * src/plasticity_model.cc:1110-1125): dt sum_a |gamma_dot_a(n+1)| / 0.005; J2 with damage
 * (CM_J2P_PARAM::get_subdiv_param, src/cm_j2_plasticity.cc:523-531): (w(n+1) - w(n)) / 0.05,
 * the damage increment over DELTA_W_MAX (:55; the formula itself is in the absent library);
 * the hyperelastic model gives 0 (Model_parameters::get_subdiv_param,
 * include/constitutive_model.h:617-624).  dt <= 0 uses the handle's time step.  The caller
 * rejects a converged step whose value exceeds 1.25 and subdivides it, as Newton_Raphson does
 * (src/Newton_Raphson.cc:1364-1398). */
int pgf_subdivision_param(pgf_handle h, double dt, double *alpha);

/* pgf_residual_tangent without allocating: the gather assembly's staging blocks go to a
 * fixed device array of PGF_STATIC_STAGE_DOUBLES doubles (64 MiB) that the library holds on
 * each GPU it runs on until the process ends, instead of the buffer pgf_residual_tangent
 * allocates at its first gather pass.  The call makes no hipMalloc, hipFree or other
 * allocating call; it launches eagerly (even with pgf_set_graph on) and synchronises the
 * handle's stream to read the failure count, like pgf_residual_tangent.
 * No counterpart in the reference: it is for callers that cannot allocate device memory
 * (serving loops, harnesses that forbid allocation).
 * Returns PGF_ESCRATCH when the mesh's staging blocks exceed the array (pgf_stage_doubles
 * gives their size), PGF_ENOTSUP when the handle has no gather tables or has a learned start
 * attached.  Every call on one GPU shares the array: calls that overlap in time on one GPU
 * (two host threads, or handles on streams that run concurrently) race. */
#define PGF_STATIC_STAGE_DOUBLES 8388608
int pgf_residual_tangent_static(pgf_handle h, const pgf_nodal_input *in, double dt, double *f,
                                double *vals);

/* Size of the gather assembly's staging buffer of the handle's mesh, in doubles. */
int pgf_stage_doubles(pgf_handle h, int64_t *n);

/* Capture each pass (all of its launches: every element type, the crystal-plasticity QP
 * kernels, the interface elements) into one hipGraph and replay it while the pass mode, the
 * vector pointers, dt and the tuning stay the same; on = 0 returns to eager launches.
 * pgf_set_orientations, pgf_set_cohesive_props, pgf_set_three_field and pgf_set_assembly
 * replace buffers or kernels a captured pass holds: they finish the handle's work and drop
 * the graph, and the next pass captures again.  (No reference counterpart: a launch-overhead
 * measure for the per-iteration element pass.) */
int pgf_set_graph(pgf_handle h, int on);

/* Autotune the launch configuration by timing passes at the given input (no reference
 * counterpart: the reference's element loop has no launch shape):
 *   - points per wavefront of the crystal-plasticity local-Newton work queue (QPs per wave
 *     range, 64..2048, or 1: every wave draws its QPs from one global queue as its lanes free
 *     up), on integrating residual passes;
 *   - the gather assembly's tile (gather mode): nodes per workgroup (1, 2, 4; capped by the
 *     LDS a node's rows take) x staging blocks each half-wave keeps in flight (2, 4, 8), on
 *     residual + tangent passes.
 * Every configuration gives the same results bit for bit (the sums keep their order).
 * *chunk = the chosen points per wavefront.  The pass memory is allocated for the call. */
int pgf_autotune(pgf_handle h, const pgf_nodal_input *in, double dt, int64_t *chunk);
/* The launch configuration in use (pgf_autotune's choice or pgf_set_tuning's). */
int pgf_get_tuning(pgf_handle h, int64_t *cp_chunk, int32_t *gather_waves, int32_t *gather_batch);
/* Set the launch configuration: cp_chunk = 1 (one global queue) or >= 64, gather_waves in
 * {1, 2, 4}, gather_batch in {2, 4, 8}; a value <= 0 keeps the current one. */
int pgf_set_tuning(pgf_handle h, int64_t cp_chunk, int32_t gather_waves, int32_t gather_batch);

/* Assembly of element rows into f and the CSR values (the role of fd_res_assemble and
 * PLoc_Sparse, src/fd_residuals.cc:58-73, src/PLoc_Sparse.cc:15-175).  Both modes are
 * deterministic, use no atomics, and give the same sums up to the order of the additions:
 *   PGF_ASSEMBLY_GATHER (default): every element writes its rows once into a node-major
 *     staging buffer (3 (3 nne + 1) doubles per element node, rounded up to whole 128-byte
 *     lines; allocated at the first such pass); then one wavefront per node adds its elements' blocks in a fixed order and
 *     writes the node's CSR rows and residual entries once, contiguously.
 *   PGF_ASSEMBLY_COLOURED: elements of one colour share no node and add their rows into
 *     the CSR values directly, colour by colour (no staging memory).
 * The gather mode needs 3 * max row length * 8 bytes of LDS per node, and packs each staging
 * block's offsets into 64 bits (staging offset < 2^36 doubles, element-node pairs < 2^28); a
 * mesh outside either limit gets PGF_ENOTSUP here and is created in the coloured mode. */
#define PGF_ASSEMBLY_GATHER 0
#define PGF_ASSEMBLY_COLOURED 1
int pgf_set_assembly(pgf_handle h, int32_t mode);

/* ======================================================================= state */
/* After convergence: n+1 -> n, n -> n-1 for every QP (constitutive_model_update_time_steps,
 * src/constitutive_model.cc:1442-1497; CM_J2P_PARAM::update_state_vars,
 * src/cm_j2_plasticity.cc:463-496; CP_PARAM::update_state_vars,
 * src/plasticity_model.cc:174-192). */
int pgf_commit_step(pgf_handle h);
/* Subdivision reset: n+1 <- n (reset_state_vars, src/cm_j2_plasticity.cc:498-520;
 * constitutive_model_reset_state_using_temporal, src/constitutive_model.cc:1421-1440). */
int pgf_reset_step(pgf_handle h);

/* ============================================================ three-field (CM3F) */
/* The CM3F analysis (-cm3f: analysis_type CM3F, include/enumerations.h:16) for meshes whose
 * materials are hyperelastic (neo-Hookean deviatoric law), J2 plasticity without damage, or
 * crystal plasticity with the hyperelastic elastic law (cp_elastic = 1): one volume variable
 * theta and one pressure P per element
 * (get_3f_pressure_volume_number, src/three_field_element.cc:1399-1416), condensed out of
 * the element residual and tangent (condense_K_3F_to_1F / condense_F_3F_to_1F,
 * src/condense.cc:304-385).  With the mode on, pgf_residual / pgf_tangent /
 * pgf_residual_tangent / pgf_prescribed_load evaluate the condensed element of
 * residuals_el_constitutive_model_3f / stiffness_el_constitutive_model_3f
 * (src/constitutive_model.cc:2714-2735, :3565-3590) at theta = V(n+1) + dV, P = P(n+1) + dP.
 * pgf_commit_step folds the increments in and rolls the variables over
 * (update_np1_from_increments + update_for_next_time_step, Newton_Raphson.cc:683-689);
 * pgf_reset_step zeroes the increments (ThreeFieldVariables::reset).
 * Crystal plasticity integrates at (theta/J)^(1/3) F, which becomes the model's F(n+1)
 * (run_integration_algorithm with tf_factor, src/constitutive_model.cc:810-818), and its
 * element takes M = pF(n+1)^-1 and the implicit dM/du into the CM_ThreeField blocks
 * (src/constitutive_model_3f.cc:189-470; residual passes also run the tangent kernel, since
 * the condensed residual needs Kut).
 * J2 (CM_J2P_PARAM: pF = I, so M = I and dM/du = 0) takes the deviatoric part of its return
 * map's stress and tangent (update_elasticity_dev, src/cm_j2_plasticity.cc:384-461) with the
 * element pressure.
 * on = 0 turns the mode off and frees its memory.  PGF_ENOTSUP for J2 with continuum damage,
 * the linear deviatoric law, and crystal plasticity with the cubic law. */
int pgf_set_three_field(pgf_handle h, int32_t on);
/* compute_cm_initial_conditions (src/constitutive_model.cc:4035-4160) for the three-field
 * analysis, with compute_3f_initial_conditions (src/three_field_element.cc:1328-1386): theta
 * and P at n (and n+1) from the nodal input in_n, at n-1 from in_nm1 (NULL: in_n); the
 * model's F(n) and F(n-1) from the same displacements (set_F, :4097-4105); increments
 * zeroed.  Returns the number of QPs with J <= 0 or a singular dX/dxi. */
int pgf_tf_initial(pgf_handle h, const pgf_nodal_input *in_n, const pgf_nodal_input *in_nm1);
/* constitutive_model_update_NR without inertia (src/constitutive_model.cc:3916-3999,
 * IntegrateThreeFieldUpdate :595-656): ddV, ddP of every element for the Newton increment
 * du (device [ndof]), from the element data of the last integrating residual pass, which
 * must have been taken at the state the increment starts from. */
int pgf_tf_update(pgf_handle h, const double *du);
/* ThreeFieldVariables::update_increments_from_NR (include/PGFem3D_data_structure.h:197-213):
 * ddV = gamma ddV, dV += ddV (and the same for P), taken from the ddV, ddP and dV, dP that
 * the last pgf_tf_update left, so that a line-search trial is taken back exactly by
 * gamma = 0 and a second trial replaces the first (the reference scales ddV in place and
 * applies the update once). */
int pgf_tf_accept(pgf_handle h, double gamma);
/* Per-element variables, [n_elem] in mesh element order. */
#define PGF_TF_V_NP1 0
#define PGF_TF_V_N 1
#define PGF_TF_V_NM1 2
#define PGF_TF_P_NP1 3
#define PGF_TF_P_N 4
#define PGF_TF_P_NM1 5
#define PGF_TF_DV 6
#define PGF_TF_DDV 7
#define PGF_TF_DP 8
#define PGF_TF_DDP 9
int pgf_tf_get(pgf_handle h, int32_t field, double *host_out);
int pgf_tf_set(pgf_handle h, int32_t field, const double *host_in);

/* State fields, stored structure-of-arrays [component][n_qp] in QP order (element order,
 * then QP order inside the element). */
#define PGF_STATE_F_NP1 0     /* 9 comps: F at n+1 */
#define PGF_STATE_F_N 1       /* 9 comps */
#define PGF_STATE_F_NM1 2     /* 9 comps */
#define PGF_STATE_PLASTIC_NP1 3  /* J2: 6 comps Cp^-1 (xx,yy,zz,xy,yz,xz); CP: 9 comps pF */
#define PGF_STATE_PLASTIC_N 4
#define PGF_STATE_PLASTIC_NM1 5
#define PGF_STATE_SCALAR_NP1 6   /* J2: ep, gam (with damage also w, X, H, damaged flag);
                                    CP: g (slip resistance), lambda */
#define PGF_STATE_SCALAR_N 7
#define PGF_STATE_SCALAR_NM1 8
#define PGF_STATE_TAU 9          /* CP: 12 resolved shear stresses at n+1 (TENSOR_tau) */
#define PGF_STATE_TAU_N 10       /* CP: at n (TENSOR_tau_n) */
#define PGF_STATE_GDOT 11        /* CP: 12 slip rates at n+1 (TENSOR_gamma_dot) */
#define PGF_STATE_GDOT_N 12      /* CP: at n (TENSOR_gamma_dot_n) */
#define PGF_STATE_ITERS 13       /* CP: local Newton iterations of the last integration */
#define PGF_STATE_COHESIVE_NP1 14 /* interface QPs: X_eff_max, t_max at n+1, [2][n_cqp] in
                                     interface order (element order, then QP) */
#define PGF_STATE_COHESIVE_N 15   /* at n (COEL::vars, src/cohesive_element.cc:152-156) */
/* Number of components of a field (0 if the field does not exist for the mesh). */
int pgf_state_ncomp(pgf_handle h, int32_t field, int32_t *ncomp);
/* Copy a field to / from host memory laid out [comp][n_qp]. */
int pgf_get_state(pgf_handle h, int32_t field, double *host_out);
int pgf_set_state(pgf_handle h, int32_t field, const double *host_in);

/* ================================================================ linear solve */
/* Solve K x = b on the device for the tangent values vals[nnz] in the handle's pattern:
 * Jacobi-preconditioned BiCGStab from the initial x (device vectors [ndof]) to
 * ||b - K x|| <= tol ||b||.  Plays the role of the reference's per-iteration hypre solve
 * (SparseSystem::solveSystem, src/solvers/hypre/Hypre.cc; called from Newton_Raphson,
 * src/Newton_Raphson.cc:980-1012).  Returns 0 when converged, 1 when maxit was reached. */
int pgf_solve(pgf_handle h, const double *vals, const double *b, double *x, double tol, int32_t maxit,
              int32_t *iters, double *relres);

/* ================================================== learned warm start (no reference
 * counterpart: the reference has no learned component).  A small MLP (hidden widths <= 128)
 * on the matrix cores, BF16 weights and activations with FP32 accumulation, trained by the
 * library's own forward / backward / Adam kernels; it predicts the 12 slip increments of a
 * crystal-plasticity QP, from which the local Newton starts (falling back to the plain
 * start if that Newton fails).  The FP64 Newton still converges to tol_M. */
typedef struct pgf_mlp_s *pgf_mlp;
/* widths[0] inputs, widths[n_layers] outputs; GELU after every layer but the last. */
int pgf_mlp_create(int32_t n_layers, const int32_t *widths, uint64_t seed, pgf_mlp *out);
int pgf_mlp_destroy(pgf_mlp m);
/* Parameters (FP32 master copy): for each layer W [Np][Kp] row-major then b [Np], with Kp, Np
 * the widths rounded up to multiples of 16 (padding is zero). */
int pgf_mlp_param_count(pgf_mlp m, int64_t *n);
int pgf_mlp_get_params(pgf_mlp m, float *host);
int pgf_mlp_set_params(pgf_mlp m, const float *host);
int pgf_mlp_get_grads(pgf_mlp m, float *host);
/* Device pointers: X [B][widths[0]], Y [B][widths[n_layers]] (FP32). */
int pgf_mlp_forward(pgf_mlp m, const float *X, int64_t B, float *Y);
/* The same forward pass as one launch with the activations in registers and one layer's
 * weights at a time in LDS, no pre-activations kept: the kernel the learned warm start runs.
 * Device pointers as pgf_mlp_forward; at most 8 layers. */
int pgf_mlp_infer(pgf_mlp m, const float *X, int64_t B, float *Y);
/* Mean-squared-error forward + backward over a batch (B a multiple of 16); gradients are
 * added to the gradient buffer, which pgf_mlp_adam_step consumes and clears. */
int pgf_mlp_backward(pgf_mlp m, const float *X, const float *T, int64_t B, float *loss);
int pgf_mlp_adam_step(pgf_mlp m, float lr, float beta1, float beta2, float eps);

/* Crystal-plasticity features and targets of the last integration, device QP order, FP32:
 * X [n_qp][34] = (F(n) pF(n)^-1 (9), F(n+1) - F(n) (9), the 12 tau_a(n)/g(n), grain orientation
 * quaternion (w, x, y, z) with w >= 0), T [n_qp][12] = dt * gamma_dot(n+1). */
int pgf_cp_log_samples(pgf_handle h, float *X, float *T);
/* Attach a trained MLP (NULL detaches): integrating passes then build the features, run the
 * MLP and start the local Newton from its prediction.  Normalisation (host arrays): the MLP
 * sees (x - shift) * scale and predicts targets times out_scale. */
int pgf_set_surrogate(pgf_handle h, pgf_mlp m, const float *shift, const float *scale, const float *out_scale);

#ifdef __cplusplus
}
#endif
#endif /* PGFEM3D_AMD_H */
