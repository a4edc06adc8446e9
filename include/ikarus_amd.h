/* ikarus_amd.h — C ABI of the MI355X nonlinear-mechanics hot path.
 *
 * Plain C: pointers, sizes, enums. No C++, no torch types. One handle = one mesh with its elements,
 * Dirichlet flags, loads and device state; one HIP stream per handle; a handle is not thread-safe.
 * The library owns all device memory it allocates. Every function returns IK_OK (0) or a negative
 * IK_E* code; ik_last_error() gives the message.
 *
 * Each entry point names the Ikarus v0.4.9 interface it replaces (paths relative to the Ikarus tree).
 */
#ifndef IKARUS_AMD_H
#define IKARUS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ik_handle ik_handle;

/* ---- error codes ----------------------------------------------------------------------------- */
enum {
  IK_OK         = 0,
  IK_EINVAL     = -1, /* bad argument / shape (InvalidStateException, interface.hh:187-193)      */
  IK_EHIP       = -2, /* HIP runtime failure                                                     */
  IK_ENEGDET    = -3, /* det C <= 0 in an element; replaces abort() in materialhelpers.hh:120-126 */
  IK_ENOTIMPL   = -4, /* unsupported affordance / option (DUNE_THROW(NotImplemented))           */
  IK_ENOTBOUND  = -5, /* requirement not set before matrix()/vector() (interface.hh:187-193)     */
  IK_ENOCONV    = -6, /* linear solver did not reach its tolerance                               */
  IK_ESINGULAR  = -7  /* EAS: D of an element is singular to working precision (a pivot of the
                         unpivoted elimination is zero or not finite); the reference inverts D by
                         pivoted LU (enhancedassumedstrains.hh:243,294,344) and has no such error.
                         IK_PCG_TWOLEVEL_FP32_IR: its coarse operator is not positive definite */
};

/* DBCOption (ikarus/assembler/dirichletbcenforcement.hh:8) */
enum { IK_DBC_RAW = 0, IK_DBC_FULL = 1, IK_DBC_REDUCED = 2 };
/* Materials (neohooke.hh:38, svk.hh:37) */
enum { IK_NEOHOOKE = 0, IK_SVK = 1 };
/* storage / arithmetic precision of the tangent */
enum { IK_FP64 = 0, IK_FP32 = 1 };
/* where a pointer lives */
enum { IK_HOST = 0, IK_DEVICE = 1 };
/* linear solvers (si_ConjugateGradient, linearsolver.cpp:23-24, and the mixed-precision forms; the two-level
 * preconditioner has no counterpart in the reference) */
enum { IK_PCG_JACOBI_FP64 = 0, IK_PCG_BJACOBI_FP32_IR = 1, IK_PCG_TWOLEVEL_FP32_IR = 2 };
/* load-step predictors: constant (loadcontrol.inl:55-57), secant/linear extrapolation
 * (pathfollowingfunctions.hh:188-195 semantics), learned per-node MLP (no reference counterpart) */
enum { IK_PRED_CONSTANT = 0, IK_PRED_LINEAR = 1, IK_PRED_MLP = 2 };
/* stress measures of ik_stress_at (ResultTypes PK2Stress, kirchhoffStress, cauchyStress) */
enum { IK_PK2 = 0, IK_KIRCHHOFF = 1, IK_CAUCHY = 2 };

/* ---- problem description (replaces makeBasis + makeFE(skills(nonLinearElastic, eas, traction))
 *      + DirichletValues + makeSparseFlatAssembler: tests/src/testcantileverbeam.hh:102-151) ---- */
/* dim 2 is the plane-strain Q1 element (planeStrain(material), vanishingstrain.hh:93-120,148-198, 2x2 Gauss):
 * X [num_nodes*2], conn [num_elements*4], dirichlet [num_nodes*2], face_axis 0|1 names the loaded edge,
 * traction[2] is ignored, point_dof / every vector argument uses the 2D DOFs (2*node + component). The device
 * evaluates it as the mesh's one-layer hex8 extrusion of thickness sqrt(2) with u_z = 0, which has the same
 * residual norm and Newton iterates; ik_matrix_csc, ik_element, ik_energy and ik_stress_at (Voigt [11,22,12])
 * return the 2D quantities. Not for dim 2: EAS modes, ik_matrix_bsr, IK_PRED_MLP, ik_path_following
 * (IK_ENOTIMPL). */
typedef struct {
  int32_t dim;                 /* 3: trilinear hex8 with 2x2x2 Gauss (nonlinearelastic.hh:112-125); 2: see above */
  int64_t num_nodes;
  int64_t num_elements;
  const double* X;             /* [num_nodes*3] reference coordinates, host                          */
  const int32_t* conn;         /* [num_elements*8] Dune reference-cube corner order, host            */
  int32_t material;            /* IK_NEOHOOKE | IK_SVK                                                */
  double lambda, mu;           /* Lame parameters (physicshelper.hh:276-281)                          */
  int32_t eas_modes;           /* 0 (NonLinearElastic) or 21 (EAS GL E21, linearandglstrains.hh:280)  */
  const uint8_t* dirichlet;    /* [num_nodes*3] fixed-DOF flags (dirichletvalues.hh:86-131), host     */
  /* Neumann traction t = lambda * traction on element faces (traction.hh:106-138) */
  int64_t num_faces;
  const int32_t* face_elem;    /* [num_faces] element of the face                                    */
  const int8_t* face_axis;     /* [num_faces] reference axis normal to the face (0,1,2)              */
  const int8_t* face_side;     /* [num_faces] 0: xi_axis = 0, 1: xi_axis = 1                          */
  double traction[3];
  /* point loads R[dof] += coef * lambda (AssemblerManipulator callbacks,
   * assemblermanipulatorfuser.hh:243; tests/src/testcantileverbeam.hh:66-80) */
  int64_t num_point_loads;
  const int64_t* point_dof;
  const double* point_coef;
} ik_problem_desc;

/* Create / destroy. Builds the CSR/BSR pattern and element block-slot maps on the first matrix call
 * like SparseFlatAssembler::preProcessSparseMatrix (simpleassemblers.inl:290-293). */
int ik_create(ik_handle** h, const ik_problem_desc* desc);
void ik_destroy(ik_handle* h);
const char* ik_last_error(const ik_handle* h);
/* first element whose det C <= 0 in the last failing call, or -1 */
int64_t ik_bad_element(const ik_handle* h);

/* FlatAssemblerBase::size / reducedSize (interface.hh:68-74) */
int64_t ik_size(const ik_handle* h);
int64_t ik_reduced_size(const ik_handle* h);

/* FERequirement (u, lambda) binding (ferequirements.hh:223-407; interface.hh:150-155).
 * u is the full DOF vector (FlatInterleaved: dof = 3*node + component). */
int ik_set_requirement(ik_handle* h, const double* u, double lambda, int where);
int ik_get_requirement(ik_handle* h, double* u, double* lambda, int where);

/* Inhomogeneous Dirichlet values u_D(lambda) = lambda * ubar (DirichletValues::
 * evaluateInhomogeneousBoundaryCondition, dirichletvalues.hh:214-257; ubar is nonzero on constrained DOFs
 * only, else IK_EINVAL). Newton then adds step_size * K_raw(u) ubar (constrained entries zeroed) to its first
 * residual (newtonraphson.hh:215-218, utils::obtainForcesDueToIDBC, functionhelper.hh:170-184) and sets the
 * constrained DOFs to lambda * ubar after every correction (nonlinearsolverfactory.hh:45-54). NULL clears. */
int ik_set_dirichlet_values(ik_handle* h, const double* ubar, int where);

/* VectorAssembler::vector(req, forces, dbc) (interface.hh:355-365; simpleassemblers.inl:60-118).
 * R has size() entries (RAW/FULL) or reducedSize() entries (REDUCED); R = NULL assembles into the
 * handle's device residual without copying it out. */
int ik_vector(ik_handle* h, int dbc, double* R, int where);

/* ScalarAssembler::scalar(req, mechanicalPotentialEnergy) (interface.hh:283-330): the sum over elements
 * of NonLinearElastic's energy (nonlinearelastic.hh:403-408) and the traction skill's (traction.hh:71-104)
 * at the bound requirement. EAS elements have no potential (enhancedassumedstrains.hh:302-309):
 * IK_ENOTIMPL. Point loads are not part of it (as for the reference's assembler without a scalar
 * manipulator). */
int ik_energy(ik_handle* h, double* energy);

/* NonLinearElastic::calculateAt<PK2Stress | kirchhoffStress | cauchyStress>(req, xi)
 * (nonlinearelastic.hh:237-270) for every element at the bound requirement: out[e*6 + v], Voigt
 * [11,22,33,23,13,12] without shear factors, xi in the reference cube [0,1]^3. Plain elements only
 * (IK_ENOTIMPL for EAS). */
int ik_stress_at(ik_handle* h, int kind, const double xi[3], double* out, int where);

/* MatrixAssembler::matrix(req, stiffness, dbc) (interface.hh:430-440; simpleassemblers.inl:121-204).
 * Assembles into library-owned 3x3-block storage on the device (FP64, or FP32 for IK_FP32; the EAS
 * element computes in FP64 and its FP32 matrix is the FP64 one rounded entry by entry). */
int ik_matrix(ik_handle* h, int dbc, int precision);

/* Residual and tangent at the bound requirement in one call: the pair residualFunction_(x),
 * jacobianFunction_(x) that NewtonRaphson evaluates at every state (newtonraphson.hh:204-205, 242-243),
 * i.e. ik_vector(h, dbc, R, where) followed by ik_matrix(h, dbc, precision) with identical results up to
 * summation order. In FP64 one element pass computes K_e and R_e together; IK_FP32 pairs the FP32 tangent
 * with the FP64 residual (two passes). */
int ik_assemble(ik_handle* h, int dbc, int precision, double* R, int where);

/* Eigen ColMajor CSC layout of the last assembled matrix (simpleassemblers.inl:207-251):
 * pass outer/inner/values = NULL to query n and nnz. outer: n+1, inner: nnz (int32), values: nnz (FP64). */
int ik_matrix_csc(ik_handle* h, int dbc, int64_t* n, int64_t* nnz, int32_t* outer, int32_t* inner,
                  double* values);

/* Device view of the block storage: block rows = nodes, 3x3 row-major blocks (values FP64 or FP32). rowptr /
 * colidx give the block CSR pattern; block b of it (rowptr order) holds values[9*block_pos[b] .. +9): the
 * values are stored in 64-row slices (sliced ELL) with zero padding between the blocks. Any output may be NULL. */
int ik_matrix_bsr(ik_handle* h, int64_t* num_block_rows, int64_t* nnzb, const int32_t** rowptr,
                  const int32_t** colidx, const int32_t** block_pos, const void** values, int* precision);

/* calculateMatrix / calculateVector for one element at the bound requirement (mixin.hh:137-152):
 * Ke row-major 24x24 (local order 3a+k, fehelper.hh:146-152), Re 24. Host outputs, either may be NULL. */
int ik_element(ik_handle* h, int64_t e, double* Ke, double* Re);

/* EAS state update on CORRECTION_UPDATED: alpha -= D^{-1}(Rtilde + L du_e)
 * (enhancedassumedstrains.hh:225-248, 350-359). du is the full correction vector. */
int ik_update_state(ik_handle* h, const double* du, int where);
/* EAS internal variables, [num_elements * eas_modes] (enhancedassumedstrains.hh:208) */
int ik_get_alpha(ik_handle* h, double* alpha, int where);
int ik_set_alpha(ik_handle* h, const double* alpha, int where);

/* ---- VTK output (Ikarus::Vtk::Writer, io/vtkwriter.hh:37-146) ------------------------------------ */
/* DataTag::asPointData / asCellData (io/vtkwriter.hh:80-100) */
enum { IK_VTK_POINT_DATA = 0, IK_VTK_CELL_DATA = 1 };
typedef struct {
  const char* name;      /* array name (no '"', '<' or '&')                                              */
  const double* values;  /* host: point data [nodes x components] (a DOF vector has dim components),
                            cell data [elements x components]                                            */
  int32_t components;
  int32_t cell;          /* IK_VTK_POINT_DATA | IK_VTK_CELL_DATA                                          */
} ik_vtk_field;
/* Writes an ASCII XML UnstructuredGrid (.vtu) of the mesh in reference coordinates (VTK_HEXAHEDRON, or VTK_QUAD
 * for plane strain), the given fields (Writer::addInterpolation, :125-140; vectors of dim 2 are padded to 3
 * components) and, for every bit k of `results` (1 << IK_PK2 | 1 << IK_KIRCHHOFF | 1 << IK_CAUCHY), the stress of
 * ResultFunction (io/resultfunction.hh:55-110) at the bound requirement: at each element centre as cell data, or
 * at each element corner averaged over the elements sharing the node as point data (result_tag). */
int ik_write_vtu(ik_handle* h, const char* path, const ik_vtk_field* fields, int32_t num_fields, int32_t results,
                 int32_t result_tag);

/* ---- linear solve (LinearSolverTemplate::analyzePattern/factorize/solve, linearsolver.hh:50-235) -- */
typedef struct {
  int32_t kind;        /* IK_PCG_JACOBI_FP64 | IK_PCG_BJACOBI_FP32_IR | IK_PCG_TWOLEVEL_FP32_IR  */
  double rtol;         /* ||r|| <= rtol ||b||   (Eigen ConjugateGradient semantics)         */
  int32_t max_iter;    /* <= 0: 2n (Eigen default)                                         */
  double inner_rtol;   /* IR form: an FP32 sweep stops at half the relative reduction the
                          outer rtol still needs, or at inner_rtol if that is larger      */
  int32_t max_refine;  /* maximum FP64 refinement sweeps (IR form)                         */
} ik_solver_opts;

typedef struct {
  int32_t iterations;    /* total PCG iterations                      */
  int32_t refinements;   /* FP64 refinement sweeps (IR form)          */
  double rel_residual;   /* final ||b - K x|| / ||b|| in FP64         */
  int32_t converged;
} ik_solver_info;

/* Solve K x = rhs with the last assembled (FULL) matrix. With the IR forms the FP64 refinement residual b - K x
 * uses the FP32 matrix, so the solve converges to the solution of the FP32-rounded system and rel_residual /
 * converged refer to it; against the FP64 tangent the residual has a floor near eps32 ||K|| ||x||. (Newton is
 * not affected: its residual is assembled in FP64.)
 * IK_PCG_TWOLEVEL_FP32_IR preconditions the FP32 PCG with 3x3 block Jacobi plus an additive coarse correction
 * z += P (P^T K P)^-1 P^T r over box aggregates of nodes (about 512, each spanned by its six rigid-body modes).
 * The coarse inverse is built from the FP32 tangent of the first such solve on the handle and rebuilt when a
 * solve needs more than twice (+10) the inner iterations of the first solve after the last build, so later
 * solves reuse it while it still preconditions well. IK_ESINGULAR if the coarse operator is not positive
 * definite (an indefinite tangent). */
int ik_solve(ik_handle* h, const double* rhs, double* x, int where, const ik_solver_opts* opts,
             ik_solver_info* info);

/* ---- Newton-Raphson (newtonraphson.hh:196-257) and LoadControl (loadcontrol.inl:21-57) -------- */
typedef struct {
  double tol;            /* absolute ||R||_2 tolerance (NRSettings, newtonraphson.hh:27-32) */
  int32_t max_iter;
  int32_t min_iter;
  int32_t precision;     /* tangent precision IK_FP64 | IK_FP32                             */
  ik_solver_opts linear;
  double forcing;        /* > 0: inexact Newton, Eisenstat-Walker choice 2 with gamma =
                            forcing, eta <= eta_max, floor max(linear.rtol, tol/(2||R_k||));
                            0: every linear solve to linear.rtol                        */
  double eta_max;        /* upper bound of the forcing term (0: 0.1)                    */
} ik_nr_settings;

typedef struct {
  int32_t success;
  int32_t iterations;
  double residual_norm;
  double correction_norm;
  int64_t linear_iterations;
} ik_nr_info;

/* NewtonRaphson::solve(x, stepSize) on the bound requirement; state stays on the device. */
int ik_newton_solve(ik_handle* h, const ik_nr_settings* s, double step_size, ik_nr_info* info);

typedef struct {
  int32_t load_steps;
  double tbegin, tend;
  int32_t predictor;      /* IK_PRED_*                                                      */
} ik_lc_config;

typedef struct {
  int32_t success;
  int32_t total_iterations;  /* Newton iterations, including those of failed learned-predictor attempts */
  int64_t linear_iterations;
  int32_t solves;         /* number of Newton solves run (initial + load steps)              */
  int32_t fallbacks;      /* IK_PRED_MLP: load steps whose Newton solve failed from the MLP prediction
                             and was solved again from the secant prediction                   */
} ik_control_info;

/* LoadControl::run. iterations/residuals: optional arrays of load_steps+1 entries (the solve that ended each
 * step). The run stops at the first failed solve (loadcontrol.inl:46-47). With IK_PRED_MLP, a step whose solve
 * from the MLP prediction fails (IK_ENEGDET, IK_ESINGULAR, a residual norm that is not finite, or max_iter) is
 * restored to u_n, alpha_n, lambda_n and solved from the secant prediction instead: plain Newton from the
 * secant start stays the fallback. A Newton solve whose residual norm is not finite counts as failed (the
 * reference's loop, newtonraphson.hh:219-252, ends on it and reports success). */
int ik_load_control(ik_handle* h, const ik_lc_config* c, const ik_nr_settings* s, ik_control_info* info,
                    int32_t* iterations, double* residuals);

typedef struct {
  int32_t steps;     /* PathFollowing steps (the start-state solve comes on top)        */
  double step_size;  /* arc length s of ArcLength                                       */
  int32_t target_iterations;  /* AdaptiveStepSizing::IterationBased (adaptivestepsizing.hh:48-90): from the
                                 second step on, s <- sqrt(target / its) s when the last solve took
                                 its > target iterations; 0: NoOp                                       */
} ik_pf_config;

/* PathFollowing::run with ArcLength (pathfollowing.inl:28-77; pathfollowingfunctions.hh:88-190), each
 * solve NewtonRaphsonWithSubsidiaryFunction::solve (newtonraphsonwithscalarsubsidiaryfunction.hh:190-307):
 * Fext0 = -R(0, 1) at the start of each solve (u = 0, the rest of the state kept: an EAS element's alpha is not
 * reset, as in the reference), then per iteration K dR = -R and K dL = Fext0 - F_idbc with the FULL tangent and
 * s->linear, the load increment from the linearized arc-length equation, the EAS state update on the correction
 * and the inhomogeneous Dirichlet DOFs set to lambda ubar. With inhomogeneous Dirichlet values the arc length
 * includes their increment Dl ubar (idbcIncrement, controlroutines/common.hh:63-76) and F_idbc = K_raw(u) ubar
 * (constrained entries zeroed). The requirement must start at lambda = 0 in equilibrium (the start solve has no
 * load direction yet, IK_EINVAL otherwise). iterations/residuals/lambdas: optional arrays of steps+1 entries (per
 * solve). s->forcing is not used. Not for plane-strain handles (IK_ENOTIMPL). */
int ik_path_following(ik_handle* h, const ik_pf_config* c, const ik_nr_settings* s, ik_control_info* info,
                      int32_t* iterations, double* residuals, double* lambdas);

/* ---- learned per-node predictor (no reference counterpart) ------------------------------------ */
/* MLP 14 -> 128 -> 128 -> 128 -> 3, SiLU, BF16 operands on MFMA with FP32 accumulation. Parameters are
 * FP32 master copies laid out W1[128x14], b1[128], W2[128x128], b2, W3[128x128], b3, W4[3x128], b4[3]
 * (row-major, W maps inputs to outputs). Features per node for the load step n -> n+1:
 *   [X/L (3), t/40 (3), lambda_{n+1}, 40 dlambda, v_n/0.1 (3), v_{n-1}/0.1 (3)], v = du/dlambda;
 * output y: v_{n+1} = v_n + 0.1 y and u_pred = u_n + dlambda v_{n+1} on free DOFs. */
int64_t ik_mlp_num_params(void);
/* parameters used by IK_PRED_MLP inside ik_load_control */
int ik_mlp_set_params(ik_handle* h, const float* params, int where);
int ik_mlp_get_params(ik_handle* h, float* params, int where);
/* traction per unit load factor and length scale L used by the features (defaults: the problem's
 * traction, L = 1) */
int ik_mlp_set_load(ik_handle* h, const double traction[3], double char_length);
/* The prediction IK_PRED_MLP makes at the start of a load step, on its own: u <- u + dlam (v_n + 0.1 y) on the
 * free DOFs of the bound requirement, with v_n = du_n / dlam_prev (0 if dlam_prev <= 0), v_{n-1} = du_n1 /
 * dlam_prev2 (0 if dlam_prev2 <= 0) and y the network's output for lambda_{n+1} = lam_next. du_n, du_n1: the
 * last two increments of u (ndof values each, host or device). One fused kernel; lambda is not changed.
 * IK_EINVAL without parameters, IK_ENOTBOUND without requirement, IK_ENOTIMPL on plane-strain handles. */
int ik_mlp_predict(ik_handle* h, const double* du_n, const double* du_n1, double lam_next, double dlam,
                   double dlam_prev, double dlam_prev2, int where);

/* Training context: forward, mean-squared-error gradient and Adam steps with the same kernels. */
typedef struct ik_mlp ik_mlp;
int ik_mlp_create(ik_mlp** m, uint64_t seed);  /* He-uniform init from SplitMix64(seed) */
void ik_mlp_destroy(ik_mlp* m);
const char* ik_mlp_last_error(const ik_mlp* m);
int ik_mlp_params(ik_mlp* m, float* params, int where);
int ik_mlp_load(ik_mlp* m, const float* params, int where);
/* x: [n x 14] features, y: [n x 3] outputs (row-major FP32) */
int ik_mlp_forward(ik_mlp* m, const float* x, int64_t n, float* y, int where);
/* Allocation-free forward, for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc); it has no counterpart in the reference. x [n x 14] and
 * y [n x 3] are device pointers, 1 <= n <= IK_MLP_NOALLOC_MAX_ROWS (else IK_EINVAL). The hidden activations live in
 * a fixed 8 MiB scratch array per device (2 x IK_MLP_NOALLOC_MAX_ROWS x 128 FP32) that every context on that
 * device shares: calls that overlap in time on one device race. Results equal ik_mlp_forward's bit for bit. */
#define IK_MLP_NOALLOC_MAX_ROWS 8192
int ik_mlp_forward_noalloc(ik_mlp* m, const float* x, int64_t n, float* y);
/* loss = mean((y - t)^2) over n x 3 entries; grad (nullable): d loss / d params */
int ik_mlp_loss_grad(ik_mlp* m, const float* x, const float* t, int64_t n, double* loss, float* grad, int where);
/* loss_grad followed by one Adam step (beta 0.9 / 0.999, eps 1e-8) with learning rate lr */
int ik_mlp_train_step(ik_mlp* m, const float* x, const float* t, int64_t n, float lr, double* loss, int where);

/* ---- timing of the hot kernels on the handle's stream (HIP events) ----------------------------- */
enum { IK_KERNEL_ASSEMBLE_KR = 0, IK_KERNEL_ASSEMBLE_R = 1, IK_KERNEL_SPMV = 2, IK_KERNEL_MLP = 3,
       IK_KERNEL_GATHER_K = 4, IK_KERNEL_COUNT = 5 };
int ik_timing_enable(ik_handle* h, int on);
int ik_timing_read(ik_handle* h, int kernel, int64_t* launches, double* total_ms);
int ik_timing_reset(ik_handle* h);
int ik_synchronize(ik_handle* h);
/* Tangent staging layout (no counterpart in the reference; for bandwidth accounting): staged slots (3x3
 * blocks written by the element kernel and read by the gather), upper blocks the gather assembles, and
 * upper blocks the element kernel writes straight into K. Builds the pattern if needed. */
int ik_assembly_layout(ik_handle* h, int64_t* staged_slots, int64_t* gathered_blocks, int64_t* direct_blocks);

#ifdef __cplusplus
}
#endif
#endif /* IKARUS_AMD_H */
