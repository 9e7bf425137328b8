/* astro_abi.h -- C ABI of the MI355X (gfx950) AStrO hot-path library
 * (misc_amd/libastro_hip.so).
 *
 * Two layers:
 *
 * 1. Routine-for-routine replacements of the reference's Fortran solvers and
 *    book-keeping, with the reference's argument lists and index conventions
 *    (1-based values, (0:n) prefix ranges as n+1 ints, column value 0 = unused
 *    slot, column-major arrays).  All pointers are HOST pointers; each call is
 *    synchronous.  A Fortran host binds them with ISO_C_BINDING (scalars by
 *    VALUE, arrays by reference).  Every function returns an astro_status.
 *
 * 2. A device-resident model context (astro_ctx_*) that keeps mesh, matrices,
 *    profile factor and Krylov workspace in HBM between calls, for the
 *    assemble -> profile -> factor -> GMRES sequence of one solve, without
 *    host round trips.
 *
 * No torch types cross this boundary; the library allocates device memory
 * with hipMalloc and owns it until the matching free/destroy call.
 */
#ifndef ASTRO_ABI_H
#define ASTRO_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ASTRO_OK = 0,
    ASTRO_ERR_ARG = 1,          /* invalid argument or size */
    ASTRO_ERR_DEVICE = 2,       /* HIP runtime error */
    ASTRO_ERR_ZERO_PIVOT = 3,   /* D(i) == 0 in the LDL^T factor */
    ASTRO_ERR_NO_CONVERGE = 4,  /* GMRES hit its cycle limit */
    ASTRO_ERR_TIMEOUT = 5,      /* a device-side bounded wait expired */
    ASTRO_ERR_UNSUPPORTED = 6   /* element type / option not implemented */
} astro_status;

/* Last error message of the calling thread ("" if none). */
const char* astro_last_error(void);
/* Library version string. */
const char* astro_version(void);

/* ---- Layer 1: reference routine replacements --------------------------- */

/* In-place row-skyline LDL^T.  Replaces getSparseLDLFact
 * (Working/AStrO_solvers.f90:10-78): ALTri(ASize) holds rows i = 1..n as
 * columns i-len+1..i with the diagonal last, ARange(0:n) the prefix ends.  On
 * return the strict lower part holds L (unit diagonal implicit) and the
 * diagonal holds D. */
int astro_getSparseLDLFact(double* ALTri, int64_t ASize, const int32_t* ARange, int32_t n);

/* Solve A x = b with the factor of astro_getSparseLDLFact.  Replaces
 * solveSparseLDLFact (Working/AStrO_solvers.f90:83-126); zVec (Ldim) receives
 * the forward-substitution intermediate z = L^-1 b as in the reference. */
int astro_solveSparseLDLFact(double* soln, const double* rhs, const double* Lmat, int64_t LSize,
                             const int32_t* LRange, int32_t Ldim, double* zVec);

/* Left-preconditioned GMRES(numVecs) on (A + C^T C) x = b with the factored
 * profile P as preconditioner.  Replaces gMRes (Working/AStrO_solvers.f90:128-297);
 * xVec is the initial guess on entry.  The stopping rule is |P^-1 (b - A x)| <=
 * max(rtol * |P^-1 (b - A x0)|, atol); the reference's rule is atol = 1e-12,
 * rtol = 0 (pass rtol <= 0 to keep it).  At most maxIt/numVecs + 1 restart
 * cycles, as in the reference. */
int astro_gMRes(double* xVec, const double* Amat, const int32_t* Acols, const int32_t* ARange, int64_t ASize,
                int32_t ADim, const double* Cmat, const int32_t* Ccols, const int32_t* CRange, int64_t CSize,
                int32_t CDim, const double* Pmat, const int32_t* PRange, int64_t PSize, const double* bVec,
                int32_t numVecs, int32_t maxIt, double rtol, double atol);

/* Profile from CSR plus the penalty C^T C.  Replaces convertToLTri
 * (Working/AStrO_bookKeeping.f90:754-792); lTri is overwritten. */
int astro_convertToLTri(double* lTri, const int32_t* lTRange, int64_t lTSize, const double* mainMat,
                        const int32_t* mainCols, const int32_t* mainRange, int64_t mainSize, int32_t mainDim,
                        const double* cMat, const int32_t* cCols, const int32_t* cRange, int64_t cSize,
                        int32_t cDim);

/* ---- Host symbolic analysis (no device work) ----------------------------- */

typedef struct astro_prep astro_prep;

/* ---- Layer 2: device-resident model context ----------------------------- */

typedef struct astro_ctx astro_ctx;

/* Model description (host pointers, copied at create).  Mirrors the model
 * arrays of Working/AStrO_globalData.f90:5-58 after readModelInput:
 *   nodeList(3,numNodes), elementList(8,numEls) 1-based with 0 padding,
 *   elementType(numEls) (8 = brick8), per-element engineering constants
 *   eprops(9,numEls) = (E1,E2,E3,nu12,nu13,nu23,G12,G13,G23), density(numEls),
 *   conductivity cond(6,numEls) = (k11,k22,k33,k12,k13,k23) (may be NULL),
 *   constraint terms (eqn, dof 1..7, coef, node-or-set) with the node lists of
 *   set terms, and nodal loads (node id, 7 components). */
typedef struct {
    int32_t numNodes, numEls;
    const double* nodeList;
    const int32_t* elementList;
    const int32_t* elementType;
    const double* eprops;
    const double* density;
    const double* cond;
    int32_t numMpcTerms;
    const int32_t* mpcEqn;
    const int32_t* mpcDof;
    const double* mpcCoef;
    const int32_t* mpcIsSet;
    const int32_t* mpcNodeStart; /* numMpcTerms + 1 */
    const int32_t* mpcNodes;
    int32_t numMPC;
    const double* mpcRHS;
    int32_t numLoads;
    const int32_t* loadNode;
    const double* loadVal; /* 7 per load */
    /* Shell sections (shell4 elements; may be null for solid models), the
     * section/layup data of Working/AStrO_input.f90's sections block resolved
     * per element: numLayers layers of (E1, E2, nu12, G12, thickness, angle in
     * radians) per element, the section z offset per element, and the two
     * orientation vectors (x axis, a vector in the x-y plane) per element. */
    int32_t numLayers;
    const double* layers;     /* numEls * numLayers * 6 */
    const double* zOffset;    /* numEls */
    const double* orient;     /* numEls * 6 */
    /* Specific heat per element (may be null: no capacitance); the transient
     * thermal path uses density * specHeat. */
    const double* specHeat;   /* numEls */
    /* Thermal expansion per element (may be null: none), the material's
     * "expansion" (Working/AStrO_input.f90:751-760, r_getMaterialThermExp,
     * Working/AStrO_r_designPropertyFunctions.f90:105-127) in the strain's
     * Voigt order 11, 22, 33, 12, 13, 23 with engineering shear.  brick8:
     * the element stress is C (strain - T alpha) with T the nodal
     * temperature interpolated to the point (nodeTemp, e.g. from a thermal
     * solve on the same context). */
    const double* thermExp;   /* numEls * 6 */
} astro_model;

/* Options of the analysis (solver parameters of the *solve command). */
typedef struct {
    int32_t maxBW;            /* <= 0: reference rule from numNodes */
    int32_t literalMaxBW;     /* 1: 27*N*N in 32-bit wraparound as the reference; 0: 64-bit */
    int32_t thermLiteralD8;   /* 1: thermal scatter never fills a row's last slot, as the reference */
    int32_t numVecs;          /* GMRES restart length; <= 0: reference rule (n/2 below 200, else 100) */
    double rtol;              /* GMRES relative tolerance on |P^-1 r| (default 1e-12 when <= 0) */
    int32_t maxCycles;        /* GMRES restart cycles cap; <= 0: 1000 */
    int32_t coupledProfile;   /* 0: profile rows of one component, as the reference (the factor is a
                                 component-block preconditioner); 1: rows reach every component of the
                                 node and its neighbours (the factor solves the system directly) */
    int32_t precision;        /* 0: FP64 profile and factor, the reference's arithmetic; 1: mixed: the
                                 profile and L stored in FP32 and the factor's trailing updates on FP32
                                 matrix cores with FP32 accumulation, while the pivots D, the diagonal
                                 blocks' factor, the sweeps' arithmetic, the operator and the GMRES
                                 vectors stay FP64 and every restart forms the true residual in
                                 double-double (the restarts are iterative refinement) */
    int32_t stepGraphs;       /* 1: dynamic_run records the Newmark step (residual, loads, one GMRES
                                 cycle with its Givens work on the device, judge, update) into HIP
                                 graphs and replays one per step; 0: eager launches per step */
} astro_options;

/* Host-only analysisPrep (Working/AStrO_bookKeeping.f90:794-1039): level
 * ordering, DOF numbering, CSR ranges and first-touch columns, profile ranges,
 * constraint matrices.  Needs no GPU.  Arrays are read with
 * astro_prep_get_i32 / astro_prep_get_f64 under the names of
 * astro_ctx_get_i32 plus the double arrays "elMPCMat", "elMPCRHS". */
int astro_prep_create(const astro_model* model, const astro_options* opts, astro_prep** out);
int astro_prep_destroy(astro_prep* prep);
int64_t astro_prep_scalar(const astro_prep* prep, const char* name);
int astro_prep_get_i32(const astro_prep* prep, const char* name, int32_t* host, int64_t count);
int astro_prep_get_f64(const astro_prep* prep, const char* name, double* host, int64_t count);

/* Create a context on the current HIP device: symbolic analysis on the host
 * (analysisPrep, Working/AStrO_bookKeeping.f90:794-1039) and upload. */
int astro_ctx_create(const astro_model* model, const astro_options* opts, astro_ctx** out);
int astro_ctx_destroy(astro_ctx* ctx);

/* Scalars: "elMatDim", "elMatSize", "elMatLTSize", "numSolidNds", "solverMaxBW",
 * "elMPCDim", "thermMatSize", "thermMatLTSize", "gmresSteps", "gmresCycles",
 * "numColors", "nodeBlockNeighbours" (neighbour slots per node of the
 * node-block form of the elastic operator that GMRES applies; 0 when it
 * applies the CSR).  Returns -1 for an unknown name. */
int64_t astro_ctx_scalar(const astro_ctx* ctx, const char* name);
double astro_ctx_scalar_f(const astro_ctx* ctx, const char* name);

/* Copy a context array to host memory: int arrays "currentRank", "nDofIndex",
 * "elMatRange", "elMatCols", "elMatLTRange", "elMPCMatRange", "elMPCMatCols",
 * "thermMatRange", "thermMatCols", "thermMatLTRange"; double arrays
 * "elasticMat", "elMatLT", "elMPCMat", "elMPCRHS", "elasticLoad", "nodeDisp",
 * "thermMat", "thermMatLT", "thermalLoad", "nodeTemp".  count is the element
 * count of the host buffer; it must equal the array length.  "elMatLT" and
 * "thermMatLT" (profile, or factor after astro_ctx_factor) cross in the
 * reference's packed row-skyline layout of elMatLTRange / thermMatLTRange;
 * on the device their rows are padded to 128-byte lines, and the copy
 * repacks them on the host. */
int astro_ctx_get_i32(astro_ctx* ctx, const char* name, int32_t* host, int64_t count);
int astro_ctx_get_f64(astro_ctx* ctx, const char* name, double* host, int64_t count);
/* Overwrite a double array of the context from host memory (profiles in the
 * packed layout, as above). */
int astro_ctx_set_f64(astro_ctx* ctx, const char* name, const double* host, int64_t count);

/* Elastic hot path, all on the device, asynchronous on the context stream
 * unless stated: */
/* getElasticSolnLoad(1) for brick8 at u = 0 (Working/AStrO_commandFunctions.f90:260-351):
 * element tangents scattered into the CSR values.  massFactor adds
 * massFactor * M_e (Newmark: 1/(beta dt^2)). */
int astro_ctx_assemble_elastic(astro_ctx* ctx, double massFactor);
/* scaleElasticMPC (Working/AStrO_commandFunctions.f90:537-566). */
int astro_ctx_scale_mpc(astro_ctx* ctx);
/* convertToLTri for the elastic profile. */
int astro_ctx_build_profile(astro_ctx* ctx);
/* getSparseLDLFact on the elastic profile (in place).  Synchronous; returns
 * ASTRO_ERR_ZERO_PIVOT if a pivot is zero. */
int astro_ctx_factor(astro_ctx* ctx);
/* stepSolve for a linear static elastic step: RHS (nodal + constraint loads,
 * Working/AStrO_commandFunctions.f90:353-535), GMRES from delDisp, nodeDisp +=
 * delDisp.  Synchronous. */
int astro_ctx_solve(astro_ctx* ctx);
/* Zero delDisp and nodeDisp, and once the thermal system exists its twins
 * (the state a fresh job starts from, Working/AStrO_bookKeeping.f90:960,
 * Working/AStrO_commandFunctions.f90:653-679). */
int astro_ctx_reset(astro_ctx* ctx);
/* Whole solve() sequence: assemble, scale, profile, factor, solve.  The step
 * right-hand side holds -R(u) of the current nodeDisp (see below), so a second
 * call from a converged state returns the same displacements. */
int astro_ctx_static_solve(astro_ctx* ctx);
/* getElasticSolnLoad(0) on a zeroed load (Working/AStrO_commandFunctions.f90:
 * 260-351, as the reference's reactionForce output calls it,
 * Working/AStrO_runJob.f90:228-233): "elasticLoad" = -R(u) for the current
 * nodeDisp, with R(u) = K u, K the assembled tangent at rest (linear static
 * model).  Reactions are -elasticLoad at the constrained nodes
 * (Working/AStrO_output.f90:122-141).  Synchronous. */
int astro_ctx_elastic_soln_load(astro_ctx* ctx);
/* Element stress and strain recovery (r_getElStress, solid branch,
 * Working/AStrO_r_elementEqns.f90:924-976, 1700-1713) at the natural point
 * sp[3] (each in [-1, 1]) of every element for the current nodeDisp:
 * Green-Lagrange strain of the element's own nodal displacements and
 * stress = C strain (no thermal strain), Voigt (11, 22, 33, 12, 13, 23) with
 * engineering shear; stress, strain: 6 * numEls doubles, element order.
 * brick8 models (ASTRO_ERR_UNSUPPORTED for shell4).  Synchronous. */
int astro_ctx_element_stress(astro_ctx* ctx, const double* sp, double* stress, double* strain);
/* Heat flux (-k grad T) and temperature gradient of every element at the
 * natural point sp[3] for the current nodeTemp (r_getElFlux,
 * Working/AStrO_r_elementEqns.f90:1746-1799); flux, grad: 3 * numEls doubles.
 * brick8 models with a thermal system.  Synchronous. */
int astro_ctx_element_flux(astro_ctx* ctx, const double* sp, double* flux, double* grad);
/* (dR/dT)^T lam of a brick8 model with thermal expansion: lam elMatDim
 * doubles, out numNodes doubles indexed like nodeTemp (by rank).  The
 * thermal adjoint's coupling term, dLdt -= dRdT^T dispAdj, of
 * solveDynamicAdjoint (Working/AStrO_commandFunctions.f90:960-999), with the
 * element's intended dRdT = -sum_g detJ B0^T C alpha N^T (r_getElRuk's loop
 * writes a stale row, D10).  Synchronous. */
int astro_ctx_dRdT_product(astro_ctx* ctx, const double* lam, double* out);
/* Co-rotational internal force of a shell4 model (the shell branch of
 * r_getElRuk with nonLinearGeom, Working/AStrO_r_elementEqns.f90:2245-2423,
 * kinematics of r_getInstOrient / r_getInstDof / r_getShellDef, :681-891):
 * the element frame follows the mean nodal rotation by a finite rotation,
 * R = dE/du by complex step on the element's 32 DOFs.  u: elMatDim doubles
 * (null: the current nodeDisp); iters (0..64) chord steps on the element's
 * 8 internal DOFs, which start from the linear recovery; force: elMatDim
 * doubles, the condensed element forces at the nodal DOFs; elEnergy:
 * numEls doubles (or null).  No counterpart to the element's own internal
 * DOF recovery in the reference (it carries them at zero).  Synchronous. */
int astro_ctx_shell_internal_force(astro_ctx* ctx, const double* u, int32_t iters, double* force,
                                   double* elEnergy);
/* Tangent at the displacement u (elMatDim doubles; null: the current
 * nodeDisp) into the CSR values elasticMat: the assembly of
 * getElasticSolnLoad(1) with nonLinearGeom (Working/AStrO_commandFunctions.f90:
 * 260-351, element dRdU of r_getElRuk, Working/AStrO_r_elementEqns.f90:
 * 2206-2533).  Each element matrix is the derivative of the element's
 * internal force at u: shell4, astro_ctx_shell_internal_force's condensed
 * force (central differences of the complex-step gradient, internal DOFs at
 * element equilibrium after intIters (0..64) chord steps); brick8, the
 * Green-Lagrange force (material and geometric parts, exact).  Synchronous. */
int astro_ctx_nonlinear_tangent(astro_ctx* ctx, const double* u, int32_t intIters);
/* Nonlinear static solve, stepSolve's elastic loop with nonLinearGeom
 * (Working/AStrO_commandFunctions.f90:704-747): Newton steps
 * (K(u) + C^T C) du = F - R(u) + C^T (rhs - C u), R the internal force at u
 * (co-rotational shell4 with intIters internal-DOF steps, Green-Lagrange
 * brick8) and K(u) the tangent above, assembled and factored at every step,
 * until max|du| <= tol max|u| or maxIt steps (ASTRO_ERR_NO_CONVERGE if du
 * stops being finite).  brick8 models with thermal expansion take the
 * thermal strain at the current nodeTemp in the force and the tangent.  Starts
 * from the current nodeDisp; *iters (may be null): the steps taken.
 * Synchronous. */
int astro_ctx_nonlinear_solve(astro_ctx* ctx, int32_t maxIt, double tol, int32_t intIters, int32_t* iters);

/* Newmark dynamics, the dynamic branch of solve()
 * (Working/AStrO_commandFunctions.f90:762-810; element residual
 * r_getElRu/r_getElRuk/r_getElRum, Working/AStrO_r_elementEqns.f90:2659-2696,
 * 2424-2458, 2628-2653), with each element gathering its own nodes' solution
 * (the reference's r_getElementSolution indexes by local node number).
 * dynamic_init: tangent K + M / (beta dt^2) at rest, penalty scaling, profile,
 * factor; state (u, v, a and their previous values) zeroed.
 * dynamic_step: one step with the nodal loads times loadScale; GMRES from the
 * previous increment to the context's rtol.
 * dynamic_run: nSteps steps with loadScale[k]; scalar "dynMs" = its time.
 * Arrays "nodeVel", "nodeAcc", "prevDisp" are readable with astro_ctx_get_f64. */
int astro_ctx_dynamic_init(astro_ctx* ctx, double beta, double gamma, double dt);
int astro_ctx_dynamic_step(astro_ctx* ctx, double loadScale);
int astro_ctx_dynamic_run(astro_ctx* ctx, int nSteps, const double* loadScale);

/* Design sensitivities by complex step (shell4 layer thickness), replacing
 * getdRelasticdD's per-variable loop (Working/AStrO_commandFunctions.f90:
 * 1172-1375, element-id and DOF indexing corrected) and, for compliance,
 * getTotaldLdD (:1377-1466).  set_design: element e's layer dvLayer (0-based)
 * thickness changes by elCoef[e] * d[elDV[e]] (elDV < 0: none; elCoef null:
 * 1); h = complex step.  dRdD: elMatDim x nDV column-major at the current
 * displacement.  compliance_gradient: dC/dd (nDV) of C = F . u by the adjoint
 * lambda = u.  All design variables are evaluated in one launch per element
 * colour. */
int astro_ctx_set_design(astro_ctx* ctx, int nDV, const int32_t* elDV, const double* elCoef, int dvLayer, double h);
int astro_ctx_dRdD(astro_ctx* ctx, double* dRdD);
int astro_ctx_compliance_gradient(astro_ctx* ctx, double* grad);
/* The same gradient by direct complex-step solves, all design variables in
 * one persistent launch (replaces the per-variable loop of getdRelasticdD +
 * getTotaldLdD, Working/AStrO_commandFunctions.f90:1172-1466): variable k's
 * system (K + C^T C)(d + i h e_k) u_k = F is assembled, factored as a
 * complex-symmetric LDL^T and solved in complex128, grad[k] =
 * Im(F . u_k) / h and compliance[k] (nDV, may be null) = Re(F . u_k).  The
 * instances use a node-major DOF order of their own and start their complex
 * arithmetic at the first row their variable touches.  Needs
 * astro_ctx_set_design; assembles the current design itself.  Memory:
 * 16 bytes x profile entries x nDV.  Scalar "csBatchMs": the launch's time. */
int astro_ctx_cs_batch_gradient(astro_ctx* ctx, double* grad, double* compliance);
/* The batched kernel's autotuner: every (row block, column chunk) tile that
 * fits in LDS is run twice on the context's batch and timed on the second
 * run; the fastest is kept for later astro_ctx_cs_batch_gradient calls.
 * tiles[t] = RB * 100 + CB and ms[t] (ms < 0: does not fit) for t < count. */
int astro_ctx_cs_batch_autotune(astro_ctx* ctx, int32_t* tiles, float* ms, int32_t count);
/* Selects the tile RB * 100 + CB for later calls (ASTRO_ERR_ARG if it is not
 * one of the kernel's tiles). */
int astro_ctx_cs_batch_tile(astro_ctx* ctx, int32_t tile);

/* ---- Design variables, objectives and the adjoint gradient (brick8) ------
 * Design variables as the reference's readDesignVarInput leaves them
 * (Working/AStrO_input.f90:1099-1408, Working/AStrO_globalData.f90): per
 * variable a category and 1-based component, the element map elToDRange
 * (numEls + 1 prefix offsets) / elToD (1-based variable ids) / elToCoef and
 * the node map ndToDRange (numNodes + 1) / ndToD / ndToCoef.  A property is
 * its model value plus sum(coef * d) over the variables that list the element
 * or node (r_getMaterialStiffness, r_getMaterialThermCond, r_getMaterialDensity,
 * r_getNodeCoord, r_getNodeDLoad, Working/AStrO_r_designPropertyFunctions.f90:
 * 8-80, 129-160, 82-102, 945-991): the model passed to astro_ctx_create holds
 * the properties at the current design, and derivatives are taken there by a
 * complex step of size h on every property the variable moves.
 * Categories: ASTRO_DV_MODULUS / _POISSON / _SHEAR ('elastic' modulus,
 * poissonRatio, shearModulus; component 1..3 = E1..E3, nu12 nu13 nu23,
 * G12 G13 G23), _DENSITY, _THERMAL_COND (component 1..6 = k11 k22 k33 k12 k13
 * k23), _NODE_COORD (component 1..3), _NODE_LOAD ('nodeLoad' elastic,
 * component 1..6; R = K u - F, so dR/dd = -coef at the node's DOF). */
enum {
    ASTRO_DV_MODULUS = 1, ASTRO_DV_POISSON = 2, ASTRO_DV_SHEAR = 3, ASTRO_DV_DENSITY = 4,
    ASTRO_DV_THERMAL_COND = 5, ASTRO_DV_NODE_COORD = 6, ASTRO_DV_NODE_LOAD = 7
};
typedef struct {
    int32_t numDV;
    const int32_t* category;   /* numDV */
    const int32_t* component;  /* numDV, 1-based */
    const int32_t* elToDRange; /* numEls + 1, may be null (no element map) */
    const int32_t* elToD;
    const double* elToCoef;
    const int32_t* ndToDRange; /* numNodes + 1, may be null (no node map) */
    const int32_t* ndToD;
    const double* ndToCoef;
    double h;                  /* complex step (the reference's compStep is 1e-14i) */
} astro_design;
/* Sets (replaces) the context's design variables.  brick8 models. */
int astro_ctx_set_design_vars(astro_ctx* ctx, const astro_design* design);
/* dR/dd of the linear static residual R = K(d) u - F(d) at the current
 * nodeDisp (getdRelasticdD, Working/AStrO_commandFunctions.f90:1172-1375,
 * with each element scattering into its own DOFs, D9 corrected): dRdD is
 * elMatDim x numDV column-major.  Element terms by complex step on the
 * device, one thread per (element, variable) pair.  Synchronous. */
int astro_ctx_dRdD_vars(astro_ctx* ctx, double* dRdD);
/* dRt/dd of the thermal residual Rt = K_t(d) T at the current nodeTemp
 * (getdRthermaldD, Working/AStrO_commandFunctions.f90:1016-1170):
 * numNodes x numDV column-major (node rank order).  Needs the thermal system
 * (astro_ctx_assemble_thermal).  Synchronous. */
int astro_ctx_dRtdD_vars(astro_ctx* ctx, double* dRtdD);

/* Objective terms (Working/AStrO_objective.f90:9-470; input
 * Working/AStrO_input.f90:1410-1576): term value = coef * sum over the set of
 * (q - tgt)^expo ('powerNorm'), coef * (sum q - tgt)^expo ('volumeIntegral')
 * or coef * (sum q v / sum v - tgt)^expo ('volumeAverage'), v the element
 * volume (2x2x2 Gauss).  q per category: displacement (component of nodeDisp
 * at the set's nodes), temperature (nodeTemp), stress / strain (component of
 * the Green-Lagrange strain, or C times it, at the element centre:
 * r_getSolidStress, Working/AStrO_r_elementEqns.f90:924-976), strain energy
 * (sum over the Gauss points of stress . strain / 2 times the weight;
 * 'powerNorm' uses it per unit volume), mass (rho v), volume.  items: node ids
 * (displacement, temperature) or element ids, 1-based, no repeats; tgt: numTgt
 * = 0 ('none'), 1 ('single') or numItems ('vector'; the volume operators read
 * its first entry). */
enum {
    ASTRO_OBJ_DISPLACEMENT = 1, ASTRO_OBJ_TEMPERATURE = 2, ASTRO_OBJ_STRESS = 3, ASTRO_OBJ_STRAIN = 4,
    ASTRO_OBJ_STRAIN_ENERGY = 5, ASTRO_OBJ_MASS = 6, ASTRO_OBJ_VOLUME = 7
};
enum { ASTRO_OP_POWER_NORM = 1, ASTRO_OP_VOLUME_INTEGRAL = 2, ASTRO_OP_VOLUME_AVERAGE = 3 };
typedef struct {
    int32_t category, op, component;
    double coef, expo;
    int32_t numItems;
    const int32_t* items;
    int32_t numTgt;
    const double* tgt;
} astro_objective_term;
/* getObjective: objVal[t] per term at the current nodeDisp / nodeTemp.
 * Element terms on the device, one thread per element.  Synchronous. */
int astro_ctx_objective(astro_ctx* ctx, int32_t numTerms, const astro_objective_term* terms, double* objVal);
/* getTotaldLdD for a static step (Working/AStrO_commandFunctions.f90:
 * 1377-1412): dL/dd = dObj/dd - lambda . dR/dd - lambda_t . dRt/dd with the
 * adjoints (K + C^T C) lambda = dObj/du and the thermal twin
 * (solveDynamicAdjoint, :917-1014; GMRES on the factored system of the last
 * static / thermal solve).  dObj/du (getdObjdU) and the explicit dObj/dd
 * (getdObjdD, complex step) on the device.  dLdD: numDV.  Synchronous. */
int astro_ctx_objective_gradient(astro_ctx* ctx, int32_t numTerms, const astro_objective_term* terms, double* dLdD);
/* Gradient of a Newmark objective J = sum over steps i = 1..nSteps of the
 * terms active at t_i = i dt (termWindow: 2 doubles per term, the activeTime
 * window; null: every step), the dynamic branch of getTotaldLdD with
 * augmentdLdu and solveDynamicAdjoint (Working/AStrO_commandFunctions.f90:
 * 819-956, 1416-1464).  U, A: the forward's displacements and accelerations,
 * (nSteps + 1) x elMatDim each, row 0 the start (host); loadScale: the nodal
 * load factor of each step (null: 1).  Needs astro_ctx_dynamic_init (its
 * factored tangent K + M / (beta dt^2) serves every adjoint step) and
 * astro_ctx_set_design_vars.  objSum (may be null): J.  Synchronous. */
int astro_ctx_dynamic_objective_gradient(astro_ctx* ctx, int32_t nSteps, const double* U, const double* A,
                                         const double* loadScale, const double* termWindow, int32_t numTerms,
                                         const astro_objective_term* terms, double* objSum, double* dLdD);

/* Thermal twins (Working/AStrO_commandFunctions.f90:16-74, 229-258, 689-703,
 * 763-769). */
int astro_ctx_assemble_thermal(astro_ctx* ctx);
int astro_ctx_thermal_solve(astro_ctx* ctx);
/* getThermalSolnLoad(0) on a zeroed load (Working/AStrO_commandFunctions.f90:16-74):
 * "thermalLoad" = -K T for the current nodeTemp (reactionHeatGen is
 * -thermalLoad).  Synchronous; needs astro_ctx_assemble_thermal first. */
int astro_ctx_thermal_soln_load(astro_ctx* ctx);
/* Transient heat conduction, the thermal branch of solve() with dynamic = 1
 * (Working/AStrO_commandFunctions.f90:763-769, 689-703, 791-803; element
 * residual r_getElRt / r_getElRtm, Working/AStrO_r_elementEqns.f90:2126-2205):
 * Tdot = (1 / gamma) ((T - Tp) / dt - (1 - gamma) Tdp), element residual
 * K T + C Tdot with the capacitance C = rho c N N^T, tangent
 * K + C / (gamma dt), factored once by thermal_dynamic_init (the state
 * starts at T = Tdot = 0).  A step: RHS -R(T) + C^T (rhs - C T) of the
 * temperature constraints, GMRES, T += delta, then Tdot and the previous
 * state.  brick8 models; "nodeTemp", "nodeTdot" readable.  Synchronous. */
int astro_ctx_thermal_dynamic_init(astro_ctx* ctx, double gamma, double dt);
int astro_ctx_thermal_dynamic_run(astro_ctx* ctx, int nSteps);

/* Preconditioner apply on device vectors of length elMatDim:
 * x = P^-1 b via the two triangular sweeps.  b and x must be distinct buffers
 * (ASTRO_ERR_ARG otherwise). */
int astro_ctx_precond_apply(astro_ctx* ctx, const double* d_b, double* d_x);
/* Operator apply y = (A + C^T C) x on device vectors. */
int astro_ctx_operator_apply(astro_ctx* ctx, const double* d_x, double* d_y);

/* Per-launch timing of the preconditioner sweeps with events on the context
 * stream (on = 1 resets the totals).  astro_ctx_kernel_stat names:
 * "sweep_fwd_ms", "sweep_bwd_ms" (summed device time), "sweep_fwd_count",
 * "sweep_bwd_count", "profile_tiles", "panels_per_block", "split_k". */
int astro_ctx_set_timing(astro_ctx* ctx, int on);
double astro_ctx_kernel_stat(const astro_ctx* ctx, const char* name);

/* |b - (A + C^T C) u| / |b| of the whole static system, b = nodal loads +
 * C^T rhs, at the current nodeDisp u; the residual is formed in double-double
 * on the device (synchronous). */
int astro_ctx_true_residual(astro_ctx* ctx, double* relRes);

/* ---- Learned block preconditioner --------------------------------------
 * No counterpart in the reference (its Newmark loop, Working/AStrO_commandFunctions.f90:
 * 781-815, preconditions GMRES with the skyline LDL^T alone).  A graph network
 * over the mesh's node graph (encoder, `layers` message-passing layers of
 * `hidden` FP32 units with mean aggregation, linear head) maps per-node
 * features of the assembled operator A (its normalised 3 x 3 diagonal block
 * D_i, log stiffness, constraint flag, neighbour coupling) to symmetric
 * blocks B_i = L_i^-T (I + Y_i) L_i^-1, D_i = L_i L_i^T.  With the blocks
 * attached, GMRES uses the two-stage preconditioner z = P^-1 r, z += B (r - A z);
 * the plain LDL^T then judges the result (a plain-preconditioned GMRES with
 * absolute target rtol |P^-1 r0| returns at once when the target is met) and
 * continues as fallback when it is not.  brick8 meshes.
 *
 * Training data are generated by the solver: astro_ctx_gnn_collect, on a
 * context whose operator is factored (after astro_ctx_dynamic_init), forms
 * nRhs seeded pairs (b, x) of A x = b and keeps r1 = b - A P^-1 b, the
 * residual the LDL^T stage leaves, with sum e1^T r1, e1 = x - P^-1 b =
 * A^-1 r1 its error, and a copy of A; the loss of a field is the energy norm
 * of the error the block stage leaves, sum (e1 - B r1)^T A (e1 - B r1) /
 * sum e1^T A e1 over its pairs.  pairs = ASTRO_GNN_PAIRS_SOLUTION: x standard
 * normal, b = A x (the errors cover the whole spectrum, so the loss sees the
 * modes a block stage can amplify); ASTRO_GNN_PAIRS_SOLVE: b standard normal,
 * x = A^-1 b by LDL^T-preconditioned GMRES to rtol (the errors are A^-1-
 * smoothed and the loss is blind to amplified stiff modes).
 * astro_gnn_train runs Adam (lr, epochs over the collected fields in seeded
 * order) with the library's forward and backward kernels; lossPerEpoch
 * (epochs, may be null) receives the mean field loss of each epoch.
 * astro_gnn_loss_grad evaluates one field's loss and gradient (grad: the
 * parameter count, may be null); astro_gnn_get_field / _field_size /
 * _field_blocks expose a field's inputs, statistics and blocks.
 * astro_ctx_gnn_get reads "features" (numNodes x 9), "G" and "blocks"
 * (numNodes x 9) of a context. */
typedef struct astro_gnn astro_gnn;
int astro_gnn_create(int hidden, int layers, uint64_t seed, astro_gnn** out);
int astro_gnn_destroy(astro_gnn* gnn);
int64_t astro_gnn_num_params(const astro_gnn* gnn);
int astro_gnn_num_fields(const astro_gnn* gnn);
int astro_gnn_get_params(astro_gnn* gnn, float* host, int64_t count);
int astro_gnn_set_params(astro_gnn* gnn, const float* host, int64_t count);
enum { ASTRO_GNN_PAIRS_SOLUTION = 0, ASTRO_GNN_PAIRS_SOLVE = 1 };
int astro_ctx_gnn_collect(astro_ctx* ctx, astro_gnn* gnn, int nRhs, uint64_t seed, double rtol, int32_t pairs);
int astro_gnn_train(astro_gnn* gnn, int epochs, double lr, uint64_t seed, double* lossPerEpoch);
int astro_gnn_loss_grad(astro_gnn* gnn, int field, double* loss, float* grad);
int astro_gnn_field_size(const astro_gnn* gnn, int field, int32_t* numNodes, int32_t* numAdj, int32_t* numDofs,
                         int32_t* samples);
int astro_gnn_get_field(astro_gnn* gnn, int field, int32_t* range, int32_t* adj, float* X, double* G, double* R1,
                        double* eTr);
int astro_gnn_field_blocks(astro_gnn* gnn, int field, double* B);
int astro_ctx_gnn_attach(astro_ctx* ctx, astro_gnn* gnn);
int astro_ctx_gnn_detach(astro_ctx* ctx);
int astro_ctx_gnn_get(astro_ctx* ctx, const char* name, double* host, int64_t count);

/* Allocation-free variant of astro_ctx_true_residual.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * For a caller-provided device displacement d_u (elMatDim doubles) it writes
 * the double-double-accurate residual r = b - (A + C^T C) u to the caller's
 * device buffer d_r (distinct from d_u) and |r| / |b| to *relRes, with
 * b = nodal loads + C^T rhs, using the same kernels and reduction order as
 * astro_ctx_true_residual (so at u = nodeDisp the two return the same value).
 * It calls no hipMalloc/hipFree: its working vectors live in a static device
 * array of 4.5 MiB that each GPU holds until the process ends.  Limits:
 * elMatDim <= 524288 and elMPCDim <= 32768 (ASTRO_ERR_ARG beyond).  Every call
 * on one GPU shares that array, so calls that overlap on one GPU (from
 * several threads or contexts) race; the array keeps the last call's
 * contents.  Synchronous on the context stream. */
int astro_ctx_true_residual_noalloc(astro_ctx* ctx, const double* d_u, double* d_r, double* relRes);

/* Wait for all work queued on the context stream. */
int astro_ctx_sync(astro_ctx* ctx);
/* Per-phase device times (ms) of the last astro_ctx_static_solve, measured
 * with events on the context stream: "assemble", "scale", "profile",
 * "factor", "solve". */
double astro_ctx_phase_ms(const astro_ctx* ctx, const char* phase);

#ifdef __cplusplus
}
#endif

#endif /* ASTRO_ABI_H */
