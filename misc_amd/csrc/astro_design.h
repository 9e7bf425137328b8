// astro_design.h -- kernels of the design-variable, objective and adjoint
// path (astro_design.hip), shared with the host orchestration (astro_abi.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace astro {

// Direction record of one (element, design variable) pair: the derivative of
// each property of the element with respect to the variable (the sum of the
// coefficients of the variable's map entries that reach the element).  The
// complex step perturbs every property p to p + i h dp at once.
enum {
    DVD_EP = 0,     // 9 engineering constants E1 E2 E3 nu12 nu13 nu23 G12 G13 G23
    DVD_COND = 9,   // 6 conductivities k11 k22 k33 k12 k13 k23
    DVD_RHO = 15,   // density
    DVD_X = 16,     // 8 nodes x 3 coordinates, node-major
    DVD_STRIDE = 40
};

// Quantity of an element objective term (the ASTRO_OBJ_* element categories).
enum { OBJQ_STRESS = 3, OBJQ_STRAIN = 4, OBJQ_STRAIN_ENERGY = 5, OBJQ_MASS = 6, OBJQ_VOLUME = 7 };

__global__ void k_dv_pair_force(int nPairs, const int* pairEl, const int* pairDV, const double* dir, const int* conn,
                                const double* X, const double* eprops, const int* rank, int nS, const double* u,
                                const double* acc, const double* rho, const double* lam, double h, double* lamDot,
                                double* dRdD, int64_t n);
__global__ void k_dv_pair_flux(int nPairs, const int* pairEl, const int* pairDV, const double* dir, const int* conn,
                               const double* X, const double* cond, const int* rank, const double* T,
                               const double* lam, double h, double* lamDot, double* dRtdD, int64_t n);
__global__ void k_obj_elements(int quantity, int comp, int nItems, const int* items, const int* conn, const double* X,
                               const int* elMat, const double* elScale, const double* matC, const double* rho,
                               const int* rank, int nS, const double* u, double* q, double* v, double* dqdu);
__global__ void k_obj_pairs(int quantity, int comp, int nItems, const int* pairIdx, const int* pairEl,
                            const double* dir, const int* conn, const double* X, const double* eprops,
                            const double* rho, const int* rank, int nS, const double* u, double h, double* dq,
                            double* dv);
__global__ void k_obj_nodes(int nItems, const int* dofs, const double* tgt, int tgtMode, double coef, double expo,
                            const double* x, double* val, double* dLdx);
__global__ void k_brick8_flux(int E, const int* conn, const double* X, const double* cond, const int* rank,
                              const double* T, double s0, double s1, double s2, double* flux, double* grad);
__global__ void k_obj_scatter_du(int count, const int* slot, const int* items, const double* w, const double* dqdu,
                                 const int* conn, const int* rank, int nS, double* dLdu);

}  // namespace astro
