// astro_kernels.h -- launch descriptors and kernel declarations shared by the
// device code (astro_kernels.hip) and its host orchestration (astro_abi.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace astro {

// One 64-column panel of one independent diagonal block of the profile:
// columns [p, pe), rows [p, iend) may reach into it; smin = min first column
// of the panel's own rows.  p < 0 marks "no panel" for a block that has fewer
// panels than the longest one.
struct PanelInfo {
    int p, pe, iend, smin;
};

// One 64-row tile of a triangular sweep in processing (ticket) order: rows
// [r0, r1) of chain `chain` (an independent diagonal block); depFirst is the
// ticket index of the first tile it depends on (== own index if none).
// depLast: ticket index of the immediately neighbouring tile of the chain
// (preceding for the forward sweep, following for the backward sweep) when
// the tile depends on it, else -1; it is handled last, on the critical path.
struct SweepTile {
    int r0, r1, chain, depFirst, depLast;
};

__global__ void k_brick8_tangent(const int* perm, int count, const int* conn, const double* X, const int* elMat,
                                 const double* elScale, const double* matC, const double* elRho, double massFactor,
                                 const int* rank, int nS, const int* csrRange, const uint8_t* elPos, double* val);
__global__ void k_shell4_tangent(const int* perm, int count, const int* conn, const double* X, const double* elABD,
                                 const double* elAlpha, const int* rank, int nSolid, int nSec, const int* csrRange,
                                 const uint8_t* elPos, double* val);
__global__ void k_shell4_dtangent_cs(const int* perm, int count, const int* conn, const double* X, const double* elABD,
                                     const double* elABDim, const double* elAlpha, const int* rank, int nSolid, int nSec,
                                     const double* u, const int* elDV, double invH, int64_t n, double* dRdD,
                                     double* gPart);
__global__ void k_brick8_residual(const int* perm, int count, const int* conn, const double* X, const int* elMat,
                                  const double* elScale, const double* matC, const double* elRho, const int* rank,
                                  int nS, const double* u, const double* pu, const double* pv, const double* pa,
                                  double c1, double c2, double dt, double beta, double* load);
__global__ void k_newmark_update(int n, double* u, const double* du, double* pu, double* pv, double* pa, double* v,
                                 double* a, double c1, double c2, double dt, double beta, double gamma);
__global__ void k_brick8_conduct(const int* perm, int count, const int* conn, const double* X, const double* elCond,
                                 const int* rank, const int* csrRange, const uint8_t* thPos, double* val);
// k ranges of step U: all columns before the panel, only those before the
// previous panel ("old", which does not wait for the previous panel's
// factor), or only the previous panel's 64 ("new").
enum { LDL_K_ALL = 0, LDL_K_OLD = 1, LDL_K_NEW = 2 };

// Blocked left-looking LDL^T: step U alone (kMode 0 all columns before the
// panel, 1 all but the previous panel, 2 the previous panel only) and the fused
// step (FT of panelIdx, "old" U of panelIdx + 1).
__global__ void k_ldl_update_mfma(const PanelInfo* panels, int panelIdx, int numPanelsPerBlock, int nSplit, int kMode,
                                  int slot0, const double* lt, const int64_t* rowoff, const int* sfirst,
                                  const double* diag, double* scratch, int scratchRows);
__global__ void k_ldl_step(const PanelInfo* panels, int panelIdx, int numPanelsPerBlock, int numBlocks, int numTiles, int nSplit,
                           double* lt, const int64_t* rowoff, const int* sfirst, double* diag, const double* scratchFT,
                           double* scratchU, int scratchRows, int* zeroPivot);
__global__ void k_max_abs(const double* v, int64_t n, unsigned long long* out);
__global__ void k_scale_mpc(double* cmat, double* rhs, const int* crange, int cdim, const unsigned long long* maxBits);
__global__ void k_fill(double* p, int64_t n, double v);
__global__ void k_fill_i32(int* p, int64_t n, int v);
__global__ void k_fill_sentinel(double* p, int64_t n);
__global__ void k_csr_to_profile(const double* val, const int* cols, const int* range, int n, const int* sfirst,
                                 const int64_t* rowoff, double* lt);
__global__ void k_penalty_to_profile(const double* cmat, const int* ccols, const int* crange, int cdim,
                                     const int* sfirst, const int64_t* rowoff, double* lt);
__global__ void k_sweep_forward(const SweepTile* tiles, int numTiles, const double* lt, const int64_t* rowoff,
                                const int* sfirst, const double* b, double* z, int* flags, int* ticket, int epoch,
                                long spinLimit, int* timeout);
__global__ void k_sweep_backward(const SweepTile* tiles, int numTiles, const double* lt, const int64_t* rowoff,
                                 const int* sfirst, const double* z, double* x, int* flags, int* ticket, int epoch,
                                 long spinLimit, int* timeout);
__global__ void k_reset_ticket(int* ticket);
__global__ void k_spmv(const double* val, const int* cols, const int* range, int n, const double* x, double* y);
__global__ void k_mpc_apply(const double* cmat, const int* ccols, const int* crange, int cdim, const double* x,
                            double* y);
__global__ void k_nodal_loads(const int* dofs, const double* vals, int count, double scale, double* load);
__global__ void k_constraint_load(const double* cmat, const int* ccols, const int* crange, const double* crhs,
                                  int cdim, const double* u, double* load);
__global__ void k_multidot(const double* V, int64_t ldv, int nqTotal, const double* w, int n, double* partial,
                           int pstride);
__global__ void k_sum_partials(const double* partial, int nblk, int pstride, int nq, double* out);
__global__ void k_sub_combination(const double* V, int64_t ldv, int nq, const double* h, double* w, int n,
                                  double sign);
__global__ void k_scale_copy(const double* x, double* y, int n, const double* den);
__global__ void k_axpby(const double* x, const double* y, double* out, int n, double a, double b);
__global__ void k_count_zero_cols(const int* cols, int64_t nnz, int* out);

}  // namespace astro
