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
// [r0, r1).  Tiles are numbered by blk: the chains (independent diagonal
// blocks) one after another, each in row order, so the tiles of a chain are
// consecutive 64-row blocks with consecutive blk.  In sweep direction (up for
// the forward sweep, down for the backward sweep) the 1st..3rd neighbours are
// the "near" tiles; nNear says how many neighbours exist, counted up to 4
// (the 3rd and 4th neighbours' coupling entries are read into registers at
// the start).  The tiles beyond them that the tile couples to are "far": blk
// from farBlk (the farthest one; -1 for none) up to the 5th neighbour.  lim: the chain's last row + 1, which bounds
// a ragged final tile.  Completion flags are indexed by blk.
struct SweepTile {
    int r0, r1, blk, nNear, farBlk, lim;
};

// Per-tile 64x64 blocks derived from the factor.  With L_tt the unit-lower
// diagonal block, F_t = L_tt^-1 L(t, t-1) and H_t = L_tt^-1 L(t, t-2), the
// forward sweep z = L^-1 b obeys, with g_t = L_tt^-1 (b_t - sum over tiles
// s <= t-3 of L(t, s) z_s),
//   z_t = g_t - F_t g_{t-1} + (F_t H_{t-1}) z_{t-3} + (F_t F_{t-1} - H_t) z_{t-2},
// so a tile waits on the tile two back, not on its neighbour, and only the
// last product sits on the critical path.  The backward sweep is the same
// with G_t = L_tt^-T L(t+1, t)^T, HB_t = L_tt^-T L(t+2, t)^T and the tiles
// after t.  Missing neighbours count as zero blocks.  Tile rows [r0, r1);
// rows [pm1, r0) and [pm2, pm1) are tiles t-1 and t-2 (empty: none), rows
// [r1, q1) and [q1, q2) tiles t+1 and t+2; prevBlk / nextBlk index the
// blocks of t-1 / t+1 (-1: none).  Each block is column-major (entry (i, c)
// at c * 64 + i) and zero outside the rows that exist.
struct SweepBlock {
    int r0, r1, pm1, pm2, q1, q2, prevBlk, nextBlk;
};
enum {
    SWEEP_DINV = 0, SWEEP_F = 1, SWEEP_H = 2, SWEEP_FH = 3, SWEEP_K = 4,
    SWEEP_DINVT = 5, SWEEP_G = 6, SWEEP_HB = 7, SWEEP_GHB = 8, SWEEP_KB = 9,
    SWEEP_KINDS = 10, SWEEP_BLOCK = 64 * 64
};


__global__ void k_brick8_tangent(const int* perm, int count, const int* conn, const double* X, const int* elMat,
                                 const double* elScale, const double* matC, const double* elRho, double massFactor,
                                 const int* rank, int nS, const int* csrRange, const uint8_t* elPos, double* val);
__global__ void k_brick8_tangent_gl(const int* perm, int count, const int* conn, const double* X, const int* elMat,
                                    const double* elScale, const double* matC, const int* rank, int nS, const double* u,
                                    const double* tExp, const double* T, const int* csrRange, const uint8_t* elPos,
                                    double* val);
__global__ void k_shell4_tangent(const int* perm, int count, const int* conn, const double* X, const double* elABD,
                                 const double* elAlpha, const int* rank, int nSolid, int nSec, const int* csrRange,
                                 const uint8_t* elPos, double* val);
__global__ void k_shell4_corot(const int* perm, int count, const int* conn, const double* X, const double* elABD,
                               const double* elAlpha, const int* rank, int nSolid, int nSec, const double* u, double h,
                               int iters, double* force, double* energy);
__global__ void k_shell4_corot_tangent(const int* perm, int count, const int* conn, const double* X,
                                       const double* elABD, const double* elAlpha, const int* rank, int nSolid,
                                       int nSec, const double* u, double h, int iters, double eps,
                                       const int* csrRange, const uint8_t* elPos, double* val);
__global__ void k_shell4_dtangent_cs(const int* perm, int count, const int* conn, const double* X, const double* elABD,
                                     const double* elABDim, const double* elAlpha, const int* rank, int nSolid, int nSec,
                                     const double* u, const int* elDV, double invH, int64_t n, double* dRdD,
                                     double* gPart, double* kim, int* kimDof);
__global__ void k_brick8_stress(int E, const int* conn, const double* X, const int* elMat, const double* elScale,
                                const double* matC, const int* rank, int nS, const double* u, const double* tExp,
                                const double* T, double s0, double s1, double s2, double* stress, double* strain);
__global__ void k_brick8_thermal_strain(int mode, const int* perm, int count, const int* conn, const double* X,
                                        const int* elMat, const double* elScale, const double* matC,
                                        const double* tExp, const int* rank, int nS, const double* T,
                                        const double* lam, double scale, double* out);
__global__ void k_brick8_residual(const int* perm, int count, const int* conn, const double* X, const int* elMat,
                                  const double* elScale, const double* matC, const double* elRho, const int* rank,
                                  int nS, const double* u, const double* pu, const double* pv, const double* pa,
                                  double c1, double c2, double dt, double beta, const double* tExp, const double* T,
                                  double* load, const int* gate);
__global__ void k_newmark_update(int n, double* u, const double* du, double* pu, double* pv, double* pa, double* v,
                                 double* a, double c1, double c2, double dt, double beta, double gamma, const int* gate);
__global__ void k_thermal_rate(int n, const double* T, const double* Tp, const double* Tdp, double gamma, double dt,
                               double cm, double* w);
__global__ void k_thermal_update(int n, double* T, const double* dT, double* Tp, double* Tdp, double* Tdot,
                                 double gamma, double dt);
__global__ void k_brick8_conduct(const int* perm, int count, const int* conn, const double* X, const double* elCond,
                                 double condFactor, const double* elCap, double capFactor, const int* rank,
                                 const int* csrRange, const uint8_t* thPos, double* val);
__global__ void k_brick8_thermal_residual(const int* perm, int count, const int* conn, const double* X,
                                          const double* elCond, const double* elCap, const int* rank, const double* T,
                                          const double* Tdot, double* load);
// k ranges of step U: all columns before the panel, only those before the
// previous panel ("old", which does not wait for the previous panel's
// factor), or only the previous panel's 64 ("new").
enum { LDL_K_ALL = 0, LDL_K_OLD = 1, LDL_K_NEW = 2 };

// Blocked left-looking LDL^T: step U alone (kMode 0 all columns before the
// panel, 1 all but the previous panel, 2 the previous panel only) and the fused
// step (FT of panelIdx, "old" U of panelIdx + 1).
// The profile kernels are templates on the storage type LT of the profile
// (and of step U's partial sums): double for the FP64 path, float for the
// mixed-precision factor (FP32 L and step U on FP32 matrix cores, FP64
// pivots in diag, FP64 diagonal-block factor and vectors).  Instantiated for
// both in astro_kernels.hip.
template <typename LT>
__global__ void k_ldl_update_mfma(const PanelInfo* panels, int panelIdx, int numPanelsPerBlock, int nSplit, int kMode,
                                  int slot0, const LT* lt, const int64_t* rowoff, const int* sfirst,
                                  const int* winMin, const double* diag, LT* scratch, int scratchRows);
template <typename LT>
__global__ void k_ldl_lead(const PanelInfo* panels, int panelIdx, int numPanelsPerBlock, int numBlocks, int numTiles, int nSplit,
                           LT* lt, const int64_t* rowoff, const int* sfirst, const int* winMin, double* diag,
                           LT* bufQ, LT* bufQ1, int scratchRows, int* zeroPivot, double* ftX);
template <typename LT>
__global__ void k_ldl_rows(const PanelInfo* panels, int panelIdx, int numPanelsPerBlock, int numBlocks, int numTiles, int nSplit,
                           LT* lt, const int64_t* rowoff, const int* sfirst, const LT* bufQ, int scratchRows,
                           const double* ftX);
__global__ void k_max_abs(const double* v, int64_t n, unsigned long long* out);
__global__ void k_scale_mpc(double* cmat, double* rhs, const int* crange, int cdim, const unsigned long long* maxBits);
__global__ void k_fill(double* p, int64_t n, double v);
__global__ void k_fill_i32(int* p, int64_t n, int v);
__global__ void k_fill_sentinel(double* p, int64_t n, const int* gate);
template <typename T>
__global__ void k_fill_t(T* p, int64_t n, T v);
template <typename LT>
__global__ void k_csr_to_profile(const double* val, const int* cols, const int* range, int n, const int* sfirst,
                                 const int64_t* rowoff, LT* lt);
template <typename LT>
__global__ void k_penalty_plan(const double* cmat, int nEnt, const int* entRow, const int* entCol, const int* start,
                               const int* pa, const int* pb, const int64_t* rowoff, LT* lt);
template <typename LT>
__global__ void k_sweep_blocks(const SweepBlock* blocks, const LT* lt, const int64_t* rowoff, const int* sfirst,
                               double* out);
__global__ void k_sweep_blocks_pair(const SweepBlock* blocks, double* out);
template <typename LT>
__global__ void k_sweep_forward(const SweepTile* tiles, int numTiles, const LT* lt, const int64_t* rowoff,
                                const int* sfirst, const double* blocks, const double* b, double* z, double* g,
                                int* flags, int* ticket, int epoch, long spinLimit, int* timeout, const int* gate);
// diag: FP64 pivots, read by the FP32 instantiation (the FP64 one reads them
// from the profile's diagonal)
template <typename LT>
__global__ void k_sweep_backward(const SweepTile* tiles, int numTiles, const LT* lt, const int64_t* rowoff,
                                 const int* sfirst, const double* blocks, const double* z, double* x, double* g,
                                 int* flags, int* ticket, int epoch, long spinLimit, int* timeout, const double* diag, const int* gate);
__global__ void k_reset_ticket(int* ticket, const int* gate);
__global__ void k_spmv(const double* val, const int* cols, const int* range, int n, const double* x, double* y, const int* gate);
__global__ void k_spmv_nodes(const double* bv, const int* nb, int K, int N, const double* x, double* y,
                             const int* gate);
__global__ void k_gather_values(const double* val, const int* map, int64_t count, double* bv);
__global__ void k_node_values(const double* val, const int* range, const uint8_t* jd, int K, int N, double* bv);
__global__ void k_mpc_dd(const double* cmat, const int* ccols, const int* crange, int cdim, const double* x,
                         double* t, const int* gate);
__global__ void k_residual_dd(const double* val, const int* cols, const int* range, int n, const int* ctRange,
                              const int* ctPos, const int* ctRow, const double* cmat, const double* t,
                              const double* b, const double* x, double* r, const int* gate);
__global__ void k_mpc_apply(const double* cmat, const int* ccols, const int* crange, int cdim, const double* x,
                            double* y, const int* gate);
__global__ void k_nodal_loads(const int* dofs, const double* vals, int count, double scale, double* load, const int* gate);
__global__ void k_constraint_load(const double* cmat, const int* ccols, const int* crange, const double* crhs,
                                  int cdim, const double* u, double* load, const int* gate);
__global__ void k_multidot(const double* V, int64_t ldv, int nqTotal, const double* w, int n, double* partial,
                           int pstride, const int* gate);
__global__ void k_sum_partials(const double* partial, int nblk, int pstride, int nq, double* out, const int* gate);
__global__ void k_sub_combination(const double* V, int64_t ldv, int nq, const double* h, double* w, int n,
                                  double sign, const int* gate);
// Device-side GMRES cycle of a captured time step (layout in astro_kernels.hip).
enum { GMD_RES = 0, GMD_TARGET = 1, GMD_RES0 = 2, GMD_HNORM = 3, GMD_INVH = 4, GMD_INVRES = 5, GMD_JUDGE = 6, GMD_R = 8 };
enum { GMI_ACTIVE = 0, GMI_K = 1, GMI_RUN = 2, GMI_STEP = 3, GMI_FAILED = 4, GMI_COUNT = 8 };
// doubles of gd for restart length m
inline size_t gmd_size(int m) { return GMD_R + (size_t)(m + 1) * m + 2 * (size_t)m + (m + 1) + m; }
inline size_t gmd_y(int m) { return GMD_R + (size_t)(m + 1) * m + 2 * (size_t)m + (m + 1); }
__global__ void k_gm_begin(const double* nrm2, double rtol, double atol, double* gd, int* gi, int m);
__global__ void k_gm_givens(int j, const double* h1, const double* h2, const double* hn, double* gd, int* gi, int m);
__global__ void k_gm_lsq(double* gd, const int* gi, int m);
__global__ void k_sub_combination_dn(const double* V, int64_t ldv, const int* nq, const double* h, double* w, int n,
                                     double sign, const int* gate);
__global__ void k_gm_judge_target(const double* nrm2, double rtol, double* gd, const int* gi);
__global__ void k_gm_check(const double* nrm2, double* gd, int* gi, int slot);
__global__ void k_gm_advance(int* gi);
__global__ void k_scale_by(const double* x, double* y, int n, const double* s, const int* gate);
__global__ void k_fill_gated(double* p, int64_t n, double v, const int* gate);
__global__ void k_nodal_loads_dev(const int* dofs, const double* vals, int count, const double* scales,
                                  const int* stepIdx, double* load, const int* gate);
__global__ void k_scale_copy(const double* x, double* y, int n, const double* den);
__global__ void k_axpby(const double* x, const double* y, double* out, int n, double a, double b, const int* gate);
__global__ void k_count_zero_cols(const int* cols, int64_t nnz, int* out);

}  // namespace astro
