// astro_batch.hip -- complex-step design sensitivities as a batch of complex128
// systems in one persistent launch.
//
// For design variable k the perturbed operator is
//   A_k = A(d + i h e_k) = A_re + i Im(A_k),
// A_re = K + C^T C of the current design (shared by all variables; with
// h = 1e-30 the real part of every complex-step product equals the real
// product in FP64) and Im(A_k) the imaginary parts of the condensed complex
// tangents of variable k's elements.  Each instance assembles A_k in its own
// profile, factors it as a complex-symmetric LDL^T (transpose, not conjugate),
// solves A_k u_k = F and forms the compliance C_k = F . u_k; then
// dC/dd_k = Im(C_k) / h and Re(C_k) is the compliance itself.  The rows of
// A_k before r_k, the first row any of variable k's elements touches, are
// real, so their factor is real too: an instance runs real arithmetic below
// r_k and complex arithmetic from there on (the instances are uneven).
//
// The profile is that of a node-major DOF order (nodes by their level rank,
// the six DOFs of a node together), so that a row reaches back one node level
// plus one node: for the 128 x 16 shell plate 210 columns at most.  One
// workgroup factors one instance: left-looking LDL^T over blocks of RB rows,
// the columns before a block in chunks of CB (a product from LDS then a small
// triangular solve per row), the block's own diagonal part sequentially; then
// the blocked forward and backward substitutions.  (RB, CB) is the tile shape
// the autotuner chooses (astro_ctx_cs_batch_autotune).  Two refinement steps
// follow: the residual F - A_k u_k with A_re's products and sums in
// double-double (k_csb_residual, k_csb_residual_el), solved with the
// instance's factor by the same kernel in solve-only mode.
//
// The reference evaluates the same derivative per variable with complex-step
// element residuals and one real adjoint (getdRelasticdD,
// Working/AStrO_commandFunctions.f90:1172-1375; getTotaldLdD :1376-1466).
#include <hip/hip_runtime.h>

#include "astro_batch.h"
#include "astro_dd.h"

namespace astro {

struct cplx {
    double r, i;
};

// CX = false: real arithmetic (the imaginary parts stay zero)
template <bool CX>
__device__ __forceinline__ cplx c_mul(cplx a, cplx b) {
    if (CX) return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
    return {a.r * b.r, 0.0};
}
// acc - a b
template <bool CX>
__device__ __forceinline__ cplx c_fms(cplx acc, cplx a, cplx b) {
    if (CX) return {acc.r - a.r * b.r + a.i * b.i, acc.i - a.r * b.i - a.i * b.r};
    return {acc.r - a.r * b.r, 0.0};
}
template <bool CX>
__device__ __forceinline__ cplx c_inv(cplx a) {
    if (CX) {
        const double d = a.r * a.r + a.i * a.i;
        return d == 0.0 ? cplx{0.0, 0.0} : cplx{a.r / d, -a.i / d};
    }
    return {a.r == 0.0 ? 0.0 : 1.0 / a.r, 0.0};
}

__device__ __forceinline__ int min_first(const int* __restrict__ first, int i0, int i1) {
    int s = first[i0];
    for (int i = i0 + 1; i < i1; ++i) s = min(s, first[i]);
    return s;
}

__global__ void k_csb_zero(double* __restrict__ p, int64_t count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// The real profile of A_re in the node-major order: CSR entries (reference
// 1-based columns, old DOF order) mapped through inv.
__global__ void k_csb_real_profile(int n, const int* __restrict__ range, const int* __restrict__ cols,
                                   const double* __restrict__ val, const int* __restrict__ inv,
                                   const int* __restrict__ first, const int64_t* __restrict__ rowoff,
                                   double* __restrict__ prof) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    const int I = inv[row];
    for (int q = range[row]; q < range[row + 1]; ++q) {
        const int c = cols[q];
        if (c < 1 || c > n) continue;  // unused slot
        const int J = inv[c - 1];
        if (J <= I) prof[rowoff[I] + J - first[I]] += val[q];
    }
}

// + C^T C: constraint row r adds c_a c_b at (a, b); constraints may share a
// DOF, so the adds are atomic.
__global__ void k_csb_penalty(int cd, int n, const double* __restrict__ cmat, const int* __restrict__ ccols,
                              const int* __restrict__ crange, const int* __restrict__ inv,
                              const int* __restrict__ first, const int64_t* __restrict__ rowoff,
                              double* __restrict__ prof) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= cd) return;
    for (int a = crange[r]; a < crange[r + 1]; ++a)
        for (int b = crange[r]; b < crange[r + 1]; ++b) {
            if (ccols[a] < 1 || ccols[a] > n || ccols[b] < 1 || ccols[b] > n) continue;
            const int I = inv[ccols[a] - 1], J = inv[ccols[b] - 1];
            if (J <= I) atomicAdd(&prof[rowoff[I] + J - first[I]], cmat[a] * cmat[b]);
        }
}

// Factor rows [i0, i0 + RB) of one instance's profile in place (L below the
// diagonal, D on it).  LDS: G (RB x W) the block rows' D-scaled values
// g_ij = L_ij D_j, Lc (CB x W) the rows of the current column chunk, Dv (W)
// 1 / D of the columns before the block, Db / Dbi (RB) the block's D and 1 / D.
template <int RB, int CB, bool CX>
__device__ void csb_factor_block(int n, int W, const int* __restrict__ first, const int64_t* __restrict__ rowoff,
                                 cplx* __restrict__ prof, int i0, cplx* G, cplx* Lc, cplx* Dv, cplx* Db, cplx* Dbi,
                                 int* status) {
    const int tid = threadIdx.x;
    const int i1 = min(i0 + RB, n), nr = i1 - i0;
    const int s = min_first(first, i0, i1);
    const int W0 = i1 - s;
    const cplx zero = {0.0, 0.0};
    for (int idx = tid; idx < nr * W0; idx += CSB_THREADS) {
        const int r = idx / W0, c = idx - r * W0, col = s + c, i = i0 + r;
        G[r * W + c] = (col >= first[i] && col <= i) ? prof[rowoff[i] + col - first[i]] : zero;
    }
    for (int c = tid; c < i0 - s; c += CSB_THREADS) {
        const int j = s + c;
        Dv[c] = c_inv<CX>(prof[rowoff[j] + j - first[j]]);
    }
    __syncthreads();
    // the columns before the block, CB at a time
    for (int j0 = s; j0 < i0; j0 += CB) {
        const int cw = min(CB, i0 - j0), wl = j0 + cw - s, cj = j0 - s;
        for (int idx = tid; idx < cw * wl; idx += CSB_THREADS) {
            const int q = idx / wl, c = idx - q * wl, j = j0 + q, col = s + c;
            Lc[q * W + c] = (col >= first[j] && col < j) ? prof[rowoff[j] + col - first[j]] : zero;
        }
        __syncthreads();
        {
            const int r = tid / CB, q = tid - (tid / CB) * CB;
            if (r < nr && q < cw) {
                cplx acc = G[r * W + cj + q];
                const cplx* gr = G + r * W;
                const cplx* lq = Lc + q * W;
                for (int c = 0; c < cj; ++c) acc = c_fms<CX>(acc, gr[c], lq[c]);
                G[r * W + cj + q] = acc;
            }
        }
        __syncthreads();
        if (tid < nr) {
            cplx* gr = G + tid * W + cj;
            for (int q = 0; q < cw; ++q) {
                cplx v = gr[q];
                const cplx* lq = Lc + q * W + cj;
                for (int q2 = 0; q2 < q; ++q2) v = c_fms<CX>(v, gr[q2], lq[q2]);
                gr[q] = v;
            }
        }
        __syncthreads();
    }
    // the block's own columns: the product over the earlier columns, then
    // the rows one after the other
    const int c0 = i0 - s;
    for (int idx = tid; idx < nr * nr; idx += CSB_THREADS) {
        const int r = idx / nr, q = idx - r * nr;
        if (q > r) continue;
        cplx acc = G[r * W + c0 + q];
        const cplx* gr = G + r * W;
        const cplx* gq = G + q * W;
        for (int c = 0; c < c0; ++c) acc = c_fms<CX>(acc, c_mul<CX>(gr[c], Dv[c]), gq[c]);
        G[r * W + c0 + q] = acc;
    }
    __syncthreads();
    for (int q = 0; q < nr; ++q) {
        const int r = tid;
        if (r == q) {
            cplx v = G[q * W + c0 + q];
            for (int q2 = 0; q2 < q; ++q2)
                v = c_fms<CX>(v, c_mul<CX>(G[q * W + c0 + q2], Dbi[q2]), G[q * W + c0 + q2]);
            Db[q] = v;
            Dbi[q] = c_inv<CX>(v);
            if (v.r == 0.0 && v.i == 0.0) *status = 1;
        } else if (r > q && r < nr) {
            cplx v = G[r * W + c0 + q];
            for (int q2 = 0; q2 < q; ++q2)
                v = c_fms<CX>(v, c_mul<CX>(G[r * W + c0 + q2], Dbi[q2]), G[q * W + c0 + q2]);
            G[r * W + c0 + q] = v;
        }
        __syncthreads();
    }
    for (int idx = tid; idx < nr * W0; idx += CSB_THREADS) {
        const int r = idx / W0, c = idx - r * W0, col = s + c, i = i0 + r;
        if (col >= first[i] && col < i)
            prof[rowoff[i] + col - first[i]] = c_mul<CX>(G[r * W + c], col < i0 ? Dv[c] : Dbi[col - i0]);
        else if (col == i)
            prof[rowoff[i] + col - first[i]] = Db[r];
    }
    __syncthreads();
}

// One workgroup per instance (instance order[blockIdx.x]).
template <int RB, int CB>
__global__ __launch_bounds__(CSB_THREADS) void k_csb_batch(CsbArgs P) {
    extern __shared__ cplx lds[];
    const int W = P.W, n = P.n, tid = threadIdx.x;
    const int inst = P.order[blockIdx.x];
    const int rk = P.instStart[inst];
    cplx* __restrict__ prof = reinterpret_cast<cplx*>(P.work) + (size_t)inst * (size_t)P.total;
    cplx* __restrict__ z = reinterpret_cast<cplx*>(P.vec) + (size_t)inst * (size_t)n;
    cplx* __restrict__ u = reinterpret_cast<cplx*>(P.sol) + (size_t)inst * (size_t)n;
    const int* __restrict__ first = P.first;
    const int64_t* __restrict__ rowoff = P.rowoff;
    cplx* G = lds;
    cplx* Lc = G + RB * W;
    cplx* Dv = Lc + CB * W;
    cplx* Db = Dv + W;
    cplx* Dbi = Db + RB;
    if (P.mode == 0) {
    // assembly: A_re, then the imaginary parts of the instance's elements
    for (int64_t q = tid; q < P.total; q += CSB_THREADS) prof[q] = {P.profRe[q], 0.0};
    __syncthreads();
    for (int k = P.elStart[inst]; k < P.elStart[inst + 1]; ++k) {
        const int e = P.elList[k];
        const double* ke = P.kim + 576 * (size_t)e;
        const int* de = P.kimDof + 24 * (size_t)e;
        for (int t = tid; t < 576; t += CSB_THREADS) {
            const int a = t / 24, b = t - (t / 24) * 24;
            const int I = P.inv[de[a] - 1], J = P.inv[de[b] - 1];
            if (J <= I) prof[rowoff[I] + J - first[I]].i += ke[t];
        }
        __syncthreads();
    }
    // factor: real blocks below r_k, complex from the block holding r_k on
    for (int i0 = 0; i0 < n; i0 += RB) {
        if (i0 + RB <= rk)
            csb_factor_block<RB, CB, false>(n, W, first, rowoff, prof, i0, G, Lc, Dv, Db, Dbi, P.status);
        else
            csb_factor_block<RB, CB, true>(n, W, first, rowoff, prof, i0, G, Lc, Dv, Db, Dbi, P.status);
    }
    for (int i = tid; i < n; i += CSB_THREADS) z[i] = {P.F[i], 0.0};
    }  // mode 0; in mode 1 z holds the residual of u
    // forward substitution L z = F, blocks of RB rows: the part of each
    // block row's sum over earlier blocks by RB groups of threads, then the
    // block's own columns by wave 0, lane q holding z of block row q, rows
    // in order with a shuffle reduction (RB <= 32)
    cplx* red = lds;              // CSB_THREADS entries (the factor's LDS is free now)
    cplx* Lb = lds + CSB_THREADS;  // RB x RB: L of the block's own columns
    cplx* zb = Lb + RB * RB;       // RB
    constexpr int TPR = CSB_THREADS / RB;  // threads per row
    static_assert(RB <= 32, "the in-block solves use one half-wave");
    __syncthreads();
    for (int i0 = 0; i0 < n; i0 += RB) {
        const int i1 = min(i0 + RB, n), nr = i1 - i0;
        {
            const int r = tid / TPR, l = tid - r * TPR;
            cplx acc = {0.0, 0.0};
            if (r < nr) {
                const int i = i0 + r;
                const cplx* Li = prof + rowoff[i] - first[i];
                for (int p = first[i] + l; p < i0; p += TPR) acc = c_fms<true>(acc, Li[p], z[p]);
            }
            red[tid] = acc;
        }
        for (int idx = tid; idx < nr * nr; idx += CSB_THREADS) {
            const int r = idx / nr, q = idx - r * nr, i = i0 + r, col = i0 + q;
            Lb[r * RB + q] = (q < r && col >= first[i]) ? prof[rowoff[i] + col - first[i]] : cplx{0.0, 0.0};
        }
        __syncthreads();
        if (tid < nr) {
            cplx v = z[i0 + tid];
            for (int l = 0; l < TPR; ++l) v = {v.r + red[tid * TPR + l].r, v.i + red[tid * TPR + l].i};
            zb[tid] = v;
        }
        __syncthreads();
        if (tid < 32) {
            cplx zq = {0.0, 0.0};
            for (int r = 0; r < nr; ++r) {
                cplx t = tid < r ? c_mul<true>(Lb[r * RB + tid], zq) : cplx{0.0, 0.0};
                for (int off = 16; off > 0; off >>= 1) {
                    t.r += __shfl_xor(t.r, off, 32);
                    t.i += __shfl_xor(t.i, off, 32);
                }
                if (tid == r) zq = {zb[r].r - t.r, zb[r].i - t.i};
            }
            if (tid < nr) z[i0 + tid] = zq;
        }
        __syncthreads();
    }
    for (int i = tid; i < n; i += CSB_THREADS) z[i] = c_mul<true>(z[i], c_inv<true>(prof[rowoff[i] + i - first[i]]));
    __syncthreads();
    // backward substitution L^T x = z from the last block up: the block's
    // own rows by wave 0 (last row first), then its rows' columns before it,
    // in parallel
    for (int i1 = n; i1 > 0; i1 -= RB) {
        const int i0 = max(i1 - RB, 0), nr = i1 - i0;
        for (int idx = tid; idx < nr * nr; idx += CSB_THREADS) {
            const int r = idx / nr, q = idx - r * nr, i = i0 + r, col = i0 + q;
            Lb[r * RB + q] = (q < r && col >= first[i]) ? prof[rowoff[i] + col - first[i]] : cplx{0.0, 0.0};
        }
        __syncthreads();
        if (tid < 32) {
            cplx xq = tid < nr ? z[i0 + tid] : cplx{0.0, 0.0};
            for (int r = nr - 1; r >= 0; --r) {
                // x_r = z_r - sum_{q > r} L(q, r) x_q
                cplx t = (tid > r && tid < nr) ? c_mul<true>(Lb[tid * RB + r], xq) : cplx{0.0, 0.0};
                for (int off = 16; off > 0; off >>= 1) {
                    t.r += __shfl_xor(t.r, off, 32);
                    t.i += __shfl_xor(t.i, off, 32);
                }
                if (tid == r) xq = {xq.r - t.r, xq.i - t.i};
            }
            if (tid < nr) z[i0 + tid] = xq;
        }
        __syncthreads();
        const int s = min_first(first, i0, i1);
        for (int p = s + tid; p < i0; p += CSB_THREADS) {
            cplx v = z[p];
            for (int j = i0; j < i1; ++j)
                if (first[j] <= p) v = c_fms<true>(v, prof[rowoff[j] + p - first[j]], z[j]);
            z[p] = v;
        }
        __syncthreads();
    }
    // u = z (mode 0) or u += z (mode 1), then C = F . u
    cplx acc = {0.0, 0.0};
    for (int i = tid; i < n; i += CSB_THREADS) {
        const cplx v = P.mode == 0 ? z[i] : cplx{u[i].r + z[i].r, u[i].i + z[i].i};
        u[i] = v;
        acc = {acc.r + P.F[i] * v.r, acc.i + P.F[i] * v.i};
    }
    red[tid] = acc;
    __syncthreads();
    for (int w = CSB_THREADS / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] = {red[tid].r + red[tid + w].r, red[tid].i + red[tid + w].i};
        __syncthreads();
    }
    if (tid == 0) {
        P.grad[inst] = red[0].i * P.invH;
        P.compliance[inst] = red[0].r;
    }
}

// r_k = F - A_k u_k for every instance (blockIdx.y): A_re u_k here, A_re as the CSR of K in
// the old DOF order plus C^T C through the transposed constraint structure
// (ctRange / ctPos / ctRow over old DOFs), products and sums of both the
// real and the imaginary part in double-double; 32 lanes per row.
__global__ __launch_bounds__(256) void k_csb_residual(CsbRes R) {
    const int inst = blockIdx.y;
    const int I = (blockIdx.x * blockDim.x + threadIdx.x) >> 5;
    const int l = threadIdx.x & 31;
    if (I >= R.n) return;
    const cplx* __restrict__ u = reinterpret_cast<const cplx*>(R.sol) + (size_t)inst * (size_t)R.n;
    const int row = R.order[I];
    double rh = 0.0, rl = 0.0, ih = 0.0, il = 0.0;
    for (int q = R.range[row] + l; q < R.range[row + 1]; q += 32) {
        const int c = R.cols[q];
        if (c < 1 || c > R.n) continue;
        const cplx x = u[R.inv[c - 1]];
        dd_add_prod(rh, rl, R.val[q], x.r);
        dd_add_prod(ih, il, R.val[q], x.i);
    }
    if (R.ctRange) {
        for (int k = R.ctRange[row] + l; k < R.ctRange[row + 1]; k += 32) {
            const double a = R.cmat[R.ctPos[k]];
            const int cr = R.ctRow[k];
            double th = 0.0, tl = 0.0, sh = 0.0, sl = 0.0;  // (C u)_cr
            for (int b = R.crange[cr]; b < R.crange[cr + 1]; ++b) {
                const int c = R.ccols[b];
                if (c < 1 || c > R.n) continue;
                const cplx x = u[R.inv[c - 1]];
                dd_add_prod(th, tl, R.cmat[b], x.r);
                dd_add_prod(sh, sl, R.cmat[b], x.i);
            }
            dd_add_prod(rh, rl, a, th);
            dd_add_prod(rh, rl, a, tl);
            dd_add_prod(ih, il, a, sh);
            dd_add_prod(ih, il, a, sl);
        }
    }
    for (int off = 16; off > 0; off >>= 1) {
        const double a = __shfl_xor(rh, off, 32), b = __shfl_xor(rl, off, 32);
        const double c = __shfl_xor(ih, off, 32), d = __shfl_xor(il, off, 32);
        dd_add(rh, rl, a, b);
        dd_add(ih, il, c, d);
    }
    if (l == 0) {
        // F - (A_re u) + the element terms, rounded once
        const double* ed = R.ed + 4 * ((size_t)inst * (size_t)R.n + I);
        double ph = ed[0], pl = ed[1], qh = ed[2], ql = ed[3];
        dd_add(ph, pl, R.F[I], 0.0);
        dd_add(ph, pl, -rh, -rl);
        dd_add(qh, ql, -ih, -il);
        cplx* __restrict__ r = reinterpret_cast<cplx*>(R.vec) + (size_t)inst * (size_t)R.n;
        r[I] = {ph + pl, qh + ql};
    }
}

// The imaginary parts of the instance's element tangents, first (one
// workgroup per instance, elements in turn since they share nodes), in
// double-double into ed (4 doubles per row): with A = A_re + i A_im,
// F - A u has the terms +A_im Im u (real part) and -A_im Re u (imaginary).
__global__ __launch_bounds__(64) void k_csb_residual_el(CsbRes R) {
    const int inst = blockIdx.x, a = threadIdx.x;
    const cplx* __restrict__ u = reinterpret_cast<const cplx*>(R.sol) + (size_t)inst * (size_t)R.n;
    double* __restrict__ ed = R.ed + 4 * (size_t)inst * (size_t)R.n;
    for (int i = a; i < 4 * R.n; i += 64) ed[i] = 0.0;
    __syncthreads();
    for (int k = R.elStart[inst]; k < R.elStart[inst + 1]; ++k) {
        const int e = R.elList[k];
        if (a < 24) {
            const double* ke = R.kim + 576 * (size_t)e + 24 * a;
            const int* de = R.kimDof + 24 * (size_t)e;
            const int I = R.inv[de[a] - 1];
            double rh = ed[4 * I], rl = ed[4 * I + 1], ih = ed[4 * I + 2], il = ed[4 * I + 3];
            for (int b = 0; b < 24; ++b) {
                const cplx x = u[R.inv[de[b] - 1]];
                dd_add_prod(rh, rl, ke[b], x.i);
                dd_add_prod(ih, il, -ke[b], x.r);
            }
            ed[4 * I] = rh;
            ed[4 * I + 1] = rl;
            ed[4 * I + 2] = ih;
            ed[4 * I + 3] = il;
        }
        __syncthreads();
    }
}

hipError_t csb_residual(const CsbRes& R, int nInst, hipStream_t s) {
    hipLaunchKernelGGL(k_csb_residual_el, dim3(nInst), dim3(64), 0, s, R);
    hipLaunchKernelGGL(k_csb_residual, dim3((R.n * 32 + 255) / 256, nInst), dim3(256), 0, s, R);
    return hipGetLastError();
}

static const int kTiles[][2] = {{16, 16}, {8, 16}, {8, 32}, {16, 8}, {32, 8}};

int csb_num_tiles() { return (int)(sizeof(kTiles) / sizeof(kTiles[0])); }
void csb_tile(int t, int* rb, int* cb) {
    *rb = kTiles[t][0];
    *cb = kTiles[t][1];
}

// W: the LDS row stride (cplx) for the widest block of tile t; 4 W = 4 mod
// 32 dwords, so that the column-chunk reads of 8 consecutive rows hit
// different banks.
int csb_stride(int t, int maxRow) { return ((maxRow + kTiles[t][0] + 7) / 8) * 8 + 1; }

size_t csb_lds_bytes(int t, int maxRow) {
    const int RB = kTiles[t][0], CB = kTiles[t][1], W = csb_stride(t, maxRow);
    size_t cells = (size_t)(RB + CB + 1) * W + 2 * RB;
    const size_t solveCells = CSB_THREADS + (size_t)RB * RB + RB;
    if (cells < solveCells) cells = solveCells;
    return cells * sizeof(cplx);
}

template <int RB, int CB>
static hipError_t launch_tile(const CsbArgs& a, int nInst, size_t lds, hipStream_t s) {
    hipError_t e = hipFuncSetAttribute((const void*)k_csb_batch<RB, CB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_csb_batch<RB, CB>), dim3(nInst), dim3(CSB_THREADS), lds, s, a);
    return hipGetLastError();
}

hipError_t csb_launch(const CsbArgs& a, int tile, int nInst, hipStream_t s) {
    const size_t lds = csb_lds_bytes(tile, a.maxRow);
    switch (tile) {
        case 0: return launch_tile<16, 16>(a, nInst, lds, s);
        case 1: return launch_tile<8, 16>(a, nInst, lds, s);
        case 2: return launch_tile<8, 32>(a, nInst, lds, s);
        case 3: return launch_tile<16, 8>(a, nInst, lds, s);
        case 4: return launch_tile<32, 8>(a, nInst, lds, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t csb_real_profile(const int* range, const int* cols, const double* val, int n, const double* cmat,
                            const int* ccols, const int* crange, int cd, const int* inv, const int* first,
                            const int64_t* rowoff, double* prof, int64_t total, hipStream_t s) {
    hipLaunchKernelGGL(k_csb_zero, dim3(1024), dim3(256), 0, s, prof, total);
    hipLaunchKernelGGL(k_csb_real_profile, dim3((n + 255) / 256), dim3(256), 0, s, n, range, cols, val, inv, first,
                       rowoff, prof);
    if (cd > 0)
        hipLaunchKernelGGL(k_csb_penalty, dim3((cd + 255) / 256), dim3(256), 0, s, cd, n, cmat, ccols, crange, inv,
                           first, rowoff, prof);
    return hipGetLastError();
}

}  // namespace astro
