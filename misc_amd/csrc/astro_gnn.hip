// astro_gnn.hip -- the learned block preconditioner's graph network: FP32
// forward, backward and Adam kernels, feature and statistics kernels, and the
// training driver behind the astro_gnn_* entry points (include/astro_abi.h).
// The network is described in astro_gnn.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>

#include "astro_gnn.h"
#include "astro_kernels.h"

namespace astro {

// ---------------------------------------------------------------------------
// Dense layers: out[i][c] = act(sum_k A[i][k] W(c, k) + sum_k A2[i][k] W2(c, k)
// + bias[c]) for c < M <= 64, K, K2 <= 64; W(c, k) = W[c K + k] (wT = 0) or
// W[k M + c] (wT = 1, the transposed use of a weight in the backward pass).
// 64 rows per workgroup; thread (r = tid / 4, g = tid % 4) forms row r,
// columns 16 g .. 16 g + 15, from LDS tiles of the rows and the weights.
__global__ __launch_bounds__(256) void k_gnn_dense(int N, const float* __restrict__ A, int K,
                                                   const float* __restrict__ W, const float* __restrict__ A2, int K2,
                                                   const float* __restrict__ W2, int wT,
                                                   const float* __restrict__ bias, int M, int relu, int accumulate,
                                                   float* __restrict__ out) {
    __shared__ float As[64][65];
    __shared__ float Ws[64][65];
    const int tid = threadIdx.x, r0 = blockIdx.x * 64;
    const int r = tid >> 2, cg = (tid & 3) * 16;
    float acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
    for (int pass = 0; pass < 2; ++pass) {
        const float* Ap = pass == 0 ? A : A2;
        const float* Wp = pass == 0 ? W : W2;
        const int Kp = pass == 0 ? K : K2;
        if (!Ap) continue;
        for (int id = tid; id < 64 * Kp; id += 256) {
            const int rr = id / Kp, k = id % Kp;
            As[rr][k] = r0 + rr < N ? Ap[(size_t)(r0 + rr) * Kp + k] : 0.0f;
        }
        for (int id = tid; id < M * Kp; id += 256) {
            const int c = id / Kp, k = id % Kp;
            Ws[c][k] = wT ? Wp[(size_t)k * M + c] : Wp[(size_t)c * Kp + k];
        }
        __syncthreads();
        for (int k = 0; k < Kp; ++k) {
            const float av = As[r][k];
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (cg + j < M) acc[j] += av * Ws[cg + j][k];
        }
        __syncthreads();
    }
    const int row = r0 + r;
    if (row >= N) return;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int c = cg + j;
        if (c >= M) continue;
        float v = acc[j] + (bias ? bias[c] : 0.0f);
        if (relu) v = v > 0.0f ? v : 0.0f;
        float* o = out + (size_t)row * M + c;
        *o = accumulate ? *o + v : v;
    }
}

// Weight gradient partials: partial[b][c K + k] = sum over the rows i of
// block b of dz[i][c] A[i][k] (and, withBias, partial[b][M K + c] =
// sum dz[i][c]); rowsPerBlock rows per workgroup in chunks of 32 staged in
// LDS.  The partials are summed in a fixed order by k_gnn_sum.
constexpr int GNN_WG_ROWS = 256;
__global__ __launch_bounds__(256) void k_gnn_wgrad(int N, const float* __restrict__ dz, int M,
                                                   const float* __restrict__ A, int K, int withBias,
                                                   float* __restrict__ partial) {
    __shared__ float Ds[32][65];
    __shared__ float As[32][65];
    const int tid = threadIdx.x, b = blockIdx.x;
    const int stride = M * K + (withBias ? M : 0);
    const int i0 = b * GNN_WG_ROWS, i1 = min(N, i0 + GNN_WG_ROWS);
    float acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0f;
    float bacc = 0.0f;
    for (int c0 = i0; c0 < i1; c0 += 32) {
        for (int id = tid; id < 32 * M; id += 256) {
            const int rr = id / M, c = id % M;
            Ds[rr][c] = c0 + rr < i1 ? dz[(size_t)(c0 + rr) * M + c] : 0.0f;
        }
        for (int id = tid; id < 32 * K; id += 256) {
            const int rr = id / K, k = id % K;
            As[rr][k] = c0 + rr < i1 ? A[(size_t)(c0 + rr) * K + k] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int o = tid + 256 * q;
            if (o < M * K) {
                const int c = o / K, k = o % K;
                float s = 0.0f;
                for (int rr = 0; rr < 32; ++rr) s += Ds[rr][c] * As[rr][k];
                acc[q] += s;
            }
        }
        if (withBias && tid < M)
            for (int rr = 0; rr < 32; ++rr) bacc += Ds[rr][tid];
        __syncthreads();
    }
    float* pb = partial + (size_t)b * stride;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int o = tid + 256 * q;
        if (o < M * K) pb[o] = acc[q];
    }
    if (withBias && tid < M) pb[M * K + tid] = bacc;
}

// out[j] = sum over blocks b (in order) of partial[b stride + j], j < len.
__global__ void k_gnn_sum(const float* __restrict__ partial, int nblk, int stride, int len, float* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= len) return;
    float s = 0.0f;
    for (int b = 0; b < nblk; ++b) s += partial[(size_t)b * stride + j];
    out[j] = s;
}

// out[i][f] = (accumulate ? out[i][f] : 0) + rs_i sum_{j in N(i)} cs_j h[j][f]
// (rs, cs null: 1), H <= 64: one wave per node, lane = feature.
__global__ __launch_bounds__(256) void k_gnn_aggregate(int N, int H, const float* __restrict__ h,
                                                       const int* __restrict__ range, const int* __restrict__ adj,
                                                       const float* __restrict__ rowScale,
                                                       const float* __restrict__ colScale, int accumulate,
                                                       float* __restrict__ out) {
    const int node = blockIdx.x * 4 + (threadIdx.x >> 6), f = threadIdx.x & 63;
    if (node >= N || f >= H) return;
    float s = 0.0f;
    for (int q = range[node]; q < range[node + 1]; ++q) {
        const int j = adj[q];
        s += (colScale ? colScale[j] : 1.0f) * h[(size_t)j * H + f];
    }
    if (rowScale) s *= rowScale[node];
    float* o = out + (size_t)node * H + f;
    *o = accumulate ? *o + s : s;
}

// dz = dh where the relu output h is positive, else 0.
__global__ void k_gnn_relu_grad(int64_t n, const float* __restrict__ dh, const float* __restrict__ h,
                                float* __restrict__ dz) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dz[i] = h[i] > 0.0f ? dh[i] : 0.0f;
}

// Head, one thread per node: y = W_out h + b_out, Y = sym(y),
// B = G (I + Y) G^T (FP64), stored row-major in Bout (N x 9).
__global__ __launch_bounds__(256) void k_gnn_head(int N, int H, const float* __restrict__ h,
                                                  const float* __restrict__ Wout, const float* __restrict__ bout,
                                                  const double* __restrict__ G, double* __restrict__ Bout) {
    __shared__ float Wsh[GNN_OUT * GNN_HMAX + GNN_OUT];
    const int tid = threadIdx.x;
    for (int id = tid; id < GNN_OUT * H; id += 256) Wsh[id] = Wout[id];
    if (tid < GNN_OUT) Wsh[GNN_OUT * H + tid] = bout[tid];
    __syncthreads();
    const int i = blockIdx.x * 256 + tid;
    if (i >= N) return;
    float y[GNN_OUT];
#pragma unroll
    for (int c = 0; c < GNN_OUT; ++c) {
        float s = Wsh[GNN_OUT * H + c];
        for (int k = 0; k < H; ++k) s += Wsh[c * H + k] * h[(size_t)i * H + k];
        y[c] = s;
    }
    const double Mm[3][3] = {{1.0 + y[0], y[3], y[4]}, {y[3], 1.0 + y[1], y[5]}, {y[4], y[5], 1.0 + y[2]}};
    double g[3][3], T[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) g[a][b] = G[(size_t)i * 9 + 3 * a + b];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) T[a][b] = g[a][0] * Mm[0][b] + g[a][1] * Mm[1][b] + g[a][2] * Mm[2][b];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            Bout[(size_t)i * 9 + 3 * a + b] = T[a][0] * g[b][0] + T[a][1] * g[b][1] + T[a][2] * g[b][2];
}

// q = B r node by node (every DOF belongs to a node of the brick mesh).
__global__ void k_gnn_block_set(int N, const int* __restrict__ nodeDofs, const double* __restrict__ B,
                                const double* __restrict__ r, double* __restrict__ q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int d[3];
    double rv[3];
    for (int a = 0; a < 3; ++a) {
        d[a] = nodeDofs[3 * i + a];
        rv[a] = d[a] >= 0 ? r[d[a]] : 0.0;
    }
    for (int a = 0; a < 3; ++a)
        if (d[a] >= 0) {
            const double* b = B + (size_t)i * 9 + 3 * a;
            q[d[a]] = b[0] * rv[0] + b[1] * rv[1] + b[2] * rv[2];
        }
}

// One pair's share of the energy-norm loss, with q = B r and w = A q:
// lossNode_i += -2 r_i . q_i + q_i . w_i and gB_i += -2 (r - w)_i r_i^T
// (d/dB_i of -2 r^T B r + (B r)^T A (B r), A symmetric).
__global__ void k_gnn_energy_acc(int N, const int* __restrict__ nodeDofs, const double* __restrict__ r,
                                 const double* __restrict__ q, const double* __restrict__ w, double* __restrict__ gB,
                                 double* __restrict__ lossNode) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double rv[3], qv[3], wv[3];
    for (int a = 0; a < 3; ++a) {
        const int d = nodeDofs[3 * i + a];
        rv[a] = d >= 0 ? r[d] : 0.0;
        qv[a] = d >= 0 ? q[d] : 0.0;
        wv[a] = d >= 0 ? w[d] : 0.0;
    }
    double l = 0.0;
    for (int a = 0; a < 3; ++a) {
        l += -2.0 * rv[a] * qv[a] + qv[a] * wv[a];
        for (int b = 0; b < 3; ++b) gB[(size_t)i * 9 + 3 * a + b] += -2.0 * (rv[a] - wv[a]) * rv[b];
    }
    lossNode[i] += l;
}

// dl/dy from dl/dB = scale gB: dl/dM = G^T (dl/dB) G, M = I + Y symmetric.
__global__ void k_gnn_head_bwd(int N, const double* __restrict__ G, const double* __restrict__ gB, double scale,
                               float* __restrict__ dy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double g[3][3], t2[3][3], gM[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) g[a][b] = G[(size_t)i * 9 + 3 * a + b];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double v = 0.0;
            for (int c = 0; c < 3; ++c) v += gB[(size_t)i * 9 + 3 * a + c] * g[c][b];
            t2[a][b] = scale * v;
        }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) gM[a][b] = g[0][a] * t2[0][b] + g[1][a] * t2[1][b] + g[2][a] * t2[2][b];
    float* d = dy + (size_t)i * GNN_OUT;
    d[0] = (float)gM[0][0];
    d[1] = (float)gM[1][1];
    d[2] = (float)gM[2][2];
    d[3] = (float)(gM[0][1] + gM[1][0]);
    d[4] = (float)(gM[0][2] + gM[2][0]);
    d[5] = (float)(gM[1][2] + gM[2][1]);
}

// part[b] = sum of x over block b's 256 entries (fixed tree order).
__global__ __launch_bounds__(256) void k_gnn_block_sum(int N, const double* __restrict__ x, double* __restrict__ part) {
    __shared__ double red[256];
    const int tid = threadIdx.x, i = blockIdx.x * 256 + tid;
    red[tid] = i < N ? x[i] : 0.0;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    if (tid == 0) part[blockIdx.x] = red[0];
}

// *acc += (eTr + sum of the block partials, in order) / eTr (one thread).
__global__ void k_gnn_sum_loss(const double* __restrict__ part, int nblk, double eTr, double* __restrict__ acc) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[b];
    *acc += (eTr + s) / eTr;
}

// Adam (Kingma & Ba) with bias corrections bc1 = 1 - b1^t, bc2 = 1 - b2^t.
__global__ void k_gnn_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    }
}

__global__ void k_gnn_fill_f(float* __restrict__ p, int64_t n, float v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// Features, statistics, the block stage and the RHS generator.

__global__ void k_gnn_features(int N, const int* __restrict__ nodeDofs, const int* __restrict__ dofNode,
                               const double* __restrict__ val, const int* __restrict__ cols,
                               const int* __restrict__ range, const int* __restrict__ ctRange,
                               const int* __restrict__ ctPos, const int* __restrict__ ctRow,
                               const double* __restrict__ cmat, float* __restrict__ X, double* __restrict__ G) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int d[3];
    for (int a = 0; a < 3; ++a) d[a] = nodeDofs[3 * i + a];
    double D[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double coup = 0.0;
    bool constrained = false;
    for (int a = 0; a < 3; ++a) {
        if (d[a] < 0) continue;
        for (int q = range[d[a]]; q < range[d[a] + 1]; ++q) {
            const int c = cols[q] - 1;
            if (c < 0) continue;
            const double v = val[q];
            if (dofNode[c] == i) {
                for (int b = 0; b < 3; ++b)
                    if (c == d[b]) D[a][b] += v;
            } else {
                coup += v * v;
            }
        }
        if (ctRange) {
            for (int k = ctRange[d[a]]; k < ctRange[d[a] + 1]; ++k) {
                constrained = true;
                const int rr = ctRow[k];
                const double ca = cmat[ctPos[k]];
                for (int b = 0; b < 3; ++b) {
                    if (d[b] < 0) continue;
                    for (int k2 = ctRange[d[b]]; k2 < ctRange[d[b] + 1]; ++k2)
                        if (ctRow[k2] == rr) D[a][b] += ca * cmat[ctPos[k2]];
                }
            }
        }
    }
    for (int a = 0; a < 3; ++a)
        if (d[a] < 0) D[a][a] = 1.0;
    const double s = (D[0][0] + D[1][1] + D[2][2]) / 3.0;
    double fro = 0.0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) fro += D[a][b] * D[a][b];
    float* x = X + (size_t)i * GNN_F;
    x[0] = (float)(D[0][0] / s);
    x[1] = (float)(D[1][1] / s);
    x[2] = (float)(D[2][2] / s);
    x[3] = (float)(D[0][1] / s);
    x[4] = (float)(D[0][2] / s);
    x[5] = (float)(D[1][2] / s);
    x[6] = (float)log10(s);
    x[7] = constrained ? 1.0f : 0.0f;
    x[8] = (float)(sqrt(coup / fro) / 8.0);
    // Cholesky D = L L^T and G = L^-T
    const double L00 = sqrt(D[0][0]);
    const double L10 = D[1][0] / L00, L20 = D[2][0] / L00;
    const double L11 = sqrt(D[1][1] - L10 * L10);
    const double L21 = (D[2][1] - L20 * L10) / L11;
    const double L22 = sqrt(D[2][2] - L20 * L20 - L21 * L21);
    const double I00 = 1.0 / L00, I11 = 1.0 / L11, I22 = 1.0 / L22;
    const double I10 = -L10 * I00 / L11;
    const double I21 = -L21 * I11 / L22;
    const double I20 = -(L20 * I00 + L21 * I10) / L22;
    double* g = G + (size_t)i * 9;
    g[0] = I00;
    g[1] = I10;
    g[2] = I20;
    g[3] = 0.0;
    g[4] = I11;
    g[5] = I21;
    g[6] = 0.0;
    g[7] = 0.0;
    g[8] = I22;
}

__global__ void k_gnn_center_log(int N, float* __restrict__ X, float mu) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float* x = X + (size_t)i * GNN_F;
    x[6] = x[7] > 0.5f ? 0.0f : fminf(4.0f, fmaxf(-4.0f, x[6] - mu));
}

__global__ void k_node_block_apply(int N, const int* __restrict__ nodeDofs, const double* __restrict__ B,
                                   const double* __restrict__ t, double* __restrict__ z, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int d[3];
    double tv[3];
    for (int a = 0; a < 3; ++a) {
        d[a] = nodeDofs[3 * i + a];
        tv[a] = d[a] >= 0 ? t[d[a]] : 0.0;
    }
    for (int a = 0; a < 3; ++a)
        if (d[a] >= 0) {
            const double* b = B + (size_t)i * 9 + 3 * a;
            z[d[a]] += b[0] * tv[0] + b[1] * tv[1] + b[2] * tv[2];
        }
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Box-Muller on two counter-based uniforms per entry.
__global__ void k_normal_fill(double* __restrict__ x, int64_t n, uint64_t seed) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t a = splitmix64(seed * 0x100000001B3ull + 2 * (uint64_t)i);
        const uint64_t b = splitmix64(seed * 0x100000001B3ull + 2 * (uint64_t)i + 1);
        const double u1 = ((a >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = (b >> 11) * (1.0 / 9007199254740992.0);
        x[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
}

// ---------------------------------------------------------------------------
// Host side.

int GnnGraph::build(const std::vector<int>& range0, const std::vector<int>& adj0, hipStream_t s) {
    N = (int)range0.size() - 1;
    hRange = range0;
    hAdj = adj0;
    std::vector<float> di(N);
    for (int i = 0; i < N; ++i) {
        const int deg = range0[i + 1] - range0[i];
        di[i] = deg > 0 ? 1.0f / (float)deg : 0.0f;
    }
    ASTRO_TRY(range.upload(range0, s));
    ASTRO_TRY(adj.upload(adj0, s));
    ASTRO_TRY(degInv.upload(di, s));
    return ASTRO_OK;
}

static uint64_t host_splitmix(uint64_t& state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int GnnNet::init(int hidden, int layers, uint64_t seed, hipStream_t s) {
    if (hidden < 1 || hidden > GNN_HMAX || layers < 1 || layers > 8)
        return fail(ASTRO_ERR_ARG, "graph network: hidden must be in [1, 64] and layers in [1, 8]");
    H = hidden;
    L = layers;
    size_t o = 0;
    offWin = o;
    o += (size_t)H * GNN_F;
    offBin = o;
    o += H;
    offWs.assign(L + 1, 0);
    offWn.assign(L + 1, 0);
    offB.assign(L + 1, 0);
    for (int l = 1; l <= L; ++l) {
        offWs[l] = o;
        o += (size_t)H * H;
        offWn[l] = o;
        o += (size_t)H * H;
        offB[l] = o;
        o += H;
    }
    offWout = o;
    o += (size_t)GNN_OUT * H;
    offBout = o;
    o += GNN_OUT;
    nParams = o;
    // Glorot-uniform weights, zero biases; the head starts at a tenth of that
    // scale with Y = -I/2 from its bias, so the initial blocks are close to
    // half the diagonal-block inverses: damped block Jacobi, which does not
    // amplify the stiff modes that the undamped form does for 3-D elasticity
    std::vector<float> p(nParams, 0.0f);
    uint64_t st = seed;
    auto fillW = [&](size_t off, int fanOut, int fanIn, double gain) {
        const double lim = gain * std::sqrt(6.0 / (fanIn + fanOut));
        for (size_t k = 0; k < (size_t)fanOut * fanIn; ++k) {
            const double u = (host_splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            p[off + k] = (float)((2.0 * u - 1.0) * lim);
        }
    };
    fillW(offWin, H, GNN_F, 1.0);
    for (int l = 1; l <= L; ++l) {
        fillW(offWs[l], H, H, 1.0);
        fillW(offWn[l], H, H, 1.0);
    }
    fillW(offWout, GNN_OUT, H, 0.1);
    for (int d = 0; d < 3; ++d) p[offBout + d] = -0.5f;
    ASTRO_TRY(params.upload(p, s));
    ASTRO_TRY(grad.alloc(nParams));
    ASTRO_TRY(m.alloc(nParams));
    ASTRO_TRY(v.alloc(nParams));
    hipLaunchKernelGGL(k_gnn_fill_f, dim3(grid_for((int64_t)nParams, 256, 1024)), dim3(256), 0, s, m.p,
                       (int64_t)nParams, 0.0f);
    hipLaunchKernelGGL(k_gnn_fill_f, dim3(grid_for((int64_t)nParams, 256, 1024)), dim3(256), 0, s, v.p,
                       (int64_t)nParams, 0.0f);
    adamT = 0;
    ASTRO_TRY(lossAcc.alloc(1));
    HIP_TRY(hipGetLastError());
    return ASTRO_OK;
}

int GnnNet::reserve(int N, hipStream_t s) {
    (void)s;
    if (N <= capN) return ASTRO_OK;
    capN = N;
    for (int l = 0; l <= L; ++l) ASTRO_TRY(h[l].alloc((size_t)N * H));
    for (int l = 1; l <= L; ++l) ASTRO_TRY(a[l].alloc((size_t)N * H));
    ASTRO_TRY(dh.alloc((size_t)N * H));
    ASTRO_TRY(dh2.alloc((size_t)N * H));
    ASTRO_TRY(dz.alloc((size_t)N * H));
    ASTRO_TRY(u.alloc((size_t)N * H));
    ASTRO_TRY(dy.alloc((size_t)N * GNN_OUT));
    const int nblk = (N + GNN_WG_ROWS - 1) / GNN_WG_ROWS;
    ASTRO_TRY(partial.alloc((size_t)nblk * (GNN_HMAX * GNN_HMAX + GNN_HMAX)));
    ASTRO_TRY(lossPart.alloc((size_t)(N + 255) / 256));
    ASTRO_TRY(B.alloc((size_t)N * 9));
    ASTRO_TRY(gB.alloc((size_t)N * 9));
    ASTRO_TRY(lossNode.alloc((size_t)N));
    return ASTRO_OK;
}

static void dense(hipStream_t s, int N, const float* A, int K, const float* W, const float* A2, int K2,
                  const float* W2, int wT, const float* bias, int M, int relu, int accumulate, float* out) {
    hipLaunchKernelGGL(k_gnn_dense, dim3((N + 63) / 64), dim3(256), 0, s, N, A, K, W, A2, K2, W2, wT, bias, M, relu,
                       accumulate, out);
}

// dW (M x K) and, withBias, db (M) of dz (N x M) against A (N x K).
static void wgrad(hipStream_t s, int N, const float* dz, int M, const float* A, int K, float* dW, float* db,
                  float* partial) {
    const int nblk = (N + GNN_WG_ROWS - 1) / GNN_WG_ROWS;
    const int withBias = db != nullptr;
    const int stride = M * K + (withBias ? M : 0);
    hipLaunchKernelGGL(k_gnn_wgrad, dim3(nblk), dim3(256), 0, s, N, dz, M, A, K, withBias, partial);
    hipLaunchKernelGGL(k_gnn_sum, dim3((M * K + 255) / 256), dim3(256), 0, s, (const float*)partial, nblk, stride,
                       M * K, dW);
    if (withBias)
        hipLaunchKernelGGL(k_gnn_sum, dim3(1), dim3(256), 0, s, (const float*)(partial + M * K), nblk, stride, M, db);
}

int GnnNet::forward(const GnnGraph& g, const float* X, const double* G, double* Bout, hipStream_t s) {
    const int N = g.N;
    ASTRO_TRY(reserve(N, s));
    const float* P = params.p;
    dense(s, N, X, GNN_F, P + offWin, nullptr, 0, nullptr, 0, P + offBin, H, 1, 0, h[0].p);
    for (int l = 1; l <= L; ++l) {
        hipLaunchKernelGGL(k_gnn_aggregate, dim3((N + 3) / 4), dim3(256), 0, s, N, H, (const float*)h[l - 1].p,
                           (const int*)g.range.p, (const int*)g.adj.p, (const float*)g.degInv.p, (const float*)nullptr,
                           0, a[l].p);
        dense(s, N, h[l - 1].p, H, P + offWs[l], a[l].p, H, P + offWn[l], 0, P + offB[l], H, 1, 0, h[l].p);
    }
    if (Bout)
        hipLaunchKernelGGL(k_gnn_head, dim3((N + 255) / 256), dim3(256), 0, s, N, H, (const float*)h[L].p,
                           P + offWout, P + offBout, G, Bout);
    HIP_TRY(hipGetLastError());
    return ASTRO_OK;
}

int GnnNet::loss_grad(const GnnField& f, double* lossDev, hipStream_t s) {
    const GnnGraph& g = f.graph;
    const int N = g.N, n = f.n;
    ASTRO_TRY(reserve(N, s));
    if (q.n < (size_t)n) {
        ASTRO_TRY(q.alloc(n));
        ASTRO_TRY(w.alloc(n));
    }
    ASTRO_TRY(forward(g, f.X.p, f.G.p, B.p, s));
    const float* P = params.p;
    float* Gr = grad.p;
    const int nb = (N + 255) / 256;
    // the energy-norm loss over the field's pairs and its gradient in B
    hipLaunchKernelGGL(k_fill, dim3(grid_for((int64_t)N * 9, 256, 4096)), dim3(256), 0, s, gB.p, (int64_t)N * 9, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(grid_for(N, 256, 4096)), dim3(256), 0, s, lossNode.p, (int64_t)N, 0.0);
    for (int k = 0; k < f.samples; ++k) {
        const double* r = f.R1.p + (size_t)k * n;
        hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, q.p, (int64_t)n, 0.0);
        hipLaunchKernelGGL(k_gnn_block_set, dim3(grid_for(N, 256)), dim3(256), 0, s, N, (const int*)f.nodeDofs.p,
                           (const double*)B.p, r, q.p);
        hipLaunchKernelGGL(k_spmv, dim3(grid_for((int64_t)n * 32, 256, 1 << 30)), dim3(256), 0, s,
                           (const double*)f.val.p, (const int*)f.cols.p, (const int*)f.range.p, n, (const double*)q.p,
                           w.p, (const int*)nullptr);
        if (f.cdim > 0)
            hipLaunchKernelGGL(k_mpc_apply, dim3(grid_for(f.cdim, 256)), dim3(256), 0, s, (const double*)f.cmat.p,
                               (const int*)f.ccols.p, (const int*)f.crange.p, f.cdim, (const double*)q.p, w.p, (const int*)nullptr);
        hipLaunchKernelGGL(k_gnn_energy_acc, dim3(grid_for(N, 256)), dim3(256), 0, s, N, (const int*)f.nodeDofs.p, r,
                           (const double*)q.p, (const double*)w.p, gB.p, lossNode.p);
    }
    hipLaunchKernelGGL(k_gnn_block_sum, dim3(nb), dim3(256), 0, s, N, (const double*)lossNode.p, lossPart.p);
    hipLaunchKernelGGL(k_gnn_sum_loss, dim3(1), dim3(64), 0, s, (const double*)lossPart.p, nb, f.eTr, lossDev);
    hipLaunchKernelGGL(k_gnn_head_bwd, dim3(grid_for(N, 256)), dim3(256), 0, s, N, (const double*)f.G.p,
                       (const double*)gB.p, 1.0 / f.eTr, dy.p);
    // head
    wgrad(s, N, dy.p, GNN_OUT, h[L].p, H, Gr + offWout, Gr + offBout, partial.p);
    dense(s, N, dy.p, GNN_OUT, P + offWout, nullptr, 0, nullptr, 1, nullptr, H, 0, 0, dh.p);
    // message-passing layers, last to first
    for (int l = L; l >= 1; --l) {
        hipLaunchKernelGGL(k_gnn_relu_grad, dim3(grid_for((int64_t)N * H, 256, 4096)), dim3(256), 0, s,
                           (int64_t)N * H, (const float*)dh.p, (const float*)h[l].p, dz.p);
        wgrad(s, N, dz.p, H, h[l - 1].p, H, Gr + offWs[l], Gr + offB[l], partial.p);
        wgrad(s, N, dz.p, H, a[l].p, H, Gr + offWn[l], nullptr, partial.p);
        // dh_{l-1} = dz W_s + agg^T(dz W_n), agg^T(u)_i = sum_{j in N(i)} u_j / deg_j
        dense(s, N, dz.p, H, P + offWs[l], nullptr, 0, nullptr, 1, nullptr, H, 0, 0, dh2.p);
        dense(s, N, dz.p, H, P + offWn[l], nullptr, 0, nullptr, 1, nullptr, H, 0, 0, u.p);
        hipLaunchKernelGGL(k_gnn_aggregate, dim3((N + 3) / 4), dim3(256), 0, s, N, H, (const float*)u.p,
                           (const int*)g.range.p, (const int*)g.adj.p, (const float*)nullptr,
                           (const float*)g.degInv.p, 1, dh2.p);
        std::swap(dh.p, dh2.p);
    }
    hipLaunchKernelGGL(k_gnn_relu_grad, dim3(grid_for((int64_t)N * H, 256, 4096)), dim3(256), 0, s, (int64_t)N * H,
                       (const float*)dh.p, (const float*)h[0].p, dz.p);
    wgrad(s, N, dz.p, H, f.X.p, GNN_F, Gr + offWin, Gr + offBin, partial.p);
    HIP_TRY(hipGetLastError());
    return ASTRO_OK;
}

int GnnNet::adam(double lr, hipStream_t s) {
    ++adamT;
    const double b1 = 0.9, b2 = 0.999;
    const float bc1 = (float)(1.0 - std::pow(b1, (double)adamT)), bc2 = (float)(1.0 - std::pow(b2, (double)adamT));
    hipLaunchKernelGGL(k_gnn_adam, dim3(grid_for((int64_t)nParams, 256, 1024)), dim3(256), 0, s, (int64_t)nParams,
                       params.p, (const float*)grad.p, m.p, v.p, (float)lr, (float)b1, (float)b2, 1e-8f, bc1, bc2);
    HIP_TRY(hipGetLastError());
    return ASTRO_OK;
}

}  // namespace astro

using namespace astro;

astro_gnn::~astro_gnn() {
    if (s) (void)hipStreamDestroy(s);
}

extern "C" {

int astro_gnn_create(int hidden, int layers, uint64_t seed, astro_gnn** out) {
    if (!out) return fail(ASTRO_ERR_ARG, "null argument");
    *out = nullptr;
    std::unique_ptr<astro_gnn> g(new astro_gnn);
    HIP_TRY(hipStreamCreateWithFlags(&g->s, hipStreamNonBlocking));
    ASTRO_TRY(g->net.init(hidden, layers, seed, g->s));
    HIP_TRY(hipStreamSynchronize(g->s));
    *out = g.release();
    return ASTRO_OK;
}

int astro_gnn_destroy(astro_gnn* g) {
    if (g) {
        (void)hipStreamSynchronize(g->s);
        delete g;
    }
    return ASTRO_OK;
}

int64_t astro_gnn_num_params(const astro_gnn* g) { return g ? (int64_t)g->net.nParams : -1; }

int astro_gnn_num_fields(const astro_gnn* g) { return g ? (int)g->fields.size() : -1; }

int astro_gnn_get_params(astro_gnn* g, float* host, int64_t count) {
    if (!g || !host) return fail(ASTRO_ERR_ARG, "null argument");
    if (count != (int64_t)g->net.nParams) return fail(ASTRO_ERR_ARG, "parameter count mismatch");
    HIP_TRY(hipMemcpyAsync(host, g->net.params.p, sizeof(float) * count, hipMemcpyDeviceToHost, g->s));
    HIP_TRY(hipStreamSynchronize(g->s));
    return ASTRO_OK;
}

int astro_gnn_set_params(astro_gnn* g, const float* host, int64_t count) {
    if (!g || !host) return fail(ASTRO_ERR_ARG, "null argument");
    if (count != (int64_t)g->net.nParams) return fail(ASTRO_ERR_ARG, "parameter count mismatch");
    HIP_TRY(hipMemcpyAsync(g->net.params.p, host, sizeof(float) * count, hipMemcpyHostToDevice, g->s));
    HIP_TRY(hipStreamSynchronize(g->s));
    return ASTRO_OK;
}

int astro_gnn_loss_grad(astro_gnn* g, int field, double* loss, float* gradHost) {
    if (!g || !loss) return fail(ASTRO_ERR_ARG, "null argument");
    if (field < 0 || field >= (int)g->fields.size()) return fail(ASTRO_ERR_ARG, "field index out of range");
    GnnNet& net = g->net;
    const double zero = 0.0;
    HIP_TRY(hipMemcpyAsync(net.lossAcc.p, &zero, sizeof(double), hipMemcpyHostToDevice, g->s));
    ASTRO_TRY(net.loss_grad(*g->fields[field], net.lossAcc.p, g->s));
    HIP_TRY(hipMemcpyAsync(loss, net.lossAcc.p, sizeof(double), hipMemcpyDeviceToHost, g->s));
    if (gradHost)
        HIP_TRY(hipMemcpyAsync(gradHost, net.grad.p, sizeof(float) * net.nParams, hipMemcpyDeviceToHost, g->s));
    HIP_TRY(hipStreamSynchronize(g->s));
    return ASTRO_OK;
}

int astro_gnn_train(astro_gnn* g, int epochs, double lr, uint64_t seed, double* lossPerEpoch) {
    if (!g) return fail(ASTRO_ERR_ARG, "null argument");
    const int nf = (int)g->fields.size();
    if (nf < 1) return fail(ASTRO_ERR_ARG, "no training fields collected");
    if (epochs < 0 || !(lr > 0.0)) return fail(ASTRO_ERR_ARG, "bad epoch count or learning rate");
    GnnNet& net = g->net;
    DBuf<double> acc;
    ASTRO_TRY(acc.alloc((size_t)std::max(1, epochs)));
    std::vector<double> zeros(std::max(1, epochs), 0.0);
    HIP_TRY(hipMemcpyAsync(acc.p, zeros.data(), sizeof(double) * zeros.size(), hipMemcpyHostToDevice, g->s));
    std::vector<int> order(nf);
    std::iota(order.begin(), order.end(), 0);
    uint64_t st = seed;
    for (int e = 0; e < epochs; ++e) {
        // Fisher-Yates shuffle of the field order, deterministic in seed
        for (int i = nf - 1; i > 0; --i) std::swap(order[i], order[host_splitmix(st) % (uint64_t)(i + 1)]);
        for (int k = 0; k < nf; ++k) {
            ASTRO_TRY(net.loss_grad(*g->fields[order[k]], acc.p + e, g->s));
            ASTRO_TRY(net.adam(lr, g->s));
        }
    }
    HIP_TRY(hipStreamSynchronize(g->s));
    if (lossPerEpoch && epochs > 0) {
        HIP_TRY(hipMemcpy(lossPerEpoch, acc.p, sizeof(double) * epochs, hipMemcpyDeviceToHost));
        for (int e = 0; e < epochs; ++e) lossPerEpoch[e] /= nf;
    }
    return ASTRO_OK;
}

int astro_gnn_field_size(const astro_gnn* g, int field, int32_t* N, int32_t* nAdj, int32_t* nDofs,
                         int32_t* samples) {
    if (!g || !N || !nAdj || !nDofs || !samples) return fail(ASTRO_ERR_ARG, "null argument");
    if (field < 0 || field >= (int)g->fields.size()) return fail(ASTRO_ERR_ARG, "field index out of range");
    const GnnField& f = *g->fields[field];
    *N = f.graph.N;
    *nAdj = (int32_t)f.graph.hAdj.size();
    *nDofs = f.n;
    *samples = f.samples;
    return ASTRO_OK;
}

int astro_gnn_get_field(astro_gnn* g, int field, int32_t* range, int32_t* adj, float* X, double* G, double* R1,
                        double* eTr) {
    if (!g) return fail(ASTRO_ERR_ARG, "null argument");
    if (field < 0 || field >= (int)g->fields.size()) return fail(ASTRO_ERR_ARG, "field index out of range");
    const GnnField& f = *g->fields[field];
    const int N = f.graph.N;
    if (range) std::memcpy(range, f.graph.hRange.data(), sizeof(int) * (N + 1));
    if (adj) std::memcpy(adj, f.graph.hAdj.data(), sizeof(int) * f.graph.hAdj.size());
    if (X) HIP_TRY(hipMemcpyAsync(X, f.X.p, sizeof(float) * N * GNN_F, hipMemcpyDeviceToHost, g->s));
    if (G) HIP_TRY(hipMemcpyAsync(G, f.G.p, sizeof(double) * N * 9, hipMemcpyDeviceToHost, g->s));
    if (R1)
        HIP_TRY(hipMemcpyAsync(R1, f.R1.p, sizeof(double) * (size_t)f.n * f.samples, hipMemcpyDeviceToHost, g->s));
    if (eTr) *eTr = f.eTr;
    HIP_TRY(hipStreamSynchronize(g->s));
    return ASTRO_OK;
}

int astro_gnn_field_blocks(astro_gnn* g, int field, double* B) {
    if (!g || !B) return fail(ASTRO_ERR_ARG, "null argument");
    if (field < 0 || field >= (int)g->fields.size()) return fail(ASTRO_ERR_ARG, "field index out of range");
    const GnnField& f = *g->fields[field];
    DBuf<double> Bd;
    ASTRO_TRY(Bd.alloc((size_t)f.graph.N * 9));
    ASTRO_TRY(g->net.forward(f.graph, f.X.p, f.G.p, Bd.p, g->s));
    HIP_TRY(hipMemcpyAsync(B, Bd.p, sizeof(double) * f.graph.N * 9, hipMemcpyDeviceToHost, g->s));
    HIP_TRY(hipStreamSynchronize(g->s));
    return ASTRO_OK;
}

}  // extern "C"
