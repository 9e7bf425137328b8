// astro_gnn.h -- the learned block preconditioner: a graph network over the
// mesh's node graph that maps per-node features of the assembled matrix to a
// symmetric 3 x 3 block B_i per node; GMRES then applies the two-stage
// preconditioner z = P^-1 r, z += B (r - A z), with P the skyline LDL^T.
// Kernels and the training driver live in astro_gnn.hip; the context-facing
// entry points (feature extraction from a context's matrix, training-pair
// collection, attaching) in astro_abi.hip.
#pragma once
#include <memory>

#include "astro_internal.h"

namespace astro {

constexpr int GNN_F = 9;    // node input features
constexpr int GNN_OUT = 6;  // symmetric 3 x 3 correction Y
constexpr int GNN_HMAX = 64;

// Node graph of one mesh: symmetric adjacency without self loops (CSR).
struct GnnGraph {
    int N = 0;
    DBuf<int> range, adj;
    DBuf<float> degInv;  // 1 / degree (0 for an isolated node)
    std::vector<int> hRange, hAdj;
    int build(const std::vector<int>& range0, const std::vector<int>& adj0, hipStream_t s);
};

// One training field: the inputs of the network, its operator A (the CSR of
// K and the penalty rows C), and the solver-generated pairs: for each random
// b, x = A^-1 b, z0 = P^-1 b (the LDL^T stage), the residual r1 = b - A z0
// and the error e1 = x - z0 = A^-1 r1 it leaves.  The loss is the energy
// norm of the error the block stage leaves,
//   l = sum_s (e1 - B r1)^T A (e1 - B r1) / sum_s e1^T A e1
//     = sum_s (e1^T r1 - 2 r1^T q + q^T A q) / sum_s e1^T r1,   q = B r1,
// which needs r1 and sum e1^T r1 only.
struct GnnField {
    GnnGraph graph;
    int n = 0, cdim = 0;
    DBuf<float> X;                  // N x GNN_F
    DBuf<double> G;                 // N x 9: L_i^-T (row-major), D_i = L_i L_i^T the diagonal block
    DBuf<int> nodeDofs;             // N x 3
    DBuf<double> val, cmat;         // operator: CSR values, penalty rows
    DBuf<int> cols, range, ccols, crange;
    DBuf<double> R1;                // samples x n
    double eTr = 0.0;               // sum over samples of e1^T r1
    int samples = 0;
};

// Encoder (F -> H, relu), L message-passing layers
//   h_l = relu(W_s h_{l-1} + W_n mean_{j in N(i)} h_{l-1, j} + b),
// then the head y = W_out h_L + b_out (H -> 6), Y_i = sym(y), and
// B_i = G_i (I + Y_i) G_i^T.  Parameters are one FP32 vector.
struct GnnNet {
    int H = 64, L = 3;
    size_t nParams = 0;
    size_t offWin = 0, offBin = 0, offWout = 0, offBout = 0;
    std::vector<size_t> offWs, offWn, offB;
    DBuf<float> params, grad, m, v;
    int64_t adamT = 0;
    // workspace (grown to the largest graph seen)
    int capN = 0;
    DBuf<float> h[9], a[9];  // h[0..L], a[1..L] (L <= 8)
    DBuf<float> dh, dh2, dz, u, dy, partial;
    DBuf<double> lossPart, lossAcc, B, gB, lossNode, q, w;
    int init(int hidden, int layers, uint64_t seed, hipStream_t s);
    int reserve(int N, hipStream_t s);
    // forward pass; Bout (N x 9, may be null) receives the blocks
    int forward(const GnnGraph& g, const float* X, const double* G, double* Bout, hipStream_t s);
    // forward + loss + backward into grad (the loss is added to *lossDev)
    int loss_grad(const GnnField& f, double* lossDev, hipStream_t s);
    int adam(double lr, hipStream_t s);
};

// Features of node i from the assembled operator A = K + C^T C (CSR of K with
// the reference's 1-based columns, C by column as in the double-double
// residual): with D_i the 3 x 3 diagonal block and s_i = tr(D_i) / 3,
//   X = (D11, D22, D33, D12, D13, D23) / s, log10 s (centred later), a flag
//       for nodes with a constrained DOF, and |A(i, others)|_F / |D_i|_F / 8;
// G_i = L_i^-T of the Cholesky factor D_i = L_i L_i^T.  nodeDofs: 3 per
// node, 0-based (-1: none); dofNode: node of each DOF.
__global__ void k_gnn_features(int N, const int* nodeDofs, const int* dofNode, const double* val, const int* cols,
                               const int* range, const int* ctRange, const int* ctPos, const int* ctRow,
                               const double* cmat, float* X, double* G);
// X[:, 6] -= mu, clamped to [-4, 4], for the unconstrained nodes; 0 for the others.
__global__ void k_gnn_center_log(int N, float* X, float mu);

// z[dofs of i] += B_i t[dofs of i] (B: N x 9 row-major).
__global__ void k_node_block_apply(int N, const int* nodeDofs, const double* B, const double* t, double* z, const int* gate);
// Deterministic standard normal vector from (seed, index).
__global__ void k_normal_fill(double* x, int64_t n, uint64_t seed);

}  // namespace astro

// The handle behind astro_gnn* (include/astro_abi.h): the network, its
// training fields and the stream it works on.
struct astro_gnn {
    astro::GnnNet net;
    std::vector<std::unique_ptr<astro::GnnField>> fields;
    hipStream_t s = nullptr;
    ~astro_gnn();
};
