// astro_batch.h -- the batched complex-step sensitivity kernel (astro_batch.hip):
// every design variable's perturbed system assembled, factored and solved as
// a complex128 instance, all instances in one launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace astro {

constexpr int CSB_THREADS = 256;

struct CsbArgs {
    int n = 0;                        // DOFs
    int W = 0;                        // LDS row stride of the tile (csb_stride)
    int maxRow = 0;                   // longest profile row
    int64_t total = 0;                // profile entries (per instance)
    const int* first = nullptr;       // node-major profile: first column of each row
    const int64_t* rowoff = nullptr;  // row offsets (n + 1)
    const int* inv = nullptr;         // old DOF (0-based) -> node-major index
    const double* profRe = nullptr;   // A_re in the profile
    const double* F = nullptr;        // loads, node-major
    const int* order = nullptr;       // instance of each workgroup (most work first)
    const int* instStart = nullptr;   // r_k: first row with an imaginary part
    const int* elStart = nullptr;     // elements of instance k: elList[elStart[k] .. elStart[k + 1])
    const int* elList = nullptr;
    const double* kim = nullptr;      // Im of each element's condensed complex tangent, 24 x 24 row-major
    const int* kimDof = nullptr;      // its 24 DOFs (1-based, old order)
    double* work = nullptr;           // instance profiles, complex interleaved, total each
    double* vec = nullptr;            // instance right-hand sides / corrections, complex interleaved, n each
    double* sol = nullptr;            // instance solutions, complex interleaved, n each
    int mode = 0;                     // 0: assemble, factor, solve F; 1: solve vec with the factor, sol += it
    double invH = 0.0;
    double* grad = nullptr;           // Im(F . u_k) / h
    double* compliance = nullptr;     // Re(F . u_k)
    int* status = nullptr;            // 1: a zero pivot
};

// The refinement residual r_k = F - A_k u_k of every instance (into vec).
struct CsbRes {
    int n = 0;
    const int* order = nullptr;  // node-major index -> old DOF (0-based)
    const int* inv = nullptr;
    const int* range = nullptr;  // CSR of K, old order, 1-based columns
    const int* cols = nullptr;
    const double* val = nullptr;
    const double* cmat = nullptr;  // constraint rows and their transpose over old DOFs
    const int* ccols = nullptr;
    const int* crange = nullptr;
    const int* ctRange = nullptr;
    const int* ctPos = nullptr;
    const int* ctRow = nullptr;
    const double* F = nullptr;
    const int* elStart = nullptr;
    const int* elList = nullptr;
    const double* kim = nullptr;
    const int* kimDof = nullptr;
    const double* sol = nullptr;
    double* ed = nullptr;  // the element terms per row, double-double (4 doubles), n each
    double* vec = nullptr;
};
hipError_t csb_residual(const CsbRes& R, int nInst, hipStream_t s);

int csb_num_tiles();
void csb_tile(int t, int* rb, int* cb);
int csb_stride(int t, int maxRow);
size_t csb_lds_bytes(int t, int maxRow);
hipError_t csb_launch(const CsbArgs& a, int tile, int nInst, hipStream_t s);
hipError_t csb_real_profile(const int* range, const int* cols, const double* val, int n, const double* cmat,
                            const int* ccols, const int* crange, int cd, const int* inv, const int* first,
                            const int64_t* rowoff, double* prof, int64_t total, hipStream_t s);

}  // namespace astro
