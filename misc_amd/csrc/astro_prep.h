// astro_prep.h -- host-side symbolic analysis of a model (integer book-keeping).
//
// Reproduces the reference's analysisPrep (Working/AStrO_bookKeeping.f90:794-1039)
// bit for bit: node adjacency in insertion order, level ordering, DOF numbering,
// CSR capacities and first-touch column order, skyline profile lengths and the
// constraint matrices.  It also derives what the device kernels need and the
// reference never stores: per-element scatter positions and an element colouring
// with no two elements of one colour sharing a node.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace astro {

struct ModelDesc {
    int numNodes = 0;
    int numEls = 0;
    const double* nodeList = nullptr;  // 3*numNodes, column-major
    const int* elementList = nullptr;  // 8*numEls, 1-based, 0 = unused
    const int* elementType = nullptr;  // numEls
    // constraint terms: equation, dof (1-7), coefficient, node or node set
    int numMpcTerms = 0;
    const int* mpcEqn = nullptr;
    const int* mpcDof = nullptr;
    const double* mpcCoef = nullptr;
    const int* mpcIsSet = nullptr;
    const int* mpcNodeStart = nullptr;  // numMpcTerms + 1 offsets into mpcNodes
    const int* mpcNodes = nullptr;
    int numMPC = 0;
    const double* mpcRHS = nullptr;
};

struct Prep {
    int numNodes = 0, numEls = 0, numSolidNds = 0, numSectionNds = 0;
    int elMatDim = 0, solverMaxBW = 0;
    std::vector<int> nCRange, nodalConn;           // adjacency with the reference's slot layout
    std::vector<int> currentRank, originalRank;    // 1-based ranks
    std::vector<int> nDofIndex;                    // 6*numNodes
    std::vector<int> elMatRange, elMatCols;        // CSR (0:n) ranges, 1-based columns
    std::vector<int> elMatLTRange;                 // profile (0:n) ranges
    std::vector<int> thermMatRange, thermMatCols, thermMatLTRange;
    bool thermLiteralD8 = true;
    int elMPCDim = 0, thermMPCDim = 0;
    std::vector<int> elMPCMatRange, elMPCMatCols, thermMPCMatRange, thermMPCMatCols;
    std::vector<double> elMPCMat, elMPCRHS, thermMPCMat, thermMPCRHS;
    // device helpers (brick8 models)
    std::vector<uint8_t> elPos;   // 192 per element: slot of column (node b, comp d) in node a's row list
    std::vector<uint8_t> thPos;   // 64 per element: slot of node b in node a's thermal row
    std::vector<int> colorPerm;   // elements sorted by colour
    std::vector<int> colorStart;  // numColors + 1
};

// solverMaxBW rule of the *solve command (Working/AStrO_runJob.f90:107-114):
// smallest b with b^3 >= 27*N*N.  literal = true forms 27*N*N in 32-bit
// two's-complement arithmetic as the reference's default INTEGER does.
int solver_max_bw(int numNodes, bool literal);

// Full analysis.  maxBW <= 0 selects solver_max_bw(numNodes, literalMaxBW).
// Returns "" on success or an error message.
std::string analysis_prep(const ModelDesc& m, int maxBW, bool literalMaxBW, bool thermLiteralD8, Prep& out);

}  // namespace astro
