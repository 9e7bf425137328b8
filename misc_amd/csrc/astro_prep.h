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
    // device helpers (all-brick8 or all-shell4 models)
    bool shellMesh = false;       // every element is a shell4 (4 nodes, 6 DOFs per node)
    std::vector<uint8_t> elPos;   // 192 per element: slot of column (node b, comp d) in node a's row list,
                                  // at (a*8+b)*3+d for bricks and (a*4+b)*6+d for shells
    std::vector<uint8_t> thPos;   // 64 per element: slot of node b in node a's thermal row
    std::vector<int> colorPerm;   // elements sorted by colour
    std::vector<int> colorStart;  // numColors + 1
};

// solverMaxBW rule of the *solve command (Working/AStrO_runJob.f90:107-114):
// smallest b with b^3 >= 27*N*N.  literal = true forms 27*N*N in 32-bit
// two's-complement arithmetic as the reference's default INTEGER does.
int solver_max_bw(int numNodes, bool literal);

// Full analysis.  maxBW <= 0 selects solver_max_bw(numNodes, literalMaxBW).
// coupledProfile widens every profile row to all components of the node and
// its neighbours (the reference keeps one component per row).
// Returns "" on success or an error message.
std::string analysis_prep(const ModelDesc& m, int maxBW, bool literalMaxBW, bool thermLiteralD8, Prep& out,
                          bool coupledProfile = false);

// Shell section data (astro_shell.cpp): frame rows = local axes from the two
// orientation vectors aligned to the element normal; alpha[i + 3 j].  Layup
// stiffness ABD (9 x 9 column-major) from nL layers of (E1, E2, nu12, G12,
// thickness, angle in radians) and the z offset.
void shell_frame(const double* orient6, const double* xyz34, double* alpha);
void shell_section_abd(int nL, const double* layers, double zOff, double* ABD);
// The same with the thickness of `layer` carrying the imaginary part hc
// (complex step); real and imaginary parts of ABD separately.
void shell_section_abd_cs(int nL, const double* layers, double zOff, int layer, double hc, double* ABDre,
                          double* ABDim);

}  // namespace astro
