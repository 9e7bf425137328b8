// astro_prep.cpp -- host-side symbolic analysis (see astro_prep.h).
#include "astro_prep.h"

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstring>

namespace astro {

namespace {

int ndof_of_type(int eT) { return (eT == 41 || eT == 3 || eT == 2) ? 6 : 3; }

// Level ordering: breadth-first from the first node of the DOF type with the
// fewest adjacency slots, neighbours in adjacency insertion order; when the
// queue runs dry the nearest unvisited node of the same type restarts it.
// Working/AStrO_bookKeeping.f90:233-331.  (Its block-split branch needs a
// block of numNodes + 1 nodes and never fires.)
void level_sort(const ModelDesc& m, Prep& p, int& nextAvail, int stopLength, int ndof,
                std::vector<char>& inserted) {
    const int N = m.numNodes;
    auto typeOf = [&](int nd) { return p.nDofIndex[6 * (size_t)(nd - 1)]; };
    auto nearest = [&](int from) {
        double best = 1e100;
        int st = 0;
        const double* f = m.nodeList + 3 * (size_t)(from - 1);
        for (int i = 1; i <= N; ++i) {
            if (inserted[i] || typeOf(i) != ndof) continue;
            const double* q = m.nodeList + 3 * (size_t)(i - 1);
            double d0 = q[0] - f[0], d1 = q[1] - f[1], d2 = q[2] - f[2];
            double d = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            if (d < best) {
                best = d;
                st = i;
            }
        }
        return st;
    };
    int minCon = 1000, stNd = 0;
    for (int i = 1; i <= N; ++i) {
        if (typeOf(i) != ndof) continue;
        int c = p.nCRange[i] - p.nCRange[i - 1];
        if (c < minCon) {
            minCon = c;
            stNd = i;
        }
    }
    p.originalRank[nextAvail] = stNd;
    int cur = stNd, curInd = nextAvail;
    ++nextAvail;
    inserted[stNd] = 1;
    while (nextAvail <= stopLength) {
        for (int s = p.nCRange[cur - 1]; s < p.nCRange[cur]; ++s) {
            int nb = p.nodalConn[s];
            if (nb != 0 && !inserted[nb]) {
                p.originalRank[nextAvail++] = nb;
                inserted[nb] = 1;
            }
        }
        ++curInd;
        if (curInd == nextAvail && nextAvail <= stopLength) {
            int st = nearest(p.originalRank[nextAvail - 1]);
            p.originalRank[nextAvail] = st;
            curInd = nextAvail;
            inserted[st] = 1;
            ++nextAvail;
        }
        cur = p.originalRank[curInd];
    }
}

// Constraint matrices, one row per expanded equation: an equation naming node
// sets gets as many rows as its largest set, a set of one node repeats that
// node on every row, a plain node id likewise.  Columns are DOF numbers
// (elastic) or node ranks (thermal).  Working/AStrO_bookKeeping.f90:474-612, 614-752.
void build_mpc(const ModelDesc& m, const Prep& p, bool thermal, int& dim, std::vector<int>& range,
               std::vector<int>& cols, std::vector<double>& mat, std::vector<double>& rhs) {
    std::vector<int> nEqns(m.numMPC + 1, 0), nTrms(m.numMPC + 1, 0);
    auto used = [&](int t) { return thermal ? m.mpcDof[t] == 7 : m.mpcDof[t] <= 6; };
    for (int t = 0; t < m.numMpcTerms; ++t) {
        if (!used(t)) continue;
        int e = m.mpcEqn[t];
        nTrms[e] += 1;
        int sz = m.mpcIsSet[t] ? m.mpcNodeStart[t + 1] - m.mpcNodeStart[t] : 1;
        nEqns[e] = std::max(nEqns[e], sz);
    }
    dim = 0;
    size_t size = 0;
    for (int e = 1; e <= m.numMPC; ++e) {
        dim += nEqns[e];
        size += (size_t)nEqns[e] * nTrms[e];
    }
    range.assign(dim + 1, 0);
    cols.assign(size, 0);
    mat.assign(size, 0.0);
    rhs.assign(dim, 0.0);
    for (int e = 2; e <= m.numMPC; ++e) nEqns[e] += nEqns[e - 1];
    for (int e = 1; e <= m.numMPC; ++e)
        for (int r = nEqns[e - 1] + 1; r <= nEqns[e]; ++r) range[r] = nTrms[e];
    for (int r = 2; r <= dim; ++r) range[r] += range[r - 1];
    auto colOf = [&](int node, int dof) {
        return thermal ? p.currentRank[node - 1] : p.nDofIndex[6 * (size_t)(node - 1) + dof - 1];
    };
    auto put = [&](int row, int col, double c) {
        for (int s = range[row - 1]; s < range[row]; ++s) {
            if (cols[s] == 0) {
                cols[s] = col;
                mat[s] = c;
                return;
            }
            if (cols[s] == col) return;
        }
    };
    for (int t = 0; t < m.numMpcTerms; ++t) {
        if (!used(t)) continue;
        int e = m.mpcEqn[t];
        int row = nEqns[e - 1] + 1;
        int a = m.mpcNodeStart[t], b = m.mpcNodeStart[t + 1];
        if (!m.mpcIsSet[t] || b - a == 1) {
            int col = colOf(m.mpcNodes[a], m.mpcDof[t]);
            for (; row <= nEqns[e]; ++row) {
                put(row, col, m.mpcCoef[t]);
                rhs[row - 1] = m.mpcRHS[e - 1];
            }
        } else {
            for (int k = a; k < b; ++k, ++row) {
                put(row, colOf(m.mpcNodes[k], m.mpcDof[t]), m.mpcCoef[t]);
                rhs[row - 1] = m.mpcRHS[e - 1];
            }
        }
    }
}

}  // namespace

int solver_max_bw(int numNodes, bool literal) {
    long long target;
    if (literal) {
        uint32_t v = 27u * (uint32_t)numNodes;
        v *= (uint32_t)numNodes;
        target = (int32_t)v;
    } else {
        target = 27LL * numNodes * numNodes;
    }
    long long b = 1;
    while (b * b * b < target) ++b;
    return (int)b;
}

std::string analysis_prep(const ModelDesc& m, int maxBW, bool literalMaxBW, bool thermLiteralD8, Prep& p,
                          bool coupledProfile) {
    const int N = m.numNodes, E = m.numEls;
    if (N <= 0 || E <= 0) return "empty model";
    for (int e = 0; e < E; ++e) {
        for (int a = 0; a < 8; ++a) {
            int nd = m.elementList[8 * (size_t)e + a];
            if (nd < 0 || nd > N) return "element node id out of range";
        }
    }
    p = Prep{};
    p.numNodes = N;
    p.numEls = E;
    p.thermLiteralD8 = thermLiteralD8;
    p.solverMaxBW = maxBW > 0 ? maxBW : solver_max_bw(N, literalMaxBW);
    p.nDofIndex.assign(6 * (size_t)N, 0);
    p.currentRank.assign(N, 0);
    p.originalRank.assign(N + 2, 0);
    // adjacency slots: one per (element, other node) pair, filled first-come
    // without duplicates, the rest left 0 (:817-882)
    p.nCRange.assign(N + 1, 0);
    for (int e = 0; e < E; ++e) {
        const int* el = m.elementList + 8 * (size_t)e;
        int nn = 0;
        for (int a = 0; a < 8; ++a) nn += el[a] != 0;
        int nd = ndof_of_type(m.elementType[e]);
        for (int a = 0; a < 8; ++a) {
            if (!el[a]) continue;
            int others = 0;
            for (int b = 0; b < 8; ++b) others += (el[b] != 0 && el[b] != el[a]);
            p.nCRange[el[a]] += others;
            p.nDofIndex[6 * (size_t)(el[a] - 1)] = nd;
        }
        (void)nn;
    }
    for (int i = 1; i <= N; ++i) p.nCRange[i] += p.nCRange[i - 1];
    p.nodalConn.assign(p.nCRange[N], 0);
    std::vector<int> fill(N + 1, 0);  // filled slots per node
    for (int e = 0; e < E; ++e) {
        const int* el = m.elementList + 8 * (size_t)e;
        for (int a = 0; a < 8; ++a) {
            int na = el[a];
            if (!na) continue;
            int base = p.nCRange[na - 1];
            for (int b = 0; b < 8; ++b) {
                int nb = el[b];
                if (!nb || nb == na) continue;
                bool have = false;
                for (int s = 0; s < fill[na]; ++s)
                    if (p.nodalConn[base + s] == nb) {
                        have = true;
                        break;
                    }
                if (!have) p.nodalConn[base + fill[na]++] = nb;
            }
        }
    }
    for (int i = 1; i <= N; ++i)
        if (p.nDofIndex[6 * (size_t)(i - 1)] == 3) ++p.numSolidNds;
    p.numSectionNds = N - p.numSolidNds;
    std::vector<char> inserted(N + 1, 0);
    int nextAvail = 1;
    if (p.numSolidNds > 0) level_sort(m, p, nextAvail, p.numSolidNds, 3, inserted);
    if (nextAvail <= N) level_sort(m, p, nextAvail, N, 6, inserted);
    for (int i = 1; i <= N; ++i) {
        int nd = p.originalRank[i];
        if (nd < 1 || nd > N) return "level ordering failed (node without elements?)";
        p.currentRank[nd - 1] = i;
    }
    for (int i = 1; i <= N; ++i) {
        int* d = &p.nDofIndex[6 * (size_t)(i - 1)];
        int r = p.currentRank[i - 1];
        if (d[0] == 3)
            for (int c = 0; c < 3; ++c) d[c] = r + c * p.numSolidNds;
        else if (d[0] == 6)
            for (int c = 0; c < 6; ++c) d[c] = r + c * p.numSectionNds + 2 * p.numSolidNds;
    }
    p.elMatDim = 3 * p.numSolidNds + 6 * p.numSectionNds;
    const int n = p.elMatDim;
    auto dofOf = [&](int node, int c) { return p.nDofIndex[6 * (size_t)(node - 1) + c]; };
    auto ndofOf = [&](int node) {
        int k = 0;
        for (int c = 0; c < 6; ++c) k += dofOf(node, c) != 0;
        return k;
    };
    // CSR capacity per row: own DOFs + DOFs of every adjacent node (:346-373)
    p.elMatRange.assign(n + 1, 0);
    for (int i = 1; i <= N; ++i) {
        int cap = ndofOf(i);
        for (int s = p.nCRange[i - 1]; s < p.nCRange[i]; ++s)
            if (p.nodalConn[s]) cap += ndofOf(p.nodalConn[s]);
        for (int c = 0; c < 6; ++c)
            if (dofOf(i, c)) p.elMatRange[dofOf(i, c)] = cap;
    }
    {
        int64_t acc = 0;
        for (int r = 1; r <= n; ++r) {
            acc += p.elMatRange[r];
            if (acc > INT32_MAX) return "elastic CSR capacity exceeds 2^31 - 1 entries";
            p.elMatRange[r] = (int)acc;
        }
    }
    // profile lengths: same-component lower-ranked neighbours within solverMaxBW
    // (:380-409); with coupledProfile every component of the node and of those
    // neighbours
    p.elMatLTRange.assign(n + 1, 0);
    for (int i = 1; i <= N; ++i)
        for (int c = 0; c < 6; ++c) {
            int r = dofOf(i, c);
            if (!r) continue;
            int len = 1;
            if (coupledProfile)  // the node's own lower components
                for (int c2 = 0; c2 < 6; ++c2) {
                    int q = dofOf(i, c2);
                    if (q != 0 && q <= r) len = std::max(len, r - q + 1);
                }
            for (int s = p.nCRange[i - 1]; s < p.nCRange[i]; ++s) {
                int nb = p.nodalConn[s];
                if (!nb) continue;
                int dr = p.currentRank[i - 1] - p.currentRank[nb - 1];
                if (dr < 0 || dr > p.solverMaxBW) continue;
                for (int c2 = 0; c2 < 6; ++c2) {
                    if (!coupledProfile && c2 != c) continue;
                    int q = dofOf(nb, c2);
                    if (q != 0 && q <= r) len = std::max(len, r - q + 1);
                }
            }
            p.elMatLTRange[r] = len;
        }
    // prefix sums in 64 bits: the reference keeps the ranges in default
    // INTEGER, so a profile past 2^31 - 1 entries cannot cross the boundary
    {
        int64_t acc = 0;
        for (int r = 1; r <= n; ++r) {
            acc += p.elMatLTRange[r];
            if (acc > INT32_MAX) return "elastic profile exceeds 2^31 - 1 entries";
            p.elMatLTRange[r] = (int)acc;
        }
    }
    // thermal capacity and profile, indexed by node rank (:413-472)
    p.thermMatRange.assign(N + 1, 0);
    p.thermMatLTRange.assign(N + 1, 0);
    for (int i = 1; i <= N; ++i) {
        int r = p.currentRank[i - 1];
        int cap = 1, len = 1;
        for (int s = p.nCRange[i - 1]; s < p.nCRange[i]; ++s) {
            int nb = p.nodalConn[s];
            if (!nb) continue;
            ++cap;
            int dr = r - p.currentRank[nb - 1];
            if (dr >= 0 && dr <= p.solverMaxBW) len = std::max(len, dr + 1);
        }
        p.thermMatRange[r] = cap;
        p.thermMatLTRange[r] = len;
    }
    {
        int64_t accR = 0, accL = 0;
        for (int r = 1; r <= N; ++r) {
            accR += p.thermMatRange[r];
            accL += p.thermMatLTRange[r];
            if (accR > INT32_MAX || accL > INT32_MAX) return "thermal matrix or profile exceeds 2^31 - 1 entries";
            p.thermMatRange[r] = (int)accR;
            p.thermMatLTRange[r] = (int)accL;
        }
    }
    // First-touch column order.  The assembly loop visits elements in order and,
    // for every row of an element, the element's DOFs in dofTable order
    // (component-major); a row's first free slot takes each new column
    // (Working/AStrO_commandFunctions.f90:323-347).  All rows of one node see the
    // same sequence, so one list per node serves its rows.  Alongside, record the
    // slot of (node b, component d) in node a's list for every element.
    bool allBrick = true, allShell = true;
    for (int e = 0; e < E; ++e) {
        allBrick &= (m.elementType[e] == 8);
        allShell &= (m.elementType[e] == 41);
    }
    // element node count and DOFs per node of the all-brick / all-shell4 mesh
    const int enn = allShell ? 4 : 8, edof = allShell ? 6 : 3;
    p.shellMesh = allShell && E > 0;
    p.elMatCols.assign(p.elMatRange[n], 0);
    p.thermMatCols.assign(p.thermMatRange[N], 0);
    if (allBrick || allShell) {
        p.elPos.assign(192 * (size_t)E, 0);
        p.thPos.assign(64 * (size_t)E, 0);
        // per node: slot of each (adjacent-or-self node, component); looked up by
        // the neighbour's index in the node's adjacency (self = 0)
        std::vector<int> used(N + 1, 0), thUsed(N + 1, 0);
        std::vector<int16_t> slot((size_t)N * 27 * 3, -1), thSlot((size_t)N * 27, -1);
        auto nbIndex = [&](int na, int nb) {
            if (na == nb) return 0;
            int base = p.nCRange[na - 1], q = 1;
            for (int s = base; s < p.nCRange[na]; ++s, ++q)
                if (p.nodalConn[s] == nb) return q;
            return -1;
        };
        if (allShell) slot.assign((size_t)N * 27 * 6, -1);
        for (int e = 0; e < E; ++e) {
            const int* el = m.elementList + 8 * (size_t)e;
            for (int a = 0; a < enn; ++a) {
                int na = el[a];
                int row0 = dofOf(na, 0);
                int rbeg = p.elMatRange[row0 - 1], cap = p.elMatRange[row0] - rbeg;
                int trow = p.currentRank[na - 1];
                int tbeg = p.thermMatRange[trow - 1], tcap = p.thermMatRange[trow] - tbeg;
                int q[8];
                for (int b = 0; b < enn; ++b) {
                    q[b] = nbIndex(na, el[b]);
                    if (q[b] < 0 || q[b] >= 27) return "adjacency larger than a brick mesh allows";
                }
                // columns in dofTable order: component-major over the element's nodes
                for (int d = 0; d < edof; ++d)
                    for (int b = 0; b < enn; ++b) {
                        int16_t& s = slot[((size_t)(na - 1) * 27 + q[b]) * edof + d];
                        if (s < 0) {
                            if (used[na] >= cap || used[na] > 255) return "CSR row capacity exceeded";
                            s = (int16_t)used[na]++;
                            for (int c = 0; c < edof; ++c)
                                p.elMatCols[p.elMatRange[dofOf(na, c) - 1] + s] = dofOf(el[b], d);
                        }
                        p.elPos[192 * (size_t)e + (a * enn + b) * edof + d] = (uint8_t)s;
                    }
                if (allShell) continue;
                // thermal row: columns are node ranks, one per node; with the
                // reference's `.lt.` probe bound (:57) the last slot never fills
                int tlimit = thermLiteralD8 ? tcap - 1 : tcap;
                for (int b = 0; b < 8; ++b) {
                    int16_t& s = thSlot[(size_t)(na - 1) * 27 + q[b]];
                    if (s < 0) {
                        if (thUsed[na] < tlimit) {
                            s = (int16_t)thUsed[na]++;
                            p.thermMatCols[tbeg + s] = p.currentRank[el[b] - 1];
                        } else {
                            s = 255;  // dropped column
                        }
                    }
                    p.thPos[64 * (size_t)e + a * 8 + b] = (uint8_t)s;
                }
            }
        }
        // greedy colouring: smallest colour unused by every node of the element
        std::vector<uint64_t> nodeMask(N + 1, 0);
        std::vector<int> color(E);
        int ncol = 0;
        for (int e = 0; e < E; ++e) {
            const int* el = m.elementList + 8 * (size_t)e;
            uint64_t busy = 0;
            for (int a = 0; a < enn; ++a) busy |= nodeMask[el[a]];
            int c = 0;
            while (c < 64 && (busy >> c) & 1ull) ++c;
            if (c == 64) return "element colouring needs more than 64 colours";
            color[e] = c;
            ncol = std::max(ncol, c + 1);
            for (int a = 0; a < enn; ++a) nodeMask[el[a]] |= 1ull << c;
        }
        p.colorStart.assign(ncol + 1, 0);
        for (int e = 0; e < E; ++e) p.colorStart[color[e] + 1]++;
        for (int c = 0; c < ncol; ++c) p.colorStart[c + 1] += p.colorStart[c];
        p.colorPerm.assign(E, 0);
        std::vector<int> at(p.colorStart.begin(), p.colorStart.end() - 1);
        for (int e = 0; e < E; ++e) p.colorPerm[at[color[e]]++] = e;
    }
    build_mpc(m, p, false, p.elMPCDim, p.elMPCMatRange, p.elMPCMatCols, p.elMPCMat, p.elMPCRHS);
    build_mpc(m, p, true, p.thermMPCDim, p.thermMPCMatRange, p.thermMPCMatCols, p.thermMPCMat, p.thermMPCRHS);
    return "";
}

}  // namespace astro
