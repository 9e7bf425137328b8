// astro_abi.hip -- host orchestration and the C ABI (include/astro_abi.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <climits>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/astro_abi.h"
#include "astro_batch.h"
#include "astro_design.h"
#include "astro_gnn.h"
#include "astro_internal.h"
#include "astro_kernels.h"
#include "astro_prep.h"

namespace astro {

static thread_local std::string g_lastError;

int fail(int code, const std::string& msg) {
    g_lastError = msg;
    return code;
}

// ---------------------------------------------------------------------------
// Factor + sweeps of one row-skyline profile whose entries are stored as LT:
// double (the FP64 path) or float (the mixed-precision factor: FP32 L and
// step-U partials on FP32 matrix cores; pivots, the diagonal-block factor,
// the sweep blocks and every vector stay FP64).
// The penalty C^T C of one constraint structure in profile order: the
// entries (row, column) inside the profile that some constraint row reaches,
// each with its products in convertToLTri's order (constraint row, then its
// term pairs; Working/AStrO_bookKeeping.f90:776-790).  k_penalty_plan adds
// them with one thread per entry.  Built on the host from the constraint
// structure and the profile's first columns.
struct PenaltyPlan {
    int entries = 0;
    DBuf<int> row, col, start, pa, pb;
    int build(const int* crange, const int* ccols, int cd, const std::vector<int>& first, hipStream_t s) {
        struct Term {
            int64_t key;
            int order, a, b;
        };
        std::vector<Term> t;
        const int n = (int)first.size();
        int order = 0;
        for (int r = 0; r < cd; ++r)
            for (int q2 = crange[r]; q2 < crange[r + 1]; ++q2) {
                const int i4 = ccols[q2] - 1;
                if (i4 < 0 || i4 >= n) continue;
                for (int q3 = crange[r]; q3 < crange[r + 1]; ++q3) {
                    const int i5 = ccols[q3] - 1;
                    if (i5 >= first[i4] && i5 <= i4) t.push_back({(int64_t)i4 * n + i5, order++, q2, q3});
                }
            }
        std::sort(t.begin(), t.end(), [](const Term& x, const Term& y) {
            return x.key != y.key ? x.key < y.key : x.order < y.order;
        });
        std::vector<int> hr, hc, hs = {0}, ha, hb;
        for (size_t q = 0; q < t.size(); ++q) {
            if (q == 0 || t[q].key != t[q - 1].key) {
                if (q > 0) hs.push_back((int)q);
                hr.push_back((int)(t[q].key / n));
                hc.push_back((int)(t[q].key % n));
            }
            ha.push_back(t[q].a);
            hb.push_back(t[q].b);
        }
        if (!t.empty()) hs.push_back((int)t.size());
        entries = (int)hr.size();
        ASTRO_TRY(row.upload(hr, s));
        ASTRO_TRY(col.upload(hc, s));
        ASTRO_TRY(start.upload(hs, s));
        ASTRO_TRY(pa.upload(ha, s));
        ASTRO_TRY(pb.upload(hb, s));
        return ASTRO_OK;
    }
    template <typename LT>
    void launch(const double* cmat, const int64_t* rowoff, LT* lt, hipStream_t s) const {
        if (entries > 0)
            hipLaunchKernelGGL(k_penalty_plan<LT>, dim3(grid_for(entries, 256, INT_MAX)), dim3(256), 0, s, cmat,
                               entries, row.p, col.p, start.p, pa.p, pb.p, rowoff, lt);
    }
};

template <typename LT>
struct DeviceProfileT {
    // rows start on 128-byte lines: 16 columns of FP64, 32 of FP32
    static constexpr int kAlign = 128 / (int)sizeof(LT);
    int n = 0;
    int64_t size = 0;        // entries of the device array lt (gaps included)
    int64_t packedSize = 0;  // entries of the reference's packed profile
    bool aligned = false;    // rows padded: see setup
    std::vector<int> hS;     // first column per row
    PenaltyPlan pen;         // the constraint penalty of the owning system (plan_penalty)
    std::vector<int64_t> hOff;  // device row offsets: entry (r, c) at lt[hOff[r] + c]
    DBuf<LT> lt;
    DBuf<int64_t> rowoff;
    DBuf<int> sfirst;
    DBuf<int> winMin;  // min first column over rows [r, r + 64) of r's diagonal block
    // factor
    int nBlocks = 0, panelsPerBlock = 0, splitK = 1, scratchRows = 64, maxTiles = 1;
    DBuf<PanelInfo> panels;
    DBuf<LT> scratch;
    DBuf<double> diag;  // D of the factor, written by step FT, read by step U
    DBuf<int> zeroPivot;
    DBuf<double> ftX;  // per block: L_PP^-1 (64 x 64) and 1 / D of the current panel's diagonal block
    // sweeps
    int numTiles = 0;
    DBuf<SweepTile> fwdTiles, bwdTiles;
    DBuf<SweepBlock> blockInfo;  // per tile, indexed by SweepTile::blk
    DBuf<double> blocks;         // SWEEP_KINDS blocks of 64 x 64 per tile
    bool blocksReady = false;    // blocks match the current factor in lt
    bool factored = false;       // lt holds a factor computed here (diag is valid)
    DBuf<double> gnear;          // per sweep: each tile's L_tt^-1 (or ^-T) part, for its neighbour
    DBuf<int> fwdFlags, bwdFlags, ticket, timeout;
    int fwdEpoch = 0, bwdEpoch = 0;
    DBuf<double> z;
    long spinLimit = 4000000;
    // optional per-launch timing of the two sweeps (events on the launch stream)
    bool timing = false;
    std::vector<hipEvent_t> evPool;
    size_t evUsed = 0;
    double fwdMs = 0, bwdMs = 0;
    long fwdCount = 0, bwdCount = 0;
    std::vector<int> evKind;  // 0 = forward, 1 = backward, per event pair

    ~DeviceProfileT() {
        for (auto e : evPool) (void)hipEventDestroy(e);
    }
    hipEvent_t ev_next() {
        if (evUsed >= evPool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            evPool.push_back(e);
        }
        return evPool[evUsed++];
    }
    // Fold recorded event pairs into the totals (stream must be idle).
    void collect() {
        for (size_t k = 0; k + 1 < evUsed; k += 2) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, evPool[k], evPool[k + 1]) == hipSuccess) {
                if (evKind[k / 2] == 0) {
                    fwdMs += ms;
                    ++fwdCount;
                } else {
                    bwdMs += ms;
                    ++bwdCount;
                }
            }
        }
        evUsed = 0;
        evKind.clear();
    }

    // alignRows: lay the rows out with gaps so that, inside each independent
    // diagonal block starting at row b0, every column c = b0 (mod kAlign) of
    // every row starts a 128-byte line.  The factor's 32-column chunks and
    // the sweeps' 64-column segments then read whole lines (about 1 % faster
    // end to end on the 64x64x32 brick).  Without it lt is the reference's
    // packed array, so it can be copied to and from the host as is.
    // Dynamic LDS (bytes, unused by the code) requested with each sweep
    // launch so that a sweep workgroup holds its CU alone: more than half the
    // CU's LDS in all.  With two per CU, a tile waiting on its neighbour's
    // values shares the CU's memory pipeline with another tile streaming its
    // far rows, and every poll waits behind that tile's loads: one per CU took
    // the sweeps on the 64x64x32 brick from 3.64 / 4.15 ms to 3.37 / 3.58 ms
    // (forward / backward).
    unsigned sweepLdsFwd = 0, sweepLdsBwd = 0;
    int sweep_lds_reserve() {
        int dev = 0, ldsPerCU = 0;
        HIP_TRY(hipGetDevice(&dev));
        HIP_TRY(hipDeviceGetAttribute(&ldsPerCU, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
        hipFuncAttributes fa{}, ba{};
        HIP_TRY(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(k_sweep_forward<LT>)));
        HIP_TRY(hipFuncGetAttributes(&ba, reinterpret_cast<const void*>(k_sweep_backward<LT>)));
        const long half = ldsPerCU / 2 + 1;
        sweepLdsFwd = (unsigned)std::max(0L, half - (long)fa.sharedSizeBytes);
        sweepLdsBwd = (unsigned)std::max(0L, half - (long)ba.sharedSizeBytes);
        return ASTRO_OK;
    }

    int setup(const int* LRange, int dim, hipStream_t s, bool alignRows = false) {
        n = dim;
        packedSize = size = LRange[n];
        aligned = alignRows;
        std::vector<int64_t>& off = hOff;
        off.assign(n, 0);
        hS.resize(n);
        for (int r = 0; r < n; ++r) {
            int len = LRange[r + 1] - LRange[r];
            if (len < 1 || len > r + 1) return fail(ASTRO_ERR_ARG, "profile row " + std::to_string(r + 1) + " has invalid length");
            hS[r] = r - len + 1;
            off[r] = (int64_t)LRange[r + 1] - 1 - r;
        }
        ASTRO_TRY(sfirst.upload(hS, s));
        ASTRO_TRY(sweep_lds_reserve());
        // independent diagonal blocks: a boundary b has no row >= b reaching
        // below b; small blocks are merged so launches stay few
        std::vector<int> sufMin(n + 1, INT32_MAX);
        for (int r = n - 1; r >= 0; --r) sufMin[r] = std::min(sufMin[r + 1], hS[r]);
        std::vector<int> bounds = {0};
        const int minBlock = 16384;
        for (int b = 1; b < n; ++b)
            if (sufMin[b] >= b && b - bounds.back() >= minBlock && n - b >= minBlock && bounds.size() < 64)
                bounds.push_back(b);
        bounds.push_back(n);
        nBlocks = (int)bounds.size() - 1;
        if (aligned) {
            // row r's first stored entry goes at the first position after
            // row r - 1 at which hOff[r] + b0 = 0 (mod kAlign); gaps stay zero
            int64_t end = 0;
            for (int k = 0; k < nBlocks; ++k) {
                const int b0 = bounds[k];
                for (int r = b0; r < bounds[k + 1]; ++r) {
                    int64_t v = end - hS[r];
                    const int64_t rem = ((v + b0) % kAlign + kAlign) % kAlign;
                    if (rem) v += kAlign - rem;
                    off[r] = v;
                    end = v + r + 1;
                }
            }
            size = end;
        }
        ASTRO_TRY(rowoff.upload(off, s));
        {
            // window minimum of the first columns over the next 64 rows of the block
            std::vector<int> wm(n);
            for (int k = 0; k < nBlocks; ++k) {
                const int b0 = bounds[k], b1 = bounds[k + 1];
                for (int r = b0; r < b1; ++r) {
                    int m = INT32_MAX;
                    for (int q = r; q < std::min(r + 64, b1); ++q) m = std::min(m, hS[q]);
                    wm[r] = m;
                }
            }
            ASTRO_TRY(winMin.upload(wm, s));
        }
        // panels
        panelsPerBlock = 0;
        for (int k = 0; k < nBlocks; ++k)
            panelsPerBlock = std::max(panelsPerBlock, (bounds[k + 1] - bounds[k] + 63) / 64);
        std::vector<PanelInfo> hp((size_t)nBlocks * panelsPerBlock, PanelInfo{-1, -1, -1, 0});
        scratchRows = 64;
        for (int k = 0; k < nBlocks; ++k) {
            const int b0 = bounds[k], b1 = bounds[k + 1];
            const int np = (b1 - b0 + 63) / 64;
            std::vector<int> maxRow(np, -1);
            for (int r = b0; r < b1; ++r) {
                int pk = (hS[r] - b0) / 64;
                maxRow[pk] = std::max(maxRow[pk], r);
            }
            int run = -1;
            for (int q = 0; q < np; ++q) {
                run = std::max(run, maxRow[q]);
                PanelInfo P;
                P.p = b0 + 64 * q;
                P.pe = std::min(P.p + 64, b1);
                P.iend = std::max(run, P.pe - 1) + 1;
                int sm = INT32_MAX;
                for (int j = P.p; j < P.pe; ++j) sm = std::min(sm, hS[j]);
                P.smin = sm;
                hp[(size_t)k * panelsPerBlock + q] = P;
                scratchRows = std::max(scratchRows, ((P.iend - P.p) + 63) / 64 * 64);
            }
        }
        maxTiles = scratchRows / 64;
        // about 768 old-update workgroups per launch (4 splits per tile at the
        // 64x64x32 block): fewer and longer ones lose the balance, more add
        // partial sums to every gather (measured 3: 201 ms, 4: 186 ms,
        // 6: 187 ms, 8: 196 ms for the whole factor)
        splitK = std::max(1, std::min(16, (768 + nBlocks * maxTiles - 1) / std::max(1, nBlocks * maxTiles)));
        ASTRO_TRY(panels.upload(hp, s));
        // two buffers (panel parity), each with splitK "old" partials and one "new"
        ASTRO_TRY(scratch.alloc(2 * (size_t)(splitK + 1) * nBlocks * scratchRows * 64));
        ASTRO_TRY(zeroPivot.alloc(1));
        ASTRO_TRY(ftX.alloc((size_t)std::max(1, nBlocks) * (64 * 64 + 64)));
        ASTRO_TRY(diag.alloc((size_t)std::max(1, n)));
        // sweep tiles, chains interleaved in ticket order
        std::vector<std::vector<int>> chainTiles(nBlocks);  // tile start rows per chain
        int maxCount = 0;
        for (int k = 0; k < nBlocks; ++k) {
            for (int r = bounds[k]; r < bounds[k + 1]; r += 64) chainTiles[k].push_back(r);
            maxCount = std::max(maxCount, (int)chainTiles[k].size());
        }
        numTiles = 0;
        for (auto& c : chainTiles) numTiles += (int)c.size();
        std::vector<SweepTile> f, bw;
        // block index of a tile: chains in order, tiles in row order
        std::vector<int> blkBase(nBlocks + 1, 0);
        std::vector<SweepBlock> hb;
        for (int k = 0; k < nBlocks; ++k) {
            blkBase[k + 1] = blkBase[k] + (int)chainTiles[k].size();
            const int cnt = (int)chainTiles[k].size();
            auto rowEnd = [&](int q) { return std::min(chainTiles[k][q] + 64, bounds[k + 1]); };
            for (int q = 0; q < cnt; ++q) {
                SweepBlock B;
                B.r0 = chainTiles[k][q];
                B.r1 = rowEnd(q);
                B.pm1 = q >= 1 ? chainTiles[k][q - 1] : B.r0;
                B.pm2 = q >= 2 ? chainTiles[k][q - 2] : B.pm1;
                B.q1 = q + 1 < cnt ? rowEnd(q + 1) : B.r1;
                B.q2 = q + 2 < cnt ? rowEnd(q + 2) : B.q1;
                B.prevBlk = q >= 1 ? blkBase[k] + q - 1 : -1;
                B.nextBlk = q + 1 < cnt ? blkBase[k] + q + 1 : -1;
                hb.push_back(B);
            }
        }
        // tickets: chains interleaved by position, forward upwards, backward
        // downwards; each tile records its chain position's neighbours
        for (int q = 0; q < maxCount; ++q)
            for (int k = 0; k < nBlocks; ++k) {
                const int cnt = (int)chainTiles[k].size();
                if (q >= cnt) continue;
                const int b0 = bounds[k], b1 = bounds[k + 1];
                const int r0 = chainTiles[k][q], r1 = std::min(r0 + 64, b1);
                int sm = r0;
                for (int r = r0; r < r1; ++r) sm = std::min(sm, hS[r]);
                const int td = (sm - b0) / 64;
                f.push_back(SweepTile{r0, r1, blkBase[k] + q, std::min(4, q), td <= q - 5 ? blkBase[k] + td : -1, b1});
            }
        std::vector<std::vector<int>> reachBack(nBlocks);
        for (int k = 0; k < nBlocks; ++k) {
            const int b0 = bounds[k], b1 = bounds[k + 1];
            const int cnt = (int)chainTiles[k].size();
            // last tile holding a row whose profile reaches tile q
            std::vector<int> maxRowByTile(cnt, -1);
            for (int r = b0; r < b1; ++r) {
                int tc = (hS[r] - b0) / 64;
                maxRowByTile[tc] = std::max(maxRowByTile[tc], r);
            }
            int run = -1;
            reachBack[k].resize(cnt);
            for (int q = 0; q < cnt; ++q) {
                run = std::max(run, maxRowByTile[q]);
                reachBack[k][q] = std::max(q, (run - b0) / 64);
            }
        }
        for (int q = 0; q < maxCount; ++q)
            for (int k = 0; k < nBlocks; ++k) {
                const int cnt = (int)chainTiles[k].size();
                const int t = cnt - 1 - q;
                if (t < 0) continue;
                const int r0 = chainTiles[k][t], r1 = std::min(r0 + 64, bounds[k + 1]);
                const int td = reachBack[k][t];
                bw.push_back(SweepTile{r0, r1, blkBase[k] + t, std::min(4, cnt - 1 - t),
                                       td >= t + 5 ? blkBase[k] + td : -1, bounds[k + 1]});
            }
        ASTRO_TRY(fwdTiles.upload(f, s));
        ASTRO_TRY(bwdTiles.upload(bw, s));
        ASTRO_TRY(blockInfo.upload(hb, s));
        ASTRO_TRY(blocks.alloc((size_t)numTiles * SWEEP_KINDS * SWEEP_BLOCK));
        ASTRO_TRY(gnear.alloc(n));
        blocksReady = false;
        ASTRO_TRY(fwdFlags.alloc(numTiles));
        ASTRO_TRY(bwdFlags.alloc(numTiles));
        HIP_TRY(hipMemsetAsync(fwdFlags.p, 0, sizeof(int) * numTiles, s));
        HIP_TRY(hipMemsetAsync(bwdFlags.p, 0, sizeof(int) * numTiles, s));
        fwdEpoch = bwdEpoch = 0;
        ASTRO_TRY(ticket.alloc(1));
        ASTRO_TRY(timeout.alloc(1));
        HIP_TRY(hipMemsetAsync(timeout.p, 0, sizeof(int), s));
        ASTRO_TRY(z.alloc(n));
        ASTRO_TRY(lt.alloc(size));
        return ASTRO_OK;
    }

    // The profile (or factor) in the reference's packed layout, packedSize
    // doubles; through a host staging copy when the rows are aligned or the
    // storage is FP32.  An FP32 factor reads back its FP64 pivots on the
    // diagonal.
    int get_packed(double* host, hipStream_t s) {
        if (!aligned && sizeof(LT) == sizeof(double)) {
            HIP_TRY(hipMemcpyAsync(host, lt.p, sizeof(double) * size, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            return ASTRO_OK;
        }
        std::vector<LT> tmp(size);
        HIP_TRY(hipMemcpyAsync(tmp.data(), lt.p, sizeof(LT) * size, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        int64_t pos = 0;
        for (int r = 0; r < n; ++r) {
            const int64_t len = r - hS[r] + 1;
            const LT* src = tmp.data() + hOff[r] + hS[r];
            for (int64_t q = 0; q < len; ++q) host[pos + q] = (double)src[q];
            pos += len;
        }
        if (sizeof(LT) != sizeof(double) && factored) {
            std::vector<double> d(n);
            HIP_TRY(hipMemcpyAsync(d.data(), diag.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            pos = 0;
            for (int r = 0; r < n; ++r) {
                pos += r - hS[r] + 1;
                host[pos - 1] = d[r];
            }
        }
        return ASTRO_OK;
    }
    int set_packed(const double* host, hipStream_t s) {
        blocksReady = false;
        factored = false;
        if (!aligned && sizeof(LT) == sizeof(double)) {
            HIP_TRY(hipMemcpyAsync(lt.p, host, sizeof(double) * size, hipMemcpyHostToDevice, s));
            HIP_TRY(hipStreamSynchronize(s));
            return ASTRO_OK;
        }
        std::vector<LT> tmp(size, (LT)0);
        int64_t pos = 0;
        for (int r = 0; r < n; ++r) {
            const int64_t len = r - hS[r] + 1;
            LT* dst = tmp.data() + hOff[r] + hS[r];
            for (int64_t q = 0; q < len; ++q) dst[q] = (LT)host[pos + q];
            pos += len;
        }
        HIP_TRY(hipMemcpyAsync(lt.p, tmp.data(), sizeof(LT) * size, hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
        return ASTRO_OK;
    }

    // Left-looking blocked factor, two launches per panel.  Synchronous.
    int factor(hipStream_t s) {
        HIP_TRY(hipMemsetAsync(zeroPivot.p, 0x7f, sizeof(int), s));
        // Panel q needs W = sum_{k < p} L(i,k) D_k L(P,k): the "old" part
        // (columns before panel q-1, nSplit partials) is computed by launch
        // L_{q-1}, the "new" part (panel q-1's columns) by L_q, which also
        // factors panel q's diagonal block; R_q then forms the panel's other
        // rows.  Partial sums live in scratch buffer q & 1: slots
        // 0..splitK-1 old, slot splitK new.
        const size_t bufSize = (size_t)(splitK + 1) * nBlocks * scratchRows * 64;
        auto buf = [&](int q) { return scratch.p + (size_t)(q & 1) * bufSize; };
        const dim3 gOld(maxTiles * splitK, nBlocks), gLead(nBlocks * maxTiles * (splitK + 1));
        const dim3 gRows(2 * nBlocks * std::max(1, maxTiles - 1));
        // panel 0 has no old part: its slots are written as zeros
        hipLaunchKernelGGL(k_ldl_update_mfma<LT>, gOld, dim3(256), 0, s, panels.p, 0, panelsPerBlock, splitK,
                           LDL_K_OLD, 0, (const LT*)lt.p, rowoff.p, sfirst.p, winMin.p, diag.p, buf(0), scratchRows);
        for (int q = 0; q < panelsPerBlock; ++q) {
            hipLaunchKernelGGL(k_ldl_lead<LT>, gLead, dim3(256), 0, s, panels.p, q, panelsPerBlock, nBlocks, maxTiles,
                               splitK, lt.p, rowoff.p, sfirst.p, winMin.p, diag.p, buf(q), buf(q + 1), scratchRows,
                               zeroPivot.p, ftX.p);
            if (maxTiles > 1)
                hipLaunchKernelGGL(k_ldl_rows<LT>, gRows, dim3(256), 0, s, panels.p, q, panelsPerBlock, nBlocks,
                                   maxTiles, splitK, lt.p, rowoff.p, sfirst.p, (const LT*)buf(q), scratchRows, ftX.p);
        }
        HIP_TRY(hipGetLastError());
        factored = true;
        ASTRO_TRY(prepare_blocks(s));
        int zp = 0;
        HIP_TRY(hipMemcpyAsync(&zp, zeroPivot.p, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (zp != 0x7f7f7f7f) return fail(ASTRO_ERR_ZERO_PIVOT, "zero pivot at profile row " + std::to_string(zp));
        return ASTRO_OK;
    }

    // The sweeps' per-tile blocks, derived from the factor now in lt.  Every
    // writer of lt other than factor() clears blocksReady.
    int prepare_blocks(hipStream_t s) {
        if (numTiles > 0) {
            hipLaunchKernelGGL(k_sweep_blocks<LT>, dim3(numTiles), dim3(256), 0, s, blockInfo.p, (const LT*)lt.p,
                               rowoff.p, sfirst.p, blocks.p);
            hipLaunchKernelGGL(k_sweep_blocks_pair, dim3(numTiles), dim3(256), 0, s, blockInfo.p, blocks.p);
        }
        HIP_TRY(hipGetLastError());
        blocksReady = true;
        return ASTRO_OK;
    }

    // x = P^-1 b (forward then backward sweep); z keeps L^-1 b.  Async.
    // gate: as DeviceOperator::apply.  Every call takes fresh epochs, so the
    // calls recorded into a graph each poll their own flag value.
    int apply(const double* b, double* x, hipStream_t s, const int* gate = nullptr) {
        if (!blocksReady) ASTRO_TRY(prepare_blocks(s));
        ++fwdEpoch;
        // the sweeps poll the neighbouring tile's values: mark them unproduced
        hipLaunchKernelGGL(k_fill_sentinel, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, z.p, (int64_t)n, gate);
        hipLaunchKernelGGL(k_fill_sentinel, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, x, (int64_t)n, gate);
        hipLaunchKernelGGL(k_fill_sentinel, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, gnear.p, (int64_t)n, gate);
        hipLaunchKernelGGL(k_reset_ticket, dim3(1), dim3(1), 0, s, ticket.p, gate);
        if (timing) {
            evKind.push_back(0);
            HIP_TRY(hipEventRecord(ev_next(), s));
        }
        hipLaunchKernelGGL(k_sweep_forward<LT>, dim3(numTiles), dim3(256), sweepLdsFwd, s, fwdTiles.p, numTiles,
                           (const LT*)lt.p, rowoff.p, sfirst.p, (const double*)blocks.p, b, z.p, gnear.p, fwdFlags.p,
                           ticket.p, fwdEpoch, spinLimit, timeout.p, gate);
        if (timing) HIP_TRY(hipEventRecord(ev_next(), s));
        ++bwdEpoch;
        hipLaunchKernelGGL(k_fill_sentinel, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, gnear.p, (int64_t)n, gate);
        hipLaunchKernelGGL(k_reset_ticket, dim3(1), dim3(1), 0, s, ticket.p, gate);
        if (timing) {
            evKind.push_back(1);
            HIP_TRY(hipEventRecord(ev_next(), s));
        }
        hipLaunchKernelGGL(k_sweep_backward<LT>, dim3(numTiles), dim3(256), sweepLdsBwd, s, bwdTiles.p, numTiles,
                           (const LT*)lt.p, rowoff.p, sfirst.p, (const double*)blocks.p, (const double*)z.p, x,
                           gnear.p, bwdFlags.p, ticket.p, bwdEpoch, spinLimit, timeout.p, (const double*)diag.p, gate);
        if (timing) HIP_TRY(hipEventRecord(ev_next(), s));
        HIP_TRY(hipGetLastError());
        return ASTRO_OK;
    }

    // Reports (and clears, so that later calls on the context start clean) a
    // bounded sweep wait that expired since the last check.
    int check_timeout(hipStream_t s) {
        int t = 0;
        HIP_TRY(hipMemcpyAsync(&t, timeout.p, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (timing) collect();
        if (t) {
            hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, timeout.p, (int64_t)1, 0);
            HIP_TRY(hipStreamSynchronize(s));
            return fail(ASTRO_ERR_TIMEOUT, "triangular sweep wait exceeded its spin limit");
        }
        return ASTRO_OK;
    }
};

// The constraint matrix by column: for DOF i (0-based), entries k in
// [range[i], range[i + 1]) name the position pos[k] (0-based) of one of its
// coefficients in the constraint arrays and that coefficient's row row[k].
struct MpcTranspose {
    std::vector<int> range, pos, row;
    void build(const int* crange, const int* ccols, int cdim, int n) {
        range.assign(n + 1, 0);
        for (int r = 0; r < cdim; ++r)
            for (int q = crange[r]; q < crange[r + 1]; ++q)
                if (ccols[q] >= 1 && ccols[q] <= n) ++range[ccols[q]];
        for (int i = 0; i < n; ++i) range[i + 1] += range[i];
        pos.assign(range[n], 0);
        row.assign(range[n], 0);
        std::vector<int> fill(range.begin(), range.end() - 1);
        for (int r = 0; r < cdim; ++r)
            for (int q = crange[r]; q < crange[r + 1]; ++q)
                if (ccols[q] >= 1 && ccols[q] <= n) {
                    const int k = fill[ccols[q] - 1]++;
                    pos[k] = q;
                    row[k] = r;
                }
    }
};

// Device copy of an MpcTranspose plus the double-double C x buffer.
struct DeviceMpcT {
    DBuf<int> range, pos, row;
    DBuf<double> t;
    int upload(const int* crange, const int* ccols, int cdim, int n, hipStream_t s) {
        MpcTranspose T;
        T.build(crange, ccols, cdim, n);
        ASTRO_TRY(range.upload(T.range, s));
        ASTRO_TRY(pos.upload(T.pos, s));
        ASTRO_TRY(row.upload(T.row, s));
        ASTRO_TRY(t.alloc(2 * (size_t)std::max(cdim, 1)));
        return ASTRO_OK;
    }
};

using DeviceProfile = DeviceProfileT<double>;
using DeviceProfile32 = DeviceProfileT<float>;

// CSR operator with optional penalty rows (device arrays, 1-based columns).
struct DeviceOperator {
    int n = 0, cdim = 0;
    const double* val = nullptr;
    const int* cols = nullptr;
    const int* range = nullptr;
    const double* cmat = nullptr;
    const int* ccols = nullptr;
    const int* crange = nullptr;
    // constraint transpose and C x scratch for true residuals (set when cdim > 0)
    const int* ctRange = nullptr;
    const int* ctPos = nullptr;
    const int* ctRow = nullptr;
    double* ctT = nullptr;
    // node-block form of A (brick8 models, k_spmv_nodes): nbK neighbour
    // slots per node over nbN node slots; 0 = the CSR form
    int nbK = 0, nbN = 0;
    const int* nb = nullptr;
    const double* bv = nullptr;
    void set_transpose(DeviceMpcT& T) {
        ctRange = T.range.p;
        ctPos = T.pos.p;
        ctRow = T.row.p;
        ctT = T.t.p;
    }
    // gate (device int, may be null): every kernel returns at once while *gate == 0
    int apply(const double* x, double* y, hipStream_t s, const int* gate = nullptr) const {
        if (nbK > 0)
            hipLaunchKernelGGL(k_spmv_nodes, dim3((nbN + 255) / 256), dim3(256), 0, s, bv, nb, nbK, nbN, x, y, gate);
        else
            hipLaunchKernelGGL(k_spmv, dim3(grid_for((int64_t)n * 32, 256, 1 << 30)), dim3(256), 0, s, val, cols,
                               range, n, x, y, gate);
        if (cdim > 0)
            hipLaunchKernelGGL(k_mpc_apply, dim3(grid_for(cdim, 256)), dim3(256), 0, s, cmat, ccols, crange, cdim, x,
                               y, gate);
        HIP_TRY(hipGetLastError());
        return ASTRO_OK;
    }
    // r = b - (A x + C^T C x), carried in double-double and rounded once.
    int residual(const double* b, const double* x, double* r, hipStream_t s, const int* gate = nullptr) const {
        const bool pen = cdim > 0;
        if (pen && !ctRange) return fail(ASTRO_ERR_ARG, "operator has constraints but no transpose");
        if (pen)
            hipLaunchKernelGGL(k_mpc_dd, dim3(grid_for(cdim, 256)), dim3(256), 0, s, cmat, ccols, crange, cdim, x,
                               ctT, gate);
        hipLaunchKernelGGL(k_residual_dd, dim3(grid_for((int64_t)n * 32, 256, 1 << 30)), dim3(256), 0, s, val, cols,
                           range, n, pen ? ctRange : nullptr, ctPos, ctRow, cmat, ctT, b, x, r, gate);
        HIP_TRY(hipGetLastError());
        return ASTRO_OK;
    }
};

// Restarted, left-preconditioned GMRES with classical Gram-Schmidt and
// re-orthogonalisation (CGS2) in place of the reference's modified
// Gram-Schmidt: the same Arnoldi relation, two multi-vector passes per step
// instead of j dependent dot/axpy pairs.  Standard Givens rotations and a
// beta*e1 right-hand side; a cycle ends when the rotated residual estimate
// meets the target.  The residual b - A x that starts each cycle is formed
// in double-double (DeviceOperator::residual).  Working/AStrO_solvers.f90:128-297
// is the reference form.
struct Gmres {
    int n = 0, m = 0;
    DBuf<double> V, t, t2, r, partial, hdev;
    double* hHost = nullptr;  // pinned
    int nblk = 0;
    int steps = 0, cycles = 0;
    double res0 = 0, res = 0;
    ~Gmres() {
        if (hHost) (void)hipHostFree(hHost);
    }
    int setup(int dim, int numVecs) {
        n = dim;
        m = numVecs;
        ASTRO_TRY(V.alloc((size_t)(m + 1) * n));
        ASTRO_TRY(t.alloc(n));
        ASTRO_TRY(t2.alloc(n));
        ASTRO_TRY(r.alloc(n));
        nblk = std::min(1024, std::max(1, n / 2048));
        ASTRO_TRY(partial.alloc((size_t)nblk * (m + 8)));
        ASTRO_TRY(hdev.alloc(3 * (m + 8)));
        if (!hHost) HIP_TRY(hipHostMalloc(&hHost, sizeof(double) * 3 * (m + 8)));
        return ASTRO_OK;
    }
    // out[0..nq) = V[0..nq)^T w  (device)
    void multidot(const double* Vb, int nq, const double* w, double* out, hipStream_t s, const int* gate = nullptr) {
        hipLaunchKernelGGL(k_multidot, dim3(nblk, (nq + 7) / 8), dim3(256), 0, s, Vb, (int64_t)n, nq, w, n,
                           partial.p, m + 8, gate);
        hipLaunchKernelGGL(k_sum_partials, dim3(nq), dim3(64), 0, s, partial.p, nblk, m + 8, nq, out, gate);
    }
    double* Vcol(int q) { return V.p + (size_t)q * n; }

    // out = |P^-1 (b - A x)|, the residual formed in double-double.  Synchronous.
    template <class Prof>
    int residual_norm(const DeviceOperator& A, Prof& P, const double* b, const double* x, double& out,
                      hipStream_t s) {
        ASTRO_TRY(A.residual(b, x, t2.p, s));
        ASTRO_TRY(P.apply(t2.p, r.p, s));
        multidot(r.p, 1, r.p, hdev.p, s);
        HIP_TRY(hipMemcpyAsync(hHost, hdev.p, sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        ASTRO_TRY(P.check_timeout(s));
        out = std::sqrt(hHost[0]);
        return ASTRO_OK;
    }

    template <class Prof>
    int solve(const DeviceOperator& A, Prof& P, const double* b, double* x, double rtol, double atol,
              int maxCycles, hipStream_t s) {
        steps = cycles = 0;
        const char* ev = std::getenv("ASTRO_VERBOSE");
        const bool verbose = ev && ev[0] == '1';
        // the true residual at every restart is exact to double-double, so the
        // restarts refine the iterate past the FP64 residual floor
        auto residual = [&](double& out) -> int {
            ASTRO_TRY(A.residual(b, x, t2.p, s));
            ASTRO_TRY(P.apply(t2.p, r.p, s));
            multidot(r.p, 1, r.p, hdev.p, s);
            HIP_TRY(hipMemcpyAsync(hHost, hdev.p, sizeof(double), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            ASTRO_TRY(P.check_timeout(s));
            out = std::sqrt(hHost[0]);
            return ASTRO_OK;
        };
        ASTRO_TRY(residual(res));
        res0 = res;
        const double target = std::max(rtol * res0, atol);
        std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
        auto h = [&](int i, int j) -> double& { return H[(size_t)i + (size_t)(m + 1) * j]; };
        while (res > target && cycles < maxCycles) {
            hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, r.p, r.p, Vcol(0), n,
                               1.0 / res, 0.0, (const int*)nullptr);
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = res;
            int k = 0;
            for (int j = 0; j < m; ++j) {
                double* w = Vcol(j + 1);
                ASTRO_TRY(A.apply(Vcol(j), t.p, s));
                ASTRO_TRY(P.apply(t.p, w, s));
                // CGS2: two projections against V[0..j]
                double* h1 = hdev.p;
                double* h2 = hdev.p + (m + 8);
                double* hn = hdev.p + 2 * (m + 8);
                multidot(V.p, j + 1, w, h1, s);
                hipLaunchKernelGGL(k_sub_combination, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, V.p, (int64_t)n,
                                   j + 1, h1, w, n, -1.0, (const int*)nullptr);
                multidot(V.p, j + 1, w, h2, s);
                hipLaunchKernelGGL(k_sub_combination, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, V.p, (int64_t)n,
                                   j + 1, h2, w, n, -1.0, (const int*)nullptr);
                multidot(w, 1, w, hn, s);
                HIP_TRY(hipMemcpyAsync(hHost, hdev.p, sizeof(double) * 3 * (m + 8), hipMemcpyDeviceToHost, s));
                HIP_TRY(hipStreamSynchronize(s));
                ASTRO_TRY(P.check_timeout(s));
                const double hnorm = std::sqrt(hHost[2 * (m + 8)]);
                for (int i = 0; i <= j; ++i) h(i, j) = hHost[i] + hHost[(m + 8) + i];
                h(j + 1, j) = hnorm;
                ++steps;
                k = j + 1;
                // Givens
                for (int i = 0; i < j; ++i) {
                    double a = cs[i] * h(i, j) + sn[i] * h(i + 1, j);
                    double bb = -sn[i] * h(i, j) + cs[i] * h(i + 1, j);
                    h(i, j) = a;
                    h(i + 1, j) = bb;
                }
                const double rr = std::hypot(h(j, j), h(j + 1, j));
                cs[j] = rr == 0.0 ? 1.0 : h(j, j) / rr;
                sn[j] = rr == 0.0 ? 0.0 : h(j + 1, j) / rr;
                h(j, j) = rr;
                h(j + 1, j) = 0.0;
                g[j + 1] = -sn[j] * g[j];
                g[j] = cs[j] * g[j];
                if (hnorm == 0.0 || std::fabs(g[j + 1]) <= target) break;
                hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, w, w, w, n, 1.0 / hnorm,
                                   0.0, (const int*)nullptr);
            }
            for (int i = k - 1; i >= 0; --i) {
                double v = g[i];
                for (int q = i + 1; q < k; ++q) v -= h(i, q) * y[q];
                y[i] = v / h(i, i);
            }
            HIP_TRY(hipMemcpyAsync(hdev.p, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_sub_combination, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, V.p, (int64_t)n, k,
                               hdev.p, x, n, 1.0, (const int*)nullptr);
            ++cycles;
            ASTRO_TRY(residual(res));
            if (verbose)
                std::fprintf(stderr, "gmres: cycle %d steps %d res %.3e (target %.3e)\n", cycles, steps, res, target);
        }
        if (res > target) return fail(ASTRO_ERR_NO_CONVERGE, "GMRES did not reach the tolerance in " +
                                                                 std::to_string(cycles) + " cycles");
        return ASTRO_OK;
    }
};

// The learned two-stage preconditioner: x = P^-1 b, then x += B (b - A x)
// with B the graph network's 3 x 3 node blocks.  A fixed linear operator, so
// plain (left-preconditioned) GMRES applies.
template <class Prof>
struct TwoStagePrec {
    Prof& P;
    const DeviceOperator& A;
    const double* B;
    const int* nodeDofs;
    int N, n;
    double* t1;
    double* t2;
    int apply(const double* b, double* x, hipStream_t s, const int* gate = nullptr) {
        ASTRO_TRY(P.apply(b, x, s, gate));
        ASTRO_TRY(A.apply(x, t1, s, gate));
        hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, b, (const double*)t1, t2, n, 1.0,
                           -1.0, gate);
        hipLaunchKernelGGL(k_node_block_apply, dim3(grid_for(N, 256)), dim3(256), 0, s, N, nodeDofs, B,
                           (const double*)t2, x, gate);
        HIP_TRY(hipGetLastError());
        return ASTRO_OK;
    }
    int check_timeout(hipStream_t s) { return P.check_timeout(s); }
};

// Host LU inverse without pivoting of a 6x6 compliance matrix (column-major),
// the procedure of r_GetInvLU (Working/AStrO_r_designPropertyFunctions.f90:996-1073).
static void compliance_inverse(const double* ep, double* C) {
    double S[36] = {0};
    auto s = [&](int r, int c) -> double& { return S[r + 6 * c]; };
    s(0, 0) = 1.0 / ep[0];
    s(0, 1) = -ep[3] / ep[0];
    s(0, 2) = -ep[4] / ep[0];
    s(1, 1) = 1.0 / ep[1];
    s(1, 2) = -ep[5] / ep[1];
    s(2, 2) = 1.0 / ep[2];
    s(3, 3) = 1.0 / ep[6];
    s(4, 4) = 1.0 / ep[7];
    s(5, 5) = 1.0 / ep[8];
    for (int i = 1; i < 6; ++i)
        for (int j = 0; j < i; ++j) s(i, j) = s(j, i);
    for (int i = 0; i < 6; ++i) {
        for (int j = i; j < 6; ++j)
            for (int k = 0; k < i; ++k) s(i, j) -= s(i, k) * s(k, j);
        double f = 1.0 / s(i, i);
        for (int j = i + 1; j < 6; ++j) {
            for (int k = 0; k < i; ++k) s(j, i) -= s(j, k) * s(k, i);
            s(j, i) *= f;
        }
    }
    for (int c = 0; c < 6; ++c) {
        double z[6], x[6];
        for (int i = 0; i < 6; ++i) {
            z[i] = (i == c) ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) z[i] -= s(i, k) * z[k];
        }
        for (int i = 5; i >= 0; --i) {
            double dp = 0;
            for (int k = i + 1; k < 6; ++k) dp += s(i, k) * x[k];
            x[i] = (z[i] - dp) / s(i, i);
        }
        for (int i = 0; i < 6; ++i) C[i + 6 * c] = x[i];
    }
}

}  // namespace astro

using namespace astro;

// ===========================================================================
struct astro_prep {
    Prep p;
};

// Context
struct astro_ctx {
    hipStream_t s = nullptr;
    int dev = 0;  // the HIP device the context lives on
    Prep prep;
    astro_options opt{};
    int numVecs = 100;
    // model on device
    DBuf<double> X, matC, elScale, elRho, elCond, elCap;
    DBuf<double> elTExp;  // thermal expansion, 6 per element (brick8), empty when the model has none
    bool hasTExp = false;
    DBuf<double> elABD, elAlpha;  // shell4 models: layup stiffness 81, frame 9 per element
    int nLayers = 0;
    std::vector<double> hLayers, hZOff;  // host copy of the shell sections
    // design variables (complex-step sensitivities)
    int nDV = 0;
    std::vector<int> hElDV;
    DBuf<int> elDV;
    DBuf<double> elABDim, dRdD, gPart;
    double csH = 1e-30;
    // batched complex-step solves (astro_ctx_cs_batch_gradient): node-major
    // profile, instances, and their device arrays
    struct CsBatch {
        bool ready = false;
        int nInst = 0, maxRow = 0, tile = 0;
        int64_t total = 0;
        std::vector<int> inv, first, instStart;
        std::vector<int64_t> rowoff;
        DBuf<int> dInv, dDofOrder, dFirst, dOrder, dInstStart, dElStart, dElList, status;
        DBuf<int64_t> dRowoff;
        DBuf<double> profRe, F, work, vec, sol, ed, grad, comp;
        float lastMs = 0.0f, tileMs[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        double flops = 0.0;  // flop model of one launch: sum over instances and rows of len^2 (x4 complex)
    } csb;
    DBuf<double> csbKim;
    DBuf<int> csbKimDof;
    // design variables of astro_ctx_set_design_vars (brick8): the (element,
    // variable) pairs sorted by element colour, then variable, then element,
    // with their direction records; per element its pairs; the node-load
    // entries (0-based DOF, coefficient, variable); the adjoint vectors
    struct DesignVars {
        bool ready = false;
        int nDV = 0;
        double h = 1e-30;
        std::vector<int> cat, comp;
        std::vector<int> pairEl, pairDV, colorPairStart, elPairStart, elPairList;
        std::vector<int> loadDof, loadDV;
        std::vector<double> loadCoef;
        DBuf<int> dPairEl, dPairDV;
        DBuf<double> dDir, dEprops;
    } dvs;
    std::vector<double> hEprops;         // per element engineering constants as created
    std::vector<int> hConn;              // elementList(8, numEls) as created
    std::vector<int> elColor;            // colour of each element
    DBuf<double> adjU, adjT, dLdu, dLdt;
    double csrMassFactor = 0.0;          // massFactor of the last elastic assembly
    bool thermalSteady = false;          // the thermal CSR holds the conductivity alone
    DBuf<int> conn, elMat, rank, perm;
    DBuf<uint8_t> elPos, thPos;
    std::vector<int> colorStart;
    // elastic system
    DBuf<int> csrRange, csrCols, mpcRange, mpcCols, loadDofs;
    DBuf<double> csrVal, mpcMat, mpcRhs, loadVals, load, delDisp, nodeDisp;
    DBuf<unsigned long long> maxBits;
    DeviceProfile prof;
    DeviceProfile32 prof32;  // the elastic profile in FP32 storage (precision = 1)
    bool mixed = false;      // precision = 1: the FP32 profiles are the ones in use
    DeviceOperator op;
    // node-block operator: neighbour slots; per CSR slot its block position
    // 3 j + d (0xff: unused) for the tiled copy (K <= 32), else per block
    // entry its CSR position
    DBuf<int> nbIdx, nbMap;
    DBuf<uint8_t> nbJD;
    DBuf<double> nbVal;
    DeviceMpcT mpcT;
    Gmres gm;
    // thermal system
    DBuf<int> tRange, tCols, tMpcRange, tMpcCols;
    DBuf<double> tVal, tMpcMat, tMpcRhs, tLoad, tDel, nodeTemp;
    // transient thermal: capacitance CSR, the rate state, a work vector
    DBuf<double> prevTemp, prevTdot, nodeTdot, tWork;
    bool thDynReady = false;
    double thGamma = 0.5, thDt = 0.0;
    int64_t thDynSteps = 0, thDynGmresSteps = 0;
    DeviceProfile tprof;
    DeviceProfile32 tprof32;
    DeviceOperator top;
    DeviceMpcT tMpcT;
    Gmres tgm;
    bool thermalReady = false;
    // Newmark state
    bool dynReady = false;
    double nmBeta = 0.25, nmGamma = 0.5, delT = 0.0;
    DBuf<double> prevDisp, prevVel, prevAcc, nodeVel, nodeAcc;
    int64_t dynSteps = 0, dynGmresSteps = 0;
    double dynMs = 0.0;
    // Newmark steps replayed from HIP graphs (stepGraphs = 1): the device
    // GMRES scalars (gmd, gmi), a step's three residual norms, the load
    // factors, and pinned snapshots of gmi (two, for one step of lookahead)
    DBuf<double> gmd, gmNrm, dynScales;
    DBuf<int> gmi;
    int* gmiHost = nullptr;
    hipEvent_t gmEv[2] = {};
    int64_t dynGraphSteps = 0, dynEagerSteps = 0, dynGraphs = 0;
    int dynGraphCap = 0;
    // learned block preconditioner (graph network over the node graph)
    bool gnnReady = false, gnnOn = false;
    std::vector<int> gnnRange, gnnAdj;  // node graph, 0-based
    GnnGraph gnnGraph;
    DBuf<int> nodeDofs, dofNode;
    DBuf<float> gnnX;
    DBuf<double> gnnG, gnnB, gnnT1, gnnT2;
    int64_t gnnFallbackSteps = 0;
    int lastSolveSteps = 0;
    // timing
    hipEvent_t ev[8] = {};
    double phaseMs[6] = {0, 0, 0, 0, 0, 0};
    ~astro_ctx() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : gmEv)
            if (e) (void)hipEventDestroy(e);
        if (gmiHost) (void)hipHostFree(gmiHost);
        if (s) (void)hipStreamDestroy(s);
    }
};

// Call f with the elastic (thermal) profile in use: FP64, or FP32 storage
// with precision = 1.
template <class F>
static int with_prof(astro_ctx* c, F f) {
    return c->mixed ? f(c->prof32) : f(c->prof);
}
template <class F>
static int with_tprof(astro_ctx* c, F f) {
    return c->mixed ? f(c->tprof32) : f(c->tprof);
}

// GMRES on the elastic system for the step load b from x.  Plain: the
// skyline LDL^T preconditioner.  With the learned blocks attached: GMRES with
// the two-stage preconditioner, then the plain LDL^T as judge and fallback: a
// plain-preconditioned GMRES whose absolute target is rtol |P^-1 r0| of the
// step's starting residual returns at once when the learned solve meets it
// and otherwise iterates on.  lastSolveSteps counts the Arnoldi steps of both.
static int ctx_gmres(astro_ctx* c, const double* b, double* x) {
    return with_prof(c, [&](auto& P) -> int {
        const double rtol = c->opt.rtol;
        if (!c->gnnOn) {
            const int rc = c->gm.solve(c->op, P, b, x, rtol, 0.0, c->opt.maxCycles, c->s);
            c->lastSolveSteps = c->gm.steps;
            return rc;
        }
        double res0 = 0.0;
        ASTRO_TRY(c->gm.residual_norm(c->op, P, b, x, res0, c->s));
        TwoStagePrec<std::decay_t<decltype(P)>> T{P,        c->op, c->gnnB.p,  c->nodeDofs.p, c->prep.numNodes,
                                                  c->prep.elMatDim, c->gnnT1.p, c->gnnT2.p};
        const int rc1 = c->gm.solve(c->op, T, b, x, rtol, 0.0, c->opt.maxCycles, c->s);
        if (rc1 != ASTRO_OK && rc1 != ASTRO_ERR_NO_CONVERGE) return rc1;
        const int s1 = c->gm.steps;
        const int rc2 = c->gm.solve(c->op, P, b, x, 0.0, rtol * res0, c->opt.maxCycles, c->s);
        c->gnnFallbackSteps += c->gm.steps;
        c->lastSolveSteps = s1 + c->gm.steps;
        return rc2;
    });
}

extern "C" {

const char* astro_last_error(void) { return g_lastError.c_str(); }
const char* astro_version(void) { return "astro-mi355x 0.1 (gfx950, FP64)"; }

// The node-block form of the elastic operator (k_spmv_nodes) for models whose
// DOFs are three components per solid node, DOF = rank + c nS: from the CSR
// structure, each node's neighbours (ascending node slot) and, per block
// entry, the CSR position that holds it.  Left in
// the CSR form when a node's three rows do not share one neighbour list, a
// node has more than 64 neighbours, or the padded blocks would hold more than
// twice the CSR's slots (boundary nodes have fewer neighbours: 1.98x on a
// 3x2x2 block, 1.04x on 64x64x32).
static int build_node_operator(astro_ctx* c) {
    const auto& p = c->prep;
    const int n = p.elMatDim, N = p.numSolidNds;
    c->op.nbK = 0;
    if (p.shellMesh || N <= 0 || n != 3 * N) return ASTRO_OK;
    const int* R = p.elMatRange.data();
    const int* Cc = p.elMatCols.data();
    std::vector<int> pos(N, -1);
    int K = 0;
    for (int a = 0; a < N; ++a) {
        int m = 0;
        for (int q = R[a]; q < R[a + 1]; ++q) {
            if (Cc[q] <= 0) continue;
            const int b = (Cc[q] - 1) % N;
            if (pos[b] < 0) pos[b] = m++;
        }
        for (int q = R[a]; q < R[a + 1]; ++q)
            if (Cc[q] > 0) pos[(Cc[q] - 1) % N] = -1;
        K = std::max(K, m);
    }
    if (K == 0 || K > 64 || 9.0 * K * N > 2.0 * (double)p.elMatCols.size()) return ASTRO_OK;
    std::vector<int> nbh((size_t)K * N, -1), map((size_t)9 * K * N, -1), lst;
    for (int a = 0; a < N; ++a) {
        // neighbours in ascending slot order: lanes of consecutive slots then
        // read x of neighbouring nodes
        lst.clear();
        for (int q = R[a]; q < R[a + 1]; ++q) {
            if (Cc[q] <= 0) continue;
            const int b = (Cc[q] - 1) % N;
            if (pos[b] < 0) {
                pos[b] = 0;
                lst.push_back(b);
            }
        }
        std::sort(lst.begin(), lst.end());
        const int m = (int)lst.size();
        for (int j = 0; j < m; ++j) {
            pos[lst[j]] = j;
            nbh[(size_t)j * N + a] = lst[j];
        }
        bool ok = true;
        for (int cc = 0; cc < 3 && ok; ++cc) {
            const int row = a + cc * N;
            for (int q = R[row]; q < R[row + 1]; ++q) {
                if (Cc[q] <= 0) continue;
                const int b = (Cc[q] - 1) % N, d = (Cc[q] - 1) / N;
                const int j = pos[b];
                if (j < 0 || map[(size_t)(9 * j + 3 * cc + d) * N + a] >= 0) {
                    ok = false;
                    break;
                }
                map[(size_t)(9 * j + 3 * cc + d) * N + a] = q;
            }
        }
        for (int j = 0; j < m; ++j) pos[nbh[(size_t)j * N + a]] = -1;
        if (!ok) return ASTRO_OK;  // rows of one node with different neighbours: CSR form
    }
    ASTRO_TRY(c->nbIdx.upload(nbh, c->s));
    if (K <= 32) {
        std::vector<uint8_t> jd(p.elMatCols.size(), 0xff);
        for (size_t t = 0; t < map.size(); ++t) {
            if (map[t] < 0) continue;
            const size_t blockRow = t / N;  // 9 j + 3 c + d
            const int j = (int)(blockRow / 9), d = (int)(blockRow % 3);
            jd[map[t]] = (uint8_t)(3 * j + d);
        }
        ASTRO_TRY(c->nbJD.upload(jd, c->s));
        HIP_TRY(hipStreamSynchronize(c->s));  // jd is a local array
    } else {
        ASTRO_TRY(c->nbMap.upload(map, c->s));
    }
    ASTRO_TRY(c->nbVal.alloc(map.size()));
    HIP_TRY(hipStreamSynchronize(c->s));  // nbh and map are local arrays
    c->op.nbK = K;
    c->op.nbN = N;
    c->op.nb = c->nbIdx.p;
    c->op.bv = c->nbVal.p;
    return ASTRO_OK;
}

// Copies the CSR values into the node-block form after every write to them.
static int refresh_node_operator(astro_ctx* c) {
    const int K = c->op.nbK, N = c->op.nbN;
    if (K <= 0) return ASTRO_OK;
    if (K <= 32) {
        hipLaunchKernelGGL(k_node_values, dim3((N + 63) / 64, 3), dim3(256), 0, c->s, c->csrVal.p, c->csrRange.p,
                           c->nbJD.p, K, N, c->nbVal.p);
    } else {
        const int64_t cnt = (int64_t)c->nbMap.n;
        hipLaunchKernelGGL(k_gather_values, dim3(grid_for(cnt, 256, 8192)), dim3(256), 0, c->s, c->csrVal.p,
                           c->nbMap.p, cnt, c->nbVal.p);
    }
    HIP_TRY(hipGetLastError());
    return ASTRO_OK;
}

static int ctx_upload_model(astro_ctx* c, const astro_model* m) {
    hipStream_t s = c->s;
    const Prep& p = c->prep;
    const int E = m->numEls, N = m->numNodes;
    ASTRO_TRY(c->X.upload(m->nodeList, 3 * (size_t)N, s));
    ASTRO_TRY(c->conn.upload(m->elementList, 8 * (size_t)E, s));
    // materials: rows of constants normalised by E1 are deduplicated; the
    // element keeps E1 as a scale (C is linear in the moduli at fixed ratios)
    std::vector<double> mats;
    std::vector<int> elMat(E);
    std::vector<double> scale(E), rho(E, 0.0), cond(6 * (size_t)E, 0.0), cap(E, 0.0);
    std::vector<std::vector<double>> keys;
    for (int e = 0; e < E; ++e) {
        const double* ep = m->eprops ? m->eprops + 9 * (size_t)e : nullptr;
        std::vector<double> k(9, 0.0);
        double E1 = 1.0;
        if (ep) {
            E1 = ep[0];
            for (int i = 0; i < 9; ++i) k[i] = ep[i];
            for (int i : {0, 1, 2, 6, 7, 8}) k[i] /= E1;
        }
        int id = -1;
        for (size_t q = keys.size(); q-- > 0;)
            if (keys[q] == k) {
                id = (int)q;
                break;
            }
        if (id < 0) {
            if (keys.size() > 4096) return fail(ASTRO_ERR_UNSUPPORTED, "more than 4096 distinct material ratios");
            id = (int)keys.size();
            keys.push_back(k);
            double C[36];
            if (ep) {
                compliance_inverse(k.data(), C);
            } else {
                std::fill(C, C + 36, 0.0);
            }
            mats.insert(mats.end(), C, C + 36);
        }
        elMat[e] = id;
        scale[e] = E1;
        if (m->density) rho[e] = m->density[e];
        if (m->specHeat) cap[e] = rho[e] * m->specHeat[e];
        if (m->cond)
            for (int i = 0; i < 6; ++i) cond[6 * (size_t)e + i] = m->cond[6 * (size_t)e + i];
    }
    if (p.shellMesh) {
        if (m->numLayers < 1 || !m->layers || !m->orient)
            return fail(ASTRO_ERR_ARG, "shell4 elements need section layers and orientation");
        std::vector<double> abd(81 * (size_t)E), al(9 * (size_t)E);
        for (int e = 0; e < E; ++e) {
            double xyz[12];
            for (int a = 0; a < 4; ++a) {
                const int nd = m->elementList[8 * (size_t)e + a];
                if (nd < 1 || nd > N) return fail(ASTRO_ERR_ARG, "shell4 node out of range");
                for (int q = 0; q < 3; ++q) xyz[q + 3 * a] = m->nodeList[3 * (size_t)(nd - 1) + q];
            }
            shell_frame(m->orient + 6 * (size_t)e, xyz, &al[9 * (size_t)e]);
            shell_section_abd(m->numLayers, m->layers + 6 * (size_t)m->numLayers * e, m->zOffset ? m->zOffset[e] : 0.0,
                              &abd[81 * (size_t)e]);
        }
        ASTRO_TRY(c->elABD.upload(abd, s));
        ASTRO_TRY(c->elAlpha.upload(al, s));
        c->nLayers = m->numLayers;
        c->hLayers.assign(m->layers, m->layers + 6 * (size_t)m->numLayers * E);
        c->hZOff.assign(E, 0.0);
        if (m->zOffset) c->hZOff.assign(m->zOffset, m->zOffset + E);
    }
    if (m->thermExp && !p.shellMesh) {
        for (size_t k = 0; k < 6 * (size_t)E && !c->hasTExp; ++k) c->hasTExp = m->thermExp[k] != 0.0;
        if (c->hasTExp) {
            ASTRO_TRY(c->elTExp.upload(m->thermExp, 6 * (size_t)E, s));
            // the elastic residual reads nodal temperatures: zero until a
            // thermal solve or the caller sets them
            ASTRO_TRY(c->nodeTemp.alloc(N));
            HIP_TRY(hipMemsetAsync(c->nodeTemp.p, 0, sizeof(double) * N, s));
        }
    }
    if (m->eprops) c->hEprops.assign(m->eprops, m->eprops + 9 * (size_t)E);
    c->hConn.assign(m->elementList, m->elementList + 8 * (size_t)E);
    c->elColor.assign(E, 0);
    for (size_t k = 0; k + 1 < p.colorStart.size(); ++k)
        for (int q = p.colorStart[k]; q < p.colorStart[k + 1]; ++q) c->elColor[p.colorPerm[q]] = (int)k;
    ASTRO_TRY(c->matC.upload(mats, s));
    ASTRO_TRY(c->elMat.upload(elMat, s));
    ASTRO_TRY(c->elScale.upload(scale, s));
    ASTRO_TRY(c->elRho.upload(rho, s));
    ASTRO_TRY(c->elCond.upload(cond, s));
    ASTRO_TRY(c->elCap.upload(cap, s));
    ASTRO_TRY(c->rank.upload(p.currentRank, s));
    ASTRO_TRY(c->elPos.upload(p.elPos, s));
    ASTRO_TRY(c->thPos.upload(p.thPos, s));
    ASTRO_TRY(c->perm.upload(p.colorPerm, s));
    c->colorStart = p.colorStart;
    const int n = p.elMatDim;
    ASTRO_TRY(c->csrRange.upload(p.elMatRange, s));
    ASTRO_TRY(c->csrCols.upload(p.elMatCols, s));
    ASTRO_TRY(c->csrVal.alloc(p.elMatCols.size()));
    ASTRO_TRY(c->mpcRange.upload(p.elMPCMatRange, s));
    ASTRO_TRY(c->mpcCols.upload(p.elMPCMatCols, s));
    ASTRO_TRY(c->mpcMat.upload(p.elMPCMat, s));
    ASTRO_TRY(c->mpcRhs.upload(p.elMPCRHS, s));
    // nodal loads expanded to (dof, value) pairs (node-set loads arrive expanded)
    std::vector<int> ld;
    std::vector<double> lv;
    for (int l = 0; l < m->numLoads; ++l) {
        int nd = m->loadNode[l];
        if (nd < 1 || nd > N) return fail(ASTRO_ERR_ARG, "load node out of range");
        for (int q = 0; q < 6; ++q) {
            int d = p.nDofIndex[6 * (size_t)(nd - 1) + q];
            double v = m->loadVal[7 * (size_t)l + q];
            if (d != 0 && v != 0.0) {
                ld.push_back(d);
                lv.push_back(v);
            }
        }
    }
    ASTRO_TRY(c->loadDofs.upload(ld, s));
    ASTRO_TRY(c->loadVals.upload(lv, s));
    ASTRO_TRY(c->load.alloc(n));
    ASTRO_TRY(c->delDisp.alloc(n));
    ASTRO_TRY(c->nodeDisp.alloc(n));
    HIP_TRY(hipMemsetAsync(c->delDisp.p, 0, sizeof(double) * n, s));
    HIP_TRY(hipMemsetAsync(c->nodeDisp.p, 0, sizeof(double) * n, s));
    ASTRO_TRY(c->maxBits.alloc(1));
    ASTRO_TRY(with_prof(c, [&](auto& P) {
        ASTRO_TRY(P.setup(p.elMatLTRange.data(), n, s, true));
        return P.pen.build(p.elMPCMatRange.data(), p.elMPCMatCols.data(), p.elMPCDim, P.hS, s);
    }));
    c->op.n = n;
    c->op.cdim = p.elMPCDim;
    c->op.val = c->csrVal.p;
    c->op.cols = c->csrCols.p;
    c->op.range = c->csrRange.p;
    c->op.cmat = c->mpcMat.p;
    c->op.ccols = c->mpcCols.p;
    c->op.crange = c->mpcRange.p;
    ASTRO_TRY(c->mpcT.upload(p.elMPCMatRange.data(), p.elMPCMatCols.data(), p.elMPCDim, n, s));
    c->op.set_transpose(c->mpcT);
    ASTRO_TRY(build_node_operator(c));
    ASTRO_TRY(c->gm.setup(n, c->numVecs));
    return ASTRO_OK;
}

static ModelDesc to_desc(const astro_model* m) {
    ModelDesc d;
    d.numNodes = m->numNodes;
    d.numEls = m->numEls;
    d.nodeList = m->nodeList;
    d.elementList = m->elementList;
    d.elementType = m->elementType;
    d.numMpcTerms = m->numMpcTerms;
    d.mpcEqn = m->mpcEqn;
    d.mpcDof = m->mpcDof;
    d.mpcCoef = m->mpcCoef;
    d.mpcIsSet = m->mpcIsSet;
    d.mpcNodeStart = m->mpcNodeStart;
    d.mpcNodes = m->mpcNodes;
    d.numMPC = m->numMPC;
    d.mpcRHS = m->mpcRHS;
    return d;
}

int astro_ctx_create(const astro_model* m, const astro_options* o, astro_ctx** out) {
    if (!m || !out) return fail(ASTRO_ERR_ARG, "null argument");
    *out = nullptr;
    std::unique_ptr<astro_ctx> c(new astro_ctx);
    if (o) c->opt = *o;
    if (c->opt.rtol <= 0) c->opt.rtol = 1e-12;
    if (c->opt.maxCycles <= 0) c->opt.maxCycles = 1000;
    if (c->opt.precision < 0 || c->opt.precision > 1) return fail(ASTRO_ERR_ARG, "precision must be 0 or 1");
    c->mixed = c->opt.precision == 1;
    HIP_TRY(hipGetDevice(&c->dev));
    HIP_TRY(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    for (auto& e : c->ev) HIP_TRY(hipEventCreate(&e));
    ModelDesc d = to_desc(m);
    std::string err = analysis_prep(d, c->opt.maxBW, c->opt.literalMaxBW != 0, c->opt.thermLiteralD8 != 0, c->prep,
                                    c->opt.coupledProfile != 0);
    if (!err.empty()) return fail(ASTRO_ERR_ARG, err);
    for (int e = 0; e < m->numEls; ++e)
        if (m->elementType[e] != (c->prep.shellMesh ? 41 : 8))
            return fail(ASTRO_ERR_UNSUPPORTED, "device path supports all-brick8 or all-shell4 meshes");
    const int n = c->prep.elMatDim;
    c->numVecs = c->opt.numVecs > 0 ? c->opt.numVecs : (n < 200 ? n / 2 : 100);
    if (c->numVecs < 1 || c->numVecs > 255) return fail(ASTRO_ERR_ARG, "numVecs must be in [1, 255]");
    ASTRO_TRY(ctx_upload_model(c.get(), m));
    HIP_TRY(hipStreamSynchronize(c->s));
    *out = c.release();
    return ASTRO_OK;
}

int astro_prep_create(const astro_model* m, const astro_options* o, astro_prep** out) {
    if (!m || !out) return fail(ASTRO_ERR_ARG, "null argument");
    *out = nullptr;
    astro_options opt{};
    if (o) opt = *o;
    std::unique_ptr<astro_prep> p(new astro_prep);
    std::string err = analysis_prep(to_desc(m), opt.maxBW, opt.literalMaxBW != 0, opt.thermLiteralD8 != 0, p->p,
                                    opt.coupledProfile != 0);
    if (!err.empty()) return fail(ASTRO_ERR_ARG, err);
    *out = p.release();
    return ASTRO_OK;
}

int astro_prep_destroy(astro_prep* p) {
    delete p;
    return ASTRO_OK;
}

int64_t astro_prep_scalar(const astro_prep* pp, const char* name) {
    if (!pp || !name) return -1;
    const Prep& p = pp->p;
    std::string s(name);
    if (s == "elMatDim") return p.elMatDim;
    if (s == "elMatSize") return (int64_t)p.elMatCols.size();
    if (s == "elMatLTSize") return p.elMatLTRange.empty() ? 0 : p.elMatLTRange.back();
    if (s == "numSolidNds") return p.numSolidNds;
    if (s == "solverMaxBW") return p.solverMaxBW;
    if (s == "elMPCDim") return p.elMPCDim;
    if (s == "elMPCSize") return (int64_t)p.elMPCMat.size();
    if (s == "thermMPCDim") return p.thermMPCDim;
    if (s == "thermMatSize") return (int64_t)p.thermMatCols.size();
    if (s == "thermMatLTSize") return p.thermMatLTRange.empty() ? 0 : p.thermMatLTRange.back();
    if (s == "numColors") return (int64_t)p.colorStart.size() - 1;
    return -1;
}

static const std::vector<int>* prep_int_array(const Prep& p, const std::string& s) {
    if (s == "currentRank") return &p.currentRank;
    if (s == "nDofIndex") return &p.nDofIndex;
    if (s == "elMatRange") return &p.elMatRange;
    if (s == "elMatCols") return &p.elMatCols;
    if (s == "elMatLTRange") return &p.elMatLTRange;
    if (s == "elMPCMatRange") return &p.elMPCMatRange;
    if (s == "elMPCMatCols") return &p.elMPCMatCols;
    if (s == "thermMatRange") return &p.thermMatRange;
    if (s == "thermMatCols") return &p.thermMatCols;
    if (s == "thermMatLTRange") return &p.thermMatLTRange;
    if (s == "thermMPCMatRange") return &p.thermMPCMatRange;
    if (s == "thermMPCMatCols") return &p.thermMPCMatCols;
    return nullptr;
}

int astro_prep_get_i32(const astro_prep* pp, const char* name, int32_t* host, int64_t count) {
    if (!pp || !name || !host) return fail(ASTRO_ERR_ARG, "null argument");
    const std::vector<int>* v = prep_int_array(pp->p, name);
    if (!v) return fail(ASTRO_ERR_ARG, std::string("unknown int array ") + name);
    if ((int64_t)v->size() != count) return fail(ASTRO_ERR_ARG, std::string("size mismatch for ") + name);
    if (count) std::memcpy(host, v->data(), sizeof(int) * count);
    return ASTRO_OK;
}

int astro_prep_get_f64(const astro_prep* pp, const char* name, double* host, int64_t count) {
    if (!pp || !name || !host) return fail(ASTRO_ERR_ARG, "null argument");
    std::string s(name);
    const std::vector<double>* v = nullptr;
    if (s == "elMPCMat") v = &pp->p.elMPCMat;
    if (s == "elMPCRHS") v = &pp->p.elMPCRHS;
    if (s == "thermMPCMat") v = &pp->p.thermMPCMat;
    if (s == "thermMPCRHS") v = &pp->p.thermMPCRHS;
    if (!v) return fail(ASTRO_ERR_ARG, std::string("unknown double array ") + name);
    if ((int64_t)v->size() != count) return fail(ASTRO_ERR_ARG, std::string("size mismatch for ") + name);
    if (count) std::memcpy(host, v->data(), sizeof(double) * count);
    return ASTRO_OK;
}

int astro_ctx_destroy(astro_ctx* c) {
    if (c) {
        (void)hipStreamSynchronize(c->s);
        delete c;
    }
    return ASTRO_OK;
}

int64_t astro_ctx_scalar(const astro_ctx* c, const char* name) {
    if (!c || !name) return -1;
    std::string s(name);
    const Prep& p = c->prep;
    if (s == "elMatDim") return p.elMatDim;
    if (s == "elMatSize") return (int64_t)p.elMatCols.size();
    if (s == "elMatLTSize") return p.elMatLTRange.empty() ? 0 : p.elMatLTRange.back();
    if (s == "numSolidNds") return p.numSolidNds;
    if (s == "solverMaxBW") return p.solverMaxBW;
    if (s == "elMPCDim") return p.elMPCDim;
    if (s == "elMPCSize") return (int64_t)p.elMPCMat.size();
    if (s == "thermMatSize") return (int64_t)p.thermMatCols.size();
    if (s == "thermMatLTSize") return p.thermMatLTRange.empty() ? 0 : p.thermMatLTRange.back();
    if (s == "gmresSteps") return c->lastSolveSteps;
    if (s == "gnnFallbackSteps") return c->gnnFallbackSteps;
    if (s == "gnnAttached") return c->gnnOn ? 1 : 0;
    if (s == "gmresCycles") return c->gm.cycles;
    if (s == "thermGmresSteps") return c->tgm.steps;
    if (s == "numColors") return (int64_t)c->colorStart.size() - 1;
    if (s == "profileBlocks") return c->mixed ? c->prof32.nBlocks : c->prof.nBlocks;
    if (s == "precision") return c->mixed ? 1 : 0;
    if (s == "numVecs") return c->numVecs;
    if (s == "dynSteps") return c->dynSteps;
    if (s == "dynGmresSteps") return c->dynGmresSteps;
    if (s == "thDynSteps") return c->thDynSteps;
    if (s == "thDynGmresSteps") return c->thDynGmresSteps;
    if (s == "dynGraphSteps") return c->dynGraphSteps;
    if (s == "dynEagerSteps") return c->dynEagerSteps;
    if (s == "dynGraphs") return c->dynGraphs;
    if (s == "dynGraphCap") return c->dynGraphCap;
    if (s == "nodeBlockNeighbours") return c->op.nbK;
    return -1;
}

double astro_ctx_scalar_f(const astro_ctx* c, const char* name) {
    if (!c || !name) return NAN;
    std::string s(name);
    if (s == "gmresRes") return c->gm.res;
    if (s == "gmresRes0") return c->gm.res0;
    if (s == "thermGmresRes") return c->tgm.res;
    if (s == "dynMs") return c->dynMs;
    if (s == "csBatchMs") return c->csb.lastMs;
    if (s == "csBatchFlops") return c->csb.flops;
    if (s == "csBatchEntries") return (double)c->csb.total;
    if (s == "csBatchMaxRow") return c->csb.maxRow;
    return NAN;
}

static const std::vector<int>* host_int_array(const astro_ctx* c, const std::string& s) {
    const Prep& p = c->prep;
    if (s == "currentRank") return &p.currentRank;
    if (s == "nDofIndex") return &p.nDofIndex;
    if (s == "elMatRange") return &p.elMatRange;
    if (s == "elMatCols") return &p.elMatCols;
    if (s == "elMatLTRange") return &p.elMatLTRange;
    if (s == "elMPCMatRange") return &p.elMPCMatRange;
    if (s == "elMPCMatCols") return &p.elMPCMatCols;
    if (s == "thermMatRange") return &p.thermMatRange;
    if (s == "thermMatCols") return &p.thermMatCols;
    if (s == "thermMatLTRange") return &p.thermMatLTRange;
    return nullptr;
}

int astro_ctx_get_i32(astro_ctx* c, const char* name, int32_t* host, int64_t count) {
    if (!c || !name || !host) return fail(ASTRO_ERR_ARG, "null argument");
    const std::vector<int>* v = host_int_array(c, name);
    if (!v) return fail(ASTRO_ERR_ARG, std::string("unknown int array ") + name);
    if ((int64_t)v->size() != count) return fail(ASTRO_ERR_ARG, std::string("size mismatch for ") + name);
    std::memcpy(host, v->data(), sizeof(int) * count);
    return ASTRO_OK;
}

static double* dev_double_array(astro_ctx* c, const std::string& s, int64_t& len) {
    auto pick = [&](DBuf<double>& b, int64_t l) {
        len = l;
        return b.p;
    };
    const Prep& p = c->prep;
    if (s == "elasticMat") return pick(c->csrVal, (int64_t)p.elMatCols.size());
    if (s == "elMatLT") {
        len = c->mixed ? c->prof32.packedSize : c->prof.packedSize;
        return nullptr;  // profiles cross through get_packed / set_packed
    }
    if (s == "elMPCMat") return pick(c->mpcMat, (int64_t)p.elMPCMat.size());
    if (s == "elMPCRHS") return pick(c->mpcRhs, (int64_t)p.elMPCRHS.size());
    if (s == "elasticLoad") return pick(c->load, p.elMatDim);
    if (s == "nodeDisp") return pick(c->nodeDisp, p.elMatDim);
    if (s == "delDisp") return pick(c->delDisp, p.elMatDim);
    if (s == "thermMat") return pick(c->tVal, (int64_t)p.thermMatCols.size());
    if (s == "thermMatLT") {
        len = c->mixed ? c->tprof32.packedSize : c->tprof.packedSize;
        return nullptr;
    }
    if (s == "thermalLoad") return pick(c->tLoad, p.numNodes);
    if (s == "nodeTemp") return pick(c->nodeTemp, p.numNodes);
    if (s == "nodeTdot") return pick(c->nodeTdot, p.numNodes);
    if (s == "nodeVel") return pick(c->nodeVel, p.elMatDim);
    if (s == "nodeAcc") return pick(c->nodeAcc, p.elMatDim);
    if (s == "prevDisp") return pick(c->prevDisp, p.elMatDim);
    len = -1;
    return nullptr;
}

int astro_ctx_get_f64(astro_ctx* c, const char* name, double* host, int64_t count) {
    if (!c || !name || !host) return fail(ASTRO_ERR_ARG, "null argument");
    int64_t len = 0;
    double* d = dev_double_array(c, name, len);
    if (len < 0) return fail(ASTRO_ERR_ARG, std::string("unknown double array ") + name);
    if (len != count) return fail(ASTRO_ERR_ARG, std::string("size mismatch for ") + name);
    const std::string nm(name);
    if (nm == "elMatLT") return with_prof(c, [&](auto& P) { return P.get_packed(host, c->s); });
    if (nm == "thermMatLT") {
        if (!c->thermalReady) return fail(ASTRO_ERR_ARG, "array not built yet: thermMatLT");
        return with_tprof(c, [&](auto& P) { return P.get_packed(host, c->s); });
    }
    if (!d) return fail(ASTRO_ERR_ARG, std::string("array not built yet: ") + name);
    HIP_TRY(hipMemcpyAsync(host, d, sizeof(double) * count, hipMemcpyDeviceToHost, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    return ASTRO_OK;
}

int astro_ctx_set_f64(astro_ctx* c, const char* name, const double* host, int64_t count) {
    if (!c || !name || !host) return fail(ASTRO_ERR_ARG, "null argument");
    int64_t len = 0;
    double* d = dev_double_array(c, name, len);
    if (len < 0) return fail(ASTRO_ERR_ARG, std::string("unknown double array ") + name);
    if (len != count) return fail(ASTRO_ERR_ARG, std::string("size mismatch for ") + name);
    const std::string nm(name);
    if (nm == "elMatLT") return with_prof(c, [&](auto& P) { return P.set_packed(host, c->s); });
    if (nm == "thermMatLT") {
        if (!c->thermalReady) return fail(ASTRO_ERR_ARG, "array not built yet: thermMatLT");
        return with_tprof(c, [&](auto& P) { return P.set_packed(host, c->s); });
    }
    if (!d) return fail(ASTRO_ERR_ARG, std::string("unknown double array ") + name);
    HIP_TRY(hipMemcpyAsync(d, host, sizeof(double) * count, hipMemcpyHostToDevice, c->s));
    if (nm == "elasticMat") ASTRO_TRY(refresh_node_operator(c));
    HIP_TRY(hipStreamSynchronize(c->s));
    return ASTRO_OK;
}

int astro_ctx_assemble_elastic(astro_ctx* c, double massFactor) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    const int64_t nnz = (int64_t)c->prep.elMatCols.size();
    c->csrMassFactor = massFactor;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(nnz, 256, 8192)), dim3(256), 0, c->s, c->csrVal.p, nnz, 0.0);
    const int nS = c->prep.numSolidNds;
    if (c->prep.shellMesh) {
        if (massFactor != 0.0) return fail(ASTRO_ERR_UNSUPPORTED, "shell4 inertia is not part of the device path");
        for (size_t k = 0; k + 1 < c->colorStart.size(); ++k) {
            const int cnt = c->colorStart[k + 1] - c->colorStart[k];
            if (cnt <= 0) continue;
            hipLaunchKernelGGL(k_shell4_tangent, dim3(cnt), dim3(64), 0, c->s, c->perm.p + c->colorStart[k], cnt,
                               c->conn.p, c->X.p, c->elABD.p, c->elAlpha.p, c->rank.p, nS, c->prep.numSectionNds,
                               c->csrRange.p, c->elPos.p, c->csrVal.p);
        }
        HIP_TRY(hipGetLastError());
        return ASTRO_OK;
    }
    for (size_t k = 0; k + 1 < c->colorStart.size(); ++k) {
        const int cnt = c->colorStart[k + 1] - c->colorStart[k];
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(k_brick8_tangent, dim3(cnt), dim3(64), 0, c->s, c->perm.p + c->colorStart[k], cnt,
                           c->conn.p, c->X.p, c->elMat.p, c->elScale.p, c->matC.p, c->elRho.p, massFactor, c->rank.p,
                           nS, c->csrRange.p, c->elPos.p, c->csrVal.p);
    }
    HIP_TRY(hipGetLastError());
    return refresh_node_operator(c);
}

int astro_ctx_scale_mpc(astro_ctx* c) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    const int64_t nnz = (int64_t)c->prep.elMatCols.size();
    HIP_TRY(hipMemsetAsync(c->maxBits.p, 0, sizeof(unsigned long long), c->s));
    hipLaunchKernelGGL(k_max_abs, dim3(grid_for(nnz, 256, 2048)), dim3(256), 0, c->s, c->csrVal.p, nnz, c->maxBits.p);
    const int cd = c->prep.elMPCDim;
    if (cd > 0)
        hipLaunchKernelGGL(k_scale_mpc, dim3(grid_for(cd, 256)), dim3(256), 0, c->s, c->mpcMat.p, c->mpcRhs.p,
                           c->mpcRange.p, cd, c->maxBits.p);
    HIP_TRY(hipGetLastError());
    return ASTRO_OK;
}

// convertToLTri on the device: CSR values plus the penalty C^T C into the
// profile of storage type LT (Working/AStrO_bookKeeping.f90:754-792).
extern "C++" {
template <typename LT>
static int profile_from_csr(DeviceProfileT<LT>& P, hipStream_t s, const double* val, const int* cols,
                            const int* range, int n, const double* cmat, const int* ccols, const int* crange,
                            int cd) {
    P.blocksReady = false;
    P.factored = false;
    hipLaunchKernelGGL(k_fill_t<LT>, dim3(grid_for(P.size, 256, 16384)), dim3(256), 0, s, P.lt.p, P.size, (LT)0);
    hipLaunchKernelGGL(k_csr_to_profile<LT>, dim3(grid_for((int64_t)n * 32, 256, 1 << 30)), dim3(256), 0, s, val,
                       cols, range, n, (const int*)P.sfirst.p, (const int64_t*)P.rowoff.p, P.lt.p);
    if (cd > 0) P.pen.launch(cmat, (const int64_t*)P.rowoff.p, P.lt.p, s);
    HIP_TRY(hipGetLastError());
    return ASTRO_OK;
}
}  // extern "C++"

int astro_ctx_build_profile(astro_ctx* c) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    return with_prof(c, [&](auto& P) {
        return profile_from_csr(P, c->s, c->csrVal.p, c->csrCols.p, c->csrRange.p, c->prep.elMatDim, c->mpcMat.p,
                                c->mpcCols.p, c->mpcRange.p, c->prep.elMPCDim);
    });
}

int astro_ctx_factor(astro_ctx* c) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    return with_prof(c, [&](auto& P) { return P.factor(c->s); });
}

// load = -R(u) for the linear static model, R(u) = K u with K the tangent at
// rest in the CSR: getElasticSolnLoad(0) (Working/AStrO_commandFunctions.f90:
// 260-351) on a zeroed load.  The reference evaluates Green-Lagrange strains
// there even without nLGeom (D20); the linear model's internal force is K u.
static void ctx_minus_internal_force(hipStream_t s, const DBuf<double>& val, const DBuf<int>& cols,
                                     const DBuf<int>& range, int n, const double* u, double* load) {
    hipLaunchKernelGGL(k_spmv, dim3(grid_for((int64_t)n * 32, 256, 1 << 30)), dim3(256), 0, s, val.p, cols.p,
                       range.p, n, u, load, (const int*)nullptr);
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, load, load, load, n, -1.0, 0.0, (const int*)nullptr);
}

// Right-hand side of a static step (stepSolve, Working/AStrO_commandFunctions.f90:
// 716-723): -R(u) + nodal loads + C^T (rhs - C u).  With the internal force in
// it, a second solve from a converged state leaves the state unchanged.
// load += scale * the thermal load of the current nodeTemp (models with
// thermal expansion; F - R(u, T) at u = 0 holds it).
static int ctx_thermal_load(astro_ctx* c, double scale, double* load) {
    if (!c->hasTExp) return ASTRO_OK;
    for (size_t k = 0; k + 1 < c->colorStart.size(); ++k) {
        const int cnt = c->colorStart[k + 1] - c->colorStart[k];
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(k_brick8_thermal_strain, dim3((cnt + 127) / 128), dim3(128), 0, c->s, 0,
                           c->perm.p + c->colorStart[k], cnt, c->conn.p, c->X.p, c->elMat.p, c->elScale.p, c->matC.p,
                           c->elTExp.p, c->rank.p, c->prep.numSolidNds, (const double*)c->nodeTemp.p,
                           (const double*)nullptr, scale, load);
    }
    HIP_TRY(hipGetLastError());
    return ASTRO_OK;
}

static int ctx_rhs(astro_ctx* c) {
    const int n = c->prep.elMatDim;
    ctx_minus_internal_force(c->s, c->csrVal, c->csrCols, c->csrRange, n, c->nodeDisp.p, c->load.p);
    ASTRO_TRY(ctx_thermal_load(c, 1.0, c->load.p));
    const int nl = (int)c->loadDofs.n;
    if (nl > 0)
        hipLaunchKernelGGL(k_nodal_loads, dim3(grid_for(nl, 256)), dim3(256), 0, c->s, c->loadDofs.p, c->loadVals.p,
                           nl, 1.0, c->load.p, (const int*)nullptr);
    const int cd = c->prep.elMPCDim;
    if (cd > 0)
        hipLaunchKernelGGL(k_constraint_load, dim3(grid_for(cd, 256)), dim3(256), 0, c->s, c->mpcMat.p, c->mpcCols.p,
                           c->mpcRange.p, c->mpcRhs.p, cd, c->nodeDisp.p, c->load.p, (const int*)nullptr);
    HIP_TRY(hipGetLastError());
    return ASTRO_OK;
}

int astro_ctx_solve(astro_ctx* c) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    ASTRO_TRY(ctx_rhs(c));
    const int n = c->prep.elMatDim;
    ASTRO_TRY(ctx_gmres(c, c->load.p, c->delDisp.p));
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 256, 4096)), dim3(256), 0, c->s, c->nodeDisp.p, c->delDisp.p,
                       c->nodeDisp.p, n, 1.0, 1.0, (const int*)nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->s));
    return ASTRO_OK;
}

int astro_ctx_reset(astro_ctx* c) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    const int n = c->prep.elMatDim;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, c->s, c->delDisp.p, (int64_t)n, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, c->s, c->nodeDisp.p, (int64_t)n, 0.0);
    if (c->thermalReady) {
        const int N = c->prep.numNodes;
        hipLaunchKernelGGL(k_fill, dim3(grid_for(N, 256, 4096)), dim3(256), 0, c->s, c->tDel.p, (int64_t)N, 0.0);
        hipLaunchKernelGGL(k_fill, dim3(grid_for(N, 256, 4096)), dim3(256), 0, c->s, c->nodeTemp.p, (int64_t)N, 0.0);
    }
    HIP_TRY(hipGetLastError());
    return ASTRO_OK;
}

// Element stress and strain at the natural point sp of every element for the
// current nodeDisp (r_getElStress with layer = 0, the solid branch,
// Working/AStrO_r_elementEqns.f90:924-976, 1700-1713).  stress, strain:
// host arrays of 6 numEls (Voigt 11, 22, 33, 12, 13, 23).  brick8 models.
int astro_ctx_element_stress(astro_ctx* c, const double* sp, double* stress, double* strain) {
    if (!c || !sp || !stress || !strain) return fail(ASTRO_ERR_ARG, "null argument");
    if (c->prep.shellMesh) return fail(ASTRO_ERR_UNSUPPORTED, "stress recovery covers brick8 elements");
    for (int k = 0; k < 3; ++k)
        if (!(std::fabs(sp[k]) <= 1.0)) return fail(ASTRO_ERR_ARG, "natural coordinates must lie in [-1, 1]");
    const int E = c->prep.numEls;
    DBuf<double> ds, de;
    ASTRO_TRY(ds.alloc(6 * (size_t)E));
    ASTRO_TRY(de.alloc(6 * (size_t)E));
    hipLaunchKernelGGL(k_brick8_stress, dim3(grid_for(E, 256)), dim3(256), 0, c->s, E, c->conn.p, c->X.p, c->elMat.p,
                       c->elScale.p, c->matC.p, c->rank.p, c->prep.numSolidNds, c->nodeDisp.p,
                       c->hasTExp ? (const double*)c->elTExp.p : nullptr, (const double*)c->nodeTemp.p, sp[0], sp[1],
                       sp[2], ds.p, de.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(stress, ds.p, sizeof(double) * 6 * (size_t)E, hipMemcpyDeviceToHost, c->s));
    HIP_TRY(hipMemcpyAsync(strain, de.p, sizeof(double) * 6 * (size_t)E, hipMemcpyDeviceToHost, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    return ASTRO_OK;
}

// Heat flux -k grad T and the temperature gradient of every element at the
// natural point sp for the current nodeTemp (r_getElFlux,
// Working/AStrO_r_elementEqns.f90:1746-1799).  flux, grad: host arrays of
// 3 numEls.  brick8 models with a thermal system.
int astro_ctx_element_flux(astro_ctx* c, const double* sp, double* flux, double* grad) {
    if (!c || !sp || !flux || !grad) return fail(ASTRO_ERR_ARG, "null argument");
    if (c->prep.shellMesh) return fail(ASTRO_ERR_UNSUPPORTED, "heat flux recovery covers brick8 elements");
    if (!c->thermalReady) return fail(ASTRO_ERR_ARG, "the thermal system has not been set up");
    for (int k = 0; k < 3; ++k)
        if (!(std::fabs(sp[k]) <= 1.0)) return fail(ASTRO_ERR_ARG, "natural coordinates must lie in [-1, 1]");
    const int E = c->prep.numEls;
    DBuf<double> df, dg;
    ASTRO_TRY(df.alloc(3 * (size_t)E));
    ASTRO_TRY(dg.alloc(3 * (size_t)E));
    hipLaunchKernelGGL(k_brick8_flux, dim3(grid_for(E, 256, INT_MAX)), dim3(256), 0, c->s, E, c->conn.p, c->X.p,
                       c->elCond.p, c->rank.p, c->nodeTemp.p, sp[0], sp[1], sp[2], df.p, dg.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(flux, df.p, sizeof(double) * 3 * (size_t)E, hipMemcpyDeviceToHost, c->s));
    HIP_TRY(hipMemcpyAsync(grad, dg.p, sizeof(double) * 3 * (size_t)E, hipMemcpyDeviceToHost, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    return ASTRO_OK;
}

// Co-rotational internal force of a shell4 model at the displacement u
// (host, elMatDim; null: the current nodeDisp): the shell branch of
// r_getElRuk with nonlinear geometry, Working/AStrO_r_elementEqns.f90:2245-2423.
// force: host elMatDim, the condensed element forces summed at the nodal DOFs;
// elEnergy: host numEls (or null), each element's strain energy.  iters chord
// steps bring the element's internal DOFs to equilibrium.
int astro_ctx_shell_internal_force(astro_ctx* c, const double* u, int32_t iters, double* force, double* elEnergy) {
    if (!c || !force) return fail(ASTRO_ERR_ARG, "null argument");
    if (!c->prep.shellMesh) return fail(ASTRO_ERR_UNSUPPORTED, "the co-rotational force covers shell4 elements");
    if (iters < 0 || iters > 64) return fail(ASTRO_ERR_ARG, "internal-DOF iterations must lie in [0, 64]");
    const int64_t n = c->prep.elMatDim;
    const int E = c->prep.numEls;
    DBuf<double> du, df, de;
    const double* ud = c->nodeDisp.p;
    if (u) {
        ASTRO_TRY(du.alloc((size_t)n));
        HIP_TRY(hipMemcpyAsync(du.p, u, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
        ud = du.p;
    }
    ASTRO_TRY(df.alloc((size_t)n));
    ASTRO_TRY(de.alloc((size_t)E));
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, c->s, df.p, n, 0.0);
    for (size_t k = 0; k + 1 < c->colorStart.size(); ++k) {
        const int cnt = c->colorStart[k + 1] - c->colorStart[k];
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(k_shell4_corot, dim3(cnt), dim3(64), 0, c->s, c->perm.p + c->colorStart[k], cnt,
                           c->conn.p, c->X.p, c->elABD.p, c->elAlpha.p, c->rank.p, c->prep.numSolidNds,
                           c->prep.numSectionNds, ud, 1e-30, (int)iters, df.p, de.p);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(force, df.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
    if (elEnergy) HIP_TRY(hipMemcpyAsync(elEnergy, de.p, sizeof(double) * E, hipMemcpyDeviceToHost, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    return ASTRO_OK;
}

// Relative step of the central differences in k_shell4_corot_tangent: the
// Hessian's truncation error grows as its square, its rounding error as
// 1e-16 |R| over it; both stay near 1e-10 of the tangent at this step.
static constexpr double kShellTangentStep = 1e-6;

// The tangent at u into the CSR values: co-rotational for shell4 models,
// Green-Lagrange for brick8 models (whose node-block operator is refreshed).
static int launch_nonlinear_tangent(astro_ctx* c, const double* u, int intIters) {
    const int64_t nnz = (int64_t)c->prep.elMatCols.size();
    hipLaunchKernelGGL(k_fill, dim3(grid_for(nnz, 256, 8192)), dim3(256), 0, c->s, c->csrVal.p, nnz, 0.0);
    for (size_t q = 0; q + 1 < c->colorStart.size(); ++q) {
        const int cnt = c->colorStart[q + 1] - c->colorStart[q];
        if (cnt <= 0) continue;
        if (c->prep.shellMesh)
            hipLaunchKernelGGL(k_shell4_corot_tangent, dim3(cnt), dim3(64), 0, c->s, c->perm.p + c->colorStart[q],
                               cnt, c->conn.p, c->X.p, c->elABD.p, c->elAlpha.p, c->rank.p, c->prep.numSolidNds,
                               c->prep.numSectionNds, u, 1e-30, intIters, kShellTangentStep, c->csrRange.p,
                               c->elPos.p, c->csrVal.p);
        else
            hipLaunchKernelGGL(k_brick8_tangent_gl, dim3(cnt), dim3(64), 0, c->s, c->perm.p + c->colorStart[q], cnt,
                               c->conn.p, c->X.p, c->elMat.p, c->elScale.p, c->matC.p, c->rank.p,
                               c->prep.numSolidNds, u, c->hasTExp ? (const double*)c->elTExp.p : nullptr,
                               (const double*)c->nodeTemp.p, c->csrRange.p, c->elPos.p, c->csrVal.p);
    }
    HIP_TRY(hipGetLastError());
    return c->prep.shellMesh ? ASTRO_OK : refresh_node_operator(c);
}

// load = -R(u): the co-rotational shell force (intIters chord steps on the
// internal DOFs; rv: elMatDim scratch) or the Green-Lagrange brick force
static void launch_minus_internal_force(astro_ctx* c, const double* u, int intIters, double* rv, double* load) {
    const int n = c->prep.elMatDim;
    hipStream_t s = c->s;
    if (c->prep.shellMesh) {
        hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, rv, (int64_t)n, 0.0);
        for (size_t q = 0; q + 1 < c->colorStart.size(); ++q) {
            const int cnt = c->colorStart[q + 1] - c->colorStart[q];
            if (cnt <= 0) continue;
            hipLaunchKernelGGL(k_shell4_corot, dim3(cnt), dim3(64), 0, s, c->perm.p + c->colorStart[q], cnt,
                               c->conn.p, c->X.p, c->elABD.p, c->elAlpha.p, c->rank.p, c->prep.numSolidNds,
                               c->prep.numSectionNds, u, 1e-30, intIters, rv, (double*)nullptr);
        }
        hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, (const double*)rv,
                           (const double*)rv, load, n, -1.0, 0.0, (const int*)nullptr);
        return;
    }
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, load, (int64_t)n, 0.0);
    for (size_t q = 0; q + 1 < c->colorStart.size(); ++q) {
        const int cnt = c->colorStart[q + 1] - c->colorStart[q];
        if (cnt <= 0) continue;
        // c1 = 0: no inertia term
        hipLaunchKernelGGL(k_brick8_residual, dim3(cnt), dim3(64), 0, s, c->perm.p + c->colorStart[q], cnt,
                           c->conn.p, c->X.p, c->elMat.p, c->elScale.p, c->matC.p, c->elRho.p, c->rank.p,
                           c->prep.numSolidNds, u, u, u, u, 0.0, 0.0, 0.0, 0.0, c->hasTExp ? (const double*)c->elTExp.p : nullptr,
                           (const double*)c->nodeTemp.p, load, (const int*)nullptr);
    }
}

// Tangent at the displacement u (host, elMatDim; null: the current nodeDisp)
// into the CSR values (elasticMat): getElasticSolnLoad(1)'s assembly with
// nonlinear geometry (Working/AStrO_commandFunctions.f90:260-351), the
// element matrix the derivative of the element's internal force at u
// (k_shell4_corot_tangent, k_brick8_tangent_gl).
int astro_ctx_nonlinear_tangent(astro_ctx* c, const double* u, int32_t intIters) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    if (intIters < 0 || intIters > 64) return fail(ASTRO_ERR_ARG, "internal-DOF iterations must lie in [0, 64]");
    const int64_t n = c->prep.elMatDim;
    DBuf<double> du;
    const double* ud = c->nodeDisp.p;
    if (u) {
        ASTRO_TRY(du.upload(u, (size_t)n, c->s));
        ud = du.p;
    }
    ASTRO_TRY(launch_nonlinear_tangent(c, ud, intIters));
    HIP_TRY(hipStreamSynchronize(c->s));
    return ASTRO_OK;
}

// Nonlinear static solve, stepSolve's elastic loop with nonLinearGeom
// (Working/AStrO_commandFunctions.f90:704-747): up to maxIt Newton steps
// u += du with (K(u) + C^T C) du = F - R(u) + C^T (rhs - C u), R the internal
// force at u (co-rotational shell4 with intIters chord steps on the internal
// DOFs, Green-Lagrange brick8) and K(u) its derivative, assembled and
// factored at every step as the reference does, until max|du| <= tol max|u|.
// The penalty rows are scaled once, from the tangent at rest.  Starts from the
// current nodeDisp; *iters: the steps taken.
int astro_ctx_nonlinear_solve(astro_ctx* c, int32_t maxIt, double tol, int32_t intIters, int32_t* iters) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    if (c->hasTExp && c->prep.shellMesh) return fail(ASTRO_ERR_UNSUPPORTED, "shell4 models carry no expansion");
    if (maxIt < 1 || !(tol > 0.0) || intIters < 0 || intIters > 64) return fail(ASTRO_ERR_ARG, "bad iteration limits");
    const int n = c->prep.elMatDim;
    hipStream_t s = c->s;
    ASTRO_TRY(astro_ctx_assemble_elastic(c, 0.0));
    ASTRO_TRY(astro_ctx_scale_mpc(c));
    DBuf<double> rv;
    ASTRO_TRY(rv.alloc((size_t)n));
    std::vector<double> hdu(n), hu(n);
    int k = 0;
    for (; k < maxIt; ++k) {
        ASTRO_TRY(launch_nonlinear_tangent(c, c->nodeDisp.p, intIters));
        ASTRO_TRY(astro_ctx_build_profile(c));
        ASTRO_TRY(astro_ctx_factor(c));
        // load = F - R(u) + C^T (rhs - C u)
        launch_minus_internal_force(c, c->nodeDisp.p, intIters, rv.p, c->load.p);
        const int nl = (int)c->loadDofs.n;
        if (nl > 0)
            hipLaunchKernelGGL(k_nodal_loads, dim3(grid_for(nl, 256)), dim3(256), 0, s, c->loadDofs.p,
                               c->loadVals.p, nl, 1.0, c->load.p, (const int*)nullptr);
        const int cd = c->prep.elMPCDim;
        if (cd > 0)
            hipLaunchKernelGGL(k_constraint_load, dim3(grid_for(cd, 256)), dim3(256), 0, s, c->mpcMat.p,
                               c->mpcCols.p, c->mpcRange.p, c->mpcRhs.p, cd, c->nodeDisp.p, c->load.p,
                               (const int*)nullptr);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, c->delDisp.p, (int64_t)n, 0.0);
        ASTRO_TRY(ctx_gmres(c, c->load.p, c->delDisp.p));
        hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, c->nodeDisp.p, c->delDisp.p,
                           c->nodeDisp.p, n, 1.0, 1.0, (const int*)nullptr);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(hdu.data(), c->delDisp.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(hu.data(), c->nodeDisp.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        double mdu = 0.0, mu = 0.0;
        for (int i = 0; i < n; ++i) {
            mdu = std::max(mdu, std::fabs(hdu[i]));
            mu = std::max(mu, std::fabs(hu[i]));
        }
        if (!std::isfinite(mdu)) return fail(ASTRO_ERR_NO_CONVERGE, "nonlinear solve diverged");
        if (mdu <= tol * mu) {
            ++k;
            break;
        }
    }
    if (iters) *iters = k;
    return ASTRO_OK;
}

// (dR/dT)^T lam for a displacement-sized lam (host, elMatDim) into out
// (host, numNodes, indexed by rank like nodeTemp): the coupling term of the
// thermal adjoint (solveDynamicAdjoint's dLdt -= dRdT^T dispAdj,
// Working/AStrO_commandFunctions.f90:960-999) with the element's intended
// dRdT (D10).  brick8 models with thermal expansion.
int astro_ctx_dRdT_product(astro_ctx* c, const double* lam, double* out) {
    if (!c || !lam || !out) return fail(ASTRO_ERR_ARG, "null argument");
    if (!c->hasTExp) return fail(ASTRO_ERR_UNSUPPORTED, "the model has no thermal expansion");
    const int n = c->prep.elMatDim, N = c->prep.numNodes;
    DBuf<double> dl, dout;
    ASTRO_TRY(dl.upload(lam, (size_t)n, c->s));
    ASTRO_TRY(dout.alloc((size_t)N));
    hipLaunchKernelGGL(k_fill, dim3(grid_for(N, 256, 4096)), dim3(256), 0, c->s, dout.p, (int64_t)N, 0.0);
    for (size_t k = 0; k + 1 < c->colorStart.size(); ++k) {
        const int cnt = c->colorStart[k + 1] - c->colorStart[k];
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(k_brick8_thermal_strain, dim3((cnt + 127) / 128), dim3(128), 0, c->s, 1,
                           c->perm.p + c->colorStart[k], cnt, c->conn.p, c->X.p, c->elMat.p, c->elScale.p, c->matC.p,
                           c->elTExp.p, c->rank.p, c->prep.numSolidNds, (const double*)nullptr,
                           (const double*)dl.p, 1.0, dout.p);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, dout.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    return ASTRO_OK;
}

int astro_ctx_elastic_soln_load(astro_ctx* c) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    ctx_minus_internal_force(c->s, c->csrVal, c->csrCols, c->csrRange, c->prep.elMatDim, c->nodeDisp.p, c->load.p);
    ASTRO_TRY(ctx_thermal_load(c, 1.0, c->load.p));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->s));
    return ASTRO_OK;
}

int astro_ctx_static_solve(astro_ctx* c) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    HIP_TRY(hipEventRecord(c->ev[0], c->s));
    ASTRO_TRY(astro_ctx_assemble_elastic(c, 0.0));
    HIP_TRY(hipEventRecord(c->ev[1], c->s));
    ASTRO_TRY(astro_ctx_scale_mpc(c));
    HIP_TRY(hipEventRecord(c->ev[2], c->s));
    ASTRO_TRY(astro_ctx_build_profile(c));
    HIP_TRY(hipEventRecord(c->ev[3], c->s));
    ASTRO_TRY(astro_ctx_factor(c));
    HIP_TRY(hipEventRecord(c->ev[4], c->s));
    ASTRO_TRY(astro_ctx_solve(c));
    HIP_TRY(hipEventRecord(c->ev[5], c->s));
    HIP_TRY(hipEventSynchronize(c->ev[5]));
    for (int k = 0; k < 5; ++k) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev[k], c->ev[k + 1]));
        c->phaseMs[k] = ms;
    }
    return ASTRO_OK;
}

// Newmark dynamics: the dynamic branch of solve (Working/AStrO_commandFunctions.f90:
// 762-810).  Set-up assembles the tangent K + M / (beta dt^2) at u = 0, scales
// the penalty rows by it, builds the profile and factors it once; the state
// starts at rest.
int astro_ctx_dynamic_init(astro_ctx* c, double beta, double gamma, double dt) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    if (!(beta > 0.0) || !(dt > 0.0) || !(gamma >= 0.0))
        return fail(ASTRO_ERR_ARG, "Newmark parameters need beta > 0, gamma >= 0, dt > 0");
    if (c->prep.shellMesh) return fail(ASTRO_ERR_UNSUPPORTED, "Newmark dynamics of shell4 meshes is not part of the device path");
    if (c->hasTExp) return fail(ASTRO_ERR_UNSUPPORTED, "Newmark dynamics with thermal expansion is not part of the device path");
    const int n = c->prep.elMatDim;
    c->nmBeta = beta;
    c->nmGamma = gamma;
    c->delT = dt;
    for (DBuf<double>* b : {&c->prevDisp, &c->prevVel, &c->prevAcc, &c->nodeVel, &c->nodeAcc})
        if (b->n != (size_t)n) ASTRO_TRY(b->alloc((size_t)n));
    for (DBuf<double>* b : {&c->prevDisp, &c->prevVel, &c->prevAcc, &c->nodeVel, &c->nodeAcc, &c->nodeDisp, &c->delDisp})
        hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, c->s, b->p, (int64_t)n, 0.0);
    ASTRO_TRY(astro_ctx_assemble_elastic(c, 1.0 / (beta * dt * dt)));
    ASTRO_TRY(astro_ctx_scale_mpc(c));
    ASTRO_TRY(astro_ctx_build_profile(c));
    ASTRO_TRY(astro_ctx_factor(c));
    c->dynSteps = c->dynGmresSteps = 0;
    c->dynGraphSteps = c->dynEagerSteps = c->dynGraphs = 0;
    c->dynMs = 0.0;
    c->dynReady = true;
    return ASTRO_OK;
}

// One Newmark step (stepSolve :704-745 with maxNLIt = 1, then :792-806):
// load = loadScale * F - sum_e R_e(u) + C^T (rhs - C u); GMRES from the
// previous increment; u += du; acceleration, velocity and the previous state.
int astro_ctx_dynamic_step(astro_ctx* c, double loadScale) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    if (!c->dynReady) return fail(ASTRO_ERR_ARG, "astro_ctx_dynamic_init has not been called");
    const int n = c->prep.elMatDim, nS = c->prep.numSolidNds;
    const double beta = c->nmBeta, dt = c->delT, c1 = 0.5 / beta, c2 = 2.0 / (dt * dt);
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, c->s, c->load.p, (int64_t)n, 0.0);
    for (size_t k = 0; k + 1 < c->colorStart.size(); ++k) {
        const int cnt = c->colorStart[k + 1] - c->colorStart[k];
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(k_brick8_residual, dim3(cnt), dim3(64), 0, c->s, c->perm.p + c->colorStart[k], cnt,
                           c->conn.p, c->X.p, c->elMat.p, c->elScale.p, c->matC.p, c->elRho.p, c->rank.p, nS,
                           c->nodeDisp.p, c->prevDisp.p, c->prevVel.p, c->prevAcc.p, c1, c2, dt, beta, (const double*)nullptr, (const double*)nullptr, c->load.p,
                           (const int*)nullptr);
    }
    const int nl = (int)c->loadDofs.n;
    if (nl > 0)
        hipLaunchKernelGGL(k_nodal_loads, dim3(grid_for(nl, 256)), dim3(256), 0, c->s, c->loadDofs.p, c->loadVals.p,
                           nl, loadScale, c->load.p, (const int*)nullptr);
    const int cd = c->prep.elMPCDim;
    if (cd > 0)
        hipLaunchKernelGGL(k_constraint_load, dim3(grid_for(cd, 256)), dim3(256), 0, c->s, c->mpcMat.p, c->mpcCols.p,
                           c->mpcRange.p, c->mpcRhs.p, cd, c->nodeDisp.p, c->load.p, (const int*)nullptr);
    HIP_TRY(hipGetLastError());
    ASTRO_TRY(ctx_gmres(c, c->load.p, c->delDisp.p));
    hipLaunchKernelGGL(k_newmark_update, dim3(grid_for(n, 256, 4096)), dim3(256), 0, c->s, n, c->nodeDisp.p,
                       c->delDisp.p, c->prevDisp.p, c->prevVel.p, c->prevAcc.p, c->nodeVel.p, c->nodeAcc.p, c1, c2, dt,
                       beta, c->nmGamma, (const int*)nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->s));
    c->dynSteps += 1;
    c->dynGmresSteps += c->lastSolveSteps;
    return ASTRO_OK;
}

// ---------------------------------------------------------------------------
// Newmark steps as HIP graphs.  One graph holds a whole step: the load
// (the eager step's kernels, the load factor read from a device table at the
// device step index), the starting residual, one GMRES cycle of at most K
// Arnoldi steps whose Givens rotations and least squares run on the device
// (k_gm_*), the judge (the true residual after the cycle against the
// cycle's target; with learned blocks against rtol |P^-1 r0| of the plain
// LDL^T, the eager path's judge), the Newmark update and the step index.
// Arnoldi steps past convergence are skipped through the device flag
// active; a step that misses its target clears run, so that the steps
// queued behind it do nothing, and the host finishes that step eagerly (a
// plain LDL^T GMRES to the same absolute target, from the cycle's iterate)
// before it replays again.  The arithmetic of a replayed step is the eager
// step's, kernel for kernel; only the rotations' rounding (device fma and
// hypot) differs.

// Records one step into the stream s, which is capturing.  P: the LDL^T
// profile; Q: the cycle's preconditioner (P itself, or the two-stage form
// with judge = true).
extern "C++" {
template <class Prof, class Pc>
static int record_newmark_step(astro_ctx* c, Prof& P, Pc& Q, bool judge, int K, hipStream_t s) {
    Gmres& g = c->gm;
    const int n = c->prep.elMatDim, nS = c->prep.numSolidNds, m = g.m;
    const double beta = c->nmBeta, dt = c->delT, c1 = 0.5 / beta, c2 = 2.0 / (dt * dt), rtol = c->opt.rtol;
    int* gi = c->gmi.p;
    double* gd = c->gmd.p;
    double* nrm = c->gmNrm.p;  // starting (cycle), starting (judge), after the cycle
    const int* run = gi + GMI_RUN;
    const int* act = gi + GMI_ACTIVE;
    hipLaunchKernelGGL(k_fill_gated, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, c->load.p, (int64_t)n, 0.0, run);
    for (size_t k = 0; k + 1 < c->colorStart.size(); ++k) {
        const int cnt = c->colorStart[k + 1] - c->colorStart[k];
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(k_brick8_residual, dim3(cnt), dim3(64), 0, s, c->perm.p + c->colorStart[k], cnt,
                           c->conn.p, c->X.p, c->elMat.p, c->elScale.p, c->matC.p, c->elRho.p, c->rank.p, nS,
                           c->nodeDisp.p, c->prevDisp.p, c->prevVel.p, c->prevAcc.p, c1, c2, dt, beta, (const double*)nullptr, (const double*)nullptr, c->load.p, run);
    }
    const int nl = (int)c->loadDofs.n;
    if (nl > 0)
        hipLaunchKernelGGL(k_nodal_loads_dev, dim3(grid_for(nl, 256)), dim3(256), 0, s, c->loadDofs.p,
                           c->loadVals.p, nl, (const double*)c->dynScales.p, (const int*)(gi + GMI_STEP), c->load.p,
                           run);
    const int cd = c->prep.elMPCDim;
    if (cd > 0)
        hipLaunchKernelGGL(k_constraint_load, dim3(grid_for(cd, 256)), dim3(256), 0, s, c->mpcMat.p, c->mpcCols.p,
                           c->mpcRange.p, c->mpcRhs.p, cd, c->nodeDisp.p, c->load.p, run);
    // starting residual (double-double), preconditioned
    ASTRO_TRY(c->op.residual(c->load.p, c->delDisp.p, g.t2.p, s, run));
    if (judge) {
        ASTRO_TRY(P.apply(g.t2.p, g.r.p, s, run));
        g.multidot(g.r.p, 1, g.r.p, nrm + 1, s, run);
        hipLaunchKernelGGL(k_gm_judge_target, dim3(1), dim3(64), 0, s, (const double*)(nrm + 1), rtol, gd,
                           (const int*)gi);
    }
    ASTRO_TRY(Q.apply(g.t2.p, g.r.p, s, run));
    g.multidot(g.r.p, 1, g.r.p, nrm, s, run);
    hipLaunchKernelGGL(k_gm_begin, dim3(1), dim3(64), 0, s, (const double*)nrm, rtol, 0.0, gd, gi, m);
    hipLaunchKernelGGL(k_scale_by, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, (const double*)g.r.p, g.Vcol(0), n,
                       (const double*)(gd + GMD_INVRES), act);
    double* h1 = g.hdev.p;
    double* h2 = g.hdev.p + (m + 8);
    double* hn = g.hdev.p + 2 * (m + 8);
    for (int j = 0; j < K; ++j) {
        double* w = g.Vcol(j + 1);
        ASTRO_TRY(c->op.apply(g.Vcol(j), g.t.p, s, act));
        ASTRO_TRY(Q.apply(g.t.p, w, s, act));
        g.multidot(g.V.p, j + 1, w, h1, s, act);
        hipLaunchKernelGGL(k_sub_combination, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, (const double*)g.V.p,
                           (int64_t)n, j + 1, (const double*)h1, w, n, -1.0, act);
        g.multidot(g.V.p, j + 1, w, h2, s, act);
        hipLaunchKernelGGL(k_sub_combination, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, (const double*)g.V.p,
                           (int64_t)n, j + 1, (const double*)h2, w, n, -1.0, act);
        g.multidot(w, 1, w, hn, s, act);
        hipLaunchKernelGGL(k_gm_givens, dim3(1), dim3(64), 0, s, j, (const double*)h1, (const double*)h2,
                           (const double*)hn, gd, gi, m);
        hipLaunchKernelGGL(k_scale_by, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, (const double*)w, w, n,
                           (const double*)(gd + GMD_INVH), act);
    }
    hipLaunchKernelGGL(k_gm_lsq, dim3(1), dim3(64), 0, s, gd, (const int*)gi, m);
    hipLaunchKernelGGL(k_sub_combination_dn, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, (const double*)g.V.p,
                       (int64_t)n, (const int*)(gi + GMI_K), (const double*)(gd + gmd_y(m)), c->delDisp.p, n, 1.0,
                       run);
    // the judge
    ASTRO_TRY(c->op.residual(c->load.p, c->delDisp.p, g.t2.p, s, run));
    if (judge)
        ASTRO_TRY(P.apply(g.t2.p, g.r.p, s, run));
    else
        ASTRO_TRY(Q.apply(g.t2.p, g.r.p, s, run));
    g.multidot(g.r.p, 1, g.r.p, nrm + 2, s, run);
    hipLaunchKernelGGL(k_gm_check, dim3(1), dim3(64), 0, s, (const double*)(nrm + 2), gd, gi,
                       judge ? (int)GMD_JUDGE : (int)GMD_TARGET);
    hipLaunchKernelGGL(k_newmark_update, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, n, c->nodeDisp.p,
                       c->delDisp.p, c->prevDisp.p, c->prevVel.p, c->prevAcc.p, c->nodeVel.p, c->nodeAcc.p, c1, c2, dt,
                       beta, c->nmGamma, run);
    hipLaunchKernelGGL(k_gm_advance, dim3(1), dim3(64), 0, s, gi);
    HIP_TRY(hipGetLastError());
    return ASTRO_OK;
}
}  // extern "C++"

// The Arnoldi cap of a step graph for a recent need of s steps.
static int graph_cap(int s, int m) { return std::min(m, std::max(4, s + std::max(3, s / 4))); }

// Finishes step f eagerly after its replay missed the target: GMRES with the
// plain LDL^T from the cycle's iterate to the replay's absolute target, then
// the Newmark update.  cycleSteps: the replay's Arnoldi steps.
static int finish_step_eagerly(astro_ctx* c, int cycleSteps) {
    double gdh[8];
    HIP_TRY(hipMemcpyAsync(gdh, c->gmd.p, sizeof(gdh), hipMemcpyDeviceToHost, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    const double target = c->gnnOn ? gdh[GMD_JUDGE] : gdh[GMD_TARGET];
    ASTRO_TRY(with_prof(c, [&](auto& P) -> int {
        return c->gm.solve(c->op, P, c->load.p, c->delDisp.p, 0.0, target, c->opt.maxCycles, c->s);
    }));
    if (c->gnnOn) c->gnnFallbackSteps += c->gm.steps;
    const int n = c->prep.elMatDim;
    const double beta = c->nmBeta, dt = c->delT, c1 = 0.5 / beta, c2 = 2.0 / (dt * dt);
    hipLaunchKernelGGL(k_newmark_update, dim3(grid_for(n, 256, 4096)), dim3(256), 0, c->s, n, c->nodeDisp.p,
                       c->delDisp.p, c->prevDisp.p, c->prevVel.p, c->prevAcc.p, c->nodeVel.p, c->nodeAcc.p, c1, c2, dt,
                       beta, c->nmGamma, (const int*)nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->s));
    c->dynSteps += 1;
    c->dynEagerSteps += 1;
    c->lastSolveSteps = cycleSteps + c->gm.steps;
    c->dynGmresSteps += c->lastSolveSteps;
    return ASTRO_OK;
}

// Steps 1..nSteps-1 of a run from graphs (step 0 runs eagerly: it readies
// the sweeps and gives the first Arnoldi count).  The host keeps one step
// of lookahead: it launches step k + 1 before it reads step k's snapshot of
// gmi.  Graphs live for one run (they hold the context's buffers and the
// preconditioner's state as recorded) and are cached by their cap K.
static int dynamic_run_graphs(astro_ctx* c, int nSteps, const double* loadScale) {
    ASTRO_TRY(astro_ctx_dynamic_step(c, loadScale[0]));
    const int m = c->gm.m;
    if (c->gmd.n < gmd_size(m)) ASTRO_TRY(c->gmd.alloc(gmd_size(m)));
    if (c->gmi.n < (size_t)GMI_COUNT) ASTRO_TRY(c->gmi.alloc(GMI_COUNT));
    if (c->gmNrm.n < 3) ASTRO_TRY(c->gmNrm.alloc(3));
    ASTRO_TRY(c->dynScales.upload(loadScale, (size_t)nSteps, c->s));
    if (!c->gmiHost) HIP_TRY(hipHostMalloc((void**)&c->gmiHost, sizeof(int) * 2 * GMI_COUNT));
    for (auto& e : c->gmEv)
        if (!e) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    struct Execs {
        std::map<int, hipGraphExec_t> x;
        ~Execs() {
            for (auto& kv : x) (void)hipGraphExecDestroy(kv.second);
        }
    } execs;
    auto graph_for = [&](int K, hipGraphExec_t& out) -> int {
        auto it = execs.x.find(K);
        if (it != execs.x.end()) {
            out = it->second;
            return ASTRO_OK;
        }
        HIP_TRY(hipStreamSynchronize(c->s));
        HIP_TRY(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
        const int rc = with_prof(c, [&](auto& P) -> int {
            if (!c->gnnOn) return record_newmark_step(c, P, P, false, K, c->s);
            TwoStagePrec<std::decay_t<decltype(P)>> T{P,        c->op, c->gnnB.p,  c->nodeDofs.p, c->prep.numNodes,
                                                      c->prep.elMatDim, c->gnnT1.p, c->gnnT2.p};
            return record_newmark_step(c, P, T, true, K, c->s);
        });
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->s, &g);
        if (rc != ASTRO_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        HIP_TRY(e);
        hipGraphExec_t x = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_TRY(ei);
        execs.x[K] = x;
        c->dynGraphs += 1;
        out = x;
        return ASTRO_OK;
    };
    auto set_state = [&](int step) -> int {
        int* h = c->gmiHost;  // both snapshot slots are idle here
        for (int q = 0; q < GMI_COUNT; ++q) h[q] = 0;
        h[GMI_RUN] = 1;
        h[GMI_STEP] = step;
        h[GMI_FAILED] = -1;
        HIP_TRY(hipMemcpyAsync(c->gmi.p, h, sizeof(int) * GMI_COUNT, hipMemcpyHostToDevice, c->s));
        HIP_TRY(hipStreamSynchronize(c->s));
        return ASTRO_OK;
    };
    int K = graph_cap(c->lastSolveSteps, m);
    std::vector<int> recent;  // Arnoldi counts of the last replayed steps
    int k = 1, pending = -1;
    ASTRO_TRY(set_state(k));
    while (k < nSteps || pending >= 0) {
        int launched = -1;
        if (k < nSteps) {
            hipGraphExec_t x = nullptr;
            ASTRO_TRY(graph_for(K, x));
            HIP_TRY(hipGraphLaunch(x, c->s));
            HIP_TRY(hipMemcpyAsync(c->gmiHost + GMI_COUNT * (k & 1), c->gmi.p, sizeof(int) * GMI_COUNT,
                                   hipMemcpyDeviceToHost, c->s));
            HIP_TRY(hipEventRecord(c->gmEv[k & 1], c->s));
            launched = k++;
        }
        if (pending >= 0) {
            HIP_TRY(hipEventSynchronize(c->gmEv[pending & 1]));
            const int* h = c->gmiHost + GMI_COUNT * (pending & 1);
            if (!h[GMI_RUN]) {
                if (h[GMI_FAILED] != pending) return fail(ASTRO_ERR_DEVICE, "step graph state out of order");
                // the launch queued behind the missed step did nothing
                HIP_TRY(hipStreamSynchronize(c->s));
                const int cycleSteps = h[GMI_K];
                ASTRO_TRY(finish_step_eagerly(c, cycleSteps));
                K = graph_cap(cycleSteps + c->gm.steps, m);
                recent.clear();
                k = pending + 1;
                pending = -1;
                if (k < nSteps) ASTRO_TRY(set_state(k));
                continue;
            }
            c->dynSteps += 1;
            c->dynGraphSteps += 1;
            c->lastSolveSteps = h[GMI_K];
            c->dynGmresSteps += h[GMI_K];
            recent.push_back(h[GMI_K]);
            if (recent.size() > 16) recent.erase(recent.begin());
            // a cap far above the recent need costs skipped launches: narrow it
            const int need = *std::max_element(recent.begin(), recent.end());
            if (recent.size() == 16 && graph_cap(need, m) * 2 < K) K = graph_cap(need, m);
        }
        pending = launched;
    }
    c->dynGraphCap = K;
    return with_prof(c, [&](auto& P) -> int { return P.check_timeout(c->s); });
}

// nSteps Newmark steps with the load factors loadScale[0..nSteps-1]; the
// elapsed time of the run is kept as scalar "dynMs".  With stepGraphs = 1
// the steps after the first replay HIP graphs (dynamic_run_graphs).
int astro_ctx_dynamic_run(astro_ctx* c, int nSteps, const double* loadScale) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    if (nSteps < 0 || (nSteps > 0 && !loadScale)) return fail(ASTRO_ERR_ARG, "bad step count or load factors");
    if (!c->dynReady) return fail(ASTRO_ERR_ARG, "astro_ctx_dynamic_init has not been called");
    bool timing = false;
    with_prof(c, [&](auto& P) -> int {
        timing = P.timing;
        return ASTRO_OK;
    });
    HIP_TRY(hipEventRecord(c->ev[6], c->s));
    if (c->opt.stepGraphs && !timing && nSteps > 1) {
        ASTRO_TRY(dynamic_run_graphs(c, nSteps, loadScale));
    } else {
        for (int k = 0; k < nSteps; ++k) ASTRO_TRY(astro_ctx_dynamic_step(c, loadScale[k]));
    }
    HIP_TRY(hipEventRecord(c->ev[7], c->s));
    HIP_TRY(hipEventSynchronize(c->ev[7]));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev[6], c->ev[7]));
    c->dynMs = ms;
    return ASTRO_OK;
}

// Design variables for complex-step sensitivities of a shell4 model: element
// e's layer dvLayer thickness changes by elCoef[e] * d[elDV[e]] (elDV < 0: no
// variable; elCoef null: 1).  h is the complex step.  The imaginary parts of
// the perturbed layups are formed here, once.
int astro_ctx_set_design(astro_ctx* c, int nDV, const int32_t* elDV, const double* elCoef, int dvLayer, double h) {
    if (!c || !elDV) return fail(ASTRO_ERR_ARG, "null argument");
    if (!c->prep.shellMesh) return fail(ASTRO_ERR_UNSUPPORTED, "design sensitivities cover shell4 layer thickness");
    if (nDV < 1 || dvLayer < 0 || dvLayer >= c->nLayers || !(h > 0.0))
        return fail(ASTRO_ERR_ARG, "bad design-variable count, layer or step");
    const int E = c->prep.numEls;
    std::vector<double> im(81 * (size_t)E, 0.0), re(81);
    c->hElDV.assign(elDV, elDV + E);
    for (int e = 0; e < E; ++e) {
        const int k = elDV[e];
        if (k >= nDV) return fail(ASTRO_ERR_ARG, "design variable index out of range");
        if (k < 0) continue;
        const double coef = elCoef ? elCoef[e] : 1.0;
        shell_section_abd_cs(c->nLayers, &c->hLayers[6 * (size_t)c->nLayers * e], c->hZOff[e], dvLayer, h * coef,
                             re.data(), &im[81 * (size_t)e]);
    }
    c->nDV = nDV;
    c->csH = h;
    c->csb.ready = false;
    ASTRO_TRY(c->elDV.upload(c->hElDV, c->s));
    ASTRO_TRY(c->elABDim.upload(im, c->s));
    ASTRO_TRY(c->dRdD.alloc((size_t)c->prep.elMatDim * nDV));
    ASTRO_TRY(c->gPart.alloc((size_t)E));
    HIP_TRY(hipStreamSynchronize(c->s));
    return ASTRO_OK;
}

static int ctx_sensitivities(astro_ctx* c) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    if (c->nDV < 1) return fail(ASTRO_ERR_ARG, "astro_ctx_set_design has not been called");
    const int64_t n = c->prep.elMatDim;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n * c->nDV, 256, 8192)), dim3(256), 0, c->s, c->dRdD.p, n * c->nDV,
                       0.0);
    for (size_t k = 0; k + 1 < c->colorStart.size(); ++k) {
        const int cnt = c->colorStart[k + 1] - c->colorStart[k];
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(k_shell4_dtangent_cs, dim3(cnt), dim3(64), 0, c->s, c->perm.p + c->colorStart[k], cnt,
                           c->conn.p, c->X.p, c->elABD.p, c->elABDim.p, c->elAlpha.p, c->rank.p, c->prep.numSolidNds,
                           c->prep.numSectionNds, c->nodeDisp.p, c->elDV.p, 1.0 / c->csH, n, c->dRdD.p, c->gPart.p,
                           (double*)nullptr, (int*)nullptr);
    }
    HIP_TRY(hipGetLastError());
    return ASTRO_OK;
}

// dR/dd at the current displacement, all design variables in one launch per
// element colour: column k = sum over the elements of variable k of
// Im(Kc(t + i h coef) u_e) / h.  Working/AStrO_commandFunctions.f90:1172-1375.
int astro_ctx_dRdD(astro_ctx* c, double* dRdD) {
    if (!dRdD) return fail(ASTRO_ERR_ARG, "null argument");
    ASTRO_TRY(ctx_sensitivities(c));
    HIP_TRY(hipMemcpyAsync(dRdD, c->dRdD.p, sizeof(double) * c->dRdD.n, hipMemcpyDeviceToHost, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    return ASTRO_OK;
}

// Gradient of the compliance C = F . u (loads independent of d) by the
// adjoint lambda = u: dC/dd_k = -u . dR/dd_k, summed per variable over its
// elements in element order.  Working/AStrO_commandFunctions.f90:1377-1466.
int astro_ctx_compliance_gradient(astro_ctx* c, double* grad) {
    if (!grad) return fail(ASTRO_ERR_ARG, "null argument");
    ASTRO_TRY(ctx_sensitivities(c));
    std::vector<double> g(c->prep.numEls);
    HIP_TRY(hipMemcpyAsync(g.data(), c->gPart.p, sizeof(double) * g.size(), hipMemcpyDeviceToHost, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    for (int k = 0; k < c->nDV; ++k) grad[k] = 0.0;
    for (size_t e = 0; e < g.size(); ++e)
        if (c->hElDV[e] >= 0) grad[c->hElDV[e]] += g[e];
    return ASTRO_OK;
}

// ---------------------------------------------------------------------------
// Design variables, objective terms and the adjoint gradient (brick8).
//
// set_design_vars turns the reference's maps (elToD over elements, ndToD over
// nodes) into (element, variable) pairs: for every element, each variable
// that moves one of its properties or node coordinates, with a direction
// record holding d(property)/d(variable) (the coefficients, summed where a
// variable is listed twice).  Pairs are sorted by element colour, then
// variable, then element, so that one launch per colour can add pair
// contributions into a dense dR/dd without two writers on one entry.
int astro_ctx_set_design_vars(astro_ctx* c, const astro_design* d) {
    if (!c || !d) return fail(ASTRO_ERR_ARG, "null argument");
    if (c->prep.shellMesh)
        return fail(ASTRO_ERR_UNSUPPORTED,
                    "astro_ctx_set_design_vars covers brick8 models (shell4 layer thickness: astro_ctx_set_design)");
    if (c->hEprops.empty()) return fail(ASTRO_ERR_ARG, "the model carries no engineering constants");
    const int nDV = d->numDV;
    if (nDV < 1 || !d->category || !d->component || !(d->h > 0.0))
        return fail(ASTRO_ERR_ARG, "bad design-variable count, arrays or step");
    if ((d->elToDRange && (!d->elToD)) || (d->ndToDRange && !d->ndToD))
        return fail(ASTRO_ERR_ARG, "a design map has a range array but no variable ids");
    const int E = c->prep.numEls, N = c->prep.numNodes;
    auto& V = c->dvs;
    V.ready = false;
    V.cat.assign(d->category, d->category + nDV);
    V.comp.assign(d->component, d->component + nDV);
    for (int k = 0; k < nDV; ++k) {
        int maxc = 0;
        switch (V.cat[k]) {
            case ASTRO_DV_MODULUS:
            case ASTRO_DV_POISSON:
            case ASTRO_DV_SHEAR:
            case ASTRO_DV_NODE_COORD: maxc = 3; break;
            case ASTRO_DV_THERMAL_COND:
            case ASTRO_DV_NODE_LOAD: maxc = 6; break;
            case ASTRO_DV_DENSITY: maxc = INT_MAX; break;
            default: return fail(ASTRO_ERR_ARG, "unknown design-variable category " + std::to_string(V.cat[k]));
        }
        if (V.cat[k] != ASTRO_DV_DENSITY && (V.comp[k] < 1 || V.comp[k] > maxc))
            return fail(ASTRO_ERR_ARG, "design variable " + std::to_string(k + 1) + ": component out of range");
    }
    if (d->elToDRange && (d->elToDRange[0] != 0 || d->elToDRange[E] < d->elToDRange[0]))
        return fail(ASTRO_ERR_ARG, "elToDRange must start at 0 and not decrease");
    if (d->ndToDRange && (d->ndToDRange[0] != 0 || d->ndToDRange[N] < d->ndToDRange[0]))
        return fail(ASTRO_ERR_ARG, "ndToDRange must start at 0 and not decrease");
    std::vector<int> slot(nDV, -1);
    std::vector<std::pair<int, std::vector<double>>> loc;
    struct Rec {
        int color, k, e;
        size_t at;
    };
    std::vector<Rec> recs;
    std::vector<double> dirs;  // DVD_STRIDE per record, in discovery order
    for (int e = 0; e < E; ++e) {
        loc.clear();
        auto get = [&](int k) -> double* {
            if (slot[k] < 0) {
                slot[k] = (int)loc.size();
                loc.emplace_back(k, std::vector<double>(DVD_STRIDE, 0.0));
            }
            return loc[slot[k]].second.data();
        };
        if (d->elToDRange) {
            for (int j = d->elToDRange[e]; j < d->elToDRange[e + 1]; ++j) {
                const int k = d->elToD[j] - 1;
                if (k < 0 || k >= nDV) return fail(ASTRO_ERR_ARG, "elToD holds a variable id out of range");
                const double cf = d->elToCoef ? d->elToCoef[j] : 1.0;
                const int cp = V.comp[k] - 1;
                switch (V.cat[k]) {
                    case ASTRO_DV_MODULUS: get(k)[DVD_EP + cp] += cf; break;
                    case ASTRO_DV_POISSON: get(k)[DVD_EP + 3 + cp] += cf; break;
                    case ASTRO_DV_SHEAR: get(k)[DVD_EP + 6 + cp] += cf; break;
                    case ASTRO_DV_DENSITY: get(k)[DVD_RHO] += cf; break;
                    case ASTRO_DV_THERMAL_COND: get(k)[DVD_COND + cp] += cf; break;
                    default: break;  // node categories act through the node map only
                }
            }
        }
        if (d->ndToDRange) {
            for (int a = 0; a < 8; ++a) {
                const int nd = c->hConn[8 * (size_t)e + a] - 1;
                for (int j = d->ndToDRange[nd]; j < d->ndToDRange[nd + 1]; ++j) {
                    const int k = d->ndToD[j] - 1;
                    if (k < 0 || k >= nDV) return fail(ASTRO_ERR_ARG, "ndToD holds a variable id out of range");
                    if (V.cat[k] == ASTRO_DV_NODE_COORD)
                        get(k)[DVD_X + 3 * a + V.comp[k] - 1] += d->ndToCoef ? d->ndToCoef[j] : 1.0;
                }
            }
        }
        for (auto& kv : loc) {
            slot[kv.first] = -1;
            bool any = false;
            for (double x : kv.second) any = any || x != 0.0;
            if (!any) continue;
            recs.push_back({c->elColor[e], kv.first, e, dirs.size()});
            dirs.insert(dirs.end(), kv.second.begin(), kv.second.end());
        }
    }
    std::sort(recs.begin(), recs.end(), [](const Rec& a, const Rec& b) {
        if (a.color != b.color) return a.color < b.color;
        if (a.k != b.k) return a.k < b.k;
        return a.e < b.e;
    });
    const int nP = (int)recs.size();
    const int nColors = (int)c->colorStart.size() - 1;
    V.pairEl.resize(nP);
    V.pairDV.resize(nP);
    V.colorPairStart.assign(nColors + 1, 0);
    std::vector<double> dsorted((size_t)nP * DVD_STRIDE);
    for (int p = 0; p < nP; ++p) {
        V.pairEl[p] = recs[p].e;
        V.pairDV[p] = recs[p].k;
        V.colorPairStart[recs[p].color + 1]++;
        std::copy(dirs.begin() + recs[p].at, dirs.begin() + recs[p].at + DVD_STRIDE,
                  dsorted.begin() + (size_t)p * DVD_STRIDE);
    }
    for (int q = 0; q < nColors; ++q) V.colorPairStart[q + 1] += V.colorPairStart[q];
    V.elPairStart.assign(E + 1, 0);
    for (int p = 0; p < nP; ++p) V.elPairStart[V.pairEl[p] + 1]++;
    for (int e = 0; e < E; ++e) V.elPairStart[e + 1] += V.elPairStart[e];
    V.elPairList.assign(nP, 0);
    {
        std::vector<int> fill(V.elPairStart.begin(), V.elPairStart.end() - 1);
        for (int p = 0; p < nP; ++p) V.elPairList[fill[V.pairEl[p]]++] = p;
    }
    // node loads: R = K u - F(d), F(dof) += coef d (r_getNodeDLoad)
    V.loadDof.clear();
    V.loadDV.clear();
    V.loadCoef.clear();
    if (d->ndToDRange) {
        for (int nd = 0; nd < N; ++nd)
            for (int j = d->ndToDRange[nd]; j < d->ndToDRange[nd + 1]; ++j) {
                const int k = d->ndToD[j] - 1;
                if (V.cat[k] != ASTRO_DV_NODE_LOAD) continue;
                const int dof = c->prep.nDofIndex[6 * (size_t)nd + V.comp[k] - 1];
                if (dof <= 0) continue;
                V.loadDof.push_back(dof - 1);
                V.loadDV.push_back(k);
                V.loadCoef.push_back(d->ndToCoef ? d->ndToCoef[j] : 1.0);
            }
    }
    V.nDV = nDV;
    V.h = d->h;
    ASTRO_TRY(V.dPairEl.upload(V.pairEl, c->s));
    ASTRO_TRY(V.dPairDV.upload(V.pairDV, c->s));
    ASTRO_TRY(V.dDir.upload(dsorted, c->s));
    ASTRO_TRY(V.dEprops.upload(c->hEprops, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    V.ready = true;
    return ASTRO_OK;
}

static int dv_require(astro_ctx* c) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    if (!c->dvs.ready) return fail(ASTRO_ERR_ARG, "astro_ctx_set_design_vars has not been called");
    return ASTRO_OK;
}

// Pair products of the elastic (force) or thermal (flux) kernel: with col,
// the dense n x nDV derivative (one launch per colour); with lam, per-pair
// lambda . f into out (one launch).
static int dv_pairs_launch(astro_ctx* c, bool thermal, const double* state, const double* lam, double* out,
                           double* col, int64_t n, const double* acc = nullptr) {
    auto& V = c->dvs;
    auto launch = [&](int p0, int cnt) {
        if (cnt <= 0) return;
        if (thermal)
            hipLaunchKernelGGL(k_dv_pair_flux, dim3(grid_for(cnt, 128, INT_MAX)), dim3(128), 0, c->s, cnt,
                               V.dPairEl.p + p0, V.dPairDV.p + p0, V.dDir.p + (size_t)p0 * DVD_STRIDE, c->conn.p,
                               c->X.p, c->elCond.p, c->rank.p, state, lam, V.h, out ? out + p0 : nullptr, col, n);
        else
            hipLaunchKernelGGL(k_dv_pair_force, dim3(grid_for(cnt, 128, INT_MAX)), dim3(128), 0, c->s, cnt,
                               V.dPairEl.p + p0, V.dPairDV.p + p0, V.dDir.p + (size_t)p0 * DVD_STRIDE, c->conn.p,
                               c->X.p, V.dEprops.p, c->rank.p, c->prep.numSolidNds, state, acc,
                               (const double*)c->elRho.p, lam, V.h, out ? out + p0 : nullptr, col, n);
    };
    if (col) {
        for (size_t q = 0; q + 1 < V.colorPairStart.size(); ++q)
            launch(V.colorPairStart[q], V.colorPairStart[q + 1] - V.colorPairStart[q]);
    } else {
        launch(0, (int)V.pairEl.size());
    }
    HIP_TRY(hipGetLastError());
    return ASTRO_OK;
}

int astro_ctx_dRdD_vars(astro_ctx* c, double* dRdD) {
    ASTRO_TRY(dv_require(c));
    if (c->hasTExp) return fail(ASTRO_ERR_UNSUPPORTED, "design derivatives of a model with thermal expansion");
    if (!dRdD) return fail(ASTRO_ERR_ARG, "null argument");
    const int64_t n = c->prep.elMatDim, cols = c->dvs.nDV;
    DBuf<double> D;
    ASTRO_TRY(D.alloc((size_t)(n * cols)));
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n * cols, 256, 8192)), dim3(256), 0, c->s, D.p, n * cols, 0.0);
    ASTRO_TRY(dv_pairs_launch(c, false, c->nodeDisp.p, nullptr, nullptr, D.p, n));
    HIP_TRY(hipMemcpyAsync(dRdD, D.p, sizeof(double) * n * cols, hipMemcpyDeviceToHost, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    for (size_t i = 0; i < c->dvs.loadDof.size(); ++i)
        dRdD[(int64_t)c->dvs.loadDV[i] * n + c->dvs.loadDof[i]] -= c->dvs.loadCoef[i];
    return ASTRO_OK;
}

int astro_ctx_dRtdD_vars(astro_ctx* c, double* dRtdD) {
    ASTRO_TRY(dv_require(c));
    if (!dRtdD) return fail(ASTRO_ERR_ARG, "null argument");
    if (!c->thermalReady) return fail(ASTRO_ERR_ARG, "the thermal system has not been assembled");
    const int64_t n = c->prep.numNodes, cols = c->dvs.nDV;
    DBuf<double> D;
    ASTRO_TRY(D.alloc((size_t)(n * cols)));
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n * cols, 256, 8192)), dim3(256), 0, c->s, D.p, n * cols, 0.0);
    ASTRO_TRY(dv_pairs_launch(c, true, c->nodeTemp.p, nullptr, nullptr, D.p, n));
    HIP_TRY(hipMemcpyAsync(dRtdD, D.p, sizeof(double) * n * cols, hipMemcpyDeviceToHost, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    return ASTRO_OK;
}

// One objective term evaluated on the device: its value and, with grad, its
// contributions to dL/du (or dL/dt) on the device and to the explicit dL/dd
// (host accumulation in pair order, so the result is deterministic).
static int obj_term(astro_ctx* c, const astro_objective_term& t, double& value, bool grad, std::vector<double>* dLdD,
                    bool& usesU, bool& usesT) {
    const int E = c->prep.numEls, N = c->prep.numNodes;
    const int ni = t.numItems;
    if (ni < 1 || !t.items) return fail(ASTRO_ERR_ARG, "an objective term needs a non-empty set");
    if (t.numTgt != 0 && t.numTgt != 1 && t.numTgt != ni)
        return fail(ASTRO_ERR_ARG, "objective target count must be 0, 1 or the set size");
    if (t.numTgt > 0 && !t.tgt) return fail(ASTRO_ERR_ARG, "objective targets missing");
    const int tgtMode = t.numTgt == 0 ? 0 : (t.numTgt == 1 ? 1 : 2);
    auto tgtAt = [&](int i) { return tgtMode == 0 ? 0.0 : (tgtMode == 1 ? t.tgt[0] : t.tgt[i]); };
    hipStream_t s = c->s;
    value = 0.0;
    if (t.category == ASTRO_OBJ_DISPLACEMENT || t.category == ASTRO_OBJ_TEMPERATURE) {
        if (t.op != ASTRO_OP_POWER_NORM) return fail(ASTRO_ERR_ARG, "node terms take the powerNorm operator");
        const bool temp = t.category == ASTRO_OBJ_TEMPERATURE;
        if (temp && !c->thermalReady) return fail(ASTRO_ERR_ARG, "temperature term without a thermal system");
        if (!temp && (t.component < 1 || t.component > 6)) return fail(ASTRO_ERR_ARG, "displacement component 1..6");
        std::vector<int> dofs(ni);
        std::vector<char> seen(N, 0);
        for (int i = 0; i < ni; ++i) {
            const int nd = t.items[i] - 1;
            if (nd < 0 || nd >= N) return fail(ASTRO_ERR_ARG, "objective node out of range");
            if (seen[nd]) return fail(ASTRO_ERR_ARG, "an objective node set lists a node twice");
            seen[nd] = 1;
            if (temp) {
                dofs[i] = c->prep.currentRank[nd] - 1;
            } else {
                const int d1 = c->prep.nDofIndex[6 * (size_t)nd + t.component - 1];
                if (d1 <= 0) return fail(ASTRO_ERR_ARG, "objective node has no such displacement component");
                dofs[i] = d1 - 1;
            }
        }
        DBuf<int> dd;
        DBuf<double> tg, val;
        ASTRO_TRY(dd.upload(dofs, s));
        if (tgtMode) ASTRO_TRY(tg.upload(t.tgt, (size_t)t.numTgt, s));
        ASTRO_TRY(val.alloc(ni));
        hipLaunchKernelGGL(k_obj_nodes, dim3(grid_for(ni, 256, INT_MAX)), dim3(256), 0, s, ni, dd.p, tg.p, tgtMode,
                           t.coef, t.expo, temp ? c->nodeTemp.p : c->nodeDisp.p, val.p,
                           grad ? (temp ? c->dLdt.p : c->dLdu.p) : nullptr);
        HIP_TRY(hipGetLastError());
        std::vector<double> hv(ni);
        HIP_TRY(hipMemcpyAsync(hv.data(), val.p, sizeof(double) * ni, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        double sum = 0.0;
        for (double x : hv) sum += x;
        value = t.coef * sum;
        (temp ? usesT : usesU) = true;
        return ASTRO_OK;
    }
    const int q = t.category;
    if (q != ASTRO_OBJ_STRESS && q != ASTRO_OBJ_STRAIN && q != ASTRO_OBJ_STRAIN_ENERGY && q != ASTRO_OBJ_MASS &&
        q != ASTRO_OBJ_VOLUME)
        return fail(ASTRO_ERR_ARG, "unknown objective category " + std::to_string(q));
    if ((q == ASTRO_OBJ_STRESS || q == ASTRO_OBJ_STRAIN) && (t.component < 1 || t.component > 6))
        return fail(ASTRO_ERR_ARG, "stress / strain component 1..6");
    const bool pn = t.op == ASTRO_OP_POWER_NORM, vi = t.op == ASTRO_OP_VOLUME_INTEGRAL,
               va = t.op == ASTRO_OP_VOLUME_AVERAGE;
    const bool opOk = (q == ASTRO_OBJ_MASS || q == ASTRO_OBJ_VOLUME) ? vi
                      : (q == ASTRO_OBJ_STRAIN_ENERGY)              ? (pn || vi || va)
                                                                    : (pn || va);
    if (!opOk) return fail(ASTRO_ERR_ARG, "operator not defined for this objective category");
    std::vector<int> items(ni);
    std::vector<char> seen(E, 0);
    for (int i = 0; i < ni; ++i) {
        const int e = t.items[i] - 1;
        if (e < 0 || e >= E) return fail(ASTRO_ERR_ARG, "objective element out of range");
        if (seen[e]) return fail(ASTRO_ERR_ARG, "an objective element set lists an element twice");
        seen[e] = 1;
        items[i] = e;
    }
    const bool needU = q != ASTRO_OBJ_MASS && q != ASTRO_OBJ_VOLUME;
    DBuf<int> di;
    DBuf<double> dq, dv, dqdu;
    ASTRO_TRY(di.upload(items, s));
    ASTRO_TRY(dq.alloc(ni));
    ASTRO_TRY(dv.alloc(ni));
    if (grad && needU) ASTRO_TRY(dqdu.alloc(24 * (size_t)ni));
    hipLaunchKernelGGL(k_obj_elements, dim3(grid_for(ni, 128, INT_MAX)), dim3(128), 0, s, q, t.component, ni, di.p,
                       c->conn.p, c->X.p, c->elMat.p, c->elScale.p, c->matC.p, c->elRho.p, c->rank.p,
                       c->prep.numSolidNds, c->nodeDisp.p, dq.p, dv.p, grad && needU ? dqdu.p : nullptr);
    HIP_TRY(hipGetLastError());
    std::vector<double> hq(ni), hv(ni);
    HIP_TRY(hipMemcpyAsync(hq.data(), dq.p, sizeof(double) * ni, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(hv.data(), dv.p, sizeof(double) * ni, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const bool perVol = pn && q == ASTRO_OBJ_STRAIN_ENERGY;  // strain energy per unit volume
    double S = 0.0, Vt = 0.0;
    std::vector<double> w(ni, 0.0);
    if (pn) {
        for (int i = 0; i < ni; ++i) {
            const double x = perVol ? hq[i] / hv[i] : hq[i];
            S += std::pow(x - tgtAt(i), t.expo);
            w[i] = t.coef * t.expo * std::pow(x - tgtAt(i), t.expo - 1.0) / (perVol ? hv[i] : 1.0);
        }
        value = t.coef * S;
    } else if (vi) {
        for (int i = 0; i < ni; ++i) S += hq[i];
        value = t.coef * std::pow(S - tgtAt(0), t.expo);
        const double f = t.coef * t.expo * std::pow(S - tgtAt(0), t.expo - 1.0);
        for (int i = 0; i < ni; ++i) w[i] = f;
    } else {
        for (int i = 0; i < ni; ++i) {
            S += hq[i] * hv[i];
            Vt += hv[i];
        }
        value = t.coef * std::pow(S / Vt - tgtAt(0), t.expo);
        const double f = t.coef * t.expo * std::pow(S / Vt - tgtAt(0), t.expo - 1.0);
        for (int i = 0; i < ni; ++i) w[i] = f * hv[i] / Vt;
    }
    if (!grad) return ASTRO_OK;
    if (needU) {
        usesU = true;
        // dL/du += w_i dq_i/du, one launch per element colour
        std::vector<int> slots;
        std::vector<int> start(c->colorStart.size(), 0);
        for (int i = 0; i < ni; ++i) start[c->elColor[items[i]] + 1]++;
        for (size_t k = 0; k + 1 < start.size(); ++k) start[k + 1] += start[k];
        slots.assign(ni, 0);
        {
            std::vector<int> fill(start.begin(), start.end() - 1);
            for (int i = 0; i < ni; ++i) slots[fill[c->elColor[items[i]]]++] = i;
        }
        DBuf<int> dslot;
        DBuf<double> dw;
        ASTRO_TRY(dslot.upload(slots, s));
        ASTRO_TRY(dw.upload(w, s));
        for (size_t k = 0; k + 1 < start.size(); ++k) {
            const int cnt = start[k + 1] - start[k];
            if (cnt <= 0) continue;
            hipLaunchKernelGGL(k_obj_scatter_du, dim3(grid_for((int64_t)cnt * 24, 256, INT_MAX)), dim3(256), 0, s,
                               cnt, dslot.p + start[k], di.p, dw.p, dqdu.p, c->conn.p, c->rank.p, c->prep.numSolidNds,
                               c->dLdu.p);
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(s));  // the launches read dslot and dw, freed at scope end
    }
    if (!c->dvs.ready || !dLdD) {
        HIP_TRY(hipStreamSynchronize(s));
        return ASTRO_OK;
    }
    // explicit dL/dd: the pairs of the set's elements, in item order
    auto& V = c->dvs;
    std::vector<int> pidx, pitem;
    for (int i = 0; i < ni; ++i)
        for (int j = V.elPairStart[items[i]]; j < V.elPairStart[items[i] + 1]; ++j) {
            pidx.push_back(V.elPairList[j]);
            pitem.push_back(i);
        }
    const int np = (int)pidx.size();
    if (np > 0) {
        DBuf<int> dp;
        DBuf<double> ddq, ddv;
        ASTRO_TRY(dp.upload(pidx, s));
        ASTRO_TRY(ddq.alloc(np));
        ASTRO_TRY(ddv.alloc(np));
        hipLaunchKernelGGL(k_obj_pairs, dim3(grid_for(np, 128, INT_MAX)), dim3(128), 0, s, q, t.component, np, dp.p,
                           V.dPairEl.p, V.dDir.p, c->conn.p, c->X.p, V.dEprops.p, c->elRho.p, c->rank.p,
                           c->prep.numSolidNds, c->nodeDisp.p, V.h, ddq.p, ddv.p);
        HIP_TRY(hipGetLastError());
        std::vector<double> gq(np), gv(np);
        HIP_TRY(hipMemcpyAsync(gq.data(), ddq.p, sizeof(double) * np, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(gv.data(), ddv.p, sizeof(double) * np, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        std::vector<double>& g = *dLdD;
        if (pn) {
            for (int j = 0; j < np; ++j) {
                const int i = pitem[j], k = V.pairDV[pidx[j]];
                // w_i already carries 1 / v for the per-volume strain energy
                const double dx = perVol ? (gq[j] - hq[i] / hv[i] * gv[j]) : gq[j];
                g[k] += w[i] * dx;
            }
        } else if (vi) {
            for (int j = 0; j < np; ++j) g[V.pairDV[pidx[j]]] += w[pitem[j]] * gq[j];
        } else {
            const double f = t.coef * t.expo * std::pow(S / Vt - tgtAt(0), t.expo - 1.0);
            std::vector<double> dS(V.nDV, 0.0), dV(V.nDV, 0.0);
            for (int j = 0; j < np; ++j) {
                const int i = pitem[j], k = V.pairDV[pidx[j]];
                dS[k] += gq[j] * hv[i] + hq[i] * gv[j];
                dV[k] += gv[j];
            }
            for (int k = 0; k < V.nDV; ++k)
                if (dS[k] != 0.0 || dV[k] != 0.0) g[k] += f * (dS[k] * Vt - S * dV[k]) / (Vt * Vt);
        }
    }
    return ASTRO_OK;
}

// Objective terms whose element quantity holds the stress (stress, strain
// energy) of a model with thermal expansion: C (strain - T alpha) is not in
// the objective kernels.
static int texp_terms_check(const astro_ctx* c, int32_t numTerms, const astro_objective_term* terms) {
    if (!c->hasTExp) return ASTRO_OK;
    for (int t = 0; t < numTerms; ++t)
        if (terms[t].category == ASTRO_OBJ_STRESS || terms[t].category == ASTRO_OBJ_STRAIN_ENERGY)
            return fail(ASTRO_ERR_UNSUPPORTED, "stress and strain-energy terms of a model with thermal expansion");
    return ASTRO_OK;
}

int astro_ctx_objective(astro_ctx* c, int32_t numTerms, const astro_objective_term* terms, double* objVal) {
    if (!c || !terms || !objVal || numTerms < 1) return fail(ASTRO_ERR_ARG, "null or empty argument");
    if (c->prep.shellMesh) return fail(ASTRO_ERR_UNSUPPORTED, "objective terms cover brick8 models");
    ASTRO_TRY(texp_terms_check(c, numTerms, terms));
    bool uu = false, ut = false;
    for (int t = 0; t < numTerms; ++t) ASTRO_TRY(obj_term(c, terms[t], objVal[t], false, nullptr, uu, ut));
    return ASTRO_OK;
}

int astro_ctx_objective_gradient(astro_ctx* c, int32_t numTerms, const astro_objective_term* terms, double* dLdD) {
    ASTRO_TRY(dv_require(c));
    if (!terms || !dLdD || numTerms < 1) return fail(ASTRO_ERR_ARG, "null or empty argument");
    if (c->hasTExp)
        return fail(ASTRO_ERR_UNSUPPORTED, "design gradients of a model with thermal expansion");
    const int n = c->prep.elMatDim, N = c->prep.numNodes, nDV = c->dvs.nDV;
    hipStream_t s = c->s;
    if (c->dLdu.n != (size_t)n) ASTRO_TRY(c->dLdu.alloc(n));
    if (c->adjU.n != (size_t)n) ASTRO_TRY(c->adjU.alloc(n));
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, c->dLdu.p, (int64_t)n, 0.0);
    if (c->thermalReady) {
        if (c->dLdt.n != (size_t)N) ASTRO_TRY(c->dLdt.alloc(N));
        if (c->adjT.n != (size_t)N) ASTRO_TRY(c->adjT.alloc(N));
        hipLaunchKernelGGL(k_fill, dim3(grid_for(N, 256, 4096)), dim3(256), 0, s, c->dLdt.p, (int64_t)N, 0.0);
    }
    std::vector<double> g(nDV, 0.0);
    bool usesU = false, usesT = false;
    for (int t = 0; t < numTerms; ++t) {
        double v = 0.0;
        ASTRO_TRY(obj_term(c, terms[t], v, true, &g, usesU, usesT));
    }
    auto& V = c->dvs;
    const int nP = (int)V.pairEl.size();
    DBuf<double> lamDot;
    ASTRO_TRY(lamDot.alloc(std::max(nP, 1)));
    std::vector<double> hl(nP);
    if (usesU) {
        // solveDynamicAdjoint, static branch: (K + C^T C) lambda = dL/du on the
        // factored tangent of the last static solve
        bool ok = false;
        ASTRO_TRY(with_prof(c, [&](auto& P) {
            ok = P.factored;
            return ASTRO_OK;
        }));
        if (!ok || c->csrMassFactor != 0.0)
            return fail(ASTRO_ERR_ARG, "the elastic adjoint needs the factored static system (static_solve first)");
        hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, c->adjU.p, (int64_t)n, 0.0);
        const int rc = ctx_gmres(c, c->dLdu.p, c->adjU.p);
        if (rc != ASTRO_OK) return rc;
        ASTRO_TRY(dv_pairs_launch(c, false, c->nodeDisp.p, c->adjU.p, lamDot.p, nullptr, n));
        HIP_TRY(hipMemcpyAsync(hl.data(), lamDot.p, sizeof(double) * nP, hipMemcpyDeviceToHost, s));
        std::vector<double> lam(n);
        HIP_TRY(hipMemcpyAsync(lam.data(), c->adjU.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int p = 0; p < nP; ++p) g[V.pairDV[p]] -= hl[p];
        for (size_t i = 0; i < V.loadDof.size(); ++i) g[V.loadDV[i]] += V.loadCoef[i] * lam[V.loadDof[i]];
    }
    if (usesT) {
        bool ok = false;
        ASTRO_TRY(with_tprof(c, [&](auto& P) {
            ok = P.factored;
            return ASTRO_OK;
        }));
        if (!ok || !c->thermalSteady)
            return fail(ASTRO_ERR_ARG, "the thermal adjoint needs the factored steady system (thermal_solve first)");
        hipLaunchKernelGGL(k_fill, dim3(grid_for(N, 256, 4096)), dim3(256), 0, s, c->adjT.p, (int64_t)N, 0.0);
        const int rc = with_tprof(c, [&](auto& P) {
            return c->tgm.solve(c->top, P, c->dLdt.p, c->adjT.p, c->opt.rtol, 0.0, c->opt.maxCycles, s);
        });
        if (rc != ASTRO_OK) return rc;
        ASTRO_TRY(dv_pairs_launch(c, true, c->nodeTemp.p, c->adjT.p, lamDot.p, nullptr, N));
        HIP_TRY(hipMemcpyAsync(hl.data(), lamDot.p, sizeof(double) * nP, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int p = 0; p < nP; ++p) g[V.pairDV[p]] -= hl[p];
    }
    for (int k = 0; k < nDV; ++k) dLdD[k] = g[k];
    return ASTRO_OK;
}


// Gradient of a Newmark objective, the dynamic branch of getTotaldLdD
// (Working/AStrO_commandFunctions.f90:1416-1464) with augmentdLdu (:819-915)
// and solveDynamicAdjoint (:917-956): J = sum over steps i = 1..nSteps of
// the terms active at t_i = i dt, on the states U_i, A_i the forward left
// (host, (nSteps + 1) x elMatDim each, row 0 the start).  Backward from the
// last step: dL/du_i = dObj/du(U_i) plus, from step i + 1, c1 c2 M lam +
// accAdj; dL/dv_i and dL/da_i likewise (M lam by the residual kernel with a
// zero displacement and lam in the previous-state slot, as augmentdLdu does);
// velAdj = dL/dv, accAdj = -(dL/da + dt gamma velAdj) / (dt^2 beta),
// (K + M / (beta dt^2) + C^T C) lam_i = dL/du - accAdj on the factored
// Newmark tangent; dL/dd += dObj/dd(U_i) - lam_i . dR_i/dd with the inertia
// at A_i and the nodal loads scaled by loadScale[i - 1].  termWindow: 2 per
// term, the activeTime window (null: always active).
int astro_ctx_dynamic_objective_gradient(astro_ctx* c, int32_t nSteps, const double* U, const double* A,
                                         const double* loadScale, const double* termWindow, int32_t numTerms,
                                         const astro_objective_term* terms, double* objSum, double* dLdD) {
    ASTRO_TRY(dv_require(c));
    if (!U || !A || !terms || !dLdD || numTerms < 1 || nSteps < 1) return fail(ASTRO_ERR_ARG, "null or empty argument");
    if (!c->dynReady) return fail(ASTRO_ERR_ARG, "astro_ctx_dynamic_init has not been called");
    const int n = c->prep.elMatDim, nS = c->prep.numSolidNds, nDV = c->dvs.nDV;
    const double beta = c->nmBeta, gamma = c->nmGamma, dt = c->delT, c1 = 0.5 / beta, c2 = 2.0 / (dt * dt);
    hipStream_t s = c->s;
    if (c->dLdu.n != (size_t)n) ASTRO_TRY(c->dLdu.alloc(n));
    DBuf<double> lam, lamN, accAdj, accAdjN, velAdj, velAdjN, dLdv, dLda, zero, accBuf, lamDot;
    for (DBuf<double>* b : {&lam, &lamN, &accAdj, &accAdjN, &velAdj, &velAdjN, &dLdv, &dLda, &zero, &accBuf}) {
        ASTRO_TRY(b->alloc(n));
        hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, b->p, (int64_t)n, 0.0);
    }
    // the current and the next step's adjoint vectors (swapped per step)
    double *pLam = lam.p, *pLamN = lamN.p, *pAcc = accAdj.p, *pAccN = accAdjN.p, *pVel = velAdj.p,
           *pVelN = velAdjN.p;
    auto& V = c->dvs;
    const int nP = (int)V.pairEl.size();
    ASTRO_TRY(lamDot.alloc(std::max(nP, 1)));
    std::vector<double> g(nDV, 0.0), hl(nP), hlam(n);
    double total = 0.0;
    // load -= R_e(u = 0, previous state (pu, pv, pa)) over the element colours
    auto inertia = [&](const double* pu, const double* pv, const double* pa, double* load) {
        for (size_t k = 0; k + 1 < c->colorStart.size(); ++k) {
            const int cnt = c->colorStart[k + 1] - c->colorStart[k];
            if (cnt <= 0) continue;
            hipLaunchKernelGGL(k_brick8_residual, dim3(cnt), dim3(64), 0, s, c->perm.p + c->colorStart[k], cnt,
                               c->conn.p, c->X.p, c->elMat.p, c->elScale.p, c->matC.p, c->elRho.p, c->rank.p, nS,
                               (const double*)zero.p, pu, pv, pa, c1, c2, dt, beta, (const double*)nullptr,
                               (const double*)nullptr, load, (const int*)nullptr);
        }
    };
    auto axpby = [&](double* y, const double* x, double a, double b) {  // y = a y + b x
        hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, y, x, y, n, a, b,
                           (const int*)nullptr);
    };
    for (int i = nSteps; i >= 1; --i) {
        const double t = dt * i;
        HIP_TRY(hipMemcpyAsync(c->nodeDisp.p, U + (size_t)i * n, sizeof(double) * n, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(accBuf.p, A + (size_t)i * n, sizeof(double) * n, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, c->dLdu.p, (int64_t)n, 0.0);
        hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, dLdv.p, (int64_t)n, 0.0);
        hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, dLda.p, (int64_t)n, 0.0);
        bool usesU = false, usesT = false;
        for (int q = 0; q < numTerms; ++q) {
            if (termWindow && !(termWindow[2 * q] <= t && t <= termWindow[2 * q + 1])) continue;
            double v = 0.0;
            ASTRO_TRY(obj_term(c, terms[q], v, true, &g, usesU, usesT));
            total += v;
        }
        if (usesT) return fail(ASTRO_ERR_UNSUPPORTED, "temperature terms of a Newmark objective");
        if (i < nSteps) {
            // augmentdLdu: the next step's residual through this step's state
            inertia(pLamN, zero.p, zero.p, c->dLdu.p);  // + c1 c2 M lam
            axpby(c->dLdu.p, pAccN, 1.0, 1.0);
            inertia(zero.p, pLamN, zero.p, dLdv.p);  // + c1 c2 dt M lam
            axpby(dLdv.p, pAccN, 1.0, dt);
            axpby(dLdv.p, pVelN, 1.0, 1.0);
            inertia(zero.p, zero.p, pLamN, dLda.p);  // + c1 (1 - 2 beta) M lam
            axpby(dLda.p, pAccN, 1.0, 0.5 * dt * dt * (1.0 - 2.0 * beta));
            axpby(dLda.p, pVelN, 1.0, dt * (1.0 - gamma));
        }
        // solveDynamicAdjoint
        HIP_TRY(hipMemcpyAsync(pVel, dLdv.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(pAcc, dLda.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        axpby(pAcc, pVel, -1.0 / (dt * dt * beta), -gamma / (dt * beta));
        axpby(c->dLdu.p, pAcc, 1.0, -1.0);
        hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, pLam, (int64_t)n, 0.0);
        const int rc = ctx_gmres(c, c->dLdu.p, pLam);
        if (rc != ASTRO_OK) return rc;
        // - lam . dR_i/dd: stiffness and inertia per pair, the scaled node loads
        ASTRO_TRY(dv_pairs_launch(c, false, c->nodeDisp.p, pLam, lamDot.p, nullptr, n, accBuf.p));
        HIP_TRY(hipMemcpyAsync(hl.data(), lamDot.p, sizeof(double) * nP, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(hlam.data(), pLam, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int p = 0; p < nP; ++p) g[V.pairDV[p]] -= hl[p];
        const double sc = loadScale ? loadScale[i - 1] : 1.0;
        for (size_t k = 0; k < V.loadDof.size(); ++k) g[V.loadDV[k]] += sc * V.loadCoef[k] * hlam[V.loadDof[k]];
        std::swap(pLam, pLamN);
        std::swap(pAcc, pAccN);
        std::swap(pVel, pVelN);
    }
    // the context keeps the forward's last state
    HIP_TRY(hipMemcpyAsync(c->nodeDisp.p, U + (size_t)nSteps * n, sizeof(double) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int k = 0; k < nDV; ++k) dLdD[k] = g[k];
    if (objSum) *objSum = total;
    return ASTRO_OK;
}

// ---------------------------------------------------------------------------
// Batched complex-step sensitivities: every design variable's perturbed
// system as a complex128 instance (the kernel and its algorithm are in
// astro_batch.hip).  Set-up, once per design: the node-major DOF order
// (nodes by level rank, the DOFs of a node together) and its profile over
// the CSR and penalty patterns, the loads in that order, the element
// stage's DOF lists, and per variable its elements and first row r_k.
static int csb_element_stage(astro_ctx* c) {
    const int64_t n = c->prep.elMatDim;
    for (size_t k = 0; k + 1 < c->colorStart.size(); ++k) {
        const int cnt = c->colorStart[k + 1] - c->colorStart[k];
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(k_shell4_dtangent_cs, dim3(cnt), dim3(64), 0, c->s, c->perm.p + c->colorStart[k], cnt,
                           c->conn.p, c->X.p, c->elABD.p, c->elABDim.p, c->elAlpha.p, c->rank.p, c->prep.numSolidNds,
                           c->prep.numSectionNds, c->nodeDisp.p, c->elDV.p, 1.0 / c->csH, n, c->dRdD.p, c->gPart.p,
                           c->csbKim.p, c->csbKimDof.p);
    }
    HIP_TRY(hipGetLastError());
    return ASTRO_OK;
}

static int csb_setup(astro_ctx* c) {
    auto& B = c->csb;
    if (B.ready) return ASTRO_OK;
    const Prep& p = c->prep;
    const int n = p.elMatDim, N = p.numNodes, E = p.numEls, nDV = c->nDV;
    std::vector<int> nodes(N);
    for (int i = 0; i < N; ++i) nodes[i] = i;
    std::stable_sort(nodes.begin(), nodes.end(),
                     [&](int a, int b) { return p.currentRank[a] < p.currentRank[b]; });
    std::vector<int> order;
    order.reserve(n);
    for (int nd : nodes)
        for (int q = 0; q < 6; ++q) {
            const int d = p.nDofIndex[6 * (size_t)nd + q];
            if (d > 0) order.push_back(d - 1);
        }
    if ((int)order.size() != n) return fail(ASTRO_ERR_ARG, "DOF map does not cover the system");
    B.inv.assign(n, -1);
    for (int k = 0; k < n; ++k) {
        if (order[k] < 0 || order[k] >= n || B.inv[order[k]] >= 0) return fail(ASTRO_ERR_ARG, "DOF map is not a permutation");
        B.inv[order[k]] = k;
    }
    B.first.resize(n);
    for (int i = 0; i < n; ++i) B.first[i] = i;
    auto touch = [&](int a, int b) {
        const int I = std::max(a, b), J = std::min(a, b);
        B.first[I] = std::min(B.first[I], J);
    };
    for (int r = 0; r < n; ++r)
        for (int q = p.elMatRange[r]; q < p.elMatRange[r + 1]; ++q) {
            const int col = p.elMatCols[q];
            if (col >= 1 && col <= n) touch(B.inv[r], B.inv[col - 1]);
        }
    for (int r = 0; r < p.elMPCDim; ++r)
        for (int a = p.elMPCMatRange[r]; a < p.elMPCMatRange[r + 1]; ++a)
            for (int b = p.elMPCMatRange[r]; b < p.elMPCMatRange[r + 1]; ++b) {
                const int ca = p.elMPCMatCols[a], cb = p.elMPCMatCols[b];
                if (ca >= 1 && ca <= n && cb >= 1 && cb <= n) touch(B.inv[ca - 1], B.inv[cb - 1]);
            }
    B.rowoff.assign(n + 1, 0);
    B.maxRow = 0;
    for (int i = 0; i < n; ++i) {
        const int len = i - B.first[i] + 1;
        B.maxRow = std::max(B.maxRow, len);
        B.rowoff[i + 1] = B.rowoff[i] + len;
    }
    B.total = B.rowoff[n];
    const size_t ldsMax = 160 * 1024;
    if (csb_lds_bytes(0, B.maxRow) > ldsMax)
        return fail(ASTRO_ERR_UNSUPPORTED, "node-major profile rows of " + std::to_string(B.maxRow) +
                                               " entries exceed the batched kernel's LDS tiles");
    // loads in the node-major order
    std::vector<int> ld(c->loadDofs.n);
    std::vector<double> lv(c->loadVals.n), F(n, 0.0);
    if (!ld.empty()) {
        HIP_TRY(hipMemcpyAsync(ld.data(), c->loadDofs.p, sizeof(int) * ld.size(), hipMemcpyDeviceToHost, c->s));
        HIP_TRY(hipMemcpyAsync(lv.data(), c->loadVals.p, sizeof(double) * lv.size(), hipMemcpyDeviceToHost, c->s));
        HIP_TRY(hipStreamSynchronize(c->s));
    }
    for (size_t k = 0; k < ld.size(); ++k) F[B.inv[ld[k] - 1]] += lv[k];
    // element DOF lists from the element stage
    ASTRO_TRY(c->csbKim.alloc(576 * (size_t)E));
    ASTRO_TRY(c->csbKimDof.alloc(24 * (size_t)E));
    std::vector<int> kd(24 * (size_t)E, 0);
    HIP_TRY(hipMemcpyAsync(c->csbKimDof.p, kd.data(), sizeof(int) * kd.size(), hipMemcpyHostToDevice, c->s));
    ASTRO_TRY(csb_element_stage(c));
    HIP_TRY(hipMemcpyAsync(kd.data(), c->csbKimDof.p, sizeof(int) * kd.size(), hipMemcpyDeviceToHost, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    std::vector<int> elStart(nDV + 1, 0), elList, instStart(nDV, n);
    for (int e = 0; e < E; ++e)
        if (c->hElDV[e] >= 0) ++elStart[c->hElDV[e] + 1];
    for (int k = 0; k < nDV; ++k) elStart[k + 1] += elStart[k];
    elList.resize(elStart[nDV]);
    {
        std::vector<int> fill(elStart.begin(), elStart.end() - 1);
        for (int e = 0; e < E; ++e) {
            const int k = c->hElDV[e];
            if (k < 0) continue;
            elList[fill[k]++] = e;
            for (int a = 0; a < 24; ++a) {
                const int d = kd[24 * (size_t)e + a];
                if (d < 1 || d > n) return fail(ASTRO_ERR_ARG, "element DOF out of range in the batched set-up");
                instStart[k] = std::min(instStart[k], B.inv[d - 1]);
            }
        }
    }
    std::vector<int> instOrder(nDV);
    for (int k = 0; k < nDV; ++k) instOrder[k] = k;
    std::stable_sort(instOrder.begin(), instOrder.end(), [&](int a, int b) { return instStart[a] < instStart[b]; });
    B.instStart = instStart;
    B.nInst = nDV;
    B.flops = 0.0;
    for (int k = 0; k < nDV; ++k)
        for (int i = 0; i < n; ++i) {
            const double len = (double)(i - B.first[i]);
            B.flops += (i >= instStart[k] ? 4.0 : 1.0) * len * len;
        }
    ASTRO_TRY(B.dInv.upload(B.inv, c->s));
    ASTRO_TRY(B.dDofOrder.upload(order, c->s));
    ASTRO_TRY(B.dFirst.upload(B.first, c->s));
    ASTRO_TRY(B.dRowoff.upload(B.rowoff, c->s));
    ASTRO_TRY(B.dOrder.upload(instOrder, c->s));
    ASTRO_TRY(B.dInstStart.upload(instStart, c->s));
    ASTRO_TRY(B.dElStart.upload(elStart, c->s));
    ASTRO_TRY(B.dElList.upload(elList, c->s));
    ASTRO_TRY(B.F.upload(F, c->s));
    ASTRO_TRY(B.profRe.alloc((size_t)B.total));
    ASTRO_TRY(B.work.alloc(2 * (size_t)B.total * nDV));
    ASTRO_TRY(B.vec.alloc(2 * (size_t)n * nDV));
    ASTRO_TRY(B.sol.alloc(2 * (size_t)n * nDV));
    ASTRO_TRY(B.ed.alloc(4 * (size_t)n * nDV));
    ASTRO_TRY(B.grad.alloc(nDV));
    ASTRO_TRY(B.comp.alloc(nDV));
    ASTRO_TRY(B.status.alloc(1));
    HIP_TRY(hipStreamSynchronize(c->s));
    B.ready = true;
    return ASTRO_OK;
}

// One batched evaluation with tile t: the real assembly and penalty scaling
// of the current design, the element stage, A_re in the node-major profile,
// then the instances' launch and kCsbRefine refinement steps (residual
// launches and the kernel in solve-only mode), timed with events on the
// context stream.  The direct solve leaves an error of about cond eps (1e-9
// on the 128 x 16 plate); a double-double residual step takes it to the
// rounding of the system's own entries.
static constexpr int kCsbRefine = 1;
static int csb_run(astro_ctx* c, int t) {
    auto& B = c->csb;
    const Prep& p = c->prep;
    if (csb_lds_bytes(t, B.maxRow) > 160 * 1024) return fail(ASTRO_ERR_ARG, "tile does not fit in LDS");
    ASTRO_TRY(astro_ctx_assemble_elastic(c, 0.0));
    ASTRO_TRY(astro_ctx_scale_mpc(c));
    ASTRO_TRY(csb_element_stage(c));
    HIP_TRY(csb_real_profile(c->csrRange.p, c->csrCols.p, c->csrVal.p, p.elMatDim, c->mpcMat.p, c->mpcCols.p,
                             c->mpcRange.p, p.elMPCDim, B.dInv.p, B.dFirst.p, B.dRowoff.p, B.profRe.p, B.total, c->s));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->s, B.status.p, (int64_t)1, 0);
    CsbArgs a;
    a.n = p.elMatDim;
    a.W = csb_stride(t, B.maxRow);
    a.maxRow = B.maxRow;
    a.total = B.total;
    a.first = B.dFirst.p;
    a.rowoff = B.dRowoff.p;
    a.inv = B.dInv.p;
    a.profRe = B.profRe.p;
    a.F = B.F.p;
    a.order = B.dOrder.p;
    a.instStart = B.dInstStart.p;
    a.elStart = B.dElStart.p;
    a.elList = B.dElList.p;
    a.kim = c->csbKim.p;
    a.kimDof = c->csbKimDof.p;
    a.work = B.work.p;
    a.vec = B.vec.p;
    a.sol = B.sol.p;
    a.invH = 1.0 / c->csH;
    a.grad = B.grad.p;
    a.compliance = B.comp.p;
    a.status = B.status.p;
    CsbRes R;
    R.n = p.elMatDim;
    R.order = B.dDofOrder.p;
    R.inv = B.dInv.p;
    R.range = c->csrRange.p;
    R.cols = c->csrCols.p;
    R.val = c->csrVal.p;
    R.cmat = c->mpcMat.p;
    R.ccols = c->mpcCols.p;
    R.crange = c->mpcRange.p;
    R.ctRange = p.elMPCDim > 0 ? c->mpcT.range.p : nullptr;
    R.ctPos = c->mpcT.pos.p;
    R.ctRow = c->mpcT.row.p;
    R.F = B.F.p;
    R.elStart = B.dElStart.p;
    R.elList = B.dElList.p;
    R.kim = c->csbKim.p;
    R.kimDof = c->csbKimDof.p;
    R.sol = B.sol.p;
    R.ed = B.ed.p;
    R.vec = B.vec.p;
    HIP_TRY(hipEventRecord(c->ev[6], c->s));
    a.mode = 0;
    HIP_TRY(csb_launch(a, t, B.nInst, c->s));
    for (int it = 0; it < kCsbRefine; ++it) {
        HIP_TRY(csb_residual(R, B.nInst, c->s));
        a.mode = 1;
        HIP_TRY(csb_launch(a, t, B.nInst, c->s));
    }
    HIP_TRY(hipEventRecord(c->ev[7], c->s));
    HIP_TRY(hipEventSynchronize(c->ev[7]));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev[6], c->ev[7]));
    B.lastMs = ms;
    int st = 0;
    HIP_TRY(hipMemcpy(&st, B.status.p, sizeof(int), hipMemcpyDeviceToHost));
    if (st) return fail(ASTRO_ERR_ZERO_PIVOT, "zero pivot in a batched complex factor");
    return ASTRO_OK;
}

// Compliance gradient by direct complex-step solves, all design variables in
// one launch: grad[k] = Im(F . u(d + i h e_k)) / h; compliance (may be null)
// receives Re(F . u_k), each instance's compliance.
int astro_ctx_cs_batch_gradient(astro_ctx* c, double* grad, double* compliance) {
    if (!c || !grad) return fail(ASTRO_ERR_ARG, "null argument");
    if (c->nDV < 1) return fail(ASTRO_ERR_ARG, "astro_ctx_set_design has not been called");
    ASTRO_TRY(csb_setup(c));
    ASTRO_TRY(csb_run(c, c->csb.tile));
    HIP_TRY(hipMemcpy(grad, c->csb.grad.p, sizeof(double) * c->nDV, hipMemcpyDeviceToHost));
    if (compliance) HIP_TRY(hipMemcpy(compliance, c->csb.comp.p, sizeof(double) * c->nDV, hipMemcpyDeviceToHost));
    return ASTRO_OK;
}

// The autotuner: times every tile shape that fits in LDS on the context's
// own batch (after one untimed run of each) and keeps the fastest for later
// calls.  tiles / ms (capacity count): RB * 100 + CB and the kernel time
// of each tile (ms < 0: does not fit).
int astro_ctx_cs_batch_autotune(astro_ctx* c, int32_t* tiles, float* ms, int32_t count) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    if (c->nDV < 1) return fail(ASTRO_ERR_ARG, "astro_ctx_set_design has not been called");
    ASTRO_TRY(csb_setup(c));
    auto& B = c->csb;
    const int nt = csb_num_tiles();
    int best = -1;
    for (int t = 0; t < nt; ++t) {
        B.tileMs[t] = -1.0f;
        if (csb_lds_bytes(t, B.maxRow) > 160 * 1024) continue;
        ASTRO_TRY(csb_run(c, t));
        ASTRO_TRY(csb_run(c, t));
        B.tileMs[t] = B.lastMs;
        if (best < 0 || B.tileMs[t] < B.tileMs[best]) best = t;
    }
    if (best < 0) return fail(ASTRO_ERR_UNSUPPORTED, "no tile shape fits");
    B.tile = best;
    for (int t = 0; t < nt && t < count; ++t) {
        int rb = 0, cb = 0;
        csb_tile(t, &rb, &cb);
        if (tiles) tiles[t] = rb * 100 + cb;
        if (ms) ms[t] = B.tileMs[t];
    }
    return ASTRO_OK;
}

int astro_ctx_cs_batch_tile(astro_ctx* c, int32_t tile) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    for (int t = 0; t < csb_num_tiles(); ++t) {
        int rb = 0, cb = 0;
        csb_tile(t, &rb, &cb);
        if (rb * 100 + cb == tile) {
            c->csb.tile = t;
            return ASTRO_OK;
        }
    }
    return fail(ASTRO_ERR_ARG, "unknown tile " + std::to_string(tile));
}

double astro_ctx_phase_ms(const astro_ctx* c, const char* phase) {
    if (!c || !phase) return NAN;
    static const char* names[] = {"assemble", "scale", "profile", "factor", "solve"};
    for (int k = 0; k < 5; ++k)
        if (!std::strcmp(phase, names[k])) return c->phaseMs[k];
    return NAN;
}

// ---- thermal ----------------------------------------------------------------
static int ctx_thermal_setup(astro_ctx* c) {
    if (c->thermalReady) return ASTRO_OK;
    const Prep& p = c->prep;
    hipStream_t s = c->s;
    ASTRO_TRY(c->tRange.upload(p.thermMatRange, s));
    ASTRO_TRY(c->tCols.upload(p.thermMatCols, s));
    ASTRO_TRY(c->tVal.alloc(p.thermMatCols.size()));
    ASTRO_TRY(c->tMpcRange.upload(p.thermMPCMatRange, s));
    ASTRO_TRY(c->tMpcCols.upload(p.thermMPCMatCols, s));
    ASTRO_TRY(c->tMpcMat.upload(p.thermMPCMat, s));
    ASTRO_TRY(c->tMpcRhs.upload(p.thermMPCRHS, s));
    const int N = p.numNodes;
    ASTRO_TRY(c->tLoad.alloc(N));
    ASTRO_TRY(c->tDel.alloc(N));
    ASTRO_TRY(c->nodeTemp.alloc(N));
    HIP_TRY(hipMemsetAsync(c->tDel.p, 0, sizeof(double) * N, s));
    HIP_TRY(hipMemsetAsync(c->nodeTemp.p, 0, sizeof(double) * N, s));
    ASTRO_TRY(with_tprof(c, [&](auto& P) {
        ASTRO_TRY(P.setup(p.thermMatLTRange.data(), N, s, true));
        return P.pen.build(p.thermMPCMatRange.data(), p.thermMPCMatCols.data(), p.thermMPCDim, P.hS, s);
    }));
    c->top.n = N;
    c->top.cdim = p.thermMPCDim;
    c->top.val = c->tVal.p;
    c->top.cols = c->tCols.p;
    c->top.range = c->tRange.p;
    c->top.cmat = c->tMpcMat.p;
    c->top.ccols = c->tMpcCols.p;
    c->top.crange = c->tMpcRange.p;
    ASTRO_TRY(c->tMpcT.upload(p.thermMPCMatRange.data(), p.thermMPCMatCols.data(), p.thermMPCDim, N, s));
    c->top.set_transpose(c->tMpcT);
    int nv = c->opt.numVecs > 0 ? c->opt.numVecs : (N < 200 ? N / 2 : 100);
    ASTRO_TRY(c->tgm.setup(N, nv));
    c->thermalReady = true;
    return ASTRO_OK;
}

int astro_ctx_assemble_thermal(astro_ctx* c) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    if (c->prep.shellMesh) return fail(ASTRO_ERR_UNSUPPORTED, "thermal analysis of shell4 meshes is not part of the device path");
    ASTRO_TRY(ctx_thermal_setup(c));
    const int64_t nnz = (int64_t)c->prep.thermMatCols.size();
    c->thermalSteady = true;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(nnz, 256, 8192)), dim3(256), 0, c->s, c->tVal.p, nnz, 0.0);
    for (size_t k = 0; k + 1 < c->colorStart.size(); ++k) {
        const int cnt = c->colorStart[k + 1] - c->colorStart[k];
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(k_brick8_conduct, dim3(cnt), dim3(64), 0, c->s, c->perm.p + c->colorStart[k], cnt,
                           c->conn.p, c->X.p, c->elCond.p, 1.0, (const double*)nullptr, 0.0, c->rank.p, c->tRange.p,
                           c->thPos.p, c->tVal.p);
    }
    HIP_TRY(hipGetLastError());
    return ASTRO_OK;
}

// Thermal solve: scaleThermalMPC, convertToLTri, factor, RHS (constraint load),
// GMRES, nodeTemp += delta.  Working/AStrO_commandFunctions.f90:763-769, 689-703.
int astro_ctx_thermal_solve(astro_ctx* c) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    ASTRO_TRY(ctx_thermal_setup(c));
    hipStream_t s = c->s;
    const Prep& p = c->prep;
    const int64_t nnz = (int64_t)p.thermMatCols.size();
    const int N = p.numNodes, cd = p.thermMPCDim;
    HIP_TRY(hipMemsetAsync(c->maxBits.p, 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_max_abs, dim3(grid_for(nnz, 256, 2048)), dim3(256), 0, s, c->tVal.p, nnz, c->maxBits.p);
    if (cd > 0)
        hipLaunchKernelGGL(k_scale_mpc, dim3(grid_for(cd, 256)), dim3(256), 0, s, c->tMpcMat.p, c->tMpcRhs.p,
                           c->tMpcRange.p, cd, c->maxBits.p);
    return with_tprof(c, [&](auto& P) -> int {
        ASTRO_TRY(profile_from_csr(P, s, c->tVal.p, c->tCols.p, c->tRange.p, N, c->tMpcMat.p, c->tMpcCols.p,
                                   c->tMpcRange.p, cd));
        ASTRO_TRY(P.factor(s));
        ctx_minus_internal_force(s, c->tVal, c->tCols, c->tRange, N, c->nodeTemp.p, c->tLoad.p);
        if (cd > 0)
            hipLaunchKernelGGL(k_constraint_load, dim3(grid_for(cd, 256)), dim3(256), 0, s, c->tMpcMat.p,
                               c->tMpcCols.p, c->tMpcRange.p, c->tMpcRhs.p, cd, c->nodeTemp.p, c->tLoad.p, (const int*)nullptr);
        ASTRO_TRY(c->tgm.solve(c->top, P, c->tLoad.p, c->tDel.p, c->opt.rtol, 0.0, c->opt.maxCycles, s));
        hipLaunchKernelGGL(k_axpby, dim3(grid_for(N, 256, 4096)), dim3(256), 0, s, c->nodeTemp.p, c->tDel.p,
                           c->nodeTemp.p, N, 1.0, 1.0, (const int*)nullptr);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(s));
        return ASTRO_OK;
    });
}

// Transient thermal set-up: the tangent K + C / (gamma dt) into the thermal
// CSR (the steps recompute K and C element by element), penalty scaling by the
// tangent, profile and factor; the state starts at T = Tdot = 0.
int astro_ctx_thermal_dynamic_init(astro_ctx* c, double gamma, double dt) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    if (!(gamma > 0.0) || !(dt > 0.0)) return fail(ASTRO_ERR_ARG, "need gamma > 0 and dt > 0");
    if (c->prep.shellMesh) return fail(ASTRO_ERR_UNSUPPORTED, "transient thermal covers brick8 meshes");
    ASTRO_TRY(ctx_thermal_setup(c));
    hipStream_t s = c->s;
    const Prep& p = c->prep;
    const int64_t nnz = (int64_t)p.thermMatCols.size();
    const int N = p.numNodes, cd = p.thermMPCDim;
    const double cm = 1.0 / (gamma * dt);
    for (DBuf<double>* b : {&c->prevTemp, &c->prevTdot, &c->nodeTdot, &c->tWork})
        if (b->n != (size_t)N) ASTRO_TRY(b->alloc(N));
    for (DBuf<double>* b : {&c->prevTemp, &c->prevTdot, &c->nodeTdot, &c->nodeTemp, &c->tDel})
        hipLaunchKernelGGL(k_fill, dim3(grid_for(N, 256, 4096)), dim3(256), 0, s, b->p, (int64_t)N, 0.0);
    c->thermalSteady = false;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(nnz, 256, 8192)), dim3(256), 0, s, c->tVal.p, nnz, 0.0);
    for (size_t k = 0; k + 1 < c->colorStart.size(); ++k) {
        const int cnt = c->colorStart[k + 1] - c->colorStart[k];
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(k_brick8_conduct, dim3(cnt), dim3(64), 0, s, c->perm.p + c->colorStart[k], cnt, c->conn.p,
                           c->X.p, c->elCond.p, 1.0, (const double*)c->elCap.p, cm, c->rank.p, c->tRange.p,
                           c->thPos.p, c->tVal.p);
    }
    HIP_TRY(hipMemsetAsync(c->maxBits.p, 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_max_abs, dim3(grid_for(nnz, 256, 2048)), dim3(256), 0, s, c->tVal.p, nnz, c->maxBits.p);
    if (cd > 0)
        hipLaunchKernelGGL(k_scale_mpc, dim3(grid_for(cd, 256)), dim3(256), 0, s, c->tMpcMat.p, c->tMpcRhs.p,
                           c->tMpcRange.p, cd, c->maxBits.p);
    HIP_TRY(hipGetLastError());
    ASTRO_TRY(with_tprof(c, [&](auto& P) -> int {
        ASTRO_TRY(profile_from_csr(P, s, c->tVal.p, c->tCols.p, c->tRange.p, N, c->tMpcMat.p, c->tMpcCols.p,
                                   c->tMpcRange.p, cd));
        return P.factor(s);
    }));
    c->thGamma = gamma;
    c->thDt = dt;
    c->thDynSteps = c->thDynGmresSteps = 0;
    c->thDynReady = true;
    return ASTRO_OK;
}

// nSteps transient steps (stepSolve's thermal branch, then the rate update).
int astro_ctx_thermal_dynamic_run(astro_ctx* c, int nSteps) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    if (!c->thDynReady) return fail(ASTRO_ERR_ARG, "astro_ctx_thermal_dynamic_init has not been called");
    if (nSteps < 0) return fail(ASTRO_ERR_ARG, "bad step count");
    hipStream_t s = c->s;
    const Prep& p = c->prep;
    const int N = p.numNodes, cd = p.thermMPCDim;
    const double g = c->thGamma, dt = c->thDt;
    return with_tprof(c, [&](auto& P) -> int {
        for (int k = 0; k < nSteps; ++k) {
            // load = -(K T + C Tdot(T)) element by element + C^T (rhs - C T)
            hipLaunchKernelGGL(k_thermal_rate, dim3(grid_for(N, 256, 4096)), dim3(256), 0, s, N, c->nodeTemp.p,
                               c->prevTemp.p, c->prevTdot.p, g, dt, 0.0, c->tWork.p);
            hipLaunchKernelGGL(k_fill, dim3(grid_for(N, 256, 4096)), dim3(256), 0, s, c->tLoad.p, (int64_t)N, 0.0);
            for (size_t q = 0; q + 1 < c->colorStart.size(); ++q) {
                const int cnt = c->colorStart[q + 1] - c->colorStart[q];
                if (cnt <= 0) continue;
                hipLaunchKernelGGL(k_brick8_thermal_residual, dim3(cnt), dim3(64), 0, s,
                                   c->perm.p + c->colorStart[q], cnt, c->conn.p, c->X.p, c->elCond.p,
                                   (const double*)c->elCap.p, c->rank.p, c->nodeTemp.p, c->tWork.p, c->tLoad.p);
            }
            if (cd > 0)
                hipLaunchKernelGGL(k_constraint_load, dim3(grid_for(cd, 256)), dim3(256), 0, s, c->tMpcMat.p,
                                   c->tMpcCols.p, c->tMpcRange.p, c->tMpcRhs.p, cd, c->nodeTemp.p, c->tLoad.p,
                                   (const int*)nullptr);
            HIP_TRY(hipGetLastError());
            // GMRES from the previous step's increment, as stepSolve passes delDisp
            ASTRO_TRY(c->tgm.solve(c->top, P, c->tLoad.p, c->tDel.p, c->opt.rtol, 0.0, c->opt.maxCycles, s));
            hipLaunchKernelGGL(k_thermal_update, dim3(grid_for(N, 256, 4096)), dim3(256), 0, s, N, c->nodeTemp.p,
                               c->tDel.p, c->prevTemp.p, c->prevTdot.p, c->nodeTdot.p, g, dt);
            HIP_TRY(hipGetLastError());
            c->thDynSteps += 1;
            c->thDynGmresSteps += c->tgm.steps;
        }
        HIP_TRY(hipStreamSynchronize(s));
        return ASTRO_OK;
    });
}

int astro_ctx_thermal_soln_load(astro_ctx* c) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    if (!c->thermalReady) return fail(ASTRO_ERR_ARG, "the thermal system has not been assembled");
    ctx_minus_internal_force(c->s, c->tVal, c->tCols, c->tRange, c->prep.numNodes, c->nodeTemp.p, c->tLoad.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->s));
    return ASTRO_OK;
}

// ---- learned block preconditioner --------------------------------------------
// Node graph, DOF maps and buffers, once per context.
static int ctx_gnn_setup(astro_ctx* c) {
    if (c->gnnReady) return ASTRO_OK;
    if (c->prep.shellMesh) return fail(ASTRO_ERR_UNSUPPORTED, "the learned preconditioner covers brick8 meshes");
    const Prep& p = c->prep;
    const int N = p.numNodes, n = p.elMatDim;
    std::vector<int> nd(3 * (size_t)N, -1), dn(n, -1);
    for (int i = 0; i < N; ++i)
        for (int a = 0; a < 3; ++a) {
            const int d = p.nDofIndex[6 * (size_t)i + a];
            if (d > 0) {
                nd[3 * (size_t)i + a] = d - 1;
                dn[d - 1] = i;
            }
        }
    c->gnnRange.assign(N + 1, 0);
    c->gnnAdj.clear();
    for (int i = 0; i < N; ++i) {
        for (int q = p.nCRange[i]; q < p.nCRange[i + 1]; ++q) {
            const int nb = p.nodalConn[q];
            if (nb > 0 && nb - 1 != i) c->gnnAdj.push_back(nb - 1);
        }
        c->gnnRange[i + 1] = (int)c->gnnAdj.size();
    }
    ASTRO_TRY(c->nodeDofs.upload(nd, c->s));
    ASTRO_TRY(c->dofNode.upload(dn, c->s));
    ASTRO_TRY(c->gnnGraph.build(c->gnnRange, c->gnnAdj, c->s));
    ASTRO_TRY(c->gnnX.alloc((size_t)N * GNN_F));
    ASTRO_TRY(c->gnnG.alloc((size_t)N * 9));
    ASTRO_TRY(c->gnnB.alloc((size_t)N * 9));
    ASTRO_TRY(c->gnnT1.alloc(n));
    ASTRO_TRY(c->gnnT2.alloc(n));
    c->gnnReady = true;
    return ASTRO_OK;
}

// Network inputs of the current operator (the assembled CSR and the scaled
// penalty rows): features, centred log stiffness, and G = L^-T per node.
static int ctx_gnn_features(astro_ctx* c) {
    const int N = c->prep.numNodes, cd = c->prep.elMPCDim;
    hipLaunchKernelGGL(k_gnn_features, dim3(grid_for(N, 256)), dim3(256), 0, c->s, N, (const int*)c->nodeDofs.p,
                       (const int*)c->dofNode.p, (const double*)c->csrVal.p, (const int*)c->csrCols.p,
                       (const int*)c->csrRange.p, cd > 0 ? (const int*)c->mpcT.range.p : (const int*)nullptr,
                       (const int*)c->mpcT.pos.p, (const int*)c->mpcT.row.p, (const double*)c->mpcMat.p, c->gnnX.p,
                       c->gnnG.p);
    HIP_TRY(hipGetLastError());
    std::vector<float> X((size_t)N * GNN_F);
    HIP_TRY(hipMemcpyAsync(X.data(), c->gnnX.p, sizeof(float) * X.size(), hipMemcpyDeviceToHost, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    double sum = 0.0;
    int cnt = 0;
    for (int i = 0; i < N; ++i)
        if (X[(size_t)i * GNN_F + 7] < 0.5f) {
            sum += X[(size_t)i * GNN_F + 6];
            ++cnt;
        }
    const float mu = cnt ? (float)(sum / cnt) : 0.0f;
    hipLaunchKernelGGL(k_gnn_center_log, dim3(grid_for(N, 256)), dim3(256), 0, c->s, N, c->gnnX.p, mu);
    HIP_TRY(hipGetLastError());
    return ASTRO_OK;
}

int astro_ctx_gnn_collect(astro_ctx* c, astro_gnn* g, int nRhs, uint64_t seed, double rtol, int32_t pairs) {
    if (!c || !g) return fail(ASTRO_ERR_ARG, "null argument");
    if (pairs != ASTRO_GNN_PAIRS_SOLUTION && pairs != ASTRO_GNN_PAIRS_SOLVE)
        return fail(ASTRO_ERR_ARG, "pairs must be ASTRO_GNN_PAIRS_SOLUTION or ASTRO_GNN_PAIRS_SOLVE");
    if (nRhs < 1 || (pairs == ASTRO_GNN_PAIRS_SOLVE && !(rtol > 0.0)))
        return fail(ASTRO_ERR_ARG, "need nRhs >= 1 (and rtol > 0 for solved pairs)");
    if (!with_prof(c, [&](auto& P) { return P.factored ? 1 : 0; }))
        return fail(ASTRO_ERR_ARG, "collect needs a factored operator (astro_ctx_dynamic_init or a static factor)");
    ASTRO_TRY(ctx_gnn_setup(c));
    ASTRO_TRY(ctx_gnn_features(c));
    hipStream_t s = c->s;
    const int N = c->prep.numNodes, n = c->prep.elMatDim;
    std::unique_ptr<GnnField> f(new GnnField);
    ASTRO_TRY(f->graph.build(c->gnnRange, c->gnnAdj, s));
    ASTRO_TRY(f->X.alloc((size_t)N * GNN_F));
    ASTRO_TRY(f->G.alloc((size_t)N * 9));
    HIP_TRY(hipMemcpyAsync(f->X.p, c->gnnX.p, sizeof(float) * N * GNN_F, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(f->G.p, c->gnnG.p, sizeof(double) * N * 9, hipMemcpyDeviceToDevice, s));
    // the field's operator A = K + C^T C and DOF map, for the energy-norm loss
    const Prep& p = c->prep;
    const int cd = p.elMPCDim;
    f->n = n;
    f->cdim = cd;
    ASTRO_TRY(f->nodeDofs.alloc((size_t)N * 3));
    ASTRO_TRY(f->val.alloc(p.elMatCols.size()));
    HIP_TRY(hipMemcpyAsync(f->nodeDofs.p, c->nodeDofs.p, sizeof(int) * N * 3, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(f->val.p, c->csrVal.p, sizeof(double) * p.elMatCols.size(), hipMemcpyDeviceToDevice, s));
    ASTRO_TRY(f->cols.upload(p.elMatCols, s));
    ASTRO_TRY(f->range.upload(p.elMatRange, s));
    ASTRO_TRY(f->cmat.alloc(p.elMPCMat.size()));
    if (!p.elMPCMat.empty())
        HIP_TRY(hipMemcpyAsync(f->cmat.p, c->mpcMat.p, sizeof(double) * p.elMPCMat.size(), hipMemcpyDeviceToDevice, s));
    ASTRO_TRY(f->ccols.upload(p.elMPCMatCols, s));
    ASTRO_TRY(f->crange.upload(p.elMPCMatRange, s));
    ASTRO_TRY(f->R1.alloc((size_t)n * nRhs));
    DBuf<double> b, x, z0, t, e1;
    for (DBuf<double>* v : {&b, &x, &z0, &t, &e1}) ASTRO_TRY(v->alloc(n));
    f->eTr = 0.0;
    for (int k = 0; k < nRhs; ++k) {
        // a pair of A x = b: x random and b = A x, or b random and x = A^-1 b by
        // LDL^T-preconditioned GMRES; after the LDL^T stage z0 = P^-1 b the
        // residual is r1 = b - A z0 and the error e1 = x - z0
        const uint64_t sk = seed * 1000003ull + (uint64_t)k;
        if (pairs == ASTRO_GNN_PAIRS_SOLUTION) {
            hipLaunchKernelGGL(k_normal_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, x.p, (int64_t)n, sk);
            ASTRO_TRY(c->op.apply(x.p, b.p, s));
        } else {
            hipLaunchKernelGGL(k_normal_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, b.p, (int64_t)n, sk);
            hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, x.p, (int64_t)n, 0.0);
        }
        ASTRO_TRY(with_prof(c, [&](auto& P) -> int {
            if (pairs == ASTRO_GNN_PAIRS_SOLVE) ASTRO_TRY(c->gm.solve(c->op, P, b.p, x.p, rtol, 0.0, c->opt.maxCycles, s));
            ASTRO_TRY(P.apply(b.p, z0.p, s));
            return P.check_timeout(s);
        }));
        double* r1 = f->R1.p + (size_t)k * n;
        ASTRO_TRY(c->op.apply(z0.p, t.p, s));
        hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, (const double*)b.p,
                           (const double*)t.p, r1, n, 1.0, -1.0, (const int*)nullptr);
        hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, (const double*)x.p,
                           (const double*)z0.p, e1.p, n, 1.0, -1.0, (const int*)nullptr);
        c->gm.multidot(e1.p, 1, r1, c->gm.hdev.p, s);
        HIP_TRY(hipMemcpyAsync(c->gm.hHost, c->gm.hdev.p, sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        f->eTr += c->gm.hHost[0];
    }
    if (!(f->eTr > 0.0)) return fail(ASTRO_ERR_ARG, "collected pairs carry no error after the LDL^T stage");
    f->samples = nRhs;
    g->fields.push_back(std::move(f));
    return ASTRO_OK;
}

int astro_ctx_gnn_attach(astro_ctx* c, astro_gnn* g) {
    if (!c || !g) return fail(ASTRO_ERR_ARG, "null argument");
    ASTRO_TRY(ctx_gnn_setup(c));
    ASTRO_TRY(ctx_gnn_features(c));
    HIP_TRY(hipStreamSynchronize(g->s));
    ASTRO_TRY(g->net.forward(c->gnnGraph, c->gnnX.p, c->gnnG.p, c->gnnB.p, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    c->gnnOn = true;
    return ASTRO_OK;
}

int astro_ctx_gnn_detach(astro_ctx* c) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    c->gnnOn = false;
    return ASTRO_OK;
}

int astro_ctx_gnn_get(astro_ctx* c, const char* name, double* host, int64_t count) {
    if (!c || !name || !host) return fail(ASTRO_ERR_ARG, "null argument");
    ASTRO_TRY(ctx_gnn_setup(c));
    const std::string nm(name);
    const int64_t N = c->prep.numNodes;
    if (nm == "features") {
        if (count != N * GNN_F) return fail(ASTRO_ERR_ARG, "size mismatch for features");
        ASTRO_TRY(ctx_gnn_features(c));
        std::vector<float> X(count);
        HIP_TRY(hipMemcpyAsync(X.data(), c->gnnX.p, sizeof(float) * count, hipMemcpyDeviceToHost, c->s));
        HIP_TRY(hipStreamSynchronize(c->s));
        for (int64_t i = 0; i < count; ++i) host[i] = X[i];
        return ASTRO_OK;
    }
    DBuf<double>* src = nm == "G" ? &c->gnnG : (nm == "blocks" ? &c->gnnB : nullptr);
    if (!src) return fail(ASTRO_ERR_ARG, "unknown learned-preconditioner array " + nm);
    if (count != N * 9) return fail(ASTRO_ERR_ARG, "size mismatch for " + nm);
    HIP_TRY(hipMemcpyAsync(host, src->p, sizeof(double) * count, hipMemcpyDeviceToHost, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    return ASTRO_OK;
}

int astro_ctx_precond_apply(astro_ctx* c, const double* d_b, double* d_x) {
    if (!c || !d_b || !d_x) return fail(ASTRO_ERR_ARG, "null argument");
    // the sweeps mark x unproduced before the forward sweep reads b
    if (d_b == d_x) return fail(ASTRO_ERR_ARG, "precond_apply needs distinct b and x buffers");
    return with_prof(c, [&](auto& P) -> int {
        ASTRO_TRY(P.apply(d_b, d_x, c->s));
        return P.check_timeout(c->s);
    });
}

int astro_ctx_operator_apply(astro_ctx* c, const double* d_x, double* d_y) {
    if (!c || !d_x || !d_y) return fail(ASTRO_ERR_ARG, "null argument");
    ASTRO_TRY(c->op.apply(d_x, d_y, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    return ASTRO_OK;
}

int astro_ctx_set_timing(astro_ctx* c, int on) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    HIP_TRY(hipStreamSynchronize(c->s));
    with_prof(c, [&](auto& P) {
        P.collect();
        P.timing = on != 0;
        P.fwdMs = P.bwdMs = 0;
        P.fwdCount = P.bwdCount = 0;
        return ASTRO_OK;
    });
    return ASTRO_OK;
}

extern "C++" {
template <class Prof>
static double profile_stat(const Prof& P, const std::string& s) {
    if (s == "sweep_fwd_ms") return P.fwdMs;
    if (s == "sweep_bwd_ms") return P.bwdMs;
    if (s == "sweep_fwd_count") return (double)P.fwdCount;
    if (s == "sweep_bwd_count") return (double)P.bwdCount;
    if (s == "profile_tiles") return (double)P.numTiles;
    if (s == "panels_per_block") return (double)P.panelsPerBlock;
    if (s == "split_k") return (double)P.splitK;
    return NAN;
}
}  // extern "C++"

double astro_ctx_kernel_stat(const astro_ctx* c, const char* name) {
    if (!c || !name) return NAN;
    const std::string s(name);
    return c->mixed ? profile_stat(c->prof32, s) : profile_stat(c->prof, s);
}

int astro_ctx_true_residual(astro_ctx* c, double* relRes) {
    if (!c || !relRes) return fail(ASTRO_ERR_ARG, "null argument");
    const int n = c->prep.elMatDim;
    Gmres& g = c->gm;
    // b = nodal loads + C^T rhs, the right-hand side of the whole system
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, c->s, g.t.p, (int64_t)n, 0.0);
    const int nl = (int)c->loadDofs.n, cd = c->prep.elMPCDim;
    if (nl > 0)
        hipLaunchKernelGGL(k_nodal_loads, dim3(grid_for(nl, 256)), dim3(256), 0, c->s, c->loadDofs.p, c->loadVals.p,
                           nl, 1.0, g.t.p, (const int*)nullptr);
    if (cd > 0)
        hipLaunchKernelGGL(k_constraint_load, dim3(grid_for(cd, 256)), dim3(256), 0, c->s, c->mpcMat.p, c->mpcCols.p,
                           c->mpcRange.p, c->mpcRhs.p, cd, (const double*)nullptr, g.t.p, (const int*)nullptr);
    ASTRO_TRY(c->op.residual(g.t.p, c->nodeDisp.p, g.t2.p, c->s));
    g.multidot(g.t2.p, 1, g.t2.p, g.hdev.p, c->s);
    g.multidot(g.t.p, 1, g.t.p, g.hdev.p + 1, c->s);
    HIP_TRY(hipMemcpyAsync(g.hHost, g.hdev.p, 2 * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIP_TRY(hipStreamSynchronize(c->s));
    *relRes = std::sqrt(g.hHost[0]) / std::sqrt(g.hHost[1]);
    return ASTRO_OK;
}

}  // extern "C"

// ---- allocation-free variant of astro_ctx_true_residual ---------------------
// Its working vectors live in one file-scope device array instead of buffers
// the library allocates: the right-hand side b (n doubles), C u as
// double-double pairs (2 cdim), and the block partials and results of the two
// norms.  Each GPU holds its own copy of the array for the life of the
// process; host code reaches it only through the address hipGetSymbolAddress
// returns on that device.
namespace {
constexpr int64_t kNoAllocMaxN = int64_t(1) << 19;     // rows (DOFs)
constexpr int64_t kNoAllocMaxCdim = int64_t(1) << 15;  // constraint rows
constexpr int64_t kNoAllocMaxBlk = 1024;               // reduction blocks (Gmres::nblk <= 1024)
constexpr int64_t kNoAllocCx = 0;
constexpr int64_t kNoAllocPartR = kNoAllocCx + 2 * kNoAllocMaxCdim;
constexpr int64_t kNoAllocPartB = kNoAllocPartR + kNoAllocMaxBlk;
constexpr int64_t kNoAllocNorms = kNoAllocPartB + kNoAllocMaxBlk;
constexpr int64_t kNoAllocB = kNoAllocNorms + 8;
constexpr int64_t kNoAllocDoubles = kNoAllocB + kNoAllocMaxN;  // 4.5 MiB
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noAllocMutex;
double* g_noAllocPtr[kNoAllocMaxDevices] = {};
}  // namespace

static __device__ __align__(256) double scratch[kNoAllocDoubles];

// Device address of `scratch` on device dev (which must be current), looked
// up once per device.
static int noalloc_scratch(int dev, double** out) {
    if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(ASTRO_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_noAllocMutex);
    if (!g_noAllocPtr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ASTRO_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_noAllocPtr[dev] = static_cast<double*>(p);
    }
    *out = g_noAllocPtr[dev];
    return ASTRO_OK;
}

static int true_residual_noalloc(astro_ctx* c, const double* d_u, double* d_r, double* relRes) {
    const int n = c->prep.elMatDim, cd = c->prep.elMPCDim, nl = (int)c->loadDofs.n;
    double* base = nullptr;
    ASTRO_TRY(noalloc_scratch(c->dev, &base));
    double* cx = base + kNoAllocCx;
    double* partR = base + kNoAllocPartR;
    double* partB = base + kNoAllocPartB;
    double* norms = base + kNoAllocNorms;
    double* b = base + kNoAllocB;
    hipStream_t s = c->s;
    // b = nodal loads + C^T rhs; every scratch element read below is written
    // first, so nothing is carried over from the previous call
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, b, (int64_t)n, 0.0);
    if (nl > 0)
        hipLaunchKernelGGL(k_nodal_loads, dim3(grid_for(nl, 256)), dim3(256), 0, s, c->loadDofs.p, c->loadVals.p, nl,
                           1.0, b, (const int*)nullptr);
    if (cd > 0) {
        hipLaunchKernelGGL(k_constraint_load, dim3(grid_for(cd, 256)), dim3(256), 0, s, c->mpcMat.p, c->mpcCols.p,
                           c->mpcRange.p, c->mpcRhs.p, cd, (const double*)nullptr, b, (const int*)nullptr);
        hipLaunchKernelGGL(k_mpc_dd, dim3(grid_for(cd, 256)), dim3(256), 0, s, c->mpcMat.p, c->mpcCols.p,
                           c->mpcRange.p, cd, d_u, cx, (const int*)nullptr);
    }
    // the kernels and reduction order of astro_ctx_true_residual
    hipLaunchKernelGGL(k_residual_dd, dim3(grid_for((int64_t)n * 32, 256, 1 << 30)), dim3(256), 0, s, c->csrVal.p,
                       c->csrCols.p, c->csrRange.p, n, cd > 0 ? c->mpcT.range.p : (const int*)nullptr, c->mpcT.pos.p,
                       c->mpcT.row.p, c->mpcMat.p, (const double*)cx, (const double*)b, d_u, d_r, (const int*)nullptr);
    const int nblk = c->gm.nblk;
    hipLaunchKernelGGL(k_multidot, dim3(nblk, 1), dim3(256), 0, s, (const double*)d_r, (int64_t)n, 1,
                       (const double*)d_r, n, partR, 1, (const int*)nullptr);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, s, (const double*)partR, nblk, 1, 1, norms, (const int*)nullptr);
    hipLaunchKernelGGL(k_multidot, dim3(nblk, 1), dim3(256), 0, s, (const double*)b, (int64_t)n, 1, (const double*)b,
                       n, partB, 1, (const int*)nullptr);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, s, (const double*)partB, nblk, 1, 1, norms + 1, (const int*)nullptr);
    HIP_TRY(hipGetLastError());
    // staged through the context's pinned buffer, allocated at create
    HIP_TRY(hipMemcpyAsync(c->gm.hHost, norms, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *relRes = std::sqrt(c->gm.hHost[0]) / std::sqrt(c->gm.hHost[1]);
    return ASTRO_OK;
}

extern "C" {

int astro_ctx_true_residual_noalloc(astro_ctx* c, const double* d_u, double* d_r, double* relRes) {
    if (!c || !d_u || !d_r || !relRes) return fail(ASTRO_ERR_ARG, "null argument");
    if (d_u == d_r) return fail(ASTRO_ERR_ARG, "true_residual_noalloc needs distinct u and r buffers");
    const int64_t n = c->prep.elMatDim, cd = c->prep.elMPCDim;
    if (n > kNoAllocMaxN)
        return fail(ASTRO_ERR_ARG, "true_residual_noalloc: " + std::to_string(n) + " DOFs exceed its scratch limit of " +
                                       std::to_string(kNoAllocMaxN));
    if (cd > kNoAllocMaxCdim)
        return fail(ASTRO_ERR_ARG, "true_residual_noalloc: " + std::to_string(cd) +
                                       " constraint rows exceed its scratch limit of " + std::to_string(kNoAllocMaxCdim));
    if (c->gm.nblk > kNoAllocMaxBlk) return fail(ASTRO_ERR_ARG, "true_residual_noalloc: too many reduction blocks");
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != c->dev) HIP_TRY(hipSetDevice(c->dev));
    const int rc = true_residual_noalloc(c, d_u, d_r, relRes);
    if (cur != c->dev) HIP_TRY(hipSetDevice(cur));
    return rc;
}

int astro_ctx_sync(astro_ctx* c) {
    if (!c) return fail(ASTRO_ERR_ARG, "null context");
    HIP_TRY(hipStreamSynchronize(c->s));
    return ASTRO_OK;
}

// ===========================================================================
// Layer 1: reference routine replacements on host arrays.

static int check_range(const int32_t* R, int32_t n, int64_t size, const char* what) {
    if (n < 0) return fail(ASTRO_ERR_ARG, std::string(what) + ": negative dimension");
    if (R[0] != 0) return fail(ASTRO_ERR_ARG, std::string(what) + ": range(0) must be 0");
    for (int i = 1; i <= n; ++i)
        if (R[i] < R[i - 1]) return fail(ASTRO_ERR_ARG, std::string(what) + ": ranges not monotone");
    if ((int64_t)R[n] > size) return fail(ASTRO_ERR_ARG, std::string(what) + ": range end exceeds array size");
    return ASTRO_OK;
}

int astro_getSparseLDLFact(double* ALTri, int64_t ASize, const int32_t* ARange, int32_t n) {
    if (!ALTri || !ARange) return fail(ASTRO_ERR_ARG, "null argument");
    ASTRO_TRY(check_range(ARange, n, ASize, "getSparseLDLFact"));
    if (n == 0) return ASTRO_OK;
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int rc;
    {
        DeviceProfile P;
        rc = P.setup(ARange, n, s);
        if (rc == ASTRO_OK && hipMemcpyAsync(P.lt.p, ALTri, sizeof(double) * P.size, hipMemcpyHostToDevice, s) != hipSuccess)
            rc = fail(ASTRO_ERR_DEVICE, "upload of ALTri failed");
        if (rc == ASTRO_OK) rc = P.factor(s);
        if (rc == ASTRO_OK || rc == ASTRO_ERR_ZERO_PIVOT) {
            if (hipMemcpyAsync(ALTri, P.lt.p, sizeof(double) * P.size, hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                rc = fail(ASTRO_ERR_DEVICE, "download of the factor failed");
        }
    }
    (void)hipStreamDestroy(s);
    return rc;
}

int astro_solveSparseLDLFact(double* soln, const double* rhs, const double* Lmat, int64_t LSize,
                             const int32_t* LRange, int32_t Ldim, double* zVec) {
    if (!soln || !rhs || !Lmat || !LRange) return fail(ASTRO_ERR_ARG, "null argument");
    ASTRO_TRY(check_range(LRange, Ldim, LSize, "solveSparseLDLFact"));
    if (Ldim == 0) return ASTRO_OK;
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int rc;
    {
        DeviceProfile P;
        DBuf<double> b, x;
        rc = P.setup(LRange, Ldim, s);
        if (rc == ASTRO_OK) rc = b.upload(rhs, Ldim, s);
        if (rc == ASTRO_OK) rc = x.alloc(Ldim);
        if (rc == ASTRO_OK && hipMemcpyAsync(P.lt.p, Lmat, sizeof(double) * P.size, hipMemcpyHostToDevice, s) != hipSuccess)
            rc = fail(ASTRO_ERR_DEVICE, "upload failed");
        if (rc == ASTRO_OK) rc = P.apply(b.p, x.p, s);
        if (rc == ASTRO_OK) rc = P.check_timeout(s);
        if (rc == ASTRO_OK) {
            if (hipMemcpyAsync(soln, x.p, sizeof(double) * Ldim, hipMemcpyDeviceToHost, s) != hipSuccess ||
                (zVec && hipMemcpyAsync(zVec, P.z.p, sizeof(double) * Ldim, hipMemcpyDeviceToHost, s) != hipSuccess) ||
                hipStreamSynchronize(s) != hipSuccess)
                rc = fail(ASTRO_ERR_DEVICE, "download failed");
        }
    }
    (void)hipStreamDestroy(s);
    return rc;
}

int astro_gMRes(double* xVec, const double* Amat, const int32_t* Acols, const int32_t* ARange, int64_t ASize,
                int32_t ADim, const double* Cmat, const int32_t* Ccols, const int32_t* CRange, int64_t CSize,
                int32_t CDim, const double* Pmat, const int32_t* PRange, int64_t PSize, const double* bVec,
                int32_t numVecs, int32_t maxIt, double rtol, double atol) {
    if (!xVec || !Amat || !Acols || !ARange || !Pmat || !PRange || !bVec) return fail(ASTRO_ERR_ARG, "null argument");
    if (numVecs < 1 || numVecs > 255) return fail(ASTRO_ERR_ARG, "numVecs must be in [1, 255]");
    ASTRO_TRY(check_range(ARange, ADim, ASize, "gMRes A"));
    ASTRO_TRY(check_range(PRange, ADim, PSize, "gMRes P"));
    if (CDim > 0) {
        if (!Cmat || !Ccols || !CRange) return fail(ASTRO_ERR_ARG, "null constraint arrays");
        ASTRO_TRY(check_range(CRange, CDim, CSize, "gMRes C"));
    }
    for (int64_t q = 0; q < ARange[ADim]; ++q)
        if (Acols[q] < 0 || Acols[q] > ADim) return fail(ASTRO_ERR_ARG, "A column out of range");
    for (int64_t q = 0; CDim > 0 && q < CRange[CDim]; ++q)
        if (Ccols[q] < 0 || Ccols[q] > ADim) return fail(ASTRO_ERR_ARG, "C column out of range");
    if (ADim == 0) return ASTRO_OK;
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int rc;
    {
        DeviceProfile P;
        DBuf<double> av, cv, b, x;
        DBuf<int> ac, ar, cc, cr;
        DeviceMpcT ct;
        Gmres gm;
        DeviceOperator op;
        rc = P.setup(PRange, ADim, s);
        if (rc == ASTRO_OK && hipMemcpyAsync(P.lt.p, Pmat, sizeof(double) * P.size, hipMemcpyHostToDevice, s) != hipSuccess)
            rc = fail(ASTRO_ERR_DEVICE, "upload failed");
        if (rc == ASTRO_OK) rc = av.upload(Amat, ARange[ADim], s);
        if (rc == ASTRO_OK) rc = ac.upload(Acols, ARange[ADim], s);
        if (rc == ASTRO_OK) rc = ar.upload(ARange, ADim + 1, s);
        if (rc == ASTRO_OK && CDim > 0) rc = cv.upload(Cmat, CRange[CDim], s);
        if (rc == ASTRO_OK && CDim > 0) rc = cc.upload(Ccols, CRange[CDim], s);
        if (rc == ASTRO_OK && CDim > 0) rc = cr.upload(CRange, CDim + 1, s);
        if (rc == ASTRO_OK && CDim > 0) rc = ct.upload(CRange, Ccols, CDim, ADim, s);
        if (rc == ASTRO_OK) rc = b.upload(bVec, ADim, s);
        if (rc == ASTRO_OK) rc = x.upload(xVec, ADim, s);
        if (rc == ASTRO_OK) rc = gm.setup(ADim, numVecs);
        if (rc == ASTRO_OK) {
            op.n = ADim;
            op.cdim = CDim > 0 ? CDim : 0;
            op.val = av.p;
            op.cols = ac.p;
            op.range = ar.p;
            op.cmat = cv.p;
            op.ccols = cc.p;
            op.crange = cr.p;
            if (CDim > 0) op.set_transpose(ct);
            const int maxCycles = maxIt / numVecs + 1;
            if (rtol <= 0 && atol <= 0) atol = (double)1e-12f;
            rc = gm.solve(op, P, b.p, x.p, rtol > 0 ? rtol : 0.0, atol, maxCycles, s);
        }
        if (rc == ASTRO_OK || rc == ASTRO_ERR_NO_CONVERGE) {
            if (hipMemcpyAsync(xVec, x.p, sizeof(double) * ADim, hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                rc = fail(ASTRO_ERR_DEVICE, "download failed");
        }
    }
    (void)hipStreamDestroy(s);
    return rc;
}

int astro_convertToLTri(double* lTri, const int32_t* lTRange, int64_t lTSize, const double* mainMat,
                        const int32_t* mainCols, const int32_t* mainRange, int64_t mainSize, int32_t mainDim,
                        const double* cMat, const int32_t* cCols, const int32_t* cRange, int64_t cSize,
                        int32_t cDim) {
    if (!lTri || !lTRange || !mainMat || !mainCols || !mainRange) return fail(ASTRO_ERR_ARG, "null argument");
    ASTRO_TRY(check_range(lTRange, mainDim, lTSize, "convertToLTri profile"));
    ASTRO_TRY(check_range(mainRange, mainDim, mainSize, "convertToLTri CSR"));
    if (cDim > 0) {
        if (!cMat || !cCols || !cRange) return fail(ASTRO_ERR_ARG, "null constraint arrays");
        ASTRO_TRY(check_range(cRange, cDim, cSize, "convertToLTri C"));
        for (int64_t q = 0; q < cRange[cDim]; ++q)
            if (cCols[q] < 0 || cCols[q] > mainDim) return fail(ASTRO_ERR_ARG, "C column out of range");
    }
    if (mainDim == 0) return ASTRO_OK;
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int rc = ASTRO_OK;
    {
        std::vector<int> sf(mainDim);
        std::vector<int64_t> off(mainDim);
        for (int r = 0; r < mainDim; ++r) {
            int len = lTRange[r + 1] - lTRange[r];
            if (len < 1 || len > r + 1) rc = fail(ASTRO_ERR_ARG, "invalid profile row length");
            sf[r] = r - len + 1;
            off[r] = (int64_t)lTRange[r + 1] - 1 - r;
        }
        DBuf<double> lt, mv, cv;
        DBuf<int> mc, mr, cc, cr, sfd;
        DBuf<int64_t> offd;
        const int64_t sz = lTRange[mainDim];
        if (rc == ASTRO_OK) rc = lt.alloc(sz);
        if (rc == ASTRO_OK) rc = sfd.upload(sf, s);
        if (rc == ASTRO_OK) rc = offd.upload(off, s);
        if (rc == ASTRO_OK) rc = mv.upload(mainMat, mainRange[mainDim], s);
        if (rc == ASTRO_OK) rc = mc.upload(mainCols, mainRange[mainDim], s);
        if (rc == ASTRO_OK) rc = mr.upload(mainRange, mainDim + 1, s);
        if (rc == ASTRO_OK && cDim > 0) rc = cv.upload(cMat, cRange[cDim], s);
        if (rc == ASTRO_OK && cDim > 0) rc = cc.upload(cCols, cRange[cDim], s);
        if (rc == ASTRO_OK && cDim > 0) rc = cr.upload(cRange, cDim + 1, s);
        PenaltyPlan pen;
        if (rc == ASTRO_OK && cDim > 0) rc = pen.build(cRange, cCols, cDim, sf, s);
        if (rc == ASTRO_OK) {
            hipLaunchKernelGGL(k_fill, dim3(grid_for(sz, 256, 16384)), dim3(256), 0, s, lt.p, sz, 0.0);
            hipLaunchKernelGGL(k_csr_to_profile<double>, dim3(grid_for((int64_t)mainDim * 32, 256, 1 << 30)), dim3(256), 0, s,
                               mv.p, mc.p, mr.p, mainDim, sfd.p, offd.p, lt.p);
            if (cDim > 0) pen.launch(cv.p, offd.p, lt.p, s);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(lTri, lt.p, sizeof(double) * sz, hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                rc = fail(ASTRO_ERR_DEVICE, "convertToLTri device error");
            for (int64_t q = sz; q < lTSize; ++q) lTri[q] = 0.0;
        }
    }
    (void)hipStreamDestroy(s);
    return rc;
}

}  // extern "C"
