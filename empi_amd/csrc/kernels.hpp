// Device-side data layout shared by the host plan and the kernels.
//
// HBM layout of one batch (all arrays flat, offsets per unit):
//   residual   FP64 [unit][channel][sample]        the MP residual
//   cellbest   u64  [unit][scale][t][g]            best product of one cell
//              (high 32 bits: FP32 product bits, low 32 bits: ~bin)
//   tilebest   u64  [unit][scale][tile]            best of a workgroup's cells
//              (high: product bits, low: ~(bin << kTileCellBits | cell in tile))
//   atoms      empi_atom [item][channel][iteration]
//   energies   FP64 [item][row][iteration + 1]
// A "unit" is one MP loop: one channel of an SMP item, the channel sum of an
// MMP2 item, or all channels of an MMP1/MMP3 item.  A "cell" is the part of
// one map column (time index t) that one workgroup slot pair transforms:
// sub-transforms k2 = {g, P-g} (g = 0 holds {0, P/2}).
#pragma once

#include <cstdint>

#include "empi_hip.h"

namespace empi {

// Low bits of a tile-best key holding the cell within the tile (at most 256
// cells per tile); the bin takes the other 24 bits.
constexpr int kTileCellBits = 8;

struct ScaleDev {
    double s, fs, tMax;
    long long cells_before;   // cell offset of this scale inside its unit
    int N, Nfft, M, T, F, Q, P, G;
    int lgQ, lgNfft;
    float alpha;              // s*fs/Nfft: s*f4 = j*alpha
    float Kp;                 // 2*sqrt(2)/(s*fs)
    float gA;                 // pi/(s*fs)^2
    int Wmax;
    int fold;                 // Q-point sections of the packed window input (>= 1)
    float invKp;              // 1/K'
    int JT;                   // bins j < JT carry E = exp(-2 pi (j alpha)^2) > 0 in FP32
    const float2* etab;       // [JT]: ((1 - E)/K', 2E/K') in FP32 from FP64
    float csnorm;             // max discrete/continuous atom-norm ratio: P <= csnorm * window energy
    // proposal bound: bins j >= JB have E_j <= 1e-4; L(E) the Lipschitz
    // constant of sqrt(P) in |u| = sqrt(K') |X| at window term E:
    // gH = L(E_JB), lL = L(E_1), gL = 1/sqrt(1 - E_1) (the largest over 1 <= j < JB)
    int JB;
    float gH, gL, lL;
    float sgA;                // sqrt(gA log2 e): the window is exp2(-(sgA u)^2)
};

struct UnitDev {
    long long res_off;        // residual offset (doubles)
    long long cell_off;       // cellbest offset
    long long in_off;         // input offset (floats) of the item's channel 0
    int C;                    // channels of the unit
    int N;
    int in_channel;           // first input channel read by this unit
    int in_sum;               // number of input channels summed into one (MMP2), else 0
    int scale0, L;            // scales of its dictionary in the ScaleDev array
    int constraint;           // 1 = MMP1
    int iterMax;
    double fs;
    double residualEnergy;
    empi_atom* atoms;         // first atom slot of the unit's channel 0
    int atom_cstride;         // atoms between consecutive channels
    int pad;
    double* energies;         // [iterMax + 1]
    double2* fitv;            // scratch: C complex values of the chosen column
    long long us_off;         // first UnitScale of the unit
    long long reg_off;        // first region of the unit (proposal mode)
    int B;                    // 256-sample blocks of the unit
    int regtab_off;           // the dictionary's region table: [scale][block]
};

struct UnitState {
    int active;
    int iteration;
    int status;
    int pad;
    double totalEnergy;
    double residEnergy;
    double tA, sA;            // last subtracted atom, seconds
    float xscale;             // power-of-two input scaling of the FP32 maps
    int pad2;
    int sel_k, sel_f, sel_t;  // the chosen entry, handed from k_select to the per-channel kernels
    int sel_pad;
};

struct Job {                  // one (unit, scale) pair in a kernel class
    int unit;
    int scale;                // global ScaleDev index
    long long cell_off;       // absolute cellbest offset of this scale
    long long tile_off;       // absolute tilebest offset of this scale
    int tiles;                // ceil(T*G / cells per workgroup)
    int lscale;               // scale index inside the unit's dictionary
};

// One contiguous run of tiles of one job to re-evaluate; cells outside
// [tlo, thi] (and, for atom items, outside the reference's overlap test) are
// left alone.
struct WorkItem {
    int job;
    int tile;                 // first tile
    int tiles;
    int tlo, thi;             // time-index range
    int atom;                 // 1: apply |t - tA| <= 3 (s + sA) per cell
    int region;               // region index (proposal items), else -1
};

// Work list of one kernel class, rebuilt every iteration:
// work[i] = tiles of items before i (work[*count] = total).
struct ClassDev {
    const Job* jobs;
    WorkItem* items;
    int* work;
    int* count;               // items in the list
    int njobs;
    int max_items;
    int ncell;                // cells per workgroup of the class
    int cls;
};

// time-index range of one (scale, 256-sample block) region of a dictionary
struct RegionDev {
    int tlo, thi;             // thi < tlo: no column of the scale falls in the block
};

// per (unit, scale): where the selection finds the scale's tile bests
struct UnitScale {
    long long tile_off;
    int tiles;
    int ncell;
    int cls_slot;             // ClassDev slot of the scale's kernel class
    int job;                  // job index inside that class
};

}  // namespace empi
