// HIP kernels of the DEM timestep (gfx950).  Included by wdm_lib.hip only.
//
// Layout in HBM (one scene):
//   node arrays, N entries, 32-byte records:  posr = (x, y, z, radius|halfThick),
//     velm = (vx, vy, vz, mass), angI = (wx, wy, wz, inertia), ori = (w, x, y, z),
//     force/torque = (x, y, z, 0), lastPos = (x, y, z, 0); flags u32.
//   particle arrays, P entries: ptype i32 (shape | axis<<4 | (sense+1)<<6), mask u32,
//     mat i32, facet nodes int4; stored bounds bmin/bmax double4.
//   contacts, C entries sorted by key = (min id << 32 | max id): ids int2 (functor
//     order), info u32, packed cellDist i32, state row i32; state (trsf columns 0-1,
//     local force, phys): the rows handed out at a full run (real and touching contacts,
//     dense in key order) component-major T[NSTATE][cap], the rows after them
//     contact-major T[NSTRIDE] each; contact point per state row (FP64); work list of the
//     ContactLoop: contacts given a row at the full run first (item == row), then the other
//     non-dormant ones (item i owns row i), dormant contacts left out; per item the record
//     (ids, info, cellDist) and (row, contact), and the gather record gfc[2i] = (global
//     force, cellDist), gfc[2i+1] = (contact point, uN).
//   CSR: off[P+1], entries u32 (contact index | isB << 31), per particle sorted by
//     the other particle's id.
// Batches: every kernel takes a table of per-scene descriptors (views of the arrays
// above) and a block prefix table; block b of a launch works on scene t with
// bprefix[t] <= b < bprefix[t+1], so many small scenes share one launch and each scene
// sees the same block layout (and the same reduction order) as when it runs alone.
#pragma once
#include "wdm_math.h"

#define WDM_BLOCK 256

// In the device pass, the array pointers held in the descriptor structs below are
// declared in the global address space, so the compiler emits global (not flat) memory
// instructions for them; the host pass sees plain pointers of the same size and layout.
#if defined(__HIP_DEVICE_COMPILE__)
#define GP __attribute__((address_space(1)))
#else
#define GP
#endif
template <typename T>
__host__ __device__ inline GP T* gpc(T* p) { return (GP T*)p; }

// contact info bits
#define CI_TYPE_MASK 3u
#define CI_SS 0u
#define CI_WS 1u
#define CI_FS 2u
#define CI_NONE 3u
#define CI_REAL 4u
#define CI_PENDING 8u
#define CI_OVERLAP 16u
#define CI_DEAD 32u
#define CI_DORMANT 64u  // cannot touch before the next full run (gap > 2 verletDist at this run)
#define CI_NEAR 128u    // surface gap below NEAR_GAP verletDist at this run: early in the work list

// contact state components
#define S_T 0      // trsf columns 0 and 1 (x, y, z each); column 2 is their cross product
#define S_F 6      // local force, 3 entries
#define S_P 9      // phys, 5 entries
#define S_L 14     // L6Geom.lens, 2 entries (written at creation, read by DynDt)
#define NSTATE 16
#define NSTRIDE 16  // values per contact-major state row: 128 bytes in FP64

// device error bits
#define ERR_NAN 1u
#define ERR_OVERSIZE 2u
#define ERR_CELLDIST 4u

// node flags
#define F_DOF_ALL 63u
#define F_DOF_RXRYRZ 56u
#define F_CLUMPED 64u
#define F_CLUMP 128u
#define F_ENERGY_SKIP 256u
#define F_GRAVITY_SKIP 512u
#define F_DAMPING_SKIP 2048u

// energy slots (same order as the ABI)
#define E_KIN 0
#define E_ELAST 1
#define E_PLAST 2
#define E_GRAV 3
#define E_DAMP 4
#define E_VISCN 5
#define E_VISCT 6
#define E_BROKEN 7
#define E_N 8

__device__ __forceinline__ int ptShape(int32_t t) { return t & 15; }
__device__ __forceinline__ int ptAxis(int32_t t) { return (t >> 4) & 3; }
__device__ __forceinline__ int ptSense(int32_t t) { return ((t >> 6) & 3) - 1; }

WDM_HD int32_t packCd(int a, int b, int c) { return (a & 1023) | ((b & 1023) << 10) | ((c & 1023) << 20); }
WDM_HD int unpack10(int32_t v, int sh) { int x = (v >> sh) & 1023; return x >= 512 ? x - 1024 : x; }

struct NodeView {
    GP double4 *posr, *velm, *angI, *ori, *force, *torque, *lastPos;
    GP uint32_t* flags;
    GP uint8_t* checked;  // node belongs to a particle with a bound (Verlet check)
    GP int32_t* master;   // clump master node of a clumped node, -1 otherwise (null: no clumps)
    int nn;
};
// clumps (ClumpData, pkg/dem/Clump.hpp): members of clump c are mem[off[c]..off[c+1]),
// with their position and orientation relative to the master node
struct ClumpView {
    const GP int32_t* off;
    const GP int32_t* mem;
    const GP double4* relPos;   // (x, y, z, 0)
    const GP double4* relOri;   // (w, x, y, z)
    const GP int32_t* nodeClump;  // clump index of a master node, -1 otherwise
    GP double4* ft;               // per clump: master force, torque with the members' sums (2 records)
    int nClumps, nMembers;
};
struct PartView {
    GP int32_t* ptype;
    GP uint32_t* mask;
    GP int32_t* mat;
    GP int4* fnodes;
    GP double4 *bmin, *bmax;
    int np;
};
template <typename T>
struct ConView {
    GP uint64_t* key;
    GP int2* ids;
    GP uint32_t* info;
    GP int32_t* slot;  // row of the contact's geometry/physics state, -1 when it has none
    int32_t* counter;  // state rows handed out at the last full run (real and touching contacts)
    GP T* st2;         // state of the rows from *counter on (items that were not touching at the full run)
    GP int32_t* cd;
    GP T* st;
    GP double4* cpt;   // per state row: contact point and uN of the last ContactLoop (x, y, z, uN; mixed: mixedCpt)
    GP int32_t* itm;   // per contact: its item in the ContactLoop's work list, -1 when dormant
    // per item, so that the ContactLoop reads them contiguously: (idA, idB, info, cellDist)
    // and (state row, contact); info and the row are written to both places
    GP int4* irec;
    GP int2* iaux;
    int32_t* nItems;   // items in the work list
    GP double4* gfc;   // per item: gather record (mixed: mixedRecA/B/C)
    int64_t cap;
};

// State value k of row `row`.  Rows handed out at the full run (nR of them: real and
// touching contacts, dense in key order) are component-major, so a wavefront of them reads
// each value with one contiguous access; the rows after them belong to items that may
// become real later, scattered among items that stay apart, and are contact-major, so that
// such a contact reads one 128-byte line instead of a line per value.
template <typename T>
__device__ __forceinline__ GP T* stAt(const ConView<T>& cv, int64_t row, int k, int64_t nR) {
    return row < nR ? cv.st + (size_t)k * cv.cap + row : cv.st2 + (size_t)(row - nR) * NSTRIDE + k;
}

__device__ __forceinline__ M3 ldM3(const GP M3& x) {
    M3 r;
    for (int k = 0; k < 9; k++) r.m[k] = x.m[k];
    return r;
}

// cell quantities that change every step (periodic scenes; zero/identity otherwise)
struct CellArgs {
    // Leapfrog (pkg/dem/Leapfrog.cpp:113-126)
    M3 ImLL4hInv, IpLL4h, LmL, dGradV, midGradV, nextGradV, gradV;
    double deltaSpinVec[3], spinVec[3];
    M3 hSize, shiftVelM;  // shiftVelM: gradV*pprevHsize (GRADV2) or gradV*hSize (HOMO_VEL*), else 0
    M3 gatherHSize;       // hSize when the gathered contact forces were computed (ContactLoop of the previous step)
    double cellSize[3];
};

// per-scene engine attributes and per-step flags
struct StepArgs {
    double dt, damping;
    double g[3];
    int hasGravity;
    int periodic, homoDeform;  // homoDeform: -1 aperiodic
    int reallyTrackEnergy, trackEnergy;
    // collider
    double verletDist, verletD2;
    int checkVerlet;
    uint32_t stamp;  // nonzero tag of this step: the Verlet trigger word holds it when a node moved too far
    int gatherOn;  // node forces come from the contacts of the previous ContactLoop
    int hasClumps;
    // contact
    double distFactor;
    int law, noSlip, noBreak, noFrict, noAttraction, trsfRenorm;
    int64_t step;
    uint32_t loneMask;
    double viscAlpha;  // -sqrt(5) log(en) / sqrt(log(en)^2 + pi^2), or 0 when en is outside (0,1)
    double poisson, gamma;
    int hertzDamped;
    GP double4* mat;  // material table (density, young, ktDivKn, tanPhi)
    const GP CellArgs* cell;
};

struct GridArgs {
    double origin[3];
    double w[3];   // cell width per axis
    int n[3];      // cells per axis
    int periodic;
    double L[3];     // periodic cell size
    double invL[3];  // 1/L (prefilter only)
    float Lf[3], invLf[3];  // the same in FP32 (FP32 prefilter)
    float pfMargin;         // FP32 prefilter slack: far above its rounding (~1e-7 of the grid span)
};


// force gather views (CSR contact -> particle, realness mask, gather records)
struct GatherView {
    const GP int64_t* off;
    const GP uint32_t* ent;
    const GP double4* gfc;
    const GP uint64_t* realMask;
    int64_t gcap;  // items the gather record buffer holds
    int mixed;     // FP32 gather records (mixed precision)
};
// Mixed precision keeps the FP64 records' buffers (contact point: one double4 per state row;
// gather record: two double4 per item) but fills them in FP32, in fewer bytes:
//   contact point: float4 (contact point - position of B in A's image, uN) per row;
//   gather record: 36 bytes per item in three arrays over the buffer: float4 (global force,
//     branch of A x), float4 (branch of A y z, branch of B x y), float (branch of B z),
//     a branch being the contact point minus the node position (B's in A's image).
__host__ __device__ inline GP float4* mixedCpt(GP double4* cpt) { return (GP float4*)cpt; }
__host__ __device__ inline GP float4* mixedRecA(GP double4* gfc) { return (GP float4*)gfc; }
__host__ __device__ inline GP float4* mixedRecB(GP double4* gfc, int64_t cap) { return (GP float4*)gfc + cap; }
__host__ __device__ inline GP float* mixedRecC(GP double4* gfc, int64_t cap) { return (GP float*)((GP float4*)gfc + 2 * cap); }

// One scene of a batch.  Kernels read their views from a table of these, so one launch
// covers many scenes; blocks are mapped to scenes by a prefix table (mapBlock).
struct SceneDesc {
    StepArgs a;
    NodeView n;
    PartView p;
    GatherView gv;      // contacts of the last ContactLoop (forces gathered by the next Leapfrog)
    const GP int32_t* binned;  // particles with finite bounds (spheres, facets)
    const GP int32_t* unb;     // particles with infinite bounds (walls)
    int nbinned, nunb;
    GP double* E;       // energy accumulators [E_N]
    uint32_t* words;    // [0] Verlet trigger, [1] created energy slots, [2] device errors
    ClumpView cl;
};

// Several steps of a batch launched back to back without the host (one captured graph):
// the per-step fields of each scene's StepArgs come from this table, one row per step.
// A scene whose Verlet trigger fires stops there: its trigger word then holds a stamp at
// or above `gate` (the stamp of the window's first step), its later steps in the window
// do nothing, and the host runs its collider before it resumes.
struct StepVar {
    int64_t step;
    const GP CellArgs* cell;
    uint32_t stamp;  // stamp of this step
    uint32_t gate;   // stamp of the window's first step
    int32_t reallyTrackEnergy;
    int32_t idle;    // the scene has no step at this position of the window
};
__device__ __forceinline__ uint32_t loadWord(const uint32_t* w) {
    return __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Leapfrog side of a step (clump sums, Leapfrog, clump members, their energies): runs unless
// the trigger fired at an earlier step of the window.  The word only moves from below the
// gate to this step's stamp while the step runs, so every block decides alike.
__device__ __forceinline__ bool leapActive(const StepVar& v, const uint32_t* words) {
    if (v.idle) return false;
    const uint32_t w = loadWord(words);
    return w < v.gate || w == v.stamp;
}
// ContactLoop side: runs when no trigger fired in the window up to and including this step
__device__ __forceinline__ bool contactActive(const StepVar& v, const uint32_t* words) {
    return !v.idle && loadWord(words) < v.gate;
}
__device__ __forceinline__ void applyStepVar(StepArgs& a, const StepVar& v) {
    a.step = v.step;
    a.cell = v.cell;
    a.stamp = v.stamp;
    a.reallyTrackEnergy = v.reallyTrackEnergy;
}

template <typename T>
struct ConDesc {
    ConView<T> cv;      // current contact list
    ConView<T> cvNew;   // merge target of a collider full run
    int64_t nc;
    GP uint64_t* realMask;  // one bit per contact, written by the contact kernel
    int32_t desc;       // index in the SceneDesc table
};
// one scene of a batched collider full run: its grid and its segments in the batch-wide
// scratch arrays (bound entries in cell order, candidate pairs, old and new contacts)
// The collider's grid of a scene has nLvl levels of cells, level l of cell width lw[l]
// (twice the width of level l-1), holding the bounds whose largest extent is at most
// lthr[l] = lw[l] / 1.001 (and above lthr[l-1]); bounds wider than the top level's
// threshold (facets larger than every sphere) are "large" and kept out of the cells.
// Cell keys of the scene: lbase[l] + (z n_y + y) n_x + x at level l, and lbase[nLvl] ==
// nCells for the large bounds.  g holds what the levels share (origin, periodic sizes,
// FP32 slack) and level 0's cells.
#define MAXLVL 4
struct FRScene {
    GridArgs g;
    int64_t cellBase, binStart, candStart, nCand, oldStart, nOld, newStart, nNew;
    int64_t rowBase;  // first row of the scene in batch-wide per-particle arrays (np + 1 rows per scene)
    GP int64_t* csrOff;
    GP uint32_t* entV;
    double wLarge;  // bounds wider than this (facets larger than every sphere) skip the grid cells
    double lw[MAXLVL][3];
    double lthr[MAXLVL];
    int32_t ln[MAXLVL][3];
    int32_t lbase[MAXLVL + 1];
    int32_t nLvl;
    int32_t nCells;      // grid cells of the scene; cell nCells lists its large bounds
    int32_t largePairs;  // large bounds may pair with each other (masks allow it)
    int32_t largeListed; // the pass lists large bounds in the cells they cover (k_large_cells)
    int32_t desc;
};

// blocks of one launch spread over several scenes: scene t owns blocks
// [bprefix[t], bprefix[t+1]) and its elements are numbered from 0 inside them
struct Map {
    const int64_t* bprefix;
    int n;
};
// Every thread searches (the same uniform loads, no barrier); the result is uniform, so
// the descriptor fields that follow are read with scalar loads.
__device__ __forceinline__ int mapBlockAt(const Map& m, int64_t b, int64_t& lblk) {
    int lo = 0, hi = m.n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (m.bprefix[mid] <= b) lo = mid; else hi = mid - 1;
    }
    lblk = b - m.bprefix[lo];
    return lo;
}
__device__ __forceinline__ int mapBlock(const Map& m, int64_t& lblk) { return mapBlockAt(m, blockIdx.x, lblk); }
// XCD-chunked block order: the hardware hands the blocks of a launch to the 8 XCDs round
// robin (block b to XCD b % 8), so consecutive blocks, which work on neighbouring particles
// and contacts, land in different L2 caches.  With a chunk of k blocks, XCD x takes logical
// blocks [(8 g + x) k, (8 g + x + 1) k) from its g-th group of k launched blocks, so k
// neighbouring blocks share one L2 while the 8 XCDs still advance through the work together
// (k = 0: plain order; k < 0: one contiguous eighth per XCD).  Returns -1 for padding blocks.
#ifndef XCD_CHUNK_LEAP
#define XCD_CHUNK_LEAP (-1)
#endif
#ifndef XCD_CHUNK_CON
#define XCD_CHUNK_CON 0
#endif
__host__ __device__ inline int64_t xcdChunkOf(int64_t n, int64_t k) { return k < 0 ? (n + 7) >> 3 : k; }
__device__ __forceinline__ int64_t xcdChunkBlock(int64_t n, int64_t k) {
    const int64_t b = blockIdx.x;
    if (k == 0) return b < n ? b : -1;
    k = xcdChunkOf(n, k);
    const int64_t r = b >> 3;
    const int64_t l = ((r / k) * 8 + (b & 7)) * k + r % k;
    return l < n ? l : -1;
}
__host__ __device__ inline int64_t xcdChunkGrid(int64_t n, int64_t k) {
    if (k == 0) return n;
    k = xcdChunkOf(n, k);
    return (n + 8 * k - 1) / (8 * k) * (8 * k);
}

// ---------------- block reductions (fixed tree: deterministic) ----------------
// bits: (1 << energy slot) for "slot created" (EnergyTracker::add, core/EnergyTracker.cpp:10-17),
// VERLET_BIT for the Verlet trigger (the trigger word then receives the step's stamp).
#define VERLET_BIT 0x80000000u
template <int NV>
__device__ void blockSumWrite(double (&v)[NV], uint32_t orbits, double* partial, uint32_t* trigOut, uint32_t stamp, uint32_t* creOut,
                              int64_t blk = -1) {
    __shared__ double sh[NV][WDM_BLOCK];
    __shared__ uint32_t shb[WDM_BLOCK];
    int t = threadIdx.x;
    for (int k = 0; k < NV; k++) sh[k][t] = v[k];
    shb[t] = orbits;
    __syncthreads();
    for (int s = WDM_BLOCK / 2; s > 0; s >>= 1) {
        if (t < s) {
            for (int k = 0; k < NV; k++) sh[k][t] = sh[k][t] + sh[k][t + s];
            shb[t] |= shb[t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        const size_t pb = blk >= 0 ? (size_t)blk : (size_t)blockIdx.x;
        if (partial) for (int k = 0; k < NV; k++) partial[pb * NV + k] = sh[k][0];
        uint32_t b = shb[0];
        if ((b & VERLET_BIT) && trigOut) atomicMax(trigOut, stamp);
        if ((b & ~VERLET_BIT) && creOut) atomicOr(creOut, b & ~VERLET_BIT);
    }
}

// ---------------- force application (ContactLoop::applyForceUninodal, ContactLoop.cpp:261-267)
// on top of the force reset of Leapfrog (Leapfrog.cpp:240-245).  A uninodal particle's
// node sums its real contacts in ascending order of the other particle's id (CSR
// order); F = trsf*Flocal and the contact point come from the contact's gather
// record, the periodic shift of the B side from the cellDist packed in its w slot.
#ifndef GATHER_CHUNK
#define GATHER_CHUNK 2
#endif
__device__ __forceinline__ void gatherForces(const StepArgs& a, const GatherView& gv, int64_t e0, int64_t e1, bool sums,
                                             uint32_t fl, double m, const D3& pos, D3& F, D3& T) {
    // gravity acts on a clump's master only (ClumpData::applyToMembers resets the members)
    F = (a.hasGravity && !(fl & (F_GRAVITY_SKIP | F_CLUMPED))) ? mk3(a.g[0] * m, a.g[1] * m, a.g[2] * m) : mk3(0., 0., 0.);
    T = mk3(0., 0., 0.);
    if (!sums) return;
    // GATHER_CHUNK entries at a time: their loads are issued together (entry, realness
    // word, gather record), then summed in entry order
    for (int64_t eb = e0; eb < e1; eb += GATHER_CHUNK) {
        uint32_t v[GATHER_CHUNK];
        bool real[GATHER_CHUNK];
        double4 fg[GATHER_CHUNK], cp[GATHER_CHUNK];
#pragma unroll
        for (int u = 0; u < GATHER_CHUNK; u++) v[u] = eb + u < e1 ? gv.ent[eb + u] : 0u;
#pragma unroll
        for (int u = 0; u < GATHER_CHUNK; u++) {
            const uint32_t c = v[u] & 0x7fffffffu;
            real[u] = eb + u < e1 && ((gv.realMask[c >> 6] >> (c & 63)) & 1ull);  // non-real: no force (ContactLoop.cpp:188)
        }
        if (gv.mixed) {  // FP32 records: force and this node's branch
            const GP float4* RA = (const GP float4*)gv.gfc;
            const GP float4* RB = RA + gv.gcap;
            const GP float* RC = (const GP float*)(RA + 2 * gv.gcap);
            float4 ra[GATHER_CHUNK], rb[GATHER_CHUNK];
            float rc[GATHER_CHUNK];
#pragma unroll
            for (int u = 0; u < GATHER_CHUNK; u++)
                if (real[u]) {
                    const size_t c = v[u] & 0x7fffffffu;
                    ra[u] = RA[c];
                    rb[u] = RB[c];
                    rc[u] = (v[u] >> 31) ? RC[c] : 0.f;
                }
#pragma unroll
            for (int u = 0; u < GATHER_CHUNK; u++) {
                if (!real[u]) continue;
                const bool isB = (v[u] >> 31) != 0;
                const double sign = isB ? -1. : 1.;
                const D3 Fs = mk3((double)ra[u].x * sign, (double)ra[u].y * sign, (double)ra[u].z * sign);
                const D3 xc = isB ? mk3((double)rb[u].z, (double)rb[u].w, (double)rc[u]) : mk3((double)ra[u].w, (double)rb[u].x, (double)rb[u].y);
                F = F + Fs;
                T = T + cross3(xc, Fs);
            }
            continue;
        }
#pragma unroll
        for (int u = 0; u < GATHER_CHUNK; u++)
            if (real[u]) {
                const size_t c = v[u] & 0x7fffffffu;
                fg[u] = gv.gfc[2 * c];
                cp[u] = gv.gfc[2 * c + 1];
            }
#pragma unroll
        for (int u = 0; u < GATHER_CHUNK; u++) {
            if (!real[u]) continue;
            const bool isB = (v[u] >> 31) != 0;
            const double sign = isB ? -1. : 1.;
            D3 Fs = mk3(fg[u].x * sign, fg[u].y * sign, fg[u].z * sign);
            D3 ref = pos;
            if (isB && a.periodic) {
                int32_t q = (int32_t)fg[u].w;
                ref = pos + mv3(ldM3(a.cell->gatherHSize), mk3((double)unpack10(q, 0), (double)unpack10(q, 10), (double)unpack10(q, 20)));
            }
            D3 xc = xyz(cp[u]) - ref;
            F = F + Fs;
            T = T + cross3(xc, Fs);
        }
    }
}
// CSR range of node i's contacts, and whether they are summed (uninodal particles: spheres
// and walls; facet nodes and clump members get none)
__device__ __forceinline__ void gatherRange(const PartView& p, const GatherView& gv, int i, int64_t& e0, int64_t& e1, bool& sums) {
    e0 = e1 = 0;
    sums = false;
    if (i >= p.np) return;
    const int sh = ptShape(p.ptype[i]);
    e0 = gv.off[i];
    e1 = gv.off[i + 1];
    sums = sh == 1 || sh == 2;
}
__device__ __forceinline__ void gatherNode(const StepArgs& a, const PartView& p, const GatherView& gv, int i, uint32_t fl, double m,
                                           const D3& pos, D3& F, D3& T) {
    int64_t e0, e1;
    bool sums;
    gatherRange(p, gv, i, e0, e1, sums);
    gatherForces(a, gv, e0, e1, sums, fl, m, pos, F, T);
}

// node forces into the arrays: the reset values plus (sums != 0) the last ContactLoop's sums
__global__ void __launch_bounds__(WDM_BLOCK) k_gather(const SceneDesc* D, const int32_t* idx, Map m, int sums) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const SceneDesc& S = D[idx ? idx[t] : t];
    const StepArgs a = S.a;
    const NodeView n = S.n;
    const PartView p = S.p;
    const GatherView gv = S.gv;
    int i = (int)(lb * WDM_BLOCK + threadIdx.x);
    if (i >= n.nn) return;
    D3 F, T;
    if (sums) gatherNode(a, p, gv, i, n.flags[i], n.velm[i].w, xyz(n.posr[i]), F, T);
    else gatherForces(a, gv, 0, 0, false, n.flags[i], n.velm[i].w, xyz(n.posr[i]), F, T);
    n.force[i] = make_double4(F.x, F.y, F.z, 0.);
    n.torque[i] = make_double4(T.x, T.y, T.z, 0.);
}

// ---------------- Leapfrog::run, pkg/dem/Leapfrog.cpp:103-255 ----------------
// Per node: accelerations with blocked DOFs, 2nd-order non-viscous damping, GRADV2
// velocity map, spherical rotation, energy terms, and the Verlet displacement check
// of InsertionSortCollider::updateBboxes_doFullRun (:340-350) on the new position.
// Partial energies per block: [nonviscDamp, grav, kinetic].  With GATHER (scenes whose
// last ContactLoop has not been materialised) the node's force and torque are summed
// from the contacts of that ContactLoop here instead of being read from the arrays.
template <int GATHER>
__global__ void __launch_bounds__(WDM_BLOCK) k_leapfrog(const SceneDesc* D, const int32_t* idx, Map m, double* epart, const StepVar* sv) {
    const int64_t blk = xcdChunkBlock(m.bprefix[m.n], XCD_CHUNK_LEAP);
    if (blk < 0) return;
    int64_t lb;
    const int t = mapBlockAt(m, blk, lb);
    const int di = idx ? idx[t] : t;
    const SceneDesc& S = D[di];
    StepArgs a = S.a;
    if (sv) {
        const StepVar v = sv[di];
        if (!leapActive(v, S.words)) return;
        applyStepVar(a, v);
    }
    const NodeView n = S.n;
    const PartView p = S.p;
    const GatherView gv = S.gv;
    int i = (int)(lb * WDM_BLOCK + threadIdx.x);
    double ev[3] = {0., 0., 0.};
    uint32_t bits = 0;
    if (i < n.nn) {
        // every load the node may need is issued here, ahead of the branches on their
        // values, so that they travel together; the gather's chain starts from the CSR range
        const uint32_t fl = n.flags[i];
        const double4 pr = n.posr[i], vm = n.velm[i], aw = n.angI[i], q0 = n.ori[i], lpos = n.lastPos[i];
        const bool chk = a.checkVerlet && n.checked[i];
        int64_t e0 = 0, e1 = 0;
        bool sums = false;
        if (GATHER) gatherRange(p, gv, i, e0, e1, sums);
        D3 pos = xyz(pr);
        if (!(fl & F_CLUMPED)) {  // clumped nodes move with their master (k_clump_members)
            const double dt = a.dt, damping = a.damping;
            D3 vel = xyz(vm), angVel = xyz(aw);
            const double mass = vm.w, inertia = aw.w;
            D3 g = mk3(a.g[0], a.g[1], a.g[2]);
            bool damp = (damping != 0. && !(fl & F_DAMPING_SKIP));
            bool blockedNone = (fl & F_DOF_ALL) == 0, blockedAll = (fl & F_DOF_ALL) == F_DOF_ALL;
            // force and torque are read by the accelerations and by the damping dissipation
            // only; a node with every DOF blocked (a wall) skips its contact sum otherwise
            const bool needFT = !blockedAll || (a.reallyTrackEnergy && damp && !(fl & F_ENERGY_SKIP));
            D3 f = mk3(0., 0., 0.), t = mk3(0., 0., 0.);
            if (needFT) {
                if (fl & F_CLUMP) {  // summed with the members by k_clump_forces
                    const int c = S.cl.nodeClump[i];
                    f = xyz(S.cl.ft[2 * c]);
                    t = xyz(S.cl.ft[2 * c + 1]);
                } else if (GATHER) gatherForces(a, gv, e0, e1, sums, fl, mass, pos, f, t);
                else { f = xyz(n.force[i]); t = xyz(n.torque[i]); }
            }
            if (a.reallyTrackEnergy) {
                if (damp && !(fl & F_ENERGY_SKIP)) {
                    ev[0] = dot3(abs3(vel), abs3(f)) * damping * dt + dot3(abs3(angVel), abs3(t)) * damping * dt;
                    if (ev[0] != 0.) bits |= 1u << E_DAMP;  // ZeroDontCreate
                }
                if (a.hasGravity && !(fl & F_GRAVITY_SKIP)) {
                    double gr = 0;
                    if (blockedNone) gr = -dot3(g, vel) * mass * dt;
                    else {
                        if (!(fl & 1u)) gr -= g.x * vel.x * mass * dt;
                        if (!(fl & 2u)) gr -= g.y * vel.y * mass * dt;
                        if (!(fl & 4u)) gr -= g.z * vel.z * mass * dt;
                    }
                    ev[1] = gr;
                    bits |= 1u << E_GRAV;
                }
            }
            D3 pfv = mk3(0., 0., 0.), pfav = mk3(0., 0., 0.);
            D3 la = mk3(0., 0., 0.), aa = mk3(0., 0., 0.);
            const bool gradv2 = a.homoDeform == 4;
            if (!blockedAll) {
                if (blockedNone) la = f / mass;
                else {
                    if (!(fl & 1u)) la.x += f.x / mass;
                    if (!(fl & 2u)) la.y += f.y / mass;
                    if (!(fl & 4u)) la.z += f.z / mass;
                }
                if (a.periodic) {
                    switch (a.homoDeform) {  // Cell::pprevFluctVel, core/Cell.cpp:36-50
                        case 2: case 3: pfv = vel - mv3(ldM3(a.cell->gradV), pos); break;
                        case 4: pfv = vel - mv3(ldM3(a.cell->gradV), pos - vel * (dt / 2)); break;
                        default: pfv = vel;
                    }
                    pfav = gradv2 ? angVel - mk3(a.cell->spinVec[0], a.cell->spinVec[1], a.cell->spinVec[2]) : angVel;
                } else {
                    pfv = vel;
                    pfav = angVel;
                }
                if (damp) {
                    la.x *= 1 - damping * sgn(f.x * (pfv.x + 0.5 * dt * la.x));
                    la.y *= 1 - damping * sgn(f.y * (pfv.y + 0.5 * dt * la.y));
                    la.z *= 1 - damping * sgn(f.z * (pfv.z + 0.5 * dt * la.z));
                }
                if (gradv2) vel = mv3(ldM3(a.cell->ImLL4hInv), (mv3(ldM3(a.cell->LmL), pos) + mv3(ldM3(a.cell->IpLL4h), vel)) + la * dt);
                else vel = vel + la * dt;
                if (inertia != 0.) {
                    if (blockedNone) aa = t / inertia;
                    else {
                        if (!(fl & 8u)) aa.x += t.x / inertia;
                        if (!(fl & 16u)) aa.y += t.y / inertia;
                        if (!(fl & 32u)) aa.z += t.z / inertia;
                    }
                    if (damp) {
                        aa.x *= 1 - damping * sgn(t.x * (pfav.x + 0.5 * dt * aa.x));
                        aa.y *= 1 - damping * sgn(t.y * (pfav.y + 0.5 * dt * aa.y));
                        aa.z *= 1 - damping * sgn(t.z * (pfav.z + 0.5 * dt * aa.z));
                    }
                    angVel = angVel + aa * dt;
                    if (gradv2) angVel = angVel - mk3(a.cell->deltaSpinVec[0], a.cell->deltaSpinVec[1], a.cell->deltaSpinVec[2]);
                }
            } else {
                if (gradv2) vel = mv3(ldM3(a.cell->ImLL4hInv), mv3(ldM3(a.cell->LmL), pos) + mv3(ldM3(a.cell->IpLL4h), vel));
            }
            if (a.homoDeform >= 0 && !gradv2) {  // applyPeriodicCorrections, :258-272
                if (a.homoDeform == 2 || a.homoDeform == 3) {
                    if (a.homoDeform == 3) vel = vel + mv3(ldM3(a.cell->midGradV), vel - la * (dt / 2.)) * dt;
                    vel = vel + mv3(ldM3(a.cell->dGradV), pos);
                } else if (a.homoDeform == 1) pos = pos + mv3(ldM3(a.cell->nextGradV), pos) * dt;
            }
            if (a.reallyTrackEnergy && !(fl & F_ENERGY_SKIP)) {  // doKineticEnergy, :75-97
                D3 cfv = pfv + la * (.5 * dt), cfav = pfav + aa * (.5 * dt);
                double Etrans = .5 * mass * sqn3(cfv);
                double Erot = 0.5 * dot3(cfav, mk3(inertia * cfav.x, inertia * cfav.y, inertia * cfav.z));
                if (isnan(Erot) || isinf(inertia)) Erot = 0;
                ev[2] = Etrans + Erot;
                bits |= 1u << E_KIN;
            }
            pos = pos + vel * dt;  // leapfrogTranslate, :275-278
            double4 q = q0;  // leapfrogSphericalRotate, :280-288
            if (!iszero3(angVel)) {
                double angle = wsqrt(sqn3(angVel));
                D3 ax = angVel / angle;
                double ha = 0.5 * (angle * dt);
                double s = sin(ha), cw = cos(ha);
                double qx = s * ax.x, qy = s * ax.y, qz = s * ax.z;
                double4 r;
                r.x = ((cw * q.x - qx * q.y) - qy * q.z) - qz * q.w;
                r.y = ((cw * q.y + qx * q.x) + qy * q.w) - qz * q.z;
                r.z = ((cw * q.z + qy * q.x) + qz * q.y) - qx * q.w;
                r.w = ((cw * q.w + qz * q.x) + qx * q.z) - qy * q.y;
                q = r;
            }
            double qn = wsqrt(((q.x * q.x + q.y * q.y) + q.z * q.z) + q.w * q.w);
            q.x = q.x / qn; q.y = q.y / qn; q.z = q.z / qn; q.w = q.w / qn;
            n.ori[i] = q;
            n.posr[i] = make_double4(pos.x, pos.y, pos.z, pr.w);
            n.velm[i] = make_double4(vel.x, vel.y, vel.z, mass);
            n.angI[i] = make_double4(angVel.x, angVel.y, angVel.z, inertia);
        }
        if (chk && !(fl & F_CLUMPED)) {  // members: checked after k_clump_members
            D3 d = xyz(lpos) - pos;
            if (sqn3(d) > a.verletD2) bits |= VERLET_BIT;
        }
    }
    if (!a.reallyTrackEnergy) {  // no energy sums: only the Verlet trigger, one vote per wavefront
        if (__any((bits & VERLET_BIT) != 0) && (threadIdx.x & 63) == 0) atomicMax(S.words, a.stamp);
        return;
    }
    blockSumWrite<3>(ev, bits, epart, S.words, a.stamp, S.words + 1, blk);
}

// ClumpData::forceTorqueFromMembers, pkg/dem/Clump.cpp:265-273, one thread per clump before
// Leapfrog: the master's own force and torque, then each member's added in member order;
// only for masters whose Leapfrog reads them (free DOFs, or damping dissipation tracked)
__global__ void __launch_bounds__(WDM_BLOCK) k_clump_forces(const SceneDesc* D, const int32_t* idx, Map m, const StepVar* sv) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const int di = idx ? idx[t] : t;
    const SceneDesc& S = D[di];
    StepArgs a = S.a;
    if (sv) {
        const StepVar v = sv[di];
        if (!leapActive(v, S.words)) return;
        applyStepVar(a, v);
    }
    const NodeView n = S.n;
    const PartView p = S.p;
    const GatherView gv = S.gv;
    const ClumpView cl = S.cl;
    const int c = (int)(lb * WDM_BLOCK + threadIdx.x);
    if (c >= cl.nClumps) return;
    int M = -1;
    for (int i = cl.off[c]; i < cl.off[c + 1] && M < 0; i++) M = n.master[cl.mem[i]];
    if (M < 0) return;  // clump without members: nothing to add
    const uint32_t fl = n.flags[M];
    const bool damp = a.damping != 0. && !(fl & F_DAMPING_SKIP);
    const bool blockedAll = (fl & F_DOF_ALL) == F_DOF_ALL;
    if (!(!blockedAll || (a.reallyTrackEnergy && damp && !(fl & F_ENERGY_SKIP)))) return;
    const D3 P = xyz(n.posr[M]);
    D3 f, tq;
    if (a.gatherOn) gatherNode(a, p, gv, M, fl, n.velm[M].w, P, f, tq);
    else { f = xyz(n.force[M]); tq = xyz(n.torque[M]); }
    for (int k = cl.off[c]; k < cl.off[c + 1]; k++) {
        const int mn = cl.mem[k];
        const D3 xm = xyz(n.posr[mn]);
        D3 fm, tm;
        if (a.gatherOn) gatherNode(a, p, gv, mn, n.flags[mn], 0., xm, fm, tm);
        else { fm = xyz(n.force[mn]); tm = xyz(n.torque[mn]); }
        f = f + fm;
        tq = tq + (tm + cross3(xm - P, fm));
    }
    cl.ft[2 * c] = make_double4(f.x, f.y, f.z, 0.);
    cl.ft[2 * c + 1] = make_double4(tq.x, tq.y, tq.z, 0.);
}

// quaternions are stored (w, x, y, z) in double4 (x, y, z, w); products and
// quaternion*vector follow Eigen 3's scalar formulas
__device__ __forceinline__ double4 qmul4(const double4& a, const double4& b) {
    double4 r;
    r.x = ((a.x * b.x - a.y * b.y) - a.z * b.z) - a.w * b.w;
    r.y = ((a.x * b.y + a.y * b.x) + a.z * b.w) - a.w * b.z;
    r.z = ((a.x * b.z + a.z * b.x) + a.w * b.y) - a.y * b.w;
    r.w = ((a.x * b.w + a.w * b.x) + a.y * b.z) - a.z * b.y;
    return r;
}
__device__ __forceinline__ D3 qrot4(const double4& q, const D3& v) {
    D3 u = mk3(q.y, q.z, q.w);
    D3 t = cross3(u, v);
    t = t + t;
    return (v + t * q.x) + cross3(u, t);
}
// ClumpData::applyToMembers, pkg/dem/Clump.cpp:276-289, after Leapfrog moved the masters,
// then the collider's Verlet displacement check (InsertionSortCollider.cpp:340-350) on the
// members' new positions
__global__ void __launch_bounds__(WDM_BLOCK) k_clump_members(const SceneDesc* D, const int32_t* idx, Map m, const StepVar* sv) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const int di = idx ? idx[t] : t;
    const SceneDesc& S = D[di];
    StepArgs a = S.a;
    if (sv) {
        const StepVar v = sv[di];
        if (!leapActive(v, S.words)) return;
        applyStepVar(a, v);
    }
    const NodeView n = S.n;
    const ClumpView cl = S.cl;
    const int k = (int)(lb * WDM_BLOCK + threadIdx.x);
    bool trig = false;
    if (k < cl.nMembers) {
        const int mn = cl.mem[k];
        const int M = n.master[mn];
        const double4 P = n.posr[M], Q = n.ori[M], V = n.velm[M], W = n.angI[M];
        const D3 pos = xyz(P) + qrot4(Q, xyz(cl.relPos[k]));
        const double4 o = qmul4(Q, cl.relOri[k]);
        const D3 vel = xyz(V) + cross3(xyz(W), pos - xyz(P));
        const double4 pr = n.posr[mn], vm = n.velm[mn], aw = n.angI[mn];
        n.posr[mn] = make_double4(pos.x, pos.y, pos.z, pr.w);
        n.ori[mn] = o;
        n.velm[mn] = make_double4(vel.x, vel.y, vel.z, vm.w);
        n.angI[mn] = make_double4(W.x, W.y, W.z, aw.w);
        if (a.checkVerlet && n.checked[mn]) trig = sqn3(xyz(n.lastPos[mn]) - pos) > a.verletD2;
    }
    if (__ballot(trig) != 0ull && (threadIdx.x & 63) == 0) atomicMax(S.words, a.stamp);
}

// Per-block partial energies of each scene summed in a fixed order (a strided sequential
// pass per thread, then a tree): the same for a scene alone or in any batch.
// Resettable slots (kinetic from Leapfrog, elast from the laws) are assigned, the others
// accumulated: EnergyTracker::resetResettables at the step prologue followed by the
// engine's adds.
__device__ __forceinline__ int energySlot(int nv, int k) {
    if (nv == 3) return k == 0 ? E_DAMP : (k == 1 ? E_GRAV : E_KIN);
    const int s[5] = {E_ELAST, E_PLAST, E_VISCN, E_VISCT, E_BROKEN};
    return s[k];
}
// one block per (scene, slot); desc[t * stride] (or t when desc is null) is the SceneDesc
// of the launch's t-th scene; spec: skip scenes whose Verlet trigger fired this step; sv:
// the step's row of a window (gated like the kernel whose partials are summed)
__global__ void __launch_bounds__(WDM_BLOCK) k_energy_finalize(const SceneDesc* D, const int32_t* desc, int stride, Map m,
                                                              const double* part, int nv, int spec, const StepVar* sv) {
    __shared__ double sh[WDM_BLOCK];
    const int t = blockIdx.x / nv, k = blockIdx.x % nv;
    const int di = desc ? desc[(size_t)t * stride] : t;
    const SceneDesc& S = D[di];
    const bool leap = nv == 3;
    if (!S.a.trackEnergy) return;
    int really = S.a.reallyTrackEnergy;
    if (sv) {
        const StepVar v = sv[di];
        if (!(leap ? leapActive(v, S.words) : contactActive(v, S.words))) return;
        really = v.reallyTrackEnergy;
    } else if (spec && loadWord(S.words) == S.a.stamp) return;
    const int slot = energySlot(nv, k);
    if (leap && !really) {
        if (slot == E_KIN && threadIdx.x == 0) S.E[E_KIN] = 0.;
        return;
    }
    double s = 0;
    for (int64_t b = m.bprefix[t] + threadIdx.x; b < m.bprefix[t + 1]; b += WDM_BLOCK) s += part[(size_t)b * nv + k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = WDM_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (slot == E_KIN || slot == E_ELAST) S.E[slot] = sh[0];
        else S.E[slot] += sh[0];
    }
}

// ---------------- bounds: Bo1_Sphere_Aabb (pkg/dem/Sphere.cpp:50-79), Bo1_Wall_Aabb
// (pkg/dem/Wall.cpp:24-33), Bo1_Facet_Aabb (pkg/dem/Facet.cpp:177-200), enlarged by
// verletDist and with node positions saved (InsertionSortCollider.cpp:376-414) ----------------
// relaxMode: plain bounds (distFactor 1, no Verlet enlargement) for the overlap relaxation
__global__ void __launch_bounds__(WDM_BLOCK) k_bounds(const SceneDesc* D, const FRScene* FR, Map m, int relaxMode) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const SceneDesc& S = D[FR[t].desc];
    const NodeView n = S.n;
    const PartView p = S.p;
    const double distFactor = relaxMode ? 1. : S.a.distFactor, vd = relaxMode ? 0. : S.a.verletDist;
    int i = (int)(lb * WDM_BLOCK + threadIdx.x);
    if (i >= p.np) return;
    int32_t pt = p.ptype[i];
    int sh = ptShape(pt);
    const double inf = __builtin_inf();
    D3 mn, mx;
    if (sh == 1) {
        double4 pr = n.posr[i];
        double hs = (distFactor > 0 ? distFactor : 1.) * pr.w;
        mn = mk3(pr.x - hs, pr.y - hs, pr.z - hs);
        mx = mk3(pr.x + hs, pr.y + hs, pr.z + hs);
        n.lastPos[i] = make_double4(pr.x, pr.y, pr.z, 0.);
    } else if (sh == 2) {
        double4 pr = n.posr[i];
        int ax = ptAxis(pt);
        mn = mk3(-inf, -inf, -inf);
        mx = mk3(inf, inf, inf);
        double c = ax == 0 ? pr.x : (ax == 1 ? pr.y : pr.z);
        setcomp3(mn, ax, c);
        setcomp3(mx, ax, c);
        n.lastPos[i] = make_double4(pr.x, pr.y, pr.z, 0.);
    } else if (sh == 3) {
        int4 fn = p.fnodes[i];
        double ht = n.posr[i].w;
        double4 x0 = n.posr[fn.x], x1 = n.posr[fn.y], x2 = n.posr[fn.z];
        mn = mk3(x0.x - ht, x0.y - ht, x0.z - ht);
        mx = mk3(x0.x + ht, x0.y + ht, x0.z + ht);
        double4 xs[2] = {x1, x2};
        for (int k = 0; k < 2; k++) {
            mn.x = fmin(mn.x, xs[k].x - ht); mx.x = fmax(mx.x, xs[k].x + ht);
            mn.y = fmin(mn.y, xs[k].y - ht); mx.y = fmax(mx.y, xs[k].y + ht);
            mn.z = fmin(mn.z, xs[k].z - ht); mx.z = fmax(mx.z, xs[k].z + ht);
        }
        n.lastPos[fn.x] = make_double4(x0.x, x0.y, x0.z, 0.);
        n.lastPos[fn.y] = make_double4(x1.x, x1.y, x1.z, 0.);
        n.lastPos[fn.z] = make_double4(x2.x, x2.y, x2.z, 0.);
    } else {
        return;
    }
    if (vd > 0) {
        mx.x += vd; mx.y += vd; mx.z += vd;
        mn.x -= vd; mn.y -= vd; mn.z -= vd;
    }
    p.bmin[i] = make_double4(mn.x, mn.y, mn.z, 0.);
    p.bmax[i] = make_double4(mx.x, mx.y, mx.z, 0.);
}

// extent reduction over binned (finite-bound) particles: min corner, max corner, max extent;
// and per scene a histogram of the largest extents of a sample of the bounds in quarter octaves
// (bin = floor(4 log2 ext) + 128, clamped to [0, EXT_BINS)), from which the host picks the
// grid levels
#define EXT_BINS 256
#define EXT_SAMPLE 8  // the bounds of every 8th block of a scene enter the histogram (few same-address atomics)
__device__ __forceinline__ int extBin(double ext) {
    if (!(ext > 0)) return 0;
    const double b = floor(4. * log2(ext)) + 128.;
    return b < 0. ? 0 : (b > EXT_BINS - 1 ? EXT_BINS - 1 : (int)b);
}
__global__ void __launch_bounds__(WDM_BLOCK) k_extent(const SceneDesc* D, const FRScene* FR, Map m, double* part, uint32_t* hist) {
    __shared__ double sh[8][WDM_BLOCK / 64];
    __shared__ uint32_t hs[EXT_BINS];
    int64_t lb;
    const int st = mapBlock(m, lb);
    const SceneDesc& S = D[FR[st].desc];
    const PartView p = S.p;
    int t = threadIdx.x;
    int i = (int)(lb * WDM_BLOCK + t);
    const double inf = __builtin_inf();
    double v[8] = {inf, inf, inf, -inf, -inf, -inf, 0., 0.};
    for (int k = t; k < EXT_BINS; k += WDM_BLOCK) hs[k] = 0;
    __syncthreads();
    if (i < S.nbinned) {
        int id = S.binned[i];
        double4 a = p.bmin[id], b = p.bmax[id];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = b.x; v[4] = b.y; v[5] = b.z;
        v[6] = fmax(fmax(b.x - a.x, b.y - a.y), b.z - a.z);
        if (ptShape(p.ptype[id]) == 1) v[7] = v[6];
        if (lb % EXT_SAMPLE == 0) atomicAdd(&hs[extBin(v[6])], 1u);
    }
    __syncthreads();
    for (int k = t; k < EXT_BINS; k += WDM_BLOCK)
        if (hs[k]) atomicAdd(&hist[(size_t)st * EXT_BINS + k], hs[k]);
    // minima and maxima over each wavefront, then over the block's wavefronts
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const double y = __shfl_xor(v[k], o, 64);
            v[k] = k < 3 ? fmin(v[k], y) : fmax(v[k], y);
        }
    }
    if ((t & 63) == 0)
        for (int k = 0; k < 8; k++) sh[k][t >> 6] = v[k];
    __syncthreads();
    if (t < 8) {
        double r = sh[t][0];
        for (int w = 1; w < WDM_BLOCK / 64; w++) r = t < 3 ? fmin(r, sh[t][w]) : fmax(r, sh[t][w]);
        part[(size_t)blockIdx.x * 8 + t] = r;
    }
}
// per scene (one block each): min corner, max corner, max extent, max sphere extent over its blocks
__global__ void __launch_bounds__(WDM_BLOCK) k_extent_finalize(Map m, const double* part, double* out) {
    __shared__ double sh[8][WDM_BLOCK];
    const int t = blockIdx.x, tid = threadIdx.x;
    const double inf = __builtin_inf();
    double v[8] = {inf, inf, inf, -inf, -inf, -inf, 0., 0.};
    for (int64_t b = m.bprefix[t] + tid; b < m.bprefix[t + 1]; b += WDM_BLOCK) {
        for (int k = 0; k < 3; k++) v[k] = fmin(v[k], part[(size_t)b * 8 + k]);
        for (int k = 3; k < 8; k++) v[k] = fmax(v[k], part[(size_t)b * 8 + k]);
    }
    for (int k = 0; k < 8; k++) sh[k][tid] = v[k];
    __syncthreads();
    for (int w = WDM_BLOCK / 2; w > 0; w >>= 1) {
        if (tid < w) {
            for (int k = 0; k < 3; k++) sh[k][tid] = fmin(sh[k][tid], sh[k][tid + w]);
            for (int k = 3; k < 8; k++) sh[k][tid] = fmax(sh[k][tid], sh[k][tid + w]);
        }
        __syncthreads();
    }
    if (tid < 8) out[(size_t)t * 8 + tid] = sh[tid][0];
}

__global__ void k_fill_i32(int32_t* a, int64_t n, int32_t v) {
    int64_t i = blockIdx.x * (int64_t)WDM_BLOCK + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------- scans within a wavefront and a workgroup ----------------
template <typename V>
__device__ __forceinline__ V waveInclusiveScan(V x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const V y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}
// exclusive prefix of x over the block's threads in thread order, and the block total;
// ws: one LDS slot per wavefront
__device__ __forceinline__ int64_t blockExclusiveScan(int64_t x, int64_t* ws, int64_t& total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int64_t inc = waveInclusiveScan(x);
    if (lane == 63) ws[wv] = inc;
    __syncthreads();
    int64_t off = 0, tot = 0;
    for (int k = 0; k < nw; k++) {
        const int64_t s = ws[k];
        off += k < wv ? s : 0;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return off + inc - x;
}
// ---------------- exclusive prefix sums: tile sums, then each tile scanned from the sum of
// the tiles before it (read by the tile's workgroup from the tile sums, or, for many tiles,
// from their prefix computed by one workgroup in between) ----------------
#define SCAN_ITEMS 8
#define SCAN_TILE (WDM_BLOCK * SCAN_ITEMS)
#define SCAN_DIRECT_TILES 4096
#define SCAN_TOP 1024
struct ScanId {
    __device__ int64_t operator()(int64_t x) const { return x; }
};
// a work-list flag: bit 0 a listed contact, bit 1 a near one; counted as 1 in the low 32
// bits and, for a near one, 1 in the high 32 bits
struct ScanNearSplit {
    __device__ int64_t operator()(int64_t x) const { return (x & 1) + ((x >> 1) << 32); }
};
// the SCAN_ITEMS items of this thread (zero past n) through op; one 8-byte load for bytes
template <typename TIn, class Op>
__device__ __forceinline__ void scanLoad(const TIn* in, int64_t n, int64_t base, Op op, int64_t v[SCAN_ITEMS]) {
    if (sizeof(TIn) == 1 && base + SCAN_ITEMS <= n && !((uintptr_t)(in + base) & 7)) {
        const uint64_t w = *(const uint64_t*)(in + base);
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; k++) v[k] = op((int64_t)((w >> (8 * k)) & 0xff));
        return;
    }
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) v[k] = base + k < n ? op((int64_t)in[base + k]) : 0;
}
template <typename TIn, class Op>
__global__ void __launch_bounds__(WDM_BLOCK) k_scan_sums(const TIn* in, int64_t n, int64_t* tsum, Op op) {
    __shared__ int64_t ws[WDM_BLOCK / 64];
    int64_t v[SCAN_ITEMS];
    scanLoad(in, n, (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS, op, v);
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) s += v[k];
    int64_t tot;
    (void)blockExclusiveScan(s, ws, tot);
    if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}
// one workgroup: tsum[0..nt) to its exclusive prefix
__global__ void __launch_bounds__(SCAN_TOP) k_scan_top(int64_t* tsum, int64_t nt) {
    __shared__ int64_t ws[SCAN_TOP / 64];
    int64_t carry = 0;
    for (int64_t b0 = 0; b0 < nt; b0 += SCAN_TOP) {
        const int64_t i = b0 + threadIdx.x;
        const int64_t x = i < nt ? tsum[i] : 0;
        int64_t tot;
        const int64_t ex = blockExclusiveScan(x, ws, tot);
        if (i < nt) tsum[i] = carry + ex;
        carry += tot;
    }
}
// prefixed: tsum holds the exclusive prefix of the tile sums; otherwise the tile sums, and
// the workgroup adds those before its tile
template <typename TIn, typename TOut, class Op>
__global__ void __launch_bounds__(WDM_BLOCK) k_scan_tiles(const TIn* in, int64_t n, const int64_t* tsum, int prefixed, TOut* out, Op op) {
    __shared__ int64_t ws[WDM_BLOCK / 64];
    const int64_t tile = blockIdx.x, base = tile * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int64_t off;
    if (prefixed) off = tsum[tile];
    else {
        int64_t s = 0;
        for (int64_t j = threadIdx.x; j < tile; j += WDM_BLOCK) s += tsum[j];
        int64_t tot;
        (void)blockExclusiveScan(s, ws, tot);
        off = tot;
    }
    int64_t v[SCAN_ITEMS];
    scanLoad(in, n, base, op, v);
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) s += v[k];
    int64_t tot;
    int64_t run = off + blockExclusiveScan(s, ws, tot);
    if (tile == (int64_t)gridDim.x - 1 && threadIdx.x == 0) out[n] = (TOut)(off + tot);
    // through LDS, so that each store instruction writes consecutive items (rows padded to
    // SCAN_ITEMS + 1 against bank conflicts)
    __shared__ TOut lo[WDM_BLOCK * (SCAN_ITEMS + 1)];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        lo[threadIdx.x * (SCAN_ITEMS + 1) + k] = (TOut)run;
        run += v[k];
    }
    __syncthreads();
    const int64_t t0 = tile * SCAN_TILE;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        const int e = k * WDM_BLOCK + threadIdx.x;
        if (t0 + e < n) out[t0 + e] = lo[(e / SCAN_ITEMS) * (SCAN_ITEMS + 1) + e % SCAN_ITEMS];
    }
}

// ---------------- bucket sort: 64-bit items grouped by a bucket index, each bucket in
// ascending item order (the high 32 bits of the items of one bucket are distinct) ----------------
// The producer counts its items per bucket with atomics (the count returned is the item's
// slot) and scatters them as records (item, b) to tmp[off[b] + slot] (one 16-byte store), off
// the exclusive prefix of the counts.  k_bucket_rank then places each item of a bucket of at most BKT_SMALL items
// by counting the smaller items of its bucket; longer buckets are listed and sorted by
// k_bucket_long, one workgroup each (bitonic sort in LDS, merged through global memory
// when a bucket exceeds BKT_LDS items).  The writer w(b, position, item) stores the result.
#define BKT_SMALL 64
#define BKT_LDS 8192
#define BKT_THREADS 1024
struct __align__(16) BItem {
    uint64_t x;  // the item
    uint32_t b;  // its bucket
    uint32_t pad;
};
__device__ __forceinline__ void putItem(BItem* tmp, int64_t pos, uint64_t x, uint32_t b) {
    *(ulonglong2*)&tmp[pos] = make_ulonglong2(x, (unsigned long long)b);
}
// n: items, or (total non-null) at most n items, *total of them
template <typename OffT, class W>
__global__ void __launch_bounds__(WDM_BLOCK) k_bucket_rank(const BItem* tmp, const OffT* off, int64_t n, const OffT* total,
                                                          uint32_t* longList, uint32_t* longCnt, W w) {
    const int64_t q = blockIdx.x * (int64_t)WDM_BLOCK + threadIdx.x;
    if (q >= (total ? (int64_t)*total : n)) return;
    const BItem it = tmp[q];
    const uint32_t b = it.b;
    const int64_t s = (int64_t)off[b], e = (int64_t)off[b + 1];
    if (e - s > BKT_SMALL) {
        if (q == s) longList[atomicAdd(longCnt, 1u)] = b;
        return;
    }
    int r = 0;
    for (int64_t k = s; k < e; k++) r += tmp[k].x < it.x ? 1 : 0;
    w(b, s + r, it.x);
}
__device__ __forceinline__ void bitonicLds(uint64_t* sh, int P) {
    for (int k = 2; k <= P; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t a = sh[i], b = sh[ixj];
                    if ((a > b) == ((i & k) == 0)) {
                        sh[i] = b;
                        sh[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
}
// aux: scratch of twice as many items as tmp holds
template <typename OffT, class W>
__global__ void __launch_bounds__(BKT_THREADS) k_bucket_long(const BItem* tmp, uint64_t* aux, int64_t n, const OffT* off,
                                                            const uint32_t* longList, const uint32_t* longCnt, W w) {
    __shared__ uint64_t sh[BKT_LDS];
    const uint32_t nl = *longCnt;
    for (uint32_t li = blockIdx.x; li < nl; li += gridDim.x) {
        const uint32_t b = longList[li];
        const int64_t s = (int64_t)off[b], len = (int64_t)off[b + 1] - s;
        for (int64_t c0 = 0; c0 < len; c0 += BKT_LDS) {
            const int cl = (int)(len - c0 < BKT_LDS ? len - c0 : BKT_LDS);
            int P = 1;
            while (P < cl) P <<= 1;
            for (int k = threadIdx.x; k < P; k += BKT_THREADS) sh[k] = k < cl ? tmp[s + c0 + k].x : ~0ull;
            __syncthreads();
            bitonicLds(sh, P);
            for (int k = threadIdx.x; k < cl; k += BKT_THREADS) aux[s + c0 + k] = sh[k];
            __syncthreads();
        }
        // sorted runs of width wd merged pairwise: an item's place is its index in its run
        // plus the number of smaller items in the other run
        uint64_t* src = aux + s;
        uint64_t* dst = aux + n + s;
        for (int64_t wd = BKT_LDS; wd < len; wd <<= 1) {
            for (int64_t k = threadIdx.x; k < len; k += BKT_THREADS) {
                const int64_t r0 = (k / wd) * wd, p0 = ((k / wd) ^ 1) * wd;
                const uint64_t x = src[k];
                int64_t pos = k;
                if (p0 < len) {
                    int64_t lo = p0, hi = p0 + wd < len ? p0 + wd : len;
                    while (lo < hi) {
                        const int64_t mid = (lo + hi) >> 1;
                        if (src[mid] < x) lo = mid + 1; else hi = mid;
                    }
                    pos = (r0 < p0 ? r0 : p0) + (k - r0) + (lo - p0);
                }
                dst[pos] = x;
            }
            __syncthreads();
            uint64_t* t2 = src;
            src = dst;
            dst = t2;
        }
        for (int64_t k = threadIdx.x; k < len; k += BKT_THREADS) w(b, s + k, src[k]);
        __syncthreads();
    }
}
// scene of a batch-wide row (rowBase) or cell key (cellBase): the last scene starting at or before it
__device__ __forceinline__ int sceneOfRow(const FRScene* FR, int nT, int64_t r) {
    int lo = 0, hi = nT - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (FR[mid].rowBase <= r) lo = mid; else hi = mid - 1;
    }
    return lo;
}
__device__ __forceinline__ int sceneOfCell(const FRScene* FR, int nT, int64_t k) {
    int lo = 0, hi = nT - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (FR[mid].cellBase <= k) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ---------------- collider grid ----------------
// cell of coordinate x along axis d at level l (clamped to the grid)
__device__ __forceinline__ int cellCoordL(const FRScene& F, int l, int d, double x) {
    const int n = F.ln[l][d];
    int c;
    if (F.g.periodic) {
        double u = x / F.g.L[d];
        double f = u - floor(u);
        c = (int)(f * n);
    } else {
        c = (int)floor((x - F.g.origin[d]) / F.lw[l][d]);
    }
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}
// level of a bound of largest extent ext (nLvl: large)
__device__ __forceinline__ int boundLevel(const FRScene& F, double ext) {
    int l = 0;
    while (l < F.nLvl && ext > F.lthr[l]) l++;
    return l;
}
// FP32 box of a bound: centre (wrapped into the periodic cell, or relative to the grid
// origin) and the largest half extent.  The pair search scans a neighbour cell's entries
// contiguously through these 16 bytes and reads a bound's exact FP64 corners only when the
// FP32 test cannot rule the pair out.
__device__ __forceinline__ float4 cellBox(const FRScene& F, const double4& a, const double4& b) {
    const double c3[3] = {.5 * (a.x + b.x), .5 * (a.y + b.y), .5 * (a.z + b.z)};
    double r[3];
    for (int d = 0; d < 3; d++) {
        if (F.g.periodic) r[d] = c3[d] - F.g.L[d] * floor(c3[d] * F.g.invL[d]);
        else r[d] = c3[d] - F.g.origin[d];
    }
    const double h = fmax(fmax(b.x - a.x, b.y - a.y), b.z - a.z) * .5;
    return make_float4((float)r[0], (float)r[1], (float)r[2], (float)h);
}
// Bound entries by cell: key = scene's first cell + cell of the bound's minimum corner at
// its level (or the scene's large list); the count per key gives each entry its slot.  A
// large bound in a pass that does not list large bounds raises *planFlag (its grid was laid
// out for smaller bounds).
__global__ void __launch_bounds__(WDM_BLOCK) k_cell_count(const SceneDesc* D, const FRScene* FR, Map m, uint32_t* ekey, uint32_t* eslot,
                                                         uint32_t* cnt, uint32_t* planFlag) {
    int64_t lb;
    const FRScene& F = FR[mapBlock(m, lb)];
    const SceneDesc& S = D[F.desc];
    const int i = (int)(lb * WDM_BLOCK + threadIdx.x);
    if (i >= S.nbinned) return;
    const int id = S.binned[i];
    const double4 a = S.p.bmin[id], b = S.p.bmax[id];
    const int l = boundLevel(F, fmax(fmax(b.x - a.x, b.y - a.y), b.z - a.z));
    int64_t cell = F.nCells;
    if (l < F.nLvl) {
        const int cx = cellCoordL(F, l, 0, a.x), cy = cellCoordL(F, l, 1, a.y), cz = cellCoordL(F, l, 2, a.z);
        cell = F.lbase[l] + ((int64_t)cz * F.ln[l][1] + cy) * F.ln[l][0] + cx;
    } else if (planFlag && !F.largeListed) {
        atomicOr(planFlag, 1u);
    }
    const uint32_t key = (uint32_t)(F.cellBase + cell);
    ekey[F.binStart + i] = key;
    eslot[F.binStart + i] = atomicAdd(&cnt[key], 1u);
}
// item (id << 32) of each entry at its cell's offset plus its slot
__global__ void __launch_bounds__(WDM_BLOCK) k_cell_scatter(const SceneDesc* D, const FRScene* FR, Map m, const uint32_t* ekey,
                                                           const uint32_t* eslot, const int32_t* cbeg, BItem* tmp) {
    int64_t lb;
    const FRScene& F = FR[mapBlock(m, lb)];
    const SceneDesc& S = D[F.desc];
    const int i = (int)(lb * WDM_BLOCK + threadIdx.x);
    if (i >= S.nbinned) return;
    const uint32_t key = ekey[F.binStart + i];
    const int64_t pos = (int64_t)cbeg[key] + eslot[F.binStart + i];
    putItem(tmp, pos, (uint64_t)(uint32_t)S.binned[i] << 32, key);
}
// the sorted entries, by (scene, cell, id): particle id, cell key, FP32 box, and a copy of
// the bound's corners, shape and mask (the exact pair test reads a partner by entry, without
// going through its id)
struct CellOut {
    const SceneDesc* D;
    const FRScene* FR;
    int nT;
    int32_t* cvals;
    uint32_t* ckeys;
    float4* boxes;
    double4 *eb0, *eb1;
    int2* einfo;
    __device__ void operator()(uint32_t b, int64_t pos, uint64_t x) const {
        const FRScene& F = FR[sceneOfCell(FR, nT, b)];
        const PartView& p = D[F.desc].p;
        const int id = (int)(x >> 32);
        const double4 a0 = p.bmin[id], a1 = p.bmax[id];
        cvals[pos] = id;
        ckeys[pos] = b;
        boxes[pos] = cellBox(F, a0, a1);
        if (eb0) {
            eb0[pos] = a0;
            eb1[pos] = a1;
            einfo[pos] = make_int2(ptShape(p.ptype[id]), (int)p.mask[id]);
        }
    }
};

// InsertionSortCollider::cellWrap / cellWrapRel (InsertionSortCollider.cpp:650-660)
__device__ __forceinline__ double cellWrapP(double x, double x0, double x1, int& period) {
    double xNorm = (x - x0) / (x1 - x0);
    period = (int)floor(xNorm);
    return x0 + (xNorm - period) * (x1 - x0);
}
__device__ __forceinline__ double cellWrapRel(double x, double x0, double x1) {
    double xNorm = (x - x0) / (x1 - x0);
    return (xNorm - floor(xNorm)) * (x1 - x0);
}
// spatialOverlap (:52-58) and spatialOverlapPeri (:903-957) on the stored bounds
__device__ __forceinline__ bool overlapAperiodic(const double4& a0, const double4& a1, const double4& b0, const double4& b1) {
    return (a0.x <= b1.x) && (a1.x >= b0.x) && (a0.y <= b1.y) && (a1.y >= b0.y) && (a0.z <= b1.z) && (a1.z >= b0.z);
}
__device__ __forceinline__ bool overlapPeri(const double4& a0, const double4& a1, const double4& b0, const double4& b1,
                                            const double* L, int periods[3], uint32_t* err) {
    const double am[3] = {a0.x, a0.y, a0.z}, aM[3] = {a1.x, a1.y, a1.z};
    const double bm[3] = {b0.x, b0.y, b0.z}, bM[3] = {b1.x, b1.y, b1.z};
    for (int ax = 0; ax < 3; ax++) {
        double dim = L[ax];
        double m1 = am[ax], m2 = bm[ax];
        if (isinf(m1) || isinf(m2)) { periods[ax] = 0; continue; }
        double wMn = (cellWrapRel(m1, m2, m2 + dim) < cellWrapRel(m2, m1, m1 + dim)) ? m2 : m1;
        int pmn1, pmx1, pmn2, pmx2;
        double mn1 = cellWrapP(am[ax], wMn, wMn + dim, pmn1), mx1 = cellWrapP(aM[ax], wMn, wMn + dim, pmx1);
        double mn2 = cellWrapP(bm[ax], wMn, wMn + dim, pmn2), mx2 = cellWrapP(bM[ax], wMn, wMn + dim, pmx2);
        if ((pmn1 != pmx1) || (pmn2 != pmx2)) { atomicOr(err, ERR_OVERSIZE); return false; }
        periods[ax] = pmn1 - pmn2;
        if (!(mn1 <= mx2 && mx1 >= mn2)) return false;
    }
    return true;
}

// functor order: Wall/Facet first, otherwise the lower id
WDM_HD void canonical(int i, int j, int si, int sj, int& A, int& B, int& type) {
    if (si == 1 && sj == 1) { A = i < j ? i : j; B = i < j ? j : i; type = CI_SS; return; }
    if ((si == 2 || si == 3) && sj == 1) { A = i; B = j; type = si == 2 ? CI_WS : CI_FS; return; }
    if ((sj == 2 || sj == 3) && si == 1) { A = j; B = i; type = sj == 2 ? CI_WS : CI_FS; return; }
    A = i < j ? i : j; B = i < j ? j : i; type = CI_NONE;
}

// Collider::mayCollide, pkg/dem/Collision.cpp:19-41; master: clump master per node (null
// when the scene has no clumps)
// ma, mb: the masks of a and b
__device__ __forceinline__ bool mayCollideM(const PartView& p, const GP int32_t* master, int a, int b, int sa, int sb, uint32_t ma, uint32_t mb,
                                            uint32_t loneMask) {
    if (a == b || sa == 0 || sb == 0) return false;
    if (!(ma & mb)) return false;
    if ((ma & mb & loneMask) != 0) return false;
    if (sa == 3 || sb == 3 || master) {
        int na[3], nb[3], ka, kb;
        if (sa == 3) { int4 f = p.fnodes[a]; na[0] = f.x; na[1] = f.y; na[2] = f.z; ka = 3; } else { na[0] = a; ka = 1; }
        if (sb == 3) { int4 f = p.fnodes[b]; nb[0] = f.x; nb[1] = f.y; nb[2] = f.z; kb = 3; } else { nb[0] = b; kb = 1; }
        for (int x = 0; x < ka; x++)
            for (int y = 0; y < kb; y++) {
                if (na[x] == nb[y]) return false;  // shared node
                if (master) {                      // nodes of the same clump
                    int mx = master[na[x]];
                    if (mx >= 0 && mx == master[nb[y]]) return false;
                }
            }
    }
    return true;
}
__device__ __forceinline__ bool mayCollideD(const PartView& p, const GP int32_t* master, int a, int b, int sa, int sb, uint32_t loneMask) {
    return mayCollideM(p, master, a, b, sa, sb, p.mask[a], p.mask[b], loneMask);
}

// FP64 prefilter of the exact test: minimum-image distance of the bound centres
__device__ __forceinline__ bool periPrefilter(const double4& a0, const double4& a1, const double4& b0, const double4& b1, const GridArgs& g) {
    // centres and half extents; reject when the minimum-image gap exceeds the half
    // extents by more than 1e-9 L on some axis (rounding here is ~1e-16 L)
    const double ca[3] = {.5 * (a0.x + a1.x), .5 * (a0.y + a1.y), .5 * (a0.z + a1.z)};
    const double cb[3] = {.5 * (b0.x + b1.x), .5 * (b0.y + b1.y), .5 * (b0.z + b1.z)};
    const double ha[3] = {.5 * (a1.x - a0.x), .5 * (a1.y - a0.y), .5 * (a1.z - a0.z)};
    const double hb[3] = {.5 * (b1.x - b0.x), .5 * (b1.y - b0.y), .5 * (b1.z - b0.z)};
    for (int d = 0; d < 3; d++) {
        double L = g.L[d];
        double dx = cb[d] - ca[d];
        dx = dx - L * rint(dx * g.invL[d]);
        if (fabs(dx) > ha[d] + hb[d] + 1e-9 * L) return false;
    }
    return true;
}

// cell index along axis d at level l without wrapping or clamping (large bounds' cell ranges)
__device__ __forceinline__ int cellCoordRawL(const FRScene& F, int l, int d, double x) {
    if (F.g.periodic) return (int)floor(x / F.g.L[d] * F.ln[l][d]);
    return (int)floor((x - F.g.origin[d]) / F.lw[l][d]);
}
__device__ __forceinline__ int wrapCell(int x, int n) { return ((x % n) + n) % n; }

// cells of level l listing a large bound: those its bound covers, two more at the low side
// and one at the high side (a bound of the level starts at most one cell before it; one
// cell of rounding margin each way), each cell once in a periodic grid
__device__ __forceinline__ void largeRangeL(const FRScene& F, int l, const double4& a0, const double4& a1, int c0[3], int nc[3]) {
    const double lo3[3] = {a0.x, a0.y, a0.z}, hi3[3] = {a1.x, a1.y, a1.z};
    for (int d = 0; d < 3; d++) {
        const int n = F.ln[l][d];
        int lo = cellCoordRawL(F, l, d, lo3[d]) - 2, hi = cellCoordRawL(F, l, d, hi3[d]) + 1;
        if (!F.g.periodic) { lo = lo < 0 ? 0 : lo; hi = hi > n - 1 ? n - 1 : hi; }
        c0[d] = lo;
        nc[d] = hi >= lo ? hi - lo + 1 : 0;
        if (F.g.periodic && nc[d] > n) { nc[d] = n; c0[d] = 0; }
    }
}
// Large bounds are listed in the cells of every level: the number of cells listing each sorted
// entry (0 for the bounds in cells)
__global__ void __launch_bounds__(WDM_BLOCK) k_large_cells(const SceneDesc* D, const FRScene* FR, Map m, const uint32_t* ckeys,
                                                          const int32_t* cvals, uint32_t* counts) {
    int64_t lb;
    const FRScene& F = FR[mapBlock(m, lb)];
    const SceneDesc& S = D[F.desc];
    const int li = (int)(lb * WDM_BLOCK + threadIdx.x);
    if (li >= S.nbinned) return;
    const int64_t i = F.binStart + li;
    if ((int)(ckeys[i] - F.cellBase) != F.nCells) {
        counts[i] = 0;
        return;
    }
    const int id = cvals[i];
    const double4 a0 = S.p.bmin[id], a1 = S.p.bmax[id];
    uint32_t tot = 0;
    for (int l = 0; l < F.nLvl; l++) {
        int c0[3], nc[3];
        largeRangeL(F, l, a0, a1, c0, nc);
        tot += (uint32_t)(nc[0] * nc[1] * nc[2]);
    }
    counts[i] = tot;
}
// The listings themselves, one thread each (a facet covers hundreds of cells over the levels):
// listing o belongs to the entry i with offs[i] <= o < offs[i + 1] (offs: the exclusive prefix
// sum of the counts over the nb sorted entries of the nT scenes); its level and cell follow
// from o - offs[i] in the order level, z, y, x.  Writes the cell key, the particle id and the
// slot of the count per cell.
__global__ void __launch_bounds__(WDM_BLOCK) k_large_list(const SceneDesc* D, const FRScene* FR, int nT, const int32_t* cvals, int64_t nb,
                                                         const int64_t* offs, int64_t nl, uint32_t* lkeys, int32_t* lids, uint32_t* lslot,
                                                         uint32_t* lcnt) {
    const int64_t o = blockIdx.x * (int64_t)WDM_BLOCK + threadIdx.x;
    if (o >= nl) return;
    int64_t lo = 0, hi = nb;  // offs[lo] <= o < offs[hi]
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (offs[mid] <= o) lo = mid;
        else hi = mid;
    }
    const int64_t i = lo;
    int tl = 0, th = nT;  // FR[tl].binStart <= i < FR[th].binStart
    while (th - tl > 1) {
        const int mid = (tl + th) >> 1;
        if (FR[mid].binStart <= i) tl = mid;
        else th = mid;
    }
    const FRScene& F = FR[tl];
    const SceneDesc& S = D[F.desc];
    const int id = cvals[i];
    const double4 a0 = S.p.bmin[id], a1 = S.p.bmax[id];
    int64_t r = o - offs[i];
    for (int l = 0; l < F.nLvl; l++) {
        int c0[3], nc[3];
        largeRangeL(F, l, a0, a1, c0, nc);
        const int64_t cnt = (int64_t)nc[0] * nc[1] * nc[2];
        if (r >= cnt) {
            r -= cnt;
            continue;
        }
        const int nx = F.ln[l][0], ny = F.ln[l][1], nz = F.ln[l][2];
        const int xx = (int)(r % nc[0]), yy = (int)((r / nc[0]) % nc[1]), zz = (int)(r / ((int64_t)nc[0] * nc[1]));
        const int z = F.g.periodic ? wrapCell(c0[2] + zz, nz) : c0[2] + zz;
        const int y = F.g.periodic ? wrapCell(c0[1] + yy, ny) : c0[1] + yy;
        const int x = F.g.periodic ? wrapCell(c0[0] + xx, nx) : c0[0] + xx;
        const uint32_t key = (uint32_t)(F.cellBase + F.lbase[l] + ((int64_t)z * ny + y) * nx + x);
        lkeys[o] = key;
        lids[o] = id;
        lslot[o] = atomicAdd(&lcnt[key], 1u);
        return;
    }
}
// each listing at its cell's offset plus its slot (the order within a cell does not matter:
// the candidates are sorted afterwards)
__global__ void __launch_bounds__(WDM_BLOCK) k_large_scatter(const uint32_t* lkeys, const int32_t* lids, const uint32_t* lslot, int64_t n,
                                                            const int32_t* lbeg, int32_t* lvals) {
    const int64_t o = blockIdx.x * (int64_t)WDM_BLOCK + threadIdx.x;
    if (o >= n) return;
    lvals[lbeg[lkeys[o]] + lslot[o]] = lids[o];
}

// one candidate of bound `id` (corners a0, a1, shape sa) with particle j: canonical order,
// the collider's exact overlap (with periods) and mayCollide; a binned j must pass the
// FP64 prefilter first (it has finite bounds)
// the same with j's bounds (b0, b1), shape sb and the masks ma, mb given
__device__ __forceinline__ bool pairTestB(const SceneDesc& S, const GridArgs& g, int id, const double4& a0, const double4& a1, int sa,
                                          uint32_t ma, int j, const double4& b0, const double4& b1, int sb, uint32_t mb, bool binned,
                                          uint32_t* err, int per[3]) {
    if (binned) {
        if (g.periodic) { if (!periPrefilter(a0, a1, b0, b1, g)) return false; }
        else if (!overlapAperiodic(a0, a1, b0, b1)) return false;
    }
    int A, B, ty;
    canonical(id, j, sa, sb, A, B, ty);
    const bool aIsId = (A == id);
    const double4 &A0 = aIsId ? a0 : b0, &A1 = aIsId ? a1 : b1, &B0 = aIsId ? b0 : a0, &B1 = aIsId ? b1 : a1;
    per[0] = per[1] = per[2] = 0;
    const bool ov = g.periodic ? overlapPeri(A0, A1, B0, B1, g.L, per, err) : (binned || overlapAperiodic(A0, A1, B0, B1));
    return ov && mayCollideM(S.p, S.n.master, A, B, aIsId ? sa : sb, aIsId ? sb : sa, aIsId ? ma : mb, aIsId ? mb : ma, S.a.loneMask);
}
__device__ __forceinline__ bool pairTest(const SceneDesc& S, const GridArgs& g, int id, const double4& a0, const double4& a1, int sa, int j,
                                         bool binned, uint32_t* err, int per[3]) {
    const PartView& p = S.p;
    return pairTestB(S, g, id, a0, a1, sa, p.mask[id], j, p.bmin[j], p.bmax[j], ptShape(p.ptype[j]), p.mask[j], binned, err, per);
}

// Candidate pairs, one thread per bound entry (in the (scene, level, cell, id) order of the
// sorted entries).  An entry visits the entries of its own cell after it and of the 13
// "forward" neighbour cells of its level, so every unordered pair of entries of one level
// is visited once (two bounds of a level overlap only if their minimum corners lie in the
// same or adjacent cells: their extents are below the cell width by 0.1 %); and, for every
// finer level, the cells in which the minimum corner of an overlapping bound of that level
// can lie, so every pair of entries of two levels is visited once, by the coarser one.
// Each entry also tests every unbinned particle (walls) and the large bounds listed in its
// cell (k_large_cells); an entry of the large list pairs with the later ones when their
// masks allow it.  Conservative tests reject most candidates cheaply (FP32 boxes, then an
// FP64 minimum-image test on bound centres); the decision itself is the collider's own
// closed-interval 3-axis overlap of the stored bounds plus mayCollide, so the set equals the
// reference's and the grid only prunes.
// Two passes with the same blocks (wavefront w of both works on the same 64 entries):
// k_pair_scan runs its lanes' partner lists flattened into one loop, one partner per lane
// per iteration (a lane whose current cell is used up moves to its next non-empty one), and
// writes the (entry, partner entry) pairs that pass the FP32 box test to the wavefront's
// region; it carries no FP64 state, so many wavefronts stay resident to cover the loads.
// k_pair_exact then runs the exact test over each region with every lane busy, and tests
// its entries against the walls and the listed large bounds.
// Output of k_pair_exact: each accepted pair goes to the wavefront's own region (item max
// id << 32 | packed periods, row, slot) and counts in rowCnt[rowBase + min id], whose
// returned value is its slot in that row.  A wavefront w owns entries [w cap, (w+1) cap)
// of a region, with their number in cnt[w]; a count above cap (the pairs past it are
// neither written nor counted in their rows, so the rows hold at most nw cap pairs)
// raises *wmax, and the host runs the passes again with larger regions.
__device__ __forceinline__ bool stencilCellL(const FRScene& F, int l, int cx, int cy, int cz, int q, int& cell) {
    // half-stencil offset q: (0,0,0), (1,0,0), (-1..1,1,0), (-1..1,-1..1,1)
    int dx, dy, dz;
    if (q == 0) { dx = 0; dy = 0; dz = 0; }
    else if (q == 1) { dx = 1; dy = 0; dz = 0; }
    else if (q < 5) { dx = q - 3; dy = 1; dz = 0; }
    else { dx = (q - 5) % 3 - 1; dy = (q - 5) / 3 - 1; dz = 1; }
    const int nx = F.ln[l][0], ny = F.ln[l][1], nz = F.ln[l][2];
    const bool per = F.g.periodic != 0;
    // periodic axes with fewer than 3 cells hold a single cell (no neighbours); aperiodic axes always scan +-1
    if ((dx && per && nx < 3) || (dy && per && ny < 3) || (dz && per && nz < 3)) return false;
    int z = cz + dz, y = cy + dy, x = cx + dx;
    if (per) {  // one step across the periodic boundary at most (|d| <= 1, n >= 3)
        x = x < 0 ? x + nx : (x >= nx ? x - nx : x);
        y = y < 0 ? y + ny : (y >= ny ? y - ny : y);
        z = z < 0 ? z + nz : (z >= nz ? z - nz : z);
    } else if (z < 0 || z >= nz || y < 0 || y >= ny || x < 0 || x >= nx) return false;
    cell = F.lbase[l] + (z * ny + y) * nx + x;
    return true;
}
// Pair scan in XCD chunks of PAIR_CHUNK blocks (xcdChunkBlock): the entries of a chunk are a
// contiguous range of cells, whose neighbours mostly share that XCD's L2, while a stretch of
// costly entries (the coarser levels come last) is spread over all XCDs.  The launch is
// padded to a multiple of 8 PAIR_CHUNK blocks.
#define PAIR_CHUNK 16
// this wavefront's index (regions, counts), its scene and its block within the scene; false
// for padding.  The logical blocks run backwards, so that the costly entries of the coarser
// levels (last in each scene) start first and do not trail the launch.
__device__ __forceinline__ bool pairWave(const Map& m, int64_t& gw, int& t, int64_t& lb) {
    const int64_t lbk = (int64_t)gridDim.x - 1 - xcdChunkBlock(gridDim.x, PAIR_CHUNK);
    if (lbk >= m.bprefix[m.n]) return false;
    t = mapBlockAt(m, lbk, lb);
    gw = lbk * (WDM_BLOCK / 64) + (threadIdx.x >> 6);
    return true;
}
__global__ void __launch_bounds__(WDM_BLOCK, 6) k_pair_scan(const SceneDesc* D, const FRScene* FR, Map m, const uint32_t* ckeys,
                                                        const int32_t* cbeg, const float4* boxes, int32_t* cnt, int32_t* wmax, int32_t cap,
                                                        int2* out) {
    int64_t gw, lb;
    int tt;
    if (!pairWave(m, gw, tt, lb)) return;
    const FRScene& F = FR[tt];
    const SceneDesc& S = D[F.desc];
    const GridArgs& g = F.g;
    GP int2* ro = gpc(out) + gw * (int64_t)cap;
    int cntW = 0;  // pairs found (the same in every lane)
    const int li = (int)(lb * WDM_BLOCK + threadIdx.x);
    const bool valid = li < S.nbinned;
    const int64_t e = F.binStart + (valid ? li : 0);  // this lane's entry
    const int64_t cb = F.cellBase;
    const float4 own = valid ? boxes[e] : make_float4(0.f, 0.f, 0.f, -1e30f);
    const int key = valid ? (int)(ckeys[e] - cb) : 0;
    int lvl = 0;  // level of the entry (nLvl: the large list)
    while (lvl < F.nLvl && key >= F.lbase[lvl + 1]) lvl++;
    const bool large = lvl >= F.nLvl;
    int cx = 0, cy = 0, cz = 0;
    if (!large) {
        const int c = key - F.lbase[lvl], nx = F.ln[lvl][0], ny = F.ln[lvl][1];
        cx = c % nx;
        cy = (c / nx) % ny;
        cz = c / (nx * ny);
    }
    // periodic wrap of the FP32 box test (zero on aperiodic grids: the wrap term vanishes)
    const float wLx = g.periodic ? g.Lf[0] : 0.f, wLy = g.periodic ? g.Lf[1] : 0.f, wLz = g.periodic ? g.Lf[2] : 0.f;
    const float wIx = g.periodic ? g.invLf[0] : 0.f, wIy = g.periodic ? g.invLf[1] : 0.f, wIz = g.periodic ? g.invLf[2] : 0.f;
    const float margin = g.pfMargin;
    // partner ranges: phases 0-13 the half stencil of the entry's level (phase 0: its own
    // cell, entries after it), then one phase per finer level scanning a box of cells
    const int nPh = large ? (F.largePairs ? 1 : 0) : 14 + lvl;
    int phase = 1;                               // phase 0 (the own cell) starts below
    int rl = 0;                                  // box scan: level,
    int r0x = 0, r0y = 0, r0z = 0, rnx = 1, rny = 1, rnz = 0;  // origin (cells, wrapped) and sizes,
    int ix = 0, iy = 0, iz = 0;                  // next cell within the box
    // the next cell of this lane's list (batch-wide index), -1 past its end
    auto nextCell = [&]() -> int64_t {
        for (;;) {
            if (iz < rnz) {
                const int nx = F.ln[rl][0], ny = F.ln[rl][1], nz = F.ln[rl][2];
                int x = r0x + ix, y = r0y + iy, z = r0z + iz;  // below 2 n: the origin is wrapped, the sizes at most n
                x -= x >= nx ? nx : 0;
                y -= y >= ny ? ny : 0;
                z -= z >= nz ? nz : 0;
                if (++ix == rnx) {
                    ix = 0;
                    if (++iy == rny) { iy = 0; ++iz; }
                }
                return cb + F.lbase[rl] + ((int64_t)z * ny + y) * nx + x;
            }
            if (phase >= nPh) return -1;
            const int ph = phase++;
            if (ph < 14) {
                int c;
                if (stencilCellL(F, lvl, cx, cy, cz, ph, c)) return cb + c;
            } else {
                // finer level rl: a bound there overlapping this one has its minimum corner in
                // [own min - lthr[rl], own max] on every axis (FP32 box, widened by the FP32 slack)
                rl = ph - 14;
                const float oc[3] = {own.x, own.y, own.z};
                int o3[3], n3[3];
                for (int d = 0; d < 3; d++) {
                    const int n = F.ln[rl][d];
                    const double lo = (double)oc[d] - (double)own.w - F.lthr[rl] - (double)g.pfMargin;
                    const double hi = (double)oc[d] + (double)own.w + (double)g.pfMargin;
                    int a, b;
                    if (g.periodic) {
                        a = (int)floor(lo * g.invL[d] * n);
                        b = (int)floor(hi * g.invL[d] * n);
                        if (b - a + 1 > n || b < a) { a = 0; b = n - 1; }
                        const int a0 = a;
                        a = wrapCell(a, n);
                        b = a + (b - a0);
                    } else {
                        a = (int)floor(lo / F.lw[rl][d]);
                        b = (int)floor(hi / F.lw[rl][d]);
                        a = a < 0 ? 0 : (a > n - 1 ? n - 1 : a);
                        b = b < 0 ? 0 : (b > n - 1 ? n - 1 : b);
                    }
                    o3[d] = a;
                    n3[d] = b - a + 1;
                }
                r0x = o3[0]; r0y = o3[1]; r0z = o3[2];
                rnx = n3[0]; rny = n3[1]; rnz = n3[2];
                if (rnx <= 0 || rny <= 0) rnz = 0;  // an empty box (non-finite bounds)
                ix = iy = iz = 0;
            }
        }
    };
    // partner entries [j, je); the next cell's range [nj, nje) is loaded one cell ahead
    int j = 0, je = 0, nj = 0, nje = 0;
    bool hasNext = false;
    if (valid && nPh > 0) {
        j = (int)e + 1;
        je = cbeg[cb + key + 1];
        const int64_t c = nextCell();
        hasNext = c >= 0;
        if (hasNext) { nj = cbeg[c]; nje = cbeg[c + 1]; }
    }
    // the FP32 test of partner box b, into this wavefront's region
    auto test = [&](bool act, const float4& b, int jj) {
        float dx = b.x - own.x, dy = b.y - own.y, dz = b.z - own.z;
        dx -= wLx * rintf(dx * wIx);
        dy -= wLy * rintf(dy * wIy);
        dz -= wLz * rintf(dz * wIz);
        const float lim = (own.w + b.w) + margin;
        const bool pass = act & !((fabsf(dx) > lim) | (fabsf(dy) > lim) | (fabsf(dz) > lim));
        const uint64_t bal = __ballot(pass);
        if (pass) {
            const int k = cntW + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            if (k < cap) ro[k] = make_int2((int)e, jj);
        }
        cntW += (int)__popcll(bal);
    };
    for (;;) {
        // lanes whose range is used up move on to their next non-empty one
        while (j >= je && hasNext) {
            j = nj;
            je = nje;
            const int64_t c = nextCell();
            hasNext = c >= 0;
            if (hasNext) { nj = cbeg[c]; nje = cbeg[c + 1]; }
        }
        const bool act = j < je, act2 = j + 1 < je;
        if (!__any(act)) break;
        // two partners per iteration, their loads issued together
        float4 pa = make_float4(0.f, 0.f, 0.f, -1e30f), pb = pa;
        if (act) pa = boxes[j];
        if (act2) pb = boxes[j + 1];
        test(act, pa, j);
        test(act2, pb, j + 1);
        j += 2;
    }
    if ((threadIdx.x & 63) == 0) {
        cnt[gw] = cntW;
        if (cntW > cap) atomicMax(wmax, cntW);
    }
}
__global__ void __launch_bounds__(WDM_BLOCK) k_pair_exact(const SceneDesc* D, const FRScene* FR, Map m, const int32_t* cvals,
                                                         const uint32_t* ckeys, const double4* eb0, const double4* eb1, const int2* einfo,
                                                         const int32_t* lbeg, const int32_t* lvals, const int32_t* cnt1, int32_t cap1,
                                                         const int2* in, int32_t* wcnt, int32_t* wmax, int32_t wcap, uint64_t* rItem,
                                                         uint32_t* rRow, uint32_t* rSlot, uint32_t* rowCnt) {
    int64_t gw, lb;
    int tt;
    if (!pairWave(m, gw, tt, lb)) return;
    const FRScene& F = FR[tt];
    const SceneDesc& S = D[F.desc];
    const PartView p = S.p;
    const GridArgs g = F.g;  // by value: the loops below store to global memory
    uint32_t* err = S.words + 2;
    GP uint64_t* rk = gpc(rItem) + gw * (int64_t)wcap;
    GP uint32_t* rr = gpc(rRow) + gw * (int64_t)wcap;
    GP uint32_t* rs = gpc(rSlot) + gw * (int64_t)wcap;
    int cntW = 0;  // pairs found (the same in every lane)
    // the pairs of the lanes with `ok`, in lane order, into this wavefront's region
    auto emit = [&](bool ok, int a, int b, const int per[3]) {
        const uint64_t bal = __ballot(ok);
        if (!bal) return;
        if (ok) {
            const int k = cntW + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            const int lo = a < b ? a : b, hi = a < b ? b : a;
            const uint32_t row = (uint32_t)(F.rowBase + lo);
            if (k < wcap) {  // only stored pairs are counted: the rows never outgrow the regions
                const uint32_t slot = atomicAdd(&rowCnt[row], 1u);
                rk[k] = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)packCd(per[0], per[1], per[2]);
                rr[k] = row;
                rs[k] = slot;
            }
        }
        cntW += (int)__popcll(bal);
    };
    // the pairs that passed the FP32 test, 64 at a time
    const int lane = threadIdx.x & 63;
    const int n1 = min(cnt1[gw], cap1);  // a region that overflowed is run again by the host
    const GP int2* ri = gpc(in) + gw * (int64_t)cap1;
    for (int b = 0; b < n1; b += 64) {
        bool ok = false;
        int oid = 0, j = 0;
        int per[3] = {0, 0, 0};
        if (b + lane < n1) {
            const int2 q = ri[b + lane];
            oid = cvals[q.x];
            j = cvals[q.y];
            const int2 ia = einfo[q.x], ib = einfo[q.y];
            ok = pairTestB(S, g, oid, eb0[q.x], eb1[q.x], ia.x, (uint32_t)ia.y, j, eb0[q.y], eb1[q.y], ib.x, (uint32_t)ib.y, true, err, per);
        }
        emit(ok, oid, j, per);
    }
    // walls, and the large bounds listed in this entry's cell, tested directly
    const int li = (int)(lb * WDM_BLOCK + threadIdx.x);
    const bool valid = li < S.nbinned;
    const int64_t e = F.binStart + (valid ? li : 0);
    const int64_t cb = F.cellBase;
    const int key = valid ? (int)(ckeys[e] - cb) : 0;
    const bool large = key >= F.nCells;
    const int id = valid ? cvals[e] : 0;
    const double4 a0 = valid ? p.bmin[id] : make_double4(0., 0., 0., 0.), a1 = valid ? p.bmax[id] : make_double4(0., 0., 0., 0.);
    const int sa = valid ? ptShape(p.ptype[id]) : 0;
    if (lbeg) {
        int ls = 0, le = 0;
        if (valid && !large) {
            ls = lbeg[cb + key];
            le = lbeg[cb + key + 1];
        }
        for (int k = 0; __any(ls + k < le); k++) {
            int per[3] = {0, 0, 0};
            const bool has = ls + k < le;
            const int jj = has ? lvals[ls + k] : 0;
            const bool ok = has && pairTest(S, g, id, a0, a1, sa, jj, true, err, per);
            emit(ok, id, jj, per);
        }
    }
    for (int u = 0; u < S.nunb; u++) {
        int per[3] = {0, 0, 0};
        const int jj = S.unb[u];
        const bool ok = valid && pairTest(S, g, id, a0, a1, sa, jj, false, err, per);
        emit(ok, id, jj, per);
    }
    if (lane == 0) {
        wcnt[gw] = cntW;
        if (cntW > wcap) atomicMax(wmax, cntW);
    }
}
// ---------------- sort-and-sweep broadphase (the alternative to the grid scan) ----------------
// The bound entries of each scene sorted by the minimum x of their FP32 boxes (radix sort of
// (scene << 32 | order-preserving bits of x - half extent)), then swept: the entry at sorted
// position i tests the entries after it whose minimum x does not exceed its maximum x, on y
// and z with the FP32 box test of k_pair_scan (x directly), and, in a periodic cell, the
// entries from the start of the scene shifted by one period (a pair that overlaps directly is
// left to the direct sweep, so each pair is emitted once).  A workgroup stages the sorted
// boxes in LDS tiles shared by its 256 consecutive positions.  Output: the pairs that pass,
// (entry, partner entry), into the same wavefront regions as k_pair_scan, which k_pair_exact
// then tests exactly, so the contact set does not depend on the broadphase.
__device__ __forceinline__ uint32_t sortableF32(float x) {
    const uint32_t u = __float_as_uint(x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__global__ void __launch_bounds__(WDM_BLOCK) k_sweep_keys(const SceneDesc* D, const FRScene* FR, Map m, const float4* boxes,
                                                         uint64_t* keys, uint32_t* vals) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const FRScene& F = FR[t];
    const int li = (int)(lb * WDM_BLOCK + threadIdx.x);
    if (li >= D[F.desc].nbinned) return;
    const int64_t e = F.binStart + li;
    const float4 b = boxes[e];
    keys[e] = ((uint64_t)(uint32_t)t << 32) | sortableF32(b.x - b.w);
    vals[e] = (uint32_t)e;
}
#define SWEEP_TILE 1024
__global__ void __launch_bounds__(WDM_BLOCK) k_pair_sweep(const SceneDesc* D, const FRScene* FR, Map m, const uint32_t* svals,
                                                         const float4* boxes, int32_t* cnt, int32_t* wmax, int32_t cap, int2* out) {
    __shared__ float4 tb[SWEEP_TILE];
    __shared__ int32_t te[SWEEP_TILE];
    int64_t lb;
    const int t = mapBlock(m, lb);
    const FRScene& F = FR[t];
    const GridArgs& g = F.g;
    const int nb = D[F.desc].nbinned;
    const int64_t gw = (int64_t)blockIdx.x * (WDM_BLOCK / 64) + (threadIdx.x >> 6);
    GP int2* ro = gpc(out) + gw * (int64_t)cap;
    int cntW = 0;
    const int li = (int)(lb * WDM_BLOCK + threadIdx.x);
    const bool valid = li < nb;
    const int64_t s0 = F.binStart, end = F.binStart + nb, si = s0 + li;
    const int e = valid ? (int)svals[si] : 0;
    const float4 own = valid ? boxes[e] : make_float4(0.f, 0.f, 0.f, -1e30f);
    const float margin = g.pfMargin;
    const float xmax = own.x + own.w + margin;
    const float wLy = g.periodic ? g.Lf[1] : 0.f, wLz = g.periodic ? g.Lf[2] : 0.f;
    const float wIy = g.periodic ? g.invLf[1] : 0.f, wIz = g.periodic ? g.invLf[2] : 0.f;
    const float Lx = g.periodic ? g.Lf[0] : 0.f;
    auto emit = [&](bool pass, int jj) {
        const uint64_t bal = __ballot(pass);
        if (pass) {
            const int k = cntW + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            if (k < cap) ro[k] = make_int2(e, jj);
        }
        cntW += (int)__popcll(bal);
    };
    // y and z with the periodic minimum image; x by the caller
    auto yz = [&](const float4& b) {
        float dy = b.y - own.y, dz = b.z - own.z;
        dy -= wLy * rintf(dy * wIy);
        dz -= wLz * rintf(dz * wIz);
        const float lim = (own.w + b.w) + margin;
        return !((fabsf(dy) > lim) | (fabsf(dz) > lim));
    };
    // direct sweep over LDS tiles of the sorted boxes after the block's first position
    bool act = valid;
    for (int64_t ts = s0 + (int64_t)lb * WDM_BLOCK + 1;; ts += SWEEP_TILE) {
        if (!__syncthreads_or(act && ts < end)) break;
        for (int k = threadIdx.x; k < SWEEP_TILE; k += WDM_BLOCK) {
            const int64_t q = ts + k;
            if (q < end) {
                const int e2 = (int)svals[q];
                tb[k] = boxes[e2];
                te[k] = e2;
            }
        }
        __syncthreads();
        if (__any(act)) {
            for (int k = 0; k < SWEEP_TILE; k++) {
                const int64_t q = ts + k;
                bool my = act && q > si;
                if (act && q >= end) { act = false; my = false; }
                const float4 b = my ? tb[k] : make_float4(0.f, 0.f, 0.f, 0.f);
                if (my && b.x - b.w > xmax) { act = false; my = false; }  // sorted: none after it overlaps
                bool pass = false;
                if (my) pass = fabsf(b.x - own.x) <= (own.w + b.w) + margin && yz(b);
                emit(pass, my ? te[k] : 0);
                if (!__any(act)) break;
            }
        }
        __syncthreads();
    }
    // periodic x: the entries from the scene's start, one period up, up to this one's maximum x
    // (every position looks: a box whose minimum x is below 0 reaches boxes whose maximum x
    // is below L; the first sorted entry ends the walk at once for most)
    if (g.periodic) {
        bool wr = valid;
        for (int64_t q = s0; __any(wr); q++) {
            bool my = wr && q < end && q != si;
            if (wr && q >= end) wr = false;
            float4 b = make_float4(0.f, 0.f, 0.f, 0.f);
            int e2 = 0;
            if (my) {
                e2 = (int)svals[q];
                b = boxes[e2];
                if (b.x - b.w + Lx > xmax) { wr = false; my = false; }
            }
            bool pass = false;
            if (my) {
                const float lim = (own.w + b.w) + margin;
                const bool direct = fabsf(b.x - own.x) <= lim;
                pass = !direct && fabsf(b.x + Lx - own.x) <= lim && yz(b);
            }
            emit(pass, e2);
        }
    }
    if ((threadIdx.x & 63) == 0) {
        cnt[gw] = cntW;
        if (cntW > cap) atomicMax(wmax, cntW);
    }
}

// each pair of the wavefront regions at its row's offset plus its slot (rows by min id); a
// region that overflowed holds its first wcap pairs (the host runs the pass again)
__global__ void __launch_bounds__(WDM_BLOCK) k_cand_scatter(const int32_t* wcnt, int64_t nw, int32_t wcap, const uint64_t* rItem,
                                                           const uint32_t* rRow, const uint32_t* rSlot, const int64_t* rowOff, BItem* tmp) {
    const int64_t w = (int64_t)blockIdx.x * (WDM_BLOCK / 64) + (threadIdx.x >> 6);
    if (w >= nw) return;
    const int n = min(wcnt[w], wcap);
    for (int k = threadIdx.x & 63; k < n; k += 64) {
        const int64_t i = w * (int64_t)wcap + k;
        const uint32_t row = rRow[i];
        const int64_t pos = rowOff[row] + rSlot[i];
        putItem(tmp, pos, rItem[i], row);
    }
}
// the sorted candidates: contact key (min id << 32 | max id) and packed periods
struct CandOut {
    const FRScene* FR;
    int nT;
    uint64_t* candK;
    int32_t* candP;
    __device__ void operator()(uint32_t b, int64_t pos, uint64_t x) const {
        const FRScene& F = FR[sceneOfRow(FR, nT, b)];
        candK[pos] = ((uint64_t)(b - F.rowBase) << 32) | (x >> 32);
        candP[pos] = (int32_t)(uint32_t)x;
    }
};

__device__ __forceinline__ int64_t lowerBound64(const uint64_t* a, int64_t n, uint64_t k) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] < k) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Rows: the candidates (batch-wide, by scene then contact key) of min id p of scene t are
// [rowOff[rowBase + p], rowOff[rowBase + p + 1]); the old contacts of min id p are
// [oldRow[rowBase + p], oldRow[rowBase + p + 1]) of the scene's old list.
// oldRow of a scene: first old contact of each min id (binary search in the sorted keys)
template <typename T>
__global__ void __launch_bounds__(WDM_BLOCK) k_old_rows(const SceneDesc* D, const ConDesc<T>* C, const FRScene* FR, Map m, int64_t* oldRow) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const FRScene& F = FR[t];
    const int np = D[F.desc].p.np;
    const int64_t pi = lb * WDM_BLOCK + threadIdx.x;
    if (pi > np) return;
    oldRow[F.rowBase + pi] = lowerBound64(C[t].cv.key, F.nOld, (uint64_t)pi << 32);
}
// position of contact key k among the candidates of the scene (>= k; its row holds every
// candidate of that min id)
__device__ __forceinline__ int64_t candLowerBound(const FRScene& F, const int64_t* rowOff, const uint64_t* candK, uint64_t k) {
    const int64_t row = F.rowBase + (int64_t)(k >> 32);
    const int64_t rs = rowOff[row], re = rowOff[row + 1];
    return rs + lowerBound64(candK + rs, re - rs, k);
}

// old contacts that stay only because they are real (their bounds no longer overlap)
template <typename T>
__global__ void __launch_bounds__(WDM_BLOCK) k_flag_old(const ConDesc<T>* C, const FRScene* FR, Map m, const int64_t* rowOff,
                                                       const uint64_t* candK, uint8_t* flag) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const FRScene& F = FR[t];
    const ConView<T> oc = C[t].cv;
    int64_t j = lb * WDM_BLOCK + threadIdx.x;
    if (j >= F.nOld) return;
    uint32_t inf = oc.info[j];
    int f = 0;
    if ((inf & CI_REAL) && !(inf & CI_DEAD)) {
        const uint64_t k = oc.key[j];
        const int64_t row = F.rowBase + (int64_t)(k >> 32);
        const int64_t l = candLowerBound(F, rowOff, candK, k);
        f = (l < rowOff[row + 1] && candK[l] == k) ? 0 : 1;
    }
    flag[F.oldStart + j] = f;
}

template <typename T>
__device__ __forceinline__ void copyContact(const ConView<T>& src, int64_t s, const ConView<T>& dst, int64_t d) {
    dst.ids[d] = src.ids[s];
    dst.cd[d] = src.cd[s];
}

// merged list = candidates (bounds overlap) U old real contacts (InsertionSortCollider keeps
// real contacts, :92-96); an old contact that is not dead keeps its state, ids order and cellDist
// Dormancy: between full runs every node moves by at most verletDist from its position at
// the run (InsertionSortCollider.cpp:340-350), so a sphere-sphere or wall-sphere pair in
// the same periodic image whose surface gap exceeds 2 verletDist now stays apart until the
// next full run; the contact kernel skips it without changing any result.
__device__ __forceinline__ bool dormantPair(const NodeView& n, const PartView& p, int A, int B, uint32_t ty, int32_t cdp, double vd,
                                            double df) {
    if (cdp != 0) return false;
    double4 pa = n.posr[A], pb = n.posr[B];
    if (ty == CI_SS) {
        D3 r = xyz(pb) - xyz(pa);
        double R = fabs(df) * (pa.w + pb.w);
        return wsqrt(sqn3(r)) - R > 2 * vd * (1 + 1e-6) + 1e-12 * R;
    }
    if (ty == CI_WS) {
        int ax = ptAxis(p.ptype[A]);
        double dist = comp3(xyz(pb), ax) - comp3(xyz(pa), ax);
        return fabs(dist) - pb.w > 2 * vd * (1 + 1e-6) + 1e-12 * pb.w;
    }
    return false;
}

// ooRank: exclusive prefix (batch-wide) of the flags of k_flag_old; a scene's old contacts
// with index < j that stay only because they are real number ooRank[oldStart + j] - ooRank[oldStart]
// A pair that touches now becomes real in the ContactLoop that follows the full run
// (Cg2_Sphere_Sphere_L6Geom / Cg2_Wall_Sphere_L6Geom accept it): it gets its state row at
// the full run, in key order, instead of one handed out in execution order by the law.
// Only the row order depends on this test.
__device__ __forceinline__ bool touchesNow(const StepArgs& a, const NodeView& n, const PartView& p, int A, int B, uint32_t ty, int32_t cdp) {
    double4 pa = n.posr[A], pb = n.posr[B];
    if (ty == CI_SS) {
        D3 p2 = xyz(pb);
        if (a.periodic) p2 = p2 + mv3(ldM3(a.cell->hSize), mk3((double)unpack10(cdp, 0), (double)unpack10(cdp, 10), (double)unpack10(cdp, 20)));
        D3 r = p2 - xyz(pa);
        double df = fabs(a.distFactor) * (pa.w + pb.w);
        return sqn3(r) - df * df <= 0;
    }
    if (ty == CI_WS) {
        int ax = ptAxis(p.ptype[A]);
        return fabs(comp3(xyz(pb), ax) - comp3(xyz(pa), ax)) <= pb.w;
    }
    return false;
}

// Work-list order of the contacts that enter it without a state row: near pairs (surface gap
// below NEAR_GAP verletDist at the full run) first, then the others, each group in key order.
// Nearly every contact that becomes real before the next full run is a near pair (on the
// 1M-sphere periodic packing 99.7 % of them had a gap under 0.05 verletDist, against half of
// the other non-dormant pairs), so the wavefronts of the far pairs keep to the short path.
// Only the order of the items depends on this test.
#define NEAR_GAP 0.05
__device__ __forceinline__ bool nearPair(const StepArgs& a, const NodeView& n, const PartView& p, int A, int B, uint32_t ty, int32_t cdp) {
    double4 pa = n.posr[A], pb = n.posr[B];
    const double lim = NEAR_GAP * a.verletDist;
    if (ty == CI_SS) {
        D3 p2 = xyz(pb);
        if (a.periodic) p2 = p2 + mv3(ldM3(a.cell->hSize), mk3((double)unpack10(cdp, 0), (double)unpack10(cdp, 10), (double)unpack10(cdp, 20)));
        return wsqrt(sqn3(p2 - xyz(pa))) - fabs(a.distFactor) * (pa.w + pb.w) < lim;
    }
    if (ty == CI_WS) {
        int ax = ptAxis(p.ptype[A]);
        return fabs(comp3(xyz(pb), ax) - comp3(xyz(pa), ax)) - pb.w < lim;
    }
    return true;  // facets: no gap test here
}

template <typename T>
__global__ void __launch_bounds__(WDM_BLOCK) k_emit_cand(const SceneDesc* D, const ConDesc<T>* C, const FRScene* FR, Map m,
                                                        const int64_t* ooRank, const int64_t* oldRow, const uint64_t* ckey,
                                                        const int32_t* cper, int32_t* oslot, uint8_t* rflag) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const FRScene& F = FR[t];
    const SceneDesc& S = D[F.desc];
    const NodeView n = S.n;
    const PartView p = S.p;
    const ConView<T> oldc = C[t].cv, nc = C[t].cvNew;
    const int64_t nold = F.nOld;
    const double vd = S.a.verletDist, df = S.a.distFactor;
    int64_t i = lb * WDM_BLOCK + threadIdx.x;
    if (i >= F.nCand) return;
    const uint64_t k = ckey[F.candStart + i];
    const int64_t row = F.rowBase + (int64_t)(k >> 32);
    const int64_t os = oldRow[row];
    const int64_t l = os + lowerBound64(oldc.key + os, oldRow[row + 1] - os, k);
    int64_t d = i + (ooRank[F.oldStart + l] - ooRank[F.oldStart]);
    int64_t gd = F.newStart + d;
    nc.key[d] = k;
    if (l < nold && oldc.key[l] == k && !(oldc.info[l] & CI_DEAD)) {
        copyContact(oldc, l, nc, d);
        uint32_t inf = oldc.info[l];
        bool r = (inf & CI_REAL) != 0;
        uint32_t ni = (inf & ~(CI_PENDING | CI_DEAD | CI_DORMANT | CI_NEAR)) | CI_OVERLAP;
        bool fresh = false;  // not real, touching: new row
        if (!r) {
            int2 id = oldc.ids[l];
            if (dormantPair(n, p, id.x, id.y, ni & CI_TYPE_MASK, oldc.cd[l], vd, df)) ni |= CI_DORMANT;
            else {
                fresh = touchesNow(S.a, n, p, id.x, id.y, ni & CI_TYPE_MASK, oldc.cd[l]);
                if (!fresh && nearPair(S.a, n, p, id.x, id.y, ni & CI_TYPE_MASK, oldc.cd[l])) ni |= CI_NEAR;
            }
        }
        nc.info[d] = ni;
        oslot[gd] = r ? oldc.slot[l] : (fresh ? -2 : -1);
        rflag[gd] = (r || fresh) ? 1 : 0;
    } else {
        int lo = (int)(k >> 32), hi = (int)(k & 0xffffffffu);
        int A, B, ty;
        canonical(lo, hi, ptShape(p.ptype[lo]), ptShape(p.ptype[hi]), A, B, ty);
        nc.ids[d] = make_int2(A, B);
        int32_t pc = cper[F.candStart + i];  // periods computed in canonical order
        nc.cd[d] = pc;
        const bool dorm = dormantPair(n, p, A, B, (uint32_t)ty, pc, vd, df);
        const bool fresh = !dorm && touchesNow(S.a, n, p, A, B, (uint32_t)ty, pc);
        const bool nearp = !dorm && !fresh && nearPair(S.a, n, p, A, B, (uint32_t)ty, pc);
        nc.info[d] = (uint32_t)ty | CI_OVERLAP | (dorm ? CI_DORMANT : 0u) | (nearp ? CI_NEAR : 0u);
        oslot[gd] = fresh ? -2 : -1;
        rflag[gd] = fresh ? 1 : 0;
    }
}
template <typename T>
__global__ void __launch_bounds__(WDM_BLOCK) k_emit_oor(const ConDesc<T>* C, const FRScene* FR, Map m, const uint8_t* flag,
                                                       const int64_t* ooRank, const int64_t* rowOff, const uint64_t* ckey, int32_t* oslot,
                                                       uint8_t* rflag) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const FRScene& F = FR[t];
    const ConView<T> oldc = C[t].cv, nc = C[t].cvNew;
    int64_t j = lb * WDM_BLOCK + threadIdx.x;
    if (j >= F.nOld || !flag[F.oldStart + j]) return;
    uint64_t k = oldc.key[j];
    int64_t d = (candLowerBound(F, rowOff, ckey, k) - F.candStart) + (ooRank[F.oldStart + j] - ooRank[F.oldStart]);
    int64_t gd = F.newStart + d;
    nc.key[d] = k;
    copyContact(oldc, j, nc, d);
    nc.info[d] = oldc.info[j] & ~(CI_PENDING | CI_DEAD | CI_OVERLAP | CI_DORMANT | CI_NEAR);
    oslot[gd] = oldc.slot[j];
    rflag[gd] = 1;
}

// dense state rows for the real (and touching) contacts of the merged list, in key order
// (rscan: exclusive batch-wide prefix of rflag); oslot -2: row without state to copy
// nrflag: for a contact that enters the work list without a state row (not dormant), bit 0,
// and bit 1 when it is a near pair; its scan (ScanNearSplit) counts both (NR_ALL, NR_NEAR)
// srcRow[newStart + row]: the old state row each new row copies (-2: none, a touching pair)
template <typename T>
__global__ void __launch_bounds__(WDM_BLOCK) k_copy_state(const ConDesc<T>* C, const FRScene* FR, Map m, const int32_t* oslot,
                                                         const int64_t* rscan, uint8_t* nrflag, int32_t* srcRow) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const FRScene& F = FR[t];
    const ConView<T> nc = C[t].cvNew;
    int64_t d = lb * WDM_BLOCK + threadIdx.x;
    if (d >= F.nNew) return;
    int32_t os = oslot[F.newStart + d];
    if (os == -1) {
        nc.slot[d] = -1;
        const uint32_t inf = nc.info[d];
        nrflag[F.newStart + d] = (inf & CI_DORMANT) ? 0 : ((inf & CI_NEAR) ? 3 : 1);
        return;
    }
    nrflag[F.newStart + d] = 0;
    int64_t ns = rscan[F.newStart + d] - rscan[F.newStart];
    nc.slot[d] = (int32_t)ns;
    srcRow[F.newStart + ns] = os;
}
// the state of the new rows, one thread per row (the rows of a scene number at most its contacts)
template <typename T>
__global__ void __launch_bounds__(WDM_BLOCK) k_copy_rows(const ConDesc<T>* C, const FRScene* FR, Map m, const int64_t* rscan,
                                                        const int32_t* srcRow) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const FRScene& F = FR[t];
    const ConView<T> oldc = C[t].cv, nc = C[t].cvNew;
    const int64_t ns = lb * WDM_BLOCK + threadIdx.x;
    if (ns >= rscan[F.newStart + F.nNew] - rscan[F.newStart]) return;
    const int32_t os = srcRow[F.newStart + ns];
    if (os < 0) return;
    const int64_t nROld = *oldc.counter;
    for (int k = 0; k < NSTATE; k++) nc.st[(size_t)k * nc.cap + ns] = *stAt(oldc, os, k, nROld);
    if constexpr (sizeof(T) == 8) nc.cpt[ns] = oldc.cpt[os];
    else mixedCpt(nc.cpt)[ns] = mixedCpt(oldc.cpt)[os];
}
#define NR_ALL(v) ((v) & 0xffffffffLL)
#define NR_NEAR(v) ((v) >> 32)
// work list of the ContactLoop: a contact given a state row at the full run (real or
// touching) is item `row`, the other non-dormant ones follow, near pairs first, each group
// in key order (nrscan: exclusive batch-wide prefix of nrflag), and take the row of their item
template <typename T>
__global__ void __launch_bounds__(WDM_BLOCK) k_items(const ConDesc<T>* C, const FRScene* FR, Map m, const int64_t* rscan,
                                                    const int64_t* nrscan) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const FRScene& F = FR[t];
    const ConView<T> nc = C[t].cvNew;
    int64_t d = lb * WDM_BLOCK + threadIdx.x;
    if (d >= F.nNew) return;
    const int32_t row = nc.slot[d];
    int64_t item = -1;
    if (row >= 0) item = row;
    else if (!(nc.info[d] & CI_DORMANT)) {
        const int64_t nRows = rscan[F.newStart + F.nNew] - rscan[F.newStart];
        const int64_t nNear = NR_NEAR(nrscan[F.newStart + F.nNew] - nrscan[F.newStart]);
        const int64_t r = nrscan[F.newStart + d] - nrscan[F.newStart];
        item = nRows + ((nc.info[d] & CI_NEAR) ? NR_NEAR(r) : nNear + (NR_ALL(r) - NR_NEAR(r)));
        nc.slot[d] = (int32_t)item;  // its state row if it becomes real before the next full run
    }
    nc.itm[d] = (int32_t)item;
    if (item >= 0) {
        const int2 id = nc.ids[d];
        nc.irec[item] = make_int4(id.x, id.y, (int32_t)nc.info[d], nc.cd[d]);
        nc.iaux[item] = make_int2((int32_t)item, (int32_t)d);
    }
}
// state rows handed out at the full run (real and touching contacts) and items of each scene
template <typename T>
__global__ void k_set_counters(const ConDesc<T>* C, const FRScene* FR, int n, const int64_t* rscan, const int64_t* nrscan) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const FRScene& F = FR[t];
    const int64_t nRows = rscan[F.newStart + F.nNew] - rscan[F.newStart];
    *C[t].cvNew.counter = (int32_t)nRows;
    *C[t].cvNew.nItems = (int32_t)(nRows + NR_ALL(nrscan[F.newStart + F.nNew] - nrscan[F.newStart]));
}

// CSR of the new lists (contact -> particle): per particle its contacts in list order (the
// contacts come in (min id, max id) order, so the other id ascends: first those where the
// particle is the max id, then those where it is the min id), entry value work-list item |
// isB << 31.  Dormant contacts cannot become real before the next full run and dead ones
// never again: they get no entries.  Each entry counts in cnt[rowBase + particle] (the slot
// returned) and is sorted within its particle by the contact index (bucket item c << 32 |
// value); bucket b == UINT32_MAX marks an entry that is left out.
template <typename T>
__global__ void __launch_bounds__(WDM_BLOCK) k_csr_count(const ConDesc<T>* C, const FRScene* FR, Map m, uint32_t* eB, uint32_t* eS,
                                                        uint64_t* eI, uint32_t* cnt) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const FRScene& F = FR[t];
    const ConView<T> nc = C[t].cvNew;
    const int64_t c = lb * WDM_BLOCK + threadIdx.x;
    const int64_t g = F.newStart + c;
    bool act = c < F.nNew;
    if (act && (nc.info[c] & (CI_DORMANT | CI_DEAD))) {
        eB[2 * g] = UINT32_MAX;
        eB[2 * g + 1] = UINT32_MAX;
        act = false;
    }
    int2 id = make_int2(0, 0);
    if (act) id = nc.ids[c];
    const uint32_t b0 = act ? (uint32_t)(F.rowBase + id.x) : UINT32_MAX, b1 = (uint32_t)(F.rowBase + id.y);
    // side A: the contacts come sorted by min id, so lanes in a row mostly share b0; one
    // atomic per run of equal b0 (lanes of a run take consecutive slots)
    const int lane = threadIdx.x & 63;
    const uint32_t prev = (uint32_t)__shfl_up((int)b0, 1, 64);
    const bool head = lane == 0 || prev != b0;
    const uint64_t heads = __ballot(head);
    const uint64_t upTo = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    const int hl = 63 - __clzll(heads & upTo);
    const uint64_t after = heads & ~upTo;
    const int runEnd = after ? __ffsll((long long)after) - 1 : 64;
    uint32_t base = 0;
    if (head && act) base = atomicAdd(&cnt[b0], (uint32_t)(runEnd - hl));
    base = (uint32_t)__shfl((int)base, hl, 64);
    if (!act) return;
    const uint32_t it = (uint32_t)nc.itm[c];
    eB[2 * g] = b0;
    eS[2 * g] = base + (uint32_t)(lane - hl);
    eI[2 * g] = ((uint64_t)c << 32) | it;
    eB[2 * g + 1] = b1;
    eS[2 * g + 1] = atomicAdd(&cnt[b1], 1u);
    eI[2 * g + 1] = ((uint64_t)c << 32) | (it | 0x80000000u);
}
__global__ void __launch_bounds__(WDM_BLOCK) k_csr_scatter(const uint32_t* eB, const uint32_t* eS, const uint64_t* eI, int64_t n,
                                                          const int64_t* off, BItem* tmp) {
    const int64_t k = blockIdx.x * (int64_t)WDM_BLOCK + threadIdx.x;
    if (k >= n) return;
    const uint32_t b = eB[k];
    if (b == UINT32_MAX) return;
    const int64_t pos = off[b] + eS[k];
    putItem(tmp, pos, eI[k], b);
}
// the sorted entries into each scene's own CSR array
struct CsrOut {
    const FRScene* FR;
    int nT;
    const int64_t* off;
    __device__ void operator()(uint32_t b, int64_t pos, uint64_t x) const {
        const FRScene& F = FR[sceneOfRow(FR, nT, b)];
        F.entV[pos - off[F.rowBase]] = (uint32_t)x;
    }
};
__global__ void __launch_bounds__(WDM_BLOCK) k_csr_offsets(const SceneDesc* D, const FRScene* FR, Map m, const int64_t* off) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const FRScene& F = FR[t];
    const int np = D[F.desc].p.np;
    int i = (int)(lb * WDM_BLOCK + threadIdx.x);
    if (i > np) return;
    F.csrOff[i] = off[F.rowBase + i] - off[F.rowBase];
}

// ---------------- work-list relayout between full runs ----------------
// Contacts that become real after a full run sit among the items after the rows (the items
// that were apart at the full run), each on its own lines (work-list record, contact-major
// state row, gather record), and the ContactLoop slows down as they accumulate.  The relayout
// moves the real ones among those items to the front of that part of the list (in their
// order), right after the rows, with component-major state rows (the row block grows by
// them), and the others after them (in their order); the rows stay in place.  The list's
// other buffer, idle between full runs, stages the moved items.  Only memory order changes:
// the contacts, their order and every stored value stay as they are.  Map m: per scene the
// blocks of its nc contacts; the flat index of item j of scene t is m.bprefix[t] * WDM_BLOCK + j.
struct RlScene {
    GP uint32_t* entV;        // CSR entries: work-list item | isB << 31
    const GP int64_t* csrOff; // csrOff[np]: entries in use
    int32_t np;
};
// items after the rows: 1 real, 2 not; 0 for the rows and past the work list (ScanNearSplit
// counts 1 and 2 apart)
template <typename T>
__global__ void __launch_bounds__(WDM_BLOCK) k_rl_flags(const ConDesc<T>* C, Map m, uint8_t* flag) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const ConView<T> cv = C[t].cv;
    const int64_t j = lb * WDM_BLOCK + threadIdx.x;
    uint8_t f = 0;
    if (j >= *cv.counter && j < *cv.nItems) f = ((uint32_t)cv.irec[j].z & CI_REAL) ? 1 : 2;
    flag[(int64_t)blockIdx.x * WDM_BLOCK + threadIdx.x] = f;
}
// the new place of item j >= nR (the rows): nR + its rank among the real ones, or nR + the
// number of real ones + its rank among the others
__device__ __forceinline__ int64_t rlPlace(const Map& m, int t, const int64_t* scan, int64_t j, int64_t nR, bool real, int64_t& nNew) {
    const int64_t g0 = m.bprefix[t] * WDM_BLOCK;
    nNew = (scan[m.bprefix[t + 1] * WDM_BLOCK] - scan[g0]) & 0xffffffffll;
    const int64_t sj = scan[g0 + j] - scan[g0];
    return nR + (real ? (sj & 0xffffffffll) : nNew + (sj >> 32));
}
// stage item j (>= nR) at k = its new place - nR in the other buffer b: work-list record, and
// for a real one its state row (contact-major, as it is), contact point and gather record;
// the contact's item and row point to the new place
template <typename T>
__global__ void __launch_bounds__(WDM_BLOCK) k_rl_out(const ConDesc<T>* C, Map m, const int64_t* scan, int32_t* map) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const ConView<T> a = C[t].cv, b = C[t].cvNew;
    const int64_t j = lb * WDM_BLOCK + threadIdx.x;
    const int64_t nR = *a.counter;
    if (j < nR || j >= *a.nItems) return;
    const int4 rec = a.irec[j];
    const int c = a.iaux[j].y;
    const bool real = ((uint32_t)rec.z & CI_REAL) != 0;
    int64_t nNew;
    const int64_t nj = rlPlace(m, t, scan, j, nR, real, nNew), k = nj - nR;
    b.irec[k] = rec;
    b.iaux[k] = make_int2((int32_t)nj, c);
    a.itm[c] = (int32_t)nj;
    a.slot[c] = (int32_t)nj;
    map[m.bprefix[t] * WDM_BLOCK + j] = (int32_t)nj;
    if (!real) return;
    for (int q = 0; q < NSTATE; q++) b.st2[(size_t)k * NSTRIDE + q] = *stAt(a, j, q, nR);
    if constexpr (sizeof(T) == 8) {
        b.cpt[k] = a.cpt[j];
        b.gfc[2 * k] = a.gfc[2 * j];
        b.gfc[2 * k + 1] = a.gfc[2 * j + 1];
    } else {
        mixedCpt(b.cpt)[k] = mixedCpt(a.cpt)[j];
        mixedRecA(b.gfc)[k] = mixedRecA(a.gfc)[j];
        mixedRecB(b.gfc, b.cap)[k] = mixedRecB(a.gfc, a.cap)[j];
        mixedRecC(b.gfc, b.cap)[k] = mixedRecC(a.gfc, a.cap)[j];
    }
}
// the staged items back at nR + k; the real ones (k < nNew) into the grown component-major block
template <typename T>
__global__ void __launch_bounds__(WDM_BLOCK) k_rl_in(const ConDesc<T>* C, Map m, const int64_t* scan) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const ConView<T> a = C[t].cv, b = C[t].cvNew;
    const int64_t k = lb * WDM_BLOCK + threadIdx.x;
    const int64_t nR = *a.counter;
    if (k >= *a.nItems - nR) return;
    const int64_t g0 = m.bprefix[t] * WDM_BLOCK;
    const int64_t nNew = (scan[m.bprefix[t + 1] * WDM_BLOCK] - scan[g0]) & 0xffffffffll;
    const int64_t nj = nR + k;
    a.irec[nj] = b.irec[k];
    a.iaux[nj] = b.iaux[k];
    if (k >= nNew) return;
    for (int q = 0; q < NSTATE; q++) *stAt(a, nj, q, nR + nNew) = b.st2[(size_t)k * NSTRIDE + q];
    if constexpr (sizeof(T) == 8) {
        a.cpt[nj] = b.cpt[k];
        a.gfc[2 * nj] = b.gfc[2 * k];
        a.gfc[2 * nj + 1] = b.gfc[2 * k + 1];
    } else {
        mixedCpt(a.cpt)[nj] = mixedCpt(b.cpt)[k];
        mixedRecA(a.gfc)[nj] = mixedRecA(b.gfc)[k];
        mixedRecB(a.gfc, a.cap)[nj] = mixedRecB(b.gfc, b.cap)[k];
        mixedRecC(a.gfc, a.cap)[nj] = mixedRecC(b.gfc, b.cap)[k];
    }
}
// realness words: the rows' bits stay, the moved real items' bits set, the others' cleared
template <typename T>
__global__ void __launch_bounds__(WDM_BLOCK) k_rl_mask(const ConDesc<T>* C, Map m, const int64_t* scan) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const ConView<T> a = C[t].cv;
    const int64_t w = lb * WDM_BLOCK + threadIdx.x;  // items [64 w, 64 w + 64)
    if (w >= (m.bprefix[t + 1] - m.bprefix[t]) * (WDM_BLOCK / 64)) return;
    const int64_t g0 = m.bprefix[t] * WDM_BLOCK;
    const int64_t nR = *a.counter;
    const int64_t nNew = (scan[m.bprefix[t + 1] * WDM_BLOCK] - scan[g0]) & 0xffffffffll;
    auto below = [](int64_t n) { return n >= 64 ? ~0ull : (n <= 0 ? 0ull : ((1ull << n) - 1)); };
    const uint64_t keep = below(nR - 64 * w), upto = below(nR + nNew - 64 * w);
    C[t].realMask[w] = (C[t].realMask[w] & keep) | (upto & ~keep);
}
// CSR entries of moved items through the item map (their order, by the other particle's id, stays)
template <typename T>
__global__ void __launch_bounds__(WDM_BLOCK) k_rl_csr(const ConDesc<T>* C, const RlScene* R, Map mE, Map m, const int32_t* map) {
    int64_t lb;
    const int t = mapBlock(mE, lb);
    const RlScene& r = R[t];
    const int64_t e = lb * WDM_BLOCK + threadIdx.x;
    if (e >= r.csrOff[r.np]) return;
    const uint32_t v = r.entV[e], item = v & 0x7fffffffu;
    if ((int64_t)item < *C[t].cv.counter) return;
    r.entV[e] = (uint32_t)map[m.bprefix[t] * WDM_BLOCK + item] | (v & 0x80000000u);
}
// the row block grown by the moved real items
template <typename T>
__global__ void k_rl_counter(const ConDesc<T>* C, Map m, const int64_t* scan, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t g0 = m.bprefix[t] * WDM_BLOCK;
    const int64_t nNew = (scan[m.bprefix[t + 1] * WDM_BLOCK] - scan[g0]) & 0xffffffffll;
    *C[t].cv.counter = (int32_t)(*C[t].cv.counter + nNew);
}

// ---------------- ContactLoop::run, pkg/dem/ContactLoop.cpp:93-259 ----------------
// One thread per contact: pending removal decided by the stored-bounds overlap
// (ContactContainer::removePending), geometry (Cg2_*_L6Geom + handleSpheresLikeContact,
// pkg/dem/L6Geom.cpp:51-217), physics on creation (Cp2_*), law (Law2_*), global force.

template <typename T>
struct Geo {
    V3T<T> normal;
    D3 contPt;
    T uN, r1, r2;
    D3 pos1, pos2, vel1, vel2, angVel1, angVel2;
};

// L6Geom::setInitialLocalCoords, pkg/dem/L6Geom.cpp:27-38
template <typename T>
__device__ __forceinline__ void initialLocalCoords(const V3T<T>& locX, T* tr) {
    V3T<T> e = (fabs(locX.y) < fabs(locX.z)) ? mk3<T>(0, 1, 0) : mk3<T>(0, 0, 1);
    V3T<T> locY = cross3(locX, e);
    locY = locY - locX * dot3(locY, locX);
    locY = normalized3(locY);
    V3T<T> locZ = cross3(locX, locY);
    tr[0] = locX.x; tr[1] = locY.x; tr[2] = locZ.x;
    tr[3] = locX.y; tr[4] = locY.y; tr[5] = locZ.y;
    tr[6] = locX.z; tr[7] = locY.z; tr[8] = locZ.z;
}
template <typename T> __device__ __forceinline__ V3T<T> trCol(const T* tr, int j) { return mk3<T>(tr[j], tr[3 + j], tr[6 + j]); }
template <typename T> __device__ __forceinline__ void trSetCol(T* tr, int j, const V3T<T>& c) { tr[j] = c.x; tr[3 + j] = c.y; tr[6 + j] = c.z; }

// GEOM: the Cg2 functors compiled in (G_SS | G_WS | G_FS), so that a scene of spheres only
// does not carry the registers of the facet geometry
#define G_SS 1
#define G_WS 2
#define G_FS 4
template <typename T, int LAW, int GEOM>
__global__ void __launch_bounds__(WDM_BLOCK) k_contacts(const SceneDesc* D, const ConDesc<T>* C, Map m, double* epart, int spec,
                                                       const StepVar* sv) {
    const int64_t blk = xcdChunkBlock(m.bprefix[m.n], XCD_CHUNK_CON);
    if (blk < 0) return;
    int64_t lb;
    const int tt = mapBlockAt(m, blk, lb);
    const ConDesc<T>& CD = C[tt];
    const SceneDesc& S = D[CD.desc];
    StepArgs a = S.a;
    const NodeView n = S.n;
    const PartView p = S.p;
    const ConView<T> cv = CD.cv;
    uint32_t* err = S.words + 2;
    uint32_t* creOut = S.words + 1;
    uint64_t* realMask = CD.realMask;
    if (sv) {  // a step of a window: stopped scenes skip it
        const StepVar v = sv[CD.desc];
        if (!contactActive(v, S.words)) return;
        applyStepVar(a, v);
    } else if (spec && loadWord(S.words) == a.stamp) return;  // the collider runs first: launched again afterwards
    // work list: dormant contacts are not in it; contacts with a state row come first, as
    // item == row, so their state is read and written contiguously
    const int64_t nItems = *cv.nItems;
    const int64_t nR = *cv.counter;
    const int64_t item = lb * WDM_BLOCK + threadIdx.x;
    double ev[5] = {0., 0., 0., 0., 0.};  // elast, plast, viscN, viscT, broken
    uint32_t bits = 0;
    bool realOut = false;  // contact is real after this step's law
    if (item < nItems) do {
        const int4 rec = cv.irec[item];
        const int2 aux = cv.iaux[item];
        const int64_t c = aux.y;
        GP int32_t* infoItem = (GP int32_t*)(cv.irec + item) + 2;
        uint32_t info = (uint32_t)rec.z;
        if (info & (CI_DEAD | CI_DORMANT)) break;
        uint32_t ty = info & CI_TYPE_MASK;
        if (info & CI_PENDING) {  // ContactContainer::removePending (ContactContainer.hpp:78-93)
            info &= ~CI_PENDING;
            if (!(info & CI_OVERLAP)) info |= CI_DEAD;
            cv.info[c] = info;
            *infoItem = (int32_t)info;
            if (info & CI_DEAD) break;
        }
        if (ty == CI_NONE) break;
        const bool real = (info & CI_REAL) != 0;
        const int2 id = make_int2(rec.x, rec.y);
        const double dt = a.dt;
        double4 pa = n.posr[id.x], pb = n.posr[id.y];
        D3 shift2 = mk3(0., 0., 0.);
        int32_t cdp = 0;
        if (a.periodic) {
            cdp = rec.w;
            shift2 = mv3(ldM3(a.cell->hSize), mk3((double)unpack10(cdp, 0), (double)unpack10(cdp, 10), (double)unpack10(cdp, 20)));
        }
        T tr[9];
        int32_t row = real ? aux.x : -1;
        // L6Geom.trsf: only columns 0 and 1 are read by the update (L6Geom.cpp:110-118
        // recomputes column 2 as their cross product), so only they are stored
        if (real) {
            for (int k = 0; k < 3; k++) {
                tr[3 * k] = *stAt(cv, row, S_T + k, nR);
                tr[3 * k + 1] = *stAt(cv, row, S_T + 3 + k, nR);
                tr[3 * k + 2] = (T)0;
            }
        }
        Geo<T> G;
        if ((GEOM & G_SS) && ty == CI_SS) {  // Cg2_Sphere_Sphere_L6Geom::go, pkg/dem/Sphere.cpp:88-109
            D3 p2 = xyz(pb) + shift2;
            D3 relPos = p2 - xyz(pa);
            double r1 = pa.w, r2 = pb.w;
            double df = fabs(a.distFactor) * (r1 + r2);
            double sq = sqn3(relPos);
            if (sq - df * df > 0 && !real) break;
            double dist = wsqrt(sq);
            double uN = dist - (r1 + r2);
            D3 normal = relPos / dist;
            G.normal = cvt3<T>(normal);
            G.contPt = xyz(pa) + normal * (r1 + 0.5 * uN);
            G.uN = (T)uN; G.r1 = (T)r1; G.r2 = (T)r2;
            G.pos1 = xyz(pa); G.pos2 = p2;
            double4 va = n.velm[id.x], vb = n.velm[id.y], wa = n.angI[id.x], wb = n.angI[id.y];
            G.vel1 = xyz(va); G.vel2 = xyz(vb); G.angVel1 = xyz(wa); G.angVel2 = xyz(wb);
        } else if ((GEOM & G_WS) && ty == CI_WS) {  // Cg2_Wall_Sphere_L6Geom::go, pkg/dem/Wall.cpp:45-79
            int32_t wt = p.ptype[id.x];
            int ax = ptAxis(wt), sn = ptSense(wt);
            double rad = pb.w;
            D3 wallPos = xyz(pa);
            D3 spherePos = xyz(pb) + shift2;
            double dist = comp3(spherePos, ax) - comp3(wallPos, ax);
            if (!real && fabs(dist) > rad) break;
            D3 contPt = spherePos;
            setcomp3(contPt, ax, comp3(wallPos, ax));
            D3 normal = mk3(0., 0., 0.);
            double nax;
            if (sn == 0) nax = real ? (double)tr[3 * ax] : (dist > 0 ? 1. : -1.);
            else nax = (sn == 1 ? 1. : -1);
            setcomp3(normal, ax, nax);
            double uN = nax * dist - rad;
            G.normal = cvt3<T>(normal);
            G.contPt = contPt;
            G.uN = (T)uN; G.r1 = (T)(-rad); G.r2 = (T)rad;
            G.pos1 = wallPos; G.pos2 = spherePos;
            double4 va = n.velm[id.x], vb = n.velm[id.y], wa = n.angI[id.x], wb = n.angI[id.y];
            G.vel1 = xyz(va); G.vel2 = xyz(vb); G.angVel1 = xyz(wa); G.angVel2 = xyz(wb);
        } else if (!(GEOM & G_FS)) {
            break;  // not in this scene (the host picks GEOM from the scene's shapes)
        } else {  // Cg2_Facet_Sphere_L6Geom::go, pkg/dem/Facet.cpp:235-288
            int4 fn = p.fnodes[id.x];
            double halfThick = pa.w, rad = pb.w;
            D3 n0 = xyz(n.posr[fn.x]), n1 = xyz(n.posr[fn.y]), n2 = xyz(n.posr[fn.z]);
            D3 sC = xyz(pb) + shift2;
            D3 nn = cross3(n1 - n0, n2 - n0);
            D3 fNormal = normalized3(nn);
            double planeDist = dot3(sC - n0, fNormal);
            if (fabs(planeDist) > (rad + halfThick) && !real) break;
            D3 fC = sC - fNormal * planeDist;
            D3 o0 = cross3(n1 - n0, nn), o1 = cross3(n2 - n1, nn), o2 = cross3(n0 - n2, nn);
            double l0 = dot3(o0, fC - n0), l1 = dot3(o1, fC - n1), l2 = dot3(o2, fC - n2);
            int w = (l0 > 0 ? 1 : 0) + (l1 > 0 ? 2 : 0) + (l2 > 0 ? 4 : 0);
            D3 contPt;
            auto seg = [](const D3& P, const D3& A, const D3& B) -> D3 {  // CompUtils::closestSegmentPt, CompUtils.cpp:57-63
                D3 BA = B - A;
                if (sqn3(BA) == 0.) return A;
                double u = (dot3(P, BA) - dot3(A, BA)) / sqn3(BA);
                return A + BA * fmin(1., fmax(0., u));
            };
            if (w == 0) contPt = fC;
            else if (w == 1) contPt = seg(fC, n0, n1);
            else if (w == 2) contPt = seg(fC, n1, n2);
            else if (w == 4) contPt = seg(fC, n2, n0);
            else if (w == 3) contPt = n1;
            else if (w == 5) contPt = n0;
            else if (w == 6) contPt = n2;
            else { atomicOr(err, ERR_NAN); break; }
            D3 normal = sC - contPt;
            double rs = rad + halfThick;
            if (sqn3(normal) > rs * rs && !real) break;
            double dist = wsqrt(sqn3(normal));
            normal = normal / dist;
            if (halfThick > 0) contPt = contPt + normal * halfThick;
            double uN = dist - rad - halfThick;
            // Facet::interpolatePtLinAngVel (Facet.cpp:158-174) with CompUtils::triangleBarycentrics (CompUtils.cpp:92-110)
            D3 vB = n1 - n0, vC = n2 - n0, vX = contPt - n0;
            double m00 = dot3(vB, vB), m01 = dot3(vC, vB), m10 = dot3(vB, vC), m11 = dot3(vC, vC);
            double invdet = 1. / (m00 * m11 - m10 * m01);
            double i00 = m11 * invdet, i01 = -m01 * invdet, i10 = -m10 * invdet, i11 = m00 * invdet;
            double rr0 = dot3(vX, vB), rr1 = dot3(vX, vC);
            double u = i00 * rr0 + i01 * rr1, v = i10 * rr0 + i11 * rr1;
            double b0 = 1 - u - v, b1 = u, b2 = v;
            D3 v0 = xyz(n.velm[fn.x]), v1 = xyz(n.velm[fn.y]), v2 = xyz(n.velm[fn.z]);
            D3 linVel = (v0 * b0 + v1 * b1) + v2 * b2;
            D3 av = (cross3(n0 - contPt, v0) + cross3(n1 - contPt, v1)) + cross3(n2 - contPt, v2);
            G.normal = cvt3<T>(normal);
            G.contPt = contPt;
            G.uN = (T)uN; G.r1 = (T)fmax(halfThick, rad); G.r2 = (T)rad;
            G.pos1 = contPt; G.pos2 = sC;
            G.vel1 = linVel; G.angVel1 = av;
            double4 vb = n.velm[id.y], wb = n.angI[id.y];
            G.vel2 = xyz(vb); G.angVel2 = xyz(wb);
        }
        // handleSpheresLikeContact, pkg/dem/L6Geom.cpp:51-217
        V3T<T> gvel, gangVel;
        T lens0 = 0, lens1 = 0, contA = 0;
        if (!real) {
            initialLocalCoords(G.normal, tr);
            lens0 = fabs(G.r1) + G.uN / 2;
            lens1 = fabs(G.r2) + G.uN / 2;
            if (lens0 < 0) lens0 = lens1;
            if (lens1 < 0) lens1 = lens0;
            T rr = (G.r1 > 0 && G.r2 > 0) ? fmin(G.r1, G.r2) : (G.r1 > 0 ? G.r1 : G.r2);
            contA = (T)3.14159265358979323846 * (rr * rr);
            gvel = mk3<T>(0, 0, 0);
            gangVel = mk3<T>(0, 0, 0);
        } else {
            const V3T<T> currNormal = G.normal, prevNormal = trCol(tr, 0);
            D3 shiftVel2 = mk3(0., 0., 0.);
            if (a.periodic) shiftVel2 = mv3(ldM3(a.cell->shiftVelM), mk3((double)unpack10(cdp, 0), (double)unpack10(cdp, 10), (double)unpack10(cdp, 20)));
            D3 prevContPt;
            if constexpr (sizeof(T) == 8) prevContPt = xyz(cv.cpt[row]);
            else {
                // B's position at the last ContactLoop (Leapfrog moved it by its velocity times
                // dt since) plus the branch stored then
                const float4 r = mixedCpt(cv.cpt)[row];
                prevContPt = (G.pos2 - (G.vel2 + shiftVel2) * dt) + mk3((double)r.x, (double)r.y, (double)r.z);
            }
            V3T<T> normRotVec = cross3(prevNormal, currNormal);
            V3T<T> midNormal = (prevNormal + currNormal) * (T).5;
            if (wsqrt(sqn3(midNormal)) == 0) midNormal = currNormal;
            else midNormal = normalized3(midNormal);
            V3T<T> normTwistVec = ((midNormal * (T)dt) * (T).5) * dot3(midNormal, cvt3<T>(G.angVel1 + G.angVel2));
            V3T<T> rt = normRotVec + normTwistVec;
            V3T<T> pc1 = trCol(tr, 1);
            V3T<T> mid1 = pc1 - cross3(pc1, rt) / (T)2.;
            V3T<T> mid2 = cross3(midNormal, mid1);
            V3T<T> cur1 = pc1 - cross3(mid1, rt);
            V3T<T> cur0 = currNormal;
            if (a.trsfRenorm > 0 && (a.step % a.trsfRenorm) == 0) {
                cur0 = normalized3(cur0);
                cur1 = cur1 - cur0 * dot3(cur1, cur0);
                cur1 = normalized3(cur1);
            }
            V3T<T> cur2 = cross3(cur0, cur1);
            D3 midContPt = (prevContPt + G.contPt) * .5;
            D3 midPos1 = G.pos1 - G.vel1 * (dt / 2.);
            D3 midPos2 = G.pos2 - (G.vel2 + shiftVel2) * (dt / 2.);
            D3 c1x = midContPt - midPos1, c2x = midContPt - midPos2;
            V3T<T> relVel;
            if constexpr (sizeof(T) == 8) {
                relVel = ((G.vel2 + shiftVel2) + cross3(G.angVel2, c2x)) - (G.vel1 + cross3(G.angVel1, c1x));
            } else {
                // mixed precision: translational difference in FP64, rotational terms in FP32
                V3T<T> dv = cvt3<T>((G.vel2 + shiftVel2) - G.vel1);
                relVel = dv + (cross3(cvt3<T>(G.angVel2), cvt3<T>(c2x)) - cross3(cvt3<T>(G.angVel1), cvt3<T>(c1x)));
            }
            trSetCol(tr, 0, cur0);
            trSetCol(tr, 1, cur1);
            trSetCol(tr, 2, cur2);
            // midTrsf^T * v: columns (midNormal, mid1, mid2) dotted with v
            gvel = mk3<T>(dot3(midNormal, relVel), dot3(mid1, relVel), dot3(mid2, relVel));
            V3T<T> dw = cvt3<T>(G.angVel2 - G.angVel1);
            gangVel = mk3<T>(dot3(midNormal, dw), dot3(mid1, dw), dot3(mid2, dw));
        }
        // physics (on creation) and law
        T F[3];
        T ph[5];
        bool keep = true;
        const T uN = G.uN;
        if (!real) {
            double4 m1 = a.mat[p.mat[id.x]], m2 = a.mat[p.mat[id.y]];
            if (LAW == 0) {  // Cp2_FrictMat_FrictPhys::updateFrictPhys, pkg/dem/FrictMat.cpp:21-44
                T l0 = lens0, l1 = lens1, A = contA;
                T kn = (T)1 / ((T)1 / ((T)m1.y * A / l0) + (T)1 / ((T)m2.y * A / l1));
                ph[0] = kn;
                ph[1] = (T).5 * ((T)m1.z + (T)m2.z) * kn;
                ph[2] = fmin((T)m1.w, (T)m2.w);
                ph[3] = 0; ph[4] = 0;
            } else {  // Cp2_FrictMat_HertzPhys::go, pkg/dem/Hertz.cpp:9-50
                T r1 = lens0, r2 = lens1;
                T E1 = (T)m1.y, E2 = (T)m2.y, nu1 = (T)a.poisson, nu2 = (T)a.poisson;
                T K = (T)(4 / 3.) * (T)1. / (((T)1 - nu1 * nu1) / E1 + ((T)1 - nu2 * nu2) / E2);
                T R = (T)1. / ((T)1. / r1 + (T)1. / r2);
                T kn0 = K * wsqrt(R);
                T G1 = E1 / ((T)2 * ((T)1 + nu1)), G2 = E2 / ((T)2 * ((T)1 + nu2));
                T Gm = (T).5 * (G1 + G2);
                T nu = (T).5 * (nu1 + nu2);
                ph[0] = K; ph[1] = R;
                ph[2] = fmin((T)m1.w, (T)m2.w);
                ph[3] = (T)2 * wsqrt((T)4 * R) * Gm / ((T)2 - nu);
                T asq = 0;
                if (a.hertzDamped) {
                    double ma = n.velm[id.x].w, mb = n.velm[id.y].w;
                    double mbar = (ma <= 0 && mb > 0) ? mb : ((ma > 0 && mb <= 0) ? ma : (ma * mb) / (ma + mb));
                    asq = fmax((T)0., (T)a.viscAlpha * wsqrt((T)mbar * kn0));
                }
                ph[4] = asq;
            }
            F[0] = F[1] = F[2] = 0;
        } else {
            for (int k = 0; k < 5; k++) ph[k] = *stAt(cv, row, S_P + k, nR);
            for (int k = 0; k < 3; k++) F[k] = *stAt(cv, row, S_F + k, nR);
        }
        if (LAW == 0) {  // Law2_L6Geom_FrictPhys_IdealElPl::go, pkg/dem/IdealElPl.cpp:12-122
            const T kn = ph[0], kt = ph[1], tanPhi = ph[2];
            if (uN > 0 && !a.noBreak) {
                if (a.noSlip && a.trackEnergy) {
                    T Fn = kn * uN;
                    T Ft0 = F[1] + (T)dt * kt * gvel.y, Ft1 = F[2] + (T)dt * kt * gvel.z;
                    ev[4] = (double)((T).5 * (Fn * Fn / kn + (Ft0 * Ft0 + Ft1 * Ft1) / kt));
                    bits |= 1u << E_BROKEN;
                }
                keep = false;
            } else {
                F[0] = kn * uN;
                if (a.noFrict || tanPhi == 0.) { F[1] = 0; F[2] = 0; }
                else {
                    F[1] += (T)dt * kt * gvel.y;
                    F[2] += (T)dt * kt * gvel.z;
                    T maxFt = fabs(F[0]) * tanPhi;
                    T ftsq = F[1] * F[1] + F[2] * F[2];
                    if (ftsq > maxFt * maxFt && !a.noSlip) {
                        T FtNorm = wsqrt(ftsq);
                        T ratio = maxFt / FtNorm;
                        if (FtNorm == (T)0.) ratio = 0.;
                        if (a.trackEnergy) {
                            T dissip = (kt != (T)0. ? (maxFt) * (FtNorm - maxFt) / kt : (T)0.);
                            if (dissip != (T)0.) { ev[1] = (double)dissip; bits |= 1u << E_PLAST; }
                        }
                        F[1] *= ratio; F[2] *= ratio;
                    }
                    if (isnan(F[0]) || isnan(F[1]) || isnan(F[2])) atomicOr(err, ERR_NAN);
                }
                if (a.trackEnergy) {
                    T ftsq = F[1] * F[1] + F[2] * F[2];
                    T el = (T)0.5 * ((F[0] * F[0]) / kn + (kt != (T)0. ? ftsq / kt : (T)0.));
                    if (isnan(el)) el = 0.;
                    ev[0] = (double)el;
                    bits |= 1u << E_ELAST;
                }
            }
        } else {  // Law2_L6Geom_HertzPhys_DMT::go, pkg/dem/Hertz.cpp:58-182, alpha==0 branch
            const T K = ph[0], R = ph[1], tanPhi = ph[2], kt0 = ph[3], asq = ph[4];
            const T velN = gvel.x;
            T kn0 = K * wsqrt(R);
            if (uN > 0) keep = false;
            else {
                T s = wsqrt(-uN);
                T Fne = -kn0 * ((s * s) * s);
                T Fna = (T)(4 * 3.14159265358979323846) * R * (T)a.gamma;
                T eta = (asq > 0 ? asq * wsqrt(wsqrt(fmax(-uN, (T)0.))) : (T)0.);
                T Fnv = eta * velN;
                if (a.noAttraction && Fne + Fnv > 0) Fnv = -Fne;
                T Fn = Fne + Fna + Fnv;
                F[0] = Fn;
                if (a.trackEnergy) { T v = Fnv * velN * (T)dt; if (v != (T)0.) { ev[2] = (double)v; bits |= 1u << E_VISCN; } }
                T kt = kt0 * wsqrt(uN < 0 ? -uN : (T)0);
                T Ft0 = (T)dt * kt * gvel.y, Ft1 = (T)dt * kt * gvel.z;
                T maxFt = fmax((T)0., Fn) * tanPhi;
                T ftsq = Ft0 * Ft0 + Ft1 * Ft1;
                if (ftsq > maxFt * maxFt) {
                    T FtNorm = wsqrt(ftsq);
                    T ratio = maxFt / FtNorm;
                    if (a.trackEnergy) {
                        T v = ((T).5 * (FtNorm - maxFt) + maxFt) * (FtNorm - maxFt) / kt;
                        if (v != (T)0.) { ev[1] = (double)v; bits |= 1u << E_PLAST; }
                    }
                    Ft0 *= ratio; Ft1 *= ratio;
                } else {
                    Ft0 += eta * gvel.y; Ft1 += eta * gvel.z;
                    if (a.trackEnergy) { T v = eta * (gvel.y * gvel.y + gvel.z * gvel.z) * (T)dt; if (v != (T)0.) { ev[3] = (double)v; bits |= 1u << E_VISCT; } }
                }
                F[1] = Ft0; F[2] = Ft1;
                if (a.trackEnergy && uN < 0) {
                    T s5 = (((s * s) * s) * s) * s;
                    T el = kn0 * (T)(2 / 5.) * s5 + (T)0.5 * (Ft0 * Ft0 + Ft1 * Ft1) / kt;
                    if (!isnan(el)) ev[0] = (double)el;
                    bits |= 1u << E_ELAST;
                }
            }
        }
        if (!keep) {  // ContactContainer::requestRemoval -> Contact::reset (ContactContainer.cpp:164-171)
            const uint32_t ni = (info & ~CI_REAL) | CI_PENDING;
            cv.info[c] = ni;
            *infoItem = (int32_t)ni;
            break;
        }
        // store geometry/physics state; every item owns the state row of its own index, so
        // contacts becoming real between full runs read and write their state contiguously
        if (!real) row = aux.x;
        for (int k = 0; k < 3; k++) {
            *stAt(cv, row, S_T + k, nR) = tr[3 * k];
            *stAt(cv, row, S_T + 3 + k, nR) = tr[3 * k + 1];
        }
        for (int k = 0; k < 3; k++) *stAt(cv, row, S_F + k, nR) = F[k];
        if (!real) {
            for (int k = 0; k < 5; k++) *stAt(cv, row, S_P + k, nR) = ph[k];
            *stAt(cv, row, S_L, nR) = lens0;
            *stAt(cv, row, S_L + 1, nR) = lens1;
            cv.info[c] = info | CI_REAL;
            *infoItem = (int32_t)(info | CI_REAL);
        }
        // global force trsf*F (Contact::getForceTorqueBranch, Contact.cpp:24-34)
        V3T<T> fl = mk3<T>(F[0], F[1], F[2]);
        V3T<T> fg = mk3<T>(dot3(mk3<T>(tr[0], tr[1], tr[2]), fl), dot3(mk3<T>(tr[3], tr[4], tr[5]), fl), dot3(mk3<T>(tr[6], tr[7], tr[8]), fl));
        if constexpr (sizeof(T) == 8) {
            cv.gfc[2 * item] = make_double4((double)fg.x, (double)fg.y, (double)fg.z, (double)cdp);
            cv.gfc[2 * item + 1] = make_double4(G.contPt.x, G.contPt.y, G.contPt.z, (double)uN);
            cv.cpt[row] = make_double4(G.contPt.x, G.contPt.y, G.contPt.z, (double)uN);
        } else {
            const D3 bA = G.contPt - G.pos1, bB = G.contPt - G.pos2;
            mixedRecA(cv.gfc)[item] = make_float4(fg.x, fg.y, fg.z, (float)bA.x);
            mixedRecB(cv.gfc, cv.cap)[item] = make_float4((float)bA.y, (float)bA.z, (float)bB.x, (float)bB.y);
            mixedRecC(cv.gfc, cv.cap)[item] = (float)bB.z;
            mixedCpt(cv.cpt)[row] = make_float4((float)bB.x, (float)bB.y, (float)bB.z, (float)uN);
        }
        realOut = true;
    } while (0);
    // one bit per item, one 64-bit word per wavefront: the gather skips non-real contacts
    uint64_t bal = __ballot(realOut);
    int64_t w0 = lb * WDM_BLOCK + (threadIdx.x & ~63);
    if ((threadIdx.x & 63) == 0 && w0 < nItems) realMask[w0 >> 6] = bal;
    if (a.trackEnergy) blockSumWrite<5>(ev, bits, epart, nullptr, 0u, creOut, blk);
}

// Mixed precision: the contact points of a list's live real contacts between their stored
// form (branch of B in A's image, FP32) and absolute FP64 points (DIR 0: to abs[row], DIR 1:
// from abs[row]), with B's position and the cell's hSize (hs; between steps both as at the
// last ContactLoop)
template <int DIR>
__global__ void __launch_bounds__(WDM_BLOCK) k_cpt_mixed(const double4* posr, ConView<float> cv, int64_t n, int periodic, M3 hs,
                                                        double4* abs) {
    const int64_t i = blockIdx.x * (int64_t)WDM_BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint32_t info = cv.info[i];
    const int32_t row = cv.slot[i];
    if (!(info & CI_REAL) || (info & CI_DEAD) || row < 0 || row >= cv.cap) return;
    const int2 id = cv.ids[i];
    D3 p2 = xyz(posr[id.y]);
    if (periodic) {
        const int32_t cdp = cv.cd[i];
        p2 = p2 + mv3(hs, mk3((double)unpack10(cdp, 0), (double)unpack10(cdp, 10), (double)unpack10(cdp, 20)));
    }
    if (DIR == 0) {
        const float4 r = mixedCpt(cv.cpt)[row];
        abs[row] = make_double4(p2.x + (double)r.x, p2.y + (double)r.y, p2.z + (double)r.z, (double)r.w);
    } else {
        const double4 q = abs[row];
        mixedCpt(cv.cpt)[row] = make_float4((float)(q.x - p2.x), (float)(q.y - p2.y), (float)(q.z - p2.z), (float)q.w);
    }
}

// ---------------- DynDt, pkg/dem/DynDt.cpp:16-70 ----------------
// nodalStiffAdd (:16-36) for the contacts of particle q: every real contact adds
// n2*(kn-kt)+kt to the translational and lens^2*kt*(n2y+n2z, n2z+n2x, n2x+n2y) to the
// rotational stiffness, n2 the squared components of the contact normal (trsf column 0,
// which the contact node's orientation rotates the x axis to) and lens the L6Geom branch of
// q's side.  Contacts come in contact-map order (other particle's id ascending: the CSR
// order).  HertzPhys stiffnesses are those of its last law call (Hertz.cpp:69,157).
template <typename T, int LAW>
__device__ __forceinline__ void stiffAdd(const ConView<T>& cv, int64_t nR, const GatherView& gv, int q, double (&kt)[3],
                                         double (&kr)[3]) {
    const int64_t e0 = gv.off[q], e1 = gv.off[q + 1];
    for (int64_t e = e0; e < e1; e++) {
        const uint32_t v = gv.ent[e];
        const uint32_t c = v & 0x7fffffffu;  // work-list item, which owns state row c
        if (!((gv.realMask[c >> 6] >> (c & 63)) & 1ull)) continue;  // !C->isReal()
        const int ix = (int)(v >> 31);
        const double nx = (double)*stAt(cv, c, S_T + 0, nR), ny = (double)*stAt(cv, c, S_T + 1, nR), nz = (double)*stAt(cv, c, S_T + 2, nR);
        double kn, kT;
        if (LAW == 0) {
            kn = (double)*stAt(cv, c, S_P + 0, nR);
            kT = (double)*stAt(cv, c, S_P + 1, nR);
        } else {
            double uN;
            if constexpr (sizeof(T) == 8) uN = gv.gfc[2 * (size_t)c + 1].w;
            else uN = (double)mixedCpt(cv.cpt)[c].w;
            const double kn0 = (double)*stAt(cv, c, S_P + 0, nR) * sqrt((double)*stAt(cv, c, S_P + 1, nR));
            kn = (3 / 2.) * kn0 * sqrt(-uN);
            kT = (double)*stAt(cv, c, S_P + 3, nR) * sqrt(uN < 0 ? -uN : 0.);
        }
        const double l = (double)*stAt(cv, c, S_L + ix, nR);
        const double n2x = nx * nx, n2y = ny * ny, n2z = nz * nz;
        kt[0] = kt[0] + (n2x * (kn - kT) + kT);
        kt[1] = kt[1] + (n2y * (kn - kT) + kT);
        kt[2] = kt[2] + (n2z * (kn - kT) + kT);
        const double lk = (l * l) * kT;
        kr[0] = kr[0] + lk * (n2y + n2z);
        kr[1] = kr[1] + lk * (n2z + n2x);
        kr[2] = kr[2] + lk * (n2x + n2y);
    }
}
// Per node of DemField.nodes (clump members are not among them; the dummy node of a facet
// particle is no node of the reference): nodalCritDtSq (:39-58), the squared critical step
// 2 min(m/|k_t,i|, I/|k_r,i|) over the free DOFs, infinite for a node with every DOF
// blocked; a clump master adds its members' contacts after its own (:48-51).  The scene's
// minimum (critDt_stiffness, :61-70) is an exact min, reduced per block and then with an
// atomic min on the bit pattern (non-negative doubles order like their bits), so it does
// not depend on the order.  out[t] must hold +inf before the launch.
template <typename T, int LAW>
__global__ void __launch_bounds__(WDM_BLOCK) k_dyndt(const SceneDesc* D, const ConDesc<T>* C, Map m, unsigned long long* out) {
    int64_t lb;
    const int t = mapBlock(m, lb);
    const ConDesc<T>& CD = C[t];
    const SceneDesc& S = D[CD.desc];
    const NodeView n = S.n;
    const PartView p = S.p;
    const GatherView gv = S.gv;
    const ClumpView cl = S.cl;
    const ConView<T> cv = CD.cv;
    const int64_t nR = *cv.counter;
    const int i = (int)(lb * WDM_BLOCK + threadIdx.x);
    double r = INFINITY;
    if (i < n.nn) {
        const uint32_t fl = n.flags[i];
        const int sh = i < p.np ? ptShape(p.ptype[i]) : 0;
        const bool own = sh == 1 || sh == 2;  // uninodal particle on this node
        const bool isNode = !(fl & F_CLUMPED) && (own || i >= p.np);
        if (isNode && (fl & F_DOF_ALL) != F_DOF_ALL) {
            double kt[3] = {0., 0., 0.}, kr[3] = {0., 0., 0.};
            if (own) stiffAdd<T, LAW>(cv, nR, gv, i, kt, kr);
            if ((fl & F_CLUMP) && cl.nodeClump) {
                const int c = cl.nodeClump[i];
                for (int k = cl.off[c]; k < cl.off[c + 1]; k++) {
                    const int q = cl.mem[k];
                    if (q < p.np) {
                        const int sq = ptShape(p.ptype[q]);
                        if (sq == 1 || sq == 2) stiffAdd<T, LAW>(cv, nR, gv, q, kt, kr);
                    }
                }
            }
            const double mass = n.velm[i].w, inertia = n.angI[i].w;
            double ret = INFINITY;
            for (int a = 0; a < 3; a++)
                if (kt[a] != 0 && mass > 0. && !(fl & (1u << a))) { const double x = mass / fabs(kt[a]); if (x < ret) ret = x; }
            for (int a = 0; a < 3; a++)
                if (kr[a] != 0 && inertia > 0. && !(fl & (8u << a))) { const double x = inertia / fabs(kr[a]); if (x < ret) ret = x; }
            r = 2 * ret;
        }
    }
    __shared__ double sm[WDM_BLOCK];
    sm[threadIdx.x] = r;
    __syncthreads();
    for (int s = WDM_BLOCK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { const double o = sm[threadIdx.x + s]; if (o < sm[threadIdx.x]) sm[threadIdx.x] = o; }
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMin(out + t, (unsigned long long)__double_as_longlong(sm[0]));
}
__global__ void k_fill_u64(unsigned long long* a, int n, unsigned long long v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------- particles removed and inserted in place ----------------
// A removed particle (DemField::removeParticle, pkg/dem/Particle.cpp:445-493) keeps its slot as
// a hole: no shape, its node fully blocked, massless, at rest and skipped by the energy terms,
// so Leapfrog changes nothing for it; its contacts are dropped.  ParticleContainer reuses the
// holes first (ParticleContainer.cpp:23-72), which wdm_insert_spheres does.
// Kernel parameters are plain pointers: an address-space-qualified parameter type is mangled
// differently in the device pass, so the host would launch a name the code object lacks.
#define F_INERT (F_DOF_ALL | F_ENERGY_SKIP | F_GRAVITY_SKIP | F_DAMPING_SKIP)
__global__ void k_remove_particles(int32_t* ptype, uint32_t* flags, double4* velm, double4* angI, uint8_t* checked, const int32_t* ids,
                                   int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = ids[k];
    ptype[i] = 0;
    flags[i] = F_INERT;
    velm[i] = make_double4(0., 0., 0., 0.);
    angI[i] = make_double4(0., 0., 0., 0.);
    checked[i] = 0;
}
// contacts with a removed particle (ContactContainer::remove, ContactContainer.cpp:124-150)
__global__ void k_kill_contacts(uint32_t* info, const int2* ids, int64_t nc, const int32_t* ptype) {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int2 id = ids[c];
    if (ptShape(ptype[id.x]) == 0 || ptShape(ptype[id.y]) == 0) info[c] = CI_DEAD;
}
// a new sphere in slot `id`: node at rest in its given state, force and torque zero (a new
// DemData's; gravity enters through the force reset of its first Leapfrog, Leapfrog.cpp:240-245)
struct InsRec {
    double4 posr, velm, angI;
    int32_t id, mat;
    uint32_t mask;
    int32_t pad;
};
__global__ void k_insert_spheres(NodeView n, PartView p, const InsRec* rec, int cnt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    const InsRec r = rec[k];
    const int i = r.id;
    n.posr[i] = r.posr;
    n.velm[i] = r.velm;
    n.angI[i] = r.angI;
    n.ori[i] = make_double4(1., 0., 0., 0.);
    n.force[i] = make_double4(0., 0., 0., 0.);
    n.torque[i] = make_double4(0., 0., 0., 0.);
    n.lastPos[i] = make_double4(r.posr.x, r.posr.y, r.posr.z, 0.);
    n.flags[i] = 0u;
    n.checked[i] = 1;
    p.ptype[i] = 1;  // Sphere
    p.mask[i] = r.mask;
    p.mat[i] = r.mat;
    p.fnodes[i] = make_int4(i, -1, -1, 0);
}
__global__ void k_set_masks(uint32_t* pmask, const int32_t* ids, const uint32_t* masks, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) pmask[ids[k]] = masks[k];
}
// BoxOutlet::isInside (pkg/dem/Outlet.hpp:56, AlignedBox3r::contains: closed box) and the
// particle tests of Outlet::run (Outlet.cpp:16-40) for uninodal particles: 1 when the
// particle is to be deleted (markMask 0) or marked
__device__ __forceinline__ bool boxSelected(const double4* posr, const int32_t* ptype, const uint32_t* pmask, int i, double3 lo,
                                            double3 hi, int inside, uint32_t mask, uint32_t markMask) {
    const int sh = ptShape(ptype[i]);
    if (!(sh == 1 || sh == 2)) return false;
    const double4 q = posr[i];
    const bool in = q.x >= lo.x && q.x <= hi.x && q.y >= lo.y && q.y <= hi.y && q.z >= lo.z && q.z <= hi.z;
    const uint32_t m = pmask[i];
    return (inside != 0) == in && !(mask && !(mask & m)) && !(markMask != 0 && (markMask & m) == markMask);
}
__global__ void k_select_box(const double4* posr, const int32_t* ptype, const uint32_t* pmask, int np, double3 lo, double3 hi,
                             int inside, uint32_t mask, uint32_t markMask, uint8_t* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    out[i] = boxSelected(posr, ptype, pmask, i, lo, hi, inside, mask, markMask) ? 1 : 0;
}
// The same selection compacted on the device (ascending ids, no host pass): per block of
// WDM_BLOCK particles the number selected, one workgroup's exclusive scan of those counts,
// then each block writes its ids at its offset in thread order.
__global__ void __launch_bounds__(WDM_BLOCK) k_select_count(const double4* posr, const int32_t* ptype, const uint32_t* pmask, int np,
                                                           double3 lo, double3 hi, int inside, uint32_t mask, uint32_t markMask,
                                                           int32_t* counts) {
    const int i = blockIdx.x * WDM_BLOCK + threadIdx.x;
    const bool s = i < np && boxSelected(posr, ptype, pmask, i, lo, hi, inside, mask, markMask);
    const int c = __syncthreads_count(s ? 1 : 0);
    if (threadIdx.x == 0) counts[blockIdx.x] = c;
}
// one workgroup of 1024 threads: offs[b] = sum of counts[0..b), *total = sum of all (nb <= 1024 * SCAN_PER_THREAD)
#define SCAN_PER_THREAD 16
__global__ void __launch_bounds__(1024) k_select_scan(const int32_t* counts, int nb, int32_t* offs, int32_t* total) {
    __shared__ int32_t part[1024];
    const int t = threadIdx.x, b0 = t * SCAN_PER_THREAD;
    int32_t v[SCAN_PER_THREAD], s = 0;
    for (int k = 0; k < SCAN_PER_THREAD; k++) {
        v[k] = b0 + k < nb ? counts[b0 + k] : 0;
        s += v[k];
    }
    part[t] = s;
    __syncthreads();
    for (int w = 1; w < 1024; w <<= 1) {  // inclusive Hillis-Steele scan of the thread sums
        const int32_t add = t >= w ? part[t - w] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    int32_t o = part[t] - s;  // exclusive prefix of this thread's first block
    for (int k = 0; k < SCAN_PER_THREAD; k++) {
        if (b0 + k < nb) offs[b0 + k] = o;
        o += v[k];
    }
    if (t == 1023) *total = part[t];
}
__global__ void __launch_bounds__(WDM_BLOCK) k_select_write(const double4* posr, const int32_t* ptype, const uint32_t* pmask, int np,
                                                           double3 lo, double3 hi, int inside, uint32_t mask, uint32_t markMask,
                                                           const int32_t* offs, int32_t* ids) {
    __shared__ int32_t waveBase[WDM_BLOCK / 64];
    const int i = blockIdx.x * WDM_BLOCK + threadIdx.x;
    const bool s = i < np && boxSelected(posr, ptype, pmask, i, lo, hi, inside, mask, markMask);
    const uint64_t bal = __ballot(s);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) waveBase[wv] = __popcll(bal);
    __syncthreads();
    int base = offs[blockIdx.x];
    for (int k = 0; k < wv; k++) base += waveBase[k];
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
    if (s) ids[base + rank] = i;
}

// ---------------- SpherePack::makeCloud (lib/sphere-pack/SpherePack.cpp:135-248) ----------------
// Random loose packing, sphere after sphere in the reference's order: each sphere tries up
// to maxTry positions drawn from the generator and takes the first that overlaps no sphere
// placed before it, (rj + r)^2 >= |cj - c|^2 being an overlap; the periodic distance is the
// smaller of the two wrapped differences per axis (SpherePack::cellWrapRel, SpherePack.hpp:32-35).
// The generator is boost::minstd_rand (Park-Miller, x <- 48271 x mod 2^31-1, values 1..2^31-2)
// read through boost::uniform_real(0,1): u = (x-1)/(2^31-3), redrawn when u reaches 1.  One
// workgroup runs the sequence: a sphere's candidate is tested by the threads against the
// spheres in the 27 grid cells around it (cells at least 2 rmax wide, so they hold every
// sphere that can overlap it), which gives the same decisions as testing all placed spheres.
struct CloudArgs {
    double mn[3], size[3];
    double h[3];      // grid cell widths
    int nc[3];        // grid cells per axis
    int periodic;
    int num;          // > 0: radii[i] for sphere i (the reference's deterministic sequence); < 0: drawn
    const double* radii;
    // radius rule for num < 0 (SpherePack.cpp:191-209): mode 0 uniform/pow3 in rMean(1 +- rRelFuzz), 1 PSD
    int mode, distributeMass, nPsd;
    double rMean, rRelFuzz;
    const double* psdR;   // PSD radii
    const double* psdC;   // cumulative (psdCumm, or psdCumm2 with distributeMass)
    // pow(r, -2) of the interpolation ends, from the host's pow: mode 0 rMean(1 -+ rRelFuzz),
    // mode 1 one per PSD radius
    double m2lo, m2hi;
    const double* psdRm2;
    int64_t cap;      // spheres the output holds
    int maxTry;
};
WDM_HD uint32_t mstdNext(uint32_t x) { return (uint32_t)(((uint64_t)x * 48271u) % 2147483647u); }
WDM_HD double mstdUniform(uint32_t& x) {
    for (;;) {
        x = mstdNext(x);
        const double u = (double)(x - 1u) / 2147483645.;
        if (u < 1.) return u;
    }
}
// SpherePack::pow3Interp (SpherePack.hpp:86)
WDM_HD double pow3Interp(double x, double a, double b) { return pow(x * (pow(b, -2.) - pow(a, -2.)) + pow(a, -2.), -1. / 2); }
// SpherePack::psdGetPiece (SpherePack.cpp:283-291)
WDM_HD int psdGetPiece(double x, const double* cumm, int sz, double& norm) {
    int i = 0;
    while (i < sz && cumm[i] <= x) i++;
    if ((i == sz - 1) && cumm[i] <= x) { norm = 1.; return sz - 2; }
    i--;
    norm = (x - cumm[i]) / (cumm[i + 1] - cumm[i]);
    return i;
}
// Sign of 1 - (y + h)^2 t, exact: the products split into exact pairs with fma, the terms
// summed with error-free additions (the leading 1 - y*y*t is exact by Sterbenz' lemma, y*y*t
// being near 1); the remaining rounding lies far below how close t^(-1/2) can come to a
// midpoint of two doubles, so the sign is right.
__device__ __forceinline__ int rsqrtMidSign(double y, double h, double t) {
    const double p = y * y, pe = fma(y, y, -p);
    const double s1 = p * t, s1e = fma(p, t, -s1);
    const double s2 = pe * t, s2e = fma(pe, t, -s2);
    const double v = y * t, ve = fma(y, t, -v);
    double hi = 1. - s1, lo = 0.;
    const double terms[6] = {-s1e, -s2, -2. * h * v, -s2e, -2. * h * ve, -(h * h) * t};
    for (int k = 0; k < 6; k++) {
        const double x = terms[k], s = hi + x, bp = s - hi;
        lo += (hi - (s - bp)) + (x - bp);
        hi = s;
    }
    const double r = hi + lo;
    return r > 0. ? 1 : (r < 0. ? -1 : 0);
}
// t^(-1/2) correctly rounded (t positive, finite, normal): 1/sqrt(t) is within two ulps; each
// step moves to the neighbour while t^(-1/2) lies beyond the midpoint toward it.
__device__ double rsqrtRN(double t) {
    double y = 1. / sqrt(t);
    for (int k = 0; k < 4; k++) {
        const double yu = __longlong_as_double(__double_as_longlong(y) + 1);
        if (rsqrtMidSign(y, (yu - y) * .5, t) > 0) { y = yu; continue; }
        const double yd = __longlong_as_double(__double_as_longlong(y) - 1);
        if (rsqrtMidSign(yd, (y - yd) * .5, t) < 0) { y = yd; continue; }
        break;
    }
    return y;
}
// pow3Interp on the device: pow(x * (b^-2 - a^-2) + a^-2, -1/2) with a^-2, b^-2 from the host's
// pow and the outer power correctly rounded; the reference's libm pow misrounds it now and then
// (glibc: about 4 in 10^4 draws, by one ulp), so draws agree with it to an ulp.
__device__ __forceinline__ double pow3InterpM2(double x, double am2, double bm2) { return rsqrtRN(x * (bm2 - am2) + am2); }
__device__ double cloudRadius(const CloudArgs& A, double rnd) {
    if (A.mode == 0) return A.distributeMass ? pow3InterpM2(rnd, A.m2lo, A.m2hi) : A.rMean * (2 * (rnd - .5) * A.rRelFuzz + 1);
    double norm;
    const int piece = psdGetPiece(rnd, A.psdC, A.nPsd, norm);
    return A.distributeMass ? pow3InterpM2(norm, A.psdRm2[piece], A.psdRm2[piece + 1])
                            : A.psdR[piece] + norm * (A.psdR[piece + 1] - A.psdR[piece]);
}
__device__ __forceinline__ double cloudWrapRel(double x, double x0, double x1) {
    const double xNorm = (x - x0) / (x1 - x0);
    return (xNorm - floor(xNorm)) * (x1 - x0);
}
__device__ __forceinline__ int cloudCell(const CloudArgs& A, int d, double x) {
    int c = (int)floor((x - A.mn[d]) / A.h[d]);
    return c < 0 ? 0 : (c >= A.nc[d] ? A.nc[d] - 1 : c);
}
// Placement (SpherePack.cpp:211-246) evaluated speculatively, one window of tries at a time.
// The reference tries candidates one after the other: try t of sphere i takes the next three
// draws (for num < 0 a sphere first takes one draw for its radius), and is accepted when it
// overlaps no sphere placed before it.  Which sphere a try belongs to depends on how many
// earlier tries were accepted.  A window fixes the draws of its next tries by jumping the
// generator ahead (x_k = x_0 48271^k mod 2^31-1) and evaluates, in parallel, each try under a
// few hypotheses on that count:
//   num > 0: try j of the window with a = lo_j .. lo_j+3 acceptances before it (sphere i+a,
//            radius radii[i+a], draws 3j..3j+2), lo_j following the last window's rate;
//   num < 0: try j with none accepted before it (sphere i, draws s0+3j..) and with all
//            accepted before it (sphere i+j, radius draw s0+4j-1, draws s0+4j..).
// Each candidate is tested against the cell grid of the spheres placed before the window by 9
// threads (one row of 3 neighbour cells each), and against every candidate of an earlier try
// of the window (a bit mask per candidate).  One wave then walks the tries in order, exactly as
// the reference does, taking the candidate of the hypothesis that holds; it overlaps the
// window's accepted spheres when its mask meets theirs.  The window ends where no evaluated
// hypothesis holds.  Results (spheres, their order, the generator state, the failing sphere)
// are those of the sequential loop; only the number of rounds changes.
#define CLOUD_TPC 9                         // threads per candidate: one row of 3 neighbour cells each
#define CLOUD_SLOTS 112                     // candidates per window
#define CLOUD_H 4                           // hypotheses per try (num > 0)
#define CLOUD_W (CLOUD_SLOTS / CLOUD_H)     // 28 tries per window (num > 0)
#define CLOUD_W2 (CLOUD_SLOTS / 2)          // 56 tries per family (num < 0)
#define CLOUD_Q 232                         // draws generated per window (num < 0 uses at most 1 + 4 x 56 + 3)
__device__ __forceinline__ bool cloudOverlap(const CloudArgs& A, const double4 p, const double4 c) {
    // p placed before c (SpherePack.cpp:222-229): periodic distance is the shorter way round
    double dr[3];
    const double pc[3] = {p.x, p.y, p.z}, cc[3] = {c.x, c.y, c.z};
    for (int d = 0; d < 3; d++)
        dr[d] = A.periodic ? fmin(cloudWrapRel(cc[d], pc[d], pc[d] + A.size[d]), cloudWrapRel(pc[d], cc[d], cc[d] + A.size[d])) : pc[d] - cc[d];
    const double rr = p.w + c.w;
    return rr * rr >= (dr[0] * dr[0] + dr[1] * dr[1]) + dr[2] * dr[2];
}
__device__ __forceinline__ double cloudU(uint32_t x) { return (double)(x - 1u) / 2147483645.; }
__device__ __forceinline__ int cloudLo(double p, int j) {
    const int cj = (int)floor(p * j + .5);
    const int lo = cj - 1 < j - (CLOUD_H - 1) ? cj - 1 : j - (CLOUD_H - 1);
    return lo < 0 ? 0 : lo;
}
__device__ __forceinline__ int cloudTry(const CloudArgs& A, int s) {  // the window's try a slot evaluates
    return A.num > 0 ? s / CLOUD_H : (s < CLOUD_W2 ? s : s - (CLOUD_W2 - 1));
}
// out[0]: spheres placed; out[1]: 1 when a sphere found no free spot in maxTry tries
__global__ void __launch_bounds__(1024) k_make_cloud(CloudArgs A, uint32_t* state, int32_t* head, int32_t* next, double4* sph, int64_t* out) {
    __shared__ uint32_t pw[CLOUD_Q];   // 48271^(k+1) mod 2^31-1
    __shared__ uint32_t xs[CLOUD_Q];   // generator state after good draw k of the window
    __shared__ double4 cnd[CLOUD_SLOTS];
    __shared__ unsigned long long ovm[CLOUD_SLOTS][2];  // candidates of earlier tries it overlaps
    __shared__ int on[CLOUD_SLOTS], hit[CLOUD_SLOTS];
    __shared__ int4 ccell[CLOUD_SLOTS];  // cell of each candidate; .w its try, INT_MAX when the slot is off
    __shared__ int loJ[CLOUD_W];         // lo_j of the window's tries (num > 0)
    __shared__ uint32_t st;
    __shared__ int64_t placed;
    __shared__ int tries, needR, fin, failed;
    __shared__ double rCur, rate;
    const int tid = threadIdx.x;
    const uint32_t M = 2147483647u;
    if (tid < CLOUD_Q) {
        uint64_t b = 48271u, r = 1;
        for (int e = tid + 1; e; e >>= 1) {
            if (e & 1) r = r * b % M;
            b = b * b % M;
        }
        pw[tid] = (uint32_t)r;
    }
    if (tid == 0) {
        st = *state;
        placed = 0;
        tries = 0;
        needR = 1;
        fin = 0;
        failed = 0;
        rCur = 0.;
        rate = 1.;
    }
    const int64_t lim = A.cap;  // num when num > 0
    __syncthreads();
    for (;;) {
        if (fin) break;
        // draws of the window; a state of 2^31-2 maps to 1 and is drawn again (rare: sequential)
        const uint32_t x0 = st;
        bool bad = false;
        if (tid < CLOUD_Q) {
            const uint32_t x = (uint32_t)((uint64_t)x0 * pw[tid] % M);
            xs[tid] = x;
            bad = x == M - 1;
        }
        if (__syncthreads_or(bad)) {
            if (tid == 0) {
                uint32_t y = x0;
                for (int k = 0; k < CLOUD_Q; k++) {
                    do y = mstdNext(y); while (y == M - 1);
                    xs[k] = y;
                }
            }
            __syncthreads();
        }
        const int64_t i0 = placed;
        const int s0 = A.num > 0 ? 0 : needR;
        const double p = rate;
        // candidates
        if (tid < CLOUD_SLOTS) {
            const int s = tid;
            int ok = 0, off = 0;
            double r = 0.;
            if (A.num > 0) {
                const int j = s / CLOUD_H, a = cloudLo(p, j) + s % CLOUD_H;
                ok = a <= j && i0 + a < lim;
                if (ok) { r = A.radii[i0 + a]; off = 3 * j; }
            } else {
                int ri = -1;  // draw of the radius, -1: the current sphere's
                if (s < CLOUD_W2) {  // none accepted before try s
                    ok = i0 < lim;
                    off = s0 + 3 * s;
                    if (needR) ri = 0;
                } else {  // all accepted before try j
                    const int j = s - (CLOUD_W2 - 1);
                    ok = i0 + j < lim;
                    off = s0 + 4 * j;
                    ri = off - 1;
                }
                if (ok) r = ri >= 0 ? cloudRadius(A, cloudU(xs[ri])) : rCur;
            }
            if (ok) {
                double c[3];
                for (int d = 0; d < 3; d++) {
                    const double u = cloudU(xs[off + d]);
                    c[d] = A.periodic ? A.mn[d] + A.size[d] * u : (A.mn[d] + r) + (A.size[d] - 2 * r) * u;
                }
                cnd[s] = make_double4(c[0], c[1], c[2], r);
                ccell[s] = make_int4(cloudCell(A, 0, c[0]), cloudCell(A, 1, c[1]), cloudCell(A, 2, c[2]), cloudTry(A, s));
            } else ccell[s] = make_int4(0, 0, 0, INT_MAX);
            if (s < CLOUD_W) loJ[s] = cloudLo(p, s);
            on[s] = ok;
            hit[s] = 0;
            ovm[s][0] = 0;
            ovm[s][1] = 0;
        }
        __syncthreads();
        // grid test: thread (slot, row of 3 neighbour cells); the three list heads are loaded together
        if (tid < CLOUD_SLOTS * CLOUD_TPC) {
            const int s = tid / CLOUD_TPC, row = tid % CLOUD_TPC;
            if (on[s]) {
                const double4 cd = cnd[s];
                const double ccd[3] = {cd.x, cd.y, cd.z};
                const int oo[3] = {0, row % 3 - 1, row / 3 - 1};
                int cell[3];
                bool inr = true;
                for (int d = 1; d < 3; d++) {
                    int q = cloudCell(A, d, ccd[d]) + oo[d];
                    // periodic axes of 3+ cells wrap; with fewer cells (and aperiodic axes) each cell
                    // in range is visited once
                    if (A.periodic && A.nc[d] >= 3) q = (q + A.nc[d]) % A.nc[d];
                    else if (q < 0 || q >= A.nc[d]) inr = false;
                    cell[d] = q;
                }
                if (inr) {
                    const int cx = cloudCell(A, 0, ccd[0]);
                    const int64_t base = ((int64_t)cell[2] * A.nc[1] + cell[1]) * A.nc[0];
                    int cur[3];
                    for (int k = 0; k < 3; k++) {
                        int q = cx + k - 1;
                        if (A.periodic && A.nc[0] >= 3) q = (q + A.nc[0]) % A.nc[0];
                        cur[k] = (q >= 0 && q < A.nc[0]) ? head[base + q] : -1;
                    }
                    bool ov = false;
                    for (int k = 0; k < 3 && !ov; k++)
                        for (int j = cur[k]; j >= 0 && !ov; j = next[j]) ov = cloudOverlap(A, sph[j], cd);
                    if (ov) hit[s] = 1;
                }
            }
        }
        // window test: thread (slot, 16 slots of earlier tries)
        if (tid < CLOUD_SLOTS * 7) {
            const int s = tid / 7, ch = tid % 7;
            if (on[s]) {
                const int4 cs = ccell[s];
                unsigned long long bits = 0;
                for (int h = 0; h < 2; h++) {  // 8 slots at a time: their loads are issued together
                    int4 cq[8];
#pragma unroll
                    for (int b = 0; b < 8; b++) {
                        const int q = ch * 16 + h * 8 + b;
                        cq[b] = ccell[q < CLOUD_SLOTS ? q : CLOUD_SLOTS - 1];
                    }
                    unsigned near = 0;
#pragma unroll
                    for (int b = 0; b < 8; b++) {
                        const int q = ch * 16 + h * 8 + b;
                        // an earlier try's candidate; spheres that can touch lie in neighbouring cells
                        // (cells are at least 2 r_max wide)
                        const int dc[3] = {cq[b].x - cs.x, cq[b].y - cs.y, cq[b].z - cs.z};
                        bool nb = q < CLOUD_SLOTS && cq[b].w < cs.w;
                        for (int d = 0; d < 3; d++) {
                            const int ad = dc[d] < 0 ? -dc[d] : dc[d];
                            nb = nb && (ad <= 1 || (A.periodic && (A.nc[d] < 3 || ad == A.nc[d] - 1)));
                        }
                        near |= nb ? 1u << b : 0u;
                    }
                    if (near) {
                        const double4 cd = cnd[s];
                        for (int b = 0; b < 8; b++) {
                            const int q = ch * 16 + h * 8 + b;
                            if (((near >> b) & 1u) && cloudOverlap(A, cnd[q], cd)) bits |= 1ull << (q & 63);
                        }
                    }
                }
                if (bits) atomicOr(&ovm[s][ch >> 2], bits);
            }
        }
        __syncthreads();
        // walk the tries in order (wave 0; every value below is uniform across the wave except
        // `mine` and the per-lane copies of the window's tests, which the walk reads by lane)
        if (tid < 64) {
            const int lane = tid;
            const bool hasB = lane + 64 < CLOUD_SLOTS;
            const int hA = hit[lane], hB = hasB ? hit[lane + 64] : 1;
            const int myLo = lane < CLOUD_W ? loJ[lane] : 0;
            const unsigned long long mA0 = ovm[lane][0], mA1 = ovm[lane][1];
            const unsigned long long mB0 = hasB ? ovm[lane + 64][0] : 0ull, mB1 = hasB ? ovm[lane + 64][1] : 0ull;
            int mine = 0;  // slot of the window's accepted sphere number `lane`
            unsigned long long acc0 = 0, acc1 = 0;  // accepted slots
            // the walk's state is uniform: kept in scalar registers
            int na = 0, t = __builtin_amdgcn_readfirstlane(tries), used = 0, nr = __builtin_amdgcn_readfirstlane(needR), fl = 0, fn = 0,
                triesUsed = 0, rcSlot = -1;
            const int64_t i0s = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(i0 >> 32)) << 32) |
                                          (uint32_t)__builtin_amdgcn_readfirstlane((int)i0));
            const int s0s = __builtin_amdgcn_readfirstlane(s0);
            auto rl = [](uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); };
            auto rl64 = [&](unsigned long long v, int l) {
                return ((unsigned long long)rl((uint32_t)(v >> 32), l) << 32) | rl((uint32_t)v, l);
            };
            auto overlapsAccepted = [&](int s) {  // s is uniform: its lane's copy is read with v_readlane
                const int l = s & 63;
                const bool b = s >= 64;
                const int h = (int)rl((uint32_t)(b ? hB : hA), l);
                const unsigned long long m0 = rl64(b ? mB0 : mA0, l), m1 = rl64(b ? mB1 : mA1, l);
                return h != 0 || (m0 & acc0) != 0 || (m1 & acc1) != 0;
            };
            auto accept = [&](int s) {
                if (lane == na) mine = s;
                na++;
                if (s < 64) acc0 |= 1ull << s;
                else acc1 |= 1ull << (s - 64);
            };
            // num > 0 without overlaps between the window's candidates (the usual case), and no
            // sphere that could reach maxTry in the window: the walk's only state is na, and try j
            // maps na in its band lo_j..lo_j+3 to na + (candidate hit ? 0 : 1), or ends the walk.
            // The maps are composed by an inclusive scan over the wave (try j on lane j), which gives
            // na after every try; the walk stops at the first try whose map ends it.
            bool fastDone = false;
            int insSlot = -1;
            int64_t insIdx = 0;
            if (A.num > 0 && !__any((mA0 | mA1 | mB0 | mB1) != 0ull) && t + CLOUD_W < A.maxTry) {
                const int j = lane;
                uint32_t T = 0xffffffffu;  // 4 entries of int8: na after try j for na = startLo + k, -1 when it ends
                int startLo = myLo;          // lo of the first try of the lane's composed range
                if (j < CLOUD_W) {
                    T = 0;
                    for (int k = 0; k < 4; k++) {
                        const int nin = myLo + k;
                        int o = -1;
                        if (nin <= j && i0s + nin < lim) o = nin + (hit[j * CLOUD_H + k] ? 0 : 1);
                        T |= (uint32_t)(o & 0xff) << (8 * k);
                    }
                }
                for (int off = 1; off < 32; off <<= 1) {  // 28 tries fit in 32 lanes
                    const uint32_t P = (uint32_t)__shfl_up((int)T, off);
                    const int pLo = __shfl_up(startLo, off);
                    if (j >= off && j < CLOUD_W) {
                        uint32_t R = 0;
                        for (int k = 0; k < 4; k++) {
                            const int pk = (int)(int8_t)((P >> (8 * k)) & 0xffu);
                            int o = -1;
                            if (pk >= 0 && pk - startLo >= 0 && pk - startLo < 4) o = (int)(int8_t)((T >> (8 * (pk - startLo))) & 0xffu);
                            R |= (uint32_t)(o & 0xff) << (8 * k);
                        }
                        T = R;
                        startLo = pLo;
                    }
                }
                const int X = (int)(int8_t)(T & 0xffu);  // na after try j (the range starts at try 0, na 0)
                int Xp = __shfl_up(X, 1);
                if (j == 0) Xp = 0;
                const unsigned long long stopM = __ballot(j < CLOUD_W && X < 0);
                const int jstar = stopM ? __ffsll((long long)stopM) - 1 : CLOUD_W;
                const int naF = jstar == 0 ? 0 : __builtin_amdgcn_readlane(X, jstar - 1);
                const bool accepted = j < jstar && X > Xp;
                const unsigned long long accM = __ballot(accepted);
                if (accepted) {
                    insSlot = j * CLOUD_H + (Xp - myLo);
                    insIdx = i0s + Xp;
                }
                na = naF;
                used = 3 * jstar;
                triesUsed = jstar;
                t = accM ? jstar - 1 - (63 - __clzll((long long)accM)) : t + jstar;
                fn = i0s + naF >= lim ? 1 : 0;
                fastDone = true;
            }
            if (fastDone) {
            } else if (A.num > 0) {
                for (int j = 0; j < CLOUD_W; j++) {
                    if (i0s + na >= lim) { fn = 1; break; }
                    const int lo = __builtin_amdgcn_readlane(myLo, j);
                    if (na < lo || na >= lo + CLOUD_H) break;
                    const int s = j * CLOUD_H + (na - lo);
                    const bool ov = overlapsAccepted(s);
                    used = 3 * (j + 1);
                    triesUsed = j + 1;
                    if (!ov) {
                        accept(s);
                        t = 0;
                    } else if (++t == A.maxTry) {
                        fl = 1;
                        fn = 1;
                        break;
                    }
                }
                if (!fn && i0s + na >= lim) fn = 1;
            } else {
                int fam = -1;  // -1 undecided (try 0), 0 none accepted, 1 all accepted
                for (int j = 0;; j++) {
                    if (i0s + na >= lim) { fn = 1; break; }
                    if (fam <= 0 && j >= CLOUD_W2) break;
                    if (fam == 1 && j > CLOUD_W2) break;
                    const int s = fam == 1 ? CLOUD_W2 - 1 + j : j;
                    const bool ov = overlapsAccepted(s);
                    triesUsed = j + 1;
                    if (!ov) {
                        accept(s);
                        t = 0;
                        nr = 1;
                        rcSlot = -1;
                        if (fam == 1) used = s0s + 4 * j + 3;
                        else used = s0s + 3 * (j + 1);
                        if (fam == 0) break;  // the next sphere's radius draw follows a run of misses
                        fam = 1;
                    } else {
                        if (fam == 1) {  // first try of sphere i0+j missed
                            used = s0s + 4 * j + 3;
                            nr = 0;
                            rcSlot = s;
                            if (++t == A.maxTry) { fl = 1; fn = 1; }
                            break;
                        }
                        fam = 0;
                        used = s0s + 3 * (j + 1);
                        nr = 0;
                        rcSlot = s;
                        if (++t == A.maxTry) { fl = 1; fn = 1; break; }
                    }
                }
            }
            const double rc = rcSlot >= 0 ? cnd[rcSlot].w : rCur;  // the radius of a sphere left mid-tries
            if (!fastDone && lane < na) {
                insSlot = mine;
                insIdx = i0s + lane;
            }
            // spheres accepted in the window go into the grid
            if (insSlot >= 0) {
                const double4 c = cnd[insSlot];
                const int64_t idx = insIdx;
                sph[idx] = c;
                const int64_t cell = ((int64_t)cloudCell(A, 2, c.z) * A.nc[1] + cloudCell(A, 1, c.y)) * A.nc[0] + cloudCell(A, 0, c.x);
                next[idx] = atomicExch(&head[cell], (int32_t)idx);
            }
            if (lane == 0) {
                if (used > 0) st = xs[used - 1];
                placed = i0 + na;
                tries = t;
                needR = nr;
                rCur = rc;
                fin = fn;
                failed = fl;
                if (triesUsed > 0) rate = (double)na / triesUsed;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    if (tid == 0) {
        out[0] = placed;
        out[1] = failed;
        *state = st;
    }
}

// ---------------- regular packings: pack.regularOrtho / pack.regularHexa ----------------
// Lattice points in the order of itertools.product(xx, yy, zz) (z fastest), kept where the
// predicate holds with the sphere radius as pad (py/pack.py:323-350); predicates
// inAlignedBox, inSphere, inCylinder (py/_packPredicates.cpp:96-104, 136-148, 189-205).
struct LatticeArgs {
    int hexa, kind;   // kind: 0 inAlignedBox(mn, mx), 1 inSphere(centre, r), 2 inCylinder(c1, c2, r)
    int64_t n[3];     // points per axis
    double start[3], a1[3];  // ortho: numpy.arange(start, stop, step) holds start, start+step, then start + i (a1 - start)
    double mn[3], radius, a, hy, hz;  // hexa
    double p[7];      // predicate: box mn3 mx3 / centre3 r / c1_3 c2_3 r
};
WDM_HD double arangeAt(const LatticeArgs& L, int d, int64_t i) {
    if (i == 0) return L.start[d];
    if (i == 1) return L.a1[d];
    return L.start[d] + (double)i * (L.a1[d] - L.start[d]);
}
WDM_HD void latticePoint(const LatticeArgs& L, int64_t idx, double& x, double& y, double& z) {
    const int64_t k = idx % L.n[2], t = idx / L.n[2], j = t % L.n[1], i = t / L.n[1];
    if (!L.hexa) {
        x = arangeAt(L, 0, i);
        y = arangeAt(L, 1, j);
        z = arangeAt(L, 2, k);
        return;
    }
    x = (L.mn[0] + L.radius) + (double)i * L.a;
    y = (L.mn[1] + L.radius) + (double)j * L.hy;
    z = (L.mn[2] + L.radius) + (double)k * L.hz;
    if (j % 2 == 0) x += (k % 2 == 0) ? L.a / 2. : -L.a / 2.;
    if (k % 2 != 0) {
        x += L.a / 2.;
        y += L.hy / 2.;
    }
}
WDM_HD bool latticePred(const LatticeArgs& L, double x, double y, double z) {
    const double pad = L.radius;
    const double* p = L.p;
    if (L.kind == 0)
        return p[0] + pad <= x && p[3] - pad >= x && p[1] + pad <= y && p[4] - pad >= y && p[2] + pad <= z && p[5] - pad >= z;
    if (L.kind == 1) {
        const double dx = x - p[0], dy = y - p[1], dz = z - p[2];
        return sqrt((dx * dx + dy * dy) + dz * dz) - pad <= p[3] - pad;
    }
    const double cx = p[3] - p[0], cy = p[4] - p[1], cz = p[5] - p[2];
    const double ht = sqrt((cx * cx + cy * cy) + cz * cz);
    const double u = (((x * cx + y * cy) + z * cz) - ((p[0] * cx + p[1] * cy) + p[2] * cz)) / (ht * ht);
    if ((u * ht < 0 + pad) || (u * ht > ht - pad)) return false;
    const double ax = x - p[0], ay = y - p[1], az = z - p[2], bx = x - p[3], by = y - p[4], bz = z - p[5];
    const double qx = ay * bz - az * by, qy = az * bx - ax * bz, qz = ax * by - ay * bx;
    const double axisDist = sqrt((qx * qx + qy * qy) + qz * qz) / ht;
    return !(axisDist > p[6] - pad);
}
__global__ void __launch_bounds__(WDM_BLOCK) k_lattice_flags(LatticeArgs L, int64_t total, uint8_t* flags) {
    const int64_t idx = (int64_t)blockIdx.x * WDM_BLOCK + threadIdx.x;
    if (idx >= total) return;
    double x, y, z;
    latticePoint(L, idx, x, y, z);
    flags[idx] = latticePred(L, x, y, z) ? 1 : 0;
}
__global__ void __launch_bounds__(WDM_BLOCK) k_lattice_write(LatticeArgs L, const int64_t* sel, const int64_t* nSel, double* out) {
    const int64_t i = (int64_t)blockIdx.x * WDM_BLOCK + threadIdx.x;
    if (i >= *nSel) return;
    double x, y, z;
    latticePoint(L, sel[i], x, y, z);
    out[3 * i] = x;
    out[3 * i + 1] = y;
    out[3 * i + 2] = z;
}

// ---------------- scene generation: Jacobi overlap relaxation ----------------
// Each sphere moves by half of every overlap with its neighbours (found through the
// same cell grid); all displacements are computed from the old positions, so the
// result does not depend on thread order.  Walls then push spheres back to their side.
// The grid has one level here (its cells given by g.n), entries sorted by (cell, id).
__global__ void __launch_bounds__(WDM_BLOCK) k_relax(const SceneDesc* D, const FRScene* FR, Map m, const uint32_t* ckeys, const int32_t* cvals,
                                                    const int32_t* cbeg, double4* disp) {
    int64_t lb;
    const FRScene& F = FR[mapBlock(m, lb)];
    const SceneDesc& S = D[F.desc];
    const NodeView n = S.n;
    const GridArgs g = F.g;  // by value: the loops below store to global memory
    int li = (int)(lb * WDM_BLOCK + threadIdx.x);
    if (li >= S.nbinned) return;
    const int64_t i = F.binStart + li;
    int id = cvals[i];
    uint32_t ck = (uint32_t)(ckeys[i] - F.cellBase);
    if ((int)ck == F.nCells || ptShape(S.p.ptype[id]) != 1) {  // spheres only
        disp[id] = make_double4(0., 0., 0., 0.);
        return;
    }
    int cx = ck % g.n[0], cy = (ck / g.n[0]) % g.n[1], cz = ck / (g.n[0] * g.n[1]);
    double4 pi = n.posr[id];
    D3 d = mk3(0., 0., 0.);
    // periodic axes with fewer than 3 cells hold a single cell (no neighbours); aperiodic axes always scan +-1
    int rx = (!g.periodic || g.n[0] >= 3) ? 1 : 0, ry = (!g.periodic || g.n[1] >= 3) ? 1 : 0, rz = (!g.periodic || g.n[2] >= 3) ? 1 : 0;
    for (int dz = -rz; dz <= rz; dz++) {
        int z = cz + dz;
        if (g.periodic) z = (z + g.n[2]) % g.n[2]; else if (z < 0 || z >= g.n[2]) continue;
        for (int dy = -ry; dy <= ry; dy++) {
            int y = cy + dy;
            if (g.periodic) y = (y + g.n[1]) % g.n[1]; else if (y < 0 || y >= g.n[1]) continue;
            for (int dx = -rx; dx <= rx; dx++) {
                int x = cx + dx;
                if (g.periodic) x = (x + g.n[0]) % g.n[0]; else if (x < 0 || x >= g.n[0]) continue;
                int64_t cell = F.cellBase + (z * g.n[1] + y) * g.n[0] + x;
                const int s = cbeg[cell], e = cbeg[cell + 1];
                for (int k = s; k < e; k++) {
                    int j = cvals[k];
                    if (j == id) continue;
                    double4 pj = n.posr[j];
                    D3 r = xyz(pi) - xyz(pj);
                    if (g.periodic) {
                        r.x -= g.L[0] * rint(r.x / g.L[0]);
                        r.y -= g.L[1] * rint(r.y / g.L[1]);
                        r.z -= g.L[2] * rint(r.z / g.L[2]);
                    }
                    double dist = sqrt(sqn3(r));
                    double ov = (pi.w + pj.w) - dist;
                    if (ov > 0 && dist > 0) d = d + r * (0.5 * ov / dist);
                }
            }
        }
    }
    disp[id] = make_double4(d.x, d.y, d.z, 0.);
}
__global__ void __launch_bounds__(WDM_BLOCK) k_relax_apply(const SceneDesc* D, const FRScene* FR, Map m, const double4* disp) {
    int64_t lb;
    const SceneDesc& S = D[FR[mapBlock(m, lb)].desc];
    const NodeView n = S.n;
    int i = (int)(lb * WDM_BLOCK + threadIdx.x);
    if (i >= S.nbinned) return;
    int id = S.binned[i];
    double4 pr = n.posr[id];
    double4 d = disp[id];
    double x[3] = {pr.x + d.x, pr.y + d.y, pr.z + d.z};
    if (ptShape(S.p.ptype[id]) == 1) {  // walls keep spheres on their side
        for (int u = 0; u < S.nunb; u++) {
            const int w = S.unb[u];
            const int32_t pt = S.p.ptype[w];
            const int ax = ptAxis(pt), se = ptSense(pt);
            const double4 pw = n.posr[w];
            const double c = ax == 0 ? pw.x : (ax == 1 ? pw.y : pw.z);
            if (se > 0) x[ax] = fmax(x[ax], c + pr.w);
            else if (se < 0) x[ax] = fmin(x[ax], c - pr.w);
        }
    }
    n.posr[id] = make_double4(x[0], x[1], x[2], pr.w);
}
