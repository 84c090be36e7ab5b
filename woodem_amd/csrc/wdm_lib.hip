// libwoodem_hip: host side of the MI355X DEM timestep.
//
// C++ engine classes named after the reference's plugins (Leapfrog,
// InsertionSortCollider, ContactLoop, Scene::doOneStep) drive the kernels of
// wdm_kernels.h on one HIP stream per context; the extern "C" functions at the
// bottom are the drop-in boundary declared in include/woodem_hip.h.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <deque>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/woodem_hip.h"
#include "wdm_kernels.h"

// Device memory of wdm_select_box_noalloc: per-block counts and offsets of the selection, the
// total and the selected ids, for scenes of up to NOALLOC_MAX_PARTICLES (4,194,304) particles
// (16.1 MiB).  One copy per GPU, held until the process ends.  In host code the name is a
// stand-in, not a device address: the host uses the address hipGetSymbolAddress gives for
// each device (scratchInit).  Calls that overlap on one GPU share it and race.
#define NOALLOC_MAX_BLOCKS (1024 * SCAN_PER_THREAD)
#define NOALLOC_MAX_PARTICLES (NOALLOC_MAX_BLOCKS * WDM_BLOCK)
#define NOALLOC_OFF_OFFS (4 * NOALLOC_MAX_BLOCKS)
#define NOALLOC_OFF_TOTAL (8 * NOALLOC_MAX_BLOCKS)
#define NOALLOC_OFF_IDS (8 * NOALLOC_MAX_BLOCKS + 256)
#define NOALLOC_BYTES (NOALLOC_OFF_IDS + 4 * (size_t)NOALLOC_MAX_PARTICLES)
static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];

namespace {

struct WdmError : std::runtime_error {
    int code;
    WdmError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw WdmError(WDM_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));             \
    } while (0)

inline unsigned nblk(int64_t n) { return (unsigned)((n + WDM_BLOCK - 1) / WDM_BLOCK); }
inline int bitsFor(uint64_t v) { int b = 0; while (b < 64 && (v >> b) != 0) b++; return b == 0 ? 1 : b; }

// ---------------- host 3x3 algebra for the periodic cell (left-to-right sums) ----------------
struct HM3 { double m[3][3]; };
HM3 hzero() { HM3 r; for (auto& a : r.m) for (double& x : a) x = 0; return r; }
HM3 heye() { HM3 r = hzero(); r.m[0][0] = r.m[1][1] = r.m[2][2] = 1; return r; }
HM3 hadd(const HM3& a, const HM3& b) { HM3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[i][j] = a.m[i][j] + b.m[i][j]; return r; }
HM3 hsub(const HM3& a, const HM3& b) { HM3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[i][j] = a.m[i][j] - b.m[i][j]; return r; }
HM3 hscale(const HM3& a, double s) { HM3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[i][j] = a.m[i][j] * s; return r; }
HM3 hdivs(const HM3& a, double s) { HM3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[i][j] = a.m[i][j] / s; return r; }
HM3 htrans(const HM3& a) { HM3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[i][j] = a.m[j][i]; return r; }
HM3 hmul(const HM3& a, const HM3& b) {
    HM3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = (a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j];
    return r;
}
double hcof(const HM3& a, int i, int j) {
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return a.m[i1][j1] * a.m[i2][j2] - a.m[i1][j2] * a.m[i2][j1];
}
double hdet(const HM3& a) { return (hcof(a, 0, 0) * a.m[0][0] + hcof(a, 1, 0) * a.m[1][0]) + hcof(a, 2, 0) * a.m[2][0]; }
// cofactors times 1/det (the algorithm of Eigen 3's 3x3 inverse, used by core/Cell.cpp)
HM3 hinv(const HM3& a) {
    double inv = 1.0 / hdet(a);
    HM3 r;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[i][j] = hcof(a, j, i) * inv;
    return r;
}
bool hequal(const HM3& a, const HM3& b) { for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) if (a.m[i][j] != b.m[i][j]) return false; return true; }
M3 toM3(const HM3& a) { M3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[3 * i + j] = a.m[i][j]; return r; }
void leviCivitaH(const HM3& m, double out[3]) { out[0] = m.m[1][2] - m.m[2][1]; out[1] = m.m[2][0] - m.m[0][2]; out[2] = m.m[0][1] - m.m[1][0]; }

// ---------------- device buffers ----------------
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
};

struct Timing {
    bool on = false;
    std::map<std::string, std::pair<double, int64_t>> acc;
    // event pairs (start, end) of launches not read yet, and their kernel names: events are read
    // when the times are asked for, so no timed launch waits for the queue to drain and a start
    // event completes when the GPU reaches the kernel, not while it idles for the launch
    std::vector<hipEvent_t> ev;
    std::vector<const char*> pend;
    int open = 0;  // timers whose end event is not recorded yet
};

struct Shared;
struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t budget = 0, used = 0;
    std::string err;
    std::map<int32_t, struct Scene*> scenes;
    int32_t nextId = 0;
    Timing timing;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Shared* sh = nullptr;  // batch scratch, descriptor tables, per-scene words
    // launch-shape knobs (results do not depend on them): steps per window, windows replayed
    // as captured graphs, collider cell width over the largest sphere bound, collider grid
    // levels (0: chosen per scene from its bound sizes; WDM_GRID_LEVELS in the environment)
    int window = 16;
    int graphs = 1;
    double cellFactor = 1.;
    int gridLevels = 0;
    // pairs a wavefront of the pair search may find: passing the FP32 box test, passing the
    // exact test (each grows when a pass needs more)
    int32_t scanCap = 1280, waveCap = 640;
    // broadphase of the collider's full run: 0 the multi-level grid scan, 1 sort and sweep
    // (radix sort of the bound entries by minimum x, LDS-tiled sweep); the contact set is the same
    int broadphase = 0;
    // steps between relayouts of a scene's work list (its real items moved to the front) after
    // its last full run or relayout, 0: never; only lists of at least relayoutMin contacts
    // (smaller ones stay in cache, and scenes of a batch would cut each other's windows).
    // Memory order only: results do not change
    int relayout = 32;
    int64_t relayoutMin = 1 << 20;
    hipError_t scratchErr = hipSuccess;  // looking up the device address of `scratch` at wdm_create
};

void* devAlloc(Ctx* c, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (c->budget && c->used + bytes > c->budget) throw WdmError(WDM_E_OOM, "HBM budget exceeded");
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) throw WdmError(WDM_E_OOM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    c->used += bytes;
    return p;
}
void devFree(Ctx* c, void* p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    c->used -= std::min(c->used, bytes == 0 ? (size_t)8 : bytes);
}
template <typename T>
void ensure(Ctx* c, DBuf<T>& b, size_t n, bool keep = false) {
    if (b.n >= n && b.p) return;
    // headroom: sizes that creep up between collider passes reallocate rarely
    size_t nn = std::max<size_t>(n + n / 8, b.n + b.n / 2);
    nn = std::max<size_t>(nn, 16);
    T* p = (T*)devAlloc(c, nn * sizeof(T));
    if (keep && b.p && b.n) HIPCHK(hipMemcpyAsync(p, b.p, b.n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    if (b.p) { HIPCHK(hipStreamSynchronize(c->stream)); devFree(c, b.p, b.n * sizeof(T)); }
    b.p = p;
    b.n = nn;
}
template <typename T>
void release(Ctx* c, DBuf<T>& b) { if (b.p) devFree(c, b.p, b.n * sizeof(T)); b.p = nullptr; b.n = 0; }

template <typename T>
struct ConBuf {
    DBuf<uint64_t> key;
    DBuf<int2> ids;
    DBuf<uint32_t> info;
    DBuf<int32_t> cd, slot, counter, itm, nItems;
    DBuf<int4> irec;
    DBuf<int2> iaux;
    DBuf<T> st, st2;
    DBuf<double4> cpt, gfc;
    int64_t cap = 0;
    void reserve(Ctx* c, int64_t n) {
        if (n <= cap && key.p) return;
        int64_t nc = std::max<int64_t>(n + n / 3, 1024);
        free(c);
        ensure(c, key, nc); ensure(c, ids, nc); ensure(c, info, nc); ensure(c, cd, nc);
        ensure(c, slot, nc); ensure(c, counter, 1); ensure(c, itm, nc); ensure(c, irec, nc); ensure(c, iaux, nc); ensure(c, nItems, 1);
        ensure(c, st, (size_t)NSTATE * nc); ensure(c, st2, (size_t)NSTRIDE * nc); ensure(c, cpt, nc); ensure(c, gfc, 2 * (size_t)nc);
        HIPCHK(hipMemsetAsync(nItems.p, 0, sizeof(int32_t), c->stream));
        HIPCHK(hipMemsetAsync(counter.p, 0, sizeof(int32_t), c->stream));
        cap = nc;
    }
    ConView<T> view() const {
        ConView<T> v; v.key = gpc(key.p); v.ids = gpc(ids.p); v.info = gpc(info.p); v.cd = gpc(cd.p); v.slot = gpc(slot.p);
        v.counter = counter.p; v.st = gpc(st.p); v.st2 = gpc(st2.p); v.cpt = gpc(cpt.p); v.itm = gpc(itm.p); v.irec = gpc(irec.p); v.iaux = gpc(iaux.p); v.nItems = nItems.p;
        v.gfc = gpc(gfc.p); v.cap = cap; return v;
    }
    void free(Ctx* c) {
        release(c, key); release(c, ids); release(c, info); release(c, cd); release(c, st); release(c, st2); release(c, gfc);
        release(c, slot); release(c, counter); release(c, itm); release(c, irec); release(c, iaux); release(c, nItems); release(c, cpt); cap = 0;
    }
};

// ---------------- staged host->device tables ----------------
// A device table written by stream-ordered copies from two alternating pinned staging
// buffers: kernels queued before an upload read the previous contents, kernels queued
// after it the new ones, and the host never overwrites a staging buffer whose copy is
// still in flight.
struct Staged {
    uint8_t* dev = nullptr;
    uint8_t* host[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    size_t cap = 0;
    int k = 0;
    // device base address valid for an upload of `bytes`
    uint8_t* prepare(Ctx* c, size_t bytes) {
        if (bytes <= cap && dev) return dev;
        size_t nc = std::max<size_t>(bytes + bytes / 2, 4096);
        HIPCHK(hipStreamSynchronize(c->stream));
        free(c);
        dev = (uint8_t*)devAlloc(c, nc);
        for (int i = 0; i < 2; i++) {
            HIPCHK(hipHostMalloc((void**)&host[i], nc, hipHostMallocDefault));
            HIPCHK(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
            HIPCHK(hipEventRecord(ev[i], c->stream));
        }
        cap = nc;
        return dev;
    }
    uint8_t* upload(Ctx* c, const void* src, size_t bytes) {
        prepare(c, bytes);
        k ^= 1;
        HIPCHK(hipEventSynchronize(ev[k]));
        if (bytes) std::memcpy(host[k], src, bytes);
        if (bytes) HIPCHK(hipMemcpyAsync(dev, host[k], bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipEventRecord(ev[k], c->stream));
        return dev;
    }
    void free(Ctx* c) {
        if (dev) devFree(c, dev, cap);
        for (int i = 0; i < 2; i++) {
            if (host[i]) (void)hipHostFree(host[i]);
            if (ev[i]) (void)hipEventDestroy(ev[i]);
            host[i] = nullptr;
            ev[i] = nullptr;
        }
        dev = nullptr;
        cap = 0;
    }
};

// packs several host arrays into one upload (64-byte aligned members)
struct Blob {
    std::vector<uint8_t> b;
    size_t add(const void* p, size_t n) {
        size_t o = (b.size() + 63) & ~(size_t)63;
        b.resize(o + n);
        if (n) std::memcpy(b.data() + o, p, n);
        return o;
    }
    template <typename T>
    size_t addv(const std::vector<T>& v) { return add(v.data(), v.size() * sizeof(T)); }
    template <typename T>
    T* at(size_t off) { return (T*)(b.data() + off); }
};

// ---------------- per-context state shared by the scenes ----------------
struct Shared {
    DBuf<uint32_t> words;  // 4 per scene slot: Verlet trigger, created energy slots, device errors, 0
    uint32_t* hostWords = nullptr;
    size_t hostWordsCap = 0;
    std::vector<int32_t> freeSlots;
    int32_t nslots = 0;
    Staged stDesc, stMapNode, stMapCon, stCon, stFR, stDyn;
    Staged stMapP, stMapB;
    Staged stRunA, stRunO, stRunN;  // collider full run: the tables of its phases
    Staged stRl;                    // work-list relayout: its tables
    DBuf<uint8_t> rlFlag;
    DBuf<int64_t> rlScan;
    DBuf<int32_t> rlMap;
    void* pin = nullptr;            // pinned readback memory
    size_t pinCap = 0;
    // pinned copy of a full run's bound extents and histograms, read by the host after the
    // run's pair-count synchronisation: the grids of the scenes' next full run
    uint8_t* extMail = nullptr;
    size_t extMailCap = 0;
    hipEvent_t evWords = nullptr;  // recorded after the copy of the words to the host
    // energy partials per block
    DBuf<double> epartL, epartC;
    // collider full-run scratch, batch-wide
    DBuf<double> extPart, extOut;
    DBuf<uint32_t> hist;                 // bound extents per scene, quarter octaves
    DBuf<uint32_t> ekey, eslot, ccnt;    // bound entries: cell key and slot; count per cell
    DBuf<uint64_t> swK, swK2;            // sort-and-sweep broadphase: keys (scene, minimum x) and their sorted copy
    DBuf<uint32_t> swV, swV2;            // entries, and in sorted order
    DBuf<int32_t> cbeg;                  // first entry of each cell (exclusive prefix of ccnt)
    DBuf<uint32_t> ckeys2;               // sorted entries: cell key, particle id, FP32 box
    DBuf<int32_t> cvals2, oslot, srcRow;
    DBuf<uint8_t> ooFlag, rflag, nrflag;  // 0/1 flags (nrflag: bit 1 a near pair) of the merge's prefix sums
    DBuf<float4> cbox;
    DBuf<double4> eb0, eb1;  // sorted entries: copies of the bound corners, (shape, mask)
    DBuf<int2> einfo;
    // bucket sort scratch: items, their buckets, merge buffer, long buckets
    DBuf<BItem> btmp;
    DBuf<uint64_t> baux;
    DBuf<uint32_t> longList, longCnt;
    DBuf<int64_t> tsum;  // prefix sums: per-tile sums
    // candidate pairs: per wavefront of k_pairs its count and region; per row (scene, min id)
    // the count and offsets
    DBuf<int32_t> wcnt, wmax, scnt;
    DBuf<int2> sPair;  // pairs passing the FP32 test: (entry, partner entry)
    DBuf<uint64_t> rItem;
    DBuf<uint32_t> rRow, rSlot, rowCnt;
    DBuf<int64_t> rowOff, oldRow;
    DBuf<int64_t> ooRank, rscan, starts, nrscan;
    DBuf<uint64_t> candK;
    DBuf<int32_t> candP;
    // CSR entries: bucket (row of the particle), slot, item; count per row and offsets
    DBuf<uint32_t> eB, eS, csrCnt;
    DBuf<uint64_t> eI;
    DBuf<int64_t> csrScan;
    // large bounds listed per cell: per entry its listing count and offset; the listings
    DBuf<uint32_t> lcntE, lkeys, lslot, lcnt;
    DBuf<int64_t> loffE;
    DBuf<int32_t> lids, lbeg, lvals;
    DBuf<uint8_t> temp;
    DBuf<double4> disp;
    DBuf<double4> cptTmp;  // absolute contact points of a mixed-precision list (download, upload)
    DBuf<unsigned long long> ddOut;  // DynDt: per scene, bits of the minimum squared critical step
    // window of steps: its tables (device and pinned staging) and the last captured graph
    DBuf<uint8_t> winDev;
    uint8_t* winHost = nullptr;
    size_t winHostCap = 0;
    hipGraph_t winGraph = nullptr;  // the captured graph, kept as long as its executable
    hipGraphExec_t winExec = nullptr;
    std::vector<uint64_t> winKey;
    int64_t graphLaunches = 0, graphCaptures = 0, windows = 0;
    // The window graph goes before any memory its kernel arguments point to (the window tables,
    // the energy partial sums) or any scene is freed, so that no executable graph outlives what
    // it references: the queue is drained, then the executable and its graph are destroyed.
    void dropGraph(Ctx* c) {
        if (!winExec && !winGraph) return;
        (void)hipStreamSynchronize(c->stream);
        if (winExec) (void)hipGraphExecDestroy(winExec);
        if (winGraph) (void)hipGraphDestroy(winGraph);
        winExec = nullptr;
        winGraph = nullptr;
        winKey.clear();
    }
    void freeAll(Ctx* c) {
        dropGraph(c);
        release(c, winDev);
        if (winHost) (void)hipHostFree(winHost);
        winHost = nullptr;
        winHostCap = 0;
        release(c, words);
        if (hostWords) (void)hipHostFree(hostWords);
        hostWords = nullptr;
        hostWordsCap = 0;
        release(c, rlFlag); release(c, rlScan); release(c, rlMap);
        for (Staged* s : {&stDesc, &stMapNode, &stMapCon, &stCon, &stFR, &stDyn, &stMapP, &stMapB, &stRunA, &stRunO, &stRunN, &stRl})
            s->free(c);
        release(c, epartL); release(c, epartC); release(c, extPart); release(c, extOut); release(c, hist);
        release(c, ekey); release(c, eslot); release(c, ccnt); release(c, cbeg); release(c, ckeys2); release(c, cvals2); release(c, cbox); release(c, eb0); release(c, eb1); release(c, einfo);
        release(c, btmp); release(c, baux); release(c, longList); release(c, longCnt); release(c, tsum);
        release(c, wcnt); release(c, wmax); release(c, scnt); release(c, sPair); release(c, rItem); release(c, rRow); release(c, rSlot); release(c, rowCnt); release(c, rowOff);
        release(c, oldRow); release(c, ooFlag); release(c, oslot); release(c, srcRow);
        release(c, ooRank); release(c, rflag); release(c, rscan); release(c, starts); release(c, nrflag); release(c, nrscan); release(c, candK);
        release(c, candP); release(c, eB); release(c, eS); release(c, csrCnt); release(c, eI); release(c, csrScan);
        release(c, temp); release(c, disp); release(c, cptTmp); release(c, ddOut);
        release(c, lcntE); release(c, lkeys); release(c, lslot); release(c, lcnt); release(c, loffE); release(c, lids); release(c, lbeg);
        release(c, lvals); release(c, swK); release(c, swK2); release(c, swV); release(c, swV2);
        if (pin) (void)hipHostFree(pin);
        pin = nullptr;
        pinCap = 0;
        if (extMail) (void)hipHostFree(extMail);
        extMail = nullptr;
        extMailCap = 0;
        if (evWords) (void)hipEventDestroy(evWords);
        evWords = nullptr;
    }
};

// ---------------- Scene (core/Scene.hpp:27) with its DemField (pkg/dem/Particle.hpp:278) ----------------
struct Scene {
    Ctx* ctx;
    int32_t np, nn, nmat;
    int32_t id = -1;    // scene id of the ABI
    int32_t slot = -1;  // index of the scene's words in Shared::words
    int32_t failed = 0;    // WDM_E_* of the device error that stopped the scene (0: running)
    std::string failText;
    wdm_params P{};
    bool paramsSet = false, nodesSet = false, partsSet = false, matsSet = false;
    // host copies of static particle data
    std::vector<int32_t> ptypeH;
    std::vector<double> radiusH;
    std::vector<int32_t> binnedH, unbH;
    std::vector<uint32_t> flagsH;
    bool hasFacets = false;
    std::vector<uint32_t> facetMasks;  // distinct masks of the facets
    // two facets may pass Collider::mayCollide's mask tests
    bool facetMasksPair(uint32_t lone) const {
        for (uint32_t a : facetMasks)
            for (uint32_t b : facetMasks)
                if ((a & b) && !(a & b & lone)) return true;
        return false;
    }
    // cell (core/Cell.hpp)
    bool periodic = false;
    HM3 hSize = heye(), gradV = hzero(), nextGradV = hzero(), pprevHsize = heye(), trsf = heye();
    double size[3] = {1, 1, 1}, spinVec[3] = {0, 0, 0};
    int homoDeform = 4;
    // counters
    int64_t step = 0, nFullRuns = 0;
    double time = 0, nextDt = NAN;  // Scene.time, Scene.nextDt (core/Scene.hpp)
    wdm_dyndt dd{};                 // DynDt after ContactLoop (enabled 0: not in the engine list)
    bool contactsRun = false;       // a ContactLoop ran on the current contact list
    int64_t layoutStep = 0;         // step of the last full run or work-list relayout
    int64_t relayouts = 0;
    std::vector<int32_t> facetNodesH, nodeMasterH;  // nodes of facets; clump master of each node (-1)
    // ParticleContainer ids (pkg/dem/ParticleContainer.cpp:23-72): particles in use end before
    // parEnd; holes below it are reused first, in the order they were freed
    int32_t parEnd = 0;
    std::deque<int32_t> freeIds;
    double verletDist = 0;
    bool verletResolved = false;
    bool needFull = true;
    // bound extents of the last full run (min corner, max corner, max extent, max sphere
    // extent) and their histogram: the next full run lays its grid out from them without
    // waiting for its own (extValid: set by a full run, cleared by anything that needs one)
    bool extValid = false;
    double extLast[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<uint32_t> histLast;
    bool lastLeapfrogVerletChecked = false;
    bool forcesFromContacts = false;  // node forces are the sums of the last ContactLoop (not materialised)
    // a Leapfrog ran after the last ContactLoop: the node forces are its reset values (m g, no
    // torque; Leapfrog.cpp:240-245), which the arrays do not hold yet
    bool forcesReset = false;
    HM3 hSizeCL = heye();             // cell hSize at the last ContactLoop (branch vectors of the B side)
    // nodes
    DBuf<double4> posr, velm, angI, ori, force, torque, lastPos;
    DBuf<uint32_t> flags;
    DBuf<uint8_t> checked;
    // particles
    DBuf<int32_t> ptype, pmat, binned, unb;
    DBuf<uint32_t> pmask;
    DBuf<int4> fnodes;
    DBuf<double4> bmin, bmax, mat;
    // contacts, double-buffered (merge target of the next full run)
    ConBuf<double> cb64[2];
    ConBuf<float> cb32[2];
    int cur = 0;
    int64_t nc = 0;
    // CSR contact -> particle, realness bits of the last ContactLoop
    DBuf<int64_t> csrOff;
    DBuf<uint32_t> entV;
    DBuf<uint64_t> realMask;
    // energy accumulators
    DBuf<double> E;
    // clumps
    int32_t nClumps = 0, nMembers = 0;
    bool clumpFree = false;  // a clump master with a free DOF
    int64_t clumpFlagged = 0;  // nodes flagged CLUMP or CLUMPED at upload
    bool clumpsSet = true;     // flags and uploaded clumps agree
    DBuf<int32_t> master, nodeClump, cOff, cMem;
    DBuf<double4> cRelPos, cRelOri, cFT;

    Scene(Ctx* c, int32_t np_, int32_t nn_, int32_t nm_) : ctx(c), np(np_), nn(nn_), nmat(nm_) {}

    bool mixed() const { return P.precision == WDM_PREC_MIXED; }

    NodeView nview() {
        NodeView v; v.posr = gpc(posr.p); v.velm = gpc(velm.p); v.angI = gpc(angI.p); v.ori = gpc(ori.p); v.force = gpc(force.p);
        v.torque = gpc(torque.p); v.lastPos = gpc(lastPos.p); v.flags = gpc(flags.p); v.checked = gpc(checked.p);
        v.master = nClumps > 0 ? gpc(master.p) : nullptr; v.nn = nn; return v;
    }
    GatherView gview() {
        GatherView g;
        g.off = gpc(csrOff.p); g.ent = gpc(entV.p); g.realMask = gpc(realMask.p);
        g.gfc = gpc(mixed() ? cb32[cur].gfc.p : cb64[cur].gfc.p);
        g.mixed = mixed() ? 1 : 0;
        g.gcap = mixed() ? cb32[cur].cap : cb64[cur].cap;
        return g;
    }
    PartView pview() {
        PartView v; v.ptype = gpc(ptype.p); v.mask = gpc(pmask.p); v.mat = gpc(pmat.p); v.fnodes = gpc(fnodes.p); v.bmin = gpc(bmin.p);
        v.bmax = gpc(bmax.p); v.np = np; return v;
    }

    void freeAll() {
        release(ctx, posr); release(ctx, velm); release(ctx, angI); release(ctx, ori); release(ctx, force); release(ctx, torque);
        release(ctx, lastPos); release(ctx, flags); release(ctx, checked); release(ctx, ptype); release(ctx, pmat); release(ctx, binned);
        release(ctx, unb); release(ctx, pmask); release(ctx, fnodes); release(ctx, bmin); release(ctx, bmax); release(ctx, mat);
        for (int i = 0; i < 2; i++) { cb64[i].free(ctx); cb32[i].free(ctx); }
        release(ctx, csrOff); release(ctx, entV); release(ctx, realMask); release(ctx, E);
        release(ctx, master); release(ctx, nodeClump); release(ctx, cOff); release(ctx, cMem); release(ctx, cRelPos); release(ctx, cRelOri); release(ctx, cFT);
    }
};

// grows a buffer that the window graph's kernel arguments point to: the graph goes first
template <typename T>
void ensureGraphArg(Ctx* c, DBuf<T>& b, size_t n) {
    if (!b.p || b.n < n) c->sh->dropGraph(c);
    ensure(c, b, n);
}

void timingFlush(Ctx* c) {  // reads the pending event pairs into the per-kernel totals
    Timing& T = c->timing;
    for (size_t k = 0; k < T.pend.size(); k++) {
        (void)hipEventSynchronize(T.ev[2 * k + 1]);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, T.ev[2 * k], T.ev[2 * k + 1]);
        auto& a = T.acc[T.pend[k]];
        a.first += ms;
        a.second += 1;
    }
    T.pend.clear();
    (void)hipGetLastError();  // an event error must not surface at the next kernel launch
}
struct KTimer {  // HIP-event timing of one launch (or phase, timers may nest) on the context stream
    Ctx* c; const char* name; bool on; size_t k = 0;
    KTimer(Ctx* c_, const char* n) : c(c_), name(n), on(c_->timing.on) {
        if (!on) return;
        Timing& T = c->timing;
        if (T.pend.size() >= 4096 && T.open == 0) timingFlush(c);
        k = T.pend.size();  // this timer's slot, reserved now so that nested timers take later ones
        T.pend.push_back(name);
        T.open++;
        while (T.ev.size() < 2 * k + 2) {
            hipEvent_t e = nullptr;
            HIPCHK(hipEventCreate(&e));
            T.ev.push_back(e);
        }
        (void)hipEventRecord(T.ev[2 * k], c->stream);
    }
    ~KTimer() {
        if (!on) return;
        (void)hipEventRecord(c->timing.ev[2 * k + 1], c->stream);
        c->timing.open--;
    }
};

void checkLaunch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw WdmError(WDM_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
}

// temp storage for hipcub calls
void* tempFor(Ctx* c, size_t bytes) { ensure(c, c->sh->temp, bytes); return c->sh->temp.p; }

void fillU32(Ctx* c, uint32_t* a, int64_t n) {
    if (n > 0) hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n)), dim3(WDM_BLOCK), 0, c->stream, (int32_t*)a, n, 0);
}
// out[i] = sum(op(in[0..i))) for i in [0, n], on the context stream
template <typename TIn, typename TOut, class Op = ScanId>
void scanExclusive(Ctx* c, const TIn* in, TOut* out, int64_t n, Op op = Op()) {
    Shared& X = *c->sh;
    hipStream_t st = c->stream;
    const int64_t nt = std::max<int64_t>((n + SCAN_TILE - 1) / SCAN_TILE, 1);
    if (nt >= INT32_MAX) throw WdmError(WDM_E_ARG, "prefix sum too long");
    ensure(c, X.tsum, (size_t)nt);
    hipLaunchKernelGGL((k_scan_sums<TIn, Op>), dim3((unsigned)nt), dim3(WDM_BLOCK), 0, st, in, n, X.tsum.p, op);
    const int prefixed = nt > SCAN_DIRECT_TILES;
    if (prefixed) hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_TOP), 0, st, X.tsum.p, nt);
    hipLaunchKernelGGL((k_scan_tiles<TIn, TOut, Op>), dim3((unsigned)nt), dim3(WDM_BLOCK), 0, st, in, n, (const int64_t*)X.tsum.p, prefixed,
                       out, op);
    checkLaunch();
}
// the second half of a bucket sort: items scattered into btmp (n of them, or at
// most n with *total of them on the device; off the exclusive prefix of the bucket counts)
// placed in order through the writer w
#define BKT_GRID 128
template <typename OffT, class W>
void bucketFinish(Ctx* c, int64_t n, const OffT* off, const W& w, const OffT* total = nullptr) {
    Shared& X = *c->sh;
    if (n <= 0) return;
    ensure(c, X.longList, (size_t)(n / (BKT_SMALL + 1) + 1)); ensure(c, X.longCnt, 1); ensure(c, X.baux, 2 * (size_t)n);
    fillU32(c, X.longCnt.p, 1);
    hipLaunchKernelGGL((k_bucket_rank<OffT, W>), dim3(nblk(n)), dim3(WDM_BLOCK), 0, c->stream, (const BItem*)X.btmp.p, off, n, total,
                       X.longList.p, X.longCnt.p, w);
    hipLaunchKernelGGL((k_bucket_long<OffT, W>), dim3(BKT_GRID), dim3(BKT_THREADS), 0, c->stream, (const BItem*)X.btmp.p, X.baux.p, n, off,
                       (const uint32_t*)X.longList.p, (const uint32_t*)X.longCnt.p, w);
    checkLaunch();
}

// pinned host memory for readbacks (grows; the stream is drained before it is replaced)
void* pinnedRead(Ctx* c, size_t bytes) {
    Shared& X = *c->sh;
    if (bytes > X.pinCap) {
        HIPCHK(hipStreamSynchronize(c->stream));
        if (X.pin) (void)hipHostFree(X.pin);
        X.pin = nullptr;
        X.pinCap = 0;
        HIPCHK(hipHostMalloc(&X.pin, std::max<size_t>(bytes, 4096), hipHostMallocDefault));
        X.pinCap = std::max<size_t>(bytes, 4096);
    }
    return X.pin;
}
// values of a batch-wide exclusive prefix at each scene's first entry (FR[t].rowBase for
// field 0, FR[t].oldStart for field 1) and at `last`
__global__ void k_seg_starts(const FRScene* FR, int nT, int field, int64_t last, const int64_t* scan, int64_t* out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > nT) return;
    const int64_t idx = t == nT ? last : (field == 0 ? FR[t].rowBase : FR[t].oldStart);
    out[t] = scan[idx];
}
// the same read back (one synchronisation); `extraBytes` of device memory at `extra` are read
// back with them, into *extraOut
std::vector<int64_t> segmentStarts(Ctx* c, const int64_t* scan, const FRScene* FRd, int nT, int field, int64_t last,
                                   const void* extra = nullptr, void* extraOut = nullptr, size_t extraBytes = 0) {
    Shared& X = *c->sh;
    ensure(c, X.starts, (size_t)nT + 1);
    hipLaunchKernelGGL(k_seg_starts, dim3((nT + 64) / 64), dim3(64), 0, c->stream, FRd, nT, field, last, scan, X.starts.p);
    checkLaunch();
    const size_t sb = ((size_t)nT + 1) * sizeof(int64_t);
    uint8_t* h = (uint8_t*)pinnedRead(c, sb + extraBytes);
    HIPCHK(hipMemcpyAsync(h, X.starts.p, sb, hipMemcpyDeviceToHost, c->stream));
    if (extraBytes) HIPCHK(hipMemcpyAsync(h + sb, extra, extraBytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<int64_t> out(nT + 1);
    std::memcpy(out.data(), h, sb);
    if (extraBytes) std::memcpy(extraOut, h + sb, extraBytes);
    return out;
}
// segmentStarts of two prefixes (scan0 at each scene's rowBase, scan1 at its oldStart; last0,
// last1 their ends) and `extraBytes` at `extra`, with one synchronisation
void segmentStarts2(Ctx* c, const FRScene* FRd, int nT, const int64_t* scan0, int64_t last0, const int64_t* scan1, int64_t last1,
                    std::vector<int64_t>& out0, std::vector<int64_t>& out1, const void* extra, void* extraOut, size_t extraBytes) {
    Shared& X = *c->sh;
    ensure(c, X.starts, 2 * ((size_t)nT + 1));
    hipLaunchKernelGGL(k_seg_starts, dim3((nT + 64) / 64), dim3(64), 0, c->stream, FRd, nT, 0, last0, scan0, X.starts.p);
    if (scan1) hipLaunchKernelGGL(k_seg_starts, dim3((nT + 64) / 64), dim3(64), 0, c->stream, FRd, nT, 1, last1, scan1, X.starts.p + nT + 1);
    checkLaunch();
    const size_t sb = ((size_t)nT + 1) * sizeof(int64_t);
    uint8_t* h = (uint8_t*)pinnedRead(c, 2 * sb + extraBytes);
    HIPCHK(hipMemcpyAsync(h, X.starts.p, scan1 ? 2 * sb : sb, hipMemcpyDeviceToHost, c->stream));
    if (extraBytes) HIPCHK(hipMemcpyAsync(h + 2 * sb, extra, extraBytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    out0.assign(nT + 1, 0);
    out1.assign(nT + 1, 0);
    std::memcpy(out0.data(), h, sb);
    if (scan1) std::memcpy(out1.data(), h + sb, sb);
    if (extraBytes) std::memcpy(extraOut, h + 2 * sb, extraBytes);
}
// several host tables in one staged upload
struct Upload {
    Blob b;
    uint8_t* dev = nullptr;
    template <typename T>
    size_t add(const std::vector<T>& v) { return b.addv(v); }
    void send(Ctx* c, Staged& s) { dev = s.upload(c, b.b.data(), b.b.size()); }
    template <typename T>
    const T* at(size_t off) const { return (const T*)(dev + off); }
    Map map(size_t off, int n) const { Map m; m.bprefix = at<int64_t>(off); m.n = n; return m; }
};

Map uploadMap(Ctx* c, Staged& s, const std::vector<int64_t>& bprefix) {
    Map m;
    m.bprefix = (const int64_t*)s.upload(c, bprefix.data(), bprefix.size() * sizeof(int64_t));
    m.n = (int)bprefix.size() - 1;
    return m;
}
// block prefix of per-scene element counts
template <typename F>
std::vector<int64_t> blockPrefix(int n, F&& count) {
    std::vector<int64_t> bp(n + 1, 0);
    for (int t = 0; t < n; t++) bp[t + 1] = bp[t] + (int64_t)nblk(count(t));
    return bp;
}

// ---------------- Cell (core/Cell.cpp:28-120) ----------------
void cellIntegrateAndUpdate(Scene& S, double dt) {
    HM3 trsfInc = hscale(S.gradV, dt);
    S.trsf = hadd(S.trsf, hmul(trsfInc, S.trsf));
    HM3 prev = S.hSize;
    if (S.homoDeform == 4) {
        HM3 gh = hdivs(hscale(S.gradV, dt), 2.);
        S.hSize = hmul(hmul(hinv(hsub(heye(), gh)), hadd(heye(), gh)), S.hSize);
    } else S.hSize = hadd(S.hSize, hmul(trsfInc, S.hSize));
    S.pprevHsize = hscale(hadd(prev, S.hSize), .5);
    if (hdet(S.hSize) == 0) throw WdmError(WDM_E_DEGENERATE, "Cell is degenerate (zero volume).");
    for (int i = 0; i < 3; i++) {
        double s2 = (S.hSize.m[0][i] * S.hSize.m[0][i] + S.hSize.m[1][i] * S.hSize.m[1][i]) + S.hSize.m[2][i] * S.hSize.m[2][i];
        S.size[i] = std::sqrt(s2);
    }
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++)
        if (i != j && S.hSize.m[i][j] != 0) throw WdmError(WDM_E_UNSUPPORTED, "sheared periodic cells are not supported");
}
void cellSetNextGradV(Scene& S) {
    S.gradV = S.nextGradV;
    HM3 W = hscale(hsub(S.gradV, htrans(S.gradV)), .5);
    leviCivitaH(W, S.spinVec);
    for (double& x : S.spinVec) x = x * .5;
}

void makeCellArgs(Scene& S, CellArgs& ca) {
    std::memset(&ca, 0, sizeof(ca));
    const double dt = S.P.dt;
    HM3 pprevL = S.gradV, nnextL = S.nextGradV;
    HM3 ImL = hzero(), IpL = hzero(), LmL = hzero();
    double dsv[3] = {0, 0, 0};
    if (S.periodic && S.homoDeform == 4) {
        ImL = hinv(hsub(heye(), hdivs(hscale(hadd(nnextL, pprevL), dt), 4.)));
        IpL = hadd(heye(), hdivs(hscale(hadd(nnextL, pprevL), dt), 4.));
        LmL = hsub(nnextL, pprevL);
        double l1[3], l2[3];
        leviCivitaH(hscale(hsub(pprevL, htrans(pprevL)), .5), l1);
        leviCivitaH(hscale(hsub(nnextL, htrans(nnextL)), .5), l2);
        for (int k = 0; k < 3; k++) dsv[k] = l1[k] * -.5 + l2[k] * .5;
    }
    ca.ImLL4hInv = toM3(ImL); ca.IpLL4h = toM3(IpL); ca.LmL = toM3(LmL);
    ca.dGradV = toM3(hsub(nnextL, pprevL));
    ca.midGradV = toM3(hscale(hadd(pprevL, nnextL), .5));
    ca.nextGradV = toM3(nnextL); ca.gradV = toM3(pprevL);
    for (int k = 0; k < 3; k++) { ca.deltaSpinVec[k] = dsv[k]; ca.spinVec[k] = S.spinVec[k]; ca.cellSize[k] = S.size[k]; }
    ca.hSize = toM3(S.hSize);
    ca.gatherHSize = toM3(S.hSizeCL);
    if (S.periodic && S.homoDeform == 4) ca.shiftVelM = toM3(hmul(S.gradV, S.pprevHsize));
    else if (S.periodic && (S.homoDeform == 2 || S.homoDeform == 3)) ca.shiftVelM = toM3(hmul(S.gradV, S.hSize));
    else ca.shiftVelM = toM3(hzero());
}

// nonzero tag of a step, written into the Verlet trigger word by a triggering Leapfrog
uint32_t stampOf(int64_t step) { return (uint32_t)(step % 0x7fffffffLL) + 1u; }

StepArgs makeStepArgs(Scene& S) {
    StepArgs a;
    std::memset(&a, 0, sizeof(a));
    const wdm_params& P = S.P;
    a.dt = P.dt; a.damping = P.damping;
    for (int k = 0; k < 3; k++) a.g[k] = P.gravity[k];
    a.hasGravity = !(P.gravity[0] == 0 && P.gravity[1] == 0 && P.gravity[2] == 0);
    a.periodic = S.periodic;
    a.homoDeform = S.periodic ? S.homoDeform : -1;
    a.trackEnergy = P.trackEnergy;
    a.reallyTrackEnergy = P.trackEnergy && (!S.periodic || S.step > 0 || hequal(S.gradV, heye()));
    a.verletDist = S.verletDist;
    a.verletD2 = S.verletDist * S.verletDist;
    a.checkVerlet = S.lastLeapfrogVerletChecked ? 1 : 0;
    a.stamp = stampOf(S.step);
    a.gatherOn = S.forcesFromContacts ? 1 : 0;
    a.hasClumps = S.nClumps > 0 ? 1 : 0;
    a.distFactor = P.distFactor;
    a.law = P.law; a.noSlip = P.noSlip; a.noBreak = P.noBreak; a.noFrict = P.noFrict; a.noAttraction = P.noAttraction;
    a.trsfRenorm = P.trsfRenorm;
    a.step = S.step;
    a.loneMask = P.loneMask;
    a.poisson = P.poisson; a.gamma = P.gamma;
    double en = P.en;
    if (en > 0 && en < 1.) {
        double le = std::log(en);
        const double PI = 3.14159265358979323846;
        a.viscAlpha = -std::sqrt(5) * le / (std::sqrt(le * le + PI * PI));
        a.hertzDamped = 1;
    } else { a.viscAlpha = 0; a.hertzDamped = 0; }
    a.mat = gpc(S.mat.p);
    a.cell = nullptr;
    return a;
}

SceneDesc makeDesc(Scene& S) {
    SceneDesc d;
    std::memset(&d, 0, sizeof(d));
    d.a = makeStepArgs(S);
    d.n = S.nview();
    d.p = S.pview();
    d.gv = S.gview();
    d.binned = gpc((const int32_t*)S.binned.p);
    d.unb = gpc((const int32_t*)S.unb.p);
    d.nbinned = (int)S.binnedH.size();
    d.nunb = (int)S.unbH.size();
    d.E = gpc(S.E.p);
    d.words = S.ctx->sh->words.p + 4 * (size_t)S.slot;
    if (S.nClumps > 0) {
        d.cl.off = gpc((const int32_t*)S.cOff.p);
        d.cl.mem = gpc((const int32_t*)S.cMem.p);
        d.cl.relPos = gpc((const double4*)S.cRelPos.p);
        d.cl.relOri = gpc((const double4*)S.cRelOri.p);
        d.cl.nodeClump = gpc((const int32_t*)S.nodeClump.p);
        d.cl.ft = gpc(S.cFT.p);
        d.cl.nClumps = S.nClumps;
        d.cl.nMembers = S.nMembers;
    }
    return d;
}

void resolveVerlet(Scene& S) {
    if (S.verletResolved) return;
    double vd = S.P.verletDist;
    if (vd < 0) {
        double minR = INFINITY;
        for (int p = 0; p < S.np; p++) if ((S.ptypeH[p] & 15) == WDM_SHAPE_SPHERE) minR = std::min(S.radiusH[p], minR);
        vd = std::isinf(minR) ? 0 : std::fabs(vd) * minR;
    }
    S.verletDist = vd;
    S.verletResolved = true;
}

// the scenes one engine call works on, with their descriptor table on the device
struct Batch {
    Ctx* c;
    std::vector<Scene*> ss;
    const SceneDesc* D = nullptr;
    Batch(Ctx* c_, std::vector<Scene*> s) : c(c_), ss(std::move(s)) {}
    int n() const { return (int)ss.size(); }
    // descriptor + cell tables of the scenes' current state (one upload)
    void upload() {
        Shared& X = *c->sh;
        int n = (int)ss.size();
        std::vector<SceneDesc> d(n);
        std::vector<CellArgs> ca(n);
        for (int i = 0; i < n; i++) { d[i] = makeDesc(*ss[i]); makeCellArgs(*ss[i], ca[i]); }
        Blob b;
        size_t od = b.addv(d), oc = b.addv(ca);
        uint8_t* base = X.stDesc.prepare(c, b.b.size());
        SceneDesc* dh = b.at<SceneDesc>(od);
        for (int i = 0; i < n; i++) dh[i].a.cell = gpc((const CellArgs*)(base + oc) + i);
        X.stDesc.upload(c, b.b.data(), b.b.size());
        D = (const SceneDesc*)(base + od);
    }
    void ensureDesc() { if (!D) upload(); }
};

// the reference's exception for a device error word (IdealElPl.cpp:76-80, InsertionSortCollider.cpp:945-949)
int deviceErrorCode(uint32_t e) { return (e & ERR_NAN) ? WDM_E_NAN : ((e & ERR_OVERSIZE) ? WDM_E_OVERSIZE : WDM_E_HIP); }
const char* deviceErrorText(uint32_t e) {
    if (e & ERR_NAN) return "NaN force in contact";
    if (e & ERR_OVERSIZE) return "Particle larger than half of the cell size encountered.";
    return "device error";
}
// A scene whose kernels flagged an error stops, as a Woo scene's run thread stops on an
// exception (core/Scene.cpp:90-95): it takes no more steps and keeps its state for reading
// until it is freed; the other scenes of a batch go on.  Marks the scenes of `ss` whose words
// (last copied to the host) carry an error; returns those newly stopped.
std::vector<Scene*> markFailed(Ctx* c, const std::vector<Scene*>& ss) {
    std::vector<Scene*> out;
    for (Scene* S : ss) {
        const uint32_t e = c->sh->hostWords[4 * (size_t)S->slot + 2];
        if (e && !S->failed) {
            S->failed = deviceErrorCode(e);
            S->failText = deviceErrorText(e);
            out.push_back(S);
        }
    }
    return out;
}
// copy every scene's words to the host (synchronising), mark the scenes of `ss` with a device
// error as stopped and, with `raise`, raise the first one's error
std::vector<Scene*> readWords(Ctx* c, const std::vector<Scene*>& ss, bool raise = true) {
    Shared& X = *c->sh;
    size_t n = 4 * (size_t)X.nslots;
    if (n) HIPCHK(hipMemcpyAsync(X.hostWords, X.words.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<Scene*> f = markFailed(c, ss);
    if (raise)
        for (Scene* S : ss)
            if (S->failed) throw WdmError(S->failed, S->failText);
    return f;
}

// ---------------- Leapfrog::run (pkg/dem/Leapfrog.cpp:103-255) ----------------
// The tables of one Leapfrog run over a batch (scenes per gather mode, clump scenes,
// block prefixes), packed into a Blob; launched once per step, or once per step of a
// window with that step's StepVar rows.
struct LeapPlan {
    std::vector<int32_t> idx[2], cfIdx, cmIdx;
    std::vector<int64_t> bp[2], cfBp, cmBp;
    size_t oIdx[2] = {0, 0}, oBp[2] = {0, 0}, oCfIdx = 0, oCfBp = 0, oCmIdx = 0, oCmBp = 0;
    bool anyEnergy = false;
    const uint8_t* base = nullptr;

    void build(Batch& B, Blob& b) {
        Ctx* c = B.c;
        const int n = B.n();
        for (int i = 0; i < n; i++) {
            Scene* S = B.ss[i];
            resolveVerlet(*S);
            S->lastLeapfrogVerletChecked = (!S->needFull && S->verletDist > 0);
            anyEnergy |= S->P.trackEnergy != 0;
            idx[S->forcesFromContacts ? 1 : 0].push_back(i);
            // clump masters' sums only where Leapfrog reads them (free DOFs, damping dissipation)
            if (S->nClumps > 0 && (S->clumpFree || (S->P.trackEnergy && S->P.damping != 0))) cfIdx.push_back(i);
            if (S->nMembers > 0) cmIdx.push_back(i);
        }
        for (int g = 0; g < 2; g++) {
            bp[g] = blockPrefix((int)idx[g].size(), [&](int t) { return B.ss[idx[g][t]]->nn; });
            oIdx[g] = b.addv(idx[g]);
            oBp[g] = b.addv(bp[g]);
        }
        cfBp = blockPrefix((int)cfIdx.size(), [&](int t) { return B.ss[cfIdx[t]]->nClumps; });
        cmBp = blockPrefix((int)cmIdx.size(), [&](int t) { return B.ss[cmIdx[t]]->nMembers; });
        oCfIdx = b.addv(cfIdx); oCfBp = b.addv(cfBp);
        oCmIdx = b.addv(cmIdx); oCmBp = b.addv(cmBp);
        ensureGraphArg(c, c->sh->epartL, (size_t)(bp[0].back() + bp[1].back()) * 3 + 3);
    }
    Map map(size_t o, size_t n) const {
        Map m;
        m.bprefix = (const int64_t*)(base + o);
        m.n = (int)n;
        return m;
    }
    void sig(std::vector<uint64_t>& k) const {
        for (int g = 0; g < 2; g++) k.insert(k.end(), {idx[g].size(), (uint64_t)bp[g].back(), oIdx[g], oBp[g]});
        k.insert(k.end(), {cfIdx.size(), (uint64_t)cfBp.back(), oCfIdx, oCfBp, cmIdx.size(), (uint64_t)cmBp.back(), oCmIdx, oCmBp,
                           (uint64_t)anyEnergy});
    }
    void launch(Ctx* c, const SceneDesc* D, const StepVar* sv) const {
        Shared& X = *c->sh;
        hipStream_t st = c->stream;
        // clump masters' force and torque with their members' (read by the Leapfrog below)
        if (cfBp.back() > 0) {
            hipLaunchKernelGGL(k_clump_forces, dim3((unsigned)cfBp.back()), dim3(WDM_BLOCK), 0, st, D, (const int32_t*)(base + oCfIdx),
                               map(oCfBp, cfIdx.size()), sv);
            checkLaunch();
        }
        const int64_t nb0 = bp[0].back();
        for (int g = 0; g < 2; g++) {
            const int ng = (int)idx[g].size();
            if (ng == 0) continue;
            const Map m = map(oBp[g], ng);
            const int32_t* id = (const int32_t*)(base + oIdx[g]);
            double* ep = X.epartL.p + (g ? 3 * nb0 : 0);
            const int64_t nbk = bp[g].back();
            if (nbk > 0) {
                KTimer t(c, "leapfrog");
                const dim3 grid((unsigned)xcdChunkGrid(nbk, XCD_CHUNK_LEAP));
                if (g) hipLaunchKernelGGL(k_leapfrog<1>, grid, dim3(WDM_BLOCK), 0, st, D, id, m, ep, sv);
                else hipLaunchKernelGGL(k_leapfrog<0>, grid, dim3(WDM_BLOCK), 0, st, D, id, m, ep, sv);
                checkLaunch();
            }
            if (anyEnergy) {
                hipLaunchKernelGGL(k_energy_finalize, dim3(3 * ng), dim3(WDM_BLOCK), 0, st, D, id, 1, m, (const double*)ep, 3, 0, sv);
                checkLaunch();
            }
        }
        // clump members follow their masters (ClumpData::applyToMembers)
        if (cmBp.back() > 0) {
            KTimer t(c, "clump_members");
            hipLaunchKernelGGL(k_clump_members, dim3((unsigned)cmBp.back()), dim3(WDM_BLOCK), 0, st, D, (const int32_t*)(base + oCmIdx),
                               map(oCmBp, cmIdx.size()), sv);
            checkLaunch();
        }
    }
};
void materialiseForces(Scene& S, bool sums);
struct Leapfrog {
    // uploads the batch's descriptors and the plan's tables, then launches; afterwards the node
    // forces are the reset values (Leapfrog.cpp:240-245) until a ContactLoop adds its sums
    static void run(Batch& B) {
        Ctx* c = B.c;
        for (Scene* S : B.ss)
            if (S->forcesReset) {  // a second Leapfrog without a ContactLoop reads the reset forces
                materialiseForces(*S, false);
                S->forcesReset = false;
            }
        LeapPlan lp;
        Blob b;
        lp.build(B, b);
        B.upload();
        lp.base = c->sh->stMapNode.upload(c, b.b.data(), b.b.size());
        lp.launch(c, B.D, nullptr);
        for (Scene* S : B.ss) {
            S->forcesFromContacts = false;
            S->forcesReset = true;
        }
    }
};

// ---------------- InsertionSortCollider::run (pkg/dem/InsertionSortCollider.cpp:445-646) ----------------
struct InsertionSortCollider {
    // Grid levels of one scene (e: its extent reduction, hs: its histogram of bound extents):
    // level l of nominal cell width W / 2^(nLvl-1-l), W the largest sphere bound times 1.001
    // and the cell factor; the number of levels is the one with the smallest estimate of
    // the pair search's work (box tests and cell visits, from the bound counts per level)
    // whose cells stay within max(4096, 2 nb) (levels > 0 forces that many).  A single level
    // that does not fit widens its cells.
    static void gridLevels(const Scene& S, const double* e, const uint32_t* hs, int nb, double cellFactor, int levels, FRScene& F) {
        double ext = e[7] > 0 ? e[7] : e[6];
        if (!(ext > 0) || !std::isfinite(ext)) ext = 1e-12;
        const double W = ext * 1.001 * cellFactor;
        const double maxCells = (double)std::max<int64_t>(4096, 2 * (int64_t)nb);
        // cells per axis and per-axis widths of a level of nominal width w; the total cells
        auto dims = [&](double w, int n[3], double wd[3]) -> double {
            double tot = 1;
            for (int d = 0; d < 3; d++) {
                double k;
                if (S.periodic) {
                    k = std::floor(S.size[d] / w);
                    if (!(k >= 3)) k = 1;
                } else {
                    k = std::floor((e[3 + d] - e[d]) / w) + 1;
                    if (!(k >= 1)) k = 1;
                }
                k = std::min(k, 1e9);
                n[d] = (int)k;
                wd[d] = S.periodic ? S.size[d] / k : w;
                tot *= k;
            }
            return tot;
        };
        double V = 1;
        for (int d = 0; d < 3; d++) V *= S.periodic ? S.size[d] : (e[3 + d] - e[d]) + W;
        V = std::max(V, 1e-300);
        const double VISIT = 1.5;  // a visited cell, in box tests
        static const std::vector<double> edge = [] {  // bin k holds extents in [edge[k], edge[k + 1])
            std::vector<double> v(EXT_BINS + 1);
            for (int k = 0; k <= EXT_BINS; k++) v[k] = std::pow(2., (k - 128) / 4.);
            return v;
        }();
        int best = 0;
        double bestCost = 0;
        const int lo = levels > 0 ? std::min(levels, MAXLVL) : 1, hi = levels > 0 ? std::min(levels, MAXLVL) : MAXLVL;
        for (int nL = lo; nL <= hi; nL++) {
            double w[MAXLVL], cnt[MAXLVL] = {0, 0, 0, 0}, cells = 0, cvol[MAXLVL];
            for (int l = 0; l < nL; l++) {
                w[l] = W / std::ldexp(1., nL - 1 - l);
                int n[3];
                double wd[3];
                cells += dims(w[l], n, wd);
                cvol[l] = wd[0] * wd[1] * wd[2];
            }
            if (cells > maxCells) continue;
            for (int k = 0; k < EXT_BINS; k++) {
                if (!hs[k]) continue;
                const double bl = edge[k], bh = edge[k + 1];  // the bin's extents
                if (bl > W / 1.001) continue;  // large: listed apart
                int l = 0;
                while (l < nL - 1 && bh > w[l] / 1.001) l++;
                cnt[l] += (double)hs[k] * EXT_SAMPLE;
            }
            double cost = 0;
            for (int l = 0; l < nL; l++) {
                cost += cnt[l] * (13.5 * cnt[l] / V * cvol[l] + 14 * VISIT);
                for (int l2 = 0; l2 < l; l2++) {
                    const double s = w[l] + 2 * w[l2], r = w[l] / w[l2] + 2;
                    cost += cnt[l] * (cnt[l2] / V * s * s * s + VISIT * r * r * r);
                }
            }
            if (!best || cost < bestCost) { best = nL; bestCost = cost; }
        }
        const int nL = best ? best : 1;
        for (int l = 0; l < nL; l++) {
            double wl = W / std::ldexp(1., nL - 1 - l);
            if (!best) {  // one level, widened until it fits
                for (int it = 0; it < 200; it++) {
                    int n[3];
                    double wd[3];
                    if (dims(wl, n, wd) <= maxCells) break;
                    wl *= 1.26;
                }
            }
            double wd[3];
            dims(wl, F.ln[l], wd);
            for (int d = 0; d < 3; d++) F.lw[l][d] = wd[d];
            F.lthr[l] = wl / 1.001;
        }
        F.nLvl = nL;
    }

    // Grids of the scenes L[t] from their bound extents, then the bound entries of all of
    // them sorted by (scene, cell key, id) with each cell's entry range [cbeg[k], cbeg[k + 1])
    // (levels: 0 lets each scene choose).  The extents of this run are reduced on the device
    // and copied to the pinned mailbox X.extMail.  lagged: the grids come from the extents of
    // the scenes' previous full run (Scene::extLast) and nothing waits for this run's; a bound
    // wider than its grid's top level in a scene whose large bounds are not listed
    // (largeListed 0) would lose its pairs, so k_cell_count raises *planFlag for it and the
    // caller lays the grids out again from this run's extents.  Otherwise the host waits for
    // this run's extents (one synchronisation).  A bound outside a lagged aperiodic grid falls
    // into its edge cells (cell coordinates are clamped), which keeps every pair.
    static void buildGrids(Batch& B, const std::vector<int>& L, std::vector<FRScene>& F, const std::vector<int64_t>& bpB, const Map& mB,
                           const FRScene*& FRd, int64_t& nbTot, double cellFactor, int levels, bool entryCopies = true,
                           bool lagged = false, uint32_t* planFlag = nullptr, int largeListed = 0) {
        Ctx* c = B.c;
        Shared& X = *c->sh;
        hipStream_t st = c->stream;
        int nT = (int)L.size();
        const int64_t nblkB = bpB[nT];
        ensure(c, X.extPart, (size_t)nblkB * 8 + 8);
        ensure(c, X.extOut, (size_t)nT * 8);
        ensure(c, X.hist, (size_t)nT * EXT_BINS);
        const size_t eb = (size_t)nT * 8 * sizeof(double), hb = (size_t)nT * EXT_BINS * sizeof(uint32_t);
        if (X.extMailCap < eb + hb) {  // grown before this run's copy (the stream is drained first)
            HIPCHK(hipStreamSynchronize(st));
            if (X.extMail) (void)hipHostFree(X.extMail);
            X.extMail = nullptr;
            X.extMailCap = 0;
            const size_t cap = std::max<size_t>(2 * (eb + hb), 4096);
            HIPCHK(hipHostMalloc((void**)&X.extMail, cap, hipHostMallocDefault));
            X.extMailCap = cap;
        }
        fillU32(c, X.hist.p, (int64_t)nT * EXT_BINS);
        if (nblkB > 0) {
            hipLaunchKernelGGL(k_extent, dim3((unsigned)nblkB), dim3(WDM_BLOCK), 0, st, B.D, FRd, mB, X.extPart.p, X.hist.p);
            checkLaunch();
        }
        hipLaunchKernelGGL(k_extent_finalize, dim3(nT), dim3(WDM_BLOCK), 0, st, mB, (const double*)X.extPart.p, X.extOut.p);
        checkLaunch();
        HIPCHK(hipMemcpyAsync(X.extMail, X.extOut.p, eb, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(X.extMail + eb, X.hist.p, hb, hipMemcpyDeviceToHost, st));
        std::vector<double> h((size_t)nT * 8);
        std::vector<uint32_t> hh((size_t)nT * EXT_BINS);
        if (lagged) {
            for (int t = 0; t < nT; t++) {
                const Scene& S = *B.ss[L[t]];
                std::memcpy(&h[(size_t)t * 8], S.extLast, sizeof(S.extLast));
                std::memcpy(&hh[(size_t)t * EXT_BINS], S.histLast.data(), EXT_BINS * sizeof(uint32_t));
            }
        } else {
            HIPCHK(hipStreamSynchronize(st));
            std::memcpy(h.data(), X.extMail, eb);
            std::memcpy(hh.data(), X.extMail + eb, hb);
        }
        int64_t cellBase = 0;
        nbTot = 0;
        for (int t = 0; t < nT; t++) {
            Scene& S = *B.ss[L[t]];
            FRScene& Ft = F[t];
            GridArgs& g = Ft.g;
            std::memset(&g, 0, sizeof(g));
            g.periodic = S.periodic;
            for (int d = 0; d < 3; d++) { g.L[d] = S.size[d]; g.invL[d] = 1. / S.size[d]; }
            Ft.nLvl = 1;
            for (int d = 0; d < 3; d++) { Ft.ln[0][d] = 1; Ft.lw[0][d] = 1; }
            Ft.lthr[0] = INFINITY;
            const int nb = (int)S.binnedH.size();
            const double* e = &h[(size_t)t * 8];
            if (nb > 0) gridLevels(S, e, &hh[(size_t)t * EXT_BINS], nb, cellFactor, levels, Ft);
            double span = 0;
            for (int d = 0; d < 3; d++) {
                g.n[d] = Ft.ln[0][d];
                g.w[d] = Ft.lw[0][d];
                g.origin[d] = (S.periodic || nb == 0) ? 0. : e[d];
                g.Lf[d] = (float)g.L[d];
                g.invLf[d] = (float)g.invL[d];
                span = std::max(span, S.periodic ? g.L[d] : Ft.ln[Ft.nLvl - 1][d] * Ft.lw[Ft.nLvl - 1][d]);
            }
            g.pfMargin = (float)(4e-6 * span);
            Ft.lbase[0] = 0;
            int64_t nc = 0;
            for (int l = 0; l < Ft.nLvl; l++) {
                nc += (int64_t)Ft.ln[l][0] * Ft.ln[l][1] * Ft.ln[l][2];
                if (nc >= INT32_MAX) throw WdmError(WDM_E_ARG, "collider grid too large (2^31 cells)");
                Ft.lbase[l + 1] = (int32_t)nc;
            }
            Ft.cellBase = cellBase;
            Ft.binStart = nbTot;
            Ft.nCells = (int32_t)nc;
            Ft.largePairs = S.facetMasksPair(S.P.loneMask) ? 1 : 0;
            Ft.largeListed = largeListed;
            cellBase += nc + 1;  // + the list of large bounds
            nbTot += nb;
        }
        if (cellBase >= INT32_MAX || nbTot >= INT32_MAX) throw WdmError(WDM_E_ARG, "collider batch too large (2^31 cells or bounds)");
        FRd = (const FRScene*)X.stFR.upload(c, F.data(), F.size() * sizeof(FRScene));
        const int64_t G = cellBase;
        const size_t nb1 = (size_t)std::max<int64_t>(nbTot, 1);
        ensure(c, X.ekey, nb1); ensure(c, X.eslot, nb1); ensure(c, X.ckeys2, nb1); ensure(c, X.cvals2, nb1); ensure(c, X.cbox, nb1);
        ensure(c, X.btmp, nb1);
        ensure(c, X.ccnt, (size_t)G); ensure(c, X.cbeg, (size_t)G + 1);
        fillU32(c, X.ccnt.p, G);
        if (nblkB > 0) {
            hipLaunchKernelGGL(k_cell_count, dim3((unsigned)nblkB), dim3(WDM_BLOCK), 0, st, B.D, FRd, mB, X.ekey.p, X.eslot.p, X.ccnt.p,
                               planFlag);
            checkLaunch();
        }
        scanExclusive(c, (const uint32_t*)X.ccnt.p, X.cbeg.p, G);
        if (nblkB > 0) {
            hipLaunchKernelGGL(k_cell_scatter, dim3((unsigned)nblkB), dim3(WDM_BLOCK), 0, st, B.D, FRd, mB, (const uint32_t*)X.ekey.p,
                               (const uint32_t*)X.eslot.p, (const int32_t*)X.cbeg.p, X.btmp.p);
            checkLaunch();
        }
        if (entryCopies) { ensure(c, X.eb0, nb1); ensure(c, X.eb1, nb1); ensure(c, X.einfo, nb1); }
        bucketFinish(c, nbTot, (const int32_t*)X.cbeg.p,
                     CellOut{B.D, FRd, nT, X.cvals2.p, X.ckeys2.p, X.cbox.p, entryCopies ? X.eb0.p : nullptr, X.eb1.p, X.einfo.p});
    }

    // list every large bound (facets wider than the top level's cells) in the cells it
    // covers at every level; false when the batch has none (one synchronisation for the
    // listing count)
    static bool listLarge(Batch& B, const std::vector<int64_t>& bpB, const Map& mB, const FRScene* FRd, int64_t nbTot, int64_t G) {
        Ctx* c = B.c;
        Shared& X = *c->sh;
        hipStream_t st = c->stream;
        const int nT = (int)bpB.size() - 1;
        if (bpB[nT] == 0) return false;
        ensure(c, X.lcntE, (size_t)nbTot); ensure(c, X.loffE, (size_t)nbTot + 1);
        hipLaunchKernelGGL(k_large_cells, dim3((unsigned)bpB[nT]), dim3(WDM_BLOCK), 0, st, B.D, FRd, mB, (const uint32_t*)X.ckeys2.p,
                           (const int32_t*)X.cvals2.p, X.lcntE.p);
        checkLaunch();
        scanExclusive(c, (const uint32_t*)X.lcntE.p, X.loffE.p, nbTot);
        int64_t nl = 0;
        int64_t* pin = (int64_t*)pinnedRead(c, sizeof(int64_t));
        HIPCHK(hipMemcpyAsync(pin, X.loffE.p + nbTot, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        nl = *pin;
        if (nl == 0) return false;
        if (nl >= INT32_MAX) throw WdmError(WDM_E_ARG, "too many large-bound cell entries");
        ensure(c, X.lkeys, (size_t)nl); ensure(c, X.lids, (size_t)nl); ensure(c, X.lslot, (size_t)nl); ensure(c, X.lvals, (size_t)nl);
        ensure(c, X.lcnt, (size_t)G); ensure(c, X.lbeg, (size_t)G + 1);
        fillU32(c, X.lcnt.p, G);
        hipLaunchKernelGGL(k_large_list, dim3(nblk(nl)), dim3(WDM_BLOCK), 0, st, B.D, FRd, nT, (const int32_t*)X.cvals2.p, nbTot,
                           (const int64_t*)X.loffE.p, nl, X.lkeys.p, X.lids.p, X.lslot.p, X.lcnt.p);
        checkLaunch();
        scanExclusive(c, (const uint32_t*)X.lcnt.p, X.lbeg.p, G);
        hipLaunchKernelGGL(k_large_scatter, dim3(nblk(nl)), dim3(WDM_BLOCK), 0, st, (const uint32_t*)X.lkeys.p, (const int32_t*)X.lids.p,
                           (const uint32_t*)X.lslot.p, nl, (const int32_t*)X.lbeg.p, X.lvals.p);
        checkLaunch();
        return true;
    }

    template <typename T>
    static ConBuf<T>* bufs(Scene& S);

    template <typename T>
    static void fullRun(Batch& B, const std::vector<int>& L) {
        Ctx* c = B.c;
        Shared& X = *c->sh;
        hipStream_t st = c->stream;
        const int nT = (int)L.size();
        if (nT == 0) return;
        KTimer tAll(c, "full_run");  // the whole pass on the stream, host synchronisations included
        auto S_ = [&](int t) -> Scene& { return *B.ss[L[t]]; };

        // per-particle rows of the batch: scene t's particles p at rowBase + p (and one row
        // past its last particle)
        std::vector<FRScene> F(nT);
        std::memset(F.data(), 0, F.size() * sizeof(FRScene));
        int64_t nRows = 0;
        for (int t = 0; t < nT; t++) {
            F[t].desc = L[t];
            F[t].rowBase = nRows;
            nRows += (int64_t)S_(t).np + 1;
        }
        if (nRows >= INT32_MAX) throw WdmError(WDM_E_ARG, "collider batch too large (2^31 particles)");
        std::vector<int64_t> bpP = blockPrefix(nT, [&](int t) { return S_(t).np; });
        std::vector<int64_t> bpB = blockPrefix(nT, [&](int t) { return (int64_t)S_(t).binnedH.size(); });
        Upload uA;
        const size_t oFA = uA.add(F), oP = uA.add(bpP), oB = uA.add(bpB);
        uA.send(c, X.stRunA);
        const FRScene* FRd = uA.at<FRScene>(oFA);
        const Map mP = uA.map(oP, nT), mB = uA.map(oB, nT);
        if (bpP[nT] > 0) {
            KTimer tm(c, "bounds");
            hipLaunchKernelGGL(k_bounds, dim3((unsigned)bpP[nT]), dim3(WDM_BLOCK), 0, st, B.D, FRd, mP, 0);
            checkLaunch();
        }
        // old contact lists (kept only because they are real, or matched by a candidate)
        std::vector<ConDesc<T>> C(nT);
        std::memset(C.data(), 0, C.size() * sizeof(ConDesc<T>));
        int64_t nOldTot = 0;
        for (int t = 0; t < nT; t++) {
            Scene& S = S_(t);
            ConBuf<T>* cb = bufs<T>(S);
            if (cb[S.cur].cap == 0) cb[S.cur].reserve(c, 1);
            C[t].cv = cb[S.cur].view();
            C[t].nc = S.nc;
            C[t].desc = L[t];
            F[t].oldStart = nOldTot;
            F[t].nOld = S.nc;
            nOldTot += S.nc;
        }
        std::vector<int64_t> bpO = blockPrefix(nT, [&](int t) { return F[t].nOld; });
        std::vector<int64_t> bp1 = blockPrefix(nT, [&](int t) { return (int64_t)S_(t).np + 1; });
        ensure(c, X.ooFlag, (size_t)std::max<int64_t>(nOldTot, 1));
        ensure(c, X.ooRank, (size_t)nOldTot + 1);
        ensure(c, X.oldRow, (size_t)nRows + 1);
        bool anyFacets = false;
        for (int t = 0; t < nT; t++) anyFacets |= S_(t).hasFacets;
        // the grids come from the extents of the scenes' last full run when every scene has
        // them (no synchronisation); wmax[2] reports a grid that was laid out too small
        bool lagged = true;
        for (int t = 0; t < nT; t++) lagged &= S_(t).extValid && !S_(t).needFull;
        ensure(c, X.wmax, 4);
        fillU32(c, (uint32_t*)X.wmax.p, 3);

        // candidate pairs: one thread per bound entry, each wavefront writing into its own
        // region and counting its stored pairs per row (scene, min id); the pairs are then
        // bucket-sorted by row and max id, so that they come in (scene, contact key) order,
        // and the old contacts that no candidate matches are flagged, before the host reads
        // the counts (one synchronisation); a wavefront that found more pairs than its region
        // holds makes the host run it all again with larger regions, and a lagged grid laid
        // out for smaller bounds makes it lay the grids out again from this run's extents
        int64_t ncandTot = 0;
        std::vector<int64_t> noor(nT, 0);
        const ConDesc<T>* Cd = nullptr;
        Map mO, m1;
        int64_t nbTot = 0;
        for (int attempt = 0;; attempt++) {
            {
                KTimer tg(c, "grid");
                buildGrids(B, L, F, bpB, mB, FRd, nbTot, c->cellFactor, c->gridLevels, true, lagged, (uint32_t*)X.wmax.p + 2,
                           anyFacets ? 1 : 0);
            }
            const bool sweep = c->broadphase == 1;
            const bool large = !sweep && anyFacets && listLarge(B, bpB, mB, FRd, nbTot, F[nT - 1].cellBase + F[nT - 1].nCells + 1);
            if (sweep && nbTot > 0) {  // the entries of each scene by the minimum x of their boxes
                KTimer tm(c, "sweep_sort");
                const size_t nb1 = (size_t)nbTot;
                ensure(c, X.swK, nb1); ensure(c, X.swK2, nb1); ensure(c, X.swV, nb1); ensure(c, X.swV2, nb1);
                hipLaunchKernelGGL(k_sweep_keys, dim3((unsigned)bpB[nT]), dim3(WDM_BLOCK), 0, st, B.D, FRd, mB, (const float4*)X.cbox.p,
                                   X.swK.p, X.swV.p);
                checkLaunch();
                size_t tb = 0;
                const int hiBit = 32 + bitsFor((uint64_t)nT);
                HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, X.swK.p, X.swK2.p, X.swV.p, X.swV2.p, (int)nbTot, 0, hiBit, st));
                void* tmp = tempFor(c, tb);
                HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, X.swK.p, X.swK2.p, X.swV.p, X.swV2.p, (int)nbTot, 0, hiBit, st));
            }
            Upload uO;
            const size_t oFO = uO.add(F), oCO = uO.add(C), oO = uO.add(bpO), o1 = uO.add(bp1);
            uO.send(c, X.stRunO);
            FRd = uO.at<FRScene>(oFO);
            Cd = uO.at<ConDesc<T>>(oCO);
            mO = uO.map(oO, nT);
            m1 = uO.map(o1, nT);  // per scene its rows (np + 1 threads)
            {
                KTimer to(c, "old_rows");
                hipLaunchKernelGGL(k_old_rows<T>, dim3((unsigned)bp1[nT]), dim3(WDM_BLOCK), 0, st, B.D, Cd, FRd, m1, X.oldRow.p);
                checkLaunch();
            }
            const int64_t nBlk = bpB[nT];  // 4 wavefronts per block
            const int64_t nW = 4 * nBlk;
            ensure(c, X.wcnt, (size_t)std::max<int64_t>(nW, 1)); ensure(c, X.scnt, (size_t)std::max<int64_t>(nW, 1));
            ensure(c, X.rowCnt, (size_t)nRows); ensure(c, X.rowOff, (size_t)nRows + 1);
            std::vector<int64_t> cs, os;
            auto grow = [](int32_t wm) { return (int32_t)std::min<int64_t>(((int64_t)wm + wm / 8 + 63) / 64 * 64, INT32_MAX / 2); };
            const int32_t scap = c->scanCap, wcap = c->waveCap;
            const size_t nreg = (size_t)std::max<int64_t>(nW * wcap, 1);
            ensure(c, X.sPair, (size_t)std::max<int64_t>(nW * scap, 1));
            ensure(c, X.rItem, nreg); ensure(c, X.rRow, nreg); ensure(c, X.rSlot, nreg);
            ensure(c, X.candK, nreg); ensure(c, X.candP, nreg); ensure(c, X.btmp, nreg);
            if (attempt > 0) fillU32(c, (uint32_t*)X.wmax.p, 2);
            fillU32(c, X.rowCnt.p, nRows);
            if (nW > 0) {
                KTimer tm(c, "pairs");
                const int64_t pad = 8 * PAIR_CHUNK;
                const dim3 grid((unsigned)((nBlk + pad - 1) / pad * pad));
                if (sweep)
                    hipLaunchKernelGGL(k_pair_sweep, dim3((unsigned)nBlk), dim3(WDM_BLOCK), 0, st, B.D, FRd, mB, (const uint32_t*)X.swV2.p,
                                       (const float4*)X.cbox.p, X.scnt.p, X.wmax.p, scap, X.sPair.p);
                else
                    hipLaunchKernelGGL(k_pair_scan, grid, dim3(WDM_BLOCK), 0, st, B.D, FRd, mB, (const uint32_t*)X.ckeys2.p,
                                       (const int32_t*)X.cbeg.p, (const float4*)X.cbox.p, X.scnt.p, X.wmax.p, scap, X.sPair.p);
                hipLaunchKernelGGL(k_pair_exact, grid, dim3(WDM_BLOCK), 0, st, B.D, FRd, mB, (const int32_t*)X.cvals2.p,
                                   (const uint32_t*)X.ckeys2.p, (const double4*)X.eb0.p, (const double4*)X.eb1.p, (const int2*)X.einfo.p,
                                   large ? (const int32_t*)X.lbeg.p : nullptr, (const int32_t*)X.lvals.p,
                                   (const int32_t*)X.scnt.p, scap, (const int2*)X.sPair.p, X.wcnt.p, X.wmax.p + 1, wcap, X.rItem.p,
                                   X.rRow.p, X.rSlot.p, X.rowCnt.p);
                checkLaunch();
            }
            scanExclusive(c, (const uint32_t*)X.rowCnt.p, X.rowOff.p, nRows);
            if (nW > 0) {
                KTimer tm(c, "pairs_sort");
                hipLaunchKernelGGL(k_cand_scatter, dim3((unsigned)((nW + 3) / 4)), dim3(WDM_BLOCK), 0, st, (const int32_t*)X.wcnt.p, nW,
                                   wcap, (const uint64_t*)X.rItem.p, (const uint32_t*)X.rRow.p, (const uint32_t*)X.rSlot.p,
                                   (const int64_t*)X.rowOff.p, X.btmp.p);
                checkLaunch();
                bucketFinish(c, (int64_t)nreg, (const int64_t*)X.rowOff.p, CandOut{FRd, nT, X.candK.p, X.candP.p},
                             (const int64_t*)X.rowOff.p + nRows);
            }
            {
                KTimer to(c, "old_flags");
                if (nOldTot > 0) {
                    hipLaunchKernelGGL(k_flag_old<T>, dim3((unsigned)bpO[nT]), dim3(WDM_BLOCK), 0, st, Cd, FRd, mO,
                                       (const int64_t*)X.rowOff.p, (const uint64_t*)X.candK.p, X.ooFlag.p);
                    checkLaunch();
                }
                scanExclusive(c, (const uint8_t*)X.ooFlag.p, X.ooRank.p, nOldTot);
            }
            int32_t wm[3] = {0, 0, 0};
            segmentStarts2(c, FRd, nT, X.rowOff.p, nRows, nOldTot > 0 ? X.ooRank.p : nullptr, nOldTot, cs, os, X.wmax.p, wm,
                           sizeof(wm));  // synchronises
            if (wm[2] != 0) {  // a lagged grid too small for a bound: lay the grids out from this run's extents
                if (!lagged) throw WdmError(WDM_E_STATE, "collider: bound wider than its grid");
                lagged = false;
                fillU32(c, (uint32_t*)X.wmax.p, 3);
                continue;
            }
            if (wm[0] > scap || wm[1] > wcap) {
                if (attempt > 2) throw WdmError(WDM_E_STATE, "collider: pairs per wavefront changed between passes");
                if (wm[0] > scap) c->scanCap = grow(wm[0]);
                if (wm[1] > wcap) c->waveCap = grow(wm[1]);
                continue;
            }
            for (int t = 0; t < nT; t++) { F[t].candStart = cs[t]; F[t].nCand = cs[t + 1] - cs[t]; }
            ncandTot = cs[nT];
            if (nOldTot > 0)
                for (int t = 0; t < nT; t++) noor[t] = os[t + 1] - os[t];
            // this run's extents (in the mailbox since the synchronisation) lay out the next run's grids
            const size_t eb = (size_t)nT * 8 * sizeof(double);
            for (int t = 0; t < nT; t++) {
                Scene& S = S_(t);
                std::memcpy(S.extLast, X.extMail + (size_t)t * 8 * sizeof(double), sizeof(S.extLast));
                S.histLast.resize(EXT_BINS);
                std::memcpy(S.histLast.data(), X.extMail + eb + (size_t)t * EXT_BINS * sizeof(uint32_t), EXT_BINS * sizeof(uint32_t));
                S.extValid = true;
            }
            break;
        }

        // merged lists
        int64_t nNewTot = 0, noorTot = 0;
        for (int t = 0; t < nT; t++) {
            Scene& S = S_(t);
            ConBuf<T>* cb = bufs<T>(S);
            int64_t nnew = F[t].nCand + noor[t];
            noorTot += noor[t];
            cb[1 - S.cur].reserve(c, std::max<int64_t>(nnew, 1));
            if (cb[S.cur].cap == 0) cb[S.cur].reserve(c, 1);
            if (S.nc == 0 && cb[S.cur].cap < cb[1 - S.cur].cap) {
                // first full run: the list that holds nothing yet gets the same capacity, so
                // that the next full run (its merge target) does not reallocate
                HIPCHK(hipStreamSynchronize(st));
                cb[S.cur].reserve(c, std::max<int64_t>(nnew, 1));
            }
            ensure(c, S.entV, (size_t)std::max<int64_t>(2 * nnew, 1));
            C[t].cv = cb[S.cur].view();
            C[t].cvNew = cb[1 - S.cur].view();
            F[t].newStart = nNewTot;
            F[t].nNew = nnew;
            F[t].csrOff = gpc(S.csrOff.p);
            F[t].entV = gpc(S.entV.p);
            nNewTot += nnew;
        }
        ensure(c, X.oslot, std::max<int64_t>(nNewTot, 1)); ensure(c, X.srcRow, std::max<int64_t>(nNewTot, 1));
        ensure(c, X.rflag, std::max<int64_t>(nNewTot, 1));
        ensure(c, X.rscan, (size_t)nNewTot + 1);
        ensure(c, X.nrflag, std::max<int64_t>(nNewTot, 1)); ensure(c, X.nrscan, (size_t)nNewTot + 1);
        std::vector<int64_t> bpC = blockPrefix(nT, [&](int t) { return F[t].nCand; });
        std::vector<int64_t> bpN = blockPrefix(nT, [&](int t) { return F[t].nNew; });
        Upload uN;
        const size_t oFN = uN.add(F), oCN = uN.add(C), oC = uN.add(bpC), oN = uN.add(bpN);
        uN.send(c, X.stRunN);
        FRd = uN.at<FRScene>(oFN);
        Cd = uN.at<ConDesc<T>>(oCN);
        const Map mC = uN.map(oC, nT), mN = uN.map(oN, nT);
        {
            KTimer tm(c, "merge");
            if (ncandTot > 0)
                hipLaunchKernelGGL(k_emit_cand<T>, dim3((unsigned)bpC[nT]), dim3(WDM_BLOCK), 0, st, B.D, Cd, FRd, mC, (const int64_t*)X.ooRank.p,
                                   (const int64_t*)X.oldRow.p, (const uint64_t*)X.candK.p, (const int32_t*)X.candP.p, X.oslot.p, X.rflag.p);
            if (noorTot > 0)
                hipLaunchKernelGGL(k_emit_oor<T>, dim3((unsigned)bpO[nT]), dim3(WDM_BLOCK), 0, st, Cd, FRd, mO, (const uint8_t*)X.ooFlag.p,
                                   (const int64_t*)X.ooRank.p, (const int64_t*)X.rowOff.p, (const uint64_t*)X.candK.p, X.oslot.p, X.rflag.p);
            checkLaunch();
            scanExclusive(c, (const uint8_t*)X.rflag.p, X.rscan.p, nNewTot);
            if (nNewTot > 0)
                hipLaunchKernelGGL(k_copy_state<T>, dim3((unsigned)bpN[nT]), dim3(WDM_BLOCK), 0, st, Cd, FRd, mN, (const int32_t*)X.oslot.p,
                                   (const int64_t*)X.rscan.p, X.nrflag.p, X.srcRow.p);
            if (nNewTot > 0)
                hipLaunchKernelGGL(k_copy_rows<T>, dim3((unsigned)bpN[nT]), dim3(WDM_BLOCK), 0, st, Cd, FRd, mN, (const int64_t*)X.rscan.p,
                                   (const int32_t*)X.srcRow.p);
            checkLaunch();
            // ContactLoop work list: rows first, then the other non-dormant contacts
            scanExclusive(c, (const uint8_t*)X.nrflag.p, X.nrscan.p, nNewTot, ScanNearSplit());
            if (nNewTot > 0)
                hipLaunchKernelGGL(k_items<T>, dim3((unsigned)bpN[nT]), dim3(WDM_BLOCK), 0, st, Cd, FRd, mN, (const int64_t*)X.rscan.p,
                                   (const int64_t*)X.nrscan.p);
            hipLaunchKernelGGL(k_set_counters<T>, dim3((nT + 63) / 64), dim3(64), 0, st, Cd, FRd, nT, (const int64_t*)X.rscan.p,
                               (const int64_t*)X.nrscan.p);
            checkLaunch();
        }
        for (int t = 0; t < nT; t++) {
            Scene& S = S_(t);
            S.cur = 1 - S.cur;
            S.nc = F[t].nNew;
            S.nFullRuns++;
            S.layoutStep = S.step + 1;  // the work list is laid out for the steps from the next one
            S.needFull = false;
        }
        // CSR (contact -> particle): entries counted per particle row, then bucket-sorted by
        // contact index within each row
        {
            KTimer tm(c, "csr");
            const int64_t ne = 2 * nNewTot;
            const size_t ne1 = (size_t)std::max<int64_t>(ne, 1);
            ensure(c, X.eB, ne1); ensure(c, X.eS, ne1); ensure(c, X.eI, ne1); ensure(c, X.btmp, ne1);
            ensure(c, X.csrCnt, (size_t)nRows); ensure(c, X.csrScan, (size_t)nRows + 1);
            fillU32(c, X.csrCnt.p, nRows);
            if (nNewTot > 0) {
                hipLaunchKernelGGL(k_csr_count<T>, dim3((unsigned)bpN[nT]), dim3(WDM_BLOCK), 0, st, Cd, FRd, mN, X.eB.p, X.eS.p, X.eI.p,
                                   X.csrCnt.p);
                checkLaunch();
            }
            scanExclusive(c, (const uint32_t*)X.csrCnt.p, X.csrScan.p, nRows);
            if (ne > 0) {
                hipLaunchKernelGGL(k_csr_scatter, dim3(nblk(ne)), dim3(WDM_BLOCK), 0, st, (const uint32_t*)X.eB.p, (const uint32_t*)X.eS.p,
                                   (const uint64_t*)X.eI.p, ne, (const int64_t*)X.csrScan.p, X.btmp.p);
                checkLaunch();
                bucketFinish(c, ne, (const int64_t*)X.csrScan.p, CsrOut{FRd, nT, (const int64_t*)X.csrScan.p},
                             (const int64_t*)X.csrScan.p + nRows);
            }
            hipLaunchKernelGGL(k_csr_offsets, dim3((unsigned)bp1[nT]), dim3(WDM_BLOCK), 0, st, B.D, FRd, m1, (const int64_t*)X.csrScan.p);
            checkLaunch();
        }
    }

    // work-list relayout of the scenes L: the real items after the rows moved right after
    // them (component-major state rows, the row block grows), the others after those, CSR
    // entries remapped; the list's other buffer stages the moved items
    template <typename T>
    static void relayout(Batch& B, const std::vector<int>& L) {
        Ctx* c = B.c;
        Shared& X = *c->sh;
        hipStream_t st = c->stream;
        const int n = (int)L.size();
        if (n == 0) return;
        std::vector<ConDesc<T>> C(n);
        std::memset(C.data(), 0, C.size() * sizeof(ConDesc<T>));
        std::vector<RlScene> R(n);
        for (int t = 0; t < n; t++) {
            Scene& S = *B.ss[L[t]];
            ConBuf<T>* cb = bufs<T>(S);
            cb[1 - S.cur].reserve(c, S.nc);
            C[t].cv = cb[S.cur].view();
            C[t].cvNew = cb[1 - S.cur].view();
            C[t].nc = S.nc;
            C[t].realMask = gpc(S.realMask.p);
            C[t].desc = L[t];
            R[t].entV = gpc(S.entV.p);
            R[t].csrOff = gpc(S.csrOff.p);
            R[t].np = S.np;
        }
        const std::vector<int64_t> bp = blockPrefix(n, [&](int t) { return C[t].nc; });
        const std::vector<int64_t> bpE = blockPrefix(n, [&](int t) { return 2 * C[t].nc; });
        const int64_t nf = bp[n] * WDM_BLOCK;
        ensure(c, X.rlFlag, (size_t)std::max<int64_t>(nf, 1));
        ensure(c, X.rlScan, (size_t)nf + 1);
        ensure(c, X.rlMap, (size_t)std::max<int64_t>(nf, 1));
        Upload u;
        const size_t oC = u.add(C), oR = u.add(R), oB = u.add(bp), oE = u.add(bpE);
        u.send(c, X.stRl);
        const ConDesc<T>* Cd = u.at<ConDesc<T>>(oC);
        const Map m = u.map(oB, n), mE = u.map(oE, n);
        if (bp[n] > 0) {
            KTimer tm(c, "relayout");
            const int64_t* scan = (const int64_t*)X.rlScan.p;
            hipLaunchKernelGGL(k_rl_flags<T>, dim3((unsigned)bp[n]), dim3(WDM_BLOCK), 0, st, Cd, m, X.rlFlag.p);
            checkLaunch();
            scanExclusive(c, (const uint8_t*)X.rlFlag.p, X.rlScan.p, nf, ScanNearSplit());
            hipLaunchKernelGGL(k_rl_out<T>, dim3((unsigned)bp[n]), dim3(WDM_BLOCK), 0, st, Cd, m, scan, X.rlMap.p);
            hipLaunchKernelGGL(k_rl_in<T>, dim3((unsigned)bp[n]), dim3(WDM_BLOCK), 0, st, Cd, m, scan);
            hipLaunchKernelGGL(k_rl_mask<T>, dim3((unsigned)bp[n]), dim3(WDM_BLOCK), 0, st, Cd, m, scan);
            checkLaunch();
            if (bpE[n] > 0)
                hipLaunchKernelGGL(k_rl_csr<T>, dim3((unsigned)bpE[n]), dim3(WDM_BLOCK), 0, st, Cd, u.at<RlScene>(oR), mE, m,
                                   (const int32_t*)X.rlMap.p);
            hipLaunchKernelGGL(k_rl_counter<T>, dim3((n + 63) / 64), dim3(64), 0, st, Cd, m, scan, n);
            checkLaunch();
        }
        for (int t = 0; t < n; t++) {
            Scene& S = *B.ss[L[t]];
            S.layoutStep = S.step;
            S.relayouts++;
        }
    }
    static void relayoutList(Batch& B) {
        std::vector<int> L64, L32;
        for (int i = 0; i < B.n(); i++) (B.ss[i]->mixed() ? L32 : L64).push_back(i);
        relayout<double>(B, L64);
        relayout<float>(B, L32);
    }

    // scenes that need a full run this step: after an upload, with verletDist 0, or when
    // the Leapfrog of this step moved a node by more than verletDist (words read by the caller)
    static std::vector<int> fullList(Batch& B, bool wordsRead) {
        std::vector<int> L;
        for (int i = 0; i < B.n(); i++) {
            Scene& S = *B.ss[i];
            resolveVerlet(S);
            bool full = S.needFull || S.verletDist == 0;
            if (!full && S.lastLeapfrogVerletChecked && wordsRead) full = B.c->sh->hostWords[4 * (size_t)S.slot] == stampOf(S.step);
            if (full) L.push_back(i);
        }
        return L;
    }
    static void runList(Batch& B, const std::vector<int>& L) {
        if (L.empty()) return;  // pending removals are decided inside the contact kernel
        std::vector<int> L64, L32;
        for (int i : L) (B.ss[i]->mixed() ? L32 : L64).push_back(i);
        B.ensureDesc();
        fullRun<double>(B, L64);
        fullRun<float>(B, L32);
    }
    static void run(Batch& B) {
        bool anyCheck = false;
        for (Scene* S : B.ss) anyCheck |= (!S->needFull && S->lastLeapfrogVerletChecked);
        if (anyCheck) readWords(B.c, B.ss);  // the Verlet decision of updateBboxes_doFullRun (:340-350)
        runList(B, fullList(B, anyCheck));
    }
};
template <>
ConBuf<double>* InsertionSortCollider::bufs<double>(Scene& S) { return S.cb64; }
template <>
ConBuf<float>* InsertionSortCollider::bufs<float>(Scene& S) { return S.cb32; }

// ---------------- ContactLoop::run (pkg/dem/ContactLoop.cpp:93-259) ----------------
// Scenes grouped by precision, law and the geometry functors they need, one launch per
// group; the groups' ConDesc tables and block prefixes packed into a Blob.
struct ContactPlan {
    struct Group {
        int mixed, law, geom, n;
        size_t oC, oBp;
        int64_t blocks;
        bool anyEnergy;
    };
    std::vector<Group> gs;
    const uint8_t* base = nullptr;

    template <typename T>
    void addGroup(Batch& B, const std::vector<int>& L, int law, int geom, Blob& b) {
        Ctx* c = B.c;
        const int n = (int)L.size();
        if (n == 0) return;
        std::vector<ConDesc<T>> C(n);
        std::memset(C.data(), 0, C.size() * sizeof(ConDesc<T>));
        Group g;
        g.mixed = sizeof(T) == 4;
        g.law = law;
        g.geom = geom;
        g.n = n;
        g.anyEnergy = false;
        for (int t = 0; t < n; t++) {
            Scene& S = *B.ss[L[t]];
            ConBuf<T>* cb = InsertionSortCollider::bufs<T>(S);
            if (cb[S.cur].cap == 0) cb[S.cur].reserve(c, 1);
            ensure(c, S.realMask, (size_t)nblk(std::max<int64_t>(S.nc, 1)) * (WDM_BLOCK / 64) + 1);
            C[t].cv = cb[S.cur].view();
            C[t].nc = S.nc;
            C[t].realMask = gpc(S.realMask.p);
            C[t].desc = L[t];
            g.anyEnergy |= S.P.trackEnergy != 0;
        }
        std::vector<int64_t> bp = blockPrefix(n, [&](int t) { return C[t].nc; });
        g.oC = b.addv(C);
        g.oBp = b.addv(bp);
        g.blocks = bp[n];
        ensureGraphArg(c, c->sh->epartC, (size_t)bp[n] * 5 + 5);
        gs.push_back(g);
    }
    void build(Batch& B, const std::vector<int>& L, Blob& b) {
        std::vector<int> g[2][2][3];  // [mixed][law][geometry: spheres, + walls, + facets]
        for (int i : L) {
            Scene& S = *B.ss[i];
            if (S.needFull) throw WdmError(WDM_E_STATE, "ContactLoop before the collider's first run");
            g[S.mixed() ? 1 : 0][S.P.law == WDM_LAW_HERTZ_DMT ? 1 : 0][S.hasFacets ? 2 : (S.unbH.empty() ? 0 : 1)].push_back(i);
        }
        const int geoms[3] = {G_SS, G_SS | G_WS, G_SS | G_WS | G_FS};
        for (int q = 0; q < 3; q++) {
            addGroup<double>(B, g[0][0][q], WDM_LAW_IDEALELPL, geoms[q], b);
            addGroup<double>(B, g[0][1][q], WDM_LAW_HERTZ_DMT, geoms[q], b);
            addGroup<float>(B, g[1][0][q], WDM_LAW_IDEALELPL, geoms[q], b);
            addGroup<float>(B, g[1][1][q], WDM_LAW_HERTZ_DMT, geoms[q], b);
        }
    }
    void sig(std::vector<uint64_t>& k) const {
        for (const Group& g : gs)
            k.insert(k.end(), {(uint64_t)g.mixed, (uint64_t)g.law, (uint64_t)g.geom, (uint64_t)g.n, g.oC, g.oBp, (uint64_t)g.blocks,
                               (uint64_t)g.anyEnergy});
    }
    template <typename T, int LAW, int GEOM>
    void launchT(Ctx* c, const SceneDesc* D, const Group& g, int spec, const StepVar* sv) const {
        Map m;
        m.bprefix = (const int64_t*)(base + g.oBp);
        m.n = g.n;
        const ConDesc<T>* Cd = (const ConDesc<T>*)(base + g.oC);
        if (g.blocks > 0) {
            KTimer tm(c, "contacts");
            hipLaunchKernelGGL((k_contacts<T, LAW, GEOM>), dim3((unsigned)xcdChunkGrid(g.blocks, XCD_CHUNK_CON)), dim3(WDM_BLOCK), 0, c->stream, D, Cd, m,
                               c->sh->epartC.p, spec, sv);
            checkLaunch();
        }
        if (g.anyEnergy) {
            const int32_t* descs = (const int32_t*)((const uint8_t*)Cd + offsetof(ConDesc<T>, desc));
            hipLaunchKernelGGL(k_energy_finalize, dim3(5 * g.n), dim3(WDM_BLOCK), 0, c->stream, D, descs, (int)(sizeof(ConDesc<T>) / 4), m,
                               (const double*)c->sh->epartC.p, 5, spec, sv);
            checkLaunch();
        }
    }
    template <typename T, int LAW>
    void launchG(Ctx* c, const SceneDesc* D, const Group& g, int spec, const StepVar* sv) const {
        if (g.geom == G_SS) launchT<T, LAW, G_SS>(c, D, g, spec, sv);
        else if (g.geom == (G_SS | G_WS)) launchT<T, LAW, G_SS | G_WS>(c, D, g, spec, sv);
        else launchT<T, LAW, G_SS | G_WS | G_FS>(c, D, g, spec, sv);
    }
    // spec: launched before the collider's decision is known; a scene whose Verlet trigger
    // fired skips it and is launched again after its full run
    void launch(Ctx* c, const SceneDesc* D, bool spec, const StepVar* sv) const {
        for (const Group& g : gs) {
            const bool h = g.law == WDM_LAW_HERTZ_DMT;
            if (g.mixed) { if (h) launchG<float, 1>(c, D, g, spec, sv); else launchG<float, 0>(c, D, g, spec, sv); }
            else { if (h) launchG<double, 1>(c, D, g, spec, sv); else launchG<double, 0>(c, D, g, spec, sv); }
        }
    }
};
struct ContactLoop {
    static void runList(Batch& B, const std::vector<int>& L, bool spec) {
        if (L.empty()) return;
        ContactPlan cp;
        Blob b;
        cp.build(B, L, b);
        B.ensureDesc();
        cp.base = B.c->sh->stCon.upload(B.c, b.b.data(), b.b.size());
        cp.launch(B.c, B.D, spec, nullptr);
    }
    static void done(Batch& B) {
        for (Scene* S : B.ss) {
            S->forcesFromContacts = true;  // applied by the next Leapfrog (or materialised on download)
            S->forcesReset = false;
            S->contactsRun = true;
            S->hSizeCL = S->hSize;
        }
    }
    static void run(Batch& B) {
        std::vector<int> all(B.n());
        for (int i = 0; i < B.n(); i++) all[i] = i;
        runList(B, all, false);
        done(B);
    }
};

// work-list relayout (memory order only), due at the start of a step c->relayout steps after
// the scene's last full run or relayout, when the current list has a ContactLoop's state
bool relayoutable(const Ctx* c, const Scene& S) {
    return S.contactsRun && !S.needFull && !S.failed && S.nc > 0 && S.nc >= c->relayoutMin;
}
bool relayoutDue(const Ctx* c, const Scene& S) { return c->relayout > 0 && relayoutable(c, S) && S.step - S.layoutStep >= c->relayout; }
void relayoutDueScenes(Ctx* c, const std::vector<Scene*>& ss) {
    std::vector<Scene*> due;
    for (Scene* S : ss)
        if (relayoutDue(c, *S)) due.push_back(S);
    if (due.empty()) return;
    Batch B(c, due);
    InsertionSortCollider::relayoutList(B);
}
// steps a window may take before a scene's next relayout is due (windows end there, so that
// every way of stepping a scene relayouts it at the same steps)
int64_t stepsToRelayout(const Ctx* c, const Scene& S) {
    if (c->relayout <= 0 || !relayoutable(c, S)) return INT64_MAX;
    return std::max<int64_t>(1, c->relayout - (S.step - S.layoutStep));
}
void checkReady(Scene& S) {
    if (S.failed)
        throw WdmError(WDM_E_STATE, "scene " + std::to_string(S.id) + " was stopped by a device error (" + S.failText + "): free it");
    if (!(S.paramsSet && S.nodesSet && S.partsSet && S.matsSet)) throw WdmError(WDM_E_STATE, "scene not fully uploaded");
    if (!S.clumpsSet) throw WdmError(WDM_E_STATE, "clump flags of the nodes and the uploaded clumps disagree (wdm_upload_clumps)");
}
// Scene::doOneStep prologue (core/Scene.cpp:371-389): cell update
void stepBeginCell(Scene& S) {
    checkReady(S);
    if (S.periodic) cellIntegrateAndUpdate(S, S.P.dt);
}
// Scene::doOneStep epilogue (core/Scene.cpp:395-400)
void stepEnd(Scene& S) {
    if (S.periodic) cellSetNextGradV(S);
    S.step++;
    S.time += S.P.dt;
    if (!std::isnan(S.nextDt)) { S.P.dt = S.nextDt; S.nextDt = NAN; }
}
// ---------------- DynDt (pkg/dem/DynDt.cpp:16-97) ----------------
// PeriodicEngine::isActivated, core/Engine.cpp:169-195, without the wall-clock criterion
bool ddActivate(wdm_dyndt& d, int64_t stepNow, double virtNow) {
    const bool initNow = d.stepLast < 0;
    if (initNow) { d.virtLast = virtNow; d.stepLast = stepNow; }
    if ((d.nDo < 0 || d.nDone < d.nDo) &&
        ((d.virtPeriod > 0 && virtNow - d.virtLast >= d.virtPeriod) ||
         (d.stepPeriod > 0 && ((stepNow - d.stepLast >= d.stepPeriod) || (d.stepModulo && stepNow % d.stepPeriod == 0))))) {
        d.virtPrev = d.virtLast; d.virtLast = virtNow;
        d.stepPrev = d.stepLast; d.stepLast = stepNow;
        if (initNow && !d.initRun) return false;
        d.nDone++;
        return true;
    }
    return false;
}
struct DynDtEngine {
    // facet nodes get their stiffness from the facets' contacts, which are not summed here:
    // they must be fully blocked, or clumped to a fully blocked master
    static void checkSupported(const Scene& S) {
        for (int32_t q : S.facetNodesH) {
            uint32_t f = S.flagsH[q];
            if ((f & WDM_CLUMP_CLUMPED) && q < (int32_t)S.nodeMasterH.size() && S.nodeMasterH[q] >= 0) f = S.flagsH[S.nodeMasterH[q]];
            else if (f & WDM_CLUMP_CLUMPED) continue;  // clumps not uploaded yet: checked again later
            if ((f & 63u) != 63u) throw WdmError(WDM_E_UNSUPPORTED, "DynDt: facet nodes with free DOFs are outside the implemented path");
        }
    }
    template <typename T, int LAW>
    static void launchGroup(Batch& B, const std::vector<int>& L, unsigned long long* out) {
        Ctx* c = B.c;
        const int n = (int)L.size();
        if (n == 0) return;
        std::vector<ConDesc<T>> C(n);
        std::memset(C.data(), 0, C.size() * sizeof(ConDesc<T>));
        for (int t = 0; t < n; t++) {
            Scene& S = *B.ss[L[t]];
            ConBuf<T>* cb = InsertionSortCollider::bufs<T>(S);
            C[t].cv = cb[S.cur].view();
            C[t].nc = S.nc;
            C[t].realMask = gpc(S.realMask.p);
            C[t].desc = L[t];
        }
        std::vector<int64_t> bp = blockPrefix(n, [&](int t) { return B.ss[L[t]]->nn; });
        Blob b;
        const size_t oC = b.addv(C), oB = b.addv(bp);
        const uint8_t* base = c->sh->stDyn.upload(c, b.b.data(), b.b.size());
        Map m;
        m.bprefix = (const int64_t*)(base + oB);
        m.n = n;
        if (bp[n] > 0) {
            KTimer tm(c, "dyndt");
            hipLaunchKernelGGL((k_dyndt<T, LAW>), dim3((unsigned)bp[n]), dim3(WDM_BLOCK), 0, c->stream, B.D, (const ConDesc<T>*)(base + oC), m, out);
            checkLaunch();
        }
    }
    // DynDt::critDt_stiffness (DynDt.cpp:61-70) of each scene, from the contacts of its last
    // ContactLoop (+inf before the first one); one launch per precision and law, one synchronisation
    static std::vector<double> crit(Ctx* c, const std::vector<Scene*>& ss) {
        Shared& X = *c->sh;
        const int n = (int)ss.size();
        std::vector<double> out(n, INFINITY);
        std::vector<int> g[2][2];
        for (int i = 0; i < n; i++) {
            checkSupported(*ss[i]);
            if (!ss[i]->needFull && ss[i]->contactsRun) g[ss[i]->mixed() ? 1 : 0][ss[i]->P.law == WDM_LAW_HERTZ_DMT ? 1 : 0].push_back(i);
        }
        if (g[0][0].empty() && g[0][1].empty() && g[1][0].empty() && g[1][1].empty()) return out;
        Batch B(c, ss);
        B.upload();
        ensure(c, X.ddOut, (size_t)n);
        hipLaunchKernelGGL(k_fill_u64, dim3((n + 63) / 64), dim3(64), 0, c->stream, X.ddOut.p, n, 0x7ff0000000000000ULL);
        checkLaunch();
        // the groups' scenes take consecutive output slots, group after group
        std::vector<int> slotOf(n, -1);
        int next = 0;
        for (int mx = 0; mx < 2; mx++)
            for (int lw = 0; lw < 2; lw++) {
                const std::vector<int>& L = g[mx][lw];
                if (L.empty()) continue;
                unsigned long long* o = X.ddOut.p + next;
                for (int i : L) slotOf[i] = next++;
                if (mx) { if (lw) launchGroup<float, 1>(B, L, o); else launchGroup<float, 0>(B, L, o); }
                else { if (lw) launchGroup<double, 1>(B, L, o); else launchGroup<double, 0>(B, L, o); }
            }
        std::vector<unsigned long long> h(n);
        unsigned long long* pin = (unsigned long long*)pinnedRead(c, n * sizeof(unsigned long long));
        HIPCHK(hipMemcpyAsync(pin, X.ddOut.p, n * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        std::memcpy(h.data(), pin, n * sizeof(unsigned long long));
        for (int i = 0; i < n; i++) {
            if (slotOf[i] < 0) continue;
            double v;
            std::memcpy(&v, &h[slotOf[i]], sizeof(double));
            out[i] = std::sqrt(v);
        }
        return out;
    }
    // DynDt::run (DynDt.cpp:72-97) for scenes whose engine fired at step dd.stepLast
    static void run(Ctx* c, const std::vector<Scene*>& ss) {
        if (ss.empty()) return;
        const std::vector<double> cr = crit(c, ss);
        for (size_t k = 0; k < ss.size(); k++) {
            Scene& S = *ss[k];
            wdm_dyndt& d = S.dd;
            const double crDt = cr[k];
            if (std::isinf(crDt)) continue;  // no timestep computed: keep the current value
            const int64_t nSteps = d.stepLast - d.stepPrev;
            const double maxNewDt = S.P.dt * std::pow(1. + d.maxRelInc, (double)nSteps);
            const double cand = crDt * d.dtSafety;
            const double nextDt = maxNewDt < cand ? maxNewDt : cand;  // std::min(cand, maxNewDt)
            if (!d.dryRun) {
                if (nextDt != S.P.dt) { d.dt = NAN; S.nextDt = nextDt; }
            } else d.dt = nextDt;
        }
    }
    // the engine loop's turn of DynDt (after ContactLoop) at the scenes' current step
    static void fire(Ctx* c, const std::vector<Scene*>& ss) {
        std::vector<Scene*> L;
        for (Scene* S : ss)
            if (S->dd.enabled && ddActivate(S->dd, S->step, S->time)) L.push_back(S);
        run(c, L);
    }
};

// One timestep of every scene of the batch, engines in Woo's order, each launched once for
// all scenes.  The ContactLoop of the scenes that can skip the collider is queued right
// behind Leapfrog, before the host has read the Verlet triggers; scenes whose trigger
// fired skip it on the device and run the collider and the ContactLoop afterwards.  The
// host waits only for Leapfrog, so the GPU stays busy while it decides.
void doOneStep(Batch& B) {
    Ctx* c = B.c;
    Shared& X = *c->sh;
    for (Scene* S : B.ss) stepBeginCell(*S);
    Leapfrog::run(B);
    std::vector<int> spec;
    for (int i = 0; i < B.n(); i++)
        if (B.ss[i]->lastLeapfrogVerletChecked) spec.push_back(i);
    if (!spec.empty()) {
        HIPCHK(hipMemcpyAsync(X.hostWords, X.words.p, 4 * (size_t)X.nslots * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipEventRecord(X.evWords, c->stream));
        ContactLoop::runList(B, spec, true);
        HIPCHK(hipEventSynchronize(X.evWords));
        markFailed(c, B.ss);  // a scene with a device error stops here; the others go on
    }
    std::vector<int> full;
    for (int i : InsertionSortCollider::fullList(B, !spec.empty()))
        if (!B.ss[i]->failed) full.push_back(i);
    InsertionSortCollider::runList(B, full);
    ContactLoop::runList(B, full, false);
    ContactLoop::done(B);
    std::vector<Scene*> live;
    for (Scene* S : B.ss)
        if (!S->failed) live.push_back(S);
    DynDtEngine::fire(c, live);
    for (Scene* S : B.ss) stepEnd(*S);
}

// ---------------- windows: several steps launched together ----------------
// Between two collider runs a scene's step is Leapfrog then ContactLoop with no host
// decision, so K steps of a batch are launched back to back (one captured graph, replayed
// while the batch's tables keep their shape) and the host looks at the Verlet triggers
// once per window.  A scene whose trigger fires stops at that step (StepVar gate); after
// the window the host finishes those steps (collider, ContactLoop) for all stopped
// scenes in one batched pass.  Each scene runs the same kernels on the same data in the
// same order as step by step, so its trajectory does not change.

// cell state before a step's prologue (periodic scenes), to rewind a scene that stopped
struct CellSnap {
    HM3 hSize, gradV, pprevHsize, trsf, hSizeCL;
    double size[3], spinVec[3];
};
CellSnap snapCell(const Scene& S) {
    CellSnap z;
    z.hSize = S.hSize; z.gradV = S.gradV; z.pprevHsize = S.pprevHsize; z.trsf = S.trsf; z.hSizeCL = S.hSizeCL;
    for (int k = 0; k < 3; k++) { z.size[k] = S.size[k]; z.spinVec[k] = S.spinVec[k]; }
    return z;
}
void restoreCell(Scene& S, const CellSnap& z) {
    S.hSize = z.hSize; S.gradV = z.gradV; S.pprevHsize = z.pprevHsize; S.trsf = z.trsf; S.hSizeCL = z.hSizeCL;
    for (int k = 0; k < 3; k++) { S.size[k] = z.size[k]; S.spinVec[k] = z.spinVec[k]; }
}

// a scene can join a window of K steps: collider state current, forces from its contacts,
// a Verlet distance, and stamps that do not wrap inside the window
bool windowable(Scene& S, int K) {
    resolveVerlet(S);
    if (S.needFull || !(S.verletDist > 0) || !S.forcesFromContacts) return false;
    return (S.step % 0x7fffffffLL) + K < 0x7fffffffLL;
}

// K steps of the scenes ss (scene i has left[i] >= 1 steps to go); returns the steps each
// completed
std::vector<int64_t> runWindow(Ctx* c, const std::vector<Scene*>& ss, const std::vector<int64_t>& left, int K) {
    Shared& X = *c->sh;
    hipStream_t st = c->stream;
    const int n = (int)ss.size();
    Batch B(c, ss);
    for (Scene* S : ss) checkReady(*S);
    LeapPlan lp;
    ContactPlan cp;
    Blob b;
    lp.build(B, b);  // sets the Verlet check flag of every scene (all are checked here)
    std::vector<int> all(n);
    for (int i = 0; i < n; i++) all[i] = i;
    cp.build(B, all, b);
    // descriptors from the state at the window's start; the per-step fields come from the rows
    std::vector<SceneDesc> d(n);
    for (int i = 0; i < n; i++) d[i] = makeDesc(*ss[i]);
    std::vector<int64_t> s0(n);
    std::vector<uint32_t> gate(n);
    for (int i = 0; i < n; i++) { s0[i] = ss[i]->step; gate[i] = stampOf(ss[i]->step); }
    // the host side of each step (cell prologue, cell of the ContactLoop, epilogue), run ahead
    std::vector<StepVar> sv((size_t)K * n);
    std::memset(sv.data(), 0, sv.size() * sizeof(StepVar));
    std::vector<CellArgs> cells;
    std::vector<size_t> cellOf((size_t)K * n, 0), cell0(n, 0);
    std::vector<std::vector<CellSnap>> snaps(n);
    // Scene.time and the DynDt state before each step's engine loop (to rewind a scene that
    // stops); a scene whose DynDt fires at step j ends its window there (cutAt)
    std::vector<std::vector<double>> tSnap(n);
    std::vector<std::vector<wdm_dyndt>> ddSnap(n);
    std::vector<int> cutAt(n, -1);
    for (int j = 0; j < K; j++) {
        for (int i = 0; i < n; i++) {
            Scene& S = *ss[i];
            StepVar& v = sv[(size_t)j * n + i];
            v.gate = gate[i];
            if (j >= left[i] || (cutAt[i] >= 0 && j > cutAt[i])) { v.idle = 1; continue; }
            if (S.periodic) snaps[i].push_back(snapCell(S));
            stepBeginCell(S);
            if (S.periodic || j == 0) {
                CellArgs ca;
                makeCellArgs(S, ca);
                cells.push_back(ca);
            }
            cellOf[(size_t)j * n + i] = (S.periodic || j == 0) ? cells.size() - 1 : cellOf[i];
            if (j == 0) cell0[i] = cells.size() - 1;
            v.step = S.step;
            v.stamp = stampOf(S.step);
            v.reallyTrackEnergy = S.P.trackEnergy && (!S.periodic || S.step > 0 || hequal(S.gradV, heye()));
            S.hSizeCL = S.hSize;  // ContactLoop::done
            tSnap[i].push_back(S.time);
            if (S.dd.enabled) {
                ddSnap[i].push_back(S.dd);
                if (ddActivate(S.dd, S.step, S.time)) cutAt[i] = j;
            }
            stepEnd(S);
        }
    }
    const size_t oD = b.addv(d), oCells = b.addv(cells), oSV = b.addv(sv);
    const size_t bytes = b.b.size();
    ensureGraphArg(c, X.winDev, bytes);
    if (X.winHostCap < bytes) {
        HIPCHK(hipStreamSynchronize(st));
        if (X.winHost) (void)hipHostFree(X.winHost);
        X.winHost = nullptr;
        X.winHostCap = 0;
        size_t cap = bytes + bytes / 2;
        HIPCHK(hipHostMalloc((void**)&X.winHost, cap, hipHostMallocDefault));
        X.winHostCap = cap;
    }
    uint8_t* base = X.winDev.p;
    const CellArgs* cellsD = (const CellArgs*)(base + oCells);
    SceneDesc* dh = b.at<SceneDesc>(oD);
    for (int i = 0; i < n; i++) dh[i].a.cell = gpc(cellsD + cell0[i]);
    StepVar* svh = b.at<StepVar>(oSV);
    for (size_t k = 0; k < (size_t)K * n; k++)
        if (!svh[k].idle) svh[k].cell = gpc(cellsD + cellOf[k]);
    std::memcpy(X.winHost, b.b.data(), bytes);  // the previous window has finished (synchronised below)
    HIPCHK(hipMemcpyAsync(base, X.winHost, bytes, hipMemcpyHostToDevice, st));
    lp.base = base;
    cp.base = base;
    const SceneDesc* D = (const SceneDesc*)(base + oD);
    const StepVar* SV = (const StepVar*)(base + oSV);
    auto launchAll = [&]() {
        for (int j = 0; j < K; j++) {
            lp.launch(c, D, SV + (size_t)j * n);
            cp.launch(c, D, true, SV + (size_t)j * n);
        }
    };
    if (c->graphs && !c->timing.on) {
        std::vector<uint64_t> key = {(uint64_t)K, (uint64_t)n, (uint64_t)(uintptr_t)base, oD, oSV,
                                     (uint64_t)(uintptr_t)X.epartL.p, (uint64_t)(uintptr_t)X.epartC.p};
        lp.sig(key);
        cp.sig(key);
        if (!X.winExec || key != X.winKey) {
            X.dropGraph(c);
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
            try {
                launchAll();
            } catch (...) {
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPCHK(hipStreamEndCapture(st, &g));
            hipError_t e = hipGraphInstantiate(&X.winExec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(g);
                X.winExec = nullptr;
                throw WdmError(WDM_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
            }
            X.winGraph = g;
            X.winKey = key;
            X.graphCaptures++;
        }
        HIPCHK(hipGraphLaunch(X.winExec, st));
        X.graphLaunches++;
    } else launchAll();
    X.windows++;
    readWords(c, ss, false);  // scenes with a device error stop; the others finish the window
    // where each scene stopped
    std::vector<int64_t> done(n);
    std::vector<Scene*> stopped, fired;
    for (int i = 0; i < n; i++) {
        Scene& S = *ss[i];
        if (S.failed) {
            done[i] = left[i];
            continue;
        }
        const uint32_t w = X.hostWords[4 * (size_t)S.slot];
        int64_t active = std::min<int64_t>(K, left[i]);
        if (cutAt[i] >= 0) active = std::min<int64_t>(active, cutAt[i] + 1);
        if (w >= gate[i]) {
            const int64_t jf = (int64_t)(w - gate[i]);
            if (jf >= active) throw WdmError(WDM_E_STATE, "Verlet trigger outside the window");
            S.step = s0[i] + jf;
            S.time = tSnap[i][jf];
            if (S.dd.enabled) S.dd = ddSnap[i][jf];
            if (S.periodic) restoreCell(S, snaps[i][jf]);
            stepBeginCell(S);  // mid-step: Leapfrog done, collider and ContactLoop to run
            stopped.push_back(&S);
            done[i] = jf + 1;
        } else {
            done[i] = active;
            if (cutAt[i] >= 0) fired.push_back(&S);  // DynDt fired at the scene's last step (its epilogue ran ahead)
        }
    }
    if (!stopped.empty()) {
        Batch F(c, stopped);
        std::vector<int> L(stopped.size());
        for (size_t i = 0; i < stopped.size(); i++) L[i] = (int)i;
        InsertionSortCollider::runList(F, L);
        ContactLoop::runList(F, L, false);
        ContactLoop::done(F);
    }
    // DynDt of every scene whose engine fires at its last step of the window, in one pass
    const size_t nAhead = fired.size();
    for (Scene* S : stopped)
        if (S->dd.enabled && ddActivate(S->dd, S->step, S->time)) fired.push_back(S);
    DynDtEngine::run(c, fired);
    for (size_t k = 0; k < nAhead; k++)
        if (!std::isnan(fired[k]->nextDt)) { fired[k]->P.dt = fired[k]->nextDt; fired[k]->nextDt = NAN; }
    for (Scene* S : stopped) stepEnd(*S);
    return done;
}

// DemData.force/torque into the arrays: after a ContactLoop the gravity reset plus the contact
// sums (sums), after a Leapfrog the reset alone
void materialiseForces(Scene& S, bool sums) {
    Ctx* c = S.ctx;
    Batch B(c, {&S});
    B.upload();
    std::vector<int64_t> bp = blockPrefix(1, [&](int) { return S.nn; });
    Map m = uploadMap(c, c->sh->stMapNode, bp);
    if (bp[1] == 0) return;
    KTimer t(c, "gather");
    hipLaunchKernelGGL(k_gather, dim3((unsigned)bp[1]), dim3(WDM_BLOCK), 0, c->stream, B.D, (const int32_t*)nullptr, m, sums ? 1 : 0);
    checkLaunch();
}
// the arrays receive the node forces, whichever engine ran last
void materialiseAll(Scene& S) {
    if (S.forcesFromContacts) materialiseForces(S, true);
    else if (S.forcesReset) materialiseForces(S, false);
    S.forcesReset = false;
}

Scene& getScene(Ctx* c, int32_t id) {
    auto it = c->scenes.find(id);
    if (it == c->scenes.end()) throw WdmError(WDM_E_ARG, "unknown scene id " + std::to_string(id));
    return *it->second;
}

template <typename F>
int guard(Ctx* c, F&& f) {
    if (!c) return WDM_E_ARG;
    try {
        HIPCHK(hipSetDevice(c->device));
        f();
        return WDM_OK;
    } catch (WdmError& e) {
        c->err = e.what();
        return e.code;
    } catch (std::exception& e) {
        c->err = e.what();
        return WDM_E_HIP;
    }
}

template <typename T>
void downloadContacts(Scene& S, ConBuf<T>* cb, int64_t cap, int32_t* idA, int32_t* idB, int32_t* cellDist3, uint8_t* isReal,
                      double* forceLocal3, double* contPt3, double* trsf9, double* uN, int64_t* nOut, double* phys5 = nullptr,
                      double* lens2 = nullptr) {
    Ctx* c = S.ctx;
    ConBuf<T>& b = cb[S.cur];
    int64_t n = S.nc;
    // only the rows and items in use are copied (rows of the full run: nR; items: nI; the
    // buffers' capacity can be far larger), into host buffers that are not cleared first
    int32_t nR = 0, nI = 0;
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(&nR, b.counter.p, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(&nI, b.nItems.p, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (nR < 0 || nI < 0 || nR > b.cap || nI > b.cap) throw WdmError(WDM_E_STATE, "contact list counters out of range");
    }
    // rows of the items after the full run's rows (none right after a checkpoint upload, which
    // hands out rows only and builds the items at the next full run)
    const int32_t nTail = nI > nR ? nI - nR : 0;
    const int64_t nRows = (int64_t)nR + nTail;
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    std::unique_ptr<int2[]> ids(new int2[nn]);
    std::unique_ptr<uint32_t[]> info(new uint32_t[nn]);
    std::unique_ptr<int32_t[]> cd(new int32_t[nn]), slot(new int32_t[nn]);
    std::unique_ptr<T[]> st(new T[(size_t)NSTATE * std::max<int32_t>(nR, 1)]), st2(new T[(size_t)NSTRIDE * std::max<int32_t>(nTail, 1)]);
    // contact point and uN of each state row (written by the ContactLoop, or by a checkpoint upload)
    std::unique_ptr<double4[]> cpt(new double4[(size_t)std::max<int64_t>(nRows, 1)]);
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(ids.get(), b.ids.p, n * sizeof(int2), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(info.get(), b.info.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(cd.get(), b.cd.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(slot.get(), b.slot.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        if (nR > 0)  // component-major: value q of the first nR rows
            HIPCHK(hipMemcpy2DAsync(st.get(), (size_t)nR * sizeof(T), b.st.p, (size_t)b.cap * sizeof(T), (size_t)nR * sizeof(T), NSTATE,
                                    hipMemcpyDeviceToHost, c->stream));
        if (nTail > 0) HIPCHK(hipMemcpyAsync(st2.get(), b.st2.p, (size_t)NSTRIDE * nTail * sizeof(T), hipMemcpyDeviceToHost, c->stream));
        if (nRows > 0) {
            const double4* src = b.cpt.p;
            if constexpr (sizeof(T) == 4) {  // stored as branches: absolute points first
                Shared& X = *c->sh;
                ensure(c, X.cptTmp, (size_t)b.cap);
                hipLaunchKernelGGL(k_cpt_mixed<0>, dim3(nblk(n)), dim3(WDM_BLOCK), 0, c->stream, (const double4*)S.posr.p, b.view(), n,
                                   S.periodic ? 1 : 0, toM3(S.hSize), X.cptTmp.p);
                checkLaunch();
                src = X.cptTmp.p;
            }
            HIPCHK(hipMemcpyAsync(cpt.get(), src, (size_t)nRows * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    // state value q of a row (the layout of the kernels' stAt, compacted to the rows in use)
    auto stv = [&](int64_t row, int q) -> T {
        if (row < 0 || row >= (int64_t)nR + nTail) throw WdmError(WDM_E_STATE, "contact state row out of range");
        return row < nR ? st[(size_t)q * nR + row] : st2[(size_t)(row - nR) * NSTRIDE + q];
    };
    int64_t k = 0;
    for (int64_t i = 0; i < n && k < cap; i++) {
        if (info[i] & CI_DEAD) continue;
        bool real = (info[i] & CI_REAL) != 0;
        if (idA) idA[k] = ids[i].x;
        if (idB) idB[k] = ids[i].y;
        if (cellDist3) { cellDist3[3 * k] = unpack10(cd[i], 0); cellDist3[3 * k + 1] = unpack10(cd[i], 10); cellDist3[3 * k + 2] = unpack10(cd[i], 20); }
        if (isReal) isReal[k] = real ? 1 : 0;
        int64_t row = slot[i];
        for (int d = 0; d < 3; d++) {
            if (forceLocal3) forceLocal3[3 * k + d] = real ? (double)stv(row, S_F + d) : 0.;
        }
        if (phys5) for (int d = 0; d < 5; d++) phys5[5 * k + d] = real ? (double)stv(row, S_P + d) : NAN;
        if (lens2) for (int d = 0; d < 2; d++) lens2[2 * k + d] = real ? (double)stv(row, S_L + d) : NAN;
        // contact point and uN exist for real contacts only (L6Geom): those of the state row, kept
        // across full runs and checkpoint uploads; zeros (and NaN uN) otherwise
        if (real && (row < 0 || row >= nRows)) throw WdmError(WDM_E_STATE, "contact state row out of range");
        const double4 rec = real ? cpt[row] : make_double4(0., 0., 0., NAN);
        if (contPt3) { contPt3[3 * k] = rec.x; contPt3[3 * k + 1] = rec.y; contPt3[3 * k + 2] = rec.z; }
        if (uN) uN[k] = rec.w;
        if (trsf9) {  // columns 0 and 1 stored, column 2 = col0 x col1 in the state's precision
            if (real && row >= 0) {
                T c0[3], c1[3];
                for (int q = 0; q < 3; q++) { c0[q] = stv(row, S_T + q); c1[q] = stv(row, S_T + 3 + q); }
                const T c2[3] = {c0[1] * c1[2] - c0[2] * c1[1], c0[2] * c1[0] - c0[0] * c1[2], c0[0] * c1[1] - c0[1] * c1[0]};
                for (int q = 0; q < 3; q++) {
                    trsf9[9 * k + 3 * q] = (double)c0[q];
                    trsf9[9 * k + 3 * q + 1] = (double)c1[q];
                    trsf9[9 * k + 3 * q + 2] = (double)c2[q];
                }
            } else for (int q = 0; q < 9; q++) trsf9[9 * k + q] = NAN;
        }
        k++;
    }
    if (nOut) *nOut = k;
}

// real contacts of a checkpoint become the scene's contact list (state rows in key order);
// the next step runs a collider full run, which keeps them and adds the potential contacts
template <typename T>
void uploadContacts(Scene& S, ConBuf<T>* cb, int64_t n, const int32_t* idA, const int32_t* idB, const int32_t* cellDist3,
                    const double* trsf9, const double* forceLocal3, const double* phys5, const double* lens2, const double* contPt3) {
    Ctx* c = S.ctx;
    std::vector<uint64_t> key(n);
    std::vector<int64_t> ord(n);
    for (int64_t k = 0; k < n; k++) {
        const int a = idA[k], b = idB[k];
        if (a < 0 || b < 0 || a >= S.np || b >= S.np || a == b) throw WdmError(WDM_E_ARG, "contact ids out of range");
        const int sa = S.ptypeH[a] & 15, sb = S.ptypeH[b] & 15;
        int A, B, ty;
        canonical(a, b, sa, sb, A, B, ty);
        if (ty == (int)CI_NONE || A != a || B != b) throw WdmError(WDM_E_ARG, "contact " + std::to_string(k) + ": ids not in functor order or no functor");
        for (int q = 0; q < 3; q++)
            if (cellDist3[3 * k + q] < -511 || cellDist3[3 * k + q] > 511) throw WdmError(WDM_E_ARG, "cellDist out of range");
        const int lo = std::min(a, b), hi = std::max(a, b);
        key[k] = ((uint64_t)(uint32_t)lo << 32) | (uint32_t)hi;
        ord[k] = k;
    }
    std::sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) { return key[x] < key[y]; });
    for (int64_t k = 1; k < n; k++)
        if (key[ord[k]] == key[ord[k - 1]]) throw WdmError(WDM_E_ARG, "duplicate contact");
    ConBuf<T>& b = cb[S.cur];
    b.reserve(c, std::max<int64_t>(n, 1));
    const int64_t cap = b.cap;
    std::vector<uint64_t> hk(n);
    std::vector<int2> ids(n);
    std::vector<uint32_t> info(n);
    std::vector<int32_t> cd(n), slot(n);
    std::vector<T> st((size_t)NSTATE * cap, (T)0);
    std::vector<double4> cpt(std::max<int64_t>(n, 1));
    for (int64_t r = 0; r < n; r++) {
        const int64_t k = ord[r];
        hk[r] = key[k];
        ids[r] = make_int2(idA[k], idB[k]);
        int A, B, ty;
        canonical(idA[k], idB[k], S.ptypeH[idA[k]] & 15, S.ptypeH[idB[k]] & 15, A, B, ty);
        info[r] = (uint32_t)ty | CI_REAL;
        cd[r] = packCd(cellDist3[3 * k], cellDist3[3 * k + 1], cellDist3[3 * k + 2]);
        slot[r] = (int32_t)r;
        const double* t = trsf9 + 9 * k;
        for (int q = 0; q < 3; q++) {
            st[(size_t)(S_T + q) * cap + r] = (T)t[3 * q];          // column 0
            st[(size_t)(S_T + 3 + q) * cap + r] = (T)t[3 * q + 1];  // column 1
            st[(size_t)(S_F + q) * cap + r] = (T)forceLocal3[3 * k + q];
        }
        for (int q = 0; q < 5; q++) st[(size_t)(S_P + q) * cap + r] = (T)phys5[5 * k + q];
        for (int q = 0; q < 2; q++) st[(size_t)(S_L + q) * cap + r] = (T)lens2[2 * k + q];
        // uN is not part of a checkpoint: unknown (NaN) until the next ContactLoop computes it
        cpt[r] = make_double4(contPt3[3 * k], contPt3[3 * k + 1], contPt3[3 * k + 2], NAN);
    }
    hipStream_t st0 = c->stream;
    const int32_t nR = (int32_t)n, zero = 0;
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(b.key.p, hk.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice, st0));
        HIPCHK(hipMemcpyAsync(b.ids.p, ids.data(), n * sizeof(int2), hipMemcpyHostToDevice, st0));
        HIPCHK(hipMemcpyAsync(b.info.p, info.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, st0));
        HIPCHK(hipMemcpyAsync(b.cd.p, cd.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st0));
        HIPCHK(hipMemcpyAsync(b.slot.p, slot.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st0));
        HIPCHK(hipMemcpyAsync(b.st.p, st.data(), st.size() * sizeof(T), hipMemcpyHostToDevice, st0));
        if constexpr (sizeof(T) == 8) HIPCHK(hipMemcpyAsync(b.cpt.p, cpt.data(), n * sizeof(double4), hipMemcpyHostToDevice, st0));
        else {  // stored as branches of B
            Shared& X = *c->sh;
            ensure(c, X.cptTmp, (size_t)cap);
            HIPCHK(hipMemcpyAsync(X.cptTmp.p, cpt.data(), n * sizeof(double4), hipMemcpyHostToDevice, st0));
        }
    }
    HIPCHK(hipMemcpyAsync(b.counter.p, &nR, sizeof(int32_t), hipMemcpyHostToDevice, st0));
    HIPCHK(hipMemcpyAsync(b.nItems.p, &zero, sizeof(int32_t), hipMemcpyHostToDevice, st0));
    if constexpr (sizeof(T) == 4) {
        if (n > 0) {
            hipLaunchKernelGGL(k_cpt_mixed<1>, dim3(nblk(n)), dim3(WDM_BLOCK), 0, st0, (const double4*)S.posr.p, b.view(), n, S.periodic ? 1 : 0,
                               toM3(S.hSize), c->sh->cptTmp.p);
            checkLaunch();
        }
    }
    HIPCHK(hipStreamSynchronize(st0));
    S.nc = n;
    S.needFull = true;
    S.contactsRun = false;
}
template <typename T>
int64_t countLive(Scene& S, ConBuf<T>* cb) {
    ConBuf<T>& b = cb[S.cur];
    int64_t n = S.nc;
    if (n == 0) return 0;
    std::vector<uint32_t> info(n);
    HIPCHK(hipMemcpyAsync(info.data(), b.info.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, S.ctx->stream));
    HIPCHK(hipStreamSynchronize(S.ctx->stream));
    int64_t k = 0;
    for (uint32_t x : info) if (!(x & CI_DEAD)) k++;
    return k;
}

// scene generation: Jacobi overlap relaxation sweeps
void relax(Scene& S, int iters) {
    Ctx* c = S.ctx;
    Shared& X = *c->sh;
    if (S.binnedH.empty()) return;
    ensure(c, X.disp, std::max(S.nn, 1));
    Batch B(c, {&S});
    B.upload();
    std::vector<int> L = {0};
    std::vector<int64_t> bpP = blockPrefix(1, [&](int) { return S.np; });
    std::vector<int64_t> bpB = blockPrefix(1, [&](int) { return (int64_t)S.binnedH.size(); });
    Map mP = uploadMap(c, X.stMapP, bpP), mB = uploadMap(c, X.stMapB, bpB);
    for (int it = 0; it < iters; it++) {
        std::vector<FRScene> F(1);
        std::memset(F.data(), 0, sizeof(FRScene));
        F[0].desc = 0;
        const FRScene* FRd = (const FRScene*)X.stFR.upload(c, F.data(), sizeof(FRScene));
        hipLaunchKernelGGL(k_bounds, dim3((unsigned)bpP[1]), dim3(WDM_BLOCK), 0, c->stream, B.D, FRd, mP, 1);
        checkLaunch();
        int64_t nbTot = 0;
        InsertionSortCollider::buildGrids(B, L, F, bpB, mB, FRd, nbTot, 1., 1, false);
        hipLaunchKernelGGL(k_relax, dim3((unsigned)bpB[1]), dim3(WDM_BLOCK), 0, c->stream, B.D, FRd, mB, (const uint32_t*)X.ckeys2.p,
                           (const int32_t*)X.cvals2.p, (const int32_t*)X.cbeg.p, X.disp.p);
        hipLaunchKernelGGL(k_relax_apply, dim3((unsigned)bpB[1]), dim3(WDM_BLOCK), 0, c->stream, B.D, FRd, mB, (const double4*)X.disp.p);
        checkLaunch();
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    S.needFull = true;
        S.contactsRun = false;
}

// scene words slot (Verlet trigger, created energy slots, device errors)
int32_t takeSlot(Ctx* c) {
    Shared& X = *c->sh;
    int32_t s;
    if (!X.freeSlots.empty()) { s = X.freeSlots.back(); X.freeSlots.pop_back(); }
    else {
        s = X.nslots;
        size_t need = 4 * ((size_t)s + 1);
        ensure(c, X.words, need, true);
        if (X.hostWordsCap < X.words.n) {
            HIPCHK(hipStreamSynchronize(c->stream));
            if (X.hostWords) (void)hipHostFree(X.hostWords);
            X.hostWords = nullptr;
            HIPCHK(hipHostMalloc((void**)&X.hostWords, X.words.n * sizeof(uint32_t), hipHostMallocDefault));
            X.hostWordsCap = X.words.n;
        }
        X.nslots = s + 1;
    }
    HIPCHK(hipMemsetAsync(X.words.p + 4 * (size_t)s, 0, 4 * sizeof(uint32_t), c->stream));
    return s;
}

// ---------------- the selection of an outlet without allocation ----------------
// The address of `scratch` on each device (declared at the top of this file), looked up at
// wdm_create; kernels and copies use it, never the array's name.
#define NOALLOC_MAX_DEVICES 64
std::mutex g_scratchMutex;
uint8_t* g_scratchPtr[NOALLOC_MAX_DEVICES] = {};
// the device address of `scratch` on the current device `dev` (looked up once per device)
hipError_t scratchInit(int dev) {
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratchMutex);
    if (g_scratchPtr[dev]) return hipSuccess;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratchPtr[dev] = (uint8_t*)p;
    return hipSuccess;
}
uint8_t* scratchOf(int dev) {
    std::lock_guard<std::mutex> lk(g_scratchMutex);
    return (dev >= 0 && dev < NOALLOC_MAX_DEVICES) ? g_scratchPtr[dev] : nullptr;
}

}  // namespace

struct wdm_ctx : Ctx {};

extern "C" {

int wdm_select_box_noalloc(wdm_ctx* c, int32_t sid, const double* lo3, const double* hi3, int32_t inside, uint32_t mask, uint32_t markMask,
                           int32_t cap, int32_t* ids, int32_t* nOut) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (!lo3 || !hi3 || !nOut || cap < 0 || (cap > 0 && !ids)) throw WdmError(WDM_E_ARG, "box arrays");
        if (S.np > NOALLOC_MAX_PARTICLES)
            throw WdmError(WDM_E_ARG, "wdm_select_box_noalloc: " + std::to_string(S.np) + " particles, more than its limit of " +
                                          std::to_string(NOALLOC_MAX_PARTICLES) + " (use wdm_select_box)");
        uint8_t* base = scratchOf(c->device);
        if (!base)
            throw WdmError(WDM_E_HIP, std::string("wdm_select_box_noalloc: hipGetSymbolAddress of its device memory failed: ") +
                                          hipGetErrorString(c->scratchErr));
        int32_t* counts = (int32_t*)base;
        int32_t* offs = (int32_t*)(base + NOALLOC_OFF_OFFS);
        int32_t* total = (int32_t*)(base + NOALLOC_OFF_TOTAL);
        int32_t* dIds = (int32_t*)(base + NOALLOC_OFF_IDS);
        const int nb = (int)nblk(S.np);
        const double3 lo = make_double3(lo3[0], lo3[1], lo3[2]), hi = make_double3(hi3[0], hi3[1], hi3[2]);
        hipStream_t st = c->stream;
        // every count, offset and the total are written by the kernels below before they are read
        if (nb > 0)
            hipLaunchKernelGGL(k_select_count, dim3(nb), dim3(WDM_BLOCK), 0, st, (const double4*)S.posr.p, (const int32_t*)S.ptype.p,
                               (const uint32_t*)S.pmask.p, S.np, lo, hi, inside, mask, markMask, counts);
        hipLaunchKernelGGL(k_select_scan, dim3(1), dim3(1024), 0, st, (const int32_t*)counts, nb, offs, total);
        if (nb > 0)
            hipLaunchKernelGGL(k_select_write, dim3(nb), dim3(WDM_BLOCK), 0, st, (const double4*)S.posr.p, (const int32_t*)S.ptype.p,
                               (const uint32_t*)S.pmask.p, S.np, lo, hi, inside, mask, markMask, (const int32_t*)offs, dIds);
        checkLaunch();
        int32_t n = 0;
        HIPCHK(hipMemcpyAsync(&n, total, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (n < 0 || n > S.np) throw WdmError(WDM_E_STATE, "selection count out of range");
        const int32_t k = std::min(n, cap);
        if (k > 0) {
            HIPCHK(hipMemcpyAsync(ids, dIds, (size_t)k * sizeof(int32_t), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
        }
        *nOut = n;
    });
}

int wdm_create(wdm_ctx** out, int device, size_t budget) {
    if (!out) return WDM_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return WDM_E_HIP;
    if (device < 0 || device >= n) return WDM_E_ARG;
    wdm_ctx* c = new wdm_ctx();
    c->device = device;
    c->budget = budget;
    c->sh = new Shared();
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&c->sh->evWords, hipEventDisableTiming) != hipSuccess) {
        delete c->sh;
        delete c;
        return WDM_E_HIP;
    }
    // the device address of the no-allocation entry point's memory, once per device, with that
    // device current; a failure is kept and reported by that entry point only
    c->scratchErr = scratchInit(device);
    if (const char* lv = std::getenv("WDM_GRID_LEVELS")) c->gridLevels = std::max(0, std::min(MAXLVL, std::atoi(lv)));
    *out = c;
    return WDM_OK;
}

void wdm_destroy(wdm_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->scenes) { kv.second->freeAll(); delete kv.second; }
    c->scenes.clear();
    c->sh->freeAll(c);
    delete c->sh;
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (hipEvent_t e : c->timing.ev) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* wdm_last_error(wdm_ctx* c) { return c ? c->err.c_str() : "null context"; }

size_t wdm_scene_footprint(int32_t n_par, int32_t n_node, double cpp) {
    size_t nodeB = (size_t)n_node * (7 * 32 + 4 + 1);
    size_t parB = (size_t)n_par * (4 + 4 + 4 + 16 + 64 + 8 + 16);
    double ncon = cpp * n_par * 1.34;
    size_t conB = (size_t)(ncon * 2 * (8 + 8 + 4 + 4 + 4 + 4 + (NSTATE + NSTRIDE) * 8 + 32 + 64));
    size_t csrB = (size_t)(ncon * 2 * (4 + 1));
    size_t scratch = (size_t)n_par * 64 + (size_t)(ncon * 48);
    return nodeB + parB + conB + csrB + scratch + (1u << 20);
}

int wdm_scene_alloc(wdm_ctx* c, int32_t np, int32_t nn, int32_t nm, int32_t* sid) {
    return guard(c, [&] {
        if (np < 0 || nn < np || nm < 1 || !sid) throw WdmError(WDM_E_ARG, "bad sizes (need n_node >= n_par, n_mat >= 1)");
        if (np >= (1 << 30)) throw WdmError(WDM_E_ARG, "too many particles");
        Scene* S = new Scene(c, np, nn, nm);
        try {
            ensure(c, S->posr, std::max(nn, 1)); ensure(c, S->velm, std::max(nn, 1)); ensure(c, S->angI, std::max(nn, 1));
            ensure(c, S->ori, std::max(nn, 1)); ensure(c, S->force, std::max(nn, 1)); ensure(c, S->torque, std::max(nn, 1));
            ensure(c, S->lastPos, std::max(nn, 1)); ensure(c, S->flags, std::max(nn, 1)); ensure(c, S->checked, std::max(nn, 1));
            ensure(c, S->ptype, std::max(np, 1)); ensure(c, S->pmat, std::max(np, 1)); ensure(c, S->pmask, std::max(np, 1));
            ensure(c, S->fnodes, std::max(np, 1)); ensure(c, S->bmin, std::max(np, 1)); ensure(c, S->bmax, std::max(np, 1));
            ensure(c, S->mat, nm);
            ensure(c, S->E, WDM_E_COUNT);
            ensure(c, S->csrOff, (size_t)np + 1);
            ensure(c, S->entV, 1);
            ensure(c, S->realMask, 1);
            HIPCHK(hipMemsetAsync(S->E.p, 0, sizeof(double) * WDM_E_COUNT, c->stream));
            HIPCHK(hipMemsetAsync(S->csrOff.p, 0, sizeof(int64_t) * ((size_t)np + 1), c->stream));
            HIPCHK(hipMemsetAsync(S->realMask.p, 0, sizeof(uint64_t), c->stream));
            S->slot = takeSlot(c);
            HIPCHK(hipStreamSynchronize(c->stream));
        } catch (...) {
            if (S->slot >= 0) c->sh->freeSlots.push_back(S->slot);
            S->freeAll();
            delete S;
            throw;
        }
        *sid = c->nextId++;
        S->id = *sid;
        c->scenes[*sid] = S;
    });
}

int wdm_scene_error(wdm_ctx* c, int32_t sid, int32_t* code) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (!code) throw WdmError(WDM_E_ARG, "null output");
        *code = S.failed;
        if (S.failed) c->err = S.failText;
    });
}

int wdm_scene_free(wdm_ctx* c, int32_t sid) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        HIPCHK(hipStreamSynchronize(c->stream));
        c->sh->dropGraph(c);  // no window graph outlives the memory of a scene it stepped
        S.freeAll();
        c->sh->freeSlots.push_back(S.slot);
        c->scenes.erase(sid);
        delete &S;
    });
}

int wdm_upload_nodes(wdm_ctx* c, int32_t sid, int32_t n, const double* pos3, const double* ori4, const double* vel3, const double* angVel3,
                     const double* mass, const double* inertia3, const uint32_t* flags, const double* force3, const double* torque3) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (n != S.nn || !pos3 || !ori4 || !vel3 || !angVel3 || !mass || !inertia3 || !flags) throw WdmError(WDM_E_ARG, "node arrays");
        std::vector<double4> pr(n), vm(n), aw(n), q(n), f(n), t(n);
        std::vector<double> rad(n, 0.);
        if (S.partsSet) for (int p = 0; p < S.np; p++) rad[p] = S.radiusH[p];
        int64_t nclf = 0;
        for (int i = 0; i < n; i++) {
            if (flags[i] & (WDM_CLUMP_CLUMP | WDM_CLUMP_CLUMPED)) nclf++;
            const double* I = inertia3 + 3 * i;
            bool asph = !(I[0] == I[1] && I[1] == I[2]);
            if (asph && (flags[i] & 56u) != 56u) throw WdmError(WDM_E_UNSUPPORTED, "aspherical integrator is not supported (node " + std::to_string(i) + ")");
            pr[i] = make_double4(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], rad[i]);
            vm[i] = make_double4(vel3[3 * i], vel3[3 * i + 1], vel3[3 * i + 2], mass[i]);
            aw[i] = make_double4(angVel3[3 * i], angVel3[3 * i + 1], angVel3[3 * i + 2], asph ? 0. : I[0]);
            q[i] = make_double4(ori4[4 * i], ori4[4 * i + 1], ori4[4 * i + 2], ori4[4 * i + 3]);
            f[i] = make_double4(force3 ? force3[3 * i] : 0., force3 ? force3[3 * i + 1] : 0., force3 ? force3[3 * i + 2] : 0., 0.);
            t[i] = make_double4(torque3 ? torque3[3 * i] : 0., torque3 ? torque3[3 * i + 1] : 0., torque3 ? torque3[3 * i + 2] : 0., 0.);
        }
        hipStream_t st = c->stream;
        HIPCHK(hipMemcpyAsync(S.posr.p, pr.data(), n * sizeof(double4), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(S.velm.p, vm.data(), n * sizeof(double4), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(S.angI.p, aw.data(), n * sizeof(double4), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(S.ori.p, q.data(), n * sizeof(double4), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(S.force.p, f.data(), n * sizeof(double4), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(S.torque.p, t.data(), n * sizeof(double4), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(S.flags.p, flags, n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        S.nodesSet = true;
        S.needFull = true;
        S.contactsRun = false;
        S.forcesFromContacts = false;
        S.forcesReset = false;
        S.clumpFlagged = nclf;
        S.flagsH.assign(flags, flags + n);
        S.nClumps = S.nMembers = 0;  // clumps are uploaded again after the nodes
        S.clumpsSet = nclf == 0;
    });
}

int wdm_upload_particles(wdm_ctx* c, int32_t sid, int32_t n, const int32_t* shape, const int32_t* nodes3, const double* radius,
                         const int32_t* axis, const int32_t* sense, const uint32_t* mask, const int32_t* mat) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (n != S.np || !shape || !nodes3 || !radius || !axis || !sense || !mask || !mat) throw WdmError(WDM_E_ARG, "particle arrays");
        std::vector<int32_t> pt(n);
        std::vector<int4> fn(n);
        std::vector<uint8_t> chk(S.nn, 0);
        S.binnedH.clear(); S.unbH.clear();
        for (int p = 0; p < n; p++) {
            int sh = shape[p];
            if (sh < 0 || sh > 3) throw WdmError(WDM_E_ARG, "bad shape code");
            if (mat[p] < 0 || mat[p] >= S.nmat) throw WdmError(WDM_E_ARG, "bad material index");
            if (sh == WDM_SHAPE_SPHERE || sh == WDM_SHAPE_WALL) {
                if (nodes3[3 * p] != p) throw WdmError(WDM_E_ARG, "uninodal particle " + std::to_string(p) + " must use node index == particle index");
                chk[p] = 1;
            }
            if (sh == WDM_SHAPE_WALL && (axis[p] < 0 || axis[p] > 2 || sense[p] < -1 || sense[p] > 1)) throw WdmError(WDM_E_ARG, "bad wall axis/sense");
            if (sh == WDM_SHAPE_FACET) {
                for (int k = 0; k < 3; k++) {
                    int q = nodes3[3 * p + k];
                    if (q < S.np || q >= S.nn) throw WdmError(WDM_E_ARG, "facet nodes must have index >= n_par");
                    chk[q] = 1;
                }
                fn[p] = make_int4(nodes3[3 * p], nodes3[3 * p + 1], nodes3[3 * p + 2], 0);
            } else fn[p] = make_int4(p, -1, -1, 0);
            pt[p] = sh | ((axis[p] & 3) << 4) | (((sense[p] + 1) & 3) << 6);
            if (sh == WDM_SHAPE_SPHERE || sh == WDM_SHAPE_FACET) S.binnedH.push_back(p);
            else if (sh == WDM_SHAPE_WALL) S.unbH.push_back(p);
        }
        S.ptypeH = pt;
        S.radiusH.assign(radius, radius + n);
        S.facetNodesH.clear();
        for (int p = 0; p < n; p++)
            if (shape[p] == WDM_SHAPE_FACET)
                for (int k = 0; k < 3; k++) S.facetNodesH.push_back(nodes3[3 * p + k]);
        S.parEnd = 0;
        for (int p = 0; p < n; p++)
            if (shape[p] != WDM_SHAPE_NONE) S.parEnd = p + 1;
        S.freeIds.clear();
        for (int p = 0; p < S.parEnd; p++)
            if (shape[p] == WDM_SHAPE_NONE) S.freeIds.push_back(p);
        S.hasFacets = false;
        S.facetMasks.clear();
        for (int p = 0; p < n; p++)
            if (shape[p] == WDM_SHAPE_FACET) {
                S.hasFacets = true;
                if (std::find(S.facetMasks.begin(), S.facetMasks.end(), mask[p]) == S.facetMasks.end()) S.facetMasks.push_back(mask[p]);
            }
        hipStream_t st = c->stream;
        HIPCHK(hipMemcpyAsync(S.ptype.p, pt.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(S.pmask.p, mask, n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(S.pmat.p, mat, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(S.fnodes.p, fn.data(), n * sizeof(int4), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(S.checked.p, chk.data(), S.nn * sizeof(uint8_t), hipMemcpyHostToDevice, st));
        ensure(c, S.binned, std::max<size_t>(S.binnedH.size(), 1));
        ensure(c, S.unb, std::max<size_t>(S.unbH.size(), 1));
        if (!S.binnedH.empty()) HIPCHK(hipMemcpyAsync(S.binned.p, S.binnedH.data(), S.binnedH.size() * 4, hipMemcpyHostToDevice, st));
        if (!S.unbH.empty()) HIPCHK(hipMemcpyAsync(S.unb.p, S.unbH.data(), S.unbH.size() * 4, hipMemcpyHostToDevice, st));
        // radius lives in posr.w of the particle's node
        if (S.nodesSet) {
            std::vector<double4> pr(S.nn);
            HIPCHK(hipMemcpyAsync(pr.data(), S.posr.p, S.nn * sizeof(double4), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            for (int p = 0; p < n; p++) pr[p].w = radius[p];
            HIPCHK(hipMemcpyAsync(S.posr.p, pr.data(), S.nn * sizeof(double4), hipMemcpyHostToDevice, st));
        }
        HIPCHK(hipStreamSynchronize(st));
        S.partsSet = true;
        S.needFull = true;
        S.contactsRun = false;
        S.verletResolved = false;
        S.nc = 0;
    });
}

int wdm_upload_materials(wdm_ctx* c, int32_t sid, int32_t n, const double* density, const double* young, const double* ktDivKn,
                         const double* tanPhi) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (n != S.nmat || !density || !young || !ktDivKn || !tanPhi) throw WdmError(WDM_E_ARG, "material arrays");
        std::vector<double4> m(n);
        for (int i = 0; i < n; i++) m[i] = make_double4(density[i], young[i], ktDivKn[i], tanPhi[i]);
        HIPCHK(hipMemcpyAsync(S.mat.p, m.data(), n * sizeof(double4), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        S.matsSet = true;
    });
}

int wdm_upload_clumps(wdm_ctx* c, int32_t sid, int32_t nc, const int32_t* master, const int32_t* off, const int32_t* members,
                      const double* relPos3, const double* relOri4) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (!S.nodesSet) throw WdmError(WDM_E_STATE, "upload nodes before clumps");
        if (nc < 0 || (nc > 0 && (!master || !off || !members || !relPos3 || !relOri4))) throw WdmError(WDM_E_ARG, "clump arrays");
        const int nn = S.nn;
        std::vector<int32_t> nodeMaster(nn, -1), nodeClump(nn, -1);
        if (nc > 0 && off[0] != 0) throw WdmError(WDM_E_ARG, "clump offsets must start at 0");
        for (int k = 0; k < nc; k++) {
            int M = master[k];
            if (M < 0 || M >= nn || !(S.flagsH[M] & WDM_CLUMP_CLUMP) || nodeClump[M] >= 0) throw WdmError(WDM_E_ARG, "bad clump master node");
            if (off[k + 1] < off[k]) throw WdmError(WDM_E_ARG, "clump offsets must not decrease");
            nodeClump[M] = k;
            for (int j = off[k]; j < off[k + 1]; j++) {
                int m = members[j];
                if (m < 0 || m >= nn || !(S.flagsH[m] & WDM_CLUMP_CLUMPED) || nodeMaster[m] >= 0 || (S.flagsH[m] & WDM_CLUMP_CLUMP))
                    throw WdmError(WDM_E_ARG, "bad clump member node " + std::to_string(m));
                nodeMaster[m] = M;
            }
        }
        for (int i = 0; i < nn; i++) {
            if ((S.flagsH[i] & WDM_CLUMP_CLUMP) && nodeClump[i] < 0) throw WdmError(WDM_E_ARG, "clump master node without a clump");
            if ((S.flagsH[i] & WDM_CLUMP_CLUMPED) && nodeMaster[i] < 0) throw WdmError(WDM_E_ARG, "clumped node without a master");
        }
        const int nm = nc > 0 ? off[nc] : 0;
        std::vector<double4> rp(nm), ro(nm);
        for (int j = 0; j < nm; j++) {
            rp[j] = make_double4(relPos3[3 * j], relPos3[3 * j + 1], relPos3[3 * j + 2], 0.);
            ro[j] = make_double4(relOri4[4 * j], relOri4[4 * j + 1], relOri4[4 * j + 2], relOri4[4 * j + 3]);
        }
        hipStream_t st = c->stream;
        if (nc > 0) {
            ensure(c, S.master, nn); ensure(c, S.nodeClump, nn); ensure(c, S.cOff, (size_t)nc + 1);
            ensure(c, S.cMem, std::max(nm, 1)); ensure(c, S.cRelPos, std::max(nm, 1)); ensure(c, S.cRelOri, std::max(nm, 1));
            ensure(c, S.cFT, 2 * (size_t)nc);
            HIPCHK(hipMemcpyAsync(S.master.p, nodeMaster.data(), nn * sizeof(int32_t), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(S.nodeClump.p, nodeClump.data(), nn * sizeof(int32_t), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(S.cOff.p, off, ((size_t)nc + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
            if (nm > 0) {
                HIPCHK(hipMemcpyAsync(S.cMem.p, members, nm * sizeof(int32_t), hipMemcpyHostToDevice, st));
                HIPCHK(hipMemcpyAsync(S.cRelPos.p, rp.data(), nm * sizeof(double4), hipMemcpyHostToDevice, st));
                HIPCHK(hipMemcpyAsync(S.cRelOri.p, ro.data(), nm * sizeof(double4), hipMemcpyHostToDevice, st));
            }
            HIPCHK(hipStreamSynchronize(st));
        }
        S.nClumps = nc;
        S.nMembers = nm;
        S.nodeMasterH = nodeMaster;
        S.clumpFree = false;
        for (int k = 0; k < nc; k++) S.clumpFree |= (S.flagsH[master[k]] & 63u) != 63u;
        S.clumpsSet = true;
        S.needFull = true;
        S.contactsRun = false;
    });
}

int wdm_set_cell(wdm_ctx* c, int32_t sid, int32_t periodic, const double* hSize9, const double* gradV9, const double* nextGradV9,
                 int32_t homoDeform) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (homoDeform < 0 || homoDeform > 4) throw WdmError(WDM_E_ARG, "homoDeform must be in 0..4");
        S.periodic = periodic != 0;
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
            S.hSize.m[i][j] = hSize9 ? hSize9[3 * i + j] : (i == j);
            S.gradV.m[i][j] = gradV9 ? gradV9[3 * i + j] : 0.;
            S.nextGradV.m[i][j] = nextGradV9 ? nextGradV9[3 * i + j] : 0.;
        }
        S.homoDeform = homoDeform;
        S.trsf = heye();
        S.pprevHsize = S.hSize;
        HM3 W = hscale(hsub(S.gradV, htrans(S.gradV)), .5);
        leviCivitaH(W, S.spinVec);
        for (double& x : S.spinVec) x = x * .5;
        if (S.periodic) cellIntegrateAndUpdate(S, 0);  // Cell::postLoad
        S.needFull = true;
        S.contactsRun = false;
    });
}

int wdm_set_params(wdm_ctx* c, int32_t sid, const wdm_params* p) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (!p) throw WdmError(WDM_E_ARG, "null params");
        if (p->law != WDM_LAW_IDEALELPL && p->law != WDM_LAW_HERTZ_DMT) throw WdmError(WDM_E_ARG, "unknown law");
        if (p->law == WDM_LAW_HERTZ_DMT && (p->gamma != 0 || p->alpha != 0)) throw WdmError(WDM_E_UNSUPPORTED, "adhesive Hertz (gamma/alpha != 0) is not supported");
        if (p->precision != WDM_PREC_FP64 && p->precision != WDM_PREC_MIXED) throw WdmError(WDM_E_ARG, "unknown precision");
        if (S.paramsSet && p->precision != S.P.precision && S.nc > 0) throw WdmError(WDM_E_STATE, "precision cannot change after contacts exist");
        S.verletResolved = false;
        S.P = *p;
        S.paramsSet = true;
    });
}

int wdm_step_begin(wdm_ctx* c, int32_t sid) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        checkReady(S);
        relayoutDueScenes(c, {&S});
        stepBeginCell(S);
        if (S.P.trackEnergy)  // EnergyTracker::resetResettables: kinetic and elast
            HIPCHK(hipMemsetAsync(S.E.p + WDM_E_KINETIC, 0, sizeof(double) * 2, c->stream));
    });
}
int wdm_leapfrog_run(wdm_ctx* c, int32_t sid) {
    return guard(c, [&] {
        Batch B(c, {&getScene(c, sid)});
        Leapfrog::run(B);
    });
}
int wdm_collider_run(wdm_ctx* c, int32_t sid) {
    return guard(c, [&] {
        Batch B(c, {&getScene(c, sid)});
        InsertionSortCollider::run(B);
    });
}
int wdm_contact_loop_run(wdm_ctx* c, int32_t sid) {
    return guard(c, [&] {
        Batch B(c, {&getScene(c, sid)});
        ContactLoop::run(B);
    });
}
int wdm_step_end(wdm_ctx* c, int32_t sid) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        stepEnd(S);
        readWords(c, {&S});
    });
}
int wdm_dyndt_run(wdm_ctx* c, int32_t sid) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        DynDtEngine::fire(c, {&S});
    });
}
int wdm_set_dyndt(wdm_ctx* c, int32_t sid, const wdm_dyndt* d) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (!d) throw WdmError(WDM_E_ARG, "null DynDt attributes");
        if (d->enabled) {
            if (1. + d->maxRelInc == 1.) throw WdmError(WDM_E_ARG, "DynDt: maxRelInc too small (1.0+maxRelInc==1.0)");  // DynDt::postLoad
            if (!(d->dtSafety > 0)) throw WdmError(WDM_E_ARG, "DynDt: dtSafety must be positive");
            DynDtEngine::checkSupported(S);
        }
        S.dd = *d;
    });
}
int wdm_get_dyndt(wdm_ctx* c, int32_t sid, wdm_dyndt* d) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (!d) throw WdmError(WDM_E_ARG, "null DynDt output");
        *d = S.dd;
    });
}
int wdm_dyndt_crit(wdm_ctx* c, int32_t sid, double* crit) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (!crit) throw WdmError(WDM_E_ARG, "null output");
        *crit = DynDtEngine::crit(c, {&S})[0];
    });
}
int wdm_get_time(wdm_ctx* c, int32_t sid, double* time, double* dt) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (time) *time = S.time;
        if (dt) *dt = S.P.dt;
    });
}

int wdm_step(wdm_ctx* c, const int32_t* scenes, int32_t ns, int32_t nsteps) {
    return guard(c, [&] {
        if (ns < 0 || nsteps < 0 || (ns > 0 && !scenes)) throw WdmError(WDM_E_ARG, "bad scene list");
        std::vector<Scene*> ss;
        for (int i = 0; i < ns; i++) {
            Scene* S = &getScene(c, scenes[i]);
            if (std::find(ss.begin(), ss.end(), S) != ss.end()) throw WdmError(WDM_E_ARG, "scene listed twice");
            checkReady(*S);
            ss.push_back(S);
        }
        if (c->timing.on) { c->timing.pend.clear(); c->timing.acc.clear(); }
        if (ss.empty()) return;
        std::vector<int64_t> left(ss.size(), nsteps);
        for (;;) {
            std::vector<Scene*> act;
            std::vector<int64_t> la;
            std::vector<size_t> ai;
            int64_t most = 0;
            for (size_t i = 0; i < ss.size(); i++)
                if (left[i] > 0 && !ss[i]->failed) {
                    act.push_back(ss[i]); la.push_back(left[i]); ai.push_back(i); most = std::max(most, left[i]);
                }
            if (act.empty()) break;
            relayoutDueScenes(c, act);
            for (Scene* S : act) most = std::min(most, std::max<int64_t>(1, stepsToRelayout(c, *S)));
            const int K = (int)std::min<int64_t>(c->window, most);
            bool win = K > 1;
            for (Scene* S : act) win = win && windowable(*S, K);
            if (win) {
                std::vector<int64_t> done = runWindow(c, act, la, K);
                for (size_t k = 0; k < act.size(); k++) left[ai[k]] -= done[k];
            } else {
                Batch B(c, act);
                doOneStep(B);
                for (size_t k = 0; k < act.size(); k++) left[ai[k]] -= 1;
            }
        }
        readWords(c, ss, false);
        // every scene without an error has taken its steps; report the ones that stopped
        std::string ids;
        int code = 0;
        for (Scene* S : ss)
            if (S->failed) {
                ids += (ids.empty() ? "" : ", ") + std::to_string(S->id);
                if (!code) code = S->failed;
            }
        if (code) throw WdmError(code, "scene(s) " + ids + " stopped by a device error (wdm_scene_error); the other scenes took their steps");
    });
}

int wdm_download_nodes(wdm_ctx* c, int32_t sid, double* pos3, double* ori4, double* vel3, double* angVel3, double* force3, double* torque3) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        int n = S.nn;
        if (force3 || torque3) materialiseAll(S);
        std::vector<double4> a(n);
        auto get = [&](DBuf<double4>& b, double* out, int comps, bool quat) {
            if (!out) return;
            HIPCHK(hipMemcpyAsync(a.data(), b.p, n * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int i = 0; i < n; i++) {
                out[comps * i] = a[i].x; out[comps * i + 1] = a[i].y; out[comps * i + 2] = a[i].z;
                if (quat) out[comps * i + 3] = a[i].w;
            }
        };
        get(S.posr, pos3, 3, false);
        get(S.ori, ori4, 4, true);
        get(S.velm, vel3, 3, false);
        get(S.angI, angVel3, 3, false);
        get(S.force, force3, 3, false);
        get(S.torque, torque3, 3, false);
    });
}

int wdm_num_contacts(wdm_ctx* c, int32_t sid, int64_t* n) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (!n) throw WdmError(WDM_E_ARG, "null");
        *n = S.mixed() ? countLive(S, S.cb32) : countLive(S, S.cb64);
    });
}

int wdm_contact_work(wdm_ctx* c, int32_t sid, int64_t* active, int64_t* rows) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        int32_t w[2] = {0, 0};
        const int32_t* pa = S.mixed() ? S.cb32[S.cur].nItems.p : S.cb64[S.cur].nItems.p;
        const int32_t* pr = S.mixed() ? S.cb32[S.cur].counter.p : S.cb64[S.cur].counter.p;
        if (pa) HIPCHK(hipMemcpyAsync(&w[0], pa, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        if (pr) HIPCHK(hipMemcpyAsync(&w[1], pr, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (active) *active = w[0];
        if (rows) *rows = w[1];
    });
}

int wdm_get_contacts(wdm_ctx* c, int32_t sid, int64_t cap, int32_t* idA, int32_t* idB, int32_t* cellDist3, uint8_t* isReal, double* forceLocal3,
                     double* contPt3, double* trsf9, double* uN, int64_t* nOut) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (S.mixed()) downloadContacts(S, S.cb32, cap, idA, idB, cellDist3, isReal, forceLocal3, contPt3, trsf9, uN, nOut);
        else downloadContacts(S, S.cb64, cap, idA, idB, cellDist3, isReal, forceLocal3, contPt3, trsf9, uN, nOut);
    });
}

int wdm_get_contact_state(wdm_ctx* c, int32_t sid, int64_t cap, double* phys5, double* lens2, int64_t* nOut) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (S.mixed()) downloadContacts(S, S.cb32, cap, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nOut, phys5, lens2);
        else downloadContacts(S, S.cb64, cap, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nOut, phys5, lens2);
    });
}
int wdm_upload_contacts(wdm_ctx* c, int32_t sid, int64_t n, const int32_t* idA, const int32_t* idB, const int32_t* cellDist3,
                        const double* trsf9, const double* forceLocal3, const double* phys5, const double* lens2, const double* contPt3) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (!S.partsSet || !S.paramsSet) throw WdmError(WDM_E_STATE, "upload particles and parameters before contacts");
        if (n < 0 || (n > 0 && (!idA || !idB || !cellDist3 || !trsf9 || !forceLocal3 || !phys5 || !lens2 || !contPt3)))
            throw WdmError(WDM_E_ARG, "contact arrays");
        if (n >= INT32_MAX) throw WdmError(WDM_E_ARG, "too many contacts");
        if (S.mixed()) uploadContacts(S, S.cb32, n, idA, idB, cellDist3, trsf9, forceLocal3, phys5, lens2, contPt3);
        else uploadContacts(S, S.cb64, n, idA, idB, cellDist3, trsf9, forceLocal3, phys5, lens2, contPt3);
    });
}
// ---------------- particles removed and inserted in place ----------------
int wdm_remove_particles(wdm_ctx* c, int32_t sid, int32_t n, const int32_t* ids) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        checkReady(S);
        if (n < 0 || (n > 0 && !ids)) throw WdmError(WDM_E_ARG, "particle ids");
        std::vector<char> seen(S.np, 0);
        for (int k = 0; k < n; k++) {
            const int id = ids[k];
            if (id < 0 || id >= S.np) throw WdmError(WDM_E_ARG, "particle id out of range");
            const int sh = S.ptypeH[id] & 15;
            if (sh == WDM_SHAPE_NONE) throw WdmError(WDM_E_ARG, "#" + std::to_string(id) + " does not exist");
            if (sh == WDM_SHAPE_FACET) throw WdmError(WDM_E_UNSUPPORTED, "removing facets is outside the implemented path");
            if (S.flagsH[id] & (WDM_CLUMP_CLUMPED | WDM_CLUMP_CLUMP))
                throw WdmError(WDM_E_UNSUPPORTED, "#" + std::to_string(id) + ": a node is clumped, remove the clump itself instead!");  // Particle.cpp:451
            if (seen[id]++) throw WdmError(WDM_E_ARG, "particle listed twice");
        }
        if (n == 0) return;
        // forces of the last ContactLoop stay with the partners (Woo's ContactLoop applied them
        // before the removal): the next Leapfrog gathers them through the current CSR and
        // realness bits, which the dropped contacts keep until the next full run
        hipStream_t st = c->stream;
        Shared& X = *c->sh;
        const int32_t* dIds = (const int32_t*)X.stDyn.upload(c, ids, n * sizeof(int32_t));
        hipLaunchKernelGGL(k_remove_particles, dim3((n + 255) / 256), dim3(256), 0, st, S.ptype.p, S.flags.p, S.velm.p, S.angI.p,
                           S.checked.p, dIds, n);
        checkLaunch();
        const int64_t nc = S.nc;
        if (nc > 0) {
            uint32_t* info = S.mixed() ? S.cb32[S.cur].info.p : S.cb64[S.cur].info.p;
            const int2* cids = S.mixed() ? S.cb32[S.cur].ids.p : S.cb64[S.cur].ids.p;
            hipLaunchKernelGGL(k_kill_contacts, dim3(nblk(nc)), dim3(WDM_BLOCK), 0, st, info, cids, nc, (const int32_t*)S.ptype.p);
            checkLaunch();
        }
        for (int k = 0; k < n; k++) {
            S.ptypeH[ids[k]] = 0;
            S.flagsH[ids[k]] = 63u | WDM_ENERGY_SKIP | WDM_GRAVITY_SKIP | WDM_DAMPING_SKIP;
            S.freeIds.push_back(ids[k]);
        }
        auto keep = [&](std::vector<int32_t>& v) { v.erase(std::remove_if(v.begin(), v.end(), [&](int32_t q) { return seen[q] != 0; }), v.end()); };
        keep(S.binnedH);
        keep(S.unbH);
        if (!S.binnedH.empty()) HIPCHK(hipMemcpyAsync(S.binned.p, S.binnedH.data(), S.binnedH.size() * 4, hipMemcpyHostToDevice, st));
        if (!S.unbH.empty()) HIPCHK(hipMemcpyAsync(S.unb.p, S.unbH.data(), S.unbH.size() * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        // the particle set changed: the collider runs a full run (InsertionSortCollider.cpp:475-492)
        S.needFull = true;
        S.contactsRun = false;
    });
}
int wdm_insert_spheres(wdm_ctx* c, int32_t sid, int32_t n, const double* pos3, const double* vel3, const double* angVel3,
                       const double* radius, const double* mass, const double* inertia, const int32_t* mat, const uint32_t* mask,
                       int32_t* idsOut) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        checkReady(S);
        if (n < 0 || (n > 0 && (!pos3 || !radius || !mass || !inertia || !mat || !mask || !idsOut))) throw WdmError(WDM_E_ARG, "sphere arrays");
        if ((int64_t)S.freeIds.size() + (S.np - S.parEnd) < n)
            throw WdmError(WDM_E_ARG, "no free particle slot (allocate the scene with spare particles of shape NONE)");
        for (int k = 0; k < n; k++) {
            if (!(radius[k] > 0)) throw WdmError(WDM_E_ARG, "sphere radius must be positive");
            if (mat[k] < 0 || mat[k] >= S.nmat) throw WdmError(WDM_E_ARG, "bad material index");
        }
        if (n == 0) return;
        // the node forces (of the last ContactLoop, or Leapfrog's reset) go to the arrays, where
        // the new nodes get zero
        materialiseAll(S);
        S.forcesFromContacts = false;
        std::vector<InsRec> rec(n);
        for (int k = 0; k < n; k++) {
            int32_t id;
            if (!S.freeIds.empty()) { id = S.freeIds.front(); S.freeIds.pop_front(); }
            else id = S.parEnd++;
            idsOut[k] = id;
            InsRec& r = rec[k];
            r.posr = make_double4(pos3[3 * k], pos3[3 * k + 1], pos3[3 * k + 2], radius[k]);
            r.velm = vel3 ? make_double4(vel3[3 * k], vel3[3 * k + 1], vel3[3 * k + 2], mass[k]) : make_double4(0., 0., 0., mass[k]);
            r.angI = angVel3 ? make_double4(angVel3[3 * k], angVel3[3 * k + 1], angVel3[3 * k + 2], inertia[k]) : make_double4(0., 0., 0., inertia[k]);
            r.id = id; r.mat = mat[k]; r.mask = mask[k]; r.pad = 0;
            S.ptypeH[id] = WDM_SHAPE_SPHERE;
            S.radiusH[id] = radius[k];
            S.flagsH[id] = 0;
            S.binnedH.push_back(id);
        }
        hipStream_t st = c->stream;
        const InsRec* d = (const InsRec*)c->sh->stDyn.upload(c, rec.data(), n * sizeof(InsRec));
        hipLaunchKernelGGL(k_insert_spheres, dim3((n + 255) / 256), dim3(256), 0, st, S.nview(), S.pview(), d, n);
        checkLaunch();
        ensure(c, S.binned, std::max<size_t>(S.binnedH.size(), 1));
        HIPCHK(hipMemcpyAsync(S.binned.p, S.binnedH.data(), S.binnedH.size() * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        S.needFull = true;
        S.contactsRun = false;
    });
}
int wdm_get_free_ids(wdm_ctx* c, int32_t sid, int32_t cap, int32_t* ids, int32_t* n, int32_t* parEnd) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (!n || !parEnd || cap < 0 || (cap > 0 && !ids)) throw WdmError(WDM_E_ARG, "output arrays");
        int32_t k = 0;
        for (int32_t q : S.freeIds) { if (k < cap) ids[k] = q; k++; }
        *n = k;
        *parEnd = S.parEnd;
    });
}
int wdm_set_free_ids(wdm_ctx* c, int32_t sid, int32_t n, const int32_t* ids, int32_t parEnd) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (n < 0 || (n > 0 && !ids) || parEnd < 0 || parEnd > S.np) throw WdmError(WDM_E_ARG, "free ids");
        std::vector<char> hole(S.np, 0);
        for (int p = 0; p < parEnd; p++) hole[p] = (S.ptypeH[p] & 15) == WDM_SHAPE_NONE;
        for (int p = parEnd; p < S.np; p++)
            if ((S.ptypeH[p] & 15) != WDM_SHAPE_NONE) throw WdmError(WDM_E_ARG, "a particle in use lies past par_end");
        int32_t cnt = 0;
        for (int k = 0; k < n; k++) {
            if (ids[k] < 0 || ids[k] >= parEnd || hole[ids[k]] != 1) throw WdmError(WDM_E_ARG, "free ids must list each hole below par_end once");
            hole[ids[k]] = 2;
            cnt++;
        }
        for (int p = 0; p < parEnd; p++)
            if (hole[p] == 1) throw WdmError(WDM_E_ARG, "free ids must list each hole below par_end once");
        S.freeIds.assign(ids, ids + n);
        S.parEnd = parEnd;
        (void)cnt;
    });
}
int wdm_set_masks(wdm_ctx* c, int32_t sid, int32_t n, const int32_t* ids, const uint32_t* masks) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (n < 0 || (n > 0 && (!ids || !masks))) throw WdmError(WDM_E_ARG, "mask arrays");
        for (int k = 0; k < n; k++)
            if (ids[k] < 0 || ids[k] >= S.np || (S.ptypeH[ids[k]] & 15) == WDM_SHAPE_NONE) throw WdmError(WDM_E_ARG, "particle id");
        if (n == 0) return;
        Blob b;
        const size_t oi = b.add(ids, n * sizeof(int32_t)), om = b.add(masks, n * sizeof(uint32_t));
        const uint8_t* base = c->sh->stDyn.upload(c, b.b.data(), b.b.size());
        hipLaunchKernelGGL(k_set_masks, dim3((n + 255) / 256), dim3(256), 0, c->stream, S.pmask.p, (const int32_t*)(base + oi),
                           (const uint32_t*)(base + om), n);
        checkLaunch();
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}
int wdm_select_box(wdm_ctx* c, int32_t sid, const double* lo3, const double* hi3, int32_t inside, uint32_t mask, uint32_t markMask,
                   int32_t cap, int32_t* ids, int32_t* nOut) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (!lo3 || !hi3 || !nOut || cap < 0 || (cap > 0 && !ids)) throw WdmError(WDM_E_ARG, "box arrays");
        ensure(c, c->sh->temp, (size_t)std::max(S.np, 1));
        uint8_t* flag = c->sh->temp.p;
        if (S.np > 0)
            hipLaunchKernelGGL(k_select_box, dim3(nblk(S.np)), dim3(WDM_BLOCK), 0, c->stream, (const double4*)S.posr.p,
                               (const int32_t*)S.ptype.p, (const uint32_t*)S.pmask.p, S.np, make_double3(lo3[0], lo3[1], lo3[2]),
                               make_double3(hi3[0], hi3[1], hi3[2]), inside, mask, markMask, flag);
        checkLaunch();
        std::vector<uint8_t> h(S.np);
        if (S.np > 0) HIPCHK(hipMemcpyAsync(h.data(), flag, S.np, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        int32_t k = 0;
        for (int i = 0; i < S.np; i++)
            if (h[i]) { if (k < cap) ids[k] = i; k++; }
        *nOut = k;
    });
}
// ---------------- SpherePack::makeCloud on the device ----------------
// The set-up of SpherePack::makeCloud (lib/sphere-pack/SpherePack.cpp:135-187): radius mode,
// mean radius from the porosity, PSD radii with the mass-weighted cumulative curve and its
// scaling, and the deterministic radius sequence for num > 0 (:191-209); placement runs in
// k_make_cloud; a sphere without a free spot makes the reference retry with the porosity
// raised by a tenth of what is left (:236-241), the generator going on where it stopped.
int wdm_make_cloud(wdm_ctx* c, const double* mn3, const double* mx3, double rMean, double rRelFuzz, int32_t num, int32_t periodic,
                   double porosity, const double* psdSizes, const double* psdCumm, int32_t nPsd, int32_t distributeMass,
                   uint32_t* rngState, int64_t cap, double* c3, double* rOut, int64_t* nOut, double* appliedPsdScaling) {
    return guard(c, [&] {
        if (!mn3 || !mx3 || !rngState || !nOut || cap <= 0 || !c3 || !rOut || (nPsd > 0 && (!psdSizes || !psdCumm)))
            throw WdmError(WDM_E_ARG, "makeCloud arguments");
        if (*rngState == 0 || *rngState >= 2147483647u) throw WdmError(WDM_E_ARG, "generator state must be in 1..2^31-2");
        if (num > 0 && num > cap) throw WdmError(WDM_E_ARG, "cap below num");
        if (cap >= INT32_MAX) throw WdmError(WDM_E_ARG, "cap too large");
        const double PI = 3.14159265358979323846;
        double size[3], mn[3];
        for (int d = 0; d < 3; d++) { mn[d] = mn3[d]; size[d] = mx3[d] - mn3[d]; }
        const double volume = size[0] * (size[1] * size[2]);  // det of diag(size), Eigen's cofactor order
        if (!volume) throw WdmError(WDM_E_ARG, "The box defined as null volume. Define at least maxCorner of the box, or hSize if periodic.");
        if (porosity <= 0) porosity = .5;
        uint32_t state = *rngState;
        double scaling = 1.;
        int64_t n = 0;
        std::vector<double4> sphH;
        for (int attempt = 0;; attempt++) {
            if (attempt > 200) throw WdmError(WDM_E_STATE, "makeCloud: no packing after 200 porosity increases");
            int mode = -1;  // 0 rMean, 1 num, 2 PSD
            bool err = false;
            double rM = attempt == 0 ? rMean : -1.;
            if (rM > 0) mode = 0;
            else if (num > 0 && nPsd == 0) {
                mode = 1;
                rM = std::pow(volume * (1 - porosity) / (PI * (4 / 3.) * (1 + rRelFuzz * rRelFuzz) * num), 1 / 3.);
            }
            std::vector<double> psdR, psdC2;
            if (nPsd > 0) {
                err = mode >= 0;
                mode = 2;
                if (nPsd <= 1) throw WdmError(WDM_E_ARG, "SpherePack.makeCloud: psdSizes must have at least 2 items");
                if (psdCumm[0] != 0. && psdCumm[nPsd - 1] != 1.) throw WdmError(WDM_E_ARG, "SpherePack.makeCloud: first and last items of psdCumm *must* be exactly 0 and 1.");
                for (int i = 0; i < nPsd; i++) {
                    psdR.push_back(.5 * psdSizes[i]);
                    if (distributeMass) {
                        if (i == 0) psdC2.push_back(0);
                        else psdC2.push_back(psdC2[i - 1] + 3.0 * volume * (1 - porosity) / PI * (psdCumm[i] - psdCumm[i - 1]) / (psdSizes[i] - psdSizes[i - 1]) *
                                                                  (std::pow(psdSizes[i - 1], -2) - std::pow(psdSizes[i], -2)));
                    }
                    if (i > 0 && (psdSizes[i - 1] > psdSizes[i] || psdCumm[i - 1] > psdCumm[i]))
                        throw WdmError(WDM_E_ARG, "SpherePack:makeCloud: psdSizes and psdCumm must be both non-decreasing.");
                }
                if (distributeMass) {
                    if (num > 0 && psdC2[nPsd - 1] < num) {
                        scaling = std::pow(psdC2[nPsd - 1] / num, 1. / 3.);
                        for (int i = 0; i < nPsd; i++) psdR[i] *= scaling;
                    }
                    for (int i = 1; i < nPsd; i++) psdC2[i] /= psdC2[nPsd - 1];
                }
            }
            if (err || mode < 0)
                throw WdmError(WDM_E_ARG, "SpherePack.makeCloud: at least one of rMean, porosity, psdSizes & psdCumm arguments must be specified. rMean can't be combined with psdSizes.");
            const double* cum = distributeMass ? psdC2.data() : psdCumm;
            auto radius = [&](double rnd) -> double {
                if (mode != 2) return distributeMass ? pow3Interp(rnd, rM * (1 - rRelFuzz), rM * (1 + rRelFuzz)) : rM * (2 * (rnd - .5) * rRelFuzz + 1);
                double norm;
                const int piece = psdGetPiece(rnd, cum, nPsd, norm);
                return distributeMass ? pow3Interp(norm, psdR[piece], psdR[piece + 1]) : psdR[piece] + norm * (psdR[piece + 1] - psdR[piece]);
            };
            std::vector<double> radii;
            double rmax = 0;
            if (num > 0) {
                radii.resize(num);
                for (int i = 0; i < num; i++) { radii[i] = radius(((double)num - (double)i + 0.5) / ((double)num + 1.)); rmax = std::max(rmax, radii[i]); }
            } else if (mode != 2) rmax = rM * (1 + std::fabs(rRelFuzz));
            else for (double q : psdR) rmax = std::max(rmax, q);
            if (!(rmax > 0)) throw WdmError(WDM_E_ARG, "makeCloud: radii must be positive");
            CloudArgs A;
            std::memset(&A, 0, sizeof(A));
            int64_t cells = 1;
            for (int d = 0; d < 3; d++) {
                A.mn[d] = mn[d];
                A.size[d] = size[d];
                A.nc[d] = std::max(1, (int)std::floor(size[d] / (2 * rmax)));
            }
            for (;;) {
                cells = (int64_t)A.nc[0] * A.nc[1] * A.nc[2];
                if (cells <= (1 << 25)) break;
                for (int d = 0; d < 3; d++) A.nc[d] = std::max(1, A.nc[d] / 2);
            }
            for (int d = 0; d < 3; d++) A.h[d] = size[d] / A.nc[d];
            A.periodic = periodic ? 1 : 0;
            A.num = num;
            A.mode = mode == 2 ? 1 : 0;
            A.distributeMass = distributeMass ? 1 : 0;
            A.nPsd = nPsd;
            A.rMean = rM;
            A.rRelFuzz = rRelFuzz;
            A.cap = num > 0 ? num : cap;
            A.maxTry = 1000;
            A.m2lo = std::pow(rM * (1 - rRelFuzz), -2.);
            A.m2hi = std::pow(rM * (1 + rRelFuzz), -2.);
            std::vector<double> psdRm2(psdR.size());
            for (size_t q = 0; q < psdR.size(); q++) psdRm2[q] = std::pow(psdR[q], -2.);
            Blob b;
            const size_t oR = b.addv(radii), oPR = b.addv(psdR), oPM = b.addv(psdRm2);
            std::vector<double> cumv(cum ? cum : psdR.data(), (cum ? cum : psdR.data()) + (nPsd > 0 ? nPsd : 0));
            const size_t oPC = b.addv(cumv);
            const size_t oS = b.add(&state, sizeof(state));
            DBuf<uint8_t> tb, cellsB, nextB, sphB, outB;
            ensure(c, tb, std::max<size_t>(b.b.size(), 64));
            ensure(c, cellsB, (size_t)cells * 4);
            ensure(c, nextB, (size_t)A.cap * 4);
            ensure(c, sphB, (size_t)A.cap * sizeof(double4));
            ensure(c, outB, 16);
            hipStream_t st = c->stream;
            HIPCHK(hipMemcpyAsync(tb.p, b.b.data(), b.b.size(), hipMemcpyHostToDevice, st));
            A.radii = (const double*)(tb.p + oR);
            A.psdR = (const double*)(tb.p + oPR);
            A.psdRm2 = (const double*)(tb.p + oPM);
            A.psdC = (const double*)(tb.p + oPC);
            hipLaunchKernelGGL(k_fill_i32, dim3(nblk(cells)), dim3(WDM_BLOCK), 0, st, (int32_t*)cellsB.p, cells, -1);
            hipLaunchKernelGGL(k_make_cloud, dim3(1), dim3(1024), 0, st, A, (uint32_t*)(tb.p + oS), (int32_t*)cellsB.p, (int32_t*)nextB.p,
                               (double4*)sphB.p, (int64_t*)outB.p);
            checkLaunch();
            int64_t out[2];
            HIPCHK(hipMemcpyAsync(out, outB.p, sizeof(out), hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(&state, tb.p + oS, sizeof(state), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            n = out[0];
            const bool retry = out[1] && num > 0 && mode != 0;
            if (!retry && n > 0) {
                sphH.resize(n);
                HIPCHK(hipMemcpy(sphH.data(), sphB.p, n * sizeof(double4), hipMemcpyDeviceToHost));
            }
            release(c, tb); release(c, cellsB); release(c, nextB); release(c, sphB); release(c, outB);
            if (!retry) break;
            porosity = porosity + (1 - porosity) / 10.;
        }
        if (n > cap) throw WdmError(WDM_E_ARG, "cap below the spheres placed");
        for (int64_t i = 0; i < n; i++) {
            c3[3 * i] = sphH[i].x; c3[3 * i + 1] = sphH[i].y; c3[3 * i + 2] = sphH[i].z;
            rOut[i] = sphH[i].w;
        }
        *nOut = n;
        *rngState = state;
        if (appliedPsdScaling) *appliedPsdScaling = scaling;
    });
}
// pack.regularOrtho / pack.regularHexa (py/pack.py:323-350) for the predicates inAlignedBox,
// inSphere and inCylinder: the predicate's aabb (py/_packPredicates.cpp), the lattice over it
// as the reference lays it out, the points where the predicate holds (pad = radius) in the
// reference's order.
int wdm_regular_pack(wdm_ctx* c, int32_t hexa, int32_t kind, const double* params, double radius, double gap, int64_t cap, double* c3,
                     int64_t* nOut) {
    return guard(c, [&] {
        if (!params || !nOut || cap < 0) throw WdmError(WDM_E_ARG, "regular packing arguments");
        if (kind < 0 || kind > 2) throw WdmError(WDM_E_ARG, "predicate kind must be 0 (inAlignedBox), 1 (inSphere) or 2 (inCylinder)");
        LatticeArgs L;
        std::memset(&L, 0, sizeof(L));
        L.hexa = hexa ? 1 : 0;
        L.kind = kind;
        const int np = kind == 0 ? 6 : (kind == 1 ? 4 : 7);
        for (int q = 0; q < np; q++) L.p[q] = params[q];
        double mn[3], mx[3];
        if (kind == 0) {
            for (int d = 0; d < 3; d++) { mn[d] = params[d]; mx[d] = params[3 + d]; }
        } else if (kind == 1) {
            for (int d = 0; d < 3; d++) { mn[d] = params[d] - params[3]; mx[d] = params[d] + params[3]; }
        } else {  // inCylinder::aabb
            const double* A = params;
            const double* B = params + 3;
            const double cx = B[0] - A[0], cy = B[1] - A[1], cz = B[2] - A[2];
            const double ht = std::sqrt((cx * cx + cy * cy) + cz * cz);
            if (!(ht > 0)) throw WdmError(WDM_E_ARG, "inCylinder: the axis has zero length");
            const double k[3] = {std::sqrt((A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2])) / ht,
                                 std::sqrt((A[0] - B[0]) * (A[0] - B[0]) + (A[2] - B[2]) * (A[2] - B[2])) / ht,
                                 std::sqrt((A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1])) / ht};
            for (int d = 0; d < 3; d++) {
                mn[d] = std::min(A[d], B[d]) - params[6] * k[d];
                mx[d] = std::max(A[d], B[d]) + params[6] * k[d];
            }
        }
        for (int d = 0; d < 3; d++)
            if (!std::isfinite(mn[d]) || !std::isfinite(mx[d])) throw WdmError(WDM_E_ARG, "Aabb of the predicate must not be infinite");
        L.radius = radius;
        int64_t total = 1;
        if (!L.hexa) {
            const double step = 2 * radius + gap;
            for (int d = 0; d < 3; d++) {
                L.start[d] = mn[d] + radius;
                L.a1[d] = L.start[d] + step;
                const double len = std::ceil(((mx[d] - radius) - L.start[d]) / step);  // numpy.arange's length
                L.n[d] = (len > 0 && std::isfinite(len)) ? (int64_t)len : 0;
            }
        } else {
            L.a = 2 * radius + gap;
            L.hy = L.a * std::sqrt(3.) / 2.;
            L.hz = L.a * std::sqrt(6.) / 3.;
            const double h[3] = {L.a, L.hy, L.hz};
            for (int d = 0; d < 3; d++) {
                L.mn[d] = mn[d];
                const double q = (mx[d] - mn[d]) / h[d];
                if (!(q >= 0) || !std::isfinite(q)) throw WdmError(WDM_E_ARG, "regularHexa: lattice spacing");
                L.n[d] = (int64_t)q + 1;  // range(0, int(dim/h) + 1)
            }
        }
        for (int d = 0; d < 3; d++) {
            if (L.n[d] > ((int64_t)1 << 31)) throw WdmError(WDM_E_ARG, "lattice too large");
            total *= L.n[d];
            if (total > INT32_MAX) throw WdmError(WDM_E_ARG, "lattice too large (more than 2^31-1 points)");
        }
        *nOut = 0;
        if (total == 0) return;
        hipStream_t st = c->stream;
        DBuf<uint8_t> flags;
        DBuf<int64_t> sel, nsel;
        DBuf<double> outB;
        ensure(c, flags, (size_t)total);
        ensure(c, sel, (size_t)total);
        ensure(c, nsel, 1);
        hipLaunchKernelGGL(k_lattice_flags, dim3(nblk(total)), dim3(WDM_BLOCK), 0, st, L, total, flags.p);
        checkLaunch();
        hipcub::CountingInputIterator<int64_t> it(0);
        size_t bytes = 0;
        HIPCHK(hipcub::DeviceSelect::Flagged(nullptr, bytes, it, flags.p, sel.p, nsel.p, (int)total, st));
        void* t = tempFor(c, bytes);
        HIPCHK(hipcub::DeviceSelect::Flagged(t, bytes, it, flags.p, sel.p, nsel.p, (int)total, st));
        int64_t n = 0;
        HIPCHK(hipMemcpyAsync(&n, nsel.p, sizeof(n), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (!c3) {  // count only
            release(c, flags); release(c, sel); release(c, nsel);
            *nOut = n;
            return;
        }
        if (n > cap) {
            release(c, flags); release(c, sel); release(c, nsel);
            *nOut = n;
            throw WdmError(WDM_E_ARG, "cap below the spheres of the lattice");
        }
        if (n > 0) {
            ensure(c, outB, (size_t)n * 3);
            hipLaunchKernelGGL(k_lattice_write, dim3(nblk(n)), dim3(WDM_BLOCK), 0, st, L, (const int64_t*)sel.p, (const int64_t*)nsel.p, outB.p);
            checkLaunch();
            HIPCHK(hipMemcpyAsync(c3, outB.p, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
        }
        release(c, flags); release(c, sel); release(c, nsel); release(c, outB);
        *nOut = n;
    });
}
int wdm_set_energy(wdm_ctx* c, int32_t sid, const double* values, const int32_t* created) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (!values || !created) throw WdmError(WDM_E_ARG, "energy arrays");
        uint32_t cre = 0;
        for (int k = 0; k < WDM_E_COUNT; k++) cre |= created[k] ? (1u << k) : 0u;
        HIPCHK(hipMemcpyAsync(S.E.p, values, WDM_E_COUNT * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->sh->words.p + 4 * (size_t)S.slot + 1, &cre, sizeof(cre), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}
int wdm_set_time(wdm_ctx* c, int32_t sid, int64_t step, double time) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (step < 0) throw WdmError(WDM_E_ARG, "negative step");
        S.step = step;
        S.time = time;
        const uint32_t z = 0;  // Verlet trigger: stamps of the old step count are meaningless now
        HIPCHK(hipMemcpyAsync(c->sh->words.p + 4 * (size_t)S.slot, &z, sizeof(z), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}
int wdm_get_energy(wdm_ctx* c, int32_t sid, double* values, int32_t* created) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        double e[WDM_E_COUNT];
        uint32_t cre = 0;
        HIPCHK(hipMemcpyAsync(e, S.E.p, sizeof(e), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(&cre, c->sh->words.p + 4 * (size_t)S.slot + 1, sizeof(cre), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int k = 0; k < WDM_E_COUNT; k++) {
            if (values) values[k] = e[k];
            if (created) created[k] = (cre >> k) & 1u;
        }
    });
}

int wdm_get_stats(wdm_ctx* c, int32_t sid, int64_t* step, int64_t* nfr, double* vd) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (step) *step = S.step;
        if (nfr) *nfr = S.nFullRuns;
        if (vd) *vd = S.verletResolved ? S.verletDist : NAN;
    });
}

int wdm_get_cell(wdm_ctx* c, int32_t sid, double* hSize9, double* gradV9) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
            if (hSize9) hSize9[3 * i + j] = S.hSize.m[i][j];
            if (gradV9) gradV9[3 * i + j] = S.gradV.m[i][j];
        }
    });
}

int wdm_sync(wdm_ctx* c) { return guard(c, [&] { HIPCHK(hipStreamSynchronize(c->stream)); }); }

int wdm_set_tuning(wdm_ctx* c, int32_t window, int32_t graphs, double cellFactor) {
    return guard(c, [&] {
        if (window < 1 || window > 4096) throw WdmError(WDM_E_ARG, "window must be in [1, 4096] steps");
        if (!(cellFactor >= 1. && cellFactor <= 16.)) throw WdmError(WDM_E_ARG, "cell_factor must be in [1, 16]");
        c->window = window;
        c->graphs = graphs != 0;
        c->cellFactor = cellFactor;
    });
}
int wdm_set_broadphase(wdm_ctx* c, int32_t mode) {
    return guard(c, [&] {
        if (mode != 0 && mode != 1) throw WdmError(WDM_E_ARG, "broadphase must be 0 (grid) or 1 (sort and sweep)");
        c->broadphase = mode;
    });
}
int wdm_get_broadphase(wdm_ctx* c, int32_t* mode) {
    return guard(c, [&] {
        if (mode) *mode = c->broadphase;
    });
}
int wdm_set_relayout(wdm_ctx* c, int32_t steps, int64_t min_contacts) {
    return guard(c, [&] {
        if (steps < 0 || min_contacts < 0) throw WdmError(WDM_E_ARG, "relayout steps and min_contacts must be >= 0 (steps 0: never)");
        c->relayout = steps;
        c->relayoutMin = min_contacts;
    });
}
int wdm_get_relayout(wdm_ctx* c, int32_t sid, int32_t* steps, int64_t* min_contacts, int64_t* relayouts) {
    return guard(c, [&] {
        if (steps) *steps = c->relayout;
        if (min_contacts) *min_contacts = c->relayoutMin;
        if (relayouts) *relayouts = sid >= 0 ? getScene(c, sid).relayouts : 0;
    });
}
int wdm_get_tuning(wdm_ctx* c, int32_t* window, int32_t* graphs, double* cellFactor, int64_t* windows, int64_t* captures,
                   int64_t* launches) {
    return guard(c, [&] {
        if (window) *window = c->window;
        if (graphs) *graphs = c->graphs;
        if (cellFactor) *cellFactor = c->cellFactor;
        if (windows) *windows = c->sh->windows;
        if (captures) *captures = c->sh->graphCaptures;
        if (launches) *launches = c->sh->graphLaunches;
    });
}
int wdm_enable_timing(wdm_ctx* c, int32_t on) {
    return guard(c, [&] {
        timingFlush(c);
        c->timing.on = on != 0;
        c->timing.acc.clear();
    });
}

int wdm_kernel_times(wdm_ctx* c, int32_t cap, char* names, double* ms, int64_t* launches, int32_t* nOut) {
    return guard(c, [&] {
        timingFlush(c);
        int k = 0;
        for (auto& kv : c->timing.acc) {
            if (k >= cap) break;
            if (names) { std::memset(names + 64 * k, 0, 64); std::strncpy(names + 64 * k, kv.first.c_str(), 63); }
            if (ms) ms[k] = kv.second.first;
            if (launches) launches[k] = kv.second.second;
            k++;
        }
        if (nOut) *nOut = k;
    });
}

int wdm_relax_overlaps(wdm_ctx* c, int32_t sid, int32_t iters) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (!(S.nodesSet && S.partsSet)) throw WdmError(WDM_E_STATE, "upload nodes and particles first");
        relax(S, iters);
    });
}

}  // extern "C"
