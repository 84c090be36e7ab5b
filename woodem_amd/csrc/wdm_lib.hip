// libwoodem_hip: host side of the MI355X DEM timestep.
//
// C++ engine classes named after the reference's plugins (Leapfrog,
// InsertionSortCollider, ContactLoop, Scene::doOneStep) drive the kernels of
// wdm_kernels.h on one HIP stream per context; the extern "C" functions at the
// bottom are the drop-in boundary declared in include/woodem_hip.h.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/woodem_hip.h"
#include "wdm_kernels.h"

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
struct Ctx;
struct DevMem {
    Ctx* ctx = nullptr;
    std::vector<std::pair<void*, size_t>> owned;
    size_t bytes = 0;
};

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
};

struct Timing {
    bool on = false;
    std::map<std::string, std::pair<double, int64_t>> acc;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t budget = 0, used = 0;
    std::string err;
    std::map<int32_t, struct Scene*> scenes;
    int32_t nextId = 0;
    Timing timing;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
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
    size_t nn = std::max<size_t>(n, b.n + b.n / 2);
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
    DBuf<int32_t> cd, slot, counter;
    DBuf<T> st;
    DBuf<double4> gfc;
    int64_t cap = 0;
    void reserve(Ctx* c, int64_t n) {
        if (n <= cap && key.p) return;
        int64_t nc = std::max<int64_t>(n + n / 3, 1024);
        release(c, key); release(c, ids); release(c, info); release(c, cd); release(c, st); release(c, gfc);
        release(c, slot); release(c, counter);
        ensure(c, key, nc); ensure(c, ids, nc); ensure(c, info, nc); ensure(c, cd, nc);
        ensure(c, slot, nc); ensure(c, counter, 1);
        ensure(c, st, (size_t)NSTATE * nc); ensure(c, gfc, 2 * (size_t)nc);
        cap = nc;
    }
    ConView<T> view() const {
        ConView<T> v; v.key = key.p; v.ids = ids.p; v.info = info.p; v.cd = cd.p; v.slot = slot.p; v.counter = counter.p;
        v.st = st.p; v.gfc = gfc.p; v.cap = cap; return v;
    }
    void free(Ctx* c) {
        release(c, key); release(c, ids); release(c, info); release(c, cd); release(c, st); release(c, gfc);
        release(c, slot); release(c, counter); cap = 0;
    }
};

// ---------------- Scene (core/Scene.hpp:27) with its DemField (pkg/dem/Particle.hpp:278) ----------------
struct Scene {
    Ctx* ctx;
    int32_t np, nn, nmat;
    wdm_params P{};
    bool paramsSet = false, nodesSet = false, partsSet = false, matsSet = false;
    // host copies of static particle data
    std::vector<int32_t> ptypeH;
    std::vector<double> radiusH;
    std::vector<int32_t> binnedH, unbH;
    // cell (core/Cell.hpp)
    bool periodic = false;
    HM3 hSize = heye(), gradV = hzero(), nextGradV = hzero(), pprevHsize = heye(), trsf = heye();
    double size[3] = {1, 1, 1}, spinVec[3] = {0, 0, 0};
    int homoDeform = 4;
    // counters
    int64_t step = 0, nFullRuns = 0;
    double verletDist = 0;
    bool verletResolved = false;
    bool needFull = true;
    bool lastLeapfrogVerletChecked = false;
    bool forcesFromContacts = false;  // node forces are the sums of the last ContactLoop (not materialised)
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
    // contacts, double-buffered
    ConBuf<double> cb64[2];
    ConBuf<float> cb32[2];
    int cur = 0;
    int64_t nc = 0;
    // broadphase scratch
    DBuf<uint32_t> ckeys, ckeys2;
    DBuf<int32_t> cvals, cvals2, cstart, cend;
    DBuf<int64_t> counts, offs, ooFlag64, ooRank;
    DBuf<int32_t> ooFlag, cnt32, oslot;
    DBuf<int64_t> rflag, rscan;
    DBuf<uint64_t> candK, candK2;
    DBuf<int32_t> candP, candP2;
    DBuf<double> extPart;
    DBuf<uint8_t> temp;
    // CSR
    DBuf<uint64_t> entK, entK2;
    DBuf<uint32_t> entV, entV2;
    DBuf<int64_t> csrOff;
    DBuf<uint64_t> realMask;
    // energy, flags
    DBuf<double> E, epartL, epartC;
    DBuf<int32_t> eslots;
    DBuf<uint32_t> orL, ecre, err;
    uint32_t* hostWords = nullptr;  // pinned: [0] leapfrog or-bits, [1] error bits, [2] contact or-bits
    uint32_t createdBits = 0;
    // relax scratch
    DBuf<double4> disp;

    Scene(Ctx* c, int32_t np_, int32_t nn_, int32_t nm_) : ctx(c), np(np_), nn(nn_), nmat(nm_) {}

    bool mixed() const { return P.precision == WDM_PREC_MIXED; }

    NodeView nview() {
        NodeView v; v.posr = posr.p; v.velm = velm.p; v.angI = angI.p; v.ori = ori.p; v.force = force.p; v.torque = torque.p;
        v.lastPos = lastPos.p; v.flags = flags.p; v.checked = checked.p; v.nn = nn; return v;
    }
    GatherView gview() {
        GatherView g;
        g.off = csrOff.p; g.ent = entV2.p; g.realMask = realMask.p;
        g.gfc = (P.precision == WDM_PREC_MIXED) ? cb32[cur].gfc.p : cb64[cur].gfc.p;
        return g;
    }
    PartView pview() {
        PartView v; v.ptype = ptype.p; v.mask = pmask.p; v.mat = pmat.p; v.fnodes = fnodes.p; v.bmin = bmin.p; v.bmax = bmax.p; v.np = np; return v;
    }

    void freeAll() {
        release(ctx, posr); release(ctx, velm); release(ctx, angI); release(ctx, ori); release(ctx, force); release(ctx, torque);
        release(ctx, lastPos); release(ctx, flags); release(ctx, checked); release(ctx, ptype); release(ctx, pmat); release(ctx, binned);
        release(ctx, unb); release(ctx, pmask); release(ctx, fnodes); release(ctx, bmin); release(ctx, bmax); release(ctx, mat);
        for (int i = 0; i < 2; i++) { cb64[i].free(ctx); cb32[i].free(ctx); }
        release(ctx, ckeys); release(ctx, ckeys2); release(ctx, cvals); release(ctx, cvals2); release(ctx, cstart); release(ctx, cend);
        release(ctx, counts); release(ctx, offs); release(ctx, ooFlag64); release(ctx, ooRank); release(ctx, ooFlag); release(ctx, cnt32); release(ctx, oslot); release(ctx, rflag); release(ctx, rscan);
        release(ctx, candK); release(ctx, candK2); release(ctx, candP); release(ctx, candP2); release(ctx, extPart); release(ctx, temp);
        release(ctx, entK); release(ctx, entK2); release(ctx, entV); release(ctx, entV2); release(ctx, csrOff); release(ctx, realMask);
        release(ctx, E); release(ctx, epartL); release(ctx, epartC); release(ctx, eslots); release(ctx, orL); release(ctx, ecre); release(ctx, err);
        release(ctx, disp);
        if (hostWords) { (void)hipHostFree(hostWords); hostWords = nullptr; }
    }
};

struct KTimer {  // HIP-event timing of one launch on the context stream
    Ctx* c; const char* name; bool on;
    KTimer(Ctx* c_, const char* n) : c(c_), name(n), on(c_->timing.on) { if (on) (void)hipEventRecord(c->ev0, c->stream); }
    ~KTimer() {
        if (!on) return;
        (void)hipEventRecord(c->ev1, c->stream);
        (void)hipEventSynchronize(c->ev1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
        auto& a = c->timing.acc[name];
        a.first += ms; a.second += 1;
    }
};

void checkLaunch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw WdmError(WDM_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
}

// temp storage for hipcub calls
void* tempFor(Scene& S, size_t bytes) { ensure(S.ctx, S.temp, bytes); return S.temp.p; }

template <typename K, typename V>
void sortPairs(Scene& S, K* kin, K* kout, V* vin, V* vout, int64_t n, int endBit) {
    if (n <= 0) return;
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, (int)n, 0, endBit, S.ctx->stream));
    void* t = tempFor(S, bytes);
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(t, bytes, kin, kout, vin, vout, (int)n, 0, endBit, S.ctx->stream));
}
// out[0] = 0, out[i+1] = sum(in[0..i])
void scanInclusiveShift(Scene& S, const int64_t* in, int64_t* out, int64_t n) {
    HIPCHK(hipMemsetAsync(out, 0, sizeof(int64_t), S.ctx->stream));
    if (n <= 0) return;
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out + 1, (int)n, S.ctx->stream));
    void* t = tempFor(S, bytes);
    HIPCHK(hipcub::DeviceScan::InclusiveSum(t, bytes, in, out + 1, (int)n, S.ctx->stream));
}
int64_t readI64(Scene& S, const int64_t* p) {
    int64_t v = 0;
    HIPCHK(hipMemcpyAsync(&v, p, sizeof(v), hipMemcpyDeviceToHost, S.ctx->stream));
    HIPCHK(hipStreamSynchronize(S.ctx->stream));
    return v;
}

__global__ void k_i64_to_i32_one(const int64_t* a, int32_t* b) { *b = (int32_t)*a; }
__global__ void k_i32_to_i64(const int32_t* a, int64_t* b, int64_t n) {
    int64_t i = blockIdx.x * (int64_t)WDM_BLOCK + threadIdx.x;
    if (i < n) b[i] = a[i];
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

StepArgs makeArgs(Scene& S) {
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
    const double dt = P.dt;
    HM3 pprevL = S.gradV, nnextL = S.nextGradV;
    HM3 ImL = hzero(), IpL = hzero(), LmL = hzero();
    double dsv[3] = {0, 0, 0};
    if (a.homoDeform == 4) {
        ImL = hinv(hsub(heye(), hdivs(hscale(hadd(nnextL, pprevL), dt), 4.)));
        IpL = hadd(heye(), hdivs(hscale(hadd(nnextL, pprevL), dt), 4.));
        LmL = hsub(nnextL, pprevL);
        double l1[3], l2[3];
        leviCivitaH(hscale(hsub(pprevL, htrans(pprevL)), .5), l1);
        leviCivitaH(hscale(hsub(nnextL, htrans(nnextL)), .5), l2);
        for (int k = 0; k < 3; k++) dsv[k] = l1[k] * -.5 + l2[k] * .5;
    }
    a.ImLL4hInv = toM3(ImL); a.IpLL4h = toM3(IpL); a.LmL = toM3(LmL);
    a.dGradV = toM3(hsub(nnextL, pprevL));
    a.midGradV = toM3(hscale(hadd(pprevL, nnextL), .5));
    a.nextGradV = toM3(nnextL); a.gradV = toM3(pprevL);
    for (int k = 0; k < 3; k++) { a.deltaSpinVec[k] = dsv[k]; a.spinVec[k] = S.spinVec[k]; a.cellSize[k] = S.size[k]; }
    a.hSize = toM3(S.hSize);
    a.gatherHSize = toM3(S.hSizeCL);
    if (S.periodic && S.homoDeform == 4) a.shiftVelM = toM3(hmul(S.gradV, S.pprevHsize));
    else if (S.periodic && (S.homoDeform == 2 || S.homoDeform == 3)) a.shiftVelM = toM3(hmul(S.gradV, S.hSize));
    else a.shiftVelM = toM3(hzero());
    a.verletD2 = S.verletDist * S.verletDist;
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
    a.mat = S.mat.p;
    return a;
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

void finalizeEnergy(Scene& S, DBuf<double>& part, int nblocks, int nv, int slotOffset) {
    hipLaunchKernelGGL(k_energy_finalize, dim3(1), dim3(64), 0, S.ctx->stream, part.p, nblocks, nv, S.eslots.p + slotOffset, S.E.p);
    checkLaunch();
}

// ---------------- Leapfrog::run (pkg/dem/Leapfrog.cpp:103-255) ----------------
struct Leapfrog {
    static void run(Scene& S) {
        Ctx* c = S.ctx;
        resolveVerlet(S);
        StepArgs a = makeArgs(S);
        a.checkVerlet = (!S.needFull && S.verletDist > 0) ? 1 : 0;
        S.lastLeapfrogVerletChecked = a.checkVerlet;
        int nb = (int)nblk(S.nn);
        ensure(c, S.epartL, (size_t)nb * 3 + 3);
        HIPCHK(hipMemsetAsync(S.orL.p, 0, sizeof(uint32_t), c->stream));
        {
            KTimer t(c, "leapfrog");
            if (S.forcesFromContacts)
                hipLaunchKernelGGL(k_leapfrog<true>, dim3(nb), dim3(WDM_BLOCK), 0, c->stream, a, S.nview(), S.pview(), S.gview(), S.epartL.p, S.orL.p, S.ecre.p);
            else
                hipLaunchKernelGGL(k_leapfrog<false>, dim3(nb), dim3(WDM_BLOCK), 0, c->stream, a, S.nview(), S.pview(), S.gview(), S.epartL.p, S.orL.p, S.ecre.p);
            checkLaunch();
        }
        if (a.reallyTrackEnergy) finalizeEnergy(S, S.epartL, nb, 3, 0);
    }
};

// ---------------- InsertionSortCollider::run (pkg/dem/InsertionSortCollider.cpp:445-646) ----------------
struct InsertionSortCollider {
    static void buildGrid(Scene& S, GridArgs& g, int& nbinned) {
        Ctx* c = S.ctx;
        int nb = (int)S.binnedH.size();
        nbinned = nb;
        std::memset(&g, 0, sizeof(g));
        g.periodic = S.periodic;
        for (int d = 0; d < 3; d++) { g.n[d] = 1; g.w[d] = 1; g.L[d] = S.size[d]; g.invL[d] = 1. / S.size[d]; }
        if (nb == 0) return;
        int nbk = (int)nblk(nb);
        ensure(c, S.extPart, (size_t)nbk * 7);
        hipLaunchKernelGGL(k_extent, dim3(nbk), dim3(WDM_BLOCK), 0, c->stream, S.pview(), (const int32_t*)S.binned.p, nb, S.extPart.p);
        checkLaunch();
        std::vector<double> h((size_t)nbk * 7);
        HIPCHK(hipMemcpyAsync(h.data(), S.extPart.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY}, ext = 0;
        for (int b = 0; b < nbk; b++) {
            for (int d = 0; d < 3; d++) { mn[d] = std::min(mn[d], h[7 * b + d]); mx[d] = std::max(mx[d], h[7 * b + 3 + d]); }
            ext = std::max(ext, h[7 * b + 6]);
        }
        if (!(ext > 0) || !std::isfinite(ext)) ext = 1e-12;
        double w = ext * 1.001;
        const int64_t maxCells = std::max<int64_t>(4096, 2 * (int64_t)nb);
        for (int it = 0; it < 200; it++) {
            int64_t tot = 1;
            for (int d = 0; d < 3; d++) {
                if (S.periodic) {
                    int n = (int)std::floor(S.size[d] / w);
                    if (n < 3) n = 1;
                    g.n[d] = n;
                    g.w[d] = S.size[d] / n;
                    g.origin[d] = 0;
                } else {
                    double e = mx[d] - mn[d];
                    int n = (int)std::floor(e / w) + 1;
                    if (n < 1) n = 1;
                    g.n[d] = n;
                    g.w[d] = w;
                    g.origin[d] = mn[d];
                }
                tot *= g.n[d];
            }
            if (tot <= maxCells) break;
            w *= 1.26;
        }
        int64_t G = (int64_t)g.n[0] * g.n[1] * g.n[2];
        ensure(c, S.ckeys, nb); ensure(c, S.ckeys2, nb); ensure(c, S.cvals, nb); ensure(c, S.cvals2, nb);
        ensure(c, S.cstart, G); ensure(c, S.cend, G);
        hipLaunchKernelGGL(k_cell_keys, dim3(nbk), dim3(WDM_BLOCK), 0, c->stream, S.pview(), g, (const int32_t*)S.binned.p, nb, S.ckeys.p, S.cvals.p);
        checkLaunch();
        sortPairs(S, S.ckeys.p, S.ckeys2.p, S.cvals.p, S.cvals2.p, nb, bitsFor((uint64_t)G));
        hipLaunchKernelGGL(k_fill_i32, dim3(nblk(G)), dim3(WDM_BLOCK), 0, c->stream, S.cstart.p, G, -1);
        hipLaunchKernelGGL(k_cell_ranges, dim3(nbk), dim3(WDM_BLOCK), 0, c->stream, (const uint32_t*)S.ckeys2.p, nb, S.cstart.p, S.cend.p);
        checkLaunch();
    }

    template <typename T>
    static void fullRun(Scene& S, ConBuf<T>* cb) {
        Ctx* c = S.ctx;
        hipStream_t st = c->stream;
        {
            KTimer t(c, "bounds");
            hipLaunchKernelGGL(k_bounds, dim3(nblk(S.np)), dim3(WDM_BLOCK), 0, st, S.nview(), S.pview(), S.P.distFactor, S.verletDist);
            checkLaunch();
        }
        GridArgs g;
        int nb;
        buildGrid(S, g, nb);
        int64_t ncand = 0;
        if (nb > 0) {
            ensure(c, S.counts, nb); ensure(c, S.offs, (size_t)nb + 1);
            ensure(c, S.cnt32, nb);
            int nbk = (int)nblk(nb);
            {
                KTimer t(c, "pairs_count");
                hipLaunchKernelGGL(k_pairs<0>, dim3(nbk), dim3(WDM_BLOCK), 0, st, S.pview(), g, (const uint32_t*)S.ckeys2.p, (const int32_t*)S.cvals2.p, nb,
                                   (const int32_t*)S.cstart.p, (const int32_t*)S.cend.p, (const int32_t*)S.unb.p, (int)S.unbH.size(), S.P.loneMask,
                                   S.cnt32.p, (const int64_t*)nullptr, (uint64_t*)nullptr, (int32_t*)nullptr, S.err.p);
                checkLaunch();
            }
            hipLaunchKernelGGL(k_i32_to_i64, dim3(nbk), dim3(WDM_BLOCK), 0, st, (const int32_t*)S.cnt32.p, S.counts.p, (int64_t)nb);
            checkLaunch();
            scanInclusiveShift(S, S.counts.p, S.offs.p, nb);
            ncand = readI64(S, S.offs.p + nb);
            ensure(c, S.candK, ncand); ensure(c, S.candK2, ncand); ensure(c, S.candP, ncand); ensure(c, S.candP2, ncand);
            if (ncand > 0) {
                KTimer t(c, "pairs_write");
                hipLaunchKernelGGL(k_pairs<1>, dim3(nbk), dim3(WDM_BLOCK), 0, st, S.pview(), g, (const uint32_t*)S.ckeys2.p, (const int32_t*)S.cvals2.p, nb,
                                   (const int32_t*)S.cstart.p, (const int32_t*)S.cend.p, (const int32_t*)S.unb.p, (int)S.unbH.size(), S.P.loneMask,
                                   (int32_t*)nullptr, (const int64_t*)S.offs.p, S.candK.p, S.candP.p, S.err.p);
                checkLaunch();
                sortPairs(S, S.candK.p, S.candK2.p, S.candP.p, S.candP2.p, ncand, 32 + bitsFor((uint64_t)S.np));
            }
        }
        // merge with the previous contact list
        ConBuf<T>& oldb = cb[S.cur];
        ConBuf<T>& newb = cb[1 - S.cur];
        int64_t nold = S.nc;
        ensure(c, S.ooFlag, std::max<int64_t>(nold, 1)); ensure(c, S.ooFlag64, std::max<int64_t>(nold, 1)); ensure(c, S.ooRank, (size_t)nold + 1);
        int64_t noor = 0;
        if (nold > 0) {
            hipLaunchKernelGGL(k_flag_old, dim3(nblk(nold)), dim3(WDM_BLOCK), 0, st, (const uint64_t*)oldb.key.p, (const uint32_t*)oldb.info.p, nold,
                               (const uint64_t*)S.candK2.p, ncand, S.ooFlag.p);
            hipLaunchKernelGGL(k_i32_to_i64, dim3(nblk(nold)), dim3(WDM_BLOCK), 0, st, (const int32_t*)S.ooFlag.p, S.ooFlag64.p, nold);
            checkLaunch();
            scanInclusiveShift(S, S.ooFlag64.p, S.ooRank.p, nold);
            noor = readI64(S, S.ooRank.p + nold);
        } else {
            HIPCHK(hipMemsetAsync(S.ooRank.p, 0, sizeof(int64_t), st));
        }
        int64_t nnew = ncand + noor;
        newb.reserve(c, std::max<int64_t>(nnew, 1));
        if (oldb.cap == 0) oldb.reserve(c, 1);
        ensure(c, S.oslot, std::max<int64_t>(nnew, 1)); ensure(c, S.rflag, std::max<int64_t>(nnew, 1));
        ensure(c, S.rscan, (size_t)nnew + 1);
        {
            KTimer t(c, "merge");
            if (ncand > 0)
                hipLaunchKernelGGL(k_emit_cand<T>, dim3(nblk(ncand)), dim3(WDM_BLOCK), 0, st, S.nview(), S.pview(), oldb.view(), nold,
                                   (const int64_t*)S.ooRank.p, (const uint64_t*)S.candK2.p, (const int32_t*)S.candP2.p, ncand, newb.view(),
                                   S.oslot.p, S.rflag.p, S.verletDist, S.P.distFactor);
            if (noor > 0)
                hipLaunchKernelGGL(k_emit_oor<T>, dim3(nblk(nold)), dim3(WDM_BLOCK), 0, st, oldb.view(), nold, (const int32_t*)S.ooFlag.p,
                                   (const int64_t*)S.ooRank.p, (const uint64_t*)S.candK2.p, ncand, newb.view(), S.oslot.p, S.rflag.p);
            checkLaunch();
            scanInclusiveShift(S, S.rflag.p, S.rscan.p, nnew);
            if (nnew > 0)
                hipLaunchKernelGGL(k_copy_state<T>, dim3(nblk(nnew)), dim3(WDM_BLOCK), 0, st, oldb.view(), newb.view(), nnew,
                                   (const int32_t*)S.oslot.p, (const int64_t*)S.rscan.p);
            checkLaunch();
            // next free row = number of real contacts carried over
            hipLaunchKernelGGL(k_i64_to_i32_one, dim3(1), dim3(1), 0, st, (const int64_t*)S.rscan.p + nnew, newb.counter.p);
            checkLaunch();
        }
        S.cur = 1 - S.cur;
        S.nc = nnew;
        // CSR (contact -> particle)
        int64_t ne = 2 * nnew;
        ensure(c, S.entK, std::max<int64_t>(ne, 1)); ensure(c, S.entK2, std::max<int64_t>(ne, 1));
        ensure(c, S.entV, std::max<int64_t>(ne, 1)); ensure(c, S.entV2, std::max<int64_t>(ne, 1));
        ensure(c, S.csrOff, (size_t)S.np + 1);
        {
            KTimer t(c, "csr");
            if (nnew > 0) {
                hipLaunchKernelGGL(k_csr_entries, dim3(nblk(nnew)), dim3(WDM_BLOCK), 0, st, (const int2*)newb.ids.p, nnew, S.entK.p, S.entV.p);
                checkLaunch();
                sortPairs(S, S.entK.p, S.entK2.p, S.entV.p, S.entV2.p, ne, 32 + bitsFor((uint64_t)S.np));
            }
            hipLaunchKernelGGL(k_csr_offsets, dim3(nblk(S.np + 1)), dim3(WDM_BLOCK), 0, st, (const uint64_t*)S.entK2.p, ne, S.np, S.csrOff.p);
            checkLaunch();
        }
        S.nFullRuns++;
        S.needFull = false;
    }

    static void run(Scene& S) {
        Ctx* c = S.ctx;
        resolveVerlet(S);
        bool full = S.needFull || S.verletDist == 0;
        if (!full && S.lastLeapfrogVerletChecked) {
            // Verlet decision of updateBboxes_doFullRun (:340-350), computed by the Leapfrog kernel
            HIPCHK(hipMemcpyAsync(S.hostWords, S.orL.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipMemcpyAsync(S.hostWords + 1, S.err.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            if (S.hostWords[1]) throwDeviceError(S.hostWords[1]);
            full = (S.hostWords[0] & VERLET_BIT) != 0;
        }
        if (!full) return;  // pending removals are decided inside the contact kernel
        if (S.mixed()) fullRun<float>(S, S.cb32);
        else fullRun<double>(S, S.cb64);
    }

    static void throwDeviceError(uint32_t e) {
        if (e & ERR_NAN) throw WdmError(WDM_E_NAN, "NaN force in contact");
        if (e & ERR_OVERSIZE) throw WdmError(WDM_E_OVERSIZE, "Particle larger than half of the cell size encountered.");
        if (e & ERR_CAPACITY) throw WdmError(WDM_E_STATE, "contact state rows exhausted");
        throw WdmError(WDM_E_HIP, "device error");
    }
};

// ---------------- ContactLoop::run (pkg/dem/ContactLoop.cpp:93-259) ----------------
struct ContactLoop {
    template <typename T>
    static void runT(Scene& S, ConBuf<T>* cb) {
        Ctx* c = S.ctx;
        StepArgs a = makeArgs(S);
        ConBuf<T>& b = cb[S.cur];
        if (b.cap == 0) b.reserve(c, 1);
        int64_t nc = S.nc;
        unsigned nbk = nblk(std::max<int64_t>(nc, 1));
        ensure(c, S.epartC, (size_t)nbk * 5 + 5);
        ensure(c, S.realMask, (size_t)nbk * (WDM_BLOCK / 64) + 1);
        if (nc == 0) HIPCHK(hipMemsetAsync(S.realMask.p, 0, sizeof(uint64_t), c->stream));
        if (nc > 0) {
            KTimer t(c, "contacts");
            if (S.P.law == WDM_LAW_HERTZ_DMT)
                hipLaunchKernelGGL((k_contacts<T, 1>), dim3(nbk), dim3(WDM_BLOCK), 0, c->stream, a, S.nview(), S.pview(), b.view(), nc, S.epartC.p, S.ecre.p, S.err.p, S.realMask.p);
            else
                hipLaunchKernelGGL((k_contacts<T, 0>), dim3(nbk), dim3(WDM_BLOCK), 0, c->stream, a, S.nview(), S.pview(), b.view(), nc, S.epartC.p, S.ecre.p, S.err.p, S.realMask.p);
            checkLaunch();
        }
        if (a.trackEnergy && nc > 0) finalizeEnergy(S, S.epartC, (int)nbk, 5, 3);
        S.forcesFromContacts = true;  // applied by the next Leapfrog (or materialised on download)
        S.hSizeCL = S.hSize;
    }
    static void run(Scene& S) {
        if (S.needFull) throw WdmError(WDM_E_STATE, "ContactLoop before the collider's first run");
        if (S.mixed()) runT<float>(S, S.cb32); else runT<double>(S, S.cb64);
    }
};

void stepBegin(Scene& S) {
    if (!(S.paramsSet && S.nodesSet && S.partsSet && S.matsSet)) throw WdmError(WDM_E_STATE, "scene not fully uploaded");
    if (S.periodic) cellIntegrateAndUpdate(S, S.P.dt);
    if (S.P.trackEnergy) {  // EnergyTracker::resetResettables: kinetic and elast
        HIPCHK(hipMemsetAsync(S.E.p + WDM_E_KINETIC, 0, sizeof(double) * 2, S.ctx->stream));
    }
}
void stepEnd(Scene& S) {
    if (S.periodic) cellSetNextGradV(S);
    S.step++;
}
void doOneStep(Scene& S) {
    stepBegin(S);
    Leapfrog::run(S);
    InsertionSortCollider::run(S);
    ContactLoop::run(S);
    stepEnd(S);
}
void syncAndCheck(Scene& S) {
    Ctx* c = S.ctx;
    HIPCHK(hipMemcpyAsync(S.hostWords + 1, S.err.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(S.hostWords + 0, S.orL.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (S.hostWords[1]) InsertionSortCollider::throwDeviceError(S.hostWords[1]);
}

// DemData.force/torque after ContactLoop: gravity reset plus the contact sums
void materialiseForces(Scene& S) {
    StepArgs a = makeArgs(S);
    KTimer t(S.ctx, "gather");
    hipLaunchKernelGGL(k_gather, dim3(nblk(S.nn)), dim3(WDM_BLOCK), 0, S.ctx->stream, a, S.nview(), S.pview(), S.gview());
    checkLaunch();
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
                      double* forceLocal3, double* contPt3, double* trsf9, double* uN, int64_t* nOut) {
    Ctx* c = S.ctx;
    ConBuf<T>& b = cb[S.cur];
    int64_t n = S.nc;
    std::vector<int2> ids(n);
    std::vector<uint32_t> info(n);
    std::vector<int32_t> cd(n), slot(n);
    std::vector<T> st((size_t)NSTATE * std::max<int64_t>(b.cap, 1));
    std::vector<double4> gfc(2 * (size_t)n);
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(ids.data(), b.ids.p, n * sizeof(int2), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(info.data(), b.info.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(cd.data(), b.cd.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(slot.data(), b.slot.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(st.data(), b.st.p, (size_t)NSTATE * b.cap * sizeof(T), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(gfc.data(), b.gfc.p, 2 * n * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
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
            if (forceLocal3) forceLocal3[3 * k + d] = real ? (double)st[(size_t)(S_F + d) * b.cap + row] : 0.;
        }
        if (contPt3) { contPt3[3 * k] = gfc[2 * i + 1].x; contPt3[3 * k + 1] = gfc[2 * i + 1].y; contPt3[3 * k + 2] = gfc[2 * i + 1].z; }
        if (uN) uN[k] = real ? gfc[2 * i + 1].w : NAN;
        if (trsf9) for (int q = 0; q < 9; q++) trsf9[9 * k + q] = (real && row >= 0) ? (double)st[(size_t)(S_T + q) * b.cap + row] : NAN;
        k++;
    }
    if (nOut) *nOut = k;
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
    int nb = (int)S.binnedH.size();
    ensure(c, S.disp, std::max(S.nn, 1));
    for (int it = 0; it < iters; it++) {
        hipLaunchKernelGGL(k_bounds, dim3(nblk(S.np)), dim3(WDM_BLOCK), 0, c->stream, S.nview(), S.pview(), 1.0, 0.0);
        checkLaunch();
        GridArgs g;
        int nbn;
        InsertionSortCollider::buildGrid(S, g, nbn);
        if (nb == 0) return;
        hipLaunchKernelGGL(k_relax, dim3(nblk(nb)), dim3(WDM_BLOCK), 0, c->stream, S.nview(), S.pview(), g, (const uint32_t*)S.ckeys2.p,
                           (const int32_t*)S.cvals2.p, nb, (const int32_t*)S.cstart.p, (const int32_t*)S.cend.p, S.disp.p);
        hipLaunchKernelGGL(k_relax_apply, dim3(nblk(nb)), dim3(WDM_BLOCK), 0, c->stream, S.nview(), (const int32_t*)S.binned.p, nb, (const double4*)S.disp.p);
        checkLaunch();
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    S.needFull = true;
}

}  // namespace

struct wdm_ctx : Ctx {};

extern "C" {

int wdm_create(wdm_ctx** out, int device, size_t budget) {
    if (!out) return WDM_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return WDM_E_HIP;
    if (device < 0 || device >= n) return WDM_E_ARG;
    wdm_ctx* c = new wdm_ctx();
    c->device = device;
    c->budget = budget;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return WDM_E_HIP;
    }
    *out = c;
    return WDM_OK;
}

void wdm_destroy(wdm_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->scenes) { kv.second->freeAll(); delete kv.second; }
    c->scenes.clear();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* wdm_last_error(wdm_ctx* c) { return c ? c->err.c_str() : "null context"; }

size_t wdm_scene_footprint(int32_t n_par, int32_t n_node, double cpp) {
    size_t nodeB = (size_t)n_node * (7 * 32 + 4 + 1);
    size_t parB = (size_t)n_par * (4 + 4 + 4 + 16 + 64 + 8 + 16);
    double ncon = cpp * n_par * 1.34;
    size_t conB = (size_t)(ncon * 2 * (8 + 8 + 4 + 4 + NSTATE * 8 + 64));
    size_t csrB = (size_t)(ncon * 2 * 2 * (8 + 4));
    size_t scratch = (size_t)n_par * 64 + (size_t)(ncon * 24);
    return nodeB + parB + conB + csrB + scratch + (1u << 20);
}

int wdm_scene_alloc(wdm_ctx* c, int32_t np, int32_t nn, int32_t nm, int32_t* sid) {
    return guard(c, [&] {
        if (np < 0 || nn < np || nm < 1 || !sid) throw WdmError(WDM_E_ARG, "bad sizes (need n_node >= n_par, n_mat >= 1)");
        Scene* S = new Scene(c, np, nn, nm);
        try {
            ensure(c, S->posr, std::max(nn, 1)); ensure(c, S->velm, std::max(nn, 1)); ensure(c, S->angI, std::max(nn, 1));
            ensure(c, S->ori, std::max(nn, 1)); ensure(c, S->force, std::max(nn, 1)); ensure(c, S->torque, std::max(nn, 1));
            ensure(c, S->lastPos, std::max(nn, 1)); ensure(c, S->flags, std::max(nn, 1)); ensure(c, S->checked, std::max(nn, 1));
            ensure(c, S->ptype, std::max(np, 1)); ensure(c, S->pmat, std::max(np, 1)); ensure(c, S->pmask, std::max(np, 1));
            ensure(c, S->fnodes, std::max(np, 1)); ensure(c, S->bmin, std::max(np, 1)); ensure(c, S->bmax, std::max(np, 1));
            ensure(c, S->mat, nm);
            ensure(c, S->E, WDM_E_COUNT); ensure(c, S->eslots, 8); ensure(c, S->orL, 1); ensure(c, S->ecre, 1); ensure(c, S->err, 1);
            ensure(c, S->csrOff, (size_t)np + 1);
            HIPCHK(hipMemsetAsync(S->E.p, 0, sizeof(double) * WDM_E_COUNT, c->stream));
            HIPCHK(hipMemsetAsync(S->orL.p, 0, 4, c->stream));
            HIPCHK(hipMemsetAsync(S->ecre.p, 0, 4, c->stream));
            HIPCHK(hipMemsetAsync(S->err.p, 0, 4, c->stream));
            HIPCHK(hipMemsetAsync(S->csrOff.p, 0, sizeof(int64_t) * ((size_t)np + 1), c->stream));
            int32_t slots[8] = {WDM_E_NONVISCDAMP, WDM_E_GRAV, WDM_E_KINETIC, WDM_E_ELAST, WDM_E_PLAST, WDM_E_VISCN, WDM_E_VISCT, WDM_E_BROKEN};
            HIPCHK(hipMemcpyAsync(S->eslots.p, slots, sizeof(slots), hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipHostMalloc((void**)&S->hostWords, 16, hipHostMallocDefault));
            std::memset(S->hostWords, 0, 16);
            HIPCHK(hipStreamSynchronize(c->stream));
        } catch (...) {
            S->freeAll();
            delete S;
            throw;
        }
        *sid = c->nextId++;
        c->scenes[*sid] = S;
    });
}

int wdm_scene_free(wdm_ctx* c, int32_t sid) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        HIPCHK(hipStreamSynchronize(c->stream));
        S.freeAll();
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
        for (int i = 0; i < n; i++) {
            if (flags[i] & (WDM_CLUMP_CLUMP | WDM_CLUMP_CLUMPED)) throw WdmError(WDM_E_UNSUPPORTED, "clumps are not supported on this path");
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
        S.forcesFromContacts = false;
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
        if (S.paramsSet && p->verletDist != S.P.verletDist) S.verletResolved = false;
        S.P = *p;
        S.paramsSet = true;
    });
}

int wdm_step_begin(wdm_ctx* c, int32_t sid) { return guard(c, [&] { stepBegin(getScene(c, sid)); }); }
int wdm_leapfrog_run(wdm_ctx* c, int32_t sid) { return guard(c, [&] { Leapfrog::run(getScene(c, sid)); }); }
int wdm_collider_run(wdm_ctx* c, int32_t sid) { return guard(c, [&] { InsertionSortCollider::run(getScene(c, sid)); }); }
int wdm_contact_loop_run(wdm_ctx* c, int32_t sid) { return guard(c, [&] { ContactLoop::run(getScene(c, sid)); }); }
int wdm_step_end(wdm_ctx* c, int32_t sid) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        stepEnd(S);
        syncAndCheck(S);
    });
}

int wdm_step(wdm_ctx* c, const int32_t* scenes, int32_t ns, int32_t nsteps) {
    return guard(c, [&] {
        if (ns < 0 || nsteps < 0 || (ns > 0 && !scenes)) throw WdmError(WDM_E_ARG, "bad scene list");
        std::vector<Scene*> ss;
        for (int i = 0; i < ns; i++) ss.push_back(&getScene(c, scenes[i]));
        if (c->timing.on) c->timing.acc.clear();
        for (int k = 0; k < nsteps; k++)
            for (Scene* S : ss) doOneStep(*S);
        for (Scene* S : ss) syncAndCheck(*S);
    });
}

int wdm_download_nodes(wdm_ctx* c, int32_t sid, double* pos3, double* ori4, double* vel3, double* angVel3, double* force3, double* torque3) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        int n = S.nn;
        if (S.forcesFromContacts && (force3 || torque3)) materialiseForces(S);
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

int wdm_get_contacts(wdm_ctx* c, int32_t sid, int64_t cap, int32_t* idA, int32_t* idB, int32_t* cellDist3, uint8_t* isReal, double* forceLocal3,
                     double* contPt3, double* trsf9, double* uN, int64_t* nOut) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        if (S.mixed()) downloadContacts(S, S.cb32, cap, idA, idB, cellDist3, isReal, forceLocal3, contPt3, trsf9, uN, nOut);
        else downloadContacts(S, S.cb64, cap, idA, idB, cellDist3, isReal, forceLocal3, contPt3, trsf9, uN, nOut);
    });
}

int wdm_get_energy(wdm_ctx* c, int32_t sid, double* values, int32_t* created) {
    return guard(c, [&] {
        Scene& S = getScene(c, sid);
        double e[WDM_E_COUNT];
        uint32_t cre = 0;
        HIPCHK(hipMemcpyAsync(e, S.E.p, sizeof(e), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(&cre, S.ecre.p, sizeof(cre), hipMemcpyDeviceToHost, c->stream));
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

int wdm_enable_timing(wdm_ctx* c, int32_t on) { return guard(c, [&] { c->timing.on = on != 0; c->timing.acc.clear(); }); }

int wdm_kernel_times(wdm_ctx* c, int32_t cap, char* names, double* ms, int64_t* launches, int32_t* nOut) {
    return guard(c, [&] {
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
