/*
 * CPU oracle for the Woo DEM timestep -- TEST INFRASTRUCTURE ONLY (see dem_oracle.h).
 *
 * Restatement of the reference's CPU arithmetic, operation by operation.  Built
 * with -ffp-contract=off so that no a*b+c is fused.  Conventions that the
 * reference leaves to its Eigen build (and that this restatement fixes):
 *   - 3-term sums (dot products, matrix-vector rows) are evaluated left to right,
 *     (a0*b0 + a1*b1) + a2*b2;
 *   - pow(x,2) is x*x, pow(sqrt(x),3) is (s*s)*s, pow(sqrt(x),5) is (((s*s)*s)*s)*s;
 *   - the contact-to-global rotation uses the L6Geom.trsf matrix directly
 *     (the reference goes through Quaternionr(trsf), equal to a few ulp);
 *   - contacts are kept in canonical order: Wall/Facet first (functor order,
 *     pkg/dem/Wall.hpp:54-55, Facet.hpp:79-80), otherwise the lower particle id
 *     first (the reference's order comes from insertion-sort history,
 *     pkg/dem/InsertionSortCollider.cpp:87,190);
 *   - forces are applied deterministically (ContactLoop.deterministic,
 *     pkg/dem/ContactLoop.cpp:250-257), contacts visited in (min id, max id) order,
 *     so each particle sums its contacts in ascending order of the other id.
 *   - energies are summed sequentially in node order / contact order;
 *   - quaternion products and quaternion*vector follow Eigen 3's scalar formulas
 *     (the reference's build may vectorise them, equal to a few ulp).
 * The "parity unpinned" items are listed in the module docstring of
 * tests/oracle_lib.py.
 */
#include "dem_oracle.h"

#include <algorithm>
#include <deque>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double INF = std::numeric_limits<double>::infinity();
const double PI = 3.14159265358979323846;

/* ---------- small fixed-size algebra (left-to-right evaluation) ---------- */
struct V3 { double x[3]; double& operator[](int i) { return x[i]; } const double& operator[](int i) const { return x[i]; } };
struct M3 { double m[3][3]; };
struct Qt { double w, x, y, z; };

inline V3 v3(double a, double b, double c) { V3 r; r[0] = a; r[1] = b; r[2] = c; return r; }
inline V3 vzero() { return v3(0, 0, 0); }
inline V3 vadd(const V3& a, const V3& b) { return v3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline V3 vsub(const V3& a, const V3& b) { return v3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline V3 vmul(const V3& a, double s) { return v3(a[0] * s, a[1] * s, a[2] * s); }
inline V3 vdiv(const V3& a, double s) { return v3(a[0] / s, a[1] / s, a[2] / s); }
inline double vdot(const V3& a, const V3& b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
inline double vsqn(const V3& a) { return vdot(a, a); }
inline double vnorm(const V3& a) { return std::sqrt(vsqn(a)); }
inline V3 vcross(const V3& a, const V3& b) {
    return v3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
inline bool viszero(const V3& a) { return a[0] == 0 && a[1] == 0 && a[2] == 0; }
/* Eigen normalize(): divide by sqrt(squaredNorm) when positive */
inline V3 vnormalized(const V3& a) { double z = vsqn(a); if (z > 0) return vdiv(a, std::sqrt(z)); return a; }
inline V3 vabs(const V3& a) { return v3(std::fabs(a[0]), std::fabs(a[1]), std::fabs(a[2])); }

inline M3 mzero() { M3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[i][j] = 0; return r; }
inline M3 meye() { M3 r = mzero(); r.m[0][0] = r.m[1][1] = r.m[2][2] = 1; return r; }
inline V3 mcol(const M3& a, int j) { return v3(a.m[0][j], a.m[1][j], a.m[2][j]); }
inline void msetcol(M3& a, int j, const V3& c) { a.m[0][j] = c[0]; a.m[1][j] = c[1]; a.m[2][j] = c[2]; }
inline V3 mvmul(const M3& a, const V3& v) {
    return v3((a.m[0][0] * v[0] + a.m[0][1] * v[1]) + a.m[0][2] * v[2],
              (a.m[1][0] * v[0] + a.m[1][1] * v[1]) + a.m[1][2] * v[2],
              (a.m[2][0] * v[0] + a.m[2][1] * v[1]) + a.m[2][2] * v[2]);
}
/* transpose(a) * v == column dots */
inline V3 mtvmul(const M3& a, const V3& v) { return v3(vdot(mcol(a, 0), v), vdot(mcol(a, 1), v), vdot(mcol(a, 2), v)); }
inline M3 mmul(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = (a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j];
    return r;
}
inline M3 madd(const M3& a, const M3& b) { M3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[i][j] = a.m[i][j] + b.m[i][j]; return r; }
inline M3 msub(const M3& a, const M3& b) { M3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[i][j] = a.m[i][j] - b.m[i][j]; return r; }
inline M3 mscale(const M3& a, double s) { M3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[i][j] = a.m[i][j] * s; return r; }
inline M3 mdivs(const M3& a, double s) { M3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[i][j] = a.m[i][j] / s; return r; }
inline M3 mtrans(const M3& a) { M3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[i][j] = a.m[j][i]; return r; }
inline bool mequal(const M3& a, const M3& b) { for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) if (a.m[i][j] != b.m[i][j]) return false; return true; }
inline double mcof(const M3& a, int i, int j) {
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return a.m[i1][j1] * a.m[i2][j2] - a.m[i1][j2] * a.m[i2][j1];
}
/* 3x3 inverse by cofactors times 1/det (Eigen 3 compute_inverse_size3 algorithm) */
inline M3 minverse(const M3& a) {
    double c0 = mcof(a, 0, 0), c1 = mcof(a, 1, 0), c2 = mcof(a, 2, 0);
    double det = (c0 * a.m[0][0] + c1 * a.m[1][0]) + c2 * a.m[2][0];
    double inv = 1.0 / det;
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = mcof(a, j, i) * inv;
    return r;
}
inline double mdet(const M3& a) {
    double c0 = mcof(a, 0, 0), c1 = mcof(a, 1, 0), c2 = mcof(a, 2, 0);
    return (c0 * a.m[0][0] + c1 * a.m[1][0]) + c2 * a.m[2][0];
}
/* lib/base/Math.hpp:222-228 */
inline V3 leviCivita(const M3& m) { return v3(m.m[1][2] - m.m[2][1], m.m[2][0] - m.m[0][2], m.m[0][1] - m.m[1][0]); }

inline Qt qmul(const Qt& a, const Qt& b) {
    Qt r;
    r.w = ((a.w * b.w - a.x * b.x) - a.y * b.y) - a.z * b.z;
    r.x = ((a.w * b.x + a.x * b.w) + a.y * b.z) - a.z * b.y;
    r.y = ((a.w * b.y + a.y * b.w) + a.z * b.x) - a.x * b.z;
    r.z = ((a.w * b.z + a.z * b.w) + a.x * b.y) - a.y * b.x;
    return r;
}
/* Quaternion * vector as Eigen 3 evaluates it (Quaternion::_transformVector):
   t = 2 (q.vec x v); v' = (v + w t) + q.vec x t */
inline V3 qrotate(const Qt& q, const V3& v) {
    V3 u = v3(q.x, q.y, q.z);
    V3 t = vcross(u, v);
    t = vadd(t, t);
    return vadd(vadd(v, vmul(t, q.w)), vcross(u, t));
}
inline Qt qnormalized(const Qt& q) {
    double n = std::sqrt(((q.w * q.w + q.x * q.x) + q.y * q.y) + q.z * q.z);
    Qt r; r.w = q.w / n; r.x = q.x / n; r.y = q.y / n; r.z = q.z / n; return r;
}

/* Mathr::Sign, lib/base/Math.hpp:162 */
inline double Sign(double f) { if (f < 0) return -1; if (f > 0) return 1; return 0; }

/* ---------- data model ---------- */
/* DemData flags, pkg/dem/Particle.hpp:177-182 */
enum : uint32_t {
    DOF_X = 1, DOF_Y = 2, DOF_Z = 4, DOF_RX = 8, DOF_RY = 16, DOF_RZ = 32,
    CLUMP_CLUMPED = 64, CLUMP_CLUMP = 128, ENERGY_SKIP = 256, GRAVITY_SKIP = 512, TRACER_SKIP = 1024, DAMPING_SKIP = 2048,
    DOF_ALL = 63, DOF_RXRYRZ = 56
};
inline uint32_t axisDOF(int axis, bool rot) { return 1u << (axis + (rot ? 3 : 0)); }

enum { HOMO_NONE = 0, HOMO_POS = 1, HOMO_VEL = 2, HOMO_VEL_2ND = 3, HOMO_GRADV2 = 4 };

struct Mat { double density, young, ktDivKn, tanPhi; };

/* Contact + L6Geom + FrictPhys/HertzPhys, pkg/dem/Contact.hpp:47-115, L6Geom.hpp:10-26, Hertz.hpp:5-21 */
struct Contact {
    int idA, idB;
    int cd[3];
    long stepCreated;
    double minDist00Sq;
    bool geom, phys;
    // L6Geom
    M3 trsf; V3 contPt; double uN; double lens[2]; double contA; V3 vel, angVel;
    // FrictPhys
    double kn, kt, tanPhi; V3 force, torque;
    // HertzPhys
    double K, R, kt0, alpha_sqrtMK, gamma, alpha;
    bool isReal() const { return geom && phys; }
};

/* InsertionSortCollider::Bounds, pkg/dem/InsertionSortCollider.hpp:120-141 */
struct Bounds {
    double coord; int id; int period; unsigned hasBB : 1; unsigned isMin : 1; unsigned isInf : 1;
    bool operator<(const Bounds& b) const { if (id == b.id && coord == b.coord) return isMin; return coord < b.coord; }
    bool operator>(const Bounds& b) const { if (id == b.id && coord == b.coord) return !isMin; return coord > b.coord; }
};
/* VecBounds, pkg/dem/InsertionSortCollider.hpp:145-168 */
struct VecBounds {
    int axis = -1; std::vector<Bounds> vec; double cellDim = 0; long size = 0; long loIdx = 0;
    Bounds& operator[](long i) { return vec[i]; }
    long norm(long i) const { if (i < 0) i += size; return i % size; }
};

struct Aabb { V3 mn, mx; V3 lastPos[3]; int nNodes = 0; double maxD2 = 0; bool has = false; };

struct Oracle {
    std::string err;
    orc_params P{};
    // nodes
    int nn = 0;
    std::vector<V3> pos, vel, angVel, force, torque, inertia;
    std::vector<Qt> ori;
    std::vector<double> mass;
    std::vector<uint32_t> flags;
    // clumps (ClumpData, pkg/dem/Clump.hpp): master node, members with relPos/relOri
    std::vector<int> clumpMaster, clumpOff, clumpMem;  // per clump; offsets; member nodes
    std::vector<V3> clumpRelPos;
    std::vector<Qt> clumpRelOri;
    std::vector<int> nodeMaster, nodeClump;  // per node: master node of a member (-1), clump index of a master (-1)
    // particles
    int np = 0;
    std::vector<int> shape, nodeIx, axis, sense, matId;  // nodeIx: 3 per particle
    std::vector<double> radius;
    std::vector<uint32_t> mask;
    std::vector<Mat> mats;
    // cell
    bool periodic = false;
    M3 hSize = meye(), gradV = mzero(), nextGradV = mzero(), pprevHsize = meye(), trsf = meye();
    V3 cellSize = v3(1, 1, 1), spinVec = vzero();
    int homoDeform = HOMO_GRADV2;
    // scene
    long step = 0;
    // energies
    double E[ORC_E_N] = {0};
    int Ecreated[ORC_E_N] = {0};
    // collider
    bool contactsDirty = true;
    bool colPeriodic = false;
    double verletDistActual = NAN;
    bool verletResolved = false;
    long nFullRuns = 0;
    long sortChunks = -1, sortPasses = 0;  // the last parallel insertion sort (sortChunks, InsertionSortCollider.hpp:264)
    VecBounds BB[3];
    std::vector<double> minima, maxima;
    std::vector<Aabb> aabb;
    // contacts: key (min<<32|max) -> contact
    std::map<uint64_t, Contact> con;
    std::vector<uint64_t> pending;
    // contacts to make / remove after the sort, one list per OpenMP thread
    // (InsertionSortCollider.hpp:94-113: rremoveContacts / rmakeContacts per thread)
    std::vector<std::vector<Contact>> makeLater;
    std::vector<std::vector<uint64_t>> removeLater;
    int tid() const {
#ifdef _OPENMP
        return omp_get_thread_num();
#else
        return 0;
#endif
    }
    static int maxThreads() {
#ifdef _OPENMP
        return omp_get_max_threads();
#else
        return 1;
#endif
    }
    // deterministic force application: per-particle list of (other id, contact*) sorted by other id
    bool adjDirty = true;
    std::vector<Contact*> lin;
    std::vector<long> adjOff;
    std::vector<Contact*> adjC;

    int nodeOf(int p, int k = 0) const { return nodeIx[3 * p + k]; }
    int numNodes(int p) const { return shape[p] == ORC_SHAPE_FACET ? 3 : (shape[p] == ORC_SHAPE_NONE ? 0 : 1); }
    static uint64_t key(int a, int b) { int lo = a < b ? a : b, hi = a < b ? b : a; return ((uint64_t)(uint32_t)lo << 32) | (uint32_t)hi; }

    /* ---------- energy tracker, core/EnergyTracker.cpp:10-17 ---------- */
    void eAdd(int slot, double val, bool zeroDontCreate) {
        if (!Ecreated[slot]) { if (val == 0. && zeroDontCreate) return; Ecreated[slot] = 1; }
        if (std::isnan(val)) return;
        E[slot] += val;
    }

    /* ---------- Cell, core/Cell.cpp:28-120 ---------- */
    void cellIntegrateAndUpdate(double dt) {
        M3 trsfInc = mscale(gradV, dt);
        trsf = madd(trsf, mmul(trsfInc, trsf));
        M3 prevHsize = hSize;
        if (homoDeform == HOMO_GRADV2) {
            M3 gh = mdivs(mscale(gradV, dt), 2.);
            hSize = mmul(mmul(minverse(msub(meye(), gh)), madd(meye(), gh)), hSize);
        } else hSize = madd(hSize, mmul(trsfInc, hSize));
        pprevHsize = mscale(madd(prevHsize, hSize), .5);
        if (mdet(hSize) == 0) throw std::runtime_error("Cell is degenerate (zero volume).");
        for (int i = 0; i < 3; i++) cellSize[i] = vnorm(mcol(hSize, i));
        if (hSize.m[0][1] != 0 || hSize.m[0][2] != 0 || hSize.m[1][0] != 0 || hSize.m[1][2] != 0 || hSize.m[2][0] != 0 || hSize.m[2][1] != 0)
            throw std::runtime_error("sheared periodic cells are outside the restated path");
    }
    void cellSetNextGradV() {  // core/Cell.cpp:53-58
        gradV = nextGradV;
        M3 W = mscale(msub(gradV, mtrans(gradV)), .5);
        spinVec = vmul(leviCivita(W), .5);
    }
    V3 intrShiftPos(const int cd[3]) const { return mvmul(hSize, v3((double)cd[0], (double)cd[1], (double)cd[2])); }
    V3 intrShiftVel(const int cd[3]) const {  // core/Cell.cpp:28-34
        V3 c = v3((double)cd[0], (double)cd[1], (double)cd[2]);
        switch (homoDeform) {
            case HOMO_VEL: case HOMO_VEL_2ND: return mvmul(mmul(gradV, hSize), c);
            case HOMO_GRADV2: return mvmul(mmul(gradV, pprevHsize), c);
            default: return vzero();
        }
    }
    V3 pprevFluctVel(const V3& p, const V3& v, double dt) const {  // core/Cell.cpp:36-50
        switch (homoDeform) {
            case HOMO_NONE: case HOMO_POS: return v;
            case HOMO_VEL: case HOMO_VEL_2ND: return vsub(v, mvmul(gradV, p));
            case HOMO_GRADV2: return vsub(v, mvmul(gradV, vsub(p, vmul(v, dt / 2))));
        }
        return v;
    }
    V3 pprevFluctAngVel(const V3& w) const { return homoDeform == HOMO_GRADV2 ? vsub(w, spinVec) : w; }

    /* ---------- Leapfrog::run, pkg/dem/Leapfrog.cpp:103-255 ---------- */
    void leapfrog() {
        const double dt = P.dt;
        const double damping = P.damping;
        int hd = periodic ? homoDeform : -1;
        M3 ImLL4hInv = mzero(), IpLL4h = mzero(), LmL = mzero(); V3 deltaSpinVec = vzero();
        M3 dGradV = msub(nextGradV, gradV), midGradV = mscale(madd(gradV, nextGradV), .5);
        if (hd == HOMO_GRADV2) {
            const M3& pprevL = gradV; const M3& nnextL = nextGradV;
            ImLL4hInv = minverse(msub(meye(), mdivs(mscale(madd(nnextL, pprevL), dt), 4.)));
            IpLL4h = madd(meye(), mdivs(mscale(madd(nnextL, pprevL), dt), 4.));
            LmL = msub(nnextL, pprevL);
            deltaSpinVec = vadd(vmul(leviCivita(mscale(msub(pprevL, mtrans(pprevL)), .5)), -.5),
                                vmul(leviCivita(mscale(msub(nnextL, mtrans(nnextL)), .5)), .5));
        }
        const bool reallyTrackEnergy = P.trackEnergy && (!periodic || step > 0 || mequal(gradV, meye()));
        V3 g = v3(P.gravity[0], P.gravity[1], P.gravity[2]);
        bool hasGravity = !viszero(g);
        std::vector<double> eDamp(nn, 0.), eGrav(nn, 0.), eKin(nn, 0.);
        std::vector<char> hDamp(nn, 0), hGrav(nn, 0), hKin(nn, 0);
        #pragma omp parallel for schedule(guided)
        for (int i = 0; i < nn; i++) {
            uint32_t fl = flags[i];
            if (fl & CLUMP_CLUMPED) continue;  // integrated through the clump's master node
            const int ci = (fl & CLUMP_CLUMP) ? nodeClump[i] : -1;
            bool damp = (damping != 0. && !(fl & DAMPING_SKIP));
            if (ci >= 0) {  // ClumpData::forceTorqueFromMembers, Clump.cpp:265-273
                for (int k = clumpOff[ci]; k < clumpOff[ci + 1]; k++) {
                    int mnode = clumpMem[k];
                    force[i] = vadd(force[i], force[mnode]);
                    torque[i] = vadd(torque[i], vadd(torque[mnode], vcross(vsub(pos[mnode], pos[i]), force[mnode])));
                }
            }
            V3 f = force[i], t = torque[i];
            bool blockedNone = (fl & DOF_ALL) == 0, blockedAll = (fl & DOF_ALL) == DOF_ALL;
            if (reallyTrackEnergy) {
                if (damp && !(fl & ENERGY_SKIP)) {  // :52-63
                    double d = vdot(vabs(vel[i]), vabs(f)) * damping * dt + vdot(vabs(angVel[i]), vabs(t)) * damping * dt;
                    eDamp[i] = d; hDamp[i] = 1;
                }
                if (hasGravity && !(fl & GRAVITY_SKIP)) {  // :66-73
                    double gr = 0;
                    if (blockedNone) gr = -vdot(g, vel[i]) * mass[i] * dt;
                    else { for (int ax = 0; ax < 3; ax++) if (!(fl & axisDOF(ax, false))) gr -= g[ax] * vel[i][ax] * mass[i] * dt; }
                    eGrav[i] = gr; hGrav[i] = 1;
                }
            }
            V3 pprevFV = vzero(), pprevFAV = vzero();
            V3 linAccel = vzero(), angAccel = vzero();
            if (!blockedAll) {
                // computeAccel, :36-41
                if (blockedNone) linAccel = vdiv(f, mass[i]);
                else { linAccel = vzero(); for (int k = 0; k < 3; k++) if (!(fl & axisDOF(k, false))) linAccel[k] += f[k] / mass[i]; }
                if (periodic) { pprevFV = pprevFluctVel(pos[i], vel[i], dt); pprevFAV = pprevFluctAngVel(angVel[i]); }
                else { pprevFV = vel[i]; pprevFAV = angVel[i]; }
                if (damp) for (int k = 0; k < 3; k++) linAccel[k] *= 1 - damping * Sign(f[k] * (pprevFV[k] + 0.5 * dt * linAccel[k]));  // :32-34
                if (hd == HOMO_GRADV2) vel[i] = mvmul(ImLL4hInv, vadd(vadd(mvmul(LmL, pos[i]), mvmul(IpLL4h, vel[i])), vmul(linAccel, dt)));
                else vel[i] = vadd(vel[i], vmul(linAccel, dt));
                if (!viszero(inertia[i])) {
                    bool aspherical = !(inertia[i][0] == inertia[i][1] && inertia[i][1] == inertia[i][2]);
                    if (aspherical && (fl & DOF_RXRYRZ) != DOF_RXRYRZ) { /* rejected at upload */ }
                    // computeAngAccel, :42-47
                    if (blockedNone) angAccel = v3(t[0] / inertia[i][0], t[1] / inertia[i][1], t[2] / inertia[i][2]);
                    else { angAccel = vzero(); for (int k = 0; k < 3; k++) if (!(fl & axisDOF(k, true))) angAccel[k] += t[k] / inertia[i][k]; }
                    if (damp) for (int k = 0; k < 3; k++) angAccel[k] *= 1 - damping * Sign(t[k] * (pprevFAV[k] + 0.5 * dt * angAccel[k]));
                    angVel[i] = vadd(angVel[i], vmul(angAccel, dt));
                    if (hd == HOMO_GRADV2) angVel[i] = vsub(angVel[i], deltaSpinVec);
                }
            } else {
                if (hd == HOMO_GRADV2) vel[i] = mvmul(ImLL4hInv, vadd(mvmul(LmL, pos[i]), mvmul(IpLL4h, vel[i])));
            }
            if (hd >= 0 && hd != HOMO_GRADV2) {  // applyPeriodicCorrections, :258-272
                if (hd == HOMO_VEL || hd == HOMO_VEL_2ND) {
                    if (hd == HOMO_VEL_2ND) vel[i] = vadd(vel[i], vmul(mvmul(midGradV, vsub(vel[i], vmul(linAccel, dt / 2.))), dt));
                    vel[i] = vadd(vel[i], mvmul(dGradV, pos[i]));
                } else if (hd == HOMO_POS) pos[i] = vadd(pos[i], vmul(mvmul(nextGradV, pos[i]), dt));
            }
            if (reallyTrackEnergy && !(fl & ENERGY_SKIP)) {  // doKineticEnergy, :75-97
                V3 cfv = vadd(pprevFV, vmul(linAccel, .5 * dt)), cfav = vadd(pprevFAV, vmul(angAccel, .5 * dt));
                double Etrans = .5 * mass[i] * vsqn(cfv);
                double Erot = 0.5 * vdot(cfav, v3(inertia[i][0] * cfav[0], inertia[i][1] * cfav[1], inertia[i][2] * cfav[2]));
                double mxI = std::max(std::max(inertia[i][0], inertia[i][1]), inertia[i][2]);
                if (std::isnan(Erot) || std::isinf(mxI)) Erot = 0;
                eKin[i] = Etrans + Erot; hKin[i] = 1;
            }
            // leapfrogTranslate, :275-278
            pos[i] = vadd(pos[i], vmul(vel[i], dt));
            // leapfrogSphericalRotate, :280-288
            V3 ax = angVel[i];
            if (!viszero(ax)) {
                double angle = vnorm(ax); ax = vdiv(ax, angle);
                double ha = 0.5 * (angle * dt);
                double s = std::sin(ha);
                Qt q; q.w = std::cos(ha); q.x = s * ax[0]; q.y = s * ax[1]; q.z = s * ax[2];
                ori[i] = qmul(q, ori[i]);
            }
            ori[i] = qnormalized(ori[i]);
            // reset, :240-245
            force[i] = (hasGravity && !(fl & GRAVITY_SKIP)) ? vmul(g, mass[i]) : vzero();
            torque[i] = vzero();
            if (ci >= 0) {  // ClumpData::applyToMembers, Clump.cpp:276-289 (reset: members' force and torque zeroed)
                for (int k = clumpOff[ci]; k < clumpOff[ci + 1]; k++) {
                    int mnode = clumpMem[k];
                    pos[mnode] = vadd(pos[i], qrotate(ori[i], clumpRelPos[k]));
                    ori[mnode] = qmul(ori[i], clumpRelOri[k]);
                    vel[mnode] = vadd(vel[i], vcross(angVel[i], vsub(pos[mnode], pos[i])));
                    angVel[mnode] = angVel[i];
                    force[mnode] = vzero();
                    torque[mnode] = vzero();
                }
            }
        }
        if (reallyTrackEnergy) {
            for (int i = 0; i < nn; i++) {
                if (hDamp[i]) eAdd(ORC_E_NONVISCDAMP, eDamp[i], true);
                if (hGrav[i]) eAdd(ORC_E_GRAV, eGrav[i], false);
                if (hKin[i]) eAdd(ORC_E_KINETIC, eKin[i], false);
            }
        }
    }

    /* ---------- bounds: Bo1_Sphere_Aabb (Sphere.cpp:50-79), Bo1_Wall_Aabb (Wall.cpp:24-33), Bo1_Facet_Aabb (Facet.cpp:177-200) ---------- */
    void computeBound(int p) {
        Aabb& a = aabb[p];
        a.has = true;
        if (shape[p] == ORC_SHAPE_SPHERE) {
            double hs = (P.distFactor > 0 ? P.distFactor : 1.) * radius[p];
            const V3& x = pos[nodeOf(p)];
            // non-sheared periodic cell: unshearPt(pos)==pos, shearAlignedExtents(h)==h
            a.mn = v3(x[0] - hs, x[1] - hs, x[2] - hs);
            a.mx = v3(x[0] + hs, x[1] + hs, x[2] + hs);
        } else if (shape[p] == ORC_SHAPE_WALL) {
            a.mn = v3(-INF, -INF, -INF); a.mx = v3(INF, INF, INF);
            a.mn[axis[p]] = a.mx[axis[p]] = pos[nodeOf(p)][axis[p]];
        } else if (shape[p] == ORC_SHAPE_FACET) {
            double ht = radius[p];
            const V3& x0 = pos[nodeOf(p, 0)];
            a.mn = v3(x0[0] - ht, x0[1] - ht, x0[2] - ht); a.mx = v3(x0[0] + ht, x0[1] + ht, x0[2] + ht);
            for (int k = 1; k < 3; k++) {
                const V3& x = pos[nodeOf(p, k)];
                for (int d = 0; d < 3; d++) { a.mn[d] = std::min(a.mn[d], x[d] - ht); a.mx[d] = std::max(a.mx[d], x[d] + ht); }
            }
        } else a.has = false;
    }

    /* InsertionSortCollider::updateBboxes_doFullRun, InsertionSortCollider.cpp:281-417 */
    bool updateBboxes_doFullRun() {
        if (!verletResolved) {  // :290-298
            double vd = P.verletDist;
            if (vd < 0) {
                double minR = INF;
                for (int p = 0; p < np; p++) if (shape[p] == ORC_SHAPE_SPHERE) minR = std::min(radius[p], minR);
                vd = std::isinf(minR) ? 0 : std::fabs(vd) * minR;
            }
            verletDistActual = vd; verletResolved = true;
        }
        const double vd = verletDistActual;
        bool recompute = false;
        if (vd == 0) recompute = true;
        else {
            for (int p = 0; p < np; p++) {
                if (shape[p] == ORC_SHAPE_NONE) continue;
                const Aabb& a = aabb[p];
                if (!a.has) { recompute = true; break; }
                double d2 = 0;
                for (int k = 0; k < a.nNodes; k++) d2 = std::max(d2, vsqn(vsub(a.lastPos[k], pos[nodeOf(p, k)])));
                if (d2 > a.maxD2) { recompute = true; break; }  // moveDueToRot2==0: all bounds here set maxRot=-1
            }
        }
        if (!recompute) return false;
        // :374, serial at the first full run
        #pragma omp parallel for num_threads(nFullRuns > 0 ? maxThreads() : 1)
        for (int p = 0; p < np; p++) {
            if (shape[p] == ORC_SHAPE_NONE) continue;
            computeBound(p);
            Aabb& a = aabb[p];
            a.nNodes = numNodes(p);
            for (int k = 0; k < a.nNodes; k++) a.lastPos[k] = pos[nodeOf(p, k)];
            a.maxD2 = vd * vd;
            if (vd > 0) {
                for (int d = 0; d < 3; d++) { a.mx[d] += vd; a.mn[d] -= vd; }
            }
        }
        return true;
    }

    /* Collider::mayCollide, pkg/dem/Collision.cpp:19-41 */
    bool mayCollide(int a, int b) const {
        if (a == b || shape[a] == ORC_SHAPE_NONE || shape[b] == ORC_SHAPE_NONE) return false;
        if (!(mask[a] & mask[b])) return false;
        if ((mask[a] & mask[b] & P.loneMask) != 0) return false;
        for (int i = 0; i < numNodes(a); i++)
            for (int j = 0; j < numNodes(b); j++) {
                int na = nodeOf(a, i), nb = nodeOf(b, j);
                if (na == nb) return false;  // shared node
                if ((flags[na] & CLUMP_CLUMPED) && (flags[nb] & CLUMP_CLUMPED) && nodeMaster[na] == nodeMaster[nb]) return false;  // same clump
            }
        return true;
    }

    /* InsertionSortCollider::spatialOverlap, :52-58 */
    bool spatialOverlap(int id1, int id2) const {
        return (minima[3 * id1 + 0] <= maxima[3 * id2 + 0]) && (maxima[3 * id1 + 0] >= minima[3 * id2 + 0]) &&
               (minima[3 * id1 + 1] <= maxima[3 * id2 + 1]) && (maxima[3 * id1 + 1] >= minima[3 * id2 + 1]) &&
               (minima[3 * id1 + 2] <= maxima[3 * id2 + 2]) && (maxima[3 * id1 + 2] >= minima[3 * id2 + 2]);
    }
    /* cellWrap / cellWrapRel, :650-660 */
    static double cellWrap(double x, double x0, double x1, int& period) {
        double xNorm = (x - x0) / (x1 - x0);
        period = (int)std::floor(xNorm);
        return x0 + (xNorm - period) * (x1 - x0);
    }
    static double cellWrapRel(double x, double x0, double x1) {
        double xNorm = (x - x0) / (x1 - x0);
        return (xNorm - std::floor(xNorm)) * (x1 - x0);
    }
    /* spatialOverlapPeri, :903-957 */
    bool spatialOverlapPeri(int id1, int id2, int periods[3]) const {
        for (int axis = 0; axis < 3; axis++) {
            double dim = cellSize[axis];
            double m1 = minima[3 * id1 + axis], m2 = minima[3 * id2 + axis];
            if (std::isinf(m1) || std::isinf(m2)) { periods[axis] = 0; continue; }
            double wMn = (cellWrapRel(m1, m2, m2 + dim) < cellWrapRel(m2, m1, m1 + dim)) ? m2 : m1;
            int pmn1, pmx1, pmn2, pmx2;
            double mn1 = cellWrap(minima[3 * id1 + axis], wMn, wMn + dim, pmn1), mx1 = cellWrap(maxima[3 * id1 + axis], wMn, wMn + dim, pmx1);
            double mn2 = cellWrap(minima[3 * id2 + axis], wMn, wMn + dim, pmn2), mx2 = cellWrap(maxima[3 * id2 + axis], wMn, wMn + dim, pmx2);
            if ((pmn1 != pmx1) || (pmn2 != pmx2)) throw std::runtime_error("Particle larger than half of the cell size encountered.");
            periods[axis] = (int)(pmn1 - pmn2);
            if (!(mn1 <= mx2 && mx1 >= mn2)) return false;
        }
        return true;
    }
    bool boundsOverlap(int a, int b) const { int per[3]; return periodic ? spatialOverlapPeri(a, b, per) : spatialOverlap(a, b); }

    Contact* find(int a, int b) { auto it = con.find(key(a, b)); return it == con.end() ? nullptr : &it->second; }

    /* makeContactLater, InsertionSortCollider.hpp:106-113, followed by the functor-order
       swap of ContactLoop.cpp:145-152 / Contact::swapOrder (Contact.cpp:12-16), done here
       once at creation so that contacts are stored in canonical order. */
    void makeContactLater(int id1, int id2, const int cd[3]) {
        Contact c{};
        c.idA = id1; c.idB = id2;
        for (int k = 0; k < 3; k++) c.cd[k] = cd[k];
        c.stepCreated = step; c.minDist00Sq = -1; c.geom = c.phys = false;
        bool swap;
        int sa = shape[id1], sb = shape[id2];
        if (sa == sb) swap = id1 > id2;
        else swap = (sb == ORC_SHAPE_WALL || sb == ORC_SHAPE_FACET) && sa == ORC_SHAPE_SPHERE;
        if (swap) { std::swap(c.idA, c.idB); for (int k = 0; k < 3; k++) c.cd[k] = -c.cd[k]; }
        makeLater[tid()].push_back(c);
    }
    /* handleBoundInversion, :63-98 */
    void handleBoundInversion(int id1, int id2, bool separating) {
        bool overlap = separating ? false : spatialOverlap(id1, id2);
        Contact* C = find(id1, id2);
        bool hasCon = C != nullptr;
        if (!overlap && !hasCon) return;
        if (overlap && hasCon) return;
        if (overlap) { if (!mayCollide(id1, id2)) return; int z[3] = {0, 0, 0}; makeContactLater(id1, id2, z); return; }
        if (!C->isReal()) removeLater[tid()].push_back(key(id1, id2));
    }
    /* handleBoundInversionPeri, :836-874 */
    void handleBoundInversionPeri(int id1, int id2, bool separating) {
        int periods[3] = {0, 0, 0};
        bool overlap = separating ? false : spatialOverlapPeri(id1, id2, periods);
        Contact* C = find(id1, id2);
        bool hasCon = C != nullptr;
        if (!overlap && !hasCon) return;
        if (overlap && hasCon) return;
        if (overlap) { if (!mayCollide(id1, id2)) return; makeContactLater(id1, id2, periods); return; }
        if (!C->isReal()) removeLater[tid()].push_back(key(id1, id2));
    }
    /* insertionSort_part, :162-196: bounds [iBegin, iEnd) sorted from iStart on; a part that
       starts inside its range (iStart != iBegin) stops at the first ordered couple */
    void insertionSortPart(VecBounds& v, bool doCollide, long iBegin, long iEnd, long iStart) {
        const bool earlyStop = iBegin != iStart;
        for (long i = std::max(iStart, iBegin + 1); i < iEnd; i++) {
            if (!(v[i - 1] > v[i])) {
                if (earlyStop) return;
                continue;
            }
            const Bounds viInit = v[i]; long j = i - 1;
            while (j >= iBegin && v[j] > viInit) {
                v[j + 1] = v[j];
                if (viInit.isMin != v[j].isMin && doCollide && viInit.hasBB && v[j].hasBB && viInit.id != v[j].id)
                    handleBoundInversion(viInit.id, v[j].id, !viInit.isMin);
                j--;
            }
            v[j + 1] = viInit;
        }
    }
    /* InsertionSortCollider::insertionSort, :100-158, with the default ompTuneSort (1, 1000, 0)
       (InsertionSortCollider.hpp:263): one chunk per thread, at least 2000 bounds per chunk;
       even passes sort the chunks between splits0, odd passes those between splits1 (half a
       chunk further), each part from its split on, until every inner split is ordered.  Each
       inverted couple is met once, whichever part meets it, so the contacts made and removed
       are those of the serial sort. */
    void insertionSort(VecBounds& v, bool doCollide) {
        const int TH = maxThreads();
        const long minPerCore = 2 * 1000;
        size_t chunks = (size_t)TH;
        if ((long)(v.size / (long)chunks) < minPerCore)
            chunks = (size_t)TH * (size_t)std::ceil(std::max(1L, v.size / minPerCore) * 1. / TH);
        const size_t chunkSize = v.size / chunks;
        if (chunks == 1 || chunkSize < 100) { insertionSortPart(v, doCollide, 0, v.size, 0); return; }
        sortChunks = (long)chunks;
        std::vector<size_t> splits0(chunks + 1), splits1(chunks);
        for (size_t i = 0; i < chunks; i++) { splits0[i] = i * chunkSize; splits1[i] = i * chunkSize + chunkSize / 2; }
        splits0[chunks] = v.size;
        bool isOk = false;
        for (size_t pass = 0; !isOk; pass++) {
            const bool even = (pass % 2) == 0;
            const std::vector<size_t>& s = even ? splits0 : splits1;
            const long nPart = (long)(even ? chunks : chunks - 1);
            #pragma omp parallel for schedule(static)
            for (long chunk = 0; chunk < nPart; chunk++) {
                const size_t start = pass == 0 ? s[chunk] : (even ? splits1[chunk] : splits0[chunk + 1]);
                insertionSortPart(v, doCollide, (long)s[chunk], (long)s[chunk + 1], (long)start);
            }
            isOk = true;
            for (size_t chunk = even ? 1 : 0; chunk < chunks; chunk++) {
                if (v[(long)s[chunk] - 1] > v[(long)s[chunk]]) { isOk = false; break; }
            }
            sortPasses = std::max(sortPasses, (long)pass + 1);
        }
    }
    /* insertionSortPeri_part (serial, partial=false), :718-791 */
    void insertionSortPeri(VecBounds& v, bool doCollide) {
        long& loIdx = v.loIdx;
        for (long _i = 0; _i < v.size; _i++) {
            const long i = v.norm(_i);
            const long i_1 = v.norm(i - 1);
            if (i == loIdx && v[i].coord < 0) { v[i].period -= 1; v[i].coord += v.cellDim; loIdx = v.norm(loIdx + 1); }
            const double iCmpCoord = v[i].coord + (i == loIdx ? v.cellDim : 0);
            if (v[i_1].coord <= iCmpCoord) continue;
            long j = i_1; Bounds vi = v[i];
            const bool viHasBB = vi.hasBB, viIsMin = vi.isMin, viIsInf = vi.isInf;
            while (v[j].coord > vi.coord + (v.norm(j + 1) == loIdx ? v.cellDim : 0)) {
                long j1 = v.norm(j + 1);
                Bounds& vNew = v[j1];
                vNew = v[j];
                if (j == loIdx && vi.coord < 0) { vi.period -= 1; vi.coord += v.cellDim; loIdx = v.norm(loIdx + 1); }
                else if (j1 == loIdx) { vNew.period += 1; vNew.coord -= v.cellDim; loIdx = v.norm(loIdx - 1); }
                if (viIsMin != v[j].isMin && doCollide && viHasBB && v[j].hasBB) {
                    if (vi.id != vNew.id) handleBoundInversionPeri(vi.id, vNew.id, (!viIsMin && !viIsInf && !v[j].isInf));
                }
                j = v.norm(j - 1);
            }
            v[v.norm(j + 1)] = vi;
        }
    }
    /* ContactContainer::removePending with shouldBeRemoved, ContactContainer.hpp:78-93, InsertionSortCollider.hpp:209-214 */
    void removePending() {
        for (uint64_t k : pending) {
            auto it = con.find(k);
            if (it == con.end()) continue;
            if (!boundsOverlap(it->second.idA, it->second.idB)) { con.erase(it); adjDirty = true; }
        }
        pending.clear();
    }
    /* makeRemoveContactLater_process, :20-48 */
    void makeRemoveLaterProcess() {
        for (auto& l : removeLater) {
            for (uint64_t k : l) { if (con.erase(k)) adjDirty = true; }
            l.clear();
        }
        for (auto& l : makeLater) {
            for (const Contact& c : l) {
                uint64_t k = key(c.idA, c.idB);
                if (con.count(k)) continue;
                con.emplace(k, c); adjDirty = true;
            }
            l.clear();
        }
    }

    /* InsertionSortCollider::run, :445-646 */
    void collider() {
        bool fullRun = false;
        if (contactsDirty) { fullRun = true; contactsDirty = false; }
        if ((long)BB[0].size != 2L * np) fullRun = true;
        if (periodic != colPeriodic) { for (int i = 0; i < 3; i++) { BB[i].vec.clear(); BB[i].size = 0; } colPeriodic = periodic; fullRun = true; }
        if ((long)aabb.size() != np) aabb.assign(np, Aabb());
        bool runBboxes = updateBboxes_doFullRun();
        if (runBboxes) fullRun = true;
        if (!fullRun) { removePending(); return; }
        if ((int)makeLater.size() < maxThreads()) { makeLater.resize(maxThreads()); removeLater.resize(maxThreads()); }
        nFullRuns++;
        long nPar = np;
        bool doInitSort = false;
        if (BB[0].size != 2 * nPar) {
            if (2 * nPar < BB[0].size) { for (int i = 0; i < 3; i++) { BB[i].vec.clear(); BB[i].size = 0; } }
            else if (2 * nPar - BB[0].size > 200 || BB[0].size == 0) doInitSort = true;
            for (int i = 0; i < 3; i++) {
                BB[i].axis = i;
                for (long id = (long)BB[i].vec.size() / 2; id < nPar; id++) {
                    Bounds b0{}; b0.coord = 0; b0.id = (int)id; b0.period = 0; b0.isMin = 1;
                    Bounds b1{}; b1.coord = 0; b1.id = (int)id; b1.period = 0; b1.isMin = 0;
                    BB[i].vec.push_back(b0); BB[i].vec.push_back(b1);
                }
                BB[i].size = (long)BB[i].vec.size();
            }
        }
        minima.assign(3 * nPar, 0.); maxima.assign(3 * nPar, 0.);
        if (periodic) for (int i = 0; i < 3; i++) BB[i].cellDim = cellSize[i];
        // copy bounds, :514-548
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < 2 * nPar; i++) {
            for (int j = 0; j < 3; j++) {
                VecBounds& BBj = BB[j];
                Bounds& b = BBj[i];
                const int id = b.id;
                const Aabb& a = aabb[id];
                if (shape[id] != ORC_SHAPE_NONE) {
                    b.hasBB = a.has;
                    b.coord = (a.has ? (b.isMin ? a.mn[j] : a.mx[j]) : pos[nodeOf(id)][j]) - (periodic ? BBj.cellDim * b.period : 0.);
                    b.isInf = false;
                    if (periodic && std::isinf(b.coord)) { b.period = (b.coord < 0 ? 0 : 1); b.coord = 0.; b.isInf = true; }
                } else {  // vanished particle (:534-541): no bound; -inf (0 if periodic) at an initial sort, else kept
                    b.hasBB = false;
                    if (doInitSort) b.coord = periodic ? 0. : -INF;
                }
                if (doInitSort && periodic) { if (!b.isInf) { int per; b.coord = cellWrap(b.coord, 0, BBj.cellDim, per); b.period = per; } }
            }
        }
        // copy minima/maxima, :552-560
        for (int id = 0; id < nPar; id++) {
            const Aabb& a = aabb[id];
            if (a.has) { for (int d = 0; d < 3; d++) { minima[3 * id + d] = a.mn[d]; maxima[3 * id + d] = a.mx[d]; } }
            else if (shape[id] != ORC_SHAPE_NONE) { for (int d = 0; d < 3; d++) minima[3 * id + d] = maxima[3 * id + d] = pos[nodeOf(id)][d]; }
        }
        removePending();
        if (!doInitSort) {
            if (!periodic) for (int i = 0; i < 3; i++) insertionSort(BB[i], true);
            else for (int i = 0; i < 3; i++) insertionSortPeri(BB[i], true);
        } else {
            for (int i = 0; i < 3; i++) { BB[i].loIdx = 0; std::sort(BB[i].vec.begin(), BB[i].vec.end()); }
            VecBounds& V = BB[0];  // sortAxis=0
            if (!periodic) {
                for (long i = 0; i < 2 * nPar; i++) {
                    if (!(V[i].isMin && V[i].hasBB)) continue;
                    const int iid = V[i].id;
                    for (long j = i + 1; j < 2 * nPar && V[j].id != iid; j++) {
                        if (!V[j].isMin) continue;
                        handleBoundInversion(iid, V[j].id, false);
                    }
                }
            } else {
                for (long i = 0; i < 2 * nPar; i++) {
                    if (!(V[i].isMin && V[i].hasBB)) continue;
                    const int iid = V[i].id;
                    long cnt = 0;
                    for (long j = V.norm(i + 1); V[j].id != iid; j = V.norm(j + 1)) {
                        if (!V[j].isMin) continue;
                        handleBoundInversionPeri(iid, V[j].id, false);
                        if (cnt++ > 2 * nPar) throw std::logic_error("loop in the initial sort");
                    }
                }
            }
        }
        makeRemoveLaterProcess();
    }

    /* ---------- L6Geom, pkg/dem/L6Geom.cpp:27-217 ---------- */
    static void setInitialLocalCoords(Contact& C, const V3& locX) {  // :27-38
        V3 locY = vcross(locX, std::fabs(locX[1]) < std::fabs(locX[2]) ? v3(0, 1, 0) : v3(0, 0, 1));
        locY = vsub(locY, vmul(locX, vdot(locY, locX)));
        locY = vnormalized(locY);
        V3 locZ = vcross(locX, locY);
        msetcol(C.trsf, 0, locX); msetcol(C.trsf, 1, locY); msetcol(C.trsf, 2, locZ);
    }
    void handleSpheresLikeContact(Contact& C, const V3& pos1, const V3& vel1, const V3& angVel1, const V3& pos2, const V3& vel2,
                                  const V3& angVel2, const V3& normal, const V3& contPt, double uN, double r1, double r2) {
        if (!C.geom) {  // :53-72
            C.geom = true;
            C.trsf = meye();
            setInitialLocalCoords(C, normal);
            C.uN = uN;
            C.lens[0] = std::fabs(r1) + uN / 2; C.lens[1] = std::fabs(r2) + uN / 2;
            if (C.lens[0] < 0) C.lens[0] = C.lens[1];
            if (C.lens[1] < 0) C.lens[1] = C.lens[0];
            double rr = (r1 > 0 && r2 > 0) ? std::min(r1, r2) : (r1 > 0 ? r1 : r2);
            C.contA = PI * (rr * rr);
            C.contPt = contPt;
            C.vel = vzero(); C.angVel = vzero();
            return;
        }
        const double dt = P.dt;
        const V3 currNormal = normal; const V3 prevNormal = mcol(C.trsf, 0);
        const V3 prevContPt = C.contPt;
        V3 shiftVel2 = periodic ? intrShiftVel(C.cd) : vzero();
        V3 normRotVec = vcross(prevNormal, currNormal);
        V3 midNormal = vmul(vadd(prevNormal, currNormal), .5);
        if (vnorm(midNormal) == 0) midNormal = currNormal;
        else midNormal = vnormalized(midNormal);
        V3 normTwistVec = vmul(vmul(vmul(midNormal, dt), .5), vdot(midNormal, vadd(angVel1, angVel2)));
        const M3 prevTrsf = C.trsf;
        M3 midTrsf;
        V3 rt = vadd(normRotVec, normTwistVec);
        msetcol(midTrsf, 0, midNormal);
        msetcol(midTrsf, 1, vsub(mcol(prevTrsf, 1), vdiv(vcross(mcol(prevTrsf, 1), rt), 2.)));
        msetcol(midTrsf, 2, vcross(mcol(midTrsf, 0), mcol(midTrsf, 1)));
        M3 currTrsf;
        msetcol(currTrsf, 0, currNormal);
        msetcol(currTrsf, 1, vsub(mcol(prevTrsf, 1), vcross(mcol(midTrsf, 1), rt)));
        msetcol(currTrsf, 2, vcross(mcol(currTrsf, 0), mcol(currTrsf, 1)));
        if (P.trsfRenorm > 0 && (step % P.trsfRenorm) == 0) {
            V3 c0 = vnormalized(mcol(currTrsf, 0));
            V3 c1 = mcol(currTrsf, 1);
            c1 = vsub(c1, vmul(c0, vdot(c1, c0)));
            c1 = vnormalized(c1);
            msetcol(currTrsf, 0, c0); msetcol(currTrsf, 1, c1); msetcol(currTrsf, 2, vcross(c0, c1));
        }
        V3 midContPt = vmul(vadd(prevContPt, contPt), .5);
        V3 midPos1 = vsub(pos1, vmul(vel1, dt / 2.));
        V3 midPos2 = vsub(pos2, vmul(vadd(vel2, shiftVel2), dt / 2.));
        V3 c1x = vsub(midContPt, midPos1);
        V3 c2x = vsub(midContPt, midPos2);
        V3 relVel = vsub(vadd(vadd(vel2, shiftVel2), vcross(angVel2, c2x)), vadd(vel1, vcross(angVel1, c1x)));
        C.trsf = currTrsf;
        C.vel = mtvmul(midTrsf, relVel);
        C.angVel = mtvmul(midTrsf, vsub(angVel2, angVel1));
        C.uN = uN;
        C.contPt = contPt;
    }
    /* Cg2_Sphere_Sphere_L6Geom::go, pkg/dem/Sphere.cpp:88-109 */
    bool cg2SphereSphere(Contact& C, const V3& shift2) {
        int a = C.idA, b = C.idB;
        double r1 = radius[a], r2 = radius[b];
        int na = nodeOf(a), nb = nodeOf(b);
        V3 p2 = vadd(pos[nb], shift2);
        V3 relPos = vsub(p2, pos[na]);
        double df = std::fabs(P.distFactor) * (r1 + r2);
        double unDistSq = vsqn(relPos) - df * df;
        if (unDistSq > 0 && !C.isReal()) return false;
        double dist = vnorm(relPos);
        double uN = dist - (r1 + r2);
        V3 normal = vdiv(relPos, dist);
        V3 contPt = vadd(pos[na], vmul(normal, r1 + 0.5 * uN));
        handleSpheresLikeContact(C, pos[na], vel[na], angVel[na], p2, vel[nb], angVel[nb], normal, contPt, uN, r1, r2);
        return true;
    }
    /* Cg2_Wall_Sphere_L6Geom::go, pkg/dem/Wall.cpp:45-79 */
    bool cg2WallSphere(Contact& C, const V3& shift2) {
        int w = C.idA, s = C.idB;
        double rad = radius[s]; int ax = axis[w]; int sn = sense[w];
        const V3& wallPos = pos[nodeOf(w)];
        V3 spherePos = vadd(pos[nodeOf(s)], shift2);
        double dist = spherePos[ax] - wallPos[ax];
        if (!C.isReal() && std::fabs(dist) > rad) return false;
        V3 contPt = spherePos; contPt[ax] = wallPos[ax];
        V3 normal = vzero();
        if (sn == 0) {
            if (!C.geom) normal[ax] = dist > 0 ? 1. : -1.;
            else normal[ax] = C.trsf.m[ax][0];
        } else normal[ax] = (sn == 1 ? 1. : -1);
        double uN = normal[ax] * dist - rad;
        handleSpheresLikeContact(C, wallPos, vel[nodeOf(w)], angVel[nodeOf(w)], spherePos, vel[nodeOf(s)], angVel[nodeOf(s)], normal, contPt, uN, -rad, rad);
        return true;
    }
    /* CompUtils::closestSegmentPt, lib/base/CompUtils.cpp:57-63 */
    static V3 closestSegmentPt(const V3& Pt, const V3& A, const V3& B) {
        V3 BA = vsub(B, A);
        if (vsqn(BA) == 0.) return A;
        double u = (vdot(Pt, BA) - vdot(A, BA)) / vsqn(BA);
        return vadd(A, vmul(BA, std::min(1., std::max(0., u))));
    }
    /* CompUtils::triangleBarycentrics, lib/base/CompUtils.cpp:92-110 (2x2 inverse as
       adjugate times 1/det, Eigen 3's compute_inverse for 2x2) */
    static V3 triangleBarycentrics(const V3& x, const V3& A, const V3& B, const V3& Cc) {
        V3 vB = vsub(B, A), vC = vsub(Cc, A), vX = vsub(x, A);
        double m00 = vdot(vB, vB), m01 = vdot(vC, vB), m10 = vdot(vB, vC), m11 = vdot(vC, vC);
        double invdet = 1. / (m00 * m11 - m10 * m01);
        double i00 = m11 * invdet, i01 = -m01 * invdet, i10 = -m10 * invdet, i11 = m00 * invdet;
        double r0 = vdot(vX, vB), r1 = vdot(vX, vC);
        double u = i00 * r0 + i01 * r1, v = i10 * r0 + i11 * r1;
        return v3(1 - u - v, u, v);
    }
    /* Cg2_Facet_Sphere_L6Geom::go, pkg/dem/Facet.cpp:235-288; getNormal :71-74; getOuterVectors :110-115;
       interpolatePtLinAngVel :158-174 (fakeVel==0) */
    bool cg2FacetSphere(Contact& C, const V3& shift2) {
        int f = C.idA, s = C.idB;
        const V3& n0 = pos[nodeOf(f, 0)]; const V3& n1 = pos[nodeOf(f, 1)]; const V3& n2 = pos[nodeOf(f, 2)];
        double halfThick = radius[f], rad = radius[s];
        V3 sC = vadd(pos[nodeOf(s)], shift2);
        V3 fNormal = vnormalized(vcross(vsub(n1, n0), vsub(n2, n0)));
        double planeDist = vdot(vsub(sC, n0), fNormal);
        if (std::fabs(planeDist) > (rad + halfThick) && !C.isReal()) return false;
        V3 fC = vsub(sC, vmul(fNormal, planeDist));
        V3 nn = vcross(vsub(n1, n0), vsub(n2, n0));
        V3 outVec[3] = {vcross(vsub(n1, n0), nn), vcross(vsub(n2, n1), nn), vcross(vsub(n0, n2), nn)};
        const V3* nd[3] = {&n0, &n1, &n2};
        double ll[3];
        for (int i = 0; i < 3; i++) ll[i] = vdot(outVec[i], vsub(fC, *nd[i]));
        int w = (ll[0] > 0 ? 1 : 0) + (ll[1] > 0 ? 2 : 0) + (ll[2] > 0 ? 4 : 0);
        V3 contPt;
        switch (w) {
            case 0: contPt = fC; break;
            case 1: contPt = closestSegmentPt(fC, n0, n1); break;
            case 2: contPt = closestSegmentPt(fC, n1, n2); break;
            case 4: contPt = closestSegmentPt(fC, n2, n0); break;
            case 3: contPt = n1; break;
            case 5: contPt = n0; break;
            case 6: contPt = n2; break;
            default: throw std::logic_error("Impossible sphere-facet intersection");
        }
        V3 normal = vsub(sC, contPt);
        double rs = rad + halfThick;
        if (vsqn(normal) > rs * rs && !C.isReal()) return false;
        double dist = vnorm(normal);
        normal = vdiv(normal, dist);
        if (halfThick > 0) contPt = vadd(contPt, vmul(normal, halfThick));
        double uN = dist - rad - halfThick;
        V3 bc = triangleBarycentrics(contPt, n0, n1, n2);
        const V3& v0 = vel[nodeOf(f, 0)]; const V3& v1 = vel[nodeOf(f, 1)]; const V3& v2 = vel[nodeOf(f, 2)];
        V3 linVel = vadd(vadd(vmul(v0, bc[0]), vmul(v1, bc[1])), vmul(v2, bc[2]));
        V3 av = vadd(vadd(vcross(vsub(n0, contPt), v0), vcross(vsub(n1, contPt), v1)), vcross(vsub(n2, contPt), v2));
        handleSpheresLikeContact(C, contPt, linVel, av, sC, vel[nodeOf(s)], angVel[nodeOf(s)], normal, contPt, uN, std::max(halfThick, rad), rad);
        return true;
    }

    /* Cp2_FrictMat_FrictPhys::updateFrictPhys, pkg/dem/FrictMat.cpp:21-44 */
    void cp2Frict(Contact& C) {
        const Mat& m1 = mats[matId[C.idA]]; const Mat& m2 = mats[matId[C.idB]];
        double l0 = C.lens[0], l1 = C.lens[1], A = C.contA;
        C.kn = 1 / (1 / (m1.young * A / l0) + 1 / (m2.young * A / l1));
        C.kt = .5 * (m1.ktDivKn + m2.ktDivKn) * C.kn;
        C.tanPhi = std::min(m1.tanPhi, m2.tanPhi);
        C.force = vzero(); C.torque = vzero();
        C.phys = true;
    }
    /* Cp2_FrictMat_HertzPhys::go, pkg/dem/Hertz.cpp:9-50 */
    void cp2Hertz(Contact& C) {
        const Mat& m1 = mats[matId[C.idA]]; const Mat& m2 = mats[matId[C.idB]];
        double r1 = C.lens[0], r2 = C.lens[1];
        double E1 = m1.young, E2 = m2.young, nu1 = P.poisson, nu2 = P.poisson;
        C.K = (4 / 3.) * 1. / ((1 - nu1 * nu1) / E1 + (1 - nu2 * nu2) / E2);
        C.R = 1. / (1. / r1 + 1. / r2);
        double kn0 = C.K * std::sqrt(C.R);
        C.gamma = P.gamma; C.alpha = P.alpha;
        double G1 = E1 / (2 * (1 + nu1)), G2 = E2 / (2 * (1 + nu2));
        double G = .5 * (G1 + G2);
        double nu = .5 * (nu1 + nu2);
        C.kt0 = 2 * std::sqrt(4 * C.R) * G / (2 - nu);
        C.tanPhi = std::min(m1.tanPhi, m2.tanPhi);
        double en = P.en;
        if (en > 0 && en < 1.) {
            double ma = mass[nodeOf(C.idA)], mb = mass[nodeOf(C.idB)];
            double mbar = (ma <= 0 && mb > 0) ? mb : ((ma > 0 && mb <= 0) ? ma : (ma * mb) / (ma + mb));
            double le = std::log(en);
            double viscAlpha = -std::sqrt(5) * le / (std::sqrt(le * le + PI * PI));
            C.alpha_sqrtMK = std::max(0., viscAlpha * std::sqrt(mbar * kn0));
        } else C.alpha_sqrtMK = 0.0;
        C.kn = 0; C.kt = 0;
        C.force = vzero(); C.torque = vzero();
        C.phys = true;
    }
    /* Law2_L6Geom_FrictPhys_IdealElPl::go, pkg/dem/IdealElPl.cpp:12-122 (iniEqlb=false, relRollStiff=0) */
    bool lawIdealElPl(Contact& C, double* e) {
        double uN = C.uN;
        const double dt = P.dt;
        if (uN > 0 && !P.noBreak) {
            if (P.noSlip && P.trackEnergy) {
                double Fn = C.kn * uN;
                double Ft0 = C.force[1] + dt * C.kt * C.vel[1], Ft1 = C.force[2] + dt * C.kt * C.vel[2];
                e[ORC_E_BROKEN] += .5 * (Fn * Fn / C.kn + (Ft0 * Ft0 + Ft1 * Ft1) / C.kt);
                e[ORC_E_N + ORC_E_BROKEN] = 1;
            }
            return false;
        }
        C.torque = vzero();
        C.force[0] = C.kn * uN;
        if (P.noFrict || C.tanPhi == 0.) { C.force[1] = 0; C.force[2] = 0; }
        else {
            C.force[1] += dt * C.kt * C.vel[1];
            C.force[2] += dt * C.kt * C.vel[2];
            double maxFt = std::fabs(C.force[0]) * C.tanPhi;
            double ftsq = C.force[1] * C.force[1] + C.force[2] * C.force[2];
            if (ftsq > maxFt * maxFt && !P.noSlip) {
                double FtNorm = std::sqrt(ftsq);
                double ratio = maxFt / FtNorm;
                if (FtNorm == 0.) ratio = 0.;
                if (P.trackEnergy) {
                    double dissip = (C.kt != 0. ? (maxFt) * (FtNorm - maxFt) / C.kt : 0.);
                    if (dissip != 0.) { e[ORC_E_PLAST] += dissip; e[ORC_E_N + ORC_E_PLAST] = 1; }
                }
                C.force[1] *= ratio; C.force[2] *= ratio;
            }
            if (std::isnan(C.force[0]) || std::isnan(C.force[1]) || std::isnan(C.force[2]))
                throw std::runtime_error("NaN force in contact");
        }
        if (P.trackEnergy) {
            double ftsq = C.force[1] * C.force[1] + C.force[2] * C.force[2];
            double elast = 0.5 * ((C.force[0] * C.force[0]) / C.kn + (C.kt != 0. ? ftsq / C.kt : 0.));
            if (std::isnan(elast)) elast = 0.;
            e[ORC_E_ELAST] += elast; e[ORC_E_N + ORC_E_ELAST] = 1;
        }
        return true;
    }
    /* Law2_L6Geom_HertzPhys_DMT::go, pkg/dem/Hertz.cpp:58-182 (alpha==0 branch) */
    bool lawHertz(Contact& C, double* e) {
        const double dt = P.dt;
        C.torque = vzero();
        const double velN = C.vel[0];
        double kn0 = C.K * std::sqrt(C.R);
        C.kn = (3 / 2.) * kn0 * std::sqrt(-C.uN);
        if (C.uN > 0) return false;
        double s = std::sqrt(-C.uN);
        double Fne = -kn0 * ((s * s) * s);
        double Fna = 4 * PI * C.R * C.gamma;
        double eta = (C.alpha_sqrtMK > 0 ? C.alpha_sqrtMK * std::sqrt(std::sqrt(std::max(-C.uN, 0.))) : 0.);
        double Fnv = eta * velN;
        if (C.alpha == 0. && P.noAttraction && Fne + Fnv > 0) Fnv = -Fne;
        double Fn = Fne + Fna + Fnv;
        C.force[0] = Fn;
        if (P.trackEnergy) { double v = Fnv * velN * dt; if (v != 0. || e[ORC_E_N + ORC_E_VISCN]) { e[ORC_E_VISCN] += v; e[ORC_E_N + ORC_E_VISCN] = 1; } }
        C.kt = C.kt0 * std::sqrt(C.uN < 0 ? -C.uN : 0);
        double Ft0 = dt * C.kt * C.vel[1], Ft1 = dt * C.kt * C.vel[2];
        double maxFt = std::max(0., Fn) * C.tanPhi;
        double ftsq = Ft0 * Ft0 + Ft1 * Ft1;
        if (ftsq > maxFt * maxFt) {
            double FtNorm = std::sqrt(ftsq);
            double ratio = maxFt / FtNorm;
            if (P.trackEnergy) { double v = (.5 * (FtNorm - maxFt) + maxFt) * (FtNorm - maxFt) / C.kt; if (v != 0.) { e[ORC_E_PLAST] += v; e[ORC_E_N + ORC_E_PLAST] = 1; } }
            Ft0 *= ratio; Ft1 *= ratio;
        } else {
            Ft0 += eta * C.vel[1]; Ft1 += eta * C.vel[2];
            if (P.trackEnergy) { double v = eta * (C.vel[1] * C.vel[1] + C.vel[2] * C.vel[2]) * dt; if (v != 0.) { e[ORC_E_VISCT] += v; e[ORC_E_N + ORC_E_VISCT] = 1; } }
        }
        C.force[1] = Ft0; C.force[2] = Ft1;
        if (P.trackEnergy && C.uN < 0) {
            double s5 = (((s * s) * s) * s) * s;
            double el = kn0 * (2 / 5.) * s5 + 0.5 * (Ft0 * Ft0 + Ft1 * Ft1) / C.kt;
            if (!std::isnan(el)) e[ORC_E_ELAST] += el;
            e[ORC_E_N + ORC_E_ELAST] = 1;
        }
        return true;
    }

    void rebuildAdjacency() {
        lin.clear(); lin.reserve(con.size());
        for (auto& kv : con) lin.push_back(&kv.second);
        adjOff.assign(np + 1, 0);
        for (Contact* c : lin) { adjOff[c->idA + 1]++; adjOff[c->idB + 1]++; }
        for (int p = 0; p < np; p++) adjOff[p + 1] += adjOff[p];
        adjC.assign(adjOff[np], nullptr);
        std::vector<long> fill(adjOff.begin(), adjOff.end() - 1);
        // contacts in (min,max) order: each particle receives them in ascending order of the other id
        for (Contact* c : lin) {
            int lo = std::min(c->idA, c->idB), hi = std::max(c->idA, c->idB);
            adjC[fill[hi]++] = c;
            (void)lo;
        }
        for (Contact* c : lin) {
            int lo = std::min(c->idA, c->idB);
            adjC[fill[lo]++] = c;
        }
        // per particle: entries where p is max come first (other<p, ascending), then where p is min (other>p, ascending)
        adjDirty = false;
    }

    /* ContactLoop::run, pkg/dem/ContactLoop.cpp:93-259, with deterministic force application */
    void contactLoop() {
        pending.clear();  // removeAllPending: the collider already processed them
        if (adjDirty) rebuildAdjacency();
        const long nc = (long)lin.size();
        std::vector<double> ebuf((size_t)nc * 2 * ORC_E_N, 0.);
        std::vector<char> broke(nc, 0);
        std::string perr;
        #pragma omp parallel for schedule(guided)
        for (long i = 0; i < nc; i++) {
            try {
                Contact& C = *lin[i];
                double* e = &ebuf[(size_t)i * 2 * ORC_E_N];
                int sa = shape[C.idA], sb = shape[C.idB];
                bool fresh = (C.stepCreated == step);
                if (!C.isReal() && fresh) {
                    if (sa == ORC_SHAPE_SPHERE && sb == ORC_SHAPE_SPHERE) {
                        double x = std::fabs(P.distFactor * (radius[C.idA] + radius[C.idB]));
                        C.minDist00Sq = x * x;
                    } else C.minDist00Sq = -1;
                }
                V3 shift2 = periodic ? intrShiftPos(C.cd) : vzero();
                if (P.dist00 && !C.isReal() && !fresh && C.minDist00Sq > 0 &&
                    vsqn(vsub(pos[nodeOf(C.idA)], vadd(pos[nodeOf(C.idB)], shift2))) > C.minDist00Sq) continue;
                bool geomCreated;
                if (sa == ORC_SHAPE_SPHERE && sb == ORC_SHAPE_SPHERE) geomCreated = cg2SphereSphere(C, shift2);
                else if (sa == ORC_SHAPE_WALL && sb == ORC_SHAPE_SPHERE) geomCreated = cg2WallSphere(C, shift2);
                else if (sa == ORC_SHAPE_FACET && sb == ORC_SHAPE_SPHERE) geomCreated = cg2FacetSphere(C, shift2);
                else geomCreated = false;
                if (!geomCreated) continue;
                if (!C.phys) C.stepCreated = step;
                if (!C.phys) { if (P.law == ORC_LAW_HERTZ_DMT) cp2Hertz(C); else cp2Frict(C); }
                bool keep = (P.law == ORC_LAW_HERTZ_DMT) ? lawHertz(C, e) : lawIdealElPl(C, e);
                if (!keep) { C.geom = C.phys = false; broke[i] = 1; }  // requestRemoval: Contact::reset, ContactContainer.cpp:164-171
            } catch (std::exception& ex) {
                #pragma omp critical
                perr = ex.what();
            }
        }
        if (!perr.empty()) throw std::runtime_error(perr);
        for (long i = 0; i < nc; i++) if (broke[i]) pending.push_back(key(lin[i]->idA, lin[i]->idB));
        if (P.trackEnergy) {
            for (long i = 0; i < nc; i++) {
                const double* e = &ebuf[(size_t)i * 2 * ORC_E_N];
                for (int k = 0; k < ORC_E_N; k++) if (e[ORC_E_N + k]) eAdd(k, e[k], k != ORC_E_ELAST && k != ORC_E_BROKEN);
            }
        }
        // deterministic force application (applyForceUninodal, ContactLoop.cpp:261-267; Contact::getForceTorqueBranch, Contact.cpp:24-34)
        #pragma omp parallel for schedule(guided)
        for (int p = 0; p < np; p++) {
            if (numNodes(p) != 1) continue;
            int nd = nodeOf(p);
            V3 F0 = force[nd], T0 = torque[nd];
            for (long k = adjOff[p]; k < adjOff[p + 1]; k++) {
                const Contact& C = *adjC[k];
                if (!C.isReal()) continue;
                bool isPA = (C.idA == p);
                double sign = isPA ? 1. : -1.;
                V3 F = vmul(mvmul(C.trsf, C.force), sign);
                V3 xc = vsub(C.contPt, (!isPA && periodic) ? vadd(pos[nd], intrShiftPos(C.cd)) : pos[nd]);
                F0 = vadd(F0, F);
                T0 = vadd(T0, vcross(xc, F));
            }
            force[nd] = F0; torque[nd] = T0;
        }
    }

    /* ---------- particles removed and inserted in place ---------- */
    /* ParticleContainer ids (pkg/dem/ParticleContainer.cpp:23-72): holes reused first, in the
       order they were freed (freeIds is a FIFO list), then the id past the last one in use */
    int parEnd = 0;
    std::deque<int> freeIds;
    void initIds() {
        parEnd = 0;
        for (int p = 0; p < np; p++) if (shape[p] != ORC_SHAPE_NONE) parEnd = p + 1;
        freeIds.clear();
        for (int p = 0; p < parEnd; p++) if (shape[p] == ORC_SHAPE_NONE) freeIds.push_back(p);
    }
    /* DemField::removeParticle, pkg/dem/Particle.cpp:445-493: the particle's contacts are
       removed (ContactContainer::remove); the slot stays as a hole whose node is fully
       blocked, massless, at rest and skipped by the energy terms, so Leapfrog changes nothing
       for it (the reference removes the node from DemField.nodes) */
    void removeParticle(int id) {
        for (auto it = con.begin(); it != con.end();) {
            if (it->second.idA == id || it->second.idB == id) it = con.erase(it);
            else ++it;
        }
        pending.erase(std::remove_if(pending.begin(), pending.end(), [&](uint64_t k) { return (int)(k >> 32) == id || (int)(k & 0xffffffffu) == id; }),
                      pending.end());
        shape[id] = ORC_SHAPE_NONE;
        const int nd = nodeOf(id);
        flags[nd] = DOF_ALL | ENERGY_SKIP | GRAVITY_SKIP | DAMPING_SKIP;
        vel[nd] = vzero(); angVel[nd] = vzero(); mass[nd] = 0; inertia[nd] = vzero();
        aabb[id] = Aabb();
        freeIds.push_back(id);
        adjDirty = true;
        contactsDirty = true;  // the particle count changed: full run (InsertionSortCollider.cpp:475-492)
    }
    /* ParticleContainer::insert of a new sphere: its node starts with zero force and torque,
       as a new DemData */
    int insertSphere(const V3& p, const V3& v, const V3& w, double r, double m, double I, int mat, uint32_t msk) {
        int id;
        if (!freeIds.empty()) { id = freeIds.front(); freeIds.pop_front(); }
        else id = parEnd++;
        shape[id] = ORC_SHAPE_SPHERE;
        nodeIx[3 * id] = id; nodeIx[3 * id + 1] = -1; nodeIx[3 * id + 2] = -1;
        radius[id] = r; axis[id] = 0; sense[id] = 0; mask[id] = msk; matId[id] = mat;
        pos[id] = p; vel[id] = v; angVel[id] = w;
        ori[id] = Qt{1, 0, 0, 0};
        mass[id] = m; inertia[id] = v3(I, I, I);
        flags[id] = 0;
        force[id] = vzero(); torque[id] = vzero();
        aabb[id] = Aabb();
        adjDirty = true;
        contactsDirty = true;
        return id;
    }

    /* ---------- DynDt, pkg/dem/DynDt.cpp:16-97 ---------- */
    orc_dyndt dd{};
    double time = 0, nextDt = NAN;
    /* inner loop of DynDt::nodalStiffAdd (:19-34) over the contacts of particle p, in
       contact-map order (other id ascending); the contact node's orientation is
       Quaternionr(trsf) (L6Geom.cpp:69,216), which rotates the x axis to trsf column 0 */
    void stiffAddParticle(int p, V3& kt, V3& kr) {
        if (adjDirty) rebuildAdjacency();
        for (long k = adjOff[p]; k < adjOff[p + 1]; k++) {
            const Contact& C = *adjC[k];
            if (!C.isReal()) continue;
            const int ix = (C.idA == p) ? 0 : 1;
            const V3 n = mcol(C.trsf, 0);
            const V3 n2 = v3(n[0] * n[0], n[1] * n[1], n[2] * n[2]);
            for (int a = 0; a < 3; a++) kt[a] = kt[a] + (n2[a] * (C.kn - C.kt) + C.kt);
            const double lk = (C.lens[ix] * C.lens[ix]) * C.kt;
            kr[0] = kr[0] + lk * (n2[1] + n2[2]);
            kr[1] = kr[1] + lk * (n2[2] + n2[0]);
            kr[2] = kr[2] + lk * (n2[0] + n2[1]);
        }
    }
    bool uninodalAt(int i) const { return i < np && (shape[i] == ORC_SHAPE_SPHERE || shape[i] == ORC_SHAPE_WALL) && nodeOf(i) == i; }
    /* DynDt::nodalCritDtSq, :39-58 */
    double nodalCritDtSq(int i) {
        const uint32_t fl = flags[i];
        if ((fl & DOF_ALL) == DOF_ALL) return INF;
        V3 kt = vzero(), kr = vzero();
        if (uninodalAt(i)) stiffAddParticle(i, kt, kr);
        if (fl & CLUMP_CLUMP) {
            const int c = nodeClump[i];
            for (int k = clumpOff[c]; k < clumpOff[c + 1]; k++)
                if (uninodalAt(clumpMem[k])) stiffAddParticle(clumpMem[k], kt, kr);
        }
        double ret = INF;
        for (int a = 0; a < 3; a++) if (kt[a] != 0 && mass[i] > 0. && !(fl & axisDOF(a, false))) ret = std::min(ret, mass[i] / std::abs(kt[a]));
        for (int a = 0; a < 3; a++) if (kr[a] != 0 && inertia[i][a] > 0. && !(fl & axisDOF(a, true))) ret = std::min(ret, inertia[i][a] / std::abs(kr[a]));
        return 2 * ret;
    }
    /* DynDt::critDt_stiffness, :61-70, over DemField.nodes (clump members are not among
       them; facet vertex nodes must be blocked or clumped: their facets' contacts are not
       restated) */
    double critDtStiffness() {
        std::vector<char> facetNode(nn, 0);
        for (int p = 0; p < np; p++)
            if (shape[p] == ORC_SHAPE_FACET) for (int k = 0; k < 3; k++) facetNode[nodeOf(p, k)] = 1;
        double ret = INF;
        for (int i = 0; i < nn; i++) {
            if (flags[i] & CLUMP_CLUMPED) {
                if (facetNode[i] && (flags[nodeMaster[i]] & DOF_ALL) != DOF_ALL) throw std::runtime_error("DynDt: facet nodes with free DOFs are not restated");
                continue;
            }
            if (i < np && !uninodalAt(i)) continue;  // no node of its own (facet particles)
            if (facetNode[i] && (flags[i] & DOF_ALL) != DOF_ALL) throw std::runtime_error("DynDt: facet nodes with free DOFs are not restated");
            ret = std::min(ret, nodalCritDtSq(i));
        }
        return std::sqrt(ret);
    }
    /* PeriodicEngine::isActivated, core/Engine.cpp:169-195 (no wall-clock criterion) */
    bool ddActivate() {
        const long stepNow = step;
        const double virtNow = time;
        const bool initNow = dd.stepLast < 0;
        if (initNow) { dd.virtLast = virtNow; dd.stepLast = stepNow; }
        if ((dd.nDo < 0 || dd.nDone < dd.nDo) &&
            ((dd.virtPeriod > 0 && virtNow - dd.virtLast >= dd.virtPeriod) ||
             (dd.stepPeriod > 0 && ((stepNow - dd.stepLast >= dd.stepPeriod) || (dd.stepModulo && stepNow % dd.stepPeriod == 0))))) {
            dd.virtPrev = dd.virtLast; dd.virtLast = virtNow;
            dd.stepPrev = dd.stepLast; dd.stepLast = stepNow;
            if (initNow && !dd.initRun) return false;
            dd.nDone++;
            return true;
        }
        return false;
    }
    /* DynDt::run, :72-97 */
    void dynDtRun() {
        const double crDt = critDtStiffness();
        if (std::isinf(crDt)) return;
        const long nSteps = step - dd.stepPrev;
        const double maxNewDt = P.dt * std::pow(1. + dd.maxRelInc, nSteps);
        const double nd = std::min(crDt * dd.dtSafety, maxNewDt);
        if (!dd.dryRun) {
            if (nd != P.dt) { dd.dt = NAN; nextDt = nd; }
        } else dd.dt = nd;
    }

    /* Scene::doOneStep, core/Scene.cpp:319-424 */
    void doOneStep() {
        if (periodic) cellIntegrateAndUpdate(P.dt);
        if (P.trackEnergy) { E[ORC_E_KINETIC] = 0; E[ORC_E_ELAST] = 0; }  // resetResettables
        leapfrog();
        collider();
        contactLoop();
        if (dd.enabled && ddActivate()) dynDtRun();
        if (periodic) cellSetNextGradV();
        step++;
        time += P.dt;
        if (!std::isnan(nextDt)) { P.dt = nextDt; nextDt = NAN; }
    }
};

Oracle* O(void* h) { return static_cast<Oracle*>(h); }

/* ---------- SpherePack::makeCloud, lib/sphere-pack/SpherePack.cpp:135-248 (no hSize) ----------
   The generator: boost::minstd_rand (x <- 48271 x mod 2^31-1) through boost::uniform_real(0,1),
   u = (x - 1)/(2^31 - 3), redrawn when it reaches 1 (Boost's pre-1.47 uniform_real; the
   reference pins no Boost version). */
struct CloudGen {
    uint32_t x;
    double rnd() {
        for (;;) {
            x = (uint32_t)(((uint64_t)x * 48271u) % 2147483647u);
            double u = (double)(x - 1u) / 2147483645.;
            if (u < 1.) return u;
        }
    }
};
struct Sph { V3 c; double r; };
double cloudPow3Interp(double x, double a, double b) { return std::pow(x * (std::pow(b, -2) - std::pow(a, -2)) + std::pow(a, -2), -1. / 2); }  // SpherePack.hpp:86
int cloudGetPiece(double x, const std::vector<double>& cumm, double& norm) {  // SpherePack.cpp:283-291
    int sz = (int)cumm.size(), i = 0;
    while (i < sz && cumm[i] <= x) i++;
    if ((i == sz - 1) && cumm[i] <= x) { i = sz - 2; norm = 1.; return i; }
    i--;
    norm = (x - cumm[i]) / (cumm[i + 1] - cumm[i]);
    return i;
}
double cloudWrapRel(double x, double x0, double x1) { double xNorm = (x - x0) / (x1 - x0); return (xNorm - std::floor(xNorm)) * (x1 - x0); }  // SpherePack.hpp:32-35
long cloudMake(std::vector<Sph>& pack, CloudGen& g, V3 mn, V3 mx, double rMean, double rRelFuzz, int num, bool periodic, double porosity,
               const std::vector<double>& psdSizes, const std::vector<double>& psdCumm, bool distributeMass, long cap, double& applied, int depth) {
    if (depth > 200) throw std::runtime_error("makeCloud: no packing after 200 porosity increases");
    std::vector<double> psdRadii, psdCumm2;
    V3 size = vsub(mx, mn);
    double volume = size[0] * (size[1] * size[2]);
    if (!volume) throw std::invalid_argument("The box defined as null volume.");
    int mode = -1; bool err = false;
    if (porosity <= 0) porosity = 0.5;
    if (rMean > 0) mode = 0;
    else if (num > 0 && psdSizes.size() == 0) { mode = 1; rMean = std::pow(volume * (1 - porosity) / (PI * (4 / 3.) * (1 + rRelFuzz * rRelFuzz) * num), 1 / 3.); }
    if (psdSizes.size() > 0) {
        err = (mode >= 0); mode = 2;
        if (psdSizes.size() <= 1) throw std::invalid_argument("psdSizes must have at least 2 items");
        for (size_t i = 0; i < psdSizes.size(); i++) {
            psdRadii.push_back(.5 * psdSizes[i]);
            if (distributeMass) {
                if (i == 0) psdCumm2.push_back(0);
                else psdCumm2.push_back(psdCumm2[i - 1] + 3.0 * volume * (1 - porosity) / PI * (psdCumm[i] - psdCumm[i - 1]) / (psdSizes[i] - psdSizes[i - 1]) *
                                                          (std::pow(psdSizes[i - 1], -2) - std::pow(psdSizes[i], -2)));
            }
        }
        if (distributeMass) {
            if (num > 0 && psdCumm2.back() < num) {
                applied = std::pow(psdCumm2.back() / num, 1. / 3.);
                for (double& q : psdRadii) q *= applied;
            }
            for (size_t i = 1; i < psdSizes.size(); i++) psdCumm2[i] /= psdCumm2.back();
        }
    }
    if (err || mode < 0) throw std::invalid_argument("makeCloud: radius specification");
    const int maxTry = 1000;
    double r = 0;
    for (long i = 0; (i < num) || (num < 0); i++) {
        if (num < 0 && i >= cap) return i;  // the stop at the output's capacity (the product's rule)
        double norm, rand;
        if (num > 0) rand = ((double)num - (double)i + 0.5) / ((double)num + 1.);
        else rand = g.rnd();
        if (mode != 2) r = distributeMass ? cloudPow3Interp(rand, rMean * (1 - rRelFuzz), rMean * (1 + rRelFuzz)) : rMean * (2 * (rand - .5) * rRelFuzz + 1);
        else if (distributeMass) { int piece = cloudGetPiece(rand, psdCumm2, norm); r = cloudPow3Interp(norm, psdRadii[piece], psdRadii[piece + 1]); }
        else { int piece = cloudGetPiece(rand, psdCumm, norm); r = psdRadii[piece] + norm * (psdRadii[piece + 1] - psdRadii[piece]); }
        int t;
        for (t = 0; t < maxTry; ++t) {
            V3 c;
            if (!periodic) { for (int axis = 0; axis < 3; axis++) c[axis] = mn[axis] + r + (size[axis] - 2 * r) * g.rnd(); }
            else { for (int axis = 0; axis < 3; axis++) c[axis] = g.rnd(); for (int axis = 0; axis < 3; axis++) c[axis] = mn[axis] + size[axis] * c[axis]; }
            bool overlap = false;
            for (const Sph& q : pack) {
                V3 dr;
                for (int axis = 0; axis < 3; axis++)
                    dr[axis] = periodic ? std::min(cloudWrapRel(c[axis], q.c[axis], q.c[axis] + size[axis]), cloudWrapRel(q.c[axis], c[axis], c[axis] + size[axis]))
                                        : q.c[axis] - c[axis];
                if ((q.r + r) * (q.r + r) >= vsqn(dr)) { overlap = true; break; }
            }
            if (!overlap) { pack.push_back(Sph{c, r}); break; }
        }
        if (t == maxTry) {
            if (num > 0 && mode != 0) {
                double nextPoro = porosity + (1 - porosity) / 10.;
                pack.clear();
                return cloudMake(pack, g, mn, mx, -1., rRelFuzz, num, periodic, nextPoro, psdSizes, psdCumm, distributeMass, cap, applied, depth + 1);
            }
            return i;
        }
    }
    return (long)pack.size();
}

// ---- regular packings (py/pack.py:323-350) and their predicates (py/_packPredicates.cpp) ----
// numpy.arange(start, stop, step) for doubles: length ceil((stop-start)/step); a[0] = start,
// a[1] = start + step, a[i] = start + i*(a[1] - a[0]) (numpy's DOUBLE_fill)
std::vector<double> npArange(double start, double stop, double step) {
    std::vector<double> a;
    const double len = std::ceil((stop - start) / step);
    if (!(len > 0)) return a;
    const long n = (long)len;
    a.resize(n);
    a[0] = start;
    if (n > 1) a[1] = start + step;
    for (long i = 2; i < n; i++) a[i] = a[0] + (double)i * (a[1] - a[0]);
    return a;
}
struct LatPred {
    int kind;
    double p[7];
    bool operator()(V3 pt, double pad) const {
        if (kind == 0)  // inAlignedBox (:136-148)
            return p[0] + pad <= pt[0] && p[3] - pad >= pt[0] && p[1] + pad <= pt[1] && p[4] - pad >= pt[1] && p[2] + pad <= pt[2] &&
                   p[5] - pad >= pt[2];
        if (kind == 1) {  // inSphere (:96-104)
            const V3 d = vsub(pt, v3(p[0], p[1], p[2]));
            return std::sqrt(vsqn(d)) - pad <= p[3] - pad;
        }
        // inCylinder (:189-205)
        const V3 c1 = v3(p[0], p[1], p[2]), c2 = v3(p[3], p[4], p[5]), c12 = vsub(c2, c1);
        const double ht = std::sqrt(vsqn(c12));
        const double u = (vdot(pt, c12) - vdot(c1, c12)) / (ht * ht);
        if ((u * ht < 0 + pad) || (u * ht > ht - pad)) return false;
        const double axisDist = std::sqrt(vsqn(vcross(vsub(pt, c1), vsub(pt, c2)))) / ht;
        if (axisDist > p[6] - pad) return false;
        return true;
    }
    void aabb(V3& mn, V3& mx) const {
        if (kind == 0) { mn = v3(p[0], p[1], p[2]); mx = v3(p[3], p[4], p[5]); return; }
        if (kind == 1) { mn = v3(p[0] - p[3], p[1] - p[3], p[2] - p[3]); mx = v3(p[0] + p[3], p[1] + p[3], p[2] + p[3]); return; }
        const V3 A = v3(p[0], p[1], p[2]), B = v3(p[3], p[4], p[5]);
        const double ht = std::sqrt(vsqn(vsub(B, A)));
        auto sq = [](double x) { return x * x; };
        const V3 k = v3(std::sqrt(sq(A[1] - B[1]) + sq(A[2] - B[2])) / ht, std::sqrt(sq(A[0] - B[0]) + sq(A[2] - B[2])) / ht,
                        std::sqrt(sq(A[0] - B[0]) + sq(A[1] - B[1])) / ht);
        for (int d = 0; d < 3; d++) { mn[d] = std::min(A[d], B[d]) - p[6] * k[d]; mx[d] = std::max(A[d], B[d]) + p[6] * k[d]; }
    }
};
std::vector<V3> regularOrtho(const LatPred& pred, double radius, double gap) {  // py/pack.py:323-332
    V3 mn, mx;
    pred.aabb(mn, mx);
    std::vector<double> ax[3];
    for (int i = 0; i < 3; i++) ax[i] = npArange(mn[i] + radius, mx[i] - radius, 2 * radius + gap);
    std::vector<V3> ret;
    for (double x : ax[0])
        for (double y : ax[1])
            for (double z : ax[2])
                if (pred(v3(x, y, z), radius)) ret.push_back(v3(x, y, z));
    return ret;
}
std::vector<V3> regularHexa(const LatPred& pred, double radius, double gap) {  // py/pack.py:334-350
    const double a = 2 * radius + gap;
    const double hy = a * std::sqrt(3.) / 2., hz = a * std::sqrt(6.) / 3.;
    V3 mn, mx;
    pred.aabb(mn, mx);
    const V3 dim = vsub(mx, mn);
    const long ni = (long)(dim[0] / a) + 1, nj = (long)(dim[1] / hy) + 1, nk = (long)(dim[2] / hz) + 1;
    std::vector<V3> ret;
    for (long i = 0; i < ni; i++)
        for (long j = 0; j < nj; j++)
            for (long k = 0; k < nk; k++) {
                double x = mn[0] + radius + (double)i * a, y = mn[1] + radius + (double)j * hy, z = mn[2] + radius + (double)k * hz;
                if (j % 2 == 0) x += (k % 2 == 0) ? a / 2. : -a / 2.;
                if (k % 2 != 0) { x += a / 2.; y += hy / 2.; }
                if (pred(v3(x, y, z), radius)) ret.push_back(v3(x, y, z));
            }
    return ret;
}

}  // namespace

extern "C" {

void* orc_create(void) {
    Oracle* o = new Oracle();
    o->P.dt = 1e-5; o->P.damping = 0.2; o->P.verletDist = -.05; o->P.distFactor = 1; o->P.poisson = .2; o->P.en = NAN;
    o->P.law = ORC_LAW_IDEALELPL; o->P.noAttraction = 1; o->P.trsfRenorm = 100; o->P.dist00 = 1; o->P.loneMask = 2;
    return o;
}
void orc_destroy(void* h) { delete O(h); }
const char* orc_last_error(void* h) { return O(h)->err.c_str(); }

int orc_set_nodes(void* h, int32_t n, const double* pos3, const double* ori4, const double* vel3, const double* angVel3,
                  const double* mass, const double* inertia3, const uint32_t* flags, const double* force3, const double* torque3) {
    Oracle* o = O(h);
    o->nn = n;
    o->pos.resize(n); o->vel.resize(n); o->angVel.resize(n); o->force.resize(n); o->torque.resize(n); o->inertia.resize(n);
    o->ori.resize(n); o->mass.resize(n); o->flags.resize(n);
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            o->pos[i][k] = pos3[3 * i + k]; o->vel[i][k] = vel3[3 * i + k]; o->angVel[i][k] = angVel3[3 * i + k];
            o->inertia[i][k] = inertia3[3 * i + k];
            o->force[i][k] = force3 ? force3[3 * i + k] : 0.; o->torque[i][k] = torque3 ? torque3[3 * i + k] : 0.;
        }
        o->ori[i].w = ori4[4 * i + 0]; o->ori[i].x = ori4[4 * i + 1]; o->ori[i].y = ori4[4 * i + 2]; o->ori[i].z = ori4[4 * i + 3];
        o->mass[i] = mass[i]; o->flags[i] = flags[i];
        bool asph = !(o->inertia[i][0] == o->inertia[i][1] && o->inertia[i][1] == o->inertia[i][2]);
        if (asph && (flags[i] & DOF_RXRYRZ) != DOF_RXRYRZ) { o->err = "aspherical integrator is outside the restated path"; return -1; }
    }
    o->nodeMaster.assign(n, -1);
    o->nodeClump.assign(n, -1);
    o->clumpMaster.clear(); o->clumpOff.assign(1, 0); o->clumpMem.clear(); o->clumpRelPos.clear(); o->clumpRelOri.clear();
    return 0;
}
/* ClumpData (pkg/dem/Clump.hpp): for clump c, master node master[c], members
   members[off[c]..off[c+1]) with relPos (3 per member) and relOri (w, x, y, z per member);
   node flags must mark masters CLUMP_CLUMP and members CLUMP_CLUMPED */
int orc_set_clumps(void* h, int32_t nc, const int32_t* master, const int32_t* off, const int32_t* members, const double* relPos3,
                   const double* relOri4) {
    Oracle* o = O(h);
    o->nodeMaster.assign(o->nn, -1);
    o->nodeClump.assign(o->nn, -1);
    o->clumpMaster.assign(master, master + nc);
    o->clumpOff.assign(off, off + nc + 1);
    int nm = nc > 0 ? off[nc] : 0;
    o->clumpMem.assign(members, members + nm);
    o->clumpRelPos.resize(nm); o->clumpRelOri.resize(nm);
    for (int c = 0; c < nc; c++) {
        int M = master[c];
        if (M < 0 || M >= o->nn || !(o->flags[M] & CLUMP_CLUMP) || o->nodeClump[M] >= 0) { o->err = "bad clump master"; return -1; }
        o->nodeClump[M] = c;
        for (int k = off[c]; k < off[c + 1]; k++) {
            int m = members[k];
            if (m < 0 || m >= o->nn || !(o->flags[m] & CLUMP_CLUMPED) || o->nodeMaster[m] >= 0) { o->err = "bad clump member"; return -1; }
            o->nodeMaster[m] = M;
            o->clumpRelPos[k] = v3(relPos3[3 * k], relPos3[3 * k + 1], relPos3[3 * k + 2]);
            o->clumpRelOri[k] = Qt{relOri4[4 * k], relOri4[4 * k + 1], relOri4[4 * k + 2], relOri4[4 * k + 3]};
        }
    }
    for (int i = 0; i < o->nn; i++) {
        if ((o->flags[i] & CLUMP_CLUMP) && o->nodeClump[i] < 0) { o->err = "clump master without members"; return -1; }
        if ((o->flags[i] & CLUMP_CLUMPED) && o->nodeMaster[i] < 0) { o->err = "clumped node without master"; return -1; }
    }
    return 0;
}
int orc_set_particles(void* h, int32_t n, const int32_t* shape, const int32_t* nodes3, const double* radius, const int32_t* axis,
                      const int32_t* sense, const uint32_t* mask, const int32_t* mat) {
    Oracle* o = O(h);
    o->np = n;
    o->shape.assign(shape, shape + n); o->nodeIx.assign(nodes3, nodes3 + 3 * n); o->radius.assign(radius, radius + n);
    o->axis.assign(axis, axis + n); o->sense.assign(sense, sense + n); o->mask.assign(mask, mask + n); o->matId.assign(mat, mat + n);
    o->contactsDirty = true; o->con.clear(); o->adjDirty = true; o->aabb.assign(n, Aabb());
    o->initIds();
    return 0;
}
int orc_set_materials(void* h, int32_t n, const double* density, const double* young, const double* ktDivKn, const double* tanPhi) {
    Oracle* o = O(h);
    o->mats.resize(n);
    for (int i = 0; i < n; i++) o->mats[i] = Mat{density[i], young[i], ktDivKn[i], tanPhi[i]};
    return 0;
}
int orc_set_cell(void* h, int32_t periodic, const double* hSize9, const double* gradV9, const double* nextGradV9, int32_t homoDeform) {
    Oracle* o = O(h);
    try {
        o->periodic = periodic != 0;
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
            o->hSize.m[i][j] = hSize9[3 * i + j]; o->gradV.m[i][j] = gradV9[3 * i + j]; o->nextGradV.m[i][j] = nextGradV9[3 * i + j];
        }
        o->homoDeform = homoDeform;
        o->trsf = meye(); o->pprevHsize = o->hSize;
        M3 W = mscale(msub(o->gradV, mtrans(o->gradV)), .5);
        o->spinVec = vmul(leviCivita(W), .5);
        if (o->periodic) o->cellIntegrateAndUpdate(0);  // Cell::postLoad
    } catch (std::exception& e) { o->err = e.what(); return -1; }
    return 0;
}
int orc_set_params(void* h, const orc_params* p) {
    Oracle* o = O(h);
    o->P = *p;
    if (p->law == ORC_LAW_HERTZ_DMT && (p->gamma != 0 || p->alpha != 0)) { o->err = "adhesive Hertz (gamma/alpha != 0) is outside the restated path"; return -1; }
    o->verletResolved = false;
#ifdef _OPENMP
    if (p->nThreads > 0) omp_set_num_threads(p->nThreads);
#endif
    return 0;
}
int orc_step(void* h, int32_t n) {
    Oracle* o = O(h);
#ifdef _OPENMP
    // the thread count is the OpenMP runtime's, shared by every oracle of the process: each
    // oracle sets its own before it steps
    if (o->P.nThreads > 0) omp_set_num_threads(o->P.nThreads);
#endif
    for (int i = 0; i < o->nn; i++) {
        if (((o->flags[i] & CLUMP_CLUMP) && o->nodeClump[i] < 0) || ((o->flags[i] & CLUMP_CLUMPED) && o->nodeMaster[i] < 0)) {
            o->err = "clump flags set but clumps not given (orc_set_clumps)";
            return -1;
        }
    }
    try { for (int i = 0; i < n; i++) o->doOneStep(); }
    catch (std::exception& e) { o->err = e.what(); return -1; }
    return 0;
}
int orc_get_nodes(void* h, double* pos3, double* ori4, double* vel3, double* angVel3, double* force3, double* torque3) {
    Oracle* o = O(h);
    for (int i = 0; i < o->nn; i++) {
        for (int k = 0; k < 3; k++) {
            if (pos3) pos3[3 * i + k] = o->pos[i][k];
            if (vel3) vel3[3 * i + k] = o->vel[i][k];
            if (angVel3) angVel3[3 * i + k] = o->angVel[i][k];
            if (force3) force3[3 * i + k] = o->force[i][k];
            if (torque3) torque3[3 * i + k] = o->torque[i][k];
        }
        if (ori4) { ori4[4 * i] = o->ori[i].w; ori4[4 * i + 1] = o->ori[i].x; ori4[4 * i + 2] = o->ori[i].y; ori4[4 * i + 3] = o->ori[i].z; }
    }
    return 0;
}
int64_t orc_num_contacts(void* h) { return (int64_t)O(h)->con.size(); }
/* The real contacts of a checkpoint become the contact map (Contact with L6Geom and FrictPhys
   or HertzPhys, as ObjectIO restores them: pkg/dem/L6Geom.hpp, FrictPhys/HertzPhys fields):
   ids in functor order, cellDist, trsf (row-major), local force, phys (FrictPhys: kn, kt,
   tanPhi; HertzPhys: K, R, tanPhi, kt0, alpha_sqrtMK), lens, contact point.  uN, the
   velocities and the Hertz stiffnesses are recomputed by the next ContactLoop; the next step
   runs a collider full run. */
int orc_set_contacts(void* h, int64_t n, const int32_t* idA, const int32_t* idB, const int32_t* cellDist3, const double* trsf9,
                     const double* forceLocal3, const double* phys5, const double* lens2, const double* contPt3) {
    Oracle* o = O(h);
    o->con.clear();
    for (int64_t k = 0; k < n; k++) {
        const int a = idA[k], b = idB[k];
        if (a < 0 || b < 0 || a >= o->np || b >= o->np || a == b) { o->err = "contact ids out of range"; return -1; }
        Contact c{};
        c.idA = a; c.idB = b;
        for (int q = 0; q < 3; q++) c.cd[q] = cellDist3[3 * k + q];
        c.stepCreated = o->step - 1;
        c.minDist00Sq = -1;
        c.geom = c.phys = true;
        for (int r = 0; r < 3; r++)
            for (int s = 0; s < 3; s++) c.trsf.m[r][s] = trsf9[9 * k + 3 * r + s];
        for (int q = 0; q < 3; q++) { c.force[q] = forceLocal3[3 * k + q]; c.contPt[q] = contPt3[3 * k + q]; }
        c.torque = vzero();
        c.vel = vzero(); c.angVel = vzero();
        c.lens[0] = lens2[2 * k]; c.lens[1] = lens2[2 * k + 1];
        c.uN = NAN;
        const double* ph = phys5 + 5 * k;
        if (o->P.law == ORC_LAW_HERTZ_DMT) {
            c.K = ph[0]; c.R = ph[1]; c.tanPhi = ph[2]; c.kt0 = ph[3]; c.alpha_sqrtMK = ph[4];
            c.gamma = o->P.gamma; c.alpha = o->P.alpha;
            c.kn = 0; c.kt = 0;
        } else {
            c.kn = ph[0]; c.kt = ph[1]; c.tanPhi = ph[2];
        }
        if (!o->con.emplace(Oracle::key(a, b), c).second) { o->err = "duplicate contact"; return -1; }
    }
    o->contactsDirty = true;
    return 0;
}
/* phys (the layout of orc_set_contacts) and lens of the contacts in orc_get_contacts order
   (NaN for contacts that are not real) */
int orc_get_contact_phys(void* h, int64_t cap, double* phys5, double* lens2) {
    Oracle* o = O(h);
    int64_t i = 0;
    for (auto& kv : o->con) {
        if (i >= cap) break;
        const Contact& c = kv.second;
        const bool r = c.isReal();
        double ph[5] = {c.kn, c.kt, c.tanPhi, 0., 0.};
        if (o->P.law == ORC_LAW_HERTZ_DMT) { ph[0] = c.K; ph[1] = c.R; ph[2] = c.tanPhi; ph[3] = c.kt0; ph[4] = c.alpha_sqrtMK; }
        for (int q = 0; q < 5; q++) phys5[5 * i + q] = r ? ph[q] : NAN;
        for (int q = 0; q < 2; q++) lens2[2 * i + q] = r ? c.lens[q] : NAN;
        i++;
    }
    return 0;
}
/* Scene::step and Scene::time of a checkpoint */
int orc_set_time(void* h, int64_t step, double time) {
    Oracle* o = O(h);
    o->step = (long)step;
    o->time = time;
    return 0;
}
int orc_get_contacts(void* h, int64_t cap, int32_t* idA, int32_t* idB, int32_t* cellDist3, uint8_t* isReal, double* forceLocal3,
                     double* contPt3, double* trsf9, double* uN, int64_t* nOut) {
    Oracle* o = O(h);
    int64_t i = 0;
    for (auto& kv : o->con) {
        if (i >= cap) break;
        const Contact& c = kv.second;
        if (idA) idA[i] = c.idA;
        if (idB) idB[i] = c.idB;
        for (int k = 0; k < 3; k++) {
            if (cellDist3) cellDist3[3 * i + k] = c.cd[k];
            if (forceLocal3) forceLocal3[3 * i + k] = c.isReal() ? c.force[k] : 0.;
            if (contPt3) contPt3[3 * i + k] = c.contPt[k];
        }
        if (trsf9) for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) trsf9[9 * i + 3 * a + b] = c.trsf.m[a][b];
        if (uN) uN[i] = c.isReal() ? c.uN : NAN;
        if (isReal) isReal[i] = c.isReal() ? 1 : 0;
        i++;
    }
    if (nOut) *nOut = i;
    return 0;
}
int orc_get_energy(void* h, double* out, int32_t* created) {
    Oracle* o = O(h);
    for (int k = 0; k < ORC_E_N; k++) { if (out) out[k] = o->E[k]; if (created) created[k] = o->Ecreated[k]; }
    return 0;
}
int orc_get_stats(void* h, int64_t* step, int64_t* nFullRuns, double* verletDist) {
    Oracle* o = O(h);
    if (step) *step = o->step;
    if (nFullRuns) *nFullRuns = o->nFullRuns;
    if (verletDist) *verletDist = o->verletDistActual;
    return 0;
}
int orc_get_sort_stats(void* h, int64_t* sortChunks, int64_t* sortPasses) {
    Oracle* o = O(h);
    if (sortChunks) *sortChunks = o->sortChunks;
    if (sortPasses) *sortPasses = o->sortPasses;
    return 0;
}
int orc_get_cell(void* h, double* hSize9, double* gradV9) {
    Oracle* o = O(h);
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) { if (hSize9) hSize9[3 * i + j] = o->hSize.m[i][j]; if (gradV9) gradV9[3 * i + j] = o->gradV.m[i][j]; }
    return 0;
}
int64_t orc_bruteforce_pairs(void* h, int64_t cap, int32_t* idMin, int32_t* idMax) {
    Oracle* o = O(h);
    int64_t n = 0;
    if ((long)o->minima.size() != 3L * o->np) return -1;
    for (int a = 0; a < o->np; a++)
        for (int b = a + 1; b < o->np; b++) {
            if (!o->aabb[a].has || !o->aabb[b].has) continue;
            bool ov;
            try { ov = o->boundsOverlap(a, b); } catch (...) { return -2; }
            if (!ov || !o->mayCollide(a, b)) continue;
            if (n < cap) { idMin[n] = a; idMax[n] = b; }
            n++;
        }
    return n;
}
int orc_remove_particles(void* h, int32_t n, const int32_t* ids) {
    Oracle* o = O(h);
    std::vector<char> seen(o->np, 0);
    for (int k = 0; k < n; k++) {
        const int id = ids[k];
        if (id < 0 || id >= o->np || o->shape[id] == ORC_SHAPE_NONE) { o->err = "particle does not exist"; return -1; }
        if (o->shape[id] == ORC_SHAPE_FACET) { o->err = "removing facets is not restated"; return -1; }
        if (o->flags[o->nodeOf(id)] & (CLUMP_CLUMPED | CLUMP_CLUMP)) { o->err = "a node is clumped, remove the clump itself instead!"; return -1; }
        if (seen[id]++) { o->err = "particle listed twice"; return -1; }
    }
    for (int k = 0; k < n; k++) o->removeParticle(ids[k]);
    return 0;
}
int orc_insert_spheres(void* h, int32_t n, const double* pos3, const double* vel3, const double* angVel3, const double* radius,
                       const double* mass, const double* inertia, const int32_t* mat, const uint32_t* mask, int32_t* idsOut) {
    Oracle* o = O(h);
    if ((long)o->freeIds.size() + (o->np - o->parEnd) < n) { o->err = "no free particle slot"; return -1; }
    for (int k = 0; k < n; k++) {
        V3 p = v3(pos3[3 * k], pos3[3 * k + 1], pos3[3 * k + 2]);
        V3 v = vel3 ? v3(vel3[3 * k], vel3[3 * k + 1], vel3[3 * k + 2]) : vzero();
        V3 w = angVel3 ? v3(angVel3[3 * k], angVel3[3 * k + 1], angVel3[3 * k + 2]) : vzero();
        idsOut[k] = o->insertSphere(p, v, w, radius[k], mass[k], inertia[k], mat[k], mask[k]);
    }
    return 0;
}
int orc_set_masks(void* h, int32_t n, const int32_t* ids, const uint32_t* masks) {
    for (int k = 0; k < n; k++) O(h)->mask[ids[k]] = masks[k];
    return 0;
}
int orc_get_free_ids(void* h, int32_t cap, int32_t* ids, int32_t* n, int32_t* parEnd) {
    Oracle* o = O(h);
    int k = 0;
    for (int q : o->freeIds) { if (k < cap) ids[k] = q; k++; }
    *n = k;
    *parEnd = o->parEnd;
    return 0;
}
int orc_make_cloud(const double* mn3, const double* mx3, double rMean, double rRelFuzz, int32_t num, int32_t periodic, double porosity,
                   const double* psdSizes, const double* psdCumm, int32_t nPsd, int32_t distributeMass, uint32_t* state, int64_t cap,
                   double* c3, double* r, int64_t* nOut, double* appliedPsdScaling) {
    try {
        std::vector<Sph> pack;
        CloudGen g{*state};
        double applied = 1.;
        std::vector<double> ps(psdSizes, psdSizes + (nPsd > 0 ? nPsd : 0)), pc(psdCumm, psdCumm + (nPsd > 0 ? nPsd : 0));
        cloudMake(pack, g, v3(mn3[0], mn3[1], mn3[2]), v3(mx3[0], mx3[1], mx3[2]), rMean, rRelFuzz, num, periodic != 0, porosity, ps, pc,
                  distributeMass != 0, cap, applied, 0);
        if ((int64_t)pack.size() > cap) return -1;
        for (size_t i = 0; i < pack.size(); i++) {
            for (int d = 0; d < 3; d++) c3[3 * i + d] = pack[i].c[d];
            r[i] = pack[i].r;
        }
        *nOut = (int64_t)pack.size();
        *state = g.x;
        if (appliedPsdScaling) *appliedPsdScaling = applied;
    } catch (std::exception&) { return -1; }
    return 0;
}
int orc_regular_pack(int32_t hexa, int32_t kind, const double* params, double radius, double gap, int64_t cap, double* c3, int64_t* nOut) {
    LatPred pr;
    pr.kind = kind;
    for (int q = 0; q < 7; q++) pr.p[q] = q < (kind == 0 ? 6 : (kind == 1 ? 4 : 7)) ? params[q] : 0.;
    const std::vector<V3> pts = hexa ? regularHexa(pr, radius, gap) : regularOrtho(pr, radius, gap);
    *nOut = (int64_t)pts.size();
    if ((int64_t)pts.size() > cap) return -1;
    for (size_t i = 0; i < pts.size(); i++)
        for (int d = 0; d < 3; d++) c3[3 * i + d] = pts[i][d];
    return 0;
}
int orc_set_dyndt(void* h, const orc_dyndt* d) {
    if (d->enabled && 1. + d->maxRelInc == 1.) { O(h)->err = "DynDt: maxRelInc too small (1.0+maxRelInc==1.0)"; return -1; }
    O(h)->dd = *d;
    return 0;
}
int orc_get_dyndt(void* h, orc_dyndt* d) { *d = O(h)->dd; return 0; }
int orc_dyndt_crit(void* h, double* critDt) {
    try { *critDt = O(h)->critDtStiffness(); }
    catch (std::exception& e) { O(h)->err = e.what(); return -1; }
    return 0;
}
int orc_get_time(void* h, double* time, double* dt) { *time = O(h)->time; *dt = O(h)->P.dt; return 0; }
double orc_pwave_dt(void* h) {
    Oracle* o = O(h);
    double dt = INF;
    for (int p = 0; p < o->np; p++) {
        if (o->shape[p] != ORC_SHAPE_SPHERE) continue;
        const Mat& m = o->mats[o->matId[p]];
        dt = std::min(dt, o->radius[p] / (1. * std::sqrt(m.young / m.density)));
    }
    return dt;
}

}  // extern "C"
