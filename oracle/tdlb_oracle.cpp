// tdlb oracle -- CPU restatement of the reference ComputeUnit hot path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product links or calls this file:
// it is loaded by tests/, by __graft_entry__.smoke() as the checker and by
// bench.py's cpu_baseline leg.  It restates, in plain C++, the algorithm of
// TurbulentDynamics/tdLBCpp's CPU ComputeUnit, single-threaded like the
// reference unless orc_set_threads asks for more (then collision and moments,
// whose cells are independent, run on OpenMP threads; every result is the same):
//
//   indexing / layouts     tdlbcpp/src/ComputeUnit.hpp:25-40,453-476; Field.hpp:18-144
//   initialise             tdlbcpp/src/ComputeUnit.hpp:174-190; QVec.hpp:146-151
//   moments                tdlbcpp/src/CollisionEgglesSomers.hpp:185-256
//   collision (+LES)       tdlbcpp/src/CollisionEgglesSomers.hpp:26-181; QVec.hpp:125-134
//   Simple push / pull     tdlbcpp/src/StreamingSimple.hpp:18-126
//   Esotwist accessors     tdlbcpp/src/StreamingEsoteric.hpp:14-180
//   bounce-back            tdlbcpp/src/Boundary.hpp:13-250
//   IB forcing             tdlbcpp/src/Forcing.hpp:18-176
//   calcNuAndRe_m          tdlbcpp/src/Params/FlowParams.hpp:65-70
//   doubleResolutionFullCU tdlbcpp/src/ComputeUnit.hpp:229-316,480-502
//   savePlaneXZ records    tdlbcpp/src/ComputeUnitOutput.hpp:354-447
//   calcVorticityXZ field  tdlbcpp/src/ComputeUnitOutput.hpp:113-160
//
// Arithmetic keeps the reference's literal types: in the float build the
// double literals (0.5, 1.0, 2.0, 1.5, 24.0, ...) promote sub-expressions to
// double exactly as C++ does in the reference, and the terms of every sum are
// added in the reference's order.  Build with -ffp-contract=off so that no
// multiply-add is fused.
//
// The checks that pin each function against the reference's fixtures and the
// mathematics live in tests/test_oracle.py.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <type_traits>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace orc {

enum { IJKL = 0, LIJK = 1 };
enum { SIMPLE = 0, ESOTWIST = 1 };

// slot indices (Header.h:70-89)
enum { Q01, Q02, Q03, Q04, Q05, Q06, Q07, Q08, Q09, Q10, Q11, Q12, Q13, Q14, Q15, Q16, Q17, Q18 };
enum { M01, M02, M03, M04, M05, M06, M07, M08, M09, M10, M11, M12, M13, M14, M15, M16, M17, M18 };

static const int EX[18] = {1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
static const int EY[18] = {0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
static const int EZ[18] = {0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
static inline int opp(int d) { return d ^ 1; }

// Flow parameters used on the path (FlowParams.hpp:27-61), in the CU's type.
template <typename T>
struct Flow {
    T initialRho, reMNonDimensional, uav, cs0, g3, nu, fx0, fy0, fz0, Re_m, Re_f, uf, alpha, beta;
    bool useLES;
};

// PosPolar subset read by forcing (Forcing.hpp:76-78,81,126-128).
struct Point {
    int64_t i, j, k;
    double iFrac, kFrac, uDelta, vDelta, wDelta;
};

struct Base {
    virtual ~Base() {}
    virtual void initialise(double rho) = 0;
    virtual void collision() = 0;
    virtual void moments() = 0;
    virtual void push() = 0;
    virtual void pull() = 0;
    virtual void bounceBack() = 0;
    virtual void forcing(const Point* g, int64_t n, double alfa, double beta) = 0;
    virtual void* field(int which, size_t* bytes) = 0;
    virtual bool& even() = 0;
    virtual int64_t dims(int which) = 0;
    virtual void cells(int op, void* buf) = 0;
    virtual void doubleResolution() = 0;
    virtual void planeXZ(int64_t cut, const uint8_t* exclude, int64_t i0, int64_t k0, void* q, int64_t* nq, void* f,
                         int64_t* nf) = 0;
    virtual void vorticityXZ(int64_t cut, const uint8_t* exclude, void* out) = 0;
    virtual void vorticityXY(int64_t cut, const uint8_t* exclude, void* out) = 0;
    virtual void setFlow(const double* v, int useLES) = 0;
};

template <typename T>
Flow<T> makeFlow(const double* v, int useLES);

template <typename T, int LAYOUT, int STREAM>
struct CU : Base {
    int64_t x, y, z, ghost;
    int64_t xg, yg, zg, xg0, yg0, zg0, xg1, yg1, zg1;
    size_t size;
    bool evenStep;
    Flow<T> flow;
    std::vector<T> Q, F, Nu;
    std::vector<uint8_t> O;

    CU(int64_t x_, int64_t y_, int64_t z_, int64_t g_, const Flow<T>& fl)
        : x(x_), y(y_), z(z_), ghost(g_), evenStep(true), flow(fl) {
        // initGridParams, ComputeUnit.hpp:25-40
        xg = x + 2 * ghost; yg = y + 2 * ghost; zg = z + 2 * ghost;
        xg0 = xg - 1; yg0 = yg - 1; zg0 = zg - 1;
        xg1 = xg - 2; yg1 = yg - 2; zg1 = zg - 2;
        size = size_t(xg) * yg * zg;
        Q.assign(size * 18, T(0));
        F.assign(size * 3, T(0));
        Nu.assign(size, T(0));
        O.assign(size, 0);
    }

    bool& even() override { return evenStep; }
    // the ComputeUnit's `flow` member replaced (e.g. nu after calcNuAndRe_m)
    void setFlow(const double* v, int useLES) override { flow = makeFlow<T>(v, useLES); }
    int64_t dims(int w) override {
        const int64_t v[] = {xg, yg, zg, (int64_t)size};
        return v[w];
    }

    inline int64_t index(int64_t i, int64_t j, int64_t k) const { return i * (yg * zg) + (j * zg) + k; }

    // IJKL: q[index*18 + l]; LIJK: q[l*size + index]  (Field.hpp:18-42, 45-83, 130-144)
    inline T& slot(int64_t n, int l) {
        if (LAYOUT == IJKL) return Q[size_t(n) * 18 + l];
        return Q[size_t(l) * size + size_t(n)];
    }

    // ---- access to one cell's 18-vector: Simple = plain, Esotwist = twisted ----
    inline int64_t wrap1(int64_t a, int64_t a1) const { return a == a1 ? 1 : a + 1; }
    // cell that stores direction d for cell (i,j,k) under the esoteric twist
    inline int64_t home(int64_t i, int64_t j, int64_t k, int d) const {
        int64_t ii = EX[d] < 0 ? wrap1(i, xg1) : i;
        int64_t jj = EY[d] < 0 ? wrap1(j, yg1) : j;
        int64_t kk = EZ[d] < 0 ? wrap1(k, zg1) : k;
        return index(ii, jj, kk);
    }

    void readCell(int64_t i, int64_t j, int64_t k, T q[18]) {
        if (STREAM == SIMPLE) {
            int64_t n = index(i, j, k);
            for (int l = 0; l < 18; l++) q[l] = slot(n, l);
            return;
        }
        // StreamingEsoteric.hpp:29-81
        for (int d = 0; d < 18; d++) q[d] = slot(home(i, j, k, d), evenStep ? d : opp(d));
    }
    void writeCell(int64_t i, int64_t j, int64_t k, const T q[18]) {
        if (STREAM == SIMPLE) {
            int64_t n = index(i, j, k);
            for (int l = 0; l < 18; l++) slot(n, l) = q[l];
            return;
        }
        // StreamingEsoteric.hpp:82-130: populations land in the opposite slot of the
        // opposite direction's home cell (even) or in their own slot there (odd).
        for (int d = 0; d < 18; d++)
            slot(home(i, j, k, opp(d)), evenStep ? opp(d) : d) = q[d];
    }
    void writeMomentsCell(int64_t i, int64_t j, int64_t k, const T q[18]) {
        if (STREAM == SIMPLE) { writeCell(i, j, k, q); return; }
        // StreamingEsoteric.hpp:131-179: the inverse of readCell at the same parity
        for (int d = 0; d < 18; d++) slot(home(i, j, k, d), evenStep ? d : opp(d)) = q[d];
    }

    // ---------------------------------------------------------------- initialise
    void initialise(double rhoIn) override {
        T rho = T(rhoIn);
        for (int64_t i = 0; i <= xg0; i++)
            for (int64_t j = 0; j <= yg0; j++)
                for (int64_t k = 0; k <= zg0; k++) {
                    int64_t n = index(i, j, k);
                    slot(n, 0) = rho;  // moment space: (rho, 0, ..., 0)
                    for (int l = 1; l < 18; l++) slot(n, l) = T(0.0);
                    F[3 * n + 0] = T(0.0); F[3 * n + 1] = T(0.0); F[3 * n + 2] = T(0.0);
                    Nu[n] = T(0.0);
                    O[n] = 0;
                }
    }

    // ---------------------------------------------------------------- moments
    static void momentsOf(const T* q, T* m) {
        m[M01] = q[Q01] + q[Q03] + q[Q02] + q[Q04] + q[Q05] + q[Q06] + q[Q07] + q[Q14] + q[Q08] + q[Q13]
               + q[Q09] + q[Q16] + q[Q10] + q[Q15] + q[Q11] + q[Q18] + q[Q12] + q[Q17];
        m[M02] = q[Q01] - q[Q02] + q[Q07] - q[Q14] - q[Q08] + q[Q13] + q[Q09] - q[Q16] - q[Q10] + q[Q15];
        m[M03] = q[Q03] - q[Q04] + q[Q07] + q[Q14] - q[Q08] - q[Q13] + q[Q11] - q[Q18] - q[Q12] + q[Q17];
        m[M04] = q[Q05] - q[Q06] + q[Q09] + q[Q16] - q[Q10] - q[Q15] + q[Q11] + q[Q18] - q[Q12] - q[Q17];
        m[M05] = -q[Q03] - q[Q04] - q[Q05] - q[Q06] + q[Q07] + q[Q14] + q[Q08] + q[Q13] + q[Q09] + q[Q16]
               + q[Q10] + q[Q15];
        m[M06] = q[Q07] - q[Q14] + q[Q08] - q[Q13];
        m[M07] = -q[Q01] - q[Q02] - q[Q05] - q[Q06] + q[Q07] + q[Q14] + q[Q08] + q[Q13] + q[Q11] + q[Q18]
               + q[Q12] + q[Q17];
        m[M08] = q[Q09] - q[Q16] + q[Q10] - q[Q15];
        m[M09] = q[Q11] - q[Q18] + q[Q12] - q[Q17];
        m[M10] = -q[Q01] - q[Q03] - q[Q02] - q[Q04] + q[Q09] + q[Q16] + q[Q10] + q[Q15] + q[Q11] + q[Q18]
               + q[Q12] + q[Q17];
        m[M11] = -q[Q01] + q[Q02] + 2 * q[Q07] - 2 * q[Q14] - 2 * q[Q08] + 2 * q[Q13] - q[Q09] + q[Q16]
               + q[Q10] - q[Q15];
        m[M12] = -q[Q03] + q[Q04] + 2 * q[Q07] + 2 * q[Q14] - 2 * q[Q08] - 2 * q[Q13] - q[Q11] + q[Q18]
               + q[Q12] - q[Q17];
        m[M13] = -3 * q[Q01] + 3 * q[Q02] + 3 * q[Q09] - 3 * q[Q16] - 3 * q[Q10] + 3 * q[Q15];
        m[M14] = -3 * q[Q03] + 3 * q[Q04] + 3 * q[Q11] - 3 * q[Q18] - 3 * q[Q12] + 3 * q[Q17];
        m[M15] = -2 * q[Q05] + 2 * q[Q06] + q[Q09] + q[Q16] - q[Q10] - q[Q15] + q[Q11] + q[Q18] - q[Q12]
               - q[Q17];
        m[M16] = -3 * q[Q09] - 3 * q[Q16] + 3 * q[Q10] + 3 * q[Q15] + 3 * q[Q11] + 3 * q[Q18] - 3 * q[Q12]
               - 3 * q[Q17];
        // the two 4th-order moments are evaluated in double (0.5 and 1.5 are double literals)
        m[M17] = 0.5 * q[Q01] + 0.5 * q[Q03] + 0.5 * q[Q02] + 0.5 * q[Q04] - q[Q05] - q[Q06] - q[Q07]
               - q[Q14] - q[Q08] - q[Q13] + 0.5 * q[Q09] + 0.5 * q[Q16] + 0.5 * q[Q10] + 0.5 * q[Q15]
               + 0.5 * q[Q11] + 0.5 * q[Q18] + 0.5 * q[Q12] + 0.5 * q[Q17];
        m[M18] = 1.5 * q[Q01] - 1.5 * q[Q03] + 1.5 * q[Q02] - 1.5 * q[Q04] - 1.5 * q[Q09] - 1.5 * q[Q16]
               - 1.5 * q[Q10] - 1.5 * q[Q15] + 1.5 * q[Q11] + 1.5 * q[Q18] + 1.5 * q[Q12] + 1.5 * q[Q17];
    }

    // cells are independent (each reads and writes only its own slots), so the
    // threaded build splits the i-planes; results do not depend on the split
    void moments() override {
#pragma omp parallel for schedule(static)
        for (int64_t i = 1; i <= xg1; i++)
            for (int64_t j = 1; j <= yg1; j++)
                for (int64_t k = 1; k <= zg1; k++) {
                    T q[18], m[18];
                    readCell(i, j, k, q);
                    momentsOf(q, m);
                    writeMomentsCell(i, j, k, m);
                }
    }

    // ---------------------------------------------------------------- collision
    // One cell of the Eggels-Somers collision; m holds moments on entry,
    // populations on exit.  nu is the cell's eddy viscosity (read & written with LES).
    void collideCell(T* m, const T* f, T& nuCell, T b, T c) {
        // velocity with half the force (QVec.hpp:125-134)
        T ux = (1.0 / m[M01]) * (m[M02] + 0.5 * f[0]);
        T uy = (1.0 / m[M01]) * (m[M03] + 0.5 * f[1]);
        T uz = (1.0 / m[M01]) * (m[M04] + 0.5 * f[2]);

        T a[18];
        if (flow.useLES) {
            T fct = 3.0 / (m[M01] * (1.0 + 6.0 * (nuCell + flow.nu)));
            T dudx = fct * ((m[M02] + 0.5 * f[0] * ux - m[M05]));
            T dvdy = fct * ((m[M03] + 0.5 * f[1] * uy - m[M07]));
            T dwdz = fct * ((m[M04] + 0.5 * f[2] * uz - m[M10]));
            T divv = dudx + dvdy + dwdz;
            T dudypdvdx = 2 * fct * ((m[M03]) + 0.5 * f[1] * ux - m[M06]);
            T dudzpdwdx = 2 * fct * ((m[M04]) + 0.5 * f[2] * ux - m[M08]);
            T dvdzpdwdy = 2 * fct * ((m[M04]) + 0.5 * f[2] * uy - m[M09]);
            T sh = 2 * dudx * dudx + 2 * dvdy * dvdy + 2 * dwdz * dwdz + dudypdvdx * dudypdvdx
                 + dudzpdwdx * dudzpdwdx + dvdzpdwdy * dvdzpdwdy - (2.0 / 3.0) * divv * divv;
            nuCell = flow.cs0 * flow.cs0 * std::sqrt(std::fabs(sh));
            T nut = nuCell + flow.nu;
            b = 1.0 / (1.0 + 6 * nut);
            c = 1.0 - 6 * nut;
        }

        a[M01] = m[M01];
        a[M02] = m[M02] + f[0];
        a[M03] = m[M03] + f[1];
        a[M04] = m[M04] + f[2];
        a[M05] = (2.0 * (m[M02] + 0.5 * f[0]) * ux - m[M05] * c) * b;
        a[M06] = (2.0 * (m[M02] + 0.5 * f[0]) * uy - m[M06] * c) * b;
        a[M07] = (2.0 * (m[M03] + 0.5 * f[1]) * uy - m[M07] * c) * b;
        a[M08] = (2.0 * (m[M02] + 0.5 * f[0]) * uz - m[M08] * c) * b;
        a[M09] = (2.0 * (m[M03] + 0.5 * f[1]) * uz - m[M09] * c) * b;
        a[M10] = (2.0 * (m[M04] + 0.5 * f[2]) * uz - m[M10] * c) * b;
        for (int l = M11; l <= M16; l++) a[l] = -flow.g3 * m[l];
        a[M17] = 0.0;
        a[M18] = 0.0;
        for (int l = 0; l < 18; l++) a[l] /= 24.0;

        // f = E.alpha (Eggels & Somers 1995 eq. 12), terms in the reference's order
        m[Q01] = 2 * a[M01] + 4 * a[M02] + 3 * a[M05] - 3 * a[M07] - 3 * a[M10] - 2 * a[M11] - 2 * a[M13]
               + 2 * a[M17] + 2 * a[M18];
        m[Q02] = 2 * a[M01] - 4 * a[M02] + 3 * a[M05] - 3 * a[M07] - 3 * a[M10] + 2 * a[M11] + 2 * a[M13]
               + 2 * a[M17] + 2 * a[M18];
        m[Q03] = 2 * a[M01] + 4 * a[M03] - 3 * a[M05] + 3 * a[M07] - 3 * a[M10] - 2 * a[M12] - 2 * a[M14]
               + 2 * a[M17] - 2 * a[M18];
        m[Q04] = 2 * a[M01] - 4 * a[M03] - 3 * a[M05] + 3 * a[M07] - 3 * a[M10] + 2 * a[M12] + 2 * a[M14]
               + 2 * a[M17] - 2 * a[M18];
        m[Q05] = 2 * a[M01] + 4 * a[M04] - 3 * a[M05] - 3 * a[M07] + 3 * a[M10] - 4 * a[M15] - 4 * a[M17];
        m[Q06] = 2 * a[M01] - 4 * a[M04] - 3 * a[M05] - 3 * a[M07] + 3 * a[M10] + 4 * a[M15] - 4 * a[M17];
        m[Q07] = a[M01] + 2 * a[M02] + 2 * a[M03] + 1.5 * a[M05] + 6 * a[M06] + 1.5 * a[M07] - 1.5 * a[M10]
               + 2 * a[M11] + 2 * a[M12] - 2 * a[M17];
        m[Q14] = a[M01] - 2 * a[M02] + 2 * a[M03] + 1.5 * a[M05] - 6 * a[M06] + 1.5 * a[M07] - 1.5 * a[M10]
               - 2 * a[M11] + 2 * a[M12] - 2 * a[M17];
        m[Q08] = a[M01] - 2 * a[M02] - 2 * a[M03] + 1.5 * a[M05] + 6 * a[M06] + 1.5 * a[M07] - 1.5 * a[M10]
               - 2 * a[M11] - 2 * a[M12] - 2 * a[M17];
        m[Q13] = a[M01] + 2 * a[M02] - 2 * a[M03] + 1.5 * a[M05] - 6 * a[M06] + 1.5 * a[M07] - 1.5 * a[M10]
               + 2 * a[M11] - 2 * a[M12] - 2 * a[M17];
        m[Q09] = a[M01] + 2 * a[M02] + 2 * a[M04] + 1.5 * a[M05] - 1.5 * a[M07] + 6 * a[M08] + 1.5 * a[M10]
               - a[M11] + a[M13] + a[M15] - a[M16] + a[M17] - a[M18];
        m[Q16] = a[M01] - 2 * a[M02] + 2 * a[M04] + 1.5 * a[M05] - 1.5 * a[M07] - 6 * a[M08] + 1.5 * a[M10]
               + a[M11] - a[M13] + a[M15] - a[M16] + a[M17] - a[M18];
        m[Q10] = a[M01] - 2 * a[M02] - 2 * a[M04] + 1.5 * a[M05] - 1.5 * a[M07] + 6 * a[M08] + 1.5 * a[M10]
               + a[M11] - a[M13] - a[M15] + a[M16] + a[M17] - a[M18];
        m[Q15] = a[M01] + 2 * a[M02] - 2 * a[M04] + 1.5 * a[M05] - 1.5 * a[M07] - 6 * a[M08] + 1.5 * a[M10]
               - a[M11] + a[M13] - a[M15] + a[M16] + a[M17] - a[M18];
        m[Q11] = a[M01] + 2 * a[M03] + 2 * a[M04] - 1.5 * a[M05] + 1.5 * a[M07] + 6 * a[M09] + 1.5 * a[M10]
               - a[M12] + a[M14] + a[M15] + a[M16] + a[M17] + a[M18];
        m[Q18] = a[M01] - 2 * a[M03] + 2 * a[M04] - 1.5 * a[M05] + 1.5 * a[M07] - 6 * a[M09] + 1.5 * a[M10]
               + a[M12] - a[M14] + a[M15] + a[M16] + a[M17] + a[M18];
        m[Q12] = a[M01] - 2 * a[M03] - 2 * a[M04] - 1.5 * a[M05] + 1.5 * a[M07] + 6 * a[M09] + 1.5 * a[M10]
               + a[M12] - a[M14] - a[M15] - a[M16] + a[M17] + a[M18];
        m[Q17] = a[M01] + 2 * a[M03] - 2 * a[M04] - 1.5 * a[M05] + 1.5 * a[M07] - 6 * a[M09] + 1.5 * a[M10]
               - a[M12] + a[M14] - a[M15] - a[M16] + a[M17] + a[M18];
    }

    void collision() override {
        T b = 1.0 / (1.0 + 6 * flow.nu);
        T c = 1.0 - 6 * flow.nu;
#pragma omp parallel for schedule(static)
        for (int64_t i = 1; i <= xg1; i++)
            for (int64_t j = 1; j <= yg1; j++)
                for (int64_t k = 1; k <= zg1; k++) {
                    T m[18];
                    int64_t n = index(i, j, k);
                    T f[3] = {F[3 * n + 0], F[3 * n + 1], F[3 * n + 2]};
                    readCell(i, j, k, m);
                    collideCell(m, f, Nu[n], b, c);
                    writeCell(i, j, k, m);
                }
    }

    // ---------------------------------------------------------------- streaming
    // In-place push: slot d of interior cell c moves to c+e_d.  Directions whose
    // first non-zero component is negative sweep forward, the others backward,
    // so every source is read before it is overwritten (StreamingSimple.hpp:76-126).
    void push() override {
        if (STREAM == ESOTWIST) { evenStep = !evenStep; return; }
        static const int fwd[9] = {Q02, Q04, Q06, Q08, Q10, Q12, Q14, Q16, Q18};
        static const int bwd[9] = {Q01, Q03, Q05, Q07, Q09, Q11, Q13, Q15, Q17};
        for (int64_t i = 1; i <= xg1; i++)
            for (int64_t j = 1; j <= yg1; j++)
                for (int64_t k = 1; k <= zg1; k++)
                    for (int d : fwd) slot(index(i + EX[d], j + EY[d], k + EZ[d]), d) = slot(index(i, j, k), d);
        for (int64_t i = xg1; i >= 1; i--)
            for (int64_t j = yg1; j >= 1; j--)
                for (int64_t k = zg1; k >= 1; k--)
                    for (int d : bwd) slot(index(i + EX[d], j + EY[d], k + EZ[d]), d) = slot(index(i, j, k), d);
    }

    // In-place pull into interior cells from c-e_d (StreamingSimple.hpp:18-66).
    void pull() override {
        if (STREAM == ESOTWIST) { evenStep = !evenStep; return; }
        static const int fwd[9] = {Q02, Q04, Q06, Q08, Q10, Q12, Q14, Q16, Q18};
        static const int bwd[9] = {Q01, Q03, Q05, Q07, Q09, Q11, Q13, Q15, Q17};
        for (int64_t i = 1; i <= xg1; i++)
            for (int64_t j = 1; j <= yg1; j++)
                for (int64_t k = 1; k <= zg1; k++)
                    for (int d : fwd) slot(index(i, j, k), d) = slot(index(i - EX[d], j - EY[d], k - EZ[d]), d);
        for (int64_t i = xg1; i >= 1; i--)
            for (int64_t j = yg1; j >= 1; j--)
                for (int64_t k = zg1; k >= 1; k--)
                    for (int d : bwd) slot(index(i, j, k), d) = slot(index(i - EX[d], j - EY[d], k - EZ[d]), d);
    }

    // ---------------------------------------------------------------- bounce-back
    // Ghost slot d of a face cell g takes the opposite population of g+e_d.
    // Faces in the reference order Right, Left, Up, Down, Backward, Forward,
    // then the 12 edge lines (Boundary.hpp:17-35, 41-250).
    void bbCell(int64_t i, int64_t j, int64_t k, int d) {
        slot(index(i, j, k), d) = slot(index(i + EX[d], j + EY[d], k + EZ[d]), opp(d));
    }
    void bounceBack() override {
        if (STREAM == ESOTWIST) return;  // Boundary.hpp:13-14
        static const int right[5] = {Q01, Q07, Q13, Q09, Q15}, left[5] = {Q02, Q08, Q14, Q10, Q16};
        static const int up[5] = {Q03, Q07, Q14, Q11, Q17}, down[5] = {Q04, Q08, Q13, Q12, Q18};
        static const int back[5] = {Q05, Q09, Q16, Q11, Q18}, fwd[5] = {Q06, Q10, Q15, Q12, Q17};
        for (int64_t j = 1; j <= yg1; j++) for (int64_t k = 1; k <= zg1; k++) for (int d : right) bbCell(0, j, k, d);
        for (int64_t j = 1; j <= yg1; j++) for (int64_t k = 1; k <= zg1; k++) for (int d : left) bbCell(xg0, j, k, d);
        for (int64_t i = 1; i <= xg1; i++) for (int64_t k = 1; k <= zg1; k++) for (int d : up) bbCell(i, 0, k, d);
        for (int64_t i = 1; i <= xg1; i++) for (int64_t k = 1; k <= zg1; k++) for (int d : down) bbCell(i, yg0, k, d);
        for (int64_t i = 1; i <= xg1; i++) for (int64_t j = 1; j <= yg1; j++) for (int d : back) bbCell(i, j, 0, d);
        for (int64_t i = 1; i <= xg1; i++) for (int64_t j = 1; j <= yg1; j++) for (int d : fwd) bbCell(i, j, zg0, d);
        // edges (Boundary.hpp:41-132)
        for (int64_t k = 1; k <= zg1; k++) bbCell(0, 0, k, Q07);
        for (int64_t k = 1; k <= zg1; k++) bbCell(xg0, 0, k, Q14);
        for (int64_t k = 1; k <= zg1; k++) bbCell(0, yg0, k, Q13);
        for (int64_t k = 1; k <= zg1; k++) bbCell(xg0, yg0, k, Q08);
        for (int64_t j = 1; j <= yg1; j++) bbCell(0, j, 0, Q09);
        for (int64_t j = 1; j <= yg1; j++) bbCell(0, j, zg0, Q15);
        for (int64_t j = 1; j <= yg1; j++) bbCell(xg0, j, zg0, Q10);
        for (int64_t j = 1; j <= yg1; j++) bbCell(xg0, j, 0, Q16);
        for (int64_t i = 1; i <= xg1; i++) bbCell(i, 0, 0, Q11);
        for (int64_t i = 1; i <= xg1; i++) bbCell(i, 0, zg0, Q17);
        for (int64_t i = 1; i <= xg1; i++) bbCell(i, yg0, zg0, Q12);
        for (int64_t i = 1; i <= xg1; i++) bbCell(i, yg0, 0, Q18);
    }

    // ---------------------------------------------------------------- forcing
    // Roma-Peskin-Berger 3-point smoothed delta, weight per axis (Forcing.hpp:18-35)
    static T weight(T xin) {
        T w = 0.;
        T xa = std::fabs(xin);
        if (xa <= 1.5) {
            if (xa <= 0.5) w = (1.0 / 3.0) * (1.0 + std::sqrt(-3.0 * xa * xa + 1.0));
            else w = (1.0 / 6.0) * (5.0 - (3.0 * xa) - std::sqrt((-3.0 * ((1.0 - xa) * (1.0 - xa))) + 1.0));
        }
        return w;
    }

    void forcing(const Point* geom, int64_t n, double alfaIn, double betaIn) override {
        const T alfa = T(alfaIn), beta = T(betaIn);
        for (int64_t i = 1; i <= xg1; i++)
            for (int64_t j = 1; j <= yg1; j++)
                for (int64_t k = 1; k <= zg1; k++) O[index(i, j, k)] = 0;

        T q[18];
        for (int64_t p = 0; p < n; p++) {
            const Point& g = geom[p];
            const T iFrac = T(g.iFrac), kFrac = T(g.kFrac);
            const T uD = T(g.uDelta), vD = T(g.vDelta), wD = T(g.wDelta);
            T ppp[3][3];
            for (int kk = -1; kk <= 1; kk++)
                for (int ii = -1; ii <= 1; ii++) {
                    T hx = -iFrac + T(ii);
                    T hz = -kFrac + T(kk);
                    ppp[ii + 1][kk + 1] = weight(hx) * weight(hz);
                }
            int64_t i = g.i + ghost, j = g.j + ghost, k = g.k + ghost;
            T rhoSum = 0.0, xSum = 0.0, ySum = 0.0, zSum = 0.0;
            for (int k1 = -1; k1 <= 1; k1++)
                for (int i1 = -1; i1 <= 1; i1++) {
                    int64_t i2 = i + i1, k2 = k + k1;
                    if (i2 == 0) i2 = xg1;
                    if (i2 == xg0) i2 = 1;
                    if (k2 == 0) k2 = zg1;
                    if (k2 == zg0) k2 = 1;
                    readCell(i2, j, k2, q);
                    int64_t c = index(i2, j, k2);
                    T rho = q[M01];
                    T xx = q[M02] + 0.5 * F[3 * c + 0];
                    T yy = q[M03] + 0.5 * F[3 * c + 1];
                    T zz = q[M04] + 0.5 * F[3 * c + 2];
                    rhoSum += ppp[i1 + 1][k1 + 1] * rho;
                    xSum += ppp[i1 + 1][k1 + 1] * xx;
                    ySum += ppp[i1 + 1][k1 + 1] * yy;
                    zSum += ppp[i1 + 1][k1 + 1] * zz;
                }
            xSum -= rhoSum * uD;
            ySum -= rhoSum * vD;
            zSum -= rhoSum * wD;
            for (int k1 = -1; k1 <= 1; k1++)
                for (int i1 = -1; i1 <= 1; i1++) {
                    int64_t i2 = i + i1, k2 = k + k1;
                    if (i2 == 0) i2 = xg1;
                    if (i2 == xg0) i2 = 1;
                    if (k2 == 0) k2 = zg1;
                    if (k2 == zg0) k2 = 1;
                    int64_t c = index(i2, j, k2);
                    T fx = F[3 * c + 0], fy = F[3 * c + 1], fz = F[3 * c + 2];
                    F[3 * c + 0] = alfa * fx - beta * ppp[i1 + 1][k1 + 1] * xSum;
                    F[3 * c + 1] = alfa * fy - beta * ppp[i1 + 1][k1 + 1] * ySum;
                    F[3 * c + 2] = alfa * fz - beta * ppp[i1 + 1][k1 + 1] * zSum;
                    O[c] = 1;
                }
        }
        for (int64_t i = 1; i <= xg1; i++)
            for (int64_t j = 1; j <= yg1; j++)
                for (int64_t k = 1; k <= zg1; k++) {
                    int64_t c = index(i, j, k);
                    if (O[c] == 0) { F[3 * c + 0] = 0.0; F[3 * c + 1] = 0.0; F[3 * c + 2] = 0.0; }
                    else O[c] = 0;
                }
    }

    // Cell-accessor access for the tests, over all interior cells in i,j,k order
    // with 18 values per cell: op 0 reads (AccessField::read), op 1 writes as
    // moments (writeMoments), op 2 writes as post-collision populations (write).
    void cells(int op, void* buf) override {
        T* all = static_cast<T*>(buf);
        size_t p = 0;
        for (int64_t i = 1; i <= xg1; i++)
            for (int64_t j = 1; j <= yg1; j++)
                for (int64_t k = 1; k <= zg1; k++, p += 18) {
                    if (op == 0) readCell(i, j, k, &all[p]);
                    else if (op == 1) writeMomentsCell(i, j, k, &all[p]);
                    else writeCell(i, j, k, &all[p]);
                }
    }

    // doubleResolutionFullCU + copyFieldsWithScaling + indexPreviousResolution
    // (ComputeUnit.hpp:229-316, 480-502): every cell of the doubled grid, ghosts
    // included, takes Q and Nu of old cell (i/2 + i%2, j/2 + j%2, k/2 + k%2);
    // F, O are zero (initialise(0)).  The reference copies 19 slots of an
    // 18-slot record; the 19th write lands in the next cell's slot 0 (IJKL,
    // rewritten by that cell's own copy) or past the array (LIJK), so inside
    // the array the result is the 18-slot copy done here.
    void doubleResolution() override {
        const int64_t f = 2;
        const std::vector<T> oQ = Q, oNu = Nu;
        const size_t osize = size;
        x *= f; y *= f; z *= f;
        xg = x + 2 * ghost; yg = y + 2 * ghost; zg = z + 2 * ghost;
        xg0 = xg - 1; yg0 = yg - 1; zg0 = zg - 1;
        xg1 = xg - 2; yg1 = yg - 2; zg1 = zg - 2;
        size = size_t(xg) * yg * zg;
        Q.assign(size * 18, T(0));
        F.assign(size * 3, T(0));
        Nu.assign(size, T(0));
        O.assign(size, 0);
        const int64_t fromYG = y / f + ghost * 2, fromZG = z / f + ghost * 2;
        for (int64_t i = 0; i <= xg0; i++)
            for (int64_t j = 0; j <= yg0; j++)
                for (int64_t k = 0; k <= zg0; k++) {
                    const int64_t fi = i / f + i % f, fj = j / f + j % f, fk = k / f + k % f;
                    const size_t from = size_t(fi * (fromYG * fromZG) + fj * fromZG + fk);
                    const int64_t n = index(i, j, k);
                    for (int l = 0; l < 18; l++)
                        slot(n, l) = LAYOUT == IJKL ? oQ[from * 18 + l] : oQ[size_t(l) * osize + from];
                    Nu[n] = oNu[from];
                }
    }

    // savePlaneXZ<float, 4> (ComputeUnitOutput.hpp:354-447): records of the
    // cells of plane j = cut (i, k interior, i-major) that are not excluded:
    // {i0+i-1, j-1, k0+k-1} and the first 4 raw slots of the cell (the storage
    // at index(i,j,k), not the streaming accessor); F records where F is not
    // "not zero" (Force::isNotZero, QVec.hpp:54-56) means all three components
    // non-zero, carry {i0+i-1, j, k0+k-1} (no -1 on j, as the reference writes it).
    struct Rec4 { uint16_t i, j, k; float q[4]; };
    struct Rec3 { uint16_t i, j, k; float q[3]; };
    void planeXZ(int64_t cut, const uint8_t* exclude, int64_t i0, int64_t k0, void* qv, int64_t* nq, void* fv,
                 int64_t* nf) override {
        Rec4* q = static_cast<Rec4*>(qv);
        Rec3* fr = static_cast<Rec3*>(fv);
        int64_t a = 0, b = 0;
        for (int64_t i = 1; i <= xg1; i++)
            for (int64_t k = 1; k <= zg1; k++) {
                const int64_t n = index(i, cut, k);
                if (exclude && exclude[n]) continue;
                Rec4 r{};
                r.i = uint16_t(i0 + i - 1); r.j = uint16_t(cut - 1); r.k = uint16_t(k0 + k - 1);
                for (int l = 0; l < 4; l++) r.q[l] = float(slot(n, l));
                q[a++] = r;
                if (F[3 * n] != 0.0 && F[3 * n + 1] != 0.0 && F[3 * n + 2] != 0.0) {
                    Rec3 s{};
                    s.i = uint16_t(i0 + i - 1); s.j = uint16_t(cut); s.k = uint16_t(k0 + k - 1);
                    for (int l = 0; l < 3; l++) s.q[l] = float(F[3 * n + l]);
                    fr[b++] = s;
                }
            }
        *nq = a;
        *nf = b;
    }

    // calcVorticityXZ / calcVorticityXY (ComputeUnitOutput.hpp:113-160,
    // 198-245) before the picture is made: log|curl u|^2 from central
    // differences of QVec::velocity() (no force term) of the accessor reads
    // around the cell; excluded cells NaN.  axis 0: plane j = cut, x*z values
    // (i-major); axis 1: plane k = cut, x*y values (i-major).
    void vorticityXZ(int64_t cut, const uint8_t* exclude, void* outv) override {
        if constexpr (std::is_floating_point<T>::value) vorticity(0, cut, exclude, static_cast<T*>(outv));
    }
    void vorticityXY(int64_t cut, const uint8_t* exclude, void* outv) override {
        if constexpr (std::is_floating_point<T>::value) vorticity(1, cut, exclude, static_cast<T*>(outv));
    }
    void vorticity(int axis, int64_t cut, const uint8_t* exclude, T* out) {
        T qa[18], qb[18];
        auto vel = [&](int64_t i, int64_t j, int64_t k, T& ux, T& uy, T& uz) {
            // the twisted read of a far-side ghost cell leaves the array in the
            // reference (home = ghost + 1): undefined there, NaN here
            if (STREAM == ESOTWIST && (i == xg1 + 1 || j == yg1 + 1 || k == zg1 + 1)) {
                ux = uy = uz = T(NAN);
                return;
            }
            readCell(i, j, k, qa);
            ux = qa[M02] / qa[M01];
            uy = qa[M03] / qa[M01];
            uz = qa[M04] / qa[M01];
        };
        (void)qb;
        const int64_t nb = axis == 0 ? zg1 : yg1;
        for (int64_t i = 1; i <= xg1; i++)
            for (int64_t c = 1; c <= nb; c++) {
                const int64_t j = axis == 0 ? cut : c, k = axis == 0 ? c : cut;
                T* o = &out[(i - 1) * nb + (c - 1)];
                const int64_t n = index(i, j, k);
                if (exclude && exclude[n]) { *o = T(NAN); continue; }
                T x5, y5, z5, x6, y6, z6, x3, y3, z3, x4, y4, z4, x1, y1, z1, x2, y2, z2;
                vel(i, j, k + 1, x5, y5, z5);
                vel(i, j, k - 1, x6, y6, z6);
                const T uxy = T(0.5) * (x5 - x6);
                vel(i, j + 1, k, x3, y3, z3);
                vel(i, j - 1, k, x4, y4, z4);
                const T uxz = T(0.5) * (x3 - x4);
                vel(i + 1, j, k, x1, y1, z1);
                vel(i - 1, j, k, x2, y2, z2);
                const T uyx = T(0.5) * (y1 - y2);
                const T uyz = T(0.5) * (y3 - y4);
                const T uzx = T(0.5) * (z1 - z2);
                const T uzy = T(0.5) * (z5 - z6);
                const T uxyuyx = uxy - uyx;
                const T uyzuzy = uyz - uzy;
                const T uzxuxz = uzx - uxz;
                *o = T(std::log(T(uyzuzy * uyzuzy + uzxuxz * uzxuxz + uxyuyx * uxyuyx)));
            }
    }

    void* field(int which, size_t* bytes) override {
        switch (which) {
            case 0: *bytes = Q.size() * sizeof(T); return Q.data();
            case 1: *bytes = F.size() * sizeof(T); return F.data();
            case 2: *bytes = Nu.size() * sizeof(T); return Nu.data();
            case 3: *bytes = O.size(); return O.data();
        }
        *bytes = 0;
        return nullptr;
    }
};

// Pure data-movement instantiation for the reference's integer-tag tests
// (tests/StreamingSimpleTest.cpp uses T = unsigned long): collision is not defined.
template <int LAYOUT, int STREAM>
struct TagCU : CU<uint64_t, LAYOUT, STREAM> {
    using CU<uint64_t, LAYOUT, STREAM>::CU;
    void collision() override {}
    void moments() override {}
    void forcing(const Point*, int64_t, double, double) override {}
    void initialise(double) override {}
};

template <typename T>
Flow<T> makeFlow(const double* v, int useLES) {
    Flow<T> f;
    f.initialRho = T(v[0]); f.reMNonDimensional = T(v[1]); f.uav = T(v[2]); f.cs0 = T(v[3]);
    f.g3 = T(v[4]); f.nu = T(v[5]); f.fx0 = T(v[6]); f.fy0 = T(v[7]); f.fz0 = T(v[8]);
    f.Re_m = T(v[9]); f.Re_f = T(v[10]); f.uf = T(v[11]); f.alpha = T(v[12]); f.beta = T(v[13]);
    f.useLES = useLES != 0;
    return f;
}

template <typename T>
Base* make(int layout, int stream, int64_t x, int64_t y, int64_t z, int64_t g, const Flow<T>& f) {
    if (layout == IJKL && stream == SIMPLE) return new CU<T, IJKL, SIMPLE>(x, y, z, g, f);
    if (layout == IJKL && stream == ESOTWIST) return new CU<T, IJKL, ESOTWIST>(x, y, z, g, f);
    if (layout == LIJK && stream == SIMPLE) return new CU<T, LIJK, SIMPLE>(x, y, z, g, f);
    return new CU<T, LIJK, ESOTWIST>(x, y, z, g, f);
}

}  // namespace orc

extern "C" {

// dtype: 0 float, 1 double, 2 uint64 tags (data movement only)
// flow: initialRho, reMNonDimensional, uav, cs0, g3, nu, fx0, fy0, fz0, Re_m, Re_f, uf, alpha, beta
void* orc_create(int dtype, int layout, int streaming, int64_t x, int64_t y, int64_t z, int64_t ghost,
                 const double* flow, int useLES) {
    using namespace orc;
    static const double dflt[14] = {8.0, 7300.0, 0.1, 0.12, 0.8, 0.0, 0, 0, 0, 0, 0, 0, 0.97, 1.9};
    const double* v = flow ? flow : dflt;
    if (dtype == 0) return make<float>(layout, streaming, x, y, z, ghost, makeFlow<float>(v, useLES));
    if (dtype == 1) return make<double>(layout, streaming, x, y, z, ghost, makeFlow<double>(v, useLES));
    Flow<uint64_t> f{};
    if (layout == IJKL && streaming == SIMPLE) return new TagCU<IJKL, SIMPLE>(x, y, z, ghost, f);
    if (layout == IJKL) return new TagCU<IJKL, ESOTWIST>(x, y, z, ghost, f);
    if (streaming == SIMPLE) return new TagCU<LIJK, SIMPLE>(x, y, z, ghost, f);
    return new TagCU<LIJK, ESOTWIST>(x, y, z, ghost, f);
}
void orc_destroy(void* h) { delete static_cast<orc::Base*>(h); }
// threads for collision and moments (1 = the reference's single-threaded path)
void orc_set_threads(int n) { omp_set_num_threads(n < 1 ? 1 : n); }
int orc_get_threads(void) { return omp_get_max_threads(); }
void orc_initialise(void* h, double rho) { static_cast<orc::Base*>(h)->initialise(rho); }
void orc_collision(void* h) { static_cast<orc::Base*>(h)->collision(); }
void orc_moments(void* h) { static_cast<orc::Base*>(h)->moments(); }
void orc_stream_push(void* h) { static_cast<orc::Base*>(h)->push(); }
void orc_stream_pull(void* h) { static_cast<orc::Base*>(h)->pull(); }
void orc_bounce_back(void* h) { static_cast<orc::Base*>(h)->bounceBack(); }
void orc_forcing(void* h, const orc::Point* pts, int64_t n, double alfa, double beta) {
    static_cast<orc::Base*>(h)->forcing(pts, n, alfa, beta);
}
// field: 0 Q, 1 F, 2 Nu, 3 O.  Copies raw bytes in the CU's in-memory layout.
int64_t orc_field_bytes(void* h, int field) {
    size_t b = 0;
    static_cast<orc::Base*>(h)->field(field, &b);
    return (int64_t)b;
}
void orc_get(void* h, int field, void* out) {
    size_t b = 0;
    void* p = static_cast<orc::Base*>(h)->field(field, &b);
    std::memcpy(out, p, b);
}
void orc_set(void* h, int field, const void* in) {
    size_t b = 0;
    void* p = static_cast<orc::Base*>(h)->field(field, &b);
    std::memcpy(p, in, b);
}
int orc_even_step(void* h) { return static_cast<orc::Base*>(h)->even() ? 1 : 0; }
void orc_set_even_step(void* h, int e) { static_cast<orc::Base*>(h)->even() = (e != 0); }
int64_t orc_dim(void* h, int which) { return static_cast<orc::Base*>(h)->dims(which); }
void orc_double_resolution(void* h) { static_cast<orc::Base*>(h)->doubleResolution(); }
void orc_set_flow(void* h, const double* flow, int useLES) { static_cast<orc::Base*>(h)->setFlow(flow, useLES); }
// records of 24 bytes {u16 i, j, k; f32 q[4]} and 20 bytes {u16 i, j, k; f32 q[3]}; x*z capacity each
void orc_plane_xz(void* h, int64_t cut, const uint8_t* exclude, int64_t i0, int64_t k0, void* q, int64_t* nq,
                  void* f, int64_t* nf) {
    static_cast<orc::Base*>(h)->planeXZ(cut, exclude, i0, k0, q, nq, f, nf);
}
void orc_vorticity(void* h, int axis, int64_t cut, const uint8_t* exclude, void* out) {
    if (axis == 0) static_cast<orc::Base*>(h)->vorticityXZ(cut, exclude, out);
    else static_cast<orc::Base*>(h)->vorticityXY(cut, exclude, out);
}
void orc_cells(void* h, int op, void* buf) { static_cast<orc::Base*>(h)->cells(op, buf); }

// calcNuAndRe_m (FlowParams.hpp:65-70) in float: returns nu, writes Re_m.
float orc_calc_nu_f32(float reMNonDimensional, float uav, int impellerBladeOuterRadius, float* reM) {
    float Re_m = reMNonDimensional * M_PI / 2.0;
    float nu = uav * static_cast<float>(impellerBladeOuterRadius) / Re_m;
    if (reM) *reM = Re_m;
    return nu;
}
double orc_calc_nu_f64(double reMNonDimensional, double uav, int impellerBladeOuterRadius, double* reM) {
    double Re_m = reMNonDimensional * M_PI / 2.0;
    double nu = uav * static_cast<double>(impellerBladeOuterRadius) / Re_m;
    if (reM) *reM = Re_m;
    return nu;
}
float orc_weight_f32(float x) { return orc::CU<float, 0, 0>::weight(x); }
double orc_weight_f64(double x) { return orc::CU<double, 0, 0>::weight(x); }

// Whole moments / collision of one cell, for closed-form checks.
void orc_moments_of_f64(const double* q, double* m) { orc::CU<double, 0, 0>::momentsOf(q, m); }
void orc_moments_of_f32(const float* q, float* m) { orc::CU<float, 0, 0>::momentsOf(q, m); }

}  // extern "C"
