// State of a ComputeUnit handle (private to the library: capi.cpp, batch.cpp, server.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tdlb.h"
#include "kernels.h"

namespace tdlb {
namespace detail {

// records the message for tdlb_last_error() and returns code
int fail(int code, const std::string& msg);

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(TDLB_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));              \
    } while (0)

#define CHK(expr)                  \
    do {                           \
        int r_ = (expr);           \
        if (r_ != TDLB_OK) return r_; \
    } while (0)

enum Phase { PH_EXPLICIT = 0, PH_LAZY = 1 };

struct DevPoints {
    PointsDev d{};
    int64_t cap = 0;
    size_t tsize = 4;
    void release() {
        void* ptrs[] = {d.i, d.j, d.k, d.ifrac, d.kfrac, d.u, d.v, d.w};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        d = PointsDev{};
        cap = 0;
    }
    int reserve(int64_t n, size_t ts) {
        if (n <= cap && ts == tsize) return TDLB_OK;
        release();
        tsize = ts;
        const int64_t c = std::max<int64_t>(n, 256);
        void** ip[] = {(void**)&d.i, (void**)&d.j, (void**)&d.k};
        for (void** p : ip)
            if (hipMalloc(p, sizeof(int32_t) * c) != hipSuccess) return fail(TDLB_ERR_NOMEM, "IB point buffers");
        void** fp[] = {&d.ifrac, &d.kfrac, &d.u, &d.v, &d.w};
        for (void** p : fp)
            if (hipMalloc(p, ts * c) != hipSuccess) return fail(TDLB_ERR_NOMEM, "IB point buffers");
        cap = c;
        return TDLB_OK;
    }
};

// Population arrays in LIJK: a k-row holds the interior cells k = 1..zg1 only,
// rows back to back (ghost rows in i and j included), and the two k-ghost
// cells of every row sit in a side region after the rows.  The stream-collide
// kernels never touch a k-ghost (Esotwist wraps periodically, the Simple pull
// reads interior neighbours only), so every line they move holds interior
// cells only, and at zg1 = 512 every row starts a 128 B line.  Slot planes of
// 16 MiB and more start on 1 MiB boundaries (planes of 18 concurrent streams
// whose starts fall at arbitrary offsets measured 4 % slower at 512^3), smaller
// ones on 128 B lines.  IJKL keeps the dense index.
inline void set_q_geometry(Grid& g, int dlay, size_t ssize) {
    g.qside = 0;
    if (dlay == LayoutIJKL) {
        g.qz = (uint32_t)g.zg;
        g.qyz = g.yz;
        g.qoff = 0;
        g.qsize = g.size;
        return;
    }
    g.qz = (uint32_t)g.zg1;
    g.qyz = (uint32_t)((uint64_t)g.yg * g.qz);
    g.qoff = 0xffffffffu;  // (i, j, k) at i*qyz + j*qz + k - 1 (modulo 2^32)
    const uint64_t line = 128 / ssize;  // cells of the stored type in a 128 B line
    g.qside = (uint32_t)(((uint64_t)g.xg * g.qyz + line - 1) / line * line);
    const uint64_t cells = (uint64_t)g.qside + 2ull * g.xg * g.yg;
    const uint64_t align = cells * ssize >= (16ull << 20) ? (1ull << 20) / ssize : line;
    g.qsize = (cells + align - 1) / align * align;
}

// Stencil cells of a point list with a moments cache: the sorted unique cells,
// and for every point the 9 positions of its stencil cells in that list.
// A bound tank's static set also carries its plane tasks (PlaneDesc): cells
// grouped by plane, points ordered plane by plane and level by level.
struct IndexedSet {
    uint32_t* cells = nullptr;
    int32_t* sidx = nullptr;
    void* cache = nullptr;
    PlaneDesc* planes = nullptr;
    int64_t* lev_off = nullptr;
    PlaneBlock* blocks = nullptr;
    uint32_t* bar = nullptr;
    int64_t n_cells = 0, n_points = 0;
    int32_t n_planes = 0, n_blocks = 0, n_teams = 0, max_levels = 0;
    std::vector<float> plane_cost;  // host: duration estimate of each plane's workgroups (planes order)
    std::vector<PlaneBlock> block_list;  // host copy of blocks
    std::vector<PlaneDesc> plane_list;   // host copy of planes
    void release() {
        void* ptrs[] = {cells, sidx, cache, planes, lev_off, blocks, bar};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        cells = nullptr;
        sidx = nullptr;
        cache = nullptr;
        planes = nullptr;
        lev_off = nullptr;
        blocks = nullptr;
        bar = nullptr;
        n_cells = n_points = 0;
        n_planes = n_blocks = n_teams = max_levels = 0;
        plane_cost.clear();
        block_list.clear();
        plane_list.clear();
    }
};

// Host-side copy of points in the CU's type, structure of arrays.
struct HostPoints {
    std::vector<int32_t> i, j, k;
    std::vector<uint8_t> ifrac, kfrac, u, v, w;
    void resize(size_t n, size_t ts) {
        i.resize(n); j.resize(n); k.resize(n);
        ifrac.resize(n * ts); kfrac.resize(n * ts); u.resize(n * ts); v.resize(n * ts); w.resize(n * ts);
    }
};

template <typename T>
void put(std::vector<uint8_t>& v, size_t idx, double x) {
    T t = (T)x;
    std::memcpy(v.data() + idx * sizeof(T), &t, sizeof(T));
}

}  // namespace detail
}  // namespace tdlb

using namespace tdlb;
using namespace tdlb::detail;

struct tdlb_cu {
    tdlb_cu_params prm{};
    tdlb_flow flow{};
    Grid g{};
    int64_t xg = 0, yg = 0, zg = 0;
    size_t tsize = 4;  // arithmetic type: F, Nu, points, explicit Q
    size_t ssize = 4;  // stored population type
    hipStream_t stream = nullptr;
    int device = 0;

    void* Q[2] = {nullptr, nullptr};
    int cur = 0;
    void* M = nullptr;   // explicit Q of the reduced-precision storage (float)
    float rho0 = 1.0f;   // rest density of the deviation storage
    double* red = nullptr;  // FP64 partial sums + result of tdlb_cu_reductions
    void* F = nullptr;
    void* Nu = nullptr;
    uint8_t* O = nullptr;
    uint64_t* fmask[2] = {nullptr, nullptr};
    uint64_t* gmask = nullptr;
    uint64_t* static_bits = nullptr;
    uint64_t* static_base = nullptr;
    int fm = 0;
    size_t mask_words = 0;

    bool even = true;
    Phase phase = PH_EXPLICIT;
    bool o_dirty = false;         // O may hold non-zero interior values
    bool static_in_prev = false;  // every static stencil cell is set in fmask[fm]

    // bound Rushton tank
    bool bound = false;
    tdlb_rushton_params rp{};
    DevPoints pts;                       // static points [0, n_static), blades after
    IndexedSet static_set;               // stencil cells of the static points
    void* rot_cache = nullptr;           // 9 moment entries per blade point
    int64_t n_static = 0, n_rot = 0;
    int32_t n_blade_classes = 0;         // blade colour classes (one forcing level each)
    BladeLayout blade{};
    double alfa = 0.97, beta = 1.9;

    // one-launch Esotwist step (launch_eso_pipe), set up at bind when the tank allows it
    bool pipe = false;
    bool xcc_rr = false;  // launches dispatch workgroup b to XCD b mod kXcds (checked at bind)
    int32_t* rowcell = nullptr;
    uint32_t* pctr = nullptr;

    DevPoints api_pts;  // points of the generic forcing call
    IndexedSet api_set;

    // event timing of tdlb_cu_run (tdlb_cu_profile)
    bool prof = false;
    std::vector<hipEvent_t> events;        // pool, recorded in pairs
    std::vector<std::pair<int, int>> coll, force;
    size_t ev_used = 0;

    // cells per slot of the kernels' population arrays (rows padded in LIJK)
    size_t q_cells() const { return dlay() == LayoutIJKL ? (size_t)g.size : (size_t)g.qsize; }
    size_t qb_bytes() const { return (size_t)g.size * kQ * tsize; }  // Q as the reference holds it (boundary)
    size_t q_bytes() const { return q_cells() * kQ * tsize; }        // explicit Q on the device
    size_t qs_bytes() const { return q_cells() * kQ * ssize; }       // one population lattice
    // Q at the boundary differs from the device array (layout or row padding)
    bool q_converted() const {
        return blay() != dlay() || g.qz != (uint32_t)g.zg || g.qoff != 0 || g.qside != 0 ||
               (size_t)g.qsize != (size_t)g.size;
    }
    size_t f_bytes() const { return (size_t)g.size * 3 * tsize; }
    size_t nu_bytes() const { return (size_t)g.size * (shifted() ? ssize : tsize); }  // Nu on the device
    size_t nub_bytes() const { return (size_t)g.size * tsize; }                       // Nu at the boundary
    size_t mask_bytes() const { return mask_words * sizeof(uint64_t); }
    bool simple() const { return prm.streaming == TDLB_STREAM_SIMPLE; }
    // layout of the lattice in HBM (what the kernels address) and of Q at the
    // boundary (download, upload, checkpoint); they differ for TDLB_LAYOUT_IJKL_SOA
    int dlay() const { return prm.layout == TDLB_LAYOUT_IJKL ? LayoutIJKL : LayoutLIJK; }
    int blay() const { return prm.layout == TDLB_LAYOUT_LIJK ? LayoutLIJK : LayoutIJKL; }
    bool shifted() const {
        return prm.precision == TDLB_PREC_F16 || prm.precision == TDLB_PREC_BF16 || prm.precision == TDLB_PREC_E3M12;
    }
    int prec() const { return prm.precision; }            // kernels on the stored populations
    int eprec() const { return prm.precision == TDLB_PREC_F64 ? 1 : 0; }  // kernels on the explicit Q
    void* eq() const { return shifted() ? M : Q[cur]; }   // the explicit Q array

    // the forcing planes of the bound tank
    ForceGeom fgeom() const {
        ForceGeom f{};
        f.pts = pts.d;
        f.n_static = n_static;
        f.n_rot = n_rot;
        f.cells = static_set.cells;
        f.sidx = static_set.sidx;
        f.scache = static_set.cache;
        f.rcache = rot_cache;
        f.planes = static_set.planes;
        f.n_planes = static_set.n_planes;
        f.lev_off = static_set.lev_off;
        f.blocks = static_set.blocks;
        f.n_blocks = static_set.n_blocks;
        f.bar = static_set.bar;
        f.local_teams = xcc_rr ? 1 : 0;
        f.blade = blade;
        return f;
    }

    CollideArgs cargs() const {
        CollideArgs c;
        c.nu = flow.nu;
        c.cs0 = flow.cs0;
        c.g3 = flow.g3;
        c.use_les = flow.use_les;
        return c;
    }
};
