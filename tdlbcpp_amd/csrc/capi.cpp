// C ABI of the MI355X ComputeUnit (declared in include/tdlb.h).
//
// State of a handle.  Q[cur] holds the ComputeUnit's Q array in the reference's
// layout.  Two phases:
//   explicit  Q[cur] is byte for byte what the reference's Q would hold;
//   lazy      left by the fused main-loop path: Q[cur] holds post-collision
//             populations (Simple) or streamed populations (Esotwist) and the
//             forcing of the step has been applied; the reference's Q is
//             moments() of the streamed populations.  Any call that reads or
//             changes Q as the reference sees it first materialises it.
// Simple uses two lattices: the reference's in-place push equals a pull into
// the other buffer (StreamingSimple.hpp:76-126).  In the lazy phase the ghost
// layers of both buffers hold the slots the reference never rewrites from the
// interior (bounce-back rims, never-written slots); materialising writes the
// rest of the ghost layer (pushed and bounced-back slots) from the interior.
//
// F occupancy: fmask[fm] has a bit per cell where F may be non-zero; F of every
// other cell is zero in the reference even if the array still holds an old
// value (forcing's "F = 0 where O == 0", Forcing.hpp:141-151).
//
// Reduced-precision storage (TDLB_PREC_F16 / BF16 / E3M12): the lazy phase
// keeps the populations as 16-bit scaled deviations k_d (f_d - w_d rho0) and
// computes in float; the explicit Q is a float array M (the FP32 reference
// layout), written by materialise and read by the first collision of a run.
// The per-call operations (collision, streaming, ...) need fp32 or fp64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/tdlb.h"
#include "handle.h"
#include "kernels.h"
#include "rushton.h"

namespace tdlb {
namespace detail {
thread_local std::string g_err;
int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
}  // namespace detail
}  // namespace tdlb

namespace {


int set_device(tdlb_cu* cu) {
    HIPCHK(hipSetDevice(cu->device));
    return TDLB_OK;
}

// ---- scratch of tdlb_cu_reductions_noalloc: the partial sums of the fixed
// reduction grid (kReduceBlocks rows of kReduceVals) and the result row,
// 65,600 bytes.  Every GPU the library runs on holds its own copy until the
// process ends; the host never passes the array's name to a kernel (in host
// code it is a stand-in, not a device address) but the address
// hipGetSymbolAddress returns on each device.
constexpr size_t kScratchDoubles = (size_t)(kReduceBlocks + 1) * kReduceVals;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kScratchDevices] = {};

// the address of scratch on `device`, resolved once per device with that
// device current (tdlb_cu_create calls this)
int resolve_scratch(int device) {
    if (device < 0 || device >= kScratchDevices)
        return fail(TDLB_ERR_UNSUPPORTED, "device id beyond " + std::to_string(kScratchDevices - 1));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return TDLB_OK;
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(TDLB_ERR_HIP, "hipGetSymbolAddress returned no address for the reduction scratch");
    g_scratch[device] = (double*)p;
    return TDLB_OK;
}

// torque region of the reductions: cells within blade_outer_radius + 2 of the
// bound tank's axis (none when no geometry is bound)
ReduceArgs reduce_args(const tdlb_cu* cu) {
    ReduceArgs ra{};
    ra.ghost = (int)cu->prm.ghost;
    ra.r2 = -1.0;
    if (cu->bound) {
        const double r = cu->rp.blade_outer_radius + 2.0;
        ra.cx = cu->rp.center_x;
        ra.cz = cu->rp.center_z;
        ra.r2 = r * r;
    }
    return ra;
}

void release(tdlb_cu* cu) {
    void* ptrs[] = {cu->Q[0], cu->Q[1], cu->M, cu->F, cu->Nu, cu->O, cu->fmask[0], cu->fmask[1],
                    cu->gmask, cu->static_bits, cu->static_base, cu->red, cu->rowcell, cu->pctr};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    cu->pts.release();
    cu->api_pts.release();
    cu->static_set.release();
    cu->api_set.release();
    if (cu->rot_cache) (void)hipFree(cu->rot_cache);
    for (hipEvent_t e : cu->events) (void)hipEventDestroy(e);
    cu->events.clear();
    if (cu->stream) (void)hipStreamDestroy(cu->stream);
}

// record a timing event on the handle's stream; returns its pool index or -1
int prof_mark(tdlb_cu* cu) {
    if (!cu->prof) return -1;
    constexpr size_t kMaxEvents = 1 << 16;
    if (cu->ev_used == cu->events.size()) {
        if (cu->events.size() >= kMaxEvents) return -1;
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        cu->events.push_back(e);
    }
    hipEvent_t e = cu->events[cu->ev_used];
    if (hipEventRecord(e, cu->stream) != hipSuccess) return -1;
    return (int)cu->ev_used++;
}

// ---- phase handling
// The explicit Q of the 16-bit storage is allocated on first use.  A fresh one
// holds what the reference's Q holds in the slots the lattice never rewrites:
// the initial moments, and for Simple the bounce-back rims settled from them.
int ensure_m(tdlb_cu* cu) {
    if (cu->M) return TDLB_OK;
    if (hipMalloc(&cu->M, cu->q_bytes()) != hipSuccess) {
        cu->M = nullptr;
        return fail(TDLB_ERR_NOMEM, "explicit Q of the 16-bit storage");
    }
    HIPCHK(launch_fill_q(cu->eprec(), cu->dlay(), cu->M, cu->g, cu->rho0, cu->stream));
    if (cu->simple()) HIPCHK(launch_bounce_back((int)cu->tsize, cu->dlay(), cu->M, cu->g, cu->stream));
    return TDLB_OK;
}

int materialise(tdlb_cu* cu) {
    if (cu->phase != PH_LAZY) return TDLB_OK;
    if (cu->shifted()) CHK(ensure_m(cu));
    const auto c = cu->cargs();
    if (cu->simple()) {
        void* src = cu->Q[cu->cur];
        void* dst = cu->shifted() ? cu->M : cu->Q[1 - cu->cur];
        HIPCHK(launch_simple(cu->prec(), cu->dlay(), MODE_MOM_PULL, src, dst, dst, cu->F, cu->Nu,
                             cu->fmask[cu->fm], cu->g, c, cu->rho0, cu->stream));
        if (cu->shifted()) {
            HIPCHK(launch_ghost_fill_decode(cu->prec(), cu->dlay(), src, dst, cu->g, cu->rho0, cu->stream));
        } else {
            HIPCHK(launch_ghost_fill((int)cu->ssize, cu->dlay(), src, dst, cu->g, cu->stream));
            cu->cur = 1 - cu->cur;
        }
    } else {
        HIPCHK(launch_esotwist(cu->prec(), cu->dlay(), MODE_MOM, cu->even ? 1 : 0, cu->Q[0], cu->eq(), cu->F,
                               cu->Nu, cu->fmask[cu->fm], cu->g, c, cu->rho0, cu->stream));
    }
    cu->phase = PH_EXPLICIT;
    return TDLB_OK;
}

// Where density and momentum of a cell come from in the current phase.
struct CellSource {
    int prec, momsrc, even;
    const void* Q;
};
CellSource cell_source(const tdlb_cu* cu) {
    if (cu->phase == PH_EXPLICIT) return {cu->eprec(), MOMSRC_STORED, cu->even ? 1 : 0, cu->eq()};
    return {cu->prec(), cu->simple() ? MOMSRC_PULL_POPS : MOMSRC_ESO_POPS, cu->even ? 1 : 0, cu->Q[cu->cur]};
}

// the per-call operations work on the reference's Q directly
int need_full_precision(const tdlb_cu* cu, const char* what) {
    if (cu->shifted())
        return fail(TDLB_ERR_UNSUPPORTED, std::string(what) + " needs precision fp32 or fp64 (the 16-bit "
                                          "population storage runs through tdlb_cu_run only)");
    return TDLB_OK;
}

// ---- IB levels: a point's level is one more than the highest level of any
// earlier point whose 3x3 (i,k) stencil shares a cell with it, so the points of
// a level are independent and levels run in order (Forcing.hpp:73-135 is a
// sequential loop over the list).
void stencil_cells(const tdlb_cu* cu, int64_t i, int64_t j, int64_t k, uint64_t* cells) {
    int t = 0;
    for (int k1 = -1; k1 <= 1; k1++)
        for (int i1 = -1; i1 <= 1; i1++) {
            int64_t i2 = i + i1, k2 = k + k1;
            if (i2 == 0) i2 = cu->g.xg1;
            if (i2 == cu->g.xg1 + 1) i2 = 1;
            if (k2 == 0) k2 = cu->g.zg1;
            if (k2 == cu->g.zg1 + 1) k2 = 1;
            cells[t++] = ((uint64_t)i2 * cu->yg + (uint64_t)j) * cu->zg + (uint64_t)k2;
        }
}

int check_points(const tdlb_cu* cu, const tdlb_ib_point* p, int64_t n) {
    const int64_t gh = cu->prm.ghost;
    for (int64_t q = 0; q < n; q++) {
        const int64_t i = p[q].i + gh, j = p[q].j + gh, k = p[q].k + gh;
        if (i < 1 || i > cu->g.xg1 || j < 1 || j > cu->g.yg1 || k < 1 || k > cu->g.zg1)
            return fail(TDLB_ERR_INVALID, "IB point " + std::to_string(q) + " outside the interior: (" +
                                              std::to_string(p[q].i) + "," + std::to_string(p[q].j) + "," +
                                              std::to_string(p[q].k) + ")");
    }
    return TDLB_OK;
}

// levels of a list; returns the permutation ordering points by level (stable)
// and CSR offsets (offsets[0] = 0 .. offsets[L] = n)
void compute_levels(const tdlb_cu* cu, const tdlb_ib_point* p, int64_t n, std::vector<int64_t>& order,
                    std::vector<int64_t>& offsets, int base_level = 0) {
    std::unordered_map<uint64_t, int32_t> last;
    last.reserve((size_t)n * 4 + 16);
    std::vector<int32_t> lev(n);
    int32_t maxl = -1;
    uint64_t cells[9];
    const int64_t gh = cu->prm.ghost;
    for (int64_t q = 0; q < n; q++) {
        stencil_cells(cu, p[q].i + gh, p[q].j + gh, p[q].k + gh, cells);
        int32_t l = base_level;
        for (uint64_t c : cells) {
            auto it = last.find(c);
            if (it != last.end()) l = std::max(l, it->second + 1);
        }
        for (uint64_t c : cells) last[c] = l;
        lev[q] = l;
        maxl = std::max(maxl, l);
    }
    const int nl = maxl + 1 - base_level;
    offsets.assign(std::max(nl, 0) + 1, 0);
    for (int64_t q = 0; q < n; q++) offsets[lev[q] - base_level + 1]++;
    for (int l = 0; l < nl; l++) offsets[l + 1] += offsets[l];
    order.resize(n);
    std::vector<int64_t> fill(offsets.begin(), offsets.end() - 1);
    for (int64_t q = 0; q < n; q++) order[fill[lev[q] - base_level]++] = q;
}

void to_host_points(const tdlb_cu* cu, const tdlb_ib_point* p, const std::vector<int64_t>& order, HostPoints& h) {
    const size_t n = order.size();
    h.resize(n, cu->tsize);
    const int64_t gh = cu->prm.ghost;
    for (size_t a = 0; a < n; a++) {
        const tdlb_ib_point& q = p[order[a]];
        h.i[a] = (int32_t)(q.i + gh);
        h.j[a] = (int32_t)(q.j + gh);
        h.k[a] = (int32_t)(q.k + gh);
        if (cu->tsize == 4) {
            put<float>(h.ifrac, a, q.i_frac); put<float>(h.kfrac, a, q.k_frac);
            put<float>(h.u, a, q.u_delta); put<float>(h.v, a, q.v_delta); put<float>(h.w, a, q.w_delta);
        } else {
            put<double>(h.ifrac, a, q.i_frac); put<double>(h.kfrac, a, q.k_frac);
            put<double>(h.u, a, q.u_delta); put<double>(h.v, a, q.v_delta); put<double>(h.w, a, q.w_delta);
        }
    }
}

int upload_points(tdlb_cu* cu, DevPoints& dp, int64_t offset, const HostPoints& h) {
    const size_t n = h.i.size();
    if (n == 0) return TDLB_OK;
    const size_t ts = cu->tsize;
    HIPCHK(hipMemcpyAsync(dp.d.i + offset, h.i.data(), n * 4, hipMemcpyHostToDevice, cu->stream));
    HIPCHK(hipMemcpyAsync(dp.d.j + offset, h.j.data(), n * 4, hipMemcpyHostToDevice, cu->stream));
    HIPCHK(hipMemcpyAsync(dp.d.k + offset, h.k.data(), n * 4, hipMemcpyHostToDevice, cu->stream));
    HIPCHK(hipMemcpyAsync((char*)dp.d.ifrac + offset * ts, h.ifrac.data(), n * ts, hipMemcpyHostToDevice, cu->stream));
    HIPCHK(hipMemcpyAsync((char*)dp.d.kfrac + offset * ts, h.kfrac.data(), n * ts, hipMemcpyHostToDevice, cu->stream));
    HIPCHK(hipMemcpyAsync((char*)dp.d.u + offset * ts, h.u.data(), n * ts, hipMemcpyHostToDevice, cu->stream));
    HIPCHK(hipMemcpyAsync((char*)dp.d.v + offset * ts, h.v.data(), n * ts, hipMemcpyHostToDevice, cu->stream));
    HIPCHK(hipMemcpyAsync((char*)dp.d.w + offset * ts, h.w.data(), n * ts, hipMemcpyHostToDevice, cu->stream));
    HIPCHK(hipStreamSynchronize(cu->stream));  // host staging buffers end with the caller
    return TDLB_OK;
}

// Build the unique-cell moment cache description of points (in device order)
// and upload it; sidx entries are indexed by the point's position + offset.
int build_indexed_set(tdlb_cu* cu, const tdlb_ib_point* p, const std::vector<int64_t>& order, IndexedSet& set) {
    set.release();
    const int64_t n = (int64_t)order.size();
    if (n == 0) return TDLB_OK;
    std::vector<uint32_t> all(9 * n);
    uint64_t cells[9];
    const int64_t gh = cu->prm.ghost;
    for (int64_t a = 0; a < n; a++) {
        const tdlb_ib_point& q = p[order[a]];
        stencil_cells(cu, q.i + gh, q.j + gh, q.k + gh, cells);
        for (int s = 0; s < 9; s++) all[9 * a + s] = (uint32_t)cells[s];
    }
    std::vector<uint32_t> uniq(all);
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    std::vector<int32_t> sidx(9 * n);
    for (int64_t e = 0; e < 9 * n; e++)
        sidx[e] = (int32_t)(std::lower_bound(uniq.begin(), uniq.end(), all[e]) - uniq.begin());
    set.n_cells = (int64_t)uniq.size();
    set.n_points = n;
    if (hipMalloc((void**)&set.cells, uniq.size() * 4) != hipSuccess ||
        hipMalloc((void**)&set.sidx, sidx.size() * 4) != hipSuccess ||
        hipMalloc(&set.cache, uniq.size() * 4 * cu->tsize) != hipSuccess)
        return fail(TDLB_ERR_NOMEM, "IB moment cache");
    HIPCHK(hipMemcpyAsync(set.cells, uniq.data(), uniq.size() * 4, hipMemcpyHostToDevice, cu->stream));
    HIPCHK(hipMemcpyAsync(set.sidx, sidx.data(), sidx.size() * 4, hipMemcpyHostToDevice, cu->stream));
    HIPCHK(hipStreamSynchronize(cu->stream));
    return TDLB_OK;
}

// Point ranges of one forcing pass sharing a source of stencil moments.
struct ForceSegment {
    const IndexedSet* set = nullptr;  // unique-cell cache (sidx by point index)
    void* direct_cache = nullptr;     // or 9 entries per point from cache_base
    int64_t cache_base = 0, n_points = 0;
    std::vector<std::pair<int64_t, int64_t>> levels;
};

// One forcing pass: levels of every segment in order.  prepare: point ranges
// whose stencil cells may be missing from the previous occupancy mask.  base:
// bits the new mask starts from.
int forcing_pass(tdlb_cu* cu, const DevPoints& dp, const std::vector<ForceSegment>& segs,
                 const std::vector<std::pair<int64_t, int64_t>>& prepare, const uint64_t* base, int momsrc) {
    uint64_t* prev = cu->fmask[cu->fm];
    uint64_t* next = cu->fmask[1 - cu->fm];
    const int even = cu->even ? 1 : 0;
    // stored moments live in the explicit Q, populations in the lattice
    const bool stored = momsrc == MOMSRC_STORED;
    const void* Qsrc = stored ? cu->eq() : cu->Q[cu->cur];
    const int qp = stored ? cu->eprec() : cu->prec();
    HIPCHK(launch_copy(next, base, cu->mask_bytes(), cu->stream));
    for (auto& r : prepare)
        HIPCHK(launch_force_prepare(cu->eprec(), dp.d, r.first, r.second - r.first, cu->F, prev, next, cu->g,
                                    cu->stream));
    // stencils wrap onto themselves below 3 interior cells: keep the plain kernel there
    const bool tiny = cu->g.xg1 < 3 || cu->g.zg1 < 3;
    if (!tiny) {
        for (auto& sg : segs) {
            if (sg.set)
                HIPCHK(launch_ib_moments_cells(qp, cu->dlay(), cu->prm.streaming, momsrc, even, sg.set->cells,
                                               sg.set->n_cells, sg.set->cache, Qsrc, cu->g, cu->rho0, cu->stream));
            else
                HIPCHK(launch_ib_moments_points(qp, cu->dlay(), cu->prm.streaming, momsrc, even, dp.d,
                                                sg.cache_base, sg.n_points, sg.direct_cache, Qsrc, cu->g, cu->rho0,
                                                cu->stream));
        }
    }
    for (auto& sg : segs)
        for (auto& r : sg.levels) {
            if (tiny)
                HIPCHK(launch_force_level(qp, cu->dlay(), cu->prm.streaming, momsrc, even, dp.d, r.first,
                                          r.second - r.first, cu->F, Qsrc, cu->g, cu->alfa, cu->beta, cu->rho0,
                                          cu->stream));
            else
                HIPCHK(launch_force_level_cached(cu->eprec(), dp.d, r.first, r.second - r.first,
                                                 sg.set ? sg.set->sidx : nullptr, sg.cache_base,
                                                 sg.set ? sg.set->cache : sg.direct_cache, cu->F, cu->g, cu->alfa,
                                                 cu->beta, cu->stream));
        }
    if (cu->o_dirty) {
        HIPCHK(launch_zero_interior_o(cu->O, cu->g, cu->stream));
        cu->o_dirty = false;
    }
    cu->fm = 1 - cu->fm;
    return TDLB_OK;
}

// cost (estimated microseconds) above which a plane gets a team of workgroups,
// one per this much; a wall plane of a 512^3 tank is ~100
constexpr double kTeamCost = 120.0;
constexpr int32_t kMaxTeam = 16;
// TDLB_FORCE_TEAM_COST overrides kTeamCost (tests give small tanks teams)
double team_cost() {
    const char* v = std::getenv("TDLB_FORCE_TEAM_COST");
    const double c = v ? std::atof(v) : 0.0;
    return c > 0.0 ? c : kTeamCost;
}

// Plane tasks of a bound tank's static list (ForceGeom, kernels.h).  Levels as
// compute_levels over the list order; the device order is plane by plane, then
// level by level (stable), so the points of one plane and level are
// contiguous.  Every interior plane with static or blade points is a task;
// tasks are ordered by an estimate of their duration, longest first, so the
// disk plane starts at once.
int build_plane_set(tdlb_cu* cu, const std::vector<tdlb_ib_point>& stat, const BladeEnum& be, int64_t blade_j0,
                    std::vector<int64_t>& order, IndexedSet& set) {
    set.release();
    const int64_t n = (int64_t)stat.size();
    const int64_t gh = cu->prm.ghost;
    std::vector<int64_t> lorder, loffs;
    compute_levels(cu, stat.data(), n, lorder, loffs);
    std::vector<int32_t> lev(n);
    for (size_t l = 0; l + 1 < loffs.size(); l++)
        for (int64_t a = loffs[l]; a < loffs[l + 1]; a++) lev[lorder[a]] = (int32_t)l;
    order.resize(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        return stat[a].j != stat[b].j ? stat[a].j < stat[b].j : lev[a] < lev[b];
    });
    const int n_classes = be.n_blades > 0 ? 5 * be.n_bcol : 0;
    const int64_t blade_pts = (int64_t)be.n_blades * be.n_radial;  // per blade plane
    std::vector<PlaneDesc> planes;
    std::vector<double> cost;
    std::vector<int64_t> lev_off;
    std::vector<uint32_t> cells;
    std::vector<int32_t> sidx(9 * (size_t)n);
    std::vector<uint32_t> pc;
    uint64_t sc[9];
    int32_t max_levels = 0;
    int64_t a0 = 0;
    for (int64_t j = 1; j <= cu->g.yg1; j++) {
        int64_t a1 = a0;
        while (a1 < n && stat[order[a1]].j + gh == j) a1++;
        const bool blades = n_classes > 0 && j >= blade_j0 && j < blade_j0 + be.n_planes;
        if (a1 == a0 && !blades) {
            a0 = a1;
            continue;
        }
        PlaneDesc pd{};
        pd.j = (int32_t)j;
        pd.cell_begin = (int64_t)cells.size();
        pc.clear();
        for (int64_t a = a0; a < a1; a++) {
            const tdlb_ib_point& q = stat[order[a]];
            stencil_cells(cu, q.i + gh, q.j + gh, q.k + gh, sc);
            for (int t = 0; t < 9; t++) pc.push_back((uint32_t)sc[t]);
        }
        std::vector<uint32_t> uq(pc);
        std::sort(uq.begin(), uq.end());
        uq.erase(std::unique(uq.begin(), uq.end()), uq.end());
        for (size_t e = 0; e < pc.size(); e++)
            sidx[9 * (size_t)a0 + e] =
                (int32_t)(pd.cell_begin + (std::lower_bound(uq.begin(), uq.end(), pc[e]) - uq.begin()));
        cells.insert(cells.end(), uq.begin(), uq.end());
        pd.cell_end = (int64_t)cells.size();
        const int32_t nl = a1 > a0 ? lev[order[a1 - 1]] + 1 : 0;
        pd.n_levels = nl;
        pd.lev = (int64_t)lev_off.size();
        int64_t a = a0, passes = 0;
        for (int32_t l = 0; l < nl; l++) {
            lev_off.push_back(a);
            const int64_t b = a;
            while (a < a1 && lev[order[a]] == l) a++;
            passes += (a - b + kForcePlanePoints - 1) / kForcePlanePoints;
        }
        lev_off.push_back(a1);
        max_levels = std::max(max_levels, nl);
        int64_t gather = pd.cell_end - pd.cell_begin;
        if (blades) {
            for (int c = 0; c < n_classes; c++) {
                const int64_t cls = blade_class_size(be, c) / std::max(be.n_planes, 1);
                passes += (cls + kForcePlanePoints - 1) / kForcePlanePoints;
            }
            gather += 9 * blade_pts;
        }
        planes.push_back(pd);
        cost.push_back(3.0 * (double)(passes + n_classes) + (double)gather / 1024.0);
        a0 = a1;
    }
    if (a0 != n) return fail(TDLB_ERR_INVALID, "static IB points outside the interior planes");
    // A plane whose passes would keep one workgroup far longer than a wall
    // plane (the impeller disk's) runs on a team; teams come first in the
    // launch, then the solo planes, longest first.
    std::vector<size_t> by(planes.size());
    std::iota(by.begin(), by.end(), 0);
    std::stable_sort(by.begin(), by.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
    std::vector<PlaneDesc> sorted(planes.size());
    std::vector<PlaneBlock> blocks;
    const double tcost = team_cost();
    set.plane_cost.resize(planes.size());
    int32_t n_teams = 0;
    for (size_t t = 0; t < by.size(); t++) {
        PlaneDesc pd = planes[by[t]];
        const double c = cost[by[t]];
        pd.team = (int32_t)std::min<double>(kMaxTeam, std::max(1.0, std::ceil(c / tcost)));
        pd.bar = pd.team > 1 ? n_teams++ : -1;
        sorted[t] = pd;
        set.plane_cost[t] = (float)(c / pd.team);
    }
    // a team's workgroups go kXcds apart (launches dispatch workgroup b to XCD
    // b mod kXcds), so one XCD runs the whole team; the other slots take the
    // solo planes, longest first
    {
        int64_t row[kXcds] = {0};
        std::vector<PlaneBlock> slot;
        for (size_t t = 0; t < sorted.size(); t++) {
            if (sorted[t].team == 1) continue;
            int x = 0;
            for (int y = 1; y < kXcds; y++)
                if (row[y] < row[x]) x = y;
            for (int32_t r = 0; r < sorted[t].team; r++) {
                const size_t at = (size_t)x + (size_t)kXcds * (size_t)(row[x] + r);
                if (slot.size() <= at) slot.resize(at + 1, PlaneBlock{-1, 0});
                slot[at] = PlaneBlock{(int32_t)t, r};
            }
            row[x] += sorted[t].team;
        }
        size_t next_solo = 0;
        auto solo = [&]() -> PlaneBlock {
            while (next_solo < sorted.size() && sorted[next_solo].team != 1) next_solo++;
            return next_solo < sorted.size() ? PlaneBlock{(int32_t)next_solo++, 0} : PlaneBlock{-1, 0};
        };
        for (auto& b : slot)
            if (b.plane < 0) b = solo();
        for (PlaneBlock b = solo(); b.plane >= 0; b = solo()) slot.push_back(b);
        blocks = slot;
    }
    set.n_cells = (int64_t)cells.size();
    set.n_points = n;
    set.n_planes = (int32_t)sorted.size();
    set.n_blocks = (int32_t)blocks.size();
    set.n_teams = n_teams;
    set.max_levels = max_levels;
    set.block_list = blocks;
    set.plane_list = sorted;
    auto upload = [&](void** dst, const void* src, size_t bytes) -> int {
        if (bytes == 0) return TDLB_OK;
        if (hipMalloc(dst, bytes) != hipSuccess) return fail(TDLB_ERR_NOMEM, "IB plane tables");
        HIPCHK(hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, cu->stream));
        return TDLB_OK;
    };
    CHK(upload((void**)&set.cells, cells.data(), cells.size() * 4));
    CHK(upload((void**)&set.sidx, sidx.data(), sidx.size() * 4));
    CHK(upload((void**)&set.planes, sorted.data(), sorted.size() * sizeof(PlaneDesc)));
    CHK(upload((void**)&set.lev_off, lev_off.data(), lev_off.size() * sizeof(int64_t)));
    CHK(upload((void**)&set.blocks, blocks.data(), blocks.size() * sizeof(PlaneBlock)));
    if (n_teams > 0) {
        const size_t bb = (size_t)n_teams * kBarWords * sizeof(uint32_t);
        if (hipMalloc((void**)&set.bar, bb) != hipSuccess) return fail(TDLB_ERR_NOMEM, "IB team barriers");
        HIPCHK(launch_zero(set.bar, bb, cu->stream));
    }
    if (!cells.empty() && hipMalloc(&set.cache, cells.size() * 4 * cu->tsize) != hipSuccess)
        return fail(TDLB_ERR_NOMEM, "IB moment cache");
    HIPCHK(hipStreamSynchronize(cu->stream));  // host tables go out of scope
    return TDLB_OK;
}

// Whether this process's GPU can run the one-launch step: workgroups of a
// launch must see every XCC id (each XCD's work queue needs workers); checked
// once per device.  The step is opt-in (TDLB_PIPE=1): at 512^3 its rows run
// slower than the separate stream-collide kernel.
bool pipe_device_ok(tdlb_cu* cu, bool honour_env = true) {
    static std::mutex mu;
    static std::unordered_map<int, bool> known;
    const char* env = std::getenv("TDLB_PIPE");
    if (honour_env && !(env && std::atoi(env) == 1)) return false;
    std::lock_guard<std::mutex> lk(mu);
    auto it = known.find(cu->device);
    if (it != known.end()) return it->second;
    const bool ok = pipe_xcc_check(cu->stream) == 1;
    known[cu->device] = ok;
    return ok;
}

// Tables of the one-launch step (PipeArgs): the first static stencil cell of
// every row in the plane-grouped cell list, and the forcing-level workgroups
// of each XCD's plane range, teams first.
int setup_pipe(tdlb_cu* cu) {
    cu->pipe = false;
    for (void* p : {(void*)cu->rowcell, (void*)cu->pctr})
        if (p) (void)hipFree(p);
    cu->rowcell = nullptr;
    cu->pctr = nullptr;
    const Grid& g = cu->g;
    const bool shape = !cu->simple() && cu->dlay() == LayoutLIJK && cu->prm.precision != TDLB_PREC_F64 &&
                       g.xg1 >= 4 && g.yg1 >= 3 * kXcds && g.zg1 >= 64;
    if (!shape || !pipe_device_ok(cu)) return TDLB_OK;
    const IndexedSet& set = cu->static_set;
    const int rw = g.xg1 + 2;
    std::vector<int32_t> rowcell((size_t)(g.yg1 + 2) * rw, 0);
    std::vector<uint32_t> cells((size_t)set.n_cells);
    if (!cells.empty())
        HIPCHK(hipMemcpyAsync(cells.data(), set.cells, cells.size() * 4, hipMemcpyDeviceToHost, cu->stream));
    HIPCHK(hipStreamSynchronize(cu->stream));
    // plane j's cells lie in [first of j, first of j+1); rows by cell index within
    std::vector<int64_t> pb((size_t)g.yg1 + 2, -1), pe((size_t)g.yg1 + 2, -1);
    for (const PlaneDesc& pd : set.plane_list) {
        pb[pd.j] = pd.cell_begin;
        pe[pd.j] = pd.cell_end;
    }
    for (int j = 0; j <= g.yg1 + 1; j++) {
        const int64_t b = pb[j] < 0 ? 0 : pb[j], e = pe[j] < 0 ? 0 : pe[j];
        for (int i = 1; i <= g.xg1 + 1; i++) {
            const uint32_t first = (uint32_t)i * g.yz + (uint32_t)j * (uint32_t)g.zg;
            const int64_t at = std::lower_bound(cells.begin() + b, cells.begin() + e, first) - cells.begin();
            rowcell[(size_t)j * rw + i] = (int32_t)(i == g.xg1 + 1 ? e : at);
        }
    }
    const size_t words = pipe_counter_words(g);
    if (hipMalloc((void**)&cu->rowcell, rowcell.size() * 4) != hipSuccess ||
        hipMalloc((void**)&cu->pctr, words * 4) != hipSuccess)
        return fail(TDLB_ERR_NOMEM, "one-launch step tables");
    HIPCHK(hipMemcpyAsync(cu->rowcell, rowcell.data(), rowcell.size() * 4, hipMemcpyHostToDevice, cu->stream));
    HIPCHK(launch_zero(cu->pctr, words * 4, cu->stream));
    HIPCHK(hipStreamSynchronize(cu->stream));
    cu->pipe = true;
    return TDLB_OK;
}

// One step of a bound Esotwist tank in one launch: blade points of the step
// generated before, static stencil cells prepared when the previous mask may
// lack them.  Same arithmetic and order as launch_esotwist + bound_forcing.
int pipe_step(tdlb_cu* cu) {
    uint64_t* prev = cu->fmask[cu->fm];
    uint64_t* next = cu->fmask[1 - cu->fm];
    if (!cu->static_in_prev)
        HIPCHK(launch_force_prepare(cu->eprec(), cu->pts.d, 0, cu->n_static, cu->F, prev, next, cu->g, cu->stream));
    PipeArgs a{};
    a.Q = cu->Q[0];
    a.F = cu->F;
    a.Nu = cu->Nu;
    a.prev = prev;
    a.next = next;
    a.static_base = cu->static_base;
    a.fg = cu->fgeom();
    a.rowcell = cu->rowcell;
    a.ctr = cu->pctr;
    HIPCHK(launch_zero(cu->pctr, pipe_counter_words(cu->g) * 4, cu->stream));
    const int hint = line_rows_bytes(cu->g, cu->ssize) ? 1 : 0;
    HIPCHK(launch_eso_pipe(cu->prec(), cu->even ? 1 : 0, hint, a, cu->g, cu->cargs(), cu->rho0, cu->alfa, cu->beta,
                           0, row_lanes(cu->g), cu->stream));
    if (cu->o_dirty) {
        HIPCHK(launch_zero_interior_o(cu->O, cu->g, cu->stream));
        cu->o_dirty = false;
    }
    cu->even = !cu->even;
    cu->fm = 1 - cu->fm;
    cu->static_in_prev = true;
    return TDLB_OK;
}

// A team barrier of the forcing planes that gave up waiting leaves its flag
// set (kBarSpinLimit): report it once and clear it.  Never expected (teams are
// resident together); results after it are not the reference's.
int check_team_barriers(tdlb_cu* cu) {
    if (cu->pctr) {
        uint32_t f[kXcds];
        for (int x = 0; x < kXcds; x++)
            HIPCHK(hipMemcpy(&f[x], cu->pctr + 64 * x + 4, 4, hipMemcpyDeviceToHost));
        for (int x = 0; x < kXcds; x++)
            if (f[x]) {
                HIPCHK(launch_zero(cu->pctr, pipe_counter_words(cu->g) * 4, cu->stream));
                HIPCHK(hipStreamSynchronize(cu->stream));
                return fail(TDLB_ERR_HIP, "one-launch step: a wait timed out (results are invalid)");
            }
    }
    const IndexedSet& set = cu->static_set;
    if (!set.bar || set.n_teams <= 0) return TDLB_OK;
    std::vector<uint32_t> w((size_t)set.n_teams * kBarWords);
    HIPCHK(hipMemcpy(w.data(), set.bar, w.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    uint32_t bad = 0;
    for (int32_t t = 0; t < set.n_teams; t++) bad |= w[(size_t)t * kBarWords + 2];
    if (!bad) return TDLB_OK;
    if (bad & 2u) {
        HIPCHK(launch_zero(set.bar, w.size() * sizeof(uint32_t), cu->stream));
        HIPCHK(hipStreamSynchronize(cu->stream));
        return fail(TDLB_ERR_HIP, "IB forcing: a plane team ran on more than one XCD (forcing results are invalid)");
    }
    // a workgroup that gave up left the arrival count behind: restart every barrier
    HIPCHK(launch_zero(set.bar, w.size() * sizeof(uint32_t), cu->stream));
    HIPCHK(hipStreamSynchronize(cu->stream));
    return fail(TDLB_ERR_HIP, "IB forcing: a plane team barrier timed out (forcing results are invalid)");
}

// One forcing pass of the bound tank: the next occupancy mask starts from the
// static bits (static_base), the static stencil cells are prepared only when
// the previous mask may lack them, and every plane task runs in one launch.
int bound_forcing(tdlb_cu* cu, int momsrc, bool prepare_static) {
    uint64_t* prev = cu->fmask[cu->fm];
    uint64_t* next = cu->fmask[1 - cu->fm];
    const bool stored = momsrc == MOMSRC_STORED;
    const void* Qsrc = stored ? cu->eq() : cu->Q[cu->cur];
    const int qp = stored ? cu->eprec() : cu->prec();
    HIPCHK(launch_copy(next, cu->static_base, cu->mask_bytes(), cu->stream));
    if (prepare_static)
        HIPCHK(launch_force_prepare(cu->eprec(), cu->pts.d, 0, cu->n_static, cu->F, prev, next, cu->g, cu->stream));
    HIPCHK(launch_force_planes(qp, cu->dlay(), cu->prm.streaming, momsrc, cu->even ? 1 : 0, cu->fgeom(), Qsrc, cu->F,
                               prev, next, cu->g, cu->alfa, cu->beta, cu->rho0, cu->stream));
    if (cu->o_dirty) {
        HIPCHK(launch_zero_interior_o(cu->O, cu->g, cu->stream));
        cu->o_dirty = false;
    }
    cu->fm = 1 - cu->fm;
    cu->static_in_prev = true;
    return TDLB_OK;
}

int gen_blades(tdlb_cu* cu, double angle, double omega) {
    BladeStep st{};
    Rushton R(cu->rp);
    R.blade_trig(angle, st.cosb, st.sinb);
    st.omega = omega;
    HIPCHK(launch_gen_blades(cu->prec(), cu->blade, st, cu->pts.d, cu->n_static, cu->stream));
    return TDLB_OK;
}

// reference precision of an angular increment (main.cpp uses useQVecPrecision)
double to_prec(const tdlb_cu* cu, double v) { return cu->tsize == 4 ? (double)(float)v : v; }

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char* tdlb_last_error(void) { return g_err.c_str(); }
int tdlb_version(void) { return 1; }

int tdlb_device_count(int* count) {
    if (!count) return fail(TDLB_ERR_INVALID, "null count");
    HIPCHK(hipGetDeviceCount(count));
    return TDLB_OK;
}

int tdlb_probe_planes(int32_t device_id, int32_t elem_bytes, int64_t cells, int64_t plane_stride_bytes,
                      int32_t planes, int32_t nontemporal, int32_t passes, double* ms_per_pass) {
    if (!ms_per_pass) return fail(TDLB_ERR_INVALID, "null argument");
    if (elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8) return fail(TDLB_ERR_INVALID, "elem_bytes: 2, 4 or 8");
    if (cells < 1 || planes < 1 || planes > 20 || passes < 1)
        return fail(TDLB_ERR_INVALID, "cells >= 1, planes 1..20, passes >= 1");
    if ((uint64_t)cells > (1ull << 40)) return fail(TDLB_ERR_INVALID, "cells too large");
    const uint64_t align = 1ull << 20;
    // default stride: the plane rounded up to 1 MiB
    const uint64_t plane_bytes = plane_stride_bytes > 0 ? (uint64_t)plane_stride_bytes
                                                        : ((uint64_t)cells * elem_bytes + align - 1) / align * align;
    if (plane_bytes < (uint64_t)cells * elem_bytes || plane_bytes % 128 != 0)
        return fail(TDLB_ERR_INVALID, "plane_stride_bytes: at least cells * elem_bytes, a multiple of 128");
    HIPCHK(hipSetDevice(device_id < 0 ? 0 : device_id));
    void* base = nullptr;
    if (hipMalloc(&base, plane_bytes * planes) != hipSuccess) return fail(TDLB_ERR_NOMEM, "probe planes");
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = launch_zero(base, plane_bytes * planes, s);
    const uint64_t stride = plane_bytes / elem_bytes;
    for (int w = 0; w < 2 && e == hipSuccess; w++)
        e = launch_probe_planes(elem_bytes, base, stride, (uint64_t)cells, planes, nontemporal, s);
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    for (int p = 0; p < passes && e == hipSuccess; p++)
        e = launch_probe_planes(elem_bytes, base, stride, (uint64_t)cells, planes, nontemporal, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (s) (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(base);
    if (e != hipSuccess) return fail(TDLB_ERR_HIP, std::string("probe: ") + hipGetErrorString(e));
    *ms_per_pass = (double)ms / passes;
    return TDLB_OK;
}

int tdlb_flow_calc_nu_re_m(tdlb_flow* f, int32_t outer_radius, int32_t precision) {
    if (!f) return fail(TDLB_ERR_INVALID, "null flow");
    if (precision == TDLB_PREC_F32 || precision == TDLB_PREC_F16 || precision == TDLB_PREC_BF16 ||
        precision == TDLB_PREC_E3M12) {
        const float reM = (float)f->re_m_nondimensional, uav = (float)f->uav;
        const float Re_m = reM * M_PI / 2.0;
        const float nu = uav * static_cast<float>(outer_radius) / Re_m;
        f->re_m = Re_m;
        f->nu = nu;
    } else if (precision == TDLB_PREC_F64) {
        const double Re_m = f->re_m_nondimensional * M_PI / 2.0;
        f->re_m = Re_m;
        f->nu = f->uav * static_cast<double>(outer_radius) / Re_m;
    } else {
        return fail(TDLB_ERR_INVALID, "precision");
    }
    return TDLB_OK;
}

int tdlb_cu_create(const tdlb_cu_params* p, const tdlb_flow* f, tdlb_cu_t* out) {
    if (!p || !f || !out) return fail(TDLB_ERR_INVALID, "null argument");
    if (p->ghost != 1) return fail(TDLB_ERR_UNSUPPORTED, "ghost must be 1 (GridParams::multiStep default)");
    if (p->x < 1 || p->y < 1 || p->z < 1) return fail(TDLB_ERR_INVALID, "grid must have at least one interior cell");
    if (p->layout != TDLB_LAYOUT_IJKL && p->layout != TDLB_LAYOUT_LIJK && p->layout != TDLB_LAYOUT_IJKL_SOA)
        return fail(TDLB_ERR_INVALID, "layout");
    if (p->streaming != TDLB_STREAM_SIMPLE && p->streaming != TDLB_STREAM_ESOTWIST)
        return fail(TDLB_ERR_INVALID, "streaming");
    if (p->precision < TDLB_PREC_F32 || p->precision > TDLB_PREC_E3M12)
        return fail(TDLB_ERR_UNSUPPORTED, "precision");
    if (f->collision != 0) return fail(TDLB_ERR_UNSUPPORTED, "only the EgglesSomers collision is implemented");
    const int64_t xg = p->x + 2, yg = p->y + 2, zg = p->z + 2;
    const uint64_t size = (uint64_t)xg * yg * zg;
    // cell indices and the Esotwist wrap offsets are 32-bit (slot addresses are
    // 64-bit): up to 2^31 cells incl. ghosts, a 1288^3 grid
    if (size >= (1ull << 31))
        return fail(TDLB_ERR_UNSUPPORTED, "grid too large: at most 2^31 cells including ghosts");

    tdlb_cu* cu = new tdlb_cu();
    cu->prm = *p;
    cu->flow = *f;
    cu->xg = xg; cu->yg = yg; cu->zg = zg;
    cu->g.xg = (int32_t)xg; cu->g.yg = (int32_t)yg; cu->g.zg = (int32_t)zg;
    cu->g.xg1 = (int32_t)(xg - 2); cu->g.yg1 = (int32_t)(yg - 2); cu->g.zg1 = (int32_t)(zg - 2);
    cu->g.yz = (uint32_t)(yg * zg);
    cu->g.mask_w = (uint32_t)((zg + 63) / 64);
    cu->g.size = size;
    cu->tsize = p->precision == TDLB_PREC_F64 ? 8 : 4;
    cu->ssize = p->precision == TDLB_PREC_F64 ? 8 : p->precision == TDLB_PREC_F32 ? 4 : 2;
    set_q_geometry(cu->g, cu->dlay(), cu->ssize);
    if (cu->q_cells() >= (1ull << 32)) {
        delete cu;
        return fail(TDLB_ERR_UNSUPPORTED, "grid too large for 32-bit cell indexing");
    }
    cu->rho0 = (float)f->initial_rho;
    cu->mask_words = (size_t)xg * yg * cu->g.mask_w;
    cu->device = p->device_id < 0 ? 0 : p->device_id;
    cu->alfa = f->alpha;
    cu->beta = f->beta;

    auto bad = [&](int code, const std::string& m) {
        release(cu);
        delete cu;
        return fail(code, m);
    };
    if (hipSetDevice(cu->device) != hipSuccess) return bad(TDLB_ERR_HIP, "hipSetDevice failed");
    if (resolve_scratch(cu->device) != TDLB_OK) {
        const std::string m = g_err;
        return bad(TDLB_ERR_HIP, m);
    }
    if (hipStreamCreateWithFlags(&cu->stream, hipStreamNonBlocking) != hipSuccess)
        return bad(TDLB_ERR_HIP, "hipStreamCreate failed");
    const int nq = cu->simple() ? 2 : 1;
    for (int b = 0; b < nq; b++)
        if (hipMalloc(&cu->Q[b], cu->qs_bytes()) != hipSuccess) return bad(TDLB_ERR_NOMEM, "hipMalloc Q");
    if (hipMalloc(&cu->F, cu->f_bytes()) != hipSuccess) return bad(TDLB_ERR_NOMEM, "hipMalloc F");
    if (hipMalloc(&cu->Nu, cu->nu_bytes()) != hipSuccess) return bad(TDLB_ERR_NOMEM, "hipMalloc Nu");
    if (hipMalloc((void**)&cu->O, size) != hipSuccess) return bad(TDLB_ERR_NOMEM, "hipMalloc O");
    uint64_t** masks[] = {&cu->fmask[0], &cu->fmask[1], &cu->gmask};
    for (uint64_t** m : masks)
        if (hipMalloc((void**)m, cu->mask_bytes()) != hipSuccess) return bad(TDLB_ERR_NOMEM, "hipMalloc mask");
    // a fresh ComputeUnit holds zero F/Nu/O (its arrays are unset in the reference; zero is a valid state)
    if (launch_zero(cu->F, cu->f_bytes(), cu->stream) != hipSuccess ||
        launch_zero(cu->Nu, cu->nu_bytes(), cu->stream) != hipSuccess ||
        launch_zero(cu->O, size, cu->stream) != hipSuccess ||
        launch_zero(cu->fmask[0], cu->mask_bytes(), cu->stream) != hipSuccess ||
        launch_zero(cu->fmask[1], cu->mask_bytes(), cu->stream) != hipSuccess ||
        launch_zero(cu->gmask, cu->mask_bytes(), cu->stream) != hipSuccess)
        return bad(TDLB_ERR_HIP, "initial fill failed");
    for (int b = 0; b < nq; b++)
        if (launch_zero(cu->Q[b], cu->qs_bytes(), cu->stream) != hipSuccess) return bad(TDLB_ERR_HIP, "Q fill failed");
    // the 16-bit lattice starts as the rest state at rho0 (all deviations zero)
    if (cu->shifted()) cu->phase = PH_LAZY;
    if (hipStreamSynchronize(cu->stream) != hipSuccess) return bad(TDLB_ERR_HIP, "synchronise failed");
    *out = cu;
    return TDLB_OK;
}

int tdlb_cu_destroy(tdlb_cu_t cu) {
    if (!cu) return TDLB_OK;
    (void)hipSetDevice(cu->device);
    if (cu->stream) (void)hipStreamSynchronize(cu->stream);
    release(cu);
    delete cu;
    return TDLB_OK;
}

int tdlb_cu_set_flow(tdlb_cu_t cu, const tdlb_flow* f) {
    if (!cu || !f) return fail(TDLB_ERR_INVALID, "null argument");
    if (f->collision != 0) return fail(TDLB_ERR_UNSUPPORTED, "only the EgglesSomers collision is implemented");
    CHK(set_device(cu));
    CHK(materialise(cu));
    cu->flow = *f;
    return TDLB_OK;
}

int tdlb_cu_initialise(tdlb_cu_t cu, double rho) {
    if (!cu) return fail(TDLB_ERR_INVALID, "null handle");
    CHK(set_device(cu));
    const double r = to_prec(cu, rho);
    if (cu->shifted()) {
        // rest state at rho0 = rho: every stored deviation is zero
        cu->rho0 = (float)r;
        const int nq = cu->simple() ? 2 : 1;
        for (int b = 0; b < nq; b++) HIPCHK(launch_zero(cu->Q[b], cu->qs_bytes(), cu->stream));
        if (cu->M) {
            HIPCHK(hipStreamSynchronize(cu->stream));
            (void)hipFree(cu->M);
            cu->M = nullptr;
        }
    } else {
        const int nq = cu->simple() ? 2 : 1;
        for (int b = 0; b < nq; b++)
            HIPCHK(launch_fill_q(cu->prec(), cu->dlay(), cu->Q[b], cu->g, r, cu->stream));
    }
    HIPCHK(launch_zero(cu->F, cu->f_bytes(), cu->stream));
    HIPCHK(launch_zero(cu->Nu, cu->nu_bytes(), cu->stream));
    HIPCHK(launch_zero(cu->O, cu->g.size, cu->stream));
    HIPCHK(launch_zero(cu->fmask[0], cu->mask_bytes(), cu->stream));
    HIPCHK(launch_zero(cu->fmask[1], cu->mask_bytes(), cu->stream));
    HIPCHK(launch_zero(cu->gmask, cu->mask_bytes(), cu->stream));
    // the 16-bit storage starts from its populations: the rest populations have
    // the moments initialise writes (rho0, 0, ..., 0)
    cu->phase = cu->shifted() ? PH_LAZY : PH_EXPLICIT;
    cu->o_dirty = false;
    cu->static_in_prev = false;
    if (cu->bound) HIPCHK(launch_copy(cu->static_base, cu->static_bits, cu->mask_bytes(), cu->stream));
    return TDLB_OK;
}

int tdlb_cu_collision(tdlb_cu_t cu) {
    if (!cu) return fail(TDLB_ERR_INVALID, "null handle");
    CHK(need_full_precision(cu, "collision"));
    CHK(set_device(cu));
    CHK(materialise(cu));
    const auto c = cu->cargs();
    void* q = cu->Q[cu->cur];
    if (cu->simple())
        HIPCHK(launch_simple(cu->prec(), cu->dlay(), MODE_COLL, q, q, q, cu->F, cu->Nu, cu->fmask[cu->fm],
                             cu->g, c, cu->rho0, cu->stream));
    else
        HIPCHK(launch_esotwist(cu->prec(), cu->dlay(), MODE_COLL, cu->even ? 1 : 0, q, q, cu->F, cu->Nu,
                               cu->fmask[cu->fm], cu->g, c, cu->rho0, cu->stream));
    return TDLB_OK;
}

int tdlb_cu_moments(tdlb_cu_t cu) {
    if (!cu) return fail(TDLB_ERR_INVALID, "null handle");
    CHK(need_full_precision(cu, "moments"));
    CHK(set_device(cu));
    CHK(materialise(cu));
    const auto c = cu->cargs();
    void* q = cu->Q[cu->cur];
    if (cu->simple())
        HIPCHK(launch_simple(cu->prec(), cu->dlay(), MODE_MOM, q, q, q, cu->F, cu->Nu, cu->fmask[cu->fm], cu->g,
                             c, cu->rho0, cu->stream));
    else
        HIPCHK(launch_esotwist(cu->prec(), cu->dlay(), MODE_MOM, cu->even ? 1 : 0, q, q, cu->F, cu->Nu,
                               cu->fmask[cu->fm], cu->g, c, cu->rho0, cu->stream));
    return TDLB_OK;
}

int tdlb_cu_stream_push(tdlb_cu_t cu) {
    if (!cu) return fail(TDLB_ERR_INVALID, "null handle");
    CHK(need_full_precision(cu, "streaming"));
    CHK(set_device(cu));
    CHK(materialise(cu));
    if (!cu->simple()) {
        cu->even = !cu->even;
        return TDLB_OK;
    }
    HIPCHK(launch_push_all((int)cu->tsize, cu->dlay(), cu->Q[cu->cur], cu->Q[1 - cu->cur], cu->g, cu->stream));
    cu->cur = 1 - cu->cur;
    return TDLB_OK;
}

int tdlb_cu_stream_pull(tdlb_cu_t cu) {
    if (!cu) return fail(TDLB_ERR_INVALID, "null handle");
    CHK(need_full_precision(cu, "streaming"));
    CHK(set_device(cu));
    CHK(materialise(cu));
    if (!cu->simple()) {
        cu->even = !cu->even;
        return TDLB_OK;
    }
    HIPCHK(launch_pull_all((int)cu->tsize, cu->dlay(), cu->Q[cu->cur], cu->Q[1 - cu->cur], cu->g, cu->stream));
    cu->cur = 1 - cu->cur;
    return TDLB_OK;
}

int tdlb_cu_bounce_back(tdlb_cu_t cu) {
    if (!cu) return fail(TDLB_ERR_INVALID, "null handle");
    CHK(need_full_precision(cu, "bounce-back"));
    CHK(set_device(cu));
    CHK(materialise(cu));
    if (!cu->simple()) return TDLB_OK;  // Boundary.hpp:13-14
    HIPCHK(launch_bounce_back((int)cu->tsize, cu->dlay(), cu->Q[cu->cur], cu->g, cu->stream));
    return TDLB_OK;
}

int tdlb_cu_forcing(tdlb_cu_t cu, const tdlb_ib_point* points, int64_t n, double alfa, double beta) {
    if (!cu || (n > 0 && !points) || n < 0) return fail(TDLB_ERR_INVALID, "bad arguments");
    CHK(need_full_precision(cu, "forcing"));
    CHK(set_device(cu));
    CHK(check_points(cu, points, n));
    CHK(materialise(cu));
    std::vector<int64_t> order, offsets;
    compute_levels(cu, points, n, order, offsets);
    HostPoints h;
    to_host_points(cu, points, order, h);
    CHK(cu->api_pts.reserve(n, cu->tsize));
    CHK(upload_points(cu, cu->api_pts, 0, h));
    CHK(build_indexed_set(cu, points, order, cu->api_set));
    std::vector<ForceSegment> segs(1);
    segs[0].set = &cu->api_set;
    segs[0].n_points = n;
    for (size_t l = 0; l + 1 < offsets.size(); l++) segs[0].levels.push_back({offsets[l], offsets[l + 1]});
    std::vector<std::pair<int64_t, int64_t>> prep;
    if (n > 0) prep.push_back({0, n});
    const double a0 = cu->alfa, b0 = cu->beta;
    cu->alfa = alfa;
    cu->beta = beta;
    int r = forcing_pass(cu, cu->api_pts, segs, prep, cu->gmask, MOMSRC_STORED);
    cu->alfa = a0;
    cu->beta = b0;
    cu->static_in_prev = false;
    return r;
}

int tdlb_cu_field_bytes(tdlb_cu_t cu, int field, size_t* bytes) {
    if (!cu || !bytes) return fail(TDLB_ERR_INVALID, "null argument");
    switch (field) {
        case TDLB_FIELD_Q: *bytes = cu->qb_bytes(); break;
        case TDLB_FIELD_F: *bytes = cu->f_bytes(); break;
        case TDLB_FIELD_NU: *bytes = cu->nub_bytes(); break;
        case TDLB_FIELD_O: *bytes = cu->g.size; break;
        default: return fail(TDLB_ERR_INVALID, "field");
    }
    return TDLB_OK;
}

int tdlb_cu_download(tdlb_cu_t cu, int field, void* host, size_t bytes) {
    size_t need = 0;
    CHK(tdlb_cu_field_bytes(cu, field, &need));
    if (!host || bytes != need) return fail(TDLB_ERR_INVALID, "download: buffer size must be " + std::to_string(need));
    CHK(set_device(cu));
    CHK(materialise(cu));
    if (field == TDLB_FIELD_F) {
        void* tmp = nullptr;
        if (hipMalloc(&tmp, need) != hipSuccess) return fail(TDLB_ERR_NOMEM, "download F staging");
        hipError_t e = launch_masked_f(cu->prec(), cu->F, cu->fmask[cu->fm], tmp, cu->g, cu->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(host, tmp, need, hipMemcpyDeviceToHost, cu->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(cu->stream);
        (void)hipFree(tmp);
        if (e != hipSuccess) return fail(TDLB_ERR_HIP, std::string("download F: ") + hipGetErrorString(e));
        return TDLB_OK;
    }
    if (field == TDLB_FIELD_Q && cu->q_converted()) {
        void* tmp = nullptr;
        if (hipMalloc(&tmp, need) != hipSuccess) return fail(TDLB_ERR_NOMEM, "download Q staging");
        hipError_t e = launch_q_boundary((int)cu->tsize, cu->dlay(), cu->blay(), 1, cu->eq(), tmp, cu->g, cu->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(host, tmp, need, hipMemcpyDeviceToHost, cu->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(cu->stream);
        (void)hipFree(tmp);
        if (e != hipSuccess) return fail(TDLB_ERR_HIP, std::string("download Q: ") + hipGetErrorString(e));
        return TDLB_OK;
    }
    if (field == TDLB_FIELD_NU && cu->shifted()) {
        void* tmp = nullptr;
        if (hipMalloc(&tmp, need) != hipSuccess) return fail(TDLB_ERR_NOMEM, "download Nu staging");
        hipError_t e = launch_nu_convert(cu->prec(), 1, cu->Nu, tmp, cu->g.size, cu->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(host, tmp, need, hipMemcpyDeviceToHost, cu->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(cu->stream);
        (void)hipFree(tmp);
        if (e != hipSuccess) return fail(TDLB_ERR_HIP, std::string("download Nu: ") + hipGetErrorString(e));
        return TDLB_OK;
    }
    const void* src = field == TDLB_FIELD_Q ? cu->eq() : field == TDLB_FIELD_NU ? cu->Nu : (const void*)cu->O;
    HIPCHK(hipMemcpyAsync(host, src, need, hipMemcpyDeviceToHost, cu->stream));
    HIPCHK(hipStreamSynchronize(cu->stream));
    return TDLB_OK;
}

int tdlb_cu_upload(tdlb_cu_t cu, int field, const void* host, size_t bytes) {
    size_t need = 0;
    CHK(tdlb_cu_field_bytes(cu, field, &need));
    if (!host || bytes != need) return fail(TDLB_ERR_INVALID, "upload: buffer size must be " + std::to_string(need));
    CHK(set_device(cu));
    CHK(materialise(cu));
    void* dst = field == TDLB_FIELD_Q ? cu->eq() : field == TDLB_FIELD_F ? cu->F
              : field == TDLB_FIELD_NU ? cu->Nu : (void*)cu->O;
    if (field == TDLB_FIELD_Q && cu->q_converted()) {
        void* tmp = nullptr;
        if (hipMalloc(&tmp, need) != hipSuccess) return fail(TDLB_ERR_NOMEM, "upload Q staging");
        hipError_t e = hipMemcpyAsync(tmp, host, need, hipMemcpyHostToDevice, cu->stream);
        if (e == hipSuccess) e = launch_q_boundary((int)cu->tsize, cu->dlay(), cu->blay(), 0, tmp, dst, cu->g, cu->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(cu->stream);
        (void)hipFree(tmp);
        if (e != hipSuccess) return fail(TDLB_ERR_HIP, std::string("upload Q: ") + hipGetErrorString(e));
        return TDLB_OK;
    }
    if (field == TDLB_FIELD_NU && cu->shifted()) {
        void* tmp = nullptr;
        if (hipMalloc(&tmp, need) != hipSuccess) return fail(TDLB_ERR_NOMEM, "upload Nu staging");
        hipError_t e = hipMemcpyAsync(tmp, host, need, hipMemcpyHostToDevice, cu->stream);
        if (e == hipSuccess) e = launch_nu_convert(cu->prec(), 0, tmp, cu->Nu, cu->g.size, cu->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(cu->stream);
        (void)hipFree(tmp);
        if (e != hipSuccess) return fail(TDLB_ERR_HIP, std::string("upload Nu: ") + hipGetErrorString(e));
        return TDLB_OK;
    }
    HIPCHK(hipMemcpyAsync(dst, host, need, hipMemcpyHostToDevice, cu->stream));
    if (field == TDLB_FIELD_F) {
        HIPCHK(launch_mask_from_f(cu->prec(), cu->F, cu->fmask[cu->fm], cu->gmask, cu->g, cu->stream));
        cu->static_in_prev = false;
        if (cu->bound) {
            // static_base = static stencil bits | ghost bits with non-zero F
            std::vector<uint64_t> a(cu->mask_words), b(cu->mask_words);
            HIPCHK(hipMemcpyAsync(a.data(), cu->static_bits, cu->mask_bytes(), hipMemcpyDeviceToHost, cu->stream));
            HIPCHK(hipMemcpyAsync(b.data(), cu->gmask, cu->mask_bytes(), hipMemcpyDeviceToHost, cu->stream));
            HIPCHK(hipStreamSynchronize(cu->stream));
            for (size_t w = 0; w < a.size(); w++) a[w] |= b[w];
            HIPCHK(hipMemcpyAsync(cu->static_base, a.data(), cu->mask_bytes(), hipMemcpyHostToDevice, cu->stream));
        }
    }
    if (field == TDLB_FIELD_O) cu->o_dirty = true;
    HIPCHK(hipStreamSynchronize(cu->stream));
    return TDLB_OK;
}

int tdlb_cu_get_even_step(tdlb_cu_t cu, int* even) {
    if (!cu || !even) return fail(TDLB_ERR_INVALID, "null argument");
    CHK(set_device(cu));
    CHK(materialise(cu));
    *even = cu->even ? 1 : 0;
    return TDLB_OK;
}

int tdlb_cu_set_even_step(tdlb_cu_t cu, int even) {
    if (!cu) return fail(TDLB_ERR_INVALID, "null handle");
    CHK(set_device(cu));
    CHK(materialise(cu));
    cu->even = even != 0;
    return TDLB_OK;
}

int tdlb_cu_dims(tdlb_cu_t cu, int64_t* xg, int64_t* yg, int64_t* zg, int64_t* size) {
    if (!cu) return fail(TDLB_ERR_INVALID, "null handle");
    if (xg) *xg = cu->xg;
    if (yg) *yg = cu->yg;
    if (zg) *zg = cu->zg;
    if (size) *size = (int64_t)cu->g.size;
    return TDLB_OK;
}

int tdlb_cu_synchronize(tdlb_cu_t cu) {
    if (!cu) return fail(TDLB_ERR_INVALID, "null handle");
    CHK(set_device(cu));
    HIPCHK(hipStreamSynchronize(cu->stream));
    return check_team_barriers(cu);
}

int tdlb_cu_stream(tdlb_cu_t cu, void** s) {
    if (!cu || !s) return fail(TDLB_ERR_INVALID, "null argument");
    *s = (void*)cu->stream;
    return TDLB_OK;
}

// ---------------------------------------------------------------- checkpoint
static std::string ckpt_id(const tdlb_cu* cu) {
    return std::to_string(cu->prm.idi) + "." + std::to_string(cu->prm.idj) + "." + std::to_string(cu->prm.idk);
}

int tdlb_cu_checkpoint_write(tdlb_cu_t cu, const char* dir, const char* unit_name, const char* all_params_json) {
    if (!cu || !dir || !unit_name) return fail(TDLB_ERR_INVALID, "null argument");
    const std::string id = ckpt_id(cu);
    if (all_params_json) {
        const std::string path = std::string(dir) + "/AllParams." + id + "." + unit_name + ".json";
        FILE* fp = std::fopen(path.c_str(), "w");
        if (!fp) return fail(TDLB_ERR_IO, "cannot write " + path);
        const size_t n = std::strlen(all_params_json);
        const bool ok = std::fwrite(all_params_json, 1, n, fp) == n;
        std::fclose(fp);
        if (!ok) return fail(TDLB_ERR_IO, "short write " + path);
    }
    const int fields[3] = {TDLB_FIELD_Q, TDLB_FIELD_F, TDLB_FIELD_NU};
    const char* names[3] = {"Q", "F", "Nu"};
    for (int a = 0; a < 3; a++) {
        size_t bytes = 0;
        CHK(tdlb_cu_field_bytes(cu, fields[a], &bytes));
        std::vector<char> buf(bytes);
        CHK(tdlb_cu_download(cu, fields[a], buf.data(), bytes));
        const std::string path = std::string(dir) + "/checkpoint_grid." + id + "." + unit_name + "." + names[a];
        FILE* fp = std::fopen(path.c_str(), "wb");
        if (!fp) return fail(TDLB_ERR_IO, "cannot write " + path);
        const bool ok = std::fwrite(buf.data(), 1, bytes, fp) == bytes;
        std::fclose(fp);
        if (!ok) return fail(TDLB_ERR_IO, "short write " + path);
    }
    return TDLB_OK;
}

int tdlb_cu_checkpoint_read(tdlb_cu_t cu, const char* dir, const char* unit_name, int* missing) {
    if (!cu || !dir || !unit_name) return fail(TDLB_ERR_INVALID, "null argument");
    const std::string id = ckpt_id(cu);
    const int fields[3] = {TDLB_FIELD_Q, TDLB_FIELD_F, TDLB_FIELD_NU};
    const char* names[3] = {"Q", "F", "Nu"};
    int miss = 0;
    for (int a = 0; a < 3; a++) {
        size_t bytes = 0;
        CHK(tdlb_cu_field_bytes(cu, fields[a], &bytes));
        const std::string path = std::string(dir) + "/checkpoint_grid." + id + "." + unit_name + "." + names[a];
        FILE* fp = std::fopen(path.c_str(), "rb");
        if (!fp) {
            miss |= 1 << a;
            continue;
        }
        std::vector<char> buf(bytes);
        const size_t got = std::fread(buf.data(), 1, bytes, fp);
        const bool extra = std::fgetc(fp) != EOF;
        std::fclose(fp);
        if (got != bytes || extra)
            return fail(TDLB_ERR_IO, path + ": expected " + std::to_string(bytes) + " bytes");
        CHK(tdlb_cu_upload(cu, fields[a], buf.data(), bytes));
    }
    if (missing) *missing = miss;
    return TDLB_OK;
}

// ---------------------------------------------------------------- geometry
int tdlb_rushton_defaults(int64_t tank_diameter, int64_t height, double tip_speed, int64_t startup_steps,
                          tdlb_rushton_params* out) {
    if (!out || tank_diameter < 6 || height < 1) return fail(TDLB_ERR_INVALID, "tank too small");
    rushton_defaults(tank_diameter, height, out);
    out->tip_speed = tip_speed;
    out->startup_steps = startup_steps;
    return TDLB_OK;
}

int tdlb_rushton_generate(const tdlb_rushton_params* p, int group, double angle, double delta_angle,
                          tdlb_ib_point* out, int64_t cap, int64_t* n) {
    if (!p || !n) return fail(TDLB_ERR_INVALID, "null argument");
    Rushton R(*p);
    std::vector<tdlb_ib_point> v;
    if (group == TDLB_GEOM_FIXED) R.fixed(v);
    else if (group == TDLB_GEOM_ROTATING_NON_UPDATING) R.rotating_non_updating(delta_angle, v);
    else if (group == TDLB_GEOM_ROTATING) R.rotating(angle, delta_angle, v);
    else return fail(TDLB_ERR_INVALID, "group");
    *n = (int64_t)v.size();
    if (out && cap > 0) std::memcpy(out, v.data(), sizeof(tdlb_ib_point) * std::min<int64_t>(cap, *n));
    return TDLB_OK;
}

int tdlb_rushton_step_increment(const tdlb_rushton_params* p, uint64_t step, double* delta) {
    if (!p || !delta) return fail(TDLB_ERR_INVALID, "null argument");
    *delta = Rushton(*p).step_increment(step);
    return TDLB_OK;
}

int tdlb_rushton_solid_mask(const tdlb_rushton_params* p, int64_t xg, int64_t yg, int64_t zg, int64_t ghost,
                            uint8_t* out) {
    if (!p || !out) return fail(TDLB_ERR_INVALID, "null argument");
    Rushton(*p).solid_mask(xg, yg, zg, ghost, out);
    return TDLB_OK;
}

// ---------------------------------------------------------------- fused main loop
int tdlb_cu_bind_rushton(tdlb_cu_t cu, const tdlb_rushton_params* p, uint64_t step, double angle, double alfa,
                         double beta) {
    if (!cu || !p) return fail(TDLB_ERR_INVALID, "null argument");
    CHK(set_device(cu));
    // the 16-bit storage reads density and momentum from its populations instead
    if (!cu->shifted()) CHK(materialise(cu));
    Rushton R(*p);
    const BladeEnum be = R.blade_enum();
    if (p->n_blades < 1 || p->n_blades > 16) return fail(TDLB_ERR_INVALID, "n_blades must be 1..16");
    // blade colour classes never overlap (blade_colours: a colour per blade in small tanks)
    // blades are generated on the device at any angle: their whole swept disk and
    // the planes they occupy must lie inside the interior
    {
        const double ro = p->blade_outer_radius;
        const double lo_i = std::floor(p->center_x - ro + 0.5), hi_i = std::floor(p->center_x + ro + 0.5);
        const double lo_k = std::floor(p->center_z - ro + 0.5), hi_k = std::floor(p->center_z + ro + 0.5);
        if (lo_i < 0 || hi_i > (double)(cu->prm.x - 1) || lo_k < 0 || hi_k > (double)(cu->prm.z - 1) ||
            p->blade_plane_lo < 0 || p->blade_plane_lo + p->blade_planes > cu->prm.y || p->blade_planes < 1 ||
            p->blade_inner_radius < 0)
            return fail(TDLB_ERR_INVALID, "impeller does not fit inside the ComputeUnit interior");
    }

    // main.cpp:277-303: increment for the current step, fixed, non-updating, rotating
    const double delta = to_prec(cu, R.step_increment(step));
    std::vector<tdlb_ib_point> stat, rot;
    R.fixed(stat);
    R.rotating_non_updating(delta, stat);
    R.rotating(angle, delta, rot);
    CHK(check_points(cu, stat.data(), (int64_t)stat.size()));
    CHK(check_points(cu, rot.data(), (int64_t)rot.size()));

    cu->n_static = (int64_t)stat.size();
    cu->n_rot = (int64_t)rot.size();
    cu->n_blade_classes = cu->n_rot > 0 ? 5 * be.n_bcol : 0;
    std::vector<int64_t> order;
    CHK(build_plane_set(cu, stat, be, p->blade_plane_lo + cu->prm.ghost, order, cu->static_set));
    HostPoints hs;
    to_host_points(cu, stat.data(), order, hs);
    CHK(cu->pts.reserve(cu->n_static + cu->n_rot, cu->tsize));
    CHK(upload_points(cu, cu->pts, 0, hs));
    if (cu->rot_cache) (void)hipFree(cu->rot_cache);
    cu->rot_cache = nullptr;
    if (cu->n_rot > 0 && hipMalloc(&cu->rot_cache, (size_t)cu->n_rot * 9 * 4 * cu->tsize) != hipSuccess)
        return fail(TDLB_ERR_NOMEM, "blade moment cache");

    cu->blade.n_blades = p->n_blades;
    cu->blade.n_bcol = be.n_bcol;
    cu->blade.j_lo = (int32_t)p->blade_plane_lo;
    cu->blade.n_planes = be.n_planes;
    cu->blade.n_radial = be.n_radial;
    cu->blade.r_inner = p->blade_inner_radius;
    cu->blade.cx = p->center_x;
    cu->blade.cz = p->center_z;
    cu->blade.ghost = (int32_t)cu->prm.ghost;
    cu->rp = *p;

    // static occupancy bits: run the prepare kernel on an empty mask
    if (!cu->static_bits && hipMalloc((void**)&cu->static_bits, cu->mask_bytes()) != hipSuccess)
        return fail(TDLB_ERR_NOMEM, "static mask");
    if (!cu->static_base && hipMalloc((void**)&cu->static_base, cu->mask_bytes()) != hipSuccess)
        return fail(TDLB_ERR_NOMEM, "static mask");
    HIPCHK(launch_zero(cu->static_bits, cu->mask_bytes(), cu->stream));
    // prepare() zeroes F where the previous mask is clear: give it an all-ones mask
    void* ones = nullptr;
    if (hipMalloc(&ones, cu->mask_bytes()) != hipSuccess) return fail(TDLB_ERR_NOMEM, "mask scratch");
    std::vector<uint64_t> allset(cu->mask_words, ~0ull);
    hipError_t e = hipMemcpyAsync(ones, allset.data(), cu->mask_bytes(), hipMemcpyHostToDevice, cu->stream);
    if (e == hipSuccess)
        e = launch_force_prepare(cu->prec(), cu->pts.d, 0, cu->n_static, cu->F, (const uint64_t*)ones,
                                 cu->static_bits, cu->g, cu->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(cu->stream);
    (void)hipFree(ones);
    if (e != hipSuccess) return fail(TDLB_ERR_HIP, std::string("static mask: ") + hipGetErrorString(e));
    {
        std::vector<uint64_t> a(cu->mask_words), b(cu->mask_words);
        HIPCHK(hipMemcpyAsync(a.data(), cu->static_bits, cu->mask_bytes(), hipMemcpyDeviceToHost, cu->stream));
        HIPCHK(hipMemcpyAsync(b.data(), cu->gmask, cu->mask_bytes(), hipMemcpyDeviceToHost, cu->stream));
        HIPCHK(hipStreamSynchronize(cu->stream));
        for (size_t w = 0; w < a.size(); w++) a[w] |= b[w];
        HIPCHK(hipMemcpyAsync(cu->static_base, a.data(), cu->mask_bytes(), hipMemcpyHostToDevice, cu->stream));
        HIPCHK(hipStreamSynchronize(cu->stream));
    }
    cu->bound = true;
    cu->alfa = alfa;
    cu->beta = beta;
    cu->xcc_rr = pipe_device_ok(cu, false);
    CHK(setup_pipe(cu));

    // initial forcing of main.cpp:305 on the current density and momentum
    CHK(gen_blades(cu, angle, delta));
    CHK(bound_forcing(cu, cell_source(cu).momsrc, true));
    return TDLB_OK;
}

int tdlb_cu_run(tdlb_cu_t cu, int64_t nsteps, uint64_t* step, double* angle) {
    if (!cu || !step || !angle || nsteps < 0) return fail(TDLB_ERR_INVALID, "bad arguments");
    if (!cu->bound) return fail(TDLB_ERR_STATE, "no geometry bound (tdlb_cu_bind_rushton)");
    CHK(set_device(cu));
    Rushton R(cu->rp);
    const auto c = cu->cargs();
    for (int64_t s = 0; s < nsteps; s++) {
        *step += 1;                                             // running.incrementStep()
        const double delta = to_prec(cu, R.step_increment(*step));
        *angle += delta;                                        // running.angle += delta
        CHK(gen_blades(cu, *angle, delta));                     // updateRotatingGeometry
        if (cu->pipe && cu->phase == PH_LAZY) {
            const int c0 = prof_mark(cu);
            CHK(pipe_step(cu));
            const int c1 = prof_mark(cu);
            if (c0 >= 0 && c1 >= 0) {
                cu->coll.push_back({c0, c1});
                cu->force.push_back({c1, c1});
            }
            continue;
        }
        const int c0 = prof_mark(cu);
        if (cu->phase == PH_EXPLICIT) {
            void* q = cu->Q[cu->cur];
            void* e = cu->eq();
            if (cu->simple()) {
                // settle the ghost layer once (bounce-back rims); the 16-bit lattices never read theirs
                HIPCHK(launch_bounce_back((int)cu->tsize, cu->dlay(), e, cu->g, cu->stream));
                if (!cu->shifted())
                    HIPCHK(launch_ghost_copy((int)cu->ssize, cu->dlay(), q, cu->Q[1 - cu->cur], cu->g,
                                             cu->stream));
                HIPCHK(launch_simple(cu->prec(), cu->dlay(), MODE_COLL, q, q, e, cu->F, cu->Nu,
                                     cu->fmask[cu->fm], cu->g, c, cu->rho0, cu->stream));
            } else {
                HIPCHK(launch_esotwist(cu->prec(), cu->dlay(), MODE_COLL, cu->even ? 1 : 0, q, e, cu->F, cu->Nu,
                                       cu->fmask[cu->fm], cu->g, c, cu->rho0, cu->stream));
                cu->even = !cu->even;
            }
            cu->phase = PH_LAZY;
        } else {
            if (cu->simple()) {
                HIPCHK(launch_simple(cu->prec(), cu->dlay(), MODE_FUSED, cu->Q[cu->cur], cu->Q[1 - cu->cur],
                                     nullptr, cu->F, cu->Nu, cu->fmask[cu->fm], cu->g, c, cu->rho0, cu->stream));
                cu->cur = 1 - cu->cur;
            } else {
                HIPCHK(launch_esotwist(cu->prec(), cu->dlay(), MODE_FUSED, cu->even ? 1 : 0, cu->Q[0], nullptr,
                                       cu->F, cu->Nu, cu->fmask[cu->fm], cu->g, c, cu->rho0, cu->stream));
                cu->even = !cu->even;
            }
        }
        const int c1 = prof_mark(cu);
        CHK(bound_forcing(cu, cu->simple() ? MOMSRC_PULL_POPS : MOMSRC_ESO_POPS, !cu->static_in_prev));
        const int c2 = prof_mark(cu);
        if (c0 >= 0 && c1 >= 0) cu->coll.push_back({c0, c1});
        if (c1 >= 0 && c2 >= 0) cu->force.push_back({c1, c2});
    }
    return TDLB_OK;
}

int tdlb_cu_profile(tdlb_cu_t cu, int enable) {
    if (!cu) return fail(TDLB_ERR_INVALID, "null handle");
    CHK(set_device(cu));
    HIPCHK(hipStreamSynchronize(cu->stream));
    cu->prof = enable != 0;
    cu->ev_used = 0;
    cu->coll.clear();
    cu->force.clear();
    return TDLB_OK;
}

int tdlb_cu_profile_read(tdlb_cu_t cu, double* collide_ms, int64_t* collide_launches, double* forcing_ms,
                         int64_t* forcing_passes) {
    if (!cu) return fail(TDLB_ERR_INVALID, "null handle");
    CHK(set_device(cu));
    HIPCHK(hipStreamSynchronize(cu->stream));
    auto sum = [&](const std::vector<std::pair<int, int>>& v, double* ms) -> int {
        double t = 0.0;
        for (auto& pr : v) {
            float e = 0.f;
            HIPCHK(hipEventElapsedTime(&e, cu->events[pr.first], cu->events[pr.second]));
            t += e;
        }
        if (ms) *ms = t;
        return TDLB_OK;
    };
    CHK(sum(cu->coll, collide_ms));
    CHK(sum(cu->force, forcing_ms));
    if (collide_launches) *collide_launches = (int64_t)cu->coll.size();
    if (forcing_passes) *forcing_passes = (int64_t)cu->force.size();
    return TDLB_OK;
}

int tdlb_cu_rotating_points(tdlb_cu_t cu, tdlb_ib_point* out, int64_t cap, int64_t* n) {
    if (!cu || !n) return fail(TDLB_ERR_INVALID, "null argument");
    if (!cu->bound) return fail(TDLB_ERR_STATE, "no geometry bound");
    CHK(set_device(cu));
    *n = cu->n_rot;
    const int64_t m = std::min<int64_t>(cap, cu->n_rot);
    if (!out || m <= 0) return TDLB_OK;
    const size_t ts = cu->tsize;
    HostPoints h;
    h.resize(cu->n_rot, ts);
    const int64_t o = cu->n_static;
    HIPCHK(hipMemcpyAsync(h.i.data(), cu->pts.d.i + o, cu->n_rot * 4, hipMemcpyDeviceToHost, cu->stream));
    HIPCHK(hipMemcpyAsync(h.j.data(), cu->pts.d.j + o, cu->n_rot * 4, hipMemcpyDeviceToHost, cu->stream));
    HIPCHK(hipMemcpyAsync(h.k.data(), cu->pts.d.k + o, cu->n_rot * 4, hipMemcpyDeviceToHost, cu->stream));
    void* src[] = {cu->pts.d.ifrac, cu->pts.d.kfrac, cu->pts.d.u, cu->pts.d.v, cu->pts.d.w};
    std::vector<uint8_t>* dst[] = {&h.ifrac, &h.kfrac, &h.u, &h.v, &h.w};
    for (int a = 0; a < 5; a++)
        HIPCHK(hipMemcpyAsync(dst[a]->data(), (char*)src[a] + o * ts, cu->n_rot * ts, hipMemcpyDeviceToHost, cu->stream));
    HIPCHK(hipStreamSynchronize(cu->stream));
    auto get = [&](const std::vector<uint8_t>& v, int64_t q) -> double {
        if (ts == 4) { float f; std::memcpy(&f, v.data() + q * 4, 4); return f; }
        double d; std::memcpy(&d, v.data() + q * 8, 8); return d;
    };
    const int64_t gh = cu->prm.ghost;
    for (int64_t q = 0; q < m; q++) {
        out[q].i = h.i[q] - gh;
        out[q].j = h.j[q] - gh;
        out[q].k = h.k[q] - gh;
        out[q].i_frac = get(h.ifrac, q);
        out[q].k_frac = get(h.kfrac, q);
        out[q].u_delta = get(h.u, q);
        out[q].v_delta = get(h.v, q);
        out[q].w_delta = get(h.w, q);
    }
    return TDLB_OK;
}

int tdlb_cu_solid_mask(tdlb_cu_t cu, uint8_t* out, size_t bytes) {
    if (!cu || !out) return fail(TDLB_ERR_INVALID, "null argument");
    if (!cu->bound) return fail(TDLB_ERR_STATE, "no geometry bound");
    if (bytes != cu->g.size) return fail(TDLB_ERR_INVALID, "mask buffer must be xg*yg*zg bytes");
    CHK(set_device(cu));
    uint8_t* tmp = nullptr;
    if (hipMalloc((void**)&tmp, bytes) != hipSuccess) return fail(TDLB_ERR_NOMEM, "mask staging");
    hipError_t e = launch_solid_mask(cu->rp.center_x, cu->rp.center_z, cu->rp.tank_radius, tmp, cu->g,
                                     (int)cu->prm.ghost, cu->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, bytes, hipMemcpyDeviceToHost, cu->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(cu->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(TDLB_ERR_HIP, std::string("solid mask: ") + hipGetErrorString(e));
    return TDLB_OK;
}

int tdlb_cu_reductions(tdlb_cu_t cu, double* out) {
    if (!cu || !out) return fail(TDLB_ERR_INVALID, "null argument");
    CHK(set_device(cu));
    if (!cu->red && hipMalloc((void**)&cu->red, sizeof(double) * (kReduceBlocks + 1) * kReduceVals) != hipSuccess)
        return fail(TDLB_ERR_NOMEM, "reduction buffer");
    const ReduceArgs ra = reduce_args(cu);
    const CellSource cs = cell_source(cu);
    double* res = cu->red + (size_t)kReduceBlocks * kReduceVals;
    HIPCHK(launch_reduce(cs.prec, cu->dlay(), cu->prm.streaming, cs.momsrc, cs.even, cs.Q, cu->F,
                         cu->fmask[cu->fm], cu->g, ra, cu->rho0, cu->red, res, cu->stream));
    HIPCHK(hipMemcpyAsync(out, res, sizeof(double) * kReduceVals, hipMemcpyDeviceToHost, cu->stream));
    HIPCHK(hipStreamSynchronize(cu->stream));
    return TDLB_OK;
}

int tdlb_cu_reductions_noalloc(tdlb_cu_t cu, double* out_device) {
    if (!cu || !out_device) return fail(TDLB_ERR_INVALID, "null argument");
    static_assert(kScratchDoubles >= (size_t)(kReduceBlocks + 1) * kReduceVals,
                  "the reduction scratch holds every partial row of the fixed grid and the result");
    if (cu->device < 0 || cu->device >= kScratchDevices || !g_scratch[cu->device])
        return fail(TDLB_ERR_STATE, "reduction scratch not resolved on this device");
    CHK(set_device(cu));
    const ReduceArgs ra = reduce_args(cu);
    // every partial row and the result are written before they are read, so
    // the shared scratch needs no clearing between calls
    double* partial = g_scratch[cu->device];
    const CellSource cs = cell_source(cu);
    HIPCHK(launch_reduce(cs.prec, cu->dlay(), cu->prm.streaming, cs.momsrc, cs.even, cs.Q, cu->F,
                         cu->fmask[cu->fm], cu->g, ra, cu->rho0, partial, out_device, cu->stream));
    return TDLB_OK;
}

int tdlb_cu_macroscopic(tdlb_cu_t cu, void* out, size_t bytes) {
    if (!cu || !out) return fail(TDLB_ERR_INVALID, "null argument");
    const size_t need = (size_t)cu->g.xg1 * cu->g.yg1 * cu->g.zg1 * 4 * cu->tsize;
    if (bytes != need) return fail(TDLB_ERR_INVALID, "macroscopic: buffer size must be " + std::to_string(need));
    CHK(set_device(cu));
    void* tmp = nullptr;
    if (hipMalloc(&tmp, need) != hipSuccess) return fail(TDLB_ERR_NOMEM, "macroscopic staging");
    const CellSource cs = cell_source(cu);
    hipError_t e = launch_macroscopic(cs.prec, cu->dlay(), cu->prm.streaming, cs.momsrc, cs.even, cs.Q, cu->F,
                                      cu->fmask[cu->fm], cu->g, cu->rho0, tmp, cu->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, need, hipMemcpyDeviceToHost, cu->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(cu->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(TDLB_ERR_HIP, std::string("macroscopic: ") + hipGetErrorString(e));
    return TDLB_OK;
}

int tdlb_cu_double_resolution(tdlb_cu_t cu) {
    if (!cu) return fail(TDLB_ERR_INVALID, "null handle");
    CHK(set_device(cu));
    CHK(materialise(cu));
    tdlb_cu_params np = cu->prm;
    np.x *= 2;
    np.y *= 2;
    np.z *= 2;
    np.resolution += 2;  // resolution += factor (ComputeUnit.hpp:262)
    tdlb_cu* n = nullptr;
    CHK(tdlb_cu_create(&np, &cu->flow, &n));
    auto drop = [&](int rc) {
        tdlb_cu_destroy(n);
        return rc;
    };
    n->rho0 = cu->rho0;
    n->even = cu->even;
    n->alfa = cu->alfa;
    n->beta = cu->beta;
    if (n->shifted()) {
        const int rc = ensure_m(n);
        if (rc != TDLB_OK) return drop(rc);
    }
    hipError_t e = hipStreamSynchronize(cu->stream);
    if (e == hipSuccess)
        e = launch_regrid((int)cu->tsize, cu->dlay(), cu->eq(), n->eq(), (int)(cu->nu_bytes() / cu->g.size), cu->Nu,
                          n->Nu, cu->g, n->g, n->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
    if (e != hipSuccess) return drop(fail(TDLB_ERR_HIP, std::string("regrid: ") + hipGetErrorString(e)));
    n->phase = PH_EXPLICIT;
    // the handle keeps its address; the old arrays leave with n.  F, O and the
    // masks start at zero and no geometry is bound (the caller binds the tank
    // of the new size, as main.cpp:360-399 regenerates it)
    std::swap(*cu, *n);
    tdlb_cu_destroy(n);
    return TDLB_OK;
}

int tdlb_cu_output_exclude(tdlb_cu_t cu, uint8_t* out, size_t bytes) {
    if (!cu || !out) return fail(TDLB_ERR_INVALID, "null argument");
    if (!cu->bound) return fail(TDLB_ERR_STATE, "no geometry bound");
    if (bytes != cu->g.size) return fail(TDLB_ERR_INVALID, "exclude mask must be xg*yg*zg bytes");
    CHK(tdlb_cu_solid_mask(cu, out, bytes));
    const int64_t n = cu->n_static + cu->n_rot;
    std::vector<int32_t> i(n), j(n), k(n);
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(i.data(), cu->pts.d.i, n * 4, hipMemcpyDeviceToHost, cu->stream));
        HIPCHK(hipMemcpyAsync(j.data(), cu->pts.d.j, n * 4, hipMemcpyDeviceToHost, cu->stream));
        HIPCHK(hipMemcpyAsync(k.data(), cu->pts.d.k, n * 4, hipMemcpyDeviceToHost, cu->stream));
        HIPCHK(hipStreamSynchronize(cu->stream));
    }
    for (int64_t q = 0; q < n; q++) out[((uint64_t)i[q] * cu->yg + (uint64_t)j[q]) * cu->zg + (uint64_t)k[q]] = 1;
    return TDLB_OK;
}

// upload an optional exclude mask (size bytes) into a device buffer
static int upload_exclude(tdlb_cu* cu, const uint8_t* exclude, uint8_t** dev) {
    *dev = nullptr;
    if (!exclude) return TDLB_OK;
    if (hipMalloc((void**)dev, cu->g.size) != hipSuccess) return fail(TDLB_ERR_NOMEM, "exclude mask");
    const hipError_t e = hipMemcpyAsync(*dev, exclude, cu->g.size, hipMemcpyHostToDevice, cu->stream);
    if (e != hipSuccess) {
        (void)hipFree(*dev);
        *dev = nullptr;
        return fail(TDLB_ERR_HIP, std::string("exclude mask: ") + hipGetErrorString(e));
    }
    return TDLB_OK;
}

int tdlb_cu_plane_xz(tdlb_cu_t cu, int64_t cut, const uint8_t* exclude, void* qvec, int64_t* n_qvec, void* f3,
                     int64_t* n_f3) {
    if (!cu || !qvec || !f3 || !n_qvec || !n_f3) return fail(TDLB_ERR_INVALID, "null argument");
    if (cut < 1 || cut > cu->g.yg1) return fail(TDLB_ERR_INVALID, "cut must be an interior j");
    CHK(set_device(cu));
    CHK(materialise(cu));
    const int64_t cells = (int64_t)cu->g.xg1 * cu->g.zg1;
    uint8_t* ex = nullptr;
    CHK(upload_exclude(cu, exclude, &ex));
    PlaneRec4* dq = nullptr;
    PlaneRec3* df = nullptr;
    uint8_t* dflag = nullptr;
    std::vector<PlaneRec4> hq(cells);
    std::vector<PlaneRec3> hf(cells);
    std::vector<uint8_t> hflag(cells);
    hipError_t e = hipMalloc((void**)&dq, cells * sizeof(PlaneRec4));
    if (e == hipSuccess) e = hipMalloc((void**)&df, cells * sizeof(PlaneRec3));
    if (e == hipSuccess) e = hipMalloc((void**)&dflag, cells);
    if (e == hipSuccess)
        e = launch_plane_xz(cu->eprec(), cu->dlay(), cu->eq(), cu->F, cu->fmask[cu->fm], ex, cu->g, (int)cut,
                            cu->prm.i0, cu->prm.k0, dq, df, dflag, cu->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hq.data(), dq, cells * sizeof(PlaneRec4), hipMemcpyDeviceToHost, cu->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hf.data(), df, cells * sizeof(PlaneRec3), hipMemcpyDeviceToHost, cu->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hflag.data(), dflag, cells, hipMemcpyDeviceToHost, cu->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(cu->stream);
    for (void* p : {(void*)dq, (void*)df, (void*)dflag, (void*)ex})
        if (p) (void)hipFree(p);
    if (e != hipSuccess) return fail(TDLB_ERR_HIP, std::string("plane: ") + hipGetErrorString(e));
    int64_t a = 0, b = 0;
    auto* oq = static_cast<PlaneRec4*>(qvec);
    auto* of = static_cast<PlaneRec3*>(f3);
    for (int64_t t = 0; t < cells; t++) {
        if (hflag[t] & 1) oq[a++] = hq[t];
        if (hflag[t] & 2) of[b++] = hf[t];
    }
    *n_qvec = a;
    *n_f3 = b;
    return TDLB_OK;
}

int tdlb_cu_vorticity(tdlb_cu_t cu, int32_t axis, int64_t cut, const uint8_t* exclude, void* out, size_t bytes) {
    if (!cu || !out) return fail(TDLB_ERR_INVALID, "null argument");
    if (axis != 0 && axis != 1) return fail(TDLB_ERR_INVALID, "axis: 0 (plane j = cut) or 1 (plane k = cut)");
    const int64_t lim = axis == 0 ? cu->g.yg1 : cu->g.zg1;
    if (cut < 1 || cut > lim) return fail(TDLB_ERR_INVALID, "cut must be an interior index");
    const size_t need = (size_t)cu->g.xg1 * (axis == 0 ? cu->g.zg1 : cu->g.yg1) * cu->tsize;
    if (bytes != need) return fail(TDLB_ERR_INVALID, "vorticity: buffer size must be " + std::to_string(need));
    CHK(set_device(cu));
    CHK(materialise(cu));  // ghost neighbours are read as the reference holds them
    uint8_t* ex = nullptr;
    CHK(upload_exclude(cu, exclude, &ex));
    void* tmp = nullptr;
    hipError_t e = hipMalloc(&tmp, need);
    if (e == hipSuccess)
        e = launch_vorticity(cu->eprec(), cu->dlay(), cu->prm.streaming, MOMSRC_STORED, cu->even ? 1 : 0, cu->eq(),
                             cu->g, axis, (int)cut, ex, tmp, cu->rho0, cu->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, need, hipMemcpyDeviceToHost, cu->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(cu->stream);
    if (tmp) (void)hipFree(tmp);
    if (ex) (void)hipFree(ex);
    if (e != hipSuccess) return fail(TDLB_ERR_HIP, std::string("vorticity: ") + hipGetErrorString(e));
    return TDLB_OK;
}

int tdlb_cu_one_launch(tdlb_cu_t cu, int32_t* on) {
    if (!cu || !on) return fail(TDLB_ERR_INVALID, "null argument");
    *on = cu->bound && cu->pipe ? 1 : 0;
    return TDLB_OK;
}

int tdlb_cu_one_launch_phases(tdlb_cu_t cu, double* us, int32_t n) {
    if (!cu || !us || n < 3 * kXcds) return fail(TDLB_ERR_INVALID, "bad arguments");
    if (!cu->pctr) return fail(TDLB_ERR_STATE, "no one-launch step set up");
    CHK(set_device(cu));
    HIPCHK(hipStreamSynchronize(cu->stream));
    int khz = 0;
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, cu->device));
    uint64_t t[kXcds][4];
    for (int x = 0; x < kXcds; x++)
        HIPCHK(hipMemcpy(t[x], cu->pctr + 64 * x + 8, sizeof(t[x]), hipMemcpyDeviceToHost));
    uint64_t t0 = ~0ull;
    for (int x = 0; x < kXcds; x++)
        if (t[x][0]) t0 = std::min(t0, t[x][0]);
    for (int x = 0; x < kXcds; x++)
        for (int p = 0; p < 3; p++)
            us[3 * x + p] = (t[x][p + 1] && t0 != ~0ull && khz > 0)
                                ? (double)(int64_t)(t[x][p + 1] - t0) * 1e3 / (double)khz
                                : -1.0;
    return TDLB_OK;
}

int tdlb_cu_forcing_info(tdlb_cu_t cu, int64_t* n_static, int64_t* n_rotating, int64_t* n_levels) {
    if (!cu) return fail(TDLB_ERR_INVALID, "null handle");
    if (n_static) *n_static = cu->n_static;
    if (n_rotating) *n_rotating = cu->n_rot;
    if (n_levels) *n_levels = (int64_t)cu->static_set.max_levels + cu->n_blade_classes;
    return TDLB_OK;
}

}  // extern "C"
