// Batches of ComputeUnits on one GPU (no counterpart in the reference, which
// runs one ComputeUnit per process).
//
// A batch holds bound handles of one precision family (float arithmetic),
// streaming and layout.  tdlb_batch_run advances every member with one launch
// per kernel of the main-loop step for the whole batch, the step captured once
// in a HIP graph and replayed.  Each member runs exactly the arithmetic of
// tdlb_cu_run on the same handle (same device functions, same order), so a
// tank's result does not depend on the batch it ran in.
//
// Tile shapes and the block size of the fused kernel come from an autotuner
// that times every candidate (3 tiles x 256 / 512 threads) per grid shape on
// scratch lattices of that shape, or on a handle's own arrays before it is
// initialised (tdlb_batch_tune), never on a member's state; the timings are
// cached per shape.  A run takes the block size with the lower total over its
// members and each member's fastest tile at that size.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/tdlb.h"
#include "handle.h"
#include "kernels.h"
#include "rushton.h"

namespace {

struct Member {
    tdlb_cu* cu = nullptr;
    uint64_t step = 0;   // RunningParams.step after the member's last completed step
    double angle = 0.0;  // RunningParams.angle
    int64_t remaining = 0;
    int32_t tile = TILE_ROW64;
    float ms[2][kTiles] = {};  // autotuned fused-kernel time per block size (256, 512) and tile
    // filled per run
    int64_t call = 0;     // steps left for this member in the current tdlb_batch_run
    int64_t n_steps = 0;
    std::vector<double> angles;  // angle after each step of the run
};

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    int grow(size_t n) {  // returns 1 when reallocated
        if (n <= cap) return 0;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t c = std::max<size_t>(n, 64) * 3 / 2;
        if (hipMalloc((void**)&p, c * sizeof(T)) != hipSuccess) return -1;
        cap = c;
        return 1;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

using TileKey = std::tuple<int64_t, int64_t, int64_t, int, int, int>;
struct TileTimes {
    float ms[2][kTiles];
};
std::mutex g_tile_mu;
std::map<TileKey, TileTimes> g_tiles;  // fused-kernel timings per grid shape, precision, layout, streaming
constexpr int kBlockSizes[2] = {256, 512};
constexpr int kGraphSteps = 8;  // steps in one replay of the long graph

}  // namespace

struct tdlb_batch {
    int prec = 0, layout = 0, streaming = 0, device = 0;
    hipStream_t stream = nullptr;
    std::vector<Member> members;
    int dlayout() const { return layout == TDLB_LAYOUT_IJKL ? LayoutIJKL : LayoutLIJK; }  // layout in HBM
    bool use_graph = true;
    bool autotune = true;

    DevBuf<MemberDev> d_mem;
    DevBuf<int64_t> d_tab;
    DevBuf<PlaneTask> d_tasks;
    DevBuf<BladeStep> d_steps;
    DevBuf<int32_t> d_bm;  // member of each fused-kernel block

    // plan of the last run and its graph
    BatchPlan plan{};
    std::vector<std::pair<tdlb_cu*, int32_t>> key;  // members (with budget) and their tiles
    std::vector<int64_t> launch_sig;                 // launch sizes and arguments of the captured step
    bool buffers_moved = true;
    hipGraph_t graph[2] = {nullptr, nullptr};     // one step, kGraphSteps steps
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int64_t launches_per_step = 0;
    int32_t bs = 256;                                // fused-kernel block size of the plan
    int32_t n_cu = 256;                              // compute units of the device
    std::vector<int> active_idx;  // members in the plan (indices into members)
};

namespace {

double to_prec_f(const tdlb_cu* cu, double v) { return cu->tsize == 4 ? (double)(float)v : v; }

int set_device_of(const tdlb_cu* cu) {
    HIPCHK(hipSetDevice(cu->device));
    return TDLB_OK;
}

void drop_graph(tdlb_batch* b) {
    for (int a = 0; a < 2; a++) {
        if (b->exec[a]) (void)hipGraphExecDestroy(b->exec[a]);
        if (b->graph[a]) (void)hipGraphDestroy(b->graph[a]);
        b->exec[a] = nullptr;
        b->graph[a] = nullptr;
    }
}

MemberDev describe(const tdlb_batch* b, const Member& m) {
    const tdlb_cu* cu = m.cu;
    MemberDev md{};
    md.g = cu->g;
    if (b->prec == TDLB_PREC_F32)
        md.P = make_collide_params<float, double>((float)cu->flow.nu, (float)cu->flow.cs0, (float)cu->flow.g3,
                                                  cu->flow.use_les);
    else
        md.P = make_collide_params<float, float>((float)cu->flow.nu, (float)cu->flow.cs0, (float)cu->flow.g3,
                                                 cu->flow.use_les);
    md.rho0 = cu->rho0;
    md.alfa = (float)cu->alfa;
    md.beta = (float)cu->beta;
    md.tile = m.tile;
    md.Q[0] = cu->Q[0];
    md.Q[1] = cu->Q[1];
    md.F = static_cast<float*>(cu->F);
    md.Nu = cu->Nu;
    md.fmask[0] = cu->fmask[0];
    md.fmask[1] = cu->fmask[1];
    md.static_base = cu->static_base;
    md.fg = cu->fgeom();
    md.fg.local_teams = 0;  // the batch orders workgroups by member, not by XCD
    md.even = cu->even ? 1 : 0;
    md.cur = cu->cur;
    md.fm = cu->fm;
    return md;
}

// Timings of the fused kernel for cu's grid shape: cached per shape, else (with
// autotuning on) timed for every tile and block size on scratch lattices of
// that shape, or on the handle's own arrays when own (zeroed first: the handle
// is about to be initialised).  Zero timings (autotuning off) select the row
// tile at 256 threads.
int tune_shape(tdlb_batch* b, tdlb_cu* cu, bool own, TileTimes* out, hipStream_t s) {
    const TileKey key{cu->g.xg1, cu->g.yg1, cu->g.zg1, b->prec, b->layout, b->streaming};
    {
        std::lock_guard<std::mutex> lk(g_tile_mu);
        auto it = g_tiles.find(key);
        if (it != g_tiles.end()) {
            *out = it->second;
            return TDLB_OK;
        }
    }
    *out = TileTimes{};
    if (!b->autotune) return TDLB_OK;
    const size_t nq = cu->simple() ? 2 : 1;
    void* q[2] = {nullptr, nullptr};
    void *F = nullptr, *Nu = nullptr, *fm = nullptr, *dmem = nullptr, *dtab = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = TDLB_OK;
    auto cleanup = [&]() {
        if (!own)
            for (void* p : {q[0], q[1], F, Nu, fm})
                if (p) (void)hipFree(p);
        for (void* p : {dmem, dtab})
            if (p) (void)hipFree(p);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    bool ok = true;
    if (own) {
        q[0] = cu->Q[0];
        q[1] = cu->Q[nq - 1];
        F = cu->F;
        Nu = cu->Nu;
        fm = cu->fmask[0];
    } else {
        for (size_t a = 0; a < nq && ok; a++) ok = hipMalloc(&q[a], cu->qs_bytes()) == hipSuccess;
        ok = ok && hipMalloc(&F, cu->f_bytes()) == hipSuccess && hipMalloc(&Nu, cu->nu_bytes()) == hipSuccess &&
             hipMalloc(&fm, cu->mask_bytes()) == hipSuccess;
    }
    ok = ok && hipMalloc(&dmem, sizeof(MemberDev)) == hipSuccess &&
         hipMalloc(&dtab, 2 * sizeof(int64_t)) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
         hipEventCreate(&e1) == hipSuccess;
    if (!ok) {
        cleanup();
        return fail(TDLB_ERR_NOMEM, "autotune lattices");
    }
    for (size_t a = 0; a < nq; a++) ok = ok && launch_zero(q[a], cu->qs_bytes(), s) == hipSuccess;
    ok = ok && launch_zero(F, cu->f_bytes(), s) == hipSuccess && launch_zero(Nu, cu->nu_bytes(), s) == hipSuccess &&
         launch_zero(fm, cu->mask_bytes(), s) == hipSuccess;
    TileTimes tt{};
    for (int bi = 0; bi < 2 && ok; bi++)
        for (int32_t tile = 0; tile < kTiles && ok; tile++) {
            MemberDev md{};
            md.g = cu->g;
            md.P = b->prec == TDLB_PREC_F32 ? make_collide_params<float, double>(1e-3f, 0.1f, 0.8f, 1)
                                            : make_collide_params<float, float>(1e-3f, 0.1f, 0.8f, 1);
            md.rho0 = 1.0f;
            md.tile = tile;
            md.Q[0] = q[0];
            md.Q[1] = q[nq - 1];
            md.F = static_cast<float*>(F);
            md.Nu = Nu;
            md.fmask[0] = md.fmask[1] = static_cast<uint64_t*>(fm);
            md.active = 1;
            md.n_steps = 1;
            const int64_t tab[2] = {0, batch_coll_blocks(cu->g, tile, kBlockSizes[bi])};
            BatchPlan p{};
            p.mem = static_cast<const MemberDev*>(dmem);
            p.n_members = 1;
            p.off_coll = static_cast<const int64_t*>(dtab);
            p.tot_coll = tab[1];
            p.bs = kBlockSizes[bi];
            ok = hipMemcpyAsync(dmem, &md, sizeof(md), hipMemcpyHostToDevice, s) == hipSuccess &&
                 hipMemcpyAsync(dtab, tab, sizeof(tab), hipMemcpyHostToDevice, s) == hipSuccess;
            for (int w = 0; w < 2 && ok; w++)
                ok = launch_batch_fused(b->prec, b->dlayout(), b->streaming, p, s) == hipSuccess;
            ok = ok && hipEventRecord(e0, s) == hipSuccess;
            constexpr int reps = 5;
            for (int w = 0; w < reps && ok; w++)
                ok = launch_batch_fused(b->prec, b->dlayout(), b->streaming, p, s) == hipSuccess;
            ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
            float ms = 0.f;
            ok = ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
            tt.ms[bi][tile] = ms / reps;
        }
    if (!ok) rc = fail(TDLB_ERR_HIP, "autotune launch failed");
    (void)hipStreamSynchronize(s);
    cleanup();
    if (rc != TDLB_OK) return rc;
    std::lock_guard<std::mutex> lk(g_tile_mu);
    g_tiles[key] = tt;
    *out = tt;
    return TDLB_OK;
}

// A member needs one step on its own handle first when its state is not the
// fused path's steady state (explicit moments, static stencils not in the mask,
// O written by the caller).
int settle(tdlb_batch* b, Member& m) {
    tdlb_cu* cu = m.cu;
    if (m.remaining <= 0) return TDLB_OK;
    if (cu->phase == PH_LAZY && cu->static_in_prev && !cu->o_dirty) return TDLB_OK;
    CHK(tdlb_cu_run(cu, 1, &m.step, &m.angle));
    HIPCHK(hipStreamSynchronize(cu->stream));
    m.remaining -= 1;
    (void)b;
    return TDLB_OK;
}

// Descriptors, prefix tables, level segments and blade tables of a run of K steps.
int build_plan(tdlb_batch* b, int64_t K, bool* structure_changed) {
    std::vector<int> idx;
    for (size_t a = 0; a < b->members.size(); a++)
        if (b->members[a].call > 0) idx.push_back((int)a);
    const int M = (int)idx.size();
    // block size with the lower total of the members' fastest tiles, then each member's tile at it
    {
        double tot[2] = {0, 0};
        for (int a : idx)
            for (int bi = 0; bi < 2; bi++) {
                const float* t = b->members[a].ms[bi];
                tot[bi] += *std::min_element(t, t + kTiles);
            }
        const int bi = tot[1] < tot[0] ? 1 : 0;
        b->bs = kBlockSizes[bi];
        for (int a : idx) {
            const float* t = b->members[a].ms[bi];
            b->members[a].tile = (int32_t)(std::min_element(t, t + kTiles) - t);
        }
    }
    std::vector<std::pair<tdlb_cu*, int32_t>> key;
    for (int a : idx) key.push_back({b->members[a].cu, b->members[a].tile});

    std::vector<MemberDev> mem(M);
    std::vector<int64_t> tab(3 * (size_t)(M + 1), 0);
    int64_t* off_coll = tab.data();
    int64_t* off_rot = off_coll + (M + 1);
    int64_t* off_words = off_rot + (M + 1);
    std::vector<BladeStep> steps;
    std::vector<size_t> step_base(M);
    // forcing workgroups: the teams of every member first (a team's workgroups
    // must be resident together), then every solo plane, longest first
    std::vector<PlaneTask> team_tasks;
    std::vector<std::pair<float, PlaneTask>> tk;
    int64_t solo_cells = 0;  // static cells of the largest solo plane
    for (int r = 0; r < M; r++) {
        Member& m = b->members[idx[r]];
        tdlb_cu* cu = m.cu;
        m.n_steps = std::min<int64_t>(m.call, K);
        off_coll[r + 1] = off_coll[r] + batch_coll_blocks(cu->g, m.tile, b->bs);
        off_rot[r + 1] = off_rot[r] + cu->n_rot;
        off_words[r + 1] = off_words[r] + (int64_t)cu->mask_words;
        const auto& bl = cu->static_set.block_list;
        const auto& pl = cu->static_set.plane_list;
        for (int32_t b = 0; b < (int32_t)bl.size(); b++) {
            if (bl[b].plane < 0) continue;  // an empty slot of the single-tank launch
            const PlaneDesc& pd = pl[bl[b].plane];
            if (pd.team > 1) {
                team_tasks.push_back(PlaneTask{r, b});
            } else {
                tk.push_back({cu->static_set.plane_cost[bl[b].plane], PlaneTask{r, b}});
                solo_cells = std::max<int64_t>(solo_cells, pd.cell_end - pd.cell_begin);
            }
        }
        // blade orientation of each step (the single-handle loop computes the same)
        Rushton R(cu->rp);
        m.angles.resize(m.n_steps);
        const size_t base = steps.size();
        step_base[r] = base;
        steps.resize(base + m.n_steps);
        double angle = m.angle;
        for (int64_t s = 0; s < m.n_steps; s++) {
            const double delta = to_prec_f(cu, R.step_increment(m.step + 1 + s));
            angle += delta;
            m.angles[s] = angle;
            BladeStep& st = steps[base + s];
            std::memset(&st, 0, sizeof(st));
            R.blade_trig(angle, st.cosb, st.sinb);
            st.omega = delta;
        }
        mem[r] = describe(b, m);
        mem[r].n_steps = m.n_steps;
        mem[r].active = m.n_steps > 0 ? 1 : 0;
        mem[r].done = 0;
    }
    std::stable_sort(tk.begin(), tk.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
    std::vector<PlaneTask> tasks(team_tasks);
    for (const auto& t : tk) tasks.push_back(t.second);
    const int64_t tot_coll = off_coll[M], tot_rot = off_rot[M], tot_words = off_words[M];
    std::vector<int32_t> blk_member((size_t)tot_coll);
    for (int r = 0; r < M; r++)
        std::fill(blk_member.begin() + off_coll[r], blk_member.begin() + off_coll[r + 1], (int32_t)r);

    int moved = 0;
    int g;
    if ((g = b->d_mem.grow(std::max(M, 1))) < 0) return fail(TDLB_ERR_NOMEM, "batch descriptors");
    moved |= g;
    if ((g = b->d_tab.grow(tab.size())) < 0) return fail(TDLB_ERR_NOMEM, "batch tables");
    moved |= g;
    if ((g = b->d_tasks.grow(std::max<size_t>(tasks.size(), 1))) < 0) return fail(TDLB_ERR_NOMEM, "batch tasks");
    moved |= g;
    if (b->d_steps.grow(std::max<size_t>(steps.size(), 1)) < 0) return fail(TDLB_ERR_NOMEM, "batch blade table");
    if ((g = b->d_bm.grow(std::max<size_t>(blk_member.size(), 1))) < 0) return fail(TDLB_ERR_NOMEM, "batch block table");
    moved |= g;
    if (!blk_member.empty())
        HIPCHK(hipMemcpyAsync(b->d_bm.p, blk_member.data(), blk_member.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                              b->stream));
    for (int r = 0; r < M; r++) mem[r].steps = b->d_steps.p + step_base[r];
    if (M > 0) HIPCHK(hipMemcpyAsync(b->d_mem.p, mem.data(), M * sizeof(MemberDev), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d_tab.p, tab.data(), tab.size() * sizeof(int64_t), hipMemcpyHostToDevice, b->stream));
    if (!tasks.empty())
        HIPCHK(hipMemcpyAsync(b->d_tasks.p, tasks.data(), tasks.size() * sizeof(PlaneTask), hipMemcpyHostToDevice,
                              b->stream));
    if (!steps.empty())
        HIPCHK(hipMemcpyAsync(b->d_steps.p, steps.data(), steps.size() * sizeof(BladeStep), hipMemcpyHostToDevice,
                              b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));  // host vectors go out of scope

    BatchPlan& p = b->plan;
    p.mem = b->d_mem.p;
    p.mem_rw = b->d_mem.p;
    p.n_members = M;
    p.off_coll = b->d_tab.p;
    p.coll_member = b->d_bm.p;
    p.off_rot = b->d_tab.p + (M + 1);
    p.off_words = b->d_tab.p + 2 * (M + 1);
    p.tasks = b->d_tasks.p;
    // small forcing workgroups only with enough tasks to keep every CU busy
    // several times over (a batch of one tank runs faster with 1024 lanes)
    p.force_bs = kForcePlaneBlock;
    if ((int64_t)tasks.size() >= 4 * (int64_t)b->n_cu) {
        double cells = 0.0;
        for (int r = 0; r < M; r++) cells += (double)b->members[idx[r]].cu->static_set.n_cells;
        p.force_bs = forcing_block_for(cells / (double)tasks.size());
    }
    if (const char* v = std::getenv("TDLB_BATCH_FORCE_BLOCK")) {
        const int t = std::atoi(v);
        if (t >= 64 && t <= kForcePlaneBlock && t % 64 == 0) p.force_bs = t;
    }
    // The force values of a solo plane stay in LDS through its levels when they
    // fit a share of the CU's LDS that keeps the workgroups resident together
    // that the registers allow (at most 128 VGPRs: 1024 lanes per CU); larger
    // planes keep them in HBM.  All batch formats compute in 4-byte floats.
    // The launch-sized LDS stays within 64 KB, which needs no kernel attribute.
    {
        const int64_t per_cu = std::max<int64_t>(1, kForcePlaneBlock / p.force_bs);
        const int64_t bytes = (int64_t)std::min<size_t>({kPlaneLdsBytes, kCuLdsBytes / (size_t)per_cu, 64 * 1024});
        p.force_lds_cells = std::min<int64_t>(solo_cells, bytes / (3 * (int64_t)sizeof(float)));
    }
    p.tot_coll = tot_coll;
    p.tot_rot = tot_rot;
    p.tot_words = tot_words;
    p.n_tasks = (int64_t)tasks.size();
    p.bs = b->bs;
    // launches: blades, fused, mask, forcing planes, advance
    b->launches_per_step = 5;
    // A captured step bakes in every launch size and kernel argument: recapture
    // whenever any of them differs (member identity alone is not enough, a freed
    // handle's address can come back for a tank of another size).
    std::vector<int64_t> sig = {M, tot_coll, tot_rot, tot_words, p.n_tasks, p.bs, p.force_bs, p.force_lds_cells,
                                (int64_t)(uintptr_t)b->d_mem.p, (int64_t)(uintptr_t)b->d_tab.p,
                                (int64_t)(uintptr_t)b->d_tasks.p, (int64_t)(uintptr_t)b->d_bm.p};
    *structure_changed = moved || sig != b->launch_sig || key != b->key || b->buffers_moved;
    b->buffers_moved = false;
    b->key = key;
    b->launch_sig = sig;
    b->active_idx = idx;
    return TDLB_OK;
}

int record_steps(tdlb_batch* b, int steps) {
    HIPCHK(launch_batch_step(b->prec, b->dlayout(), b->streaming, b->plan, steps, b->stream));
    return TDLB_OK;
}

// graphs of one step and of kGraphSteps steps (the step state lives on the device)
int capture(tdlb_batch* b) {
    drop_graph(b);
    for (int a = 0; a < 2; a++) {
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int r = record_steps(b, a == 0 ? 1 : kGraphSteps);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (r != TDLB_OK) {
            if (g) (void)hipGraphDestroy(g);
            return r;
        }
        if (e != hipSuccess) return fail(TDLB_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        b->graph[a] = g;
        HIPCHK(hipGraphInstantiate(&b->exec[a], g, nullptr, nullptr, 0));
    }
    return TDLB_OK;
}

Member* find_member(tdlb_batch* b, tdlb_cu* cu) {
    for (auto& m : b->members)
        if (m.cu == cu) return &m;
    return nullptr;
}

}  // namespace

extern "C" {

int tdlb_batch_create(int32_t precision, int32_t layout, int32_t streaming, int32_t device_id, tdlb_batch_t* out) {
    if (!out) return fail(TDLB_ERR_INVALID, "null argument");
    if (precision != TDLB_PREC_F32 && precision != TDLB_PREC_F16 && precision != TDLB_PREC_BF16 &&
        precision != TDLB_PREC_E3M12)
        return fail(TDLB_ERR_UNSUPPORTED, "batches run the float-arithmetic precisions (fp32, fp16, bf16, e3m12)");
    if (layout != TDLB_LAYOUT_IJKL && layout != TDLB_LAYOUT_LIJK && layout != TDLB_LAYOUT_IJKL_SOA)
        return fail(TDLB_ERR_INVALID, "layout");
    if (streaming != TDLB_STREAM_SIMPLE && streaming != TDLB_STREAM_ESOTWIST) return fail(TDLB_ERR_INVALID, "streaming");
    tdlb_batch* b = new tdlb_batch();
    b->prec = precision;
    b->layout = layout;
    b->streaming = streaming;
    b->device = device_id < 0 ? 0 : device_id;
    if (hipSetDevice(b->device) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return fail(TDLB_ERR_HIP, "batch stream");
    }
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device) == hipSuccess && ncu > 0)
        b->n_cu = ncu;
    *out = b;
    return TDLB_OK;
}

int tdlb_batch_destroy(tdlb_batch_t b) {
    if (!b) return TDLB_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    b->d_mem.release();
    b->d_tab.release();
    b->d_tasks.release();
    b->d_steps.release();
    b->d_bm.release();
    (void)hipStreamDestroy(b->stream);
    delete b;
    return TDLB_OK;
}

int tdlb_batch_set_options(tdlb_batch_t b, int32_t use_graph, int32_t autotune) {
    if (!b) return fail(TDLB_ERR_INVALID, "null batch");
    b->use_graph = use_graph != 0;
    b->autotune = autotune != 0;
    b->buffers_moved = true;  // recapture on the next run
    return TDLB_OK;
}

int tdlb_batch_add(tdlb_batch_t b, tdlb_cu_t cu, uint64_t step, double angle, int64_t nsteps) {
    if (!b || !cu || nsteps < 0) return fail(TDLB_ERR_INVALID, "bad arguments");
    if (cu->prm.precision != b->prec || cu->prm.layout != b->layout || cu->prm.streaming != b->streaming)
        return fail(TDLB_ERR_INVALID, "member precision, layout and streaming must match the batch");
    if (cu->device != b->device) return fail(TDLB_ERR_INVALID, "member is on another device");
    if (!cu->bound) return fail(TDLB_ERR_STATE, "member has no bound geometry (tdlb_cu_bind_rushton)");
    if (cu->g.xg1 < 3 || cu->g.zg1 < 3) return fail(TDLB_ERR_UNSUPPORTED, "members need at least 3 cells in x and z");
    if (find_member(b, cu)) return fail(TDLB_ERR_INVALID, "handle is already in the batch");
    HIPCHK(hipSetDevice(cu->device));
    Member m;
    m.cu = cu;
    m.step = step;
    m.angle = angle;
    m.remaining = nsteps;
    TileTimes tt{};
    CHK(tune_shape(b, cu, false, &tt, b->stream));
    std::memcpy(m.ms, tt.ms, sizeof(m.ms));
    b->members.push_back(m);
    b->buffers_moved = true;  // new member: recapture
    return TDLB_OK;
}

int tdlb_batch_tune(tdlb_batch_t b, tdlb_cu_t cu) {
    if (!b || !cu) return fail(TDLB_ERR_INVALID, "null argument");
    if (cu->prm.precision != b->prec || cu->prm.layout != b->layout || cu->prm.streaming != b->streaming)
        return fail(TDLB_ERR_INVALID, "precision, layout and streaming must match the batch");
    if (cu->g.xg1 < 3 || cu->g.zg1 < 3) return fail(TDLB_ERR_UNSUPPORTED, "members need at least 3 cells in x and z");
    CHK(set_device_of(cu));
    TileTimes tt{};
    return tune_shape(b, cu, true, &tt, cu->stream);  // the arrays it wrote are reset by tdlb_cu_initialise
}

int tdlb_batch_shape(tdlb_batch_t b, int32_t* block_size, int32_t* graph_steps, int32_t* forcing_block) {
    if (!b) return fail(TDLB_ERR_INVALID, "null batch");
    if (block_size) *block_size = b->bs;
    if (graph_steps) *graph_steps = b->use_graph ? kGraphSteps : 1;
    if (forcing_block) *forcing_block = b->plan.force_bs;
    return TDLB_OK;
}

int tdlb_batch_remove(tdlb_batch_t b, tdlb_cu_t cu) {
    if (!b || !cu) return fail(TDLB_ERR_INVALID, "null argument");
    for (size_t a = 0; a < b->members.size(); a++)
        if (b->members[a].cu == cu) {
            b->members.erase(b->members.begin() + (long)a);
            return TDLB_OK;
        }
    return fail(TDLB_ERR_INVALID, "handle is not in the batch");
}

int tdlb_batch_member(tdlb_batch_t b, tdlb_cu_t cu, uint64_t* step, double* angle, int64_t* remaining) {
    if (!b || !cu) return fail(TDLB_ERR_INVALID, "null argument");
    Member* m = find_member(b, cu);
    if (!m) return fail(TDLB_ERR_INVALID, "handle is not in the batch");
    if (step) *step = m->step;
    if (angle) *angle = m->angle;
    if (remaining) *remaining = m->remaining;
    return TDLB_OK;
}

int tdlb_batch_run(tdlb_batch_t b, int64_t nsteps, int64_t* steps_run) {
    if (!b || nsteps < 0) return fail(TDLB_ERR_INVALID, "bad arguments");
    HIPCHK(hipSetDevice(b->device));
    if (steps_run) *steps_run = 0;
    if (nsteps == 0) return TDLB_OK;
    // every member takes min(nsteps, remaining) steps; one whose state is not
    // the steady fused state takes the first of them alone
    int64_t most = 0, K = 0;
    for (auto& m : b->members) {
        HIPCHK(hipStreamSynchronize(m.cu->stream));
        m.call = std::min<int64_t>(m.remaining, nsteps);
        most = std::max(most, m.call);
        const int64_t before = m.remaining;
        if (m.call > 0) CHK(settle(b, m));
        m.call -= before - m.remaining;
        K = std::max(K, m.call);
    }
    if (steps_run) *steps_run = most;
    if (K <= 0) return TDLB_OK;
    bool changed = false;
    CHK(build_plan(b, K, &changed));
    if (b->plan.n_members == 0) return TDLB_OK;
    if (b->use_graph) {
        if (changed || !b->exec[0]) CHK(capture(b));
        int64_t s = 0;
        for (; s + kGraphSteps <= K; s += kGraphSteps) HIPCHK(hipGraphLaunch(b->exec[1], b->stream));
        for (; s < K; s++) HIPCHK(hipGraphLaunch(b->exec[0], b->stream));
    } else {
        for (int64_t s = 0; s < K; s++) CHK(record_steps(b, 1));
    }
    std::vector<MemberDev> mem(b->plan.n_members);
    HIPCHK(hipMemcpyAsync(mem.data(), b->d_mem.p, mem.size() * sizeof(MemberDev), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (size_t r = 0; r < mem.size(); r++) {
        Member& m = b->members[b->active_idx[r]];
        tdlb_cu* cu = m.cu;
        const int64_t done = mem[r].done;
        cu->even = mem[r].even != 0;
        cu->cur = mem[r].cur;
        cu->fm = mem[r].fm;
        if (done > 0) {
            m.step += (uint64_t)done;
            m.angle = m.angles[done - 1];
            m.remaining -= done;
        }
        m.call = 0;
    }
    return TDLB_OK;
}

int tdlb_batch_info(tdlb_batch_t b, int64_t* n_members, int64_t* launches_per_step, int32_t* tiles, int64_t cap) {
    if (!b) return fail(TDLB_ERR_INVALID, "null batch");
    if (n_members) *n_members = (int64_t)b->members.size();
    if (launches_per_step) *launches_per_step = b->launches_per_step;
    if (tiles)
        for (int64_t a = 0; a < cap && a < (int64_t)b->members.size(); a++) tiles[a] = b->members[a].tile;
    return TDLB_OK;
}

}  // extern "C"
