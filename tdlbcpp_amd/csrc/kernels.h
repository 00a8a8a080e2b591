// Host-side launchers of the gfx950 kernels (implemented in kernels.hip).
// Pointers are device pointers; every launcher enqueues on `s` and returns the
// launch status.  prec: 0 float, 1 double, 2 fp16 storage, 3 bf16 storage (the
// last two compute in float and store Nu in their 16-bit format; F and point
// arrays are float).
// layout: LayoutIJKL/LayoutLIJK.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include "device.h"
#include "es_collide.h"

namespace tdlb {

enum StepMode : int {
    MODE_COLL = 0,      // moments in Q -> collide -> populations (ComputeUnit::collision)
    MODE_MOM = 1,       // populations in Q -> moments (ComputeUnit::moments)
    MODE_FUSED = 2,     // streamed populations -> moments -> collide -> populations
    MODE_MOM_PULL = 3,  // Simple only: pull-stream from src -> moments into dst
};

// How IB forcing obtains density and momentum of a stencil cell.
enum MomSource : int {
    MOMSRC_STORED = 0,     // Q holds moments (read through the streaming accessor)
    MOMSRC_ESO_POPS = 1,   // Esotwist: Q holds streamed populations at the current parity
    MOMSRC_PULL_POPS = 2,  // Simple: Q holds post-collision populations, stream by pull
};

struct CollideArgs {
    double nu, cs0, g3;  // converted to the CU's type on the device side
    int use_les;
};

// IB points on the device, structure of arrays in the CU's type.
struct PointsDev {
    int32_t* i;  // cell index including the ghost offset
    int32_t* j;
    int32_t* k;
    void* ifrac;  // T arrays
    void* kfrac;
    void* u;
    void* v;
    void* w;
};

// Blade layout for on-device generation of the rotating (updating) IB points.
struct BladeLayout {
    int32_t n_blades;
    int32_t n_bcol;          // blade colours (blade_colours of the host generator)
    int32_t j_lo, n_planes;  // first plane (pre-ghost j) and number of planes
    int32_t n_radial;        // points per blade per plane
    double r_inner;          // radius of the first point; spacing 1
    double cx, cz;           // axis position (pre-ghost grid coordinates)
    int32_t ghost;
};
struct BladeStep {
    double cosb[16], sinb[16];  // per-blade orientation this step
    double omega;               // angular increment this step (rad / step)
};

hipError_t launch_fill_q(int prec, int layout, void* Q, const Grid& g, double rho, hipStream_t s);
hipError_t launch_zero(void* p, size_t bytes, hipStream_t s);
// One pass of the access-pattern probe: `planes` (1..20) planes of n values of
// elem_bytes (2, 4, 8) bytes, `stride` values apart from base, each value read
// and written back (bits flipped) by one lane; nt: non-temporal hints.
hipError_t launch_probe_planes(int elem_bytes, void* base, uint64_t stride, uint64_t n, int planes, int nt,
                               hipStream_t s);
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);

// Populations in Q (stored type), moments in mout (arithmetic type, Q's shape
// and addressing).  For prec 0/1 mout is Q itself.  For the reduced-precision
// storage (prec 2 fp16, 3 bf16: populations kept as k_d (f_d - w_d rho0))
// MODE_COLL reads the moments from mout and MODE_MOM / MODE_MOM_PULL write them
// there; F is float, Nu in the 16-bit format.
hipError_t launch_esotwist(int prec, int layout, int mode, int even, void* Q, void* mout, const void* F, void* Nu,
                           const uint64_t* fmask, const Grid& g, const CollideArgs& c, float rho0, hipStream_t s);
hipError_t launch_simple(int prec, int layout, int mode, const void* src, void* dst, void* mout, const void* F,
                         void* Nu, const uint64_t* fmask, const Grid& g, const CollideArgs& c, float rho0,
                         hipStream_t s);

// element size 4 or 8: pure data movement on every cell
hipError_t launch_push_all(int esize, int layout, const void* src, void* dst, const Grid& g, hipStream_t s);
hipError_t launch_pull_all(int esize, int layout, const void* src, void* dst, const Grid& g, hipStream_t s);
hipError_t launch_bounce_back(int esize, int layout, void* Q, const Grid& g, hipStream_t s);
hipError_t launch_ghost_fill(int esize, int layout, const void* src, void* dst, const Grid& g, hipStream_t s);
// Q (elements of esize bytes, 4 or 8) between the kernels' population array
// (layout dlayout, padded rows) and the dense size x 18 array of the boundary
// (layout blayout): to_boundary != 0 reads the former and writes the latter.
hipError_t launch_q_boundary(int esize, int dlayout, int blayout, int to_boundary, const void* src, void* dst,
                             const Grid& g, hipStream_t s);
// ghost fill from reduced-precision populations (prec 2/3) into a float array
hipError_t launch_ghost_fill_decode(int prec, int layout, const void* src, void* dst, const Grid& g, float rho0,
                                    hipStream_t s);
hipError_t launch_ghost_copy(int esize, int layout, const void* src, void* dst, const Grid& g, hipStream_t s);

hipError_t launch_mask_from_f(int prec, const void* F, uint64_t* mask, uint64_t* gmask, const Grid& g,
                              hipStream_t s);
hipError_t launch_masked_f(int prec, const void* F, const uint64_t* mask, void* out, const Grid& g, hipStream_t s);
hipError_t launch_zero_interior_o(uint8_t* O, const Grid& g, hipStream_t s);

hipError_t launch_force_prepare(int prec, const PointsDev& p, int64_t begin, int64_t count, void* F,
                                const uint64_t* prev, uint64_t* next, const Grid& g, hipStream_t s);
hipError_t launch_force_level(int prec, int layout, int streaming, int momsrc, int even, const PointsDev& p,
                              int64_t begin, int64_t count, void* F, const void* Q, const Grid& g, double alfa,
                              double beta, float rho0, hipStream_t s);

// Density and momentum of IB stencil cells, 4 values of T per entry {rho, jx, jy, jz}.
// cells: sorted cell indices (one entry per cell).
hipError_t launch_ib_moments_cells(int prec, int layout, int streaming, int momsrc, int even, const uint32_t* cells,
                                   int64_t n, void* cache, const void* Q, const Grid& g, float rho0, hipStream_t s);
// one entry per (point, stencil cell): entry 9*t + s for point begin + t
hipError_t launch_ib_moments_points(int prec, int layout, int streaming, int momsrc, int even, const PointsDev& p,
                                    int64_t begin, int64_t count, void* cache, const void* Q, const Grid& g,
                                    float rho0, hipStream_t s);
// One forcing level reading stencil moments from a cache.  sidx: per point 9 cache
// indices (static lists); null: the entries of point begin+t are 9*(t + cache_base) ...
hipError_t launch_force_level_cached(int prec, const PointsDev& p, int64_t begin, int64_t count,
                                     const int32_t* sidx, int64_t cache_base, const void* cache, void* F,
                                     const Grid& g, double alfa, double beta, hipStream_t s);
// ---- IB forcing of a bound tank by j planes.  Stencils span i and k at the
// point's j, so points of different planes never share a cell and each plane
// is an independent task: a workgroup (or a team of workgroups for a plane
// with many points, such as the impeller disk's) runs a plane's blade
// occupancy bits, the density and momentum of its stencil cells, then its
// static levels and blade colour classes in list order, with a barrier
// between levels instead of a kernel launch.
struct PlaneDesc {
    int32_t j;                     // plane index including the ghost offset
    int32_t n_levels;              // static levels of the plane
    int64_t cell_begin, cell_end;  // its unique static stencil cells in ForceGeom::cells / scache
    int64_t lev;                   // ForceGeom::lev_off[lev .. lev + n_levels]: point offsets of its levels
    int32_t team;                  // workgroups running the plane together
    int32_t bar;                   // its barrier in ForceGeom::bar (team > 1), else -1
};
// workgroup -> (plane, rank in the plane's team); plane -1: an empty slot
struct PlaneBlock {
    int32_t plane, rank;
};
struct ForceGeom {
    PointsDev pts;           // static points [0, n_static) plane by plane, level by level; blades after
    int64_t n_static, n_rot;
    const uint32_t* cells;   // unique static stencil cells, grouped by plane
    const int32_t* sidx;     // per static point: the 9 positions of its stencil cells in cells
    void* scache;            // {rho, jx, jy, jz} in T per entry of cells
    void* rcache;            // {rho, jx, jy, jz} in T x 9 per blade point
    const PlaneDesc* planes;
    int32_t n_planes;
    const PlaneBlock* blocks;  // team workgroups first, then the longest tasks first
    int32_t n_blocks;
    const int64_t* lev_off;
    uint32_t* bar;           // kBarWords words per team: arrivals, generation, flags, XCC ids seen
    BladeLayout blade;       // n_blades 0: no blades
    int32_t local_teams;     // 1: a team's workgroups sit kXcds apart in blocks, so one XCD runs them
                             // (launches deal workgroups to the XCDs in turn)
};
constexpr int kBarWords = 4;
// points one pass of a plane task handles: 16 wavefronts of 21 groups of 3 lanes
constexpr int kForcePlanePoints = 336;
constexpr int kForcePlaneBlock = 1024;  // threads of a plane task's workgroup, at most
// LDS of a single tank's plane-task workgroup for the force values of a solo
// plane's static cells: 8192 cells in FP32 (6570 per plane at 512^3).  The
// workgroup's registers already hold a CU to one workgroup.
constexpr size_t kPlaneLdsBytes = 96 * 1024;
constexpr size_t kCuLdsBytes = 160 * 1024;  // LDS of a gfx950 CU
// A single tank's plane keeps its force values in LDS only from this many
// static cells on: with few planes per CU (tanks to about 256^3) the pass is
// not bandwidth-bound and the copy in and out costs more than the levels save
// (64 tanks of 48..160^3 one at a time: 3.5 % slower with every plane in LDS;
// 512^3, 6570 cells per plane: 4.5 % faster).  Batches, whose thousands of
// tasks keep every CU gathering, use LDS for every plane that fits.
constexpr int64_t kPlaneLdsMinCells = 4096;
// Workgroup of a batch's forcing planes for a mean of `cells` static stencil
// cells per task: a lane per ~4 cells, a power of two from 128 to
// kForcePlaneBlock lanes (batches of 48^3 tanks: 128 lanes 1.7x faster than
// 1024; the mixed 48..160 batch: 256 lanes 9 % faster; 160^3 tanks: 256
// lanes 2 % faster, 512 lanes 4 %).  A batch has thousands of tasks; a single
// tank's few hundred planes leave CUs idle anyway and run faster with 1024
// lanes (512^3: 0.653 against 0.734 ms per pass with 512 lanes).
inline int forcing_block_for(double cells) {
    int t = 128;
    while (t < kForcePlaneBlock && 2.0 * t <= cells / 4.0) t *= 2;
    return t;
}
// One forcing pass of a bound tank (every plane of fg): next must already hold
// the static occupancy bits; blade stencil cells missing from prev get F = 0
// and their bit in next.
hipError_t launch_force_planes(int prec, int layout, int streaming, int momsrc, int even, const ForceGeom& fg,
                               const void* Q, void* F, const uint64_t* prev, uint64_t* next, const Grid& g,
                               double alfa, double beta, float rho0, hipStream_t s);

// ---- One Esotwist main-loop step in one persistent launch (LIJK, float
// arithmetic): the fused stream-collide of every row, the forcing gather of
// each row soon after the rows it reads are written (lines still in L2 or the
// memory-side cache), then the forcing levels of every plane.  Each XCD takes
// a contiguous range of j planes from its own work queue (the hardware XCC id
// picks the queue), so the rows a gather reads were written through the same
// L2; only the two edge planes of a range wait for the neighbouring XCDs.
constexpr int kXcds = 8;
struct PipeArgs {
    void* Q;                      // the lattice (stored type)
    void* F;                      // float
    void* Nu;                     // stored Nu type
    const uint64_t* prev;         // occupancy read by this step's collision
    uint64_t* next;               // occupancy after this step's forcing
    const uint64_t* static_base;  // next starts from these bits
    ForceGeom fg;
    const int32_t* rowcell;       // fg.cells index of the first static stencil cell of row (i, j):
                                  // rowcell[j * (xg1 + 2) + i], i = 1 .. xg1 + 1 (end of the plane)
    uint32_t* ctr;                // pipe_counter_words(g) words, zero before the launch
};
inline size_t pipe_counter_words(const Grid& g) { return (size_t)kXcds * (64 + (size_t)g.xg1 + 2); }
// planes [jlo, jhi] of XCD x: the interior planes split into kXcds contiguous ranges
TDLB_HD inline void pipe_range(int yg1, int x, int& jlo, int& jhi) {
    const int base = yg1 / kXcds, rem = yg1 % kXcds;
    jlo = 1 + x * base + (x < rem ? x : rem);
    jhi = jlo + base + (x < rem ? 1 : 0) - 1;
}
// prec 0, 2, 3, 4 (float arithmetic); hint: the streaming cache hint of the
// fused kernel (line_rows); blocks: the persistent grid
hipError_t launch_eso_pipe(int prec, int even, int hint, const PipeArgs& a, const Grid& g, const CollideArgs& c,
                           float rho0, double alfa, double beta, uint32_t blocks, uint32_t lanes, hipStream_t s);
// lanes per row block of the stream-collide kernels for this grid (the pipe's block size)
uint32_t row_lanes(const Grid& g);
// every k-row of the population planes (stored values of ssize bytes) is whole 128 B lines
bool line_rows_bytes(const Grid& g, size_t ssize);
// 1 when workgroups of a launch see XCC ids 0 .. kXcds - 1 (each XCD's queue has workers)
int pipe_xcc_check(hipStream_t s);

// FP64 sums over the interior cells: {rho, jx, jy, jz, Fx, Fy, Fz, torque}.  The
// torque sum x*Fz - z*Fx runs over cells whose centre (x, z) relative to the
// axis (cx, cz) lies within r2 (squared radius; negative: none).
constexpr int kReduceVals = 8;
constexpr int kReduceBlocks = 1024;
struct ReduceArgs {
    double cx, cz, r2;
    int ghost;
};
// partial: kReduceBlocks * kReduceVals doubles; out: kReduceVals doubles (device)
hipError_t launch_reduce(int prec, int layout, int streaming, int momsrc, int even, const void* Q, const void* F,
                         const uint64_t* fmask, const Grid& g, const ReduceArgs& ra, float rho0, double* partial,
                         double* out, hipStream_t s);
// {rho, ux, uy, uz} per interior cell in T, interior cells in (i, j, k) order
hipError_t launch_macroscopic(int prec, int layout, int streaming, int momsrc, int even, const void* Q,
                              const void* F, const uint64_t* fmask, const Grid& g, float rho0, void* out,
                              hipStream_t s);

// doubleResolutionFullCU: Q (esize 4/8) and a per-cell array (nu_esize 2/4/8)
// of the old grid og copied into the doubled grid ng
hipError_t launch_regrid(int esize, int layout, const void* oldQ, void* newQ, int nu_esize, const void* oldNu,
                         void* newNu, const Grid& og, const Grid& ng, hipStream_t s);
// n Nu values between the 16-bit storage of prec 2/3 and float (to_float != 0:
// decode src into the float dst; else encode)
hipError_t launch_nu_convert(int prec, int to_float, const void* src, void* dst, uint64_t n, hipStream_t s);

// records of savePlaneXZ<float, 4> / <float, 3> (tDiskGrid, Output.hpp:22-25:
// three uint16 then the floats at offset 8); the two alignment bytes are
// written as zero here (the reference leaves them unset)
struct PlaneRec4 {
    uint16_t i, j, k, pad;
    float q[4];
};
struct PlaneRec3 {
    uint16_t i, j, k, pad;
    float q[3];
};
// one record slot and one flag byte per plane cell (x * z, i-major); prec 0/1 (explicit Q type)
hipError_t launch_plane_xz(int prec, int layout, const void* Q, const void* F, const uint64_t* fmask,
                           const uint8_t* exclude, const Grid& g, int cut, int64_t i0, int64_t k0, PlaneRec4* q,
                           PlaneRec3* f, uint8_t* flags, hipStream_t s);
// vorticity field of plane j = cut (axis 0) or k = cut (axis 1), T per cell
hipError_t launch_vorticity(int prec, int layout, int streaming, int momsrc, int even, const void* Q, const Grid& g,
                            int axis, int cut, const uint8_t* exclude, void* out, float rho0, hipStream_t s);

hipError_t launch_gen_blades(int prec, const BladeLayout& bl, const BladeStep& st, const PointsDev& out,
                             int64_t offset, hipStream_t s);
hipError_t launch_solid_mask(double cx, double cz, double radius, uint8_t* out, const Grid& g, int ghost,
                             hipStream_t s);

// ---------------------------------------------------------------- batches of ComputeUnits
// Many bound ComputeUnits (float arithmetic: prec 0, 2, 3; one streaming and
// layout) advance together: every kernel of a main-loop step is one launch over
// all members, blocks mapped to members through prefix tables.  Per-member
// state that changes from step to step (parity, lattice, occupancy mask, step
// index) lives in MemberDev on the device and is advanced by a kernel, so one
// step is a fixed launch sequence that can be captured in a HIP graph.
enum Tile : int32_t {
    TILE_ROW64 = 0,  // a wavefront takes 64 consecutive k of one (i, j) row
    TILE_PLANE = 1,  // a block takes consecutive cells of one i plane (ghost cells idle)
    TILE_FLAT = 2,   // a block takes consecutive interior cells of the tank, rows back to back
};
constexpr int kTiles = 3;

struct MemberDev {
    Grid g;
    CollideParams<float> P;
    float rho0;
    float alfa, beta;
    int32_t tile;
    void* Q[2];  // Esotwist: Q[0]; Simple: the two lattices
    float* F;
    void* Nu;  // in the member's Nu storage type (float, or its 16-bit format)
    uint64_t* fmask[2];
    const uint64_t* static_base;
    ForceGeom fg;            // IB points, stencil caches and plane tasks of the member
    const BladeStep* steps;  // blade orientation of step done + 1, done + 2, ...
    int64_t n_steps;         // steps this member may still take in the batch
    // advanced on the device after every step
    int32_t even, cur, fm, active;
    int64_t done;
};

// one workgroup of the batch's forcing launch: member and its PlaneBlock
struct PlaneTask {
    int32_t member, block;
};

// Device tables of one batch step.  off_* are member prefix sums (n_members + 1
// entries) of blocks / threads; tasks: every plane of every member, heaviest
// first.
struct BatchPlan {
    const MemberDev* mem;
    MemberDev* mem_rw;
    int32_t n_members;
    const int64_t* off_coll;   // blocks of the fused kernel
    const int32_t* coll_member;  // member of each block of the fused kernel (or null: search off_coll)
    const int64_t* off_rot;    // blade points
    const int64_t* off_words;  // occupancy-mask words
    const PlaneTask* tasks;
    int64_t tot_coll, tot_rot, tot_words, n_tasks;
    int32_t bs;  // threads per block of the fused kernel (256 or 512)
    int32_t force_bs;  // threads per workgroup of the forcing pass (64 .. kForcePlaneBlock, a multiple of 64)
    int64_t force_lds_cells;  // force values of a solo plane task kept in LDS, at most (0: none)
};

// Enqueue `steps` main-loop steps of every active member: per step blade
// generation, the fused stream-collide, the static occupancy bits, the forcing
// planes, advance.
hipError_t launch_batch_step(int prec, int layout, int streaming, const BatchPlan& plan, int steps, hipStream_t s);
// blocks of the fused kernel for one member with the given tile and block size
int64_t batch_coll_blocks(const Grid& g, int tile, int bs);
// The fused kernel alone over every member (autotuning; advances no state).
hipError_t launch_batch_fused(int prec, int layout, int streaming, const BatchPlan& plan, hipStream_t s);

}  // namespace tdlb
