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
    TILE_PLANE = 1,  // a block takes 256 consecutive cells of one i plane (ghost cells idle)
};

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
    PointsDev pts;  // static points [0, n_static), blades [n_static, n_static + n_rot)
    int64_t n_static, n_rot;
    const uint32_t* scells;  // unique stencil cells of the static points
    const int32_t* sidx;     // 9 per static point into scells / scache
    int64_t n_scells;
    void* scache;  // Mom4<float> per unique static cell
    void* rcache;  // Mom4<float> x 9 per blade point
    BladeLayout blade;
    const BladeStep* steps;  // blade orientation of step done + 1, done + 2, ...
    int64_t n_steps;         // steps this member may still take in the batch
    // advanced on the device after every step
    int32_t even, cur, fm, active;
    int64_t done;
};

// points [begin, begin + count) of member `member` in one forcing level
struct LevelSeg {
    int32_t member, pad;
    int64_t begin, count;
};

// Device tables of one batch step.  off_* are member prefix sums (n_members + 1
// entries) of blocks / threads; level l of the static (blade) lists is
// segs[lv_*[l] .. lv_*[l + 1]) with thread prefix sums in seg_cum.
struct BatchPlan {
    const MemberDev* mem;
    MemberDev* mem_rw;
    int32_t n_members;
    const int64_t* off_coll;    // blocks of the fused kernel
    const int64_t* off_rot;     // blade points
    const int64_t* off_words;   // occupancy-mask words
    const int64_t* off_scells;  // unique static stencil cells
    const LevelSeg* segs;
    const int64_t* seg_cum;
    int64_t tot_coll, tot_rot, tot_words, tot_scells;
    // host copies of the level structure (launch sizes)
    int32_t n_slevels, n_rlevels;
    const int64_t* lv_static;   // host: n_slevels + 1 segment offsets
    const int64_t* lv_blade;    // host: n_rlevels + 1 segment offsets
    const int64_t* lv_threads;  // host: threads of static levels then blade levels
};

// Enqueue one main-loop step of every active member (fused stream-collide,
// blade generation, occupancy masks, moment caches, forcing levels, advance).
hipError_t launch_batch_step(int prec, int layout, int streaming, const BatchPlan& plan, hipStream_t s);
// blocks of the fused kernel for one member with the given tile
int64_t batch_coll_blocks(const Grid& g, int tile);
// The fused kernel alone over every member (autotuning; advances no state).
hipError_t launch_batch_fused(int prec, int layout, int streaming, const BatchPlan& plan, hipStream_t s);

}  // namespace tdlb
